"""Benchmark: single-vector matrix-free H·v on the 64-spin pyrochlore ice basis.

Workload (one GPU): pyrochlore supercell diag(2,2,4) on the fcc primitive
vectors (64 spins), Q=0 sector (all 2-in-2-out states, enumerated here),
H = -sum_hex g_s (O_hex + O_hex^dagger) - sum_i (B.z_sl(i)) S^z_i with
g_s ~ U[0.2, 1.0] per hexagon sublattice and |B| = 0.1 uniform on the sphere
(seeded), FP64 vectors.  A step is one y = H x over the whole basis with x, y
and the basis resident in HBM.

Prints one JSON line: basis states/s (value), the apply kernel's roofline
figures from HIP events, Lanczos and eigensolver rates, and the CPU baseline
(the oracle's OpenMP push apply on this host, same workload).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, AMD published
PMC_SUMMARY = os.path.join(ROOT, "profiles", "apply_64spin_pmc.json")
CALIB_BYTES = 64 << 20  # the calibration copy reads and writes this many bytes


def measured_traffic():
    """HBM bytes per apply from the committed rocprofv3 PMC summary
    (FETCH_SIZE + WRITE_SIZE of the apply kernel, each divided by what the
    same counter reads for a copy of known size), or None."""
    try:
        with open(PMC_SUMMARY) as f:
            return float(json.load(f)["traffic_bytes_per_apply"])
    except (OSError, KeyError, ValueError):
        return None


def build_workload(seed=1234, n_threads=16):
    import qsi_subspace_ed_amd as q
    jd = q.load_lattice(q.lattice_path("pyro_2,0,0_0,2,0_0,0,4.json"))
    rng = np.random.default_rng(seed)
    g = rng.uniform(0.2, 1.0, size=4)
    b = rng.normal(size=3)
    B = 0.1 * b / np.linalg.norm(b)
    t0 = time.perf_counter()
    states = q.enumerate_basis(jd, 0, n_threads=n_threads)
    t_enum = time.perf_counter() - t0
    H = q.build_hamiltonian_g(jd, [-x for x in g])
    q.add_local_field(H, jd, B)
    return jd, states, H, g, B, t_enum


def cpu_baseline(states, H, x, budget_s=12.0, threads=None):
    """Oracle push apply (the reference's OpenMP algorithm) on this host."""
    from oracle import pyoracle as orc
    lib = orc.load()
    if threads:
        lib.oracle_set_threads(int(threads))
    nthr = lib.oracle_num_threads()
    ob = orc.Basis(states)
    D = states.shape[0]
    y = np.zeros(D)
    t0 = time.perf_counter()
    orc.evaluate_add(ob, H, x, y)  # warm-up
    t_first = time.perf_counter() - t0
    times = []
    n = max(1, min(5, int(budget_s / max(t_first, 1e-3))))
    for _ in range(n):
        y[:] = 0.0
        t0 = time.perf_counter()
        orc.evaluate_add(ob, H, x, y)
        times.append(time.perf_counter() - t0)
    t = float(np.median(times))
    return {"value": D / t, "unit": "basis-states/s", "cores": int(nthr), "kind": "port",
            "sample": f"{n} full applies (after 1 warm-up) of the same 64-spin workload, "
                      f"median {t * 1e3:.1f} ms per apply"}, y


def bench_ftlm(R=64, M=100, seed=77, plain_samples=4):
    """Batched FTLM on the 48-spin diag(2,2,3) lattice: sectors
    n_spinon_pairs = 0 and 1 in one block launch, R random-sign start vectors
    each, M Lanczos steps, FP64, g_s ~ U[0.2, 1], B = 0.  Compared with the
    plain path: one single-vector Lanczos solve of M steps per sample, one
    after another (timed on plain_samples samples per sector)."""
    import torch
    import qsi_subspace_ed_amd as q
    from qsi_subspace_ed_amd import ftlm
    jd = q.load_lattice(q.lattice_path("pyro_2,0,0_0,2,0_0,0,3.json"))
    rng = np.random.default_rng(seed)
    g = rng.uniform(0.2, 1.0, size=4)
    H = q.build_hamiltonian_g(jd, [-x for x in g])
    bases = [q.enumerate_basis(jd, n, n_threads=16) for n in (0, 1)]
    ops = [q.GpuOp(q.GpuBasis(b, 48), H) for b in bases]
    V0s = [torch.from_numpy(ftlm.random_sign_block(b.shape[0], R, seed + k)).cuda()
           for k, b in enumerate(bases)]
    ftlm.lanczos_block(ops, V0s, 3)  # warm-up (graph capture, code load)
    t0 = time.perf_counter()
    al, be, dev_ms = ftlm.lanczos_block(ops, V0s, M, use_graph=True)
    t_batched = time.perf_counter() - t0
    _, _, dev_ms_nograph = ftlm.lanczos_block(ops, V0s, M, use_graph=False)
    # thermodynamics of sector 0 from the batch (a sanity figure)
    temps = np.array([0.05, 0.1, 0.2, 0.5, 1.0])
    th = ftlm.thermodynamics(ftlm.quadrature(al[0], be[0]), bases[0].shape[0], temps)
    # plain: single-vector Lanczos solves, one at a time
    plain = []
    for op in ops:
        st = q.Settings(krylov_dim=M, abs_tol=0.0, rel_tol=0.0, min_iterations=10**9,
                        max_iterations=M, x0_seed=0, calc_eigenvector=False)
        op.lanczos_eigval0(st)  # warm-up
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for k in range(plain_samples):
            st.x0_seed = k + 1
            op.lanczos_eigval0(st)
        plain.append((time.perf_counter() - t0) / plain_samples)
    n_solves = 2 * R
    t_plain = R * sum(plain)  # 2R solves one after another
    return {
        "workload": "48-spin diag(2,2,3), sectors n_spinon_pairs=0,1 (D=%d,%d), R=%d, M=%d"
                    % (bases[0].shape[0], bases[1].shape[0], R, M),
        "batched_solves_per_s": n_solves / t_batched,
        "batched_device_ms": dev_ms, "batched_device_ms_no_graph": dev_ms_nograph,
        "batched_wall_s": t_batched,
        "plain_solves_per_s": n_solves / t_plain,
        "plain_ms_per_solve": [1e3 * v for v in plain],
        "speedup": t_plain / t_batched,
        "E_of_T_sector0": {"T": temps.tolist(), "E": th["E"].tolist(), "C": th["C"].tolist()},
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--lanczos-steps", type=int, default=200)
    ap.add_argument("--seed", type=int, default=1234)
    ap.add_argument("--calib-copy", action="store_true",
                    help="also run a 64 MiB device copy (read+write) for counter calibration")
    ap.add_argument("--no-solvers", action="store_true", help="skip the Lanczos/eigsh extras")
    ap.add_argument("--variant", default="auto",
                    help="apply kernel: auto, lane, lane2, lane4, lane8, window, hash, hash4")
    ap.add_argument("--bucket-bits", type=int, default=0, help="top-bit table width K (0 = auto)")
    ap.add_argument("--no-ftlm", action="store_true", help="skip the batched FTLM comparison")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 "
                                   "(multi-GPU sharding is design-only)"}))
        sys.exit(2)

    import torch
    import qsi_subspace_ed_amd as q
    from qsi_subspace_ed_amd import _lib
    import ctypes as C

    torch.cuda.set_device(0)
    jd, states, H, g, B, t_enum = build_workload(args.seed)
    D = states.shape[0]
    print(f"[bench] 64-spin Q=0 basis D={D} enumerated in {t_enum:.2f} s", file=sys.stderr,
          flush=True)
    basis = q.GpuBasis(states, 64, bucket_bits=args.bucket_bits,
                       hash_index=args.variant.startswith("hash"))
    op = q.GpuOp(basis, H)
    op.set_variant(args.variant)
    nnz = C.c_int64(0)
    _lib.check(_lib.load().qsi_op_connectivity(op._h, C.byref(nnz), None, None, None, None))
    nnz = nnz.value
    rng = np.random.default_rng(args.seed + 1)
    xh = rng.standard_normal(D)
    xh /= np.linalg.norm(xh)
    x = torch.from_numpy(xh).cuda()
    y = torch.zeros(D, dtype=torch.float64, device="cuda")

    for _ in range(args.warmup):
        op.evaluate(x, y)
    torch.cuda.synchronize()
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    ev0.record()
    for _ in range(args.steps):
        op.evaluate(x, y)
    ev1.record()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    wall = t1 - t0
    kern_s = ev0.elapsed_time(ev1) / 1e3 / args.steps
    print(f"[bench] {args.steps} applies: {wall * 1e3 / args.steps:.3f} ms/step wall, "
          f"{kern_s * 1e6:.1f} us/apply (events)", file=sys.stderr, flush=True)

    S = 8  # bytes per state (<= 64 spins)
    bytes_apply = D * (S + 8 + 8) + 8 * nnz
    achieved = bytes_apply / kern_s / 1e9

    if args.calib_copy:
        a = torch.full((CALIB_BYTES // 8,), 1.0, dtype=torch.float64, device="cuda")
        c = torch.empty_like(a)
        for _ in range(3):
            torch.mul(a, 1.0, out=c)
        torch.cuda.synchronize()
        del a, c

    lanczos = None
    eigsh = None
    if not args.no_solvers and args.lanczos_steps > 0:
        # Lanczos steps (apply + fused dot + axpy/norm + scale) on the device;
        # two lengths so that the start-vector setup cancels out
        def run_lanczos(n):
            st = q.Settings(krylov_dim=15, abs_tol=0.0, rel_tol=0.0, min_iterations=10**9,
                            max_iterations=n, x0_seed=0, calc_eigenvector=False)
            torch.cuda.synchronize()
            t = time.perf_counter()
            op.lanczos_eigval0(st)
            return time.perf_counter() - t
        n1, n2 = max(1, args.lanczos_steps // 4), args.lanczos_steps
        run_lanczos(n1)
        t1 = run_lanczos(n1)
        t2 = run_lanczos(n2)
        lanczos = {"steps": n2, "steps_per_s": (n2 - n1) / max(t2 - t1, 1e-9),
                   "ms_per_step": (t2 - t1) / (n2 - n1) * 1e3}
        # one mfsparse-style solve: 5 lowest, ncv=15, tol 1e-10
        te0 = time.perf_counter()
        evals, _, nconv, napp = op.eigsh(nev=5, ncv=15, tol=1e-10, seed=0,
                                         raise_on_noconv=False)
        te = time.perf_counter() - te0
        eigsh = {"nev": 5, "ncv": 15, "tol": 1e-10, "seconds": te, "nconv": nconv,
                 "applies": napp, "evals": [float(v) for v in evals]}

    ftlm_res = None
    if not args.no_ftlm and not args.no_solvers:
        ftlm_res = bench_ftlm()
        print(f"[bench] FTLM batched {ftlm_res['batched_solves_per_s']:.0f} solves/s vs plain "
              f"{ftlm_res['plain_solves_per_s']:.0f} ({ftlm_res['speedup']:.1f}x)",
              file=sys.stderr, flush=True)

    out = {
        "metric": "H.v basis-states/s (64-spin Q=0 single-vector matrix-free apply)",
        "value": D * args.steps / wall,
        "unit": "basis-states/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded g_s ~ U[0.2,1], |B|=0.1 random direction; basis enumerated)",
        "config": {"workload": "pyro diag(2,2,4) 64 spins, Q=0, single-vector H.v",
                   "n_spins": 64, "basis_dim": int(D), "nnz_offdiag": int(nnz),
                   "g": [float(v) for v in g], "B": [float(v) for v in B]},
        "hv_per_s": args.steps / wall,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": measured_traffic()},
        "apply_kernel_us": kern_s * 1e6,
        "apply_variant": args.variant,
        "bucket_bits": basis.bucket_bits(),
        "algorithmic_bytes_per_apply": int(bytes_apply),
        "lanczos": lanczos,
        "eigsh": eigsh,
        "ftlm_batched": ftlm_res,
    }
    if not args.no_cpu:
        cb, y_cpu = cpu_baseline(states, H, xh)
        out["cpu_baseline"] = cb
        yg = y.cpu().numpy()
        out["parity_rel_err_vs_cpu"] = float(np.linalg.norm(yg - y_cpu) / np.linalg.norm(y_cpu))
        out["speedup_vs_cpu"] = out["value"] / cb["value"]
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
