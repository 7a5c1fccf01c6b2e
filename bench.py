"""Benchmark: single-vector matrix-free H·v on the 96-spin pyrochlore ice basis.

Default workload (one GPU): pyrochlore supercell diag(2,4,3) on the fcc
primitive vectors (96 spins), Q=0 sector (all 2-in-2-out states, 2.0e9 of them,
enumerated and sorted on the device), H = -sum_hex g_s (O_hex + O_hex^dagger)
- sum_i (B.z_sl(i)) S^z_i with g_s ~ U[0.2, 1.0] per hexagon sublattice and
|B| = 0.05 uniform on the sphere (seeded), FP64 vectors.  A step is one
y = H x over the whole basis with x, y and the basis resident in HBM.

Prints one JSON line: basis states/s (value), the apply kernel's roofline
figures from HIP events, a parity check of the same apply against the oracle
(the reference's push algorithm) over a bounded sample of source rows, 200
Lanczos steps with Cullum-Willoughby-selected Ritz values and residuals, and
the CPU baseline (the oracle's OpenMP push apply on this host over the same
sample of rows).  The 64-spin diag(2,2,4) apply, its Lanczos/eigensolver
rates, batched FTLM (48 spins, two sectors) and served solves follow under
"spin64" (skip them with --no-extras).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
                       [--workload 96spin|64spin] [--no-extras]
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
PMC_SUMMARY_96 = os.path.join(ROOT, "profiles", "apply_96spin_pmc.json")

# the apply kernel each variant name launches (single vector, no Z strings in
# the off-diagonal terms, no dot product), as rocprofv3 names it
_VARIANT_KERNEL = {"lane": ("lane", 1), "lane2": ("lane", 2), "lane4": ("lane", 4),
                   "lane8": ("lane", 8), "low2": ("low", 2), "low4": ("low", 4),
                   "low8": ("low", 8), "slot2": ("slot", 2), "slot3": ("slot", 3),
                   "slot4": ("slot", 4), "window": ("win", 4), "hash": ("hash", 1),
                   "hash4": ("hash", 4)}


def variant_kernel(variant, nw, cached_pipe=None, low_mode=0):
    """Kernel-name prefix of an apply variant (cached: by gather depth; low
    keys: by how candidates are confirmed, qsi_op_verify_lowkeys' mode)."""
    if variant == "cached":
        return f"k_apply_cached<{nw}, false, {cached_pipe},"
    kind, g = _VARIANT_KERNEL[variant]
    if kind == "low":
        return f"k_apply_low<{nw}, false, false, {g}, {int(low_mode or 0)}>"
    return f"k_apply_{kind}<{nw}, false, false, {g}>"


def measured_traffic(path, kernel):
    """Bytes per apply that cross the L2/fabric interface (to MALL or HBM)
    for the apply kernel whose name starts with `kernel`, from the committed
    rocprofv3 counter summary: read requests counted by size, 32*RDREQ_32B +
    64*RDREQ_64B + 128*RDREQ_128B, plus writes, 64*WRREQ_64B +
    32*(WRREQ - WRREQ_64B); None if absent."""
    try:
        with open(path) as f:
            kernels = json.load(f)["kernels"]
    except (OSError, KeyError, ValueError):
        return None
    for name, rec in kernels.items():
        if name.startswith(kernel):
            return float(rec["traffic_bytes_per_apply"])
    return None


def stream_copy_ceiling(gib=2, reps=10):
    """Measured HBM ceiling for a plain streaming kernel on this device:
    torch's device copy of a `gib` GiB FP64 buffer, read + write bytes per
    second (HIP events), beside the 8 TB/s published peak."""
    import torch
    n = (gib << 30) // 8
    a = torch.ones(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    for _ in range(2):
        b.copy_(a)
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        b.copy_(a)
    e1.record()
    torch.cuda.synchronize()
    sec = e0.elapsed_time(e1) / 1e3 / reps
    gbps = 2 * n * 8 / sec / 1e9
    del a, b
    torch.cuda.empty_cache()
    return {"GBps": gbps, "frac_of_peak": gbps / HBM_PEAK_GBS,
            "method": f"torch copy_ of a {gib} GiB FP64 buffer, read + write bytes, "
                      f"mean of {reps}"}


def cached_kernel_bytes(D, nnz, S):
    """Bytes the cached-connectivity apply itself moves: per row the state
    (diagonal), x_i, two row offsets (one shared) and y_i; per nonzero the
    4-byte column, the 1-byte class/sign and the 8-byte gathered x_j."""
    return D * (S + 24) + 13 * nnz


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


EIGS64_REF = os.path.join(ROOT, "tests", "golden", "eigs64_bench.json")


def cpu_eigs_reference(g, B, D):
    """The committed CPU eigenvalues of this workload (tests/golden/
    make_eigs64.py: the oracle's sparse H and ARPACK), when the file is there
    and was made for the same g, B and basis size; else None."""
    try:
        with open(EIGS64_REF) as f:
            ref = json.load(f)
    except OSError:
        return None
    same = (ref.get("dim") == int(D) and np.array_equal(ref.get("g"), [float(v) for v in g])
            and np.array_equal(ref.get("B"), [float(v) for v in B]))
    return ref if same else None


def host_cpu():
    """The host's CPU (model, sockets, cores, threads, clock), logical core
    count and memory, for the CPU baseline: lscpu when present, else
    /proc/cpuinfo; MemTotal from /proc/meminfo."""
    info = {"cpu_model": "unknown", "nproc": os.cpu_count()}
    import subprocess
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        keys = {"Model name": "cpu_model", "Socket(s)": "sockets",
                "Core(s) per socket": "cores_per_socket", "Thread(s) per core": "threads_per_core",
                "CPU max MHz": "cpu_max_mhz", "L3 cache": "l3_cache"}
        for line in out.splitlines():
            k, _, v = line.partition(":")
            if k.strip() in keys and v.strip():
                info[keys[k.strip()]] = v.strip()
    except (OSError, subprocess.SubprocessError):
        pass
    if info["cpu_model"] == "unknown":
        try:
            with open("/proc/cpuinfo") as f:
                for line in f:
                    if line.startswith("model name"):
                        info["cpu_model"] = line.split(":", 1)[1].strip()
                        break
        except OSError:
            pass
    try:
        with open("/proc/meminfo") as f:
            for line in f:
                if line.startswith("MemTotal"):
                    info["host_mem_gib"] = round(int(line.split()[1]) / 2**20, 1)
                    break
    except (OSError, ValueError):
        pass
    return info


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
                      f"median {t * 1e3:.1f} ms per apply", **host_cpu()}, y


def bench_ftlm(R=64, M=100, seed=77, plain_samples=4):
    """Batched FTLM on the 48-spin diag(2,2,3) lattice: sectors
    n_spinon_pairs = 0 and 1 in one block launch, R random-sign start vectors
    each, M Lanczos steps, FP64, g_s ~ U[0.2, 1], B = 0, with and without the
    operators' connectivity cached.  Compared with the plain path: one
    single-vector Lanczos solve of M steps per sample, one after another
    (timed on plain_samples samples per sector), matrix-free and cached."""
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
    _, _, dev_ms_search = ftlm.lanczos_block(ops, V0s, M, use_graph=True)
    t_batched_search = time.perf_counter() - t0
    # connectivity cached (the FTLM run applies each operator R*M times)
    tc0 = time.perf_counter()
    for op in ops:
        op.cache_connectivity(True)
    t_cache = time.perf_counter() - tc0
    # block configuration chosen by the autotuner on the multi-sector launch
    tuned = [ftlm.autotune(ops, R, reps=3)]
    ftlm.lanczos_block(ops, V0s, 3)
    t0 = time.perf_counter()
    al, be, dev_ms = ftlm.lanczos_block(ops, V0s, M, use_graph=True)
    t_batched = time.perf_counter() - t0
    _, _, dev_ms_nograph = ftlm.lanczos_block(ops, V0s, M, use_graph=False)
    # thermodynamics of sector 0 from the batch (a sanity figure)
    temps = np.array([0.05, 0.1, 0.2, 0.5, 1.0])
    th = ftlm.thermodynamics(ftlm.quadrature(al[0], be[0]), bases[0].shape[0], temps)
    # plain: single-vector Lanczos solves, one at a time, with the cached
    # connectivity (the faster single-vector kernel), then matrix-free
    def time_plain():
        out = []
        for op in ops:
            st = q.Settings(krylov_dim=M, abs_tol=0.0, rel_tol=0.0, min_iterations=10**9,
                            max_iterations=M, x0_seed=0, calc_eigenvector=False)
            op.lanczos_eigval0(st)  # warm-up
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for k in range(plain_samples):
                st.x0_seed = k + 1
                op.lanczos_eigval0(st)
            out.append((time.perf_counter() - t0) / plain_samples)
        return out
    plain_c = time_plain()
    for op in ops:
        op.cache_connectivity(False)
    plain = time_plain()
    n_solves = 2 * R
    t_plain = R * sum(plain)  # 2R solves one after another
    t_plain_c = R * sum(plain_c)
    return {
        "workload": "48-spin diag(2,2,3), sectors n_spinon_pairs=0,1 (D=%d,%d), R=%d, M=%d"
                    % (bases[0].shape[0], bases[1].shape[0], R, M),
        "batched_solves_per_s": n_solves / t_batched,
        "batched_device_ms": dev_ms, "batched_device_ms_no_graph": dev_ms_nograph,
        "batched_wall_s": t_batched,
        "connectivity_cache_build_s": t_cache,
        "autotuned_block_cfg": [{"cfg": c, "ms_per_step": m} for c, m in tuned],
        "batched_search_solves_per_s": n_solves / t_batched_search,
        "batched_search_device_ms": dev_ms_search,
        "plain_solves_per_s": n_solves / t_plain,
        "plain_ms_per_solve": [1e3 * v for v in plain],
        "plain_cached_solves_per_s": n_solves / t_plain_c,
        "plain_cached_ms_per_solve": [1e3 * v for v in plain_c],
        "speedup": t_plain / t_batched,
        "speedup_vs_plain_cached": t_plain_c / t_batched,
        "E_of_T_sector0": {"T": temps.tolist(), "E": th["E"].tolist(), "C": th["C"].tolist()},
    }


def bench_serve(states64, n_req=40, max_cols=16, nev=5, ncv=15, tol=1e-10, seed=99):
    """Served solves: n_req requests at random (g, B) points (g_s ~ U[0, 1]
    per hexagon sublattice with H = -sum g_s (O + O^dagger), |B| ~ U[0, 0.2]
    in a uniformly random direction) spread over six spaces (the 32-, 48- and
    64-spin sectors n_spinon_pairs = 0 and 1), FP64,
    the nev lowest eigenvalues to relative residual tol with ncv Krylov
    vectors (the mfsparse defaults).
      plain   one qsi_eigsh per request, one after another (the operator of
              each point built beforehand, not timed), started from the same
              device-generated Gaussian vector as the served request
      served  SolveServer: all spaces resident, continuous batching of up to
              max_cols requests per space; after one untimed pass of the mix,
              closed loop (all requests queued at t = 0) for throughput, then
              Poisson arrivals at 70% of that rate for p50/p99 latency."""
    import torch
    import qsi_subspace_ed_amd as q
    from qsi_subspace_ed_amd import serve
    names = {32: "pyro_2,0,0_0,2,0_0,0,2.json", 48: "pyro_2,0,0_0,2,0_0,0,3.json",
             64: "pyro_2,0,0_0,2,0_0,0,4.json"}
    jds = {W: q.load_lattice(q.lattice_path(n)) for W, n in names.items()}
    spaces = []
    for W in (32, 48):
        for n in (0, 1):
            spaces.append((f"{W}/n{n}", W, n, q.enumerate_basis(jds[W], n, n_threads=16)))
    spaces.append(("64/n0", 64, 0, states64))
    spaces.append(("64/n1", 64, 1, q.enumerate_basis(jds[64], 1, n_threads=16)))
    rng = np.random.default_rng(seed)
    loc = q.get_loc_z()
    reqs = []
    for k in range(n_req):
        key, W, n, states = spaces[int(rng.integers(len(spaces)))]
        g = list(-rng.uniform(0.0, 1.0, size=4))
        B = rng.standard_normal(3)
        B *= rng.uniform(0.0, 0.2) / np.linalg.norm(B)
        h = [float(B @ loc[s]) for s in range(4)]
        reqs.append((key, W, g, h, 1000 + k))
    by_key = {s[0]: s for s in spaces}
    # plain: one solve at a time
    bases = {s[0]: q.GpuBasis(s[3], s[1]) for s in spaces}
    # the plain path solves an unslotted operator: the point's coefficients
    # are baked into its own term list
    plain_ops = []
    for key, W, g, h, sd in reqs:
        Hs = serve.slotted_hamiltonian(jds[W])
        c = serve.point_coeffs(g, h)
        Hp = q.SymbolicOpSum()
        for (_, o, s) in list(Hs.off_diag_terms) + list(Hs.diagonal_terms):
            Hp.add_term(float(c[s]), o)
        plain_ops.append(q.GpuOp(bases[key], Hp))
    from qsi_subspace_ed_amd.device import randn
    def run_plain():
        plain_ops[0].eigsh(nev=nev, ncv=ncv, tol=tol, seed=1)  # warm-up
        evals, times = [], []
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for op, r in zip(plain_ops, reqs):
            ts = time.perf_counter()
            v0 = randn(op.rows(), r[4])
            ev, _, _, _ = op.eigsh(nev=nev, ncv=ncv, tol=tol, v0=v0, raise_on_noconv=False)
            times.append(time.perf_counter() - ts)
            evals.append(ev)
        return evals, times, time.perf_counter() - t0
    # matrix-free applies (the mfsparse path), then with each operator's
    # connectivity cached (built beforehand, not timed)
    plain_evals, plain_t, t_plain = run_plain()
    for op in plain_ops:
        op.cache_connectivity(True)
    plain_c_evals, plain_c_t, t_plain_c = run_plain()
    del plain_ops
    # served
    free, _ = torch.cuda.mem_get_info()
    srv = serve.SolveServer(int(0.8 * free), max_cols=max_cols, nev=nev, ncv=ncv, tol=tol)
    for key, W, n, states in spaces:
        srv.add_space(key, jds[W], states, W)

    def requests():
        return [serve.Request(key, serve.point_coeffs(g, h), seed=sd, tag=i)
                for i, (key, W, g, h, sd) in enumerate(reqs)]
    # warm-up: one untimed pass of the same mix (uploads, connectivity caches,
    # engine creation and tuning, code load), as a running server would be
    srv.run(requests())
    srv.stats = serve.ServeStats()
    closed = srv.run(requests())
    st_closed = srv.stats.summary()
    rate = 0.7 * st_closed["solves_per_s"]
    arrivals = np.cumsum(rng.exponential(1.0 / rate, size=n_req))
    srv.stats = serve.ServeStats()
    opened = srv.run(requests(), arrivals)
    st_open = srv.stats.summary()
    srv.close()
    errs = []
    for r in closed + opened:
        pe = plain_evals[r.tag]
        errs.append(float(np.max(np.abs(r.result.evals - pe) / np.maximum(1.0, np.abs(pe)))))
        pc = plain_c_evals[r.tag]
        errs.append(float(np.max(np.abs(r.result.evals - pc) / np.maximum(1.0, np.abs(pc)))))
    dims = {s[0]: int(s[3].shape[0]) for s in spaces}
    return {
        "workload": "%d requests, nev=%d ncv=%d tol=%g, spaces %s" % (n_req, nev, ncv, tol, dims),
        "max_cols": max_cols,
        "plain_solves_per_s": n_req / t_plain,
        "plain_p50_ms": float(np.percentile(plain_t, 50) * 1e3),
        "plain_p99_ms": float(np.percentile(plain_t, 99) * 1e3),
        "plain_cached_solves_per_s": n_req / t_plain_c,
        "plain_cached_p50_ms": float(np.percentile(plain_c_t, 50) * 1e3),
        "plain_cached_p99_ms": float(np.percentile(plain_c_t, 99) * 1e3),
        "speedup": st_closed["solves_per_s"] * t_plain_c / n_req,
        "speedup_basis": "plain solves on the same cached connectivity as the engines",
        "served_closed_loop": st_closed,
        "served_poisson": dict(st_open, arrival_rate_per_s=rate),
        "speedup_vs_plain_matrix_free": st_closed["solves_per_s"] * t_plain / n_req,
        "max_rel_eig_diff_vs_plain": max(errs),
        "requests_per_space": {k: sum(1 for r in reqs if r[0] == k) for k in by_key},
    }


def cpu_sample_rate(ob, D, H, x, budget_s=15.0, threads=None):
    """Oracle push apply (the reference's OpenMP algorithm) over a bounded
    sample of source rows of a basis too large to apply whole on the CPU in
    the bench's time: rows [0, n) with n grown (by 4x from 2^20) until one run
    takes about a tenth of the budget, then the median of 3 runs.  Returns
    the rate in rows/s, the thread count, a description of the sample, n,
    and y of one run from zero (the pushes of rows [0, n) only, for the
    parity check)."""
    from oracle import pyoracle as orc
    lib = orc.load()
    if threads:
        lib.oracle_set_threads(int(threads))
    nthr = lib.oracle_num_threads()
    y = np.zeros(D)
    n = 1 << 20
    while True:
        t0 = time.perf_counter()
        orc.evaluate_add_rows(ob, H, x, y, 0, n)
        t = time.perf_counter() - t0
        if t > budget_s / 10 or n >= D:
            break
        n = min(D, n * 4)
    times = []
    for k in range(3):
        if k == 0:
            y[:] = 0.0
        t0 = time.perf_counter()
        orc.evaluate_add_rows(ob, H, x, y, 0, n)
        times.append(time.perf_counter() - t0)
        if k == 0:
            y_one = y.copy()
    t = float(np.median(times))
    return n / t, nthr, (f"rows [0, {n}) of the {D}-state basis (its first {100 * n / D:.2f}%), "
                         f"median of 3 runs, {t:.2f} s each"), n, y_one


def time_variants(op, x, y, names, reps=2):
    """HIP-event time (ms) of one apply with each named kernel variant."""
    import torch
    out = {}
    for v in names:
        op.set_variant(v)
        op.evaluate(x, y)
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            op.evaluate(x, y)
        e1.record()
        torch.cuda.synchronize()
        out[v] = e0.elapsed_time(e1) / reps
    return out


def lanczos_lowest(op, M, nev, seed):
    """M steps of the 3-term recurrence (no reorthogonalisation) from a
    seeded Gaussian vector and the nev lowest Cullum-Willoughby-selected Ritz
    values with their residual estimates."""
    import qsi_subspace_ed_amd as q
    import torch
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    al, be = op.lanczos_coeffs(M, seed=seed)
    t_lz = time.perf_counter() - t0
    vals, res, mult = q.lanczos_ritz(al, be)
    T = np.diag(al) + np.diag(be[:-1], 1) + np.diag(be[:-1], -1)
    plain = np.linalg.eigvalsh(T)
    return {"steps": int(len(al)), "seconds": t_lz, "ms_per_step": t_lz / len(al) * 1e3,
            "lowest": [float(v) for v in vals[:nev]],
            "residuals": [float(r) for r in res[:nev]],
            "multiplicity_in_T": [int(m) for m in mult[:nev]],
            "selection": "Cullum-Willoughby (copies merged, spurious values dropped)",
            "plain_lowest_ritz_values": [float(v) for v in plain[:nev]]}


def bench_96(args):
    """The HBM-scale single-vector apply: 96-spin diag(2,4,3), Q=0, the basis
    enumerated and sorted on the device, g_s ~ U[0.2, 1] and |B| = 0.05 in a
    random direction (seeded), FP64; the kernel picked by the autotuner;
    parity against the oracle over a sample of source rows; 200 Lanczos
    steps for the 5 lowest eigenvalues."""
    import torch
    import qsi_subspace_ed_amd as q
    from qsi_subspace_ed_amd import _lib
    import ctypes as C
    torch.cuda.set_device(0)
    jd = q.load_lattice(q.lattice_path("pyro_2,0,0_0,4,0_0,0,3.json"))
    t0 = time.perf_counter()
    basis = q.enumerate_basis_device(jd, 0, bucket_bits=args.bucket_bits)
    t_enum = time.perf_counter() - t0
    D = basis.dim()
    print(f"[bench] 96-spin Q=0 basis D={D} enumerated on the device in {t_enum:.1f} s",
          file=sys.stderr, flush=True)
    rng = np.random.default_rng(args.seed)
    g = rng.uniform(0.2, 1.0, size=4)
    b = rng.normal(size=3)
    B = 0.05 * b / np.linalg.norm(b)
    H = q.build_hamiltonian_g(jd, [-v for v in g])
    q.add_local_field(H, jd, B)
    if not args.no_slots:
        ts = time.perf_counter()
        basis.build_slots()
        print(f"[bench] slot index built in {time.perf_counter() - ts:.1f} s", file=sys.stderr,
              flush=True)
    if not args.no_lowkeys:
        ts = time.perf_counter()
        basis.build_lowkeys()
        print(f"[bench] low keys built in {time.perf_counter() - ts:.1f} s", file=sys.stderr,
              flush=True)
    op = q.GpuOp(basis, H)
    lowkeys_mode = None
    if basis.has_lowkeys():
        ts = time.perf_counter()
        lowkeys_mode = op.verify_lowkeys()
        print(f"[bench] low-key check mode for the operator: {lowkeys_mode} "
              f"({time.perf_counter() - ts:.1f} s)", file=sys.stderr, flush=True)
    nnz = C.c_int64(0)
    _lib.check(_lib.load().qsi_op_connectivity(op._h, C.byref(nnz), None, None, None, None))
    nnz = nnz.value
    x = torch.randn(D, dtype=torch.float64, device="cuda", generator=torch.Generator(
        device="cuda").manual_seed(args.seed))
    x /= torch.linalg.norm(x)
    y = torch.zeros_like(x)
    variants_ms = time_variants(op, x, y, ["lane4"] + ([] if args.no_slots else
                                                       ["slot2", "slot4", "slot3"])
                                + ([] if args.no_lowkeys else ["low2", "low4", "low8"]))
    tuned = op.autotune(0, reps=1, x=x, y=y) if args.variant == "auto" else None
    if tuned is None:
        op.set_variant(args.variant)
    print(f"[bench] autotuned variant {tuned}; per variant (ms) {variants_ms}", file=sys.stderr,
          flush=True)
    steps = max(1, min(args.steps, 50))
    for _ in range(max(1, min(args.warmup, 3))):
        op.evaluate(x, y)
    torch.cuda.synchronize()
    # HIP events on the stream the library launches on (torch's current one)
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record()
    for _ in range(steps):
        op.evaluate(x, y)
    e1.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kern_s = e0.elapsed_time(e1) / 1e3 / steps
    print(f"[bench] {steps} applies: {wall * 1e3 / steps:.1f} ms/step wall, "
          f"{kern_s * 1e3:.1f} ms/apply (events)", file=sys.stderr, flush=True)
    S = 16
    bytes_apply = D * (S + 8 + 8) + 8 * nnz
    achieved = bytes_apply / kern_s / 1e9
    out = {
        "metric": "H.v basis-states/s (96-spin Q=0 single-vector matrix-free apply)",
        "value": D * steps / wall,
        "unit": "basis-states/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded g_s ~ U[0.2,1], |B|=0.05 random direction; basis enumerated "
                "on the device)",
        "config": {"workload": "pyro diag(2,4,3) 96 spins, Q=0, single-vector H.v",
                   "n_spins": 96, "basis_dim": int(D), "nnz_offdiag": int(nnz),
                   "g": [float(v) for v in g], "B": [float(v) for v in B]},
        "hv_per_s": steps / wall,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": measured_traffic(PMC_SUMMARY_96, variant_kernel(
                         tuned[0] if tuned else args.variant, 2,
                         low_mode=lowkeys_mode))},
        "apply_kernel_ms": kern_s * 1e3,
        "apply_variant": tuned[0] if tuned else args.variant,
        "autotuned": tuned is not None,
        "variants_ms": variants_ms,
        "bucket_bits": basis.bucket_bits(),
        "slot_index": basis.has_slots(),
        "low_keys": basis.has_lowkeys(),
        "low_keys_mode": lowkeys_mode,
        "enumeration_s": t_enum,
        "algorithmic_bytes_per_apply": int(bytes_apply),
    }
    # parity and CPU baseline on a bounded sample of source rows [0, n): the
    # oracle pushes those rows; the device pulls with x zero outside them
    if not args.no_cpu:
        from oracle import pyoracle as orc
        states = basis.states()
        tb = time.perf_counter()
        ob = orc.Basis(states)
        t_ob = time.perf_counter() - tb
        basis._states_host = None
        del states
        xh = x.cpu().numpy()
        rate, nthr, sample, n_rows, y_cpu = cpu_sample_rate(ob, D, H, xh)
        del ob
        xm = x.clone()
        xm[n_rows:] = 0.0
        ym = torch.zeros_like(x)
        op.evaluate(xm, ym)
        yc = torch.from_numpy(y_cpu).cuda()
        del y_cpu
        err = float(torch.linalg.norm(ym - yc) / torch.linalg.norm(yc))
        out["parity"] = {"rows": int(n_rows), "rel_err_vs_oracle": err,
                         "max_abs_diff": float((ym - yc).abs().max()),
                         "method": "oracle push of source rows [0, n) vs device apply of x "
                                   "zeroed outside them, whole output vector"}
        del xm, ym, yc
        print(f"[bench] parity over {n_rows} source rows: rel err {err:.2e}", file=sys.stderr,
              flush=True)
        out["cpu_baseline"] = {"value": rate, "unit": "basis-states/s", "cores": int(nthr),
                               "kind": "port", "sample": sample + f"; oracle basis built in "
                               f"{t_ob:.0f} s", **host_cpu()}
        out["speedup_vs_cpu"] = out["value"] / rate
        torch.cuda.empty_cache()
    # the same apply through the cached connectivity, when it fits in HBM
    # (the slot index and low keys are dropped first: the cached kernel does
    # not search)
    cached = None
    if not args.no_cached:
        basis.build_slots(False)
        basis.build_lowkeys(False)
        torch.cuda.empty_cache()
    free, _ = torch.cuda.mem_get_info()
    need = 5 * nnz + 8 * (D + 1)
    if not args.no_cached and free > 1.05 * need + 2 * 8 * D:
        torch.cuda.synchronize()
        tc0 = time.perf_counter()
        op.cache_connectivity(True)
        t_build = time.perf_counter() - tc0
        yc = torch.zeros_like(y)
        # the cached kernel's gather depth (and whether it wins) by the autotuner
        tuned_c = op.autotune(0, reps=1, x=x, y=yc)
        op.evaluate(x, yc)
        torch.cuda.synchronize()
        ec0 = torch.cuda.Event(enable_timing=True)
        ec1 = torch.cuda.Event(enable_timing=True)
        tw0 = time.perf_counter()
        ec0.record()
        for _ in range(steps):
            op.evaluate(x, yc)
        ec1.record()
        torch.cuda.synchronize()
        wall_c = time.perf_counter() - tw0
        kc = ec0.elapsed_time(ec1) / 1e3 / steps
        cached = {"states_per_s": D * steps / wall_c, "apply_kernel_ms": kc * 1e3,
                  "build_s": t_build, "device_bytes": int(need), "autotuned": tuned_c[0],
                  "cached_pipe": op.cached_pipe(),
                  "algorithmic_GBps": bytes_apply / kc / 1e9,
                  "kernel_bytes_per_apply": cached_kernel_bytes(D, nnz, S),
                  "kernel_GBps": cached_kernel_bytes(D, nnz, S) / kc / 1e9,
                  "traffic": measured_traffic(PMC_SUMMARY_96,
                                              variant_kernel("cached", 2, op.cached_pipe())),
                  "bit_identical_to_search": bool(torch.equal(y, yc))}
        print(f"[bench] cached connectivity: {kc * 1e3:.1f} ms/apply (built in {t_build:.1f} s)",
              file=sys.stderr, flush=True)
        del yc
    elif not args.no_cached:
        cached = {"skipped": f"needs {need / 1e9:.0f} GB, {free / 1e9:.0f} GB free"}
    out["cached_connectivity"] = cached
    del x, y
    torch.cuda.empty_cache()  # the Lanczos vectors are allocated by the library
    if args.lanczos_steps > 0:
        lz = lanczos_lowest(op, args.lanczos_steps, 5, args.seed)
        lz["apply"] = "cached" if cached and "skipped" not in cached else "search"
        out["lanczos"] = lz
        print(f"[bench] {lz['steps']} Lanczos steps in {lz['seconds']:.1f} s, lowest "
              f"{lz['lowest']}", file=sys.stderr, flush=True)
    op.close()
    basis.close()
    torch.cuda.empty_cache()
    out["stream_copy_ceiling"] = stream_copy_ceiling()
    return out


def bench_64(args):
    """The 64-spin diag(2,2,4) Q=0 apply (search and cached connectivity),
    Lanczos and eigensolver rates, the CPU baseline on full applies, batched
    FTLM (48 spins) and served solves; returns the JSON record.  An apply
    takes ~0.35 ms here, so at least 200 are timed after 10 warm-ups."""
    steps = max(args.steps, 200)
    warmup = max(args.warmup, 10)
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
    if not args.no_slots:
        basis.build_slots()
    if not args.no_lowkeys:
        basis.build_lowkeys()
    op = q.GpuOp(basis, H)
    lowkeys_mode = op.verify_lowkeys() if basis.has_lowkeys() else None
    tuned = None
    if args.variant == "auto":
        tuned = op.autotune(0, reps=3)  # fastest searching kernel on this workload
    else:
        op.set_variant(args.variant)
    nnz = C.c_int64(0)
    _lib.check(_lib.load().qsi_op_connectivity(op._h, C.byref(nnz), None, None, None, None))
    nnz = nnz.value
    rng = np.random.default_rng(args.seed + 1)
    xh = rng.standard_normal(D)
    xh /= np.linalg.norm(xh)
    x = torch.from_numpy(xh).cuda()
    y = torch.zeros(D, dtype=torch.float64, device="cuda")

    for _ in range(warmup):
        op.evaluate(x, y)
    torch.cuda.synchronize()
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    ev0.record()
    for _ in range(steps):
        op.evaluate(x, y)
    ev1.record()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    wall = t1 - t0
    kern_s = ev0.elapsed_time(ev1) / 1e3 / steps
    print(f"[bench] {steps} applies: {wall * 1e3 / steps:.3f} ms/step wall, "
          f"{kern_s * 1e6:.1f} us/apply (events)", file=sys.stderr, flush=True)

    S = 8  # bytes per state (<= 64 spins)
    bytes_apply = D * (S + 8 + 8) + 8 * nnz
    achieved = bytes_apply / kern_s / 1e9

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
        ref = cpu_eigs_reference(g, B, D)
        if ref is not None:
            w = np.asarray(ref["evals"][:5])
            eigsh["cpu_reference"] = {
                "method": ref["method"], "evals": [float(v) for v in w],
                "max_rel_diff": float(np.max(np.abs(np.asarray(evals) - w) / np.abs(w)))}

    # the same workload with the connectivity cached on the device
    cached = None
    if not args.no_cached:
        torch.cuda.synchronize()
        tc0 = time.perf_counter()
        cnnz = op.cache_connectivity(True)
        torch.cuda.synchronize()
        t_build = time.perf_counter() - tc0
        tuned_c = op.autotune(0, reps=3)  # the cached kernel's gather depth
        if tuned_c[0] != "cached":
            op.set_variant("cached")
        yc = torch.zeros_like(y)
        for _ in range(warmup):
            op.evaluate(x, yc)
        torch.cuda.synchronize()
        ec0 = torch.cuda.Event(enable_timing=True)
        ec1 = torch.cuda.Event(enable_timing=True)
        tw0 = time.perf_counter()
        ec0.record()
        for _ in range(steps):
            op.evaluate(x, yc)
        ec1.record()
        torch.cuda.synchronize()
        wall_c = time.perf_counter() - tw0
        kc = ec0.elapsed_time(ec1) / 1e3 / steps
        cached = {"states_per_s": D * steps / wall_c, "apply_kernel_us": kc * 1e6,
                  "build_s": t_build, "nnz": int(cnnz), "autotuned": tuned_c[0],
                  "cached_pipe": op.cached_pipe(),
                  "device_bytes": int(cnnz) * 5 + 8 * (D + 1),
                  "algorithmic_GBps": bytes_apply / kc / 1e9,
                  "kernel_bytes_per_apply": cached_kernel_bytes(D, cnnz, S),
                  "kernel_GBps": cached_kernel_bytes(D, cnnz, S) / kc / 1e9,
                  "traffic": measured_traffic(PMC_SUMMARY,
                                              variant_kernel("cached", 1, op.cached_pipe())),
                  "bit_identical_to_search": bool(torch.equal(y, yc))}
        if not args.no_solvers and args.lanczos_steps > 0:
            n1, n2 = max(1, args.lanczos_steps // 4), args.lanczos_steps
            run_lanczos(n1)
            t1c = run_lanczos(n1)
            t2c = run_lanczos(n2)
            cached["lanczos_ms_per_step"] = (t2c - t1c) / (n2 - n1) * 1e3
            te0 = time.perf_counter()
            evc, _, nconv_c, napp_c = op.eigsh(nev=5, ncv=15, tol=1e-10, seed=0,
                                               raise_on_noconv=False)
            cached["eigsh"] = {"seconds": time.perf_counter() - te0, "applies": napp_c,
                               "nconv": nconv_c, "evals": [float(v) for v in evc]}
        print(f"[bench] cached connectivity: {kc * 1e6:.1f} us/apply "
              f"(built in {t_build * 1e3:.1f} ms, {cnnz} nonzeros)", file=sys.stderr, flush=True)
        op.cache_connectivity(False)

    ftlm_res = None
    if not args.no_ftlm and not args.no_solvers:
        ftlm_res = bench_ftlm()
        print(f"[bench] FTLM batched {ftlm_res['batched_solves_per_s']:.0f} solves/s vs plain "
              f"{ftlm_res['plain_solves_per_s']:.0f} ({ftlm_res['speedup']:.1f}x)",
              file=sys.stderr, flush=True)

    serve_res = None
    if not args.no_serve and not args.no_solvers:
        serve_res = bench_serve(states)
        print(f"[bench] served {serve_res['served_closed_loop']['solves_per_s']:.1f} solves/s vs "
              f"plain cached {serve_res['plain_cached_solves_per_s']:.1f} "
              f"({serve_res['speedup']:.1f}x), "
              f"Poisson p50/p99 {serve_res['served_poisson']['p50_ms']:.0f}/"
              f"{serve_res['served_poisson']['p99_ms']:.0f} ms", file=sys.stderr, flush=True)

    out = {
        "metric": "H.v basis-states/s (64-spin Q=0 single-vector matrix-free apply)",
        "value": D * steps / wall,
        "unit": "basis-states/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warmup,
        "ms_per_step": wall * 1e3 / steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded g_s ~ U[0.2,1], |B|=0.1 random direction; basis enumerated)",
        "config": {"workload": "pyro diag(2,2,4) 64 spins, Q=0, single-vector H.v",
                   "n_spins": 64, "basis_dim": int(D), "nnz_offdiag": int(nnz),
                   "g": [float(v) for v in g], "B": [float(v) for v in B]},
        "hv_per_s": steps / wall,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": measured_traffic(PMC_SUMMARY, variant_kernel(
                         tuned[0] if tuned else args.variant, 1,
                         low_mode=lowkeys_mode))},
        "apply_kernel_us": kern_s * 1e6,
        "apply_variant": tuned[0] if tuned else args.variant,
        "low_keys_mode": lowkeys_mode,
        "autotuned": tuned is not None,
        "bucket_bits": basis.bucket_bits(),
        "algorithmic_bytes_per_apply": int(bytes_apply),
        "lanczos": lanczos,
        "eigsh": eigsh,
        "cached_connectivity": cached,
        "ftlm_batched": ftlm_res,
        "served": serve_res,
        "stream_copy_ceiling": stream_copy_ceiling(),
    }
    if not args.no_cpu:
        cb, y_cpu = cpu_baseline(states, H, xh)
        out["cpu_baseline"] = cb
        yg = y.cpu().numpy()
        out["parity_rel_err_vs_cpu"] = float(np.linalg.norm(yg - y_cpu) / np.linalg.norm(y_cpu))
        out["speedup_vs_cpu"] = out["value"] / cb["value"]
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--lanczos-steps", type=int, default=200)
    ap.add_argument("--seed", type=int, default=1234)
    ap.add_argument("--no-solvers", action="store_true", help="skip the Lanczos/eigsh extras")
    ap.add_argument("--variant", default="auto",
                    help="apply kernel: auto, lane, lane2, lane4, lane8, low2, low4, low8, "
                         "slot2, slot3, slot4, window, hash, hash4")
    ap.add_argument("--bucket-bits", type=int, default=0, help="bucket key width K in bits (0 = automatic)")
    ap.add_argument("--no-ftlm", action="store_true", help="skip the batched FTLM comparison")
    ap.add_argument("--no-serve", action="store_true", help="skip the served-solves comparison")
    ap.add_argument("--no-cached", action="store_true",
                    help="skip the cached-connectivity apply figures")
    ap.add_argument("--workload", default="96spin", choices=["64spin", "96spin"],
                    help="96spin: the default line (the HBM-scale apply); 64spin: the 64-spin "
                         "apply and its extras alone")
    ap.add_argument("--no-extras", action="store_true",
                    help="96spin: skip the 64-spin, FTLM and served sections")
    ap.add_argument("--no-lowkeys", action="store_true",
                    help="do not build the low keys (no QSI_APPLY_LOW* variants)")
    ap.add_argument("--no-slots", action="store_true",
                    help="do not build the slot index (search by bucket bisection only)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "the benchmark runs on one GPU; --gpus must be 1 "
                                   "(multi-GPU sharding is design-only)"}))
        sys.exit(2)

    if args.workload == "96spin":
        out = bench_96(args)
        if not args.no_extras:
            out["spin64"] = bench_64(args)
    else:
        out = bench_64(args)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
