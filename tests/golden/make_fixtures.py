"""Regenerate the committed golden fixtures from the reference's own data files.

Only data is taken: lattice JSON files (inputs) and the HDF5 datasets the
reference's tests compare against (basis files, the `mfeig0` eigen-result of
test/test_correct.sh and its observables file).  HDF5 is read with the
`h5dump` tool (raw little-endian binary dump of one dataset), so no reference
code is imported or run.

Usage (in a checkout that has the reference mounted):
    python tests/golden/make_fixtures.py /root/reference

Outputs (under tests/golden/ unless noted):
    qsi_subspace_ed_amd/lattices/<name>.json
                                    lattice specs (copied byte for byte; the
                                    package ships them as workload inputs)
    basis_<W>.npy / .npz            sorted (D, 2) uint64 [lo, hi] ice bases
    eigs48_jpm0.02_B0.1.npz         E0 + eigenvector (+ .out twin values and the
                                    ring, ring_ring, partial_vol_sl0..3 observables)
    MANIFEST.json                   sha256 of every fixture + provenance
"""
import hashlib
import json
import os
import shutil
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

LATTICES = {
    # name in fixture dir            : path in reference
    "pyro_1,0,0_0,1,0_0,0,2.json": "test/lattice_files/pyro_1,0,0_0,1,0_0,0,2.json",
    "pyro_2,0,0_0,2,0_0,0,2.json": "test/lattice_files/pyro_2,0,0_0,2,0_0,0,2.json",
    "pyro_2,0,0_0,2,0_0,0,3.json": "test/lattice_files/pyro_2,0,0_0,2,0_0,0,3.json",
    "pyro_2,0,0_0,2,0_0,0,4.json": "bench/test_data/pyro_2,0,0_0,2,0_0,0,4.json",
    "pyro_2_2_2x0,4,4b4,0,4b4,4,0b1.json": "test/data/pyro_2_2_2x0,4,4b4,0,4b4,4,0b1.json",
}

BASES = {
    # out name      : (reference h5 file, dataset, rows)
    "basis_8.npy": ("test/data/pyro_1,0,0_0,1,0_0,0,2.reference.basis.h5", "/basis"),
    "basis_32.npy": ("test/data/pyro_2_2_2x0,4,4b4,0,4b4,4,0b1.reference.basis.h5", "/basis"),
    "basis_48.npz": ("test/lattice_files/pyro_2_2_3x0,4,4b4,0,4b4,4,0b1.0test.h5", "/basis"),
}

EIGS = "test/data/REFERENCE_Jpm=0.0200%Bx=0.1000%By=0.1000%Bz=0.1000%.eigs.h5"
EIGS_OUT = "test/data/REFERENCE_Jpm=0.0200%Bx=0.1000%By=0.1000%Bz=0.1000%.eigs.out.h5"


def _h5dump_raw(h5file, dset, dtype):
    h5dump = shutil.which("h5dump") or "/opt/conda/bin/h5dump"
    env = dict(os.environ)
    env["LD_LIBRARY_PATH"] = "/opt/conda/lib:" + env.get("LD_LIBRARY_PATH", "")
    with tempfile.TemporaryDirectory() as td:
        out = os.path.join(td, "raw.bin")
        subprocess.run([h5dump, "-d", dset, "-b", "LE", "-o", out, h5file],
                       check=True, stdout=subprocess.DEVNULL, env=env)
        return np.fromfile(out, dtype=dtype)


def _sha(path):
    h = hashlib.sha256()
    with open(path, "rb") as f:
        h.update(f.read())
    return h.hexdigest()


def main(ref_root):
    latdir = os.path.join(os.path.dirname(os.path.dirname(HERE)), "qsi_subspace_ed_amd", "lattices")
    os.makedirs(latdir, exist_ok=True)
    manifest = {"source": "reference data files (no code run)", "files": {}}
    for name, rel in LATTICES.items():
        dst = os.path.join(latdir, name)
        shutil.copyfile(os.path.join(ref_root, rel), dst)
        manifest["files"]["qsi_subspace_ed_amd/lattices/" + name] = {"from": rel, "sha256": _sha(dst)}

    for name, (rel, dset) in BASES.items():
        raw = _h5dump_raw(os.path.join(ref_root, rel), dset, "<u8").reshape(-1, 2)
        dst = os.path.join(HERE, name)
        if name.endswith(".npz"):
            np.savez_compressed(dst, basis=raw)
        else:
            np.save(dst, raw)
        manifest["files"][name] = {"from": rel + ":" + dset, "rows": int(raw.shape[0]),
                                   "sha256": _sha(dst)}

    e0 = _h5dump_raw(os.path.join(ref_root, EIGS), "/eigenvalues", "<f8")
    vec = _h5dump_raw(os.path.join(ref_root, EIGS), "/eigenvectors", "<f8")
    e0_out = _h5dump_raw(os.path.join(ref_root, EIGS_OUT), "/eigenvalues", "<f8")
    ring = _h5dump_raw(os.path.join(ref_root, EIGS_OUT), "/ring", "<f8")
    ring_ring = _h5dump_raw(os.path.join(ref_root, EIGS_OUT), "/ring_ring", "<f8").reshape(48, 48)
    partial_vol = np.stack([_h5dump_raw(os.path.join(ref_root, EIGS_OUT), f"/partial_vol_sl{sl}", "<f8")
                            for sl in range(4)])
    dst = os.path.join(HERE, "eigs48_jpm0.02_B0.1.npz")
    np.savez_compressed(dst, eigenvalues=e0, eigenvector=vec, eigenvalues_out=e0_out,
                        ring=ring, ring_ring=ring_ring, partial_vol=partial_vol)
    manifest["files"]["eigs48_jpm0.02_B0.1.npz"] = {
        "from": [EIGS, EIGS_OUT],
        "produced_by": "test/test_correct.sh:8 (mfeig0, Jpm=0.02, B=(0.1,0.1,0.1), -k 20, "
                       "--rng_seed 11, --atol -9 --rtol -9)",
        "sha256": _sha(dst)}

    with open(os.path.join(HERE, "MANIFEST.json"), "w") as f:
        json.dump(manifest, f, indent=1, sort_keys=True)
    print(json.dumps(manifest, indent=1))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
