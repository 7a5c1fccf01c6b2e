"""HDF5 files in the reference's layouts (basis files, .eigs.h5, .out.h5).

Round trips through the HDF5 C library, and the file structure as h5dump
prints it compared with the structure of the reference's own files
(test/data/*.reference.basis.h5 and REFERENCE_*.eigs.h5, whose h5dump -H
headers are restated below with the sizes substituted)."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from conftest import load_basis
from qsi_subspace_ed_amd import h5io

try:
    h5io._lib()
    HAVE_H5 = True
except RuntimeError:
    HAVE_H5 = False

pytestmark = pytest.mark.skipif(not HAVE_H5, reason="no HDF5 C library in this image")

H5DUMP = shutil.which("h5dump") or ("/opt/conda/bin/h5dump" if os.path.exists("/opt/conda/bin/h5dump")
                                    else None)

STRING = """      DATATYPE  H5T_STRING {
         STRSIZE H5T_VARIABLE;
         STRPAD H5T_STR_NULLTERM;
         CSET H5T_CSET_ASCII;
         CTYPE H5T_C_S1;
      }
      DATASPACE  SCALAR"""


def _header(path):
    env = dict(os.environ)
    env["LD_LIBRARY_PATH"] = "/opt/conda/lib:" + env.get("LD_LIBRARY_PATH", "")
    out = subprocess.run([H5DUMP, "-H", str(path)], check=True, capture_output=True, text=True,
                         env=env).stdout
    return out.split("\n", 1)[1]  # drop the line naming the file


def test_basis_round_trip(tmp_path):
    b = load_basis(32)
    p = tmp_path / "b.h5"
    h5io.write_basis(p, b)
    assert np.array_equal(h5io.read_basis(p), b)
    if H5DUMP:
        want = """GROUP "/" {
   DATASET "basis" {
      DATATYPE  H5T_STD_U64LE
      DATASPACE  SIMPLE { ( 2970, 2 ) / ( 2970, 2 ) }
   }
}
}
"""
        assert _header(p) == want


def test_eigs_round_trip_and_layout(tmp_path):
    rng = np.random.default_rng(1)
    D, k = 87546, 2
    V = rng.standard_normal((D, k))
    ev = np.array([-1.5, -1.25])
    p = tmp_path / "x.eigs.h5"
    h5io.write_eigs(p, ev, V, "/some/dir/pyro_2,0,0_0,2,0_0,0,3.json")
    ev2, V2, lat, dn = h5io.read_eigs(p)
    assert np.array_equal(ev2, ev) and np.array_equal(V2, V)
    assert lat == "pyro_2,0,0_0,2,0_0,0,3.json" and dn == "basis"
    # the eigenvectors dataset holds the column-major buffer under dims (D, k)
    raw = h5io.read_dataset(p, "eigenvectors")
    assert raw.shape == (D, k)
    assert np.array_equal(raw.reshape(-1)[:D], V[:, 0])
    if H5DUMP:
        want = f"""GROUP "/" {{
   DATASET "dset_name" {{
{STRING}
   }}
   DATASET "eigenvalues" {{
      DATATYPE  H5T_IEEE_F64LE
      DATASPACE  SIMPLE {{ ( 2 ) / ( 2 ) }}
   }}
   DATASET "eigenvectors" {{
      DATATYPE  H5T_IEEE_F64LE
      DATASPACE  SIMPLE {{ ( 87546, 2 ) / ( 87546, 2 ) }}
   }}
   DATASET "latfile_json" {{
{STRING}
   }}
}}
}}
"""
        assert _header(p) == want


def test_observables_file(tmp_path):
    rng = np.random.default_rng(2)
    ds = {"ring": rng.standard_normal((48, 1, 1)), "ring_ring": rng.standard_normal((48, 48, 1, 1))}
    for sl in range(4):
        ds[f"partial_vol_sl{sl}"] = rng.standard_normal((24, 1, 1))
    p = tmp_path / "x.eigs.out.h5"
    h5io.write_observables(p, [-0.0129], ds, "pyro_2,0,0_0,2,0_0,0,3.json")
    for name, a in ds.items():
        assert np.array_equal(h5io.read_dataset(p, name), a)
    assert np.array_equal(h5io.read_dataset(p, "eigenvalues"), [-0.0129])
