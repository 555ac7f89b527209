"""The Fortran drop-in boundary: a Fortran program (nexd-3d_amd/fortran/
nexd_drive.f90) calls the library through the iso_c_binding module with the
reference's column-major, 1-based arrays; its seismograms must match the
oracle to <= 1e-10."""
import os
import subprocess

import numpy as np
import pytest

from nexd_amd import fdump, layout
from oracle import oracle
from tests.cases import CASES

pytestmark = pytest.mark.gpu

DRIVER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "nexd-3d_amd", "fortran",
                      "nexd_drive")


@pytest.mark.parametrize("name", ["cube3_n3_moment", "layered4_n4_att_pml_rk3"])
def test_fortran_caller_matches_oracle(name, tmp_path):
    if not os.access(DRIVER, os.X_OK):
        pytest.fail("nexd_drive is not built; run __graft_entry__.build()")
    builder, kw = CASES[name]
    p = builder(**kw)
    arrs = layout.reference_arrays(p)
    fin, fout = str(tmp_path / "ops.bin"), str(tmp_path / "seis.bin")
    fdump.write(fin, arrs)
    print(f"[gpu fortran] {name} nelem={p.nelem} nt={p.nt}", flush=True)
    r = subprocess.run([DRIVER, fin, fout], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    s = fdump.read_seismograms(fout, arrs["nrec"], p.nt)
    ref, _, _ = oracle.run(arrs, compat=3)
    err = np.max(np.linalg.norm(s - ref, axis=-1) / np.maximum(np.linalg.norm(ref, axis=-1), 1e-300))
    assert err <= 1e-10
