"""The Fortran driver (reference subroutine interfaces over the C-ABI): it must compile and
link against libfdmamr on the CPU; on the GPU its residual history must equal the oracle's
(it drives the same FP64 path through ISO_C_BINDING)."""
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FDIR = os.path.join(ROOT, "fdm-cell-based-amr-cartesian_amd", "fortran")
EXE = os.path.join(FDIR, "fdmamr_driver")


def build_driver():
    subprocess.run(["make", "-s", "-C", FDIR], check=True, capture_output=True)
    return EXE


def test_driver_builds_and_links():
    exe = build_driver()
    assert os.path.exists(exe)
    out = subprocess.run(["ldd", exe], capture_output=True, text=True).stdout
    assert "libfdmamr.so" in out


@pytest.mark.gpu
def test_driver_residuals_match_oracle(tmp_path):
    exe = build_driver()
    nml = tmp_path / "c.nml"
    nml.write_text("&fdmamr\n max_steps = 6, Pstep_Res = 1\n/\n")
    subprocess.run([exe, str(nml)], cwd=tmp_path, check=True, capture_output=True, timeout=120)
    rows = np.loadtxt(tmp_path / "Output" / "Residual.dat")
    assert len(rows) == 6
    o = O.Oracle()
    for r in rows[:6]:
        assert o.step() == 0
        rmax, rrms = o.residual()
        assert int(r[0]) == o.scalars()["it"]
        assert np.allclose(r[4:8], rmax, rtol=1e-9)
        assert np.allclose(r[8:12], rrms, rtol=1e-9)
    # accuracy_error.dat in the reference's layout (output_ADT.F90:35,65,67-70): a header, then
    # it, nResgrid, L1, P1, L2, P2, Loo, Poo of rho against the exact vortex, then the line of
    # it, nResgrid, L1, L2, Loo and two groups the reference never fills (zeros here)
    lines = (tmp_path / "Output" / "accuracy_error.dat").read_text().splitlines()
    assert len(lines) == 3
    assert lines[0].split() == ["nstep", "Ngrids", "L1_rho", "P1_rho", "L2_rho", "P2_rho", "Loo_rho", "Poo_rho"]
    a = lines[1].split()
    b = lines[2].split()
    L = o.leaves()
    g = O.GAMMA_REF
    rr = (L["x"] - 5) ** 2 + (L["y"] - 5) ** 2
    dT = -(g - 1) * 25 * np.exp(1 - rr) / (8 * g * np.pi ** 2)
    err = L["prim"][:, 0] - (1 + dT) ** (1 / (g - 1))
    area = (0.5 / 2.0 ** L["lvl"]) ** 2
    l1 = np.sum(np.abs(err) * area) / area.sum()
    l2 = np.sqrt(np.sum(err ** 2 * area) / area.sum())
    loo = np.max(np.abs(err))
    assert int(a[0]) == 6 and int(a[1]) == o.n_leaf
    assert np.allclose([float(v) for v in a[2:]], [l1, 0.0, l2, 0.0, loo, 0.0], rtol=2e-6, atol=0)
    assert int(b[0]) == 6 and int(b[1]) == o.n_leaf and len(b) == 13
    assert np.allclose([float(v) for v in b[2:5]], [l1, l2, loo], rtol=2e-6)
    assert [float(v) for v in b[5:]] == [0.0] * 8
