"""The Fortran free energy (module scft_energy, used by wormlike_MDE's free_energy) on the
CPU through fe_check: two blocks against the numpy restatement of the reference's AB
expression (SCFT_Density_FE.f90:296-298, :326-333), three blocks against the oracle, and
the homogeneous closed form."""
import os
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FE_CHECK = os.path.join(ROOT, "wormlike-chain-triblock_amd", "fortran", "fe_check")
sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))

pytestmark = pytest.mark.skipif(not os.path.exists(FE_CHECK),
                                reason="Fortran fe_check not built (no flang)")


def fortran_field_energy(tmp_path, w, phi, chi, box):
    nblk, nz, ny, nx = w.shape
    inp, out = tmp_path / "fe_in.bin", tmp_path / "fe_out.bin"
    with open(inp, "wb") as f:
        f.write(np.array([nx, ny, nz, nblk], dtype=np.int32).tobytes())
        f.write(np.asarray(box, dtype=np.float64).tobytes())
        f.write(np.asarray(chi, dtype=np.float64).reshape(9).tobytes())
        f.write(np.ascontiguousarray(w, dtype=np.float64).tobytes())
        f.write(np.ascontiguousarray(phi, dtype=np.float64).tobytes())
    r = subprocess.run([FE_CHECK, str(inp), str(out)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    return float(np.fromfile(out)[0])


def test_two_blocks_match_the_reference_expression(tmp_path):
    import np_restatement as NR
    shape, box = (6, 4, 8), (3.0, 2.0, 4.5)
    rng = np.random.default_rng(31)
    w = 3.0 * rng.standard_normal((2,) + shape)
    phi = 0.5 + 0.3 * rng.standard_normal((2,) + shape)
    chi = np.zeros((3, 3))
    chi[0, 1] = 17.0
    Q = 1.0  # fe_check returns the field part: F + ln Q
    ref = NR.free_energy_ab(w[0], w[1], phi[0], phi[1], Q, 17.0, box)
    assert fortran_field_energy(tmp_path, w, phi, chi, box) == pytest.approx(ref, rel=1e-13, abs=1e-13)


def test_three_blocks_match_the_oracle(tmp_path, oracle):
    shape, box = (4, 6, 8), (2.0, 3.0, 4.0)
    rng = np.random.default_rng(32)
    w = 2.0 * rng.standard_normal((3,) + shape)
    phi = 0.33 + 0.2 * rng.standard_normal((3,) + shape)
    chi = np.triu(rng.uniform(10, 30, (3, 3)), 1)
    orc = oracle.Oracle(shape, box, (4, 5, 3), (1, 1, 1), (1.0, 1.0, 1.0))
    ref = orc.free_energy(w, phi, 1.0, chi)
    orc.close()
    assert fortran_field_energy(tmp_path, w, phi, chi, box) == pytest.approx(ref, rel=1e-13, abs=1e-13)


def test_homogeneous_melt(tmp_path):
    shape, box = (4, 4, 4), (2.0, 2.0, 2.0)
    f = np.array([0.2, 0.5, 0.3])
    chi = np.zeros((3, 3))
    chi[0, 1], chi[0, 2], chi[1, 2] = 12.0, 25.0, 18.0
    phi = np.stack([np.full(shape, v) for v in f])
    wc = np.array([1.5, -0.5, 3.0])
    w = np.stack([np.full(shape, v) for v in wc])
    fh = chi[0, 1] * f[0] * f[1] + chi[0, 2] * f[0] * f[2] + chi[1, 2] * f[1] * f[2]
    assert fortran_field_energy(tmp_path, w, phi, chi, box) == pytest.approx(fh - wc @ f, abs=1e-13)
