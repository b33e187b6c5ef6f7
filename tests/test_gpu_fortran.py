"""The Fortran binding (module wormlike_MDE: MDE_q / MDE_qstar / density / free_energy)
driven by the simple-mixing SCFT program, against the same loop on the CPU oracle."""
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "wormlike-chain-triblock_amd", "fortran", "scft_simple_mixing")


def write_input(path, shape, box, Nblk, L, lam, chi, step, niter, w):
    nz, ny, nx = shape
    nb = len(Nblk)
    hdr = np.zeros(13, dtype=np.int32)
    hdr[:5] = [nx, ny, nz, nb, sum(Nblk)]
    hdr[5:5 + nb] = Nblk
    hdr[8:8 + nb] = L
    hdr[11] = niter
    lam3 = np.zeros(3)
    lam3[:nb] = lam
    with open(path, "wb") as f:
        f.write(hdr.tobytes())
        f.write(np.asarray(box, dtype=np.float64).tobytes())
        f.write(lam3.tobytes())
        f.write(np.asarray(chi, dtype=np.float64).reshape(9).tobytes())
        f.write(np.float64(step).tobytes())
        # Fortran w_blk(0:N-1, 1:nblk): column-major == [nblk][N] C order
        f.write(np.ascontiguousarray(w, dtype=np.float64).tobytes())


def oracle_loop(oracle, shape, box, Nblk, L, lam, chi, step, niter, w):
    orc = oracle.Oracle(shape, box, Nblk, L, lam)
    f = np.array(Nblk, dtype=float) / sum(Nblk)
    nb = len(Nblk)
    w = w.copy()
    rows = []
    for it in range(niter):
        phi, Q, _ = orc.density(orc.forward(w), orc.backward(w))
        F = orc.free_energy(w, phi, Q, chi)
        xi = w.mean(0)
        t = np.zeros_like(w)
        for a in range(nb):
            for b in range(nb):
                if b != a:
                    t[a] += chi[a][b] * (phi[b] - f[b])
            t[a] += xi - w[a]
        rows.append([it + 1, F, np.log(Q), np.abs(t).max()] + [phi[a].mean() for a in range(nb)])
        w = w + step * t
    return np.array(rows), w


@pytest.mark.skipif(not os.path.exists(EXE), reason="Fortran binding not built (no flang)")
def test_fortran_drop_in_scft_loop(wlc, oracle, tmp_path):
    shape, box = (16, 16, 16), (3.0, 3.0, 3.0)
    Nblk, L, lam = (4, 5, 3), (3, 2, 3), (1.0, 0.5, 2.0)
    chi = np.array([[0.0, 20.0, 15.0], [20.0, 0.0, 25.0], [15.0, 25.0, 0.0]])
    step, niter = 0.05, 3
    rng = np.random.default_rng(4)
    w = 0.5 * rng.standard_normal((3,) + shape)
    inp, out, wout = tmp_path / "in.bin", tmp_path / "out.txt", tmp_path / "w.bin"
    write_input(inp, shape, box, Nblk, L, lam, chi, step, niter, w)
    r = subprocess.run([EXE, str(inp), str(out), str(wout)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    got = np.loadtxt(out).reshape(niter, -1)
    ref, wref = oracle_loop(oracle, shape, box, Nblk, L, lam, chi, step, niter, w)
    assert np.array_equal(got[:, 0], ref[:, 0])
    assert np.abs(got[:, 1] - ref[:, 1]).max() < 1e-9                  # F, kT per chain
    assert np.abs(got[:, 2] - ref[:, 2]).max() < 1e-10                 # ln Q
    assert np.abs(got[:, 3:] - ref[:, 3:]).max() < 1e-9 * max(1.0, np.abs(ref[:, 3:]).max())
    wfin = np.fromfile(wout, dtype=np.float64).reshape(w.shape)
    assert np.abs(wfin - wref).max() < 1e-9 * np.abs(wref).max()
