"""The SCFT force of the field update (module scft_iterate, device kb_mix_force; restated by
the numpy Mixer of the Fortran GPU tests): for two blocks it is the reference's AB force
chiN (phi_B - f_B) + (w_A + w_B)/2 - w_A (SCFT_iterate.f90:145-148); for three, its common
mode is the incompressibility deviation, sum_a t_a = 3 kappa (sum_b phi_b - 1)."""
import os
import sys

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
from test_gpu_fortran import Mixer  # noqa: E402


class _Vol:
    @staticmethod
    def volume_integral(a, box):
        return float(a.sum()) * np.prod(box) / a.size


def _force(w, phi, chi, f):
    m = Mixer(_Vol, (1.0, 1.0, 1.0), n_simple=10, n_hist=3, lam_s=1.0, lam_a=0.1)
    wn, _ = m.update(1, w, phi, chi, f)
    return wn - w  # lam_s = 1: the simple-mixing step is the force


def test_two_blocks_is_the_reference_force():
    rng = np.random.default_rng(1)
    w = rng.standard_normal((2, 4, 4, 4))
    phi = 0.5 + 0.2 * rng.standard_normal((2, 4, 4, 4))
    chiN, f = 17.0, (0.37, 0.63)
    chi = np.array([[0.0, chiN], [chiN, 0.0]])
    t = _force(w, phi, chi, f)
    p = 0.5 * (w[0] + w[1])
    assert np.allclose(t[0], chiN * (phi[1] - f[1]) + p - w[0], atol=1e-13)
    assert np.allclose(t[1], chiN * (phi[0] - f[0]) + p - w[1], atol=1e-13)


def test_three_blocks_common_mode_is_incompressibility():
    rng = np.random.default_rng(2)
    w = rng.standard_normal((3, 4, 4, 4))
    phi = 0.33 + 0.2 * rng.standard_normal((3, 4, 4, 4))
    chi = np.array([[0.0, 12.0, 27.0], [12.0, 0.0, 18.0], [27.0, 18.0, 0.0]])
    f = (0.25, 0.4, 0.35)
    t = _force(w, phi, chi, f)
    kappa = (12.0 + 27.0 + 18.0) / 6
    assert np.allclose(t.sum(0), 3 * kappa * (phi.sum(0) - 1.0), atol=1e-12)
    # incompressible and at the Flory-Huggins fields: no force
    phi0 = np.stack([np.full((4, 4, 4), v) for v in f])
    w0 = np.stack([np.full((4, 4, 4), sum(chi[a][b] * 0.0 for b in range(3))) for a in range(3)])
    assert np.abs(_force(w0, phi0, chi, f)).max() < 1e-13
