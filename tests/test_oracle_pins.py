"""Pins of the CPU oracle (oracle/wlc_oracle.c) -- no GPU needed.

The reference ships no tests, fixtures or golden vectors for the MDE path, and it
cannot be built here (MPI Fortran over FFTW/2DECOMP/MKL), so every function of the
oracle is pinned by one of: numpy/scipy library routines, the independent numpy
restatement in tests/golden (committed fixtures), closed forms, or invariants.
"""
import os
import sys

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


# --- FFT (2dcomp_fft_mod.f90:39-73, wormlike_MDE_2dcomp_mod.f90:357-455) ----------------

@pytest.mark.parametrize("n", [1, 2, 3, 4, 6, 8, 12, 16, 24, 32, 48, 64, 96, 128, 5, 10])
def test_fft1_matches_numpy(oracle, n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    f = oracle.fft1(x, -1)
    assert np.allclose(f, np.fft.fft(x), rtol=0, atol=1e-13 * max(1, np.abs(f).max()))
    b = oracle.fft1(x, +1)
    assert np.allclose(b, np.fft.ifft(x) * n, rtol=0, atol=1e-13 * max(1, np.abs(b).max()))


@pytest.mark.parametrize("shape", [(4, 6, 8), (6, 12, 6), (8, 8, 8), (24, 6, 4), (2, 2, 2)])
def test_rfft3_roundtrip_and_numpy(oracle, shape):
    rng = np.random.default_rng(1)
    a = rng.standard_normal(shape)
    s = oracle.rfft3(a)
    ref = np.fft.rfftn(a)
    assert np.abs(s - ref).max() < 1e-12 * np.abs(ref).max()
    back = oracle.irfft3(s, shape[2]) / a.size
    assert np.abs(back - a).max() < 1e-13


def test_irfft3_drops_imag_of_self_conjugate_x_planes(oracle):
    """A non-Hermitian spectrum (what the G^-1 apply produces at kx = 0 and kx = nx/2)
    follows numpy.irfftn: c2c inverse over z and y, then c2r along x."""
    rng = np.random.default_rng(2)
    shape = (4, 6, 8)
    spec = rng.standard_normal((4, 6, 5)) + 1j * rng.standard_normal((4, 6, 5))
    got = oracle.irfft3(spec, 8)
    ref = np.fft.irfftn(spec, s=shape, axes=(0, 1, 2)) * np.prod(shape)
    assert np.abs(got - ref).max() < 1e-12 * np.abs(ref).max()


# --- SH algebra (SPH_mod.f90:11-233, init_array_2dcomp_mod.f90:166-182) -----------------

def test_three_j_known_values(oracle):
    assert oracle.three_j(1, 1, 0, 0, 0, 0) == pytest.approx(-1 / np.sqrt(3), abs=1e-15)
    assert oracle.three_j(2, 1, 1, 0, 0, 0) == pytest.approx(np.sqrt(2 / 15), abs=1e-15)
    assert oracle.three_j(1, 1, 1, 1, -1, 0) == pytest.approx(1 / np.sqrt(6), abs=1e-15)
    assert oracle.three_j(2, 2, 2, 0, 0, 0) == pytest.approx(-np.sqrt(2 / 35), abs=1e-15)
    assert oracle.three_j(1, 1, 3, 0, 0, 0) == 0.0            # triangle rule
    assert oracle.three_j(2, 1, 1, 1, 1, 0) == 0.0            # m sum
    # orthogonality: sum_{m1,m2} (2 l3 + 1) 3j^2 = 1
    for l1, l2, l3 in [(2, 3, 4), (4, 4, 2), (6, 5, 1)]:
        for m3 in range(-l3, l3 + 1):
            s = sum(oracle.three_j(l1, l2, l3, m1, -m1 - m3, m3) ** 2
                    for m1 in range(-l1, l1 + 1) if abs(m1 + m3) <= l2)
            assert (2 * l3 + 1) * s == pytest.approx(1.0, abs=1e-13)


def test_rmatrices_against_quadrature_fixture(oracle):
    """The restated triple product reproduces <Y_i|u_a|Y_j> of real harmonics that
    keep the Condon-Shortley phase, by Gauss-Legendre quadrature (fixture)."""
    R = np.load(os.path.join(GOLDEN, "rmatrices_L8.npz"))["R"]
    Ro = oracle.rmatrices(8)
    assert np.abs(R - Ro).max() < 1e-13
    assert Ro[2][0, 2] == pytest.approx(1 / np.sqrt(3), abs=1e-15)     # R_z[(0,0),(1,0)]
    assert Ro[0][0, 3] == pytest.approx(-1 / np.sqrt(3), abs=1e-15)    # R_x[(0,0),(1,1)]


@pytest.mark.parametrize("L", [1, 3, 6, 8])
def test_rmatrices_invariants(oracle, L):
    R = oracle.rmatrices(L)
    M = (L + 1) ** 2
    ls = np.array([l for l in range(L + 1) for m in range(-l, l + 1)])
    for a in range(3):
        assert np.abs(R[a] - R[a].T).max() < 1e-15
        nz = np.argwhere(np.abs(R[a]) > 1e-14)
        assert np.all(np.abs(ls[nz[:, 0]] - ls[nz[:, 1]]) == 1)
    # |u|^2 = 1 on the harmonics whose products stay inside the truncation
    Mi = L * L
    S = sum(R[a] @ R[a] for a in range(3))
    assert np.abs(S[:Mi, :Mi] - np.eye(Mi)).max() < 1e-14
    # multiplication operators commute away from the truncation edge
    Mc = L * L
    for a in range(3):
        for b in range(3):
            C = R[a] @ R[b] - R[b] @ R[a]
            assert np.abs(C[:Mc, :Mc]).max() < 1e-14
    assert M == R.shape[1]


def test_fp32_literal_mode_is_the_documented_perturbation(oracle):
    Ri = oracle.rmatrices(4, literal=0)
    Rl = oracle.rmatrices(4, literal=1)
    d = np.abs(Ri - Rl).max()
    assert 1e-10 < d < 1e-6


# --- G(k) and its inverse (Gmatrix_2dcomp_mod.f90:384-500) -------------------------------

def _G(oracle, L, cd, sigma, ds, lam, k):
    R = oracle.rmatrices(L)
    ls = np.array([l for l in range(L + 1) for m in range(-l, l + 1)])
    return np.diag(cd + ds * lam * ls * (ls + 1)).astype(complex) + 1j * sigma * ds * (
        k[0] * R[0] + k[1] * R[1] + k[2] * R[2])


@pytest.mark.parametrize("L", [0, 2, 4, 6])
def test_ginv_matches_numpy_inverse(oracle, L):
    rng = np.random.default_rng(L)
    for _ in range(4):
        k = rng.standard_normal(3) * 40
        for cd in (1.0, 1.5, 11 / 6):
            for sigma in (1.0, -1.0):
                Gi = oracle.ginv(L, cd, sigma, 0.005, 0.8, k)
                ref = np.linalg.inv(_G(oracle, L, cd, sigma, 0.005, 0.8, k))
                assert np.abs(Gi - ref).max() < 1e-13 * np.abs(ref).max()
                assert np.abs(Gi - Gi.T).max() < 1e-13          # complex symmetric


def test_ginv_closed_forms(oracle):
    L, ds, lam = 5, 0.01, 1.3
    ls = np.array([l for l in range(L + 1) for m in range(-l, l + 1)])
    Gi = oracle.ginv(L, 11 / 6, 1.0, ds, lam, (0.0, 0.0, 0.0))
    assert np.allclose(Gi, np.diag(1.0 / (11 / 6 + ds * lam * ls * (ls + 1))), atol=1e-15, rtol=0)
    k = np.array([3.0, -2.0, 5.0])
    gp = oracle.ginv(L, 1.5, +1.0, ds, lam, k)
    gm = oracle.ginv(L, 1.5, -1.0, ds, lam, k)
    assert np.abs(gm - np.conj(gp)).max() < 1e-14             # G*(k) = conj G(k)
    gneg = oracle.ginv(L, 1.5, +1.0, ds, lam, -k)
    assert np.abs(gneg - gm).max() < 1e-14                    # G(-k) = G*(k)


# --- quadrature (utility_mod.f90:238-298, 346-374) ----------------------------------------

@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 6, 7, 33, 34, 66, 67])
def test_contour_weights_integrate_polynomials(oracle, n):
    ds = 1.0 / n
    w = oracle.simpson_weights(n, ds)
    x = np.arange(n + 1) * ds
    deg = 1 if n == 1 else 3
    for p in range(deg + 1):
        assert w @ x ** p == pytest.approx(1.0 / (p + 1), abs=1e-13)
    if n % 2 == 0 and n >= 2:        # the reference's rule exactly
        ref = np.ones(n + 1)
        ref[1:n:2] = 4
        ref[2:n:2] = 2
        assert np.allclose(w, ds / 3 * ref, atol=1e-16, rtol=0)


def test_volume_integral(oracle):
    shape, box = (8, 6, 10), (2.0, 3.0, 1.5)
    nz, ny, nx = shape
    assert oracle.volume_integral(np.ones(shape), box) == pytest.approx(np.prod(box), rel=1e-14)
    x = (np.arange(nx) * box[0] / nx)[None, None, :]
    f = np.broadcast_to(np.cos(2 * np.pi * x / box[0]), shape)
    assert abs(oracle.volume_integral(f, box)) < 1e-13
    rng = np.random.default_rng(3)
    g = rng.standard_normal(shape)
    c = lambda n: np.where(np.arange(n) % 2 == 0, 2 / 3, 4 / 3)
    W = c(nz)[:, None, None] * c(ny)[None, :, None] * c(nx)[None, None, :] * np.prod(box) / g.size
    assert oracle.volume_integral(g, box) == pytest.approx((W * g).sum(), rel=1e-13)


# --- sweeps -------------------------------------------------------------------------------

@pytest.mark.parametrize("name", ["ab_even", "abc_odd_radix3", "homopolymer"])
def test_sweep_matches_independent_restatement(oracle, name):
    f = np.load(os.path.join(GOLDEN, f"{name}.npz"))
    orc = oracle.Oracle(tuple(f["shape"]), tuple(f["box"]), tuple(f["Nblk"]), tuple(f["L"]),
                        tuple(f["lam"]))
    q = orc.forward(f["w"])
    qs = orc.backward(f["w"])
    phi, Q, td = orc.density(q, qs)
    Ns = orc.Ns
    tol = 1e-12
    for got, ref in [(orc.slice(q, Ns), f["q_end"]), (orc.slice(q, int(f["s_mid"])), f["q_mid"]),
                     (orc.slice(qs, 0), f["qs_0"]), (phi, f["phi"])]:
        assert np.abs(got - ref).max() <= tol * np.abs(ref).max()
    assert Q == pytest.approx(float(f["Q"]), rel=tol)
    assert td == pytest.approx(float(f["totden"]), rel=tol)


def _uniform_recurrence(Nblk, wconst, literal=0):
    Ns = sum(Nblk)
    ds = 1.0 / Ns
    start = np.concatenate([[0], np.cumsum(Nblk)[:-1]])
    third = float(np.float32(1.0) / np.float32(3.0)) if literal else 1.0 / 3.0
    c3 = float(np.float32(11.0) / np.float32(6.0)) if literal else 11.0 / 6.0

    def blk(s):
        for b in range(len(Nblk)):
            if s <= start[b] + Nblk[b]:
                return b

    q = [1.0]
    for s in range(1, Ns + 1):
        W = wconst[blk(s)]
        o = min(s, 3)
        h = q[::-1]
        if o == 1:
            P, c = (1 - ds * W) * h[0], 1.0
        elif o == 2:
            P, c = 2 * h[0] - 0.5 * h[1] - ds * W * (2 * h[0] - h[1]), 1.5
        else:
            P, c = (3 * h[0] - 1.5 * h[1] + third * h[2] - ds * W * (3 * h[0] - 3 * h[1] + h[2]), c3)
        v = P / c
        # literal mode: the G-applied spectrum is staged through default-real cmplx()
        # (wormlike_MDE_2dcomp_mod.f90:429); at k = 0 that rounds the value to FP32
        q.append(float(np.float32(v)) if literal else v)
    return np.array(q)


@pytest.mark.parametrize("literal", [0, 1])
def test_uniform_field_closed_form(oracle, literal):
    """Constant w per block keeps q uniform and isotropic: q_0(s) follows the scalar
    BDF recurrence with G = c (k = 0, l = 0); with w = 0 it stays exactly 1."""
    shape, box, Nblk, L, lam = (4, 4, 4), (2.0, 2.0, 2.0), (5, 6, 5), (2, 3, 2), (1.0, 0.5, 2.0)
    wc = [0.7, -1.2, 2.5]
    w = np.stack([np.full(shape, v) for v in wc])
    orc = oracle.Oracle(shape, box, Nblk, L, lam, literal=literal)
    q = orc.forward(w)
    ref = _uniform_recurrence(Nblk, wc, literal)
    for s in range(orc.Ns + 1):
        sl = orc.slice(q, s)
        assert np.abs(sl[0] - ref[s]).max() < 1e-13 * max(1, abs(ref[s]))
        assert np.abs(sl[1:]).max() < 1e-14
    w0 = np.zeros((3,) + shape)
    q0 = orc.forward(w0)
    last = orc.slice(q0, orc.Ns)[0]
    assert np.abs(last - 1).max() < 1e-14
    if literal:  # the FP32 literals and staging perturb the result at the 1e-8 level
        dev = abs(orc.slice(q, orc.Ns)[0].mean() - _uniform_recurrence(Nblk, wc, 0)[-1])
        assert 1e-10 < dev < 1e-5


def test_w_zero_gives_Q_one_and_phi_fraction(oracle):
    shape, box, Nblk = (6, 4, 8), (3.0, 2.0, 4.0), (7, 5, 4)
    orc = oracle.Oracle(shape, box, Nblk, (3, 2, 4), (1.0, 2.0, 0.5))
    w = np.zeros((3,) + shape)
    phi, Q, td = orc.density(orc.forward(w), orc.backward(w))
    assert Q == pytest.approx(1.0, abs=1e-14)
    f = np.array(Nblk) / sum(Nblk)
    for b in range(3):
        assert np.abs(phi[b] - f[b]).max() < 1e-13
    assert td == pytest.approx(np.prod(box), rel=1e-13)


def test_homopolymer_reversal_identity(oracle):
    """One block: q*_i(s) = (-1)^{l_i} q_i(Ns - s) exactly (parity maps G(sigma=+1) to
    G(sigma=-1) and the BDF start-up is mirrored)."""
    f = np.load(os.path.join(GOLDEN, "homopolymer.npz"))
    orc = oracle.Oracle(tuple(f["shape"]), tuple(f["box"]), tuple(f["Nblk"]), tuple(f["L"]),
                        tuple(f["lam"]))
    q = orc.forward(f["w"])
    qs = orc.backward(f["w"])
    L = int(f["L"][0])
    par = np.array([(-1) ** l for l in range(L + 1) for m in range(-l, l + 1)])[:, None, None, None]
    for s in range(orc.Ns + 1):
        a = orc.slice(qs, s)
        b = par * orc.slice(q, orc.Ns - s)
        assert np.abs(a - b).max() < 1e-13 * np.abs(b).max()


def test_translation_invariance(oracle):
    """Shifting w by an even number of cells (the periodic Simpson weights alternate
    2/3, 4/3) leaves Q unchanged and shifts phi."""
    shape, box, Nblk = (4, 6, 8), (2.0, 3.0, 4.0), (4, 4, 5)
    rng = np.random.default_rng(5)
    w = rng.standard_normal((3,) + shape)
    orc = oracle.Oracle(shape, box, Nblk, (2, 2, 3), (1.0, 1.0, 0.5))
    phi, Q, _ = orc.density(orc.forward(w), orc.backward(w))
    ws = np.roll(w, (2, -2, 4), axis=(1, 2, 3))
    phis, Qs, _ = orc.density(orc.forward(ws), orc.backward(ws))
    assert Qs == pytest.approx(Q, rel=1e-13)
    assert np.abs(np.roll(phi, (2, -2, 4), axis=(1, 2, 3)) - phis).max() < 1e-12


# --- free energy (SCFT_Density_FE.f90:273-342; ABC form of the same terms) ----------------

def _fe_ctx(oracle, shape, box, nblk):
    Nblk = (4, 5, 3)[:nblk]
    return oracle.Oracle(shape, box, Nblk, (1,) * nblk, (1.0,) * nblk)


@pytest.mark.parametrize("nblk", [2, 3])
def test_free_energy_homogeneous_closed_form(oracle, nblk):
    """Uniform fields w_b and phi_b = f_b: F = -ln Q + sum_{a<b} chi_ab f_a f_b - sum_b w_b f_b
    (the xi term vanishes because sum_b f_b = 1); w = 0, Q = 1 gives the Flory-Huggins
    energy of the disordered melt exactly."""
    shape, box = (4, 6, 8), (2.0, 3.0, 4.0)
    orc = _fe_ctx(oracle, shape, box, nblk)
    f = np.array([0.25, 0.45, 0.30][:nblk])
    f = f / f.sum()
    chi = np.zeros((3, 3))
    chi[0, 1], chi[0, 2], chi[1, 2] = 13.0, 21.0, 17.5
    phi = np.stack([np.full(shape, v) for v in f])
    fh = sum(chi[a, b] * f[a] * f[b] for a in range(nblk) for b in range(a + 1, nblk))
    assert orc.free_energy(np.zeros((nblk,) + shape), phi, 1.0, chi) == pytest.approx(fh, abs=1e-13)
    wc = np.array([0.8, -1.7, 2.9][:nblk])
    w = np.stack([np.full(shape, v) for v in wc])
    Q = 0.37
    ref = -np.log(Q) + fh - (wc * f).sum()
    assert orc.free_energy(w, phi, Q, chi) == pytest.approx(ref, abs=1e-13)
    orc.close()


def test_free_energy_two_blocks_is_the_reference_form(oracle):
    """Two blocks on random (w, phi, Q), phi not normalised: the oracle's ABC form
    (pair sum, xi = mean field) equals the reference's AB expression
    chiN RA RB - wA RA - wB RB + (wA + wB)/2 (RA + RB - 1), restated in numpy."""
    sys.path.insert(0, GOLDEN)
    import np_restatement as NR
    shape, box = (6, 4, 8), (3.0, 2.0, 4.5)
    orc = _fe_ctx(oracle, shape, box, 2)
    rng = np.random.default_rng(21)
    for _ in range(3):
        w = 3.0 * rng.standard_normal((2,) + shape)
        phi = 0.5 + 0.3 * rng.standard_normal((2,) + shape)
        Q = float(rng.uniform(0.1, 5.0))
        chiN = float(rng.uniform(5.0, 30.0))
        chi = np.zeros((3, 3))
        chi[0, 1] = chiN
        ref = NR.free_energy_ab(w[0], w[1], phi[0], phi[1], Q, chiN, box)
        assert orc.free_energy(w, phi, Q, chi) == pytest.approx(ref, rel=1e-13, abs=1e-13)
    orc.close()


def test_free_energy_gauge_invariance(oracle):
    """A constant c added to every field multiplies Q by exp(-c) and leaves F unchanged,
    for any phi (a wrong sign of the xi or w.phi term breaks this when sum_b phi_b != 1);
    with Q held fixed F moves by exactly -c."""
    shape, box = (4, 6, 8), (2.0, 3.0, 4.0)
    orc = _fe_ctx(oracle, shape, box, 3)
    rng = np.random.default_rng(22)
    w = 2.0 * rng.standard_normal((3,) + shape)
    phi = 0.33 + 0.2 * rng.standard_normal((3,) + shape)
    chi = np.triu(rng.uniform(10, 30, (3, 3)), 1)
    Q = 1.7
    F = orc.free_energy(w, phi, Q, chi)
    for c in (-3.0, 0.5, 7.25):
        assert orc.free_energy(w + c, phi, Q * np.exp(-c), chi) == pytest.approx(F, abs=1e-12)
        assert orc.free_energy(w + c, phi, Q, chi) == pytest.approx(F - c, abs=1e-12)
    orc.close()


def test_free_energy_pair_terms_are_additive(oracle):
    """F is linear in chi with slope <phi_a phi_b> for each pair (a < b only: the lower
    triangle of chi is not read)."""
    shape, box = (4, 4, 6), (2.0, 2.0, 3.0)
    orc = _fe_ctx(oracle, shape, box, 3)
    rng = np.random.default_rng(23)
    w = rng.standard_normal((3,) + shape)
    phi = 0.33 + 0.2 * rng.standard_normal((3,) + shape)
    F0 = orc.free_energy(w, phi, 1.0, np.zeros((3, 3)))
    for a, b in [(0, 1), (0, 2), (1, 2)]:
        chi = np.zeros((3, 3))
        chi[a, b] = 4.0
        slope = oracle.volume_integral(phi[a] * phi[b], box) / np.prod(box)
        assert orc.free_energy(w, phi, 1.0, chi) - F0 == pytest.approx(4.0 * slope, rel=1e-12)
        chi_lower = np.zeros((3, 3))
        chi_lower[b, a] = 4.0
        assert orc.free_energy(w, phi, 1.0, chi_lower) == pytest.approx(F0, abs=1e-13)
    orc.close()


def test_free_energy_of_disordered_sweep(oracle):
    """End to end through the oracle's sweeps: w = 0 gives Q = 1 and phi = f, so F is the
    disordered melt's sum_{a<b} chi_ab f_a f_b."""
    shape, box, Nblk = (4, 4, 6), (2.0, 2.0, 3.0), (5, 7, 4)
    orc = oracle.Oracle(shape, box, Nblk, (2, 3, 2), (1.0, 0.5, 2.0))
    w = np.zeros((3,) + shape)
    phi, Q, _ = orc.density(orc.forward(w), orc.backward(w))
    chi = np.zeros((3, 3))
    chi[0, 1], chi[0, 2], chi[1, 2] = 20.0, 15.0, 25.0
    f = np.array(Nblk) / sum(Nblk)
    fh = sum(chi[a, b] * f[a] * f[b] for a in range(3) for b in range(a + 1, 3))
    assert orc.free_energy(w, phi, Q, chi) == pytest.approx(fh, abs=1e-12)
    orc.close()
