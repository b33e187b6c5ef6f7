"""Independent numpy/scipy restatement of the reference MDE path (test infrastructure).

It shares no code with oracle/wlc_oracle.c and is built from different ingredients,
so agreement between the two pins both:
  - R_x, R_y, R_z by Gauss-Legendre x uniform-phi quadrature of real spherical
    harmonics assembled from scipy's complex Y_l^m (Condon-Shortley phase), instead
    of the reference's Wigner-3j case analysis (SPH_mod.f90:11-98,
    init_array_2dcomp_mod.f90:166-182);
  - the 3D transforms by numpy.fft.rfftn / irfftn (pocketfft), whose inverse drops
    the imaginary parts of the kx = 0 and kx = nx/2 planes;
  - G(k)^-1 P by numpy.linalg.solve on the dense G of Gmatrix_2dcomp_mod.f90:470-495;
  - BDF1/2/3 steps, junction padding and the w-per-target-block rule of
    wormlike_MDE_2dcomp_mod.f90:323-937 / :1035-1532;
  - the contour Simpson rule (utility_mod.f90:346-374; odd block lengths: Simpson plus
    a 3/8 closing panel) and the periodic volume rule (utility_mod.f90:238-298).
"""
import numpy as np
from scipy.special import sph_harm_y


def basis(L):
    return [(l, m) for l in range(L + 1) for m in range(-l, l + 1)]


def real_sh(L, theta, phi):
    """Real harmonics Y_i(theta, phi) in the reference's order i = l^2 + l + m.

    The reference's triple product (SPH_mod.f90:11-98) corresponds to real harmonics
    that keep the Condon-Shortley phase: (-1)^m times the usual real combination of
    scipy's Y_l^{+-m}, so Y_{1,1} ~ -x and Y_{1,-1} ~ -y (R_x(0,(1,1)) = -1/sqrt3)."""
    out = []
    for l, m in basis(L):
        a = abs(m)
        if m == 0:
            y = sph_harm_y(l, 0, theta, phi).real
        elif m > 0:
            y = ((sph_harm_y(l, -a, theta, phi) + (-1) ** a * sph_harm_y(l, a, theta, phi))
                 / np.sqrt(2)).real
        else:
            y = (1j * (sph_harm_y(l, -a, theta, phi) - (-1) ** a * sph_harm_y(l, a, theta, phi))
                 / np.sqrt(2)).real
        out.append((-1) ** a * y)
    return np.array(out)


def rmatrices_quadrature(L):
    """<Y_i | u_a | Y_j> by quadrature exact for the polynomial degree involved."""
    nth = L + 3
    nph = 2 * L + 6
    x, wx = np.polynomial.legendre.leggauss(nth)
    theta = np.arccos(x)
    phi = 2 * np.pi * np.arange(nph) / nph
    TH, PH = np.meshgrid(theta, phi, indexing="ij")
    W = (wx[:, None] * np.full(nph, 2 * np.pi / nph)[None, :])
    Y = real_sh(L, TH, PH)                                   # [M, nth, nph]
    u = [np.sin(TH) * np.cos(PH), np.sin(TH) * np.sin(PH), np.cos(TH)]
    R = np.array([np.einsum("ipq,jpq,pq->ij", Y, Y, ua * W) for ua in u])
    return R


def signed(j, n):
    return np.where(j <= n // 2, j, j - n)


def kgrid(shape, box):
    nz, ny, nx = shape
    kx = signed(np.arange(nx // 2 + 1), nx) * 2 * np.pi / box[0]
    ky = signed(np.arange(ny), ny) * 2 * np.pi / box[1]
    kz = signed(np.arange(nz), nz) * 2 * np.pi / box[2]
    KZ, KY, KX = np.meshgrid(kz, ky, kx, indexing="ij")
    return KX, KY, KZ


def contour_weights(n, ds):
    w = np.zeros(n + 1)
    if n == 1:
        w[:] = ds / 2
        return w
    ns = n if n % 2 == 0 else n - 3
    if ns > 0:
        c = np.ones(ns + 1)
        c[1:ns:2] = 4.0
        c[2:ns:2] = 2.0
        w[:ns + 1] += ds / 3 * c
    if n % 2:
        w[ns:ns + 4] += 3 * ds / 8 * np.array([1.0, 3.0, 3.0, 1.0])
    return w


def volume_weights(shape, box):
    nz, ny, nx = shape
    c = lambda n: np.where(np.arange(n) % 2 == 0, 2.0 / 3.0, 4.0 / 3.0)
    dv = box[0] / nx * box[1] / ny * box[2] / nz
    return dv * c(nz)[:, None, None] * c(ny)[None, :, None] * c(nx)[None, None, :]


class Chain:
    def __init__(self, shape, box, Nblk, L, lam):
        self.shape, self.box = tuple(shape), tuple(box)
        self.Nblk, self.L, self.lam = list(Nblk), list(L), list(lam)
        self.nb = len(Nblk)
        self.Ns = int(sum(Nblk))
        self.ds = 1.0 / self.Ns
        self.start = np.concatenate([[0], np.cumsum(Nblk)[:-1]]).astype(int)
        self.M = [(l + 1) ** 2 for l in L]
        self.R = rmatrices_quadrature(max(L))
        self.K = kgrid(self.shape, self.box)

    def block_of(self, s):
        for b in range(self.nb):
            if s <= self.start[b] + self.Nblk[b]:
                return b
        return self.nb - 1

    def Mslice(self, s):
        return self.M[self.block_of(s)]

    def _G(self, blk, cdelta, sigma):
        M = self.M[blk]
        ls = np.array([l for l, m in basis(self.L[blk])])
        KX, KY, KZ = self.K
        R = self.R[:, :M, :M]
        G = (np.einsum("zyx,ij->zyxij", KX, R[0]) + np.einsum("zyx,ij->zyxij", KY, R[1])
             + np.einsum("zyx,ij->zyxij", KZ, R[2])) * (1j * sigma * self.ds)
        G += np.diag(cdelta + self.ds * self.lam[blk] * ls * (ls + 1))
        return G

    def step(self, w, direction, s, hist):
        Ns = self.Ns
        order = min(s, 3) if direction > 0 else min(Ns - s, 3)
        blk = self.block_of(s)
        M = self.M[blk]
        a = {1: [1.0], 2: [2.0, -0.5], 3: [3.0, -1.5, 1.0 / 3.0]}[order]
        b = {1: [1.0], 2: [2.0, -1.0], 3: [3.0, -3.0, 1.0]}[order]
        cdelta = {1: 1.0, 2: 1.5, 3: 11.0 / 6.0}[order]
        P = np.zeros((M,) + self.shape)
        for j in range(order):
            h = hist[j]
            m = min(M, h.shape[0])
            P[:m] += (a[j] - self.ds * w[blk] * b[j]) * h[:m]
        Ph = np.fft.rfftn(P, axes=(1, 2, 3))                     # [M, z, y, kx]
        G = self._G(blk, cdelta, 1.0 if direction > 0 else -1.0)
        rhs = np.moveaxis(Ph, 0, -1)[..., None]
        sol = np.linalg.solve(G, rhs)[..., 0]
        qh = np.moveaxis(sol, -1, 0)
        return np.fft.irfftn(qh, s=self.shape, axes=(1, 2, 3))

    def forward(self, w):
        q = [None] * (self.Ns + 1)
        q0 = np.zeros((self.Mslice(0),) + self.shape)
        q0[0] = 1.0
        q[0] = q0
        for s in range(1, self.Ns + 1):
            hist = [q[s - j - 1] if s - j - 1 >= 0 else None for j in range(3)]
            q[s] = self.step(w, +1, s, hist)
        return q

    def backward(self, w):
        qs = [None] * (self.Ns + 1)
        qn = np.zeros((self.Mslice(self.Ns),) + self.shape)
        qn[0] = 1.0
        qs[self.Ns] = qn
        for s in range(self.Ns - 1, -1, -1):
            hist = [qs[s + j + 1] if s + j + 1 <= self.Ns else None for j in range(3)]
            qs[s] = self.step(w, -1, s, hist)
        return qs

    def density(self, q, qs):
        Wv = volume_weights(self.shape, self.box)
        V = np.prod(self.box)
        Q = (Wv * q[self.Ns][0]).sum() / V
        rho = np.zeros((self.nb,) + self.shape)
        for b in range(self.nb):
            cw = contour_weights(self.Nblk[b], self.ds)
            for j in range(self.Nblk[b] + 1):
                s = self.start[b] + j
                lim = min(self.Mslice(s), self.M[b])
                rho[b] += cw[j] * (q[s][:lim] * qs[s][:lim]).sum(0)
        totden = (Wv * rho.sum(0)).sum()
        return V * rho / totden, Q, totden


def free_energy_ab(wA, wB, RA, RB, Q, chiN, box):
    """Free energy per chain of the AB melt as the reference writes it
    (SCFT_Density_FE.f90:296-298, :326-333): -ln Q plus the periodic-Simpson volume
    average of  chiN RA RB - wA RA - wB RB + (wA + wB)/2 (RA + RB - 1)."""
    ar = chiN * RA * RB - wA * RA - wB * RB + 0.5 * (wA + wB) * (RA + RB - 1.0)
    Wv = volume_weights(wA.shape, box)
    return -np.log(Q) + (Wv * ar).sum() / np.prod(box)
