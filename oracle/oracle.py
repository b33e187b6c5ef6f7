"""ctypes front end of the CPU oracle (TEST INFRASTRUCTURE ONLY).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this
module, and only as the checker (or the timed CPU baseline).  The product path
never imports it.  The C restatement it loads is oracle/wlc_oracle.c; see that
file's header for the reference file:line each function follows.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


class OrcParams(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
        ("box", ctypes.c_double * 3),
        ("nblk", ctypes.c_int), ("Ns", ctypes.c_int),
        ("Nblk", ctypes.c_int * 3), ("L", ctypes.c_int * 3),
        ("lam", ctypes.c_double * 3),
        ("literal", ctypes.c_int), ("nthreads", ctypes.c_int),
    ]


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        dp = ctypes.POINTER(ctypes.c_double)
        L.orc_create.restype = ctypes.c_void_p
        L.orc_create.argtypes = [ctypes.POINTER(OrcParams)]
        L.orc_destroy.argtypes = [ctypes.c_void_p]
        L.orc_slice_M.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.orc_slice_offset.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.orc_slice_offset.restype = ctypes.c_longlong
        L.orc_store_size.argtypes = [ctypes.c_void_p]
        L.orc_store_size.restype = ctypes.c_longlong
        L.orc_block_of.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.orc_forward.argtypes = [ctypes.c_void_p, dp, dp]
        L.orc_backward.argtypes = [ctypes.c_void_p, dp, dp]
        L.orc_density.argtypes = [ctypes.c_void_p, dp, dp, dp, dp, dp]
        L.orc_free_energy.argtypes = [ctypes.c_void_p, dp, dp, ctypes.c_double, dp]
        L.orc_free_energy.restype = ctypes.c_double
        L.orc_step.argtypes = [ctypes.c_void_p, dp, ctypes.c_int, ctypes.c_int,
                               ctypes.POINTER(dp), dp]
        L.orc_three_j.argtypes = [ctypes.c_int] * 6
        L.orc_three_j.restype = ctypes.c_double
        L.orc_triple_product.argtypes = [ctypes.c_int] * 7
        L.orc_triple_product.restype = ctypes.c_double
        L.orc_rmatrices.argtypes = [ctypes.c_int, ctypes.c_int, dp, dp, dp]
        L.orc_ginv.argtypes = [ctypes.c_int] + [ctypes.c_double] * 7 + [ctypes.c_int, dp]
        L.orc_rfft3.argtypes = [ctypes.c_int] * 3 + [dp, dp]
        L.orc_irfft3.argtypes = [ctypes.c_int] * 3 + [dp, dp]
        L.orc_fft1.argtypes = [ctypes.c_int, dp, ctypes.c_int]
        L.orc_simpson_weights.argtypes = [ctypes.c_int, ctypes.c_double, dp]
        L.orc_volume_integral.argtypes = [ctypes.c_int] * 3 + [dp, dp]
        L.orc_volume_integral.restype = ctypes.c_double
        L.orc_time_bdf3_steps.argtypes = [ctypes.c_void_p, dp, ctypes.c_int]
        L.orc_time_bdf3_steps.restype = ctypes.c_double
        L.orc_set_threads.argtypes = [ctypes.c_int]
        L.orc_set_threads.restype = None
        _LIB = L
    return _LIB


def _p(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def three_j(l1, l2, l3, m1, m2, m3):
    return lib().orc_three_j(l1, l2, l3, m1, m2, m3)


def triple_product(l1, l2, l3, m1, m2, m3, literal=0):
    return lib().orc_triple_product(l1, l2, l3, m1, m2, m3, literal)


def rmatrices(L, literal=0):
    M = (L + 1) ** 2
    R = np.zeros((3, M, M))
    lib().orc_rmatrices(L, literal, _p(R[0]), _p(R[1]), _p(R[2]))
    return R


def ginv(L, cdelta, sigma, ds, lam, k, literal=0):
    M = (L + 1) ** 2
    out = np.zeros((M, M, 2))
    rc = lib().orc_ginv(L, cdelta, sigma, ds, lam, k[0], k[1], k[2], literal, _p(out))
    if rc:
        raise RuntimeError("singular G")
    return out[..., 0] + 1j * out[..., 1]


def rfft3(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    nz, ny, nx = a.shape
    out = np.zeros((nz, ny, nx // 2 + 1, 2))
    lib().orc_rfft3(nx, ny, nz, _p(a), _p(out))
    return out[..., 0] + 1j * out[..., 1]


def irfft3(spec, nx):
    nz, ny, nxh = spec.shape
    s = np.ascontiguousarray(np.stack([spec.real, spec.imag], -1))
    out = np.zeros((nz, ny, nx))
    lib().orc_irfft3(nx, ny, nz, _p(s), _p(out))
    return out


def fft1(x, sign=-1):
    x = np.ascontiguousarray(np.stack([np.real(x), np.imag(x)], -1), dtype=np.float64)
    lib().orc_fft1(x.shape[0], _p(x), sign)
    return x[:, 0] + 1j * x[:, 1]


def simpson_weights(n, ds):
    w = np.zeros(n + 1)
    lib().orc_simpson_weights(n, ds, _p(w))
    return w


def volume_integral(f, box):
    f = np.ascontiguousarray(f, dtype=np.float64)
    nz, ny, nx = f.shape
    b = np.asarray(box, dtype=np.float64)
    return lib().orc_volume_integral(nx, ny, nz, _p(b), _p(f))


class Oracle:
    """One chain configuration: grid, box, blocks (lengths, Lmax, lambda)."""

    def __init__(self, shape, box, Nblk, L, lam, literal=0, nthreads=0):
        nz, ny, nx = shape
        p = OrcParams()
        p.nx, p.ny, p.nz = nx, ny, nz
        for a in range(3):
            p.box[a] = float(box[a])
        nb = len(Nblk)
        p.nblk = nb
        p.Ns = int(sum(Nblk))
        for b in range(nb):
            p.Nblk[b] = int(Nblk[b])
            p.L[b] = int(L[b])
            p.lam[b] = float(lam[b])
        p.literal = literal
        p.nthreads = nthreads
        self.params = p
        self.shape = (nz, ny, nx)
        self.N = nx * ny * nz
        self.nblk = nb
        self.Ns = p.Ns
        self.box = tuple(float(b) for b in box)
        self.h = lib().orc_create(ctypes.byref(p))
        if not self.h:
            raise ValueError("bad oracle parameters")

    def close(self):
        if self.h:
            lib().orc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def slice_M(self, s):
        return lib().orc_slice_M(self.h, s)

    def block_of(self, s):
        return lib().orc_block_of(self.h, s)

    def new_store(self):
        return np.zeros(lib().orc_store_size(self.h))

    def slice(self, store, s):
        off = lib().orc_slice_offset(self.h, s)
        M = self.slice_M(s)
        return store[off:off + M * self.N].reshape((M,) + self.shape)

    def forward(self, w):
        w = np.ascontiguousarray(w, dtype=np.float64)
        q = self.new_store()
        if lib().orc_forward(self.h, _p(w), _p(q)):
            raise RuntimeError("oracle forward failed")
        return q

    def backward(self, w):
        w = np.ascontiguousarray(w, dtype=np.float64)
        q = self.new_store()
        if lib().orc_backward(self.h, _p(w), _p(q)):
            raise RuntimeError("oracle backward failed")
        return q

    def density(self, q, qs):
        phi = np.zeros((self.nblk,) + self.shape)
        Q = ctypes.c_double()
        td = ctypes.c_double()
        lib().orc_density(self.h, _p(q), _p(qs), _p(phi), ctypes.byref(Q), ctypes.byref(td))
        return phi, Q.value, td.value

    def free_energy(self, w, phi, Q, chi):
        w = np.ascontiguousarray(w, dtype=np.float64)
        chi = np.ascontiguousarray(chi, dtype=np.float64).reshape(3, 3)
        return lib().orc_free_energy(self.h, _p(w), _p(np.ascontiguousarray(phi)), Q, _p(chi))

    def step(self, w, direction, s, hist):
        w = np.ascontiguousarray(w, dtype=np.float64)
        dp = ctypes.POINTER(ctypes.c_double)
        arr = (dp * 3)()
        keep = []
        for j in range(3):
            if hist[j] is None:
                arr[j] = ctypes.cast(None, dp)
            else:
                h = np.ascontiguousarray(hist[j], dtype=np.float64)
                keep.append(h)
                arr[j] = _p(h)
        M = self.slice_M(s)
        out = np.zeros((M,) + self.shape)
        if lib().orc_step(self.h, _p(w), direction, s, arr, _p(out)):
            raise RuntimeError("oracle step failed")
        return out

    @staticmethod
    def set_threads(n):
        lib().orc_set_threads(int(n))

    def time_bdf3_steps(self, w, nsteps):
        w = np.ascontiguousarray(w, dtype=np.float64)
        return lib().orc_time_bdf3_steps(self.h, _p(w), nsteps)
