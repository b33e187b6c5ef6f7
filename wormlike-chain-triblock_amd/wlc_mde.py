"""Python host binding of libwlc_mde.so (the MI355X MDE propagator solve).

Mirrors the reference's MDE module interface (module wormlike_MDE,
wormlike_MDE_2dcomp_mod.f90:227-244): a `WormlikeMDE` object holds what the
reference keeps in module globals (grid, box, blocks, fields w) and exposes

    MDE_q()       forward propagator   (wormlike_MDE_2dcomp_mod.f90:231-956)
    MDE_qstar()   backward propagator  (wormlike_MDE_2dcomp_mod.f90:961-1557)
    density()     Q, totden, phi       (SCFT_Density_FE.f90:3-159)

All work runs on the GPU through the C ABI declared in include/wlc_mde.h.  There
is no CPU fallback: constructing a solver raises when the library or a GPU is
missing.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libwlc_mde.so")

WLC_OK = 0
_STATUS = {-1: "WLC_EINVAL", -2: "WLC_ENOMEM", -3: "WLC_EHIP", -4: "WLC_ESTATE",
           -5: "WLC_ENONFINITE"}


class WlcError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{_STATUS.get(code, code)}: {msg}")
        self.code = code


class WlcParams(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
        ("box", ctypes.c_double * 3),
        ("nblk", ctypes.c_int), ("Ns", ctypes.c_int),
        ("Nblk", ctypes.c_int * 3), ("Lmax", ctypes.c_int * 3),
        ("lam", ctypes.c_double * 3),
        ("keep_qstar", ctypes.c_int), ("ckpt_every", ctypes.c_int),
        ("spectrum_fp32", ctypes.c_int),
        ("reserved", ctypes.c_int * 5),
    ]


EXPORTS = [
    "wlc_mde_create", "wlc_mde_set_box", "wlc_mde_set_w", "wlc_mde_set_w_device",
    "wlc_mde_forward", "wlc_mde_backward", "wlc_mde_density", "wlc_mde_density_device",
    "wlc_mde_get_q", "wlc_mde_slice_M", "wlc_mde_block_of", "wlc_mde_sync",
    "wlc_mde_stream", "wlc_mde_set_timing", "wlc_mde_get_timing", "wlc_mde_event_record",
    "wlc_mde_event_elapsed", "wlc_mde_destroy", "wlc_mde_last_error", "wlc_test_solve",
    "wlc_mde_set_spectrum_fp32", "wlc_mde_forward_stored",
]

_lib = None


def load(path=LIB_PATH):
    """Load libwlc_mde.so and declare its signatures (no GPU call is made)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise FileNotFoundError(
            f"{path} is missing: build it first (python -c 'import __graft_entry__ as g; g.build()')")
    L = ctypes.CDLL(path)
    dp = ctypes.POINTER(ctypes.c_double)
    vp = ctypes.c_void_p
    L.wlc_mde_create.argtypes = [ctypes.POINTER(WlcParams), ctypes.POINTER(vp)]
    L.wlc_mde_set_box.argtypes = [vp, dp]
    L.wlc_mde_set_w.argtypes = [vp, dp]
    L.wlc_mde_set_w_device.argtypes = [vp, vp]
    L.wlc_mde_forward.argtypes = [vp]
    L.wlc_mde_backward.argtypes = [vp]
    L.wlc_mde_density.argtypes = [vp, dp, dp, dp]
    L.wlc_mde_density_device.argtypes = [vp, vp, dp, dp]
    L.wlc_mde_get_q.argtypes = [vp, ctypes.c_int, ctypes.c_int, dp]
    L.wlc_mde_slice_M.argtypes = [vp, ctypes.c_int]
    L.wlc_mde_block_of.argtypes = [vp, ctypes.c_int]
    L.wlc_mde_forward_stored.argtypes = [vp]
    L.wlc_mde_sync.argtypes = [vp]
    L.wlc_mde_stream.argtypes = [vp]
    L.wlc_mde_stream.restype = vp
    L.wlc_mde_set_timing.argtypes = [vp, ctypes.c_int]
    L.wlc_mde_set_spectrum_fp32.argtypes = [vp, ctypes.c_int]
    L.wlc_mde_get_timing.argtypes = [vp, dp, ctypes.POINTER(ctypes.c_longlong), ctypes.c_int]
    L.wlc_mde_event_record.argtypes = [vp, ctypes.c_int]
    L.wlc_mde_event_elapsed.argtypes = [vp, ctypes.c_int, ctypes.c_int, dp]
    L.wlc_mde_destroy.argtypes = [vp]
    L.wlc_mde_destroy.restype = None
    L.wlc_mde_last_error.restype = ctypes.c_char_p
    L.wlc_test_solve.argtypes = [ctypes.c_int, ctypes.c_int, dp, ctypes.c_double,
                                 ctypes.c_double, ctypes.c_double, ctypes.c_double, dp]
    _lib = L
    return L


def _check(rc):
    if rc != WLC_OK:
        raise WlcError(rc, _lib.wlc_mde_last_error().decode())
    return rc


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def block_lengths(Ns, fractions):
    """NA = NINT(fA Ns), NB = NINT(fB Ns), last block takes the rest
    (initialize_mod.f90:141-145)."""
    n = [int(np.floor(f * Ns + 0.5)) for f in fractions[:-1]]
    n.append(Ns - sum(n))
    return n


class WormlikeMDE:
    """One chain architecture on one grid; all state lives on the GPU."""

    def __init__(self, shape, box, Nblk, Lmax, lam, keep_qstar=False, ckpt_every=0,
                 spectrum_fp32=False):
        load()
        nz, ny, nx = shape
        p = WlcParams()
        p.nx, p.ny, p.nz = nx, ny, nz
        for a in range(3):
            p.box[a] = float(box[a])
        nb = len(Nblk)
        p.nblk = nb
        p.Ns = int(sum(Nblk))
        for b in range(nb):
            p.Nblk[b] = int(Nblk[b])
            p.Lmax[b] = int(Lmax[b])
            p.lam[b] = float(lam[b])
        p.keep_qstar = 1 if keep_qstar else 0
        p.ckpt_every = int(ckpt_every)
        p.spectrum_fp32 = 1 if spectrum_fp32 else 0
        self.params = p
        self.shape = (nz, ny, nx)
        self.N = nx * ny * nz
        self.nblk = nb
        self.Ns = p.Ns
        self.box = tuple(float(b) for b in box)
        h = ctypes.c_void_p()
        _check(_lib.wlc_mde_create(ctypes.byref(p), ctypes.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            _lib.wlc_mde_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # --- reference-shaped interface ------------------------------------------------
    def set_w(self, w):
        w = np.ascontiguousarray(w, dtype=np.float64)
        if w.shape != (self.nblk,) + self.shape:
            raise ValueError(f"w must have shape {(self.nblk,) + self.shape}")
        _check(_lib.wlc_mde_set_w(self.h, _dp(w)))

    def set_w_device(self, ptr):
        _check(_lib.wlc_mde_set_w_device(self.h, ctypes.c_void_p(ptr)))

    def set_box(self, box):
        b = np.asarray(box, dtype=np.float64)
        _check(_lib.wlc_mde_set_box(self.h, _dp(b)))
        self.box = tuple(float(x) for x in b)

    def MDE_q(self):
        _check(_lib.wlc_mde_forward(self.h))

    def MDE_qstar(self):
        _check(_lib.wlc_mde_backward(self.h))

    def density(self, want_phi=True):
        phi = np.zeros((self.nblk,) + self.shape) if want_phi else None
        Q = ctypes.c_double()
        td = ctypes.c_double()
        _check(_lib.wlc_mde_density(self.h, _dp(phi) if want_phi else None,
                                    ctypes.byref(Q), ctypes.byref(td)))
        return phi, Q.value, td.value

    def density_device(self, phi_ptr=None):
        Q = ctypes.c_double()
        td = ctypes.c_double()
        _check(_lib.wlc_mde_density_device(self.h, ctypes.c_void_p(phi_ptr) if phi_ptr else None,
                                           ctypes.byref(Q), ctypes.byref(td)))
        return Q.value, td.value

    def slice_M(self, s):
        return _lib.wlc_mde_slice_M(self.h, s)

    def block_of(self, s):
        return _lib.wlc_mde_block_of(self.h, s)

    def get_q(self, direction, s):
        M = self.slice_M(s)
        out = np.zeros((M,) + self.shape)
        _check(_lib.wlc_mde_get_q(self.h, direction, s, _dp(out)))
        return out

    def sync(self):
        _check(_lib.wlc_mde_sync(self.h))

    # --- measurement ---------------------------------------------------------------
    def forward_stored(self):
        """Number of leading forward slices kept whole (Ns + 1 with the full store)."""
        n = _lib.wlc_mde_forward_stored(self.h)
        if n < 0:
            _check(n)
        return n

    def set_spectrum_fp32(self, on):
        _check(_lib.wlc_mde_set_spectrum_fp32(self.h, 1 if on else 0))

    def set_timing(self, enable):
        """0/False off, 1/True live (sweeps stay concurrent), 2 sweeps serialized."""
        _check(_lib.wlc_mde_set_timing(self.h, int(enable)))

    def get_timing(self):
        ms = (ctypes.c_double * 5)()
        cnt = (ctypes.c_longlong * 5)()
        _check(_lib.wlc_mde_get_timing(self.h, ms, cnt, 5))
        return list(ms), list(cnt)

    def event_record(self, slot):
        _check(_lib.wlc_mde_event_record(self.h, slot))

    def event_elapsed(self, a, b):
        ms = ctypes.c_double()
        _check(_lib.wlc_mde_event_elapsed(self.h, a, b, ctypes.byref(ms)))
        return ms.value


def test_solve(L, kvecs, cdelta, sigma, ds, dslam, v):
    """k-aligned solve of G(k) x = v on the GPU for each row of kvecs."""
    load()
    kv = np.ascontiguousarray(kvecs, dtype=np.float64)
    nk = kv.shape[0]
    M = (L + 1) ** 2
    buf = np.ascontiguousarray(np.stack([np.real(v), np.imag(v)], -1).reshape(nk, M, 2))
    _check(_lib.wlc_test_solve(L, nk, _dp(kv), cdelta, sigma, ds, dslam, _dp(buf)))
    return buf[..., 0] + 1j * buf[..., 1]
