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


class ScftOpts(ctypes.Structure):
    _fields_ = [
        ("max_iter", ctypes.c_int), ("n_simple", ctypes.c_int), ("n_hist", ctypes.c_int),
        ("form", ctypes.c_int), ("lam_simple", ctypes.c_double),
        ("lam_anderson", ctypes.c_double), ("tol", ctypes.c_double),
        ("adaptive", ctypes.c_int), ("no_compact", ctypes.c_int), ("reserved", ctypes.c_int * 2),
        ("lam_max", ctypes.c_double),
    ]


class ScftStatus(ctypes.Structure):
    _fields_ = [
        ("iterations", ctypes.c_int), ("converged", ctypes.c_int), ("done", ctypes.c_int),
        ("pad", ctypes.c_int), ("F", ctypes.c_double), ("lnQ", ctypes.c_double),
        ("resid", ctypes.c_double * 3), ("restarts", ctypes.c_int), ("pad2", ctypes.c_int),
        ("time_ms", ctypes.c_double),
        ("elapsed_ms", ctypes.c_double),
    ]


EXPORTS = [
    "wlc_mde_create", "wlc_mde_set_box", "wlc_mde_set_w", "wlc_mde_set_w_device",
    "wlc_mde_forward", "wlc_mde_backward", "wlc_mde_density", "wlc_mde_density_device",
    "wlc_mde_get_q", "wlc_mde_slice_M", "wlc_mde_block_of", "wlc_mde_sync",
    "wlc_mde_stream", "wlc_mde_set_timing", "wlc_mde_get_timing", "wlc_mde_event_record",
    "wlc_mde_event_elapsed", "wlc_mde_destroy", "wlc_mde_last_error", "wlc_test_solve",
    "wlc_mde_set_spectrum_fp32", "wlc_mde_forward_stored", "wlc_mde_step",
    "wlc_mde_step_noalloc", "wlc_batch_create", "wlc_batch_size", "wlc_batch_launches",
    "wlc_batch_set_w", "wlc_batch_get_w", "wlc_batch_sweep", "wlc_batch_density",
    "wlc_batch_get_q", "wlc_batch_iterate", "wlc_batch_status", "wlc_batch_destroy",
    "wlc_unet_create", "wlc_unet_resize", "wlc_unet_num_params", "wlc_unet_layer", "wlc_unet_forward",
    "wlc_unet_gradients", "wlc_unet_adam", "wlc_unet_train_step", "wlc_unet_get",
    "wlc_unet_set_params", "wlc_unet_destroy", "wlc_test_conv3d",
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
    L.wlc_mde_step.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.POINTER(vp), vp, vp]
    L.wlc_mde_step_noalloc.argtypes = [ctypes.POINTER(WlcParams), ctypes.c_int, ctypes.c_int,
                                       ctypes.POINTER(vp), vp, vp, vp]
    L.wlc_batch_create.argtypes = [ctypes.POINTER(WlcParams), dp, ctypes.c_int,
                                   ctypes.POINTER(ScftOpts), ctypes.POINTER(vp)]
    L.wlc_batch_size.argtypes = [vp]
    L.wlc_batch_launches.argtypes = [vp, ctypes.c_int]
    L.wlc_batch_set_w.argtypes = [vp, ctypes.c_int, dp]
    L.wlc_batch_get_w.argtypes = [vp, ctypes.c_int, dp]
    L.wlc_batch_sweep.argtypes = [vp]
    L.wlc_batch_density.argtypes = [vp, ctypes.c_int, dp, dp, dp]
    L.wlc_batch_get_q.argtypes = [vp, ctypes.c_int, ctypes.c_int, dp]
    L.wlc_batch_iterate.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
    L.wlc_batch_status.argtypes = [vp, ctypes.c_int, ctypes.POINTER(ScftStatus)]
    L.wlc_batch_destroy.argtypes = [vp]
    L.wlc_batch_destroy.restype = None
    fp = ctypes.POINTER(ctypes.c_float)
    L.wlc_unet_create.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_ulonglong, ctypes.POINTER(vp)]
    L.wlc_unet_resize.argtypes = [vp, ctypes.c_int, ctypes.c_int]
    L.wlc_unet_num_params.argtypes = [vp, ctypes.c_int]
    L.wlc_unet_num_params.restype = ctypes.c_longlong
    L.wlc_unet_layer.argtypes = [vp, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
    L.wlc_unet_forward.argtypes = [vp, fp, fp]
    L.wlc_unet_gradients.argtypes = [vp, fp, fp, fp]
    L.wlc_unet_adam.argtypes = [vp, ctypes.c_float]
    L.wlc_unet_train_step.argtypes = [vp, fp, fp, ctypes.c_float, fp]
    L.wlc_unet_get.argtypes = [vp, ctypes.c_int, fp]
    L.wlc_unet_set_params.argtypes = [vp, fp]
    L.wlc_unet_destroy.argtypes = [vp]
    L.wlc_unet_destroy.restype = None
    L.wlc_test_conv3d.argtypes = [ctypes.c_int] * 6 + [fp] * 8
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

    def step_device(self, direction, s, hist_ptrs, w_ptr, out_ptr):
        """One contour step on device buffers (dense [M][nz][ny][nx] slices, integer device
        addresses; hist_ptrs[j] = slice s - direction (j + 1), None beyond the BDF order),
        queued on the handle's stream."""
        _check(_lib.wlc_mde_step(self.h, direction, s, _ptr3(hist_ptrs), ctypes.c_void_p(w_ptr),
                                 ctypes.c_void_p(out_ptr)))

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


def _ptr3(ptrs):
    arr = (ctypes.c_void_p * 3)()
    for j in range(3):
        arr[j] = ptrs[j] if j < len(ptrs) and ptrs[j] else None
    return arr


def make_params(shape, box, Nblk, Lmax, lam, spectrum_fp32=False):
    """wlc_params of a chain on a grid (shape = (nz, ny, nx))."""
    nz, ny, nx = shape
    p = WlcParams()
    p.nx, p.ny, p.nz = nx, ny, nz
    for a in range(3):
        p.box[a] = float(box[a])
    p.nblk = len(Nblk)
    p.Ns = int(sum(Nblk))
    for b in range(len(Nblk)):
        p.Nblk[b] = int(Nblk[b])
        p.Lmax[b] = int(Lmax[b])
        p.lam[b] = float(lam[b])
    p.spectrum_fp32 = 1 if spectrum_fp32 else 0
    return p


def step_noalloc(params, direction, s, hist_ptrs, w_ptr, out_ptr, stream=None):
    """wlc_mde_step_noalloc: one contour step with no handle and no allocation (the scratch
    is the library's fixed device array), queued on `stream` (an integer hipStream_t)."""
    load()
    _check(_lib.wlc_mde_step_noalloc(ctypes.byref(params), direction, s, _ptr3(hist_ptrs),
                                     ctypes.c_void_p(w_ptr), ctypes.c_void_p(out_ptr),
                                     ctypes.c_void_p(stream)))


def scft_opts(max_iter=500, n_simple=20, n_hist=10, form=0, lam_simple=0.05, lam_anderson=0.1,
              tol=1e-6, adaptive=False, lam_max=1.0, compact=True):
    o = ScftOpts()
    o.max_iter, o.n_simple, o.n_hist, o.form = max_iter, n_simple, n_hist, form
    o.lam_simple, o.lam_anderson, o.tol = lam_simple, lam_anderson, tol
    o.adaptive, o.lam_max = 1 if adaptive else 0, lam_max
    o.no_compact = 0 if compact else 1
    return o


class Batch:
    """Many independent SCFT instances advanced together on the GPU (wlc_batch_*).

    cases: list of dicts with shape (nz, ny, nx), box, Nblk, L, lam and chi (3x3, chi_ab N);
    opts: ScftOpts (scft_opts()) or None for the defaults."""

    def __init__(self, cases, opts=None):
        load()
        n = len(cases)
        self.n = n
        P = (WlcParams * n)()
        chi = np.zeros((n, 9))
        self.shapes, self.nblk = [], []
        for i, c in enumerate(cases):
            P[i] = make_params(c["shape"], c["box"], c["Nblk"], c["L"], c["lam"],
                               c.get("spectrum_fp32", False))
            chi[i] = np.asarray(c["chi"], dtype=np.float64).reshape(9)
            self.shapes.append(tuple(c["shape"]))
            self.nblk.append(len(c["Nblk"]))
        h = ctypes.c_void_p()
        _check(_lib.wlc_batch_create(P, _dp(chi), n, ctypes.byref(opts) if opts is not None else None,
                                     ctypes.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            _lib.wlc_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def launches(self, with_update=True):
        return _lib.wlc_batch_launches(self.h, 1 if with_update else 0)

    def set_w(self, i, w):
        w = np.ascontiguousarray(w, dtype=np.float64)
        if w.shape != (self.nblk[i],) + self.shapes[i]:
            raise ValueError("w has the wrong shape")
        _check(_lib.wlc_batch_set_w(self.h, i, _dp(w)))

    def get_w(self, i):
        w = np.zeros((self.nblk[i],) + self.shapes[i])
        _check(_lib.wlc_batch_get_w(self.h, i, _dp(w)))
        return w

    def sweep(self):
        _check(_lib.wlc_batch_sweep(self.h))

    def density(self, i):
        phi = np.zeros((self.nblk[i],) + self.shapes[i])
        Q, td = ctypes.c_double(), ctypes.c_double()
        _check(_lib.wlc_batch_density(self.h, i, _dp(phi), ctypes.byref(Q), ctypes.byref(td)))
        return phi, Q.value, td.value

    def get_q(self, i, s, M):
        out = np.zeros((M,) + self.shapes[i])
        _check(_lib.wlc_batch_get_q(self.h, i, s, _dp(out)))
        return out

    def iterate(self, niter, use_graph=True):
        act = ctypes.c_int()
        _check(_lib.wlc_batch_iterate(self.h, int(niter), 1 if use_graph else 0, ctypes.byref(act)))
        return act.value

    def status(self, i):
        st = ScftStatus()
        _check(_lib.wlc_batch_status(self.h, i, ctypes.byref(st)))
        return dict(iterations=st.iterations, converged=bool(st.converged), done=bool(st.done),
                    F=st.F, lnQ=st.lnQ, resid=list(st.resid), restarts=st.restarts,
                    time_ms=st.time_ms,
                    elapsed_ms=st.elapsed_ms)


def _fp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


class UNet:
    """The warm-start network (wlc_unet_*): x [B][n][n][n][11] -> y [B][n][n][n][3]."""

    def __init__(self, n, batch, seed=0):
        load()
        h = ctypes.c_void_p()
        _check(_lib.wlc_unet_create(int(n), int(batch), int(seed), ctypes.byref(h)))
        self.h, self.n, self.B = h, n, batch
        self.np = _lib.wlc_unet_num_params(h, 1)

    def close(self):
        if getattr(self, "h", None):
            _lib.wlc_unet_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def resize(self, n, batch):
        _check(_lib.wlc_unet_resize(self.h, int(n), int(batch)))
        self.n, self.B = n, batch

    def num_params(self, padded=False):
        return _lib.wlc_unet_num_params(self.h, 1 if padded else 0)

    def layers(self):
        out = []
        for i in range(11):
            info = (ctypes.c_int * 8)()
            _check(_lib.wlc_unet_layer(self.h, i, info))
            out.append(dict(zip(("cin", "cout", "cin_real", "cout_real", "level", "relu", "woff", "boff"),
                                list(info))))
        return out

    def _x(self, x):
        x = np.ascontiguousarray(x, dtype=np.float32)
        if x.shape != (self.B, self.n, self.n, self.n, 11):
            raise ValueError("x must be [B][n][n][n][11]")
        return x

    def forward(self, x):
        x = self._x(x)
        y = np.zeros((self.B, self.n, self.n, self.n, 3), dtype=np.float32)
        _check(_lib.wlc_unet_forward(self.h, _fp(x), _fp(y)))
        return y

    def gradients(self, x, t):
        x = self._x(x)
        t = np.ascontiguousarray(t, dtype=np.float32)
        loss = ctypes.c_float()
        _check(_lib.wlc_unet_gradients(self.h, _fp(x), _fp(t), ctypes.byref(loss)))
        return loss.value

    def adam(self, lr):
        _check(_lib.wlc_unet_adam(self.h, float(lr)))

    def train_step(self, x, t, lr):
        x = self._x(x)
        t = np.ascontiguousarray(t, dtype=np.float32)
        loss = ctypes.c_float()
        _check(_lib.wlc_unet_train_step(self.h, _fp(x), _fp(t), float(lr), ctypes.byref(loss)))
        return loss.value

    def params(self):
        p = np.zeros(self.np, dtype=np.float32)
        _check(_lib.wlc_unet_get(self.h, 0, _fp(p)))
        return p

    def grads(self):
        g = np.zeros(self.np, dtype=np.float32)
        _check(_lib.wlc_unet_get(self.h, 1, _fp(g)))
        return g

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        _check(_lib.wlc_unet_set_params(self.h, _fp(p)))


def test_conv3d(x, w, b, dy):
    """One device convolution and its gradients (wlc_test_conv3d): x [B,D,H,W,cin],
    w [cout,27,cin], b [cout], dy [B,D,H,W,cout] -> y, dx, dw, db (float32)."""
    load()
    B, D, H, W, cin = x.shape
    cout = w.shape[0]
    f = lambda a: np.ascontiguousarray(a, dtype=np.float32)
    x, w, b, dy = f(x), f(w), f(b), f(dy)
    y = np.zeros((B, D, H, W, cout), np.float32)
    dx = np.zeros_like(x)
    dw = np.zeros_like(w)
    db = np.zeros_like(b)
    _check(_lib.wlc_test_conv3d(B, D, H, W, cin, cout, _fp(x), _fp(w), _fp(b), _fp(dy), _fp(y), _fp(dx),
                                _fp(dw), _fp(db)))
    return y, dx, dw, db


def test_solve(L, kvecs, cdelta, sigma, ds, dslam, v):
    """k-aligned solve of G(k) x = v on the GPU for each row of kvecs."""
    load()
    kv = np.ascontiguousarray(kvecs, dtype=np.float64)
    nk = kv.shape[0]
    M = (L + 1) ** 2
    buf = np.ascontiguousarray(np.stack([np.real(v), np.imag(v)], -1).reshape(nk, M, 2))
    _check(_lib.wlc_test_solve(L, nk, _dp(kv), cdelta, sigma, ds, dslam, _dp(buf)))
    return buf[..., 0] + 1j * buf[..., 1]
