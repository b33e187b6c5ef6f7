"""One contour step on caller device buffers (wlc_mde_step) against the oracle's step
(oracle/wlc_oracle.c orc_step, the restatement of one iteration of MDE_q / MDE_qstar's
contour loop, wormlike_MDE_2dcomp_mod.f90:323-455, :584-650, :1035-1110): every BDF order,
both directions, a junction where the basis grows and one where it shrinks."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CHAIN = dict(shape=(8, 12, 16), box=(3.0, 2.5, 4.0), Nblk=(5, 4, 7), L=(3, 2, 4),
             lam=(1.0, 0.5, 2.0))




def _history(orc, w, direction, s, rng):
    """Random history slices with the harmonics of their own blocks."""
    Ns = orc.Ns
    order = min(s, 3) if direction > 0 else min(Ns - s, 3)
    hist = []
    for j in range(3):
        sh = s - direction * (j + 1)
        if j < order:
            hist.append(rng.standard_normal((orc.slice_M(sh),) + orc.shape))
        else:
            hist.append(None)
    return hist


@pytest.mark.parametrize("direction,s", [(+1, 1), (+1, 2), (+1, 3), (+1, 6), (+1, 5), (+1, 10),
                                         (+1, 16), (-1, 15), (-1, 14), (-1, 13), (-1, 9),
                                         (-1, 5), (-1, 0)])
def test_step_matches_oracle(wlc, oracle, direction, s):
    from hipmem import DeviceArray
    rng = np.random.default_rng(100 + 17 * s + direction)
    c = CHAIN
    w = rng.standard_normal((3,) + c["shape"])
    orc = oracle.Oracle(c["shape"], c["box"], c["Nblk"], c["L"], c["lam"])
    hist = _history(orc, w, direction, s, rng)
    ref = orc.step(w, direction, s, hist)
    orc.close()
    g = wlc.WormlikeMDE(c["shape"], c["box"], c["Nblk"], c["L"], c["lam"])
    dh = [DeviceArray(h) if h is not None else None for h in hist]
    dw = DeviceArray(w)
    out = DeviceArray(shape=ref.shape)
    out.fill_nan()
    g.step_device(direction, s, [d.ptr if d is not None else None for d in dh], dw.ptr, out.ptr)
    g.sync()
    got = out.get()
    g.close()
    assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()
