"""wlc_mde_step_noalloc (no handle, no allocation: the scratch is the library's fixed device
array) against wlc_mde_step of a handle on the same inputs, on two calls in a row, the second
with a smaller chain and grid, so the second call runs on the array the first one left
behind; both also against the oracle's step (oracle/wlc_oracle.c orc_step)."""
import numpy as np
import pytest

CALLS = [
    # 16 x 24 x 32 grid, three blocks, a BDF3 step that crosses the A|B junction (Lmax 4 -> 3)
    dict(shape=(16, 24, 32), box=(3.0, 2.5, 4.0), Nblk=(6, 5, 7), L=(4, 3, 5), lam=(1.0, 0.5, 2.0),
         direction=+1, s=8),
    # smaller: 8 x 12 x 16, two blocks, backward BDF2 step
    dict(shape=(8, 12, 16), box=(2.0, 3.0, 2.5), Nblk=(5, 4), L=(2, 3), lam=(0.7, 1.5),
         direction=-1, s=7),
]


def _run(wlc, oracle, c, seed):
    from hipmem import DeviceArray, Stream
    rng = np.random.default_rng(seed)
    nb = len(c["Nblk"])
    w = rng.standard_normal((nb,) + c["shape"])
    orc = oracle.Oracle(c["shape"], c["box"], c["Nblk"], c["L"], c["lam"])
    d, s = c["direction"], c["s"]
    order = min(s, 3) if d > 0 else min(orc.Ns - s, 3)
    hist = [rng.standard_normal((orc.slice_M(s - d * (j + 1)),) + c["shape"]) if j < order else None
            for j in range(3)]
    ref = orc.step(w, d, s, hist)
    orc.close()
    dh = [DeviceArray(h) if h is not None else None for h in hist]
    ptrs = [t.ptr if t is not None else None for t in dh]
    dw = DeviceArray(w)
    out_na = DeviceArray(shape=ref.shape)
    out_h = DeviceArray(shape=ref.shape)
    out_na.fill_nan()
    out_h.fill_nan()
    st = Stream()
    p = wlc.make_params(c["shape"], c["box"], c["Nblk"], c["L"], c["lam"])
    wlc.step_noalloc(p, d, s, ptrs, dw.ptr, out_na.ptr, st.ptr)
    st.sync()
    st.close()
    g = wlc.WormlikeMDE(c["shape"], c["box"], c["Nblk"], c["L"], c["lam"])
    g.step_device(d, s, ptrs, dw.ptr, out_h.ptr)
    g.sync()
    g.close()
    return out_na.get(), out_h.get(), ref


@pytest.mark.gpu
def test_noalloc_step_matches_the_handle_step_twice(wlc, oracle):
    print("no-alloc step: two calls, 16x24x32 forward then 8x12x16 backward", flush=True)
    for k, c in enumerate(CALLS):
        na, hd, ref = _run(wlc, oracle, c, 50 + k)
        # same kernels on the same inputs: bit-identical to the handle's step
        assert np.array_equal(na, hd), f"call {k}"
        assert np.abs(na - ref).max() <= 1e-12 * np.abs(ref).max(), f"call {k}"


def test_noalloc_rejects_steps_beyond_the_scratch(wlc):
    """64^3 at Lmax 8: the half spectrum alone (191 MB) exceeds the 128 MiB array; the call
    fails with WLC_EINVAL before touching the device."""
    p = wlc.make_params((64, 64, 64), (5.0, 5.0, 5.0), (10, 10), (8, 8), (1.0, 1.0))
    with pytest.raises(wlc.WlcError) as e:
        wlc.step_noalloc(p, +1, 5, [1, 1, 1], 1, 1, None)
    assert e.value.code == -1
    assert "scratch" in str(e.value)
