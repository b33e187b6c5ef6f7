"""What the trained warm-start network gains over starting from the seed fields, at a size the
test suite can afford: the network is trained on the batched solver's converged fields of
draws at grid edge 16 (Lmax 4), then held-out draws (another seed) are solved from their seed
fields and from the network's proposal.  The proposal must reach the residual in fewer
iterations on average, converge at least as often, and give the same free energy (the same
phase) wherever both runs converge."""
import os
import sys
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "wormlike-chain-triblock_amd"))
import warmstart as WS  # noqa: E402

TOL = 1e-4
CAP = 320


def _log(m):
    print("# " + m, flush=True)


def test_warm_start_beats_seed_start_on_held_out_draws(wlc):
    t0 = time.perf_counter()
    data, info = WS.generate(wlc, 64, 21, (16,), (4,), chunk=64, max_iter=CAP, log=_log)
    assert 16 in data and len(data[16][0]) >= 24, info
    net, hist, _ = WS.train(wlc, data, 40, batch=4, lr=1e-3, seed=0, log=_log)
    assert hist[-1] < 0.2 * hist[0]
    cases = WS.draw_at(12, 22, (16,), (4,))
    init = WS.predict(net, cases)
    net.close()
    seed_res, _ = WS.solve(wlc, cases, max_iter=CAP, tol=TOL, log=_log)
    warm_res, _ = WS.solve(wlc, cases, init=init, max_iter=CAP, tol=TOL, log=_log)
    it_seed = np.array([r["iterations"] for r in seed_res], dtype=float)
    it_warm = np.array([r["iterations"] for r in warm_res], dtype=float)
    conv_seed = np.array([bool(r["converged"]) for r in seed_res])
    conv_warm = np.array([bool(r["converged"]) for r in warm_res])
    _log(f"seed start: {conv_seed.sum()} converged, mean {it_seed.mean():.1f} iterations; "
         f"warm start: {conv_warm.sum()} converged, mean {it_warm.mean():.1f} iterations; "
         f"{time.perf_counter() - t0:.1f} s")
    assert conv_warm.sum() >= conv_seed.sum()
    assert it_warm.mean() < 0.7 * it_seed.mean()
    both = conv_seed & conv_warm
    for s, w, ok in zip(seed_res, warm_res, both):
        if ok:
            # residual 1e-4 leaves F uncertain at about that level; a different phase differs by far more
            assert abs(s["F"] - w["F"]) < 2e-3, (s["F"], w["F"])
