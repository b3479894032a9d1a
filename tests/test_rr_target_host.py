"""Host logic of RR_target's speculative bisection (no GPU): with the RR it
calls routed to the oracle, every tree depth gives the sequential bisection's
result -- the same threshold, the same RR bits -- on searches that stop by
the epsilon test, by the interval test and by running out of iterations.
Also the argument checks the reference makes."""
import numpy as np
import pytest

import accrqa_amd
from pyoracle import rr_target


@pytest.fixture
def on_oracle(monkeypatch, oracle):
    calls = []

    def rr(x, taus, embs, thr, dist):
        calls.append(len(thr))
        out, e = oracle.RR(x, taus, embs, np.asarray(thr, dtype=x.dtype), dist)
        assert e == 0
        return out.reshape(len(taus), len(embs), len(thr))

    monkeypatch.setattr(accrqa_amd, "RR", rr)
    return calls


def test_speculative_walk_equals_sequential(on_oracle, oracle, capsys):
    rng = np.random.default_rng(61)
    x = np.cumsum(rng.standard_normal(500)).astype(np.float32)
    x = (x - x.mean()) / x.std()
    cases = [(0.1, 1, 0.01, 20, 0, 10), (0.35, 2, 0.001, 20, 0, 10), (0.05, 1, 1e-7, 9, 0, 4),
             (0.5, 2, 0.02, 20, 0.5, 3.0), (0.2, 1, 0.01, 20, 5, 10), (0.9, 1, 1e-6, 13, 0, 10)]
    for target, dist, eps, it, lo, hi in cases:
        want_t, want_rr, seen = rr_target(oracle, x, 2, 3, target, dist, eps, it, lo, hi)
        for depth in (1, 2, 3, 5, 7):
            on_oracle.clear()
            got_t, got_rr = accrqa_amd.RR_target(x, 2, 3, target, dist, eps, it, lo, hi, depth=depth)
            assert got_t == want_t and got_rr.tobytes() == want_rr.tobytes(), (target, dist, eps, depth)
            # one pass for the bounds, then one per `depth` levels reached
            assert len(on_oracle) == 1 + -(-len(seen) // depth), (len(on_oracle), len(seen), depth)
    out = capsys.readouterr().out
    assert "Warning: Bisection search stopped after 9 iterations" in out


def test_argument_checks():
    x = np.zeros(100, np.float32)
    for kw in ({"tau": 0}, {"emb": 0}, {"target_RR": 1.5}, {"epsilon": 0.0}, {"max_iter": 0},
               {"threshold_min": 2.0, "threshold_max": 1.0}):
        args = dict(input_data=x, tau=1, emb=2, target_RR=0.2)
        args.update(kw)
        with pytest.raises(TypeError):
            accrqa_amd.RR_target(**args)
    with pytest.raises(TypeError):
        accrqa_amd.RR_target(list(x), 1, 2, 0.2)
    with pytest.raises(TypeError):
        accrqa_amd.RR_target(x.astype(np.float64), 1, 2, 0.2)
