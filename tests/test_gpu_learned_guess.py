"""The learned Stokes guess end to end at a small size: data from the plain
method's solves, U-Net training by the GPU kernels, and the host Stokes solve (FGMRES +
field split) started from the plain, previous and learned guesses."""
import numpy as np
import pytest

from fvmic2d_amd import learned_guess as lg
from fvmic2d_amd.lib import Context

pytestmark = pytest.mark.gpu


def test_learned_guess_pipeline_small():
    print("CASE learned guess 33x33, 64 train / 6 test", flush=True)
    N = 33
    raw_tr = lg.generate(N, list(range(64)), workers=0)
    raw_te = lg.generate(N, list(range(1000, 1006)), workers=0)
    ctx = Context()
    tr, te = lg.Dataset(ctx, N, raw_tr), lg.Dataset(ctx, N, raw_te)
    X, T = tr.arrays()
    model = lg.LearnedGuess(ctx, levels=3, base=16, seed=0)
    hist = model.train(X, T, epochs=30, batch=8, lr=3e-3)
    assert hist[-1] < 0.5 * hist[0], hist
    # held-out states: the predicted dynamic pressure explains part of its variance
    Xt, Tt = te.arrays()
    pred = model.predict(Xt)
    assert np.mean((pred - Tt) ** 2) < np.mean(Tt ** 2)
    rows = lg.evaluate(te, model)
    s = lg.summarize(rows)
    print(s, flush=True)
    for r in rows:
        assert all(np.isfinite(v) for v in r["r0"].values())
        assert all(0 <= v <= 3000 for v in r["its"].values())
