"""Comparison helpers for the parity tests."""
import numpy as np

RTOL = 1e-12  # FP64 parity bar: relative 1e-12 against the oracle


def assert_bitwise(gpu, ref, what):
    g, r = np.asarray(gpu), np.asarray(ref)
    assert g.shape == r.shape, what
    same = (g == r) | (np.isnan(g) & np.isnan(r)) if g.dtype.kind == "f" else (g == r)
    n_bad = int((~same).sum())
    assert n_bad == 0, "%s: %d of %d entries differ (first at %s)" % (
        what, n_bad, g.size, np.argwhere(~same)[:3].tolist())


def rel_err(gpu, ref, scale=None):
    """max |g-r| / max(|r|, scale), over entries that are finite in the reference."""
    g, r = np.asarray(gpu, dtype=np.float64), np.asarray(ref, dtype=np.float64)
    fin = np.isfinite(r)
    d = np.abs(g[fin] - r[fin])
    den = np.abs(r[fin])
    if scale is not None:
        den = np.maximum(den, np.broadcast_to(scale, r.shape)[fin])
    with np.errstate(invalid="ignore", divide="ignore"):
        e = np.where(den > 0, d / den, np.where(d == 0, 0.0, np.inf))
    return float(e.max()) if e.size else 0.0


def assert_close(gpu, ref, what, rtol=RTOL, scale=None):
    """Same NaN/Inf mask, then |g-r| <= rtol * max(|r|, scale) elementwise.

    ``scale`` is the magnitude the value was summed from (e.g. the projection of
    |f|), so a node where a mixed-sign sum cancels is judged against the size of
    its terms (the mixed criterion for cancelling sums)."""
    g, r = np.asarray(gpu, dtype=np.float64), np.asarray(ref, dtype=np.float64)
    assert g.shape == r.shape, what
    nm_g, nm_r = ~np.isfinite(g), ~np.isfinite(r)
    assert np.array_equal(nm_g, nm_r), "%s: non-finite masks differ (%d vs %d)" % (
        what, nm_g.sum(), nm_r.sum())
    assert np.array_equal(g[nm_r & np.isinf(r)], r[nm_r & np.isinf(r)]), what + ": inf values differ"
    e = rel_err(g, r, scale)
    assert e <= rtol, "%s: max relative error %.3e > %.1e" % (what, e, rtol)
    return e
