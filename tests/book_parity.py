"""Tolerance for the book's per-replica Greek sums against the book oracle.

Pathwise Greeks jump where the payoff's derivative does (a strike or barrier
crossing, a tie of the lookback's running minimum), so a point that sits on such
an edge to FP32 rounding can take the other side on the GPU.  A replica's sum may
therefore differ from the oracle's by the jump of one or two points, never more:
  * replicas within rounding (2e-4 per point + 1e-4 relative) are equal;
  * at most two replicas per Greek differ beyond rounding;
  * such a difference is at most twice the largest per-point Greek of that
    replica over the range (one flipped point moves the sum by at most the
    point's fine-minus-coarse Greek on either side), the per-point values
    taken from the oracle evaluating every point of the range on its own.
"""
import numpy as np


def check_greek_sums(got, ref, n, per_point):
    """got, ref: [R][3] sums over n points; per_point(i) -> [R][4] oracle sums of
    the single point first + i (price difference, then the three Greeks)."""
    got, ref = np.asarray(got, float), np.asarray(ref, float)
    d = np.abs(got - ref)
    bad = d > 2e-4 * n + 1e-4 * np.abs(ref)
    assert bad.sum(axis=0).max() <= 2, (got, ref)
    if not bad.any():
        return
    pts = np.array([np.asarray(per_point(i), float)[:, 1:] for i in range(n)])   # [n][R][3]
    jump = 2.0 * np.abs(pts).max(axis=0)                                          # [R][3]
    for r, k in zip(*np.nonzero(bad)):
        assert d[r, k] <= 2.0 * jump[r, k], ("replica", r, "greek", k, d[r, k], jump[r, k], got, ref)
    # the oracle's points add up to its sums (the per-point evaluation is the same computation)
    assert np.allclose(pts.sum(axis=0), ref, rtol=1e-9, atol=1e-9 * n)
