"""Computed parallel-hole GRF tables (gendrf.c:221-427): host code of the
library, checked without a GPU against properties the collimator model
fixes.  The transfer function is (2 J1(pi a d f)/(pi a d f))^2 (hole
geometry, a = (z + gap + L)/L) times the intrinsic Gaussian and the bin
sampling sinc, so the per-axis second moment of the response is
(a d)^2/8 + (FWHM/2.3548)^2 + bin^2/12.  The reference's own tables come
from libmip's FFT, which is not available: the values are pinned by these
moments, symmetry, normalisation and the truncation rule, not bit for bit."""
import numpy as np
import pytest

import libirl_amd as L

LEHR = dict(col_thick=2.405, hole_diam=0.111, gap=0.0, intrinsic=0.32)


def _table(max_err, dists, bin_size=0.22, noffs=24, **kw):
    c = dict(LEHR)
    c.update(kw)
    return L.gendrf(c["col_thick"], c["hole_diam"], c["gap"], c["intrinsic"], max_err, bin_size,
                    1.0, dists[0], len(dists), noffs)


def test_normalised_and_symmetric():
    for p in _table(0.05, [1.0, 5.0, 10.0, 20.0]):
        assert abs(p.sum(dtype=np.float64) - 1.0) < 1e-5
        assert np.array_equal(p, p[::-1, :]) and np.array_equal(p, p[:, ::-1])
        assert np.array_equal(p, p.T)


@pytest.mark.parametrize("dist", [4.0, 8.0, 15.0, 25.0])
def test_second_moment_matches_the_collimator_model(dist):
    b = 0.22
    p = _table(0.0, [dist], bin_size=b, noffs=40)[0].astype(np.float64)
    h = (p.shape[1] - 1) // 2
    x = (np.arange(p.shape[1]) - h) * b
    var = float((p.sum(axis=0) * x * x).sum())
    a = (dist + LEHR["gap"] + LEHR["col_thick"]) / LEHR["col_thick"]
    want = (a * LEHR["hole_diam"]) ** 2 / 8 + (LEHR["intrinsic"] / 2.3548) ** 2 + b * b / 12
    assert abs(var / want - 1) < 0.04, (var, want)


def test_gap_enters_like_distance():
    p1 = _table(0.0, [10.0], noffs=30, gap=2.0)[0]
    p2 = _table(0.0, [12.0], noffs=30, gap=0.0)[0]
    assert np.allclose(p1, p2, atol=1e-6)


def test_truncation_keeps_the_smallest_square():
    full = _table(0.0, [12.0], noffs=24)[0].astype(np.float64)
    cut = _table(0.05, [12.0], noffs=24)[0]
    h, k = (full.shape[0] - 1) // 2, (cut.shape[0] - 1) // 2
    area = lambda m: full[h - m:h + m + 1, h - m:h + m + 1].sum()
    assert area(k) >= 0.95 * full.sum() - 1e-9
    assert area(k - 1) <= 0.95 * full.sum() + 1e-9
    assert abs(cut.sum(dtype=np.float64) - 1.0) < 1e-5


def test_non_positive_distance_is_a_delta():
    for d in (-0.5, 0.0):
        q = _table(0.05, [d])[0]
        assert q.shape == (1, 1) and q[0, 0] == 1.0
