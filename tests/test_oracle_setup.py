"""Pins the oracle's setup-side restatements with known answers worked out
by hand: the initial estimate (initial_est.c:29-180: average, circle,
contour and attenuation-map support) and the measured-bin masks of the norm
and ratio images (osemrecon.c:92-184)."""
import math

import numpy as np

from libirl_amd import PrjView
from oracle import oracle as O

# Circle of IrlOsem's initial estimate at N = 8 (initial_est.c:136-156): a
# pixel stays when the corner farthest from the centre (4, 4) lies within
# radius 4, except that row and column 4 use their near edge (fX = 0).
CIRCLE8 = ["....X...",
           "..XXXX..",
           ".XXXXXX.",
           ".XXXXXX.",
           "XXXXXXXX",
           ".XXXXXX.",
           "..XXXX..",
           "....X..."]


def _views(angles, cfcr):
    out = []
    for a in angles:
        v = PrjView()
        v.Angle, v.Left, v.Right, v.CFCR = a, 0.0, 0.0, cfcr
        out.append(v)
    return out


def test_average_and_circle():
    n, nz, V = 8, 3, 4
    prj = np.full((V, nz, n), 2.0, np.float32)
    x, warned = O.initial_estimate(n, nz, 0.5, _views([0.0] * V, 100.0), prj)
    assert not warned
    # sum(prj) / (N^2 Nz pi/4 V) = 4 p / (pi N) for constant projections p
    avg = 4 * 2.0 / (math.pi * n)
    mask = np.array([[c == "X" for c in row] for row in CIRCLE8])
    for s in range(nz):
        np.testing.assert_allclose(x[:, s, :], np.where(mask, avg, 0.0), rtol=1e-6)


def test_contour_zeroes_behind_each_detector_face():
    # radius 1 cm / 0.5 cm = 2 pixels.  Angle 0 (cos 1, sin 0): columns with
    # ix - 8 > 2 are behind the face; angle pi/2 (sin 1): rows iy - 8 > 2;
    # angle pi (cos -1): columns with 8 - ix > 2 (initial_est.c:160-180).
    n, nz = 16, 2
    prj = np.ones((1, nz, n), np.float32)
    for ang, zero in ((0.0, lambda r, c: c >= 11), (math.pi / 2, lambda r, c: r >= 11),
                      (math.pi, lambda r, c: c <= 5)):
        x, _ = O.initial_estimate(n, nz, 0.5, _views([ang], 1.0), prj, contour=True)
        full, _ = O.initial_estimate(n, nz, 0.5, _views([ang], 1.0), prj, contour=False)
        for r in range(n):
            for c in range(n):
                want = 0.0 if zero(r, c) else full[r, 0, c]
                assert x[r, 0, c] == want and x[r, 1, c] == want, (ang, r, c)
    # several views: the intersection of the half-planes
    x, _ = O.initial_estimate(n, nz, 0.5, _views([0.0, math.pi / 2, math.pi], 1.0),
                              np.ones((3, nz, n), np.float32), contour=True)
    nz_rows, nz_cols = np.nonzero(x[:, 0, :])
    assert nz_rows.max() == 10 and nz_cols.min() == 6 and nz_cols.max() == 10


def test_attenuation_map_support():
    # the reordered, not yet scaled map against thresh * pixel width
    n, nz, pw = 8, 2, 0.5
    atn = np.full((n, nz, n), 0.06, np.float32)
    atn[3, 1, 4] = 0.04
    atn[4, 0, 2] = 0.0499
    atn[2, 0, 3] = np.float32(0.1) * np.float32(pw)   # exactly the threshold: kept
    prj = np.ones((2, nz, n), np.float32)
    x, warned = O.initial_estimate(n, nz, pw, _views([0.0, 1.0], 100.0), prj, atn=atn,
                                   thresh=0.1)
    assert not warned
    ref, _ = O.initial_estimate(n, nz, pw, _views([0.0, 1.0], 100.0), prj)
    want = ref.copy()
    want[3, 1, 4] = 0.0
    want[4, 0, 2] = 0.0
    np.testing.assert_array_equal(x, want)
    # a negative threshold is reported and ignored (initial_est.c:87-95)
    x, warned = O.initial_estimate(n, nz, pw, _views([0.0, 1.0], 100.0), prj, atn=atn,
                                   thresh=-0.1)
    assert warned
    np.testing.assert_array_equal(x, ref)


def test_measured_bin_masks():
    # 16 bins, Left = -3.5: pixel i covers bins floor(i - 3.5) .. floor(i - 2.5).
    # Norm mask: measured where the upper bin > 0 (i >= 4), the lower < 16,
    # and stop once the lower bin passes Right = 10.2 (i = 15).  Ratio mask:
    # the same lower limits, no Right limit (osemrecon.c:140-184).
    nz, n = 3, 16
    meas = np.ones((nz, n))
    meas[1, 6] = -1.0
    one = O.measured_to_one(meas, -3.5, 10.2)
    want = np.zeros((nz, n))
    want[:, 4:15] = 1.0
    want[1, 6] = 0.0
    np.testing.assert_array_equal(one, want)
    prj = np.full((nz, n), 5.0)
    z = O.nonmeasured_to_zero(prj, meas, -3.5)
    want = np.full((nz, n), 5.0)
    want[:, :4] = 0.0
    want[1, 6] = 0.0
    np.testing.assert_array_equal(z, want)
    # Left = 2.25: pixels 14 and 15 start at bins 16 and 17, past the last
    one = O.measured_to_one(np.ones((nz, n)), 2.25, 20.0)
    assert one[:, :14].all() and not one[:, 14:].any()
    z = O.nonmeasured_to_zero(prj, np.ones((nz, n)), 2.25)
    assert (z[:, :14] == 5.0).all() and not z[:, 14:].any()
    # full range: every bin measured
    assert O.measured_to_one(np.ones((nz, n)), 0.0, float(n)).all()
