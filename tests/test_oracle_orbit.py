"""Pinning of the oracle's ComputeOrbit (computeorbit.c:21-150) by geometry:
a centred disk of radius R gives R + margin at every angle (to a pixel), an
off-centre disk follows its centre's projection on the detector axis, the
L-mode pairing takes the larger radius of views v and v+V/2, the threshold
zeroes pixels in place, and an unknown orbit type returns 2."""
import numpy as np
import pytest

from oracle import oracle as O


def _disk(n, r, cx=0.0, cy=0.0, val=1.0):
    c = (n - 1) / 2.0
    j, i = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    return np.where((i - c - cx) ** 2 + (j - c - cy) ** 2 <= r * r, val, 0.0).astype(np.float32)


ANGLES = (-2 * np.pi * np.arange(24) / 24).astype(np.float32)


def test_circular_orbit_is_the_margin():
    img = _disk(32, 8)
    rc, cf = O.compute_orbit(img, 0.4, 25.0, ANGLES, 0.1, 0)
    assert rc == 0 and np.all(cf == np.float32(25.0))


def test_unknown_orbit_type():
    rc, _ = O.compute_orbit(_disk(32, 8), 0.4, 2.0, ANGLES, 0.1, 3)
    assert rc == 2


@pytest.mark.parametrize("n,r", [(32, 8), (64, 20), (48, 5)])
def test_centred_disk_radius(n, r):
    pix, margin = 0.5, 3.0
    rc, cf = O.compute_orbit(_disk(n, r), pix, margin, ANGLES, 0.5, 1)
    assert rc == 0
    assert np.all(np.abs(cf - (r * pix + margin)) <= 1.5 * pix), cf


def test_off_centre_disk_follows_the_centre():
    n, r, pix, margin = 64, 12, 0.5, 2.0
    cx, cy = 6.0, -4.0          # offset below the radius: the edge never passes the centre
    rc, cf = O.compute_orbit(_disk(n, r, cx, cy), pix, margin, ANGLES, 0.5, 1)
    assert rc == 0
    # radius varies with the angle by the centre's offset along the viewing
    # direction; the spread equals twice the offset length (the search stops
    # at the centre row, so an edge beyond the centre would clamp)
    spread = (cf.max() - cf.min()) / pix
    assert abs(spread - 2 * np.hypot(cx, cy)) <= 2.5, spread
    # opposite views see the centre on opposite sides: their radii sum to 2R
    opp = cf[:12] + cf[12:]
    assert np.all(np.abs(opp - 2 * (r * pix + margin)) <= 2 * 1.5 * pix), opp


def test_lmode_pairs_take_the_larger_radius():
    n, pix = 64, 0.5
    img = _disk(n, 8, 12.0, 0.0)
    _, c1 = O.compute_orbit(img.copy(), pix, 2.0, ANGLES, 0.5, 1)
    _, c2 = O.compute_orbit(img.copy(), pix, 2.0, ANGLES, 0.5, 2)
    h = len(ANGLES) // 2
    want = np.maximum(c1[:h], c1[h:])
    assert np.array_equal(c2[:h], want) and np.array_equal(c2[h:], want)


def test_threshold_zeroes_in_place_and_per_view_margins():
    img = _disk(32, 8, val=2.0) + np.float32(0.1)
    _, cf = O.compute_orbit(img, 0.4, 0.0, ANGLES, -0.5, 1,
                            margins=np.arange(len(ANGLES), dtype=np.float32))
    assert img.min() == 0.0 and np.all((img == 0) | (img == np.float32(2.1)))
    base = cf - np.arange(len(ANGLES), dtype=np.float32)
    assert np.all(np.abs(base - 8 * 0.4) <= 1.5 * 0.4)
