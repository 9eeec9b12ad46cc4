"""Pinning of the oracle's three-shear rotation (fastrotate.c:330-589).

The reference ships no vectors for vShearsRotate, so the restatement is
checked against properties fixed by the mathematics and by the already
pinned bilinear rotation:
  * angle 0 is the identity on the circle trimmed by one pixel per row;
  * three linear-interpolation shears reproduce a linear function exactly,
    so inside the circle the result equals the linear function sampled at
    the rotated coordinates -- the same map as rot3dpar's (direction,
    centre, and each of the four 90/180 degree folds);
  * near multiples of 90 degrees the rotation is a pure index permutation,
    so an arbitrary (non-smooth) image agrees with rot3dpar there;
  * Project/BackProject with bFastRotate compose the pinned pieces.
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests.cases import rel_l2


def _inner(n, margin):
    c = (n - 1) / 2.0
    j, i = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    return (i - c) ** 2 + (j - c) ** 2 < (n / 2.0 - margin) ** 2


@pytest.mark.parametrize("real", ["f32", "f64"])
@pytest.mark.parametrize("n", [16, 32, 48])
def test_angle_zero_is_identity_on_trimmed_circle(real, n):
    rng = np.random.default_rng(0)
    nz = 3
    v = rng.random((n, nz, n))
    out = O.shears_rotate(0.0, v, nz, n, real=real)
    s, e = O.shears_extents(n)
    for j in range(n):
        keep = np.zeros(n, bool)
        keep[s[j] + 1:e[j] - 1] = True
        want = np.where(keep[None, :], v[j].astype(np.float32 if real == "f32" else np.float64),
                        0.0)
        assert np.array_equal(out[j], want), j


def test_extents_closed_form():
    # iLen = 2 floor(sqrt(c^2 - d^2)) + 2 (fastrotate.c:584-586), centred
    for n in (8, 16, 31, 64):
        s, e = O.shears_extents(n)
        c = (n - 1) / 2.0
        for j in range(n):
            ln = 2 * int(np.floor(np.sqrt(c * c - (j - c) ** 2))) + 2
            assert s[j] == int((n - ln) / 2) and e[j] - s[j] == ln   # C division truncates


# angles exercising every fold: (b90, b180) = (0,0), (1,0), (0,1), (1,1)
FOLD_ANGLES = [0.3, -0.6, 1.2, 2.0, 2.6, 3.5, -2.5, 4.4, -1.3, 5.9, 0.5 * np.pi + 1.0]


@pytest.mark.parametrize("angle", FOLD_ANGLES)
def test_linear_function_rotates_exactly(angle):
    n, nz = 48, 2
    c = (n - 1) / 2.0
    j, i = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    coef = [(0.7, -0.4, 3.0), (-0.2, 0.9, 5.0)]
    vol = np.zeros((n, nz, n))
    for z, (a, b, d) in enumerate(coef):
        vol[:, z, :] = a * i + b * j + d
    out = O.shears_rotate(angle, vol, nz, n, real="f64")
    sn, cs = np.float32(np.sin(np.float32(angle))), np.float32(np.cos(np.float32(angle)))
    X, Y = i - c, j - c
    fx, fy = X * cs - Y * sn + c, X * sn + Y * cs + c
    m = _inner(n, 3)
    for z, (a, b, d) in enumerate(coef):
        want = a * fx + b * fy + d
        err = np.abs(out[:, z, :] - want)[m].max()
        assert err < 1e-4 * np.abs(want[m]).max(), (angle, z, err)


@pytest.mark.parametrize("k", [1, 2, 3, -1, -2, 4])
def test_quarter_turns_agree_with_bilinear_rotation(k):
    # arbitrary image: only an index permutation can agree with rot3dpar
    n, nz = 32, 3
    rng = np.random.default_rng(k + 10)
    v = rng.random((n, nz, n))
    ang = k * 0.5 * np.pi
    a = O.shears_rotate(ang, v, nz, n, real="f64")
    b = O.rot3dpar(ang, v, 0, nz, n, n, real="f64")
    m = _inner(n, 2)
    for z in range(nz):
        assert np.abs(a[:, z, :] - b[:, z, :])[m].max() < 1e-4, z


def test_shear_and_bilinear_agree_on_smooth_images():
    n, nz = 64, 2
    c = (n - 1) / 2.0
    j, i = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    img = np.exp(-((i - c - 5) ** 2 + (j - c + 3) ** 2) / 80.0)
    v = np.repeat(img[:, None, :], nz, axis=1)
    m = _inner(n, 3)
    for ang in (0.4, 1.9, -2.7, 4.0):
        a = O.shears_rotate(ang, v, nz, n, real="f64")
        b = O.rot3dpar(ang, v, 0, nz, n, n, real="f64")
        assert rel_l2(a[:, 0, :][m], b[:, 0, :][m]) < 1e-2, ang


def test_project_and_backproject_compose_with_the_shear_rotation():
    n, nz, pw = 24, 4, 0.5
    rng = np.random.default_rng(2)
    x = rng.random((n, nz, n))
    O.init_prj(nz, n, n, 7, real="f64")
    ang = 0.9
    p = O.project(nz, n, n, 0, pw, ang, 10.0, None, None, x, real="f64", fast=1)
    rot = O.shears_rotate(0.5 * np.pi + np.float32(ang), x, nz, n, real="f64")
    assert np.allclose(p, rot.sum(axis=0), rtol=1e-12, atol=1e-12)
    ratio = rng.random((nz, n))
    img = np.zeros((n, nz, n))
    O.backproject(nz, n, n, 0, pw, ang, 10.0, None, None, ratio, img, real="f64", fast=1)
    back = O.shears_rotate(-0.5 * np.pi - np.float32(ang), np.broadcast_to(ratio, (n, nz, n)),
                           nz, n, real="f64")
    back = np.where(back < 1e-5, 0.0, back)
    assert np.allclose(img, back, rtol=1e-12, atol=1e-12)
