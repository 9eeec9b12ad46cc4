"""Pins the CPU oracle (oracle/irl_oracle.c) with checks that do not restate
its own formulas: closed forms, independent library routines (scipy),
invariants, and compositions of separately pinned pieces.  The reference
ships no test vectors and cannot be built here (its libmip dependency is
absent), so these are what fix the oracle."""
import numpy as np
import pytest
from scipy.signal import convolve2d

from oracle import oracle as O
from tests.cases import Problem, rel_l2


def circle_extents(n):
    """Row extents of the inscribed circle (rot3dpar.c:98-101), by the
    geometric definition: a row keeps the pixels whose centres lie within
    half-pixel-rounded chord length of the circle of radius n/2."""
    head, num = [], []
    R = n / 2.0
    for y in range(n):
        half = int(np.sqrt(R * R - (R - y - 0.5) ** 2) + 0.5)
        num.append(2 * half)
        head.append((n - 2 * half) // 2)
    return np.array(head), np.array(num)


def inside_mask(n):
    head, num = circle_extents(n)
    m = np.zeros((n, n), bool)
    for y in range(n):
        m[y, head[y]:head[y] + num[y]] = True
    return m


# ----------------------------------------------------------------- rotation
def test_rot_angle_zero_is_identity_inside_circle():
    rng = np.random.default_rng(0)
    n, nz = 18, 3
    v = rng.random((n, nz, n))
    r = O.rot3dpar(0.0, v, 0, nz, n, n)
    m = inside_mask(n)[:, None, :].repeat(nz, 1)
    assert np.array_equal(r[m], v[m])
    assert np.all(r[~m] == 0)


def test_rot_quarter_turn_matches_rot90():
    # angle pi/2 maps output (Y, X) to input (OldY, OldX) = (X, n-1-Y) up to the
    # 4e-8 residual of float(cos(pi/2)); pins direction and centre.
    rng = np.random.default_rng(1)
    n, nz = 16, 2
    v = rng.integers(0, 10, size=(n, nz, n)).astype(np.float64)
    r = O.rot3dpar(np.pi / 2, v, 0, nz, n, n)
    exp = np.empty_like(v)
    for Y in range(n):
        for X in range(n):
            exp[Y, :, X] = v[X, :, n - 1 - Y]
    m = inside_mask(n)
    assert np.abs(r - exp)[m[:, None, :].repeat(nz, 1)].max() < 1e-4


@pytest.mark.parametrize("angle", [0.37, 1.9, -2.6, 4.0])
def test_rot_reproduces_linear_functions(angle):
    # Bilinear interpolation is exact on linear functions: inside the circle,
    # away from the border taps, output = f(rotated coordinates).
    n, nz = 24, 1
    a, b, c = 0.7, -1.3, 5.0
    yy, xx = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    f = (a * xx + b * yy + c)[:, None, :]
    r = O.rot3dpar(angle, f, 0, nz, n, n)[:, 0, :]
    ax = (n - 1) / 2.0
    s, co = np.sin(np.float32(angle)), np.cos(np.float32(angle))
    X, Y = xx - ax, yy - ax
    fx = X * co - Y * s + ax
    fy = X * s + Y * co + ax
    ok = inside_mask(n) & (fx >= 0) & (fy >= 0) & (fx <= n - 2) & (fy <= n - 2)
    assert ok.sum() > 100
    assert np.abs(r - (a * fx + b * fy + c))[ok].max() < 1e-4


# ------------------------------------------------------------- convolution
def test_conv_matches_scipy_same_with_asymmetric_psf():
    rng = np.random.default_rng(2)
    a = rng.random((21, 26))
    p = rng.random((7, 5)).astype(np.float32)
    np.testing.assert_allclose(O.conv2d_ypad(a, p), convolve2d(a, p.astype(np.float64), "same"),
                               rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("pad,avg", [(2, 1), (3, 2), (5, 4), (9, 3), (2, 40)])
def test_conv_axial_padding_closed_form(pad, avg):
    # pad2d.c semantics: p = min(pad, cy) rows on each end hold the mean of
    # the first / last min(avg, ny) rows; x is zero padded.
    rng = np.random.default_rng(3)
    ny, nx = 12, 10
    a = rng.random((ny, nx))
    psf = rng.random((9, 5)).astype(np.float32)
    cy = 4
    p_eff = min(pad, cy)
    av = min(avg, ny)
    top = a[:av].mean(0)
    bot = a[ny - av:].mean(0)
    big = np.vstack([np.repeat(top[None], p_eff, 0), a, np.repeat(bot[None], p_eff, 0)])
    full = convolve2d(big, psf.astype(np.float64), "same")[p_eff:p_eff + ny]
    np.testing.assert_allclose(O.conv2d_ypad(a, psf, pad, avg), full, rtol=1e-12, atol=1e-12)


def test_delta_psf_is_identity():
    rng = np.random.default_rng(4)
    a = rng.random((9, 11))
    d = np.zeros((5, 7), np.float32)
    d[2, 3] = 1
    assert np.array_equal(O.conv2d_ypad(a, d), a)


# ----------------------------------------------------------------- subsets
def test_subset_order_closed_form():
    s = O.subsets(120, 15)
    assert s.shape == (8, 15)
    assert list(s[0]) == list(range(0, 120, 8))
    assert list(s[1]) == list(range(4, 120, 8))
    assert list(s[2]) == list(range(1, 120, 8))
    assert list(s[3]) == list(range(5, 120, 8))
    assert sorted(s.reshape(-1)) == list(range(120))
    assert O.subsets(60, 15).tolist()[1][0] == 2
    assert O.subsets(60, 7) == -1     # 60 % 7 != 0
    assert O.subsets(60, 20) == -2    # 3 subsets: odd
    assert O.subsets(60, 60).tolist() == [list(range(60))]


# --------------------------------------------------------------- projector
def test_drf_plane_index_closed_form():
    # start = pw/2, spacing = pw: row r lies at CFCR - N*pw/2 + (r+1/2)pw from
    # the face, which is table entry round(CFCR/pw - N/2 + r).
    pw, n = 0.5, 32
    drf = O.DrfTable(pw / 2, pw, [(2, 2)] * 200, [np.ones((3, 3), np.float32)] * 200)
    L = O.lib("f64")
    for cfcr in (10.0, 12.25, 30.0):
        for r in range(n):
            k = L.orc_drf_index(drf.handle("f64"), cfcr, pw, n, r)
            dist = cfcr - n * pw / 2 + (r + 0.5) * pw
            exp = -1 if dist <= 0 else min(199, max(0, int(np.floor(cfcr / pw - n / 2 + r + 0.5))))
            assert k == exp


def test_project_uniform_attenuation_closed_form():
    # x = 1 and mu = m everywhere, near-zero rotation: each bin's column is a
    # run of K in-circle rows, so p = sum_k exp(-m (k + 1/2)).
    n, nz, m = 20, 2, 0.13
    x = np.ones((n, nz, n))
    mu = np.full((n, nz, n), m)
    O.init_prj(nz, n, n, 7)
    # fAngle = -pi/2 makes the rotation angle 0.5*pi + fAngle ~ -1e-8
    p = O.project(nz, n, n, 1, 0.5, np.float32(-np.pi / 2), 20.0, None, mu, x)
    inside = inside_mask(n)
    K = inside.sum(0)  # in-circle rows per bin column
    exp = np.array([np.exp(-m * (np.arange(k) + 0.5)).sum() for k in K])
    assert np.abs(p - exp[None, :]).max() < 1e-4


def test_project_no_model_is_row_sum_of_rotation():
    rng = np.random.default_rng(5)
    n, nz = 16, 3
    x = rng.random((n, nz, n))
    ang = np.float32(0.8)
    O.init_prj(nz, n, n, 7)
    p = O.project(nz, n, n, 0, 0.5, ang, 12.0, None, None, x)
    rot = O.rot3dpar(np.float32(0.5 * np.pi + float(ang)), x, 0, nz, n, n)
    np.testing.assert_allclose(p, rot.sum(0), rtol=1e-12, atol=1e-12)


def _gauss_table(pw, n, half=3):
    from libirl_amd import synth
    offs, psfs = synth.gaussian_drf_table(pw, pw / 2, 80, pw, half)
    return O.DrfTable(pw / 2, pw, offs, psfs), psfs


def test_project_atn_drf_composition():
    # Project == sum over rows of scipy-convolved (rotated x * exp(-S)) with
    # the plane's table entry; rotation pinned above, convolution by scipy.
    rng = np.random.default_rng(6)
    n, nz, pw, cfcr = 16, 6, 0.5, 9.0
    x = rng.random((n, nz, n))
    mu = rng.random((n, nz, n)) * 0.1
    drf, psfs = _gauss_table(pw, n)
    ang = np.float32(-0.9)
    O.init_prj(nz, n, n, 7)
    p = O.project(nz, n, n, 3, pw, ang, cfcr, drf, mu, x)
    a = np.float32(0.5 * np.pi + float(ang))
    rx, rm = O.rot3dpar(a, x, 0, nz, n, n), O.rot3dpar(a, mu, 0, nz, n, n)
    L = O.lib("f64")
    S = np.zeros((nz, n))
    exp = np.zeros((nz, n))
    for r in range(n):
        S = S + 0.5 * rm[r]
        plane = rx[r] * np.exp(-S)
        k = L.orc_drf_index(drf.handle("f64"), cfcr, pw, n, r)
        if k >= 0:
            plane = convolve2d(plane, psfs[k].astype(np.float64), "same")
        exp += plane
        S = S + 0.5 * rm[r]
    np.testing.assert_allclose(p, exp, rtol=1e-10, atol=1e-12)


def test_backproject_composition_and_clamp():
    rng = np.random.default_rng(7)
    n, nz, pw, cfcr = 16, 6, 0.5, 9.0
    ratio = rng.random((nz, n))
    mu = rng.random((n, nz, n)) * 0.1
    drf, psfs = _gauss_table(pw, n)
    ang = np.float32(2.2)
    img0 = rng.random((n, nz, n)) * 2e-5  # some voxels fall under EPS after adding
    img0[0, 0, 0] = -1.0                  # forced below EPS
    O.init_prj(nz, n, n, 7)
    img = O.backproject(nz, n, n, 3, pw, ang, cfcr, drf, mu, ratio, img0.copy())
    a = np.float32(0.5 * np.pi + float(ang))
    rm = O.rot3dpar(a, mu, 0, nz, n, n)
    L = O.lib("f64")
    S = np.zeros((nz, n))
    B = np.zeros((n, nz, n))
    for r in range(n):
        S = S + 0.5 * rm[r]
        plane = ratio * np.exp(-S)
        k = L.orc_drf_index(drf.handle("f64"), cfcr, pw, n, r)
        if k >= 0:
            plane = convolve2d(plane, psfs[k].astype(np.float64), "same")
        B[r] = plane
        S = S + 0.5 * rm[r]
    back = O.rot3dpar(np.float32(-float(ang) - 0.5 * np.pi), B, 1, nz, n, n)
    exp = img0 + back
    exp[exp < 1e-5] = 0
    np.testing.assert_allclose(img, exp, rtol=1e-10, atol=1e-14)


# -------------------------------------------------------------------- OSEM
def test_osem_noise_free_fixed_point():
    # data = Project(x*) with the same model => the ratio is 1 wherever the
    # projection is positive, U equals the sensitivity sum and x* is
    # stationary.
    pb = Problem(n=24, nz=8, nviews=8, per=2, iters=1, noise=False, half=3)
    x_star = pb.act
    pb.x0 = x_star.copy()
    x = pb.oracle_osem("f64")
    inside = np.broadcast_to(inside_mask(24)[None], x.shape)
    # data are stored as float32, so the ratio is 1 to within float rounding
    assert rel_l2(x[inside], x_star[inside]) < 1e-6


def test_osem_f32_matches_f64():
    pb = Problem(n=24, nz=8, nviews=8, per=2, iters=2, half=3)
    assert rel_l2(pb.oracle_osem("f32"), pb.oracle_osem("f64")) < 1e-5


def test_osem_single_subset_reduces_to_mlem_update():
    # One subset holding every view is MLEM: x1 = x0 * BP(meas/P x0) / BP(1),
    # composed from the pinned projector and backprojector.
    pb = Problem(n=16, nz=4, nviews=4, per=4, iters=1, half=2)
    n, nz = pb.n, pb.nz
    x0 = O.reorder(nz, n, n, pb.x0)
    O.init_prj(nz, n, n, 7)
    U = np.zeros((n, nz, n))
    Nn = np.zeros((n, nz, n))
    for v, vw in enumerate(pb.views):
        O.backproject(nz, n, n, 3, pb.pix, vw.Angle, vw.CFCR, pb.drf, pb.atn_int,
                      np.ones((nz, n)), Nn)
    for v, vw in enumerate(pb.views):
        p = O.project(nz, n, n, 3, pb.pix, vw.Angle, vw.CFCR, pb.drf, pb.atn_int, x0)
        r = np.where(p > 1e-20, pb.prj[v] / np.where(p > 1e-20, p, 1), 0)
        O.backproject(nz, n, n, 3, pb.pix, vw.Angle, vw.CFCR, pb.drf, pb.atn_int, r, U)
    exp = x0 * U * np.where(Nn < 1e-6, 0, 1 / np.where(Nn < 1e-6, 1, Nn))
    exp_ext = O.reorder(n, nz, n, exp)
    got_ext = pb.oracle_osem("f64")
    assert rel_l2(got_ext, exp_ext) < 1e-12
