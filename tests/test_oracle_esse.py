"""Pins the oracle's ESSE restatement (krnlprj3d.c) with closed forms: delta
kernels (kernel centre, depth direction, x mirroring, density factor),
collapse/expand normalisation, and the zero-scatter-factor switch."""
import numpy as np

from oracle import oracle as O

MU0, PIX = 0.15, 0.5
# The reference forms the density factor 1/(mu0 * pix) in float
# (krnlprj3d.c:709, 941), so results match the closed forms to float rounding.


def _delta_kernel(ndepths=9, noffs=3, depth_off=0, x_off=0):
    k = np.zeros((ndepths, noffs, noffs), np.float32)
    k[ndepths // 2 - 1 - depth_off, 0, x_off] = 1.0
    return k, np.full_like(k, MU0)


def test_centre_delta_gives_density_weighted_activity():
    rng = np.random.default_rng(0)
    n, nz = 16, 8
    act = rng.random((n, nz, n))
    mp = rng.random((n, nz, n)) * 0.1
    e = O.Esse(*_delta_kernel(), MU0, PIX, n, nz)
    # krnlMu == mu0 everywhere: muavg = mu0, the first-order kernel vanishes
    # and exp(d (mu0 - muavg)) = 1, leaving act * map / (mu0 * pix)
    np.testing.assert_allclose(e.effsrc(act, mp), act * mp / (MU0 * PIX), rtol=1e-6)


def test_shifted_delta_fixes_depth_direction_and_mirroring():
    rng = np.random.default_rng(1)
    n, nz = 16, 8
    act = rng.random((n, nz, n))
    mp = np.full((n, nz, n), 0.05)
    e = O.Esse(*_delta_kernel(depth_off=2, x_off=1), MU0, PIX, n, nz)
    shifted = np.zeros_like(act)
    shifted[2:, :, 1:] += act[:-2, :, :-1]   # depth +2, bin +1
    shifted[2:, :, :-1] += act[:-2, :, 1:]   # depth +2, bin -1 (mirrored quadrant)
    np.testing.assert_allclose(e.effsrc(act, mp), shifted * mp / (MU0 * PIX), rtol=1e-6,
                               atol=1e-14)


def test_collapse_expand_preserves_a_constant():
    n, nz, c0 = 16, 8, 0.7
    act = np.full((n, nz, n), c0)
    mp = np.full((n, nz, n), 0.05)
    for fac in (2, 4):
        e = O.Esse(*_delta_kernel(), MU0, PIX, n, nz, collapse=fac)
        np.testing.assert_allclose(e.effsrc(act, mp), act * mp / (MU0 * PIX), rtol=1e-6)


def test_depth_correction_closed_form():
    # uniform mu: muavg from a kernel whose mu kernel is m1 everywhere is m1,
    # so eff = act * exp(d (mu0 - m1)) * density with d = (k + 1/2) mu / mu0
    n, nz, m1, mu = 16, 4, 0.12, 0.06
    k, _ = _delta_kernel()
    km = np.full_like(k, m1)
    e = O.Esse(k, km, MU0, PIX, n, nz)
    act = np.ones((n, nz, n))
    mp = np.full((n, nz, n), mu)
    d = (np.arange(n) + 0.5) * mu / MU0
    exp = np.exp(d * (MU0 - m1))[:, None, None] * mu / (MU0 * PIX)
    np.testing.assert_allclose(e.effsrc(act, mp), np.broadcast_to(exp, act.shape), rtol=1e-6)


def test_zero_scatter_factor_disables_the_model():
    k, km = _delta_kernel()
    e = O.Esse(k, km, MU0, PIX, 16, 8, scatfac=0.0)
    assert not e.h
    out = e.effsrc(np.ones((16, 8, 16)), np.ones((16, 8, 16)))
    assert np.all(out == 0)
