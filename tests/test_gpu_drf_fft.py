"""The FP32 DRF convolution by 2-D FFT (libirl_amd/csrc/drf_fft.hip) inside
the reconstruction: against the spatial kernel of the same reconstruction
(equal to FP32 rounding of the transforms) for every supported transform
length, and against the FP64 oracle with the reference's bFFTconvolve
tables, with axial padding, with rows added unconvolved, and without
attenuation (the backprojector then reuses one input spectrum per block)."""
import numpy as np
import pytest

from tests.cases import Problem, rel_l2

pytestmark = pytest.mark.gpu


def _recon(pb, mode, fast_rotate=0):
    import libirl_amd as L
    L.set_adv_options(fast_rotate, 0)
    prev = L.set_drf_fft(mode)
    try:
        return L.Recon(pb.parms(), pb.options(), pb.views, pb.drf_file, None, None, pb.mu, pb.prj,
                       pb.x0)
    finally:
        L.set_drf_fft(prev)


def _run(r, iters):
    for _ in range(iters):
        for s in range(r.num_subsets):
            r.subiteration(s)
    return r.estimate()


# (N, Nz, PSF half width, FFT length): 2 N - 1 taps wide PSFs need N + 2 half <= P
@pytest.mark.parametrize("n,nz,half,P", [(16, 8, 3, 32), (32, 16, 5, 64), (64, 24, 15, 128),
                                         (128, 32, 15, 160)])
def test_fft_equals_spatial_kernel(n, nz, half, P):
    print("FFT DRF vs spatial N=%d Nz=%d half=%d" % (n, nz, half), flush=True)
    pb = Problem(n=n, nz=nz, nviews=8, per=2, iters=2, half=half, noise=False,
                 pix=0.5 if n <= 64 else 0.33)
    a = _recon(pb, "spatial")
    b = _recon(pb, "fft")
    assert a.drf_algorithms()[:2] == ("spatial", "spatial")
    assert b.drf_algorithms() == ("fft", "fft", P)
    e = rel_l2(_run(b, 2), _run(a, 2))
    print("rel L2 FFT vs spatial after 2 iterations %.3g" % e)
    assert e <= 1e-5
    a.close()
    b.close()


# ragged sizes: slice counts that are not multiples of a line's 8 threads
# (a register straddles the last slice; 7 slices: fewer than one register),
# bin-pair counts that leave the last wave's line slots partly idle (15, 23,
# 31 pairs; NumPixels is even, as the library requires)
@pytest.mark.parametrize("n,nz,half", [(30, 13, 3), (46, 21, 5), (62, 7, 4)])
def test_fft_equals_spatial_kernel_ragged_sizes(n, nz, half):
    print("FFT DRF vs spatial, ragged N=%d Nz=%d half=%d" % (n, nz, half), flush=True)
    pb = Problem(n=n, nz=nz, nviews=8, per=2, iters=2, half=half, noise=False)
    a = _recon(pb, "spatial")
    b = _recon(pb, "fft")
    assert b.drf_algorithms()[:2] == ("fft", "fft")
    ga, gb = _run(a, 2), _run(b, 2)
    e = rel_l2(gb, ga)
    print("rel L2 FFT vs spatial %.3g" % e)
    assert e <= 1e-5
    if n <= 32:
        ref = pb.oracle_osem("f64")
        e_ref = rel_l2(gb, ref)
        print("rel L2 vs FP64 oracle %.3g" % e_ref)
        assert e_ref <= 1e-4
    a.close()
    b.close()


@pytest.mark.parametrize("atn", [True, False])
def test_irlosem_with_fft_tables_matches_oracle(atn):
    # bFFTconvolve tables: the reference's Fourier branch (no axial pad)
    print("IrlOsem, bFFTconvolve tables, atn=%s" % atn, flush=True)
    import libirl_amd as L
    pb = Problem(atn=atn, fft=True, pad=2, avg=2)
    ref = pb.oracle_osem("f64")
    L.set_adv_options(0, 0)
    recon = pb.x0.copy()
    rc = L.irl_osem(pb.parms(), pb.options(), pb.views, pb.drf_file, None, None,
                    None if pb.mu is None else pb.mu.copy(), pb.prj.copy(), recon)
    assert rc == 0, L.error_string()
    e = rel_l2(recon, ref)
    print("rel L2 vs FP64 oracle %.3g" % e)
    assert e <= 1e-4


@pytest.mark.parametrize("pad,avg", [(2, 1), (3, 2)])
def test_fft_kernel_with_axial_pad_matches_oracle(pad, avg):
    print("FFT DRF with axial pad %d avg %d" % (pad, avg), flush=True)
    pb = Problem(n=32, nz=16, nviews=8, per=2, iters=2, half=5, pad=pad, avg=avg, noise=False)
    ref = pb.oracle_osem("f64")
    r = _recon(pb, "fft")
    assert r.drf_algorithms()[:2] == ("fft", "fft")
    e = rel_l2(_run(r, 2), ref)
    print("rel L2 vs FP64 oracle %.3g" % e)
    assert e <= 1e-4
    r.close()


def test_fft_kernel_with_rows_beyond_the_detector():
    # a detector closer than half the volume: rows at distance <= 0 are added
    # unconvolved
    print("FFT DRF, rows at distance <= 0", flush=True)
    n, pix = 32, 0.5
    pb = Problem(n=n, nz=16, nviews=8, per=2, iters=2, half=5, noise=False,
                 cfcr=n * pix / 2 - 2.0)
    ref = pb.oracle_osem("f64")
    a = _recon(pb, "spatial")
    b = _recon(pb, "fft")
    ga, gb = _run(a, 2), _run(b, 2)
    e_sp, e_ref = rel_l2(gb, ga), rel_l2(gb, ref)
    print("rel L2 vs spatial %.3g, vs FP64 oracle %.3g" % (e_sp, e_ref))
    assert e_sp <= 1e-5 and e_ref <= 1e-4
    a.close()
    b.close()


def _asymmetric_table(pb, half, ramp_x, ramp_y):
    """Replaces the problem's DRF table by its Gaussians times linear ramps
    (the spectra of the FFT kernel are then complex)."""
    import os
    from libirl_amd import synth
    from oracle import oracle as O
    offs, psfs, start = synth.drf_for_orbit(pb.drf_file, pb.pix, pb.n, pb.views, half)
    r = np.arange(2 * half + 1) - half
    ramp = np.outer(1.0 + ramp_y * r / half, 1.0 + ramp_x * r / half)
    psfs = [(p * ramp / (p * ramp).sum()).astype(np.float32) for p in psfs]
    pb.drf_file = os.path.join(pb.tmpdir, "asym.drft")
    synth.write_drft(pb.drf_file, pb.pix, pb.pix, start, offs, psfs)
    pb.drf = O.DrfTable(start, pb.pix, offs, psfs)


# an x ramp only (symmetric in y) and both ramps: complex spectra either way
@pytest.mark.parametrize("n,nz,half,ramps", [(32, 16, 5, (0.4, 0.0)), (128, 32, 15, (0.3, 0.5))])
def test_asymmetric_psfs(n, nz, half, ramps):
    print("FFT DRF, asymmetric PSFs N=%d ramps %s" % (n, ramps), flush=True)
    pb = Problem(n=n, nz=nz, nviews=8, per=2, iters=2, half=half, noise=False,
                 pix=0.5 if n <= 64 else 0.33)
    _asymmetric_table(pb, half, *ramps)
    a = _recon(pb, "spatial")
    b = _recon(pb, "fft")
    ga, gb = _run(a, 2), _run(b, 2)
    e = rel_l2(gb, ga)
    print("rel L2 FFT vs spatial %.3g" % e)
    assert e <= 1e-5
    if n <= 32:
        ref = pb.oracle_osem("f64")
        e_ref = rel_l2(gb, ref)
        print("rel L2 vs FP64 oracle %.3g" % e_ref)
        assert e_ref <= 1e-4
    a.close()
    b.close()


@pytest.mark.parametrize("fast_rotate", [0, 1])
def test_fft_projector_with_both_rotations(fast_rotate):
    # the bilinear (rot3dpar) and the three-shear rotation feed the same
    # FFT kernels
    print("FFT DRF, fast rotate %d" % fast_rotate, flush=True)
    pb = Problem(n=64, nz=24, nviews=8, per=2, iters=2, half=15, noise=False)
    a = _recon(pb, "spatial", fast_rotate)
    b = _recon(pb, "fft", fast_rotate)
    e = rel_l2(_run(b, 2), _run(a, 2))
    print("rel L2 FFT vs spatial %.3g" % e)
    assert e <= 1e-5
    a.close()
    b.close()


def test_timed_choice_runs_one_of_the_two():
    print("FFT DRF timed choice", flush=True)
    pb = Problem(n=64, nz=32, nviews=8, per=2, iters=1, half=15, noise=False)
    a = _recon(pb, "spatial")
    b = _recon(pb, "auto")
    prj, bck, P = b.drf_algorithms()
    print("chosen", prj, bck, P)
    assert P == 128 and prj in ("spatial", "fft") and bck in ("spatial", "fft")
    e = rel_l2(_run(b, 1), _run(a, 1))
    assert e <= 1e-5
    a.close()
    b.close()


def test_sizes_beyond_the_transforms_keep_the_spatial_kernel():
    pb = Problem(n=144, nz=8, nviews=4, per=2, iters=1, half=15, noise=False, pix=0.3,
                 atn=False)
    r = _recon(pb, "fft")
    assert r.drf_algorithms() == ("spatial", "spatial", 0)
    r.close()
