"""Seeded synthetic SPECT inputs for the benchmark workloads and the tests.

Everything here is plain numpy and deterministic in its seed:
  * ellipsoid activity phantoms with hot/cold inserts and the matching
    attenuation map (body mu plus a low-mu "lung" ellipsoid),
  * view tables for circular and non-circular (ellipse) orbits,
  * Gaussian distance-dependent DRF tables written as DRFT files
    (drftabio.c:21-46 layout),
  * Poisson noise at a given total count.
Volumes are external order [z][y][x]; projections [view][slice][bin].
"""
import struct

import numpy as np

from . import PrjView


def ellipsoid_mask(shape, center, radii, angle=0.0):
    """Boolean mask of an ellipsoid in a [z][y][x] grid (voxel units), rotated
    by 'angle' about the slice axis."""
    nz, ny, nx = shape
    z, y, x = np.meshgrid(np.arange(nz) + 0.5, np.arange(ny) + 0.5, np.arange(nx) + 0.5,
                          indexing="ij")
    dx, dy, dz = x - center[2], y - center[1], z - center[0]
    c, s = np.cos(angle), np.sin(angle)
    u = c * dx + s * dy
    w = -s * dx + c * dy
    return (u / radii[2]) ** 2 + (w / radii[1]) ** 2 + (dz / radii[0]) ** 2 <= 1.0


def phantom(n, nz, seed, n_inserts=(3, 5), body_mu=0.15, lung_mu=0.05):
    """Activity and attenuation (1/cm) volumes of the seeded ellipsoid class."""
    rng = np.random.default_rng(seed)
    shape = (nz, n, n)
    c = (nz / 2, n / 2, n / 2)
    body_r = (0.42 * nz, 0.33 * n, 0.40 * n)
    body = ellipsoid_mask(shape, c, body_r)
    act = np.where(body, 1.0, 0.0).astype(np.float32)
    mu = np.where(body, body_mu, 0.0).astype(np.float32)
    lung_c = (nz / 2, n / 2 - 0.08 * n, n / 2 + 0.18 * n)
    lung = ellipsoid_mask(shape, lung_c, (0.25 * nz, 0.12 * n, 0.10 * n)) & body
    mu[lung] = lung_mu
    k = int(rng.integers(n_inserts[0], n_inserts[1] + 1))
    for _ in range(k):
        r = rng.uniform(0.04, 0.09, size=3) * np.array([nz, n, n])
        off = rng.uniform(-0.5, 0.5, size=3) * np.array([body_r[0], body_r[1], body_r[2]])
        m = ellipsoid_mask(shape, (c[0] + off[0], c[1] + off[1], c[2] + off[2]), r,
                           rng.uniform(0, np.pi)) & body
        act[m] = rng.uniform(0.0, 8.0)
    return act, mu


def views_circular(nviews, cfcr, n):
    """Views over 360 deg with negative rotation (irl.h:128-133)."""
    return [PrjView(np.float32(-2.0 * np.pi * v / nviews), 0.0, float(n), float(cfcr))
            for v in range(nviews)]


def views_ellipse(nviews, n, seed, rmin=20.0, rmax=28.0):
    """Non-circular body-contour orbit: CFCR follows a seeded ellipse with
    semi-axes in [rmin, rmax] cm."""
    rng = np.random.default_rng(seed + 7919)
    a, b = sorted(rng.uniform(rmin, rmax, size=2))
    out = []
    for v in range(nviews):
        th = -2.0 * np.pi * v / nviews
        r = a * b / np.sqrt((b * np.cos(th)) ** 2 + (a * np.sin(th)) ** 2)
        out.append(PrjView(np.float32(th), 0.0, float(n), float(r)))
    return out


def gaussian_drf_table(pixsize, start, ndist, spacing, half, fwhm0=0.3, slope=0.02):
    """Gaussian DRF with FWHM linear in distance, (2*half+1)^2 support, each
    plane normalised to sum 1.  Taps below the normal float range (< 2^-126,
    far Gaussian tails of the narrow planes) are stored as 0: as subnormal
    operands they would only slow CPU float arithmetic 10-100x.  Returns
    (offs, psfs) with offs=(half+1)."""
    psfs = []
    ax = (np.arange(2 * half + 1) - half) * pixsize
    for d in range(ndist):
        dist = start + d * spacing
        fwhm = fwhm0 + slope * dist
        sig = fwhm / (2.0 * np.sqrt(2.0 * np.log(2.0)))
        g = np.exp(-(ax ** 2) / (2 * sig * sig))
        p = np.outer(g, g)
        p = (p / p.sum()).astype(np.float32)
        p[np.abs(p) < np.finfo(np.float32).tiny] = 0.0
        psfs.append(p)
    return [(half + 1, half + 1)] * ndist, psfs


def write_drft(fname, pixsize, spacing, start, offs, psfs, title="gaussian drf",
               col="synthetic"):
    """Writes a DRFT file (drftabio.c:21-46)."""
    with open(fname, "wb") as f:
        f.write(b"DRFT")
        f.write(struct.pack("<i", 1))
        f.write(title.encode()[:79].ljust(80, b"\0"))
        f.write(col.encode()[:19].ljust(20, b"\0"))
        f.write(struct.pack("<fffii", pixsize, spacing, start, len(psfs), 0))
        for (ox, oy), p in zip(offs, psfs):
            assert p.shape == (2 * oy - 1, 2 * ox - 1)
            f.write(struct.pack("<ii", ox, oy))
            f.write(np.ascontiguousarray(p, dtype="<f4").tobytes())


def drf_for_orbit(fname, pixsize, n, views, half):
    """DRFT file covering every distance the orbit needs (psDrfSetup range)."""
    cf = [v.CFCR for v in views]
    start = pixsize / 2
    max_dist = max(cf) + n // 2 * pixsize - pixsize / 2
    ndist = int(np.ceil((max_dist - start) / pixsize)) + 2
    offs, psfs = gaussian_drf_table(pixsize, start, ndist, pixsize, half)
    write_drft(fname, pixsize, pixsize, start, offs, psfs)
    return offs, psfs, start


def poisson(prj, total_counts, seed):
    """Scales prj to total_counts and draws seeded Poisson counts."""
    rng = np.random.default_rng(seed + 104729)
    s = prj.sum(dtype=np.float64)
    lam = prj.astype(np.float64) * (total_counts / s if s > 0 else 0.0)
    return rng.poisson(np.clip(lam, 0, None)).astype(np.float32)


def write_krnl(fname, krnl, krnl_mu, mu0, pixsize, hi_e=154.0, lo_e=126.0, pri_fac=1.0,
               scat_fac=1.0, cntr_pix_fac=0.0, title="esse kernel", isotope="Tc99m"):
    """Writes a KRNL file (krnlio.c:30-60 layout): krnl and krnl_mu are
    [nDists][nXYoffs][nXYoffs] one-quadrant kernels."""
    krnl = np.ascontiguousarray(krnl, dtype="<f4")
    krnl_mu = np.ascontiguousarray(krnl_mu, dtype="<f4")
    nd, n, n2 = krnl.shape
    assert n == n2 and krnl_mu.shape == krnl.shape
    with open(fname, "wb") as f:
        f.write(b"KRNL")
        f.write(struct.pack("<i", 1))
        f.write(title.encode()[:79].ljust(80, b"\0"))
        f.write(isotope.encode()[:19].ljust(20, b"\0"))
        f.write(struct.pack("<7f", mu0, hi_e, lo_e, pixsize, pri_fac, scat_fac, cntr_pix_fac))
        f.write(struct.pack("<ii", n, nd))
        f.write(krnl.tobytes())
        f.write(krnl_mu.tobytes())


def gaussian_mixture_krnl(noffs, ndepths, pixsize, amps=(0.02, 0.006, 0.0015),
                          sigmas=(0.8, 2.0, 4.5), mu0=0.15, mu_spread=0.02):
    """Seeded-free Gaussian-mixture ESSE kernel on one quadrant of offsets
    (depth index d has offset ndepths//2 - 1 - d, krnlprj3d.c:92-97) and a
    smoothly varying average-attenuation kernel around mu0 (1/cm)."""
    off = np.arange(noffs) * pixsize
    dep = (ndepths // 2 - 1 - np.arange(ndepths)) * pixsize
    r2 = off[None, :, None] ** 2 + off[None, None, :] ** 2 + dep[:, None, None] ** 2
    k = np.zeros((ndepths, noffs, noffs))
    for a, s in zip(amps, sigmas):
        k += a * np.exp(-r2 / (2 * s * s))
    kmu = mu0 * (1.0 - mu_spread * np.sqrt(r2) / (np.sqrt(r2).max() + 1e-9))
    return k.astype(np.float32), kmu.astype(np.float32)


def krnl_full_sum(krnl):
    """Sum of the kernel after the reference mirrors its quadrant onto
    +-(nXYoffs-1) in both lateral axes (krnlprj3d.c:43-127): offset 0 once,
    the others twice per axis."""
    k = np.asarray(krnl, np.float64)
    w = np.full(k.shape[1], 2.0)
    w[0] = 1.0
    return float(np.einsum("dij,i,j->", k, w, w))


def scale_krnl_to_fraction(krnl, fraction):
    """Scales a kernel so that a uniform activity in a uniform mu0 medium has
    the given scatter fraction eff/(primary + eff) = S/(1 + S), S the
    mirrored kernel sum (the effective source of such a medium is S times
    the activity)."""
    s = krnl_full_sum(krnl)
    target = fraction / (1.0 - fraction)
    return (np.asarray(krnl, np.float64) * (target / s)).astype(np.float32)


def esse_kernel_file(fname, cfg, fraction=0.3):
    """KRNL file of a workload configuration's ESSE kernel (Gaussian mixture,
    mu0 = 0.15/cm, krnl_offs x krnl_offs x krnl_depths one-quadrant support)
    scaled to the given scatter fraction."""
    k, km = gaussian_mixture_krnl(cfg["krnl_offs"], cfg["krnl_depths"], cfg["pix"])
    k = scale_krnl_to_fraction(k, fraction)
    write_krnl(fname, k, km, 0.15, cfg["pix"])
    return fname


# Workload definitions of the benchmark configurations.
CONFIGS = {
    "c1": dict(n=64, nz=64, pix=0.4, views=60, vs=15, iters=2, orbit="circular", cfcr=25.0,
               inserts=(3, 5), half=7, counts=5e5),
    "c2": dict(n=128, nz=128, pix=0.33, views=120, vs=15, iters=4, orbit="ellipse",
               inserts=(6, 10), half=15, counts=2e6),
    "c3": dict(n=128, nz=128, pix=0.33, views=120, vs=15, iters=4, orbit="ellipse",
               inserts=(6, 10), half=15, counts=2e6, krnl_offs=17, krnl_depths=33,
               srf_iters=4),
}
