"""CPU restatement (float64, numpy) of the scatter CNN of libirl_amd.

TEST INFRASTRUCTURE ONLY: imported by tests/ and bench.py as the checker of
the HIP forward/backward kernels (libirl_amd/csrc/cnn.hip).  The product
never imports it.

The reference has no network: the CNN stands in for the two FFT convolution
terms of its ESSE effective source (krnlprj3d.c:826-948), E0 = K * act and
E1 = K(mu - muavg) * act, which enter the per-view source as

    eff_v = (E0 - d_v E1) exp(d_v fMu) map fDens  (krnlprj3d.c:836-838, 903-920, 946-947)
          ~ (E0 + d_v (fMu E0 - E1)) map fDens            (first order in d_v fMu)

with d_v the view's attenuation depth.  For a kernel that is symmetric about
the rotation axis both terms are the same in every view's frame, so the
network predicts them once per sub-iteration in the unrotated frame.

Tensors are [Y][X][Z][C] (channels fastest) on the half-resolution grid.
Convolution weights are [27][Cout][Cin] with tap = 9(dy+1) + 3(dx+1) + (dz+1);
layer l reads its input at offsets dil_l * (dy, dx, dz), zero outside.
"""
import numpy as np

# (Cin, Cout, dilation, relu)
LAYERS = ((2, 32, 1, True), (32, 32, 2, True), (32, 32, 4, True), (32, 2, 8, False))


def taps():
    return [(dy, dx, dz) for dy in (-1, 0, 1) for dx in (-1, 0, 1) for dz in (-1, 0, 1)]


def _shift(a, dy, dx, dz):
    """s[y, x, z] = a[y + dy, x + dx, z + dz], zero outside."""
    Y, X, Z = a.shape[:3]
    s = np.zeros_like(a)
    ys, yd = (slice(dy, Y), slice(0, Y - dy)) if dy >= 0 else (slice(0, Y + dy), slice(-dy, Y))
    xs, xd = (slice(dx, X), slice(0, X - dx)) if dx >= 0 else (slice(0, X + dx), slice(-dx, X))
    zs, zd = (slice(dz, Z), slice(0, Z - dz)) if dz >= 0 else (slice(0, Z + dz), slice(-dz, Z))
    if min(ys.stop - ys.start, xs.stop - xs.start, zs.stop - zs.start) > 0:
        s[yd, xd, zd] = a[ys, xs, zs]
    return s


def conv(inp, W, b, dil):
    """out[v][co] = b[co] + sum_tap sum_ci W[tap][co][ci] in[v + dil*off(tap)][ci]."""
    out = np.broadcast_to(np.asarray(b, np.float64), inp.shape[:3] + (W.shape[1],)).copy()
    for t, (dy, dx, dz) in enumerate(taps()):
        out += _shift(inp, dil * dy, dil * dx, dil * dz) @ W[t].T
    return out


def conv_grad_input(dout, W, dil):
    """Adjoint of conv in its input: din[v][ci] = sum_tap sum_co W[tap][co][ci] dout[v - dil*off][co]."""
    din = np.zeros(dout.shape[:3] + (W.shape[2],))
    for t, (dy, dx, dz) in enumerate(taps()):
        din += _shift(dout, -dil * dy, -dil * dx, -dil * dz) @ W[t]
    return din


def conv_grad_weights(inp, dout, dil):
    """dW[tap][co][ci] = sum_v dout[v][co] in[v + dil*off][ci]; db[co] = sum_v dout[v][co]."""
    co, ci = dout.shape[3], inp.shape[3]
    dW = np.zeros((27, co, ci))
    d2 = dout.reshape(-1, co)
    for t, (dy, dx, dz) in enumerate(taps()):
        dW[t] = d2.T @ _shift(inp, dil * dy, dil * dx, dil * dz).reshape(-1, ci)
    return dW, d2.sum(axis=0)


def init_params(seed=0):
    """He-normal weights, zero biases (the HIP trainer's initialisation is its own)."""
    rng = np.random.default_rng(seed)
    ps = []
    for ci, co, _, _ in LAYERS:
        ps.append((rng.normal(0, np.sqrt(2.0 / (27 * ci)), (27, co, ci)), np.zeros(co)))
    return ps


def forward(params, inp):
    """Activations [input, a1, a2, a3, out] (a_l after the ReLU where the layer has one)."""
    acts = [np.asarray(inp, np.float64)]
    for (W, b), (_, _, dil, relu) in zip(params, LAYERS):
        a = conv(acts[-1], np.asarray(W, np.float64), b, dil)
        acts.append(np.maximum(a, 0.0) if relu else a)
    return acts


def mse(pred, tgt):
    """Loss mean((pred - tgt)^2) over every element, and its gradient."""
    d = pred - tgt
    return float(np.mean(d * d)), 2.0 * d / d.size


def backward(params, acts, dout):
    """Gradients [(dW, db)] of the layers for d(loss)/d(out) = dout."""
    grads = [None] * len(LAYERS)
    g = dout
    for l in range(len(LAYERS) - 1, -1, -1):
        W = np.asarray(params[l][0], np.float64)
        dil, relu = LAYERS[l][2], LAYERS[l][3]
        if relu:
            g = g * (acts[l + 1] > 0)
        grads[l] = conv_grad_weights(acts[l], g, dil)
        if l > 0:
            g = conv_grad_input(g, W, dil)
    return grads


def pool2(v):
    """Mean over 2x2x2 blocks of a [Y][X][Z] volume (even sizes)."""
    Y, X, Z = v.shape
    return v.reshape(Y // 2, 2, X // 2, 2, Z // 2, 2).mean(axis=(1, 3, 5))


def _up_axis(c, axis):
    n = c.shape[axis]
    j = np.arange(2 * n)
    u = j / 2.0 - 0.25
    i0 = np.floor(u).astype(int)
    w = u - i0
    lo = np.clip(i0, 0, n - 1)
    hi = np.clip(i0 + 1, 0, n - 1)
    w = np.where(u < 0, 0.0, w)
    shape = [1] * c.ndim
    shape[axis] = 2 * n
    w = w.reshape(shape)
    return np.take(c, lo, axis=axis) * (1 - w) + np.take(c, hi, axis=axis) * w


def upsample2(c):
    """Trilinear upsampling of a [y][x][z] grid by 2: fine sample j sits at
    coarse coordinate j/2 - 1/4 (cell centres), clamped at the edges."""
    for a in range(3):
        c = _up_axis(c, a)
    return c


def network_input(x, dens, sx):
    """[Y/2][X/2][Z/2][2]: pooled x / sx and pooled density (map*fDens)."""
    return np.stack([pool2(x) / sx, pool2(dens)], axis=-1)


def targets(E0, E1, fMu, sx, kappa):
    """Training targets: pooled E0/sx and kappa * pooled (fMu E0 - E1)/sx."""
    return np.stack([pool2(E0) / sx, kappa * pool2(fMu * E0 - E1) / sx], axis=-1)


def sources(x, dens, out, sx, kappa, primary=1.0, cntr=0.0):
    """Unrotated-frame source volumes from the network output:
    g0 = primary x + (sx up(o0) + cntr x) dens,  g1 = (sx/kappa) up(o1) dens,
    with dens = map * fDens.  Per view: att = (rot g0 + d rot g1) afac."""
    e0 = sx * upsample2(out[..., 0])
    e1 = (sx / kappa) * upsample2(out[..., 1])
    return primary * x + (e0 + cntr * x) * dens, e1 * dens
