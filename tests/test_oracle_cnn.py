"""The scatter-CNN oracle (oracle/cnn.py) pinned by checks that do not restate
it: brute-force convolution loops (tap order, offset direction, dilation,
zero padding), the adjoint identity of the input gradient, central finite
differences of the loss for every layer's weights and biases, and exact
reproduction of linear fields by pooling followed by upsampling."""
import numpy as np
import pytest

from oracle import cnn as C


def _brute_conv(inp, W, b, dil):
    Y, X, Z, ci = inp.shape
    co = W.shape[1]
    out = np.zeros((Y, X, Z, co))
    for y in range(Y):
        for x in range(X):
            for z in range(Z):
                acc = np.array(b, np.float64).copy()
                t = 0
                for dy in (-1, 0, 1):
                    for dx in (-1, 0, 1):
                        for dz in (-1, 0, 1):
                            yy, xx, zz = y + dil * dy, x + dil * dx, z + dil * dz
                            if 0 <= yy < Y and 0 <= xx < X and 0 <= zz < Z:
                                acc += W[t] @ inp[yy, xx, zz]
                            t += 1
                out[y, x, z] = acc
    return out


@pytest.mark.parametrize("dil", [1, 2, 3])
def test_conv_equals_brute_force_loops(dil):
    rng = np.random.default_rng(dil)
    inp = rng.normal(size=(4, 5, 6, 3))
    W = rng.normal(size=(27, 2, 3))
    b = rng.normal(size=2)
    np.testing.assert_allclose(C.conv(inp, W, b, dil), _brute_conv(inp, W, b, dil), rtol=1e-12,
                               atol=1e-12)


def test_single_tap_is_a_shift():
    # tap (dy, dx, dz) = (1, 0, -1) reads in[y + d, x, z - d]
    inp = np.random.default_rng(0).normal(size=(6, 5, 7, 1))
    W = np.zeros((27, 1, 1))
    W[9 * 2 + 3 * 1 + 0, 0, 0] = 1.0
    out = C.conv(inp, W, [0.0], 2)
    assert np.allclose(out[1:4, :, 2:, 0], inp[3:6, :, 0:5, 0])
    assert np.all(out[4:, :, :, 0] == 0) and np.all(out[:, :, :2, 0] == 0)


@pytest.mark.parametrize("dil", [1, 4])
def test_input_gradient_is_the_adjoint(dil):
    rng = np.random.default_rng(3)
    x = rng.normal(size=(5, 6, 7, 4))
    y = rng.normal(size=(5, 6, 7, 3))
    W = rng.normal(size=(27, 3, 4))
    lhs = np.sum(C.conv(x, W, np.zeros(3), dil) * y)
    rhs = np.sum(x * C.conv_grad_input(y, W, dil))
    assert abs(lhs - rhs) < 1e-9 * abs(lhs)


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(5)
    params = C.init_params(seed=2)
    # nonzero biases so that every ReLU region is exercised
    params = [(W, rng.normal(0, 0.1, size=b.shape)) for W, b in params]
    inp = rng.normal(size=(5, 4, 6, 2))
    tgt = rng.normal(size=(5, 4, 6, 2))

    def loss(ps):
        return C.mse(C.forward(ps, inp)[-1], tgt)[0]

    acts = C.forward(params, inp)
    _, d = C.mse(acts[-1], tgt)
    grads = C.backward(params, acts, d)
    eps = 1e-6
    for l in range(len(C.LAYERS)):
        W, b = params[l]
        for _ in range(6):
            idx = tuple(rng.integers(0, s) for s in W.shape)
            Wp, Wm = W.copy(), W.copy()
            Wp[idx] += eps
            Wm[idx] -= eps
            pp = [p if k != l else (Wp, b) for k, p in enumerate(params)]
            pm = [p if k != l else (Wm, b) for k, p in enumerate(params)]
            fd = (loss(pp) - loss(pm)) / (2 * eps)
            assert abs(fd - grads[l][0][idx]) <= 1e-6 * max(1.0, abs(fd)) + 1e-9, (l, idx)
        k = int(rng.integers(0, b.shape[0]))
        bp, bm = b.copy(), b.copy()
        bp[k] += eps
        bm[k] -= eps
        fd = (loss([p if q != l else (W, bp) for q, p in enumerate(params)]) -
              loss([p if q != l else (W, bm) for q, p in enumerate(params)])) / (2 * eps)
        assert abs(fd - grads[l][1][k]) <= 1e-6 * max(1.0, abs(fd)) + 1e-9, (l, k)


def test_pool_then_upsample_reproduces_linear_fields():
    y, x, z = np.meshgrid(np.arange(8), np.arange(6), np.arange(10), indexing="ij")
    f = 0.3 + 0.7 * (y + 0.5) - 1.1 * (x + 0.5) + 0.45 * (z + 0.5)
    up = C.upsample2(C.pool2(f))
    assert up.shape == f.shape
    np.testing.assert_allclose(up[1:-1, 1:-1, 1:-1], f[1:-1, 1:-1, 1:-1], rtol=1e-12, atol=1e-12)
    # at the faces the samples beyond the last cell centre are clamped
    np.testing.assert_allclose(C.upsample2(np.ones((3, 3, 3))), np.ones((6, 6, 6)))
    # pooling is the block mean
    assert np.isclose(C.pool2(f)[1, 2, 3], f[2:4, 4:6, 6:8].mean())
