"""CPU tests pinning the FP64 U-Net oracle (tests/unet_oracle.py): finite differences of its
backprop, adjoint pairs, and the parameter count of the network in unet_net.hpp."""
import numpy as np

import unet_oracle as U


def _rand_params(seed, scale=1.0):
    rng = np.random.default_rng(seed)
    p = np.zeros(U.n_params())
    off = 0
    for t, c1, c2, co, _ in U.SPECS:
        nw = t * co * (c1 + c2)
        p[off:off + nw] = rng.standard_normal(nw) * scale * np.sqrt(2.0 / (t * (c1 + c2)))
        off += nw
        p[off:off + co] = 0.05 * rng.standard_normal(co)
        off += co
    return p


def test_param_count():
    assert U.n_params() == 1466320


def test_adjoint_pairs():
    rng = np.random.default_rng(0)
    x = rng.standard_normal((8, 8, 8, 3))
    y = rng.standard_normal((4, 4, 4, 3))
    assert np.isclose((U.pool(x) * y).sum(), (x * U.pool_bwd(y)).sum(), rtol=1e-12)
    assert np.isclose((U.up(y) * x).sum(), (y * U.up_bwd(x)).sum(), rtol=1e-12)
    W = rng.standard_normal((27, 5, 3))
    yy = rng.standard_normal((8, 8, 8, 5))
    _, _, dx = U.conv_bwd(x, W, yy)
    assert np.isclose((U.conv(x, W, 0.0) * yy).sum(), (x * dx).sum(), rtol=1e-12)


def test_conv_matches_direct_sum():
    rng = np.random.default_rng(1)
    x = rng.standard_normal((4, 3, 5, 2))
    W = rng.standard_normal((27, 3, 2))
    y = U.conv(x, W, np.zeros(3))
    for (i, j, k) in [(0, 0, 0), (3, 2, 4), (1, 1, 2)]:
        ref = np.zeros(3)
        for t in range(27):
            di, dj, dk = t // 9 - 1, (t // 3) % 3 - 1, t % 3 - 1
            a, b, c = i + di, j + dj, k + dk
            if 0 <= a < 4 and 0 <= b < 3 and 0 <= c < 5:
                ref += W[t] @ x[a, b, c]
        assert np.allclose(y[i, j, k], ref, rtol=1e-12)


def test_backprop_matches_finite_differences():
    params = _rand_params(2)
    rng = np.random.default_rng(3)
    x = rng.standard_normal((8, 8, 8, 16))
    dout = rng.standard_normal((8, 8, 8, 16))
    A = U.forward(params, x)
    g = U.backward(params, A, dout)
    f = lambda p: (U.forward(p, x)['OUT'] * dout).sum()
    # one weight and one bias from a spread of layers
    offs, off = [], 0
    for li, (t, c1, c2, co, _) in enumerate(U.SPECS):
        nw = t * co * (c1 + c2)
        if li in (0, 3, 7, 8, 12, 14):
            offs += [off + (nw // 3), off + nw + co // 2]
        off += nw + co
    for q in offs:
        h = 1e-5
        pp, pm = params.copy(), params.copy()
        pp[q] += h
        pm[q] -= h
        fd = (f(pp) - f(pm)) / (2 * h)
        assert abs(fd - g[q]) <= 1e-6 * max(1.0, abs(fd)), (q, fd, g[q])
