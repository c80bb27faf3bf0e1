"""The CPU oracle of the learned correction network: its parameter layout
agrees with the package's, its autograd gradient agrees with central finite
differences (float64, no rounding), odd image sizes pool and upsample
consistently, and a zero output layer gives a zero correction."""
import numpy as np
import torch

import oracle.cnn as ocnn
from earthbox_gpu import learned


def _params(seed, zero_out=False):
    rng = np.random.default_rng(seed)
    lay, n = ocnn.layout()
    p = np.zeros(n)
    for (w, b, cin, cout, ge), (_, _, cir, cor, _, _) in zip(lay, ocnn.LAYERS):
        W = np.zeros((9, cout, cin))
        W[:, :cor, :cir] = rng.uniform(-1, 1, (9, cor, cir)) * np.sqrt(3.0 / (9 * cir))
        if zero_out and not ge:
            W[...] = 0.0
        p[w:w + 9 * cin * cout] = W.ravel()
        p[b:b + cor] = 0.1 * rng.standard_normal(cor)
        if zero_out and not ge:
            p[b:b + cout] = 0.0
    return p


def test_layout_matches_package():
    assert ocnn.layout()[1] == learned.param_layout()[1]
    assert ocnn.LAYERS == learned.LAYERS


def test_zero_output_layer_gives_zero():
    x = np.random.default_rng(1).standard_normal((2, 13, 10, 7))
    out = ocnn.forward(_params(2, zero_out=True), x)
    assert out.shape == (2, 13, 10, 3) and np.all(out == 0.0)


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(3)
    p = _params(4)
    x = rng.standard_normal((1, 9, 7, 7))
    t = rng.standard_normal((1, 9, 7, 3))
    loss, g = ocnn.loss_and_grad(p, x, t, emulate_bf16=False)
    lay, _ = ocnn.layout()
    idx = []
    for (w, b, cin, cout, ge), (_, _, cir, cor, _, _) in zip(lay, ocnn.LAYERS):
        idx += [w + (4 * cout + 0) * cin + 1, b]  # centre tap weight, first bias
    for k in idx:
        h = 1e-6 * max(1.0, abs(p[k]))
        pp, pm = p.copy(), p.copy()
        pp[k] += h
        pm[k] -= h
        fd = (ocnn.loss_and_grad(pp, x, t, False)[0] - ocnn.loss_and_grad(pm, x, t, False)[0]) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 * max(1e-3, abs(fd)) + 1e-10, (k, fd, g[k])


def test_pool_and_upsample_odd_sizes():
    x = torch.arange(2 * 5 * 7, dtype=torch.float64).reshape(1, 2, 5, 7)
    p = ocnn._pool(x)
    assert p.shape == (1, 2, 3, 4)
    assert p[0, 0, 2, 3] == x[0, 0, 4, 6] / 4.0  # lone corner pixel, divisor 4
    u = ocnn._up(p, 5, 7)
    assert u.shape == (1, 2, 5, 7) and u[0, 1, 4, 6] == p[0, 1, 2, 3]
