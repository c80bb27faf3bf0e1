"""CPU oracle of the learned preconditioner correction network.

TEST INFRASTRUCTURE ONLY: imported by tests/ (and nothing in the product).
The reference has no learned component; this is a float64 restatement of the
network the library implements in csrc/ebg_learned.hip, written with torch on
the CPU so autograd supplies the gradients the GPU backward kernels must match:
  - the flat parameter vector: per convolution, weights [9][cout][cin]
    (tap = ky*3 + kx, offset (ky-1, kx-1) in (h, w)) then bias [cout], with
    the padded channel counts of LAYERS;
  - the U-Net: in, a0 | pool | e1a, e1b | pool | e2a, e2b | pool | e3a, e3b |
    up+skip | d2a, d2b | up+skip | d1a, d1b | up+skip | d0a, d0b | out,
    GELU (tanh form) after every convolution but the last;
  - 2x2 average pooling with zero padding at odd edges, nearest upsampling;
  - with emulate_bf16 the values are rounded to bfloat16 where the device
    stores BF16 (input image, weights, activations, pooled and skip-summed
    tensors); the rounding is a straight-through identity for gradients.
Loss: mean squared error over the three output channels.
"""
from __future__ import annotations

import numpy as np
import torch
import torch.nn.functional as F

LAYERS = ([(16, 32, 7, 32, 0, True), (32, 32, 32, 32, 0, True)] +
          [(32, 32, 32, 32, lv, True) for lv in (1, 1, 2, 2, 3, 3, 2, 2, 1, 1, 0, 0)] +
          [(32, 16, 32, 3, 0, False)])


def layout():
    out, off = [], 0
    for cin, cout, cir, cor, lv, ge in LAYERS:
        w = off
        off += 9 * cin * cout
        out.append((w, off, cin, cout, ge))
        off += cout
    return out, off


def _bf16(x: torch.Tensor, on: bool) -> torch.Tensor:
    if not on:
        return x
    return x + (x.to(torch.bfloat16).to(torch.float64) - x).detach()


def _gelu(x):
    c = 0.7978845608028654
    return 0.5 * x * (1.0 + torch.tanh(c * (x + 0.044715 * x ** 3)))


def _pool(x):  # NCHW, zero-padded to even size, 2x2 mean with divisor 4
    H, W = x.shape[2], x.shape[3]
    x = F.pad(x, (0, W % 2, 0, H % 2))
    return F.avg_pool2d(x, 2, divisor_override=4)


def _up(x, H, W):
    return x.repeat_interleave(2, dim=2).repeat_interleave(2, dim=3)[:, :, :H, :W]


def forward_t(params: torch.Tensor, x: torch.Tensor, emulate_bf16: bool = True) -> torch.Tensor:
    """x: [B][H][W][7] float64 tensor -> out [B][H][W][3]."""
    lay, n = layout()
    assert params.numel() == n
    r = lambda t: _bf16(t, emulate_bf16)

    def conv(k, t):
        w_off, b_off, cin, cout, ge = lay[k]
        w = params[w_off:w_off + 9 * cin * cout].reshape(9, cout, cin)
        w = r(w).permute(1, 2, 0).reshape(cout, cin, 3, 3)
        b = params[b_off:b_off + cout]
        pre = F.conv2d(t, w, b, padding=1)
        return r(_gelu(pre)) if ge else pre

    B, H, W, _ = x.shape
    t = torch.zeros((B, 16, H, W), dtype=torch.float64)
    t[:, :7] = r(x.permute(0, 3, 1, 2))
    x0 = conv(1, conv(0, t))
    x1 = conv(3, conv(2, r(_pool(x0))))
    x2 = conv(5, conv(4, r(_pool(x1))))
    x3 = conv(7, conv(6, r(_pool(x2))))
    y2 = conv(9, conv(8, r(_up(x3, x2.shape[2], x2.shape[3]) + x2)))
    y1 = conv(11, conv(10, r(_up(y2, x1.shape[2], x1.shape[3]) + x1)))
    y0 = conv(13, conv(12, r(_up(y1, x0.shape[2], x0.shape[3]) + x0)))
    out = conv(14, y0)
    return out[:, :3].permute(0, 2, 3, 1)


def forward(params: np.ndarray, x: np.ndarray, emulate_bf16: bool = True) -> np.ndarray:
    with torch.no_grad():
        return forward_t(torch.from_numpy(np.asarray(params, np.float64)),
                         torch.from_numpy(np.asarray(x, np.float64)), emulate_bf16).numpy()


def loss_and_grad(params: np.ndarray, x: np.ndarray, tgt: np.ndarray, emulate_bf16: bool = True):
    p = torch.from_numpy(np.asarray(params, np.float64).copy()).requires_grad_(True)
    out = forward_t(p, torch.from_numpy(np.asarray(x, np.float64)), emulate_bf16)
    loss = torch.mean((out - torch.from_numpy(np.asarray(tgt, np.float64))) ** 2)
    loss.backward()
    return float(loss), p.grad.numpy().copy()
