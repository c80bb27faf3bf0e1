"""Learned correction of the multigrid preconditioner (ebg_cnn_* in the C ABI).

The network (csrc/ebg_learned.hip) is a fully convolutional four-level U-Net
with 3x3 convolutions, 32 channels, GELU, BF16 weights/activations and FP32
accumulation on the MFMA units.  It maps (log10 viscosity, r, z = M r) on the
cell grid to a correction of z; the solve keeps the corrected direction only
where it lowers the FP64 residual ||r - L z|| (ebg_stokes_solve with
learned=1), so the plain V-cycle is the fallback and the FP64 residual the judge.

Training pairs come from the plain preconditioner itself: for an error field e
the library forms r = L e and z = M r on the device (ebg_cnn_make_sample), and
the target is (e - z) / RMS(z).  Training (forward, backward,
AdamW) runs entirely in the library's kernels; torch only holds device memory.
"""
from __future__ import annotations

import ctypes as C
import pathlib

import numpy as np
import torch

from ._lib import check, lib
from .device import ptr

# (cin, cout, cin_real, cout_real, level, gelu) per convolution, in the order of
# the library's flat parameter vector (weights [9][cout][cin], then bias [cout])
LAYERS = ([(16, 32, 7, 32, 0, True), (32, 32, 32, 32, 0, True)] +
          [(32, 32, 32, 32, lv, True) for lv in (1, 1, 2, 2, 3, 3, 2, 2, 1, 1, 0, 0)] +
          [(32, 16, 32, 3, 0, False)])
N_IN, N_OUT = 7, 3
WEIGHTS_FILE = pathlib.Path(__file__).resolve().parent / "data" / "learned_2d_config2.npy"


def param_layout():
    """[(w_off, b_off, cin, cout, cin_real, cout_real, level, gelu)], total."""
    out, off = [], 0
    for cin, cout, cir, cor, lv, ge in LAYERS:
        w = off
        off += 9 * cin * cout
        out.append((w, off, cin, cout, cir, cor, lv, ge))
        off += cout
    return out, off


def effective_param_count() -> int:
    """Parameters that are not channel padding."""
    return sum(9 * cir * cor + cor for _, _, _, _, cir, cor, _, _ in param_layout()[0])


class LearnedCorrection:
    """The context's network: parameters, inference, loss/gradients, AdamW."""

    def __init__(self, ctx, seed: int = 1):
        self.ctx = ctx
        check(lib.ebg_cnn_create(ctx.handle, int(seed)), "ebg_cnn_create")
        n = C.c_int64()
        check(lib.ebg_cnn_num_params(ctx.handle, C.byref(n)), "ebg_cnn_num_params")
        self.num_params = n.value
        if self.num_params != param_layout()[1]:
            raise RuntimeError("library parameter layout does not match LAYERS")

    def get_params(self) -> np.ndarray:
        a = np.empty(self.num_params, dtype=np.float32)
        check(lib.ebg_cnn_get_params(self.ctx.handle, a.ctypes.data), "ebg_cnn_get_params")
        return a

    def set_params(self, a: np.ndarray) -> None:
        a = np.ascontiguousarray(a, dtype=np.float32)
        if a.size != self.num_params:
            raise ValueError(f"expected {self.num_params} parameters, got {a.size}")
        check(lib.ebg_cnn_set_params(self.ctx.handle, a.ctypes.data), "ebg_cnn_set_params")

    def save(self, path) -> None:
        pathlib.Path(path).parent.mkdir(parents=True, exist_ok=True)
        np.save(path, self.get_params())

    def load(self, path=WEIGHTS_FILE) -> None:
        self.set_params(np.load(path, allow_pickle=False))

    @staticmethod
    def _shape(img: torch.Tensor, ch: int):
        if img.dim() != 4 or img.shape[3] != ch or img.dtype != torch.float32 or not img.is_cuda:
            raise ValueError(f"expected a CUDA float32 tensor [B][H][W][{ch}]")
        return int(img.shape[0]), int(img.shape[1]), int(img.shape[2])

    def forward(self, img: torch.Tensor) -> torch.Tensor:
        B, H, W = self._shape(img, N_IN)
        img = img.contiguous()
        out = torch.empty((B, H, W, N_OUT), dtype=torch.float32, device=img.device)
        check(lib.ebg_cnn_forward(self.ctx.handle, B, H, W, ptr(img), ptr(out)), "ebg_cnn_forward")
        return out

    def loss(self, img: torch.Tensor, tgt: torch.Tensor) -> float:
        B, H, W = self._shape(img, N_IN)
        self._shape(tgt, N_OUT)
        v = C.c_double()
        check(lib.ebg_cnn_loss(self.ctx.handle, B, H, W, ptr(img.contiguous()), ptr(tgt.contiguous()),
                               C.byref(v)), "ebg_cnn_loss")
        return v.value

    def loss_and_grad(self, img: torch.Tensor, tgt: torch.Tensor, host_grad: bool = False):
        B, H, W = self._shape(img, N_IN)
        self._shape(tgt, N_OUT)
        v = C.c_double()
        g = np.empty(self.num_params, dtype=np.float32) if host_grad else None
        check(lib.ebg_cnn_loss_and_grad(self.ctx.handle, B, H, W, ptr(img.contiguous()),
                                        ptr(tgt.contiguous()), C.byref(v),
                                        g.ctypes.data if host_grad else None), "ebg_cnn_loss_and_grad")
        return v.value, g

    def adamw_step(self, lr=1e-3, betas=(0.9, 0.999), eps=1e-8, weight_decay=1e-4) -> None:
        check(lib.ebg_cnn_adamw_step(self.ctx.handle, lr, betas[0], betas[1], eps, weight_decay),
              "ebg_cnn_adamw_step")

    def make_sample(self, solver, e: torch.Tensor, in_img: torch.Tensor, tgt_img: torch.Tensor,
                    **opts) -> None:
        """One training sample of the solver's current system from the error
        vector e (reference order, device FP64)."""
        from .stokes import default_opts
        g = solver.prob.grid
        H, W = g.xnum - 1, g.ynum - 1
        if tuple(in_img.shape) != (H, W, N_IN) or tuple(tgt_img.shape) != (H, W, N_OUT):
            raise ValueError("sample images must be [xnum-1][ynum-1][7] and [...][3]")
        o = default_opts(**opts)
        check(lib.ebg_cnn_make_sample(self.ctx.handle, ptr(e, np.float64, solver.prob.N), C.byref(o),
                                      ptr(in_img), ptr(tgt_img)), "ebg_cnn_make_sample")


def smooth_error(g, rng: np.random.Generator, scales, corr_cells=(2.0, 40.0)) -> np.ndarray:
    """Random smooth error vector in the reference's unknown order (per cell
    [vx, vy, p']): Gaussian random fields with a random correlation length of
    corr_cells cells, component RMS values `scales`."""
    ny1, nx1 = g.ynum - 1, g.xnum - 1
    out = np.empty((nx1, ny1, 3))
    kx = np.fft.fftfreq(nx1)[:, None]
    ky = np.fft.fftfreq(ny1)[None, :]
    for c in range(3):
        L = float(np.exp(rng.uniform(np.log(corr_cells[0]), np.log(corr_cells[1]))))
        f = np.real(np.fft.ifft2(np.fft.fft2(rng.standard_normal((nx1, ny1))) *
                                 np.exp(-0.5 * (2 * np.pi) ** 2 * (kx ** 2 + ky ** 2) * L ** 2)))
        out[:, :, c] = scales[c] * f / max(f.std(), 1e-300)
    return out.reshape(-1)


def component_rms(S: np.ndarray) -> np.ndarray:
    a = S.reshape(-1, 3)
    return np.sqrt(np.mean(a * a, axis=0)) + 1e-300


def random_crops(in_img: torch.Tensor, tgt_img: torch.Tensor, n: int, ch: int, cw: int,
                 rng: np.random.Generator):
    """n random ch x cw crops of one sample image pair (device tensors)."""
    H, W = in_img.shape[0], in_img.shape[1]
    xs, ys = [], []
    for _ in range(n):
        h0 = int(rng.integers(0, H - ch + 1))
        w0 = int(rng.integers(0, W - cw + 1))
        xs.append(in_img[h0:h0 + ch, w0:w0 + cw])
        ys.append(tgt_img[h0:h0 + ch, w0:w0 + cw])
    return torch.stack(xs), torch.stack(ys)


def build_dataset(ctx, net: LearnedCorrection, field_seeds, samples_per_field: int, crops_per_sample: int,
                  crop=(128, 64), xnum=1001, ynum=501, seed=0, log=print, **opts):
    """Training crops from full-size layered random systems: for each field
    the mixed FGMRES solve gives the solution scale, then random smooth
    errors e give (input, target) images through the library, cropped."""
    from . import workloads
    from .device import to_device
    from .stokes import StokesSolverGPU, bc_mirror_factors
    rng = np.random.default_rng(seed)
    X, Y = [], []
    sol = StokesSolverGPU(ctx)
    for fs in field_seeds:
        prob = workloads.layered_random(xnum, ynum, 1.5e6 * (xnum - 1) / 1000, 6.0e5 * (ynum - 1) / 500,
                                        seed=int(fs))
        sol.set_system(prob)
        bcd = bc_mirror_factors(prob)
        S, st = sol.solve_system(precision=1, rtol=1e-6, max_iterations=400, bc_d=bcd, **opts)
        scales = component_rms(S.cpu().numpy())
        H, W = xnum - 1, ynum - 1
        in_img = torch.empty((H, W, N_IN), dtype=torch.float32, device="cuda")
        tgt_img = torch.empty((H, W, N_OUT), dtype=torch.float32, device="cuda")
        for _ in range(samples_per_field):
            e = to_device(smooth_error(prob.grid, rng, scales))
            net.make_sample(sol, e, in_img, tgt_img, bc_d=bcd, **opts)
            x, y = random_crops(in_img, tgt_img, crops_per_sample, crop[0], crop[1], rng)
            X.append(x)
            Y.append(y)
        log(f"field seed {fs}: solve iterations {st['iterations']}, samples {samples_per_field}")
    return torch.cat(X), torch.cat(Y)


def krylov_vectors(sol, nvec: int, **opts):
    """The first nvec orthonormal Krylov vectors v_k of the flexible GMRES
    iteration with the multigrid preconditioner (the vectors the
    preconditioner is applied to), built from the library's operator and
    preconditioner; torch only does the vector algebra."""
    b = sol.rhs()
    V = [b / torch.linalg.vector_norm(b)]
    for _ in range(nvec - 1):
        w = sol.apply(sol.precondition(V[-1], **opts))
        for _ in range(2):  # classical Gram-Schmidt, twice
            for v in V:
                w -= torch.dot(w, v) * v
        V.append(w / torch.linalg.vector_norm(w))
    return V


def build_krylov_dataset(ctx, net: LearnedCorrection, field_seeds, nvec: int, crops_per_sample: int,
                         crop=(128, 64), xnum=1001, ynum=501, seed=0, log=print, **opts):
    """Training crops whose inputs are the preconditioner's actual inputs: for
    each layered random field, the first nvec Krylov vectors v_k; the exact
    error e_k = L^-1 v_k comes from a tight solve with v_k as right-hand side,
    and ebg_cnn_make_sample forms (L e_k ~ v_k, z = M v_k, e_k - z)."""
    from . import workloads
    from .stokes import StokesSolverGPU, bc_mirror_factors
    rng = np.random.default_rng(seed)
    X, Y = [], []
    sol = StokesSolverGPU(ctx)
    H, W = xnum - 1, ynum - 1
    in_img = torch.empty((H, W, N_IN), dtype=torch.float32, device="cuda")
    tgt_img = torch.empty((H, W, N_OUT), dtype=torch.float32, device="cuda")
    for fs in field_seeds:
        prob = workloads.layered_random(xnum, ynum, 1.5e6 * (xnum - 1) / 1000, 6.0e5 * (ynum - 1) / 500,
                                        seed=int(fs))
        sol.set_system(prob)
        o = dict(opts, bc_d=bc_mirror_factors(prob))
        its = []
        for v in krylov_vectors(sol, nvec, **o):
            e, st = sol.solve_rhs(v, rtol=1e-9, max_iterations=600, **o)
            its.append(st["iterations"])
            net.make_sample(sol, e, in_img, tgt_img, **o)
            x, y = random_crops(in_img, tgt_img, crops_per_sample, crop[0], crop[1], rng)
            X.append(x)
            Y.append(y)
        log(f"field seed {fs}: {nvec} Krylov samples, solve iterations {min(its)}-{max(its)}")
    return torch.cat(X), torch.cat(Y)


def train(net: LearnedCorrection, X: torch.Tensor, Y: torch.Tensor, steps: int, batch: int = 32,
          lr: float = 1e-3, weight_decay: float = 1e-4, seed: int = 0, log=print, log_every: int = 100):
    """AdamW on minibatches of (X, Y); returns the loss history."""
    g = torch.Generator(device="cpu").manual_seed(seed)
    hist = []
    n = X.shape[0]
    for it in range(steps):
        idx = torch.randint(0, n, (batch,), generator=g).to(X.device)
        loss, _ = net.loss_and_grad(X.index_select(0, idx), Y.index_select(0, idx))
        net.adamw_step(lr=lr, weight_decay=weight_decay)
        hist.append(loss)
        if log and (it % log_every == 0 or it == steps - 1):
            log(f"step {it}: loss {loss:.5f}")
    return hist
