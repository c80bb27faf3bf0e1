"""Host-side mirror of the reference's hot-path entry points, driving the C ABI.

Method names follow the reference Julia functions they replace (without the
trailing '!'); arguments are device tensors (see device.py for the layout).
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib
from ._lib import check, lib
from .device import ptr
from .grid import StaggeredGrid2D


def _np_ptr(a: np.ndarray) -> int:
    return a.ctypes.data


class GPUContext:
    """One HIP stream plus device-resident grid state (ebg_ctx)."""

    def __init__(self, device_index: int = 0):
        if not torch.cuda.is_available():
            raise RuntimeError("EarthBox GPU path needs a HIP device")
        handle = C.c_void_p()
        check(lib.ebg_ctx_create(C.byref(handle), device_index), "ebg_ctx_create")
        self._h = handle
        self.device = torch.device("cuda", device_index)
        self.grid: StaggeredGrid2D | None = None
        # Run on torch's current stream (the null stream unless the caller
        # selected another) so tensor copies and library kernels are ordered.
        stream = torch.cuda.current_stream(self.device).cuda_stream
        check(lib.ebg_ctx_set_stream(self._h, C.c_void_p(stream) if stream else None),
              "ebg_ctx_set_stream")

    @property
    def handle(self) -> C.c_void_p:
        return self._h

    def close(self) -> None:
        if getattr(self, "_h", None) is not None and self._h.value:
            lib.ebg_ctx_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def synchronize(self) -> None:
        check(lib.ebg_ctx_synchronize(self._h), "ebg_ctx_synchronize")

    # ---- grid -------------------------------------------------------------
    def set_grid(self, g: StaggeredGrid2D) -> None:
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in
                (g.gridx_b, g.gridy_b, g.xstp_b, g.ystp_b,
                 g.gridx_vy, g.xstp_vy, g.gridy_vx, g.ystp_vx)]
        check(lib.ebg_grid2d_set(self._h, g.xnum, g.ynum, *[_np_ptr(a) for a in arrs],
                                 g.xstpavg, g.ystpavg), "ebg_grid2d_set")
        self.grid = g

    def _need_grid(self) -> StaggeredGrid2D:
        if self.grid is None:
            raise RuntimeError("set_grid() must be called first")
        return self.grid

    # ---- markers ------------------------------------------------------------
    def get_marker_inside_flags(self, x: torch.Tensor, y: torch.Tensor,
                                out: torch.Tensor | None = None) -> torch.Tensor:
        """GridFuncs.get_marker_inside_flags (src/utils/GridFuncs.jl:54-80)."""
        self._need_grid()
        n = x.numel()
        if out is None:
            out = torch.empty(n, dtype=torch.int8, device=x.device)
        check(lib.ebg_marker_inside_flags(self._h, n, ptr(x, np.float64), ptr(y, np.float64, n),
                                          ptr(out, np.int8, n)), "ebg_marker_inside_flags")
        return out

    def calc_upper_left_basic_grid_cell_indices_for_markers(
            self, x, y, inside_flags, marker_xn, marker_yn) -> None:
        """MarkerGridMapping.jl:99-129 (bit-exact, 1-based Int32)."""
        self._need_grid()
        n = x.numel()
        check(lib.ebg_marker_cell_indices(
            self._h, n, ptr(x, np.float64), ptr(y, np.float64, n),
            ptr(inside_flags, np.int8, n) if inside_flags is not None else None,
            ptr(marker_xn, np.int32, n), ptr(marker_yn, np.int32, n)),
            "ebg_marker_cell_indices")

    def _rk4(self, order, timestep, x, y, xn, yn, flags, vx1, vy1, esp,
             mvx, mvy, mspin, sxx, sxy, update_location):
        g = self._need_grid()
        n = x.numel()
        ny, nx = g.ynum, g.xnum
        check(lib.ebg_rk4_advect(
            self._h, n, int(order), float(timestep),
            ptr(x, np.float64, n), ptr(y, np.float64, n),
            ptr(xn, np.int32, n), ptr(yn, np.int32, n), ptr(flags, np.int8, n),
            ptr(vx1, np.float64, (ny + 1) * nx), ptr(vy1, np.float64, ny * (nx + 1)),
            ptr(esp, np.float64, ny * nx),
            ptr(mvx, np.float64, n), ptr(mvy, np.float64, n), ptr(mspin, np.float64, n),
            ptr(sxx, np.float64, n), ptr(sxy, np.float64, n), int(update_location)),
            "ebg_rk4_advect")

    def update_marker_velocity_and_spin_using_runge_kutta(
            self, runge_kutta_order_max, timestep, inside_flags, marker_x, marker_y,
            marker_xn, marker_yn, vx1, vy1, esp, marker_vx, marker_vy, marker_spin) -> None:
        """VelocityAndSpin.jl:45-129: marker velocity and spin only."""
        self._rk4(runge_kutta_order_max, timestep, marker_x, marker_y, marker_xn, marker_yn,
                  inside_flags, vx1, vy1, esp, marker_vx, marker_vy, marker_spin,
                  None, None, 0)

    def advect_markers_and_rotate_stress_tensor(
            self, runge_kutta_order_max, timestep, inside_flags, marker_x, marker_y,
            marker_xn, marker_yn, vx1, vy1, esp, marker_vx=None, marker_vy=None,
            marker_spin=None, marker_sxx=None, marker_sxy=None) -> None:
        """Advection.jl:109-114 fused: RK4 velocity/spin, location, stress rotation."""
        self._rk4(runge_kutta_order_max, timestep, marker_x, marker_y, marker_xn, marker_yn,
                  inside_flags, vx1, vy1, esp, marker_vx, marker_vy, marker_spin,
                  marker_sxx, marker_sxy, 1)

    RK4_EXACT, RK4_TOLERANCE = 0, 1

    def set_rk4_arithmetic(self, mode: int) -> None:
        """RK4_EXACT: the reference's FP64 evaluation order, bit-identical;
        RK4_TOLERANCE: FMAs in nominal uniform cell coordinates with the grid
        patch of each wavefront staged in LDS (used when both axes are uniform
        to 1e-9 of a step)."""
        check(lib.ebg_rk4_set_arithmetic(self._h, int(mode)), "ebg_rk4_set_arithmetic")

    def set_rk4_tiling(self, on: bool) -> None:
        """Tolerance arithmetic only: True (default) serves the stage reads of
        each wavefront from an LDS copy of its grid patch, False reads global
        memory; the results are identical."""
        check(lib.ebg_rk4_set_tiling(self._h, int(bool(on))), "ebg_rk4_set_tiling")

    SCATTER_SERIAL_ORDER, SCATTER_CELL_ORDER = 0, 1

    def set_scatter_order(self, order: int) -> None:
        """SCATTER_SERIAL_ORDER (default): node sums in the reference's serial
        marker order, bit-identical; SCATTER_CELL_ORDER: per-cell partial sums
        in marker order added per node in a fixed order (deterministic,
        within a few ulps of the sum of the absolute terms)."""
        check(lib.ebg_scatter_set_order(self._h, int(order)), "ebg_scatter_set_order")

    def set_stokes_graphs(self, on: bool) -> None:
        """True (default): the 2-D preconditioner's V-cycle is replayed from a
        captured HIP graph; False launches its kernels one by one (identical
        results)."""
        check(lib.ebg_stokes_set_graphs(self._h, int(bool(on))), "ebg_stokes_set_graphs")

    def rk4_tolerance_kernel_available(self) -> bool:
        return lib.ebg_rk4_tolerance_kernel_available(self._h) == 1

    def calculate_spin(self, vx1, vy1, esp) -> None:
        """GridStrainRate.calculate_spin! (GridStrainRate.jl:161-182)."""
        g = self._need_grid()
        ny, nx = g.ynum, g.xnum
        check(lib.ebg_calculate_spin(self._h, ptr(vx1, np.float64, (ny + 1) * nx),
                                     ptr(vy1, np.float64, ny * (nx + 1)),
                                     ptr(esp, np.float64, ny * nx)), "ebg_calculate_spin")
