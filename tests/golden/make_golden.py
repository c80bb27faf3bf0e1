"""Write the golden fixtures under tests/golden/ (run from the repo root):

    python tests/golden/make_golden.py

The reference ships no function-level vectors and cannot run here (no Julia),
so these fixtures are the oracle's outputs on tiny seeded inputs, frozen:
  rk4_9x7.npz       RK4 advect (markers, fields, dt) -> positions, velocities, spin
  stokes_9x7.npz    assembled system (COO in reference order), RHS and the
                    SuperLU solution of a 9 x 7-node free-slip system
  stokes_ext_17x11.npz  the same for prescribed boundary velocities with
                    interface stabilization
They pin the oracle against silent change (tests/test_golden.py) and give the
GPU tests fixed expected outputs.

stokes_sinker3d_97_stats.json is not written by this script: it holds the
final-solution statistics the reference's authors recorded in the source of
its 97^3 Stokes sinker demo (multigrid_tests/stokes_sinker/StokesSinker.jl:60-64),
transcribed as data."""
import pathlib
import sys

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
ROOT = HERE.parents[1]
sys.path[:0] = [str(ROOT), str(ROOT / "earthbox.jl_amd"), str(ROOT / "tests")]

import oracle  # noqa: E402
from earthbox_gpu import workloads  # noqa: E402
from earthbox_gpu.grid import uniform_grid  # noqa: E402
from stokes_cases import random_case  # noqa: E402


def rk4_fixture():
    g = uniform_grid(9, 7, 1.2e5, 9.0e4)
    x, y = workloads.marker_positions(g, 3, 3, 11)
    vx1, vy1 = workloads.random_velocity_fields(g, 12, vmax=1e-9)
    esp = oracle.calculate_spin(g, vx1, vy1)
    dt = 0.5 * g.xstpavg / 1e-9
    flags = oracle.inside_flags(x, y, g.limits)
    xn, yn = oracle.cell_indices_basic(x, y, flags, g)
    ex, ey, evx, evy, esp_m, _, _ = oracle.advect_markers_and_rotate_stress_tensor(
        4, dt, g, x, y, xn, yn, flags, vx1, vy1, esp)
    np.savez(HERE / "rk4_9x7.npz", xnum=g.xnum, ynum=g.ynum, xsize=g.xsize, ysize=g.ysize,
             x=x, y=y, vx1=vx1, vy1=vy1, esp=esp, dt=dt, xn=xn, yn=yn, flags=flags,
             out_x=ex, out_y=ey, out_vx=evx, out_vy=evy, out_spin=esp_m)


def stokes_fixture(name, **case):
    pr = random_case(**case)
    Li, Lj, Lv, R = oracle.stokes_build(pr)
    S = oracle.direct_solve(pr.N, Li, Lj, Lv, R)
    np.savez(HERE / name, case=np.array(repr(case)), Li=Li, Lj=Lj, Lv=Lv, R=R, S=S)


if __name__ == "__main__":
    rk4_fixture()
    stokes_fixture("stokes_9x7.npz", nx=9, ny=7, seed=1)
    stokes_fixture("stokes_ext_17x11.npz", nx=17, ny=11, seed=3, bc="ext", stab=1)
    print("wrote", sorted(p.name for p in HERE.glob("*.npz")))
