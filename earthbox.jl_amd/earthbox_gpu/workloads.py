"""Seeded synthetic inputs for the hot path (markers, velocity fields, models).

All generators use numpy.random.default_rng(seed) so the GPU path, the CPU
oracle and the benchmark see the same bits.
"""
from __future__ import annotations

import numpy as np

from .grid import StaggeredGrid2D


def marker_positions(g: StaggeredGrid2D, nmx: int, nmy: int, seed: int,
                     jitter: bool = True) -> tuple[np.ndarray, np.ndarray]:
    """Markers nmx x nmy per cell, ordered x-major then y as the reference
    (markers/coordinates/init_manager/types/MarkersRandomized.jl:5-38 with
    mxnum=(xnum-1)*nmx, mxstep=xsize/mxnum, ModelDataContainer.jl:355-392).

    Coordinates are i*step - step/2 + (r - 0.5)*step.  Unlike the reference,
    which reuses one random draw for both axes, x and y use independent draws.
    """
    mxnum = (g.xnum - 1) * nmx
    mynum = (g.ynum - 1) * nmy
    mxstep = g.xsize / mxnum
    mystep = g.ysize / mynum
    rng = np.random.default_rng(seed)
    ix = np.repeat(np.arange(1, mxnum + 1, dtype=np.float64), mynum)
    iy = np.tile(np.arange(1, mynum + 1, dtype=np.float64), mxnum)
    if jitter:
        rx = rng.random(mxnum * mynum)
        ry = rng.random(mxnum * mynum)
    else:
        rx = np.full(mxnum * mynum, 0.5)
        ry = np.full(mxnum * mynum, 0.5)
    x = ix * mxstep - mxstep / 2.0 + (rx - 0.5) * mxstep
    y = iy * mystep - mystep / 2.0 + (ry - 0.5) * mystep
    return x, y


def random_velocity_fields(g: StaggeredGrid2D, seed: int, vmax: float = 1e-9):
    """Smooth-ish random vx1 ((ynum+1) x xnum) and vy1 (ynum x (xnum+1)),
    Fortran-ordered, used for advection parity tests."""
    rng = np.random.default_rng(seed)
    ny, nx = g.ynum, g.xnum
    def field(shape):
        base = rng.standard_normal(shape)
        # cheap smoothing so neighbouring nodes correlate
        for _ in range(2):
            base = (base + np.roll(base, 1, 0) + np.roll(base, 1, 1)) / 3.0
        return np.asfortranarray(vmax * base / max(1e-300, np.abs(base).max()))
    return field((ny + 1, nx)), field((ny, nx + 1))


def solid_body_rotation_fields(g: StaggeredGrid2D, omega: float):
    """Rigid rotation about the box centre sampled on the staggered nodes
    (the closed-form field of kinematics/Kinematics.jl:14-43):
    vx = -omega*(y - yc), vy = omega*(x - xc)."""
    xc = 0.5 * (g.gridx_b[0] + g.gridx_b[-1])
    yc = 0.5 * (g.gridy_b[0] + g.gridy_b[-1])
    vx1 = np.asfortranarray(-omega * (g.gridy_vx[:, None] - yc) * np.ones((1, g.xnum)))
    vy1 = np.asfortranarray(omega * (g.gridx_vy[None, :] - xc) * np.ones((g.ynum, 1)))
    return vx1, vy1


def _rhs_viscous(g: StaggeredGrid2D, rho1: np.ndarray, rho1_vy: np.ndarray, gravity_y: float,
                 stabilization: int) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
    """RX1, RY1, RC1 for a purely viscous, stress-free state (gravity only),
    i.e. solver/rhs/StokesContinuityRhs.jl:30-190 with sxx0 = sxy0 = 0 and
    gravity_x = 0."""
    ny, nx = g.ynum, g.xnum
    RX1 = np.zeros((ny + 1, nx), order="F")
    RY1 = np.zeros((ny, nx + 1), order="F")
    RC1 = np.zeros((ny - 1, nx - 1), order="F")
    # RY1[i, j] for i in 2..ynum-1, j in 2..xnum (1-based)
    if stabilization == 1:
        RY1[1:ny - 1, 1:nx] = -gravity_y * rho1_vy[1:ny - 1, 1:nx]
    else:
        RY1[1:ny - 1, 1:nx] = -gravity_y * (rho1[1:ny - 1, 1:nx] + rho1[1:ny - 1, 0:nx - 1]) / 2.0
    # RX1 = -gravity_x*(...) = 0 written as -0.0*(...) by the reference; zeros suffice.
    return RX1, RY1, RC1


def falling_block(xnum: int = 101, ynum: int = 101, size: float = 1.0e6,
                  block: float = 1.0e5, eta_bg: float = 1e21, eta_block: float = 1e27,
                  rho_bg: float = 3200.0, rho_block: float = 3300.0, gravity_y: float = 10.0,
                  stabilization: int = 0, timestep: float = 0.0):
    """Nodal falling-block Stokes problem (free-slip box, centred square block):
    viscosity and density sampled at the staggered nodes."""
    from .grid import uniform_grid
    from .stokes_problem import StokesProblem, free_slip_bcs
    g = uniform_grid(xnum, ynum, size, size)
    xc = yc = 0.5 * size
    def inside(X, Y):
        return (np.abs(X - xc) <= block / 2) & (np.abs(Y - yc) <= block / 2)
    Xb, Yb = np.meshgrid(g.gridx_b, g.gridy_b)
    Xp, Yp = np.meshgrid(g.gridx_pr, g.gridy_pr)
    Xv, Yv = np.meshgrid(g.gridx_vy, g.gridy_b)
    etas = np.where(inside(Xb, Yb), eta_block, eta_bg)
    etan = np.where(inside(Xp, Yp), eta_block, eta_bg)
    rho1 = np.where(inside(Xb, Yb), rho_block, rho_bg)
    rho1_vy = np.where(inside(Xv, Yv), rho_block, rho_bg)
    RX1, RY1, RC1 = _rhs_viscous(g, rho1, rho1_vy, gravity_y, stabilization)
    return StokesProblem(grid=g, etas0=etas, etan0=etan, rho1_vy=rho1_vy, RX1=RX1, RY1=RY1,
                         RC1=RC1, bcs=free_slip_bcs(g), pressure_bc_mode=0, pressure_bc=0.0,
                         timestep=timestep, gravity_y=gravity_y,
                         iuse_interface_stabilization=stabilization)


def layered_random(xnum: int = 1001, ynum: int = 501, xsize: float = 1.5e6, ysize: float = 6.0e5,
                   seed: int = 2, corr_len: float = 2.0e4, gravity_y: float = 10.0,
                   log_eta_range=(18.0, 25.0), amplitude: float = 0.5):
    """Layered log10-viscosity with a Gaussian random perturbation (correlation
    length corr_len, FFT-filtered, standard deviation `amplitude` decades)
    clipped to log_eta_range, density
    2700-3300 tied to the layers; free-slip walls."""
    from .grid import uniform_grid
    from .stokes_problem import StokesProblem, free_slip_bcs
    g = uniform_grid(xnum, ynum, xsize, ysize)
    rng = np.random.default_rng(seed)
    # perturbation on a fine uniform lattice covering the domain
    nxf, nyf = 2 * xnum, 2 * ynum
    noise = rng.standard_normal((nyf, nxf))
    kx = np.fft.fftfreq(nxf, d=xsize / nxf)
    ky = np.fft.fftfreq(nyf, d=ysize / nyf)
    K2 = ky[:, None] ** 2 + kx[None, :] ** 2
    filt = np.exp(-0.5 * (2 * np.pi) ** 2 * K2 * corr_len ** 2)
    field = np.real(np.fft.ifft2(np.fft.fft2(noise) * filt))
    field /= field.std()
    layer_depths = np.array([0.0, 0.05, 0.20, 0.45, 1.01]) * ysize
    layer_logeta = np.array([23.0, 21.5, 24.0, 19.5])
    layer_rho = np.array([2700.0, 2900.0, 3300.0, 3250.0])
    def sample(X, Y):
        ix = np.clip((X / xsize * nxf).astype(int), 0, nxf - 1)
        iy = np.clip((Y / ysize * nyf).astype(int), 0, nyf - 1)
        lay = np.clip(np.searchsorted(layer_depths, Y, side="right") - 1, 0, 3)
        loge = np.clip(layer_logeta[lay] + amplitude * field[iy, ix], *log_eta_range)
        return 10.0 ** loge, layer_rho[lay] + 30.0 * field[iy, ix]
    Xb, Yb = np.meshgrid(g.gridx_b, g.gridy_b)
    Xp, Yp = np.meshgrid(g.gridx_pr, g.gridy_pr)
    Xv, Yv = np.meshgrid(g.gridx_vy, g.gridy_b)
    etas, rho1 = sample(Xb, Yb)
    etan, _ = sample(Xp, Yp)
    _, rho1_vy = sample(Xv, Yv)
    RX1, RY1, RC1 = _rhs_viscous(g, rho1, rho1_vy, gravity_y, 0)
    return StokesProblem(grid=g, etas0=etas, etan0=etan, rho1_vy=rho1_vy, RX1=RX1, RY1=RY1,
                         RC1=RC1, bcs=free_slip_bcs(g), pressure_bc_mode=0, pressure_bc=0.0,
                         timestep=0.0, gravity_y=gravity_y, iuse_interface_stabilization=0)


def _axis_kinds(n: int, size: float) -> dict:
    """Node coordinates of one uniform axis by kind: basic, staggered with
    ghosts, pressure (cell centres), shear centre."""
    h = size / (n - 1)
    b = np.concatenate([[0.0], np.cumsum(np.full(n - 1, h))])
    mid = 0.5 * (b[1:] + b[:-1])
    v = np.concatenate([[b[0] - h / 2], mid, [b[-1] + h / 2]])
    return {"b": b, "v": v, "p": mid, "s": mid}


def spheres3d(n: int = 257, size: float = 1.0, nspheres: int = 8, seed: int = 3,
              radius=(0.05, 0.12), log_contrast=(2.0, 6.0), drho: float = 0.01,
              gravity: float = 1.0, device=None) -> dict:
    """Seeded 3-D inclusion problem on an n^3-node unit cube: nspheres spheres
    of radius in `radius`, viscosity 10^U(log_contrast) times the background
    of 1, density 1 +- drho on a background of 1, body force -gravity*rho in z
    (RZ), free-slip walls.  Fields are column-major torch tensors with the
    3-D multigrid's level shapes (on `device`, default the CPU)."""
    import torch
    rng = np.random.default_rng(seed)
    cen = rng.uniform(0.2, 0.8, size=(nspheres, 3)) * size
    rad = rng.uniform(*radius, size=nspheres) * size
    eta_s = 10.0 ** rng.uniform(*log_contrast, size=nspheres)
    sgn = rng.choice([-1.0, 1.0], size=nspheres)
    ax = _axis_kinds(n, size)
    dev = torch.device("cpu") if device is None else device
    kinds = {"vx": "vbv", "vy": "bvv", "vz": "vvb", "pr": "ppp", "etaxy": "bbs",
             "etaxz": "sbb", "etayz": "bsb"}

    def coords(field):
        ky, kx, kz = kinds[field]
        t = lambda k: torch.as_tensor(ax[k], dtype=torch.float64, device=dev)
        return t(ky), t(kx), t(kz)

    def inside(field, s):
        y, x, z = coords(field)
        cy, cx, cz = cen[s][1], cen[s][0], cen[s][2]
        d2 = ((z - cz) ** 2)[:, None, None] + ((x - cx) ** 2)[None, :, None] + ((y - cy) ** 2)[None, None, :]
        return (d2 <= rad[s] ** 2).permute(2, 1, 0)

    def julia_zeros(field):
        y, x, z = coords(field)
        return torch.zeros((z.numel(), x.numel(), y.numel()), dtype=torch.float64, device=dev).permute(2, 1, 0)

    out = {"dims": (n, n, n, size, size, size), "spheres": {"center": cen, "radius": rad,
                                                             "eta": eta_s, "drho_sign": sgn}}
    for f, name in (("pr", "etan"), ("etaxy", "etaxy"), ("etaxz", "etaxz"), ("etayz", "etayz")):
        eta = julia_zeros(f) + 1.0
        for s in range(nspheres):
            eta = torch.where(inside(f, s), torch.clamp(eta, min=float(eta_s[s])), eta)
        out[name] = eta
    rho = julia_zeros("vz") + 1.0
    for s in range(nspheres):
        rho = rho + inside("vz", s).to(torch.float64) * (sgn[s] * drho)
    out["RZ"] = -gravity * rho
    out["RX"] = julia_zeros("vx")
    out["RY"] = julia_zeros("vy")
    out["RC"] = julia_zeros("pr")
    return out


def _sinker_axis(n: int, size: float) -> dict:
    """One uniform axis as the multigrid's level grid builds it: basic nodes
    by running sum of the average step (Spacing.jl:12, Initialization.jl:27-43),
    cell centres from the pressure steps (PressureGridTools.jl:34-86) and the
    shear-centre nodes (SxyGridTools.jl:30-40)."""
    h = size / (n - 1)
    b = np.zeros(n)
    for i in range(1, n):
        b[i] = b[i - 1] + h
    stp = b[1:] - b[:-1]
    stp_p = (b[2:] - b[:-2]) / 2.0
    p = np.zeros(n - 1)
    p[0] = (b[0] + b[1]) / 2.0
    for i in range(1, n - 1):
        p[i] = p[i - 1] + stp_p[i - 1]
    s = b[:-1] + stp / 2.0
    s[0] = b[0] + stp[0] / 2.0
    s[-1] = b[-1] - stp[-1] / 2.0
    return {"b": b, "p": p, "s": s, "stp_p": stp_p}


# The reference's own multigrid demo (multigrid_tests/stokes_sinker/StokesSinker.jl:29-160,
# model_data/ModelParametersManager.jl): a dense, stiff cube (normalised
# 0.4..0.6 in every direction) sinking in a 100 km box.
SINKER = dict(density_block=3100.0, density_medium=3000.0, viscosity_block=1e26,
              viscosity_medium=1e20, gravity=9.81, pressure_bc=0.0, block=(0.4, 0.6))
# V-cycle and relaxation settings of the demo's runs (StokesSinker.jl:40-101)
SINKER_RUNS = {
    "3d_smoke": dict(n=17, nvcycles=30, nvcycles_viscosity_jump=15, nviscosity_jumps=11.475031046555408,
                     finest_sweeps=2, convergence_criterion=1e-4, cap=12),
    "3d": dict(n=97, nvcycles=250, nvcycles_viscosity_jump=15, nviscosity_jumps=11.475031046555408,
               finest_sweeps=5, convergence_criterion=1e-20, cap=48),
    "2d": dict(n=97, nvcycles=500, nvcycles_viscosity_jump=17, nviscosity_jumps=12.0, finest_sweeps=5,
               convergence_criterion=1e-5, cap=48),
}
SINKER_RELAXATION = dict(theta_s=0.9, relax_v=1.0, theta_c=0.3, relax_p=1.0)


def _in_block(*c):
    lo, hi = SINKER["block"]
    m = True
    for a in c:
        m = m & (lo <= a) & (a <= hi)
    return m


def stokes_sinker3d(n: int = 97, size: float = 1.0e5, viscosity_block: float | None = None) -> dict:
    """Initial state of the 3-D Stokes sinker on an n^3-node cube
    (StokesSinker.jl:111-139 with initial_conditions/Density.jl:11-25, 45-58,
    Viscosity.jl:11-40, 99-160, RightParts.jl:10-36, 62-71, Pressure.jl:10-40):
    basic-node density, the four viscosity arrays, RX = RZ = RC = 0,
    RY = -g * (mean of the four basic densities around the vy node), the
    hydrostatic pressure guess integrated down from pressure_bc, zero
    velocity guesses, and the residual scales of ScalingFactors.jl:3-18.
    Arrays are numpy column-major with the level shapes (y, x, z)."""
    P = dict(SINKER)
    if viscosity_block is not None:
        P["viscosity_block"] = viscosity_block
    a = _sinker_axis(n, size)
    g = P["gravity"]
    F = lambda sh: np.zeros(sh, order="F")

    def field(ky, kx, kz, val_in, val_out):
        y, x, z = a[ky] / size, a[kx] / size, a[kz] / size
        m = _in_block(y[:, None, None], x[None, :, None], z[None, None, :])
        return np.asfortranarray(np.where(m, val_in, val_out))

    rho = field("b", "b", "b", P["density_block"], P["density_medium"])
    vb, vm = P["viscosity_block"], P["viscosity_medium"]
    out = {"rho": rho, "etan": field("p", "p", "p", vb, vm), "etaxy": field("b", "b", "s", vb, vm),
           "etaxz": field("s", "b", "b", vb, vm), "etayz": field("b", "s", "b", vb, vm)}
    ny = nx = nz = n
    RY = F((ny, nx + 1, nz + 1))
    avg = (((rho[:, 1:, 1:] + rho[:, :-1, 1:]) + rho[:, 1:, :-1]) + rho[:, :-1, :-1]) / 4.0
    RY[1:ny - 1, 1:nx, 1:nz] = -g * avg[1:ny - 1]
    out.update(RX=F((ny + 1, nx, nz + 1)), RY=RY, RZ=F((ny + 1, nx + 1, nz)), RC=F((ny - 1, nx - 1, nz - 1)))
    pr = F((ny - 1, nx - 1, nz - 1))
    pr[0] = P["pressure_bc"]
    for i in range(1, ny - 1):
        pr[i] = pr[i - 1] + a["stp_p"][i - 1] * g * avg[i]
    out.update(pr=pr, vx=F((ny + 1, nx, nz + 1)), vy=F((ny, nx + 1, nz + 1)), vz=F((ny + 1, nx + 1, nz)))
    xstpavg = size / (n - 1)
    out["stokesscale"] = P["density_block"] * g
    out["continscale"] = P["density_block"] * g / P["viscosity_medium"] * xstpavg
    out["dims"] = (n, n, n, size, size, size)
    out["pressure_bc"] = P["pressure_bc"]
    return out


def stokes_sinker2d(n: int = 97, size: float = 1.0e5, viscosity_block: float | None = None) -> dict:
    """The 2-D Stokes sinker (StokesSinker.jl:84-101, 140-157 with
    Density.jl:27-43, 60-72, Viscosity.jl:42-60, 75-97, 99-110,
    RightParts.jl:38-60, 73-81, Pressure.jl:42-70): density and etas on basic
    nodes, etan on cell centres, RX = RC = 0, RY = -g * (mean of the two basic
    densities left and right of the vy node), hydrostatic pressure guess."""
    P = dict(SINKER)
    if viscosity_block is not None:
        P["viscosity_block"] = viscosity_block
    a = _sinker_axis(n, size)
    g = P["gravity"]

    def field(ky, kx, val_in, val_out):
        y, x = a[ky] / size, a[kx] / size
        return np.asfortranarray(np.where(_in_block(y[:, None], x[None, :]), val_in, val_out))

    rho = field("b", "b", P["density_block"], P["density_medium"])
    vb, vm = P["viscosity_block"], P["viscosity_medium"]
    ny = nx = n
    out = {"rho": rho, "etas": field("b", "b", vb, vm), "etan": field("p", "p", vb, vm)}
    RY = np.zeros((ny, nx + 1), order="F")
    avg = (rho[:, 1:] + rho[:, :-1]) / 2.0
    RY[1:ny - 1, 1:nx] = -g * avg[1:ny - 1]
    out.update(RX=np.zeros((ny + 1, nx), order="F"), RY=RY, RC=np.zeros((ny - 1, nx - 1), order="F"))
    pr = np.zeros((ny - 1, nx - 1), order="F")
    pr[0] = P["pressure_bc"]
    for i in range(1, ny - 1):
        pr[i] = pr[i - 1] + a["stp_p"][i - 1] * g * avg[i]
    out.update(pr=pr, vx=np.zeros((ny + 1, nx), order="F"), vy=np.zeros((ny, nx + 1), order="F"))
    xstpavg = size / (n - 1)
    out["stokesscale"] = P["density_block"] * g
    out["continscale"] = P["density_block"] * g / P["viscosity_medium"] * xstpavg
    out["dims"] = (n, n, size, size)
    out["pressure_bc"] = P["pressure_bc"]
    return out


RIFT_LAYERS = (  # bottom depth (m), flow viscosity (Pa s), density (kg/m3)
    (2.0e4, 1e24, 2700.0),   # upper crust
    (3.5e4, 1e21, 2900.0),   # lower crust
    (1.0e5, 1e23, 3300.0),   # lithospheric mantle
    (np.inf, 1e19, 3250.0),  # asthenosphere
)


def rift2d(xnum: int = 2001, ynum: int = 801, xsize: float = 5.0e5, ysize: float = 1.6e5,
           seed: int = 5, extension: float = 0.5 / 100 / 3.15576e7, gravity_y: float = 9.81,
           timestep: float = 3.15576e12):
    """Synthetic visco-elasto-plastic rift state on 2001 x 801 nodes over
    500 x 160 km: four layers (upper crust, lower crust, lithospheric mantle,
    asthenosphere) with flow viscosity clipped to [1e18, 1e25], shear modulus
    3e10, cohesion 20 MPa and friction 30 degrees (5 MPa in a seeded weak
    patch of the upper crust), the sides pulled apart at +-0.5 cm/yr with
    the outflow replaced through the bottom, free slip elsewhere.  Returns
    (template StokesProblem, grid state for the Picard loop as numpy arrays,
    plastic-yield scalars)."""
    from .grid import uniform_grid
    from .stokes_problem import StokesProblem, free_slip_bcs
    g = uniform_grid(xnum, ynum, xsize, ysize)
    rng = np.random.default_rng(seed)

    def layer_props(Y):
        eta = np.empty_like(Y)
        rho = np.empty_like(Y)
        top = 0.0
        for bottom, e, r in RIFT_LAYERS:
            m = (Y >= top) & (Y < bottom)
            eta[m] = e
            rho[m] = r
            top = bottom
        return eta, rho
    Xb, Yb = np.meshgrid(g.gridx_b, g.gridy_b)
    Xp, Yp = np.meshgrid(g.gridx_pr, g.gridy_pr)
    Xv, Yv = np.meshgrid(g.gridx_vy, g.gridy_b)
    eta_b, rho_b = layer_props(Yb)
    eta_p, _ = layer_props(Yp)
    _, rho_v = layer_props(Yv)
    # log-normal perturbation of the flow viscosity, +-0.3 decades, seeded
    eta_b = np.clip(eta_b * 10 ** (0.3 * rng.standard_normal(eta_b.shape)), 1e18, 1e25)
    eta_p = np.clip(eta_p * 10 ** (0.3 * rng.standard_normal(eta_p.shape)), 1e18, 1e25)
    xc = xsize / 2 + 1e4 * (rng.random() - 0.5)
    weak = (np.abs(Xb - xc) < 5e3) & (Yb > 5e3) & (Yb < 1.5e4)
    F = np.asfortranarray
    ny, nx = ynum, xnum
    state = {"etas1": F(eta_b), "eta_flow": F(eta_b.copy()), "rho1": F(rho_b), "rho1_vy": F(rho_v),
             "etan1": F(eta_p), "mus1": F(np.full((ny, nx), 3e10)), "mun1": F(np.full((ny - 1, nx - 1), 3e10)),
             "sxy1": F(np.zeros((ny, nx))), "sxx1": F(np.zeros((ny - 1, nx - 1))),
             "plastics": F(np.zeros((ny, nx))), "plasticn": F(np.zeros((ny - 1, nx - 1))),
             "plastic_yield": F(np.zeros((ny, nx))), "cohesion_grid": F(np.where(weak, 5e6, 2e7)),
             "fric_degrees_grid": F(np.full((ny, nx), 30.0)), "dilatation_grid": F(np.zeros((ny - 1, nx - 1))),
             "eii_plastic_pressure": F(np.zeros((ny - 1, nx - 1))), "yield_error": F(np.zeros((ny, nx)))}
    bcs = free_slip_bcs(g, vx_left=-extension, vx_right=extension,
                        vy_bottom=-2.0 * extension * ysize / xsize)
    RX1, RY1, RC1 = _rhs_viscous(g, rho_b, rho_v, gravity_y, 0)
    template = StokesProblem(grid=g, etas0=state["etas1"], etan0=state["etan1"], rho1_vy=state["rho1_vy"],
                             RX1=RX1, RY1=RY1, RC1=RC1, bcs=bcs, pressure_bc_mode=0, pressure_bc=0.0,
                             timestep=timestep, gravity_y=gravity_y, iuse_interface_stabilization=0)
    params = dict(timestep=timestep, timesum=timestep, ntimestep=1, viscosity_min=1e18, viscosity_max=1e25,
                  iuse_fluid_pressure_for_yield=0)
    return template, state, params


# Four creeping layers of the rifting model: dislocation creep (power law,
# n = 3-3.5) with diffusion and Peierls creep in the lithospheric mantle, and
# Drucker-Prager parameters with linear strain weakening (cohesion 20 -> 4
# MPa, friction 30 -> 5 degrees over plastic strain 0 -> 1).  Columns follow
# the reference's materials.arrays (1-based): mat_flow 1 viscosity, 2
# A_dis (1/s/MPa^n), 3 n, 4 Ea (kJ/mol), 5 Va (J/MPa/mol), 6-9 Peierls
# (A, m, n, sigma_p MPa), 10-12 diffusion (A, Ea, Va); mat_plastic cohesion
# initial/final, sine friction initial/final, plastic strain initial/final.
RIFT_MATERIALS = {
    # name: (bottom depth m, density kg/m3 at 273 K, mat_flow row, mat_plastic row)
    "upper_crust": (2.0e4, 2700.0, (0.0, 5.0e-6, 3.0, 190.0, 0.0, 0.0, 1.0, 2.0, 9100.0, 0.0, 0.0, 0.0),
                    (2e7, 4e6, 0.5, 0.0872, 0.0, 1.0)),
    "lower_crust": (3.5e4, 2900.0, (0.0, 3.3e-4, 3.2, 238.0, 0.0, 0.0, 1.0, 2.0, 9100.0, 0.0, 0.0, 0.0),
                    (2e7, 4e6, 0.5, 0.0872, 0.0, 1.0)),
    "lithospheric_mantle": (1.0e5, 3300.0,
                            (0.0, 2.5e4, 3.5, 532.0, 8.0, 5.7e11, 1.0, 2.0, 9100.0, 1.5e-9, 300.0, 4.0),
                            (2e7, 4e6, 0.6, 0.1, 0.0, 1.0)),
    "asthenosphere": (np.inf, 3300.0, (0.0, 2.0e3, 3.5, 471.0, 4.0, 0.0, 1.0, 2.0, 9100.0, 0.0, 0.0, 0.0),
                      (2e7, 4e6, 0.6, 0.1, 0.0, 1.0)),
}


def rift2d_markers(xnum: int = 2001, ynum: int = 801, nm: int = 5, xsize: float = 5.0e5, ysize: float = 1.6e5,
                   seed: int = 5, extension: float = 0.5 / 100 / 3.15576e7, gravity_y: float = 9.81,
                   timestep: float = 3.15576e12):
    """Marker-based visco-elasto-plastic rifting model on xnum x ynum nodes over
    xsize x ysize (2001 x 801 over 500 x 160 km, 5 x 5 markers per cell:
    40M markers): four layers (upper crust 0-20 km, lower crust 20-35 km,
    lithospheric mantle 35-100 km, asthenosphere) whose boundaries are
    perturbed by a seeded +-1 km undulation; a seeded geotherm (273 K at the
    surface, 823 K at the Moho, 1573 K at 100 km, then 0.4 K/km, plus a
    seeded +-10 K perturbation); densities with thermal expansion 3e-5/K; the
    creep laws and strain-weakening Drucker-Prager parameters of
    RIFT_MATERIALS; shear modulus 3e10; a seeded 10 km wide weak seed of
    plastic strain 0.5-1.0 in the upper crust; sides pulled apart at +-0.5
    cm/yr with the outflow replaced through the bottom, free slip elsewhere.
    Markers start stress-free with the background extension strain rate and
    lithostatic pressure (the synthetic initial state).

    Returns (template StokesProblem with the grid, boundary conditions and
    gravity; marker arrays (numpy); material tables (numpy); scalars)."""
    from .grid import uniform_grid
    from .stokes_problem import StokesProblem, free_slip_bcs
    g = uniform_grid(xnum, ynum, xsize, ysize)
    rng = np.random.default_rng(seed)
    x, y = marker_positions(g, nm, nm, seed)
    n = x.size
    names = list(RIFT_MATERIALS)
    # layer boundaries with a seeded undulation (four Fourier modes, +-1 km)
    k = np.arange(1, 5)
    amp = rng.uniform(-250.0, 250.0, (3, 4))
    ph = rng.uniform(0, 2 * np.pi, (3, 4))
    matid = np.full(n, 4, dtype=np.int16)
    for li in (2, 1, 0):
        bottom = RIFT_MATERIALS[names[li]][0]
        und = (amp[li][:, None] * np.sin(2 * np.pi * k[:, None] * x[None, :] / xsize + ph[li][:, None])).sum(0)
        matid[y < bottom + und] = li + 1
    # geotherm (piecewise linear) with a seeded perturbation
    TK = np.interp(y, [0.0, 3.5e4, 1.0e5, 1e9], [273.0, 823.0, 1573.0, 1573.0 + 0.4e-3 * (1e9 - 1e5)])
    TK += 10.0 * np.sin(2 * np.pi * x / xsize * 3 + rng.uniform(0, 2 * np.pi)) * np.exp(-y / 5e4)
    rho0 = np.array([RIFT_MATERIALS[nm_][1] for nm_ in names])
    rho = rho0[matid - 1] * (1.0 - 3e-5 * (TK - 273.0))
    # weak seed of plastic strain in the upper crust
    xc = xsize / 2 + 1e4 * (rng.random() - 0.5)
    strain_plastic = np.where((np.abs(x - xc) < 5e3) & (y < 2.0e4), rng.uniform(0.5, 1.0, n), 0.0)
    eb = 2.0 * extension / xsize
    mats = {"mat_flow_type": np.ones(4, dtype=np.int64),
            "mat_flow": np.asfortranarray([RIFT_MATERIALS[nm_][2] for nm_ in names]),
            "mat_plastic": np.asfortranarray([RIFT_MATERIALS[nm_][3] for nm_ in names]),
            "mat_mu": np.full(4, 3e10)}
    markers = {"x": x, "y": y, "matid": matid, "TK": TK, "rho": rho,
               "pr": 3000.0 * gravity_y * np.clip(y, 0.0, None),
               "sxx": np.zeros(n), "sxy": np.zeros(n), "exx": np.full(n, eb), "exy": np.zeros(n),
               "sr_ratio": np.ones(n), "GII": np.zeros(n), "strain_plastic": strain_plastic,
               "strain_rate_plastic": np.zeros(n), "fric_ini": mats["mat_plastic"][matid - 1, 2].copy(),
               "cohesion": np.zeros(n), "fric": np.zeros(n), "eta_flow": np.zeros(n), "preexp": np.zeros(n),
               "eta": np.zeros(n), "pfailure": np.zeros(n, np.float32)}
    ny, nx = ynum, xnum
    bcs = free_slip_bcs(g, vx_left=-extension, vx_right=extension,
                        vy_bottom=-2.0 * extension * ysize / xsize)
    z = lambda s: np.zeros(s, order="F")
    template = StokesProblem(grid=g, etas0=z((ny, nx)), etan0=z((ny - 1, nx - 1)), rho1_vy=z((ny, nx + 1)),
                             RX1=z((ny + 1, nx)), RY1=z((ny, nx + 1)), RC1=z((ny - 1, nx - 1)), bcs=bcs,
                             pressure_bc_mode=0, pressure_bc=0.0, timestep=timestep, gravity_y=gravity_y,
                             iuse_interface_stabilization=0)
    params = dict(timestep=timestep, temperature_top=273.0, viscosity_min=1e18, viscosity_max=1e25,
                  powerlaw_stress_min=1e4, gas_constant=8.314, iuse_viscous_strain_soft=0, vsoftfac=1.0,
                  itype_global=1, marker_cell_displ_max=0.5, plastic_healing_rate=0.0,
                  iuse_fluid_pressure_for_yield=0, nglobal=10, tolerance_picard=1e-3)
    return template, markers, mats, params


# thermal constants of the rift layers: conductivity W/m/K, rho*Cp J/m3/K,
# radiogenic heat production W/m3, thermal expansivity 1/K
RIFT_THERMAL = {
    "upper_crust": (2.5, 2700.0 * 1000.0, 1.5e-6, 3e-5),
    "lower_crust": (2.1, 2900.0 * 1000.0, 0.4e-6, 3e-5),
    "lithospheric_mantle": (3.0, 3300.0 * 1000.0, 0.0, 3e-5),
    "asthenosphere": (3.0, 3300.0 * 1000.0, 0.0, 3e-5),
}


def rift2d_thermal(template, markers: dict, max_temp_change: float = 20.0):
    """Heat-equation inputs of the rifting model: per-marker conductivity,
    rho*Cp and adiabatic term (expansivity x temperature) from RIFT_THERMAL
    (held at their initial values: the temperature-dependent rock-property
    update is not part of this path), the material heat production table,
    thermal boundary conditions (273 K at the top, the unperturbed geotherm at
    the bottom, 1597 K at 160 km, no heat flux through the sides) and the heat-solver
    scalars.  Returns (marker arrays, mats, bc, params), numpy."""
    g = template.grid
    names = list(RIFT_THERMAL)
    tab = np.array([RIFT_THERMAL[k] for k in names])
    mid = markers["matid"].astype(np.int64) - 1
    therm = {"kt": tab[mid, 0].copy(), "rhocp": tab[mid, 1].copy(), "ha": tab[mid, 3] * markers["TK"]}
    nx, ny = g.xnum, g.ynum
    t_bottom = float(np.interp(g.ysize, [0.0, 3.5e4, 1.0e5, 1e9],
                               [273.0, 823.0, 1573.0, 1573.0 + 0.4e-3 * (1e9 - 1e5)]))
    btopt = np.zeros((nx, 2), order="F")
    bbottomt = np.zeros((nx, 2), order="F")
    btopt[:, 0] = 273.0
    bbottomt[:, 0] = t_bottom
    bleftt = np.zeros((ny, 2), order="F")
    brightt = np.zeros((ny, 2), order="F")
    bleftt[:, 1] = 1.0
    brightt[:, 1] = 1.0
    bc = {"btopt": btopt, "bbottomt": bbottomt, "bleftt": bleftt, "brightt": brightt}
    params = dict(iuse_heat=1, max_temp_change=max_temp_change, iuse_adiabatic_heating=0, iuse_shear_heating=1,
                  subgrid_diff_coef_temp=0.0, temperature_top=273.0)
    return therm, {"mat_hr": tab[:, 2].copy()}, bc, params
