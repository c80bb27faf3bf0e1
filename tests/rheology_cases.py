"""Material tables and seeded marker states for the marker rheology tests:
four creeping layers (upper crust, lower crust, lithospheric mantle with
Peierls and diffusion creep, asthenosphere), power-law exponents 3-3.5, and a
constant-viscosity material, with linear strain weakening."""
import numpy as np

NMAT = 5
# mat_flow columns (1-based in the reference): 1 viscosity, 2 A_dis (1/s/MPa^n),
# 3 n, 4 Ea (kJ/mol), 5 Va (J/MPa/mol), 6 A_peierls, 7 m, 8 n, 9 s_peierls (MPa),
# 10 A_diff, 11 Ea_diff, 12 Va_diff
MAT_FLOW = np.asfortranarray(np.array([
    [0.0, 5.0e-6, 3.0, 190.0, 0.0, 0.0, 1.0, 2.0, 9100.0, 0.0, 0.0, 0.0],
    [0.0, 3.3e-4, 3.2, 238.0, 0.0, 0.0, 1.0, 2.0, 9100.0, 0.0, 0.0, 0.0],
    [0.0, 2.5e4, 3.5, 532.0, 8.0, 5.7e11, 1.0, 2.0, 9100.0, 1.5e-9, 300.0, 4.0],
    [0.0, 2.0e3, 3.5, 471.0, 4.0, 0.0, 1.0, 2.0, 9100.0, 0.0, 0.0, 0.0],
    [1.0e21, 0.0, 1.0, 0.0, 0.0, 0.0, 1.0, 2.0, 9100.0, 0.0, 0.0, 0.0],
]))
MAT_FLOW_TYPE = np.array([1, 1, 1, 2, 0], dtype=np.int64)
# mat_plastic columns: cohesion initial/final, sine friction initial/final, strain initial/final
MAT_PLASTIC = np.asfortranarray(np.array([
    [2.0e7, 4.0e6, 0.5, 0.1, 0.0, 1.0],
    [2.0e7, 4.0e6, 0.5, 0.1, 0.5, 1.5],
    [1.0e7, 1.0e7, 0.6, 0.6, 0.0, 1.0],
    [1.0e6, 1.0e6, 0.0, 0.0, 0.0, 1.0],
    [0.0, 0.0, 0.0, 0.0, 0.0, 1.0],
]))
MAT_MU = np.array([3e10, 3e10, 6.7e10, 6.7e10, 1e20])
MATS = dict(mat_flow_type=MAT_FLOW_TYPE, mat_flow=MAT_FLOW, mat_plastic=MAT_PLASTIC, mat_mu=MAT_MU)
PARAMS = dict(timestep=3.15576e12, temperature_top=273.0, viscosity_min=1e18, viscosity_max=1e25,
              powerlaw_stress_min=1e4, gas_constant=8.314, iuse_viscous_strain_soft=1, vsoftfac=10.0,
              itype_global=1, ntimestep=3)


def marker_states(n, seed, melt_damage=False):
    rng = np.random.default_rng(seed)
    m = {"matid": rng.integers(1, NMAT + 1, n).astype(np.int16),
         "flags": np.where(rng.random(n) < 0.02, -1, 1).astype(np.int8),
         "strain_plastic": rng.uniform(0.0, 2.0, n), "fric_ini": rng.uniform(0.4, 0.6, n),
         "sxx": rng.uniform(-3e7, 3e7, n), "sxy": rng.uniform(-3e7, 3e7, n),
         "exx": rng.uniform(-3e-14, 3e-14, n), "exy": rng.uniform(-3e-14, 3e-14, n),
         "TK": rng.uniform(250.0, 1750.0, n), "pr": rng.uniform(0.0, 4e9, n),
         "GII": rng.uniform(0.0, 2.0, n), "sr_ratio": rng.uniform(0.5, 2.0, n),
         "eta": np.where(rng.random(n) < 0.2, 0.0, 10 ** rng.uniform(19, 23, n)),
         "cohesion": np.full(n, -1.0), "fric": np.full(n, -1.0), "eta_flow": np.full(n, -1.0),
         "preexp": np.full(n, -1.0)}
    if melt_damage:
        m["melt_damage"] = rng.uniform(1.0, 3.0, n)
    return m
