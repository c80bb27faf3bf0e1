"""Marker states for the marker plastic-yielding step of the Picard loop over
markers: a mix of markers far above, near and below their yield stress, with
zero-stress, zero-strength, outside, already-failed and mobile-wall markers."""
import numpy as np

MAT_MU = np.array([3e10, 2.5e10, 6.7e10, 1e10])
# material ids of the sandbox domains in these cases
AIR, WALL, PLATE = 2, 3, 4

PARAMS = {"ntimestep": 3, "timestep": 3.15576e11, "y_sealevel": 1.0e4, "viscosity_min": 1e18,
          "viscosity_max": 1e25, "yield_stress_min": 0.0, "yield_stress_max": 1e32,
          "iuse_fluid_pressure_for_yield": 0, "no_yielding_in_mobile_wall": 0,
          "matid_sticky_air": -1, "matid_mobile_wall": -1, "matid_plate_extension": -1}


def markers(n, seed):
    """Random markers: viscosities 1e19..1e25 and strain rates 1e-16..1e-13
    put both forecasts on either side of Mohr-Coulomb yield stresses of a
    few MPa to a few hundred MPa."""
    rng = np.random.default_rng(seed)
    flags = np.ones(n, np.int8)
    flags[rng.random(n) < 0.05] = -1
    m = {
        "flags": flags, "y": rng.uniform(0.0, 1e5, n),
        "matid": rng.integers(1, MAT_MU.size + 1, n).astype(np.int16), "mat_mu": MAT_MU,
        "cohesion": rng.uniform(1e6, 4e7, n), "fric": rng.uniform(0.0, 0.6, n),
        "exx": rng.choice([-1, 1], n) * 10 ** rng.uniform(-16, -13, n),
        "exy": rng.choice([-1, 1], n) * 10 ** rng.uniform(-16, -13, n),
        "sr_ratio": rng.uniform(0.5, 2.0, n), "pr": rng.uniform(0.0, 1e9, n),
        "sxx": rng.uniform(-3e8, 3e8, n), "sxy": rng.uniform(-3e8, 3e8, n),
        "eta": 10 ** rng.uniform(19, 25, n),
        "pfailure": (rng.random(n) < 0.2).astype(np.float32),
    }
    k = max(n // 50, 1)
    m["cohesion"][:k] = 0.0          # no strength: yielding not applicable
    m["fric"][:k] = 0.0
    m["sxx"][k:2 * k] = 0.0          # zero old stress
    m["sxy"][k:2 * k] = 0.0
    return m
