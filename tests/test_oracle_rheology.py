"""Pins of the oracle's marker rheology (CPU only), from closed forms the
flow laws and rules fix:
- failure properties: strain below the weakening interval keeps the initial
  cohesion and the marker's friction, above it the final values (the final
  friction scaled by the marker's initial friction when both material
  frictions are positive), the interval's midpoint gives the averages, melt
  damage divides both (FailureProperties.jl);
- composite viscosity: a constant-viscosity material is not clipped; without
  elasticity (mu dt >> eta) the bisection finds the power-law stress
  sigma = (e_dot r / (A exp(-(Ea + Va P)/RT)))^(1/n) and eta = sigma / (2 e_dot r);
  in the elastic limit (mu dt << eta) the stress stays the marker's and
  eta = sigma / (2 A exp(...) sigma^n); pure diffusion creep gives
  eta = 1e6 / (2 A exp(...)) at any stress; at the Peierls stress the Peierls
  rate is A_p sigma_p^2; Peierls creep is off at 1200 K and above; the
  temperature floor; viscous strain softening of the pre-exponential; the
  clip to [viscosity_min, viscosity_max] (CompositeViscosity.jl,
  CompositeCreep.jl, ViscousStrainSoftening.jl);
- viscoplastic viscosity: the itype_global / ntimestep table
  (ViscoplasticViscosity.jl:21-61);
- grid -> marker viscoplastic viscosity: no yielding corner, equal yielding
  corners, one yielding corner (GridViscoplasticityToMarkers.jl);
- marker strain: |e| dt added, plastic strain rate 0.5 sii (1/eta - 1/eta_flow),
  healing and the clamps at zero (MarkersStrain.jl).
"""
import numpy as np
import pytest

import oracle
from earthbox_gpu.grid import uniform_grid
from rheology_cases import MATS, PARAMS, marker_states

R, TTOP = PARAMS["gas_constant"], PARAMS["temperature_top"]


def _one(mat_row, flow_type, plastic_row=(1e7, 1e7, 0.5, 0.5, 0.0, 1.0), mu=3e10, **marker):
    mats = dict(mat_flow_type=np.array([flow_type], np.int64), mat_flow=np.asfortranarray([mat_row]),
                mat_plastic=np.asfortranarray([plastic_row]), mat_mu=np.array([mu]))
    m = {"matid": np.array([1], np.int16), "flags": np.array([1], np.int8), "strain_plastic": [0.0],
         "fric_ini": [0.5], "sxx": [1e6], "sxy": [0.0], "exx": [1e-15], "exy": [0.0], "TK": [1500.0],
         "pr": [1e9], "GII": [0.0], "sr_ratio": [1.0], "eta": [0.0]}
    m.update({k: [v] for k, v in marker.items()})
    m = {k: np.asarray(v) for k, v in m.items()}
    return m, mats


def _law(A=1e3, n=3.5, Ea=400.0, Va=5.0, Ap=0.0, sp=9100.0, Ad=0.0, Ead=0.0, Vad=0.0, visc=0.0):
    return [visc, A, n, Ea, Va, Ap, 1.0, 2.0, sp, Ad, Ead, Vad]


def _expterm(Ea, Va, T, P_pa):
    return np.exp(-(Ea * 1000.0 + Va * P_pa * 1e-6) / (R * T))


def _run(m, mats, **par):
    return oracle.marker_rheology(m, mats, dict(PARAMS, **par))


# ---------------------------------------------------------------- failure properties
@pytest.mark.parametrize("strain,exp_c,exp_f", [(0.2, 2e7, 0.45), (2.0, 4e6, 0.45 * 0.2),
                                                (1.0, 1.2e7, 0.5 * (0.45 + 0.09))])
def test_strain_weakening_closed_form(strain, exp_c, exp_f):
    # weakening between plastic strain 0.5 and 1.5; marker friction 0.45
    m, mats = _one(_law(), 1, plastic_row=(2e7, 4e6, 0.5, 0.1, 0.5, 1.5), strain_plastic=strain, fric_ini=0.45)
    out = _run(m, mats)
    assert out["cohesion"][0] == pytest.approx(exp_c, rel=1e-15)
    assert out["fric"][0] == pytest.approx(exp_f, rel=1e-15)


def test_final_friction_without_initial_material_friction_and_melt_damage():
    m, mats = _one(_law(), 1, plastic_row=(2e7, 4e6, 0.0, 0.1, 0.0, 1.0), strain_plastic=5.0, fric_ini=0.45)
    out = _run(m, mats)
    assert out["fric"][0] == 0.1 and out["cohesion"][0] == 4e6
    m["melt_damage"] = np.array([2.0])
    out = _run(m, mats)
    assert out["fric"][0] == 0.05 and out["cohesion"][0] == 2e6


# ---------------------------------------------------------------- flow viscosity
def test_constant_viscosity_material_is_not_clipped():
    m, mats = _one(_law(visc=1e30), 0)
    out = _run(m, mats)
    assert out["eta_flow"][0] == 1e30


@pytest.mark.parametrize("s0", [5e6, 3e7])
def test_power_law_without_elasticity(s0):
    A, n, Ea, Va, T, P, e, r = 1e3, 3.5, 400.0, 5.0, 1500.0, 1e9, 2e-15, 1.5
    m, mats = _one(_law(A, n, Ea, Va), 1, mu=1e40, sxx=s0, sxy=0.0, exx=e, exy=0.0, TK=T, pr=P, sr_ratio=r)
    out = _run(m, mats, iuse_viscous_strain_soft=0, viscosity_min=1e10, viscosity_max=1e40)
    s_star = (e * r / (A * _expterm(Ea, Va, T, P))) ** (1.0 / n) * 1e6
    eta_star = s_star / (2.0 * e * r)
    assert abs(s_star - s0) < s0, "the bracket must hold the answer"
    # bisection: 20 halvings of a bracket narrower than the start stress, or 1 Pa
    tol = max(1.0, 2 * s0 / 2 ** 20) / s_star * 3.5
    assert out["eta_flow"][0] == pytest.approx(eta_star, rel=tol)


def test_power_law_elastic_limit_keeps_marker_stress():
    A, n, Ea, Va, T, P, s = 1e3, 3.5, 400.0, 5.0, 1500.0, 1e9, 2e7
    m, mats = _one(_law(A, n, Ea, Va), 1, mu=1e-20, sxx=0.6 * s, sxy=0.8 * s, TK=T, pr=P)
    out = _run(m, mats, iuse_viscous_strain_soft=0, viscosity_min=1e10, viscosity_max=1e40)
    eta = s / (2.0 * A * _expterm(Ea, Va, T, P) * (s * 1e-6) ** n)
    assert out["eta_flow"][0] == pytest.approx(eta, rel=1e-12)


def test_diffusion_creep_is_newtonian():
    Ad, Ead, Vad, T, P = 1.5e-9, 300.0, 4.0, 1600.0, 2e9
    eta = 1e6 / (2.0 * Ad * _expterm(Ead, Vad, T, P))
    for s in (1e5, 3e7):
        m, mats = _one(_law(A=0.0, Ad=Ad, Ead=Ead, Vad=Vad), 1, sxx=s, TK=T, pr=P)
        out = _run(m, mats, iuse_viscous_strain_soft=0, viscosity_min=1e10, viscosity_max=1e40)
        assert out["eta_flow"][0] == pytest.approx(eta, rel=1e-12)


def test_peierls_rate_at_the_peierls_stress():
    """At sigma = sigma_p the Peierls exponential is exp(0): e = A_p sigma_p^2,
    so in the elastic limit eta = 1e6 / (2 A_p sigma_p) (dislocation creep
    negligible); at 1200 K Peierls creep is off."""
    Ap, sp = 5.7e-15, 1000.0
    m, mats = _one(_law(A=1e-30, Ap=Ap, sp=sp), 1, mu=1e-20, sxx=sp * 1e6, TK=1000.0)
    out = _run(m, mats, iuse_viscous_strain_soft=0, viscosity_min=1e5, viscosity_max=1e40)
    assert out["eta_flow"][0] == pytest.approx(1e6 / (2.0 * Ap * sp), rel=1e-12)
    m["TK"] = np.array([1200.0])
    out_off = _run(m, mats, iuse_viscous_strain_soft=0, viscosity_min=1e5, viscosity_max=1e40)
    m2, mats2 = _one(_law(A=1e-30), 1, mu=1e-20, sxx=sp * 1e6, TK=1200.0)
    out_dis = _run(m2, mats2, iuse_viscous_strain_soft=0, viscosity_min=1e5, viscosity_max=1e40)
    assert out_off["eta_flow"][0] == out_dis["eta_flow"][0]


def test_temperature_floor_and_clip():
    m, mats = _one(_law(), 1, TK=100.0)
    a = _run(m, mats, viscosity_max=1e40)["eta_flow"][0]
    m["TK"] = np.array([TTOP])
    b = _run(m, mats, viscosity_max=1e40)["eta_flow"][0]
    assert a == b
    assert _run(m, mats)["eta_flow"][0] == PARAMS["viscosity_max"]
    m2, mats2 = _one(_law(A=1e10, Ea=10.0), 1, TK=1700.0, sxx=3e7)
    assert _run(m2, mats2)["eta_flow"][0] == PARAMS["viscosity_min"]


@pytest.mark.parametrize("strain,factor", [(0.0, 1.0), (0.25, 1.0 + 9 * 0.25), (1.0, 10.0), (3.0, 10.0)])
def test_viscous_strain_softening_of_preexponential(strain, factor):
    m, mats = _one(_law(A=2.0), 1, plastic_row=(1e7, 1e7, 0.5, 0.5, 0.0, 1.0), GII=strain)
    out = _run(m, mats, vsoftfac=10.0, iuse_viscous_strain_soft=1)
    assert out["preexp"][0] == pytest.approx(2.0 * factor, rel=1e-15)


# ---------------------------------------------------------------- viscoplastic viscosity
@pytest.mark.parametrize("itype,nstep,eta0,takes_flow", [(0, 5, 7e20, True), (1, 0, 7e20, True),
                                                         (1, 5, 0.0, True), (1, 5, 7e20, False)])
def test_viscoplastic_viscosity_rule(itype, nstep, eta0, takes_flow):
    m, mats = _one(_law(visc=3e21), 0, eta=eta0)
    out = _run(m, mats, itype_global=itype, ntimestep=nstep)
    assert out["eta"][0] == (3e21 if takes_flow else eta0)


def test_outside_markers_untouched():
    m = marker_states(200, 3)
    out = _run(m, MATS)
    out_ = m["flags"] != 1
    for k in ("cohesion", "fric", "eta_flow", "preexp"):
        assert np.all(out[k][out_] == -1.0)
    assert np.array_equal(out["eta"][out_], m["eta"][out_])
    assert np.all(out["eta_flow"][~out_] >= PARAMS["viscosity_min"])


# ---------------------------------------------------------------- grid -> marker
def _grid_case(yield_corners, etas_val, eta_flow, dx=0.25, dy=0.75):
    g = uniform_grid(5, 4, 4.0, 3.0)
    py = np.zeros((4, 5), order="F")
    et = np.full((4, 5), 1e30, order="F")
    for (r, c), v in yield_corners.items():
        py[r, c] = 1.0
        et[r, c] = etas_val if v is None else v
    m = {"x": np.array([1.0 + dx]), "y": np.array([1.0 + dy]), "flags": np.array([1], np.int8),
         "marker_xn": np.array([2], np.int32), "marker_yn": np.array([2], np.int32),
         "eta_flow": np.array([eta_flow]), "eta": np.array([0.0]), "pfailure": np.array([0.0], np.float32)}
    return oracle.marker_viscoplastic_from_grid(g, m, py, et)


def test_grid_viscosity_to_markers_rules():
    eta, pf = _grid_case({}, 0.0, 5e21)
    assert eta[0] == 5e21 and pf[0] == 0.0
    corners = {(1, 1): None, (2, 1): None, (1, 2): None, (2, 2): None}
    eta, pf = _grid_case(corners, 2e20, 5e21)
    assert eta[0] == pytest.approx(2e20, rel=1e-15) and pf[0] == 1.0
    eta, pf = _grid_case(corners, 9e21, 5e21)
    assert eta[0] == 5e21 and pf[0] == 0.0
    # one yielding corner (lower right, weight dx dy): numerator and
    # denominator share the weight, so eta is that corner's viscosity
    eta, pf = _grid_case({(2, 2): 3e20}, 0.0, 5e21)
    assert eta[0] == pytest.approx(3e20, rel=1e-15) and pf[0] == 1.0


# ---------------------------------------------------------------- marker strain
def test_marker_strain_closed_form():
    m = {"flags": np.array([1, 1, 1, -1], np.int8), "exx": np.array([3e-15, 0.0, 1e-15, 1.0]),
         "exy": np.array([4e-15, 0.0, 0.0, 1.0]), "sxx": np.array([6e6, 6e6, 6e6, 6e6]),
         "sxy": np.array([8e6, 8e6, 8e6, 8e6]), "eta": np.array([1e20, 1e22, 1e20, 1e20]),
         "eta_flow": np.array([1e21, 1e21, 1e21, 1e21]), "pfailure": np.array([1, 1, 0, 1], np.float32),
         "GII": np.array([0.5, 0.5, 0.5, 0.5]), "strain_plastic": np.array([0.1, 0.1, 0.1, 0.1]),
         "strain_rate_plastic": np.full(4, -1.0)}
    dt, heal = 1e13, 1e-15
    gii, ps, psr = oracle.update_marker_strain(m, dt, heal)
    assert gii[0] == pytest.approx(0.5 + dt * 5e-15, rel=1e-15) and gii[1] == 0.5 and gii[3] == 0.5
    rate = 0.5 * 1e7 * (1 / 1e20 - 1 / 1e21)
    assert psr[0] == pytest.approx(rate, rel=1e-15)
    assert ps[0] == pytest.approx(0.1 + dt * rate - heal * dt, rel=1e-14)
    assert psr[1] == 0.0 and ps[1] == pytest.approx(0.1 - heal * dt)   # eta above eta_flow: no plastic rate
    assert psr[2] == -1.0 and ps[2] == 0.1                              # not failing: untouched
    assert ps[3] == 0.1                                                 # outside
    m["strain_plastic"][0] = 0.0
    _, ps, _ = oracle.update_marker_strain(dict(m, eta=np.full(4, 1e21)), dt, heal)
    assert ps[0] == 0.0                                                 # healing clamps at zero
