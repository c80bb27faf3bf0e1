"""Pins of the scatter restatement (CPU only): partition of unity, constant
and linear reproduction, the exclusive rule's tie at a cell centre, backup
fallback, harmonic averaging."""
import numpy as np

import oracle
from earthbox_gpu import workloads
from earthbox_gpu.grid import uniform_grid
from scatter_cases import FIELDS, marker_set, previous_fields


def test_partition_of_unity_and_constants():
    g, m = marker_set(21, 11, 4, 3)
    n_in = int((oracle.inside_flags(m["x"], m["y"], g.limits) == 1).sum())
    const = dict(m)
    for k in ("rho", "cohesion", "eta_flow", "eta", "sxy", "sxx", "meltfrac"):
        const[k] = np.full(m["x"].size, 2.5)
    const["fric"] = np.full(m["x"].size, 0.5)
    const["matid"] = np.full(m["x"].size, 2, np.int16)
    out = oracle.scatter_stokes(g, const, previous_fields(g, 3))
    # each marker's four inclusive weights sum to one
    assert abs(out["wtnodes"].sum() - n_in) < 1e-9 * n_in
    assert abs(out["wtnodes_vy"].sum() - n_in) < 1e-9 * n_in
    for k in ("rho1", "etas1", "etan1", "eta_flow", "sxy1", "sxx1", "cohesion_grid",
              "extractable_meltfrac_grid", "rho1_vy"):
        assert np.allclose(out[k], 2.5, rtol=1e-14), k
    assert np.allclose(out["fric_degrees_grid"], 30.0, rtol=1e-13)   # asin(0.5) = 30 degrees
    assert np.allclose(out["mus1"], 6e10, rtol=1e-14) and np.allclose(out["mun1"], 6e10, rtol=1e-14)


def test_linear_field_reproduced_at_interior_nodes():
    """Regular (unjittered) markers are symmetric about interior basic nodes,
    so the inclusive average of a linear field equals its nodal value."""
    g, m = marker_set(13, 9, 4, 5, jitter=False, outside=False)
    lin = dict(m)
    lin["rho"] = 3.0 + 2e-5 * m["x"] - 1e-5 * m["y"]
    out = oracle.scatter_stokes(g, lin, previous_fields(g, 5))
    X, Y = np.meshgrid(g.gridx_b, g.gridy_b)
    exact = 3.0 + 2e-5 * X - 1e-5 * Y
    assert np.allclose(out["rho1"][1:-1, 1:-1], exact[1:-1, 1:-1], rtol=1e-12)


def test_exclusive_rule_tie_at_cell_centre_and_backup():
    """A marker exactly at a cell centre (dx = dy = 0.5) feeds all four nodes
    under the exclusive rule (WeightFuncs.jl:122-150); nodes without markers
    keep their previous value."""
    g = uniform_grid(5, 4, 4.0e4, 3.0e4)
    m = {"x": np.array([0.5 * (g.gridx_b[1] + g.gridx_b[2])]),
         "y": np.array([0.5 * (g.gridy_b[1] + g.gridy_b[2])]),
         "eta": np.array([7.0]), "rho": np.array([1.0]), "mat_mu": np.array([1.0]),
         "matid": np.array([1], np.int16)}
    prev = previous_fields(g, 1)
    out = oracle.scatter_stokes(g, m, prev)
    assert np.count_nonzero(out["wtetas"]) == 4
    assert np.allclose(out["wtetas"][1:3, 1:3], 0.25)
    fed = out["wtetas"] != 0
    assert np.all(out["etas1"][fed] == 7.0)
    assert np.array_equal(out["etas1"][~fed], prev["etas1"][~fed])
    assert out["etan1"][1, 1] == 7.0 and out["wtetan"][1, 1] == 1.0


def test_harmonic_shear_modulus():
    g, m = marker_set(9, 7, 4, 8)
    out = oracle.scatter_stokes(g, m, previous_fields(g, 8))
    # harmonic mean lies between the smallest and largest modulus
    fed = out["wtetas"] != 0
    assert out["mus1"][fed].min() >= 1e10 * (1 - 1e-12) and out["mus1"][fed].max() <= 6e10 * (1 + 1e-12)
