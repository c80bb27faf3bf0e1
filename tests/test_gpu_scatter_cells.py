"""GPU parity of the cell-ordered scatter (EBG_SCATTER_CELL_ORDER) against the
oracle's serial loops.

Bar: the cell order sums each cell's markers in marker order and then the
neighbouring cells' partial sums in a fixed order, so a node sum differs from
the serial order only by the grouping of its additions.  For every averaged
output and weight sum, |got - serial| <= 1e-13 * (the same average of the
absolute property values), node by node (where a node has no weight the
average is its previous value, kept within the same bound).  Pressure-node outputs take the
markers of one cell in marker order and must be bit-identical, as must the
marker cell indices.  Two runs give bit-identical results (no floating-point
atomics), and the serial order is restored by ebg_scatter_set_order.
"""
import numpy as np
import pytest
import torch

import oracle
from earthbox_gpu.context import GPUContext
from earthbox_gpu.device import to_device, to_host
from earthbox_gpu.scatter import interpolate_marker_stokes_parameters_to_grid
from scatter_cases import FIELDS, marker_set, previous_fields
from test_gpu_scatter import _shuffled_crowded

pytestmark = pytest.mark.gpu

PRESSURE = ("etan1", "mun1", "sxx1", "plasticn", "dilatation_grid", "extractable_meltfrac_grid", "wtetan")
WEIGHTS = ("wtnodes", "wtetas", "wtetan", "wtnodes_vy")
SIGNED = ("sxy", "sxx")


@pytest.fixture
def cell_ctx(gpu_ctx):
    gpu_ctx.set_scatter_order(GPUContext.SCATTER_CELL_ORDER)
    yield gpu_ctx
    gpu_ctx.set_scatter_order(GPUContext.SCATTER_SERIAL_ORDER)


def _run(ctx, g, m, prev, outputs):
    ctx.set_grid(g)
    dm = {k: to_device(v) for k, v in m.items()}
    df = {k: to_device(prev[k]) for k in outputs}
    n = m["x"].size
    for k in WEIGHTS:
        df[k] = to_device(np.zeros_like(prev["rho1"] if k in ("wtnodes", "wtetas") else
                                        prev["etan1"] if k == "wtetan" else prev["rho1_vy"], order="F"))
    df["marker_xn"] = torch.zeros(n, dtype=torch.int32, device="cuda")
    df["marker_yn"] = torch.zeros(n, dtype=torch.int32, device="cuda")
    interpolate_marker_stokes_parameters_to_grid(ctx, dm, df)
    ctx.synchronize()
    return {k: to_host(v) for k, v in df.items()}


def _check(ctx, g, m, prev, outputs=FIELDS):
    exp = oracle.scatter_stokes(g, m, prev)
    mabs = {k: (np.abs(v) if k in SIGNED else v) for k, v in m.items()}
    bound = oracle.scatter_stokes(g, mabs, prev)   # the same averages of |value|
    got = _run(ctx, g, m, prev, outputs)
    again = _run(ctx, g, m, prev, outputs)
    for k in got:
        assert np.array_equal(got[k], again[k]), f"{k}: not deterministic"
    inside = oracle.inside_flags(m["x"], m["y"], g.limits) == 1
    assert np.array_equal(got["marker_xn"][inside], exp["marker_xn"][inside])
    assert np.array_equal(got["marker_yn"][inside], exp["marker_yn"][inside])
    for k in tuple(outputs) + WEIGHTS:
        if k in PRESSURE:
            assert np.array_equal(got[k], exp[k]), k
            continue
        b = np.abs(bound[k])
        err = np.abs(got[k] - exp[k])
        assert np.all(err <= 1e-13 * b), (k, float((err / np.maximum(b, 1e-300)).max()))


@pytest.mark.parametrize("nx,ny,nm,seed", [(9, 7, 3, 1), (41, 23, 5, 2), (101, 51, 6, 3)])
def test_scatter_cell_order(cell_ctx, nx, ny, nm, seed):
    print(f"cell-order scatter nx={nx} ny={ny} nm={nm}", flush=True)
    g, m = marker_set(nx, ny, nm, seed)
    _check(cell_ctx, g, m, previous_fields(g, seed))


@pytest.mark.parametrize("nx,ny,nm,seed", [(12, 9, 4, 4), (61, 31, 5, 5)])
def test_scatter_cell_order_shuffled_crowded_ties(cell_ctx, nx, ny, nm, seed):
    """Random marker order, three cells holding a quarter of the markers
    (lists of hundreds to thousands: the merged register runs), markers on
    nodes, centres and vy nodes."""
    print(f"cell-order scatter shuffled/crowded nx={nx} ny={ny} nm={nm}", flush=True)
    g, m = _shuffled_crowded(nx, ny, nm, seed)
    _check(cell_ctx, g, m, previous_fields(g, seed))


def test_scatter_cell_order_partial_properties(cell_ctx):
    """The density / viscosity / material-table pass and the viscosity-only
    pass (both compiled for their field sets), and a set of signed fields."""
    print("cell-order scatter partial properties", flush=True)
    g, m = marker_set(41, 23, 5, 6)
    prev = previous_fields(g, 6)
    part = {k: m[k] for k in ("x", "y", "rho", "eta", "matid", "mat_mu")}
    _check(cell_ctx, g, part, prev)
    _check(cell_ctx, g, {k: m[k] for k in ("x", "y", "eta")}, prev, outputs=("etas1", "etan1"))
    _check(cell_ctx, g, {k: m[k] for k in ("x", "y", "fric", "pfailure", "sxx", "sxy")}, prev)


def test_scatter_cell_order_list_lengths(cell_ctx):
    """Cells holding 17-32, 33-64 and more than 64 markers in one wavefront
    (every sorting path), in shuffled order."""
    print("cell-order scatter list lengths", flush=True)
    g, m = marker_set(33, 21, 2, 8)
    rng = np.random.default_rng(8)
    n = m["x"].size
    extra = []
    for count, (cx, cy) in ((20, (3, 3)), (48, (10, 5)), (150, (20, 12))):
        xs = g.gridx_b[cx - 1] + rng.random(count) * g.xstp_b[cx - 1]
        ys = g.gridy_b[cy - 1] + rng.random(count) * g.ystp_b[cy - 1]
        extra.append((xs, ys))
    add = sum(len(e[0]) for e in extra)
    for k in m:
        if k == "mat_mu":
            continue
        if k == "x":
            m[k] = np.concatenate([m[k]] + [e[0] for e in extra])
        elif k == "y":
            m[k] = np.concatenate([m[k]] + [e[1] for e in extra])
        else:
            m[k] = np.concatenate([m[k], m[k][rng.integers(0, n, add)]])
    perm = rng.permutation(m["x"].size)
    for k in m:
        if k != "mat_mu":
            m[k] = np.ascontiguousarray(m[k][perm])
    _check(cell_ctx, g, m, previous_fields(g, 8))


def test_scatter_cell_order_heat(cell_ctx):
    """Heat-transport scatter in the cell order against the serial oracle on a
    graded grid with serpentinization heat and an empty patch (all fields
    positive: relative bound 1e-13, backups exact), deterministic; the
    thermal boundary conditions then act on the boundary rows as usual."""
    print("cell-order heat scatter", flush=True)
    from earthbox_gpu import heat
    from earthbox_gpu.grid import graded_grid
    from heat_cases import MAT_HR, heat_markers, random_bc
    nx, ny, seed = 41, 23, 2
    g = graded_grid(nx, ny, 2e5, 1e5, seed=seed)
    m = heat_markers(g, 4, seed)
    empty = np.isin(m["marker_xn"], [3, 4]) & np.isin(m["marker_yn"], [2, 3])
    m["flags"][empty] = -1
    rng = np.random.default_rng(seed)
    prev = {k: np.asfortranarray(rng.uniform(1, 2, (ny, nx))) for k in ("tk2", "kt1", "rhocp1", "hr1", "ha1")}
    bc = random_bc(g, seed)
    exp = oracle.scatter_heat(g, m, prev, MAT_HR, m["serp"], bc)
    cell_ctx.set_grid(g)
    dm = {"x": to_device(m["x"]), "y": to_device(m["y"]), "inside_flags": to_device(m["flags"]),
          "TK": to_device(m["TK"]), "kt": to_device(m["kt"]), "rhocp": to_device(m["rhocp"]),
          "ha": to_device(m["ha"]), "matid": to_device(m["matid"]),
          "serpentinization_heat_production": to_device(m["serp"])}
    outs = []
    for _ in range(2):
        fields = {k: to_device(np.asfortranarray(v)) for k, v in
                  {"tk1": prev["tk2"], "kt1": prev["kt1"], "rhocp1": prev["rhocp1"], "hr1": prev["hr1"],
                   "ha1": prev["ha1"], "wtnodes": np.zeros((ny, nx), order="F")}.items()}
        heat.interpolate_marker_heat_parameters_to_grid(cell_ctx, dm, {"mat_hr": to_device(MAT_HR)}, fields,
                                                        {k: to_device(np.asfortranarray(v) if np.ndim(v) == 2
                                                                      else v) for k, v in bc.items()})
        cell_ctx.synchronize()
        outs.append({k: to_host(v) for k, v in fields.items()})
    assert (exp["wtnodes"] == 0).any()
    for k in outs[0]:
        assert np.array_equal(outs[0][k], outs[1][k]), k
        np.testing.assert_allclose(outs[0][k], exp[k], rtol=1e-13, atol=0, err_msg=k)
