"""GPU parity of the marker -> grid Stokes scatter (bit-exact; the friction
angle, which goes through the device asin, within 1e-15 relative)."""
import numpy as np
import pytest
import torch

import oracle
from earthbox_gpu.context import GPUContext
from earthbox_gpu.device import to_device, to_host
from earthbox_gpu.scatter import interpolate_marker_stokes_parameters_to_grid
from scatter_cases import FIELDS, marker_set, previous_fields

pytestmark = pytest.mark.gpu


def _check(gpu_ctx, g, m, prev, outputs=FIELDS):
    """Run the GPU scatter with the marker arrays in m and the requested
    outputs, and compare with the oracle's serial loops."""
    exp = oracle.scatter_stokes(g, m, prev)
    gpu_ctx.set_grid(g)
    dm = {k: to_device(v) for k, v in m.items()}
    df = {k: to_device(prev[k]) for k in outputs}
    n = m["x"].size
    for k in ("wtnodes", "wtetas", "wtetan", "wtnodes_vy"):
        df[k] = to_device(np.zeros_like(exp[k], order="F"))
    df["marker_xn"] = torch.zeros(n, dtype=torch.int32, device="cuda")
    df["marker_yn"] = torch.zeros(n, dtype=torch.int32, device="cuda")
    interpolate_marker_stokes_parameters_to_grid(gpu_ctx, dm, df)
    gpu_ctx.synchronize()
    for k in tuple(outputs) + ("wtnodes", "wtetas", "wtetan", "wtnodes_vy"):
        got = to_host(df[k])
        if k == "fric_degrees_grid" and "fric" in m:
            np.testing.assert_allclose(got, exp[k], rtol=1e-15, atol=0)
        else:
            assert np.array_equal(got, exp[k]), k
    inside = oracle.inside_flags(m["x"], m["y"], g.limits) == 1
    assert np.array_equal(to_host(df["marker_xn"])[inside], exp["marker_xn"][inside])
    assert np.array_equal(to_host(df["marker_yn"])[inside], exp["marker_yn"][inside])


def _shuffled_crowded(nx, ny, nm, seed):
    """Markers in random order, a quarter of them crowded into three cells,
    and some exactly on basic nodes, cell centres and vy-grid nodes (the
    weight-rule ties at dx, dy = 0, 0.5, 1)."""
    g, m = marker_set(nx, ny, nm, seed)
    rng = np.random.default_rng(seed + 11)
    n = m["x"].size
    crowd = rng.choice(n, n // 4, replace=False)
    for k, (cx, cy) in enumerate(((1, 1), (nx // 2, ny // 2), (nx - 2, ny - 2))):
        sel = crowd[k::3]
        m["x"][sel] = g.gridx_b[cx - 1] + rng.random(sel.size) * g.xstp_b[cx - 1]
        m["y"][sel] = g.gridy_b[cy - 1] + rng.random(sel.size) * g.ystp_b[cy - 1]
    ties = rng.choice(n, 60, replace=False)
    jx = rng.integers(0, nx, 60)
    jy = rng.integers(0, ny, 60)
    m["x"][ties[:20]] = g.gridx_b[jx[:20]]
    m["y"][ties[:20]] = g.gridy_b[jy[:20]]
    m["x"][ties[20:40]] = g.gridx_vy[1:-1][np.minimum(jx[20:40], nx - 2)]
    m["y"][ties[20:40]] = g.gridy_b[jy[20:40]]
    cxx = np.minimum(jx[40:], nx - 2)
    m["x"][ties[40:]] = g.gridx_b[cxx] + 0.5 * g.xstp_b[cxx]
    cyy = np.minimum(jy[40:], ny - 2)
    m["y"][ties[40:]] = g.gridy_b[cyy] + 0.5 * g.ystp_b[cyy]
    perm = rng.permutation(n)
    for k in m:
        if k != "mat_mu":
            m[k] = np.ascontiguousarray(m[k][perm])
    return g, m


@pytest.mark.parametrize("nx,ny,nm,seed", [(9, 7, 3, 1), (41, 23, 5, 2), (101, 51, 6, 3)])
def test_scatter_bit_exact(gpu_ctx, nx, ny, nm, seed):
    print(f"scatter case nx={nx} ny={ny} nm={nm}", flush=True)
    g, m = marker_set(nx, ny, nm, seed)
    prev = previous_fields(g, seed)
    _check(gpu_ctx, g, m, prev)


@pytest.mark.parametrize("nx,ny,nm,seed", [(12, 9, 4, 4), (61, 31, 5, 5)])
def test_scatter_shuffled_crowded_ties_bit_exact(gpu_ctx, nx, ny, nm, seed):
    print(f"scatter shuffled/crowded case nx={nx} ny={ny} nm={nm}", flush=True)
    g, m = _shuffled_crowded(nx, ny, nm, seed)
    _check(gpu_ctx, g, m, previous_fields(g, seed))


def test_scatter_partial_properties_bit_exact(gpu_ctx):
    """Only density, viscosity and the material table given (the other sums
    are zero), every output requested; then only three outputs requested."""
    print("scatter partial properties", flush=True)
    g, m = marker_set(41, 23, 5, 6)
    part = {k: m[k] for k in ("x", "y", "rho", "eta", "matid", "mat_mu")}
    prev = previous_fields(g, 6)
    _check(gpu_ctx, g, part, prev)
    _check(gpu_ctx, g, part, prev, outputs=("rho1_vy", "etan1", "mus1"))
    nomat = {k: m[k] for k in ("x", "y", "fric", "pfailure", "sxx")}
    _check(gpu_ctx, g, nomat, prev)
