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


@pytest.mark.parametrize("nx,ny,nm,seed", [(9, 7, 3, 1), (41, 23, 5, 2), (101, 51, 6, 3)])
def test_scatter_bit_exact(gpu_ctx, nx, ny, nm, seed):
    print(f"scatter case nx={nx} ny={ny} nm={nm}", flush=True)
    g, m = marker_set(nx, ny, nm, seed)
    prev = previous_fields(g, seed)
    exp = oracle.scatter_stokes(g, m, prev)
    gpu_ctx.set_grid(g)
    dm = {k: to_device(v) for k, v in m.items()}
    df = {k: to_device(prev[k]) for k in FIELDS}
    n = m["x"].size
    for k in ("wtnodes", "wtetas", "wtetan", "wtnodes_vy"):
        df[k] = to_device(np.zeros_like(exp[k], order="F"))
    df["marker_xn"] = torch.zeros(n, dtype=torch.int32, device="cuda")
    df["marker_yn"] = torch.zeros(n, dtype=torch.int32, device="cuda")
    interpolate_marker_stokes_parameters_to_grid(gpu_ctx, dm, df)
    gpu_ctx.synchronize()
    for k in FIELDS + ("wtnodes", "wtetas", "wtetan", "wtnodes_vy"):
        got = to_host(df[k])
        if k == "fric_degrees_grid":
            np.testing.assert_allclose(got, exp[k], rtol=1e-15, atol=0)
        else:
            assert np.array_equal(got, exp[k]), k
    inside = oracle.inside_flags(m["x"], m["y"], g.limits) == 1
    assert np.array_equal(to_host(df["marker_xn"])[inside], exp["marker_xn"][inside])
    assert np.array_equal(to_host(df["marker_yn"])[inside], exp["marker_yn"][inside])
