"""GPU parity of the grid strain rate, forecast stress and stress change with
the oracle (bit-exact)."""
import numpy as np
import pytest

import oracle
from earthbox_gpu.device import to_device, to_host
from earthbox_gpu.viscoelastic import update_grid_strain_rate_and_stress
from strain_cases import strain_fields

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("nx,ny,seed", [(9, 7, 1), (61, 31, 2), (1001, 501, 3)])
def test_strain_rate_and_stress_bit_exact(gpu_ctx, nx, ny, seed):
    print(f"strain rate and stress nx={nx} ny={ny}", flush=True)
    g, vx1, vy1, f = strain_fields(nx, ny, seed)
    dt = 3.15e11
    exp = oracle.grid_strain_rate_and_stress(g, vx1, vy1, f, dt)
    gpu_ctx.set_grid(g)
    d = {k: to_device(v) for k, v in f.items()}
    d["vx1"] = to_device(vx1)
    d["vy1"] = to_device(vy1)
    out = update_grid_strain_rate_and_stress(gpu_ctx, d, dt)
    gpu_ctx.synchronize()
    for k in ("exx", "exy", "eii", "sxy2", "sxx2", "dsxy", "dsxx"):
        assert np.array_equal(to_host(out[k]), exp[k]), k
