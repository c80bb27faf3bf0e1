"""GPU parity of the viscoelastic terms and right-hand sides (K7) with the
oracle: bit-exact except RC1, whose sine goes through the device math
library (within 1e-15 relative)."""
import numpy as np
import pytest

import oracle
from earthbox_gpu._lib import EBGError
from earthbox_gpu.device import to_device, to_host
from earthbox_gpu.viscoelastic import calc_viscoelastic_terms_and_rhs
from ve_cases import ve_fields

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("nx,ny,seed,stab,dil", [(9, 7, 1, 0, True), (41, 23, 2, 1, True),
                                                 (101, 51, 3, 0, False), (257, 129, 4, 1, True)])
def test_viscoelastic_rhs_matches_oracle(gpu_ctx, nx, ny, seed, stab, dil):
    print(f"viscoelastic rhs nx={nx} ny={ny} stab={stab} dilatation={dil}", flush=True)
    g, f = ve_fields(nx, ny, seed, dilatation=dil)
    dt, gx, gy = 3.15e11, 0.7, 9.81
    exp = oracle.stokes_viscoelastic_rhs(g, f, dt, gx, gy, stab)
    gpu_ctx.set_grid(g)
    out = calc_viscoelastic_terms_and_rhs(gpu_ctx, {k: to_device(v) for k, v in f.items()}, dt, gx, gy, stab)
    gpu_ctx.synchronize()
    for k in ("etas0", "sxy0", "etan0", "sxx0", "RX1", "RY1"):
        assert np.array_equal(to_host(out[k]), exp[k]), k
    np.testing.assert_allclose(to_host(out["RC1"]), exp["RC1"], rtol=1e-15, atol=0)


def test_viscoelastic_rhs_argument_errors(gpu_ctx):
    g, f = ve_fields(9, 7, 5)
    gpu_ctx.set_grid(g)
    d = {k: to_device(v) for k, v in f.items()}
    d.pop("rho1_vy")
    with pytest.raises(EBGError, match="rho1_vy"):
        calc_viscoelastic_terms_and_rhs(gpu_ctx, d, 1.0, 0.0, 9.81, 1)
    d2 = {k: to_device(v) for k, v in f.items()}
    d2.pop("eii_plastic_pressure")
    with pytest.raises(EBGError, match="together"):
        calc_viscoelastic_terms_and_rhs(gpu_ctx, d2, 1.0, 0.0, 9.81, 0)
