"""The allocation-free plastic-yield entry point
(ebg_update_nodes_for_plastic_yielding_noalloc) against the allocating one,
through the C ABI: same kernels and reduction order, so every output array,
the yield count and the RMS yield error are bit-identical.  Two calls in a row
on the same device-wide scratch, the second on a smaller grid."""
import numpy as np
import pytest

from earthbox_gpu.device import to_device, to_host
from earthbox_gpu.plastic import update_nodes_for_plastic_yielding
from plastic_cases import PARAMS, plastic_fields

pytestmark = pytest.mark.gpu


def _both(ctx, nx, ny, seed, wall, ext):
    g, f, ye = plastic_fields(nx, ny, seed)
    par = dict(PARAMS, iuse_fluid_pressure_for_yield=seed % 2, no_yielding_in_mobile_wall=wall,
               no_yielding_in_plate_extension=ext)
    ctx.set_grid(g)
    df = {k: to_device(v) for k, v in f.items()}
    res = []
    for noalloc in (False, True):
        dye = to_device(ye)
        out, n, err = update_nodes_for_plastic_yielding(ctx, df, dye, par, bool(wall), bool(ext),
                                                        noalloc=noalloc)
        ctx.synchronize()
        res.append(({k: to_host(v) for k, v in out.items()}, to_host(dye), n, err))
    return res


def test_noalloc_matches_allocating_entry_point(gpu_ctx):
    for nx, ny, seed, wall, ext in ((401, 201, 11, 1, 1), (61, 31, 12, 0, 0)):
        print(f"plastic yielding noalloc vs allocating nx={nx} ny={ny}", flush=True)
        (oa, yea, na, ea), (ob, yeb, nb, eb) = _both(gpu_ctx, nx, ny, seed, wall, ext)
        assert na > 0 and nb == na
        assert eb == ea
        for k in oa:
            assert np.array_equal(ob[k], oa[k]), k
        assert np.array_equal(yeb, yea)
