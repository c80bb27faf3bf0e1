"""GPU parity of the nodal plastic-yield update with the oracle.  Flags and
the yield count are exact.  Per-node values differ only through the device
math library's sine, arcsine and cosine (a few ulp of the yield stress), which
the return map can amplify near the yield surface: the viscosity is held to
1e-12 relative, the plastic strain rate and the yield errors to 1e-11 of
their largest magnitude; the RMS yield error, summed in another order, to
1e-12."""
import numpy as np
import pytest

import oracle
from earthbox_gpu.device import to_device, to_host
from earthbox_gpu.plastic import (interpolate_basic_viscoplastic_viscosity_to_pressure_grid,
                                  plastic_strain_rate_at_pressure_nodes, update_nodes_for_plastic_yielding)
from plastic_cases import PARAMS, plastic_fields

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("nx,ny,seed,fluid,wall,ext", [(9, 7, 1, 0, 0, 0), (61, 31, 2, 1, 1, 0),
                                                       (257, 129, 3, 0, 1, 1), (1001, 501, 4, 1, 0, 1)])
def test_plastic_nodes_match_oracle(gpu_ctx, nx, ny, seed, fluid, wall, ext):
    print(f"plastic nodes nx={nx} ny={ny} fluid={fluid} wall={wall} ext={ext}", flush=True)
    g, f, ye = plastic_fields(nx, ny, seed)
    par = dict(PARAMS, iuse_fluid_pressure_for_yield=fluid, no_yielding_in_mobile_wall=wall,
               no_yielding_in_plate_extension=ext)
    ev, psr, pyn, ye2, n, err = oracle.update_nodes_for_plastic_yielding(g, f, par, ye)
    gpu_ctx.set_grid(g)
    dye = to_device(ye)
    out, gn, gerr = update_nodes_for_plastic_yielding(gpu_ctx, {k: to_device(v) for k, v in f.items()}, dye, par,
                                                      bool(wall), bool(ext))
    gpu_ctx.synchronize()
    assert gn == n and n > 0
    np.testing.assert_array_equal(to_host(out["plastic_yield_new"]), pyn)
    np.testing.assert_allclose(to_host(out["etavp_new"]), ev, rtol=1e-12, atol=0)
    np.testing.assert_allclose(to_host(out["plastic_strain_rate"]), psr, rtol=0, atol=1e-11 * np.abs(psr).max())
    np.testing.assert_allclose(to_host(dye), ye2, rtol=0, atol=1e-11 * np.abs(ye2).max())
    np.testing.assert_allclose(gerr, err, rtol=1e-12)


def test_no_yielding_before_the_first_step(gpu_ctx):
    g, f, ye = plastic_fields(41, 23, 5)
    gpu_ctx.set_grid(g)
    dye = to_device(ye)
    out, n, err = update_nodes_for_plastic_yielding(gpu_ctx, {k: to_device(v) for k, v in f.items()}, dye,
                                                    dict(PARAMS, ntimestep=0))
    assert n == 0 and err == 0.0
    np.testing.assert_array_equal(to_host(out["etavp_new"]), f["eta_flow"])
    np.testing.assert_array_equal(to_host(dye), ye)
    assert not to_host(out["plastic_yield_new"]).any()


@pytest.mark.parametrize("nx,ny,seed", [(12, 9, 6), (301, 151, 7)])
def test_picard_grid_updates_bit_exact(gpu_ctx, nx, ny, seed):
    print(f"harmonic viscosity and plastic strain rate nx={nx} ny={ny}", flush=True)
    g, f, _ = plastic_fields(nx, ny, seed)
    rng = np.random.default_rng(seed)
    plastics = np.asfortranarray((rng.random((ny, nx)) < 0.1).astype(float))
    plasticn = np.asfortranarray((rng.random((ny - 1, nx - 1)) < 0.1).astype(float))
    pyield = np.asfortranarray((rng.random((ny, nx)) < 0.1).astype(float))
    etan1 = np.asfortranarray(rng.uniform(1e20, 1e22, (ny - 1, nx - 1)))
    eii = np.asfortranarray(rng.uniform(0, 1e-14, (ny, nx)))
    exp_eta = oracle.harmonic_normal_viscosity(g, f["etas1"], plastics, plasticn, pyield, etan1)
    exp_eii = oracle.plastic_strain_rate_at_pressure_nodes(g, eii)
    gpu_ctx.set_grid(g)
    d_eta = to_device(etan1)
    interpolate_basic_viscoplastic_viscosity_to_pressure_grid(gpu_ctx, to_device(f["etas1"]), to_device(plastics),
                                                              to_device(plasticn), to_device(pyield), d_eta)
    d_ep = to_device(np.zeros((ny - 1, nx - 1), order="F"))
    plastic_strain_rate_at_pressure_nodes(gpu_ctx, to_device(eii), d_ep)
    gpu_ctx.synchronize()
    assert np.array_equal(to_host(d_eta), exp_eta)
    assert np.array_equal(to_host(d_ep), exp_eii)
