"""GPU parity of the grid -> marker gathers after the solve (marker strain
rate, strain-rate ratio, marker pressure): bit-exact against the oracle's
serial loops."""
import numpy as np
import pytest
import torch

import oracle
from earthbox_gpu.device import to_device, to_host
from earthbox_gpu.marker_gather import (OUTPUTS, STRESS, apply_marker_stress_correction_for_remaining_grid_stress,
                                        calculate_marker_pressure, calculate_marker_strain_rate,
                                        update_markers_after_stokes)
from gather_cases import BASIC, PRESSURE, gather_case

pytestmark = pytest.mark.gpu


def _device_case(g, m, f, with_flags):
    dm = {"x": to_device(m["x"]), "y": to_device(m["y"]), "marker_xn": to_device(m["xn"]),
          "marker_yn": to_device(m["yn"]), "eta": to_device(m["eta"]), "matid": to_device(m["matid"]),
          "mat_mu": to_device(m["mat_mu"])}
    if with_flags:
        dm["inside_flags"] = to_device(m["inside"])
    df = {k: to_device(f[k]) for k in PRESSURE + BASIC}
    return dm, df


def _prev(n, seed):
    rng = np.random.default_rng(seed + 5)
    return {k: rng.standard_normal(n) for k in OUTPUTS}


@pytest.mark.parametrize("nx,ny,n,seed,flags", [(9, 7, 300, 1, True), (41, 23, 20000, 2, False),
                                                (101, 51, 200000, 3, True), (64, 65, 50000, 4, False)])
def test_marker_gathers_bit_exact(gpu_ctx, nx, ny, n, seed, flags):
    print(f"marker gather case nx={nx} ny={ny} n={n} flags={flags}", flush=True)
    g, m, f = gather_case(nx, ny, n, seed)
    dt = 3.15576e12
    prev = _prev(n, seed)
    exp = oracle.marker_strain_rate_and_pressure(g, m, f, dt, prev=prev)
    gpu_ctx.set_grid(g)
    dm, df = _device_case(g, m, f, flags)
    out = {k: to_device(prev[k]) for k in OUTPUTS}
    update_markers_after_stokes(gpu_ctx, dm, df, dt, out)
    gpu_ctx.synchronize()
    for k in OUTPUTS:
        assert np.array_equal(to_host(out[k]), exp[k]), k
    # again with the stress correction fused into the same pass
    rng = np.random.default_rng(seed)
    s0 = {"marker_sxx": 1e6 * rng.standard_normal(n), "marker_sxy": 1e6 * rng.standard_normal(n)}
    exp = oracle.marker_strain_rate_and_pressure(g, dict(m, **s0), f, dt, prev=prev)
    out = {k: to_device(v) for k, v in list(prev.items()) + list(s0.items())}
    update_markers_after_stokes(gpu_ctx, dm, df, dt, out, OUTPUTS + STRESS)
    gpu_ctx.synchronize()
    for k in OUTPUTS + STRESS:
        assert np.array_equal(to_host(out[k]), exp[k]), k


def test_marker_gathers_subsets_bit_exact(gpu_ctx):
    """Pressure only and strain rate only leave the other outputs alone; the
    ratio alone equals the fused pass's ratio."""
    print("marker gather subsets", flush=True)
    g, m, f = gather_case(33, 21, 30000, 6)
    dt = 1.0e12
    exp = oracle.marker_strain_rate_and_pressure(g, m, f, dt)
    gpu_ctx.set_grid(g)
    dm, df = _device_case(g, m, f, False)
    pr = calculate_marker_pressure(gpu_ctx, dm, {"pr1": df["pr1"]})
    sr = calculate_marker_strain_rate(gpu_ctx, dm, {"exx": df["exx"], "exy": df["exy"]})
    ra = update_markers_after_stokes(gpu_ctx, dm, df, dt, outputs=("marker_sr_ratio",))
    rng = np.random.default_rng(9)
    s0 = {"marker_sxx": rng.standard_normal(m["x"].size), "marker_sxy": rng.standard_normal(m["x"].size)}
    st = {k: to_device(v) for k, v in s0.items()}
    apply_marker_stress_correction_for_remaining_grid_stress(gpu_ctx, dm, {"dsxx": df["dsxx"], "dsxy": df["dsxy"]}, st)
    exp_s = oracle.marker_strain_rate_and_pressure(g, dict(m, **s0), f, dt)
    gpu_ctx.synchronize()
    assert np.array_equal(to_host(st["marker_sxx"]), exp_s["marker_sxx"])
    assert np.array_equal(to_host(st["marker_sxy"]), exp_s["marker_sxy"])
    assert set(pr) == {"marker_pr"} and set(sr) == {"marker_exx", "marker_exy"}
    assert np.array_equal(to_host(pr["marker_pr"]), exp["marker_pr"])
    assert np.array_equal(to_host(sr["marker_exx"]), exp["marker_exx"])
    assert np.array_equal(to_host(sr["marker_exy"]), exp["marker_exy"])
    assert np.array_equal(to_host(ra["marker_sr_ratio"]), exp["marker_sr_ratio"])


def test_marker_gathers_argument_errors(gpu_ctx):
    from earthbox_gpu._lib import EBGError
    g, m, f = gather_case(9, 7, 100, 7)
    gpu_ctx.set_grid(g)
    dm, df = _device_case(g, m, f, True)
    with pytest.raises(EBGError):
        update_markers_after_stokes(gpu_ctx, dm, {"pr1": df["pr1"]}, 1.0, outputs=("marker_exx",))
    with pytest.raises(EBGError):
        update_markers_after_stokes(gpu_ctx, dict(dm, eta=None), df, 1.0, outputs=("marker_sr_ratio",))
    empty = {k: v[:0] for k, v in dm.items() if k != "mat_mu"}
    empty["mat_mu"] = dm["mat_mu"]
    r = update_markers_after_stokes(gpu_ctx, empty, df, 1.0)
    assert all(v.numel() == 0 for v in r.values())
    torch.cuda.synchronize()
