"""The reference's Stokes sinker demo inputs (earthbox_gpu.workloads.stokes_sinker3d
/ stokes_sinker2d) against what the demo's definitions fix by counting and
closed form (StokesSinker.jl:29-160 and its initial_conditions/)."""
import numpy as np

from earthbox_gpu import workloads


def test_sinker3d_block_counts_rhs_and_pressure_guess():
    n = 97
    S = workloads.stokes_sinker3d(n)
    # basic nodes at (i-1)/96 in [0.4, 0.6]: i-1 = 39..57 -> 19 per axis;
    # cell centres (i-1/2)/96: i = 39..58 -> 20; shear centres are the same
    # midpoints on a uniform axis
    assert int((S["rho"] == 3100.0).sum()) == 19 ** 3
    assert int((S["etan"] == 1e26).sum()) == 20 ** 3
    assert int((S["etaxy"] == 1e26).sum()) == 19 * 19 * 20
    assert int((S["etaxz"] == 1e26).sum()) == 20 * 19 * 19
    assert int((S["etayz"] == 1e26).sum()) == 19 * 20 * 19
    g = 9.81
    RY = S["RY"]
    assert np.all(RY[0] == 0) and np.all(RY[-1] == 0) and np.all(RY[:, 0] == 0) and np.all(RY[:, :, 0] == 0)
    inner = RY[1:-1, 1:-1, 1:-1]
    vals = set(np.unique(inner / -g).round(6))
    # four-node means of 3000/3100: 3000, 3025, 3050, 3100 (3075 needs three
    # block nodes around a vy node, impossible for an axis-aligned cube)
    assert vals == {3000.0, 3025.0, 3050.0, 3100.0}
    assert int((inner == -g * 3100.0).sum()) == 19 * 18 * 18   # every block row, 18 x 18 node pairs
    for k in ("RX", "RZ", "RC", "vx", "vy", "vz"):
        assert not S[k].any(), k
    # hydrostatic guess: a column away from the block integrates 3000 kg/m^3
    dy = 1e5 / 96
    col = S["pr"][:, 0, 0]
    np.testing.assert_allclose(col, np.arange(n - 1) * dy * g * 3000.0, rtol=1e-12, atol=1e-6)
    assert S["pr"][-1].max() > col[-1]            # the dense block adds weight below it
    np.testing.assert_allclose(S["stokesscale"], 3100.0 * g)
    np.testing.assert_allclose(S["continscale"], 3100.0 * g / 1e20 * dy)


def test_sinker2d_block_counts_rhs_and_pressure_guess():
    n = 97
    S = workloads.stokes_sinker2d(n)
    assert int((S["rho"] == 3100.0).sum()) == 19 ** 2
    assert int((S["etas"] == 1e26).sum()) == 19 ** 2
    assert int((S["etan"] == 1e26).sum()) == 20 ** 2
    g = 9.81
    inner = S["RY"][1:-1, 1:-1]
    assert set(np.unique(inner / -g).round(6)) == {3000.0, 3050.0, 3100.0}
    assert not S["RX"].any() and not S["RC"].any()
    dy = 1e5 / 96
    np.testing.assert_allclose(S["pr"][:, 0], np.arange(n - 1) * dy * g * 3000.0, rtol=1e-12, atol=1e-6)
