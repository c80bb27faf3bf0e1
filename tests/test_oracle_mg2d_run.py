"""Pins of the restated 2-D stand-alone multigrid driver (oracle/mg.py
Hierarchy.run, oracle/mg2d.c ebo_mg_mean_residuals).

What fixes each piece independently of its own formulas:
  - mean residuals: alternating +-c on the interior and huge values in the
    excluded slots give exactly log10(c/scale);
  - continuation: constant minimum viscosity before the first jump, maximum
    equal to min * (max/min)^(jumps/njumps) after `jumps` jumps, the full
    field after njumps jumps, one jump every nvcycles_viscosity_jump cycles;
  - stopping rule: with a loose criterion the run stops at the first cycle
    past nvcycles_viscosity_jump * njumps (and never before the third);
  - fixed point: a converged run ends where plain V-cycles on the full
    viscosity end;
  - multimulti: with one cycle per jump and 1.9999 jumps the maximum
    viscosity oversteps the full maximum at the second jump of each ramp, so
    global updates fall on cycles 3, 6, 9, ... (derived from the rule, not
    from the code); the run cannot stop before the second update and stops
    there under any criterion below the 1e38 placeholder; the recorded global
    residuals are the mean residuals of level 0's own residual (recomputing
    them after the run gives the last record bit for bit, the ghost and
    pressure passes being idempotent); a converged run ends at the plain
    V-cycle fixed point.
"""
import numpy as np

import oracle.mg as mg
from earthbox_gpu import workloads

RNG = np.random.default_rng(7)


def _sinker_hierarchy(n=33, eta_block=1e23):
    S = workloads.stokes_sinker2d(n, viscosity_block=eta_block)
    H = mg.Hierarchy(n, n, 1e5, 1e5)
    f = H.levels[0]
    for k in ("etas", "etan", "RX", "RY", "RC", "vx", "vy", "pr"):
        getattr(f, k)[...] = S[k]
    return H, S


def test_mean_residuals_closed_form():
    H, _ = _sinker_hierarchy(17)
    f = H.levels[0]
    ny, nx = f.ynum, f.xnum
    c = {"rx": 3.0, "ry": 0.25, "rc": 0.04}
    sl = {"rx": (slice(1, ny), slice(1, nx - 1)), "ry": (slice(1, ny - 1), slice(1, nx)),
          "rc": (slice(None), slice(None))}
    for k, s in sl.items():
        a = getattr(f, k)
        a[...] = 1e30
        sign = np.where(RNG.uniform(size=a[s].shape) < 0.5, -1.0, 1.0)
        a[s] = c[k] * sign
    out = H.mean_residuals(2.0, 0.5)
    np.testing.assert_allclose(out, [np.log10(3.0 / 2.0), np.log10(0.25 / 2.0), np.log10(0.04 / 0.5)],
                               rtol=0, atol=1e-14)


def _visc_range(H, k=0):
    f = H.levels[k]
    return min(f.etas.min(), f.etan.min()), max(f.etas.max(), f.etan.max())


def test_run_continuation_schedule_and_stopping_rule():
    kw = dict(nviscosity_jumps=3.0, nvcycles_viscosity_jump=2, finest_sweeps=2, cap=16)
    H, S = _sinker_hierarchy(17)
    lo, hi = 1e20, 1e23
    n, conv, hist = H.run(1, 1e300, **kw)
    assert (n, conv) == (1, False) and hist.shape == (1, 3)
    for lv in H.levels:
        assert np.all(lv.etas == lo) and np.all(lv.etan == lo)
    step = 10 ** (np.log10(hi / lo) / 3.0)
    for cycles, jumps in ((2, 1), (3, 1), (4, 2)):
        H, _ = _sinker_hierarchy(17)
        H.run(cycles, 1e300, **kw)
        np.testing.assert_allclose(_visc_range(H), [lo, lo * step ** jumps], rtol=1e-13)
    H, _ = _sinker_hierarchy(17)
    n, conv, hist = H.run(50, 1e300, **kw)
    assert (n, conv) == (7, True) and hist.shape == (7, 3)
    np.testing.assert_allclose(H.levels[0].etas, S["etas"], rtol=1e-12)
    np.testing.assert_allclose(H.levels[0].etan, S["etan"], rtol=1e-12)
    # before the third cycle the largest residual counts as 1e38: with a
    # criterion below that the third cycle is the earliest stop, above it the
    # first cycle past the ramp is
    H, _ = _sinker_hierarchy(17, eta_block=1e20)
    n, conv, _ = H.run(50, 1e30, nviscosity_jumps=1.0, nvcycles_viscosity_jump=1)
    assert (n, conv) == (3, True)
    H, _ = _sinker_hierarchy(17, eta_block=1e20)
    n, conv, _ = H.run(50, 1e300, nviscosity_jumps=1.0, nvcycles_viscosity_jump=1)
    assert (n, conv) == (2, True)


def test_converged_run_reaches_the_plain_vcycle_fixed_point():
    kw = dict(finest_sweeps=3, cap=16)
    H, _ = _sinker_hierarchy(33, eta_block=1e22)
    n, conv, hist = H.run(400, 1e-9, nviscosity_jumps=2.0, nvcycles_viscosity_jump=5, **kw)
    assert conv and 10 < n < 400
    P, _ = _sinker_hierarchy(33, eta_block=1e22)
    P.restrict_viscosity()
    sm = mg.smoothing_counts(len(P.levels), 3, cap=16)
    for _ in range(400):
        P.vcycle(sm)
    for k in ("vx", "vy", "pr"):
        a, b = getattr(H.levels[0], k), getattr(P.levels[0], k)
        assert np.abs(a - b).max() <= 1e-6 * np.abs(b).max(), k


def test_multimulti_global_update_schedule():
    kw = dict(nviscosity_jumps=1.9999, nvcycles_viscosity_jump=1, finest_sweeps=2, cap=16)
    H, _ = _sinker_hierarchy(17)
    n, conv, hist, gh = H.run_multimulti(9, 0.0, **kw)
    assert (n, conv) == (9, False) and gh.shape == (3, 3) and hist.shape == (9, 3)
    for cycles, updates in ((2, 0), (3, 1), (5, 1), (6, 2)):
        H, _ = _sinker_hierarchy(17)
        _, _, _, gh = H.run_multimulti(cycles, 0.0, **kw)
        assert gh.shape[0] == updates, (cycles, gh.shape)
    H, _ = _sinker_hierarchy(17)
    n, conv, _, gh = H.run_multimulti(50, 1e30, **kw)
    assert (n, conv) == (6, True) and gh.shape[0] == 2
    # the placeholder before two updates is 1e38: a criterion above it stops at once
    H, _ = _sinker_hierarchy(17)
    n, conv, _, gh = H.run_multimulti(50, 1e39, **kw)
    assert (n, conv, gh.shape[0]) == (1, True, 0)


def test_multimulti_global_residuals_are_level0_residuals():
    H, S = _sinker_hierarchy(17)
    sc = dict(stokesscale=S["stokesscale"], continscale=S["continscale"])
    _, _, _, gh = H.run_multimulti(7, 0.0, nviscosity_jumps=1.9999, nvcycles_viscosity_jump=1, finest_sweeps=2,
                                   cap=16, **sc)
    lev0 = H.level0
    rx = H.levels[0].RX.copy()
    H.level0_residual(lev0, 0.0)
    out = np.zeros(3)
    mg._sigs().ebo_mg_mean_residuals(mg.C.byref(lev0.cs()), sc["stokesscale"], sc["continscale"], mg.p(out))
    assert np.array_equal(out, gh[-1])
    # level 1's right-hand side is level 0's residual
    assert np.array_equal(H.levels[0].RX, rx)


def test_multimulti_converges_to_the_plain_vcycle_fixed_point():
    H, S = _sinker_hierarchy(33, eta_block=1e22)
    n, conv, hist, gh = H.run_multimulti(600, 1e-9, nviscosity_jumps=5.9999, nvcycles_viscosity_jump=3,
                                         stokesscale=S["stokesscale"], continscale=S["continscale"],
                                         finest_sweeps=5, cap=48)
    assert conv and n < 600 and gh.shape[0] >= 2
    P, _ = _sinker_hierarchy(33, eta_block=1e22)
    P.restrict_viscosity()
    sm = mg.smoothing_counts(len(P.levels), 5, cap=48)
    for _ in range(400):
        P.vcycle(sm)
    for k in ("vx", "vy", "pr"):
        a, b = getattr(H.level0, k), getattr(P.levels[0], k)
        assert np.abs(a - b).max() <= 1e-6 * np.abs(b).max(), k
