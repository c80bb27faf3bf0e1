"""One model time step composed from the oracle's restatements with a sparse
direct solve, in the order of the device driver (earthbox_gpu.timeloop),
i.e. the reference's time loop on the hot path, with the Picard loop over
nodes (itype_global 1) or over markers (itype_global 0)."""
import dataclasses

import numpy as np

import oracle
from earthbox_gpu.picard import check_stop_global_loop_markers, check_stop_global_loop_nodes

GRID_FIELDS = ("rho1", "rho1_vy", "etas1", "etan1", "eta_flow", "mus1", "mun1", "sxy1", "sxx1", "plastics",
               "plasticn", "cohesion_grid", "fric_degrees_grid", "dilatation_grid", "extractable_meltfrac_grid")


def initial_grid(g):
    ny, nx = g.ynum, g.xnum
    from earthbox_gpu.scatter import grid_field_shape
    f = {k: np.zeros(grid_field_shape(k, nx, ny), order="F") for k in GRID_FIELDS}
    f.update(plastic_yield=np.zeros((ny, nx), order="F"), yield_error=np.zeros((ny, nx), order="F"),
             eii_plastic_pressure=np.zeros((ny - 1, nx - 1), order="F"))
    for k in ("tk1", "tk2", "kt1", "rhocp1", "hr1", "ha1"):
        f[k] = np.zeros((ny, nx), order="F")
    return f


def picard(g, template, s, par, dt, timesum, step, nglobal, tol, old):
    """PlasticityLoopNodes.picard_loop with the oracle and a direct solve."""
    ppar = dict(timestep=dt, timesum=timesum, ntimestep=step, viscosity_min=par["viscosity_min"],
                viscosity_max=par["viscosity_max"], iuse_fluid_pressure_for_yield=0)
    log = []
    for it in range(1, nglobal + 1):
        if it > 1:
            s["etan1"] = oracle.harmonic_normal_viscosity(g, s["etas1"], s["plastics"], s["plasticn"],
                                                          s["plastic_yield"], s["etan1"])
        ve = oracle.stokes_viscoelastic_rhs(g, s, dt, 0.0, template.gravity_y, 0)
        prob = dataclasses.replace(template, etas0=ve["etas0"], etan0=ve["etan0"], RX1=ve["RX1"], RY1=ve["RY1"],
                                   RC1=ve["RC1"], rho1_vy=s["rho1_vy"], timestep=dt,
                                   pscale=2.0 * ve["etan0"][0, 0] / (g.xstpavg + g.ystpavg))
        Li, Lj, Lv, R = oracle.stokes_build(prob)
        S = oracle.direct_solve(prob.N, Li, Lj, Lv, R)
        vx1, vy1, pr1 = oracle.stokes_unpack(prob, S)
        ss = oracle.grid_strain_rate_and_stress(g, vx1, vy1, s, dt)
        v = np.concatenate([vx1.ravel(order="F"), vy1.ravel(order="F")])
        dv = oracle.get_L2_norms(v, old if old is not None else np.zeros_like(v))[1]
        old = v
        f = {k: s[k] for k in oracle.PLASTIC_FIELDS if k in s}
        f.update(sxy2=ss["sxy2"], sxx2=ss["sxx2"], pr1=pr1)
        ev, psr, pyn, ye, n, err = oracle.update_nodes_for_plastic_yielding(g, f, ppar, s["yield_error"])
        s["yield_error"] = ye
        if n > 0:
            s["etas1"], s["plastic_yield"] = ev, pyn
            s["eii_plastic_pressure"] = oracle.plastic_strain_rate_at_pressure_nodes(g, psr)
        log.append({"nnode_yield": n, "global_yield_error": err, "dvxy_rel_L2": dv})
        if check_stop_global_loop_nodes(n, it, nglobal, err, dv, tol):
            break
    return vx1, vy1, pr1, log


def marker_plastic_params(par, step, dt):
    """The scalars of the marker yielding step, with the device driver's defaults."""
    out = {k: par.get(k, d) for k, d in (("y_sealevel", 0.0), ("yield_stress_min", 0.0), ("yield_stress_max", 1e32),
                                         ("no_yielding_in_mobile_wall", 0), ("iuse_fluid_pressure_for_yield", 0),
                                         ("matid_sticky_air", -1), ("matid_mobile_wall", -1),
                                         ("matid_plate_extension", -1))}
    out.update(ntimestep=step, timestep=dt, viscosity_min=par["viscosity_min"], viscosity_max=par["viscosity_max"],
               forecast=par.get("itype_plasticity", 0))
    return out


def picard_markers(g, template, s, m, par, nglobal, tol, old=None, iupdate_timestep=0, displ_max=0.5):
    """PlasticityLoopMarkers.picard_loop with the oracle and a direct solve.
    s (grid fields) and m (markers: x, y, inside_flags, marker_xn, marker_yn,
    matid, mat_mu, eta, sxx, sxy, exx, exy, sr_ratio, pr, cohesion, fric,
    pfailure) are updated in place; par: marker_plastic_params.  Returns
    (vx1, vy1, pr1, final dt, log)."""
    dt = par["timestep"]
    log = []
    for it in range(1, nglobal + 1):
        nfailed = None
        if it > 1:
            pm = dict(m, flags=m["inside_flags"])
            m["sxx"], m["sxy"], m["eta"], m["pfailure"] = oracle.update_marker_plastic_yielding(
                pm, dict(par, timestep=dt))
            nfailed = int((m["pfailure"] == 1).sum())
            sc = oracle.scatter_stokes(g, {k: m[k] for k in ("x", "y", "inside_flags", "eta")},
                                       {"etas1": s["etas1"], "etan1": s["etan1"]})
            s["etas1"], s["etan1"] = sc["etas1"], sc["etan1"]
            s["etan1"] = oracle.harmonic_normal_viscosity(g, s["etas1"], s["plastics"], s["plasticn"],
                                                          s["plastic_yield"], s["etan1"])
        ve = oracle.stokes_viscoelastic_rhs(g, s, dt, 0.0, template.gravity_y, 0)
        prob = dataclasses.replace(template, etas0=ve["etas0"], etan0=ve["etan0"], RX1=ve["RX1"], RY1=ve["RY1"],
                                   RC1=ve["RC1"], rho1_vy=s["rho1_vy"], timestep=dt,
                                   pscale=2.0 * ve["etan0"][0, 0] / (g.xstpavg + g.ystpavg))
        Li, Lj, Lv, R = oracle.stokes_build(prob)
        S = oracle.direct_solve(prob.N, Li, Lj, Lv, R)
        vx1, vy1, pr1 = oracle.stokes_unpack(prob, S)
        if iupdate_timestep:
            dt, _ = oracle.calculate_new_timestep(dt, displ_max, g.xstpavg, g.ystpavg, vx1, vy1)
        ss = oracle.grid_strain_rate_and_stress(g, vx1, vy1, s, dt)
        gm = {"x": m["x"], "y": m["y"], "xn": m["marker_xn"], "yn": m["marker_yn"], "inside": m["inside_flags"],
              "eta": m["eta"], "matid": m["matid"], "mat_mu": m["mat_mu"]}
        gf = {"exx": ss["exx"], "exy": ss["exy"], "pr1": pr1, "sxx2": ss["sxx2"], "sxy2": ss["sxy2"],
              "dsxx": ss["dsxx"], "dsxy": ss["dsxy"]}
        prev = {"marker_exx": m["exx"], "marker_exy": m["exy"], "marker_pr": m["pr"],
                "marker_sr_ratio": m["sr_ratio"]}
        ga = oracle.marker_strain_rate_and_pressure(g, gm, gf, dt, prev)
        m["exx"], m["exy"], m["pr"], m["sr_ratio"] = (ga[k] for k in ("marker_exx", "marker_exy", "marker_pr",
                                                                      "marker_sr_ratio"))
        v = np.concatenate([vx1.ravel(order="F"), vy1.ravel(order="F")])
        dv = oracle.get_L2_norms(v, old if old is not None else np.zeros_like(v))[1]
        old = v
        log.append({"iglobal": it, "nmarker_failed": nfailed, "dvxy_rel_L2": dv, "timestep": dt})
        if check_stop_global_loop_markers(it, nglobal, dv, tol):
            break
    return vx1, vy1, pr1, dt, log


def time_step(template, m, mats, par, grid, dt, timesum, step, vold, thermal=None):
    """m, grid: numpy dicts updated in place; returns (new dt, vx1, vy1, log).
    thermal: (mats with mat_hr, thermal boundary conditions, heat scalars)
    adds the heat path (m then holds kt, rhocp, ha; grid holds tk2)."""
    g = template.grid
    flags = oracle.inside_flags(m["x"], m["y"], g.limits)
    if thermal is not None:   # heat-field scatter with fresh cells, then the boundary conditions
        tmats, tbc, hpar = thermal
        hm = {k: m[k] for k in ("x", "y", "TK", "kt", "rhocp", "ha", "matid")}
        hm["flags"] = flags
        prev = {"tk2": grid["tk2"], "kt1": grid["kt1"], "rhocp1": grid["rhocp1"], "hr1": grid["hr1"],
                "ha1": grid["ha1"]}
        grid.update({k: v for k, v in oracle.scatter_heat(g, hm, prev, tmats["mat_hr"], None, tbc).items()
                     if k in ("tk1", "kt1", "rhocp1", "hr1", "ha1")})
    r = oracle.marker_rheology(dict(m, flags=flags), mats, dict(par, timestep=dt, ntimestep=step))
    m.update({k: r[k] for k in ("cohesion", "fric", "eta_flow", "preexp", "eta")})
    props = {k: m[k] for k in ("x", "y", "rho", "cohesion", "fric", "eta_flow", "eta", "sxy", "sxx", "pfailure",
                               "matid")}
    props.update(inside_flags=flags, mat_mu=mats["mat_mu"])
    sc = oracle.scatter_stokes(g, props, {k: grid[k] for k in GRID_FIELDS})
    grid.update({k: sc[k] for k in GRID_FIELDS})
    inside = flags == 1
    m["marker_xn"] = np.where(inside, sc["marker_xn"], m.get("marker_xn", 0)).astype(np.int32)
    m["marker_yn"] = np.where(inside, sc["marker_yn"], m.get("marker_yn", 0)).astype(np.int32)
    grid["etan1"] = oracle.harmonic_normal_viscosity(g, grid["etas1"], grid["plastics"], grid["plasticn"],
                                                     grid["plastic_yield"], grid["etan1"])
    if int(par.get("itype_global", 1)) == 1:
        vx1, vy1, pr1, log = picard(g, template, grid, par, dt, timesum, step, int(par["nglobal"]),
                                    float(par["tolerance_picard"]), vold)
        m["eta"], m["pfailure"] = oracle.marker_viscoplastic_from_grid(
            g, dict(m, flags=flags), grid["plastic_yield"], grid["etas1"])
    else:
        mm = dict(m, inside_flags=flags, mat_mu=mats["mat_mu"])
        vx1, vy1, pr1, _, log = picard_markers(g, template, grid, mm, marker_plastic_params(par, step, dt),
                                               int(par["nglobal"]), float(par["tolerance_picard"]), vold)
        m.update({k: mm[k] for k in ("eta", "sxx", "sxy", "pfailure", "exx", "exy", "sr_ratio", "pr")})
    dt_new, _ = oracle.calculate_new_timestep(dt, par["marker_cell_displ_max"], g.xstpavg, g.ystpavg, vx1, vy1)
    ss = oracle.grid_strain_rate_and_stress(g, vx1, vy1, grid, dt_new)
    gm = {"x": m["x"], "y": m["y"], "xn": m["marker_xn"], "yn": m["marker_yn"], "inside": flags, "eta": m["eta"],
          "matid": m["matid"], "mat_mu": mats["mat_mu"], "marker_sxx": m["sxx"], "marker_sxy": m["sxy"]}
    fg = {"exx": ss["exx"], "exy": ss["exy"], "pr1": pr1, "sxx2": ss["sxx2"], "sxy2": ss["sxy2"],
          "dsxx": ss["dsxx"], "dsxy": ss["dsxy"]}
    prev = {"marker_exx": m["exx"], "marker_exy": m["exy"], "marker_sr_ratio": m["sr_ratio"], "marker_pr": m["pr"]}
    o = oracle.marker_strain_rate_and_pressure(g, gm, fg, dt_new, prev)
    m.update(exx=o["marker_exx"], exy=o["marker_exy"], sr_ratio=o["marker_sr_ratio"], pr=o["marker_pr"],
             sxx=o["marker_sxx"], sxy=o["marker_sxy"])
    if thermal is not None:
        hm = {"x": m["x"], "y": m["y"], "flags": flags, "marker_xn": m["marker_xn"], "marker_yn": m["marker_yn"],
              "TK": m["TK"]}
        hf = {k: grid[k] for k in ("tk1", "kt1", "rhocp1", "hr1", "ha1", "rho1", "etas1", "etan1")}
        hf.update(vx1=vx1, vy1=vy1, sxy2=ss["sxy2"], sxx2=ss["sxx2"])
        hr = oracle.update_marker_temp_for_grid_and_subgrid_changes(
            g, tbc, hm, hf, dt_new, hpar["max_temp_change"], hpar.get("subgrid_diff_coef_temp", 0.0), 0.0,
            template.gravity_y, hpar.get("iuse_adiabatic_heating", 0), hpar.get("iuse_shear_heating", 0))
        m["TK"] = hr["TK"]
        grid["tk2"] = hr["tk2"]
    esp = oracle.calculate_spin(g, vx1, vy1)
    x, y, _, _, _, sxx, sxy = oracle.advect_markers_and_rotate_stress_tensor(
        4, dt_new, g, m["x"], m["y"], m["marker_xn"], m["marker_yn"], flags, vx1, vy1, esp, m["sxx"], m["sxy"])
    m.update(x=x, y=y, sxx=sxx, sxy=sxy)
    m["GII"], m["strain_plastic"], m["strain_rate_plastic"] = oracle.update_marker_strain(
        dict(m, flags=flags), dt_new, par.get("plastic_healing_rate", 0.0))
    return dt_new, vx1, vy1, log
