"""Benchmark: simulated kyr/s of the N=350 laterally variable ELVA configuration on one MI355X.

Workload: lv_elva n=350 (dx = 18 km over a 6300 km square, log10 viscosity in [19, 23] with a
500 km correlation length, lithosphere 50-250 km, 3-6 seeded ice domes under a sawtooth glacial
cycle, elastic + viscous + sea-surface responses on), adaptive BS3 with reltol 1e-5, abstol 1e-6,
sparse elastic / sea-surface update every 10 yr. Inputs are generated, then resident on the device.

A step is one 10-yr window of simulated time (all adaptive BS3 steps inside it plus its one sparse
update). value = simulated kyr / wall-clock s of the plain solver over --steps windows after
--warmup windows. The same line carries the metric's other half, "surrogate": the
surrogate-with-fallback mode over 0 -> --window-kyr (default 10 kyr; 500-yr macro steps, a
plain-solver check of every 10th, fallback above 1 m; predictor RKC2 of the same right-hand side)
with its kyr/s, the plain solver's kyr/s over the same window, their ratio and the largest
|u_surrogate - u_plain| at the 5-kyr marks.

roofline: the dominant kernel is the persistent integration kernel. Its algorithmic traffic is
171 B per cell per accepted BS3 step (FP64 state, FP32 derivatives and fields, counted once per
unit); achieved = that traffic / the kernel time measured with HIP events on the engine's stream.
cpu_baseline: the FP64 CPU restatement of the reference path (oracle/, OpenMP) on this host, on
the same workload, over a bounded window.

Other workloads (--workload): implicit is the disc benchmark of the reference's docs on the
N=350 grid with the implicit step (RigidLithosphere, Euler with dt_min = 100 yr, 1-kyr windows;
the roofline counts IMPLICIT_BYTES_PER_CELL_STEP per cell per step). elra is the relaxed-mantle model (ELRA) on the same N=350 grid:
laterally constant 88 km lithosphere with the elastic response, the 2D relaxation time of the
ELRA example (weak mantle, a Gaussian low-viscosity anomaly), a 1000 m ice cylinder of radius
1000 km switched on over the first year, laterally variable sea surface; a step is a 100-yr window
and the roofline counts ELRA_BYTES_PER_CELL_RHS per cell per RHS call. lv_elva_hr is the N=1024
high-resolution instance (W = 3000 km,
dx = 5.86 km, field seeds 5001/5002); ensemble is the uneven ensemble of 256 members with sides
drawn from {64, 96, 128, 192, 256}, advanced together by fi_batch_integrate (value = member-kyr/s,
the sum over members of simulated kyr per wall-clock second); elra_ensemble is 256 ELRA members
with the same sides, each with its own relaxation-time anomaly (workloads.elra_ensemble), in 1-kyr
windows; inversion is the unscented Kalman inversion of two regional upper-mantle viscosities on
the N=256 disc benchmark (value = UKI iterations/s, each 5 forward runs to 3 kyr on the ensemble
engine plus the host update); surrogate is the N=350 workload run
in 500-yr macro steps by the trained surrogate with a plain-solver check every 10 macro steps and
fallback above 1 m (--validate adds the plain run of the same window for the error).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [ROOT, os.path.join(ROOT, "fastisostasy.jl_amd")]

from fastiso import _abi  # noqa: E402

HEADLINE_METRIC = "simulated kyr/s at N=350 LV-ELVA: plain vs surrogate at \u22641 m max |\u0394u| error"
BYTES_PER_CELL_STEP = 171.0
# ELRA, per cell per RHS call, counted once per unit: rows phase reads the base state (8) and on
# average 5/3 FP32 derivative slots (6.7), mask, bedrock offset, reference H_af, litho anomaly,
# u_ref and the two ice snapshots (25), writes the padded half spectrum (8), the loads for the
# sparse update (8) and on the last stage u_{n+1} (8/3); the column phase reads and writes the
# padded spectrum (16); the last phase reads it (8), the stage displacement (8) and 1/tau (4),
# writes dudt and u_eq (8) and on the last stage reads k1..k3 and u_n for the error norm (20/3)
ELRA_BYTES_PER_CELL_RHS = 8 + 20 / 3 + 25 + 8 + 8 + 8 / 3 + 16 + 8 + 8 + 4 + 8 + 20 / 3
# implicit step, per cell per step: rows phase reads u (8) and the per-cell fields (mask,
# bedrock offset, two ice snapshots, reference H_af, litho anomaly, u_ref: 25), writes the u and
# load half spectra (16) and the loads for the sparse update (8); the column phase reads both
# spectra (16) and writes one (8); the last phase reads it (8) and u (8) and writes u (8)
IMPLICIT_BYTES_PER_CELL_STEP = 8 + 25 + 16 + 8 + 16 + 8 + 8 + 8 + 8
HBM_PEAK_GBS = 8000.0
# measured HBM-side traffic of the persistent kernel at N=350 (rocprofv3 --pmc FETCH_SIZE and
# WRITE_SIZE, calibrated on the copy probe); scaled to this run's accepted steps per launch
TRAFFIC_N350 = os.path.join(ROOT, "profiles", "fi_persistent_lv_elva_n350_traffic.json")
TRAFFIC_N1024 = os.path.join(ROOT, "profiles", "fi_persistent_lv_elva_hr_n1024_traffic.json")
TRAFFIC_ENS = os.path.join(ROOT, "profiles", "fi_batch_persistent_ensemble256_traffic.json")


def traffic_fields(path, per_unit_key, units, launches):
    """roofline.traffic (HBM bytes per launch) from a committed PMC summary: its bytes per unit
    (accepted step or cell-step, rocprofv3 FETCH_SIZE / WRITE_SIZE calibrated on a copy kernel)
    times this run's units per launch."""
    if not os.path.exists(path) or launches <= 0:
        return {"traffic": None}
    with open(path) as f:
        per = json.load(f)[per_unit_key]
    return {"traffic": per * units / launches, "traffic_unit": "B per launch",
            "traffic_source": os.path.relpath(path, ROOT) + f" ({per_unit_key} from rocprofv3 --pmc FETCH_SIZE, "
                              "WRITE_SIZE, scaled to this run's units per launch)"}


def make_sim(n, engine, hr=False, elra=False, implicit=False):
    import fastiso
    from fastiso import workloads
    if implicit:
        w = workloads.disc_benchmark(n=n, t_end=1e6, t_out=[])
        w["solidearth"] = fastiso.SolidEarth(w["domain"], lithosphere=fastiso.RigidLithosphere(),
                                             layer_boundaries=[88e3], layer_viscosities=[1e21],
                                             rho_uppermantle=3300.0, rho_litho=0.0)
        w["opts"] = fastiso.SolverOptions(diffeq=fastiso.DiffEqOptions(alg="Euler", dt_min=100.0))
    elif elra:
        w = workloads.elra(n=n, W=3.15e6, t_end=50e3, tau="weak", lithosphere="constant", t_out=[])
        w["sealevel"] = fastiso.RegionalSeaLevel(surface=fastiso.LaterallyVariableSeaSurface())
    elif hr:
        w = workloads.lv_elva(n=n, W=3e6, t_end=120e3, seeds=(5001, 5002, 1003), t_out=[])
    else:
        w = workloads.lv_elva(n=n, t_end=120e3, t_out=[])
    return w, fastiso.Simulation(w["domain"], w["bcs"], w["sealevel"], w["solidearth"], w["t_span"],
                                 nout=w["nout"], opts=w["opts"], c=w["c"], engine=engine)


def ensemble_inputs(nmem, elra=False):
    import fastiso
    from fastiso import workloads
    ws = workloads.elra_ensemble(nmem=nmem) if elra else workloads.ensemble(nmem=nmem, t_end=20e3)
    return [fastiso.Simulation(w["domain"], w["bcs"], w["sealevel"], w["solidearth"], w["t_span"],
                               nout=w["nout"], opts=w["opts"], c=w["c"], engine=None).inputs for w in ws]


def bench_ensemble(a):
    import fastiso
    t0 = time.perf_counter()
    elra = a.workload == "elra_ensemble"
    inputs = ensemble_inputs(a.members, elra)
    ens = fastiso.Ensemble(inputs, team_size=a.team_size)
    ens.integrator_init()
    setup_s = time.perf_counter() - t0
    print(f"ensemble of {len(inputs)} members set up in {setup_s:.1f} s", file=sys.stderr, flush=True)
    tune_s = None
    if a.autotune == "all":
        t1 = time.perf_counter()
        ens.autotune_fft(attempts=20)
        tune_s = time.perf_counter() - t1
        print(f"FFT radix orders tuned in {tune_s:.1f} s", file=sys.stderr, flush=True)
    cells = [inp["cfg"].nx * inp["cfg"].ny for inp in inputs]
    t = 0.0
    for k in range(a.warmup):
        t += a.window
        ens.integrate(t, fsal=False)
        print(f"warm-up window {k + 1}/{a.warmup} done: {ens.info()}", file=sys.stderr, flush=True)
    acc0 = [ens[m].stats()["n_accept"] for m in range(len(inputs))]
    rhs0 = [ens[m].stats()["n_rhs"] for m in range(len(inputs))]
    i0 = ens.info()
    t0 = time.perf_counter()
    for k in range(a.steps):
        t += a.window
        ens.integrate(t, fsal=False)
        print(f"window {k + 1}/{a.steps} done", file=sys.stderr, flush=True)
    el = time.perf_counter() - t0
    i1 = ens.info()
    acc = [ens[m].stats()["n_accept"] - acc0[m] for m in range(len(inputs))]
    rhs = [ens[m].stats()["n_rhs"] - rhs0[m] for m in range(len(inputs))]
    cell_steps = float(sum(c * n for c, n in zip(cells, acc)))
    dev_s = (i1["device_ms"] - i0["device_ms"]) / 1e3
    member_kyr = len(inputs) * a.steps * a.window / 1e3
    if elra:
        alg_bytes = ELRA_BYTES_PER_CELL_RHS * float(sum(c * n for c, n in zip(cells, rhs)))
    else:
        alg_bytes = BYTES_PER_CELL_STEP * cell_steps
    achieved = alg_bytes / dev_s / 1e9 if dev_s > 0 else None
    sizes = sorted(set(inp["cfg"].nx for inp in inputs))
    kind = "ELRA" if elra else "LV-ELVA"
    ens.close()
    return {
        "metric": f"simulated member-kyr/s, {len(inputs)}-member uneven {kind} ensemble (plain solver)",
        "value": member_kyr / el,
        "unit": "member-kyr/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": el / a.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 fft/stencil-moments, f64 state/stencil/error-norm",
        "data": "synthetic (seeded fields)",
        "config": {"workload": a.workload, "members": len(inputs), "sizes": sizes, "total_cells": int(sum(cells)),
                   "window_yr": a.window, "integrator": "BS3 reltol 1e-5 abstol 1e-6",
                   "accepted_steps_total": int(sum(acc)), "rhs_calls_total": int(sum(rhs)), "cell_steps": cell_steps,
                   "team_size": i1["team_size"], "n_teams": i1["n_teams"], "autotuned": a.team_size == 0,
                   "wg_threads": i1["wg_threads"],
                   "kernel_launches": int(i1["n_launch"] - i0["n_launch"]), "setup_s": setup_s,
                   "fft_plans_tuned_s": tune_s},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": (achieved / HBM_PEAK_GBS) if achieved else None,
                     **(traffic_fields(TRAFFIC_ENS, "bytes_per_cell_step", cell_steps,
                                       int(i1["n_launch"] - i0["n_launch"])) if not elra else {"traffic": None})},
    }


def host_cores():
    """Host cores this process may use: OMP_NUM_THREADS where the host sets it (a one-GPU job's
    share of the box), else the CPU affinity set."""
    env = os.environ.get("OMP_NUM_THREADS")
    return int(env) if env else len(os.sched_getaffinity(0))


def cpu_baseline(n, window, sample_s, hr=False, elra=False, implicit=False, start=None):
    """The FP64 oracle on this host, on all cores and on one core, each for about sample_s of
    wall time. With start = (t, u) both legs begin from the GPU leg's state at the start of its
    timed windows (fi_set_field + fi_set_time, then a 1/100-window untimed lead-in so the step
    size settles), so they sample the same part of the run; otherwise after one window from rest.
    The implicit step takes fixed steps of dt_min and refuses fi_set_time (its calls advance u), so
    its legs start from rest: a step costs the same whatever the state."""
    if implicit:
        start = None
    from oracle import pyref
    w, sim = make_sim(n, engine=None, hr=hr, elra=elra, implicit=implicit)
    legs = []
    for cores in (host_cores(), 1):
        pyref.set_threads(cores)
        e = pyref.RefEngine.from_inputs(sim.inputs)
        e.integrator_init()
        if start is not None:
            e.set_field("u", start[1])
            e.set_time(start[0])
            e.integrate(start[0] + window / 100)
        else:
            e.integrate(window)
        s0 = e.stats()
        t0 = time.perf_counter()
        last = t0
        while time.perf_counter() - t0 < sample_s:
            e.integrate(e.stats()["t"] + window / 10)
            if time.perf_counter() - last > 10:
                last = time.perf_counter()
                print(f"cpu baseline ({cores} cores): t = {e.stats()['t']:.2f} yr", file=sys.stderr, flush=True)
        el = time.perf_counter() - t0
        st = e.stats()
        sim_yr = st["t"] - s0["t"]
        legs.append({"value": sim_yr / 1e3 / el, "unit": "kyr/s", "cores": cores, "kind": "port",
                     "sample": f"oracle FP64, n={n}, t={s0['t']:.2f}->{st['t']:.2f} yr "
                               + ("from the GPU state at the start of the timed windows" if start is not None
                                  else f"after a {window:.0f}-yr warm-up")
                               + f", {el:.1f} s wall, {st['n_accept'] - s0['n_accept']} accepted steps"})
        print(f"cpu baseline ({cores} cores): {legs[-1]['value']:.3g} kyr/s", file=sys.stderr, flush=True)
    pyref.set_threads(legs[0]["cores"])
    out = dict(legs[0])
    out["one_core"] = legs[1]
    return out


WEIGHTS = os.path.join(ROOT, "fastisostasy.jl_amd", "fastiso", "data", "fno_lv_elva_n350.npz")


def surrogate_leg(n, window_kyr, predictor="rkc", rkc_dt=250.0, check_every=10, macro_dt=500.0, tol=1.0,
                  weights=None, validate=True, hr=False):
    """Surrogate-with-fallback over [0, window_kyr] of the N=n lv_elva workload: macro steps of
    macro_dt years in periods of check_every (check_every - 1 predicted steps, then a plain-solver
    check of the next one), the period rolled back and integrated by the plain solver when the
    check deviates by more than tol. With validate the plain solver integrates the same window in
    a second simulation (stopping at the same macro-step times); its wall time gives the plain
    kyr/s of the same window, and max |u_surrogate - u_plain| is taken at every 5-kyr mark. Without
    it the plain solver's rate is that of the run's own plain checks (plain_kyr_per_s_in_run). hr:
    the N=1024 workload's fields and domain."""
    import numpy as np
    from fastiso import surrogate as S
    w, sim = make_sim(n, engine="hip", hr=hr)
    sur = None
    if weights is not None:
        d = np.load(weights, allow_pickle=False)
        cfgv = [int(v) for v in d["cfg"]]
        sur = S.Surrogate(in_channels=cfgv[0], width=cfgv[1], modes_x=cfgv[2], modes_y=cfgv[3], layers=cfgv[4],
                          proj_hidden=cfgv[5])
        sur.set_grid(n, n)
        sur.set_params(d["theta"])
    macro = S.lv_macro(w, macro_dt)
    eng = sim.engine
    eng.integrator_init()
    runner = S.SurrogateRunner(sim, sur, macro, check_every=check_every, tol=tol, predictor=predictor,
                               rkc_dt=rkc_dt)
    t_end = window_kyr * 1e3
    marks = {}

    def prog(t1, kind, dev):
        if abs(t1 / 5e3 - round(t1 / 5e3)) < 1e-9:
            marks[t1] = eng.get_field("u")
        print(f"surrogate run at {t1:.0f} yr: {kind} dev={dev}", file=sys.stderr, flush=True)

    t0 = time.perf_counter()
    res = runner.run(t_end, t0=0.0, progress=prog)
    el = time.perf_counter() - t0
    devs = [dv for _, kind, dv in runner.log if dv is not None]
    out = {"value": t_end / 1e3 / el, "unit": "kyr/s", "window_kyr": window_kyr, "predictor": predictor,
           "rkc_max_substep_yr": rkc_dt if predictor == "rkc" else None, "macro_step_yr": macro_dt,
           "check_every": check_every, "tolerance_m": tol, "wall_s": el,
           "max_check_deviation_m": max(devs) if devs else None, **res}
    if res["plain_s_in_run"] > 0:
        out["plain_kyr_per_s_in_run"] = res["plain_yr_in_run"] / 1e3 / res["plain_s_in_run"]
        out["speedup_vs_plain_in_run"] = out["value"] / out["plain_kyr_per_s_in_run"]
    if sur is not None:
        out["network_device_ms"] = sur.device_ms()
        sur.close()
    sim.close()
    if validate:
        w2, sim2 = make_sim(n, engine="hip", hr=hr)
        e2 = sim2.engine
        e2.integrator_init()
        errs = {}
        t0 = time.perf_counter()
        tk = 0.0
        while tk < t_end - 1e-9:
            tk = min(tk + macro_dt, t_end)
            e2.integrate(tk)
            if tk in marks:
                errs[f"{tk:.0f}"] = float(np.abs(e2.get_field("u") - marks[tk]).max())
                print(f"plain reference at {tk:.0f} yr: max |u_sur - u_plain| = {errs[f'{tk:.0f}']:.4f} m",
                      file=sys.stderr, flush=True)
        el_p = time.perf_counter() - t0
        sim2.close()
        out["plain_kyr_per_s_same_window"] = t_end / 1e3 / el_p
        out["plain_wall_s"] = el_p
        out["speedup_vs_plain"] = el_p / el
        out["max_abs_u_error_vs_plain_m"] = errs
        out["max_abs_u_error_m"] = max(errs.values()) if errs else None
    return out


def bench_rkc(a):
    """RKC2 macro steps of 500 yr (sub-steps of at most --rkc-dt yr) at N=350 from the BS3 state
    at 100 yr: the first call estimates the spectral radius and captures the chain's HIP graph,
    the next --steps calls are timed one by one (FASTISO_RKC_GRAPH=0 launches the chain
    directly)."""
    w, sim = make_sim(a.n, engine="hip")
    e = sim.engine
    e.integrator_init()
    e.integrate(100.0)
    t = 100.0
    first = None
    times, graphs = [], []
    steps = a.steps or 10
    for k in range(steps + 1):
        t += 500.0
        t0 = time.perf_counter()
        info = e.integrate_rkc(t, a.rkc_dt)
        el = time.perf_counter() - t0
        if k == 0:
            first = {"wall_ms": el * 1e3, **info}
        else:
            times.append(el * 1e3)
            graphs.append(info["graph"])
    sim.close()
    ms = sorted(times)[len(times) // 2]
    return {"metric": f"RKC2 500-yr macro steps at N={a.n} LV-ELVA (ms per macro step, median)", "value": ms,
            "unit": "ms", "n_gpus": 1, "steps": steps, "warmup": 1, "higher_is_better": False,
            "dtype": "f32 fft/stencil-moments, f64 state", "data": "synthetic (seeded fields)",
            "config": {"workload": f"rkc_n{a.n}", "rkc_max_substep_yr": a.rkc_dt, "substeps": info["substeps"],
                       "stages": info["stages"], "rhs_per_macro_step": info["n_rhs"], "graph_states": graphs,
                       "ms_per_macro_step": times, "first_call": first,
                       "graph_disabled": os.environ.get("FASTISO_RKC_GRAPH") == "0"}}


def bench_surrogate(a):
    """--workload surrogate: the surrogate leg alone, over --window-kyr."""
    leg = surrogate_leg(a.n, a.window_kyr, predictor=a.predictor, rkc_dt=a.rkc_dt, check_every=a.check_every,
                        weights=a.weights if a.predictor != "rkc" or a.with_net else None, validate=a.validate,
                        hr=a.n == 1024)
    return {
        "metric": f"simulated kyr/s at N={a.n} LV-ELVA (surrogate with plain-solver fallback)",
        "value": leg["value"], "unit": "kyr/s", "n_gpus": 1, "steps": leg["checks"], "warmup": 0,
        "ms_per_step": leg["wall_s"] / max(leg["checks"], 1) * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32 fft/stencil-moments, f64 state (RKC2 predictor); plain-solver checks",
        "data": "synthetic (seeded fields)",
        "config": {"workload": f"surrogate_lv_elva_n{a.n}", "n": a.n, **leg},
    }


def inversion_setup(n):
    """Two-region upper-mantle viscosity inversion on the disc benchmark (n x n, W = 3000 km):
    log10 viscosity inside and outside r = 1000 km (truth 21.0 / 20.6), displacement data at 1, 2
    and 3 kyr within r < 1500 km, prior mean 20.7 / 20.9 with sd 0.5; 5 sigma points per UKI
    iteration."""
    import numpy as np
    import fastiso
    from fastiso import inversion, workloads
    w = workloads.disc_benchmark(n=n, t_end=3000.0)
    d = w["domain"]

    def builder(eta):
        se = fastiso.SolidEarth(d, layer_boundaries=[88e3], layer_viscosities=[eta], rho_uppermantle=3300.0,
                                rho_litho=0.0)
        return fastiso.Simulation(d, w["bcs"], w["sealevel"], se, (0.0, 3000.0), nout=fastiso.NativeOutput(),
                                  opts=w["opts"], c=w["c"], engine=None).inputs

    red = inversion.RegionalViscosityReduction(np.full((d.nx, d.ny), 21.0), [d.R < 1.0e6, d.R >= 1.0e6], builder)
    return red, d, (1000.0, 2000.0, 3000.0)


def bench_inversion(a):
    import numpy as np
    import fastiso
    from fastiso import inversion
    n = a.inv_n
    red, d, tobs = inversion_setup(n)
    log = []

    class RecordingEnsemble(fastiso.Ensemble):
        def __init__(self, inputs_list, **kw):
            super().__init__(inputs_list, **kw)
            self._cells = [inp["cfg"].nx * inp["cfg"].ny for inp in inputs_list]

        def close(self):
            if getattr(self, "h", None):
                i = self.info()
                cs = sum(c * self[m].stats()["n_accept"] for m, c in enumerate(self._cells))
                log.append((i["device_ms"], cs, i["team_size"]))
            super().close()

    truth = fastiso.Ensemble([red.reconstruct(np.array([21.0, 20.6]))])
    truth.integrator_init()
    Y = []
    for tk in tobs:
        truth.integrate(tk)
        Y.append(truth[0].get_field("u"))
    truth.close()
    data = inversion.InversionData(list(tobs), Y, d.R < 1.5e6)
    prior = inversion.GaussianPrior([20.7, 20.9], [0.5, 0.5])
    cfg = inversion.InversionConfig("unscented", N_iter=a.warmup + a.steps, alpha_reg=1.0, update_freq=1,
                                    scale_obscov=1e-6)
    prob = inversion.inversion_problem(cfg, data, red, prior, make_batch=RecordingEnsemble, seed=5)

    def iteration():
        U = prob.process.sigma_points()
        prob.G_ens = prob.forward(prob.prior.to_physical(U))
        prob.process.update(prob.G_ens, prob.y_obs, prob.Gamma)
        prob.out.append(prob.prior.to_physical(prob.process.u_mean[-1]))

    for k in range(a.warmup):
        iteration()
        print(f"warm-up iteration {k + 1}/{a.warmup} done", file=sys.stderr, flush=True)
    n0 = len(log)
    t0 = time.perf_counter()
    for k in range(a.steps):
        iteration()
        print(f"iteration {k + 1}/{a.steps} done: mean {prob.out[-1]}", file=sys.stderr, flush=True)
    el = time.perf_counter() - t0
    dev_s = sum(x[0] for x in log[n0:]) / 1e3
    cell_steps = float(sum(x[1] for x in log[n0:]))
    achieved = BYTES_PER_CELL_STEP * cell_steps / dev_s / 1e9 if dev_s > 0 else None
    out = {
        "metric": f"UKI iterations/s, two-region viscosity inversion on the N={n} disc (5 sigma points, 3 kyr)",
        "value": a.steps / el, "unit": "iterations/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": el / a.steps * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 fft/stencil-moments, f64 state/stencil/error-norm", "data": "synthetic (disc benchmark)",
        "config": {"workload": f"inversion_disc_n{n}", "n": n, "members_per_iteration": prob.process.N_ens,
                   "obs_times_yr": list(tobs), "device_s": dev_s, "cell_steps": cell_steps,
                   "team_size": log[-1][2] if log else None, "final_mean_log10_eta": [float(x) for x in prob.out[-1]],
                   "truth_log10_eta": [21.0, 20.6]},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": (achieved / HBM_PEAK_GBS) if achieved else None, "traffic": None},
    }
    if not a.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline_inversion(n, prob, red)
    return out


def cpu_baseline_inversion(n, prob, red):
    """One UKI forward map (the 5 sigma-point runs, 0 -> 3 kyr) on the FP64 oracle ensemble."""
    threads = str(host_cores())
    from oracle import pyref
    pyref.set_threads(int(threads))
    U = prob.process.sigma_points()
    phys = prob.prior.to_physical(U)
    inputs = [red.reconstruct(phys[:, j]) for j in range(phys.shape[1])]
    t0 = time.perf_counter()
    ens = pyref.RefEnsemble(inputs)
    ens.integrator_init()
    for tk in prob.data.t:
        ens.integrate(tk)
    el = time.perf_counter() - t0
    return {"value": 1.0 / el, "unit": "iterations/s", "cores": int(threads), "kind": "port",
            "sample": f"oracle FP64 ensemble, one iteration's {len(inputs)} forward runs at n={n}, 0 -> 3 kyr "
                      f"(the UKI update, host numpy in both legs, takes milliseconds), {el:.1f} s wall"}


def bench_fno_train(a):
    """Device training throughput of the surrogate network at N=350 (width 32, 16x16 modes, 4
    layers, 5 input channels) on a synthetic set of 64 samples: Adam steps of batch 8."""
    import numpy as np
    from fastiso import surrogate as S
    n = a.n
    rng = np.random.default_rng(7)
    X = rng.standard_normal((64, n * n, 5)).astype(np.float32)
    Y = rng.standard_normal((64, n * n)).astype(np.float32)
    sur = S.Surrogate(in_channels=5)
    sur.set_grid(n, n)
    sur.set_data(X, Y)
    sur.train(a.warmup, batch=8, lr=1e-4, seed=1)
    ms0 = sur.device_ms()
    t0 = time.perf_counter()
    sur.train(a.steps, batch=8, lr=1e-4, seed=2)
    el = time.perf_counter() - t0
    dev = (sur.device_ms() - ms0) / 1e3
    sur.close()
    samples = a.steps * 8
    return {
        "metric": f"surrogate training samples/s at N={n} (forward + backward + Adam)",
        "value": samples / el, "unit": "samples/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": el / a.steps * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "bf16 mfma, fp32 accumulate", "data": "synthetic",
        "config": {"workload": f"fno_train_n{n}", "n": n, "batch": 8, "params": sur.n_params,
                   "device_s": dev, "ms_per_sample_device": dev / samples * 1e3},
    }


def bench_hbm_copy(a):
    """A plain streaming kernel of known size: out = 1.0 * x on 2**28 FP32 values (1 GiB read,
    1 GiB written per launch), --steps launches after --warmup. It calibrates the HBM traffic
    counters (rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE per dispatch against the known bytes) and
    shows the copy bandwidth the box attains."""
    import torch
    n = 1 << 28
    x = torch.ones(n, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    for _ in range(a.warmup):
        torch.mul(x, 1.0, out=y)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(a.steps):
        torch.mul(x, 1.0, out=y)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / a.steps
    moved = 2.0 * n * 4
    return {"metric": "HBM copy bandwidth (counter calibration kernel)", "value": moved / ms / 1e6, "unit": "GB/s",
            "n_gpus": 1, "steps": a.steps, "warmup": a.warmup, "ms_per_step": ms, "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
            "config": {"workload": "hbm_copy", "bytes_read_per_launch": n * 4, "bytes_written_per_launch": n * 4}}


def profile_suite(a):
    """The workloads whose kernels profiles/ covers, in one process, so that one profiler pass
    (rocprofv3 --kernel-trace, or one --pmc counter set) collects all of them: the HBM copy
    calibration kernel, the N=350 plain solver (3 windows after 1), the 256-member ensemble (2
    windows after 1, at the tuner's team size), the N=1024 instance (1 window after 1) and the RKC
    predictor (one 500-yr macro step at N=350). One JSON line per part. The parts run the engine's default FFT plans and
    grid size (no autotuning): the profiles describe those kernels, and their dispatches stay
    free of tuning launches."""
    import copy
    parts = []
    a = copy.copy(a)
    a.autotune = "none"
    b = copy.copy(a)
    b.steps, b.warmup = 5, 2
    parts.append(bench_hbm_copy(b))
    for wl, steps, warm in (("lv_elva", 3, 1), ("lv_elva_hr", 1, 1)):
        b = copy.copy(a)
        b.workload, b.steps, b.warmup, b.no_cpu_baseline, b.no_surrogate, b.n = wl, steps, warm, True, True, 350
        b.window = 5.0 if wl == "lv_elva_hr" else 10.0
        parts.append(run_plain(b))
    b = copy.copy(a)
    # the team size the tuner settles on for this ensemble (32 workgroups of 256 threads), fixed
    # so that the profiled launches are whole windows rather than the tuner's trials
    b.workload, b.steps, b.warmup, b.window, b.team_size = "ensemble", 2, 1, 10.0, 32
    parts.append(bench_ensemble(b))
    w, sim = make_sim(350, engine="hip")
    e = sim.engine
    e.integrator_init()
    e.integrate(100.0)
    t0 = time.perf_counter()
    info = e.integrate_rkc(600.0, 250.0)
    el = time.perf_counter() - t0
    sim.close()
    parts.append({"workload": "rkc_macro_step_n350", "wall_s": el, **info})
    return parts


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None, help="timed windows (default 200; 3 for lv_elva_hr)")
    ap.add_argument("--warmup", type=int, default=None, help="untimed windows (default 5; 1 for lv_elva_hr)")
    ap.add_argument("--n", type=int, default=350)
    ap.add_argument("--window", type=float, default=None, help="window in yr (default 10; 5 for lv_elva_hr)")
    ap.add_argument("--cpu-sample-s", type=float, default=15.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--workload", choices=["lv_elva", "lv_elva_hr", "elra", "implicit", "ensemble", "elra_ensemble", "inversion",
                                               "surrogate", "rkc", "fno_train", "hbm_copy", "profile_suite"],
                    default="lv_elva")
    ap.add_argument("--weights", default=WEIGHTS)
    ap.add_argument("--window-kyr", type=float, default=10.0, help="surrogate window (kyr)")
    ap.add_argument("--validate", action="store_true")
    ap.add_argument("--predictor", choices=["rkc", "net", "extrapolation"], default="rkc")
    ap.add_argument("--with-net", action="store_true", help="RKC predictor corrected by the network (--weights)")
    ap.add_argument("--rkc-dt", type=float, default=250.0, help="largest RKC sub-step of the predictor (yr)")
    ap.add_argument("--check-every", type=int, default=10)
    ap.add_argument("--no-surrogate", action="store_true", help="default workload: skip the surrogate leg")
    ap.add_argument("--autotune", choices=["none", "workgroups", "all"], default="all",
                    help="tune the grid size and FFT radix orders (ensembles: the radix orders) before timing")
    ap.add_argument("--members", type=int, default=256)
    ap.add_argument("--team-size", type=int, default=0)
    ap.add_argument("--inv-n", type=int, default=256, help="grid side of the inversion workload")
    a = ap.parse_args()
    hr_defaults = a.workload == "lv_elva_hr"
    elra = a.workload == "elra"
    implicit = a.workload == "implicit"
    elra_ens = a.workload == "elra_ensemble"
    inv = a.workload == "inversion"
    if a.steps is None:
        a.steps = 3 if hr_defaults else 4 if inv else 10 if (elra_ens or a.workload == "rkc") else 50 if (elra or implicit) else 200
    if a.warmup is None:
        a.warmup = 1 if (hr_defaults or inv) else 8 if elra_ens else 2 if (elra or implicit) else 5
    if a.window is None:
        a.window = 5.0 if hr_defaults else 100.0 if elra else 1000.0 if (implicit or elra_ens) else 10.0
    if a.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU runs are not part of it",
                          "n_gpus": a.gpus}))
        sys.exit(2)
    if a.workload == "inversion":
        print(json.dumps(bench_inversion(a)), flush=True)
        return
    if a.workload in ("ensemble", "elra_ensemble"):
        print(json.dumps(bench_ensemble(a)), flush=True)
        return
    if a.workload == "surrogate":
        print(json.dumps(bench_surrogate(a)), flush=True)
        return
    if a.workload == "rkc":
        print(json.dumps(bench_rkc(a)), flush=True)
        return
    if a.workload == "fno_train":
        print(json.dumps(bench_fno_train(a)), flush=True)
        return
    if a.workload == "hbm_copy":
        print(json.dumps(bench_hbm_copy(a)), flush=True)
        return
    if a.workload == "profile_suite":
        for part in profile_suite(a):
            print(json.dumps(part), flush=True)
        return
    print(json.dumps(run_plain(a)), flush=True)


def run_plain(a):
    """The single-context workloads (lv_elva, lv_elva_hr, elra, implicit): returns the bench line."""
    elra = a.workload == "elra"
    implicit = a.workload == "implicit"
    hr = a.workload == "lv_elva_hr"
    if hr and a.n == 350:
        a.n = 1024
    headline = a.workload == "lv_elva" and a.n == 350 and not a.no_surrogate
    # the engine owns its device memory and stream; every fi_integrate call returns after its
    # stream has synchronised, so wall-clock brackets below time completed device work
    w, sim = make_sim(a.n, engine="hip", hr=hr, elra=elra, implicit=implicit)
    eng = sim.engine
    eng.integrator_init()
    t = 0.0
    for _ in range(a.warmup):
        t += a.window
        eng.integrate(t)
    tune = None
    if a.autotune != "none":
        # timed on this run's own state after the warm-up (restored afterwards)
        what = _abi.FI_TUNE_WORKGROUPS | (_abi.FI_TUNE_FFT_PLANS if a.autotune == "all" else 0)
        info = eng.autotune(what, attempts=60)
        tune = {"what": a.autotune, "workgroups": info["workgroups"],
                "wg_candidates": info["wg_candidates"], "wg_us_per_attempt": info["wg_us"],
                "plans": [_abi.plan_radices(c) for c in info["plans"]],
                "plan_default_us_per_attempt": info["plan_default_us"],
                "plan_best_us_per_attempt": info["plan_best_us"]}
        print(f"autotune: {tune}", file=sys.stderr, flush=True)
    start = (eng.stats()["t"], eng.get_field("u"))
    s0 = eng.stats()
    t0 = time.perf_counter()
    for k in range(a.steps):
        t += a.window
        eng.integrate(t)
        if hr:   # windows of the N=1024 grid take tens of seconds each
            print(f"window {k + 1}/{a.steps} done", file=sys.stderr, flush=True)
    el = time.perf_counter() - t0
    s1 = eng.stats()
    kyr = a.steps * a.window / 1e3
    N = a.n * a.n
    acc = s1["n_accept_timed"] - s0["n_accept_timed"]
    dev_s = (s1["device_ms"] - s0["device_ms"]) / 1e3
    launches = s1["n_launch"] - s0["n_launch"]
    rhs = s1["n_rhs"] - s0["n_rhs"]
    if elra:
        achieved = ELRA_BYTES_PER_CELL_RHS * N * rhs / dev_s / 1e9 if dev_s > 0 else None
    elif implicit:
        achieved = IMPLICIT_BYTES_PER_CELL_STEP * N * acc / dev_s / 1e9 if dev_s > 0 else None
    else:
        achieved = BYTES_PER_CELL_STEP * N * acc / dev_s / 1e9 if dev_s > 0 else None
    tr = {"traffic": None}
    if a.n == 350 and not (hr or elra or implicit):
        tr = traffic_fields(TRAFFIC_N350, "bytes_per_accepted_step", acc, launches)
    elif hr and a.n == 1024:
        tr = traffic_fields(TRAFFIC_N1024, "bytes_per_accepted_step", acc, launches)
    engine_name, engine_lib = _abi.engine_identity(eng.lib)
    out = {
        "metric": (f"simulated kyr/s at N={a.n} " + ("ELRA (plain solver)" if elra else "implicit LC step (Euler, dt_min 100 yr)"
                                                      if implicit else "LV-ELVA (plain solver)")) if not headline else HEADLINE_METRIC,
        "value": kyr / el,
        "unit": "kyr/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": el / a.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 fft/stencil-moments, f64 state/stencil/error-norm",
        "data": "synthetic (seeded fields)",
        "config": {"workload": f"{a.workload}_n{a.n}", "n": a.n, "dx_km": 2 * w["domain"].Wx / a.n / 1e3,
                   "window_yr": a.window,
                   "integrator": "Euler dt_min 100 yr" if implicit else "BS3 reltol 1e-5 abstol 1e-6", "accepted_steps": int(s1["n_accept"] - s0["n_accept"]),
                   "rejected_steps": int(s1["n_reject"] - s0["n_reject"]),
                   "rhs_calls": int(s1["n_rhs"] - s0["n_rhs"]), "sim_t_end_yr": s1["t"],
                   "kernel_launches": int(launches), "kernel_ms_per_launch": dev_s * 1e3 / max(launches, 1),
                   "us_per_accepted_step": dev_s * 1e6 / max(acc, 1)},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": (achieved / HBM_PEAK_GBS) if achieved else None, **tr},
    }
    if tune is not None:
        out["config"]["autotune"] = tune
    out["config"]["engine"] = engine_name
    out["config"]["engine_library"] = os.path.relpath(engine_lib, ROOT) if engine_lib else None
    sim.close()
    if headline:
        # the metric's second half: the surrogate-with-fallback mode over a bounded window, against
        # the plain solver over the same window (value above = the plain solver's kyr/s)
        out["surrogate"] = surrogate_leg(a.n, a.window_kyr, predictor=a.predictor, rkc_dt=a.rkc_dt,
                                         check_every=a.check_every,
                                         weights=a.weights if (a.predictor != "rkc" or a.with_net) else None)
    if not a.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(a.n, a.window, a.cpu_sample_s, hr, elra, implicit, start=start)
    return out


if __name__ == "__main__":
    main()
