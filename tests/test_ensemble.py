"""Uneven ensembles (fi_create_batch / fi_batch_integrate).

The ensemble engine splits the device into teams of workgroups that integrate members taken from
a queue. Every cross-workgroup sum runs over fixed per-row-pair partials, so a member's trajectory
cannot depend on the team size or on which team ran it: members are compared bit for bit with
single contexts (fi_integrate with the same output times) and across team sizes, and within the
trajectory tolerance with the FP64 oracle.
"""
import numpy as np
import pytest

from fastiso import workloads

from _cases import DRYRUN, sim_inputs

T_OUT = (40.0, 100.0)


def small_ensemble(nmem=6, t_end=100.0):
    ws = workloads.ensemble(nmem=nmem, sizes=(32, 48, 64), t_end=t_end)
    return [sim_inputs(w).inputs for w in ws]


def with_tout(inp, tout):
    d = dict(inp)
    d["tout"] = list(tout)
    return d


# ---- CPU: workload generator and the oracle's ensemble ----------------------------------------
def test_ensemble_workload_is_seeded_and_uneven():
    ws = workloads.ensemble(nmem=40, t_end=20e3)
    sizes = [w["domain"].nx for w in ws]
    assert set(sizes) <= {64, 96, 128, 192, 256} and len(set(sizes)) >= 4
    again = workloads.ensemble(nmem=3, t_end=20e3)
    for a, b in zip(ws, again):
        assert a["domain"].nx == b["domain"].nx
        np.testing.assert_array_equal(a["logeta"], b["logeta"])
    # one shared dome field: members of the same size carry the same load, fields differ
    by_n = {}
    for w in ws:
        by_n.setdefault(w["domain"].nx, []).append(w)
    pairs = [v for v in by_n.values() if len(v) >= 2]
    assert pairs
    a, b = pairs[0][:2]
    np.testing.assert_array_equal(a["H_dome"], b["H_dome"])
    assert not np.array_equal(a["logeta"], b["logeta"])


def test_oracle_ensemble_equals_single_contexts(ref):
    inputs = small_ensemble(nmem=4)
    ens = ref.RefEnsemble(inputs)
    ens.integrator_init()
    for t in T_OUT:
        ens.integrate(t)
    for m, inp in enumerate(inputs):
        e = ref.RefEngine.from_inputs(with_tout(inp, T_OUT))
        e.integrator_init()
        e.integrate(T_OUT[-1], lambda t, k: None)
        np.testing.assert_array_equal(e.get_field("u"), ens[m].get_field("u"))
        assert e.stats()["n_accept"] == ens[m].stats()["n_accept"]


# ---- GPU --------------------------------------------------------------------------------------
def hip_ensemble(inputs, team_size):
    if DRYRUN:
        from oracle import pyref
        return pyref.RefEnsemble(inputs)
    import fastiso
    return fastiso.Ensemble(inputs, team_size=team_size)


def hip_single(inp):
    if DRYRUN:
        from oracle import pyref
        return pyref.RefEngine.from_inputs(inp)
    from fastiso.engine import Engine
    from fastiso.simulation import push_inputs
    return push_inputs(Engine(inp["cfg"]), inp)


def run_ensemble(inputs, team_size):
    ens = hip_ensemble(inputs, team_size)
    ens.integrator_init()
    for t in T_OUT:
        ens.integrate(t)
    out = [(ens[m].get_field("u"), ens[m].get_field("ue"), ens[m].stats()) for m in range(len(inputs))]
    info = None if DRYRUN else ens.info()
    return out, info


@pytest.mark.gpu
def test_ensemble_members_equal_single_contexts_bitwise():
    inputs = small_ensemble()
    res, info = run_ensemble(inputs, team_size=4)
    for m, inp in enumerate(inputs):
        e = hip_single(with_tout(inp, T_OUT))
        e.integrator_init()
        e.integrate(T_OUT[-1], lambda t, k: None)
        u, ue, st = res[m]
        np.testing.assert_array_equal(u, e.get_field("u"), err_msg=f"member {m}")
        np.testing.assert_array_equal(ue, e.get_field("ue"), err_msg=f"member {m}")
        s1 = e.stats()
        assert (st["n_accept"], st["n_reject"], st["n_rhs"]) == (s1["n_accept"], s1["n_reject"], s1["n_rhs"])
        assert st["t"] == T_OUT[-1]


@pytest.mark.gpu
@pytest.mark.parametrize("team_size", [1, 3, 0])
def test_ensemble_independent_of_team_size(team_size):
    inputs = small_ensemble()
    ref_res, _ = run_ensemble(inputs, team_size=8)
    res, info = run_ensemble(inputs, team_size=team_size)
    for m in range(len(inputs)):
        np.testing.assert_array_equal(res[m][0], ref_res[m][0], err_msg=f"member {m}")
    if info is not None:
        assert info["n_members"] == len(inputs) and info["n_launch"] >= 1
        if team_size:
            assert info["tuned"] == 1 and info["team_size"] == team_size
        # with team_size 0 this short run may end before every candidate size has been timed


@pytest.mark.gpu
def test_ensemble_builds_of_256_and_512_threads_agree_bitwise(monkeypatch):
    """Members up to 256 cells wide run the 256-thread build of the ensemble kernel (two
    workgroups per CU); FASTISO_BATCH_WG_THREADS=512 forces the 512-thread build, which wider
    members always use. Both sum in the canonical 512-lane order, so members agree bit for bit;
    an ensemble with a member wider than 256 cells runs the 512-thread build."""
    inputs = small_ensemble()
    res256, info256 = run_ensemble(inputs, team_size=4)
    monkeypatch.setenv("FASTISO_BATCH_WG_THREADS", "512")
    res512, info512 = run_ensemble(inputs, team_size=4)
    monkeypatch.delenv("FASTISO_BATCH_WG_THREADS")
    for m in range(len(inputs)):
        np.testing.assert_array_equal(res256[m][0], res512[m][0], err_msg=f"member {m}")
        np.testing.assert_array_equal(res256[m][1], res512[m][1], err_msg=f"member {m}")
        s0, s1 = res256[m][2], res512[m][2]
        assert (s0["n_accept"], s0["n_reject"], s0["n_rhs"]) == (s1["n_accept"], s1["n_reject"], s1["n_rhs"])
    if not DRYRUN:
        assert info256["wg_threads"] == 256 and info512["wg_threads"] == 512
        wide = [sim_inputs(w).inputs for w in workloads.ensemble(nmem=1, sizes=(300,), t_end=10.0)]
        ens = hip_ensemble(inputs[:1] + wide, 4)
        ens.integrator_init()
        ens.integrate(1.0)
        assert ens.info()["wg_threads"] == 512, ens.info()
        ens.close()


def scaled_reltol(inp, factor):
    import ctypes
    cfg = type(inp["cfg"])()
    ctypes.memmove(ctypes.byref(cfg), ctypes.byref(inp["cfg"]), ctypes.sizeof(cfg))
    cfg.reltol *= factor
    d = dict(inp)
    d["cfg"] = cfg
    return d


@pytest.mark.gpu
def test_ensemble_vs_oracle():
    """Early in the glacial cycle |u| is a few cm and two valid BS3 integrations differ at the
    level of the step-size tolerance: the bound is three times the oracle's own change when its
    reltol moves from 1e-5 to 0.9e-5, plus 1e-4 of max |u|."""
    from oracle import pyref
    inputs = small_ensemble()
    res, _ = run_ensemble(inputs, team_size=0)
    ref = pyref.RefEnsemble(inputs)
    ref.integrator_init()
    alt = pyref.RefEnsemble([scaled_reltol(inp, 0.9) for inp in inputs])
    alt.integrator_init()
    for t in T_OUT:
        ref.integrate(t)
        alt.integrate(t)
    for m in range(len(inputs)):
        b, eb = ref[m].get_field("u"), ref[m].get_field("ue")
        tol_u = 3 * np.abs(b - alt[m].get_field("u")).max() + 1e-4 * np.abs(b).max() + 1e-6
        tol_e = 3 * np.abs(eb - alt[m].get_field("ue")).max() + 1e-4 * np.abs(eb).max() + 1e-6
        u, ue, st = res[m]
        assert np.abs(u - b).max() <= tol_u, (m, np.abs(u - b).max(), tol_u)
        assert np.abs(ue - eb).max() <= tol_e, (m, np.abs(ue - eb).max(), tol_e)
        assert st["count_sparse_updates"] == ref[m].stats()["count_sparse_updates"]


@pytest.mark.gpu
def test_ensemble_mixing_distorted_members_equals_single_contexts_bitwise():
    """With one distorted member the whole launch runs the kernel with per-cell grid steps; the
    undistorted members take K = 1 there, and every member still equals its own single context."""
    inputs = small_ensemble(nmem=4)
    d = inputs[1]["cfg"]
    nx, ny = d.nx, d.ny
    x = np.linspace(-1.0, 1.0, nx)[:, None]
    y = np.linspace(-1.0, 1.0, ny)[None, :]
    inputs[1] = dict(inputs[1], fields=dict(inputs[1]["fields"], distortion=0.96 + 0.05 * (x ** 2 + y ** 2)))
    res, _ = run_ensemble(inputs, team_size=4)
    for m, inp in enumerate(inputs):
        e = hip_single(with_tout(inp, T_OUT))
        e.integrator_init()
        e.integrate(T_OUT[-1], lambda t, k: None)
        np.testing.assert_array_equal(res[m][0], e.get_field("u"), err_msg=f"member {m}")


@pytest.mark.gpu
def test_ensemble_wide_distorted_members_equal_single_contexts_bitwise():
    """Members wider than 512 cells keep FP32 plate moments; with a distorted one the launch runs
    that path's distortion instance (1/K in registers), K = 1 for the other member. Both equal
    their single contexts bit for bit over a short window."""
    ws = workloads.ensemble(nmem=2, sizes=(525,), t_end=1.0)
    inputs = [sim_inputs(w).inputs for w in ws]
    d = inputs[0]["cfg"]
    x = np.linspace(-1.0, 1.0, d.nx)[:, None]
    y = np.linspace(-1.0, 1.0, d.ny)[None, :]
    inputs[0] = dict(inputs[0], fields=dict(inputs[0]["fields"], distortion=0.97 + 0.04 * (x ** 2 + y ** 2)))
    tout = (0.2, 0.4)
    ens = hip_ensemble(inputs, 0)
    ens.integrator_init()
    for t in tout:
        ens.integrate(t)
    res = [(ens[m].get_field("u"), ens[m].stats()) for m in range(len(inputs))]
    if not DRYRUN:
        ens.close()
    for m, inp in enumerate(inputs):
        e = hip_single(with_tout(inp, tout))
        e.integrator_init()
        e.integrate(tout[-1], lambda t, k: None)
        np.testing.assert_array_equal(res[m][0], e.get_field("u"), err_msg=f"member {m}")
        assert res[m][1]["n_accept"] == e.stats()["n_accept"] > 0
        if not DRYRUN:
            e.close()


# ---- relaxed-mantle (ELRA) ensembles ------------------------------------------------------------
def elra_ensemble(sizes=(32, 40, 48, 36), t_end=2000.0):
    import fastiso
    out = []
    for m, n in enumerate(sizes):
        w = workloads.elra(n=n, W=3e6, t_end=20e3, tau=2000.0 + 500.0 * m, lithosphere="constant",
                           t_out=[])
        w["sealevel"] = fastiso.RegionalSeaLevel(surface=fastiso.LaterallyVariableSeaSurface())
        out.append(sim_inputs(w).inputs)
    return out


def test_oracle_elra_ensemble_equals_single_contexts():
    from oracle import pyref
    inputs = elra_ensemble()
    ens = pyref.RefEnsemble(inputs)
    ens.integrator_init()
    ens.integrate(600.0)
    for m, inp in enumerate(inputs):
        e = pyref.RefEngine.from_inputs(with_tout(inp, [600.0]))
        e.integrator_init()
        e.integrate(600.0, lambda t, k: None)
        np.testing.assert_array_equal(ens[m].get_field("u"), e.get_field("u"))


@pytest.mark.gpu
@pytest.mark.parametrize("team_size", [2, 0])
def test_gpu_elra_ensemble_members_equal_single_contexts_bitwise(team_size):
    """Relaxed-mantle members (uneven sides, their own tau) in one ensemble kernel: each equals its
    single-context run bit for bit, whatever the team size; and the oracle's ensemble to 2e-3 of
    max |u| (the oracle against itself at reltol 0.7e-5 and 1.3e-5 differs by up to 1.2e-3 on these
    members at 600 yr)."""
    from oracle import pyref
    inputs = elra_ensemble()
    ens = hip_ensemble(inputs, team_size)
    ens.integrator_init()
    ens.integrate(300.0)
    ens.integrate(600.0)
    ref = pyref.RefEnsemble(inputs)
    ref.integrator_init()
    ref.integrate(300.0)
    ref.integrate(600.0)
    for m, inp in enumerate(inputs):
        e = hip_single(with_tout(inp, [300.0, 600.0]))
        e.integrator_init()
        e.integrate(600.0, lambda t, k: None)
        np.testing.assert_array_equal(ens[m].get_field("u"), e.get_field("u"), err_msg=f"member {m}")
        np.testing.assert_array_equal(ens[m].get_field("u_eq"), e.get_field("u_eq"), err_msg=f"member {m}")
        a, b = ens[m].get_field("u"), ref[m].get_field("u")
        assert np.abs(a - b).max() <= 2e-3 * np.abs(b).max(), m


@pytest.mark.gpu
def test_gpu_elra_ensemble_with_distorted_member_equals_single_contexts_bitwise():
    """An ELRA member with map distortion (its load enters the convolution times K^2, read per cell
    in the batched rows phase) beside undistorted ones, on a team size that batches several row
    pairs and columns per item: every member equals its single-context run bit for bit."""
    inputs = elra_ensemble()
    d = inputs[2]["cfg"]
    x = np.linspace(-1.0, 1.0, d.nx)[:, None]
    y = np.linspace(-1.0, 1.0, d.ny)[None, :]
    inputs[2] = dict(inputs[2], fields=dict(inputs[2]["fields"], distortion=0.95 + 0.06 * (x ** 2 + y ** 2)))
    ens = hip_ensemble(inputs, 2)
    ens.integrator_init()
    ens.integrate(300.0)
    ens.integrate(600.0)
    for m, inp in enumerate(inputs):
        e = hip_single(with_tout(inp, [300.0, 600.0]))
        e.integrator_init()
        e.integrate(600.0, lambda t, k: None)
        np.testing.assert_array_equal(ens[m].get_field("u"), e.get_field("u"), err_msg=f"member {m}")
        np.testing.assert_array_equal(ens[m].get_field("u_eq"), e.get_field("u_eq"), err_msg=f"member {m}")
    # the distortion is in effect: the same member without it ends elsewhere
    e0 = hip_single(with_tout(elra_ensemble()[2], [300.0, 600.0]))
    e0.integrator_init()
    e0.integrate(600.0, lambda t, k: None)
    assert np.abs(e0.get_field("u") - ens[2].get_field("u")).max() > 1e-4 * np.abs(e0.get_field("u")).max()


# ---- ensembles mixing member kinds ------------------------------------------------------------
def implicit_member(n=32):
    """The disc benchmark under a laterally constant lithosphere: the implicit step (Euler,
    dt_min 20 yr), which no ensemble kernel runs."""
    import fastiso
    w = workloads.disc_benchmark(n=n, t_end=1000.0, t_out=[])
    d = w["domain"]
    w["solidearth"] = fastiso.SolidEarth(d, lithosphere=fastiso.LaterallyConstantLithosphere(),
                                         layer_boundaries=[88e3], layer_viscosities=[1e21], rho_uppermantle=3300.0,
                                         rho_litho=0.0)
    w["opts"] = fastiso.SolverOptions(diffeq=fastiso.DiffEqOptions(alg="Euler", dt_min=20.0))
    return sim_inputs(w).inputs


def short_elra_member(n=36, t_end=150.0):
    """An ELRA member whose ice forcing ends at t_end (no flat extension)."""
    w = workloads.elra(n=n, W=3e6, t_end=t_end, tau=2500.0, lithosphere="constant", t_out=[])
    return sim_inputs(w).inputs


def mixed_ensemble():
    """LV-ELVA, ELRA and implicit-step members interleaved, so that the composite's member indices
    differ from its sub-ensembles' own."""
    lv = small_ensemble(nmem=2)
    el = elra_ensemble(sizes=(32, 40))
    return [lv[0], el[0], implicit_member(), lv[1], el[1], short_elra_member()]


def test_oracle_mixed_ensemble_equals_single_contexts():
    from oracle import pyref
    inputs = mixed_ensemble()
    ens = pyref.RefEnsemble(inputs)
    ens.integrator_init()
    for t in T_OUT:
        ens.integrate(t)
    for m, inp in enumerate(inputs):
        e = pyref.RefEngine.from_inputs(with_tout(inp, T_OUT))
        e.integrator_init()
        e.integrate(T_OUT[-1], lambda t, k: None)
        np.testing.assert_array_equal(ens[m].get_field("u"), e.get_field("u"), err_msg=f"member {m}")


@pytest.mark.gpu
def test_gpu_mixed_ensemble_members_equal_single_contexts_bitwise():
    """An ensemble of LV-ELVA, ELRA and implicit-step members: one sub-ensemble per ensemble
    kernel plus the implicit member through its own integration. Every member equals its
    single-context run bit for bit; errors name the caller's member index, from the composite's
    own checks and from a sub-ensemble's device error (the ELRA member whose ice ends at 150 yr)."""
    import fastiso
    inputs = mixed_ensemble()
    ens = hip_ensemble(inputs, 0)
    ens.integrator_init()
    for t in T_OUT:
        ens.integrate(t)
    for m, inp in enumerate(inputs):
        e = hip_single(with_tout(inp, T_OUT))
        e.integrator_init()
        e.integrate(T_OUT[-1], lambda t, k: None)
        np.testing.assert_array_equal(ens[m].get_field("u"), e.get_field("u"), err_msg=f"member {m}")
        assert ens[m].stats()["t"] == T_OUT[-1]
    assert np.abs(ens[2].get_field("u")).max() > 0
    if not DRYRUN:
        assert ens.info()["n_members"] == len(inputs)
    err = RuntimeError if DRYRUN else fastiso.EngineError
    t = T_OUT[-1]
    with pytest.raises(err, match="member 2"):
        ens.integrate_each([t, t, t - 1.0, t, t, t])
    with pytest.raises(err, match="member 5.*TimeInterpolation2D"):
        ens.integrate_each([t, t, t, t, t, 400.0])


@pytest.mark.gpu
@pytest.mark.parametrize("kind", ["elra", "lv"])
def test_gpu_ensemble_repeatable_across_runs_and_team_sizes(kind):
    """Members of the benchmark ensembles (sides 64..256) batch several row pairs and columns per
    item, as many as a team needs to cover them in one round; for ELRA members the sparse update's
    pointwise phase hands the row pairs out the same way (the rows phase after it reads those cells
    behind a workgroup barrier only). Two runs at one team size and a run at another agree bit for
    bit."""
    if DRYRUN:
        pytest.skip("the team mapping is the HIP ensemble kernel's")
    import fastiso
    ws = workloads.elra_ensemble(nmem=48) if kind == "elra" else workloads.ensemble(nmem=48, t_end=20e3)
    inputs = [sim_inputs(w).inputs for w in ws]
    assert {inp["cfg"].nx for inp in inputs} >= {64, 256}
    win = 1000.0 if kind == "elra" else 10.0
    runs = []
    for ts in ((16, 16, 4) if kind == "elra" else (32, 32, 8)):
        ens = fastiso.Ensemble(inputs, team_size=ts)
        ens.integrator_init()
        for t in (win, 2 * win):
            ens.integrate(t, fsal=False)
        runs.append([(ens[m].get_field("u"), ens[m].stats()["n_accept"]) for m in range(len(inputs))])
        ens.close()
    for r in runs[1:]:
        for m in range(len(inputs)):
            np.testing.assert_array_equal(runs[0][m][0], r[m][0], err_msg=f"member {m}")
            assert runs[0][m][1] == r[m][1], m


@pytest.mark.gpu
def test_gpu_elra_ensemble_tuner_settles_on_short_windows():
    """The ELRA members take far fewer than the tuner's 64 attempts per 100-yr window, so its
    capped trials run into the window end: it must switch to timing one whole window per
    candidate and fix a team size within a handful of windows. The members stay bitwise equal to
    a fixed-team-size run."""
    inputs = elra_ensemble()
    tuned = hip_ensemble(inputs, 0)
    fixed = hip_ensemble(inputs, 2)
    tuned.integrator_init()
    fixed.integrator_init()
    t = 0.0
    for _ in range(12):
        t += 100.0
        tuned.integrate(t)
        fixed.integrate(t)
    for m in range(len(inputs)):
        np.testing.assert_array_equal(tuned[m].get_field("u"), fixed[m].get_field("u"), err_msg=f"member {m}")
        assert tuned[m].stats()["n_accept"] / 12 < 64, "windows are not shorter than a trial"
    if not DRYRUN:
        t += 100.0
        tuned.integrate(t)
        info = tuned.info()
        assert info["tuned"] == 1 and info["team_size"] in info["cand_size"], info
        # members up to 256 cells wide run the 256-thread build, whose candidates start at 64
        assert info["wg_threads"] == 256 and info["cand_size"][0] == 64, info
        # the size in use is the fastest timed candidate, and later launches use it
        timed = [(r, sz) for sz, r in zip(info["cand_size"], info["cand_rate"]) if r > 0]
        assert timed and info["team_size"] == max(timed)[1], info
        assert info["last_team_size"] == info["team_size"], info


@pytest.mark.gpu
def test_ensemble_members_held_at_their_time_across_tuner_launches():
    """Members whose target equals their current time sit out a call while the others run
    capped tuner trials (several launches per call): every member stays bitwise equal to a single
    context stopped at the same times."""
    inputs = small_ensemble(t_end=1200.0)
    nm = len(inputs)
    plan = []   # per call: target of each member; 200-yr windows exceed the 64-attempt trials
    cur = [0.0] * nm
    for k in range(1, 7):
        tgt = [cur[m] if (m + k) % 3 == 0 else 200.0 * k for m in range(nm)]
        plan.append(tgt)
        cur = tgt
    ens = hip_ensemble(inputs, 0)
    ens.integrator_init()
    for tgt in plan:
        ens.integrate_each(tgt, fsal=True)
    for m, inp in enumerate(inputs):
        stops = sorted({tgt[m] for tgt in plan if tgt[m] > 0})
        e = hip_single(with_tout(inp, stops))
        e.integrator_init()
        e.integrate(stops[-1], lambda t, k: None)
        np.testing.assert_array_equal(ens[m].get_field("u"), e.get_field("u"), err_msg=f"member {m}")
        s0, s1 = ens[m].stats(), e.stats()
        assert (s0["t"], s0["n_accept"], s0["n_reject"]) == (s1["t"], s1["n_accept"], s1["n_reject"]), m
    if not DRYRUN:
        info = ens.info()
        assert info["n_launch"] > len(plan), info


@pytest.mark.gpu
def test_gpu_full_ensemble_one_window_sampled_members_bitwise():
    """The 256-member uneven ensemble of the ensemble benchmark (sides 64..256), one 10-yr window
    in one launch sequence with the tuner on; one sampled member of every side equals its single
    context bit for bit."""
    import bench
    if DRYRUN:
        pytest.skip("256 oracle members take minutes on the CPU")
    inputs = bench.ensemble_inputs(256)
    ens = hip_ensemble(inputs, 0)
    ens.integrator_init()
    ens.integrate(10.0)
    sides = {}
    for m, inp in enumerate(inputs):
        sides.setdefault(inp["cfg"].nx, m)
    assert sorted(sides) == [64, 96, 128, 192, 256], sorted(sides)
    for n, m in sorted(sides.items()):
        e = hip_single(with_tout(inputs[m], [10.0]))
        e.integrator_init()
        e.integrate(10.0, lambda t, k: None)
        np.testing.assert_array_equal(ens[m].get_field("u"), e.get_field("u"), err_msg=f"member {m} (side {n})")
        s0, s1 = ens[m].stats(), e.stats()
        assert (s0["n_accept"], s0["n_reject"], s0["n_rhs"]) == (s1["n_accept"], s1["n_reject"], s1["n_rhs"])
        e.close()
    info = ens.info()
    assert info["n_members"] == 256 and info["n_launch"] >= 1
    ens.close()
