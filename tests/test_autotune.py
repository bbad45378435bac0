"""Autotuner (fi_autotune, fi_set_fft_plans, fi_set_workgroups, fi_batch_autotune_fft).

* Plan codes (CPU): radices round-trip through the 4-bit code.
* The grid size is bitwise neutral: every cross-workgroup sum runs over fixed per-row-pair
  partials, so a run at any workgroup count equals the default one bit for bit.
* Any radix order is a correct transform: the RHS with reversed and rotated plans matches the
  FP64 oracle at the tolerances of the other RHS parity tests, on odd, even and mixed grids.
* The tuner restores the state: a run tuned half way equals a run that pins the chosen plans
  and grid size at the same point, bit for bit.
* Ensembles: after fi_batch_autotune_fft every member equals a single context with its plans.
"""
import numpy as np
import pytest

from fastiso import _abi

from _cases import DRYRUN, hip_sim, lv_case

RHS_MAX = 2e-6


def test_plan_code_round_trip():
    for radices in ([2, 5, 5, 7], [8, 8, 8, 2], [3], [7, 5]):
        assert _abi.plan_radices(_abi.plan_code(radices)) == radices


def rhs_rel(e, ref, d, times=(0.0, 3.0, 10.0)):
    rng = np.random.default_rng(4)
    u = -50 * np.exp(-(d.R / 1e6) ** 2) + 0.01 * rng.standard_normal((d.nx, d.ny))
    worst = 0.0
    for t in times:
        dg, _ = e.rhs(t, u)
        dr, _ = ref.rhs(t, u)
        worst = max(worst, np.abs(dg - dr).max() / np.abs(dr).max())
    return worst


def other_orders(code):
    """Other plans of the same length: the reversed and the rotated-by-one radix orders, and the
    first radix 8 or 4 split into 2 x 4 or 2 x 2 (when they differ from the plan)."""
    r = _abi.plan_radices(code)
    split = list(r)
    for i, x in enumerate(split):
        if x in (8, 4):
            split[i:i + 1] = [2, x // 2]
            break
    out = []
    for v in (r[::-1], r[1:] + r[:1], split):
        c = _abi.plan_code(v)
        if c != code and c not in out:
            out.append(c)
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("n", [35, 48, 64, 350])
def test_gpu_any_radix_order_matches_oracle(n):
    from oracle import pyref
    w = lv_case(n=n, seed=7)
    sim = hip_sim(w)
    ref = pyref.RefEngine.from_inputs(sim.inputs)
    e = sim.engine
    plans = e.fft_plans()
    assert _abi.plan_radices(plans[0]) and np.prod(_abi.plan_radices(plans[0])) == n
    ox, oy = other_orders(plans[0]), other_orders(plans[1])
    assert ox and oy
    for k in range(max(len(ox), len(oy))):
        cx, cy = ox[k % len(ox)], oy[::-1][k % len(oy)]
        e.set_fft_plans([cx, cy, 0, 0])
        assert e.fft_plans()[:2] == [cx, cy]
        m = rhs_rel(e, ref, w["domain"])
        assert m < RHS_MAX, (n, _abi.plan_radices(cx), _abi.plan_radices(cy), m)
    with pytest.raises(RuntimeError):
        e.set_fft_plans([_abi.plan_code([2, 2]), 0, 0, 0])   # product is not the length
    sim.close()


def run(w, t1, t2, tune=None, pin=None, g=None):
    sim = hip_sim(w)
    e = sim.engine
    e.integrator_init()
    e.integrate(t1)
    info = None
    if tune is not None:
        info = e.autotune(tune, attempts=12)
    if pin is not None:
        e.set_fft_plans(pin)
    if g is not None:
        e.set_workgroups(g)
    e.integrate(t2)
    out = e.get_field("u"), e.stats()
    sim.close()
    return out, info


@pytest.mark.gpu
def test_gpu_workgroup_count_is_bitwise_neutral():
    w = lv_case(n=48, seed=3, t_end=60.0)
    (u0, s0), _ = run(w, 5.0, 30.0)
    for g in (1, 5, 13):
        (u, s), _ = run(w, 5.0, 30.0, g=g)
        np.testing.assert_array_equal(u, u0, err_msg=f"g={g}")
        assert s["n_accept"] == s0["n_accept"]


@pytest.mark.gpu
def test_gpu_tuned_run_equals_pinned_run():
    w = lv_case(n=48, seed=3, t_end=60.0)
    (u_t, s_t), info = run(w, 5.0, 30.0, tune=_abi.FI_TUNE_WORKGROUPS | _abi.FI_TUNE_FFT_PLANS)
    if DRYRUN:
        return
    assert info["workgroups"] >= 1 and len(info["wg_candidates"]) >= 2
    assert info["n_plan_candidates"][0] >= 2 and info["plan_best_us"][0] <= info["plan_default_us"][0]
    (u_p, s_p), _ = run(w, 5.0, 30.0, pin=info["plans"], g=info["workgroups"])
    np.testing.assert_array_equal(u_t, u_p)
    assert (s_t["n_accept"], s_t["n_reject"], s_t["n_rhs"]) == (s_p["n_accept"], s_p["n_reject"], s_p["n_rhs"])
    assert s_t["t"] == 30.0


@pytest.mark.gpu
def test_gpu_ensemble_autotune_fft_members_equal_single_contexts():
    from test_ensemble import T_OUT, hip_ensemble, hip_single, small_ensemble, with_tout
    inputs = small_ensemble(nmem=4)
    ens = hip_ensemble(inputs, 4)
    ens.integrator_init()
    if not DRYRUN:
        ens.autotune_fft(attempts=8)
    for t in T_OUT:
        ens.integrate(t)
    for m, inp in enumerate(inputs):
        e = hip_single(with_tout(inp, T_OUT))
        e.set_fft_plans(ens[m].fft_plans())
        e.integrator_init()
        e.integrate(T_OUT[-1], lambda t, k: None)
        np.testing.assert_array_equal(ens[m].get_field("u"), e.get_field("u"), err_msg=f"member {m}")
