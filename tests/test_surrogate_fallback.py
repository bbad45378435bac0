"""The surrogate-with-fallback driver on the device at N=350 (the lv_elva workload), RKC predictor:
a check step whose prediction is perturbed beyond the 1-m tolerance rolls the period back to its
start, and the plain solver integrates it -- the state at the period's end equals, bit for bit,
a plain run that stops at the same macro-step times. The next period passes its check with the
unperturbed predictor."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_gpu_forced_fallback_equals_plain_run_bitwise():
    import bench
    from fastiso import surrogate as S
    w, sim = bench.make_sim(350, engine="hip")
    _, ref = bench.make_sim(350, engine="hip")
    macro = S.lv_macro(w, 50.0)
    eng = sim.engine
    eng.integrator_init()
    seen = {}

    def progress(t1, kind, dev):
        print(f"t = {t1:.0f} yr: {kind} {dev}", flush=True)
        if kind == "plain":
            seen[round(t1)] = eng.get_field("u")

    runner = S.SurrogateRunner(sim, None, macro, check_every=4, tol=1.0, predictor="rkc", rkc_dt=25.0,
                               perturb=lambda k, du: du + 2.0 if k == 1 else du)
    res = runner.run(400.0, t0=0.0, progress=progress)
    assert res["fallbacks"] == 1 and res["checks"] == 2 and res["surrogate_steps"] == 3, (res, runner.log)
    e2 = ref.engine
    e2.integrator_init()
    for t in (50.0, 100.0, 150.0, 200.0):
        e2.integrate(t)
    np.testing.assert_array_equal(seen[200], e2.get_field("u"))
    # the second period's unperturbed check is within the tolerance (RKC against BS3)
    devs = [d for _, k, d in runner.log if k.startswith("check-rkc") and not k.endswith("fallback")]
    assert devs and max(devs) < 0.05, runner.log
    sim.close()
    ref.close()
