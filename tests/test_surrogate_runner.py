"""Host logic of the surrogate-with-fallback driver (fastiso.surrogate.SurrogateRunner) on a scripted
engine, no GPU: the plain solver is replaced by the exact solution of du/dt = g(t), which
continues from whatever state a surrogate step wrote.

* A quadratic trajectory has increments linear in the step index, so the order-2 extrapolation
  is exact and every check passes. That holds only when each period's extrapolation is seeded
  with the increments of two consecutive macro steps; seeding it with two checks ten steps apart
  makes the next check fail.
* A kink in g(t) makes the check deviate by a known amount: above the tolerance the period is
  rolled back and integrated by the plain solver, so the final state equals the plain one.
"""
import numpy as np

from fastiso.surrogate import SurrogateRunner


class ScriptedEngine:
    def __init__(self, G, shape=(6, 5)):
        self.G = G                      # antiderivative of the rate g(t), per cell
        self.t = 0.0
        self.u = np.zeros(shape)
        self.n_integrate = 0

    def stats(self):
        return {"t": self.t}

    def get_field(self, name):
        return self.u.copy() if name == "u" else np.zeros_like(self.u)

    def set_field(self, name, v):
        assert name == "u"
        self.u = np.array(v, dtype=np.float64)

    def set_time(self, t):
        self.t = float(t)

    def integrate(self, t1):
        self.u = self.u + (self.G(t1) - self.G(self.t))
        self.t = float(t1)
        self.n_integrate += 1


class Macro:
    def __init__(self, dt):
        self.macro_dt = dt


class Sim:
    def __init__(self, eng):
        self.engine = eng


def run(G, t_end, dt=100.0, tol=1.0, check_every=10, order=2):
    eng = ScriptedEngine(G)
    r = SurrogateRunner(Sim(eng), None, Macro(dt), check_every=check_every, tol=tol, order=order,
                        predictor="extrapolation")
    res = r.run(t_end, t0=0.0)
    return eng, r, res


def test_quadratic_trajectory_passes_every_check():
    w = np.linspace(0.5, 2.0, 30).reshape(6, 5)
    a, b = 1e-5 * w, 1e-3 * w[::-1]

    def G(t):
        return a * t * t + b * t

    eng, r, res = run(G, 4200.0)
    # 2 warm plain steps, then 4 periods of 10 macro steps (9 surrogate + 1 check)
    assert res["fallbacks"] == 0, r.log
    assert res["surrogate_steps"] == 36 and res["plain_steps"] == 6
    devs = [d for _, kind, d in r.log if d is not None]
    assert len(devs) == 4 and max(devs) < 1e-9
    np.testing.assert_allclose(eng.u, G(4200.0), rtol=0, atol=1e-9)


def test_order_one_extrapolation_error_is_the_increment_change():
    # du/dt = 2 a t: increments grow by 2 a dt^2 per step; order 1 keeps the last increment
    a = 1e-4

    def G(t):
        return np.full((6, 5), a * t * t)

    eng, r, res = run(G, 1200.0, tol=1e9, order=1)
    dev = [d for _, kind, d in r.log if d is not None][0]
    # the check step's extrapolation lags the plain increment by 10 increment changes (warm step
    # 2 to check step 12)
    assert abs(dev - 10 * 2 * a * 100.0 ** 2) < 1e-9


def test_kink_above_tolerance_falls_back_to_the_plain_solution():
    # rate 0.01 m/yr until 1500 yr, then 0.05 m/yr: the period containing the kink misses its
    # check by (0.05 - 0.01) * 100 = 4 m > 1 m and is integrated by the plain solver
    def G(t):
        return np.full((6, 5), 0.01 * min(t, 1500.0) + 0.05 * max(t - 1500.0, 0.0))

    eng, r, res = run(G, 2200.0)
    assert res["fallbacks"] == 1
    kinds = [kind for _, kind, _ in r.log]
    assert "check-extrapolation-fallback" in kinds
    np.testing.assert_allclose(eng.u, G(2200.0), rtol=0, atol=1e-9)


class CheckpointEngine(ScriptedEngine):
    """The scripted engine with checkpoints and an RKC step that is exact up to a fixed bias."""

    def __init__(self, G, bias=0.0, shape=(6, 5)):
        super().__init__(G, shape)
        self.bias = bias
        self.n_restore = 0

    def checkpoint(self):
        return (self.u.copy(), self.t, self.n_integrate)

    def restore(self, blob):
        self.u, self.t = blob[0].copy(), blob[1]
        self.n_restore += 1

    def integrate_rkc(self, t1, dt_max):
        self.u = self.u + (self.G(t1) - self.G(self.t)) + self.bias
        self.t = float(t1)
        return {"n_rhs": 10}


def test_rkc_predictor_periods_and_exact_fallback():
    """RKC predictor: 9 predicted steps and one checked step per period; a check perturbed beyond
    the tolerance rolls the period back to its start and integrates it with the plain solver."""
    def G(t):
        return np.full((6, 5), 0.01 * t + 1e-6 * t * t)

    eng = CheckpointEngine(G, bias=1e-3)
    bumped = []

    def perturb(k, du):
        if k == 2:
            bumped.append(k)
            return du + 2.0
        return du

    r = SurrogateRunner(Sim(eng), None, Macro(100.0), check_every=10, tol=1.0, predictor="rkc", perturb=perturb)
    res = r.run(3000.0, t0=0.0)
    assert bumped == [2] and res["fallbacks"] == 1 and res["checks"] == 3
    assert res["surrogate_steps"] == 18 and res["plain_steps"] == 2 + 10
    # periods 1 and 3 carry 9 steps of RKC bias each; period 2 was integrated by the plain solver
    np.testing.assert_allclose(eng.u, G(3000.0) + 18e-3, rtol=0, atol=1e-9)
    kinds = [k for _, k, _ in r.log]
    assert kinds.count("check-rkc") == 2 and kinds.count("check-rkc-fallback") == 1
