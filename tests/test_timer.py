"""Timer history (src/simulation.jl:28-52): t_computation! appends, at the end of every
update_diagnostics! call whose time is past t_span[0], the call's time to t_vec and the elapsed
wall time to t_computation; run! leaves t_computation relative to its first entry
(src/simulation.jl:260).

Pins: the recorded times are exactly the right-hand-side times of the run, which an RHS-by-RHS
driver knows (fi_rhs at given times); the kept count obeys the capacity; t_computation starts at
0 and never decreases; an integrated run holds every accepted step's end time among its stage
times, and as many calls as the engine counted past t0. Oracle on the CPU, HIP engine on the GPU.
"""
import numpy as np
import pytest

from _cases import hip_sim, lv_case, sim_inputs


def oracle_sim(w, cap):
    from oracle import pyref
    inputs = sim_inputs(w).inputs
    inputs["timer_capacity"] = cap
    return pyref.RefEngine.from_inputs(inputs)


def check_rhs_times(e, cap):
    u = np.zeros((e.nx, e.ny))
    times = [0.0, 0.5, 3.0, 3.0, 7.25, 20.0]
    for t in times:
        e.rhs(t, u)
    tv, sec, n = e.timer_history(cap)
    past = [t for t in times if t > 0.0]
    assert n == min(len(past), cap - 1)           # call 0 (t = 0) is recorded, then dropped as t0
    np.testing.assert_array_equal(tv, past[:n])
    assert sec[0] == 0.0 and np.all(np.diff(sec) >= 0)


def check_run(e, w, cap):
    e.integrator_init()
    e.integrate(w["t_span"][1])
    tv, sec, n = e.timer_history(cap)
    n_rhs = e.stats()["n_rhs"]
    # the calls at t0 (initial derivative, FSAL start) are dropped; every later one is kept
    assert n_rhs - 3 <= n < n_rhs
    assert np.all(tv > 0.0) and np.all(tv <= w["t_span"][1])
    assert sec[0] == 0.0 and np.all(np.diff(sec) >= 0)
    assert w["t_span"][1] in set(tv)                # the last step's FSAL call at t_end
    return tv, n


def test_oracle_timer_history():
    w = lv_case(n=16, W=3e6, t_end=50.0)
    check_rhs_times(oracle_sim(w, 100), 100)
    check_rhs_times(oracle_sim(w, 4), 4)           # capacity 4: calls 0-3 recorded
    check_run(oracle_sim(w, 1 << 16), w, 1 << 16)


def test_timer_is_off_for_zero_capacity():
    w = lv_case(n=16, W=3e6, t_end=50.0)
    e = oracle_sim(w, 0)
    e.rhs(1.0, np.zeros((16, 16)))
    assert e.timer_history(10)[2] == 0


@pytest.mark.gpu
def test_gpu_timer_history():
    w = lv_case(n=32, W=3e6, t_end=50.0)
    sim = hip_sim(w)
    check_rhs_times(sim.engine, sim.timer.capacity)
    sim.close()
    sim = hip_sim(w)
    tv, n = check_run(sim.engine, w, sim.timer.capacity)
    np.testing.assert_array_equal(sim.timer.t_vec, tv)
    assert len(sim.timer.t_computation) == n == sim.timer.n_total
    sim.close()
