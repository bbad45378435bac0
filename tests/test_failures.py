"""Failure handling forced from the host, on both back-ends: a NaN in the state and a step size
driven below its minimum end the integration with an error code (no exception crosses the ABI,
no kernel faults). OrdinaryDiffEq stops on a NaN error norm with retcode DtNaN (its PI
controller proposes a NaN step), and on a step below its minimum with DtLessThanMin."""
import numpy as np
import pytest

import fastiso
from _cases import hip_sim, lv_case, sim_inputs


def oracle_engine(w):
    from oracle import pyref
    return pyref.RefEngine.from_inputs(sim_inputs(w).inputs)


def hip_engine(w):
    return hip_sim(w).engine


def nan_case(make):
    w = lv_case(n=32, t_end=100.0)
    e = make(w)
    e.integrator_init()
    e.integrate(2.0)
    u = e.get_field("u")
    u[7, 11] = np.nan
    e.set_field("u", u)
    return e


def tiny_step_case(make):
    """A target 4e-15 yr past the current time: the step clipped to it is below the minimum step
    1e-14 max(1, |t|) of both back-ends."""
    w = lv_case(n=32, t_end=100.0)
    e = make(w)
    e.integrator_init()
    e.integrate(2.0)
    return e


def test_oracle_nan_state_is_nonfinite():
    e = nan_case(oracle_engine)
    r0 = e.stats()["n_reject"]
    with pytest.raises(RuntimeError, match=r"error -4:"):
        e.integrate(4.0)
    assert e.stats()["n_reject"] == r0 + 1


def test_oracle_dt_below_minimum():
    e = tiny_step_case(oracle_engine)
    a0 = e.stats()["n_accept"]
    with pytest.raises(RuntimeError, match=r"error -7:"):
        e.integrate(2.0 + 4e-15)
    st = e.stats()
    assert st["n_accept"] == a0 and st["t"] == 2.0


@pytest.mark.gpu
def test_gpu_nan_state_is_nonfinite():
    e = nan_case(hip_engine)
    r0 = e.stats()["n_reject"]
    with pytest.raises(fastiso.engine.EngineError) as ex:
        e.integrate(4.0)
    assert ex.value.code == fastiso._abi.FI_ERR_NONFINITE
    assert e.stats()["n_reject"] == r0 + 1
    e.close()


@pytest.mark.gpu
def test_gpu_dt_below_minimum():
    e = tiny_step_case(hip_engine)
    a0 = e.stats()["n_accept"]
    with pytest.raises(fastiso.engine.EngineError) as ex:
        e.integrate(2.0 + 4e-15)
    assert ex.value.code == fastiso._abi.FI_ERR_DT_MIN
    st = e.stats()
    assert st["n_accept"] == a0 and st["t"] == 2.0
    e.close()


@pytest.mark.gpu
def test_gpu_nan_member_in_ensemble_is_nonfinite():
    """An ensemble member with a NaN state stops the batch with FI_ERR_NONFINITE naming it."""
    ws = [lv_case(n=32, seed=s, t_end=100.0) for s in (21, 31)]
    ens = fastiso.Ensemble([sim_inputs(w).inputs for w in ws], team_size=2)
    ens.integrator_init()
    ens.integrate(2.0, fsal=True)
    u = ens[1].get_field("u")
    u[3, 5] = np.nan
    ens[1].set_field("u", u)
    with pytest.raises(fastiso.engine.EngineError) as ex:
        ens.integrate(4.0, fsal=True)
    assert ex.value.code == fastiso._abi.FI_ERR_NONFINITE and "member 1" in str(ex.value)
    ens.close()
