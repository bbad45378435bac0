"""fi_batch_integrate_each_noalloc (the ensemble launch without device allocation, its per-launch
buffers in one static device array) against fi_batch_integrate_each, through the C ABI: members
bitwise equal, on two calls in a row, the second on a smaller ensemble."""
import ctypes as C

import numpy as np
import pytest

import fastiso
from fastiso import _abi
from _cases import lv_case, sim_inputs


def member_inputs(sides, seed0):
    return [sim_inputs(lv_case(n=n, seed=seed0 + 10 * k, t_end=200.0)).inputs for k, n in enumerate(sides)]


def integrate_each(ens, targets, noalloc):
    tt = np.ascontiguousarray(targets, dtype=np.float64)
    f = ens.lib.fi_batch_integrate_each_noalloc if noalloc else ens.lib.fi_batch_integrate_each
    ens._check(f(ens.h, tt.ctypes.data_as(C.POINTER(C.c_double)), 1))


def run_pair(sides, seed0, targets_list, team_size_ref):
    inputs = member_inputs(sides, seed0)
    ref = fastiso.Ensemble(inputs, team_size=team_size_ref)
    new = fastiso.Ensemble(inputs, team_size=0)
    try:
        for e in (ref, new):
            e.integrator_init()
        for targets in targets_list:
            integrate_each(ref, targets, noalloc=False)
            integrate_each(new, targets, noalloc=True)
            for m in range(len(inputs)):
                for f in ("u", "ue", "dz_ss", "dudt"):
                    np.testing.assert_array_equal(new[m].get_field(f), ref[m].get_field(f),
                                                  err_msg=f"member {m} field {f}")
                sr, sn = ref[m].stats(), new[m].stats()
                assert (sn["t"], sn["n_accept"], sn["n_reject"], sn["n_rhs"]) == \
                       (sr["t"], sr["n_accept"], sr["n_reject"], sr["n_rhs"])
        return new.info()
    finally:
        ref.close()
        new.close()


@pytest.mark.gpu
def test_gpu_noalloc_batch_equals_allocating_entry_point_twice():
    print("noalloc call 1: 6 members, sides 64/48/32, uneven targets", flush=True)
    info = run_pair([64, 48, 32, 64, 48, 32], 101, [[20.0, 15.0, 20.0, 12.0, 20.0, 20.0],
                                                      [40.0, 40.0, 33.0, 40.0, 40.0, 40.0]], 4)
    assert info["n_launch"] >= 2
    print("noalloc call 2: 3 members, sides 40/32/24", flush=True)
    run_pair([40, 32, 24], 211, [[10.0, 10.0, 10.0]], 2)


@pytest.mark.gpu
def test_gpu_noalloc_refuses_more_members_than_its_scratch():
    """A batch larger than the static array's member maximum is refused with FI_ERR_INVALID
    before anything is launched (1025 members of 8 x 8)."""
    n = 1025
    cfgs = (_abi.FiConfig * n)()
    lib = _abi.load_engine()
    for m in range(n):
        lib.fi_default_config(C.byref(cfgs[m]), 8, 8, 3e6, 3e6)
    h = C.c_void_p()
    assert lib.fi_create_batch(cfgs, n, C.byref(h)) == 0
    try:
        tt = np.zeros(n)
        # members not initialised: the size check comes first for the static variant
        st = lib.fi_batch_integrate_each_noalloc(h, tt.ctypes.data_as(C.POINTER(C.c_double)), 0)
        assert st == _abi.FI_ERR_INVALID
        assert b"static scratch maximum" in lib.fi_batch_last_error(h)
    finally:
        lib.fi_batch_destroy(h)
