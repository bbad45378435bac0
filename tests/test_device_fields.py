"""fi_set_field_dev / fi_get_field_dev: fields on device-resident FP64 arrays (the reference's
array-type seam, src/convenience_types.jl:2-29). On the GPU the displacement moves device to
device and a run continued from a device-written state equals one continued from the same state
written through the host, bit for bit; on the oracle the "device" arrays are host arrays."""
import ctypes as C

import numpy as np
import pytest

from fastiso import _abi
from _cases import hip_sim, lv_case, sim_inputs


def test_oracle_device_field_calls_take_host_arrays(ref):
    e = ref.RefEngine.from_inputs(sim_inputs(lv_case(n=16, t_end=100.0)).inputs)
    e.integrator_init()
    e.integrate(2.0)
    L = ref.lib()
    L.fi_set_field_dev.argtypes = [C.c_void_p, C.c_int32, C.c_void_p, C.c_size_t]
    L.fi_get_field_dev.argtypes = [C.c_void_p, C.c_int32, C.c_void_p, C.c_size_t]
    u = np.asfortranarray(e.get_field("u") + 0.25)
    assert L.fi_set_field_dev(e.h, ref.FIELDS["u"], u.ctypes.data, u.size) == 0
    out = np.empty_like(u)
    assert L.fi_get_field_dev(e.h, ref.FIELDS["u"], out.ctypes.data, out.size) == 0
    np.testing.assert_array_equal(out, u)


class Hip:
    """hipMalloc / hipMemcpy / hipFree of the HIP runtime, for device arrays in tests."""

    def __init__(self):
        self.L = C.CDLL("libamdhip64.so")
        self.L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.L.hipFree.argtypes = [C.c_void_p]

    def alloc(self, nbytes):
        p = C.c_void_p()
        assert self.L.hipMalloc(C.byref(p), nbytes) == 0
        return p

    def h2d(self, d, a):
        assert self.L.hipMemcpy(d, a.ctypes.data, a.nbytes, 1) == 0

    def d2h(self, a, d):
        assert self.L.hipMemcpy(a.ctypes.data, d, a.nbytes, 2) == 0


@pytest.mark.gpu
def test_gpu_displacement_through_device_arrays():
    hip = Hip()
    w = lv_case(n=32, t_end=100.0)
    a, b = hip_sim(w).engine, hip_sim(w).engine
    for e in (a, b):
        e.integrator_init()
        e.integrate(2.0)
    lib, N = a.lib, a.N
    d = hip.alloc(N * 8)
    try:
        host = np.empty(N)
        assert lib.fi_get_field_dev(a.h, _abi.FIELDS["u"], d, N) == 0
        hip.d2h(host, d)
        np.testing.assert_array_equal(host.reshape((a.nx, a.ny), order="F"), a.get_field("u"))
        u2 = a.get_field("u") * 1.01 + 0.5
        hip.h2d(d, np.ascontiguousarray(u2.ravel(order="F")))
        assert lib.fi_set_field_dev(a.h, _abi.FIELDS["u"], d, N) == 0
        b.set_field("u", u2)
        np.testing.assert_array_equal(a.get_field("u"), u2)
        for e in (a, b):
            e.integrate(4.0)
        np.testing.assert_array_equal(a.get_field("u"), b.get_field("u"))
        # a static input set from the device before fi_init is staged through the host
        cfg = _abi.FiConfig()
        lib.fi_default_config(C.byref(cfg), 16, 16, 3e6, 3e6)
        h = C.c_void_p()
        assert lib.fi_create(C.byref(cfg), C.byref(h)) == 0
        D = np.linspace(1e22, 2e22, 256)
        d2 = hip.alloc(256 * 8)
        hip.h2d(d2, D)
        assert lib.fi_set_field_dev(h, _abi.FIELDS["litho_rigidity"], d2, 256) == 0
        back = np.empty(256)
        assert lib.fi_get_field(h, _abi.FIELDS["litho_rigidity"], back.ctypes.data_as(C.POINTER(C.c_double)), 256) == 0
        np.testing.assert_array_equal(back, D)
        lib.fi_destroy(h)
        hip.L.hipFree(d2)
    finally:
        hip.L.hipFree(d)
        a.close()
        b.close()
