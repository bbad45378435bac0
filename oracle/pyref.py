"""ORACLE -- test infrastructure only.

ctypes wrapper of oracle/lib/libfastiso_ref.so (the FP64 CPU restatement). Only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg import this module.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libfastiso_ref.so")

FIELDS = {
    "litho_rigidity": 1, "effective_viscosity": 2, "scaled_pseudodiff_inv": 3, "maskactive": 4,
    "green_elastic": 5, "green_geoid": 6, "ref_u": 7, "ref_ue": 8, "ref_dz_ss": 9, "ref_z_b": 10,
    "ref_H_ice": 11, "distortion": 12, "green_viscous": 13, "tau": 14, "litho_thickness": 15,
    "u": 32, "dudt": 33, "ue": 34, "dz_ss": 35, "z_ss": 36, "z_b": 37, "H_ice": 38, "H_af": 39,
    "H_water": 40, "maskgrounded": 41, "maskocean": 42, "canom_ice": 43, "canom_seawater": 44,
    "canom_litho": 45, "canom_mantle": 46, "canom_load": 47, "canom_full": 48, "u_eq": 49, "u_x": 50, "u_y": 51,
}


class RefStats(C.Structure):
    _fields_ = [
        ("t", C.c_double), ("dt", C.c_double),
        ("n_rhs", C.c_int64), ("n_accept", C.c_int64), ("n_reject", C.c_int64),
        ("count_sparse_updates", C.c_int64), ("n_output", C.c_int64),
        ("V_af", C.c_double), ("V_den", C.c_double), ("V_pov", C.c_double), ("delta_V", C.c_double),
        ("z_bsl", C.c_double), ("last_EEst", C.c_double),
        ("device_ms", C.c_double), ("n_launch", C.c_int64), ("n_accept_timed", C.c_int64),
        ("A_ocean", C.c_double),
    ]


class RefRkcInfo(C.Structure):
    _fields_ = [("rho", C.c_double), ("h", C.c_double), ("substeps", C.c_int32), ("stages", C.c_int32),
                ("n_rhs", C.c_int64), ("graph", C.c_int32), ("reserved", C.c_int32)]


class RefTuneInfo(C.Structure):
    _fields_ = [("workgroups", C.c_int32), ("n_wg_candidates", C.c_int32), ("wg_candidates", C.c_int32 * 8),
                ("wg_us", C.c_double * 8), ("plans", C.c_uint64 * 4), ("n_plan_candidates", C.c_int32 * 4),
                ("plan_default_us", C.c_double * 4), ("plan_best_us", C.c_double * 4)]


CB = C.CFUNCTYPE(None, C.c_void_p, C.c_double, C.c_int32)
_dp = C.POINTER(C.c_double)
_lib = None


def build():
    subprocess.check_call(["make", "-s", "-C", HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        P = C.c_void_p
        sig = {
            "fi_abi_version": (C.c_int32, []),
            "fi_create": (C.c_int, [P, C.POINTER(P)]),
            "fi_destroy": (None, [P]),
            "fi_last_error": (C.c_char_p, [P]),
            "fi_set_field": (C.c_int, [P, C.c_int32, _dp, C.c_size_t]),
            "fi_get_field": (C.c_int, [P, C.c_int32, _dp, C.c_size_t]),
            "fi_set_ice_snapshots": (C.c_int, [P, _dp, C.c_int32, _dp, C.c_int32]),
            "fi_set_sea_surface_snapshots": (C.c_int, [P, _dp, C.c_int32, _dp, C.c_int32]),
            "fi_set_timer_history": (C.c_int, [P, C.c_int64]),
            "fi_get_timer_history": (C.c_int, [P, _dp, _dp, C.c_int64, C.POINTER(C.c_int64)]),
            "fi_set_ice_separable": (C.c_int, [P, _dp, _dp, C.c_int32, _dp, C.c_int32]),
            "fi_set_bsl": (C.c_int, [P, C.c_int32, C.c_int32, C.c_double, _dp, _dp, C.c_int32, C.c_int32]),
            "fi_set_output_times": (C.c_int, [P, _dp, C.c_int32]),
            "fi_init": (C.c_int, [P, C.c_double, C.c_double]),
            "fi_rhs": (C.c_int, [P, C.c_double, _dp, _dp]),
            "fi_sparse_update": (C.c_int, [P, C.c_double]),
            "fi_integrator_init": (C.c_int, [P]),
            "fi_integrate": (C.c_int, [P, C.c_double, CB, C.c_void_p]),
            "fi_get_stats": (C.c_int, [P, C.POINTER(RefStats)]),
            "fi_integrate_rkc": (C.c_int, [P, C.c_double, C.c_double, C.c_double, C.POINTER(RefRkcInfo)]),
            "fi_autotune": (C.c_int, [P, C.c_int32, C.c_int32, C.POINTER(RefTuneInfo)]),
            "fi_set_fft_plans": (C.c_int, [P, C.POINTER(C.c_uint64)]),
            "fi_get_fft_plans": (C.c_int, [P, C.POINTER(C.c_uint64)]),
            "fi_set_workgroups": (C.c_int, [P, C.c_int32]),
            "fi_batch_autotune_fft": (C.c_int, [P, C.c_int32]),
            "fi_set_time": (C.c_int, [P, C.c_double]),
            "fi_checkpoint_size": (C.c_int, [P, C.POINTER(C.c_size_t)]),
            "fi_checkpoint_save": (C.c_int, [P, C.c_void_p, C.c_size_t]),
            "fi_checkpoint_load": (C.c_int, [P, C.c_void_p, C.c_size_t]),
            "fi_create_batch": (C.c_int, [P, C.c_int32, C.POINTER(P)]),
            "fi_batch_member": (P, [P, C.c_int32]),
            "fi_batch_destroy": (None, [P]),
            "fi_batch_last_error": (C.c_char_p, [P]),
            "fi_batch_integrate": (C.c_int, [P, C.c_double, C.c_int32]),
            "fi_batch_integrate_each": (C.c_int, [P, _dp, C.c_int32]),
            "ref_fft": (None, [C.c_int, _dp, _dp, _dp, _dp, C.c_int]),
            "ref_derivatives": (None, [C.c_int, C.c_int, C.c_double, C.c_double, _dp, _dp, _dp, _dp, _dp]),
            "ref_derivatives_k": (None, [C.c_int, C.c_int, C.c_double, C.c_double, _dp, _dp, _dp, _dp, _dp, _dp]),
            "ref_conv_cropped": (None, [C.c_int, C.c_int, _dp, _dp, _dp]),
            "ref_samesize_conv": (None, [C.c_int, C.c_int, _dp, _dp, C.c_double, _dp]),
            "ref_nextfastfft": (C.c_int, [C.c_int]),
            "ref_gauss_legendre": (None, [C.c_int, _dp, _dp]),
            "ref_pseudodiff": (None, [C.c_int, C.c_int, C.c_double, C.c_double, _dp]),
            "ref_elastic_green": (None, [C.c_int, C.c_int, C.c_double, C.c_double, _dp, _dp, C.c_int, C.c_int, _dp]),
            "ref_geoid_green": (None, [C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double, _dp]),
            "ref_set_threads": (None, [C.c_int]),
            "ref_max_threads": (C.c_int, []),
            "ref_solidearth": (None, [C.c_int, C.c_int, C.c_double, C.c_double, C.c_int, _dp, _dp, _dp,
                                      C.c_double, C.c_double, C.c_double, C.c_int, _dp, _dp, _dp, _dp]),
            "ref_analytic_solution": (C.c_double, [C.c_double] * 9 + [C.c_int]),
            "ref_bs3_linear": (C.c_double, [C.c_double] * 6 + [C.POINTER(C.c_int64)]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _flat(X):
    return np.ascontiguousarray(np.asarray(X, dtype=np.float64).ravel(order="F"))


def _p(a):
    return a.ctypes.data_as(_dp)


def _shape(v, nx, ny):
    return v.reshape((nx, ny), order="F")


class RefEngine:
    """One context of the CPU oracle, fed with the same inputs as the HIP engine."""

    def __init__(self, cfg):
        self.L = lib()
        self.nx, self.ny = cfg.nx, cfg.ny
        self.N = cfg.nx * cfg.ny
        h = C.c_void_p()
        st = self.L.fi_create(C.byref(cfg), C.byref(h))
        if st != 0:
            raise RuntimeError(f"oracle fi_create failed: {st}")
        self.h = h
        self._cb = None

    @classmethod
    def from_inputs(cls, inputs, handle=None):
        e = cls(inputs["cfg"]) if handle is None else cls._member(inputs["cfg"], handle)
        for name, X in inputs["fields"].items():
            e.set_field(name, X)
        ice = inputs["ice"]
        if ice[0] == "snapshots":
            t = np.asarray(ice[1], dtype=np.float64)
            H = np.concatenate([_flat(x) for x in ice[2]])
            e._ck(e.L.fi_set_ice_snapshots(e.h, _p(t), len(t), _p(H), int(ice[3])))
        elif ice[0] == "separable":
            t = np.asarray(ice[1], dtype=np.float64)
            a = np.asarray(ice[2], dtype=np.float64)
            H = _flat(ice[3])
            e._ck(e.L.fi_set_ice_separable(e.h, _p(t), _p(a), len(t), _p(H), int(ice[4])))
        if inputs.get("sea_surface") is not None:
            t = np.asarray(inputs["sea_surface"][0], dtype=np.float64)
            X = np.concatenate([_flat(x) for x in inputs["sea_surface"][1]])
            e._ck(e.L.fi_set_sea_surface_snapshots(e.h, _p(t), len(t), _p(X), int(inputs["sea_surface"][2])))
        if handle is None and inputs.get("timer_capacity"):
            e._ck(e.L.fi_set_timer_history(e.h, int(inputs["timer_capacity"])))
        if "bsl" in inputs:
            e.set_bsl(*inputs["bsl"])
        tout = np.asarray(inputs["tout"], dtype=np.float64)
        e._ck(e.L.fi_set_output_times(e.h, _p(tout), len(tout)))
        e._ck(e.L.fi_init(e.h, *inputs["t_span"]))
        return e

    @classmethod
    def _member(cls, cfg, handle):
        e = cls.__new__(cls)
        e.L = lib()
        e.nx, e.ny = cfg.nx, cfg.ny
        e.N = cfg.nx * cfg.ny
        e.h = C.c_void_p(handle)
        e._cb = None
        e._owned = False
        return e

    def __del__(self):
        if getattr(self, "h", None):
            if getattr(self, "_owned", True):
                self.L.fi_destroy(self.h)
            self.h = None

    def close(self):
        self.__del__()

    def timer_history(self, capacity):
        t = np.empty(max(int(capacity), 1))
        sec = np.empty_like(t)
        n = C.c_int64(0)
        self._ck(self.L.fi_get_timer_history(self.h, _p(t), _p(sec), int(capacity), C.byref(n)))
        m = min(n.value, int(capacity))
        return t[:m].copy(), sec[:m].copy(), n.value

    def _ck(self, st):
        if st != 0:
            raise RuntimeError(f"oracle error {st}: {self.L.fi_last_error(self.h).decode()}")

    def set_bsl(self, rule, update, A_ref, x=None, y=None, flat_bc=False):
        xs = np.asarray(x if x is not None else [], dtype=np.float64)
        ys = np.asarray(y if y is not None else [], dtype=np.float64)
        self._ck(self.L.fi_set_bsl(self.h, int(rule), int(update), float(A_ref), _p(xs) if len(xs) else None,
                                   _p(ys) if len(ys) else None, len(xs), int(flat_bc)))

    def set_field(self, name, X):
        v = _flat(X)
        self._ck(self.L.fi_set_field(self.h, FIELDS[name], _p(v), self.N))

    def get_field(self, name):
        v = np.empty(self.N)
        self._ck(self.L.fi_get_field(self.h, FIELDS[name], _p(v), self.N))
        return _shape(v, self.nx, self.ny)

    def rhs(self, t, u):
        uu = _flat(u).copy()
        du = np.empty(self.N)
        self._ck(self.L.fi_rhs(self.h, float(t), _p(uu), _p(du)))
        return _shape(du, self.nx, self.ny), _shape(uu, self.nx, self.ny)

    def sparse_update(self, t):
        self._ck(self.L.fi_sparse_update(self.h, float(t)))

    def integrator_init(self):
        self._ck(self.L.fi_integrator_init(self.h))

    def integrate(self, t_target, callback=None):
        cb = CB() if callback is None else CB(lambda user, t, k: callback(t, k))
        self._cb = cb
        self._ck(self.L.fi_integrate(self.h, float(t_target), cb, None))

    def set_time(self, t):
        self._ck(self.L.fi_set_time(self.h, float(t)))

    def integrate_rkc(self, t_target, dt_max, rho=0.0):
        info = RefRkcInfo()
        self._ck(self.L.fi_integrate_rkc(self.h, float(t_target), float(dt_max), float(rho), C.byref(info)))
        return {k: getattr(info, k) for k, _ in info._fields_}

    def autotune(self, what=3, attempts=40):
        info = RefTuneInfo()
        self._ck(self.L.fi_autotune(self.h, int(what), int(attempts), C.byref(info)))
        return {"workgroups": info.workgroups, "plans": list(info.plans), "wg_candidates": [], "wg_us": []}

    def fft_plans(self):
        a = (C.c_uint64 * 4)()
        self._ck(self.L.fi_get_fft_plans(self.h, a))
        return list(a)

    def set_fft_plans(self, plans):
        self._ck(self.L.fi_set_fft_plans(self.h, (C.c_uint64 * 4)(*[int(x) for x in plans])))

    def set_workgroups(self, g):
        self._ck(self.L.fi_set_workgroups(self.h, int(g)))

    def checkpoint(self):
        n = C.c_size_t()
        self._ck(self.L.fi_checkpoint_size(self.h, C.byref(n)))
        buf = C.create_string_buffer(n.value)
        self._ck(self.L.fi_checkpoint_save(self.h, buf, n.value))
        return buf.raw

    def restore(self, blob):
        buf = C.create_string_buffer(blob, len(blob))
        self._ck(self.L.fi_checkpoint_load(self.h, buf, len(blob)))

    def stats(self):
        s = RefStats()
        self._ck(self.L.fi_get_stats(self.h, C.byref(s)))
        return {k: getattr(s, k) for k, _ in s._fields_}


class RefEnsemble:
    """fi_create_batch of the oracle: its members integrated one after the other."""

    def __init__(self, inputs_list):
        self.L = lib()
        n = len(inputs_list)
        cfg_t = type(inputs_list[0]["cfg"])
        cfgs = (cfg_t * n)(*[inp["cfg"] for inp in inputs_list])
        h = C.c_void_p()
        st = self.L.fi_create_batch(C.cast(cfgs, C.c_void_p), n, C.byref(h))
        if st != 0:
            raise RuntimeError(f"oracle fi_create_batch failed: {st}")
        self.h = h
        self.members = [RefEngine.from_inputs(inp, handle=self.L.fi_batch_member(h, m))
                        for m, inp in enumerate(inputs_list)]

    def __getitem__(self, m):
        return self.members[m]

    def __len__(self):
        return len(self.members)

    def integrator_init(self):
        for e in self.members:
            e.integrator_init()

    def integrate(self, t_target, fsal=True):
        st = self.L.fi_batch_integrate(self.h, float(t_target), int(bool(fsal)))
        if st != 0:
            raise RuntimeError(f"oracle error {st}: {self.L.fi_batch_last_error(self.h).decode()}")

    def info(self):
        return {"n_members": len(self.members)}

    def close(self):
        self.__del__()

    def integrate_each(self, t_targets, fsal=True):
        tt = np.ascontiguousarray(t_targets, dtype=np.float64)
        st = self.L.fi_batch_integrate_each(self.h, _p(tt), int(bool(fsal)))
        if st != 0:
            raise RuntimeError(f"oracle error {st}: {self.L.fi_batch_last_error(self.h).decode()}")

    def __del__(self):
        if getattr(self, "h", None):
            for e in self.members:
                e.h = None
            self.L.fi_batch_destroy(self.h)
            self.h = None


# ---- ref_* hooks ---------------------------------------------------------------------------
def fft(x, sign=-1):
    x = np.asarray(x, dtype=np.complex128)
    n = len(x)
    re, im = np.ascontiguousarray(x.real), np.ascontiguousarray(x.imag)
    ore, oim = np.empty(n), np.empty(n)
    lib().ref_fft(n, _p(re), _p(im), _p(ore), _p(oim), sign)
    return ore + 1j * oim


def derivatives(u, dx, dy):
    nx, ny = u.shape
    uf = _flat(u)
    outs = [np.empty(nx * ny) for _ in range(4)]
    lib().ref_derivatives(nx, ny, dx, dy, _p(uf), *[_p(o) for o in outs])
    return [_shape(o, nx, ny) for o in outs]


def derivatives_k(u, dx, dy, K):
    """update_second_derivatives! with per-cell steps K dx, K dy: (uxx, uyy, ux, uxy)."""
    nx, ny = u.shape
    uf, kf = _flat(u), _flat(K)
    outs = [np.empty(nx * ny) for _ in range(4)]
    lib().ref_derivatives_k(nx, ny, dx, dy, _p(kf), _p(uf), *[_p(o) for o in outs])
    return [_shape(o, nx, ny) for o in outs]


def conv_cropped(kernel, inp):
    nx, ny = kernel.shape
    out = np.empty((2 * nx - 1) * (2 * ny - 1))
    lib().ref_conv_cropped(nx, ny, _p(_flat(kernel)), _p(_flat(inp)), _p(out))
    return _shape(out, 2 * nx - 1, 2 * ny - 1)


def samesize_conv(kernel, inp, x_border=0.0):
    nx, ny = kernel.shape
    out = np.empty(nx * ny)
    lib().ref_samesize_conv(nx, ny, _p(_flat(kernel)), _p(_flat(inp)), x_border, _p(out))
    return _shape(out, nx, ny)


def gauss_legendre(n):
    x, w = np.empty(n), np.empty(n)
    lib().ref_gauss_legendre(n, _p(x), _p(w))
    return x, w


def pseudodiff(nx, ny, Wx, Wy):
    out = np.empty(nx * ny)
    lib().ref_pseudodiff(nx, ny, Wx, Wy, _p(out))
    return _shape(out, nx, ny)


def set_threads(n):
    """OpenMP threads of the oracle's loops from now on."""
    lib().ref_set_threads(int(n))


def max_threads():
    return int(lib().ref_max_threads())


def elastic_green(nx, ny, Wx, Wy, rm, ge, nq=4):
    out = np.empty(nx * ny)
    rm = np.ascontiguousarray(rm, dtype=np.float64)
    ge = np.ascontiguousarray(ge, dtype=np.float64)
    lib().ref_elastic_green(nx, ny, Wx, Wy, _p(rm), _p(ge), len(rm), nq, _p(out))
    return _shape(out, nx, ny)


def geoid_green(nx, ny, Wx, Wy, r_pole=6357e3, mE=5.972e24):
    out = np.empty(nx * ny)
    lib().ref_geoid_green(nx, ny, Wx, Wy, r_pole, mE, _p(out))
    return _shape(out, nx, ny)


def solidearth(nx, ny, Wx, Wy, lb, lv, maskactive=None, E=6.6e10, nu=0.28, nu_m=0.28, compressible=True):
    nl = len(lb)
    N = nx * ny
    lbf = np.concatenate([_flat(np.broadcast_to(np.asarray(x, float), (nx, ny))) for x in lb])
    lvf = np.concatenate([_flat(np.broadcast_to(np.asarray(x, float), (nx, ny))) for x in lv])
    m = _flat(np.ones((nx, ny)) if maskactive is None else maskactive)
    outs = [np.empty(N) for _ in range(4)]
    lib().ref_solidearth(nx, ny, Wx, Wy, nl, _p(lbf), _p(lvf), _p(m), E, nu, nu_m, int(compressible),
                         *[_p(o) for o in outs])
    return [_shape(o, nx, ny) for o in outs]


def analytic_solution(r, t_sec, rho_ice, g, H0, R0, rho_um, D, eta, nq=5):
    return lib().ref_analytic_solution(r, t_sec, rho_ice, g, H0, R0, rho_um, D, eta, nq)


def bs3_linear(lam, u0, t_end, reltol=1e-5, abstol=1e-6, dt_fixed=0.0):
    n = C.c_int64()
    u = lib().ref_bs3_linear(lam, u0, t_end, reltol, abstol, dt_fixed, C.byref(n))
    return u, n.value
