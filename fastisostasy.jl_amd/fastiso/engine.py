"""Thin object wrapper over one fi_ctx of the HIP engine."""
import ctypes as C

import numpy as np

from . import _abi


class EngineError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"fastiso engine error {code}: {msg}")
        self.code = code


def _f64(a, n=None):
    a = np.ascontiguousarray(np.asarray(a, dtype=np.float64).reshape(-1, order="F") if np.ndim(a) == 2
                             else np.asarray(a, dtype=np.float64).ravel())
    if n is not None and a.size != n:
        raise ValueError(f"expected {n} values, got {a.size}")
    return a


def _ptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def fortran_flat(X):
    """(nx, ny) array -> column-major flat FP64 vector (element (i, j) at i + j*nx)."""
    return np.ascontiguousarray(np.asarray(X, dtype=np.float64).ravel(order="F"))


def from_flat(v, nx, ny):
    return np.asarray(v).reshape((nx, ny), order="F")


class Engine:
    """One simulation context on the HIP engine (libfastiso_hip.so)."""

    def __init__(self, cfg, lib=None):
        self.lib = lib if lib is not None else _abi.load_engine()
        self.cfg = cfg
        self.nx, self.ny = cfg.nx, cfg.ny
        self.N = cfg.nx * cfg.ny
        h = C.c_void_p()
        st = self.lib.fi_create(C.byref(cfg), C.byref(h))
        if st != 0:
            raise EngineError(st, "fi_create failed")
        self.h = h
        self._cb = None

    @classmethod
    def member(cls, lib, cfg, handle):
        """Wrap a context owned by an ensemble (fi_batch_member); closing it is the ensemble's job."""
        e = cls.__new__(cls)
        e.lib, e.cfg = lib, cfg
        e.nx, e.ny = cfg.nx, cfg.ny
        e.N = cfg.nx * cfg.ny
        e.h = C.c_void_p(handle)
        e._cb = None
        e._owned = False
        return e

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
        return False

    def close(self):
        if getattr(self, "h", None):
            if getattr(self, "_owned", True):
                self.lib.fi_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, st):
        if st != 0:
            msg = self.lib.fi_last_error(self.h)
            raise EngineError(st, msg.decode() if msg else "")

    def set_field(self, name, X):
        v = fortran_flat(X) if np.ndim(X) == 2 else _f64(X, self.N)
        self._check(self.lib.fi_set_field(self.h, _abi.FIELDS[name], _ptr(v), self.N))

    def get_field(self, name):
        v = np.empty(self.N, dtype=np.float64)
        self._check(self.lib.fi_get_field(self.h, _abi.FIELDS[name], _ptr(v), self.N))
        return from_flat(v, self.nx, self.ny)

    def set_ice_snapshots(self, t, H_list, flat_bc=False):
        t = _f64(t)
        H = np.concatenate([fortran_flat(H) for H in H_list])
        self._check(self.lib.fi_set_ice_snapshots(self.h, _ptr(t), len(t), _ptr(H), int(flat_bc)))

    def set_sea_surface_snapshots(self, t, dz_list, flat_bc=False):
        """fi_set_sea_surface_snapshots: the dz_ss snapshots of an ImposedSeaSurface."""
        t = _f64(t)
        X = np.concatenate([fortran_flat(x) for x in dz_list])
        self._check(self.lib.fi_set_sea_surface_snapshots(self.h, _ptr(t), len(t), _ptr(X), int(flat_bc)))

    def set_timer_history(self, capacity):
        """fi_set_timer_history: record (t, device wall clock) of the first `capacity` RHS calls."""
        self._check(self.lib.fi_set_timer_history(self.h, int(capacity)))

    def timer_history(self, capacity):
        """fi_get_timer_history: (t_vec, t_computation, n_total) as Timer holds them."""
        t = np.empty(max(int(capacity), 1))
        sec = np.empty_like(t)
        n = C.c_int64(0)
        self._check(self.lib.fi_get_timer_history(self.h, _ptr(t), _ptr(sec), int(capacity), C.byref(n)))
        m = min(n.value, int(capacity))
        return t[:m].copy(), sec[:m].copy(), n.value

    def set_ice_separable(self, t, a, H_shape, flat_bc=False):
        t, a, H = _f64(t), _f64(a), fortran_flat(H_shape)
        self._check(self.lib.fi_set_ice_separable(self.h, _ptr(t), _ptr(a), len(t), _ptr(H), int(flat_bc)))

    def set_bsl(self, rule, update, A_ref, x=None, y=None, flat_bc=False):
        """fi_set_bsl: the BSL rule (0 constant, 1 constant ocean surface, 2 piecewise constant,
        3 imposed), the update mechanism (0 internal, 1 external) and the rule's table."""
        xs = _f64(x if x is not None else [])
        ys = _f64(y if y is not None else [])
        self._check(self.lib.fi_set_bsl(self.h, int(rule), int(update), float(A_ref),
                                        _ptr(xs) if len(xs) else None, _ptr(ys) if len(ys) else None,
                                        len(xs), int(flat_bc)))

    def set_output_times(self, t):
        t = _f64(t)
        self._check(self.lib.fi_set_output_times(self.h, _ptr(t), len(t)))

    def init(self, t0, t_end):
        self._check(self.lib.fi_init(self.h, float(t0), float(t_end)))

    def rhs(self, t, u):
        """update_diagnostics! on host arrays; returns (dudt, u_after_bc)."""
        uu = fortran_flat(u).copy()
        du = np.empty(self.N, dtype=np.float64)
        self._check(self.lib.fi_rhs(self.h, float(t), _ptr(uu), _ptr(du)))
        return from_flat(du, self.nx, self.ny), from_flat(uu, self.nx, self.ny)

    def sparse_update(self, t):
        self._check(self.lib.fi_sparse_update(self.h, float(t)))

    def integrator_init(self):
        self._check(self.lib.fi_integrator_init(self.h))

    def integrate(self, t_target, callback=None):
        if callback is None:
            cb = _abi.OUTPUT_CB()
        else:
            cb = _abi.OUTPUT_CB(lambda user, t, k: callback(t, k))
        self._cb = cb
        self._check(self.lib.fi_integrate(self.h, float(t_target), cb, None))

    def checkpoint(self):
        """The dynamic state as bytes (fi_checkpoint_save)."""
        n = C.c_size_t()
        self._check(self.lib.fi_checkpoint_size(self.h, C.byref(n)))
        buf = C.create_string_buffer(n.value)
        self._check(self.lib.fi_checkpoint_save(self.h, buf, n.value))
        return buf.raw

    def restore(self, blob):
        """Restore a checkpoint() of this library and configuration (fi_checkpoint_load)."""
        buf = C.create_string_buffer(blob, len(blob))
        self._check(self.lib.fi_checkpoint_load(self.h, buf, len(blob)))

    def integrate_rkc(self, t_target, dt_max, rho=0.0):
        """fi_integrate_rkc: RKC2 sub-steps of at most dt_max years to t_target (rho > 0: the
        spectral radius to size the stages with, else the context's estimate). Returns the
        fi_rkc_info fields."""
        info = _abi.FiRkcInfo()
        self._check(self.lib.fi_integrate_rkc(self.h, float(t_target), float(dt_max), float(rho), C.byref(info)))
        return info.as_dict()

    def autotune(self, what=_abi.FI_TUNE_WORKGROUPS | _abi.FI_TUNE_FFT_PLANS, attempts=40):
        """fi_autotune: time the persistent grid size and the FFT radix orders on this run (the
        state is restored) and keep the fastest; returns the fi_tune_info fields."""
        info = _abi.FiTuneInfo()
        self._check(self.lib.fi_autotune(self.h, int(what), int(attempts), C.byref(info)))
        return info.as_dict()

    def fft_plans(self):
        a = (C.c_uint64 * 4)()
        self._check(self.lib.fi_get_fft_plans(self.h, a))
        return list(a)

    def set_fft_plans(self, plans):
        a = (C.c_uint64 * 4)(*[int(x) for x in plans])
        self._check(self.lib.fi_set_fft_plans(self.h, a))

    def set_workgroups(self, g):
        self._check(self.lib.fi_set_workgroups(self.h, int(g)))

    def set_time(self, t):
        """Move the integrator time to t without integrating (after set_field("u", ...))."""
        self._check(self.lib.fi_set_time(self.h, float(t)))

    def stats(self):
        s = _abi.FiStats()
        self._check(self.lib.fi_get_stats(self.h, C.byref(s)))
        return s.as_dict()
