"""Ensembles of independent simulations on uneven grids, advanced together on one device.

The reference runs an ensemble (the inversion / ensemble-Kalman path) as one deep copy of the
Simulation per member integrated in a threaded loop
(ext/FastIsostasyEnsembleKalmanProcessesExt.jl:103-122). Here the members live in one fi_batch:
every fi_batch_integrate call advances all of them in one persistent launch, the device split
into teams of workgroups that take members from a queue, largest first.
"""
import ctypes as C

from . import _abi
from .engine import Engine, EngineError
from .simulation import push_inputs


class Ensemble:
    """Ensemble(inputs_list; team_size=0): inputs_list holds Simulation(..., engine=None).inputs
    per member. team_size 0 picks the team size by timing the first launches (results do not
    depend on it)."""

    def __init__(self, inputs_list, team_size=0, lib=None):
        self.lib = lib if lib is not None else _abi.load_engine()
        n = len(inputs_list)
        cfgs = (_abi.FiConfig * n)(*[inp["cfg"] for inp in inputs_list])
        h = C.c_void_p()
        st = self.lib.fi_create_batch(cfgs, n, C.byref(h))
        if st != 0:
            raise EngineError(st, "fi_create_batch failed")
        self.h = h
        self.members = []
        for m, inp in enumerate(inputs_list):
            e = Engine.member(self.lib, inp["cfg"], self.lib.fi_batch_member(self.h, m))
            push_inputs(e, inp, timer=False)
            self.members.append(e)
        self._check(self.lib.fi_batch_set_team_size(self.h, int(team_size)))

    def _check(self, st):
        if st != 0:
            msg = self.lib.fi_batch_last_error(self.h)
            raise EngineError(st, msg.decode() if msg else "")

    def __len__(self):
        return len(self.members)

    def __getitem__(self, m):
        return self.members[m]

    def integrator_init(self):
        for e in self.members:
            e.integrator_init()

    def autotune_fft(self, attempts=40):
        """fi_batch_autotune_fft: FFT radix orders tuned on one member per kind and grid."""
        self._check(self.lib.fi_batch_autotune_fft(self.h, int(attempts)))

    def integrate(self, t_target, fsal=True):
        """Advance every member to t_target; fsal re-evaluates each member's FSAL derivative
        there, as after an output callback."""
        self._check(self.lib.fi_batch_integrate(self.h, float(t_target), int(bool(fsal))))

    def integrate_each(self, t_targets, fsal=True):
        """Advance member m to t_targets[m]."""
        import numpy as np
        tt = np.ascontiguousarray(t_targets, dtype=np.float64)
        self._check(self.lib.fi_batch_integrate_each(self.h, tt.ctypes.data_as(C.POINTER(C.c_double)),
                                                     int(bool(fsal))))

    def set_team_size(self, team_size):
        """Workgroups per member team; 0 picks it by timing the next launches."""
        self._check(self.lib.fi_batch_set_team_size(self.h, int(team_size)))

    def info(self):
        s = _abi.FiBatchInfo()
        self._check(self.lib.fi_batch_get_info(self.h, C.byref(s)))
        return s.as_dict()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
        return False

    def close(self):
        if getattr(self, "h", None):
            for e in self.members:
                e.h = None
            self.lib.fi_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
