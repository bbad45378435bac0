"""The learned surrogate (include/fastiso_surrogate.h) and its use with the plain solver.

Surrogate wraps the device network: forward, loss gradient, training on a device-resident set.
The network maps [u, load anomaly, log10 viscosity, lithosphere thickness] (scaled) to the viscous
displacement increment over one macro step (scaled). It has no counterpart in the reference.
"""
import ctypes as C

import numpy as np

from . import _abi
from .engine import EngineError

_fp = C.POINTER(C.c_float)


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _p(a):
    return a.ctypes.data_as(_fp)


def cells(F):
    """(nx, ny) field -> flat cell order i + j nx."""
    return np.asarray(F).ravel(order="F")


def stack_inputs(fields):
    """list of (nx, ny) fields -> (nx ny, channels) cell-major input of the network."""
    return _f32(np.stack([cells(f) for f in fields], axis=1))


class Surrogate:
    def __init__(self, width=32, modes_x=16, modes_y=16, layers=4, proj_hidden=64, in_channels=4, seed=12345,
                 device=0, lib=None):
        self.lib = lib if lib is not None else _abi.load_engine()
        cfg = _abi.FnoConfig()
        self.lib.fi_fno_default_config(C.byref(cfg))
        cfg.width, cfg.modes_x, cfg.modes_y, cfg.layers = width, modes_x, modes_y, layers
        cfg.proj_hidden, cfg.in_channels, cfg.seed, cfg.device = proj_hidden, in_channels, seed, device
        self.cfg = cfg
        h = C.c_void_p()
        st = self.lib.fi_surrogate_create(C.byref(cfg), C.byref(h))
        if st != 0:
            raise EngineError(st, "fi_surrogate_create failed")
        self.h = h
        self.n_params = int(self.lib.fi_surrogate_num_params(self.h))
        self.grid = None
        self.scales = None

    def config_dict(self):
        return {k: getattr(self.cfg, k) for k in ("in_channels", "width", "modes_x", "modes_y", "layers",
                                                  "proj_hidden")}

    def _check(self, st):
        if st != 0:
            msg = self.lib.fi_surrogate_last_error(self.h)
            raise EngineError(st, msg.decode() if msg else "")

    def close(self):
        if getattr(self, "h", None):
            self.lib.fi_surrogate_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_grid(self, nx, ny):
        self._check(self.lib.fi_surrogate_set_grid(self.h, int(nx), int(ny)))
        self.grid = (nx, ny)

    def get_params(self):
        out = np.empty(self.n_params, dtype=np.float32)
        self._check(self.lib.fi_surrogate_get_params(self.h, _p(out), self.n_params))
        return out

    def set_params(self, theta):
        theta = _f32(theta)
        self._check(self.lib.fi_surrogate_set_params(self.h, _p(theta), self.n_params))

    def forward(self, x):
        """x: (cells, channels) -> y (cells,)"""
        x = _f32(x)
        y = np.empty(x.shape[0], dtype=np.float32)
        self._check(self.lib.fi_surrogate_forward(self.h, _p(x), _p(y)))
        return y

    def loss_grad(self, x, t):
        x, t = _f32(x), _f32(t)
        g = np.empty(self.n_params, dtype=np.float32)
        loss = C.c_double()
        self._check(self.lib.fi_surrogate_loss_grad(self.h, _p(x), _p(t), C.byref(loss), _p(g)))
        return loss.value, g

    def set_data(self, X, Y):
        """X: (samples, cells, channels), Y: (samples, cells)"""
        X, Y = _f32(X), _f32(Y)
        self._check(self.lib.fi_surrogate_set_data(self.h, X.shape[0], _p(X), _p(Y)))

    def train(self, steps, batch=4, lr=1e-3, lr_min=1e-5, seed=1):
        losses = np.empty(max(steps, 1), dtype=np.float64)
        self._check(self.lib.fi_surrogate_train(self.h, int(steps), int(batch), float(lr), float(lr_min), int(seed),
                                                losses.ctypes.data_as(C.POINTER(C.c_double))))
        return losses[:steps]

    def device_ms(self):
        return float(self.lib.fi_surrogate_device_ms(self.h))

    def save(self, path):
        np.savez(path, theta=self.get_params(), cfg=np.array(list(self.config_dict().values())),
                 scales=np.array(self.scales if self.scales is not None else []))

    def load(self, path):
        d = np.load(path, allow_pickle=False)
        self.set_params(d["theta"])
        if d["scales"].size:
            self.scales = tuple(float(v) for v in d["scales"])
