"""The learned surrogate (include/fastiso_surrogate.h) and its use with the plain solver.

Surrogate wraps the device network: forward, loss gradient, training on a device-resident set.
The network maps [u, load anomaly, log10 viscosity, lithosphere thickness] (scaled) to the viscous
displacement increment over one macro step (scaled). It has no counterpart in the reference.
"""
import ctypes as C
import time

import numpy as np

from . import _abi
from .engine import EngineError

_fp = C.POINTER(C.c_float)


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _p(a):
    return a.ctypes.data_as(_fp)


def augment(X, Y, nx, ny):
    """The 8 symmetries of the square grid (flips of x and y, and the transpose), applied to every
    channel: the LV-ELVA operator on a square domain is invariant under them."""
    if nx != ny:
        return X, Y
    Xs, Ys = [X], [Y]
    Xg = X.reshape(X.shape[0], ny, nx, X.shape[2])   # [sample, j, i, channel]
    Yg = Y.reshape(Y.shape[0], ny, nx)
    for tr in (False, True):
        for fx in (False, True):
            for fy in (False, True):
                if not (tr or fx or fy):
                    continue
                a, b = Xg, Yg
                if tr:
                    a, b = a.transpose(0, 2, 1, 3), b.transpose(0, 2, 1)
                if fx:
                    a, b = a[:, :, ::-1], b[:, :, ::-1]
                if fy:
                    a, b = a[:, ::-1], b[:, ::-1]
                Xs.append(np.ascontiguousarray(a).reshape(X.shape))
                Ys.append(np.ascontiguousarray(b).reshape(Y.shape))
    return np.concatenate(Xs), np.concatenate(Ys)


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

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
        return False

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

    def loss_grad_n(self, X, T):
        """loss and gradient summed over the samples X: (samples, cells, channels), T: (samples, cells),
        which go through the network side by side"""
        X, T = _f32(X), _f32(T)
        g = np.empty(self.n_params, dtype=np.float32)
        loss = C.c_double()
        self._check(self.lib.fi_surrogate_loss_grad_n(self.h, int(X.shape[0]), _p(X), _p(T), C.byref(loss), _p(g)))
        return loss.value, g

    def set_data(self, X, Y):
        """X: (samples, cells, channels), Y: (samples, cells)"""
        X, Y = _f32(X), _f32(Y)
        self._check(self.lib.fi_surrogate_set_data(self.h, X.shape[0], _p(X), _p(Y)))

    def train(self, steps, batch=4, lr=1e-3, lr_min=1e-5, seed=1, wd=0.0):
        losses = np.empty(max(steps, 1), dtype=np.float64)
        self._check(self.lib.fi_surrogate_train(self.h, int(steps), int(batch), float(lr), float(lr_min), float(wd),
                                                int(seed), losses.ctypes.data_as(C.POINTER(C.c_double))))
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


# ---- the LV-ELVA macro step -------------------------------------------------------------------
# Fixed physical scales of the network's inputs and output.
U_SCALE = 500.0          # m, displacement
S_SCALE = 1.0e6          # kg m^-2, load anomaly (about 1100 m of ice)
D_SCALE = 2.0            # m per macro step, correction of the extrapolated increment
R_SCALE = 0.02           # m per macro step, correction of the RKC increment
LOGETA_MID, LOGETA_HALF = 21.0, 2.0
H_MID, H_HALF = 150e3, 100e3


class MacroStep:
    """Inputs and scaling of the surrogate for one simulation, and the spin-up step of its
    training states.

    The network input is [u, load anomaly, log10 viscosity, lithosphere thickness, previous
    increment] (scaled), the load anomaly being the surface load at the step midpoint plus the
    lithospheric column anomaly rho_litho (ue - ue_ref) of the current elastic displacement
    (src/loads.jl:41-44), which is of the size of the disequilibrium that drives the viscous flow.
    The increment over a macro step of macro_dt years is the previous increment (linear
    extrapolation in time) plus the network's correction:
        du = du_prev + D_SCALE * network(input)

    baseline() (used only to spin up training states) blends exact steps of laterally
    homogeneous Earths: for each (eta_j, D_l) of a lattice
    spanning the simulation's effective viscosity and rigidity, mode by mode
        du_k = (u_eq,k - u_k)(1 - exp(-beta_k dt / (2 eta_j kappa_k))),  u_eq,k = -g sigma_k / beta_k,
        beta_k = rho_um g + D_l kappa_k^4
    (the LV-ELVA operator of src/deformation.jl:91-131 with constant fields), and every cell takes
    the bilinear blend in (log10 eta, log10 D) of its own values; the result is corner-corrected
    like the plain solver's displacement."""

    def __init__(self, domain, solidearth, c, load_fn, macro_dt=500.0, n_eta=10, n_D=3):
        self.domain, self.c, self.macro_dt = domain, c, float(macro_dt)
        self.load_fn = load_fn
        eta = np.asarray(solidearth.effective_viscosity, dtype=np.float64)
        D = np.asarray(solidearth.litho_rigidity, dtype=np.float64)
        self.logeta = np.log10(eta)
        self.h = None
        self.rho_um = solidearth.rho_uppermantle
        self.rho_litho = solidearth.rho_litho
        nx, ny = domain.nx, domain.ny
        kx = np.pi / domain.Wx * np.fft.fftfreq(nx, d=1.0 / nx)
        ky = np.pi / domain.Wy * np.fft.rfftfreq(ny, d=1.0 / ny)
        self.kappa = np.sqrt(kx[:, None] ** 2 + ky[None, :] ** 2)
        logD = np.log10(np.maximum(D, 1.0))
        self.le = np.linspace(self.logeta.min(), self.logeta.max(), n_eta) if np.ptp(self.logeta) > 0 else \
            np.array([self.logeta.min()])
        self.lD = np.linspace(logD.min(), logD.max(), n_D) if np.ptp(logD) > 0 else np.array([logD.min()])
        self.weights = [self._lin_weights(self.logeta, self.le), self._lin_weights(logD, self.lD)]

    @staticmethod
    def _lin_weights(v, grid):
        """Per cell: (index of the lower lattice point, weight of the upper one)."""
        if grid.size == 1:
            return np.zeros(v.shape, dtype=int), np.zeros(v.shape)
        i = np.clip(np.searchsorted(grid, v, side="right") - 1, 0, grid.size - 2)
        w = np.clip((v - grid[i]) / (grid[i + 1] - grid[i]), 0.0, 1.0)
        return i, w

    def load(self, t, ue):
        """Load anomaly driving the step from t: surface load at the midpoint + rho_litho ue."""
        return self.load_fn(t + 0.5 * self.macro_dt) + self.rho_litho * ue

    def features(self, u, ue, du_prev, t, h_field):
        sig = self.load(t, ue)
        return stack_inputs([u / U_SCALE, sig / S_SCALE, (self.logeta - LOGETA_MID) / LOGETA_HALF,
                             (h_field - H_MID) / H_HALF, du_prev / D_SCALE])

    def _homogeneous(self, U, S, logeta, logD, dt):
        beta = self.rho_um * self.c.g + 10.0 ** logD * self.kappa ** 4
        with np.errstate(divide="ignore", invalid="ignore"):
            ueq = -self.c.g * S / beta
            rate = beta / (2 * 10.0 ** logeta * self.kappa) * self.c.seconds_per_year
            fac = -np.expm1(-rate * dt)
        dU = (ueq - U) * fac
        dU[0, 0] = 0.0
        return np.fft.irfft2(dU, s=self.kappa.shape[:1] + (self.domain.ny,))

    def baseline(self, u, ue, t):
        """Blended homogeneous-Earth increment over [t, t + macro_dt] with the midpoint load."""
        sig = self.load(t, ue)
        U, S = np.fft.rfft2(u), np.fft.rfft2(sig)
        (ie, we), (iD, wD) = self.weights
        du = np.zeros_like(u)
        for a in range(self.le.size):
            for b in range(self.lD.size):
                wa = np.where(ie == a, 1 - we, 0.0) + np.where(ie == a - 1, we, 0.0) if self.le.size > 1 else 1.0
                wb = np.where(iD == b, 1 - wD, 0.0) + np.where(iD == b - 1, wD, 0.0) if self.lD.size > 1 else 1.0
                w = wa * wb
                if np.ndim(w) and not np.any(w):
                    continue
                du += w * self._homogeneous(U, S, self.le[a], self.lD[b], self.macro_dt)
        return du - 0.25 * (du[0, 0] + du[-1, 0] + du[0, -1] + du[-1, -1])

    def equilibrium(self, t):
        """Equilibrium displacement under the load at t with the mean rigidity."""
        S = np.fft.rfft2(self.load_fn(t))
        beta = self.rho_um * self.c.g + 10.0 ** self.lD.mean() * self.kappa ** 4
        U = -self.c.g * S / beta
        U[0, 0] = 0.0
        u = np.fft.irfft2(U, s=self.kappa.shape[:1] + (self.domain.ny,))
        return u - 0.25 * (u[0, 0] + u[-1, 0] + u[0, -1] + u[-1, -1])


def lv_macro(w, macro_dt=500.0):
    """MacroStep of an lv_elva workload (separable dome load: sigma = rho_ice a(t) H_dome)."""
    c = w["c"]
    it = w["bcs"].ice_thickness
    tk, ak = np.asarray(it.t_vec), np.asarray(it.a_vec)
    Hd = np.asarray(it.H_shape)

    def load(t):
        return c.rho_ice * np.interp(t, tk, ak) * Hd

    m = MacroStep(w["domain"], w["solidearth"], c, load, macro_dt)
    m.h = np.asarray(w["h"])
    return m


def predict(sur, macro, u, ue, du_prev, t):
    """Surrogate increment of u over [t, t + macro_dt] (sur None: extrapolation alone)."""
    if sur is None:
        return du_prev.copy()
    x = macro.features(u, ue, du_prev, t, macro.h)
    r = sur.forward(x).astype(np.float64)
    nx, ny = u.shape
    return du_prev + D_SCALE * r.reshape((nx, ny), order="F")


class SurrogateRunner:
    """Surrogate-with-fallback integration of a HIP Simulation in macro steps of macro_dt years.

    The run proceeds in periods of check_every macro steps. The predictor takes the first
    check_every - 1 steps; for the last one it predicts the step, the state is restored to the
    step's start and the plain solver integrates it. If max |u_plain - u_predicted| <= tol the
    period is kept (continuing from the plain state); otherwise the run restores the state of the
    period's start and the plain solver integrates the whole period (the fallback). With an engine
    that has checkpoints (fi_checkpoint_save / fi_checkpoint_load) every restore is exact, so a
    fallen-back period equals the plain run bit for bit.

    predictor:
      "rkc"            the step is integrated by RKC2 sub-steps of at most rkc_dt years
                       (fi_integrate_rkc): the same right-hand side, stable at any step, so its
                       cost grows like the square root of the stiffness instead of linearly; with a
                       network (sur) the increment is corrected by it (inputs: u, load, log10
                       viscosity, lithosphere thickness, RKC increment);
      "net"            extrapolation of the previous increments in time plus the network's
                       correction (needs sur);
      "extrapolation"  constant (order 1) or linearly changing (order 2) increments.
    The first `warm` steps (default 0 for "rkc", 2 otherwise) are plain: they give the
    extrapolating predictors their first increments. With select, each check also evaluates the
    predictor's alternative (RKC without the network, or extrapolation without it) and the next
    period uses whichever was closer to the plain solver. perturb(period, du) may modify the
    check step's prediction (tests force a fallback with it)."""

    def __init__(self, sim, sur, macro, check_every=10, tol=1.0, predictor="rkc", rkc_dt=250.0, select=True,
                 warm=None, order=2, perturb=None):
        if predictor not in ("rkc", "net", "extrapolation"):
            raise ValueError(f"unknown predictor {predictor}")
        if predictor == "net" and sur is None:
            raise ValueError("the net predictor needs a network")
        self.sim, self.sur, self.macro = sim, sur, macro
        self.check_every, self.tol = int(check_every), float(tol)
        self.predictor, self.rkc_dt = predictor, float(rkc_dt)
        self.warm = int(warm) if warm is not None else (0 if predictor == "rkc" else 2)
        self.order = int(order)
        self.select = select and sur is not None
        self.use_net = sur is not None
        self.perturb = perturb
        self.log = []
        self.rkc_rhs = 0
        # wall time and simulated years of the plain-solver integrations of run() (checks and
        # fallbacks): the plain solver's rate on the run's own states
        self.plain_s = 0.0
        self.plain_yr = 0.0
        # output scale of a network correcting the RKC increment (its saved scales: U, S, scale,
        # macro_dt, base 1 = RKC); a network for the extrapolated increment is not one
        sc = tuple(getattr(sur, "scales", None) or ()) if sur is not None else ()
        self.net_scale = float(sc[2]) if len(sc) > 2 else R_SCALE
        if predictor == "rkc" and sur is not None and not (len(sc) > 4 and sc[4] == 1.0):
            raise ValueError("the network was not trained on RKC increments (its scales lack base 1)")

    # ---- engine state -------------------------------------------------------------------------
    def _save(self, eng):
        if hasattr(eng, "checkpoint"):
            return ("ck", eng.checkpoint())
        return ("u", eng.get_field("u"), eng.stats()["t"])

    @staticmethod
    def _restore(eng, st):
        if st[0] == "ck":
            eng.restore(st[1])
        else:
            eng.set_field("u", st[1])
            eng.set_time(st[2])

    # ---- one predicted macro step from the engine's state (advances the engine) ------------------
    def _step(self, eng, kind, net, inc, t, t1):
        u0 = eng.get_field("u")
        if kind == "rkc":
            info = eng.integrate_rkc(t1, self.rkc_dt)
            self.rkc_rhs += int(info["n_rhs"]) if isinstance(info, dict) else 0
            u1 = eng.get_field("u")
            du = u1 - u0
            if net:
                ue = eng.get_field("ue")
                x = self.macro.features(u0, ue, du, t, self.macro.h)
                r = self.sur.forward(x).astype(np.float64).reshape(u0.shape, order="F")
                du = du + self.net_scale * r
                eng.set_field("u", u0 + du)
                eng.set_time(t1)
            return du
        if net:
            du = predict(self.sur, self.macro, u0, eng.get_field("ue"), inc[-1], t)
        else:
            du = 2.0 * inc[-1] - inc[-2] if self.order == 2 else inc[-1].copy()
        eng.set_field("u", u0 + du)
        eng.set_time(t1)
        return du

    def run(self, t_end, t0=None, progress=None):
        eng = self.sim.engine
        dt = self.macro.macro_dt
        t = eng.stats()["t"] if t0 is None else t0
        n_plain = n_sur = n_fallback = n_wasted = 0
        hist = []
        kind = self.predictor
        name = {True: "net", False: "extrapolation"} if kind != "rkc" else {True: "rkc-net", False: "rkc"}

        def note(t1, k, dev):
            self.log.append((t1, k, dev))
            if progress:
                progress(t1, k, dev)

        def plain_to(t1):
            u0 = eng.get_field("u")
            ta, w0 = eng.stats()["t"], time.perf_counter()
            eng.integrate(t1)
            u1 = eng.get_field("u")
            self.plain_s += time.perf_counter() - w0
            self.plain_yr += t1 - ta
            return u1 - u0

        n_period = 0
        while t < t_end - 1e-9:
            t1 = min(t + dt, t_end)
            remaining = int(round((t_end - t) / dt))
            period = min(self.check_every, remaining)   # the last period may be shorter
            if len(hist) < self.warm or period < 2 or t1 - t < dt:
                hist.append(plain_to(t1))
                n_plain += 1
                note(t1, "plain", None)
                t = t1
                continue
            n_period += 1
            start = self._save(eng)
            inc, tp = list(hist[-2:]), t
            du_start = hist[-1] if hist else None
            for _ in range(period - 1):
                du = self._step(eng, kind, self.use_net, inc, tp, tp + dt)
                inc = [inc[-1], du] if inc else [du]
                tp += dt
                note(tp, "surrogate", None)
            # the check step: the predictor (and its alternative), then the plain solver
            at_check = self._save(eng)
            cands = [self.use_net] + ([not self.use_net] if self.select else [])
            dus = {}
            for net in cands:
                self._restore(eng, at_check)
                dus[net] = self._step(eng, kind, net, inc, tp, tp + dt)
            if self.perturb is not None:
                dus[self.use_net] = self.perturb(n_period, dus[self.use_net])
            self._restore(eng, at_check)
            du_p = plain_to(tp + dt)
            devs = {net: float(np.abs(du_p - d).max()) for net, d in dus.items()}
            dev = devs[self.use_net]
            used = name[self.use_net]
            if self.select:
                self.use_net = devs[True] <= devs[False]
            if dev <= self.tol:
                n_sur += period - 1
                n_plain += 1
                if kind != "rkc" and du_start is not None:
                    # seed the next period's extrapolation with the plain check increment and the
                    # slope of the plain increments across the period (first plain increment to the
                    # check, period steps apart); the slope between the last surrogate step and
                    # the check would carry the check's one-off relaxation of the surrogate's
                    # error, amplified period-fold
                    slope = (du_p - du_start) / period
                    hist.append(du_p - slope)
                hist.append(du_p)
                note(tp + dt, f"check-{used}", dev)
            else:
                # fallback: back to the start of the period, integrated by the plain solver
                n_fallback += 1
                n_wasted += period - 1
                note(tp + dt, f"check-{used}-fallback", dev)
                self._restore(eng, start)
                tq = t
                for _ in range(period):
                    hist.append(plain_to(tq + dt))
                    tq += dt
                    n_plain += 1
                    note(tq, "plain", None)
            t = tp + dt
        return {"predictor": self.predictor, "plain_steps": n_plain, "surrogate_steps": n_sur,
                "fallbacks": n_fallback, "discarded_surrogate_steps": n_wasted, "rkc_rhs_calls": self.rkc_rhs,
                "network_periods": sum(1 for _, k, _ in self.log if k.startswith("check-net")
                                       or k.startswith("check-rkc-net")),
                "checks": sum(1 for _, k, _ in self.log if k.startswith("check-")),
                "plain_yr_in_run": self.plain_yr, "plain_s_in_run": self.plain_s}


def generate_rkc_pairs(n_members, seed0, steps_per_member, macro_dt=500.0, n=350, W=3.15e6, t_range=(5e3, 115e3),
                       rkc_dt=250.0, team_size=0, spin=4, progress=None):
    """Training pairs for a network that corrects the RKC macro step: member m gets C1-style
    fields and domes (seeds seed0 + m, seed0 + 100000 + m, seed0 + 200000 + m) and a random time
    t_m of the glacial cycle; its state there is spun up by spin RKC macro steps from a random
    fraction of the equilibrium. Then, macro step by macro step, every member takes an RKC step
    (fi_integrate_rkc from a checkpoint, which is restored) and the ensemble engine integrates
    all members with the plain solver; each gives a pair (input with the RKC increment, scaled
    residual (du_plain - du_rkc) / R_SCALE). Returns (X, Y, meta): X (pairs, cells, 5),
    Y (pairs, cells), meta (member, t, max |du_plain|, max |du_plain - du_rkc|)."""
    from . import workloads
    from .ensemble import Ensemble
    from .simulation import Simulation
    rng = np.random.default_rng(seed0)
    macros, inputs, starts = [], [], []
    for m in range(n_members):
        w = workloads.lv_elva(n=n, W=W, t_end=120e3, seeds=(seed0 + m, seed0 + 100000 + m, seed0 + 200000 + m),
                              t_out=[])
        macros.append(lv_macro(w, macro_dt))
        inputs.append(Simulation(w["domain"], w["bcs"], w["sealevel"], w["solidearth"], w["t_span"],
                                 nout=w["nout"], opts=w["opts"], c=w["c"], engine=None).inputs)
        starts.append((float(rng.uniform(*t_range)), float(rng.uniform(0.5, 1.0))))
    ens = Ensemble(inputs, team_size=team_size)
    ens.integrator_init()
    for m, (ts, f) in enumerate(starts):
        t = ts - spin * macro_dt
        ens[m].set_field("u", f * macros[m].equilibrium(t))
        ens[m].set_time(t)
        ens[m].integrate_rkc(ts, rkc_dt)
    X, Y, meta = [], [], []
    for s in range(steps_per_member):
        base = []
        for m in range(n_members):
            e = ens[m]
            t_m = starts[m][0] + s * macro_dt
            ck = e.checkpoint()
            u0, ue0 = e.get_field("u"), e.get_field("ue")
            e.integrate_rkc(t_m + macro_dt, rkc_dt)
            base.append((u0, ue0, e.get_field("u") - u0, t_m))
            e.restore(ck)
        ens.integrate_each([starts[m][0] + (s + 1) * macro_dt for m in range(n_members)], fsal=False)
        for m in range(n_members):
            u0, ue0, du_r, t_m = base[m]
            du = ens[m].get_field("u") - u0
            X.append(macros[m].features(u0, ue0, du_r, t_m, macros[m].h))
            Y.append(cells((du - du_r) / R_SCALE).astype(np.float32))
            meta.append((m, t_m, float(np.abs(du).max()), float(np.abs(du - du_r).max())))
        if progress:
            progress(s, ens.info())
    ens.close()
    return np.array(X), np.array(Y), meta


def generate_pairs(n_members, seed0, steps_per_member, macro_dt=500.0, n=350, W=3.15e6, t_range=(5e3, 115e3),
                   team_size=0, spin=6, skip=1, progress=None):
    """Training pairs of the plain solver: member m gets C1-style fields and domes (seeds seed0 + m,
    seed0 + 100000 + m, seed0 + 200000 + m) and a random time t_m of the glacial cycle. Its state at
    t_m is spun up by spin baseline macro steps from a random fraction of the equilibrium, then the
    ensemble engine integrates every member steps_per_member macro steps; step k >= skip gives a
    pair (input with the increment of step k - 1, scaled correction du_k - du_{k-1}): the first
    steps after the spin-up carry its transient. Returns
    (X, Y, meta): X (pairs, cells, 5), Y (pairs, cells)."""
    from . import workloads
    from .ensemble import Ensemble
    from .simulation import Simulation
    rng = np.random.default_rng(seed0)
    ws, macros, inputs, starts = [], [], [], []
    for m in range(n_members):
        w = workloads.lv_elva(n=n, W=W, t_end=120e3, seeds=(seed0 + m, seed0 + 100000 + m, seed0 + 200000 + m),
                              t_out=[])
        ws.append(w)
        macros.append(lv_macro(w, macro_dt))
        inputs.append(Simulation(w["domain"], w["bcs"], w["sealevel"], w["solidearth"], w["t_span"],
                                 nout=w["nout"], opts=w["opts"], c=w["c"], engine=None).inputs)
        starts.append((float(rng.uniform(*t_range)), float(rng.uniform(0.5, 1.0))))
    ens = Ensemble(inputs, team_size=team_size)
    ens.integrator_init()
    for m, (ts, f) in enumerate(starts):
        t = ts - spin * macro_dt
        u = f * macros[m].equilibrium(t)
        ens[m].set_field("u", u)
        ens[m].set_time(t)
        for _ in range(spin):
            u = u + macros[m].baseline(u, ens[m].get_field("ue"), t)
            t += macro_dt
            ens[m].set_field("u", u)
            ens[m].set_time(t)
    X, Y, meta = [], [], []
    states = [(ens[m].get_field("u"), ens[m].get_field("ue")) for m in range(n_members)]
    prev = [None] * n_members
    for s in range(steps_per_member):
        ens.integrate_each([starts[m][0] + (s + 1) * macro_dt for m in range(n_members)], fsal=False)
        for m in range(n_members):
            t_m = starts[m][0] + s * macro_dt
            (u0, ue0), u1 = states[m], ens[m].get_field("u")
            du = u1 - u0
            if prev[m] is not None and s >= skip:
                X.append(macros[m].features(u0, ue0, prev[m], t_m, macros[m].h))
                Y.append(cells((du - prev[m]) / D_SCALE).astype(np.float32))
                meta.append((m, t_m, float(np.abs(du).max()), float(np.abs(du - prev[m]).max())))
            prev[m] = du
            states[m] = (u1, ens[m].get_field("ue"))
        if progress:
            progress(s, ens.info())
    ens.close()
    return np.array(X), np.array(Y), meta
