"""Inversion of solid-Earth parameters with unscented Kalman inversion (UKI), members advanced
together by the ensemble engine.

The reference (src/inversion.jl and ext/FastIsostasyEnsembleKalmanProcessesExt.jl:39-167) builds
an InversionProblem from a Simulation, an InversionConfig, InversionData and a user
ParameterReduction, and iterates EnsembleKalmanProcesses' Unscented process: sigma points in the
unconstrained parameter space, one forward run per sigma point (a deep copy of the simulation
per thread, :103-122), then update_ensemble!. Here every iteration's sigma points run as one
fi_batch (fastiso.Ensemble), and the Kalman update is restated below.

UKI restated from Huang, Schneider & Stuart, "Iterated Kalman methodology for inverse problems"
(J. Comput. Phys. 463, 2022), in the form EnsembleKalmanProcesses 2.x (the reference's compat
bound, Project.toml:49) implements as `Unscented` with its defaults (symmetric sigma points,
modified unscented transform):
* prediction: m^ = r + a (m - r), C^ = a^2 C + S_w, with a = alpha_reg, r the prior mean and
  S_w = (2 - a^2) times the prior covariance (re-set to (2 - a^2) C every update_freq iterations
  when update_freq > 0);
* 2N+1 sigma points m^ and m^ +- c [sqrt C^]_i, c = a_ut sqrt(N), a_ut = min(sqrt(4/N), 1);
* the predicted observation is G at the central point; covariances use the 2N outer points with
  weight 1 / (2 a_ut^2 N); the data covariance gets 2 Gamma (Sigma_nu_scale = 2);
* analysis: m = m^ + C_ug C_gg^-1 (y - g^), C = C^ - C_ug C_gg^-1 C_ug^T;
* error: (y - g^)^T Gamma^-1 (y - g^).
The observation is the mean of n_samples noisy copies of the truth with the diagonal covariance
scale_obscov / 2 (uncorrelated_obs_covariance, :91-94). Parameters are Gaussian in an
unconstrained space and mapped to physical values by a per-parameter transform (identity,
bounded, or bounded below); EKP's constrained_gaussian moment matching is not restated: the
caller gives the unconstrained mean and standard deviation.
"""
import numpy as np


class InversionConfig:
    """InversionConfig(method, N_ens, N_iter, n_samples; alpha_reg, update_freq, scale_obscov)
    (src/inversion.jl:26-41). Only the unscented method exists, as in the reference."""

    def __init__(self, method="unscented", N_ens=None, N_iter=5, n_samples=1, alpha_reg=1.0, update_freq=1,
                 scale_obscov=1000.0):
        if str(method).lower() not in ("unscented", "uki"):
            raise NotImplementedError("only the unscented Kalman inversion is available")
        self.method, self.N_ens, self.N_iter, self.n_samples = "unscented", N_ens, int(N_iter), int(n_samples)
        self.alpha_reg, self.update_freq, self.scale_obscov = float(alpha_reg), int(update_freq), float(scale_obscov)


class InversionData:
    """InversionData(t, Y, mask) (src/inversion.jl:57-69): fields Y[k] at times t[k], observed on mask."""

    def __init__(self, t, Y, mask):
        self.t = [float(x) for x in t]
        self.Y = [np.asarray(y, dtype=np.float64) for y in Y]
        self.nY = len(self.Y)
        self.mask = np.asarray(mask, dtype=bool)
        self.countmask = int(self.mask.sum())
        self.y = np.concatenate([y[self.mask] for y in self.Y])


class Transform:
    """Unconstrained -> physical map of one parameter: identity, bounded (lower, upper) with
    phi = upper - (upper - lower) / (exp(u) + 1), or bounded below with phi = lower + exp(u)."""

    def __init__(self, lower=None, upper=None):
        self.lower, self.upper = lower, upper

    def __call__(self, u):
        u = np.asarray(u, dtype=np.float64)
        if self.lower is not None and self.upper is not None:
            return self.upper - (self.upper - self.lower) / (np.exp(u) + 1.0)
        if self.lower is not None:
            return self.lower + np.exp(u)
        if self.upper is not None:
            return self.upper - np.exp(-u)
        return u


class GaussianPrior:
    """Independent Gaussian priors in the unconstrained space with a transform per parameter."""

    def __init__(self, mean, std, transforms=None, names=None):
        self.mean = np.asarray(mean, dtype=np.float64)
        self.cov = np.diag(np.asarray(std, dtype=np.float64) ** 2)
        n = self.mean.size
        self.transforms = list(transforms) if transforms is not None else [Transform() for _ in range(n)]
        self.names = list(names) if names is not None else [f"theta_{i}" for i in range(n)]

    def to_physical(self, U):
        """Columns of U (N_par x N_ens, or one vector) -> physical parameters."""
        U = np.asarray(U, dtype=np.float64)
        out = np.empty_like(U)
        for i, tr in enumerate(self.transforms):
            out[i] = tr(U[i])
        return out


class Unscented:
    """The unscented Kalman inversion process (see the module docstring)."""

    def __init__(self, mean, cov, alpha_reg=1.0, update_freq=0):
        self.r = np.asarray(mean, dtype=np.float64).copy()
        self.u_mean = [self.r.copy()]
        self.uu_cov = [np.asarray(cov, dtype=np.float64).copy()]
        self.alpha_reg = float(alpha_reg)
        self.update_freq = int(update_freq)
        self.S_w = (2.0 - self.alpha_reg ** 2) * self.uu_cov[0]
        self.N = self.r.size
        self.a_ut = min(np.sqrt(4.0 / self.N), 1.0)
        self.c = self.a_ut * np.sqrt(self.N)
        self.w_cov = 1.0 / (2.0 * self.a_ut ** 2 * self.N)
        self.iter = 0
        self.error = []
        self.g_mean = []
        self._pred = None

    @property
    def N_ens(self):
        return 2 * self.N + 1

    def sigma_points(self):
        """Prediction step and the 2N+1 sigma points (columns) of this iteration."""
        a = self.alpha_reg
        m = a * self.u_mean[-1] + (1.0 - a) * self.r
        C = a ** 2 * self.uu_cov[-1] + self.S_w
        L = np.linalg.cholesky(0.5 * (C + C.T))
        U = np.empty((self.N, self.N_ens))
        U[:, 0] = m
        for i in range(self.N):
            U[:, 1 + i] = m + self.c * L[:, i]
            U[:, 1 + self.N + i] = m - self.c * L[:, i]
        self._pred = (m, C, U)
        return U

    def update(self, G, y, Gamma):
        """update_ensemble!(uki, G) with the forward evaluations G (N_obs x N_ens) of the sigma
        points of sigma_points(). Gamma is the observation covariance: an N_obs x N_obs matrix,
        or the vector of its diagonal. With a diagonal the gain is formed through the Woodbury
        identity: C_gg = A A^T + R with A = sqrt(w) dG (rank <= 2N) and R = 2 Gamma, so
        K = C_ug C_gg^-1 = sqrt(w) dU M^-1 A^T R^-1 and K C_ug^T = w dU (I - M^-1) dU^T with
        M = I + A^T R^-1 A (2N x 2N); no N_obs x N_obs matrix is formed. For these sigma points
        w dU dU^T = L L^T is the predicted covariance C (2 c^2 w = 1), so the new covariance
        C - K C_ug^T is formed as w dU M^-1 dU^T: a posterior far tighter than the prediction
        (small Gamma) keeps its relative accuracy instead of being the difference of two nearly
        equal matrices."""
        m, C, U = self._pred
        G = np.asarray(G, dtype=np.float64)
        Gamma = np.asarray(Gamma, dtype=np.float64)
        g = G[:, 0]
        dU = U[:, 1:] - m[:, None]
        dG = G[:, 1:] - g[:, None]
        d = y - g
        if Gamma.ndim == 1:
            sw = np.sqrt(self.w_cov)
            A = sw * dG
            BR = A.T / (2.0 * Gamma)[None, :]          # A^T R^-1
            M = np.eye(A.shape[1]) + BR @ A
            m_new = m + sw * dU @ np.linalg.solve(M, BR @ d)
            C_new = self.w_cov * dU @ np.linalg.solve(M, dU.T)
            err = float(d @ (d / Gamma))
        else:
            C_ug = self.w_cov * dU @ dG.T
            C_gg = self.w_cov * dG @ dG.T + 2.0 * Gamma
            K = np.linalg.solve(C_gg, C_ug.T).T
            m_new = m + K @ d
            C_new = C - K @ C_ug.T
            err = float(d @ np.linalg.solve(Gamma, d))
        self.u_mean.append(m_new)
        self.uu_cov.append(0.5 * (C_new + C_new.T))
        self.g_mean.append(g)
        self.error.append(err)
        self.iter += 1
        if self.update_freq > 0 and self.iter % self.update_freq == 0:
            self.S_w = (2.0 - self.alpha_reg ** 2) * self.uu_cov[-1]


class ParameterReduction:
    """ParameterReduction (src/inversion.jl:1-8). Subclasses give nparams and
    reconstruct(params) -> Simulation inputs (reconstruct!, src/inversion.jl:129-133); the
    default extract_output (src/inversion.jl:136-140) is the viscous displacement on the data
    mask at each data time."""

    nparams = 0

    def reconstruct(self, params):
        raise NotImplementedError

    def extract_output(self, engine, data, k):
        return engine.get_field("u")[data.mask]


class RegionalViscosityReduction(ParameterReduction):
    """log10 of the upper-mantle viscosity on each of a list of region masks, above a base
    Earth: builder(eta_field) must return Simulation(...).inputs for that viscosity field."""

    def __init__(self, base_log10_eta, regions, builder):
        self.base = np.asarray(base_log10_eta, dtype=np.float64)
        self.regions = [np.asarray(r, dtype=bool) for r in regions]
        self.builder = builder
        self.nparams = len(self.regions)

    def field(self, params):
        f = self.base.copy()
        for r, p in zip(self.regions, params):
            f[r] = p
        return 10.0 ** f

    def reconstruct(self, params):
        return self.builder(self.field(params))


class InversionProblem:
    """inversion_problem(sim, config, data, reduction, priors) (ext :39-89) and solve!
    (ext :97-139): make_batch(inputs_list) returns an ensemble (fastiso.Ensemble on the device)
    whose members run the sigma points of one iteration."""

    def __init__(self, config, data, reduction, prior, make_batch, seed=0, t0=0.0):
        self.config, self.data, self.reduction, self.prior = config, data, reduction, prior
        self.make_batch = make_batch
        self.t0 = float(t0)
        nobs = data.countmask * data.nY
        # observation noise covariance (scale_obscov / 2) I, kept as its diagonal
        self.Gamma = np.full(nobs, config.scale_obscov / 2.0)
        rng = np.random.default_rng(seed)
        yn = data.y[:, None] + rng.normal(0.0, np.sqrt(config.scale_obscov / 2.0), (nobs, config.n_samples))
        self.y_obs = yn.mean(axis=1)
        self.process = Unscented(prior.mean, prior.cov, config.alpha_reg, config.update_freq)
        if config.N_ens is not None and config.N_ens != self.process.N_ens:
            raise ValueError(f"the unscented process runs 2 N_par + 1 = {self.process.N_ens} members")
        self.out = [prior.to_physical(self.process.u_mean[-1])]
        self.G_ens = None
        # every iteration runs the same number of members: once a batch has chosen its team size
        # by timing, the later batches take it instead of timing the candidates again
        self.team_size = 0

    def forward(self, params_phys):
        """One forward run per column of params_phys, all members in one ensemble; returns G
        (N_obs x N_ens): extract_output at every data time, stacked in time order."""
        inputs = [self.reduction.reconstruct(params_phys[:, j]) for j in range(params_phys.shape[1])]
        batch = self.make_batch(inputs)
        try:
            if self.team_size and hasattr(batch, "set_team_size"):
                batch.set_team_size(self.team_size)
            batch.integrator_init()
            G = [[] for _ in inputs]
            for k, t in enumerate(self.data.t):
                if t > self.t0:
                    batch.integrate(t)
                for j in range(len(inputs)):
                    G[j].append(self.reduction.extract_output(batch[j], self.data, k))
            if not self.team_size and hasattr(batch, "info"):
                info = batch.info()
                if info.get("tuned") and info.get("team_size"):
                    self.team_size = int(info["team_size"])
            return np.stack([np.concatenate(g) for g in G], axis=1)
        finally:
            close = getattr(batch, "close", None)
            if close is not None:
                close()

    def solve(self, verbose=False):
        for n in range(self.config.N_iter):
            U = self.process.sigma_points()
            phys = self.prior.to_physical(U)
            self.G_ens = self.forward(phys)
            self.process.update(self.G_ens, self.y_obs, self.Gamma)
            self.out.append(self.prior.to_physical(self.process.u_mean[-1]))
            if verbose:
                print(f"iteration {n + 1}: error {self.process.error[-1]:.4g}, mean {self.out[-1]}")
        return self

    @property
    def error(self):
        return list(self.process.error)

    def extract_inversion(self):
        """extract_inversion (ext :147-152): physical mean, the mean prediction and its absolute
        error against the (noise-free) data."""
        p = self.prior.to_physical(self.process.u_mean[-1])
        Gx = self.process.g_mean[-1]
        return p, Gx, np.abs(self.data.y - Gx)


def inversion_problem(config, data, reduction, prior, make_batch=None, seed=0, t0=0.0):
    if make_batch is None:
        from .ensemble import Ensemble
        make_batch = Ensemble
    return InversionProblem(config, data, reduction, prior, make_batch, seed=seed, t0=t0)
