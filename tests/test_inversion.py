"""Unscented Kalman inversion driving the ensemble engine (the inversion path,
ext/FastIsostasyEnsembleKalmanProcessesExt.jl:39-167, src/inversion.jl).

EnsembleKalmanProcesses (the reference's dependency, compat 2.0) is absent, so the process is
pinned on what the mathematics fixes for a linear forward map G(theta) = A theta, where the
unscented transform is exact:
* with alpha_reg = 1 the first update is the Bayesian posterior of the Gaussian prior: mean
  m0 + C0 A^T (A C0 A^T + Gamma)^-1 (y - A m0) and twice the posterior covariance (the prediction
  doubles the prior, the data covariance enters as 2 Gamma). A wrong sigma-point weight or
  spread fails it;
* with update_freq = 1 the iterates converge to the least-squares solution
  (A^T Gamma^-1 A)^-1 A^T Gamma^-1 y with covariance (A^T Gamma^-1 A)^-1 (the precision halves
  and gains A^T Gamma^-1 A / 2 per iteration).
Step-by-step parity with EnsembleKalmanProcesses itself is unpinned.

The driver then recovers known upper-mantle viscosities from synthetic displacement data of the
disc benchmark, and a known ELRA relaxation time from an ELRA ensemble: with the oracle's
ensemble on the CPU and with the HIP ensemble on the GPU.
"""
import numpy as np
import pytest

import fastiso
from fastiso import inversion, workloads

from _cases import DRYRUN


def _linear(seed=0, npar=3, nobs=7):
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((nobs, npar))
    m0 = rng.standard_normal(npar)
    S = rng.standard_normal((npar, npar))
    C0 = S @ S.T + npar * np.eye(npar)
    Gamma = np.diag(0.1 + rng.random(nobs))
    y = A @ rng.standard_normal(npar) + 0.05 * rng.standard_normal(nobs)
    return A, m0, C0, Gamma, y


@pytest.mark.parametrize("diagonal", [False, True])
def test_uki_first_step_is_the_gaussian_posterior(diagonal):
    """Dense Gamma, and Gamma given as its diagonal (the Woodbury form of the gain)."""
    A, m0, C0, Gamma, y = _linear()
    p = inversion.Unscented(m0, C0, alpha_reg=1.0, update_freq=0)
    U = p.sigma_points()
    assert U.shape == (3, 7)
    p.update(A @ U, y, np.diag(Gamma) if diagonal else Gamma)
    K = C0 @ A.T @ np.linalg.inv(A @ C0 @ A.T + Gamma)
    np.testing.assert_allclose(p.u_mean[-1], m0 + K @ (y - A @ m0), rtol=1e-10, atol=1e-12)
    np.testing.assert_allclose(p.uu_cov[-1], 2 * (C0 - K @ A @ C0), rtol=1e-9, atol=1e-12)
    d = y - A @ m0
    assert p.error[-1] == pytest.approx(d @ np.linalg.solve(Gamma, d))


def test_uki_converges_to_least_squares():
    A, m0, C0, Gamma, y = _linear(seed=1, npar=4, nobs=9)
    p = inversion.Unscented(m0, C0, alpha_reg=1.0, update_freq=1)
    for _ in range(60):
        p.update(A @ p.sigma_points(), y, Gamma)
    Gi = np.linalg.inv(Gamma)
    P = A.T @ Gi @ A
    m_ls = np.linalg.solve(P, A.T @ Gi @ y)
    np.testing.assert_allclose(p.u_mean[-1], m_ls, rtol=1e-9, atol=1e-10)
    np.testing.assert_allclose(p.uu_cov[-1], np.linalg.inv(P), rtol=1e-7, atol=1e-12)


def test_uki_diagonal_and_dense_noise_agree_on_a_nonlinear_map():
    """The Woodbury gain equals the dense solve when the forward map is not linear (the unscented
    transform is then not exact), over several iterations, with more observations than members."""
    rng = np.random.default_rng(3)
    npar, nobs = 3, 40
    A = rng.standard_normal((nobs, npar))
    gam = 0.05 + rng.random(nobs)
    y = np.tanh(A @ np.array([0.3, -0.2, 0.5])) + 0.01 * rng.standard_normal(nobs)
    m0, C0 = np.zeros(npar), np.eye(npar)
    pd = inversion.Unscented(m0, C0, alpha_reg=0.9, update_freq=1)
    pv = inversion.Unscented(m0, C0, alpha_reg=0.9, update_freq=1)
    for _ in range(5):
        pd.update(np.tanh(A @ pd.sigma_points()), y, np.diag(gam))
        pv.update(np.tanh(A @ pv.sigma_points()), y, gam)
    np.testing.assert_allclose(pv.u_mean[-1], pd.u_mean[-1], rtol=1e-10, atol=1e-12)
    np.testing.assert_allclose(pv.uu_cov[-1], pd.uu_cov[-1], rtol=1e-9, atol=1e-13)
    np.testing.assert_allclose(pv.error, pd.error, rtol=1e-12)


def test_uki_diagonal_noise_keeps_a_tight_posterior_accurate():
    """Observation noise 1e8 times below the prior: the update's covariance must still equal the
    information form (C^-1 + A^T (2 Gamma)^-1 A)^-1 of the linear-Gaussian posterior (C the
    predicted covariance), which involves no cancellation; a covariance formed as C - K C_ug^T
    loses most of its digits here."""
    rng = np.random.default_rng(4)
    npar, nobs = 2, 3000
    A = rng.standard_normal((nobs, npar)) * 5.0
    gam = np.full(nobs, 5e-7)
    y = A @ np.array([0.4, -0.7]) + 1e-4 * rng.standard_normal(nobs)
    C0 = np.array([[0.25, 0.05], [0.05, 0.25]])
    p = inversion.Unscented(np.zeros(npar), C0, alpha_reg=1.0, update_freq=1)
    for _ in range(3):
        U = p.sigma_points()
        C = p._pred[1]
        p.update(A @ U, y, gam)
        C_info = np.linalg.inv(np.linalg.inv(C) + A.T @ (A / (2.0 * gam)[:, None]))
        np.testing.assert_allclose(p.uu_cov[-1], C_info, rtol=1e-6)
        m_info = p._pred[0] + C_info @ (A.T @ ((y - A @ p._pred[0]) / (2.0 * gam)))
        np.testing.assert_allclose(p.u_mean[-1], m_info, rtol=1e-8, atol=1e-12)


def test_transforms_and_observation_noise():
    u = np.linspace(-30, 30, 101)
    b = inversion.Transform(19.0, 23.0)(u)
    assert np.all(np.diff(b) > 0) and b.min() >= 19.0 and b.max() <= 23.0
    assert inversion.Transform(19.0, 23.0)(0.0) == pytest.approx(21.0)
    assert np.all(inversion.Transform(lower=2.0)(u) > 2.0)
    np.testing.assert_array_equal(inversion.Transform()(u), u)
    with pytest.raises(NotImplementedError):
        inversion.InversionConfig("ensemble kalman sampler")


def _disc_problem(n, regions_fn, truth, prior_mean, make_batch, seed=5, t=(1000.0, 2000.0, 3000.0)):
    w = workloads.disc_benchmark(n=n, t_end=max(t))
    d = w["domain"]

    def builder(eta):
        se = fastiso.SolidEarth(d, layer_boundaries=[88e3], layer_viscosities=[eta], rho_uppermantle=3300.0,
                                rho_litho=0.0)
        return fastiso.Simulation(d, w["bcs"], w["sealevel"], se, (0.0, max(t)), nout=fastiso.NativeOutput(),
                                  opts=w["opts"], c=w["c"], engine=None).inputs

    regions = regions_fn(d)
    red = inversion.RegionalViscosityReduction(np.full((d.nx, d.ny), 21.0), regions, builder)
    mask = d.R < 1.5e6
    ens = make_batch([red.reconstruct(np.asarray(truth))])
    ens.integrator_init()
    Y = []
    for tk in t:
        ens.integrate(tk)
        Y.append(ens[0].get_field("u"))
    close = getattr(ens, "close", None)
    if close:
        close()
    data = inversion.InversionData(list(t), Y, mask)
    prior = inversion.GaussianPrior(prior_mean, [0.5] * len(truth))
    cfg = inversion.InversionConfig("unscented", N_iter=4, n_samples=1, alpha_reg=1.0, update_freq=1,
                                    scale_obscov=1e-6)
    return inversion.inversion_problem(cfg, data, red, prior, make_batch=make_batch, seed=seed)


def test_oracle_ensemble_inversion_recovers_viscosity():
    from oracle import pyref
    prob = _disc_problem(16, lambda d: [np.ones((d.nx, d.ny), dtype=bool)], [21.0], [20.6], pyref.RefEnsemble)
    prob.solve()
    p, Gx, err = prob.extract_inversion()
    assert abs(p[0] - 21.0) < 0.02, p
    assert prob.error[-1] < prob.error[0]
    assert len(prob.out) == 5


@pytest.mark.gpu
def test_gpu_ensemble_inversion_recovers_two_viscosities():
    if DRYRUN:
        from oracle import pyref
        make = pyref.RefEnsemble
    else:
        make = fastiso.Ensemble
    prob = _disc_problem(32, lambda d: [d.R < 1.0e6, d.R >= 1.0e6], [21.0, 20.6], [20.7, 20.9], make)
    prob.solve()
    p, Gx, err = prob.extract_inversion()
    assert np.abs(p - np.array([21.0, 20.6])).max() < 0.03, p
    assert prob.error[-1] < 1e-2 * prob.error[0], prob.error


class _LogTauReduction(inversion.ParameterReduction):
    """log10 of a uniform ELRA relaxation time (the parameter of the reference's ELRA examples)."""

    nparams = 1

    def __init__(self, n=32):
        self.n = n

    def reconstruct(self, params):
        w = workloads.elra(n=self.n, W=3e6, t_end=20e3, tau=10.0 ** float(params[0]), lithosphere="rigid",
                           t_out=[])
        return fastiso.Simulation(w["domain"], w["bcs"], w["sealevel"], w["solidearth"], w["t_span"],
                                  nout=w["nout"], opts=w["opts"], c=w["c"], engine=None).inputs


def _tau_problem(make_batch, truth=np.log10(3000.0), t=(500.0, 1000.0, 2000.0)):
    red = _LogTauReduction()
    ens = make_batch([red.reconstruct([truth])])
    ens.integrator_init()
    Y = []
    for tk in t:
        ens.integrate(tk)
        Y.append(ens[0].get_field("u"))
    close = getattr(ens, "close", None)
    if close:
        close()
    d = fastiso.RegionalDomain(3e6, 3e6, 32, 32)
    data = inversion.InversionData(list(t), Y, d.R < 1.5e6)
    prior = inversion.GaussianPrior([np.log10(2000.0)], [0.3])
    cfg = inversion.InversionConfig("unscented", N_iter=4, alpha_reg=1.0, update_freq=1, scale_obscov=1e-6)
    return inversion.inversion_problem(cfg, data, red, prior, make_batch=make_batch)


def test_oracle_elra_ensemble_inversion_recovers_tau():
    from oracle import pyref
    prob = _tau_problem(pyref.RefEnsemble).solve()
    p, _, _ = prob.extract_inversion()
    assert abs(p[0] - np.log10(3000.0)) < 0.01, p


@pytest.mark.gpu
def test_gpu_elra_ensemble_inversion_recovers_tau():
    if DRYRUN:
        from oracle import pyref
        make = pyref.RefEnsemble
    else:
        make = fastiso.Ensemble
    prob = _tau_problem(make).solve()
    p, _, _ = prob.extract_inversion()
    assert abs(p[0] - np.log10(3000.0)) < 0.01, p
    assert prob.error[-1] < 1e-2 * prob.error[0], prob.error
