"""Golden fixtures (tests/golden/, written by tests/golden/make_golden.py from the FP64 oracle).

CPU: the inputs the seeded generators rebuild are the ones the fixtures were made from (sha256);
the oracle reproduces every fixture; the product's host set-up (Green tables in
fastiso.greens, the analytic disc solution in fastiso.analytic) equals them; the convolution
fixture also equals a direct summation.
GPU: the fixtures' two right-hand sides on the HIP engine (disc benchmark at n = 32 and 64,
laterally variable case at n = 64) against the stored dudt, corner-corrected u, elastic
displacement and sea surface of the second, at the tolerances of the other RHS parity tests.
"""
import os
import sys

import numpy as np
import pytest

import fastiso
from fastiso.analytic import analytic_solution
from fastiso.greens import elastic_green, geoid_green, load_farrell_table

from _cases import hip_sim

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
sys.path.insert(0, GOLD)
import make_golden as mg  # noqa: E402

RHS_MAX = 2e-6


def load(name):
    with np.load(os.path.join(GOLD, name), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def rhs_case(name):
    if name == "rhs_lv_n64.npz":
        return mg.lv_case64(), mg.RHS_T_LV
    return mg.disc_case(32 if "n32" in name else 64), mg.RHS_T


RHS_FILES = ["rhs_disc_n32.npz", "rhs_disc_n64.npz", "rhs_lv_n64.npz"]


@pytest.mark.parametrize("n", [16, 64])
def test_green_tables(n):
    from oracle import pyref
    g = load(f"green_n{n}.npz")
    rm, ge = load_farrell_table()
    assert str(g["farrell_sha"]) == mg.sha(rm, ge)
    d = fastiso.RegionalDomain(mg.W, mg.W, n, n)
    for got in (pyref.elastic_green(n, n, mg.W, mg.W, rm, ge, 4), elastic_green(d)):
        np.testing.assert_allclose(got, g["elastic"], rtol=1e-13, atol=0)
    for got in (pyref.geoid_green(n, n, mg.W, mg.W), geoid_green(d, fastiso.PhysicalConstants())):
        np.testing.assert_allclose(got, g["geoid"], rtol=1e-13, atol=0)


@pytest.mark.parametrize("n", [16, 64])
def test_padded_convolution(n):
    from oracle import pyref
    g = load(f"conv_n{n}.npz")
    kernel, inp = mg.conv_inputs(n)
    assert str(g["inputs_sha"]) == mg.sha(kernel, inp)
    np.testing.assert_allclose(pyref.samesize_conv(kernel, inp, 0.0), g["out"], rtol=0,
                               atol=1e-12 * np.abs(g["out"]).max())
    if n == 16:
        np.testing.assert_allclose(mg.direct_samesize(kernel, inp), g["out"], rtol=0,
                                   atol=1e-12 * np.abs(g["out"]).max())


@pytest.mark.parametrize("name", RHS_FILES)
def test_oracle_rhs(name):
    g = load(name)
    w, t = rhs_case(name)
    f = mg.rhs_fixture(w, t, int(g["seed"]), float(g["amp"]))
    assert str(f["inputs_sha"]) == str(g["inputs_sha"])
    for k in ("dudt", "u_bc", "ue", "dz_ss"):
        scale = max(np.abs(g[k]).max(), 1e-30)
        np.testing.assert_allclose(f[k], g[k], rtol=0, atol=1e-10 * scale, err_msg=k)
    assert np.abs(g["dudt"]).max() > 0 and np.abs(g["ue"]).max() > 0


def test_analytic_disc():
    from oracle import pyref
    g = load("analytic_disc.npz")
    w = mg.disc_case(64)
    c, se = w["c"], w["solidearth"]
    for i, t in enumerate(g["t"]):
        u = analytic_solution(g["r"], t * c.seconds_per_year, c, se, 1e3, 1e6)
        np.testing.assert_allclose(u, g["u"][i], rtol=1e-9, atol=1e-9)
        for j, r in enumerate(g["r"]):
            v = pyref.analytic_solution(r, t * c.seconds_per_year, c.rho_ice, c.g, 1e3, 1e6, se.rho_uppermantle,
                                        float(se.litho_rigidity.flat[0]), float(se.effective_viscosity.flat[0]))
            assert v == g["u"][i, j]
    # the load depresses the centre and the depression deepens with time
    assert np.all(g["u"][:, 0] < 0) and np.all(np.diff(g["u"][:, 0]) < 0)


@pytest.mark.gpu
@pytest.mark.parametrize("name", RHS_FILES)
def test_gpu_rhs_against_golden(name):
    g = load(name)
    w, t = rhs_case(name)
    sim = hip_sim(w)
    for tk in t:
        dudt, ubc = sim.engine.rhs(tk, g["u_in"])
    m = np.abs(dudt - g["dudt"]).max() / np.abs(g["dudt"]).max()
    assert m < RHS_MAX, (name, m)
    np.testing.assert_allclose(ubc, g["u_bc"], rtol=0, atol=1e-9 * np.abs(g["u_bc"]).max())
    for k in ("ue", "dz_ss"):
        ref = g[k]
        got = sim.engine.get_field(k)
        if np.abs(ref).max() == 0:
            assert np.abs(got).max() == 0, k
        else:
            # FP32 padded transforms of a seeded, noisy load: 1e-5 of the field's maximum
            assert np.abs(got - ref).max() <= 1e-5 * np.abs(ref).max(), (name, k)
    sim.close()
