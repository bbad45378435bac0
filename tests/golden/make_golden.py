"""Write the golden fixtures of tests/test_golden.py from the FP64 oracle (test infrastructure).

    python tests/golden/make_golden.py

Each fixture is a small .npz of expected outputs for inputs the seeded generators of
fastiso.workloads rebuild (their sha256 is stored beside them, so a generator change shows up
as a mismatch, not as a silent new answer):
  green_n16.npz, green_n64.npz  elastic Green table (Farrell table A3, 4-point Gauss-Legendre
                                cells) and the sea-surface Green function of a 2 W = 6000 km grid
  conv_n16.npz, conv_n64.npz    same-size padded convolution with the extended corner BC of a
                                seeded kernel and input, also computed here by direct summation
  rhs_disc_n32.npz, rhs_disc_n64.npz, rhs_lv_n64.npz
                                two right-hand sides at a seeded u (the second one's dudt, the
                                corner-corrected u, and the elastic displacement and sea surface of
                                its sparse update) of the disc benchmark and of a laterally
                                variable case
  analytic_disc.npz             the disc benchmark's analytic displacement at selected r and t
The oracle itself is pinned by tests/test_oracle_numerics.py (closed forms, brute force,
library routines); these files pin its outputs against drift and give the GPU tests a reference
that needs no oracle run.
"""
import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "fastisostasy.jl_amd"))
sys.path.insert(0, os.path.join(ROOT, "tests"))

import fastiso  # noqa: E402
from fastiso import workloads  # noqa: E402
from fastiso.greens import load_farrell_table  # noqa: E402
from oracle import pyref  # noqa: E402

W = 3e6
# two right-hand sides each: the first one's sparse update reads the loads of no earlier call
# (zero), the second one's those of the first, so its elastic displacement and sea surface are
# the padded convolutions of a real load
RHS_T = (0.5, 10.0)        # disc case: during the one-year ice ramp, then with the full disc
RHS_T_LV = (100.0, 110.0)


def sha(*arrays):
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a, dtype=np.float64).tobytes())
    return h.hexdigest()


def seeded_u(d, seed, amp):
    rng = np.random.default_rng(seed)
    return -amp * np.exp(-(d.R / 1e6) ** 2) + 0.05 * rng.standard_normal((d.nx, d.ny))


def conv_inputs(n):
    rng = np.random.default_rng(100 + n)
    d = fastiso.RegionalDomain(W, W, n, n)
    kernel = np.exp(-(d.R / 8e5) ** 2) * (1.0 + 0.1 * rng.standard_normal((n, n)))
    inp = rng.standard_normal((n, n))
    return kernel, inp


def direct_samesize(kernel, inp, x_border=0.0):
    """Full (2n-1)^2 linear convolution by direct summation, extended corner BC, then the crop
    [i1:i2, j1:j2] (src/convolutions.jl:134-154,196-204)."""
    n = kernel.shape[0]
    full = np.zeros((2 * n - 1, 2 * n - 1))
    for i in range(n):
        for j in range(n):
            full[i:i + n, j:j + n] += inp[i, j] * kernel
    full -= (full[0, 0] + full[-1, 0] + full[0, -1] + full[-1, -1]) / 4 - x_border
    m = n // 2
    i1 = m if n % 2 == 0 else m + 1
    i2 = 2 * n - 1 - m
    return full[i1 - 1:i2, i1 - 1:i2]


def disc_case(n):
    return workloads.disc_benchmark(n=n, t_end=50e3, t_out=[])


def lv_case64():
    from _cases import lv_case
    return lv_case(n=64, seed=11)


def rhs_fixture(w, times, seed, amp):
    sim = fastiso.Simulation(w["domain"], w["bcs"], w["sealevel"], w["solidearth"], w["t_span"],
                             nout=w["nout"], opts=w.get("opts"), c=w.get("c"), engine=None,
                             z_b_ref=w.get("z_b_ref"))
    e = pyref.RefEngine.from_inputs(sim.inputs)
    u = seeded_u(w["domain"], seed, amp)
    for t in times:
        dudt, ubc = e.rhs(t, u)
    return dict(u_in=u, dudt=dudt, u_bc=ubc, ue=e.get_field("ue"), dz_ss=e.get_field("dz_ss"),
                t=np.asarray(times, dtype=np.float64), seed=np.int64(seed), amp=np.float64(amp),
                inputs_sha=np.array(sha(*[np.asarray(v) for v in sim.inputs["fields"].values()], u)))


def analytic_fixture():
    w = disc_case(64)
    c, se = w["c"], w["solidearth"]
    r = np.array([0.0, 2.5e5, 5e5, 9e5, 1.1e6, 1.5e6, 2.5e6])
    t = np.array([1e3, 5e3, 1e4, 5e4])
    u = np.array([[pyref.analytic_solution(ri, ti * c.seconds_per_year, c.rho_ice, c.g, 1e3, 1e6,
                                           se.rho_uppermantle, float(se.litho_rigidity.flat[0]),
                                           float(se.effective_viscosity.flat[0])) for ri in r] for ti in t])
    return dict(r=r, t=t, u=u)


def main():
    rm, ge = load_farrell_table()
    for n in (16, 64):
        eg = pyref.elastic_green(n, n, W, W, rm, ge, 4)
        gg = pyref.geoid_green(n, n, W, W)
        np.savez_compressed(os.path.join(HERE, f"green_n{n}.npz"), elastic=eg, geoid=gg, W=W,
                            farrell_sha=np.array(sha(rm, ge)))
        kernel, inp = conv_inputs(n)
        out = pyref.samesize_conv(kernel, inp, 0.0)
        direct = direct_samesize(kernel, inp)
        assert np.abs(out - direct).max() <= 1e-12 * np.abs(direct).max(), n
        np.savez_compressed(os.path.join(HERE, f"conv_n{n}.npz"), out=out,
                            inputs_sha=np.array(sha(kernel, inp)))
    for n in (32, 64):
        np.savez_compressed(os.path.join(HERE, f"rhs_disc_n{n}.npz"), **rhs_fixture(disc_case(n), RHS_T, 7, 30.0))
    np.savez_compressed(os.path.join(HERE, "rhs_lv_n64.npz"), **rhs_fixture(lv_case64(), RHS_T_LV, 8, 80.0))
    np.savez_compressed(os.path.join(HERE, "analytic_disc.npz"), **analytic_fixture())
    for f in sorted(os.listdir(HERE)):
        if f.endswith(".npz"):
            print(f, os.path.getsize(os.path.join(HERE, f)))


if __name__ == "__main__":
    main()
