"""What the trained surrogate network gains over the method without it.

Without the network the macro step is extrapolated in time (du = du_prev); the network adds a
learned correction. On plain-solver pairs of two members whose seeds were not used in training
(generated here by the ensemble engine at N=350), the corrected increment must be closer to the
plain solver than the extrapolated one, in the mean over pairs of the RMS error.
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
WEIGHTS = os.path.join(ROOT, "fastisostasy.jl_amd", "fastiso", "data", "fno_lv_elva_n350.npz")


def test_network_beats_extrapolation_on_held_out_members():
    from fastiso import surrogate as S
    X, Y, meta = S.generate_pairs(2, 90000, 4, 500.0, n=350, skip=2)
    d = np.load(WEIGHTS, allow_pickle=False)
    c = [int(v) for v in d["cfg"]]
    sur = S.Surrogate(in_channels=c[0], width=c[1], modes_x=c[2], modes_y=c[3], layers=c[4], proj_hidden=c[5])
    sur.set_grid(350, 350)
    sur.set_params(d["theta"])
    rms_x, rms_n = [], []
    for x, y in zip(X, Y):
        r = sur.forward(x)
        rms_x.append(S.D_SCALE * np.sqrt(np.mean(y.astype(np.float64) ** 2)))
        rms_n.append(S.D_SCALE * np.sqrt(np.mean((y - r).astype(np.float64) ** 2)))
    print("rms extrapolation", np.round(rms_x, 4), "rms network", np.round(rms_n, 4))
    assert np.mean(rms_n) < np.mean(rms_x), (rms_x, rms_n)
