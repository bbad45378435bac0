"""Extract the Farrell (1972, table A3) elastic load-response table that the reference ships as
src/input/elasticgreencoeffs_farrell1972.jld2 into elasticgreencoeffs_farrell1972.csv.

The JLD2 file holds three little-endian float64 arrays of 42 entries: the angular distance
(deg) at byte 622, the distance rm (m) at byte 1024 and the coefficient GE at byte 1426.
They are the columns the reference reads from its remote text copy of the same table
(src/dataloaders.jl:257-273: rm = column 2, GE = column 3). Plain data, read as bytes.

Run once: python make_farrell_csv.py /path/to/reference/src/input/elasticgreencoeffs_farrell1972.jld2
"""
import os
import sys

import numpy as np


def main(path):
    raw = open(path, "rb").read()
    cols = [np.frombuffer(raw[o:o + 42 * 8], dtype="<f8") for o in (622, 1024, 1426)]
    theta, rm, ge = cols
    assert rm[0] == 0.0 and np.all(np.diff(rm) > 0) and np.all(np.diff(theta) > 0)
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "elasticgreencoeffs_farrell1972.csv")
    with open(out, "w") as f:
        f.write("theta_deg,rm_m,GE\n")
        for a, b, c in zip(theta, rm, ge):
            f.write(f"{float(a)!r},{float(b)!r},{float(c)!r}\n")
    print("wrote", out)


if __name__ == "__main__":
    main(sys.argv[1])
