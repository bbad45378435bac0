"""Train the LV-ELVA macro-step surrogate on plain-solver trajectories.

    python -m fastiso.train_surrogate --out fno.npz [--base rkc --members 12 --steps-per-member 8 ...]

1. Training pairs: C1-style members (seeds from --seed) integrated by the ensemble engine, macro
   step by macro step, from random phases of the glacial cycle; the network's target is the
   plain increment minus the base predictor's (RKC macro step, generate_rkc_pairs; or the
   extrapolated increment, generate_pairs).
2. Training: Adam on the device (fi_surrogate_train) on the scaled corrections.
3. Evaluation on held-out members: max and RMS error of the increment with the base predictor
   alone and with the network, printed with the weights' path.
"""
import argparse
import json
import sys
import time

import numpy as np

from . import surrogate as S


def evaluate(sur, X, Y, scale=S.D_SCALE):
    """Errors (m) of the base predictor alone (extrapolation or RKC) and of base + network on
    scaled pairs (X, Y)."""
    out = []
    for x, y in zip(X, Y):
        r = sur.forward(x) if sur is not None else 0.0 * y
        e_net = scale * (y - r)
        e_base = scale * y
        out.append((float(np.abs(e_base).max()), float(np.sqrt(np.mean(e_base ** 2))),
                    float(np.abs(e_net).max()), float(np.sqrt(np.mean(e_net ** 2)))))
    return np.array(out)


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", required=True)
    ap.add_argument("--n", type=int, default=350)
    ap.add_argument("--members", type=int, default=64)
    ap.add_argument("--steps-per-member", type=int, default=5)
    ap.add_argument("--val-members", type=int, default=8)
    ap.add_argument("--seed", type=int, default=10000)
    ap.add_argument("--macro-dt", type=float, default=500.0)
    ap.add_argument("--train-steps", type=int, default=2000)
    ap.add_argument("--batch", type=int, default=8)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--width", type=int, default=32)
    ap.add_argument("--modes", type=int, default=16)
    ap.add_argument("--layers", type=int, default=4)
    ap.add_argument("--wd", type=float, default=1e-4)
    ap.add_argument("--skip", type=int, default=2)
    ap.add_argument("--base", choices=["extrapolation", "rkc"], default="rkc",
                    help="the predictor the network corrects")
    ap.add_argument("--rkc-dt", type=float, default=250.0)
    a = ap.parse_args(argv)
    scale = S.R_SCALE if a.base == "rkc" else S.D_SCALE

    def log(*msg):
        print(*msg, file=sys.stderr, flush=True)

    t0 = time.perf_counter()
    def pairs(members, seed, label):
        prog = lambda s, i: log(f"{label} pairs: macro step {s + 1} done ({time.perf_counter() - t0:.0f} s)")
        if a.base == "rkc":
            return S.generate_rkc_pairs(members, seed, a.steps_per_member, a.macro_dt, n=a.n, rkc_dt=a.rkc_dt,
                                        progress=prog)
        return S.generate_pairs(members, seed, a.steps_per_member, a.macro_dt, n=a.n, skip=a.skip, progress=prog)

    X, Y, meta = pairs(a.members, a.seed, "train")
    Xv, Yv, meta_v = pairs(a.val_members, a.seed + 5000, "validation")
    gen_s = time.perf_counter() - t0
    log(f"{len(X)} training and {len(Xv)} validation pairs in {gen_s:.0f} s")
    sur = S.Surrogate(width=a.width, modes_x=a.modes, modes_y=a.modes, layers=a.layers, seed=a.seed,
                      in_channels=X.shape[2])
    sur.set_grid(a.n, a.n)
    Xa, Ya = S.augment(X, Y, a.n, a.n)
    log(f"{len(Xa)} training samples after the 8 grid symmetries")
    sur.set_data(Xa, Ya)
    del Xa, Ya
    base_v = evaluate(None, Xv, Yv, scale)
    t1 = time.perf_counter()
    losses = []
    best = None
    chunk = 200
    done = 0
    while done < a.train_steps:
        k = min(chunk, a.train_steps - done)
        # one cosine schedule over all steps, run in chunks so progress is printed
        frac0, frac1 = done / a.train_steps, (done + k) / a.train_steps
        lr0 = 1e-5 + 0.5 * (a.lr - 1e-5) * (1 + np.cos(np.pi * frac0))
        lr1 = 1e-5 + 0.5 * (a.lr - 1e-5) * (1 + np.cos(np.pi * frac1))
        losses += list(sur.train(k, batch=a.batch, lr=lr0, lr_min=lr1, seed=a.seed + done, wd=a.wd))
        done += k
        ev = evaluate(sur, Xv, Yv, scale)
        score = float(np.mean(ev[:, 3]))   # mean over held-out pairs of the RMS error
        if best is None or score < best[0]:
            best = (score, done, sur.get_params())
        log(f"step {done}: train loss {np.mean(losses[-k:]):.4g}  val max err {ev[:, 2].max():.4f} m "
            f"({a.base} {base_v[:, 0].max():.4f})  val rms {ev[:, 3].mean():.5f} m ({a.base} {base_v[:, 1].mean():.5f})"
            f"  {time.perf_counter() - t1:.0f} s")
    train_s = time.perf_counter() - t1
    sur.set_params(best[2])   # early stopping on the held-out members
    log(f"kept the parameters of step {best[1]}")
    ev = evaluate(sur, Xv, Yv, scale)
    sur.scales = (S.U_SCALE, S.S_SCALE, scale, a.macro_dt, 1.0 if a.base == "rkc" else 0.0)
    sur.save(a.out)
    summary = {
        "base": a.base, "pairs": len(X), "val_pairs": len(Xv), "n": a.n, "macro_dt": a.macro_dt, "generation_s": gen_s,
        "train_steps": a.train_steps, "batch": a.batch, "train_s": train_s,
        "surrogate_device_ms": sur.device_ms(),
        "val_base_max_m": base_v[:, 0].tolist(), "val_base_rms_m": base_v[:, 1].tolist(),
        "val_network_max_m": ev[:, 2].tolist(), "val_network_rms_m": ev[:, 3].tolist(),
        "val_increment_max_m": [m[2] for m in meta_v], "best_step": best[1],
        "val_pass_1m_base": int(np.sum(base_v[:, 0] <= 1.0)), "val_pass_1m_network": int(np.sum(ev[:, 2] <= 1.0)),
        "loss_first": float(np.mean(losses[:20])), "loss_last": float(np.mean(losses[-20:])),
    }
    print(json.dumps(summary), flush=True)


if __name__ == "__main__":
    main()
