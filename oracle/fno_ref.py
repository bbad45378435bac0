"""ORACLE -- test infrastructure only (tests/ import it; nothing in the package does).

FP64 numpy restatement of the surrogate network of include/fastiso_surrogate.h: forward pass,
loss, and the gradient of the loss with respect to the flat parameter vector by explicit
backpropagation. The surrogate has no counterpart in the reference, so it is pinned by
mathematics instead of reference vectors:
  * spectral_conv_fft() computes the truncated spectral convolution a second way, with numpy's
    rfft2/irfft2 (a library routine) instead of the DFT-matrix products used here and on the
    device;
  * the gradient is checked against central finite differences of loss() in the tests.
Layout: fields are (nx, ny) arrays (x first); the flat cell index is i + j nx as on the device.
"""
import numpy as np
from scipy.special import erf


def default_config():
    return dict(in_channels=4, width=32, modes_x=16, modes_y=16, layers=4, proj_hidden=64)


def param_layout(cfg):
    Ci, C, mx, my, L, H = (cfg[k] for k in ("in_channels", "width", "modes_x", "modes_y", "layers", "proj_hidden"))
    shapes = [("P", (Ci, C)), ("p", (C,))]
    for l in range(L):
        shapes += [(f"W{l}", (C, C)), (f"b{l}", (C,)), (f"R{l}", (mx, 2 * my, C, C, 2))]
    shapes += [("Q1", (C, H)), ("q1", (H,)), ("Q2", (H,)), ("q2", (1,))]
    out, o = {}, 0
    for name, shp in shapes:
        n = int(np.prod(shp))
        out[name] = (o, shp)
        o += n
    return out, o


def unpack(theta, cfg):
    lay, n = param_layout(cfg)
    assert theta.size == n
    return {k: theta[o:o + int(np.prod(s))].reshape(s) for k, (o, s) in lay.items()}


def kept_ky(ny, my):
    return np.array(list(range(my)) + list(range(ny - my, ny)))


def dft_mats(nx, ny, cfg):
    """The DFT matrices of the truncated transform (same definitions as the device)."""
    mx, my = cfg["modes_x"], cfg["modes_y"]
    i = np.arange(nx)
    k = np.arange(mx)
    thx = 2 * np.pi * np.outer(k, i) / nx                   # (mx, nx)
    Ex = np.empty((2 * mx, nx))
    Ex[0::2], Ex[1::2] = np.cos(thx), -np.sin(thx)
    w = np.where(k == 0, 1.0, 2.0) / (nx * ny)
    Bx = np.empty((nx, 2 * mx))
    Bx[:, 0::2], Bx[:, 1::2] = (w * np.cos(thx).T), (-w * np.sin(thx).T)
    ky = kept_ky(ny, my)
    thy = 2 * np.pi * np.outer(ky, np.arange(ny)) / ny       # (2my, ny)
    Ar = np.empty((4 * my, ny))
    Ai = np.empty((4 * my, ny))
    Ar[0::2], Ar[1::2] = np.cos(thy), -np.sin(thy)
    Ai[0::2], Ai[1::2] = np.sin(thy), np.cos(thy)
    return Ex, Ar, Ai, Bx


def gelu(z):
    return 0.5 * z * (1 + erf(z / np.sqrt(2)))


def gelu_d(z):
    return 0.5 * (1 + erf(z / np.sqrt(2))) + z * np.exp(-0.5 * z * z) / np.sqrt(2 * np.pi)


# Activations are kept as (P, C) arrays with P = nx ny cells in the order i + j nx, i.e. the
# array V[j, i, c] reshaped; spectral quantities as in the device comments.
def _spec_fwd(v, R, M, nx, ny, C):
    Ex, Ar, Ai, Bx = M
    V = v.reshape(ny, nx, C)
    Fx = np.einsum("qi,jic->jqc", Ex, V)                     # (ny, 2mx, C)
    Fr, Fi = Fx[:, 0::2, :], Fx[:, 1::2, :]                  # (ny, mx, C)
    G = np.einsum("rj,jkc->krc", Ar, Fr) + np.einsum("rj,jkc->krc", Ai, Fi)   # (mx, 4my, C)
    Gr, Gi = G[:, 0::2, :], G[:, 1::2, :]                    # (mx, 2my, C)
    Rr, Ri = R[..., 0], R[..., 1]                            # (mx, 2my, C, C)
    Hr = np.einsum("kmc,kmco->kmo", Gr, Rr) - np.einsum("kmc,kmco->kmo", Gi, Ri)
    Hi = np.einsum("kmc,kmco->kmo", Gr, Ri) + np.einsum("kmc,kmco->kmo", Gi, Rr)
    H = np.empty_like(G)
    H[:, 0::2, :], H[:, 1::2, :] = Hr, Hi
    Fyr = np.einsum("rj,krc->jkc", Ar, H)
    Fyi = np.einsum("rj,krc->jkc", Ai, H)
    Fy = np.empty_like(Fx)
    Fy[:, 0::2, :], Fy[:, 1::2, :] = Fyr, Fyi
    S = np.einsum("iq,jqc->jic", Bx, Fy)
    return S.reshape(nx * ny, C), (G, Gr, Gi)


def spectral_conv_fft(v, R, nx, ny, cfg):
    """The same truncated spectral convolution through numpy's rfft2 / irfft2."""
    mx, my, C = cfg["modes_x"], cfg["modes_y"], cfg["width"]
    V = v.reshape(ny, nx, C)
    Z = np.fft.rfft2(V, axes=(0, 1))                         # (ny, nx//2+1, C): ky, kx
    ky = kept_ky(ny, my)
    Rc = R[..., 0] + 1j * R[..., 1]                          # (mx, 2my, C, C)
    Out = np.zeros_like(Z)
    for m, k_y in enumerate(ky):
        Out[k_y, :mx, :] = np.einsum("kc,kco->ko", Z[k_y, :mx, :], Rc[:, m])
    return np.fft.irfft2(Out, s=(ny, nx), axes=(0, 1)).reshape(nx * ny, C)


def forward(theta, x, nx, ny, cfg, keep=False):
    """x: (P, in_channels) cell-major; returns y (P,) (and the tape when keep)."""
    p = unpack(theta, cfg)
    L, C = cfg["layers"], cfg["width"]
    M = dft_mats(nx, ny, cfg)
    v = x @ p["P"] + p["p"]
    tape = {"v": [v], "z": [], "G": []}
    for l in range(L):
        S, (G, _, _) = _spec_fwd(v, p[f"R{l}"], M, nx, ny, C)
        z = v @ p[f"W{l}"] + p[f"b{l}"] + S
        v = gelu(z) if l < L - 1 else z
        tape["z"].append(z)
        tape["G"].append(G)
        tape["v"].append(v)
    qp = v @ p["Q1"] + p["q1"]
    q = gelu(qp)
    y = q @ p["Q2"] + p["q2"][0]
    if keep:
        tape.update(qp=qp, q=q, M=M, p=p)
        return y, tape
    return y


def loss(theta, x, t, nx, ny, cfg):
    y = forward(theta, x, nx, ny, cfg)
    return float(np.mean((y - t) ** 2))


def loss_grad(theta, x, t, nx, ny, cfg):
    """Loss 1/P sum (y - t)^2 and its gradient (flat, same layout as theta)."""
    y, tp = forward(theta, x, nx, ny, cfg, keep=True)
    p, (Ex, Ar, Ai, Bx) = tp["p"], tp["M"]
    L, C, mx, my = cfg["layers"], cfg["width"], cfg["modes_x"], cfg["modes_y"]
    P = nx * ny
    lay, n = param_layout(cfg)
    g = {k: np.zeros(s) for k, (o, s) in lay.items()}
    dy = 2.0 * (y - t) / P
    g["Q2"] = tp["q"].T @ dy
    g["q2"] = np.array([dy.sum()])
    dq = np.outer(dy, p["Q2"]) * gelu_d(tp["qp"])
    g["Q1"] = tp["v"][L].T @ dq
    g["q1"] = dq.sum(0)
    dv = dq @ p["Q1"].T
    for l in range(L - 1, -1, -1):
        dz = dv if l == L - 1 else dv * gelu_d(tp["z"][l])
        vl = tp["v"][l]
        g[f"b{l}"] = dz.sum(0)
        g[f"W{l}"] = vl.T @ dz
        dv = dz @ p[f"W{l}"].T
        # spectral path, reversed
        dZ = dz.reshape(ny, nx, C)
        dFy = np.einsum("iq,jic->jqc", Bx, dZ)               # (ny, 2mx, C)
        dFyr, dFyi = dFy[:, 0::2, :], dFy[:, 1::2, :]
        dH = np.einsum("rj,jkc->krc", Ar, dFyr) + np.einsum("rj,jkc->krc", Ai, dFyi)
        dHr, dHi = dH[:, 0::2, :], dH[:, 1::2, :]
        G = tp["G"][l]
        Gr, Gi = G[:, 0::2, :], G[:, 1::2, :]
        R = p[f"R{l}"]
        Rr, Ri = R[..., 0], R[..., 1]
        dR = np.empty_like(R)
        dR[..., 0] = np.einsum("kmo,kmc->kmco", dHr, Gr) + np.einsum("kmo,kmc->kmco", dHi, Gi)
        dR[..., 1] = -np.einsum("kmo,kmc->kmco", dHr, Gi) + np.einsum("kmo,kmc->kmco", dHi, Gr)
        g[f"R{l}"] = dR
        dGr = np.einsum("kmo,kmco->kmc", dHr, Rr) + np.einsum("kmo,kmco->kmc", dHi, Ri)
        dGi = -np.einsum("kmo,kmco->kmc", dHr, Ri) + np.einsum("kmo,kmco->kmc", dHi, Rr)
        dG = np.empty_like(G)
        dG[:, 0::2, :], dG[:, 1::2, :] = dGr, dGi
        dFr = np.einsum("rj,krc->jkc", Ar, dG)
        dFi = np.einsum("rj,krc->jkc", Ai, dG)
        dFx = np.empty((ny, 2 * mx, C))
        dFx[:, 0::2, :], dFx[:, 1::2, :] = dFr, dFi
        dv = dv + np.einsum("qi,jqc->jic", Ex, dFx).reshape(P, C)
    g["p"] = dv.sum(0)
    g["P"] = x.T @ dv
    flat = np.zeros(n)
    for k, (o, s) in lay.items():
        flat[o:o + int(np.prod(s))] = np.asarray(g[k]).ravel()
    return float(np.mean((y - t) ** 2)), flat


def init_params(cfg, seed):
    """A seeded parameter vector with the device's initialisation ranges (values differ)."""
    rng = np.random.default_rng(seed)
    lay, n = param_layout(cfg)
    th = np.zeros(n)
    C, Ci, H = cfg["width"], cfg["in_channels"], cfg["proj_hidden"]
    fan = {"P": Ci, "p": Ci, "Q1": C, "q1": C, "Q2": H, "q2": H}
    for k, (o, s) in lay.items():
        m = int(np.prod(s))
        if k.startswith("R"):
            th[o:o + m] = rng.uniform(0, 1, m) / (C * C)
        else:
            a = 1 / np.sqrt(fan.get(k, C))
            th[o:o + m] = rng.uniform(-a, a, m)
    return th
