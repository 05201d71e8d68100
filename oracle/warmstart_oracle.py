"""FP64 reference of the learned warm-start network (test infrastructure only).

Only the tests may use this module: it checks the GPU forward and backward
kernels.  The network (no reference counterpart; the reference has no learned
component):

    h0      = relu(F We + be)                                F: (N, 32)
    layer l = 0..2:
      agg_h = mean_{j in N(i)} h_j                            (mesh neighbours)
      x     = [h | agg_h | eagg]                              (N, 160)
      pre   = x Wl + bl
      h     = h + relu(pre)
    out     = h Wd + bd                                       (N, 32), first 3 used

eagg (N, 32) holds the mean edge features (first 4 used).  Weights are stored
as (K, NOUT) matrices.  The parameter vector order is
[We, be, W0, b0, W1, b1, W2, b2, Wd, bd].

The warm-start pipeline around it (no reference counterpart either):
  * galerkin_projection: x = V c, c the eigen-truncated solution of
    (V^T A V) c = V^T b over the eigen-directions above 1e-12 of the largest;
  * linesearch_loss_grad: L = 1 - (d.r)^2 / ((d.Ad)(e.r)) and dL/dd, the
    training loss of the network's direction d (r = b - A x_b, e = x* - x_b).

Pinned by tests/test_oracle_warmstart.py: backward vs central finite
differences of forward; the mean aggregation vs a dense adjacency product;
the projection vs the exact solution when it lies in span(V) and vs energy
minimality under perturbation; the line-search gradient vs central finite
differences, L = 0 for d parallel to e and L = 1 for d A-orthogonal to e.
"""

from __future__ import annotations

import numpy as np

FIN, H, XL, NOUT, NL = 32, 64, 160, 32, 3
SHAPES = [("We", (FIN, H)), ("be", (H,))]
for _l in range(NL):
    SHAPES += [(f"W{_l}", (XL, H)), (f"b{_l}", (H,))]
SHAPES += [("Wd", (H, NOUT)), ("bd", (NOUT,))]
NPARAM = sum(int(np.prod(s)) for _, s in SHAPES)


def unpack(theta):
    out, o = {}, 0
    for name, shp in SHAPES:
        k = int(np.prod(shp))
        out[name] = np.asarray(theta[o:o + k], np.float64).reshape(shp)
        o += k
    return out


def pack(p):
    return np.concatenate([np.asarray(p[name], np.float64).reshape(-1) for name, _ in SHAPES])


def init_params(seed=0, scale=1.0):
    rng = np.random.default_rng(seed)
    p = {}
    for name, shp in SHAPES:
        if name.startswith("W"):
            fan_in = shp[0]
            p[name] = rng.normal(0.0, scale * np.sqrt(2.0 / fan_in), shp)
        else:
            p[name] = np.zeros(shp)
    p["Wd"] *= 0.1
    return pack(p)


def mean_agg(h, ro, co):
    n = ro.shape[0] - 1
    deg = np.diff(ro)
    rows = np.repeat(np.arange(n), deg)
    out = np.zeros_like(h)
    np.add.at(out, rows, h[co])
    return out / np.maximum(deg, 1)[:, None]


def mean_agg_T(g, ro, co):
    """Adjoint of mean_agg: out_j = sum_{i: j in N(i)} g_i / deg_i."""
    n = ro.shape[0] - 1
    deg = np.diff(ro)
    rows = np.repeat(np.arange(n), deg)
    out = np.zeros_like(g)
    np.add.at(out, co, (g / np.maximum(deg, 1)[:, None])[rows])
    return out


def forward(theta, F, eagg, ro, co, keep=False):
    p = unpack(theta)
    pre_e = F @ p["We"] + p["be"]
    h = np.maximum(pre_e, 0.0)
    cache = {"F": F, "pre_e": pre_e, "h": [h], "agg": [], "pre": []}
    for l in range(NL):
        agg = mean_agg(h, ro, co)
        x = np.concatenate([h, agg, eagg], 1)
        pre = x @ p[f"W{l}"] + p[f"b{l}"]
        h = h + np.maximum(pre, 0.0)
        cache["agg"].append(agg)
        cache["pre"].append(pre)
        cache["h"].append(h)
    out = h @ p["Wd"] + p["bd"]
    return (out, cache) if keep else out


def backward(theta, cache, dout, eagg, ro, co):
    """Gradient of sum(dout * out) w.r.t. the parameter vector."""
    p = unpack(theta)
    g = {}
    h3 = cache["h"][NL]
    g["Wd"] = h3.T @ dout
    g["bd"] = dout.sum(0)
    dh = dout @ p["Wd"].T
    for l in reversed(range(NL)):
        h, agg, pre = cache["h"][l], cache["agg"][l], cache["pre"][l]
        dpre = dh * (pre > 0)
        x = np.concatenate([h, agg, eagg], 1)
        g[f"W{l}"] = x.T @ dpre
        g[f"b{l}"] = dpre.sum(0)
        dx = dpre @ p[f"W{l}"].T
        dh = dh + dx[:, :H] + mean_agg_T(dx[:, H:2 * H], ro, co)
    dpre_e = dh * (cache["pre_e"] > 0)
    g["We"] = cache["F"].T @ dpre_e
    g["be"] = dpre_e.sum(0)
    return pack(g)


def loss_and_grad(theta, F, eagg, target, valid, ro, co):
    """L = sum over valid rows of |out[:, :3] - target|^2 / n_valid."""
    out, cache = forward(theta, F, eagg, ro, co, keep=True)
    nv = max(int(valid.sum()), 1)
    diff = (out[:, :3] - target) * valid[:, None]
    L = float((diff ** 2).sum() / nv)
    dout = np.zeros_like(out)
    dout[:, :3] = 2.0 * diff / nv
    return L, backward(theta, cache, dout, eagg, ro, co), out


def galerkin_projection(A, b, V, cutoff=1e-12):
    """x = V c minimising x^T A x / 2 - b^T x over span(V) (A SPD, V (n, k)).

    Eigen-truncated: directions of V^T A V below `cutoff` times its largest
    eigenvalue are dropped.  Returns (x, c, kept)."""
    V = np.asarray(V, np.float64)
    AV = A @ V
    G = V.T @ AV
    G = 0.5 * (G + G.T)
    g = V.T @ b
    lam, U = np.linalg.eigh(G)
    keep = lam > cutoff * lam.max()
    c = U[:, keep] @ ((U[:, keep].T @ g) / lam[keep])
    return V @ c, c, int(keep.sum())


def linesearch_loss_grad(d, r, Ad, e_dot_r):
    """L = 1 - (d.r)^2 / ((d.Ad)(e.r)): the share of the A-norm error e of
    the base guess left after x_b + alpha d, alpha = d.r / d.Ad (r = A e).
    Returns (L, dL/dd)."""
    u = float(d @ r)
    w = float(d @ Ad)
    E = float(e_dot_r)
    L = 1.0 - u * u / (w * E)
    grad = -(2.0 * u / (w * E)) * (r - (u / w) * Ad)
    return L, grad
