"""CPU oracle of the learned guidance network (test infrastructure only).

Only the tests, the smoke check and the benchmark's CPU baseline use this
module, as the checker of the GPU network in tessera_amd/csrc/gnn.hip; the
product path never calls it.

The network has no upstream kernel counterpart. Its structure follows the
template of python/gnn_guide.py:81-103 (IsingGNN: node embedding,
edge-conditioned messages summed at the destination, node update, per-node
readout), with this library's own layer algebra:

    x_i   = (h_i, deg_i / 3)
    H0    = relu(x W_in + b_in)
    layer: [P Q] = H W_sd;  a_uv = relu(P_u + Q_v + J_uv w_J + b_m);
           M_v = sum_{u in adj(v)} a_uv;  Z = [H M] W_u + b_u;  H <- H + relu(Z)
    z_i   = tanh(H_i . w_z + c_z)
    s_g   = sigmoid(mean_{i in g} H_i . w_s + c_s)
    loss  = mean_i (z_i - tz_i)^2 + mean_g (s_g - ts_g)^2

Everything is FP64. With bf16=True the operands of the four dense products
(H, [H M], dZ, dPQ and the weights) are rounded to BF16 (round to nearest
even) as the MFMA path rounds them, so the GPU can be compared tightly.

Parity: the analytic gradient of `loss_and_grad` is pinned by central finite
differences of `forward` (tests/test_gnn_oracle.py); Adam by its closed form
on a one-parameter quadratic.
"""
from __future__ import annotations

import numpy as np

F = 64
L = 3
W_IN, B_IN = 0, 2 * F
LAYER0 = B_IN + F
W_SD, W_J, B_M = 0, F * 2 * F, F * 2 * F + F
W_U = B_M + F
B_U = W_U + 2 * F * F
LAYER_SIZE = B_U + F
HEAD = LAYER0 + L * LAYER_SIZE
W_Z, C_Z = HEAD, HEAD + F
W_S, C_S = C_Z + 1, C_Z + 1 + F
COUNT = C_S + 1


def layer_base(l: int) -> int:
    return LAYER0 + l * LAYER_SIZE


def bf16(x: np.ndarray) -> np.ndarray:
    """Round to BF16 (nearest even) and return as float64."""
    u = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(w: np.ndarray):
    w = np.asarray(w, np.float64)
    p = dict(W_in=w[W_IN:W_IN + 2 * F].reshape(2, F), b_in=w[B_IN:B_IN + F], layers=[],
             w_z=w[W_Z:W_Z + F], c_z=w[C_Z], w_s=w[W_S:W_S + F], c_s=w[C_S])
    for l in range(L):
        b = layer_base(l)
        p["layers"].append(dict(W_sd=w[b + W_SD:b + W_SD + 2 * F * F].reshape(F, 2 * F),
                                w_J=w[b + W_J:b + W_J + F], b_m=w[b + B_M:b + B_M + F],
                                W_u=w[b + W_U:b + W_U + 2 * F * F].reshape(2 * F, F),
                                b_u=w[b + B_U:b + B_U + F]))
    return p


def _edges(batch):
    """Directed message list (src u, dst v, J) from the symmetric adjacency."""
    adj_off = batch["adj_off"]
    n = len(adj_off) - 1
    dst = np.repeat(np.arange(n), np.diff(adj_off))
    src = np.asarray(batch["adj_nbr"], np.int64)
    return src, dst, np.asarray(batch["adj_j"], np.float64)


def forward(w: np.ndarray, batch: dict, bf: bool = False, cache: dict | None = None):
    """Per-node z and per-graph s of a packed batch (see tessera_amd.learned.pack)."""
    r = bf16 if bf else (lambda x: np.asarray(x, np.float64))
    p = unpack(w)
    x = np.asarray(batch["feat"], np.float64).reshape(-1, 2)
    n = x.shape[0]
    src, dst, J = _edges(batch)
    pre = x @ p["W_in"] + p["b_in"]
    H = np.maximum(pre, 0.0)
    c = dict(pre=pre, H=[H], X=[], PQ=[], Z=[])
    for lp in p["layers"]:
        Hb = r(H)
        PQ = Hb @ r(lp["W_sd"])
        P, Q = PQ[:, :F], PQ[:, F:]
        A = np.maximum(P[src] + Q[dst] + J[:, None] * lp["w_J"] + lp["b_m"], 0.0)
        M = np.zeros((n, F))
        np.add.at(M, dst, A)
        X = np.concatenate([Hb, r(M)], axis=1)
        Z = X @ r(lp["W_u"]) + lp["b_u"]
        H = H + np.maximum(Z, 0.0)
        c["X"].append(X)
        c["PQ"].append(PQ)
        c["Z"].append(Z)
        c["H"].append(H)
    z = np.tanh(H @ p["w_z"] + p["c_z"])
    off = np.asarray(batch["node_off"], np.int64)
    pool = np.stack([H[off[g]:off[g + 1]].mean(axis=0) for g in range(len(off) - 1)])
    s = 1.0 / (1.0 + np.exp(-(pool @ p["w_s"] + p["c_s"])))
    c.update(pool=pool, z=z, s=s, src=src, dst=dst, J=J, off=off, x=x)
    if cache is not None:
        cache.update(c)
    return z, s


def loss_of(w, batch, bf=False) -> float:
    z, s = forward(w, batch, bf)
    return float(np.mean((z - batch["tz"]) ** 2) + np.mean((s - batch["ts"]) ** 2))


def loss_and_grad(w: np.ndarray, batch: dict, bf: bool = False, plant: str | None = None):
    """Loss and its gradient with respect to the flat parameter vector.

    `plant` injects a known mistake for the tests that show the gradient check
    catches it: "drop_residual" (dH_l misses the residual dH_l+1 term) or
    "drop_source" (the message gradient misses its source-node side dP)."""
    r = bf16 if bf else (lambda x: np.asarray(x, np.float64))
    p = unpack(w)
    c = {}
    z, s = forward(w, batch, bf, c)
    tz = np.asarray(batch["tz"], np.float64)
    ts = np.asarray(batch["ts"], np.float64)
    n, B = len(z), len(s)
    loss = float(np.mean((z - tz) ** 2) + np.mean((s - ts) ** 2))
    g = np.zeros(COUNT)
    H3 = c["H"][-1]
    dzp = 2.0 * (z - tz) / n * (1.0 - z * z)
    g[W_Z:W_Z + F] = dzp @ H3
    g[C_Z] = dzp.sum()
    G = np.outer(dzp, p["w_z"])
    dsp = 2.0 * (s - ts) / B * s * (1.0 - s)
    g[W_S:W_S + F] = dsp @ c["pool"]
    g[C_S] = dsp.sum()
    off = c["off"]
    for k in range(B):
        G[off[k]:off[k + 1]] += dsp[k] * p["w_s"] / (off[k + 1] - off[k])
    src, dst, J = c["src"], c["dst"], c["J"]
    for l in reversed(range(L)):
        lp = p["layers"][l]
        b = layer_base(l)
        X, PQ, Z = c["X"][l], c["PQ"][l], c["Z"][l]
        dZ = G * (Z > 0)
        g[b + B_U:b + B_U + F] = dZ.sum(axis=0)
        dZb = r(dZ)
        g[b + W_U:b + W_U + 2 * F * F] = (X.T @ dZb).ravel()
        dX = dZb @ r(lp["W_u"]).T
        dM = dX[:, F:]
        pre = PQ[src, :F] + PQ[dst, F:] + J[:, None] * lp["w_J"] + lp["b_m"]
        da = dM[dst] * (pre > 0)
        dP = np.zeros((n, F))
        dQ = np.zeros((n, F))
        if plant != "drop_source":
            np.add.at(dP, src, da)
        np.add.at(dQ, dst, da)
        g[b + W_J:b + W_J + F] = (da * J[:, None]).sum(axis=0)
        g[b + B_M:b + B_M + F] = da.sum(axis=0)
        dPQb = r(np.concatenate([dP, dQ], axis=1))
        g[b + W_SD:b + W_SD + 2 * F * F] = (X[:, :F].T @ dPQb).ravel()
        G = (0.0 if plant == "drop_residual" else G) + dX[:, :F] + dPQb @ r(lp["W_sd"]).T
    dpre = G * (c["pre"] > 0)
    x = c["x"]
    g[W_IN:W_IN + F] = x[:, 0] @ dpre
    g[W_IN + F:W_IN + 2 * F] = x[:, 1] @ dpre
    g[B_IN:B_IN + F] = dpre.sum(axis=0)
    return loss, g


def adam(w, g, m, v, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step (t counts from 1); returns new (w, m, v)."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    w = w - lr * (m / (1 - b1 ** t)) / (np.sqrt(v / (1 - b2 ** t)) + eps)
    return w, m, v
