"""numpy oracle of the learned message-passing preconditioner -- TEST
INFRASTRUCTURE ONLY (tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg may use it, only as the checker).

The network is the reference's message-passing layer
(``zoomy_jax/gnn_blueprint/graph_mp_multibranch.py:9-16``:
h' = act(h Ws + sum_{j->i} h_j Wn + b)) made linear in the residual so that
it can sit inside a Krylov method: no bias, and the activation replaced by a
static gate that depends only on operator features,

    h_{l+1} = sigmoid(phi U_l) * ([h_l | sum_{j in N(i)} h_j] W_l),

on the graph of 2x2 cell blocks of the fine grid (node (a, b) = cells
(2a + di, 2b + dj), slot s = 2 di + dj; 4-neighbour adjacency, zero outside
the grid).  Input h_0 = (r / diag) W_in (4 -> 64), output z = h_3 W_out
(64 -> 4).  phi (16 per node): cE/diag, cN/diag and (sum of |off-diagonal|)
/diag of the four cells, then 1, then three zeros.

``bf16=True`` rounds to bfloat16 (round to nearest even) exactly where the
device kernels store or feed bfloat16: the stored h_l, the aggregated
neighbour sum, phi, the layer weights W_l and U_l, and the backward operands
dy and the gate-gradient.  Parameters are kept in FP32 ("master weights").
"""
from __future__ import annotations

import numpy as np

HID = 64
NL = 3
NF = 16

# flat parameter order (FP32): Win (4, 64) | W_l (128, 64), U_l (16, 64) for
# l = 0, 1, 2 | Wout (64, 4)
SHAPES = [("Win", (4, HID))]
for _l in range(NL):
    SHAPES += [(f"W{_l}", (2 * HID, HID)), (f"U{_l}", (NF, HID))]
SHAPES += [("Wout", (HID, 4))]
N_PARAMS = sum(int(np.prod(s)) for _, s in SHAPES)


def unpack(flat):
    out, o = {}, 0
    for name, s in SHAPES:
        k = int(np.prod(s))
        out[name] = np.asarray(flat[o:o + k], np.float64).reshape(s)
        o += k
    return out


def pack(p):
    return np.concatenate([np.asarray(p[n], np.float64).ravel()
                           for n, _ in SHAPES]).astype(np.float32)


def init_params(seed=0, zero_out=True):
    """Glorot-like init; W_out = 0 makes the network start as the zero map."""
    rng = np.random.default_rng(seed)
    p = {"Win": rng.standard_normal((4, HID)) * 0.5}
    for l in range(NL):
        p[f"W{l}"] = rng.standard_normal((2 * HID, HID)) / np.sqrt(2 * HID)
        p[f"U{l}"] = rng.standard_normal((NF, HID)) * 0.1
    p["Wout"] = np.zeros((HID, 4)) if zero_out else \
        rng.standard_normal((HID, 4)) * 0.05
    return pack(p)


def bf16(x):
    """Round to bfloat16 (nearest even), returned as float64."""
    a = np.ascontiguousarray(np.asarray(x, np.float32))
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000).astype(np.uint32)
    return r.view(np.float32).astype(np.float64)


def _id(x):
    return np.asarray(x, np.float64)


def blocks(a):
    """(nx, ny) cell field -> (nx/2, ny/2, 4) node field (slot 2 di + dj)."""
    return np.stack([a[0::2, 0::2], a[0::2, 1::2], a[1::2, 0::2],
                     a[1::2, 1::2]], -1)


def unblocks(o):
    nxn, nyn, _ = o.shape
    z = np.zeros((2 * nxn, 2 * nyn))
    z[0::2, 0::2] = o[..., 0]
    z[0::2, 1::2] = o[..., 1]
    z[1::2, 0::2] = o[..., 2]
    z[1::2, 1::2] = o[..., 3]
    return z


def features(diag, cE, cN):
    diag, cE, cN = (np.asarray(a, np.float64) for a in (diag, cE, cN))
    cW = np.zeros_like(cE)
    cW[1:] = cE[:-1]
    cS = np.zeros_like(cN)
    cS[:, 1:] = cN[:, :-1]
    tot = (np.abs(cE) + np.abs(cN) + np.abs(cW) + np.abs(cS)) / diag
    nxn, nyn = diag.shape[0] // 2, diag.shape[1] // 2
    phi = np.zeros((nxn, nyn, NF))
    phi[..., 0:4] = blocks(cE / diag)
    phi[..., 4:8] = blocks(cN / diag)
    phi[..., 8:12] = blocks(tot)
    phi[..., 12] = 1.0
    return phi


def aggregate(h):
    agg = np.zeros_like(h)
    agg[1:] += h[:-1]
    agg[:-1] += h[1:]
    agg[:, 1:] += h[:, :-1]
    agg[:, :-1] += h[:, 1:]
    return agg


def sigmoid(x):
    return 1.0 / (1.0 + np.exp(-x))


def forward(flat, r, diag, phi, bf=True):
    """z = N(r) on an (nx, ny) field; returns (z, cache)."""
    P = unpack(flat)
    q = bf16 if bf else _id
    rh = blocks(np.asarray(r, np.float64) / diag)
    h = q(rh @ P["Win"])
    phq = q(phi)
    cache = {"rh": rh, "h": [h], "A": [], "y": [], "g": [], "phi": phq}
    for l in range(NL):
        A = np.concatenate([h, q(aggregate(h))], -1)
        y = A @ q(P[f"W{l}"])
        g = sigmoid(phq @ q(P[f"U{l}"]))
        h = q(g * y)
        cache["A"].append(A)
        cache["y"].append(y)
        cache["g"].append(g)
        cache["h"].append(h)
    o = h @ P["Wout"]
    return unblocks(o), cache


def backward(flat, cache, dz, diag, bf=True):
    """Gradients of sum(dz * z) w.r.t. the parameters (flat FP32 order) and
    w.r.t. the input r."""
    P = unpack(flat)
    q = bf16 if bf else _id
    do = blocks(np.asarray(dz, np.float64))
    G = {}
    h3 = cache["h"][NL]
    G["Wout"] = np.einsum("abk,abs->ks", h3, do)
    dh = do @ P["Wout"].T
    for l in reversed(range(NL)):
        g, y, A = cache["g"][l], cache["y"][l], cache["A"][l]
        dy = q(dh * g)
        dpre = q(dh * y * g * (1.0 - g))
        G[f"W{l}"] = np.einsum("abk,abn->kn", A, dy)
        G[f"U{l}"] = np.einsum("abk,abn->kn", cache["phi"], dpre)
        dA = dy @ q(P[f"W{l}"]).T
        dh = dA[..., :HID] + aggregate(dA[..., HID:])
    G["Win"] = np.einsum("abs,abk->sk", cache["rh"], dh)
    drh = dh @ P["Win"].T
    dr = unblocks(drh) / diag
    return pack(G).astype(np.float64), dr


def adam(flat, grad, m, v, t, lr=3e-3, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step (Kingma & Ba 2015, bias-corrected); returns new
    (params, m, v)."""
    m = b1 * m + (1 - b1) * grad
    v = b2 * v + (1 - b2) * grad * grad
    mh = m / (1 - b1 ** t)
    vh = v / (1 - b2 ** t)
    return flat - lr * mh / (np.sqrt(vh) + eps), m, v


# ------------------------------------------------- the learned V-cycle
def learned_vcycle(flat, H, r, bf=True):
    """z = M r on level 0 of the hierarchy H (oracle.chorin.Hierarchy):
    x = V r (the plain V-cycle from level 0), z = x + N(r - S x).  Returns
    (z, cache)."""
    from oracle.chorin import apply, vcycle
    coef = H.levels[0]["coef"]
    x = vcycle(H, r, level=0)
    u, cache = forward(flat, r - apply(coef, x), coef[0], features(*coef), bf)
    return x + u, cache


def learned_vcycle_backward(flat, H, cache, g, bf=True):
    """Parameter gradient of sum(g * z) for learned_vcycle: the network input
    r - S V r does not depend on the parameters."""
    G, _ = backward(flat, cache, g, H.levels[0]["coef"][0], bf)
    return G
