"""CPU oracle of the learned eigenstrain predictor (test infrastructure only: imported by tests/,
never by the product). A float64 numpy restatement of ibemproject_amd/csrc/learned.hip: the kNN
graph on integer-quantised centres, the edge and node features, the network's forward pass and its
gradient by hand-derived backpropagation. The backpropagation itself is pinned by central finite
differences of the forward pass (tests/test_gnn_oracle.py); the kNN by a float brute force on
well-separated points.
"""
from __future__ import annotations

import numpy as np

NNB, FE, FES, FN, FNS, H, OUT = 16, 6, 8, 3, 4, 128, 360
IN2 = H + FN
QBITS = 28
SHAPES = [("W1", (FE, H)), ("b1", (H,)), ("W2", (IN2, H)), ("b2", (H,)), ("W3", (H, H)), ("b3", (H,)),
          ("W4", (H, OUT)), ("b4", (OUT,))]


def unpack(p):
    out, off = {}, 0
    for name, shp in SHAPES:
        k = int(np.prod(shp))
        out[name] = np.asarray(p[off:off + k], dtype=np.float64).reshape(shp)
        off += k
    return out


def pack(d):
    return np.concatenate([d[name].reshape(-1) for name, _ in SHAPES])


def knn(centres, k=NNB):
    """Neighbour lists by exact integer squared distance of rint(c 2^28), ties to the smaller index."""
    q = np.rint(np.asarray(centres, dtype=np.float64) * float(1 << QBITS)).astype(np.int64)
    n = q.shape[0]
    nbr = -np.ones((n, k), dtype=np.int32)
    for i in range(n):
        d = ((q - q[i]) ** 2).sum(1).astype(np.uint64)
        order = sorted((int(d[j]), j) for j in range(n) if j != i)
        for s, (_, j) in enumerate(order[:k]):
            nbr[i, s] = j
    return nbr


def features(centres, radii, nbr, E1_over_E0, nu1):
    c = np.asarray(centres, dtype=np.float64)
    r = np.asarray(radii, dtype=np.float64)
    n = c.shape[0]
    edge = np.zeros((n, NNB, FES))
    node = np.zeros((n, FNS))
    for i in range(n):
        wall = min(min(c[i, a], 1.0 - c[i, a]) for a in range(3))
        node[i, :3] = [np.log(E1_over_E0), nu1, min(wall / r[i], 10.0) / 10.0]
        for s in range(NNB):
            j = nbr[i, s]
            if j < 0:
                continue
            d = (c[j] - c[i]) / r[i]
            dn = np.sqrt((d ** 2).sum())
            sj = r[j] / r[i]
            q3 = sj ** 3 / dn ** 3
            edge[i, s, :6] = [d[0] / dn, d[1] / dn, d[2] / dn, sj, q3, q3 / dn]
            edge[i, s, 6] = 1.0
    return edge, node


def forward(P, edge, node):
    """Network outputs (S, 360) and the intermediate values."""
    e = edge[:, :, :FE]
    m = edge[:, :, 6]
    z1 = np.maximum(e @ P["W1"] + P["b1"], 0.0)                      # (S, 16, H)
    cnt = np.maximum(m.sum(1), 1.0)
    a = (m[:, :, None] * z1).sum(1) / cnt[:, None]
    in2 = np.concatenate([a, node[:, :FN]], axis=1)
    z2 = np.maximum(in2 @ P["W2"] + P["b2"], 0.0)
    z3 = np.maximum(z2 @ P["W3"] + P["b3"], 0.0)
    y = z3 @ P["W4"] + P["b4"]
    return y, dict(e=e, m=m, z1=z1, cnt=cnt, in2=in2, z2=z2, z3=z3)


def loss_grad(p, edge, node, target):
    """Mean squared error over all outputs and its gradient (packed like the parameters)."""
    P = unpack(p)
    y, c = forward(P, edge, node)
    S = y.shape[0]
    d = y - target
    loss = (d ** 2).mean()
    dy = 2.0 * d / d.size
    g = {"W4": c["z3"].T @ dy, "b4": dy.sum(0)}
    dz3 = (dy @ P["W4"].T) * (c["z3"] > 0)
    g["W3"], g["b3"] = c["z2"].T @ dz3, dz3.sum(0)
    dz2 = (dz3 @ P["W3"].T) * (c["z2"] > 0)
    g["W2"], g["b2"] = c["in2"].T @ dz2, dz2.sum(0)
    din2 = dz2 @ P["W2"].T
    da = din2[:, :H] / c["cnt"][:, None]
    dz1 = c["m"][:, :, None] * da[:, None, :] * (c["z1"] > 0)
    g["W1"] = np.einsum("snk,snh->kh", c["e"], dz1)
    g["b1"] = dz1.sum((0, 1))
    assert S == target.shape[0]
    return loss, pack(g)
