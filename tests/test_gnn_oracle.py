"""The CPU oracle of the learned predictor pinned on its own terms (no GPU): hand-derived gradients
against central finite differences of the forward pass, and the integer kNN against a float brute
force where no two distances tie."""
import os
import sys

import numpy as np

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import gnn_oracle as G  # noqa: E402


def random_case(seed, S=5):
    rng = np.random.default_rng(seed)
    p = rng.standard_normal(sum(int(np.prod(s)) for _, s in G.SHAPES)) * 0.2
    edge = rng.standard_normal((S, G.NNB, G.FES))
    edge[:, :, 6] = (rng.uniform(size=(S, G.NNB)) > 0.2).astype(float)
    edge[0, :, 6] = 0.0  # a sphere without neighbours
    node = rng.standard_normal((S, G.FNS))
    tgt = rng.standard_normal((S, G.OUT))
    return p, edge, node, tgt


def test_gradient_by_central_differences():
    p, edge, node, tgt = random_case(1)
    _, g = G.loss_grad(p, edge, node, tgt)
    rng = np.random.default_rng(2)
    off = 0
    for name, shp in G.SHAPES:
        k = int(np.prod(shp))
        for u in rng.choice(k, size=min(k, 6), replace=False):
            q = off + u
            h = 1e-6
            pp, pm = p.copy(), p.copy()
            pp[q] += h
            pm[q] -= h
            fd = (G.loss_grad(pp, edge, node, tgt)[0] - G.loss_grad(pm, edge, node, tgt)[0]) / (2 * h)
            assert abs(fd - g[q]) <= 1e-6 * max(1.0, abs(g[q])) + 1e-9, (name, u, fd, g[q])
        off += k


def test_knn_matches_float_brute_force():
    rng = np.random.default_rng(4)
    c = rng.uniform(0.1, 0.9, size=(60, 3))
    nbr = G.knn(c)
    d = ((c[:, None, :] - c[None, :, :]) ** 2).sum(-1)
    np.fill_diagonal(d, np.inf)
    want = np.argsort(d, axis=1, kind="stable")[:, :G.NNB]
    assert np.array_equal(nbr, want)


def test_knn_ties_and_padding():
    # four points on a square: two equidistant neighbours, the smaller index first; fewer than 16
    c = np.array([[0.5, 0.5, 0.5], [0.6, 0.5, 0.5], [0.5, 0.6, 0.5], [0.6, 0.6, 0.5]])
    nbr = G.knn(c)
    assert list(nbr[0, :3]) == [1, 2, 3] and (nbr[:, 3:] == -1).all()
