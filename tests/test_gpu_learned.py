"""Learned initial guess (ibem3d_gnn_*) on the GPU, against the CPU oracle (oracle/gnn_oracle.py):
the neighbour lists bit for bit, the features, the loss and its gradient from the hand-written
backward kernels; BF16-MFMA inference against FP32; training lowers the loss; and end to end, what
the trained network gains on held-out RVEs: its guess against the dilute guess, both judged by the
LU, and block-Jacobi GMRES from its guess against the same solve from zero."""
import os
import sys

import numpy as np
import pytest

from ibemproject_amd import engine as E
from ibemproject_amd import learned as Lr
from ibemproject_amd.generate import make_batch, make_batch_rve

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import gnn_oracle as G  # noqa: E402

pytestmark = pytest.mark.gpu


def case(msg):
    print(f"[case] {msg}", flush=True)


@pytest.mark.parametrize("num", [5, 120])
def test_graph_matches_oracle(num):
    prob = make_batch_rve(41, num, nmin=num, nmax=num)[0]
    case(f"graph num={prob.num}")
    nbr, edge, node = Lr.graph(prob)
    want = G.knn(prob.centres)
    assert np.array_equal(nbr, want)
    we, wn = G.features(prob.centres, prob.radii, want, prob.E1 / prob.E0, prob.nu1)
    assert np.allclose(edge, we, rtol=1e-6, atol=1e-6) and np.allclose(node, wn, rtol=1e-6, atol=1e-7)


def test_gradient_matches_oracle():
    rng = np.random.default_rng(5)
    S = 96
    edge = rng.standard_normal((S, 16, 8)).astype(np.float32)
    edge[:, :, 6] = (rng.uniform(size=(S, 16)) > 0.3).astype(np.float32)
    edge[:, :, 7] = 0.0
    edge[3, :, 6] = 0.0
    node = rng.standard_normal((S, 4)).astype(np.float32)
    node[:, 3] = 0.0
    tgt = rng.standard_normal((S, 360)).astype(np.float32)
    case(f"loss and gradient, {S} spheres")
    g = Lr.Gnn(seed=7)
    g.set_data(edge, node, tgt, normalise=False)
    idx = rng.permutation(S)[:64].astype(np.int32)
    loss, grad = g.loss_grad(idx)
    p = g.params().astype(np.float64)
    rl, rg = G.loss_grad(p, edge[idx].astype(np.float64), node[idx].astype(np.float64),
                         tgt[idx].astype(np.float64))
    print("loss", loss, rl, "grad rel", np.linalg.norm(grad - rg) / np.linalg.norm(rg))
    assert abs(loss - rl) <= 1e-5 * abs(rl)
    assert np.linalg.norm(grad - rg) <= 1e-4 * np.linalg.norm(rg)
    off = 0
    for name, shp in G.SHAPES:  # every block of the gradient, not only the largest
        k = int(np.prod(shp))
        assert np.linalg.norm(grad[off:off + k] - rg[off:off + k]) <= 1e-4 * max(np.linalg.norm(rg[off:off + k]), 1e-12), name
        off += k
    # BF16 inference on MFMA against FP32
    y32 = g.predict(edge, node, bf16=False)
    y16 = g.predict(edge, node, bf16=True)
    yo, _ = G.forward(G.unpack(p), edge.astype(np.float64), node.astype(np.float64))
    e32 = np.linalg.norm(y32 - yo) / np.linalg.norm(yo)
    e16 = np.linalg.norm(y16 - yo) / np.linalg.norm(yo)
    print("fp32 inference vs oracle", e32, "bf16", e16)
    assert e32 < 1e-5 and e16 < 2e-2
    g.close()


@pytest.fixture(scope="module")
def trained():
    """A network trained on the eigenstrains of 600 batched RVEs (8-96 spheres)."""
    probs = [p for p, _ in make_batch(17, 600, nmin=8, nmax=96, workers=16)]
    case(f"training set: {len(probs)} RVEs, {sum(p.num for p in probs)} spheres")
    edge, node, tgt, r = Lr.training_set(probs, refine_steps=4, wave_max=16, mem_budget_gb=100.0,
                                         use_graph=False, host_threads=8)
    assert (r["rel"] <= 1e-10).all()
    g = Lr.Gnn(seed=3)
    g.set_data(edge, node, tgt)
    hist = g.train(epochs=80, batch=256, lr=2e-3)
    print("loss per epoch (units of each order's RMS)", hist[0], hist[9], hist[-1])
    yield g, hist
    g.close()


def test_training_lowers_the_loss(trained):
    _, hist = trained
    assert hist[-1] < 0.6 * hist[0]


def order_errors(x, ref, num):
    o = np.tile(Lr.ORDER, num)
    return [float(np.linalg.norm(x[o == k] - ref[o == k]) / np.linalg.norm(ref[o == k])) for k in (0, 1, 2)]


@pytest.mark.parametrize("seed", [101, 102])
def test_learned_guess_gain(trained, seed):
    """What the trained network gains over the method without it, on held-out RVEs (not in the
    training set) under their macro strain. Without the network every sphere answers the macro
    strain as if it were alone (the dilute guess); the network's guess must be closer to the LU's
    eigenstrains in every coefficient order, since it has learned the neighbours' interaction. As
    the initial guess of block-Jacobi GMRES the scaled start (the multiple of the guess that
    minimises the start residual) must never cost iterations, and both solves must match the LU to
    1e-8. The iterations it saves are printed, not asserted: on this operator the start residual
    stays near one unless a guess is accurate to about 1e-3 in every order."""
    g, _ = trained
    prob, eps6 = make_batch_rve(seed, 0, nmin=60, nmax=90)
    case(f"held-out RVE seed {seed}: num={prob.num} n={prob.n60}")
    _, edge, node = Lr.graph(prob)
    y = g.predict(edge, node, bf16=True)
    x_guess = Lr.guess(y, prob.radii, eps6)
    yd = Lr.dilute_response(prob.E0, prob.nu0, prob.E1, prob.nu1)
    x_dilute = Lr.guess(np.tile(yd, (prob.num, 1)), prob.radii, eps6)
    eng = E.Engine(prob)
    eng.assemble()
    eng.solve(E.SOLVE_FP64_LU)
    _, _, x_lu = eng.getUTVector()
    rel_c, it_c = eng.solve(E.SOLVE_BJ_GMRES, 1e-10)
    _, _, x_c = eng.getUTVector()
    eng.set_initial_guess(x_guess)
    rel_w, it_w = eng.solve(E.SOLVE_BJ_GMRES | E.SOLVE_REUSE, 1e-10)
    r0, alpha = eng.warm_start_info()
    _, _, x_w = eng.getUTVector()
    eng.close()
    nb = 3 * prob.NN
    e_net = order_errors(x_guess, x_lu[nb:], prob.num)
    e_dil = order_errors(x_dilute, x_lu[nb:], prob.num)
    print(f"error vs LU by order (uniform, linear, quadratic): network {np.round(e_net, 3)} "
          f"dilute {np.round(e_dil, 3)}; start residual {r0:.3e} (alpha {alpha:.3f}); "
          f"iterations cold {it_c} warm {it_w}")
    assert all(a < b for a, b in zip(e_net, e_dil))
    for x in (x_c, x_w):
        assert np.linalg.norm(x - x_lu) / np.linalg.norm(x_lu) < 1e-8
    assert rel_c <= 1e-10 and rel_w <= 1e-10
    assert r0 <= 1.0 and it_w <= it_c
