"""Learned initial guess of the eigenstrain unknowns (the ibem3d_gnn_* kernels).

Per sphere, a message-passing MLP over its 16 nearest neighbours predicts the linear map from the
macro strain to the sphere's 60 eigenstrain coefficients: 360 outputs, coefficient c under the
unit macro strain t (pair order 00 01 02 11 12 22, e_ij = e_ji = 1) at index 6 c + t, scaled by
r^o (r the sphere's radius, o the coefficient's order: 0 for the 6 uniform, 1 for the 18 linear,
2 for the 36 quadratic coefficients), so the targets are dimensionless. Training data come from
the batched solver (six unit loads per RVE, one factorisation each); the prediction, contracted
with a macro strain and scaled back, is the initial guess of the block-Jacobi GMRES
(IBEM3D_SOLVE_BJ_GMRES), whose result the LU judges (fallback and error).
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import engine as E

NNB = 16
ORDER = np.array([0] * 6 + [1] * 18 + [2] * 36)
_fp = ctypes.POINTER(ctypes.c_float)


def _f(a):
    return a.ctypes.data_as(_fp)


def graph(prob, device: int = 0):
    """(nbr (num, 16) int32, edge (num, 16, 8) float32, node (num, 4) float32) of one RVE."""
    L = E.load()
    num = prob.num
    cen = np.ascontiguousarray(prob.centres, dtype=np.float64).reshape(-1, 3)
    rad = np.ascontiguousarray(prob.radii, dtype=np.float64)
    nbr = np.zeros((num, NNB), dtype=np.int32)
    edge = np.zeros((num, NNB, 8), dtype=np.float32)
    node = np.zeros((num, 4), dtype=np.float32)
    rc = L.ibem3d_gnn_graph(E._d(cen), E._d(rad), num, prob.E1 / prob.E0, prob.nu1, device,
                            nbr.ctypes.data_as(E._ip), _f(edge), _f(node))
    if rc != 0:
        raise E.EngineError(L.ibem3d_gnn_last_error().decode())
    return nbr, edge, node


def targets(eig, radii):
    """(num, 360) float32 training targets from a sphere block of the batch's eigenstrains
    (num, 60, 6): coefficient c scaled by r^order(c), index 6 c + t."""
    scale = np.asarray(radii)[:, None] ** ORDER[None, :]
    return (eig * scale[:, :, None]).reshape(len(radii), 360).astype(np.float32)


def guess(y, radii, eps6):
    """Eigenstrain unknowns (60 num, the 60 layout) from predictions y (num, 360) and the macro
    strain eps6 (6 components, pair order)."""
    m = np.asarray(y, dtype=np.float64).reshape(-1, 60, 6) @ np.asarray(eps6, dtype=np.float64)
    return (m / np.asarray(radii)[:, None] ** ORDER[None, :]).reshape(-1)


class Gnn:
    """The network's parameters, Adam state and training set on one GPU. The targets are trained
    in units of their coefficient order's RMS over the training set (the uniform coefficients are
    an order of magnitude larger than the scaled linear and quadratic ones, which an unweighted
    loss would leave unlearned); predictions are returned in the targets' units."""

    def __init__(self, device: int = 0, seed: int = 1):
        self.norm = np.ones(360, dtype=np.float32)
        self.L = E.load()
        self.g = self.L.ibem3d_gnn_create(device, seed)
        if not self.g:
            raise E.EngineError(self.L.ibem3d_gnn_last_error().decode())
        self.P = int(self.L.ibem3d_gnn_param_count())

    def close(self):
        if getattr(self, "g", None):
            self.L.ibem3d_gnn_destroy(self.g)
            self.g = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc, what):
        if rc != 0:
            raise E.EngineError(f"{what}: {self.L.ibem3d_gnn_last_error().decode()} (code {rc})")

    def params(self) -> np.ndarray:
        p = np.zeros(self.P, dtype=np.float32)
        self._chk(self.L.ibem3d_gnn_get_params(self.g, _f(p)), "get_params")
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == self.P
        self._chk(self.L.ibem3d_gnn_set_params(self.g, _f(p)), "set_params")

    def set_data(self, edge, node, target, normalise: bool = True):
        target = np.asarray(target, dtype=np.float32)
        if normalise:
            t3 = target.reshape(-1, 60, 6)
            rms = np.array([np.sqrt(np.mean(t3[:, ORDER == o] ** 2)) for o in (0, 1, 2)])
            rms[rms == 0] = 1.0
            self.norm = np.repeat(rms[ORDER], 6).astype(np.float32)
        self._edge = np.ascontiguousarray(edge, dtype=np.float32)
        self._node = np.ascontiguousarray(node, dtype=np.float32)
        self._tgt = np.ascontiguousarray(target / self.norm, dtype=np.float32)
        self.S = self._node.shape[0]
        self._chk(self.L.ibem3d_gnn_set_data(self.g, _f(self._edge), _f(self._node), _f(self._tgt), self.S),
                  "set_data")

    def loss_grad(self, idx):
        idx = np.ascontiguousarray(idx, dtype=np.int32)
        loss = ctypes.c_float()
        grad = np.zeros(self.P, dtype=np.float32)
        self._chk(self.L.ibem3d_gnn_loss_grad(self.g, idx.ctypes.data_as(E._ip), idx.size, ctypes.byref(loss),
                                              _f(grad)), "loss_grad")
        return float(loss.value), grad

    def train(self, epochs: int, batch: int = 512, lr: float = 1e-3, seed: int = 0):
        rng = np.random.default_rng(seed)
        perms = np.stack([rng.permutation(self.S) for _ in range(epochs)]).astype(np.int32)
        hist = np.zeros(epochs, dtype=np.float32)
        self._chk(self.L.ibem3d_gnn_train(self.g, perms.ctypes.data_as(E._ip), epochs, batch, lr, _f(hist)),
                  "train")
        return hist

    def predict(self, edge, node, bf16: bool = True) -> np.ndarray:
        edge = np.ascontiguousarray(edge, dtype=np.float32)
        node = np.ascontiguousarray(node, dtype=np.float32)
        S = node.shape[0]
        y = np.zeros((S, 360), dtype=np.float32)
        self._chk(self.L.ibem3d_gnn_predict(self.g, _f(edge), _f(node), S, 1 if bf16 else 0, _f(y)), "predict")
        return y * self.norm


def training_set(problems, device: int = 0, **batch_kw):
    """Solve the problems (six unit loads each, batched) and build (edge, node, target) arrays over
    all their spheres."""
    r = E.batch_stiffness(problems, eigenstrains=True, device=device, **batch_kw)
    edges, nodes, tgts = [], [], []
    for p, eig in zip(problems, r["eig"]):
        if p.num == 0:
            continue
        _, e, f = graph(p, device)
        edges.append(e)
        nodes.append(f)
        tgts.append(targets(eig, p.radii))
    return np.concatenate(edges), np.concatenate(nodes), np.concatenate(tgts), r


def dilute_response(E0: float, nu0: float, E1: float, nu1: float, device: int = 0, radius: float = 0.05,
                    m: int = 8):
    """The guess without the network: every sphere answers the macro strain as if it were alone.
    One sphere of the given radius at the centre of the unit cube (volume fraction 5e-4 at r = 0.05)
    is solved for the six unit macro strains by the batched solver; returns its (360,) scaled
    response in the targets' layout, to be tiled over the spheres of an RVE."""
    from .formats import Problem
    from .generate import cube_mesh, macro_strain_bc, unit_strain_tensor
    nodes, conn, _ = cube_mesh(m)
    bc, tbc = macro_strain_bc(nodes, conn, unit_strain_tensor(np.eye(6)[0]))
    prob = Problem("dilute", E0, nu0, E1, nu1, nodes, conn, bc, tbc, np.array([[0.5, 0.5, 0.5]]),
                   np.array([radius]), np.zeros((0, 3)))
    r = E.batch_stiffness([prob], eigenstrains=True, device=device, refine_steps=4, wave_max=1,
                          use_graph=False)
    return targets(r["eig"][0], prob.radii)[0]


def macro_strain_of_uniaxial_traction(E0: float, nu0: float, sigma: float = 1.0):
    """The matrix's strain under a uniaxial stress sigma along x: the macro-strain input for the
    traction-loaded RVE (pair order 00 01 02 11 12 22)."""
    return np.array([1.0, 0.0, 0.0, -nu0, 0.0, -nu0]) * sigma / E0
