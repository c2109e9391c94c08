"""ctypes binding of the C oracle (liboracle.so) -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this module, as the checker
or the reported CPU baseline. The product path (ibemproject_amd) never imports it.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


class Material(ctypes.Structure):
    _fields_ = [("E0", ctypes.c_double), ("nu0", ctypes.c_double), ("E1", ctypes.c_double),
                ("nu1", ctypes.c_double)]


class CProblem(ctypes.Structure):
    _fields_ = [("NN", ctypes.c_int), ("NE", ctypes.c_int), ("num", ctypes.c_int),
                ("nump", ctypes.c_int), ("mat", Material), ("nodes", _dp), ("conn", _ip),
                ("bc", _dp), ("tbc", _dp), ("centres", _dp), ("radii", _dp), ("points", _dp)]


PHASES = dict(addBEM=1, addFieldToBEM=2, incl0=4, incl1=8, incl2=16, bem0=32, bem1=64, bem2=128)
ALL = 255


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (make -C oracle oracle)")
        L = ctypes.CDLL(path)
        L.orc_D.restype = ctypes.c_double
        L.orc_D.argtypes = [ctypes.POINTER(Material), ctypes.c_int, _ip, _dp, ctypes.c_double]
        L.orc_green.restype = ctypes.c_double
        L.orc_green.argtypes = [ctypes.POINTER(Material), ctypes.c_int, _ip, _dp]
        L.orc_igreen.restype = ctypes.c_double
        L.orc_igreen.argtypes = [ctypes.POINTER(Material), ctypes.c_int, _ip, _dp, ctypes.c_double]
        L.orc_integ3.argtypes = [ctypes.POINTER(Material), _ip, _dp, ctypes.c_int, _dp, _dp]
        L.orc_integ3_stress.argtypes = [ctypes.POINTER(Material), ctypes.c_int, ctypes.c_int, _dp,
                                        _dp, _dp, _dp]
        for f in ("orc_assemble78",):
            getattr(L, f).argtypes = [ctypes.POINTER(CProblem), ctypes.c_int, _dp, _dp]
        L.orc_apply_bc.argtypes = [ctypes.POINTER(CProblem), _dp, _dp, _dp, _dp]
        L.orc_column_map.argtypes = [ctypes.POINTER(CProblem), _ip, _ip]
        L.orc_reduce_78_to_60.argtypes = [ctypes.POINTER(CProblem), _dp, _dp, _dp, _dp]
        L.orc_expand_60_to_78.argtypes = [ctypes.POINTER(CProblem), _dp, _dp]
        L.orc_fullpiv_solve.argtypes = [ctypes.c_int, _dp, _dp, _dp]
        L.orc_partialpiv_solve.argtypes = [ctypes.c_int, _dp, _dp, _dp]
        L.orc_get_ut.argtypes = [ctypes.POINTER(CProblem), _dp, _dp, _dp]
        L.orc_stress.argtypes = [ctypes.POINTER(CProblem), _dp, _dp, _dp]
        _LIB = L
    return _LIB


def _d(a):
    return a.ctypes.data_as(_dp)


def _i(a):
    return a.ctypes.data_as(_ip)


def material(E0, nu0, E1, nu1):
    return Material(E0, nu0, E1, nu1)


def D(mat, kind, idx, x, a):
    g = np.ascontiguousarray(idx, dtype=np.int32)
    xx = np.ascontiguousarray(x, dtype=np.float64)
    return lib().orc_D(ctypes.byref(mat), kind, _i(g), _d(xx), float(a))


def green(mat, order, idx, x):
    g = np.ascontiguousarray(idx, dtype=np.int32)
    xx = np.ascontiguousarray(x, dtype=np.float64)
    return lib().orc_green(ctypes.byref(mat), order, _i(g), _d(xx))


def igreen(mat, kind, idx, x, a):
    g = np.ascontiguousarray(idx, dtype=np.int32)
    xx = np.ascontiguousarray(x, dtype=np.float64)
    return lib().orc_igreen(ctypes.byref(mat), kind, _i(g), _d(xx), float(a))


def integ3(mat, sctrb, nodes):
    nn = nodes.shape[0]
    s = np.ascontiguousarray(sctrb, dtype=np.int32)
    nd = np.ascontiguousarray(nodes, dtype=np.float64)
    dU = np.zeros((3 * nn, 12))
    dT = np.zeros((3 * nn, 12))
    lib().orc_integ3(ctypes.byref(mat), _i(s), _d(nd), nn, _d(dU), _d(dT))
    return dU, dT


def integ3_stress(mat, order, elcor, point, post=0):
    rows = (6, 18, 54)[order]
    el = np.ascontiguousarray(elcor, dtype=np.float64)
    pt = np.ascontiguousarray(point, dtype=np.float64)
    su = np.zeros((rows, 12))
    st = np.zeros((rows, 12))
    lib().orc_integ3_stress(ctypes.byref(mat), order, post, _d(el), _d(pt), _d(su), _d(st))
    return su, st


class OracleProblem:
    """Holds contiguous copies of a formats.Problem for the C oracle."""

    def __init__(self, prob):
        self.prob = prob
        self.nodes = np.ascontiguousarray(prob.nodes, dtype=np.float64)
        self.conn = np.ascontiguousarray(prob.conn, dtype=np.int32)
        self.bc = np.ascontiguousarray(prob.bc, dtype=np.float64)
        self.tbc = np.ascontiguousarray(prob.tbc, dtype=np.float64)
        self.centres = np.ascontiguousarray(prob.centres, dtype=np.float64).reshape(-1, 3)
        self.radii = np.ascontiguousarray(prob.radii, dtype=np.float64)
        self.points = np.ascontiguousarray(prob.points, dtype=np.float64).reshape(-1, 3)
        self.c = CProblem(prob.NN, prob.NE, prob.num, prob.nump,
                          Material(prob.E0, prob.nu0, prob.E1, prob.nu1), _d(self.nodes),
                          _i(self.conn), _d(self.bc), _d(self.tbc), _d(self.centres),
                          _d(self.radii), _d(self.points))
        self.n78 = prob.n78
        self.n60 = prob.n60

    def assemble78(self, phases=ALL):
        H = np.zeros((self.n78, self.n78), order="F")
        G = np.zeros((self.n78, 12 * self.prob.NE), order="F")
        lib().orc_assemble78(ctypes.byref(self.c), phases, _d(H), _d(G))
        return H, G

    def apply_bc(self, H, G):
        A = np.zeros((self.n78, self.n78), order="F")
        b = np.zeros(self.n78)
        rc = lib().orc_apply_bc(ctypes.byref(self.c), _d(np.asfortranarray(H)),
                                _d(np.asfortranarray(G)), _d(A), _d(b))
        if rc != 0:
            raise ValueError("BC fold: #displacement unknowns + #traction keys != 3NN")
        return A, b

    def column_map(self):
        nc = np.zeros(3 * self.prob.NN, dtype=np.int32)
        sc = np.zeros(12 * self.prob.NE, dtype=np.int32)
        t = lib().orc_column_map(ctypes.byref(self.c), _i(nc), _i(sc))
        return t, nc, sc

    def reduce60(self, A78, b78):
        A = np.zeros((self.n60, self.n60), order="F")
        b = np.zeros(self.n60)
        lib().orc_reduce_78_to_60(ctypes.byref(self.c), _d(np.asfortranarray(A78)),
                                  _d(np.ascontiguousarray(b78)), _d(A), _d(b))
        return A, b

    def expand78(self, x60):
        x = np.zeros(self.n78)
        lib().orc_expand_60_to_78(ctypes.byref(self.c), _d(np.ascontiguousarray(x60)), _d(x))
        return x

    def get_ut(self, X78):
        U = np.zeros(self.n78)
        T = np.zeros(12 * self.prob.NE)
        lib().orc_get_ut(ctypes.byref(self.c), _d(np.ascontiguousarray(X78)), _d(U), _d(T))
        return U, T

    def stress(self, U, T):
        s = np.zeros((self.prob.nump, 6))
        lib().orc_stress(ctypes.byref(self.c), _d(np.ascontiguousarray(U)),
                         _d(np.ascontiguousarray(T)), _d(s))
        return s


def average_fields(prob, U, T):
    """Volume averages from a boundary solution, restated from the divergence theorem (no reference
    counterpart): <s_ij> = (1/V) int t_i x_j dS (symmetrised), <e_ij> = (1/V) int (u_i n_j + u_j n_i)/2
    dS, V = (1/3) int x.n dS, on each element's 8 x 8 Gauss-Legendre rule with the bilinear quad map
    of BEMElasticHelper3D.cpp:191-274 (Serendip_func, Normal_Jac, Cartesian). Returns (sigma6,
    eps6, V) in the order 00 01 02 11 12 22."""
    xg, wg = np.polynomial.legendre.leggauss(8)
    S = np.zeros((3, 3))
    E = np.zeros((3, 3))
    V = 0.0
    nodes = np.asarray(prob.nodes, dtype=np.float64).reshape(-1, 3)
    conn = np.asarray(prob.conn, dtype=np.int64).reshape(-1, 4) - 1
    U = np.asarray(U)
    T = np.asarray(T).reshape(-1, 4, 3)
    for e in range(conn.shape[0]):
        X = nodes[conn[e]]          # 4 x 3
        ue = U[: 3 * prob.NN].reshape(-1, 3)[conn[e]]
        te = T[e]
        for a, xs in enumerate(xg):
            for b, et in enumerate(xg):
                N = 0.25 * np.array([(1 - xs) * (1 - et), (1 + xs) * (1 - et), (1 + xs) * (1 + et),
                                     (1 - xs) * (1 + et)])
                dxs = 0.25 * np.array([-(1 - et), (1 - et), (1 + et), -(1 + et)])
                det = 0.25 * np.array([-(1 - xs), -(1 + xs), (1 + xs), (1 - xs)])
                vx, ve = dxs @ X, det @ X
                nv = np.cross(vx, ve)
                jac = np.linalg.norm(nv)
                nv = nv / jac
                x = N @ X
                w = wg[a] * wg[b] * jac
                t = N @ te
                u = N @ ue
                S += w * np.outer(t, x)
                E += w * 0.5 * (np.outer(u, nv) + np.outer(nv, u))
                V += w * (x @ nv) / 3.0
    S = 0.5 * (S + S.T) / V
    E = E / V
    idx = [(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)]
    return np.array([S[i, j] for i, j in idx]), np.array([E[i, j] for i, j in idx]), V


def average_eigenstrain(prob, U, V):
    """(1/V) sum_s V_s (E_ij + a_s^2/5 sum_k E_ij,kk): the integral over each sphere of the quadratic
    eigenstrain polynomial E_ij + E_ij,k x_k + E_ij,kl x_k x_l (78-layout unknowns,
    configElastic3D.cpp:160-195), using int x_k dV = 0 and int x_k x_l dV = V_s a^2/5 d_kl."""
    U = np.asarray(U)
    acc = np.zeros(6)
    for h in range(prob.num):
        a = float(prob.radii[h])
        Vs = 4.0 / 3.0 * np.pi * a ** 3
        E = U[3 * prob.NN + 78 * h: 3 * prob.NN + 78 * (h + 1)]
        quad = E[24:].reshape(6, 3, 3)
        acc += Vs * (E[:6] + a * a / 5.0 * np.trace(quad, axis1=1, axis2=2))
    return acc / V


def fullpiv_solve(A, b):
    n = A.shape[0]
    M = np.array(A, dtype=np.float64, order="F", copy=True)
    x = np.zeros(n)
    rc = lib().orc_fullpiv_solve(n, _d(M), _d(np.ascontiguousarray(b, dtype=np.float64)), _d(x))
    if rc != 0:
        raise np.linalg.LinAlgError("singular")
    return x


def partialpiv_solve(A, b):
    n = A.shape[0]
    M = np.array(A, dtype=np.float64, order="F", copy=True)
    x = np.zeros(n)
    rc = lib().orc_partialpiv_solve(n, _d(M), _d(np.ascontiguousarray(b, dtype=np.float64)), _d(x))
    if rc != 0:
        raise np.linalg.LinAlgError("singular")
    return x


def solve_problem(prob, layout=78):
    """Whole oracle pipeline: assemble (78), fold BCs, solve, unpack, stresses.

    layout 78 solves the reference's system with complete pivoting; layout 60 reduces it to the
    60-unknown system and solves with partial pivoting. Returns dict of arrays (78-layout U)."""
    op = OracleProblem(prob)
    H, G = op.assemble78()
    A78, b78 = op.apply_bc(H, G)
    if layout == 78:
        X78 = fullpiv_solve(A78, b78)
        A60 = b60 = X60 = None
    else:
        A60, b60 = op.reduce60(A78, b78)
        X60 = partialpiv_solve(A60, b60)
        X78 = op.expand78(X60)
    U, T = op.get_ut(X78)
    s = op.stress(U, T) if prob.nump > 0 else np.zeros((0, 6))
    return dict(H=H, G=G, A78=A78, b78=b78, X78=X78, A60=A60, b60=b60, X60=X60, U=U, T=T,
                stress=s)
