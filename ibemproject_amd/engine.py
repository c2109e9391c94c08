"""ctypes binding of the HIP engine's C ABI (include/ibem3d.h).

The binding loads the in-tree ``libibem3d.so`` and raises if it is missing or if a call fails; there
is no CPU fallback. Method names follow the reference plugin phases (iBEMSystemService.h:58-99,
systemSolverService.h:41-45, postProcessorElastic3D.cpp:525).
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from .formats import Problem

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libibem3d.so")

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)

SOLVE_FP64_LU = 0
SOLVE_MIXED_FP32 = 1
SOLVE_MIXED_FP16 = 2     # FP32 LU, trailing updates on three FP16 products of hi/lo splits
SOLVE_MIXED_FP16_1 = 3   # FP32 LU, trailing updates on one FP16 product (power-of-two scaled operands)
SOLVE_MIXED_BF16 = 4     # FP32 LU, trailing updates on one BF16 product
SOLVE_BJ_GMRES = 5       # FP64 GMRES preconditioned by block Jacobi (no n x n factor), LU fallback
SOLVE_BJ_GMRES16 = 6     # the same, its directions from inner block-Jacobi GMRES on a 16-bit copy of A
SOLVE_MIXED_BF16X3 = 7   # split BF16 (hi/lo) trailing-update products + FP64 refinement
SOLVE_GMRES = 0x100  # flag for the mixed modes: GMRES-IR from the start
SOLVE_REUSE = 0x200  # flag: reuse the factors of the last factorisation in the same mode
SOLVE_COMPACT = 0x400  # flag: the factor takes A's place (16-bit factor copy for the mixed modes)
TUNE_BIG_MIN_TILES = 1
TUNE_PANEL_CU_PERCENT = 2
TUNE_INNER_ITERS = 3
TUNE_INNER_TOL_EXP = 4
TUNE_ROWS_POINTWISE = 5
TUNE_BJ_NOPIV = 6
TUNE_BJ_OVERLAP = 7
TUNE_STRESS_BESIDE = 8
TUNE_BJ_FP32 = 9
TUNE_RESIDUAL_UPDATE = 10


def set_tuning(key: int, value: int) -> None:
    L = load()
    if L.ibem3d_set_tuning(key, value) != 0:
        raise EngineError(f"set_tuning({key}, {value}): {L.ibem3d_last_error(None).decode()}")


def get_tuning(key: int) -> int:
    return int(load().ibem3d_get_tuning(key))


class BatchOptions(ctypes.Structure):
    """ibem3d_batch_options: mode, in-graph refinement steps, RVEs per wave, wave memory budget
    (GB), graph capture on/off."""
    _fields_ = [("mode", ctypes.c_int), ("refine_steps", ctypes.c_int), ("wave_max", ctypes.c_int),
                ("mem_budget_gb", ctypes.c_double), ("use_graph", ctypes.c_int),
                ("host_threads", ctypes.c_int)]


class _CProblem(ctypes.Structure):
    _fields_ = [("NN", ctypes.c_int), ("NE", ctypes.c_int), ("num", ctypes.c_int),
                ("nump", ctypes.c_int), ("E0", ctypes.c_double), ("nu0", ctypes.c_double),
                ("E1", ctypes.c_double), ("nu1", ctypes.c_double), ("nodes", _dp), ("conn", _ip),
                ("bc", _dp), ("tbc", _dp), ("centres", _dp), ("radii", _dp), ("points", _dp)]


_LIB = None

EXPORTS = ["ibem3d_create", "ibem3d_destroy", "ibem3d_last_error", "ibem3d_size", "ibem3d_add_bem",
           "ibem3d_add_field", "ibem3d_add_incl", "ibem3d_add_bem_rows", "ibem3d_apply_bc",
           "ibem3d_assemble", "ibem3d_reset", "ibem3d_solve", "ibem3d_get_ut", "ibem3d_stress",
           "ibem3d_get_system", "ibem3d_solve_info", "ibem3d_timings", "ibem3d_reset_timings",
           "ibem3d_read_inputs", "ibem3d_inputs_problem", "ibem3d_inputs_title",
           "ibem3d_free_inputs", "ibem3d_io_error", "ibem3d_write_stress", "ibem3d_average_fields",
           "ibem3d_set_known_values", "ibem3d_average_eigenstrain", "ibem3d_residual",
           "ibem3d_update_known_values", "ibem3d_batch_create", "ibem3d_batch_destroy",
           "ibem3d_batch_last_error", "ibem3d_batch_stiffness", "ibem3d_set_tuning",
           "ibem3d_get_tuning", "ibem3d_set_initial_guess", "ibem3d_solve_fell_back", "ibem3d_inner_iterations",
           "ibem3d_rows_moment_form", "ibem3d_bj_nopiv",
           "ibem3d_warm_start_info",
           "ibem3d_gnn_last_error", "ibem3d_gnn_param_count", "ibem3d_gnn_graph", "ibem3d_gnn_create",
           "ibem3d_gnn_destroy", "ibem3d_gnn_get_params", "ibem3d_gnn_set_params", "ibem3d_gnn_set_data",
           "ibem3d_gnn_loss_grad", "ibem3d_gnn_train", "ibem3d_gnn_predict", "ibem3d_gnn_predict_noalloc"]


class EngineError(RuntimeError):
    pass


def load(path: str = LIB_PATH):
    """Load libibem3d.so (raises when it is not built)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(path):
        raise EngineError(f"{path} is not built; run __graft_entry__.build() or "
                          "python ibemproject_amd/build.py")
    L = ctypes.CDLL(path)
    L.ibem3d_create.restype = ctypes.c_void_p
    L.ibem3d_create.argtypes = [ctypes.POINTER(_CProblem), ctypes.c_int]
    L.ibem3d_destroy.argtypes = [ctypes.c_void_p]
    L.ibem3d_last_error.restype = ctypes.c_char_p
    L.ibem3d_last_error.argtypes = [ctypes.c_void_p]
    L.ibem3d_size.restype = ctypes.c_long
    L.ibem3d_size.argtypes = [ctypes.c_void_p]
    for f in ("ibem3d_add_bem", "ibem3d_add_field", "ibem3d_apply_bc", "ibem3d_assemble",
              "ibem3d_reset"):
        getattr(L, f).argtypes = [ctypes.c_void_p]
    for f in ("ibem3d_add_incl", "ibem3d_add_bem_rows"):
        getattr(L, f).argtypes = [ctypes.c_void_p, ctypes.c_int]
    L.ibem3d_solve.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_double]
    L.ibem3d_get_ut.argtypes = [ctypes.c_void_p, _dp, _dp, _dp]
    L.ibem3d_stress.argtypes = [ctypes.c_void_p, _dp]
    L.ibem3d_get_system.argtypes = [ctypes.c_void_p, _dp, _dp]
    L.ibem3d_solve_info.argtypes = [ctypes.c_void_p, _dp, _ip]
    L.ibem3d_timings.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_int, _dp, ctypes.c_int]
    L.ibem3d_reset_timings.argtypes = [ctypes.c_void_p]
    L.ibem3d_read_inputs.restype = ctypes.c_void_p
    L.ibem3d_read_inputs.argtypes = [ctypes.c_char_p]
    L.ibem3d_inputs_problem.restype = ctypes.POINTER(_CProblem)
    L.ibem3d_inputs_problem.argtypes = [ctypes.c_void_p]
    L.ibem3d_inputs_title.restype = ctypes.c_char_p
    L.ibem3d_inputs_title.argtypes = [ctypes.c_void_p]
    L.ibem3d_free_inputs.argtypes = [ctypes.c_void_p]
    L.ibem3d_io_error.restype = ctypes.c_char_p
    L.ibem3d_io_error.argtypes = []
    L.ibem3d_write_stress.argtypes = [ctypes.c_char_p, _dp, ctypes.c_int, ctypes.c_int]
    L.ibem3d_average_fields.argtypes = [ctypes.c_void_p, _dp, _dp, _dp]
    L.ibem3d_set_known_values.argtypes = [ctypes.c_void_p, _dp, _dp]
    L.ibem3d_average_eigenstrain.argtypes = [ctypes.c_void_p, _dp]
    L.ibem3d_residual.argtypes = [ctypes.c_void_p, _dp]
    L.ibem3d_update_known_values.argtypes = [ctypes.c_void_p, _dp, _dp]
    L.ibem3d_set_initial_guess.argtypes = [ctypes.c_void_p, _dp]
    L.ibem3d_solve_fell_back.argtypes = [ctypes.c_void_p]
    L.ibem3d_rows_moment_form.argtypes = [ctypes.c_void_p]
    L.ibem3d_bj_nopiv.argtypes = [ctypes.c_void_p]
    L.ibem3d_inner_iterations.restype = ctypes.c_long
    L.ibem3d_inner_iterations.argtypes = [ctypes.c_void_p]
    L.ibem3d_warm_start_info.argtypes = [ctypes.c_void_p, _dp, _dp]
    L.ibem3d_set_tuning.argtypes = [ctypes.c_int, ctypes.c_long]
    L.ibem3d_get_tuning.restype = ctypes.c_long
    L.ibem3d_get_tuning.argtypes = [ctypes.c_int]
    L.ibem3d_batch_create.restype = ctypes.c_void_p
    L.ibem3d_batch_create.argtypes = [ctypes.POINTER(_CProblem), ctypes.c_int, ctypes.c_int]
    L.ibem3d_batch_destroy.argtypes = [ctypes.c_void_p]
    L.ibem3d_batch_last_error.restype = ctypes.c_char_p
    L.ibem3d_batch_last_error.argtypes = []
    L.ibem3d_batch_stiffness.argtypes = [ctypes.c_void_p, ctypes.POINTER(BatchOptions), ctypes.c_double,
                                         _dp, _dp, _dp, _ip, _dp, _dp]
    _fp = ctypes.POINTER(ctypes.c_float)
    L.ibem3d_gnn_last_error.restype = ctypes.c_char_p
    L.ibem3d_gnn_last_error.argtypes = []
    L.ibem3d_gnn_param_count.restype = ctypes.c_long
    L.ibem3d_gnn_param_count.argtypes = []
    L.ibem3d_gnn_graph.argtypes = [_dp, _dp, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                   _ip, _fp, _fp]
    L.ibem3d_gnn_create.restype = ctypes.c_void_p
    L.ibem3d_gnn_create.argtypes = [ctypes.c_int, ctypes.c_ulonglong]
    L.ibem3d_gnn_destroy.argtypes = [ctypes.c_void_p]
    L.ibem3d_gnn_get_params.argtypes = [ctypes.c_void_p, _fp]
    L.ibem3d_gnn_set_params.argtypes = [ctypes.c_void_p, _fp]
    L.ibem3d_gnn_set_data.argtypes = [ctypes.c_void_p, _fp, _fp, _fp, ctypes.c_long]
    L.ibem3d_gnn_loss_grad.argtypes = [ctypes.c_void_p, _ip, ctypes.c_long, _fp, _fp]
    L.ibem3d_gnn_train.argtypes = [ctypes.c_void_p, _ip, ctypes.c_int, ctypes.c_long, ctypes.c_float, _fp]
    L.ibem3d_gnn_predict.argtypes = [ctypes.c_void_p, _fp, _fp, ctypes.c_long, ctypes.c_int, _fp]
    L.ibem3d_gnn_predict_noalloc.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_long,
                                             ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
    _LIB = L
    return L


def _d(a):
    return a.ctypes.data_as(_dp)


def read_inputs_native(directory: str) -> Problem:
    """Parse a problem directory with the library's own reader (ibem3d_read_inputs)."""
    L = load()
    h = L.ibem3d_read_inputs(os.fsencode(directory))
    if not h:
        raise EngineError(L.ibem3d_io_error().decode())
    try:
        c = L.ibem3d_inputs_problem(h).contents
        NN, NE, num, nump = c.NN, c.NE, c.num, c.nump

        def arr(ptr, n, dt=np.float64):
            return np.ctypeslib.as_array(ptr, shape=(n,)).astype(dt, copy=True) if n else np.zeros(0, dt)
        return Problem(L.ibem3d_inputs_title(h).decode(), c.E0, c.nu0, c.E1, c.nu1,
                       arr(c.nodes, 3 * NN).reshape(NN, 3), arr(c.conn, 4 * NE, np.int32).reshape(NE, 4),
                       arr(c.bc, 6 * NN).reshape(3 * NN, 2), arr(c.tbc, 24 * NE).reshape(NE, 24),
                       arr(c.centres, 3 * num).reshape(num, 3), arr(c.radii, num),
                       arr(c.points, 3 * nump).reshape(nump, 3))
    finally:
        L.ibem3d_free_inputs(h)


def write_stress_native(path: str, s: np.ndarray, full: bool = False) -> None:
    """Write stresses with the library's writer (reference format, or %.17g with full=True)."""
    L = load()
    a = np.ascontiguousarray(s, dtype=np.float64).reshape(-1, 6)
    if L.ibem3d_write_stress(os.fsencode(path), _d(a), a.shape[0], int(full)) != 0:
        raise EngineError(L.ibem3d_io_error().decode())


def _c_problem(prob: Problem):
    """The ibem3d_problem of prob and the arrays it points into (keep them alive while it is used)."""
    keep = dict(
        nodes=np.ascontiguousarray(prob.nodes, dtype=np.float64),
        conn=np.ascontiguousarray(prob.conn, dtype=np.int32),
        bc=np.ascontiguousarray(prob.bc, dtype=np.float64),
        tbc=np.ascontiguousarray(prob.tbc, dtype=np.float64),
        centres=np.ascontiguousarray(prob.centres, dtype=np.float64).reshape(-1, 3),
        radii=np.ascontiguousarray(prob.radii, dtype=np.float64),
        points=np.ascontiguousarray(prob.points, dtype=np.float64).reshape(-1, 3),
    )
    k = keep
    cp = _CProblem(prob.NN, prob.NE, prob.num, prob.nump, prob.E0, prob.nu0, prob.E1, prob.nu1,
                   _d(k["nodes"]), k["conn"].ctypes.data_as(_ip), _d(k["bc"]), _d(k["tbc"]),
                   _d(k["centres"]), _d(k["radii"]), _d(k["points"]))
    return keep, cp


def batch_stiffness(problems, mode: int = SOLVE_MIXED_FP16, refine_steps: int = 4, wave_max: int = 16,
                    mem_budget_gb: float = 200.0, use_graph: bool = True, tol: float = 1e-10,
                    device: int = 0, host_threads: int = 1, eigenstrains: bool = False):
    """Apparent stiffness of many RVEs (ibem3d_batch_stiffness): problems whose boundary
    displacements are all prescribed; six unit macro strains per RVE on one factorisation. Returns
    dict(C (count, 6, 6) from the inclusions, C_bnd (count, 6, 6) from the tractions, rel (count, 6),
    host_refined (count, 6), times {setup, capture, instantiate, run, host, waves}; with eigenstrains:
    eig, a list of (num, 60, 6) arrays, each sphere's eigenstrain unknowns under the six unit
    macro strains)."""
    L = load()
    count = len(problems)
    keeps, cps = [], (_CProblem * count)()
    for i, p in enumerate(problems):
        k, cp = _c_problem(p)
        keeps.append(k)
        cps[i] = cp
    b = L.ibem3d_batch_create(cps, count, device)
    if not b:
        raise EngineError(L.ibem3d_batch_last_error().decode())
    C = np.zeros((count, 6, 6))
    Cb = np.zeros((count, 6, 6))
    rel = np.zeros((count, 6))
    hr = np.zeros((count, 6), dtype=np.int32)
    tm = np.zeros(6)
    eig = np.zeros(sum(360 * p.num for p in problems)) if eigenstrains else None
    opt = BatchOptions(mode, refine_steps, wave_max, mem_budget_gb, 1 if use_graph else 0, host_threads)
    try:
        rc = L.ibem3d_batch_stiffness(b, ctypes.byref(opt), tol, _d(C), _d(Cb), _d(rel),
                                      hr.ctypes.data_as(_ip), _d(tm), _d(eig) if eig is not None else None)
        if rc != 0:
            raise EngineError(f"batch_stiffness: {L.ibem3d_batch_last_error().decode()} (code {rc})")
    finally:
        L.ibem3d_batch_destroy(b)
    names = ("setup", "capture", "instantiate", "run", "host", "waves")
    out = dict(C=C, C_bnd=Cb, rel=rel, host_refined=hr, times=dict(zip(names, tm.tolist())))
    if eig is not None:
        blocks, off = [], 0
        for p in problems:
            k = 360 * p.num
            blocks.append(eig[off:off + k].reshape(6, p.num, 60).transpose(1, 2, 0).copy())
            off += k
        out["eig"] = blocks
    return out


class Engine:
    """One RVE on one GPU: owns the device system and its solution."""

    def __init__(self, prob: Problem, device: int = 0):
        self.L = load()
        self.prob = prob
        self._keep, cp = _c_problem(prob)
        self.ctx = self.L.ibem3d_create(ctypes.byref(cp), device)
        if not self.ctx:
            raise EngineError(self.L.ibem3d_last_error(None).decode())
        self.n = int(self.L.ibem3d_size(self.ctx))

    def close(self):
        if getattr(self, "ctx", None):
            self.L.ibem3d_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc: int, what: str):
        if rc != 0:
            raise EngineError(f"{what}: {self.L.ibem3d_last_error(self.ctx).decode()} (code {rc})")

    # ---- reference phases ----
    def addBEM(self):
        self._chk(self.L.ibem3d_add_bem(self.ctx), "addBEM")

    def addFieldToBEM(self):
        self._chk(self.L.ibem3d_add_field(self.ctx), "addFieldToBEM")

    def addFluxEquivalentInclusion(self, order: int = 0):
        self._chk(self.L.ibem3d_add_incl(self.ctx, order), "addFluxEquivalentInclusion")

    def addFluxEquivalentBEM(self, order: int = 0):
        self._chk(self.L.ibem3d_add_bem_rows(self.ctx, order), "addFluxEquivalentBEM")

    def ApplyBC(self):
        self._chk(self.L.ibem3d_apply_bc(self.ctx), "ApplyBC")

    def assemble(self):
        self._chk(self.L.ibem3d_assemble(self.ctx), "assemble")

    def reset(self):
        self._chk(self.L.ibem3d_reset(self.ctx), "reset")

    def solve(self, mode: int = SOLVE_FP64_LU, tol: float = 1e-10):
        self._chk(self.L.ibem3d_solve(self.ctx, mode, tol), "solve")
        rel = ctypes.c_double()
        it = ctypes.c_int()
        self.L.ibem3d_solve_info(self.ctx, ctypes.byref(rel), ctypes.byref(it))
        return rel.value, it.value

    def set_initial_guess(self, x_eig):
        """Eigenstrain unknowns (60 num, 60 layout) the block-Jacobi GMRES starts from; None clears."""
        if x_eig is None:
            self._chk(self.L.ibem3d_set_initial_guess(self.ctx, None), "set_initial_guess")
            return
        v = np.ascontiguousarray(x_eig, dtype=np.float64).reshape(-1)
        if v.size != 60 * self.prob.num:
            raise ValueError("set_initial_guess: expected 60 num values")
        self._chk(self.L.ibem3d_set_initial_guess(self.ctx, _d(v)), "set_initial_guess")

    def warm_start_info(self):
        """(relative residual of the scaled guess, its scale alpha); rel = -1 when none was used."""
        r, a = ctypes.c_double(), ctypes.c_double()
        self.L.ibem3d_warm_start_info(self.ctx, ctypes.byref(r), ctypes.byref(a))
        return r.value, a.value

    def fell_back(self) -> bool:
        return bool(self.L.ibem3d_solve_fell_back(self.ctx))

    def bj_nopiv(self) -> int:
        """The block-Jacobi boundary factor: 2 without row interchanges (natural column order) in
        FP32, 1 without interchanges in FP64, 0 with partial pivoting or none yet."""
        return int(self.L.ibem3d_bj_nopiv(self.ctx))

    def rows_moment_form(self) -> bool:
        """True when the boundary->inclusion rows are integrated in moment form (planar elements)."""
        return bool(self.L.ibem3d_rows_moment_form(self.ctx))

    def inner_iterations(self) -> int:
        """Inner iterations on the 16-bit A of the last SOLVE_BJ_GMRES16 solve (-1: ran without it)."""
        return int(self.L.ibem3d_inner_iterations(self.ctx))

    def residual(self) -> float:
        """Relative residual ||b - A x|| / ||b|| of the current solution (after a compact FP64 solve
        the system is assembled again first)."""
        rel = ctypes.c_double()
        self._chk(self.L.ibem3d_residual(self.ctx, ctypes.byref(rel)), "residual")
        return rel.value

    def getUTVector(self):
        p = self.prob
        U = np.zeros(p.n78)
        T = np.zeros(12 * p.NE)
        X = np.zeros(self.n)
        self._chk(self.L.ibem3d_get_ut(self.ctx, _d(U), _d(T), _d(X)), "getUTVector")
        return U, T, X

    def getFlux(self):
        s = np.zeros((self.prob.nump, 6))
        self._chk(self.L.ibem3d_stress(self.ctx, _d(s)), "getFlux")
        return s

    def average_fields(self):
        """Volume averages (stress, strain: 6 values each, 00 01 02 11 12 22; enclosed volume) from
        the boundary solution -- the effective-property post-step (no reference counterpart)."""
        sig = np.zeros(6)
        eps = np.zeros(6)
        vol = ctypes.c_double()
        self._chk(self.L.ibem3d_average_fields(self.ctx, _d(sig), _d(eps), ctypes.byref(vol)),
                  "average_fields")
        return sig, eps, vol.value

    def set_known_values(self, bc_values, tbc_values):
        """New prescribed values (3 NN displacements, 12 NE tractions) with the known/unknown pattern
        of the problem the engine was created with; clears the assembled system (assemble again,
        then solve with SOLVE_REUSE to keep the factors)."""
        bv = np.ascontiguousarray(bc_values, dtype=np.float64).reshape(-1)
        tv = np.ascontiguousarray(tbc_values, dtype=np.float64).reshape(-1)
        if bv.size != 3 * self.prob.NN or tv.size != 12 * self.prob.NE:
            raise ValueError("set_known_values: expected 3 NN and 12 NE values")
        self._chk(self.L.ibem3d_set_known_values(self.ctx, _d(bv), _d(tv)), "set_known_values")

    def update_known_values(self, bc_values, tbc_values):
        """New prescribed values with b assembled again and the matrix (or its in-place factor)
        kept: solve next with SOLVE_REUSE."""
        bv = np.ascontiguousarray(bc_values, dtype=np.float64).reshape(-1)
        tv = np.ascontiguousarray(tbc_values, dtype=np.float64).reshape(-1)
        if bv.size != 3 * self.prob.NN or tv.size != 12 * self.prob.NE:
            raise ValueError("update_known_values: expected 3 NN and 12 NE values")
        self._chk(self.L.ibem3d_update_known_values(self.ctx, _d(bv), _d(tv)), "update_known_values")

    def average_eigenstrain(self):
        """Volume average of the equivalent eigenstrain (6 values, 00 01 02 11 12 22)."""
        e = np.zeros(6)
        self._chk(self.L.ibem3d_average_eigenstrain(self.ctx, _d(e)), "average_eigenstrain")
        return e

    def system(self):
        A = np.zeros((self.n, self.n), order="F")
        b = np.zeros(self.n)
        self._chk(self.L.ibem3d_get_system(self.ctx, _d(A), _d(b)), "get_system")
        return A, b

    def rhs(self):
        b = np.zeros(self.n)
        self._chk(self.L.ibem3d_get_system(self.ctx, None, _d(b)), "get_system")
        return b

    def timings(self) -> dict:
        names = ctypes.create_string_buffer(8192)
        ms = np.zeros(128)
        cnt = self.L.ibem3d_timings(self.ctx, names, 8192, _d(ms), 128)
        out = {}
        raw = names.raw
        pos = 0
        for i in range(cnt):
            end = raw.index(b"\0", pos)
            out[raw[pos:end].decode()] = float(ms[i])
            pos = end + 1
        return out

    def reset_timings(self):
        self.L.ibem3d_reset_timings(self.ctx)


def run_reference_phases(eng: Engine):
    """iBEMSystemService::run order (iBEMSystemService.h:58-67) then ApplyBC."""
    eng.addBEM()
    eng.addFieldToBEM()
    for o in range(3):
        eng.addFluxEquivalentInclusion(o)
    for o in range(3):
        eng.addFluxEquivalentBEM(o)
    eng.ApplyBC()


PAIRS = [(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)]


def unit_strain(t: int) -> np.ndarray:
    """The symmetric unit tensor of pair t (00 01 02 11 12 22): e_ij = e_ji = 1."""
    i, j = PAIRS[t]
    e = np.zeros((3, 3))
    e[i, j] = e[j, i] = 1.0
    return e


def isotropic_stress(E: float, nu: float, eps6) -> np.ndarray:
    """s_ij = lam d_ij tr(e) + 2 mu e_ij for tensor components in pair order."""
    lam = E * nu / ((1 + nu) * (1 - 2 * nu))
    mu = E / (2 * (1 + nu))
    eps6 = np.asarray(eps6, dtype=np.float64)
    tr = eps6[0] + eps6[3] + eps6[5]
    return np.array([lam * tr * (i == j) + 2 * mu * eps6[t] for t, (i, j) in enumerate(PAIRS)])


def effective_stiffness(prob: Problem, mode: int = SOLVE_MIXED_FP16, tol: float = 1e-10,
                        device: int = 0):
    """Apparent stiffness of an RVE from six affine displacement loads u = e_t . x on every face
    (e_t the unit tensor of pair t), one factorisation reused for the other five loads (no reference
    counterpart). Returns (K_inc, K_bnd, info): 6 x 6 matrices of tensor components (column t = the
    average stress for e_t, pair order 00 01 02 11 12 22), K_inc from the inclusions,
    <s> = C0 : (<e> - <e*>), K_bnd from the boundary tractions, <s> = (1/V) int t x dS; info holds
    the relative residual and refinement iterations of each load."""
    from .generate import macro_strain_bc
    nodes = np.asarray(prob.nodes, dtype=np.float64).reshape(-1, 3)
    conn = np.asarray(prob.conn).reshape(-1, 4)
    bc0, tbc0 = macro_strain_bc(nodes, conn, unit_strain(0))
    p0 = Problem(prob.title, prob.E0, prob.nu0, prob.E1, prob.nu1, nodes, conn, bc0, tbc0,
                 prob.centres, prob.radii, np.zeros((0, 3)))
    eng = Engine(p0, device)
    K_inc = np.zeros((6, 6))
    K_bnd = np.zeros((6, 6))
    info = []
    try:
        for t in range(6):
            if t > 0:
                bc, tbc = macro_strain_bc(nodes, conn, unit_strain(t))
                eng.update_known_values(bc[:, 1], tbc[:, 12:])  # b only; A and its factors kept
            else:
                eng.assemble()
            rel, it = eng.solve(mode | (SOLVE_REUSE if t > 0 else 0), tol)
            info.append((rel, it))
            sig_b, eps_b, _ = eng.average_fields()
            eps_star = eng.average_eigenstrain()
            K_inc[:, t] = isotropic_stress(prob.E0, prob.nu0, eps_b - eps_star)
            K_bnd[:, t] = sig_b
    finally:
        eng.close()
    return K_inc, K_bnd, info
