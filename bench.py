"""Benchmark: one iBEM elastic-3D RVE solve on one MI355X.

A step is one pass of the hot path over one RVE whose inputs are already resident on the device:
assembly of the dense system (all phases), the solve to a relative FP64 residual of 1e-10, and the
stresses at the RVE's 1000 points. The timed solve is the mixed-precision path: an FP32 LU whose
trailing updates run as single FP16 MFMA products of operands scaled per row / column by powers of
two (FP32 accumulation), then FP64 GMRES-IR against the FP64 system. On the 2,500-sphere RVE the
FP64 A takes 205.8 GB of the 288 GB, so the solve uses compact storage: the FP32 factor is built in
A's memory, kept as a per-tile-scaled FP16 copy beside A, and A is assembled again for the FP64
residual. The plain FP64 LU (in place of A) is timed beside it, with the FP16 split-product (three
products) LU; every result states its error against the FP64 LU.

Workload (default, rve2500_m24): 2500 spheres, radii U[0.015, 0.022], surface gap 0.01, wall
clearance 0.02, in a unit cube meshed with 24 x 24 four-node quads per face; E0 = 1, nu0 = 0.3,
E1 = 10, nu1 = 0.25; traction 1 on x = 1 with symmetric supports; n = 3 * 3458 + 60 * 2500 =
160,374 unknowns.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from ibemproject_amd.engine import (SOLVE_BJ_GMRES, SOLVE_BJ_GMRES16, SOLVE_FP64_LU,  # noqa: E402
                                    SOLVE_MIXED_BF16, SOLVE_MIXED_BF16X3, SOLVE_MIXED_FP16, SOLVE_MIXED_FP16_1,
                                    SOLVE_MIXED_FP32, TUNE_RESIDUAL_UPDATE, Engine, get_tuning)
from ibemproject_amd.generate import make_workload  # noqa: E402

MI355X_PEAK = {"f16_mfma_tflops": 2500.0, "f32_mfma_tflops": 157.3, "f64_mfma_tflops": 78.6,
               "hbm_gbs": 8000.0}
NB = 1024  # LU block size (solver.hip IBEM_LU_NB)
DEFAULT_WORKLOAD = "rve2500_m24"
MODES = {"bj_gmres16": SOLVE_BJ_GMRES16, "bj_gmres": SOLVE_BJ_GMRES, "fp16_1": SOLVE_MIXED_FP16_1, "fp16x3": SOLVE_MIXED_FP16, "bf16": SOLVE_MIXED_BF16, "bf16x3": SOLVE_MIXED_BF16X3,
         "fp32": SOLVE_MIXED_FP32}
# the trailing-update kernel of each mode (k_gemm_f16x3_256 / k_gemm_f16x3 instantiated for the
# product type) and its MFMA products per algorithmic multiply-add
GEMM_KERNEL = {"fp16_1": ("k_gemm_f16x3_256<kProdF16> (LU trailing update: persistent 256x256 tiles, "
                          "direct global-to-LDS loads, one v_mfma_f32_16x16x32_f16 product of operands "
                          "scaled per row / column by powers of two; k_gemm_f16x3<..., kProdF16> "
                          "128x128 for updates under 512 tiles)", 1),
               "fp16x3": ("k_gemm_f16x3_256<kProdF16x3> (three FP16 products of hi/lo splits)", 3),
               "bf16": ("k_gemm_f16x3_256<kProdBF16> (one v_mfma_f32_16x16x32_bf16 product)", 1),
               "bf16x3": ("k_gemm_f16x3_256<kProdBF16x3> (three v_mfma_f32_16x16x32_bf16 products of "
                          "hi/lo BF16 splits)", 3),
               "fp32": ("k_gemm_sub<float> (FP32 MFMA)", 1)}
# HBM bytes per trailing-update launch (FETCH_SIZE x f + WRITE_SIZE x 1.0 per dispatch, averaged over
# one LU's launches, f the factor that makes the counter read 1 GiB for a 1 GiB copy in the same
# session), from the counter runs recorded under profiles/
GEMM_TRAFFIC_BYTES_PER_LAUNCH = {("rve1000_m16", "fp16x3"): 1.6204e10, ("rve1000_m16", "fp16_1"): 9.92e9}


def trailing_update_flops(n: int) -> tuple[float, int]:
    """Algorithmic flops of the blocked LU's trailing-matrix updates (2 m^2 kb per step, m the
    trailing size) and the number of trailing-update GEMM launches: two per step (the next panel's
    columns, then the rest), one when the next panel is the whole trailing matrix."""
    flops = 0.0
    launches = 0
    for k in range(0, n, NB):
        kb = min(NB, n - k)
        m = n - k - kb
        if m > 0:
            flops += 2.0 * m * m * kb
            launches += 1 + (1 if m > min(NB, m) else 0)
    return flops, launches


def trailing_update_bytes(n: int, operand_bytes: int) -> float:
    """Algorithmic bytes per trailing-update launch, averaged over one LU: the 16-bit operands
    (operand_bytes per element: 2 for one product, 4 for hi + lo) read once, C read and written
    once in FP32."""
    total, launches = 0.0, 0
    for k in range(0, n, NB):
        kb = min(NB, n - k)
        m = n - k - kb
        if m <= 0:
            continue
        kb2 = min(NB, m)
        for nc in ([kb2, m - kb2] if m > kb2 else [kb2]):
            total += operand_bytes * (m * kb + kb * nc) + 8.0 * m * nc
            launches += 1
    return total / launches if launches else 0.0


def progress(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def calibrate_copy(gib: int = 1, reps: int = 3) -> None:
    """A plain device copy of known size (torch's elementwise kernel, gib GiB read and written per
    launch): the reference reading for the FETCH_SIZE / WRITE_SIZE counters."""
    import torch
    x = torch.ones((gib << 30) // 4, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    for i in range(reps):
        torch.mul(x, 1.0, out=y)
        torch.cuda.synchronize()
        progress(f"copy {i + 1}/{reps}: {gib} GiB read, {gib} GiB written")


def run_step(eng: Engine, mode: int, tol: float = 1e-10):
    """reset, every assembly phase, the solve to tol, the stresses at the postprocess points."""
    eng.reset()
    eng.assemble()
    rel, it = eng.solve(mode, tol)
    s = eng.getFlux()
    return rel, it, s


def cpu_model() -> str:
    try:
        for ln in open("/proc/cpuinfo"):
            if ln.startswith("model name"):
                return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(prob, workload: str, ns: int = 16, ms: int = 3):
    """Reference CPU path (oracle/_ref/libibemref.so, the reference's own sources compiled here) on a
    bounded sample of the workload, extrapolated to the full RVE: the assembly phases on the RVE's
    first ns spheres with an ms x ms-per-face mesh (each phase scaled by its pair count), Eigen's
    FullPivLU (the reference's solver, one thread) at three sizes with the exponent fitted."""
    import ctypes
    lib = os.path.join(ROOT, "oracle", "_ref", "libibemref.so")
    if not os.path.exists(lib):
        return None
    from ibemproject_amd.formats import Problem, write_problem
    from ibemproject_amd.generate import cube_mesh, symmetric_support_bc
    L = ctypes.CDLL(lib)
    L.ref_run.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int, ctypes.c_int]
    L.ref_fullpiv_solve.restype = ctypes.c_double
    dp = ctypes.POINTER(ctypes.c_double)
    L.ref_fullpiv_solve.argtypes = [ctypes.c_int, dp, dp, dp]
    nodes, conn, face = cube_mesh(ms)
    bc, tbc = symmetric_support_bc(nodes, conn, face)
    sample = Problem(prob.title, prob.E0, prob.nu0, prob.E1, prob.nu1, nodes, conn, bc, tbc,
                     prob.centres[:ns], prob.radii[:ns], prob.points[:4])
    t0 = time.time()
    with tempfile.TemporaryDirectory() as d:
        din = os.path.join(d, "in")
        dout = os.path.join(d, "out")
        os.makedirs(dout)
        write_problem(sample, din)
        # the reference prints progress to stdout; keep the benchmark's stdout to its JSON line
        sys.stdout.flush()
        saved = os.dup(1)
        devnull = os.open(os.devnull, os.O_WRONLY)
        os.dup2(devnull, 1)
        try:
            rc = L.ref_run(din.encode(), dout.encode(), 0, 1)
        finally:
            ctypes.CDLL(None).fflush(None)
            os.dup2(saved, 1)
            os.close(devnull)
            os.close(saved)
        if rc != 0:
            return None
        ph = {}
        for ln in open(os.path.join(dout, "times.txt")):
            k, v = ln.split()
            ph[k] = float(v)
    wall_asm = time.time() - t0
    rng = np.random.default_rng(0)
    lu_pts = []
    for nlu in (2048, 2896, 4096):  # every size past the last-level cache of one core's share
        A = np.asfortranarray(rng.standard_normal((nlu, nlu)) + nlu * np.eye(nlu))
        b = rng.standard_normal(nlu)
        x = np.zeros(nlu)
        lu_pts.append((nlu, L.ref_fullpiv_solve(nlu, A.ctypes.data_as(dp), b.ctypes.data_as(dp),
                                                 x.ctypes.data_as(dp))))
    expo, lnc = np.polyfit(np.log([p[0] for p in lu_pts]), np.log([p[1] for p in lu_pts]), 1)
    num, n78 = prob.num, prob.n78
    ne_s, nn_s = sample.NE, sample.NN
    t_lu = float(np.exp(lnc) * n78 ** expo)
    t_lu3 = float(lu_pts[-1][1] * (n78 / lu_pts[-1][0]) ** 3)  # the operation count's n^3 from the largest
    est_asm = (ph["addBEM"] * (prob.NE * prob.NN) / (ne_s * nn_s)
               + ph["addFieldToBEM"] * (prob.NN * num) / (nn_s * ns)
               + (ph["incl0"] + ph["incl1"] + ph["incl2"]) * (num / ns) ** 2
               + (ph["bem0"] + ph["bem1"] + ph["bem2"]) * (prob.NE * num) / (ne_s * ns))
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    fair = fair_cpu_lu(prob.n60)
    return {"value": est_asm + t_lu, "unit": "s per RVE solve (extrapolated)", "cores": cores,
            "cpu_model": cpu_model(), "kind": "reference",
            "sample": (f"reference assembly phases (OpenMP, {cores} threads) on {ns} of the {workload} "
                       f"spheres with a {ms}x{ms}-per-face mesh ({wall_asm:.1f} s wall), each phase "
                       f"scaled by its pair count (element-node, node-sphere, sphere-sphere, "
                       f"element-sphere) to the full RVE: {est_asm:.0f} s; Eigen FullPivLU (the "
                       f"reference's solver, 1 thread) timed at n = "
                       f"{', '.join(str(p[0]) for p in lu_pts)}, fitted t ~ n^{expo:.2f} and "
                       f"extrapolated to the reference's n78 = {n78} unknowns: {t_lu:.3g} s; not run "
                       f"end to end"),
            "phases_s": ph, "assembly_s_extrapolated": est_asm,
            "fullpiv_lu": {"points": lu_pts, "exponent": expo, "s_at_n78": t_lu, "n78": n78,
                           "s_at_n78_by_n3": t_lu3},
            "fair_lu": fair}


def fair_cpu_lu(n60: int, nlu: int = 6000) -> dict:
    """The "fair CPU" LU beside the reference's unthreaded full-pivot one: LAPACK's blocked
    partial-pivot getrf (scipy, all host threads) at n = nlu, scaled by (n60 / nlu)^3 to the
    60-unknown system the GPU factors."""
    from scipy.linalg import lu_factor
    from threadpoolctl import threadpool_limits
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    rng = np.random.default_rng(1)
    A = np.asfortranarray(rng.standard_normal((nlu, nlu)) + nlu * np.eye(nlu))
    with threadpool_limits(limits=threads):  # the BLAS pool would otherwise size itself to the host
        lu_factor(A[:500, :500].copy(order="F"))  # start the threads
        t0 = time.perf_counter()
        lu_factor(A, overwrite_a=True, check_finite=False)
        t = time.perf_counter() - t0
    return {"n": nlu, "s": t, "gflops": 2.0 * nlu ** 3 / 3 / t / 1e9, "threads": threads,
            "s_at_n60": t * (n60 / nlu) ** 3, "n60": n60,
            "note": "LAPACK dgetrf (blocked, partial pivoting) through scipy, extrapolated by n^3"}


def timed(eng: Engine, hip, fn):
    eng.reset_timings()
    hip.hipDeviceSynchronize()
    t0 = time.perf_counter()
    out = fn()
    hip.hipDeviceSynchronize()
    return out, time.perf_counter() - t0, eng.timings()


def side_step(eng: Engine, hip, mode: int, lu_key: str, flops: float, x_ref, s_ref):
    """One untimed-region step of another mode: its time, LU rate and error against x_ref / s_ref."""
    (rel, it, s), t, tm = timed(eng, hip, lambda: run_step(eng, mode))
    if rel < 0:  # compact FP64: A holds the factor; assembled again for the residual (not timed)
        rel = eng.residual()
    _, _, x = eng.getUTVector()
    lu = tm.get(lu_key)
    g = tm.get("lu.gemm", 0.0)
    return {"step_s": t, "rel_residual": rel, "refinement_iterations": it, "lu_ms": lu,
            "lu_tflops": 2.0 * len(x) ** 3 / 3 / (lu * 1e-3) / 1e12 if lu else None,
            "trailing_gemm_tflops": flops / (g * 1e-3) / 1e12 if g else None,
            "x_relerr_vs_ref": float(np.linalg.norm(x - x_ref) / np.linalg.norm(x_ref)),
            "stress_relerr_vs_ref": float(np.abs(s - s_ref).max() / np.abs(s_ref).max()),
            "phase_ms": tm}, x, s


def contrasts(args, hip):
    """Particle/matrix stiffness contrasts E1/E0 = 2, 10, 100 on the workload's geometry: the timed
    mode and the plain FP64 LU on each, with the error of the first against the second."""
    out = []
    for ratio in (2.0, 10.0, 100.0):
        prob = make_workload(args.workload, seed=args.seed, E1=ratio)
        eng = Engine(prob)
        (rel, it, s), t, tm = timed(eng, hip, lambda: run_step(eng, MODES[args.mode]))
        _, _, x = eng.getUTVector()
        inner = eng.inner_iterations() if args.mode == "bj_gmres16" else None
        (rel64, _, s64), t64, _ = timed(eng, hip, lambda: run_step(eng, SOLVE_FP64_LU))
        rel64 = eng.residual()
        _, _, x64 = eng.getUTVector()
        eng.close()
        rec = {"E1_over_E0": ratio, "nu1": prob.nu1, "step_s": t, "rel_residual": rel,
               "iterations": it, "inner_iterations": inner, "fp64_step_s": t64, "fp64_rel_residual": rel64,
               "x_relerr_vs_fp64_lu": float(np.linalg.norm(x - x64) / np.linalg.norm(x64)),
               "stress_relerr_vs_fp64_lu": float(np.abs(s - s64).max() / np.abs(s64).max())}
        progress(f"contrast {ratio:g}: {rec}")
        out.append(rec)
    print(json.dumps({"workload": args.workload, "mode": args.mode, "contrasts": out}), flush=True)


def batch_leg(args, hip):
    """The batched workload: 512 seeded RVEs (8-256 spheres, 8 x 8 quads per face, random moduli),
    the six-load apparent stiffness of each on one factorisation, waves of RVEs captured as one
    graph each; the schedule from the autotuner (tuned on a 24-RVE sample, outside the timed
    region). One step = the whole batch."""
    from ibemproject_amd import autotune
    from ibemproject_amd.engine import batch_stiffness
    from ibemproject_amd.generate import make_batch
    count = args.batch_count
    t0 = time.perf_counter()
    items = make_batch(args.seed, count, workers=min(16, os.cpu_count() or 1))
    probs = [p for p, _ in items]
    dirs = np.array([e for _, e in items])
    progress(f"batch of {count} generated in {time.perf_counter() - t0:.1f} s")
    budget = args.batch_budget_gb
    tuned = autotune.tune(probs, mem_budget_gb=budget, log=progress)
    cfg = tuned["best"]
    mode = autotune.MODES[cfg["mode"]]
    kw = dict(refine_steps=cfg["refine_steps"], wave_max=cfg["wave_max"], mem_budget_gb=budget,
              use_graph=cfg["use_graph"], host_threads=cfg["host_threads"])
    for i in range(args.warmup):
        batch_stiffness(probs, mode, **kw)
        progress(f"warmup pass {i + 1}/{args.warmup} done")
    walls, res = [], None
    for i in range(args.steps):
        hip.hipDeviceSynchronize()
        w0 = time.perf_counter()
        res = batch_stiffness(probs, mode, **kw)
        hip.hipDeviceSynchronize()
        walls.append(time.perf_counter() - w0)
        progress(f"pass {i + 1}/{args.steps}: {walls[-1]:.2f} s")
    wall = sum(walls) / len(walls)
    n = np.array([p.n60 for p in probs], dtype=float)
    lu_flops = float((2.0 * n ** 3 / 3.0).sum())
    C = res["C"]
    resp = np.einsum("kij,kj->ki", C, dirs)  # stress response to each RVE's random unit macro strain
    energy = np.einsum("ki,ki->k", resp, dirs)
    out = {
        "metric": "batched RVE apparent stiffness (six loads per RVE, one factorisation each)",
        "value": count / wall, "unit": "RVE/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": wall * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None,
        "dtype": f"f64 assembly and residual; {cfg['mode']} LU; f64 refinement",
        "data": "synthetic (seeded RSA spheres, random moduli and macro-strain directions)",
        "config": {"workload": f"batch{count}", "rves": count, "spheres": [int(min(p.num for p in probs)),
                                                                           int(max(p.num for p in probs))],
                   "unknowns": [int(n.min()), int(n.max())], "sum_unknowns_sq_GB_fp64": float(8 * (n ** 2).sum() / 1e9),
                   "schedule": cfg, "parallelism": "single GPU"},
        "tuning": {"sample": tuned["sample"], "trials": len(tuned["trials"]),
                   "best_sample_wall_s": tuned["best_wall_s"]},
        "times": res["times"], "max_rel_residual": float(res["rel"].max()),
        "host_refined_loads": int(res["host_refined"].sum()),
        "lu_tflops_effective": lu_flops / wall / 1e12,
        # whole-batch rate: the algorithmic 2n^3/3 of every RVE's LU over the batch's wall time (all
        # kernels of the batch, assembly and refinement included), against the LU's product peak
        "roofline": {"bound": "mfma", "achieved": lu_flops / wall / 1e12,
                     "peak": 157.3 if cfg["mode"] == "fp32" else 2500.0, "unit": "TFLOP/s",
                     "frac": lu_flops / wall / 1e12 / (157.3 if cfg["mode"] == "fp32" else 2500.0),
                     "traffic": None, "scope": "whole batch, all kernels"},
        "stiffness": {"C11_mean": float(C[:, 0, 0].mean()), "C11_min": float(C[:, 0, 0].min()),
                      "C11_max": float(C[:, 0, 0].max()),
                      "directional_modulus_mean": float(energy.mean()),
                      "symmetry_max": float(np.abs(C - C.transpose(0, 2, 1)).max() / np.abs(C).max())},
    }
    print(json.dumps(out), flush=True)


def learned_leg(args, hip):
    """The learned initial guess: a network trained (by its own HIP kernels) on the eigenstrains of
    seeded batched RVEs (8-256 spheres, six unit loads each, from the batched solver), then a
    held-out RVE of the 2,500-sphere geometry (contrast 10, traction load, seed != the training
    seeds) solved by block-Jacobi GMRES from zero and from the network's guess (BF16 inference), the
    mixed LU beside them; every solution is compared with the LU's."""
    from ibemproject_amd import learned as Lr
    from ibemproject_amd.generate import make_batch
    t0 = time.perf_counter()
    probs = [p for p, _ in make_batch(args.seed + 1000, args.train_rves, workers=min(16, os.cpu_count() or 1))]
    kw = dict(refine_steps=3, wave_max=32, mem_budget_gb=200.0, use_graph=False, host_threads=8)
    edge, node, tgt, r = Lr.training_set(probs, **kw)
    t_data = time.perf_counter() - t0
    progress(f"training set: {len(probs)} RVEs, {len(tgt)} spheres in {t_data:.1f} s")
    g = Lr.Gnn(seed=3)
    g.set_data(edge, node, tgt)
    t1 = time.perf_counter()
    hist = g.train(epochs=args.train_epochs, batch=256, lr=2e-3)
    hip.hipDeviceSynchronize()
    t_train = time.perf_counter() - t1
    progress(f"trained {args.train_epochs} epochs in {t_train:.1f} s, loss {hist[0]:.4g} -> {hist[-1]:.4g}")
    prob = make_workload(args.workload if args.workload != "learned" else DEFAULT_WORKLOAD, seed=args.seed + 7)
    eng = Engine(prob)
    n = eng.n
    run_step(eng, SOLVE_BJ_GMRES)  # loads every kernel, allocates the workspaces
    (rel_c, it_c, s_c), t_cold, tm_c = timed(eng, hip, lambda: run_step(eng, SOLVE_BJ_GMRES))
    _, _, x_c = eng.getUTVector()
    t2 = time.perf_counter()
    _, e1, n1 = Lr.graph(prob)
    y = g.predict(e1, n1, bf16=True)
    eps = Lr.macro_strain_of_uniaxial_traction(prob.E0, prob.nu0)
    xg = Lr.guess(y, prob.radii, eps)
    t_inf = time.perf_counter() - t2
    xd = Lr.guess(np.tile(Lr.dilute_response(prob.E0, prob.nu0, prob.E1, prob.nu1), (prob.num, 1)),
                  prob.radii, eps)

    def warm(x0):
        eng.set_initial_guess(x0)
        return run_step(eng, SOLVE_BJ_GMRES)
    (rel_w, it_w, s_w), t_warm, _ = timed(eng, hip, lambda: warm(xg))
    r0, alpha = eng.warm_start_info()
    fb = eng.fell_back()
    _, _, x_w = eng.getUTVector()
    (_, it_d, _), _, _ = timed(eng, hip, lambda: warm(xd))
    r0_d, alpha_d = eng.warm_start_info()
    # the accuracy a warm start needs on this operator: the LU's own eigenstrains with relative noise
    nb = 3 * prob.NN
    rng = np.random.default_rng(5)
    noisy = {}
    for lvl in (1e-3, 1e-2):
        _, itn, _ = warm(x_c[nb:] * (1 + lvl * rng.standard_normal(n - nb)))
        noisy[str(lvl)] = {"iterations": itn, "start_rel_residual": eng.warm_start_info()[0]}
    eng.set_initial_guess(None)
    (rel_l, it_l, s_l), t_lu, _ = timed(eng, hip, lambda: run_step(eng, SOLVE_MIXED_FP16_1))
    _, _, x_l = eng.getUTVector()
    eng.close()

    def by_order(x):
        o = np.tile(Lr.ORDER, prob.num)
        return [float(np.linalg.norm(x[o == k] - x_l[nb:][o == k]) / np.linalg.norm(x_l[nb:][o == k]))
                for k in (0, 1, 2)]
    out = {
        "metric": "learned initial guess: block-Jacobi GMRES iterations and time to 1e-10, with and "
                  "without the network, beside the mixed LU",
        "value": t_warm + t_inf, "unit": "s", "n_gpus": 1, "higher_is_better": False,
        "config": {"workload": prob.title, "unknowns": n, "spheres": prob.num, "contrast": prob.E1 / prob.E0,
                   "train_rves": len(probs), "train_spheres": int(len(tgt)), "epochs": args.train_epochs},
        "training": {"data_s": t_data, "train_s": t_train, "loss_first": float(hist[0]),
                     "loss_last": float(hist[-1]), "max_rel_residual": float(r["rel"].max())},
        "inference_s": t_inf,
        "guess_err_vs_lu_by_order": {"network": by_order(xg), "dilute": by_order(xd)},
        "warm_start": {"alpha": alpha, "start_rel_residual": r0, "dilute_alpha": alpha_d,
                       "dilute_start_rel_residual": r0_d, "dilute_iterations": it_d},
        "lu_with_noise": noisy,
        "cold": {"step_s": t_cold, "iterations": it_c, "rel_residual": rel_c,
                 "x_relerr_vs_lu": float(np.linalg.norm(x_c - x_l) / np.linalg.norm(x_l))},
        "warm": {"step_s": t_warm, "iterations": it_w, "rel_residual": rel_w, "fell_back": fb,
                 "x_relerr_vs_lu": float(np.linalg.norm(x_w - x_l) / np.linalg.norm(x_l)),
                 "stress_relerr_vs_lu": float(np.abs(s_w - s_l).max() / np.abs(s_l).max())},
        "mixed_lu": {"step_s": t_lu, "iterations": it_l, "rel_residual": rel_l},
    }
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default=DEFAULT_WORKLOAD)
    ap.add_argument("--mode", default="bj_gmres16", choices=sorted(MODES))
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-compare", action="store_true", help="skip the FP16 split-product step")
    ap.add_argument("--contrasts", action="store_true",
                    help="run the E1/E0 = 2, 10, 100 sweep (timed mode and FP64 LU) and exit")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a 1 GiB device copy (counter calibration) and exit")
    ap.add_argument("--batch-count", type=int, default=512)
    ap.add_argument("--learned", action="store_true",
                    help="train the eigenstrain network and compare its warm start (then exit)")
    ap.add_argument("--train-rves", type=int, default=2000)
    ap.add_argument("--train-epochs", type=int, default=200)
    ap.add_argument("--batch-budget-gb", type=float, default=240.0)
    args = ap.parse_args()
    if args.calibrate_copy:
        calibrate_copy()
        return
    if args.gpus != 1:
        print(f"bench.py: this engine solves one RVE on one GPU; --gpus {args.gpus} is not "
              "supported (a multi-GPU run would shard the LU block-cyclically by 1024-column panels)",
              file=sys.stderr)
        sys.exit(2)
    import ctypes
    if args.workload.startswith("batch") or args.learned:
        # the batch's latency-bound small solves overlap across streams: give the process 16
        # hardware queues (HIP's default is 4) before the runtime starts
        os.environ.setdefault("GPU_MAX_HW_QUEUES", "16")
    hip = ctypes.CDLL("libamdhip64.so")
    if args.contrasts:
        contrasts(args, hip)
        return
    if args.workload.startswith("batch"):
        batch_leg(args, hip)
        return
    if args.learned:
        learned_leg(args, hip)
        return

    prob = make_workload(args.workload, seed=args.seed)
    eng = Engine(prob)
    n = eng.n
    progress(f"workload {args.workload}: n = {n}, mode {args.mode}")
    MODE = MODES[args.mode]
    for i in range(args.warmup):
        run_step(eng, MODE)
        progress(f"warmup step {i + 1}/{args.warmup} done")
    eng.reset_timings()
    hip.hipDeviceSynchronize()
    t0 = time.perf_counter()
    rels = []
    for _ in range(args.steps):
        rel, it, s = run_step(eng, MODE)
        rels.append((rel, it))
    hip.hipDeviceSynchronize()
    t1 = time.perf_counter()
    step_s = (t1 - t0) / args.steps
    progress(f"{args.steps} timed steps done: {step_s:.3f} s per step")
    tm = eng.timings()
    _, _, x = eng.getUTVector()
    true_rel = eng.residual()  # b - A x of the last step by an FP64 product, outside the timed region
    inner_its = eng.inner_iterations() if args.mode == "bj_gmres16" else 0  # before the side steps
    flops, launches = trailing_update_flops(n)
    gemm_ms = tm.get("lu.gemm", 0.0) / args.steps
    achieved = flops / (gemm_ms * 1e-3) / 1e12 if gemm_ms > 0 else None
    asm_ms = sum(v for k, v in tm.items() if k.startswith("assemble.")) / args.steps
    lu_ms = tm.get("solve.lu_f32", 0.0) / args.steps

    fp64 = None
    if not args.no_fp64:
        fp64, x64, s64 = side_step(eng, hip, SOLVE_FP64_LU, "solve.lu_f64", flops, x, s)
        progress(f"FP64 LU step: {fp64['step_s']:.3f} s")
    bj = args.mode in ("bj_gmres", "bj_gmres16")
    cmp_mode = "fp16_1" if bj or args.mode == "fp16x3" else "fp16x3"
    other = None
    if not args.no_compare:
        ref_x, ref_s = (x64, s64) if fp64 else (x, s)
        other, _, _ = side_step(eng, hip, MODES[cmp_mode], "solve.lu_f32", flops, ref_x, ref_s)
        progress(f"{cmp_mode} step: {other['step_s']:.3f} s")

    base = None if args.no_cpu_baseline else cpu_baseline(prob, args.workload)
    if base:
        progress(f"CPU baseline: {base['value']:.4g} s (extrapolated)")
    gemv = None
    if bj:
        # the timed step's dominant kernel: a matrix-vector product over A. FP64 (k_gemv_part +
        # k_gemv_done): one per outer GMRES iteration and one per restart cycle (its true residual;
        # the cycle that meets the tolerance checks it from its own products instead, unless that
        # is switched off); FP16 copy (k_gemv16_part + k_gemv_done_n, bj_gmres16): one per inner
        # iteration
        its = rels[-1][1]
        inner = inner_its
        cycles = max(1, -(-its // 100))
        g64_launches = its + cycles - (1 if get_tuning(TUNE_RESIDUAL_UPDATE) else 0)
        g64_ms = tm.get("solve.residual", 0.0) / args.steps / g64_launches
        g64_bytes = 8.0 * n * n + 8.0 * 33 * n  # A once, x per column chunk, partial sums
        g16_ms = tm.get("solve.gemv_m16", 0.0) / args.steps / inner if inner > 0 else 0.0
        ld16 = (n + 255) // 256 * 256
        g16_bytes = 2.0 * ld16 * n + 4.0 * (ld16 // 256) * ((n + 31) // 32) + 8.0 * 33 * n
        use16 = inner > 0 and g16_ms * inner >= g64_ms * g64_launches
        gms, gbytes, glaunch = (g16_ms, g16_bytes, inner) if use16 else (g64_ms, g64_bytes, g64_launches)
        gemv = {"kernel": ("k_gemv16_part + k_gemv_done_n (FP16 copy of A, power-of-two scale per 256x32 "
                           "tile: eight rows per thread, 16-byte loads, FP32 sums per tile, FP64 across)"
                           if use16 else
                           "k_gemv_part + k_gemv_done (FP64 A x: 32 column chunks, two rows per thread, "
                           "16-byte loads; fixed-order partial sums)"),
                "bound": "hbm", "achieved": gbytes / (gms * 1e-3) / 1e9, "peak": MI355X_PEAK["hbm_gbs"],
                "unit": "GB/s", "frac": gbytes / (gms * 1e-3) / 1e9 / MI355X_PEAK["hbm_gbs"],
                "traffic": None, "traffic_algorithmic": gbytes, "launches_per_step": glaunch,
                "avg_launch_ms": gms,
                "fp64_product": {"avg_launch_ms": g64_ms, "launches_per_step": g64_launches,
                                 "achieved_gbs": g64_bytes / (g64_ms * 1e-3) / 1e9 if g64_ms else None},
                "fp16_product": ({"avg_launch_ms": g16_ms, "launches_per_step": inner,
                                  "achieved_gbs": g16_bytes / (g16_ms * 1e-3) / 1e9}
                                 if inner > 0 else None),
                "note": "achieved = algorithmic bytes per product (the matrix read once, plus x and the "
                        "partial sums) / the average product time (HIP events around each product on "
                        "the context's stream)"}
    lu_side = other if bj else None
    lu_mode = cmp_mode if bj else args.mode
    lu_gemm_ms = (lu_side["phase_ms"].get("lu.gemm", 0.0) if lu_side else gemm_ms)
    lu_achieved = flops / (lu_gemm_ms * 1e-3) / 1e12 if lu_gemm_ms else None
    kname, prods = GEMM_KERNEL[lu_mode]
    op_bytes = 4 if lu_mode in ("fp16x3", "bf16x3", "fp32") else 2
    peak = MI355X_PEAK["f32_mfma_tflops"] if lu_mode == "fp32" else MI355X_PEAK["f16_mfma_tflops"]
    lu_roof = {"kernel": kname, "bound": "mfma", "achieved": lu_achieved, "peak": peak,
               "unit": "TFLOP/s", "frac": lu_achieved / peak if lu_achieved else None,
               "traffic": GEMM_TRAFFIC_BYTES_PER_LAUNCH.get((args.workload, lu_mode)),
               "traffic_unit": "bytes per launch (HBM, counter-measured, averaged over the LU's launches)",
               "traffic_algorithmic": trailing_update_bytes(n, op_bytes),
               "launches_per_step": launches,
               "avg_launch_ms": lu_gemm_ms / launches if launches else None,
               "flops_per_step": flops,
               "mfma_issued_tflops": prods * lu_achieved if lu_achieved else None,
               "note": f"{lu_mode} LU step: achieved = algorithmic trailing-update flops (2 m^2 {NB} "
                       "per LU step) / the trailing-update GEMM time of the step (HIP events around "
                       "each launch on the stream it runs on)"}
    solver = {"bj_gmres": "fp64 GMRES right-preconditioned by block Jacobi (boundary-block LU, inverse "
                          "60x60 sphere blocks) to 1e-10, mixed LU as fallback",
              "bj_gmres16": "fp64 flexible GMRES (fp64 residuals and products) whose directions come from "
                            "inner block-Jacobi GMRES on a per-tile-scaled fp16 copy of A, to 1e-10; mixed "
                            "LU as fallback"}.get(args.mode, f"{args.mode} LU + fp64 GMRES-IR / iterative "
                                                             "refinement to 1e-10")
    dtype = {"bj_gmres": "f64 (assembly, block-Jacobi preconditioner, GMRES, residual)",
             "bj_gmres16": ("f64 assembly, preconditioner, outer GMRES and residual; inner products with an "
                            "f16 copy of A (per-tile pow2 scaling, f32 sums)"),
             "fp16_1": ("f64 assembly, residual and solution update; f32 LU with f16 mfma products "
                        "(one per multiply-add, per-row/column pow2 scaling) and f32 accumulation; "
                        "f16 factor copy (per-tile pow2 scaling)")}.get(args.mode, args.mode)
    out = {
        "metric": "RVE solve time at 1e-10 FP64 residual vs plain FP64 LU; assembly/LU TFLOP/s",
        "value": step_s,
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": step_s * 1e3,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": dtype,
        "data": "synthetic (seeded RSA spheres, generated mesh)",
        "config": {"workload": args.workload, "spheres": prob.num, "nodes": prob.NN,
                   "elements": prob.NE, "unknowns": n, "points": prob.nump,
                   "solver": solver, "parallelism": "single GPU"},
        "rel_residual": rels[-1][0],
        "true_rel_residual": true_rel,
        "iterations": rels[-1][1],
        "inner_iterations": inner_its if args.mode == "bj_gmres16" else None,
        "error_vs_fp64_lu": ({"x_rel": float(np.linalg.norm(x - x64) / np.linalg.norm(x64)),
                              "stress_rel": float(np.abs(s - s64).max() / np.abs(s64).max())}
                             if fp64 else None),
        "assembly_ms": asm_ms,
        "rows_form": "moment" if eng.rows_moment_form() else "pointwise",
        "lu_ms": lu_ms or None,
        "lu_tflops_effective": 2.0 * n ** 3 / 3 / (lu_ms * 1e-3) / 1e12 if lu_ms else None,
        "phase_ms": {k: v / args.steps for k, v in tm.items()},
        "fp64_lu_step": fp64,
        f"{cmp_mode}_step": other,
        "speedup_vs_fp64_lu": (fp64["step_s"] / step_s) if fp64 else None,
        "roofline": gemv if gemv else lu_roof,
        "lu_roofline": lu_roof if gemv else None,
        "cpu_baseline": base,
    }
    print(json.dumps(out), flush=True)
    eng.close()


if __name__ == "__main__":
    main()
