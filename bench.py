"""Benchmark: one iBEM elastic-3D RVE solve on one MI355X.

A step is one pass of the hot path over one RVE whose inputs are already resident on the device:
assembly of the dense system (all phases), the solve to a relative residual of 1e-10, and the
stresses at the RVE's 1000 points. The timed solve is the mixed-precision path: an FP32 LU whose
trailing updates run on FP16 MFMA with scaled hi/lo split operands, then FP64 iterative refinement
against the FP64 system. One step with the plain FP64 LU, and one with the FP32-MFMA LU, are timed
beside it for the comparison the metric asks for.

Workload (default): 1000 spheres, radii U[0.02, 0.03], surface gap 0.01, wall clearance 0.02, in a
unit cube meshed with 16 x 16 four-node quads per face; E0 = 1, nu0 = 0.3, E1 = 10, nu1 = 0.25;
traction 1 on x = 1 with symmetric supports; n = 3 * 1538 + 60 * 1000 = 64,614 unknowns.
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

from ibemproject_amd.engine import SOLVE_FP64_LU, SOLVE_MIXED_FP16, SOLVE_MIXED_FP32, Engine  # noqa: E402
from ibemproject_amd.generate import make_workload  # noqa: E402

MI355X_PEAK = {"f16_mfma_tflops": 2500.0, "f32_mfma_tflops": 157.3, "f64_mfma_tflops": 78.6,
               "hbm_gbs": 8000.0}
NB = 1024  # LU block size (solver.hip IBEM_LU_NB)
# HBM bytes per trailing-update launch on rve1000_m16 (FETCH_SIZE x 2.667 + WRITE_SIZE x 1.0 per
# dispatch, averaged over one LU's 125 launches of k_gemm_f16x3_256 / k_gemm_f16x3; the factors make
# the counters read 1 GiB for a 1 GiB copy), from the counter runs recorded in
# profiles/k_gemm_f16x3_256__rve1000_m16.md
GEMM_TRAFFIC_BYTES_PER_LAUNCH = 1.6204e10


def trailing_update_flops(n: int) -> tuple[float, int]:
    """Algorithmic flops of the blocked LU's trailing-matrix updates (2 m^2 kb per step, m the
    trailing size) and the number of split-GEMM launches (k_gemm_f16x3_256, or k_gemm_f16x3 for
    small updates): two per step (the next panel's columns, then the rest), one when the next panel
    is the whole trailing matrix."""
    flops = 0.0
    launches = 0
    for k in range(0, n, NB):
        kb = min(NB, n - k)
        m = n - k - kb
        if m > 0:
            flops += 2.0 * m * m * kb
            launches += 1 + (1 if m > min(NB, m) else 0)
    return flops, launches


def trailing_update_bytes(n: int) -> float:
    """Algorithmic bytes per split-GEMM launch, averaged over one LU: the split operands (hi + lo,
    2 bytes each) read once, C read and written once in FP32 (4 m kb + 4 kb nc + 8 m nc)."""
    total, launches = 0.0, 0
    for k in range(0, n, NB):
        kb = min(NB, n - k)
        m = n - k - kb
        if m <= 0:
            continue
        kb2 = min(NB, m)
        for nc in ([kb2, m - kb2] if m > kb2 else [kb2]):
            total += 4.0 * m * kb + 4.0 * kb * nc + 8.0 * m * nc
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
    eng.reset()
    eng.assemble()
    rel, it = eng.solve(mode, tol)
    s = eng.getFlux()
    return rel, it, s


def cpu_baseline(prob, workload: str):
    """Reference CPU path (oracle/_ref/libibemref.so, the reference's own sources compiled here)
    on a bounded sample of the workload, extrapolated to the full RVE."""
    import ctypes
    lib = os.path.join(ROOT, "oracle", "_ref", "libibemref.so")
    if not os.path.exists(lib):
        return None
    from ibemproject_amd.formats import Problem, write_problem
    L = ctypes.CDLL(lib)
    L.ref_run.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int, ctypes.c_int]
    L.ref_fullpiv_solve.restype = ctypes.c_double
    dp = ctypes.POINTER(ctypes.c_double)
    L.ref_fullpiv_solve.argtypes = [ctypes.c_int, dp, dp, dp]
    # sample: the RVE's first 3 spheres on a 6 x 6-per-face mesh with the same boundary conditions;
    # per-pair costs do not depend on the mesh density, so each phase is scaled by its pair count
    from ibemproject_amd.generate import cube_mesh, symmetric_support_bc
    ns, ms = 3, 6
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
    nlu = 1200
    rng = np.random.default_rng(0)
    A = rng.standard_normal((nlu, nlu)) + nlu * np.eye(nlu)
    A = np.asfortranarray(A)
    b = rng.standard_normal(nlu)
    x = np.zeros(nlu)
    t_lu = L.ref_fullpiv_solve(nlu, A.ctypes.data_as(dp), b.ctypes.data_as(dp), x.ctypes.data_as(dp))
    wall = time.time() - t0
    num = prob.num
    n78 = prob.n78
    ne_s, nn_s = sample.NE, sample.NN
    est = (ph["addBEM"] * (prob.NE * prob.NN) / (ne_s * nn_s)
           + ph["addFieldToBEM"] * (prob.NN * num) / (nn_s * ns)
           + (ph["incl0"] + ph["incl1"] + ph["incl2"]) * (num / ns) ** 2
           + (ph["bem0"] + ph["bem1"] + ph["bem2"]) * (prob.NE * num) / (ne_s * ns)
           + t_lu * (n78 / nlu) ** 3)
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    fair = fair_cpu_lu(prob.n60)
    return {"value": est, "unit": "s per RVE solve (extrapolated)", "cores": cores,
            "kind": "reference",
            "sample": (f"reference assembly phases (OpenMP {cores} threads) on {ns} of the {workload} "
                       f"spheres with a {ms}x{ms}-per-face mesh, {wall:.1f} s wall, each phase scaled "
                       f"by its pair count (element-node, node-sphere, sphere-sphere, "
                       f"element-sphere) to the full RVE; Eigen FullPivLU (1 thread) timed at "
                       f"n={nlu} ({t_lu:.2f} s) and scaled by (n78/{nlu})^3 with n78={n78}; "
                       f"extrapolated, not run end to end"),
            "phases_s": ph, "fullpiv_s_at_n": [nlu, t_lu], "fair_lu": fair}


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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="rve1000_m16")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a 1 GiB device copy (counter calibration) and exit")
    args = ap.parse_args()
    if args.calibrate_copy:
        calibrate_copy()
        return
    if args.gpus != 1:
        print(f"bench.py: this engine solves one RVE on one GPU; --gpus {args.gpus} is not "
              "supported (a multi-GPU run would shard the LU block-cyclically by 1024-column panels)",
              file=sys.stderr)
        sys.exit(2)

    prob = make_workload(args.workload, seed=args.seed)
    eng = Engine(prob)
    n = eng.n
    progress(f"workload {args.workload}: n = {n}")
    MODE = SOLVE_MIXED_FP16
    for i in range(args.warmup):
        run_step(eng, MODE)
        progress(f"warmup step {i + 1}/{args.warmup} done")
    eng.reset_timings()
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    hip.hipDeviceSynchronize()
    t0 = time.perf_counter()
    rels = []
    for _ in range(args.steps):
        rel, it, s = run_step(eng, MODE)
        rels.append((rel, it))
    hip.hipDeviceSynchronize()
    t1 = time.perf_counter()
    progress(f"{args.steps} timed steps done: {(t1 - t0) / args.steps:.3f} s per step")
    tm = eng.timings()
    step_s = (t1 - t0) / args.steps
    flops, launches = trailing_update_flops(n)
    gemm_ms = tm.get("lu.gemm", 0.0) / args.steps
    achieved = flops / (gemm_ms * 1e-3) / 1e12 if gemm_ms > 0 else None
    asm_ms = sum(v for k, v in tm.items() if k.startswith("assemble.")) / args.steps
    lu_ms = tm.get("solve.lu_f32", 0.0) / args.steps

    def one_step(mode, lu_key):
        eng.reset_timings()
        hip.hipDeviceSynchronize()
        f0 = time.perf_counter()
        rel_o, it_o, s_o = run_step(eng, mode)
        hip.hipDeviceSynchronize()
        f1 = time.perf_counter()
        t_o = eng.timings()
        g_o = t_o.get("lu.gemm", 0.0)
        return {"step_s": f1 - f0, "rel_residual": rel_o, "refinement_iterations": it_o,
                "lu_ms": t_o.get(lu_key), "lu_tflops": 2.0 * n ** 3 / 3 / (t_o.get(lu_key) * 1e-3) / 1e12,
                "trailing_gemm_tflops": flops / (g_o * 1e-3) / 1e12 if g_o else None,
                "stress_maxdiff_vs_timed": float(np.abs(s_o - s).max() / np.abs(s_o).max()),
                "phase_ms": t_o}

    fp64 = None if args.no_fp64 else one_step(SOLVE_FP64_LU, "solve.lu_f64")
    if fp64:
        progress(f"FP64 LU step: {fp64['step_s']:.3f} s")
    fp32 = None if args.no_fp64 else one_step(SOLVE_MIXED_FP32, "solve.lu_f32")
    if fp32:
        progress(f"FP32-MFMA mixed step: {fp32['step_s']:.3f} s")

    base = None if args.no_cpu_baseline else cpu_baseline(prob, args.workload)
    if base:
        progress(f"CPU baseline: {base['value']:.1f} s (extrapolated)")
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
        "dtype": "f64+f32 (fp16 mfma split products)",
        "data": "synthetic (seeded RSA spheres, generated mesh)",
        "config": {"workload": args.workload, "spheres": prob.num, "nodes": prob.NN,
                   "elements": prob.NE, "unknowns": n, "points": prob.nump,
                   "solver": "fp32 LU (trailing updates on fp16 MFMA, scaled hi/lo split) + fp64 "
                             "iterative refinement to 1e-10",
                   "parallelism": "single GPU"},
        "rel_residual": rels[-1][0],
        "refinement_iterations": rels[-1][1],
        "assembly_ms": asm_ms,
        "lu_ms": lu_ms,
        "lu_tflops_effective": 2.0 * n ** 3 / 3 / (lu_ms * 1e-3) / 1e12 if lu_ms else None,
        "phase_ms": {k: v / args.steps for k, v in tm.items()},
        "fp64_lu_step": fp64,
        "fp32_mixed_step": fp32,
        "speedup_vs_fp64_lu": (fp64["step_s"] / step_s) if fp64 else None,
        "roofline": {"kernel": "k_gemm_f16x3_256 (LU trailing update: persistent 256x256 tiles, "
                               "direct global-to-LDS loads, v_mfma_f32_16x16x32_f16, three products "
                               "per block; k_gemm_f16x3 128x128 for updates under 512 tiles)",
                     "bound": "mfma", "achieved": achieved,
                     "peak": MI355X_PEAK["f16_mfma_tflops"], "unit": "TFLOP/s",
                     "frac": achieved / MI355X_PEAK["f16_mfma_tflops"] if achieved else None,
                     "traffic": GEMM_TRAFFIC_BYTES_PER_LAUNCH if args.workload == "rve1000_m16" else None,
                     "traffic_unit": "bytes per launch (HBM, counter-measured, averaged over the "
                                     "step's launches)",
                     "traffic_algorithmic": trailing_update_bytes(n),
                     "launches_per_step": launches,
                     "avg_launch_ms": gemm_ms / launches if launches else None,
                     "flops_per_step": flops,
                     "mfma_issued_tflops": 3 * achieved if achieved else None,
                     "note": f"achieved = algorithmic trailing-update flops (2 m^2 {NB} per LU "
                             "step) / the split-GEMM time of the step (both kernels, HIP events "
                             "around each launch); the kernels issue three fp16 MFMA products per "
                             "algorithmic one"},
        "cpu_baseline": base,
    }
    print(json.dumps(out), flush=True)
    eng.close()


if __name__ == "__main__":
    main()
