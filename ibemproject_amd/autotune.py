"""Autotuner of the batched RVE stiffness (ibem3d_batch_stiffness).

The batch's schedule has choices whose best values depend on the RVE sizes and the device: the
LU's product type (FP16 split products, single FP16 products, FP32), the number of classic
refinement steps run inside the graph (too few send loads to the slower host refinement), the
number of RVEs per wave (graph branches that run concurrently), graph capture or direct launches
(enqueued from one or several host threads), and two kernel-level tunables (the tile count from
which trailing updates take the 256 x 256 kernel, and the share of CUs the bulk update leaves to
the concurrent panel). `tune` measures
candidate settings on a sample of the batch that spans its size range, one choice at a time (a
coordinate search: mode, refinement steps, wave size, graph, kernel tunables), keeps a setting only
when every load converged inside the graph or its time improves, and returns the best with the
measurements. The result can be cached in a JSON file keyed by device and workload.
"""
from __future__ import annotations

import json
import os
import time

import numpy as np

from . import engine as E

MODES = {"fp16x3": E.SOLVE_MIXED_FP16, "fp16_1": E.SOLVE_MIXED_FP16_1, "fp32": E.SOLVE_MIXED_FP32}


def sample(problems, k: int = 24):
    """k problems evenly spread over the batch sorted by size (n)."""
    order = sorted(range(len(problems)), key=lambda i: problems[i].n60)
    if len(order) <= k:
        return [problems[i] for i in order]
    idx = np.linspace(0, len(order) - 1, k).round().astype(int)
    return [problems[order[i]] for i in idx]


def measure(probs, cfg: dict, mem_budget_gb: float) -> dict:
    E.set_tuning(E.TUNE_BIG_MIN_TILES, cfg["big_min_tiles"])
    E.set_tuning(E.TUNE_PANEL_CU_PERCENT, cfg["panel_cu_percent"])
    t0 = time.perf_counter()
    r = E.batch_stiffness(probs, MODES[cfg["mode"]], refine_steps=cfg["refine_steps"],
                          wave_max=cfg["wave_max"], mem_budget_gb=mem_budget_gb,
                          use_graph=cfg["use_graph"], host_threads=cfg["host_threads"])
    wall = time.perf_counter() - t0
    return {"cfg": dict(cfg), "wall_s": wall, "host_refined": int(r["host_refined"].sum()),
            "max_rel": float(r["rel"].max()), "times": r["times"]}


DEFAULT = {"mode": "fp16x3", "refine_steps": 4, "wave_max": 16, "use_graph": True, "host_threads": 1,
           "big_min_tiles": 512, "panel_cu_percent": 100}

SPACE = [
    ("mode", ["fp16x3", "fp16_1", "fp32"]),
    ("refine_steps", [2, 3, 4, 5, 6]),
    ("wave_max", [4, 8, 16, 32, 64]),
    ("use_graph", [True, False]),
    ("host_threads", [1, 4, 8, 16]),
    ("big_min_tiles", [128, 256, 512, 1024]),
    ("panel_cu_percent", [0, 50, 100, 200]),
]


def tune(problems, mem_budget_gb: float = 200.0, k: int = 24, log=print) -> dict:
    """Coordinate search over SPACE on a sample of k problems. A candidate's score is its wall time;
    candidates that leave loads to the host refinement are scored with that time too (the host
    refinement is part of the cost). Returns {"best": cfg, "trials": [...], "sample": k}."""
    probs = sample(problems, k)
    best = dict(DEFAULT)
    trials = []
    cur = measure(probs, best, mem_budget_gb)  # also loads every kernel once
    cur = measure(probs, best, mem_budget_gb)
    trials.append(cur)
    log(f"[tune] start {best}: {cur['wall_s']:.3f} s, host-refined loads {cur['host_refined']}")
    for key, values in SPACE:
        for v in values:
            if v == best[key]:
                continue
            cand = dict(best)
            cand[key] = v
            m = measure(probs, cand, mem_budget_gb)
            trials.append(m)
            log(f"[tune] {key}={v}: {m['wall_s']:.3f} s, host-refined loads {m['host_refined']}")
            if m["max_rel"] <= 1e-10 and m["wall_s"] < cur["wall_s"]:
                best, cur = cand, m
        log(f"[tune] {key} -> {best[key]}")
    E.set_tuning(E.TUNE_BIG_MIN_TILES, best["big_min_tiles"])
    E.set_tuning(E.TUNE_PANEL_CU_PERCENT, best["panel_cu_percent"])
    return {"best": best, "best_wall_s": cur["wall_s"], "trials": trials, "sample": len(probs)}


def cached(path: str, key: str):
    try:
        with open(path) as f:
            return json.load(f).get(key)
    except (OSError, ValueError):
        return None


def store(path: str, key: str, result: dict) -> None:
    data = {}
    try:
        with open(path) as f:
            data = json.load(f)
    except (OSError, ValueError):
        pass
    data[key] = result
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    with open(path, "w") as f:
        json.dump(data, f, indent=1)
