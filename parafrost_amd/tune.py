"""Autotuner for the simplifier's launch and algorithm choices (pfg_tuning).

A tuning never changes a result, only the time: the occurrence-table build (radix
sort of (literal, reference) keys, or counting with an atomic fill and a sort of each
list), the block count of the clause-streaming passes (their XCD windows) and, apart,
of LCVE's record-building pass, the lane group per elected variable in subsumption,
the LDS arena of a BVE detection block (occupancy against staging rounds), and whether
the LCVE election rounds are launched one by one or replayed from a captured HIP
graph. tune() measures a workload's step
under candidate tunings -- coordinate descent over the knobs in SPACE, every candidate
timed over a few steps after a warm-up -- and records the best, with every
measurement, in tuning/<name>.json next to this module; apply() reads that record back
onto a simplifier. The record names the device it was measured on and its
architecture (gfx target and compute-unit count); apply() takes it on any device of
the same architecture, whatever name the runtime reports for it.
"""
from __future__ import annotations

import json
import os
import time

HERE = os.path.dirname(os.path.abspath(__file__))
TUNING_DIR = os.path.join(HERE, "tuning")

# candidate values per knob, in the order coordinate descent visits them
SPACE = {
    "ot_build": [0, 1, 2],
    "clause_grid": [1024, 2048, 4096, 8192],
    "sub_group": [8, 16, 32, 64],
    "lcve_graph": [0, 1],
    "ve_arena": [8192, 4096],
    "init_grid": [0, 512, 1024],
}


def _device_name() -> str:
    try:
        import torch
        return torch.cuda.get_device_name(0)
    except Exception:  # noqa: BLE001 -- no device: the record still names the workload
        return "unknown"


def _device_arch() -> str:
    """gfx target and compute units, e.g. 'gfx950/256': the key records match on"""
    try:
        import torch
        p = torch.cuda.get_device_properties(0)
        return f"{p.gcnArchName.split(':')[0]}/{p.multi_processor_count}"
    except Exception:  # noqa: BLE001 -- no device
        return "unknown"


def _path(name: str) -> str:
    return os.path.join(TUNING_DIR, f"{name}.json")


def measure(step, sync, reps: int) -> float:
    """seconds per step: one untimed step, then the mean of reps timed steps"""
    step()
    sync()
    t0 = time.perf_counter()
    for _ in range(reps):
        step()
    sync()
    return (time.perf_counter() - t0) / reps


def tune(simplifier, step, sync, name: str, reps: int = 2, passes: int = 1, log=None, workload: str = "") -> dict:
    """Coordinate descent over SPACE on `simplifier` (a Simplifier or BatchSimplifier
    whose step() runs the workload once); sync() waits for the device. Returns and
    records {"tuning": best, "seconds": best time, "trials": [...]}."""
    best = simplifier.tuning()
    best_s = measure(step, sync, reps)
    trials = [{"tuning": dict(best), "seconds": best_s}]
    if log:
        log(f"tune {name}: start {best} {best_s * 1e3:.2f} ms")
    for _ in range(passes):
        for knob, values in SPACE.items():
            for v in values:
                if best[knob] == v:
                    continue
                cand = dict(best, **{knob: v})
                simplifier.set_tuning(**cand)
                s = measure(step, sync, reps)
                trials.append({"tuning": dict(cand), "seconds": s})
                if log:
                    log(f"tune {name}: {knob}={v} {s * 1e3:.2f} ms")
                if s < best_s:
                    best, best_s = cand, s
            simplifier.set_tuning(**best)
    rec = {"name": name, "workload": workload, "device": _device_name(), "arch": _device_arch(), "tuning": best,
           "seconds": best_s, "trials": trials}
    os.makedirs(TUNING_DIR, exist_ok=True)
    with open(_path(name), "w") as f:
        json.dump(rec, f, indent=1)
    return rec


def load(name: str) -> dict | None:
    p = _path(name)
    if not os.path.exists(p):
        return None
    with open(p) as f:
        return json.load(f)


def apply(simplifier, name: str) -> dict | None:
    """Set the recorded tuning of `name` when it was measured on a device of this
    architecture; returns the record applied, or None."""
    rec = load(name)
    if rec is None or rec.get("arch") not in (_device_arch(), "unknown"):
        return None
    knobs = {k: v for k, v in rec["tuning"].items()
             if k in ("ot_build", "clause_grid", "sub_group", "stage_words", "lcve_graph", "ve_arena", "init_grid")}
    simplifier.set_tuning(**knobs)
    return rec
