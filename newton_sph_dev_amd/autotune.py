"""Autotuner for the SPH step's launch configuration.

Tunables (none of them changes a single bit of the results, which the GPU tests check):
  search         neighbour-search kernel variant: list quads per particle staged in LDS,
                 which also sets its workgroups per CU (SolverSPH.SEARCH_VARIANTS)
  gather_block   workgroup size of the gather kernels: a 256-particle chunk is split over
                 1, 2 or 4 workgroups of 256 / 128 / 64 threads (the work partition: every
                 world is cut into 256-particle chunks, so worlds of any size spread evenly)
  mode           "eager" with a host convergence check every `check_every` iterations, or
                 "graph": the step replayed from HIP graphs with `graph_slots` captured Jacobi
                 iterations (one host read after the slots; eager iterations finish the solve
                 if the slots were too few)
  solve          the pressure solve's schedule (batches of at least MIN_WORLDS_SOLVE worlds):
                 "passes" (one launch per pass over all worlds) or "worlds" (one persistent
                 launch whose workgroups of `solve_block` threads take whole worlds from a
                 queue, largest first, after `solve_passes` per-pass Jacobi iterations), eager
                 or captured: the work partition between worlds of uneven size
The tuner times each candidate over a few steps of the caller's own solver and state
(ping-pong, so the physics keeps advancing) and keeps the fastest, in three rounds: the search
variant first (eager, the current gather block), then gather block x mode with that variant,
then the solve schedules with the best of those.
Results are persisted as JSON keyed by GPU architecture, fluid-count bucket and world count.
"""

from __future__ import annotations

import json
import os
import time
from pathlib import Path

import numpy as np
import torch

DEFAULT_CACHE = Path(os.environ.get("NSPH_AUTOTUNE_CACHE",
                                    Path(__file__).resolve().parent / "data" / "autotune.json"))


def _arch() -> str:
    try:
        return torch.cuda.get_device_properties(0).gcnArchName.split(":")[0]
    except Exception:
        return "unknown"


def workload_key(solver) -> str:
    nf = max(solver.fluid_count, 1)
    bucket = 1 << int(np.ceil(np.log2(nf)))
    return f"{_arch()}|fluid<={bucket}|worlds={solver.W}|warm={solver.warm_start}"


def search_candidates(gather_block: int) -> list[dict]:
    return [{"gather_block": gather_block, "search": v, "mode": "eager", "check_every": 1}
            for v in _variants()]


def _variants():
    from .solvers import SolverSPH

    return SolverSPH.SEARCH_VARIANTS


def candidates(mean_iterations: float, search=None) -> list[dict]:
    k = max(2, int(np.ceil(mean_iterations)))
    search = search if search is not None else _variants()[0]
    out = []
    for gb in (256, 128, 64):
        out.append({"gather_block": gb, "search": search, "mode": "eager", "check_every": 1})
        out.append({"gather_block": gb, "search": search, "mode": "eager", "check_every": 2})
        for slots in sorted({k, k + 1}):
            out.append({"gather_block": gb, "search": search, "mode": "graph", "graph_slots": slots})
    return out


# a whole-world schedule needs enough worlds to fill the GPU's workgroup slots
MIN_WORLDS_SOLVE = 64


def solve_candidates(base: dict, solver) -> list[dict]:
    """Whole-world schedules on top of `base` (the best per-pass configuration)."""
    if solver.W < MIN_WORLDS_SOLVE:
        return []
    out = []
    for block in (256, 512, 1024):
        for k in sorted({0, max(solver.min_iters, 1)}):
            for mode in ("eager", "graph"):
                c = {"gather_block": base["gather_block"], "search": base.get("search", _variants()[0]),
                     "mode": mode, "solve": "worlds", "solve_block": block, "solve_passes": k}
                if mode == "graph":
                    c["graph_slots"] = None
                else:
                    c["check_every"] = 1
                out.append(c)
    return out


def apply(solver, cfg: dict) -> None:
    solver.set_launch(cfg["gather_block"], int(cfg.get("search", _variants()[0])))
    solver.set_solve(cfg.get("solve", "passes"), cfg.get("solve_block", 256), cfg.get("solve_passes", 0))
    if cfg["mode"] == "graph":
        solver.use_graph = True
        solver.graph_slots = cfg["graph_slots"]
    else:
        solver.use_graph = False
        solver.check_every = cfg["check_every"]
    solver.clear_graphs()


def _time(solver, a, b, dt, steps) -> tuple[float, object, object]:
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        solver.step(a, b, None, None, dt)
        a, b = b, a
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / steps, a, b


def tune(solver, state_a, state_b, dt: float, steps: int = 6, cache: Path | None = DEFAULT_CACHE,
         force: bool = False, log=None) -> dict:
    """Pick the fastest launch configuration for this solver and apply it. Returns the
    chosen config with its measured ms per step (and the per-candidate table)."""
    key = workload_key(solver)
    table = {}
    if cache is not None and Path(cache).exists():
        try:
            table = json.loads(Path(cache).read_text())
        except ValueError:
            table = {}
    if not force and key in table:
        apply(solver, table[key])
        return table[key]
    a, b = state_a, state_b
    # iteration estimate from one eager step
    solver.use_graph = False
    solver.step(a, b, None, None, dt)
    a, b = b, a
    its = solver.last_report["max_iterations"] if solver.last_report else 2
    results = []

    def run(cands):
        nonlocal a, b
        out = []
        for cfg in cands:
            apply(solver, cfg)
            _, a, b = _time(solver, a, b, dt, 1)  # warm (captures graphs)
            ms, a, b = _time(solver, a, b, dt, steps)
            out.append({**cfg, "ms_per_step": ms * 1e3})
            if log:
                log(f"autotune {cfg} -> {ms * 1e3:.3f} ms")
        results.extend(out)
        return min(out, key=lambda r: r["ms_per_step"])

    search = run(search_candidates(solver.gather_block))["search"]
    base = run(candidates(its, search))
    sc = solve_candidates(base, solver)
    if sc:
        run(sc)
    best = min(results, key=lambda r: r["ms_per_step"])
    chosen = {k: v for k, v in best.items()}
    apply(solver, chosen)
    chosen["candidates"] = results
    if cache is not None:
        table[key] = chosen
        Path(cache).parent.mkdir(parents=True, exist_ok=True)
        Path(cache).write_text(json.dumps(table, indent=1))
    return chosen
