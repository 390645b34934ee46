"""Benchmark: SPH dam break, 1,048,576 fluid particles, one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W]
                    [--workload dam_break_1m|dam_break_32k|batched_4096|batched_1024|xpbd_granular_1m|semi_implicit_1m]
                    [--warm-start zero|prev] [--no-cpu-baseline] [--no-extra]

Prints ONE JSON line. A step = one SolverSPH.step (hash/sort, neighbour lists, density,
advection, IISPH pressure solve to 0.1 % mean compression, integration) over the whole
scene; value = fluid particle-updates/s over the K timed steps (inputs resident in HBM).
roofline: the kernel class with the largest GPU time, its algorithmic bytes per launch
(compulsory unique bytes, each array once, neighbour re-reads not counted, list bytes
counted) divided by its mean launch time from HIP events recorded on the solver's stream
during the timed steps. cpu_baseline: the same-order FP32 CPU restatement (oracle) of the
same step, OpenMP over all host cores, on a bounded sample of steps.
other_configs (default run of the 1M workload, skipped with --no-extra): the other benchmark
configurations, each with its own config and numbers -- the 32,768-particle dam break, 4096
batched worlds, the learned warm start against the zero and previous-pressure warm starts on
the 1M dam break at dt 1e-4 and 1e-3, and 1024 batched worlds with the learned warm start
replayed from HIP graphs (latency percentiles).
"""

from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import platform
import sys
import time

import numpy as np

HBM_PEAK_GBS = 8000.0
# What limits each kernel class, from rocprofv3 counters on MI355X (profiles/pmc_dam_break_1m.md):
# VALU busy (SQ_ACTIVE_INST_VALU) and L1 line accesses per cycle (TCP_TOTAL_CACHE_ACCESSES).
LIMITERS = {
    "neighbors_density": "valu (106 % busy at 4 cycles per wave64 instruction; TA 71 %, L1 0.72 lines/cycle)",
    "pressure_update": "l1 line accesses (0.98 per CU cycle; TA 69 %, VALU 74 %)",
    "advect": "l1 line accesses and valu (0.80 per CU cycle; TA 57 %, VALU 88 %)",
    "source": "l1 line accesses (0.96 per CU cycle; TA 65 %, VALU 81 %)",
    "pressure_acc": "l1 line accesses and valu (0.71 per CU cycle; VALU 93 %)",
    "final": "l1 line accesses and valu (0.65 per CU cycle; VALU 83 %)",
}
PROBE_STEPS = 10  # eager steps with every kernel class timed, before the timed steps
PMC_SUMMARY = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "pmc_traffic.json")


def measured_traffic(workload: str, kernel: str):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 PMC summary
    (FETCH_SIZE + WRITE_SIZE, scaled by the copy-kernel calibration), or None."""
    try:
        with open(PMC_SUMMARY) as f:
            d = json.load(f)
        return d[workload]["kernels"][kernel]["bytes_per_launch"]
    except (OSError, KeyError, ValueError):
        return None


def sort_passes(params) -> int:
    """Radix passes of the cell sort (8-, 9- or 10-bit digits, whichever needs fewest): keys
    count the cells of every world's box (cell width w = H (1 + skin) (1 + 1/128) along x and
    y, w / 2 along z, at least 3 per axis) times 4 x-slabs per cell."""
    total = 0
    for w in range(params.n_worlds):
        cw = params.smoothing_length[w] * (1.0 + params.skin) * (1.0 + 1.0 / 128.0)
        widths = np.array([cw, cw, 0.5 * cw])
        d = np.maximum(np.ceil((params.domain_hi[w] - params.domain_lo[w]) / widths) + 1.0, 3.0)
        total += int(np.prod(d))
    bits = max(int(np.ceil(np.log2(max(total * 4, 2)))), 1)
    return min((bits + 7) // 8, (bits + 8) // 9, (bits + 9) // 10)


def algorithmic_bytes(nbar: float, passes: int = 3, fused_source: bool = True, iters: float = 2.0,
                      solve_passes: int = 0) -> dict[str, float]:
    """Compulsory bytes per fluid particle and launch, per kernel class (nbar = mean
    fluid-fluid list length; 4 B per list entry + 4 B count).  `passes` radix passes of
    20 B each; the source pass also writes iteration 0's pressure acceleration (12 B) unless
    the MLP warm start runs between them.  The whole-world solve runs, in one launch, the
    source pass (after advection, which runs as its own launch with the MLP warm start),
    `iters` pressure updates, iters - 1 pressure accelerations and the forces; after
    `solve_passes` per-pass iterations it runs only the remaining iterations (an upper bound:
    worlds that stopped earlier skip it)."""
    L = 4.0 * nbar + 4.0
    t = {
        "build": 28 + passes * 20 + 92 + 16,
        "neighbors_density": 16 + 4 + 4 + 16 + L,
        # xs, vs, pws -> avisc, vadv (v_adv, p0 / rho^2)
        "advect": 16 + 16 + 4 + 32 + L,
        # xs, vadv, gb, rho, pws -> src, p, xpr (x, p / rho^2) [, ap]
        "source": 16 + 16 + 16 + 4 + 4 + 4 + 4 + 16 + (12 if fused_source else 0) + L,
        # xs, xpr, gb -> ap (three floats)
        "pressure_acc": 16 + 16 + 16 + 12 + L,
        # xs, ap, gb, src, aii, p, rho -> p, xpr
        "pressure_update": 16 + 12 + 16 + 4 + 4 + 4 + 4 + 4 + 16 + L,
        "mlp_warm_start": 4 + 4 + 4 + 8 + 16 + 4 + 4 + 16 + 16 + 4,
        "reduce": 0.0,
        # xs, xpr, gb, avisc, p -> fp (f, p)
        "final": 16 + 16 + 16 + 16 + 4 + 16 + L,
        # inv, fid, fworld, fp, q_in, qd_in, inv_mass -> q, qd, f, pprev (particle order); the
        # same launch copies the non-fluid particles through, whose bytes are not counted
        "integrate": 4 + 4 + 4 + 16 + 12 + 12 + 4 + 40,
    }
    if solve_passes > 0:
        rest = max(iters - solve_passes, 0.0)
        t["solve_worlds"] = rest * (t["pressure_update"] + t["pressure_acc"])
    else:
        t["solve_worlds"] = ((t["advect"] if fused_source else 0.0) + t["source"] + iters * t["pressure_update"]
                             + max(iters - 1.0, 0.0) * t["pressure_acc"] + t["final"])
    return t


def host_cpu() -> tuple[int, str]:
    """Cores this process may run on and the CPU model name."""
    try:
        cores = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        cores = os.cpu_count() or 1
    model = platform.processor() or platform.machine()
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return cores, model


def cpu_baseline(model, solver, q0, qd0, dt, budget_s: float = 20.0) -> dict:
    from oracle import oracle as O

    pb = O.problem_from_model(model, solver)
    # the restatement keeps no candidates between steps; it searches every step with radius H
    # (skin 0), the cheaper of its two searches
    pb.skin, pb.max_skin_neighbors = 0.0, 0
    nt, cpu_model = host_cpu()
    q, qd, pp = pb.q0, pb.qd0, None
    steps = 0
    t0 = time.perf_counter()
    while True:
        r = O.step(pb, q, qd, pprev=pp, dt=dt, precision="f32", nthreads=nt)
        q, qd, pp = r["q"], r["qd"], r["pprev"]
        steps += 1
        if time.perf_counter() - t0 > budget_s / 2 or steps >= 50:
            break
    el = time.perf_counter() - t0
    nf = solver.fluid_count
    # one thread: the first step again
    t1 = time.perf_counter()
    O.step(pb, pb.q0, pb.qd0, dt=dt, precision="f32", nthreads=1)
    el1 = time.perf_counter() - t1
    return {
        "value": nf * steps / el,
        "unit": "particle-updates/s",
        "cores": nt,
        "kind": "port",
        "cpu_model": cpu_model,
        "sample": f"{steps} steps of the same workload from the initial state, same-order FP32 C "
                  f"restatement searching neighbours every step (OpenMP, {nt} threads on {nt} host cores, "
                  f"{cpu_model})",
        "single_thread": {"value": nf / el1, "unit": "particle-updates/s", "cores": 1,
                          "sample": "the first step of the same workload, one thread"},
    }


CONTACT_WORKLOADS = ("xpbd_granular_1m", "semi_implicit_1m")


def bench_contacts(args) -> None:
    """The reference's own particle-contact paths on the HIP grid: a 100^3 block of grains
    (radius 1 cm) falling onto a floor of static grains, 1,044,100 particles.  xpbd_granular_1m
    steps SolverXPBD (4 iterations: grid build, particle contacts, deltas); semi_implicit_1m
    builds model.particle_grid each step (as a user of the reference does) and steps
    SolverSemiImplicit (penalty contacts, integration).  One step = one solver.step."""
    import torch

    from newton_sph_dev_amd import ModelBuilder, SolverSemiImplicit, SolverXPBD

    n, r = 100, 0.01
    b = ModelBuilder()
    b.add_particle_grid(pos=(0.1, 0.1, 0.05), rot=(0.0, 0.0, 0.0, 1.0), vel=(0.0, 0.0, 0.0), dim_x=n, dim_y=n,
                        dim_z=n, cell_x=2.1 * r, cell_y=2.1 * r, cell_z=2.1 * r, mass=0.01, jitter=0.2 * r,
                        radius_mean=r)
    g = np.stack(np.meshgrid(np.arange(2 * n + 10), np.arange(2 * n + 10), indexing="ij"), -1).reshape(-1, 2)
    floor = np.zeros((g.shape[0], 3))
    floor[:, :2] = g * (2.0 * r)
    b.add_particles(floor, np.zeros_like(floor), np.zeros(floor.shape[0]), np.full(floor.shape[0], r))
    m = b.finalize("cuda:0")
    m.particle_mu = 0.5
    xpbd = args.workload == "xpbd_granular_1m"
    solver = SolverXPBD(m, iterations=4) if xpbd else SolverSemiImplicit(m)
    s0, s1 = m.state(), m.state()
    dt = args.dt if args.dt > 0 else 1.0 / 60.0 / 8

    def step():
        nonlocal s0, s1
        s0.clear_forces()
        if not xpbd:
            m.particle_grid.build(s0.particle_q, radius=m.particle_max_radius * 2.0 + m.particle_cohesion)
        solver.step(s0, s1, None, None, dt)
        s0, s1 = s1, s0

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    lat = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ta = time.perf_counter()
        step()
        torch.cuda.synchronize()
        lat.append((time.perf_counter() - ta) * 1e3)
    el = time.perf_counter() - t0
    npart = m.particle_count
    print(json.dumps({
        "metric": "particle-updates/s (" + ("XPBD" if xpbd else "semi-implicit") + " particle step, 1M grains)",
        "value": npart * args.steps / el, "unit": "particle-updates/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": el / args.steps * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic (seeded grain lattice)",
        "config": {"workload": args.workload, "particles": npart, "dt": dt, "iterations": 4 if xpbd else None},
        "latency_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99))},
        "roofline": None, "cpu_baseline": None}))


def _timed_steps(solver, s0, s1, dt, steps: int):
    """K synchronised steps: (ms per step, per-step latencies in ms, iterations, states)."""
    import torch

    lat, its = [], []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        ta = time.perf_counter()
        solver.step(s0, s1, None, None, dt)
        if not solver.report:
            torch.cuda.synchronize()
        lat.append((time.perf_counter() - ta) * 1e3)
        its.append(solver.last_report["max_iterations"] if solver.last_report else None)
        s0, s1 = s1, s0
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) * 1e3 / steps, lat, its, (s0, s1)


def _pct(lat):
    return {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99))}


def extra_configs(log) -> dict:
    """The other benchmark configurations, each measured here with its own config."""
    import torch

    from newton_sph_dev_amd import SolverSPH, scenes
    from newton_sph_dev_amd.learned import WarmStartMLP

    data = os.path.join(os.path.dirname(os.path.abspath(__file__)), "newton_sph_dev_amd", "data")
    out = {}

    def free():
        torch.cuda.synchronize()
        torch.cuda.empty_cache()

    # C1: 32,768-particle dam break, 200 steps from rest
    log("C1 dam_break_32k")
    cfg = scenes.DAM_BREAK_32K
    m = scenes.build_dam_break(cfg).finalize("cuda:0")
    sol = SolverSPH(m, cfg.support, warm_start="prev")
    ms, lat, its, _ = _timed_steps(sol, m.state(), m.state(), scenes.DT, 200)
    out["C1_dam_break_32k"] = {
        "config": {"workload": cfg.name, "fluid_particles": sol.fluid_count, "steps": 200, "dt": scenes.DT,
                   "warm_start": "prev", "from": "rest"},
        "value": sol.fluid_count * 1e3 / ms, "unit": "particle-updates/s", "ms_per_step": ms,
        "latency_ms": _pct(lat), "mean_iterations": float(np.mean(its))}
    del sol, m
    free()

    # C4: learned warm start on the 1M dam break against zero and previous-pressure warm starts
    cfg = scenes.DAM_BREAK_1M
    m = scenes.build_dam_break(cfg).finalize("cuda:0")
    c4 = {"config": {"workload": cfg.name, "fluid_particles": None, "tolerance": 1e-3, "warmup_steps": 10,
                     "timed_steps": 20, "from": "rest"}, "runs": {}}
    for dt, net in ((1e-4, "warmstart_mlp.npy"), (1e-3, "warmstart_mlp_dt1e-3.npy")):
        for mode in ("zero", "prev", "mlp"):
            log(f"C4 dt={dt} {mode}")
            sol = SolverSPH(m, cfg.support, warm_start="zero" if mode == "mlp" else mode)
            if mode == "mlp":
                sol.set_mlp(WarmStartMLP.load(os.path.join(data, net)))
            a, b = m.state(), m.state()
            _, _, _, (a, b) = _timed_steps(sol, a, b, dt, 10)
            ms, lat, its, _ = _timed_steps(sol, a, b, dt, 20)
            c4["runs"][f"dt={dt:g}/{mode}"] = {"ms_per_step": ms, "mean_iterations": float(np.mean(its)),
                                               "network": net if mode == "mlp" else None}
            c4["config"]["fluid_particles"] = sol.fluid_count
            del sol, a, b
            free()
        r = c4["runs"]
        k = f"dt={dt:g}"
        c4[f"speedup_mlp_vs_zero_{k}"] = r[f"{k}/zero"]["ms_per_step"] / r[f"{k}/mlp"]["ms_per_step"]
        c4[f"speedup_mlp_vs_prev_{k}"] = r[f"{k}/prev"]["ms_per_step"] / r[f"{k}/mlp"]["ms_per_step"]
    out["C4_learned_warm_start_1m"] = c4
    del m
    free()

    # C3: 4096 batched worlds (log-uniform 256..8192 fluid particles)
    log("C3 batched_4096 (building)")
    bcfg = scenes.BATCH_4096
    b, H = scenes.build_batched(bcfg)
    m = b.finalize("cuda:0")
    del b
    # the whole-world schedule (each workgroup iterates whole worlds to their own stop test in
    # one persistent launch, no host read inside the step), and the per-pass one beside it
    c3 = None
    for solve, block in (("worlds", 512), ("passes", 256)):
        sol = SolverSPH(m, H, warm_start="prev", solve=solve, solve_block=block)
        a, c = m.state(), m.state()
        _, _, _, (a, c) = _timed_steps(sol, a, c, scenes.DT, 3)
        ms, lat, its, _ = _timed_steps(sol, a, c, scenes.DT, 10)
        if c3 is None:
            c3 = {
                "config": {"workload": bcfg.name, "worlds": bcfg.n_worlds, "fluid_particles": sol.fluid_count,
                           "boundary_particles": sol.boundary_count, "steps": 10, "warmup": 3, "dt": scenes.DT,
                           "warm_start": "prev",
                           "launch": f"eager; whole-world solve: one persistent launch of {block}-thread "
                                     "workgroups taking whole worlds, largest first, each iterated to its own "
                                     "stop test on the device; one report read per step"},
                "value": sol.fluid_count * 1e3 / ms, "unit": "particle-updates/s", "ms_per_step": ms,
                "latency_ms": _pct(lat), "mean_iterations": float(np.mean(its))}
        else:
            c3["per_pass_schedule"] = {"ms_per_step": ms, "latency_ms": _pct(lat),
                                       "launch": "eager, one launch per pass over all worlds, one host read of "
                                                 "the still-iterating count per step"}
        del sol, a, c
        free()
    out["C3_batched_4096"] = c3
    del m
    free()

    # C5: 1024 batched worlds (uniform 4096..65536), learned warm start, HIP graph replays
    log("C5 batched_1024 (building)")
    bcfg = scenes.BATCH_1024
    b, H = scenes.build_batched(bcfg)
    m = b.finalize("cuda:0")
    del b
    mlp = WarmStartMLP.load(os.path.join(data, "warmstart_mlp.npy"))
    c5 = None
    # graph replays of the whole-world schedule after 2 per-pass iterations (no host read inside
    # the step), and of the per-pass schedule with captured iteration slots beside it
    for solve in ("worlds", "passes"):
        sol = SolverSPH(m, H, solve=solve, solve_block=512, solve_passes=2)
        sol.set_mlp(mlp)
        a, c = m.state(), m.state()
        _, _, its0, (a, c) = _timed_steps(sol, a, c, scenes.DT, 6)
        slots = max(its0[-3:]) + 1  # the first steps from rest need more; a replay finishes eagerly if short
        sol.use_graph, sol.graph_slots = True, slots
        _, _, _, (a, c) = _timed_steps(sol, a, c, scenes.DT, 2)  # captures both directions
        ms, lat, its, _ = _timed_steps(sol, a, c, scenes.DT, 20)
        if c5 is None:
            c5 = {
                "config": {"workload": bcfg.name, "worlds": bcfg.n_worlds, "fluid_particles": sol.fluid_count,
                           "boundary_particles": sol.boundary_count, "steps": 20, "dt": scenes.DT,
                           "warm_start": "mlp",
                           "launch": "HIP graph replay; 2 Jacobi iterations as launches over all worlds, then "
                                     "one persistent launch of 512-thread workgroups iterating each world still "
                                     "iterating to its own stop test; no host read inside the step, one report "
                                     "read per step"},
                "value": sol.fluid_count * 1e3 / ms, "unit": "particle-updates/s", "ms_per_step": ms,
                "latency_ms": _pct(lat), "mean_iterations": float(np.mean(its))}
        else:
            c5["per_pass_schedule"] = {"ms_per_step": ms, "latency_ms": _pct(lat),
                                       "launch": f"HIP graph replay, {slots} captured Jacobi slots; one host "
                                                 "read of the still-iterating count after the slots per step"}
        del sol, a, c
        free()
    out["C5_batched_1024_mlp_graph"] = c5
    del m
    free()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="dam_break_1m")
    ap.add_argument("--warm-start", default="prev")
    ap.add_argument("--check-every", type=int, default=1)
    ap.add_argument("--skin", type=float, default=None,
                    help="relative skin of the kept neighbour candidates (default: the solver's)")
    ap.add_argument("--dt", type=float, default=0.0, help="time step (0: the workload's 1e-4 s)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-extra", action="store_true", help="skip the other benchmark configurations")
    ap.add_argument("--graph", action="store_true", help="replay captured HIP graphs")
    ap.add_argument("--graph-slots", type=int, default=0,
                    help="Jacobi iterations captured in the graph (0: warm-up maximum + 1)")
    ap.add_argument("--solve", default="passes", choices=("passes", "worlds"),
                    help="pressure-solve schedule: launches per pass, or the persistent whole-world launch")
    ap.add_argument("--solve-block", type=int, default=512, help="workgroup size of the whole-world launch")
    ap.add_argument("--solve-passes", type=int, default=0,
                    help="Jacobi iterations run per pass over all worlds before the whole-world launch")
    ap.add_argument("--autotune", action="store_true", help="tune the launch configuration first")
    ap.add_argument("--mlp", default=os.path.join(os.path.dirname(os.path.abspath(__file__)),
                                                  "newton_sph_dev_amd", "data", "warmstart_mlp.npy"),
                    help="weights for --warm-start mlp")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="run one 256 MiB device copy before the timed steps (PMC byte-count calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on exactly one GPU; --gpus must be 1"}))
        sys.exit(2)

    import torch

    from newton_sph_dev_amd import SolverSPH, _native as N, scenes

    if args.workload in CONTACT_WORKLOADS:
        bench_contacts(args)
        return

    if args.workload.startswith("batched"):
        cfg = {"batched_4096": scenes.BATCH_4096, "batched_1024": scenes.BATCH_1024}[args.workload]
        print(f"[bench] building {cfg.name}", file=sys.stderr, flush=True)
        builder, support = scenes.build_batched(cfg)
        model = builder.finalize("cuda:0")
        cfg.support = float(np.mean(support))
    else:
        cfg = {"dam_break_1m": scenes.DAM_BREAK_1M, "dam_break_32k": scenes.DAM_BREAK_32K}[args.workload]
        print(f"[bench] building {cfg.name}", file=sys.stderr, flush=True)
        model = scenes.build_dam_break(cfg).finalize("cuda:0")
        support = cfg.support
    solver = SolverSPH(model, support, warm_start="zero" if args.warm_start == "mlp" else args.warm_start,
                       check_every=args.check_every, skin=args.skin, solve=args.solve,
                       solve_block=args.solve_block, solve_passes=args.solve_passes)
    if args.warm_start == "mlp":
        from newton_sph_dev_amd.learned import WarmStartMLP

        solver.set_mlp(WarmStartMLP.load(args.mlp))
    s0, s1 = model.state(), model.state()
    q0 = model.particle_q.clone()
    qd0 = model.particle_qd.clone()
    dt = args.dt if args.dt > 0 else scenes.DT
    print(f"[bench] fluid={solver.fluid_count} boundary={solver.boundary_count}; warmup {args.warmup}",
          file=sys.stderr, flush=True)
    warm_iters = []
    for _ in range(args.warmup):
        solver.step(s0, s1, None, None, dt)
        warm_iters.append(solver.last_report["max_iterations"])
        s0, s1 = s1, s0
    torch.cuda.synchronize()
    if args.graph:
        solver.use_graph = True
        solver.graph_slots = args.graph_slots or (max(warm_iters, default=2) + 1)
    tuned = None
    if args.autotune:
        from newton_sph_dev_amd import autotune

        tuned = autotune.tune(solver, s0, s1, dt, force=True,
                              log=lambda m: print(f"[bench] {m}", file=sys.stderr, flush=True))
        tuned = {k: v for k, v in tuned.items() if k != "candidates"}
        torch.cuda.synchronize()

    if args.calibrate_copy:
        src = torch.ones(64 * 1024 * 1024, dtype=torch.float32, device="cuda:0")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    lib = N.load()
    nk = lib.nsph_kernel_count()
    names = [lib.nsph_kernel_name(k).decode() for k in range(nk)]
    ms = (C.c_double * nk)()
    cnt = (C.c_int * nk)()
    # eager pass with every class timed, to find the dominant kernel class (graph replays
    # carry no per-kernel events); long enough that its iteration counts, which set the
    # pressure classes' launches per step, are those of the timed steps
    graph_mode = solver.use_graph
    solver.use_graph = False
    N.call("nsph_solver_set_timing", solver._handle, -1)
    probe_iters = []
    for _ in range(PROBE_STEPS):
        solver.step(s0, s1, None, None, dt)
        probe_iters.append(solver.last_report["sum_iterations"] / max(solver.W, 1))
        s0, s1 = s1, s0
    N.call("nsph_solver_timing", solver._handle, ms, cnt)
    solver.use_graph = graph_mode
    per_all = {names[k]: (ms[k], cnt[k]) for k in range(nk)}
    counts = solver.fields()["neighbor_count"]
    nf = solver.fluid_count
    nbar = float(counts.sum().item()) / max(nf, 1)
    ab = algorithmic_bytes(nbar, sort_passes(solver.params), fused_source=args.warm_start != "mlp",
                           iters=float(np.mean(probe_iters)),
                           solve_passes=solver.solve_passes if solver.solve == "worlds" else 0)
    dom = max((k for k in names if ab.get(k, 0.0) > 0), key=lambda k: per_all[k][0])
    # timed region: events only around the dominant class
    N.call("nsph_solver_set_timing", solver._handle, 1 << names.index(dom))
    iters = []
    lat = []
    searched = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        t1 = time.perf_counter()
        solver.step(s0, s1, None, None, dt)  # the step report synchronises: t is the step latency
        lat.append(time.perf_counter() - t1)
        iters.append(solver.last_report["max_iterations"])
        searched.append(solver.last_report["searched_worlds"])
        s0, s1 = s1, s0
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    N.call("nsph_solver_timing", solver._handle, ms, cnt)
    N.call("nsph_solver_set_timing", solver._handle, 0)
    per = {names[k]: (ms[k], cnt[k]) for k in range(nk)}
    roofline_src = "HIP events around every launch of the class in the timed steps"
    if graph_mode or per[dom][1] == 0 or per[dom][0] <= 0.0:
        # graph replays carry no per-kernel events (only the eager finish of a solve whose
        # captured slots were too few does): use the eager all-class pass
        per = per_all
        roofline_src = f"HIP events of a {PROBE_STEPS}-step eager pass (the timed steps replay HIP graphs)"
    avg_s = per[dom][0] / max(per[dom][1], 1) / 1e3
    achieved = ab[dom] * nf / avg_s / 1e9
    value = nf * args.steps / el
    # per-class table from the all-class pass (event brackets on every launch there)
    kern = {k: {"ms_avg": round(per_all[k][0] / max(per_all[k][1], 1), 4),
                "launches_per_step": per_all[k][1] / PROBE_STEPS,
                "GBps": round(ab[k] * nf / (per_all[k][0] / max(per_all[k][1], 1) / 1e3) / 1e9, 1)
                if ab.get(k, 0.0) > 0 and per_all[k][1] else None}
            for k in names}
    launches_per_step = {k: per_all[k][1] / PROBE_STEPS for k in names}
    step_bytes = sum(ab.get(k, 0.0) * nf * launches_per_step[k] for k in names)
    out = {
        "metric": "particle-updates/s (1M SPH) + % HBM peak",
        "value": value,
        "unit": "particle-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": el * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded dam-break lattice)",
        "config": {
            "workload": cfg.name,
            "fluid_particles": nf,
            "boundary_particles": solver.boundary_count,
            "support_radius": cfg.support,
            "dt": dt,
            "tolerance": solver.tolerance,
            "warm_start": args.warm_start,
            "worlds": solver.W,
            "mean_iterations": float(np.mean(iters)),
            "skin": solver.skin,
            "searched_steps": int(np.count_nonzero(searched)),
            "mean_fluid_neighbors": nbar,
        },
        "roofline": {
            "bound": "hbm",
            "kernel": dom,
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": measured_traffic(cfg.name, dom),
            "traffic_source": "profiles/pmc_traffic.json: rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE of this kernel "
                              "(copy-kernel calibrated), committed with the kernel it measures; not measured in "
                              "this run",
            "limiter": LIMITERS.get(dom),
            "algorithmic_bytes_per_launch": ab[dom] * nf,
            "source": roofline_src,
        },
        "step_hbm_frac": step_bytes / (el / args.steps) / 1e9 / HBM_PEAK_GBS,
        "latency_ms": {"p50": float(np.percentile(lat, 50) * 1e3), "p99": float(np.percentile(lat, 99) * 1e3)},
        "launch": tuned or {"gather_block": solver.gather_block, "graph": solver.use_graph,
                            "graph_slots": solver.graph_slots if solver.use_graph else None,
                            "check_every": solver.check_every, "solve": solver.solve,
                            "solve_block": solver.solve_block if solver.solve == "worlds" else None,
                            "solve_passes": solver.solve_passes if solver.solve == "worlds" else None},
        "kernels": kern,
    }
    if not args.no_cpu_baseline:
        print("[bench] cpu baseline", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(model, solver, q0, qd0, dt)
    if not args.no_extra and cfg.name == "dam_break_1m" and args.workload == "dam_break_1m":
        del solver, model, s0, s1
        out["other_configs"] = extra_configs(lambda msg: print(f"[bench] {msg}", file=sys.stderr, flush=True))
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
