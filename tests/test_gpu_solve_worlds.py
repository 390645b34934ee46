"""The whole-world solve schedule (one persistent launch; each workgroup solves whole worlds,
iterating each to its own stop test on the device), alone or after k per-pass iterations over
all worlds, on the GPU.

Bars: for every workgroup size and k it gives the same bits as the per-pass schedule
(positions, velocities, SPH force, pressure, per-world iteration counts, final errors and stop
states) and as the same-order FP32 restatement; with the learned warm start; with a moving
boundary; with worlds stopping at max_iters; with worlds of 0, 1, 3 and 5 fluid particles; and
captured in HIP graphs, whose replays (no host read) equal eager steps bit for bit.
"""

from __future__ import annotations

import ctypes
from pathlib import Path

import numpy as np
import pytest
import torch

from newton_sph_dev_amd import SolverSPH, scenes
from newton_sph_dev_amd import _native as N
from newton_sph_dev_amd.learned import WarmStartMLP
from oracle import oracle as O

from test_gpu_batched import _tiny_batch
from test_gpu_moving import _piston_model

pytestmark = pytest.mark.gpu
DEV = "cuda:0"
MLP_PATH = Path(__file__).resolve().parents[1] / "newton_sph_dev_amd" / "data" / "warmstart_mlp.npy"


def _print(*a):
    print(*a, flush=True)


def _np(t):
    return t.detach().cpu().numpy()


def _steps(m, H, n, dt=scenes.DT, mlp=None, solve="passes", block=256, passes=0, **kw):
    """n steps from the model's state; returns per-step snapshots of everything the step makes."""
    s = SolverSPH(m, H, solve=solve, solve_block=block, solve_passes=passes, **kw)
    if mlp is not None:
        s.set_mlp(mlp)
    a, b = m.state(), m.state()
    out = []
    for _ in range(n):
        s.step(a, b, None, None, dt)
        torch.cuda.synchronize()
        ws = s.world_stats()
        out.append({"q": _np(b.particle_q), "qd": _np(b.particle_qd), "f": _np(s.sph_force),
                    "p": _np(s.fields()["pressure"]), "iters": ws["iterations"].copy(), "err": ws["density_error"].copy(),
                    "state": ws["state"].copy(), "report": dict(s.last_report)})
        a, b = b, a
    return s, out


def _same(ref, got, label):
    assert len(ref) == len(got)
    for k, (r, g) in enumerate(zip(ref, got)):
        for key in ("q", "qd", "f", "p", "iters", "err", "state"):
            assert np.array_equal(r[key], g[key]), f"{label}: step {k}, {key} differs"
        for key in ("max_iterations", "sum_iterations", "unconverged_worlds", "max_density_error",
                    "max_neighbor_count"):
            assert r["report"][key] == g["report"][key], f"{label}: step {k}, report {key}"


@pytest.mark.parametrize("block,passes", [(256, 0), (512, 0), (1024, 0), (512, 1), (256, 2)])
def test_solve_worlds_equals_passes_and_restatement(block, passes):
    cfg = scenes.small_batch(20, seed=11)
    _print(f"20 batched worlds, 3 steps: whole-world solve ({block} threads, after {passes} per-pass "
           "iterations) vs per-pass schedule and FP32")
    b, H = scenes.build_batched(cfg)
    m = b.finalize(DEV)
    _, ref = _steps(m, H, 3, warm_start="prev")
    s, got = _steps(m, H, 3, warm_start="prev", solve="worlds", block=block, passes=passes)
    _same(ref, got, f"block {block}, passes {passes}")
    it = ctypes.c_int()
    N.call("nsph_solver_iterations_launched", s._handle, it)
    assert it.value == -1  # no per-pass iteration launches
    pb = O.problem_from_model(m, s)
    r = O.step(pb, pb.q0, pb.qd0, precision="f32")
    assert np.array_equal(r["q"], got[0]["q"]) and np.array_equal(r["qd"], got[0]["qd"])
    assert np.array_equal(r["f"], got[0]["f"])
    assert np.array_equal(r["iters"], got[0]["iters"])


def test_solve_worlds_tiny_and_empty_worlds():
    _print("worlds of 0, 1, 120, 3, 5 and 210 fluid particles: whole-world solve vs per-pass, 3 steps")
    b, H = _tiny_batch()
    m = b.finalize(DEV)
    _, ref = _steps(m, H, 3)
    for block, passes in ((256, 0), (1024, 0), (512, 2)):
        _, got = _steps(m, H, 3, solve="worlds", block=block, passes=passes)
        _same(ref, got, f"block {block}, passes {passes}")


def test_solve_worlds_stops_at_max_iters():
    """dt = 1e-3 from rest with the zero warm start and max_iters = 3: worlds stop unconverged
    (state 2) inside the persistent launch exactly where the per-pass schedule stops them."""
    cfg = scenes.small_batch(8, seed=9)
    _print("8 worlds at dt = 1e-3, max_iters 3: whole-world solve vs per-pass")
    b, H = scenes.build_batched(cfg)
    m = b.finalize(DEV)
    _, ref = _steps(m, H, 3, dt=1e-3, warm_start="zero", max_iters=3)
    assert any((r["state"] == 2).any() for r in ref), "the scene must hit max_iters"
    for passes in (0, 2, 3, 5):
        _, got = _steps(m, H, 3, dt=1e-3, warm_start="zero", max_iters=3, solve="worlds", block=512, passes=passes)
        _same(ref, got, f"max_iters 3, passes {passes}")


@pytest.mark.parametrize("block", [256, 512])
def test_solve_worlds_learned_warm_start(block):
    cfg = scenes.small_batch(12, seed=3)
    _print(f"12 worlds, shipped network as warm start, 3 steps at dt 1e-4 and 5e-4: whole-world solve ({block})")
    b, H = scenes.build_batched(cfg)
    m = b.finalize(DEV)
    mlp = WarmStartMLP.load(str(MLP_PATH))
    for dt in (scenes.DT, 5e-4):
        _, ref = _steps(m, H, 3, dt=dt, mlp=mlp)
        for passes in (0, 1):
            _, got = _steps(m, H, 3, dt=dt, mlp=mlp, solve="worlds", block=block, passes=passes)
            _same(ref, got, f"mlp, dt {dt}, passes {passes}")


def test_solve_worlds_moving_boundary():
    cfg = scenes.small_dam_break(10, 8, 12)
    m, n_wall = _piston_model(cfg)
    _print(f"moving wall ({n_wall} particles): whole-world solve vs per-pass, 6 steps")
    dom = ((-0.05, -0.05, -0.05), (0.35, 0.35, 0.35))
    _, ref = _steps(m, cfg.support, 6, warm_start="prev", domain=dom)
    for passes in (0, 2):
        _, got = _steps(m, cfg.support, 6, warm_start="prev", domain=dom, solve="worlds", passes=passes)
        _same(ref, got, f"moving boundary, passes {passes}")


def test_solve_worlds_graph_replay_equals_eager():
    """Two captured whole-world steps (a -> b, b -> a) replayed 10 times in ping-pong: every
    replay equals the same eager per-pass step bit for bit, with eager torch work and a
    synchronisation between replays; the learned warm start's graphs likewise."""
    cfg = scenes.small_batch(10, seed=5)
    b, H = scenes.build_batched(cfg)
    m = b.finalize(DEV)
    mlp = WarmStartMLP.load(str(MLP_PATH))
    for use_mlp, passes in ((False, 0), (True, 0), (False, 2), (True, 1)):
        _print(f"whole-world solve in HIP graphs (mlp={use_mlp}, passes={passes}): 10 replays vs eager steps")
        _, ref = _steps(m, H, 10, warm_start="prev", mlp=mlp if use_mlp else None)
        sg = SolverSPH(m, H, warm_start="prev", solve="worlds", solve_block=512, solve_passes=passes)
        if use_mlp:
            sg.set_mlp(mlp)
        a, c = m.state(), m.state()
        stream = torch.cuda.Stream()
        with torch.cuda.stream(stream):
            g_ac = sg.capture(a, c, scenes.DT)
            g_ca = sg.capture(c, a, scenes.DT)
        for k in range(10):
            with torch.cuda.stream(stream):
                sg.replay(g_ac if k % 2 == 0 else g_ca)
            torch.cuda.synchronize()
            out = c if k % 2 == 0 else a
            assert np.array_equal(_np(out.particle_q), ref[k]["q"]), f"replay {k}"
            assert np.array_equal(_np(out.particle_qd), ref[k]["qd"]), f"replay {k}"
            assert np.array_equal(_np(out.sph.pressure), ref[k]["p"]), f"replay {k}"
            assert np.array_equal(sg.world_stats()["iterations"], ref[k]["iters"]), f"replay {k}"
            scratch = out.particle_q.clone().mul_(2.0).sum(dim=0)  # eager kernels between replays
            torch.cuda.synchronize()
            del scratch
        # a graph captured under one schedule refuses to run under another
        for mode, k in ((0, 0), (1, passes + 1)):
            N.call("nsph_solver_set_solve", sg._handle, mode, 256, k)
            with pytest.raises(N.NsphError):
                N.call("nsph_solver_graph_step", sg._handle, int(g_ac), None)
        N.call("nsph_solver_set_solve", sg._handle, 1, 512, passes)


def test_set_solve_rejects_bad_arguments():
    cfg = scenes.small_dam_break(6, 6, 6)
    m = scenes.build_dam_break(cfg).finalize(DEV)
    s = SolverSPH(m, cfg.support)
    for mode, block, k in ((2, 256, 0), (1, 128, 0), (0, 2048, 0), (1, 256, -1)):
        with pytest.raises(N.NsphError):
            N.call("nsph_solver_set_solve", s._handle, mode, block, k)
    with pytest.raises(ValueError):
        s.set_solve("persistent")


def test_autotune_searches_solve_schedules(tmp_path):
    """With enough worlds the tuner's third round times the whole-world schedules (three
    workgroup sizes, with and without per-pass iterations first, eager and captured) and
    applies the fastest configuration it found."""
    from newton_sph_dev_amd import autotune

    cfg = scenes.small_batch(64, seed=2)
    _print("autotune over 64 small worlds, solve schedules included")
    b, H = scenes.build_batched(cfg)
    m = b.finalize(DEV)
    s = SolverSPH(m, H, warm_start="prev")
    a, c = m.state(), m.state()
    best = autotune.tune(s, a, c, scenes.DT, steps=1, cache=tmp_path / "tune.json")
    sched = [r for r in best["candidates"] if r.get("solve") == "worlds"]
    assert len(sched) == 12
    assert {(r["solve_block"], r["solve_passes"]) for r in sched} == {(bk, k) for bk in (256, 512, 1024) for k in (0, 2)}
    assert s.solve == best.get("solve", "passes")
    if s.solve == "worlds":
        assert (s.solve_block, s.solve_passes) == (best["solve_block"], best["solve_passes"])
