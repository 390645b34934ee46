"""Batched independent worlds and captured (HIP graph) steps on the GPU.

Bars: every world of a batch equals the same world run alone, bit for bit (keys are
world-segmented, per-world convergence, world-local FP64 reduction order); the batch equals
the same-order FP32 restatement bit for bit; graph replay equals the eager step bit for bit,
including when the captured iteration slots are too few and the replay finishes eagerly.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

from newton_sph_dev_amd import SolverSPH, scenes
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _print(*a):
    print(*a, flush=True)


def _run(model, H, steps, **kw):
    s = SolverSPH(model, H, **kw)
    a, b = model.state(), model.state()
    iters = []
    for _ in range(steps):
        s.step(a, b, None, None, scenes.DT)
        iters.append(s.world_stats()["iterations"].copy())
        a, b = b, a
    torch.cuda.synchronize()
    return s, a, iters


def test_batched_worlds_equal_single_world_runs():
    cfg = scenes.small_batch(12)
    _print("12 batched worlds (256..2048 particles) vs each world alone, 3 steps")
    b, H = scenes.build_batched(cfg)
    m = b.finalize("cuda:0")
    s, st, iters = _run(m, H, 3, warm_start="prev")
    q = st.particle_q.cpu().numpy()
    qd = st.particle_qd.cpu().numpy()
    world = m.particle_world.cpu().numpy()
    assert s.world_stats()["state"].tolist() == [1] * 12
    for w in range(cfg.n_worlds):
        mw = scenes.extract_world(m, w).finalize("cuda:0")
        sw, stw, itw = _run(mw, H[w], 3, warm_start="prev")
        assert np.array_equal(stw.particle_q.cpu().numpy(), q[world == w]), w
        assert np.array_equal(stw.particle_qd.cpu().numpy(), qd[world == w]), w
        assert [int(i[0]) for i in itw] == [int(i[w]) for i in iters], w


def test_batched_worlds_equal_fp32_restatement():
    cfg = scenes.small_batch(20, seed=11)
    _print("20 batched worlds: one step vs the FP32 restatement")
    b, H = scenes.build_batched(cfg)
    m = b.finalize("cuda:0")
    s = SolverSPH(m, H)
    a, c = m.state(), m.state()
    s.step(a, c, None, None, scenes.DT)
    torch.cuda.synchronize()
    pb = O.problem_from_model(m, s)
    r = O.step(pb, pb.q0, pb.qd0, precision="f32")
    assert np.array_equal(r["q"], c.particle_q.cpu().numpy())
    assert np.array_equal(r["f"], s.sph_force.cpu().numpy())
    assert np.array_equal(r["iters"], s.world_stats()["iterations"])


@pytest.mark.parametrize("slots", [None, 1])
def test_graph_replay_equals_eager(slots):
    """Two captured steps (a -> b, b -> a) replayed 12 times in ping-pong; after every replay
    the state equals the same eager step's bit for bit, with a synchronisation and eager
    torch work on the output between replays."""
    cfg = scenes.small_dam_break(14, 12, 16)
    _print(f"graph replay (iter_slots={slots}) vs eager, 12 steps ping-pong, checked per replay")
    m = scenes.build_dam_break(cfg).finalize("cuda:0")
    se = SolverSPH(m, cfg.support, warm_start="prev")
    x, y = m.state(), m.state()
    ref = []
    for _ in range(12):
        se.step(x, y, None, None, scenes.DT)
        x, y = y, x
        ref.append((x.particle_q.cpu().numpy(), x.particle_qd.cpu().numpy()))
    sg = SolverSPH(m, cfg.support, warm_start="prev")
    a, b = m.state(), m.state()
    stream = torch.cuda.Stream()
    with torch.cuda.stream(stream):
        g_ab = sg.capture(a, b, scenes.DT, iter_slots=slots)
        g_ba = sg.capture(b, a, scenes.DT, iter_slots=slots)
    for k in range(12):
        with torch.cuda.stream(stream):
            sg.replay(g_ab if k % 2 == 0 else g_ba)
        torch.cuda.synchronize()
        out = b if k % 2 == 0 else a
        assert np.array_equal(out.particle_q.cpu().numpy(), ref[k][0]), f"replay {k}"
        assert np.array_equal(out.particle_qd.cpu().numpy(), ref[k][1]), f"replay {k}"
        assert sg.last_report["unconverged_worlds"] == 0
        scratch = out.particle_q.clone().mul_(2.0).sum(dim=0)  # eager kernels between replays
        torch.cuda.synchronize()
        del scratch


def test_launch_configs_do_not_change_results():
    cfg = scenes.small_batch(10, seed=5)
    _print("gather block 64/128/256, every search variant and graph mode give identical steps")
    b, H = scenes.build_batched(cfg)
    m = b.finalize("cuda:0")
    outs = []
    configs = [(256, False, 6), (128, False, 6), (64, False, 8), (128, True, 4)]
    configs += [(256, False, v) for v in SolverSPH.SEARCH_VARIANTS[1:]]
    for gb, graph, search in configs:
        s = SolverSPH(m, H, warm_start="prev", gather_block=gb, use_graph=graph, graph_slots=2)
        s.set_launch(gb, search)
        a, c = m.state(), m.state()
        for _ in range(4):
            s.step(a, c, None, None, scenes.DT)
            a, c = c, a
        torch.cuda.synchronize()
        outs.append((a.particle_q.cpu().numpy(), a.particle_qd.cpu().numpy()))
    for q, qd in outs[1:]:
        assert np.array_equal(q, outs[0][0]) and np.array_equal(qd, outs[0][1])


def test_check_schedules_do_not_change_results():
    """The step's end is enqueued after min_iters iterations behind a device-side "every world
    converged" guard; at dt = 1e-3 with the zero warm start every step needs more iterations,
    so the guarded end is skipped and enqueued again.  Host check intervals 1 and 3 and no
    host check at all (0) must give identical steps and iteration counts."""
    cfg = scenes.small_batch(6, seed=9)
    _print("check_every 0/1/3 at dt = 1e-3: identical steps")
    b, H = scenes.build_batched(cfg)
    m = b.finalize("cuda:0")
    outs = []
    for ce in (0, 1, 3):
        s = SolverSPH(m, H, warm_start="zero", check_every=ce, min_iters=2)
        a, c = m.state(), m.state()
        iters = []
        for _ in range(3):
            s.step(a, c, None, None, 1e-3)
            iters.append(s.last_report["max_iterations"])
            a, c = c, a
        torch.cuda.synchronize()
        outs.append((a.particle_q.cpu().numpy(), a.particle_qd.cpu().numpy(), iters))
    assert min(outs[0][2]) > 2
    for q, qd, it in outs[1:]:
        assert it == outs[0][2]
        assert np.array_equal(q, outs[0][0]) and np.array_equal(qd, outs[0][1])


def test_autotune_picks_a_config(tmp_path):
    from newton_sph_dev_amd import autotune

    cfg = scenes.small_dam_break(16, 12, 16)
    _print("autotune on a small dam break")
    m = scenes.build_dam_break(cfg).finalize("cuda:0")
    s = SolverSPH(m, cfg.support, warm_start="prev")
    a, c = m.state(), m.state()
    best = autotune.tune(s, a, c, scenes.DT, steps=2, cache=tmp_path / "tune.json")
    assert best["gather_block"] in (64, 128, 256) and best["mode"] in ("eager", "graph")
    assert best["search"] in SolverSPH.SEARCH_VARIANTS and s.search_variant == best["search"]
    assert len({r["search"] for r in best["candidates"]}) == len(SolverSPH.SEARCH_VARIANTS)
    assert (tmp_path / "tune.json").exists()
    again = autotune.tune(s, a, c, scenes.DT, cache=tmp_path / "tune.json")
    assert again["gather_block"] == best["gather_block"]


def _tiny_batch():
    """Worlds of 0, 1, 3 and 5 fluid particles and two ordinary blocks (120 and 210), each in
    its own walled box: partial waves and quads, lanes past a world's end, and a world with
    no fluid at all."""
    from newton_sph_dev_amd import ModelBuilder
    from newton_sph_dev_amd.sim import ParticleFlags

    sp = 0.02
    rng = np.random.default_rng(5)
    b = ModelBuilder()
    box = (10 * sp, 8 * sp, 8 * sp)
    for nf in (0, 1, 120, 3, 5, 210):
        b.begin_world()
        if nf in (120, 210):
            dims = (6, 5, 4) if nf == 120 else (7, 6, 5)
            b.add_particle_grid(pos=(0.5 * sp, 0.5 * sp, 0.5 * sp), rot=(0.0, 0.0, 0.0, 1.0), vel=(0.0, 0.0, 0.0),
                                dim_x=dims[0], dim_y=dims[1], dim_z=dims[2], cell_x=sp, cell_y=sp, cell_z=sp,
                                mass=scenes.REST_DENSITY * sp ** 3, jitter=0.1 * sp, radius_mean=0.5 * sp)
        elif nf:
            pos = 0.5 * np.array(box) + rng.uniform(-1.5 * sp, 1.5 * sp, (nf, 3))
            b.add_particles(pos, np.zeros_like(pos), np.full(nf, scenes.REST_DENSITY * sp ** 3),
                            np.full(nf, 0.5 * sp), np.full(nf, int(ParticleFlags.ACTIVE)))
        bp = scenes.box_boundary_points(sp, box, 2)
        b.add_particles(bp, np.zeros_like(bp), np.zeros(bp.shape[0]), np.full(bp.shape[0], 0.5 * sp),
                        np.full(bp.shape[0], int(ParticleFlags.ACTIVE)))
        b.end_world()
    return b, np.full(6, 2 * sp)


def test_tiny_and_empty_worlds_in_a_batch():
    """A batch with worlds of 0, 1, 3 and 5 fluid particles: one step equals the FP32
    restatement bit for bit, and three steps of every world equal that world run alone."""
    _print("6 batched worlds with 0, 1, 120, 3, 5 and 210 fluid particles")
    b, H = _tiny_batch()
    m = b.finalize("cuda:0")
    s = SolverSPH(m, H)
    a, c = m.state(), m.state()
    s.step(a, c, None, None, scenes.DT)
    torch.cuda.synchronize()
    pb = O.problem_from_model(m, s)
    r = O.step(pb, pb.q0, pb.qd0, precision="f32")
    assert np.array_equal(r["q"], c.particle_q.cpu().numpy())
    assert np.array_equal(r["qd"], c.particle_qd.cpu().numpy())
    assert np.array_equal(r["iters"], s.world_stats()["iterations"])
    s3, st, iters = _run(m, H, 3)
    q = st.particle_q.cpu().numpy()
    world = m.particle_world.cpu().numpy()
    for w in range(len(H)):
        mw = scenes.extract_world(m, w).finalize("cuda:0")
        sw, stw, itw = _run(mw, H[w], 3)
        assert np.array_equal(stw.particle_q.cpu().numpy(), q[world == w]), w
        assert [int(i[0]) for i in itw] == [int(i[w]) for i in iters], w
