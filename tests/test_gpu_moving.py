"""Moving boundary particles on the GPU (zero-mass particles with a velocity: a piston wall).

Bars: every step of a dam break whose left wall moves into the fluid is bit-exact against the
same-order FP32 restatement (fluid x, v, p; the wall's own motion by the reference
integrator); graph replays of such steps equal the eager steps bit for bit.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

from newton_sph_dev_amd import SolverSPH, scenes
from oracle import oracle as O

pytestmark = pytest.mark.gpu
DEV = "cuda:0"
PISTON = (0.6, 0.0, 0.0)  # m/s, the left wall moving into the fluid


def _print(*a):
    print(*a, flush=True)


def _np(t):
    return t.detach().cpu().numpy()


def _piston_model(cfg):
    m = scenes.build_dam_break(cfg).finalize(DEV)
    wall = (m.particle_inv_mass == 0) & (m.particle_q[:, 0] < 0.0)
    m.particle_qd[wall] = torch.tensor(PISTON, device=DEV)
    return m, int(wall.sum())


def test_piston_steps_bitexact_vs_restatement():
    cfg = scenes.small_dam_break(10, 8, 12)
    m, n_wall = _piston_model(cfg)
    _print(f"dam break 10x8x12 with a moving left wall ({n_wall} particles at {PISTON} m/s): 8 steps vs FP32")
    s = SolverSPH(m, cfg.support, warm_start="prev", domain=((-0.05, -0.05, -0.05), (0.35, 0.35, 0.35)))
    assert s.boundary_count == int((m.particle_inv_mass == 0).sum())
    pb = O.problem_from_model(m, s)
    a, b = m.state(), m.state()
    q, qd, pp = pb.q0, pb.qd0, None
    for k in range(8):
        s.step(a, b, None, None, scenes.DT)
        torch.cuda.synchronize()
        r = O.step(pb, q, qd, pprev=pp, precision="f32")
        assert np.array_equal(r["q"], _np(b.particle_q)), k
        assert np.array_equal(r["qd"], _np(b.particle_qd)), k
        assert np.array_equal(r["p"], _np(b.sph.pressure)), k
        assert np.array_equal(r["iters"], s.world_stats()["iterations"]), k
        q, qd, pp = r["q"], r["qd"], r["pprev"]
        a, b = b, a
    # the wall moved 8 dt at its velocity, and the fluid next to it is compressed
    wall = (pb.inv_mass == 0) & (pb.q0[:, 0] < 0.0)
    moved = _np(a.particle_q)[wall, 0] - pb.q0[wall, 0]
    assert np.allclose(moved, 8 * scenes.DT * PISTON[0], rtol=1e-3)


def test_all_boundary_moving_bitexact_vs_restatement():
    """Every boundary particle moving (the whole box translating slowly, so no static boundary
    particle exists): the particles next to the walls are finished by the boundary pass all the
    same, bit-exact against the FP32 restatement for 4 steps."""
    cfg = scenes.small_dam_break(10, 8, 12)
    m = scenes.build_dam_break(cfg).finalize(DEV)
    walls = m.particle_inv_mass == 0
    m.particle_qd[walls] = torch.tensor((0.05, 0.0, 0.0), device=DEV)
    _print(f"dam break 10x8x12, all {int(walls.sum())} boundary particles moving: 4 steps vs FP32")
    s = SolverSPH(m, cfg.support, warm_start="prev", domain=((-0.05, -0.05, -0.05), (0.35, 0.35, 0.35)))
    pb = O.problem_from_model(m, s)
    a, b = m.state(), m.state()
    q, qd, pp = pb.q0, pb.qd0, None
    for k in range(4):
        s.step(a, b, None, None, scenes.DT)
        torch.cuda.synchronize()
        r = O.step(pb, q, qd, pprev=pp, precision="f32")
        assert np.array_equal(r["q"], _np(b.particle_q)), k
        assert np.array_equal(r["qd"], _np(b.particle_qd)), k
        assert np.array_equal(r["rho"], _np(s.fields()["density"])), k
        q, qd, pp = r["q"], r["qd"], r["pprev"]
        a, b = b, a


def test_piston_graph_replay_equals_eager():
    cfg = scenes.small_dam_break(10, 8, 12)
    m, _ = _piston_model(cfg)
    _print("moving wall: 6 graph replays (ping-pong) vs eager steps")
    dom = ((-0.05, -0.05, -0.05), (0.35, 0.35, 0.35))
    se = SolverSPH(m, cfg.support, warm_start="prev", domain=dom)
    x, y = m.state(), m.state()
    ref = []
    for _ in range(6):
        se.step(x, y, None, None, scenes.DT)
        x, y = y, x
        ref.append(_np(x.particle_q))
    sg = SolverSPH(m, cfg.support, warm_start="prev", domain=dom)
    a, b = m.state(), m.state()
    stream = torch.cuda.Stream()
    with torch.cuda.stream(stream):
        g_ab = sg.capture(a, b, scenes.DT)
        g_ba = sg.capture(b, a, scenes.DT)
    for k in range(6):
        with torch.cuda.stream(stream):
            sg.replay(g_ab if k % 2 == 0 else g_ba)
        torch.cuda.synchronize()
        out = b if k % 2 == 0 else a
        assert np.array_equal(_np(out.particle_q), ref[k]), f"replay {k}"
