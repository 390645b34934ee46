"""Kept neighbour candidates (skin) on the GPU.

Bars:
  - steps that keep their candidates and steps that search anew, mixed along a trajectory,
    are bit-exact against the same-order FP32 restatement fed the same search positions, and
    the State's search positions equal the restatement's;
  - the lists of a kept step are the O(N^2) brute-force sets of the current positions;
  - a solver whose own candidates are not those of state_in's search positions (another
    run's state) steps exactly like the run the state came from;
  - rebuild_lists() searches around the current positions: the restatement without search
    positions; skin 0 (the default) is the restatement with skin 0 and leaves the State's
    search positions alone;
  - the kept-candidate capacity overflow is reported as a CapacityError.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

from newton_sph_dev_amd import SolverSPH, scenes
from oracle import oracle as O

pytestmark = pytest.mark.gpu
DEV = "cuda:0"
DT = 4e-4  # a fast-falling column: the candidates are searched anew every few steps


def _print(*a):
    print(*a, flush=True)


def _np(t):
    return t.detach().cpu().numpy()


def test_kept_and_searched_steps_bitexact_vs_restatement():
    cfg = scenes.small_dam_break(10, 8, 12)
    _print(f"dam break 10x8x12, skin 0.1, 20 steps at dt {DT}: GPU vs FP32 restatement step by step")
    m = scenes.build_dam_break(cfg).finalize(DEV)
    s = SolverSPH(m, cfg.support, warm_start="prev", skin=0.1)
    pb = O.problem_from_model(m, s)
    assert pb.skin == 0.1
    fluid = pb.inv_mass > 0
    a, b = m.state(), m.state()
    q, qd, pp, xb = pb.q0, pb.qd0, None, None
    kept = searched = 0
    for k in range(20):
        s.step(a, b, None, None, DT)
        torch.cuda.synchronize()
        r = O.step(pb, q, qd, pprev=pp, dt=DT, precision="f32", xb=xb)
        assert np.array_equal(r["q"], _np(b.particle_q)), k
        assert np.array_equal(r["qd"], _np(b.particle_qd)), k
        assert np.array_equal(r["p"], _np(b.sph.pressure)), k
        assert np.array_equal(r["xb"][fluid], _np(b.sph.search_q)[fluid]), k
        assert np.array_equal(r["iters"], s.world_stats()["iterations"]), k
        if xb is not None and np.array_equal(r["xb"][fluid], xb[fluid]):
            kept += 1
            # a kept step's lists are the brute-force sets of the step's input positions
            counts, idx = s.neighbors()
            counts, idx = _np(counts), _np(idx)
            fid = np.nonzero(fluid)[0]
            bc, bi = O.neighbors_brute(q[fid], None, np.float32(cfg.support ** 2), cap=96)
            for j, i in enumerate(fid):
                assert sorted(idx[: counts[i], i].tolist()) == sorted(fid[bi[j, : bc[j]]].tolist()), (k, i)
        else:
            searched += 1
        q, qd, pp, xb = r["q"], r["qd"], r["pprev"], r["xb"]
        a, b = b, a
    _print(f"  {kept} steps kept their candidates, {searched} searched anew")
    assert kept >= 5 and searched >= 2


def test_foreign_search_positions_are_followed():
    """Solver B, warmed up on another trajectory, steps run A's state 6 -> 7 -> 8 exactly
    like A (its own kept candidates are not those of A's search positions)."""
    cfg = scenes.small_dam_break(10, 8, 12)
    _print("run A 8 steps; solver B (warmed on other states) steps A's state 6 twice")
    m = scenes.build_dam_break(cfg).finalize(DEV)
    sa = SolverSPH(m, cfg.support, warm_start="prev", skin=0.1)
    states = [m.state()]
    for _ in range(8):
        nxt = m.state()
        sa.step(states[-1], nxt, None, None, DT)
        states.append(nxt)
    torch.cuda.synchronize()
    sb = SolverSPH(m, cfg.support, warm_start="prev", skin=0.1)
    x, y = m.state(), m.state()
    x.particle_q.add_(0.003)
    for _ in range(3):
        sb.step(x, y, None, None, DT)
        x, y = y, x
    out1, out2 = m.state(), m.state()
    sb.step(states[6], out1, None, None, DT)
    sb.step(out1, out2, None, None, DT)
    torch.cuda.synchronize()
    for got, ref in ((out1, states[7]), (out2, states[8])):
        assert torch.equal(got.particle_q, ref.particle_q)
        assert torch.equal(got.particle_qd, ref.particle_qd)
        assert torch.equal(got.sph.search_q, ref.sph.search_q)


def test_rebuild_lists_and_skin_zero():
    cfg = scenes.small_dam_break(10, 8, 12)
    _print("rebuild_lists() == restatement without search positions; skin 0 == restatement skin 0")
    m = scenes.build_dam_break(cfg).finalize(DEV)
    s = SolverSPH(m, cfg.support, warm_start="zero", skin=0.1)
    a, b, c = m.state(), m.state(), m.state()
    for _ in range(2):
        s.step(a, b, None, None, DT)
        a, b = b, a
    s.rebuild_lists()
    s.step(a, c, None, None, DT)
    torch.cuda.synchronize()
    pb = O.problem_from_model(m, s)
    r = O.step(pb, _np(a.particle_q), _np(a.particle_qd), dt=DT, precision="f32")
    assert np.array_equal(r["q"], _np(c.particle_q))
    fluid = pb.inv_mass > 0
    assert np.array_equal(_np(c.sph.search_q)[fluid], _np(a.particle_q)[fluid])

    s0 = SolverSPH(m, cfg.support, warm_start="prev")
    pb0 = O.problem_from_model(m, s0)
    assert pb0.skin == 0.0 and s0.max_skin_neighbors == 0  # the default
    a, b = m.state(), m.state()
    q, qd, pp = pb0.q0, pb0.qd0, None
    for k in range(6):
        s0.step(a, b, None, None, DT)
        r = O.step(pb0, q, qd, pprev=pp, dt=DT, precision="f32")
        torch.cuda.synchronize()
        assert np.array_equal(r["q"], _np(b.particle_q)), k
        q, qd, pp = r["q"], r["qd"], r["pprev"]
        a, b = b, a
    # without a skin the State's search positions (NaN in a new State) are not written
    assert bool(torch.isnan(a.sph.search_q).all()) and bool(torch.isnan(b.sph.search_q).all())


def test_candidate_capacity_overflow_is_reported():
    from newton_sph_dev_amd import _native as N

    cfg = scenes.small_dam_break(10, 8, 12)
    _print("max_skin_neighbors = max_neighbors = 24: CapacityError")
    m = scenes.build_dam_break(cfg).finalize(DEV)
    s = SolverSPH(m, cfg.support, max_neighbors=24, skin=0.1, max_skin_neighbors=24)
    a, b = m.state(), m.state()
    with pytest.raises(N.CapacityError):
        s.step(a, b, None, None, DT)
    torch.cuda.synchronize()
    assert s.last_report["neighbor_overflow"] == 1
