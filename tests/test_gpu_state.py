"""State-level behaviour of the SPH step on the GPU.

Bars:
  - forces the caller puts in state_in.particle_f are integrated with the SPH force, as
    integrate_particles integrates particle_f (solver.py:24-46): bit-exact against the FP32
    restatement given the same forces;
  - the step's pressure travels in state.sph.pressure (SolverSPH.register_custom_attributes),
    so a run resumed from a recorded frame equals the uninterrupted run bit for bit with the
    previous-pressure and the learned warm starts, not only the zero one;
  - a captured step (HIP graph) records the warm start it was captured with: changing the
    warm start or the MLP drops the captured steps, and replaying a stale handle is an error.
"""

from __future__ import annotations

from pathlib import Path

import numpy as np
import pytest
import torch

from newton_sph_dev_amd import SolverSPH, State, ViewerFile, _native as N, scenes
from newton_sph_dev_amd.learned import WarmStartMLP
from oracle import oracle as O

pytestmark = pytest.mark.gpu
DEV = "cuda:0"
MLP_PATH = Path(__file__).resolve().parents[1] / "newton_sph_dev_amd" / "data" / "warmstart_mlp.npy"


def _print(*a):
    print(*a, flush=True)


def test_caller_force_is_integrated_like_the_restatement():
    cfg = scenes.small_dam_break(10, 8, 12)
    _print("dam break 10x8x12 with a random caller force on every particle, 3 steps vs FP32 restatement")
    m = scenes.build_dam_break(cfg).finalize(DEV)
    s = SolverSPH(m, cfg.support, warm_start="prev")
    a, b = m.state(), m.state()
    gen = torch.Generator(DEV).manual_seed(9)
    pb = O.problem_from_model(m, s)
    q, qd, pp, xb = pb.q0, pb.qd0, None, None
    for k in range(3):
        a.particle_f.normal_(0.0, 2e-3, generator=gen)  # fluid mass is 1.95e-3 kg: accelerations ~1 m/s^2
        f = a.particle_f.cpu().numpy()
        s.step(a, b, None, None, scenes.DT)
        torch.cuda.synchronize()
        r = O.step(pb, q, qd, pprev=pp, precision="f32", f_in=f, xb=xb)
        assert np.array_equal(r["q"], b.particle_q.cpu().numpy()), k
        assert np.array_equal(r["qd"], b.particle_qd.cpu().numpy()), k
        assert np.array_equal(r["p"], b.sph.pressure.cpu().numpy()), k
        q, qd, pp, xb = r["q"], r["qd"], r["pprev"], r["xb"]
        a, b = b, a
    # the force changed the motion (compared with the same steps without it)
    r0 = O.step(pb, pb.q0, pb.qd0, precision="f32")
    assert not np.array_equal(r0["qd"], O.step(pb, pb.q0, pb.qd0, precision="f32", f_in=f)["qd"])


def test_pressure_is_carried_in_the_state():
    cfg = scenes.small_dam_break(10, 8, 12)
    m = scenes.build_dam_break(cfg).finalize(DEV)
    s = SolverSPH(m, cfg.support, warm_start="prev")
    a, b = m.state(), m.state()
    assert a.sph.pressure.shape == (m.particle_count,)
    s.step(a, b, None, None, scenes.DT)
    torch.cuda.synchronize()
    assert torch.equal(b.sph.pressure, s.fields()["pressure"])
    assert float(b.sph.pressure.max()) > 0.0
    # the next step's warm start is 0.5 * the state's pressure, whatever the solver last saw
    c = m.state()
    s.step(b, c, None, None, scenes.DT)
    p0 = s.warm_pressure()
    assert torch.equal(p0, 0.5 * b.sph.pressure)
    b.sph.pressure.zero_()
    s.step(b, c, None, None, scenes.DT)
    assert float(s.warm_pressure().abs().max()) == 0.0
    # a State without the attribute falls back to the solver's own copy of the last pressure
    plain = State()
    plain.particle_q, plain.particle_qd, plain.particle_f = (b.particle_q.clone(), b.particle_qd.clone(),
                                                             b.particle_f.clone())
    out = State()
    out.particle_q, out.particle_qd, out.particle_f = (torch.empty_like(plain.particle_q),
                                                       torch.empty_like(plain.particle_qd), None)
    last = s.fields()["pressure"]
    s.step(plain, out, None, None, scenes.DT)
    assert torch.equal(s.warm_pressure(), 0.5 * last)


@pytest.mark.parametrize("warm", ["prev", "mlp"])
def test_resume_from_recording_is_bit_exact_with_warm_start(tmp_path, warm):
    cfg = scenes.small_dam_break(12, 10, 14)
    _print(f"record 10 steps ({warm} warm start), resume from frame 4 of the .bin recording")
    m = scenes.build_dam_break(cfg).finalize(DEV)

    def solver():
        s = SolverSPH(m, cfg.support, warm_start="prev" if warm == "prev" else "zero")
        if warm == "mlp":
            s.set_mlp(WarmStartMLP.load(str(MLP_PATH)))
        return s

    s = solver()
    rec = ViewerFile(str(tmp_path / "run.bin"), auto_save=False)
    rec.set_model(m)
    a, b = m.state(), m.state()
    for _ in range(10):
        s.step(a, b, None, None, scenes.DT * 3)
        rec.log_state(b)
        a, b = b, a
    torch.cuda.synchronize()
    final = a.particle_q.clone(), a.particle_qd.clone(), a.sph.pressure.clone()
    assert float(final[2].max()) > 0.0
    rec.save_recording()

    play = ViewerFile(str(tmp_path / "run.bin"), auto_save=False)
    play.load_recording()
    st = m.state()
    play.load_state(st, 4)  # the state after 5 steps, pressure included
    assert float(st.sph.pressure.max()) > 0.0
    s2 = solver()  # a fresh solver: nothing but the State carries the previous pressure
    a2, b2 = st, m.state()
    for _ in range(5):
        s2.step(a2, b2, None, None, scenes.DT * 3)
        a2, b2 = b2, a2
    torch.cuda.synchronize()
    assert torch.equal(a2.particle_q, final[0]) and torch.equal(a2.particle_qd, final[1])
    assert torch.equal(a2.sph.pressure, final[2])


def test_captured_steps_follow_warm_start_changes():
    cfg = scenes.small_dam_break(12, 10, 14)
    _print("capture with the MLP warm start, detach the MLP, step again: equals an eager zero-warm-start step")
    m = scenes.build_dam_break(cfg).finalize(DEV)
    s = SolverSPH(m, cfg.support, use_graph=True, graph_slots=3)
    s.set_mlp(WarmStartMLP.load(str(MLP_PATH)))
    a, b = m.state(), m.state()
    s.step(a, b, None, None, scenes.DT)  # captures (mlp) and replays
    h = next(iter(s._graphs.values()))
    s.set_mlp(None)  # warm start -> zero; the captured step is dropped
    assert not s._graphs
    with pytest.raises(N.NsphError):
        s.replay(h)
    s.step(a, b, None, None, scenes.DT)  # captured again with the zero warm start
    torch.cuda.synchronize()
    e = SolverSPH(m, cfg.support, warm_start="zero")
    c = m.state()
    e.step(a, c, None, None, scenes.DT)
    torch.cuda.synchronize()
    assert torch.equal(b.particle_q, c.particle_q) and torch.equal(b.particle_qd, c.particle_qd)
    # a handle captured directly and then invalidated by a warm-start change is refused
    h2 = s.capture(a, b, scenes.DT, 3)
    N.call("nsph_solver_set_warm_start", s._handle, N.WARM_PREV)
    with pytest.raises(N.NsphError):
        s.replay(h2)
