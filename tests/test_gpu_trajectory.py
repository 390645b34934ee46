"""Trajectories on the GPU: resume from a recording, and drift against the FP64 oracle.

Bars:
  - a run resumed from a recorded frame (ViewerFile `.bin`, zero warm start so the state is
    the whole input) equals the uninterrupted run bit for bit;
  - after 50 steps of a small dam break positions stay within DRIFT_TOL of the FP64 oracle run
    with the same per-step iteration counts (state stored in FP32 between steps, as the API
    stores it), and after the 200 steps of the 32,768-particle dam break's benchmark
    configuration within DRIFT_TOL_32K (largest) and DRIFT_RMS_32K (rms); the drift measured on
    MI355X is printed.
"""

from __future__ import annotations

import numpy as np
import torch

import pytest

from newton_sph_dev_amd import SolverSPH, ViewerFile, scenes
from oracle import oracle as O

pytestmark = pytest.mark.gpu

DRIFT_STEPS = 50
DRIFT_TOL = 1e-5  # m, 8e-4 of the 0.0125 m spacing (small dam break, 50 steps)
# 32,768-particle dam break over its 200 benchmark steps: FP32 against FP64 rounding grows
# through the iterations' p >= 0 clamp at the free surface, so a few surface particles drift
# more than the bulk (measured on MI355X: max 1.18e-4 m, rms 3.3e-6 m)
DRIFT_TOL_32K = 2.5e-4  # m, 2 % of the 0.0125 m spacing, on the largest drift
DRIFT_RMS_32K = 1e-5  # m


def _print(*a):
    print(*a, flush=True)


def test_resume_from_recording_is_bit_exact(tmp_path):
    cfg = scenes.small_dam_break(12, 10, 14)
    _print("record 10 steps, resume from frame 4 of the .bin recording")
    m = scenes.build_dam_break(cfg).finalize("cuda:0")
    s = SolverSPH(m, cfg.support, warm_start="zero")
    rec = ViewerFile(str(tmp_path / "run.bin"), auto_save=False)
    rec.set_model(m)
    a, b = m.state(), m.state()
    for _ in range(10):
        s.step(a, b, None, None, scenes.DT)
        rec.log_state(b)
        a, b = b, a
    torch.cuda.synchronize()
    final = a.particle_q.clone(), a.particle_qd.clone()
    rec.save_recording()

    play = ViewerFile(str(tmp_path / "run.bin"), auto_save=False)
    play.load_recording()
    assert play.get_frame_count() == 10 and play.has_model()
    st = m.state()
    play.load_state(st, 4)  # the state after 5 steps
    s2 = SolverSPH(m, cfg.support, warm_start="zero")
    a2, b2 = st, m.state()
    for _ in range(5):
        s2.step(a2, b2, None, None, scenes.DT)
        a2, b2 = b2, a2
    torch.cuda.synchronize()
    assert torch.equal(a2.particle_q, final[0]) and torch.equal(a2.particle_qd, final[1])


@pytest.mark.parametrize("scene,steps", [("12x10x14", DRIFT_STEPS), ("32k", 200)])
def test_trajectory_drift_vs_fp64_oracle(scene, steps):
    """The small dam break over 50 steps, and the 32,768-particle dam break over the 200 steps
    of its benchmark configuration, against the FP64 oracle."""
    cfg = scenes.small_dam_break(12, 10, 14) if scene == "12x10x14" else scenes.DAM_BREAK_32K
    _print(f"dam break {scene}: {steps} steps, GPU vs FP64 oracle with the same iteration counts")
    m = scenes.build_dam_break(cfg).finalize("cuda:0")
    s = SolverSPH(m, cfg.support, warm_start="prev")
    a, b = m.state(), m.state()
    iters = []
    for _ in range(steps):
        s.step(a, b, None, None, scenes.DT)
        iters.append(s.world_stats()["iterations"].copy())
        a, b = b, a
    torch.cuda.synchronize()
    q_gpu = a.particle_q.cpu().numpy().astype(np.float64)
    pb = O.problem_from_model(m, s)
    q, qd, pp, xb = pb.q0, pb.qd0, None, None
    for k in range(steps):
        r = O.step(pb, q, qd, pprev=pp, precision="f64", forced_iters=iters[k], xb=xb)
        q, qd, pp, xb = r["q"], r["qd"], r["pprev"], r["xb"]
    d = np.linalg.norm(q_gpu - q.astype(np.float64), axis=1)
    rms = float(np.sqrt(np.mean(d * d)))
    _print(f"  drift after {steps} steps: max {d.max():.3e} m, rms {rms:.3e} m")
    if scene == "32k":
        assert d.max() <= DRIFT_TOL_32K and rms <= DRIFT_RMS_32K
    else:
        assert d.max() <= DRIFT_TOL
