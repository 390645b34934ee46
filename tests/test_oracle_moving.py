"""Pins the moving boundary of the CPU oracle without a GPU.

A zero-mass ACTIVE particle with a velocity in the model is a moving boundary particle: the
reference integrator moves it with its velocity (solver.py:24-46), and the SPH step hashes it
from state_in every step and counts its velocity in the source term,
psi sum_b (v_i - v_b) . gradW_ib.  Checked against what does not depend on that code:
  - Galilean invariance: fluid and walls translating together with one velocity u step like
    the same scene at rest (density identical, pressure and relative motion equal to
    rounding), while walls held at rest under the translating fluid give a clearly different
    pressure; a sign error in the boundary velocity term fails it;
  - the moving particles are taken where state_in puts them: densities equal those of the
    same walls made static at that place;
  - the moving particles themselves follow the reference integrator (x + v dt, v kept).
"""

from __future__ import annotations

import numpy as np

from newton_sph_dev_amd import scenes
from newton_sph_dev_amd.solvers import SPHParams
from oracle import oracle as O

U = np.array([0.35, -0.2, 0.15], np.float32)
DT = 1e-4


def _problem():
    cfg = scenes.small_dam_break(8, 7, 9)
    m = scenes.build_dam_break(cfg).finalize("cpu")
    prm = SPHParams.resolve(m, cfg.support, domain=((-0.1, -0.1, -0.1), (0.4, 0.4, 0.4)))
    return cfg, O.problem_from_model(m, prm)


def _moving(pb, walls: bool):
    """The problem with every zero-mass particle given velocity U in the model (walls=True)."""
    pm = O.Problem(**{k: getattr(pb, k) for k in pb.__dataclass_fields__})
    pm.qd0 = pb.qd0.copy()
    if walls:
        pm.qd0[pb.inv_mass == 0] = U
    return pm


def test_translating_scene_steps_like_the_scene_at_rest():
    cfg, pb = _problem()
    fluid = pb.inv_mass > 0
    wall = ~fluid
    rest = O.step(pb, pb.q0, pb.qd0, dt=DT, precision="f64", forced_iters=np.array([4]))
    pm = _moving(pb, walls=True)
    qd_in = pb.qd0.copy()
    qd_in[:] = U  # fluid and walls translate together
    mov = O.step(pm, pb.q0, qd_in, dt=DT, precision="f64", forced_iters=np.array([4]))
    assert np.array_equal(mov["rho"], rest["rho"])
    pmax = float(np.abs(rest["p"]).max())
    assert pmax > 10.0
    assert np.max(np.abs(mov["p"] - rest["p"])) <= 1e-5 * pmax
    assert np.max(np.abs((mov["qd"] - U)[fluid] - rest["qd"][fluid])) <= 1e-5
    assert np.max(np.abs((mov["q"] - U * np.float32(DT))[fluid] - rest["q"][fluid])) <= 1e-6
    # the walls moved with their velocity (reference integrator), and kept it
    assert np.array_equal(mov["q"][wall], (pb.q0[wall] + U * np.float32(DT)).astype(np.float32))
    assert np.array_equal(mov["qd"][wall], np.broadcast_to(U, mov["qd"][wall].shape))
    # walls at rest under the same translating fluid: the boundary term then sees the relative
    # motion, and the pressure differs by far more than the rounding above
    still = O.step(_moving(pb, walls=False), pb.q0, np.where(fluid[:, None], U, 0).astype(np.float32),
                   dt=DT, precision="f64", forced_iters=np.array([4]))
    assert np.max(np.abs(still["p"] - rest["p"])) >= 1e-2 * pmax


def test_moving_particles_are_taken_from_state_in():
    cfg, pb = _problem()
    wall = pb.inv_mass == 0
    left = wall & (pb.q0[:, 0] < 0.0)
    shift = np.float32(0.3 * cfg.spacing)
    # the left wall moving (velocity in the model) and placed 0.3 spacings inward in state_in
    pm = _moving(pb, walls=False)
    pm.qd0[left] = (1.0, 0.0, 0.0)
    q_in = pb.q0.copy()
    q_in[left, 0] += shift
    qd_in = pm.qd0.copy()
    r_mov = O.step(pm, q_in, qd_in, dt=DT, precision="f32", forced_iters=np.array([2]))
    # the same wall static at the shifted place
    ps = _moving(pb, walls=False)
    ps.q0 = q_in.copy()
    r_static = O.step(ps, q_in, pb.qd0, dt=DT, precision="f32", forced_iters=np.array([2]))
    assert np.array_equal(r_mov["rho"], r_static["rho"])
    # and the shift mattered: the fluid next to the wall is denser than with the wall in place
    r0 = O.step(pb, pb.q0, pb.qd0, dt=DT, precision="f32", forced_iters=np.array([2]))
    assert float(np.sum(r_mov["rho"])) > float(np.sum(r0["rho"])) + 1.0
