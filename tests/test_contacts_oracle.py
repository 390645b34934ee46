"""Pins of the particle-contact oracle (oracle/contacts_oracle.py) on the CPU.

Each check is fixed by something other than the oracle's own formula: closed-form
two-particle corrections and forces, Newton's third law / momentum conservation of pairwise
corrections, free fall, and the inactive / static-particle special cases.  A dropped term,
wrong sign or swapped index in the restatement fails one of them.
"""

from __future__ import annotations

import numpy as np
import pytest

from oracle import contacts_oracle as K


def _pair(sep=0.9, r=0.5, v0=(0, 0, 0), v1=(0, 0, 0), w=(1.0, 1.0)):
    x = np.array([[0.0, 0.0, 0.0], [sep, 0.0, 0.0]])
    v = np.array([v0, v1], float)
    return x, v, np.array(w), np.array([r, r]), np.array([1, 1])


def test_two_particle_normal_correction_closed_form():
    x, v, w, r, fl = _pair()
    d = K.xpbd_particle_contacts(x, v, w, r, fl, mu=0.0, cohesion=0.0, dt=0.01, relaxation=0.9)
    # overlap 0.1 shared by two unit inverse masses: each moves 0.05 * relaxation apart
    assert np.allclose(d, [[-0.045, 0, 0], [0.045, 0, 0]], atol=1e-15)


def test_static_partner_takes_no_correction():
    x, v, w, r, fl = _pair(w=(1.0, 0.0))
    d = K.xpbd_particle_contacts(x, v, w, r, fl, mu=0.0, cohesion=0.0, dt=0.01, relaxation=1.0)
    # the dynamic particle takes the whole overlap, the static one nothing
    assert np.allclose(d, [[-0.1, 0, 0], [0, 0, 0]], atol=1e-15)


def test_friction_correction_closed_form():
    dt = 0.01
    # tangential slip of 10 m/s along z: |vt| dt = 0.1 > mu |err| = 0.05, so Coulomb caps it
    x, v, w, r, fl = _pair(v0=(0, 0, 10.0))
    d = K.xpbd_particle_contacts(x, v, w, r, fl, mu=0.5, cohesion=0.0, dt=dt, relaxation=1.0)
    # normal part -0.05 along x; friction lambda_f = max(0.5 * -0.1, -0.1) = -0.05 along +z
    assert np.allclose(d[0], [-0.05, 0, -0.025], atol=1e-15)
    assert np.allclose(d[1], [0.05, 0, 0.025], atol=1e-15)
    # slow slip: |vt| dt = 0.001 < mu |err|, friction removes exactly the slip
    x, v, w, r, fl = _pair(v0=(0, 0, 0.1))
    d = K.xpbd_particle_contacts(x, v, w, r, fl, mu=0.5, cohesion=0.0, dt=dt, relaxation=1.0)
    assert np.allclose(d[0], [-0.05, 0, -0.0005], atol=1e-15)


def test_equal_velocities_give_no_friction():
    x, v, w, r, fl = _pair(v0=(0, 0, 10.0), v1=(0, 0, 10.0))
    d = K.xpbd_particle_contacts(x, v, w, r, fl, mu=1.0, cohesion=0.0, dt=0.01, relaxation=1.0)
    assert d[0, 2] == 0.0 and d[1, 2] == 0.0


def test_separated_and_inactive_pairs():
    x, v, w, r, fl = _pair(sep=1.2)
    assert not np.any(K.xpbd_particle_contacts(x, v, w, r, fl, 0.5, 0.0, 0.01, 0.9))
    # cohesion reaches across the 0.2 gap and pulls the pair together
    d = K.xpbd_particle_contacts(x, v, w, r, fl, 0.0, 0.25, 0.01, 1.0)
    assert np.allclose(d, [[0.1, 0, 0], [-0.1, 0, 0]], atol=1e-15)
    x, v, w, r, fl = _pair()
    fl = np.array([1, 0])
    assert not np.any(K.xpbd_particle_contacts(x, v, w, r, fl, 0.5, 0.0, 0.01, 0.9))


def test_random_pile_corrections_conserve_momentum():
    rng = np.random.default_rng(3)
    n = 60
    x = rng.uniform(0, 1.0, (n, 3))
    v = rng.normal(0, 1.0, (n, 3))
    r = np.full(n, 0.09)
    w = np.full(n, 2.0)  # equal masses, so sum(m * delta) = 0 for pairwise corrections
    fl = np.ones(n, int)
    d = K.xpbd_particle_contacts(x, v, w, r, fl, mu=0.3, cohesion=0.0, dt=0.01, relaxation=0.7)
    assert np.count_nonzero(np.linalg.norm(d, axis=1)) > 10
    assert np.allclose(d.sum(axis=0), 0.0, atol=1e-13)
    f = K.eval_particle_contact(x, v, r, fl, 1e3, 1e2, 1e2, 0.3, 0.0, np.zeros((n, 3)))
    assert np.allclose(f.sum(axis=0), 0.0, atol=1e-9)


def test_penalty_force_closed_form():
    x, v, w, r, fl = _pair()
    f = K.eval_particle_contact(x, v, r, fl, ke=1e3, kd=1e2, kf=1e2, mu=0.5, cohesion=0.0,
                                f_in=np.ones((2, 3)))
    # static overlap 0.1: fn = -0.1 ke, force pushes the pair apart with |f| = 100
    assert np.allclose(f, [[-100.0, 0, 0], [100.0, 0, 0]], atol=1e-12)
    # approaching at 1 m/s adds damping kd * vn; sliding at 10 m/s adds Coulomb friction
    x, v, w, r, fl = _pair(v0=(1.0, 0, 10.0))
    f = K.eval_particle_contact(x, v, r, fl, 1e3, 1e2, 1e2, 0.5, 0.0, np.zeros((2, 3)))
    # n = -x, vrel = (1,0,10): vn = -1 -> fn = -100 - 100 = -200; vs = 10, ft = min(1000, 100)
    assert np.allclose(f[0], [-200.0, 0, -100.0], atol=1e-12)
    assert np.allclose(f[1], -f[0], atol=1e-12)
    # inactive particles keep the force passed in
    f = K.eval_particle_contact(x, v, r, np.array([1, 0]), 1e3, 1e2, 1e2, 0.5, 0.0, np.ones((2, 3)))
    assert np.array_equal(f[1], [1, 1, 1]) and np.array_equal(f[0], [0, 0, 0])


def test_integrate_free_fall_and_clamp():
    x = np.array([[0.0, 1.0, 0.0], [0.0, 1.0, 0.0], [5.0, 5.0, 5.0]])
    v = np.zeros((3, 3))
    g = np.array([[0.0, -9.81, 0.0]])
    x1, v1 = K.integrate_particles(x, v, None, [1.0, 0.0, 1.0], [1, 1, 0], None, g, 0.1, 1e5)
    assert np.allclose(v1[0], [0, -0.981, 0]) and np.allclose(x1[0], [0, 1.0 - 0.0981, 0])
    assert np.array_equal(v1[1], [0, 0, 0]) and np.array_equal(x1[1], x[1])  # zero inverse mass
    assert np.array_equal(x1[2], x[2])  # inactive
    x1, v1 = K.integrate_particles(x, v, None, [1.0, 1.0, 1.0], [1, 1, 1], None, g, 1.0, 2.0)
    assert np.allclose(np.linalg.norm(v1, axis=1), 2.0)


def test_apply_deltas_velocity_from_positions():
    x0 = np.zeros((2, 3))
    xp = np.array([[0.1, 0, 0], [0, 0, 0]])
    d = np.array([[0.05, 0, 0], [1.0, 0, 0]])
    xo, vo = K.apply_particle_deltas(x0, xp, [1, 0], d, 0.5, 1e5, np.full((2, 3), 7.0), np.full((2, 3), 8.0))
    assert np.allclose(xo[0], [0.15, 0, 0]) and np.allclose(vo[0], [0.3, 0, 0])
    assert np.array_equal(xo[1], [7, 7, 7]) and np.array_equal(vo[1], [8, 8, 8])


@pytest.mark.parametrize("iters", [1, 4])
def test_xpbd_step_separates_overlap(iters):
    x, v, w, r, fl = _pair()
    q, qd = K.xpbd_step(x, v, None, w, r, fl, None, np.zeros((1, 3)), 0.01, 1e5, 0.0, 0.0, iters)
    sep = q[1, 0] - q[0, 0]
    # each iteration removes a 0.9 fraction of the remaining overlap
    assert np.isclose(1.0 - sep, 0.1 * 0.1 ** iters)
    assert np.allclose(qd[0], (q[0] - x[0]) / 0.01)
