"""Pins of the particle-shape oracle (oracle/shapes_oracle.py) on the CPU: closed-form
distances, gradients vs finite differences, contact points on the surface, the reference's
world-filter scenario, and closed-form corrections and forces for a particle on a plane."""

from __future__ import annotations

import numpy as np
import pytest

from oracle import shapes_oracle as K

ID = (0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 1.0)


def test_closed_form_distances():
    assert K.sdf(K.SPHERE, (3.0, 4.0, 0.0), (2.0, 0, 0)) == pytest.approx(3.0)
    assert K.sdf(K.SPHERE, (0.5, 0.0, 0.0), (2.0, 0, 0)) == pytest.approx(-1.5)
    h = (1.0, 2.0, 3.0)
    assert K.sdf(K.BOX, (2.5, 0.0, 0.0), h) == pytest.approx(1.5)  # face
    assert K.sdf(K.BOX, (2.0, 3.0, 0.0), h) == pytest.approx(np.sqrt(2.0))  # edge
    assert K.sdf(K.BOX, (2.0, 3.0, 5.0), h) == pytest.approx(np.sqrt(1 + 1 + 4))  # corner
    assert K.sdf(K.BOX, (0.5, 0.0, 0.0), h) == pytest.approx(-0.5)  # interior: nearest face x
    assert K.sdf(K.CAPSULE, (0.0, 0.0, 3.0), (0.5, 1.0, 0)) == pytest.approx(1.5)  # above the cap
    assert K.sdf(K.CAPSULE, (2.0, 0.0, 0.3), (0.5, 1.0, 0)) == pytest.approx(1.5)  # beside the side
    assert K.sdf(K.CYLINDER, (2.0, 0.0, 0.0), (0.5, 1.0, 0)) == pytest.approx(1.5)
    assert K.sdf(K.CYLINDER, (0.0, 0.0, 2.0), (0.5, 1.0, 0)) == pytest.approx(1.0)
    assert K.sdf(K.CYLINDER, (1.5, 0.0, 2.0), (0.5, 1.0, 0)) == pytest.approx(np.sqrt(2.0))  # rim
    assert K.sdf(K.CYLINDER, (0.0, 0.0, 0.0), (0.5, 1.0, 0)) == pytest.approx(-0.5)
    assert K.sdf(K.PLANE, (5.0, -3.0, 0.25), (0.0, 0.0, 0)) == 0.25  # infinite plane
    assert K.sdf(K.PLANE, (3.0, 0.0, 0.25), (4.0, 2.0, 0)) == pytest.approx(1.0)  # finite quad (Chebyshev)


@pytest.mark.parametrize("kind,scale", [(K.SPHERE, (0.7, 0, 0)), (K.BOX, (0.5, 0.8, 0.3)),
                                        (K.CAPSULE, (0.4, 0.6, 0)), (K.CYLINDER, (0.5, 0.4, 0))])
def test_gradients_match_finite_differences(kind, scale):
    rng = np.random.default_rng(kind)
    checked = 0
    for _ in range(400):
        p = rng.uniform(-1.5, 1.5, 3)
        e = 1e-6
        fd = np.array([(K.sdf(kind, p + e * u, scale) - K.sdf(kind, p - e * u, scale)) / (2 * e) for u in np.eye(3)])
        if abs(np.linalg.norm(fd) - 1.0) > 1e-4:
            continue  # a kink of the distance (medial axis) within e
        assert np.allclose(K.sdf_grad(kind, p, scale), fd, atol=1e-4), p
        checked += 1
    assert checked > 200


@pytest.mark.parametrize("kind,scale", [(K.SPHERE, (0.7, 0, 0)), (K.BOX, (0.5, 0.8, 0.3)),
                                        (K.CAPSULE, (0.4, 0.6, 0)), (K.CYLINDER, (0.5, 0.4, 0))])
def test_contact_points_lie_on_the_surface(kind, scale):
    rng = np.random.default_rng(10 + kind)
    q = np.array([0.2, -0.3, 0.5, 0.8])
    q /= np.linalg.norm(q)
    shape = {"xform": (0.3, -0.2, 0.1, *q), "type": kind, "scale": scale, "world": -1, "flags": K.COLLIDE_PARTICLES}
    x = rng.uniform(-1.5, 1.5, (300, 3))
    c = K.soft_contacts(x, np.full(300, 0.1), np.ones(300, int), None, [shape], margin=0.3)
    assert len(c) > 20
    for (i, s, bx, bv, n) in c:
        local = K.qrot_inv(q, bx - np.array(shape["xform"][:3]))
        assert abs(K.sdf(kind, local, scale)) < 1e-9
        assert abs(np.linalg.norm(n) - 1.0) < 1e-9


def test_world_filter_of_the_reference_scenario():
    # test_environment_group_collision.py:74-112: particle in world 0, sphere in world 1,
    # global box: one contact, with the box (shape 1)
    shapes = [{"xform": ID, "type": K.SPHERE, "scale": (0.2, 0, 0), "world": 1, "flags": 7},
              {"xform": (0.0, 0.2, 0.0, 0, 0, 0, 1), "type": K.BOX, "scale": (0.5, 0.1, 0.5), "world": -1,
               "flags": 7}]
    c = K.soft_contacts([[0.0, 0.0, 0.0]], [0.1], [1], [0], shapes, margin=0.01)
    assert [(i, s) for (i, s, *_) in c] == [(0, 1)]


def _ground(mu=0.5):
    return [{"xform": ID, "type": K.PLANE, "scale": (0.0, 0.0, 0.0), "world": -1, "flags": K.COLLIDE_PARTICLES}]


def test_xpbd_plane_correction_closed_form():
    # particle of radius 0.1 sunk 0.04 into the plane z = 0, sliding at 2 m/s
    x = np.array([[0.0, 0.0, 0.06]])
    v = np.array([[2.0, 0.0, -1.0]])
    c = K.soft_contacts(x, [0.1], [1], None, _ground(), margin=0.01)
    d = K.xpbd_shape_deltas(x, v, [1.0], [0.1], [1], c, shape_mu=[1.0], particle_mu=0.5, particle_ka=0.0,
                            dt=0.01, relaxation=0.9)
    # normal: -n c = +0.04 z; friction: mu = 0.75, lambda_f = max(0.75 * -0.04, -2 * 0.01) = -0.02 along +x
    assert np.allclose(d[0], [-0.02 * 0.9, 0.0, 0.04 * 0.9])


def test_penalty_force_closed_form():
    x = np.array([[0.0, 0.0, 0.06]])
    v = np.array([[0.0, 0.0, -1.0]])
    c = K.soft_contacts(x, [0.1], [1], None, _ground(), margin=0.01)
    f = K.shape_forces(x, v, [0.1], [1], c, [2.5e3], [100.0], [1e3], [1.0], 1e3, 10.0, 1e3, 0.5, 0.0,
                       np.zeros((1, 3)))
    # ke = 1750, kd = 55: f = -(n c ke + n min(vn, 0) kd) = +(0.04 * 1750 + 55) z
    assert np.allclose(f[0], [0.0, 0.0, 0.04 * 1750 + 55.0])
