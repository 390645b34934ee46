"""CPU oracle (test infrastructure only) for particle-shape soft contacts against static
shapes.  FP64 numpy restatement of
  signed distances and gradients   newton/_src/geometry/kernels.py:132-600
                                   (plane, sphere, box, capsule and cylinder along local Z)
  create_soft_contacts             newton/_src/geometry/kernels.py:1000-1141
  solve_particle_shape_contacts    newton/_src/solvers/xpbd/kernels.py:114-222
  eval_particle_body_contact       newton/_src/solvers/semi_implicit/kernels_contact.py:144-262
with the contact list ordered by (particle, shape), as the GPU writes it.

Pinned in tests/test_shapes_oracle.py: closed-form distances (sphere, box faces/edges/
corners/interior, capsule caps and side, cylinder side/caps/rim, infinite and finite planes),
gradients against finite differences of the distance, contact points lying on the surface,
the world filter of the reference's own test (test_environment_group_collision.py:74-112),
and closed-form XPBD corrections and penalty forces for a particle on a plane.
"""

from __future__ import annotations

import numpy as np

PLANE, SPHERE, CAPSULE, CYLINDER, BOX = 1, 3, 4, 6, 7
COLLIDE_PARTICLES = 1 << 2
ACTIVE = 1


def qrot(q, v):
    """Rotate v by the unit quaternion q = (x, y, z, w)."""
    u = np.asarray(q[:3], np.float64)
    w = float(q[3])
    v = np.asarray(v, np.float64)
    t = 2.0 * np.cross(u, v)
    return v + w * t + np.cross(u, t)


def qrot_inv(q, v):
    return qrot((-q[0], -q[1], -q[2], q[3]), v)


def _unit(v, fallback):
    n = np.linalg.norm(v)
    return v / n if n > 1e-8 else np.asarray(fallback, np.float64)


def sdf(kind, p, scale):
    p = np.asarray(p, np.float64)
    x, y, z = p
    if kind == SPHERE:
        return np.linalg.norm(p) - scale[0]
    if kind == BOX:
        q = np.abs(p) - np.asarray(scale[:3])
        return np.linalg.norm(np.maximum(q, 0.0)) + min(max(q[0], q[1], q[2]), 0.0)
    if kind == CAPSULE:
        r, hh = scale[0], scale[1]
        if z > hh:
            return np.linalg.norm([x, y, z - hh]) - r
        if z < -hh:
            return np.linalg.norm([x, y, z + hh]) - r
        return np.hypot(x, y) - r
    if kind == CYLINDER:
        r, hh = scale[0], scale[1]
        dx, dy = np.hypot(x, y) - r, abs(z) - hh
        return min(max(dx, dy), 0.0) + np.hypot(max(dx, 0.0), max(dy, 0.0))
    if kind == PLANE:
        w, l = 0.5 * scale[0], 0.5 * scale[1]
        if w > 0.0 and l > 0.0:
            return max(abs(x) - w, abs(y) - l, abs(z))
        return z
    raise ValueError(kind)


def sdf_grad(kind, p, scale):
    p = np.asarray(p, np.float64)
    x, y, z = p
    if kind == SPHERE:
        return _unit(p, (0, 0, 1))
    if kind == BOX:
        h = np.asarray(scale[:3], np.float64)
        q = np.abs(p) - h
        if np.any(q > 0.0):
            v = p - np.clip(p, -h, h)
            n = np.linalg.norm(v)
            return v / n if n > 0 else np.zeros(3)
        s = np.sign(p)
        if (q[0] > q[1] and q[0] > q[2]) or (q[1] == 0.0 and q[2] == 0.0):
            return np.array([s[0], 0.0, 0.0])
        if (q[1] > q[0] and q[1] > q[2]) or (q[0] == 0.0 and q[2] == 0.0):
            return np.array([0.0, s[1], 0.0])
        return np.array([0.0, 0.0, s[2]])
    if kind == CAPSULE:
        hh = scale[1]
        if z > hh:
            return _unit(np.array([x, y, z - hh]), (0, 0, 1))
        if z < -hh:
            return _unit(np.array([x, y, z + hh]), (0, 0, -1))
        return _unit(np.array([x, y, 0.0]), (0, 0, 1))
    if kind == CYLINDER:
        r, hh = scale[0], scale[1]
        v = np.array([x, y, 0.0])
        lv = np.linalg.norm(v)
        radial = v / lv if lv > 1e-8 else np.array([0.0, 0.0, 1.0])
        axial = np.array([0.0, 0.0, np.sign(z)])
        dx, dy = lv - r, abs(z) - hh
        if dx > 0.0 and dy > 0.0:
            return _unit(radial * dx + axial * dy, radial)
        return radial if dx > dy else axial
    if kind == PLANE:
        return np.array([0.0, 0.0, 1.0])
    raise ValueError(kind)


def soft_contacts(x, radius, flags, world, shapes, margin):
    """Contacts as the list of (particle, shape, body_pos, body_vel, normal), ordered by
    (particle, shape).  shapes: list of dicts with xform (7), type, scale (3), world, flags."""
    out = []
    for i in range(len(x)):
        if (flags[i] & ACTIVE) == 0:
            continue
        pw = -1 if world is None else int(world[i])
        for s, sh in enumerate(shapes):
            if (sh["flags"] & COLLIDE_PARTICLES) == 0:
                continue
            if pw != -1 and sh["world"] != -1 and pw != sh["world"]:
                continue
            pos = np.asarray(sh["xform"][:3], np.float64)
            q = sh["xform"][3:]
            local = qrot_inv(q, np.asarray(x[i], np.float64) - pos)
            d = sdf(sh["type"], local, sh["scale"])
            if d < margin + radius[i]:
                n = sdf_grad(sh["type"], local, sh["scale"])
                out.append((i, s, qrot(q, local - n * d) + pos, np.zeros(3), qrot(q, n)))
    return out


def xpbd_shape_deltas(x, v, inv_mass, radius, flags, contacts, shape_mu, particle_mu, particle_ka, dt,
                      relaxation):
    """Per particle, the sum over its contacts of w1 * delta_total (xpbd/kernels.py:114-222, static shapes)."""
    d = np.zeros((len(x), 3))
    for (i, s, bx, bv, n) in contacts:
        if (flags[i] & ACTIVE) == 0:
            continue
        c = float(np.dot(n, np.asarray(x[i]) - bx)) - radius[i]
        if c > particle_ka:
            continue
        mu = 0.5 * (particle_mu + shape_mu[s])
        rv = np.asarray(v[i]) - bv
        vn = float(np.dot(n, rv))
        vt = rv - n * vn
        w1 = inv_mass[i]
        if w1 == 0.0:
            continue
        lam_f = max(mu * c, -np.linalg.norm(vt) * dt)
        vt_n = np.linalg.norm(vt)
        delta_f = (vt / vt_n if vt_n > 0 else np.zeros(3)) * lam_f
        d[i] += w1 * ((delta_f - n * c) / w1 * relaxation)
    return d


def shape_forces(x, v, radius, flags, contacts, s_ke, s_kd, s_kf, s_mu, ke, kd, kf, mu, ka, f_in):
    """particle_f after subtracting the penalty contact forces (kernels_contact.py:144-262)."""
    f = np.array(f_in, np.float64, copy=True)
    for (i, s, bx, bv, n) in contacts:
        if (flags[i] & ACTIVE) == 0:
            continue
        c = float(np.dot(n, np.asarray(x[i]) - bx)) - radius[i]
        if c > ka:
            continue
        k_e, k_d = 0.5 * (ke + s_ke[s]), 0.5 * (kd + s_kd[s])
        k_f, m_u = 0.5 * (kf + s_kf[s]), 0.5 * (mu + s_mu[s])
        rv = np.asarray(v[i]) - bv
        vn = float(np.dot(n, rv))
        vt = rv - n * vn
        vt_n = np.linalg.norm(vt)
        ft = (vt / vt_n if vt_n > 0 else np.zeros(3)) * min(k_f * vt_n, abs(m_u * c * k_e))
        f[i] -= n * c * k_e + (n * min(vn, 0.0) * k_d + ft)
    return f
