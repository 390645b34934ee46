"""CPU oracle (test infrastructure only) for the reference's particle kernels that consume the
hash grid.  FP64 numpy restatement with brute-force candidate sets: every kernel filters its
candidates by distance (err <= cohesion), and the grid query radius covers every such pair,
so the contributing pairs are the same as through any grid; only the FP summation order
differs from the GPU's cell order, hence tolerance (not bit) parity.

  integrate_particles      newton/_src/solvers/solver.py:10-47
  xpbd_particle_contacts   newton/_src/solvers/xpbd/kernels.py:225-289
  apply_particle_deltas    newton/_src/solvers/xpbd/kernels.py:776-806
  particle_force           newton/_src/solvers/semi_implicit/kernels_contact.py:11-32
  eval_particle_contact    newton/_src/solvers/semi_implicit/kernels_contact.py:35-84
  xpbd_step                particle path of newton/_src/solvers/xpbd/solver_xpbd.py:258-441, 504-505, 657-660
                           (shape contacts: positional projection only, solver_xpbd.py:386-418)
  semi_implicit_step       particle path of newton/_src/solvers/semi_implicit/solver_semi_implicit.py:120-177

Pinned in tests/test_contacts_oracle.py by closed forms (two-particle normal and friction
corrections, penalty force magnitude, free fall), Newton's third law / momentum conservation
of the pairwise corrections, and the inactive / zero-mass special cases.
warp's normalize() returns the zero vector for a zero-length input (warp 1.x vec.h); that
case is restated as such.
"""

from __future__ import annotations

import numpy as np

ACTIVE = 1


def integrate_particles(x, v, f, inv_mass, flags, world, gravity, dt, v_max):
    """solver.py:10-47.  Returns (x_new, v_new); inactive particles keep x and v (v_out is
    left untouched by the reference, which here means: the v passed in)."""
    x = np.asarray(x, np.float64)
    v = np.asarray(v, np.float64)
    f = np.zeros_like(x) if f is None else np.asarray(f, np.float64)
    w = np.asarray(inv_mass, np.float64)
    g = np.asarray(gravity, np.float64).reshape(-1, 3)
    wd = np.zeros(len(x), int) if world is None else np.maximum(np.asarray(world), 0)
    active = (np.asarray(flags) & ACTIVE) != 0
    step = (-w < 0.0).astype(np.float64)  # wp.step(-inv_mass): 1 for a dynamic particle
    v1 = v + (f * w[:, None] + g[wd] * step[:, None]) * dt
    mag = np.linalg.norm(v1, axis=1)
    scale = np.where(mag > v_max, v_max / np.where(mag > 0, mag, 1.0), 1.0)
    v1 = v1 * scale[:, None]
    x1 = x + v1 * dt
    return np.where(active[:, None], x1, x), np.where(active[:, None], v1, v)


def _pairs(x, flags, i):
    act = (np.asarray(flags) & ACTIVE) != 0
    js = np.nonzero(act)[0]
    return js[js != i]


def xpbd_particle_contacts(x, v, inv_mass, radius, flags, mu, cohesion, dt, relaxation):
    """kernels.py:225-289: per active particle i, the summed normal + friction correction
    over active j != i with err = |x_i - x_j| - r_i - r_j <= cohesion and w_i + w_j > 0,
    times w_i * relaxation.  Inactive particles get zero."""
    x = np.asarray(x, np.float64)
    v = np.asarray(v, np.float64)
    w = np.asarray(inv_mass, np.float64)
    r = np.asarray(radius, np.float64)
    out = np.zeros_like(x)
    for i in range(len(x)):
        if (flags[i] & ACTIVE) == 0:
            continue
        js = _pairs(x, flags, i)
        n = x[i] - x[js]
        d = np.linalg.norm(n, axis=1)
        err = d - r[i] - r[js]
        denom = w[i] + w[js]
        sel = (err <= cohesion) & (denom > 0.0)
        if not np.any(sel):
            continue
        n, d, err, denom, jj = n[sel], d[sel], err[sel], denom[sel], js[sel]
        n = n / d[:, None]
        vrel = v[i] - v[jj]
        delta_n = n * err[:, None]
        vn = np.sum(n * vrel, axis=1)
        vt = vrel - n * vn[:, None]
        vtl = np.linalg.norm(vt, axis=1)
        lam_f = np.maximum(mu * err, -vtl * dt)
        tdir = np.where(vtl[:, None] > 0.0, vt / np.where(vtl > 0, vtl, 1.0)[:, None], 0.0)
        delta_f = tdir * lam_f[:, None]
        out[i] = np.sum((delta_f - delta_n) / denom[:, None], axis=0) * w[i] * relaxation
    return out


def apply_particle_deltas(x_orig, x_pred, flags, deltas, dt, v_max, x_out, v_out):
    """kernels.py:776-806; inactive entries of x_out / v_out are left as passed in."""
    x_orig = np.asarray(x_orig, np.float64)
    xn = np.asarray(x_pred, np.float64) + np.asarray(deltas, np.float64)
    vn = (xn - x_orig) / dt
    mag = np.linalg.norm(vn, axis=1)
    vn = vn * np.where(mag > v_max, v_max / np.where(mag > 0, mag, 1.0), 1.0)[:, None]
    act = ((np.asarray(flags) & ACTIVE) != 0)[:, None]
    return np.where(act, xn, x_out), np.where(act, vn, v_out)


def particle_force(n, v, c, k_n, k_d, k_f, k_mu):
    """kernels_contact.py:11-32 for one contact (vectors of shape (3,))."""
    vn = float(np.dot(n, v))
    jn = c * k_n
    jd = min(vn, 0.0) * k_d
    fn = jn + jd
    vt = v - n * vn
    vs = float(np.linalg.norm(vt))
    if vs > 0.0:
        vt = vt / vs
    ft = min(vs * k_f, k_mu * abs(fn))
    return -n * fn - vt * ft


def eval_particle_contact(x, v, radius, flags, ke, kd, kf, mu, cohesion, f_in):
    """kernels_contact.py:35-84: f[i] of every active particle is overwritten by its summed
    penalty contact force; inactive entries keep f_in."""
    x = np.asarray(x, np.float64)
    v = np.asarray(v, np.float64)
    r = np.asarray(radius, np.float64)
    f = np.array(f_in, np.float64, copy=True)
    for i in range(len(x)):
        if (flags[i] & ACTIVE) == 0:
            continue
        acc = np.zeros(3)
        for j in _pairs(x, flags, i):
            n = x[i] - x[j]
            d = float(np.linalg.norm(n))
            err = d - r[i] - r[j]
            if err <= cohesion:
                acc += particle_force(n / d, v[i] - v[j], err, ke, kd, kf, mu)
        f[i] = acc
    return f


def xpbd_step(x, v, f, inv_mass, radius, flags, world, gravity, dt, v_max, mu, cohesion,
              iterations, relaxation=0.9, shape_contacts=None, shape_mu=None, soft_contact_mu=0.5,
              particle_adhesion=0.0):
    """Particle path of one SolverXPBD.step: integrate, then `iterations` rounds of
    particle-shape corrections (shape_contacts: the (particle, shape, body_pos, body_vel,
    normal) list of Model.collide, fixed during the step, projected against the current
    positions as solve_particle_shape_contacts does, xpbd/kernels.py:114-222, launched at
    solver_xpbd.py:386-418) plus particle-particle contacts, then apply_deltas against the
    pre-step positions.  f is only read (integrate_particles).  Returns (x, v)."""
    from . import shapes_oracle as S

    x0 = np.asarray(x, np.float64)
    xp, vp = integrate_particles(x, v, f, inv_mass, flags, world, gravity, dt, v_max)
    # the reference's ping-pong buffers: first target is state_in's (x0, v0)
    bufs = [(xp, vp), (x0.copy(), np.asarray(v, np.float64).copy())]
    cur = 0
    q, qd = xp, vp
    for _ in range(iterations):
        d = xpbd_particle_contacts(q, qd, inv_mass, radius, flags, mu, cohesion, dt, relaxation)
        if shape_contacts:
            d = d + S.xpbd_shape_deltas(q, qd, inv_mass, radius, flags, shape_contacts, shape_mu,
                                        soft_contact_mu, particle_adhesion, dt, relaxation)
        tgt = 1 - cur
        q, qd = apply_particle_deltas(x0, q, flags, d, dt, v_max, bufs[tgt][0], bufs[tgt][1])
        bufs[tgt] = (q, qd)
        cur = tgt
    return q, qd


def semi_implicit_step(x, v, f, inv_mass, radius, flags, world, gravity, dt, v_max, ke, kd, kf,
                       mu, cohesion):
    """Particle path of one SolverSemiImplicit.step with a built grid: contact forces
    overwrite f of active particles, then integrate.  Returns (x, v, f)."""
    f1 = eval_particle_contact(x, v, radius, flags, ke, kd, kf, mu, cohesion, f)
    x1, v1 = integrate_particles(x, v, f1, inv_mass, flags, world, gravity, dt, v_max)
    return x1, v1, f1
