"""Particle-shape soft contacts on the GPU against the FP64 oracle and reference scenarios.

Bars: the contact list (particle, shape) equals the oracle's, excluding pairs within 1e-5 m
of the margin threshold, where FP32 and FP64 may disagree; contact points and normals match
within 2e-6; XPBD corrections and penalty forces match within FP32 rounding; the reference's
world-filter test; a particle dropped on a ground plane comes to rest on it; SPH fluid flowing
past a static sphere stays out of it.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

from newton_sph_dev_amd import ModelBuilder, SolverSemiImplicit, SolverSPH, SolverXPBD, scenes
from oracle import shapes_oracle as K

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def _print(*a):
    print(*a, flush=True)


def _np(t):
    return t.detach().cpu().numpy().astype(np.float64)


def _shape_dicts(m):
    out = []
    for s in range(m.shape_count):
        out.append({"xform": tuple(_np(m.shape_transform)[s]), "type": int(m.shape_type[s]),
                    "scale": tuple(_np(m.shape_scale)[s]), "world": int(m.shape_world[s]),
                    "flags": int(m.shape_flags[s])})
    return out


def _scene(n=4000, seed=0):
    rng = np.random.default_rng(seed)
    b = ModelBuilder()
    b.begin_world()
    x = rng.uniform(-1.2, 1.2, (n // 2, 3))
    b.add_particles(x.tolist(), [(0.0, 0.0, 0.0)] * len(x), [1.0] * len(x), rng.uniform(0.01, 0.05, len(x)).tolist())
    b.add_shape_sphere(-1, ((0.3, -0.2, 0.1), (0.0, 0.0, 0.0, 1.0)), radius=0.4)
    b.end_world()
    b.begin_world()
    x = rng.uniform(-1.2, 1.2, (n // 2, 3))
    b.add_particles(x.tolist(), [(0.0, 0.0, 0.0)] * len(x), [1.0] * len(x), rng.uniform(0.01, 0.05, len(x)).tolist())
    q = np.array([0.2, -0.3, 0.5, 0.8])
    q /= np.linalg.norm(q)
    b.add_shape_box(-1, ((-0.2, 0.3, 0.0), tuple(q)), hx=0.3, hy=0.5, hz=0.2)
    b.end_world()
    b.add_shape_capsule(-1, ((0.5, 0.5, -0.4), (0.0, 0.0, 0.0, 1.0)), radius=0.2, half_height=0.3)
    b.add_shape_cylinder(-1, ((-0.6, -0.5, 0.3), tuple(q)), radius=0.25, half_height=0.2)
    b.add_shape_plane(plane=(0.0, 0.0, 1.0, 0.9), width=1.0, length=1.6)  # finite quad at z = -0.9
    b.add_ground_plane(height=-1.1)
    return b.finalize(DEV)


def test_soft_contacts_match_oracle():
    m = _scene()
    s = m.state()
    _print(f"soft contacts: {m.particle_count} particles x {m.shape_count} shapes")
    c = m.collide(s, soft_contact_margin=0.02)
    torch.cuda.synchronize()
    count = int(c.soft_contact_count[0])
    x, r = _np(s.particle_q), _np(m.particle_radius)
    ref = K.soft_contacts(x, r, m.particle_flags.cpu().numpy(), m.particle_world.cpu().numpy(), _shape_dicts(m), 0.02)
    assert count == len(ref) or abs(count - len(ref)) <= 3
    got = {(int(i), int(sh)): k for k, (i, sh) in
           enumerate(zip(c.soft_contact_particle[:count].tolist(), c.soft_contact_shape[:count].tolist()))}
    shapes = _shape_dicts(m)
    near = set()
    for i in range(len(x)):
        for sh_i, sh in enumerate(shapes):
            local = K.qrot_inv(sh["xform"][3:], x[i] - np.array(sh["xform"][:3]))
            if abs(K.sdf(sh["type"], local, sh["scale"]) - (0.02 + r[i])) < 1e-5:
                near.add((i, sh_i))
    want = {(i, s_) for (i, s_, *_) in ref}
    assert {p for p in got if p not in near} == {p for p in want if p not in near}
    assert len(want) > 300 and len({s_ for (_, s_) in want}) == m.shape_count
    bp, nn = _np(c.soft_contact_body_pos), _np(c.soft_contact_normal)
    for (i, s_, bx, bv, n) in ref:
        if (i, s_) in got:
            k = got[(i, s_)]
            assert np.allclose(bp[k], bx, atol=2e-6) and np.allclose(nn[k], n, atol=2e-6), (i, s_)
    # ordered by (particle, shape) and particle ranges consistent with the start array
    pairs = list(zip(c.soft_contact_particle[:count].tolist(), c.soft_contact_shape[:count].tolist()))
    assert pairs == sorted(pairs)
    start = c.soft_contact_start.cpu().numpy()
    assert start[-1] == count and np.all(np.diff(start) >= 0)


def test_particle_shape_collision_filtering():
    """newton/tests/test_environment_group_collision.py:74-112 with its assertions."""
    b = ModelBuilder()
    b.begin_world()
    b.add_particle(pos=(0, 0, 0), vel=(0, 0, 0), mass=1.0, radius=0.1)
    b.end_world()
    b.begin_world()
    b.add_shape_sphere(body=-1, xform=((0, 0, 0), (0, 0, 0, 1)), radius=0.2)
    b.end_world()
    b.add_shape_box(body=-1, xform=((0, 0.2, 0), (0, 0, 0, 1)), hx=0.5, hy=0.1, hz=0.5)
    model = b.finalize(DEV)
    state = model.state()
    contacts = model.contacts()
    model.collide(state, contacts)
    assert int(contacts.soft_contact_count.cpu()[0]) == 1
    assert int(contacts.soft_contact_shape.cpu()[0]) == 1


def test_xpbd_and_penalty_consumers_match_oracle():
    m = _scene(n=2000, seed=3)
    s = m.state()
    rng = np.random.default_rng(5)
    s.particle_qd.copy_(torch.from_numpy(rng.normal(0, 1.0, (m.particle_count, 3)).astype(np.float32)).to(DEV))
    c = m.collide(s, soft_contact_margin=0.0)
    torch.cuda.synchronize()
    from newton_sph_dev_amd.particle_solvers import shape_contact_deltas
    from newton_sph_dev_amd import _native as N

    d = torch.zeros_like(s.particle_q)
    shape_contact_deltas(m, c, s.particle_q, s.particle_qd, 1e-3, 0.8, d)
    f = torch.full_like(s.particle_q, 0.5)
    N.call("nsph_eval_shape_contacts", N.context(0), m.particle_count, N.ptr(s.particle_q), N.ptr(s.particle_qd),
           N.ptr(m.particle_radius), N.ptr(m.particle_flags), N.ptr(m.shape_material_ke), N.ptr(m.shape_material_kd),
           N.ptr(m.shape_material_kf), N.ptr(m.shape_material_mu), m.soft_contact_ke, m.soft_contact_kd,
           m.soft_contact_kf, m.soft_contact_mu, 0.0, N.ptr(c.soft_contact_start), c.soft_contact_max,
           N.ptr(c.soft_contact_shape), N.ptr(c.soft_contact_body_pos), N.ptr(c.soft_contact_body_vel),
           N.ptr(c.soft_contact_normal), N.ptr(f))
    torch.cuda.synchronize()
    count = int(c.soft_contact_count[0])
    # the oracle consumes the GPU's own contact list (generation is checked above)
    ref_c = [(int(i), int(sh), bp, bv, n) for i, sh, bp, bv, n in
             zip(c.soft_contact_particle[:count].tolist(), c.soft_contact_shape[:count].tolist(),
                 _np(c.soft_contact_body_pos[:count]), _np(c.soft_contact_body_vel[:count]),
                 _np(c.soft_contact_normal[:count]))]
    x, v = _np(s.particle_q), _np(s.particle_qd)
    fl = m.particle_flags.cpu().numpy()
    dr = K.xpbd_shape_deltas(x, v, _np(m.particle_inv_mass), _np(m.particle_radius), fl, ref_c,
                             _np(m.shape_material_mu), m.soft_contact_mu, 0.0, 1e-3, 0.8)
    assert np.count_nonzero(np.abs(dr).sum(axis=1)) > 100
    assert np.allclose(_np(d), dr, rtol=0, atol=2e-7)
    fr = K.shape_forces(x, v, _np(m.particle_radius), fl, ref_c, _np(m.shape_material_ke), _np(m.shape_material_kd),
                        _np(m.shape_material_kf), _np(m.shape_material_mu), m.soft_contact_ke, m.soft_contact_kd,
                        m.soft_contact_kf, m.soft_contact_mu, 0.0, np.full((m.particle_count, 3), 0.5))
    assert np.allclose(_np(f), fr, rtol=0, atol=2e-5 * np.abs(fr).max())


def test_xpbd_step_with_shape_contacts_matches_oracle():
    """SolverXPBD.step with particle-shape contacts: positional projection only
    (solver_xpbd.py:386-418), no penalty force, state_in.particle_f only read.  The oracle
    projects the contact list Model.collide produced (its generation is checked above)."""
    from oracle import contacts_oracle as C

    rng = np.random.default_rng(11)
    b = ModelBuilder()
    r = 0.05
    # two well separated layers: one dipping into the plane, one under the sphere; no
    # projection can bring two particles within contact distance (the grid is built before
    # the projections, solver_xpbd.py:315-319, so pairs a projection creates across cells are
    # not candidates in the reference either; the brute-force oracle would see them)
    g = np.stack(np.meshgrid(np.arange(6), np.arange(6), np.arange(2), indexing="ij"), -1).reshape(-1, 3)
    x = (g * np.array([0.3, 0.3, 0.27]) + np.array([-0.75, -0.75, -0.02])) + rng.uniform(-0.01, 0.01, g.shape)
    n = x.shape[0]
    b.add_particles(x.tolist(), (rng.normal(0.0, 0.5, (n, 3))).tolist(), [1.0] * n, [r] * n)
    b.add_ground_plane(height=0.0)
    b.add_shape_sphere(-1, ((0.15, 0.15, 0.33), (0.0, 0.0, 0.0, 1.0)), radius=0.12)
    m = b.finalize(DEV)
    m.particle_mu = 0.3
    sol = SolverXPBD(m, iterations=3)
    s0, s1 = m.state(), m.state()
    s0.particle_f.normal_(0.0, 0.05, generator=torch.Generator(DEV).manual_seed(3))
    contacts = m.contacts()
    m.collide(s0, contacts, soft_contact_margin=0.01)
    x0, v0, f0 = _np(s0.particle_q), _np(s0.particle_qd), _np(s0.particle_f)
    flags = m.particle_flags.cpu().numpy()
    cnt = int(contacts.soft_contact_count.item())
    cl = [(int(i), int(sh), bp, bv, nr) for i, sh, bp, bv, nr in zip(
        contacts.soft_contact_particle[:cnt].cpu().numpy(), contacts.soft_contact_shape[:cnt].cpu().numpy(),
        _np(contacts.soft_contact_body_pos[:cnt]), _np(contacts.soft_contact_body_vel[:cnt]),
        _np(contacts.soft_contact_normal[:cnt]))]
    _print(f"SolverXPBD, {n} particles on a plane and under a sphere, {cnt} shape contacts")
    assert cnt >= 36 and any(c[1] == 1 for c in cl)
    dt = 1.0 / 240.0
    f_before = s0.particle_f.clone()
    sol.step(s0, s1, None, contacts, dt)
    torch.cuda.synchronize()
    assert torch.equal(s0.particle_f, f_before), "SolverXPBD must not write state_in.particle_f"
    inv_mass = m.particle_inv_mass.cpu().numpy().astype(np.float64)
    radius = m.particle_radius.cpu().numpy().astype(np.float64)
    xr, vr = C.xpbd_step(x0, v0, f0, inv_mass, radius, flags, None, _np(m.gravity).reshape(-1, 3), dt,
                         m.particle_max_velocity, 0.3, m.particle_cohesion, 3, shape_contacts=cl,
                         shape_mu=_np(m.shape_material_mu), soft_contact_mu=float(m.soft_contact_mu),
                         particle_adhesion=float(m.particle_adhesion))
    e = np.abs(_np(s1.particle_q) - xr).max(axis=1)
    i = int(np.argmax(e))
    _print(f"  max |x - oracle| {e[i]:.2e} at particle {i} (x0 {x0[i]}, contacts "
           f"{sum(1 for c in cl if c[0] == i)}, gpu {_np(s1.particle_q)[i]}, oracle {xr[i]})")
    assert e[i] < 2e-6
    assert np.abs(_np(s1.particle_qd) - vr).max() < 2e-6 / dt
    # the projection moves every particle that touched the plane out of it (relaxation 0.9)
    assert _np(s1.particle_q)[:, 2].min() > 0.5 * r


@pytest.mark.parametrize("solver", ["xpbd", "semi_implicit"])
def test_particle_comes_to_rest_on_ground_plane(solver):
    _print(f"{solver}: particle dropped on a ground plane")
    b = ModelBuilder(up_axis="Y")
    r = 0.1
    b.add_particle(pos=(0.0, 0.5, 0.0), vel=(0.0, 0.0, 0.0), mass=1.0, radius=r)
    b.add_ground_plane()
    m = b.finalize(DEV)
    sol = SolverXPBD(m, iterations=4) if solver == "xpbd" else SolverSemiImplicit(m)
    s0, s1 = m.state(), m.state()
    contacts = m.contacts()
    dt = 1.0 / 600.0
    for _ in range(1200):
        m.collide(s0, contacts)
        s0.clear_forces()
        sol.step(s0, s1, None, contacts, dt)
        s0, s1 = s1, s0
    q, qd = s0.particle_q.cpu().numpy()[0], s0.particle_qd.cpu().numpy()[0]
    _print(f"  rest height {q[1]:.4f} (radius {r}), |v| {np.linalg.norm(qd):.2e}")
    assert abs(q[1] - r) < 0.02 and np.linalg.norm(qd) < 0.05


def test_sph_fluid_stays_out_of_a_sphere_obstacle():
    cfg = scenes.small_dam_break(12, 10, 14)
    b = scenes.build_dam_break(cfg)
    centre = (0.22, 0.12, 0.03)
    b.add_shape_sphere(-1, (centre, (0.0, 0.0, 0.0, 1.0)), radius=0.05)
    m = b.finalize(DEV)
    _print("SPH dam break flowing onto a static sphere, 300 steps")
    s = SolverSPH(m, cfg.support, warm_start="prev")
    a, c = m.state(), m.state()
    contacts = m.contacts()
    fluid = (m.particle_inv_mass > 0).cpu().numpy()
    for _ in range(300):
        m.collide(a, contacts, soft_contact_margin=cfg.support)
        s.step(a, c, None, contacts, scenes.DT * 5)
        a, c = c, a
    q = a.particle_q.cpu().numpy()[fluid].astype(np.float64)
    r = m.particle_radius.cpu().numpy()[fluid]
    d = np.linalg.norm(q - np.array(centre), axis=1) - 0.05 - r
    _print(f"  min clearance {d.min():.2e} m over {int(fluid.sum())} fluid particles")
    assert np.isfinite(q).all() and d.min() > -0.2 * cfg.spacing
