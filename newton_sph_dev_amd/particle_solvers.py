"""Particle branches of the reference's XPBD and semi-implicit solvers on the HIP HashGrid.

These are the reference's own consumers of ``model.particle_grid``; with them a Newton user
can keep running particle-particle contact scenes (granular piles, the two-particle friction
cases of the reference's tests) on the MI355X grid:
  SolverXPBD         newton/_src/solvers/xpbd/solver_xpbd.py:84-201, 258-441, 504-505, 657-660
                     kernels xpbd/kernels.py:225-289 (contacts), 776-806 (deltas)
  SolverSemiImplicit newton/_src/solvers/semi_implicit/solver_semi_implicit.py:95-177
                     kernel semi_implicit/kernels_contact.py:11-84, launch 559-579
Only particles are simulated: the models built by ModelBuilder here have no bodies, shapes,
springs, triangles or tetrahedra, so those branches of the reference solvers never run.
"""

from __future__ import annotations

import torch

from . import _native as N
from .sim import Contacts, Control, Model, State
from .solvers import SolverBase


def _ctx(model: Model):
    ctx = N.context(model.device.index or 0)
    N.set_stream(ctx, torch.cuda.current_stream(model.device).cuda_stream)
    return ctx


def _grid_ready(grid) -> bool:
    return grid is not None and getattr(grid, "_handle", None) is not None


def shape_contact_deltas(model: Model, contacts: Contacts, q: torch.Tensor, qd: torch.Tensor, dt: float,
                         relaxation: float, deltas: torch.Tensor) -> None:
    """Add the particle-shape contact corrections of xpbd/kernels.py:114-222 into deltas
    (static shapes; particle_mu = model.soft_contact_mu, ka = model.particle_adhesion as
    solver_xpbd.py:386-418 passes them)."""
    N.call("nsph_xpbd_shape_contacts", _ctx(model), model.particle_count, N.ptr(q), N.ptr(qd),
           N.ptr(model.particle_inv_mass), N.ptr(model.particle_radius), N.ptr(model.particle_flags),
           N.ptr(model.shape_material_mu), float(model.soft_contact_mu), float(model.particle_adhesion),
           N.ptr(contacts.soft_contact_start), contacts.soft_contact_max, N.ptr(contacts.soft_contact_shape),
           N.ptr(contacts.soft_contact_body_pos), N.ptr(contacts.soft_contact_body_vel),
           N.ptr(contacts.soft_contact_normal), float(dt), float(relaxation), N.ptr(deltas))


class SolverXPBD(SolverBase):
    """Position-based particle solver (particle subset of solver_xpbd.py).

    One step: integrate the particles into state_out (with state_in.particle_f, which the
    solver only reads), build ``model.particle_grid`` over the predicted positions with radius
    2 * particle_max_radius + particle_cohesion, then `iterations` times: zero the deltas,
    project particle-shape contacts (the contacts of Model.collide; positional only, as
    solver_xpbd.py:386-418 -- no penalty force) and particle-particle contacts (normal +
    Coulomb friction on the relative velocity, relaxation `soft_contact_relaxation`) and
    apply the deltas (v = (x_new - x_init) / dt, clamped).  The position buffers ping-pong between
    state_out and state_in as the reference does without gradients (solver_xpbd.py:157-201),
    so state_in's particle arrays are overwritten; the final result is in state_out.
    """

    def __init__(self, model: Model, iterations: int = 2, soft_body_relaxation: float = 0.9,
                 soft_contact_relaxation: float = 0.9, **kwargs):
        super().__init__(model)
        self.iterations = int(iterations)
        self.soft_body_relaxation = float(soft_body_relaxation)
        self.soft_contact_relaxation = float(soft_contact_relaxation)
        self._particle_delta_counter = 0
        self.particle_q_init = None
        if model.particle_count > 1 and _grid_ready(model.particle_grid):
            model.particle_grid.reserve(model.particle_count)

    def _apply_particle_deltas(self, model: Model, state_in: State, state_out: State,
                               particle_deltas: torch.Tensor, dt: float):
        if self._particle_delta_counter == 0:
            particle_q = state_out.particle_q
            new_q, new_qd = state_in.particle_q, state_in.particle_qd
        else:
            particle_q = state_in.particle_q
            new_q, new_qd = state_out.particle_q, state_out.particle_qd
        self._particle_delta_counter = 1 - self._particle_delta_counter
        N.call("nsph_apply_particle_deltas", _ctx(model), model.particle_count, N.ptr(self.particle_q_init),
               N.ptr(particle_q), N.ptr(model.particle_flags), N.ptr(particle_deltas), float(dt),
               float(model.particle_max_velocity), N.ptr(new_q), N.ptr(new_qd))
        return new_q, new_qd

    def step(self, state_in: State, state_out: State, control: Control | None,
             contacts: Contacts | None, dt: float) -> None:
        model = self.model
        n = model.particle_count
        self._particle_delta_counter = 0
        if not n:
            return
        particle_q, particle_qd = state_out.particle_q, state_out.particle_qd
        self.particle_q_init = state_in.particle_q.clone()
        particle_deltas = torch.empty_like(state_out.particle_qd)
        self.integrate_particles(model, state_in, state_out, dt)
        grid = model.particle_grid
        use_grid = n > 1 and _grid_ready(grid)
        if use_grid:
            grid.build(state_out.particle_q, radius=model.particle_max_radius * 2.0 + model.particle_cohesion)
        shapes = contacts is not None and model.shape_count > 0 and contacts.soft_contact_max > 0
        for _ in range(self.iterations):
            particle_deltas.zero_()
            if shapes:
                shape_contact_deltas(model, contacts, particle_q, particle_qd, dt, self.soft_contact_relaxation,
                                     particle_deltas)
            if use_grid and model.particle_max_radius > 0.0:
                N.set_stream(grid._ctx, torch.cuda.current_stream(model.device).cuda_stream)
                N.call("nsph_xpbd_particle_contacts", grid._handle, n, N.ptr(particle_q), N.ptr(particle_qd),
                       N.ptr(model.particle_inv_mass), N.ptr(model.particle_radius),
                       N.ptr(model.particle_flags), float(model.particle_mu), float(model.particle_cohesion),
                       float(model.particle_max_radius), float(dt), self.soft_contact_relaxation,
                       N.ptr(particle_deltas))
            particle_q, particle_qd = self._apply_particle_deltas(model, state_in, state_out, particle_deltas, dt)
        if particle_q.data_ptr() != state_out.particle_q.data_ptr():
            state_out.particle_q.copy_(particle_q)
            state_out.particle_qd.copy_(particle_qd)


class SolverSemiImplicit(SolverBase):
    """Force-based particle solver (particle subset of solver_semi_implicit.py).

    One step: particle-particle penalty contacts into state_in.particle_f (the reference's
    kernel overwrites the force of every active particle), particle-shape penalty contacts
    subtracted from it (kernels_contact.py:144-262), then integrate_particles.  As in
    the reference, the contact forces need ``model.particle_grid`` built by the caller
    (``model.particle_grid.build(state.particle_q, radius)``); an unbuilt grid adds no
    contact force.
    """

    def __init__(self, model: Model, angular_damping: float = 0.05, friction_smoothing: float = 1.0,
                 joint_attach_ke: float = 1.0e4, joint_attach_kd: float = 1.0e2,
                 enable_tri_contact: bool = True, **kwargs):
        super().__init__(model)
        self.angular_damping = float(angular_damping)
        self.friction_smoothing = float(friction_smoothing)
        self.joint_attach_ke = float(joint_attach_ke)
        self.joint_attach_kd = float(joint_attach_kd)
        self.enable_tri_contact = bool(enable_tri_contact)

    def step(self, state_in: State, state_out: State, control: Control | None,
             contacts: Contacts | None, dt: float) -> None:
        model = self.model
        n = model.particle_count
        if not n:
            return
        grid = model.particle_grid
        if n > 1 and _grid_ready(grid) and grid.built:
            N.set_stream(grid._ctx, torch.cuda.current_stream(model.device).cuda_stream)
            N.call("nsph_eval_particle_contact", grid._handle, n, N.ptr(state_in.particle_q),
                   N.ptr(state_in.particle_qd), N.ptr(model.particle_radius), N.ptr(model.particle_flags),
                   float(model.particle_ke), float(model.particle_kd), float(model.particle_kf),
                   float(model.particle_mu), float(model.particle_cohesion), float(model.particle_max_radius),
                   N.ptr(state_in.particle_f))
        if contacts is not None and model.shape_count > 0 and contacts.soft_contact_max > 0:
            N.call("nsph_eval_shape_contacts", _ctx(model), n, N.ptr(state_in.particle_q), N.ptr(state_in.particle_qd),
                   N.ptr(model.particle_radius), N.ptr(model.particle_flags), N.ptr(model.shape_material_ke),
                   N.ptr(model.shape_material_kd), N.ptr(model.shape_material_kf), N.ptr(model.shape_material_mu),
                   float(model.soft_contact_ke), float(model.soft_contact_kd), float(model.soft_contact_kf),
                   float(model.soft_contact_mu), float(model.particle_adhesion), N.ptr(contacts.soft_contact_start),
                   contacts.soft_contact_max, N.ptr(contacts.soft_contact_shape), N.ptr(contacts.soft_contact_body_pos),
                   N.ptr(contacts.soft_contact_body_vel), N.ptr(contacts.soft_contact_normal),
                   N.ptr(state_in.particle_f))
        self.integrate_particles(model, state_in, state_out, dt)
