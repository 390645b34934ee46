"""Solver API: SolverBase (newton/_src/solvers/solver.py:177-366) and SolverSPH.

SolverSPH keeps the reference contract ``step(state_in, state_out, control, contacts, dt)``
(solver.py:301-316) and runs the whole particle step in libnsph (gfx950 HIP kernels).
"""

from __future__ import annotations

import ctypes as C
import enum
from collections.abc import Sequence
from dataclasses import dataclass

import numpy as np
import torch

from . import _native as N
from .sim import AttributeAssignment, AttributeFrequency, Contacts, Control, Model, ModelBuilder, State


class SolverNotifyFlags(enum.IntFlag):
    """newton/_src/solvers/flags.py"""

    JOINT_PROPERTIES = 1 << 0
    JOINT_DOF_PROPERTIES = 1 << 1
    BODY_PROPERTIES = 1 << 2
    BODY_INERTIAL_PROPERTIES = 1 << 3
    SHAPE_PROPERTIES = 1 << 4
    MODEL_PROPERTIES = 1 << 5
    CONSTRAINT_PROPERTIES = 1 << 6
    TENDON_PROPERTIES = 1 << 7
    ACTUATOR_PROPERTIES = 1 << 8


class SolverBase:
    """Generic base class for solvers (solver.py:177-366)."""

    def __init__(self, model: Model):
        self.model = model

    @property
    def device(self) -> torch.device:
        return self.model.device

    def step(self, state_in: State, state_out: State, control: Control | None,
             contacts: Contacts | None, dt: float) -> None:
        raise NotImplementedError()

    def notify_model_changed(self, flags: int) -> None:
        pass

    def update_contacts(self, contacts: Contacts, state: State | None = None) -> None:
        raise NotImplementedError()

    def integrate_particles(self, model: Model, state_in: State, state_out: State, dt: float) -> None:
        """Semi-implicit Euler of the particles (solver.py:266-299, kernel solver.py:10-47):
        v1 = v0 + (f w + g step(-w)) dt clamped to particle_max_velocity, x1 = x0 + v1 dt;
        inactive particles copy x and leave v_out untouched."""
        n = model.particle_count
        if not n:
            return
        ctx = N.context(model.device.index or 0)
        N.set_stream(ctx, torch.cuda.current_stream(model.device).cuda_stream)
        g = model.gravity.reshape(-1, 3).contiguous()
        N.call("nsph_particles_integrate", ctx, n, N.ptr(state_in.particle_q), N.ptr(state_in.particle_qd),
               N.ptr(state_in.particle_f), N.ptr(model.particle_inv_mass), N.ptr(model.particle_flags),
               N.ptr(model.particle_world), N.ptr(g), int(g.shape[0]), float(dt),
               float(model.particle_max_velocity), N.ptr(state_out.particle_q), N.ptr(state_out.particle_qd))

    @classmethod
    def register_custom_attributes(cls, builder: ModelBuilder) -> None:
        pass


_WARM = {"zero": N.WARM_ZERO, "prev": N.WARM_PREV, "mlp": N.WARM_MLP}
# skin > 0: neighbour candidates are searched within H (1 + skin) and kept between steps until
# some particle of the world has moved more than H skin / 2 from its search position.  Off by
# default: on MI355X the cell search costs about what filtering the kept candidates does.
DEFAULT_SKIN = 0.0


def _per_world(value, W: int, name: str) -> np.ndarray:
    arr = np.asarray(value, dtype=np.float64).reshape(-1)
    if arr.size == 1:
        arr = np.full(W, float(arr[0]))
    if arr.size != W:
        raise ValueError(f"{name} needs 1 or {W} values, got {arr.size}")
    return arr


@dataclass
class SPHParams:
    """Resolved SolverSPH parameters (host side, no GPU needed)."""

    n_worlds: int
    smoothing_length: np.ndarray
    boundary_volume: np.ndarray
    domain_lo: np.ndarray  # (W, 3)
    domain_hi: np.ndarray
    rest_density: float
    viscosity: float
    omega: float
    tolerance: float
    min_iters: int
    max_iters: int
    max_neighbors: int
    warm_start: str
    skin: float = DEFAULT_SKIN
    max_skin_neighbors: int = 0

    @staticmethod
    def resolve(model: Model, smoothing_length, *, rest_density: float = 1000.0,
                viscosity: float = 1.0e-4, boundary_volume=None, domain=None, omega: float = 0.5,
                tolerance: float = 1.0e-3, min_iters: int = 2, max_iters: int = 100,
                max_neighbors: int = 96, warm_start: str = "zero", skin: float | None = None,
                max_skin_neighbors: int | None = None) -> "SPHParams":
        if warm_start not in _WARM:
            raise ValueError(f"warm_start must be one of {sorted(_WARM)}")
        skin = DEFAULT_SKIN if skin is None else float(skin)
        if not (skin == 0.0 or 0.01 <= skin <= 1.0):
            raise ValueError(f"skin must be 0 or in [0.01, 1] (got {skin})")
        if max_skin_neighbors is None:
            # the candidates within H (1 + skin) of a uniform fluid, with a 25 % margin
            max_skin_neighbors = int(np.ceil(max_neighbors * (1.0 + skin) ** 3 * 1.25)) if skin > 0 else 0
        max_skin_neighbors = int(max_skin_neighbors)
        if skin > 0 and not (max_neighbors <= max_skin_neighbors <= 8192):
            raise ValueError("max_skin_neighbors must be in [max_neighbors, 8192]")
        W = max(int(model.world_count), 1)
        h = _per_world(smoothing_length, W, "smoothing_length")
        if np.any(h <= 0):
            raise ValueError("smoothing_length must be positive")
        vb = _per_world((h * 0.5) ** 3 if boundary_volume is None else boundary_volume, W,
                        "boundary_volume")
        if domain is not None:
            lo = np.asarray(domain[0], dtype=np.float64).reshape(-1, 3)
            hi = np.asarray(domain[1], dtype=np.float64).reshape(-1, 3)
            if lo.shape[0] == 1:
                lo = np.repeat(lo, W, axis=0)
                hi = np.repeat(hi, W, axis=0)
            if lo.shape != (W, 3) or hi.shape != (W, 3):
                raise ValueError("domain must be (lo, hi) with 1 or world_count boxes")
        else:
            lo = np.zeros((W, 3))
            hi = np.ones((W, 3))
            if model.particle_count:
                q = model.particle_q.detach().cpu().numpy().astype(np.float64)
                wd = model.particle_world.detach().cpu().numpy()
                wd = np.where(wd < 0, 0, wd)
                order = np.argsort(wd, kind="stable")
                bounds = np.searchsorted(wd[order], np.arange(W + 1))
                for w in range(W):
                    sel = q[order[bounds[w]:bounds[w + 1]]]
                    if sel.shape[0]:
                        pad = 2.0 * h[w]
                        lo[w] = sel.min(axis=0) - pad
                        hi[w] = sel.max(axis=0) + pad
        return SPHParams(W, h, vb, lo, hi, float(rest_density), float(viscosity), float(omega),
                         float(tolerance), int(min_iters), int(max_iters), int(max_neighbors),
                         warm_start, skin, max_skin_neighbors)


class SolverSPH(SolverBase):
    """Incompressible SPH (IISPH) particle step on MI355X.

    Fluid particles are ACTIVE with mass > 0; ACTIVE zero-mass particles are boundary
    samples (walls, obstacles) with pseudo-mass rest_density * boundary_volume: static when
    the model gives them no velocity (hashed once from the model), moving otherwise (pistons,
    paddles: the reference integrator moves them with their velocity, and each step hashes
    them from state_in and counts their velocity in the density change); inactive particles
    are carried unchanged (solver.py:27-29).  Worlds never interact.

    Args:
        model: the Model (on a ROCm device).
        smoothing_length: kernel support radius H [m], scalar or one per world.
        rest_density: rho_0 [kg/m^3].
        viscosity: kinematic viscosity nu [m^2/s] of the explicit SPH viscosity term.
        boundary_volume: rest volume of one boundary particle, default (H/2)^3.
        domain: (lo, hi) box per world for the cell tables (particles outside still work,
            they are clamped into border cells); default: bounding box of the world's
            particles at creation, padded by 2 H.
        omega: relaxed Jacobi factor. tolerance: mean compression / rho_0 stop test.
        min_iters, max_iters: Jacobi iteration bounds (the step never stops silently at
            max_iters: `last_report["unconverged_worlds"]` counts such worlds).
        max_neighbors: capacity of the fluid-fluid neighbour list per particle.
        skin: relative skin of the kept neighbour candidates, 0 (default: every step sorts
            and searches) or in [0.01, 1]: a world's candidates are searched within
            H (1 + skin) of each particle's search position and reused by later steps until
            one of its particles is farther than about H skin / 2 from its own, so those steps
            skip the sort and the cell search.  The lists are the same sets either way (in
            the sorted order of the last search); the search positions travel with the State
            (``sph.search_q``), so the step stays a function of the State alone.
        max_skin_neighbors: capacity of the kept candidates per particle (default
            max_neighbors (1 + skin)^3 1.25).
        warm_start: "zero", "prev" (0.5 * previous pressure) or "mlp".
        check_every: host convergence check interval in iterations; 0 launches max_iters
            early-exit iterations with no host read.
        report: collect a step report (one host sync per step) into `last_report`.
        solve: "passes" (one launch per pass over all worlds) or "worlds" (one persistent
            launch in which each workgroup of `solve_block` threads solves whole worlds,
            iterating each to its own stop test on the device; for batches of many worlds),
            after `solve_passes` Jacobi iterations run per pass over all worlds (default 0).
            "auto" (the default): "worlds" after min_iters per-pass iterations for models of
            at least AUTO_WORLDS worlds, "passes" otherwise.  Results are identical; see
            set_solve.

    Walls are boundary particles.  Static collision shapes (ModelBuilder.add_shape_*) act
    through `contacts` (from Model.collide): after the SPH step each fluid particle in contact
    is projected out of the shapes with the XPBD particle-shape correction.

    State: state_in.particle_f (forces the caller applied) is integrated together with the SPH
    force, as integrate_particles integrates particle_f (solver.py:24-46); the step writes
    neither particle_f array (clear_forces() between steps, as the reference's loops do), and
    the SPH force of the last step is `sph_force` (N x 3, particle order).  With the custom
    attributes of register_custom_attributes, each State carries the step's pressure in
    ``state.sph.pressure`` and the search positions of its kept neighbour candidates in
    ``state.sph.search_q`` (particle order): the step reads the warm-start input and the
    search positions from state_in and writes the new ones to state_out, so a recorded State
    resumes bit for bit whatever the warm start and the skin.  Without them the solver keeps
    its own copy of the last step's.
    """

    def __init__(self, model: Model, smoothing_length: float | Sequence[float], *,
                 rest_density: float = 1000.0, viscosity: float = 1.0e-4,
                 boundary_volume: float | Sequence[float] | None = None,
                 domain: tuple | None = None, omega: float = 0.5, tolerance: float = 1.0e-3,
                 min_iters: int = 2, max_iters: int = 100, max_neighbors: int = 96,
                 warm_start: str = "zero", check_every: int = 1, report: bool = True,
                 skin: float | None = None, max_skin_neighbors: int | None = None,
                 gather_block: int = 256, use_graph: bool = False, graph_slots: int | None = None,
                 solve: str = "auto", solve_block: int = 512, solve_passes: int | None = None):
        super().__init__(model)
        if model.device.type != "cuda":
            raise RuntimeError("SolverSPH runs on a ROCm GPU only (model.device is %s)" % model.device)
        prm = SPHParams.resolve(model, smoothing_length, rest_density=rest_density,
                                viscosity=viscosity, boundary_volume=boundary_volume, domain=domain,
                                omega=omega, tolerance=tolerance, min_iters=min_iters,
                                max_iters=max_iters, max_neighbors=max_neighbors,
                                warm_start=warm_start, skin=skin,
                                max_skin_neighbors=max_skin_neighbors)
        self.params = prm
        self.W = prm.n_worlds
        for name in ("smoothing_length", "boundary_volume", "rest_density", "viscosity", "omega",
                     "tolerance", "min_iters", "max_iters", "max_neighbors", "warm_start", "skin",
                     "max_skin_neighbors",
                     "domain_lo", "domain_hi"):
            setattr(self, name, getattr(prm, name))
        self.check_every = int(check_every)
        self.report = bool(report)
        self.shape_relaxation = 1.0
        self.last_report: dict | None = None
        self.gather_block = int(gather_block)
        self.search_variant = self.SEARCH_VARIANTS[0]
        if solve != "auto" and solve not in self.SOLVE_MODES:
            raise ValueError(f"solve must be 'auto' or one of {self.SOLVE_MODES}")
        if solve == "auto":
            # batches of many worlds: iteration control on the device, the first min_iters
            # iterations as launches over all worlds (their L2 locality), the rest per world
            solve = "worlds" if self.W >= self.AUTO_WORLDS else "passes"
            if solve_passes is None:
                solve_passes = max(int(min_iters), 1)
        self.solve = solve
        self.solve_block = int(solve_block)
        self.solve_passes = int(solve_passes or 0)
        self.use_graph = bool(use_graph)
        self.graph_slots = graph_slots
        self._graphs: dict = {}
        self._handle = None
        self._create()

    # ---- set-up ---------------------------------------------------------------------------
    def _create(self):
        m = self.model
        dev = m.device.index or 0
        self._ctx = N.context(dev)
        N.set_stream(self._ctx, torch.cuda.current_stream(m.device).cuda_stream)
        W = self.W
        g = m.gravity.detach().cpu().numpy().astype(np.float32).reshape(-1, 3)
        if g.shape[0] < W:
            g = np.concatenate([g, np.repeat(g[:1], W - g.shape[0], axis=0)])
        self._gravity = np.ascontiguousarray(g[:W])
        self._h = np.ascontiguousarray(self.smoothing_length)
        self._vb = np.ascontiguousarray(self.boundary_volume)
        self._lo = np.ascontiguousarray(self.domain_lo.reshape(-1))
        self._hi = np.ascontiguousarray(self.domain_hi.reshape(-1))
        md = N.ModelDesc()
        md.n_particles = m.particle_count
        md.n_worlds = W
        md.d_q = N.ptr(m.particle_q)
        md.d_qd = N.ptr(m.particle_qd)
        md.d_mass = N.ptr(m.particle_mass)
        md.d_inv_mass = N.ptr(m.particle_inv_mass)
        md.d_flags = N.ptr(m.particle_flags)
        md.d_world = N.ptr(m.particle_world)
        md.h_gravity = self._gravity.ctypes.data_as(C.POINTER(C.c_float))
        md.v_max = float(m.particle_max_velocity)
        pr = N.SphParams()
        pr.h_support = self._h.ctypes.data_as(C.POINTER(C.c_double))
        pr.h_boundary_volume = self._vb.ctypes.data_as(C.POINTER(C.c_double))
        pr.h_domain_lo = self._lo.ctypes.data_as(C.POINTER(C.c_double))
        pr.h_domain_hi = self._hi.ctypes.data_as(C.POINTER(C.c_double))
        pr.rest_density = self.rest_density
        pr.viscosity = self.viscosity
        pr.omega = self.omega
        pr.tolerance = self.tolerance
        pr.min_iters = self.min_iters
        pr.max_iters = self.max_iters
        pr.max_neighbors = self.max_neighbors
        pr.skin = self.skin
        pr.max_skin_neighbors = self.max_skin_neighbors
        pr.warm_start = _WARM[self.warm_start] if self.warm_start != "mlp" else N.WARM_ZERO
        self._mlp = None
        h = C.c_void_p()
        N.call("nsph_solver_create", self._ctx, C.byref(md), C.byref(pr), C.byref(h))
        self._handle = h
        N.call("nsph_solver_set_launch", h, self.gather_block)
        N.call("nsph_solver_set_search", h, self.search_variant)
        N.call("nsph_solver_set_solve", h, self.SOLVE_MODES.index(self.solve), self.solve_block, self.solve_passes)
        # SPH force (viscosity + pressure) of the last step, particle order; the reference's
        # solvers leave state_out.particle_f to the caller, so the step does too
        self.sph_force = torch.zeros((m.particle_count, 3), dtype=torch.float32, device=m.device)
        nf, nb = C.c_int(), C.c_int()
        N.call("nsph_solver_num_fluid", h, C.byref(nf), C.byref(nb))
        self.fluid_count = nf.value
        self.boundary_count = nb.value

    @classmethod
    def register_custom_attributes(cls, builder: ModelBuilder) -> None:
        """Per-particle State attribute ``sph:pressure`` (solver.py:358-366 hook; custom
        attributes as builder.py:1266): the pressure of the last step, the warm-start input
        of the next, carried in the State so that recordings and resumes keep it; and
        ``sph:search_q`` (N x 3): the positions around which the step's kept neighbour
        candidates were searched, which make the lists, like the pressure, part of the State."""
        builder.add_custom_attribute(ModelBuilder.CustomAttribute(
            name="pressure", frequency=AttributeFrequency.PARTICLE, dtype=torch.float32, default=0.0,
            assignment=AttributeAssignment.STATE, namespace="sph"))
        # positions around which the kept neighbour candidates were searched (NaN: none yet)
        builder.add_custom_attribute(ModelBuilder.CustomAttribute(
            name="search_q", frequency=AttributeFrequency.PARTICLE, dtype=torch.float32,
            default=(float("nan"),) * 3, assignment=AttributeAssignment.STATE, namespace="sph"))

    def set_mlp(self, mlp) -> None:
        """Attach a learned warm start (newton_sph_dev_amd.learned.WarmStartMLP) and switch
        warm_start to "mlp"; None detaches it and falls back to "zero"."""
        N.call("nsph_solver_set_mlp", self._handle, None if mlp is None else mlp._handle)
        self._mlp = mlp
        self.set_warm_start("mlp" if mlp is not None else "zero")

    def set_warm_start(self, mode: str) -> None:
        if mode not in _WARM:
            raise ValueError(f"warm_start must be one of {sorted(_WARM)}")
        if mode == "mlp" and self._mlp is None:
            raise ValueError("warm_start='mlp' needs set_mlp() first")
        N.call("nsph_solver_set_warm_start", self._handle, _WARM[mode])
        self.warm_start = mode
        self.clear_graphs()

    def clear_graphs(self) -> None:
        """Free every captured step (they hold the warm start, MLP and launch configuration
        of their capture)."""
        for h in self._graphs.values():
            if self._handle is not None:
                N.call("nsph_solver_graph_destroy", self._handle, int(h))
        self._graphs.clear()
        self._graph_refs = []

    def notify_model_changed(self, flags: int) -> None:
        """MODEL_PROPERTIES: re-create (re-classify particles, re-hash the static boundary);
        captured steps are dropped with the old solver."""
        if flags & SolverNotifyFlags.MODEL_PROPERTIES:
            self._graphs.clear()
            self._graph_refs = []
            self._destroy()
            self._create()

    # ---- step ---------------------------------------------------------------------------------
    @staticmethod
    def _pressure_of(state: State):
        ns = getattr(state, "sph", None)
        return getattr(ns, "pressure", None) if ns is not None else None

    @staticmethod
    def _search_q_of(state: State):
        ns = getattr(state, "sph", None)
        return getattr(ns, "search_q", None) if ns is not None else None

    def _io(self, state_in: State, state_out: State) -> tuple:
        m = self.model
        for t in (state_in.particle_q, state_in.particle_qd, state_out.particle_q, state_out.particle_qd):
            if t is None or t.dtype != torch.float32 or t.device != m.device or not t.is_contiguous():
                raise ValueError("state arrays must be contiguous float32 tensors on the model device")
        p_in, p_out = self._pressure_of(state_in), self._pressure_of(state_out)
        xb_in, xb_out = self._search_q_of(state_in), self._search_q_of(state_out)
        for t in (state_in.particle_f, state_out.particle_f, p_in, p_out, xb_in, xb_out):
            if t is not None and (t.dtype != torch.float32 or t.device != m.device or not t.is_contiguous()):
                raise ValueError("particle_f / sph.pressure / sph.search_q must be contiguous float32 tensors "
                                 "on the model device")
        si = N.StateIn(N.ptr(state_in.particle_q), N.ptr(state_in.particle_qd), N.ptr(state_in.particle_f),
                       N.ptr(p_in), N.ptr(xb_in))
        so = N.StateOut(N.ptr(state_out.particle_q), N.ptr(state_out.particle_qd), N.ptr(self.sph_force),
                        N.ptr(p_out), N.ptr(xb_out))
        return si, so

    def step(self, state_in: State, state_out: State, control: Control | None,
             contacts: Contacts | None, dt: float) -> None:
        m = self.model
        if m.particle_count == 0:
            return
        N.set_stream(self._ctx, torch.cuda.current_stream(m.device).cuda_stream)
        si, so = self._io(state_in, state_out)
        if self.use_graph:
            key = (si.d_q, si.d_qd, si.d_f, si.d_p, si.d_xb, so.d_q, so.d_qd, so.d_f, so.d_p, so.d_xb, float(dt))
            h = self._graphs.get(key)
            if h is None:
                h = self.capture(state_in, state_out, dt, self.graph_slots)
                self._graphs[key] = h
            self.replay(h)
        else:
            rep = N.StepReport() if self.report else None
            N.call("nsph_solver_step", self._handle, C.byref(si), C.byref(so), float(dt), self.check_every,
                   C.byref(rep) if rep is not None else None)
            self._check_report(rep)
        if contacts is not None and m.shape_count > 0 and getattr(contacts, "soft_contact_max", 0) > 0:
            self._shape_projection(state_out, contacts, dt)

    def _shape_projection(self, state_out: State, contacts: Contacts, dt: float) -> None:
        """Particle-shape contacts of Model.collide applied after the SPH step, the way
        SolverXPBD projects them (solver_xpbd.py:386-418, xpbd/kernels.py:114-222): one
        projection with relaxation `shape_relaxation`; particles with a correction move by it
        and gain correction / dt of velocity."""
        from .particle_solvers import shape_contact_deltas

        m = self.model
        deltas = torch.zeros_like(state_out.particle_q)
        shape_contact_deltas(m, contacts, state_out.particle_q, state_out.particle_qd, dt, self.shape_relaxation,
                             deltas)
        N.call("nsph_apply_shape_deltas", self._ctx, m.particle_count, N.ptr(m.particle_flags), N.ptr(deltas),
               float(dt), float(m.particle_max_velocity), N.ptr(state_out.particle_q), N.ptr(state_out.particle_qd))

    def _check_report(self, rep) -> None:
        if rep is not None:
            self.last_report = rep.as_dict()
            if rep.status == N.NSPH_ERR_CAPACITY:
                raise N.CapacityError(rep.status, f"a neighbour list exceeded max_neighbors={self.max_neighbors} "
                                                  f"(max count {rep.max_neighbor_count})")
            if rep.status == N.NSPH_ERR_NONFINITE:
                raise N.NsphError(rep.status, "non-finite position or velocity after the step")

    # ---- captured steps (HIP graphs) ------------------------------------------------------------
    def capture(self, state_in: State, state_out: State, dt: float, iter_slots: int | None = None) -> int:
        """Capture one step state_in -> state_out into HIP graphs and return its handle.

        The graph holds the prologue and `iter_slots` Jacobi iterations (default: max_iters,
        fully asynchronous) plus the epilogue; with fewer slots the replay reads the
        convergence count once and finishes eagerly if needed. The tensors must stay alive
        and keep their storage; dt is fixed at capture."""
        N.set_stream(self._ctx, torch.cuda.current_stream(self.model.device).cuda_stream)
        slots = self.max_iters if iter_slots is None else int(iter_slots)
        si, so = self._io(state_in, state_out)
        h = C.c_int()
        N.call("nsph_solver_graph_capture", self._handle, C.byref(si), C.byref(so), float(dt), slots, C.byref(h))
        self._graph_refs = getattr(self, "_graph_refs", []) + [(state_in, state_out)]
        return h.value

    def replay(self, handle: int) -> None:
        """Run a captured step on the current stream (report as for step())."""
        N.set_stream(self._ctx, torch.cuda.current_stream(self.model.device).cuda_stream)
        rep = N.StepReport() if self.report else None
        N.call("nsph_solver_graph_step", self._handle, int(handle), C.byref(rep) if rep is not None else None)
        self._check_report(rep)

    # neighbour-search kernel variants: list quads per particle staged in LDS (the first is the default)
    SEARCH_VARIANTS = (6, 4, 8, 2)

    def set_launch(self, gather_block: int, search: int | None = None) -> None:
        """Gather-kernel workgroup size (64/128/256) and, optionally, the neighbour-search
        variant (LDS list quads, one of SEARCH_VARIANTS); results do not depend on either."""
        N.call("nsph_solver_set_launch", self._handle, int(gather_block))
        self.gather_block = int(gather_block)
        if search is not None:
            N.call("nsph_solver_set_search", self._handle, int(search))
            self.search_variant = int(search)
        self.clear_graphs()

    # pressure-solve schedules: one launch per pass over all worlds, or one persistent launch in
    # which each workgroup solves whole worlds (nsph_solver_set_solve)
    SOLVE_MODES = ("passes", "worlds")
    AUTO_WORLDS = 64  # worlds from which solve="auto" takes the whole-world schedule

    def set_solve(self, mode: str, block: int | None = None, passes: int | None = None) -> None:
        """Schedule of the pressure solve: "passes" (one launch per pass over all worlds, one
        host read of the still-iterating count per step) or "worlds" (one persistent launch,
        each workgroup of `block` = 256/512/1024 threads iterating whole worlds to their own
        stop test, largest world first; with `passes` = k > 0 the first k Jacobi iterations run
        as per-pass launches over all worlds and the persistent launch finishes the worlds
        still iterating).  "worlds" keeps iteration control on the device: no host read, and a
        captured step needs no iteration slots beyond the k.  Results do not depend on it."""
        if mode not in self.SOLVE_MODES:
            raise ValueError(f"solve must be one of {self.SOLVE_MODES}")
        block = self.solve_block if block is None else int(block)
        passes = self.solve_passes if passes is None else int(passes)
        N.call("nsph_solver_set_solve", self._handle, self.SOLVE_MODES.index(mode), block, passes)
        self.solve, self.solve_block, self.solve_passes = mode, block, passes
        self.clear_graphs()

    def collect_report(self) -> dict:
        rep = N.StepReport()
        N.call("nsph_solver_report", self._handle, C.byref(rep))
        self.last_report = rep.as_dict()
        return self.last_report

    # ---- inspection -----------------------------------------------------------------------------
    def fields(self) -> dict[str, torch.Tensor]:
        """Density, pressure and fluid-fluid neighbour count of the last step, particle order."""
        n = self.model.particle_count
        dev = self.model.device
        rho = torch.empty(n, dtype=torch.float32, device=dev)
        p = torch.empty(n, dtype=torch.float32, device=dev)
        cnt = torch.empty(n, dtype=torch.int32, device=dev)
        N.call("nsph_solver_fields", self._handle, N.ptr(rho), N.ptr(p), N.ptr(cnt))
        return {"density": rho, "pressure": p, "neighbor_count": cnt}

    def neighbors(self) -> tuple[torch.Tensor, torch.Tensor]:
        """Fluid-fluid neighbour lists of the last step: counts[N], idx[max_neighbors, N] (ids)."""
        n = self.model.particle_count
        dev = self.model.device
        counts = torch.empty(n, dtype=torch.int32, device=dev)
        idx = torch.full((self.max_neighbors, n), -1, dtype=torch.int32, device=dev)
        N.call("nsph_solver_neighbors", self._handle, N.ptr(counts), N.ptr(idx))
        return counts, idx

    def world_stats(self) -> dict[str, np.ndarray]:
        W = self.W
        it = np.zeros(W, dtype=np.int32)
        err = np.zeros(W, dtype=np.float64)
        st = np.zeros(W, dtype=np.int32)
        N.call("nsph_solver_world_stats", self._handle, it.ctypes.data_as(C.POINTER(C.c_int)),
               err.ctypes.data_as(C.POINTER(C.c_double)), st.ctypes.data_as(C.POINTER(C.c_int)))
        return {"iterations": it, "density_error": err, "state": st}

    def rebuild_lists(self) -> None:
        """Search every world's neighbour candidates anew at the next step (after particles
        were moved by other means than the step, e.g. a teleport of state_in)."""
        N.call("nsph_solver_rebuild_lists", self._handle)

    def reset_pressure(self) -> None:
        """Forget the solver's own copy of the last pressure (States carrying sph.pressure
        keep theirs)."""
        N.call("nsph_solver_reset_pressure", self._handle)

    def warm_pressure(self) -> torch.Tensor:
        """Pressure the last step's Jacobi iterations started from (particle order)."""
        p0 = torch.empty(self.model.particle_count, dtype=torch.float32, device=self.model.device)
        N.call("nsph_solver_warm_pressure", self._handle, N.ptr(p0))
        return p0

    def _destroy(self):
        h = getattr(self, "_handle", None)
        if h is not None and N._lib is not None:
            try:
                N._lib.nsph_solver_destroy(h)
            except Exception:
                pass
        self._handle = None

    def __del__(self):
        self._destroy()
