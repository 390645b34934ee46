"""Python wrapper of the CPU oracle (liboracle.so) -- TEST INFRASTRUCTURE ONLY.

Used by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as the checker /
CPU baseline. The product (newton_sph_dev_amd) never imports this module.
"""

from __future__ import annotations

import ctypes as C
import subprocess
from dataclasses import dataclass
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB


class _Problem(C.Structure):
    _fields_ = [
        ("n", C.c_int),
        ("n_worlds", C.c_int),
        ("q0", C.c_void_p),
        ("qd0", C.c_void_p),
        ("mass", C.c_void_p),
        ("inv_mass", C.c_void_p),
        ("flags", C.c_void_p),
        ("world", C.c_void_p),
        ("gravity", C.c_void_p),
        ("v_max", C.c_float),
        ("support", C.c_void_p),
        ("boundary_volume", C.c_void_p),
        ("domain_lo", C.c_void_p),
        ("domain_hi", C.c_void_p),
        ("rest_density", C.c_double),
        ("viscosity", C.c_double),
        ("omega", C.c_double),
        ("tolerance", C.c_double),
        ("min_iters", C.c_int),
        ("max_iters", C.c_int),
        ("max_neighbors", C.c_int),
        ("warm_start", C.c_int),
        ("skin", C.c_double),
        ("max_skin_neighbors", C.c_int),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        _lib = C.CDLL(str(LIB))
        for name in ("orc_step_f32", "orc_step_f64"):
            getattr(_lib, name).restype = C.c_int
        _lib.orc_num_threads.restype = C.c_int
    return _lib


def _p(a: np.ndarray | None):
    return C.c_void_p(0) if a is None else C.c_void_p(a.ctypes.data)


@dataclass
class Problem:
    """Numpy copy of a particle Model plus the SPH parameters (all arrays contiguous)."""

    q0: np.ndarray
    qd0: np.ndarray
    mass: np.ndarray
    inv_mass: np.ndarray
    flags: np.ndarray
    world: np.ndarray
    gravity: np.ndarray
    support: np.ndarray
    boundary_volume: np.ndarray
    domain_lo: np.ndarray
    domain_hi: np.ndarray
    v_max: float = 1e5
    rest_density: float = 1000.0
    viscosity: float = 1e-4
    omega: float = 0.5
    tolerance: float = 1e-3
    min_iters: int = 2
    max_iters: int = 100
    max_neighbors: int = 64
    warm_start: int = 0
    skin: float = 0.0
    max_skin_neighbors: int = 0

    @property
    def n(self) -> int:
        return int(self.q0.shape[0])

    @property
    def n_worlds(self) -> int:
        return int(self.support.shape[0])

    def _struct(self) -> _Problem:
        for name in ("q0", "qd0", "mass", "inv_mass", "gravity"):
            setattr(self, name, np.ascontiguousarray(getattr(self, name), dtype=np.float32))
        for name in ("flags", "world"):
            setattr(self, name, np.ascontiguousarray(getattr(self, name), dtype=np.int32))
        for name in ("support", "boundary_volume", "domain_lo", "domain_hi"):
            setattr(self, name, np.ascontiguousarray(getattr(self, name), dtype=np.float64))
        s = _Problem()
        s.n, s.n_worlds = self.n, self.n_worlds
        s.q0, s.qd0, s.mass, s.inv_mass = _p(self.q0), _p(self.qd0), _p(self.mass), _p(self.inv_mass)
        s.flags, s.world, s.gravity = _p(self.flags), _p(self.world), _p(self.gravity)
        s.v_max = self.v_max
        s.support, s.boundary_volume = _p(self.support), _p(self.boundary_volume)
        s.domain_lo, s.domain_hi = _p(self.domain_lo), _p(self.domain_hi)
        s.rest_density, s.viscosity, s.omega, s.tolerance = (
            self.rest_density, self.viscosity, self.omega, self.tolerance)
        s.min_iters, s.max_iters, s.max_neighbors, s.warm_start = (
            self.min_iters, self.max_iters, self.max_neighbors, self.warm_start)
        s.skin, s.max_skin_neighbors = float(self.skin), int(self.max_skin_neighbors)
        return s


def problem_from_model(model, solver_like) -> Problem:
    """Problem from a finalized Model (any device) and an object carrying the SolverSPH
    parameters (smoothing_length, boundary_volume, domain_lo/hi, rest_density, ...)."""
    def cpu(t):
        return t.detach().cpu().numpy()

    W = len(solver_like.smoothing_length)
    g = cpu(model.gravity).astype(np.float32).reshape(-1, 3)
    if g.shape[0] < W:
        g = np.concatenate([g, np.repeat(g[:1], W - g.shape[0], axis=0)])
    warm = {"zero": 0, "prev": 1, "mlp": 2}[getattr(solver_like, "warm_start", "zero")]
    return Problem(
        q0=cpu(model.particle_q), qd0=cpu(model.particle_qd), mass=cpu(model.particle_mass),
        inv_mass=cpu(model.particle_inv_mass), flags=cpu(model.particle_flags),
        world=cpu(model.particle_world), gravity=g[:W],
        support=np.asarray(solver_like.smoothing_length, dtype=np.float64),
        boundary_volume=np.asarray(solver_like.boundary_volume, dtype=np.float64),
        domain_lo=np.asarray(solver_like.domain_lo, dtype=np.float64).reshape(-1),
        domain_hi=np.asarray(solver_like.domain_hi, dtype=np.float64).reshape(-1),
        v_max=float(model.particle_max_velocity), rest_density=solver_like.rest_density,
        viscosity=solver_like.viscosity, omega=solver_like.omega, tolerance=solver_like.tolerance,
        min_iters=solver_like.min_iters, max_iters=solver_like.max_iters,
        max_neighbors=solver_like.max_neighbors, warm_start=warm,
        skin=float(getattr(solver_like, "skin", 0.0)),
        max_skin_neighbors=int(getattr(solver_like, "max_skin_neighbors", 0)))


def step(pb: Problem, q_in, qd_in, pprev=None, dt: float = 1e-4, precision: str = "f64",
         forced_iters=None, nthreads: int = 0, f_in=None, xb=None) -> dict:
    """One SPH step on the CPU. Returns q, qd, f, rho, p, xb (particle order), neighbour
    counts, per-world iterations and errors, the updated pprev (fluid-list order) and the status.
    f_in: optional caller force per particle (N x 3), added to the SPH force before the
    reference integrator (solver.py:24-46 integrates the total particle_f).
    xb: optional search positions of the input state (N x 3; the previous step's r["xb"]): with
    pb.skin > 0 a world whose particles all lie within the rebuild distance of them takes its
    neighbour candidates from them, as the solver does; None searches every world anew."""
    L = lib()
    s = pb._struct()
    n, W = pb.n, pb.n_worlds
    q_in = np.ascontiguousarray(q_in, dtype=np.float32)
    qd_in = np.ascontiguousarray(qd_in, dtype=np.float32)
    nf = int(np.count_nonzero(((pb.flags & 1) != 0) & (pb.inv_mass > 0)))
    pp = np.zeros(nf + 1, np.float32) if pprev is None else np.ascontiguousarray(
        np.concatenate([np.asarray(pprev, np.float32).reshape(-1), [0.0]]), dtype=np.float32)
    out = {
        "q": np.zeros((n, 3), np.float32), "qd": np.zeros((n, 3), np.float32),
        "f": np.zeros((n, 3), np.float32), "rho": np.zeros(n, np.float32),
        "p": np.zeros(n, np.float32), "xb": np.zeros((n, 3), np.float32), "nbr_count": np.zeros(n, np.int32),
        "iters": np.zeros(W, np.int32), "err": np.zeros(W, np.float64),
    }
    fi = None if forced_iters is None else np.ascontiguousarray(forced_iters, dtype=np.int32)
    nt = nthreads if nthreads > 0 else L.orc_num_threads()
    fn = L.orc_step_f32 if precision == "f32" else L.orc_step_f64
    fin = None if f_in is None else np.ascontiguousarray(f_in, dtype=np.float32).reshape(-1)
    xbi = None if xb is None else np.ascontiguousarray(xb, dtype=np.float32).reshape(-1)
    st = fn(C.byref(s), _p(q_in), _p(qd_in), _p(fin), _p(xbi), _p(pp), C.c_float(dt), _p(fi), _p(out["q"]),
            _p(out["qd"]), _p(out["f"]), _p(out["rho"]), _p(out["p"]), _p(out["xb"]), _p(out["nbr_count"]),
            _p(out["iters"]), _p(out["err"]), C.c_int(nt))
    out["status"] = st
    out["pprev"] = pp[:nf].copy()
    return out


def neighbors_brute(x: np.ndarray, flags: np.ndarray | None, H2: float, cap: int = 96,
                    nthreads: int = 0):
    """O(N^2) FP32 neighbour sets. Returns (counts[n], idx[n, cap])."""
    L = lib()
    x = np.ascontiguousarray(x, dtype=np.float32)
    n = x.shape[0]
    fl = None if flags is None else np.ascontiguousarray(flags, dtype=np.int32)
    counts = np.zeros(n, np.int32)
    idx = np.full((n, cap), -1, np.int32)
    nt = nthreads if nthreads > 0 else L.orc_num_threads()
    st = L.orc_neighbors_brute(_p(x), _p(fl), C.c_int(n), C.c_float(H2), C.c_int(cap), _p(counts),
                               _p(idx), C.c_int(nt))
    if st:
        raise RuntimeError(f"brute-force neighbour list exceeded cap={cap}")
    return counts, idx


def solver_neighbors(pb: Problem, q_in, xb=None):
    """Fluid-fluid lists as the solver builds them (ids, slot-major [cap, n]); xb as in step()."""
    L = lib()
    s = pb._struct()
    n = pb.n
    q_in = np.ascontiguousarray(q_in, dtype=np.float32)
    xbi = None if xb is None else np.ascontiguousarray(xb, dtype=np.float32).reshape(-1)
    counts = np.zeros(n, np.int32)
    idx = np.full((pb.max_neighbors, n), -1, np.int32)
    st = L.orc_solver_neighbors(C.byref(s), _p(q_in), _p(xbi), _p(counts), _p(idx))
    return counts, idx, st


def sorted_order(pb: Problem, q_in, xb=None):
    """Fluid ids in the solver's sorted order; xb as in step()."""
    L = lib()
    s = pb._struct()
    q_in = np.ascontiguousarray(q_in, dtype=np.float32)
    xbi = None if xb is None else np.ascontiguousarray(xb, dtype=np.float32).reshape(-1)
    ids = np.zeros(pb.n + 1, np.int32)
    nf, nb = C.c_int(), C.c_int()
    L.orc_sorted_order(C.byref(s), _p(q_in), _p(xbi), _p(ids), C.byref(nf), C.byref(nb))
    return ids[: nf.value].copy(), nf.value, nb.value


def kernel_f64(H: float, r2: float) -> tuple[float, float]:
    """Cubic spline of support H at squared distance r2: (W, (dW/dr)/r) in FP64."""
    L = lib()
    W, gf = C.c_double(), C.c_double()
    L.orc_kernel_f64(C.c_double(H), C.c_double(r2), C.byref(W), C.byref(gf))
    return W.value, gf.value


def viscosity(pb: Problem, q_in, qd_in, precision: str = "f64") -> tuple[np.ndarray, np.ndarray]:
    """Viscous acceleration a_visc (N x 3) and density (N) of one state, particle order, from
    the same density and viscosity passes as the step (0 for non-fluid particles)."""
    L = lib()
    s = pb._struct()
    q_in = np.ascontiguousarray(q_in, dtype=np.float32)
    qd_in = np.ascontiguousarray(qd_in, dtype=np.float32)
    av = np.zeros((pb.n, 3), np.float64)
    rho = np.zeros(pb.n, np.float64)
    fn = L.orc_viscosity_f32 if precision == "f32" else L.orc_viscosity_f64
    st = fn(C.byref(s), _p(q_in), _p(qd_in), _p(av), _p(rho))
    if st:
        raise RuntimeError(f"orc_viscosity failed ({st})")
    return av, rho


def operator_f64(pb: Problem, q_in, qd_in, p_fluid, dt: float = 1e-4) -> dict:
    """IISPH operator pieces in FP64 (fluid-list order): rho, aii, s, Ap = A p, ap = a^p(p)."""
    L = lib()
    s = pb._struct()
    q_in = np.ascontiguousarray(q_in, dtype=np.float32)
    qd_in = np.ascontiguousarray(qd_in, dtype=np.float32)
    p = np.ascontiguousarray(p_fluid, dtype=np.float64)
    nf = p.shape[0]
    out = {k: np.zeros(nf, np.float64) for k in ("rho", "aii", "s", "Ap")}
    out["ap"] = np.zeros((nf, 3), np.float64)
    st = L.orc_operator_f64(C.byref(s), _p(q_in), _p(qd_in), C.c_float(dt), _p(p), _p(out["rho"]),
                            _p(out["aii"]), _p(out["s"]), _p(out["Ap"]), _p(out["ap"]))
    if st:
        raise RuntimeError(f"orc_operator_f64 failed ({st})")
    return out
