"""ctypes binding of libnsph.so (the gfx950 HIP library).

The library is loaded from this package directory; there is no fallback. Every call
returns an int status; `check()` turns a non-zero status into an exception carrying
nsph_last_error().
"""

from __future__ import annotations

import ctypes as C
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libnsph.so"

NSPH_OK = 0
NSPH_ERR_INVALID = 1
NSPH_ERR_HIP = 2
NSPH_ERR_CAPACITY = 3
NSPH_ERR_NOT_BUILT = 4
NSPH_ERR_NONFINITE = 5
NSPH_ERR_NOT_CONVERGED = 6

WARM_ZERO = 0
WARM_PREV = 1
WARM_MLP = 2


class NsphError(RuntimeError):
    """Error status returned by libnsph."""

    def __init__(self, status: int, message: str):
        super().__init__(f"libnsph status {status}: {message}")
        self.status = status


class CapacityError(NsphError):
    """A capacity (reserve / neighbour list) was exceeded."""


class NotBuiltError(NsphError):
    """The grid was queried before build()."""


class ModelDesc(C.Structure):
    _fields_ = [
        ("n_particles", C.c_int),
        ("n_worlds", C.c_int),
        ("d_q", C.c_void_p),
        ("d_qd", C.c_void_p),
        ("d_mass", C.c_void_p),
        ("d_inv_mass", C.c_void_p),
        ("d_flags", C.c_void_p),
        ("d_world", C.c_void_p),
        ("h_gravity", C.POINTER(C.c_float)),
        ("v_max", C.c_float),
    ]


class SphParams(C.Structure):
    _fields_ = [
        ("h_support", C.POINTER(C.c_double)),
        ("h_boundary_volume", C.POINTER(C.c_double)),
        ("h_domain_lo", C.POINTER(C.c_double)),
        ("h_domain_hi", C.POINTER(C.c_double)),
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


class StepReport(C.Structure):
    _fields_ = [
        ("status", C.c_int),
        ("max_iterations", C.c_int),
        ("sum_iterations", C.c_int),
        ("unconverged_worlds", C.c_int),
        ("max_neighbor_count", C.c_int),
        ("neighbor_overflow", C.c_int),
        ("searched_worlds", C.c_int),
        ("max_density_error", C.c_double),
    ]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


class StateIn(C.Structure):
    """nsph_state_in: device pointers of state_in's particle arrays (q, qd, f, sph pressure,
    sph search positions)."""

    _fields_ = [("d_q", C.c_void_p), ("d_qd", C.c_void_p), ("d_f", C.c_void_p), ("d_p", C.c_void_p),
                ("d_xb", C.c_void_p)]


class StateOut(C.Structure):
    """nsph_state_out: device pointers of state_out's particle arrays (q, qd, f, sph pressure,
    sph search positions)."""

    _fields_ = [("d_q", C.c_void_p), ("d_qd", C.c_void_p), ("d_f", C.c_void_p), ("d_p", C.c_void_p),
                ("d_xb", C.c_void_p)]


_vp = C.c_void_p
_i = C.c_int
_f = C.c_float
_pi = C.POINTER(C.c_int)
_pd = C.POINTER(C.c_double)

# name -> (argtypes); every function returns c_int except nsph_last_error
SIGNATURES: dict[str, list] = {
    "nsph_ctx_create": [_i, _vp, C.POINTER(_vp)],
    "nsph_ctx_set_stream": [_vp, _vp],
    "nsph_ctx_synchronize": [_vp],
    "nsph_ctx_destroy": [_vp],
    "nsph_abi_version": [],
    "nsph_grid_create": [_vp, _i, _i, _i, C.POINTER(_vp)],
    "nsph_grid_reserve": [_vp, _i],
    "nsph_grid_build": [_vp, _vp, _i, _f],
    "nsph_grid_point_ids": [_vp, _vp],
    "nsph_grid_num_points": [_vp, _pi],
    "nsph_grid_neighbors": [_vp, _vp, _f, _i, _vp, _vp],
    "nsph_grid_query": [_vp, _vp, _i, _f, _i, _vp, _vp],
    "nsph_grid_destroy": [_vp],
    "nsph_grid_device_view": [_vp, _vp],
    "nsph_grid_count": [_vp, _vp, _i, _f, _vp],
    "nsph_solver_create": [_vp, C.POINTER(ModelDesc), C.POINTER(SphParams), C.POINTER(_vp)],
    "nsph_solver_step": [_vp, C.POINTER(StateIn), C.POINTER(StateOut), _f, _i, C.POINTER(StepReport)],
    "nsph_solver_report": [_vp, C.POINTER(StepReport)],
    "nsph_solver_graph_capture": [_vp, C.POINTER(StateIn), C.POINTER(StateOut), _f, _i, _pi],
    "nsph_solver_graph_step": [_vp, _i, C.POINTER(StepReport)],
    "nsph_solver_graph_destroy": [_vp, _i],
    "nsph_solver_warm_pressure": [_vp, _vp],
    "nsph_solver_fields": [_vp, _vp, _vp, _vp],
    "nsph_solver_world_stats": [_vp, _pi, _pd, _pi],
    "nsph_solver_neighbors": [_vp, _vp, _vp],
    "nsph_solver_set_warm_start": [_vp, _i],
    "nsph_solver_reset_pressure": [_vp],
    "nsph_solver_rebuild_lists": [_vp],
    "nsph_solver_num_fluid": [_vp, _pi, _pi],
    "nsph_solver_set_timing": [_vp, _i],
    "nsph_solver_set_launch": [_vp, _i],
    "nsph_solver_set_search": [_vp, _i],
    "nsph_solver_set_solve": [_vp, _i, _i, _i],
    "nsph_solver_timing": [_vp, _pd, _pi],
    "nsph_kernel_count": [],
    "nsph_solver_iterations_launched": [_vp, _pi],
    "nsph_solver_destroy": [_vp],
    "nsph_particles_integrate": [_vp, _i, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _i, _f, _f, _vp, _vp],
    "nsph_xpbd_particle_contacts": [_vp, _i, _vp, _vp, _vp, _vp, _vp, _f, _f, _f, _f, _f, _vp],
    "nsph_apply_particle_deltas": [_vp, _i, _vp, _vp, _vp, _vp, _f, _f, _vp, _vp],
    "nsph_eval_particle_contact": [_vp, _i, _vp, _vp, _vp, _vp, _f, _f, _f, _f, _f, _f, _vp],
    "nsph_soft_contacts": [_vp, _i, _vp, _vp, _vp, _vp, _i, _vp, _vp, _vp, _vp, _vp, _f, _i, _vp, _vp, _vp, _vp,
                           _vp, _vp, _vp, _vp],
    "nsph_soft_contacts_static": [_vp, _i, _vp, _vp, _vp, _vp, _i, _vp, _vp, _vp, _vp, _vp, _f, _i, _vp, _vp, _vp,
                                  _vp, _vp, _vp, _vp],
    "nsph_xpbd_shape_contacts": [_vp, _i, _vp, _vp, _vp, _vp, _vp, _vp, _f, _f, _vp, _i, _vp, _vp, _vp, _vp, _f,
                                 _f, _vp],
    "nsph_eval_shape_contacts": [_vp, _i, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _f, _f, _f, _f, _f, _vp, _i, _vp,
                                 _vp, _vp, _vp, _vp],
    "nsph_apply_shape_deltas": [_vp, _i, _vp, _vp, _f, _f, _vp, _vp],
    "nsph_mlp_num_params": [],
    "nsph_mlp_create": [_vp, C.POINTER(_vp)],
    "nsph_mlp_destroy": [_vp],
    "nsph_mlp_set_params": [_vp, C.POINTER(C.c_float)],
    "nsph_mlp_get_params": [_vp, C.POINTER(C.c_float)],
    "nsph_mlp_forward": [_vp, _vp, _i, _vp],
    "nsph_mlp_grad": [_vp, _vp, _vp, _i, C.POINTER(C.c_float), _vp],
    "nsph_mlp_adam": [_vp, _f, _f, _f, _f],
    "nsph_mfma_selftest": [_vp, C.POINTER(C.c_float), C.POINTER(C.c_float), C.POINTER(C.c_float)],
    "nsph_solver_set_mlp": [_vp, _vp],
    "nsph_solver_mlp_features": [_vp, _vp, _vp],
}

_lib: C.CDLL | None = None


def load() -> C.CDLL:
    """Load libnsph.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it with `python -m newton_sph_dev_amd.build` "
            "(the SPH path has no CPU fallback)"
        )
    lib = C.CDLL(str(LIB_PATH))
    for name, args in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = C.c_int
    lib.nsph_last_error.argtypes = []
    lib.nsph_last_error.restype = C.c_char_p
    lib.nsph_kernel_name.argtypes = [C.c_int]
    lib.nsph_kernel_name.restype = C.c_char_p
    _lib = lib
    return lib


def last_error() -> str:
    msg = load().nsph_last_error()
    return msg.decode() if msg else ""


def check(status: int) -> None:
    if status == NSPH_OK:
        return
    msg = last_error()
    if status == NSPH_ERR_CAPACITY:
        raise CapacityError(status, msg)
    if status == NSPH_ERR_NOT_BUILT:
        raise NotBuiltError(status, msg)
    raise NsphError(status, msg)


def call(name: str, *args) -> None:
    check(getattr(load(), name)(*args))


_contexts: dict[int, C.c_void_p] = {}


def context(device: int) -> C.c_void_p:
    """One library context per HIP device (stream is set per call)."""
    ctx = _contexts.get(device)
    if ctx is None:
        ctx = C.c_void_p()
        call("nsph_ctx_create", device, None, C.byref(ctx))
        _contexts[device] = ctx
    return ctx


def set_stream(ctx: C.c_void_p, stream_ptr: int) -> None:
    call("nsph_ctx_set_stream", ctx, C.c_void_p(stream_ptr))


def ptr(t) -> C.c_void_p:
    """Device pointer of a torch tensor (None -> NULL)."""
    if t is None:
        return C.c_void_p(0)
    if not t.is_contiguous():
        raise ValueError("libnsph needs contiguous tensors")
    return C.c_void_p(t.data_ptr())
