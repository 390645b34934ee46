"""HashGrid: periodic spatial hash grid on the GPU (replaces warp.HashGrid).

Usage as in Newton: ``Model.particle_grid = HashGrid(128, 128, 128)``
(newton/_src/sim/builder.py:9851-9855), ``grid.reserve(n)`` (solver_xpbd.py:136-139),
``grid.build(points, radius)`` (solver_xpbd.py:315-319); kernels then visit the points
in cell-sorted order (``point_ids``, = hash_grid_point_id) and query neighbours
(kernels_contact.py:54-73).  Here the query is exposed as whole-array calls that fill
fixed-capacity neighbour lists: ``neighbors`` (grid points against themselves, ACTIVE
filter and j != i as the reference consumers do) and ``query`` (arbitrary points).
"""

from __future__ import annotations

import ctypes as C

import torch

from . import _native as N


class HashGrid:
    def __init__(self, dim_x: int, dim_y: int, dim_z: int, device: str | torch.device | None = None):
        self.dim_x, self.dim_y, self.dim_z = int(dim_x), int(dim_y), int(dim_z)
        self.device = torch.device(device) if device is not None else torch.device("cuda:0")
        self.reserved = 0
        self.built = False
        self.radius = 0.0
        self._points = None
        self._handle = None
        if self.device.type == "cuda":
            self._open()

    def _open(self):
        ctx = N.context(self.device.index or 0)
        h = C.c_void_p()
        N.call("nsph_grid_create", ctx, self.dim_x, self.dim_y, self.dim_z, C.byref(h))
        self._ctx = ctx
        self._handle = h

    def _require(self):
        if self._handle is None:
            raise RuntimeError("HashGrid needs a ROCm (cuda) device; the CPU has no grid implementation")
        N.set_stream(self._ctx, torch.cuda.current_stream(self.device).cuda_stream)

    @property
    def id(self) -> int:
        """Opaque handle value (the reference passes grid.id to kernels)."""
        return int(self._handle.value) if self._handle is not None else 0

    def reserve(self, num_points: int) -> None:
        self._require()
        N.call("nsph_grid_reserve", self._handle, int(num_points))
        self.reserved = max(self.reserved, int(num_points))

    def build(self, points: torch.Tensor, radius: float) -> None:
        self._require()
        pts = points.reshape(-1, 3).contiguous()
        if pts.dtype != torch.float32 or pts.device != self.device:
            raise ValueError("points must be a float32 (N, 3) tensor on the grid's device")
        N.call("nsph_grid_build", self._handle, N.ptr(pts), int(pts.shape[0]), float(radius))
        self.reserved = max(self.reserved, int(pts.shape[0]))
        self.radius = float(radius)
        self.built = True
        self._points = pts

    @property
    def num_points(self) -> int:
        n = C.c_int()
        N.call("nsph_grid_num_points", self._handle, C.byref(n))
        return n.value

    def point_ids(self) -> torch.Tensor:
        """Point ids in cell-sorted order (hash_grid_point_id for tid = 0..n-1)."""
        self._require()
        out = torch.empty(self.num_points, dtype=torch.int32, device=self.device)
        N.call("nsph_grid_point_ids", self._handle, N.ptr(out))
        return out

    def neighbors(self, radius: float | None = None, flags: torch.Tensor | None = None,
                  max_neighbors: int = 64) -> tuple[torch.Tensor, torch.Tensor]:
        """(counts[n], idx[max_neighbors, n]) for the built points: j != i, ACTIVE (if flags
        given) and |x_i - x_j|^2 < radius^2 in FP32."""
        self._require()
        n = self.num_points
        r = self.radius if radius is None else float(radius)
        counts = torch.empty(n, dtype=torch.int32, device=self.device)
        idx = torch.full((max_neighbors, n), -1, dtype=torch.int32, device=self.device)
        N.call("nsph_grid_neighbors", self._handle, N.ptr(flags), r, int(max_neighbors),
               N.ptr(counts), N.ptr(idx))
        return counts, idx

    def query(self, points: torch.Tensor, radius: float, max_neighbors: int = 64):
        """(counts[nq], idx[max_neighbors, nq]) of grid points within radius of each query point."""
        self._require()
        q = points.reshape(-1, 3).contiguous()
        nq = int(q.shape[0])
        counts = torch.empty(nq, dtype=torch.int32, device=self.device)
        idx = torch.full((max_neighbors, nq), -1, dtype=torch.int32, device=self.device)
        N.call("nsph_grid_query", self._handle, N.ptr(q), nq, float(radius), int(max_neighbors),
               N.ptr(counts), N.ptr(idx))
        return counts, idx

    def count(self, points: torch.Tensor, radius: float) -> torch.Tensor:
        """Number of grid points within radius of each query point (the library's consumer
        of the in-kernel query of include/nsph_device.h)."""
        self._require()
        q = points.reshape(-1, 3).contiguous()
        nq = int(q.shape[0])
        counts = torch.empty(nq, dtype=torch.int32, device=self.device)
        N.call("nsph_grid_count", self._handle, N.ptr(q), nq, float(radius), N.ptr(counts))
        return counts

    def __del__(self):
        h = getattr(self, "_handle", None)
        if h is not None and N._lib is not None:
            try:
                N._lib.nsph_grid_destroy(h)
            except Exception:
                pass
            self._handle = None
