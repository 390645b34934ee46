"""Particle subset of Newton's sim data model on PyTorch-ROCm tensors.

Mirrors the names, argument meaning and error behaviour of
  ModelBuilder  newton/_src/sim/builder.py:795 (particles: 7186-7280, 8037-8137,
                world scopes 2815-2907, finalize particle section 9815-9855, gravity 10655-10668)
  Model         newton/_src/sim/model.py:140-187, 827-851
  State         newton/_src/sim/state.py:101-108, 156-229
  Control       newton/_src/sim/control.py:9
  ParticleFlags newton/_src/geometry/flags.py:8-14
Vectors are (N, 3) float32 tensors instead of wp.array(dtype=vec3).
"""

from __future__ import annotations

import enum
from collections.abc import Sequence
from dataclasses import dataclass, replace
from typing import Any

import numpy as np
import torch


class ParticleFlags(enum.IntEnum):
    """Flags for particle properties (newton/_src/geometry/flags.py:8-14)."""

    ACTIVE = 1 << 0


class ShapeFlags(enum.IntFlag):
    """Flags for shape properties (newton/_src/geometry/flags.py:18-33)."""

    VISIBLE = 1 << 0
    COLLIDE_SHAPES = 1 << 1
    COLLIDE_PARTICLES = 1 << 2
    SITE = 1 << 3


class GeoType(enum.IntEnum):
    """Shape geometry types (newton/_src/geometry/types.py:38-80)."""

    NONE = 0
    PLANE = 1
    HFIELD = 2
    SPHERE = 3
    CAPSULE = 4
    ELLIPSOID = 5
    CYLINDER = 6
    BOX = 7
    MESH = 8
    CONE = 9
    CONVEX_MESH = 10


def quat_between_vectors(a, b) -> tuple[float, float, float, float]:
    """Unit quaternion (x, y, z, w) rotating direction a onto direction b."""
    a = np.asarray(a, np.float64) / np.linalg.norm(a)
    b = np.asarray(b, np.float64) / np.linalg.norm(b)
    c = float(np.dot(a, b))
    if c < -1.0 + 1e-12:
        # opposite directions: half turn about any axis orthogonal to a
        axis = np.cross(a, (1.0, 0.0, 0.0))
        if np.linalg.norm(axis) < 1e-6:
            axis = np.cross(a, (0.0, 1.0, 0.0))
        axis /= np.linalg.norm(axis)
        return (float(axis[0]), float(axis[1]), float(axis[2]), 0.0)
    v = np.cross(a, b)
    q = np.array([v[0], v[1], v[2], 1.0 + c])
    q /= np.linalg.norm(q)
    return tuple(float(x) for x in q)


def _as_transform(xform) -> tuple[float, ...]:
    """(px, py, pz, qx, qy, qz, qw) from None, a 7-sequence or a (position, quaternion) pair."""
    if xform is None:
        return (0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 1.0)
    if len(xform) == 2:
        p, q = xform
        return tuple(float(c) for c in (*p, *q))
    if len(xform) == 7:
        return tuple(float(c) for c in xform)
    raise ValueError("a transform is (position, quaternion) or 7 numbers (px, py, pz, qx, qy, qz, qw)")


class Axis(enum.IntEnum):
    """Axes of 3D space (newton/_src/core/types.py:81-130)."""

    X = 0
    Y = 1
    Z = 2

    @classmethod
    def from_string(cls, axis_str: str) -> "Axis":
        a = axis_str.lower()
        if a in ("x", "y", "z"):
            return cls("xyz".index(a))
        raise ValueError(f"Invalid axis string: {axis_str}")

    @classmethod
    def from_any(cls, value) -> "Axis":
        if isinstance(value, cls):
            return value
        if isinstance(value, str):
            return cls.from_string(value)
        if isinstance(value, (int, np.integer)):
            return cls(int(value))
        raise TypeError(f"Cannot convert {type(value).__name__} to Axis")


def quat_to_matrix(q: Sequence[float]) -> np.ndarray:
    """Rotation matrix of a unit quaternion given as (x, y, z, w), in float64 (shape frames)."""
    x, y, z, w = (float(c) for c in q)
    return np.array(
        [
            [1.0 - 2.0 * (y * y + z * z), 2.0 * (x * y - z * w), 2.0 * (x * z + y * w)],
            [2.0 * (x * y + z * w), 1.0 - 2.0 * (x * x + z * z), 2.0 * (y * z - x * w)],
            [2.0 * (x * z - y * w), 2.0 * (y * z + x * w), 1.0 - 2.0 * (x * x + y * y)],
        ],
        dtype=np.float64,
    )


def quat_to_matrix_f32(q: Sequence[float]) -> np.ndarray:
    """Rotation matrix of a unit quaternion (x, y, z, w) in float32, as the reference's
    add_particle_grid takes it from wp.quat_to_matrix (builder.py:8090): column k is the
    rotation of the k-th unit vector by v c + u d + (u x v)(2 w), with c = 2 w^2 - 1,
    d = 2 (u . v), u = (x, y, z), every operation rounded to float32, no fused multiply-add.
    Whether the reference's host build contracts these products is not known here, so a
    non-identity rotation may differ from it in the last float32 bit of an entry."""
    f = np.float32
    x, y, z, w = (f(c) for c in q)
    two = f(2.0)
    c = two * w * w - f(1.0)
    w2 = w * two
    cols = []
    for v in ((f(1), f(0), f(0)), (f(0), f(1), f(0)), (f(0), f(0), f(1))):
        d = two * ((x * v[0] + y * v[1]) + z * v[2])
        cr = (y * v[2] - z * v[1], z * v[0] - x * v[2], x * v[1] - y * v[0])
        cols.append([(v[k] * c + (x, y, z)[k] * d) + cr[k] * w2 for k in range(3)])
    return np.array(cols, dtype=np.float32).T


def _identity_quat(q: Sequence[float]) -> bool:
    return tuple(float(c) for c in q) == (0.0, 0.0, 0.0, 1.0)


class AttributeAssignment(enum.IntEnum):
    """Which object holds a custom attribute (model.py:47-62)."""

    MODEL = 0
    STATE = 1
    CONTROL = 2
    CONTACT = 3


class AttributeFrequency(enum.IntEnum):
    """Entity count a custom attribute follows (model.py:64-100; the particle-relevant subset
    keeps the reference's values)."""

    ONCE = 0
    PARTICLE = 9
    WORLD = 15


class AttributeNamespace:
    """Container of namespaced custom attributes (model.py:106-125)."""

    def __init__(self, name: str):
        self._name = name

    def __repr__(self):
        return f"AttributeNamespace('{self._name}', attributes={[k for k in self.__dict__ if not k.startswith('_')]})"


def _attach(obj, name: str, namespace: str | None, value) -> None:
    if namespace:
        if not hasattr(obj, namespace):
            setattr(obj, namespace, AttributeNamespace(namespace))
        setattr(getattr(obj, namespace), name, value)
    else:
        setattr(obj, name, value)


class ModelBuilder:
    """Builds a particle Model (particle methods of newton/_src/sim/builder.py) with static
    collision shapes (body = -1) for particle-shape contacts."""

    @dataclass
    class CustomAttribute:
        """A per-entity attribute added to the Model at finalize (builder.py:522-570): values
        default to `default` (scalar or fixed-size vector) and are set per particle through the
        `custom_attributes` argument of add_particle / add_particles / add_particle_grid
        (dict of key -> value or per-particle values).  `dtype` is a torch dtype."""

        name: str
        frequency: AttributeFrequency = AttributeFrequency.PARTICLE
        dtype: Any = torch.float32
        default: Any = 0
        assignment: AttributeAssignment = AttributeAssignment.MODEL
        namespace: str | None = None

        @property
        def key(self) -> str:
            return f"{self.namespace}:{self.name}" if self.namespace else self.name

    @dataclass
    class ShapeConfig:
        """Shape material and collision settings (builder.py:231-300); the fields used by
        particle-shape contacts, with the reference's defaults."""

        density: float = 1000.0
        ke: float = 2.5e3
        kd: float = 100.0
        kf: float = 1000.0
        ka: float = 0.0
        mu: float = 1.0
        restitution: float = 0.0
        margin: float = 0.0
        gap: float | None = None
        is_solid: bool = True
        collision_group: int = 1
        collision_filter_parent: bool = True
        has_shape_collision: bool = True
        has_particle_collision: bool = True
        is_visible: bool = True
        is_site: bool = False

        def copy(self) -> "ModelBuilder.ShapeConfig":
            return replace(self)

        def mark_as_site(self) -> None:
            self.is_site = True
            self.is_visible = False
            self.has_shape_collision = False
            self.has_particle_collision = False
            self.density = 0.0

        @property
        def flags(self) -> int:
            f = 0
            f |= int(ShapeFlags.VISIBLE) if self.is_visible else 0
            f |= int(ShapeFlags.COLLIDE_SHAPES) if self.has_shape_collision else 0
            f |= int(ShapeFlags.COLLIDE_PARTICLES) if self.has_particle_collision else 0
            f |= int(ShapeFlags.SITE) if self.is_site else 0
            return f

    def __init__(self, up_axis: Axis | int | str = Axis.Z, gravity: float = -9.81):
        self.up_axis = Axis.from_any(up_axis)
        self.gravity = float(gravity)
        self.world_count = 0
        self.world_gravity: list[tuple[float, float, float]] = []
        self._current_world = -1
        self.default_particle_radius = 0.1
        self.particle_max_velocity = 1e5
        # particle data in chunks (float64 until finalize, as the reference lists hold)
        self._q: list[np.ndarray] = []
        self._qd: list[np.ndarray] = []
        self._mass: list[np.ndarray] = []
        self._radius: list[np.ndarray] = []
        self._flags: list[np.ndarray] = []
        self._world: list[np.ndarray] = []
        self._count = 0
        self.default_shape_cfg = ModelBuilder.ShapeConfig()
        self.default_site_cfg = ModelBuilder.ShapeConfig()
        self.default_site_cfg.mark_as_site()
        self._shapes: list[dict] = []
        self.custom_attributes: dict[str, ModelBuilder.CustomAttribute] = {}
        self._custom_values: dict[str, list[tuple[int, Any]]] = {}

    # ---- custom attributes (builder.py:1266-1340) ----------------------------------------
    def add_custom_attribute(self, attribute: "ModelBuilder.CustomAttribute") -> None:
        """Define a custom attribute; registering the same key again with the same frequency,
        dtype and assignment is a no-op, with a different one a ValueError."""
        key = attribute.key
        old = self.custom_attributes.get(key)
        if old is not None:
            if (old.frequency, old.dtype, old.assignment) != (attribute.frequency, attribute.dtype,
                                                               attribute.assignment):
                raise ValueError(f"Custom attribute '{key}' already exists with an incompatible specification")
            return
        if attribute.frequency not in (AttributeFrequency.PARTICLE, AttributeFrequency.WORLD, AttributeFrequency.ONCE):
            raise ValueError(f"Custom attribute '{key}': unsupported frequency {attribute.frequency!r}")
        self.custom_attributes[key] = attribute
        self._custom_values[key] = []

    def _record_custom(self, custom_attributes: dict | None, first: int, n: int) -> None:
        for key, value in (custom_attributes or {}).items():
            if key not in self.custom_attributes:
                raise AttributeError(f"Custom attribute '{key}' is not defined (add_custom_attribute)")
            if self.custom_attributes[key].frequency != AttributeFrequency.PARTICLE:
                raise ValueError(f"Custom attribute '{key}' is not per particle")
            self._custom_values[key].append((first, n, value))

    # ---- worlds (builder.py:2815-2907) -------------------------------------------------
    @property
    def current_world(self) -> int:
        return self._current_world

    @property
    def up_vector(self) -> tuple[float, float, float]:
        v = [0.0, 0.0, 0.0]
        v[int(self.up_axis)] = 1.0
        return tuple(v)

    def begin_world(self, label: str | None = None, attributes: dict | None = None, gravity=None):
        if self._current_world != -1:
            raise RuntimeError(
                f"Cannot begin a new world: already in world context (current_world={self._current_world}). "
                "Call end_world() first to close the current world context."
            )
        self._current_world = self.world_count
        self.world_count += 1
        if gravity is not None:
            self.world_gravity.append(tuple(float(g) for g in gravity))
        else:
            u = self.up_vector
            self.world_gravity.append((u[0] * self.gravity, u[1] * self.gravity, u[2] * self.gravity))

    def end_world(self):
        if self._current_world == -1:
            raise RuntimeError("Cannot end world: not currently in a world context (current_world is already -1).")
        self._current_world = -1

    # ---- particles (builder.py:7186-7280, 8037-8137) ------------------------------------
    @property
    def particle_count(self) -> int:
        return self._count

    @property
    def particle_q(self) -> np.ndarray:
        return np.concatenate(self._q) if self._q else np.zeros((0, 3))

    def _append(self, pos, vel, mass, radius, flags):
        n = pos.shape[0]
        self._q.append(np.asarray(pos, dtype=np.float64).reshape(n, 3))
        self._qd.append(np.asarray(vel, dtype=np.float64).reshape(n, 3))
        self._mass.append(np.asarray(mass, dtype=np.float64).reshape(n))
        self._radius.append(np.asarray(radius, dtype=np.float64).reshape(n))
        self._flags.append(np.asarray(flags, dtype=np.int64).reshape(n))
        self._world.append(np.full(n, self._current_world, dtype=np.int32))
        self._count += n

    def add_particle(self, pos, vel, mass: float, radius: float | None = None,
                     flags: int = ParticleFlags.ACTIVE, custom_attributes: dict | None = None) -> int:
        if radius is None:
            radius = self.default_particle_radius
        self._append(np.array([pos], dtype=np.float64), np.array([vel], dtype=np.float64),
                     [mass], [radius], [int(flags)])
        self._record_custom(custom_attributes, self._count - 1, 1)
        return self._count - 1

    def add_particles(self, pos, vel, mass, radius=None, flags=None, custom_attributes: dict | None = None):
        pos = np.asarray(pos, dtype=np.float64).reshape(-1, 3)
        n = pos.shape[0]
        vel = np.asarray(vel, dtype=np.float64).reshape(n, 3)
        mass = np.asarray(mass, dtype=np.float64).reshape(n)
        if radius is None:
            radius = np.full(n, self.default_particle_radius)
        if flags is None:
            flags = np.full(n, int(ParticleFlags.ACTIVE))
        self._append(pos, vel, mass, np.asarray(radius, dtype=np.float64).reshape(n),
                     np.asarray(flags, dtype=np.int64).reshape(n))
        self._record_custom(custom_attributes, self._count - n, n)

    def add_particle_grid(self, pos, rot, vel, dim_x: int, dim_y: int, dim_z: int,
                          cell_x: float, cell_y: float, cell_z: float, mass: float, jitter: float,
                          radius_mean: float | None = None, radius_std: float = 0.0,
                          flags: list[int] | int | None = None,
                          custom_attributes: dict[str, Any] | None = None):
        """Regular lattice of particles (builder.py:8037-8137).

        Lattice order is numpy meshgrid's default 'xy' indexing: y slowest, then x, then z
        fastest (builder.py:8084-8087). The rotation matrix is float32 (wp.quat_to_matrix,
        builder.py:8090), applied to the float64 points with the translation in float64
        (builder.py:8091); jitter (rng.random - 0.5) * jitter from
        numpy.random.default_rng(42 + particle_count) (builder.py:8095-8096); radii
        radius_mean (+ standard_normal * radius_std) from the same generator (8098-8103).
        """
        px = np.arange(dim_x) * cell_x
        py = np.arange(dim_y) * cell_y
        pz = np.arange(dim_z) * cell_z
        points = np.stack(np.meshgrid(px, py, pz)).reshape(3, -1).T
        rot_mat = quat_to_matrix_f32(rot)
        points = points @ rot_mat.reshape(3, 3).T + np.array(pos, dtype=np.float64)
        velocity = np.broadcast_to(np.array(vel, dtype=np.float64).reshape(1, 3), points.shape)
        rng = np.random.default_rng(42 + self._count)
        points += (rng.random(points.shape) - 0.5) * jitter
        if radius_mean is None:
            radius_mean = self.default_particle_radius
        radii = np.full(points.shape[0], fill_value=radius_mean, dtype=np.float64)
        if radius_std > 0.0:
            radii += rng.standard_normal(radii.shape) * radius_std
        n = points.shape[0]
        masses = np.full(n, mass, dtype=np.float64)
        fl = np.full(n, int(ParticleFlags.ACTIVE) if flags is None else 0, dtype=np.int64)
        if flags is not None:
            fl[:] = int(flags) if np.isscalar(flags) else np.asarray(flags)
        self.add_particles(points, velocity, masses, radii, fl, custom_attributes=custom_attributes)

    # ---- finalize (builder.py:9745, 9815-9855, 10655-10668) -----------------------------
    # ---- static shapes (builder.py:5449-5990) ----------------------------------------------
    @property
    def shape_count(self) -> int:
        return len(self._shapes)

    def add_shape(self, body: int, type: int, xform=None, cfg: "ModelBuilder.ShapeConfig | None" = None,
                  scale=(1.0, 1.0, 1.0), is_static: bool = False, label: str | None = None,
                  custom_attributes: dict | None = None, color=None) -> int:
        if body != -1:
            raise NotImplementedError("only static shapes (body=-1) are supported: rigid bodies are not modelled")
        cfg = self.default_shape_cfg if cfg is None else cfg
        self._shapes.append({
            "xform": _as_transform(xform), "type": int(type), "scale": tuple(float(c) for c in scale),
            "world": self._current_world, "flags": cfg.flags, "ke": cfg.ke, "kd": cfg.kd, "kf": cfg.kf,
            "ka": cfg.ka, "mu": cfg.mu, "restitution": cfg.restitution,
            "label": label if label is not None else f"shape_{len(self._shapes)}",
        })
        return len(self._shapes) - 1

    def _shape_cfg(self, cfg, as_site: bool):
        if cfg is None:
            return self.default_site_cfg if as_site else self.default_shape_cfg
        if as_site:
            cfg = cfg.copy()
            cfg.mark_as_site()
        return cfg

    def add_shape_plane(self, plane=(0.0, 0.0, 1.0, 0.0), xform=None, width: float = 10.0, length: float = 10.0,
                        body: int = -1, cfg=None, color=None, label: str | None = None,
                        custom_attributes: dict | None = None) -> int:
        """Plane with normal along its local Z; from the equation a x + b y + c z + d = 0 when
        xform is None.  width/length 0 make it infinite for collision."""
        if xform is None:
            n = np.asarray(plane[:3], np.float64)
            norm = np.linalg.norm(n)
            n = n / norm
            pos = -(plane[3] / norm) * n
            xform = (tuple(pos), quat_between_vectors((0.0, 0.0, 1.0), n))
        return self.add_shape(body, GeoType.PLANE, xform, self._shape_cfg(cfg, False), (width, length, 0.0),
                              is_static=True, label=label)

    def add_ground_plane(self, height: float = 0.0, cfg=None, color=None, label: str | None = None) -> int:
        return self.add_shape_plane(plane=(*self.up_vector, -height), width=0.0, length=0.0, cfg=cfg,
                                    label=label or "ground_plane")

    def add_shape_sphere(self, body: int, xform=None, radius: float = 1.0, cfg=None, as_site: bool = False,
                         color=None, label: str | None = None, custom_attributes: dict | None = None) -> int:
        return self.add_shape(body, GeoType.SPHERE, xform, self._shape_cfg(cfg, as_site), (radius, 0.0, 0.0),
                              label=label)

    def add_shape_box(self, body: int, xform=None, hx: float = 0.5, hy: float = 0.5, hz: float = 0.5, cfg=None,
                      as_site: bool = False, color=None, label: str | None = None,
                      custom_attributes: dict | None = None) -> int:
        return self.add_shape(body, GeoType.BOX, xform, self._shape_cfg(cfg, as_site), (hx, hy, hz), label=label)

    def add_shape_capsule(self, body: int, xform=None, radius: float = 1.0, half_height: float = 0.5, cfg=None,
                          as_site: bool = False, color=None, label: str | None = None,
                          custom_attributes: dict | None = None) -> int:
        """Capsule along its local Z axis."""
        return self.add_shape(body, GeoType.CAPSULE, xform, self._shape_cfg(cfg, as_site),
                              (radius, half_height, 0.0), label=label)

    def add_shape_cylinder(self, body: int, xform=None, radius: float = 1.0, half_height: float = 0.5, cfg=None,
                           as_site: bool = False, color=None, label: str | None = None,
                           custom_attributes: dict | None = None) -> int:
        """Cylinder along its local Z axis."""
        return self.add_shape(body, GeoType.CYLINDER, xform, self._shape_cfg(cfg, as_site),
                              (radius, half_height, 0.0), label=label)

    def finalize(self, device: str | torch.device | None = None, requires_grad: bool = False) -> Model:
        if device is None:
            device = "cuda:0" if torch.cuda.is_available() else "cpu"
        device = torch.device(device)
        m = Model(device)
        n = self._count
        q = np.concatenate(self._q) if n else np.zeros((0, 3))
        qd = np.concatenate(self._qd) if n else np.zeros((0, 3))
        mass = np.concatenate(self._mass).astype(np.float32) if n else np.zeros(0, np.float32)
        radius = np.concatenate(self._radius) if n else np.zeros(0)
        flags = np.concatenate(self._flags) if n else np.zeros(0, np.int64)
        world = np.concatenate(self._world) if n else np.zeros(0, np.int32)
        inv_mass = np.divide(1.0, mass, out=np.zeros_like(mass), where=mass != 0.0)
        W = max(self.world_count, 1)
        m.world_count = self.world_count
        m.particle_q = torch.from_numpy(q.astype(np.float32)).to(device)
        m.particle_qd = torch.from_numpy(qd.astype(np.float32)).to(device)
        m.particle_mass = torch.from_numpy(mass).to(device)
        m.particle_inv_mass = torch.from_numpy(inv_mass.astype(np.float32)).to(device)
        m.particle_radius = torch.from_numpy(radius.astype(np.float32)).to(device)
        m.particle_flags = torch.from_numpy(flags.astype(np.int32)).to(device)
        m.particle_world = torch.from_numpy(world.astype(np.int32)).to(device)
        m.particle_max_radius = float(np.max(radius)) if n else 0.0
        m.particle_max_velocity = float(self.particle_max_velocity)
        m.particle_world_start = torch.from_numpy(_world_starts(world, self.world_count)).to(device)
        if self.world_gravity:
            g = np.array(self.world_gravity, dtype=np.float32)
        else:
            u = self.up_vector
            g = np.array([[u[0] * self.gravity, u[1] * self.gravity, u[2] * self.gravity]] * W,
                         dtype=np.float32)
        m.gravity = torch.from_numpy(g.reshape(-1, 3)).to(device)
        m.up_axis = self.up_axis
        S = len(self._shapes)
        m.shape_count = S
        if S:
            col = lambda k, dt: torch.tensor([sh[k] for sh in self._shapes], dtype=dt, device=device)  # noqa: E731
            m.shape_transform = col("xform", torch.float32)
            m.shape_body = torch.full((S,), -1, dtype=torch.int32, device=device)
            m.shape_type = col("type", torch.int32)
            m.shape_scale = col("scale", torch.float32)
            m.shape_world = col("world", torch.int32)
            m.shape_flags = col("flags", torch.int32)
            m.shape_material_ke = col("ke", torch.float32)
            m.shape_material_kd = col("kd", torch.float32)
            m.shape_material_kf = col("kf", torch.float32)
            m.shape_material_ka = col("ka", torch.float32)
            m.shape_material_mu = col("mu", torch.float32)
            m.shape_material_restitution = col("restitution", torch.float32)
            m.shape_label = [sh["label"] for sh in self._shapes]
        for key, attr in self.custom_attributes.items():
            count = {AttributeFrequency.PARTICLE: n, AttributeFrequency.WORLD: W,
                     AttributeFrequency.ONCE: 1}[attr.frequency]
            dflt = torch.as_tensor(attr.default, dtype=attr.dtype)
            t = dflt.expand((count,) + tuple(dflt.shape)).clone()
            for first, cnt, value in self._custom_values[key]:
                t[first:first + cnt] = torch.as_tensor(np.asarray(value), dtype=attr.dtype)
            m.add_attribute(attr.name, t.to(device), attr.frequency, attr.assignment, attr.namespace)
        if n > 1 and m.particle_max_radius > 0.0:
            from .hashgrid import HashGrid

            m.particle_grid = HashGrid(128, 128, 128, device=device)
        else:
            m.particle_grid = None
        return m


def _world_starts(world: np.ndarray, world_count: int) -> np.ndarray:
    """particle_world_start of builder.py:9545-9604: [start_0 .. start_{W-1}, global tail, total]."""
    n = world.shape[0]
    front = 0
    for w in world:
        if w == -1:
            front += 1
        else:
            break
    starts = np.zeros(world_count + 2, dtype=np.int32)
    counts = np.bincount(world[world >= 0], minlength=world_count) if world_count else np.zeros(0, int)
    acc = front
    for w in range(world_count):
        starts[w] = acc
        acc += int(counts[w])
    starts[world_count] = acc
    starts[world_count + 1] = n
    return starts


class Model:
    """Particle part of newton.Model (newton/_src/sim/model.py:140-187)."""

    AttributeAssignment = AttributeAssignment
    AttributeFrequency = AttributeFrequency
    AttributeNamespace = AttributeNamespace

    def __init__(self, device: torch.device | str = "cpu"):
        self.device = torch.device(device)
        self.attribute_frequency: dict[str, AttributeFrequency] = {}
        self.attribute_assignment: dict[str, AttributeAssignment] = {}
        self.requires_grad = False
        self.world_count = 0
        self.particle_q = None
        self.particle_qd = None
        self.particle_mass = None
        self.particle_inv_mass = None
        self.particle_radius = None
        self.particle_max_radius = 0.0
        self.particle_ke = 1.0e3
        self.particle_kd = 1.0e2
        self.particle_kf = 1.0e2
        self.particle_mu = 0.5
        self.particle_cohesion = 0.0
        self.particle_adhesion = 0.0
        self.particle_grid = None
        self.particle_flags = None
        self.particle_max_velocity = 1e5
        self.particle_world = None
        self.particle_world_start = None
        self.gravity = None
        self.up_axis = Axis.Z
        self.body_count = 0
        self.shape_count = 0
        self.shape_transform = None
        self.shape_body = None
        self.shape_type = None
        self.shape_scale = None
        self.shape_world = None
        self.shape_flags = None
        self.shape_material_ke = None
        self.shape_material_kd = None
        self.shape_material_kf = None
        self.shape_material_ka = None
        self.shape_material_mu = None
        self.shape_material_restitution = None
        self.shape_label = []
        self.soft_contact_ke = 1.0e3
        self.soft_contact_kd = 10.0
        self.soft_contact_kf = 1.0e3
        self.soft_contact_mu = 0.5
        self.soft_contact_restitution = 0.0

    @property
    def particle_count(self) -> int:
        return 0 if self.particle_q is None else int(self.particle_q.shape[0])

    def add_attribute(self, name: str, attrib, frequency, assignment=None, namespace: str | None = None) -> None:
        """Add a custom attribute (model.py:1112-1160); STATE ones are cloned into every
        state(), as model.py:875 does."""
        if isinstance(attrib, torch.Tensor) and attrib.device != self.device:
            raise AttributeError(f"Attribute '{name}' device mismatch (model={self.device}, got={attrib.device})")
        holder = getattr(self, namespace, None) if namespace else self
        if holder is not None and hasattr(holder, name):
            raise AttributeError(f"Attribute already exists: {namespace + '.' if namespace else ''}{name}")
        _attach(self, name, namespace, attrib)
        key = f"{namespace}:{name}" if namespace else name
        self.attribute_frequency[key] = frequency
        if assignment is not None:
            self.attribute_assignment[key] = assignment

    def _add_custom_attributes(self, destination, assignment: AttributeAssignment) -> None:
        for key in self.attribute_frequency:
            if self.attribute_assignment.get(key, AttributeAssignment.MODEL) != assignment:
                continue
            namespace, name = key.split(":", 1) if ":" in key else (None, key)
            src = getattr(getattr(self, namespace), name) if namespace else getattr(self, name)
            _attach(destination, name, namespace, src.clone() if isinstance(src, torch.Tensor) else src)

    def state(self, requires_grad: bool | None = None) -> State:
        s = State()
        if self.particle_count:
            s.particle_q = self.particle_q.clone()
            s.particle_qd = self.particle_qd.clone()
            s.particle_f = torch.zeros_like(self.particle_qd)
        self._add_custom_attributes(s, AttributeAssignment.STATE)
        return s

    def control(self, requires_grad: bool | None = None, clone_variables: bool = True) -> Control:
        return Control()

    def contacts(self, soft_contact_max: int | None = None) -> Contacts:
        """Contact buffers for collide(): soft_contact_max defaults to shape_count *
        particle_count, as the reference's collision pipeline allocates (collide.py:814-815)."""
        if soft_contact_max is None:
            soft_contact_max = self.shape_count * self.particle_count
        return Contacts(int(soft_contact_max), self.particle_count, self.device)

    def set_gravity(self, gravity, world: int | None = None) -> None:
        """Runtime gravity change (model.py:923-958): one vector for every world, one per
        world ((world_count, 3)), or one vector for `world` only.  Solvers that keep a copy
        need notify_model_changed(SolverNotifyFlags.MODEL_PROPERTIES) afterwards."""
        g = np.asarray(gravity, dtype=np.float32)
        if world is not None:
            if g.shape != (3,):
                raise ValueError("Expected single gravity vector (3,) when world is specified")
            if world < 0 or world >= self.world_count:
                raise IndexError(f"world {world} out of range [0, {self.world_count})")
            self.gravity[world] = torch.from_numpy(g).to(self.gravity.device)
        elif g.ndim == 1:
            self.gravity[:] = torch.from_numpy(g.reshape(1, 3)).to(self.gravity.device)
        else:
            if len(g) != self.world_count:
                raise ValueError(f"Expected {self.world_count} gravity vectors, got {len(g)}")
            self.gravity.copy_(torch.from_numpy(g.reshape(-1, 3)).to(self.gravity.device))

    def collide(self, state: State, contacts: Contacts | None = None, *, soft_contact_margin: float = 0.01,
                **kwargs) -> Contacts:
        """Contact generation (model.py:996-1030): particle-shape soft contacts against the
        static shapes (collide.py:1194-1230; there are no rigid bodies, hence no rigid
        contacts).  Particle-particle contacts are found by the solvers through particle_grid."""
        if contacts is None:
            contacts = self.contacts()
        n = self.particle_count
        if self.shape_count == 0 or n == 0 or contacts.soft_contact_max == 0:
            contacts.clear()
            return contacts
        from . import _native as N

        ctx = N.context(self.device.index or 0)
        N.set_stream(ctx, torch.cuda.current_stream(self.device).cuda_stream)
        N.call("nsph_soft_contacts", ctx, n, N.ptr(state.particle_q), N.ptr(self.particle_radius),
               N.ptr(self.particle_flags), N.ptr(self.particle_world), self.shape_count, N.ptr(self.shape_transform),
               N.ptr(self.shape_type), N.ptr(self.shape_scale), N.ptr(self.shape_world), N.ptr(self.shape_flags),
               float(soft_contact_margin), contacts.soft_contact_max, N.ptr(contacts._work),
               N.ptr(contacts.soft_contact_start), N.ptr(contacts.soft_contact_count),
               N.ptr(contacts.soft_contact_particle), N.ptr(contacts.soft_contact_shape),
               N.ptr(contacts.soft_contact_body_pos), N.ptr(contacts.soft_contact_body_vel),
               N.ptr(contacts.soft_contact_normal))
        return contacts


class State:
    """Particle part of newton.State (newton/_src/sim/state.py:101-108, 156-229)."""

    def __init__(self) -> None:
        self.particle_q: torch.Tensor | None = None
        self.particle_qd: torch.Tensor | None = None
        self.particle_f: torch.Tensor | None = None

    @property
    def particle_count(self) -> int:
        return 0 if self.particle_q is None else int(self.particle_q.shape[0])

    def clear_forces(self) -> None:
        if self.particle_count:
            self.particle_f.zero_()

    def assign(self, other: State) -> None:
        """Copy the arrays of another state into this one (state.py:169-229), custom
        namespaced attributes included."""
        for name in ("particle_q", "particle_qd", "particle_f"):
            dst, src = getattr(self, name), getattr(other, name)
            if (dst is None) != (src is None):
                raise ValueError(f"State array '{name}' is present in only one of the two states.")
            if dst is not None:
                dst.copy_(src)
        ns_self = {k: v for k, v in self.__dict__.items() if isinstance(v, AttributeNamespace)}
        ns_other = {k: v for k, v in other.__dict__.items() if isinstance(v, AttributeNamespace)}
        for ns in ns_self.keys() | ns_other.keys():
            a, b = ns_self.get(ns), ns_other.get(ns)
            if a is None or b is None:
                held = (b if a is None else a).__dict__.values()
                if any(isinstance(v, torch.Tensor) for v in held):
                    raise ValueError(f"State namespace '{ns}' holds arrays in only one of the two states.")
                continue
            for name in a.__dict__.keys() | b.__dict__.keys():
                if name.startswith("_"):
                    continue
                dst, src = a.__dict__.get(name), b.__dict__.get(name)
                if isinstance(dst, torch.Tensor) != isinstance(src, torch.Tensor):
                    raise ValueError(f"State array '{ns}.{name}' is present in only one of the two states.")
                if isinstance(dst, torch.Tensor):
                    dst.copy_(src)


class Control:
    """newton.Control (no particle controls are used by the SPH step)."""


class Contacts:
    """Particle-shape soft contacts (newton/_src/sim/contacts.py:90-350, soft-contact part).

    Arrays of capacity soft_contact_max: soft_contact_particle, soft_contact_shape,
    soft_contact_body_pos / _body_vel / _normal (world frame for static shapes);
    soft_contact_count[0] = contacts found (may exceed the capacity).  Contacts are ordered by
    (particle, shape); soft_contact_start[i] .. soft_contact_start[i + 1] are particle i's."""

    def __init__(self, soft_contact_max: int = 0, particle_count: int = 0, device: torch.device | str = "cpu"):
        dev = torch.device(device)
        self.soft_contact_max = int(soft_contact_max)
        self.rigid_contact_max = 0
        m = max(self.soft_contact_max, 0)
        self.soft_contact_count = torch.zeros(1, dtype=torch.int32, device=dev)
        self.soft_contact_particle = torch.full((m,), -1, dtype=torch.int32, device=dev)
        self.soft_contact_shape = torch.full((m,), -1, dtype=torch.int32, device=dev)
        self.soft_contact_body_pos = torch.zeros((m, 3), dtype=torch.float32, device=dev)
        self.soft_contact_body_vel = torch.zeros((m, 3), dtype=torch.float32, device=dev)
        self.soft_contact_normal = torch.zeros((m, 3), dtype=torch.float32, device=dev)
        self.soft_contact_start = torch.zeros(particle_count + 1, dtype=torch.int32, device=dev)
        # nsph_soft_contacts' workspace: a count per particle and one per 256 particles
        self._work = torch.zeros(max(particle_count + (particle_count + 255) // 256, 1), dtype=torch.int32,
                                 device=dev)

    def clear(self) -> None:
        self.soft_contact_count.zero_()
        self.soft_contact_start.zero_()
        self.soft_contact_particle.fill_(-1)
        self.soft_contact_shape.fill_(-1)
