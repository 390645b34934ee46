"""Trajectory recording and playback in the file format of Newton's ViewerFile.

Mirrors the names, arguments and file layout of
  ViewerFile, RingBuffer          newton/_src/viewer/viewer_file.py:33-121, 1021-1432
  array records and their cache   viewer_file.py:124-218, 276-392, 531-602, 899-1019
so a recording written here loads in the reference viewer and the reference's recordings
(particle arrays) load here:
  - `.json`: the serialized tree written with `json.dump`; numpy payloads are JSON strings
    of nested lists, with "dtype" = str(dtype).
  - `.bin`: the same tree as CBOR; numpy payloads are raw C-order bytes ("binary_data"),
    with "dtype" = dtype.str.
  - Device arrays (torch tensors here, warp arrays in the reference) are
    {"__type__": "warp.array", "__dtype__": "<class 'warp.types.vec3f'>", "cache_index": k,
    "data": <numpy record>}; a repeated array is {"__type__": "warp.array_ref",
    "cache_index": k}; objects are {"__type__", "__module__", "attributes"}; primitives
    {"__type__", "value"}; containers {"__type__", "items"}.
The file's root is {"model": <Model>, "states": [<{name: array}> per frame]}.
"""

from __future__ import annotations

import json
import os
from collections.abc import Iterable, Mapping
from pathlib import Path
from typing import Any

import numpy as np
import torch

from . import cbor
from .sim import AttributeNamespace

__all__ = ["RingBuffer", "ViewerFile"]


class RingBuffer:
    """List-like history that keeps only the last `capacity` items (oldest first)."""

    def __init__(self, capacity: int = 100):
        if capacity < 1:
            raise ValueError("capacity must be >= 1")
        self.capacity = int(capacity)
        self._items: list = []
        self._head = 0  # physical index of the oldest item once full

    def append(self, item) -> None:
        if len(self._items) < self.capacity:
            self._items.append(item)
        else:
            self._items[self._head] = item
            self._head = (self._head + 1) % self.capacity

    def __len__(self) -> int:
        return len(self._items)

    def _phys(self, index) -> int:
        if not isinstance(index, int):
            raise TypeError("Index must be an integer")
        if not 0 <= index < len(self._items):
            raise IndexError(f"Index {index} out of range [0, {len(self._items)})")
        return (self._head + index) % len(self._items) if len(self._items) == self.capacity else index

    def __getitem__(self, index):
        return self._items[self._phys(index)]

    def __setitem__(self, index, value) -> None:
        self._items[self._phys(index)] = value

    def __iter__(self):
        for i in range(len(self._items)):
            yield self[i]

    def clear(self) -> None:
        self._items = []
        self._head = 0

    def to_list(self) -> list:
        return list(self)

    def from_list(self, items) -> None:
        self.clear()
        for x in items:
            self.append(x)


# ---- array records --------------------------------------------------------------------------

_SCALAR = {np.dtype(np.float32): ("float32", "f"), np.dtype(np.float64): ("float64", "d"),
           np.dtype(np.int32): ("int32", "i"), np.dtype(np.int64): ("int64", "l"),
           np.dtype(np.uint32): ("uint32", "u"), np.dtype(np.uint64): ("uint64", "ul"),
           np.dtype(np.int16): ("int16", "h"), np.dtype(np.uint16): ("uint16", "uh"),
           np.dtype(np.int8): ("int8", "b"), np.dtype(np.uint8): ("uint8", "ub"),
           np.dtype(np.bool_): ("bool", "")}


def _warp_dtype_of(a: np.ndarray) -> str:
    """The warp element type a warp array of this numpy layout has: (N, 2..4) floats are
    vecNf/vecNd, (N, 3, 3) float32 mat33f, otherwise the scalar type."""
    name, suffix = _SCALAR.get(a.dtype, ("float32", "f"))
    if a.ndim == 2 and a.shape[1] in (2, 3, 4) and a.dtype.kind == "f":
        return f"<class 'warp.types.vec{a.shape[1]}{suffix}'>"
    if a.ndim == 3 and a.shape[1:] == (3, 3) and a.dtype.kind == "f":
        return f"<class 'warp.types.mat33{suffix}'>"
    return f"<class 'warp.types.{name}'>"


def _ndarray_record(a: np.ndarray, fmt: str) -> dict:
    if fmt == "json":
        return {"__type__": "numpy.ndarray", "dtype": str(a.dtype), "shape": list(a.shape),
                "data": json.dumps(a.tolist())}
    a = np.ascontiguousarray(a)
    return {"__type__": "numpy.ndarray", "dtype": a.dtype.str, "shape": list(a.shape), "order": "C",
            "binary_data": a.tobytes(order="C")}


def _ndarray_from_record(rec: dict, fmt: str) -> np.ndarray:
    dtype = np.dtype(rec["dtype"])
    shape = tuple(rec["shape"])
    if "binary_data" in rec:
        return np.frombuffer(rec["binary_data"], dtype=dtype).reshape(shape, order=rec.get("order", "C")).copy()
    data = rec["data"]
    if isinstance(data, str):
        data = json.loads(data)
    return np.array(data, dtype=dtype).reshape(shape)


class _Cache:
    """Index per array identity (1, 2, ...): the first occurrence carries the payload."""

    def __init__(self):
        self.index: dict[int, int] = {}
        self.values: dict[int, Any] = {}

    def lookup_or_add(self, key: int) -> tuple[int, bool]:
        if key in self.index:
            return self.index[key], False
        k = len(self.index) + 1
        self.index[key] = k
        return k, True


def _serialize(obj, fmt: str, cache: _Cache, visiting: set):
    from .hashgrid import HashGrid

    if isinstance(obj, torch.Tensor):
        k, first = cache.lookup_or_add(2 << 60 | (obj.data_ptr() if obj.numel() else id(obj)))
        a = obj.detach().cpu().numpy()
        if not first:
            return {"__type__": "warp.array_ref", "__dtype__": _warp_dtype_of(a), "cache_index": k}
        return {"__type__": "warp.array", "__dtype__": _warp_dtype_of(a), "cache_index": k,
                "data": _ndarray_record(a, fmt)}
    if isinstance(obj, HashGrid):
        return {"__type__": "warp.HashGrid", "data": None}
    if isinstance(obj, torch.device):
        return {"__type__": "wp.Device", "data": None}
    if callable(obj):
        return {"__type__": "callable", "data": None}
    if id(obj) in visiting:
        return "<circular_reference>"
    if isinstance(obj, (str, int, float, bool)) or obj is None:
        value = obj if type(obj) in (str, int, float, bool) or obj is None else (
            int(obj) if isinstance(obj, int) else float(obj) if isinstance(obj, float) else str(obj))
        return {"__type__": type(obj).__name__, "value": value}
    if isinstance(obj, np.number):
        return {"__type__": f"numpy.{type(obj).__name__}", "value": obj.item()}
    if isinstance(obj, np.ndarray):
        return _ndarray_record(obj, fmt)
    visiting.add(id(obj))
    try:
        if isinstance(obj, Mapping):
            return {"__type__": type(obj).__name__,
                    "items": {str(k): _serialize(v, fmt, cache, visiting) for k, v in obj.items()}}
        if isinstance(obj, Iterable) and not isinstance(obj, (bytes, bytearray)):
            return {"__type__": type(obj).__name__, "items": [_serialize(v, fmt, cache, visiting) for v in obj]}
        if hasattr(obj, "__dict__"):
            return {"__type__": type(obj).__name__, "__module__": type(obj).__module__,
                    "attributes": {a: _serialize(v, fmt, cache, visiting) for a, v in vars(obj).items()}}
        raise ValueError(f"Cannot serialize object of type {type(obj)}")
    finally:
        visiting.discard(id(obj))


class _Ref:
    def __init__(self, k: int):
        self.k = k


def _deserialize(data, fmt: str, cache: _Cache, device: torch.device):
    if not isinstance(data, dict) or "__type__" not in data:
        return data
    t = data["__type__"]
    if t == "warp.array":
        a = _ndarray_from_record(data["data"], fmt)
        v = torch.from_numpy(a).to(device)
        if data.get("cache_index") is not None:
            cache.values[int(data["cache_index"])] = v
        return v
    if t == "warp.array_ref":
        k = int(data["cache_index"])
        return cache.values[k] if k in cache.values else _Ref(k)
    if t in ("warp.HashGrid", "warp.Mesh", "callable"):
        return None
    if t in ("str", "int", "float", "bool", "NoneType"):
        return data["value"]
    if t == "numpy.ndarray":
        return _ndarray_from_record(data, fmt)
    if t.startswith("numpy."):
        return getattr(np, t.split(".")[-1])(data["value"])
    if t == "dict":
        return {k: _deserialize(v, fmt, cache, device) for k, v in data["items"].items()}
    if t in ("list", "tuple", "set"):
        items = [_deserialize(v, fmt, cache, device) for v in data["items"]]
        return tuple(items) if t == "tuple" else (set(items) if t == "set" else items)
    if "attributes" in data:
        return {a: _deserialize(v, fmt, cache, device) for a, v in data["attributes"].items()}
    return data["value"] if "value" in data else data


def _resolve(obj, cache: _Cache):
    if isinstance(obj, _Ref):
        return cache.values[obj.k]
    if isinstance(obj, dict):
        return {k: _resolve(v, cache) for k, v in obj.items()}
    if isinstance(obj, list):
        return [_resolve(v, cache) for v in obj]
    if isinstance(obj, tuple):
        return tuple(_resolve(v, cache) for v in obj)
    return obj


def _format_of(path: str) -> str:
    ext = os.path.splitext(path.lower())[1]
    if ext == ".json":
        return "json"
    if ext == ".bin":
        return "cbor"
    raise ValueError(f"Unsupported file extension '{ext}'. Supported extensions: .json, .bin")


class ViewerFile:
    """Records Model and State snapshots to `.json` or `.bin` and plays them back.

    Same constructor and methods as Newton's ViewerFile (viewer_file.py:1021-1432); the
    rendering calls of the viewer interface are accepted and ignored."""

    def __init__(self, output_path: str, auto_save: bool = True, save_interval: int = 100,
                 max_history_size: int | None = None, device: str | torch.device | None = None):
        if not output_path:
            raise ValueError("output_path must be a non-empty path")
        self.output_path = Path(output_path)
        self.auto_save = bool(auto_save)
        self.save_interval = int(save_interval)
        self.history = [] if max_history_size is None else RingBuffer(max_history_size)
        self.raw_model = None
        self.deserialized_model: dict | None = None
        self.model = None
        self.time = 0.0
        self._frame_count = 0
        self._model_recorded = False
        self._device = torch.device(device) if device is not None else (
            torch.device("cuda:0") if torch.cuda.is_available() else torch.device("cpu"))

    # -- viewer interface ------------------------------------------------------------------
    def set_model(self, model, max_worlds: int | None = None) -> None:
        self.model = model
        if model is not None and not self._model_recorded:
            self.record_model(model)
            self._model_recorded = True

    def begin_frame(self, time: float) -> None:
        self.time = float(time)

    def end_frame(self) -> None:
        pass

    def is_running(self) -> bool:
        return True

    def is_paused(self) -> bool:
        return False

    def should_step(self) -> bool:
        return True

    def log_state(self, state) -> None:
        self.record(state)
        self._frame_count += 1
        if self.auto_save and self._frame_count % self.save_interval == 0:
            self._save_recording()

    def log_mesh(self, *args, **kwargs) -> None:
        pass

    def log_instances(self, *args, **kwargs) -> None:
        pass

    def log_lines(self, *args, **kwargs) -> None:
        pass

    def log_points(self, *args, **kwargs) -> None:
        pass

    def log_array(self, *args, **kwargs) -> None:
        pass

    def log_scalar(self, *args, **kwargs) -> None:
        pass

    def apply_forces(self, state) -> None:
        pass

    def close(self) -> None:
        if self._frame_count > 0:
            self._save_recording()

    # -- recording ---------------------------------------------------------------------------
    def record(self, state) -> None:
        """Snapshot (clone) every array attribute of `state` (viewer_file.py:1123-1133), and
        the arrays of its custom-attribute namespaces under "namespace:name" keys (the
        solver state a resume needs, such as the SPH pressure sph:pressure)."""
        frame = {}
        for k, v in vars(state).items():
            if isinstance(v, torch.Tensor):
                frame[k] = v.clone()
            elif isinstance(v, AttributeNamespace):
                for a, t in vars(v).items():
                    if isinstance(t, torch.Tensor):
                        frame[f"{k}:{a}"] = t.clone()
        self.history.append(frame)

    def playback(self, state, frame_id: int) -> None:
        """Restore a snapshot into `state` (viewer_file.py:1135-1149): the arrays `state` has,
        namespaced ones included."""
        if not 0 <= frame_id < len(self.history):
            print(f"Warning: frame_id {frame_id} is out of bounds. Playback skipped.")
            return
        for name, value in self.history[frame_id].items():
            if ":" in name:
                ns, a = name.split(":", 1)
                holder = getattr(state, ns, None)
                if holder is not None and hasattr(holder, a):
                    setattr(holder, a, value)
            elif hasattr(state, name):
                setattr(state, name, value)

    def record_model(self, model) -> None:
        self.raw_model = model

    def playback_model(self, model) -> None:
        """Copy the recorded model's attributes that `model` also has (arrays moved to the
        model's device; read-only properties and devices are left alone)."""
        if not self.deserialized_model:
            print("Warning: No model data to playback.")
            return
        dev = getattr(model, "device", self._device)
        for name, value in self.deserialized_model.items():
            if name.startswith("_") or name == "device" or not hasattr(model, name):
                continue
            if isinstance(getattr(type(model), name, None), property):
                continue
            if isinstance(value, dict):
                continue
            if isinstance(value, torch.Tensor):
                value = value.to(dev)
            try:
                setattr(model, name, value)
            except (AttributeError, TypeError):
                pass

    def save_recording(self, file_path: str | None = None, verbose: bool = False) -> None:
        self._save_recording(file_path, verbose=verbose)

    def _save_recording(self, file_path: str | None = None, verbose: bool = False) -> None:
        path = str(file_path if file_path is not None else self.output_path)
        if "." not in os.path.basename(path):
            path += ".json"
        fmt = _format_of(path)
        states = self.history.to_list() if isinstance(self.history, RingBuffer) else list(self.history)
        tree = _serialize({"model": self.raw_model, "states": states}, fmt, _Cache(), set())
        if fmt == "json":
            with open(path, "w") as f:
                json.dump(tree, f, indent=2)
        else:
            with open(path, "wb") as f:
                f.write(cbor.dumps(tree))
        if verbose:
            print(f"Recording saved to {path} ({self._frame_count} frames)")

    def load_recording(self, file_path: str | None = None, verbose: bool = False) -> None:
        path = str(file_path if file_path is not None else self.output_path)
        if "." not in os.path.basename(path) and os.path.exists(path + ".json"):
            path += ".json"
        fmt = _format_of(path)
        if fmt == "json":
            with open(path) as f:
                tree = json.load(f)
        else:
            with open(path, "rb") as f:
                tree = cbor.loads(f.read())
        cache = _Cache()
        raw = _resolve(_deserialize(tree, fmt, cache, self._device), cache)
        self.deserialized_model = raw["model"]
        if isinstance(self.history, RingBuffer):
            self.history.from_list(raw["states"])
        else:
            self.history = list(raw["states"])
        self._frame_count = len(self.history)
        if verbose:
            print(f"Loaded recording with {self._frame_count} frames from {path}")

    def get_frame_count(self) -> int:
        return self._frame_count

    def has_model(self) -> bool:
        return self.deserialized_model is not None

    def load_model(self, model) -> None:
        self.playback_model(model)

    def load_state(self, state, frame_id: int) -> None:
        self.playback(state, frame_id)
