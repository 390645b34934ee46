"""Viewers for headless runs: ViewerNull (newton/_src/viewer/viewer_null.py) and ViewerFile.

ViewerNull counts frames, stops after `num_frames`, and in benchmark mode times the frames
after `benchmark_start_frame` warm-up frames (synchronising the device around the timed
window), as the reference's null viewer does.
"""

from __future__ import annotations

import time as _time

import torch

from .recorder import ViewerFile

__all__ = ["ViewerFile", "ViewerNull"]


def _sync() -> None:
    if torch.cuda.is_available():
        torch.cuda.synchronize()


class ViewerNull:
    def __init__(self, num_frames: int = 1000, benchmark: bool = False, benchmark_timeout: float | None = None,
                 benchmark_start_frame: int = 3):
        self.num_frames = int(num_frames)
        self.frame_count = 0
        self.benchmark = bool(benchmark) or benchmark_timeout is not None
        self.benchmark_timeout = benchmark_timeout
        self.benchmark_start_frame = int(benchmark_start_frame)
        self._bench_start_time: float | None = None
        self._bench_frames = 0
        self._bench_elapsed = 0.0
        self.model = None
        self.time = 0.0

    def set_model(self, model, max_worlds: int | None = None) -> None:
        self.model = model

    def begin_frame(self, time: float) -> None:
        self.time = float(time)

    def end_frame(self) -> None:
        self.frame_count += 1
        if self.benchmark:
            if self.frame_count == self.benchmark_start_frame:
                _sync()
                self._bench_start_time = _time.perf_counter()
            elif self._bench_start_time is not None:
                _sync()
                self._bench_frames = self.frame_count - self.benchmark_start_frame
                self._bench_elapsed = _time.perf_counter() - self._bench_start_time

    def is_running(self) -> bool:
        if self.frame_count >= self.num_frames:
            return False
        if (self.benchmark_timeout is not None and self._bench_start_time is not None
                and self._bench_elapsed >= self.benchmark_timeout):
            return False
        return True

    def is_paused(self) -> bool:
        return False

    def should_step(self) -> bool:
        return True

    def benchmark_result(self) -> dict | None:
        if not self.benchmark:
            return None
        if self._bench_frames == 0 or self._bench_elapsed == 0.0:
            return {"fps": 0.0, "frames": 0, "elapsed": 0.0}
        return {"fps": self._bench_frames / self._bench_elapsed, "frames": self._bench_frames,
                "elapsed": self._bench_elapsed}

    def log_state(self, state) -> None:
        pass

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
        pass
