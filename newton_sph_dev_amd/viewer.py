"""Headless viewers: ViewerNull (the interface of newton/_src/viewer/viewer_null.py:31-165)
and ViewerFile (recorder.py).

ViewerNull renders nothing.  It ends a run after `num_frames` frames, and in benchmark mode
measures frames per second over a window that opens once `benchmark_start_frame` frames have
ended (the warm-up) and closes at the last frame or after `benchmark_timeout` seconds.  Each
window edge waits for the device, so queued GPU work is inside the window it belongs to.
"""

from __future__ import annotations

import time
from dataclasses import dataclass

import torch

from .recorder import ViewerFile

__all__ = ["ViewerFile", "ViewerNull"]


@dataclass
class _TimingWindow:
    """Frames ended and seconds elapsed since the window opened (None until it opens)."""

    opened_at: float | None = None
    open_frame: int = 0
    frames: int = 0
    seconds: float = 0.0

    def open(self, frame: int) -> None:
        self.opened_at, self.open_frame = time.perf_counter(), frame

    def extend(self, frame: int) -> None:
        self.frames = frame - self.open_frame
        self.seconds = time.perf_counter() - self.opened_at


def _device_idle() -> None:
    if torch.cuda.is_available():
        torch.cuda.synchronize()


class ViewerNull:
    """A viewer that draws nothing; see the module docstring for the benchmark window."""

    def __init__(self, num_frames: int = 1000, benchmark: bool = False, benchmark_timeout: float | None = None,
                 benchmark_start_frame: int = 3):
        self.num_frames = int(num_frames)
        self.benchmark_timeout = benchmark_timeout
        # a timeout implies benchmarking, as in the reference's null viewer
        self.benchmark = bool(benchmark) or benchmark_timeout is not None
        self.benchmark_start_frame = int(benchmark_start_frame)
        self.frame_count = 0
        self.model = None
        self.time = 0.0
        self._window = _TimingWindow()

    # ---- frame loop ---------------------------------------------------------------------
    def set_model(self, model, max_worlds: int | None = None) -> None:
        self.model = model

    def begin_frame(self, time: float) -> None:
        self.time = float(time)

    def end_frame(self) -> None:
        self.frame_count += 1
        if not self.benchmark:
            return
        w = self._window
        if w.opened_at is None:
            if self.frame_count == self.benchmark_start_frame:
                _device_idle()
                w.open(self.frame_count)
        else:
            _device_idle()
            w.extend(self.frame_count)

    def is_running(self) -> bool:
        if self.frame_count >= self.num_frames:
            return False
        w = self._window
        timed_out = (self.benchmark_timeout is not None and w.opened_at is not None
                     and w.seconds >= self.benchmark_timeout)
        return not timed_out

    def is_paused(self) -> bool:
        return False

    def should_step(self) -> bool:
        return True

    def benchmark_result(self) -> dict | None:
        """{"fps", "frames", "elapsed"} of the timed window, None without benchmarking."""
        if not self.benchmark:
            return None
        w = self._window
        if w.frames == 0 or w.seconds == 0.0:
            return {"fps": 0.0, "frames": 0, "elapsed": 0.0}
        return {"fps": w.frames / w.seconds, "frames": w.frames, "elapsed": w.seconds}

    # ---- logging: nothing is drawn or stored --------------------------------------------
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
