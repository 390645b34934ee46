"""Example harness (newton/examples/__init__.py behaviour): parser defaults, the null
viewer's frame count and benchmark window, the run loop's test hooks and NaN check, and the
examples themselves on the GPU in test mode."""

from __future__ import annotations

import subprocess
import sys

import numpy as np
import pytest
import torch

import newton_sph_dev_amd.examples as ex
from newton_sph_dev_amd import ModelBuilder
from newton_sph_dev_amd.viewer import ViewerNull


def test_parser_defaults_and_benchmark_flag():
    args = ex.default_args()
    assert args.viewer == "null" and args.num_frames == 100 and args.test is False and args.benchmark is False
    p = ex.create_parser()
    assert p.parse_args(["--benchmark"]).benchmark is None
    assert p.parse_args(["--benchmark", "5"]).benchmark == 5


def test_null_viewer_counts_frames_and_times_after_warmup():
    v = ViewerNull(num_frames=6, benchmark=True)
    frames = 0
    while v.is_running():
        v.begin_frame(0.0)
        v.end_frame()
        frames += 1
    assert frames == 6
    r = v.benchmark_result()
    assert r["frames"] == 3 and r["elapsed"] > 0.0 and r["fps"] > 0.0
    assert ViewerNull(num_frames=2).benchmark_result() is None


class _Dummy:
    def __init__(self, viewer, nan=False):
        self.viewer = viewer
        self.steps = 0
        self.finals = 0
        m = ModelBuilder().finalize("cpu")
        self.model = m
        self.state_0 = type("S", (), {})()
        self.state_0.particle_q = torch.full((3, 3), float("nan") if nan else 0.0)

    def step(self):
        self.steps += 1

    def render(self):
        self.viewer.begin_frame(0.0)
        self.viewer.end_frame()

    def test_final(self):
        self.finals += 1


def test_run_loop_hooks_and_nan_check(capsys):
    args = ex.create_parser().parse_args(["--num-frames", "5", "--test", "--benchmark"])
    args.viewer = "null"
    d = _Dummy(ViewerNull(num_frames=5, benchmark=True))
    ex.run(d, args)
    assert d.steps == 5 and d.finals == 1
    assert "Benchmark:" in capsys.readouterr().out
    with pytest.raises(ValueError, match="NaN members"):
        ex.run(_Dummy(ViewerNull(num_frames=2), nan=True), args)


def test_particle_state_check():
    s = type("S", (), {})()
    s.particle_q = torch.tensor([[0.0, 0.0, 1.0], [0.0, 0.0, -1.0]])
    s.particle_qd = torch.zeros(2, 3)
    ex.test_particle_state(s, "z above -2", lambda q, qd: q[2] > -2)
    with pytest.raises(ValueError, match="above zero"):
        ex.test_particle_state(s, "above zero", lambda q, qd: q[2] > 0)


def test_list_examples():
    out = subprocess.run([sys.executable, "-m", "newton_sph_dev_amd.examples", "--list"], capture_output=True,
                         text=True, check=True).stdout
    assert "sph_dam_break" in out and "xpbd_granular" in out


@pytest.mark.gpu
@pytest.mark.parametrize("name,extra", [("sph_dam_break", ["--graph"]), ("sph_dam_break", []),
                                        ("xpbd_granular", ["--grains-per-side", "6"])])
def test_example_runs_in_test_mode(name, extra, tmp_path):
    import importlib

    print(f"example {name} {extra}", flush=True)
    mod = importlib.import_module(f"newton_sph_dev_amd.examples.example_{name}")
    args = mod.create_parser().parse_args(["--num-frames", "6", "--test", "--viewer", "file", "--output-path",
                                           str(tmp_path / "rec.bin"), *extra])
    viewer = ex._FileViewer(args.output_path, args.num_frames)
    example = mod.Example(viewer, args)
    ex.run(example, args)
    assert (tmp_path / "rec.bin").exists()
    assert np.isfinite(example.state_0.particle_q.cpu().numpy()).all()
