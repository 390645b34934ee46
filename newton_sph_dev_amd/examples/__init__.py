"""Example harness with the reference's command line and run loop.

Mirrors newton/examples/__init__.py: `create_parser` (the common options; `--viewer`
offers the headless "null" and "file" viewers), `init(parser)` -> (viewer, args) with
`--benchmark [SECONDS]` forcing the null viewer, `run(example, args)` (step / render until
the viewer stops, `test_post_step` / `test_final` hooks under `--test`, the "Benchmark: X
FPS (N frames in T s)" line, and the finiteness check of the example's states and model),
`test_particle_state`, `get_examples` and `main` (`python -m newton_sph_dev_amd.examples
<name> [options]`, `--list`).
"""

from __future__ import annotations

import os
import sys

import torch

from ..viewer import ViewerFile, ViewerNull


def create_parser():
    import argparse

    p = argparse.ArgumentParser(add_help=True, formatter_class=argparse.ArgumentDefaultsHelpFormatter)
    p.add_argument("--device", type=str, default=None, help="Device for the model (default cuda:0).")
    p.add_argument("--viewer", type=str, default="null", choices=["null", "file"],
                   help="Viewer: null (no output) or file (ViewerFile recording at --output-path).")
    p.add_argument("--output-path", type=str, default="output.bin", help="Recording path for the file viewer.")
    p.add_argument("--num-frames", type=int, default=100, help="Total number of frames.")
    p.add_argument("--headless", action=argparse.BooleanOptionalAction, default=True,
                   help="Accepted for compatibility; every viewer here is headless.")
    p.add_argument("--test", action=argparse.BooleanOptionalAction, default=False,
                   help="Whether to run the example in test mode.")
    p.add_argument("--quiet", action=argparse.BooleanOptionalAction, default=False,
                   help="Suppress informational output.")
    p.add_argument("--benchmark", type=int, default=False, nargs="?", const=None, metavar="SECONDS",
                   help="Benchmark mode: FPS after a warm-up; with SECONDS, stop after that long or "
                        "--num-frames, whichever comes first.")
    p.add_argument("--realtime", action="store_true", default=False, help="Accepted for compatibility.")
    return p


def default_args(parser=None):
    return (parser or create_parser()).parse_args([])


def init(parser=None):
    """Parse arguments and create the viewer. Returns (viewer, args)."""
    if parser is None:
        parser = create_parser()
        args = parser.parse_known_args()[0]
    else:
        args = parser.parse_args()
    if args.device:
        torch.cuda.set_device(torch.device(args.device))
    if args.benchmark is not False:
        args.viewer = "null"
    if args.viewer == "null":
        viewer = ViewerNull(num_frames=args.num_frames, benchmark=args.benchmark is not False,
                            benchmark_timeout=args.benchmark or None)
    elif args.viewer == "file":
        viewer = _FileViewer(args.output_path, args.num_frames)
    else:
        raise ValueError(f"Invalid viewer: {args.viewer}")
    return viewer, args


class _FileViewer(ViewerFile):
    """ViewerFile that stops after num_frames (the reference's file viewer runs until the
    example loop ends)."""

    def __init__(self, path: str, num_frames: int):
        super().__init__(path, auto_save=False)
        self.num_frames = int(num_frames)
        self.frame_count = 0

    def end_frame(self) -> None:
        self.frame_count += 1

    def is_running(self) -> bool:
        return self.frame_count < self.num_frames


def _format_fps(fps: float) -> str:
    if fps >= 10:
        return f"{fps:.1f}"
    if fps >= 1:
        return f"{fps:.2f}"
    return f"{fps:.3f}"


def find_nan_members(obj) -> list[str]:
    """Names of tensor attributes of obj holding a NaN."""
    out = []
    for name, v in vars(obj).items():
        if isinstance(v, torch.Tensor) and v.is_floating_point() and bool(torch.isnan(v).any()):
            out.append(name)
    return out


def test_particle_state(state, test_name: str, test_fn, indices=None) -> None:
    """Raise ValueError unless test_fn(q, qd) holds for every (selected) particle."""
    q = state.particle_q.detach().cpu().numpy()
    qd = state.particle_qd.detach().cpu().numpy()
    idx = range(q.shape[0]) if indices is None else indices
    bad = [i for i in idx if not test_fn(q[i], qd[i])]
    if bad:
        raise ValueError(f'Test "{test_name}" failed for {len(bad)} particles, e.g. {bad[:10]}')


def run(example, args) -> None:
    viewer = example.viewer
    perform_test = args is not None and args.test
    test_post_step = perform_test and hasattr(example, "test_post_step")
    test_final = perform_test and hasattr(example, "test_final")
    while viewer.is_running():
        if viewer.should_step():
            example.step()
        if test_post_step:
            example.test_post_step()
        example.render()
    if perform_test:
        if test_final:
            example.test_final()
        elif not test_post_step:
            raise NotImplementedError("Example does not have a test_final or test_post_step method")
    if isinstance(viewer, _FileViewer):
        viewer.save_recording()
    viewer.close()
    if hasattr(viewer, "benchmark_result"):
        result = viewer.benchmark_result()
        if result is not None:
            print(f"Benchmark: {_format_fps(result['fps'])} FPS ({result['frames']} frames in "
                  f"{result['elapsed']:.2f}s)")
    if perform_test:
        for name in ("state_0", "state_1", "model"):
            if hasattr(example, name):
                nan = find_nan_members(getattr(example, name))
                if nan:
                    raise ValueError(f"NaN members found in {name}: {nan}")


def get_source_directory() -> str:
    return os.path.dirname(os.path.abspath(__file__))


def get_examples() -> dict[str, str]:
    out = {}
    for f in sorted(os.listdir(get_source_directory())):
        if f.startswith("example_") and f.endswith(".py"):
            out[f[8:-3]] = f"newton_sph_dev_amd.examples.{f[:-3]}"
    return out


def main() -> None:
    import runpy

    examples = get_examples()
    if len(sys.argv) >= 2 and sys.argv[1] in ("-h", "--help"):
        print("Usage: python -m newton_sph_dev_amd.examples <example_name> [options]")
        print("       python -m newton_sph_dev_amd.examples --list")
        sys.exit(0)
    if len(sys.argv) >= 2 and sys.argv[1] == "--list":
        print("Available examples:")
        for name in examples:
            print(f"  {name}")
        sys.exit(0)
    name = sys.argv[1] if len(sys.argv) >= 2 else "sph_dam_break"
    if name not in examples:
        print(f"Error: Unknown example '{name}'\n")
        sys.exit(1)
    sys.argv = [sys.argv[0], *sys.argv[2:]]
    runpy.run_module(examples[name], run_name="__main__")

