"""SPH dam break: a water column collapsing in an open box (IISPH on the HIP kernels).

Command: python -m newton_sph_dev_amd.examples sph_dam_break [--scale 1m] [--benchmark]
The frame (1/fps) is split into substeps of --dt; when the substep count is even the two
ping-pong directions are captured once as HIP graphs and replayed every frame.
"""

from __future__ import annotations

import numpy as np

import newton_sph_dev_amd.examples as ex
from newton_sph_dev_amd import SolverSPH, scenes

SCALES = {"small": scenes.small_dam_break(16, 14, 20), "32k": scenes.DAM_BREAK_32K, "1m": scenes.DAM_BREAK_1M}


class Example:
    def __init__(self, viewer, args):
        self.viewer = viewer
        self.cfg = SCALES[args.scale]
        self.sim_dt = float(args.dt)
        self.fps = float(args.fps)
        self.sim_substeps = max(1, int(round(1.0 / (self.fps * self.sim_dt))))
        self.frame_dt = self.sim_substeps * self.sim_dt
        self.sim_time = 0.0
        device = args.device or "cuda:0"
        self.model = scenes.build_dam_break(self.cfg).finalize(device)
        self.solver = SolverSPH(self.model, self.cfg.support, warm_start=args.warm_start)
        self.state_0, self.state_1 = self.model.state(), self.model.state()
        self.viewer.set_model(self.model)
        self.graphs = None
        if args.graph and self.sim_substeps % 2 == 0:
            self.graphs = (self.solver.capture(self.state_0, self.state_1, self.sim_dt, iter_slots=args.graph_slots),
                           self.solver.capture(self.state_1, self.state_0, self.sim_dt, iter_slots=args.graph_slots))

    def simulate(self):
        for _ in range(self.sim_substeps):
            self.solver.step(self.state_0, self.state_1, None, None, self.sim_dt)
            self.state_0, self.state_1 = self.state_1, self.state_0

    def step(self):
        if self.graphs is not None:
            for k in range(self.sim_substeps):
                self.solver.replay(self.graphs[k % 2])
        else:
            self.simulate()
        self.sim_time += self.frame_dt

    def render(self):
        self.viewer.begin_frame(self.sim_time)
        self.viewer.log_state(self.state_0)
        self.viewer.end_frame()

    def test_final(self):
        box = np.asarray(self.cfg.box, dtype=np.float64)
        h = self.cfg.support
        ex.test_particle_state(self.state_0, "all particles stay in the box",
                               lambda q, qd: bool(np.all(q > -2.0 * h) and np.all(q[:2] < box[:2] + 2.0 * h)))
        rep = self.solver.last_report or {}
        if rep.get("max_density_error", 0.0) > self.solver.tolerance:
            raise ValueError(f"density error {rep['max_density_error']:.2e} above the tolerance")


def create_parser():
    p = ex.create_parser()
    p.add_argument("--scale", choices=sorted(SCALES), default="small", help="Dam-break size.")
    p.add_argument("--dt", type=float, default=scenes.DT, help="Substep [s].")
    p.add_argument("--fps", type=float, default=500.0, help="Frames per simulated second.")
    p.add_argument("--warm-start", choices=["zero", "prev"], default="prev", help="Pressure warm start.")
    p.add_argument("--graph", action="store_true", help="Replay captured HIP graphs.")
    p.add_argument("--graph-slots", type=int, default=3, help="Jacobi iterations captured per step.")
    return p


if __name__ == "__main__":
    viewer, args = ex.init(create_parser())
    ex.run(Example(viewer, args), args)
