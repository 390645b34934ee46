"""Granular pile with XPBD particle-particle contacts (the reference's own grid consumers on
the HIP hash grid): a block of grains falls onto a floor of static grains.

Command: python -m newton_sph_dev_amd.examples xpbd_granular [--benchmark]
"""

from __future__ import annotations

import numpy as np

import newton_sph_dev_amd.examples as ex
from newton_sph_dev_amd import ModelBuilder, SolverXPBD


class Example:
    def __init__(self, viewer, args):
        self.viewer = viewer
        self.fps = float(args.fps)
        self.sim_substeps = int(args.substeps)
        self.frame_dt = 1.0 / self.fps
        self.sim_dt = self.frame_dt / self.sim_substeps
        self.sim_time = 0.0
        r = self.radius = 0.01
        b = ModelBuilder()
        n = int(args.grains_per_side)
        b.add_particle_grid(pos=(0.1, 0.1, 0.05), rot=(0.0, 0.0, 0.0, 1.0), vel=(0.0, 0.0, 0.0), dim_x=n, dim_y=n,
                            dim_z=n, cell_x=2.1 * r, cell_y=2.1 * r, cell_z=2.1 * r, mass=0.01, jitter=0.2 * r,
                            radius_mean=r)
        # a static floor of grains (zero mass)
        g = np.stack(np.meshgrid(np.arange(2 * n + 10), np.arange(2 * n + 10), indexing="ij"), -1).reshape(-1, 2)
        floor = np.zeros((g.shape[0], 3))
        floor[:, :2] = g * (2.0 * r)
        b.add_particles(floor, np.zeros_like(floor), np.zeros(floor.shape[0]), np.full(floor.shape[0], r))
        device = args.device or "cuda:0"
        self.model = b.finalize(device)
        self.model.particle_mu = 0.5
        self.solver = SolverXPBD(self.model, iterations=int(args.iterations))
        self.state_0, self.state_1 = self.model.state(), self.model.state()
        self.viewer.set_model(self.model)

    def step(self):
        for _ in range(self.sim_substeps):
            self.state_0.clear_forces()
            self.solver.step(self.state_0, self.state_1, None, None, self.sim_dt)
            self.state_0, self.state_1 = self.state_1, self.state_0
        self.sim_time += self.frame_dt

    def render(self):
        self.viewer.begin_frame(self.sim_time)
        self.viewer.log_state(self.state_0)
        self.viewer.end_frame()

    def test_final(self):
        r = self.radius
        ex.test_particle_state(self.state_0, "grains rest on or above the floor", lambda q, qd: q[2] > -r)


def create_parser():
    p = ex.create_parser()
    p.add_argument("--fps", type=float, default=60.0)
    p.add_argument("--substeps", type=int, default=8)
    p.add_argument("--iterations", type=int, default=4)
    p.add_argument("--grains-per-side", type=int, default=12)
    return p


if __name__ == "__main__":
    viewer, args = ex.init(create_parser())
    ex.run(Example(viewer, args), args)
