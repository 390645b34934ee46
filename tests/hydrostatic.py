"""Hydrostatic column shared by the oracle and GPU tests (not a test module).

A 8 x 8 x 24 fluid lattice at spacing 0.02 m in a box of static boundary particles, settled
for 600 steps of 1 ms with the IISPH solve converged to 1e-4 mean compression and viscosity
0.05 (to damp the start-up sloshing).  Checked on the central interior (two spacings from the
side walls, three below the surface and above the floor): gravity is balanced by the SPH
force, the density is rho0, the pressure rises with depth at rho0 g, the fluid is at rest.
"""

from __future__ import annotations

import numpy as np

from newton_sph_dev_amd import ModelBuilder, ParticleFlags, scenes

S, NX, NY, NZ = 0.02, 8, 8, 24
DT, STEPS = 1e-3, 600
PARAMS = dict(warm_start="prev", viscosity=0.05, tolerance=1e-4, max_iters=500)
G, RHO0 = 9.81, 1000.0


def column(device: str):
    b = ModelBuilder()
    b.add_particle_grid(pos=(0.5 * S, 0.5 * S, 0.5 * S), rot=(0.0, 0.0, 0.0, 1.0), vel=(0.0, 0.0, 0.0),
                        dim_x=NX, dim_y=NY, dim_z=NZ, cell_x=S, cell_y=S, cell_z=S, mass=RHO0 * S ** 3,
                        jitter=0.0, radius_mean=0.5 * S)
    bp = scenes.box_boundary_points(S, (NX * S, NY * S, (NZ + 6) * S), 2)
    b.add_particles(bp, np.zeros_like(bp), np.zeros(bp.shape[0]), np.full(bp.shape[0], 0.5 * S),
                    np.full(bp.shape[0], int(ParticleFlags.ACTIVE)))
    return b.finalize(device)


def metrics(q, qd, p, rho, f, mass, fluid):
    """Central-interior figures of a settled column (all inputs in particle order)."""
    q, qd, p, rho, f, mass = (np.asarray(a, np.float64) for a in (q, qd, p, rho, f, mass))
    x, y, z = q[fluid, 0], q[fluid, 1], q[fluid, 2]
    depth = z.max() + 0.5 * S - z
    sel = ((x > 2 * S) & (x < (NX - 2) * S) & (y > 2 * S) & (y < (NY - 2) * S) &
           (depth > 3 * S) & (depth < depth.max() - 3 * S))
    az = f[fluid, 2][sel] / mass[fluid][sel]
    slope = np.polyfit(depth[sel], p[fluid][sel], 1)[0]
    return {
        "n": int(sel.sum()),
        "az_rel": float(abs(az.mean() - G) / G),
        "rho_rel": float(np.abs(rho[fluid][sel] / RHO0 - 1.0).max()),
        "slope_rel": float(slope / (RHO0 * G)),
        "speed": float(np.linalg.norm(qd[fluid][sel], axis=1).max()),
    }


def check(mt):
    assert mt["n"] > 250
    assert mt["az_rel"] < 0.03, mt          # SPH force balances gravity
    assert mt["rho_rel"] < 1e-3, mt         # incompressible to 0.1 % per particle
    assert 0.85 < mt["slope_rel"] < 1.15, mt  # dp/d(depth) = rho0 g
    assert mt["speed"] < 0.1, mt             # at rest
