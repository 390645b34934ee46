"""Physics of the GPU step on a hydrostatic column (tests/hydrostatic.py): 600 steps, the
trajectory bit-exact against the FP32 restatement, and the settled column at rest with
gravity balanced by the SPH force, density rho0 and pressure rising at rho0 g with depth."""

from __future__ import annotations

import numpy as np
import pytest
import torch

import hydrostatic as HC
from newton_sph_dev_amd import SolverSPH
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def test_hydrostatic_column():
    print(f"hydrostatic column {HC.NX}x{HC.NY}x{HC.NZ}: {HC.STEPS} steps of {HC.DT} s", flush=True)
    m = HC.column("cuda:0")
    s = SolverSPH(m, 2 * HC.S, **HC.PARAMS)
    a, b = m.state(), m.state()
    iters = []
    for _ in range(HC.STEPS):
        s.step(a, b, None, None, HC.DT)
        iters.append(s.last_report["max_iterations"])
        a, b = b, a
    torch.cuda.synchronize()
    q, qd, f = (t.cpu().numpy() for t in (a.particle_q, a.particle_qd, s.sph_force))
    fields = {k: v.cpu().numpy() for k, v in s.fields().items()}
    pb = O.problem_from_model(m, s)
    mt = HC.metrics(q, qd, fields["pressure"], fields["density"], f, pb.mass, pb.inv_mass > 0)
    print(f"  iterations mean {np.mean(iters):.1f} max {max(iters)}; {mt}", flush=True)
    HC.check(mt)
    # the whole trajectory, step by step, equals the FP32 restatement
    qo, qdo, pp, xb = pb.q0, pb.qd0, None, None
    for _ in range(HC.STEPS):
        r = O.step(pb, qo, qdo, pprev=pp, dt=HC.DT, precision="f32", xb=xb)
        qo, qdo, pp, xb = r["q"], r["qd"], r["pprev"], r["xb"]
    assert np.array_equal(qo, q) and np.array_equal(qdo, qd)
