"""Committed fixtures (tests/golden, written by tests/golden/make_fixtures.py) as regression
pins: generator bytes per scene (SHA-256), the FP64 oracle's one-step dam break, its
brute-force neighbour lists, and the initial neighbour-count histograms.  The GPU test
checks the HIP step against the fixture's FP64 values at the 1e-5 bar."""

from __future__ import annotations

import hashlib
import json
import os

import numpy as np
import pytest

from newton_sph_dev_amd import scenes
from newton_sph_dev_amd.solvers import SPHParams
from oracle import oracle as O

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _sha(q):
    return hashlib.sha256(np.ascontiguousarray(q, dtype=np.float32).tobytes()).hexdigest()


@pytest.mark.parametrize("name", ["dam_break_12x10x14", "dam_break_32k", "dam_break_1m"])
def test_generator_bytes(name):
    ref = json.load(open(os.path.join(HERE, "generator_sha256.json")))[name]
    cfg = {"dam_break_12x10x14": scenes.small_dam_break(12, 10, 14), "dam_break_32k": scenes.DAM_BREAK_32K,
           "dam_break_1m": scenes.DAM_BREAK_1M}[name]
    m = scenes.build_dam_break(cfg).finalize("cpu")
    assert m.particle_count == ref["particles"]
    assert _sha(m.particle_q.numpy()) == ref["sha256"]


def test_batched_generator_bytes():
    ref = json.load(open(os.path.join(HERE, "generator_sha256.json")))
    for cfg in (scenes.small_batch(12), scenes.small_batch(20, seed=11)):
        b, _ = scenes.build_batched(cfg)
        m = b.finalize("cpu")
        assert _sha(m.particle_q.numpy()) == ref[cfg.name + f"_seed{cfg.seed}"]["sha256"]


def _small():
    cfg = scenes.small_dam_break(12, 10, 14)
    m = scenes.build_dam_break(cfg).finalize("cpu")
    return cfg, m, O.problem_from_model(m, SPHParams.resolve(m, cfg.support, warm_start="zero"))


def test_oracle_reproduces_the_one_step_fixture():
    fx = np.load(os.path.join(HERE, "dambreak_small.npz"))
    cfg, m, pb = _small()
    assert np.array_equal(pb.q0, fx["q0"])
    r = O.step(pb, pb.q0, pb.qd0, precision="f64")
    fid = fx["fluid_ids"]
    assert np.array_equal(r["iters"], fx["iters"])
    assert np.allclose(r["rho"][fid], fx["rho"], rtol=1e-12, atol=0)
    assert np.allclose(r["f"][fid], fx["f"], rtol=1e-10, atol=1e-12)
    assert np.array_equal(r["q"], fx["q1"])
    H2 = np.float32(cfg.support * cfg.support)
    bc, bi = O.neighbors_brute(pb.q0[fid], None, H2, cap=96)
    indices = np.concatenate([np.sort(fid[bi[k, : bc[k]]]) for k in range(fid.size)])
    assert np.array_equal(np.concatenate([[0], np.cumsum(bc)]), fx["nbr_indptr"])
    assert np.array_equal(indices, fx["nbr_indices"])


def test_initial_neighbour_histograms():
    ref = json.load(open(os.path.join(HERE, "nbr_hist.json")))
    for cfg in (scenes.small_dam_break(12, 10, 14), scenes.DAM_BREAK_32K):
        m = scenes.build_dam_break(cfg).finalize("cpu")
        pb = O.problem_from_model(m, SPHParams.resolve(m, cfg.support, warm_start="zero"))
        cnt, _, _ = O.solver_neighbors(pb, pb.q0)
        fl = ((pb.flags & 1) != 0) & (pb.inv_mass > 0)
        h = np.bincount(cnt[fl])
        assert {str(k): int(v) for k, v in enumerate(h) if v} == ref[cfg.name]


@pytest.mark.gpu
def test_gpu_step_against_the_fixture():
    import torch

    from newton_sph_dev_amd import SolverSPH

    fx = np.load(os.path.join(HERE, "dambreak_small.npz"))
    cfg = scenes.small_dam_break(12, 10, 14)
    m = scenes.build_dam_break(cfg).finalize("cuda:0")
    s = SolverSPH(m, cfg.support, warm_start="zero")
    a, b = m.state(), m.state()
    s.step(a, b, None, None, scenes.DT)
    torch.cuda.synchronize()
    fid = fx["fluid_ids"]
    f = s.fields()
    rho = f["density"].cpu().numpy()[fid].astype(np.float64)
    assert np.max(np.abs(rho - fx["rho"]) / fx["rho"]) <= 1e-5
    force = s.sph_force.cpu().numpy()[fid].astype(np.float64)
    assert np.abs(force - fx["f"]).max() / np.abs(fx["f"]).max() <= 1e-5
    counts, idx = (t.cpu().numpy() for t in s.neighbors())
    got = np.concatenate([np.sort(idx[: counts[i], i]) for i in fid])
    assert np.array_equal(got, fx["nbr_indices"])
