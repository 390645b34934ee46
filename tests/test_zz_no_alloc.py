"""nsph_soft_contacts_static, the allocation-free variant of nsph_soft_contacts.

Called through the C ABI on two calls in a row (the second with fewer particles, so that the
static workspace still holds the first call's larger contents) and compared with
nsph_soft_contacts given its own workspace: contact counts, starts and every contact record
are equal (the two share their kernels, so equality is exact).  Sizes above the workspace's
maximum are refused before any GPU work.
"""

from __future__ import annotations

import ctypes as C

import numpy as np
import pytest
import torch

from newton_sph_dev_amd import ModelBuilder, _native as N

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def _scene(n: int, seed: int):
    rng = np.random.default_rng(seed)
    b = ModelBuilder()
    x = rng.uniform(-1.0, 1.0, (n, 3))
    b.add_particles(x.tolist(), [(0.0, 0.0, 0.0)] * n, [1.0] * n, rng.uniform(0.01, 0.05, n).tolist())
    b.add_shape_sphere(-1, ((0.2, -0.1, 0.0), (0.0, 0.0, 0.0, 1.0)), radius=0.5)
    q = np.array([0.1, 0.3, -0.2, 0.9])
    q /= np.linalg.norm(q)
    b.add_shape_box(-1, ((-0.4, 0.3, 0.2), tuple(q)), hx=0.3, hy=0.4, hz=0.2)
    b.add_ground_plane(height=-0.7)
    return b.finalize(DEV)


def _call(fn: str, m, work=None):
    n = m.particle_count
    cap = 3 * n
    out = {"start": torch.full((n + 1,), -7, dtype=torch.int32, device=DEV),
           "count": torch.full((1,), -7, dtype=torch.int32, device=DEV),
           "particle": torch.full((cap,), -7, dtype=torch.int32, device=DEV),
           "shape": torch.full((cap,), -7, dtype=torch.int32, device=DEV),
           "pos": torch.full((cap, 3), -7.0, device=DEV), "vel": torch.full((cap, 3), -7.0, device=DEV),
           "normal": torch.full((cap, 3), -7.0, device=DEV)}
    ctx = N.context(0)
    N.set_stream(ctx, torch.cuda.current_stream().cuda_stream)
    args = [ctx, n, N.ptr(m.particle_q), N.ptr(m.particle_radius), N.ptr(m.particle_flags), N.ptr(m.particle_world),
            m.shape_count, N.ptr(m.shape_transform), N.ptr(m.shape_type), N.ptr(m.shape_scale), N.ptr(m.shape_world),
            N.ptr(m.shape_flags), 0.01, cap]
    if work is not None:
        args.append(N.ptr(work))
    args += [N.ptr(out["start"]), N.ptr(out["count"]), N.ptr(out["particle"]), N.ptr(out["shape"]),
             N.ptr(out["pos"]), N.ptr(out["vel"]), N.ptr(out["normal"])]
    N.call(fn, *args)
    torch.cuda.synchronize()
    return {k: v.cpu().numpy() for k, v in out.items()}


def test_static_workspace_equals_caller_workspace_on_two_calls():
    for n, seed in ((200_000, 1), (30_000, 2)):  # the second call is smaller
        print(f"soft contacts, {n} particles: static workspace vs caller workspace", flush=True)
        m = _scene(n, seed)
        work = torch.zeros(n + (n + 255) // 256, dtype=torch.int32, device=DEV)
        ref = _call("nsph_soft_contacts", m, work)
        got = _call("nsph_soft_contacts_static", m)
        total = int(ref["count"][0])
        assert total > n // 4
        assert int(got["count"][0]) == total
        assert np.array_equal(got["start"], ref["start"])
        for k in ("particle", "shape", "pos", "vel", "normal"):
            assert np.array_equal(got[k][:total], ref[k][:total]), k
            assert np.array_equal(got[k][total:], ref[k][total:]), k  # slots past the total untouched


def test_static_workspace_refuses_oversized_calls():
    lib = N.load()
    ctx = N.context(0)
    too_many = 4_000_001
    start = torch.zeros(1, dtype=torch.int32, device=DEV)
    count = torch.zeros(1, dtype=torch.int32, device=DEV)
    # refused on the host, before any array is read or written
    st = lib.nsph_soft_contacts_static(ctx, too_many, None, None, None, None, 1, None, None, None, None, None,
                                       C.c_float(0.01), 0, N.ptr(start), N.ptr(count), None, None, None, None, None)
    assert st == N.NSPH_ERR_CAPACITY
    assert "static workspace" in N.last_error()
