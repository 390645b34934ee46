"""Host-side logic without a GPU: the particle generator (builder.py:8037-8137), world
starts (builder.py:9545-9604), finalize (builder.py:9815-9855), parameter resolution, scene
sizes, and the C ABI library's exported symbols."""

from __future__ import annotations

import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

from newton_sph_dev_amd import ModelBuilder, ParticleFlags, scenes
from newton_sph_dev_amd.sim import _world_starts
from newton_sph_dev_amd.solvers import SPHParams

ROOT = Path(__file__).resolve().parents[1]


def test_grid_order_is_y_slowest_x_then_z_fastest():
    b = ModelBuilder()
    b.add_particle_grid(pos=(0, 0, 0), rot=(0, 0, 0, 1), vel=(0, 0, 0), dim_x=2, dim_y=3, dim_z=2,
                        cell_x=1.0, cell_y=10.0, cell_z=100.0, mass=1.0, jitter=0.0)
    q = b.particle_q
    expected = []
    for iy in range(3):
        for ix in range(2):
            for iz in range(2):
                expected.append((ix * 1.0, iy * 10.0, iz * 100.0))
    assert np.array_equal(q, np.array(expected))


def test_grid_jitter_bounds_and_seed_depends_on_particle_count():
    s = 0.0125
    kw = dict(pos=(0, 0, 0), rot=(0, 0, 0, 1), vel=(0, 0, 0), dim_x=4, dim_y=5, dim_z=6,
              cell_x=s, cell_y=s, cell_z=s, mass=1.0, jitter=0.1 * s)
    a = ModelBuilder()
    a.add_particle_grid(**kw)
    lattice = ModelBuilder()
    lattice.add_particle_grid(**{**kw, "jitter": 0.0})
    off = a.particle_q - lattice.particle_q
    assert np.all(np.abs(off) <= 0.05 * s) and np.abs(off).max() > 0.04 * s
    # a second grid after n particles is jittered like a first grid after n dummy particles
    b = ModelBuilder()
    b.add_particle_grid(**kw)
    b.add_particle_grid(**kw)
    c = ModelBuilder()
    c.add_particles(np.zeros((120, 3)), np.zeros((120, 3)), np.ones(120))
    c.add_particle_grid(**kw)
    assert np.array_equal(b.particle_q[120:], c.particle_q[120:])
    assert not np.array_equal(b.particle_q[120:], b.particle_q[:120])


def test_grid_rotation_is_the_float32_quaternion_matrix():
    """builder.py:8090-8091 rotates with Warp's float32 quat_to_matrix and translates in float64.
    Pinned here: a 90-degree turn about z maps the grid's x axis onto +y and y onto -x, each
    entry within one float32 ulp of the exact matrix, with the float32 roundings visible (the
    cosine of the float32 half-angle quaternion is not exactly 0); the points are the float64
    product of the lattice with that float32 matrix; the identity quaternion gives the lattice
    unchanged.  Whether Warp's host build contracts the products, which would move an entry by
    one ulp, cannot be checked without Warp (parity unpinned at that last bit)."""
    from newton_sph_dev_amd.sim import quat_to_matrix_f32

    h = np.float32(np.sqrt(0.5))
    M = quat_to_matrix_f32((0.0, 0.0, h, h))
    assert M.dtype == np.float32
    exact = np.array([[0.0, -1.0, 0.0], [1.0, 0.0, 0.0], [0.0, 0.0, 1.0]])
    ulp = np.spacing(np.float32(1.0))
    assert np.all(np.abs(M.astype(np.float64) - exact) <= ulp)
    assert M[0, 0] != 0.0  # float32 half angle: cos 90 deg comes out as a rounding residue
    b = ModelBuilder()
    b.add_particle_grid(pos=(1.0, 2.0, 3.0), rot=(0.0, 0.0, h, h), vel=(0, 0, 0), dim_x=3, dim_y=2, dim_z=2,
                        cell_x=0.5, cell_y=0.25, cell_z=0.125, mass=1.0, jitter=0.0)
    lat = np.stack(np.meshgrid(np.arange(3) * 0.5, np.arange(2) * 0.25, np.arange(2) * 0.125)).reshape(3, -1).T
    want = lat @ M.astype(np.float64).T + np.array([1.0, 2.0, 3.0])
    assert np.array_equal(b.particle_q, want)
    assert np.allclose(b.particle_q[:, 1] - 2.0, lat[:, 0], atol=1e-6)  # x axis -> +y
    assert np.array_equal(quat_to_matrix_f32((0.0, 0.0, 0.0, 1.0)), np.eye(3, dtype=np.float32))


def test_finalize_casts_and_inverse_mass():
    b = ModelBuilder()
    b.add_particles([(0.1, 0.2, 0.3), (1.0 / 3.0, 0.0, 0.0)], [(0, 0, 0), (0, 0, 0)], [2.0, 0.0])
    m = b.finalize("cpu")
    assert m.particle_q.dtype.is_floating_point and m.particle_q.numpy().dtype == np.float32
    assert m.particle_q.numpy()[1, 0] == np.float32(1.0 / 3.0)
    assert np.array_equal(m.particle_inv_mass.numpy(), np.array([0.5, 0.0], np.float32))
    assert m.particle_flags.numpy().tolist() == [int(ParticleFlags.ACTIVE)] * 2
    assert m.gravity.numpy().tolist() == [[0.0, 0.0, np.float32(-9.81)]]


def test_world_starts_layout():
    # worlds [-1, -1, 0, 0, 1, -1]: front globals 2, world 0 at 2, world 1 at 4, tail 5, total 6
    w = np.array([-1, -1, 0, 0, 1, -1], np.int32)
    assert _world_starts(w, 2).tolist() == [2, 4, 5, 6]
    b = ModelBuilder()
    with pytest.raises(RuntimeError):
        b.end_world()
    b.begin_world(gravity=(0.0, 0.0, -1.0))
    with pytest.raises(RuntimeError):
        b.begin_world()
    b.add_particle((0, 0, 0), (0, 0, 0), 1.0)
    b.end_world()
    m = b.finalize("cpu")
    assert m.world_count == 1 and m.particle_world.numpy().tolist() == [0]
    assert m.gravity.numpy().tolist() == [[0.0, 0.0, -1.0]]


def test_dam_break_sizes():
    assert scenes.DAM_BREAK_32K.n_fluid == 32768
    assert scenes.DAM_BREAK_1M.n_fluid == 1048576
    bp = scenes.box_boundary_points(0.0125, (1.0, 1.0, 1.0), 2)
    assert bp.shape[0] == 84 * 84 * 82 - 80 * 80 * 80
    bp = scenes.box_boundary_points(0.004, (2.0, 2.0, 1.0), 2)
    assert bp.shape[0] == 504 * 504 * 252 - 500 * 500 * 250
    # boundary points never overlap the interior lattice
    s = 0.0125
    idx = np.round(scenes.box_boundary_points(s, (0.25, 0.25, 0.25), 2) / s - 0.5).astype(int)
    inside = np.all((idx >= 0) & (idx < 20), axis=1)
    assert not inside.any()


def test_params_resolution():
    b = scenes.build_dam_break(scenes.small_dam_break(4, 4, 4))
    m = b.finalize("cpu")
    p = SPHParams.resolve(m, 0.025)
    assert p.n_worlds == 1 and p.boundary_volume[0] == pytest.approx(0.0125 ** 3)
    q = m.particle_q.numpy()
    assert np.allclose(p.domain_lo[0], q.min(axis=0) - 0.05)
    with pytest.raises(ValueError):
        SPHParams.resolve(m, [0.1, 0.2])
    with pytest.raises(ValueError):
        SPHParams.resolve(m, 0.025, warm_start="bogus")


def test_solver_refuses_cpu_model():
    from newton_sph_dev_amd import SolverSPH

    m = scenes.build_dam_break(scenes.small_dam_break(4, 4, 4)).finalize("cpu")
    with pytest.raises(RuntimeError):
        SolverSPH(m, 0.025)


def _header_functions() -> list[str]:
    text = (ROOT / "include" / "nsph.h").read_text()
    return re.findall(r"^(?:int|const char\*)\s+(nsph_\w+)\(", text, flags=re.M)


def test_abi_exports_every_declared_symbol():
    from newton_sph_dev_amd import _native as N

    lib = N.load()
    names = _header_functions()
    assert len(names) >= 25
    for name in names:
        assert hasattr(lib, name), name
    assert lib.nsph_abi_version() == 3
    # every declared function has a ctypes signature in the binding
    for name in names:
        assert name in N.SIGNATURES or name in ("nsph_last_error", "nsph_kernel_name"), name


def test_device_header_compiles_as_c_and_hip():
    """include/nsph_device.h: its C part (the view struct) for a host C compiler, its device
    query functions for hipcc."""
    import shutil
    import subprocess

    hdr = ROOT / "include" / "nsph_device.h"
    subprocess.run(["gcc", "-fsyntax-only", "-Wall", "-Werror", "-x", "c", str(hdr)], check=True)
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not Path(hipcc).exists():
        pytest.skip("hipcc not available")
    subprocess.run([hipcc, "--offload-arch=gfx950", "-fsyntax-only", "-Wall", "-Werror",
                    "-Wno-unused-command-line-argument", "-x", "hip",
                    str(hdr)], check=True)


def test_abi_errors_without_device():
    """Without a visible GPU the context cannot be created; the error is a status code plus
    a message, never a crash."""
    import torch

    from newton_sph_dev_amd import _native as N

    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    lib = N.load()
    out = C.c_void_p()
    st = lib.nsph_ctx_create(0, None, C.byref(out))
    assert st != 0
    assert len(N.last_error()) > 0
    assert lib.nsph_solver_step(None, None, None, C.c_float(1e-4), 0, None) == 1


def test_bench_accounts_every_kernel_class():
    """bench.py's algorithmic-byte table names every kernel class the library times."""
    import bench
    from newton_sph_dev_amd import _native as N

    lib = N.load()
    names = [lib.nsph_kernel_name(k).decode() for k in range(lib.nsph_kernel_count())]
    table = bench.algorithmic_bytes(28.0)
    assert sorted(table) == sorted(names)
