"""Learned warm start for the IISPH pressure solve.

A per-particle MLP 6 -> 64 -> 64 -> 1 (BF16 weights, FP32 accumulation, hand-written gfx950
MFMA kernels in libnsph: forward, gradient, Adam) predicts the correction
(p - 0.5 p_prev) / p_unit to the IISPH previous-pressure warm start from
  density error, velocity divergence, neighbour count, previous pressure, |v| dt / H and
  distance to the nearest boundary particle
(exact definitions in csrc/sph.hip, mlp_features).  The solver starts the Jacobi iterations
from p0 = max(0.5 p_prev + clamp(y, +-8) p_unit, 0) for particles with at least
GATE_NEIGHBOURS fluid + boundary neighbours (0.5 p_prev elsewhere: at the free surface and
in splashes p_unit is large and a network output would become a pressure spike), and always
finishes with the plain Jacobi iterations to the same tolerance, so the plain solver stays
the fallback and what decides convergence.

Training data are solver snapshots: after each step of seeded rollouts of the dam-break
generator with random block extents, the features and the target (converged pressure
correction / p_unit) of every fluid particle the network is applied to are exported on the
GPU.
"""

from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np
import torch

from . import _native as N
from . import scenes
from .sim import ModelBuilder, ParticleFlags
from .solvers import SolverSPH

N_PARAM = 4673
N_FEAT = 6
X_STRIDE = 8
GATE_NEIGHBOURS = 24.0  # MLP_GATE_NEIGHBOURS of the warm-start kernel


def init_params(seed: int = 0) -> np.ndarray:
    """He-style initial weights, zero biases (flat layout of nsph_mlp_set_params)."""
    rng = np.random.default_rng(seed)
    p = np.zeros(N_PARAM, dtype=np.float32)
    o1, o2, o3 = 0, 6 * 64 + 64, 6 * 64 + 64 + 64 * 64 + 64
    p[o1:o1 + 6 * 64] = rng.standard_normal(6 * 64) * np.sqrt(2.0 / 6)
    p[o2:o2 + 64 * 64] = rng.standard_normal(64 * 64) * np.sqrt(2.0 / 64)
    p[o3:o3 + 64] = rng.standard_normal(64) * np.sqrt(1.0 / 64)
    return p


class WarmStartMLP:
    """Handle on an nsph_mlp (parameters live on the GPU)."""

    def __init__(self, device: str | torch.device = "cuda:0", seed: int = 0,
                 params: np.ndarray | None = None):
        self.device = torch.device(device)
        if self.device.type != "cuda":
            raise RuntimeError("WarmStartMLP runs on a ROCm GPU only")
        self._ctx = N.context(self.device.index or 0)
        N.set_stream(self._ctx, torch.cuda.current_stream(self.device).cuda_stream)
        h = C.c_void_p()
        N.call("nsph_mlp_create", self._ctx, C.byref(h))
        self._handle = h
        self.set_params(init_params(seed) if params is None else params)

    def _stream(self):
        N.set_stream(self._ctx, torch.cuda.current_stream(self.device).cuda_stream)

    def set_params(self, p: np.ndarray) -> None:
        self._stream()
        p = np.ascontiguousarray(p, dtype=np.float32).reshape(-1)
        if p.size != N_PARAM:
            raise ValueError(f"expected {N_PARAM} parameters, got {p.size}")
        N.call("nsph_mlp_set_params", self._handle, p.ctypes.data_as(C.POINTER(C.c_float)))

    def params(self) -> np.ndarray:
        self._stream()
        p = np.zeros(N_PARAM, dtype=np.float32)
        N.call("nsph_mlp_get_params", self._handle, p.ctypes.data_as(C.POINTER(C.c_float)))
        return p

    def forward(self, X: torch.Tensor) -> torch.Tensor:
        """y for feature rows X (n, 8) float32 (columns 6, 7 ignored)."""
        self._stream()
        X = X.contiguous()
        y = torch.empty(X.shape[0], dtype=torch.float32, device=self.device)
        N.call("nsph_mlp_forward", self._handle, N.ptr(X), int(X.shape[0]), N.ptr(y))
        return y

    def grad(self, X: torch.Tensor, T: torch.Tensor) -> tuple[float, torch.Tensor]:
        """Mean squared error of y against T and its gradient (flat, float32)."""
        self._stream()
        X, T = X.contiguous(), T.contiguous()
        g = torch.empty(N_PARAM + 1, dtype=torch.float32, device=self.device)
        loss = C.c_float()
        N.call("nsph_mlp_grad", self._handle, N.ptr(X), N.ptr(T), int(X.shape[0]), C.byref(loss), N.ptr(g))
        return loss.value, g[:N_PARAM]

    def adam(self, lr: float = 1e-3, beta1: float = 0.9, beta2: float = 0.999, eps: float = 1e-8) -> None:
        self._stream()
        N.call("nsph_mlp_adam", self._handle, C.c_float(lr), C.c_float(beta1), C.c_float(beta2), C.c_float(eps))

    def train(self, X: torch.Tensor, T: torch.Tensor, epochs: int = 10, batch: int = 65536,
              lr: float = 2e-3, lr_min: float = 5e-5, seed: int = 0, log=None) -> list[float]:
        """Adam over shuffled mini-batches with a cosine learning-rate decay from lr to
        lr_min; returns the mean loss per epoch."""
        import math

        gen = torch.Generator(device="cpu").manual_seed(seed)
        n = X.shape[0]
        hist = []
        total = epochs * ((n + batch - 1) // batch)
        it = 0
        for ep in range(epochs):
            perm = torch.randperm(n, generator=gen).to(self.device)
            tot, cnt = 0.0, 0
            for b0 in range(0, n, batch):
                idx = perm[b0:b0 + batch]
                loss, _ = self.grad(X[idx], T[idx])
                rate = lr_min + 0.5 * (lr - lr_min) * (1.0 + math.cos(math.pi * it / max(total - 1, 1)))
                self.adam(rate)
                it += 1
                tot += loss * idx.numel()
                cnt += idx.numel()
            hist.append(tot / max(cnt, 1))
            if log:
                log(f"epoch {ep}: loss {hist[-1]:.5f}")
        return hist

    def save(self, path: str) -> None:
        np.save(path, self.params())

    @staticmethod
    def load(path: str, device="cuda:0") -> "WarmStartMLP":
        return WarmStartMLP(device, params=np.load(path, allow_pickle=False))

    def __del__(self):
        h = getattr(self, "_handle", None)
        if h is not None and N._lib is not None:
            try:
                N._lib.nsph_mlp_destroy(h)
            except Exception:
                pass
            self._handle = None


def features(solver: SolverSPH) -> tuple[torch.Tensor, torch.Tensor]:
    """(X, T) of the solver's last step: features (Nf, 8) and target (p - 0.5 p_prev) / p_unit."""
    dev = solver.model.device
    nf = solver.fluid_count
    X = torch.empty((nf, X_STRIDE), dtype=torch.float32, device=dev)
    T = torch.empty(nf, dtype=torch.float32, device=dev)
    N.call("nsph_solver_mlp_features", solver._handle, N.ptr(X), N.ptr(T))
    return X, T


@dataclass
class SnapshotConfig:
    worlds: int = 64
    steps: int = 120
    spacing: float = 0.0125
    extent: tuple[float, float] = (0.1, 0.3)
    seed: int = 2024
    dts: tuple = (scenes.DT,)
    tolerance: float = 1e-4


def build_training_worlds(cfg: SnapshotConfig) -> tuple[ModelBuilder, np.ndarray]:
    """Dam-break worlds of the config-1 generator (spacing, jitter 0.1 s, H = 2 s, two-layer
    walls) with random block extents, one world each."""
    rng = np.random.default_rng(cfg.seed)
    b = ModelBuilder()
    s = cfg.spacing
    for _ in range(cfg.worlds):
        ext = rng.uniform(cfg.extent[0], cfg.extent[1], size=3)
        dims = [max(2, int(round(e / s))) for e in ext]
        box = tuple(float(round(2.5 * dims[a] * s / s) * s) if a == 0 else float(round(1.5 * dims[a] * s / s) * s)
                    for a in range(3))
        b.begin_world()
        b.add_particle_grid(pos=(0.5 * s,) * 3, rot=(0.0, 0.0, 0.0, 1.0), vel=(0.0, 0.0, 0.0),
                            dim_x=dims[0], dim_y=dims[1], dim_z=dims[2], cell_x=s, cell_y=s, cell_z=s,
                            mass=scenes.REST_DENSITY * s ** 3, jitter=0.1 * s, radius_mean=0.5 * s)
        bp = scenes.box_boundary_points(s, box, 2)
        b.add_particles(bp, np.zeros_like(bp), np.zeros(bp.shape[0]), np.full(bp.shape[0], 0.5 * s),
                        np.full(bp.shape[0], int(ParticleFlags.ACTIVE)))
        b.end_world()
    return b, np.full(cfg.worlds, 2.0 * s)


def collect_snapshots(cfg: SnapshotConfig, device="cuda:0", keep_every: int = 1,
                      max_samples: int | None = None):
    """Roll the plain solver (previous-pressure warm start, tight tolerance) over batched
    training worlds; after each step keep every fluid particle's (features, target).
    Returns X (n, 8), T (n,) on the device and the number of world snapshots."""
    b, H = build_training_worlds(cfg)
    m = b.finalize(device)
    Xs, Ts = [], []
    snaps = 0
    steps = max(1, cfg.steps // len(cfg.dts))
    for dt in cfg.dts:
        solver = SolverSPH(m, H, warm_start="prev", tolerance=cfg.tolerance, max_iters=300)
        a, c = m.state(), m.state()
        for k in range(steps):
            solver.step(a, c, None, None, dt)
            a, c = c, a
            if k % keep_every == 0:
                X, T = features(solver)
                # the warm start applies the network only in the bulk (see GATE_NEIGHBOURS)
                bulk = X[:, 2] * 32.0 >= GATE_NEIGHBOURS
                Xs.append(X[bulk])
                Ts.append(T[bulk])
                snaps += cfg.worlds
    X = torch.cat(Xs)
    T = torch.cat(Ts)
    if max_samples is not None and X.shape[0] > max_samples:
        sel = torch.randperm(X.shape[0], generator=torch.Generator().manual_seed(cfg.seed))[:max_samples].to(X.device)
        X, T = X[sel], T[sel]
    return X, T, snaps


def evaluate(model_builder_fn, H, mlp: WarmStartMLP | None, steps: int = 40, warmup: int = 5,
             dt: float = scenes.DT, tolerance: float = 1e-3, min_iters: int = 2,
             modes=("zero", "prev", "mlp"), device="cuda:0") -> dict:
    """Iterations and wall-clock per step to the same tolerance for each warm start, from the
    same initial state (each mode rolls its own trajectory)."""
    import time

    out = {}
    for mode in modes:
        if mode == "mlp" and mlp is None:
            continue
        m = model_builder_fn().finalize(device)
        s = SolverSPH(m, H, warm_start="zero" if mode == "mlp" else mode, tolerance=tolerance,
                      min_iters=min_iters, max_iters=500)
        if mode == "mlp":
            s.set_mlp(mlp)
        a, b = m.state(), m.state()
        for _ in range(warmup):
            s.step(a, b, None, None, dt)
            a, b = b, a
        torch.cuda.synchronize()
        its, errs = [], []
        t0 = time.perf_counter()
        for _ in range(steps):
            s.step(a, b, None, None, dt)
            its.append(s.last_report["max_iterations"])
            errs.append(s.last_report["max_density_error"])
            a, b = b, a
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
        out[mode] = {"mean_iterations": float(np.mean(its)), "ms_per_step": el * 1e3 / steps,
                     "max_density_error": float(np.max(errs)),
                     "unconverged_steps": 0}
    return out


def main(argv=None):
    import argparse
    import json
    import sys
    import time

    ap = argparse.ArgumentParser(description="train / evaluate the learned warm start")
    ap.add_argument("--worlds", type=int, default=80)
    ap.add_argument("--steps", type=int, default=250)
    ap.add_argument("--epochs", type=int, default=6)
    ap.add_argument("--samples", type=int, default=16_000_000)
    ap.add_argument("--batch", type=int, default=65536)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--out", default="")
    ap.add_argument("--load", default="")
    ap.add_argument("--eval-steps", type=int, default=30)
    ap.add_argument("--eval-dt", type=float, nargs="*", default=[scenes.DT, 5e-4, 1e-3])
    ap.add_argument("--train-dt", type=float, nargs="*", default=[scenes.DT, 5e-4, 1e-3])
    ap.add_argument("--min-iters", type=int, nargs="*", default=[2, 1])
    args = ap.parse_args(argv)
    log = lambda msg: print(f"[learned] {msg}", file=sys.stderr, flush=True)  # noqa: E731
    report = {}
    if args.load:
        mlp = WarmStartMLP.load(args.load)
    else:
        cfg = SnapshotConfig(worlds=args.worlds, steps=args.steps, dts=tuple(args.train_dt))
        t0 = time.perf_counter()
        X, T, snaps = collect_snapshots(cfg, max_samples=args.samples)
        report["snapshots"] = snaps
        report["samples"] = int(X.shape[0])
        report["collect_s"] = time.perf_counter() - t0
        log(f"{snaps} snapshots, {X.shape[0]} samples in {report['collect_s']:.1f} s")
        mlp = WarmStartMLP(seed=0)
        t0 = time.perf_counter()
        hist = mlp.train(X, T, epochs=args.epochs, batch=args.batch, lr=args.lr, log=log)
        torch.cuda.synchronize()
        report["train_s"] = time.perf_counter() - t0
        report["loss"] = hist
        report["target_var"] = float(T.var().item())
        if args.out:
            mlp.save(args.out)
    cfg1m = scenes.DAM_BREAK_1M
    for dt in args.eval_dt:
        for mi in args.min_iters:
            key = f"dam_break_1m dt={dt:g} min_iters={mi}"
            log(f"evaluate {key}")
            try:
                report[key] = evaluate(lambda: scenes.build_dam_break(cfg1m), cfg1m.support, mlp,
                                       steps=args.eval_steps, dt=dt, min_iters=mi)
            except N.NsphError as e:
                report[key] = {"error": str(e)}
    print(json.dumps(report))


if __name__ == "__main__":
    main()
