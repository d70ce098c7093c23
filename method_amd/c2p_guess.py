"""Training data and training driver for the learned cons2prim starting guess.

The network (8 conserved inputs -> log W, log p/E) is trained on the GPU by
the engine's own kernels (``engine.Mlp``). Training states combine seeded
synthetic primitives over the ranges

    rho     log-uniform in [1e-4, 10]
    p/rho   log-uniform in [1e-3, 1e2]
    W       uniform in [1, 10], isotropic direction
    |B|^2/p log-uniform in [1e-2, 1e2], isotropic direction

with primitive states taken from the Kelvin-Helmholtz and Orszag-Tang
workloads. Conserved variables come from the engine's primsToAll kernel
(SRMHD::primsToAll, srmhd.cc:573-673), so the network sees exactly what the
cons2prim kernel will be given.
"""
from __future__ import annotations

import numpy as np

from . import engine as eng
from . import initdata as idt


def _unit_vectors(rng, n):
    v = rng.normal(size=(n, 3))
    return v / np.linalg.norm(v, axis=1, keepdims=True)


def synthetic_prims(n: int, seed: int) -> np.ndarray:
    """n x 8 SRMHD primitives (rho, vx, vy, vz, p, Bx, By, Bz) over the training ranges."""
    rng = np.random.default_rng(seed)
    rho = 10.0 ** rng.uniform(-4.0, 1.0, n)
    p = rho * 10.0 ** rng.uniform(-3.0, 2.0, n)
    W = rng.uniform(1.0, 10.0, n)
    v = np.sqrt(1.0 - 1.0 / W ** 2)[:, None] * _unit_vectors(rng, n)
    B = np.sqrt(p * 10.0 ** rng.uniform(-2.0, 2.0, n))[:, None] * _unit_vectors(rng, n)
    return np.concatenate([rho[:, None], v, p[:, None], B], axis=1)


def workload_prims(n_kh: int, n_ot: int, seed: int) -> np.ndarray:
    """Primitive states of the workloads: a 3-D KH state after a few steps and
    the Orszag-Tang state (its rho, v, p, B)."""
    rng = np.random.default_rng(seed + 1)
    out = []
    if n_kh:
        g = idt.kh3d_grid(64)
        with eng.Engine(eng.Config(nx=64, ny=64, nz=64, xmin=-0.5, xmax=0.5, ymin=-1.0, ymax=1.0, zmin=0.0,
                                   zmax=1.0, weno_order=5)) as e:
            e.set_prims(idt.kh_prims(g, seed=seed))
            for _ in range(5):
                e.update_time()
            _, p, _ = e.download()
        cells = p[:, 4:-4, 4:-4, 4:-4].reshape(8, -1).T
        out.append(cells[rng.integers(0, cells.shape[0], n_kh)])
    if n_ot:
        go = idt.ot2d_grid(256)
        po = idt.ot_prims_resistive(go)[:8, 4:-4, 4:-4, 0].reshape(8, -1).T
        out.append(po[rng.integers(0, po.shape[0], n_ot)])
    return np.concatenate(out, axis=0) if out else np.zeros((0, 8))


def cons_from_prims(prims8: np.ndarray, gamma: float = 4.0 / 3.0):
    """SRMHD cons (n x 9) of n primitive states via the engine's primsToAll."""
    n = prims8.shape[0]
    ny = 64
    nx = max(4, -(-n // ny))
    grid = np.zeros((8, nx * ny))
    grid[:, :n] = prims8.T
    grid[0, n:] = 1.0
    grid[4, n:] = 1.0
    padded = np.zeros((8, nx + 8, ny + 8, 1))
    padded[:, 4:-4, 4:-4, 0] = grid.reshape(8, nx, ny)
    with eng.Engine(eng.Config(nx=nx, ny=ny, nz=0, xmin=0.0, xmax=1.0, ymin=0.0, ymax=1.0, gamma=gamma)) as e:
        e.set_prims(padded)
        c, _, _ = e.download()
    return c[:, 4:-4, 4:-4, 0].reshape(9, -1)[:, :n].T.copy()


def training_set(n_synth: int, n_kh: int, n_ot: int, seed: int):
    """(cons8, (W, p)) arrays of the training states, shuffled."""
    prims = np.concatenate([synthetic_prims(n_synth, seed), workload_prims(n_kh, n_ot, seed)], axis=0)
    rng = np.random.default_rng(seed + 2)
    prims = prims[rng.permutation(prims.shape[0])]
    cons = cons_from_prims(prims)
    W = 1.0 / np.sqrt(1.0 - np.sum(prims[:, 1:4] ** 2, axis=1))
    return cons[:, :8].copy(), np.stack([W, prims[:, 4]], axis=1)


def train_guess(n_synth: int = 1 << 18, n_kh: int = 1 << 18, n_ot: int = 1 << 17, steps: int = 12000,
                batch: int = 8192, lr: float = 5e-3, seed: int = 7):
    """Trains the network on the GPU with a step-decayed Adam learning rate;
    returns (params, output scales, final loss). By default 40 % of the
    samples are Kelvin-Helmholtz states, 20 % Orszag-Tang states and 40 %
    synthetic states over the full ranges."""
    cons8, wp = training_set(n_synth, n_kh, n_ot, seed)
    m = eng.Mlp(seed=seed)
    m.set_data(cons8, wp)
    loss = m.train(steps // 2, batch, lr)
    loss = m.train(steps // 4, batch, lr * 0.3)
    loss = m.train(steps - steps // 2 - steps // 4, batch, lr * 0.05)
    p, n = m.params()
    m.close()
    return p, n, loss
