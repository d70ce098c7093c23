"""Seeded initial primitive states for the workloads the engine is measured on.

The reference builds initial data inside ``InitialFunc`` subclasses that write
``d->prims`` over the whole ghost-padded grid (Project/CPU/Src/initFunc.cc:10-40);
``Simulation::set`` then derives cons/aux and applies the boundary conditions
(Project/CPU/Src/simulation.cc:80-101). The states below are new variants of the
reference's Kelvin-Helmholtz (initFunc.cc:512-563) and Orszag-Tang
(initFunc.cc:229-271) set-ups with the sizes and profiles of the workloads, so
they are generated here, once, and the same array is handed to every engine.

Random perturbations come from a counter-based generator (splitmix64 of the
seed and the cell's linear index) so the same field can be produced by numpy,
by C++ and on a GPU without sharing generator state.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(z: np.ndarray) -> np.ndarray:
    """splitmix64 finaliser (Steele, Lea & Flood 2014) on a uint64 array."""
    z = np.asarray(z, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
    return z ^ (z >> np.uint64(31))


def uniform01(seed: int, index: np.ndarray, stream: int) -> np.ndarray:
    """U[0,1) doubles: 53 high bits of splitmix64(seed + golden*(2*index+stream+1))."""
    idx = np.asarray(index, dtype=np.uint64)
    with np.errstate(over="ignore"):
        key = np.uint64(seed & 0xFFFFFFFFFFFFFFFF) + _GOLDEN * (np.uint64(2) * idx + np.uint64(stream + 1))
    bits = splitmix64(key) >> np.uint64(11)
    return bits.astype(np.float64) * (1.0 / 9007199254740992.0)


@dataclass
class Grid:
    """Padded grid of the reference ``Data`` (simData.cc:61-97, simulation.cc:34-61)."""

    nx: int
    ny: int
    nz: int
    xmin: float
    xmax: float
    ymin: float
    ymax: float
    zmin: float = 0.0
    zmax: float = 1.0
    Ng: int = 4

    @property
    def dims(self) -> int:
        return 1 if self.ny == 0 else (2 if self.nz == 0 else 3)

    @property
    def shape(self) -> tuple[int, int, int]:
        Nx = self.nx + 2 * self.Ng
        Ny = self.ny + 2 * self.Ng if self.ny > 0 else 1
        Nz = self.nz + 2 * self.Ng if self.nz > 0 else 1
        return Nx, Ny, Nz

    def centres(self):
        """Cell centres x[i] = xmin + (i + 0.5 - Ng) dx (simulation.cc:53-61)."""
        Nx, Ny, Nz = self.shape
        dx = (self.xmax - self.xmin) / self.nx
        x = self.xmin + (np.arange(Nx) + 0.5 - self.Ng) * dx
        if self.ny > 0:
            dy = (self.ymax - self.ymin) / self.ny
            y = self.ymin + (np.arange(Ny) + 0.5 - self.Ng) * dy
        else:
            y = np.zeros(1)
        if self.nz > 0:
            dz = (self.zmax - self.zmin) / self.nz
            z = self.zmin + (np.arange(Nz) + 0.5 - self.Ng) * dz
        else:
            z = np.zeros(1)
        return x, y, z


def kh_prims(grid: Grid, seed: int = 1234, shear: float = 0.5, bz: float = 0.1, width: float = 0.01,
             rho_in: float = 0.55, rho_out: float = 0.45, amp: float = 0.01, pressure: float = 1.0) -> np.ndarray:
    """Ideal SRMHD Kelvin-Helmholtz shear layers at |y| = 0.5 (8 primitives).

    rho = rho_in inside |y|<0.5 and rho_out outside, joined by tanh((|y|-0.5)/width);
    vx = shear * tanh((|y|-0.5)/width) (the two-branch form of initFunc.cc:540-549);
    vy (and vz in 3-D) = amp * (2u-1) with u from the seeded counter generator;
    p = pressure, Bz = bz, Bx = By = 0.
    """
    Nx, Ny, Nz = grid.shape
    _, y, _ = grid.centres()
    prims = np.zeros((8, Nx, Ny, Nz), dtype=np.float64)
    t = np.tanh((np.abs(y) - 0.5) / width)
    half_sum = 0.5 * (rho_in + rho_out)
    half_diff = 0.5 * (rho_in - rho_out)
    prims[0] = (half_sum - half_diff * t)[None, :, None]
    prims[1] = (shear * t)[None, :, None]
    n = np.arange(Nx * Ny * Nz, dtype=np.uint64).reshape(Nx, Ny, Nz)
    prims[2] = amp * (2.0 * uniform01(seed, n, 0) - 1.0)
    if grid.dims == 3:
        prims[3] = amp * (2.0 * uniform01(seed, n, 1) - 1.0)
    prims[4] = pressure
    prims[7] = bz
    return prims


def ot_prims_resistive(grid: Grid, rho: float = 1.0, pressure: float = 1.0) -> np.ndarray:
    """Resistive SRRMHD Orszag-Tang variant on [0,1]^2 (11 primitives).

    v = (-sin 2pi y, sin 2pi x) / (2 sqrt 2) so that max|v| = 0.5,
    B = (-sin 2pi y, sin 4pi x, 0), E = -v x B; cf. initFunc.cc:229-271 for the
    reference's ideal form.
    """
    Nx, Ny, Nz = grid.shape
    x, y, _ = grid.centres()
    X = x[:, None, None]
    Y = y[None, :, None]
    two_pi = 2.0 * math.pi
    s = 1.0 / (2.0 * math.sqrt(2.0))
    prims = np.zeros((11, Nx, Ny, Nz), dtype=np.float64)
    vx = -s * np.sin(two_pi * Y) + 0.0 * X
    vy = s * np.sin(two_pi * X) + 0.0 * Y
    bx = -np.sin(two_pi * Y) + 0.0 * X
    by = np.sin(2.0 * two_pi * X) + 0.0 * Y
    prims[0] = rho
    prims[1] = vx
    prims[2] = vy
    prims[4] = pressure
    prims[5] = bx
    prims[6] = by
    # E = -v x B with vz = Bz = 0
    prims[8] = 0.0
    prims[9] = 0.0
    prims[10] = -(vx * by - vy * bx)
    return prims


# Workload grids -------------------------------------------------------------

def current_sheet_prims(grid: Grid, sigma: float, direction: int = 0) -> np.ndarray:
    """CurrentSheetSingleFluid (initFunc.cc:191-226): rho 1, p 50, at rest, and
    one field component B0 erf(s sqrt(sigma) / 2) along axis `direction`
    (By(x), Bz(y) or Bx(z)), the resistive current sheet of the REGIME
    examples; gamma 2 on [-3, 3] along that axis."""
    p = np.zeros((8,) + grid.shape)
    p[0], p[4] = 1.0, 50.0
    centres = grid.centres()[direction]
    erf = np.vectorize(math.erf)
    b = erf(0.5 * centres * math.sqrt(sigma))
    shape = [1, 1, 1]
    shape[direction] = -1
    p[(6, 7, 5)[direction]] = b.reshape(shape)
    return p


def kh2d_grid(n: int = 64) -> Grid:
    """2-D KH grid on [-0.5,0.5]x[-1,1] (n x n interior cells, Ng=4)."""
    return Grid(n, n, 0, -0.5, 0.5, -1.0, 1.0)


def kh3d_grid(n: int = 256) -> Grid:
    """3-D KH grid on [-0.5,0.5]x[-1,1]x[0,1] (n^3 interior cells, Ng=4)."""
    return Grid(n, n, n, -0.5, 0.5, -1.0, 1.0, 0.0, 1.0)


def ot2d_grid(n: int = 1024) -> Grid:
    """2-D Orszag-Tang grid on [0,1]^2 (n x n interior cells, Ng=4)."""
    return Grid(n, n, 0, 0.0, 1.0, 0.0, 1.0)


def kh_ensemble(n: int = 64, seed: int = 2024, sizes=(128, 256, 384, 512, 768)):
    """Members of the batched Kelvin-Helmholtz ensemble: independent nx, ny drawn
    uniformly from `sizes`, shear velocity U[0.2, 0.8], Bz U[0.01, 0.5] and a
    per-member perturbation seed, all from one seeded generator. Returns a list
    of (nx, ny, shear, bz, seed)."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n):
        nx = int(rng.choice(sizes))
        ny = int(rng.choice(sizes))
        shear = float(rng.uniform(0.2, 0.8))
        bz = float(rng.uniform(0.01, 0.5))
        out.append((nx, ny, shear, bz, int(seed * 1000 + k)))
    return out
