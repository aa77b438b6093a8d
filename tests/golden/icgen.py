"""Deterministic initial data for golden vectors and parity tests.

Values come from a SplitMix64 hash of (seed, variable, cell or block index) and are
dyadic rationals, so every value is exact in float32 and identical on any machine.
SplitMix64: Steele, Lea, Flood, "Fast splittable pseudorandom number generators",
OOPSLA 2014.
"""
import numpy as np

_GOLD = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(x):
    x = np.asarray(x, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = x + _GOLD
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
    return z ^ (z >> np.uint64(31))


def _hash(seed, var, idx):
    key = (np.uint64(seed) << np.uint64(48)) ^ (np.uint64(var) << np.uint64(40)) ^ idx.astype(np.uint64)
    return splitmix64(key)


def _fields(h_rho, h_u, h_v, h_T):
    rho = 0.5 + (h_rho % np.uint64(384)).astype(np.float64) / 32.0          # [0.5, 12.47]
    u = ((h_u >> np.uint64(16)) % np.uint64(257)).astype(np.float64) / 128.0 - 1.0   # [-1, 1]
    v = ((h_v >> np.uint64(16)) % np.uint64(257)).astype(np.float64) / 128.0 - 1.0
    T = 0.5 + ((h_T >> np.uint64(32)) % np.uint64(193)).astype(np.float64) / 128.0  # [0.5, 2]
    return rho, u, v, T


def random_1d(n, seed, block=1):
    """(3, n) primitives rho, u, T; constant over blocks of `block` cells."""
    idx = np.arange(n, dtype=np.int64) // block
    rho, u, _, T = _fields(_hash(seed, 0, idx), _hash(seed, 1, idx), _hash(seed, 2, idx), _hash(seed, 3, idx))
    return np.stack([rho, u, T])


def random_2d(nx, ny, seed, block=1, y_uniform=False):
    """(4, nx*ny) primitives rho, u, v, T (index i*ny+j).

    block: cells per block in each direction.  y_uniform: the field depends on i
    only and v = 0 (a 1D problem embedded in 2D)."""
    i = np.repeat(np.arange(nx, dtype=np.int64), ny) // block
    j = np.tile(np.arange(ny, dtype=np.int64), nx) // block
    idx = i if y_uniform else i * 65536 + j
    rho, u, v, T = _fields(_hash(seed, 0, idx), _hash(seed, 1, idx), _hash(seed, 2, idx), _hash(seed, 3, idx))
    if y_uniform:
        v = np.zeros_like(v)
    return np.stack([rho, u, v, T])


def sod_2d_y_uniform(nx, ny, rhoL=10.0, rhoR=1.0):
    """Sod data along x embedded in 2D: rho = rhoL for i < nx/2 else rhoR, u = v = 0, T = 1."""
    i = np.repeat(np.arange(nx), ny)
    rho = np.where(i < 0.5 * nx, rhoL, rhoR)
    z = np.zeros(nx * ny)
    return np.stack([rho, z, z.copy(), np.ones(nx * ny)])


def sample_indices(nx, ny, count=256, seed=7):
    """Fixed cell sample for 2D digests: corners, edge midpoints, plus hashed cells."""
    fixed = [0, ny - 1, (nx - 1) * ny, nx * ny - 1, ny // 2, (nx // 2) * ny, (nx // 2) * ny + ny // 2,
             (nx - 1) * ny + ny // 2, (nx // 2) * ny + ny - 1]
    h = splitmix64(np.arange(count, dtype=np.uint64) + np.uint64(seed << 32))
    rnd = (h % np.uint64(nx * ny)).astype(np.int64)
    return np.unique(np.concatenate([np.array(fixed, dtype=np.int64), rnd]))


def circle_2d(nx, ny, cx=0.5, cy=0.5, r=0.25, rho_in=10.0, T_in=1.0, rho_out=1.0, T_out=1.0):
    """rho_in where the cell centre ((i+1/2)/nx, (j+1/2)/ny) lies strictly inside the circle."""
    x = (np.repeat(np.arange(nx), ny) + 0.5) * (1.0 / nx) - cx
    y = (np.tile(np.arange(ny), nx) + 0.5) * (1.0 / ny) - cy
    inside = x * x + y * y < r * r
    z = np.zeros(nx * ny)
    return np.stack([np.where(inside, rho_in, rho_out), z, z.copy(), np.where(inside, T_in, T_out)])
