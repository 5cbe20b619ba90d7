"""Synthetic, seeded inputs for the solver's workloads (geometry, porosity fields, initial states).

All generators use numpy's PCG64 (numpy.random.default_rng(seed)), so a (size, seed) pair gives
the same input on every machine.  Layouts match the C-ABI: cell index i = x + nx*(y + ny*z),
arrays shaped (nz, ny, nx).
"""
from __future__ import annotations

import numpy as np

DEFAULT_SEED = 20261202


def porous_spheres(n: int, porosity: float = 0.8, rmin: int = 4, rmax: int = 12, seed: int = DEFAULT_SEED):
    """Periodic n^3 box; solid spheres with integer radius uniform in [rmin, rmax] and uniformly
    random centres are added (overlaps allowed, wrapping periodically) until the fluid fraction
    drops to `porosity`.  Returns a bool array (nz, ny, nx), True = solid."""
    rng = np.random.default_rng(seed)
    solid = np.zeros((n, n, n), dtype=bool)
    target = int(round((1.0 - porosity) * n ** 3))
    count = 0
    while count < target:
        c = rng.integers(0, n, size=3)
        r = int(rng.integers(rmin, rmax + 1))
        off = np.arange(-r, r + 1)
        d2 = off[:, None, None] ** 2 + off[None, :, None] ** 2 + off[None, None, :] ** 2
        ball = d2 <= r * r
        zi = (c[2] + off) % n
        yi = (c[1] + off) % n
        xi = (c[0] + off) % n
        sub = solid[np.ix_(zi, yi, xi)]
        new = ball & ~sub
        if 2 * r + 1 > n:  # a ball wider than the box wraps onto itself
            sub |= ball
            solid[np.ix_(zi, yi, xi)] = sub
            count = int(solid.sum())
            continue
        solid[np.ix_(zi, yi, xi)] = sub | ball
        count += int(new.sum())
    return solid


def channel_walls(nx: int, ny: int, nz: int):
    """Solid walls on the y and z faces (a square duct, periodic along x)."""
    solid = np.zeros((nz, ny, nx), dtype=bool)
    solid[0, :, :] = solid[-1, :, :] = True
    solid[:, 0, :] = solid[:, -1, :] = True
    return solid


def fourier_porosity(nx: int, ny: int, nz: int, modes: int = 8, seed: int = DEFAULT_SEED):
    """Smooth porosity field: sum of `modes` random Fourier modes, rescaled and clipped to [0, 1]."""
    rng = np.random.default_rng(seed)
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    a = np.zeros((nz, ny, nx))
    for _ in range(modes):
        k = rng.integers(1, 4, size=3) * 2 * np.pi / np.array([nx, ny, nz])
        phase = rng.uniform(0, 2 * np.pi)
        amp = rng.uniform(0.5, 1.0)
        a += amp * np.sin(k[0] * x + k[1] * y + k[2] * z + phase)
    a = 0.5 + 0.5 * a / np.abs(a).max()
    return np.clip(a, 0.0, 1.0)


# D3Q19 directions (z-major), as in the solver
_C = np.array([[0, -1, -1], [-1, 0, -1], [0, 0, -1], [1, 0, -1], [0, 1, -1], [-1, -1, 0], [0, -1, 0], [1, -1, 0],
               [-1, 0, 0], [0, 0, 0], [1, 0, 0], [-1, 1, 0], [0, 1, 0], [1, 1, 0], [0, -1, 1], [-1, 0, 1],
               [0, 0, 1], [1, 0, 1], [0, 1, 1]])


def torus_table(n: int, dims, taus, first: int = 0):
    """Patch table entries for an (a, b, c) torus of edge-n patches: patch (i, j, k) -> id
    first + i + a*(j + b*k); its neighbour along c_q is the patch at (i, j, k) + c_q (periodic)."""
    a, b, c = dims
    out = []
    for k in range(c):
        for j in range(b):
            for i in range(a):
                nbr = []
                for q in range(19):
                    ii, jj, kk = (i + _C[q, 0]) % a, (j + _C[q, 1]) % b, (k + _C[q, 2]) % c
                    nbr.append(first + ii + a * (jj + b * kk))
                out.append((n, float(taus[len(out)]), nbr))
    return out


def uneven_patches(count: int = 4096, edges=(8, 16, 24, 32), tau_range=(0.55, 1.5), seed: int = DEFAULT_SEED):
    """The batched-patch workload: `count` cubic patches with edges drawn uniformly from `edges`
    and tau uniform in tau_range.  Patches of equal edge are joined into 2x2x2 tori (the remainder
    of each edge class into one ring), so every face and edge exchange crosses between patches.
    Returns (table, tori) where tori lists (first_id, dims, n) per torus."""
    rng = np.random.default_rng(seed)
    n_of = rng.choice(np.asarray(edges), size=count)
    taus = rng.uniform(tau_range[0], tau_range[1], size=count)
    table, tori = [], []
    for n in edges:
        members = int((n_of == n).sum())
        k = 0
        while k < members:
            size = 8 if members - k >= 8 else members - k
            dims = (2, 2, 2) if size == 8 else (size, 1, 1)
            first = len(table)
            table += torus_table(int(n), dims, taus[first:first + size], first)
            tori.append((first, dims, int(n)))
            k += size
    return table, tori


def taylor_green(n: int, u0: float):
    """Taylor-Green vortex on a periodic n^3 box (k = 2 pi / n): rho and u per cell."""
    k = 2 * np.pi / n
    # broadcast 1-D coordinates (same values and operation order as full meshgrids, a fraction of
    # the memory: 1024^3 fits in ~45 GB)
    a = np.arange(n)
    x, y, z = a[None, None, :], a[None, :, None], a[:, None, None]
    u = np.zeros((n, n, n, 3))
    u[..., 0] = u0 * np.sin(k * x) * np.cos(k * y) * np.cos(k * z)
    u[..., 1] = -u0 * np.cos(k * x) * np.sin(k * y) * np.cos(k * z)
    rho = 1.0 + 3.0 * u0 ** 2 / 16.0 * (np.cos(2 * k * x) + np.cos(2 * k * y)) * (np.cos(2 * k * z) + 2.0)
    return rho, u
