"""Synthetic specimens for the benchmark configurations.

All positions in angstrom.  Random numbers come from numpy's Philox
bit generator seeded with 0x4D534C + job index, so every structure is
reproducible on any host.
"""
from __future__ import annotations

import numpy as np

SEED = 0x4D534C

# fractional bases
DIAMOND = np.array([(0, 0, 0), (0, .5, .5), (.5, 0, .5), (.5, .5, 0),
                    (.25, .25, .25), (.25, .75, .75), (.75, .25, .75), (.75, .75, .25)])


def supercell(basis_frac, Zs, a, nx, ny, nz):
    """Tile a cubic cell of edge a; returns (Z int32, x, y, z float32)."""
    basis_frac = np.asarray(basis_frac, float)
    Zs = np.asarray(Zs, np.int32)
    ii, jj, kk = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    cells = np.stack([ii.ravel(), jj.ravel(), kk.ravel()], 1).astype(float)
    pos = (cells[:, None, :] + basis_frac[None, :, :]).reshape(-1, 3) * a
    Z = np.tile(Zs, len(cells))
    return Z, pos[:, 0].astype(np.float32), pos[:, 1].astype(np.float32), pos[:, 2].astype(np.float32)


def silicon(nx, ny, nz, a=5.43):
    return supercell(DIAMOND, [14] * 8, a, nx, ny, nz)


def srtio3(nx, ny, nz, a=3.905):
    basis = [(0, 0, 0), (.5, .5, .5), (.5, .5, 0), (.5, 0, .5), (0, .5, .5)]
    return supercell(basis, [38, 22, 8, 8, 8], a, nx, ny, nz)


def random_atoms(n, box, zrange=(6, 79), dmin=1.4, seed=SEED, max_tries=200000):
    """Random sequential addition with a minimum pair distance (cell list)."""
    rng = np.random.Generator(np.random.Philox(seed))
    box = np.asarray(box, float)
    cell = dmin
    ncell = np.maximum((box // cell).astype(int), 1)
    grid = {}
    pts = []
    tries = 0
    while len(pts) < n and tries < max_tries * 10:
        tries += 1
        p = rng.random(3) * box
        c = tuple((p // cell).astype(int) % ncell)
        ok = True
        for dx in (-1, 0, 1):
            for dy in (-1, 0, 1):
                for dz in (-1, 0, 1):
                    for q in grid.get(((c[0] + dx) % ncell[0], (c[1] + dy) % ncell[1], (c[2] + dz) % ncell[2]), ()):
                        if np.sum((p - q) ** 2) < dmin * dmin:
                            ok = False
                            break
                    if not ok:
                        break
                if not ok:
                    break
            if not ok:
                break
        if ok:
            grid.setdefault(c, []).append(p)
            pts.append(p)
    pts = np.array(pts)
    Z = rng.integers(zrange[0], zrange[1] + 1, size=len(pts)).astype(np.int32)
    return Z, pts[:, 0].astype(np.float32), pts[:, 1].astype(np.float32), pts[:, 2].astype(np.float32)


def jittered_atoms(n, box, zrange=(6, 79), dmin=1.4, seed=SEED):
    """n atoms spread uniformly over box with every pair at least dmin apart,
    in O(n): a random subset of the sites of a cubic lattice of spacing a
    (the largest spacing that still offers n sites), each site moved by an
    independent uniform offset in [-(a - dmin)/2, (a - dmin)/2] per axis, so
    atoms on different sites differ by at least dmin along some axis."""
    rng = np.random.Generator(np.random.Philox(seed))
    box = np.asarray(box, float)
    a = (box.prod() / n) ** (1.0 / 3.0)
    while np.prod(np.maximum(np.floor(box / a), 1)) < n:
        a *= 0.99
    if a <= dmin:
        raise ValueError(f"{n} atoms do not fit in {box.tolist()} at a minimum distance of {dmin}")
    m = np.maximum(np.floor(box / a), 1).astype(int)
    sites = rng.choice(int(np.prod(m)), size=n, replace=False)
    idx = np.stack(np.unravel_index(sites, m), 1).astype(float)
    pos = (idx + 0.5) * a + rng.uniform(-(a - dmin) / 2, (a - dmin) / 2, size=(n, 3))
    Z = rng.integers(zrange[0], zrange[1] + 1, size=n).astype(np.int32)
    return Z, pos[:, 0].astype(np.float32), pos[:, 1].astype(np.float32), pos[:, 2].astype(np.float32)


def amorphous_carbon(box=(102.4, 102.4, 200.0), density_per_A3=0.0602, dmin=1.4, seed=SEED):
    n = int(round(density_per_A3 * box[0] * box[1] * box[2]))
    return random_atoms(n, box, (6, 6), dmin, seed)
