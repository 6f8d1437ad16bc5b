"""Seeded synthetic structures for the SOAP workloads, plus a minimal xyz reader.

The water clusters and boxes follow one recipe: O sites on an 'ij'-ordered
cubic lattice, each coordinate jittered uniformly, each molecule rotated by a
uniform random rotation (normalised Gaussian quaternion).  Local frame: O at
the origin, H at (+-0.9572 sin(52.26 deg), 0, 0.9572 cos(52.26 deg)), i.e.
O-H 0.9572 A and H-O-H 104.52 deg.  Atom order is O, H, H per molecule.

All systems are returned as dicts accepted by ``SOAP.create``:
``{"positions": (N,3) float64, "atomic_numbers": (N,) int64[, "cell", "pbc"]}``.
"""

from __future__ import annotations

import math
from typing import Dict, Optional, Sequence, Tuple

import numpy as np

OH = 0.9572
HOH_HALF = math.radians(104.52 / 2.0)
WATER_MASS = 18.015          # g/mol
AVOGADRO_E24 = 0.60221       # N_A * 1e-24 (A^3 -> cm^3)


def _quat_to_rot(q: np.ndarray) -> np.ndarray:
    q = q / np.linalg.norm(q, axis=1, keepdims=True)
    w, x, y, z = q[:, 0], q[:, 1], q[:, 2], q[:, 3]
    R = np.empty((q.shape[0], 3, 3))
    R[:, 0, 0] = 1 - 2 * (y * y + z * z)
    R[:, 0, 1] = 2 * (x * y - z * w)
    R[:, 0, 2] = 2 * (x * z + y * w)
    R[:, 1, 0] = 2 * (x * y + z * w)
    R[:, 1, 1] = 1 - 2 * (x * x + z * z)
    R[:, 1, 2] = 2 * (y * z - x * w)
    R[:, 2, 0] = 2 * (x * z - y * w)
    R[:, 2, 1] = 2 * (y * z + x * w)
    R[:, 2, 2] = 1 - 2 * (x * x + y * y)
    return R


def _water_sites(shape: Tuple[int, int, int], spacing: float, n_mol: Optional[int]) -> np.ndarray:
    gx, gy, gz = np.meshgrid(np.arange(shape[0]), np.arange(shape[1]), np.arange(shape[2]), indexing="ij")
    sites = np.stack([gx.reshape(-1), gy.reshape(-1), gz.reshape(-1)], axis=1).astype(np.float64) * spacing
    if n_mol is not None:
        sites = sites[:n_mol]
    return sites


def water(shape: Tuple[int, int, int], spacing: float = 3.1, jitter: float = 0.3, seed: int = 0,
          n_mol: Optional[int] = None, cell_length: Optional[float] = None) -> Dict[str, np.ndarray]:
    """(H2O)_n on a jittered cubic lattice; periodic when ``cell_length`` is given."""
    rng = np.random.default_rng(seed)
    O = _water_sites(shape, spacing, n_mol)
    m = O.shape[0]
    O = O + rng.uniform(-jitter, jitter, size=(m, 3))
    R = _quat_to_rot(rng.normal(size=(m, 4)))
    local = np.array([[OH * math.sin(HOH_HALF), 0.0, OH * math.cos(HOH_HALF)],
                      [-OH * math.sin(HOH_HALF), 0.0, OH * math.cos(HOH_HALF)]])
    H = O[:, None, :] + np.einsum("mij,hj->mhi", R, local)
    pos = np.concatenate([O[:, None, :], H], axis=1).reshape(-1, 3)
    Z = np.tile(np.array([8, 1, 1], dtype=np.int64), m)
    out = {"positions": np.ascontiguousarray(pos), "atomic_numbers": Z}
    if cell_length is not None:
        out["positions"] = np.mod(pos, cell_length)
        out["cell"] = np.eye(3) * cell_length
        out["pbc"] = np.array([True, True, True])
    return out


def water_cluster_300(seed: int = 0):
    """Non-periodic (H2O)_100: 5x5x4 lattice, 300 atoms."""
    return water((5, 5, 4), seed=seed)


def water_cluster_3000(seed: int = 0):
    """Non-periodic (H2O)_1000: 10x10x10 lattice, 3,000 atoms."""
    return water((10, 10, 10), seed=seed)


def water_box_length(n_mol: int, density: float = 1.0) -> float:
    """Cubic cell edge (A) of n_mol waters at ``density`` g/cm^3."""
    return (n_mol * WATER_MASS / (AVOGADRO_E24 * density)) ** (1.0 / 3.0)


def water_box(n_side: int = 32, seed: int = 0, density: float = 1.0):
    """Periodic bulk water, n_side^3 molecules at ``density`` (32 -> 98,304 atoms, L=99.338 A)."""
    L = water_box_length(n_side ** 3, density)
    return water((n_side, n_side, n_side), spacing=L / n_side, seed=seed, cell_length=L)


def water_cluster_1m(seed: int = 0):
    """Non-periodic (H2O)_333334: first 333,334 sites of a 70x70x69 lattice (1,000,002 atoms)."""
    return water((70, 70, 69), seed=seed, n_mol=333334)


def random_box(n_atoms: int = 20000, fractions: Sequence[float] = (0.50, 0.25, 0.10, 0.15),
               Z: Sequence[int] = (1, 6, 7, 8), density: float = 0.10, d_min: float = 0.9,
               seed: int = 0):
    """Periodic random multi-species box by sequential rejection sampling (minimum image).

    Counts per species are round(fraction * n_atoms) (the last species takes the
    remainder), the species labels are shuffled, and each atom is accepted when
    its minimum-image distance to every accepted atom is >= d_min.
    """
    rng = np.random.default_rng(seed)
    L = (n_atoms / density) ** (1.0 / 3.0)
    counts = [int(round(f * n_atoms)) for f in fractions[:-1]]
    counts.append(n_atoms - sum(counts))
    labels = np.concatenate([np.full(c, z, dtype=np.int64) for c, z in zip(counts, Z)])
    rng.shuffle(labels)
    nb = max(1, int(L // d_min))
    h = L / nb
    grid: Dict[Tuple[int, int, int], list] = {}
    pos = np.empty((n_atoms, 3))
    k = 0
    d2min = d_min * d_min
    while k < n_atoms:
        p = rng.uniform(0.0, L, size=3)
        ci = tuple(int(v) for v in np.minimum((p / h).astype(int), nb - 1))
        ok = True
        for dx in (-1, 0, 1):
            for dy in (-1, 0, 1):
                for dz in (-1, 0, 1):
                    key = ((ci[0] + dx) % nb, (ci[1] + dy) % nb, (ci[2] + dz) % nb)
                    for j in grid.get(key, ()):
                        d = pos[j] - p
                        d -= L * np.round(d / L)
                        if d @ d < d2min:
                            ok = False
                            break
                    if not ok:
                        break
                if not ok:
                    break
            if not ok:
                break
        if ok:
            pos[k] = p
            grid.setdefault(ci, []).append(k)
            k += 1
    return {"positions": pos, "atomic_numbers": labels, "cell": np.eye(3) * L,
            "pbc": np.array([True, True, True])}


SYMBOLS = {"H": 1, "He": 2, "Li": 3, "Be": 4, "B": 5, "C": 6, "N": 7, "O": 8, "F": 9, "Ne": 10,
           "Na": 11, "Mg": 12, "Al": 13, "Si": 14, "P": 15, "S": 16, "Cl": 17, "Ar": 18,
           "K": 19, "Ca": 20, "Fe": 26, "Cu": 29, "Zn": 30, "Br": 35, "I": 53}


def read_xyz(path: str):
    """Read a multi-frame plain xyz file (count line, comment line, 'sym x y z' lines)."""
    frames = []
    with open(path) as fh:
        lines = fh.read().splitlines()
    i = 0
    while i < len(lines):
        if not lines[i].strip():
            i += 1
            continue
        n = int(lines[i].split()[0])
        body = lines[i + 2:i + 2 + n]
        sym = [b.split()[0] for b in body]
        pos = np.array([[float(v) for v in b.split()[1:4]] for b in body], dtype=np.float64)
        Z = np.array([SYMBOLS[s] if s in SYMBOLS else int(s) for s in sym], dtype=np.int64)
        frames.append({"positions": pos, "atomic_numbers": Z})
        i += 2 + n
    return frames
