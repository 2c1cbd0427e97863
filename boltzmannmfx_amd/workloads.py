"""Synthetic inputs and force parameters for the standard workloads.

Force parameters are BMX's cell_force deck keys in fpar order
(src/chemistry/bmx_cell_interaction.cpp:84-93):
  [boundary_width W_A, stiffness mu, z_wall, wall_boundary_width W_Z,
   wall_stiffness xi, gravity g, bond_strength k_b, bond_cutoff, viscous_drag]

Sphere colonies use the values of the bacteria deck
exec/nlev_large_test/input_real_nlev (mu = xi = 1e6, W_Z = 1e-4, g = 0.01);
the growth surface is the bottom of the box (z_wall = prob_lo[2]).  The
adhesion band W_A is 5e-5 as in that deck for the 2,000-sphere case and
0.1 * mean radius (6e-6) for the large colonies.  Hyphae use the fungi deck
exec/fungi/input_fungi (mu = 1e6, xi = 1e5, W_A = W_Z = 1e-4,
viscous_drag = 200, g = 0.01) with bond_strength 1e-2.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import List, Tuple

import numpy as np

from . import (FUNGI, RECORD_DTYPE, Geometry, generate_hyphae, generate_spheres,
               make_geometry, mean_sphere_volume)

R_MIN, R_MAX = 0.5e-4, 0.7e-4
DT = 0.25


def sphere_fpar(W_A: float, z_wall: float = 0.0) -> np.ndarray:
    return np.array([W_A, 1.0e6, z_wall, 1.0e-4, 1.0e6, 0.01, 1.0, 5.0e-4, 1.0])


def hyphae_fpar(z_wall: float = 0.0, k_bond: float = 1.0e-2) -> np.ndarray:
    return np.array([1.0e-4, 1.0e6, z_wall, 1.0e-4, 1.0e5, 0.01, k_bond, 5.0e-4, 200.0])


@dataclass
class Colony:
    recs: np.ndarray
    prob_lo: np.ndarray
    prob_hi: np.ndarray
    periodic: Tuple[int, int, int]
    fpar: np.ndarray
    dt: float = DT
    n_cell: Tuple[int, int, int] = (1, 1, 1)
    skin: float = 2.0e-5
    rebuild_every: int = 10
    meta: dict = field(default_factory=dict)

    def geometry(self) -> Geometry:
        return make_geometry(self.prob_lo, self.prob_hi, self.periodic, self.n_cell)


def sphere_colony(n: int, phi: float, seed: int, W_A: float = 6.0e-6,
                  overlap_free: bool = True, skin: float = 2.0e-5) -> Colony:
    """n spheres, r ~ U[0.5e-4, 0.7e-4] cm, 2:2:1 box sized to volume fraction
    phi, periodic in x/y, growth surface at the box bottom."""
    recs, lo, hi = generate_spheres(n, R_MIN, R_MAX, phi, seed, (2.0, 2.0, 1.0),
                                    (1, 1, 0), overlap_free)
    return Colony(recs, lo, hi, (1, 1, 0), sphere_fpar(W_A, z_wall=lo[2]),
                  skin=skin, meta={"n": n, "phi": phi, "seed": seed, "W_A": W_A})


def c1_small(seed: int = 21893) -> Colony:
    """2,000 spheres at volume fraction 0.2 (2.65e-3 x 2.65e-3 x 1.33e-3 cm)."""
    return sphere_colony(2000, 0.2, seed, W_A=5.0e-5)


# Verlet policy of the large sphere workloads (the 1M colony and the
# 256-colony ensemble): lists rebuilt every 10 steps, the interval the
# 1M-sphere workload is specified with, with a skin of 4e-6 cm.  Every call
# checks the summed per-step displacement maxima against skin / 2 and reports
# an interval too long for the skin; the largest displacement over 40 steps
# of the 1M colony is 1.68e-6 cm, so 10 steps stay well inside 2e-6.
C2_SKIN = 4.0e-6
C2_REBUILD_EVERY = 10


def c2_large(n: int = 1_000_000, seed: int = 7) -> Colony:
    """1,000,000 spheres at volume fraction 0.3, adhesion band 0.1 * r."""
    col = sphere_colony(n, 0.3, seed, W_A=6.0e-6, skin=C2_SKIN)
    col.rebuild_every = C2_REBUILD_EVERY
    return col


def c3_hyphae(n: int = 200_000, nroots: int = 2000, seed: int = 43343,
              coverage: float = 0.3, mesh_dx: float = 2.0e-3) -> Colony:
    """Branching hyphae: segments 30e-4 cm long, radius 2e-4 cm, branch
    probability 0.05 per segment at 30-60 degrees, lying on the growth
    surface z = 0 of a periodic x/y box sized so the segments' footprint
    covers `coverage` of it.  n_cell sets the reference-mesh bins (dx =
    mesh_dx, the fungi deck's finest level)."""
    L, r = 30.0e-4, 2.0e-4
    side = math.sqrt(n * L * 2.0 * r / coverage)
    ncx = max(4, int(round(side / mesh_dx)))
    side = ncx * mesh_dx
    height = 8 * mesh_dx
    lo = np.array([0.0, 0.0, 0.0])
    hi = np.array([side, side, height])
    recs = generate_hyphae(n, nroots, L, r, 0.05, 30.0, 60.0, seed, lo, hi, 0.0)
    return Colony(recs, lo, hi, (1, 1, 0), hyphae_fpar(0.0), n_cell=(ncx, ncx, 8),
                  skin=1.0e-3, meta={"n": n, "nroots": nroots, "seed": seed})


def c4_ensemble(ncol: int = 256, seed: int = 2024, nmin: int = 500,
                nmax: int = 50_000, skin: float = C2_SKIN) -> List[Colony]:
    """Colonies with cell counts log-uniform in [nmin, nmax] and volume
    fraction uniform in [0.1, 0.35], each in its own periodic x/y box."""
    rng = np.random.default_rng(seed)
    sizes = np.exp(rng.uniform(math.log(nmin), math.log(nmax), ncol)).astype(int)
    phis = rng.uniform(0.1, 0.35, ncol)
    seeds = rng.integers(1, 2**62, ncol)
    cols = [sphere_colony(int(n), float(p), int(s), skin=skin)
            for n, p, s in zip(sizes, phis, seeds)]
    for c in cols:
        c.rebuild_every = C2_REBUILD_EVERY
    return cols


def concat_colonies(cols: List[Colony]):
    """Stacked records + colony_start offsets + per-colony geometries."""
    recs = np.concatenate([c.recs for c in cols]).astype(RECORD_DTYPE)
    starts = np.zeros(len(cols) + 1, dtype=np.int64)
    starts[1:] = np.cumsum([len(c.recs) for c in cols])
    return recs, starts, [c.geometry() for c in cols]


def bin_dims_verlet(col: Colony, radius: float, pct: int = 100):
    """Bins per axis the library picks for a Verlet radius (for the oracle)."""
    edge = radius * pct / 100.0
    L = np.asarray(col.prob_hi) - np.asarray(col.prob_lo)
    return tuple(max(1, int(math.floor(L[d] / edge))) for d in range(3))


def verlet_radius(col: Colony) -> float:
    rec = col.recs
    W_A = col.fpar[0]
    rl = 0.0
    types = rec["idata"][:, 5]
    if np.any(types == 0):
        rl = max(rl, 2.0 * rec["rdata"][types == 0, 0].max() + W_A)
    if np.any(types == FUNGI):
        rl = max(rl, 1.5 * (2.0 * rec["rdata"][types == FUNGI, 1].max()) + W_A)
    return rl + col.skin
