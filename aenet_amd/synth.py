"""Seeded synthetic structures for the workloads this package is measured on.

Lattice vectors are rows (ASE convention); positions are Cartesian in Å;
types are 1-based indices into the species list of the potential set.
All generators use numpy's PCG64 with fixed seeds so that every run (and the
oracle) sees identical inputs.
"""
import math

import numpy as np

FCC_BASIS = np.array([[0.0, 0.0, 0.0], [0.5, 0.5, 0.0], [0.5, 0.0, 0.5], [0.0, 0.5, 0.5]])
BCC_BASIS = np.array([[0.0, 0.0, 0.0], [0.5, 0.5, 0.5]])


class Structure:
    def __init__(self, cell, positions, types, pbc=True, label=""):
        self.cell = np.asarray(cell, dtype=np.float64).reshape(3, 3)
        self.positions = np.ascontiguousarray(positions, dtype=np.float64).reshape(-1, 3)
        self.types = np.ascontiguousarray(types, dtype=np.int32).reshape(-1)
        self.pbc = bool(pbc)
        self.label = label

    def __len__(self):
        return len(self.types)

    def copy(self):
        return Structure(self.cell.copy(), self.positions.copy(), self.types.copy(), self.pbc, self.label)


def lattice(basis, a, reps):
    n1, n2, n3 = reps
    grid = np.array([[i, j, k] for i in range(n1) for j in range(n2) for k in range(n3)], dtype=np.float64)
    frac = (grid[:, None, :] + basis[None, :, :]).reshape(-1, 3)
    cell = np.diag([a * n1, a * n2, a * n3]).astype(np.float64)
    return cell, frac * a


def fcc(a, reps):
    return lattice(FCC_BASIS, a, reps)


def displaced(pos, sigma, seed):
    rng = np.random.Generator(np.random.PCG64(seed))
    return pos + rng.normal(0.0, sigma, size=pos.shape)


def c1_fcc108():
    """Single-species 108-atom FCC cell, a = 3.61 Å, 3x3x3 conventional,
    Gaussian displacements sigma = 0.1 Å (seed 1)."""
    cell, pos = fcc(3.61, (3, 3, 3))
    return Structure(cell, displaced(pos, 0.1, 1), np.ones(len(pos), np.int32), True, "fcc108")


def binary_fcc(a, reps, sigma, seed_species, seed_disp, p=0.5):
    cell, pos = fcc(a, reps)
    rng = np.random.Generator(np.random.PCG64(seed_species))
    types = np.where(rng.random(len(pos)) < p, 1, 2).astype(np.int32)
    return Structure(cell, displaced(pos, sigma, seed_disp), types, True)


def c2_alloy4000():
    """4000-atom binary A/B FCC cell, a = 3.8 Å, 10x10x10 conventional,
    species Bernoulli(0.5) (seed 11), displacements sigma = 0.08 Å (seed 12)."""
    s = binary_fcc(3.8, (10, 10, 10), 0.08, 11, 12)
    s.label = "alloy4000"
    return s


def alloy_replicas(nrep, seed0=1000):
    """Independent 4000-atom alloy cells of the same shape (new species draw
    and displacements per replica)."""
    out = []
    for r in range(nrep):
        s = binary_fcc(3.8, (10, 10, 10), 0.08, seed0 + 2 * r, seed0 + 2 * r + 1)
        s.label = "alloy4000-%d" % r
        out.append(s)
    return out


def mc_cell(seed_species, a=3.61):
    """256-atom binary A/B FCC cell, 4x4x4 conventional."""
    cell, pos = fcc(a, (4, 4, 4))
    rng = np.random.Generator(np.random.PCG64(seed_species))
    types = np.where(rng.random(len(pos)) < 0.5, 1, 2).astype(np.int32)
    return Structure(cell, pos, types, True, "mc256")


def _supercell_reps(n_target, per_cell):
    """3 integers (n1 <= n2 <= n3 ... any order) whose product times per_cell
    is closest to n_target, favouring near-cubic shapes."""
    best = None
    m = max(1, int(round(n_target / per_cell)))
    for n1 in range(1, m + 1):
        for n2 in range(n1, m + 1):
            n3 = max(1, int(round(m / (n1 * n2))))
            if n3 < n2:
                continue
            tot = n1 * n2 * n3 * per_cell
            score = abs(tot - n_target) + 0.01 * (n3 - n1)
            if best is None or score < best[0]:
                best = (score, (n1, n2, n3))
    return best[1] if best else (1, 1, 1)


def batch_uneven(nstruct=2048, ntypes=3, seed_sizes=21, seed_comp=22, seed_disp=23):
    """Batched uneven structures: atom counts log-uniform in [8, 512]
    (seed 21), a mix of periodic bulk FCC (a = 3.6 Å) / BCC (a = 2.87 Å),
    slabs with 12 Å vacuum and isolated clusters; species drawn with
    Dirichlet(1,1,1) compositions (seed 22); sigma = 0.1 Å."""
    rs = np.random.Generator(np.random.PCG64(seed_sizes))
    rc = np.random.Generator(np.random.PCG64(seed_comp))
    rd = np.random.Generator(np.random.PCG64(seed_disp))
    sizes = np.rint(np.exp(rs.uniform(math.log(8), math.log(512), nstruct))).astype(int)
    kinds = rs.integers(0, 4, nstruct)
    out = []
    for n, kind in zip(sizes, kinds):
        if kind == 0:       # bulk FCC
            reps = _supercell_reps(n, 4)
            cell, pos = fcc(3.6, reps)
            pbc = True
        elif kind == 1:     # bulk BCC
            reps = _supercell_reps(n, 2)
            cell, pos = lattice(BCC_BASIS, 2.87, reps)
            pbc = True
        elif kind == 2:     # FCC slab, vacuum along c
            reps = _supercell_reps(n, 4)
            cell, pos = fcc(3.6, reps)
            cell = cell.copy()
            cell[2, 2] += 12.0
            pbc = True
        else:               # isolated FCC cluster: the n lattice sites nearest the origin
            m = int(math.ceil((n / 4.0) ** (1.0 / 3.0))) + 2
            _, big = fcc(3.6, (2 * m, 2 * m, 2 * m))
            big = big - big.mean(axis=0)
            order = np.argsort(np.sum(big * big, axis=1), kind="stable")
            pos = big[order[:n]]
            cell = np.eye(3)
            pbc = False
        comp = rc.dirichlet(np.ones(ntypes))
        types = (rc.choice(ntypes, size=len(pos), p=comp) + 1).astype(np.int32)
        pos = pos + rd.normal(0.0, 0.1, size=pos.shape)
        out.append(Structure(cell, pos, types, pbc, ["fcc", "bcc", "slab", "cluster"][kind]))
    return out
