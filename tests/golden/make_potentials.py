"""Write the synthetic ænet potentials the workloads use (tests/golden/potentials/).

Run from the repository root:  python tests/golden/make_potentials.py
(``... make_potentials.py mixed`` writes only the mixed Behler-Parrinello set)

Weights are drawn N(0, 1/fan_in) with fixed seeds (fan_in = number of nodes
of the previous layer); normalization statistics (sfval_avg = mean,
sfval_cov = mean of squares, sfval_min/max) are the FP64 statistics of the
unscaled descriptors over seeded calibration cells, computed by the oracle
(oracle/_build/liboracle.so).  The files are written in the reference's
.nn format so that the reference library loads them unchanged.
"""
import os
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

from aenet_amd import nnfile, synth  # noqa: E402
from oracle_lib import Oracle  # noqa: E402

OUT = os.path.join(ROOT, "tests", "golden", "potentials")


def random_network(arch, seed, acts=None):
    rng = np.random.Generator(np.random.PCG64(seed))
    blocks = []
    for l in range(len(arch) - 1):
        n1, n2 = arch[l], arch[l + 1]
        # column-major (n2, n1+1) block, last column = bias
        blocks.append(rng.normal(0.0, 1.0 / np.sqrt(n1), size=(n1 + 1) * n2))
    if acts is None:
        acts = [1] * (len(arch) - 2) + [0]
    return np.concatenate(blocks), acts


def make_pot(atomtype, setup, arch, seed, types, e_atoms, acts=None):
    W, fa = random_network(arch, seed, acts)
    pot = dict(setup)
    pot.update({
        "nnodes": arch, "f_a": fa, "W": W, "description": "synthetic potential (seed %d)" % seed,
        "sfval_min": np.zeros(setup["nsf"]), "sfval_max": np.zeros(setup["nsf"]),
        "sfval_avg": np.zeros(setup["nsf"]), "sfval_cov": np.ones(setup["nsf"]),
        "ts_file": "synthetic", "ts_normalized": True, "ts_scale": 1.0, "ts_shift": 0.0,
        "ts_types": list(types), "ts_E_atom": list(e_atoms), "ts_natoms": 0, "ts_nstrucs": 0,
        "ts_E": [0.0, 0.0, 0.0],
    })
    pot["atomtype"] = atomtype
    return pot


def calibrate(pots, types, structures):
    """Fill sfval_* of every potential from unscaled descriptors of the
    atoms of its species in the calibration structures."""
    with tempfile.TemporaryDirectory() as d:
        files = {}
        for t in types:
            files[t] = os.path.join(d, t + ".nn")
            nnfile.write_nn(files[t], pots[t])
        orc = Oracle(types, files)
        vals = {t: [] for t in types}
        for s in structures:
            nb = orc.neighbors(s.cell, s.pbc, s.types, s.positions)
            L = nb["latt"].T  # columns = lattice vectors
            for i in range(len(s)):
                coo_i = L @ nb["frac"][i]
                # recompute the centre exactly as the neighbor list frames it
                n = nb["nnb"][i]
                G = orc.descriptors(int(s.types[i]), coo_i, nb["coo"][i, :n], nb["type"][i, :n],
                                    scaled=False)
                vals[types[s.types[i] - 1]].append(G)
    for t in types:
        G = np.array(vals[t])
        p = pots[t]
        p["sfval_min"] = G.min(axis=0)
        p["sfval_max"] = G.max(axis=0)
        p["sfval_avg"] = G.mean(axis=0)
        p["sfval_cov"] = (G * G).mean(axis=0)
        p["neval"] = len(G)


def write_set(name, types, pots):
    paths = {}
    for t in types:
        path = os.path.join(OUT, "%s.%s.nn" % (name, t))
        nnfile.write_nn(path, pots[t])
        paths[t] = path
    return paths


def cheb_set(name, types, e_atoms, rr, rn, ar, an, hidden, seed0, calib):
    pots = {}
    for k, t in enumerate(types):
        st = nnfile.chebyshev_setup(t, types, rr, rn, ar, an)
        arch = [st["nsf"]] + list(hidden) + [1]
        pots[t] = make_pot(t, st, arch, seed0 + k, types, e_atoms)
    calibrate(pots, types, calib)
    return write_set(name, types, pots)


def bp_functions(types):
    fns = []
    etas2 = np.geomspace(0.003, 0.5, 8)
    for t2 in types:
        for eta in etas2:
            fns.append({"G": 2, "type2": t2, "Rc": 6.0, "Rs": 0.0, "eta": float(eta)})
    for a in range(len(types)):
        for b in range(a, len(types)):
            for zeta in (1.0, 2.0, 4.0):
                for lam in (-1.0, 1.0):
                    for eta in (0.0003, 0.03, 0.09):
                        fns.append({"G": 4, "type2": types[a], "type3": types[b], "Rc": 6.0,
                                    "lambda": lam, "zeta": zeta, "eta": eta})
    return fns


def mixed_functions(types):
    """Behler-Parrinello functions of every kind (G1-G5) with mixed cutoffs:
    G2 with Rs = 0.5 at 4.5 and 6.0 A, G3 at 5.5 A, and G4/G5 of one type
    pair at 4.5, 5.0 and 6.0 A."""
    fns = []
    for t2 in types:
        fns.append({"G": 1, "type2": t2, "Rc": 5.0})
        for eta, rc in ((0.01, 6.0), (0.1, 6.0), (0.05, 4.5)):
            fns.append({"G": 2, "type2": t2, "Rc": rc, "Rs": 0.5, "eta": eta})
        fns.append({"G": 3, "type2": t2, "Rc": 5.5, "kappa": 1.3})
    for a in range(len(types)):
        for b in range(a, len(types)):
            for lam, zeta, eta, rc in ((1.0, 1.0, 0.003, 6.0), (-1.0, 2.0, 0.01, 6.0),
                                       (1.0, 4.0, 0.02, 4.5), (-1.0, 1.0, 0.005, 4.5)):
                fns.append({"G": 4, "type2": types[a], "type3": types[b], "Rc": rc,
                            "lambda": lam, "zeta": zeta, "eta": eta})
            for lam, zeta, eta, rc in ((1.0, 2.0, 0.01, 5.0), (-1.0, 1.0, 0.002, 6.0)):
                fns.append({"G": 5, "type2": types[a], "type3": types[b], "Rc": rc,
                            "lambda": lam, "zeta": zeta, "eta": eta})
    return fns


def mixed_structures():
    """The two structures the mixed set is checked on: a 108-atom binary
    bulk cell and a 40-atom isolated cluster."""
    bulk = synth.binary_fcc(3.7, (3, 3, 3), 0.1, 41, 42)
    bulk.label = "bulk108"
    _, big = synth.fcc(3.6, (6, 6, 6))
    big = big - big.mean(axis=0)
    pos = big[np.argsort(np.sum(big * big, axis=1), kind="stable")[:40]]
    rng = np.random.Generator(np.random.PCG64(43))
    pos = pos + rng.normal(0.0, 0.1, size=pos.shape)
    cluster = synth.Structure(np.eye(3), pos, rng.integers(1, 3, len(pos)).astype(np.int32), False,
                              "cluster40")
    return {"bulk108": bulk, "cluster40": cluster}


def mixed_set():
    types = ["A", "B"]
    fns = mixed_functions(types)
    pots = {}
    for k, t in enumerate(types):
        st = nnfile.behler_setup(t, types, fns)
        pots[t] = make_pot(t, st, [st["nsf"], 12, 1], 60 + k, types, [-1000.0, -1500.0])
    calibrate(pots, types, list(mixed_structures().values()))
    return write_set("mixed", types, pots)


def chebhi_set():
    """A/B with typespin, radial order 10 / 6.0 A and angular order 12 /
    4.0 A (48 features), 48-16-16-1: angular orders above 7 take the atom
    kernel's (16, 16) register layout."""
    cal = [synth.binary_fcc(3.61, (3, 3, 3), 0.1, 17, 18)]
    return cheb_set("chebhi", ["A", "B"], [-1000.0, -1500.0], 6.0, 10, 4.0, 12, (16, 16), 70, cal)


def chebwide_set():
    """A/B with typespin, radial order 8 / 5.0 A and angular order 4 / 3.5 A
    (28 features), 28-100-80-1: hidden layers wider than the 64 lanes of a
    wavefront."""
    cal = [synth.binary_fcc(3.61, (3, 3, 3), 0.1, 21, 22)]
    return cheb_set("chebwide", ["A", "B"], [-1000.0, -1500.0], 5.0, 8, 3.5, 4, (100, 80), 80, cal)


def main():
    os.makedirs(OUT, exist_ok=True)
    if sys.argv[1:] == ["mixed"]:
        mixed_set()
        print("wrote the mixed set to", OUT)
        return
    if sys.argv[1:] == ["chebhi"]:
        chebhi_set()
        print("wrote the chebhi set to", OUT)
        return
    if sys.argv[1:] == ["chebwide"]:
        chebwide_set()
        print("wrote the chebwide set to", OUT)
        return
    # fcc108: single species, Chebyshev radial 16 / 6.0 A, angular 4 / 4.0 A, 22-20-20-1
    c, p = synth.fcc(3.61, (3, 3, 3))
    cal = [synth.Structure(c, synth.displaced(p, 0.1, 3), np.ones(len(p), np.int32))]
    cheb_set("fcc108", ["Cu"], [-1000.0], 6.0, 16, 4.0, 4, (20, 20), 2, cal)
    # alloy: A/B with typespin, radial 18 / 6.5 A, angular 6 / 4.5 A, 52-30-30-1
    cal = [synth.binary_fcc(3.8, (5, 5, 5), 0.08, 13, 14)]
    cheb_set("alloy", ["A", "B"], [-1000.0, -1500.0], 6.5, 18, 4.5, 6, (30, 30), 4, cal)
    # teacher for Monte Carlo: fcc108 descriptor shapes for A/B (44 features), 44-20-20-1
    cal = [synth.binary_fcc(3.61, (4, 4, 4), 0.1, 15, 16)]
    cheb_set("teacher", ["A", "B"], [-1000.0, -1500.0], 6.0, 16, 4.0, 4, (20, 20), 6, cal)
    # student: radial-only Chebyshev order 8 / 5.0 A (angular order 0, Rc 0), 20-8-1
    cheb_set("student", ["A", "B"], [-1000.0, -1500.0], 5.0, 8, 0.0, 0, (8,), 30, cal)
    # Behler-Parrinello, 3 species: 8 G2 per species + 18 G4 per species pair = 132
    types = ["A", "B", "C"]
    pots = {}
    fns = bp_functions(types)
    for k, t in enumerate(types):
        st = nnfile.behler_setup(t, types, fns)
        pots[t] = make_pot(t, st, [st["nsf"], 25, 25, 1], 8 + k, types, [-1000.0, -1500.0, -800.0])
    cal = []
    for seed in (24, 25):
        c, p = synth.fcc(3.6, (3, 3, 3))
        rng = np.random.Generator(np.random.PCG64(seed))
        cal.append(synth.Structure(c, synth.displaced(p, 0.1, seed + 10), rng.integers(1, 4, len(p))))
    calibrate(pots, types, cal)
    write_set("bp3", types, pots)
    # Behler-Parrinello with every function kind and mixed cutoffs, 2 species
    mixed_set()
    # Chebyshev with an angular order above the radial one
    chebhi_set()
    # Chebyshev with hidden layers wider than a wavefront
    chebwide_set()
    print("wrote potentials to", OUT)


if __name__ == "__main__":
    main()
