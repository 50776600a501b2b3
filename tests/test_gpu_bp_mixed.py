"""Behler-Parrinello potentials with every function kind (G1-G5) and mixed
cutoffs -- functions of one neighbor type, or of one type pair, at different
radii, G2 with Rs = 0.5 -- against the reference build's numbers (fixtures
ref_08_mixed_bulk108 / ref_09_mixed_cluster40, recorded from the reference's
sf_fingerprint, src/ext/symmfunc.f90:311-431, by tests/golden/make_fixtures.py)
and against the FP64 oracle, which equals the reference bit for bit on them
(tests/test_oracle.py).

The atom kernel tabulates each neighbor's cutoff at the radius of the first
angular function of the central species and computes the others per triplet;
these potentials take both routes, G5 blocks skip the r_jk filter, and
segments of one kind and cutoff are flushed in between.  Bars: per-atom
energies within 1e-5 eV/atom, forces within 1e-4 eV/A.
"""
import os

import numpy as np
import pytest

from conftest import pot_files

pytestmark = pytest.mark.gpu

E_TOL = 1e-5
F_TOL = 1e-4
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


@pytest.mark.parametrize("name", ["ref_08_mixed_bulk108", "ref_09_mixed_cluster40"])
def test_bp_mixed_kinds_and_cutoffs(name, device):
    from aenet_amd.core import ANNPotentials
    from oracle_lib import Oracle
    types = ["A", "B"]
    files = pot_files("mixed", types)
    g = np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False)
    print("case", name, len(g["atom_types"]), "atoms", flush=True)
    cell, pbc, pos, at = g["cell"], bool(g["pbc"]), g["positions"], g["atom_types"]
    orc = Oracle(types, files).structure(cell, pbc, at, pos)
    assert np.array_equal(orc["energies"], g["energies"])   # the oracle is the reference here
    ann = ANNPotentials(files)
    try:
        r = ann.structure_energy_and_forces(cell, pbc, pos, list(at))
    finally:
        ann.release()
    dE = np.abs(r["energies"] - g["energies"]).max()
    dF = np.abs(r["forces"] - g["forces"]).max()
    print("max |dE| %.3e eV/atom  max |dF| %.3e eV/A  (|F| up to %.2f)" % (dE, dF, np.abs(g["forces"]).max()))
    assert np.abs(g["forces"]).max() > 0.1  # the forces are not trivially zero
    assert dE < E_TOL
    assert dF < F_TOL


def _bp3_variant(tmp_path, kind):
    """bp3 potentials changed so that the kernel's angular fast path (blocks
    of one cutoff, one kind and lambda = +-1, functions sign-sorted) meets
    the cases bp3 itself does not hold: 'g5' turns every G4 into G5 (no r_jk
    cutoff, S = r_ij^2 + r_ik^2); 'signs' gives the (A, A) block lambda = -1
    only and the (B, B) block lambda = +1 only.  bp3's blocks are complete
    grids of (lambda, zeta) pairs and etas, which the kernel evaluates by
    factors (BPAngGrid); 'signs' makes two blocks repeat functions, and
    'nogrid' moves the first G4 of every block to an eta of its own, so
    that those blocks take the per-function fast path."""
    from aenet_amd.nnfile import read_nn, write_nn
    files = {}
    for t in "ABC":
        p = read_nn(pot_files("bp3", "ABC")[t])
        sf = np.array(p["sf"])
        par = np.array(p["sfparam"], dtype=np.float64).reshape(int(p["nsf"]), -1)
        env = np.array(p["sfenv"]).reshape(int(p["nsf"]), 2)
        if kind == "g5":
            sf[sf == 4] = 5
        elif kind == "nogrid":
            seen = set()
            for i in range(len(sf)):
                if sf[i] == 4 and tuple(sorted(env[i])) not in seen:
                    seen.add(tuple(sorted(env[i])))
                    par[i, 3] = 0.05
        else:
            for i in range(len(sf)):
                if sf[i] == 4 and tuple(env[i]) == (1, 1):
                    par[i, 1] = -1.0
                elif sf[i] == 4 and tuple(env[i]) == (2, 2):
                    par[i, 1] = 1.0
        p["sf"] = sf
        p["sfparam"] = par
        path = str(tmp_path / ("%s.%s.nn" % (kind, t)))
        write_nn(path, p)
        files[t] = path
    return files


@pytest.mark.parametrize("kind", ["g5", "signs", "nogrid"])
def test_bp_angular_fast_path_variants(kind, tmp_path, device):
    """Energies and forces of the fast-path variants against the FP64
    oracle: a displaced 108-atom FCC cell of three species and a 40-atom
    cluster.  Bars as above."""
    from aenet_amd import synth
    from aenet_amd.core import ANNPotentials
    from oracle_lib import Oracle
    types = ["A", "B", "C"]
    files = _bp3_variant(tmp_path, kind)
    rng = np.random.Generator(np.random.PCG64(7))
    cell, pos = synth.fcc(3.6, (3, 3, 3))
    cases = [(cell, True, synth.displaced(pos, 0.1, 8))]
    big_cell, big = synth.fcc(3.6, (4, 4, 4))
    big = big - big.mean(axis=0)
    order = np.argsort(np.sum(big * big, axis=1), kind="stable")
    cases.append((np.eye(3) * 30.0, False, synth.displaced(big[order[:40]], 0.1, 9)))
    orc = Oracle(types, files)
    ann = ANNPotentials(files)
    try:
        for cell_, pbc, p in cases:
            at = rng.integers(1, 4, len(p)).astype(np.int32)
            print("case", kind, len(p), "atoms pbc", pbc, flush=True)
            ref = orc.structure(cell_, pbc, at, p)
            r = ann.structure_energy_and_forces(cell_, pbc, p, [types[i - 1] for i in at])
            dE = np.abs(r["energies"] - ref["energies"]).max()
            dF = np.abs(r["forces"] - ref["forces"]).max()
            print("max |dE| %.3e eV/atom  max |dF| %.3e eV/A  (|F| up to %.2f)" % (dE, dF, np.abs(ref["forces"]).max()))
            assert np.abs(ref["forces"]).max() > 0.1
            assert dE < E_TOL
            assert dF < F_TOL
    finally:
        ann.release()
