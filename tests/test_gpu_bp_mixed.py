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
