"""Behler-Parrinello potentials with every function kind (G1-G5) and mixed
cutoffs -- functions of one neighbor type, or of one type pair, at different
radii -- against the FP64 oracle (sf_fingerprint and its G1-G5 updates,
src/ext/symmfunc.f90:311-431, restated in oracle/aenet_oracle.c).

The atom kernel tabulates each neighbor's cutoff at the radius of the first
angular function of the central species and computes the others per triplet;
these potentials take both routes, G5 blocks skip the r_jk filter, and
segments of one kind and cutoff are flushed in between.  Bars: per-atom
energies within 1e-5 eV/atom, forces within 1e-4 eV/A.
"""
import os
import sys
import tempfile

import numpy as np
import pytest

from conftest import ROOT

pytestmark = pytest.mark.gpu

sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))

E_TOL = 1e-5
F_TOL = 1e-4


def _functions(types):
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


def _structures():
    from aenet_amd import synth
    bulk = synth.binary_fcc(3.7, (3, 3, 3), 0.1, 41, 42)
    _, big = synth.fcc(3.6, (6, 6, 6))
    big = big - big.mean(axis=0)
    pos = big[np.argsort(np.sum(big * big, axis=1), kind="stable")[:40]]
    rng = np.random.Generator(np.random.PCG64(43))
    pos = pos + rng.normal(0.0, 0.1, size=pos.shape)
    cluster = synth.Structure(np.eye(3), pos, rng.integers(1, 3, len(pos)).astype(np.int32), False)
    return {"bulk108": bulk, "cluster40": cluster}


@pytest.fixture(scope="module")
def mixed_set():
    import make_potentials as mp
    from aenet_amd import nnfile
    types = ["A", "B"]
    fns = _functions(types)
    pots = {}
    for k, t in enumerate(types):
        st = nnfile.behler_setup(t, types, fns)
        pots[t] = mp.make_pot(t, st, [st["nsf"], 12, 1], 60 + k, types, [-1000.0, -1500.0])
    mp.calibrate(pots, types, list(_structures().values()))
    with tempfile.TemporaryDirectory() as d:
        files = {}
        for t in types:
            files[t] = os.path.join(d, "mixed.%s.nn" % t)
            nnfile.write_nn(files[t], pots[t])
        yield types, files


@pytest.mark.parametrize("name", ["bulk108", "cluster40"])
def test_bp_mixed_kinds_and_cutoffs(mixed_set, name, device):
    from aenet_amd.core import ANNPotentials
    from oracle_lib import Oracle
    types, files = mixed_set
    s = _structures()[name]
    print("case", name, len(s.types), "atoms", flush=True)
    ref = Oracle(types, files).structure(s.cell, s.pbc, s.types, s.positions)
    ann = ANNPotentials(files)
    try:
        r = ann.structure_energy_and_forces(s.cell, s.pbc, s.positions, list(s.types))
    finally:
        ann.release()
    dE = np.abs(r["energies"] - ref["energies"]).max()
    dF = np.abs(r["forces"] - ref["forces"]).max()
    print("max |dE| %.3e eV/atom  max |dF| %.3e eV/A  (|F| up to %.2f)" % (dE, dF, np.abs(ref["forces"]).max()))
    assert np.abs(ref["forces"]).max() > 0.1  # the forces are not trivially zero
    assert dE < E_TOL
    assert dF < F_TOL
