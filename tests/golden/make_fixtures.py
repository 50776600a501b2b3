"""Record golden vectors from the reference library built from its own
sources (oracle/_ref/libaenet_ref.so; build with `make -C oracle ref`).

Run from the repository root:  python tests/golden/make_fixtures.py

Each fixture holds the inputs (cell, pbc, types, Cartesian positions) and
the reference's outputs: per-atom energies, forces, total energy and the
full neighbor lists (lcl_nbdist_cart).  They pin the oracle and the GPU
path where the reference itself is not available (e.g. on the GPU box).
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

from aenet_amd import synth  # noqa: E402
from conftest import TI2O4_CELL, TI2O4_POS, TI2O4_TYPES, TI_O, pot_files  # noqa: E402
from oracle_lib import Reference  # noqa: E402

OUT = os.path.join(ROOT, "tests", "golden")


def cases():
    b = synth.batch_uneven(40)
    return [
        ("fcc108", ["Cu"], synth.c1_fcc108()),
        ("alloy", ["A", "B"], synth.binary_fcc(3.8, (4, 4, 4), 0.08, 1, 2)),
        ("bp3", ["A", "B", "C"], b[0]),
        ("bp3", ["A", "B", "C"], b[3]),
        ("bp3", ["A", "B", "C"], b[4]),
        ("bp3", ["A", "B", "C"], b[10]),
        ("teacher", ["A", "B"], synth.binary_fcc(3.61, (3, 3, 3), 0.1, 5, 6)),
        ("student", ["A", "B"], synth.binary_fcc(3.61, (3, 3, 3), 0.1, 5, 6)),
    ]


def record(path, name, types, s):
    r = Reference(types, pot_files(name, types))
    e = r.structure(s.cell, s.pbc, s.types, s.positions)
    nb = r.neighbors(s.cell, s.pbc, s.types, s.positions)
    m = int(nb["nnb"].max())
    np.savez_compressed(path, potential=name, types=np.array(types), cell=s.cell, pbc=s.pbc,
                        atom_types=s.types, positions=s.positions, Etot=e["Etot"], Ecoh=e["Ecoh"],
                        energies=e["energies"], forces=e["forces"], nnb=nb["nnb"], frac=nb["frac"],
                        latt=nb["latt"], nb_coo=nb["coo"][:, :m], nb_dist=nb["dist"][:, :m],
                        nb_list=nb["list"][:, :m], nb_type=nb["type"][:, :m])


def main():
    if not Reference.available():
        raise SystemExit("build the reference first: make -C oracle ref")
    r = Reference(["Ti", "O"], TI_O)
    e = r.structure(TI2O4_CELL, True, TI2O4_TYPES, TI2O4_POS)
    np.savez_compressed(os.path.join(OUT, "ti2o4_reference.npz"), Etot=e["Etot"], energies=e["energies"],
                        forces=e["forces"], golden_test_ase=-4990.310783539421)
    for k, (name, types, s) in enumerate(cases()):
        label = s.label or name
        path = os.path.join(OUT, "ref_%02d_%s_%s.npz" % (k, name, label))
        record(path, name, types, s)
        print("wrote", path)


if __name__ == "__main__":
    main()
