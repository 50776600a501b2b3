"""Record golden vectors from the reference library built from its own
sources (oracle/_ref/libaenet_ref.so; build with `make -C oracle ref`).

Run from the repository root:  python tests/golden/make_fixtures.py [names]
(names: ``alloy4000``, ``sfb`` and/or two-digit fixture numbers; default all)

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
sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))

from aenet_amd import synth  # noqa: E402
from conftest import TI2O4_CELL, TI2O4_POS, TI2O4_TYPES, TI_O, pot_files  # noqa: E402
from oracle_lib import Reference  # noqa: E402
from make_potentials import mixed_structures  # noqa: E402

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
        ("mixed", ["A", "B"], mixed_structures()["bulk108"]),
        ("mixed", ["A", "B"], mixed_structures()["cluster40"]),
        ("chebhi", ["A", "B"], synth.binary_fcc(3.61, (3, 3, 3), 0.1, 19, 20)),
        ("chebwide", ["A", "B"], synth.binary_fcc(3.61, (3, 3, 3), 0.1, 23, 24)),
    ]


def record(path, name, types, s, nb_every=1):
    """Energies and forces of every atom; neighbor lists of every
    ``nb_every``-th atom (``nb_atoms`` holds their 0-based indices)."""
    r = Reference(types, pot_files(name, types))
    e = r.structure(s.cell, s.pbc, s.types, s.positions)
    nb = r.neighbors(s.cell, s.pbc, s.types, s.positions)
    sel = np.arange(0, len(s), nb_every)
    m = int(nb["nnb"].max())
    np.savez_compressed(path, potential=name, types=np.array(types), cell=s.cell, pbc=s.pbc,
                        atom_types=s.types, positions=s.positions, Etot=e["Etot"], Ecoh=e["Ecoh"],
                        energies=e["energies"], forces=e["forces"], nnb=nb["nnb"], frac=nb["frac"],
                        latt=nb["latt"], nb_atoms=sel, nb_coo=nb["coo"][sel, :m],
                        nb_dist=nb["dist"][sel, :m], nb_list=nb["list"][sel, :m],
                        nb_type=nb["type"][sel, :m])


def record_sfb(path):
    """aenet_sfb_eval and aenet_sfb_reconstruct_radial of the reference
    (src/aenet.f90:681-799) on one seeded two-species environment."""
    import ctypes as C
    lib = Reference.lib()
    st = C.c_int(0)
    lib.aenet_sfb_init(2, (C.c_char_p * 2)(b"A", b"B"), 18, 6, 6.5, 4.5, C.byref(st))
    assert st.value == 0
    rng = np.random.Generator(np.random.PCG64(11))
    coo = rng.uniform(-4, 4, size=(40, 3))
    coo = np.ascontiguousarray(coo[np.linalg.norm(coo, axis=1) > 1.5])
    t = rng.integers(1, 3, len(coo)).astype(np.int32)
    nv = lib.aenet_sfb_nvalues()
    vals = np.zeros(nv)
    lib.aenet_sfb_eval(1, np.zeros(3), len(t), t, coo.reshape(-1), nv, vals, C.byref(st))
    assert st.value == 0
    nx = 41
    x = np.zeros(nx)
    y = np.zeros(nx)
    lib.aenet_sfb_reconstruct_radial(nv, vals, nx, x, y, C.byref(st))
    assert st.value == 0
    lib.aenet_sfb_final(C.byref(st))
    np.savez_compressed(path, coo=coo, types=t, values=vals, nx=nx, x=x, y=y, radial_order=18,
                        angular_order=6, radial_Rc=6.5, angular_Rc=4.5)


def main():
    if not Reference.available():
        raise SystemExit("build the reference first: make -C oracle ref")
    if sys.argv[1:]:
        # only the named fixtures, e.g. ``alloy4000`` or ``08 09``
        want = sys.argv[1:]
    else:
        want = None
    if want is None or "alloy4000" in want:
        # the 4000-atom alloy at full size; lists of every 40th atom
        path = os.path.join(OUT, "full_alloy4000.npz")
        record(path, "alloy", ["A", "B"], synth.c2_alloy4000(), nb_every=40)
        print("wrote", path)
        if want is not None:
            want = [w for w in want if w != "alloy4000"]
            if not want:
                return
    if want is None or "sfb" in want:
        path = os.path.join(OUT, "sfb_reference.npz")
        record_sfb(path)
        print("wrote", path)
        if want is not None:
            want = [w for w in want if w != "sfb"]
            if not want:
                return
    if want is None:
        r = Reference(["Ti", "O"], TI_O)
        e = r.structure(TI2O4_CELL, True, TI2O4_TYPES, TI2O4_POS)
        np.savez_compressed(os.path.join(OUT, "ti2o4_reference.npz"), Etot=e["Etot"], energies=e["energies"],
                            forces=e["forces"], golden_test_ase=-4990.310783539421)
    for k, (name, types, s) in enumerate(cases()):
        if want is not None and "%02d" % k not in want:
            continue
        label = s.label or name
        path = os.path.join(OUT, "ref_%02d_%s_%s.npz" % (k, name, label))
        record(path, name, types, s)
        print("wrote", path)


if __name__ == "__main__":
    main()
