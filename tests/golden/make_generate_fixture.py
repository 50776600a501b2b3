"""Reference training-set files for aenet_amd.generate: the reference's own
generate.x (built from its sources by ``make -C oracle ref``) run on seeded
inputs that ``write_inputs`` writes (XSF structures with energies and
forces, .stp set-ups, generate.in).

Run from the repository root:  python tests/golden/make_generate_fixture.py

Writes tests/golden/generate/{cheb,bp}.train (and cheb.train.forces, the
descriptor derivatives of FORCES).  The GPU test writes the same inputs and
compares aenet_amd.generate's output with these files.
"""
import os
import shutil
import subprocess
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))

from aenet_amd import synth, trainfile, xsf  # noqa: E402

OUT = os.path.join(ROOT, "tests", "golden", "generate")
GEN = os.path.join(ROOT, "oracle", "_ref", "generate.x")

CHEB_STP = """DESCR
  synthetic Chebyshev set-up for %s
  radial and angular terms with typespin
END DESCR

ATOM %s

ENV 2
A
B

RMIN 0.75d0

BASIS type=Chebyshev
radial_Rc = %s  radial_N = %d angular_Rc = %s  angular_N = %d
"""


def bp_stp(t):
    from make_potentials import mixed_functions
    lines = ["DESCR", "  synthetic Behler-Parrinello set-up for %s (G1-G5)" % t, "END DESCR", "",
             "ATOM %s" % t, "", "ENV 2", "A", "B", "", "RMIN 0.75d0", "", "SYMMFUNC type=Behler2011"]
    fns = mixed_functions(["A", "B"])
    lines.append("%d" % len(fns))
    for fn in fns:
        g = fn["G"]
        s = "G=%d type2=%s" % (g, fn["type2"])
        if g in (4, 5):
            s += " type3=%s" % fn["type3"]
        if g == 2:
            s += " eta=%.6f Rs=%.4f" % (fn["eta"], fn["Rs"])
        elif g == 3:
            s += " kappa=%.4f" % fn["kappa"]
        elif g in (4, 5):
            s += " eta=%.6f lambda=%.1f zeta=%.1f" % (fn["eta"], fn["lambda"], fn["zeta"])
        s += " Rc=%.4f" % fn["Rc"]
        lines.append(s)
    return "\n".join(lines) + "\n"


def structures(seed):
    """A periodic 32-atom cell, a slab and an isolated 14-atom cluster with
    seeded (unphysical) energies and forces; names A/B."""
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    s = synth.binary_fcc(3.8, (2, 2, 2), 0.08, seed + 1, seed + 2)
    out.append((s.cell, s.positions, s.types))
    cell, pos = synth.fcc(3.7, (2, 2, 1))
    cell = cell.copy()
    cell[2, 2] += 10.0
    pos = pos + rng.normal(0, 0.08, pos.shape)
    out.append((cell, pos, rng.integers(1, 3, len(pos))))
    _, big = synth.fcc(3.6, (3, 3, 3))
    big = big - big.mean(axis=0)
    cl = big[np.argsort(np.sum(big * big, axis=1), kind="stable")[:14]] + rng.normal(0, 0.1, (14, 3))
    out.append((None, cl, rng.integers(1, 3, 14)))
    res = []
    for cell, pos, t in out:
        n = len(pos)
        E = -1000.0 * n + rng.normal(0, 3.0)
        F = rng.normal(0, 0.5, (n, 3))
        res.append((cell, pos, np.asarray(t), E, F))
    return res


def write_inputs(d, kind):
    """Inputs of one run in directory d; returns the generate.in path."""
    names = ["A", "B"]
    for t in names:
        with open(os.path.join(d, "%s.stp" % t), "w") as f:
            if kind == "cheb":
                f.write(CHEB_STP % (t, t, "6.5", 12, "4.0", 4))
            else:
                f.write(bp_stp(t))
    files = []
    for k, (cell, pos, t, E, F) in enumerate(structures(70 if kind == "cheb" else 80)):
        p = "s%02d.xsf" % k
        xsf.write_xsf(os.path.join(d, p), [names[i - 1] for i in t], pos, cell=cell, forces=F, E_tot=E)
        files.append(p)
    gin = os.path.join(d, "generate.in")
    with open(gin, "w") as f:
        f.write("OUTPUT %s.train\n\nTYPES\n2\nA -1000.5  | eV\nB -1500.25 | eV\n\n" % kind)
        f.write("SETUPS\nA A.stp\nB B.stp\n\n")
        if kind == "cheb":
            f.write("FORCES\nFORCESPERCENT 100\n\n")
        f.write("FILES\n%d\n" % len(files))
        for p in files:
            f.write("%s\n" % p)
    return gin


def main():
    if not os.path.exists(GEN):
        raise SystemExit("build the reference first: make -C oracle ref")
    os.makedirs(OUT, exist_ok=True)
    for kind in ("cheb", "bp"):
        with tempfile.TemporaryDirectory() as d:
            gin = write_inputs(d, kind)
            r = subprocess.run([GEN, os.path.basename(gin)], cwd=d, capture_output=True, text=True)
            if r.returncode != 0:
                raise SystemExit(r.stdout[-2000:] + r.stderr[-2000:])
            src = os.path.join(d, kind + ".train")
            shutil.copy(src, os.path.join(OUT, kind + ".train"))
            print("wrote", os.path.join(OUT, kind + ".train"), os.path.getsize(src), "bytes")
            fsrc = os.path.join(d, kind + ".train.forces")
            if os.path.exists(fsrc):
                # the derivative file is large: keep the records of a few atoms
                f = trainfile.read_forces(fsrc)
                keep = {"nstrucs": f["nstrucs"], "max_nnb": f["max_nnb"]}
                for k, s in enumerate(f["structures"]):
                    keep["s%d_meta" % k] = np.array([s["natoms"], s["ntypes"], s["selected"]])
                    for i in (0, 5, s["natoms"] - 1):
                        a = s["atoms"][i]
                        keep["s%d_a%d_meta" % (k, i)] = np.array([a["itype"], a["nsf"], a["nnb"]])
                        keep["s%d_a%d_nblist" % (k, i)] = a["nblist"]
                        keep["s%d_a%d_deriv_i" % (k, i)] = a["deriv_i"]
                        keep["s%d_a%d_deriv_j" % (k, i)] = a["deriv_j"]
                path = os.path.join(OUT, kind + "_forces_subset.npz")
                np.savez_compressed(path, **keep)
                print("wrote", path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
