"""predict.x, the batched command-line driver (aenet_amd/csrc/predict.cpp):
predict.in keywords (src/input.f90:358-400), XSF input
(src/ext/xsflib.f90:109-208), the printed blocks (src/predict.F90:439-568)
and the RELAX loop (src/predict.F90:234-286).

- CPU: the driver's block printer against the reference's own recorded
  output of unittests/Cu-dimer-forces (tests/golden/predict_cu_dimer_blocks.txt),
  line for line;
- GPU: every structure of a FILES list in one run, the printed per-atom
  energies, forces and structure energies against the reference build's
  fixtures; a relaxation of an isolated cluster ends below F_conv by the
  oracle's own forces at the printed coordinates.
"""
import os
import re
import subprocess

import numpy as np
import pytest

from conftest import ROOT, pot_files

GOLDEN = os.path.join(ROOT, "tests", "golden")
EXE = os.path.join(ROOT, "aenet_amd", "_lib", "predict.x")


def test_printed_blocks_match_reference_output(tmp_path):
    exe = tmp_path / "fmt"
    subprocess.run(["g++", "-std=c++17", "-I" + os.path.join(ROOT, "aenet_amd", "csrc"),
                    os.path.join(ROOT, "tests", "predict_format_check.cpp"), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    ref = open(os.path.join(GOLDEN, "predict_cu_dimer_blocks.txt")).read()
    assert out.splitlines() == ref.splitlines()


def test_input_errors_are_reported(tmp_path):
    """A predict.in without TYPES is refused before any GPU work."""
    if not os.path.exists(EXE):
        pytest.skip("predict.x not built")
    f = tmp_path / "predict.in"
    f.write_text("NETWORKS\nCu Cu.nn\n")
    r = subprocess.run([EXE, str(f)], capture_output=True, text=True)
    assert r.returncode != 0 and "no atom TYPES" in r.stderr


def write_xsf(path, names, s):
    with open(path, "w") as f:
        if s.pbc:
            f.write("CRYSTAL\nPRIMVEC\n")
            for v in s.cell:
                f.write(" %.17g %.17g %.17g\n" % tuple(v))
            f.write("PRIMCOORD\n %d 1\n" % len(s.types))
        else:
            f.write("ATOMS\n")
        for t, x in zip(s.types, s.positions):
            f.write("%s %.17g %.17g %.17g\n" % ((names[t - 1],) + tuple(x)))


def write_input(path, names, files, structs, extra="", verbosity="low"):
    with open(path, "w") as f:
        f.write("TYPES\n%d\n" % len(names))
        for n in names:
            f.write("%s\n" % n)
        f.write("\nNETWORKS\n")
        for n in names:
            f.write("%s %s\n" % (n, files[n]))
        f.write("\nFORCES\nPRINT_ATOMIC_ENERGIES\nVERBOSITY %s\n%s\nFILES\n%d\n" % (verbosity, extra, len(structs)))
        for s in structs:
            f.write("%s\n" % s)


def parse_output(text):
    """Per structure: coordinates, forces, atomic energies of the first
    coordinate block, the final coordinate block, and the energies."""
    out = []
    blocks = text.split(" Cohesive energy")
    for b in blocks[:-1]:
        rows = re.findall(r"^ (\S{1,2})\s+(" + r"\s+".join([r"(-?\d+\.\d+)"] * 7) + r")\s*$", b, re.M)
        rows6 = re.findall(r"^ (\S{1,2})" + r"\s+(-?\d+\.\d+)" * 6 + r"\s*$", b, re.M)
        out.append({"rows7": [[float(v) for v in r[1].split()] for r in rows],
                    "rows6": [[float(v) for v in r[1:]] for r in rows6]})
    for k, b in enumerate(blocks[1:]):
        m = re.search(r":\s+(-?\d+\.\d+) eV\n Total energy\s+:\s+(-?\d+\.\d+) eV", " Cohesive energy" + b[:200])
        out[k]["Ecoh"], out[k]["Etot"] = float(m.group(1)), float(m.group(2))
    return out


@pytest.mark.gpu
def test_predict_batch_matches_reference(tmp_path, device):
    fx = ["ref_02_bp3_fcc", "ref_03_bp3_fcc", "ref_04_bp3_slab", "ref_05_bp3_cluster"]
    names = ["A", "B", "C"]
    paths, gs = [], []
    for k, name in enumerate(fx):
        g = np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False)
        gs.append(g)

        class S:
            pass
        s = S()
        s.cell, s.pbc, s.types, s.positions = g["cell"], bool(g["pbc"]), g["atom_types"], g["positions"]
        p = str(tmp_path / ("s%02d.xsf" % k))
        write_xsf(p, names, s)
        paths.append(p)
    inp = str(tmp_path / "predict.in")
    write_input(inp, names, pot_files("bp3", names), paths)
    print("case predict.x on", fx, flush=True)
    r = subprocess.run([EXE, inp], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    res = parse_output(r.stdout)
    assert len(res) == len(fx)
    for g, o in zip(gs, res):
        rows = np.array(o["rows7"])
        n = len(g["atom_types"])
        assert rows.shape == (n, 7)
        assert np.abs(rows[:, 3:6] - g["forces"]).max() < 1e-4 + 5e-7
        assert np.abs(rows[:, 6] - g["energies"]).max() < 1e-5 + 5e-7
        if bool(g["pbc"]):
            # printed positions are the wrapped lattice positions
            d = rows[:, :3] - g["positions"]
            frac = d @ np.linalg.inv(g["cell"])
            assert np.abs(frac - np.rint(frac)).max() < 1e-6
        else:
            assert np.abs(rows[:, :3] - g["positions"]).max() < 5e-7 + 1e-9
        assert abs(o["Etot"] - g["Etot"]) < 1e-5 * n + 5e-9
        assert abs(o["Ecoh"] - g["Ecoh"]) < 1e-5 * n + 5e-9


@pytest.mark.gpu
def test_predict_relax_ti2o4(tmp_path, device):
    """predict.x RELAX (L-BFGS, steps capped at 0.1 A as opt_optimize_coords'
    dmax, src/predict.F90:262-263) on the Ti2O4 rutile cell of
    python3/aenet/tests/test_ase.py with its atoms displaced by up to 0.15 A,
    with the trained Ti/O potentials of the reference's examples.  It
    reports convergence (src/predict.F90:269-274), the energy went down, and
    at the printed coordinates the oracle's forces are within F_conv (plus
    the per-atom bar) and its energy equals the printed one."""
    from oracle_lib import Oracle
    from conftest import TI2O4_CELL, TI2O4_POS, TI2O4_TYPES, TI_O
    names = ["Ti", "O"]
    rng = np.random.default_rng(7)

    class S:
        pass
    s = S()
    s.cell, s.pbc, s.types = TI2O4_CELL, True, TI2O4_TYPES
    s.positions = TI2O4_POS + rng.uniform(-0.15, 0.15, TI2O4_POS.shape)
    p = str(tmp_path / "ti2o4.xsf")
    write_xsf(p, names, s)
    inp = str(tmp_path / "predict.in")
    write_input(inp, names, TI_O, [p], extra="RELAX\nmethod=bfgs F_conv=0.01 steps=400\n", verbosity="normal")
    print("case predict.x RELAX on the displaced Ti2O4 cell", flush=True)
    r = subprocess.run([EXE, inp], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    assert "(optimized)" in r.stdout
    assert "converged after" in r.stdout
    res = parse_output(r.stdout)[0]
    orc = Oracle(names, TI_O)
    o0 = orc.structure(TI2O4_CELL, True, TI2O4_TYPES, s.positions)
    assert res["Ecoh"] < o0["Ecoh"] - 1e-3
    final = np.array(res["rows6"])[-len(TI2O4_TYPES):]
    x = final[:, :3]
    o = orc.structure(TI2O4_CELL, True, TI2O4_TYPES, x)
    fmax = np.abs(o["forces"]).max()
    print("relaxed: Ecoh %.6f (from %.6f), oracle max |F| %.4f eV/A at the printed coordinates"
          % (res["Ecoh"], o0["Ecoh"], fmax))
    assert fmax < 0.01 + 1e-3
    assert abs(o["Ecoh"] - res["Ecoh"]) < 1e-5 * len(x) + 1e-4
