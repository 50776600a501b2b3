"""generate.x on the GPU (aenet_amd.generate) against the reference's own
generate.x, built from its sources (oracle/_ref/generate.x) and recorded
into tests/golden/generate/ by tests/golden/make_generate_fixture.py.

- CPU: the .train reader and writer reproduce the reference's file byte for
  byte; the .stp parser gives the set-ups the reference stored in the
  footer (parameters, environment, description); the structures' records
  (names, energies, species, coordinates, forces) come out identical;
- GPU: the whole .train written by the driver against the reference's:
  identical records except the descriptor values and their statistics
  (FP32 descriptor sums, within 1e-5 of 1 + |G|), and the descriptor
  derivatives of FORCES against the reference's records.
"""
import os
import sys
import tempfile

import numpy as np
import pytest

from conftest import ROOT

GOLDEN = os.path.join(ROOT, "tests", "golden", "generate")
sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))


@pytest.mark.parametrize("kind", ["cheb", "bp"])
def test_train_file_round_trip(kind, tmp_path):
    from aenet_amd import trainfile
    path = os.path.join(GOLDEN, kind + ".train")
    t = trainfile.read_train(path)
    out = str(tmp_path / "x.train")
    w = trainfile.TrainWriter(out, t["types"], t["E_atom"], len(t["structures"]), t["normalized"], t["scale"],
                              t["shift"])
    for s in t["structures"]:
        w.structure(s["name"], s["ntypes"], s["energy"], s["itype"], s["coords"], s["forces"], s["sfval"])
    w.close(t["setups"])
    assert open(out, "rb").read() == open(path, "rb").read()


@pytest.mark.parametrize("kind", ["cheb", "bp"])
def test_setup_parser_and_structure_records(kind):
    """read_stp against the set-ups in the reference's footer, and the
    per-structure records this driver computes on the host (cohesive
    energies, species, wrapped/box coordinates, forces) against the
    reference's."""
    import make_generate_fixture as mg
    from aenet_amd import generate, trainfile, xsf
    ref = trainfile.read_train(os.path.join(GOLDEN, kind + ".train"))
    with tempfile.TemporaryDirectory() as d:
        gin = mg.write_inputs(d, kind)
        inp = generate.read_input(gin)
        assert inp["types"] == ["A", "B"] and inp["E_atom"] == [-1000.5, -1500.25]
        for t, p in zip(inp["types"], inp["setups"]):
            stp = generate.read_stp(os.path.join(d, p), inp["types"])
            r = ref["setups"][inp["types"].index(t)]
            for k in ("atomtype", "envtypes", "Rc_min", "Rc_max", "sftype", "nsf", "description"):
                assert stp[k] == r[k], k
            for k in ("sf", "sfparam", "sfenv"):
                assert np.array_equal(np.asarray(stp[k]), r[k]), k
        for f, rs in zip(inp["files"], ref["structures"]):
            s = xsf.read_xsf(os.path.join(d, f))
            gt = np.array([inp["types"].index(n) + 1 for n in s.names])[s.types - 1]
            E_coh = s.E_tot - sum(inp["E_atom"][t - 1] for t in gt)
            assert rs["name"] == f and rs["natoms"] == len(gt)
            assert np.array_equal(rs["itype"], gt)
            assert abs(E_coh - rs["energy"]) <= 1e-12 * abs(rs["energy"])
            coo, frc = generate.written_coordinates(s)
            assert np.abs(coo - rs["coords"]).max() < 1e-12
            assert np.abs(frc - rs["forces"]).max() < 1e-12


@pytest.mark.gpu
@pytest.mark.parametrize("kind", ["cheb", "bp"])
def test_generate_matches_reference(kind, device):
    import make_generate_fixture as mg
    from aenet_amd import generate, trainfile
    ref = trainfile.read_train(os.path.join(GOLDEN, kind + ".train"))
    with tempfile.TemporaryDirectory() as d:
        gin = mg.write_inputs(d, kind)
        print("case generate", kind, flush=True)
        out = generate.run(gin, log=lambda *a: None)
        got = trainfile.read_train(out)
        forces = trainfile.read_forces(out + ".forces") if kind == "cheb" else None
    for k in ("types", "normalized", "scale", "shift", "natoms_tot", "E_av", "E_min", "E_max"):
        assert got[k] == ref[k], k
    assert np.array_equal(got["E_atom"], ref["E_atom"])
    worst = 0.0
    for a, b in zip(got["structures"], ref["structures"]):
        for k in ("name", "natoms", "ntypes", "energy"):
            assert a[k] == b[k], k
        assert np.array_equal(a["itype"], b["itype"])
        assert np.abs(a["coords"] - b["coords"]).max() < 1e-12
        assert np.abs(a["forces"] - b["forces"]).max() < 1e-12
        for va, vb in zip(a["sfval"], b["sfval"]):
            assert va.shape == vb.shape
            worst = max(worst, float((np.abs(va - vb) / (1.0 + np.abs(vb))).max()))
    print("%s: max relative |dG| %.3e" % (kind, worst))
    assert worst < 1e-5
    for sa, sb in zip(got["setups"], ref["setups"]):
        assert sa["neval"] == sb["neval"]
        for k in ("sfval_min", "sfval_max", "sfval_avg", "sfval_cov"):
            assert np.allclose(sa[k], sb[k], rtol=1e-5, atol=1e-5), k
    if forces is not None:
        sub = np.load(os.path.join(GOLDEN, kind + "_forces_subset.npz"), allow_pickle=False)
        assert forces["nstrucs"] == int(sub["nstrucs"]) and forces["max_nnb"] == int(sub["max_nnb"])
        werr = 0.0
        for k, s in enumerate(forces["structures"]):
            assert np.array_equal([s["natoms"], s["ntypes"], s["selected"]], sub["s%d_meta" % k])
            for i in (0, 5, s["natoms"] - 1):
                a = s["atoms"][i]
                assert np.array_equal([a["itype"], a["nsf"], a["nnb"]], sub["s%d_a%d_meta" % (k, i)])
                assert np.array_equal(a["nblist"], sub["s%d_a%d_nblist" % (k, i)])
                for key in ("deriv_i", "deriv_j"):
                    r = sub["s%d_a%d_%s" % (k, i, key)]
                    werr = max(werr, float(np.abs(a[key] - r).max() / max(1.0, np.abs(r).max())))
        print("dG/dR records: max |d| / max(1, |dG/dR|) = %.3e" % werr)
        assert werr < 2e-5
