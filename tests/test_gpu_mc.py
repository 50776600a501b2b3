"""GPU Monte Carlo, the first-stage (student) potential and its training.

- bookkeeping: after plain and delayed-acceptance runs, the chains' running
  energies equal a fresh evaluation of the final states with the batch
  kernels, and the incrementally updated first-stage descriptors and
  energies equal the standard kernels' values for the student potential;
  species counts are conserved by swaps;
- the fit kernel's gradient (energies and forces) against central
  differences of its own loss; a short training run lowers the held-out
  errors;
- the decisions follow the plain and the delayed-acceptance rules exactly
  (recorded energy changes against acceptance frequencies), the trained
  first stage accepts more than the untrained one, and delayed acceptance
  samples the same equilibrium distribution as plain Metropolis.
"""
import numpy as np
import pytest

from conftest import pot_files

pytestmark = pytest.mark.gpu

AB = ["A", "B"]


def _models(device):
    from aenet_amd.model import Model
    teacher = Model(pot_files("teacher", AB), device=device)
    trained = Model(pot_files("student_trained", AB), device=device)
    return teacher, trained


def _cells(n, seed0=100):
    from aenet_amd import synth
    return [synth.mc_cell(seed0 + c) for c in range(n)]


def _fresh(model, cell, x, t, device):
    from aenet_amd import synth
    from aenet_amd.device import DeviceBatch
    db = DeviceBatch([synth.Structure(cell, x[c], t[c], True) for c in range(len(x))], device=device, model=model)
    E_tot, _, E_at, _ = db.eval(forces=False)
    return E_tot.cpu().numpy(), E_at.cpu().numpy()


def test_bookkeeping_and_first_stage(device):
    from aenet_amd.mc import MonteCarlo
    teacher, student = _models(device)
    cells = _cells(128)
    mc = MonteCarlo(teacher, cells, first=student, seed=3)
    mc.set(T=600.0)
    comp0 = np.array([np.bincount(c.types, minlength=3) for c in cells])
    mc.run(300, delayed=False)
    st = mc.stats(reset=True)
    assert st["full_evaluations"] == st["proposed"] and 0 < st["accepted"] < st["proposed"]
    mc.run(300, delayed=True, graph=True)
    st = mc.stats()
    print("delayed:", st, flush=True)
    assert st["accepted"] <= st["full_evaluations"] <= st["proposed"]
    x, t = mc.state()
    comp1 = np.array([np.bincount(tt, minlength=3) for tt in t])
    assert np.array_equal(comp0, comp1)
    E_run = mc.energies()
    E_new, _ = _fresh(teacher, mc.cell, x, t, device)
    print("max |E_chain - fresh| %.3e eV" % np.abs(E_run - E_new).max())
    assert np.abs(E_run - E_new).max() < 2e-3  # 256 FP32 atomic energies, summation order of the lists
    E_s, _ = mc.first_stage_state()
    _, E_s_std = _fresh(student, mc.cell, x, t, device)
    print("max |E_first incremental - standard| %.3e eV" % np.abs(E_s - E_s_std).max())
    assert np.abs(E_s - E_s_std).max() < 5e-5


def test_fit_gradient_and_training(device):
    import torch
    from aenet_amd import student as S
    from aenet_amd.mc import MonteCarlo
    from aenet_amd.model import Model
    teacher = Model(pot_files("teacher", AB), device=device)
    fresh = Model(pot_files("student", AB), device=device)
    mc = MonteCarlo(teacher, _cells(64, 500), seed=5)
    mc.set(T=600.0)
    data = S.collect(mc, teacher, 3, 50, device)
    tr, va = S.split(data, 64)
    n = data["natoms"]
    # gradient of the combined loss against central differences
    sel = torch.arange(0, 8, dtype=torch.int32, device=device)
    cE, cF = 2.0 / (8 * n), 2.0 / (3 * 8 * n)

    def loss_at(W):
        fresh.set_weights(W)
        g = torch.zeros(fresh.wsize, dtype=torch.float32, device=device)
        L = torch.zeros(2, dtype=torch.float64, device=device)
        fresh.fit_cells(tr["cell"], n, tr["coords"], tr["types"], sel, tr["E"], tr["F"], cE, cF, g, L)
        L = L.cpu().numpy()
        return 0.5 * (cE * L[0] + cF * L[1]), g

    W0 = fresh.weights()
    _, g = loss_at(W0)
    g = g.cpu().numpy()
    for k in np.argsort(-np.abs(g))[:6]:
        h = 2e-3
        Wp = W0.clone()
        Wp[k] += h
        Wm = W0.clone()
        Wm[k] -= h
        fd = (loss_at(Wp)[0] - loss_at(Wm)[0]) / (2 * h)
        print("w[%d]: kernel %.6e  fd %.6e" % (k, g[k], fd))
        assert abs(fd - g[k]) <= 2e-2 * abs(g[k]) + 1e-5
    fresh.set_weights(W0)
    e0, f0 = S.errors(fresh, va)
    S.train(fresh, tr, epochs=8)
    e1, f1 = S.errors(fresh, va)
    print("held-out RMSE: energy %.4f -> %.4f eV, force %.4f -> %.4f eV/A" % (e0, e1, f0, f1))
    assert e1 ** 2 + f1 ** 2 < e0 ** 2 + f0 ** 2


def test_fit_forward_matches_reference(device):
    """The fit kernel's own forward pass (the student's atomic energies and
    forces inside its loss) against the reference build's values for the
    student potential (tests/golden/ref_07_student_student.npz, recorded
    from aenet_atomic_energy_and_forces, src/aenet.f90:410-486).  With the
    reference values as targets and zero loss weights, the kernel's loss
    sums are sum_i (E_i - E_ref,i)^2 and sum |F - F_ref|^2; bounding each
    sum by the square of the per-atom bar bounds every atom's error by the
    bar.  The gradient stays zero."""
    import os
    import torch
    from aenet_amd.model import Model
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", "ref_07_student_student.npz"), allow_pickle=False)
    assert [str(t) for t in g["types"]] == AB and bool(g["pbc"])
    n = len(g["atom_types"])
    # the fit kernel uses minimum images: the cell must be wider than 2 Rc
    assert np.min(np.diag(g["cell"])) > 2 * 5.0
    student = Model(pot_files("student", AB), device=device)
    coords = torch.as_tensor(g["positions"].reshape(-1), dtype=torch.float64, device=device)
    types = torch.as_tensor(g["atom_types"], dtype=torch.int32, device=device)
    E_ref = torch.as_tensor(g["energies"], dtype=torch.float64, device=device)
    F_ref = torch.as_tensor(g["forces"].reshape(-1), dtype=torch.float64, device=device)
    sel = torch.zeros(1, dtype=torch.int32, device=device)
    grad = torch.zeros(student.wsize, dtype=torch.float32, device=device)
    loss = torch.zeros(2, dtype=torch.float64, device=device)
    student.fit_cells(g["cell"], n, coords, types, sel, E_ref, F_ref, 0.0, 0.0, grad, loss)
    L = loss.cpu().numpy()
    print("fit forward vs reference: sqrt(sum dE^2) %.3e eV, sqrt(sum |dF|^2) %.3e eV/A (over %d atoms)"
          % (np.sqrt(L[0]), np.sqrt(L[1]), n))
    assert np.abs(g["forces"]).max() > 0.1
    assert np.sqrt(L[0]) < 1e-5   # implies every |dE_i| < 1e-5 eV
    assert np.sqrt(L[1]) < 1e-4   # implies every force component within 1e-4 eV/A
    assert not grad.any()


def _rule_check(p, hit, label):
    """Bernoulli outcomes hit with probabilities p: total and per p-bin
    counts within 4 standard deviations."""
    p = np.asarray(p)
    hit = np.asarray(hit, dtype=float)
    z = (hit.sum() - p.sum()) / np.sqrt(max((p * (1 - p)).sum(), 1e-12))
    print("%s: %d proposals, observed %.4f expected %.4f, z = %.2f" % (label, len(p), hit.mean(), p.mean(), z))
    assert abs(z) < 4.0
    edges = np.array([0.0, 0.05, 0.2, 0.4, 0.6, 0.8, 0.95, 1.0 + 1e-12])
    for lo, hi in zip(edges[:-1], edges[1:]):
        sel = (p >= lo) & (p < hi)
        if sel.sum() < 200:
            continue
        v = (p[sel] * (1 - p[sel])).sum()
        zb = (hit[sel].sum() - p[sel].sum()) / np.sqrt(max(v, 1e-12))
        assert abs(zb) < 4.5, (label, lo, hi, zb)


def test_acceptance_rules(device):
    """Every decision against its probability: plain min(1, e^-b dE);
    delayed acceptance min(1, e^-b dE_first) then min(1, e^-b (dE_full -
    dE_first)) (Christen & Fox 2005), from the recorded energy changes; the
    trained first stage lets more proposals through to acceptance than the
    untrained network."""
    from aenet_amd.mc import K_B, MonteCarlo
    from aenet_amd.model import Model
    teacher, student = _models(device)
    untrained = Model(pot_files("student", AB), device=device)
    beta = 1.0 / (K_B * 600.0)
    acc = {}
    for name, first, delayed in (("plain", student, False), ("delayed", student, True),
                                 ("untrained", untrained, True)):
        mc = MonteCarlo(teacher, _cells(256, 300), first=first, seed=29)
        mc.set(T=600.0)
        mc.run(500)
        mc.arm_trace(800)
        t = mc.run(800, delayed=delayed)
        tr = mc.trace().reshape(-1, 4)
        st = mc.stats()
        print(name, "%.3f ms/step" % (t / 800 * 1e3), st, flush=True)
        moved = tr[:, 2] >= 0
        acc[name] = tr[moved, 3].mean()
        if not delayed:
            p = np.minimum(1.0, np.exp(-beta * tr[moved, 1]))
            _rule_check(p, tr[moved, 3], name)
        else:
            p1 = np.minimum(1.0, np.exp(-beta * tr[moved, 0]))
            _rule_check(p1, tr[moved, 2] == 1, name + " stage 1")
            passed = tr[:, 2] == 1
            p2 = np.minimum(1.0, np.exp(-beta * (tr[passed, 1] - tr[passed, 0])))
            _rule_check(p2, tr[passed, 3], name + " stage 2")
            assert np.all(tr[tr[:, 2] == 0, 3] == 0)
        del mc
    print("acceptance: plain %.4f, delayed (trained) %.4f, delayed (untrained) %.4f"
          % (acc["plain"], acc["delayed"], acc["untrained"]))
    assert acc["delayed"] > 1.2 * acc["untrained"]


def test_delayed_acceptance_samples_the_same_distribution(device):
    """Equilibrium of species swaps on fixed sites at 5000 K (a regime where
    the chains reach a stationary energy within the run): started from
    common equilibrated states, the per-chain mean energies of delayed
    acceptance and plain Metropolis agree within their statistical error,
    and so do their distributions over the chains (two-sample KS)."""
    from scipy import stats
    from aenet_amd import synth
    from aenet_amd.mc import MonteCarlo
    teacher, student = _models(device)
    C = 256
    T = 5000.0
    burn = MonteCarlo(teacher, _cells(C, 2000), seed=17)
    burn.set(T=T, p_swap=1.0)
    burn.run(20000)
    x, t = burn.state()
    start = [synth.Structure(burn.cell, x[c], t[c], True) for c in range(C)]
    del burn
    means = {}
    for name, first, delayed in (("plain", None, False), ("delayed", student, True)):
        mc = MonteCarlo(teacher, start, first=first, seed=23)
        mc.set(T=T, p_swap=1.0)
        mc.run(20000, delayed=delayed, sample_every=200)
        S = mc.samples() / 256.0
        means[name] = S.mean(axis=0)
        print("%s %s mean E/atom %.6f" % (name, mc.stats(), means[name].mean()), flush=True)
        del mc
    d = means["delayed"] - means["plain"]
    z = d.mean() / (d.std(ddof=1) / np.sqrt(C))
    ks = stats.ks_2samp(means["delayed"], means["plain"])
    print("delayed - plain: mean %.3e eV/atom, z = %.2f; KS p = %.3f" % (d.mean(), z, ks.pvalue))
    assert abs(z) < 4.0
    assert ks.pvalue > 1e-3
