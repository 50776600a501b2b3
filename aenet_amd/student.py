"""Training a small first-stage potential (the student) on the full
potential (the teacher) along the teacher's own Monte Carlo trajectories,
with this library's GPU kernels only:

- data: snapshots of plain teacher MC chains; per atom the teacher's atomic
  energy (the batch evaluation of the full potential) and the student's
  normalized descriptor (the batch descriptor kernel with the student's
  basis);
- training: squared errors of the atomic energies and of the forces over
  whole cells, gradients by the fit kernel (k_first_stage_fit: the force
  term differentiated through the network's input gradient), updates by
  the Adam kernel;
- error report: RMSE of atomic energies and forces against the teacher on
  held-out snapshots;
- then training on what the second stage of delayed acceptance judges: the
  energy change of proposals drawn from the chains' own move distribution
  (swaps and displacements from equilibrated teacher states), with the loss
  sum_p (dE_student_p - dE_teacher_p)^2.  dE_teacher comes from batch
  evaluations of the cells before and after each move; dE_student is the
  change of the student's atomic energies over the atoms whose descriptor
  or species the move changes, and its weight gradient comes from the
  network-gradient kernel with per-atom upstream factors
  2 (dE_s - dE_t) (+1 after the move, -1 before).

Run as a script to train the student of tests/golden/potentials and write
the trained networks (student_trained.A/B.nn).
"""
import os

import numpy as np
import torch

from . import nnfile
from .device import DeviceBatch
from .synth import Structure


def collect(mc, teacher, nsnap, steps_between, device="cuda"):
    """nsnap snapshots of the chains of ``mc`` (plain teacher MC,
    steps_between moves per chain apart) with the teacher's atomic energies
    and forces.  Returns a dict of device tensors: coords [S*C*N, 3],
    types [S*C*N], E [S*C*N], F [S*C*N, 3], and natoms, cell."""
    x, t = mc.state()
    structs = [Structure(mc.cell, x[c], t[c], True) for c in range(mc.nchains)]
    db = DeviceBatch(structs, device=device, model=teacher)
    X, T, E, F = [], [], [], []
    for _ in range(nsnap):
        mc.run(steps_between, delayed=False)
        x, t = mc.state()
        db.coords.view(-1).copy_(torch.as_tensor(x.reshape(-1), device=db.device))
        db.types.copy_(torch.as_tensor(t.reshape(-1), device=db.device))
        _, _, E_at, F_at = db.eval(forces=True, model=teacher)
        X.append(db.coords.clone())
        T.append(db.types.clone())
        E.append(E_at.clone())
        F.append(F_at.clone())
    return {"coords": torch.cat(X), "types": torch.cat(T), "E": torch.cat(E), "F": torch.cat(F),
            "natoms": mc.natoms, "cell": mc.cell.copy()}


def split(data, ncells_out):
    """(train, held-out): the last ncells_out cells are held out."""
    n = data["natoms"]
    k = data["types"].numel() - ncells_out * n
    a = {key: (v[:k] if torch.is_tensor(v) else v) for key, v in data.items()}
    b = {key: (v[k:] if torch.is_tensor(v) else v) for key, v in data.items()}
    return a, b


def errors(student, data, chunk=512):
    """RMSE of the atomic energies (eV) and force components (eV/A)."""
    dev = data["coords"].device
    n = data["natoms"]
    ncell = data["types"].numel() // n
    grad = torch.zeros(student.wsize, dtype=torch.float32, device=dev)
    loss = torch.zeros(2, dtype=torch.float64, device=dev)
    for c0 in range(0, ncell, chunk):
        sel = torch.arange(c0, min(ncell, c0 + chunk), dtype=torch.int32, device=dev)
        student.fit_cells(data["cell"], n, data["coords"], data["types"], sel, data["E"], data["F"], 0.0, 0.0, grad,
                          loss)
    L = loss.cpu().numpy()
    return float(np.sqrt(L[0] / (ncell * n))), float(np.sqrt(L[1] / (3 * ncell * n)))


def train(student, data, epochs=40, cells_per_step=128, lr=3e-3, w_energy=1.0, w_force=1.0, seed=0, log=None):
    """Adam on w_energy * MSE(atomic energies) + w_force * MSE(force
    components) over cells; returns (energy MSE, force MSE) per epoch."""
    dev = data["coords"].device
    n = data["natoms"]
    ncell = data["types"].numel() // n
    grad = torch.zeros(student.wsize, dtype=torch.float32, device=dev)
    m = torch.zeros_like(grad)
    v = torch.zeros_like(grad)
    loss = torch.zeros(2, dtype=torch.float64, device=dev)
    gen = torch.Generator(device=dev).manual_seed(seed)
    t = 0
    hist = []
    for ep in range(epochs):
        perm = torch.randperm(ncell, generator=gen, device=dev).to(torch.int32)
        loss.zero_()
        for c0 in range(0, ncell, cells_per_step):
            sel = perm[c0:c0 + cells_per_step].contiguous()
            na = sel.numel() * n
            grad.zero_()
            student.fit_cells(data["cell"], n, data["coords"], data["types"], sel, data["E"], data["F"],
                              2.0 * w_energy / na, 2.0 * w_force / (3 * na), grad, loss)
            t += 1
            student.adam(grad, m, v, t, lr)
        L = loss.cpu().numpy()
        hist.append((L[0] / (ncell * n), L[1] / (3 * ncell * n)))
        if log:
            log("epoch %d  energy mse %.4e eV^2  force mse %.4e (eV/A)^2" % (ep, hist[-1][0], hist[-1][1]))
    student.check()
    return hist


def propose(x, t, cell, rng, p_swap=0.5, max_disp=0.075):
    """One proposal per chain from the MC move distribution: with p_swap a
    swap of the species of two atoms of different species (the first
    uniform, the second uniform among the other species), else a
    displacement of one uniform atom by U(-max_disp, max_disp)^3.  x (C, N,
    3), t (C, N); returns the states after the moves."""
    C, N = t.shape
    x2, t2 = x.copy(), t.copy()
    for c in range(C):
        i = int(rng.integers(N))
        other = np.nonzero(t[c] != t[c, i])[0]
        if rng.random() < p_swap and other.size:
            j = int(other[rng.integers(other.size)])
            t2[c, i], t2[c, j] = t[c, j], t[c, i]
        else:
            x2[c, i] += rng.uniform(-max_disp, max_disp, 3)
    return x2, t2


def collect_moves(mc, teacher, student, nsnap, steps_between, seed=0, p_swap=0.5, max_disp=0.075,
                  device="cuda"):
    """Proposals from nsnap snapshots of the plain teacher chains of mc
    (steps_between moves per chain apart), one per chain and snapshot.
    Returns device tensors: dE [P] (teacher energy change), and per affected
    atom row (atoms whose student descriptor or species the move changes,
    before and after): G [R, nsf] normalized student descriptors, types [R],
    sign [R] (-1 before, +1 after), start [P+1] (rows of proposal p are
    start[p]:start[p+1])."""
    rng = np.random.default_rng(seed)
    x, t = mc.state()
    C, N = t.shape
    structs = [Structure(mc.cell, x[c], t[c], True) for c in range(C)]
    db = DeviceBatch(structs, device=device, model=teacher)
    dE, G, T, S, cnt = [], [], [], [], []
    for _ in range(nsnap):
        mc.run(steps_between, delayed=False)
        x, t = mc.state()
        x2, t2 = propose(x, t, mc.cell, rng, p_swap, max_disp)
        out = []
        for xs, ts in ((x, t), (x2, t2)):
            db.coords.view(-1).copy_(torch.as_tensor(xs.reshape(-1), device=db.device))
            db.types.copy_(torch.as_tensor(ts.reshape(-1), device=db.device))
            E_tot, _, _, _ = db.eval(forces=False, model=teacher)
            Gs = db.descriptors(student.nsf_max, scaled=True, model=student)
            out.append((E_tot.clone(), Gs, db.types.clone()))
        (Eb, Gb, tb), (Ea, Ga, ta) = out
        dE.append(Ea - Eb)
        changed = ((Ga != Gb).any(dim=1) | (ta != tb)).view(C, N)
        for c in range(C):
            rows = torch.nonzero(changed[c]).view(-1) + c * N
            G.append(Gb[rows])
            G.append(Ga[rows])
            T.append(tb[rows])
            T.append(ta[rows])
            k = rows.numel()
            S.append(torch.cat([torch.full((k,), -1.0, dtype=torch.float64, device=db.device),
                                torch.full((k,), 1.0, dtype=torch.float64, device=db.device)]))
            cnt.append(2 * k)
    del db
    start = torch.zeros(len(cnt) + 1, dtype=torch.int64, device=S[0].device)
    start[1:] = torch.cumsum(torch.tensor(cnt, dtype=torch.int64, device=start.device), 0)
    return {"dE": torch.cat(dE), "G": torch.cat(G).contiguous(), "types": torch.cat(T).contiguous(),
            "sign": torch.cat(S), "start": start}


def _rows(moves, sel):
    """Row indices of the proposals sel and the local proposal of each row."""
    st = moves["start"]
    a, b = st[sel], st[sel + 1]
    cnt = b - a
    local = torch.repeat_interleave(torch.arange(sel.numel(), device=sel.device), cnt)
    first = torch.cumsum(cnt, 0) - cnt
    idx = torch.repeat_interleave(a, cnt) + torch.arange(int(cnt.sum()), device=sel.device) - \
        torch.repeat_interleave(first, cnt)
    return idx, local


def move_dE(student, moves, sel):
    """The student's energy change of the proposals sel."""
    idx, local = _rows(moves, sel)
    E = torch.zeros(idx.numel(), dtype=torch.float64, device=idx.device)
    student.net_grad(moves["G"][idx].contiguous(), moves["types"][idx].contiguous(), E_out=E)
    return torch.zeros(sel.numel(), dtype=torch.float64, device=idx.device).index_add_(
        0, local, moves["sign"][idx] * E)


def move_errors(student, moves, T=600.0, chunk=4096):
    """RMSE and median |dE_s - dE_t| (eV) of the proposals, and the mean
    second-stage acceptance min(1, exp(-(dE_t - dE_s)/kT)) of those that
    the first stage would accept (each weighted by its first-stage
    acceptance min(1, exp(-dE_s/kT)))."""
    P = moves["dE"].numel()
    dev = moves["dE"].device
    d = []
    for c0 in range(0, P, chunk):
        sel = torch.arange(c0, min(P, c0 + chunk), device=dev)
        d.append(move_dE(student, moves, sel))
    dEs = torch.cat(d)
    r = (dEs - moves["dE"]).cpu().numpy()
    w = move_weights(moves["dE"], T).cpu().numpy()
    beta = 1.0 / (8.617333262e-5 * T)
    s = dEs.cpu().numpy()
    t = moves["dE"].cpu().numpy()
    w1 = np.minimum(1.0, np.exp(-beta * s))
    a2 = np.minimum(1.0, np.exp(-beta * (t - s)))
    return {"rmse": float(np.sqrt(np.mean(r * r))), "median_abs": float(np.median(np.abs(r))),
            "weighted_rmse": float(np.sqrt(np.sum(w * r * r) / np.sum(w))),
            "second_stage_acceptance": float(np.sum(w1 * a2) / np.sum(w1))}


def move_weights(dE, T=600.0, margin=0.15):
    """Relevance of a proposal to the acceptance: 1 for teacher energy
    changes up to margin (eV) above zero, decaying as exp(-(dE - margin)/kT)
    beyond (moves both stages reject whatever their error)."""
    beta = 1.0 / (8.617333262e-5 * T)
    return torch.clamp(torch.exp(-beta * (dE - margin)), max=1.0)


def train_moves(student, moves, epochs=40, batch=256, lr=1e-3, seed=0, log=None, T=600.0, margin=0.15):
    """Adam on the weighted mean of (dE_s - dE_t)^2 over proposals (weights
    move_weights); returns the weighted mean squared error per epoch."""
    dev = moves["dE"].device
    P = moves["dE"].numel()
    wts = move_weights(moves["dE"], T, margin) if margin is not None else torch.ones_like(moves["dE"])
    grad = torch.zeros(student.wsize, dtype=torch.float32, device=dev)
    m = torch.zeros_like(grad)
    v = torch.zeros_like(grad)
    gen = torch.Generator(device=dev).manual_seed(seed)
    t = 0
    hist = []
    for ep in range(epochs):
        perm = torch.randperm(P, generator=gen, device=dev)
        tot = 0.0
        for c0 in range(0, P, batch):
            sel = perm[c0:c0 + batch]
            idx, local = _rows(moves, sel)
            Gi = moves["G"][idx].contiguous()
            ti = moves["types"][idx].contiguous()
            si = moves["sign"][idx]
            E = torch.zeros(idx.numel(), dtype=torch.float64, device=dev)
            student.net_grad(Gi, ti, E_out=E)
            r = torch.zeros(sel.numel(), dtype=torch.float64, device=dev).index_add_(0, local, si * E) - \
                moves["dE"][sel]
            w = wts[sel]
            tot += float((w * r * r).sum())
            up = (2.0 / float(w.sum().clamp(min=1e-12))) * (w * r)[local] * si
            grad.zero_()
            student.net_grad(Gi, ti, upstream=up.contiguous(), grad=grad)
            t += 1
            student.adam(grad, m, v, t, lr)
        hist.append(tot / float(wts.sum()))
        if log:
            log("epoch %d  weighted dE mse %.4e eV^2" % (ep, hist[-1]))
    student.check()
    return hist


def write_trained(student, src_files, dst_files):
    """Write the student's current weights into copies of its .nn files."""
    W = student.weights().cpu().numpy().astype(np.float64)
    for k, sp in enumerate(student.atom_types):
        pot = nnfile.read_nn(src_files[sp])
        off, size = student.species_weights(k + 1)
        pot["W"] = W[off:off + size]
        nnfile.write_nn(dst_files[sp], pot)


def main():
    import argparse
    import time
    from . import synth
    from .mc import MonteCarlo
    from .model import Model
    ap = argparse.ArgumentParser(description="train the Monte Carlo first-stage potential on the teacher")
    ap.add_argument("--chains", type=int, default=512)
    ap.add_argument("--burn", type=int, default=2000)
    ap.add_argument("--snapshots", type=int, default=24)
    ap.add_argument("--between", type=int, default=200)
    ap.add_argument("--epochs", type=int, default=60)
    ap.add_argument("--batch", type=int, default=128, help="cells per Adam step of the energy/force fit")
    ap.add_argument("--move-snapshots", type=int, default=48,
                    help="snapshots of proposals (one per chain each) for the energy-change stage; 0: skip it")
    ap.add_argument("--move-epochs", type=int, default=60)
    ap.add_argument("--move-lr", type=float, default=1e-3)
    ap.add_argument("--margin", type=float, default=0.15,
                    help="proposals weighted 1 up to this teacher energy change (eV), exp(-(dE - margin)/kT) above")
    ap.add_argument("--start", default=None, help="start from these student networks (directory of "
                                                  "student_trained.A/B.nn) instead of fitting energies first")
    ap.add_argument("--out", default=None)
    args = ap.parse_args()
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    pd = os.path.join(root, "tests", "golden", "potentials")
    tf = {s: os.path.join(pd, "teacher.%s.nn" % s) for s in "AB"}
    sf = {s: os.path.join(pd, "student.%s.nn" % s) for s in "AB"}
    if args.start:
        sf = {s: os.path.join(args.start, "student_trained.%s.nn" % s) for s in "AB"}
    dev = torch.device("cuda:0")
    teacher = Model(tf, device=dev)
    student = Model(sf, device=dev)
    cells = [synth.mc_cell(1000 + c) for c in range(args.chains)]
    mc = MonteCarlo(teacher, cells, first=student, seed=11)
    mc.set(T=600.0)
    t0 = time.perf_counter()
    mc.run(args.burn)
    if not args.start:
        data = collect(mc, teacher, args.snapshots, args.between, dev)
        tr, va = split(data, 4 * args.chains)  # the last 4 snapshots are held out
        e0, f0 = errors(student, va)
        train(student, tr, epochs=args.epochs, cells_per_step=args.batch, log=print)
        e1, f1 = errors(student, va)
        print("train %.1f s; held-out RMSE: atomic energy %.4f -> %.4f eV, force %.4f -> %.4f eV/A"
              % (time.perf_counter() - t0, e0, e1, f0, f1), flush=True)
    if args.move_snapshots > 0:
        nva = max(1, args.move_snapshots // 8)
        moves = collect_moves(mc, teacher, student, args.move_snapshots + nva, args.between, seed=5, device=dev)
        P = moves["dE"].numel()
        k = args.move_snapshots * args.chains
        tr = {"dE": moves["dE"][:k], "G": moves["G"], "types": moves["types"], "sign": moves["sign"],
              "start": moves["start"][:k + 1]}
        va = {"dE": moves["dE"][k:], "G": moves["G"], "types": moves["types"], "sign": moves["sign"],
              "start": moves["start"][k:]}
        before = move_errors(student, va)
        print("proposals: %d training, %d held out; held-out before: %s" % (k, P - k, before), flush=True)
        train_moves(student, tr, epochs=args.move_epochs, lr=args.move_lr, log=print, margin=args.margin)
        after = move_errors(student, va)
        print("held-out after: %s (%.1f s)" % (after, time.perf_counter() - t0), flush=True)
    out = args.out or pd
    os.makedirs(out, exist_ok=True)
    src = {s: os.path.join(pd, "student.%s.nn" % s) for s in "AB"}
    write_trained(student, src, {s: os.path.join(out, "student_trained.%s.nn" % s) for s in "AB"})
    print("wrote", out)


if __name__ == "__main__":
    main()
