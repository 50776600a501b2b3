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
  held-out snapshots.

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
    ap.add_argument("--out", default=None)
    args = ap.parse_args()
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    pd = os.path.join(root, "tests", "golden", "potentials")
    tf = {s: os.path.join(pd, "teacher.%s.nn" % s) for s in "AB"}
    sf = {s: os.path.join(pd, "student.%s.nn" % s) for s in "AB"}
    dev = torch.device("cuda:0")
    teacher = Model(tf, device=dev)
    student = Model(sf, device=dev)
    cells = [synth.mc_cell(1000 + c) for c in range(args.chains)]
    mc = MonteCarlo(teacher, cells, first=student, seed=11)
    mc.set(T=600.0)
    t0 = time.perf_counter()
    mc.run(args.burn)
    data = collect(mc, teacher, args.snapshots, args.between, dev)
    tr, va = split(data, 4 * args.chains)  # the last 4 snapshots are held out
    e0, f0 = errors(student, va)
    train(student, tr, epochs=args.epochs, log=print)
    e1, f1 = errors(student, va)
    print("train %.1f s; held-out RMSE: atomic energy %.4f -> %.4f eV, force %.4f -> %.4f eV/A"
          % (time.perf_counter() - t0, e0, e1, f0, f1))
    out = args.out or pd
    os.makedirs(out, exist_ok=True)
    write_trained(student, sf, {s: os.path.join(out, "student_trained.%s.nn" % s) for s in "AB"})
    print("wrote", out)


if __name__ == "__main__":
    main()
