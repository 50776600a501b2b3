"""train.x on the GPU: per-species networks fitted to the energies of a
training set (``.train``), written as ``.nn`` potentials.

    python -m aenet_amd.train train.in

Input and output follow the reference program (src/train.f90:244-503):
the train.in keywords of read_InpTrain (src/input.f90:319-353: TRAININGSET,
TESTPERCENT, ITERATIONS, MAXENERGY, SAMPLING, METHOD, NETWORKS with
architectures, SAVE_ENERGIES, TIMING, DEBUG), the normalization of the
training set (ts_normalize, src/trainset.f90:1160-1271: energies scaled to
[-1, 1], structures above MAXENERGY dropped, descriptors shifted and
scaled with the set-ups' statistics, written to ``<file>.scaled``), the
random test split (decide_testing_set, src/train.f90:1437-1490), the error
function (eval_next_structure, :799-889: dE = (E_coh - sum_i E_i)/N per
structure, loss sum 0.5 dE^2), the per-epoch table (print_energies,
:1232-1245: MAE and RMSE per atom of the training and test sets) and the
networks saved after every epoch with the set-up and training-set records
(save_all_networks, :1251-1291).

Methods:
- ``adam`` (src/optimize.f90:914-1046): each epoch draws ``samplesize``
  structures, and every mini-batch of ``batchsize`` structures is one Adam
  step on the summed gradient;
- ``bfgs`` (the reference's default; it calls L-BFGS-B, src/optimize.f90:
  1587-1843): full-batch limited-memory BFGS with 20 corrections, one
  function and gradient evaluation per epoch, backtracking on the Armijo
  condition.
Energies and weight gradients come from the GPU network kernels
(aenet_amd_model_net_grad, the reference's ff_eval / ff_wderiv), for every
atom of a batch in one launch; the optimizer state lives on the GPU (Adam)
or the host (BFGS).  Weights are FP32 on the device.
"""
import os
import sys
import tempfile
import time

import numpy as np

from . import nnfile, trainfile

ACT = {"linear": 0, "tanh": 1, "sigmoid": 2, "mtanh": 3, "twist": 4}


def _content_lines(path):
    out = []
    with open(path) as f:
        for line in f:
            t = line.strip()
            if not t or t[0] in "!#%":
                continue
            out.append(t)
    return out


def _find(lines, kw):
    for i, l in enumerate(lines):
        if l.split()[0].lower() == kw:
            return i
    return -1


def _real(s):
    return float(s.replace("d", "e").replace("D", "e"))


def _kv(line, name, default):
    """io_readval on a METHOD line ('name=value')."""
    i = line.find(name)
    if i < 0:
        return default
    j = line.find("=", i + len(name))
    return _real(line[j + 1:].split()[0]) if j >= 0 else default


def read_input(path):
    """read_InpTrain (src/input.f90:319-353); the NETWORKS block with
    architectures (inp_read_networks with readarch, :797-860) is kept as
    lines."""
    L = _content_lines(path)
    inp = {"testpercent": 10.0, "iterations": 10, "maxenergy": float("inf"), "sampling": "sequential",
           "method": "bfgs", "params": {}, "save_energies": False, "timing": False, "debug": False}
    k = _find(L, "method")
    if k >= 0:
        line = L[k + 1]
        m = line.split()[0].lower()
        inp["method"] = m
        if m == "adam":
            inp["params"] = {"mu": _kv(line, "mu", 0.001), "b1": _kv(line, "b1", 0.9), "b2": _kv(line, "b2", 0.999),
                             "eps": _kv(line, "eps", 1e-8), "samplesize": _kv(line, "samplesize", 500.0),
                             "batchsize": _kv(line, "batchsize", 1.0)}
            inp["sampling"] = "random"
        elif m != "bfgs":
            raise ValueError("training method %s is not available here (bfgs, adam)" % m)
    for kw, key, conv in (("trainingset", "trainingset", str), ("testpercent", "testpercent", _real),
                          ("iterations", "iterations", int), ("maxenergy", "maxenergy", _real),
                          ("sampling", "sampling", str)):
        i = _find(L, kw)
        if i >= 0:
            inp[key] = conv(L[i].split()[1])
    for kw in ("save_energies", "timing", "debug"):
        inp[kw] = _find(L, kw) >= 0
    k = _find(L, "networks")
    if k < 0:
        raise ValueError("no NETWORKS in %s" % path)
    # one line per species of the training set follows (read once the
    # training set's species are known, as train.x does)
    inp["network_lines"] = L[k + 1:]
    inp["base"] = os.path.dirname(os.path.abspath(path))
    return inp


# ---------------------------------------------------------------------------
# training set normalization (ts_normalize, src/trainset.f90:1160-1271)
# ---------------------------------------------------------------------------

def stp_scale(stp):
    """stp_normalize's shift and scale (src/sfsetup.f90:1005-1031)."""
    avg = np.asarray(stp["sfval_avg"])
    s2 = np.asarray(stp["sfval_cov"]) - avg * avg
    s = np.sqrt(np.maximum(s2, 0.0))
    scale = np.where(s > 1.0e-10, 1.0 / np.where(s > 1.0e-10, s, 1.0), 0.0)
    return avg, scale


def normalize(ts, maxenergy, path_out=None):
    """The normalized training set (energies scale*(E - N shift), forces
    scaled alike, descriptors (G - avg) scale); structures whose scaled
    energy per atom exceeds 1.001 are dropped when MAXENERGY lies below the
    largest energy.  Written to path_out in the .train format if given."""
    E_max = min(ts["E_max"], maxenergy)
    scale = 2.0 / (E_max - ts["E_min"])
    shift = 0.5 * (E_max + ts["E_min"])
    norms = [stp_scale(s) for s in ts["setups"]]
    keep = []
    for s in ts["structures"]:
        e = scale * (s["energy"] - s["natoms"] * shift)
        if maxenergy < ts["E_max"] and e / s["natoms"] > 1.001:
            continue
        sf = []
        for i in range(s["natoms"]):
            avg, sc = norms[s["itype"][i] - 1]
            sf.append(sc * (s["sfval"][i] - avg))
        keep.append(dict(s, energy=e, forces=scale * s["forces"], sfval=sf))
    out = {"types": ts["types"], "E_atom": ts["E_atom"], "normalized": True, "scale": scale, "shift": shift,
           "structures": keep, "setups": ts["setups"]}
    w = None
    if path_out:
        w = trainfile.TrainWriter(path_out, ts["types"], ts["E_atom"], len(keep), True, scale, shift)
        for s in keep:
            w.structure(s["name"], s["ntypes"], s["energy"], s["itype"], s["coords"], s["forces"], s["sfval"])
        w.close(ts["setups"])
        out.update(natoms_tot=w.natoms_tot, E_av=w.E_av, E_min=w.E_min, E_max=w.E_max)
    else:
        e = [s["energy"] / s["natoms"] for s in keep]
        out.update(natoms_tot=sum(s["natoms"] for s in keep), E_av=float(np.mean(e)) if e else 0.0,
                   E_min=min(e) if e else 0.0, E_max=max(e) if e else 0.0)
    return out


def decide_testing_set(n, testpercent, rng):
    """decide_testing_set (src/train.f90:1437-1490) with a seeded generator
    in place of Fortran's random_number."""
    ntrain = int(np.ceil((1.0 - testpercent / 100.0) * n))
    ntest = n - ntrain
    is_test = np.zeros(n, bool)
    itrain = itest = 0
    for i in range(n):
        if itest == ntest:
            itrain += 1
        elif itrain == ntrain:
            itest += 1
            is_test[i] = True
        elif rng.random() <= testpercent / 100.0:
            itest += 1
            is_test[i] = True
        else:
            itrain += 1
    return is_test


def random_network(arch, rng):
    """ff_random_init_weights (src/ext/feedforward.f90:512-554): uniform
    weights with standard deviation 1/sqrt(inputs + 1) per layer."""
    W = []
    for l in range(len(arch) - 1):
        a = np.sqrt(12.0) / np.sqrt(arch[l] + 1.0)
        W.append((rng.random((arch[l] + 1) * arch[l + 1]) - 0.5) * a)
    return np.concatenate(W)


# ---------------------------------------------------------------------------
# the GPU side: packed atoms of a set of structures
# ---------------------------------------------------------------------------

class AtomSet:
    """Normalized descriptors, types and structure index of every atom of a
    list of structures, resident on the device."""

    def __init__(self, strucs, stride, device):
        import torch
        n = sum(s["natoms"] for s in strucs)
        G = np.zeros((max(n, 1), stride))
        t = np.ones(max(n, 1), np.int32)
        sid = np.zeros(max(n, 1), np.int64)
        a = 0
        for k, s in enumerate(strucs):
            for i in range(s["natoms"]):
                v = s["sfval"][i]
                G[a, :len(v)] = v
                t[a] = s["itype"][i]
                sid[a] = k
                a += 1
        self.n = n
        self.ns = len(strucs)
        self.G = torch.as_tensor(G, device=device)
        self.types = torch.as_tensor(t, device=device)
        self.sid = torch.as_tensor(sid, device=device)
        self.natoms = torch.as_tensor(np.array([s["natoms"] for s in strucs] or [1], np.float64), device=device)
        self.E_ref = torch.as_tensor(np.array([s["energy"] for s in strucs] or [0.0]), device=device)
        self.E = torch.zeros(max(n, 1), dtype=torch.float64, device=device)

    def errors(self, model, struct_mask=None):
        """(dE per structure) on the device: (E_ref - sum_i E_i)/N."""
        import torch
        model.net_grad(self.G, self.types, E_out=self.E)
        Es = torch.zeros(self.ns, dtype=torch.float64, device=self.G.device).index_add_(0, self.sid[:self.n],
                                                                                         self.E[:self.n])
        return (self.E_ref[:self.ns] - Es) / self.natoms[:self.ns]

    def grad(self, model, dE, grad, atom_mask=None):
        """grad += d/dW sum_s 0.5 dE_s^2 over the structures with nonzero
        dE (upstream g_i = -dE_s / N_s)."""
        g = (-dE / self.natoms[:self.ns])[self.sid[:self.n]]
        if atom_mask is not None:
            g = g * atom_mask
        model.net_grad(self.G, self.types, upstream=g.contiguous(), grad=grad)


def _write_nn(path, net_arch, acts, W, stp, tsinfo):
    pot = dict(stp)
    pot.update({"nnodes": net_arch, "f_a": acts, "W": W})
    pot.update(tsinfo)
    nnfile.write_nn(path, pot)


def run(input_path, device="cuda", log=print, seed=1):
    import torch
    from .model import Model
    inp = read_input(input_path)
    base = inp["base"]

    def rel(p):
        return p if os.path.isabs(p) else os.path.join(base, p)

    raw = trainfile.read_train(rel(inp["trainingset"]))
    if raw["setups"] is None:
        raise ValueError("no structural fingerprint basis setups in training set file")
    if raw["normalized"]:
        ts = dict(raw)
        ts["file"] = inp["trainingset"]
    else:
        scaled = rel(inp["trainingset"]) + ".scaled"
        ts = normalize(raw, inp["maxenergy"], scaled)
        ts["file"] = inp["trainingset"] + ".scaled"
        log(" The network output energy will be normalized to the interval [-1,1].")
        log("   Energy scaling factor: f = %.6f" % ts["scale"])
        log("   Atomic energy shift  : s = %.6f" % ts["shift"])
    types = list(ts["types"])
    nets = {}
    for line in inp["network_lines"][:len(types)]:
        w = line.split()
        nl = int(w[2]) if len(w) > 2 else 0
        nets[w[0]] = {"type": w[0], "file": w[1],
                      "hidden": [(int(x.split(":")[0]), x.split(":")[1]) for x in w[3:3 + nl]]}
    missing = [t for t in types if t not in nets]
    if missing:
        raise ValueError("no network for species %s" % ", ".join(missing))
    rng = np.random.Generator(np.random.PCG64(seed))
    archs, acts, W0 = [], [], []
    for k, t in enumerate(types):
        stp = ts["setups"][k]
        nf = rel(nets[t]["file"])
        if os.path.exists(nf):
            # restart from the existing network (init_networks, src/train.f90:577-581)
            p = nnfile.read_nn(nf)
            arch, act, w = [int(x) for x in p["nnodes"]], [int(x) for x in p["f_a"]], p["W"]
        else:
            arch = [stp["nsf"]] + [h for h, _ in nets[t]["hidden"]] + [1]
            act = [ACT[a.lower()] for _, a in nets[t]["hidden"]] + [0]
            w = random_network(arch, rng)
        archs.append(arch)
        acts.append(act)
        W0.append(w)
    # the device model: networks as they are trained (raw outputs, E_i = y)
    d = tempfile.mkdtemp(prefix="aenet_train_")
    files = {}
    for k, t in enumerate(types):
        files[t] = os.path.join(d, "%s.nn" % t)
        _write_nn(files[t], archs[k], acts[k], W0[k], ts["setups"][k],
                  {"ts_types": types, "ts_E_atom": [0.0] * len(types), "ts_scale": 1.0, "ts_shift": 0.0})
    model = Model(files, device=device)
    for f in files.values():
        os.remove(f)
    os.rmdir(d)
    stride = model.nsf_max
    is_test = decide_testing_set(len(ts["structures"]), inp["testpercent"], rng)
    trn = [s for s, te in zip(ts["structures"], is_test) if not te]
    tst = [s for s, te in zip(ts["structures"], is_test) if te]
    A_trn = AtomSet(trn, stride, device)
    A_tst = AtomSet(tst, stride, device)
    log(" Training structures: %d (%d atoms), test structures: %d (%d atoms)"
        % (len(trn), A_trn.n, len(tst), A_tst.n))

    def report(ep):
        dtr = A_trn.errors(model)
        out = [float(dtr.abs().mean()), float(torch.sqrt((dtr * dtr).mean()))]
        if A_tst.ns and tst:
            dte = A_tst.errors(model)
            out += [float(dte.abs().mean()), float(torch.sqrt((dte * dte).mean()))]
        else:
            out += [0.0, 0.0]
        sc = ts["scale"]
        log(" %5d  %14.6E  %14.6E  %14.6E  %14.6E <" % (ep, out[0] / sc, out[1] / sc, out[2] / sc, out[3] / sc))
        return [v / sc for v in out]

    def save(ep=None):
        Wd = model.weights().cpu().numpy().astype(np.float64)
        info = {"ts_file": ts["file"], "ts_normalized": True, "ts_scale": ts["scale"], "ts_shift": ts["shift"],
                "ts_types": types, "ts_E_atom": list(ts["E_atom"]), "ts_natoms": ts["natoms_tot"],
                "ts_nstrucs": len(ts["structures"]), "ts_E": [ts["E_min"], ts["E_max"], ts["E_av"]]}
        for k, t in enumerate(types):
            off, size = model.species_weights(k + 1)
            name = rel(nets[t]["file"]) + ("" if ep is None else "-%05d" % ep)
            _write_nn(name, archs[k], acts[k], Wd[off:off + size], ts["setups"][k], info)

    log(" %5s  %14s  %14s  %14s  %14s" % ("epoch", "MAE (train)", "RMSE (train)", "MAE (test)", "RMSE (test)"))
    hist = [report(0)]
    t0 = time.perf_counter()
    grad = torch.zeros(model.wsize, dtype=torch.float32, device=device)
    if inp["method"] == "adam":
        p = inp["params"]
        m = torch.zeros_like(grad)
        v = torch.zeros_like(grad)
        batch = max(1, min(int(p["batchsize"]), len(trn)))
        sample = max(1, int(p["samplesize"]) // batch) * batch
        sample = min(sample, len(trn))
        step = 1
        gen = torch.Generator(device=device).manual_seed(seed)
        for ep in range(1, inp["iterations"] + 1):
            # samplesize structures drawn at random (opt_schedule_epoch, 'random')
            pick = torch.randint(0, len(trn), (sample,), generator=gen, device=device)
            dE = A_trn.errors(model)
            for b0 in range(0, sample, batch):
                sel = pick[b0:b0 + batch]
                # weight each structure by how often this batch drew it
                cnt = torch.zeros(A_trn.ns, dtype=torch.float64, device=device).index_add_(
                    0, sel, torch.ones(len(sel), dtype=torch.float64, device=device))
                if b0 > 0:
                    dE = A_trn.errors(model)
                grad.zero_()
                A_trn.grad(model, dE * cnt, grad)
                model.adam(grad, m, v, step, p["mu"], p["b1"], p["b2"], p["eps"])
                step += 1
            hist.append(report(ep))
            save(ep)
    else:
        # full-batch limited-memory BFGS (20 corrections), one evaluation per epoch
        x = model.weights().double()
        mem_s, mem_y = [], []

        def fg(w):
            model.set_weights(w.float())
            dE = A_trn.errors(model)
            grad.zero_()
            A_trn.grad(model, dE, grad)
            return 0.5 * float((dE * dE).sum()), grad.double().clone()

        f, g = fg(x)
        d = -g
        alpha = 1.0 / max(float(g.abs().max()), 1e-12) * 1e-2
        for ep in range(1, inp["iterations"] + 1):
            gd = float((g * d).sum())
            xn = x + alpha * d
            fn, gn = fg(xn)
            if fn <= f + 1e-4 * alpha * gd or alpha < 1e-12:
                s_, y_ = xn - x, gn - g
                if float((s_ * y_).sum()) > 1e-12:
                    mem_s.append(s_)
                    mem_y.append(y_)
                    if len(mem_s) > 20:
                        mem_s.pop(0)
                        mem_y.pop(0)
                x, f, g = xn, fn, gn
                # two-loop recursion for the next direction
                q = g.clone()
                a_ = []
                for s_, y_ in zip(reversed(mem_s), reversed(mem_y)):
                    r_ = 1.0 / float((y_ * s_).sum())
                    ai = r_ * float((s_ * q).sum())
                    a_.append((ai, r_, s_, y_))
                    q -= ai * y_
                if mem_s:
                    q *= float((mem_s[-1] * mem_y[-1]).sum()) / float((mem_y[-1] * mem_y[-1]).sum())
                else:
                    q *= 1e-2 / max(float(g.abs().max()), 1e-12)
                for ai, r_, s_, y_ in reversed(a_):
                    b = r_ * float((y_ * q).sum())
                    q += s_ * (ai - b)
                d = -q
                if float((g * d).sum()) >= 0.0:
                    mem_s.clear()
                    mem_y.clear()
                    d = -g * (1e-2 / max(float(g.abs().max()), 1e-12))
                alpha = 1.0
            else:
                alpha *= 0.5
                model.set_weights(x.float())
            hist.append(report(ep))
            save(ep)
        model.set_weights(x.float())
    model.check()
    save()
    log("")
    log(" Training finished (%.2f s on the GPU).  Networks written to %s"
        % (time.perf_counter() - t0, ", ".join(rel(nets[t]["file"]) for t in types)))
    return hist


def main():
    if len(sys.argv) < 2:
        print("train -- network training on the GPU.\nUsage: python -m aenet_amd.train <input-file>")
        sys.exit(1)
    run(sys.argv[1])


if __name__ == "__main__":
    main()
