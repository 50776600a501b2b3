#!/usr/bin/env python3
"""Energy+force throughput of the 4000-atom binary alloy on one MI355X.

One step = one full pass of the prediction path over the structure held in
HBM: Cartesian -> lattice coordinates, cell binning, ordered FP64 neighbor
search, Chebyshev descriptors (radial 18 / 6.5 A, angular 6 / 4.5 A,
typespin), per-species 52-30-30-1 network forward and backward, forces
(64-bit fixed-point sums, converted to FP64) and per-structure energy sums.

Prints one JSON line.  ``roofline`` is for the fused atom kernel, timed with
HIP events on its own stream; ``cpu_baseline`` is the reference library
(built from its sources, oracle/_ref) run the way predict.x runs it, in
worker processes over the host cores, on a bounded sample of atoms.
"""
import argparse
import json
import multiprocessing as mp
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "atom energy+force evals/s (1 GPU, FP32 vs FP64 ref); DA-MC vs plain accepted moves/s"
FP32_PEAK = 157.3e12      # MI355X dense FP32 vector, FLOP/s
HBM_PEAK = 8.0e12         # B/s
# algorithmic FP32 FLOPs (FMA = 2) per unit of the Chebyshev path with
# N_r = 18, N_a = 6 and typespin (m = 2): per pair 24 + (N_r+1)(12+4(m-1)),
# per triplet 61 + (N_a+1)(11+4(m-1)), network 4*Wsize per atom, 3*nsf
# for the normalization
FLOP_PAIR = 24 + 19 * 16
FLOP_TRIPLET = 61 + 7 * 15
WSIZE = 53 * 30 + 31 * 30 + 31
FLOP_ATOM = 4 * WSIZE + 3 * 52


def pot_files():
    d = os.path.join(ROOT, "tests", "golden", "potentials")
    return {"A": os.path.join(d, "alloy.A.nn"), "B": os.path.join(d, "alloy.B.nn")}


# ---------------------------------------------------------------------------
# CPU baseline: the reference library, one process per core
# ---------------------------------------------------------------------------

def _cpu_worker(args):
    rank, nranks, reps = args
    import ctypes as C
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_lib import Reference
    from aenet_amd import synth
    s = synth.c2_alloy4000()
    ref = Reference(["A", "B"], pot_files())
    lib = ref.lib()
    lib.refdrv_timed.argtypes = [np.ctypeslib.ndpointer(np.float64), C.c_int, C.c_int,
                                 np.ctypeslib.ndpointer(np.int32), np.ctypeslib.ndpointer(np.float64), C.c_int,
                                 C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_double),
                                 np.ctypeslib.ndpointer(np.float64), np.ctypeslib.ndpointer(np.float64),
                                 C.POINTER(C.c_int)]
    n = len(s)
    frac = np.zeros((n, 3))
    L = np.zeros(9)
    lib.refdrv_cart_to_frac(np.ascontiguousarray(s.cell).ravel(), 1, n, np.ascontiguousarray(s.positions).ravel(),
                            frac, L)
    t_nb = t_ev = 0.0
    E = np.full(n, np.nan)
    F = np.zeros((n, 3))
    for _ in range(reps):
        tn, te, et = C.c_double(), C.c_double(), C.c_double()
        st = C.c_int()
        lib.refdrv_timed(L, 1, n, np.ascontiguousarray(s.types, np.int32), frac, rank, nranks, C.byref(tn),
                         C.byref(te), C.byref(et), E, F, C.byref(st))
        t_nb += tn.value
        t_ev += te.value
    count = reps * len(range(rank, n, nranks))
    return count, t_nb, t_ev, E, F


def cpu_baseline(procs, stride, reps):
    """Reference library on `procs` worker processes; every worker takes
    the atoms i = rank (mod procs*stride), so procs*stride sets the sample,
    and evaluates them reps times."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_lib import Reference
    if not Reference.available():
        raise RuntimeError("oracle/_ref/libaenet_ref.so not built (make -C oracle ref)")
    ctx = mp.get_context("fork")
    nranks = procs * stride
    t0 = time.time()
    with ctx.Pool(procs) as pool:
        res = pool.map(_cpu_worker, [(r, nranks, reps) for r in range(procs)])
    wall = time.time() - t0
    atoms = sum(r[0] for r in res)
    # the reference's per-atom energies of the sampled atoms and, when every
    # atom is sampled, its forces (the workers' contributions summed)
    import numpy as np
    E_ref = np.full(len(res[0][3]), np.nan)
    for r in res:
        m = ~np.isnan(r[3])
        E_ref[m] = r[3][m]
    F_ref = sum(r[4] for r in res) if stride == 1 else None
    t_eval = max(r[2] for r in res)
    t_e2e = max(r[1] + r[2] for r in res)
    cpu_s = sum(r[1] + r[2] for r in res)
    return {
        "value": atoms / t_eval,
        "unit": "atoms/s",
        "cores": procs,
        "kind": "reference",
        "sample": "%d atom evaluations: every %dth of the 4000 atoms, %d passes, %d worker processes over "
                  "disjoint atoms; value = aenet_atomic_energy_and_forces on precomputed lclist neighbor lists "
                  "(kernel-equivalent); end_to_end includes lcl_init + lcl_nbdist_cart as predict.x runs them"
                  % (atoms, stride, reps, procs),
        "end_to_end": atoms / t_e2e,
        "cpu_seconds": cpu_s,
        "wall_seconds": wall,
        "host_cores_total": os.cpu_count(),
        "host_cores_affinity": sys_cores(),
        "cores_note": "worker processes capped at 16, the host CPU share of one GPU on these machines",
        "_E_ref": E_ref,
        "_F_ref": F_ref,
    }


def sys_cores():
    try:
        return len(os.sched_getaffinity(0))
    except Exception:
        return os.cpu_count() or 1


# ---------------------------------------------------------------------------
# GPU
# ---------------------------------------------------------------------------

def pair_triplet_counts(db, rc_r=6.5, rc_a=4.5):
    import torch
    nb = db.neighbors(256)
    n = nb["nnb"].long()
    d = nb["dist"]
    idx = torch.arange(d.shape[1], device=d.device)[None, :]
    valid = idx < n[:, None]
    n_p = ((d <= rc_r) & (d > 1e-12) & valid).sum(dim=1)
    m_a = ((d <= rc_a) & (d >= 1e-12) & valid).sum(dim=1)
    n_t = m_a * (m_a - 1) // 2
    return int(n_p.sum().item()), int(n_t.sum().item())


# Behler-Parrinello batch (2048 structures of 8-512 atoms, 3 species, 8 G2
# per neighbor species and 18 G4 per species pair at Rc = 6 A, 132-25-25-1):
# per pair 24 + 14*n_G2, per contributing triplet (all legs < Rc)
# 40 + 16*n_G4, network 4*Wsize + 3*nsf per atom
BP_FLOP_PAIR = 24 + 14 * 8
BP_FLOP_TRIPLET = 40 + 16 * 18
BP_WSIZE = 133 * 25 + 26 * 25 + 26
BP_FLOP_ATOM = 4 * BP_WSIZE + 3 * 132


def launch_shape(chosen, rows):
    """The tuner's choice and every candidate it timed (aenet_amd_batch_autotune);
    tile = 1: the networks of a workgroup's atoms as one matrix-core tile."""
    return {"chosen": {"waves_per_workgroup": chosen[0], "weights_in_lds": chosen[1], "tile_network": chosen[2]},
            "candidates_us": [{"waves_per_workgroup": w, "weights_in_lds": s, "tile_network": t, "kernel_us": ms * 1e3}
                              for (w, s, t, ms) in rows]}


def bp_counts(db, rc=6.0, chunk=2048):
    """Pairs with r < rc and triplets with all three legs < rc, from the
    batch's own neighbor lists (counted on the GPU)."""
    import torch
    nb = db.neighbors(160)
    n = nb["nnb"].long()
    pairs = triplets = 0
    for a0 in range(0, db.natoms, chunk):
        coo = nb["coo"][a0:a0 + chunk]
        d = nb["dist"][a0:a0 + chunk]
        cnt = n[a0:a0 + chunk]
        valid = torch.arange(d.shape[1], device=d.device)[None, :] < cnt[:, None]
        inside = valid & (d < rc)
        pairs += int(inside.sum().item())
        djk = torch.cdist(coo, coo)
        m = inside[:, :, None] & inside[:, None, :] & (djk < rc)
        triplets += int(torch.triu(m.to(torch.int8), diagonal=1).sum().item())
    return pairs, triplets


def batched_bp(steps, warmup):
    """Throughput and atom-kernel roofline of the batched BP workload."""
    import torch
    from aenet_amd import synth
    from aenet_amd.core import ANNPotentials
    from aenet_amd.device import DeviceBatch
    d = os.path.join(ROOT, "tests", "golden", "potentials")
    ann = ANNPotentials({t: os.path.join(d, "bp3.%s.nn" % t) for t in "ABC"})
    try:
        structs = synth.batch_uneven(2048)
        db = DeviceBatch(structs)
        chosen, rows = db.autotune(forces=True, repeats=2)
        for _ in range(warmup):
            db.eval(check=False)
        db.check()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            db.eval(check=False)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        db.check()
        ms = (t1 - t0) * 1e3 / steps
        db.set_timing(True)
        for _ in range(steps):
            db.eval(check=False)
        kms, kcount = db.kernel_time()
        db.set_timing(False)
        k_s = kms / kcount * 1e-3
        n_p, n_t = bp_counts(db)
        flops = n_p * BP_FLOP_PAIR + n_t * BP_FLOP_TRIPLET + db.natoms * BP_FLOP_ATOM
        out = {
            "workload": "batch2048: 2048 structures, 8-512 atoms (log-uniform), bulk FCC/BCC, slabs, clusters; "
                        "3 species; BP 24 G2 + 108 G4 at Rc 6 A; 132-25-25-1",
            "structures": db.nstruct,
            "atoms_per_step": db.natoms,
            "ms_per_step": ms,
            "atoms_per_s": db.natoms / (ms * 1e-3),
            "kernel": "k_atom_bp<0,true,false>",
            "kernel_ms": k_s * 1e3,
            "pairs": n_p,
            "contributing_triplets": n_t,
            "flop_per_launch": flops,
            "achieved_tflops": flops / k_s / 1e12,
            "frac_fp32_peak": flops / k_s / FP32_PEAK,
            "launch_shape": launch_shape(chosen, rows),
        }
        plan = db.buckets()
        if plan is not None:
            # kernel_ms then covers the bucketing launch and one atom-kernel
            # launch per neighbor-count bucket
            out["kernel"] = "k_bucket_atoms + k_atom_bp<0,true,false> per bucket"
            out["neighbor_count_buckets"] = {
                "caps": plan["caps"], "atoms": plan["atoms"],
                "shapes": [{"waves_per_workgroup": w, "weights_in_lds": st, "tile_network": t}
                           for (w, st, t) in plan["shapes"]],
                "bucket_kernel_us": [m * 1e3 for m in plan["ms"]],
                "tuner_single_launch_us": plan["ms_single"] * 1e3, "tuner_buckets_us": plan["ms_buckets"] * 1e3}
        else:
            out["neighbor_count_buckets"] = None
        del db
        return out
    finally:
        ann.release()


def mc_throughput(steps, burn):
    """Accepted Monte Carlo moves per second, plain Metropolis with the
    teacher potential against delayed acceptance with the trained
    radial-only student as first stage: 512 chains of the 256-atom binary
    FCC cell (a = 3.61 A), 600 K, 50 % swaps / 50 % displacements
    U(+-0.075 A)^3, both from the same burned-in states."""
    from aenet_amd import synth
    from aenet_amd.mc import MonteCarlo
    from aenet_amd.model import Model
    d = os.path.join(ROOT, "tests", "golden", "potentials")
    teacher = Model({t: os.path.join(d, "teacher.%s.nn" % t) for t in "AB"})
    student = Model({t: os.path.join(d, "student_trained.%s.nn" % t) for t in "AB"})
    cells = [synth.mc_cell(100 + c) for c in range(512)]
    out = {"workload": "mc512x256: 512 chains x 256 atoms (4x4x4 FCC A/B, a=3.61 A), 600 K, 50% swaps, "
                       "50% displacements U(+-0.075 A)^3; teacher Chebyshev 16/6.0 + 4/4.0 typespin 44-20-20-1; "
                       "first stage radial-only Chebyshev 8/5.0 typespin 20-8-1 trained on teacher MC data (atomic energies and "
                       "forces, then the energy changes of proposals)",
           "steps": steps, "burn_in": burn}
    for name, first, delayed in (("plain", None, False), ("delayed", student, True)):
        mc = MonteCarlo(teacher, cells, first=first, seed=7)
        mc.set(T=600.0, max_screen=4)
        mc.run(burn)
        mc.stats(reset=True)
        t = mc.run(steps, delayed=delayed)
        st = mc.stats()
        out[name] = {"accepted_moves_per_s": st["accepted"] / t, "proposals_per_s": st["proposed"] / t,
                     "acceptance": st["accepted"] / st["proposed"],
                     "full_potential_fraction": st["full_evaluations"] / st["proposed"],
                     "second_stage_acceptance": st["accepted"] / st["full_evaluations"],
                     "proposals_per_chain_step": st["proposed"] / (len(cells) * steps),
                     "ms_per_step": t / steps * 1e3}
        if delayed:
            out[name]["max_screen"] = 4
        del mc
    out["delayed_over_plain"] = out["delayed"]["accepted_moves_per_s"] / out["plain"]["accepted_moves_per_s"]
    return out


def per_atom_abi(ann, db, cpu, natoms=400):
    """The per-atom drop-in (aenet_atomic_energy_and_forces, one call per
    atom with its neighbor list from host memory, as predict.x and the ASE
    calculator call it) on the first atoms of the alloy, against the
    reference library's per-atom rate on one core (its kernel-equivalent
    rate over the worker processes, divided by their number)."""
    import numpy as np
    nb = db.neighbors(256)
    nnb = nb["nnb"].cpu().numpy()
    coo = nb["coo"].cpu().numpy()
    lst = nb["list"].cpu().numpy()
    typ = nb["type"].cpu().numpy()
    x = db.coords.view(-1, 3).cpu().numpy()
    t_at = db.types.cpu().numpy()
    names = ["A", "B"]
    F = np.zeros((db.natoms, 3))
    # lattice-wrapped centre: the lists hold the image coordinates
    L = np.asarray(synth_cell(), dtype=np.float64)
    frac = x @ np.linalg.inv(L)
    ci = (frac - np.floor(frac)) @ L
    for i in range(20):   # warm-up
        k = nnb[i]
        ann.atomic_energy_and_forces(ci[i], names[t_at[i] - 1], i + 1, coo[i, :k],
                                     [names[t - 1] for t in typ[i, :k]], lst[i, :k], F)
    t0 = time.perf_counter()
    for i in range(natoms):
        k = nnb[i]
        ann.atomic_energy_and_forces(ci[i], names[t_at[i] - 1], i + 1, coo[i, :k],
                                     [names[t - 1] for t in typ[i, :k]], lst[i, :k], F)
    dt = time.perf_counter() - t0
    out = {"workload": "aenet_atomic_energy_and_forces once per atom, %d atoms of alloy4000, neighbor lists "
                       "in host memory (one launch and two copies per call)" % natoms,
           "gpu_atoms_per_s": natoms / dt, "us_per_call": dt / natoms * 1e6}
    if cpu and cpu.get("value"):
        out["reference_atoms_per_s_one_core"] = cpu["value"] / cpu["cores"]
    return out


def synth_cell():
    from aenet_amd import synth
    return synth.c2_alloy4000().cell


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--replicas", type=int, default=1,
                    help="independent 4000-atom cells per step (default 1: the single alloy cell)")
    ap.add_argument("--cpu-procs", type=int, default=0, help="CPU baseline worker processes (0: host cores, <=16)")
    ap.add_argument("--cpu-stride", type=int, default=1, help="sample = every (procs*stride)th atom")
    ap.add_argument("--cpu-passes", type=int, default=2, help="CPU baseline passes over the sample")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched Behler-Parrinello measurement")
    ap.add_argument("--no-mc", action="store_true", help="skip the Monte Carlo measurement")
    ap.add_argument("--mc-steps", type=int, default=3000, help="Monte Carlo steps per chain after the burn-in")
    ap.add_argument("--mc-burn", type=int, default=1000)
    ap.add_argument("--mc-only", action="store_true", help="only the Monte Carlo measurement (its own JSON line)")
    ap.add_argument("--no-per-atom", action="store_true", help="skip the per-atom C-ABI measurement")
    ap.add_argument("--launch", default=None, metavar="W,S,T",
                    help="atom-kernel launch shape instead of the tuner's: wavefronts per workgroup, weights "
                         "in LDS (0/1), tile network (0/1) (aenet_amd_batch_set_launch)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the occupancy model's launch shape instead of the measured one")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only copy a MIB-MiB device buffer 5 times (reference kernel for the "
                         "FETCH_SIZE/WRITE_SIZE counters) and exit")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        n = args.copy_calibration * (1 << 20) // 4
        x = torch.ones(n, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        for _ in range(5):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": 4 * n, "copy_bytes_written": 4 * n, "launches": 5}))
        return
    if args.gpus != 1:
        print("bench.py measures one GPU only: the prediction path runs on a single MI355X and this build "
              "has no multi-GPU mode (--gpus %d requested)" % args.gpus, file=sys.stderr)
        sys.exit(2)
    if args.mc_only:
        print(json.dumps({"mc": mc_throughput(args.mc_steps, args.mc_burn)}))
        return

    cpu = None
    if not args.no_cpu_baseline:
        # at most 16 workers: the CPU share of one GPU on these machines
        procs = args.cpu_procs or min(16, sys_cores())
        try:
            cpu = cpu_baseline(procs, args.cpu_stride, args.cpu_passes)
        except Exception as e:  # reference not built: report, do not fake
            cpu = {"value": None, "unit": "atoms/s", "cores": procs, "kind": "reference",
                   "sample": "unavailable: %s" % e}

    import torch
    from aenet_amd import synth
    from aenet_amd.core import ANNPotentials
    from aenet_amd.device import DeviceBatch

    torch.cuda.set_device(0)
    ann = ANNPotentials(pot_files())
    if args.replicas == 1:
        structs = [synth.c2_alloy4000()]
    else:
        structs = synth.alloy_replicas(args.replicas)
    db = DeviceBatch(structs)
    natoms = db.natoms
    tune = None
    if args.launch:
        w, st, tl = (int(v) for v in args.launch.split(","))
        db.set_launch(w, st, tl)
        tune = {"chosen": {"waves_per_workgroup": w, "weights_in_lds": st, "tile_network": tl}, "set": True}
    elif not args.no_autotune:
        # the atom kernel's launch shape by measurement (aenet_amd_batch_autotune)
        chosen, rows = db.autotune(forces=True, repeats=10)
        tune = launch_shape(chosen, rows)
    # eager launches (five kernels per step)
    for _ in range(args.warmup):
        db.eval(check=False)
    db.check()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        db.eval(check=False)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    db.check()
    eager_ms = (t1 - t0) * 1e3 / args.steps
    E_eager = db.E_atoms.clone()
    # the headline: one step captured in a HIP graph and replayed (the
    # step allocates and clears nothing, so it captures whole)
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        for _ in range(2):
            db.eval(check=False)
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        db.eval(check=False)
    for _ in range(args.warmup):
        graph.replay()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        graph.replay()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    db.check()
    graph_ms = (t1 - t0) * 1e3 / args.steps
    # the headline is the faster of the two ways of issuing the same step
    ms = min(graph_ms, eager_ms)
    value = natoms / (ms * 1e-3)
    if not torch.equal(db.E_atoms, E_eager):
        raise RuntimeError("graph replay and eager evaluation differ")
    E_tot = float(db.E_struct[0].item())

    # roofline of the fused atom kernel: HIP events around its launches
    db.set_timing(True)
    nt = min(args.steps, 100)
    for _ in range(nt):
        db.eval(check=False)
    kms, kcount = db.kernel_time()
    sms, scount = db.search_time()
    db.set_timing(False)
    k_avg_s = kms / kcount * 1e-3
    s_avg_s = sms / scount * 1e-3
    n_p, n_t = pair_triplet_counts(db)
    entries = int(db.neighbors(256)["nnb"].sum().item())
    # neighbor search: algorithmic bytes = the list it writes (16 B of
    # float4 R + 4 B of index per entry) + 4 B of count per atom + the
    # coordinates and types it reads once (28 B per atom)
    search_bytes = 20 * entries + 32 * natoms
    flops = n_p * FLOP_PAIR + n_t * FLOP_TRIPLET + natoms * FLOP_ATOM
    achieved = flops / k_avg_s
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "k_atom_cheb_alloy4000_traffic.json")
    tile = bool(tune and tune["chosen"].get("tile_network"))
    kname = "k_atom_cheb<19,7,0,true,%s,false>" % ("true" if tile else "false")
    if args.replicas == 1 and os.path.exists(tpath):
        t = json.load(open(tpath))["kernels"].get(kname)
        if t:
            traffic = t["read_bytes"] + t["write_bytes"]

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "atoms/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "alloy4000: 4000-atom binary A/B FCC cell (10x10x10, a=3.8 A, sigma=0.08 A), "
                        "Chebyshev radial 18/6.5 A + angular 6/4.5 A with typespin (52 features), "
                        "52-30-30-1 tanh network per species",
            "structures_per_step": len(structs),
            "atoms_per_step": natoms,
            "E_tot_eV": E_tot,
        },
        "roofline": {
            "bound": "valu",
            "pipe": "FP32 VALU: the atom kernel issues no MFMA; peak = MI355X FP32 vector 157.3 TFLOP/s",
            "kernel": kname,
            "achieved": achieved / 1e12,
            "peak": FP32_PEAK / 1e12,
            "unit": "TFLOP/s",
            "frac": achieved / FP32_PEAK,
            "traffic": traffic,
            "traffic_source": "rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE of this workload, scaled by a 256 MiB copy "
                              "(profiles/k_atom_cheb_alloy4000_traffic.md)" if traffic else None,
            "kernel_ms": k_avg_s * 1e3,
            "flop_per_launch": flops,
            "pairs": n_p,
            "triplets": n_t,
            "launch_shape": tune,
        },
        "step": {
            "timing": "the step (binning, search, atom kernel, structure sums: five launches) issued one by "
                      "one (eager) and as one captured HIP graph replayed (graph); value and ms_per_step "
                      "use the faster, here %s" % ("graph" if graph_ms <= eager_ms else "eager"),
            "graph_ms_per_step": graph_ms,
            "eager_ms_per_step": eager_ms,
        },
        "neighbor_search": {
            "kernel": "k_search_cells",
            "bound": "hbm",
            "kernel_ms": s_avg_s * 1e3,
            "list_entries": entries,
            "bytes_per_launch": search_bytes,
            "achieved_GBps": search_bytes / s_avg_s / 1e9,
            "peak_GBps": HBM_PEAK / 1e9,
            "frac": search_bytes / s_avg_s / HBM_PEAK,
        },
        "cpu_baseline": cpu,
    }
    if cpu is not None:
        # the GPU step against the reference's own numbers for the sampled
        # atoms (energies) and, when every atom is sampled, the forces
        import numpy as np
        E_ref, F_ref = cpu.pop("_E_ref", None), cpu.pop("_F_ref", None)
        if E_ref is not None:
            m = ~np.isnan(E_ref)
            Eg = db.E_atoms.cpu().numpy()
            cpu["parity"] = {
                "atoms_compared": int(m.sum()),
                "max_abs_dE_eV_per_atom": float(np.abs(Eg[m] - E_ref[m]).max()),
                "max_abs_dF_eV_per_A": (float(np.abs(db.forces.cpu().numpy() - F_ref).max())
                                        if F_ref is not None else None),
                "bars": "1e-5 eV/atom, 1e-4 eV/A",
            }
    if cpu and cpu.get("value"):
        out["gpu_over_cpu"] = value / cpu["value"]
        out["gpu_over_cpu_end_to_end"] = value / cpu["end_to_end"]
    if not args.no_per_atom:
        out["per_atom_abi"] = per_atom_abi(ann, db, cpu)
    ann.release()
    if not args.no_batch:
        out["batched_bp"] = batched_bp(max(5, min(args.steps, 20)), 2)
    if not args.no_mc:
        out["mc"] = mc_throughput(args.mc_steps, args.mc_burn)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
