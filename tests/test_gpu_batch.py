"""Batched uneven structures on the GPU: 2048 Behler-Parrinello structures of
8-512 atoms (periodic bulk FCC/BCC, slabs, isolated clusters; 3 species)
through one persistent launch of each kernel.

- a subset against the FP64 oracle at the per-atom bars (1e-5 eV/atom,
  1e-4 eV/A);
- the full batch through size-independent properties: per-structure
  energies are the sums of the atomic energies, forces of every structure
  sum to zero (translation invariance), and a structure's atomic energies
  do not depend on which batch it is evaluated in;
- the evaluation captured in a HIP graph and replayed, every replay checked
  against an eager evaluation of the same coordinates.
"""
import numpy as np
import pytest

from conftest import pot_files

pytestmark = pytest.mark.gpu

E_TOL = 1e-5
F_TOL = 1e-4
NAMES = ["A", "B", "C"]


@pytest.fixture(scope="module")
def structures():
    from aenet_amd import synth
    return synth.batch_uneven(2048)


def _subset(structs, nmax, count):
    out = []
    for s in structs:
        if len(s.types) <= nmax:
            out.append(s)
        if len(out) == count:
            break
    return out


def test_uneven_subset_matches_oracle(structures, device):
    from aenet_amd.core import ANNPotentials
    from oracle_lib import Oracle
    sub = _subset(structures, 64, 24)
    kinds = sorted({s.label for s in sub})
    print("subset: %d structures, %d atoms, kinds %s" % (len(sub), sum(len(s.types) for s in sub), kinds),
          flush=True)
    assert len(kinds) == 4
    ann = ANNPotentials(pot_files("bp3", NAMES))
    try:
        r = ann.structure_energy_and_forces([s.cell for s in sub], [s.pbc for s in sub],
                                            [s.positions for s in sub], [list(s.types) for s in sub])
    finally:
        ann.release()
    o = Oracle(NAMES, pot_files("bp3", NAMES))
    off = 0
    worst_e = worst_f = 0.0
    for k, s in enumerate(sub):
        n = len(s.types)
        ref = o.structure(s.cell, s.pbc, s.types, s.positions)
        worst_e = max(worst_e, np.abs(r["energies"][off:off + n] - ref["energies"]).max())
        worst_f = max(worst_f, np.abs(r["forces"][off:off + n] - ref["forces"]).max())
        assert abs(r["E_tot"][k] - ref["Etot"]) < E_TOL * n
        assert abs(r["E_coh"][k] - ref["Ecoh"]) < E_TOL * n
        off += n
    print("max |dE| %.3e eV/atom  max |dF| %.3e eV/A" % (worst_e, worst_f))
    assert worst_e < E_TOL
    assert worst_f < F_TOL


def test_full_batch_properties(structures, device):
    import torch
    from aenet_amd.core import ANNPotentials
    from aenet_amd.device import DeviceBatch
    ann = ANNPotentials(pot_files("bp3", NAMES))
    try:
        db = DeviceBatch(structures, device=device)
        print("batch: %d structures, %d atoms" % (db.nstruct, db.natoms), flush=True)
        E_tot, E_coh, E_at, F = db.eval()
        E_tot = E_tot.cpu().numpy()
        E_at = E_at.cpu().numpy()
        F = F.cpu().numpy()
        torch.cuda.synchronize()
        off = db.offsets
        worst_sum = 0.0
        for k in range(db.nstruct):
            e = E_at[off[k]:off[k + 1]]
            assert abs(E_tot[k] - e.sum()) <= 1e-9 * max(1.0, abs(E_tot[k]))
            fsum = np.abs(F[off[k]:off[k + 1]].sum(axis=0)).max()
            worst_sum = max(worst_sum, fsum)
        print("max |sum F| over structures %.3e eV/A" % worst_sum)
        # exactly zero in the kernels' fixed-point units (F_i takes minus the
        # integer sum of the values given to its neighbors); what remains is
        # the FP64 rounding of the sum above
        assert worst_sum < 1e-10
        assert np.all(np.isfinite(F))
        del db
        # the same structures alone or in a different batch give the same atomic energies
        picks = [0, 1, 2, 3, 100, 777, 2047]
        db2 = DeviceBatch([structures[k] for k in picks], device=device)
        _, _, e2, _ = db2.eval()
        e2 = e2.cpu().numpy()
        o2 = db2.offsets
        for i, k in enumerate(picks):
            assert np.array_equal(e2[o2[i]:o2[i + 1]], E_at[off[k]:off[k + 1]])
        del db2
    finally:
        ann.release()


def test_graph_capture_replays_match_eager(structures, device):
    import torch
    from aenet_amd.core import ANNPotentials
    from aenet_amd.device import DeviceBatch
    ann = ANNPotentials(pot_files("bp3", NAMES))
    try:
        # periodic structures: random moves keep every atom valid (an isolated
        # structure must stay inside the bounding box its batch was made for)
        sub = [s for s in structures[:400] if s.pbc]
        db = DeviceBatch(sub, device=device)
        x0 = db.coords.clone()
        s = torch.cuda.Stream(device=device)
        s.wait_stream(torch.cuda.current_stream(device))
        with torch.cuda.stream(s):
            for _ in range(2):
                db.eval(check=False)
        torch.cuda.current_stream(device).wait_stream(s)
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            db.eval(check=False)
        gen = torch.Generator(device=device).manual_seed(5)
        for rep in range(4):
            # new coordinates, replay, keep the results, then eager on the same coordinates
            db.coords.copy_(x0 + 0.02 * torch.randn(x0.shape, generator=gen, device=device, dtype=torch.float64))
            g.replay()
            torch.cuda.synchronize()
            E_g = db.E_struct.clone()
            A_g = db.E_atoms.clone()
            F_g = db.forces.clone()
            db.eval()
            torch.cuda.synchronize()
            assert torch.equal(A_g, db.E_atoms), rep
            assert torch.allclose(E_g, db.E_struct, rtol=0, atol=1e-9), rep
            assert torch.allclose(F_g, db.forces, rtol=0, atol=1e-10), rep
        db.check()
        del g
        del db
    finally:
        ann.release()


def test_autotune_keeps_results(device):
    """The measured launch shape changes nothing but the time: with the
    network per wavefront, per-atom energies are bitwise those of the
    occupancy model's shape (one wavefront computes each atom either way),
    and so are the forces (fixed-point sums, independent of their order); with the tile
    network (its FP32 products summed in another order) both within the
    per-atom bars.  Every candidate was timed (the tile network among them
    where the species share one architecture) and the kept one is the
    fastest."""
    import torch
    from aenet_amd import synth
    from aenet_amd.core import ANNPotentials
    from aenet_amd.device import DeviceBatch
    from conftest import pot_files
    cases = [(pot_files("alloy", ["A", "B"]), [synth.binary_fcc(3.8, (4, 4, 4), 0.08, 41, 42)]),
             (pot_files("bp3", ["A", "B", "C"]), synth.batch_uneven(24))]
    for pots, structs in cases:
        ann = ANNPotentials(pots)
        try:
            db = DeviceBatch(structs, device=device)
            _, _, E0, F0 = db.eval()
            E0, F0 = E0.clone(), F0.clone()
            chosen, rows = db.autotune(forces=True, repeats=2)
            print("chosen", chosen, "rows", rows)
            assert rows and chosen in [(w, s, t) for (w, s, t, _) in rows]
            assert min(rows, key=lambda r: r[3])[:3] == chosen
            assert any(t == 1 for (_, _, t, _) in rows) == (pots == pot_files("alloy", ["A", "B"]))
            _, _, E1, F1 = db.eval()
            if chosen[2] == 0:
                assert torch.equal(E0, E1)
                assert torch.equal(F0, F1)
            else:
                assert (E1 - E0).abs().max().item() < 2 * E_TOL
                assert (F1 - F0).abs().max().item() < 2 * F_TOL
        finally:
            ann.release()


def test_launch_shapes_and_tile_network(device):
    """Every launch shape of the Chebyshev atom kernel set explicitly
    (aenet_amd_batch_set_launch), with the network per wavefront and as one
    matrix-core tile per workgroup round, against the occupancy model's
    shape: per-wavefront network bitwise (energies and forces); tile network
    within the per-atom bars (FP32 products
    summed in another order).  Structures: 500 atoms (the last workgroup
    partly filled) and five 4000-atom cells (several rounds per workgroup).
    Before each checked evaluation the same shape evaluates displaced
    coordinates, so that values an evaluation fails to write would show as
    the other coordinates' values."""
    import torch
    from aenet_amd import synth
    from aenet_amd.core import ANNPotentials
    from aenet_amd.device import DeviceBatch
    ann = ANNPotentials(pot_files("alloy", ["A", "B"]))
    try:
        for structs in ([synth.binary_fcc(3.8, (5, 5, 5), 0.08, 43, 44)], synth.alloy_replicas(5)):
            db = DeviceBatch(structs, device=device)
            _, _, E0, F0 = db.eval()
            E0, F0 = E0.clone(), F0.clone()
            other = db.coords + 0.05 * torch.sin(torch.arange(db.coords.numel(), device=db.coords.device,
                                                              dtype=torch.float64)).reshape(db.coords.shape)
            for tile in (0, 1):
                for waves in (1, 2, 4, 8, 16):
                    for staged in (0, 1):
                        print("shape", len(structs), waves, staged, tile, flush=True)
                        db.set_launch(waves, staged, tile)
                        _, _, Eo, _ = db.eval(coords=other)
                        assert (Eo - E0).abs().max().item() > 1e-3
                        _, _, E1, F1 = db.eval()
                        if tile == 0:
                            assert torch.equal(E0, E1)
                            assert torch.equal(F0, F1)
                        else:
                            assert (E1 - E0).abs().max().item() < 2 * E_TOL
                            assert (F1 - F0).abs().max().item() < 2 * F_TOL
            db.set_launch(0)
            _, _, E2, _ = db.eval()
            assert torch.equal(E0, E2)
            with pytest.raises(Exception):
                db.set_launch(3, 0, 0)
            del db
    finally:
        ann.release()
    # no tile network for Behler-Parrinello potentials
    ann = ANNPotentials(pot_files("bp3", NAMES))
    try:
        db = DeviceBatch(synth.batch_uneven(8), device=device)
        with pytest.raises(Exception):
            db.set_launch(4, 0, 1)
        db.set_launch(4, 0, 0)
        del db
    finally:
        ann.release()


def test_neighbor_count_buckets(device, structures):
    """Neighbor-count buckets (aenet_amd_batch_set_buckets): every bucket's
    atoms evaluated by a launch of their own, with LDS regions of the
    bucket's cap and their own launch shape, give bitwise the energies and
    forces of the single launch (per-wavefront network; per-atom results
    and 64-bit fixed-point force sums do not depend on the launch).  For
    the 2048-structure Behler-Parrinello batch (clusters of 8 atoms to
    512-atom bulk) and an uneven two-species Chebyshev batch, with an empty
    bucket, and with other coordinates evaluated in between so that values
    a launch fails to write would show.  The tuner's plan (kept only where
    it measures faster) reproduces them too."""
    import torch
    from aenet_amd import synth
    from aenet_amd.core import ANNPotentials
    from aenet_amd.device import DeviceBatch
    cases = ((pot_files("bp3", NAMES), structures, [(1, 0, 0), (4, 0, 0), (2, 1, 0), (4, 1, 0)]),
             (pot_files("alloy", ["A", "B"]), synth.batch_uneven(400, ntypes=2),
              [(16, 1, 0), (4, 0, 0), (8, 1, 0), (1, 0, 0)]))
    for pots, structs, shapes in cases:
        ann = ANNPotentials(pots)
        try:
            db = DeviceBatch(structs, device=device)
            _, _, E0, F0 = db.eval()
            E0, F0 = E0.clone(), F0.clone()
            other = db.coords + 0.03 * torch.sin(torch.arange(db.coords.numel(), device=db.coords.device,
                                                              dtype=torch.float64)).reshape(db.coords.shape)
            for caps in ([16, 32, 64, 512], [8, 512], [24, 40, 48, 56, 60, 512]):
                sh = [shapes[k % len(shapes)] for k in range(len(caps))]
                print("buckets", db.natoms, caps, sh, flush=True)
                db.set_buckets(caps, sh)
                plan = db.buckets()
                assert plan["caps"][:-1] == caps[:-1] and plan["caps"][-1] > caps[-2]
                assert plan["atoms"] == [-1] * len(caps)
                _, _, Eo, _ = db.eval(coords=other)
                assert (Eo - E0).abs().max().item() > 1e-4
                _, _, E1, F1 = db.eval()
                assert torch.equal(E0, E1)
                assert torch.equal(F0, F1)
            db.set_buckets([], [])
            assert db.buckets() is None
            _, _, E2, F2 = db.eval()
            assert torch.equal(E0, E2) and torch.equal(F0, F2)
            with pytest.raises(Exception):
                db.set_buckets([32, 16, 512], [(1, 0, 0)] * 3)
            with pytest.raises(Exception):
                db.set_buckets([16, 512], [(3, 0, 0), (1, 0, 0)])
            chosen, _ = db.autotune(forces=True, repeats=1)
            plan = db.buckets()
            print("tuner", chosen, plan, flush=True)
            _, _, E3, F3 = db.eval()
            tile = chosen[2] == 1 or (plan is not None and any(s[2] for s in plan["shapes"]))
            if plan is not None:
                assert len(plan["caps"]) >= 2 and sum(plan["atoms"]) == db.natoms
                assert plan["ms_buckets"] < plan["ms_single"]
            if tile:
                assert (E3 - E0).abs().max().item() < 2 * E_TOL
                assert (F3 - F0).abs().max().item() < 2 * F_TOL
            else:
                assert torch.equal(E0, E3) and torch.equal(F0, F3)
            # a launch shape set by the caller drops the plan
            db.set_launch(1, 0, 0)
            assert db.buckets() is None
            del db
        finally:
            ann.release()


def test_forces_bitwise_reproducible(device, structures):
    """Forces are summed as 64-bit fixed-point integers, so the order in
    which the atomics of different atoms arrive changes nothing: repeated
    evaluations give bitwise the same forces (and energies), for the
    2048-structure Behler-Parrinello batch and for the 4000-atom Chebyshev
    alloy.  The fixed-point unit, 2^-32 eV/A, is far below the FP32
    arithmetic of the contributions: the forces stay within the per-atom
    bar of the oracle (checked on a subset in test_uneven_subset_matches_oracle
    and on the alloy in test_gpu_parity)."""
    import torch
    from aenet_amd import synth
    from aenet_amd.core import ANNPotentials
    from aenet_amd.device import DeviceBatch
    for pots, structs in ((pot_files("bp3", NAMES), structures), (pot_files("alloy", ["A", "B"]),
                                                                  [synth.c2_alloy4000()])):
        ann = ANNPotentials(pots)
        try:
            db = DeviceBatch(structs, device=device)
            _, _, E0, F0 = db.eval()
            E0, F0 = E0.clone(), F0.clone()
            assert torch.isfinite(F0).all()
            for _ in range(3):
                _, _, E1, F1 = db.eval()
                assert torch.equal(E0, E1)
                assert torch.equal(F0, F1)
            del db
        finally:
            ann.release()


def test_isolated_structure_leaving_its_box_is_replanned(device):
    """An isolated cluster in a device batch whose atom moves outward, past
    the bounding box the batch was planned with: the checked evaluation
    makes a new plan from the current coordinates and repeats, and the
    energies and forces equal the oracle's for the moved cluster (which
    takes the bounding box of the moved coordinates, geo_get_bounds)."""
    import torch
    from aenet_amd import synth
    from aenet_amd.core import ANNPotentials
    from aenet_amd.device import DeviceBatch
    from oracle_lib import Oracle
    files = pot_files("bp3", NAMES)
    ann = ANNPotentials(files)
    try:
        cl = [s for s in synth.batch_uneven(40) if not s.pbc][0]
        bulk = [s for s in synth.batch_uneven(40) if s.pbc][0]
        db = DeviceBatch([bulk, cl], device=device)
        db.eval()
        moved = cl.positions.copy()
        far = int(np.argmax(np.sum((moved - moved.mean(axis=0)) ** 2, axis=1)))
        d = moved[far] - moved.mean(axis=0)
        moved[far] += 1.5 * d / np.linalg.norm(d)      # 1.5 A outward: outside the old box
        x = db.coords.clone().view(-1, 3)
        n0 = len(bulk.types)
        x[n0:] = torch.as_tensor(moved, device=device)
        print("case cluster", len(cl.types), "atoms, atom", far, "moved out", flush=True)
        E_tot, _, E_at, F = db.eval(coords=x.view(-1).contiguous())
        o = Oracle(NAMES, files)
        ref = o.structure(cl.cell, False, cl.types, moved)
        dE = np.abs(E_at[n0:].cpu().numpy() - ref["energies"]).max()
        dF = np.abs(F[n0:].cpu().numpy() - ref["forces"]).max()
        print("moved cluster: max |dE| %.3e eV/atom  max |dF| %.3e eV/A" % (dE, dF))
        assert dE < 1e-5 and dF < 1e-4
        ref_b = o.structure(bulk.cell, True, bulk.types, bulk.positions)
        assert np.abs(E_at[:n0].cpu().numpy() - ref_b["energies"]).max() < 1e-5
    finally:
        ann.release()
