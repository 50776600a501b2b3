"""The arithmetic the cell-major neighbor search (k_search_cells,
aenet_amd/csrc/kernels.hip) rests on, checked on the CPU:

- the work list: item (c, j) -- atoms 8j..8j+7 of cell c in cell order --
  sits at slot start_c // 8 + c + j, and the slots from there to the next
  cell's first are -1; every slot of the natoms // 8 + ncells + 1 is
  written exactly once per step, whatever the cell counts;
- the FP32 prefilter: with candidate and central atom within M of the home
  cell centre, coordinates rounded to FP32 relative to the centre and the
  squared distance formed in FP32 (fused or not), the error stays below
  1e-5 M^2, the band inside which the kernel falls back to the reference's
  FP64 test.  A distance outside Rc^2 +- band therefore decides as the FP64
  test does (lcl_nbdist_cart, src/ext/lclist.f90:629-630, :654).
"""
import numpy as np

CS_WAVES = 8


def _work_list(counts):
    """The slots k_bin_small / k_cell_sort write, restated."""
    ncells = len(counts)
    starts = np.concatenate([[0], np.cumsum(counts)])
    natoms = int(starts[-1])
    cap = natoms // CS_WAVES + ncells + 1
    slots = np.full(cap, -2, np.int64)  # -2: never written
    writes = np.zeros(cap, np.int64)
    for c in range(ncells):
        base = starts[c] // CS_WAVES + c
        end = starts[c + 1] // CS_WAVES + c + 1 if c + 1 < ncells else cap
        items = (counts[c] + CS_WAVES - 1) // CS_WAVES
        assert items <= end - base
        for k in range(base, end):
            slots[k] = c if k - base < items else -1
            writes[k] += 1
    return slots, writes, starts


def test_work_list_covers_every_slot_once():
    rng = np.random.default_rng(5)
    cases = [np.zeros(7, int), np.array([1]), np.array([392, 0, 3, 4, 0, 9]), np.full(125, 32)]
    for _ in range(40):
        n = int(rng.integers(1, 300))
        cases.append(rng.integers(0, rng.choice([2, 9, 40, 400]), n))
    for counts in cases:
        slots, writes, starts = _work_list(np.asarray(counts))
        assert np.all(writes == 1)
        # the search decodes slot k of cell h as j = k - (start_h // 8 + h)
        seen = np.zeros(len(counts), np.int64)
        for k, h in enumerate(slots):
            if h < 0:
                continue
            j = k - (starts[h] // CS_WAVES + h)
            assert 0 <= j * CS_WAVES < counts[h]
            seen[h] += 1
        assert np.array_equal(seen, (np.asarray(counts) + CS_WAVES - 1) // CS_WAVES)


def _d2_f32(cand, atom, ctr, fused):
    c32 = (cand - ctr).astype(np.float32)
    a32 = (atom - ctr).astype(np.float32)
    d = (c32 - a32).astype(np.float32)
    if fused:  # FMA chain: each product-sum rounded once (emulated through FP64)
        acc = (d[:, 0].astype(np.float64) * d[:, 0]).astype(np.float32)
        for k in (1, 2):
            acc = (d[:, k].astype(np.float64) * d[:, k] + acc).astype(np.float32)
        return acc
    sq = (d * d).astype(np.float32)
    return ((sq[:, 0] + sq[:, 1]).astype(np.float32) + sq[:, 2]).astype(np.float32)


def test_fp32_prefilter_band():
    rng = np.random.default_rng(11)
    for M in (3.0, 26.2, 80.0, 400.0):
        n = 200000
        ctr = rng.uniform(-1e3, 1e3, 3)
        # candidates and the central atom anywhere within M of the centre
        cand = ctr + rng.uniform(-M, M, (n, 3)) / np.sqrt(3.0)
        atom = ctr + rng.uniform(-M, M, (n, 3)) / np.sqrt(3.0)
        exact = np.sum((cand - atom) ** 2, axis=1)
        band = 1.0e-5 * M * M
        for fused in (False, True):
            err = np.abs(_d2_f32(cand, atom, ctr, fused).astype(np.float64) - exact)
            print("M %.1f fused %d: max |d2_f32 - d2| %.3e, band %.3e" % (M, fused, err.max(), band))
            assert err.max() < band
            # the bound the kernel states, 84 M^2 2^-24, sits inside the band
            assert err.max() < 84.0 * M * M * 2.0 ** -24 < band


def test_fp32_prefilter_decisions_match_fp64():
    """Distances placed on both sides of the radius, from 1e-9 A to 1e-1 A
    away: wherever the FP32 value leaves the band, its decision is the FP64
    decision; the rest go to the FP64 test in the kernel."""
    rng = np.random.default_rng(12)
    rc2 = (6.5 + 1e-3) ** 2
    M = 6.501 + 3 * 6.58
    band = 1.0e-5 * M * M + 1.0e-6 * rc2
    n = 100000
    ctr = np.array([19.0, 19.0, 19.0])
    atom = ctr + rng.uniform(-6.58, 6.58, (n, 3)) / np.sqrt(3.0)
    u = rng.normal(size=(n, 3))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    r = np.sqrt(rc2) + rng.choice([-1, 1], n) * 10.0 ** rng.uniform(-9, -1, n)
    cand = atom + u * r[:, None]
    exact_pass = np.sum((cand - atom) ** 2, axis=1) <= rc2
    d2 = _d2_f32(cand, atom, ctr, True).astype(np.float64)
    decided = (d2 < rc2 - band) | (d2 > rc2 + band)
    assert np.all((d2[decided] < rc2 - band) == exact_pass[decided])
    # offsets are log-uniform over 1e-9..1e-1 A and the band is about 5e-4 A
    # wide in r, so about 28 % of these crowded distances decide in FP32
    print("decided in FP32: %d of %d" % (decided.sum(), n))
    assert decided.sum() > n // 5
