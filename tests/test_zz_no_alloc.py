"""aenet_structures_energy_and_forces_noalloc (device buffers from the
library's fixed scratch array, no device allocation) against
aenet_structures_energy_and_forces on the same inputs, through the C ABI:
two calls in a row, the second with a smaller input, then an input larger
than the scratch array, which must be refused with a message.  Per-atom
energies are computed by the same kernels in the same order (identical);
forces are FP64 sums whose order may differ (within 1e-10 eV/A)."""
import ctypes as C

import numpy as np
import pytest

from conftest import pot_files

pytestmark = pytest.mark.gpu


def _call(fn, structs):
    from aenet_amd import _native
    L = _native.lib()
    nat = np.array([len(s.types) for s in structs], np.int32)
    latt = np.ascontiguousarray(np.concatenate([s.cell.reshape(9) for s in structs]))
    pbc = np.array([1 if s.pbc else 0 for s in structs], np.int32)
    types = np.ascontiguousarray(np.concatenate([s.types for s in structs]).astype(np.int32))
    coords = np.ascontiguousarray(np.concatenate([s.positions for s in structs]).reshape(-1))
    n = int(nat.sum())
    E_tot = np.zeros(len(structs))
    E_coh = np.zeros(len(structs))
    E_at = np.zeros(n)
    F = np.zeros(3 * n)
    st = C.c_int(-1)
    getattr(L, fn)(len(structs), nat, latt, pbc, types, coords, E_tot.ctypes.data, E_coh.ctypes.data,
                   E_at.ctypes.data, F.ctypes.data, C.byref(st))
    return st.value, E_tot, E_coh, E_at, F.reshape(-1, 3)


def _same(structs, label):
    print("case", label, sum(len(s.types) for s in structs), "atoms", flush=True)
    a = _call("aenet_structures_energy_and_forces", structs)
    b = _call("aenet_structures_energy_and_forces_noalloc", structs)
    assert a[0] == 0 and b[0] == 0
    assert np.array_equal(a[3], b[3])
    assert np.allclose(a[1], b[1], rtol=0, atol=1e-9) and np.allclose(a[2], b[2], rtol=0, atol=1e-9)
    dF = np.abs(a[4] - b[4]).max()
    print("  max |dF| %.3e eV/A" % dF)
    assert dF < 1e-10


def test_noalloc_matches_allocating_entry_point(device):
    from aenet_amd import _native, synth
    from aenet_amd.core import ANNPotentials
    ann = ANNPotentials(pot_files("alloy", ["A", "B"]))
    try:
        _same([synth.c2_alloy4000()], "alloy4000")
        # second call, smaller: a 256-atom cell and an isolated cluster
        small = synth.binary_fcc(3.8, (4, 4, 4), 0.08, 1, 2)
        c = synth.binary_fcc(3.8, (2, 2, 2), 0.08, 7, 8)
        cl = synth.Structure(np.eye(3), c.positions, c.types, False, "cluster32")
        _same([small, cl], "alloy256+cluster32")
        # larger than the scratch array: refused with a message, nothing computed
        big = synth.alloy_replicas(10)
        st = _call("aenet_structures_energy_and_forces_noalloc", big)[0]
        msg = _native.last_error()
        print("  40000 atoms:", st, msg)
        assert st == _native.AENET_ERR_CAPACITY and "scratch" in msg
        # and the array still serves a later call
        _same([small], "alloy256 again")
    finally:
        ann.release()
