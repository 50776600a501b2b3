"""Monte Carlo of many periodic cells on the GPU (aenet_amd_mc_* of
include/aenet_amd.h): plain Metropolis with a full potential, or delayed
acceptance with a radial-only first-stage potential that screens proposals
before the full potential decides.
"""
import ctypes as C
import time

import numpy as np

from . import _native
from .core import check_status

K_B = 8.617333262e-5  # eV/K


class MonteCarlo:
    def __init__(self, full, structures, first=None, seed=1):
        """full, first: aenet_amd.model.Model (first may be None); structures:
        cells of equal size and the same lattice (rows = vectors), .types
        1-based ids of the models' species."""
        self.full, self.first = full, first
        self.nchains = len(structures)
        self.natoms = len(structures[0].types)
        cell = np.ascontiguousarray(structures[0].cell, dtype=np.float64).reshape(9)
        for s in structures:
            assert len(s.types) == self.natoms and np.allclose(np.asarray(s.cell).reshape(9), cell)
        types = np.ascontiguousarray(np.concatenate([s.types for s in structures]), dtype=np.int32)
        coords = np.ascontiguousarray(np.concatenate([s.positions for s in structures]), dtype=np.float64)
        h = C.c_void_p()
        stat = C.c_int()
        _native.lib().aenet_amd_mc_create(full.handle, first.handle if first is not None else None, self.nchains,
                                          self.natoms, cell, types, coords.reshape(-1), C.c_ulonglong(seed),
                                          C.byref(h), C.byref(stat))
        check_status(stat.value, "aenet_amd_mc_create")
        self.handle = h
        self.cell = cell.reshape(3, 3)

    def __del__(self):
        h = getattr(self, "handle", None)
        if h:
            _native.lib().aenet_amd_mc_destroy(h)
            self.handle = None

    def set(self, T=600.0, p_swap=0.5, max_disp=0.075, beta=None, max_screen=4):
        """max_screen: first-stage tries per chain in a delayed-acceptance
        step (the first that passes goes to the full potential)."""
        stat = C.c_int()
        b = beta if beta is not None else 1.0 / (K_B * T)
        _native.lib().aenet_amd_mc_set(self.handle, float(b), float(p_swap), float(max_disp), int(max_screen),
                                       C.byref(stat))
        check_status(stat.value)

    def refresh(self):
        stat = C.c_int()
        _native.lib().aenet_amd_mc_refresh(self.handle, C.byref(stat))
        check_status(stat.value)

    def run(self, nsteps, delayed=False, sample_every=0, graph=True):
        """nsteps steps (plain: one proposal per chain; delayed: up to
        max_screen first-stage tries per chain); energies sampled every
        sample_every proposals of each chain; returns the wall time in s."""
        stat = C.c_int()
        t0 = time.perf_counter()
        _native.lib().aenet_amd_mc_run(self.handle, int(nsteps), int(delayed), int(sample_every), int(graph),
                                       C.byref(stat))
        t1 = time.perf_counter()
        check_status(stat.value, "aenet_amd_mc_run")
        return t1 - t0

    def stats(self, reset=False):
        out = (C.c_longlong * 4)()
        stat = C.c_int()
        _native.lib().aenet_amd_mc_stats(self.handle, out, int(reset), C.byref(stat))
        check_status(stat.value)
        return {"proposed": out[0], "accepted": out[1], "full_evaluations": out[2], "swaps": out[3]}

    def energies(self):
        E = np.zeros(self.nchains)
        stat = C.c_int()
        _native.lib().aenet_amd_mc_energies(self.handle, E, C.byref(stat))
        check_status(stat.value)
        return E

    def samples(self):
        n = C.c_longlong()
        stat = C.c_int()
        _native.lib().aenet_amd_mc_samples(self.handle, None, C.byref(n), C.byref(stat))
        check_status(stat.value)
        out = np.zeros((n.value, self.nchains))
        if n.value:
            _native.lib().aenet_amd_mc_samples(self.handle, out.ctypes.data, C.byref(n), C.byref(stat))
            check_status(stat.value)
        return out

    def arm_trace(self, max_props):
        """Record the first max_props decisions of every chain in the next runs."""
        n = C.c_longlong()
        stat = C.c_int()
        _native.lib().aenet_amd_mc_trace(self.handle, int(max_props), None, C.byref(n), C.byref(stat))
        check_status(stat.value)

    def trace(self):
        """(proposal index, chains, 4): dE_first, dE_full, stage-1 outcome
        (1 passed, 0 rejected, -1 no record), accepted."""
        n = C.c_longlong()
        stat = C.c_int()
        _native.lib().aenet_amd_mc_trace(self.handle, -1, None, C.byref(n), C.byref(stat))
        out = np.zeros((n.value, self.nchains, 4))
        if n.value:
            _native.lib().aenet_amd_mc_trace(self.handle, -1, out.ctypes.data, C.byref(n), C.byref(stat))
            check_status(stat.value)
        return out

    def state(self):
        x = np.zeros((self.nchains, self.natoms, 3))
        t = np.zeros((self.nchains, self.natoms), np.int32)
        stat = C.c_int()
        _native.lib().aenet_amd_mc_state(self.handle, x.ctypes.data, t.ctypes.data, C.byref(stat))
        check_status(stat.value)
        return x, t

    def first_stage_state(self):
        n = self.nchains * self.natoms
        nsf = C.c_int()
        stat = C.c_int()
        _native.lib().aenet_amd_mc_first_stage_state(self.handle, None, None, C.byref(nsf), C.byref(stat))
        E = np.zeros(n)
        G = np.zeros((n, nsf.value))
        _native.lib().aenet_amd_mc_first_stage_state(self.handle, E.ctypes.data, G.ctypes.data, C.byref(nsf),
                                                     C.byref(stat))
        check_status(stat.value)
        return E, G
