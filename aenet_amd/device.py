"""Device-resident structure batches (torch tensors in HBM).

Wraps the aenet_amd_batch_* entry points: the structures' shapes are fixed
at creation; coordinates and types are device tensors handed to each
evaluation.  torch provides the device memory and the stream; all compute
is in libaenet_amd.so.
"""
import ctypes as C

import numpy as np
import torch

from . import _native
from .core import check_status


def _ptr(t):
    return C.c_void_p(t.data_ptr()) if t is not None else None


def _handle(model):
    return model.handle if model is not None else None


class DeviceBatch:
    """A fixed set of structures evaluated on the GPU.

    structures: list of objects with .cell (3x3, rows = vectors), .pbc,
    .positions (N,3) and .types (1-based ids of the loaded potentials)."""

    def __init__(self, structures, device="cuda", model=None):
        self.device = torch.device(device)
        self.nstruct = len(structures)
        self.natoms_per = np.array([len(s.types) for s in structures], dtype=np.int32)
        latt = np.ascontiguousarray(np.concatenate([np.asarray(s.cell, np.float64).reshape(9) for s in structures]))
        pbc = np.array([1 if s.pbc else 0 for s in structures], dtype=np.int32)
        coords = np.ascontiguousarray(np.concatenate([np.asarray(s.positions, np.float64) for s in structures]))
        types = np.ascontiguousarray(np.concatenate([np.asarray(s.types, np.int32) for s in structures]))
        h = C.c_void_p()
        stat = C.c_int(0)
        self.model = model
        _native.lib().aenet_amd_batch_create_model(_handle(model), self.nstruct, self.natoms_per, latt, pbc,
                                                   coords.reshape(-1), C.byref(h), C.byref(stat))
        check_status(stat.value, "aenet_amd_batch_create")
        self._h = h
        self.natoms = int(self.natoms_per.sum())
        self.offsets = np.concatenate([[0], np.cumsum(self.natoms_per)])
        self.coords = torch.tensor(coords, dtype=torch.float64, device=self.device)
        self.types = torch.tensor(types, dtype=torch.int32, device=self.device)
        self.E_atoms = torch.zeros(self.natoms, dtype=torch.float64, device=self.device)
        self.E_struct = torch.zeros(2 * self.nstruct, dtype=torch.float64, device=self.device)
        self.forces = torch.zeros((self.natoms, 3), dtype=torch.float64, device=self.device)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            _native.lib().aenet_amd_batch_destroy(h)
            self._h = None

    def _stream(self):
        return C.c_void_p(torch.cuda.current_stream(self.device).cuda_stream)

    def _model(self, model):
        return _handle(model if model is not None else self.model)

    def _check_inputs(self, coords, types):
        assert coords.is_contiguous() and coords.dtype == torch.float64 and coords.numel() == 3 * self.natoms
        assert types.is_contiguous() and types.dtype == torch.int32 and types.numel() == self.natoms

    def eval(self, coords=None, types=None, forces=True, check=True, model=None):
        """Launch one evaluation on the current torch stream; returns
        (E_tot, E_coh, E_atoms, forces) device tensors.  With check, an
        isolated structure whose atoms left the bounding box of its plan
        gets a new plan from the current coordinates and the evaluation is
        repeated (without check, keep isolated structures inside their
        boxes, e.g. under graph replay)."""
        coords = self.coords if coords is None else coords
        types = self.types if types is None else types
        self._check_inputs(coords, types)
        for attempt in range(2):
            stat = C.c_int(0)
            _native.lib().aenet_amd_batch_eval_model(self._h, self._model(model), _ptr(coords), _ptr(types),
                                                     _ptr(self.E_atoms), _ptr(self.E_struct),
                                                     _ptr(self.forces) if forces else None, self._stream(),
                                                     C.byref(stat))
            check_status(stat.value, "aenet_amd_batch_eval")
            if not check:
                break
            stat = C.c_int(0)
            _native.lib().aenet_amd_batch_check(self._h, self._stream(), C.byref(stat))
            if stat.value == _native.AENET_ERR_COORD and attempt == 0 and self.has_isolated():
                self.replan(coords)
                continue
            check_status(stat.value, "aenet_amd_batch_check")
            break
        return (self.E_struct[:self.nstruct], self.E_struct[self.nstruct:], self.E_atoms,
                self.forces if forces else None)

    def eval_active(self, active, nactive, max_active, E_out, coords=None, types=None, reuse_lists=False,
                    model=None):
        """Atomic energies of the atoms active[:nactive[0]] into E_out
        (device tensors: int32 active list, int32 count, float64 energies)."""
        coords = self.coords if coords is None else coords
        types = self.types if types is None else types
        self._check_inputs(coords, types)
        stat = C.c_int(0)
        _native.lib().aenet_amd_batch_eval_active(self._h, self._model(model), _ptr(coords), _ptr(types),
                                                  _ptr(active), _ptr(nactive), int(max_active), int(reuse_lists),
                                                  _ptr(E_out), self._stream(), C.byref(stat))
        check_status(stat.value, "aenet_amd_batch_eval_active")

    def vjp(self, g, grad_W=None, coords=None, types=None, model=None):
        """sum_i g_i dE_i/dR (returned, (natoms, 3) float64) and, when grad_W
        is given, grad_W += sum_i g_i dE_i/dW.  The coordinates must lie in
        the isolated structures' boxes (as after a checked eval of them)."""
        coords = self.coords if coords is None else coords
        types = self.types if types is None else types
        self._check_inputs(coords, types)
        g = g.to(torch.float64).contiguous()
        gR = torch.empty((self.natoms, 3), dtype=torch.float64, device=self.device)
        stat = C.c_int(0)
        _native.lib().aenet_amd_batch_vjp(self._h, self._model(model), _ptr(coords), _ptr(types), _ptr(g), _ptr(gR),
                                          _ptr(grad_W), self._stream(), C.byref(stat))
        check_status(stat.value, "aenet_amd_batch_vjp")
        return gR

    def has_isolated(self):
        return bool(_native.lib().aenet_amd_batch_has_isolated(self._h))

    def replan(self, coords=None):
        """New bounding boxes of the isolated structures from coordinates
        (default: the batch's own tensor); periodic structures keep their
        plans (aenet_amd_batch_replan)."""
        coords = self.coords if coords is None else coords
        x = np.ascontiguousarray(coords.detach().to("cpu", torch.float64).numpy().reshape(-1))
        stat = C.c_int(0)
        _native.lib().aenet_amd_batch_replan(self._h, x, C.byref(stat))
        check_status(stat.value, "aenet_amd_batch_replan")

    def check(self):
        stat = C.c_int(0)
        _native.lib().aenet_amd_batch_check(self._h, self._stream(), C.byref(stat))
        check_status(stat.value, "aenet_amd_batch_check")

    def set_timing(self, on):
        _native.lib().aenet_amd_batch_set_timing(self._h, 1 if on else 0)

    def kernel_time(self):
        """(summed ms, launches) of the atom kernel since the last call."""
        ms = C.c_double(0.0)
        n = C.c_int(0)
        stat = C.c_int(0)
        _native.lib().aenet_amd_batch_kernel_time(self._h, C.byref(ms), C.byref(n), C.byref(stat))
        check_status(stat.value, "aenet_amd_batch_kernel_time")
        return ms.value, n.value

    def search_time(self):
        """(summed ms, launches) of the neighbor search since the last call."""
        ms = C.c_double(0.0)
        n = C.c_int(0)
        stat = C.c_int(0)
        _native.lib().aenet_amd_batch_search_time(self._h, C.byref(ms), C.byref(n), C.byref(stat))
        check_status(stat.value, "aenet_amd_batch_search_time")
        return ms.value, n.value

    def neighbors(self, cap, coords=None, types=None):
        coords = self.coords if coords is None else coords
        types = self.types if types is None else types
        d = self.device
        n = self.natoms
        count = torch.zeros(n, dtype=torch.int32, device=d)
        coo = torch.zeros((n, cap, 3), dtype=torch.float64, device=d)
        dist = torch.zeros((n, cap), dtype=torch.float64, device=d)
        lst = torch.zeros((n, cap), dtype=torch.int32, device=d)
        typ = torch.zeros((n, cap), dtype=torch.int32, device=d)
        stat = C.c_int(0)
        _native.lib().aenet_amd_batch_neighbors(self._h, _ptr(coords), _ptr(types), cap, _ptr(count), _ptr(coo),
                                                _ptr(dist), _ptr(lst), _ptr(typ), self._stream(), C.byref(stat))
        check_status(stat.value, "aenet_amd_batch_neighbors")
        self.check()
        return {"nnb": count, "coo": coo, "dist": dist, "list": lst, "type": typ}

    def autotune(self, forces=True, repeats=5, model=None, coords=None, types=None):
        """Time every feasible launch shape of the atom kernel on this batch
        (aenet_amd_batch_autotune) and keep the fastest for the later
        evaluations with the model.  Returns ((waves, staged, tile), rows)
        with rows = [(waves, staged, tile, average ms), ...]; tile = 1 for
        the networks of a workgroup's atoms as one matrix-core tile."""
        coords = self.coords if coords is None else coords
        types = self.types if types is None else types
        self._check_inputs(coords, types)
        res = np.zeros(4 * 32)
        n = C.c_int(0)
        chosen = (C.c_int * 3)()
        stat = C.c_int(0)
        _native.lib().aenet_amd_batch_autotune(self._h, self._model(model), _ptr(coords), _ptr(types), int(forces),
                                               int(repeats), res.ctypes.data, C.byref(n), chosen, self._stream(),
                                               C.byref(stat))
        check_status(stat.value, "aenet_amd_batch_autotune")
        rows = [(int(res[4 * k]), int(res[4 * k + 1]), int(res[4 * k + 2]), float(res[4 * k + 3]))
                for k in range(min(n.value, 32))]
        return (chosen[0], chosen[1], chosen[2]), rows

    def set_launch(self, waves, staged=1, tile=0, model=None):
        """Use this launch shape of the atom kernel (aenet_amd_batch_set_launch);
        waves = 0 returns to the occupancy model."""
        stat = C.c_int(0)
        _native.lib().aenet_amd_batch_set_launch(self._h, self._model(model), int(waves), int(staged), int(tile),
                                                 C.byref(stat))
        check_status(stat.value, "aenet_amd_batch_set_launch")

    def set_buckets(self, caps, shapes, model=None):
        """Neighbor-count buckets of the atom kernel (aenet_amd_batch_set_buckets):
        caps ascending (the last becomes the batch's neighbor capacity), one
        (waves, staged, tile) shape per bucket; caps = [] drops the plan."""
        caps = np.ascontiguousarray(caps, dtype=np.int32)
        sh = np.ascontiguousarray(np.asarray(shapes, dtype=np.int32).reshape(-1))
        if len(caps) == 0:
            caps = np.zeros(1, dtype=np.int32)
            sh = np.zeros(3, dtype=np.int32)
            n = 0
        else:
            n = len(caps)
            if sh.size != 3 * n:
                raise ValueError("one (waves, staged, tile) shape per bucket")
        stat = C.c_int(0)
        _native.lib().aenet_amd_batch_set_buckets(self._h, self._model(model), n, caps, sh, C.byref(stat))
        check_status(stat.value, "aenet_amd_batch_set_buckets")

    def buckets(self, model=None):
        """The bucket plan in use (aenet_amd_batch_buckets): a dict with caps,
        shapes, atoms and ms per bucket and the tuner's totals (ms) for one
        launch and for the buckets; None without a plan."""
        n = C.c_int(0)
        caps = np.zeros(6, dtype=np.int32)
        shapes = np.zeros(18, dtype=np.int32)
        atoms = np.zeros(6, dtype=np.int32)
        ms = np.zeros(6)
        tot = np.zeros(2)
        stat = C.c_int(0)
        _native.lib().aenet_amd_batch_buckets(self._h, self._model(model), C.byref(n), caps.ctypes.data,
                                              shapes.ctypes.data, atoms.ctypes.data, ms.ctypes.data,
                                              tot.ctypes.data, C.byref(stat))
        check_status(stat.value, "aenet_amd_batch_buckets")
        k = n.value
        if k == 0:
            return None
        return {"caps": caps[:k].tolist(), "shapes": shapes[:3 * k].reshape(k, 3).tolist(),
                "atoms": atoms[:k].tolist(), "ms": ms[:k].tolist(), "ms_single": float(tot[0]),
                "ms_buckets": float(tot[1])}

    def descriptors(self, stride, scaled=False, coords=None, types=None, model=None):
        coords = self.coords if coords is None else coords
        types = self.types if types is None else types
        G = torch.zeros((self.natoms, stride), dtype=torch.float64, device=self.device)
        stat = C.c_int(0)
        _native.lib().aenet_amd_batch_descriptors_model(self._h, self._model(model), _ptr(coords), _ptr(types),
                                                        int(scaled), _ptr(G), stride, self._stream(), C.byref(stat))
        check_status(stat.value, "aenet_amd_batch_descriptors")
        self.check()
        return G
