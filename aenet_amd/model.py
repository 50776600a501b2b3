"""Independent potential sets on the GPU (aenet_amd_model_* of
include/aenet_amd.h) and the network-training steps that run on them.

A ``Model`` holds one .nn potential per species, like the set that
``ANNPotentials`` loads through the reference's interface, but any number
of them can live side by side: a cheap and a full potential in
delayed-acceptance Monte Carlo, or a network under training beside its
teacher.  Weights are one float32 device array in the .nn layout (per
layer a column-major (n_out, n_in + 1) block, bias last; species after
species).
"""
import ctypes as C

import numpy as np
import torch

from . import _native
from .core import check_status


def _ptr(t):
    return C.c_void_p(t.data_ptr()) if t is not None else None


class Model:
    def __init__(self, potential_files, device="cuda"):
        """potential_files: dict species -> .nn file, in type-id order."""
        self.device = torch.device(device)
        self.atom_types = list(potential_files.keys())
        names = (C.c_char_p * len(self.atom_types))(*[t.encode() for t in self.atom_types])
        files = (C.c_char_p * len(self.atom_types))(*[str(potential_files[t]).encode() for t in self.atom_types])
        h = C.c_void_p()
        stat = C.c_int(0)
        _native.lib().aenet_amd_model_load(len(self.atom_types), names, files, C.byref(h), C.byref(stat))
        check_status(stat.value, "aenet_amd_model_load")
        self.handle = h
        rc = C.c_double()
        nsf = C.c_int()
        ws = C.c_int()
        _native.lib().aenet_amd_model_info(h, C.byref(rc), C.byref(nsf), C.byref(ws), C.byref(stat))
        check_status(stat.value)
        self.Rc_max, self.nsf_max, self.wsize = rc.value, nsf.value, ws.value

    def __del__(self):
        h = getattr(self, "handle", None)
        if h:
            _native.lib().aenet_amd_model_destroy(h)
            self.handle = None

    def _stream(self):
        return C.c_void_p(torch.cuda.current_stream(self.device).cuda_stream)

    def type_id(self, species):
        return self.atom_types.index(species) + 1

    def species_weights(self, type_id):
        """(offset, size) of species type_id's network in the weight array."""
        off, n, stat = C.c_int(), C.c_int(), C.c_int()
        _native.lib().aenet_amd_model_species_weights(self.handle, type_id, C.byref(off), C.byref(n),
                                                      C.byref(stat))
        check_status(stat.value)
        return off.value, n.value

    def weights(self):
        W = torch.empty(self.wsize, dtype=torch.float32, device=self.device)
        stat = C.c_int()
        _native.lib().aenet_amd_model_weights(self.handle, _ptr(W), 0, self._stream(), C.byref(stat))
        check_status(stat.value)
        return W

    def set_weights(self, W):
        W = W.detach().to(device=self.device, dtype=torch.float32).contiguous()
        assert W.numel() == self.wsize
        stat = C.c_int()
        _native.lib().aenet_amd_model_weights(self.handle, _ptr(W), 1, self._stream(), C.byref(stat))
        check_status(stat.value)

    def descriptor_derivatives(self, envs, stride=None):
        """Descriptor values and derivatives of explicit environments
        (aenet_amd_model_descriptor_derivatives).  envs: list of (type_i,
        coo_i (3,), coo_j (n, 3), type_j (n,)).  Returns (values [nenv,
        stride], deriv_i [nenv, stride, 3], list of deriv_j [n, stride, 3])."""
        stride = int(stride or self.nsf_max)
        nenv = len(envs)
        start = np.zeros(nenv + 1, np.int32)
        for e, env in enumerate(envs):
            start[e + 1] = start[e] + len(env[3])
        ti = np.array([int(env[0]) for env in envs], np.int32)
        ci = np.ascontiguousarray(np.concatenate([np.reshape(env[1], 3) for env in envs]), dtype=np.float64)
        nb = int(start[-1])
        cj = np.zeros(3 * max(nb, 1))
        tj = np.zeros(max(nb, 1), np.int32)
        for e, env in enumerate(envs):
            if len(env[3]):
                cj[3 * start[e]:3 * start[e + 1]] = np.ascontiguousarray(env[2], dtype=np.float64).reshape(-1)
                tj[start[e]:start[e + 1]] = env[3]
        vals = np.zeros((nenv, stride))
        di = np.zeros((nenv, stride, 3))
        dj = np.zeros((max(nb, 1), stride, 3))
        stat = C.c_int()
        _native.lib().aenet_amd_model_descriptor_derivatives(self.handle, nenv, start, ti, ci, cj, tj, stride,
                                                            vals.ctypes.data, di.ctypes.data, dj.ctypes.data,
                                                            C.byref(stat))
        check_status(stat.value, "aenet_amd_model_descriptor_derivatives")
        return vals, di, [dj[start[e]:start[e + 1]] for e in range(nenv)]

    def net_grad(self, G, types, E_ref=None, weight=None, upstream=None, coef=1.0, E_out=None, grad=None,
                 loss=None):
        """Atomic energies of normalized descriptors G (natoms, stride) and,
        with grad, grad += d(sum_i g_i E_i)/dW where g_i = coef*weight_i*
        (E_i - E_ref_i) or upstream_i (include/aenet_amd.h)."""
        stat = C.c_int()
        _native.lib().aenet_amd_model_net_grad(self.handle, G.shape[0], _ptr(G), G.shape[1], _ptr(types),
                                               _ptr(E_ref), _ptr(weight), _ptr(upstream), float(coef), _ptr(E_out),
                                               _ptr(grad), _ptr(loss), self._stream(), C.byref(stat))
        check_status(stat.value, "aenet_amd_model_net_grad")

    def fit_cells(self, latt, natoms, coords, types, sel, E_ref, F_ref, coefE, coefF, grad, loss):
        """grad += gradient of coefE/2 sum (E - E_ref)^2 + coefF/2 sum |F -
        F_ref|^2 over the cells sel (device int32) of a set of cells of
        natoms atoms with one lattice (rows = vectors); loss[0:2] += the two
        sums of squares (aenet_amd_model_fit_cells)."""
        stat = C.c_int()
        L = np.ascontiguousarray(latt, dtype=np.float64).reshape(9)
        _native.lib().aenet_amd_model_fit_cells(self.handle, int(types.numel() // natoms), int(natoms), L,
                                                _ptr(coords), _ptr(types), _ptr(sel), int(sel.numel()), _ptr(E_ref),
                                                _ptr(F_ref), float(coefE), float(coefF), _ptr(grad), _ptr(loss),
                                                self._stream(), C.byref(stat))
        check_status(stat.value, "aenet_amd_model_fit_cells")

    def adam(self, grad, m, v, t, lr, b1=0.9, b2=0.999, eps=1e-8):
        stat = C.c_int()
        _native.lib().aenet_amd_model_adam(self.handle, _ptr(grad), _ptr(m), _ptr(v), int(t), float(lr), float(b1),
                                           float(b2), float(eps), self._stream(), C.byref(stat))
        check_status(stat.value, "aenet_amd_model_adam")

    def check(self):
        stat = C.c_int()
        _native.lib().aenet_amd_model_check(self.handle, self._stream(), C.byref(stat))
        check_status(stat.value, "aenet_amd_model_check")
