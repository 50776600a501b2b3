"""Python interface to the MI355X implementation of libaenet.

``ANNPotentials`` mirrors the reference's Cython class
(python3/aenet/core.pyx:38-202): same constructor argument (a dict
species -> potential file), attributes (atom_types, Rc_min, Rc_max,
nnb_max), methods (atomic_energy, atomic_energy_and_forces, get_type_id,
print_info_to_stdout) and exceptions, over the C ABI of libaenet_amd.so.
``structure_energy_and_forces`` is the batched path (whole structures on
the GPU, neighbor lists included).

As in the reference library, the loaded potential set is process-global:
one ANNPotentials instance at a time.
"""
import ctypes as C

import numpy as np

from . import _native


class AenetError(Exception):
    pass


class InitError(Exception):
    pass


class SpeciesError(Exception):
    pass


def check_status(stat, what=""):
    """Raise the exception the reference's _check_aenet_return_status
    (python3/aenet/core.pyx:189-202) raises for a status code."""
    if stat == _native.AENET_OK:
        return
    msg = _native.last_error() or what
    if stat == _native.AENET_ERR_MALLOC:
        raise MemoryError("aenetLib: Memory allocation failed. " + msg)
    if stat == _native.AENET_ERR_INIT:
        raise InitError("aenetLib: Module not initialized (or attempt to re-initialize module). " + msg)
    if stat == _native.AENET_ERR_IO:
        raise IOError("aenetLib: I/O error encountered. " + msg)
    if stat == _native.AENET_ERR_TYPE:
        raise SpeciesError("aenetLib: Invalid atomic species. " + msg)
    raise AenetError("aenetLib: error %d: %s" % (stat, msg))


def _names(types):
    return (C.c_char_p * len(types))(*[t.encode("utf-8") for t in types])


class ANNPotentials:
    """Artificial Neural Network (ANN) potentials on the GPU."""

    def __init__(self, potential_files):
        L = _native.lib()
        self.atom_types = list(potential_files.keys())
        self.potential_files = dict(potential_files)
        self.ntypes = len(self.atom_types)
        self.Rc_min = 0.0
        self.Rc_max = 0.0
        self._alive = False
        stat = C.c_int(0)
        L.aenet_init(self.ntypes, _names(self.atom_types), C.byref(stat))
        check_status(stat.value, "aenet_init")
        self._alive = True
        try:
            for t in self.atom_types:
                self._load_potential(t, self.potential_files[t])
            assert L.aenet_all_loaded()
        except Exception:
            self.release()
            raise
        self.Rc_min = _native.global_double("aenet_Rc_min")
        self.Rc_max = _native.global_double("aenet_Rc_max")
        self.nnb_max = _native.global_int("aenet_nnb_max")
        self.nsf_max = _native.global_int("aenet_nsf_max")

    def release(self):
        if self._alive:
            stat = C.c_int(0)
            _native.lib().aenet_final(C.byref(stat))
            self._alive = False

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass

    def __str__(self):
        return " ANN Potentials\n   Atomic species: " + " ".join(self.atom_types) + "\n"

    def _load_potential(self, atom_type, filename):
        stat = C.c_int(0)
        type_id = self.get_type_id(atom_type)
        _native.lib().aenet_load_potential(type_id, filename.encode("utf-8"), C.byref(stat))
        if stat.value == _native.AENET_ERR_TYPE:
            raise SpeciesError("Unknown type ID: %d" % type_id)
        if stat.value == _native.AENET_ERR_IO:
            raise IOError("File not found: %s (%s)" % (filename, _native.last_error()))
        check_status(stat.value, "aenet_load_potential")

    def print_info_to_stdout(self):
        _native.lib().aenet_print_info()

    def get_type_id(self, atom_type):
        for i, t in enumerate(self.atom_types):
            if t.strip() == str(atom_type).strip():
                return i + 1
        raise SpeciesError("Unknown atomic species: {}".format(atom_type))

    def free_atom_energy(self, atom_type):
        return _native.lib().aenet_free_atom_energy(self.get_type_id(atom_type))

    def atomic_energy(self, coo_i, type_i, coo_j, type_j):
        coo_i = np.ascontiguousarray(coo_i, dtype=np.float64).reshape(3)
        coo_j = np.ascontiguousarray(coo_j, dtype=np.float64).reshape(-1, 3)
        itype_j = np.array([self.get_type_id(str(t)) for t in type_j], dtype=np.int32)
        if len(itype_j) == 0:
            coo_j = np.zeros((1, 3))
            itype_j = np.zeros(1, np.int32)
        E = np.zeros(1)
        stat = C.c_int(0)
        _native.lib().aenet_atomic_energy(coo_i, self.get_type_id(type_i), len(type_j), coo_j, itype_j, E,
                                          C.byref(stat))
        check_status(stat.value)
        return float(E[0])

    def atomic_energy_and_forces(self, coo_i, type_i, index_i, coo_j, type_j, index_j, F):
        """F (natoms, 3) float64 C-contiguous is accumulated in place."""
        if not (isinstance(F, np.ndarray) and F.dtype == np.float64 and F.flags.c_contiguous):
            raise TypeError("F must be a C-contiguous float64 array of shape (natoms, 3)")
        coo_i = np.ascontiguousarray(coo_i, dtype=np.float64).reshape(3)
        coo_j = np.ascontiguousarray(coo_j, dtype=np.float64).reshape(-1, 3)
        itype_j = np.array([self.get_type_id(str(t)) for t in type_j], dtype=np.int32)
        index_j = np.ascontiguousarray(index_j, dtype=np.int32)
        if len(itype_j) == 0:
            coo_j = np.zeros((1, 3))
            itype_j = np.zeros(1, np.int32)
            index_j = np.ones(1, np.int32)
        E = np.zeros(1)
        stat = C.c_int(0)
        _native.lib().aenet_atomic_energy_and_forces(coo_i, self.get_type_id(type_i), int(index_i), len(type_j),
                                                     coo_j, itype_j, index_j, F.shape[0], E, F, C.byref(stat))
        check_status(stat.value)
        return float(E[0])

    def structure_energy_and_forces(self, cells, pbcs, positions, types, forces=True):
        """Energies (and forces) of one or more whole structures.

        cells: (3,3) or (nstruct,3,3), rows = lattice vectors (ignored for
        isolated structures); pbcs: bool or list; positions: (N,3) or list
        of arrays; types: species names or 1-based ids, per atom.
        Returns dict with E_tot, E_coh (per structure), energies and forces
        (per atom, concatenated)."""
        if isinstance(positions, np.ndarray) and positions.ndim == 2:
            positions = [positions]
            types = [types]
            cells = [cells]
            pbcs = [pbcs]
        ns = len(positions)
        natoms = np.array([len(p) for p in positions], dtype=np.int32)
        latt = np.ascontiguousarray(np.concatenate([np.asarray(c, np.float64).reshape(9) for c in cells]))
        pbc = np.array([1 if (np.all(p) if np.ndim(p) else p) else 0 for p in pbcs], dtype=np.int32)
        coords = np.ascontiguousarray(np.concatenate([np.asarray(p, np.float64).reshape(-1, 3) for p in positions]))
        ids = []
        for tl in types:
            for t in tl:
                ids.append(int(t) if isinstance(t, (int, np.integer)) else self.get_type_id(t))
        ids = np.ascontiguousarray(ids, dtype=np.int32)
        N = int(natoms.sum())
        E_tot = np.zeros(ns)
        E_coh = np.zeros(ns)
        E_at = np.zeros(N)
        F = np.zeros((N, 3)) if forces else None
        stat = C.c_int(0)
        _native.lib().aenet_structures_energy_and_forces(
            ns, natoms, latt, pbc, ids, coords.reshape(-1), E_tot.ctypes.data, E_coh.ctypes.data,
            E_at.ctypes.data, F.ctypes.data if forces else None, C.byref(stat))
        check_status(stat.value)
        return {"E_tot": E_tot, "E_coh": E_coh, "energies": E_at, "forces": F}

    def neighbor_lists(self, cell, pbc, positions, types, cartesian=True):
        """All neighbor lists of one structure through aenet_nbl_init /
        aenet_nbl_neighbors (lattice as given, as the reference's
        aenet_nbl_init does).  Returns (wrapped coords, list of dicts)."""
        L = _native.lib()
        ids = np.ascontiguousarray([int(t) if isinstance(t, (int, np.integer)) else self.get_type_id(t)
                                    for t in types], dtype=np.int32)
        coords = np.ascontiguousarray(positions, dtype=np.float64).reshape(-1).copy()
        latt = np.ascontiguousarray(cell, dtype=np.float64).reshape(9)
        L.aenet_nbl_init(latt, len(ids), ids, coords, bool(cartesian), bool(np.all(pbc)))
        cap = self.nnb_max
        out = []
        try:
            for i in range(len(ids)):
                n = C.c_int(cap)
                coo = np.zeros(3 * cap)
                dist = np.zeros(cap)
                lst = np.zeros(cap, np.int32)
                typ = np.zeros(cap, np.int32)
                L.aenet_nbl_neighbors(i + 1, C.byref(n), coo, dist, lst, typ)
                m = min(n.value, cap)
                out.append({"coo": coo[:3 * m].reshape(m, 3), "dist": dist[:m], "list": lst[:m], "type": typ[:m]})
        finally:
            L.aenet_nbl_final()
        return coords.reshape(-1, 3), out
