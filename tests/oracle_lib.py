"""ctypes access to the reference implementations the tests compare against
(test infrastructure only).

* ``Oracle``  -- the FP64 C restatement in oracle/aenet_oracle.c
  (oracle/_build/liboracle.so).
* ``Reference`` -- the reference's own libaenet, built from its Fortran
  sources by ``make -C oracle ref`` (oracle/_ref/libaenet_ref.so), with
  oracle/refdrv.f90 driving it the way predict.x does.

Both are test references; the product (aenet_amd) never imports this module.
"""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "_build", "liboracle.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libaenet_ref.so")

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


class Oracle:
    """FP64 restatement; one context per instance (1-based type ids)."""

    _lib = None

    @classmethod
    def lib(cls):
        if cls._lib is None:
            lib = C.CDLL(ORACLE_SO)
            lib.orc_create.restype = C.c_void_p
            lib.orc_create.argtypes = [C.c_int, C.POINTER(C.c_char_p)]
            lib.orc_destroy.argtypes = [C.c_void_p]
            lib.orc_load.argtypes = [C.c_void_p, C.c_int, C.c_char_p]
            for f in ("orc_rc_min", "orc_rc_max"):
                getattr(lib, f).restype = C.c_double
                getattr(lib, f).argtypes = [C.c_void_p]
            lib.orc_nnb_max.argtypes = [C.c_void_p]
            lib.orc_nsf.argtypes = [C.c_void_p, C.c_int]
            lib.orc_wsize.argtypes = [C.c_void_p, C.c_int]
            lib.orc_e_atom.restype = C.c_double
            lib.orc_e_atom.argtypes = [C.c_void_p, C.c_int]
            lib.orc_descriptors.argtypes = [C.c_void_p, C.c_int, _dp, C.c_int, _dp, _ip, C.c_int,
                                            _dp, C.c_void_p, C.c_void_p]
            lib.orc_network.argtypes = [C.c_void_p, C.c_int, _dp, _dp, C.c_void_p, C.c_void_p]
            lib.orc_atomic_energy.argtypes = [C.c_void_p, _dp, C.c_int, C.c_int, _dp, _ip, _dp]
            lib.orc_atomic_energy_and_forces.argtypes = [C.c_void_p, _dp, C.c_int, C.c_int, C.c_int,
                                                         _dp, _ip, _ip, C.c_int, _dp, _dp]
            lib.orc_cart_to_frac.argtypes = [_dp, C.c_int, C.c_int, _dp, _dp, _dp]
            lib.orc_neighbors.argtypes = [C.c_double, C.c_double, _dp, C.c_int, C.c_int, _ip, _dp,
                                          C.c_int, _ip, _dp, _dp, _ip, _ip]
            lib.orc_structure.argtypes = [C.c_void_p, _dp, C.c_int, C.c_int, _ip, _dp, C.c_int,
                                          C.c_int, _dp, _dp, C.c_void_p, C.c_void_p]
            lib.orc_nmax_nbdist.argtypes = [C.c_double, C.c_double]
            lib.orc_adam_step.argtypes = [C.c_int, _dp, _dp, _dp, _dp, C.c_int, C.c_double,
                                          C.c_double, C.c_double, C.c_double]
            cls._lib = lib
        return cls._lib

    def __init__(self, types, files):
        lib = self.lib()
        arr = (C.c_char_p * len(types))(*[t.encode() for t in types])
        self.types = list(types)
        self.ctx = lib.orc_create(len(types), arr)
        for i, t in enumerate(types):
            st = lib.orc_load(self.ctx, i + 1, files[t].encode())
            if st != 0:
                raise RuntimeError("oracle failed to load %s (status %d)" % (files[t], st))

    def __del__(self):
        if getattr(self, "ctx", None):
            self.lib().orc_destroy(self.ctx)
            self.ctx = None

    @property
    def Rc_min(self):
        return self.lib().orc_rc_min(self.ctx)

    @property
    def Rc_max(self):
        return self.lib().orc_rc_max(self.ctx)

    @property
    def nnb_max(self):
        return self.lib().orc_nnb_max(self.ctx)

    def nsf(self, t):
        return self.lib().orc_nsf(self.ctx, t)

    def e_atom(self, t):
        return self.lib().orc_e_atom(self.ctx, t)

    def descriptors(self, type_i, coo_i, coo_j, type_j, scaled=True, derivs=False):
        n = len(type_j)
        nsf = self.nsf(type_i)
        G = np.zeros(nsf)
        coo_j = _f64(np.reshape(coo_j, (-1, 3))) if n else np.zeros((1, 3))
        tj = _i32(type_j) if n else np.zeros(1, np.int32)
        if derivs:
            dGi = np.zeros((nsf, 3))
            dGj = np.zeros((max(n, 1), nsf, 3))
            st = self.lib().orc_descriptors(self.ctx, type_i, _f64(coo_i), n, coo_j, tj, int(scaled), G,
                                            dGi.ctypes.data, dGj.ctypes.data)
        else:
            st = self.lib().orc_descriptors(self.ctx, type_i, _f64(coo_i), n, coo_j, tj, int(scaled), G,
                                            None, None)
        assert st == 0, st
        return (G, dGi, dGj[:n]) if derivs else G

    def network(self, t, x, wsize):
        y = np.zeros(1)
        dx = np.zeros(len(x))
        dw = np.zeros(wsize)
        st = self.lib().orc_network(self.ctx, t, _f64(x), y, dx.ctypes.data, dw.ctypes.data)
        assert st == 0
        return y[0], dx, dw

    def atomic_energy_and_forces(self, coo_i, type_i, index_i, coo_j, type_j, index_j, F):
        E = np.zeros(1)
        n = len(type_j)
        st = self.lib().orc_atomic_energy_and_forces(
            self.ctx, _f64(coo_i), type_i, index_i, n, _f64(np.reshape(coo_j, (-1, 3))) if n else np.zeros(3),
            _i32(type_j) if n else np.zeros(1, np.int32), _i32(index_j) if n else np.zeros(1, np.int32),
            F.shape[0], E, F)
        assert st == 0, st
        return E[0]

    def structure(self, latt, pbc, types, cart, forces=True, rank=0, nranks=1):
        """latt rows are lattice vectors (ASE convention); types 1-based."""
        n = len(types)
        L = _f64(np.asarray(latt, dtype=np.float64))  # row-major rows == column-major columns
        e = np.zeros(1)
        t = np.zeros(1)
        Eat = np.zeros(n)
        F = np.zeros((n, 3))
        st = self.lib().orc_structure(self.ctx, L.ravel(), int(pbc), n, _i32(types), _f64(cart).ravel(),
                                      rank, nranks, e, t, Eat.ctypes.data,
                                      F.ctypes.data if forces else None)
        if st != 0:
            raise RuntimeError("oracle structure evaluation failed (%d)" % st)
        return {"Ecoh": e[0], "Etot": t[0], "energies": Eat, "forces": F}

    def neighbors(self, latt, pbc, types, cart, cap=None):
        return neighbors_via(self.lib(), "orc", self.Rc_min, self.Rc_max, latt, pbc, types, cart,
                             cap or self.nnb_max)


def neighbors_via(lib, kind, rc_min, rc_max, latt, pbc, types, cart, cap):
    n = len(types)
    L = _f64(latt).ravel()
    frac = np.zeros((n, 3))
    Lout = np.zeros(9)
    if kind == "orc":
        lib.orc_cart_to_frac(L, int(pbc), n, _f64(cart).ravel(), frac, Lout)
    else:
        lib.refdrv_cart_to_frac(L, int(pbc), n, _f64(cart).ravel(), frac, Lout)
    nnb = np.zeros(n, np.int32)
    nbcoo = np.zeros((n, cap, 3))
    nbdist = np.zeros((n, cap))
    nblist = np.zeros((n, cap), np.int32)
    nbtype = np.zeros((n, cap), np.int32)
    if kind == "orc":
        st = lib.orc_neighbors(rc_min, rc_max, Lout, int(pbc), n, _i32(types), frac, cap, nnb, nbcoo,
                               nbdist, nblist, nbtype)
    else:
        stat = np.zeros(1, np.int32)
        lib.refdrv_neighbors(Lout, int(pbc), n, _i32(types), frac, cap, nnb, nbcoo, nbdist, nblist,
                             nbtype, stat)
        st = int(stat[0])
    if st != 0:
        raise RuntimeError("neighbor list failed (%d)" % st)
    return {"frac": frac, "latt": Lout.reshape(3, 3), "nnb": nnb, "coo": nbcoo, "dist": nbdist,
            "list": nblist, "type": nbtype}


def _raise_stack_limit():
    """The flang-built reference keeps large array temporaries on the stack
    (Behler-Parrinello path); let the main thread's stack grow."""
    import resource
    soft, hard = resource.getrlimit(resource.RLIMIT_STACK)
    want = resource.RLIM_INFINITY if hard == resource.RLIM_INFINITY else hard
    if soft != want:
        resource.setrlimit(resource.RLIMIT_STACK, (want, hard))


class Reference:
    """The reference libaenet (flang build).  Module-global state: one
    potential set per process, as in the reference library itself."""

    _lib = None

    @classmethod
    def available(cls):
        return os.path.exists(REF_SO)

    @classmethod
    def lib(cls):
        if cls._lib is None:
            _raise_stack_limit()
            lib = C.CDLL(REF_SO)
            lib.aenet_init.argtypes = [C.c_int, C.POINTER(C.c_char_p), C.POINTER(C.c_int)]
            lib.aenet_final.argtypes = [C.POINTER(C.c_int)]
            lib.aenet_load_potential.argtypes = [C.c_int, C.c_char_p, C.POINTER(C.c_int)]
            lib.aenet_all_loaded.restype = C.c_bool
            lib.aenet_atomic_energy.argtypes = [_dp, C.c_int, C.c_int, _dp, _ip, _dp, C.POINTER(C.c_int)]
            lib.aenet_atomic_energy_and_forces.argtypes = [_dp, C.c_int, C.c_int, C.c_int, _dp, _ip, _ip,
                                                           C.c_int, _dp, _dp, C.POINTER(C.c_int)]
            lib.aenet_sfb_init.argtypes = [C.c_int, C.POINTER(C.c_char_p), C.c_int, C.c_int, C.c_double,
                                           C.c_double, C.POINTER(C.c_int)]
            lib.aenet_sfb_final.argtypes = [C.POINTER(C.c_int)]
            lib.aenet_sfb_eval.argtypes = [C.c_int, _dp, C.c_int, _ip, _dp, C.c_int, _dp, C.POINTER(C.c_int)]
            lib.aenet_sfb_nvalues.restype = C.c_int
            lib.aenet_sfb_reconstruct_radial.argtypes = [C.c_int, _dp, C.c_int, _dp, _dp, C.POINTER(C.c_int)]
            lib.refdrv_cart_to_frac.argtypes = [_dp, C.c_int, C.c_int, _dp, _dp, _dp]
            lib.refdrv_neighbors.argtypes = [_dp, C.c_int, C.c_int, _ip, _dp, C.c_int, _ip, _dp, _dp, _ip,
                                             _ip, _ip]
            lib.refdrv_energy.argtypes = [_dp, C.c_int, C.c_int, _ip, _dp, C.c_int, C.c_int, C.c_int, _dp,
                                          _dp, _dp, _dp, _ip]
            cls._lib = lib
        return cls._lib

    _loaded = None

    def __init__(self, types, files):
        lib = self.lib()
        key = (tuple(types), tuple(files[t] for t in types))
        if Reference._loaded != key:
            st = C.c_int(0)
            if Reference._loaded is not None:
                lib.aenet_final(C.byref(st))
            arr = (C.c_char_p * len(types))(*[t.encode() for t in types])
            lib.aenet_init(len(types), arr, C.byref(st))
            assert st.value == 0, st.value
            for i, t in enumerate(types):
                lib.aenet_load_potential(i + 1, files[t].encode(), C.byref(st))
                assert st.value == 0, (files[t], st.value)
            Reference._loaded = key
        self.types = list(types)

    @property
    def Rc_min(self):
        return C.c_double.in_dll(self.lib(), "aenet_Rc_min").value

    @property
    def Rc_max(self):
        return C.c_double.in_dll(self.lib(), "aenet_Rc_max").value

    @property
    def nnb_max(self):
        return C.c_int.in_dll(self.lib(), "aenet_nnb_max").value

    def structure(self, latt, pbc, types, cart, forces=True, rank=0, nranks=1):
        lib = self.lib()
        n = len(types)
        frac = np.zeros((n, 3))
        Lout = np.zeros(9)
        lib.refdrv_cart_to_frac(_f64(latt).ravel(), int(pbc), n, _f64(cart).ravel(), frac, Lout)
        e = np.zeros(1)
        t = np.zeros(1)
        Eat = np.zeros(n)
        F = np.zeros((n, 3))
        stat = np.zeros(1, np.int32)
        lib.refdrv_energy(Lout, int(pbc), n, _i32(types), frac, rank, nranks, int(forces), e, t, Eat, F,
                          stat)
        assert stat[0] == 0, stat[0]
        return {"Ecoh": e[0], "Etot": t[0], "energies": Eat, "forces": F}

    def neighbors(self, latt, pbc, types, cart, cap=None):
        return neighbors_via(self.lib(), "ref", self.Rc_min, self.Rc_max, latt, pbc, types, cart,
                             cap or self.nnb_max)

    def atomic_energy_and_forces(self, coo_i, type_i, index_i, coo_j, type_j, index_j, F):
        E = np.zeros(1)
        st = C.c_int(0)
        n = len(type_j)
        self.lib().aenet_atomic_energy_and_forces(_f64(coo_i), type_i, index_i, n,
                                                  _f64(np.reshape(coo_j, (-1, 3))), _i32(type_j),
                                                  _i32(index_j), F.shape[0], E, F, C.byref(st))
        assert st.value == 0
        return E[0]
