"""ctypes binding of libaenet_amd.so (the HIP implementation).

The library is built in-tree by ``python -m aenet_amd.build``.  There is no
fallback: if the shared library is missing, importing the bindings raises.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_lib", "libaenet_amd.so")

AENET_OK = 0
AENET_ERR_INIT = 1
AENET_ERR_MALLOC = 2
AENET_ERR_IO = 3
AENET_ERR_TYPE = 4
AENET_ERR_CAPACITY = 5
AENET_ERR_DEVICE = 6
AENET_ERR_ARG = 7
AENET_ERR_COORD = 8

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_pint = C.POINTER(C.c_int)
_vp = C.c_void_p

# C signatures of every function of include/aenet.h and include/aenet_amd.h
SIGNATURES = {
    "aenet_init": (None, [C.c_int, C.POINTER(C.c_char_p), _pint]),
    "aenet_final": (None, [_pint]),
    "aenet_print_info": (None, []),
    "aenet_load_potential": (None, [C.c_int, C.c_char_p, _pint]),
    "aenet_all_loaded": (C.c_bool, []),
    "aenet_free_atom_energy": (C.c_double, [C.c_int]),
    "aenet_atomic_energy": (None, [_dp, C.c_int, C.c_int, _dp, _ip, _dp, _pint]),
    "aenet_atomic_energy_and_forces": (None, [_dp, C.c_int, C.c_int, C.c_int, _dp, _ip, _ip, C.c_int, _dp,
                                              _dp, _pint]),
    "aenet_convert_atom_types": (None, [C.c_int, C.POINTER(C.c_char_p), C.c_int, _ip, _ip, _pint]),
    "aenet_sfb_init": (None, [C.c_int, C.POINTER(C.c_char_p), C.c_int, C.c_int, C.c_double, C.c_double,
                              _pint]),
    "aenet_sfb_final": (None, [_pint]),
    "aenet_sfb_nvalues": (C.c_int, []),
    "aenet_sfb_eval": (None, [C.c_int, _dp, C.c_int, _ip, _dp, C.c_int, _dp, _pint]),
    "aenet_sfb_eval_deriv": (None, [C.c_int, _dp, C.c_int, _ip, _dp, C.c_int, _dp, _vp, _vp, _pint]),
    "aenet_sfb_reconstruct_radial": (None, [C.c_int, _dp, C.c_int, _dp, _dp, _pint]),
    "aenet_nbl_init": (None, [_dp, C.c_int, _ip, _dp, C.c_bool, C.c_bool]),
    "aenet_nbl_final": (None, []),
    "aenet_nbl_neighbors": (None, [C.c_int, _pint, _dp, _dp, _ip, _ip]),
    "aenet_structures_energy_and_forces": (None, [C.c_int, _ip, _dp, _ip, _ip, _dp, _vp, _vp, _vp, _vp,
                                                  _pint]),
    "aenet_structures_energy_and_forces_noalloc": (None, [C.c_int, _ip, _dp, _ip, _ip, _dp, _vp, _vp, _vp, _vp,
                                                          _pint]),
    "aenet_amd_batch_create": (None, [C.c_int, _ip, _dp, _ip, _dp, C.POINTER(_vp), _pint]),
    "aenet_amd_batch_destroy": (None, [_vp]),
    "aenet_amd_batch_natoms": (C.c_int, [_vp]),
    "aenet_amd_batch_replan": (None, [_vp, _dp, _pint]),
    "aenet_amd_batch_has_isolated": (C.c_int, [_vp]),
    "aenet_amd_batch_eval": (None, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _pint]),
    "aenet_amd_batch_check": (None, [_vp, _vp, _pint]),
    "aenet_amd_batch_set_timing": (None, [_vp, C.c_int]),
    "aenet_amd_batch_kernel_time": (None, [_vp, C.POINTER(C.c_double), _pint, _pint]),
    "aenet_amd_batch_search_time": (None, [_vp, C.POINTER(C.c_double), _pint, _pint]),
    "aenet_amd_batch_neighbors": (None, [_vp, _vp, _vp, C.c_int, _vp, _vp, _vp, _vp, _vp, _vp, _pint]),
    "aenet_amd_batch_descriptors": (None, [_vp, _vp, _vp, C.c_int, _vp, C.c_int, _vp, _pint]),
    "aenet_amd_model_load": (None, [C.c_int, C.POINTER(C.c_char_p), C.POINTER(C.c_char_p), C.POINTER(_vp),
                                    _pint]),
    "aenet_amd_model_destroy": (None, [_vp]),
    "aenet_amd_model_info": (None, [_vp, C.POINTER(C.c_double), _pint, _pint, _pint]),
    "aenet_amd_model_species_weights": (None, [_vp, C.c_int, _pint, _pint, _pint]),
    "aenet_amd_model_weights": (None, [_vp, _vp, C.c_int, _vp, _pint]),
    "aenet_amd_model_descriptor_derivatives": (None, [_vp, C.c_int, _ip, _ip, _dp, _dp, _ip, C.c_int, _vp, _vp,
                                                      _vp, _pint]),
    "aenet_amd_batch_create_model": (None, [_vp, C.c_int, _ip, _dp, _ip, _dp, C.POINTER(_vp), _pint]),
    "aenet_amd_batch_eval_model": (None, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _pint]),
    "aenet_amd_batch_autotune": (None, [_vp, _vp, _vp, _vp, C.c_int, C.c_int, _vp, _pint, _pint, _vp, _pint]),
    "aenet_amd_batch_set_launch": (None, [_vp, _vp, C.c_int, C.c_int, C.c_int, _pint]),
    "aenet_amd_batch_set_buckets": (None, [_vp, _vp, C.c_int, _ip, _ip, _pint]),
    "aenet_amd_batch_buckets": (None, [_vp, _vp, _pint, _vp, _vp, _vp, _vp, _vp, _pint]),
    "aenet_amd_bucket_caps": (None, [_ip, C.c_int, C.c_int, _ip, _pint]),
    "aenet_amd_batch_eval_active": (None, [_vp, _vp, _vp, _vp, _vp, _vp, C.c_int, C.c_int, _vp, _vp, _pint]),
    "aenet_amd_batch_vjp": (None, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _pint]),
    "aenet_amd_batch_descriptors_model": (None, [_vp, _vp, _vp, _vp, C.c_int, _vp, C.c_int, _vp, _pint]),
    "aenet_amd_model_net_grad": (None, [_vp, C.c_int, _vp, C.c_int, _vp, _vp, _vp, _vp, C.c_double, _vp, _vp,
                                        _vp, _vp, _pint]),
    "aenet_amd_model_adam": (None, [_vp, _vp, _vp, _vp, C.c_int, C.c_double, C.c_double, C.c_double,
                                    C.c_double, _vp, _pint]),
    "aenet_amd_model_check": (None, [_vp, _vp, _pint]),
    "aenet_amd_mc_create": (None, [_vp, _vp, C.c_int, C.c_int, _dp, _ip, _dp, C.c_ulonglong, C.POINTER(_vp), _pint]),
    "aenet_amd_mc_destroy": (None, [_vp]),
    "aenet_amd_mc_set": (None, [_vp, C.c_double, C.c_double, C.c_double, C.c_int, _pint]),
    "aenet_amd_mc_refresh": (None, [_vp, _pint]),
    "aenet_amd_mc_run": (None, [_vp, C.c_int, C.c_int, C.c_int, C.c_int, _pint]),
    "aenet_amd_mc_stats": (None, [_vp, C.POINTER(C.c_longlong), C.c_int, _pint]),
    "aenet_amd_mc_energies": (None, [_vp, _dp, _pint]),
    "aenet_amd_mc_samples": (None, [_vp, _vp, C.POINTER(C.c_longlong), _pint]),
    "aenet_amd_mc_state": (None, [_vp, _vp, _vp, _pint]),
    "aenet_amd_mc_trace": (None, [_vp, C.c_longlong, _vp, C.POINTER(C.c_longlong), _pint]),
    "aenet_amd_mc_first_stage_state": (None, [_vp, _vp, _vp, _pint, _pint]),
    "aenet_amd_model_fit_cells": (None, [_vp, C.c_int, C.c_int, _dp, _vp, _vp, _vp, C.c_int, _vp, _vp, C.c_double,
                                         C.c_double, _vp, _vp, _vp, _pint]),
    "aenet_amd_last_error": (C.c_char_p, []),
}

_lib = None


class NativeLibraryMissing(ImportError):
    pass


def lib():
    """The loaded libaenet_amd.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise NativeLibraryMissing(
            "libaenet_amd.so not found at %s; build it with `python -m aenet_amd.build`" % LIB_PATH)
    L = C.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def last_error():
    msg = lib().aenet_amd_last_error()
    return msg.decode() if msg else ""


def global_int(name):
    return C.c_int.in_dll(lib(), name).value


def global_double(name):
    return C.c_double.in_dll(lib(), name).value


def exported_symbols():
    """Names declared in include/aenet.h and include/aenet_amd.h."""
    funcs = list(SIGNATURES)
    data = ["AENET_OK", "AENET_ERR_INIT", "AENET_ERR_MALLOC", "AENET_ERR_IO", "AENET_ERR_TYPE",
            "AENET_TYPELEN", "AENET_PATHLEN", "AENET_TRUE", "AENET_FALSE", "aenet_nsf_max", "aenet_nnb_max",
            "aenet_Rc_min", "aenet_Rc_max"]
    return funcs, data
