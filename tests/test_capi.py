"""The C-ABI library without a GPU: it loads, exports every symbol the
headers in include/ declare, and returns the reference's status codes on
the error paths of src/aenet.f90 (no compute calls)."""
import ctypes as C
import os
import re

import pytest

from conftest import ROOT, TI_O

from aenet_amd import _native


def _declared():
    names = set()
    for h in ("aenet.h", "aenet_amd.h"):
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names |= set(re.findall(r"\b(aenet_[a-z_]+)\s*\(", text))
        names |= set(re.findall(r"extern\s+\w+\s+(\w+)\s*;", text))
    return names


def test_library_exports_declared_symbols():
    lib = _native.lib()
    decl = _declared()
    assert len(decl) > 20
    for name in sorted(decl):
        assert hasattr(lib, name), name
    funcs, data = _native.exported_symbols()
    assert set(funcs) | set(data) >= decl


def test_headers_compile_as_c99(tmp_path):
    """include/*.h are C headers: a C99 translation unit includes both."""
    import subprocess
    src = tmp_path / "h.c"
    src.write_text('#include "aenet.h"\n#include "aenet_amd.h"\nint main(void) { return 0; }\n')
    subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-I" + os.path.join(ROOT, "include"), "-c", str(src),
                    "-o", str(tmp_path / "h.o")], check=True)


def test_status_constants():
    L = _native.lib()
    vals = {n: C.c_int.in_dll(L, n).value for n in
            ("AENET_OK", "AENET_ERR_INIT", "AENET_ERR_MALLOC", "AENET_ERR_IO", "AENET_ERR_TYPE",
             "AENET_TYPELEN", "AENET_PATHLEN")}
    # src/aenet.f90:101-108
    assert vals == {"AENET_OK": 0, "AENET_ERR_INIT": 1, "AENET_ERR_MALLOC": 2, "AENET_ERR_IO": 3,
                    "AENET_ERR_TYPE": 4, "AENET_TYPELEN": 2, "AENET_PATHLEN": 1024}
    assert C.c_bool.in_dll(L, "AENET_TRUE").value is True
    assert C.c_bool.in_dll(L, "AENET_FALSE").value is False


def test_error_paths_follow_reference():
    L = _native.lib()
    st = C.c_int(-1)
    # evaluation before aenet_init -> AENET_ERR_INIT (src/aenet.f90:380-383)
    import numpy as np
    E = np.zeros(1)
    L.aenet_atomic_energy(np.zeros(3), 1, 0, np.zeros(3), np.zeros(1, np.int32), E, C.byref(st))
    assert st.value == _native.AENET_ERR_INIT
    # load before init -> AENET_ERR_INIT (:261-264)
    L.aenet_load_potential(1, TI_O["Ti"].encode(), C.byref(st))
    assert st.value == _native.AENET_ERR_INIT
    names = (C.c_char_p * 2)(b"Ti", b"O")
    L.aenet_init(2, names, C.byref(st))
    assert st.value == 0
    try:
        # double init -> AENET_ERR_INIT (:150-153)
        L.aenet_init(2, names, C.byref(st))
        assert st.value == _native.AENET_ERR_INIT
        # type id out of range -> AENET_ERR_TYPE (:266-269)
        L.aenet_load_potential(3, TI_O["Ti"].encode(), C.byref(st))
        assert st.value == _native.AENET_ERR_TYPE
        L.aenet_load_potential(0, TI_O["Ti"].encode(), C.byref(st))
        assert st.value == _native.AENET_ERR_TYPE
        # missing file -> AENET_ERR_IO (:271-275)
        L.aenet_load_potential(1, b"/nonexistent/Ti.nn", C.byref(st))
        assert st.value == _native.AENET_ERR_IO
        # path longer than 1024 characters -> AENET_ERR_MALLOC (:305-309)
        L.aenet_load_potential(1, b"x" * 1100, C.byref(st))
        assert st.value == _native.AENET_ERR_MALLOC
        # one of two potentials loaded: not all loaded, evaluation refused
        L.aenet_load_potential(1, TI_O["Ti"].encode(), C.byref(st))
        assert st.value == 0
        assert not L.aenet_all_loaded()
        # free-atom energy of a loaded species (src/aenet.f90:337-346): the
        # E_atom of the training set record (src/potential.f90:150-151)
        from aenet_amd import nnfile
        p = nnfile.read_nn(TI_O["Ti"])
        e_ti = p["ts_E_atom"][p["ts_types"].index("Ti")]
        assert L.aenet_free_atom_energy(1) == e_ti
        assert np.isnan(L.aenet_free_atom_energy(2))      # not loaded yet
        L.aenet_atomic_energy(np.zeros(3), 1, 0, np.zeros(3), np.zeros(1, np.int32), E, C.byref(st))
        assert st.value == _native.AENET_ERR_INIT
        # type conversion by name (geo_type_conv, src/geometry.f90:440-461)
        tin = np.array([1, 2, 2, 1], np.int32)
        tout = np.zeros(4, np.int32)
        L.aenet_convert_atom_types(2, (C.c_char_p * 2)(b"O", b"Ti"), 4, tin, tout, C.byref(st))
        assert st.value == 0 and list(tout) == [2, 1, 1, 2]
        L.aenet_convert_atom_types(1, (C.c_char_p * 1)(b"Xx"), 1, np.array([1], np.int32), tout,
                                   C.byref(st))
        assert st.value == _native.AENET_ERR_TYPE
    finally:
        L.aenet_final(C.byref(st))
    assert st.value == 0


def test_nn_reader_writer_roundtrip(tmp_path):
    """The package's .nn reader on the reference's Ti potential (sizes
    parsed from python3/aenet/tests/Ti.15t-15t.nn) and a write/read
    round trip."""
    import numpy as np
    from aenet_amd import nnfile
    p = nnfile.read_nn(TI_O["Ti"])
    assert list(p["nnodes"]) == [44, 15, 15, 1]
    assert p["Wsize"] == 931 and list(p["iw"]) == [0, 675, 915, 931] and list(p["iv"]) == [0, 45, 61, 77]
    assert p["nsf"] == 44 and list(p["sfparam"][0]) == [8.0, 16.0, 6.5, 4.0]
    f = tmp_path / "Ti.nn"
    nnfile.write_nn(str(f), p)
    q = nnfile.read_nn(str(f))
    for k in ("W", "sfparam", "sfval_avg", "sfval_cov", "ts_E_atom"):
        assert np.array_equal(np.asarray(p[k]), np.asarray(q[k])), k
    assert open(TI_O["Ti"], "rb").read()[:20] == open(str(f), "rb").read()[:20]


def test_python_mirror_without_device():
    """ANNPotentials raises the reference's exception types on error paths
    (python3/aenet/core.pyx:189-202)."""
    from aenet_amd.core import ANNPotentials, InitError
    with pytest.raises(IOError):
        ANNPotentials({"Ti": "/nonexistent.nn", "O": TI_O["O"]})
    # the failed constructor released the library state
    st = C.c_int(0)
    names = (C.c_char_p * 1)(b"Ti")
    _native.lib().aenet_init(1, names, C.byref(st))
    assert st.value == 0
    _native.lib().aenet_final(C.byref(st))
    assert InitError is not None


def test_bucket_caps():
    """Neighbor-count buckets the autotuner would take (aenet_amd_bucket_caps,
    host only): powers of two from 16 below the capacity, then the
    capacity; a bucket with fewer than max(natoms/32, 2048) atoms joins the
    next larger one, and the last bucket always stays."""
    import numpy as np
    lib = _native.lib()

    def caps(counts, cap):
        c = np.ascontiguousarray(counts, dtype=np.int32)
        out = np.zeros(6, dtype=np.int32)
        n = C.c_int(0)
        lib.aenet_amd_bucket_caps(c, len(c), cap, out, C.byref(n))
        return out[:n.value].tolist()

    # clusters (10 neighbors) and bulk (80): a bucket each, the empty last stays
    assert caps([10] * 3000 + [80] * 5000, 160) == [16, 128, 160]
    # too few small atoms for a bucket of their own: they join the next one
    assert caps([10] * 1000 + [80] * 5000, 160) == [128, 160]
    # one neighbor count: one bucket up to it, then the capacity
    assert caps([80] * 8000, 128) == [128]
    assert caps([40] * 8000, 128) == [64, 128]
    # every count above the last power of two: one bucket
    assert caps([150] * 8000, 160) == [160]
    # the share rule: 1/32 of 200 000 atoms is 6250
    assert caps([20] * 6000 + [60] * 194000, 512) == [64, 512]
    assert caps([20] * 7000 + [60] * 193000, 512) == [32, 64, 512]
    # spread over every size: the small ones merged, at most six buckets
    assert caps(list(range(1, 513)) * 40, 512) == [64, 128, 256, 512]
    assert caps(list(range(1, 513)) * 200, 512) == [16, 32, 64, 128, 256, 512]
    assert caps([], 64) == [64]


def test_chebyshev_orders_beyond_the_register_layouts_are_refused(tmp_path):
    """The Chebyshev atom kernel holds a descriptor's terms in per-lane
    registers: radial order up to 23 with angular order up to 7, or both up
    to 15 (the reference sizes them at run time).  A potential beyond both
    layouts is refused when the set is complete (AENET_ERR_IO), before any
    device work."""
    from golden.make_potentials import make_pot
    from aenet_amd import nnfile
    L = _native.lib()
    st = C.c_int(-1)
    setup = nnfile.chebyshev_setup("A", ["A"], 6.0, 20, 4.0, 10)
    pot = make_pot("A", setup, [setup["nsf"], 8, 1], 5, ["A"], [-1000.0])
    path = str(tmp_path / "A.nn")
    nnfile.write_nn(path, pot)
    L.aenet_init(1, (C.c_char_p * 1)(b"A"), C.byref(st))
    assert st.value == 0
    try:
        L.aenet_load_potential(1, path.encode(), C.byref(st))
        assert st.value == _native.AENET_ERR_IO
    finally:
        L.aenet_final(C.byref(st))
