"""Reader/writer for ænet ``.nn`` potential files.

A ``.nn`` file is a Fortran unformatted sequential file (4-byte record
markers before and after every record) holding, in this order:

* the network      -- load_Network      (src/ext/feedforward.f90:291-306)
* the descriptor   -- load_Setup        (src/sfsetup.f90:585-609)
* training info    -- load_TrnSet_info  (src/trainset.f90:430-440)

36 records in total.  The layout is the reference's on-disk format and is
kept identical so that files move freely between the two codes.
"""
import struct

import numpy as np

TYPELEN = 2
ACTIVATIONS = {"linear": 0, "tanh": 1, "sigmoid": 2, "mtanh": 3, "twist": 4}


class NNFormatError(ValueError):
    pass


def _records(data):
    recs = []
    i = 0
    while i < len(data):
        if i + 4 > len(data):
            raise NNFormatError("truncated record marker")
        (n,) = struct.unpack("<i", data[i:i + 4])
        if n < 0 or i + 8 + n > len(data):
            raise NNFormatError("bad record length")
        (m,) = struct.unpack("<i", data[i + 4 + n:i + 8 + n])
        if m != n:
            raise NNFormatError("record markers disagree")
        recs.append(data[i + 4:i + 4 + n])
        i += 8 + n
    return recs


def _ints(rec, n=None):
    a = np.frombuffer(rec, dtype="<i4")
    if n is not None and a.size != n:
        raise NNFormatError("expected %d integers, found %d" % (n, a.size))
    return a.astype(np.int64)


def _dbls(rec, n=None):
    a = np.frombuffer(rec, dtype="<f8")
    if n is not None and a.size != n:
        raise NNFormatError("expected %d reals, found %d" % (n, a.size))
    return a.copy()


def _chars(rec, k):
    s = rec.decode("latin-1")
    return [s[i:i + k].rstrip(" \x00") for i in range(0, len(s), k)]


def read_nn(path):
    """Return a dict with the network, descriptor set-up and energy
    constants of a ``.nn`` file."""
    with open(path, "rb") as f:
        recs = _records(f.read())
    if len(recs) < 35:
        raise NNFormatError("%s: %d records, expected 36" % (path, len(recs)))
    r = iter(recs)
    nlayers = int(_ints(next(r), 1)[0])
    nnodes_max = int(_ints(next(r), 1)[0])
    wsize = int(_ints(next(r), 1)[0])
    nvalues = int(_ints(next(r), 1)[0])
    nnodes = _ints(next(r), nlayers)
    f_a = _ints(next(r), nlayers - 1)
    iw = _ints(next(r), nlayers)
    iv = _ints(next(r), nlayers)
    W = _dbls(next(r), wsize)
    description = next(r).decode("latin-1").rstrip()
    atomtype = _chars(next(r), TYPELEN)[0]
    nenv = int(_ints(next(r), 1)[0])
    envtypes = _chars(next(r), TYPELEN)
    rc_min = float(_dbls(next(r), 1)[0])
    rc_max = float(_dbls(next(r), 1)[0])
    sftype = next(r).decode("latin-1").rstrip()
    nsf = int(_ints(next(r), 1)[0])
    nsfparam = int(_ints(next(r), 1)[0])
    sf = _ints(next(r), nsf)
    sfparam = _dbls(next(r), nsfparam * nsf).reshape(nsf, nsfparam)
    sfenv = _ints(next(r), 2 * nsf).reshape(nsf, 2)
    neval = int(_ints(next(r), 1)[0])
    sfval_min = _dbls(next(r), nsf)
    sfval_max = _dbls(next(r), nsf)
    sfval_avg = _dbls(next(r), nsf)
    sfval_cov = _dbls(next(r), nsf) if len(recs) >= 36 else None
    ts_file = next(r).decode("latin-1").rstrip()
    ts_normalized = bool(_ints(next(r), 1)[0])
    ts_scale = float(_dbls(next(r), 1)[0])
    ts_shift = float(_dbls(next(r), 1)[0])
    ts_ntypes = int(_ints(next(r), 1)[0])
    ts_types = _chars(next(r), TYPELEN)
    ts_eatom = _dbls(next(r), ts_ntypes)
    ts_natoms = int(_ints(next(r), 1)[0])
    ts_nstrucs = int(_ints(next(r), 1)[0])
    ts_e = _dbls(next(r), 3)
    if sfval_cov is None:
        raise NNFormatError("legacy 35-record .nn files have no sfval_cov record")
    # E_scale/E_shift/E_atom as load_NNPot derives them (src/potential.f90:145-151)
    e_atom = 0.0
    for t, e in zip(ts_types, ts_eatom):
        if t.strip() == atomtype.strip():
            e_atom = float(e)
            break
    return {
        "nlayers": nlayers, "nnodes_max": nnodes_max, "Wsize": wsize, "nvalues": nvalues,
        "nnodes": nnodes, "f_a": f_a, "iw": iw, "iv": iv, "W": W,
        "description": description, "atomtype": atomtype, "envtypes": envtypes,
        "Rc_min": rc_min, "Rc_max": rc_max, "sftype": sftype, "nsf": nsf, "sf": sf,
        "sfparam": sfparam, "sfenv": sfenv, "neval": neval, "sfval_min": sfval_min,
        "sfval_max": sfval_max, "sfval_avg": sfval_avg, "sfval_cov": sfval_cov,
        "ts_file": ts_file, "ts_normalized": ts_normalized, "ts_scale": ts_scale,
        "ts_shift": ts_shift, "ts_types": ts_types, "ts_E_atom": ts_eatom,
        "ts_natoms": ts_natoms, "ts_nstrucs": ts_nstrucs, "ts_E": ts_e,
        "E_scale": 1.0 / ts_scale, "E_shift": ts_shift, "E_atom": e_atom,
    }


def _rec(payload):
    n = len(payload)
    return struct.pack("<i", n) + payload + struct.pack("<i", n)


def _pad(s, k):
    b = s.encode("latin-1")[:k]
    return b + b" " * (k - len(b))


def network_layout(arch):
    """iw/iv/Wsize/nvalues of new_Network (src/ext/feedforward.f90:97-149)."""
    arch = [int(a) for a in arch]
    nl = len(arch)
    iw = [0] * nl
    iv = [0] * nl
    wsize = 0
    nvalues = 0
    for l in range(nl - 1):
        wsize += (arch[l] + 1) * arch[l + 1]
        iw[l + 1] = wsize
        nvalues += arch[l] + 1
        iv[l + 1] = nvalues
    nvalues += arch[-1]
    return iw, iv, wsize, nvalues


def write_nn(path, pot):
    """Write a potential dict (keys as returned by :func:`read_nn`; iw/iv/
    Wsize/nvalues/nnodes_max are derived from ``nnodes``)."""
    nnodes = [int(x) for x in pot["nnodes"]]
    nl = len(nnodes)
    iw, iv, wsize, nvalues = network_layout(nnodes)
    W = np.ascontiguousarray(pot["W"], dtype="<f8")
    if W.size != wsize:
        raise NNFormatError("W has %d entries, architecture needs %d" % (W.size, wsize))
    nsf = int(pot["nsf"])
    sfparam = np.ascontiguousarray(pot["sfparam"], dtype="<f8").reshape(nsf, -1)
    nsfparam = sfparam.shape[1]
    envtypes = list(pot["envtypes"])
    ts_types = list(pot["ts_types"])
    out = b""
    out += _rec(struct.pack("<i", nl))
    out += _rec(struct.pack("<i", max(nnodes)))
    out += _rec(struct.pack("<i", wsize))
    out += _rec(struct.pack("<i", nvalues))
    out += _rec(np.asarray(nnodes, "<i4").tobytes())
    out += _rec(np.asarray(pot["f_a"], "<i4").tobytes())
    out += _rec(np.asarray(iw, "<i4").tobytes())
    out += _rec(np.asarray(iv, "<i4").tobytes())
    out += _rec(W.tobytes())
    out += _rec(_pad(pot.get("description", ""), 1024))
    out += _rec(_pad(pot["atomtype"], TYPELEN))
    out += _rec(struct.pack("<i", len(envtypes)))
    out += _rec(b"".join(_pad(t, TYPELEN) for t in envtypes))
    out += _rec(struct.pack("<d", float(pot["Rc_min"])))
    out += _rec(struct.pack("<d", float(pot["Rc_max"])))
    out += _rec(_pad(pot["sftype"], 100))
    out += _rec(struct.pack("<i", nsf))
    out += _rec(struct.pack("<i", nsfparam))
    out += _rec(np.asarray(pot["sf"], "<i4").tobytes())
    out += _rec(sfparam.tobytes())
    out += _rec(np.ascontiguousarray(pot["sfenv"], "<i4").reshape(nsf, 2).tobytes())
    out += _rec(struct.pack("<i", int(pot.get("neval", 0))))
    for k in ("sfval_min", "sfval_max", "sfval_avg", "sfval_cov"):
        out += _rec(np.ascontiguousarray(pot[k], "<f8").tobytes())
    out += _rec(_pad(pot.get("ts_file", ""), 1024))
    out += _rec(struct.pack("<i", 1 if pot.get("ts_normalized", True) else 0))
    out += _rec(struct.pack("<d", float(pot.get("ts_scale", 1.0))))
    out += _rec(struct.pack("<d", float(pot.get("ts_shift", 0.0))))
    out += _rec(struct.pack("<i", len(ts_types)))
    out += _rec(b"".join(_pad(t, TYPELEN) for t in ts_types))
    out += _rec(np.asarray(pot["ts_E_atom"], "<f8").tobytes())
    out += _rec(struct.pack("<i", int(pot.get("ts_natoms", 0))))
    out += _rec(struct.pack("<i", int(pot.get("ts_nstrucs", 0))))
    out += _rec(np.asarray(pot.get("ts_E", [0.0, 0.0, 0.0]), "<f8").tobytes())
    with open(path, "wb") as f:
        f.write(out)


def chebyshev_setup(atomtype, envtypes, radial_Rc, radial_N, angular_Rc, angular_N, rmin=1.0):
    """Descriptor part of a Chebyshev set-up as read_basis_chebyshev stores
    it (src/sfsetup.f90:1102-1158): nsf = r_N + a_N + 2, doubled for more
    than one species; parameters only in column 1 of sfparam."""
    nsf = radial_N + angular_N + 2
    if len(envtypes) > 1:
        nsf *= 2
    sfparam = np.zeros((nsf, 4))
    sfparam[0] = [radial_Rc, radial_N, angular_Rc, angular_N]
    return {
        "atomtype": atomtype, "envtypes": list(envtypes), "Rc_min": rmin,
        "Rc_max": max(radial_Rc, angular_Rc), "sftype": "Chebyshev", "nsf": nsf,
        "sf": np.zeros(nsf, np.int64), "sfparam": sfparam, "sfenv": np.zeros((nsf, 2), np.int64),
    }


def behler_setup(atomtype, envtypes, functions, rmin=1.0):
    """Behler2011 set-up (src/sfsetup.f90:1222-1280).  ``functions`` is a
    list of dicts: {"G": 2, "type2": "A", "Rc":, "Rs":, "eta":} or
    {"G": 4, "type2":, "type3":, "Rc":, "lambda":, "zeta":, "eta":}."""
    nsf = len(functions)
    sf = np.zeros(nsf, np.int64)
    sfparam = np.zeros((nsf, 4))
    sfenv = np.zeros((nsf, 2), np.int64)
    rc_max = 0.0
    for i, fn in enumerate(functions):
        g = int(fn["G"])
        sf[i] = g
        sfenv[i, 0] = envtypes.index(fn["type2"]) + 1
        if g == 1:
            sfparam[i, 0] = fn["Rc"]
        elif g == 2:
            sfparam[i, :3] = [fn["Rc"], fn["Rs"], fn["eta"]]
        elif g == 3:
            sfparam[i, :2] = [fn["Rc"], fn["kappa"]]
        elif g in (4, 5):
            sfenv[i, 1] = envtypes.index(fn["type3"]) + 1
            sfparam[i] = [fn["Rc"], fn["lambda"], fn["zeta"], fn["eta"]]
        else:
            raise ValueError("unknown symmetry function G%d" % g)
        rc_max = max(rc_max, sfparam[i, 0])
    return {
        "atomtype": atomtype, "envtypes": list(envtypes), "Rc_min": rmin, "Rc_max": rc_max,
        "sftype": "Behler2011", "nsf": nsf, "sf": sf, "sfparam": sfparam, "sfenv": sfenv,
    }
