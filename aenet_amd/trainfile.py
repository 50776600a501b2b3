"""Reader/writer for ænet ``.train`` training-set files.

A ``.train`` file is a Fortran unformatted sequential file (4-byte record
markers around every record), written by generate.x and read by train.x
(src/trainset.f90):

* header  -- ts_write_header (:739-756): nTypes, nStrucs, typeName(:)
  (2 characters each), E_atom(:), normalized (4-byte logical), scale,
  shift;
* per structure -- ts_write_structure_info (:911-953): len_trim(name),
  name, (nAtoms, nTypes), energy; then per atom ts_write_atom_info
  (:990-1006): itype, cooCart(3), forCart(3), and ts_write_sf_info
  (:1030-1044): nsf, sfval(nsf);
* footer  -- ts_write_footer (:829-865): nAtomsTot, (E_av, E_min, E_max),
  has_setups (logical) and, per species, itype and the set-up records of
  save_Setup (src/sfsetup.f90:459-505).

The layout is the reference's, so files move freely between the codes.
"""
import struct

import numpy as np

from .nnfile import NNFormatError, TYPELEN, _chars, _dbls, _ints, _pad, _rec


class RecordReader:
    def __init__(self, path):
        self.f = open(path, "rb")

    def close(self):
        self.f.close()

    def next(self):
        h = self.f.read(4)
        if len(h) < 4:
            raise NNFormatError("unexpected end of file")
        (n,) = struct.unpack("<i", h)
        data = self.f.read(n)
        t = self.f.read(4)
        if len(data) != n or len(t) != 4 or struct.unpack("<i", t)[0] != n:
            raise NNFormatError("bad record")
        return data


def _logical(rec):
    return struct.unpack("<i", rec)[0] != 0


SETUP_RECORDS = ("description", "atomtype", "nenv", "envtypes", "Rc_min", "Rc_max", "sftype", "nsf", "nsfparam",
                 "sf", "sfparam", "sfenv", "neval", "sfval_min", "sfval_max", "sfval_avg", "sfval_cov")


def setup_records(stp):
    """The 17 records of save_Setup (src/sfsetup.f90:483-499)."""
    nsf = int(stp["nsf"])
    sfparam = np.ascontiguousarray(stp["sfparam"], dtype="<f8").reshape(nsf, -1)
    out = b""
    out += _rec(_pad(stp.get("description", ""), 1024))
    out += _rec(_pad(stp["atomtype"], TYPELEN))
    out += _rec(struct.pack("<i", len(stp["envtypes"])))
    out += _rec(b"".join(_pad(t, TYPELEN) for t in stp["envtypes"]))
    out += _rec(struct.pack("<d", float(stp["Rc_min"])))
    out += _rec(struct.pack("<d", float(stp["Rc_max"])))
    out += _rec(_pad(stp["sftype"], 100))
    out += _rec(struct.pack("<i", nsf))
    out += _rec(struct.pack("<i", sfparam.shape[1]))
    out += _rec(np.asarray(stp["sf"], "<i4").tobytes())
    out += _rec(sfparam.tobytes())
    out += _rec(np.ascontiguousarray(stp["sfenv"], "<i4").reshape(nsf, 2).tobytes())
    out += _rec(struct.pack("<i", int(stp.get("neval", 0))))
    for k in ("sfval_min", "sfval_max", "sfval_avg", "sfval_cov"):
        out += _rec(np.ascontiguousarray(stp[k], "<f8").tobytes())
    return out


def read_setup(rr):
    description = rr.next().decode("latin-1").rstrip()
    atomtype = _chars(rr.next(), TYPELEN)[0]
    nenv = int(_ints(rr.next(), 1)[0])
    envtypes = _chars(rr.next(), TYPELEN)
    rc_min = float(_dbls(rr.next(), 1)[0])
    rc_max = float(_dbls(rr.next(), 1)[0])
    sftype = rr.next().decode("latin-1").rstrip()
    nsf = int(_ints(rr.next(), 1)[0])
    nsfparam = int(_ints(rr.next(), 1)[0])
    sf = _ints(rr.next(), nsf)
    sfparam = _dbls(rr.next(), nsfparam * nsf).reshape(nsf, nsfparam)
    sfenv = _ints(rr.next(), 2 * nsf).reshape(nsf, 2)
    neval = int(_ints(rr.next(), 1)[0])
    st = {"description": description, "atomtype": atomtype, "envtypes": envtypes[:nenv], "Rc_min": rc_min,
          "Rc_max": rc_max, "sftype": sftype, "nsf": nsf, "sf": sf, "sfparam": sfparam, "sfenv": sfenv,
          "neval": neval}
    for k in ("sfval_min", "sfval_max", "sfval_avg", "sfval_cov"):
        st[k] = _dbls(rr.next(), nsf)
    return st


class TrainWriter:
    """Streams a .train file: header at creation, then structures, then
    the footer at close (statistics kept as ts_write_structure_info keeps
    them)."""

    def __init__(self, path, types, E_atom, nstrucs, normalized=False, scale=1.0, shift=0.0):
        self.f = open(path, "wb")
        self.types = list(types)
        self.nstrucs = int(nstrucs)
        self.istruc = 0
        self.natoms_tot = 0
        self.E_av = self.E_min = self.E_max = 0.0
        out = _rec(struct.pack("<i", len(types)))
        out += _rec(struct.pack("<i", self.nstrucs))
        out += _rec(b"".join(_pad(t, TYPELEN) for t in types))
        out += _rec(np.asarray(E_atom, "<f8").tobytes())
        out += _rec(struct.pack("<i", 1 if normalized else 0))
        out += _rec(struct.pack("<d", float(scale)))
        out += _rec(struct.pack("<d", float(shift)))
        self.f.write(out)

    def structure(self, name, ntypes, energy, itype, coords, forces, sfvals):
        """itype: per atom, 1-based; coords/forces (n, 3); sfvals: per atom
        1-d arrays (each atom's own feature count)."""
        n = len(itype)
        self.istruc += 1
        b = name.encode("latin-1")
        out = _rec(struct.pack("<i", len(b))) + _rec(b)
        out += _rec(struct.pack("<ii", n, ntypes))
        out += _rec(struct.pack("<d", float(energy)))
        for i in range(n):
            out += _rec(struct.pack("<i", int(itype[i])))
            out += _rec(np.ascontiguousarray(coords[i], "<f8").tobytes())
            out += _rec(np.ascontiguousarray(forces[i], "<f8").tobytes())
            v = np.ascontiguousarray(sfvals[i], "<f8")
            out += _rec(struct.pack("<i", v.size))
            out += _rec(v.tobytes())
        self.f.write(out)
        # energy statistics as ts_write_structure_info (src/trainset.f90:938-949)
        e = float(energy) / n
        if self.istruc > 1:
            self.E_min = min(self.E_min, e)
            self.E_max = max(self.E_max, e)
            self.E_av = self.E_av + e / self.nstrucs
        else:
            self.E_min = self.E_max = e
            self.E_av = e / self.nstrucs
        self.natoms_tot += n

    def close(self, setups=None):
        out = _rec(struct.pack("<i", self.natoms_tot))
        out += _rec(struct.pack("<ddd", self.E_av, self.E_min, self.E_max))
        out += _rec(struct.pack("<i", 1 if setups is not None else 0))
        if setups is not None:
            for k, stp in enumerate(setups):
                out += _rec(struct.pack("<i", k + 1))
                out += setup_records(stp)
        self.f.write(out)
        self.f.close()


def read_train(path):
    """The whole file as a dict: header fields, a list of structures
    (name, natoms, ntypes, energy, itype, coords, forces, sfval) and the
    footer (natoms_tot, E_av, E_min, E_max, setups or None)."""
    rr = RecordReader(path)
    try:
        ntypes = int(_ints(rr.next(), 1)[0])
        nstrucs = int(_ints(rr.next(), 1)[0])
        types = _chars(rr.next(), TYPELEN)
        E_atom = _dbls(rr.next(), ntypes)
        normalized = _logical(rr.next())
        scale = float(_dbls(rr.next(), 1)[0])
        shift = float(_dbls(rr.next(), 1)[0])
        strucs = []
        for _ in range(nstrucs):
            ln = int(_ints(rr.next(), 1)[0])
            name = rr.next().decode("latin-1")[:ln]
            nat, nty = (int(v) for v in _ints(rr.next(), 2))
            energy = float(_dbls(rr.next(), 1)[0])
            itype = np.zeros(nat, np.int64)
            coords = np.zeros((nat, 3))
            forces = np.zeros((nat, 3))
            sfval = []
            for i in range(nat):
                itype[i] = int(_ints(rr.next(), 1)[0])
                coords[i] = _dbls(rr.next(), 3)
                forces[i] = _dbls(rr.next(), 3)
                nsf = int(_ints(rr.next(), 1)[0])
                sfval.append(_dbls(rr.next(), nsf))
            strucs.append({"name": name, "natoms": nat, "ntypes": nty, "energy": energy, "itype": itype,
                           "coords": coords, "forces": forces, "sfval": sfval})
        natoms_tot = int(_ints(rr.next(), 1)[0])
        E_av, E_min, E_max = (float(v) for v in _dbls(rr.next(), 3))
        has = _logical(rr.next())
        setups = None
        if has:
            setups = [None] * ntypes
            for _ in range(ntypes):
                it = int(_ints(rr.next(), 1)[0])
                setups[it - 1] = read_setup(rr)
    finally:
        rr.close()
    return {"types": types, "E_atom": E_atom, "normalized": normalized, "scale": scale, "shift": shift,
            "structures": strucs, "natoms_tot": natoms_tot, "E_av": E_av, "E_min": E_min, "E_max": E_max,
            "setups": setups}


def read_forces(path):
    """The descriptor-derivative file written beside a training set when
    FORCES is set (src/pytorchoutput.f90): nstrucs; per structure
    len_trim(name), name, (natoms, ntypes), selected; per atom of a selected
    structure itype, (nsf, nnb), nblist(nnb), sfderiv_i(nsf, 3) and
    sfderiv_j(nnb, nsf, 3) (column-major, returned as [nsf, 3] and
    [nnb, nsf, 3]); at the end the largest neighbor count."""
    rr = RecordReader(path)
    try:
        nstrucs = int(_ints(rr.next(), 1)[0])
        strucs = []
        for _ in range(nstrucs):
            ln = int(_ints(rr.next(), 1)[0])
            name = rr.next().decode("latin-1")[:ln]
            nat, nty = (int(v) for v in _ints(rr.next(), 2))
            sel = int(_ints(rr.next(), 1)[0])
            atoms = []
            if sel:
                for _ in range(nat):
                    it = int(_ints(rr.next(), 1)[0])
                    nsf, nnb = (int(v) for v in _ints(rr.next(), 2))
                    nbl = _ints(rr.next(), nnb)
                    di = _dbls(rr.next(), 3 * nsf).reshape(3, nsf).T
                    dj = _dbls(rr.next(), 3 * nsf * nnb).reshape(3, nsf, nnb).transpose(2, 1, 0)
                    atoms.append({"itype": it, "nsf": nsf, "nnb": nnb, "nblist": nbl, "deriv_i": di, "deriv_j": dj})
            strucs.append({"name": name, "natoms": nat, "ntypes": nty, "selected": sel, "atoms": atoms})
        max_nnb = int(_ints(rr.next(), 1)[0])
    finally:
        rr.close()
    return {"nstrucs": nstrucs, "structures": strucs, "max_nnb": max_nnb}
