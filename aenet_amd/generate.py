"""generate.x on the GPU: a training set (``.train``) of descriptor values
from reference structures, optionally with the descriptor derivatives for
force training (``.train.forces``).

    python -m aenet_amd.generate generate.in

Input and output follow the reference program (src/generate.f90:180-440):
the generate.in keywords of read_InpGenerate (src/input.f90:269-315: TYPES
with atomic energies, SETUPS, OUTPUT, FORCES, FORCESPERCENT, FILES,
TIMING, DEBUG), the .stp set-ups of read_Setup_parameters
(src/sfsetup.f90:190-318), XSF structures with energies and forces
(src/ext/xsflib.f90:109-208), the .train records of src/trainset.f90 and
the derivative records of src/pytorchoutput.f90.  Unlike the reference,
which evaluates one atom after another, the descriptors of whole chunks of
structures come from one batched GPU pass (aenet_amd_batch_descriptors),
and the derivatives from one call per chunk
(aenet_amd_model_descriptor_derivatives); the neighbor lists are the
reference's own (bit-identical lcl_nbdist_cart order).
"""
import os
import struct
import sys
import tempfile
import time

import numpy as np

from . import nnfile, trainfile, xsf

PI_GEO = 3.141592653589793   # src/constants.f90:30


# ---------------------------------------------------------------------------
# input files
# ---------------------------------------------------------------------------

def _content_lines(path):
    """aeio_readline (src/aeio.f90:187-213): trimmed lines, skipping blank
    ones and those starting with '!', '#' or '%'."""
    out = []
    with open(path) as f:
        for line in f:
            t = line.strip()
            if not t or t[0] in "!#%":
                continue
            out.append(t)
    return out


def _find(lines, kw):
    for i, l in enumerate(lines):
        if l.split()[0].lower() == kw:
            return i
    return -1


def _real(s):
    return float(s.replace("d", "e").replace("D", "e"))


def _flag(lines, kw, default=False):
    i = _find(lines, kw)
    if i < 0:
        return default
    w = lines[i].split()
    if len(w) > 1:
        v = w[1].lower().lstrip(".")
        if v[:1] in ("t", "f"):
            return v[0] == "t"
    return True


def read_input(path):
    """read_InpGenerate (src/input.f90:269-315)."""
    L = _content_lines(path)
    # FORCESPERCENT defaults to 1 (per cent), as in inp_defaults (src/input.f90:252)
    inp = {"output": "refdata.train", "forces": False, "forcespercent": 1.0, "timing": False, "debug": False}
    k = _find(L, "types")
    if k < 0:
        raise ValueError("no atom TYPES specified in file `%s'" % path)
    nt = int(L[k + 1].split()[0])
    inp["types"], inp["E_atom"] = [], []
    for i in range(nt):
        w = L[k + 2 + i].split()
        inp["types"].append(w[0])
        inp["E_atom"].append(_real(w[1]))
    k = _find(L, "setups")
    if k < 0:
        raise ValueError("no basis function SETUPS specified in file `%s'" % path)
    inp["setups"] = [None] * nt
    for i in range(nt):
        w = L[k + 1 + i].split()
        inp["setups"][inp["types"].index(w[0])] = w[1]
    inp["forces"] = _flag(L, "forces")
    i = _find(L, "forcespercent")
    if i >= 0:
        inp["forcespercent"] = _real(L[i].split()[1])
    i = _find(L, "output")
    if i >= 0:
        inp["output"] = L[i].split()[1]
    inp["timing"] = _flag(L, "timing")
    inp["debug"] = _flag(L, "debug")
    # the structure list follows FILES (case as written, src/generate.f90:268-276)
    k = _find(L, "files")
    if k < 0:
        raise ValueError("no FILES in `%s'" % path)
    n = int(L[k + 1].split()[0])
    inp["files"] = [L[k + 2 + i] for i in range(n)]
    base = os.path.dirname(os.path.abspath(path))
    inp["base"] = base
    return inp


def _readval(line, name):
    """io_readval: the first token after the '=' that follows the first
    occurrence of name (case-sensitive, src/ext/io.f90:540-620)."""
    i = line.find(name)
    if i < 0:
        return None
    j = line.find("=", i + len(name))
    if j < 0:
        raise ValueError("invalid string in readval: %s" % line)
    return line[j + 1:].split()[0]


def read_stp(path, global_types):
    """read_Setup_parameters (src/sfsetup.f90:190-318) with the Chebyshev
    basis (read_basis_chebyshev, :1102-1158) and Behler2011 functions
    (read_symmfunc_Behler2011, :1222-1280).  Returns a set-up dict in the
    layout of aenet_amd.nnfile."""
    with open(path) as f:
        raw = f.read().splitlines()
    stp = {"description": "", "Rc_min": 0.0, "envtypes": [], "nsf": 0}
    i = 0

    def next_content():
        nonlocal i
        while i < len(raw):
            t = raw[i].strip()
            i += 1
            if t and t[0] not in "!#%":
                return t
        return None

    while True:
        line = next_content()
        if line is None:
            break
        kw = line.split()[0].lower()
        if kw == "atom":
            stp["atomtype"] = line[4:].split()[0]
        elif kw == "env":
            n = int(line[3:].split()[0])
            stp["envtypes"] = [next_content().split()[0] for _ in range(n)]
        elif kw == "descr":
            d = ""
            while i < len(raw) and "end descr" not in raw[i].lower():
                if len(d) < 1024:
                    d = d + raw[i].rstrip() + "$$"
                i += 1
            i += 1
            stp["description"] = d[:1024]
        elif kw == "rmin":
            stp["Rc_min"] = _real(line[4:].split()[0])
        elif kw in ("symmfunc", "functions"):
            if stp["nsf"] > 0:
                raise ValueError("multiple basis definitions in %s" % path)
            sftype = _readval(line, "type")
            if sftype != "Behler2011":
                raise ValueError("unknown basis function type: %s" % sftype)
            nsf = int(raw[i].split()[0])
            i += 1
            fns = []
            for _ in range(nsf):
                l = raw[i]
                i += 1
                g = int(_readval(l, "G"))
                fn = {"G": g, "type2": _readval(l, "type2"), "Rc": _real(_readval(l, "Rc"))}
                if g == 2:
                    fn.update(Rs=_real(_readval(l, "Rs") or "0"), eta=_real(_readval(l, "eta") or "0"))
                elif g == 3:
                    fn.update(kappa=_real(_readval(l, "kappa") or "0"))
                elif g in (4, 5):
                    fn.update(type3=_readval(l, "type3"), **{"lambda": _real(_readval(l, "lambda") or "0")},
                              zeta=_real(_readval(l, "zeta") or "0"), eta=_real(_readval(l, "eta") or "0"))
                elif g != 1:
                    raise ValueError("symmetry function type not implemented: %d" % g)
                fns.append(fn)
            st = nnfile.behler_setup(stp.get("atomtype", ""), stp["envtypes"], fns)
            stp.update({k: st[k] for k in ("sftype", "nsf", "sf", "sfparam", "sfenv", "Rc_max")})
        elif kw == "basis":
            if stp["nsf"] > 0:
                raise ValueError("multiple basis definitions in %s" % path)
            sftype = _readval(line, "type")
            if sftype.lower() != "chebyshev":
                raise ValueError("unknown basis type: %s" % sftype)
            l = raw[i]
            i += 1
            rRc = _real(_readval(l, "radial_Rc") or "0")
            rN = int(_readval(l, "radial_N") or "0")
            aRc = _real(_readval(l, "angular_Rc") or "0")
            aN = int(_readval(l, "angular_N") or "0")
            st = nnfile.chebyshev_setup(stp.get("atomtype", ""), stp["envtypes"], rRc, rN, aRc, aN)
            stp.update({k: st[k] for k in ("sftype", "nsf", "sf", "sfparam", "sfenv", "Rc_max")})
        else:
            print("Error: Unknown keyword : %s\n       file            : %s" % (kw, path), file=sys.stderr)
    if stp["Rc_min"] == 0.0:
        print("Warning: RMIN not set in : %s\n         --> setting Rc_min = 1.0" % path, file=sys.stderr)
        stp["Rc_min"] = 1.0
    if not stp.get("Rc_max"):
        raise ValueError("Cutoff undetermined in %s" % path)
    nsf = stp["nsf"]
    stp["neval"] = 0
    for k in ("sfval_min", "sfval_max", "sfval_avg", "sfval_cov"):
        stp[k] = np.zeros(nsf)
    return stp


# ---------------------------------------------------------------------------
# geometry as geo_init_xsf and lcl_init leave it (src/geometry.f90:661-736,
# src/ext/lclist.f90:953-987)
# ---------------------------------------------------------------------------

def _recip(L):
    """geo_recip_lattice (src/geometry.f90:491-514); L rows = vectors."""
    a, b, c = L
    B = np.array([np.cross(b, c), -np.cross(a, c), np.cross(a, b)])
    V = np.dot(a, np.cross(b, c))
    return B / V * 2.0 * PI_GEO


def written_coordinates(s):
    """cooCart(iatom) of the geometry module after lcl_init: L s with s the
    lattice coordinates, wrapped into [0,1) for periodic structures and
    relative to the bounding box (geo_get_bounds) for isolated ones; and
    the forces through forLatt and back."""
    if s.pbc:
        L = s.cell.copy()
        x = s.positions
    else:
        eps = 2.0e-6
        mn, mx = s.positions.min(axis=0), s.positions.max(axis=0)
        L = np.diag(mx - mn + eps)
        x = s.positions - (mn - 0.5 * eps)
    B = _recip(L)
    frac = (x @ B.T) / (2.0 * PI_GEO)
    if s.pbc:
        frac = frac - np.floor(frac)
        frac[frac >= 1.0] -= 1.0
    F = s.forces if s.forces is not None else np.zeros_like(s.positions)
    forlatt = (F @ B.T) / (2.0 * PI_GEO)
    return frac @ L, forlatt @ L


# ---------------------------------------------------------------------------
# driver
# ---------------------------------------------------------------------------

def descriptor_model(inp, setups, device):
    """A potential set whose networks are never evaluated: each species'
    set-up with a zero one-layer network, so the GPU descriptor kernels see
    the set-up exactly as load_Setup gives it."""
    from .model import Model
    d = tempfile.mkdtemp(prefix="aenet_generate_")
    files = {}
    for t, stp in zip(inp["types"], setups):
        pot = dict(stp)
        pot.update({"nnodes": [stp["nsf"], 1], "f_a": [0], "W": np.zeros(stp["nsf"] + 1),
                    "sfval_cov": np.ones(stp["nsf"]), "sfval_avg": np.zeros(stp["nsf"]),
                    "ts_types": list(inp["types"]), "ts_E_atom": list(inp["E_atom"])})
        files[t] = os.path.join(d, "%s.nn" % t)
        nnfile.write_nn(files[t], pot)
    m = Model(files, device=device)
    for f in files.values():
        os.remove(f)
    os.rmdir(d)
    return m


def run(input_path, device="cuda", chunk_atoms=200000, log=print):
    import torch
    from .device import DeviceBatch
    inp = read_input(input_path)
    base = inp["base"]

    def rel(p):
        return p if os.path.isabs(p) else os.path.join(base, p)

    types = inp["types"]
    setups = [read_stp(rel(p), types) for p in inp["setups"]]
    for t, stp in zip(types, setups):
        if stp["atomtype"].strip() != t:
            raise ValueError("inconsistent atom type in setup: expected %s, found %s" % (t, stp["atomtype"]))
    out_path = rel(inp["output"])
    if os.path.exists(out_path):
        raise FileExistsError("file already exists: %s" % out_path)
    model = descriptor_model(inp, setups, device)
    stride = model.nsf_max
    nfiles = len(inp["files"])
    tw = trainfile.TrainWriter(out_path, types, inp["E_atom"], nfiles)
    pyo = None
    selected = np.ones(nfiles, bool)
    if inp["forces"]:
        pyo = open(out_path + ".forces", "wb")
        pyo.write(nnfile._rec(struct.pack("<i", nfiles)))
        # the fraction FORCESPERCENT of the structures, chosen by a seeded
        # shuffle (the reference draws them with Fortran's random_number)
        nsel = int(round(nfiles * inp["forcespercent"] / 100.0))
        selected = np.zeros(nfiles, bool)
        selected[np.random.Generator(np.random.PCG64(0)).permutation(nfiles)[:nsel]] = True
    max_nnb = 0
    log("#%6s  %6s  %6s  %15s  %s" % ("N", "nAtoms", "nTypes", "E/atom", "structure file (xsf)"))
    t0 = time.perf_counter()
    k = 0
    while k < nfiles:
        # a chunk of structures: read, evaluate on the GPU, write in order
        chunk = []
        na = 0
        while k < nfiles and (not chunk or na < chunk_atoms):
            s = xsf.read_xsf(rel(inp["files"][k]))
            if s.forces is None or not s.has_energy:
                raise ValueError("incomplete output data in : %s" % inp["files"][k])
            gt = np.array([types.index(nm) + 1 if nm in types else 0 for nm in s.names], np.int32)
            if (gt == 0).any():
                raise ValueError("not a valid structure: %s (additional species found)" % inp["files"][k])
            s.gtypes = gt[s.types - 1]
            chunk.append((k, s))
            na += len(s.types)
            k += 1

        class _S:
            pass
        structs = []
        for _, s in chunk:
            o = _S()
            o.cell, o.pbc, o.positions, o.types = (s.cell if s.pbc else np.eye(3)), s.pbc, s.positions, s.gtypes
            structs.append(o)
        db = DeviceBatch(structs, device=device, model=model)
        G = db.descriptors(stride, scaled=False).cpu().numpy()
        nbl = None
        if pyo is not None and any(selected[i] for i, _ in chunk):
            nbl = db.neighbors(512)
            nbl = {key: v.cpu().numpy() for key, v in nbl.items()}
        del db
        a0 = 0
        for ifile, s in chunk:
            n = len(s.types)
            if s.E_coh != 0.0:
                E_coh = s.E_coh
            else:
                E_coh = s.E_tot - sum(inp["E_atom"][t - 1] for t in s.gtypes)
            log(" %6d  %6d  %6d  %15.8E  %s" % (ifile + 1, n, len(s.names), E_coh / n, inp["files"][ifile]))
            coo, frc = written_coordinates(s)
            sf = []
            for i in range(n):
                t = int(s.gtypes[i])
                stp = setups[t - 1]
                v = G[a0 + i, :stp["nsf"]]
                sf.append(v)
                # running statistics as stp_eval keeps them (src/sfsetup.f90:960-977)
                ne = stp["neval"]
                if ne == 0:
                    stp["sfval_min"] = v.copy()
                    stp["sfval_max"] = v.copy()
                    stp["sfval_avg"] = v.copy()
                    stp["sfval_cov"] = v * v
                else:
                    stp["sfval_min"] = np.minimum(stp["sfval_min"], v)
                    stp["sfval_max"] = np.maximum(stp["sfval_max"], v)
                    stp["sfval_avg"] = (ne * stp["sfval_avg"] + v) / (ne + 1)
                    stp["sfval_cov"] = (ne * stp["sfval_cov"] + v * v) / (ne + 1)
                stp["neval"] = ne + 1
            tw.structure(inp["files"][ifile], len(s.names), E_coh, s.gtypes, coo, frc, sf)
            if pyo is not None:
                name = inp["files"][ifile].encode("latin-1")
                pyo.write(nnfile._rec(struct.pack("<i", len(name))) + nnfile._rec(name))
                pyo.write(nnfile._rec(struct.pack("<ii", n, len(s.names))))
                pyo.write(nnfile._rec(struct.pack("<i", 1 if selected[ifile] else 0)))
                if selected[ifile]:
                    envs = []
                    for i in range(n):
                        c = int(nbl["nnb"][a0 + i])
                        envs.append((int(s.gtypes[i]), coo[i], nbl["coo"][a0 + i, :c],
                                     nbl["type"][a0 + i, :c]))
                    _, di, dj = model.descriptor_derivatives(envs, stride)
                    for i in range(n):
                        t = int(s.gtypes[i])
                        nsf = setups[t - 1]["nsf"]
                        c = int(nbl["nnb"][a0 + i])
                        max_nnb = max(max_nnb, c)
                        rec = nnfile._rec(struct.pack("<i", t))
                        rec += nnfile._rec(struct.pack("<ii", nsf, c))
                        rec += nnfile._rec(np.ascontiguousarray(nbl["list"][a0 + i, :c], "<i4").tobytes())
                        # sfderiv_i_aux(nsf, 3) and sfderiv_j_aux(nnb, nsf, 3), column-major
                        rec += nnfile._rec(np.ascontiguousarray(di[i, :nsf, :].T, "<f8").tobytes())
                        rec += nnfile._rec(np.ascontiguousarray(dj[i][:, :nsf, :].transpose(2, 1, 0), "<f8").tobytes())
                        pyo.write(rec)
            a0 += n
    tw.close(setups)
    if pyo is not None:
        pyo.write(nnfile._rec(struct.pack("<i", max_nnb)))
        pyo.close()
    log("")
    log(" Training set written to %s (%d structures, %d atoms) in %.2f s"
        % (out_path, nfiles, tw.natoms_tot, time.perf_counter() - t0))
    return out_path


def main():
    if len(sys.argv) < 2:
        print("generate -- training set generation on the GPU.\nUsage: python -m aenet_amd.generate <input-file>")
        sys.exit(1)
    run(sys.argv[1])


if __name__ == "__main__":
    main()
