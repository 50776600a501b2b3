"""XSF structure files as ænet reads them (xsf_init,
src/ext/xsflib.f90:109-208): comment lines with ``total energy`` or
``cohesive energy`` carry the reference energies (``name = value``), a
``PRIMVEC`` block makes the structure periodic, atoms follow ``PRIMCOORD``
(after a line with their count) or ``ATOMS`` (isolated) up to the first
blank line, each ``symbol x y z [fx fy fz]``; forces are present when the
last atom line has at least 7 fields.  Species are numbered in their order
of first appearance."""
import numpy as np


def _readval(line, name):
    """io_readval (src/ext/io.f90:540-620): the first token after the first
    '=' that follows the first occurrence of name."""
    i = line.find(name)
    if i < 0:
        return None
    j = line.find("=", i + len(name))
    if j < 0:
        return None
    tok = line[j + 1:].split()
    return float(tok[0].replace("d", "e").replace("D", "e")) if tok else None


class XSF:
    def __init__(self):
        self.path = ""
        self.pbc = False
        self.cell = np.zeros((3, 3))       # rows = lattice vectors
        self.names = []                     # species in order of appearance
        self.types = None                   # 1-based into names
        self.positions = None               # (n, 3) Cartesian
        self.forces = None                  # (n, 3) or None
        self.E_coh = 0.0
        self.E_tot = 0.0
        self.has_energy = False

    @property
    def symbols(self):
        return [self.names[t - 1] for t in self.types]


def read_xsf(path):
    s = XSF()
    s.path = path
    with open(path) as f:
        lines = f.read().splitlines()
    i = 0
    atoms = []

    def read_atoms(k):
        rows = []
        while k < len(lines) and lines[k].strip():
            rows.append(lines[k].split())
            k += 1
        return rows, k

    while i < len(lines):
        line = lines[i]
        i += 1
        if not line.strip():
            continue
        if "cohesive energy" in line:
            v = _readval(line, "cohesive energy")
            if v is not None:
                s.E_coh = v
            s.has_energy = True
            continue
        if "total energy" in line:
            v = _readval(line, "total energy")
            if v is not None:
                s.E_tot = v
            s.has_energy = True
            continue
        t = line.strip()
        if t.startswith("#"):
            continue
        if "PRIMVEC" in t:
            s.pbc = True
            for k in range(3):
                s.cell[k] = [float(v) for v in lines[i + k].split()[:3]]
            i += 3
            continue
        if "PRIMCOORD" in t:
            i += 1      # the atom count line
            atoms, i = read_atoms(i)
            continue
        if "ATOMS" in t:
            atoms, i = read_atoms(i)
            continue
    if not atoms:
        raise ValueError("invalid input file: %s" % path)
    n = len(atoms)
    s.types = np.zeros(n, np.int32)
    s.positions = np.zeros((n, 3))
    has_f = len(atoms[-1]) >= 7
    if has_f:
        s.forces = np.zeros((n, 3))
    for k, row in enumerate(atoms):
        if row[0] not in s.names:
            s.names.append(row[0])
        s.types[k] = s.names.index(row[0]) + 1
        s.positions[k] = [float(v) for v in row[1:4]]
        if has_f:
            s.forces[k] = [float(v) for v in row[4:7]]
    return s


def write_xsf(path, symbols, positions, cell=None, forces=None, E_tot=None, E_coh=None):
    """A structure in the form read_xsf (and the reference) reads;
    cell None writes an isolated structure."""
    with open(path, "w") as f:
        if E_tot is not None:
            f.write("# total energy = %.17g eV\n" % E_tot)
        if E_coh is not None:
            f.write("# cohesive energy = %.17g eV\n" % E_coh)
        f.write("\n")
        if cell is not None:
            f.write("CRYSTAL\nPRIMVEC\n")
            for v in np.asarray(cell, dtype=np.float64):
                f.write(" %.17g %.17g %.17g\n" % tuple(v))
            f.write("PRIMCOORD\n %d 1\n" % len(symbols))
        else:
            f.write("ATOMS\n")
        for k, (el, x) in enumerate(zip(symbols, positions)):
            f.write("%-2s %.17g %.17g %.17g" % ((el,) + tuple(x)))
            if forces is not None:
                f.write(" %.17g %.17g %.17g" % tuple(forces[k]))
            f.write("\n")
