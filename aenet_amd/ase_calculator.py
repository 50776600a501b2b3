"""ASE-style calculator on the GPU path.

Drop-in for ``aenet.ase_calculator.ANNCalculator``
(python3/aenet/ase_calculator.py:26-123): same constructor
(``ANNCalculator(potentials)`` with a dict species -> .nn file), same
``implemented_properties`` ('energy', 'forces'), same results keys.  The
reference builds neighbor lists with ASE and calls the library per atom;
here whole structures go through the batched GPU path (cell-list neighbor
search on the device).  When ASE is installed the class derives from its
Calculator; otherwise it works on any object with ``get_positions()``,
``get_cell()``, ``get_pbc()`` and ``get_chemical_symbols()`` (see
:class:`Atoms` below).
"""
import numpy as np

from .core import ANNPotentials

try:  # pragma: no cover - ASE is optional
    from ase.calculators.calculator import Calculator as _Base
    from ase.calculators.calculator import all_changes
    HAS_ASE = True
except ImportError:
    HAS_ASE = False
    all_changes = ["positions", "numbers", "cell", "pbc"]

    class _Base:
        implemented_properties = []

        def __init__(self, **kwargs):
            self.results = {}
            self.atoms = None

        def get_property(self, name, atoms=None):
            if atoms is not None:
                self.calculate(atoms, [name], all_changes)
            return self.results[name]

        def get_potential_energy(self, atoms=None, force_consistent=False):
            return self.get_property("energy", atoms)

        def get_forces(self, atoms=None):
            return self.get_property("forces", atoms)

        def calculate(self, atoms=None, properties=("energy",), system_changes=all_changes):
            if atoms is not None:
                self.atoms = atoms


class Atoms:
    """Minimal stand-in for ase.Atoms (symbols, positions, cell, pbc)."""

    def __init__(self, symbols, positions, cell=None, pbc=False):
        self.symbols = list(symbols)
        self.positions = np.array(positions, dtype=np.float64).reshape(-1, 3)
        self.cell = np.zeros((3, 3)) if cell is None else np.array(cell, dtype=np.float64).reshape(3, 3)
        self.pbc = np.array([pbc] * 3 if np.ndim(pbc) == 0 else pbc, dtype=bool)
        self.calc = None

    def __len__(self):
        return len(self.symbols)

    def get_positions(self):
        return self.positions.copy()

    def get_cell(self):
        return self.cell.copy()

    def get_pbc(self):
        return self.pbc.copy()

    def get_chemical_symbols(self):
        return list(self.symbols)

    def set_calculator(self, calc):
        self.calc = calc

    def get_potential_energy(self):
        return self.calc.get_potential_energy(self)

    def get_forces(self):
        return self.calc.get_forces(self)


def periodic_cell_for(cell, pbc, positions, cutoff, margin=1.0):
    """A fully periodic cell equivalent to a partly periodic structure: the
    periodic lattice vectors are kept, and each non-periodic one becomes a
    vector perpendicular to the periodic ones, longer than the structure's
    extent along it plus the cutoff, so that no periodic image along it
    comes within the cutoff of any atom (every image displacement then has
    a component of at least that length perpendicular to the periodic
    vectors).  The atoms' environments, and so all energies and forces, are
    those of the partly periodic structure."""
    cell = np.asarray(cell, dtype=np.float64).reshape(3, 3)
    pbc = np.asarray(pbc, dtype=bool)
    per = [cell[k] for k in range(3) if pbc[k]]
    # an orthonormal basis of the directions perpendicular to the periodic vectors
    if len(per) == 2:
        n = np.cross(per[0], per[1])
        perp = [n / np.linalg.norm(n)]
    elif len(per) == 1:
        a = per[0] / np.linalg.norm(per[0])
        t = np.eye(3)[int(np.argmin(np.abs(a)))]
        e1 = np.cross(a, t)
        e1 /= np.linalg.norm(e1)
        perp = [e1, np.cross(a, e1)]
    else:
        raise ValueError("periodic_cell_for needs one or two periodic directions")
    out = cell.copy()
    x = np.asarray(positions, dtype=np.float64).reshape(-1, 3)
    j = 0
    for k in range(3):
        if not pbc[k]:
            e = perp[j]
            j += 1
            proj = x @ e
            out[k] = e * ((proj.max() - proj.min()) + cutoff + margin)
    return out


class ANNCalculator(_Base):

    implemented_properties = ["energy", "forces"]

    def __init__(self, potentials, **kwargs):
        _Base.__init__(self, **kwargs)
        self.ann = ANNPotentials(potentials)
        self.cutoff = self.ann.Rc_max
        self.cutoff2 = self.cutoff ** 2
        self.results = {}

    def release(self):
        if getattr(self, "ann", None) is not None:
            self.ann.release()
            self.ann = None

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass

    def calculate(self, atoms=None, properties=("energy",), system_changes=all_changes):
        _Base.calculate(self, atoms, properties, system_changes)
        atoms = self.atoms if atoms is None else atoms
        pbc = np.asarray(atoms.get_pbc(), dtype=bool)
        cell = np.asarray(atoms.get_cell(), dtype=np.float64).reshape(3, 3)
        if pbc.any() and not pbc.all():
            # partly periodic (slab, wire): the equivalent fully periodic
            # cell, as the reference's ASE neighbor list treats them
            # (python3/aenet/ase_calculator.py:49-56)
            cell = periodic_cell_for(cell, pbc, atoms.get_positions(), self.cutoff + 1.0e-3)
            pbc = np.ones(3, bool)
        want_forces = "forces" in properties
        r = self.ann.structure_energy_and_forces(cell, bool(pbc.all()), np.asarray(atoms.get_positions()),
                                                 atoms.get_chemical_symbols(), forces=want_forces)
        self.results["energy"] = float(r["E_tot"][0])
        self.results["energies"] = r["energies"]
        if want_forces:
            self.results["forces"] = r["forces"]
        self.atoms = atoms
