"""Dynamical-matrix files and the supercell modes built from them.

The reference reads its harmonic input through CellConstructor
(``CC.Phonons.Phonons(prefix, nqirr)``, tests/test_julia/test_julia.py:27,
Modules/DynamicalLanczos.py:322-328) and its ensembles through the sscha
package (``Ensemble.load_bin``, tests/test_julia/test_julia.py:29).  Neither
package is part of the reference tree or of this image; this module reads the
same files:

  * Quantum ESPRESSO dynamical-matrix files (``prefix1`` .. ``prefixN``, one
    per irreducible q-point, each holding the matrices of its star): the
    header (ntyp, nat, ibrav = 0, celldm(1) = alat in bohr, the cell in alat
    units, species with masses in Rydberg units, positions in alat units),
    then blocks "Dynamical Matrix in cartesian axes" with q in units of
    2 pi / alat and the 3x3 complex blocks (i, j) of the force-constant
    matrix in Ry / bohr^2, then "Diagonalizing the dynamical matrix" with the
    frequencies of the first q-point;
  * the binary ensemble layout: ``xats_pop<p>.npy`` (N, nat, 3) positions in
    Angstrom and ``forces_pop<p>.npy`` (N, nat, 3) forces in Ry / Angstrom.

From the q-points of the full grid the supercell force constants are
Phi(R_a + tau_i, R_b + tau_j) = Re (1/N_q) sum_q D_ij(q) exp(i q.(R_b - R_a)),
and the supercell modes diagonalise Phi_ij / sqrt(m_i m_j) (the role of
``Phonons.DiagonalizeSupercell`` at DL:328).  Supercell atoms are ordered
cell by cell, the last lattice index fastest, unit-cell atoms innermost: the
order of the reference's test ensembles.
"""
from __future__ import annotations

import os
import re
from fractions import Fraction
from math import gcd

import numpy as np

RY_TO_CM = 109737.31568527
BOHR_ANGSTROM = 0.52917720859


class DynFile:
    """Contents of one dynamical-matrix file (all lengths in bohr)."""

    def __init__(self, alat, cell, species, masses, ityp, tau, qpoints, dynmats, freqs_cm):
        self.alat = alat              # bohr
        self.cell = cell              # (3, 3) rows = cell vectors, bohr
        self.species = species        # names of the types
        self.masses = masses          # (ntyp,) Rydberg mass units
        self.ityp = ityp              # (nat,) type index of each atom
        self.tau = tau                # (nat, 3) positions, bohr
        self.qpoints = qpoints        # list of (3,) in 2 pi / alat
        self.dynmats = dynmats        # list of (3 nat, 3 nat) complex, Ry / bohr^2
        self.freqs_cm = freqs_cm      # frequencies printed for the first q-point (cm^-1)

    @property
    def nat(self):
        return len(self.ityp)

    def atom_masses(self):
        return np.asarray(self.masses)[self.ityp]


def _floats(line):
    return [float(x) for x in line.split()]


def read_dyn(path):
    """Reads one dynamical-matrix file (ibrav = 0, as CellConstructor writes)."""
    with open(path) as fp:
        lines = fp.read().splitlines()
    head = lines[2].split()
    ntyp, nat, ibrav = int(head[0]), int(head[1]), int(head[2])
    alat = float(head[3])
    if ibrav != 0:
        raise NotImplementedError("dynamical-matrix files with ibrav != 0 are not supported")
    k = 3
    if lines[k].strip().lower().startswith("basis"):
        k += 1
    cell = np.array([_floats(lines[k + i]) for i in range(3)]) * alat
    k += 3
    species, masses = [], []
    for i in range(ntyp):
        m = re.match(r"\s*(\d+)\s+'([^']*)'\s+(\S+)", lines[k + i])
        if not m:
            raise ValueError("%s: cannot read species line %r" % (path, lines[k + i]))
        species.append(m.group(2).strip())
        masses.append(float(m.group(3)))
    k += ntyp
    ityp, tau = [], []
    for i in range(nat):
        p = lines[k + i].split()
        ityp.append(int(p[1]) - 1)
        tau.append([float(x) for x in p[2:5]])
    k += nat
    tau = np.array(tau) * alat
    qpoints, dynmats, freqs = [], [], []
    qre = re.compile(r"q\s*=\s*\(\s*(\S+)\s+(\S+)\s+(\S+)\s*\)")
    while k < len(lines):
        ln = lines[k]
        if "Dynamical" in ln and "Matrix" in ln and "cartesian" in ln:
            k += 1
            while not qre.search(lines[k]):
                k += 1
            q = np.array([float(x) for x in qre.search(lines[k]).groups()])
            k += 1
            D = np.zeros((3 * nat, 3 * nat), complex)
            for _ in range(nat * nat):
                while not lines[k].strip():
                    k += 1
                i, j = (int(x) - 1 for x in lines[k].split()[:2])
                for r in range(3):
                    v = _floats(lines[k + 1 + r])
                    D[3 * i + r, 3 * j:3 * j + 3] = np.array(v[0::2]) + 1j * np.array(v[1::2])
                k += 4
            qpoints.append(q)
            dynmats.append(D)
            continue
        if "Diagonalizing" in ln:
            for ln2 in lines[k + 1:]:
                m = re.search(r"freq\s*\(\s*\d+\)\s*=\s*(\S+)\s*\[THz\]\s*=\s*(\S+)\s*\[cm-1\]", ln2)
                if m:
                    freqs.append(float(m.group(2)))
            break
        k += 1
    return DynFile(alat, cell, species, np.array(masses), np.array(ityp), tau, qpoints, dynmats,
                   np.array(freqs))


class DynSet:
    """The files of all irreducible q-points: the unit cell of the first, the
    q-points and matrices of every star."""

    def __init__(self, files):
        f0 = files[0]
        self.files = files
        self.alat, self.cell, self.species = f0.alat, f0.cell, f0.species
        self.masses, self.ityp, self.tau = f0.masses, f0.ityp, f0.tau
        self.qpoints = [q for f in files for q in f.qpoints]
        self.dynmats = [D for f in files for D in f.dynmats]

    @property
    def nat(self):
        return len(self.ityp)

    def atom_masses(self):
        return np.asarray(self.masses)[self.ityp]

    def q_fractional(self):
        """q-points in the reciprocal basis of the unit cell (q . a_i)."""
        return np.array([np.asarray(q) @ (self.cell / self.alat).T for q in self.qpoints])

    def supercell_dims(self):
        """Smallest supercell whose reciprocal lattice holds every q-point."""
        fr = self.q_fractional()
        dims = []
        for i in range(3):
            n = 1
            for x in fr[:, i]:
                d = Fraction(float(x)).limit_denominator(64).denominator
                n = n * d // gcd(n, d)
            dims.append(n)
        return tuple(dims)

    def check_grid(self):
        """Every point of the supercell's q grid exactly once (modulo G)."""
        dims = self.supercell_dims()
        fr = self.q_fractional() * np.array(dims)
        if not np.allclose(fr, np.round(fr), atol=1e-6):
            raise ValueError("q-points are not on the grid of supercell %s" % (dims,))
        keys = {tuple(int(v) % n for v, n in zip(np.round(f).astype(int), dims)) for f in fr}
        if len(keys) != len(fr) or len(fr) != int(np.prod(dims)):
            raise ValueError("the files hold %d distinct of the %d q-points of supercell %s"
                             % (len(keys), int(np.prod(dims)), dims))
        return dims


def read_dyn_set(prefix, nqirr):
    """``prefix1`` .. ``prefix<nqirr>`` (CC.Phonons.Phonons(prefix, nqirr))."""
    return DynSet([read_dyn("%s%d" % (prefix, i + 1)) for i in range(nqirr)])


def supercell_structure(dyn, dims=None):
    """Positions (bohr), lattice (bohr), species and masses of the supercell,
    cell by cell (last lattice index fastest), unit-cell atoms innermost."""
    dims = dyn.supercell_dims() if dims is None else tuple(dims)
    pos, spec, mass = [], [], []
    am = dyn.atom_masses()
    for i in range(dims[0]):
        for j in range(dims[1]):
            for k in range(dims[2]):
                R = i * dyn.cell[0] + j * dyn.cell[1] + k * dyn.cell[2]
                for a in range(dyn.nat):
                    pos.append(dyn.tau[a] + R)
                    spec.append(dyn.species[dyn.ityp[a]])
                    mass.append(am[a])
    lattice = np.array(dims, float)[:, None] * dyn.cell
    return np.array(pos), lattice, np.array(spec), np.array(mass)


def map_to_unit_cell(dyn, positions, tol=1e-4):
    """(unit-cell atom, lattice vector in bohr) of every supercell position."""
    Ai = np.linalg.inv(dyn.cell)
    atoms, cells = [], []
    for p in np.asarray(positions, float):
        best = None
        for a in range(dyn.nat):
            f = (p - dyn.tau[a]) @ Ai
            r = np.round(f)
            err = np.linalg.norm((f - r) @ dyn.cell)
            if best is None or err < best[0]:
                best = (err, a, r)
        if best[0] > tol * dyn.alat:
            raise ValueError("position %s is not a lattice site of the unit cell" % (p,))
        atoms.append(best[1])
        cells.append(best[2] @ dyn.cell)
    return np.array(atoms), np.array(cells)


def supercell_force_constants(dyn, positions):
    """Phi (3 nat_sc, 3 nat_sc) in Ry / bohr^2 for the supercell atoms at
    ``positions`` (bohr, any order), from the matrices of the full q grid."""
    dyn.check_grid()
    atoms, cells = map_to_unit_cell(dyn, positions)
    n = len(atoms)
    nq = len(dyn.qpoints)
    phi = np.zeros((3 * n, 3 * n), complex)
    two_pi = 2 * np.pi / dyn.alat
    for q, D in zip(dyn.qpoints, dyn.dynmats):
        ph = np.exp(1j * two_pi * (cells @ np.asarray(q)))           # e^{i q.R}
        for x in range(n):
            a = atoms[x]
            for y in range(n):
                b = atoms[y]
                phi[3 * x:3 * x + 3, 3 * y:3 * y + 3] += \
                    D[3 * a:3 * a + 3, 3 * b:3 * b + 3] * (ph[y] / ph[x])
    phi /= nq
    if np.abs(phi.imag).max() > 1e-8 * max(np.abs(phi.real).max(), 1e-300):
        raise ValueError("supercell force constants are not real: the q-points do not "
                         "close under inversion")
    phi = phi.real
    return 0.5 * (phi + phi.T)


def supercell_modes(phi, masses):
    """Frequencies (Ry, negative for unstable modes) and mass-scaled
    eigenvectors (columns) of the supercell, ascending."""
    m = np.repeat(np.asarray(masses, float), 3)
    D = phi / np.sqrt(np.outer(m, m))
    lam, V = np.linalg.eigh(0.5 * (D + D.T))
    w = np.sign(lam) * np.sqrt(np.abs(lam))
    return w, V


def q_frequencies(dyn_file, iq=0):
    """Frequencies (cm^-1) of the file's q-point iq from its matrix."""
    D = dyn_file.dynmats[iq]
    m = np.repeat(dyn_file.atom_masses(), 3)
    lam = np.linalg.eigvalsh(0.5 * (D + D.conj().T) / np.sqrt(np.outer(m, m)))
    return np.sign(lam) * np.sqrt(np.abs(lam)) * RY_TO_CM


def load_ensemble_bin(directory, population):
    """Positions (Angstrom) and forces (Ry / Angstrom) of an ensemble saved in
    the binary layout (``xats_pop<p>.npy``, ``forces_pop<p>.npy``)."""
    x = np.load(os.path.join(directory, "xats_pop%d.npy" % population), allow_pickle=False)
    f = np.load(os.path.join(directory, "forces_pop%d.npy" % population), allow_pickle=False)
    if x.shape != f.shape or x.ndim != 3 or x.shape[2] != 3:
        raise ValueError("positions and forces must both be (N, nat, 3)")
    return x, f


class SupercellModel:
    """The supercell of a DynSet in the form the Bloch-mode builder and the
    q-space symmetries read (qspace.BlochModes, qspace.q_symmetries): cubic
    L x L x L supercell, per-coordinate masses ``m``, force constants, the
    primitive cell, species, and positions taken about the first atom of the
    unit cell (the point group acts about that site)."""

    def __init__(self, dyn):
        dims = dyn.check_grid()
        if len(set(dims)) != 1:
            raise ValueError("q space needs an L x L x L supercell, the files give %s" % (dims,))
        if dyn.nat != 2:
            raise NotImplementedError("the Bloch-mode builder has six bands: two atoms per cell, "
                                      "the files have %d" % dyn.nat)
        self.L = dims[0]
        pos, lat, spec, mass = supercell_structure(dyn, dims)
        self.cell = np.asarray(dyn.cell, float)
        self.lattice = lat
        self.positions = pos - dyn.tau[0]
        self.species = spec
        self.nat = len(pos)
        self.m = np.repeat(mass, 3)
        self._phi = supercell_force_constants(dyn, pos)

    def force_constants(self):
        return self._phi
