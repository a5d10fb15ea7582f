"""Synthetic rock-salt spring-model ensembles (the benchmark workloads).

Structure: rock salt (Fm-3m), primitive FCC vectors a/2 (0,1,1), (1,0,1),
(1,1,0); cation at 0, anion at (a/2, 0, 0); an L x L x L supercell of the
primitive cell, atoms ordered cell-major (cation, anion per cell) as the
reference's supercells (Modules/DynamicalLanczos.py:1747 tiles effective
charges per cell in that order).

Potential: nearest-neighbour (cation-anion) central springs on the bond
stretch d, V(d) = k2 d^2/2 + k3 d^3/6 + k4 d^4/24, plus a harmonic
second-neighbour (same species) central spring k2nn.  Without that
stabiliser the nearest-neighbour model has 12 (16 atoms) / 48 (128 atoms)
zero modes: rigid sliding of atomic lines along a bond axis costs no energy.
With it exactly the 3 acoustic modes vanish (45 / 381 optical+acoustic
modes remain, as the workloads state).

Ensemble: the SCHA density matrix is the harmonic one of the model; each
mode amplitude is Gaussian with <q^2> = (1 + 2 n(w)) / (2 w) (hbar = 1,
Rydberg units), seeded NumPy PCG64.  Forces come from the full anharmonic
model.  Preprocessing to the mode basis follows
Modules/DynamicalLanczos.py:389-452: residual forces f - f_scha, minus the
symmetrised average force, mass scaling u sqrt(m), f / sqrt(m), projection
on the polarisation vectors.

Units: Rydberg atomic units (Ry, Bohr, mass unit = 2 m_e).
"""
from __future__ import annotations

import numpy as np

from . import symmetry as SY

RY_TO_K = 157887.32400374097
RYDBERG_EV = 13.605691932782346          # CODATA 2006 (as ase create_units("2006"))
BOHR_ANGSTROM = 0.52917720859
AMU_RY = 911.444243                       # 1 amu in Rydberg mass units (2 m_e)
RY_TO_CM = 109737.31568527                # 1 Ry in cm^-1

EV_A2 = 1.0 / RYDBERG_EV * BOHR_ANGSTROM ** 2   # eV/A^2 -> Ry/Bohr^2
EV_A3 = 1.0 / RYDBERG_EV * BOHR_ANGSTROM ** 3
EV_A4 = 1.0 / RYDBERG_EV * BOHR_ANGSTROM ** 4


def bose(w, T):
    if T > 0:
        return 1.0 / (np.exp(w * RY_TO_K / T) - 1.0)
    return np.zeros_like(w)


class RockSalt:
    """Rock-salt spring model on an L x L x L supercell of the primitive cell."""

    def __init__(self, L=2, a_angstrom=5.6, masses_amu=(22.99, 35.45), k2=1.0, k3=-2.0,
                 k4=6.0, k2nn=0.25):
        self.L = L
        a = a_angstrom / BOHR_ANGSTROM
        self.a = a
        self.cell = 0.5 * a * np.array([[0.0, 1.0, 1.0], [1.0, 0.0, 1.0], [1.0, 1.0, 0.0]])
        self.lattice = L * self.cell
        basis = np.array([[0.0, 0.0, 0.0], [0.5 * a, 0.0, 0.0]])
        pos, species = [], []
        self.translations = []
        for i in range(L):
            for j in range(L):
                for k in range(L):
                    t = i * self.cell[0] + j * self.cell[1] + k * self.cell[2]
                    self.translations.append(t)
                    for s in range(2):
                        pos.append(t + basis[s])
                        species.append(s)
        self.positions = np.array(pos)
        self.species = np.array(species)
        self.nat = len(pos)
        self.n_cells = L ** 3
        m = np.array(masses_amu)[self.species] * AMU_RY
        self.m_atoms = m
        self.m = np.repeat(m, 3)                       # per Cartesian dof (DL:334)
        self.k2, self.k3, self.k4 = k2 * EV_A2, k3 * EV_A3, k4 * EV_A4
        self.k2nn = k2nn * EV_A2
        nn = [np.array(v) * 0.5 * a for v in
              ([1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1])]
        nnn = []
        for v in ([1, 1, 0], [1, -1, 0], [-1, 1, 0], [-1, -1, 0], [1, 0, 1], [1, 0, -1],
                  [-1, 0, 1], [-1, 0, -1], [0, 1, 1], [0, 1, -1], [0, -1, 1], [0, -1, -1]):
            nnn.append(np.array(v) * 0.5 * a)
        self.bonds_nn = self._bonds(nn)
        self.bonds_nnn = self._bonds(nnn)

    def _find(self, r):
        inv = np.linalg.inv(self.lattice)
        d = (self.positions - r) @ inv
        d -= np.round(d)
        dist = np.linalg.norm(d @ self.lattice, axis=1)
        j = int(np.argmin(dist))
        assert dist[j] < 1e-8
        return j

    def _bonds(self, vecs):
        """Unique (i, j, d) with r_j = r_i + d (mod supercell); each bond once."""
        out = []
        for i in range(self.nat):
            for d in vecs:
                # keep one orientation of every bond vector
                key = tuple(np.sign(np.round(d, 8)))
                first_nonzero = next(x for x in key if x != 0)
                if first_nonzero < 0:
                    continue
                out.append((i, self._find(self.positions[i] + d), d))
        I = np.array([b[0] for b in out])
        J = np.array([b[1] for b in out])
        D = np.array([b[2] for b in out])
        return I, J, D

    # ------------------------------------------------------------------ model
    def force_constants(self):
        """Harmonic force constants Phi (3 nat x 3 nat), Ry/Bohr^2."""
        n3 = 3 * self.nat
        Phi = np.zeros((n3, n3))
        for (I, J, D), k in ((self.bonds_nn, self.k2), (self.bonds_nnn, self.k2nn)):
            for i, j, d in zip(I, J, D):
                e = d / np.linalg.norm(d)
                blk = k * np.outer(e, e)
                Phi[3 * i:3 * i + 3, 3 * i:3 * i + 3] += blk
                Phi[3 * j:3 * j + 3, 3 * j:3 * j + 3] += blk
                Phi[3 * i:3 * i + 3, 3 * j:3 * j + 3] -= blk
                Phi[3 * j:3 * j + 3, 3 * i:3 * i + 3] -= blk
        return Phi

    def forces(self, u):
        """Forces (Ry/Bohr) for displacements u (Ncfg, nat, 3) in Bohr."""
        F = np.zeros_like(u)
        for (I, J, D), harmonic_only in ((self.bonds_nn, False), (self.bonds_nnn, True)):
            r = D[None] + u[:, J] - u[:, I]
            rn = np.linalg.norm(r, axis=2)
            dl = rn - np.linalg.norm(D, axis=1)[None]
            if harmonic_only:
                dV = self.k2nn * dl
            else:
                dV = self.k2 * dl + self.k3 * dl ** 2 / 2 + self.k4 * dl ** 3 / 6
            g = (dV / rn)[..., None] * r
            for c in range(3):
                np.add.at(F[:, :, c].T, I, g[:, :, c].T)
                np.add.at(F[:, :, c].T, J, -g[:, :, c].T)
        return F

    def harmonic_modes(self):
        """(w_all, pols_all): frequencies (Ry, signed sqrt) and mass-scaled
        eigenvectors (3 nat x 3 nat) sorted ascending."""
        Phi = self.force_constants()
        sm = 1.0 / np.sqrt(self.m)
        Dm = Phi * sm[:, None] * sm[None, :]
        w2, e = np.linalg.eigh(0.5 * (Dm + Dm.T))
        w = np.sign(w2) * np.sqrt(np.abs(w2))
        return w, e

    def modes(self):
        """Non-acoustic (w, pols): drop the three mass-weighted translations."""
        w, e = self.harmonic_modes()
        t = np.zeros((3 * self.nat, 3))
        for c in range(3):
            t[c::3, c] = np.sqrt(self.m[c::3])
        t /= np.linalg.norm(t, axis=0)
        overlap = np.sum((t.T @ e) ** 2, axis=0)
        trans = overlap > 0.5
        if trans.sum() != 3:
            raise RuntimeError("expected 3 translational modes, found %d" % trans.sum())
        good = ~trans
        if np.any(w[good] <= 1e-10):
            raise RuntimeError("the model has %d extra zero modes" % int(np.sum(w[good] <= 1e-10)))
        return w[good], e[:, good]


class Ensemble:
    """Mode-basis ensemble of a RockSalt model (what the Lanczos consumes)."""

    def __init__(self, model: RockSalt, T, N, seed=0):
        self.model = model
        self.T = float(T)
        self.N = int(N)
        self.seed = seed
        w, pols = model.modes()
        self.w = w
        self.pols = pols
        self.m = model.m
        n = len(w)
        rng = np.random.Generator(np.random.PCG64(seed))
        sigma = np.sqrt((1.0 + 2.0 * bose(w, self.T)) / (2.0 * w))
        q = rng.standard_normal((self.N, n)) * sigma[None]
        u_tilde = q @ pols.T                                    # mass-scaled
        u = u_tilde / np.sqrt(self.m)[None]                    # Bohr
        nat = model.nat
        f = model.forces(u.reshape(self.N, nat, 3)).reshape(self.N, 3 * nat)
        Phi = model.force_constants()
        f_scha = -(u @ Phi.T)
        fr = f - f_scha                                         # DL:394-395
        # symmetrised average force per unit-cell atom (DL:398-409), tiled
        fm = fr.reshape(self.N, model.n_cells, 2, 3).mean(axis=(0, 1))
        rots = SY.cubic_rotations()
        fm_sym = np.mean([fm @ R.T for R in rots], axis=0)     # O_h has no invariant vector
        fr = fr - np.tile(fm_sym.ravel(), model.n_cells)[None]
        self.u_disps = u
        self.rho = np.ones(self.N)
        self.X = np.ascontiguousarray((u * np.sqrt(self.m)[None]) @ pols)      # DL:451
        self.Y = np.ascontiguousarray((fr / np.sqrt(self.m)[None]) @ pols)     # DL:452

    def symmetries(self, use_symmetries=True, gamma_only=True):
        m = self.model
        if not use_symmetries:
            n = len(self.w)
            syms = [np.ones((1, 1, 1)) for _ in range(n)]
            blocks = [np.array([i], dtype=np.int32) for i in range(n)]
            return SY.ModeSymmetries(syms, blocks, n)
        ms = SY.build_mode_symmetries(m.positions, m.lattice, m.positions[0], self.pols, self.w,
                                      m.translations, gamma_only=gamma_only)
        if ms.leakage > 1e-8:
            raise RuntimeError("symmetry not block diagonal (leakage %.2e)" % ms.leakage)
        return ms

    def ir_vector(self, pol_vec=(1.0, 0.0, 0.0)):
        """Rigid cation-vs-anion displacement: effective charges +I / -I
        (tests/test_julia/test_gamma_only.py:36-40), tiled per cell as
        prepare_ir does (DL:1744-1747)."""
        pv = np.asarray(pol_vec, float)
        z = np.array([pv, -pv]).ravel()
        return np.tile(z, self.model.n_cells)
