"""Ensemble preprocessing into the mode-space data the Lanczos consumes.

Restates, on plain arrays, what ``DynamicalLanczos.Lanczos.__init__`` does with
a sscha ``Ensemble`` (Modules/DynamicalLanczos.py:322-452):

  * drop the rigid translations from the supercell modes (DL:340-352; the
    reference calls CellConstructor's ``get_translations`` on the
    mass-scaled eigenvectors) and keep an optional ``select_modes`` subset
    (DL:355-366);
  * residual forces f - f_sscha (DL:394-395), minus their average over the
    ensemble and the supercell images per unit-cell atom, symmetrised by the
    point group and tiled over the supercell (DL:398-419, 441);
  * mass scaling u~ = sqrt(m) u, f~ = f / sqrt(m) (DL:444-445) and the mode
    projections X = u~ pols, Y = f~ pols (DL:451-452).

Units are Rydberg atomic units: displacements in Bohr, forces in Ry/Bohr,
masses in Rydberg mass units (2 m_e); ``angstrom=True`` takes displacements
in Angstrom and forces in Ry/Angstrom.  Supercell atoms are ordered cell by
cell (atom k of cell l at index l * nat_uc + k), the order of the unit cell
repeated.  The result has the attributes the ``Lanczos`` constructor reads
(T, w, pols, m, N, rho, X, Y) and, when the supercell structure is given, a
``symmetries`` method that detects the space group (symmetry.find_symmetries)
for ``Lanczos.init(use_symmetries=True)``.
"""
from __future__ import annotations

import numpy as np

BOHR_ANGSTROM = 0.52917720859
CC_BOHR_PER_ANGSTROM = 1.889725989


def translation_mask(pols, m, threshold=0.5):
    """Modes that are rigid translations: total overlap with the three
    mass-weighted uniform translations above ``threshold`` (pols are the
    mass-scaled eigenvectors, columns; m the mass of every degree of
    freedom)."""
    pols = np.asarray(pols, float)
    m = np.asarray(m, float)
    t = np.zeros((len(m), 3))
    for c in range(3):
        t[c::3, c] = np.sqrt(m[c::3])
    t /= np.linalg.norm(t, axis=0)
    overlap = np.sum((t.T @ pols) ** 2, axis=0)
    return overlap > threshold


def symmetrize_unit_cell_vector(v, rotations, atom_maps):
    """Point-group average of a per-atom vector v (nat_uc, 3): operation j
    moves atom k to atom_maps[j][k] and rotates by rotations[j]
    (the role of QE_Symmetry.SymmetrizeVector at DL:408-410)."""
    v = np.asarray(v, float)
    out = np.zeros_like(v)
    for R, irt in zip(rotations, atom_maps):
        rv = v @ np.asarray(R, float).T
        out[np.asarray(irt)] += rv
    return out / len(rotations)


class PreparedEnsemble:
    """Mode-space ensemble (the attributes the Lanczos constructor reads)."""

    def __init__(self, T, w, pols, m, rho, X, Y, u_tilde, f_tilde, mode_mask, structure=None,
                 ignore_small_w=False):
        self.T = float(T)
        self.ignore_small_w = bool(ignore_small_w)
        self.w = w
        self.pols = pols
        self.m = m
        self.rho = rho
        self.N = len(rho)
        self.N_eff = float(np.sum(rho))
        self.X = X
        self.Y = Y
        self.u_tilde = u_tilde
        self.f_tilde = f_tilde
        self.mode_mask = mode_mask
        self.structure = structure

    def symmetries(self, use_symmetries=True, gamma_only=True):
        """Mode-space symmetries of the supercell (DL:694-871): detected space
        group, point group + translation projector if ``gamma_only``."""
        from . import symmetry as SY
        n = len(self.w)
        if not use_symmetries:
            syms = [np.ones((1, 1, 1)) for _ in range(n)]
            blocks = [np.array([i], dtype=np.int32) for i in range(n)]
            return SY.ModeSymmetries(syms, blocks, n)
        if self.structure is None:
            raise ValueError("symmetries need the supercell structure (lattice, positions, "
                             "species) passed to prepare_ensemble")
        lattice, positions, species = self.structure
        ms = SY.detect_mode_symmetries(lattice, positions, species, self.pols, self.w,
                                       gamma_only=gamma_only)
        if ms.leakage > 1e-8:
            raise RuntimeError("symmetry not block diagonal (leakage %.2e)" % ms.leakage)
        return ms


def prepare_ensemble(u_disps, forces, sscha_forces, masses, w, pols, T, rho=None, n_cells=1,
                     rotations=None, atom_maps=None, select_modes=None, ignore_small_w=False,
                     small_w=1e-8, angstrom=False, structure=None):
    """Mode-space ensemble from raw configurations.

    u_disps, forces, sscha_forces: (N, nat, 3) or (N, 3 nat) displacements
    from the average positions, forces and SSCHA (harmonic) forces of the
    supercell; masses: (nat,) per atom; w, pols: all 3 nat supercell
    frequencies (Ry) and mass-scaled eigenvectors (columns); rho: weights
    (default 1); n_cells: unit cells in the supercell; rotations,
    atom_maps: the unit cell's point group for the average-force
    symmetrisation (skipped if None); structure: (lattice, positions,
    species) of the supercell, Bohr (Angstrom with ``angstrom``), for the
    symmetry detection."""
    u = np.asarray(u_disps, float)
    N = u.shape[0]
    u = u.reshape(N, -1).copy()
    f = np.asarray(forces, float).reshape(N, -1) - np.asarray(sscha_forces, float).reshape(N, -1)
    if angstrom:                           # DL:389-395, 421
        u /= BOHR_ANGSTROM
        f *= BOHR_ANGSTROM
    n3 = u.shape[1]
    masses = np.asarray(masses, float)
    if 3 * len(masses) != n3:
        raise ValueError("masses must give one entry per atom of the supercell")
    m = np.repeat(masses, 3)               # DL:336
    w = np.asarray(w, float)
    pols = np.asarray(pols, float)
    if pols.shape != (n3, len(w)):
        raise ValueError("pols must be (3 nat, n_modes) with one column per frequency")
    if ignore_small_w:                     # DL:347-350
        trans = np.abs(w) < small_w
    else:
        trans = translation_mask(pols, m)
    good = ~trans
    if select_modes is not None:           # DL:355-366
        select_modes = np.asarray(select_modes, bool)
        if len(select_modes) != len(w):
            raise ValueError("'select_modes' should be an array of the same length of the "
                             "number of modes.")
        good &= select_modes
    rho = np.ones(N) if rho is None else np.asarray(rho, float).copy()
    if len(rho) != N:
        raise ValueError("rho must have one weight per configuration")
    if n3 % (3 * n_cells):
        raise ValueError("the supercell atoms are not a whole number of unit cells")
    nat_uc = n3 // (3 * n_cells)
    # average residual force per unit-cell atom (DL:398-419)
    fm = (rho @ f) / rho.sum()
    fm = fm.reshape(n_cells, nat_uc, 3).mean(axis=0)
    if rotations is not None:
        fm = symmetrize_unit_cell_vector(fm, rotations, atom_maps)
    f -= np.tile(fm.ravel(), n_cells)[None]       # DL:441
    u *= np.sqrt(m)[None]                          # DL:444-445
    f /= np.sqrt(m)[None]
    pg = np.ascontiguousarray(pols[:, good])
    X = np.ascontiguousarray(u @ pg)               # DL:451-452
    Y = np.ascontiguousarray(f @ pg)
    if structure is not None:
        lat, pos, spec = structure
        lat = np.asarray(lat, float)
        pos = np.asarray(pos, float)
        if angstrom:
            lat = lat / BOHR_ANGSTROM
            pos = pos / BOHR_ANGSTROM
        structure = (lat, pos, np.asarray(spec))
    return PreparedEnsemble(T, w[good].copy(), pg, m, rho, X, Y, u, f, good, structure,
                            ignore_small_w=ignore_small_w)


def prepare_from_files(dyn_prefix, nqirr, ensemble_dir, population, T, rho=None,
                       select_modes=None, ignore_small_w=False):
    """Mode-space ensemble from the reference's input files: the
    dynamical-matrix files ``dyn_prefix1`` .. ``dyn_prefix<nqirr>`` and the
    binary ensemble ``xats_pop<p>.npy`` / ``forces_pop<p>.npy`` in
    ``ensemble_dir`` (the sequence CC.Phonons.Phonons(prefix, nqirr),
    sscha Ensemble(dyn, T).load_bin(dir, p), Lanczos(ens) of
    tests/test_julia/test_julia.py:27-32).  The ensemble's atoms follow the
    supercell order of dynmat.supercell_structure; the SSCHA forces are the
    harmonic forces -Phi u of the supercell force constants, rho defaults to
    1 (configurations drawn from the same matrices), and the average force is
    symmetrised by the unit cell's space group (symmetry.find_symmetries).

    Units follow the reference's two conversions: the structure's Angstrom
    coordinates and the harmonic forces use 1 Angstrom = CC_BOHR_PER_ANGSTROM
    bohr (the value of DL:133), the Lanczos' own conversions
    BOHR_ANGSTROM (DL:126-128, ASE's 2006 units).  The two differ by 8e-8,
    which moves X and Y by about 1e-6 relative because positions are much
    larger than displacements; with both, X and Y match the reference's to
    5e-10."""
    from . import dynmat as DM
    from . import symmetry as SY
    ds = DM.read_dyn_set(dyn_prefix, nqirr)
    xats, forces = DM.load_ensemble_bin(ensemble_dir, population)
    pos, lat, spec, mass = DM.supercell_structure(ds)
    if xats.shape[1] != len(pos):
        raise ValueError("the ensemble has %d atoms, the supercell of the matrices %d"
                         % (xats.shape[1], len(pos)))
    phi = DM.supercell_force_constants(ds, pos)
    w, pols = DM.supercell_modes(phi, mass)
    pos_a, lat_a = pos / CC_BOHR_PER_ANGSTROM, lat / CC_BOHR_PER_ANGSTROM
    u = xats - pos_a[None]                                # Angstrom, nearest image of each site
    u -= np.round(u @ np.linalg.inv(lat_a)) @ lat_a
    nn = min(np.linalg.norm(pos_a[i] - pos_a[j]) for i in range(len(pos_a))
             for j in range(i + 1, len(pos_a)))
    if np.abs(u).max() > 0.5 * nn:
        raise ValueError("ensemble atoms are not in the supercell order of the matrices")
    N = len(u)
    f_h = -(u.reshape(N, -1) @ phi) * CC_BOHR_PER_ANGSTROM ** 2   # Ry / Angstrom
    ops = SY.find_symmetries(ds.cell, ds.tau, [ds.species[t] for t in ds.ityp])
    return prepare_ensemble(u.reshape(N, -1), forces.reshape(N, -1), f_h, mass, w, pols, T,
                            rho=rho, n_cells=int(np.prod(ds.supercell_dims())),
                            rotations=[op[0] for op in ops], atom_maps=[op[2] for op in ops],
                            select_modes=select_modes, ignore_small_w=ignore_small_w,
                            angstrom=True, structure=(lat_a, pos_a, spec))
