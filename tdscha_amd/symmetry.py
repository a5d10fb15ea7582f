"""Space-group symmetries of rock-salt supercells, mapped into the mode basis.

Replaces what the reference obtains from spglib + CellConstructor
(Modules/DynamicalLanczos.py:694-871: ``GetSymmetriesFromSPGLIB``,
``GetSymmetriesOnModesDeg``, the gamma-only point-group/translation split at
DL:747-791).  Neither library is available here, so the operations are
generated directly for the Fm-3m rock-salt structure: the 48 proper and
improper cubic rotations about an atomic site (each is a signed permutation
matrix), times the lattice translations of the supercell.

Mode-space representation: for an operation g acting on the mass-scaled
Cartesian displacement field as P_g (atom i -> pi(i), vector rotated by R),
S_g = pols^T P_g pols.  S_g is block-diagonal over groups of degenerate
frequencies; each block is returned as f64[N_sym, d_b, d_b] exactly as the
reference's ``symmetries`` list (row = output mode within the block).
"""
from __future__ import annotations

import itertools

import numpy as np


def cubic_rotations():
    """The 48 elements of O_h as 3x3 signed permutation matrices."""
    ops = []
    for perm in itertools.permutations(range(3)):
        for signs in itertools.product((1, -1), repeat=3):
            R = np.zeros((3, 3))
            for row, col in enumerate(perm):
                R[row, col] = signs[row]
            ops.append(R)
    # identity first
    ops.sort(key=lambda R: (not np.allclose(R, np.eye(3)),))
    return ops


def _wrap_frac(fr):
    fr = fr - np.floor(fr + 1e-6)
    fr[np.abs(fr - 1.0) < 1e-6] = 0.0
    return fr


def atom_permutation(positions, lattice, R, center, t, tol=1e-4):
    """Atom map of r -> R (r - center) + center + t modulo the supercell.

    ``positions`` (nat, 3) Cartesian, ``lattice`` (3, 3) rows = supercell
    vectors.  Returns irt with image(atom i) = atom irt[i].
    """
    inv_lat = np.linalg.inv(lattice)
    frac = _wrap_frac(positions @ inv_lat)
    new = (positions - center) @ R.T + center + t
    new_frac = _wrap_frac(new @ inv_lat)
    irt = -np.ones(len(positions), dtype=np.int64)
    for i, nf in enumerate(new_frac):
        d = frac - nf
        d -= np.round(d)
        dist = np.linalg.norm(d @ lattice, axis=1)
        j = int(np.argmin(dist))
        if dist[j] > tol:
            raise ValueError("operation does not map the structure onto itself")
        irt[i] = j
    if len(set(irt.tolist())) != len(irt):
        raise ValueError("atom map is not a permutation")
    return irt


def cartesian_operator(irt, R):
    """P_g (3nat x 3nat): displacement of atom i, rotated, lands on atom irt[i]."""
    nat = len(irt)
    P = np.zeros((3 * nat, 3 * nat))
    for i, j in enumerate(irt):
        P[3 * j:3 * j + 3, 3 * i:3 * i + 3] = R
    return P


def degenerate_blocks(w, rtol=1e-6):
    """Group consecutive (sorted) frequencies that coincide within rtol."""
    blocks = []
    cur = [0]
    for k in range(1, len(w)):
        if abs(w[k] - w[k - 1]) <= rtol * max(abs(w[k]), abs(w[k - 1])):
            cur.append(k)
        else:
            blocks.append(cur)
            cur = [k]
    blocks.append(cur)
    return [np.array(b, dtype=np.int32) for b in blocks]


class ModeSymmetries:
    """Block-diagonal mode-space symmetries + translation projector.

    Attributes mirror the reference Lanczos object (DL:797-836):
    ``symmetries`` list of f64[N_sym, d_b, d_b], ``degenerate_space`` list
    of int32[d_b], ``N_degeneracy`` int32[n_blocks], ``sym_block_id``
    int32[n_modes], ``n_syms``; gamma-only extras ``trans_projector``
    (n, n), ``trans_operators``, ``trans_cart_perms`` (DL:761-791).
    """

    def __init__(self, symmetries, degenerate_space, n_modes,
                 trans_projector=None, trans_operators=None,
                 trans_cart_perms=None, leakage=0.0):
        self.symmetries = symmetries
        self.degenerate_space = degenerate_space
        self.N_degeneracy = np.array([len(b) for b in degenerate_space], dtype=np.int32)
        self.sym_block_id = -np.ones(n_modes, dtype=np.int32)
        for bid, b in enumerate(degenerate_space):
            self.sym_block_id[b] = bid
        assert np.all(self.sym_block_id >= 0)
        self.n_syms = symmetries[0].shape[0]
        self.trans_projector = trans_projector
        self.trans_operators = trans_operators
        self.trans_cart_perms = trans_cart_perms
        self.leakage = leakage

    def dense(self):
        """Dense stack S (N_sym, n, n)."""
        n = len(self.sym_block_id)
        S = np.zeros((self.n_syms, n, n))
        for sb, modes in zip(self.symmetries, self.degenerate_space):
            S[:, modes[:, None], modes[None, :]] = sb
        return S


def find_symmetries(lattice, positions, species, tol=1e-5):
    """Space-group operations of a periodic structure: the role spglib plays in
    the reference (``spglib.get_symmetry``, used by the symmetry setup at
    DL:694-871 and QL:1498-1515).

    ``lattice`` (3, 3) rows = cell vectors, ``positions`` (nat, 3) Cartesian,
    ``species`` (nat,) labels.  The rotations are the integer matrices M of
    the lattice basis (entries in {-1, 0, 1}, enough for a reduced cell)
    whose Cartesian form R = (A^-1 M A)^T is orthogonal; for each, every
    translation t (modulo the cell) that sends each atom onto an atom of the
    same species.  Returns a list of (R, t, irt), r -> R r + t, atom i ->
    irt[i], identity first.  For a supercell the pure lattice translations
    inside it appear as (I, t)."""
    A = np.asarray(lattice, float)
    Ai = np.linalg.inv(A)
    pos = np.asarray(positions, float)
    spec = np.asarray(species)
    frac = pos @ Ai
    scale = np.linalg.norm(A, axis=1).max()
    rots = []
    for e in itertools.product((-1, 0, 1), repeat=9):
        M = np.array(e, float).reshape(3, 3)
        if abs(abs(np.linalg.det(M)) - 1.0) > 1e-8:
            continue
        R = (Ai @ M @ A).T
        if np.allclose(R @ R.T, np.eye(3), atol=1e-8):
            rots.append(R)
    rots.sort(key=lambda R: (not np.allclose(R, np.eye(3)),))
    ops = []
    for R in rots:
        newf = (pos @ R.T) @ Ai                          # fractional images, before t
        for k in np.flatnonzero(spec == spec[0]):
            tf = frac[k] - newf[0]                         # candidate: atom 0 -> atom k
            img = newf + tf
            d = img[:, None, :] - frac[None, :, :]
            d -= np.round(d)
            dist = np.linalg.norm(d @ A, axis=2)
            irt = np.argmin(dist, axis=1)
            if (dist[np.arange(len(pos)), irt] > tol * scale).any():
                continue
            if (spec[irt] != spec).any() or len(set(irt.tolist())) != len(irt):
                continue
            tf = tf - np.floor(tf + 1e-9)
            ops.append((R, tf @ A, irt))
    return ops


def split_point_group(ops, lattice, tol=1e-8):
    """Pure translations T = {t : (I, t)} and one representative (R, t) per
    distinct rotation, t the shortest of its coset t + T (0 where the group
    is symmorphic about the origin)."""
    A = np.asarray(lattice, float)
    Ai = np.linalg.inv(A)
    trans = [t for R, t, _ in ops if np.allclose(R, np.eye(3), atol=tol)]
    reps = []
    for R, t, irt in ops:
        if any(np.allclose(R, Rr, atol=tol) for Rr, _, _ in reps):
            continue
        best = None
        for R2, t2, irt2 in ops:
            if not np.allclose(R2, R, atol=tol):
                continue
            f = t2 @ Ai
            f = f - np.round(f)
            n2 = np.linalg.norm(f @ A)
            if best is None or n2 < best[0] - tol:
                best = (n2, t2, irt2)
        reps.append((R, best[1], best[2]))
    return trans, reps


def build_mode_symmetries_from_ops(positions, lattice, pols, w, ops, translations=None,
                                   gamma_only=False, rtol=1e-6):
    """Mode-space symmetries from explicit operations (R, t) (r -> R r + t):
    ``ops`` are the images; with ``gamma_only`` the ``translations`` are
    returned as the projector (DL:747-791) instead of images."""
    blocks = degenerate_blocks(w, rtol)
    n = pols.shape[1]
    zero = np.zeros(3)
    mats = [[] for _ in blocks]
    leak = 0.0
    for R, t in ops:
        irt = atom_permutation(positions, lattice, R, zero, np.asarray(t, float))
        S = pols.T @ cartesian_operator(irt, R) @ pols
        mask = np.zeros((n, n), dtype=bool)
        for bi, b in enumerate(blocks):
            mats[bi].append(S[np.ix_(b, b)])
            mask[np.ix_(b, b)] = True
        leak = max(leak, float(np.max(np.abs(S[~mask]))) if (~mask).any() else 0.0)
    symmetries = [np.ascontiguousarray(np.array(m)) for m in mats]
    tp = tops = tperms = None
    if gamma_only:
        tops, tperms = [], []
        I3 = np.eye(3)
        for t in translations:
            irt = atom_permutation(positions, lattice, I3, zero, np.asarray(t, float))
            P = cartesian_operator(irt, I3)
            nat = len(irt)
            inv_perm = np.zeros(3 * nat, dtype=np.int64)
            for i, j in enumerate(irt):
                inv_perm[3 * j:3 * j + 3] = [3 * i, 3 * i + 1, 3 * i + 2]
            tperms.append(inv_perm)
            tops.append(pols.T @ P @ pols)
        tp = np.mean(tops, axis=0)
    return ModeSymmetries(symmetries, blocks, n, tp, tops, tperms, leak)


def detect_mode_symmetries(lattice, positions, species, pols, w, gamma_only=False, tol=1e-5,
                           rtol=1e-6):
    """find_symmetries + build_mode_symmetries_from_ops for a supercell:
    every space-group operation as an image, or (``gamma_only``) the point
    group representatives as images and the pure translations as the
    projector."""
    ops = find_symmetries(lattice, positions, species, tol)
    if gamma_only:
        trans, reps = split_point_group(ops, lattice)
        return build_mode_symmetries_from_ops(positions, lattice, pols, w,
                                              [(R, t) for R, t, _ in reps], trans,
                                              gamma_only=True, rtol=rtol)
    return build_mode_symmetries_from_ops(positions, lattice, pols, w,
                                          [(R, t) for R, t, _ in ops], rtol=rtol)


def build_mode_symmetries(positions, lattice, center, pols, w,
                          translations, gamma_only=False, rtol=1e-6):
    """Mode-space symmetries of a rock-salt supercell: the 48 rotations of
    O_h about ``center`` (r -> R (r - c) + c).

    ``translations``: list of Cartesian lattice translations of the primitive
    cell inside the supercell (including 0).  If ``gamma_only`` the images
    use only the 48 point-group operations and the translations are
    returned as a projector (DL:747-791); otherwise all 48 x N_cell
    operations are images.
    """
    c = np.asarray(center, float)
    if gamma_only:
        ops = [(R, c - R @ c) for R in cubic_rotations()]
    else:
        ops = [(R, c - R @ c + np.asarray(t, float)) for R in cubic_rotations()
               for t in translations]
    return build_mode_symmetries_from_ops(positions, lattice, pols, w, ops, translations,
                                          gamma_only=gamma_only, rtol=rtol)
