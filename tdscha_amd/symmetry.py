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


def build_mode_symmetries(positions, lattice, center, pols, w,
                          translations, gamma_only=False, rtol=1e-6):
    """Mode-space symmetries of a rock-salt supercell.

    ``translations``: list of Cartesian lattice translations of the primitive
    cell inside the supercell (including 0).  If ``gamma_only`` the images
    use only the 48 point-group operations (t = 0) and the translations are
    returned as a projector (DL:747-791); otherwise all 48 x N_cell
    operations are images.
    """
    rots = cubic_rotations()
    blocks = degenerate_blocks(w, rtol)
    n = pols.shape[1]
    ops = []
    if gamma_only:
        for R in rots:
            ops.append((R, np.zeros(3)))
    else:
        for R in rots:
            for t in translations:
                ops.append((R, np.asarray(t, float)))
    mats = [[] for _ in blocks]
    leak = 0.0
    for R, t in ops:
        irt = atom_permutation(positions, lattice, R, center, t)
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
            irt = atom_permutation(positions, lattice, I3, center, np.asarray(t, float))
            P = cartesian_operator(irt, I3)
            nat = len(irt)
            inv_perm = np.zeros(3 * nat, dtype=np.int64)
            for i, j in enumerate(irt):
                inv_perm[3 * j:3 * j + 3] = [3 * i, 3 * i + 1, 3 * i + 2]
            tperms.append(inv_perm)
            tops.append(pols.T @ P @ pols)
        tp = np.mean(tops, axis=0)
    return ModeSymmetries(symmetries, blocks, n, tp, tops, tperms, leak)
