"""Bloch (q-space) description of the rock-salt spring model and its ensemble.

The q-space Lanczos of the reference (Modules/QSpaceLanczos.py) works in the
basis of Bloch modes e_{q,nu}(l, alpha) = exp(2 pi i m_q . n_l / L) / sqrt(N_c)
* eps_{q,nu}(alpha), where n_l are the integer coordinates of cell l,
m_q the integer coordinates of q on the L x L x L grid, alpha the 6 mass-scaled
Cartesian degrees of freedom of the primitive cell and eps_{q,nu} the
eigenvectors of the dynamical matrix D(q) = sum_l D[(0, a), (l, b)]
exp(2 pi i m_q . n_l / L).  Gauge: eps_{-q} = conj(eps_q), and real
eigenvectors where q = -q, so that the Bloch components of a real
displacement satisfy X_{-q} = conj(X_q).

Ensemble data in this basis (QSpaceLanczos._bloch_transform_ensemble,
Modules/QSpaceLanczos.py:220-295): X_q[q, i, nu] = <e_{q,nu}, u~_i>,
Y_q[q, i, nu] = <e_{q,nu}, f~_i> with u~ = sqrt(m) u and f~ = f_res / sqrt(m)
the mass-scaled displacement and residual force of configuration i.
Frequencies w_q[nu, q] and the mask valid_modes_q[nu, q] (the three acoustic
bands at Gamma are invalid) use the reference's (n_bands, n_q) layout.
"""
from __future__ import annotations

import numpy as np


def q_grid(L):
    """Integer coordinates m (n_q, 3) of the q-points of an L^3 grid, in the
    cell order of the supercell (index (m0 L + m1) L + m2; Gamma first)."""
    return np.array([(i, j, k) for i in range(L) for j in range(L) for k in range(L)],
                    dtype=np.int64)


def q_index(m, L):
    m = np.mod(np.asarray(m), L)
    return int((m[0] * L + m[1]) * L + m[2])


def minus_q(L):
    """Index of -q for every q."""
    g = q_grid(L)
    return np.array([q_index(-m, L) for m in g], dtype=np.int64)


def pair_map(L, iq_pert):
    """q_pair_map[iq1] = iq2 with q1 + q2 = q_pert + G
    (QSpaceLanczos.build_q_pair_map, Modules/QSpaceLanczos.py:297-331) and
    the unique pairs iq1 <= iq2."""
    g = q_grid(L)
    mp = g[iq_pert]
    pm = np.array([q_index(mp - m, L) for m in g], dtype=np.int64)
    pairs = [(i, int(pm[i])) for i in range(len(g)) if i <= pm[i]]
    return pm, pairs


class BlochModes:
    """Bloch modes of a RockSalt model (supercell L x L x L)."""

    def __init__(self, model, acoustic_eps=1e-8):
        self.model = model
        L = model.L
        self.L = L
        self.n_cells = L ** 3
        self.n_bands = 6
        self.grid = q_grid(L)
        self.n_q = len(self.grid)
        # integer coordinates of every cell, in the supercell's cell order
        self.cells = q_grid(L)
        m = model.m
        Phi = model.force_constants()
        sm = 1.0 / np.sqrt(m)
        D = Phi * sm[:, None] * sm[None, :]
        nb = self.n_bands
        # D[(0, a), (l, b)] blocks
        D0 = D[:nb].reshape(nb, self.n_cells, nb)
        phase = np.exp(2j * np.pi * (self.grid @ self.cells.T) / L)       # (n_q, n_cells)
        mq = minus_q(L)
        w = np.zeros((nb, self.n_q))
        eps = np.zeros((self.n_q, nb, nb), np.complex128)               # eps[q][:, nu]
        done = np.zeros(self.n_q, bool)
        for iq in range(self.n_q):
            if done[iq]:
                continue
            Dq = np.einsum("alb,l->ab", D0, phase[iq])
            Dq = 0.5 * (Dq + Dq.conj().T)
            if mq[iq] == iq:
                w2, e = np.linalg.eigh(Dq.real)
                e = e.astype(np.complex128)
            else:
                w2, e = np.linalg.eigh(Dq)
            w[:, iq] = np.sign(w2) * np.sqrt(np.abs(w2))
            eps[iq] = e
            done[iq] = True
            if mq[iq] != iq:
                w[:, mq[iq]] = w[:, iq]
                eps[mq[iq]] = e.conj()
                done[mq[iq]] = True
        self.w_q = w
        self.eps = eps
        self.valid_modes_q = np.ones((nb, self.n_q), dtype=bool)
        self.valid_modes_q[:, 0] = np.abs(w[:, 0]) > acoustic_eps * max(1.0, np.abs(w).max())
        if self.valid_modes_q[:, 0].sum() != nb - 3:
            raise RuntimeError("expected three acoustic bands at Gamma")
        if np.any(w[self.valid_modes_q] <= 0):
            raise RuntimeError("unstable Bloch modes")
        # e_{q,nu}(l, a) as a matrix (3 nat, n_q * nb), mass-scaled coordinates
        self.phase = phase

    def basis(self):
        """E (3 nat, n_q * nb): column (q, nu) = e_{q,nu} in supercell order."""
        nb, nc = self.n_bands, self.n_cells
        E = np.einsum("ql,qab->laqb", self.phase, self.eps) / np.sqrt(nc)
        return E.reshape(nc * nb, self.n_q * nb)

    def transform(self, Z):
        """Bloch components (n_q, N, nb) of mass-scaled Cartesian rows Z (N, 3 nat)."""
        nb, nc = self.n_bands, self.n_cells
        Zc = Z.reshape(len(Z), nc, nb)
        Zq = np.einsum("ql,ila->qia", self.phase.conj(), Zc) / np.sqrt(nc)
        return np.ascontiguousarray(np.einsum("qia,qab->qib", Zq, self.eps.conj()))


def bloch_ensemble(ens, modes=None):
    """(modes, X_q, Y_q) of a rock-salt Ensemble: the mode-space rows of the
    ensemble mapped back to mass-scaled Cartesian (acoustic part excluded, as
    in the mode-space data) and Bloch transformed."""
    if modes is None:
        modes = BlochModes(ens.model)
    ut = ens.X @ ens.pols.T
    ft = ens.Y @ ens.pols.T
    return modes, modes.transform(ut), modes.transform(ft)


def q_symmetries(modes, rotations=None):
    """Point-group images in the Bloch basis (the q-space symmetry matrices of
    QSpaceLanczos._build_qspace_symmetries, Modules/QSpaceLanczos.py:1550-1633).

    For a rotation R about the cation at the origin (t = 0), an atom k of the
    primitive cell at tau_k goes to R tau_k = tau_{irt[k]} + M L_k (M the
    primitive lattice vectors as columns, L_k integer), and cell n to
    R_f n + L_k with R_f = M^-1 R M.  The Bloch components of the rotated
    mass-scaled displacement are then, for every source q-point p,

        x'[q] = D[p] x[p],  m_q = R_f^-T m_p (mod L),
        D[p]_{nu mu} = sum_k eps_q(irt[k])^H_nu R exp(-2 pi i m_q . L_k / L) eps_p(k)_mu

    with eps the Bloch eigenvectors of ``modes``.  Returns ``perm``
    (n_sym, n_q) int32, perm[j, p] = q, and ``D`` (n_sym, n_q, nb, nb)
    complex.  ``rotations`` defaults to the 48 elements of O_h (identity
    first)."""
    from .symmetry import cubic_rotations
    model = modes.model
    L, nq, nb = modes.L, modes.n_q, modes.n_bands
    if rotations is None:
        rotations = cubic_rotations()
    M = model.cell.T
    Minv = np.linalg.inv(M)
    tau = model.positions[:2]
    grid = modes.grid
    perms, Ds = [], []
    for R in rotations:
        Rf = Minv @ R @ M
        Rfi = np.round(Rf)
        if np.abs(Rf - Rfi).max() > 1e-8:
            raise ValueError("rotation is not a symmetry of the lattice")
        RfTinv = np.round(np.linalg.inv(Rfi).T).astype(np.int64)
        irt = np.zeros(2, np.int64)
        Lk = np.zeros((2, 3), np.int64)
        for k in range(2):
            r = R @ tau[k]
            for kp in range(2):
                d = Minv @ (r - tau[kp])
                if np.abs(d - np.round(d)).max() < 1e-8 and model.species[kp] == model.species[k]:
                    irt[k] = kp
                    Lk[k] = np.round(d).astype(np.int64)
                    break
            else:
                raise ValueError("rotation does not map the primitive cell onto itself")
        perm = np.zeros(nq, np.int32)
        D = np.zeros((nq, nb, nb), np.complex128)
        for p in range(nq):
            mq = np.mod(RfTinv @ grid[p], L)
            q = q_index(mq, L)
            perm[p] = q
            P = np.zeros((nb, nb), np.complex128)
            for k in range(2):
                ph = np.exp(-2j * np.pi * float(mq @ Lk[k]) / L)
                P[3 * irt[k]:3 * irt[k] + 3, 3 * k:3 * k + 3] = ph * R
            D[p] = modes.eps[q].conj().T @ P @ modes.eps[p]
        if len(set(perm.tolist())) != nq:
            raise ValueError("q-point map is not a permutation")
        perms.append(perm)
        Ds.append(D)
    return np.array(perms, dtype=np.int32), np.array(Ds)
