"""Free-energy Hessian in q space by linear solves with the static
Liouvillian (Modules/QSpaceHessian.py, class QSpaceHessian).

For each q-point and band i, L_static(q) x_i = e_i is solved (GMRES with the
harmonic preconditioner, in the basis made Hermitian by the square root of
the metric, QSH:720-956); G(q)[:, i] is the R sector of x_i and H(q) =
G(q)^-1 on the non-acoustic bands.  The static L runs in the gfx950 library
(``tdscha_qspace_apply_L_static``); the solver runs on the host as in the
reference.  ``compute_full_hessian`` solves one q-point per orbit of the
point group and rotates the result to the rest of the orbit,
H(R q) = D H(q) D^H, with the band-space matrices D of qspace.q_symmetries.
"""
from __future__ import annotations

import os

import numpy as np
import scipy.sparse.linalg

from .qlanczos import EPSILON, QSpaceLanczos
from .qspace import q_symmetries
from .symmetry import degenerate_blocks

RY_TO_K = 157887.32400374097


class QSpaceHessian:
    """QSH:52.  ``ensemble`` as for QSpaceLanczos, or an initialised
    QSpaceLanczos through from_qspace_lanczos."""

    def __init__(self, ensemble=None, verbose=True, ignore_v3=False, ignore_v4=False):
        self.verbose = verbose
        self.ignore_v3 = ignore_v3
        self.ignore_v4 = ignore_v4
        self.ignore_harmonic = False
        self.H_q_dict = {}
        self.qlanc = None
        self.irr_qpoints = None
        self._orbits = None
        if ensemble is not None:
            self._attach(QSpaceLanczos(ensemble))

    def _attach(self, qlanc):
        self.qlanc = qlanc
        self.n_q = qlanc.n_q
        self.n_bands = qlanc.n_bands
        self.w_q = qlanc.w_q
        self.T = qlanc.T

    @classmethod
    def from_qspace_lanczos(cls, qlanc, verbose=True, use_symmetries=True):
        """QSH:1071-1128: reuse the Bloch ensemble (and device context, when
        initialised) of a QSpaceLanczos."""
        h = cls(None, verbose=verbose, ignore_v3=qlanc.ignore_v3, ignore_v4=qlanc.ignore_v4)
        h._attach(qlanc)
        if not qlanc.initialized:
            qlanc.init(use_symmetries=use_symmetries)
        h._find_irreducible_qpoints(use_symmetries)
        return h

    def init(self, use_symmetries=True):
        """QSH:116-130: upload (and expand) the ensemble, find the
        irreducible q-points."""
        self.qlanc.init(use_symmetries=use_symmetries)
        self._find_irreducible_qpoints(use_symmetries)

    def _find_irreducible_qpoints(self, use_symmetries=True):
        """Orbits of the q-points under the point group (QSH:249-338, with the
        group of qspace.q_symmetries in place of spglib)."""
        if not use_symmetries:
            self._perm, self._D = None, None
            self.irr_qpoints = list(range(self.n_q))
            self._orbits = {iq: [(iq, None)] for iq in self.irr_qpoints}
            return
        perm, D = q_symmetries(self.qlanc.modes)
        self._perm, self._D = perm, D
        seen = {}
        self.irr_qpoints, self._orbits = [], {}
        for p in range(self.n_q):
            if p in seen:
                continue
            self.irr_qpoints.append(p)
            self._orbits[p] = []
            for j in range(len(perm)):
                q = int(perm[j][p])
                if q not in seen:
                    seen[q] = p
                    self._orbits[p].append((q, j))

    # ------------------------------------------------------------ layout
    def _layout(self, iq):
        ql = self.qlanc
        ql.build_q_pair_map(iq)
        nb = self.n_bands
        offs, sizes, off = [], [], nb
        for a, b in ql.unique_pairs:
            s = nb * nb if a < b else nb * (nb + 1) // 2
            offs.append(off)
            sizes.append(s)
            off += s
        self._offs, self._sizes, self._size = offs, sizes, off
        assert off == ql.engine.static_psi_size(ql.q_pair_map)

    def _get_W_block(self, psi, p):
        nb = self.n_bands
        a, b = self.qlanc.unique_pairs[p]
        raw = psi[self._offs[p]:self._offs[p] + self._sizes[p]]
        if a < b:
            return raw.reshape(nb, nb).copy()
        M = np.zeros((nb, nb), np.complex128)
        M[np.triu_indices(nb)] = raw
        il = np.tril_indices(nb, -1)
        M[il] = M.T.conj()[il]
        return M

    def _set_W_block(self, psi, p, M):
        nb = self.n_bands
        a, b = self.qlanc.unique_pairs[p]
        o = self._offs[p]
        if a < b:
            psi[o:o + nb * nb] = M.ravel()
        else:
            psi[o:o + self._sizes[p]] = M[np.triu_indices(nb)]

    def _static_mask(self):
        """QSH:434-479: metric of the static layout, masked modes at zero."""
        nb = self.n_bands
        valid = np.asarray(self.qlanc.valid_modes_q, bool)
        m = np.ones(self._size)
        m[:nb] = valid[:, self.qlanc.iq_pert]
        for p, (a, b) in enumerate(self.qlanc.unique_pairs):
            bad = ~valid[:, a][:, None] | ~valid[:, b][None, :]
            full = np.where(bad, 0.0, 2.0)
            o = self._offs[p]
            if a < b:
                m[o:o + nb * nb] = full.ravel()
            else:
                np.fill_diagonal(full, np.where(valid[:, a] & valid[:, b], 1.0, 0.0))
                m[o:o + self._sizes[p]] = full[np.triu_indices(nb)]
        return m

    # --------------------------------------------------- static quantities
    def _bose(self, iq):
        w = self.w_q[:, iq]
        v = np.asarray(self.qlanc.valid_modes_q, bool)[:, iq]
        n = np.zeros_like(w)
        if self.T > EPSILON:
            n[v] = 1.0 / (np.exp(w[v] * RY_TO_K / self.T) - 1.0)
        return n, v

    def _compute_lambda_q(self, iq1, iq2):
        """QSH:481-544."""
        w1, w2 = self.w_q[:, iq1], self.w_q[:, iq2]
        n1, v1 = self._bose(iq1)
        n2, v2 = self._bose(iq2)
        W1, W2 = np.meshgrid(w1, w2, indexing="ij")
        N1, N2 = np.meshgrid(n1, n2, indexing="ij")
        vm = np.outer(v1, v2)
        dn = np.zeros_like(W1)
        if self.T > EPSILON:
            beta = RY_TO_K / self.T
            eq = np.abs(W1 - W2) < 1e-8
            dn = np.where(eq, 0.0, (N1 - N2) / np.where(eq, 1.0, W1 - W2))
            dn = np.where(eq & vm, -beta * np.exp(np.where(eq & vm, W1, 0.0) * beta) * N1 ** 2,
                          dn)
        return np.where(vm, ((N1 + N2 + 1) / np.where(vm, W1 + W2, 1.0) - dn)
                        / (4 * np.where(vm, W1 * W2, 1.0)), 0.0)

    def _precompute_static_quantities(self):
        """QSH:564-603 (the parts the preconditioner needs; the operator
        itself computes them on the device)."""
        iq = self.qlanc.iq_pert
        w = self.w_q[:, iq]
        v = np.asarray(self.qlanc.valid_modes_q, bool)[:, iq]
        self._cached_inv_w_qp_sq = np.where(v, 1.0 / np.where(v, w, 1.0) ** 2, 0.0)
        self._cached_lambda = [self._compute_lambda_q(a, b) for a, b in self.qlanc.unique_pairs]

    def _apply_L_static_q(self, psi):
        """QSH:605-692 on the device."""
        ql = self.qlanc
        return ql.engine.apply_L_static([ql.iq_pert], [ql.q_pair_map], [psi],
                                        self.ignore_harmonic, self.ignore_v3, self.ignore_v4)[0]

    def _apply_harmonic_preconditioner(self, psi_tilde, sqrt_mask, inv_sqrt_mask):
        """QSH:694-718: the inverse of the harmonic static L."""
        nb = self.n_bands
        psi = psi_tilde * inv_sqrt_mask
        out = np.zeros_like(psi)
        out[:nb] = self._cached_inv_w_qp_sq * psi[:nb]
        for p in range(len(self.qlanc.unique_pairs)):
            self._set_W_block(out, p, self._cached_lambda[p] * self._get_W_block(psi, p))
        return out * sqrt_mask

    def _find_degenerate_blocks(self, iq):
        """QSH:340-377: non-acoustic bands grouped by equal frequency."""
        valid = np.asarray(self.qlanc.valid_modes_q, bool)[:, iq]
        blocks = []
        for blk in degenerate_blocks(self.w_q[:, iq]):
            b = [int(x) for x in blk if valid[x]]
            if b:
                blocks.append(b)
        return blocks

    # -------------------------------------------------------------- solver
    def compute_hessian_at_q(self, iq, tol=1e-6, max_iters=500, use_preconditioner=True,
                             dense_fallback=False, use_mode_symmetry=True):
        """QSH:720-956."""
        nb = self.n_bands
        self._layout(iq)
        n = self._size
        self._precompute_static_quantities()
        mask = self._static_mask()
        sqrt_mask = np.sqrt(mask)
        inv_sqrt_mask = np.zeros_like(sqrt_mask)
        nz = mask > 0
        inv_sqrt_mask[nz] = 1.0 / sqrt_mask[nz]
        counter = [0]

        def apply_L_tilde(x_tilde):
            counter[0] += 1
            return self._apply_L_static_q(x_tilde * inv_sqrt_mask) * sqrt_mask

        L_op = scipy.sparse.linalg.LinearOperator((n, n), matvec=apply_L_tilde,
                                                  dtype=np.complex128)
        M_op = None
        if use_preconditioner:
            def apply_M(x):
                return self._apply_harmonic_preconditioner(x, sqrt_mask, inv_sqrt_mask)
            M_op = scipy.sparse.linalg.LinearOperator((n, n), matvec=apply_M,
                                                      dtype=np.complex128)
        valid = np.asarray(self.qlanc.valid_modes_q, bool)[:, iq]
        non_acoustic = [nu for nu in range(nb) if valid[nu]]
        if use_mode_symmetry:
            schedule = [(blk[0], blk) for blk in self._find_degenerate_blocks(iq)]
        else:
            schedule = [(nu, [nu]) for nu in non_acoustic]
        G = np.zeros((nb, nb), np.complex128)
        L_dense = None
        for band, block in schedule:
            rhs = np.zeros(n, np.complex128)
            rhs[band] = 1.0
            rhs_t = rhs * sqrt_mask
            x0 = apply_M(rhs_t) if use_preconditioner else None
            x_t, info = scipy.sparse.linalg.gmres(L_op, rhs_t, x0=x0, rtol=tol, maxiter=max_iters,
                                                  M=M_op)
            if info != 0:
                x_t, info = scipy.sparse.linalg.bicgstab(L_op, rhs_t, x0=x_t, rtol=tol,
                                                         maxiter=max_iters, M=M_op)
                if info != 0:
                    if not dense_fallback:
                        raise RuntimeError(
                            "Iterative solvers (GMRES, BiCGSTAB) did not converge for band {} "
                            "at iq={} (info={}). Set dense_fallback=True to use a dense solve."
                            .format(band, iq, info))
                    if L_dense is None:
                        L_dense = np.zeros((n, n), np.complex128)
                        for j in range(n):
                            e = np.zeros(n, np.complex128)
                            e[j] = 1.0
                            L_dense[:, j] = apply_L_tilde(e)
                        L_dense = (L_dense + L_dense.conj().T) / 2
                    x_t = np.linalg.lstsq(L_dense, rhs_t, rcond=1e-10)[0]
            x = x_t * inv_sqrt_mask
            if len(block) == 1:
                G[:, band] = x[:nb]
                for _, other in schedule:
                    if len(other) >= 2:
                        G[other, band] = 0.0
            else:
                d = len(block)
                for j in range(d):
                    col = block[j]
                    G[col, col] = x[band]
                    for _, other in schedule:
                        if other[0] != band and len(other) == d:
                            G[other[j], col] = x[other[0]]
        G = (G + G.conj().T) / 2
        H = np.zeros((nb, nb), np.complex128)
        if non_acoustic:
            na = np.array(non_acoustic)
            H[np.ix_(na, na)] = np.linalg.inv(G[np.ix_(na, na)])
        H = (H + H.conj().T) / 2
        if self.verbose:
            print("  q-point %d: %d solves, %d static L applications" % (iq, len(schedule),
                                                                         counter[0]))
        self.H_q_dict[iq] = H
        return H

    def compute_full_hessian(self, tol=1e-6, max_iters=500, use_preconditioner=True,
                             dense_fallback=False, use_mode_symmetry=True, checkpoint=None):
        """QSH:958-1049: H(q) of every q-point, (n_q, nb, nb), from one solve
        per orbit; with ``checkpoint`` the solved q-points are saved after
        each and skipped on restart."""
        done = set()
        if checkpoint is not None:
            f = checkpoint if checkpoint.lower().endswith(".npz") else checkpoint + ".npz"
            if os.path.exists(f):
                done = set(self.load_status(checkpoint))
        for iq in self.irr_qpoints:
            if iq not in done:
                self.compute_hessian_at_q(iq, tol, max_iters, use_preconditioner,
                                          dense_fallback, use_mode_symmetry)
                if checkpoint is not None:
                    self.save_status(checkpoint)
        out = np.zeros((self.n_q, self.n_bands, self.n_bands), np.complex128)
        for p in self.irr_qpoints:
            Hp = self.H_q_dict[p]
            for q, j in self._orbits[p]:
                if j is None:
                    out[q] = Hp
                else:
                    Dm = self._D[j][p]
                    out[q] = Dm @ Hp @ Dm.conj().T
        return out

    def save_status(self, fname):
        """QSH:132-172: the Hessians solved so far."""
        if not fname.lower().endswith(".npz"):
            fname += ".npz"
        iqs = np.array(sorted(self.H_q_dict), dtype=np.int64)
        H = np.array([self.H_q_dict[i] for i in iqs]) if len(iqs) else np.zeros((0, 0, 0))
        np.savez_compressed(fname, completed_iqs=iqs, H_q=H, T=self.T, n_q=self.n_q,
                            n_bands=self.n_bands, ignore_v3=self.ignore_v3,
                            ignore_v4=self.ignore_v4)

    def load_status(self, fname):
        """QSH:174-247: returns the q-points loaded."""
        if not fname.lower().endswith(".npz"):
            fname += ".npz"
        with np.load(fname, allow_pickle=False) as d:
            if int(d["n_q"]) != self.n_q or int(d["n_bands"]) != self.n_bands:
                raise ValueError("the status file belongs to another system")
            iqs = [int(i) for i in d["completed_iqs"]]
            for i, H in zip(iqs, d["H_q"]):
                self.H_q_dict[i] = H
        return iqs

    def close(self):
        if self.qlanc is not None:
            self.qlanc.close()
