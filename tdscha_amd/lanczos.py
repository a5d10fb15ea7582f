"""Drop-in ``Lanczos`` for the TD-SCHA linear-response hot path on MI355X.

Mirrors the public surface of the reference class that this path uses
(Modules/DynamicalLanczos.py, class Lanczos, DL:182): construction from an
ensemble (DL:183-452), ``init`` / ``prepare_symmetrization`` (DL:564,
694-871), ``prepare_perturbation`` / ``prepare_mode`` / ``prepare_ir``
(DL:2023-2091, 1705-1749), ``apply_full_L`` (DL:3549-3651), ``run_FT``
(DL:5958-6491), ``get_green_function_continued_fraction`` (DL:5032-5162),
``mask_dot_wigner`` (DL:5905-5954), ``save_status`` / ``load_status`` /
``save_abc`` / ``load_abc`` (DL:3653-3839) and the standalone C++ driver's
files ``prepare_input_files`` / ``load_from_input_files`` (DL:873-1044).

Every L application, the perturbed averages, the gamma-only translational
projection, the batched device-resident Lanczos and the continued fraction
run in the gfx950 library (``libtdscha_amd.so``).  ``run_FT`` keeps the
reference's host-side recurrence (biconjugate or symmetric, same
coefficients and quirks) around device L applications, so its output is
comparable step by step with the reference; ``run_FT_batched`` is the
device-resident batched recurrence (optionally with full
reorthogonalisation) used for throughput.
"""
from __future__ import annotations

import ctypes as C
import difflib
import json
import os
import warnings

import numpy as np

from . import _lib
from .engine import Engine, precision_code

TYPE_DP = np.double
EPSILON = 1e-12
RY_TO_K = 157887.32400374097
# largest image ensemble (U and F in FP64) that symmetry_form "auto" keeps on
# the device; beyond it the symmetries are applied by rotating operators
IMAGE_BUDGET_BYTES = 64 * 2 ** 30

MODE_FAST_SERIAL = 1   # names kept for API compatibility; all run on the GPU
MODE_FAST_MPI = 2
MODE_FAST_JULIA = 3
MODE_GPU = 4


def _p(a):
    return C.c_void_p(a.ctypes.data)


def continued_fraction(a_coeffs, b_coeffs, c_coeffs, w_array, use_wigner, use_terminator=True,
                       last_average=1, smearing=0.0, smooth_ramp=0, reverse_L=False,
                       shift_value=0.0, perturbation_modulus=1.0):
    """Continued-fraction Green function (DL:5032-5162) on the GPU.

    ``c_coeffs`` is used when it has as many entries as ``b_coeffs``
    (biconjugate run), otherwise c = b (symmetric Lanczos).
    """
    w_array = np.ascontiguousarray(w_array, dtype=TYPE_DP)
    a = np.ascontiguousarray(a_coeffs, dtype=TYPE_DP)
    b = np.ascontiguousarray(b_coeffs, dtype=TYPE_DP)
    cc = None
    if c_coeffs is not None and len(c_coeffs) == len(b_coeffs) and len(b) > 0:
        cc = np.ascontiguousarray(c_coeffs, dtype=TYPE_DP)
    out = np.zeros((len(w_array), 2))
    rc = _lib.lib().tdscha_continued_fraction(
        len(a), _p(a), len(b), _p(b) if len(b) else None, None if cc is None else _p(cc),
        len(w_array), _p(w_array), int(bool(use_wigner)), int(bool(use_terminator)),
        int(last_average), float(smearing), int(smooth_ramp), int(bool(reverse_L)),
        float(shift_value), float(perturbation_modulus), _p(out))
    _lib.check(rc, "tdscha_continued_fraction")
    return out[:, 0] + 1j * out[:, 1]


class Lanczos(object):
    def __init__(self, ensemble=None, mode=None, unwrap_symmetries=False, select_modes=None,
                 use_wigner=True, lo_to_split="random"):
        self.mode = MODE_GPU if mode is None else mode
        self.verbose = False
        self.T = 0
        self.nat = 0
        self.m = []
        self.w = []
        self.pols = []
        self.n_modes = 0
        self.ignore_harmonic = False
        self.ignore_v3 = False
        self.ignore_v4 = False
        self.N = 0
        self.rho = []
        self.N_eff = 0
        self.X = []
        self.Y = []
        self.psi = []
        self.a_coeffs = []
        self.b_coeffs = []
        self.c_coeffs = []
        self.krilov_basis = []
        self.basis_P = []
        self.basis_Q = []
        self.s_norm = []
        self.arnoldi_matrix = []
        self.reverse_L = False
        self.shift_value = 0
        self.symmetries = None
        self.degenerate_space = None
        self.N_degeneracy = None
        self.sym_block_id = None
        self.n_syms = 1
        self.initialized = False
        self.perturbation_modulus = 1
        self.gamma_only = False
        self.trans_projector = None
        self.trans_operators = None
        self.trans_cart_perms = None
        self.use_wigner = use_wigner
        self.ignore_small_w = False
        self.precision = "fp64"
        # "images": the symmetric images live on the device; "operators": the
        # configurations only, symmetries applied by rotating the perturbation
        # and the results; "auto": images while they take at most
        # IMAGE_BUDGET_BYTES
        self.symmetry_form = "auto"
        self.ensemble = None
        self.engine = None
        self._sym_source = None
        self.L_linop = None
        self.M_linop = None
        self.__total_attributes__ = [item for item in self.__dict__.keys()]
        self.fixed_attributes = True
        if ensemble is None:
            return
        self.ignore_small_w = bool(getattr(ensemble, "ignore_small_w", False))
        # mode-basis quantities (DL:322-452); the ensemble object provides them
        self.ensemble = ensemble
        self.T = float(ensemble.T)
        self.w = np.asarray(ensemble.w, dtype=TYPE_DP)
        self.pols = np.asarray(ensemble.pols, dtype=TYPE_DP)
        self.m = np.asarray(ensemble.m, dtype=TYPE_DP)
        self.nat = len(self.m) // 3
        if select_modes is not None:
            select_modes = np.asarray(select_modes, dtype=bool)
            if len(select_modes) != len(self.w):
                raise ValueError("Error, 'select_modes' should be an array of the same length "
                                 "of the number of modes.")
            self.w = self.w[select_modes]
            self.pols = self.pols[:, select_modes]
        self.n_modes = len(self.w)
        self.N = int(ensemble.N)
        self.rho = np.asarray(ensemble.rho, dtype=TYPE_DP).copy()
        self.N_eff = np.sum(self.rho)
        X = np.asarray(ensemble.X, dtype=TYPE_DP)
        Y = np.asarray(ensemble.Y, dtype=TYPE_DP)
        if select_modes is not None:
            X = X[:, select_modes]
            Y = Y[:, select_modes]
        self.X = np.ascontiguousarray(X)
        self.Y = np.ascontiguousarray(Y)
        self.psi = np.zeros(self.n_modes + self.n_modes * (self.n_modes + 1), dtype=TYPE_DP)
        self._make_linops()

    def _make_linops(self):
        """L and the harmonic preconditioner as SciPy LinearOperators
        (DL:478-489, rebuilt by load_status as DL:3829-3838)."""
        import scipy.sparse.linalg
        size = len(self.psi)
        self.L_linop = scipy.sparse.linalg.LinearOperator(
            shape=(size, size), matvec=lambda psi: self.apply_full_L(np.ravel(psi)),
            rmatvec=lambda psi: self.apply_full_L(np.ravel(psi), transpose=True), dtype=TYPE_DP)
        self.M_linop = scipy.sparse.linalg.LinearOperator(
            shape=(size, size), matvec=self.apply_L1_inverse_FT,
            rmatvec=lambda psi: self.apply_L1_inverse_FT(psi, transpose=True), dtype=TYPE_DP)

    def __setattr__(self, name, value):
        """Reject unknown attributes with close-match suggestions (DL:508-526)."""
        if "fixed_attributes" in self.__dict__:
            if name in self.__total_attributes__:
                super(Lanczos, self).__setattr__(name, value)
            elif self.fixed_attributes:
                similar = str(difflib.get_close_matches(name, self.__total_attributes__))
                raise AttributeError("Error, the attribute '{}' is not a member of '{}'.\n"
                                     "Suggested similar attributes: {} ?".format(
                                         name, type(self).__name__, similar))
        else:
            super(Lanczos, self).__setattr__(name, value)

    # ------------------------------------------------------------ setup
    def reset(self):
        """DL:528-561."""
        self.psi = np.zeros(self.n_modes + self.n_modes * (self.n_modes + 1), dtype=TYPE_DP)
        self.a_coeffs = []
        self.b_coeffs = []
        self.c_coeffs = []
        self.basis_P = []
        self.basis_Q = []
        self.s_norm = []
        self.krilov_basis = []
        self.arnoldi_matrix = []

    def init(self, use_symmetries=True, symmetries=None):
        """DL:564-580: allocate psi, set up the symmetries and upload the ensemble."""
        self.psi = np.zeros(self.n_modes + self.n_modes * (self.n_modes + 1), dtype=TYPE_DP)
        self.prepare_symmetrization(no_sym=not use_symmetries, symmetries=symmetries)
        self.initialized = True

    def prepare_symmetrization(self, no_sym=False, verbose=True, symmetries=None):
        """DL:694-871.  ``symmetries`` may be a ModeSymmetries object; otherwise
        the ensemble's own builder is asked (rock-salt models)."""
        self.initialized = True
        if no_sym:
            n = self.n_modes
            self.symmetries = [np.ones((1, 1, 1), dtype=TYPE_DP)] * n
            self.N_degeneracy = np.ones(n, dtype=np.intc)
            self.degenerate_space = [np.array([i], dtype=np.intc) for i in range(n)]
            self.sym_block_id = np.arange(n).astype(np.intc)
            self.n_syms = 1
            self.trans_projector = None
            self.trans_cart_perms = None
        else:
            ms = symmetries
            if ms is None:
                if self.ensemble is None or not hasattr(self.ensemble, "symmetries"):
                    raise ImportError("no symmetry source: pass symmetries= or use no_sym")
                ms = self.ensemble.symmetries(True, gamma_only=self.gamma_only)
            self.symmetries = [np.ascontiguousarray(s, TYPE_DP) for s in ms.symmetries]
            self.degenerate_space = [np.asarray(b, dtype=np.intc) for b in ms.degenerate_space]
            self.N_degeneracy = np.array([len(b) for b in self.degenerate_space], dtype=np.intc)
            self.sym_block_id = np.asarray(ms.sym_block_id, dtype=np.intc)
            self.n_syms = self.symmetries[0].shape[0]
            if self.gamma_only and ms.trans_projector is not None:
                self.trans_projector = ms.trans_projector
                self.trans_operators = ms.trans_operators
                self.trans_cart_perms = ms.trans_cart_perms
            else:
                self.trans_projector = None
                self.trans_cart_perms = None
        self._upload()

    def rotates_operators(self):
        """Whether the engine applies the symmetries by rotating operators
        (see symmetry_form)."""
        form = self.symmetry_form
        if form not in ("auto", "images", "operators"):
            raise ValueError("symmetry_form must be 'auto', 'images' or 'operators'")
        if form == "auto":
            np_ = -(-self.n_modes // 128) * 128
            image_bytes = 16.0 * self.N * self.n_syms * np_     # U and F in FP64
            return self.n_syms > 1 and image_bytes > IMAGE_BUDGET_BYTES
        return form == "operators" and self.n_syms > 1

    def _upload(self):
        if self.engine is not None:
            self.engine.close()
        self.engine = Engine(self.X, self.Y, self.w, self.rho, self.T, self.symmetries,
                             self.degenerate_space, self.sym_block_id,
                             rotate=self.rotates_operators())
        L = _lib.lib()
        if self.trans_projector is not None:
            perms = np.ascontiguousarray(np.array(self.trans_cart_perms), dtype=np.int32)
            pols = np.ascontiguousarray(self.pols, TYPE_DP)
            tp = np.ascontiguousarray(self.trans_projector, TYPE_DP)
            rc = L.tdscha_set_translations(self.engine.ctx, pols.shape[0], _p(pols),
                                           perms.shape[0], _p(perms), _p(tp))
            _lib.check(rc, "tdscha_set_translations")

    # -------------------------------------------------------- perturbations
    def prepare_perturbation(self, vector, masses_exp=1, add=False):
        """DL:2023-2073: psi[:n] += sum_a v_a m_a^(e/2) pols[a, :]."""
        if not add:
            self.reset()
            self.psi = np.zeros(self.psi.shape, dtype=TYPE_DP)
        m_on = np.sqrt(self.m) ** masses_exp
        new_v = np.einsum("a, a, ab->b", m_on, vector, self.pols)
        self.psi[:self.n_modes] += new_v
        self.perturbation_modulus = new_v.dot(new_v)

    def prepare_mode(self, index):
        """DL:2077-2091."""
        self.reset()
        self.psi[:] = 0
        self.psi[index] = 1
        self.perturbation_modulus = 1

    def prepare_ir(self, effective_charges=None, pol_vec=np.array([1, 0, 0])):
        """DL:1705-1749 with effective charges (nat_uc, 3, 3)."""
        ec = effective_charges
        if ec is None:
            raise ValueError("Error, no effective charge found. Cannot initialize IR response")
        ec = np.asarray(ec, dtype=TYPE_DP)
        n_cells = self.nat // ec.shape[0]
        msg = ("Error, effective charges of the wrong shape: {}\nNumber of modes : {}\n"
               "Dimension of the supercell : {}".format(ec.shape, self.n_modes, n_cells))
        if ec.ndim != 3 or ec.shape[1] != 3 or ec.shape[2] != 3:
            raise ValueError(msg)
        if not self.ignore_small_w and ec.shape[0] * ec.shape[2] * n_cells != self.n_modes + 3:
            raise ValueError(msg)     # DL:1739-1741: only the 3 translations are dropped
        z_eff = np.einsum("abc, b", ec, pol_vec)
        self.prepare_perturbation(np.tile(z_eff.ravel(), n_cells), masses_exp=-1)

    # --------------------------------------------------------------- L
    def _L_flags(self, transpose):
        return (int(self.use_wigner), int(transpose), int(self.ignore_harmonic),
                int(self.ignore_v3), int(self.ignore_v4), int(bool(self.reverse_L)),
                float(self.shift_value), precision_code(self.precision))

    def apply_L_batch(self, psis, transpose=False):
        """L applied to up to 8 vectors at once (rows of ``psis``)."""
        if not self.initialized or self.engine is None:
            raise ValueError("Error, this class must be initialized before lunching the "
                             "computation:\n Use the .init()")
        psis = np.ascontiguousarray(np.atleast_2d(psis), dtype=TYPE_DP)
        out = np.empty_like(psis)
        rc = _lib.lib().tdscha_apply_L(self.engine.ctx, psis.shape[0], _p(psis), _p(out),
                                       *self._L_flags(transpose))
        _lib.check(rc, "tdscha_apply_L")
        return out

    def apply_full_L(self, target=None, force_t_0=False, force_FT=True, transpose=False,
                     fast_lanczos=True):
        """DL:3549-3651.  Overwrites self.psi with the result and returns it."""
        if force_t_0 and force_FT:
            raise ValueError("Error, only one between force_t_0 and force_FT can be True")
        if target is not None:
            self.psi = target
        out = self.apply_L_batch(self.psi[None, :], transpose=transpose)[0]
        self.psi = out
        return self.psi

    def apply_L1_inverse_FT(self, psi, transpose=False):
        """DL:2451-2549: the inverse of the harmonic 2x2 pair blocks of the
        standard representation (the preconditioner of M_linop), with the
        reference's regularisation, on the device.  Does not touch self.psi."""
        if not self.initialized or self.engine is None:
            raise ValueError("Error, this class must be initialized before lunching the "
                             "computation:\n Use the .init()")
        psi = np.ascontiguousarray(np.ravel(psi), dtype=TYPE_DP)
        if len(psi) != self.n_modes + self.n_modes * (self.n_modes + 1):
            raise ValueError("ERROR,\nThe initial vector for the Lanczos algorithm has a wrong "
                             "dimension.")
        out = np.empty_like(psi)
        rc = _lib.lib().tdscha_apply_L1_inverse(self.engine.ctx, 1, _p(psi), _p(out),
                                                int(bool(transpose)), int(self.ignore_harmonic))
        _lib.check(rc, "tdscha_apply_L1_inverse")
        return out

    def L_matvec(self, v):
        return self.apply_full_L(np.array(v, dtype=TYPE_DP))

    def L_rmatvec(self, v):
        return self.apply_full_L(np.array(v, dtype=TYPE_DP), transpose=True)

    # -------------------------------------------------------------- Lanczos
    def mask_dot_wigner(self, debug=False):
        """DL:5905-5954."""
        n = self.n_modes
        mask = np.ones(n + n * (n + 1))
        rows, cols = np.triu_indices(n)
        nw2 = len(rows)
        off = rows < cols
        mask[n:n + nw2][off] = 2
        mask[n + nw2:][off] = 2
        return mask

    def run_FT(self, n_iter, save_dir=None, save_each=5, verbose=True, n_rep_orth=0, n_ortho=10,
               flush_output=True, debug=False, prefix="LANCZOS", run_simm=None, optimized=False,
               device=False):
        """DL:5958-6491: the reference recurrence around device L applications.

        ``device=True`` runs the same recurrence (biconjugate, or symmetric
        with run_simm, n_rep_orth = 0) entirely on the device for a fresh
        chain: only the coefficients come back (a_coeffs, b_coeffs,
        c_coeffs), the Krylov bases are not kept."""
        if device:
            return self._run_FT_device(n_iter, run_simm, verbose)
        self.verbose = verbose
        if not self.initialized:
            self.prepare_symmetrization()
        if self.psi is None:
            raise ValueError("Error, you must initialize a perturbation to start the Lanczos.")
        psi_norm = np.sum(self.psi ** 2)
        if np.isnan(psi_norm) or psi_norm == 0:
            raise ValueError("Error, you must initialize a perturbation to start the Lanczos.")
        if save_dir is not None and not os.path.exists(save_dir):
            os.makedirs(save_dir)
        if run_simm is None:
            run_simm = self.use_wigner
        if run_simm and not self.use_wigner:
            raise NotImplementedError("The symmetric Lanczos works only with Wigner.")
        mask = self.mask_dot_wigner() if run_simm else np.ones(len(self.psi))
        dot = (lambda u, v: u.dot(v * mask)) if run_simm else (lambda u, v: u.dot(v))
        i_step = len(self.a_coeffs)
        if i_step == 0:
            first = self.psi / np.sqrt(dot(self.psi, self.psi))
            self.basis_Q = [first]
            self.basis_P = [first]
            self.s_norm = [1]
        else:
            self.basis_Q = list(self.basis_Q)
            self.basis_P = list(self.basis_P)
            self.s_norm = list(self.s_norm)
            self.a_coeffs = list(self.a_coeffs)
            self.b_coeffs = list(self.b_coeffs)
            self.c_coeffs = list(self.c_coeffs)
        psi_q = self.basis_Q[-1]
        psi_p = self.basis_P[-1]
        next_converged = False
        converged = False
        for i in range(i_step, i_step + n_iter):
            if not self.use_wigner or not run_simm:
                L_q = self.L_matvec(psi_q)
                p_L = self.L_rmatvec(psi_p)
            else:
                L_q = self.L_matvec(psi_q)
                p_L = np.copy(L_q)
            c_old = self.c_coeffs[-1] if len(self.c_coeffs) > 0 else 1
            p_norm = self.s_norm[-1] / c_old
            a_coeff = dot(psi_p, L_q) * p_norm
            if np.isnan(a_coeff):
                raise ValueError("Invalid value encountered during the Lanczos.")
            rk = L_q - a_coeff * psi_q
            if len(self.basis_Q) > 1:
                rk -= self.c_coeffs[-1] * self.basis_Q[-2]
            sk = p_L - a_coeff * psi_p
            if len(self.basis_P) > 1:
                old_p_norm = self.s_norm[-2] if len(self.c_coeffs) < 2 else \
                    self.s_norm[-2] / self.c_coeffs[-2]
                sk -= self.b_coeffs[-1] * self.basis_P[-2] * (old_p_norm / p_norm)
            s_norm = np.sqrt(dot(sk, sk))
            sk_tilde = sk / s_norm
            s_norm *= p_norm
            b_coeff = np.sqrt(dot(rk, rk))
            c_coeff = dot(sk_tilde, rk / b_coeff) * s_norm
            self.a_coeffs.append(a_coeff)
            if np.abs(b_coeff) < EPSILON or next_converged:
                converged = True
                break
            if np.abs(c_coeff) < EPSILON or next_converged:
                converged = True
                break
            psi_q = rk / b_coeff
            psi_p = sk_tilde.copy()
            new_q = psi_q.copy()
            new_p = psi_p.copy()
            for _k in range(n_rep_orth):
                ortho_q = 0
                ortho_p = 0
                start = 0 if n_ortho is None else max(len(self.basis_P) - n_ortho, 0)
                for j in range(start, len(self.basis_P)):
                    coeff1 = dot(self.basis_P[j], new_q)
                    coeff2 = dot(self.basis_Q[j], new_p)
                    new_q -= coeff1 * self.basis_P[j]
                    new_p -= coeff2 * self.basis_Q[j]
                    ortho_q += np.abs(coeff1)
                    ortho_p += np.abs(ortho_p)   # reference quirk (DL:6391)
                normq = np.sqrt(dot(new_q, new_q))
                if normq < EPSILON:
                    next_converged = True
                new_q /= normq
                normp = dot(new_p, new_p)        # squared norm, as DL:6412-6425
                if np.abs(normp) < EPSILON:
                    next_converged = True
                new_p /= normp
                s_norm = c_coeff / dot(new_p, new_q)
                if ortho_p < EPSILON and ortho_q < EPSILON:
                    break
            self.basis_Q.append(new_q)
            self.basis_P.append(new_p)
            psi_q = new_q.copy()
            psi_p = new_p.copy()
            self.b_coeffs.append(b_coeff)
            self.c_coeffs.append(c_coeff)
            self.s_norm.append(s_norm)
            if optimized:
                for lst in (self.basis_P, self.basis_Q, self.s_norm):
                    if len(lst) > 3:
                        lst.pop(-4)
            if verbose:
                print("a_%d = %.8e" % (i, self.a_coeffs[-1]))
                print("b_%d = %.8e" % (i, self.b_coeffs[-1]))
                print("c_%d = %.8e" % (i, self.c_coeffs[-1]))
            if save_dir is not None and (i + 1) % save_each == 0:
                self.save_status("%s/%s_STEP%d" % (save_dir, prefix, i + 1))
        return converged

    def _run_FT_device(self, n_iter, run_simm, verbose):
        if not self.initialized:
            self.prepare_symmetrization()
        if len(self.a_coeffs) > 0:
            raise NotImplementedError("the device recurrence starts a fresh chain; call reset()")
        psi_norm = np.sum(self.psi ** 2)
        if np.isnan(psi_norm) or psi_norm == 0:
            raise ValueError("Error, you must initialize a perturbation to start the Lanczos.")
        if run_simm is None:
            run_simm = self.use_wigner
        if run_simm and not self.use_wigner:
            raise NotImplementedError("The symmetric Lanczos works only with Wigner.")
        psi = np.ascontiguousarray(self.psi[None, :], dtype=TYPE_DP)
        a = np.zeros((1, n_iter))
        b = np.zeros((1, n_iter))
        c = np.zeros((1, n_iter))
        nd = np.zeros(1, dtype=np.int32)
        ms = C.c_double(0.0)
        flags = (int(self.ignore_harmonic), int(self.ignore_v3), int(self.ignore_v4))
        prec = precision_code(self.precision)
        if run_simm:
            if self.reverse_L or self.shift_value:
                raise NotImplementedError("the device symmetric recurrence has no reverse/shift")
            rc = _lib.lib().tdscha_lanczos_run(
                self.engine.ctx, 1, _p(psi), int(n_iter), None, EPSILON, 0, prec, *flags, 1,
                _p(a), _p(b), _p(nd), None, C.byref(ms))
            _lib.check(rc, "tdscha_lanczos_run")
            k = int(nd[0])
            nb = k if (k == n_iter or abs(b[0, k - 1]) >= EPSILON) else k - 1
            self.a_coeffs, self.b_coeffs = list(a[0, :k]), list(b[0, :nb])
            self.c_coeffs = list(b[0, :nb])
        else:
            rc = _lib.lib().tdscha_lanczos_run_biconjugate(
                self.engine.ctx, 1, _p(psi), int(n_iter), None, EPSILON, int(self.use_wigner),
                prec, *flags, int(bool(self.reverse_L)), float(self.shift_value), 1, _p(a),
                _p(b), _p(c), _p(nd), C.byref(ms))
            _lib.check(rc, "tdscha_lanczos_run_biconjugate")
            k = int(nd[0])
            # a chain that stopped on a small b or c leaves that slot zero
            kept = abs(b[0, k - 1]) >= EPSILON and abs(c[0, k - 1]) >= EPSILON
            nb = k if kept else k - 1
            self.a_coeffs, self.b_coeffs = list(a[0, :k]), list(b[0, :nb])
            self.c_coeffs = list(c[0, :nb])
        if verbose:
            for i in range(len(self.a_coeffs)):
                print("a_%d = %.8e" % (i, self.a_coeffs[i]))
        return len(self.b_coeffs) < len(self.a_coeffs)

    def run_FT_batched(self, psis, n_iter, caps=None, tol=EPSILON, reorth=True,
                       precision=None, use_graph=True, return_basis=False):
        """Device-resident symmetric (Wigner) Lanczos for up to 64 chains at once.

        Returns a list of (a, b) coefficient arrays per chain and the device
        time in ms.  Each chain stops at |b| < tol or at its cap.  use_graph:
        True or "step" replays a captured step, "run" captures the whole run
        as one graph, False launches kernels plainly.
        """
        graph_mode = {False: 0, True: 1, "step": 1, "run": 2}[use_graph]
        if not self.use_wigner:
            raise NotImplementedError("the batched device Lanczos is the symmetric Wigner one")
        psis = np.ascontiguousarray(np.atleast_2d(psis), dtype=TYPE_DP)
        nc = psis.shape[0]
        a = np.zeros((nc, n_iter))
        b = np.zeros((nc, n_iter))
        nd = np.zeros(nc, dtype=np.int32)
        capv = None if caps is None else np.ascontiguousarray(caps, dtype=np.int32)
        basis = None
        if return_basis:
            basis = np.zeros((nc, n_iter + 1, psis.shape[1]))
        ms = C.c_double(0.0)
        prec = precision_code(self.precision if precision is None else precision)
        rc = _lib.lib().tdscha_lanczos_run(
            self.engine.ctx, nc, _p(psis), int(n_iter), None if capv is None else _p(capv),
            float(tol), int(bool(reorth)), prec, int(self.ignore_harmonic), int(self.ignore_v3),
            int(self.ignore_v4), graph_mode, _p(a), _p(b), _p(nd),
            None if basis is None else _p(basis), C.byref(ms))
        _lib.check(rc, "tdscha_lanczos_run")
        chains = []
        for c in range(nc):
            k = int(nd[c])
            nb = k if (k == n_iter or abs(b[c, k - 1]) >= tol) else k - 1
            chains.append((a[c, :k].copy(), b[c, :nb].copy()))
        if return_basis:
            return chains, ms.value, basis
        return chains, ms.value

    # ----------------------------------------------------- Green function
    def get_green_function_continued_fraction(self, w_array, use_terminator=True,
                                              last_average=1, smearing=0.0, smooth_ramp=0):
        """DL:5032-5162, evaluated on the GPU (one thread per frequency)."""
        return continued_fraction(self.a_coeffs, self.b_coeffs, self.c_coeffs, w_array,
                                  self.use_wigner, use_terminator, last_average, smearing,
                                  smooth_ramp, self.reverse_L, self.shift_value,
                                  self.perturbation_modulus)

    # ------------------------------------------------------------- I/O
    def save_abc(self, file):
        """DL:3653-3666."""
        total_len = len(self.a_coeffs)
        abc = np.zeros((total_len, 3), dtype=np.double)
        abc[:, 0] = self.a_coeffs
        abc[:len(self.b_coeffs), 1] = self.b_coeffs
        abc[:len(self.c_coeffs), 2] = self.c_coeffs
        np.savetxt(file, abc, header="a; b; c")

    def load_abc(self, file):
        """DL:3668-3676."""
        abc = np.loadtxt(file)
        self.a_coeffs = abc[:, 0]
        self.b_coeffs = abc[:, 1]
        self.c_coeffs = abc[:, 2]

    def save_status(self, file):
        """DL:3678-3727: same npz keys (object-valued Julia keys omitted)."""
        if ".npz" not in file.lower():
            file += ".npz"
        np.savez_compressed(file, T=self.T, nat=self.nat, m=self.m, w=self.w, pols=self.pols,
                            n_modes=self.n_modes, ignore_v3=self.ignore_v3,
                            ignore_v4=self.ignore_v4, N=self.N, rho=self.rho, X=self.X, Y=self.Y,
                            psi=self.psi, a_coeffs=self.a_coeffs, b_coeffs=self.b_coeffs,
                            c_coeffs=self.c_coeffs, basis_Q=self.basis_Q, basis_P=self.basis_P,
                            s_norm=self.s_norm, krilov_basis=self.krilov_basis,
                            arnoldi_matrix=self.arnoldi_matrix, reverse=self.reverse_L,
                            shift=self.shift_value,
                            perturbation_modulus=self.perturbation_modulus,
                            use_wigner=self.use_wigner, ignore_small_w=self.ignore_small_w,
                            n_syms=self.n_syms)

    def load_status(self, file, is_file_instance=False):
        """DL:3730-3839.  Loads without pickle; object-valued keys are skipped."""
        if not is_file_instance:
            if ".npz" not in file.lower():
                file += ".npz"
            if not os.path.exists(file):
                raise IOError("Error while loading %s" % file)
        data = {}
        with np.load(file, allow_pickle=False) as d:
            for k in d.files:
                try:
                    data[k] = d[k]
                except ValueError:
                    pass
        self.T = np.double(data["T"])
        self.nat = int(data["nat"])
        self.m = data["m"]
        self.w = data["w"]
        self.pols = data["pols"]
        self.n_modes = int(data["n_modes"])
        self.ignore_v3 = bool(data["ignore_v3"])
        self.ignore_v4 = bool(data["ignore_v4"])
        self.N = int(data["N"])
        self.rho = data["rho"]
        self.N_eff = np.sum(self.rho)
        self.X = np.ascontiguousarray(data["X"])
        self.Y = np.ascontiguousarray(data["Y"])
        self.psi = data["psi"]
        self.a_coeffs = list(data["a_coeffs"])
        self.b_coeffs = list(data["b_coeffs"])
        self.c_coeffs = list(data["c_coeffs"]) if "c_coeffs" in data else []
        self.krilov_basis = data.get("krilov_basis", [])
        self.arnoldi_matrix = data.get("arnoldi_matrix", [])
        self.basis_Q = list(data["basis_Q"])
        self.basis_P = list(data["basis_P"])
        self.s_norm = list(data["s_norm"])
        self.use_wigner = bool(data["use_wigner"])
        self.ignore_small_w = bool(data.get("ignore_small_w", False))
        if "reverse" in data:
            self.reverse_L = bool(data["reverse"])
            self.shift_value = float(data["shift"])
        if "perturbation_modulus" in data:
            self.perturbation_modulus = float(data["perturbation_modulus"])
        self.n_syms = int(data.get("n_syms", 1))
        self.initialized = False
        self._make_linops()

    def prepare_input_files(self, root_name="tdscha", n_steps=100, start_from_scratch=True,
                            directory=".", run_symm=False):
        """DL:873-978: the input files of the reference's standalone C++
        driver (``<root>.json``, ``.ndegs``, ``.blockid``, ``.masses``,
        ``.freqs``, ``.rho``, ``.X.dat``, ``.Y.dat``, ``.psi``, and per
        symmetry block ``.block<i>`` and ``.symsb<i>``), in the same layout,
        so that a run prepared here can go to that driver and back."""
        if not start_from_scratch:
            raise NotImplementedError("Error, restarting from a previous calculation is not yet "
                                      "implemented.")
        if self.symmetries is None:
            raise ValueError("call init() before prepare_input_files")
        nsyms = int(np.shape(self.symmetries[0])[0])
        json_data = {"T": float(self.T), "n_steps": int(n_steps),
                     "ignore_v2": bool(self.ignore_harmonic), "ignore_v3": bool(self.ignore_v3),
                     "ignore_v4": bool(self.ignore_v4), "use_wigner": bool(self.use_wigner),
                     "run_sym": bool(run_symm),
                     "data": {"n_configs": int(self.N), "n_modes": int(self.n_modes),
                              "n_syms": nsyms, "n_blocks": len(self.symmetries),
                              "perturbation_modulus": float(self.perturbation_modulus),
                              "reverse": bool(self.reverse_L), "shift": float(self.shift_value)}}
        os.makedirs(directory, exist_ok=True)
        root = os.path.join(directory, root_name)
        with open(root + ".json", "w") as fp:
            json.dump(json_data, fp)
        np.savetxt(root + ".ndegs", self.N_degeneracy, fmt="%d")
        np.savetxt(root + ".blockid", self.sym_block_id, fmt="%d")
        np.savetxt(root + ".masses", self.m)
        np.savetxt(root + ".freqs", self.w)
        np.savetxt(root + ".rho", self.rho)
        np.savetxt(root + ".X.dat", self.X)
        np.savetxt(root + ".Y.dat", self.Y)
        np.savetxt(root + ".psi", self.psi)
        for i, (blk, sym) in enumerate(zip(self.degenerate_space, self.symmetries)):
            np.savetxt(root + ".block%d" % i, blk, fmt="%d")
            ns, b1, b2 = sym.shape
            with open(root + ".symsb%d" % i, "w") as fp:
                for isym in range(ns):
                    for k1 in range(b1):
                        fp.write("".join(" {:22.16f}".format(sym[isym, k1, k2]) for k2 in range(b2)))
                        fp.write("\n")
                    fp.write("\n")

    def load_from_input_files(self, root_name="tdscha", directory="."):
        """DL:981-1044: the coefficients (``<root>.abc``) and, when present,
        the Krylov bases and norms a run of the C++ driver wrote, with the
        run's settings from ``<root>.json``."""
        if not os.path.exists(directory):
            raise IOError("Error, the specified directory '%s' does not exist" % directory)
        abc_file = os.path.join(directory, "%s.abc" % root_name)
        if not os.path.exists(abc_file):
            raise IOError("Error, the file '%s' does not exist." % abc_file)
        abc = np.atleast_2d(np.loadtxt(abc_file, dtype=np.double))
        self.a_coeffs = abc[:, 0]
        self.b_coeffs = abc[:, 1]
        self.c_coeffs = abc[:, 2]
        for attr, suffix in (("basis_Q", "qbasis.out"), ("basis_P", "pbasis.out"),
                             ("s_norm", "snorm.out")):
            fname = os.path.join(directory, "%s.%s" % (root_name, suffix))
            if os.path.exists(fname):
                v = np.loadtxt(fname, dtype=np.double)
                setattr(self, attr, v.ravel() if attr == "s_norm" else v)
            else:
                warnings.warn("File %s not found. %s not loaded." % (fname, attr))
        with open(os.path.join(directory, root_name + ".json")) as fp:
            js = json.load(fp)
        self.perturbation_modulus = js["data"]["perturbation_modulus"]
        self.T = js["T"]
        self.ignore_harmonic = js["ignore_v2"]
        self.ignore_v3 = js["ignore_v3"]
        self.ignore_v4 = js["ignore_v4"]
        self.N = js["data"]["n_configs"]
        self.n_modes = js["data"]["n_modes"]
        self.reverse_L = js["data"]["reverse"]
        self.shift_value = js["data"]["shift"]

    def close(self):
        if self.engine is not None:
            self.engine.close()
            self.engine = None
