"""Drop-in ``QSpaceLanczos`` for the q-space (Bloch basis) Lanczos on MI355X.

Mirrors the public surface of the reference class (Modules/QSpaceLanczos.py,
class QSpaceLanczos): ``build_q_pair_map`` (QL:297-331), the psi layout
helpers (QL:333-432), ``mask_dot_wigner`` (QL:441-480), ``prepare_mode_q``
/ ``prepare_perturbation_q`` / ``reset_q`` (QL:1186-1436), ``apply_full_L``
(QL:870-893), ``run_FT`` (QL:967-1182, Hermitian, full
reorthogonalisation) and the continued fraction of the parent class.  Every
L application and the whole Lanczos run execute in the gfx950 library
(``tdscha_qctx_*`` / ``tdscha_qspace_*``); ``run_FT_batched`` runs many
chains with different perturbation q-points in one device-resident batch.

Point-group images (``init(use_symmetries=True)``, QL:1471-1633): the q-space
symmetry matrices of ``qspace.q_symmetries`` expand the ensemble on the device
into its n_sym N images (``tdscha_qctx_symmetrize``), the per-(configuration,
symmetry) rotation of JQ:496-502.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from .lanczos import continued_fraction
from .qspace import BlochModes, bloch_ensemble, pair_map as _pair_map, q_symmetries

EPSILON = 1e-12


def _p(a):
    return C.c_void_p(a.ctypes.data)


class QSpaceEngine:
    """Device copy of a q-space ensemble (X_q, Y_q, frequencies, weights)."""

    def __init__(self, X_q, Y_q, w_q, valid_modes_q, rho, T):
        X_q = np.ascontiguousarray(X_q, np.complex128)
        Y_q = np.ascontiguousarray(Y_q, np.complex128)
        self.n_q, self.N, self.n_bands = X_q.shape
        w = np.ascontiguousarray(w_q, np.float64)
        v = np.ascontiguousarray(valid_modes_q, np.int32)
        r = np.ascontiguousarray(rho, np.float64)
        h = C.c_void_p()
        self.n_syms = 1
        _lib.check(_lib.lib().tdscha_qctx_create(C.byref(h), self.n_q, self.n_bands, self.N,
                                                 _p(X_q.view(np.float64)),
                                                 _p(Y_q.view(np.float64)), _p(w), _p(v), _p(r),
                                                 float(T)), "tdscha_qctx_create")
        self.ctx = h

    def symmetrize(self, perm, D):
        """Expand the ensemble into its images under the symmetries (perm, D)
        of qspace.q_symmetries; once, before any run."""
        perm = np.ascontiguousarray(perm, np.int32)
        D = np.ascontiguousarray(D, np.complex128)
        ns = perm.shape[0]
        if perm.shape != (ns, self.n_q) or D.shape != (ns, self.n_q, self.n_bands, self.n_bands):
            raise ValueError("symmetry tables do not match the ensemble's q-points and bands")
        _lib.check(_lib.lib().tdscha_qctx_symmetrize(self.ctx, ns, _p(perm),
                                                     _p(D.view(np.float64))),
                   "tdscha_qctx_symmetrize")
        self.N *= ns
        self.n_syms = ns

    def psi_size(self, pm):
        n = C.c_int()
        pm = np.ascontiguousarray(pm, np.int32)
        _lib.check(_lib.lib().tdscha_qspace_psi_size(self.n_q, self.n_bands, _p(pm), C.byref(n)),
                   "tdscha_qspace_psi_size")
        return n.value

    def _check_lengths(self, pms, psis):
        if len(psis) != len(pms):
            raise ValueError("one pair map and one psi per chain")
        for i, p in enumerate(psis):
            n = self.psi_size(pms[i])
            if len(p) != n:
                raise ValueError("chain %d: psi of %d entries, its layout has %d"
                                 % (i, len(p), n))

    def static_psi_size(self, pm):
        n = C.c_int()
        pm = np.ascontiguousarray(pm, np.int32)
        _lib.check(_lib.lib().tdscha_qspace_static_psi_size(self.n_q, self.n_bands, _p(pm),
                                                            C.byref(n)),
                   "tdscha_qspace_static_psi_size")
        return n.value

    def apply_L_static(self, iq_pert, pair_maps, psis, ignore_harmonic=False, ignore_v3=False,
                       ignore_v4=False):
        """Static Liouvillian of the Hessian on psis in the static layout."""
        iqs = np.ascontiguousarray(iq_pert, np.int32)
        pms = np.ascontiguousarray(pair_maps, np.int32)
        for i, p in enumerate(psis):
            n = self.static_psi_size(pms[i])
            if len(p) != n:
                raise ValueError("chain %d: psi of %d entries, the static layout has %d"
                                 % (i, len(p), n))
        flat = np.ascontiguousarray(np.concatenate(psis), np.complex128)
        out = np.zeros_like(flat)
        _lib.check(_lib.lib().tdscha_qspace_apply_L_static(
            self.ctx, len(iqs), _p(iqs), _p(pms), _p(flat.view(np.float64)),
            _p(out.view(np.float64)), int(ignore_harmonic), int(ignore_v3), int(ignore_v4)),
            "tdscha_qspace_apply_L_static")
        res, pos = [], 0
        for p in psis:
            res.append(out[pos:pos + len(p)].copy())
            pos += len(p)
        return res

    def apply_L(self, iq_pert, pair_maps, psis, ignore_harmonic=False, ignore_v3=False,
                ignore_v4=False):
        iqs = np.ascontiguousarray(iq_pert, np.int32)
        pms = np.ascontiguousarray(pair_maps, np.int32)
        self._check_lengths(pms, psis)
        flat = np.ascontiguousarray(np.concatenate(psis), np.complex128)
        out = np.zeros_like(flat)
        _lib.check(_lib.lib().tdscha_qspace_apply_L(
            self.ctx, len(iqs), _p(iqs), _p(pms), _p(flat.view(np.float64)),
            _p(out.view(np.float64)), int(ignore_harmonic), int(ignore_v3), int(ignore_v4)),
            "tdscha_qspace_apply_L")
        res, pos = [], 0
        for p in psis:
            res.append(out[pos:pos + len(p)].copy())
            pos += len(p)
        return res

    def lanczos(self, iq_pert, pair_maps, psis, n_steps, caps=None, tol=EPSILON,
                ignore_harmonic=False, ignore_v3=False, ignore_v4=False, use_graph=True):
        iqs = np.ascontiguousarray(iq_pert, np.int32)
        pms = np.ascontiguousarray(pair_maps, np.int32)
        nc = len(iqs)
        self._check_lengths(pms, psis)
        flat = np.ascontiguousarray(np.concatenate(psis), np.complex128)
        a = np.zeros((nc, n_steps))
        b = np.zeros((nc, n_steps))
        c = np.zeros((nc, n_steps))
        nd = np.zeros(nc, np.int32)
        capv = None if caps is None else np.ascontiguousarray(caps, np.int32)
        ms = C.c_double(0.0)
        _lib.check(_lib.lib().tdscha_qspace_lanczos_run(
            self.ctx, nc, _p(iqs), _p(pms), _p(flat.view(np.float64)), int(n_steps),
            None if capv is None else _p(capv), float(tol), int(ignore_harmonic),
            int(ignore_v3), int(ignore_v4), int(bool(use_graph)), _p(a), _p(b), _p(c), _p(nd),
            C.byref(ms)), "tdscha_qspace_lanczos_run")
        chains = []
        for i in range(nc):
            k = int(nd[i])
            nb = k if (abs(b[i, k - 1]) >= tol) else k - 1
            chains.append((a[i, :k].copy(), b[i, :nb].copy(), c[i, :nb].copy()))
        return chains, ms.value

    def kpm(self, iq_pert, pair_maps, vms, vs, rescale, n_moments, n_have=0, moments=None,
            ignore_harmonic=False, ignore_v3=False, ignore_v4=False, use_graph=True):
        """Device Chebyshev recurrence of many chains (tdscha_qspace_kpm_run):
        returns (moments [C, n_moments], vms, vs, device ms)."""
        iqs = np.ascontiguousarray(iq_pert, np.int32)
        pms = np.ascontiguousarray(pair_maps, np.int32)
        nc = len(iqs)
        sizes = [len(x) for x in vs]
        if not (len(vms) == len(vs) == len(pms) == nc):
            raise ValueError("one pair map, vm and v per chain")
        for i in range(nc):
            n = self.psi_size(pms[i])
            if len(vms[i]) != n or len(vs[i]) != n:
                raise ValueError("chain %d: vectors of %d / %d entries, its layout has %d"
                                 % (i, len(vms[i]), len(vs[i]), n))
        if n_have and (moments is None or np.shape(moments)[0] != nc
                       or np.shape(moments)[1] < n_have):
            raise ValueError("continuing needs the first n_have moments of every chain")
        fvm = np.ascontiguousarray(np.concatenate(vms), np.complex128)
        fv = np.ascontiguousarray(np.concatenate(vs), np.complex128)
        res = np.ascontiguousarray(rescale, np.float64).reshape(nc, 2)
        mom = np.zeros((nc, n_moments))
        if n_have:
            mom[:, :n_have] = np.asarray(moments)[:, :n_have]
        ms = C.c_double(0.0)
        _lib.check(_lib.lib().tdscha_qspace_kpm_run(
            self.ctx, nc, _p(iqs), _p(pms), _p(res), int(n_moments), int(n_have), _p(mom),
            _p(fvm.view(np.float64)), _p(fv.view(np.float64)), int(ignore_harmonic),
            int(ignore_v3), int(ignore_v4), int(bool(use_graph)), C.byref(ms)),
            "tdscha_qspace_kpm_run")
        cut = np.cumsum([0] + sizes)
        return (mom, [fvm[cut[i]:cut[i + 1]].copy() for i in range(nc)],
                [fv[cut[i]:cut[i + 1]].copy() for i in range(nc)], ms.value)

    def set_tiles(self, weights_tile, d2v_chunk, reorth_jtile):
        _lib.check(_lib.lib().tdscha_qspace_set_tiles(self.ctx, int(weights_tile),
                                                      int(d2v_chunk), int(reorth_jtile)),
                   "tdscha_qspace_set_tiles")

    def set_d2v_form(self, form):
        _lib.check(_lib.lib().tdscha_qspace_set_d2v_form(self.ctx, int(form)),
                   "tdscha_qspace_set_d2v_form")

    def d2v_form(self):
        f = np.zeros(1, np.int32)
        _lib.check(_lib.lib().tdscha_qspace_get_d2v_form(self.ctx, _p(f)),
                   "tdscha_qspace_get_d2v_form")
        return int(f[0])

    def tiles(self):
        t = np.zeros(3, np.int32)
        _lib.check(_lib.lib().tdscha_qspace_get_tiles(self.ctx, _p(t)), "tdscha_qspace_get_tiles")
        return tuple(int(x) for x in t)

    def autotune(self, iq_pert, pair_maps, reps=3):
        iqs = np.ascontiguousarray(iq_pert, np.int32)
        pms = np.ascontiguousarray(pair_maps, np.int32)
        t = np.zeros(3, np.int32)
        ms = C.c_double(0.0)
        _lib.check(_lib.lib().tdscha_qspace_autotune(self.ctx, len(iqs), _p(iqs), _p(pms),
                                                     int(reps), _p(t), C.byref(ms)),
                   "tdscha_qspace_autotune")
        return tuple(int(x) for x in t), ms.value

    def close(self):
        if getattr(self, "ctx", None):
            _lib.lib().tdscha_qctx_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def wigner_mask(n_bands, unique_pairs):
    """Metric of the psi layout (QL:441-480): 1 on R, 2 on full pair blocks,
    1 / 2 on the diagonal / off-diagonal of the upper triangles."""
    nb = n_bands
    iu = np.triu_indices(nb)
    tri = np.where(iu[0] == iu[1], 1.0, 2.0)
    blocks = [np.full(nb * nb, 2.0) if a < b else tri for a, b in unique_pairs]
    return np.concatenate([np.ones(nb)] + blocks + blocks)


class QSpaceLanczos(object):
    """q-space Lanczos of a rock-salt ensemble (QSpaceLanczos, QL:116)."""

    def __init__(self, ensemble=None, modes=None, X_q=None, Y_q=None):
        self.ensemble = ensemble
        self.T = 0.0
        self.ignore_harmonic = False
        self.ignore_v3 = False
        self.ignore_v4 = False
        self.use_wigner = True
        self.psi = None
        self.a_coeffs, self.b_coeffs, self.c_coeffs = [], [], []
        self.perturbation_modulus = 1.0
        self.iq_pert = None
        self.q_pair_map = None
        self.unique_pairs = None
        self.engine = None
        self.initialized = False
        self.n_syms_qspace = 1
        if ensemble is None:
            return
        if X_q is None:
            modes, X_q, Y_q = bloch_ensemble(ensemble, modes)
        self.modes = modes
        self.L = modes.L
        self.T = float(ensemble.T)
        self.q_points = modes.grid
        self.n_q = modes.n_q
        self.n_bands = modes.n_bands
        self.w_q = modes.w_q
        self.pols_q = modes.eps
        self.valid_modes_q = modes.valid_modes_q
        self.X_q = X_q
        self.Y_q = Y_q
        self.rho = np.asarray(ensemble.rho, np.float64)
        self.N = len(self.rho)

    def init(self, use_symmetries=True):
        """QL:1658-1662."""
        self.prepare_symmetrization(no_sym=not use_symmetries)

    def prepare_symmetrization(self, no_sym=False, verbose=False, symmetries=None):
        """QL:1471-1633: upload the ensemble and, unless ``no_sym``, expand it
        into its point-group images (``symmetries`` = (perm, D), default the
        48 operations of O_h built by qspace.q_symmetries)."""
        if self.engine is not None:
            self.engine.close()
        self.engine = QSpaceEngine(self.X_q, self.Y_q, self.w_q, self.valid_modes_q, self.rho,
                                   self.T)
        self.n_syms_qspace = 1
        if not no_sym:
            perm, D = q_symmetries(self.modes) if symmetries is None else symmetries
            self.engine.symmetrize(perm, D)
            self.n_syms_qspace = len(perm)
            if verbose:
                print("Q-space: %d point-group images of %d configurations"
                      % (self.n_syms_qspace, self.N))
        self.initialized = True

    # ---------------------------------------------------------- layout
    def build_q_pair_map(self, iq_pert):
        """QL:297-331."""
        self.iq_pert = int(iq_pert)
        self.q_pair_map, self.unique_pairs = _pair_map(self.L, self.iq_pert)

    def get_psi_size(self):
        if self.q_pair_map is None:
            raise ValueError("Must call build_q_pair_map first")
        nb = self.n_bands
        return nb + 2 * sum(nb * nb if a < b else nb * (nb + 1) // 2
                            for a, b in self.unique_pairs)

    def mask_dot_wigner(self, debug=False):
        """QL:441-480."""
        return wigner_mask(self.n_bands, self.unique_pairs)

    def reset_q(self):
        self.psi = np.zeros(self.get_psi_size(), np.complex128)
        self.a_coeffs, self.b_coeffs, self.c_coeffs = [], [], []

    def prepare_mode_q(self, iq, band_index):
        """QL:1186-1205."""
        if band_index < 0 or band_index >= self.n_bands:
            raise ValueError("Invalid band index for perturbation: {}".format(band_index))
        self.build_q_pair_map(iq)
        self.reset_q()
        self.psi[band_index] = 1.0
        self.perturbation_modulus = 1.0

    def prepare_perturbation_q(self, iq, vector, add=False):
        """QL:1406-1432: project a Cartesian unit-cell vector (3 n_at) on the
        bands at iq, with the 1 / sqrt(m) mass scaling."""
        if not add:
            self.build_q_pair_map(iq)
            self.reset_q()
        m = self.ensemble.model.m[:self.n_bands]
        R1 = np.conj(self.pols_q[iq]).T @ (np.asarray(vector) / np.sqrt(m))
        self.psi[:self.n_bands] += R1
        self.perturbation_modulus = float(np.real(np.conj(R1) @ R1))

    # ---------------------------------------------------------- L and runs
    def apply_full_L(self, target=None, **kwargs):
        if target is not None:
            self.psi = np.array(target, np.complex128)
        return self.engine.apply_L([self.iq_pert], [self.q_pair_map], [self.psi],
                                   self.ignore_harmonic, self.ignore_v3, self.ignore_v4)[0]

    def run_FT(self, n_iter, verbose=False, reorthogonalize=True, tol=EPSILON, use_graph=True):
        """QL:967-1182 (reorthogonalize=True), on the device."""
        if not reorthogonalize:
            raise NotImplementedError("the device q-space Lanczos reorthogonalises fully")
        (a, b, c), = self.engine.lanczos([self.iq_pert], [self.q_pair_map], [self.psi], n_iter,
                                         tol=tol, ignore_harmonic=self.ignore_harmonic,
                                         ignore_v3=self.ignore_v3, ignore_v4=self.ignore_v4,
                                         use_graph=use_graph)[0]
        self.a_coeffs = list(a)
        self.b_coeffs = list(b)
        self.c_coeffs = list(c)
        return a, b

    def run_FT_batched(self, iq_perts, psis, n_steps, caps=None, tol=EPSILON, use_graph=True):
        """Many chains, each with its own perturbation q-point, in one device
        batch.  Returns [(a, b, c)] and the device time (ms)."""
        pms = [_pair_map(self.L, int(iq))[0] for iq in iq_perts]
        return self.engine.lanczos(iq_perts, pms, psis, n_steps, caps=caps, tol=tol,
                                   ignore_harmonic=self.ignore_harmonic,
                                   ignore_v3=self.ignore_v3, ignore_v4=self.ignore_v4,
                                   use_graph=use_graph)

    def get_green_function_continued_fraction(self, w_array, use_terminator=True,
                                              last_average=1, smearing=0.0, smooth_ramp=0):
        return continued_fraction(self.a_coeffs, self.b_coeffs, self.c_coeffs, w_array, True,
                                  use_terminator, last_average, smearing, smooth_ramp, False, 0.0,
                                  self.perturbation_modulus)

    def close(self):
        if self.engine is not None:
            self.engine.close()
            self.engine = None
