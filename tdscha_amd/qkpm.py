"""Kernel polynomial method on the q-space Liouvillian (Modules/QSpaceKPM.py,
class QSpaceKPM).

Chebyshev moments mu_n = <v0, T_n(L~) v0>_M of the rescaled operator
L~ = (L - b) / a, with v0 the normalised perturbation and M the Wigner
metric (mask_dot_wigner), two moments per L application through
mu_{2n} = 2 <t_n, t_n>_M - mu_0 and mu_{2n+1} = 2 <t_{n+1}, t_n>_M - mu_1
(QSK:260-454), and the spectral function from the damped Chebyshev series
(QSK:536-559).  The recurrence runs in the gfx950 library: every step (L
application, Chebyshev combination, the two metric products) on the device,
one graph launch per run, for one perturbation or a batch of them
(run_KPM_batched).
"""
from __future__ import annotations

import os

import numpy as np

from .qlanczos import EPSILON, QSpaceLanczos, wigner_mask

RY_TO_CM = 109737.31568527

_KPM_ATTRS = ("kpm_moments", "kpm_vector_norm", "kpm_n_moments", "kpm_lambda_min",
              "kpm_lambda_max", "kpm_rescale_a", "kpm_rescale_b", "kpm_bound_factor",
              "kpm_chebyshev_vm", "kpm_chebyshev_v")


def jackson_kernel(n_moments):
    """Jackson damping factors g_n, n = 0 .. N-1 (QSK:515-519; Weisse et al.,
    Rev. Mod. Phys. 78, 275 (2006), Eq. 71)."""
    n = np.arange(n_moments, dtype=np.float64)
    phi = np.pi / (n_moments + 1)
    return ((n_moments - n + 1) * np.cos(n * phi) + np.sin(n * phi) / np.tan(phi)) / (n_moments + 1)


def kpm_bounds(w_q, valid_modes_q, use_wigner, lambda_min=None, lambda_max=None,
               bound_factor=2.5, edge_buffer=1e-8):
    """Eigenvalue window of L (QSK:156-176): [-(2 w_max)^2, 0] in the Wigner
    representation, [0, (2 w_max)^2] otherwise, widened by (bound_factor - 1)
    of its width, then by edge_buffer."""
    if (lambda_min is None) != (lambda_max is None):
        raise ValueError("lambda_min and lambda_max must be both specified or both omitted")
    if lambda_min is None:
        max_w = np.max(np.abs(np.asarray(w_q)[np.asarray(valid_modes_q, bool)]))
        width = (2 * max_w) ** 2
        margin = (bound_factor - 1) * width
        lmin, lmax = (-width - margin, margin) if use_wigner else (-margin, width + margin)
        span = lmax - lmin
        return lmin - edge_buffer * span, lmax + edge_buffer * span
    if lambda_min >= lambda_max:
        raise ValueError("Invalid KPM bounds: lambda_min must be smaller than lambda_max")
    span = lambda_max - lambda_min
    return lambda_min - edge_buffer * span, lambda_max + edge_buffer * span


def chebyshev_spectrum(moments, w_array, rescale_a, rescale_b, vector_norm, use_wigner, g):
    """pi rho(lambda(w)) of the damped Chebyshev series (QSK:536-559)."""
    w_array = np.asarray(w_array, dtype=np.float64)
    lam = -(w_array ** 2) if use_wigner else w_array ** 2
    x = (lam - rescale_b) / rescale_a
    inside = np.abs(x) < 1
    out = np.zeros_like(w_array)
    if not inside.any():
        return out
    xin = np.clip(x[inside], -1 + 1e-15, 1 - 1e-15)
    n = len(moments)
    t0 = np.ones_like(xin)
    rho = g[0] * moments[0] * t0
    if n > 1:
        t1 = xin.copy()
        rho += 2 * g[1] * moments[1] * t1
        for i in range(2, n):
            t2 = 2 * xin * t1 - t0
            rho += 2 * g[i] * moments[i] * t2
            t0, t1 = t1, t2
    rho /= np.pi * rescale_a * np.sqrt(1 - xin ** 2)
    rho *= vector_norm ** 2
    out[inside] = np.pi * rho
    return out


class QSpaceKPM(QSpaceLanczos):
    """Kernel polynomial method using the q-space Liouvillian (QSK:17)."""

    def __init__(self, ensemble=None, modes=None, X_q=None, Y_q=None):
        super().__init__(ensemble, modes=modes, X_q=X_q, Y_q=Y_q)
        self._init_kpm_attributes()

    def _init_kpm_attributes(self):
        """QSK:56-66."""
        for k in _KPM_ATTRS:
            setattr(self, k, None)
        self.kpm_n_moments = 0

    _invalidate_kpm_cache = _init_kpm_attributes       # QSK:68-78

    def reset_q(self):
        super().reset_q()
        self._invalidate_kpm_cache()

    def _metric_dot(self, v1, v2, mask=None):
        if mask is None:
            mask = self.mask_dot_wigner(False)
        return float(np.real(np.conj(v1).dot(v2 * mask)))

    def _get_kpm_bounds(self, lambda_min=None, lambda_max=None, bound_factor=2.5,
                        edge_buffer=1e-8):
        return kpm_bounds(self.w_q, self.valid_modes_q, self.use_wigner, lambda_min, lambda_max,
                          bound_factor, edge_buffer)

    def estimate_kpm_steps(self, precision_cm, bound_factor=1.2, regularization="jackson"):
        """QSK:178-255: moments needed to resolve precision_cm at the lowest
        frequency of the perturbation's q-point, n = pi a / (2 w_min dw)."""
        if precision_cm <= 0:
            raise ValueError("precision_cm must be positive")
        if bound_factor <= 1.0:
            raise ValueError("bound_factor must be > 1.0")
        if self.iq_pert is None:
            raise ValueError("Must prepare a perturbation before estimating KPM steps")
        w_at_q = self.w_q[:, self.iq_pert]
        valid = np.asarray(self.valid_modes_q, bool)[:, self.iq_pert]
        w_min = np.min(np.abs(w_at_q[valid]))
        if w_min < EPSILON:
            raise ValueError("Smallest frequency at q-point is effectively zero; cannot estimate "
                             "steps")
        lmin, lmax = self._get_kpm_bounds(bound_factor=bound_factor, edge_buffer=1e-8)
        a = 0.5 * (lmax - lmin)
        n = int(np.ceil(np.pi * a / (2.0 * w_min * precision_cm / RY_TO_CM)))
        return max(n, 8)

    def _apply_rescaled_L(self, vec):
        return (self.apply_full_L(vec) - self.kpm_rescale_b * vec) / self.kpm_rescale_a

    def run_KPM(self, n_moments, lambda_min=None, lambda_max=None, bound_factor=1.2,
                edge_buffer=1e-8, verbose=True, use_graph=True):
        """QSK:260-454, including continuation from saved recurrence vectors
        (odd, even or single-moment restarts).  The recurrence runs on the
        device (tdscha_qspace_kpm_run), in one graph launch."""
        if n_moments < 1:
            raise ValueError("n_moments must be positive")
        if self.psi is None:
            raise ValueError("Prepare a perturbation before running KPM")
        existing = self.kpm_n_moments
        can_continue = (existing > 0 and self.kpm_moments is not None
                        and self.kpm_chebyshev_v is not None)
        if can_continue and n_moments <= existing:
            if verbose:
                print("KPM: already have {} moments, {} requested. Nothing to do.".format(
                    existing, n_moments))
            return
        if can_continue:
            vm = np.array(self.kpm_chebyshev_vm, np.complex128)
            v = np.array(self.kpm_chebyshev_v, np.complex128)
            have = existing
            prev = np.asarray(self.kpm_moments)[None, :existing]
        else:
            mask = self.mask_dot_wigner(False)
            norm_sq = self._metric_dot(self.psi, self.psi, mask)
            if norm_sq <= EPSILON or np.isnan(norm_sq):
                raise ValueError("Prepare a non-zero perturbation before running KPM")
            self.kpm_lambda_min, self.kpm_lambda_max = self._get_kpm_bounds(
                lambda_min, lambda_max, bound_factor, edge_buffer)
            self.kpm_rescale_a = 0.5 * (self.kpm_lambda_max - self.kpm_lambda_min)
            self.kpm_rescale_b = 0.5 * (self.kpm_lambda_max + self.kpm_lambda_min)
            self.kpm_vector_norm = np.sqrt(norm_sq)
            self.kpm_bound_factor = bound_factor
            vm = v = np.asarray(self.psi, np.complex128) / self.kpm_vector_norm
            have, prev = 0, None
        mom, vms, vs, ms = self.engine.kpm(
            [self.iq_pert], [self.q_pair_map], [vm], [v],
            [self.kpm_rescale_a, self.kpm_rescale_b], n_moments, have, prev,
            self.ignore_harmonic, self.ignore_v3, self.ignore_v4, use_graph)
        self.kpm_n_moments = int(n_moments)
        self.kpm_moments = mom[0]
        self.kpm_chebyshev_vm, self.kpm_chebyshev_v = vms[0], vs[0]
        if verbose:
            print("KPM completed: {} moments in {:.2f} ms on the device".format(n_moments, ms))

    def run_KPM_batched(self, iq_perts, psis, n_moments, lambda_min=None, lambda_max=None,
                        bound_factor=1.2, edge_buffer=1e-8, use_graph=True):
        """Moments of many perturbations (each at its own q-point) in one
        device run.  Returns (moments [C, n_moments], norms, (rescale_a,
        rescale_b), device ms); the spectra follow from chebyshev_spectrum."""
        from .qspace import pair_map
        lmin, lmax = self._get_kpm_bounds(lambda_min, lambda_max, bound_factor, edge_buffer)
        a, b = 0.5 * (lmax - lmin), 0.5 * (lmax + lmin)
        pms, v0s, norms = [], [], []
        for iq, psi in zip(iq_perts, psis):
            pm, pairs = pair_map(self.L, int(iq))
            mask = wigner_mask(self.n_bands, pairs)
            psi = np.asarray(psi, np.complex128)
            if len(psi) != len(mask):
                raise ValueError("psi of q-point %d has %d entries, its layout %d"
                                 % (iq, len(psi), len(mask)))
            nrm = np.sqrt(float(np.real(np.conj(psi).dot(psi * mask))))
            if nrm <= np.sqrt(EPSILON):
                raise ValueError("Prepare a non-zero perturbation before running KPM")
            pms.append(pm)
            v0s.append(psi / nrm)
            norms.append(nrm)
        mom, _, _, ms = self.engine.kpm(list(iq_perts), pms, v0s, v0s, [(a, b)] * len(pms),
                                        n_moments, 0, None, self.ignore_harmonic,
                                        self.ignore_v3, self.ignore_v4, use_graph)
        return mom, np.array(norms), (a, b), ms

    def save_kpm(self, file):
        """QSK:456-471: moments with the parameters in a one-line header."""
        if self.kpm_moments is None:
            raise ValueError("Run run_KPM before saving")
        header = ("kpm_n_moments={} kpm_vector_norm={} kpm_lambda_min={} kpm_lambda_max={} "
                  "kpm_rescale_a={} kpm_rescale_b={}").format(
            self.kpm_n_moments, self.kpm_vector_norm, self.kpm_lambda_min, self.kpm_lambda_max,
            self.kpm_rescale_a, self.kpm_rescale_b)
        np.savetxt(file, self.kpm_moments, header=header)

    def load_kpm(self, file):
        """QSK:473-513."""
        with open(file) as fp:
            head = fp.readline()
        if not head.startswith("#"):
            raise ValueError("Invalid KPM save file: missing header")
        params = {}
        for part in head[1:].split():
            if "=" in part:
                k, v = part.split("=")
                params[k] = float(v)
        self.kpm_n_moments = int(params["kpm_n_moments"])
        self.kpm_vector_norm = params["kpm_vector_norm"]
        self.kpm_lambda_min = params["kpm_lambda_min"]
        self.kpm_lambda_max = params["kpm_lambda_max"]
        self.kpm_rescale_a = params["kpm_rescale_a"]
        self.kpm_rescale_b = params["kpm_rescale_b"]
        self.kpm_moments = np.atleast_1d(np.loadtxt(file))

    def save_status(self, file):
        """QSK:80-111: KPM state with the recurrence vectors, for continuation."""
        if self.kpm_moments is None:
            raise ValueError("Run run_KPM before saving status")
        if ".npz" not in file.lower():
            file += ".npz"
        d = {k: getattr(self, k) for k in _KPM_ATTRS[:8]}
        for k in ("kpm_chebyshev_vm", "kpm_chebyshev_v"):
            if getattr(self, k) is not None:
                d[k] = getattr(self, k)
        np.savez_compressed(file, **d)

    def load_status(self, file):
        """QSK:113-145 (without pickle)."""
        if ".npz" not in file.lower():
            file += ".npz"
        if not os.path.exists(file):
            raise IOError("KPM status file not found: {}".format(file))
        with np.load(file, allow_pickle=False) as d:
            data = {k: d[k] for k in d.files}
        self.kpm_n_moments = int(data["kpm_n_moments"])
        for k in ("kpm_vector_norm", "kpm_lambda_min", "kpm_lambda_max", "kpm_rescale_a",
                  "kpm_rescale_b", "kpm_bound_factor"):
            setattr(self, k, float(data[k]))
        self.kpm_moments = data["kpm_moments"]
        self.kpm_chebyshev_vm = data.get("kpm_chebyshev_vm")
        self.kpm_chebyshev_v = data.get("kpm_chebyshev_v")

    _jackson_kernel = staticmethod(jackson_kernel)

    def _get_kpm_damping(self, regularization="jackson", damping_factors=None):
        """QSK:521-534."""
        if self.kpm_moments is None:
            raise ValueError("Run run_KPM before requesting the spectral function")
        if damping_factors is not None:
            g = np.asarray(damping_factors, dtype=np.float64)
        elif regularization is None or str(regularization).lower() == "none":
            g = np.ones(self.kpm_n_moments)
        elif str(regularization).lower() == "jackson":
            g = jackson_kernel(self.kpm_n_moments)
        else:
            raise ValueError("Unknown KPM regularization: {}".format(regularization))
        if len(g) != self.kpm_n_moments:
            raise ValueError("Damping factors must have length {}".format(self.kpm_n_moments))
        return g

    def get_spectral_function_KPM(self, w_array, regularization="jackson", damping_factors=None):
        """QSK:536-559."""
        g = self._get_kpm_damping(regularization, damping_factors)
        return chebyshev_spectrum(self.kpm_moments, w_array, self.kpm_rescale_a,
                                  self.kpm_rescale_b, self.kpm_vector_norm, self.use_wigner, g)

    @classmethod
    def from_qspace_lanczos(cls, qlanc):
        """QSK:561-631: a KPM object sharing the data and device context of an
        initialised QSpaceLanczos."""
        kpm = cls(None)
        for k, v in qlanc.__dict__.items():
            setattr(kpm, k, v)
        kpm._invalidate_kpm_cache()
        return kpm
