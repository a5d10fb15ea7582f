"""CPU oracle for the TD-SCHA Lanczos linear-response path (numpy, FP64).

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import this module, and only as the
reference for comparison.  The product package ``tdscha_amd`` never imports it.

Each function restates one piece of the reference algorithm (SSCHAcode/tdscha)
in plain numpy; the docstring of every function cites the reference file:line
it follows.  Abbreviations: DL = Modules/DynamicalLanczos.py,
LF = CModules/LanczosFunctions.c, JC = Modules/tdscha_core.jl.

Pins (what checks each function, see tests/test_oracle_*.py):
  * bose/f_ups/f_psi ............ closed form T=0 limits + LF via oracle/_ref
  * perturb_averages ............ oracle/_ref (the reference C kernel compiled
                                  from its own sources), dense explicit-tensor L
                                  at tiny n, LANCZOS_STEP10.npz golden L.q_k
  * harmonic / wigner / standard  dense L (DL:5167-5900 restated), npz golden
  * translation projection ...... full-group vs point-group+projector equality
  * run_FT (biconjugate) ........ LANCZOS_STEP10.npz coefficients and bases
  * run_FT (symmetric Wigner) ... b_n == c_n, M-orthonormal basis, harmonic
                                  single-mode closed form
  * continued_fraction .......... tests/test_continued_fraction.py closed forms
"""
from __future__ import annotations

import numpy as np

RY_TO_K = 157887.32400374097          # DL:134, LF header, JC:24
EPSILON = 1e-12                        # DL:127 (__EPSILON__)


# --------------------------------------------------------------------------
# Thermal factors
# --------------------------------------------------------------------------
def bose(w, T):
    """Bose occupation n(w) with w in Ry, T in K (DL:158-163)."""
    w = np.asarray(w, dtype=np.float64)
    if T > 0:
        return 1.0 / (np.exp(w * RY_TO_K / T) - 1.0)
    return np.zeros_like(w)


def f_ups(w, T):
    """Eigenvalue of the inverse covariance, 2w/(1+2n) (LF:14-21)."""
    return 2.0 * w / (1.0 + 2.0 * bose(w, T))


def f_psi(w, T):
    """Eigenvalue of the covariance, (1+2n)/(2w) (LF:24-31)."""
    return (1.0 + 2.0 * bose(w, T)) / (2.0 * w)


# --------------------------------------------------------------------------
# Symmetry images (block-diagonal mode-space symmetries)
# --------------------------------------------------------------------------
def dense_symmetries(n_modes, block_syms, block_modes):
    """Assemble S_j (N_sym, n, n) from the per-block matrices.

    ``block_syms[b]`` is f64[N_sym, d_b, d_b]; ``block_modes[b]`` the mode
    indices of block b.  Row mu_id / column k of block b act as
    u[mode[mu_id]] += S_b[j, mu_id, k] * x[mode[k]]  (LF:2158-2170, JC:40-59).
    """
    n_syms = block_syms[0].shape[0]
    S = np.zeros((n_syms, n_modes, n_modes))
    for sb, modes in zip(block_syms, block_modes):
        modes = np.asarray(modes)
        S[:, modes[:, None], modes[None, :]] = sb
    return S


def images(X, S):
    """Symmetry images u_ij = S_j x_i in reference order (row = i*N_sym + j).

    The big index ordering follows JC:82-83 (bigindex-1 = (i-1)*N_sym + (j-1)).
    """
    N, n = X.shape
    out = np.einsum("jab,ib->ija", S, X)
    return out.reshape(N * S.shape[0], n)


# --------------------------------------------------------------------------
# The perturbed averages (the hot kernel)
# --------------------------------------------------------------------------
def perturb_averages(X, Y, w, rho, R1, Y1, T, apply_d4, S):
    """Ensemble-averaged D3/D4 application, FP64.

    Restates get_f_average_from_Y_pert_sym_fast (LF:2082-2254),
    get_d2v_dR2_from_R_pert_sym_fast (LF:2257-2438) and
    get_d2v_dR2_from_Y_pert_sym_fast (LF:2615-2754); identical algebra in
    JC:64-208.  For every image (u, f) = (S_j x_i, S_j y_i):

        a = f_Y * u, g = f_psi * f
        f_av  += rho_i * [ -(u.Y1.u)/6 * f - (u.Y1.g + g.Y1.u)/6 * a ]
        d2v   -= rho_i * (a.R1)/3 * (a f^T + f a^T) + rho_i * (f.R1)/3 * a a^T
        d2v   += d4 * rho_i * [ (u.Y1.u)/8 (a f^T + f a^T)
                                + (g.Y1.u + u.Y1.g)/8 a a^T ]

    normalised by N_eff * N_sym.  ``S`` is the dense (N_sym, n, n) stack.
    Returns (f_pert_av f64[n], d2v_pert_av f64[n, n]).
    """
    X = np.asarray(X, np.float64)
    Y = np.asarray(Y, np.float64)
    n_syms = S.shape[0]
    fy = f_ups(w, T)
    fp = f_psi(w, T)
    n_eff = np.sum(rho)
    n = X.shape[1]
    f_av = np.zeros(n)
    d2v = np.zeros((n, n))
    for j in range(n_syms):
        U = X @ S[j].T
        F = Y @ S[j].T
        A = U * fy
        G = F * fp
        V = U @ Y1
        s1 = np.einsum("ia,ia->i", V, U)                  # u.Y1.u
        s2 = 0.5 * (np.einsum("ia,ia->i", V, G)           # (u.Y1.g + g.Y1.u)/2
                    + np.einsum("ia,ia->i", G @ Y1, U))
        s3 = A @ R1
        s4 = F @ R1
        f_av += (rho * (-s1 / 6.0)) @ F + (rho * (-s2 / 3.0)) @ A
        c1 = rho * (-s3 / 3.0 + apply_d4 * s1 / 8.0)
        c2 = rho * (-s4 / 3.0 + apply_d4 * s2 / 4.0)
        Gm = A.T @ (c1[:, None] * F)
        d2v += Gm + Gm.T + A.T @ (c2[:, None] * A)
    norm = n_eff * n_syms
    return f_av / norm, d2v / norm


# --------------------------------------------------------------------------
# psi layout helpers
# --------------------------------------------------------------------------
def upper_pairs(n):
    """Independent (row, col) pairs of the packed upper triangle.

    Row-major order over rows, col >= row; w_a = w[col], w_b = w[row]
    (DL:2339-2346: i_a is the column index, i_b the row index).
    """
    rows, cols = np.triu_indices(n)
    return rows, cols


def len_psi(n):
    return n + n * (n + 1)


def unpack_sym(vec, n):
    """Packed upper triangle (row-major) -> full symmetric matrix (DL:2858-2866)."""
    M = np.zeros((n, n))
    r, c = upper_pairs(n)
    M[r, c] = vec
    M[c, r] = vec
    return M


def pack_upper(M):
    r, c = upper_pairs(M.shape[0])
    return M[r, c].copy()


def _pair_quantities(w, T):
    n = len(w)
    r, c = upper_pairs(n)
    w_a = w[c]
    w_b = w[r]
    if T > 0:
        n_a = 1.0 / (np.exp(w_a / (T / RY_TO_K)) - 1.0)
        n_b = 1.0 / (np.exp(w_b / (T / RY_TO_K)) - 1.0)
    else:
        n_a = np.zeros_like(w_a)
        n_b = np.zeros_like(w_b)
    return w_a, w_b, n_a, n_b


# --------------------------------------------------------------------------
# Harmonic part
# --------------------------------------------------------------------------
def harmonic_L1(psi, w, T, use_wigner, transpose=False, ignore_harmonic=False):
    """apply_L1_FT (DL:2287-2445)."""
    n = len(w)
    out = np.zeros_like(psi)
    if ignore_harmonic:
        return out
    w_a, w_b, n_a, n_b = _pair_quantities(w, T)
    nw2 = len(w_a)
    sY, sA = n, n + nw2
    if not use_wigner:
        out[:n] = psi[:n] * w * w
        X_ab = -w_a ** 2 - w_b ** 2 - (2 * w_a * w_b) / ((2 * n_a + 1) * (2 * n_b + 1))
        out[sY:sA] = -X_ab * psi[sY:sA]
        Y_ab = -(8 * w_a * w_b) / ((2 * n_a + 1) * (2 * n_b + 1))
        if not transpose:
            out[sY:sA] += -Y_ab * psi[sA:]
        else:
            out[sA:] += -Y_ab * psi[sY:sA]
        X1_ab = -(2 * n_a * n_b + n_a + n_b) * (2 * n_a * n_b + n_a + n_b + 1) \
            * (2 * w_a * w_b) / ((2 * n_a + 1) * (2 * n_b + 1))
        if not transpose:
            out[sA:] += -X1_ab * psi[sY:sA]
        else:
            out[sY:sA] += -X1_ab * psi[sA:]
        Y1_ab = -w_a ** 2 - w_b ** 2 + (2 * w_a * w_b) / ((2 * n_a + 1) * (2 * n_b + 1))
        out[sA:] += -Y1_ab * psi[sA:]
    else:
        out[:n] = -(psi[:n] * w) * w
        out[sY:sA] = -(w_a ** 2 + w_b ** 2 - 2.0 * w_a * w_b) * psi[sY:sA]
        out[sA:] = -(w_a ** 2 + w_b ** 2 + 2.0 * w_a * w_b) * psi[sA:]
    return out


def harmonic_L1_inverse(psi, w, T, transpose=False, ignore_harmonic=False):
    """apply_L1_inverse_FT (DL:2451-2549), the preconditioner of M_linop:
    the inverse of the standard-representation harmonic blocks with the
    reference's regularisation (populations zeroed where 0.01 w_a RY_TO_K > T
    for both members, DL:2491-2498; determinant below 1e-12 keeps -1/X only,
    DL:2522-2531; equal thermally populated frequencies invert Y1 alone,
    DL:2533-2536)."""
    n = len(w)
    out = np.zeros_like(psi)
    if ignore_harmonic:
        return out
    r, c = upper_pairs(n)
    w_a, w_b = w[c], w[r]
    n_a = np.zeros_like(w_a)
    n_b = np.zeros_like(w_a)
    if T > 0:
        keep = ~(0.01 * w_a * RY_TO_K > T)
        n_a = np.where(keep, 1.0 / (np.exp(w_a / (T / RY_TO_K)) - 1.0), 0.0)
        n_b = np.where(keep, 1.0 / (np.exp(w_b / (T / RY_TO_K)) - 1.0), 0.0)
    out[:n] = (psi[:n] / w) / w
    d = (2 * n_a + 1) * (2 * n_b + 1)
    X = -w_a ** 2 - w_b ** 2 - (2 * w_a * w_b) / d
    Y = -(8 * w_a * w_b) / d
    X1 = -(2 * n_a * n_b + n_a + n_b) * (2 * n_a * n_b + n_a + n_b + 1) * (2 * w_a * w_b) / d
    Y1 = -w_a ** 2 - w_b ** 2 + (2 * w_a * w_b) / d
    den = X * Y1 - X1 * Y
    reg = den < EPSILON
    safe = np.where(reg, 1.0, den)
    Xn = np.where(reg, -1.0 / X, -Y1 / safe)
    Yn = np.where(reg, 0.0, Y / safe)
    X1n = np.where(reg, 0.0, X1 / safe)
    Y1n = np.where(reg, 0.0, -X / safe)
    if T > EPSILON:
        Y1n = np.where((w_a == w_b) & (n_a > EPSILON), -1.0 / Y1, Y1n)
    nw2 = len(w_a)
    sY, sA = n, n + nw2
    out[sY:sA] = Xn * psi[sY:sA]
    if not transpose:
        out[sY:sA] += Yn * psi[sA:]
        out[sA:] += X1n * psi[sY:sA]
    else:
        out[sA:] += Yn * psi[sY:sA]
        out[sY:sA] += X1n * psi[sA:]
    out[sA:] += Y1n * psi[sA:]
    return out


# --------------------------------------------------------------------------
# Wigner change of variables
# --------------------------------------------------------------------------
def chi_pm(w, T):
    """(chi_minus, chi_plus) full n x n matrices (DL:2632-2687)."""
    n = len(w)
    W = np.tile(w, (n, 1))
    if T > 1e-12:
        nn = 1.0 / (np.exp(W * 157887.32400374097 / T) - 1.0)
    else:
        nn = np.zeros((n, n))
    chi_m = (W - W.T) * (nn - nn.T) / (2.0 * W * W.T)
    chi_p = (W + W.T) * (1 + nn + nn.T) / (2.0 * W * W.T)
    return chi_m, chi_p


def wigner_coefficients(w, T):
    """Per-pair coefficients of the Wigner map (DL:2841-2853, 3326-3330).

    Y1_pair = ca * a'_pair - cb * b'_pair, with
    ca = (w_a w_b / X) sqrt(-chi^-/2), cb = (w_a w_b / X) sqrt(chi^+/2),
    X = (1+2n_a)(1+2n_b)/8.  Output: a' = +sqrt(-chi^-/2) d2v,
    b' = -sqrt(chi^+/2) d2v.
    """
    w_a, w_b, n_a, n_b = _pair_quantities(w, T)
    Xc = (1 + 2 * n_a) * (1 + 2 * n_b) / 8
    w2x = (w_a * w_b) / Xc
    chi_m = ((w_a - w_b) * (n_a - n_b)) / (2 * w_a * w_b)
    chi_p = ((w_a + w_b) * (1 + n_a + n_b)) / (2 * w_a * w_b)
    sm = np.sqrt(-0.5 * chi_m)
    sp = np.sqrt(0.5 * chi_p)
    return w2x * sm, w2x * sp, sm, sp


def wigner_Y1(psi, w, T):
    """alpha^(1) = Upsilon^(1) from (a', b') (DL:2773-2882, get_alpha=True)."""
    n = len(w)
    nw2 = n * (n + 1) // 2
    ca, cb, _, _ = wigner_coefficients(w, T)
    vec = ca * psi[n:n + nw2] - cb * psi[n + nw2:]
    return unpack_sym(vec, n)


def standard_Y1(psi, n, half_off_diagonal=False):
    """get_Y1 (DL:2887-2917)."""
    nw2 = n * (n + 1) // 2
    Y1 = unpack_sym(psi[n:n + nw2], n)
    if half_off_diagonal:
        mask = np.full((n, n), 0.5)
        np.fill_diagonal(mask, 1.0)
        Y1 *= mask
    return Y1


def standard_ReA1(psi, n, half_off_diagonal=False):
    """get_ReA1 (DL:2920-2948)."""
    nw2 = n * (n + 1) // 2
    A1 = unpack_sym(psi[n + nw2:], n)
    if half_off_diagonal:
        mask = np.full((n, n), 0.5)
        np.fill_diagonal(mask, 1.0)
        A1 *= mask
    return A1


def standard_eigs(w, T):
    """Equilibrium Upsilon and ReA eigenvalues (DL:3026-3032)."""
    nm = 0
    if T > EPSILON:
        nm = 1.0 / (np.exp(w * 157887.32400374097 / T) - 1.0)
    Y_w = 2 * w / (2 * nm + 1)
    ReA_w = 2 * w * nm * (nm + 1) / (2 * nm + 1)
    return Y_w, ReA_w


# --------------------------------------------------------------------------
# Translational projection for the gamma-only mode
# --------------------------------------------------------------------------
def translation_project(f, d2v, pols, trans_projector, inv_perms):
    """Impose translations on the perturbed averages (DL:3174-3191)."""
    f = trans_projector @ f
    d_cart = pols @ d2v @ pols.T
    acc = np.zeros_like(d_cart)
    for p in inv_perms:
        acc += d_cart[np.ix_(p, p)]
    acc /= len(inv_perms)
    return f, pols.T @ acc @ pols


# --------------------------------------------------------------------------
# Anharmonic part and the full L
# --------------------------------------------------------------------------
class OracleSystem:
    """Everything apply_full_L needs, as plain arrays.

    Mirrors the attributes of the reference Lanczos object that the L
    application reads (DL:183-452, 694-871).
    """

    def __init__(self, X, Y, w, rho, T, S=None, use_wigner=True,
                 ignore_harmonic=False, ignore_v3=False, ignore_v4=False,
                 reverse_L=False, shift_value=0.0, pols=None,
                 trans_projector=None, inv_perms=None):
        self.X = np.ascontiguousarray(X, np.float64)
        self.Y = np.ascontiguousarray(Y, np.float64)
        self.w = np.asarray(w, np.float64)
        self.rho = np.asarray(rho, np.float64)
        self.T = float(T)
        n = len(self.w)
        self.n = n
        self.S = np.eye(n)[None] if S is None else np.asarray(S)
        self.use_wigner = use_wigner
        self.ignore_harmonic = ignore_harmonic
        self.ignore_v3 = ignore_v3
        self.ignore_v4 = ignore_v4
        self.reverse_L = reverse_L
        self.shift_value = shift_value
        self.pols = pols
        self.trans_projector = trans_projector
        self.inv_perms = inv_perms

    # DL:2951-3363
    def anharmonic(self, psi, transpose=False):
        n = self.n
        R1 = psi[:n].copy()
        if not self.use_wigner:
            Y1 = standard_Y1(psi, n, half_off_diagonal=transpose)
            Y_w, ReA_w = standard_eigs(self.w, self.T)
            if transpose:
                ReA1 = standard_ReA1(psi, n, half_off_diagonal=transpose)
                cY = 2 * (np.einsum("a,b,b->ab", Y_w, Y_w, Y_w)
                          + np.einsum("a,a,b->ab", Y_w, Y_w, Y_w))
                cRA = 2 * (np.einsum("a,b,b->ab", Y_w, ReA_w, Y_w)
                           + np.einsum("a,a,b->ab", Y_w, ReA_w, Y_w))
                Y1 = Y1 * cY + ReA1 * cRA
        else:
            Y1 = wigner_Y1(psi, self.w, self.T)
        apply_d4 = 0 if self.ignore_v4 else 1
        if self.ignore_v3:
            R1[:] = 0.0
        f_av, d2v = perturb_averages(self.X, self.Y, self.w, self.rho, R1, Y1,
                                     self.T, apply_d4, self.S)
        if self.trans_projector is not None:
            f_av, d2v = translation_project(f_av, d2v, self.pols,
                                            self.trans_projector, self.inv_perms)
        out = np.zeros_like(psi)
        out[:n] = f_av
        if not self.use_wigner:
            if not transpose:
                pY = d2v * Y_w[:, None] + d2v * Y_w[None, :]
                pRA = d2v * ReA_w[:, None] + d2v * ReA_w[None, :]
            else:
                Yi = 1.0 / Y_w
                pY = 0.5 * Yi[:, None] * d2v * Yi[None, :]
                pRA = np.zeros_like(pY)
                mask = np.full(pY.shape, 2.0)
                np.fill_diagonal(mask, 1.0)
                pY *= mask
                pRA *= mask
        else:
            chi_m, chi_p = chi_pm(self.w, self.T)
            pY = np.sqrt(-0.5 * chi_m) * d2v
            pRA = -np.sqrt(0.5 * chi_p) * d2v
        nw2 = n * (n + 1) // 2
        out[n:n + nw2] = pack_upper(pY)
        out[n + nw2:] = pack_upper(pRA)
        return -out if not self.use_wigner else out

    # DL:3549-3651
    def apply_full_L(self, psi, transpose=False):
        psi = np.asarray(psi, np.float64)
        out = harmonic_L1(psi, self.w, self.T, self.use_wigner, transpose,
                          self.ignore_harmonic)
        if not self.ignore_v3:
            out = out + self.anharmonic(psi, transpose)
        if self.reverse_L:
            out = -out
        out = out + self.shift_value * psi
        return out


# --------------------------------------------------------------------------
# Dense explicit-tensor L (brute force pin for tiny n)
# --------------------------------------------------------------------------
def dense_L(sys: OracleSystem):
    """Explicit L operator from D3/D4 tensors (DL:5167-5900, symmetrize=False).

    Independent algebra for the anharmonic part: d3 = -(1/N_eff) sym6(sum
    rho X_ups X_ups Y), d4 = -(1/(4 N_eff)) sum of the four (X_ups^3 Y)
    placements.  Only for tiny n (O(n^4) memory).  No symmetry images.
    """
    n = sys.n
    w, T = sys.w, sys.T
    r, c = upper_pairs(n)
    new_i_a, new_i_b = c, r          # i_a = column (w_a), i_b = row (w_b)
    w_a, w_b, n_a, n_b = _pair_quantities(w, T)
    nw2 = len(w_a)
    L = np.zeros((n + 2 * nw2, n + 2 * nw2))
    sY, sA = n, n + nw2
    # harmonic part: the diagonal blocks of harmonic_L1
    for k in range(n + 2 * nw2):
        e = np.zeros(n + 2 * nw2)
        e[k] = 1.0
        L[:, k] = harmonic_L1(e, w, T, sys.use_wigner, False, sys.ignore_harmonic)
    n_eff = np.sum(sys.rho)
    Yw = sys.Y * sys.rho[:, None]
    Xu = sys.X * f_ups(w, T)[None, :]
    ec = np.where(new_i_a != new_i_b, 2.0, 1.0)
    if not sys.ignore_v3:
        d3n = -np.einsum("ia,ib,ic->abc", Xu, Xu, Yw) / n_eff
        d3 = (d3n + d3n.transpose(0, 2, 1) + d3n.transpose(1, 0, 2)
              + d3n.transpose(1, 2, 0) + d3n.transpose(2, 0, 1)
              + d3n.transpose(2, 1, 0)) / 6
        d3s = d3[new_i_a, new_i_b, :]
        if not sys.use_wigner:
            Zc = 2 * ((2 * n_a + 1) * w_b + (2 * n_b + 1) * w_a) / ((2 * n_a + 1) * (2 * n_b + 1))
            L[sY:sA, :sY] = -(d3s * Zc[:, None])
            Z1 = 2 * ((2 * n_a + 1) * w_b * n_b * (n_b + 1)
                      + (2 * n_b + 1) * w_a * n_a * (n_a + 1)) / ((2 * n_a + 1) * (2 * n_b + 1))
            L[sA:, :sY] = -(d3s * Z1[:, None])
            X2 = (2 * n_b + 1) * (2 * n_a + 1) / (8 * w_a * w_b)
            L[:sY, sY:sA] = -(d3s * (X2 * ec)[:, None]).T
        else:
            chm = ((n_a - n_b) * (w_a - w_b) / (2 * w_a * w_b))
            chp = ((1 + n_a + n_b) * (w_a + w_b) / (2 * w_a * w_b))
            L[:sY, sY:sA] = (d3s * (np.sqrt(-0.5 * chm) * ec)[:, None]).T
            L[:sY, sA:] = -(d3s * (np.sqrt(0.5 * chp) * ec)[:, None]).T
            L[sY:sA, :sY] = d3s * np.sqrt(-0.5 * chm)[:, None]
            L[sA:, :sY] = -(d3s * np.sqrt(0.5 * chp)[:, None])
    if not sys.ignore_v4 and not sys.ignore_v3:
        d4 = (np.einsum("ia,ib,ic,id->abcd", Xu, Xu, Xu, Yw)
              + np.einsum("ia,ib,ic,id->abcd", Xu, Xu, Yw, Xu)
              + np.einsum("ia,ib,ic,id->abcd", Xu, Yw, Xu, Xu)
              + np.einsum("ia,ib,ic,id->abcd", Yw, Xu, Xu, Xu)) / (-4 * n_eff)
        d4s = d4[new_i_a, new_i_b][:, new_i_a, new_i_b]
        if not sys.use_wigner:
            Xl = -((2 * w_a * n_b + 2 * w_b * n_a + w_a + w_b)) / (4 * (2 * n_a + 1) * (2 * n_b + 1))
            Xr = ((2 * n_a + 1) * (2 * n_b + 1)) / (w_a * w_b)
            L[sY:sA, sY:sA] += -(Xl[:, None] * d4s * (Xr * ec)[None, :])
            X1l = -(w_a * n_a * (n_a + 1) * (2 * n_b + 1)
                    + w_b * n_b * (n_b + 1) * (2 * n_a + 1)) / (4 * (2 * n_a + 1) * (2 * n_b + 1))
            L[sA:, sY:sA] += -(X1l[:, None] * d4s * (Xr * ec)[None, :])
        else:
            chp = ((1 + n_a + n_b) * (w_a + w_b) / (2 * w_a * w_b))
            chm = ((n_a - n_b) * (w_a - w_b) / (2 * w_a * w_b))
            sp = np.sqrt(0.5 * chp)
            sm = np.sqrt(-0.5 * chm)
            L[sY:sA, sY:sA] += -(sm[:, None] * d4s * (sm * ec)[None, :])
            L[sA:, sA:] += -(sp[:, None] * d4s * (sp * ec)[None, :])
            L[sY:sA, sA:] += sm[:, None] * d4s * (sp * ec)[None, :]
            L[sA:, sY:sA] += sp[:, None] * d4s * (sm * ec)[None, :]
    return L


# --------------------------------------------------------------------------
# Lanczos recurrence
# --------------------------------------------------------------------------
def mask_dot_wigner(n):
    """Metric weights: 2 on off-diagonal pairs, 1 elsewhere (DL:5905-5954)."""
    m = np.ones(len_psi(n))
    r, c = upper_pairs(n)
    nw2 = len(r)
    off = r < c
    m[n:n + nw2][off] = 2.0
    m[n + nw2:][off] = 2.0
    return m


class LanczosState:
    def __init__(self):
        self.a = []
        self.b = []
        self.c = []
        self.Q = []
        self.P = []
        self.s_norm = []
        self.psi = None


def run_FT(matvec, rmatvec, psi0, n_iter, run_simm, mask=None, n_rep_orth=0,
           n_ortho=10, state=None):
    """Biconjugate / symmetric Lanczos exactly as run_FT (DL:5958-6491).

    ``matvec``/``rmatvec`` apply L and L^T.  Includes the reference quirks
    of the Gram-Schmidt branch (``ortho_p += abs(ortho_p)``, p divided by
    its squared norm; DL:6364-6436).  Returns the LanczosState.
    """
    st = state if state is not None else LanczosState()
    if mask is None:
        mask = np.ones(len(psi0))
    dot = (lambda u, v: u.dot(v * mask)) if run_simm else (lambda u, v: u.dot(v))
    if len(st.a) == 0:
        first = psi0 / np.sqrt(dot(psi0, psi0))
        st.Q = [first]
        st.P = [first]
        st.s_norm = [1.0]
    psi_q = st.Q[-1]
    psi_p = st.P[-1]
    next_converged = False
    i_step = len(st.a)
    for _ in range(i_step, i_step + n_iter):
        L_q = matvec(psi_q)
        if run_simm:
            p_L = L_q.copy()
        else:
            p_L = rmatvec(psi_p)
        st.psi = p_L if not run_simm else L_q
        c_old = st.c[-1] if len(st.c) > 0 else 1.0
        p_norm = st.s_norm[-1] / c_old
        a_coeff = dot(psi_p, L_q) * p_norm
        if np.isnan(a_coeff):
            raise ValueError("Invalid value encountered during the Lanczos.")
        rk = L_q - a_coeff * psi_q
        if len(st.Q) > 1:
            rk = rk - st.c[-1] * st.Q[-2]
        sk = p_L - a_coeff * psi_p
        if len(st.P) > 1:
            if len(st.c) < 2:
                old_p_norm = st.s_norm[-2]
            else:
                old_p_norm = st.s_norm[-2] / st.c[-2]
            sk = sk - st.b[-1] * st.P[-2] * (old_p_norm / p_norm)
        s_norm = np.sqrt(dot(sk, sk))
        sk_tilde = sk / s_norm
        s_norm *= p_norm
        b_coeff = np.sqrt(dot(rk, rk))
        c_coeff = dot(sk_tilde, rk / b_coeff) * s_norm
        st.a.append(a_coeff)
        if abs(b_coeff) < EPSILON or next_converged:
            break
        if abs(c_coeff) < EPSILON or next_converged:
            break
        psi_q = rk / b_coeff
        psi_p = sk_tilde.copy()
        new_q = psi_q.copy()
        new_p = psi_p.copy()
        for _k in range(n_rep_orth):
            ortho_q = 0.0
            ortho_p = 0.0
            start = 0
            if n_ortho is not None:
                start = max(len(st.P) - n_ortho, 0)
            for j in range(start, len(st.P)):
                coeff1 = dot(st.P[j], new_q)
                coeff2 = dot(st.Q[j], new_p)
                new_q = new_q - coeff1 * st.P[j]
                new_p = new_p - coeff2 * st.Q[j]
                ortho_q += abs(coeff1)
                ortho_p += abs(ortho_p)
            normq = np.sqrt(dot(new_q, new_q))
            if normq < EPSILON:
                next_converged = True
            new_q = new_q / normq
            normp = dot(new_p, new_p)
            if abs(normp) < EPSILON:
                next_converged = True
            new_p = new_p / normp
            s_norm = c_coeff / dot(new_p, new_q)
            if ortho_p < EPSILON and ortho_q < EPSILON:
                break
        st.Q.append(new_q)
        st.P.append(new_p)
        psi_q = new_q.copy()
        psi_p = new_p.copy()
        st.b.append(b_coeff)
        st.c.append(c_coeff)
        st.s_norm.append(s_norm)
    return st


def run_symmetric_full_reorth(matvec, psi0, n_iter, mask, tol=1e-12):
    """Symmetric M-inner-product Lanczos with full reorthogonalisation.

    The recurrence of run_FT with run_simm (DL:6145-6240, p = q) plus a full
    classical Gram-Schmidt pass, repeated twice (CGS2), against every stored
    basis vector, the real-space analogue of the Hermitian q-space Lanczos
    reorthogonalisation (Modules/QSpaceLanczos.py:1113-1130).  Stops when
    |b| < tol.  Returns (a, b, Q).
    """
    dot = lambda u, v: u.dot(v * mask)
    q = psi0 / np.sqrt(dot(psi0, psi0))
    Q = [q]
    a, b = [], []
    for _ in range(n_iter):
        Lq = matvec(Q[-1])
        ak = dot(Q[-1], Lq)
        r = Lq - ak * Q[-1]
        if len(Q) > 1:
            r = r - b[-1] * Q[-2]
        a.append(ak)
        Qm = np.array(Q)
        for _rep in range(2):
            coef = Qm @ (r * mask)
            r = r - coef @ Qm
        bk = np.sqrt(dot(r, r))
        if abs(bk) < tol:
            break
        b.append(bk)
        Q.append(r / bk)
    return np.array(a), np.array(b), np.array(Q)


# --------------------------------------------------------------------------
# Continued fraction
# --------------------------------------------------------------------------
def continued_fraction(a_coeffs, b_coeffs, c_coeffs, w_array, use_wigner,
                       use_terminator=True, last_average=1, smearing=0.0,
                       smooth_ramp=0, reverse_L=False, shift_value=0.0,
                       perturbation_modulus=1.0):
    """get_green_function_continued_fraction (DL:5032-5162)."""
    a_coeffs = list(a_coeffs)
    b_coeffs = list(b_coeffs)
    c_coeffs = list(c_coeffs)
    w_array = np.asarray(w_array, np.float64)
    n_iters = len(a_coeffs)
    gf = np.zeros(w_array.shape, dtype=np.complex128)
    sign = -1 if reverse_L else 1
    a_mean = b_mean = c_mean = None
    nonsym = len(c_coeffs) == len(b_coeffs)
    if use_terminator:
        a_av = np.mean(a_coeffs[-last_average:])
        b_av = np.mean(b_coeffs[-last_average:])
        c_av = np.mean(c_coeffs[-last_average:]) if nonsym else b_av
        a = a_av * sign - sign * shift_value
        b = b_av * sign
        c = c_av * sign
        a_mean, b_mean, c_mean = a, b, c
        d = a - w_array ** 2 if not use_wigner else a + w_array ** 2
        disc = d ** 2 - 4 * b * c
        sq = np.sqrt(disc + 0j)
        gp = (d + sq) / (2 * b * c)
        gm = (d - sq) / (2 * b * c)
        gf[:] = np.where(np.abs(gp) < np.abs(gm), gp, gm)
    else:
        a = a_coeffs[-1] * sign - sign * shift_value
        if not use_wigner:
            gf[:] = 1 / (a - w_array ** 2 + 2j * w_array * smearing)
        else:
            gf[:] = 1 / (a + w_array ** 2 + 2j * w_array * smearing)
    for i in range(n_iters - 2, -1, -1):
        a = a_coeffs[i] * sign - sign * shift_value
        b = b_coeffs[i] * sign
        c = c_coeffs[i] * sign if nonsym else b
        if use_terminator and smooth_ramp > 0 and i >= n_iters - smooth_ramp:
            alpha = (i - n_iters + smooth_ramp + 1) / smooth_ramp
            a = (1 - alpha) * a + alpha * a_mean
            b = (1 - alpha) * b + alpha * b_mean
            c = (1 - alpha) * c + alpha * c_mean
        if not use_wigner:
            gf = 1.0 / (a - w_array ** 2 + 2j * w_array * smearing - b * c * gf)
        else:
            gf = 1.0 / (a + w_array ** 2 + 2j * w_array * smearing - b * c * gf)
    if not use_wigner:
        return gf * perturbation_modulus
    return (-np.real(gf) + 1j * np.imag(gf)) * perturbation_modulus
