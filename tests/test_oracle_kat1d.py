"""Known-answer pin of the anharmonic L on a one-dimensional quartic well (CPU).

The reference's heavy test tests/_heavy_tests/test_fast_lanczos_1D/test_1d_fast.py
gives, for V(x) = 3x^4 - 3x^2 + x^3/2, unit mass and T = 0, the 3 x 3 L in the
standard representation, psi = (R, Y, ReA) (:126-143):

    L[0,0] = w^2   L[0,1] = -d3 / (8 w^2)   L[1,0] = -4 w d3
    L[1,1] = d4 / (2 w) + 2 w^2              L[1,2] = 4 w^2

with d3 = <V'''> = 72 x0 + 3 and d4 = <V''''> = 72 (:108-115).  That test is
skipped in the reference and asserts only that L and L^T agree (:437).  Here:

* the anharmonic part of the oracle's L (the full L minus the harmonic L) equals
  the closed form's d3 and d4 terms exactly, with d3, d4 estimated from the
  same samples as -Y^2 <u^2 df> and -Y^3 <u^3 df> (Y = 2w, df = f - f_scha;
  :118-132), so the D3 and D4 weights of the kernel are pinned by a closed
  form independent of the reference's code;
* those sample estimates agree with the analytic <V'''> and <V''''> within
  their Monte Carlo error (Gaussian integration by parts at the SSCHA minimum);
* the transposed L is the transpose of L (the reference test's assertion);
* the closed form's harmonic (Y, ReA) entries are half of those of the
  reference code's standard-representation harmonic block (DL:2396-2438),
  which the oracle follows and the reference's golden run pins; only the
  diagonal R entry, w^2, is shared.
"""
import numpy as np
import scipy.optimize

from oracle import lanczos_oracle as O


def quartic_well(N=200000, seed=0):
    def V(x):
        return 3 * x ** 4 - 3 * x ** 2 + 0.5 * x ** 3

    def force(x):
        return -(12 * x ** 3 - 6 * x + 1.5 * x ** 2)

    grid = np.linspace(-5, 5, 10000)

    def free_energy(par):
        x0, phi = par
        if phi <= 0:
            return 1e6 + phi ** 2
        xx = x0 + grid * phi
        rho = np.exp(-(xx - x0) ** 2 / (2 * phi ** 2)) / np.sqrt(2 * np.pi * phi ** 2)
        return 1 / (8 * phi ** 2) + np.sum(V(xx) * rho) * (xx[1] - xx[0])

    x0, phi = scipy.optimize.minimize(free_energy, [1.0, 1.0], method="BFGS").x
    w = 1 / (2 * phi ** 2)
    x = np.random.default_rng(seed).normal(x0, phi, N)
    u = x - x0
    f = force(x) + w ** 2 * u          # residual force: minus the SSCHA force -w^2 u
    return x0, phi, w, u, f


def closed_form(w, d3, d4):
    L = np.zeros((3, 3))
    L[0, 0] = w ** 2
    L[0, 1] = -d3 / (8 * w ** 2)
    L[1, 0] = -4 * w * d3
    L[1, 1] = d4 / (2 * w) + 2 * w ** 2
    L[1, 2] = 4 * w ** 2
    return L


def oracle_L(u, f, w, transpose=False, **flags):
    s = O.OracleSystem(u[:, None], f[:, None], np.array([w]), np.ones(len(u)), 0.0, None,
                       use_wigner=False, **flags)
    return np.array([s.apply_full_L(e, transpose=transpose) for e in np.eye(3)]).T


def test_anharmonic_L_matches_the_one_dimensional_closed_form():
    x0, phi, w, u, f = quartic_well()
    Y = 2 * w
    d3 = -Y ** 2 * np.mean(u * u * f)
    d4 = -Y ** 3 * np.mean(u ** 3 * f)
    L = oracle_L(u, f, w)
    Lh = oracle_L(u, f, w, ignore_v3=True, ignore_v4=True)
    anh = L - Lh
    anh_cf = closed_form(w, d3, d4) - closed_form(w, 0.0, 0.0)
    err = np.max(np.abs(anh - anh_cf)) / np.max(np.abs(anh_cf))
    print("x0 %.6f phi %.6f w %.6f d3 %.5f d4 %.5f: anharmonic L vs closed form %.2e" % (
        x0, phi, w, d3, d4, err))
    assert err < 1e-12
    # the sample estimates against the analytic averages, within 5 standard errors
    s3 = Y ** 2 * np.std(u * u * f) / np.sqrt(len(u))
    s4 = Y ** 3 * np.std(u ** 3 * f) / np.sqrt(len(u))
    assert abs(d3 - (72 * x0 + 3)) < 5 * s3 and abs(d4 - 72) < 5 * s4
    # transposition, and the shared harmonic entry
    Lt = oracle_L(u, f, w, transpose=True)
    assert np.max(np.abs(Lt - L.T)) < 1e-12 * np.max(np.abs(L))
    assert abs(Lh[0, 0] - w ** 2) < 1e-14 * w ** 2
    assert abs(Lh[1, 2] - 2 * closed_form(w, 0, 0)[1, 2]) < 1e-12 * w ** 2
