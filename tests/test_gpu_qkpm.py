"""GPU tests of the q-space kernel polynomial method (tdscha_amd/qkpm.py),
following the reference's tests/test_qspace/test_qspace_kpm.py: moments
against the oracle's one-moment-per-L recurrence, the harmonic peak exactly at
the harmonic frequency, the KPM peak against the continued fraction, and
exact continuation from a saved status."""
import os

import numpy as np
import pytest

from oracle import qspace_oracle as QO

pytestmark = pytest.mark.gpu


def _kpm(L=2, N=60, ignore_v3=False, ignore_v4=False, sym=False, seed=5):
    from tdscha_amd.qkpm import QSpaceKPM
    from tdscha_amd.rocksalt import Ensemble, RockSalt
    ens = Ensemble(RockSalt(L=L), 300.0, N, seed=seed)
    kpm = QSpaceKPM(ens)
    kpm.ignore_v3, kpm.ignore_v4 = ignore_v3, ignore_v4
    kpm.init(use_symmetries=sym)
    band = int(np.argmax(kpm.w_q[:, 0]))          # highest Gamma band
    kpm.prepare_mode_q(0, band)
    return ens, kpm, band


@pytest.mark.parametrize("L,iq", [(2, 0), (2, 3), (3, 0), (3, 5)])
def test_moments_match_the_oracle_recurrence(L, iq):
    """Against the oracle restatement of the reference's two-moments-per-L
    run on 2x2x2 and 3x3x3 grids, at Gamma and at q_pert != Gamma; on the
    2x2x2 grid, where L is self-adjoint in the metric, also against the
    one-moment-per-L recurrence.  (On odd grids the reference's q-space L is
    not self-adjoint in its metric, so there the two-moment identity, and
    with it the reference's KPM, departs from the direct moments.)"""
    ens, kpm, band = _kpm(L=L, N=60 if L == 2 else 30)
    kpm.prepare_mode_q(iq, int(np.argmax(kpm.w_q[:, iq])))
    kpm.run_KPM(33, verbose=False)
    from tdscha_amd.qspace import pair_map
    pm, pairs = pair_map(kpm.L, iq)
    qs = QO.QSystem(kpm.X_q, kpm.Y_q, kpm.w_q, kpm.valid_modes_q, ens.rho, ens.T, pm, pairs, iq)
    args = (qs.apply_full_L, kpm.psi, qs.mask(), 33, kpm.kpm_rescale_a, kpm.kpm_rescale_b)
    mu, norm = QO.kpm_moments_doubling(*args)
    err = np.abs(kpm.kpm_moments - mu).max()
    print("L=%d iq=%d moments vs oracle: max abs diff %.2e" % (L, iq, err))
    assert abs(kpm.kpm_vector_norm - norm) < 1e-14
    assert err < 1e-11
    if L == 2:
        mu_d, _ = QO.kpm_moments_direct(*args)
        assert np.abs(kpm.kpm_moments - mu_d).max() < 1e-11
    kpm.close()


def test_harmonic_peak_exact():
    """test_qspace_kpm.py:121-143: harmonic L, bounds without margin, 256
    moments; the peak on a grid of step 0.002 w is at w."""
    ens, kpm, band = _kpm(ignore_v3=True, ignore_v4=True)
    kpm.run_KPM(256, bound_factor=1.0, verbose=False)
    w_mode = kpm.w_q[band, 0]
    w = w_mode + 0.002 * w_mode * np.arange(-100, 101)
    s = kpm.get_spectral_function_KPM(w, regularization="jackson")
    assert int(np.argmax(s)) == 100
    kpm.close()


def test_peak_against_continued_fraction():
    """test_qspace_kpm.py:76-118: the KPM peak (256 moments, Jackson) within
    8% of the continued-fraction peak, with the point-group images."""
    from tdscha_amd.qlanczos import QSpaceLanczos
    ens, kpm, band = _kpm(sym=True, N=40)
    kpm.run_KPM(256, verbose=False)
    ql = QSpaceLanczos(ens)
    ql.init(use_symmetries=True)
    ql.prepare_mode_q(0, band)
    ql.run_FT(60)
    w_mode = kpm.w_q[band, 0]
    w = np.linspace(0.3 * w_mode, 1.4 * w_mode, 81)
    cf = -np.imag(ql.get_green_function_continued_fraction(w, use_terminator=False,
                                                            smearing=0.1 * w_mode))
    sk = kpm.get_spectral_function_KPM(w)
    p_cf, p_k = w[np.argmax(cf)], w[np.argmax(sk)]
    print("peak CF %.6e KPM %.6e rel %.3e" % (p_cf, p_k, abs(p_k - p_cf) / p_cf))
    assert abs(p_k - p_cf) / p_cf < 0.08
    ql.close()
    kpm.close()


@pytest.mark.parametrize("first", [1, 4, 5])
def test_continuation_is_exact(tmp_path, first):
    """test_qspace_kpm.py:146-180: first moments, save, restore on a fresh
    object, continue to 10 -> the one-shot moments and spectrum, bit for bit
    (odd, even and the single-moment restarts)."""
    _, ref, band = _kpm()
    ref.run_KPM(10, verbose=False)
    _, a, _ = _kpm()
    a.run_KPM(first, verbose=False)
    f = os.path.join(str(tmp_path), "kpm_checkpoint")
    a.save_status(f)
    _, b, _ = _kpm()
    b.load_status(f)
    b.run_KPM(10, verbose=False)
    np.testing.assert_array_equal(b.kpm_moments, ref.kpm_moments)
    w_mode = ref.w_q[band, 0]
    w = np.linspace(0.3 * w_mode, 1.4 * w_mode, 81)
    np.testing.assert_array_equal(b.get_spectral_function_KPM(w), ref.get_spectral_function_KPM(w))
    for k in (ref, a, b):
        k.close()


def test_from_qspace_lanczos_and_kpm_file(tmp_path):
    from tdscha_amd.qkpm import QSpaceKPM
    from tdscha_amd.qlanczos import QSpaceLanczos
    from tdscha_amd.rocksalt import Ensemble, RockSalt
    ens = Ensemble(RockSalt(L=2), 300.0, 60, seed=5)
    ql = QSpaceLanczos(ens)
    ql.init(use_symmetries=False)
    band = int(np.argmax(ql.w_q[:, 0]))
    ql.prepare_mode_q(0, band)
    kpm = QSpaceKPM.from_qspace_lanczos(ql)
    kpm.run_KPM(16, verbose=False)
    _, ref, _ = _kpm()
    ref.run_KPM(16, verbose=False)
    np.testing.assert_array_equal(kpm.kpm_moments, ref.kpm_moments)
    f = str(tmp_path / "m.kpm")
    kpm.save_kpm(f)
    back = QSpaceKPM(None)
    back.load_kpm(f)
    w = np.linspace(0.5, 1.5, 11) * kpm.w_q[band, 0]
    assert np.allclose(back.get_spectral_function_KPM(w), kpm.get_spectral_function_KPM(w),
                       rtol=1e-14)
    ref.close()
    ql.close()


def test_batched_equals_single_runs_graph_and_plain():
    """run_KPM_batched over several q-points (own pair maps and layouts) =
    one run_KPM per q-point, bit for bit; graph launch = plain launches."""
    from tdscha_amd.qspace import pair_map
    ens, kpm, _ = _kpm(L=2, N=50)
    iqs = [0, 3, 5, 7]
    rng = np.random.default_rng(2)
    psis = []
    for iq in iqs:
        pm, pairs = pair_map(kpm.L, iq)
        kpm.build_q_pair_map(iq)
        p = np.zeros(kpm.get_psi_size(), complex)
        p[:kpm.n_bands] = rng.normal(size=kpm.n_bands) + 1j * rng.normal(size=kpm.n_bands)
        psis.append(p)
    mom, norms, (a, b), ms = kpm.run_KPM_batched(iqs, psis, 24)
    mom2, _, _, _ = kpm.run_KPM_batched(iqs, psis, 24, use_graph=False)
    np.testing.assert_array_equal(mom, mom2)
    for i, iq in enumerate(iqs):
        kpm.build_q_pair_map(iq)
        kpm.psi = psis[i]
        kpm._invalidate_kpm_cache()
        kpm.run_KPM(24, verbose=False)
        np.testing.assert_array_equal(kpm.kpm_moments, mom[i])
        assert kpm.kpm_rescale_a == a and kpm.kpm_rescale_b == b
        assert abs(kpm.kpm_vector_norm - norms[i]) < 1e-15 * norms[i]
    print("batched KPM: 4 chains x 24 moments in %.2f ms" % ms)
    kpm.close()


def test_argument_errors_are_reported():
    """Wrong vector lengths and bad continuation arguments raise before any
    device work; the library rejects n_have >= n_moments."""
    from tdscha_amd import _lib
    ens, kpm, band = _kpm()
    pm = kpm.q_pair_map
    v = kpm.psi / np.sqrt(kpm._metric_dot(kpm.psi, kpm.psi))
    eng = kpm.engine
    with pytest.raises(ValueError):
        eng.kpm([0], [pm], [v[:-1]], [v[:-1]], [(1.0, 0.0)], 8)
    with pytest.raises(ValueError):
        eng.kpm([0], [pm], [v], [v], [(1.0, 0.0)], 8, n_have=4, moments=None)
    with pytest.raises(_lib.TdschaError, match="n_have"):
        eng.kpm([0], [pm], [v], [v], [(1.0, 0.0)], 8, n_have=8, moments=np.ones((1, 8)))
    with pytest.raises(_lib.TdschaError, match="rescale"):
        eng.kpm([0], [pm], [v], [v], [(0.0, 0.0)], 8)
    with pytest.raises(ValueError):
        eng.apply_L_static([0], [pm], [v])            # spectral layout into the static operator
    with pytest.raises(ValueError):
        eng.apply_L([0], [pm], [v[:5]])
    kpm.close()
