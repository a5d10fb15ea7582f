"""GPU parity of the q-space L application and Hermitian Lanczos against the
CPU oracle (oracle/qspace_oracle.py), FP64 complex on both sides.

Tolerances: one L application within 1e-12 of the oracle relative to the
largest entry; Lanczos coefficients within 1e-10 relative (the device uses two
classical Gram-Schmidt passes, the oracle's run here too)."""
import numpy as np
import pytest

from oracle import qspace_oracle as QO

pytestmark = pytest.mark.gpu


def _setup(L, N, seed=5, T=300.0):
    from tdscha_amd.qlanczos import QSpaceEngine
    from tdscha_amd.qspace import bloch_ensemble
    from tdscha_amd.rocksalt import Ensemble, RockSalt
    ens = Ensemble(RockSalt(L=L), T, N, seed=seed)
    modes, Xq, Yq = bloch_ensemble(ens)
    eng = QSpaceEngine(Xq, Yq, modes.w_q, modes.valid_modes_q, ens.rho, T)
    return ens, modes, Xq, Yq, eng


def _oracle(ens, modes, Xq, Yq, iq, **kw):
    from tdscha_amd.qspace import pair_map
    pm, pairs = pair_map(modes.L, iq)
    return QO.QSystem(Xq, Yq, modes.w_q, modes.valid_modes_q, ens.rho, ens.T, pm, pairs, iq, **kw)


def _rel(x, ref):
    return float(np.max(np.abs(x - ref)) / np.max(np.abs(ref)))


@pytest.mark.parametrize("L,N,iqs", [(2, 200, [0, 3, 7]), (4, 150, [0, 1, 21, 32, 63])])
@pytest.mark.parametrize("flags", [(0, 0, 0), (0, 1, 0), (0, 0, 1), (1, 0, 0), (0, 1, 1)])
def test_apply_L_batch_matches_oracle(L, N, iqs, flags):
    ih, i3, i4 = flags
    ens, modes, Xq, Yq, eng = _setup(L, N)
    rng = np.random.default_rng(11)
    pms, psis, refs = [], [], []
    for iq in iqs:
        qs = _oracle(ens, modes, Xq, Yq, iq, ignore_harmonic=bool(ih), ignore_v3=bool(i3),
                     ignore_v4=bool(i4))
        psi = rng.normal(size=qs.psi_size) + 1j * rng.normal(size=qs.psi_size)
        psi[:qs.nb] *= 0.3
        pms.append(qs.q_pair_map)
        psis.append(psi)
        refs.append(qs.apply_full_L(psi))
        assert eng.psi_size(qs.q_pair_map) == qs.psi_size
    outs = eng.apply_L(iqs, pms, psis, ih, i3, i4)
    for iq, o, r in zip(iqs, outs, refs):
        err = _rel(o, r) if np.max(np.abs(r)) > 0 else float(np.max(np.abs(o)))
        print("L=%d iq=%d flags=%s rel err %.2e" % (L, iq, flags, err))
        assert err < 1e-12
    eng.close()


@pytest.mark.parametrize("L,N", [(2, 10), (4, 157)])
@pytest.mark.parametrize("form,chunk", [(0, 0), (0, 64), (1, 0), (1, 3), (1, 40)])
def test_d2v_forms_and_chunks_match_oracle(L, N, form, chunk):
    """Both forms of the d2v kernel (LDS slabs of every q-point, or one
    workgroup per (chain, pair) block), chunks that do not divide the
    ensemble, and a last slab that runs past its end (N = 10 with slabs of 16
    configurations at 8 q-points, N = 157 with slabs of 2 at 64)."""
    ens, modes, Xq, Yq, eng = _setup(L, N)
    eng.set_d2v_form(form)
    eng.set_tiles(64, chunk, 16)
    iqs = [0, 3, 7] if L == 2 else [0, 1, 21, 42, 63]
    rng = np.random.default_rng(23)
    pms, psis, refs = [], [], []
    for iq in iqs:
        qs = _oracle(ens, modes, Xq, Yq, iq)
        psi = rng.normal(size=qs.psi_size) + 1j * rng.normal(size=qs.psi_size)
        pms.append(qs.q_pair_map)
        psis.append(psi)
        refs.append(qs.apply_full_L(psi))
    outs = eng.apply_L(iqs, pms, psis)
    for iq, o, r in zip(iqs, outs, refs):
        err = _rel(o, r)
        print("L=%d N=%d form=%d chunk=%d iq=%d rel err %.2e" % (L, N, form, chunk, iq, err))
        assert err < 1e-12
    eng.close()


def test_lanczos_matches_oracle_chains():
    ens, modes, Xq, Yq, eng = _setup(4, 150)
    iqs = [0, 5, 42]
    pms, psis, ref = [], [], []
    rng = np.random.default_rng(2)
    for iq in iqs:
        qs = _oracle(ens, modes, Xq, Yq, iq)
        psi = np.zeros(qs.psi_size, np.complex128)
        v = rng.normal(size=qs.nb) + 1j * rng.normal(size=qs.nb)
        psi[:qs.nb] = v / np.linalg.norm(v)
        pms.append(qs.q_pair_map)
        psis.append(psi)
        ref.append(QO.hermitian_lanczos(qs.apply_full_L, psi, qs.mask(), 25, mgs=False))
    chains, ms = eng.lanczos(iqs, pms, psis, 25)
    for (a, b, c), (ao, bo, co, _) in zip(chains, ref):
        n = min(len(a), len(ao))
        print("steps", len(a), len(ao), "a %.2e b %.2e" % (np.max(np.abs(a[:n] / ao[:n] - 1)),
                                                           np.max(np.abs(b / bo[:len(b)] - 1))))
        assert len(a) == len(ao) and len(b) == len(bo)
        np.testing.assert_allclose(a, ao, rtol=1e-10)
        np.testing.assert_allclose(b, bo, rtol=1e-10)
        np.testing.assert_allclose(c, b, rtol=1e-14)
    eng.close()


def test_uneven_batch_equals_per_chain_runs_graph_and_plain():
    """Uneven caps and a tolerance stop in one batch; the same chains run one
    by one; the whole-run graph and plain launches."""
    ens, modes, Xq, Yq, eng = _setup(4, 120)
    from tdscha_amd.qspace import pair_map
    rng = np.random.default_rng(4)
    iqs = [0, 7, 13, 22, 35, 48]
    caps = [5, 30, 12, 30, 1, 20]
    pms, psis = [], []
    for iq in iqs:
        pm, _ = pair_map(4, iq)
        n = eng.psi_size(pm)
        psi = np.zeros(n, np.complex128)
        v = rng.normal(size=6) + 1j * rng.normal(size=6)
        psi[:6] = v / np.linalg.norm(v)
        pms.append(pm)
        psis.append(psi)
    eng.set_tiles(128, 2048, 16)       # fixed chunking: batch and single runs sum alike
    batch, _ = eng.lanczos(iqs, pms, psis, 30, caps=caps)
    plain, _ = eng.lanczos(iqs, pms, psis, 30, caps=caps, use_graph=False)
    for i, iq in enumerate(iqs):
        (a, b, c), = eng.lanczos([iq], [pms[i]], [psis[i]], 30, caps=[caps[i]])[0]
        assert len(batch[i][0]) == caps[i]
        np.testing.assert_allclose(batch[i][0], a, rtol=1e-12)
        np.testing.assert_allclose(batch[i][1], b, rtol=1e-12)
        assert np.array_equal(batch[i][0], plain[i][0])
        assert np.array_equal(batch[i][1], plain[i][1])
    # a harmonic-only chain from one band stops on |b| < tol after one step
    pm3 = pair_map(4, 3)[0]
    psi3 = np.zeros(eng.psi_size(pm3), np.complex128)
    psi3[2] = 1.0
    (a, b, c), = eng.lanczos([3], [pm3], [psi3], 10, ignore_v3=True, ignore_v4=True,
                             tol=1e-8)[0]
    assert len(a) == 1 and len(b) == 0
    assert a[0] == pytest.approx(-modes.w_q[2, 3] ** 2, rel=1e-13)
    eng.close()


def test_autotune_picks_valid_tiles_and_keeps_results():
    ens, modes, Xq, Yq, eng = _setup(4, 300)
    from tdscha_amd.qspace import pair_map
    iqs = list(range(0, 64, 4))
    pms = [pair_map(4, iq)[0] for iq in iqs]
    rng = np.random.default_rng(9)
    psis = [rng.normal(size=eng.psi_size(pm)) + 0j for pm in pms]
    before = eng.apply_L(iqs, pms, psis)
    tiles, ms = eng.autotune(iqs, pms, reps=2)
    print("tiles", tiles, "ms %.3f" % ms)
    assert tiles[0] in (64, 128, 256) and tiles[1] > 0 and ms > 0
    after = eng.apply_L(iqs, pms, psis)
    for x, y in zip(before, after):
        assert _rel(x, y) < 1e-12
    eng.close()


@pytest.mark.parametrize("L,N,iqs", [(2, 30, [0, 3, 7]), (4, 8, [0, 1, 21, 42])])
def test_symmetrized_apply_L_matches_oracle(L, N, iqs):
    """Point-group images expanded on the device (48 per configuration) against
    the oracle's rotated ensemble."""
    from tdscha_amd.qspace import q_symmetries
    ens, modes, Xq, Yq, eng = _setup(L, N)
    sym = q_symmetries(modes)
    eng.symmetrize(*sym)
    assert eng.N == 48 * N
    rng = np.random.default_rng(17)
    pms, psis, refs = [], [], []
    for iq in iqs:
        qs = _oracle(ens, modes, Xq, Yq, iq, symmetries=sym)
        psi = rng.normal(size=qs.psi_size) + 1j * rng.normal(size=qs.psi_size)
        pms.append(qs.q_pair_map)
        psis.append(psi)
        refs.append(qs.apply_full_L(psi))
    outs = eng.apply_L(iqs, pms, psis)
    for iq, o, r in zip(iqs, outs, refs):
        err = _rel(o, r)
        print("L=%d iq=%d symmetrised rel err %.2e" % (L, iq, err))
        assert err < 1e-12
    eng.close()


def test_symmetrized_lanczos_and_mirror_match_oracle():
    """QSpaceLanczos.init(use_symmetries=True) + run_FT against the oracle's
    Hermitian Lanczos on the rotated ensemble; a bad q-point table is
    rejected."""
    from tdscha_amd.qlanczos import QSpaceLanczos
    from tdscha_amd.qspace import q_symmetries
    from tdscha_amd.rocksalt import Ensemble, RockSalt
    ens = Ensemble(RockSalt(L=4), 300.0, 10, seed=8)
    lanc = QSpaceLanczos(ens)
    lanc.init(use_symmetries=True)
    assert lanc.n_syms_qspace == 48 and lanc.engine.N == 480
    sym = q_symmetries(lanc.modes)
    for iq, band in ((0, 4), (9, 1)):
        lanc.prepare_mode_q(iq, band)
        a, b = lanc.run_FT(12)
        qs = _oracle(ens, lanc.modes, lanc.X_q, lanc.Y_q, iq, symmetries=sym)
        psi = np.zeros(qs.psi_size, np.complex128)
        psi[band] = 1.0
        ao, bo, _, _ = QO.hermitian_lanczos(qs.apply_full_L, psi, qs.mask(), 12, mgs=False)
        print("iq %d a %s / %s" % (iq, a[:3], ao[:3]))
        assert len(a) == len(ao)
        np.testing.assert_allclose(a, ao, rtol=1e-10)
        np.testing.assert_allclose(b, bo, rtol=1e-10)
    perm, D = sym
    bad = perm[:2].copy()
    bad[1, 0] = bad[1, 1]
    with pytest.raises(RuntimeError, match="already expanded"):
        lanc.engine.symmetrize(perm, D)
    from tdscha_amd.qlanczos import QSpaceEngine
    fresh = QSpaceEngine(lanc.X_q, lanc.Y_q, lanc.w_q, lanc.valid_modes_q, lanc.rho, lanc.T)
    with pytest.raises(RuntimeError, match="not a permutation"):
        fresh.symmetrize(bad, D[:2])
    fresh.close()
    lanc.close()
