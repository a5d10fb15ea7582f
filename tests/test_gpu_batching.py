"""GPU: real-space batching beyond one perturbed-average pass, whole-run
graphs and the K-split autotuner."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _lanczos(N=400, seed=2):
    from tdscha_amd.lanczos import Lanczos
    from tdscha_amd.rocksalt import Ensemble, RockSalt
    ens = Ensemble(RockSalt(L=2), 300.0, N, seed=seed)
    lanc = Lanczos(ens)
    lanc.init(use_symmetries=False)
    return ens, lanc


def test_twelve_uneven_chains_equal_single_runs_and_graph_modes():
    """12 chains (two perturbed-average groups of 8 and 4) with caps 3-25:
    every chain equals its own single-chain run; the whole run captured as
    one graph, one step replayed and plain launches agree bit for bit."""
    ens, lanc = _lanczos()
    rng = np.random.default_rng(31)
    psis = rng.normal(size=(12, len(lanc.psi)))
    caps = np.array([3, 25, 7, 25, 11, 19, 25, 4, 16, 25, 9, 22], dtype=np.int32)
    run, _ = lanc.run_FT_batched(psis, 25, caps=caps, reorth=True, use_graph="run")
    step, _ = lanc.run_FT_batched(psis, 25, caps=caps, reorth=True, use_graph="step")
    plain, _ = lanc.run_FT_batched(psis, 25, caps=caps, reorth=True, use_graph=False)
    for c in range(12):
        assert len(run[c][0]) == caps[c]
        for other in (step, plain):
            assert np.array_equal(run[c][0], other[c][0]) and np.array_equal(run[c][1], other[c][1])
    for c in (0, 6, 11):
        single, _ = lanc.run_FT_batched(psis[c], int(caps[c]), reorth=True, use_graph="run")
        assert np.allclose(single[0][0], run[c][0], rtol=1e-13, atol=0)
        assert np.allclose(single[0][1], run[c][1], rtol=1e-13, atol=0)
    lanc.close()


def test_biconjugate_whole_run_graph():
    from tests.systems import load_snte
    from tests.test_gpu_lanczos import _product_from_fixture
    snte = load_snte()
    lanc = _product_from_fixture(snte, use_wigner=False)
    lanc.prepare_mode(10)
    lanc.run_FT(10, verbose=False, device=True)
    ref = [np.array(x) for x in (lanc.a_coeffs, lanc.b_coeffs, lanc.c_coeffs)]
    import ctypes as C
    from tdscha_amd import _lib
    psi = np.zeros((1, len(lanc.psi)))
    psi[0, 10] = 1.0
    a, b, cc = (np.zeros((1, 10)) for _ in range(3))
    nd = np.zeros(1, dtype=np.int32)
    rc = _lib.lib().tdscha_lanczos_run_biconjugate(
        lanc.engine.ctx, 1, C.c_void_p(psi.ctypes.data), 10, None, 1e-12, 0, 0, 0, 0, 0, 0, 0.0,
        2, C.c_void_p(a.ctypes.data), C.c_void_p(b.ctypes.data), C.c_void_p(cc.ctypes.data),
        C.c_void_p(nd.ctypes.data), None)
    _lib.check(rc, "tdscha_lanczos_run_biconjugate")
    assert np.array_equal(a[0, :len(ref[0])], ref[0])
    assert np.array_equal(b[0, :len(ref[1])], ref[1]) and np.array_equal(cc[0, :len(ref[2])], ref[2])
    lanc.close()


def test_autotuner_keeps_the_result():
    from oracle import c_oracle as CO
    from tdscha_amd.engine import Engine
    from tests.systems import random_block_system, random_perturbation
    s = random_block_system(130, 3000, 6, [3, 5, 2] + [1] * 120, seed=61)
    eng = Engine(s["X"], s["Y"], s["w"], s["rho"], s["T"], s["syms"], s["blocks"], s["sbi"])
    R1, Y1 = random_perturbation(130, seed=62)
    Y1 = Y1 * 1e-3
    fo, do = CO.oracle_perturb_averages(s["X"], s["Y"], s["w"], s["rho"], R1, Y1, s["T"], 1,
                                        s["syms"], s["blocks"], s["sbi"], n_threads=8)
    ks, ms = eng.autotune(3, "fp64", reps=2)
    print("fp64 autotune: ksplit %d, %.3f ms" % (ks, ms))
    assert ks % 8 == 0 and ms > 0
    f, d = eng.perturb_averages(R1, Y1, 1)
    assert np.linalg.norm(f - fo) / np.linalg.norm(fo) < 1e-12
    assert np.linalg.norm(d - do) / np.linalg.norm(do) < 1e-12
    ks, ms = eng.autotune(3, "fp16x2", reps=2)
    print("fp16x2 autotune: ksplit %d, %.3f ms" % (ks, ms))
    f, d = eng.perturb_averages_batched(R1[None], Y1[None], 1, "fp16x2")
    assert np.linalg.norm(d[0] - do) / np.linalg.norm(do) < 2e-5
    eng.close()
