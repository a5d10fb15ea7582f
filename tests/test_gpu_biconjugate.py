"""GPU: the device-resident biconjugate Lanczos (run_FT without run_simm,
Modules/DynamicalLanczos.py:6126-6240) against the reference's golden run
and the host recurrence around the same device L."""
import ctypes as C

import numpy as np
import pytest

from tests.systems import load_snte
from tests.test_gpu_lanczos import _product_from_fixture

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def snte():
    return load_snte()


def test_device_biconjugate_reproduces_golden_run(snte):
    """The run of test_julia_lf.py:24-45 (standard representation, mode 10,
    10 steps).  The tenth b*c is ill-conditioned (a 1e-16 perturbation of L
    moves it by 1e-7, see test_gpu_lanczos), hence the reference's 1e-6."""
    lanc = _product_from_fixture(snte, use_wigner=False)
    lanc.prepare_mode(10)
    lanc.run_FT(10, verbose=False, device=True)
    a = np.array(lanc.a_coeffs)
    bc = np.array(lanc.b_coeffs) * np.array(lanc.c_coeffs)
    assert len(a) == 10 and len(bc) == 10
    da = np.max(np.abs(a / snte["a_coeffs"] - 1))
    dbc = np.abs(bc / (snte["b_coeffs"] * snte["c_coeffs"]) - 1)
    print("device biconjugate vs golden: a %.2e, b*c %s" % (da, dbc))
    assert da < 1e-10 and np.max(dbc[:9]) < 1e-10 and dbc[9] < 1e-6
    assert abs(lanc.c_coeffs[4] - 6.59744344e-07) / 6.59744344e-07 < 1e-6
    lanc.close()


@pytest.mark.parametrize("use_wigner,steps", [(False, 10), (True, 12)])
def test_device_biconjugate_equals_host_recurrence(snte, use_wigner, steps):
    lanc = _product_from_fixture(snte, use_wigner=use_wigner)
    lanc.prepare_mode(7)
    lanc.run_FT(steps, verbose=False, run_simm=False)
    host = [np.array(x) for x in (lanc.a_coeffs, lanc.b_coeffs, lanc.c_coeffs)]
    lanc.reset()
    lanc.prepare_mode(7)
    lanc.run_FT(steps, verbose=False, run_simm=False, device=True)
    dev = [np.array(x) for x in (lanc.a_coeffs, lanc.b_coeffs, lanc.c_coeffs)]
    for h, d in zip(host, dev):
        assert len(h) == len(d)
        k = min(len(h), 8)      # the last steps are as ill-conditioned as in the golden run
        print("device vs host:", np.max(np.abs(d[:k] / h[:k] - 1)))
        assert np.max(np.abs(d[:k] / h[:k] - 1)) < 1e-11
        assert np.max(np.abs(d / h - 1)) < 1e-5
    lanc.close()


def test_device_biconjugate_batched_uneven_chains_and_graph(snte):
    from tdscha_amd import _lib
    lanc = _product_from_fixture(snte, use_wigner=False)
    n_len = len(lanc.psi)
    psis = np.zeros((3, n_len))
    for c, mode in enumerate((10, 3, 30)):
        psis[c, mode] = 1.0
    psis[2, 100] = 0.3
    caps = np.array([4, 10, 7], dtype=np.int32)
    L = _lib.lib()

    def run(ps, cp, use_graph):
        nc = ps.shape[0]
        a, b, cc = (np.zeros((nc, 10)) for _ in range(3))
        nd = np.zeros(nc, dtype=np.int32)
        ps = np.ascontiguousarray(ps)
        rc = L.tdscha_lanczos_run_biconjugate(
            lanc.engine.ctx, nc, C.c_void_p(ps.ctypes.data), 10,
            None if cp is None else C.c_void_p(cp.ctypes.data), 1e-12, 0, 0, 0, 0, 0, 0, 0.0,
            use_graph, C.c_void_p(a.ctypes.data), C.c_void_p(b.ctypes.data),
            C.c_void_p(cc.ctypes.data), C.c_void_p(nd.ctypes.data), None)
        _lib.check(rc, "tdscha_lanczos_run_biconjugate")
        return a, b, cc, nd

    a, b, cc, nd = run(psis, caps, 1)
    ag, bg, cg, ndg = run(psis, caps, 0)
    assert list(nd) == [4, 10, 7] and list(ndg) == list(nd)
    assert np.array_equal(a, ag) and np.array_equal(b, bg) and np.array_equal(cc, cg)
    for c in range(3):
        a1, b1, c1, n1 = run(psis[c:c + 1], None, 1)
        k = int(nd[c])
        for x1, x in ((a1, a), (b1, b), (c1, cc)):
            assert np.allclose(x1[0, :k], x[c, :k], rtol=1e-13, atol=0)
    lanc.close()
