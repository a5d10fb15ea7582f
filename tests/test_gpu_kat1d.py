"""GPU: the device L on the one-dimensional quartic well of the reference's
heavy test (closed form and oracle pin in tests/test_oracle_kat1d.py): the
standard-representation L at T = 0 equals the oracle's, its anharmonic part the
closed form with the sample estimates of d3 and d4, and the transposed
application its transpose."""
import numpy as np
import pytest

from tests.test_oracle_kat1d import closed_form, oracle_L, quartic_well

pytestmark = pytest.mark.gpu


class _Ens:
    def __init__(self, **kw):
        self.__dict__.update(kw)


def _device_L(u, f, w, transpose=False, **flags):
    from tdscha_amd.lanczos import Lanczos
    ens = _Ens(T=0.0, w=np.array([w]), pols=np.ones((1, 1)), m=np.ones(1), N=len(u),
               rho=np.ones(len(u)), X=u[:, None], Y=f[:, None])
    lanc = Lanczos(ens, use_wigner=False)
    for k, v in flags.items():
        setattr(lanc, k, v)
    lanc.init(use_symmetries=False)
    cols = [lanc.apply_full_L(e.copy(), transpose=transpose) for e in np.eye(3)]
    lanc.close()
    return np.array(cols).T


def test_device_L_on_the_quartic_well():
    x0, phi, w, u, f = quartic_well()
    Y = 2 * w
    d3 = -Y ** 2 * np.mean(u * u * f)
    d4 = -Y ** 3 * np.mean(u ** 3 * f)
    L = _device_L(u, f, w)
    Lo = oracle_L(u, f, w)
    print("device L\n", L)
    assert np.max(np.abs(L - Lo)) < 1e-12 * np.max(np.abs(Lo))
    Lh = _device_L(u, f, w, ignore_v3=True, ignore_v4=True)
    anh_cf = closed_form(w, d3, d4) - closed_form(w, 0.0, 0.0)
    err = np.max(np.abs((L - Lh) - anh_cf)) / np.max(np.abs(anh_cf))
    print("anharmonic device L vs closed form %.2e" % err)
    assert err < 1e-12
    Lt = _device_L(u, f, w, transpose=True)
    assert np.max(np.abs(Lt - L.T)) < 1e-12 * np.max(np.abs(L))
