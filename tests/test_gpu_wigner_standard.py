"""GPU: Wigner and standard representations give the same Green function
(the reference's tests/test_lanczos_fast/test_wigner_vs_nowigner.py, restated):
the reference's SnTe input files, T = 250 K, mode 10, 50 steps, the symmetric
Wigner recurrence against the biconjugate standard one; the Green function
without terminator and with smearing 0.1 w_mode at w = 0 and w = w_mode.  The
renormalised frequency sqrt(1 / Re G(0)) agrees within 0.01 cm^-1 and G(w_mode)
within 0.1% (the reference's tolerances, :26-27, 117-145), for the host
recurrence around device L applications and for the device-resident
recurrences."""
import os

import numpy as np
import pytest

from tests.systems import GOLDEN

pytestmark = pytest.mark.gpu

RY_TO_CM = 109737.31568527
MODE, STEPS = 10, 50


def _green(ens, use_wigner, device):
    from tdscha_amd.lanczos import Lanczos
    lanc = Lanczos(ens, use_wigner=use_wigner)
    lanc.init(use_symmetries=True)
    lanc.prepare_mode(MODE)
    lanc.run_FT(STEPS, run_simm=use_wigner, verbose=False, device=device)
    w_mode = lanc.w[MODE]
    g = lanc.get_green_function_continued_fraction(np.array([0.0, w_mode]), use_terminator=False,
                                                   smearing=0.1 * w_mode)
    n = len(lanc.a_coeffs)
    lanc.close()
    return g, n


@pytest.mark.parametrize("device", [False, True])
def test_wigner_and_standard_green_functions_agree(device):
    from tdscha_amd import preprocess as PP
    raw = os.path.join(GOLDEN, "snte_raw")
    ens = PP.prepare_from_files(os.path.join(raw, "dyn_gen_pop1_"), 3, raw, 1, 250.0)
    gw, nw = _green(ens, True, device)
    gs, ns = _green(ens, False, device)

    def freq(g):
        w2 = 1.0 / g[0].real
        return np.sign(w2) * np.sqrt(abs(w2)) * RY_TO_CM

    fw, fs = freq(gw), freq(gs)
    ds = abs(gw[1].imag - gs[1].imag) / max(abs(gw[1].imag), abs(gs[1].imag))
    dr = abs(gw[1].real - gs[1].real) / max(abs(gw[1].real), abs(gs[1].real))
    print("device=%s: %d / %d steps; renormalised frequency %.6f vs %.6f cm^-1; G(w_mode) "
          "Im %.2e Re %.2e" % (device, nw, ns, fw, fs, ds, dr))
    assert abs(fw - fs) < 0.01
    assert ds < 1e-3 and dr < 1e-3
