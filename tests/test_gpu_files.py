"""The reference's SnTe input files through both GPU paths: the q-space
Lanczos built from the dynamical matrices (Bloch modes of the files, Bloch
transform of the prepared ensemble, 48 point-group images) against the
real-space Lanczos of the same prepared ensemble with its detected 384
operations.  On the 2x2x2 grid every q equals -q, the two operators are the
same, so the chains agree to rounding at each of the 8 q-points (the reference's own comparison,
tests/test_qspace/test_qspace_lanczos.py:154-270, allows 1% on the Green
function)."""
import os

import numpy as np
import pytest

from tests.systems import GOLDEN

pytestmark = pytest.mark.gpu

RAW = os.path.join(GOLDEN, "snte_raw")


@pytest.mark.parametrize("iq", range(8))
def test_qspace_from_files_equals_real_space(iq):
    from tdscha_amd import dynmat as DM
    from tdscha_amd import preprocess as PP
    from tdscha_amd.lanczos import Lanczos, continued_fraction
    from tdscha_amd.qlanczos import QSpaceLanczos
    from tdscha_amd.qspace import BlochModes
    prefix = os.path.join(RAW, "dyn_gen_pop1_")
    ens = PP.prepare_from_files(prefix, 3, RAW, 1, 250.0)
    modes = BlochModes(DM.SupercellModel(DM.read_dyn_set(prefix, 3)))
    ql = QSpaceLanczos(ens, modes=modes)
    ql.init(use_symmetries=True)
    band = int(np.argmax(modes.w_q[:, iq]))
    ql.prepare_mode_q(iq, band)
    steps = 8
    aq, bq = ql.run_FT(steps)
    # the same perturbation in real space: the Bloch mode's supercell vector
    col = modes.basis()[:, iq * modes.n_bands + band]
    r = ens.pols.T @ col
    assert np.abs(r.imag).max() < 1e-12 * np.abs(r).max()
    lanc = Lanczos(ens)
    lanc.init(use_symmetries=True)
    psi = np.zeros(len(lanc.psi))
    psi[:len(r)] = r.real
    (ar, br), = lanc.run_FT_batched(np.array([psi]), steps, reorth=True)[0]
    ea = np.max(np.abs(np.asarray(aq) - ar[:len(aq)]) / np.abs(ar[:len(aq)]))
    eb = np.max(np.abs(np.asarray(bq) - br[:len(bq)]) / np.abs(br[:len(bq)]))
    w = np.array([0.0, modes.w_q[band, iq]])
    gq = ql.get_green_function_continued_fraction(w, use_terminator=False, smearing=1e-5)
    gr = continued_fraction(ar, br, None, w, True, use_terminator=False, smearing=1e-5)
    eg = np.max(np.abs(gq - gr) / np.abs(gr))
    print("iq %d band %d: a %.2e b %.2e G %.2e" % (iq, band, ea, eb, eg))
    assert ea < 1e-9 and eb < 1e-9 and eg < 1e-9
    ql.close()
    lanc.close()
