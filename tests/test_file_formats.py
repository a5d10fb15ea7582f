"""The reference's on-disk formats (DL:3653-3839): status npz files and .abc
coefficient files, on the CPU (loading needs no device)."""
import os

import numpy as np
import pytest

from tdscha_amd.lanczos import Lanczos

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "snte_step10.npz")
# keys of Lanczos.save_status (DL:3699-3727) except the Julia objects sym_julia
# and deg_julia
REF_KEYS = ["T", "nat", "m", "w", "pols", "n_modes", "ignore_v3", "ignore_v4", "N", "rho", "X",
            "Y", "psi", "a_coeffs", "b_coeffs", "c_coeffs", "basis_Q", "basis_P", "s_norm",
            "krilov_basis", "arnoldi_matrix", "reverse", "shift", "perturbation_modulus",
            "use_wigner", "ignore_small_w", "n_syms"]


def test_status_npz_loads_and_round_trips(tmp_path):
    lanc = Lanczos()
    lanc.load_status(GOLDEN)
    assert lanc.n_modes == 45 and lanc.N == 10 and lanc.X.shape == (10, 45)
    assert len(lanc.a_coeffs) == 10 and len(lanc.basis_Q) == 11
    out = str(tmp_path / "status")
    lanc.save_status(out)                     # the extension is added
    assert os.path.exists(out + ".npz")
    with np.load(out + ".npz", allow_pickle=False) as new:
        missing = [k for k in REF_KEYS if k not in new.files]
        assert not missing, missing
    back = Lanczos()
    back.load_status(out)
    for k in ("X", "Y", "w", "pols", "rho", "psi", "m"):
        np.testing.assert_array_equal(getattr(back, k), getattr(lanc, k))
    for k in ("a_coeffs", "b_coeffs", "c_coeffs", "s_norm"):
        np.testing.assert_array_equal(np.asarray(getattr(back, k)), np.asarray(getattr(lanc, k)))
    np.testing.assert_array_equal(np.asarray(back.basis_Q), np.asarray(lanc.basis_Q))
    assert back.T == lanc.T and back.use_wigner == lanc.use_wigner
    with pytest.raises(IOError):
        Lanczos().load_status(str(tmp_path / "absent"))


def test_abc_round_trip(tmp_path):
    lanc = Lanczos()
    lanc.a_coeffs = [1.5, -2.0, 3.25]
    lanc.b_coeffs = [0.5, 0.125]
    lanc.c_coeffs = [0.5, 0.25]
    f = str(tmp_path / "chain.abc")
    lanc.save_abc(f)
    with open(f) as fh:
        assert fh.readline().startswith("# a; b; c")
    back = Lanczos()
    back.load_abc(f)
    np.testing.assert_array_equal(back.a_coeffs, lanc.a_coeffs)
    np.testing.assert_array_equal(back.b_coeffs, lanc.b_coeffs + [0.0])
    np.testing.assert_array_equal(back.c_coeffs, lanc.c_coeffs + [0.0])
