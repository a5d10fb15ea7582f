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


CIN = os.path.join(os.path.dirname(__file__), "golden", "snte_c_input")


def _snte_for_c_driver():
    """The SnTe system of the fixture with the builder's symmetry blocks and
    the perturbation of prepare_mode(10), as the reference had when it wrote
    tests/test_julia/test_c_code_lanczos (test_lanczos.py:40-57)."""
    from tests.systems import load_snte
    d = load_snte()
    lanc = Lanczos()
    lanc.load_status(GOLDEN)
    ms = d["syms_full"]
    lanc.symmetries = ms.symmetries
    lanc.degenerate_space = ms.degenerate_space
    lanc.N_degeneracy = ms.N_degeneracy
    lanc.sym_block_id = ms.sym_block_id
    lanc.use_wigner = False
    lanc.prepare_mode(10)
    return lanc


def test_c_driver_input_files_match_the_reference(tmp_path):
    """prepare_input_files (DL:873-978) against the files the reference wrote
    for the same system: json, frequencies, masses, weights, blocks,
    perturbation."""
    import json
    lanc = _snte_for_c_driver()
    lanc.prepare_input_files("mode10", 5, directory=str(tmp_path))
    with open(os.path.join(CIN, "mode10.json")) as fp:
        ref = json.load(fp)
    with open(tmp_path / "mode10.json") as fp:
        mine = json.load(fp)
    assert mine == ref
    for ext in ("freqs", "masses", "rho", "psi"):
        a = np.loadtxt(tmp_path / ("mode10." + ext))
        b = np.loadtxt(os.path.join(CIN, "mode10." + ext))
        np.testing.assert_allclose(a, b, rtol=1e-15, atol=0)
    for ext in ["ndegs", "blockid"] + ["block%d" % i for i in range(9)]:
        a = np.atleast_1d(np.loadtxt(tmp_path / ("mode10." + ext), dtype=int))
        b = np.atleast_1d(np.loadtxt(os.path.join(CIN, "mode10." + ext), dtype=int))
        np.testing.assert_array_equal(a, b)
    X = np.loadtxt(tmp_path / "mode10.X.dat")
    np.testing.assert_allclose(X, lanc.X, rtol=1e-15)
    # symmetry blocks: n_syms matrices of d_b rows each, blank line between
    s0 = [l for l in open(tmp_path / "mode10.symsb5").read().splitlines() if l.strip()]
    assert len(s0) == 384 * 8
    np.testing.assert_allclose(np.array([[float(x) for x in l.split()] for l in s0[:8]]),
                               lanc.symmetries[5][0], atol=1e-15)


def test_load_from_input_files(tmp_path):
    """DL:981-1044: coefficients from <root>.abc, settings from <root>.json,
    missing basis files only warn."""
    lanc = _snte_for_c_driver()
    lanc.prepare_input_files("run", 10, directory=str(tmp_path))
    ref = Lanczos()
    ref.load_status(GOLDEN)
    ref.save_abc(str(tmp_path / "run.abc"))
    back = Lanczos()
    with pytest.warns(UserWarning):
        back.load_from_input_files("run", directory=str(tmp_path))
    np.testing.assert_array_equal(back.a_coeffs, ref.a_coeffs)
    np.testing.assert_array_equal(back.c_coeffs, ref.c_coeffs)
    assert back.T == 250 and back.N == 10 and back.n_modes == 45 and back.ignore_v4 is False
    with pytest.raises(IOError):
        back.load_from_input_files("absent", directory=str(tmp_path))
