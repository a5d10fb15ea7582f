"""Dynamical-matrix files, supercell modes and the file-based ensemble
preparation, against the reference's SnTe test data (tests/golden/snte_raw:
the dynamical-matrix files and binary ensemble of tests/test_julia/data) and
the mode-space arrays its Lanczos saved for that ensemble
(tests/golden/snte_step10.npz: w, pols, X, Y of LANCZOS_STEP10.npz)."""
import os

import numpy as np
import pytest

from tdscha_amd import dynmat as DM
from tdscha_amd import preprocess as PP
from tdscha_amd import symmetry as SY

HERE = os.path.dirname(os.path.abspath(__file__))
RAW = os.path.join(HERE, "golden", "snte_raw")
PREFIX = os.path.join(RAW, "dyn_gen_pop1_")


@pytest.fixture(scope="module")
def ref():
    return dict(np.load(os.path.join(HERE, "golden", "snte_step10.npz")))


@pytest.fixture(scope="module")
def dyn():
    return DM.read_dyn_set(PREFIX, 3)


def test_read_dyn_header_and_stars(dyn):
    f0 = dyn.files[0]
    assert f0.alat == pytest.approx(12.4) and dyn.nat == 2
    assert [s.strip() for s in dyn.species] == ["Te", "Sn"]
    assert dyn.masses == pytest.approx([116300.2852960780001013, 108197.5459835219953675], rel=1e-15)
    assert [len(f.qpoints) for f in dyn.files] == [1, 4, 3]      # Gamma, the L star, the X star
    assert dyn.supercell_dims() == (2, 2, 2) and dyn.check_grid() == (2, 2, 2)
    for f in dyn.files:
        D = f.dynmats[0]
        assert np.abs(D - D.conj().T).max() < 1e-15              # Hermitian as written


def test_frequencies_of_each_star_match_the_printed_ones(dyn):
    """The files print the frequencies of their first q-point; ours from the
    matrices are the same up to one factor for all 15 nonzero modes (the
    file's Ry -> cm^-1 and mass constants differ from ours by 4e-7)."""
    ratios = []
    for f in dyn.files:
        mine = DM.q_frequencies(f)
        nz = np.abs(f.freqs_cm) > 1.0
        if (~nz).any():
            assert np.abs(mine[~nz]).max() < 1e-4               # acoustic modes at Gamma
        ratios.append(mine[nz] / f.freqs_cm[nz])
    r = np.concatenate(ratios)
    assert len(r) == 15
    assert abs(r.mean() - 1) < 1e-6 and r.std() < 1e-9


def test_supercell_modes_match_the_reference(dyn, ref):
    pos, lat, spec, mass = DM.supercell_structure(dyn)
    assert np.allclose(np.repeat(mass, 3), ref["m"], rtol=1e-15)
    phi = DM.supercell_force_constants(dyn, pos)
    w, pols = DM.supercell_modes(phi, mass)
    assert (w[:3] ** 2).max() < 1e-14 * w.max() ** 2             # translations: w^2 at rounding
    assert np.abs(w[3:] / ref["w"] - 1).max() < 1e-13
    # mode subspaces (eigenvectors are defined up to rotations inside them)
    for blk in SY.degenerate_blocks(ref["w"]):
        P = pols[:, 3 + blk] @ pols[:, 3 + blk].T
        Pr = ref["pols"][:, blk] @ ref["pols"][:, blk].T
        assert np.abs(P - Pr).max() < 1e-10


def test_supercell_order_is_that_of_the_reference_ensemble(dyn):
    xats, forces = DM.load_ensemble_bin(RAW, 1)
    assert xats.shape == forces.shape == (10, 16, 3)
    pos, lat, _, _ = DM.supercell_structure(dyn)
    u = xats[0] * PP.CC_BOHR_PER_ANGSTROM - pos
    u -= np.round(u @ np.linalg.inv(lat)) @ lat
    assert np.abs(u).max() < 0.6                                # bohr; nearest neighbours are 6.2 apart


def test_prepared_from_files_matches_the_reference_X_Y(ref):
    """dyn files + binary ensemble -> X, Y of the reference's Lanczos
    (DL:322-452); compared through the basis-free projections X pols^T."""
    ens = PP.prepare_from_files(PREFIX, 3, RAW, 1, 250.0)
    assert len(ens.w) == 45 and ens.N == 10 and ens.T == 250.0
    for k in ("X", "Y"):
        P = getattr(ens, k) @ ens.pols.T
        Pr = ref[k] @ ref["pols"].T
        err = np.abs(P - Pr).max() / np.abs(Pr).max()
        print(k, "%.2e" % err)
        assert err < 2e-9
    ms = ens.symmetries(use_symmetries=True, gamma_only=False)
    assert ms.n_syms == 384 and ms.leakage < 1e-10
    assert sorted(len(b) for b in ms.degenerate_space) == sorted(ref["ref_ndegs"].tolist())


def test_reader_errors(tmp_path, dyn):
    # incomplete grid: the X star missing
    part = DM.DynSet(dyn.files[:2])
    with pytest.raises(ValueError, match="distinct"):
        part.check_grid()
    # ensemble of another supercell
    np.save(tmp_path / "xats_pop1.npy", np.zeros((2, 8, 3)))
    np.save(tmp_path / "forces_pop1.npy", np.zeros((2, 8, 3)))
    with pytest.raises(ValueError, match="atoms"):
        PP.prepare_from_files(PREFIX, 3, str(tmp_path), 1, 250.0)
    # a file with ibrav != 0
    lines = open(PREFIX + "1").read().splitlines()
    head = lines[2].split()
    head[2] = "2"
    lines[2] = " ".join(head)
    (tmp_path / "dyn1").write_text("\n".join(lines) + "\n")
    with pytest.raises(NotImplementedError):
        DM.read_dyn(str(tmp_path / "dyn1"))
    # q space: the incomplete grid is refused before any Bloch transform
    with pytest.raises(ValueError, match="distinct"):
        DM.SupercellModel(DM.DynSet(dyn.files[:2]))


def test_supercell_model_bloch_modes_and_point_group(dyn):
    """The files as a q-space model: the Bloch basis diagonalises the
    supercell dynamical matrix, its frequencies are those of the supercell
    modes (and at Gamma the printed ones), and the 48 operations of O_h map
    the supercell onto itself."""
    from tdscha_amd.qspace import BlochModes, q_symmetries
    sm = DM.SupercellModel(dyn)
    assert sm.L == 2 and sm.nat == 16
    bm = BlochModes(sm)
    E = bm.basis()
    assert np.abs(E.conj().T @ E - np.eye(48)).max() < 1e-13
    s = 1 / np.sqrt(sm.m)
    M = E.conj().T @ (sm.force_constants() * s[:, None] * s[None, :]) @ E
    w2 = (np.sign(bm.w_q) * bm.w_q ** 2).T.ravel()
    assert np.abs(M - np.diag(w2)).max() < 1e-13 * np.abs(w2).max()
    pos, _, _, mass = DM.supercell_structure(dyn)
    w, _ = DM.supercell_modes(sm.force_constants(), mass)
    assert np.abs(np.sort(bm.w_q.ravel())[3:] - w[3:]).max() < 1e-13 * w.max()
    g = DM.q_frequencies(dyn.files[0])
    assert np.abs(np.sort(bm.w_q[:, 0]) * DM.RY_TO_CM / g - 1)[3:].max() < 1e-12
    perm, D = q_symmetries(bm)
    assert perm.shape == (48, 8) and D.shape == (48, 8, 6, 6)
    for k in range(48):
        assert sorted(perm[k]) == list(range(8))
        for iq in range(8):
            assert np.abs(D[k, iq] @ D[k, iq].conj().T - np.eye(6)).max() < 1e-12
