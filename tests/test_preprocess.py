"""Ensemble preprocessing from raw arrays (tdscha_amd/preprocess.py) against
the rock-salt generator, which builds its mode-space data with the same steps
of DL:322-452 from the model directly."""
import numpy as np
import pytest

from tdscha_amd import preprocess as PP
from tdscha_amd import symmetry as SY
from tdscha_amd.rocksalt import Ensemble, RockSalt


def _raw(L=2, N=30, seed=4):
    model = RockSalt(L=L)
    ens = Ensemble(model, 300.0, N, seed=seed)
    u = ens.u_disps
    f = model.forces(u.reshape(N, model.nat, 3)).reshape(N, -1)
    fs = -(u @ model.force_constants().T)
    w, pols = model.harmonic_modes()
    rots = SY.cubic_rotations()
    maps = [np.arange(2)] * len(rots)          # both rock-salt sites are inversion centres
    return model, ens, u, f, fs, w, pols, rots, maps


def test_prepared_ensemble_equals_generator():
    model, ens, u, f, fs, w, pols, rots, maps = _raw()
    pe = PP.prepare_ensemble(u, f, fs, model.m_atoms, w, pols, ens.T, n_cells=model.n_cells,
                             rotations=rots, atom_maps=maps)
    assert pe.mode_mask.sum() == len(ens.w) == 3 * model.nat - 3
    np.testing.assert_allclose(pe.w, ens.w, rtol=1e-14)
    np.testing.assert_allclose(pe.X, ens.X, rtol=0, atol=1e-12 * np.abs(ens.X).max())
    np.testing.assert_allclose(pe.Y, ens.Y, rtol=0, atol=1e-12 * np.abs(ens.Y).max())
    assert pe.N == ens.N and pe.N_eff == pytest.approx(ens.N)


def test_units_weights_and_mode_selection():
    model, ens, u, f, fs, w, pols, rots, maps = _raw(N=20)
    base = PP.prepare_ensemble(u, f, fs, model.m_atoms, w, pols, ens.T, n_cells=model.n_cells)
    ang = PP.prepare_ensemble(u * PP.BOHR_ANGSTROM, f / PP.BOHR_ANGSTROM, fs / PP.BOHR_ANGSTROM,
                              model.m_atoms, w, pols, ens.T, n_cells=model.n_cells, angstrom=True)
    np.testing.assert_allclose(ang.X, base.X, rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(ang.Y, base.Y, rtol=1e-12, atol=1e-14)
    # the translations are found by overlap and by |w| alike
    small = PP.prepare_ensemble(u, f, fs, model.m_atoms, w, pols, ens.T, n_cells=model.n_cells,
                                ignore_small_w=True)
    assert np.array_equal(small.mode_mask, base.mode_mask)
    # selecting modes keeps the selected columns
    sel = np.zeros(len(w), bool)
    sel[10:20] = True
    part = PP.prepare_ensemble(u, f, fs, model.m_atoms, w, pols, ens.T, n_cells=model.n_cells,
                               select_modes=sel)
    keep = base.mode_mask[10:20].sum()
    assert part.X.shape == (20, keep)
    cols = np.flatnonzero(base.mode_mask)
    idx = [np.flatnonzero(cols == c)[0] for c in range(10, 20) if base.mode_mask[c]]
    np.testing.assert_allclose(part.X, base.X[:, idx], rtol=0, atol=1e-13 * np.abs(base.X).max())
    # doubling every weight changes nothing; the mean force follows the weights
    w2 = PP.prepare_ensemble(u, f, fs, model.m_atoms, w, pols, ens.T, rho=2 * np.ones(20),
                             n_cells=model.n_cells)
    np.testing.assert_allclose(w2.Y, base.Y, rtol=1e-13, atol=1e-15)
    with pytest.raises(ValueError):
        PP.prepare_ensemble(u, f, fs, model.m_atoms, w, pols, ens.T, n_cells=5)


def test_mode_projection_reconstructs_the_vibrational_part():
    """X pols^T / sqrt(m) is u without its rigid-translation component."""
    model, ens, u, f, fs, w, pols, rots, maps = _raw(N=6)
    pe = PP.prepare_ensemble(u, f, fs, model.m_atoms, w, pols, ens.T, n_cells=model.n_cells)
    ut = u * np.sqrt(pe.m)[None]
    trans = pols[:, ~pe.mode_mask]
    expect = ut - (ut @ trans) @ trans.T
    np.testing.assert_allclose(pe.X @ pe.pols.T, expect, atol=1e-12 * np.abs(ut).max())
    # a rigid translation has no vibrational component
    shift = np.tile([0.01, -0.02, 0.03], model.nat)[None]
    pe2 = PP.prepare_ensemble(u + shift, f, fs, model.m_atoms, w, pols, ens.T,
                              n_cells=model.n_cells)
    np.testing.assert_allclose(pe2.X, pe.X, atol=1e-12 * np.abs(pe.X).max())


def test_prepared_ensemble_detects_its_symmetries():
    """With the supercell structure the prepared ensemble finds its space group:
    the gamma-only point group + projector equal the rock-salt builder's."""
    from oracle import lanczos_oracle as LO
    model, ens, u, f, fs, w, pols, rots, maps = _raw(N=10)
    pe = PP.prepare_ensemble(u, f, fs, model.m_atoms, w, pols, ens.T, n_cells=model.n_cells,
                             rotations=rots, atom_maps=maps,
                             structure=(model.lattice, model.positions, model.species))
    det = pe.symmetries(True, gamma_only=True)
    ref = ens.symmetries(True, gamma_only=True)
    assert det.n_syms == 48 and len(det.trans_cart_perms) == model.n_cells
    np.testing.assert_allclose(det.trans_projector, ref.trans_projector, atol=1e-13)
    o1 = LO.OracleSystem(pe.X, pe.Y, pe.w, pe.rho, pe.T, det.dense(), use_wigner=True,
                         pols=pe.pols, trans_projector=det.trans_projector,
                         inv_perms=det.trans_cart_perms)
    o2 = LO.OracleSystem(ens.X, ens.Y, ens.w, ens.rho, ens.T, ref.dense(), use_wigner=True,
                         pols=ens.pols, trans_projector=ref.trans_projector,
                         inv_perms=ref.trans_cart_perms)
    v = np.random.default_rng(3).normal(size=LO.len_psi(len(pe.w)))
    a, b = o1.apply_full_L(v), o2.apply_full_L(v)
    assert np.linalg.norm(a - b) / np.linalg.norm(b) < 1e-11
    # in Angstrom the structure is converted like the displacements
    pa = PP.prepare_ensemble(u * PP.BOHR_ANGSTROM, f / PP.BOHR_ANGSTROM, fs / PP.BOHR_ANGSTROM,
                             model.m_atoms, w, pols, ens.T, n_cells=model.n_cells, angstrom=True,
                             structure=(model.lattice * PP.BOHR_ANGSTROM,
                                        model.positions * PP.BOHR_ANGSTROM, model.species))
    assert pa.symmetries(True, gamma_only=True).n_syms == 48
    with pytest.raises(ValueError):
        PP.prepare_ensemble(u, f, fs, model.m_atoms, w, pols, ens.T,
                            n_cells=model.n_cells).symmetries(True)
