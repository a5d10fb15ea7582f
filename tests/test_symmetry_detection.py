"""Space-group detection (symmetry.find_symmetries, the role of spglib in the
reference) and the mode-space symmetries built from the detected operations."""
import numpy as np
import pytest

from oracle import lanczos_oracle as LO
from tdscha_amd import symmetry as SY
from tdscha_amd.rocksalt import Ensemble, RockSalt


def _is_group(ops, lattice):
    """Closure of {R|t} modulo the lattice."""
    A = np.asarray(lattice)
    Ai = np.linalg.inv(A)

    def key(R, t):
        f = t @ Ai
        f = f - np.floor(f + 1e-9)
        return (tuple(np.round(R, 6).ravel()), tuple(np.round(f, 6)))
    keys = {key(R, t) for R, t, _ in ops}
    for R1, t1, _ in ops:
        for R2, t2, _ in ops:
            if key(R1 @ R2, R1 @ t2 + t1) not in keys:
                return False
    return True


def test_rock_salt_primitive_cell_is_symmorphic_oh():
    m = RockSalt(L=1)
    ops = SY.find_symmetries(m.cell, m.positions, m.species)
    assert len(ops) == 48
    assert np.allclose(ops[0][0], np.eye(3))
    Ai = np.linalg.inv(m.cell)
    for R, t, irt in ops:
        f = t @ Ai
        assert np.allclose(f - np.round(f), 0.0, atol=1e-9)
        assert np.array_equal(irt, [0, 1])
    assert _is_group(ops, m.cell)


def test_rock_salt_supercell_splits_into_translations_and_point_group():
    m = RockSalt(L=2)
    ops = SY.find_symmetries(m.lattice, m.positions, m.species)
    assert len(ops) == 48 * 8
    trans, reps = SY.split_point_group(ops, m.lattice)
    assert len(trans) == 8 and len(reps) == 48
    assert all(np.allclose(t, 0.0, atol=1e-9) for _, t, _ in reps)


def test_diamond_is_nonsymmorphic():
    """Diamond (two identical atoms at 0 and a/4 (1,1,1) in the FCC primitive
    cell): 48 operations, the 24 outside T_d with a fractional translation."""
    a = 6.0
    cell = 0.5 * a * np.array([[0.0, 1, 1], [1, 0, 1], [1, 1, 0]])
    pos = np.array([[0.0, 0, 0], [0.25 * a, 0.25 * a, 0.25 * a]])
    ops = SY.find_symmetries(cell, pos, [0, 0])
    assert len(ops) == 48
    Ai = np.linalg.inv(cell)
    frac = [t @ Ai for _, t, _ in ops]
    nonzero = sum(not np.allclose(f - np.round(f), 0.0, atol=1e-9) for f in frac)
    assert nonzero == 24
    assert _is_group(ops, cell)
    # a distorted cell loses most operations
    pos2 = pos + np.array([[0, 0, 0], [0.05, 0, 0]])
    assert len(SY.find_symmetries(cell, pos2, [0, 0])) < 48


@pytest.mark.parametrize("gamma_only", [False, True])
def test_detected_mode_symmetries_give_the_builder_operator(gamma_only):
    """The L application with the detected operations equals the one with the
    rock-salt builder's (group averages do not depend on the order)."""
    ens = Ensemble(RockSalt(L=2), 300.0, 20, seed=6)
    m = ens.model
    ref = ens.symmetries(use_symmetries=True, gamma_only=gamma_only)
    det = SY.detect_mode_symmetries(m.lattice, m.positions, m.species, ens.pols, ens.w,
                                    gamma_only=gamma_only)
    assert det.n_syms == ref.n_syms and det.leakage < 1e-8
    kw = {}
    kw_ref = {}
    if gamma_only:
        kw = dict(pols=ens.pols, trans_projector=det.trans_projector,
                  inv_perms=det.trans_cart_perms)
        kw_ref = dict(pols=ens.pols, trans_projector=ref.trans_projector,
                      inv_perms=ref.trans_cart_perms)
        np.testing.assert_allclose(det.trans_projector, ref.trans_projector, atol=1e-13)
    o_det = LO.OracleSystem(ens.X, ens.Y, ens.w, ens.rho, ens.T, det.dense(), use_wigner=True,
                            **kw)
    o_ref = LO.OracleSystem(ens.X, ens.Y, ens.w, ens.rho, ens.T, ref.dense(), use_wigner=True,
                            **kw_ref)
    v = np.random.default_rng(1).normal(size=LO.len_psi(len(ens.w)))
    a, b = o_det.apply_full_L(v), o_ref.apply_full_L(v)
    assert np.linalg.norm(a - b) / np.linalg.norm(b) < 1e-12
