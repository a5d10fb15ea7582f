"""The operator-rotation form of the symmetric perturbed averages (CPU).

With images b = (i, j), u_b = S_j x_i and f_b = S_j y_i (the reference
applies S_j per image, CModules/LanczosFunctions.c:2158-2170), and S_j block
diagonal over degenerate modes, the averages over the images equal

    f = sum_j S_j f~_j / N_sym,     d2v = sum_j S_j d2v~_j S_j^T / N_sym

where f~_j, d2v~_j are the averages over the configurations alone (one
identity symmetry, normalised by N_eff) for the rotated perturbation
(S_j^T R1, S_j^T Y1 S_j).  This is what the device computes with
TDSCHA_CTX_ROTATE_OPERATORS (tdscha_amd/csrc/rotate.hip); the identity is
checked here on the C oracle, for the SnTe space group of the reference's
test data and for random block systems (orthogonal and not).  Rotating the
inputs with S_j instead of S_j^T fails it.
"""
import numpy as np
import pytest

from oracle import c_oracle as CO
from tests.systems import load_snte, random_block_system, random_perturbation


def dense_syms(syms, blocks, n):
    S = np.zeros((syms[0].shape[0], n, n))
    for s, b in zip(syms, blocks):
        b = np.asarray(b)
        S[:, b[:, None], b[None, :]] = s
    return S


def rotated_form(X, Y, w, rho, R1, Y1, T, d4, syms, blocks, sbi, slip=False):
    n = X.shape[1]
    S = dense_syms(syms, blocks, n)
    ident = [np.eye(len(b))[None] for b in blocks]
    f, d = np.zeros(n), np.zeros((n, n))
    for Sj in S:
        A = Sj if slip else Sj.T
        fo, do = CO.oracle_perturb_averages(X, Y, w, rho, A @ R1, A @ Y1 @ A.T, T, d4, ident,
                                            blocks, sbi)
        f += Sj @ fo
        d += Sj @ do @ Sj.T
    return f / len(S), d / len(S)


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def _systems():
    snte = load_snte()
    ms = snte["syms_full"]
    yield ("snte", snte["X"], snte["Y"], snte["w"], snte["rho"], float(snte["T"]),
           ms.symmetries, ms.degenerate_space, ms.sym_block_id)
    s = random_block_system(40, 30, 6, [3, 5, 2] + [1] * 30, seed=5)
    yield "random-orthogonal", s["X"], s["Y"], s["w"], s["rho"], s["T"], s["syms"], s["blocks"], s["sbi"]
    # blocks that are not orthogonal: the identity does not need S_j^T = S_j^-1
    rng = np.random.default_rng(9)
    syms = [np.concatenate([np.eye(len(b))[None], rng.normal(size=(5, len(b), len(b)))])
            for b in s["blocks"]]
    yield "random-general", s["X"], s["Y"], s["w"], s["rho"], s["T"], syms, s["blocks"], s["sbi"]


@pytest.mark.parametrize("d4", [0, 1])
def test_rotated_operators_equal_the_image_average(d4):
    for name, X, Y, w, rho, T, syms, blocks, sbi in _systems():
        n = X.shape[1]
        R1, Y1 = random_perturbation(n, seed=3)
        Y1 *= 1e-3
        fi, di = CO.oracle_perturb_averages(X, Y, w, rho, R1, Y1, T, d4, syms, blocks, sbi)
        fr, dr = rotated_form(X, Y, w, rho, R1, Y1, T, d4, syms, blocks, sbi)
        print("%s: rotated vs images f %.2e d2v %.2e" % (name, rel(fr, fi), rel(dr, di)))
        assert rel(fr, fi) < 1e-12 and rel(dr, di) < 1e-12


def test_transposed_rotation_fails():
    for name, X, Y, w, rho, T, syms, blocks, sbi in list(_systems())[1:]:
        n = X.shape[1]
        R1, Y1 = random_perturbation(n, seed=4)
        Y1 *= 1e-3
        fi, di = CO.oracle_perturb_averages(X, Y, w, rho, R1, Y1, T, 1, syms, blocks, sbi)
        fr, dr = rotated_form(X, Y, w, rho, R1, Y1, T, 1, syms, blocks, sbi, slip=True)
        assert rel(fr, fi) > 1e-3 or rel(dr, di) > 1e-3
