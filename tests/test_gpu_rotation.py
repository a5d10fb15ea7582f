"""GPU: the operator-rotation form (TDSCHA_CTX_ROTATE_OPERATORS).

The context keeps the configurations only and applies each symmetry by
rotating the perturbation and the results (identity pinned on the oracle in
tests/test_rotation_form.py).  Checked here against the C oracle's image
average and against the image form on the device:
* FP64 averages of 1-8 chains, with symmetry counts that make the 8-pair
  passes span chains (n_syms = 6) and the SnTe space group (384), on blocks
  of consecutive modes of up to 24 (tiled rotation kernels) and on blocks of
  25 modes or of scattered modes (general rotation kernels);
* split precision within its tolerance;
* image ranges of whole configurations, and the error for partial ones;
* whole L applications and device Lanczos chains (16-atom rock salt, 48
  point-group images and the translational projector) against the image
  form, and the whole run captured as one graph against plain launches.
"""
import ctypes as C

import numpy as np
import pytest

from tests.systems import load_snte, random_block_system, random_perturbation

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.fixture(scope="module")
def snte():
    return load_snte()


def _relabelled(s, seed):
    """The same system with its modes relabelled by a random permutation, so
    that its degenerate blocks are no longer runs of consecutive modes (the
    general rotation kernels instead of the tiled ones)."""
    n = s["X"].shape[1]
    perm = np.random.default_rng(seed).permutation(n)   # mode m -> perm[m]
    X, Y = np.empty_like(s["X"]), np.empty_like(s["Y"])
    X[:, perm], Y[:, perm] = s["X"], s["Y"]
    w = np.empty_like(s["w"])
    w[perm] = s["w"]
    sbi = np.empty_like(s["sbi"])
    sbi[perm] = s["sbi"]
    blocks = [perm[b].astype(np.int32) for b in s["blocks"]]
    assert not all(np.all(np.diff(b) == 1) for b in blocks)
    return dict(s, X=X, Y=Y, w=w, sbi=sbi, blocks=blocks)


def _systems(snte):
    ms = snte["syms_full"]
    yield ("snte", snte["X"], snte["Y"], snte["w"], snte["rho"], float(snte["T"]),
           ms.symmetries, ms.degenerate_space, ms.sym_block_id)
    s = random_block_system(130, 700, 6, [3, 5, 2] + [1] * 120, seed=21)
    yield "random130", s["X"], s["Y"], s["w"], s["rho"], s["T"], s["syms"], s["blocks"], s["sbi"]
    # blocks of 24 modes, the largest the tiled kernels take, next to a 25-mode
    # block (general kernels)
    s = random_block_system(200, 400, 3, [24, 1, 24, 7] + [24] * 6, seed=24)
    yield "blocks24", s["X"], s["Y"], s["w"], s["rho"], s["T"], s["syms"], s["blocks"], s["sbi"]
    s = random_block_system(200, 400, 3, [25, 1, 24, 6] + [24] * 6, seed=25)
    yield "blocks25", s["X"], s["Y"], s["w"], s["rho"], s["T"], s["syms"], s["blocks"], s["sbi"]
    # n a multiple of 128: no padding panel
    s = random_block_system(128, 400, 4, [24] * 5 + [5, 3], seed=27)
    yield "n128", s["X"], s["Y"], s["w"], s["rho"], s["T"], s["syms"], s["blocks"], s["sbi"]
    s = _relabelled(random_block_system(130, 700, 6, [3, 5, 2] + [1] * 120, seed=26), 7)
    yield ("random130-relabelled", s["X"], s["Y"], s["w"], s["rho"], s["T"], s["syms"],
           s["blocks"], s["sbi"])


def test_rotated_fp64_matches_oracle_and_images(snte):
    from oracle import c_oracle as CO
    from tdscha_amd.engine import Engine
    for name, X, Y, w, rho, T, syms, blocks, sbi in _systems(snte):
        n = X.shape[1]
        rot = Engine(X, Y, w, rho, T, syms, blocks, sbi, rotate=True)
        img = Engine(X, Y, w, rho, T, syms, blocks, sbi)
        assert rot.n_images == img.n_images and rot.n_images_padded < img.n_images_padded
        for nc in (1, 3, 5, 8):
            R1 = np.stack([random_perturbation(n, seed=40 + k)[0] for k in range(nc)])
            Y1 = np.stack([random_perturbation(n, seed=40 + k)[1] for k in range(nc)]) * 1e-3
            fr, dr = rot.perturb_averages_batched(R1, Y1, 1)
            fi, di = img.perturb_averages_batched(R1, Y1, 1)
            for k in (0, nc - 1):
                fo, do = CO.oracle_perturb_averages(X, Y, w, rho, R1[k], Y1[k], T, 1, syms, blocks,
                                                    sbi, n_threads=8)
                print("%s nc=%d chain %d: rotated f %.2e d2v %.2e, images f %.2e d2v %.2e" % (
                    name, nc, k, rel(fr[k], fo), rel(dr[k], do), rel(fi[k], fo), rel(di[k], do)))
                assert rel(fr[k], fo) < 1e-12 and rel(dr[k], do) < 1e-12
            assert rel(fr, fi) < 1e-12 and rel(dr, di) < 1e-12
        rot.close()
        img.close()


@pytest.mark.parametrize("prec,tol", [("fp16x2", 2e-5), ("bf16x2", 2e-3)])
def test_rotated_split_precision(snte, prec, tol):
    from oracle import c_oracle as CO
    from tdscha_amd.engine import Engine
    s = random_block_system(130, 700, 6, [3, 5, 2] + [1] * 120, seed=22)
    rot = Engine(s["X"], s["Y"], s["w"], s["rho"], s["T"], s["syms"], s["blocks"], s["sbi"],
                 rotate=True)
    R1 = np.stack([random_perturbation(130, seed=50 + k)[0] for k in range(3)])
    Y1 = np.stack([random_perturbation(130, seed=50 + k)[1] for k in range(3)]) * 1e-3
    f, d = rot.perturb_averages_batched(R1, Y1, 1, prec)
    for k in range(3):
        fo, do = CO.oracle_perturb_averages(s["X"], s["Y"], s["w"], s["rho"], R1[k], Y1[k], s["T"],
                                            1, s["syms"], s["blocks"], s["sbi"], n_threads=8)
        print("%s chain %d: f %.2e d2v %.2e" % (prec, k, rel(f[k], fo), rel(d[k], do)))
        assert rel(f[k], fo) < tol and rel(d[k], do) < tol
    rot.close()


def test_rotated_ranges_of_whole_configurations(snte):
    from oracle import c_oracle as CO
    from tdscha_amd import _lib
    from tdscha_amd.engine import Engine
    s = random_block_system(130, 700, 6, [3, 5, 2] + [1] * 120, seed=23)
    rot = Engine(s["X"], s["Y"], s["w"], s["rho"], s["T"], s["syms"], s["blocks"], s["sbi"],
                 rotate=True)
    R1, Y1 = random_perturbation(130, seed=61)
    Y1 = Y1 * 1e-3
    f_full, d_full = rot.perturb_averages(R1, Y1, 1)
    fs, ds = np.zeros(130), np.zeros((130, 130))
    for c0, c1 in [(0, 250), (250, 251), (251, 700)]:
        start, end = c0 * 6 + 1, c1 * 6
        f, d = rot.perturb_averages_range(R1, Y1, start, end, 1)
        fo, do = CO.oracle_perturb_averages(s["X"], s["Y"], s["w"], s["rho"], R1, Y1, s["T"], 1,
                                            s["syms"], s["blocks"], s["sbi"], b_begin=start - 1,
                                            b_end=end, n_threads=8)
        assert rel(f, fo) < 1e-12 and rel(d, do) < 1e-12
        fs += f
        ds += d
    assert rel(fs, f_full) < 1e-13 and rel(ds, d_full) < 1e-13
    with pytest.raises(_lib.TdschaError, match="whole"):
        rot.perturb_averages_range(R1, Y1, 2, 600, 1)
    rot.close()


def _rocksalt_lanczos(form):
    from tdscha_amd.lanczos import Lanczos
    from tdscha_amd.rocksalt import Ensemble, RockSalt
    ens = Ensemble(RockSalt(L=2), 300.0, 400, seed=4)
    lanc = Lanczos(ens)
    lanc.gamma_only = True
    lanc.symmetry_form = form
    lanc.init(use_symmetries=True)
    assert lanc.engine.rotate == (form == "operators") and lanc.n_syms == 48
    return ens, lanc


def test_rotated_L_and_lanczos_equal_the_image_form():
    ens, rot = _rocksalt_lanczos("operators")
    _, img = _rocksalt_lanczos("images")
    rng = np.random.default_rng(3)
    V = rng.normal(size=(3, len(rot.psi)))
    Lr, Li = rot.apply_L_batch(V), img.apply_L_batch(V)
    print("L: rotated vs images %.2e" % rel(Lr, Li))
    assert rel(Lr, Li) < 1e-12
    psis = []
    for d in ((1.0, 0, 0), (0, 1.0, 0), (0, 0, 1.0)):
        rot.prepare_perturbation(ens.ir_vector(d), masses_exp=-1)
        psis.append(rot.psi.copy())
    psis = np.array(psis)
    run, _ = rot.run_FT_batched(psis, 30, reorth=True, use_graph="run")
    plain, _ = rot.run_FT_batched(psis, 30, reorth=True, use_graph=False)
    ref, _ = img.run_FT_batched(psis, 30, reorth=True, use_graph="run")
    for c in range(3):
        assert np.array_equal(run[c][0], plain[c][0]) and np.array_equal(run[c][1], plain[c][1])
        # the IR chains of this small model nearly close their Krylov space
        # after about 9 steps; compare the well-conditioned start
        k = min(len(run[c][0]), len(ref[c][0]), 8)
        ea = np.max(np.abs(run[c][0][:k] - ref[c][0][:k]) / np.abs(ref[c][0][:k]))
        eb = np.max(np.abs(run[c][1][:k] - ref[c][1][:k]) / np.abs(ref[c][1][:k]))
        print("chain %d: first %d coefficients rotated vs images: a %.2e b %.2e" % (c, k, ea, eb))
        assert ea < 1e-9 and eb < 1e-9
    rot.close()
    img.close()
