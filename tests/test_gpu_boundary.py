"""GPU: the drop-in boundary's partition, K-split and caching semantics.

* Image ranges with the index semantics of get_perturb_averages_sym
  (Modules/tdscha_core.jl:164-238): each partial equals the oracle's partial
  over the same images, normalised by the global N_eff * N_sym, and the
  partials of the rank split of Modules/DynamicalLanczos.py:3141-3164 add up
  to the full call (the reference's partition-invariance test,
  tests/test_mpi_parallel.py:115-123).
* The FP64 result does not depend on the K split of the outer GEMM.
* GetPerturbAverageSym returns fresh results after the ensemble or the
  symmetries change (in place or not).
* A device Lanczos chain that stops on |b| < tol reports one b fewer than a.
"""
import numpy as np
import pytest

from tests.systems import load_snte, random_block_system, random_perturbation

pytestmark = pytest.mark.gpu
TOL = 1e-12


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def rank_split(n_total, n_proc):
    """[start, end] (1-based, inclusive) per rank, as DL:3145-3158."""
    count, rem = divmod(n_total, n_proc)
    out = []
    for rank in range(n_proc):
        if rank < rem:
            start = rank * (count + 1)
            stop = start + count + 1
        else:
            start = rank * count + rem
            stop = start + count
        out.append((start + 1, stop))
    return out


@pytest.fixture(scope="module")
def snte():
    return load_snte()


def _systems(snte):
    ms = snte["syms_full"]
    yield ("snte384", snte["X"], snte["Y"], snte["w"], snte["rho"], 250.0, ms.symmetries,
           ms.degenerate_space, ms.sym_block_id)
    s = random_block_system(130, 301, 6, [3, 5, 2] + [1] * 120, seed=41)
    yield ("random130", s["X"], s["Y"], s["w"], s["rho"], s["T"], s["syms"], s["blocks"], s["sbi"])


@pytest.mark.parametrize("n_proc", [4, 7])
def test_image_range_partials_sum_to_full(snte, n_proc):
    from oracle import c_oracle as CO
    from tdscha_amd.engine import Engine
    for name, X, Y, w, rho, T, syms, blocks, sbi in _systems(snte):
        n = X.shape[1]
        eng = Engine(X, Y, w, rho, T, syms, blocks, sbi)
        R1, Y1 = random_perturbation(n, seed=12)
        Y1 = Y1 * 1e-3
        f_full, d_full = eng.perturb_averages(R1, Y1, 1)
        n_img = eng.n_images
        fs, ds = np.zeros(n), np.zeros((n, n))
        for start, end in rank_split(n_img, n_proc):
            f, d = eng.perturb_averages_range(R1, Y1, start, end, 1)
            fo, do = CO.oracle_perturb_averages(X, Y, w, rho, R1, Y1, T, 1, syms, blocks, sbi,
                                                b_begin=start - 1, b_end=end, n_threads=8)
            ef, ed = rel(f, fo), rel(d, do)
            print("%s range [%d, %d] of %d: rel f %.2e d2v %.2e" % (name, start, end, n_img,
                                                                   ef, ed))
            assert ef < TOL and ed < TOL
            fs += f
            ds += d
        ef, ed = rel(fs, f_full), rel(ds, d_full)
        print("%s: %d-way sum vs full call: f %.2e d2v %.2e" % (name, n_proc, ef, ed))
        assert ef < 1e-13 and ed < 1e-13
        eng.close()


def test_image_range_single_images_and_split_precision(snte):
    from oracle import c_oracle as CO
    from tdscha_amd.engine import Engine
    ms = snte["syms_full"]
    X, Y, w, rho = snte["X"], snte["Y"], snte["w"], snte["rho"]
    eng = Engine(X, Y, w, rho, 250.0, ms.symmetries, ms.degenerate_space, ms.sym_block_id)
    R1, Y1 = random_perturbation(45, seed=13)
    for start, end in [(1, 1), (385, 385), (3840, 3840), (17, 1000)]:
        f, d = eng.perturb_averages_range(R1, Y1, start, end, 1)
        fo, do = CO.oracle_perturb_averages(X, Y, w, rho, R1, Y1, 250.0, 1, ms.symmetries,
                                            ms.degenerate_space, ms.sym_block_id,
                                            b_begin=start - 1, b_end=end)
        assert rel(f, fo) < TOL and rel(d, do) < TOL
    # split precision honours the range too (its own tolerance)
    f, d = eng.perturb_averages_range(R1, Y1, 100, 2900, 1, precision="fp16x2")
    fo, do = CO.oracle_perturb_averages(X, Y, w, rho, R1, Y1, 250.0, 1, ms.symmetries,
                                        ms.degenerate_space, ms.sym_block_id,
                                        b_begin=99, b_end=2900)
    print("fp16x2 range: rel f %.2e d2v %.2e" % (rel(f, fo), rel(d, do)))
    assert rel(f, fo) < 2e-5 and rel(d, do) < 2e-5
    eng.close()


def test_image_range_argument_errors(snte):
    from tdscha_amd._lib import TdschaError
    from tdscha_amd.engine import Engine
    ms = snte["syms_full"]
    eng = Engine(snte["X"], snte["Y"], snte["w"], snte["rho"], 250.0, ms.symmetries,
                 ms.degenerate_space, ms.sym_block_id)
    R1, Y1 = random_perturbation(45, seed=14)
    for start, end in [(0, 10), (20, 10), (1, 3841)]:
        with pytest.raises(TdschaError, match="code 2"):
            eng.perturb_averages_range(R1, Y1, start, end, 1)
    eng.close()


def test_ksplit_invariance(snte):
    from oracle import c_oracle as CO
    from tdscha_amd.engine import Engine
    s = random_block_system(130, 1200, 6, [3, 5, 2] + [1] * 120, seed=42)
    eng = Engine(s["X"], s["Y"], s["w"], s["rho"], s["T"], s["syms"], s["blocks"], s["sbi"])
    R1, Y1 = random_perturbation(130, seed=15)
    Y1 = Y1 * 1e-3
    fo, do = CO.oracle_perturb_averages(s["X"], s["Y"], s["w"], s["rho"], R1, Y1, s["T"], 1,
                                        s["syms"], s["blocks"], s["sbi"], n_threads=8)
    default = eng.ksplit
    results = []
    for ks in (8, 16, 64, 256, default):
        eng.set_ksplit(ks)
        f, d = eng.perturb_averages(R1, Y1, 1)
        print("ksplit %d: rel f %.2e d2v %.2e" % (ks, rel(f, fo), rel(d, do)))
        assert rel(f, fo) < TOL and rel(d, do) < TOL
        results.append((f, d))
    for f, d in results[:-1]:
        assert rel(f, results[-1][0]) < 1e-14 and rel(d, results[-1][1]) < 1e-14
    eng.close()


def test_drop_in_cache_sees_changed_inputs():
    from oracle import c_oracle as CO
    from tdscha_amd.engine import GetPerturbAverageSym
    s = random_block_system(45, 200, 4, [3, 3, 3] + [1] * 36, seed=43)
    s2 = random_block_system(45, 200, 4, [3, 3, 3] + [1] * 36, seed=44)
    X, Y, w, rho = s["X"], s["Y"], s["w"], s["rho"].copy()
    n_deg = np.array([len(b) for b in s["blocks"]], dtype=np.intc)
    R1, Y1 = random_perturbation(45, seed=16)
    f = np.zeros(45)
    d = np.zeros((45, 45))

    def check(syms):
        GetPerturbAverageSym(X, Y, w, rho, R1, Y1, s["T"], 1, 4, syms, n_deg, s["blocks"],
                             s["sbi"], f, d)
        fo, do = CO.oracle_perturb_averages(X, Y, w, rho, R1, Y1, s["T"], 1, syms, s["blocks"],
                                            s["sbi"])
        assert rel(f, fo) < TOL and rel(d, do) < TOL

    check(s["syms"])
    rho[:50] *= 3.0           # in place: same buffer address
    check(s["syms"])
    X[7] += 0.5               # in place
    check(s["syms"])
    check(s2["syms"])         # other symmetry matrices, same n_syms and block layout


def test_preconditioner_and_linear_operators(snte):
    """apply_L1_inverse_FT on the device against the oracle (with the
    reference's regularised pairs: the SnTe frequencies at 250 K hit both
    branches), and L_linop / M_linop as the reference builds them
    (DL:478-489), used the way its static solver does (DL:3935-3945)."""
    import scipy.sparse.linalg as SLA
    from oracle import lanczos_oracle as O
    from tests.test_gpu_lanczos import _product_from_fixture
    lanc = _product_from_fixture(snte, use_wigner=False)
    rng = np.random.default_rng(23)
    v = rng.normal(size=len(lanc.psi))
    for transpose in (False, True):
        got = lanc.apply_L1_inverse_FT(v, transpose=transpose)
        ref = O.harmonic_L1_inverse(v, lanc.w, lanc.T, transpose=transpose)
        assert rel(got, ref) < 1e-13
    assert rel(lanc.M_linop.matvec(v), O.harmonic_L1_inverse(v, lanc.w, lanc.T)) < 1e-13
    assert rel(lanc.M_linop.rmatvec(v),
               O.harmonic_L1_inverse(v, lanc.w, lanc.T, transpose=True)) < 1e-13
    assert rel(lanc.L_linop.matvec(v), lanc.apply_full_L(v.copy())) < 1e-14
    assert rel(lanc.L_linop.rmatvec(v), lanc.apply_full_L(v.copy(), transpose=True)) < 1e-14
    # SciPy drives them as operators (the consumers at DL:3935-3966 do)
    B = rng.normal(size=(len(v), 2))
    LB = lanc.L_linop.matmat(B)
    for k in range(2):
        assert rel(LB[:, k], lanc.apply_full_L(B[:, k].copy())) < 1e-14
    assert isinstance(lanc.M_linop * lanc.L_linop, SLA.LinearOperator)
    lanc.close()


def test_stopped_chain_reports_one_b_fewer():
    """A chain that exhausts its Krylov space stops on |b| < tol with
    len(b) == len(a) - 1, whatever the device memory held before."""
    from tdscha_amd.lanczos import Lanczos
    from tdscha_amd.rocksalt import Ensemble, RockSalt
    ens = Ensemble(RockSalt(L=2), 300.0, 200, seed=5)
    lz = Lanczos(ens)
    lz.init(use_symmetries=False)
    n = lz.n_modes
    # a full run first leaves large coefficients in freed device memory
    rng = np.random.default_rng(17)
    psi = rng.normal(size=(3, len(lz.psi)))
    lz.run_FT_batched(psi, 12, reorth=True)
    lz.ignore_v3 = True       # harmonic L: diagonal, a single mode is invariant
    for k in (0, 7, n - 1):
        psi = np.zeros((1, len(lz.psi)))
        psi[0, k] = 1.0
        chains, _ = lz.run_FT_batched(psi, 12, reorth=True)
        a, b = chains[0]
        assert len(a) == 1 and len(b) == 0, (len(a), len(b), b)
        assert np.isfinite(a).all()
        assert abs(a[0] + lz.w[k] ** 2) <= 1e-12 * lz.w[k] ** 2
    # a two-mode start spans a two-dimensional invariant space
    psi = np.zeros((1, len(lz.psi)))
    psi[0, 3] = 1.0
    psi[0, 9] = 2.0
    chains, _ = lz.run_FT_batched(psi, 12, reorth=True)
    a, b = chains[0]
    assert len(a) == 2 and len(b) == 1 and np.isfinite(b).all() and abs(b[0]) > 0
    lz.close()
