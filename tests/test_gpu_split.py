"""Split-precision (FP16 / BF16 hi-lo on MFMA, FP32 accumulation) paths:
error against the FP64 oracle, batching, and the spectrum of split chains."""
import numpy as np
import pytest

from tests.systems import load_snte, random_block_system, random_perturbation

pytestmark = pytest.mark.gpu

# Tolerances (norm-wise relative to the FP64 oracle) for one perturbed average:
# fp16x2 products carry ~2^-22 per element plus FP32 accumulation over the
# images; bf16x2 ~2^-16 per element.
TOL = {"fp16x2": 2e-5, "bf16x2": 2e-3}


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.parametrize("prec", ["fp16x2", "bf16x2"])
@pytest.mark.parametrize("n,N,blocks,nsym", [
    (45, 2000, [1] * 45, 1),
    (130, 300, [3, 5, 2] + [1] * 120, 6),
    # n a multiple of 128: no padding column, the row-dot u.(fY R1) in the epilogue
    (128, 600, [4, 4] + [1] * 120, 3),
])
def test_split_perturbed_averages_error(prec, n, N, blocks, nsym):
    from oracle import c_oracle as CO
    from tdscha_amd.engine import Engine
    s = random_block_system(n, N, nsym, blocks, seed=n + N)
    eng = Engine(s["X"], s["Y"], s["w"], s["rho"], s["T"], s["syms"], s["blocks"], s["sbi"])
    R1, Y1 = random_perturbation(n, seed=11)
    Y1 *= 1e-3
    f, d = eng.perturb_averages_batched(R1[None], Y1[None], 1, prec)
    fo, do = CO.oracle_perturb_averages(s["X"], s["Y"], s["w"], s["rho"], R1, Y1, s["T"], 1,
                                        s["syms"], s["blocks"], s["sbi"], n_threads=8)
    ef, ed = rel(f[0], fo), rel(d[0], do)
    print("%s n=%d N=%d nsym=%d rel f %.2e d2v %.2e" % (prec, n, N, nsym, ef, ed))
    assert ef < TOL[prec] and ed < TOL[prec]
    eng.close()


@pytest.mark.parametrize("prec", ["fp16x2", "bf16x2"])
def test_split_snte_and_chain_groups(prec):
    """384-image SnTe ensemble; 6 chains = one 4-chain pass + one 2-chain pass,
    equal to single-chain calls to rounding of the FP64 reductions."""
    from oracle import c_oracle as CO
    from tdscha_amd.engine import Engine
    d = load_snte()
    ms = d["syms_full"]
    eng = Engine(d["X"], d["Y"], d["w"], d["rho"], 250.0, ms.symmetries, ms.degenerate_space,
                 ms.sym_block_id)
    R, Ys = [], []
    for seed in range(6):
        r, y = random_perturbation(45, seed=20 + seed)
        R.append(r)
        Ys.append(y)
    R, Ys = np.array(R), np.array(Ys)
    fb, db = eng.perturb_averages_batched(R, Ys, 1, prec)
    for c in (0, 3, 5):
        f1, d1 = eng.perturb_averages_batched(R[c:c + 1], Ys[c:c + 1], 1, prec)
        assert rel(fb[c], f1[0]) < 1e-6 and rel(db[c], d1[0]) < 1e-6
        fo, do = CO.oracle_perturb_averages(d["X"], d["Y"], d["w"], d["rho"], R[c], Ys[c], 250.0,
                                            1, ms.symmetries, ms.degenerate_space,
                                            ms.sym_block_id, n_threads=4)
        print("%s snte chain %d rel f %.2e d2v %.2e" % (prec, c, rel(fb[c], fo), rel(db[c], do)))
        assert rel(fb[c], fo) < TOL[prec] and rel(db[c], do) < TOL[prec]
    eng.close()


def test_split_lanczos_spectrum_error():
    """16-atom rock salt, N=2000, point group + translations, 3 IR chains,
    60 steps: fp16x2 chains vs FP64 chains on the 2000-point 0-600 cm^-1
    grid with 5 cm^-1 broadening.  Gate: L-inf relative error <= 1e-4."""
    from tdscha_amd.lanczos import Lanczos, continued_fraction
    from tdscha_amd.rocksalt import Ensemble, RockSalt
    ens = Ensemble(RockSalt(L=2), 300.0, 2000, seed=0)
    lanc = Lanczos(ens)
    lanc.gamma_only = True
    lanc.init(use_symmetries=True)
    psis = []
    for pv in ((1, 0, 0), (0, 1, 0), (0, 0, 1)):
        lanc.prepare_perturbation(ens.ir_vector(pv), masses_exp=-1)
        psis.append(lanc.psi.copy())
    psis = np.array(psis)
    ry = 109737.31568527
    wg = np.linspace(0, 600, 2000) / ry
    ch64, _ = lanc.run_FT_batched(psis, 60, reorth=True, precision="fp64")
    ch16, _ = lanc.run_FT_batched(psis, 60, reorth=True, precision="fp16x2")
    errs = []
    for (a1, b1), (a0, b0) in zip(ch16, ch64):
        s1 = -np.imag(continued_fraction(a1, b1, None, wg, True, False, 1, 5 / ry))
        s0 = -np.imag(continued_fraction(a0, b0, None, wg, True, False, 1, 5 / ry))
        errs.append(np.max(np.abs(s1 - s0)) / np.max(np.abs(s0)))
        # first 50 coefficients to the split path's accuracy
        print("a rel %.2e b rel %.2e" % (np.max(np.abs(a1[:50] / a0[:50] - 1)),
                                         np.max(np.abs(b1[:50] / b0[:50] - 1))))
    print("spectrum L-inf rel errors", errs)
    assert max(errs) < 1e-4
    lanc.close()
