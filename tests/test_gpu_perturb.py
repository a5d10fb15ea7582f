"""GPU parity: HIP perturbed averages vs the CPU oracle (FP64, <= 1e-12)."""
import numpy as np
import pytest

from tests.systems import load_snte, random_block_system, random_perturbation

pytestmark = pytest.mark.gpu
TOL = 1e-12  # norm-wise relative, FP64 path vs FP64 oracle


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.fixture(scope="module")
def snte():
    return load_snte()


@pytest.mark.parametrize("apply_d4", [0, 1])
def test_snte_full_group(snte, apply_d4):
    from oracle import c_oracle as CO
    from tdscha_amd.engine import Engine
    ms = snte["syms_full"]
    eng = Engine(snte["X"], snte["Y"], snte["w"], snte["rho"], 250.0, ms.symmetries,
                 ms.degenerate_space, ms.sym_block_id)
    for seed in (3, 4):
        R1, Y1 = random_perturbation(45, seed=seed)
        f, d = eng.perturb_averages(R1, Y1, apply_d4)
        fo, do = CO.oracle_perturb_averages(snte["X"], snte["Y"], snte["w"], snte["rho"], R1, Y1,
                                            250.0, apply_d4, ms.symmetries, ms.degenerate_space,
                                            ms.sym_block_id, n_threads=4)
        print("snte d4=%d seed=%d rel f %.2e d2v %.2e" % (apply_d4, seed, rel(f, fo), rel(d, do)))
        assert rel(f, fo) < TOL and rel(d, do) < TOL
    eng.close()


@pytest.mark.parametrize("n,N,blocks,nsym", [
    (45, 2000, [1] * 45, 1),                       # 16 atoms, N = 2000, no symmetry
    (130, 300, [3, 5, 2] + [1] * 120, 6),          # two column tiles + padding
    (256, 97, [8, 8] + [4] * 60, 3),               # exact tile multiple, odd N
])
def test_random_systems(n, N, blocks, nsym):
    from oracle import c_oracle as CO
    from tdscha_amd.engine import Engine
    s = random_block_system(n, N, nsym, blocks, seed=n + N)
    eng = Engine(s["X"], s["Y"], s["w"], s["rho"], s["T"], s["syms"], s["blocks"], s["sbi"])
    R1, Y1 = random_perturbation(n, seed=11)
    Y1 = Y1 * 1e-3
    f, d = eng.perturb_averages(R1, Y1, 1)
    fo, do = CO.oracle_perturb_averages(s["X"], s["Y"], s["w"], s["rho"], R1, Y1, s["T"], 1,
                                        s["syms"], s["blocks"], s["sbi"], n_threads=8)
    print("n=%d N=%d nsym=%d rel f %.2e d2v %.2e" % (n, N, nsym, rel(f, fo), rel(d, do)))
    assert rel(f, fo) < TOL and rel(d, do) < TOL
    assert np.max(np.abs(d - d.T)) <= 1e-14 * np.max(np.abs(d))
    eng.close()


def test_batched_entries_match_single_calls(snte):
    from tdscha_amd._lib import DeviceBuffer
    from tdscha_amd.engine import Engine
    ms = snte["syms_full"]
    eng = Engine(snte["X"], snte["Y"], snte["w"], snte["rho"], 250.0, ms.symmetries,
                 ms.degenerate_space, ms.sym_block_id)
    R, Ys = [], []
    for seed in (5, 6, 7):
        r, y = random_perturbation(45, seed=seed)
        R.append(r)
        Ys.append(y)
    R, Ys = np.array(R), np.array(Ys)
    fb, db = eng.perturb_averages_batched(R, Ys, 1, 0)
    dR, dY = DeviceBuffer.from_array(R), DeviceBuffer.from_array(Ys)
    df, dd = DeviceBuffer((3, 45)), DeviceBuffer((3, 45, 45))
    eng.perturb_averages_dev(3, dR, dY, df, dd, 1, 0)
    eng.synchronize()
    fd, ddv = df.numpy(), dd.numpy()
    for c in range(3):
        f1, d1 = eng.perturb_averages(R[c], Ys[c], 1)
        assert rel(fb[c], f1) < 1e-14 and rel(db[c], d1) < 1e-14
        assert rel(fd[c], f1) < 1e-14 and rel(ddv[c], d1) < 1e-14
    eng.close()


def test_drop_in_binding_semantics(snte):
    """GetPerturbAverageSym argument list and in-place outputs."""
    from oracle import c_oracle as CO
    from tdscha_amd.engine import GetPerturbAverageSym
    ms = snte["syms_full"]
    X, Y, w, rho = snte["X"], snte["Y"], snte["w"], snte["rho"]
    R1, Y1 = random_perturbation(45, seed=8)
    f = np.zeros(45)
    d = np.full((45, 45), 7.0)
    GetPerturbAverageSym(X, Y, w, rho, R1, Y1, 250.0, 1, 384, ms.symmetries, ms.N_degeneracy,
                         ms.degenerate_space, ms.sym_block_id, f, d)
    fo, do = CO.oracle_perturb_averages(X, Y, w, rho, R1, Y1, 250.0, 1, ms.symmetries,
                                        ms.degenerate_space, ms.sym_block_id)
    assert rel(f, fo) < TOL and rel(d, do) < TOL
    with pytest.raises(ValueError):
        GetPerturbAverageSym(X, Y, w, rho, R1, Y1, 250.0, 2, 384, ms.symmetries, ms.N_degeneracy,
                             ms.degenerate_space, ms.sym_block_id, f, d)
