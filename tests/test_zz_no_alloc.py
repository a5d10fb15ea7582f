"""GPU: the entry point that allocates nothing, tdscha_perturb_averages_noalloc,
against the context entry tdscha_perturb_averages_range on the same inputs
(FP64, <= 1e-12), two calls in a row, the second with a smaller input."""
import ctypes as C

import numpy as np
import pytest

from tests.systems import random_block_system, random_perturbation

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def packed(s):
    """Symmetry tables in the packed layout of Modules/DynamicalLanczos.py:807-822."""
    syms, blocks = s["syms"], s["blocks"]
    nb = len(blocks)
    n_syms = syms[0].shape[0]
    dmax = max(len(b) for b in blocks)
    sp = np.zeros((nb, n_syms, dmax, dmax))
    dg = np.zeros((nb, dmax), dtype=np.int32)
    for i, (sm, b) in enumerate(zip(syms, blocks)):
        d = len(b)
        sp[i, :, :d, :d] = sm
        dg[i, :d] = b
    return sp, np.array([len(b) for b in blocks], dtype=np.int32), dg, dmax


def noalloc_call(s, R1, Y1, apply_d4, start, end):
    from tdscha_amd import _lib
    from tdscha_amd._lib import DeviceBuffer
    fn = _lib.lib().tdscha_perturb_averages_noalloc
    N, n = s["X"].shape
    sp, nd, dg, dmax = packed(s)
    bufs = {k: DeviceBuffer.from_array(v) for k, v in dict(
        X=s["X"], Y=s["Y"], w=s["w"], rho=s["rho"], R1=R1, Y1=Y1, sp=sp, nd=nd, dg=dg,
        sbi=np.asarray(s["sbi"], np.int32)).items()}
    fo, do = DeviceBuffer((n,)), DeviceBuffer((n, n))
    print("noalloc call: n=%d N=%d n_syms=%d range [%d, %d]" % (n, N, sp.shape[1], start, end),
          flush=True)
    try:
        rc = fn(n, N, bufs["X"].ptr, bufs["Y"].ptr, bufs["w"].ptr, bufs["rho"].ptr,
                bufs["R1"].ptr, bufs["Y1"].ptr, float(s["T"]), int(apply_d4), sp.shape[1],
                len(s["blocks"]), dmax, bufs["sp"].ptr, bufs["nd"].ptr, bufs["dg"].ptr,
                bufs["sbi"].ptr, int(start), int(end), fo.ptr, do.ptr, None)
        _lib.check(rc, "tdscha_perturb_averages_noalloc")
        return fo.numpy(), do.numpy()   # synchronous copies after the null-stream work
    finally:
        for b in list(bufs.values()) + [fo, do]:
            b.free()


def test_noalloc_matches_range_entry_twice():
    from tdscha_amd.engine import Engine
    cases = [
        # 48000 images in several scratch chunks, a range that starts mid-configuration
        (random_block_system(130, 2000, 24, [3, 5, 2] + [1] * 120, seed=51), 1, 17, 47000),
        # a smaller input next: fewer modes, configurations and images
        (random_block_system(45, 300, 4, [3, 3, 3] + [1] * 36, seed=52), 0, 1, 1200),
    ]
    for s, d4, start, end in cases:
        n = s["X"].shape[1]
        R1, Y1 = random_perturbation(n, seed=n)
        Y1 = Y1 * 1e-3
        f, d = noalloc_call(s, R1, Y1, d4, start, end)
        eng = Engine(s["X"], s["Y"], s["w"], s["rho"], s["T"], s["syms"], s["blocks"], s["sbi"])
        fr, dr = eng.perturb_averages_range(R1, Y1, start, end, d4)
        eng.close()
        print("n=%d range [%d, %d]: rel f %.2e d2v %.2e" % (n, start, end, rel(f, fr),
                                                            rel(d, dr)), flush=True)
        assert rel(f, fr) < 1e-12 and rel(d, dr) < 1e-12
        assert np.max(np.abs(d - d.T)) <= 1e-14 * np.max(np.abs(d))


def test_noalloc_rejects_bad_arguments():
    from tdscha_amd import _lib
    from tdscha_amd._lib import TdschaError
    s = random_block_system(45, 30, 2, [1] * 45, seed=53)
    R1, Y1 = random_perturbation(45, seed=1)
    for start, end in [(0, 10), (11, 10), (1, 61)]:
        with pytest.raises(TdschaError, match="code 2"):
            noalloc_call(s, R1, Y1, 1, start, end)
    # a mode count beyond what the scratch holds
    L = _lib.lib()
    rc = L.tdscha_perturb_averages_noalloc(1000, 1, *([C.c_void_p(8)] * 6), 0.0, 1, 1, 1000, 1,
                                           *([C.c_void_p(8)] * 4), 1, 1, C.c_void_p(8),
                                           C.c_void_p(8), None)
    assert rc == 2 and b"scratch" in L.tdscha_last_error()
