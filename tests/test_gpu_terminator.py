"""GPU parity of the terminator network kernels against the CPU oracle, and
the held-out gain of the trained terminator over the standard one.

Tolerances: the kernels compute the network in FP32 (the oracle in FP64), so
outputs and gradients are compared relative to their largest entry at 1e-4 to
2e-3; the spectral loss itself runs in FP64 on the device.
"""
import os

import numpy as np
import pytest

from oracle import mlp_oracle as MO

pytestmark = pytest.mark.gpu

RY_TO_CM = 109737.31568527
K = 40
DIMS = [2 * K, 128, 128, 128, 10]


def _net(params, max_batch=64):
    from tdscha_amd.terminator import Network
    return Network(2 * K, 128, 3, 10, params=params, max_batch=max_batch)


def _params(seed):
    from tdscha_amd.terminator import init_params
    p = init_params(DIMS, np.random.default_rng(seed))
    # a larger output layer than the training initialisation, so every layer's
    # gradient is well above FP32 rounding
    return (p * 3.0).astype(np.float32)


def _relerr(x, ref):
    return float(np.max(np.abs(x - ref)) / np.max(np.abs(ref)))


def test_forward_matches_oracle_ragged_batches():
    p = _params(0)
    net = _net(p, max_batch=64)
    rng = np.random.default_rng(1)
    for batch in (1, 37, 64, 150):          # 150 > max_batch: chunked
        x = rng.normal(size=(batch, 2 * K)).astype(np.float32)
        y = net.forward(x)
        yo, _, _ = MO.forward(p.astype(np.float64), DIMS, x.astype(np.float64))
        print("batch", batch, "rel err %.2e" % _relerr(y, yo))
        assert y.shape == (batch, 10)
        assert _relerr(y, yo) < 1e-4
    net.close()


def test_mse_gradient_matches_oracle():
    p = _params(2)
    net = _net(p)
    rng = np.random.default_rng(3)
    x = rng.normal(size=(50, 2 * K)).astype(np.float32)
    t = rng.normal(size=(50, 10)).astype(np.float32)
    w = np.linspace(0.5, 2.0, 10).astype(np.float32)
    net.set_mse_weights(w)
    loss, g = net.loss_grad(0, x, target=t)
    lo, go = MO.mse_loss_grad(p.astype(np.float64), DIMS, x.astype(np.float64),
                              t.astype(np.float64), w.astype(np.float64))
    print("loss %.6g oracle %.6g grad rel err %.2e" % (loss, lo, _relerr(g, go)))
    assert loss == pytest.approx(lo, rel=1e-4)
    assert _relerr(g, go) < 1e-3
    # every layer, not only the largest entries
    off = 0
    for l in range(len(DIMS) - 1):
        n = DIMS[l + 1] * DIMS[l] + DIMS[l + 1]
        assert _relerr(g[off:off + n], go[off:off + n]) < 2e-3, l
        off += n
    net.close()


def _cf_batch(rng, n, n_w=2000, learnable=False):
    w = np.linspace(0.0, 600.0, n_w) / RY_TO_CM
    eta = 5.0 / RY_TO_CM
    chains, refs = [], []
    for _ in range(n):
        b0 = 5.3e-7 * rng.uniform(0.5, 2.0)
        a = b0 * (-18.0 + rng.normal(size=K) * 2.0)
        b = b0 * (9.0 + rng.normal(size=K))
        a[0], b[0] = -6.0 * b0, b0
        tail = np.r_[rng.uniform(-20, -15), rng.uniform(8, 10), rng.normal(size=8) * 0.5]
        if learnable:     # a tail that the chain itself determines
            tail = np.r_[np.mean(a[-10:]) / b0 + 3.0, np.mean(b[-10:]) / b0 - 1.0, np.zeros(8)]
        ea, eb = MO.extend_chain(a, b, tail, b0)
        S = MO.spectrum(ea, eb, w, eta)
        inv_s = 1.0 / np.abs(S).max()
        chains.append(np.r_[a, b, b0, inv_s])
        refs.append(S * inv_s)
    return np.array(chains), np.array(refs), w, eta


def test_spectral_loss_gradient_matches_oracle():
    rng = np.random.default_rng(4)
    chains, refs, w, eta = _cf_batch(rng, 6)
    p = _params(5)
    net = _net(p)
    out_mean = np.r_[-17.5, 9.0, np.zeros(8)]
    out_std = np.r_[2.0, 0.7, np.full(8, 1.0)]
    net.set_cf_loss(K, w, eta, out_mean, out_std, 1e3)
    x = rng.normal(size=(6, 2 * K)).astype(np.float32)
    loss, g = net.loss_grad(1, x, chains=chains, ref=refs)
    lo, go = MO.cf_batch_loss_grad(p.astype(np.float64), DIMS, x.astype(np.float64), chains,
                                   refs, w, eta, out_mean, out_std, 1e3, K)
    print("cf loss %.8g oracle %.8g grad rel err %.2e" % (loss, lo, _relerr(g, go)))
    assert loss == pytest.approx(lo, rel=1e-4)
    assert _relerr(g, go) < 2e-3
    net.close()


def test_adam_step_matches_oracle():
    p = _params(6)
    net = _net(p)
    rng = np.random.default_rng(7)
    x = rng.normal(size=(32, 2 * K)).astype(np.float32)
    t = rng.normal(size=(32, 10)).astype(np.float32)
    lr = 1e-3
    _, go = MO.mse_loss_grad(p.astype(np.float64), DIMS, x.astype(np.float64),
                             t.astype(np.float64))
    net.adam_step(0, x, target=t, lr=lr, eps=1e-8)
    p1 = net.params()
    po = MO.Adam(len(p), lr, eps=1e-8).step(p.astype(np.float64), go)
    # the first Adam step moves each parameter by lr * sign(g); compare where
    # |g| is well above the FP32 rounding of the gradient
    sel = np.abs(go) > 1e-3 * np.abs(go).max()
    print("adam: %d of %d compared, max diff %.2e" % (sel.sum(), len(p), np.max(np.abs(
        p1[sel] - po[sel]))))
    assert sel.sum() > 0.5 * len(p)
    assert np.max(np.abs(p1[sel] - po[sel])) < 1e-3 * lr + 1e-6
    # a second step continues from the kept moments
    g2 = net.loss_grad(0, x, target=t)[1]
    opt = MO.Adam(len(p), lr, eps=1e-8)
    opt.step(p.astype(np.float64), go)
    _, go2 = MO.mse_loss_grad(p1.astype(np.float64), DIMS, x.astype(np.float64),
                              t.astype(np.float64))
    net.adam_step(0, x, target=t, lr=lr, eps=1e-8)
    p2 = net.params()
    q2 = opt.step(p1.astype(np.float64), go2)
    sel2 = sel & (np.abs(go2) > 1e-3 * np.abs(go2).max())
    assert np.max(np.abs(p2[sel2] - q2[sel2])) < 2e-3 * lr + 1e-6
    assert _relerr(g2, go2) < 2e-3
    net.close()


def test_training_reduces_loss_and_is_reproducible():
    rng = np.random.default_rng(8)
    x = rng.normal(size=(512, 2 * K)).astype(np.float32)
    Wt = rng.normal(size=(2 * K, 10)) / np.sqrt(2 * K)
    t = np.tanh(x @ Wt).astype(np.float32)
    p = _params(9) / 3.0
    runs = []
    for _ in range(2):
        net = _net(p, max_batch=64)
        hist = net.train(0, x, target=t, batch=64, epochs=30, lr=1e-3, seed=11)
        runs.append(net.params())
        net.close()
    print("loss %.4g -> %.4g" % (hist[0], hist[-1]))
    assert hist[-1] < 0.1 * hist[0]
    assert np.array_equal(runs[0], runs[1])


def test_spectral_training_reduces_loss():
    rng = np.random.default_rng(12)
    chains, refs, w, eta = _cf_batch(rng, 64, n_w=500, learnable=True)
    x = (np.concatenate([chains[:, :K], chains[:, K:2 * K]], 1) / chains[:, 2 * K:2 * K + 1])
    x = ((x - x.mean(0)) / (x.std(0) + 1e-6)).astype(np.float32)
    net = _net(_params(13) / 3.0)
    net.set_cf_loss(K, w, eta, np.r_[-17.5, 9.0, np.zeros(8)], np.r_[2.0, 1.0, np.ones(8)],
                    1e3)
    hist = net.train(1, x, chains=chains, ref=refs, batch=16, epochs=40, lr=1e-3, seed=3)
    print("spectral loss %.4g -> %.4g" % (hist[0], hist[-1]))
    assert hist[-1] < 0.5 * hist[0]
    net.close()


@pytest.fixture(scope="module")
def heldout_chains():
    """Held-out potentials (seeds outside the training streams): three
    1000-step FP64 chains each and the spectra they give."""
    from tdscha_amd.lanczos import Lanczos, continued_fraction
    from tdscha_amd.rocksalt import Ensemble, RockSalt
    rng = np.random.default_rng(31337)
    w = np.linspace(0.0, 600.0, 2000) / RY_TO_CM
    eta = 5.0 / RY_TO_CM
    out = []
    for i in range(8):
        k3, k4, T = rng.uniform(-4, 0), rng.uniform(2, 10), rng.uniform(50, 600)
        ens = Ensemble(RockSalt(L=2, k3=k3, k4=k4), T, 2000, seed=7_000_000 + i)
        lanc = Lanczos(ens)
        lanc.init(use_symmetries=False)
        psis = []
        for pv in ((1, 0, 0), (0, 1, 0), (0, 0, 1)):
            lanc.prepare_perturbation(ens.ir_vector(pv), masses_exp=-1)
            psis.append(lanc.psi.copy())
        chains, _ = lanc.run_FT_batched(np.array(psis), 1000, reorth=True, precision="fp64")
        lanc.close()
        for a, b in chains:
            ref = -np.imag(continued_fraction(a, b, b, w, True, smearing=eta))
            out.append((a, b, ref))
    return out, w, eta


def _closures(term, chains, w, eta, K, k_max=None):
    """Per chain: the L-inf spectrum errors (relative to the 1000-step
    spectrum) of the standard and the learned closure after K steps, whether
    the learned tail was used, and (with k_max) the number of steps the
    standard terminator needs to reach the learned closure's error."""
    from tdscha_amd.lanczos import continued_fraction
    rows = []
    for a, b, ref in chains:
        sc = np.abs(ref).max()

        def err_std(k):
            s = -np.imag(continued_fraction(a[:k], b[:k], b[:k], w, True, smearing=eta))
            return np.abs(s - ref).max() / sc

        lrn = -np.imag(term.green_function(a[:K], b[:K], w, eta))
        e_l = np.abs(lrn - ref).max() / sc
        k_eq = None
        if k_max is not None:
            k_eq = next((k for k in range(K, k_max + 1) if err_std(k) <= e_l), k_max + 1)
        rows.append((err_std(K), e_l, term.last_used == "learned", k_eq))
    return [np.array(c) for c in zip(*rows)]


def _data(K):
    from tdscha_amd.terminator import packaged_path
    return packaged_path(K)


@pytest.mark.skipif(not os.path.exists(_data(K)), reason="trained terminator not present")
def test_learned_terminator_gains_over_standard_on_heldout_potentials(heldout_chains):
    """40-step chains closed by the learned tail versus the standard constant
    terminator, scored against 1000-step FP64 chains of the same potential."""
    from tdscha_amd.terminator import LearnedTerminator
    chains, w, eta = heldout_chains
    term = LearnedTerminator.load(_data(K))
    e_std, e_l, used, _ = _closures(term, chains, w, eta, K)
    term.close()
    ratio = np.exp(np.mean(np.log(e_l / e_std)))
    print("standard median %.3e learned median %.3e geo-mean ratio %.3f wins %d/%d used %d" % (
        np.median(e_std), np.median(e_l), ratio, (e_l < e_std).sum(), len(e_l), sum(used)))
    assert sum(used) >= len(used) // 2
    assert ratio < 0.9
    assert np.mean(e_l < e_std) >= 0.6


@pytest.mark.skipif(not os.path.exists(_data(20)), reason="trained terminator not present")
def test_short_chain_closure_saves_a_quarter_of_the_L_applications(heldout_chains):
    """20-step chains closed by the K = 20 ensemble: the standard terminator
    needs a median of at least 27 steps (25% more L applications than 20) to
    reach the same spectrum error on the held-out potentials."""
    from tdscha_amd.terminator import LearnedTerminator
    chains, w, eta = heldout_chains
    term = LearnedTerminator.load(_data(20))
    e_std, e_l, used, k_eq = _closures(term, chains, w, eta, 20, k_max=120)
    term.close()
    saving = 1.0 - 20.0 / np.median(k_eq)
    print("K = 20: standard median %.3e learned median %.3e; standard needs median %d / mean "
          "%.1f steps (saving %.0f%%), used %d/%d" % (np.median(e_std), np.median(e_l),
                                                     np.median(k_eq), np.mean(k_eq),
                                                     100 * saving, sum(used), len(used)))
    assert sum(used) >= len(used) // 2
    assert np.median(k_eq) >= 27


@pytest.mark.skipif(not all(os.path.exists(_data(k)) for k in (20, 30, 40)),
                    reason="trained terminators not present")
def test_terminator_family_dispatches_on_chain_length(heldout_chains):
    """The family closes a chain with the ensemble trained for its length and
    any other length with the reference's standard terminator."""
    from tdscha_amd.lanczos import continued_fraction
    from tdscha_amd.terminator import LearnedTerminator, TerminatorFamily
    chains, w, eta = heldout_chains
    a, b, _ = chains[0]
    fam = TerminatorFamily.packaged()
    assert fam.lengths == [20, 30, 40]
    for k in (20, 30, 40):
        single = LearnedTerminator.load(_data(k))
        g1 = fam.green_function(a[:k], b[:k], w, eta)
        g2 = single.green_function(a[:k], b[:k], w, eta)
        assert fam.last_used == single.last_used
        assert np.array_equal(g1, g2)
        single.close()
    g = fam.green_function(a[:25], b[:25], w, eta)
    assert fam.last_used == "standard"
    gs = continued_fraction(a[:25], b[:25], b[:25], w, True, smearing=eta)
    assert np.array_equal(g, gs)
    fam.close()
