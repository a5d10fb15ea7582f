"""CPU pins of the terminator-network oracle (oracle/mlp_oracle.py)."""
import numpy as np
import pytest

from oracle import lanczos_oracle as LO
from oracle import mlp_oracle as MO

RY_TO_CM = 109737.31568527


def _fd(f, x, h):
    g = np.zeros_like(x)
    for i in range(len(x)):
        xp = x.copy()
        xm = x.copy()
        xp[i] += h
        xm[i] -= h
        g[i] = (f(xp) - f(xm)) / (2 * h)
    return g


def test_forward_closed_form():
    # 2 -> 3 (SiLU) -> 1 with hand-set weights
    W0 = np.array([[1.0, -2.0], [0.5, 0.25], [-1.0, 3.0]])
    b0 = np.array([0.1, -0.2, 0.3])
    W1 = np.array([[2.0, -1.0, 0.5]])
    b1 = np.array([-0.7])
    params = np.concatenate([W0.ravel(), b0, W1.ravel(), b1])
    x = np.array([[0.3, -0.4], [1.5, 2.0]])
    y, _, _ = MO.forward(params, [2, 3, 1], x)
    for s in range(2):
        z = W0 @ x[s] + b0
        h = z / (1 + np.exp(-z))
        assert y[s, 0] == pytest.approx(float(W1[0] @ h + b1[0]), rel=1e-14)
    assert MO.n_params([2, 3, 1]) == len(params)


def test_mse_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    dims = [5, 7, 6, 3]
    p = rng.normal(size=MO.n_params(dims)) * 0.5
    x = rng.normal(size=(4, 5))
    t = rng.normal(size=(4, 3))
    w = np.array([1.0, 0.5, 2.0])
    loss, g = MO.mse_loss_grad(p, dims, x, t, w)
    gf = _fd(lambda q: MO.mse_loss_grad(q, dims, x, t, w)[0], p, 1e-6)
    np.testing.assert_allclose(g, gf, rtol=1e-6, atol=1e-9)
    # a dropped SiLU' factor or a transposed weight would fail the same check
    assert loss > 0


def synthetic_chain(rng, K, b0=5.3e-7):
    """Coefficients shaped like a 16-atom rock-salt Wigner chain (units Ry^2)."""
    a = b0 * (-18.0 + rng.normal(size=K) * 2.0)
    b = b0 * (9.0 + rng.normal(size=K) * 1.0)
    a[0] = -6.0 * b0
    b[0] = b0
    return a, b


def grid(n_w=400, start_cm=0.0):
    return np.linspace(start_cm, 600.0, n_w) / RY_TO_CM, 5.0 / RY_TO_CM


def test_cf_loss_adjoint_value_matches_reference_continued_fraction():
    rng = np.random.default_rng(1)
    K = 12
    a, b = synthetic_chain(rng, K)
    w, eta = grid()
    p = np.r_[-17.0, 8.5, rng.normal(size=8) * 0.5]
    b0 = abs(b[0])
    ea, eb = MO.extend_chain(a, b, p, b0)
    assert len(ea) == K + 5 and len(eb) == K + 5
    assert ea[-1] == pytest.approx(b0 * p[0]) and eb[-1] == pytest.approx(b0 * p[1])
    S = -np.imag(LO.continued_fraction(ea, eb, eb, w, True, use_terminator=True, smearing=eta))
    inv_s = 1.0 / np.abs(S).max()
    ref = S * inv_s * (1 + 0.01 * np.sin(np.arange(len(w))))
    l_ref = MO.cf_loss(p, a, b, b0, inv_s, ref, w, eta)
    l_adj, _ = MO.cf_loss_grad_p(p, a, b, b0, inv_s, ref, w, eta)
    assert l_adj == pytest.approx(l_ref, rel=1e-10)
    # the grid covers both sides of the terminator's band edge
    d = b0 * p[0] + w ** 2
    disc = d ** 2 - 4 * (b0 * p[1]) ** 2
    assert (disc < 0).any() and (disc > 0).any()


def test_cf_loss_gradient_matches_finite_differences():
    rng = np.random.default_rng(2)
    K = 10
    a, b = synthetic_chain(rng, K)
    w, eta = grid(300, 1.0)
    b0 = abs(b[0])
    p = np.r_[-16.0, 8.0, rng.normal(size=8) * 0.3]
    ea, eb = MO.extend_chain(a, b, p + 0.5, b0)
    S = MO.spectrum(ea, eb, w, eta)
    inv_s = 1.0 / np.abs(S).max()
    ref = S * inv_s
    _, g = MO.cf_loss_grad_p(p, a, b, b0, inv_s, ref, w, eta)
    gf = _fd(lambda q: MO.cf_loss(q, a, b, b0, inv_s, ref, w, eta), p, 1e-6)
    np.testing.assert_allclose(g, gf, rtol=2e-5, atol=1e-9 * np.abs(gf).max())


def test_cf_batch_gradient_matches_finite_differences():
    rng = np.random.default_rng(3)
    K = 6
    dims = [2 * K, 8, 8, 10]
    params = rng.normal(size=MO.n_params(dims)) * 0.3
    w, eta = grid(120, 1.0)
    chains, refs, xs = [], [], []
    for s in range(3):
        a, b = synthetic_chain(rng, K)
        b0 = abs(b[0])
        ea, eb = MO.extend_chain(a, b, np.r_[-17.0, 8.7, np.zeros(8)], b0)
        S = MO.spectrum(ea, eb, w, eta)
        inv_s = 1.0 / np.abs(S).max()
        chains.append(np.r_[a, b, b0, inv_s])
        refs.append(S * inv_s * (1 + 0.02 * rng.normal(size=len(w))))
        xs.append(np.r_[a, b] / b0 / 10)
    chains, refs, xs = np.array(chains), np.array(refs), np.array(xs)
    out_mean = np.r_[-17.0, 8.5, np.zeros(8)]
    out_std = np.r_[2.0, 1.0, np.full(8, 0.5)]
    f = lambda q: MO.cf_batch_loss_grad(q, dims, xs, chains, refs, w, eta, out_mean, out_std,
                                        1e3, K)
    loss, g = f(params)
    gf = _fd(lambda q: f(q)[0], params, 1e-6)
    np.testing.assert_allclose(g, gf, rtol=1e-4, atol=1e-7 * np.abs(gf).max())
    assert loss > 0


def test_adam_first_two_steps_closed_form():
    lr, b1, b2, eps = 1e-2, 0.9, 0.999, 1e-8
    p0 = np.array([1.0, -2.0, 0.5])
    g1 = np.array([0.3, -0.1, 2.0])
    g2 = np.array([-0.2, 0.4, 1.0])
    opt = MO.Adam(3, lr, b1, b2, eps)
    p1 = opt.step(p0, g1)
    np.testing.assert_allclose(p1, p0 - lr * g1 / (np.abs(g1) + eps), rtol=1e-14)
    p2 = opt.step(p1, g2)
    m = (0.9 * 0.1 * g1 + 0.1 * g2) / (1 - 0.81)
    v = (0.999 * 0.001 * g1 ** 2 + 0.001 * g2 ** 2) / (1 - 0.999 ** 2)
    np.testing.assert_allclose(p2, p1 - lr * m / (np.sqrt(v) + eps), rtol=1e-13)
