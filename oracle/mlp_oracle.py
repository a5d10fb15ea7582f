"""CPU oracle (test infrastructure only) for the learned terminator's network.

FP64 numpy restatement of what the library's FP32 network kernels compute:
forward pass, weighted mean-squared-error loss, the spectral continued-fraction
loss and their parameter gradients, and bias-corrected Adam.  Only the tests
use this file; the library never imports it.

The spectral loss wraps the continued fraction of the reference,
``get_green_function_continued_fraction`` (Modules/DynamicalLanczos.py:
5032-5162, restated in lanczos_oracle.continued_fraction), applied to the
chain's K coefficients extended by the predicted tail; the network itself has
no counterpart in the reference.

Pins (tests/test_mlp_oracle.py):
  * forward: a one-hidden-layer network with hand-set weights against its
    closed form (SiLU(z) = z sigmoid(z));
  * network and MSE gradients: central finite differences in FP64;
  * spectral loss value: lanczos_oracle.continued_fraction on the extended
    coefficient arrays (use_terminator=True, last_average=1), which is itself
    pinned to the reference's closed-form continued-fraction tests;
  * spectral-loss gradient (analytic adjoint below): central finite
    differences of that value;
  * Adam: the closed form of the first two steps.
"""
from __future__ import annotations

import numpy as np

from oracle import lanczos_oracle as LO

N_TAIL = 4
N_CF_OUT = 2 + 2 * N_TAIL


def layer_dims(n_in, width, n_hidden, n_out):
    return [n_in] + [width] * n_hidden + [n_out]


def unpack(params, dims):
    """[(W (out, in), b (out)), ...] views of the flat parameter array."""
    out, off = [], 0
    for l in range(len(dims) - 1):
        i, o = dims[l], dims[l + 1]
        W = params[off:off + o * i].reshape(o, i)
        off += o * i
        b = params[off:off + o]
        off += o
        out.append((W, b))
    assert off == len(params)
    return out


def n_params(dims):
    return sum(dims[l + 1] * dims[l] + dims[l + 1] for l in range(len(dims) - 1))


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def forward(params, dims, x):
    """Returns (y, acts, pres): acts[l] is the input of layer l."""
    layers = unpack(np.asarray(params, np.float64), dims)
    a = np.asarray(x, np.float64)
    acts, pres = [a], []
    for l, (W, b) in enumerate(layers):
        z = a @ W.T + b
        pres.append(z)
        a = silu(z) if l + 1 < len(layers) else z
        acts.append(a)
    return a, acts, pres


def backward(params, dims, acts, pres, dy):
    """Parameter gradient (flat) from dL/dy of the last layer."""
    layers = unpack(np.asarray(params, np.float64), dims)
    grads = [None] * len(layers)
    dz = dy
    for l in range(len(layers) - 1, -1, -1):
        W, _ = layers[l]
        grads[l] = (dz.T @ acts[l], dz.sum(0))
        if l > 0:
            dz = (dz @ W) * dsilu(pres[l - 1])
    return np.concatenate([np.concatenate([g[0].ravel(), g[1]]) for g in grads])


def mse_loss_grad(params, dims, x, t, w=None):
    y, acts, pres = forward(params, dims, x)
    B = len(x)
    w = np.ones(y.shape[1]) if w is None else np.asarray(w, np.float64)
    d = y - t
    loss = float(np.sum(w * d * d) / B)
    return loss, backward(params, dims, acts, pres, 2.0 * w * d / B)


# --------------------------------------------------------------------------
# spectral continued-fraction loss
# --------------------------------------------------------------------------
def extend_chain(a, b, p, b0):
    """a, b (K each) + tail parameters p (units of b0) -> extended (K + 5)."""
    a_inf, b_inf = p[0], p[1]
    ta = [b0 * (a_inf + p[2 + j]) for j in range(N_TAIL)] + [b0 * a_inf]
    tb = [b0 * (b_inf + p[2 + N_TAIL + j]) for j in range(N_TAIL)] + [b0 * b_inf]
    return np.r_[a, ta], np.r_[b, tb]


def spectrum(a, b, w, eta):
    """-Im G of the Wigner continued fraction with the standard terminator."""
    gf = LO.continued_fraction(a, b, b, w, True, use_terminator=True, last_average=1,
                               smearing=eta)
    return -np.imag(gf)


def cf_loss(p, a, b, b0, inv_s, ref, w, eta):
    """Mean squared deviation of the extended chain's spectrum from ref."""
    ea, eb = extend_chain(a, b, p, b0)
    r = spectrum(ea, eb, w, eta) * inv_s - ref
    return float(np.mean(r * r))


def cf_loss_grad_p(p, a, b, b0, inv_s, ref, w, eta):
    """Analytic gradient of cf_loss with respect to the ten tail parameters.

    The terminator root g solves B^2 g^2 - d g + 1 = 0 (d = A + w^2, no
    smearing), so dg = (g dd - 2 B g^2 dB) / (2 B^2 g - d); each level
    g_i = 1 / (e_i - b_i^2 g_{i+1}) gives dg_i = -g_i^2 (de_i - 2 b_i g_{i+1} db_i
    - b_i^2 dg_{i+1}); above the tail dG/dg_K = prod_i b_i^2 g_i^2.
    """
    w = np.asarray(w, np.float64)
    w2, sm = w * w, 2.0 * w * eta
    A, Bt = b0 * p[0], b0 * p[1]
    d = A + w2
    disc = d * d - 4.0 * Bt * Bt
    sq = np.sqrt(disc + 0j)
    den = 2.0 * Bt * Bt
    gp, gm = (d + sq) / den, (d - sq) / den
    use_p = np.abs(gp) < np.abs(gm)
    g = np.where(use_p, gp, gm)
    delta = np.where(use_p, sq, -sq)
    dg = np.zeros((N_CF_OUT, len(w)), np.complex128)
    ok = delta != 0
    dg[0, ok] = b0 * g[ok] / delta[ok]
    dg[1, ok] = -2.0 * Bt * b0 * g[ok] ** 2 / delta[ok]
    for j in range(N_TAIL - 1, -1, -1):
        aj = b0 * (p[0] + p[2 + j])
        bj = b0 * (p[1] + p[2 + N_TAIL + j])
        gn = 1.0 / (aj + w2 + 1j * sm - bj * bj * g)
        de = np.zeros(N_CF_OUT)
        de[0] = de[2 + j] = b0
        dbv = np.zeros(N_CF_OUT)
        dbv[1] = dbv[2 + N_TAIL + j] = b0
        dg = -(gn ** 2)[None] * (de[:, None] - 2.0 * bj * g[None] * dbv[:, None] - bj * bj * dg)
        g = gn
    D = np.ones(len(w), np.complex128)
    for i in range(len(a) - 1, -1, -1):
        gn = 1.0 / (a[i] + w2 + 1j * sm - b[i] * b[i] * g)
        D = D * b[i] * b[i] * gn * gn
        g = gn
    S = -np.imag(g)
    r = S * inv_s - ref
    dS = -np.imag(D[None] * dg)
    return float(np.mean(r * r)), (2.0 * inv_s * dS @ r) / len(w)


def cf_batch_loss_grad(params, dims, x, chains, ref, w, eta, out_mean, out_std, loss_scale, K):
    """Loss kind 1 of the library: loss_scale * mean over the batch."""
    y, acts, pres = forward(params, dims, x)
    B = len(x)
    dy = np.zeros_like(y)
    total = 0.0
    for s in range(B):
        ch = chains[s]
        a, b, b0, inv_s = ch[:K], ch[K:2 * K], ch[2 * K], ch[2 * K + 1]
        p = out_mean + out_std * y[s]
        l, gp = cf_loss_grad_p(p, a, b, b0, inv_s, ref[s], w, eta)
        total += l
        dy[s] = loss_scale / B * gp * out_std
    return loss_scale * total / B, backward(params, dims, acts, pres, dy)


# --------------------------------------------------------------------------
# Adam
# --------------------------------------------------------------------------
class Adam:
    def __init__(self, n, lr, beta1=0.9, beta2=0.999, eps=1e-8):
        self.m = np.zeros(n)
        self.v = np.zeros(n)
        self.t = 0
        self.lr, self.b1, self.b2, self.eps = lr, beta1, beta2, eps

    def step(self, p, g):
        self.t += 1
        self.m = self.b1 * self.m + (1 - self.b1) * g
        self.v = self.b2 * self.v + (1 - self.b2) * g * g
        mh = self.m / (1 - self.b1 ** self.t)
        vh = self.v / (1 - self.b2 ** self.t)
        return p - self.lr * mh / (np.sqrt(vh) + self.eps)
