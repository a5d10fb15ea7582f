"""Learned continued-fraction terminator.

The reference closes a Lanczos chain with a constant terminator: the last
coefficient pair (or the mean of the last few) is taken as the asymptotic
(a, b) of an infinite constant tail (Modules/DynamicalLanczos.py:5089-5122).
Here a small network reads the first K coefficient pairs of a chain,
normalised by |b_0|, and predicts a tail: the asymptotic (a_inf, b_inf) and
corrections to the next four a's and b's.  The tail extends the chain's
coefficients and the reference's continued fraction (with its terminator on
the predicted asymptote) gives the Green function, so fewer L applications
reach a given spectrum accuracy.

One ensemble is trained per chain length K (``TerminatorFamily``): the
shorter the chain, the more the learned tail gains over the constant one.

An ensemble of independently trained networks gives a spread; where the
spread of the predicted asymptote exceeds a threshold fixed at training time,
or the chain is not exactly K steps long, the standard constant terminator is
used instead (the chain is then handled exactly as the reference would).

The networks run and train on the gfx950 library's own FP32 kernels
(``tdscha_mlp_*``); the spectral training loss back-propagates through the
continued fraction in FP64 on the device.
"""
from __future__ import annotations

import ctypes as C
import json

import numpy as np

from . import _lib
from .lanczos import continued_fraction

N_TAIL = 4
N_OUT = 2 + 2 * N_TAIL


def _p(a):
    return C.c_void_p(a.ctypes.data)


def features(a, b, K):
    """[a_0..a_{K-1}, b_0..b_{K-1}] / |b_0| (float64; rows for 2-D input)."""
    a = np.atleast_2d(np.asarray(a, np.float64))[:, :K]
    b = np.atleast_2d(np.asarray(b, np.float64))[:, :K]
    s = np.abs(b[:, :1])
    return np.concatenate([a / s, b / s], axis=1)


def extend(a, b, p):
    """Chain coefficients (first K) + tail parameters p (units of |b_0|) ->
    extended (a, b) of length K + N_TAIL + 1 whose last pair is the asymptote."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    s = abs(b[0])
    ta = s * np.r_[p[0] + p[2:2 + N_TAIL], p[0]]
    tb = s * np.r_[p[1] + p[2 + N_TAIL:2 + 2 * N_TAIL], p[1]]
    return np.r_[a, ta], np.r_[b, tb]


def tail_targets(a_long, b_long, K, n_avg=40):
    """Supervised targets from longer chains (units of |b_0|): the mean of the
    coefficients K..K+n_avg-1 as the asymptote, and the next N_TAIL
    coefficients relative to it as corrections."""
    a_long = np.atleast_2d(a_long)
    b_long = np.atleast_2d(b_long)
    s = np.abs(b_long[:, :1])
    an = a_long / s
    bn = b_long / s
    a_inf = an[:, K:K + n_avg].mean(1)
    b_inf = bn[:, K:K + n_avg].mean(1)
    da = an[:, K:K + N_TAIL] - a_inf[:, None]
    db = bn[:, K:K + N_TAIL] - b_inf[:, None]
    return np.concatenate([a_inf[:, None], b_inf[:, None], da, db], axis=1)


def init_params(dims, rng):
    """Uniform(-1/sqrt(fan_in), 1/sqrt(fan_in)) weights and biases; the output
    layer is scaled by 0.1 so training starts near the mean tail."""
    parts = []
    for l in range(len(dims) - 1):
        fan_in = dims[l]
        lim = 1.0 / np.sqrt(fan_in)
        sc = 0.1 if l == len(dims) - 2 else 1.0
        parts.append(rng.uniform(-lim, lim, size=dims[l + 1] * fan_in) * sc)
        parts.append(rng.uniform(-lim, lim, size=dims[l + 1]) * sc)
    return np.concatenate(parts).astype(np.float32)


class Network:
    """One FP32 network held by the library (device parameters + Adam state)."""

    def __init__(self, n_in, width, n_hidden, n_out, params=None, max_batch=256):
        self.dims = [n_in] + [width] * n_hidden + [n_out]
        self.shape = (n_in, width, n_hidden, n_out)
        self.max_batch = max_batch
        h = C.c_void_p()
        init = None
        if params is not None:
            init = np.ascontiguousarray(params, dtype=np.float32)
        _lib.check(_lib.lib().tdscha_mlp_create(C.byref(h), n_in, width, n_hidden, n_out,
                                                max_batch, None if init is None else _p(init)),
                   "tdscha_mlp_create")
        self.h = h
        n = C.c_long()
        _lib.check(_lib.lib().tdscha_mlp_n_params(self.h, C.byref(n)), "tdscha_mlp_n_params")
        self.n_params = n.value
        if init is not None and len(init) != self.n_params:
            self.close()
            raise ValueError("parameter array has %d entries, the network %d" % (
                len(init), self.n_params))

    def params(self):
        out = np.empty(self.n_params, np.float32)
        _lib.check(_lib.lib().tdscha_mlp_get_params(self.h, _p(out)), "tdscha_mlp_get_params")
        return out

    def set_params(self, p, reset_optimizer=True):
        p = np.ascontiguousarray(p, dtype=np.float32)
        if len(p) != self.n_params:
            raise ValueError("expected %d parameters" % self.n_params)
        _lib.check(_lib.lib().tdscha_mlp_set_params(self.h, _p(p), int(bool(reset_optimizer))),
                   "tdscha_mlp_set_params")

    def forward(self, x):
        x = np.ascontiguousarray(np.atleast_2d(x), dtype=np.float32)
        if x.shape[1] != self.dims[0]:
            raise ValueError("input has %d features, the network %d" % (x.shape[1], self.dims[0]))
        y = np.empty((x.shape[0], self.dims[-1]), np.float32)
        _lib.check(_lib.lib().tdscha_mlp_forward(self.h, x.shape[0], _p(x), _p(y)),
                   "tdscha_mlp_forward")
        return y

    def set_mse_weights(self, w):
        w = np.ascontiguousarray(w, dtype=np.float32)
        _lib.check(_lib.lib().tdscha_mlp_set_mse_weights(self.h, _p(w)), "set_mse_weights")

    def set_cf_loss(self, K, w, smearing, out_mean, out_std, loss_scale):
        w = np.ascontiguousarray(w, dtype=np.float64)
        om = np.ascontiguousarray(out_mean, dtype=np.float64)
        os_ = np.ascontiguousarray(out_std, dtype=np.float64)
        _lib.check(_lib.lib().tdscha_mlp_set_cf_loss(self.h, int(K), len(w), _p(w),
                                                     float(smearing), _p(om), _p(os_),
                                                     float(loss_scale)), "set_cf_loss")

    @staticmethod
    def _data(kind, x, target, chains, ref):
        x = np.ascontiguousarray(np.atleast_2d(x), dtype=np.float32)
        t = ch = rf = None
        if kind == 0:
            t = np.ascontiguousarray(target, dtype=np.float32)
        else:
            ch = np.ascontiguousarray(chains, dtype=np.float64)
            rf = np.ascontiguousarray(ref, dtype=np.float64)
        return x, t, ch, rf

    def loss_grad(self, kind, x, target=None, chains=None, ref=None):
        x, t, ch, rf = self._data(kind, x, target, chains, ref)
        loss = C.c_float()
        g = np.empty(self.n_params, np.float32)
        _lib.check(_lib.lib().tdscha_mlp_loss_grad(
            self.h, int(kind), x.shape[0], _p(x), None if t is None else _p(t),
            None if ch is None else _p(ch), None if rf is None else _p(rf), C.byref(loss),
            _p(g)), "tdscha_mlp_loss_grad")
        return loss.value, g

    def adam_step(self, kind, x, target=None, chains=None, ref=None, lr=1e-3, beta1=0.9,
                  beta2=0.999, eps=1e-8):
        x, t, ch, rf = self._data(kind, x, target, chains, ref)
        loss = C.c_float()
        _lib.check(_lib.lib().tdscha_mlp_adam_step(
            self.h, int(kind), x.shape[0], _p(x), None if t is None else _p(t),
            None if ch is None else _p(ch), None if rf is None else _p(rf), lr, beta1, beta2,
            eps, C.byref(loss)), "tdscha_mlp_adam_step")
        return loss.value

    def train(self, kind, x, target=None, chains=None, ref=None, batch=64, epochs=10, lr=1e-3,
              beta1=0.9, beta2=0.999, eps=1e-8, seed=0):
        x, t, ch, rf = self._data(kind, x, target, chains, ref)
        hist = np.zeros(epochs, np.float32)
        _lib.check(_lib.lib().tdscha_mlp_train(
            self.h, int(kind), x.shape[0], _p(x), None if t is None else _p(t),
            None if ch is None else _p(ch), None if rf is None else _p(rf), int(batch),
            int(epochs), lr, beta1, beta2, eps, int(seed), _p(hist)), "tdscha_mlp_train")
        return hist

    def close(self):
        if getattr(self, "h", None):
            _lib.lib().tdscha_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class LearnedTerminator:
    """Ensemble of tail-predicting networks plus the normalisation they need.

    ``green_function(a, b, w, smearing)`` returns the Wigner Green function
    of a chain of exactly K coefficient pairs with the learned tail, or with
    the standard terminator when the ensemble disagrees (or the chain has
    another length); ``last_used`` tells which.
    """

    def __init__(self, members, x_mean, x_std, out_mean, out_std, var_threshold, K=40,
                 width=128, n_hidden=3, info=None):
        self.members = [np.asarray(m, np.float32) for m in members]
        self.x_mean = np.asarray(x_mean, np.float64)
        self.x_std = np.asarray(x_std, np.float64)
        self.out_mean = np.asarray(out_mean, np.float64)
        self.out_std = np.asarray(out_std, np.float64)
        self.var_threshold = float(var_threshold)
        self.K = int(K)
        self.width = int(width)
        self.n_hidden = int(n_hidden)
        self.info = dict(info or {})
        self._nets = None
        self.last_used = None

    # ------------------------------------------------------------- I/O
    def save(self, path):
        np.savez(path, members=np.array(self.members), x_mean=self.x_mean, x_std=self.x_std,
                 out_mean=self.out_mean, out_std=self.out_std,
                 var_threshold=np.array(self.var_threshold), K=np.array(self.K),
                 width=np.array(self.width), n_hidden=np.array(self.n_hidden),
                 info=np.array(json.dumps(self.info)))

    @classmethod
    def load(cls, path):
        d = np.load(path, allow_pickle=False)
        return cls(list(d["members"]), d["x_mean"], d["x_std"], d["out_mean"], d["out_std"],
                   float(d["var_threshold"]), int(d["K"]), int(d["width"]), int(d["n_hidden"]),
                   json.loads(str(d["info"])))

    # ------------------------------------------------------- inference
    def nets(self):
        if self._nets is None:
            self._nets = [Network(2 * self.K, self.width, self.n_hidden, N_OUT, params=m)
                          for m in self.members]
        return self._nets

    def standardise(self, a, b):
        return ((features(a, b, self.K) - self.x_mean) / self.x_std).astype(np.float32)

    def predict(self, a, b):
        """Tail parameters (n, 10) in units of |b_0| (ensemble mean) and the
        ensemble variance of the standardised asymptote (n,)."""
        x = self.standardise(a, b)
        zs = np.array([net.forward(x) for net in self.nets()], np.float64)   # (M, n, 10)
        p = self.out_mean + self.out_std * zs.mean(0)
        var = zs[:, :, :2].var(0).mean(1) if len(zs) > 1 else np.zeros(len(x))
        return p, var

    def coefficients(self, a, b):
        """(a_ext, b_ext, learned): extended coefficients for the continued
        fraction, or the chain itself when the fallback applies."""
        a = np.asarray(a, np.float64)
        b = np.asarray(b, np.float64)
        if len(a) != self.K or len(b) < self.K:
            return a, b, False
        p, var = self.predict(a, b[:self.K])
        if not np.all(np.isfinite(p)) or var[0] > self.var_threshold or p[0, 1] <= 0:
            return a, b, False
        ea, eb = extend(a, b[:self.K], p[0])
        return ea, eb, True

    def green_function(self, a, b, w, smearing, perturbation_modulus=1.0, last_average=1):
        """Wigner Green function (Modules/DynamicalLanczos.py:5032-5162) of
        the chain closed by the learned tail, or by the standard terminator
        (with ``last_average``) when the fallback applies."""
        ea, eb, used = self.coefficients(a, b)
        self.last_used = "learned" if used else "standard"
        return continued_fraction(ea, eb, eb, w, True, use_terminator=True,
                                  last_average=1 if used else last_average, smearing=smearing,
                                  perturbation_modulus=perturbation_modulus)

    def close(self):
        if self._nets is not None:
            for n in self._nets:
                n.close()
            self._nets = None


PACKAGED_LENGTHS = (20, 30, 40)


def packaged_path(K):
    """The trained ensemble for K-step chains of the 16-atom rock-salt models."""
    import os
    return os.path.join(os.path.dirname(os.path.abspath(__file__)), "data",
                        "terminator_rocksalt16_K%d.npz" % K)


class TerminatorFamily:
    """Learned terminators for several chain lengths, one trained ensemble per
    K.  The learned closure gains most on short chains, so the family lets a
    caller stop the recurrence early: a chain of K steps is closed by the K
    ensemble, a chain of a length without one by the standard terminator.
    ``last_used`` tells which closed the last chain."""

    def __init__(self, terms):
        self.terms = {t.K: t for t in terms}
        self.last_used = None

    @classmethod
    def load(cls, paths):
        return cls([LearnedTerminator.load(p) for p in paths])

    @classmethod
    def packaged(cls):
        return cls.load([packaged_path(K) for K in PACKAGED_LENGTHS])

    @property
    def lengths(self):
        return sorted(self.terms)

    def green_function(self, a, b, w, smearing, perturbation_modulus=1.0, last_average=1):
        term = self.terms.get(len(a))
        if term is None:
            self.last_used = "standard"
            return continued_fraction(np.asarray(a, np.float64), np.asarray(b, np.float64),
                                      np.asarray(b, np.float64), w, True, use_terminator=True,
                                      last_average=last_average, smearing=smearing,
                                      perturbation_modulus=perturbation_modulus)
        g = term.green_function(a, b, w, smearing, perturbation_modulus, last_average)
        self.last_used = term.last_used
        return g

    def close(self):
        for t in self.terms.values():
            t.close()


# --------------------------------------------------------------------------
# training
# --------------------------------------------------------------------------
def reference_spectra(a_long, b_long, w, smearing):
    """-Im G of long chains (standard terminator) on the GPU, normalised by
    their maxima; returns (ref (n, n_w), inv_scale (n,))."""
    ref = np.zeros((len(a_long), len(w)))
    inv = np.zeros(len(a_long))
    for i, (a, b) in enumerate(zip(a_long, b_long)):
        s = -np.imag(continued_fraction(a, b, b, w, True, use_terminator=True, smearing=smearing))
        m = np.max(np.abs(s))
        inv[i] = 1.0 / m
        ref[i] = s / m
    return ref, inv


def train_terminator(a_long, b_long, w, smearing, K=40, width=128, n_hidden=3, n_members=4,
                     stage1_epochs=300, stage2_epochs=150, batch=64, lr1=1e-3, lr2=3e-4,
                     loss_scale=1e6, var_quantile=0.99, seed=0, log=None, ref=None,
                     inv_scale=None):
    """Train an ensemble on chains longer than K + 40 steps.

    Stage 1 regresses the tail targets (mean asymptote, next coefficients);
    stage 2 minimises the spectral loss against the long chains' own spectra.
    Returns the LearnedTerminator and a dict of training curves.
    """
    a_long = np.asarray(a_long, np.float64)
    b_long = np.asarray(b_long, np.float64)
    X = features(a_long, b_long, K)
    T = tail_targets(a_long, b_long, K)
    x_mean, x_std = X.mean(0), X.std(0) + 1e-6
    out_mean = T.mean(0)
    out_std = T.std(0) + 1e-6
    Xs = ((X - x_mean) / x_std).astype(np.float32)
    Ts = ((T - out_mean) / out_std).astype(np.float32)
    if ref is None:
        ref, inv_scale = reference_spectra(a_long, b_long, w, smearing)
    s = np.abs(b_long[:, :1])
    chains = np.concatenate([a_long[:, :K], b_long[:, :K], s, inv_scale[:, None]], axis=1)
    members, curves = [], []
    dims = [2 * K] + [width] * n_hidden + [N_OUT]
    for mbr in range(n_members):
        rng = np.random.default_rng(seed + 1000 * mbr)
        net = Network(2 * K, width, n_hidden, N_OUT, params=init_params(dims, rng))
        h1 = net.train(0, Xs, target=Ts, batch=batch, epochs=stage1_epochs, lr=lr1,
                       seed=seed + 17 + mbr)
        net.set_params(net.params(), reset_optimizer=True)
        net.set_cf_loss(K, w, smearing, out_mean, out_std, loss_scale)
        h2 = net.train(1, Xs, chains=chains, ref=ref, batch=batch, epochs=stage2_epochs, lr=lr2,
                       seed=seed + 91 + mbr)
        members.append(net.params())
        curves.append({"stage1": h1.tolist(), "stage2": h2.tolist()})
        if log:
            log("member %d: stage1 %.4g -> %.4g, stage2 %.4g -> %.4g" % (
                mbr, h1[0], h1[-1], h2[0], h2[-1]))
        net.close()
    term = LearnedTerminator(members, x_mean, x_std, out_mean, out_std, np.inf, K, width,
                             n_hidden)
    _, var = term.predict(a_long[:, :K], b_long[:, :K])
    term.var_threshold = float(np.quantile(var, var_quantile))
    term.close()
    return term, curves
