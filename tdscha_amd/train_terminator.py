"""Generate Lanczos chains on the GPU and train the learned terminator.

    python -m tdscha_amd.train_terminator --out-dir DIR [--K 20 30 40]
        [--n-train 5000] [--n-heldout 200] [--train-steps 400] [--ref-steps 1000]
        [--save-chains FILE | --chains FILE]

Training data: rock-salt 2x2x2 spring models (16 atoms, 45 modes) with
k3 ~ U(-4, 0) eV/A^3, k4 ~ U(2, 10) eV/A^4, T ~ U(50, 600) K, one seeded
ensemble of 2000 configurations each, no symmetries; each potential gives
three chains (IR-like x, y, z perturbations, Wigner, full
reorthogonalisation, FP64) run as one batch by the library's device Lanczos.
Ensembles are built by CPU worker processes started before the GPU is
touched.  Held-out potentials come from a separate seed stream and are run
to ``--ref-steps``; their spectra are the reference for the standard and the learned
terminator at K steps.  One ensemble is trained per K on the same chains.
Writes DIR/terminator_rocksalt16_K<K>.npz (the trained ensemble) and
DIR/terminator_rocksalt16_K<K>.json (held-out report) for every K.
``--chains`` trains from chains written earlier by ``--save-chains``
instead of generating them (the GPU still trains).
"""
from __future__ import annotations

import argparse
import json
import multiprocessing as mp
import os
import sys
import time

import numpy as np

RY_TO_CM = 109737.31568527


def log(*a):
    print(*a, flush=True)


def potential_params(n, seed):
    rng = np.random.default_rng(seed)
    k3 = rng.uniform(-4.0, 0.0, n)
    k4 = rng.uniform(2.0, 10.0, n)
    T = rng.uniform(50.0, 600.0, n)
    return k3, k4, T


class _Ens:
    """What the Lanczos reads from an ensemble (picklable, arrays only)."""

    def __init__(self, **kw):
        self.__dict__.update(kw)


def make_ensemble(args):
    k3, k4, T, seed, N = args
    from tdscha_amd.rocksalt import Ensemble, RockSalt
    e = Ensemble(RockSalt(L=2, k3=k3, k4=k4), T, N, seed=seed)
    irs = np.array([e.ir_vector(pv) for pv in ((1, 0, 0), (0, 1, 0), (0, 0, 1))])
    return dict(T=e.T, w=e.w, pols=e.pols, m=e.m, N=e.N, rho=e.rho, X=e.X, Y=e.Y, irs=irs)


def run_chains(pool, params, seed0, N, n_steps, tag):
    """Three full-length chains per potential: arrays a, b [n_pot, 3, n_steps]
    (zero-padded where a chain stopped early) and lengths [n_pot, 3]."""
    from tdscha_amd.lanczos import Lanczos
    k3, k4, T = params
    n = len(k3)
    A = np.zeros((n, 3, n_steps))
    B = np.zeros((n, 3, n_steps))
    lens = np.zeros((n, 3), np.int32)
    jobs = [(float(k3[i]), float(k4[i]), float(T[i]), seed0 + i, N) for i in range(n)]
    t0 = time.time()
    last = t0
    for i, d in enumerate(pool.imap(make_ensemble, jobs, chunksize=2)):
        ens = _Ens(**d)
        lanc = Lanczos(ens)
        lanc.init(use_symmetries=False)
        psis = []
        for v in ens.irs:
            lanc.prepare_perturbation(v, masses_exp=-1)
            psis.append(lanc.psi.copy())
        chains, _ = lanc.run_FT_batched(np.array(psis), n_steps, reorth=True, precision="fp64")
        lanc.close()
        for c, (a, b) in enumerate(chains):
            A[i, c, :len(a)] = a
            B[i, c, :len(b)] = b
            lens[i, c] = min(len(a), len(b))
        if time.time() - last > 20 or i + 1 == n:
            last = time.time()
            log("%s: %d/%d potentials, %.1fs" % (tag, i + 1, n, last - t0))
    return A, B, lens


def spectrum(a, b, w, eta, last_average=1):
    from tdscha_amd.lanczos import continued_fraction
    return -np.imag(continued_fraction(a, b, b, w, True, use_terminator=True,
                                       last_average=last_average, smearing=eta))


def evaluate(term, A, B, lens, w, eta, K, k_max):
    """Held-out report: L-inf spectrum errors (relative to the long chain's maximum)
    of the standard terminator at K, of the learned terminator at K, and the
    number of steps the standard terminator needs to match the learned one."""
    rows = []
    for i in range(A.shape[0]):
        for c in range(3):
            n = lens[i, c]
            a, b = A[i, c, :n], B[i, c, :n]
            if n < k_max:
                continue
            ref = spectrum(a, b, w, eta)
            sc = np.max(np.abs(ref))
            err = lambda s: float(np.max(np.abs(s - ref)) / sc)
            e_std = err(spectrum(a[:K], b[:K], w, eta))
            g = term.green_function(a[:K], b[:K], w, eta)
            used = term.last_used
            e_learn = err(-np.imag(g))
            k_eq = None
            for k in range(K, k_max + 1):
                if err(spectrum(a[:k], b[:k], w, eta)) <= e_learn:
                    k_eq = k
                    break
            rows.append((i, c, e_std, e_learn, used == "learned", k_eq))
    e_std = np.array([r[2] for r in rows])
    e_l = np.array([r[3] for r in rows])
    used = np.array([r[4] for r in rows])
    k_eq = np.array([r[5] if r[5] is not None else k_max + 1 for r in rows])
    ratio = e_l / e_std
    return {
        "chains": len(rows),
        "learned_used": int(used.sum()),
        "fallback": int((~used).sum()),
        "median_err_standard": float(np.median(e_std)),
        "median_err_learned": float(np.median(e_l)),
        "geomean_ratio_learned_over_standard": float(np.exp(np.mean(np.log(ratio)))),
        "fraction_learned_better": float(np.mean(e_l < e_std)),
        "fraction_learned_worse_by_2x": float(np.mean(ratio > 2)),
        "median_steps_standard_needs": float(np.median(k_eq)),
        "mean_steps_standard_needs": float(np.mean(k_eq)),
        "per_chain": [[int(r[0]), int(r[1]), r[2], r[3], bool(r[4]),
                       None if r[5] is None else int(r[5])] for r in rows],
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out-dir", required=True)
    ap.add_argument("--n-train", type=int, default=5000)
    ap.add_argument("--n-heldout", type=int, default=200)
    ap.add_argument("--train-steps", type=int, default=400)
    ap.add_argument("--ref-steps", type=int, default=1000)
    ap.add_argument("--configs", type=int, default=2000)
    ap.add_argument("--K", type=int, nargs="+", default=[20, 30, 40])
    ap.add_argument("--members", type=int, default=4)
    ap.add_argument("--stage1-epochs", type=int, default=300)
    ap.add_argument("--stage2-epochs", type=int, default=150)
    ap.add_argument("--workers", type=int, default=12)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--save-chains", default=None, help="also write the raw chains (npz)")
    ap.add_argument("--chains", default=None,
                    help="train from chains written by --save-chains instead of generating them")
    args = ap.parse_args()
    os.makedirs(args.out_dir, exist_ok=True)

    t0 = time.time()
    if args.chains:
        d = np.load(args.chains, allow_pickle=False)
        A, B, lens, Ah, Bh, lh = (d[k] for k in ("A", "B", "lens", "Ah", "Bh", "lh"))
        args.train_steps, args.ref_steps = A.shape[-1], Ah.shape[-1]
        args.n_heldout = Ah.shape[0]
        log("chains from %s: %d training and %d held-out potentials" % (
            args.chains, A.shape[0], Ah.shape[0]))
    else:
        os.environ["OMP_NUM_THREADS"] = "1"
        os.environ["OPENBLAS_NUM_THREADS"] = "1"
        ctx = mp.get_context("spawn")
        pool = ctx.Pool(args.workers)   # started before anything touches the GPU
        n_pot = (args.n_train + 2) // 3
        ptrain = potential_params(n_pot, args.seed)
        A, B, lens = run_chains(pool, ptrain, 100000, args.configs, args.train_steps, "train")
        pheld = potential_params(args.n_heldout, args.seed + 7919)
        Ah, Bh, lh = run_chains(pool, pheld, 900000, args.configs, args.ref_steps, "held-out")
        pool.close()
        pool.join()
        if args.save_chains:
            np.savez(args.save_chains, A=A, B=B, lens=lens, ptrain=np.array(ptrain), Ah=Ah,
                     Bh=Bh, lh=lh, pheld=np.array(pheld))
    t_gen = time.time() - t0
    from tdscha_amd import terminator as TM

    # training set: full-length chains, the first n_train of them
    a_tr = A.reshape(-1, args.train_steps)
    b_tr = B.reshape(-1, args.train_steps)
    ok = lens.reshape(-1) >= args.train_steps
    a_tr, b_tr = a_tr[ok][:args.n_train], b_tr[ok][:args.n_train]
    log("training chains: %d (generation %.1fs)" % (len(a_tr), t_gen))
    w = np.linspace(0.0, 600.0, 2000) / RY_TO_CM
    eta = 5.0 / RY_TO_CM
    summary = {}
    for K in args.K:
        t1 = time.time()
        term, curves = TM.train_terminator(a_tr, b_tr, w, eta, K=K, n_members=args.members,
                                           stage1_epochs=args.stage1_epochs,
                                           stage2_epochs=args.stage2_epochs, seed=args.seed,
                                           log=log)
        t_train = time.time() - t1
        log("K = %d: trained %d members in %.1fs; variance threshold %.4g" % (
            K, args.members, t_train, term.var_threshold))
        rep = evaluate(term, Ah, Bh, lh, w, eta, K, 120)
        rep.update({
            "K": K, "n_train_chains": int(len(a_tr)), "train_steps": args.train_steps,
            "ref_steps": args.ref_steps, "n_heldout_potentials": args.n_heldout,
            "configs_per_ensemble": args.configs, "members": args.members,
            "chains_from": args.chains or "generated", "generation_s": t_gen,
            "training_s": t_train, "var_threshold": term.var_threshold,
            "saving_median": 1.0 - K / rep["median_steps_standard_needs"],
            "saving_mean": 1.0 - K / rep["mean_steps_standard_needs"],
            "grid": "2000 points 0-600 cm^-1, smearing 5 cm^-1",
            "curves": curves,
        })
        term.info = {k: v for k, v in rep.items() if k not in ("per_chain", "curves")}
        term.save(os.path.join(args.out_dir, "terminator_rocksalt16_K%d.npz" % K))
        term.close()
        with open(os.path.join(args.out_dir, "terminator_rocksalt16_K%d.json" % K), "w") as fp:
            json.dump(rep, fp)
        summary[K] = term.info
        log(json.dumps(term.info, indent=1))
    log(json.dumps({K: {k: v[k] for k in ("median_err_standard", "median_err_learned",
                                          "median_steps_standard_needs", "saving_median",
                                          "saving_mean")} for K, v in summary.items()}))
    return 0


if __name__ == "__main__":
    sys.exit(main())
