#!/usr/bin/env python3
"""Benchmark: L-applications/s of the batched Lanczos on one MI355X.

Workload: rock-salt spring model on a 4x4x4 supercell (128 atoms, 381
non-acoustic modes), N = 20000 configurations at 300 K, the 48 point-group
operations as symmetry images plus the translational projector, three IR-like
perturbations (x, y, z) run as one batch of Lanczos chains in the Wigner
representation with full reorthogonalisation.

One step = one batched Lanczos iteration = 3 L applications.  The headline
value is the split-precision (FP16 hi/lo, FP32 accumulation; recurrence FP64)
throughput; the FP64 path is timed on the same steps and the spectrum of the
split path is compared with it on a 2000-point grid (0-600 cm^-1,
5 cm^-1 broadening).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--precision fp16x2]
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

RY_TO_CM = 109737.31568527
PEAK = {"fp64": 78.6e12, "fp16x2": 2.5e15, "bf16x2": 2.5e15}
PRODUCTS = {"fp64": 1, "fp16x2": 3, "bf16x2": 3}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def build_workload(L, N, T, seed):
    from tdscha_amd.rocksalt import Ensemble, RockSalt
    t = time.time()
    ens = Ensemble(RockSalt(L=L), T, N, seed=seed)
    log("ensemble: %d atoms, %d modes, N=%d (%.1fs)" % (ens.model.nat, len(ens.w), N,
                                                       time.time() - t))
    return ens


def spectrum(a, b, w_grid, eta):
    from tdscha_amd.lanczos import continued_fraction
    gf = continued_fraction(a, b, None, w_grid, True, use_terminator=False, smearing=eta)
    return -np.imag(gf)


def copy_probe(nbytes):
    """One device-to-device copy of a known size through the library, for
    scaling the FETCH_SIZE / WRITE_SIZE counters of a profiled run."""
    from tdscha_amd import _lib
    a = _lib.DeviceBuffer((nbytes,), "uint8")
    b = _lib.DeviceBuffer((nbytes,), "uint8")
    _lib.check(_lib.lib().tdscha_memcpy(b.ptr, a.ptr, nbytes, 3), "copy")   # synchronous
    a.free()
    b.free()


def profiled_traffic(kernel, precision):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 counter
    summary of this workload (profiles/), or None."""
    path = os.path.join(ROOT, "profiles", "%s_%s_128atoms_20k.txt" % (kernel, precision))
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("hbm_bytes_per_launch:"):
            return float(line.split(":", 1)[1])
    return None


def cpu_baseline(ens, lanc, n_sample_cfg, budget_s):
    """Reference C kernels (oracle/_ref, compiled from LanczosFunctions.c) on a
    bounded sample of the same workload, single thread; falls back to the
    project's C restatement when the reference build is absent."""
    from oracle import c_oracle as CO
    rng = np.random.default_rng(7)
    n = len(ens.w)
    R1 = rng.normal(size=n)
    Y1 = rng.normal(size=(n, n)) * 1e-3
    Y1 = Y1 + Y1.T
    kind = "reference" if CO.ref_lib() is not None else "port"
    fn = CO.ref_perturb_averages if kind == "reference" else CO.oracle_perturb_averages
    # probe with one configuration, then size the sample to the time budget
    ns = 1
    while True:
        X = ens.X[:ns]
        Y = ens.Y[:ns]
        rho = ens.rho[:ns]
        t = time.time()
        fn(X, Y, ens.w, rho, R1, Y1, ens.T, 1, lanc.symmetries, lanc.degenerate_space,
           lanc.sym_block_id)
        dt = time.time() - t
        if dt > budget_s / 4 or ns >= n_sample_cfg:
            break
        ns = min(n_sample_cfg, max(ns * 2, int(ns * budget_s / 4 / max(dt, 1e-3))))
    images = ns * lanc.n_syms
    total = ens.N * lanc.n_syms
    t_full = dt * total / images
    single = {"value": 1.0 / t_full, "unit": "L-applications/s", "cores": 1, "kind": kind,
            "sample": "%d configurations x %d symmetry images = %d of %d images, one "
                      "perturbed-average pass (the %s C kernels), time scaled linearly in "
                      "images (measured %.2f s)" % (ns, lanc.n_syms, images, total,
                                                    "reference LanczosFunctions.c sym_fast"
                                                    if kind == "reference" else
                                                    "oracle restatement", dt)}
    # all host cores: the project's C restatement with OpenMP threads over
    # configurations (the reference's C kernels are serial and not safe to
    # call from several threads at once), on P times the sample
    P = max(1, min(16, os.cpu_count() or 1))
    ns_p = max(1, min(ns * P, ens.N))
    while True:
        t = time.time()
        CO.oracle_perturb_averages(ens.X[:ns_p], ens.Y[:ns_p], ens.w, ens.rho[:ns_p], R1, Y1,
                                   ens.T, 1, lanc.symmetries, lanc.degenerate_space,
                                   lanc.sym_block_id, n_threads=P)
        dtp = time.time() - t
        if dtp > budget_s / 8 or ns_p >= ens.N:
            break
        ns_p = min(ens.N, max(2 * ns_p, int(ns_p * budget_s / 4 / max(dtp, 1e-3))))
    imgs_p = ns_p * lanc.n_syms
    parallel = {"value": imgs_p / dtp / total, "unit": "L-applications/s", "cores": P,
                "kind": "port",
                "sample": "%d configurations x %d images on %d OpenMP threads (the oracle's C "
                          "restatement), one perturbed-average pass, scaled linearly in images "
                          "(measured %.2f s)" % (ns_p, lanc.n_syms, P, dtp)}
    return single, parallel


def qspace_run(ens, seed=3, tol=1e-8, cap_range=(50, 400), per_chain=True, tiles=None,
               n_apply=0):
    """Batched uneven q-space chains on the same ensemble: one chain per
    q-point of the 4x4x4 grid, each from a seeded random unit complex
    perturbation of its 6 bands, stopping at |b| < tol or at a per-q cap.
    Timed as one whole-run hipGraph launch (autotuned tiles, or `tiles`) and
    as per-chain plain launches; L-applications/s counts every a coefficient
    computed.  With n_apply > 0 it only runs that many full-batch L
    applications (for profiles)."""
    from tdscha_amd.qlanczos import QSpaceEngine
    from tdscha_amd.qspace import bloch_ensemble, pair_map
    t = time.time()
    modes, Xq, Yq = bloch_ensemble(ens)
    eng = QSpaceEngine(Xq, Yq, modes.w_q, modes.valid_modes_q, ens.rho, ens.T)
    t_setup = time.time() - t
    rng = np.random.default_rng(seed)
    nq, nb = modes.n_q, modes.n_bands
    caps = rng.integers(cap_range[0], cap_range[1] + 1, size=nq)
    iqs = np.arange(nq)
    pms, psis = [], []
    for iq in iqs:
        pm, _ = pair_map(modes.L, int(iq))
        psi = np.zeros(eng.psi_size(pm), np.complex128)
        v = rng.normal(size=nb) + 1j * rng.normal(size=nb)
        psi[:nb] = v / np.linalg.norm(v)
        pms.append(pm)
        psis.append(psi)
    if tiles is None:
        tiles, tune_ms = eng.autotune(iqs, pms, reps=2)
    else:
        eng.set_tiles(*tiles)
        tune_ms = None
    if n_apply:
        # profile mode: full-batch L applications only (kernel times come
        # from the profiler; each call here also builds its batch tables)
        for _ in range(n_apply):
            eng.apply_L(iqs, pms, psis)
        eng.close()
        return {"workload": "q-space full-batch L applications (64 chains)", "n_apply": n_apply,
                "tiles": list(tiles)}
    n_steps = int(caps.max())
    t = time.perf_counter()
    chains, dev_ms = eng.lanczos(iqs, pms, psis, n_steps, caps=caps, tol=tol, use_graph=True)
    wall = time.perf_counter() - t
    n_l = sum(len(a) for a, _, _ in chains)
    out = {"workload": "q-space rock-salt 4x4x4 (64 q-points, 6 bands), %d configurations, "
                       "64 chains (one per q, random unit complex perturbation), caps %d-%d by q, "
                       "tol %g, FP64 complex" % (ens.N, cap_range[0], cap_range[1], tol),
           "l_applications": int(n_l), "batched_graph_value": n_l / wall,
           "batched_graph_wall_s": wall, "batched_graph_device_ms": dev_ms,
           "unit": "L-applications/s", "tiles": list(tiles), "d2v_form": eng.d2v_form(),
           "autotune_best_L_ms": tune_ms,
           "setup_s": t_setup, "chains_stopped_on_tol": int(sum(len(b) < len(a)
                                                               for a, b, _ in chains))}
    # kernel polynomial method on the same batch: 256 Chebyshev moments per
    # chain (128 L applications), one graph launch
    from tdscha_amd.qkpm import kpm_bounds
    from tdscha_amd.qlanczos import wigner_mask
    lmin, lmax = kpm_bounds(modes.w_q, modes.valid_modes_q, True, bound_factor=1.2)
    v0s = []
    for iq, psi in zip(iqs, psis):
        mask = wigner_mask(nb, pair_map(modes.L, int(iq))[1])
        v0s.append(psi / np.sqrt(np.real(np.conj(psi) @ (psi * mask))))
    n_mom = 256
    t = time.perf_counter()
    _, _, _, kpm_ms = eng.kpm(iqs, pms, v0s, v0s, [(0.5 * (lmax - lmin), 0.5 * (lmax + lmin))] * nq,
                              n_mom)
    kpm_wall = time.perf_counter() - t
    out.update({"kpm_chains": nq, "kpm_moments": n_mom,
                "kpm_L_applications_per_s": nq * (n_mom // 2) / kpm_wall,
                "kpm_device_ms": kpm_ms, "kpm_wall_s": kpm_wall})
    if per_chain:
        tiles1, _ = eng.autotune(iqs[:1], pms[:1], reps=2)     # tiles for one chain at a time
        t = time.perf_counter()
        worst, worst50 = 0.0, 0.0
        for i in range(nq):
            (a, b, c), = eng.lanczos([iqs[i]], [pms[i]], [psis[i]], int(caps[i]), tol=tol,
                                     use_graph=False)[0]
            k = min(len(a), len(chains[i][0]))
            d = np.abs(a[:k] - chains[i][0][:k]) / np.max(np.abs(chains[i][0][:k]))
            worst = max(worst, float(np.max(d)))
            worst50 = max(worst50, float(np.max(d[:50])))
        wall1 = time.perf_counter() - t
        # The per-chain run sums in the order of its own tiles: the first 50
        # coefficients agree to rounding.  On this grid the chains start from
        # complex Bloch vectors, where the q-space operator (QL:398-400,
        # JQ:509-517) is neither complex-linear nor Hermitian in its metric, so
        # the Hermitian recurrence is not a projection of it and nothing damps
        # rounding differences: late coefficients of a few chains depend on the
        # summation order at the 1e-2 level past step ~200.  With equal tiles
        # the runs are bit-identical (tests/test_gpu_qspace.py).
        out.update({"per_chain_value": n_l / wall1, "per_chain_wall_s": wall1,
                    "per_chain_tiles": list(tiles1), "per_chain_d2v_form": eng.d2v_form(),
                    "per_chain_vs_batched_max_rel_diff_a_first50": worst50,
                    "per_chain_vs_batched_max_rel_diff_a": worst,
                    "batched_speedup": wall1 / wall})
    eng.close()
    return out


def rotation_run(ens, psis, lanc_img, results, args):
    """The same chains with the symmetries applied by rotating operators
    (configurations only on the device): L-applications/s of the FP64 and
    headline split paths, device bytes of the two ensembles, and how far the
    rotated L and the first coefficients sit from the image form's."""
    from tdscha_amd.lanczos import Lanczos
    lanc = Lanczos(ens)
    lanc.gamma_only = True
    lanc.symmetry_form = "operators"
    t = time.time()
    lanc.init(use_symmetries=True)
    log("rotation form: init %.1fs, rows stored %d (images form %d)" % (
        time.time() - t, lanc.engine.n_images_padded, lanc_img.engine.n_images_padded))
    nc = len(psis)
    out = {"stored_rows": lanc.engine.n_images_padded,
           "images_form_rows": lanc_img.engine.n_images_padded,
           "ensemble_bytes": 16 * lanc.engine.n_images_padded * lanc.engine.n_modes_padded,
           "images_form_bytes": 16 * lanc_img.engine.n_images_padded
           * lanc_img.engine.n_modes_padded}
    # a pass carries up to 8 (chain, symmetry) pairs
    for prec in dict.fromkeys(["fp64", args.precision]):
        lanc.engine.autotune(8, prec, reps=2)
        lanc.run_FT_batched(psis, 1, reorth=True, precision=prec)
        lanc.engine.synchronize()
        t0 = time.perf_counter()
        chains, _ = lanc.run_FT_batched(psis, args.rotation_steps, reorth=True, precision=prec)
        lanc.engine.synchronize()
        wall = time.perf_counter() - t0
        n_l = sum(len(a) for a, _ in chains)
        out[prec + "_value"] = n_l / wall
        if prec in results:
            img = results[prec]
            out[prec + "_vs_images"] = (n_l / wall) / (img["n_l"] / img["wall"])
            da = 0.0
            for (a1, _), (a0, _) in zip(chains, img["chains"]):
                k = min(len(a1), len(a0))
                da = max(da, float(np.max(np.abs(a1[:k] - a0[:k]) / np.abs(a0[:k]))))
            out[prec + "_coeff_a_max_rel_diff_vs_images"] = da
        log("rotation form %s: %d steps x %d chains %.3fs -> %.1f L/s" % (
            prec, args.rotation_steps, nc, wall, n_l / wall))
    lanc.precision = "fp64"
    lanc_img.precision = "fp64"
    Lr, Li = lanc.apply_L_batch(psis), lanc_img.apply_L_batch(psis)
    out["fp64_Lv_rel_diff_vs_images"] = float(np.max(np.linalg.norm(Lr - Li, axis=1)
                                                     / np.linalg.norm(Li, axis=1)))
    lanc.close()
    return out


def terminator_run(n_pot=8, ref_steps=1000, k_max=120):
    """The packaged learned terminators (one ensemble per chain length K) on
    fresh held-out potentials of their training family (16-atom rock salt,
    k3 ~ U(-4, 0), k4 ~ U(2, 10), T ~ U(50, 600) K, seeds outside the training
    streams): per K, the median L-inf spectrum error of the standard and the
    learned closure after K steps against the 1000-step FP64 chain, and the
    steps the standard terminator needs to reach the learned one's error."""
    from tdscha_amd.lanczos import Lanczos, continued_fraction
    from tdscha_amd.rocksalt import Ensemble, RockSalt
    from tdscha_amd.terminator import TerminatorFamily
    w = np.linspace(0.0, 600.0, 2000) / RY_TO_CM
    eta = 5.0 / RY_TO_CM
    rng = np.random.default_rng(4242)
    chains = []
    t0 = time.perf_counter()
    for i in range(n_pot):
        k3, k4, T = rng.uniform(-4, 0), rng.uniform(2, 10), rng.uniform(50, 600)
        ens = Ensemble(RockSalt(L=2, k3=k3, k4=k4), T, 2000, seed=8_000_000 + i)
        lanc = Lanczos(ens)
        lanc.init(use_symmetries=False)
        psis = []
        for pv in ((1, 0, 0), (0, 1, 0), (0, 0, 1)):
            lanc.prepare_perturbation(ens.ir_vector(pv), masses_exp=-1)
            psis.append(lanc.psi.copy())
        out, _ = lanc.run_FT_batched(np.array(psis), ref_steps, reorth=True, precision="fp64")
        lanc.close()
        chains += [(a, b) for a, b in out if len(a) >= k_max]
    fam = TerminatorFamily.packaged()
    res = {"potentials": n_pot, "chains": len(chains), "ref_steps": ref_steps,
           "grid": "2000 points 0-600 cm^-1, smearing 5 cm^-1"}
    for K in fam.lengths:
        e_std, e_l, k_eq, used = [], [], [], 0
        for a, b in chains:
            ref = -np.imag(continued_fraction(a, b, b, w, True, smearing=eta))
            sc = np.abs(ref).max()

            def err_std(k):
                g = continued_fraction(a[:k], b[:k], b[:k], w, True, smearing=eta)
                return float(np.abs(-np.imag(g) - ref).max() / sc)

            el = float(np.abs(-np.imag(fam.green_function(a[:K], b[:K], w, eta)) - ref).max() / sc)
            used += fam.last_used == "learned"
            e_std.append(err_std(K))
            e_l.append(el)
            k_eq.append(next((k for k in range(K, k_max + 1) if err_std(k) <= el), k_max + 1))
        med = float(np.median(k_eq))
        res["K%d" % K] = {"median_err_standard": float(np.median(e_std)),
                          "median_err_learned": float(np.median(e_l)),
                          "median_steps_standard_needs": med,
                          "fewer_L_applications_median": 1.0 - K / med,
                          "learned_used": int(used)}
    fam.close()
    res["wall_s"] = time.perf_counter() - t0
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--precision", default="fp16x2")
    ap.add_argument("--supercell", type=int, default=4)
    ap.add_argument("--configs", type=int, default=20000)
    ap.add_argument("--gate-steps", type=int, default=200,
                    help="chain length of the split-vs-FP64 accuracy figures (untimed runs "
                         "when --steps is shorter)")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=20.0)
    ap.add_argument("--no-qspace", action="store_true")
    ap.add_argument("--qspace-only", action="store_true",
                    help="run only the batched q-space chains and print their JSON")
    ap.add_argument("--qspace-tiles", default=None,
                    help="q-space tiles W,C,R instead of the autotuner's")
    ap.add_argument("--qspace-apply", type=int, default=0,
                    help="run only this many full-batch q-space L applications (profiles)")
    ap.add_argument("--no-per-chain", action="store_true",
                    help="skip the per-chain q-space comparison run")
    ap.add_argument("--copy-probe", type=int, default=0,
                    help="bytes of one device-to-device copy to run first (counter scaling)")
    ap.add_argument("--no-rotation", action="store_true",
                    help="skip the operator-rotation form's measurement")
    ap.add_argument("--rotation-steps", type=int, default=5)
    ap.add_argument("--no-terminator", action="store_true",
                    help="skip the learned terminator's held-out evaluation")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the library has no multi-GPU path (the ensemble sum "
              "would shard by configuration with one all-reduce of n + n^2 doubles per L "
              "application).  Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    from tdscha_amd import _lib
    from tdscha_amd.lanczos import Lanczos
    if args.copy_probe:
        copy_probe(args.copy_probe)

    ens = build_workload(args.supercell, args.configs, 300.0, seed=0)
    if args.qspace_only:
        qt = tuple(int(x) for x in args.qspace_tiles.split(",")) if args.qspace_tiles else None
        print(json.dumps(qspace_run(ens, per_chain=not args.no_per_chain, tiles=qt,
                                    n_apply=args.qspace_apply)), flush=True)
        return
    lanc = Lanczos(ens)
    lanc.gamma_only = True
    t = time.time()
    lanc.init(use_symmetries=True)
    log("init (symmetries + upload + image expansion): %.1fs, n_syms=%d, images=%d" % (
        time.time() - t, lanc.n_syms, lanc.engine.n_images))
    psis = []
    for pv in ((1, 0, 0), (0, 1, 0), (0, 0, 1)):
        lanc.prepare_perturbation(ens.ir_vector(pv), masses_exp=-1)
        psis.append(lanc.psi.copy())
    psis = np.array(psis)
    nc = len(psis)
    K, W = args.steps, args.warmup

    def timed_run(prec, steps):
        lanc.engine.synchronize()
        t0 = time.perf_counter()
        chains, dev_ms = lanc.run_FT_batched(psis, steps, reorth=True, precision=prec)
        lanc.engine.synchronize()
        return chains, time.perf_counter() - t0, dev_ms

    # K splits of the outer-product GEMMs chosen by the autotuner for this
    # batch (untimed); both split precisions share the split path's choice
    tuned = {}
    for prec in dict.fromkeys([args.precision] + ([] if args.no_fp64 else ["fp64"])):
        ks, ms = lanc.engine.autotune(len(psis), prec, reps=2)
        tuned[prec] = {"ksplit": ks, "ms_per_pass": ms}
        log("autotune %s: K split %d, %.3f ms per batched pass" % (prec, ks, ms))

    results = {}
    precs = [args.precision] + ([] if (args.no_fp64 or args.precision == "fp64") else ["fp64"])
    if args.precision == "fp16x2" and not args.no_fp64:
        precs.append("bf16x2")               # the other split variant, reported beside
    for prec in precs:
        if W > 0:
            timed_run(prec, W)
        chains, wall, dev_ms = timed_run(prec, K)
        n_l = sum(len(a) for a, _ in chains)
        results[prec] = dict(chains=chains, wall=wall, dev_ms=dev_ms, n_l=n_l)
        log("%s: %d steps x %d chains: wall %.3fs, device loop %.3fs -> %.1f L/s" % (
            prec, K, nc, wall, dev_ms / 1e3, n_l / wall))

    head = results[args.precision]
    value = head["n_l"] / head["wall"]
    ms_per_step = head["wall"] / K * 1e3

    # spectrum error of the headline path against the FP64 path
    def spec_error(chains, ref_chains=None):
        w_grid = np.linspace(0.0, 600.0, 2000) / RY_TO_CM
        eta = 5.0 / RY_TO_CM
        errs = []
        ref_chains = results["fp64"]["chains"] if ref_chains is None else ref_chains
        for (a1, b1), (a0, b0) in zip(chains, ref_chains):
            s1 = spectrum(a1, b1, w_grid, eta)
            s0 = spectrum(a0, b0, w_grid, eta)
            errs.append(float(np.max(np.abs(s1 - s0)) / np.max(np.abs(s0))))
        return max(errs), errs

    # Accuracy of the split paths against FP64 on the configured chains: 200
    # steps per chain whatever --steps is (untimed runs when the timed ones
    # were shorter), the spectrum on the 2000-point grid, the first 50
    # coefficients and one L application of each starting vector.
    spec_err = bf16 = None
    acc = {}
    gate_steps = max(K, args.gate_steps)
    if "fp64" in results and args.precision != "fp64":
        short = {p: spec_error(results[p]["chains"])[0] for p in results if p != "fp64"}
        long_chains = {}
        for prec in results:
            if K >= gate_steps:
                long_chains[prec] = results[prec]["chains"]
            else:
                t0 = time.perf_counter()
                long_chains[prec] = lanc.run_FT_batched(psis, gate_steps, reorth=True,
                                                        precision=prec)[0]
                log("%s: %d-step chains for the accuracy gate (untimed, %.1fs)" % (
                    prec, gate_steps, time.perf_counter() - t0))
        # the FP64 chains again with another K split of the outer GEMM (same
        # arithmetic, another summation order): how far FP64 itself carries
        # the coefficients of this workload
        ks0 = lanc.engine.ksplit
        lanc.engine.set_ksplit(max(8, (ks0 // 2) // 8 * 8))
        fp64_alt = lanc.run_FT_batched(psis, gate_steps, reorth=True, precision="fp64")[0]
        lanc.engine.set_ksplit(ks0)

        def parting_step(ch1, ch0, thr=1e-6):
            """First step whose a or b differs by more than thr (relative)."""
            first = gate_steps
            for (a1, b1), (a0, b0) in zip(ch1, ch0):
                k = min(len(a1), len(a0), len(b1), len(b0))
                d = np.maximum(np.abs(a1[:k] - a0[:k]) / np.abs(a0[:k]),
                               np.abs(b1[:k] - b0[:k]) / np.abs(b0[:k]))
                bad = np.flatnonzero(d > thr)
                first = min(first, int(bad[0]) if len(bad) else k)
            return first

        fp64_parting = parting_step(fp64_alt, long_chains["fp64"])
        fp64_spec = spec_error(fp64_alt, long_chains["fp64"])[0]
        log("fp64 vs fp64 with K split %d instead of %d: coefficients part (1e-6) at step %d, "
            "spectrum %.2e" % (max(8, (ks0 // 2) // 8 * 8), ks0, fp64_parting, fp64_spec))
        lanc.precision = "fp64"
        L64 = lanc.apply_L_batch(psis)
        for prec in results:
            if prec == "fp64":
                continue
            se, errs = spec_error(long_chains[prec], long_chains["fp64"])
            ea = eb = 0.0
            for (a1, b1), (a0, b0) in zip(long_chains[prec], long_chains["fp64"]):
                k = min(50, len(a1), len(a0))
                ea = max(ea, float(np.max(np.abs(a1[:k] - a0[:k]) / np.abs(a0[:k]))))
                kb = min(50, len(b1), len(b0))
                eb = max(eb, float(np.max(np.abs(b1[:kb] - b0[:kb]) / np.abs(b0[:kb]))))
            lanc.precision = prec
            Ls = lanc.apply_L_batch(psis)
            lanc.precision = "fp64"
            lv = float(np.max(np.linalg.norm(Ls - L64, axis=1) / np.linalg.norm(L64, axis=1)))
            acc[prec] = {"spectrum_linf_rel_err": se, "chain_steps": gate_steps,
                         "spectrum_linf_rel_err_per_chain": errs,
                         "spectrum_linf_rel_err_timed_chains": short[prec],
                         "timed_chain_steps": K,
                         "coeff_a_max_rel_err_first50": ea, "coeff_b_max_rel_err_first50": eb,
                         "coeff_parting_step_1e-6": parting_step(long_chains[prec],
                                                                 long_chains["fp64"]),
                         "fp64_vs_fp64_other_ksplit_parting_step_1e-6": fp64_parting,
                         "fp64_vs_fp64_other_ksplit_spectrum_linf_rel_err": fp64_spec,
                         "Lv_rel_l2_err": lv}
            log("%s vs FP64 on %d-step chains: spectrum %.2e (per chain %s), a %.2e, b %.2e "
                "(first 50), coefficients part at step %d, L.v %.2e; %d-step timed chains: "
                "spectrum %.2e" % (prec, gate_steps, se, ["%.2e" % x for x in errs], ea, eb,
                                   acc[prec]["coeff_parting_step_1e-6"], lv, K, short[prec]))
        spec_err = acc[args.precision]["spectrum_linf_rel_err"]
    if "bf16x2" in results and args.precision != "bf16x2":
        r = results["bf16x2"]
        bf16 = {"value": r["n_l"] / r["wall"], "unit": "L-applications/s",
                "speedup_vs_fp64": (r["n_l"] / r["wall"]) / (results["fp64"]["n_l"] /
                                                            results["fp64"]["wall"])}
        bf16.update(acc.get("bf16x2", {}))

    # per-kernel durations with HIP events on the library stream
    L = _lib.lib()
    prof = {}
    for prec in results:
        ms = np.zeros(4)
        rc = L.tdscha_profile_perturb_averages(lanc.engine.ctx, nc,
                                               {"fp64": 0, "fp16x2": 1, "bf16x2": 2}[prec], 5,
                                               C.c_void_p(ms.ctypes.data))
        _lib.check(rc, "profile")
        prof[prec] = ms
        log("%s kernels (ms per batched pass): weights %.3f outer %.3f reduce %.3f total %.3f"
            % (prec, *ms))
    n = lanc.n_modes
    M = lanc.engine.n_images
    pm = prof[args.precision]
    dom = int(np.argmax(pm[:2]))
    kname = ["weights_gemm", "outer_gemm"][dom]
    alg = 2.0 * M * n * n * nc                 # one contraction of 2 M n^2 per chain
    achieved = alg * PRODUCTS[args.precision] / (pm[dom] * 1e-3)
    roof = {"bound": "mfma", "kernel": kname, "achieved": achieved / 1e12,
            "peak": PEAK[args.precision] / 1e12, "unit": "TFLOP/s",
            "frac": achieved / PEAK[args.precision],
            "traffic": profiled_traffic(kname, args.precision)}

    rot = None
    if not args.no_rotation:
        rot = rotation_run(ens, psis, lanc, results, args)

    qres = None
    if not args.no_qspace:
        try:
            qres = qspace_run(ens)
            log("q-space: %s" % {k: v for k, v in qres.items() if k != "workload"})
        except Exception as e:          # reported, never hides the headline line
            qres = {"error": str(e)}
    term = None
    if not args.no_terminator:
        try:
            term = terminator_run()
            log("learned terminator: %s" % term)
        except Exception as e:          # reported, never hides the headline line
            term = {"error": str(e)}
    cpu = cpu_all = None
    if not args.no_cpu:
        cpu, cpu_all = cpu_baseline(ens, lanc, 64, args.cpu_budget)
        log("cpu baseline: %s" % cpu)
        log("cpu baseline, all cores: %s" % cpu_all)

    dtype = {"fp64": "f64", "fp16x2": "f16x2-split/f32-acc + f64 recurrence",
             "bf16x2": "bf16x2-split/f32-acc + f64 recurrence"}[args.precision]
    out = {
        "metric": "L-applications/s at 128 atoms x 20k configs x 48 syms; split-prec vs FP64, "
                  "spec err",
        "value": value,
        "unit": "L-applications/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": dtype,
        "data": "synthetic (seeded rock-salt spring-model ensemble)",
        "config": {"workload": "rock-salt spring model %dx%dx%d supercell (%d atoms, %d modes), "
                               "%d configurations, %d point-group images + translational "
                               "projector, 3 IR chains, Wigner, full reorthogonalisation" % (
                                   args.supercell, args.supercell, args.supercell,
                                   ens.model.nat, n, args.configs, lanc.n_syms),
                   "chains": nc, "images": M, "precision": args.precision},
        "fp64_value": (results["fp64"]["n_l"] / results["fp64"]["wall"]) if "fp64" in results
        else None,
        "speedup_vs_fp64": (value / (results["fp64"]["n_l"] / results["fp64"]["wall"]))
        if ("fp64" in results and args.precision != "fp64") else None,
        "spectrum_linf_rel_err": spec_err,
        "accuracy": acc.get(args.precision),
        "bf16x2": bf16,
        "kernel_ms": {p: [float(x) for x in prof[p]] for p in prof},
        "autotuned_ksplit": tuned,
        "roofline": roof,
        "cpu_baseline": cpu,
        "cpu_baseline_all_cores": cpu_all,
        "operator_rotation": rot,
        "qspace": qres,
        "learned_terminator": term,
    }
    print(json.dumps(out), flush=True)
    lanc.close()


if __name__ == "__main__":
    main()
