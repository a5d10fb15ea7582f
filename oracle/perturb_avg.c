/*
 * CPU oracle: C restatement of the ensemble-averaged D3/D4 application.
 *
 * TEST INFRASTRUCTURE ONLY (the test oracle and the CPU baseline port); the
 * product library never links it.
 *
 * Follows the reference per-image loop: for every configuration i and
 * symmetry j the image (u, f) = (S_j x_i, S_j y_i) is formed from the
 * block-diagonal symmetry (CModules/LanczosFunctions.c:2158-2170), then
 *   f_av  += rho_i [ -(u.Y1.u)/6 f - (u.Y1.g + g.Y1.u)/6 a ]      (LF:2173-2226)
 *   d2v   -= rho_i [ (a.R1)/3 (a f^T + f a^T) + (f.R1)/3 a a^T ]  (LF:2366-2407)
 *   d2v   += d4 rho_i [ (u.Y1.u)/8 (a f^T + f a^T)
 *                      + (g.Y1.u + u.Y1.g)/8 a a^T ]             (LF:2683-2725)
 * with a = f_Y u, g = f_psi f (LF:14-31), normalised by N_eff * N_sym.
 * The three reference passes are fused into one and f_Y, f_psi are
 * evaluated once per mode instead of inside the inner loops.
 *
 * Images are addressed by the reference's big index b = i * N_sym + j
 * (Modules/tdscha_core.jl:82-83); [b_begin, b_end) selects a partition as
 * Parallel.GoParallel does (Modules/DynamicalLanczos.py:3141-3164), with
 * the global normalisation kept, so partial results sum to the total.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>

#define RY_TO_K 157887.32400374097

static double bose(double w, double T) {
    return T > 0 ? 1.0 / (exp(w * RY_TO_K / T) - 1.0) : 0.0;
}

int oracle_perturb_averages(int n, int N, const double *X, const double *Y,
                            const double *w, const double *rho,
                            const double *R1, const double *Y1, double T,
                            int apply_d4, int n_syms, int n_blocks,
                            const int *block_dim, const int *block_modes,
                            const double *block_syms, const int *sym_block_id,
                            long b_begin, long b_end, int n_threads,
                            double *f_out, double *d2v_out) {
    if (n <= 0 || N <= 0 || n_syms <= 0 || n_blocks <= 0) return 1;
    if (apply_d4 != 0 && apply_d4 != 1) return 2;
    long n_img = (long)N * n_syms;
    if (b_begin < 0) b_begin = 0;
    if (b_end < 0 || b_end > n_img) b_end = n_img;

    /* per-block offsets into the concatenated arrays */
    int *mode_off = (int *)malloc(sizeof(int) * n_blocks);
    long *sym_off = (long *)malloc(sizeof(long) * n_blocks);
    int *pos_in_block = (int *)malloc(sizeof(int) * n);
    int mo = 0; long so = 0;
    for (int b = 0; b < n_blocks; ++b) {
        mode_off[b] = mo; sym_off[b] = so;
        for (int k = 0; k < block_dim[b]; ++k) pos_in_block[block_modes[mo + k]] = k;
        mo += block_dim[b];
        so += (long)n_syms * block_dim[b] * block_dim[b];
    }
    double *fy = (double *)malloc(sizeof(double) * n);
    double *fp = (double *)malloc(sizeof(double) * n);
    for (int m = 0; m < n; ++m) {
        double nb = bose(w[m], T);
        fy[m] = 2.0 * w[m] / (1.0 + 2.0 * nb);
        fp[m] = (1.0 + 2.0 * nb) / (2.0 * w[m]);
    }
    double n_eff = 0;
    for (int i = 0; i < N; ++i) n_eff += rho[i];

    if (n_threads <= 0) n_threads = 1;
    double *f_acc = (double *)calloc((size_t)n_threads * n, sizeof(double));
    double *d_acc = (double *)calloc((size_t)n_threads * n * n, sizeof(double));

    #pragma omp parallel num_threads(n_threads)
    {
        int tid = omp_get_thread_num();
        int nth = omp_get_num_threads();
        double *fa = f_acc + (size_t)tid * n;
        double *da = d_acc + (size_t)tid * n * n;
        double *u = (double *)malloc(sizeof(double) * n);
        double *f = (double *)malloc(sizeof(double) * n);
        double *a = (double *)malloc(sizeof(double) * n);
        double *g = (double *)malloc(sizeof(double) * n);
        double *Yu = (double *)malloc(sizeof(double) * n);
        double *Yg = (double *)malloc(sizeof(double) * n);
        /* contiguous chunk of the image range per thread (fixed order) */
        long len = b_end - b_begin;
        long chunk = (len + nth - 1) / nth;
        long lo = b_begin + chunk * tid;
        long hi = lo + chunk; if (hi > b_end) hi = b_end;
        for (long bi = lo; bi < hi; ++bi) {
            int i = (int)(bi / n_syms), j = (int)(bi % n_syms);
            const double *x = X + (size_t)i * n, *y = Y + (size_t)i * n;
            for (int m = 0; m < n; ++m) {
                int b = sym_block_id[m];
                int d = block_dim[b];
                const double *Sb = block_syms + sym_off[b] + (size_t)j * d * d
                                   + (size_t)pos_in_block[m] * d;
                const int *modes = block_modes + mode_off[b];
                double su = 0, sf = 0;
                for (int k = 0; k < d; ++k) {
                    su += Sb[k] * x[modes[k]];
                    sf += Sb[k] * y[modes[k]];
                }
                u[m] = su; f[m] = sf; a[m] = fy[m] * su; g[m] = fp[m] * sf;
            }
            double s1 = 0, s2 = 0, s3 = 0, s4 = 0;
            for (int m = 0; m < n; ++m) {
                double tu = 0, tg = 0;
                const double *row = Y1 + (size_t)m * n;
                for (int k = 0; k < n; ++k) { tu += row[k] * u[k]; tg += row[k] * g[k]; }
                Yu[m] = tu; Yg[m] = tg;
            }
            double s2b = 0;
            for (int m = 0; m < n; ++m) {
                s1 += u[m] * Yu[m];
                s2 += Yu[m] * g[m];     /* g.Y1.u  via (Y1 u) . g   (Y1 any) */
                s2b += u[m] * Yg[m];    /* u.Y1.g */
                s3 += a[m] * R1[m];
                s4 += f[m] * R1[m];
            }
            s2 = 0.5 * (s2 + s2b);
            double r = rho[i];
            double al = -r * s1 / 6.0, be = -r * s2 / 3.0;
            double c1 = r * (-s3 / 3.0 + apply_d4 * s1 / 8.0);
            double c2 = r * (-s4 / 3.0 + apply_d4 * s2 / 4.0);
            for (int m = 0; m < n; ++m) {
                fa[m] += al * f[m] + be * a[m];
                double *drow = da + (size_t)m * n;
                double am = a[m], fm = f[m];
                for (int k = 0; k < n; ++k)
                    drow[k] += c1 * (am * f[k] + fm * a[k]) + c2 * am * a[k];
            }
        }
        free(u); free(f); free(a); free(g); free(Yu); free(Yg);
    }
    double norm = n_eff * n_syms;
    for (int m = 0; m < n; ++m) {
        double s = 0;
        for (int t = 0; t < n_threads; ++t) s += f_acc[(size_t)t * n + m];
        f_out[m] = s / norm;
    }
    for (long e = 0; e < (long)n * n; ++e) {
        double s = 0;
        for (int t = 0; t < n_threads; ++t) s += d_acc[(size_t)t * n * n + e];
        d2v_out[e] = s / norm;
    }
    free(f_acc); free(d_acc); free(fy); free(fp);
    free(mode_off); free(sym_off); free(pos_in_block);
    return 0;
}
