/*
 * tdscha_amd — C ABI of the MI355X (gfx950) TD-SCHA Lanczos hot path.
 *
 * Plain pointers and sizes; int return codes (0 = success, 2 = invalid
 * argument, -1 = HIP/runtime failure); never exit().  The message of the last
 * failure on the calling thread is returned by tdscha_last_error().
 *
 * Reference interfaces replaced (SSCHAcode/tdscha, paths relative to its root):
 *   CModules/odd_corr_module.c:491-634  sscha_HP_odd.GetPerturbAverageSym
 *   Modules/tdscha_core.jl:212-238      get_perturb_averages_sym (Julia backend)
 *   Modules/tdscha_core.jl:29-34        init_sparse_symmetries (symmetry cache)
 *   Modules/DynamicalLanczos.py:3549-3651  Lanczos.apply_full_L
 *   Modules/DynamicalLanczos.py:5958-6491  Lanczos.run_FT (symmetric Wigner)
 */
#ifndef TDSCHA_AMD_H
#define TDSCHA_AMD_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct tdscha_ctx tdscha_ctx;

/* Message of the last failure on this thread ("" if none). */
const char* tdscha_last_error(void);

/*
 * Upload an ensemble and its mode-space symmetries; build the symmetric image
 * ensemble (S_j x_i, S_j y_i for every configuration i and symmetry j) on the
 * device.  Replaces the per-call marshalling of X, Y and the symmetry cache of
 * init_sparse_symmetries (Modules/tdscha_core.jl:29-62) and the arguments of
 * GetPerturbAverageSym (CModules/odd_corr_module.c:519-601).
 *   X, Y         [n_configs, n_modes] row-major (host), mass-scaled, mode basis
 *   w            [n_modes] frequencies (Ry), all > 0;  rho [n_configs]
 *   T            temperature (K)
 *   block_dim    [n_blocks] sizes d_b of the degenerate blocks
 *   block_modes  concatenated mode indices of the blocks (length n_modes)
 *   block_syms   n_blocks pointers to f64[n_syms, d_b, d_b] (row = output mode)
 *   sym_block_id [n_modes] block of each mode
 *   flags        reserved (0)
 */
int tdscha_ctx_create(tdscha_ctx** out, int n_modes, int n_configs, const double* X,
                      const double* Y, const double* w, const double* rho, double T, int n_syms,
                      int n_blocks, const int* block_dim, const int* block_modes,
                      const double* const* block_syms, const int* sym_block_id, int flags);
void tdscha_ctx_destroy(tdscha_ctx* ctx);

/* Padded sizes and the K-split of the outer-product GEMM. */
int tdscha_ctx_info(const tdscha_ctx* ctx, int* n_modes, int* n_modes_padded, long* n_images,
                    long* n_images_padded, int* ksplit);

/* Wait for all work queued on the context's stream. */
int tdscha_synchronize(tdscha_ctx* ctx);

/*
 * Drop-in for GetPerturbAverageSym (CModules/odd_corr_module.c:491-634):
 * host pointers, FP64.  f_out[n] is overwritten with <dV/dR>_(1); d2v_out[n,n]
 * is overwritten with the D3 part plus, if apply_d4, the D4 part
 * (LF:2316 overwrite, LF:2743 accumulate).  Y1 is used through its symmetric
 * part (the reference's weights are symmetric in Y1).
 */
int tdscha_perturb_averages(tdscha_ctx* ctx, const double* R1, const double* Y1, int apply_d4,
                            double* f_out, double* d2v_out);

/*
 * Batched device-pointer variant: n_chains (1..8) perturbations at once.
 *   R1_dev [n_chains, n], Y1_dev [n_chains, n, n]  -> f_dev [n_chains, n],
 *   d2v_dev [n_chains, n, n].
 *   precision: 0 = FP64, 1 = split FP16 hi/lo (FP32 accumulation),
 *              2 = split BF16 hi/lo (FP32 accumulation).
 *   stream: hipStream_t to run on (NULL = the context's stream).
 * Replaces get_perturb_averages_sym (Modules/tdscha_core.jl:212-238).
 */
int tdscha_perturb_averages_dev(tdscha_ctx* ctx, int n_chains, const double* R1_dev,
                                const double* Y1_dev, int apply_d4, int precision, double* f_dev,
                                double* d2v_dev, void* stream);

/* Batched host-pointer variant (same layouts as the _dev entry, synchronous). */
int tdscha_perturb_averages_batched(tdscha_ctx* ctx, int n_chains, const double* R1,
                                    const double* Y1, int apply_d4, int precision, double* f_out,
                                    double* d2v_out);

/*
 * Device memory helpers, so callers need no other GPU runtime for buffer
 * handoff.  kind: 0 host->host, 1 host->device, 2 device->host, 3 device->device.
 */
int tdscha_device_alloc(void** out, size_t bytes);
int tdscha_device_free(void* ptr);
int tdscha_memcpy(void* dst, const void* src, size_t bytes, int kind);
int tdscha_device_count(int* count);

#ifdef __cplusplus
}
#endif
#endif /* TDSCHA_AMD_H */
