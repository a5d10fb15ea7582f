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
 *   flags        0, or TDSCHA_CTX_ROTATE_OPERATORS: store the configurations
 *                only and apply the symmetries by rotating the perturbation
 *                (S_j^T R1, S_j^T Y1 S_j) and the results (S_j f, S_j d2v
 *                S_j^T), as the reference does per image on the fly
 *                (CModules/LanczosFunctions.c:2158-2170): the same flops in
 *                n_syms passes over the configurations, n_syms times less
 *                device memory.  Image ranges must then hold whole
 *                configurations.
 */
#define TDSCHA_CTX_ROTATE_OPERATORS 1
int tdscha_ctx_create(tdscha_ctx** out, int n_modes, int n_configs, const double* X,
                      const double* Y, const double* w, const double* rho, double T, int n_syms,
                      int n_blocks, const int* block_dim, const int* block_modes,
                      const double* const* block_syms, const int* sym_block_id, int flags);
void tdscha_ctx_destroy(tdscha_ctx* ctx);

/* Padded sizes and the K-split of the outer-product GEMM.  n_images counts
 * configurations x symmetries; n_images_padded the rows stored on the device
 * (configurations only, padded, with TDSCHA_CTX_ROTATE_OPERATORS). */
int tdscha_ctx_info(const tdscha_ctx* ctx, int* n_modes, int* n_modes_padded, long* n_images,
                    long* n_images_padded, int* ksplit);

/*
 * Number of image ranges the outer-product GEMM is split into (a multiple of
 * 8, one slice per XCD in turn); each range is summed in FP64 afterwards in a
 * fixed order.  Smaller ranges shorten the FP32 accumulation of the split
 * paths.  Re-allocates the partial-sum workspace.
 */
int tdscha_set_ksplit(tdscha_ctx* ctx, int ksplit);

/*
 * Autotuner of the real-space perturbed averages (no counterpart in the
 * reference): times one pass of n_chains (1..8) perturbations at the given
 * precision for each K split around the default (multiples of 8; reps
 * passes each, in two sweeps, the smaller time of a candidate counting),
 * keeps the fastest for that precision and returns it in ksplit_out with its
 * time in best_ms.  Every K split computes the same sums in another order.
 */
int tdscha_autotune(tdscha_ctx* ctx, int n_chains, int precision, int reps, int* ksplit_out,
                    double* best_ms);

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
 * Replaces get_perturb_averages_sym (Modules/tdscha_core.jl:212-238) over
 * the whole index range [1, n_configs * n_syms]; its sub-ranges are
 * tdscha_perturb_averages_range below.
 */
int tdscha_perturb_averages_dev(tdscha_ctx* ctx, int n_chains, const double* R1_dev,
                                const double* Y1_dev, int apply_d4, int precision, double* f_dev,
                                double* d2v_dev, void* stream);

/*
 * Image-range variant of tdscha_perturb_averages_dev with the index
 * semantics of get_perturb_averages_sym(..., start_index, end_index)
 * (Modules/tdscha_core.jl:164-238, called per rank at
 * Modules/DynamicalLanczos.py:3141-3164): only the images
 * bigindex = (config - 1) * n_syms + sym, 1-based, start_index <= bigindex
 * <= end_index, are summed, and the partial result is still divided by the
 * global N_eff * n_syms, so that the partials of any split of
 * [1, n_configs * n_syms] add up to the full call.  Returns 2 unless
 * 1 <= start_index <= end_index <= n_configs * n_syms (the reference's
 * @assert, JC:178-180).
 */
int tdscha_perturb_averages_range(tdscha_ctx* ctx, int n_chains, const double* R1_dev,
                                  const double* Y1_dev, int apply_d4, int precision,
                                  long start_index, long end_index, double* f_dev,
                                  double* d2v_dev, void* stream);

/*
 * Variant of tdscha_perturb_averages_range (FP64, one perturbation) that
 * allocates nothing: no context, no hipMalloc/hipFree, no allocating library
 * call.  It has no counterpart in the reference; it is for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid allocation).  Every argument is a device pointer,
 * with the arrays of get_perturb_averages_sym (Modules/tdscha_core.jl:212-238)
 * in the packed layout of Modules/DynamicalLanczos.py:807-822:
 *   X, Y [n_configs, n_modes] row-major; w [n_modes]; rho [n_configs];
 *   R1 [n_modes]; Y1 [n_modes, n_modes];
 *   syms [n_blocks, n_syms, dmax, dmax] (row = output mode);
 *   n_degeneracy [n_blocks]; degenerate_space [n_blocks, dmax];
 *   sym_block_id [n_modes];  f_out [n_modes], d2v_out [n_modes, n_modes].
 * start_index, end_index as tdscha_perturb_averages_range (1-based,
 * inclusive); normalised by the global sum(rho) * n_syms.  The work is
 * queued on `stream` (NULL = the null stream) and the call returns at once.
 * Intermediates live in one 64 MiB device array per GPU, held until the
 * process ends: n_modes up to 768 (larger gives 2).  Calls that overlap on
 * one GPU (other threads or streams) share that array and race; serialise
 * them.  The tables are trusted as the reference trusts them: entries
 * outside the tables are skipped, never read.
 */
int tdscha_perturb_averages_noalloc(int n_modes, int n_configs, const double* X, const double* Y,
                                    const double* w, const double* rho, const double* R1,
                                    const double* Y1, double T, int apply_d4, int n_syms,
                                    int n_blocks, int dmax, const double* syms,
                                    const int* n_degeneracy, const int* degenerate_space,
                                    const int* sym_block_id, long start_index, long end_index,
                                    double* f_out, double* d2v_out, void* stream);

/* Batched host-pointer variant (same layouts as the _dev entry, synchronous). */
int tdscha_perturb_averages_batched(tdscha_ctx* ctx, int n_chains, const double* R1,
                                    const double* Y1, int apply_d4, int precision, double* f_out,
                                    double* d2v_out);

/*
 * Gamma-only translational projection (Modules/DynamicalLanczos.py:747-791,
 * 3174-3191): after the point-group image average, f <- P_T f and
 * d2v <- pols^T [ (1/n_cells) sum_R P_R (pols d2v pols^T) P_R^T ] pols.
 *   pols [n_cart, n] (host), inv_perms [n_cells, n_cart] Cartesian index maps
 *   (d_cart[inv_perm[J], inv_perm[K]] is the translated matrix), trans_projector
 *   [n, n] = mean of pols^T P_R pols.  n_cells = 0 disables the projection.
 */
int tdscha_set_translations(tdscha_ctx* ctx, int n_cart, const double* pols, int n_cells,
                            const int* inv_perms, const double* trans_projector);

/*
 * L application (Modules/DynamicalLanczos.py:3549-3651) for n_vec (1..8)
 * vectors of length n + n(n+1): harmonic part (DL:2287-2445) plus the
 * anharmonic part (DL:2951-3363), then out = (reverse ? -out : out) + shift*psi.
 * use_wigner selects the Wigner (a', b') or the standard (Upsilon, ReA)
 * representation; transpose applies L^T in the standard one (ignored in the
 * Wigner one, as in the reference).  ignore_v3 drops the anharmonic part,
 * ignore_v4 the D4 term.  precision as for the perturbed averages.
 */
int tdscha_apply_L(tdscha_ctx* ctx, int n_vec, const double* psi, double* out, int use_wigner,
                   int transpose, int ignore_harmonic, int ignore_v3, int ignore_v4, int reverse,
                   double shift, int precision);
int tdscha_apply_L_dev(tdscha_ctx* ctx, int n_vec, const double* psi_dev, double* out_dev,
                       int use_wigner, int transpose, int ignore_harmonic, int ignore_v3,
                       int ignore_v4, int reverse, double shift, int precision, void* stream);

/*
 * Inverse of the harmonic part in the standard representation, the
 * preconditioner behind Lanczos.M_linop (replaces
 * Lanczos.apply_L1_inverse_FT, Modules/DynamicalLanczos.py:2451-2549, with
 * its regularisation of singular and unpopulated pair blocks): n_vec (1..8)
 * host vectors of length n + n(n+1); transpose applies the transposed
 * blocks; ignore_harmonic gives zero.
 */
int tdscha_apply_L1_inverse(tdscha_ctx* ctx, int n_vec, const double* psi, double* out,
                            int transpose, int ignore_harmonic);

/*
 * Device-resident symmetric Lanczos in the Wigner representation
 * (run_FT with run_simm, Modules/DynamicalLanczos.py:6032-6240) for n_chains
 * (1..64) starting vectors psi0 [n_chains, n + n(n+1)] (host), batched through
 * every L application (8 chains per perturbed-average pass).  Each chain
 * stops at |b| < tol or after caps[c] steps (caps may be NULL: n_steps for
 * all).  reorth = 1 adds full classical Gram-Schmidt reorthogonalisation
 * (two passes) against the whole basis.  use_graph = 1 captures one step as
 * a hipGraph and replays it; 2 captures the whole run and launches it once.
 * The same limits and use_graph values hold for
 * tdscha_lanczos_run_biconjugate below.
 * Outputs: a_out, b_out [n_chains, n_steps]; n_done [n_chains] = number of a
 * coefficients; basis_out (optional) [n_chains, n_steps + 1, len];
 * elapsed_ms (optional) device time of the step loop.
 */
int tdscha_lanczos_run(tdscha_ctx* ctx, int n_chains, const double* psi0, int n_steps,
                       const int* caps, double tol, int reorth, int precision,
                       int ignore_harmonic, int ignore_v3, int ignore_v4, int use_graph,
                       double* a_out, double* b_out, int* n_done, double* basis_out,
                       double* elapsed_ms);

/*
 * Device-resident biconjugate Lanczos (run_FT without run_simm,
 * Modules/DynamicalLanczos.py:6126-6240, n_rep_orth = 0): right vectors q
 * and left vectors p with plain dot products, L q and L^T p per step
 * (transpose semantics of tdscha_apply_L, in the standard or the Wigner
 * representation), the reference's a, b, c and s_norm bookkeeping.  Chains
 * stop when |b| or |c| < tol (a kept, b and c dropped) or at caps[c] steps
 * (caps may be NULL).  reverse / shift as tdscha_apply_L.  use_graph as
 * tdscha_lanczos_run (0 plain launches, 1 one step captured and replayed,
 * 2 the whole run captured as one graph).  a_out, b_out, c_out
 * [n_chains, n_steps]; n_done[c] = number of a coefficients.
 */
int tdscha_lanczos_run_biconjugate(tdscha_ctx* ctx, int n_chains, const double* psi0,
                                   int n_steps, const int* caps, double tol, int use_wigner,
                                   int precision, int ignore_harmonic, int ignore_v3,
                                   int ignore_v4, int reverse, double shift, int use_graph,
                                   double* a_out, double* b_out, double* c_out, int* n_done,
                                   double* elapsed_ms);

/*
 * Continued-fraction Green function with the standard terminator
 * (Modules/DynamicalLanczos.py:5032-5162) on n_w frequencies.  a [n_a],
 * b [n_b] (n_b >= n_a - 1), c [n_b] or NULL (c = b).  gf_out [n_w, 2] holds
 * (Re, Im), Wigner sign convention applied, times modulus.
 */
int tdscha_continued_fraction(int n_a, const double* a, int n_b, const double* b,
                              const double* c, int n_w, const double* w, int use_wigner,
                              int use_terminator, int last_average, double smearing,
                              int smooth_ramp, int reverse, double shift, double modulus,
                              double* gf_out);

/*
 * q-space (Bloch basis) context: replaces the ensemble arguments of
 * get_perturb_averages_qspace (Modules/tdscha_qspace.jl:639-712) and the
 * q-space attributes of QSpaceLanczos (Modules/QSpaceLanczos.py:150-295).
 *   X_q, Y_q  complex128 [n_q, n_configs, n_bands] (interleaved re, im),
 *             reference layout X_q[iq, config, nu]
 *   w_q       [n_bands, n_q] frequencies (Ry); valid_modes_q [n_bands, n_q]
 *             (0 = masked acoustic band); rho [n_configs]; T (K)
 * n_bands must be 3 or 6.  The context starts with one symmetry, the
 * identity (as QSpaceLanczos.prepare_symmetrization(no_sym=True));
 * tdscha_qctx_symmetrize below adds the point-group images.
 */
typedef struct tdscha_qctx tdscha_qctx;
int tdscha_qctx_create(tdscha_qctx** out, int n_q, int n_bands, int n_configs, const double* X_q,
                       const double* Y_q, const double* w_q, const int* valid_modes_q,
                       const double* rho, double T);
void tdscha_qctx_destroy(tdscha_qctx* ctx);

/*
 * Point-group images (replaces the sparse q-space symmetry matrices cached by
 * QSpaceLanczos.prepare_symmetrization, Modules/QSpaceLanczos.py:1471-1633,
 * and their use per (configuration, symmetry) in the fused kernel,
 * Modules/tdscha_qspace.jl:496-502).  Symmetry j maps the Bloch components of
 * source q-point p to q-point q_perm[j * n_q + p] through the n_bands x
 * n_bands complex block D[((j * n_q + p) * n_bands + nu) * n_bands + mu]
 * (interleaved re, im).  Expands the context's ensemble on the device into
 * its n_sym * n_configs images (image j * n_configs + i), normalised by
 * n_sym * sum(rho); once per context (a second call is rejected), before any
 * run.  Rows of q_perm that are not permutations are rejected.
 */
int tdscha_qctx_symmetrize(tdscha_qctx* ctx, int n_sym, const int* q_perm, const double* D);

/* psi length of the layout built from one q_pair_map (QL:313-343). */
int tdscha_qspace_psi_size(int n_q, int n_bands, const int* q_pair_map, int* psi_size);

/*
 * QSpaceLanczos.apply_full_L (Modules/QSpaceLanczos.py:870-893) for n_chains
 * (1..256) vectors, chain c perturbed at iq_pert[c] with pair map
 * q_pair_map[c, n_q] (q1 + q2 = q_pert + G, QL:297-331).  psi and out hold
 * the chains' complex vectors one after the other (interleaved re, im), each
 * in the reference layout of its own pair map.
 */
int tdscha_qspace_apply_L(tdscha_qctx* ctx, int n_chains, const int* iq_pert,
                          const int* q_pair_map, const double* psi, double* out,
                          int ignore_harmonic, int ignore_v3, int ignore_v4);

/*
 * Static Liouvillian of the free-energy Hessian (replaces
 * QSpaceHessian._apply_L_static_q, Modules/QSpaceHessian.py:605-692) on psi
 * in the static layout [R (n_bands), one W sector of the same blocks as a']:
 * R <- w^2 R - f_pert, W <- W / Lambda + d2v, with the perturbed averages of
 * R1 = R and alpha = -2 Y_w1 Y_w2 W.  Same batching as tdscha_qspace_apply_L.
 */
int tdscha_qspace_static_psi_size(int n_q, int n_bands, const int* q_pair_map, int* psi_size);
int tdscha_qspace_apply_L_static(tdscha_qctx* ctx, int n_chains, const int* iq_pert,
                                 const int* q_pair_map, const double* psi, double* out,
                                 int ignore_harmonic, int ignore_v3, int ignore_v4);

/*
 * Hermitian Lanczos with full reorthogonalisation (QSpaceLanczos.run_FT with
 * reorthogonalize=True, QL:967-1182) for n_chains chains at once, device
 * resident.  Chain c stops when |b| < tol or after caps[c] steps (caps may
 * be NULL: n_steps).  use_graph = 1 captures the whole run as one hipGraph and
 * launches it once.  a_out, b_out, c_out (may be NULL) [n_chains, n_steps];
 * n_done[c] = number of a coefficients (b has n_done[c] entries, or one fewer
 * when the chain stopped on |b| < tol).
 */
int tdscha_qspace_lanczos_run(tdscha_qctx* ctx, int n_chains, const int* iq_pert,
                              const int* q_pair_map, const double* psi0, int n_steps,
                              const int* caps, double tol, int ignore_harmonic, int ignore_v3,
                              int ignore_v4, int use_graph, double* a_out, double* b_out,
                              double* c_out, int* n_done, double* elapsed_ms);

/*
 * Kernel polynomial method for a batch of chains (replaces the loop of
 * QSpaceKPM.run_KPM, Modules/QSpaceKPM.py:260-454): Chebyshev moments
 * mu_n = <v0, T_n(L~) v0>_M of L~ = (L - rescale[2c+1]) / rescale[2c], two
 * per L application, n = 0 .. n_moments-1, every step on the device (one
 * graph launch with use_graph).  vm, v (complex, concatenated over chains in
 * each chain's psi layout) hold T_{n-1}, T_n: for a fresh run (n_have = 0)
 * both hold the metric-normalised v0 and moments[c][0] is set to 1; to
 * continue, n_have moments of moments[c][0 .. n_have) and the vm, v of that
 * run.  On return moments[c][0 .. n_moments) and the final vm, v.
 */
int tdscha_qspace_kpm_run(tdscha_qctx* ctx, int n_chains, const int* iq_pert,
                          const int* q_pair_map, const double* rescale, int n_moments,
                          int n_have, double* moments, double* vm, double* v,
                          int ignore_harmonic, int ignore_v3, int ignore_v4, int use_graph,
                          double* elapsed_ms);

/*
 * Tile choices of the q-space kernels: configurations per workgroup of the
 * weights kernel (64, 128, 256), configurations per workgroup of the d2v
 * kernel (0 = automatic) and basis vectors per workgroup of the
 * reorthogonalisation (8, 16, 32).  tdscha_qspace_autotune times one L
 * application of the given batch for every weights/d2v-chunk/d2v-form candidate (reps
 * repetitions each), keeps the fastest and returns it in tiles_out[3].
 */
int tdscha_qspace_set_tiles(tdscha_qctx* ctx, int weights_tile, int d2v_chunk, int reorth_jtile);
int tdscha_qspace_get_tiles(const tdscha_qctx* ctx, int* tiles_out);
/*
 * Form of the d2v kernel: 1 (default) stages configuration slabs of every
 * q-point in LDS and gives each thread one (chain, pair) block; 0 gives each
 * workgroup one (chain, pair) block and each thread a configuration.  Grids
 * too large for a one-configuration slab use form 0 whatever is set.  Both
 * compute the same sums in different orders.  tdscha_qspace_autotune also
 * picks the form; get returns the form in use.
 */
int tdscha_qspace_set_d2v_form(tdscha_qctx* ctx, int form);
int tdscha_qspace_get_d2v_form(const tdscha_qctx* ctx, int* form);
int tdscha_qspace_autotune(tdscha_qctx* ctx, int n_chains, const int* iq_pert,
                           const int* q_pair_map, int reps, int* tiles_out, double* best_ms);

/*
 * Learned continued-fraction terminator: a fully connected FP32 network
 * n_in -> width (n_hidden layers, SiLU) -> n_out (linear), trained and run
 * by the library's own kernels.  No counterpart in the reference, whose only
 * terminator is the constant one of Modules/DynamicalLanczos.py:5089-5122;
 * the network's outputs extend the coefficients handed to that terminator.
 * Parameters are one FP32 array [W0 (d1 x d0), b0 (d1), W1 (d2 x d1), ...],
 * W row-major [out][in].  init_params may be NULL (all zero).
 */
typedef struct tdscha_mlp tdscha_mlp;
int tdscha_mlp_create(tdscha_mlp** out, int n_in, int width, int n_hidden, int n_out,
                      int max_batch, const float* init_params);
void tdscha_mlp_destroy(tdscha_mlp* net);
int tdscha_mlp_n_params(const tdscha_mlp* net, long* n_params);
int tdscha_mlp_get_params(tdscha_mlp* net, float* params_out);
/* reset_optimizer = 1 also clears the Adam moments and step count. */
int tdscha_mlp_set_params(tdscha_mlp* net, const float* params, int reset_optimizer);
/* Inference: y [batch, n_out] from x [batch, n_in] (host arrays). */
int tdscha_mlp_forward(tdscha_mlp* net, int batch, const float* x, float* y);
/* Loss kind 0: sum_o w_o (y_o - t_o)^2 averaged over the batch (w defaults to 1). */
int tdscha_mlp_set_mse_weights(tdscha_mlp* net, const float* w);
/*
 * Loss kind 1 (n_out = 10): outputs z map to tail parameters
 * p = out_mean + out_std * z in units of b0 (a_inf, b_inf, da[4], db[4]); the
 * chain's K coefficients are extended by a_j = b0 (a_inf + da_j),
 * b_j = b0 (b_inf + db_j), j = 0..3, and the terminator (a_inf, b_inf) b0;
 * the loss is loss_scale * mean over samples and the n_w frequencies w of
 * (S(w) / max|S_ref| - ref)^2 with S = -Im G the Wigner continued fraction
 * (smearing eta) of Modules/DynamicalLanczos.py:5032-5162.
 * Per-sample data: chains [n, 2K + 2] = (a[K], b[K], b0, 1 / max|S_ref|),
 * ref [n, n_w] = S_ref / max|S_ref|.
 */
int tdscha_mlp_set_cf_loss(tdscha_mlp* net, int K, int n_w, const double* w, double smearing,
                           const double* out_mean, const double* out_std, double loss_scale);
/* Loss and parameter gradient of one batch (kind 0: target; kind 1: chains, ref). */
int tdscha_mlp_loss_grad(tdscha_mlp* net, int kind, int batch, const float* x,
                         const float* target, const double* chains, const double* ref,
                         float* loss, float* grad_out);
/* One Adam step on one batch (bias-corrected Adam, eps outside the square root). */
int tdscha_mlp_adam_step(tdscha_mlp* net, int kind, int batch, const float* x,
                         const float* target, const double* chains, const double* ref, float lr,
                         float beta1, float beta2, float eps, float* loss);
/* Minibatch Adam over a device-resident training set, reshuffled every epoch. */
int tdscha_mlp_train(tdscha_mlp* net, int kind, long n_samples, const float* x,
                     const float* target, const double* chains, const double* ref, int batch,
                     int epochs, float lr, float beta1, float beta2, float eps,
                     unsigned long long seed, float* epoch_loss);

/*
 * Time the perturbed-average pipeline with HIP events on the context stream:
 * reps passes (after one warm-up) for n_chains chains at the given precision.
 * ms_out[4] = average ms of {weights GEMM, outer-product GEMM, K-split
 * reduction, total}.
 */
int tdscha_profile_perturb_averages(tdscha_ctx* ctx, int n_chains, int precision, int reps,
                                    double* ms_out);

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
