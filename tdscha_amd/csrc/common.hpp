// Shared definitions for the tdscha_amd HIP library (gfx950 / CDNA4).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

namespace tdscha {

// ---------------------------------------------------------------- errors
void set_error(const std::string& msg);
const char* last_error();

#define TDS_CHECK_HIP(expr)                                                   \
  do {                                                                        \
    hipError_t _e = (expr);                                                   \
    if (_e != hipSuccess) {                                                   \
      ::tdscha::set_error(std::string(#expr) + ": " + hipGetErrorString(_e)); \
      return -1;                                                              \
    }                                                                         \
  } while (0)

#define TDS_REQUIRE(cond, code, msg)            \
  do {                                          \
    if (!(cond)) {                              \
      ::tdscha::set_error(msg);                 \
      return (code);                            \
    }                                           \
  } while (0)

// ------------------------------------------------------------ constants
constexpr double RY_TO_K = 157887.32400374097;
constexpr int kTile = 128;      // column tile of the weighted outer-product GEMM
constexpr int kRowTile = 64;    // image rows per workgroup of the weights GEMM
constexpr int kMaxChains = 8;   // perturbations applied in one batched call
constexpr int kPartBlocks = 1024;  // blocks of the grid-wide maxima in the split prep
// waves per work-group of the FP64 outer-product GEMM (4: 64 x 64 per wave,
// two work-groups per CU; 8: 64 x 32 per wave, two work-groups and four
// waves per SIMD, measured 17.6 ms against 22.6 ms at the 128-atom
// workload); its launches use 64 * this many threads
#ifndef TDS_F64_OUTER_WAVES
#define TDS_F64_OUTER_WAVES 8
#endif
constexpr int kF64OuterThreads = 64 * TDS_F64_OUTER_WAVES;

inline int round_up(int x, int m) { return (x + m - 1) / m * m; }
inline long round_up_l(long x, long m) { return (x + m - 1) / m * m; }

// MFMA operand vector types
// direct global -> LDS copy of 16 bytes per lane: lane i's bytes land at
// lds_base + 16 i (lds_base is wave-uniform)
typedef __attribute__((address_space(3))) void* lds_ptr_t;
typedef __attribute__((address_space(1))) void* gbl_ptr_t;
__device__ __forceinline__ void dma16(const void* g, void* lds_base) {
  __builtin_amdgcn_global_load_lds((gbl_ptr_t)g, (lds_ptr_t)lds_base, 16, 0, 0);
}

// s_waitcnt vmcnt(N) (expcnt and lgkmcnt not waited), gfx9 encoding
#define TDS_WAIT_VMCNT(N) \
  __builtin_amdgcn_s_waitcnt(((N) & 15) | (((N) >> 4) << 14) | 0x0F70)

// Work-group barrier without a memory fence.  The direct-to-LDS copies are
// completed explicitly (vmcnt) before it; this waits for the wave's own LDS
// reads and writes (lgkmcnt(0)), and the empty asm statements keep the
// compiler from moving memory accesses across it.
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_s_waitcnt(0xC07F);
  asm volatile("" ::: "memory");
  __builtin_amdgcn_s_barrier();
  asm volatile("" ::: "memory");
}

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));
typedef _Float16 v8h __attribute__((ext_vector_type(8)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));

enum Precision { PREC_FP64 = 0, PREC_FP16X2 = 1, PREC_BF16X2 = 2 };

// Per-precision image storage for the split-precision kernels.
struct SplitStore {
  void* Uh = nullptr;   // image-major blocks, hi part of the scaled U (weights A operand)
  void* Ul = nullptr;   //   lo part
  float* FA = nullptr;  // image-major FP32 blocks of F (weights epilogue)
  float* UA32 = nullptr;  // image-major FP32 blocks of U (weights epilogue)
  void* UhT = nullptr;  // mode-major blocks, hi part of the scaled U (outer GEMM)
  void* UlT = nullptr;  //   lo part
  float* XT = nullptr;  // mode-major FP32 blocks of F (outer B build)
  float* YT = nullptr;  // mode-major FP32 blocks of fY U / 2 (outer B build)
  double* su = nullptr;  // [np] power-of-two column scale of U
  double* sf = nullptr;  // [np] power-of-two column scale of F
  double* umax = nullptr;  // [np] max |U| per column (unscaled)
  double* fmax = nullptr;  // [np] max |F| per column (unscaled)
  float* colw = nullptr;   // [3, np] 1/su, 1/sf, fpsi in FP32 (weights epilogue)
  bool ready = false;
};

// ------------------------------------------------------------- context
struct Ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  // problem
  int n = 0;      // modes
  int np = 0;     // padded modes (multiple of kTile)
  int N = 0;      // configurations
  int n_syms = 0;
  long M = 0;     // images N * n_syms
  long Mp = 0;    // padded images (multiple of the K split granule)
  double T = 0;
  double n_eff = 0;
  // image range of the next perturbed-average pass, [row_lo, row_hi) in
  // image order b = i * n_syms + j (default: every image)
  long row_lo = 0;
  long row_hi = 0;
  // symmetric image ensemble, device resident
  double* U = nullptr;      // [Mp, np] S_j x_i
  double* F = nullptr;      // [Mp, np] S_j y_i
  double* rho_img = nullptr;// [Mp]
  double* fy = nullptr;     // [np]
  double* fpsi = nullptr;   // [np]
  double* w = nullptr;      // [np]
  // per-call buffers
  double* R1 = nullptr;     // [kMaxChains, np]
  double* Y1 = nullptr;     // [kMaxChains, np, np]
  double* coef = nullptr;   // [kMaxChains, Mp, 4] (c1, c2, alpha, beta)
  double* zpart = nullptr;  // [kMaxChains, ksplit, np, np]
  double* fpart = nullptr;  // [kMaxChains, ksplit, np]
  double* f_out = nullptr;  // [kMaxChains, np]
  double* d2v_out = nullptr;// [kMaxChains, np, np]
  int ksplit = 0;
  // split precision
  SplitStore split[3];
  void* Y1h = nullptr;      // [kMaxChains, np, np] (16-bit)
  void* Y1l = nullptr;
  double* ysc = nullptr;    // [kMaxChains] power-of-two scale of split Y1
  double* ysc3 = nullptr;   // [kMaxChains] scale of the folded fY R1 column
  double* bsc = nullptr;    // [kMaxChains, np] column scale of split B
  double* cmax = nullptr;   // [kMaxChains, 2] max|c1|, max|c2| (device)
  double* pmax = nullptr;   // [kMaxChains, kPartBlocks, 2] per-block maxima of the split prep
  float* epi = nullptr;     // [kMaxChains, 2, np] fY R1 and R1 (weights epilogue)
  float* coef32 = nullptr;  // [Mp/16, kMaxChains, 4, 16] FP32 coefficients for the outer GEMM
  double* wpart = nullptr;  // [kMaxChains, np/128, Mp, 4] weights row-dot partials
  int ksplit_s = 0;          // K split of the split-precision outer GEMM
  float* zpart32 = nullptr;  // [kMaxChains, ksplit_s, np, np] raw FP32 partial tiles
  double* fpart_s = nullptr; // [kMaxChains, ksplit_s, np]
  double* sbsc = nullptr;    // [kMaxChains, np] B column scales of the last call
  // Wigner / harmonic helpers
  double* pair_ca = nullptr;   // [npairs]
  double* pair_cb = nullptr;
  double* pair_sm = nullptr;
  double* pair_sp = nullptr;
  double* pair_harm_a = nullptr;
  double* pair_harm_b = nullptr;
  int* pair_row = nullptr;
  int* pair_col = nullptr;
  // standard representation helpers
  double* std_coef = nullptr;  // [8, npairs] harmonic 2x2 block coefficients
  double* inv_coef = nullptr;  // [4, npairs] inverse harmonic 2x2 blocks (preconditioner)
  double* yw = nullptr;        // [np] Y_w
  double* reaw = nullptr;      // [np] ReA_w
  // gamma-only translational projector
  int n_cells = 0;
  int n_cart = 0;
  double* pols = nullptr;      // [n_cart, n]
  int* inv_perms = nullptr;    // [n_cells, n_cart]
  // translation group orbits (when the permutations form a group): dof J
  // lies in the orbit of rep_dof[jrep[J]], reached by inv_perms[ju[J]]
  int n_rep = 0;
  int* rep_dof = nullptr;      // [n_rep]
  int* jrep = nullptr;         // [n_cart]
  int* ju = nullptr;           // [n_cart]
  double* rep_sum = nullptr;   // [kMaxChains, n_rep, n_cart]
  double* tproj = nullptr;     // [n, n]
  double* cart_tmp = nullptr;  // [kMaxChains, n_cart, n_cart] x2
  double* cart_tmp2 = nullptr;
  double* mode_tmp = nullptr;  // [kMaxChains, n, n_cart]
  // scratch for the anharmonic wrapper
  double* psi_dev = nullptr;   // [kMaxChains, len_psi] x2
  double* psi_out_dev = nullptr;
  double* f_fin = nullptr;     // [kMaxChains, n] perturbed averages (normalised)
  double* d2v_fin = nullptr;   // [kMaxChains, n, n]
  double* f_proj = nullptr;    // [kMaxChains, n] translation-projected f
  double* stage_in = nullptr;  // [kMaxChains, n + n*n] host<->device staging
  double* stage_out = nullptr;
  // host-side copies of the symmetry description
  std::vector<int> block_dim, block_off, block_modes, sym_block_id;
  // operator-rotation form (flag TDSCHA_CTX_ROTATE_OPERATORS): U, F hold the
  // configurations themselves (n_syms = 1 for the kernels) and every pass is
  // run on rotated perturbations S_j^T R1, S_j^T Y1 S_j, its results rotated
  // back and summed over the rot_nsym symmetries
  bool rotate = false;
  int rot_nsym = 1;
  int rt_dmax = 0;             // largest block
  int* rt_cols = nullptr;      // [dmax, n] modes of each mode's block (position-major)
  int* rt_d = nullptr;         // [n] size of each mode's block
  double* rt_sr = nullptr;     // [rot_nsym, dmax, n] S_j[m, cols(m)[t]]   (row m of S_j)
  double* rt_sc = nullptr;     // [rot_nsym, dmax, n] S_j[cols(m)[t], m]   (column m)
  int* rt_bs = nullptr;        // [n] first mode of each mode's block (panel kernels)
  int* rt_pstart = nullptr;    // [rt_npanel + 1] first mode of each panel, then np
  int rt_npanel = 0, rt_nreal = 0;
  bool rt_panels = false;      // every block consecutive modes, at most 24: panel kernels
  double* R1_base = nullptr;   // [kMaxChains, np] unrotated padded inputs of a call
  double* Y1_base = nullptr;   // [kMaxChains, np, np]
  double* rot_t1 = nullptr;    // [kMaxChains, np, np] rotation intermediates
  double* rot_t2 = nullptr;
  double* rot_f = nullptr;     // [kMaxChains, np]
};

int ensure_split(Ctx* c, int prec);

}  // namespace tdscha
