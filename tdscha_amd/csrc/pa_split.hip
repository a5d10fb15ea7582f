// Split-precision perturbed averages: FP16 (or BF16) hi/lo compensated
// products on MFMA (v_mfma_f32_32x32x16_{f16,bf16}) with FP32 accumulation.
//
// Every operand x is held as x ~ (hi + lo) / s with hi = T(x s), lo =
// T(x s - hi) (T = fp16 or bf16) and s a power-of-two per column, so a
// product A.B ~ (Ah.Bh + Ah.Bl + Al.Bh) / (sA sB): three MFMAs, relative
// error ~ 2^-22 (fp16) or 2^-16 (bf16) per element, FP32 accumulation.
// The per-image coefficients (c1, c2, alpha, beta) and all reductions across
// work-groups stay FP64.
//
// Operand storage is blocked: every GEMM operand is a sequence of 4 KB
// blocks of 128 rows x 16 K-elements, each block byte-for-byte the LDS image
// the MFMA fragment reads want (16-byte chunks XOR-swizzled by row).  A block
// is therefore copied global -> LDS with four 1 KB direct-to-LDS loads
// (global_load_lds_dwordx4), fully coalesced, without passing through VGPRs.
//   UA       [Mp/128][np/16] blocks    image-major (weights GEMM, K = modes)
//   UT       [np/128][Mp/16] blocks    mode-major  (outer GEMM,   K = images)
//   YB       [chain][np/128][np/16]    Y1 columns  (weights GEMM B operand)
// The outer GEMM builds its B operand from F and fY U / 2, and the weights
// epilogue reads F; none of them changes between calls, so they are kept as
// FP32 blocks XT, YT (mode-major) and FA (image-major), same element order
// as the 16-bit blocks (8 KB blocks): as many bytes as a hi/lo pair and no
// conversions per use.
//
// All chains of one pass (up to 3) share a work-group, so every image block
// is read once per pass.
#include <cmath>
#include <type_traits>

#include "common.hpp"

namespace tdscha {

int launch_pa_fp64(Ctx* c, int nc, int apply_d4, hipStream_t s, hipEvent_t* ev);

typedef float v16f __attribute__((ext_vector_type(16)));

// 32x32x16 MFMA (sustains ~2.3 PF/s in FP16 on MI355X, against ~1.4 PF/s for
// the 16x16x32 shape).  Operand layout: A[i][k]: i = lane % 32,
// k = 8 * (lane / 32) + 0..7; B likewise with j; result D[i][j]: j = lane % 32,
// i = 8 * (r / 4) + 4 * (lane / 32) + r % 4 for r = 0..15.
template <typename H>
struct Mma;
template <>
struct Mma<_Float16> {
  typedef v8h vec;
  static __device__ __forceinline__ v16f mma(vec a, vec b, v16f c) {
    return __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, c, 0, 0, 0);
  }
};
template <>
struct Mma<__bf16> {
  typedef v8bf vec;
  static __device__ __forceinline__ v16f mma(vec a, vec b, v16f c) {
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
  }
};

typedef unsigned short u16;
typedef u16 u16x8 __attribute__((ext_vector_type(8)));
typedef u16 u16x4 __attribute__((ext_vector_type(4)));

template <typename H>
__device__ __forceinline__ float to_f(u16 b) {
  H h;
  __builtin_memcpy(&h, &b, 2);
  return (float)h;
}
template <typename H>
__device__ __forceinline__ u16 from_f(float x) {
  H h = (H)x;
  u16 b;
  __builtin_memcpy(&b, &h, 2);
  return b;
}
template <typename H>
__device__ __forceinline__ u16 from_d(double x) {
  H h = (H)x;
  u16 b;
  __builtin_memcpy(&b, &h, 2);
  return b;
}

// power of two s with max * s < 2^12 (1 for max == 0)
__device__ __forceinline__ double pow2_scale(double mx) {
  if (!(mx > 0.0)) return 1.0;
  int e;
  (void)frexp(mx, &e);
  return ldexp(1.0, 12 - e);
}

__device__ __forceinline__ void atomic_max_pos(double* addr, double v) {
  // non-negative doubles order like their bit patterns
  atomicMax(reinterpret_cast<unsigned long long*>(addr), (unsigned long long)__double_as_longlong(v));
}

// ------------------------------------------------------------ block layout
constexpr int kBR = 128;            // rows of a block
constexpr int kBK = 16;             // K-elements of a block (one 32x32x16 step)
constexpr int kBE = kBR * kBK;      // 16-bit elements of a block (4 KB)
// Chunk-major: the block is two 2 KB halves (K-elements 0..7 and 8..15), each
// 128 rows x 16 bytes.  An MFMA fragment read (32 consecutive rows, one half)
// is then 512 contiguous bytes per half-wave, free of bank conflicts, and a
// linear 16-byte-per-lane pass over a block gives every wave one half.
__host__ __device__ __forceinline__ int blk_off(int row, int k) {
  return (k >> 3) * (kBR * 8) + row * 8 + (k & 7);
}
// fragment of rows 32 b .. 32 b + 31: lane reads row 32 b + lane % 32, half lane / 32
__device__ __forceinline__ int frag_off(int b, int lane) {
  return (lane >> 5) * (kBR * 8) + (32 * b + (lane & 31)) * 8;
}


// ------------------------------------------------------------ static prep
__global__ void colmax_kernel(const double* __restrict__ U, const double* __restrict__ F, long Mp,
                              int np, double* __restrict__ umx_out, double* __restrict__ fmx_out) {
  const int col = blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= np) return;
  const long rows_per = (Mp + gridDim.y - 1) / gridDim.y;
  const long lo = blockIdx.y * rows_per;
  long hi = lo + rows_per;
  if (hi > Mp) hi = Mp;
  double mu = 0.0, mf = 0.0;
  for (long r = lo; r < hi; ++r) {
    mu = fmax(mu, fabs(U[r * np + col]));
    mf = fmax(mf, fabs(F[r * np + col]));
  }
  atomic_max_pos(&umx_out[col], mu);
  atomic_max_pos(&fmx_out[col], mf);
}

__global__ void scales_kernel(const double* __restrict__ umx, const double* __restrict__ fmx,
                              int np, double* __restrict__ su, double* __restrict__ sf) {
  const int col = blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= np) return;
  su[col] = pow2_scale(umx[col]);
  sf[col] = pow2_scale(fmx[col]);
}

// hi/lo split of the images into the image-major and mode-major block
// layouts, and the FP32 mode-major blocks of F and fY U / 2 (32 x 32 tiles
// through LDS for the transpose)
template <typename H>
__global__ void split_images(const double* __restrict__ U, const double* __restrict__ F, long Mp,
                             int np, const double* __restrict__ su,
                             const double* __restrict__ sf, const double* __restrict__ fy,
                             u16* __restrict__ UAh, u16* __restrict__ UAl,
                             float* __restrict__ FA, float* __restrict__ UA32,
                             u16* __restrict__ UTh, u16* __restrict__ UTl,
                             float* __restrict__ XT, float* __restrict__ YT) {
  __shared__ u16 t[2][32][33];
  __shared__ float t32[2][32][33];
  const long r0 = (long)blockIdx.x * 32;
  const int c0 = blockIdx.y * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
  const int nkc = np / kBK;
  for (int yy = ty; yy < 32; yy += 8) {
    const long r = r0 + yy;
    const int m = c0 + tx;
    const double u0 = U[r * np + m], f0 = F[r * np + m];
    const double u = u0 * su[m];
    const u16 uh = from_d<H>(u);
    const u16 ul = from_d<H>(u - (double)to_f<H>(uh));
    const long a = ((r / kBR) * nkc + m / kBK) * kBE + blk_off((int)(r % kBR), m % kBK);
    UAh[a] = uh;
    UAl[a] = ul;
    FA[a] = (float)f0;
    UA32[a] = (float)u0;
    t[0][yy][tx] = uh;
    t[1][yy][tx] = ul;
    t32[0][yy][tx] = (float)f0;
    t32[1][yy][tx] = (float)(0.5 * fy[m] * u0);
  }
  __syncthreads();
  const long nki = Mp / kBK;
  for (int yy = ty; yy < 32; yy += 8) {
    const int m = c0 + yy;
    const long r = r0 + tx;
    const long a = ((long)(m / kBR) * nki + r / kBK) * kBE + blk_off(m % kBR, (int)(r % kBK));
    UTh[a] = t[0][tx][yy];
    UTl[a] = t[1][tx][yy];
    XT[a] = t32[0][tx][yy];
    YT[a] = t32[1][tx][yy];
  }
}

// ------------------------------------------------------------ per-call prep
// One power-of-two scale per chain for the split Y1: sy with
// max_jk |Y1[j][k] / su[k]| * sy < 2^12.  Block maxima over slices of Y1
// (grid kYmaxBlocks x chains), then one block per chain.
constexpr int kYmaxBlocks = 64;
__global__ void ymax_part_kernel(const double* __restrict__ Y1all, const double* __restrict__ su,
                                 int np, double* __restrict__ pmax) {
  const int chain = blockIdx.y;
  const double* Y = Y1all + (long)chain * np * np;
  const long nn = (long)np * np, per = (nn + gridDim.x - 1) / gridDim.x;
  const long lo = blockIdx.x * per, hi = lo + per < nn ? lo + per : nn;
  __shared__ double red[256];
  double mx = 0.0;
  for (long e = lo + threadIdx.x; e < hi; e += blockDim.x) mx = fmax(mx, fabs(Y[e] / su[e % np]));
  red[threadIdx.x] = mx;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + o]);
    __syncthreads();
  }
  if (threadIdx.x == 0) pmax[((long)chain * kPartBlocks + blockIdx.x) * 2] = red[0];
}

__global__ void ymax_kernel(const double* __restrict__ pmax, int nblocks, double* __restrict__ ysc) {
  const int chain = blockIdx.x;
  __shared__ double red[64];
  double mx = 0.0;
  for (int b = threadIdx.x; b < nblocks; b += blockDim.x)
    mx = fmax(mx, pmax[((long)chain * kPartBlocks + b) * 2]);
  red[threadIdx.x] = mx;
  __syncthreads();
  for (int o = 32; o > 0; o >>= 1) {
    if (threadIdx.x < o) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + o]);
    __syncthreads();
  }
  if (threadIdx.x == 0) ysc[chain] = pow2_scale(red[0]);
}

// When the modes are padded (n < np), the fp16 path lets the row-dot
// s3 = u . (fY R1) ride the weights GEMM: the last padding column of YB holds fY R1 / su with a scale
// of its own per chain (sy3: max_k |fY R1 / su| sy3 < 2^12), and the
// accumulator of that column is s3 sy3 (the padding of U and F keeps it out
// of the other row-dots).
__global__ void foldmax_kernel(const double* __restrict__ R1all, const double* __restrict__ su,
                               const double* __restrict__ fy, int np, double* __restrict__ ysc3) {
  const int chain = blockIdx.x;
  __shared__ double red[256];
  double mx = 0.0;
  for (int k = threadIdx.x; k < np; k += blockDim.x)
    mx = fmax(mx, fabs(fy[k] * R1all[(long)chain * np + k] / su[k]));
  red[threadIdx.x] = mx;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + o]);
    __syncthreads();
  }
  if (threadIdx.x == 0) ysc3[chain] = pow2_scale(red[0]);
}

// YB[c] block (column tile, K step) holds Y1[c][j][k] sy / su[k] split hi/lo
// (and, with ysc3 given, fY R1 sy3 / su in the last column); per-chain
// epilogue column factors (fY R1, R1) in FP32.  One block per (column j,
// chain).
template <typename H>
__global__ void prep_y(const double* __restrict__ Y1all, const double* __restrict__ R1all,
                       const double* __restrict__ su, const double* __restrict__ fy, int np,
                       const double* __restrict__ ysc, const double* __restrict__ ysc3,
                       u16* __restrict__ YBh, u16* __restrict__ YBl, float* __restrict__ epw) {
  const int j = blockIdx.x, chain = blockIdx.y;
  const double* Y = Y1all + ((long)chain * np + j) * np;
  const bool fold = ysc3 != nullptr && j == np - 1;
  const double sy = fold ? ysc3[chain] : ysc[chain];
  const int nct = np / kBR, nkc = np / kBK;
  const long base = ((long)chain * nct + j / kBR) * nkc * kBE;
  for (int k = threadIdx.x; k < np; k += blockDim.x) {
    const double v = (fold ? fy[k] * R1all[(long)chain * np + k] : Y[k]) * sy / su[k];
    const u16 h = from_d<H>(v);
    const long a = base + (long)(k / kBK) * kBE + blk_off(j % kBR, k % kBK);
    YBh[a] = h;
    YBl[a] = from_d<H>(v - (double)to_f<H>(h));
  }
  if (threadIdx.x == 0) {
    const double R = R1all[(long)chain * np + j];
    epw[((long)chain * 2 + 0) * np + j] = (float)(fy[j] * R);
    epw[((long)chain * 2 + 1) * np + j] = (float)R;
  }
}

// chain-independent epilogue column factors: 1/su, 1/sf, fpsi
__global__ void colw_kernel(const double* __restrict__ su, const double* __restrict__ sf,
                            const double* __restrict__ fpsi, int np, float* __restrict__ colw) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= np) return;
  colw[j] = (float)(1.0 / su[j]);
  colw[np + j] = (float)(1.0 / sf[j]);
  colw[2 * np + j] = (float)fpsi[j];
}

// ----------------------------------------------------------- weights GEMM
// V_c = U Y1_c for a tile of 128 images x 128 modes per chain, then per image
// the row-dot products s1 = V.u, s2 = V.(fpsi f), s3 = u.(fY R1), s4 = f.R1
// over the tile's columns (partials per column tile).
// Persistent: the grid is 8 x nper work-groups; XCD x (= blockIdx % 8) owns
// the row tiles rt = x (mod 8), so the column tiles of one row tile run
// side by side on one L2.  Each work-group walks its tiles as one flat
// stream of 16-wide K steps through a ring of 4 LDS slots (separate arrays,
// slot index fixed at compile time by a 4-fold unroll), direct-to-LDS loads
// issued 3 steps ahead.  Wave w owns the 32 columns 32 w .. 32 w + 31 of the
// tile for all 128 images (4 x 1 blocks of 32 x 32 per chain): per step it
// reads 8 image fragments and 2 NC Y1 fragments, one step ahead of its MFMAs.
template <int K>
using ic = std::integral_constant<int, K>;

template <typename H, int NC, bool FOLD>
__global__ __launch_bounds__(256, 1) void weights_split(
    const u16* __restrict__ UAh, const u16* __restrict__ UAl, const float* __restrict__ FA,
    const float* __restrict__ UA32,
    const u16* __restrict__ YBh, const u16* __restrict__ YBl,
    const float* __restrict__ colw, const float* __restrict__ epw, int np, long Mp, int chain0,
    int nrt, double* __restrict__ part) {
  typedef typename Mma<H>::vec vec;
  constexpr int kBlocks = 2 + 2 * NC;           // A hi/lo + B hi/lo per chain
  constexpr int kSlot = kBlocks * kBE;
  __shared__ __align__(16) u16 sw0[kSlot];
  __shared__ __align__(16) u16 sw1[kSlot];
  __shared__ __align__(16) u16 sw2[kSlot];
  __shared__ __align__(16) u16 sw3[kSlot];
  u16* const slots[4] = {sw0, sw1, sw2, sw3};
  // during a tile's epilogue slot 3 (its fragments were read in the tile's
  // last-but-one step; the next copies go to slots 0..2) holds the column
  // factors and the reduction
  constexpr int kCw = (3 + 2 * NC) * kBR;       // floats: 1/su, 1/sf, fpsi, (fY R1, R1) per chain
  static_assert(kCw * 4 + 4 * NC * 64 * 4 * 8 <= kSlot * 2, "epilogue scratch exceeds a slot");
  float* const cw = reinterpret_cast<float*>(sw3);
  double* const red = reinterpret_cast<double*>(sw3 + kCw * 2);   // [wave][chain][64 rows][4]

  const int nct = np / kBR, nkc = np / kBK;     // nkc is a multiple of 8
  const int xcd = blockIdx.x & 7, s = blockIdx.x >> 3, nper = gridDim.x >> 3;
  const int nrt_x = nrt > xcd ? (nrt - xcd + 7) / 8 : 0;
  const int ntile_x = nrt_x * nct;
  const int my_tiles = ntile_x > s ? (ntile_x - s + nper - 1) / nper : 0;
  const int total = my_tiles * nkc;
  if (total == 0) return;

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int lr = lane & 31, lh = lane >> 5;

  // copy step g (clamped to the stream) into slot `slot`: wave w copies
  // bytes [1 KB w, 1 KB (w + 1)) of every block, kBlocks loads per wave
  auto issue = [&](int g, u16* slot) {
    if (g > total - 1) g = total - 1;
    const int ti = g / nkc, kc = g - ti * nkc;
    const int idx = s + ti * nper;
    const int rt = xcd + 8 * (idx / nct), ct = idx % nct;
    u16* dst = slot + wave * 512;
    const long src = (long)wave * 512 + lane * 8;
    const long ab = ((long)rt * nkc + kc) * kBE + src;
    dma16(UAh + ab, dst);
    dma16(UAl + ab, dst + kBE);
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      const long bb = (((long)(chain0 + c) * nct + ct) * nkc + kc) * kBE + src;
      dma16(YBh + bb, dst + (2 + 2 * c) * kBE);
      dma16(YBl + bb, dst + (3 + 2 * c) * kBE);
    }
  };
  struct Frags {
    vec a[4][2];      // image rows 32 i .., hi / lo
    vec b[NC][2];     // Y1 columns 32 wave .., hi / lo
  };
  auto read = [&](const u16* cur, Frags& f) {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int o = frag_off(i, lane);
      f.a[i][0] = *reinterpret_cast<const vec*>(cur + o);
      f.a[i][1] = *reinterpret_cast<const vec*>(cur + kBE + o);
    }
    const int o = frag_off(wave, lane);
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      f.b[c][0] = *reinterpret_cast<const vec*>(cur + (2 + 2 * c) * kBE + o);
      f.b[c][1] = *reinterpret_cast<const vec*>(cur + (3 + 2 * c) * kBE + o);
    }
  };

  v16f acc[NC][4];   // set by the first step of every tile

  // step g (slot K % 4) with its fragments already in `cur`; reads the
  // fragments of step g + 1 into `nxt`.  Every step issues exactly kBlocks
  // loads per wave (the copy of step g + 3), so vmcnt(kBlocks) at its start
  // means the copy of step g + 1 has landed.
  // The first step of a tile starts its accumulators from the zero operand of
  // the MFMA instead of zeroed registers.
  auto step = [&](auto kk, auto prefetch, int g, Frags& cur, Frags& nxt, auto first) {
    constexpr int K = decltype(kk)::value;
    constexpr bool kFirst = decltype(first)::value;
    TDS_WAIT_VMCNT(kBlocks);
    lds_barrier();
    // slot (K + 3) % 4 was last read for step g - 1, during step g - 2 (or
    // at its own start)
    issue(g + 3, slots[(K + 3) & 3]);
    if constexpr (decltype(prefetch)::value) read(slots[(K + 1) & 3], nxt);
    // operands swapped: D = V^T block, lane holds image row 32 i + lane % 32
    // and columns 32 wave + 8 q + 4 (lane / 32) + 0..3 (q = 0..3);
    // product-major order keeps dependent MFMAs 4 NC instructions apart
#pragma unroll
    for (int pr = 0; pr < 3; ++pr)
#pragma unroll
      for (int c = 0; c < NC; ++c)
#pragma unroll
        for (int i = 0; i < 4; ++i)
          acc[c][i] = Mma<H>::mma(cur.b[c][pr == 1 ? 1 : 0], cur.a[i][pr == 2 ? 1 : 0],
                                  (kFirst && pr == 0) ? v16f{} : acc[c][i]);
  };

  Frags f0, f1;
  issue(0, sw0);
  issue(1, sw1);
  issue(2, sw2);
  TDS_WAIT_VMCNT(2 * kBlocks);
  lds_barrier();

  for (int ti = 0; ti < my_tiles; ++ti) {
    // the first step's copy was certified by the last barrier; its fragments
    // are read here, every later step's one step ahead (none across the
    // epilogue)
    const int gt = ti * nkc;
    read(sw0, f0);
    constexpr std::true_type yes{};
    constexpr std::false_type no{};
    step(ic<0>{}, yes, gt, f0, f1, yes);
    step(ic<1>{}, yes, gt + 1, f1, f0, no);
    step(ic<2>{}, yes, gt + 2, f0, f1, no);
    step(ic<3>{}, yes, gt + 3, f1, f0, no);
    for (int kc = 4; kc < nkc - 4; kc += 4) {
      step(ic<0>{}, yes, gt + kc, f0, f1, no);
      step(ic<1>{}, yes, gt + kc + 1, f1, f0, no);
      step(ic<2>{}, yes, gt + kc + 2, f0, f1, no);
      step(ic<3>{}, yes, gt + kc + 3, f1, f0, no);
    }
    step(ic<0>{}, yes, gt + nkc - 4, f0, f1, no);
    step(ic<1>{}, yes, gt + nkc - 3, f1, f0, no);
    step(ic<2>{}, yes, gt + nkc - 2, f0, f1, no);
    step(ic<3>{}, no, gt + nkc - 1, f1, f0, no);

    // ---- epilogue of tile (rt, ct).  Per image row and chain, with u, f the
    // unscaled images and V the accumulators (sy times the true V):
    //   p0 = sum V u, p1 = sum V fpsi f, p2 = sum u fY R1, p3 = sum f R1
    // over the tile's 128 columns; FP32 over 4 columns, FP64 beyond; halves
    // of a wave combined by a shuffle, the four column waves through LDS.
    const int idx = s + ti * nper;
    const int rt = xcd + 8 * (idx / nct), ct = idx % nct;
    const long r0 = (long)rt * kBR;
    const int c0 = ct * kBR;
    for (int e = tid; e < kCw; e += 256) {
      const int w = e / kBR, cc = c0 + e % kBR;
      cw[e] = w < 3 ? colw[w * np + cc] : epw[((long)chain0 * 2 + (w - 3)) * np + cc];
    }
    lds_barrier();
    // two halves of 64 rows: every wave stores its column-quarter partials,
    // then the four quarters are summed in a fixed order.  The image loads of
    // row group i + 1 (16 loads of 8 bytes per lane) are issued before row
    // group i is multiplied, so their latency overlaps the arithmetic.
    struct Rows {
      float4 u[4];      // U of the lane's 4 columns, per column quarter (FP32)
      float4 f[4];      // F (FP32)
    };
    Rows ldb[2];
    auto load_rows = [&](int i, Rows& ld) {
      const int row = 32 * i + lr;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int col = 32 * wave + 8 * q + 4 * lh;
        const long a = ((long)rt * nkc + (c0 + col) / kBK) * kBE + blk_off(row, col % kBK);
        ld.u[q] = *reinterpret_cast<const float4*>(UA32 + a);
        ld.f[q] = *reinterpret_cast<const float4*>(FA + a);
      }
    };
    load_rows(0, ldb[0]);
#pragma unroll
    for (int half = 0; half < 2; ++half) {
#pragma unroll
      for (int ii = 0; ii < 2; ++ii) {
        const int i = 2 * half + ii;
        // FP32 over the lane's 16 columns, FP64 beyond
        float p32[NC][4];
#pragma unroll
        for (int c = 0; c < NC; ++c) p32[c][0] = p32[c][1] = p32[c][2] = p32[c][3] = 0.f;
        if (i < 3) load_rows(i + 1, ldb[(i + 1) & 1]);
        const Rows& ld = ldb[i & 1];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int col = 32 * wave + 8 * q + 4 * lh;   // tile column of t = 0
          const float4 fps = *reinterpret_cast<const float4*>(cw + 2 * kBR + col);
          const float fp4[4] = {fps.x, fps.y, fps.z, fps.w};
          const float f[4] = {ld.f[q].x, ld.f[q].y, ld.f[q].z, ld.f[q].w};
          const float u[4] = {ld.u[q].x, ld.u[q].y, ld.u[q].z, ld.u[q].w};
          float g[4];
#pragma unroll
          for (int t = 0; t < 4; ++t) g[t] = f[t] * fp4[t];
#pragma unroll
          for (int c = 0; c < NC; ++c) {
            const float4 rr = *reinterpret_cast<const float4*>(cw + (4 + 2 * c) * kBR + col);
            const float rr4[4] = {rr.x, rr.y, rr.z, rr.w};
            float fr4[4] = {0.f, 0.f, 0.f, 0.f};
            if constexpr (!FOLD) {
              const float4 fr = *reinterpret_cast<const float4*>(cw + (3 + 2 * c) * kBR + col);
              fr4[0] = fr.x; fr4[1] = fr.y; fr4[2] = fr.z; fr4[3] = fr.w;
            }
#pragma unroll
            for (int t = 0; t < 4; ++t) {
              const float v = acc[c][i][4 * q + t];
              p32[c][0] += v * u[t];
              p32[c][1] += v * g[t];
              if constexpr (!FOLD) p32[c][2] += u[t] * fr4[t];
              p32[c][3] += f[t] * rr4[t];
            }
          }
        }
        if constexpr (FOLD) {
          // s3 sy3 of the row: the folded column, tile column 127 of the last
          // column tile (wave 3, q = 3, lane half 1, t = 3)
          const bool holder = ct == nct - 1 && wave == 3 && lh == 1;
#pragma unroll
          for (int c = 0; c < NC; ++c) p32[c][2] = holder ? acc[c][i][15] : 0.f;
        }
        double p[NC][4];
#pragma unroll
        for (int c = 0; c < NC; ++c)
#pragma unroll
          for (int qq = 0; qq < 4; ++qq) {
            p[c][qq] = (double)p32[c][qq];
            p[c][qq] += __shfl_xor(p[c][qq], 32);
          }
        if (lh == 0) {
#pragma unroll
          for (int c = 0; c < NC; ++c)
#pragma unroll
            for (int qq = 0; qq < 4; ++qq)
              red[((wave * NC + c) * 64 + 32 * ii + lr) * 4 + qq] = p[c][qq];
        }
      }
      lds_barrier();
      for (int e2 = tid; e2 < NC * 64 * 4; e2 += 256) {
        const int c = e2 / 256, rq = e2 % 256;   // (row, quantity) of this half
        double v = red[(0 * NC + c) * 256 + rq];
        v += red[(1 * NC + c) * 256 + rq];
        v += red[(2 * NC + c) * 256 + rq];
        v += red[(3 * NC + c) * 256 + rq];
        part[(((long)(chain0 + c) * nct + ct) * Mp + r0 + 64 * half) * 4 + rq] = v;
      }
      lds_barrier();
    }
  }
  // the copies issued past the end of the stream land before the work-group
  // releases its LDS
  TDS_WAIT_VMCNT(0);
}

// coefficients from the column-tile partials (fixed order); |c| maxima.
// coefB layout: [Mp/16][kMaxChains][4][16] floats (c1, c2, alpha, beta of
// the 16 images of a K step, all chains of a step contiguous).
__global__ void finalize_coef(const double* __restrict__ part, const double* __restrict__ rho_img,
                              const double* __restrict__ ysc, const double* __restrict__ ysc3,
                              long Mp, int nct, int apply_d4,
                              long row_lo, long row_hi, double* __restrict__ coef_all,
                              float* __restrict__ coefB, double* __restrict__ pmax) {
  const int chain = blockIdx.y;
  const double isy = 1.0 / ysc[chain];   // the V-weighted sums carry the Y1 scale
  double m1 = 0.0, m2 = 0.0;
  for (long row = (long)blockIdx.x * blockDim.x + threadIdx.x; row < Mp;
       row += (long)gridDim.x * blockDim.x) {
    double s[4] = {0, 0, 0, 0};
    for (int ct = 0; ct < nct; ++ct) {
      const double* p = part + (((long)chain * nct + ct) * Mp + row) * 4;
      s[0] += p[0]; s[1] += p[1]; s[2] += p[2]; s[3] += p[3];
    }
    s[0] *= isy;
    s[1] *= isy;
    if (ysc3) s[2] /= ysc3[chain];   // the folded column carries its own scale
    // images outside [row_lo, row_hi) carry no weight (image-range calls)
    const double rho = (row >= row_lo && row < row_hi) ? rho_img[row] : 0.0;
    const double c1 = rho * (-s[2] / 3.0 + apply_d4 * s[0] / 8.0);
    const double c2 = rho * (-s[3] / 3.0 + apply_d4 * s[1] / 4.0);
    const double al = -rho * s[0] / 6.0, be = -rho * s[1] / 3.0;
    if (coef_all) {
      double* c = coef_all + ((long)chain * Mp + row) * 4;
      c[0] = c1; c[1] = c2; c[2] = al; c[3] = be;
    }
    float* cb = coefB + (((row / kBK) * kMaxChains + chain) * 4) * kBK + row % kBK;
    cb[0] = (float)c1; cb[kBK] = (float)c2; cb[2 * kBK] = (float)al; cb[3 * kBK] = (float)be;
    m1 = fmax(m1, fabs(c1));
    m2 = fmax(m2, fabs(c2));
  }
  __shared__ double r1[256], r2[256];
  r1[threadIdx.x] = m1;
  r2[threadIdx.x] = m2;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) {
      r1[threadIdx.x] = fmax(r1[threadIdx.x], r1[threadIdx.x + o]);
      r2[threadIdx.x] = fmax(r2[threadIdx.x], r2[threadIdx.x + o]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    double* pm = pmax + ((long)chain * kPartBlocks + blockIdx.x) * 2;
    pm[0] = r1[0];
    pm[1] = r2[0];
  }
}

// One power-of-two B scale per chain: sb with max_nu (max|c1| fmx[nu] +
// max|c2| fY[nu] umx[nu] / 2) * sb < 2^12, written for every nu (the
// reduction undoes it per column).  One block per chain.
__global__ void bscale_kernel(const double* __restrict__ pmax, int nblocks,
                              const double* __restrict__ umx, const double* __restrict__ fmx,
                              const double* __restrict__ fy, int np, double* __restrict__ cmax,
                              double* __restrict__ sb_all) {
  const int chain = blockIdx.x;
  __shared__ double red[256], red2[256];
  // max |c1|, max |c2| over the coefficient blocks
  double m1 = 0.0, m2 = 0.0;
  for (int bk = threadIdx.x; bk < nblocks; bk += blockDim.x) {
    const double* pm = pmax + ((long)chain * kPartBlocks + bk) * 2;
    m1 = fmax(m1, pm[0]);
    m2 = fmax(m2, pm[1]);
  }
  red[threadIdx.x] = m1;
  red2[threadIdx.x] = m2;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) {
      red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + o]);
      red2[threadIdx.x] = fmax(red2[threadIdx.x], red2[threadIdx.x + o]);
    }
    __syncthreads();
  }
  const double a = red[0], b = 0.5 * red2[0];
  __syncthreads();
  if (threadIdx.x == 0) {
    cmax[chain * 2 + 0] = a;
    cmax[chain * 2 + 1] = red2[0];
  }
  double mx = 0.0;
  for (int nu = threadIdx.x; nu < np; nu += blockDim.x)
    mx = fmax(mx, a * fmx[nu] + b * fy[nu] * umx[nu]);
  red[threadIdx.x] = mx;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + o]);
    __syncthreads();
  }
  const double sb = pow2_scale(red[0]);
  for (int nu = threadIdx.x; nu < np; nu += blockDim.x) sb_all[(long)chain * np + nu] = sb;
}

// c1 and c2 of the coefficient blocks times the chain's B scale (exact)
__global__ void scale_coef_kernel(const double* __restrict__ sb_all, int np, long nsub,
                                  float* __restrict__ coefB) {
  const int chain = blockIdx.y;
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;   // (step, image) pair
  if (e >= nsub * kBK) return;
  const float sb = (float)sb_all[(long)chain * np];
  float* cb = coefB + ((e / kBK) * kMaxChains + chain) * 4 * kBK + e % kBK;
  cb[0] *= sb;
  cb[kBK] *= sb;
}

// ------------------------------------------------------------- outer GEMM
// Z_c[mu, nu] = sum_k U[k, mu] B_c[k, nu], B_c = c1 F + (c2/2) fY U, over a
// range of K steps (16 images each, a multiple of 4 steps per range); tile
// 128 (mu) x 128 (nu).  The A blocks (U^T of the mu tile), the FP32 blocks of
// F and fY U / 2 of the nu tile and the step's coefficients arrive by direct-to-LDS
// loads three steps ahead in rings of 4 slots (separate arrays, slot index
// fixed at compile time by a 4-fold unroll).  B_c of step t + 1 is built in
// FP32 from the raw blocks while the MFMAs of step t run, split hi/lo again
// and stored to the other of two B slots.  8 waves, two per SIMD, so one
// wave's build overlaps the other's MFMAs: wave (wr, wc) owns mu 64 wr ..
// 64 wr + 63 and nu 32 wc .. 32 wc + 31 (2 x 1 blocks of 32 x 32 per chain).
// The raw FP32 accumulators go to an FP32 partial buffer (exact); the scales
// are undone in the FP64 reduction.
template <typename H, int NC>
__global__ __launch_bounds__(512, 1) void outer_split(
    const u16* __restrict__ UTh, const u16* __restrict__ UTl, const float* __restrict__ XT,
    const float* __restrict__ YT, const float* __restrict__ coefB, int np, long Mp, int ksplit,
    int ntile, int chain0, float* __restrict__ zpart_all, double* __restrict__ fpart_all) {
  typedef typename Mma<H>::vec vec;
  typedef u16 u16x4v __attribute__((ext_vector_type(4)));
  constexpr int kA = 2 * kBE;                   // A hi/lo
  constexpr int kR = 2 * kBE;                   // floats: F, fY U / 2
  constexpr int kB = 2 * NC * kBE;              // B hi/lo per chain
  constexpr int kCoef = NC * 4 * kBK;           // floats of one step's coefficients
  __shared__ __align__(16) u16 sa0[kA];
  __shared__ __align__(16) u16 sa1[kA];
  __shared__ __align__(16) u16 sa2[kA];
  __shared__ __align__(16) u16 sa3[kA];
  __shared__ __align__(16) float sr0[kR];
  __shared__ __align__(16) float sr1[kR];
  __shared__ __align__(16) float sr2[kR];
  __shared__ __align__(16) float sr3[kR];
  __shared__ __align__(16) u16 sb0[kB];
  __shared__ __align__(16) u16 sb1[kB];
  __shared__ __align__(16) float sc0[kCoef];
  __shared__ __align__(16) float sc1[kCoef];
  __shared__ __align__(16) float sc2[kCoef];
  __shared__ __align__(16) float sc3[kCoef];
  u16* const sA[4] = {sa0, sa1, sa2, sa3};
  float* const sR[4] = {sr0, sr1, sr2, sr3};
  u16* const sB[2] = {sb0, sb1};
  float* const sC[4] = {sc0, sc1, sc2, sc3};

  const int id = blockIdx.x;
  const int xcd = id & 7, slotid = id >> 3;
  const int t2 = ntile * ntile;
  const int ks = xcd + 8 * (slotid / t2);
  const int tile = slotid % t2;
  if (ks >= ksplit) return;
  const int tm = tile / ntile, tn = tile % ntile;
  const int mu0 = tm * kBR, nu0 = tn * kBR;
  const bool diag = (tm == tn);
  // wave in a scalar register: the copy sources below are then chosen
  // without branches
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int lr = lane & 31, lh = lane >> 5;
  const int wr = wave >> 2, wc = wave & 3;

  // K steps [g0, g0 + n) in units of 4 steps
  const long nsub = Mp / kBK, nunit = nsub / 4;
  const long g0 = 4 * ((long)ks * nunit / ksplit);
  const long n = 4 * ((long)(ks + 1) * nunit / ksplit) - g0;

  // builder role: thread t owns images 4 bq .. 4 bq + 3 of row nu0 + brow of
  // every B block: element (bq / 2) * 1024 + row * 8 + 4 (bq % 2) of the
  // chunk-major block.  The two halves of a row's 8-image chunk sit on
  // adjacent lanes (bq % 2 = t % 2, row = (t / 2) % 128, chunk bq / 2 = t / 256),
  // so a wave's reads and stores cover contiguous runs without bank
  // conflicts.
  const int brow = (tid >> 1) & (kBR - 1);
  const int bq = (tid & 1) | ((tid >> 8) << 1);
  const int boff = (bq >> 1) * (kBR * 8) + brow * 8 + 4 * (bq & 1);

  const long a_tile = (long)tm * nsub * kBE;
  const long b_tile = (long)tn * nsub * kBE;
  // a step's A blocks (2 x 4 KB) and FP32 blocks (2 x 8 KB) in 24 pieces of
  // 1 KB, 3 per wave, and a 32 NC-byte share of its coefficients: 4 loads per
  // wave
  const u16* const UTw = wave < 4 ? UTh : UTl;
  auto issue = [&](long ts, int p) {
    if (ts > n - 1) ts = n - 1;
    const long offa = a_tile + (g0 + ts) * kBE;
    const long offb = b_tile + (g0 + ts) * kBE;
    const int part = (wave & 3) * 512;
    dma16(UTw + offa + part + lane * 8, sA[p] + (wave >> 2) * kBE + part);
    dma16(XT + offb + wave * 256 + lane * 4, sR[p] + wave * 256);
    dma16(YT + offb + wave * 256 + lane * 4, sR[p] + kBE + wave * 256);
    const float* cs = coefB + ((g0 + ts) * kMaxChains + chain0) * 4 * kBK + wave * NC * 8 + lane * 4;
    if (lane < 2 * NC) dma16(cs, sC[p] + wave * NC * 8);
  };

  double fl_acc[NC];
#pragma unroll
  for (int c = 0; c < NC; ++c) fl_acc[c] = 0.0;
  // B of one step from its raw blocks and coefficients; on diagonal tiles
  // also the f partial sums (weighted by `live`, 0 for a clamped step)
  auto build = [&](auto is_diag, const float* rs, const float* cs, u16* b_slot, double live) {
    constexpr bool DIAG = decltype(is_diag)::value;
    // xf = F, yu = fY U / 2, shared by the chains
    const float4 x4 = *reinterpret_cast<const float4*>(rs + boff);
    const float4 y4 = *reinterpret_cast<const float4*>(rs + kBE + boff);
    const float xf[4] = {x4.x, x4.y, x4.z, x4.w};
    const float yu[4] = {y4.x, y4.y, y4.z, y4.w};
    const float* cf = cs + 4 * bq;   // two addresses per wave: broadcast reads
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      const float4 c1 = *reinterpret_cast<const float4*>(cf + (c * 4 + 0) * kBK);
      const float4 c2 = *reinterpret_cast<const float4*>(cf + (c * 4 + 1) * kBK);
      const float c1v[4] = {c1.x, c1.y, c1.z, c1.w};
      const float c2v[4] = {c2.x, c2.y, c2.z, c2.w};
      u16x4v bh, bl;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const float b = __builtin_fmaf(c1v[e], xf[e], c2v[e] * yu[e]);
        const u16 hb = from_f<H>(b);
        bh[e] = hb;
        // b - hi is exact in FP32; one mixed-precision FMA where the type has it
        bl[e] = from_f<H>(__builtin_fmaf(to_f<H>(hb), -1.0f, b));
      }
      *reinterpret_cast<u16x4v*>(b_slot + (2 * c) * kBE + boff) = bh;
      *reinterpret_cast<u16x4v*>(b_slot + (2 * c + 1) * kBE + boff) = bl;
      if constexpr (DIAG) {
        const float4 aa = *reinterpret_cast<const float4*>(cf + (c * 4 + 2) * kBK);
        const float4 bb = *reinterpret_cast<const float4*>(cf + (c * 4 + 3) * kBK);
        const float av[4] = {aa.x, aa.y, aa.z, aa.w};
        const float bv[4] = {bb.x, bb.y, bb.z, bb.w};
        float fsum = 0.f;
#pragma unroll
        for (int e = 0; e < 4; ++e) fsum += av[e] * xf[e] + bv[e] * (2.0f * yu[e]);
        fl_acc[c] += live * (double)fsum;
      }
    }
  };

  v16f acc[NC][2];
#pragma unroll
  for (int c = 0; c < NC; ++c) acc[c][0] = acc[c][1] = v16f{};

  typedef vec AFrag[2][2];
  auto read_a = [&](const u16* ca, AFrag& fa) {
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      const int o = frag_off(2 * wr + i, lane);
      fa[i][0] = *reinterpret_cast<const vec*>(ca + o);
      fa[i][1] = *reinterpret_cast<const vec*>(ca + kBE + o);
    }
  };
  // one K step t from ring position K % 4 (B slot K % 2).  Each step issues
  // exactly 4 loads per wave (step t + 3, clamped), so vmcnt(4) at its end
  // means everything issued before it has landed: A of t + 1 and the raw
  // blocks and coefficients of t + 2.
  auto step = [&](auto is_diag, auto kk, long t) {
    constexpr int K = decltype(kk)::value;
    lds_barrier();   // B of step t stored; ring position (K + 3) % 4 free
    issue(t + 3, (K + 3) & 3);
    const u16* cb = sB[K & 1];
    vec fb[NC][2];
    const int ob = frag_off(wc, lane);
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      fb[c][0] = *reinterpret_cast<const vec*>(cb + (2 * c) * kBE + ob);
      fb[c][1] = *reinterpret_cast<const vec*>(cb + (2 * c + 1) * kBE + ob);
    }
    AFrag cur;
    read_a(sA[K], cur);
#pragma unroll
    for (int pr = 0; pr < 3; ++pr)
#pragma unroll
      for (int c = 0; c < NC; ++c)
#pragma unroll
        for (int i = 0; i < 2; ++i)
          acc[c][i] = Mma<H>::mma(cur[i][pr == 2 ? 1 : 0], fb[c][pr == 1 ? 1 : 0], acc[c][i]);
    // B of step t + 1 (clamped at the end of the range) into the other B slot
    build(is_diag, sR[(K + 1) & 3], sC[(K + 1) & 3], sB[(K + 1) & 1], t + 1 < n ? 1.0 : 0.0);
    TDS_WAIT_VMCNT(4);
  };

  auto run = [&](auto is_diag) {
    // prologue: steps 0..2 requested; B of step 0 built
    issue(0, 0);
    issue(1, 1);
    issue(2, 2);
    TDS_WAIT_VMCNT(0);
    lds_barrier();
    build(is_diag, sr0, sc0, sb0, 1.0);
    for (long t = 0; t < n; t += 4) {
      step(is_diag, ic<0>{}, t);
      step(is_diag, ic<1>{}, t + 1);
      step(is_diag, ic<2>{}, t + 2);
      step(is_diag, ic<3>{}, t + 3);
    }
  };
  if (n > 0) {
    if (diag)
      run(std::true_type{});
    else
      run(std::false_type{});
  }
  TDS_WAIT_VMCNT(0);
  lds_barrier();

  // raw partial tiles (FP32, exact copies of the accumulators)
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    float* zp = zpart_all + ((long)(chain0 + c) * ksplit + ks) * np * np;
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int mm = 64 * wr + 32 * i + 8 * (r >> 2) + 4 * lh + (r & 3);
        const int nn = 32 * wc + lr;
        zp[(long)(mu0 + mm) * np + nu0 + nn] = acc[c][i][r];
      }
  }
  if (diag) {
    double* red = reinterpret_cast<double*>(sb0);  // [NC][4 quarters][128]
#pragma unroll
    for (int c = 0; c < NC; ++c) red[(c * 4 + bq) * kBR + brow] = fl_acc[c];
    lds_barrier();
    for (int e = tid; e < NC * kBR; e += 512) {
      const int c = e / kBR, nn = e % kBR;
      const double* q = red + c * 4 * kBR + nn;
      fpart_all[((long)(chain0 + c) * ksplit + ks) * np + nu0 + nn] =
          (q[0] + q[kBR]) + (q[2 * kBR] + q[3 * kBR]);
    }
  }
}

// FP64 sum of the FP32 partial tiles (fixed order), scales undone:
// zsum[mu, nu] = sum_k zpart[k, mu, nu] / (su[mu] sb[nu])
__global__ void reduce_ksplit_s32(const float* __restrict__ zpart_all,
                                  const double* __restrict__ fpart_all,
                                  const double* __restrict__ su, const double* __restrict__ sb_all,
                                  int np, int ksplit, double* __restrict__ zsum_all,
                                  double* __restrict__ f_all, double norm) {
  const int chain = blockIdx.y;
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long nn = (long)np * np;
  const float* zp = zpart_all + (long)chain * ksplit * nn;
  if (e < nn) {
    // eight loads in flight, summed in the same order as one at a time
    double s = 0.0;
    int k = 0;
    for (; k + 8 <= ksplit; k += 8) {
      float v[8];
#pragma unroll
      for (int j = 0; j < 8; ++j) v[j] = zp[(long)(k + j) * nn + e];
#pragma unroll
      for (int j = 0; j < 8; ++j) s += (double)v[j];
    }
    for (; k < ksplit; ++k) s += (double)zp[(long)k * nn + e];
    const int mu = (int)(e / np), nu = (int)(e % np);
    zsum_all[(long)chain * nn + e] = s / (su[mu] * sb_all[(long)chain * np + nu]);
  }
  if (e < np) {
    const double* fp = fpart_all + (long)chain * ksplit * np;
    double s = 0.0;
    for (int k = 0; k < ksplit; ++k) s += fp[(long)k * np + e];
    f_all[(long)chain * np + e] = s * norm;
  }
}

// ---------------------------------------------------------------- host side
// K split of the outer GEMM: ntile^2 * ks work-groups (ks a multiple of 8, one
// range per XCD in turn) filling whole waves of 256 CUs as closely as
// possible, the smallest such ks, with at least 32 K steps per range.
static int choose_ksplit_split(int ntile, long Mp) {
  const long nsub = Mp / kBK;   // ranges are whole units of 4 steps
  const int t2 = ntile * ntile;
  int best = 8;
  double best_eff = -1.0;
  for (int ks = 8; ks <= 512; ks += 8) {
    if (ks > 8 && nsub / ks < 32) break;
    const long wgs = (long)t2 * ks;
    const long rounds = (wgs + 255) / 256;
    const double eff = (double)wgs / (double)(rounds * 256);
    if (eff > best_eff + 0.02) {
      best_eff = eff;
      best = ks;
    }
  }
  return best;
}

int ensure_split(Ctx* c, int prec) {
  static_assert(kMaxChains == 8, "coefficient block layout assumes 8 chains");
  SplitStore& st = c->split[prec];
  if (st.ready) return 0;
  const long Mp = c->Mp;
  const int np = c->np;
  TDS_REQUIRE(Mp % kBR == 0 && np % kBR == 0, 3, "split path: padded sizes must be multiples of 128");
  const size_t half_bytes = sizeof(u16) * (size_t)Mp * np;
  void** arr[] = {&st.Uh, &st.Ul, &st.UhT, &st.UlT};
  for (void** p : arr) TDS_CHECK_HIP(hipMalloc(p, half_bytes));
  TDS_CHECK_HIP(hipMalloc((void**)&st.FA, 2 * half_bytes));
  TDS_CHECK_HIP(hipMalloc((void**)&st.UA32, 2 * half_bytes));
  TDS_CHECK_HIP(hipMalloc((void**)&st.XT, 2 * half_bytes));
  TDS_CHECK_HIP(hipMalloc((void**)&st.YT, 2 * half_bytes));
  TDS_CHECK_HIP(hipMalloc((void**)&st.su, sizeof(double) * np));
  TDS_CHECK_HIP(hipMalloc((void**)&st.sf, sizeof(double) * np));
  TDS_CHECK_HIP(hipMalloc((void**)&st.umax, sizeof(double) * np));
  TDS_CHECK_HIP(hipMalloc((void**)&st.fmax, sizeof(double) * np));
  TDS_CHECK_HIP(hipMalloc((void**)&st.colw, sizeof(float) * 3 * np));
  if (!c->Y1h) {
    TDS_CHECK_HIP(hipMalloc(&c->Y1h, sizeof(u16) * kMaxChains * (size_t)np * np));
    TDS_CHECK_HIP(hipMalloc(&c->Y1l, sizeof(u16) * kMaxChains * (size_t)np * np));
    TDS_CHECK_HIP(hipMalloc((void**)&c->epi, sizeof(float) * kMaxChains * 2 * np));
    TDS_CHECK_HIP(hipMalloc((void**)&c->ysc, sizeof(double) * kMaxChains));
    TDS_CHECK_HIP(hipMalloc((void**)&c->ysc3, sizeof(double) * kMaxChains));
    TDS_CHECK_HIP(hipMalloc((void**)&c->cmax, sizeof(double) * kMaxChains * 2));
    TDS_CHECK_HIP(hipMalloc((void**)&c->pmax, sizeof(double) * kMaxChains * kPartBlocks * 2));
    TDS_CHECK_HIP(hipMalloc((void**)&c->coef32, sizeof(float) * kMaxChains * Mp * 4));
    TDS_CHECK_HIP(hipMalloc((void**)&c->wpart, sizeof(double) * kMaxChains * (np / kBR) * Mp * 4));
    const int ks = choose_ksplit_split(np / kBR, Mp);
    c->ksplit_s = ks;
    TDS_CHECK_HIP(hipMalloc((void**)&c->zpart32, sizeof(float) * kMaxChains * (size_t)ks * np * np));
    TDS_CHECK_HIP(hipMalloc((void**)&c->fpart_s, sizeof(double) * kMaxChains * (size_t)ks * np));
    TDS_CHECK_HIP(hipMalloc((void**)&c->sbsc, sizeof(double) * kMaxChains * np));
  }
  hipStream_t s = c->stream;
  TDS_CHECK_HIP(hipMemsetAsync(st.umax, 0, sizeof(double) * np, s));
  TDS_CHECK_HIP(hipMemsetAsync(st.fmax, 0, sizeof(double) * np, s));
  hipLaunchKernelGGL(colmax_kernel, dim3((np + 127) / 128, 256), dim3(128), 0, s, c->U, c->F, Mp,
                     np, st.umax, st.fmax);
  hipLaunchKernelGGL(scales_kernel, dim3((np + 127) / 128), dim3(128), 0, s, st.umax, st.fmax, np,
                     st.su, st.sf);
  hipLaunchKernelGGL(colw_kernel, dim3((np + 127) / 128), dim3(128), 0, s, st.su, st.sf, c->fpsi,
                     np, st.colw);
  dim3 g((unsigned)(Mp / 32), np / 32);
  if (prec == PREC_FP16X2)
    hipLaunchKernelGGL(split_images<_Float16>, g, dim3(256), 0, s, c->U, c->F, Mp, np, st.su,
                       st.sf, c->fy, (u16*)st.Uh, (u16*)st.Ul, st.FA, st.UA32,
                       (u16*)st.UhT, (u16*)st.UlT, st.XT, st.YT);
  else
    hipLaunchKernelGGL(split_images<__bf16>, g, dim3(256), 0, s, c->U, c->F, Mp, np, st.su,
                       st.sf, c->fy, (u16*)st.Uh, (u16*)st.Ul, st.FA, st.UA32,
                       (u16*)st.UhT, (u16*)st.UlT, st.XT, st.YT);
  TDS_CHECK_HIP(hipGetLastError());
  TDS_CHECK_HIP(hipStreamSynchronize(s));
  st.ready = true;
  return 0;
}

template <typename H, int NC, bool FOLD>
static void launch_weights(Ctx* c, const SplitStore& st, int chain0, hipStream_t s) {
  const int np = c->np;
  const int nct = np / kBR;
  const int nrt = (int)(c->Mp / kBR);
  // persistent: at most 32 work-groups per XCD, none without a tile
  const int per_xcd_tiles = (nrt + 7) / 8 * nct;
  const int nper = per_xcd_tiles < 32 ? per_xcd_tiles : 32;
  hipLaunchKernelGGL((weights_split<H, NC, FOLD>), dim3((unsigned)(8 * nper)), dim3(256), 0, s,
                     (const u16*)st.Uh, (const u16*)st.Ul, st.FA, st.UA32,
                     (const u16*)c->Y1h, (const u16*)c->Y1l, st.colw, c->epi, np, c->Mp, chain0,
                     nrt, c->wpart);
}

template <typename H, int NC>
static void launch_outer(Ctx* c, const SplitStore& st, int chain0, hipStream_t s) {
  const int ntile = c->np / kBR;
  hipLaunchKernelGGL((outer_split<H, NC>), dim3((unsigned)(ntile * ntile * c->ksplit_s)),
                     dim3(512), 0, s, (const u16*)st.UhT, (const u16*)st.UlT, st.XT, st.YT,
                     c->coef32, c->np, c->Mp,
                     c->ksplit_s, ntile, chain0, c->zpart32, c->fpart_s);
}

template <typename H>
static int launch_all(Ctx* c, int nc, int apply_d4, hipStream_t s, hipEvent_t* ev) {
  const SplitStore& st = c->split[std::is_same<H, _Float16>::value ? PREC_FP16X2 : PREC_BF16X2];
  const int np = c->np;
  const long Mp = c->Mp;
  if (ev) (void)hipEventRecord(ev[0], s);
  hipLaunchKernelGGL(ymax_part_kernel, dim3(kYmaxBlocks, nc), dim3(256), 0, s, c->Y1, st.su, np,
                     c->pmax);
  hipLaunchKernelGGL(ymax_kernel, dim3(nc), dim3(64), 0, s, c->pmax, kYmaxBlocks, c->ysc);
  // with padded modes the row-dot s3 rides the fp16 GEMM in the last padding
  // column (hi/lo products good to ~2^-22, like the FP32 sums they replace);
  // bf16 hi/lo carries ~2^-17, and s3 dominates c1, so the bf16 path keeps
  // the FP32 row-dot (its spectrum error would grow fivefold)
  const bool fold = c->n < np && std::is_same<H, _Float16>::value;
  if (fold)
    hipLaunchKernelGGL(foldmax_kernel, dim3(nc), dim3(256), 0, s, c->R1, st.su, c->fy, np,
                       c->ysc3);
  hipLaunchKernelGGL(prep_y<H>, dim3(np, nc), dim3(256), 0, s, c->Y1, c->R1, st.su, c->fy, np,
                     c->ysc, fold ? c->ysc3 : nullptr, (u16*)c->Y1h, (u16*)c->Y1l, c->epi);
  // chains in groups of up to 3 per pass
  for (int c0 = 0; c0 < nc;) {
    const int g = nc - c0 >= 3 ? 3 : nc - c0;
    if (fold) {
      switch (g) {
        case 3: launch_weights<H, 3, true>(c, st, c0, s); break;
        case 2: launch_weights<H, 2, true>(c, st, c0, s); break;
        default: launch_weights<H, 1, true>(c, st, c0, s); break;
      }
    } else {
      switch (g) {
        case 3: launch_weights<H, 3, false>(c, st, c0, s); break;
        case 2: launch_weights<H, 2, false>(c, st, c0, s); break;
        default: launch_weights<H, 1, false>(c, st, c0, s); break;
      }
    }
    c0 += g;
  }
  const int nct = np / kBR;
  const long fin_blocks_l = (Mp + 255) / 256;
  const int fin_blocks = fin_blocks_l < kPartBlocks ? (int)fin_blocks_l : kPartBlocks;
  hipLaunchKernelGGL(finalize_coef, dim3((unsigned)fin_blocks, nc), dim3(256), 0, s, c->wpart,
                     c->rho_img, c->ysc, fold ? c->ysc3 : nullptr, Mp, nct, apply_d4, c->row_lo,
                     c->row_hi,
                     (double*)nullptr, c->coef32, c->pmax);
  hipLaunchKernelGGL(bscale_kernel, dim3(nc), dim3(256), 0, s, c->pmax, fin_blocks, st.umax,
                     st.fmax, c->fy, np, c->cmax, c->sbsc);
  hipLaunchKernelGGL(scale_coef_kernel, dim3((unsigned)((Mp + 255) / 256), nc), dim3(256), 0, s,
                     c->sbsc, np, Mp / kBK, c->coef32);
  if (ev) (void)hipEventRecord(ev[1], s);
  for (int c0 = 0; c0 < nc;) {
    const int g = nc - c0 >= 3 ? 3 : nc - c0;
    switch (g) {
      case 3: launch_outer<H, 3>(c, st, c0, s); break;
      case 2: launch_outer<H, 2>(c, st, c0, s); break;
      default: launch_outer<H, 1>(c, st, c0, s); break;
    }
    c0 += g;
  }
  if (ev) (void)hipEventRecord(ev[2], s);
  const long nn = (long)np * np;
  const double norm = 1.0 / (c->n_eff * c->n_syms);
  hipLaunchKernelGGL(reduce_ksplit_s32, dim3((unsigned)((nn + 255) / 256), nc), dim3(256), 0, s,
                     c->zpart32, c->fpart_s, st.su, c->sbsc, np, c->ksplit_s, c->d2v_out,
                     c->f_out, norm);
  if (ev) (void)hipEventRecord(ev[3], s);
  TDS_CHECK_HIP(hipGetLastError());
  return 0;
}

// K split of the split-precision outer GEMM (work-group K ranges are whole
// units of 4 steps); re-allocates the partial tiles
int set_ksplit_split(Ctx* c, int ks) {
  TDS_REQUIRE(c->coef32 != nullptr, 2, "set_ksplit_split: split path not prepared");
  TDS_REQUIRE(ks >= 8 && ks % 8 == 0 && (long)ks * 4 <= c->Mp / kBK, 2,
              "set_ksplit_split: K split must be a multiple of 8 with at least 4 steps per range");
  TDS_CHECK_HIP(hipStreamSynchronize(c->stream));
  float* zp = nullptr;
  double* fp = nullptr;
  TDS_CHECK_HIP(hipMalloc((void**)&zp, sizeof(float) * kMaxChains * (size_t)ks * c->np * c->np));
  if (hipMalloc((void**)&fp, sizeof(double) * kMaxChains * (size_t)ks * c->np) != hipSuccess) {
    (void)hipFree(zp);
    set_error("set_ksplit_split: hipMalloc failed");
    return -1;
  }
  (void)hipFree(c->zpart32);
  (void)hipFree(c->fpart_s);
  c->zpart32 = zp;
  c->fpart_s = fp;
  c->ksplit_s = ks;
  return 0;
}

int default_ksplit_split(Ctx* c) { return choose_ksplit_split(c->np / kBR, c->Mp); }

int launch_pa_split(Ctx* c, int nc, int apply_d4, int prec, hipStream_t s, hipEvent_t* ev) {
  int rc = ensure_split(c, prec);
  if (rc) return rc;
  if (prec == PREC_FP16X2) return launch_all<_Float16>(c, nc, apply_d4, s, ev);
  return launch_all<__bf16>(c, nc, apply_d4, s, ev);
}

}  // namespace tdscha
