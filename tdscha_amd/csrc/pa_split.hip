// Split-precision perturbed averages: FP16 (or BF16) hi/lo compensated
// products on MFMA (v_mfma_f32_16x16x32_{f16,bf16}) with FP32 accumulation.
//
// Every operand x is held as x ~ (hi + lo) / s with hi = T(x s), lo =
// T(x s - hi) (T = fp16 or bf16) and s a power-of-two per column, so a
// product A.B ~ (Ah.Bh + Ah.Bl + Al.Bh) / (sA sB): three MFMAs, relative
// error ~ 2^-22 (fp16) or 2^-16 (bf16) per element, FP32 accumulation.
// The per-image coefficients (c1, c2, alpha, beta) and all reductions across
// work-groups stay FP64.
//
// Storage (built once per context and precision):
//   Uh, Ul, Fh, Fl     [Mp, np]  row-major images       (weights GEMM, K = modes)
//   UhT, UlT, FhT, FlT [np, Mp]  mode-major images      (outer GEMM, K = images)
// so that both GEMMs read K-contiguous 16-byte fragments: the MFMA 16x16x32
// operand layout is A[i][k]: i = lane % 16, k = 8 * (lane / 16) + 0..7 (same
// for B with j), result D[i][j]: j = lane % 16, i = 4 * (lane / 16) + r.
//
// All chains of one batch are processed inside a work-group, so each image
// tile is read from HBM once per batch (template NC = chains per pass).
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

// row-major hi/lo and the mode-major copies, through a 32x32 LDS transpose
template <typename H>
__global__ void split_images(const double* __restrict__ U, const double* __restrict__ F, long Mp,
                             int np, const double* __restrict__ su,
                             const double* __restrict__ sf, u16* __restrict__ Uh,
                             u16* __restrict__ Ul, u16* __restrict__ Fh, u16* __restrict__ Fl,
                             u16* __restrict__ UhT, u16* __restrict__ UlT,
                             u16* __restrict__ FhT, u16* __restrict__ FlT) {
  __shared__ u16 t[4][32][33];
  const long r0 = (long)blockIdx.x * 32;
  const int c0 = blockIdx.y * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
  for (int yy = ty; yy < 32; yy += 8) {
    const long r = r0 + yy;
    const int cc = c0 + tx;
    const double u = U[r * np + cc] * su[cc];
    const double f = F[r * np + cc] * sf[cc];
    const u16 uh = from_d<H>(u), fh = from_d<H>(f);
    const u16 ul = from_d<H>(u - (double)to_f<H>(uh));
    const u16 fl = from_d<H>(f - (double)to_f<H>(fh));
    Uh[r * np + cc] = uh;
    Ul[r * np + cc] = ul;
    Fh[r * np + cc] = fh;
    Fl[r * np + cc] = fl;
    t[0][yy][tx] = uh;
    t[1][yy][tx] = ul;
    t[2][yy][tx] = fh;
    t[3][yy][tx] = fl;
  }
  __syncthreads();
  for (int yy = ty; yy < 32; yy += 8) {
    const long dst = (long)(c0 + yy) * Mp + r0 + tx;
    UhT[dst] = t[0][tx][yy];
    UlT[dst] = t[1][tx][yy];
    FhT[dst] = t[2][tx][yy];
    FlT[dst] = t[3][tx][yy];
  }
}

// ------------------------------------------------------------ per-call prep
// YT[c][j][k] = Y1[c][j][k] sy[j] / su[k] split hi/lo; epilogue factors;
// resets the |c| maxima.  One block per (row j, chain).
template <typename H>
__global__ void prep_y(const double* __restrict__ Y1all, const double* __restrict__ R1all,
                       const double* __restrict__ su, const double* __restrict__ sf,
                       const double* __restrict__ fy, const double* __restrict__ fpsi, int np,
                       u16* __restrict__ YhT, u16* __restrict__ YlT, double* __restrict__ epi,
                       double* __restrict__ cmax) {
  const int j = blockIdx.x, chain = blockIdx.y;
  const double* Y = Y1all + ((long)chain * np + j) * np;
  __shared__ double red[256];
  double mx = 0.0;
  for (int k = threadIdx.x; k < np; k += blockDim.x) mx = fmax(mx, fabs(Y[k] / su[k]));
  red[threadIdx.x] = mx;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + o]);
    __syncthreads();
  }
  const double sy = pow2_scale(red[0]);
  u16* yh = YhT + ((long)chain * np + j) * np;
  u16* yl = YlT + ((long)chain * np + j) * np;
  for (int k = threadIdx.x; k < np; k += blockDim.x) {
    const double v = Y[k] * sy / su[k];
    const u16 h = from_d<H>(v);
    yh[k] = h;
    yl[k] = from_d<H>(v - (double)to_f<H>(h));
  }
  if (threadIdx.x == 0) {
    const double R = R1all[(long)chain * np + j];
    double* e = epi + ((long)chain * np + j) * 4;
    e[0] = 1.0 / (sy * su[j]);
    e[1] = fpsi[j] / (sy * sf[j]);
    e[2] = fy[j] * R / su[j];
    e[3] = R / sf[j];
    if (j == 0) {
      cmax[chain * 2 + 0] = 0.0;
      cmax[chain * 2 + 1] = 0.0;
    }
  }
}

// ----------------------------------------------------------- LDS layout
// Tiles are 128 rows x 32 (16-bit) elements: four 16-byte chunks per row,
// unpadded, chunk index XOR-swizzled with (row / 4) % 4.  Measured on gfx950
// this makes both the 32x32x16 fragment reads (32 consecutive rows x one chunk
// per half-wave) and the staging writes (4 chunks per row) conflict-free.
__device__ __forceinline__ int lds_off(int row, int chunk) {
  return row * 32 + 8 * (chunk ^ ((row >> 2) & 3));
}
constexpr int kTileElems = 128 * 32;  // u16 elements of one 128 x 32 tile

// ----------------------------------------------------------- weights GEMM
// tile 128 rows x 128 cols, 4 waves (2 x 2) of 64 x 64, one wave per SIMD;
// K = np in chunks of 32, double-buffered LDS, one barrier per chunk
constexpr int SW_BM = 128, SW_BN = 128, SW_KC = 32;

template <typename H, int NC>
__global__ __launch_bounds__(256, 1) void weights_split(
    const u16* __restrict__ Uh, const u16* __restrict__ Ul, const u16* __restrict__ Fh,
    const u16* __restrict__ Fl, const u16* __restrict__ YhT, const u16* __restrict__ YlT,
    const double* __restrict__ epi, int np, long Mp, int chain0, int nrt,
    double* __restrict__ part) {
  typedef typename Mma<H>::vec vec;
  const int nct = np / SW_BN;
  // XCD-aware: all column tiles of a row tile land on the same XCD
  const int id = blockIdx.x;
  const int xcd = id & 7, slot = id >> 3;
  const int ct = slot % nct;
  const int rt = xcd + 8 * (slot / nct);
  if (rt >= nrt) return;
  const long r0 = (long)rt * SW_BM;
  const int c0 = ct * SW_BN;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wr = wave >> 1, wc = wave & 1;  // wave rows 64*wr, cols 64*wc

  constexpr int kStage = (2 + 2 * NC) * kTileElems;
  __shared__ __align__(16) u16 smem[2 * kStage];

  v16f acc[NC][2][2];
#pragma unroll
  for (int c = 0; c < NC; ++c)
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) acc[c][i][j] = v16f{};

  // staging: 128 rows x 4 chunks = 512 slots, two per thread
  const int srow = tid >> 2, sch = tid & 3;
  u16x8 ah[2], al[2], bh[NC][2], bl[NC][2];
  auto load = [&](int k0) {
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int row = srow + 64 * h;
      ah[h] = *reinterpret_cast<const u16x8*>(Uh + (r0 + row) * np + k0 + 8 * sch);
      al[h] = *reinterpret_cast<const u16x8*>(Ul + (r0 + row) * np + k0 + 8 * sch);
#pragma unroll
      for (int c = 0; c < NC; ++c) {
        const long off = ((long)(chain0 + c) * np + c0 + row) * np + k0 + 8 * sch;
        bh[c][h] = *reinterpret_cast<const u16x8*>(YhT + off);
        bl[c][h] = *reinterpret_cast<const u16x8*>(YlT + off);
      }
    }
  };
  auto store = [&](u16* st) {
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int o = lds_off(srow + 64 * h, sch);
      *reinterpret_cast<u16x8*>(st + o) = ah[h];
      *reinterpret_cast<u16x8*>(st + kTileElems + o) = al[h];
#pragma unroll
      for (int c = 0; c < NC; ++c) {
        *reinterpret_cast<u16x8*>(st + (2 + 2 * c) * kTileElems + o) = bh[c][h];
        *reinterpret_cast<u16x8*>(st + (3 + 2 * c) * kTileElems + o) = bl[c][h];
      }
    }
  };
  load(0);
  store(smem);
  __syncthreads();
  const int nk = np / SW_KC;
  for (int kc = 0; kc < nk; ++kc) {
    u16* cur = smem + (kc & 1) * kStage;
    if (kc + 1 < nk) load((kc + 1) * SW_KC);
    // operands swapped: D = V^T block, lane holds image row 32 i + lane % 32 and
    // columns 32 j + 8 q + 4 (lane / 32) + 0..3 (q = 0..3).  All fragment reads
    // of the chunk are issued first; the MFMAs then run product-major so that
    // dependent MFMAs on one accumulator are 4 NC instructions apart.
    const int lr = lane & 31, lh = lane >> 5;
    vec fa[2][2][2], fb[2][NC][2][2];   // [ksb][block][hi/lo]
#pragma unroll
    for (int ksb = 0; ksb < 2; ++ksb) {
#pragma unroll
      for (int i = 0; i < 2; ++i) {
        const int o = lds_off(64 * wr + 32 * i + lr, 2 * ksb + lh);
        fa[ksb][i][0] = *reinterpret_cast<const vec*>(cur + o);
        fa[ksb][i][1] = *reinterpret_cast<const vec*>(cur + kTileElems + o);
      }
#pragma unroll
      for (int c = 0; c < NC; ++c)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
          const int o = lds_off(64 * wc + 32 * j + lr, 2 * ksb + lh);
          fb[ksb][c][j][0] = *reinterpret_cast<const vec*>(cur + (2 + 2 * c) * kTileElems + o);
          fb[ksb][c][j][1] = *reinterpret_cast<const vec*>(cur + (3 + 2 * c) * kTileElems + o);
        }
    }
#pragma unroll
    for (int ksb = 0; ksb < 2; ++ksb)
#pragma unroll
      for (int pr = 0; pr < 3; ++pr)
#pragma unroll
        for (int c = 0; c < NC; ++c)
#pragma unroll
          for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int i = 0; i < 2; ++i)
              acc[c][i][j] = Mma<H>::mma(fb[ksb][c][j][pr == 1 ? 1 : 0],
                                         fa[ksb][i][pr == 2 ? 1 : 0], acc[c][i][j]);
    if (kc + 1 < nk) store(smem + ((kc + 1) & 1) * kStage);
    __syncthreads();
  }

  // epilogue straight from the accumulators: U and F loaded once (8 bytes per
  // 4 columns) and shared by all chains; per-lane partials -> FP64, reduced
  // over the two half-waves by a shuffle and over the 2 column waves in LDS.
  double* red = reinterpret_cast<double*>(smem);  // [2 wc][NC][128 rows][4]
  const int lr = lane & 31, lh = lane >> 5;
  const double* e[NC];
#pragma unroll
  for (int c = 0; c < NC; ++c) e[c] = epi + ((long)(chain0 + c) * np + c0) * 4;
  typedef u16 u16x4 __attribute__((ext_vector_type(4)));
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    const int row = 64 * wr + 32 * i + lr;
    const long grow = r0 + row;
    double p[NC][4];
#pragma unroll
    for (int c = 0; c < NC; ++c) p[c][0] = p[c][1] = p[c][2] = p[c][3] = 0.0;
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int col = 64 * wc + 32 * j + 8 * q + 4 * lh;
        const u16x4 uh = *reinterpret_cast<const u16x4*>(Uh + grow * np + c0 + col);
        const u16x4 ul = *reinterpret_cast<const u16x4*>(Ul + grow * np + c0 + col);
        const u16x4 fh = *reinterpret_cast<const u16x4*>(Fh + grow * np + c0 + col);
        const u16x4 fl = *reinterpret_cast<const u16x4*>(Fl + grow * np + c0 + col);
        float u[4], f[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
          u[t] = to_f<H>(uh[t]) + to_f<H>(ul[t]);
          f[t] = to_f<H>(fh[t]) + to_f<H>(fl[t]);
        }
#pragma unroll
        for (int c = 0; c < NC; ++c) {
          float q0 = 0.f, q1 = 0.f, q2 = 0.f, q3 = 0.f;
#pragma unroll
          for (int t = 0; t < 4; ++t) {
            const double* ee = e[c] + (col + t) * 4;
            const float v = acc[c][i][j][4 * q + t];
            q0 += v * u[t] * (float)ee[0];
            q1 += v * f[t] * (float)ee[1];
            q2 += u[t] * (float)ee[2];
            q3 += f[t] * (float)ee[3];
          }
          p[c][0] += (double)q0;
          p[c][1] += (double)q1;
          p[c][2] += (double)q2;
          p[c][3] += (double)q3;
        }
      }
#pragma unroll
    for (int c = 0; c < NC; ++c)
#pragma unroll
      for (int qq = 0; qq < 4; ++qq) p[c][qq] += __shfl_xor(p[c][qq], 32);
    if (lh == 0) {
#pragma unroll
      for (int c = 0; c < NC; ++c)
#pragma unroll
        for (int qq = 0; qq < 4; ++qq) red[((wc * NC + c) * SW_BM + row) * 4 + qq] = p[c][qq];
    }
  }
  __syncthreads();
  for (int e2 = tid; e2 < NC * SW_BM; e2 += blockDim.x) {
    const int c = e2 / SW_BM, row = e2 % SW_BM;
    const double* a0 = red + ((0 * NC + c) * SW_BM + row) * 4;
    const double* a1 = red + ((1 * NC + c) * SW_BM + row) * 4;
    double* o = part + (((long)(chain0 + c) * nct + ct) * Mp + r0 + row) * 4;
    o[0] = a0[0] + a1[0];
    o[1] = a0[1] + a1[1];
    o[2] = a0[2] + a1[2];
    o[3] = a0[3] + a1[3];
  }
}

// coefficients from the column-tile partials (fixed order); |c| maxima
__global__ void finalize_coef(const double* __restrict__ part, const double* __restrict__ rho_img,
                              long Mp, int nct, int apply_d4, double* __restrict__ coef_all,
                              float* __restrict__ coef32_all, double* __restrict__ cmax) {
  const int chain = blockIdx.y;
  const long row = (long)blockIdx.x * blockDim.x + threadIdx.x;
  double m1 = 0.0, m2 = 0.0;
  if (row < Mp) {
    double s[4] = {0, 0, 0, 0};
    for (int ct = 0; ct < nct; ++ct) {
      const double* p = part + (((long)chain * nct + ct) * Mp + row) * 4;
      s[0] += p[0]; s[1] += p[1]; s[2] += p[2]; s[3] += p[3];
    }
    const double rho = rho_img[row];
    const double c1 = rho * (-s[2] / 3.0 + apply_d4 * s[0] / 8.0);
    const double c2 = rho * (-s[3] / 3.0 + apply_d4 * s[1] / 4.0);
    const double al = -rho * s[0] / 6.0, be = -rho * s[1] / 3.0;
    if (coef_all) {
      double* c = coef_all + ((long)chain * Mp + row) * 4;
      c[0] = c1; c[1] = c2; c[2] = al; c[3] = be;
    }
    float* c32 = coef32_all + (long)chain * 4 * Mp + row;   // planes [4][Mp]
    c32[0] = (float)c1; c32[Mp] = (float)c2; c32[2 * Mp] = (float)al; c32[3 * Mp] = (float)be;
    m1 = fabs(c1);
    m2 = fabs(c2);
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
    atomic_max_pos(&cmax[chain * 2 + 0], r1[0]);
    atomic_max_pos(&cmax[chain * 2 + 1], r2[0]);
  }
}

// ------------------------------------------------------------- outer GEMM
// Z_c[mu, nu] = sum_k U[k, mu] B_c[k, nu], B_c = c1 F + (c2/2) fY U, over a
// K range of images; tile 128 (mu) x 128 (nu), 4 waves (2 x 2) of 64 x 64,
// double-buffered LDS.  B_c is built in FP32 from the loaded hi/lo images and
// split again while staging.  The raw FP32 accumulators go to an FP32 partial
// buffer (exact); the scales are undone in the FP64 reduction.
constexpr int SO_T = 128, SO_KC = 32;

template <typename H, int NC>
__global__ __launch_bounds__(256, 1) void outer_split(
    const u16* __restrict__ UhT, const u16* __restrict__ UlT, const u16* __restrict__ FhT,
    const u16* __restrict__ FlT, const float* __restrict__ coef32_all,
    const double* __restrict__ cmax, const double* __restrict__ umx,
    const double* __restrict__ fmx, const double* __restrict__ su,
    const double* __restrict__ sf, const double* __restrict__ fy, int np, long Mp, int ksplit,
    int ntile, long rows_per_split, int chain0, float* __restrict__ zpart_all,
    double* __restrict__ fpart_all, double* __restrict__ sb_all) {
  typedef typename Mma<H>::vec vec;
  const int id = blockIdx.x;
  const int xcd = id & 7, slot = id >> 3;
  const int t2 = ntile * ntile;
  const int ks = xcd + 8 * (slot / t2);
  const int tile = slot % t2;
  const int tm = tile / ntile, tn = tile % ntile;
  const int mu0 = tm * SO_T, nu0 = tn * SO_T;
  const bool diag = (tm == tn);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int lr = lane & 31, lh = lane >> 5;
  const int wr = wave >> 1, wc = wave & 1;

  constexpr int kStage = (2 + 2 * NC) * kTileElems;
  __shared__ __align__(16) u16 smem[2 * kStage];

  // staging: 128 modes x 4 chunks of 8 images = 512 slots, two per thread
  const int sn = tid >> 2, sch = tid & 3;
  float gF[NC][2], gU[NC][2], inv_sf[2], fyu[2];
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int nu = nu0 + sn + 64 * h;
    inv_sf[h] = (float)(1.0 / sf[nu]);
    fyu[h] = (float)(fy[nu] / su[nu]);
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      // B scale per (chain, nu): sb with bound(B) * sb < 2^12
      const double bound = cmax[(chain0 + c) * 2 + 0] * fmx[nu] +
                           0.5 * cmax[(chain0 + c) * 2 + 1] * fy[nu] * umx[nu];
      const double sb = pow2_scale(bound);
      gF[c][h] = (float)(sb / sf[nu]);
      gU[c][h] = (float)(0.5 * sb * fy[nu] / su[nu]);
      if (ks == 0 && tm == 0 && sch == 0) sb_all[(long)(chain0 + c) * np + nu] = sb;
    }
  }

  v16f acc[NC][2][2];
#pragma unroll
  for (int c = 0; c < NC; ++c)
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) acc[c][i][j] = v16f{};

  const long kb = (long)ks * rows_per_split;
  long ke = kb + rows_per_split;
  if (ke > Mp) ke = Mp;
  double fl_acc[NC][2];
#pragma unroll
  for (int c = 0; c < NC; ++c) fl_acc[c][0] = fl_acc[c][1] = 0.0;

  u16x8 ah[2], al[2], uh[2], ul[2], fh[2], fl[2];
  float4 cq[NC][4][2];  // per chain: c1, c2, alpha, beta of this thread's 8 images
  auto load = [&](long k0) {
    const long kk = k0 + 8 * sch;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int m = sn + 64 * h;
      ah[h] = *reinterpret_cast<const u16x8*>(UhT + (long)(mu0 + m) * Mp + kk);
      al[h] = *reinterpret_cast<const u16x8*>(UlT + (long)(mu0 + m) * Mp + kk);
      if (diag) {
        uh[h] = ah[h];
        ul[h] = al[h];
      } else {
        uh[h] = *reinterpret_cast<const u16x8*>(UhT + (long)(nu0 + m) * Mp + kk);
        ul[h] = *reinterpret_cast<const u16x8*>(UlT + (long)(nu0 + m) * Mp + kk);
      }
      fh[h] = *reinterpret_cast<const u16x8*>(FhT + (long)(nu0 + m) * Mp + kk);
      fl[h] = *reinterpret_cast<const u16x8*>(FlT + (long)(nu0 + m) * Mp + kk);
    }
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      const float* base = coef32_all + (long)(chain0 + c) * 4 * Mp + kk;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        if (q >= 2 && !diag) continue;
        cq[c][q][0] = *reinterpret_cast<const float4*>(base + q * Mp);
        cq[c][q][1] = *reinterpret_cast<const float4*>(base + q * Mp + 4);
      }
    }
  };
  auto build_store = [&](u16* st) {
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int m = sn + 64 * h;
      const int o = lds_off(m, sch);
      *reinterpret_cast<u16x8*>(st + o) = ah[h];
      *reinterpret_cast<u16x8*>(st + kTileElems + o) = al[h];
      float uu[8], ff[8];
#pragma unroll
      for (int t = 0; t < 8; ++t) {
        uu[t] = to_f<H>(uh[h][t]) + to_f<H>(ul[h][t]);
        ff[t] = to_f<H>(fh[h][t]) + to_f<H>(fl[h][t]);
      }
#pragma unroll
      for (int c = 0; c < NC; ++c) {
        const float c1v[8] = {cq[c][0][0].x, cq[c][0][0].y, cq[c][0][0].z, cq[c][0][0].w,
                              cq[c][0][1].x, cq[c][0][1].y, cq[c][0][1].z, cq[c][0][1].w};
        const float c2v[8] = {cq[c][1][0].x, cq[c][1][0].y, cq[c][1][0].z, cq[c][1][0].w,
                              cq[c][1][1].x, cq[c][1][1].y, cq[c][1][1].z, cq[c][1][1].w};
        u16x8 bh, bl;
#pragma unroll
        for (int t = 0; t < 8; ++t) {
          const float b = c1v[t] * ff[t] * gF[c][h] + c2v[t] * uu[t] * gU[c][h];
          const u16 hb = from_f<H>(b);
          bh[t] = hb;
          bl[t] = from_f<H>(b - to_f<H>(hb));
        }
        *reinterpret_cast<u16x8*>(st + (2 + 2 * c) * kTileElems + o) = bh;
        *reinterpret_cast<u16x8*>(st + (3 + 2 * c) * kTileElems + o) = bl;
        if (diag) {
          const float av[8] = {cq[c][2][0].x, cq[c][2][0].y, cq[c][2][0].z, cq[c][2][0].w,
                               cq[c][2][1].x, cq[c][2][1].y, cq[c][2][1].z, cq[c][2][1].w};
          const float bv[8] = {cq[c][3][0].x, cq[c][3][0].y, cq[c][3][0].z, cq[c][3][0].w,
                               cq[c][3][1].x, cq[c][3][1].y, cq[c][3][1].z, cq[c][3][1].w};
          float fsum = 0.f;
#pragma unroll
          for (int t = 0; t < 8; ++t) fsum += av[t] * ff[t] * inv_sf[h] + bv[t] * uu[t] * fyu[h];
          fl_acc[c][h] += (double)fsum;
        }
      }
    }
  };
  const long nk = (ke > kb) ? (ke - kb) / SO_KC : 0;
  if (nk > 0) {
    load(kb);
    build_store(smem);
  }
  __syncthreads();
  for (long kc = 0; kc < nk; ++kc) {
    u16* cur = smem + (kc & 1) * kStage;
    if (kc + 1 < nk) load(kb + (kc + 1) * SO_KC);
    // per k-step: issue all fragment reads, then the MFMAs product-major
#pragma unroll
    for (int ksb = 0; ksb < 2; ++ksb) {
      vec fa[2][2], fb[NC][2][2];
#pragma unroll
      for (int i = 0; i < 2; ++i) {
        const int oo = lds_off(64 * wr + 32 * i + lr, 2 * ksb + lh);
        fa[i][0] = *reinterpret_cast<const vec*>(cur + oo);
        fa[i][1] = *reinterpret_cast<const vec*>(cur + kTileElems + oo);
      }
#pragma unroll
      for (int c = 0; c < NC; ++c)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
          const int oo = lds_off(64 * wc + 32 * j + lr, 2 * ksb + lh);
          fb[c][j][0] = *reinterpret_cast<const vec*>(cur + (2 + 2 * c) * kTileElems + oo);
          fb[c][j][1] = *reinterpret_cast<const vec*>(cur + (3 + 2 * c) * kTileElems + oo);
        }
#pragma unroll
      for (int pr = 0; pr < 3; ++pr)
#pragma unroll
        for (int c = 0; c < NC; ++c)
#pragma unroll
          for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int i = 0; i < 2; ++i)
              acc[c][i][j] = Mma<H>::mma(fa[i][pr == 2 ? 1 : 0], fb[c][j][pr == 1 ? 1 : 0],
                                         acc[c][i][j]);
    }
    if (kc + 1 < nk) build_store(smem + ((kc + 1) & 1) * kStage);
    __syncthreads();
  }
  // raw partial tiles (FP32, exact copies of the accumulators)
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    float* zp = zpart_all + ((long)(chain0 + c) * ksplit + ks) * np * np;
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j)
#pragma unroll
        for (int r = 0; r < 16; ++r) {
          const int mm = 64 * wr + 32 * i + 8 * (r >> 2) + 4 * lh + (r & 3);
          const int nn = 64 * wc + 32 * j + lr;
          zp[(long)(mu0 + mm) * np + nu0 + nn] = acc[c][i][j][r];
        }
  }
  if (diag) {
    double* red = reinterpret_cast<double*>(smem);  // [NC][128][4]
#pragma unroll
    for (int c = 0; c < NC; ++c)
#pragma unroll
      for (int h = 0; h < 2; ++h) red[(c * SO_T + sn + 64 * h) * 4 + sch] = fl_acc[c][h];
    __syncthreads();
    for (int e = tid; e < NC * SO_T; e += blockDim.x) {
      const int c = e / SO_T, nn = e % SO_T;
      const double* q = red + (c * SO_T + nn) * 4;
      fpart_all[((long)(chain0 + c) * ksplit + ks) * np + nu0 + nn] = (q[0] + q[1]) + (q[2] + q[3]);
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
    double s = 0.0;
    for (int k = 0; k < ksplit; ++k) s += (double)zp[(long)k * nn + e];
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
int ensure_split(Ctx* c, int prec) {
  SplitStore& st = c->split[prec];
  if (st.ready) return 0;
  const long Mp = c->Mp;
  const int np = c->np;
  const size_t half_bytes = sizeof(u16) * (size_t)Mp * np;
  void** arr[] = {&st.Uh, &st.Ul, &st.Fh, &st.Fl, &st.UhT, &st.UlT, &st.FhT, &st.FlT};
  for (void** p : arr) TDS_CHECK_HIP(hipMalloc(p, half_bytes));
  TDS_CHECK_HIP(hipMalloc((void**)&st.su, sizeof(double) * np));
  TDS_CHECK_HIP(hipMalloc((void**)&st.sf, sizeof(double) * np));
  TDS_CHECK_HIP(hipMalloc((void**)&st.umax, sizeof(double) * np));
  TDS_CHECK_HIP(hipMalloc((void**)&st.fmax, sizeof(double) * np));
  if (!c->Y1h) {
    TDS_CHECK_HIP(hipMalloc(&c->Y1h, sizeof(u16) * kMaxChains * (size_t)np * np));
    TDS_CHECK_HIP(hipMalloc(&c->Y1l, sizeof(u16) * kMaxChains * (size_t)np * np));
    TDS_CHECK_HIP(hipMalloc((void**)&c->epi, sizeof(double) * kMaxChains * np * 4));
    TDS_CHECK_HIP(hipMalloc((void**)&c->cmax, sizeof(double) * kMaxChains * 2));
    TDS_CHECK_HIP(hipMalloc((void**)&c->coef32, sizeof(float) * kMaxChains * Mp * 4));
    TDS_CHECK_HIP(hipMalloc((void**)&c->wpart, sizeof(double) * kMaxChains * (np / SW_BN) * Mp * 4));
    // K split of the split paths: ~1536 work-groups over the 128 x 128 tiles,
    // at least 2048 images per range, multiple of 8 (one range per XCD in turn)
    const int ntile = np / SO_T;
    long ks = (1536 + ntile * ntile - 1) / (ntile * ntile);
    ks = (ks + 7) / 8 * 8;
    while (ks > 8 && ks * 2048 > Mp) ks -= 8;
    if (ks < 8) ks = 8;
    c->ksplit_s = (int)ks;
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
  dim3 g((unsigned)(Mp / 32), np / 32);
  if (prec == PREC_FP16X2)
    hipLaunchKernelGGL(split_images<_Float16>, g, dim3(256), 0, s, c->U, c->F, Mp, np, st.su,
                       st.sf, (u16*)st.Uh, (u16*)st.Ul, (u16*)st.Fh, (u16*)st.Fl, (u16*)st.UhT,
                       (u16*)st.UlT, (u16*)st.FhT, (u16*)st.FlT);
  else
    hipLaunchKernelGGL(split_images<__bf16>, g, dim3(256), 0, s, c->U, c->F, Mp, np, st.su,
                       st.sf, (u16*)st.Uh, (u16*)st.Ul, (u16*)st.Fh, (u16*)st.Fl, (u16*)st.UhT,
                       (u16*)st.UlT, (u16*)st.FhT, (u16*)st.FlT);
  TDS_CHECK_HIP(hipGetLastError());
  TDS_CHECK_HIP(hipStreamSynchronize(s));
  st.ready = true;
  return 0;
}

template <typename H, int NC>
static void launch_group(Ctx* c, const SplitStore& st, int chain0, int apply_d4, hipStream_t s) {
  const int np = c->np;
  const long Mp = c->Mp;
  const int nct = np / SW_BN;
  const int nrt = (int)(Mp / SW_BM);
  const int nrt8 = round_up(nrt, 8);
  hipLaunchKernelGGL((weights_split<H, NC>), dim3((unsigned)(nrt8 * nct)), dim3(256), 0, s,
                     (const u16*)st.Uh, (const u16*)st.Ul, (const u16*)st.Fh, (const u16*)st.Fl,
                     (const u16*)c->Y1h, (const u16*)c->Y1l, c->epi, np, Mp, chain0, nrt,
                     c->wpart);
  (void)apply_d4;
}

template <typename H, int NC>
static void launch_outer(Ctx* c, const SplitStore& st, int chain0, long rps, hipStream_t s) {
  const int ntile = c->np / SO_T;
  hipLaunchKernelGGL((outer_split<H, NC>), dim3((unsigned)(ntile * ntile * c->ksplit_s)),
                     dim3(256), 0, s, (const u16*)st.UhT, (const u16*)st.UlT, (const u16*)st.FhT,
                     (const u16*)st.FlT, c->coef32, c->cmax, st.umax, st.fmax, st.su, st.sf,
                     c->fy, c->np, c->Mp, c->ksplit_s, ntile, rps, chain0, c->zpart32,
                     c->fpart_s, c->sbsc);
}

template <typename H>
static int launch_all(Ctx* c, int nc, int apply_d4, hipStream_t s, hipEvent_t* ev) {
  const SplitStore& st = c->split[std::is_same<H, _Float16>::value ? PREC_FP16X2 : PREC_BF16X2];
  const int np = c->np;
  const long Mp = c->Mp;
  if (ev) (void)hipEventRecord(ev[0], s);
  hipLaunchKernelGGL(prep_y<H>, dim3(np, nc), dim3(256), 0, s, c->Y1, c->R1, st.su, st.sf, c->fy,
                     c->fpsi, np, (u16*)c->Y1h, (u16*)c->Y1l, c->epi, c->cmax);
  // chains in groups of up to 3 per pass (two LDS stages of (2 + 2 NC) tiles)
  for (int c0 = 0; c0 < nc;) {
    const int g = nc - c0 >= 3 ? 3 : nc - c0;
    switch (g) {
      case 3: launch_group<H, 3>(c, st, c0, apply_d4, s); break;
      case 2: launch_group<H, 2>(c, st, c0, apply_d4, s); break;
      default: launch_group<H, 1>(c, st, c0, apply_d4, s); break;
    }
    c0 += g;
  }
  const int nct = np / SW_BN;
  hipLaunchKernelGGL(finalize_coef, dim3((unsigned)((Mp + 255) / 256), nc), dim3(256), 0, s,
                     c->wpart, c->rho_img, Mp, nct, apply_d4, (double*)nullptr, c->coef32,
                     c->cmax);
  if (ev) (void)hipEventRecord(ev[1], s);
  const long rps = round_up_l((Mp + c->ksplit_s - 1) / c->ksplit_s, SO_KC);
  for (int c0 = 0; c0 < nc;) {
    const int g = nc - c0 >= 3 ? 3 : nc - c0;
    switch (g) {
      case 3: launch_outer<H, 3>(c, st, c0, rps, s); break;
      case 2: launch_outer<H, 2>(c, st, c0, rps, s); break;
      default: launch_outer<H, 1>(c, st, c0, rps, s); break;
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

int launch_pa_split(Ctx* c, int nc, int apply_d4, int prec, hipStream_t s, hipEvent_t* ev) {
  int rc = ensure_split(c, prec);
  if (rc) return rc;
  if (prec == PREC_FP16X2) return launch_all<_Float16>(c, nc, apply_d4, s, ev);
  return launch_all<__bf16>(c, nc, apply_d4, s, ev);
}

}  // namespace tdscha
