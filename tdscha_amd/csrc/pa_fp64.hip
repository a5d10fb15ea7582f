// FP64 ensemble-averaged D3/D4 application on MFMA (v_mfma_f64_16x16x4f64).
//
// The per-image algebra of CModules/LanczosFunctions.c:2082-2754 is regrouped
// into three passes over the device-resident symmetric image ensemble
// (rows = images b = i*N_sym + j, columns = modes):
//
//   weights_f64   V = U . Y1 (tiled GEMM, K = modes) with a fused row-dot
//                 epilogue: s1 = u.Y1.u, s2 = u.Y1.g, s3 = a.R1, s4 = f.R1
//                 -> per image c1, c2 (d2v weights) and alpha, beta (force
//                 weights).
//   outer_f64     Z = U^T B with B = c1 F + (c2/2) fY U  (split-K GEMM over
//                 the images, LDS-staged image tiles), and on the diagonal
//                 tiles the force average sum alpha F + beta fY U.
//   reduce        d2v = fY Z + (fY Z)^T, f = sum of partials, scaled by
//                 1/(N_eff N_sym); fixed summation order (deterministic).
//
// MFMA f64 16x16x4 operand layout (one double per lane):
//   A[i][k]: i = lane % 16, k = lane / 16;  B[k][j]: j = lane % 16, k = lane / 16
//   D[i][j]: j = lane % 16, i = 4 * r + lane / 16  (r = register 0..3; note the
//   row interleave, unlike the FP32-output MFMAs whose row is 4 * (lane / 16) + r)
#include "common.hpp"

namespace tdscha {

// ---------------------------------------------------------------------------
// image expansion: U[b, m] = sum_k S_{blk(m)}[j, pos(m), k] X[i, modes_blk[k]]
// ---------------------------------------------------------------------------
__global__ void expand_images(const double* __restrict__ X, const double* __restrict__ Y,
                              const double* __restrict__ rho, const int* __restrict__ mode_block,
                              const int* __restrict__ mode_pos, const int* __restrict__ block_dim,
                              const int* __restrict__ block_moff, const long* __restrict__ block_soff,
                              const int* __restrict__ block_modes, const double* __restrict__ syms,
                              int n, int np, int n_syms, long M, double* __restrict__ U,
                              double* __restrict__ F, double* __restrict__ rho_img) {
  const long row = blockIdx.x;
  const int m = threadIdx.x + blockIdx.y * blockDim.x;
  if (m >= np) return;
  double su = 0.0, sf = 0.0;
  if (row < M && m < n) {
    const long i = row / n_syms;
    const int j = (int)(row % n_syms);
    const int b = mode_block[m];
    const int d = block_dim[b];
    const double* S = syms + block_soff[b] + ((long)j * d + mode_pos[m]) * d;
    const int* modes = block_modes + block_moff[b];
    const double* x = X + i * n;
    const double* y = Y + i * n;
    for (int k = 0; k < d; ++k) {
      const double s = S[k];
      su += s * x[modes[k]];
      sf += s * y[modes[k]];
    }
  }
  U[row * np + m] = su;
  F[row * np + m] = sf;
  if (m == 0) rho_img[row] = (row < M) ? rho[row / n_syms] : 0.0;
}

// ---------------------------------------------------------------------------
// Work-group order of both GEMMs: the hardware deals work-groups to the 8
// XCDs in turn (id % 8), so the chains of one tile get ids 8 apart and run on
// one XCD at about the same time, sharing its L2 for the image rows.
// ---------------------------------------------------------------------------

// weights: one work-group = 64 image rows x a 128-column pass; K = np in
// chunks of 16, double-buffered: the next chunk's global loads are in
// flight while the MFMAs consume the current one (one barrier per chunk)
constexpr int WA_KC = 16;
// LDS row padding (doubles) of the K-chunk tiles: 16 puts the four 16-lane
// groups of an MFMA operand read in distinct bank halves
#ifndef TDS_F64_PAD
#define TDS_F64_PAD 16
#endif
constexpr int kPad = TDS_F64_PAD;

__global__ __launch_bounds__(256, 2) void weights_f64(
    const double* __restrict__ U, const double* __restrict__ F,
    const double* __restrict__ Y1all, const double* __restrict__ R1all,
    const double* __restrict__ fy, const double* __restrict__ fpsi,
    const double* __restrict__ rho_img, int np, long Mp, int apply_d4, long tile0, long tile1,
    long row_lo, long row_hi, int nc, double* __restrict__ coef_all) {
  const int xcd = blockIdx.x & 7;
  const int jj = blockIdx.x >> 3;
  const int chain = jj % nc;
  const long tile = tile0 + (long)(jj / nc) * 8 + xcd;
  if (tile >= tile1) return;   // whole work-group, before any barrier
  const double* Y1 = Y1all + (long)chain * np * np;
  const double* R1 = R1all + (long)chain * np;
  double* coef = coef_all + (long)chain * Mp * 4;
  const long r0 = tile * 64;
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int lr = lane & 15;   // lane within a 16-group
  const int lg = lane >> 4;   // 16-group index

  __shared__ double sU[2][64][WA_KC + 1];
  __shared__ double sY[2][WA_KC][kTile + kPad];
  __shared__ double red[4][64][4];
  __shared__ double rowsum[64][4];

  rowsum[tid >> 2][tid & 3] = 0.0;

  double ru[4], ry[8];
  for (int c0 = 0; c0 < np; c0 += kTile) {
    v4d acc[4][2];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b) acc[a][b] = v4d{0.0, 0.0, 0.0, 0.0};

    // U[r0:r0+64, k0:k0+16] (4 per thread) and Y1[k0:k0+16, c0:c0+128] (8 per thread)
    auto load = [&](int k0) {
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int idx = tid + 256 * e;
        ru[e] = U[(r0 + (idx >> 4)) * np + k0 + (idx & 15)];
      }
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        const int idx = tid + 256 * e;
        ry[e] = Y1[(long)(k0 + (idx >> 7)) * np + c0 + (idx & 127)];
      }
    };
    auto store = [&](int b) {
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int idx = tid + 256 * e;
        sU[b][idx >> 4][idx & 15] = ru[e];
      }
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        const int idx = tid + 256 * e;
        sY[b][idx >> 7][idx & 127] = ry[e];
      }
    };
    load(0);
    store(0);
    __syncthreads();
    int cur = 0;
    for (int k0 = 0; k0 < np; k0 += WA_KC) {
      const bool more = k0 + WA_KC < np;
      if (more) load(k0 + WA_KC);
#pragma unroll
      for (int kk = 0; kk < WA_KC; kk += 4) {
        double a[4], b[2];
#pragma unroll
        for (int rb = 0; rb < 4; ++rb) a[rb] = sU[cur][16 * rb + lr][kk + lg];
#pragma unroll
        for (int cb = 0; cb < 2; ++cb) b[cb] = sY[cur][kk + lg][32 * wave + 16 * cb + lr];
#pragma unroll
        for (int rb = 0; rb < 4; ++rb)
#pragma unroll
          for (int cb = 0; cb < 2; ++cb)
            acc[rb][cb] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[rb], b[cb], acc[rb][cb], 0, 0, 0);
      }
      if (more) store(cur ^ 1);
      __syncthreads();
      cur ^= 1;
    }

    // fused epilogue: partial row dots over this wave's 32 columns
#pragma unroll
    for (int rb = 0; rb < 4; ++rb) {
      double p[4][4];
#pragma unroll
      for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int q = 0; q < 4; ++q) p[r][q] = 0.0;
#pragma unroll
      for (int cb = 0; cb < 2; ++cb) {
        const int col = c0 + 32 * wave + 16 * cb + lr;
        const double fyc = fy[col], fpc = fpsi[col], r1c = R1[col];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const long row = r0 + 16 * rb + 4 * r + lg;
          const double u = U[row * np + col];
          const double f = F[row * np + col];
          const double v = acc[rb][cb][r];
          p[r][0] += v * u;
          p[r][1] += v * f * fpc;
          p[r][2] += u * fyc * r1c;
          p[r][3] += f * r1c;
        }
      }
#pragma unroll
      for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          double v = p[r][q];
          v += __shfl_xor(v, 1);
          v += __shfl_xor(v, 2);
          v += __shfl_xor(v, 4);
          v += __shfl_xor(v, 8);
          p[r][q] = v;
        }
      if (lr == 0) {
#pragma unroll
        for (int r = 0; r < 4; ++r)
#pragma unroll
          for (int q = 0; q < 4; ++q) red[wave][16 * rb + 4 * r + lg][q] = p[r][q];
      }
    }
    __syncthreads();
    {
      const int rr = tid >> 2, q = tid & 3;
      rowsum[rr][q] += (red[0][rr][q] + red[1][rr][q]) + (red[2][rr][q] + red[3][rr][q]);
    }
  }
  __syncthreads();
  if (tid < 64) {
    const long row = r0 + tid;
    const double s1 = rowsum[tid][0];  // u.Y1.u
    const double s2 = rowsum[tid][1];  // u.Y1.g  (Y1 symmetrised on upload)
    const double s3 = rowsum[tid][2];  // a.R1
    const double s4 = rowsum[tid][3];  // f.R1
    // images outside [row_lo, row_hi) carry no weight (image-range calls)
    const double rho = (row >= row_lo && row < row_hi) ? rho_img[row] : 0.0;
    double* c = coef + row * 4;
    c[0] = rho * (-s3 / 3.0 + apply_d4 * s1 / 8.0);
    c[1] = rho * (-s4 / 3.0 + apply_d4 * s2 / 4.0);
    c[2] = -rho * s1 / 6.0;
    c[3] = -rho * s2 / 3.0;
  }
}

// ---------------------------------------------------------------------------
// outer: split-K Z[mu, nu] = sum_b U[b, mu] B[b, nu] over a range of images,
// 128 x 128 tiles, 4 waves of 64 x 64; K in chunks of 8 images,
// double-buffered like the weights GEMM (raw U, F and coefficients of the
// next chunk in registers during the MFMAs, B built when they are stored).
// Two work-groups per CU: one's barrier overlaps the other's MFMAs.
// ---------------------------------------------------------------------------
constexpr int WB_KC = 8;
constexpr int kOW = TDS_F64_OUTER_WAVES;   // 4 (64 x 64 per wave) or 8 (64 x 32 per wave)
constexpr int kOT = kF64OuterThreads;      // threads
constexpr int kWN = kOW == 4 ? 64 : 32;    // columns per wave
constexpr int kNJ = kWN / 16;              // 16-column blocks per wave
constexpr int kRS = kOT / kTile;           // image rows staged per pass of the threads
constexpr int kRE = WB_KC / kRS;           // rows per thread and chunk

__global__ __launch_bounds__(kOT, kOW == 8 ? 4 : 2) void outer_f64(
    const double* __restrict__ U, const double* __restrict__ F,
    const double* __restrict__ coef_all, const double* __restrict__ fy, int np, long Mp,
    int ksplit, int ntile, long kbase, long kend, long rows_per_split, int accumulate, int nc,
    double* __restrict__ zpart_all, double* __restrict__ fpart_all) {
  const int xcd = blockIdx.x & 7;
  const int jj = blockIdx.x >> 3;
  const int chain = jj % nc;
  const int rest = jj / nc;
  const int t2 = ntile * ntile;
  const int tile = rest % t2;
  const int ks = xcd + 8 * (rest / t2);
  const int tm = tile / ntile, tn = tile % ntile;
  const double* coef = coef_all + (long)chain * Mp * 4;
  double* zpart = zpart_all + ((long)chain * ksplit + ks) * np * np;
  double* fpart = fpart_all + ((long)chain * ksplit + ks) * np;
  const int mu0 = tm * kTile, nu0 = tn * kTile;
  const bool diag = (tm == tn);

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int lr = lane & 15, lg = lane >> 4;
  const int wm = wave / (kTile / kWN), wn = wave % (kTile / kWN);   // 64 rows x kWN columns

  __shared__ double sA[2][WB_KC][kTile + kPad];
  __shared__ double sB[2][WB_KC][kTile + kPad];
  __shared__ double sF[kRS][kTile];

  v4d acc[4][kNJ];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < kNJ; ++b) acc[a][b] = v4d{0.0, 0.0, 0.0, 0.0};

  const long kb = kbase + (long)ks * rows_per_split;
  long ke = kb + rows_per_split;
  if (ke > kend) ke = kend;
  const int col = tid & 127;
  const int rsub = tid >> 7;  // rows rsub, rsub + kRS, ... of a chunk
  const double fyn = fy[nu0 + col];
  double fl = 0.0;
  // next chunk's raw values; on diagonal tiles U of the nu columns is the
  // A value itself, and rx, ry carry alpha, beta instead
  double rc1[kRE], rc2[kRE], rum[kRE], rfn[kRE], rx[kRE], ry[kRE];
  const long rstride = (long)kRS * np;       // between a thread's rows
  auto load = [&](long k0) {
    const long row0 = k0 + rsub;
    const double* pu = U + row0 * np + col;
    const double* pf = F + row0 * np + nu0 + col;
    const double* pc = coef + row0 * 4;
#pragma unroll
    for (int e = 0; e < kRE; ++e) {
      rc1[e] = pc[4 * kRS * e];
      rc2[e] = pc[4 * kRS * e + 1];
      rum[e] = pu[e * rstride + mu0];
      rfn[e] = pf[e * rstride];
      if (diag) {
        rx[e] = pc[4 * kRS * e + 2];
        ry[e] = pc[4 * kRS * e + 3];
      } else {
        rx[e] = pu[e * rstride + nu0];
      }
    }
  };
  auto store = [&](int b) {
#pragma unroll
    for (int e = 0; e < kRE; ++e) {
      const int kk = rsub + kRS * e;
      const double un = diag ? rum[e] : rx[e];
      sA[b][kk][col] = rum[e];
      sB[b][kk][col] = rc1[e] * rfn[e] + 0.5 * rc2[e] * fyn * un;
      if (diag) fl += rx[e] * rfn[e] + ry[e] * fyn * un;
    }
  };
  if (kb < ke) {
    load(kb);
    store(0);
  }
  __syncthreads();
  int cur = 0;
  for (long k0 = kb; k0 < ke; k0 += WB_KC) {
    const bool more = k0 + WB_KC < ke;
    if (more) load(k0 + WB_KC);
#pragma unroll
    for (int kk = 0; kk < WB_KC; kk += 4) {
      double a[4], b[kNJ];
#pragma unroll
      for (int i = 0; i < 4; ++i) a[i] = sA[cur][kk + lg][64 * wm + 16 * i + lr];
#pragma unroll
      for (int j = 0; j < kNJ; ++j) b[j] = sB[cur][kk + lg][kWN * wn + 16 * j + lr];
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < kNJ; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i], b[j], acc[i][j], 0, 0, 0);
    }
    if (more) store(cur ^ 1);
    __syncthreads();
    cur ^= 1;
  }
  // write the partial tile
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < kNJ; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int mu = mu0 + 64 * wm + 16 * i + 4 * r + lg;
        const int nu = nu0 + kWN * wn + 16 * j + lr;
        // accumulate: add to the partial of earlier image chunks (each
        // element has one writer, chunks run in order)
        const double prev = accumulate ? zpart[(long)mu * np + nu] : 0.0;
        zpart[(long)mu * np + nu] = prev + acc[i][j][r];
      }
  if (diag) {
    sF[rsub][col] = fl;
    __syncthreads();
    if (tid < kTile) {
      double sum = 0.0;
#pragma unroll
      for (int q = 0; q < kRS; ++q) sum += sF[q][tid];
      const double prev = accumulate ? fpart[nu0 + tid] : 0.0;
      fpart[nu0 + tid] = prev + sum;
    }
  }
}

// sum the K-split partials (fixed order): zsum[mu, nu], f
__global__ void reduce_ksplit(const double* __restrict__ zpart_all,
                              const double* __restrict__ fpart_all, int np, int ksplit,
                              double* __restrict__ zsum_all, double* __restrict__ f_all,
                              double norm) {
  const int chain = blockIdx.y;
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long nn = (long)np * np;
  const double* zp = zpart_all + (long)chain * ksplit * nn;
  if (e < nn) {
    // eight loads in flight, summed in the same order as one at a time
    double s = 0.0;
    int k = 0;
    for (; k + 8 <= ksplit; k += 8) {
      double v[8];
#pragma unroll
      for (int j = 0; j < 8; ++j) v[j] = zp[(long)(k + j) * nn + e];
#pragma unroll
      for (int j = 0; j < 8; ++j) s += v[j];
    }
    for (; k < ksplit; ++k) s += zp[(long)k * nn + e];
    zsum_all[(long)chain * nn + e] = s;
  }
  if (e < np) {
    const double* fp = fpart_all + (long)chain * ksplit * np;
    double s = 0.0;
    for (int k = 0; k < ksplit; ++k) s += fp[(long)k * np + e];
    f_all[(long)chain * np + e] = s * norm;
  }
}

int launch_pa_fp64(Ctx* c, int nc, int apply_d4, hipStream_t s, hipEvent_t* ev) {
  const int ntile = c->np / kTile;
  // image range [row_lo, row_hi): whole 64-row tiles of the weights kernel
  // (rows outside the range get zero coefficients), and the outer GEMM's K
  // range from the 16-row step below row_lo to the step above row_hi
  const long t0 = c->row_lo / 64, t1 = (c->row_hi + 63) / 64;
  // tiles dealt 8 at a time (one per XCD), every chain of a tile on its XCD
  const long ga = round_up_l(t1 - t0, 8) * nc;
  if (ev) (void)hipEventRecord(ev[0], s);
  hipLaunchKernelGGL(weights_f64, dim3((unsigned)ga), dim3(256), 0, s, c->U, c->F, c->Y1, c->R1,
                     c->fy, c->fpsi, c->rho_img, c->np, c->Mp, apply_d4, t0, t1, c->row_lo,
                     c->row_hi, nc, c->coef);
  if (ev) (void)hipEventRecord(ev[1], s);
  const long kb = c->row_lo / WB_KC * WB_KC;
  const long ke = round_up_l(c->row_hi, WB_KC);
  const long rps = round_up_l((ke - kb + c->ksplit - 1) / c->ksplit, WB_KC);
  const long gb = (long)ntile * ntile * c->ksplit * nc;
  hipLaunchKernelGGL(outer_f64, dim3((unsigned)gb), dim3(kOT), 0, s, c->U, c->F, c->coef, c->fy,
                     c->np, c->Mp, c->ksplit, ntile, kb, ke, rps, 0, nc, c->zpart, c->fpart);
  if (ev) (void)hipEventRecord(ev[2], s);
  const long nn = (long)c->np * c->np;
  const double norm = 1.0 / (c->n_eff * c->n_syms);
  // zsum -> c->d2v_out (padded), normalised f -> c->f_out
  dim3 gr((unsigned)((nn + 255) / 256), nc);
  hipLaunchKernelGGL(reduce_ksplit, gr, dim3(256), 0, s, c->zpart, c->fpart, c->np, c->ksplit,
                     c->d2v_out, c->f_out, norm);
  if (ev) (void)hipEventRecord(ev[3], s);
  return 0;
}

}  // namespace tdscha
