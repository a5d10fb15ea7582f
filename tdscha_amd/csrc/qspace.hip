// q-space (Bloch basis) Lanczos on the device, FP64 complex.
//
// L application of Modules/QSpaceLanczos.py (apply_full_L, QL:870-893):
// harmonic part (QL:485-576) plus the perturbed averages of
// Modules/tdscha_qspace.jl:455-628 assembled into the a'/b' blocks
// (QL:718-758), for a batch of chains that each have
// their own perturbation q-point (q_pert) and hence their own pair list and
// psi layout (QL:297-343).  The Hermitian Lanczos with full
// reorthogonalisation (QL:967-1182) runs on the device for the whole batch,
// every chain stopping at its own cap or at |b| < tol.
//
// Layouts on the device:
//   X, Y     double2 [n_q][nb][N]   Bloch components, configurations fastest
//   psi      double2 [C][ld]        chain c's vector in the reference layout,
//                                   zero-padded to the common stride ld
// Kernels of one L application:
//   q_unpack   alpha1 blocks (QL:662-716) and R1 per chain and pair
//   q_weights  per configuration and chain: buffer_u = A x_partner summed
//              into total_sum and buf_f_weight, the D3/D4 weights and the
//              f_pert contribution (JQ:489-573)
//   q_d2v      per chain, pair and configuration chunk: the d2v block as a
//              sum of two rank-1 complex products per configuration
//              (JQ:575-610)
//   q_finish   fixed-order reduction of the partials, normalisation and
//              assembly of harmonic + anharmonic parts
// Point-group images (the per-(configuration, symmetry) rotation
// x_rot = S_j x of JQ:496-502) are expanded once per context by q_expand:
// the ensemble then holds n_sym N configurations, normalised by
// n_sym N_eff (JQ:621), and every kernel above runs unchanged.
// All sums are fixed-order trees or loops: results do not depend on timing.
#include <algorithm>
#include <array>
#include <cmath>
#include <vector>

#include "common.hpp"
#include "../../include/tdscha_amd.h"

namespace tdscha {

constexpr int kQMaxChains = 256;
constexpr int kQMaxQ = 512;
constexpr int kQWeightsChains = 2;   // chains per q_weights workgroup
constexpr int kD2vChains = 16;       // chains per slab q_d2v workgroup
constexpr int kSlabBytes = 65536;    // LDS of one slab q_d2v workgroup (two buffers)
constexpr int kD2vThreads = 256;     // threads (items) per slab q_d2v workgroup

struct QCtx {
  int nq = 0, nb = 0, N = 0;
  int n_sym = 1;            // point-group images expanded so far (1: none)
  double T = 0.0, neff = 0.0;
  double2* X = nullptr;     // [nq][nb][N]
  double2* Y = nullptr;
  double* rho = nullptr;    // [N]
  double* wq = nullptr;     // [nq][nb]
  double* nbose = nullptr;  // [nq][nb]
  int* valid = nullptr;     // [nq][nb]
  double* fY = nullptr;     // [nq][nb]
  double* fP = nullptr;     // [nq][nb]
  double2* S = nullptr;     // slab images of (fY X, Y) for q_d2v_slab
  int slabT = 0;            // configurations per slab (0: no slab images)
  int d2v_form = 1;         // 1: q_d2v_slab when slab images exist, 0: q_d2v
  int n_cu = 256;           // compute units of the context's device
  // tile choices (autotuner)
  int weights_tile = 128;   // configurations per workgroup of q_weights
  int d2v_chunk = 0;        // configurations per workgroup of q_d2v (0 = automatic)
  int reorth_jtile = 16;    // basis vectors per workgroup of the reorthogonalisation dots
  hipStream_t stream = nullptr;
};

// Chain tables and work buffers of one batch.
struct QBatch {
  int C = 0, MP = 0, maxp = 0, ld = 0, nch = 1, ntile = 1;
  std::vector<int> h_len, h_np;
  int* iqp = nullptr;       // [C]
  int* npairs = nullptr;    // [C]
  int* pq1 = nullptr;       // [C][MP]
  int* pq2 = nullptr;
  int* offa = nullptr;
  int* offb = nullptr;
  int* partner = nullptr;   // [C][nq]
  int* apair = nullptr;     // [C][nq]
  int* aorient = nullptr;   // [C][nq]
  int* active = nullptr;    // [C]
  double2* alpha = nullptr; // [C][nq][nb*nb] oriented alpha1 block of each q
  double2* R1 = nullptr;    // [C][nb]
  double2* wc = nullptr;    // [C][N]
  double2* wd = nullptr;
  double2* fpart = nullptr; // [C][ntile][nb]
  double2* dpart = nullptr; // [C][MP][nch][nb*nb]
  // slab form of q_d2v: (chain, pair, q1, q2) items sorted by chain, cut into
  // workgroups of at most kD2vThreads items and kD2vChains chains
  int slabT = 0, chunk = 0, nib = 0, n_items = 0;
  int4* items = nullptr;    // [n_items]
  int* blk_start = nullptr; // [nib + 1]
};

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cmulcj(double2 a, double2 b) {   // a * conj(b)
  return make_double2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
__device__ __forceinline__ double2 cj(double2 a) { return make_double2(a.x, -a.y); }
__device__ __forceinline__ double2 cadd(double2 a, double2 b) {
  return make_double2(a.x + b.x, a.y + b.y);
}
__device__ __forceinline__ double2 csc(double2 a, double s) { return make_double2(a.x * s, a.y * s); }
// acc + a b and acc + a conj(b) as four FMAs
__device__ __forceinline__ double2 cmac(double2 acc, double2 a, double2 b) {
  acc.x = __builtin_fma(a.x, b.x, acc.x);
  acc.x = __builtin_fma(-a.y, b.y, acc.x);
  acc.y = __builtin_fma(a.x, b.y, acc.y);
  acc.y = __builtin_fma(a.y, b.x, acc.y);
  return acc;
}
__device__ __forceinline__ double2 cmacj(double2 acc, double2 a, double2 b) {
  acc.x = __builtin_fma(a.x, b.x, acc.x);
  acc.x = __builtin_fma(a.y, b.y, acc.x);
  acc.y = __builtin_fma(a.y, b.x, acc.y);
  acc.y = __builtin_fma(-a.x, b.y, acc.y);
  return acc;
}

__device__ __forceinline__ int tri_index(int i, int j, int nb) {   // i <= j
  return i * nb - i * (i - 1) / 2 + (j - i);
}

// entry (i, j) of the a' or b' block at offset off: full row-major, or the
// Hermitian fill of the row-major upper triangle (QL:345-405)
__device__ __forceinline__ double2 block_entry(const double2* psi, int off, int full, int i, int j,
                                               int nb) {
  if (full) return psi[off + i * nb + j];
  if (i <= j) return psi[off + tri_index(i, j, nb)];
  return cj(psi[off + tri_index(j, i, nb)]);
}

// chi^-, chi^+ and w1 w2 / X of one band pair (QL:604-716), zero where a
// band is masked
struct PairFac {
  double chim, chip, w2x, vm;
};
__device__ __forceinline__ PairFac pair_factors(double w1, double w2, double n1, double n2,
                                                bool v1, bool v2) {
  PairFac f;
  const bool vm = v1 && v2;
  f.vm = vm ? 1.0 : 0.0;
  if (!vm) {
    f.chim = f.chip = f.w2x = 0.0;
    return f;
  }
  const double den = 2.0 * w1 * w2;
  f.chim = (w1 - w2) * (n1 - n2) / den;
  f.chip = (w1 + w2) * (1.0 + n1 + n2) / den;
  const double X = (1.0 + 2.0 * n1) * (1.0 + 2.0 * n2) / 8.0;
  f.w2x = w1 * w2 / X;
  return f;
}

// alpha1 block of the pair that q belongs to, oriented for q (JQ:519-526):
// alq[c][q][v1][v2] = alpha[v1][v2] when q is the pair's first q-point,
// conj(alpha[v2][v1]) when it is the second.  One workgroup per (q, chain).
template <int NB>
__global__ void q_unpack(const double2* __restrict__ psi, int ld, const int* __restrict__ active,
                         const int* __restrict__ pq1, const int* __restrict__ pq2,
                         const int* __restrict__ offa, const int* __restrict__ offb, int MP,
                         int nq, const int* __restrict__ apair, const int* __restrict__ aorient,
                         const double* __restrict__ wq, const double* __restrict__ nbose,
                         const int* __restrict__ valid, int ignore_v3,
                         double2* __restrict__ alq, double2* __restrict__ R1) {
  const int c = blockIdx.y, q = blockIdx.x, t = threadIdx.x;
  if (!active[c]) return;
  const double2* ps = psi + (long)c * ld;
  if (q == 0 && t < NB) R1[c * NB + t] = ignore_v3 ? make_double2(0.0, 0.0) : ps[t];
  if (t >= NB * NB) return;
  const int p = apair[c * nq + q], o = aorient[c * nq + q];
  const int i = o ? t % NB : t / NB, j = o ? t / NB : t % NB;
  const int q1 = pq1[c * MP + p], q2 = pq2[c * MP + p];
  const int full = q1 < q2;
  const double2 a = block_entry(ps, offa[c * MP + p], full, i, j, NB);
  const double2 b = block_entry(ps, offb[c * MP + p], full, i, j, NB);
  const PairFac f = pair_factors(wq[q1 * NB + i], wq[q2 * NB + j], nbose[q1 * NB + i],
                                 nbose[q2 * NB + j], valid[q1 * NB + i], valid[q2 * NB + j]);
  const double ca = f.w2x * sqrt(-0.5 * f.chim), cb = f.w2x * sqrt(0.5 * f.chip);
  const double2 al = make_double2(ca * a.x - cb * b.x, ca * a.y - cb * b.y);
  alq[((long)c * nq + q) * NB * NB + t] = o ? cj(al) : al;
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Workgroup b of a tiled launch runs on XCD b % 8.  This maps b to (tile,
// group) so that every group of one tile runs on the same XCD, tiles
// t = 8 s + x going to XCD x in order: the groups of a tile then read its
// data through one L2.  Returns false for the padding workgroups (tile >=
// ntile) of the last round.
__device__ __forceinline__ bool xcd_tile(int b, int ngroups, int ntile, int& tile, int& group) {
  const int s = b >> 3;
  group = s % ngroups;
  tile = (s / ngroups) * 8 + (b & 7);
  return tile < ntile;
}

// One thread per configuration and CB chains per workgroup: the own-q
// vectors x_q and fP_q y_q are loaded once for the CB chains, each chain
// then loads its partner x_{q_pert - q} and applies its oriented alpha block,
// which is the same for every thread (scalar loads).
template <int NB, int CB>
__global__ void __launch_bounds__(256) q_weights(
    const double2* __restrict__ X, const double2* __restrict__ Y, const double* __restrict__ rho,
    int N, int nq, const double* __restrict__ fY, const double* __restrict__ fP,
    const int* __restrict__ active, const int* __restrict__ iqp,
    const int* __restrict__ partner, const double2* __restrict__ alq,
    const double2* __restrict__ R1, int C, int ngroups, int ntile, int apply_v4,
    double2* __restrict__ wc, double2* __restrict__ wd, double2* __restrict__ fpart) {
  __shared__ double red[4][CB][2 * NB];
  int tile, g;
  if (!xcd_tile(blockIdx.x, ngroups, ntile, tile, g)) return;
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
  const int c0 = g * CB;
  unsigned act = 0;
#pragma unroll
  for (int k = 0; k < CB; ++k)
    if (c0 + k < C && active[c0 + k]) act |= 1u << k;
  if (!act) return;
  const int i = tile * blockDim.x + t;
  const bool in = i < N;
  double2 tot[CB], bfw[CB];
#pragma unroll
  for (int k = 0; k < CB; ++k) tot[k] = bfw[k] = make_double2(0.0, 0.0);
  if (in) {
    for (int q = 0; q < nq; ++q) {
      double2 xq[NB], yq[NB];
#pragma unroll
      for (int v = 0; v < NB; ++v) {
        xq[v] = X[((long)q * NB + v) * N + i];
        yq[v] = csc(Y[((long)q * NB + v) * N + i], fP[q * NB + v]);
      }
#pragma unroll
      for (int k = 0; k < CB; ++k) {
        if (!((act >> k) & 1u)) continue;
        const int c = c0 + k;
        const int q2 = partner[c * nq + q];
        const double2* A = alq + ((long)c * nq + q) * NB * NB;
        double2 x2[NB];
#pragma unroll
        for (int v = 0; v < NB; ++v) x2[v] = X[((long)q2 * NB + v) * N + i];
#pragma unroll
        for (int v1 = 0; v1 < NB; ++v1) {
          double2 bu = cmul(A[v1 * NB], x2[0]);
#pragma unroll
          for (int v2 = 1; v2 < NB; ++v2) bu = cmac(bu, A[v1 * NB + v2], x2[v2]);
          tot[k] = cmacj(tot[k], bu, xq[v1]);      // conj(x) bu
          bfw[k] = cmacj(bfw[k], yq[v1], bu);      // conj(bu) fP y
        }
      }
    }
  }
  const double r = in ? rho[i] : 0.0;
#pragma unroll
  for (int k = 0; k < CB; ++k) {
    if (!((act >> k) & 1u)) continue;
    const int c = c0 + k;
    const int qp = iqp[c];
    double2 fc[NB];
#pragma unroll
    for (int v = 0; v < NB; ++v) fc[v] = make_double2(0.0, 0.0);
    if (in) {
      double2 wR = make_double2(0.0, 0.0), wRf = make_double2(0.0, 0.0);
      double2 xp[NB], yp[NB];
#pragma unroll
      for (int v = 0; v < NB; ++v) {
        xp[v] = X[((long)qp * NB + v) * N + i];
        yp[v] = Y[((long)qp * NB + v) * N + i];
        const double2 r1 = R1[c * NB + v];
        wR = cadd(wR, csc(cmulcj(xp[v], r1), fY[qp * NB + v]));
        wRf = cadd(wRf, cmulcj(yp[v], r1));
      }
      wR = csc(wR, r / 3.0);
      wRf = csc(wRf, r / 3.0);
      const double2 w1 = csc(tot[k], -0.5 * r / 3.0);
      const double2 w2 = csc(bfw[k], -r / 3.0);
#pragma unroll
      for (int v = 0; v < NB; ++v)
        fc[v] = cadd(cmul(w1, yp[v]), csc(cmul(w2, xp[v]), fY[qp * NB + v]));
      double2 c1 = wR, c2 = wRf;
      if (apply_v4) {
        c1 = cadd(c1, csc(tot[k], -r / 8.0));
        c2 = cadd(c2, csc(bfw[k], -r / 4.0));
      }
      wc[(long)c * N + i] = c1;
      wd[(long)c * N + i] = c2;
    }
    // fixed-order reduction of the f_pert contributions: waves, then the
    // workgroup's waves in order
#pragma unroll
    for (int v = 0; v < NB; ++v) {
      const double sx = wave_sum(fc[v].x), sy = wave_sum(fc[v].y);
      if (lane == 0) {
        red[wv][k][2 * v] = sx;
        red[wv][k][2 * v + 1] = sy;
      }
    }
  }
  __syncthreads();
  if (t < CB * NB) {
    const int k = t / NB, v = t % NB;
    if ((act >> k) & 1u) {
      double sx = 0.0, sy = 0.0;
      for (int w = 0; w < (int)(blockDim.x >> 6); ++w) {
        sx += red[w][k][2 * v];
        sy += red[w][k][2 * v + 1];
      }
      fpart[((long)(c0 + k) * ntile + tile) * NB + v] = make_double2(sx, sy);
    }
  }
}

// grid (largest pair count, chains, configuration chunks); the 36 complex
// accumulators take 144 VGPRs, the cap of two waves per SIMD leaves the rest
// for the loads of one configuration
template <int NB>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) q_d2v(
    const double2* __restrict__ X, const double2* __restrict__ Y, int N,
    const double* __restrict__ fY, const int* __restrict__ active,
    const int* __restrict__ npairs, const int* __restrict__ pq1, const int* __restrict__ pq2,
    int MP, const double2* __restrict__ wc, const double2* __restrict__ wd, int chunk, int nch,
    double2* __restrict__ dpart) {
  const int p = blockIdx.x, c = blockIdx.y, ch = blockIdx.z, t = threadIdx.x;
  if (!active[c] || p >= npairs[c]) return;
  __shared__ double red[4][2 * NB * NB];
  const int q1 = pq1[c * MP + p], q2 = pq2[c * MP + p];
  double f1[NB], f2[NB];
#pragma unroll
  for (int v = 0; v < NB; ++v) {
    f1[v] = fY[q1 * NB + v];
    f2[v] = fY[q2 * NB + v];
  }
  double2 acc[NB][NB];
#pragma unroll
  for (int a = 0; a < NB; ++a)
#pragma unroll
    for (int b = 0; b < NB; ++b) acc[a][b] = make_double2(0.0, 0.0);
  const int lo = ch * chunk;
  const int hi = min(N, lo + chunk);
  for (int i = lo + t; i < hi; i += blockDim.x) {
    const double2 wci = wc[(long)c * N + i], wdi = wd[(long)c * N + i];
    double2 r11[NB], r21[NB], Q1[NB], Q2[NB];
#pragma unroll
    for (int v = 0; v < NB; ++v) {
      r11[v] = csc(X[((long)q1 * NB + v) * N + i], f1[v]);
      r21[v] = Y[((long)q1 * NB + v) * N + i];
      const double2 r12 = csc(X[((long)q2 * NB + v) * N + i], f2[v]);
      const double2 r22 = Y[((long)q2 * NB + v) * N + i];
      // -Q1 = conj(wc) r22 + conj(wd) r12 ; -Q2 = conj(wc) r12
      Q1[v] = cmac(cmul(cj(wci), r22), cj(wdi), r12);
      Q2[v] = cmul(cj(wci), r12);
    }
    // accumulates -(r11 Q1^H + r21 Q2^H); the sign is restored below
#pragma unroll
    for (int a = 0; a < NB; ++a)
#pragma unroll
      for (int b = 0; b < NB; ++b)
        acc[a][b] = cmacj(cmacj(acc[a][b], r11[a], Q1[b]), r21[a], Q2[b]);
  }
  const int lane = t & 63, wv = t >> 6;
#pragma unroll
  for (int a = 0; a < NB; ++a)
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      const double sx = -wave_sum(acc[a][b].x), sy = -wave_sum(acc[a][b].y);
      if (lane == 0) {
        red[wv][2 * (a * NB + b)] = sx;
        red[wv][2 * (a * NB + b) + 1] = sy;
      }
    }
  __syncthreads();
  if (t < NB * NB) {
    const int nw = blockDim.x >> 6;
    double sx = 0.0, sy = 0.0;
    for (int w = 0; w < nw; ++w) {
      sx += red[w][2 * t];
      sy += red[w][2 * t + 1];
    }
    dpart[(((long)c * MP + p) * nch + ch) * NB * NB + t] = make_double2(sx, sy);
  }
}

// Slab images: the configurations [s T, s T + T) of slab s as one
// contiguous block of U = 2 nq SR entries, SR = NB T + 1: entry
// (h nq + q) SR + v T + j holds fY[q][v] x_{q,v} (h = 0) or y_{q,v} (h = 1)
// of configuration s T + j, zero past N; the last entry of each row is a
// zero pad that puts the rows of neighbouring q-points on different LDS
// banks.  One thread per entry.
template <int NB>
__global__ void q_slab_build(const double2* __restrict__ X, const double2* __restrict__ Y,
                             const double* __restrict__ fY, int N, int nq, int lgT, long total,
                             double2* __restrict__ S) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= total) return;
  const int T = 1 << lgT, SR = NB * T + 1;
  const long U = 2L * nq * SR;
  const long sl = e / U;
  const int u = (int)(e - sl * U);
  const int hq = u / SR, w = u - hq * SR;
  const int h = hq >= nq, q = h ? hq - nq : hq;
  double2 val = make_double2(0.0, 0.0);
  if (w < NB * T) {
    const int v = w >> lgT, j = w & (T - 1);
    const long cfg = sl * T + j;
    if (cfg < N) {
      const long r = (long)q * NB + v;
      val = h ? Y[r * N + cfg] : csc(X[r * N + cfg], fY[r]);
    }
  }
  S[e] = val;
}

// d2v blocks from slab images in LDS.  Workgroup (item block ib,
// configuration chunk ch) copies the slab images of its chunk, and the
// weights wc, wd of its chains for the slab's configurations, into LDS with direct global -> LDS loads into two
// buffers in turn: the copy of slab k + 1 runs while the threads accumulate
// slab k.  The two buffers are distinct arrays and each half of the unrolled
// loop names its own, so that the compiler sees that the reads of one do not
// depend on the copies into the other.  Each thread accumulates the 6x6
// block of one (chain, pair) item, in configuration order.  The item blocks
// of one chunk run on one XCD and share the chunk through its L2.
constexpr int kSlabUnits = kSlabBytes / 2 / (int)sizeof(double2);

template <int NB>
__global__ void __launch_bounds__(kD2vThreads) __attribute__((amdgpu_waves_per_eu(2))) q_d2v_slab(
    const double2* __restrict__ S, int N, int nq, const int* __restrict__ active,
    const int4* __restrict__ items, const int* __restrict__ blk_start, int nib, int MP,
    const double2* __restrict__ wc, const double2* __restrict__ wd, int lgT, int chunk, int nch,
    double2* __restrict__ dpart) {
  __shared__ double2 bufA[kSlabUnits], bufB[kSlabUnits];
  int ch, ib;
  if (!xcd_tile(blockIdx.x, nib, nch, ch, ib)) return;
  const int T = 1 << lgT, SR = NB * T + 1, U = 2 * nq * SR;
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6, nwave = blockDim.x >> 6;
  const int i0 = blk_start[ib], i1 = blk_start[ib + 1];
  const int c_lo = items[i0].x, ncb = items[i1 - 1].x - c_lo + 1;
  bool any = false;
  for (int k = 0; k < ncb; ++k) any |= active[c_lo + k] != 0;
  if (!any) return;
  int4 it = make_int4(0, 0, 0, 0);
  bool mine = i0 + t < i1;
  if (mine) {
    it = items[i0 + t];
    mine = active[it.x] != 0;
  }
  const int cw = (it.x - c_lo) << (lgT + 1);
  const int nW = ncb << (lgT + 1);
  const int lo = ch * chunk, hi = min(N, lo + chunk);
  auto issue = [&](int s0, double2* dst) {
    const long sl = s0 >> lgT;
    const double2* src = S + sl * U;
    for (int k = wave; k * 64 < U; k += nwave)
      if (k * 64 + lane < U) dma16(src + k * 64 + lane, dst + k * 64);
    // weights: entry (chain k, c1 | c2, configuration j) at (2 k + h) T + j;
    // past N any finite value (the image is zero there)
    for (int k = wave; k * 64 < nW; k += nwave) {
      const int e = k * 64 + lane;
      if (e < nW) {
        const int j = e & (T - 1), h = (e >> lgT) & 1, c = c_lo + (e >> (lgT + 1));
        dma16((h ? wd : wc) + (long)c * N + min(s0 + j, N - 1), dst + U + k * 64);
      }
    }
  };
  double2 acc[NB][NB];
#pragma unroll
  for (int a = 0; a < NB; ++a)
#pragma unroll
    for (int b = 0; b < NB; ++b) acc[a][b] = make_double2(0.0, 0.0);
  auto accumulate = [&](const double2* img) {
    if (!mine) return;
    for (int j = 0; j < T; ++j) {
      const double2 wci = img[U + cw + j], wdi = img[U + cw + T + j];
      const double2* x1 = img + it.z * SR + j;
      const double2* y1 = img + (nq + it.z) * SR + j;
      const double2* x2 = img + it.w * SR + j;
      const double2* y2 = img + (nq + it.w) * SR + j;
      double2 Q1[NB], Q2[NB];
#pragma unroll
      for (int v = 0; v < NB; ++v) {
        // -Q1 = conj(wc) r22 + conj(wd) r12, -Q2 = conj(wc) r12 with
        // r12 = fY x2 (the image holds fY x) and r22 = y2
        const double2 r12 = x2[v * T], r22 = y2[v * T];
        Q1[v] = cmac(cmul(cj(wci), r22), cj(wdi), r12);
        Q2[v] = cmul(cj(wci), r12);
      }
#pragma unroll
      for (int a = 0; a < NB; ++a) {
        const double2 r11 = x1[a * T], r21 = y1[a * T];
#pragma unroll
        for (int b = 0; b < NB; ++b) acc[a][b] = cmacj(cmacj(acc[a][b], r11, Q1[b]), r21, Q2[b]);
      }
    }
  };
  issue(lo, bufA);
  for (int s0 = lo; s0 < hi; s0 += 2 * T) {
    TDS_WAIT_VMCNT(0);
    lds_barrier();                     // slab s0 in bufA; bufB free
    if (s0 + T < hi) issue(s0 + T, bufB);
    accumulate(bufA);
    if (s0 + T >= hi) break;
    TDS_WAIT_VMCNT(0);
    lds_barrier();                     // slab s0 + T in bufB; bufA free
    if (s0 + 2 * T < hi) issue(s0 + 2 * T, bufA);
    accumulate(bufB);
  }
  if (mine) {
    double2* o = dpart + (((long)it.x * MP + it.y) * nch + ch) * NB * NB;
#pragma unroll
    for (int a = 0; a < NB; ++a)
#pragma unroll
      for (int b = 0; b < NB; ++b) o[a * NB + b] = make_double2(-acc[a][b].x, -acc[a][b].y);
  }
}

// Static Liouvillian of the free-energy Hessian (QSpaceHessian,
// Modules/QSpaceHessian.py:605-692): psi = [R, one W sector].  The input of
// the perturbed averages is alpha = -2 Y_w1 Y_w2 W (QSH:675-679; Y_w = fY),
// oriented per q-point as in q_unpack.
template <int NB>
__global__ void q_unpack_static(const double2* __restrict__ psi, int ld,
                                const int* __restrict__ active, const int* __restrict__ pq1,
                                const int* __restrict__ pq2, const int* __restrict__ offa, int MP,
                                int nq, const int* __restrict__ apair,
                                const int* __restrict__ aorient, const double* __restrict__ fY,
                                int ignore_v3, double2* __restrict__ alq,
                                double2* __restrict__ R1) {
  const int c = blockIdx.y, q = blockIdx.x, t = threadIdx.x;
  if (!active[c]) return;
  const double2* ps = psi + (long)c * ld;
  if (q == 0 && t < NB) R1[c * NB + t] = ignore_v3 ? make_double2(0.0, 0.0) : ps[t];
  if (t >= NB * NB) return;
  const int p = apair[c * nq + q], o = aorient[c * nq + q];
  const int i = o ? t % NB : t / NB, j = o ? t / NB : t % NB;
  const int q1 = pq1[c * MP + p], q2 = pq2[c * MP + p];
  const double2 w = block_entry(ps, offa[c * MP + p], q1 < q2, i, j, NB);
  const double2 al = csc(w, -2.0 * fY[q1 * NB + i] * fY[q2 * NB + j]);
  alq[((long)c * nq + q) * NB * NB + t] = o ? cj(al) : al;
}

// Static two-phonon propagator (QSH:481-544): Lambda = ((n1 + n2 + 1) /
// (w1 + w2) - dn) / (4 w1 w2), dn = (n1 - n2) / (w1 - w2), or
// -beta e^{beta w} n^2 when |w1 - w2| < 1e-8; zero unless both modes count.
__device__ __forceinline__ double static_lambda(double w1, double w2, double n1, double n2,
                                                bool v1, bool v2, double beta) {
  if (!(v1 && v2)) return 0.0;
  double dn = 0.0;
  if (beta > 0.0) {
    if (fabs(w1 - w2) < 1e-8)
      dn = -beta * exp(w1 * beta) * n1 * n1;
    else
      dn = (n1 - n2) / (w1 - w2);
  }
  return ((n1 + n2 + 1.0) / (w1 + w2) - dn) / (4.0 * w1 * w2);
}

// R <- w^2 R - f_pert, W <- W / Lambda + d2v (QSH:605-692).  Block 0: R
// sector; block 1 + p: the W block of pair p.
template <int NB>
__global__ void q_finish_static(const double2* __restrict__ psi, int ld, double2* __restrict__ out,
                                const int* __restrict__ active, const int* __restrict__ iqp,
                                const int* __restrict__ npairs, const int* __restrict__ pq1,
                                const int* __restrict__ pq2, const int* __restrict__ offa, int MP,
                                const double* __restrict__ wq, const double* __restrict__ nbose,
                                const int* __restrict__ valid, double beta,
                                const double2* __restrict__ fpart, int ntile,
                                const double2* __restrict__ dpart, int nch, double inv_norm,
                                int do_harm, int do_anh) {
  const int c = blockIdx.y, blk = blockIdx.x, t = threadIdx.x;
  if (!active[c]) return;
  const double2* ps = psi + (long)c * ld;
  double2* po = out + (long)c * ld;
  if (blk == 0) {
    if (t >= NB) return;
    const int qp = iqp[c];
    double2 o = make_double2(0.0, 0.0);
    if (do_anh) {
      double2 f = make_double2(0.0, 0.0);
      for (int k = 0; k < ntile; ++k) f = cadd(f, fpart[((long)c * ntile + k) * NB + t]);
      o = csc(f, -inv_norm);
    }
    if (do_harm) {
      const double w = wq[qp * NB + t];
      o = cadd(o, csc(ps[t], w * w));
    }
    po[t] = o;
    return;
  }
  const int p = blk - 1;
  if (p >= npairs[c] || t >= NB * NB) return;
  const int i = t / NB, j = t % NB;
  const int q1 = pq1[c * MP + p], q2 = pq2[c * MP + p];
  const int full = q1 < q2;
  if (!full && i > j) return;
  const int e = offa[c * MP + p] + (full ? i * NB + j : tri_index(i, j, NB));
  double2 o = make_double2(0.0, 0.0);
  if (do_anh) {
    const double2* dp = dpart + ((long)c * MP + p) * nch * NB * NB + t;
    for (int k = 0; k < nch; ++k) o = cadd(o, dp[(long)k * NB * NB]);
    o = csc(o, inv_norm);
  }
  if (do_harm) {
    const double lam = static_lambda(wq[q1 * NB + i], wq[q2 * NB + j], nbose[q1 * NB + i],
                                     nbose[q2 * NB + j], valid[q1 * NB + i], valid[q2 * NB + j],
                                     beta);
    if (fabs(lam) > 1e-30) o = cadd(o, csc(ps[e], 1.0 / lam));
  }
  po[e] = o;
}

// block 0: R sector; block 1 + p: pair p
template <int NB>
__global__ void q_finish(const double2* __restrict__ psi, int ld, double2* __restrict__ out,
                         const int* __restrict__ active, const int* __restrict__ iqp,
                         const int* __restrict__ npairs, const int* __restrict__ pq1,
                         const int* __restrict__ pq2, const int* __restrict__ offa,
                         const int* __restrict__ offb, int MP, const double* __restrict__ wq,
                         const double* __restrict__ nbose, const int* __restrict__ valid,
                         const double2* __restrict__ fpart, int ntile,
                         const double2* __restrict__ dpart, int nch, double inv_norm,
                         int do_harm, int do_anh) {
  const int c = blockIdx.y, blk = blockIdx.x, t = threadIdx.x;
  if (!active[c]) return;
  const double2* ps = psi + (long)c * ld;
  double2* po = out + (long)c * ld;
  if (blk == 0) {
    if (t >= NB) return;
    const int qp = iqp[c];
    double2 f = make_double2(0.0, 0.0);
    for (int k = 0; k < ntile; ++k) f = cadd(f, fpart[((long)c * ntile + k) * NB + t]);
    double2 o = make_double2(0.0, 0.0);
    if (do_anh) o = csc(f, inv_norm);
    if (do_harm && valid[qp * NB + t]) {
      const double w = wq[qp * NB + t];
      o = cadd(o, csc(ps[t], -w * w));
    }
    po[t] = o;
    return;
  }
  const int p = blk - 1;
  if (p >= npairs[c] || t >= NB * NB) return;
  const int i = t / NB, j = t % NB;
  const int q1 = pq1[c * MP + p], q2 = pq2[c * MP + p];
  const int full = q1 < q2;
  if (!full && i > j) return;
  const int e = full ? i * NB + j : tri_index(i, j, NB);
  const int ea = offa[c * MP + p] + e, eb = offb[c * MP + p] + e;
  const double w1 = wq[q1 * NB + i], w2 = wq[q2 * NB + j];
  const PairFac f = pair_factors(w1, w2, nbose[q1 * NB + i], nbose[q2 * NB + j],
                                 valid[q1 * NB + i], valid[q2 * NB + j]);
  double2 oa = make_double2(0.0, 0.0), ob = make_double2(0.0, 0.0);
  if (do_anh) {
    double2 d = make_double2(0.0, 0.0);
    const double2* dp = dpart + ((long)c * MP + p) * nch * NB * NB + t;
    for (int k = 0; k < nch; ++k) d = cadd(d, dp[(long)k * NB * NB]);
    d = csc(d, inv_norm);
    oa = csc(d, sqrt(-0.5 * f.chim));
    ob = csc(d, -sqrt(0.5 * f.chip));
  }
  if (do_harm) {
    oa = cadd(oa, csc(ps[ea], -(w1 - w2) * (w1 - w2) * f.vm));
    ob = cadd(ob, csc(ps[eb], -(w1 + w2) * (w1 + w2) * f.vm));
  }
  po[ea] = oa;
  po[eb] = ob;
}

// Images of every configuration under every symmetry: image j N + i of
// q-point perm[j][p] is D[j][p] x_i[p] (JQ:496-502).  Grid
// (ceil(N / 256), n_sym, n_q), configurations fastest on both sides.
template <int NB>
__global__ void q_expand(const double2* __restrict__ X, const double2* __restrict__ Y, int N,
                         int nq, const int* __restrict__ perm, const double2* __restrict__ D,
                         double2* __restrict__ XE, double2* __restrict__ YE) {
  const int p = blockIdx.z, j = blockIdx.y;
  __shared__ double2 d[NB * NB];
  for (int e = threadIdx.x; e < NB * NB; e += blockDim.x)
    d[e] = D[((long)j * nq + p) * NB * NB + e];
  __syncthreads();
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const int q = perm[(long)j * nq + p];
  const long NE = (long)N * gridDim.y;
  double2 x[NB], y[NB];
#pragma unroll
  for (int m = 0; m < NB; ++m) {
    x[m] = X[((long)p * NB + m) * N + i];
    y[m] = Y[((long)p * NB + m) * N + i];
  }
#pragma unroll
  for (int v = 0; v < NB; ++v) {
    double2 sx = make_double2(0.0, 0.0), sy = make_double2(0.0, 0.0);
#pragma unroll
    for (int m = 0; m < NB; ++m) {
      sx = cadd(sx, cmul(d[v * NB + m], x[m]));
      sy = cadd(sy, cmul(d[v * NB + m], y[m]));
    }
    XE[((long)q * NB + v) * NE + (long)j * N + i] = sx;
    YE[((long)q * NB + v) * NE + (long)j * N + i] = sy;
  }
}

// ------------------------------------------------------------ batch setup
static void free_batch(QBatch& b) {
  void* p[] = {b.iqp,   b.npairs, b.pq1, b.pq2, b.offa, b.offb,  b.partner, b.apair,
               b.aorient, b.active, b.alpha, b.R1, b.wc,  b.wd,   b.fpart, b.dpart,
               b.items, b.blk_start};
  for (void* x : p)
    if (x) (void)hipFree(x);
  b = QBatch();
}

static int choose_chunk(const QCtx* q, int C, int MP_used) {
  if (q->d2v_chunk > 0) return q->d2v_chunk;
  // aim at ~2048 workgroups in total, at least 1024 configurations each
  long blocks = (long)C * std::max(1, MP_used);
  long nch = std::max(1L, 2048 / std::max(1L, blocks));
  int chunk = (int)((q->N + nch - 1) / nch);
  return std::max(chunk, std::min(q->N, 1024));
}

// Configurations per slab of the slab q_d2v: the largest power of two up to
// 16 whose slab image and the weights of kD2vChains chains fit one of its
// two LDS buffers (kSlabBytes / 2); 0 when one configuration does not fit.
static size_t slab_units(int nq, int nb, int T) {
  return 2 * (size_t)nq * (nb * T + 1) + 2 * (size_t)kD2vChains * T;
}

static int log2_pow2(int T) {
  int lg = 0;
  while ((1 << lg) < T) ++lg;
  return lg;
}

static int slab_configs(int nq, int nb) {
  int T = 16;
  while (T > 0 && slab_units(nq, nb, T) * sizeof(double2) > (size_t)kSlabBytes / 2) T >>= 1;
  return T;
}

// (Re)builds the slab images of the context's ensemble; without room for
// one configuration per slab the context keeps none and q_d2v runs instead.
static int build_slabs(QCtx* q) {
  if (q->S) (void)hipFree(q->S);
  q->S = nullptr;
  q->slabT = slab_configs(q->nq, q->nb);
  if (!q->slabT) return 0;
  const int lg = log2_pow2(q->slabT);
  const long nslab = (q->N + q->slabT - 1) / q->slabT;
  const long total = nslab * 2L * q->nq * (q->nb * q->slabT + 1);
  if (hipMalloc((void**)&q->S, sizeof(double2) * (size_t)total) != hipSuccess) {
    q->S = nullptr;
    set_error("q-space: allocation of the slab images failed");
    return -1;
  }
  const unsigned nblk = (unsigned)((total + 255) / 256);
  if (q->nb == 6)
    hipLaunchKernelGGL(q_slab_build<6>, dim3(nblk), dim3(256), 0, q->stream, q->X, q->Y, q->fY,
                       q->N, q->nq, lg, total, q->S);
  else
    hipLaunchKernelGGL(q_slab_build<3>, dim3(nblk), dim3(256), 0, q->stream, q->X, q->Y, q->fY,
                       q->N, q->nq, lg, total, q->S);
  if (hipGetLastError() != hipSuccess || hipStreamSynchronize(q->stream) != hipSuccess) {
    set_error("q-space: slab image kernel failed");
    return -1;
  }
  return 0;
}

// d2v chunk (configurations per workgroup) that gives `rounds` rounds of
// slab workgroups on the device (n_item_blocks workgroups per chunk, two
// resident workgroups per CU: two buffers of half the CU's LDS and 256
// threads at up to 256 VGPRs), at least 8 slabs.
static int slab_chunk(const QCtx* q, int n_item_blocks, int rounds) {
  const long slots = 2L * q->n_cu * rounds;
  const long nch = std::max(1L, slots / std::max(1, n_item_blocks));
  const long chunk = std::max<long>(8L * q->slabT, (q->N + nch - 1) / nch);
  return (int)((chunk + q->slabT - 1) / q->slabT * q->slabT);
}

// Host tables from the pair maps; ld_min forces a minimum stride.
static int build_batch(QCtx* q, int C, const int* iq_pert, const int* pair_map, QBatch& b,
                       int ld_min = 0, bool static_layout = false) {
  const int nq = q->nq, nb = q->nb;
  b.C = C;
  b.MP = nq;
  std::vector<int> h_iqp(C), h_np(C), h_pq1((size_t)C * nq, 0), h_pq2((size_t)C * nq, 0),
      h_offa((size_t)C * nq, 0), h_offb((size_t)C * nq, 0), h_part((size_t)C * nq),
      h_apair((size_t)C * nq), h_orient((size_t)C * nq), h_len(C);
  int maxp = 0;
  for (int c = 0; c < C; ++c) {
    TDS_REQUIRE(iq_pert[c] >= 0 && iq_pert[c] < nq, 2, "q-space: iq_pert out of range");
    const int* pm = pair_map + (long)c * nq;
    for (int i = 0; i < nq; ++i) {
      TDS_REQUIRE(pm[i] >= 0 && pm[i] < nq && pm[pm[i]] == i, 2,
                  "q-space: q_pair_map must be an involution on [0, n_q)");
    }
    h_iqp[c] = iq_pert[c];
    int np = 0;
    for (int i = 0; i < nq; ++i)
      if (i <= pm[i]) {
        h_pq1[(size_t)c * nq + np] = i;
        h_pq2[(size_t)c * nq + np] = pm[i];
        h_apair[(size_t)c * nq + i] = np;
        h_orient[(size_t)c * nq + i] = 0;
        if (pm[i] != i) {
          h_apair[(size_t)c * nq + pm[i]] = np;
          h_orient[(size_t)c * nq + pm[i]] = 1;
        }
        ++np;
      }
    for (int i = 0; i < nq; ++i) h_part[(size_t)c * nq + i] = pm[i];
    int off = nb;
    for (int p = 0; p < np; ++p) {
      h_offa[(size_t)c * nq + p] = off;
      off += h_pq1[(size_t)c * nq + p] < h_pq2[(size_t)c * nq + p] ? nb * nb : nb * (nb + 1) / 2;
    }
    for (int p = 0; p < np; ++p) {     // the static layout has no b' sector
      h_offb[(size_t)c * nq + p] = static_layout ? h_offa[(size_t)c * nq + p] : off;
      if (!static_layout)
        off += h_pq1[(size_t)c * nq + p] < h_pq2[(size_t)c * nq + p] ? nb * nb : nb * (nb + 1) / 2;
    }
    h_np[c] = np;
    h_len[c] = off;
    maxp = std::max(maxp, np);
  }
  int ld = ld_min;
  for (int c = 0; c < C; ++c) ld = std::max(ld, h_len[c]);
  b.ld = ld;
  b.h_len = h_len;
  b.h_np = h_np;
  b.ntile = (q->N + q->weights_tile - 1) / q->weights_tile;
  b.maxp = maxp;
  // slab items and their workgroups
  std::vector<int4> h_items;
  std::vector<int> h_bs(1, 0);
  b.slabT = q->S && q->d2v_form ? q->slabT : 0;
  if (b.slabT) {
    // items sorted by chain, cut at `cap` items or kD2vChains chains
    auto partition = [&](int cap) {
      h_items.clear();
      h_bs.assign(1, 0);
      int first_c = -1;
      for (int c = 0; c < C; ++c)
        for (int p = 0; p < h_np[c]; ++p) {
          const int cur = (int)h_items.size() - h_bs.back();
          if (cur == cap || (cur > 0 && c - first_c >= kD2vChains))
            h_bs.push_back((int)h_items.size());
          if ((int)h_items.size() == h_bs.back()) first_c = c;
          h_items.push_back(make_int4(c, p, h_pq1[(size_t)c * nq + p], h_pq2[(size_t)c * nq + p]));
        }
      h_bs.push_back((int)h_items.size());
    };
    partition(kD2vThreads);
    const int nfull = (int)h_bs.size() - 1;      // then the same count of even workgroups
    partition(((int)h_items.size() + nfull - 1) / nfull);
    b.n_items = (int)h_items.size();
    b.nib = (int)h_bs.size() - 1;
  }
  int chunk;
  if (b.slabT) {
    // automatic: one round of workgroups on the device's CUs
    chunk = q->d2v_chunk > 0 ? q->d2v_chunk : slab_chunk(q, b.nib, 1);
    chunk = (chunk + b.slabT - 1) / b.slabT * b.slabT;
  } else {
    chunk = choose_chunk(q, C, maxp);
  }
  b.nch = (q->N + chunk - 1) / chunk;
  b.chunk = chunk;
  const size_t nbb = (size_t)nb * nb;
  bool ok = hipMalloc((void**)&b.iqp, sizeof(int) * C) == hipSuccess &&
            hipMalloc((void**)&b.npairs, sizeof(int) * C) == hipSuccess &&
            hipMalloc((void**)&b.pq1, sizeof(int) * C * nq) == hipSuccess &&
            hipMalloc((void**)&b.pq2, sizeof(int) * C * nq) == hipSuccess &&
            hipMalloc((void**)&b.offa, sizeof(int) * C * nq) == hipSuccess &&
            hipMalloc((void**)&b.offb, sizeof(int) * C * nq) == hipSuccess &&
            hipMalloc((void**)&b.partner, sizeof(int) * C * nq) == hipSuccess &&
            hipMalloc((void**)&b.apair, sizeof(int) * C * nq) == hipSuccess &&
            hipMalloc((void**)&b.aorient, sizeof(int) * C * nq) == hipSuccess &&
            hipMalloc((void**)&b.active, sizeof(int) * C) == hipSuccess &&
            hipMalloc((void**)&b.alpha, sizeof(double2) * C * nq * nbb) == hipSuccess &&
            hipMalloc((void**)&b.R1, sizeof(double2) * C * nb) == hipSuccess &&
            hipMalloc((void**)&b.wc, sizeof(double2) * (size_t)C * q->N) == hipSuccess &&
            hipMalloc((void**)&b.wd, sizeof(double2) * (size_t)C * q->N) == hipSuccess &&
            hipMalloc((void**)&b.fpart, sizeof(double2) * (size_t)C * b.ntile * nb) == hipSuccess &&
            hipMalloc((void**)&b.dpart, sizeof(double2) * (size_t)C * nq * b.nch * nbb) == hipSuccess &&
            (!b.slabT ||
             (hipMalloc((void**)&b.items, sizeof(int4) * (size_t)b.n_items) == hipSuccess &&
              hipMalloc((void**)&b.blk_start, sizeof(int) * (size_t)(b.nib + 1)) == hipSuccess));
  if (!ok) {
    free_batch(b);
    set_error("q-space: device allocation failed");
    return -1;
  }
  std::vector<int> ones(C, 1);
  ok = hipMemcpy(b.iqp, h_iqp.data(), sizeof(int) * C, hipMemcpyHostToDevice) == hipSuccess &&
       hipMemcpy(b.npairs, h_np.data(), sizeof(int) * C, hipMemcpyHostToDevice) == hipSuccess &&
       hipMemcpy(b.pq1, h_pq1.data(), sizeof(int) * C * nq, hipMemcpyHostToDevice) == hipSuccess &&
       hipMemcpy(b.pq2, h_pq2.data(), sizeof(int) * C * nq, hipMemcpyHostToDevice) == hipSuccess &&
       hipMemcpy(b.offa, h_offa.data(), sizeof(int) * C * nq, hipMemcpyHostToDevice) == hipSuccess &&
       hipMemcpy(b.offb, h_offb.data(), sizeof(int) * C * nq, hipMemcpyHostToDevice) == hipSuccess &&
       hipMemcpy(b.partner, h_part.data(), sizeof(int) * C * nq, hipMemcpyHostToDevice) == hipSuccess &&
       hipMemcpy(b.apair, h_apair.data(), sizeof(int) * C * nq, hipMemcpyHostToDevice) == hipSuccess &&
       hipMemcpy(b.aorient, h_orient.data(), sizeof(int) * C * nq, hipMemcpyHostToDevice) == hipSuccess &&
       hipMemcpy(b.active, ones.data(), sizeof(int) * C, hipMemcpyHostToDevice) == hipSuccess &&
       (!b.slabT ||
        (hipMemcpy(b.items, h_items.data(), sizeof(int4) * (size_t)b.n_items,
                   hipMemcpyHostToDevice) == hipSuccess &&
         hipMemcpy(b.blk_start, h_bs.data(), sizeof(int) * (size_t)(b.nib + 1),
                   hipMemcpyHostToDevice) == hipSuccess));
  if (!ok) {
    free_batch(b);
    set_error("q-space: table upload failed");
    return -1;
  }
  return 0;
}

template <int NB>
static void launch_L_nb(QCtx* q, QBatch& b, const double2* psi, double2* out, int ignore_harmonic,
                        int ignore_v3, int ignore_v4, hipStream_t s, bool static_op) {
  const int C = b.C, MP = b.MP, nq = q->nq;
  const int do_anh = !(ignore_v3 && ignore_v4);
  if (do_anh) {
    if (static_op)
      hipLaunchKernelGGL(q_unpack_static<NB>, dim3(nq, C), dim3(64), 0, s, psi, b.ld, b.active,
                         b.pq1, b.pq2, b.offa, MP, nq, b.apair, b.aorient, q->fY, ignore_v3,
                         b.alpha, b.R1);
    else
      hipLaunchKernelGGL(q_unpack<NB>, dim3(nq, C), dim3(64), 0, s, psi, b.ld, b.active, b.pq1,
                         b.pq2, b.offa, b.offb, MP, nq, b.apair, b.aorient, q->wq, q->nbose,
                         q->valid, ignore_v3, b.alpha, b.R1);
    constexpr int CB = kQWeightsChains;
    const int ngroups = (C + CB - 1) / CB;
    const int nblk = (b.ntile + 7) / 8 * 8 * ngroups;
    hipLaunchKernelGGL((q_weights<NB, CB>), dim3(nblk), dim3(q->weights_tile), 0, s, q->X, q->Y,
                       q->rho, q->N, nq, q->fY, q->fP, b.active, b.iqp, b.partner, b.alpha, b.R1,
                       C, ngroups, b.ntile, ignore_v4 ? 0 : 1, b.wc, b.wd, b.fpart);
    if (b.slabT) {
      const int nblk2 = (b.nch + 7) / 8 * 8 * b.nib;
      hipLaunchKernelGGL(q_d2v_slab<NB>, dim3(nblk2), dim3(kD2vThreads), 0, s, q->S, q->N, nq,
                         b.active, b.items, b.blk_start, b.nib, MP, b.wc, b.wd,
                         log2_pow2(b.slabT), b.chunk, b.nch, b.dpart);
    } else {
      hipLaunchKernelGGL(q_d2v<NB>, dim3(b.maxp, C, b.nch), dim3(256), 0, s, q->X, q->Y, q->N,
                         q->fY, b.active, b.npairs, b.pq1, b.pq2, MP, b.wc, b.wd, b.chunk, b.nch,
                         b.dpart);
    }
  }
  if (static_op)
    hipLaunchKernelGGL(q_finish_static<NB>, dim3(MP + 1, C), dim3(64), 0, s, psi, b.ld, out,
                       b.active, b.iqp, b.npairs, b.pq1, b.pq2, b.offa, MP, q->wq, q->nbose,
                       q->valid, q->T > 1e-12 ? RY_TO_K / q->T : 0.0, b.fpart, b.ntile, b.dpart,
                       b.nch, 1.0 / q->neff, ignore_harmonic ? 0 : 1, do_anh);
  else
    hipLaunchKernelGGL(q_finish<NB>, dim3(MP + 1, C), dim3(64), 0, s, psi, b.ld, out, b.active,
                       b.iqp, b.npairs, b.pq1, b.pq2, b.offa, b.offb, MP, q->wq, q->nbose,
                       q->valid, b.fpart, b.ntile, b.dpart, b.nch, 1.0 / q->neff,
                       ignore_harmonic ? 0 : 1, do_anh);
}

static void launch_L(QCtx* q, QBatch& b, const double2* psi, double2* out, int ih, int i3, int i4,
                     hipStream_t s, bool static_op = false) {
  if (q->nb == 6)
    launch_L_nb<6>(q, b, psi, out, ih, i3, i4, s, static_op);
  else
    launch_L_nb<3>(q, b, psi, out, ih, i3, i4, s, static_op);
}

// ------------------------------------------------------- Hermitian Lanczos
struct QLanczos {
  int kmax = 0;
  double2* Q = nullptr;     // [C][kmax+1][ld]
  double2* qcur = nullptr;  // [C][ld]
  double2* Lq = nullptr;
  double2* r = nullptr;
  double* mask = nullptr;   // [C][ld]
  double* coef = nullptr;   // [C][kmax+1]
  double* a = nullptr;      // [C][kmax]
  double* bb = nullptr;
  double* cc = nullptr;
  double* scal = nullptr;   // [C][4]: a_k, b_k^2, |q'|^2
  int* k = nullptr;
  int* cont = nullptr;
  int* nextc = nullptr;
  int* cap = nullptr;
  int* ndone = nullptr;
};

static void free_ql(QLanczos& l) {
  void* p[] = {l.Q, l.qcur, l.Lq, l.r, l.mask, l.coef, l.a, l.bb, l.cc, l.scal, l.k, l.cont,
               l.nextc, l.cap, l.ndone};
  for (void* x : p)
    if (x) (void)hipFree(x);
  l = QLanczos();
}

// Re <x_c, y_c>_M, one workgroup per chain (vectors of a few thousand entries)
__global__ void __launch_bounds__(256) ql_dot(const double2* __restrict__ x,
                                              const double2* __restrict__ y,
                                              const double* __restrict__ mask, int ld,
                                              const int* __restrict__ flag,
                                              double* __restrict__ scal, int slot) {
  const int c = blockIdx.x, t = threadIdx.x;
  if (!flag[c]) return;
  const double2* xc = x + (long)c * ld;
  const double2* yc = y + (long)c * ld;
  const double* mc = mask + (long)c * ld;
  double s = 0.0;
  for (int e = t; e < ld; e += blockDim.x) s += mc[e] * (xc[e].x * yc[e].x + xc[e].y * yc[e].y);
  __shared__ double red[4];
  s = wave_sum(s);
  if ((t & 63) == 0) red[t >> 6] = s;
  __syncthreads();
  if (t == 0) scal[c * 4 + slot] = red[0] + red[1] + red[2] + red[3];
}

__global__ void ql_init(const double2* __restrict__ psi0, int ld, int kmax,
                        const double* __restrict__ scal, double2* __restrict__ Q,
                        double2* __restrict__ qcur) {
  const int c = blockIdx.y;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= ld) return;
  const double nrm = sqrt(scal[c * 4 + 2]);
  const double2 v = csc(psi0[(long)c * ld + e], 1.0 / nrm);
  Q[(long)c * (kmax + 1) * ld + e] = v;
  qcur[(long)c * ld + e] = v;
}

// r = L q_k - a_k q_k - c_{k-1} q_{k-1}
__global__ void ql_residual(const double2* __restrict__ Lq, const double2* __restrict__ qcur,
                            const double2* __restrict__ Q, int ld, int kmax,
                            const int* __restrict__ active, const int* __restrict__ kk,
                            const double* __restrict__ scal, const double* __restrict__ cc,
                            double* __restrict__ a, double2* __restrict__ r) {
  const int c = blockIdx.y;
  if (!active[c]) return;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  const int k = kk[c];
  const double ak = scal[c * 4 + 0];
  if (e == 0) a[c * kmax + k] = ak;
  if (e >= ld) return;
  const long o = (long)c * ld + e;
  double2 v = make_double2(Lq[o].x - ak * qcur[o].x, Lq[o].y - ak * qcur[o].y);
  if (k > 0) {
    const double cp = cc[c * kmax + k - 1];
    const double2 qm = Q[((long)c * (kmax + 1) + k - 1) * ld + e];
    v.x -= cp * qm.x;
    v.y -= cp * qm.y;
  }
  r[o] = v;
}

// stop tests of QL:1083-1095 and the step cap
__global__ void ql_decide(int C, int kmax, double tol, int* __restrict__ active,
                          int* __restrict__ cont, const int* __restrict__ nextc,
                          const int* __restrict__ kk, const int* __restrict__ cap,
                          int* __restrict__ ndone, const double* __restrict__ scal,
                          double* __restrict__ bb, double* __restrict__ cc) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= C) return;
  cont[c] = 0;
  if (!active[c]) return;
  const int k = kk[c];
  const double b = sqrt(scal[c * 4 + 1]);
  ndone[c] = k + 1;
  const bool conv = !(fabs(b) >= tol) || nextc[c];
  if (conv) {
    active[c] = 0;
    return;
  }
  bb[c * kmax + k] = b;
  cc[c * kmax + k] = b;     // c_k = Re<r/b, r/b> b = b for the Hermitian chain
  if (k + 1 >= cap[c]) {
    active[c] = 0;
    return;
  }
  cont[c] = 1;
}

__global__ void ql_scale(int ld, const int* __restrict__ cont, const double* __restrict__ scal,
                         double2* __restrict__ r) {
  const int c = blockIdx.y;
  if (!cont[c]) return;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= ld) return;
  const double inv = 1.0 / sqrt(scal[c * 4 + 1]);
  r[(long)c * ld + e] = csc(r[(long)c * ld + e], inv);
}

// coef_j = Re <Q_j, r>_M for j <= k, JT basis vectors per workgroup
template <int JT>
__global__ void __launch_bounds__(256) ql_rdots(const double2* __restrict__ Q,
                                                const double2* __restrict__ r,
                                                const double* __restrict__ mask, int ld, int kmax,
                                                const int* __restrict__ cont,
                                                const int* __restrict__ kk,
                                                double* __restrict__ coef) {
  const int c = blockIdx.x, jt = blockIdx.y, t = threadIdx.x;
  if (!cont[c]) return;
  const int k = kk[c];
  const int j0 = jt * JT;
  if (j0 > k) return;
  const int nj = min(JT, k + 1 - j0);
  double s[JT];
#pragma unroll
  for (int j = 0; j < JT; ++j) s[j] = 0.0;
  const double2* rc = r + (long)c * ld;
  const double* mc = mask + (long)c * ld;
  const double2* Qc = Q + ((long)c * (kmax + 1) + j0) * ld;
  for (int e = t; e < ld; e += blockDim.x) {
    const double2 re = rc[e];
    const double m = mc[e];
#pragma unroll
    for (int j = 0; j < JT; ++j)
      if (j < nj) {
        const double2 qv = Qc[(long)j * ld + e];
        s[j] += m * (qv.x * re.x + qv.y * re.y);
      }
  }
  __shared__ double red[4][JT];
#pragma unroll
  for (int j = 0; j < JT; ++j) {
    const double v = wave_sum(s[j]);
    if ((t & 63) == 0) red[t >> 6][j] = v;
  }
  __syncthreads();
  if (t < nj) coef[c * (kmax + 1) + j0 + t] = red[0][t] + red[1][t] + red[2][t] + red[3][t];
}

__global__ void ql_rupdate(const double2* __restrict__ Q, const double* __restrict__ coef,
                           int ld, int kmax, const int* __restrict__ cont,
                           const int* __restrict__ kk, double2* __restrict__ r) {
  const int c = blockIdx.y;
  if (!cont[c]) return;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= ld) return;
  const int k = kk[c];
  const double2* Qc = Q + (long)c * (kmax + 1) * ld + e;
  const double* cf = coef + c * (kmax + 1);
  double2 v = r[(long)c * ld + e];
  for (int j = 0; j <= k; ++j) {
    const double2 qv = Qc[(long)j * ld];
    v.x -= cf[j] * qv.x;
    v.y -= cf[j] * qv.y;
  }
  r[(long)c * ld + e] = v;
}

__global__ void ql_store(int ld, int kmax, double tol, const int* __restrict__ cont,
                         const int* __restrict__ kk, const double* __restrict__ scal,
                         int* __restrict__ nextc, const double2* __restrict__ r,
                         double2* __restrict__ Q, double2* __restrict__ qcur) {
  const int c = blockIdx.y;
  if (!cont[c]) return;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  const int k = kk[c];
  const double nrm = sqrt(scal[c * 4 + 2]);
  const bool small = !(nrm >= tol);
  if (e == 0) nextc[c] = small ? 1 : 0;
  if (e >= ld) return;
  const double2 v = small ? r[(long)c * ld + e] : csc(r[(long)c * ld + e], 1.0 / nrm);
  Q[((long)c * (kmax + 1) + k + 1) * ld + e] = v;
  qcur[(long)c * ld + e] = v;
}

__global__ void ql_advance(int C, const int* __restrict__ cont, int* __restrict__ kk) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c < C && cont[c]) kk[c] += 1;
}

__global__ void q_fill_int(int* p, int n, int v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

template <int JT>
static void reorth_pass(QLanczos& l, QBatch& b, hipStream_t s) {
  const int ld = b.ld, C = b.C;
  hipLaunchKernelGGL(ql_rdots<JT>, dim3(C, (l.kmax + JT) / JT), dim3(256), 0, s, l.Q, l.r, l.mask,
                     ld, l.kmax, l.cont, l.k, l.coef);
  hipLaunchKernelGGL(ql_rupdate, dim3((ld + 255) / 256, C), dim3(256), 0, s, l.Q, l.coef, ld,
                     l.kmax, l.cont, l.k, l.r);
}

static void ql_step(QCtx* q, QBatch& b, QLanczos& l, double tol, int ih, int i3, int i4,
                    hipStream_t s) {
  const int ld = b.ld, C = b.C;
  const dim3 gv((ld + 255) / 256, C);
  launch_L(q, b, l.qcur, l.Lq, ih, i3, i4, s);
  hipLaunchKernelGGL(ql_dot, dim3(C), dim3(256), 0, s, l.qcur, l.Lq, l.mask, ld, b.active, l.scal,
                     0);
  hipLaunchKernelGGL(ql_residual, gv, dim3(256), 0, s, l.Lq, l.qcur, l.Q, ld, l.kmax, b.active,
                     l.k, l.scal, l.cc, l.a, l.r);
  hipLaunchKernelGGL(ql_dot, dim3(C), dim3(256), 0, s, l.r, l.r, l.mask, ld, b.active, l.scal, 1);
  hipLaunchKernelGGL(ql_decide, dim3((C + 63) / 64), dim3(64), 0, s, C, l.kmax, tol, b.active,
                     l.cont, l.nextc, l.k, l.cap, l.ndone, l.scal, l.bb, l.cc);
  hipLaunchKernelGGL(ql_scale, gv, dim3(256), 0, s, ld, l.cont, l.scal, l.r);
  for (int pass = 0; pass < 2; ++pass) {
    if (q->reorth_jtile == 8)
      reorth_pass<8>(l, b, s);
    else if (q->reorth_jtile == 32)
      reorth_pass<32>(l, b, s);
    else
      reorth_pass<16>(l, b, s);
  }
  hipLaunchKernelGGL(ql_dot, dim3(C), dim3(256), 0, s, l.r, l.r, l.mask, ld, l.cont, l.scal, 2);
  hipLaunchKernelGGL(ql_store, gv, dim3(256), 0, s, ld, l.kmax, tol, l.cont, l.k, l.scal, l.nextc,
                     l.r, l.Q, l.qcur);
  hipLaunchKernelGGL(ql_advance, dim3((C + 63) / 64), dim3(64), 0, s, C, l.cont, l.k);
}

// Hermitian metric of QL:441-480 for every chain's layout
static void host_mask(const QBatch& b, int nq, int nb, const int* pair_map, int c,
                      std::vector<double>& m) {
  const int ld = b.ld;
  std::fill(m.begin() + (long)c * ld, m.begin() + (long)(c + 1) * ld, 0.0);
  double* mc = m.data() + (long)c * ld;
  for (int v = 0; v < nb; ++v) mc[v] = 1.0;
  const int* pm = pair_map + (long)c * nq;
  int off = nb;
  std::vector<std::pair<int, int>> pairs;
  for (int i = 0; i < nq; ++i)
    if (i <= pm[i]) pairs.push_back({i, pm[i]});
  for (int sector = 0; sector < 2; ++sector)
    for (auto& pr : pairs) {
      if (pr.first < pr.second) {
        for (int e = 0; e < nb * nb; ++e) mc[off + e] = 2.0;
        off += nb * nb;
      } else {
        for (int i = 0; i < nb; ++i)
          for (int j = i; j < nb; ++j) mc[off++] = (i == j) ? 1.0 : 2.0;
      }
    }
}

// ------------------------------------------------- kernel polynomial method
// One work-group per chain: d = <x, y>_M (fixed-order sum), then moment m of
// chain c is d (mode 0) or 2 d - mom[c][ref] (mode 1: mu_{2n} with ref 0,
// mu_{2n+1} with ref 1; QSK:346-390).
__global__ void __launch_bounds__(256) kpm_moment(const double2* __restrict__ x,
                                                  const double2* __restrict__ y,
                                                  const double* __restrict__ mask, int ld,
                                                  double* __restrict__ mom, int n_mom, int m,
                                                  int mode, int ref) {
  const int c = blockIdx.x, t = threadIdx.x;
  const double2* xc = x + (long)c * ld;
  const double2* yc = y + (long)c * ld;
  const double* mc = mask + (long)c * ld;
  double s = 0.0;
  for (int e = t; e < ld; e += blockDim.x) s += mc[e] * (xc[e].x * yc[e].x + xc[e].y * yc[e].y);
  __shared__ double red[4];
  s = wave_sum(s);
  if ((t & 63) == 0) red[t >> 6] = s;
  __syncthreads();
  if (t == 0) {
    const double d = red[0] + red[1] + red[2] + red[3];
    mom[(long)c * n_mom + m] = mode ? 2.0 * d - mom[(long)c * n_mom + ref] : d;
  }
}

// Chebyshev step of chain c (blockIdx.y): t = (L v - b v) / a with (a, b) =
// rescale[c]; out = t (first step) or 2 t - vm (out may alias vm).
__global__ void kpm_combine(const double2* __restrict__ Lv, const double2* __restrict__ v,
                            const double2* vm, const double* __restrict__ rescale, int ld,
                            int first, double2* out) {
  const int c = blockIdx.y, e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= ld) return;
  const long i = (long)c * ld + e;
  const double a = rescale[2 * c], b = rescale[2 * c + 1];
  const double2 lv = Lv[i], vv = v[i];
  double2 r = make_double2((lv.x - b * vv.x) / a, (lv.y - b * vv.y) / a);
  if (!first) {
    const double2 p = vm[i];
    r = make_double2(2.0 * r.x - p.x, 2.0 * r.y - p.y);
  }
  out[i] = r;
}

}  // namespace tdscha

using namespace tdscha;

extern "C" {

int tdscha_qctx_create(tdscha_qctx** out, int n_q, int n_bands, int n_configs, const double* X_q,
                       const double* Y_q, const double* w_q, const int* valid_modes_q,
                       const double* rho, double T) {
  TDS_REQUIRE(out != nullptr, 2, "tdscha_qctx_create: NULL out");
  TDS_REQUIRE(n_q >= 1 && n_q <= kQMaxQ, 2, "tdscha_qctx_create: n_q must be in [1, 512]");
  TDS_REQUIRE(n_bands == 3 || n_bands == 6, 2,
              "tdscha_qctx_create: n_bands must be 3 or 6 (1 or 2 atoms per cell)");
  TDS_REQUIRE(n_configs >= 1, 2, "tdscha_qctx_create: n_configs must be positive");
  TDS_REQUIRE(X_q && Y_q && w_q && valid_modes_q && rho, 2, "tdscha_qctx_create: NULL array");
  *out = nullptr;
  const int nq = n_q, nb = n_bands, N = n_configs;
  QCtx* q = new QCtx();
  q->nq = nq;
  q->nb = nb;
  q->N = N;
  q->T = T;
  double neff = 0.0;
  for (int i = 0; i < N; ++i) neff += rho[i];
  TDS_REQUIRE(neff > 0, 2, "tdscha_qctx_create: sum(rho) must be positive");
  q->neff = neff;
  // host copies in device layouts
  const size_t ne = (size_t)nq * nb * N;
  std::vector<double2> hx(ne), hy(ne);
  for (int iq = 0; iq < nq; ++iq)
    for (int i = 0; i < N; ++i)
      for (int v = 0; v < nb; ++v) {
        const size_t src = (((size_t)iq * N + i) * nb + v) * 2;
        const size_t dst = ((size_t)iq * nb + v) * N + i;
        hx[dst] = make_double2(X_q[src], X_q[src + 1]);
        hy[dst] = make_double2(Y_q[src], Y_q[src + 1]);
      }
  std::vector<double> hw(nq * nb), hn(nq * nb), hfy(nq * nb), hfp(nq * nb);
  std::vector<int> hv(nq * nb);
  for (int iq = 0; iq < nq; ++iq)
    for (int v = 0; v < nb; ++v) {
      const double w = w_q[v * nq + iq];
      const int ok = valid_modes_q[v * nq + iq] ? 1 : 0;
      double n = 0.0;
      if (ok && T > 1e-12) n = 1.0 / (std::exp(w * RY_TO_K / T) - 1.0);   // QL:581-602
      hw[iq * nb + v] = w;
      hn[iq * nb + v] = n;
      hv[iq * nb + v] = ok;
      // JQ:680-699
      double nw = 0.0;
      if (ok && T > 0) nw = 1.0 / (std::exp(w * RY_TO_K / T) - 1.0);
      hfy[iq * nb + v] = ok ? 2.0 * w / (1.0 + 2.0 * nw) : 0.0;
      hfp[iq * nb + v] = ok ? (1.0 + 2.0 * nw) / (2.0 * w) : 0.0;
    }
  auto fail = [&](const char* what) {
    tdscha_qctx_destroy(reinterpret_cast<tdscha_qctx*>(q));
    set_error(std::string("tdscha_qctx_create: ") + what);
    return -1;
  };
  if (hipStreamCreate(&q->stream) != hipSuccess) return fail("hipStreamCreate failed");
  {
    int dev = 0, ncu = 0;
    if (hipGetDevice(&dev) == hipSuccess &&
        hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess &&
        ncu > 0)
      q->n_cu = ncu;
  }
  bool ok = hipMalloc((void**)&q->X, sizeof(double2) * ne) == hipSuccess &&
            hipMalloc((void**)&q->Y, sizeof(double2) * ne) == hipSuccess &&
            hipMalloc((void**)&q->rho, sizeof(double) * N) == hipSuccess &&
            hipMalloc((void**)&q->wq, sizeof(double) * nq * nb) == hipSuccess &&
            hipMalloc((void**)&q->nbose, sizeof(double) * nq * nb) == hipSuccess &&
            hipMalloc((void**)&q->valid, sizeof(int) * nq * nb) == hipSuccess &&
            hipMalloc((void**)&q->fY, sizeof(double) * nq * nb) == hipSuccess &&
            hipMalloc((void**)&q->fP, sizeof(double) * nq * nb) == hipSuccess;
  if (!ok) return fail("device allocation failed");
  ok = hipMemcpy(q->X, hx.data(), sizeof(double2) * ne, hipMemcpyHostToDevice) == hipSuccess &&
       hipMemcpy(q->Y, hy.data(), sizeof(double2) * ne, hipMemcpyHostToDevice) == hipSuccess &&
       hipMemcpy(q->rho, rho, sizeof(double) * N, hipMemcpyHostToDevice) == hipSuccess &&
       hipMemcpy(q->wq, hw.data(), sizeof(double) * nq * nb, hipMemcpyHostToDevice) == hipSuccess &&
       hipMemcpy(q->nbose, hn.data(), sizeof(double) * nq * nb, hipMemcpyHostToDevice) == hipSuccess &&
       hipMemcpy(q->valid, hv.data(), sizeof(int) * nq * nb, hipMemcpyHostToDevice) == hipSuccess &&
       hipMemcpy(q->fY, hfy.data(), sizeof(double) * nq * nb, hipMemcpyHostToDevice) == hipSuccess &&
       hipMemcpy(q->fP, hfp.data(), sizeof(double) * nq * nb, hipMemcpyHostToDevice) == hipSuccess;
  if (!ok) return fail("upload failed");
  if (build_slabs(q)) return fail("slab images failed");
  *out = reinterpret_cast<tdscha_qctx*>(q);
  return 0;
}

void tdscha_qctx_destroy(tdscha_qctx* ctx) {
  if (!ctx) return;
  QCtx* q = reinterpret_cast<QCtx*>(ctx);
  if (q->stream) (void)hipStreamSynchronize(q->stream);
  void* p[] = {q->X, q->Y, q->rho, q->wq, q->nbose, q->valid, q->fY, q->fP, q->S};
  for (void* x : p)
    if (x) (void)hipFree(x);
  if (q->stream) (void)hipStreamDestroy(q->stream);
  delete q;
}

int tdscha_qctx_symmetrize(tdscha_qctx* ctx, int n_sym, const int* q_perm, const double* D) {
  TDS_REQUIRE(ctx && q_perm && D, 2, "tdscha_qctx_symmetrize: NULL argument");
  QCtx* q = reinterpret_cast<QCtx*>(ctx);
  const int nq = q->nq, nb = q->nb, N = q->N;
  TDS_REQUIRE(n_sym >= 1 && n_sym <= 65535, 2, "tdscha_qctx_symmetrize: n_sym must be in [1, 65535]");
  TDS_REQUIRE(q->n_sym == 1, 2, "tdscha_qctx_symmetrize: the ensemble is already expanded");
  TDS_REQUIRE((long)N * n_sym <= 0x7fffffffL, 2,
              "tdscha_qctx_symmetrize: n_sym * n_configs exceeds 2^31 - 1");
  for (int j = 0; j < n_sym; ++j) {
    std::vector<char> seen(nq, 0);
    for (int p = 0; p < nq; ++p) {
      const int t = q_perm[(long)j * nq + p];
      TDS_REQUIRE(t >= 0 && t < nq && !seen[t], 2,
                  "tdscha_qctx_symmetrize: q_perm row is not a permutation of the q-points");
      seen[t] = 1;
    }
  }
  const int N2 = N * n_sym;
  const size_t ne2 = (size_t)nq * nb * N2;
  std::vector<double> hrho(N), hrho2((size_t)N2);
  if (hipMemcpy(hrho.data(), q->rho, sizeof(double) * N, hipMemcpyDeviceToHost) != hipSuccess) {
    set_error("tdscha_qctx_symmetrize: rho download failed");
    return -1;
  }
  for (int j = 0; j < n_sym; ++j)
    for (int i = 0; i < N; ++i) hrho2[(size_t)j * N + i] = hrho[i];
  double2 *XE = nullptr, *YE = nullptr, *dD = nullptr;
  double* rho2 = nullptr;
  int* dperm = nullptr;
  auto release = [&]() {
    void* p[] = {XE, YE, dD, rho2, dperm};
    for (void* x : p)
      if (x) (void)hipFree(x);
  };
  bool ok = hipMalloc((void**)&XE, sizeof(double2) * ne2) == hipSuccess &&
            hipMalloc((void**)&YE, sizeof(double2) * ne2) == hipSuccess &&
            hipMalloc((void**)&rho2, sizeof(double) * N2) == hipSuccess &&
            hipMalloc((void**)&dperm, sizeof(int) * (size_t)n_sym * nq) == hipSuccess &&
            hipMalloc((void**)&dD, sizeof(double2) * (size_t)n_sym * nq * nb * nb) == hipSuccess;
  if (!ok) {
    release();
    set_error("tdscha_qctx_symmetrize: device allocation failed");
    return -1;
  }
  ok = hipMemcpy(rho2, hrho2.data(), sizeof(double) * N2, hipMemcpyHostToDevice) == hipSuccess &&
       hipMemcpy(dperm, q_perm, sizeof(int) * (size_t)n_sym * nq, hipMemcpyHostToDevice) ==
           hipSuccess &&
       hipMemcpy(dD, D, sizeof(double2) * (size_t)n_sym * nq * nb * nb, hipMemcpyHostToDevice) ==
           hipSuccess;
  if (!ok) {
    release();
    set_error("tdscha_qctx_symmetrize: upload failed");
    return -1;
  }
  const dim3 g((unsigned)((N + 255) / 256), (unsigned)n_sym, (unsigned)nq);
  if (nb == 6)
    hipLaunchKernelGGL(q_expand<6>, g, dim3(256), 0, q->stream, q->X, q->Y, N, nq, dperm, dD, XE, YE);
  else
    hipLaunchKernelGGL(q_expand<3>, g, dim3(256), 0, q->stream, q->X, q->Y, N, nq, dperm, dD, XE, YE);
  if (hipGetLastError() != hipSuccess || hipStreamSynchronize(q->stream) != hipSuccess) {
    release();
    set_error("tdscha_qctx_symmetrize: expansion kernel failed");
    return -1;
  }
  (void)hipFree(q->X);
  (void)hipFree(q->Y);
  (void)hipFree(q->rho);
  (void)hipFree(dperm);
  (void)hipFree(dD);
  q->X = XE;
  q->Y = YE;
  q->rho = rho2;
  q->N = N2;
  q->neff *= n_sym;
  q->n_sym = n_sym;
  return build_slabs(q);
}

static int apply_L_host(tdscha_qctx* ctx, int n_chains, const int* iq_pert,
                        const int* q_pair_map, const double* psi, double* out,
                        int ignore_harmonic, int ignore_v3, int ignore_v4, bool static_op);

int tdscha_qspace_apply_L(tdscha_qctx* ctx, int n_chains, const int* iq_pert,
                          const int* q_pair_map, const double* psi, double* out,
                          int ignore_harmonic, int ignore_v3, int ignore_v4) {
  return apply_L_host(ctx, n_chains, iq_pert, q_pair_map, psi, out, ignore_harmonic, ignore_v3,
                      ignore_v4, false);
}

int tdscha_qspace_apply_L_static(tdscha_qctx* ctx, int n_chains, const int* iq_pert,
                                 const int* q_pair_map, const double* psi, double* out,
                                 int ignore_harmonic, int ignore_v3, int ignore_v4) {
  return apply_L_host(ctx, n_chains, iq_pert, q_pair_map, psi, out, ignore_harmonic, ignore_v3,
                      ignore_v4, true);
}

int tdscha_qspace_static_psi_size(int n_q, int n_bands, const int* q_pair_map, int* psi_size) {
  int full = 0;
  const int rc = tdscha_qspace_psi_size(n_q, n_bands, q_pair_map, &full);
  if (rc) return rc;
  *psi_size = n_bands + (full - n_bands) / 2;
  return 0;
}

int tdscha_qspace_psi_size(int n_q, int n_bands, const int* q_pair_map, int* psi_size) {
  TDS_REQUIRE(q_pair_map && psi_size && n_q > 0 && n_bands > 0, 2,
              "tdscha_qspace_psi_size: bad argument");
  int off = n_bands;
  for (int i = 0; i < n_q; ++i) {
    TDS_REQUIRE(q_pair_map[i] >= 0 && q_pair_map[i] < n_q && q_pair_map[q_pair_map[i]] == i, 2,
                "tdscha_qspace_psi_size: q_pair_map must be an involution");
    if (i < q_pair_map[i]) off += 2 * n_bands * n_bands;
    else if (i == q_pair_map[i]) off += n_bands * (n_bands + 1);
  }
  *psi_size = off;
  return 0;
}

int tdscha_qspace_set_tiles(tdscha_qctx* ctx, int weights_tile, int d2v_chunk, int reorth_jtile) {
  TDS_REQUIRE(ctx != nullptr, 2, "tdscha_qspace_set_tiles: NULL context");
  TDS_REQUIRE(weights_tile == 64 || weights_tile == 128 || weights_tile == 256, 2,
              "weights_tile must be 64, 128 or 256");
  TDS_REQUIRE(d2v_chunk >= 0, 2, "d2v_chunk must be >= 0 (0 = automatic)");
  TDS_REQUIRE(reorth_jtile == 8 || reorth_jtile == 16 || reorth_jtile == 32, 2,
              "reorth_jtile must be 8, 16 or 32");
  QCtx* q = reinterpret_cast<QCtx*>(ctx);
  q->weights_tile = weights_tile;
  q->d2v_chunk = d2v_chunk;
  q->reorth_jtile = reorth_jtile;
  return 0;
}

int tdscha_qspace_get_d2v_form(const tdscha_qctx* ctx, int* form) {
  TDS_REQUIRE(ctx && form, 2, "tdscha_qspace_get_d2v_form: NULL argument");
  const QCtx* q = reinterpret_cast<const QCtx*>(ctx);
  *form = q->S && q->d2v_form ? 1 : 0;
  return 0;
}

int tdscha_qspace_set_d2v_form(tdscha_qctx* ctx, int form) {
  TDS_REQUIRE(ctx != nullptr, 2, "tdscha_qspace_set_d2v_form: NULL context");
  TDS_REQUIRE(form == 0 || form == 1, 2, "d2v form must be 0 or 1");
  reinterpret_cast<QCtx*>(ctx)->d2v_form = form;
  return 0;
}

int tdscha_qspace_get_tiles(const tdscha_qctx* ctx, int* tiles_out) {
  TDS_REQUIRE(ctx && tiles_out, 2, "tdscha_qspace_get_tiles: NULL argument");
  const QCtx* q = reinterpret_cast<const QCtx*>(ctx);
  tiles_out[0] = q->weights_tile;
  tiles_out[1] = q->d2v_chunk;
  tiles_out[2] = q->reorth_jtile;
  return 0;
}

static int apply_L_host(tdscha_qctx* ctx, int n_chains, const int* iq_pert,
                        const int* q_pair_map, const double* psi, double* out,
                        int ignore_harmonic, int ignore_v3, int ignore_v4, bool static_op) {
  TDS_REQUIRE(ctx && iq_pert && q_pair_map && psi && out, 2, "tdscha_qspace_apply_L: NULL argument");
  TDS_REQUIRE(n_chains >= 1 && n_chains <= kQMaxChains, 2, "n_chains must be in [1, 256]");
  QCtx* q = reinterpret_cast<QCtx*>(ctx);
  QBatch b;
  int rc = build_batch(q, n_chains, iq_pert, q_pair_map, b, 0, static_op);
  if (rc) return rc;
  const int C = n_chains, ld = b.ld;
  double2 *dpsi = nullptr, *dout = nullptr;
  std::vector<double2> hp((size_t)C * ld, make_double2(0.0, 0.0));
  long pos = 0;
  for (int c = 0; c < C; ++c) {
    for (int e = 0; e < b.h_len[c]; ++e)
      hp[(size_t)c * ld + e] = make_double2(psi[2 * (pos + e)], psi[2 * (pos + e) + 1]);
    pos += b.h_len[c];
  }
  hipStream_t s = q->stream;
  bool ok = hipMalloc((void**)&dpsi, sizeof(double2) * C * ld) == hipSuccess &&
            hipMalloc((void**)&dout, sizeof(double2) * C * ld) == hipSuccess &&
            hipMemcpy(dpsi, hp.data(), sizeof(double2) * C * ld, hipMemcpyHostToDevice) == hipSuccess &&
            hipMemset(dout, 0, sizeof(double2) * C * ld) == hipSuccess;
  if (ok) {
    launch_L(q, b, dpsi, dout, ignore_harmonic, ignore_v3, ignore_v4, s, static_op);
    ok = hipGetLastError() == hipSuccess && hipStreamSynchronize(s) == hipSuccess &&
         hipMemcpy(hp.data(), dout, sizeof(double2) * C * ld, hipMemcpyDeviceToHost) == hipSuccess;
  }
  if (dpsi) (void)hipFree(dpsi);
  if (dout) (void)hipFree(dout);
  const std::vector<int> lens = b.h_len;
  free_batch(b);
  if (!ok) {
    set_error("tdscha_qspace_apply_L: device execution failed");
    return -1;
  }
  pos = 0;
  for (int c = 0; c < C; ++c) {
    for (int e = 0; e < lens[c]; ++e) {
      out[2 * (pos + e)] = hp[(size_t)c * ld + e].x;
      out[2 * (pos + e) + 1] = hp[(size_t)c * ld + e].y;
    }
    pos += lens[c];
  }
  return 0;
}

int tdscha_qspace_lanczos_run(tdscha_qctx* ctx, int n_chains, const int* iq_pert,
                              const int* q_pair_map, const double* psi0, int n_steps,
                              const int* caps, double tol, int ignore_harmonic, int ignore_v3,
                              int ignore_v4, int use_graph, double* a_out, double* b_out,
                              double* c_out, int* n_done, double* elapsed_ms) {
  TDS_REQUIRE(ctx && iq_pert && q_pair_map && psi0 && a_out && b_out && n_done, 2,
              "tdscha_qspace_lanczos_run: NULL argument");
  TDS_REQUIRE(n_chains >= 1 && n_chains <= kQMaxChains, 2, "n_chains must be in [1, 256]");
  TDS_REQUIRE(n_steps >= 1 && n_steps <= 100000, 2, "n_steps must be in [1, 100000]");
  QCtx* q = reinterpret_cast<QCtx*>(ctx);
  QBatch b;
  int rc = build_batch(q, n_chains, iq_pert, q_pair_map, b);
  if (rc) return rc;
  const int C = n_chains, ld = b.ld, kmax = n_steps;
  QLanczos l;
  l.kmax = kmax;
  const size_t qn = (size_t)C * (kmax + 1) * ld;
  bool ok = hipMalloc((void**)&l.Q, sizeof(double2) * qn) == hipSuccess &&
            hipMalloc((void**)&l.qcur, sizeof(double2) * C * ld) == hipSuccess &&
            hipMalloc((void**)&l.Lq, sizeof(double2) * C * ld) == hipSuccess &&
            hipMalloc((void**)&l.r, sizeof(double2) * C * ld) == hipSuccess &&
            hipMalloc((void**)&l.mask, sizeof(double) * C * ld) == hipSuccess &&
            hipMalloc((void**)&l.coef, sizeof(double) * C * (kmax + 1)) == hipSuccess &&
            hipMalloc((void**)&l.a, sizeof(double) * C * kmax) == hipSuccess &&
            hipMalloc((void**)&l.bb, sizeof(double) * C * kmax) == hipSuccess &&
            hipMalloc((void**)&l.cc, sizeof(double) * C * kmax) == hipSuccess &&
            hipMalloc((void**)&l.scal, sizeof(double) * C * 4) == hipSuccess &&
            hipMalloc((void**)&l.k, sizeof(int) * C) == hipSuccess &&
            hipMalloc((void**)&l.cont, sizeof(int) * C) == hipSuccess &&
            hipMalloc((void**)&l.nextc, sizeof(int) * C) == hipSuccess &&
            hipMalloc((void**)&l.cap, sizeof(int) * C) == hipSuccess &&
            hipMalloc((void**)&l.ndone, sizeof(int) * C) == hipSuccess;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  auto fail = [&](const char* what) {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    free_ql(l);
    free_batch(b);
    set_error(std::string("tdscha_qspace_lanczos_run: ") + what);
    return -1;
  };
  if (!ok) return fail("device allocation failed");
  hipStream_t s = q->stream;
  // zero-padded starting vectors, masks, caps
  std::vector<double2> hp((size_t)C * ld, make_double2(0.0, 0.0));
  std::vector<double> hm((size_t)C * ld, 0.0);
  std::vector<int> hcap(C, n_steps);
  long pos = 0;
  for (int c = 0; c < C; ++c) {
    for (int e = 0; e < b.h_len[c]; ++e)
      hp[(size_t)c * ld + e] = make_double2(psi0[2 * (pos + e)], psi0[2 * (pos + e) + 1]);
    pos += b.h_len[c];
    host_mask(b, q->nq, q->nb, q_pair_map, c, hm);
    if (caps) hcap[c] = std::max(1, std::min(caps[c], n_steps));
  }
  ok = hipMemcpy(l.r, hp.data(), sizeof(double2) * C * ld, hipMemcpyHostToDevice) == hipSuccess &&
       hipMemcpy(l.mask, hm.data(), sizeof(double) * C * ld, hipMemcpyHostToDevice) == hipSuccess &&
       hipMemcpy(l.cap, hcap.data(), sizeof(int) * C, hipMemcpyHostToDevice) == hipSuccess &&
       hipMemset(l.Q, 0, sizeof(double2) * qn) == hipSuccess &&
       hipMemset(l.Lq, 0, sizeof(double2) * C * ld) == hipSuccess &&
       hipMemset(l.a, 0, sizeof(double) * C * kmax) == hipSuccess &&
       hipMemset(l.bb, 0, sizeof(double) * C * kmax) == hipSuccess &&
       hipMemset(l.cc, 0, sizeof(double) * C * kmax) == hipSuccess;
  if (!ok) return fail("upload failed");
  hipLaunchKernelGGL(q_fill_int, dim3((C + 63) / 64), dim3(64), 0, s, l.k, C, 0);
  hipLaunchKernelGGL(q_fill_int, dim3((C + 63) / 64), dim3(64), 0, s, l.cont, C, 1);
  hipLaunchKernelGGL(q_fill_int, dim3((C + 63) / 64), dim3(64), 0, s, l.nextc, C, 0);
  hipLaunchKernelGGL(q_fill_int, dim3((C + 63) / 64), dim3(64), 0, s, l.ndone, C, 0);
  hipLaunchKernelGGL(ql_dot, dim3(C), dim3(256), 0, s, l.r, l.r, l.mask, ld, l.cont, l.scal, 2);
  hipLaunchKernelGGL(ql_init, dim3((ld + 255) / 256, C), dim3(256), 0, s, l.r, ld, kmax, l.scal,
                     l.Q, l.qcur);
  if (hipGetLastError() != hipSuccess || hipStreamSynchronize(s) != hipSuccess)
    return fail("initialisation failed");
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
    return fail("event creation failed");
  if (use_graph) {
    // the whole run as one graph: one launch replays every step
    if (hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) != hipSuccess)
      return fail("stream capture failed to start");
    for (int step = 0; step < n_steps; ++step)
      ql_step(q, b, l, tol, ignore_harmonic, ignore_v3, ignore_v4, s);
    if (hipStreamEndCapture(s, &graph) != hipSuccess) return fail("stream capture failed");
    if (hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) != hipSuccess)
      return fail("graph instantiation failed");
  }
  (void)hipEventRecord(e0, s);
  if (use_graph) {
    if (hipGraphLaunch(exec, s) != hipSuccess) return fail("graph launch failed");
  } else {
    for (int step = 0; step < n_steps; ++step)
      ql_step(q, b, l, tol, ignore_harmonic, ignore_v3, ignore_v4, s);
  }
  (void)hipEventRecord(e1, s);
  if (hipStreamSynchronize(s) != hipSuccess || hipGetLastError() != hipSuccess)
    return fail("step execution failed");
  float ms = 0.f;
  (void)hipEventElapsedTime(&ms, e0, e1);
  if (elapsed_ms) *elapsed_ms = ms;
  ok = hipMemcpy(a_out, l.a, sizeof(double) * C * kmax, hipMemcpyDeviceToHost) == hipSuccess &&
       hipMemcpy(b_out, l.bb, sizeof(double) * C * kmax, hipMemcpyDeviceToHost) == hipSuccess &&
       (!c_out || hipMemcpy(c_out, l.cc, sizeof(double) * C * kmax, hipMemcpyDeviceToHost) == hipSuccess) &&
       hipMemcpy(n_done, l.ndone, sizeof(int) * C, hipMemcpyDeviceToHost) == hipSuccess;
  if (!ok) return fail("download failed");
  if (exec) (void)hipGraphExecDestroy(exec);
  if (graph) (void)hipGraphDestroy(graph);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  free_ql(l);
  free_batch(b);
  return 0;
}

// Time one L application of the batch for every (weights tile, d2v chunk)
// candidate and every reorthogonalisation tile; keep the fastest.
int tdscha_qspace_kpm_run(tdscha_qctx* ctx, int n_chains, const int* iq_pert,
                          const int* q_pair_map, const double* rescale, int n_moments,
                          int n_have, double* moments, double* vm, double* v,
                          int ignore_harmonic, int ignore_v3, int ignore_v4, int use_graph,
                          double* elapsed_ms) {
  TDS_REQUIRE(ctx && iq_pert && q_pair_map && rescale && moments && vm && v, 2,
              "tdscha_qspace_kpm_run: NULL argument");
  TDS_REQUIRE(n_chains >= 1 && n_chains <= kQMaxChains, 2, "n_chains must be in [1, 256]");
  TDS_REQUIRE(n_moments >= 1 && n_moments <= 200000, 2, "n_moments must be in [1, 200000]");
  TDS_REQUIRE(n_have >= 0 && n_have < n_moments, 2, "n_have must be in [0, n_moments)");
  for (int c = 0; c < n_chains; ++c)
    TDS_REQUIRE(rescale[2 * c] > 0, 2, "tdscha_qspace_kpm_run: rescale a must be positive");
  QCtx* q = reinterpret_cast<QCtx*>(ctx);
  QBatch b;
  int rc = build_batch(q, n_chains, iq_pert, q_pair_map, b);
  if (rc) return rc;
  const int C = n_chains, ld = b.ld, N = n_moments;
  const size_t nv = (size_t)C * ld;
  double2 *B0 = nullptr, *B1 = nullptr, *B2 = nullptr, *LV = nullptr;
  double *dmask = nullptr, *dmom = nullptr, *dres = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  auto release = [&]() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    void* p[] = {B0, B1, B2, LV, dmask, dmom, dres};
    for (void* x : p)
      if (x) (void)hipFree(x);
    free_batch(b);
  };
  auto fail = [&](const char* what) {
    release();
    set_error(std::string("tdscha_qspace_kpm_run: ") + what);
    return -1;
  };
  bool ok = hipMalloc((void**)&B0, sizeof(double2) * nv) == hipSuccess &&
            hipMalloc((void**)&B1, sizeof(double2) * nv) == hipSuccess &&
            hipMalloc((void**)&B2, sizeof(double2) * nv) == hipSuccess &&
            hipMalloc((void**)&LV, sizeof(double2) * nv) == hipSuccess &&
            hipMalloc((void**)&dmask, sizeof(double) * nv) == hipSuccess &&
            hipMalloc((void**)&dmom, sizeof(double) * C * N) == hipSuccess &&
            hipMalloc((void**)&dres, sizeof(double) * 2 * C) == hipSuccess;
  if (!ok) return fail("device allocation failed");
  // vectors in the padded [C][ld] layout; masks; moments so far
  std::vector<double2> hvm(nv, make_double2(0.0, 0.0)), hv(nv, make_double2(0.0, 0.0));
  std::vector<double> hmask(nv, 0.0), hmom((size_t)C * N, 0.0);
  long pos = 0;
  for (int c = 0; c < C; ++c) {
    for (int e = 0; e < b.h_len[c]; ++e) {
      hvm[(size_t)c * ld + e] = make_double2(vm[2 * (pos + e)], vm[2 * (pos + e) + 1]);
      hv[(size_t)c * ld + e] = make_double2(v[2 * (pos + e)], v[2 * (pos + e) + 1]);
    }
    pos += b.h_len[c];
    host_mask(b, q->nq, q->nb, q_pair_map, c, hmask);
    for (int m = 0; m < n_have; ++m) hmom[(size_t)c * N + m] = moments[(size_t)c * N + m];
    hmom[(size_t)c * N] = n_have ? hmom[(size_t)c * N] : 1.0;
  }
  ok = hipMemcpy(B0, hvm.data(), sizeof(double2) * nv, hipMemcpyHostToDevice) == hipSuccess &&
       hipMemcpy(B1, hv.data(), sizeof(double2) * nv, hipMemcpyHostToDevice) == hipSuccess &&
       hipMemcpy(dmask, hmask.data(), sizeof(double) * nv, hipMemcpyHostToDevice) == hipSuccess &&
       hipMemcpy(dmom, hmom.data(), sizeof(double) * C * N, hipMemcpyHostToDevice) == hipSuccess &&
       hipMemcpy(dres, rescale, sizeof(double) * 2 * C, hipMemcpyHostToDevice) == hipSuccess &&
       hipMemset(B2, 0, sizeof(double2) * nv) == hipSuccess &&
       hipMemset(LV, 0, sizeof(double2) * nv) == hipSuccess;
  if (!ok) return fail("upload failed");
  hipStream_t s = q->stream;
  const dim3 gv((ld + 255) / 256, C);
  // The recurrence of QSK:301-440: vm, v hold T_{n-1}, T_n; a step computes
  // mu_{2n} = 2 <v, v> - mu_0, then vp = 2 L~ v - vm (into the third
  // buffer) and mu_{2n+1} = 2 <vp, v> - mu_1.  All launches are enqueued
  // here; with use_graph they are captured first and replayed as one graph.
  double2* pv[3] = {B0, B1, B2};   // pv[0] = vm, pv[1] = v, pv[2] = spare
  int n = 0;
  auto L = [&](const double2* in) {
    launch_L(q, b, in, LV, ignore_harmonic, ignore_v3, ignore_v4, s);
  };
  auto rotate = [&]() {             // vm, v = v, vp
    double2* old_vm = pv[0];
    pv[0] = pv[1];
    pv[1] = pv[2];
    pv[2] = old_vm;
  };
  auto enqueue = [&]() {
    if (n_have == 0) {
      // fresh start: vm = v = v0 (given in both); T_1 = L~ v0, mu_1 = <v0, T_1>
      if (N > 1) {
        L(pv[1]);
        hipLaunchKernelGGL(kpm_combine, gv, dim3(256), 0, s, LV, pv[1], pv[0], dres, ld, 1, pv[2]);
        hipLaunchKernelGGL(kpm_moment, dim3(C), dim3(256), 0, s, pv[1], pv[2], dmask, ld, dmom,
                           N, 1, 0, 0);
        rotate();
        n = 1;
      }
    } else if (n_have == 1) {
      L(pv[1]);
      hipLaunchKernelGGL(kpm_combine, gv, dim3(256), 0, s, LV, pv[1], pv[0], dres, ld, 1, pv[2]);
      hipLaunchKernelGGL(kpm_moment, dim3(C), dim3(256), 0, s, pv[0], pv[2], dmask, ld, dmom, N,
                         1, 0, 0);
      rotate();
      n = 1;
    } else if (n_have % 2 == 0) {
      n = n_have / 2;
    } else {
      const int h = n_have / 2;      // the interrupted odd moment
      L(pv[1]);
      hipLaunchKernelGGL(kpm_combine, gv, dim3(256), 0, s, LV, pv[1], pv[0], dres, ld, 0, pv[2]);
      hipLaunchKernelGGL(kpm_moment, dim3(C), dim3(256), 0, s, pv[2], pv[1], dmask, ld, dmom, N,
                         2 * h + 1, 1, 1);
      rotate();
      n = h + 1;
    }
    if (N > 1 || n_have > 0) {
      while (2 * n < N) {
        hipLaunchKernelGGL(kpm_moment, dim3(C), dim3(256), 0, s, pv[1], pv[1], dmask, ld, dmom,
                           N, 2 * n, 1, 0);
        if (2 * n + 1 >= N) break;
        L(pv[1]);
        hipLaunchKernelGGL(kpm_combine, gv, dim3(256), 0, s, LV, pv[1], pv[0], dres, ld, 0, pv[2]);
        hipLaunchKernelGGL(kpm_moment, dim3(C), dim3(256), 0, s, pv[2], pv[1], dmask, ld, dmom,
                           N, 2 * n + 1, 1, 1);
        rotate();
        ++n;
      }
    }
  };
  if (hipStreamSynchronize(s) != hipSuccess) return fail("upload synchronisation failed");
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
    return fail("event creation failed");
  if (use_graph) {
    if (hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) != hipSuccess)
      return fail("stream capture failed to start");
    enqueue();
    if (hipStreamEndCapture(s, &graph) != hipSuccess) return fail("stream capture failed");
    if (hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) != hipSuccess)
      return fail("graph instantiation failed");
    (void)hipEventRecord(e0, s);
    if (hipGraphLaunch(exec, s) != hipSuccess) return fail("graph launch failed");
  } else {
    (void)hipEventRecord(e0, s);
    enqueue();
  }
  (void)hipEventRecord(e1, s);
  if (hipStreamSynchronize(s) != hipSuccess || hipGetLastError() != hipSuccess)
    return fail("step execution failed");
  float ms = 0.f;
  (void)hipEventElapsedTime(&ms, e0, e1);
  if (elapsed_ms) *elapsed_ms = ms;
  // download the moments and the final recurrence vectors (vm, v)
  ok = hipMemcpy(hmom.data(), dmom, sizeof(double) * C * N, hipMemcpyDeviceToHost) == hipSuccess &&
       hipMemcpy(hvm.data(), pv[0], sizeof(double2) * nv, hipMemcpyDeviceToHost) == hipSuccess &&
       hipMemcpy(hv.data(), pv[1], sizeof(double2) * nv, hipMemcpyDeviceToHost) == hipSuccess;
  if (!ok) return fail("download failed");
  std::copy(hmom.begin(), hmom.end(), moments);
  pos = 0;
  for (int c = 0; c < C; ++c) {
    for (int e = 0; e < b.h_len[c]; ++e) {
      vm[2 * (pos + e)] = hvm[(size_t)c * ld + e].x;
      vm[2 * (pos + e) + 1] = hvm[(size_t)c * ld + e].y;
      v[2 * (pos + e)] = hv[(size_t)c * ld + e].x;
      v[2 * (pos + e) + 1] = hv[(size_t)c * ld + e].y;
    }
    pos += b.h_len[c];
  }
  release();
  return 0;
}

int tdscha_qspace_autotune(tdscha_qctx* ctx, int n_chains, const int* iq_pert,
                           const int* q_pair_map, int reps, int* tiles_out, double* best_ms) {
  TDS_REQUIRE(ctx && iq_pert && q_pair_map, 2, "tdscha_qspace_autotune: NULL argument");
  TDS_REQUIRE(n_chains >= 1 && n_chains <= kQMaxChains, 2, "n_chains must be in [1, 256]");
  QCtx* q = reinterpret_cast<QCtx*>(ctx);
  reps = std::max(1, reps);
  // candidates (d2v form, weights tile, d2v chunk): the slab form with the
  // chunks that give 1 to 4 rounds of its workgroups for this batch, and the
  // per-block form with fixed chunks
  const int saved[4] = {q->weights_tile, q->d2v_chunk, q->reorth_jtile, q->d2v_form};
  std::vector<std::array<int, 3>> cand;
  for (int wt : {64, 128, 256})
    for (int chk : {1024, 2048, 4096, 8192})
      if (chk <= q->N || chk == 1024) cand.push_back({0, wt, std::min(chk, q->N)});
  if (q->S) {
    q->d2v_form = 1;
    QBatch b0;
    if (build_batch(q, n_chains, iq_pert, q_pair_map, b0)) {
      q->d2v_form = saved[3];
      return -1;
    }
    const int nib = b0.nib;
    free_batch(b0);
    std::vector<int> ch;
    for (int r = 1; r <= 4; ++r) {
      const int c = slab_chunk(q, nib, r);
      if (std::find(ch.begin(), ch.end(), c) == ch.end()) ch.push_back(c);
    }
    for (int wt : {64, 128, 256})
      for (int c : ch) cand.push_back({1, wt, c});
  }
  hipStream_t s = q->stream;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
    q->d2v_form = saved[3];
    set_error("tdscha_qspace_autotune: event creation failed");
    return -1;
  }
  double best = 1e300;
  int bw = saved[0], bc = saved[1], bf = saved[3];
  double2 *dpsi = nullptr, *dout = nullptr;
  int rc = 0;
  for (const auto& cd : cand) {
    q->d2v_form = cd[0];
    q->weights_tile = cd[1];
    q->d2v_chunk = cd[2];
    QBatch b;
    rc = build_batch(q, n_chains, iq_pert, q_pair_map, b);
    if (rc) break;
    const size_t nbytes = sizeof(double2) * (size_t)n_chains * b.ld;
    if (!dpsi) {
      if (hipMalloc((void**)&dpsi, nbytes * 2) != hipSuccess ||
          hipMemset(dpsi, 0, nbytes * 2) != hipSuccess) {
        free_batch(b);
        set_error("tdscha_qspace_autotune: allocation failed");
        rc = -1;
        break;
      }
      dout = dpsi + (size_t)n_chains * b.ld;
    }
    launch_L(q, b, dpsi, dout, 0, 0, 0, s);   // warm-up
    (void)hipEventRecord(e0, s);
    for (int r = 0; r < reps; ++r) launch_L(q, b, dpsi, dout, 0, 0, 0, s);
    (void)hipEventRecord(e1, s);
    if (hipStreamSynchronize(s) != hipSuccess || hipGetLastError() != hipSuccess) {
      free_batch(b);
      set_error("tdscha_qspace_autotune: timing run failed");
      rc = -1;
      break;
    }
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    if (ms / reps < best) {
      best = ms / reps;
      bw = cd[1];
      bc = cd[2];
      bf = cd[0];
    }
    free_batch(b);
  }
  if (dpsi) (void)hipFree(dpsi);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc) {
    q->weights_tile = saved[0];
    q->d2v_chunk = saved[1];
    q->reorth_jtile = saved[2];
    q->d2v_form = saved[3];
    return rc;
  }
  q->weights_tile = bw;
  q->d2v_chunk = bc;
  q->d2v_form = bf;
  if (tiles_out) {
    tiles_out[0] = q->weights_tile;
    tiles_out[1] = q->d2v_chunk;
    tiles_out[2] = q->reorth_jtile;
  }
  if (best_ms) *best_ms = best;
  return 0;
}

}  // extern "C"
