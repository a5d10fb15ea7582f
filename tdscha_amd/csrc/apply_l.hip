// Device-resident L application (Modules/DynamicalLanczos.py:3549-3651).
//
//   psi -> (R1, Y1)               prepare_pert   (DL:3009-3053, 2773-2882)
//   perturbed averages            run_pa         (the MFMA kernels)
//   gamma-only translations       project_*      (DL:3174-3191)
//   harmonic + output assembly    assemble_out   (DL:2287-2445, 3282-3363, 3639-3641)
//
// psi layout: [R (n) | sector A: packed upper triangle (n(n+1)/2) | sector B],
// pairs p in row-major upper-triangle order (row r <= col c), w_a = w[c],
// w_b = w[r] (DL:2339-2346).
#include <cmath>
#include <map>
#include <vector>

#include "common.hpp"
#include "../../include/tdscha_amd.h"

namespace tdscha {

int run_pa(Ctx* c, int nc, int apply_d4, int prec, double* f_dev, double* d2v_dev, hipStream_t s);

// coefficient rows of c->std_coef
enum { SC_X = 0, SC_Y = 1, SC_X1 = 2, SC_Y1 = 3 };

static double bose_h(double w, double T) {
  return T > 0 ? 1.0 / (std::exp(w / (T / RY_TO_K)) - 1.0) : 0.0;
}

int setup_pairs(Ctx* c) {
  if (c->pair_row) return 0;
  const int n = c->n;
  const long P = (long)n * (n + 1) / 2;
  std::vector<int> pr(P), pc(P);
  std::vector<double> ca(P), cb(P), sm(P), sp(P), ha(P), hb(P), sc(4 * P);
  std::vector<double> wv(n);
  TDS_CHECK_HIP(hipMemcpy(wv.data(), c->w, sizeof(double) * n, hipMemcpyDeviceToHost));
  long p = 0;
  for (int r = 0; r < n; ++r)
    for (int col = r; col < n; ++col, ++p) {
      pr[p] = r;
      pc[p] = col;
      const double wa = wv[col], wb = wv[r];
      const double na = bose_h(wa, c->T), nb = bose_h(wb, c->T);
      const double X = (1 + 2 * na) * (1 + 2 * nb) / 8;
      const double w2x = (wa * wb) / X;
      const double chm = ((wa - wb) * (na - nb)) / (2 * wa * wb);
      const double chp = ((wa + wb) * (1 + na + nb)) / (2 * wa * wb);
      sm[p] = std::sqrt(-0.5 * chm);
      sp[p] = std::sqrt(0.5 * chp);
      ca[p] = w2x * sm[p];
      cb[p] = w2x * sp[p];
      ha[p] = -(wa * wa + wb * wb - 2.0 * wa * wb);
      hb[p] = -(wa * wa + wb * wb + 2.0 * wa * wb);
      const double d = (2 * na + 1) * (2 * nb + 1);
      sc[SC_X * P + p] = -wa * wa - wb * wb - (2 * wa * wb) / d;
      sc[SC_Y * P + p] = -(8 * wa * wb) / d;
      sc[SC_X1 * P + p] = -(2 * na * nb + na + nb) * (2 * na * nb + na + nb + 1) * (2 * wa * wb) / d;
      sc[SC_Y1 * P + p] = -wa * wa - wb * wb + (2 * wa * wb) / d;
    }
  std::vector<double> yw(c->np, 0.0), rw(c->np, 0.0);
  for (int m = 0; m < n; ++m) {
    const double nm = c->T > 1e-12 ? 1.0 / (std::exp(wv[m] * RY_TO_K / c->T) - 1.0) : 0.0;
    yw[m] = 2 * wv[m] / (2 * nm + 1);
    rw[m] = 2 * wv[m] * nm * (nm + 1) / (2 * nm + 1);
  }
  auto up = [&](void** dst, const void* src, size_t bytes) -> int {
    TDS_CHECK_HIP(hipMalloc(dst, bytes));
    TDS_CHECK_HIP(hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice));
    return 0;
  };
  int rc = 0;
  rc |= up((void**)&c->pair_row, pr.data(), sizeof(int) * P);
  rc |= up((void**)&c->pair_col, pc.data(), sizeof(int) * P);
  rc |= up((void**)&c->pair_ca, ca.data(), sizeof(double) * P);
  rc |= up((void**)&c->pair_cb, cb.data(), sizeof(double) * P);
  rc |= up((void**)&c->pair_sm, sm.data(), sizeof(double) * P);
  rc |= up((void**)&c->pair_sp, sp.data(), sizeof(double) * P);
  rc |= up((void**)&c->pair_harm_a, ha.data(), sizeof(double) * P);
  rc |= up((void**)&c->pair_harm_b, hb.data(), sizeof(double) * P);
  rc |= up((void**)&c->std_coef, sc.data(), sizeof(double) * 4 * P);
  rc |= up((void**)&c->yw, yw.data(), sizeof(double) * c->np);
  rc |= up((void**)&c->reaw, rw.data(), sizeof(double) * c->np);
  if (rc) return -1;
  const long len = n + 2 * P;
  TDS_CHECK_HIP(hipMalloc((void**)&c->psi_dev, sizeof(double) * kMaxChains * len));
  TDS_CHECK_HIP(hipMalloc((void**)&c->psi_out_dev, sizeof(double) * kMaxChains * len));
  TDS_CHECK_HIP(hipMalloc((void**)&c->f_fin, sizeof(double) * kMaxChains * n));
  TDS_CHECK_HIP(hipMalloc((void**)&c->d2v_fin, sizeof(double) * kMaxChains * (long)n * n));
  return 0;
}

// psi -> padded R1 / symmetric Y1 (one thread per padded matrix element)
__global__ void prepare_pert(const double* __restrict__ psi_all, long len, int n, int np,
                             const double* __restrict__ ca, const double* __restrict__ cb,
                             const double* __restrict__ yw, const double* __restrict__ reaw,
                             int use_wigner, int transpose, double* __restrict__ R1p,
                             double* __restrict__ Y1p) {
  const int chain = blockIdx.y;
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long)np * np) return;
  const double* psi = psi_all + chain * len;
  const int a = (int)(e / np), b = (int)(e % np);
  double v = 0.0;
  if (a < n && b < n) {
    const int r = a < b ? a : b, col = a < b ? b : a;
    const long P = (long)n * (n + 1) / 2;
    const long p = (long)r * n - (long)r * (r - 1) / 2 + (col - r);
    const double pa = psi[n + p], pb = psi[n + P + p];
    if (use_wigner) {
      v = ca[p] * pa - cb[p] * pb;
    } else if (!transpose) {
      v = pa;
    } else {
      const double h = (r == col) ? 1.0 : 0.5;
      const double ya = yw[a], yb = yw[b];
      const double cY = 2.0 * (ya * yb * yb + ya * ya * yb);
      const double cRA = 2.0 * ya * yb * (reaw[a] + reaw[b]);
      v = h * pa * cY + h * pb * cRA;
    }
  }
  Y1p[(long)chain * np * np + e] = v;
  if (e < np) R1p[(long)chain * np + e] = (e < n) ? psi[e] : 0.0;
}

// final (normalised, unpadded) f and d2v from the K-split sums
__global__ void finalize_pa(const double* __restrict__ zsum_all, const double* __restrict__ fsum,
                            const double* __restrict__ fy, int n, int np, double norm,
                            double* __restrict__ f_all, double* __restrict__ d2v_all) {
  const int chain = blockIdx.y;
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e < n) f_all[(long)chain * n + e] = fsum[(long)chain * np + e];
  if (e >= (long)n * n) return;
  const int mu = (int)(e / n), nu = (int)(e % n);
  const double* z = zsum_all + (long)chain * np * np;
  d2v_all[(long)chain * n * n + e] =
      norm * (fy[mu] * z[(long)mu * np + nu] + fy[nu] * z[(long)nu * np + mu]);
}

// C = op(A) op(B), FP64, batched (blockIdx.z), on v_mfma_f64_16x16x4f64.
// A 32 x 32 output tile per work-group; the four waves split each 64-deep K
// chunk (16 rows each) and their partial tiles are summed in LDS in wave
// order at the end (fixed order).  The next K chunk is loaded into registers
// while the current one is multiplied.  ta / tb select A^T / B^T; global
// loads run along the contiguous index of each operand.  Operand layout of
// the MFMA: A[i][k]: i = lane % 16, k = lane / 16; B likewise with j;
// D[i][j]: i = 4 r + lane / 16, j = lane % 16 (r = 0..3).
__global__ __launch_bounds__(256) void gemm_f64(int M, int N, int K, const double* __restrict__ A,
                                                long lda, long sA, int ta,
                                                const double* __restrict__ B, long ldb, long sB,
                                                int tb, double* __restrict__ C, long ldc, long sC) {
  const int b = blockIdx.z;
  A += b * sA;
  B += b * sB;
  C += b * sC;
  const int m0 = blockIdx.y * 32, n0 = blockIdx.x * 32;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  __shared__ double As[64][33];
  __shared__ double Bs[64][33];
  __shared__ double red[3][32 * 32];
  double ra[8], rb[8];
  auto load = [&](int k0) {
#pragma unroll
    for (int r = 0; r < 8; ++r) {
      const int e = tid + 256 * r;
      int mm, kk;
      if (ta) { kk = e >> 5; mm = e & 31; } else { mm = e >> 6; kk = e & 63; }
      const int gm = m0 + mm, gk = k0 + kk;
      ra[r] = (gm < M && gk < K) ? (ta ? A[(long)gk * lda + gm] : A[(long)gm * lda + gk]) : 0.0;
      int nn;
      if (tb) { nn = e >> 6; kk = e & 63; } else { kk = e >> 5; nn = e & 31; }
      const int gn = n0 + nn, gk2 = k0 + kk;
      rb[r] = (gn < N && gk2 < K) ? (tb ? B[(long)gn * ldb + gk2] : B[(long)gk2 * ldb + gn]) : 0.0;
    }
  };
  auto store = [&]() {
#pragma unroll
    for (int r = 0; r < 8; ++r) {
      const int e = tid + 256 * r;
      int mm, kk;
      if (ta) { kk = e >> 5; mm = e & 31; } else { mm = e >> 6; kk = e & 63; }
      As[kk][mm] = ra[r];
      int nn;
      if (tb) { nn = e >> 6; kk = e & 63; } else { kk = e >> 5; nn = e & 31; }
      Bs[kk][nn] = rb[r];
    }
  };
  v4d acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = v4d{0, 0, 0, 0};
  load(0);
  for (int k0 = 0; k0 < K; k0 += 64) {
    store();
    __syncthreads();
    if (k0 + 64 < K) load(k0 + 64);
#pragma unroll
    for (int ks = 0; ks < 4; ++ks) {
      const int k = 16 * wave + 4 * ks + (lane >> 4);
      double a[2], bb[2];
#pragma unroll
      for (int i = 0; i < 2; ++i) a[i] = As[k][16 * i + (lane & 15)];
#pragma unroll
      for (int j = 0; j < 2; ++j) bb[j] = Bs[k][16 * j + (lane & 15)];
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i], bb[j], acc[i][j], 0, 0, 0);
    }
    __syncthreads();
  }
  if (wave > 0) {
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r)
          red[wave - 1][(16 * i + 4 * r + (lane >> 4)) * 32 + 16 * j + (lane & 15)] = acc[i][j][r];
  }
  __syncthreads();
  if (wave == 0) {
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int li = 16 * i + 4 * r + (lane >> 4), lj = 16 * j + (lane & 15);
          const double v = acc[i][j][r] + red[0][li * 32 + lj] + red[1][li * 32 + lj] +
                           red[2][li * 32 + lj];
          const int gm = m0 + li, gn = n0 + lj;
          if (gm < M && gn < N) C[(long)gm * ldc + gn] = v;
        }
  }
}

static void gemm(int batch, int M, int N, int K, const double* A, long lda, long sA, int ta,
                 const double* B, long ldb, long sB, int tb, double* C, long ldc, long sC,
                 hipStream_t s) {
  dim3 g((N + 31) / 32, (M + 31) / 32, batch);
  hipLaunchKernelGGL(gemm_f64, g, dim3(256), 0, s, M, N, K, A, lda, sA, ta, B, ldb, sB, tb, C, ldc,
                     sC);
}

// average of the permuted Cartesian matrix over the translations
__global__ void perm_average(const double* __restrict__ Ccart, const int* __restrict__ inv_perms,
                             int n_cart, int n_cells, double* __restrict__ Cavg) {
  const int chain = blockIdx.y;
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long nn = (long)n_cart * n_cart;
  if (e >= nn) return;
  const int J = (int)(e / n_cart), K = (int)(e % n_cart);
  const double* C = Ccart + chain * nn;
  double s = 0.0;
  for (int t = 0; t < n_cells; ++t) {
    const int* ip = inv_perms + (long)t * n_cart;
    s += C[(long)ip[J] * n_cart + ip[K]];
  }
  Cavg[chain * nn + e] = s / n_cells;
}

// The same average when the permutations form a group H: for dof J in the
// orbit of k = rep_dof[jrep[J]] with inv_perms[u](J) = k (u = ju[J]),
//   Cavg[J, K] = S[k][inv_perms[u](K)],  S[k][K'] = mean_v C[ip_v(k), ip_v(K')]
// (substitute t = v o u in the sum over t), so only the n_rep representative
// rows are averaged.
__global__ void perm_rep_sum(const double* __restrict__ Ccart, const int* __restrict__ inv_perms,
                             const int* __restrict__ rep_dof, int n_rep, int n_cart, int n_cells,
                             double* __restrict__ S) {
  const int chain = blockIdx.y;
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long)n_rep * n_cart) return;
  const int r = (int)(e / n_cart), K = (int)(e % n_cart);
  const int k = rep_dof[r];
  const long nn = (long)n_cart * n_cart;
  const double* C = Ccart + chain * nn;
  double s = 0.0;
#pragma unroll 8
  for (int t = 0; t < n_cells; ++t) {
    const int* ip = inv_perms + (long)t * n_cart;
    s += C[(long)ip[k] * n_cart + ip[K]];
  }
  S[(long)chain * n_rep * n_cart + e] = s / n_cells;
}

__global__ void perm_expand(const double* __restrict__ S, const int* __restrict__ inv_perms,
                            const int* __restrict__ jrep, const int* __restrict__ ju, int n_rep,
                            int n_cart, double* __restrict__ Cavg) {
  const int chain = blockIdx.y;
  const long nn = (long)n_cart * n_cart;
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nn) return;
  const int J = (int)(e / n_cart), K = (int)(e % n_cart);
  const int Kp = inv_perms[(long)ju[J] * n_cart + K];
  Cavg[chain * nn + e] = S[((long)chain * n_rep + jrep[J]) * n_cart + Kp];
}

// f <- P_T f  (P_T = mean of the mode-space translation operators); tprojT
// holds P_T transposed so that consecutive threads read consecutive entries
// out = P_T f: a block of 32 output modes, the K = n sum split over 8 thread
// rows and combined in LDS (fixed order)
__global__ __launch_bounds__(256) void project_f(const double* __restrict__ tprojT,
                                                 const double* __restrict__ f_all, int n,
                                                 double* __restrict__ out_all) {
  const int chain = blockIdx.y;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  const int m = blockIdx.x * 32 + tx;
  const double* f = f_all + (long)chain * n;
  double s = 0.0;
  if (m < n) {
#pragma unroll 8
    for (int k = ty; k < n; k += 8) s += tprojT[(long)k * n + m] * f[k];
  }
  __shared__ double red[8][32];
  red[ty][tx] = s;
  __syncthreads();
  if (ty == 0 && m < n) {
    double t = 0.0;
#pragma unroll
    for (int r = 0; r < 8; ++r) t += red[r][tx];
    out_all[(long)chain * n + m] = t;
  }
}

int project_translations(Ctx* c, int nc, hipStream_t s) {
  const int n = c->n, nc_ = c->n_cart;
  // mode_tmp = pols * d2v            (n_cart x n)
  gemm(nc, nc_, n, n, c->pols, n, 0, 0, c->d2v_fin, n, (long)n * n, 0, c->mode_tmp, n,
       (long)nc_ * n, s);
  // cart_tmp = mode_tmp * pols^T     (n_cart x n_cart)
  gemm(nc, nc_, nc_, n, c->mode_tmp, n, (long)nc_ * n, 0, c->pols, n, 0, 1, c->cart_tmp, nc_,
       (long)nc_ * nc_, s);
  const long nn = (long)nc_ * nc_;
  if (c->n_rep > 0) {
    const long nr = (long)c->n_rep * nc_;
    hipLaunchKernelGGL(perm_rep_sum, dim3((unsigned)((nr + 255) / 256), nc), dim3(256), 0, s,
                       c->cart_tmp, c->inv_perms, c->rep_dof, c->n_rep, nc_, c->n_cells,
                       c->rep_sum);
    hipLaunchKernelGGL(perm_expand, dim3((unsigned)((nn + 255) / 256), nc), dim3(256), 0, s,
                       c->rep_sum, c->inv_perms, c->jrep, c->ju, c->n_rep, nc_, c->cart_tmp2);
  } else {
    hipLaunchKernelGGL(perm_average, dim3((unsigned)((nn + 255) / 256), nc), dim3(256), 0, s,
                       c->cart_tmp, c->inv_perms, nc_, c->n_cells, c->cart_tmp2);
  }
  // mode_tmp = pols^T * Cavg         (n x n_cart)
  gemm(nc, n, nc_, nc_, c->pols, n, 0, 1, c->cart_tmp2, nc_, nn, 0, c->mode_tmp, nc_,
       (long)n * nc_, s);
  // d2v = mode_tmp * pols            (n x n)
  gemm(nc, n, n, nc_, c->mode_tmp, nc_, (long)n * nc_, 0, c->pols, n, 0, 0, c->d2v_fin, n,
       (long)n * n, s);
  hipLaunchKernelGGL(project_f, dim3((n + 31) / 32, nc), dim3(256), 0, s, c->tproj, c->f_fin, n,
                     c->f_proj);
  return 0;
}

// out = [L1 + L_anh](psi), then reverse/shift
__global__ void assemble_out(const double* __restrict__ psi_all, long len, int n,
                             const double* __restrict__ w, const double* __restrict__ f_all,
                             const double* __restrict__ d2v_all, const int* __restrict__ prow,
                             const int* __restrict__ pcol, const double* __restrict__ sm,
                             const double* __restrict__ sp, const double* __restrict__ ha,
                             const double* __restrict__ hb, const double* __restrict__ sc,
                             const double* __restrict__ yw, const double* __restrict__ reaw,
                             int use_wigner, int transpose, int harmonic, int anharmonic,
                             int reverse, double shift, double* __restrict__ out_all) {
  const int chain = blockIdx.y;
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= len) return;
  const double* psi = psi_all + chain * len;
  const double* f = f_all + (long)chain * n;
  const double* d2v = d2v_all + (long)chain * n * n;
  const long P = (long)n * (n + 1) / 2;
  double v = 0.0;
  if (e < n) {
    const double wm = w[e];
    if (harmonic) v = use_wigner ? -(psi[e] * wm) * wm : (psi[e] * wm) * wm;
    if (anharmonic) v += use_wigner ? f[e] : -f[e];
  } else {
    const bool secA = e < n + P;
    const long p = secA ? e - n : e - n - P;
    const int r = prow[p], col = pcol[p];
    const double pa = psi[n + p], pb = psi[n + P + p];
    if (use_wigner) {
      if (harmonic) v = secA ? ha[p] * pa : hb[p] * pb;
      if (anharmonic) {
        const double d = d2v[(long)r * n + col];
        v += secA ? sm[p] * d : -sp[p] * d;
      }
    } else {
      if (harmonic) {
        const double X = sc[SC_X * P + p], Y = sc[SC_Y * P + p];
        const double X1 = sc[SC_X1 * P + p], Y1 = sc[SC_Y1 * P + p];
        if (!transpose)
          v = secA ? (-X * pa - Y * pb) : (-X1 * pa - Y1 * pb);
        else
          v = secA ? (-X * pa - X1 * pb) : (-Y * pa - Y1 * pb);
      }
      if (anharmonic) {
        const double d = d2v[(long)r * n + col];
        double pert;
        if (!transpose)
          pert = secA ? d * (yw[r] + yw[col]) : d * (reaw[r] + reaw[col]);
        else
          pert = secA ? 0.5 * d / (yw[r] * yw[col]) * (r == col ? 1.0 : 2.0) : 0.0;
        v -= pert;
      }
    }
  }
  if (reverse) v = -v;
  out_all[chain * len + e] = v + shift * psi[e];
}

int apply_L_dev(Ctx* c, int nc, const double* psi, double* out, int use_wigner, int transpose,
                int ignore_harmonic, int ignore_v3, int ignore_v4, int reverse, double shift,
                int precision, hipStream_t s) {
  int rc = setup_pairs(c);
  if (rc) return rc;
  const int n = c->n, np = c->np;
  const long P = (long)n * (n + 1) / 2;
  const long len = n + 2 * P;
  const int anharmonic = ignore_v3 ? 0 : 1;
  if (anharmonic) {
    const long nn = (long)np * np;
    hipLaunchKernelGGL(prepare_pert, dim3((unsigned)((nn + 255) / 256), nc), dim3(256), 0, s, psi,
                       len, n, np, c->pair_ca, c->pair_cb, c->yw, c->reaw, use_wigner, transpose,
                       c->R1, c->Y1);
    rc = run_pa(c, nc, ignore_v4 ? 0 : 1, precision, c->f_fin, c->d2v_fin, s);
    if (rc) return rc;
    if (c->n_cells > 0) {
      rc = project_translations(c, nc, s);
      if (rc) return rc;
    }
  }
  const double* f_src = (anharmonic && c->n_cells > 0) ? c->f_proj : c->f_fin;
  hipLaunchKernelGGL(assemble_out, dim3((unsigned)((len + 255) / 256), nc), dim3(256), 0, s, psi,
                     len, n, c->w, f_src, c->d2v_fin, c->pair_row, c->pair_col, c->pair_sm,
                     c->pair_sp, c->pair_harm_a, c->pair_harm_b, c->std_coef, c->yw, c->reaw,
                     use_wigner, transpose, ignore_harmonic ? 0 : 1, anharmonic, reverse, shift,
                     out);
  TDS_CHECK_HIP(hipGetLastError());
  return 0;
}

// Inverse of the harmonic 2x2 pair blocks (Modules/DynamicalLanczos.py:2451-2549),
// evaluated once per context on the host in FP64 with the reference's
// regularisation: populations set to zero where 0.01 w_a RY_TO_K > T (the
// mask of both members uses w_a, DL:2491-2492), a determinant below 1e-12
// keeps only the inverse of the X element, and at T > 0 equal frequencies
// with a thermal population invert the Y1 element alone.
static int setup_inverse(Ctx* c) {
  if (c->inv_coef) return 0;
  const int n = c->n;
  const long P = (long)n * (n + 1) / 2;
  std::vector<double> wv(n), ic(4 * P);
  TDS_CHECK_HIP(hipMemcpy(wv.data(), c->w, sizeof(double) * n, hipMemcpyDeviceToHost));
  const double eps = 1e-12;
  long p = 0;
  for (int r = 0; r < n; ++r)
    for (int col = r; col < n; ++col, ++p) {
      const double wa = wv[col], wb = wv[r];
      double na = 0.0, nb = 0.0;
      if (c->T > 0) {
        na = 1.0 / (std::exp(wa / (c->T / RY_TO_K)) - 1.0);
        nb = 1.0 / (std::exp(wb / (c->T / RY_TO_K)) - 1.0);
        if (0.01 * wa * RY_TO_K > c->T) na = nb = 0.0;
      }
      const double d = (2 * na + 1) * (2 * nb + 1);
      const double X = -wa * wa - wb * wb - (2 * wa * wb) / d;
      const double Y = -(8 * wa * wb) / d;
      const double X1 = -(2 * na * nb + na + nb) * (2 * na * nb + na + nb + 1) * (2 * wa * wb) / d;
      const double Y1 = -wa * wa - wb * wb + (2 * wa * wb) / d;
      const double den = X * Y1 - X1 * Y;
      double xn, yn, x1n, y1n;
      if (den < eps) {       // den -> inf in the reference
        xn = -1.0 / X;
        yn = 0.0;
        x1n = 0.0;
        y1n = 0.0;
      } else {
        xn = -Y1 / den;
        yn = Y / den;
        x1n = X1 / den;
        y1n = -X / den;
      }
      if (c->T > eps && wa == wb && na > eps) y1n = -1.0 / Y1;
      ic[0 * P + p] = xn;
      ic[1 * P + p] = yn;
      ic[2 * P + p] = x1n;
      ic[3 * P + p] = y1n;
    }
  TDS_CHECK_HIP(hipMalloc((void**)&c->inv_coef, sizeof(double) * 4 * P));
  TDS_CHECK_HIP(hipMemcpy(c->inv_coef, ic.data(), sizeof(double) * 4 * P, hipMemcpyHostToDevice));
  return 0;
}

// out = L1^-1 psi (or its transpose); one thread per psi entry
__global__ void apply_l1_inverse(const double* __restrict__ psi_all, long len, int n,
                                 const double* __restrict__ w, const double* __restrict__ ic,
                                 int transpose, int ignore_harmonic, double* __restrict__ out_all) {
  const int chain = blockIdx.y;
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= len) return;
  const double* psi = psi_all + chain * len;
  double* out = out_all + chain * len;
  if (ignore_harmonic) {
    out[e] = 0.0;
    return;
  }
  if (e < n) {
    out[e] = (psi[e] / w[e]) / w[e];
    return;
  }
  const long P = (long)n * (n + 1) / 2;
  const bool secY = e < n + P;
  const long p = secY ? e - n : e - n - P;
  const double py = psi[n + p], pa = psi[n + P + p];
  const double xn = ic[p], yn = ic[P + p], x1n = ic[2 * P + p], y1n = ic[3 * P + p];
  double v;
  if (secY)
    v = xn * py + (transpose ? x1n * pa : yn * pa);
  else
    v = (transpose ? yn * py : x1n * py) + y1n * pa;
  out[e] = v;
}

}  // namespace tdscha

using namespace tdscha;

extern "C" {

int tdscha_apply_L1_inverse(tdscha_ctx* ctx, int n_vec, const double* psi, double* out,
                            int transpose, int ignore_harmonic) {
  TDS_REQUIRE(ctx != nullptr, 2, "tdscha_apply_L1_inverse: NULL context");
  TDS_REQUIRE(psi && out, 2, "tdscha_apply_L1_inverse: NULL pointer");
  TDS_REQUIRE(n_vec >= 1 && n_vec <= kMaxChains, 2, "n_vec must be in [1, 8]");
  Ctx* c = reinterpret_cast<Ctx*>(ctx);
  int rc = setup_pairs(c);
  if (rc) return rc;
  rc = setup_inverse(c);
  if (rc) return rc;
  const long len = c->n + (long)c->n * (c->n + 1);
  TDS_CHECK_HIP(hipMemcpyAsync(c->psi_dev, psi, sizeof(double) * n_vec * len,
                               hipMemcpyHostToDevice, c->stream));
  hipLaunchKernelGGL(apply_l1_inverse, dim3((unsigned)((len + 255) / 256), n_vec), dim3(256), 0,
                     c->stream, c->psi_dev, len, c->n, c->w, c->inv_coef, transpose,
                     ignore_harmonic, c->psi_out_dev);
  TDS_CHECK_HIP(hipGetLastError());
  TDS_CHECK_HIP(hipMemcpyAsync(out, c->psi_out_dev, sizeof(double) * n_vec * len,
                               hipMemcpyDeviceToHost, c->stream));
  TDS_CHECK_HIP(hipStreamSynchronize(c->stream));
  return 0;
}

int tdscha_set_translations(tdscha_ctx* ctx, int n_cart, const double* pols, int n_cells,
                            const int* inv_perms, const double* trans_projector) {
  TDS_REQUIRE(ctx != nullptr, 2, "tdscha_set_translations: NULL context");
  Ctx* c = reinterpret_cast<Ctx*>(ctx);
  if (n_cells <= 0) {
    c->n_cells = 0;
    return 0;
  }
  TDS_REQUIRE(pols && inv_perms && trans_projector && n_cart > 0, 2,
              "tdscha_set_translations: NULL argument");
  for (long e = 0; e < (long)n_cells * n_cart; ++e)
    TDS_REQUIRE(inv_perms[e] >= 0 && inv_perms[e] < n_cart, 2,
                "tdscha_set_translations: permutation index out of range");
  const int n = c->n;
  if (c->pols) {
    hipFree(c->pols); hipFree(c->inv_perms); hipFree(c->tproj); hipFree(c->cart_tmp);
    hipFree(c->cart_tmp2); hipFree(c->mode_tmp); hipFree(c->f_proj);
    c->pols = nullptr;
  }
  if (c->rep_dof) {
    hipFree(c->rep_dof); hipFree(c->jrep); hipFree(c->ju); hipFree(c->rep_sum);
    c->rep_dof = c->jrep = c->ju = nullptr;
    c->rep_sum = nullptr;
  }
  c->n_rep = 0;
  // orbit tables when the permutations form a group (closed under
  // composition); otherwise the direct average is used
  {
    std::map<std::vector<int>, int> index;
    for (int t = 0; t < n_cells; ++t)
      index[std::vector<int>(inv_perms + (long)t * n_cart, inv_perms + (long)(t + 1) * n_cart)] = t;
    bool group = (int)index.size() == n_cells;
    std::vector<int> comp(n_cart);
    for (int t1 = 0; group && t1 < n_cells; ++t1)
      for (int t2 = 0; group && t2 < n_cells; ++t2) {
        for (int i = 0; i < n_cart; ++i)
          comp[i] = inv_perms[(long)t1 * n_cart + inv_perms[(long)t2 * n_cart + i]];
        group = index.count(comp) > 0;
      }
    if (group) {
      std::vector<int> jr(n_cart), jt(n_cart), reps, rep_of(n_cart, -1);
      for (int J = 0; J < n_cart; ++J) {
        int best = n_cart, bt = 0;
        for (int t = 0; t < n_cells; ++t) {
          const int k = inv_perms[(long)t * n_cart + J];
          if (k < best) { best = k; bt = t; }
        }
        if (rep_of[best] < 0) {
          rep_of[best] = (int)reps.size();
          reps.push_back(best);
        }
        jr[J] = rep_of[best];
        jt[J] = bt;
      }
      c->n_rep = (int)reps.size();
      TDS_CHECK_HIP(hipMalloc((void**)&c->rep_dof, sizeof(int) * c->n_rep));
      TDS_CHECK_HIP(hipMalloc((void**)&c->jrep, sizeof(int) * n_cart));
      TDS_CHECK_HIP(hipMalloc((void**)&c->ju, sizeof(int) * n_cart));
      TDS_CHECK_HIP(hipMalloc((void**)&c->rep_sum,
                              sizeof(double) * kMaxChains * (long)c->n_rep * n_cart));
      TDS_CHECK_HIP(hipMemcpy(c->rep_dof, reps.data(), sizeof(int) * c->n_rep,
                              hipMemcpyHostToDevice));
      TDS_CHECK_HIP(hipMemcpy(c->jrep, jr.data(), sizeof(int) * n_cart, hipMemcpyHostToDevice));
      TDS_CHECK_HIP(hipMemcpy(c->ju, jt.data(), sizeof(int) * n_cart, hipMemcpyHostToDevice));
    }
  }
  c->n_cart = n_cart;
  TDS_CHECK_HIP(hipMalloc((void**)&c->pols, sizeof(double) * (long)n_cart * n));
  TDS_CHECK_HIP(hipMalloc((void**)&c->inv_perms, sizeof(int) * (long)n_cells * n_cart));
  TDS_CHECK_HIP(hipMalloc((void**)&c->tproj, sizeof(double) * (long)n * n));
  TDS_CHECK_HIP(hipMalloc((void**)&c->cart_tmp, sizeof(double) * kMaxChains * (long)n_cart * n_cart));
  TDS_CHECK_HIP(hipMalloc((void**)&c->cart_tmp2, sizeof(double) * kMaxChains * (long)n_cart * n_cart));
  TDS_CHECK_HIP(hipMalloc((void**)&c->mode_tmp, sizeof(double) * kMaxChains * (long)n_cart * n));
  TDS_CHECK_HIP(hipMalloc((void**)&c->f_proj, sizeof(double) * kMaxChains * n));
  TDS_CHECK_HIP(hipMemcpy(c->pols, pols, sizeof(double) * (long)n_cart * n, hipMemcpyHostToDevice));
  TDS_CHECK_HIP(hipMemcpy(c->inv_perms, inv_perms, sizeof(int) * (long)n_cells * n_cart,
                          hipMemcpyHostToDevice));
  std::vector<double> tT((size_t)n * n);
  for (int a = 0; a < n; ++a)
    for (int b = 0; b < n; ++b) tT[(size_t)b * n + a] = trans_projector[(size_t)a * n + b];
  TDS_CHECK_HIP(hipMemcpy(c->tproj, tT.data(), sizeof(double) * (long)n * n,
                          hipMemcpyHostToDevice));
  c->n_cells = n_cells;
  return 0;
}

int tdscha_apply_L_dev(tdscha_ctx* ctx, int n_vec, const double* psi_dev, double* out_dev,
                       int use_wigner, int transpose, int ignore_harmonic, int ignore_v3,
                       int ignore_v4, int reverse, double shift, int precision, void* stream) {
  TDS_REQUIRE(ctx != nullptr, 2, "tdscha_apply_L_dev: NULL context");
  TDS_REQUIRE(n_vec >= 1 && n_vec <= kMaxChains, 2, "n_vec must be in [1, 8]");
  TDS_REQUIRE(psi_dev && out_dev && psi_dev != out_dev, 2,
              "tdscha_apply_L_dev: psi and out must be distinct device buffers");
  TDS_REQUIRE(precision >= 0 && precision <= 2, 2, "precision must be 0, 1 or 2");
  Ctx* c = reinterpret_cast<Ctx*>(ctx);
  hipStream_t s = stream ? reinterpret_cast<hipStream_t>(stream) : c->stream;
  return apply_L_dev(c, n_vec, psi_dev, out_dev, use_wigner, transpose, ignore_harmonic, ignore_v3,
                     ignore_v4, reverse, shift, precision, s);
}

int tdscha_apply_L(tdscha_ctx* ctx, int n_vec, const double* psi, double* out, int use_wigner,
                   int transpose, int ignore_harmonic, int ignore_v3, int ignore_v4, int reverse,
                   double shift, int precision) {
  TDS_REQUIRE(ctx != nullptr, 2, "tdscha_apply_L: NULL context");
  TDS_REQUIRE(psi && out, 2, "tdscha_apply_L: NULL pointer");
  TDS_REQUIRE(n_vec >= 1 && n_vec <= kMaxChains, 2, "n_vec must be in [1, 8]");
  Ctx* c = reinterpret_cast<Ctx*>(ctx);
  int rc = setup_pairs(c);
  if (rc) return rc;
  const long len = c->n + (long)c->n * (c->n + 1);
  TDS_CHECK_HIP(hipMemcpyAsync(c->psi_dev, psi, sizeof(double) * n_vec * len,
                               hipMemcpyHostToDevice, c->stream));
  rc = tdscha_apply_L_dev(ctx, n_vec, c->psi_dev, c->psi_out_dev, use_wigner, transpose,
                          ignore_harmonic, ignore_v3, ignore_v4, reverse, shift, precision,
                          c->stream);
  if (rc) return rc;
  TDS_CHECK_HIP(hipMemcpyAsync(out, c->psi_out_dev, sizeof(double) * n_vec * len,
                               hipMemcpyDeviceToHost, c->stream));
  TDS_CHECK_HIP(hipStreamSynchronize(c->stream));
  return 0;
}

}  // extern "C"
