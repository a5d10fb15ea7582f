// Device-resident symmetric Lanczos for a batch of chains.
//
// The Wigner-representation recurrence of run_FT with run_simm
// (Modules/DynamicalLanczos.py:6032-6240, p = q, metric = mask_dot_wigner,
// DL:5905-5954), batched over up to 64 chains (L applied 8 at a time),
// optionally with full classical Gram-Schmidt reorthogonalisation repeated
// twice (the real-space analogue of the q-space full reorthogonalisation,
// Modules/QSpaceLanczos.py:1113-1130).  Chains stop independently at
// |b| < tol or at their own step cap; every scalar lives on the device so
// one step is a fixed kernel sequence: captured once as a hipGraph and
// replayed, or the whole run captured and launched once.  Reductions are
// two-stage with a fixed order (deterministic).
#include <cmath>
#include <vector>

#include "common.hpp"
#include "../../include/tdscha_amd.h"

namespace tdscha {

int apply_L_dev(Ctx* c, int nc, const double* psi, double* out, int use_wigner, int transpose,
                int ignore_harmonic, int ignore_v3, int ignore_v4, int reverse, double shift,
                int precision, hipStream_t s);
int setup_pairs(Ctx* c);

constexpr int kRedBlocks = 128;   // blocks per chain in the partial reductions
constexpr int kRedThreads = 256;
constexpr int kMaxRunChains = 64; // chains of one device Lanczos run

// L applied to nc chains of the run, kMaxChains at a time (the perturbed
// averages take up to kMaxChains per pass)
static void apply_L_groups(Ctx* c, int nc, long len, const double* in, double* out,
                           int use_wigner, int transpose, int ignore_harmonic, int ignore_v3,
                           int ignore_v4, int reverse, double shift, int precision,
                           hipStream_t s) {
  for (int c0 = 0; c0 < nc; c0 += kMaxChains) {
    const int g = nc - c0 < kMaxChains ? nc - c0 : kMaxChains;
    apply_L_dev(c, g, in + (long)c0 * len, out + (long)c0 * len, use_wigner, transpose,
                ignore_harmonic, ignore_v3, ignore_v4, reverse, shift, precision, s);
  }
}

struct LanczosState {
  int nc = 0, kmax = 0;
  long len = 0;
  double* Q = nullptr;       // [nc, kmax+1, len]
  double* qcur = nullptr;    // [nc, len]
  double* Lq = nullptr;      // [nc, len]
  double* r = nullptr;       // [nc, len]
  double* mask = nullptr;    // [len]
  double* part = nullptr;    // [nc, kmax+1, kRedBlocks]
  double* coef = nullptr;    // [nc, kmax+1]
  double* a = nullptr;       // [nc, kmax]
  double* b = nullptr;       // [nc, kmax]
  double* scal = nullptr;    // [nc, 4]: a_k, b_k, scratch
  int* k = nullptr;          // [nc] current step
  int* active = nullptr;     // [nc]
  int* act2 = nullptr;       // [nc] chains whose second Gram-Schmidt pass runs
  int* cap = nullptr;        // [nc]
  int* ndone = nullptr;      // [nc] number of a coefficients
  double* upart = nullptr;   // [nc, jsplit, len] basis-range partial updates (short vectors)
  int nslice = 1;            // element slices of the projections
  int jsplit = 1;            // basis-vector ranges of the update
  int jchunk = 1;
};

static void free_state(LanczosState& s) {
  void* p[] = {s.Q, s.qcur, s.Lq, s.r, s.mask, s.part, s.coef, s.a, s.b, s.scal, s.k, s.active,
               s.cap, s.ndone, s.upart, s.act2};
  for (void* x : p)
    if (x) (void)hipFree(x);
}

__global__ void build_mask(int n, long len, double* __restrict__ mask) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= len) return;
  double v = 1.0;
  if (e >= n) {
    const long P = (long)n * (n + 1) / 2;
    long p = e - n;
    if (p >= P) p -= P;
    // row r of pair p: largest r with r*n - r(r-1)/2 <= p
    long r = 0, start = 0;
    long lo = 0, hi = n - 1;
    while (lo < hi) {
      const long mid = (lo + hi + 1) / 2;
      const long st = mid * n - mid * (mid - 1) / 2;
      if (st <= p) lo = mid; else hi = mid - 1;
    }
    r = lo;
    start = r * n - r * (r - 1) / 2;
    v = (p == start) ? 1.0 : 2.0;   // the first pair of each row is the diagonal
  }
  mask[e] = v;
}

// partial sums of <x_c, y_c>_M over a grid-stride slice -> part[c, slot, blk]
__global__ void dot_partial(const double* __restrict__ x, long xs, const double* __restrict__ y,
                            long ys, const double* __restrict__ mask, long len,
                            double* __restrict__ part, long ps, int slot) {
  const int c = blockIdx.y;
  const double* xc = x + c * xs;
  const double* yc = y + c * ys;
  double s = 0.0;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < len;
       i += (long)gridDim.x * blockDim.x)
    s += xc[i] * yc[i] * mask[i];
  __shared__ double red[kRedThreads];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int o = kRedThreads / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[c * ps + (long)slot * kRedBlocks + blockIdx.x] = red[0];
}

__device__ double sum_partials(const double* p) {
  double s = 0.0;
  for (int i = 0; i < kRedBlocks; ++i) s += p[i];
  return s;
}

// a_k = <q, Lq>_M ; r = Lq - a q - b_{k-1} q_{k-1}
__global__ void lanczos_residual(const double* __restrict__ part, long ps,
                                 const double* __restrict__ Q, const double* __restrict__ qcur,
                                 const double* __restrict__ Lq, long len, int kmax,
                                 const int* __restrict__ kk, const int* __restrict__ active,
                                 double* __restrict__ a, const double* __restrict__ b,
                                 double* __restrict__ scal, double* __restrict__ r) {
  const int c = blockIdx.y;
  if (!active[c]) return;
  const int k = kk[c];
  const double ak = sum_partials(part + c * ps);   // recomputed per block: same value
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    a[c * kmax + k] = ak;
    scal[c * 4 + 0] = ak;
  }
  if (i >= len) return;
  double v = Lq[c * len + i] - ak * qcur[c * len + i];
  if (k > 0) v -= b[c * kmax + k - 1] * Q[((long)c * (kmax + 1) + (k - 1)) * len + i];
  r[c * len + i] = v;
}

__device__ __forceinline__ double wave_sum64(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// CGS projections <Q_j, r>_M for j <= k: one work-group per (element slice,
// tile of JT basis vectors, chain); partial sums part[c][j][slice]
template <int JT>
__global__ void __launch_bounds__(256) reorth_dots(const double* __restrict__ Q,
                                                   const double* __restrict__ r,
                                                   const double* __restrict__ mask, long len,
                                                   int kmax, const int* __restrict__ kk,
                                                   const int* __restrict__ active, int nslice,
                                                   double* __restrict__ part) {
  const int c = blockIdx.z, jt = blockIdx.y, sl = blockIdx.x, t = threadIdx.x;
  if (!active[c]) return;
  const int k = kk[c];
  const int j0 = jt * JT;
  if (j0 > k) return;
  const int nj = min(JT, k + 1 - j0);
  const long slice = ((len + nslice - 1) / nslice + 255) / 256 * 256;
  const long lo = (long)sl * slice;
  const long hi = min(len, lo + slice);
  double s[JT];
#pragma unroll
  for (int j = 0; j < JT; ++j) s[j] = 0.0;
  const double* rc = r + (long)c * len;
  const double* Qc = Q + ((long)c * (kmax + 1) + j0) * len;
  for (long e = lo + t; e < hi; e += 256) {
    const double rm = rc[e] * mask[e];
#pragma unroll
    for (int j = 0; j < JT; ++j)
      if (j < nj) s[j] += Qc[(long)j * len + e] * rm;
  }
  __shared__ double red[4][JT];
#pragma unroll
  for (int j = 0; j < JT; ++j) {
    const double v = wave_sum64(s[j]);
    if ((t & 63) == 0) red[t >> 6][j] = v;
  }
  __syncthreads();
  if (t < nj)
    part[((long)c * (kmax + 1) + j0 + t) * kRedBlocks + sl] =
        red[0][t] + red[1][t] + red[2][t] + red[3][t];
}

__global__ void reorth_coef(const double* __restrict__ part, int kmax, const int* __restrict__ kk,
                            const int* __restrict__ active, int nslice,
                            double* __restrict__ coef) {
  const int c = blockIdx.y;
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (!active[c] || j > kk[c]) return;
  const double* p = part + ((long)c * (kmax + 1) + j) * kRedBlocks;
  double s = 0.0;
  for (int i = 0; i < nslice; ++i) s += p[i];
  coef[c * (kmax + 1) + j] = s;
}

// r -= sum_j coef_j Q_j.  jsplit == 1: in place.  jsplit > 1 (short vectors):
// each grid slice z sums its range of basis vectors into upart[c][z], and
// reorth_finish subtracts the ranges in order.
__global__ void reorth_update(const double* __restrict__ Q, const double* __restrict__ coef,
                              long len, int kmax, const int* __restrict__ kk,
                              const int* __restrict__ active, int jsplit, int jchunk,
                              double* __restrict__ upart, double* __restrict__ r) {
  const int c = blockIdx.y, z = blockIdx.z;
  if (!active[c]) return;
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= len) return;
  const int k = kk[c];
  const int ja = z * jchunk, jb = min(k + 1, ja + jchunk);
  double v = 0.0;
  for (int j = ja; j < jb; ++j)
    v += coef[c * (kmax + 1) + j] * Q[((long)c * (kmax + 1) + j) * len + i];
  if (jsplit == 1)
    r[c * len + i] -= v;
  else
    upart[((long)c * jsplit + z) * len + i] = v;
}

__global__ void reorth_finish(const double* __restrict__ upart, long len, int jsplit,
                              const int* __restrict__ active, double* __restrict__ r) {
  const int c = blockIdx.y;
  if (!active[c]) return;
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= len) return;
  double v = 0.0;
  for (int z = 0; z < jsplit; ++z) v += upart[((long)c * jsplit + z) * len + i];
  r[c * len + i] -= v;
}

// Second Gram-Schmidt pass only where the first one removed more than half
// of the residual's squared norm (the Daniel-Gragg-Kaufman-Stewart test,
// |r'| < |r| / sqrt(2)): the norm before the pass is kept in scal[c][2], the
// norm after it is summed from partial row `slot`.
__global__ void keep_norm(const double* __restrict__ part, long ps, int slot,
                          const int* __restrict__ active, double* __restrict__ scal, int nc) {
  const int c = threadIdx.x;
  if (c >= nc || !active[c]) return;
  scal[c * 4 + 2] = sum_partials(part + c * ps + (long)slot * kRedBlocks);
}

__global__ void second_pass_flags(const double* __restrict__ part, long ps, int slot,
                                  const int* __restrict__ active,
                                  const double* __restrict__ scal, int* __restrict__ act2,
                                  int nc) {
  const int c = threadIdx.x;
  if (c >= nc) return;
  const double after = sum_partials(part + c * ps + (long)slot * kRedBlocks);
  act2[c] = (active[c] && after < 0.5 * scal[c * 4 + 2]) ? 1 : 0;
}

// b_k = ||r||_M ; stop or normalise into Q_{k+1} / qcur ; advance k
__global__ void lanczos_advance(const double* __restrict__ part, long ps, long len, int kmax,
                                double tol, int* __restrict__ kk, int* __restrict__ active,
                                const int* __restrict__ cap, int* __restrict__ ndone,
                                double* __restrict__ b, const double* __restrict__ r,
                                double* __restrict__ Q, double* __restrict__ qcur) {
  const int c = blockIdx.y;
  if (!active[c]) return;
  const int k = kk[c];
  const double bk = sqrt(sum_partials(part + c * ps));
  const bool stop = !(fabs(bk) >= tol) || (k + 1 >= cap[c]);
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (!stop && i < len) {
    const double v = r[c * len + i] / bk;
    Q[((long)c * (kmax + 1) + k + 1) * len + i] = v;
    qcur[c * len + i] = v;
  }
}

// separate tiny kernel so that every block above saw the old k/active
__global__ void lanczos_commit(const double* __restrict__ part, long ps, int kmax, double tol,
                               int* __restrict__ kk, int* __restrict__ active,
                               const int* __restrict__ cap, int* __restrict__ ndone,
                               double* __restrict__ b) {
  const int c = threadIdx.x;
  if (!active[c]) return;
  const int k = kk[c];
  const double bk = sqrt(sum_partials(part + c * ps));
  const bool small = !(fabs(bk) >= tol);
  ndone[c] = k + 1;
  if (small || k + 1 >= cap[c]) {
    if (!small) b[c * kmax + k] = bk;
    active[c] = 0;
  } else {
    b[c * kmax + k] = bk;
    kk[c] = k + 1;
  }
}

__global__ void lanczos_init(const double* __restrict__ part, long ps,
                             const double* __restrict__ psi0, long len, int kmax,
                             double* __restrict__ Q, double* __restrict__ qcur) {
  const int c = blockIdx.y;
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= len) return;
  const double nrm = sqrt(sum_partials(part + c * ps));
  const double v = psi0[c * len + i] / nrm;
  Q[((long)c * (kmax + 1)) * len + i] = v;
  qcur[c * len + i] = v;
}

static void lanczos_step(Ctx* c, LanczosState& st, int reorth, int precision, int ignore_harmonic,
                         int ignore_v3, int ignore_v4, double tol, hipStream_t s) {
  const long len = st.len;
  const int nc = st.nc, kmax = st.kmax;
  const long ps = (long)(kmax + 1) * kRedBlocks;
  const dim3 gv((unsigned)((len + 255) / 256), nc);
  const dim3 gr(kRedBlocks, nc);
  apply_L_groups(c, nc, len, st.qcur, st.Lq, 1, 0, ignore_harmonic, ignore_v3, ignore_v4, 0,
                 0.0, precision, s);
  hipLaunchKernelGGL(dot_partial, gr, dim3(kRedThreads), 0, s, st.qcur, len, st.Lq, len, st.mask,
                     len, st.part, ps, 0);
  hipLaunchKernelGGL(lanczos_residual, gv, dim3(256), 0, s, st.part, ps, st.Q, st.qcur, st.Lq,
                     len, kmax, st.k, st.active, st.a, st.b, st.scal, st.r);
  if (reorth) {
    // classical Gram-Schmidt against the whole basis; the second pass runs
    // for the chains whose first pass failed the norm test (flags in act2).
    // Partial row kmax is never a basis row, so it holds the two norms.
    constexpr int JT = 16;
    hipLaunchKernelGGL(dot_partial, gr, dim3(kRedThreads), 0, s, st.r, len, st.r, len, st.mask,
                       len, st.part, ps, kmax);
    hipLaunchKernelGGL(keep_norm, dim3(1), dim3(64), 0, s, st.part, ps, kmax, st.active, st.scal,
                       nc);
    for (int pass = 0; pass < 2; ++pass) {
      const int* act = pass == 0 ? st.active : st.act2;
      if (pass == 1) {
        hipLaunchKernelGGL(dot_partial, gr, dim3(kRedThreads), 0, s, st.r, len, st.r, len,
                           st.mask, len, st.part, ps, kmax);
        hipLaunchKernelGGL(second_pass_flags, dim3(1), dim3(64), 0, s, st.part, ps, kmax,
                           st.active, st.scal, st.act2, nc);
      }
      hipLaunchKernelGGL(reorth_dots<JT>, dim3(st.nslice, (kmax + JT) / JT, nc), dim3(256), 0, s,
                         st.Q, st.r, st.mask, len, kmax, st.k, act, st.nslice, st.part);
      hipLaunchKernelGGL(reorth_coef, dim3((kmax + 1 + 63) / 64, nc), dim3(64), 0, s, st.part,
                         kmax, st.k, act, st.nslice, st.coef);
      hipLaunchKernelGGL(reorth_update, dim3(gv.x, nc, st.jsplit), dim3(256), 0, s, st.Q,
                         st.coef, len, kmax, st.k, act, st.jsplit, st.jchunk, st.upart, st.r);
      if (st.jsplit > 1)
        hipLaunchKernelGGL(reorth_finish, gv, dim3(256), 0, s, st.upart, len, st.jsplit, act,
                           st.r);
    }
  }
  hipLaunchKernelGGL(dot_partial, gr, dim3(kRedThreads), 0, s, st.r, len, st.r, len, st.mask, len,
                     st.part, ps, 0);
  hipLaunchKernelGGL(lanczos_advance, gv, dim3(256), 0, s, st.part, ps, len, kmax, tol, st.k,
                     st.active, st.cap, st.ndone, st.b, st.r, st.Q, st.qcur);
  hipLaunchKernelGGL(lanczos_commit, dim3(1, 1), dim3(nc), 0, s, st.part, ps, kmax, tol, st.k,
                     st.active, st.cap, st.ndone, st.b);
}

__global__ void fill_int(int* p, int n, int v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ void fill_zero(double* p, long n) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.0;
}

// ---------------------------------------------------------------------------
// Biconjugate Lanczos (run_FT without run_simm, DL:6126-6240, n_rep_orth = 0):
// left and right vectors p, q with plain dot products, two L applications per
// step (L q and L^T p).  Per chain, scal holds
//   [0] c_(k-1)  [1] c_(k-2)  [2] s_norm_k  [3] s_norm_(k-1)  [4] b_(k-1)
// ---------------------------------------------------------------------------
struct BiState {
  int nc = 0, kmax = 0;
  long len = 0;
  double *q = nullptr, *qprev = nullptr, *p = nullptr, *pprev = nullptr;  // [nc, len]
  double *Lq = nullptr, *pL = nullptr, *r = nullptr, *s = nullptr;        // [nc, len]
  double* ones = nullptr;   // [len]
  double* part = nullptr;   // [nc, 4, kRedBlocks]: <p, Lq>, <s, s>, <r, r>, <s, r>
  double *a = nullptr, *b = nullptr, *c = nullptr;   // [nc, kmax]
  double* scal = nullptr;   // [nc, 8]
  int *k = nullptr, *active = nullptr, *cap = nullptr, *ndone = nullptr;
};

static void free_bi(BiState& s) {
  void* p[] = {s.q, s.qprev, s.p, s.pprev, s.Lq, s.pL, s.r, s.s, s.ones, s.part, s.a, s.b, s.c,
               s.scal, s.k, s.active, s.cap, s.ndone};
  for (void* x : p)
    if (x) (void)hipFree(x);
}

__global__ void fill_value(double* p, long n, double v) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// q = p = psi0 / |psi0| (plain norm); s_norm_0 = 1
__global__ void bi_init(const double* __restrict__ part, long ps, const double* __restrict__ psi0,
                        long len, double* __restrict__ q, double* __restrict__ p,
                        double* __restrict__ scal) {
  const int c = blockIdx.y;
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (blockIdx.x == 0 && threadIdx.x < 8) scal[c * 8 + threadIdx.x] = threadIdx.x == 2 ? 1.0 : 0.0;
  if (i >= len) return;
  const double v = psi0[c * len + i] / sqrt(sum_partials(part + c * ps));
  q[c * len + i] = v;
  p[c * len + i] = v;
}

// a = <p, Lq> p_norm;  r = Lq - a q - c_(k-1) q_prev;
// s = pL - a p - b_(k-1) p_prev (old_p_norm / p_norm)
__global__ void bi_residual(const double* __restrict__ part, long ps,
                            const double* __restrict__ q, const double* __restrict__ qprev,
                            const double* __restrict__ p, const double* __restrict__ pprev,
                            const double* __restrict__ Lq, const double* __restrict__ pL, long len,
                            int kmax, const int* __restrict__ kk, const int* __restrict__ active,
                            const double* __restrict__ scal, double* __restrict__ a_out,
                            double* __restrict__ r, double* __restrict__ s) {
  const int c = blockIdx.y;
  if (!active[c]) return;
  const int k = kk[c];
  const double* sc = scal + c * 8;
  const double c_old = k > 0 ? sc[0] : 1.0;
  const double p_norm = sc[2] / c_old;
  const double ak = sum_partials(part + c * ps) * p_norm;
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (blockIdx.x == 0 && threadIdx.x == 0) a_out[c * kmax + k] = ak;
  if (i >= len) return;
  const long o = (long)c * len + i;
  double rv = Lq[o] - ak * q[o];
  double sv = pL[o] - ak * p[o];
  if (k > 0) {
    rv -= sc[0] * qprev[o];
    const double old_p_norm = k < 2 ? sc[3] : sc[3] / sc[1];
    sv -= sc[4] * pprev[o] * (old_p_norm / p_norm);
  }
  r[o] = rv;
  s[o] = sv;
}

// <s, s>, <r, r>, <s, r> partial sums into slots 1..3
__global__ void bi_dots(const double* __restrict__ r, const double* __restrict__ s, long len,
                        const int* __restrict__ active, double* __restrict__ part, long ps) {
  const int c = blockIdx.y;
  if (!active[c]) return;
  const double* rc = r + (long)c * len;
  const double* scv = s + (long)c * len;
  double ss = 0.0, rr = 0.0, sr = 0.0;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < len;
       i += (long)gridDim.x * blockDim.x) {
    const double x = scv[i], y = rc[i];
    ss += x * x;
    rr += y * y;
    sr += x * y;
  }
  __shared__ double red[3][kRedThreads];
  red[0][threadIdx.x] = ss;
  red[1][threadIdx.x] = rr;
  red[2][threadIdx.x] = sr;
  __syncthreads();
  for (int o = kRedThreads / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o)
      for (int j = 0; j < 3; ++j) red[j][threadIdx.x] += red[j][threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x < 3) part[c * ps + (long)(1 + threadIdx.x) * kRedBlocks + blockIdx.x] = red[threadIdx.x][0];
}

// b = |r|, s_raw = |s|, c = <s / s_raw, r / b> s_raw p_norm; the chain stops
// (a kept, b and c not) when |b| or |c| is below tol, or at its cap
__device__ __forceinline__ void bi_scalars(const double* part, const double* sc, int k,
                                           double* b, double* s_raw, double* cc,
                                           double* p_norm) {
  const double c_old = k > 0 ? sc[0] : 1.0;
  *p_norm = sc[2] / c_old;
  *s_raw = sqrt(sum_partials(part + 1 * kRedBlocks));
  *b = sqrt(sum_partials(part + 2 * kRedBlocks));
  const double sr = sum_partials(part + 3 * kRedBlocks);
  *cc = (sr / (*s_raw * *b)) * (*s_raw * *p_norm);
}

__global__ void bi_advance(const double* __restrict__ part, long ps, long len, double tol,
                           const int* __restrict__ kk, const int* __restrict__ active,
                           const int* __restrict__ cap, const double* __restrict__ scal,
                           const double* __restrict__ r, const double* __restrict__ s,
                           double* __restrict__ q, double* __restrict__ qprev,
                           double* __restrict__ p, double* __restrict__ pprev) {
  const int c = blockIdx.y;
  if (!active[c]) return;
  const int k = kk[c];
  double b, s_raw, cc, p_norm;
  bi_scalars(part + c * ps, scal + c * 8, k, &b, &s_raw, &cc, &p_norm);
  const bool stop = !(fabs(b) >= tol) || !(fabs(cc) >= tol) || (k + 1 >= cap[c]);
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (stop || i >= len) return;
  const long o = (long)c * len + i;
  qprev[o] = q[o];
  q[o] = r[o] / b;
  pprev[o] = p[o];
  p[o] = s[o] / s_raw;
}

__global__ void bi_commit(const double* __restrict__ part, long ps, int kmax, double tol,
                          int* __restrict__ kk, int* __restrict__ active,
                          const int* __restrict__ cap, int* __restrict__ ndone,
                          double* __restrict__ scal, double* __restrict__ b_out,
                          double* __restrict__ c_out) {
  const int c = threadIdx.x;
  if (!active[c]) return;
  const int k = kk[c];
  double* sc = scal + c * 8;
  double b, s_raw, cc, p_norm;
  bi_scalars(part + c * ps, sc, k, &b, &s_raw, &cc, &p_norm);
  const bool small = !(fabs(b) >= tol) || !(fabs(cc) >= tol);
  ndone[c] = k + 1;
  if (!small) {
    b_out[c * kmax + k] = b;
    c_out[c * kmax + k] = cc;
    sc[1] = sc[0];
    sc[0] = cc;
    sc[3] = sc[2];
    sc[2] = s_raw * p_norm;
    sc[4] = b;
  }
  if (small || k + 1 >= cap[c])
    active[c] = 0;
  else
    kk[c] = k + 1;
}

static void bi_step(Ctx* c, BiState& st, int use_wigner, int precision, int ignore_harmonic,
                    int ignore_v3, int ignore_v4, int reverse, double shift, double tol,
                    hipStream_t s) {
  const long len = st.len;
  const int nc = st.nc;
  const long ps = 4L * kRedBlocks;
  const dim3 gv((unsigned)((len + 255) / 256), nc);
  const dim3 gr(kRedBlocks, nc);
  apply_L_groups(c, nc, len, st.q, st.Lq, use_wigner, 0, ignore_harmonic, ignore_v3, ignore_v4,
                 reverse, shift, precision, s);
  apply_L_groups(c, nc, len, st.p, st.pL, use_wigner, 1, ignore_harmonic, ignore_v3, ignore_v4,
                 reverse, shift, precision, s);
  hipLaunchKernelGGL(dot_partial, gr, dim3(kRedThreads), 0, s, st.p, len, st.Lq, len, st.ones,
                     len, st.part, ps, 0);
  hipLaunchKernelGGL(bi_residual, gv, dim3(256), 0, s, st.part, ps, st.q, st.qprev, st.p,
                     st.pprev, st.Lq, st.pL, len, st.kmax, st.k, st.active, st.scal, st.a, st.r,
                     st.s);
  hipLaunchKernelGGL(bi_dots, gr, dim3(kRedThreads), 0, s, st.r, st.s, len, st.active, st.part,
                     ps);
  hipLaunchKernelGGL(bi_advance, gv, dim3(256), 0, s, st.part, ps, len, tol, st.k, st.active,
                     st.cap, st.scal, st.r, st.s, st.q, st.qprev, st.p, st.pprev);
  hipLaunchKernelGGL(bi_commit, dim3(1), dim3(nc), 0, s, st.part, ps, st.kmax, tol, st.k,
                     st.active, st.cap, st.ndone, st.scal, st.b, st.c);
}

}  // namespace tdscha

using namespace tdscha;

extern "C" {

int tdscha_lanczos_run_biconjugate(tdscha_ctx* ctx, int n_chains, const double* psi0,
                                   int n_steps, const int* caps, double tol, int use_wigner,
                                   int precision, int ignore_harmonic, int ignore_v3,
                                   int ignore_v4, int reverse, double shift, int use_graph,
                                   double* a_out, double* b_out, double* c_out, int* n_done,
                                   double* elapsed_ms) {
  TDS_REQUIRE(ctx != nullptr, 2, "tdscha_lanczos_run_biconjugate: NULL context");
  TDS_REQUIRE(n_chains >= 1 && n_chains <= kMaxRunChains, 2, "n_chains must be in [1, 64]");
  TDS_REQUIRE(use_graph >= 0 && use_graph <= 2, 2, "use_graph must be 0, 1 or 2");
  TDS_REQUIRE(n_steps >= 1 && n_steps <= 100000, 2, "n_steps must be in [1, 100000]");
  TDS_REQUIRE(psi0 && a_out && b_out && c_out && n_done, 2,
              "tdscha_lanczos_run_biconjugate: NULL pointer");
  TDS_REQUIRE(precision >= 0 && precision <= 2, 2, "precision must be 0, 1 or 2");
  Ctx* c = reinterpret_cast<Ctx*>(ctx);
  int rc = setup_pairs(c);
  if (rc) return rc;
  if (precision != PREC_FP64) {
    rc = ensure_split(c, precision);  // allocations must precede graph capture
    if (rc) return rc;
  }
  const int n = c->n;
  BiState st;
  st.nc = n_chains;
  st.kmax = n_steps;
  st.len = n + (long)n * (n + 1);
  const long len = st.len;
  const size_t vb = sizeof(double) * n_chains * len;
  bool ok = true;
  double** vecs[] = {&st.q, &st.qprev, &st.p, &st.pprev, &st.Lq, &st.pL, &st.r, &st.s};
  for (double** v : vecs) ok = ok && hipMalloc((void**)v, vb) == hipSuccess;
  ok = ok && hipMalloc((void**)&st.ones, sizeof(double) * len) == hipSuccess &&
       hipMalloc((void**)&st.part, sizeof(double) * n_chains * 4 * kRedBlocks) == hipSuccess &&
       hipMalloc((void**)&st.a, sizeof(double) * n_chains * n_steps) == hipSuccess &&
       hipMalloc((void**)&st.b, sizeof(double) * n_chains * n_steps) == hipSuccess &&
       hipMalloc((void**)&st.c, sizeof(double) * n_chains * n_steps) == hipSuccess &&
       hipMalloc((void**)&st.scal, sizeof(double) * n_chains * 8) == hipSuccess &&
       hipMalloc((void**)&st.k, sizeof(int) * n_chains) == hipSuccess &&
       hipMalloc((void**)&st.active, sizeof(int) * n_chains) == hipSuccess &&
       hipMalloc((void**)&st.cap, sizeof(int) * n_chains) == hipSuccess &&
       hipMalloc((void**)&st.ndone, sizeof(int) * n_chains) == hipSuccess;
  if (!ok) {
    free_bi(st);
    set_error("tdscha_lanczos_run_biconjugate: device allocation failed");
    return -1;
  }
  hipStream_t s = c->stream;
  std::vector<int> hcap(n_chains, n_steps);
  if (caps)
    for (int i = 0; i < n_chains; ++i) hcap[i] = std::max(1, std::min(caps[i], n_steps));
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  auto fail = [&](const char* what) {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    free_bi(st);
    set_error(std::string("tdscha_lanczos_run_biconjugate: ") + what);
    return -1;
  };
  if (hipMemcpyAsync(st.cap, hcap.data(), sizeof(int) * n_chains, hipMemcpyHostToDevice, s) != hipSuccess ||
      hipMemcpyAsync(st.Lq, psi0, vb, hipMemcpyHostToDevice, s) != hipSuccess)
    return fail("upload failed");
  const long na = (long)n_chains * n_steps;
  hipLaunchKernelGGL(fill_int, dim3(1), dim3(64), 0, s, st.k, n_chains, 0);
  hipLaunchKernelGGL(fill_int, dim3(1), dim3(64), 0, s, st.active, n_chains, 1);
  hipLaunchKernelGGL(fill_int, dim3(1), dim3(64), 0, s, st.ndone, n_chains, 0);
  hipLaunchKernelGGL(fill_value, dim3((unsigned)((len + 255) / 256)), dim3(256), 0, s, st.ones,
                     len, 1.0);
  hipLaunchKernelGGL(fill_zero, dim3((unsigned)((na + 255) / 256)), dim3(256), 0, s, st.a, na);
  hipLaunchKernelGGL(fill_zero, dim3((unsigned)((na + 255) / 256)), dim3(256), 0, s, st.b, na);
  hipLaunchKernelGGL(fill_zero, dim3((unsigned)((na + 255) / 256)), dim3(256), 0, s, st.c, na);
  const long ps = 4L * kRedBlocks;
  // |psi0|^2 into slot 0, then q = p = psi0 / |psi0| (psi0 staged in Lq)
  hipLaunchKernelGGL(dot_partial, dim3(kRedBlocks, n_chains), dim3(kRedThreads), 0, s, st.Lq, len,
                     st.Lq, len, st.ones, len, st.part, ps, 0);
  hipLaunchKernelGGL(bi_init, dim3((unsigned)((len + 255) / 256), n_chains), dim3(256), 0, s,
                     st.part, ps, st.Lq, len, st.q, st.p, st.scal);
  if (hipGetLastError() != hipSuccess || hipStreamSynchronize(s) != hipSuccess)
    return fail("initialisation failed");
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
    return fail("event creation failed");
  const int captured = use_graph == 2 ? n_steps : 1;
  if (use_graph) {
    if (hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) != hipSuccess)
      return fail("stream capture failed to start");
    for (int step = 0; step < captured; ++step)
      bi_step(c, st, use_wigner, precision, ignore_harmonic, ignore_v3, ignore_v4, reverse, shift,
              tol, s);
    if (hipStreamEndCapture(s, &graph) != hipSuccess) return fail("stream capture failed");
    if (hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) != hipSuccess)
      return fail("graph instantiation failed");
  }
  (void)hipEventRecord(e0, s);
  for (int step = 0; step < n_steps; step += use_graph ? captured : 1) {
    if (use_graph) {
      if (hipGraphLaunch(exec, s) != hipSuccess) return fail("graph launch failed");
    } else {
      bi_step(c, st, use_wigner, precision, ignore_harmonic, ignore_v3, ignore_v4, reverse,
              shift, tol, s);
    }
  }
  (void)hipEventRecord(e1, s);
  if (hipStreamSynchronize(s) != hipSuccess || hipGetLastError() != hipSuccess)
    return fail("step execution failed");
  float ms = 0.f;
  (void)hipEventElapsedTime(&ms, e0, e1);
  if (elapsed_ms) *elapsed_ms = ms;
  ok = hipMemcpy(a_out, st.a, sizeof(double) * na, hipMemcpyDeviceToHost) == hipSuccess &&
       hipMemcpy(b_out, st.b, sizeof(double) * na, hipMemcpyDeviceToHost) == hipSuccess &&
       hipMemcpy(c_out, st.c, sizeof(double) * na, hipMemcpyDeviceToHost) == hipSuccess &&
       hipMemcpy(n_done, st.ndone, sizeof(int) * n_chains, hipMemcpyDeviceToHost) == hipSuccess;
  if (!ok) return fail("download failed");
  if (exec) (void)hipGraphExecDestroy(exec);
  if (graph) (void)hipGraphDestroy(graph);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  free_bi(st);
  return 0;
}

int tdscha_lanczos_run(tdscha_ctx* ctx, int n_chains, const double* psi0, int n_steps,
                       const int* caps, double tol, int reorth, int precision,
                       int ignore_harmonic, int ignore_v3, int ignore_v4, int use_graph,
                       double* a_out, double* b_out, int* n_done, double* basis_out,
                       double* elapsed_ms) {
  TDS_REQUIRE(ctx != nullptr, 2, "tdscha_lanczos_run: NULL context");
  TDS_REQUIRE(n_chains >= 1 && n_chains <= kMaxRunChains, 2, "n_chains must be in [1, 64]");
  TDS_REQUIRE(use_graph >= 0 && use_graph <= 2, 2, "use_graph must be 0, 1 or 2");
  TDS_REQUIRE(n_steps >= 1 && n_steps <= 100000, 2, "n_steps must be in [1, 100000]");
  TDS_REQUIRE(psi0 && a_out && b_out && n_done, 2, "tdscha_lanczos_run: NULL pointer");
  TDS_REQUIRE(precision >= 0 && precision <= 2, 2, "precision must be 0, 1 or 2");
  TDS_REQUIRE(reorth == 0 || reorth == 1, 2, "reorth must be 0 or 1");
  Ctx* c = reinterpret_cast<Ctx*>(ctx);
  int rc = setup_pairs(c);
  if (rc) return rc;
  if (precision != PREC_FP64) {
    rc = ensure_split(c, precision);  // allocations must precede graph capture
    if (rc) return rc;
  }
  const int n = c->n;
  LanczosState st;
  st.nc = n_chains;
  st.kmax = n_steps;
  st.len = n + (long)n * (n + 1);
  const long len = st.len;
  const size_t qbytes = sizeof(double) * (size_t)n_chains * (n_steps + 1) * len;
  bool ok = hipMalloc((void**)&st.Q, qbytes) == hipSuccess &&
            hipMalloc((void**)&st.qcur, sizeof(double) * n_chains * len) == hipSuccess &&
            hipMalloc((void**)&st.Lq, sizeof(double) * n_chains * len) == hipSuccess &&
            hipMalloc((void**)&st.r, sizeof(double) * n_chains * len) == hipSuccess &&
            hipMalloc((void**)&st.mask, sizeof(double) * len) == hipSuccess &&
            hipMalloc((void**)&st.part, sizeof(double) * n_chains * (n_steps + 1) * kRedBlocks) == hipSuccess &&
            hipMalloc((void**)&st.coef, sizeof(double) * n_chains * (n_steps + 1)) == hipSuccess &&
            hipMalloc((void**)&st.a, sizeof(double) * n_chains * n_steps) == hipSuccess &&
            hipMalloc((void**)&st.b, sizeof(double) * n_chains * n_steps) == hipSuccess &&
            hipMalloc((void**)&st.scal, sizeof(double) * n_chains * 4) == hipSuccess &&
            hipMalloc((void**)&st.k, sizeof(int) * n_chains) == hipSuccess &&
            hipMalloc((void**)&st.active, sizeof(int) * n_chains) == hipSuccess &&
            hipMalloc((void**)&st.act2, sizeof(int) * n_chains) == hipSuccess &&
            hipMalloc((void**)&st.cap, sizeof(int) * n_chains) == hipSuccess &&
            hipMalloc((void**)&st.ndone, sizeof(int) * n_chains) == hipSuccess;
  // projections: element slices of >= 4096 entries (at most 64); updates of
  // short vectors split over ranges of basis vectors so that enough
  // work-groups run
  st.nslice = (int)std::max(1L, std::min(64L, len / 4096));
  st.jsplit = len < 32768 ? std::max(1, std::min(16, (n_steps + 1) / 32)) : 1;
  st.jchunk = (n_steps + 1 + st.jsplit - 1) / st.jsplit;
  if (ok && st.jsplit > 1)
    ok = hipMalloc((void**)&st.upart, sizeof(double) * n_chains * st.jsplit * len) == hipSuccess;
  if (!ok) {
    free_state(st);
    set_error("tdscha_lanczos_run: device allocation failed");
    return -1;
  }
  hipStream_t s = c->stream;
  std::vector<int> hcap(n_chains, n_steps);
  if (caps)
    for (int i = 0; i < n_chains; ++i) hcap[i] = std::max(1, std::min(caps[i], n_steps));
  std::vector<double> ha(n_chains * (size_t)n_steps, 0.0), hb(n_chains * (size_t)n_steps, 0.0);
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  auto fail = [&](const char* what) {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    free_state(st);
    set_error(std::string("tdscha_lanczos_run: ") + what);
    return -1;
  };
  // psi0 staged in the residual buffer (n_chains vectors)
  if (hipMemcpyAsync(st.cap, hcap.data(), sizeof(int) * n_chains, hipMemcpyHostToDevice, s) != hipSuccess ||
      hipMemcpyAsync(st.r, psi0, sizeof(double) * n_chains * len, hipMemcpyHostToDevice, s) != hipSuccess)
    return fail("upload failed");
  hipLaunchKernelGGL(fill_int, dim3(1), dim3(64), 0, s, st.k, n_chains, 0);
  hipLaunchKernelGGL(fill_int, dim3(1), dim3(64), 0, s, st.active, n_chains, 1);
  hipLaunchKernelGGL(fill_int, dim3(1), dim3(64), 0, s, st.ndone, n_chains, 0);
  // a chain that stops on |b| < tol leaves its last b slot unwritten: the
  // coefficient arrays start at zero so that slot reads as a stop
  {
    const long na = (long)n_chains * n_steps;
    hipLaunchKernelGGL(fill_zero, dim3((unsigned)((na + 255) / 256)), dim3(256), 0, s, st.a, na);
    hipLaunchKernelGGL(fill_zero, dim3((unsigned)((na + 255) / 256)), dim3(256), 0, s, st.b, na);
  }
  hipLaunchKernelGGL(build_mask, dim3((unsigned)((len + 255) / 256)), dim3(256), 0, s, n, len,
                     st.mask);
  const long ps = (long)(n_steps + 1) * kRedBlocks;
  hipLaunchKernelGGL(dot_partial, dim3(kRedBlocks, n_chains), dim3(kRedThreads), 0, s, st.r,
                     len, st.r, len, st.mask, len, st.part, ps, 0);
  hipLaunchKernelGGL(lanczos_init, dim3((unsigned)((len + 255) / 256), n_chains), dim3(256), 0, s,
                     st.part, ps, st.r, len, n_steps, st.Q, st.qcur);
  if (hipGetLastError() != hipSuccess || hipStreamSynchronize(s) != hipSuccess)
    return fail("initialisation failed");
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
    return fail("event creation failed");
  // use_graph 1: one step captured and replayed; 2: the whole run captured
  // and launched once (stopped chains fall through every later step on
  // their device flags)
  const int captured = use_graph == 2 ? n_steps : 1;
  if (use_graph) {
    if (hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) != hipSuccess)
      return fail("stream capture failed to start");
    for (int step = 0; step < captured; ++step)
      lanczos_step(c, st, reorth, precision, ignore_harmonic, ignore_v3, ignore_v4, tol, s);
    if (hipStreamEndCapture(s, &graph) != hipSuccess) return fail("stream capture failed");
    if (hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) != hipSuccess)
      return fail("graph instantiation failed");
  }
  (void)hipEventRecord(e0, s);
  for (int step = 0; step < n_steps; step += use_graph ? captured : 1) {
    if (use_graph) {
      if (hipGraphLaunch(exec, s) != hipSuccess) return fail("graph launch failed");
    } else {
      lanczos_step(c, st, reorth, precision, ignore_harmonic, ignore_v3, ignore_v4, tol, s);
    }
  }
  (void)hipEventRecord(e1, s);
  if (hipStreamSynchronize(s) != hipSuccess || hipGetLastError() != hipSuccess)
    return fail("step execution failed");
  float ms = 0.f;
  (void)hipEventElapsedTime(&ms, e0, e1);
  if (elapsed_ms) *elapsed_ms = ms;
  ok = hipMemcpy(a_out, st.a, sizeof(double) * n_chains * n_steps, hipMemcpyDeviceToHost) == hipSuccess &&
       hipMemcpy(b_out, st.b, sizeof(double) * n_chains * n_steps, hipMemcpyDeviceToHost) == hipSuccess &&
       hipMemcpy(n_done, st.ndone, sizeof(int) * n_chains, hipMemcpyDeviceToHost) == hipSuccess;
  if (ok && basis_out)
    ok = hipMemcpy(basis_out, st.Q, qbytes, hipMemcpyDeviceToHost) == hipSuccess;
  if (!ok) return fail("download failed");
  if (exec) (void)hipGraphExecDestroy(exec);
  if (graph) (void)hipGraphDestroy(graph);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  free_state(st);
  return 0;
}

}  // extern "C"
