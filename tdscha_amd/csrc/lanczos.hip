// Device-resident symmetric Lanczos for a batch of chains.
//
// The Wigner-representation recurrence of run_FT with run_simm
// (Modules/DynamicalLanczos.py:6032-6240, p = q, metric = mask_dot_wigner,
// DL:5905-5954), batched over up to 8 chains that share every L application,
// optionally with full classical Gram-Schmidt reorthogonalisation repeated
// twice (the real-space analogue of the q-space full reorthogonalisation,
// Modules/QSpaceLanczos.py:1113-1130).  Chains stop independently at
// |b| < tol or at their own step cap; every scalar lives on the device so
// one step is a fixed kernel sequence, captured once as a hipGraph and
// replayed.  Reductions are two-stage with a fixed order (deterministic).
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
  apply_L_dev(c, nc, st.qcur, st.Lq, 1, 0, ignore_harmonic, ignore_v3, ignore_v4, 0, 0.0,
              precision, s);
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

}  // namespace tdscha

using namespace tdscha;

extern "C" {

int tdscha_lanczos_run(tdscha_ctx* ctx, int n_chains, const double* psi0, int n_steps,
                       const int* caps, double tol, int reorth, int precision,
                       int ignore_harmonic, int ignore_v3, int ignore_v4, int use_graph,
                       double* a_out, double* b_out, int* n_done, double* basis_out,
                       double* elapsed_ms) {
  TDS_REQUIRE(ctx != nullptr, 2, "tdscha_lanczos_run: NULL context");
  TDS_REQUIRE(n_chains >= 1 && n_chains <= kMaxChains, 2, "n_chains must be in [1, 8]");
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
  if (hipMemcpyAsync(st.cap, hcap.data(), sizeof(int) * n_chains, hipMemcpyHostToDevice, s) != hipSuccess ||
      hipMemcpyAsync(c->psi_dev, psi0, sizeof(double) * n_chains * len, hipMemcpyHostToDevice, s) != hipSuccess)
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
  hipLaunchKernelGGL(dot_partial, dim3(kRedBlocks, n_chains), dim3(kRedThreads), 0, s, c->psi_dev,
                     len, c->psi_dev, len, st.mask, len, st.part, ps, 0);
  hipLaunchKernelGGL(lanczos_init, dim3((unsigned)((len + 255) / 256), n_chains), dim3(256), 0, s,
                     st.part, ps, c->psi_dev, len, n_steps, st.Q, st.qcur);
  if (hipGetLastError() != hipSuccess || hipStreamSynchronize(s) != hipSuccess)
    return fail("initialisation failed");
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
    return fail("event creation failed");
  if (use_graph) {
    if (hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) != hipSuccess)
      return fail("stream capture failed to start");
    lanczos_step(c, st, reorth, precision, ignore_harmonic, ignore_v3, ignore_v4, tol, s);
    if (hipStreamEndCapture(s, &graph) != hipSuccess) return fail("stream capture failed");
    if (hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) != hipSuccess)
      return fail("graph instantiation failed");
  }
  (void)hipEventRecord(e0, s);
  for (int step = 0; step < n_steps; ++step) {
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
