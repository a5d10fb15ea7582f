// Operator-rotation form of the symmetrised perturbed averages.
//
// The reference applies every symmetry to every configuration inside its
// loop over configurations (CModules/LanczosFunctions.c:2158-2170).  The
// image form of this library materialises those images once (U = S_j x_i,
// F = S_j y_i, n_syms times the ensemble).  This form keeps only the
// configurations and rotates the small operators instead.  The S_j are block
// diagonal over degenerate modes, on which fY and fpsi are constant, so they
// commute with both, and for the image b = (i, j):
//   s1 = x_i^T (S_j^T Y1 S_j) x_i          s2 = x_i^T (S_j^T Y1 S_j) (fpsi y_i)
//   s3 = (fY x_i) . (S_j^T R1)             s4 = y_i . (S_j^T R1)
//   f   = sum_j S_j f~_j                   d2v = sum_j S_j d2v~_j S_j^T
// where f~_j and d2v~_j are the averages over the configurations alone for the
// rotated perturbation (S_j^T R1, S_j^T Y1 S_j).  The flops of the GEMMs are
// those of the image form (n_syms passes over N configurations instead of one
// over N n_syms images); the ensemble takes n_syms times less memory.  The
// (chain, symmetry) pairs run through the ordinary kernels 8 at a time and
// are summed in a fixed order (chain-major, symmetry by symmetry).
//
// The rotations are two one-sided block products each (S_j^T (Y1 S_j), and
// S_j (d2v~ S_j^T)): n^2 sum_b d_b^2 / n multiply-adds per side, against
// (sum_b d_b^2)^2 for the two-sided sum per element (blocks of up to 24 modes
// at 128 atoms).  Per mode m the tables hold its block's modes cols(m) and
// the row and column of S_j restricted to them.
#include "common.hpp"

namespace tdscha {

int launch_pa_fp64(Ctx* c, int nc, int apply_d4, hipStream_t s, hipEvent_t* ev);
int launch_pa_split(Ctx* c, int nc, int apply_d4, int prec, hipStream_t s, hipEvent_t* ev);

// position-major tables, so that neighbouring modes (lanes) read neighbouring
// entries: entry t of mode m at [t * n + m]
struct RotTab {
  const int* cols;      // [dmax, n]
  const int* d;         // [n]
  const double* sr;     // [nsym, dmax, n] S_j[m, cols(m)[t]]
  const double* sc;     // [nsym, dmax, n] S_j[cols(m)[t], m]
  int dmax;
  int n;
  __device__ __forceinline__ int col(int m, int t) const { return cols[(long)t * n + m]; }
  __device__ __forceinline__ double row_s(int j, int m, int t) const {
    return sr[((long)j * dmax + t) * n + m];
  }
  __device__ __forceinline__ double col_s(int j, int m, int t) const {
    return sc[((long)j * dmax + t) * n + m];
  }
};

// the configurations, padded to [Mp, np] (zero rows and columns beyond)
__global__ void rot_copy_configs(const double* __restrict__ X, const double* __restrict__ Y,
                                 const double* __restrict__ rho, int n, int np, long N, long Mp,
                                 double* __restrict__ U, double* __restrict__ F,
                                 double* __restrict__ rho_img) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= Mp * np) return;
  const long r = e / np;
  const int m = (int)(e % np);
  const bool in = r < N && m < n;
  U[e] = in ? X[r * n + m] : 0.0;
  F[e] = in ? Y[r * n + m] : 0.0;
  if (m == 0) rho_img[r] = r < N ? rho[r] : 0.0;
}

__global__ void rot_copy(const double* __restrict__ src, long count, double* __restrict__ dst) {
  for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < count;
       e += (long)gridDim.x * blockDim.x)
    dst[e] = src[e];
}

__global__ void rot_zero(long count, double* __restrict__ dst) {
  for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < count;
       e += (long)gridDim.x * blockDim.x)
    dst[e] = 0.0;
}

// pair k = blockIdx.y of the pass: p = p0 + k, chain p / nsym, symmetry
// p % nsym.  Grids: (ceil(np^2 / 256), g), one thread per padded element.

// T1[k] = Y1[chain] S_j:  T1[r, q] = sum_t Y1[r, cols(q)[t]] S_j[cols(q)[t], q]
__global__ void rot_in_right(RotTab tb, const double* __restrict__ Y1b, int np, int nsym, int p0,
                             double* __restrict__ T1) {
  const int k = blockIdx.y, p = p0 + k, ch = p / nsym, j = p % nsym;
  const long nn = (long)np * np;
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nn) return;
  const int r = (int)(e / np), q = (int)(e % np);
  double v = 0.0;
  if (r < tb.n && q < tb.n) {
    const double* y = Y1b + (long)ch * nn + (long)r * np;
    const int d = tb.d[q];
    for (int t = 0; t < d; ++t) v += y[tb.col(q, t)] * tb.col_s(j, q, t);
  }
  T1[(long)k * nn + e] = v;
}

// Y1o[k] = S_j^T T1[k]:  Y1o[r, q] = sum_t S_j[cols(r)[t], r] T1[cols(r)[t], q];
// R1o[k] = S_j^T R1[chain]
__global__ void rot_in_left(RotTab tb, const double* __restrict__ T1,
                            const double* __restrict__ R1b, int np, int nsym, int p0,
                            double* __restrict__ R1o, double* __restrict__ Y1o) {
  const int k = blockIdx.y, p = p0 + k, ch = p / nsym, j = p % nsym;
  const long nn = (long)np * np;
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nn) return;
  const int r = (int)(e / np), q = (int)(e % np);
  double v = 0.0, w = 0.0;
  if (r < tb.n) {
    const int d = tb.d[r];
    if (q < tb.n) {
      const double* t1 = T1 + (long)k * nn + q;
      for (int t = 0; t < d; ++t) v += tb.col_s(j, r, t) * t1[(long)tb.col(r, t) * np];
    }
    if (q == 0)
      for (int t = 0; t < d; ++t) w += tb.col_s(j, r, t) * R1b[(long)ch * np + tb.col(r, t)];
  }
  Y1o[(long)k * nn + e] = v;
  if (q == 0) R1o[(long)k * np + r] = w;
}

// T1[k] = fY Z + (fY Z)^T of the pass's raw outer sums Z (32 x 32 tiles
// through LDS, so that both reads are coalesced)
__global__ void rot_sym(const double* __restrict__ zsum, const double* __restrict__ fy, int np,
                        double* __restrict__ T1) {
  __shared__ double tile[32][33];
  const int k = blockIdx.z;
  const long nn = (long)np * np;
  const double* z = zsum + (long)k * nn;
  const int bx = blockIdx.x * 32, by = blockIdx.y * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;   // 32 x 8
  // tile of Z rows bx.., columns by.. (the transposed block)
  for (int yy = ty; yy < 32; yy += 8) tile[yy][tx] = fy[bx + yy] * z[(long)(bx + yy) * np + by + tx];
  __syncthreads();
  for (int yy = ty; yy < 32; yy += 8) {
    const int a = by + yy, b = bx + tx;
    T1[(long)k * nn + (long)a * np + b] = fy[a] * z[(long)a * np + b] + tile[tx][yy];
  }
}

// T2[k] = D~ S_j^T:  T2[a, q] = sum_t D~[a, cols(q)[t]] S_j[q, cols(q)[t]]
__global__ void rot_out_right(RotTab tb, const double* __restrict__ T1, int np, int nsym, int p0,
                              double* __restrict__ T2) {
  const int k = blockIdx.y, p = p0 + k, j = p % nsym;
  const long nn = (long)np * np;
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nn) return;
  const int a = (int)(e / np), q = (int)(e % np);
  double v = 0.0;
  if (a < tb.n && q < tb.n) {
    const double* dt = T1 + (long)k * nn + (long)a * np;
    const int d = tb.d[q];
    for (int t = 0; t < d; ++t) v += dt[tb.col(q, t)] * tb.row_s(j, q, t);
  }
  T2[(long)k * nn + e] = v;
}

// T1[k] = S_j T2[k] (padded np), F1[k] = S_j f~[k]
__global__ void rot_out_left(RotTab tb, const double* __restrict__ T2,
                             const double* __restrict__ fsum, int np, int nsym, int p0,
                             double* __restrict__ T1, double* __restrict__ F1) {
  const int k = blockIdx.y, j = (p0 + k) % nsym;
  const long nn = (long)np * np;
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nn) return;
  const int r = (int)(e / np), q = (int)(e % np);
  double v = 0.0, w = 0.0;
  if (r < tb.n) {
    const int d = tb.d[r];
    if (q < tb.n) {
      const double* t2 = T2 + (long)k * nn + q;
      for (int t = 0; t < d; ++t) v += tb.row_s(j, r, t) * t2[(long)tb.col(r, t) * np];
    }
    if (q == 0)
      for (int t = 0; t < d; ++t) w += tb.row_s(j, r, t) * fsum[(long)k * np + tb.col(r, t)];
  }
  T1[(long)k * nn + e] = v;
  if (q == 0) F1[(long)k * np + r] = w;
}

// d2v[chain] += normd T1[k], f[chain] += normf F1[k] for the pairs of the pass
// in order (one thread per element of the unpadded outputs)
__global__ void rot_accumulate(const double* __restrict__ T1, const double* __restrict__ F1,
                               int n, int np, int nsym, int p0, int g, double normf, double normd,
                               double* __restrict__ f_all, double* __restrict__ d2v_all) {
  const long nn = (long)np * np;
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long)n * n) return;
  const int r = (int)(e / n), q = (int)(e % n);
  // the pairs of one chain are consecutive: the running sums stay in
  // registers and go back when the chain changes (same order of additions)
  double t[kMaxChains], fv[kMaxChains];
#pragma unroll
  for (int k = 0; k < kMaxChains; ++k) {
    const int kk = k < g ? k : g - 1;
    t[k] = T1[(long)kk * nn + (long)r * np + q];
    fv[k] = e < n ? F1[(long)kk * np + e] : 0.0;
  }
  int ch = p0 / nsym;
  double dv = d2v_all[(long)ch * n * n + e], fa = e < n ? f_all[(long)ch * n + e] : 0.0;
#pragma unroll
  for (int k = 0; k < kMaxChains; ++k) {
    if (k < g) {
      const int c = (p0 + k) / nsym;
      if (c != ch) {
        d2v_all[(long)ch * n * n + e] = dv;
        if (e < n) f_all[(long)ch * n + e] = fa;
        ch = c;
        dv = d2v_all[(long)ch * n * n + e];
        fa = e < n ? f_all[(long)ch * n + e] : 0.0;
      }
      dv += normd * t[k];
      fa += normf * fv[k];
    }
  }
  d2v_all[(long)ch * n * n + e] = dv;
  if (e < n) f_all[(long)ch * n + e] = fa;
}

// ---- panel form, used when every block is a range of consecutive modes
// listed in ascending order and has at most kRotPanelDmax modes: bs[m] is the
// first mode of m's block.  The modes are cut into panels of whole blocks, at
// most kRotPanel modes each (a last panel holds the padding n .. np - 1).
// Because S_j is block diagonal, the (row panel, column panel) tile of
// S_j^T A S_j (or S_j A S_j^T) is L M R with M that tile of A and L, R the
// panels' diagonal parts of S_j (transposed as the side requires): two dense
// 32 x 32 x 32 FP64 MFMA products per work-group, everything through LDS.
// Grids: (npanel^2, g), 256 threads (4 waves, one 16 x 16 quadrant each).
constexpr int kRotPanel = 32;
constexpr int kRotPanelDmax = 24;
constexpr int kRotLdP = kRotPanel + 1;

struct Panels {
  const int* start;   // [npanel + 1] first mode of each panel, then np
  int npanel;         // panels, the padding panel (when np > n) last
  int nreal;          // panels holding modes below n
};

// table entry t of mode m: S_j[m, bs(m) + t] (ROW) or S_j[bs(m) + t, m]
template <bool ROW>
__device__ __forceinline__ double rot_entry(const RotTab& tb, int j, int m, int t) {
  return ROW ? tb.row_s(j, m, t) : tb.col_s(j, m, t);
}

// The sides of a tile: L[rr][a] = entry (a - bs(r) + r0) of row r = r0 + rr
// and R[b][qq] = entry (b - bs(q) + q0) of column q = q0 + qq inside the
// mode's block, zero elsewhere and past the tile (h rows, w columns): both
// are zeroed, then (after a barrier) the block entries are scattered (the
// tables hold zeros past a mode's block, so the loads take table rows and
// modes clamped into range and need no test).
__device__ __forceinline__ void panel_sides_zero(double* __restrict__ L, double* __restrict__ R) {
  for (int e = threadIdx.x; e < kRotPanel * kRotLdP; e += 256) {
    L[e] = 0.0;
    R[e] = 0.0;
  }
}
// C = A B for 32 x 32 LDS matrices (row stride kRotLdP); wave w computes the
// quadrant rows 16 (w / 2) .., columns 16 (w % 2) ..; lane holds
// C[16 (w / 2) + 4 r + lane / 16][16 (w % 2) + lane % 16] in acc[r]
__device__ __forceinline__ v4d panel_mfma(const double* __restrict__ A,
                                          const double* __restrict__ B) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int i = 16 * (w >> 1) + (lane & 15), jc = 16 * (w & 1) + (lane & 15), kk = lane >> 4;
  v4d acc = v4d{0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int ks = 0; ks < kRotPanel / 4; ++ks)
    acc = __builtin_amdgcn_mfma_f64_16x16x4f64(A[i * kRotLdP + 4 * ks + kk],
                                               B[(4 * ks + kk) * kRotLdP + jc], acc, 0, 0, 0);
  return acc;
}

// zeros on a tile that holds padding rows or columns (and the padding of the
// vector output on the first column panel)
__device__ __forceinline__ void panel_zero(int r0, int h, int q0, int w, int np, bool first_col,
                                           double* __restrict__ out, double* __restrict__ vec) {
  for (int e = threadIdx.x; e < h * w; e += 256) out[(long)(r0 + e / w) * np + q0 + e % w] = 0.0;
  if (first_col)
    for (int e = threadIdx.x; e < h; e += 256) vec[r0 + e] = 0.0;
}

// One tile of a rotation: out tile = L M R and, on the first column panel,
// vec[r0 + rr] = sum_a L[rr][a] x[r0 + a].  In (OUT false): M = A (the
// perturbation), L = S_r^T, R = S_c.  Out: M = D~ = fY A + (fY A)^T of the
// raw sums A, L = S_r, R = S_c^T.  Every global load of a thread is issued
// in one phase; the product L (M R) keeps M R in M's place.
template <bool OUT>
__device__ __forceinline__ void panel_run(const RotTab& tb, const int* __restrict__ bs,
                                          const Panels& pn, int j, const double* __restrict__ A,
                                          const double* __restrict__ fy,
                                          const double* __restrict__ x, int np,
                                          double* __restrict__ out, double* __restrict__ vec,
                                          double* Ls, double* Ms, double* Rs, double* xs) {
  const int pi = blockIdx.x / pn.npanel, pj = blockIdx.x % pn.npanel;
  const int r0 = pn.start[pi], h = pn.start[pi + 1] - r0;
  const int q0 = pn.start[pj], w = pn.start[pj + 1] - q0;
  if (pi >= pn.nreal || pj >= pn.nreal) {
    panel_zero(r0, h, q0, w, np, pj == 0, out, vec);
    return;
  }
  const int tid = threadIdx.x;
  constexpr int kM = kRotPanel * kRotPanel / 256;
  constexpr int kU = kRotPanel * kRotPanelDmax / 256;
  static_assert(kRotPanel * kRotPanelDmax == 256 * kU, "entries per thread");
  // ---- loads: the tile (direct term, and the transposed term read along
  // the rows of A), the side entries, the vector slice
  // (rows and columns clamped into the tile, so every load is unconditional)
  double mv[kM], mt[kM];
#pragma unroll
  for (int u = 0; u < kM; ++u) {
    const int e = tid + 256 * u, a = e / kRotPanel, b = e % kRotPanel;
    const int ar = a < h ? a : h - 1, bc = b < w ? b : w - 1;
    mv[u] = A[(long)(r0 + ar) * np + q0 + bc];
    if (OUT) mv[u] *= fy[r0 + ar];
    if (!(a < h && b < w)) mv[u] = 0.0;
    if (OUT) {
      // transposed term of element (b, a): A[q0 + a][r0 + b]
      const int aw = a < w ? a : w - 1, bh = b < h ? b : h - 1;
      mt[u] = fy[q0 + aw] * A[(long)(q0 + aw) * np + r0 + bh];
    }
  }
  double lv[kU], rv[kU];
  int lb[kU], ld[kU], rb[kU], rd[kU];
#pragma unroll
  for (int u = 0; u < kU; ++u) {
    const int e = tid + 256 * u, xx = e / kRotPanelDmax, t = e % kRotPanelDmax;
    const int tt = t < tb.dmax ? t : tb.dmax - 1;
    const int mr = r0 + (xx < h ? xx : h - 1), mc = q0 + (xx < w ? xx : w - 1);
    lv[u] = rot_entry<OUT>(tb, j, mr, tt);
    rv[u] = rot_entry<OUT>(tb, j, mc, tt);
    lb[u] = bs[mr];
    ld[u] = tb.d[mr];
    rb[u] = bs[mc];
    rd[u] = tb.d[mc];
  }
  const double xv0 = x[r0 + (tid < h ? tid : h - 1)];
  const double xv = tid < h ? xv0 : 0.0;
  panel_sides_zero(Ls, Rs);
  __syncthreads();
  // ---- stores
#pragma unroll
  for (int u = 0; u < kM; ++u) {
    const int e = tid + 256 * u;
    Ms[(e / kRotPanel) * kRotLdP + e % kRotPanel] = mv[u];
  }
#pragma unroll
  for (int u = 0; u < kU; ++u) {
    const int e = tid + 256 * u, xx = e / kRotPanelDmax, t = e % kRotPanelDmax;
    if (xx < h && t < ld[u]) Ls[xx * kRotLdP + lb[u] - r0 + t] = lv[u];
    if (xx < w && t < rd[u]) Rs[(rb[u] - q0 + t) * kRotLdP + xx] = rv[u];
  }
  if (tid < kRotPanel) xs[tid] = xv;
  __syncthreads();
  if (OUT) {
#pragma unroll
    for (int u = 0; u < kM; ++u) {
      const int e = tid + 256 * u, a = e / kRotPanel, b = e % kRotPanel;
      if (b < h && a < w) Ms[b * kRotLdP + a] += mt[u];
    }
    __syncthreads();
  }
  // ---- products
  const int lane = tid & 63, wv = tid >> 6;
  const int i0 = 16 * (wv >> 1) + (lane >> 4), jc = 16 * (wv & 1) + (lane & 15);
  const v4d t = panel_mfma(Ms, Rs);
  __syncthreads();
#pragma unroll
  for (int r = 0; r < 4; ++r) Ms[(i0 + 4 * r) * kRotLdP + jc] = t[r];
  __syncthreads();
  const v4d o = panel_mfma(Ls, Ms);
#pragma unroll
  for (int r = 0; r < 4; ++r)
    if (i0 + 4 * r < h && jc < w) out[(long)(r0 + i0 + 4 * r) * np + q0 + jc] = o[r];
  if (pj == 0 && tid < h) {
    double v = 0.0;
    for (int a = 0; a < h; ++a) v += Ls[tid * kRotLdP + a] * xs[a];
    vec[r0 + tid] = v;
  }
}

// Y1o[k] = S_j^T Y1[chain] S_j on one tile (zero on the padding); the tiles of
// the first column panel also write R1o[k] = S_j^T R1[chain]
__global__ __launch_bounds__(256) void rot_in_panel(RotTab tb, const int* __restrict__ bs,
                                                    Panels pn, const double* __restrict__ Y1b,
                                                    const double* __restrict__ R1b, int np,
                                                    int nsym, int p0, double* __restrict__ R1o,
                                                    double* __restrict__ Y1o) {
  __shared__ double Ls[kRotPanel * kRotLdP], Ms[kRotPanel * kRotLdP], Rs[kRotPanel * kRotLdP];
  __shared__ double xs[kRotPanel];
  const int k = blockIdx.y, p = p0 + k, ch = p / nsym, j = p % nsym;
  const long nn = (long)np * np;
  panel_run<false>(tb, bs, pn, j, Y1b + (long)ch * nn, nullptr, R1b + (long)ch * np, np,
                   Y1o + (long)k * nn, R1o + (long)k * np, Ls, Ms, Rs, xs);
}

// T1[k] = S_j D~ S_j^T on one tile, D~ = fY Z + (fY Z)^T of the pass's raw
// outer sums Z = zsum[k] (zero on the padding); the tiles of the first column
// panel also write F1[k] = S_j f~[k]
__global__ __launch_bounds__(256) void rot_out_panel(RotTab tb, const int* __restrict__ bs,
                                                     Panels pn, const double* __restrict__ zsum,
                                                     const double* __restrict__ fy,
                                                     const double* __restrict__ fsum, int np,
                                                     int nsym, int p0, double* __restrict__ T1,
                                                     double* __restrict__ F1) {
  __shared__ double Ls[kRotPanel * kRotLdP], Ms[kRotPanel * kRotLdP], Rs[kRotPanel * kRotLdP];
  __shared__ double xs[kRotPanel];
  const int k = blockIdx.y, j = (p0 + k) % nsym;
  const long nn = (long)np * np;
  panel_run<true>(tb, bs, pn, j, zsum + (long)k * nn, fy, fsum + (long)k * np, np,
                  T1 + (long)k * nn, F1 + (long)k * np, Ls, Ms, Rs, xs);
}

bool rot_panels_fit(int dmax) { return dmax <= kRotPanelDmax; }

// panel starts for consecutive blocks: whole blocks, at most kRotPanel modes
// per panel, then the padding n .. np - 1 as one more panel when np > n
// (start has room for n + 2 entries)
int rot_panels(const int* bs, const int* d, int n, int np, int* start, int* nreal) {
  int np_ = 0, m = 0;
  while (m < n) {
    start[np_++] = m;
    int e = m;
    while (e < n && e + d[e] - m <= kRotPanel) e += d[e];
    m = e > m ? e : m + d[m];   // a block wider than a panel is its own panel
  }
  *nreal = np_;
  if (np > n) start[np_++] = n;
  start[np_] = np;
  return np_;
}

int rot_setup(Ctx* c, const double* X_dev, const double* Y_dev, const double* rho_dev) {
  const long count = c->Mp * c->np;
  hipLaunchKernelGGL(rot_copy_configs, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, c->stream,
                     X_dev, Y_dev, rho_dev, c->n, c->np, (long)c->N, c->Mp, c->U, c->F, c->rho_img);
  TDS_CHECK_HIP(hipGetLastError());
  return 0;
}

// the perturbed averages of nc chains whose padded inputs are in c->R1,
// c->Y1, by nc * rot_nsym rotated passes; unpadded outputs at f_dev, d2v_dev
int run_pa_rot(Ctx* c, int nc, int apply_d4, int prec, double* f_dev, double* d2v_dev,
               hipStream_t s) {
  const int n = c->n, np = c->np, ns = c->rot_nsym;
  const long nn = (long)np * np;
  const RotTab tb{c->rt_cols, c->rt_d, c->rt_sr, c->rt_sc, c->rt_dmax, n};
  hipLaunchKernelGGL(rot_copy, dim3(64), dim3(256), 0, s, c->R1, (long)nc * np, c->R1_base);
  hipLaunchKernelGGL(rot_copy, dim3(1024), dim3(256), 0, s, c->Y1, (long)nc * nn, c->Y1_base);
  hipLaunchKernelGGL(rot_zero, dim3(64), dim3(256), 0, s, (long)nc * n, f_dev);
  hipLaunchKernelGGL(rot_zero, dim3(1024), dim3(256), 0, s, (long)nc * n * n, d2v_dev);
  TDS_CHECK_HIP(hipGetLastError());
  const int pairs = nc * ns;
  const double normf = 1.0 / ns, normd = 1.0 / (c->n_eff * ns);
  const unsigned eb = (unsigned)((nn + 255) / 256);
  const Panels pn{c->rt_pstart, c->rt_npanel, c->rt_nreal};
  const unsigned npanel2 = (unsigned)(pn.npanel * pn.npanel);
  for (int p0 = 0; p0 < pairs; p0 += kMaxChains) {
    const int g = pairs - p0 < kMaxChains ? pairs - p0 : kMaxChains;
    if (c->rt_panels) {
      hipLaunchKernelGGL(rot_in_panel, dim3(npanel2, g), dim3(256), 0, s, tb, c->rt_bs, pn,
                         c->Y1_base, c->R1_base, np, ns, p0, c->R1, c->Y1);
    } else {
      hipLaunchKernelGGL(rot_in_right, dim3(eb, g), dim3(256), 0, s, tb, c->Y1_base, np, ns, p0,
                         c->rot_t1);
      hipLaunchKernelGGL(rot_in_left, dim3(eb, g), dim3(256), 0, s, tb, c->rot_t1, c->R1_base,
                         np, ns, p0, c->R1, c->Y1);
    }
    TDS_CHECK_HIP(hipGetLastError());
    const int rc = prec == PREC_FP64 ? launch_pa_fp64(c, g, apply_d4, s, nullptr)
                                     : launch_pa_split(c, g, apply_d4, prec, s, nullptr);
    if (rc) return rc;
    if (c->rt_panels) {
      hipLaunchKernelGGL(rot_out_panel, dim3(npanel2, g), dim3(256), 0, s, tb, c->rt_bs, pn,
                         c->d2v_out, c->fy, c->f_out, np, ns, p0, c->rot_t1, c->rot_f);
    } else {
      hipLaunchKernelGGL(rot_sym, dim3(np / 32, np / 32, g), dim3(256), 0, s, c->d2v_out, c->fy,
                         np, c->rot_t1);
      hipLaunchKernelGGL(rot_out_right, dim3(eb, g), dim3(256), 0, s, tb, c->rot_t1, np, ns, p0,
                         c->rot_t2);
      hipLaunchKernelGGL(rot_out_left, dim3(eb, g), dim3(256), 0, s, tb, c->rot_t2, c->f_out, np,
                         ns, p0, c->rot_t1, c->rot_f);
    }
    hipLaunchKernelGGL(rot_accumulate, dim3((unsigned)(((long)n * n + 255) / 256)), dim3(256), 0, s,
                       c->rot_t1, c->rot_f, n, np, ns, p0, g, normf, normd, f_dev, d2v_dev);
    TDS_CHECK_HIP(hipGetLastError());
  }
  return 0;
}

}  // namespace tdscha
