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
  for (int k = 0; k < g; ++k) {
    const int ch = (p0 + k) / nsym;
    d2v_all[(long)ch * n * n + e] += normd * T1[(long)k * nn + (long)r * np + q];
    if (e < n) f_all[(long)ch * n + e] += normf * F1[(long)k * np + e];
  }
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
  for (int p0 = 0; p0 < pairs; p0 += kMaxChains) {
    const int g = pairs - p0 < kMaxChains ? pairs - p0 : kMaxChains;
    hipLaunchKernelGGL(rot_in_right, dim3(eb, g), dim3(256), 0, s, tb, c->Y1_base, np, ns, p0,
                       c->rot_t1);
    hipLaunchKernelGGL(rot_in_left, dim3(eb, g), dim3(256), 0, s, tb, c->rot_t1, c->R1_base, np,
                       ns, p0, c->R1, c->Y1);
    TDS_CHECK_HIP(hipGetLastError());
    const int rc = prec == PREC_FP64 ? launch_pa_fp64(c, g, apply_d4, s, nullptr)
                                     : launch_pa_split(c, g, apply_d4, prec, s, nullptr);
    if (rc) return rc;
    hipLaunchKernelGGL(rot_sym, dim3(np / 32, np / 32, g), dim3(256), 0, s, c->d2v_out, c->fy, np,
                       c->rot_t1);
    hipLaunchKernelGGL(rot_out_right, dim3(eb, g), dim3(256), 0, s, tb, c->rot_t1, np, ns, p0,
                       c->rot_t2);
    hipLaunchKernelGGL(rot_out_left, dim3(eb, g), dim3(256), 0, s, tb, c->rot_t2, c->f_out, np, ns,
                       p0, c->rot_t1, c->rot_f);
    hipLaunchKernelGGL(rot_accumulate, dim3((unsigned)(((long)n * n + 255) / 256)), dim3(256), 0, s,
                       c->rot_t1, c->rot_f, n, np, ns, p0, g, normf, normd, f_dev, d2v_dev);
    TDS_CHECK_HIP(hipGetLastError());
  }
  return 0;
}

}  // namespace tdscha
