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
// rotated perturbation (S_j^T R1, S_j^T Y1 S_j).  The flops are those of the
// image form (n_syms passes over N configurations instead of one over N n_syms
// images); the ensemble takes n_syms times less memory.  The (chain, symmetry)
// pairs run through the ordinary kernels 8 at a time and are summed in a fixed
// order (chain-major, symmetry by symmetry).
#include "common.hpp"

namespace tdscha {

int launch_pa_fp64(Ctx* c, int nc, int apply_d4, hipStream_t s, hipEvent_t* ev);
int launch_pa_split(Ctx* c, int nc, int apply_d4, int prec, hipStream_t s, hipEvent_t* ev);

struct RotTab {
  const int* mb;      // block of each mode
  const int* mp;      // position of each mode in its block
  const int* bd;      // block sizes
  const int* bmo;     // offset of each block's modes in bm
  const int* bm;      // block modes, concatenated
  const long* bso;    // offset of each block's matrices in syms
  const double* syms; // per block [n_syms, d, d], row = output mode
};

// S_j[m, m'] for m' the k-th mode of m's block
__device__ __forceinline__ double s_at(const RotTab& t, int j, int m, int k) {
  const int b = t.mb[m], d = t.bd[b];
  return t.syms[t.bso[b] + ((long)j * d + t.mp[m]) * d + k];
}

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

// pair k = blockIdx.y of the pass (p = p0 + k: chain p / nsym, symmetry
// p % nsym): R1o = S_j^T R1, Y1o = S_j^T Y1 S_j (padded np)
__global__ void rot_inputs(RotTab t, const double* __restrict__ R1b, const double* __restrict__ Y1b,
                           int n, int np, int nsym, int p0, double* __restrict__ R1o,
                           double* __restrict__ Y1o) {
  const int k = blockIdx.y;
  const int p = p0 + k, ch = p / nsym, j = p % nsym;
  const long nn = (long)np * np;
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nn) return;
  const int r = (int)(e / np), q = (int)(e % np);
  double v = 0.0;
  if (r < n && q < n) {
    const int br = t.mb[r], bc = t.mb[q];
    const int dr = t.bd[br], dc = t.bd[bc];
    const int* mr = t.bm + t.bmo[br];
    const int* mc = t.bm + t.bmo[bc];
    const double* Y = Y1b + (long)ch * nn;
    for (int a = 0; a < dr; ++a) {
      double u = 0.0;
      for (int b = 0; b < dc; ++b) u += Y[(long)mr[a] * np + mc[b]] * s_at(t, j, mc[b], t.mp[q]);
      v += s_at(t, j, mr[a], t.mp[r]) * u;
    }
  }
  Y1o[(long)k * nn + e] = v;
  if (e < np) {
    double w = 0.0;
    if (e < n) {
      const int br = t.mb[e], dr = t.bd[br];
      const int* mr = t.bm + t.bmo[br];
      for (int a = 0; a < dr; ++a) w += s_at(t, j, mr[a], t.mp[e]) * R1b[(long)ch * np + mr[a]];
    }
    R1o[(long)k * np + e] = w;
  }
}

// add the pass's results, rotated back, to the outputs (unpadded n):
//   f[ch]   += normf S_j f~            (f~ already divided by N_eff)
//   d2v[ch] += normd S_j (fY Z + (fY Z)^T) S_j^T   (Z the raw outer sums)
// one thread per output element, the pairs of the pass in order
__global__ void rot_accumulate(RotTab t, const double* __restrict__ zsum,
                               const double* __restrict__ fsum, const double* __restrict__ fy,
                               int n, int np, int nsym, int p0, int g, double normf, double normd,
                               double* __restrict__ f_all, double* __restrict__ d2v_all) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long nn = (long)np * np;
  if (e < (long)n * n) {
    const int r = (int)(e / n), q = (int)(e % n);
    const int br = t.mb[r], bc = t.mb[q];
    const int dr = t.bd[br], dc = t.bd[bc];
    const int* mr = t.bm + t.bmo[br];
    const int* mc = t.bm + t.bmo[bc];
    for (int k = 0; k < g; ++k) {
      const int p = p0 + k, ch = p / nsym, j = p % nsym;
      const double* z = zsum + (long)k * nn;
      double v = 0.0;
      for (int a = 0; a < dr; ++a) {
        const int ma = mr[a];
        double u = 0.0;
        for (int b = 0; b < dc; ++b) {
          const int mq = mc[b];
          const double d = fy[ma] * z[(long)ma * np + mq] + fy[mq] * z[(long)mq * np + ma];
          u += d * s_at(t, j, q, b);
        }
        v += s_at(t, j, r, a) * u;
      }
      d2v_all[(long)ch * n * n + e] += normd * v;
    }
  }
  if (e < n) {
    const int br = t.mb[e], dr = t.bd[br];
    const int* mr = t.bm + t.bmo[br];
    for (int k = 0; k < g; ++k) {
      const int p = p0 + k, ch = p / nsym, j = p % nsym;
      double w = 0.0;
      for (int a = 0; a < dr; ++a) w += s_at(t, j, (int)e, a) * fsum[(long)k * np + mr[a]];
      f_all[(long)ch * n + e] += normf * w;
    }
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
  const RotTab t{c->rt_mb, c->rt_mp, c->rt_bd, c->rt_bmo, c->rt_bm, c->rt_bso, c->rt_syms};
  hipLaunchKernelGGL(rot_copy, dim3(64), dim3(256), 0, s, c->R1, (long)nc * np, c->R1_base);
  hipLaunchKernelGGL(rot_copy, dim3(1024), dim3(256), 0, s, c->Y1, (long)nc * nn, c->Y1_base);
  hipLaunchKernelGGL(rot_zero, dim3(64), dim3(256), 0, s, (long)nc * n, f_dev);
  hipLaunchKernelGGL(rot_zero, dim3(1024), dim3(256), 0, s, (long)nc * n * n, d2v_dev);
  TDS_CHECK_HIP(hipGetLastError());
  const int pairs = nc * ns;
  const double normf = 1.0 / ns, normd = 1.0 / (c->n_eff * ns);
  for (int p0 = 0; p0 < pairs; p0 += kMaxChains) {
    const int g = pairs - p0 < kMaxChains ? pairs - p0 : kMaxChains;
    hipLaunchKernelGGL(rot_inputs, dim3((unsigned)((nn + 255) / 256), g), dim3(256), 0, s, t,
                       c->R1_base, c->Y1_base, n, np, ns, p0, c->R1, c->Y1);
    TDS_CHECK_HIP(hipGetLastError());
    const int rc = prec == PREC_FP64 ? launch_pa_fp64(c, g, apply_d4, s, nullptr)
                                     : launch_pa_split(c, g, apply_d4, prec, s, nullptr);
    if (rc) return rc;
    hipLaunchKernelGGL(rot_accumulate, dim3((unsigned)(((long)n * n + 255) / 256)), dim3(256), 0, s,
                       t, c->d2v_out, c->f_out, c->fy, n, np, ns, p0, g, normf, normd, f_dev,
                       d2v_dev);
    TDS_CHECK_HIP(hipGetLastError());
  }
  return 0;
}

}  // namespace tdscha
