// Perturbed averages that allocate nothing: the FP64 image-range pass of
// tdscha_perturb_averages_range run straight from caller-owned device
// buffers (X, Y and the packed symmetry tables), with every intermediate in
// one file-scope device array of fixed size.
//
// The images of the range are processed in chunks of Mc rows: each chunk is
// expanded into the scratch array (S_j x_i, S_j y_i, rho), weighted by the
// FP64 weights GEMM and added into the K-split partials of the FP64 outer
// GEMM (the same kernels as the context path, chunks summed in order), then
// one fixed-order reduction and the normalisation by N_eff * n_syms.
#include <mutex>

#include "common.hpp"
#include "../../include/tdscha_amd.h"

namespace tdscha {

__global__ void weights_f64(const double*, const double*, const double*, const double*,
                            const double*, const double*, const double*, int, long, int, long,
                            long, long, long, int, double*);
__global__ void outer_f64(const double*, const double*, const double*, const double*, int, long,
                          int, int, long, long, long, int, int, double*, double*);
__global__ void reduce_ksplit(const double*, const double*, int, int, double*, double*, double);

// 64 MiB, held by every device the library runs on until the process ends
constexpr long kScratchDoubles = 8L * 1024 * 1024;
static __device__ __align__(256) double scratch[kScratchDoubles];

constexpr int kMaxDevices = 64;
constexpr long kChunkMax = 16384;   // image rows per chunk at most
constexpr long kChunkGranule = 256; // chunk rows: whole weights tiles and K-split steps

static std::mutex g_mu;
static double* g_ptr[kMaxDevices] = {nullptr};

// device address of `scratch` on the current device (looked up once per device)
static int scratch_ptr(double** out) {
  int dev = 0;
  TDS_CHECK_HIP(hipGetDevice(&dev));
  TDS_REQUIRE(dev >= 0 && dev < kMaxDevices, -1, "no-allocation entry: device index out of range");
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      set_error(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
      return -1;
    }
    g_ptr[dev] = static_cast<double*>(p);
  }
  *out = g_ptr[dev];
  return 0;
}

// thermal factors fY = 2w/(1+2n), fpsi = (1+2n)/(2w) (CModules/LanczosFunctions.c:14-31),
// padded R1 and symmetric Y1 (zero outside n)
__global__ void na_prep(const double* __restrict__ w, double T, const double* __restrict__ R1,
                        const double* __restrict__ Y1, int n, int np, double* __restrict__ fy,
                        double* __restrict__ fpsi, double* __restrict__ R1p,
                        double* __restrict__ Y1p) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long)np * np) return;
  const int a = (int)(e / np), b = (int)(e % np);
  Y1p[e] = (a < n && b < n) ? 0.5 * (Y1[(long)a * n + b] + Y1[(long)b * n + a]) : 0.0;
  if (e < np) {
    double y = 0.0, p = 0.0, r = 0.0;
    if (e < n) {
      const double nb = T > 0 ? 1.0 / (exp(w[e] * RY_TO_K / T) - 1.0) : 0.0;
      y = 2.0 * w[e] / (1.0 + 2.0 * nb);
      p = (1.0 + 2.0 * nb) / (2.0 * w[e]);
      r = R1[e];
    }
    fy[e] = y;
    fpsi[e] = p;
    R1p[e] = r;
  }
}

// N_eff = sum of rho in a fixed order (one work-group)
__global__ void na_neff(const double* __restrict__ rho, int N, double* __restrict__ out) {
  __shared__ double red[256];
  double s = 0.0;
  for (int i = threadIdx.x; i < N; i += 256) s += rho[i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = red[0];
}

// images b0 .. b0 + Mc - 1 of the packed symmetry layout of
// Modules/DynamicalLanczos.py:807-822 (syms[block, sym, dmax, dmax],
// degenerate_space[block, dmax]); rows outside [lo, hi) are zero.  Entries
// that fall outside the tables are skipped rather than read.
__global__ void na_expand(const double* __restrict__ X, const double* __restrict__ Y,
                          const double* __restrict__ rho, const double* __restrict__ syms,
                          const int* __restrict__ n_deg, const int* __restrict__ deg_space,
                          const int* __restrict__ sbi, int n, int np, int n_syms, int n_blocks,
                          int dmax, long b0, long lo, long hi, double* __restrict__ U,
                          double* __restrict__ F, double* __restrict__ rho_img) {
  const long row = blockIdx.x;
  const int m = threadIdx.x + blockIdx.y * blockDim.x;
  if (m >= np) return;
  const long gb = b0 + row;
  const bool in = gb >= lo && gb < hi;
  double su = 0.0, sf = 0.0;
  long i = 0;
  if (in) {
    i = gb / n_syms;
    const int j = (int)(gb % n_syms);
    const int b = m < n ? sbi[m] : -1;
    if (b >= 0 && b < n_blocks) {
      int d = n_deg[b];
      d = d < 0 ? 0 : (d > dmax ? dmax : d);
      const int* ds = deg_space + (long)b * dmax;
      int pos = -1;
      for (int k = 0; k < d; ++k)
        if (ds[k] == m) pos = k;
      if (pos >= 0) {
        const double* S = syms + (((long)b * n_syms + j) * dmax + pos) * dmax;
        const double* x = X + i * n;
        const double* y = Y + i * n;
        for (int k = 0; k < d; ++k) {
          const int mk = ds[k];
          if (mk < 0 || mk >= n) continue;
          su += S[k] * x[mk];
          sf += S[k] * y[mk];
        }
      }
    }
  }
  U[row * np + m] = su;
  F[row * np + m] = sf;
  if (m == 0) rho_img[row] = in ? rho[i] : 0.0;
}

// d2v = (fY Z + (fY Z)^T) / (N_eff n_syms), f / (N_eff n_syms), unpadded
__global__ void na_finish(const double* __restrict__ zsum, const double* __restrict__ fsum,
                          const double* __restrict__ fy, const double* __restrict__ neff, int n,
                          int np, int n_syms, double* __restrict__ f_out,
                          double* __restrict__ d2v_out) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const double norm = 1.0 / (neff[0] * n_syms);
  if (e < n) f_out[e] = fsum[e] * norm;
  if (e >= (long)n * n) return;
  const int mu = (int)(e / n), nu = (int)(e % n);
  d2v_out[e] = norm * (fy[mu] * zsum[(long)mu * np + nu] + fy[nu] * zsum[(long)nu * np + mu]);
}

struct NaLayout {
  long fy, fpsi, R1p, Y1p, coef, U, F, rho, zpart, fpart, zsum, fsum, neff, total;
  long Mc;
  int ks;
};

static long al32(long x) { return (x + 31) / 32 * 32; }

static NaLayout na_layout(int np, int ks, long Mc) {
  NaLayout L;
  long o = 0;
  auto take = [&](long n) { const long at = o; o += al32(n); return at; };
  L.fy = take(np);
  L.fpsi = take(np);
  L.R1p = take(np);
  L.Y1p = take((long)np * np);
  L.zpart = take((long)ks * np * np);
  L.fpart = take((long)ks * np);
  L.zsum = take((long)np * np);
  L.fsum = take(np);
  L.neff = take(1);
  L.coef = take(4 * Mc);
  L.rho = take(Mc);
  L.U = take(Mc * np);
  L.F = take(Mc * np);
  L.total = o;
  L.Mc = Mc;
  L.ks = ks;
  return L;
}

// largest chunk (a multiple of the granule, at most kChunkMax and at most the
// range rounded up) that fits, with 16 K-split slices, else 8
static bool na_plan(int np, long range, NaLayout* out) {
  const long want = std::min(kChunkMax, round_up_l(range, kChunkGranule));
  for (int ks : {16, 8}) {
    const NaLayout fixed = na_layout(np, ks, 0);
    const long room = kScratchDoubles - fixed.total - 4 * 32;  // alignment slack
    if (room <= 0) continue;
    long Mc = room / (2L * np + 5) / kChunkGranule * kChunkGranule;
    Mc = std::min(Mc, want);
    if (Mc < kChunkGranule) continue;
    const NaLayout L = na_layout(np, ks, Mc);
    if (L.total > kScratchDoubles) continue;
    *out = L;
    return true;
  }
  return false;
}

}  // namespace tdscha

using namespace tdscha;

extern "C" {

int tdscha_perturb_averages_noalloc(int n_modes, int n_configs, const double* X, const double* Y,
                                    const double* w, const double* rho, const double* R1,
                                    const double* Y1, double T, int apply_d4, int n_syms,
                                    int n_blocks, int dmax, const double* syms,
                                    const int* n_degeneracy, const int* degenerate_space,
                                    const int* sym_block_id, long start_index, long end_index,
                                    double* f_out, double* d2v_out, void* stream) {
  TDS_REQUIRE(n_modes > 0 && n_configs > 0 && n_syms > 0 && n_blocks > 0 && dmax > 0, 2,
              "tdscha_perturb_averages_noalloc: sizes must be positive");
  TDS_REQUIRE(dmax <= n_modes && n_blocks <= n_modes, 2,
              "tdscha_perturb_averages_noalloc: dmax and n_blocks cannot exceed n_modes");
  TDS_REQUIRE(X && Y && w && rho && R1 && Y1 && syms && n_degeneracy && degenerate_space &&
                  sym_block_id && f_out && d2v_out,
              2, "tdscha_perturb_averages_noalloc: NULL pointer");
  TDS_REQUIRE(apply_d4 == 0 || apply_d4 == 1, 2, "the apply_d4 value passed must be either 0 or 1");
  const long M = (long)n_configs * n_syms;
  TDS_REQUIRE(start_index <= end_index, 2,
              "tdscha_perturb_averages_noalloc: start_index must not exceed end_index");
  TDS_REQUIRE(start_index >= 1, 2, "tdscha_perturb_averages_noalloc: start_index must be >= 1");
  TDS_REQUIRE(end_index <= M, 2,
              "tdscha_perturb_averages_noalloc: end_index cannot exceed n_configs * n_syms");
  const int np = round_up(n_modes, kTile);
  const long lo = start_index - 1, hi = end_index;
  NaLayout L;
  TDS_REQUIRE(na_plan(np, hi - lo, &L), 2,
              "tdscha_perturb_averages_noalloc: n_modes too large for the 64 MiB scratch "
              "(at most 768 modes)");
  double* base = nullptr;
  const int rc = scratch_ptr(&base);
  if (rc) return rc;
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  double* fy = base + L.fy;
  double* fpsi = base + L.fpsi;
  double* R1p = base + L.R1p;
  double* Y1p = base + L.Y1p;
  double* coef = base + L.coef;
  double* U = base + L.U;
  double* F = base + L.F;
  double* rimg = base + L.rho;
  double* zpart = base + L.zpart;
  double* fpart = base + L.fpart;
  double* zsum = base + L.zsum;
  double* fsum = base + L.fsum;
  double* neff = base + L.neff;
  const long nn = (long)np * np;
  hipLaunchKernelGGL(na_prep, dim3((unsigned)((nn + 255) / 256)), dim3(256), 0, s, w, T, R1, Y1,
                     n_modes, np, fy, fpsi, R1p, Y1p);
  hipLaunchKernelGGL(na_neff, dim3(1), dim3(256), 0, s, rho, n_configs, neff);
  const int ntile = np / kTile;
  const long Mc = L.Mc;
  const long rps = Mc / L.ks;   // a multiple of 16
  int chunk = 0;
  for (long b0 = lo; b0 < hi; b0 += Mc, ++chunk) {
    hipLaunchKernelGGL(na_expand, dim3((unsigned)Mc, (np + 127) / 128), dim3(128), 0, s, X, Y,
                       rho, syms, n_degeneracy, degenerate_space, sym_block_id, n_modes, np,
                       n_syms, n_blocks, dmax, b0, lo, hi, U, F, rimg);
    // one chain: work-group id = row tile (Mc / 64 tiles)
    hipLaunchKernelGGL(weights_f64, dim3((unsigned)(Mc / 64)), dim3(256), 0, s, U, F, Y1p, R1p,
                       fy, fpsi, rimg, np, Mc, apply_d4, 0L, Mc / 64, 0L, Mc, 1, coef);
    hipLaunchKernelGGL(outer_f64, dim3((unsigned)(ntile * ntile * L.ks)), dim3(kF64OuterThreads),
                       0, s, U, F,
                       coef, fy, np, Mc, L.ks, ntile, 0L, Mc, rps, chunk > 0 ? 1 : 0, 1, zpart,
                       fpart);
  }
  hipLaunchKernelGGL(reduce_ksplit, dim3((unsigned)((nn + 255) / 256), 1), dim3(256), 0, s, zpart,
                     fpart, np, L.ks, zsum, fsum, 1.0);
  const long n2 = (long)n_modes * n_modes;
  hipLaunchKernelGGL(na_finish, dim3((unsigned)((n2 + 255) / 256)), dim3(256), 0, s, zsum, fsum,
                     fy, neff, n_modes, np, n_syms, f_out, d2v_out);
  TDS_CHECK_HIP(hipGetLastError());
  return 0;
}

}  // extern "C"
