// Continued-fraction Green function with the standard terminator
// (Modules/DynamicalLanczos.py:5032-5162), one thread per frequency, FP64.
#include <hip/hip_complex.h>

#include "common.hpp"
#include "../../include/tdscha_amd.h"

namespace tdscha {

__device__ inline hipDoubleComplex cdiv1(hipDoubleComplex z) {
  const double d = z.x * z.x + z.y * z.y;
  return make_hipDoubleComplex(z.x / d, -z.y / d);
}

__device__ inline hipDoubleComplex csqrt_(hipDoubleComplex z) {
  const double r = hypot(z.x, z.y);
  double re = sqrt(0.5 * (r + z.x));
  double im = sqrt(0.5 * (r - z.x));
  if (z.y < 0) im = -im;
  return make_hipDoubleComplex(re, im);
}

// coefficient tail averages are passed in (am, bm, cm: signed means)
__global__ void cfrac_kernel(const double* __restrict__ a, const double* __restrict__ b,
                             const double* __restrict__ c, int n_iters, int has_c,
                             const double* __restrict__ w, int n_w, int use_wigner,
                             int use_terminator, double am, double bm, double cm, double smearing,
                             int smooth_ramp, double sign, double shift, double modulus,
                             double* __restrict__ out /* [n_w, 2] */) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_w) return;
  const double wi = w[i];
  const double w2 = use_wigner ? wi * wi : -wi * wi;
  hipDoubleComplex gf;
  if (use_terminator) {
    const double d = am + w2;
    const hipDoubleComplex disc = make_hipDoubleComplex(d * d - 4 * bm * cm, 0.0);
    const hipDoubleComplex sq = csqrt_(disc);
    const double den = 2 * bm * cm;
    const hipDoubleComplex gp = make_hipDoubleComplex((d + sq.x) / den, sq.y / den);
    const hipDoubleComplex gm = make_hipDoubleComplex((d - sq.x) / den, -sq.y / den);
    gf = (hypot(gp.x, gp.y) < hypot(gm.x, gm.y)) ? gp : gm;
  } else {
    const double al = a[n_iters - 1] * sign - sign * shift;
    gf = cdiv1(make_hipDoubleComplex(al + w2, 2.0 * wi * smearing));
  }
  for (int k = n_iters - 2; k >= 0; --k) {
    double ak = a[k] * sign - sign * shift;
    double bk = b[k] * sign;
    double ck = has_c ? c[k] * sign : bk;
    if (use_terminator && smooth_ramp > 0 && k >= n_iters - smooth_ramp) {
      const double alpha = (double)(k - n_iters + smooth_ramp + 1) / smooth_ramp;
      ak = (1 - alpha) * ak + alpha * am;
      bk = (1 - alpha) * bk + alpha * bm;
      ck = (1 - alpha) * ck + alpha * cm;
    }
    const double bc = bk * ck;
    gf = cdiv1(make_hipDoubleComplex(ak + w2 - bc * gf.x, 2.0 * wi * smearing - bc * gf.y));
  }
  if (use_wigner) gf.x = -gf.x;
  out[2 * i] = gf.x * modulus;
  out[2 * i + 1] = gf.y * modulus;
}

}  // namespace tdscha

using namespace tdscha;

extern "C" int tdscha_continued_fraction(int n_a, const double* a, int n_b, const double* b,
                                         const double* c, int n_w, const double* w,
                                         int use_wigner, int use_terminator, int last_average,
                                         double smearing, int smooth_ramp, int reverse,
                                         double shift, double modulus, double* gf_out) {
  TDS_REQUIRE(n_a >= 1 && n_w >= 1 && a && w && gf_out, 2,
              "tdscha_continued_fraction: bad arguments");
  TDS_REQUIRE(n_b >= n_a - 1 && (n_b == 0 || b != nullptr), 2,
              "tdscha_continued_fraction: need at least len(a) - 1 b coefficients");
  TDS_REQUIRE(!use_terminator || (last_average >= 1 && last_average <= n_a && last_average <= n_b),
              2, "tdscha_continued_fraction: last_average out of range");
  const double sign = reverse ? -1.0 : 1.0;
  double am = 0, bm = 0, cm = 0;
  if (use_terminator) {
    // np.mean over the last `last_average` entries of each list (DL:5082-5090)
    for (int k = n_a - last_average; k < n_a; ++k) am += a[k];
    for (int k = n_b - last_average; k < n_b; ++k) bm += b[k];
    am /= last_average;
    bm /= last_average;
    if (c) {
      for (int k = n_b - last_average; k < n_b; ++k) cm += c[k];
      cm /= last_average;
    } else {
      cm = bm;
    }
    am = am * sign - sign * shift;
    bm *= sign;
    cm *= sign;
  }
  const int nbb = n_b > 0 ? n_b : 1;
  double *da = nullptr, *db = nullptr, *dc = nullptr, *dw = nullptr, *dout = nullptr;
  TDS_CHECK_HIP(hipMalloc((void**)&da, sizeof(double) * n_a));
  TDS_CHECK_HIP(hipMalloc((void**)&db, sizeof(double) * nbb));
  TDS_CHECK_HIP(hipMalloc((void**)&dc, sizeof(double) * nbb));
  TDS_CHECK_HIP(hipMalloc((void**)&dw, sizeof(double) * n_w));
  TDS_CHECK_HIP(hipMalloc((void**)&dout, sizeof(double) * 2 * n_w));
  TDS_CHECK_HIP(hipMemcpy(da, a, sizeof(double) * n_a, hipMemcpyHostToDevice));
  if (n_b > 0) {
    TDS_CHECK_HIP(hipMemcpy(db, b, sizeof(double) * n_b, hipMemcpyHostToDevice));
    TDS_CHECK_HIP(hipMemcpy(dc, c ? c : b, sizeof(double) * n_b, hipMemcpyHostToDevice));
  }
  TDS_CHECK_HIP(hipMemcpy(dw, w, sizeof(double) * n_w, hipMemcpyHostToDevice));
  hipLaunchKernelGGL(cfrac_kernel, dim3((n_w + 127) / 128), dim3(128), 0, 0, da, db, dc, n_a,
                     c ? 1 : 0, dw, n_w, use_wigner, use_terminator, am, bm, cm, smearing,
                     smooth_ramp, sign, shift, modulus, dout);
  TDS_CHECK_HIP(hipGetLastError());
  TDS_CHECK_HIP(hipMemcpy(gf_out, dout, sizeof(double) * 2 * n_w, hipMemcpyDeviceToHost));
  (void)hipFree(da); (void)hipFree(db); (void)hipFree(dc); (void)hipFree(dw); (void)hipFree(dout);
  return 0;
}
