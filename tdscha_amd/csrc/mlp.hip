// Small fully connected network in FP32 for the learned continued-fraction
// terminator: n_in -> width (x n_hidden, SiLU) -> n_out (linear), with
// forward, backward and Adam on the device.  Two losses:
//   kind 0  weighted mean-squared error against FP32 targets;
//   kind 1  spectral loss: the n_out = 10 outputs are the tail parameters of a
//           continued fraction (asymptotic a_inf, b_inf and four corrections
//           each to the next a and b), appended to the first K coefficients
//           of a chain; the loss is the mean squared deviation of the
//           resulting spectrum from a reference spectrum, and its gradient
//           comes from the adjoint of the continued-fraction recurrence (FP64).
// Every reduction is a fixed-order loop or tree (no atomics), so training is
// reproducible.  Parameters are one contiguous FP32 array
//   [W0 (d1 x d0), b0 (d1), W1 (d2 x d1), b1 (d2), ...], W row-major [out][in].
#include <cmath>
#include <vector>

#include "common.hpp"
#include "../../include/tdscha_amd.h"

namespace tdscha {

constexpr int kCfOut = 10;   // a_inf, b_inf, da[4], db[4]
constexpr int kCfTail = 4;

struct Mlp {
  int n_layers = 0;              // linear layers = n_hidden + 1
  std::vector<int> dims;         // d0 = n_in, ..., d_L = n_out
  std::vector<long> w_off, b_off;
  long n_params = 0;
  int max_batch = 0, max_dim = 0;
  float* params = nullptr;       // device
  float* grads = nullptr;
  float* m1 = nullptr;           // Adam moments
  float* m2 = nullptr;
  float* act = nullptr;          // [n_layers + 1][max_batch][max_dim] layer outputs (act[0] = input)
  float* pre = nullptr;          // [n_layers][max_batch][max_dim] pre-activations
  float* dz = nullptr;           // [2][max_batch][max_dim] backward ping-pong
  float* target = nullptr;       // [max_batch][n_out] (kind 0)
  float* wout = nullptr;         // [n_out] loss weights (kind 0)
  float* loss_part = nullptr;    // [max_batch]
  int* idx = nullptr;            // [max_batch] rows of the data arrays in use
  // data arrays (device): either one batch (direct calls) or a training set
  float* data_x = nullptr;       // [n][n_in]
  float* data_t = nullptr;       // [n][n_out]       kind 0
  double* data_ch = nullptr;     // [n][2K + 2]      kind 1: a[K], b[K], b0, 1/max|S_ref|
  double* data_ref = nullptr;    // [n][n_w]         kind 1: S_ref / max|S_ref|
  long data_cap_x = 0, data_cap_t = 0, data_cap_ch = 0, data_cap_ref = 0;
  // spectral loss configuration
  int cf_K = 0, cf_nw = 0;
  double cf_eta = 0.0, cf_scale = 1.0;
  double* cf_w = nullptr;        // [n_w]
  double* cf_mean = nullptr;     // [10]
  double* cf_std = nullptr;      // [10]
  long step = 0;                 // Adam steps taken
  hipStream_t stream = nullptr;
};

__device__ __forceinline__ float silu(float z) { return z / (1.0f + expf(-z)); }
__device__ __forceinline__ float dsilu(float z) {
  const float s = 1.0f / (1.0f + expf(-z));
  return s * (1.0f + z * (1.0f - s));
}

// Z[b][o] = sum_i X[b][i] W[o][i] + bias[o]; A = act ? SiLU(Z) : Z.  16 x 16
// output tile per block, K in chunks of 16 through LDS.
__global__ void mlp_linear_fwd(const float* __restrict__ X, const float* __restrict__ W,
                               const float* __restrict__ bias, float* __restrict__ Z,
                               float* __restrict__ A, int B, int in, int out, int ld, int act) {
  __shared__ float xs[16][17], ws[16][17];
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
  const int b = blockIdx.y * 16 + ty, o = blockIdx.x * 16 + tx;
  float s = 0.f;
  for (int k0 = 0; k0 < in; k0 += 16) {
    const int xk = k0 + tx;
    xs[ty][tx] = (b < B && xk < in) ? X[(long)b * ld + xk] : 0.f;
    const int wo = blockIdx.x * 16 + ty;
    ws[ty][tx] = (wo < out && xk < in) ? W[(long)wo * in + xk] : 0.f;
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 16; ++k) s += xs[ty][k] * ws[tx][k];
    __syncthreads();
  }
  if (b < B && o < out) {
    const float z = s + bias[o];
    Z[(long)b * ld + o] = z;
    A[(long)b * ld + o] = act ? silu(z) : z;
  }
}

// kind 0: dY = 2 w (Y - T) / B, loss_part = sum_o w (Y - T)^2 / B
__global__ void mlp_mse_grad(const float* __restrict__ Y, const float* __restrict__ T,
                             const float* __restrict__ w, float* __restrict__ dY,
                             float* __restrict__ loss_part, int B, int out, int ld) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B) return;
  float l = 0.f;
  for (int o = 0; o < out; ++o) {
    const float d = Y[(long)b * ld + o] - T[(long)b * out + o];
    dY[(long)b * ld + o] = 2.0f * w[o] * d / (float)B;
    l += w[o] * d * d;
  }
  loss_part[b] = l / (float)B;
}

struct cplx {
  double r, i;
};
__device__ __forceinline__ cplx cmul(cplx a, cplx b) {
  return {a.r * b.r - a.i * b.i, a.r * b.i + a.i * b.r};
}
__device__ __forceinline__ cplx cinv(cplx z) {
  const double d = z.r * z.r + z.i * z.i;
  return {z.r / d, -z.i / d};
}
__device__ __forceinline__ cplx cscale(cplx a, double s) { return {a.r * s, a.i * s}; }

// kind 1: one block per sample, threads stride over the frequency grid.
// Tail: terminator root from (A, B) = b0 (p0, p1) as in the reference's
// terminator (no smearing in its discriminant, smaller-modulus root, ties to
// the minus root), then the four predicted levels a_j = b0 (p0 + p_{2+j}),
// b_j = b0 (p1 + p_{6+j}), then the chain's K measured levels.  The
// derivative of g with respect to the ten parameters is carried through the
// tail levels; above them dG/dg_K = prod_i b_i^2 g_i^2.
__global__ void __launch_bounds__(256) cf_loss_kernel(
    const float* __restrict__ Y, int ld, const int* __restrict__ idx,
    const double* __restrict__ chains, const double* __restrict__ ref,
    const double* __restrict__ wgrid, int n_w, int K, double eta,
    const double* __restrict__ out_mean, const double* __restrict__ out_std, double scale,
    float* __restrict__ dY, float* __restrict__ loss_part) {
  __shared__ double red[kCfOut + 1][256];
  const int bi = blockIdx.x;
  const long row = idx[bi];
  const double* ch = chains + row * (2 * K + 2);
  const double* av = ch;
  const double* bv = ch + K;
  const double b0 = ch[2 * K];
  const double inv_s = ch[2 * K + 1];
  const double* rf = ref + row * n_w;
  double p[kCfOut];
#pragma unroll
  for (int o = 0; o < kCfOut; ++o) p[o] = out_mean[o] + out_std[o] * (double)Y[(long)bi * ld + o];
  double lsum = 0.0, gacc[kCfOut];
#pragma unroll
  for (int o = 0; o < kCfOut; ++o) gacc[o] = 0.0;
  const double A = b0 * p[0], Bt = b0 * p[1];
  for (int k = threadIdx.x; k < n_w; k += blockDim.x) {
    const double om = wgrid[k];
    const double w2 = om * om, sm = 2.0 * om * eta;
    // terminator root and its derivatives
    const double d = A + w2;
    const double disc = d * d - 4.0 * Bt * Bt;
    const double den = 2.0 * Bt * Bt;
    cplx g, delta;
    if (disc < 0.0) {
      const double sq = sqrt(-disc);
      g = {d / den, -sq / den};
      delta = {0.0, -sq};
    } else {
      const double sq = sqrt(disc);
      const double gp = (d + sq) / den, gmi = (d - sq) / den;
      if (fabs(gp) < fabs(gmi)) {
        g = {gp, 0.0};
        delta = {sq, 0.0};
      } else {
        g = {gmi, 0.0};
        delta = {-sq, 0.0};
      }
    }
    cplx dg[kCfOut];
#pragma unroll
    for (int o = 0; o < kCfOut; ++o) dg[o] = {0.0, 0.0};
    if (delta.r != 0.0 || delta.i != 0.0) {
      const cplx idl = cinv(delta);
      dg[0] = cscale(cmul(g, idl), b0);
      dg[1] = cscale(cmul(cmul(g, g), idl), -2.0 * Bt * b0);
    }
    // predicted tail levels, bottom up
#pragma unroll
    for (int j = kCfTail - 1; j >= 0; --j) {
      const double aj = b0 * (p[0] + p[2 + j]);
      const double bj = b0 * (p[1] + p[6 + j]);
      const double bj2 = bj * bj;
      const cplx gn = cinv({aj + w2 - bj2 * g.r, sm - bj2 * g.i});
      const cplx mg2 = cscale(cmul(gn, gn), -1.0);
      const cplx dbt = cscale(g, -2.0 * bj * b0);   // d(e - b^2 g)/dp through b: -2 b b0 g
#pragma unroll
      for (int o = 0; o < kCfOut; ++o) {
        cplx t = cscale(dg[o], -bj2);
        if (o == 0 || o == 2 + j) t.r += b0;
        if (o == 1 || o == 6 + j) {
          t.r += dbt.r;
          t.i += dbt.i;
        }
        dg[o] = cmul(mg2, t);
      }
      g = gn;
    }
    // the chain's own levels
    cplx D = {1.0, 0.0};
    for (int i = K - 1; i >= 0; --i) {
      const double bi2 = bv[i] * bv[i];
      const cplx gn = cinv({av[i] + w2 - bi2 * g.r, sm - bi2 * g.i});
      D = cscale(cmul(D, cmul(gn, gn)), bi2);
      g = gn;
    }
    // Wigner convention: spectrum = -Im(G)
    const double r = -g.i * inv_s - rf[k];
    lsum += r * r;
    const double c = 2.0 * r * inv_s;
#pragma unroll
    for (int o = 0; o < kCfOut; ++o) {
      const cplx t = cmul(D, dg[o]);
      gacc[o] -= c * t.i;
    }
  }
  red[0][threadIdx.x] = lsum;
#pragma unroll
  for (int o = 0; o < kCfOut; ++o) red[o + 1][threadIdx.x] = gacc[o];
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
#pragma unroll
      for (int q = 0; q <= kCfOut; ++q) red[q][threadIdx.x] += red[q][threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    loss_part[bi] = (float)(scale * red[0][0] / n_w);
#pragma unroll
    for (int o = 0; o < kCfOut; ++o)
      dY[(long)bi * ld + o] = (float)(scale * red[o + 1][0] / n_w * out_std[o]);
  }
}

// dW[o][i] = sum_b dZ[b][o] Ain[b][i]; db[o] = sum_b dZ[b][o] (fixed order in b)
__global__ void mlp_linear_bwd_w(const float* __restrict__ dZ, const float* __restrict__ Ain,
                                 float* __restrict__ dW, float* __restrict__ db, int B, int in,
                                 int out, int ld) {
  __shared__ float zs[16][17], as[16][17];
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
  const int o = blockIdx.y * 16 + ty, i = blockIdx.x * 16 + tx;
  float s = 0.f, sb = 0.f;
  for (int b0 = 0; b0 < B; b0 += 16) {
    const int zb = b0 + tx;
    zs[ty][tx] = (zb < B && o < out) ? dZ[(long)zb * ld + o] : 0.f;   // zs[o][b]
    const int ab = b0 + ty;
    as[ty][tx] = (ab < B && i < in) ? Ain[(long)ab * ld + i] : 0.f;   // as[b][i]
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      s += zs[ty][k] * as[k][tx];
      sb += zs[ty][k];
    }
    __syncthreads();
  }
  if (o < out && i < in) dW[(long)o * in + i] = s;
  if (o < out && blockIdx.x == 0 && tx == 0) db[o] = sb;
}

// dZprev[b][i] = (sum_o dZ[b][o] W[o][i]) * SiLU'(Zprev[b][i])
__global__ void mlp_linear_bwd_x(const float* __restrict__ dZ, const float* __restrict__ W,
                                 const float* __restrict__ Zprev, float* __restrict__ dZprev,
                                 int B, int in, int out, int ld) {
  __shared__ float zs[16][17], ws[16][17];
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
  const int b = blockIdx.y * 16 + ty, i = blockIdx.x * 16 + tx;
  float s = 0.f;
  for (int o0 = 0; o0 < out; o0 += 16) {
    const int zo = o0 + tx;
    zs[ty][tx] = (b < B && zo < out) ? dZ[(long)b * ld + zo] : 0.f;   // zs[b][o]
    const int wo = o0 + ty;
    ws[ty][tx] = (wo < out && i < in) ? W[(long)wo * in + i] : 0.f;   // ws[o][i]
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 16; ++k) s += zs[ty][k] * ws[k][tx];
    __syncthreads();
  }
  if (b < B && i < in) dZprev[(long)b * ld + i] = s * dsilu(Zprev[(long)b * ld + i]);
}

__global__ void mlp_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                         float* __restrict__ v, long n, float lr, float b1, float b2, float eps,
                         float bc1, float bc2) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const float gi = g[e];
  const float mi = b1 * m[e] + (1.f - b1) * gi;
  const float vi = b2 * v[e] + (1.f - b2) * gi * gi;
  m[e] = mi;
  v[e] = vi;
  p[e] -= lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
}

// minibatch rows into the first layer's input (and kind-0 targets)
__global__ void mlp_gather(const float* __restrict__ dx, const float* __restrict__ dt,
                           const int* __restrict__ idx, int B, int n_in, int n_out, int ld,
                           float* __restrict__ x, float* __restrict__ t) {
  const int b = blockIdx.x;
  if (b >= B) return;
  const long r = idx[b];
  for (int i = threadIdx.x; i < n_in; i += blockDim.x) x[(long)b * ld + i] = dx[r * n_in + i];
  if (dt)
    for (int o = threadIdx.x; o < n_out; o += blockDim.x) t[(long)b * n_out + o] = dt[r * n_out + o];
}

static float* act_l(Mlp* m, int l) { return m->act + (long)l * m->max_batch * m->max_dim; }
static float* pre_l(Mlp* m, int l) { return m->pre + (long)l * m->max_batch * m->max_dim; }

static void forward(Mlp* m, int B) {
  const int ld = m->max_dim;
  for (int l = 0; l < m->n_layers; ++l) {
    const int in = m->dims[l], out = m->dims[l + 1];
    dim3 g((out + 15) / 16, (B + 15) / 16);
    hipLaunchKernelGGL(mlp_linear_fwd, g, dim3(256), 0, m->stream, act_l(m, l),
                       m->params + m->w_off[l], m->params + m->b_off[l], pre_l(m, l),
                       act_l(m, l + 1), B, in, out, ld, l + 1 < m->n_layers ? 1 : 0);
  }
}

// loss partials and the gradient with respect to the network outputs
static void loss_head(Mlp* m, int kind, int B) {
  const int ld = m->max_dim, L = m->n_layers;
  if (kind == 0) {
    hipLaunchKernelGGL(mlp_mse_grad, dim3((B + 255) / 256), dim3(256), 0, m->stream,
                       act_l(m, L), m->target, m->wout, m->dz, m->loss_part, B, m->dims[L], ld);
  } else {
    hipLaunchKernelGGL(cf_loss_kernel, dim3(B), dim3(256), 0, m->stream, act_l(m, L), ld, m->idx,
                       m->data_ch, m->data_ref, m->cf_w, m->cf_nw, m->cf_K, m->cf_eta,
                       m->cf_mean, m->cf_std, m->cf_scale / B, m->dz, m->loss_part);
  }
}

static void backward(Mlp* m, int B) {
  const int ld = m->max_dim, L = m->n_layers;
  float* dzb[2] = {m->dz, m->dz + (long)m->max_batch * m->max_dim};
  int cur = 0;
  for (int l = L - 1; l >= 0; --l) {
    const int in = m->dims[l], out = m->dims[l + 1];
    hipLaunchKernelGGL(mlp_linear_bwd_w, dim3((in + 15) / 16, (out + 15) / 16), dim3(256), 0,
                       m->stream, dzb[cur], act_l(m, l), m->grads + m->w_off[l],
                       m->grads + m->b_off[l], B, in, out, ld);
    if (l > 0) {
      hipLaunchKernelGGL(mlp_linear_bwd_x, dim3((in + 15) / 16, (B + 15) / 16), dim3(256), 0,
                         m->stream, dzb[cur], m->params + m->w_off[l], pre_l(m, l - 1),
                         dzb[1 - cur], B, in, out, ld);
      cur = 1 - cur;
    }
  }
}

static void adam(Mlp* m, float lr, float b1, float b2, float eps) {
  m->step += 1;
  const float bc1 = 1.f - powf(b1, (float)m->step), bc2 = 1.f - powf(b2, (float)m->step);
  hipLaunchKernelGGL(mlp_adam, dim3((unsigned)((m->n_params + 255) / 256)), dim3(256), 0,
                     m->stream, m->params, m->grads, m->m1, m->m2, m->n_params, lr, b1, b2, eps,
                     bc1, bc2);
}

template <typename T>
static int ensure(T** p, long* cap, long n) {
  if (*cap >= n && *p) return 0;
  if (*p) TDS_CHECK_HIP(hipFree(*p));
  *p = nullptr;
  *cap = 0;
  TDS_CHECK_HIP(hipMalloc((void**)p, sizeof(T) * (n > 0 ? n : 1)));
  *cap = n;
  return 0;
}

// Upload n rows of data for the given loss kind to the device data arrays
// (synchronous: the host arrays may be released on return).
static int upload_data(Mlp* m, int kind, long n, const float* x, const float* t,
                       const double* chains, const double* ref) {
  const int n_in = m->dims[0], n_out = m->dims[m->n_layers];
  TDS_CHECK_HIP(hipStreamSynchronize(m->stream));
  int rc = ensure(&m->data_x, &m->data_cap_x, n * n_in);
  if (rc) return rc;
  TDS_CHECK_HIP(hipMemcpy(m->data_x, x, sizeof(float) * n * n_in, hipMemcpyHostToDevice));
  if (kind == 0) {
    rc = ensure(&m->data_t, &m->data_cap_t, n * n_out);
    if (rc) return rc;
    TDS_CHECK_HIP(hipMemcpy(m->data_t, t, sizeof(float) * n * n_out, hipMemcpyHostToDevice));
  } else {
    const long nch = n * (2L * m->cf_K + 2), nref = n * (long)m->cf_nw;
    rc = ensure(&m->data_ch, &m->data_cap_ch, nch);
    if (rc) return rc;
    rc = ensure(&m->data_ref, &m->data_cap_ref, nref);
    if (rc) return rc;
    TDS_CHECK_HIP(hipMemcpy(m->data_ch, chains, sizeof(double) * nch, hipMemcpyHostToDevice));
    TDS_CHECK_HIP(hipMemcpy(m->data_ref, ref, sizeof(double) * nref, hipMemcpyHostToDevice));
  }
  return 0;
}

// one minibatch: gather rows idx[0..B), forward, loss head, backward
static void batch_pass(Mlp* m, int kind, int B) {
  const int n_in = m->dims[0], n_out = m->dims[m->n_layers];
  hipLaunchKernelGGL(mlp_gather, dim3(B), dim3(128), 0, m->stream, m->data_x,
                     kind == 0 ? m->data_t : nullptr, m->idx, B, n_in, n_out, m->max_dim,
                     act_l(m, 0), m->target);
  forward(m, B);
  loss_head(m, kind, B);
  backward(m, B);
}

static int check_kind(Mlp* m, int kind, const float* t, const double* chains, const double* ref) {
  TDS_REQUIRE(kind == 0 || kind == 1, 2, "mlp: loss kind must be 0 (MSE) or 1 (spectral)");
  if (kind == 0) {
    TDS_REQUIRE(t != nullptr, 2, "mlp: MSE loss needs targets");
  } else {
    TDS_REQUIRE(m->cf_K > 0, 2, "mlp: spectral loss not configured (tdscha_mlp_set_cf_loss)");
    TDS_REQUIRE(chains != nullptr && ref != nullptr, 2, "mlp: spectral loss needs chains and ref");
  }
  return 0;
}

static int identity_idx(Mlp* m, int B) {
  std::vector<int> id(B);
  for (int i = 0; i < B; ++i) id[i] = i;
  TDS_CHECK_HIP(hipMemcpy(m->idx, id.data(), sizeof(int) * B, hipMemcpyHostToDevice));
  return 0;
}

static int read_loss(Mlp* m, int B, float* loss) {
  std::vector<float> lp(B);
  TDS_CHECK_HIP(hipMemcpyAsync(lp.data(), m->loss_part, sizeof(float) * B, hipMemcpyDeviceToHost,
                               m->stream));
  TDS_CHECK_HIP(hipStreamSynchronize(m->stream));
  double s = 0.0;
  for (float v : lp) s += v;
  *loss = (float)s;
  return 0;
}

}  // namespace tdscha

using namespace tdscha;

extern "C" {

int tdscha_mlp_create(tdscha_mlp** out, int n_in, int width, int n_hidden, int n_out,
                      int max_batch, const float* init_params) {
  TDS_REQUIRE(out != nullptr, 2, "tdscha_mlp_create: NULL out");
  TDS_REQUIRE(n_in > 0 && width > 0 && n_hidden >= 0 && n_out > 0 && max_batch > 0, 2,
              "tdscha_mlp_create: sizes must be positive");
  TDS_REQUIRE(n_in <= 4096 && width <= 4096 && n_out <= 4096 && n_hidden <= 64 &&
                  max_batch <= 65536,
              2, "tdscha_mlp_create: sizes above the supported maxima");
  *out = nullptr;
  Mlp* m = new Mlp();
  m->n_layers = n_hidden + 1;
  m->dims.push_back(n_in);
  for (int h = 0; h < n_hidden; ++h) m->dims.push_back(width);
  m->dims.push_back(n_out);
  long off = 0;
  for (int l = 0; l < m->n_layers; ++l) {
    m->w_off.push_back(off);
    off += (long)m->dims[l + 1] * m->dims[l];
    m->b_off.push_back(off);
    off += m->dims[l + 1];
  }
  m->n_params = off;
  m->max_batch = max_batch;
  for (int d : m->dims) m->max_dim = d > m->max_dim ? d : m->max_dim;
  const long act_n = (long)(m->n_layers + 1) * max_batch * m->max_dim;
  const long pre_n = (long)m->n_layers * max_batch * m->max_dim;
  const long dz_n = 2L * max_batch * m->max_dim;
  auto fail = [&]() {
    tdscha_mlp_destroy(reinterpret_cast<tdscha_mlp*>(m));
    return -1;
  };
  if (hipStreamCreate(&m->stream) != hipSuccess ||
      hipMalloc((void**)&m->params, sizeof(float) * off) != hipSuccess ||
      hipMalloc((void**)&m->grads, sizeof(float) * off) != hipSuccess ||
      hipMalloc((void**)&m->m1, sizeof(float) * off) != hipSuccess ||
      hipMalloc((void**)&m->m2, sizeof(float) * off) != hipSuccess ||
      hipMalloc((void**)&m->act, sizeof(float) * act_n) != hipSuccess ||
      hipMalloc((void**)&m->pre, sizeof(float) * pre_n) != hipSuccess ||
      hipMalloc((void**)&m->dz, sizeof(float) * dz_n) != hipSuccess ||
      hipMalloc((void**)&m->target, sizeof(float) * max_batch * n_out) != hipSuccess ||
      hipMalloc((void**)&m->wout, sizeof(float) * n_out) != hipSuccess ||
      hipMalloc((void**)&m->loss_part, sizeof(float) * max_batch) != hipSuccess ||
      hipMalloc((void**)&m->idx, sizeof(int) * max_batch) != hipSuccess) {
    set_error("tdscha_mlp_create: device allocation failed");
    return fail();
  }
  // zero once so padded columns of the activations stay finite
  if (hipMemset(m->act, 0, sizeof(float) * act_n) != hipSuccess ||
      hipMemset(m->pre, 0, sizeof(float) * pre_n) != hipSuccess ||
      hipMemset(m->dz, 0, sizeof(float) * dz_n) != hipSuccess ||
      hipMemset(m->m1, 0, sizeof(float) * off) != hipSuccess ||
      hipMemset(m->m2, 0, sizeof(float) * off) != hipSuccess) {
    set_error("tdscha_mlp_create: hipMemset failed");
    return fail();
  }
  std::vector<float> ones(n_out, 1.0f);
  hipError_t e = hipMemcpy(m->wout, ones.data(), sizeof(float) * n_out, hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = init_params ? hipMemcpy(m->params, init_params, sizeof(float) * off, hipMemcpyHostToDevice)
                    : hipMemset(m->params, 0, sizeof(float) * off);
  if (e != hipSuccess) {
    set_error(std::string("tdscha_mlp_create: ") + hipGetErrorString(e));
    return fail();
  }
  *out = reinterpret_cast<tdscha_mlp*>(m);
  return 0;
}

void tdscha_mlp_destroy(tdscha_mlp* h) {
  if (!h) return;
  Mlp* m = reinterpret_cast<Mlp*>(h);
  if (m->stream) hipStreamSynchronize(m->stream);
  void* ptrs[] = {m->params,  m->grads,  m->m1,      m->m2,       m->act,  m->pre,
                  m->dz,      m->target, m->wout,    m->loss_part, m->idx, m->data_x,
                  m->data_t,  m->data_ch, m->data_ref, m->cf_w,    m->cf_mean, m->cf_std};
  for (void* p : ptrs)
    if (p) hipFree(p);
  if (m->stream) hipStreamDestroy(m->stream);
  delete m;
}

int tdscha_mlp_n_params(const tdscha_mlp* h, long* n) {
  TDS_REQUIRE(h && n, 2, "tdscha_mlp_n_params: NULL argument");
  *n = reinterpret_cast<const Mlp*>(h)->n_params;
  return 0;
}

int tdscha_mlp_get_params(tdscha_mlp* h, float* out) {
  TDS_REQUIRE(h && out, 2, "tdscha_mlp_get_params: NULL argument");
  Mlp* m = reinterpret_cast<Mlp*>(h);
  TDS_CHECK_HIP(hipStreamSynchronize(m->stream));
  TDS_CHECK_HIP(hipMemcpy(out, m->params, sizeof(float) * m->n_params, hipMemcpyDeviceToHost));
  return 0;
}

int tdscha_mlp_set_params(tdscha_mlp* h, const float* in, int reset_optimizer) {
  TDS_REQUIRE(h && in, 2, "tdscha_mlp_set_params: NULL argument");
  Mlp* m = reinterpret_cast<Mlp*>(h);
  TDS_CHECK_HIP(hipStreamSynchronize(m->stream));
  TDS_CHECK_HIP(hipMemcpy(m->params, in, sizeof(float) * m->n_params, hipMemcpyHostToDevice));
  if (reset_optimizer) {
    TDS_CHECK_HIP(hipMemset(m->m1, 0, sizeof(float) * m->n_params));
    TDS_CHECK_HIP(hipMemset(m->m2, 0, sizeof(float) * m->n_params));
    m->step = 0;
  }
  return 0;
}

int tdscha_mlp_set_mse_weights(tdscha_mlp* h, const float* w) {
  TDS_REQUIRE(h && w, 2, "tdscha_mlp_set_mse_weights: NULL argument");
  Mlp* m = reinterpret_cast<Mlp*>(h);
  TDS_CHECK_HIP(hipStreamSynchronize(m->stream));
  TDS_CHECK_HIP(hipMemcpy(m->wout, w, sizeof(float) * m->dims[m->n_layers], hipMemcpyHostToDevice));
  return 0;
}

int tdscha_mlp_set_cf_loss(tdscha_mlp* h, int K, int n_w, const double* w, double smearing,
                           const double* out_mean, const double* out_std, double loss_scale) {
  TDS_REQUIRE(h && w && out_mean && out_std, 2, "tdscha_mlp_set_cf_loss: NULL argument");
  Mlp* m = reinterpret_cast<Mlp*>(h);
  TDS_REQUIRE(m->dims[m->n_layers] == kCfOut, 2,
              "tdscha_mlp_set_cf_loss: the spectral loss needs 10 network outputs");
  TDS_REQUIRE(K > 0 && K <= 4096 && n_w > 0 && n_w <= (1 << 20), 2,
              "tdscha_mlp_set_cf_loss: K or n_w out of range");
  TDS_CHECK_HIP(hipStreamSynchronize(m->stream));
  if (m->cf_w) TDS_CHECK_HIP(hipFree(m->cf_w));
  m->cf_w = nullptr;
  if (!m->cf_mean) TDS_CHECK_HIP(hipMalloc((void**)&m->cf_mean, sizeof(double) * kCfOut));
  if (!m->cf_std) TDS_CHECK_HIP(hipMalloc((void**)&m->cf_std, sizeof(double) * kCfOut));
  TDS_CHECK_HIP(hipMalloc((void**)&m->cf_w, sizeof(double) * n_w));
  TDS_CHECK_HIP(hipMemcpy(m->cf_w, w, sizeof(double) * n_w, hipMemcpyHostToDevice));
  TDS_CHECK_HIP(hipMemcpy(m->cf_mean, out_mean, sizeof(double) * kCfOut, hipMemcpyHostToDevice));
  TDS_CHECK_HIP(hipMemcpy(m->cf_std, out_std, sizeof(double) * kCfOut, hipMemcpyHostToDevice));
  m->cf_K = K;
  m->cf_nw = n_w;
  m->cf_eta = smearing;
  m->cf_scale = loss_scale;
  return 0;
}

int tdscha_mlp_forward(tdscha_mlp* h, int batch, const float* x, float* y) {
  TDS_REQUIRE(h && x && y, 2, "tdscha_mlp_forward: NULL argument");
  TDS_REQUIRE(batch >= 0, 2, "tdscha_mlp_forward: negative batch");
  Mlp* m = reinterpret_cast<Mlp*>(h);
  const int n_in = m->dims[0], n_out = m->dims[m->n_layers];
  for (int b0 = 0; b0 < batch; b0 += m->max_batch) {
    const int B = batch - b0 < m->max_batch ? batch - b0 : m->max_batch;
    TDS_CHECK_HIP(hipMemcpy2DAsync(act_l(m, 0), sizeof(float) * m->max_dim,
                                   x + (long)b0 * n_in, sizeof(float) * n_in,
                                   sizeof(float) * n_in, B, hipMemcpyHostToDevice, m->stream));
    forward(m, B);
    TDS_CHECK_HIP(hipGetLastError());
    TDS_CHECK_HIP(hipMemcpy2DAsync(y + (long)b0 * n_out, sizeof(float) * n_out,
                                   act_l(m, m->n_layers), sizeof(float) * m->max_dim,
                                   sizeof(float) * n_out, B, hipMemcpyDeviceToHost, m->stream));
    TDS_CHECK_HIP(hipStreamSynchronize(m->stream));
  }
  return 0;
}

int tdscha_mlp_loss_grad(tdscha_mlp* h, int kind, int batch, const float* x, const float* target,
                         const double* chains, const double* ref, float* loss, float* grad) {
  TDS_REQUIRE(h && x && loss, 2, "tdscha_mlp_loss_grad: NULL argument");
  Mlp* m = reinterpret_cast<Mlp*>(h);
  TDS_REQUIRE(batch > 0 && batch <= m->max_batch, 2, "tdscha_mlp_loss_grad: batch out of range");
  int rc = check_kind(m, kind, target, chains, ref);
  if (rc) return rc;
  rc = upload_data(m, kind, batch, x, target, chains, ref);
  if (rc) return rc;
  rc = identity_idx(m, batch);
  if (rc) return rc;
  batch_pass(m, kind, batch);
  TDS_CHECK_HIP(hipGetLastError());
  if (grad)
    TDS_CHECK_HIP(hipMemcpyAsync(grad, m->grads, sizeof(float) * m->n_params,
                                 hipMemcpyDeviceToHost, m->stream));
  return read_loss(m, batch, loss);
}

int tdscha_mlp_adam_step(tdscha_mlp* h, int kind, int batch, const float* x, const float* target,
                         const double* chains, const double* ref, float lr, float beta1,
                         float beta2, float eps, float* loss) {
  TDS_REQUIRE(h && x, 2, "tdscha_mlp_adam_step: NULL argument");
  Mlp* m = reinterpret_cast<Mlp*>(h);
  TDS_REQUIRE(batch > 0 && batch <= m->max_batch, 2, "tdscha_mlp_adam_step: batch out of range");
  int rc = check_kind(m, kind, target, chains, ref);
  if (rc) return rc;
  rc = upload_data(m, kind, batch, x, target, chains, ref);
  if (rc) return rc;
  rc = identity_idx(m, batch);
  if (rc) return rc;
  batch_pass(m, kind, batch);
  adam(m, lr, beta1, beta2, eps);
  TDS_CHECK_HIP(hipGetLastError());
  if (loss) return read_loss(m, batch, loss);
  TDS_CHECK_HIP(hipStreamSynchronize(m->stream));
  return 0;
}

int tdscha_mlp_train(tdscha_mlp* h, int kind, long n_samples, const float* x, const float* target,
                     const double* chains, const double* ref, int batch, int epochs, float lr,
                     float beta1, float beta2, float eps, unsigned long long seed,
                     float* epoch_loss) {
  TDS_REQUIRE(h && x, 2, "tdscha_mlp_train: NULL argument");
  Mlp* m = reinterpret_cast<Mlp*>(h);
  TDS_REQUIRE(batch > 0 && batch <= m->max_batch && n_samples >= batch && epochs > 0 &&
                  n_samples <= (1L << 30),
              2, "tdscha_mlp_train: need 0 < batch <= max_batch, batch <= n_samples, epochs > 0");
  int rc = check_kind(m, kind, target, chains, ref);
  if (rc) return rc;
  rc = upload_data(m, kind, n_samples, x, target, chains, ref);
  if (rc) return rc;
  // per-epoch permutation on the host (splitmix64-driven Fisher-Yates); the
  // last n_samples % batch rows of each permutation sit the epoch out
  std::vector<int> perm(n_samples);
  unsigned long long st = seed;
  auto next = [&st]() {
    st += 0x9E3779B97F4A7C15ull;
    unsigned long long z = st;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  };
  const long n_batches = n_samples / batch;
  std::vector<float> lp(batch);
  for (int ep = 0; ep < epochs; ++ep) {
    for (long i = 0; i < n_samples; ++i) perm[i] = (int)i;
    for (long i = n_samples - 1; i > 0; --i) {
      const long j = (long)(next() % (unsigned long long)(i + 1));
      const int tmp = perm[i];
      perm[i] = perm[j];
      perm[j] = tmp;
    }
    double ep_loss = 0.0;
    for (long bi = 0; bi < n_batches; ++bi) {
      TDS_CHECK_HIP(hipMemcpyAsync(m->idx, perm.data() + bi * batch, sizeof(int) * batch,
                                   hipMemcpyHostToDevice, m->stream));
      batch_pass(m, kind, batch);
      adam(m, lr, beta1, beta2, eps);
      // one synchronisation per batch, which also keeps the pageable index
      // upload from racing the next permutation
      TDS_CHECK_HIP(hipMemcpyAsync(lp.data(), m->loss_part, sizeof(float) * batch,
                                   hipMemcpyDeviceToHost, m->stream));
      TDS_CHECK_HIP(hipStreamSynchronize(m->stream));
      for (float v : lp) ep_loss += v;
    }
    TDS_CHECK_HIP(hipGetLastError());
    if (epoch_loss) epoch_loss[ep] = (float)(ep_loss / n_batches);
  }
  TDS_CHECK_HIP(hipStreamSynchronize(m->stream));
  return 0;
}

}  // extern "C"
