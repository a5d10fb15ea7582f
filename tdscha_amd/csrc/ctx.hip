// Context lifetime, uploads and the perturbed-average C-ABI entry points.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>

#include "common.hpp"
#include "../../include/tdscha_amd.h"

namespace tdscha {

static thread_local std::string g_err;
void set_error(const std::string& msg) { g_err = msg; }
const char* last_error() { return g_err.c_str(); }

__global__ void expand_images(const double*, const double*, const double*, const int*, const int*,
                              const int*, const int*, const long*, const int*, const double*, int,
                              int, int, long, double*, double*, double*);
int launch_pa_fp64(Ctx* c, int nc, int apply_d4, hipStream_t s, hipEvent_t* ev);
int launch_pa_split(Ctx* c, int nc, int apply_d4, int prec, hipStream_t s, hipEvent_t* ev);
int set_ksplit_split(Ctx* c, int ks);
int default_ksplit_split(Ctx* c);
int rot_setup(Ctx* c, const double* X_dev, const double* Y_dev, const double* rho_dev);
bool rot_panels_fit(int dmax);
int rot_panels(const int* bs, const int* d, int n, int np, int* start, int* nreal);
int run_pa_rot(Ctx* c, int nc, int apply_d4, int prec, double* f_dev, double* d2v_dev,
               hipStream_t s);

// pad R1 (n) -> (np) and Y1 (n,n) -> symmetric (np,np); zero outside
__global__ void pad_inputs(const double* __restrict__ R1, const double* __restrict__ Y1, int n,
                           int np, double* __restrict__ R1p, double* __restrict__ Y1p) {
  const int chain = blockIdx.y;
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long)np * np) return;
  const int a = (int)(e / np), b = (int)(e % np);
  double v = 0.0;
  if (a < n && b < n) {
    const double* y = Y1 + (long)chain * n * n;
    v = 0.5 * (y[(long)a * n + b] + y[(long)b * n + a]);
  }
  Y1p[(long)chain * np * np + e] = v;
  if (e < np) R1p[(long)chain * np + e] = (e < n) ? R1[(long)chain * n + e] : 0.0;
}

__global__ void symmetrize_out(const double* __restrict__ zsum_all, const double* __restrict__ fsum,
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

static double bose(double w, double T) { return T > 0 ? 1.0 / (std::exp(w * RY_TO_K / T) - 1.0) : 0.0; }

static int choose_ksplit(int ntile, long Mp) {
  // aim at ~6 workgroups per CU for the outer GEMM, multiple of 8 (one per XCD)
  const int target = 1536;
  int ks = (target + ntile * ntile - 1) / (ntile * ntile);
  ks = round_up(ks, 8);
  ks = std::min(ks, 128);
  // at least 256 images per split
  while (ks > 8 && (long)ks * 256 > Mp) ks -= 8;
  return std::max(ks, 8);
}

// run the perturbed averages for nc chains whose padded inputs are in c->R1/c->Y1;
// outputs: unpadded f/d2v at the given device pointers
int run_pa(Ctx* c, int nc, int apply_d4, int prec, double* f_dev, double* d2v_dev, hipStream_t s) {
  if (c->rotate) return run_pa_rot(c, nc, apply_d4, prec, f_dev, d2v_dev, s);
  int rc;
  if (prec == PREC_FP64) {
    rc = launch_pa_fp64(c, nc, apply_d4, s, nullptr);
  } else {
    rc = launch_pa_split(c, nc, apply_d4, prec, s, nullptr);
  }
  if (rc) return rc;
  const double norm = 1.0 / (c->n_eff * c->n_syms);
  dim3 g((unsigned)(((long)c->n * c->n + 255) / 256), nc);
  hipLaunchKernelGGL(symmetrize_out, g, dim3(256), 0, s, c->d2v_out, c->f_out, c->fy, c->n, c->np,
                     norm, f_dev, d2v_dev);
  TDS_CHECK_HIP(hipGetLastError());
  return 0;
}

}  // namespace tdscha

using namespace tdscha;

extern "C" {

const char* tdscha_last_error(void) { return tdscha::last_error(); }

int tdscha_ctx_create(tdscha_ctx** out, int n_modes, int n_configs, const double* X,
                      const double* Y, const double* w, const double* rho, double T, int n_syms,
                      int n_blocks, const int* block_dim, const int* block_modes,
                      const double* const* block_syms, const int* sym_block_id, int flags) {
  TDS_REQUIRE(out != nullptr, 2, "tdscha_ctx_create: out is NULL");
  TDS_REQUIRE((flags & ~TDSCHA_CTX_ROTATE_OPERATORS) == 0, 2, "tdscha_ctx_create: unknown flags");
  const bool rotate = (flags & TDSCHA_CTX_ROTATE_OPERATORS) != 0;
  *out = nullptr;
  TDS_REQUIRE(n_modes > 0 && n_configs > 0 && n_syms > 0 && n_blocks > 0, 2,
              "tdscha_ctx_create: sizes must be positive");
  TDS_REQUIRE(X && Y && w && rho && block_dim && block_modes && block_syms && sym_block_id, 2,
              "tdscha_ctx_create: NULL array argument");
  TDS_REQUIRE(n_modes <= 4096, 2, "tdscha_ctx_create: n_modes above 4096 is not supported");
  // validate the block description on the host
  std::vector<int> pos(n_modes, -1);
  long tot = 0;
  for (int b = 0; b < n_blocks; ++b) {
    TDS_REQUIRE(block_dim[b] > 0 && block_dim[b] <= n_modes, 2, "tdscha_ctx_create: bad block_dim");
    TDS_REQUIRE(block_syms[b] != nullptr, 2, "tdscha_ctx_create: NULL symmetry block");
    tot += block_dim[b];
  }
  TDS_REQUIRE(tot == n_modes, 2, "tdscha_ctx_create: blocks must partition the modes");
  {
    long off = 0;
    for (int b = 0; b < n_blocks; ++b) {
      for (int k = 0; k < block_dim[b]; ++k) {
        const int m = block_modes[off + k];
        TDS_REQUIRE(m >= 0 && m < n_modes && pos[m] < 0, 2,
                    "tdscha_ctx_create: block_modes is not a permutation of the modes");
        TDS_REQUIRE(sym_block_id[m] == b, 2, "tdscha_ctx_create: sym_block_id disagrees with blocks");
        pos[m] = k;
      }
      off += block_dim[b];
    }
  }
  for (int m = 0; m < n_modes; ++m) TDS_REQUIRE(w[m] > 0, 2, "tdscha_ctx_create: frequencies must be positive");

  Ctx* c = new Ctx();
  c->n = n_modes;
  c->np = round_up(n_modes, kTile);
  c->N = n_configs;
  // the rotation form runs the kernels on the configurations alone
  c->rotate = rotate;
  c->rot_nsym = n_syms;
  c->n_syms = rotate ? 1 : n_syms;
  c->M = (long)n_configs * c->n_syms;
  c->Mp = round_up_l(c->M, 128);
  c->row_lo = 0;
  c->row_hi = c->M;
  c->T = T;
  double ne = 0;
  for (int i = 0; i < n_configs; ++i) ne += rho[i];
  c->n_eff = ne;
  TDS_REQUIRE(ne != 0.0, 2, "tdscha_ctx_create: sum of weights is zero");
  const int np = c->np;
  const long Mp = c->Mp;
  c->ksplit = choose_ksplit(np / kTile, Mp);

  auto fail = [&](void) { tdscha_ctx_destroy(reinterpret_cast<tdscha_ctx*>(c)); return -1; };
#define ALLOC(ptr, bytes)                                                 \
  if (hipMalloc((void**)&(ptr), (bytes)) != hipSuccess) {                 \
    set_error(std::string("hipMalloc failed for ") + #ptr);               \
    return fail();                                                        \
  }
  if (hipGetDevice(&c->device) != hipSuccess) { set_error("no HIP device"); delete c; return -1; }
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    set_error("hipStreamCreate failed"); delete c; return -1;
  }
  ALLOC(c->U, sizeof(double) * Mp * np);
  ALLOC(c->F, sizeof(double) * Mp * np);
  ALLOC(c->rho_img, sizeof(double) * Mp);
  ALLOC(c->fy, sizeof(double) * np);
  ALLOC(c->fpsi, sizeof(double) * np);
  ALLOC(c->w, sizeof(double) * np);
  ALLOC(c->R1, sizeof(double) * kMaxChains * np);
  ALLOC(c->Y1, sizeof(double) * kMaxChains * (long)np * np);
  ALLOC(c->coef, sizeof(double) * kMaxChains * Mp * 4);
  ALLOC(c->zpart, sizeof(double) * kMaxChains * (long)c->ksplit * np * np);
  ALLOC(c->fpart, sizeof(double) * kMaxChains * (long)c->ksplit * np);
  ALLOC(c->f_out, sizeof(double) * kMaxChains * np);
  ALLOC(c->d2v_out, sizeof(double) * kMaxChains * (long)np * np);
  ALLOC(c->stage_in, sizeof(double) * kMaxChains * ((long)n_modes + (long)n_modes * n_modes));
  ALLOC(c->stage_out, sizeof(double) * kMaxChains * ((long)n_modes + (long)n_modes * n_modes));

  // thermal factors (padded modes get zero)
  std::vector<double> hfy(np, 0.0), hfp(np, 0.0), hw(np, 0.0);
  for (int m = 0; m < n_modes; ++m) {
    const double nb = bose(w[m], T);
    hfy[m] = 2.0 * w[m] / (1.0 + 2.0 * nb);
    hfp[m] = (1.0 + 2.0 * nb) / (2.0 * w[m]);
    hw[m] = w[m];
  }
  if (hipMemcpy(c->fy, hfy.data(), sizeof(double) * np, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(c->fpsi, hfp.data(), sizeof(double) * np, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(c->w, hw.data(), sizeof(double) * np, hipMemcpyHostToDevice) != hipSuccess) {
    set_error("upload of thermal factors failed");
    return fail();
  }

  // symmetry tables
  std::vector<int> mode_block(n_modes), mode_pos(n_modes), bmoff(n_blocks);
  std::vector<long> bsoff(n_blocks);
  long soff = 0;
  int moff = 0;
  for (int b = 0; b < n_blocks; ++b) {
    bmoff[b] = moff;
    bsoff[b] = soff;
    moff += block_dim[b];
    soff += (long)n_syms * block_dim[b] * block_dim[b];
  }
  for (int m = 0; m < n_modes; ++m) {
    mode_block[m] = sym_block_id[m];
    mode_pos[m] = pos[m];
  }
  std::vector<double> hsyms(soff);
  for (int b = 0; b < n_blocks; ++b)
    std::memcpy(hsyms.data() + bsoff[b], block_syms[b],
                sizeof(double) * (size_t)n_syms * block_dim[b] * block_dim[b]);
  c->block_dim.assign(block_dim, block_dim + n_blocks);
  c->block_modes.assign(block_modes, block_modes + n_modes);
  c->sym_block_id.assign(sym_block_id, sym_block_id + n_modes);

  int *d_mb = nullptr, *d_mp = nullptr, *d_bd = nullptr, *d_bmo = nullptr, *d_bm = nullptr;
  long* d_bso = nullptr;
  double *d_syms = nullptr, *d_X = nullptr, *d_Y = nullptr, *d_rho = nullptr;
  bool ok = hipMalloc((void**)&d_mb, sizeof(int) * n_modes) == hipSuccess &&
            hipMalloc((void**)&d_mp, sizeof(int) * n_modes) == hipSuccess &&
            hipMalloc((void**)&d_bd, sizeof(int) * n_blocks) == hipSuccess &&
            hipMalloc((void**)&d_bmo, sizeof(int) * n_blocks) == hipSuccess &&
            hipMalloc((void**)&d_bm, sizeof(int) * n_modes) == hipSuccess &&
            hipMalloc((void**)&d_bso, sizeof(long) * n_blocks) == hipSuccess &&
            hipMalloc((void**)&d_syms, sizeof(double) * soff) == hipSuccess &&
            hipMalloc((void**)&d_X, sizeof(double) * (size_t)n_configs * n_modes) == hipSuccess &&
            hipMalloc((void**)&d_Y, sizeof(double) * (size_t)n_configs * n_modes) == hipSuccess &&
            hipMalloc((void**)&d_rho, sizeof(double) * n_configs) == hipSuccess;
  ok = ok && hipMemcpy(d_mb, mode_block.data(), sizeof(int) * n_modes, hipMemcpyHostToDevice) == hipSuccess;
  ok = ok && hipMemcpy(d_mp, mode_pos.data(), sizeof(int) * n_modes, hipMemcpyHostToDevice) == hipSuccess;
  ok = ok && hipMemcpy(d_bd, block_dim, sizeof(int) * n_blocks, hipMemcpyHostToDevice) == hipSuccess;
  ok = ok && hipMemcpy(d_bmo, bmoff.data(), sizeof(int) * n_blocks, hipMemcpyHostToDevice) == hipSuccess;
  ok = ok && hipMemcpy(d_bm, block_modes, sizeof(int) * n_modes, hipMemcpyHostToDevice) == hipSuccess;
  ok = ok && hipMemcpy(d_bso, bsoff.data(), sizeof(long) * n_blocks, hipMemcpyHostToDevice) == hipSuccess;
  ok = ok && hipMemcpy(d_syms, hsyms.data(), sizeof(double) * soff, hipMemcpyHostToDevice) == hipSuccess;
  ok = ok && hipMemcpy(d_X, X, sizeof(double) * (size_t)n_configs * n_modes, hipMemcpyHostToDevice) == hipSuccess;
  ok = ok && hipMemcpy(d_Y, Y, sizeof(double) * (size_t)n_configs * n_modes, hipMemcpyHostToDevice) == hipSuccess;
  ok = ok && hipMemcpy(d_rho, rho, sizeof(double) * n_configs, hipMemcpyHostToDevice) == hipSuccess;
  if (ok && rotate) {
    // the configurations only, and per mode its block's modes and the row
    // and column of every S_j restricted to them
    int dmax = 0;
    for (int b = 0; b < n_blocks; ++b) dmax = std::max(dmax, block_dim[b]);
    std::vector<int> cols((size_t)n_modes * dmax, 0), dd(n_modes, 0), bs(n_modes, 0);
    // blocks of consecutive modes in ascending order take the panel kernels
    bool consecutive = true;
    for (int b = 0; b < n_blocks; ++b) {
      const int* bm = block_modes + bmoff[b];
      for (int a = 0; a < block_dim[b]; ++a) {
        bs[bm[a]] = bm[0];
        dd[bm[a]] = block_dim[b];
        consecutive = consecutive && bm[a] == bm[0] + a;
      }
    }
    c->rt_panels = consecutive && rot_panels_fit(dmax);
    std::vector<int> pstart(n_modes + 2, 0);
    if (c->rt_panels)
      c->rt_npanel = rot_panels(bs.data(), dd.data(), n_modes, np, pstart.data(), &c->rt_nreal);
    std::vector<double> sr((size_t)n_syms * n_modes * dmax, 0.0), sc(sr.size(), 0.0);
    for (int b = 0; b < n_blocks; ++b) {
      const int d = block_dim[b];
      const int* bm = block_modes + bmoff[b];
      for (int a = 0; a < d; ++a) {
        const int m = bm[a];
        for (int t = 0; t < d; ++t) cols[(size_t)t * n_modes + m] = bm[t];
        for (int j = 0; j < n_syms; ++j)
          for (int t = 0; t < d; ++t) {
            sr[((size_t)j * dmax + t) * n_modes + m] = block_syms[b][((size_t)j * d + a) * d + t];
            sc[((size_t)j * dmax + t) * n_modes + m] = block_syms[b][((size_t)j * d + t) * d + a];
          }
      }
    }
    c->rt_dmax = dmax;
    ok = hipMalloc((void**)&c->rt_cols, sizeof(int) * cols.size()) == hipSuccess &&
         hipMalloc((void**)&c->rt_d, sizeof(int) * n_modes) == hipSuccess &&
         hipMalloc((void**)&c->rt_bs, sizeof(int) * n_modes) == hipSuccess &&
         hipMalloc((void**)&c->rt_pstart, sizeof(int) * (n_modes + 2)) == hipSuccess &&
         hipMalloc((void**)&c->rt_sr, sizeof(double) * sr.size()) == hipSuccess &&
         hipMalloc((void**)&c->rt_sc, sizeof(double) * sc.size()) == hipSuccess &&
         hipMalloc((void**)&c->R1_base, sizeof(double) * kMaxChains * np) == hipSuccess &&
         hipMalloc((void**)&c->Y1_base, sizeof(double) * kMaxChains * (long)np * np) == hipSuccess &&
         hipMalloc((void**)&c->rot_t1, sizeof(double) * kMaxChains * (long)np * np) == hipSuccess &&
         hipMalloc((void**)&c->rot_t2, sizeof(double) * kMaxChains * (long)np * np) == hipSuccess &&
         hipMalloc((void**)&c->rot_f, sizeof(double) * kMaxChains * np) == hipSuccess;
    ok = ok && hipMemcpy(c->rt_cols, cols.data(), sizeof(int) * cols.size(), hipMemcpyHostToDevice) == hipSuccess;
    ok = ok && hipMemcpy(c->rt_d, dd.data(), sizeof(int) * n_modes, hipMemcpyHostToDevice) == hipSuccess;
    ok = ok && hipMemcpy(c->rt_bs, bs.data(), sizeof(int) * n_modes, hipMemcpyHostToDevice) == hipSuccess;
    ok = ok && hipMemcpy(c->rt_pstart, pstart.data(), sizeof(int) * (n_modes + 2), hipMemcpyHostToDevice) == hipSuccess;
    ok = ok && hipMemcpy(c->rt_sr, sr.data(), sizeof(double) * sr.size(), hipMemcpyHostToDevice) == hipSuccess;
    ok = ok && hipMemcpy(c->rt_sc, sc.data(), sizeof(double) * sc.size(), hipMemcpyHostToDevice) == hipSuccess;
    ok = ok && rot_setup(c, d_X, d_Y, d_rho) == 0 && hipStreamSynchronize(c->stream) == hipSuccess;
  } else if (ok) {
    const int tpb = 128;
    dim3 g((unsigned)Mp, (np + tpb - 1) / tpb);
    hipLaunchKernelGGL(expand_images, g, dim3(tpb), 0, c->stream, d_X, d_Y, d_rho, d_mb, d_mp, d_bd,
                       d_bmo, d_bso, d_bm, d_syms, n_modes, np, n_syms, c->M, c->U, c->F,
                       c->rho_img);
    ok = hipGetLastError() == hipSuccess && hipStreamSynchronize(c->stream) == hipSuccess;
  }
  hipFree(d_mb); hipFree(d_mp); hipFree(d_bd); hipFree(d_bmo); hipFree(d_bm); hipFree(d_bso);
  hipFree(d_syms); hipFree(d_X); hipFree(d_Y); hipFree(d_rho);
  if (!ok) {
    set_error("tdscha_ctx_create: upload / image expansion failed");
    return fail();
  }
#undef ALLOC
  *out = reinterpret_cast<tdscha_ctx*>(c);
  return 0;
}

void tdscha_ctx_destroy(tdscha_ctx* ctx) {
  if (!ctx) return;
  Ctx* c = reinterpret_cast<Ctx*>(ctx);
  if (c->stream) hipStreamSynchronize(c->stream);
  void* ptrs[] = {c->U, c->F, c->rho_img, c->fy, c->fpsi, c->w, c->R1, c->Y1, c->coef, c->zpart,
                  c->fpart, c->f_out, c->d2v_out, c->Y1h, c->Y1l, c->ysc, c->ysc3, c->bsc, c->cmax,
                  c->pair_ca, c->pair_cb, c->pair_sm, c->pair_sp, c->pair_harm_a, c->pair_harm_b,
                  c->pair_row, c->pair_col, c->std_coef, c->inv_coef, c->yw, c->reaw, c->pols, c->inv_perms, c->rep_dof, c->jrep, c->ju, c->rep_sum,
                  c->tproj, c->cart_tmp, c->cart_tmp2, c->mode_tmp, c->psi_dev, c->psi_out_dev,
                  c->stage_in, c->stage_out, c->f_fin, c->d2v_fin, c->f_proj, c->epi,
                  c->coef32, c->wpart, c->zpart32, c->fpart_s, c->sbsc, c->pmax,
                  c->rt_cols, c->rt_d, c->rt_bs, c->rt_pstart, c->rt_sr, c->rt_sc, c->R1_base, c->Y1_base, c->rot_t1,
                  c->rot_t2, c->rot_f};
  for (void* p : ptrs)
    if (p) hipFree(p);
  for (auto& s : c->split) {
    void* sp[] = {s.Uh, s.Ul, s.FA, s.UA32, s.UhT, s.UlT, s.XT, s.YT, s.su, s.sf, s.umax, s.fmax, s.colw};
    for (void* p : sp)
      if (p) hipFree(p);
  }
  if (c->stream) hipStreamDestroy(c->stream);
  delete c;
}

int tdscha_ctx_info(const tdscha_ctx* ctx, int* n_modes, int* n_modes_padded, long* n_images,
                    long* n_images_padded, int* ksplit) {
  TDS_REQUIRE(ctx != nullptr, 2, "tdscha_ctx_info: NULL context");
  const Ctx* c = reinterpret_cast<const Ctx*>(ctx);
  if (n_modes) *n_modes = c->n;
  if (n_modes_padded) *n_modes_padded = c->np;
  // images of the symmetric ensemble (in the rotation form: configurations
  // times symmetries, none of them stored) and the stored rows
  if (n_images) *n_images = c->M * (c->rotate ? c->rot_nsym : 1);
  if (n_images_padded) *n_images_padded = c->Mp;
  if (ksplit) *ksplit = c->ksplit;
  return 0;
}

int tdscha_perturb_averages_dev(tdscha_ctx* ctx, int n_chains, const double* R1_dev,
                                const double* Y1_dev, int apply_d4, int precision, double* f_dev,
                                double* d2v_dev, void* stream) {
  TDS_REQUIRE(ctx != nullptr, 2, "tdscha_perturb_averages_dev: NULL context");
  TDS_REQUIRE(n_chains >= 1 && n_chains <= kMaxChains, 2,
              "tdscha_perturb_averages_dev: n_chains must be in [1, 8]");
  TDS_REQUIRE(apply_d4 == 0 || apply_d4 == 1, 2, "the apply_d4 value passed must be either 0 or 1");
  TDS_REQUIRE(precision >= 0 && precision <= 2, 2, "precision must be 0 (FP64), 1 (FP16x2) or 2 (BF16x2)");
  TDS_REQUIRE(R1_dev && Y1_dev && f_dev && d2v_dev, 2, "tdscha_perturb_averages_dev: NULL pointer");
  Ctx* c = reinterpret_cast<Ctx*>(ctx);
  hipStream_t s = stream ? reinterpret_cast<hipStream_t>(stream) : c->stream;
  const long nn = (long)c->np * c->np;
  dim3 g((unsigned)((nn + 255) / 256), n_chains);
  hipLaunchKernelGGL(pad_inputs, g, dim3(256), 0, s, R1_dev, Y1_dev, c->n, c->np, c->R1, c->Y1);
  TDS_CHECK_HIP(hipGetLastError());
  return run_pa(c, n_chains, apply_d4, precision, f_dev, d2v_dev, s);
}

int tdscha_perturb_averages_range(tdscha_ctx* ctx, int n_chains, const double* R1_dev,
                                  const double* Y1_dev, int apply_d4, int precision,
                                  long start_index, long end_index, double* f_dev,
                                  double* d2v_dev, void* stream) {
  TDS_REQUIRE(ctx != nullptr, 2, "tdscha_perturb_averages_range: NULL context");
  Ctx* c = reinterpret_cast<Ctx*>(ctx);
  // the reference's assertions (Modules/tdscha_core.jl:178-180)
  TDS_REQUIRE(start_index <= end_index, 2,
              "tdscha_perturb_averages_range: start_index must not exceed end_index");
  TDS_REQUIRE(start_index >= 1, 2, "tdscha_perturb_averages_range: start_index must be >= 1");
  const long ns = c->rotate ? c->rot_nsym : 1;
  TDS_REQUIRE(end_index <= c->M * ns, 2,
              "tdscha_perturb_averages_range: end_index cannot exceed n_configs * n_syms");
  // the rotation form sums all symmetries of a configuration in one pass
  TDS_REQUIRE(!c->rotate || ((start_index - 1) % ns == 0 && end_index % ns == 0), 2,
              "tdscha_perturb_averages_range: with rotated operators the range must hold whole "
              "configurations (start_index - 1 and end_index multiples of n_syms)");
  // kernel arguments are taken at launch, so the range applies to this call only
  c->row_lo = (start_index - 1) / ns;
  c->row_hi = end_index / ns;
  const int rc = tdscha_perturb_averages_dev(ctx, n_chains, R1_dev, Y1_dev, apply_d4, precision,
                                             f_dev, d2v_dev, stream);
  c->row_lo = 0;
  c->row_hi = c->M;
  return rc;
}

int tdscha_perturb_averages_batched(tdscha_ctx* ctx, int n_chains, const double* R1,
                                    const double* Y1, int apply_d4, int precision, double* f_out,
                                    double* d2v_out) {
  TDS_REQUIRE(ctx != nullptr, 2, "tdscha_perturb_averages_batched: NULL context");
  TDS_REQUIRE(R1 && Y1 && f_out && d2v_out, 2, "tdscha_perturb_averages_batched: NULL pointer");
  TDS_REQUIRE(n_chains >= 1 && n_chains <= kMaxChains, 2, "n_chains must be in [1, 8]");
  Ctx* c = reinterpret_cast<Ctx*>(ctx);
  const size_t n = c->n, nn = n * n;
  double* dR = c->stage_in;
  double* dY = c->stage_in + (size_t)n_chains * n;
  double* df = c->stage_out;
  double* dd = c->stage_out + (size_t)n_chains * n;
  TDS_CHECK_HIP(hipMemcpyAsync(dR, R1, sizeof(double) * n_chains * n, hipMemcpyHostToDevice, c->stream));
  TDS_CHECK_HIP(hipMemcpyAsync(dY, Y1, sizeof(double) * n_chains * nn, hipMemcpyHostToDevice, c->stream));
  int rc = tdscha_perturb_averages_dev(ctx, n_chains, dR, dY, apply_d4, precision, df, dd, c->stream);
  if (rc) return rc;
  TDS_CHECK_HIP(hipMemcpyAsync(f_out, df, sizeof(double) * n_chains * n, hipMemcpyDeviceToHost, c->stream));
  TDS_CHECK_HIP(hipMemcpyAsync(d2v_out, dd, sizeof(double) * n_chains * nn, hipMemcpyDeviceToHost, c->stream));
  TDS_CHECK_HIP(hipStreamSynchronize(c->stream));
  return 0;
}

int tdscha_perturb_averages(tdscha_ctx* ctx, const double* R1, const double* Y1, int apply_d4,
                            double* f_out, double* d2v_out) {
  return tdscha_perturb_averages_batched(ctx, 1, R1, Y1, apply_d4, PREC_FP64, f_out, d2v_out);
}

// ------------------------------------------------------------ device buffers
int tdscha_device_alloc(void** out, size_t bytes) {
  TDS_REQUIRE(out != nullptr, 2, "tdscha_device_alloc: NULL out");
  *out = nullptr;
  TDS_CHECK_HIP(hipMalloc(out, bytes ? bytes : 1));
  return 0;
}

int tdscha_device_free(void* ptr) {
  if (ptr) TDS_CHECK_HIP(hipFree(ptr));
  return 0;
}

int tdscha_memcpy(void* dst, const void* src, size_t bytes, int kind) {
  TDS_REQUIRE(kind >= 0 && kind <= 3, 2, "tdscha_memcpy: kind must be 0..3 (H2H, H2D, D2H, D2D)");
  TDS_CHECK_HIP(hipMemcpy(dst, src, bytes, static_cast<hipMemcpyKind>(kind)));
  return 0;
}

int tdscha_device_count(int* count) {
  TDS_REQUIRE(count != nullptr, 2, "tdscha_device_count: NULL out");
  *count = 0;
  if (hipGetDeviceCount(count) != hipSuccess) *count = 0;
  return 0;
}

__global__ void fill_pattern(double* R1, double* Y1, int n, int np, int nc) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long nn = (long)np * np;
  if (e >= nn * nc) return;
  const int chain = (int)(e / nn);
  const long q = e % nn;
  const int a = (int)(q / np), b = (int)(q % np);
  Y1[e] = (a < n && b < n) ? 1e-3 * cos(0.37 * (a + b) + 0.11 * a * b + chain) : 0.0;
  if (q < np) R1[(long)chain * np + q] = (q < n) ? sin(0.5 * q + chain) : 0.0;
}

int tdscha_profile_perturb_averages(tdscha_ctx* ctx, int n_chains, int precision, int reps,
                                    double* ms_out) {
  TDS_REQUIRE(ctx != nullptr && ms_out != nullptr, 2, "tdscha_profile_perturb_averages: NULL");
  TDS_REQUIRE(n_chains >= 1 && n_chains <= kMaxChains && reps >= 1, 2, "bad n_chains / reps");
  TDS_REQUIRE(precision >= 0 && precision <= 2, 2, "precision must be 0, 1 or 2");
  Ctx* c = reinterpret_cast<Ctx*>(ctx);
  hipStream_t s = c->stream;
  if (precision != PREC_FP64) {
    int rc = ensure_split(c, precision);
    if (rc) return rc;
  }
  const long tot = (long)c->np * c->np * n_chains;
  hipLaunchKernelGGL(fill_pattern, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, c->R1,
                     c->Y1, c->n, c->np, n_chains);
  hipEvent_t ev[4];
  for (int i = 0; i < 4; ++i) TDS_CHECK_HIP(hipEventCreate(&ev[i]));
  double acc[4] = {0, 0, 0, 0};
  for (int r = 0; r < reps + 1; ++r) {
    int rc = precision == PREC_FP64 ? launch_pa_fp64(c, n_chains, 1, s, ev)
                                    : launch_pa_split(c, n_chains, 1, precision, s, ev);
    if (rc) return rc;
    TDS_CHECK_HIP(hipStreamSynchronize(s));
    if (r == 0) continue;  // first pass is a warm-up
    float t01, t12, t23;
    (void)hipEventElapsedTime(&t01, ev[0], ev[1]);
    (void)hipEventElapsedTime(&t12, ev[1], ev[2]);
    (void)hipEventElapsedTime(&t23, ev[2], ev[3]);
    acc[0] += t01; acc[1] += t12; acc[2] += t23; acc[3] += t01 + t12 + t23;
  }
  for (int i = 0; i < 4; ++i) {
    ms_out[i] = acc[i] / reps;
    (void)hipEventDestroy(ev[i]);
  }
  return 0;
}

int tdscha_set_ksplit(tdscha_ctx* ctx, int ksplit) {
  TDS_REQUIRE(ctx != nullptr, 2, "tdscha_set_ksplit: NULL context");
  TDS_REQUIRE(ksplit >= 8 && ksplit <= 4096 && ksplit % 8 == 0, 2,
              "tdscha_set_ksplit: ksplit must be a multiple of 8 in [8, 4096]");
  Ctx* c = reinterpret_cast<Ctx*>(ctx);
  TDS_CHECK_HIP(hipStreamSynchronize(c->stream));
  double *zp = nullptr, *fp = nullptr;
  TDS_CHECK_HIP(hipMalloc((void**)&zp, sizeof(double) * kMaxChains * (size_t)ksplit * c->np * c->np));
  TDS_CHECK_HIP(hipMalloc((void**)&fp, sizeof(double) * kMaxChains * (size_t)ksplit * c->np));
  (void)hipFree(c->zpart);
  (void)hipFree(c->fpart);
  c->zpart = zp;
  c->fpart = fp;
  c->ksplit = ksplit;
  return 0;
}

int tdscha_autotune(tdscha_ctx* ctx, int n_chains, int precision, int reps, int* ksplit_out,
                    double* best_ms) {
  TDS_REQUIRE(ctx != nullptr, 2, "tdscha_autotune: NULL context");
  TDS_REQUIRE(n_chains >= 1 && n_chains <= kMaxChains && reps >= 1, 2,
              "tdscha_autotune: n_chains must be in [1, 8] and reps >= 1");
  TDS_REQUIRE(precision >= 0 && precision <= 2, 2, "precision must be 0, 1 or 2");
  Ctx* c = reinterpret_cast<Ctx*>(ctx);
  if (precision != PREC_FP64) {
    const int rc = ensure_split(c, precision);
    if (rc) return rc;
  }
  // candidates: the default and its neighbours (multiples of 8), each timed
  // on reps passes of this batch in two sweeps (the clock of a busy board
  // drifts); the smallest time of a candidate counts, the fastest is kept
  const int ks0 = precision == PREC_FP64 ? c->ksplit : default_ksplit_split(c);
  std::vector<int> cand;
  for (int f : {1, 2, 3, 4, 6}) {
    for (int ks : {ks0 * f / 4, ks0 * f / 2}) {
      ks = std::max(8, ks / 8 * 8);
      if (std::find(cand.begin(), cand.end(), ks) == cand.end()) cand.push_back(ks);
    }
  }
  std::sort(cand.begin(), cand.end());
  std::vector<double> tmin(cand.size(), 1e300);
  for (int sweep = 0; sweep < 2; ++sweep) {
    for (size_t i = 0; i < cand.size(); ++i) {
      const int ks = cand[i];
      int rc;
      if (precision == PREC_FP64) {
        if ((long)ks * 256 > c->Mp) continue;
        rc = tdscha_set_ksplit(ctx, ks);
      } else {
        if ((long)ks * 4 > c->Mp / 16) continue;
        rc = set_ksplit_split(c, ks);
      }
      if (rc) return rc;
      double ms[4];
      rc = tdscha_profile_perturb_averages(ctx, n_chains, precision, reps, ms);
      if (rc) return rc;
      tmin[i] = std::min(tmin[i], ms[3]);
    }
  }
  double best = 1e300;
  int best_ks = ks0;
  for (size_t i = 0; i < cand.size(); ++i)
    if (tmin[i] < best) {
      best = tmin[i];
      best_ks = cand[i];
    }
  const int rc = precision == PREC_FP64 ? tdscha_set_ksplit(ctx, best_ks)
                                        : set_ksplit_split(c, best_ks);
  if (rc) return rc;
  if (ksplit_out) *ksplit_out = best_ks;
  if (best_ms) *best_ms = best;
  return 0;
}

int tdscha_synchronize(tdscha_ctx* ctx) {
  TDS_REQUIRE(ctx != nullptr, 2, "tdscha_synchronize: NULL context");
  TDS_CHECK_HIP(hipStreamSynchronize(reinterpret_cast<Ctx*>(ctx)->stream));
  return 0;
}

}  // extern "C"
