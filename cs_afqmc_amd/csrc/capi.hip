// C ABI of the gfx950 AFQMC step library: context lifetime, error reporting and the
// entry points declared in include/cs_afqmc_amd.h.
#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "common.h"

static thread_local char g_err[512] = "";

int afqmc_set_error(int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
  return code;
}

extern "C" const char* afqmc_last_error(void) { return g_err; }

namespace {

template <typename T>
int dalloc(afqmc_ctx* c, T** p, size_t count) {
  *p = nullptr;
  if (count == 0) return AFQMC_OK;
  const size_t bytes = count * sizeof(T);
  HIP_TRY(hipMalloc((void**)p, bytes));
  c->device_bytes += (int64_t)bytes;
  return AFQMC_OK;
}

// true for memory of a device allocation on `device`; host memory, registered or not,
// and other devices' memory are false
bool is_device_ptr(const void* p, int device) {
  hipPointerAttribute_t a;
  if (hipPointerGetAttributes(&a, p) != hipSuccess) {
    (void)hipGetLastError();  // unregistered host memory reports an error: clear it
    return false;
  }
  return a.type == hipMemoryTypeDevice && a.device == device;
}

// stream-ordered on the context's stream; ctx creation synchronises that stream (only)
// before it returns, so the host arrays may be released afterwards
template <typename T>
int upload(afqmc_ctx* c, T** p, const T* host, size_t count) {
  int rc = dalloc(c, p, count);
  if (rc) return rc;
  HIP_TRY(hipMemcpyAsync(*p, host, count * sizeof(T), hipMemcpyDefault, c->stream));
  return AFQMC_OK;
}

int copy_async(void* dst, const void* src, size_t bytes, hipStream_t s) {
  HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDefault, s));
  return AFQMC_OK;
}

int zero_async(void* p, size_t bytes, hipStream_t s) {
  HIP_TRY(hipMemsetAsync(p, 0, bytes, s));
  return AFQMC_OK;
}

int sync_stream(hipStream_t s) {
  HIP_TRY(hipStreamSynchronize(s));
  return AFQMC_OK;
}

void free_ctx(afqmc_ctx* c) {
  if (!c) return;
  void* ptrs[] = {c->trial[0], c->trial[1], c->rot_h1[0], c->rot_h1[1], c->rot_chol[0],
                  c->rot_chol[1], c->chol_t, c->exp_h1[0], c->exp_h1[1], c->mf_shifts,
                  c->emi_bf[0][0], c->emi_bf[0][1], c->emi_bf[1][0], c->emi_bf[1][1], c->chol_t_bf[0],
                  c->chol_t_bf[1], c->rot_chol_bf[0][0], c->rot_chol_bf[0][1],
                  c->rot_chol_bf[1][0], c->rot_chol_bf[1][1], c->green[0], c->green[1], c->fb,
                  c->xs, c->vhs, c->terms, c->dets, c->energy_x, c->energy_acc, c->walk_tmp,
                  c->sr_cum, c->sr_idx, c->xs_bf[0], c->xs_bf[1], c->green_bf[0][0],
                  c->green_bf[0][1], c->green_bf[1][0], c->green_bf[1][1], c->vhs_f32,
                  c->sym_stat};
  prof_free(&c->prof);
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  delete c;
}

int check_w(const afqmc_ctx* c, int W) {
  if (!c) return afqmc_set_error(AFQMC_ERR_ARG, "null context");
  if (W <= 0 || W > c->d.n_walkers)
    return afqmc_set_error(AFQMC_ERR_ARG, "n_walkers=%d outside [1, %d]", W, c->d.n_walkers);
  return AFQMC_OK;
}

int check_walkers(const afqmc_ctx* c, const void* up, const void* dn) {
  if (!up) return afqmc_set_error(AFQMC_ERR_ARG, "walkers_up is NULL");
  if (c->nspin == 2 && !dn)
    return afqmc_set_error(AFQMC_ERR_ARG, "unrestricted walkers need walkers_dn");
  return AFQMC_OK;
}

int set_device(const afqmc_ctx* c) {
  HIP_TRY(hipSetDevice(c->device));
  return AFQMC_OK;
}

// fb[w][c][g] = mult * sum_s green_s[w][c][:] . rot_chol_s[g][:], as c->fb_parts split-K
// slices of (2W, G) that the consumers add in slice order
// arithmetic: FB_BF16 the split-bf16 planes of the mixed mode; otherwise the float64
// Green's function, on float64 MFMA (FB_F64) or float32 MFMA (FB_F32, the FP32 mode)
enum { FB_F64 = 0, FB_BF16 = 1, FB_F32 = 2 };
int step_fb_mode(const afqmc_ctx* c) { return c->mixed ? FB_BF16 : (c->f32 ? FB_F32 : FB_F64); }
int force_bias_gemm(afqmc_ctx* c, int W, int mode) {
  const bool bf16 = mode == FB_BF16;
  const int G = c->d.nchol;
  const int64_t slice = (int64_t)2 * W * G;
  int part = 0;
  for (int s = 0; s < c->nspin; ++s) {
    int rc;
    int sp;
    if (bf16) {
      // split-bf16: rows (w, re/im) of the padded Green planes, K = n_s * Np32
      const int kd = c->nocc[s] * c->Np32;
      sp = gemm_f64_splits(2 * W, G, kd);
      // a 64-wide grid that already fills the chip on its own but a 128-wide one that does
      // not (large W and G): split K instead, so the product runs on 128 x 128 tiles
      const int64_t big = (int64_t)((2 * W + 127) / 128) * ((G + 127) / 128);
      if (sp == 1 && big < 512 && kd >= 4096) sp = (int)std::min<int64_t>(4, (511 + big) / big);
      if (part + sp > c->fb_cap)
        return afqmc_set_error(AFQMC_ERR_STATE, "force-bias split slices exceed workspace");
      rc = launch_gemm_nt_bf16x3_f64(c->stream, 2 * W, G, kd, (float)c->mult, c->green_bf[s][0],
                                     c->green_bf[s][1], kd, c->rot_chol_bf[s][0],
                                     c->rot_chol_bf[s][1], kd, c->fb + part * slice, G, sp, slice);
    } else {
      const int kd = c->kdim[s];
      sp = gemm_f64_splits(2 * W, G, kd);
      if (part + sp > c->fb_cap)
        return afqmc_set_error(AFQMC_ERR_STATE, "force-bias split slices exceed workspace");
      if (mode == FB_F32)
        rc = launch_gemm_nt_f32_split(c->stream, 2 * W, G, kd, (double)c->mult, c->green[s], kd,
                                      c->rot_chol[s], kd, c->fb + part * slice, G, sp, slice);
      else
        rc = launch_gemm_nt_f64_split(c->stream, 2 * W, G, kd, (double)c->mult, c->green[s], kd,
                                      c->rot_chol[s], kd, c->fb + part * slice, G, sp, slice);
    }
    if (rc) return rc;
    part += sp;
  }
  c->fb_parts = part;
  return AFQMC_OK;
}

}  // namespace

extern "C" int afqmc_ctx_create(afqmc_ctx** out, const afqmc_dims* dims,
                                const afqmc_ham_host* ham, void* stream) {
  if (!out || !dims || !ham) return afqmc_set_error(AFQMC_ERR_ARG, "null argument");
  *out = nullptr;
  const afqmc_dims d = *dims;
  if (d.walker_type != AFQMC_RHF && d.walker_type != AFQMC_UHF &&
      d.walker_type != AFQMC_UHF_SPIN_CHOL)
    return afqmc_set_error(AFQMC_ERR_ARG, "walker_type=%d", d.walker_type);
  if (d.norb < 1 || d.norb > AFQMC_MAX_NORB)
    return afqmc_set_error(AFQMC_ERR_SIZE, "norb=%d outside [1, %d]", d.norb, AFQMC_MAX_NORB);
  if (d.nocc_up < 1 || d.nocc_up > AFQMC_MAX_NOCC || d.nocc_up > d.norb)
    return afqmc_set_error(AFQMC_ERR_SIZE, "nocc_up=%d outside [1, min(norb, %d)]", d.nocc_up,
                           AFQMC_MAX_NOCC);
  if (d.walker_type != AFQMC_RHF &&
      (d.nocc_dn < 1 || d.nocc_dn > AFQMC_MAX_NOCC || d.nocc_dn > d.norb))
    return afqmc_set_error(AFQMC_ERR_SIZE, "nocc_dn=%d outside [1, min(norb, %d)]", d.nocc_dn,
                           AFQMC_MAX_NOCC);
  if (d.walker_type == AFQMC_RHF && d.nocc_dn != d.nocc_up)
    return afqmc_set_error(AFQMC_ERR_ARG, "restricted walkers need nocc_dn == nocc_up");
  if (d.nchol < 1) return afqmc_set_error(AFQMC_ERR_ARG, "nchol=%d", d.nchol);
  if (d.n_walkers < 1) return afqmc_set_error(AFQMC_ERR_ARG, "n_walkers=%d", d.n_walkers);
  if (d.n_exp_terms < 1 || d.n_exp_terms > 18)
    return afqmc_set_error(AFQMC_ERR_ARG, "n_exp_terms=%d outside [1, 18]", d.n_exp_terms);
  if (!(d.dt > 0.0)) return afqmc_set_error(AFQMC_ERR_ARG, "dt must be > 0");
  if (d.precision != AFQMC_PREC_FP64 && d.precision != AFQMC_PREC_FP32 &&
      d.precision != AFQMC_PREC_BF16X3)
    return afqmc_set_error(AFQMC_ERR_ARG, "precision mode %d not supported", d.precision);
  if (!ham->chol || !ham->mo_coeff || !ham->rot_h1 || !ham->rot_chol || !ham->mf_shifts ||
      !ham->exp_h1)
    return afqmc_set_error(AFQMC_ERR_ARG, "null intermediate array");

  afqmc_ctx* c = new afqmc_ctx();
  std::memset(c, 0, sizeof(*c));
  c->d = d;
  c->stream = (hipStream_t)stream;
  hipError_t e = hipGetDevice(&c->device);
  if (e != hipSuccess) {
    delete c;
    return afqmc_set_error(AFQMC_ERR_HIP, "hipGetDevice: %s", hipGetErrorString(e));
  }
  c->nspin = d.walker_type == AFQMC_RHF ? 1 : 2;
  c->chol_spins = d.walker_type == AFQMC_UHF_SPIN_CHOL ? 2 : 1;
  c->mult = d.walker_type == AFQMC_RHF ? 2 : 1;
  c->nocc[0] = d.nocc_up;
  c->nocc[1] = d.walker_type == AFQMC_RHF ? d.nocc_up : d.nocc_dn;
  const int N = d.norb, G = d.nchol, W = d.n_walkers;
  c->kdim[0] = c->nocc[0] * N;
  c->kdim[1] = c->nocc[1] * N;
  c->h0 = ham->h0;
  c->h0_prop = ham->h0_prop;

  int rc = AFQMC_OK;
#define TRY(x)          \
  do {                  \
    rc = (x);           \
    if (rc) goto fail;  \
  } while (0)
  {
    size_t off_mo = 0, off_h1 = 0, off_rc = 0;
    for (int s = 0; s < c->nspin; ++s) {
      const int n = c->nocc[s];
      TRY(upload(c, &c->trial[s], ham->mo_coeff + off_mo, (size_t)N * n));
      TRY(upload(c, &c->rot_h1[s], ham->rot_h1 + off_h1, (size_t)n * N));
      TRY(upload(c, &c->rot_chol[s], ham->rot_chol + off_rc, (size_t)G * n * N));
      TRY(upload(c, &c->exp_h1[s], ham->exp_h1 + (size_t)s * N * N, (size_t)N * N));
      off_mo += (size_t)N * n;
      off_h1 += (size_t)n * N;
      off_rc += (size_t)G * n * N;
    }
    // transposed Cholesky vectors (N*N, G): K-contiguous B operand of the VHS GEMM,
    // transposed on the device from the (G, N*N) array (read in place when the caller
    // passes device memory, else from a temporary upload)
    {
      const size_t per = (size_t)G * N * N;
      double* raw = nullptr;
      const double* src = ham->chol;
      if (!is_device_ptr(src, c->device)) {
        rc = upload(c, &raw, ham->chol, c->chol_spins * per);
        src = raw;
      }
      if (!rc) rc = dalloc(c, &c->chol_t, c->chol_spins * per);
      for (int cs = 0; cs < c->chol_spins && !rc; ++cs)
        rc = launch_transpose_f64(c->stream, G, N * N, src + cs * per, c->chol_t + cs * per);
      (void)hipStreamSynchronize(c->stream);
      if (raw) {
        (void)hipFree(raw);
        c->device_bytes -= (int64_t)(c->chol_spins * per * sizeof(double));
      }
      if (rc) goto fail;
    }
    TRY(upload(c, &c->mf_shifts, reinterpret_cast<const cplx*>(ham->mf_shifts), (size_t)G));

    // mixed mode: split-bf16 planes of the contraction operands (zero padded to 32)
    c->mixed = d.precision == AFQMC_PREC_BF16X3;
    c->f32 = d.precision == AFQMC_PREC_FP32;
    c->vhs_nv = N * N;
    c->Gp = (G + 31) / 32 * 32;
    c->Np32 = (N + 31) / 32 * 32;
    c->taylor_bf = c->mixed;  // split-bf16 Taylor kernels (register or streaming form)
    if (c->mixed) {
      const size_t nn = (size_t)N * N;
      // V kept as its upper triangle (half the VHS product, its output and the Taylor
      // kernels' reads) when the Cholesky vectors are symmetric, as they are for real
      // orbitals
      {
        unsigned long long st[2] = {0ull, 0ull};
        TRY(dalloc(c, &c->sym_stat, 2));
        TRY(zero_async(c->sym_stat, 2 * sizeof(unsigned long long), c->stream));
        for (int cs = 0; cs < c->chol_spins; ++cs)  // maxima accumulate over the spins
          TRY(launch_sym_check(c->stream, N, G, chol_t_of(c, cs), c->sym_stat));
        TRY(copy_async(st, c->sym_stat, sizeof(st), c->stream));
        TRY(sync_stream(c->stream));
        double asym, amax;
        std::memcpy(&asym, &st[0], sizeof(double));
        std::memcpy(&amax, &st[1], sizeof(double));
        if (asym <= 1e-12 * amax) c->vhs_nv = N * (N + 1) / 2;
      }
      for (int h = 0; h < 2; ++h)
        TRY(dalloc(c, &c->chol_t_bf[h], (size_t)c->chol_spins * c->vhs_nv * c->Gp));
      for (int cs = 0; cs < c->chol_spins; ++cs) {
        if (c->vhs_nv != N * N)
          TRY(launch_split_bf16_sym(c->stream, N, G, chol_t_of(c, cs), chol_t_bf_of(c, 0, cs),
                                    chol_t_bf_of(c, 1, cs), c->Gp));
        else
          TRY(launch_split_bf16(c->stream, (int)nn, G, chol_t_of(c, cs), G,
                                chol_t_bf_of(c, 0, cs), chol_t_bf_of(c, 1, cs), c->Gp));
      }
      for (int s = 0; s < c->nspin; ++s) {
        const size_t rows = (size_t)G * c->nocc[s];
        for (int h = 0; h < 2; ++h) {
          TRY(dalloc(c, &c->rot_chol_bf[s][h], rows * c->Np32));
          TRY(dalloc(c, &c->green_bf[s][h], (size_t)W * 2 * c->nocc[s] * c->Np32));
          TRY(zero_async(c->green_bf[s][h],
                         (size_t)W * 2 * c->nocc[s] * c->Np32 * sizeof(uint16_t), c->stream));
        }
        TRY(launch_split_bf16(c->stream, (int)rows, N, c->rot_chol[s], N, c->rot_chol_bf[s][0],
                              c->rot_chol_bf[s][1], c->Np32));
      }
      for (int h = 0; h < 2; ++h) {
        TRY(dalloc(c, &c->xs_bf[h], (size_t)W * 2 * c->Gp));
        TRY(zero_async(c->xs_bf[h], (size_t)W * 2 * c->Gp * sizeof(uint16_t), c->stream));
      }
      if (c->taylor_bf) {
        // 8 floats of tail padding: the streaming split-bf16 Taylor kernel reads V rows in
        // 8-float chunks, and the chunk that straddles the end of the last row is masked;
        // at least N^2 doubles, which stage exp_h1 - I below
        c->vhs_f32_stride = (int64_t)W * 2 * c->vhs_nv + 8;
        TRY(dalloc(c, &c->vhs_f32,
                   std::max((size_t)c->chol_spins * c->vhs_f32_stride, 2 * nn)));
        for (int s = 0; s < c->nspin; ++s) {
          // exp_h1 - I as split-bf16 planes, rows padded to Np32 (bf16 Taylor kernel).  The
          // float64 matrix is staged in the VHS workspace (8 W N^2 bytes, unused until the
          // first step): a temporary allocation's hipFree would wait for the whole device.
          std::vector<double> emi(nn);
          for (size_t e = 0; e < nn; ++e)
            emi[e] = ham->exp_h1[(size_t)s * nn + e] - ((e / N == e % N) ? 1.0 : 0.0);
          double* emi_dev = reinterpret_cast<double*>(c->vhs_f32);
          TRY(copy_async(emi_dev, emi.data(), nn * sizeof(double), c->stream));
          for (int h = 0; h < 2; ++h) TRY(dalloc(c, &c->emi_bf[s][h], (size_t)N * c->Np32));
          TRY(launch_split_bf16(c->stream, N, N, emi_dev, N, c->emi_bf[s][0], c->emi_bf[s][1],
                                c->Np32));
          TRY(sync_stream(c->stream));  // the host copy `emi` is released next iteration
        }
      }
      TRY(sync_stream(c->stream));
    }

    // workspace
    for (int s = 0; s < c->nspin; ++s) TRY(dalloc(c, &c->green[s], (size_t)W * 2 * c->kdim[s]));
    c->fb_cap = 0;
    for (int s = 0; s < c->nspin; ++s) {
      const int kd = c->mixed ? c->nocc[s] * c->Np32 : c->kdim[s];
      int cap = kd / 256;
      c->fb_cap += cap < 1 ? 1 : (cap > 8 ? 8 : cap);
    }
    TRY(dalloc(c, &c->fb, (size_t)c->fb_cap * W * 2 * G));
    TRY(dalloc(c, &c->xs, (size_t)W * 2 * G));
    if (!c->taylor_bf) TRY(dalloc(c, &c->vhs, (size_t)c->chol_spins * W * 2 * N * N));
    TRY(dalloc(c, &c->terms, (size_t)W * 2));
    TRY(dalloc(c, &c->dets, (size_t)W * 2));
    {
      // exchange partial sums of the local energy: one complex per (g-tile, walker)
      size_t tiles = 0;
      for (int s = 0; s < c->nspin; ++s) tiles += (size_t)energy_g_tiles(c, s);
      TRY(dalloc(c, &c->energy_x, tiles * W));
    }
    TRY(dalloc(c, &c->walk_tmp, (size_t)W * N * (c->nocc[0] > c->nocc[1] ? c->nocc[0] : c->nocc[1])));
    TRY(dalloc(c, &c->sr_cum, (size_t)W));
    TRY(dalloc(c, &c->sr_idx, (size_t)W));
    TRY(sync_stream(c->stream));
  }
#undef TRY
  *out = c;
  return AFQMC_OK;
fail:
  (void)hipStreamSynchronize(c->stream);
  free_ctx(c);
  return rc;
}

extern "C" int afqmc_ctx_destroy(afqmc_ctx* ctx) {
  if (!ctx) return AFQMC_OK;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  free_ctx(ctx);
  return AFQMC_OK;
}

extern "C" int afqmc_set_stream(afqmc_ctx* ctx, void* stream) {
  if (!ctx) return afqmc_set_error(AFQMC_ERR_ARG, "null context");
  ctx->stream = (hipStream_t)stream;
  return AFQMC_OK;
}

extern "C" int64_t afqmc_ctx_device_bytes(const afqmc_ctx* ctx) {
  return ctx ? ctx->device_bytes : 0;
}

extern "C" int64_t afqmc_ctx_vhs_entries(const afqmc_ctx* ctx) {
  return ctx ? ctx->vhs_nv : 0;
}

extern "C" int afqmc_propagate(afqmc_ctx* c, int32_t W, void* walkers_up, void* walkers_dn,
                               double* weights, void* overlaps, const double* fields,
                               double* scalars, void* force_bias_out, int32_t flags) {
  int rc = check_w(c, W);
  if (rc) return rc;
  if ((rc = check_walkers(c, walkers_up, walkers_dn))) return rc;
  if (!weights || !overlaps || !fields || !scalars)
    return afqmc_set_error(AFQMC_ERR_ARG, "null weights/overlaps/fields/scalars");
  if ((rc = set_device(c))) return rc;
  cplx* wu = static_cast<cplx*>(walkers_up);
  cplx* wd = static_cast<cplx*>(walkers_dn);
  const int N = c->d.norb;
  // force bias: Green of the current walkers (or the one the previous step left),
  // then the rot_chol contraction
  const bool reuse = (flags & AFQMC_STEP_REUSE_GREEN) && c->green_src == walkers_up &&
                     c->green_w == W;
  c->green_src = nullptr;
  if (!reuse) {
    PhaseTimer t(c, AFQMC_PH_GREEN);
    if ((rc = launch_measure(c, W, wu, wd, !c->mixed, c->mixed))) return rc;
  }
  {
    PhaseTimer t(c, AFQMC_PH_FB_GEMM);
    if ((rc = force_bias_gemm(c, W, step_fb_mode(c)))) return rc;
  }
  {
    PhaseTimer t(c, AFQMC_PH_FIELDS);
    if ((rc = launch_fields(c, W, fields, static_cast<cplx*>(force_bias_out)))) return rc;
  }
  {
    // VHS[w] = sum_g (i sqrt(dt) x~_wg) L_g  : (2W, G) x (vhs_nv, G)^T, all N*N entries
    // or the upper triangle of the (then symmetric) V
    PhaseTimer t(c, AFQMC_PH_VHS_GEMM);
    for (int cs = 0; cs < c->chol_spins && !rc; ++cs) {
      if (c->mixed)
        rc = launch_gemm_nt_bf16x3_f32(c->stream, 2 * W, c->vhs_nv, c->Gp, 1.0f, c->xs_bf[0],
                                       c->xs_bf[1], c->Gp, chol_t_bf_of(c, 0, cs),
                                       chol_t_bf_of(c, 1, cs), c->Gp, vhs_f32_of(c, cs),
                                       c->vhs_nv, 1, 0);
      else if (c->f32)
        rc = launch_gemm_nt_f32(c->stream, 2 * W, N * N, c->d.nchol, 1.0, c->xs, c->d.nchol,
                                chol_t_of(c, cs), c->d.nchol, 0.0, vhs_of(c, cs), N * N);
      else
        rc = launch_gemm_nt_f64(c->stream, 2 * W, N * N, c->d.nchol, 1.0, c->xs, c->d.nchol,
                                chol_t_of(c, cs), c->d.nchol, 0.0, vhs_of(c, cs), N * N);
    }
    if (rc) return rc;
  }
  {
    PhaseTimer t(c, AFQMC_PH_TAYLOR);
    if ((rc = launch_taylor(c, W, wu, wd))) return rc;
  }
  {
    // new overlaps and each walker's weight update; the Green's function of the new
    // walkers is kept for the next step
    PhaseTimer t(c, AFQMC_PH_OVERLAP);
    const WeightsArgs tail = weights_args(c, W, weights, static_cast<cplx*>(overlaps), scalars);
    if ((rc = launch_measure(c, W, wu, wd, !c->mixed, c->mixed, &tail))) return rc;
    c->green_src = walkers_up;
    c->green_w = W;
  }
  // population-control shift from the summed weights
  PhaseTimer t(c, AFQMC_PH_WEIGHTS);
  return launch_pcs(c, W, weights, scalars);
}

extern "C" int afqmc_block_stats(afqmc_ctx* c, int32_t W, const void* energies,
                                 const double* weights, double* scalars, double* out) {
  int rc = check_w(c, W);
  if (rc) return rc;
  if (!energies || !weights || !scalars || !out)
    return afqmc_set_error(AFQMC_ERR_ARG, "null energies/weights/scalars/out");
  if ((rc = set_device(c))) return rc;
  return launch_block_stats(c, W, static_cast<const cplx*>(energies), weights, scalars, out);
}

extern "C" int afqmc_cs_outliers(afqmc_ctx* c, int32_t W, double* weights1, double* weights2,
                                 double z_thresh) {
  int rc = check_w(c, W);
  if (rc) return rc;
  if (!weights1 || !weights2) return afqmc_set_error(AFQMC_ERR_ARG, "null weights");
  if ((rc = set_device(c))) return rc;
  return launch_cs_outliers(c, W, weights1, weights2, z_thresh);
}

int prof_enable(ProfRec* p, int on) {
  if (on && !p->ev) {
    const int cap = 1 << 14;
    p->ev = new hipEvent_t[2 * cap];
    p->phase = new int[cap];
    for (int i = 0; i < 2 * cap; ++i) HIP_TRY(hipEventCreate(&p->ev[i]));
    p->cap = cap;
    p->n = 0;
  }
  p->on = on ? 1 : 0;
  return AFQMC_OK;
}

int prof_collect(ProfRec* p, double* ms, int64_t* counts) {
  for (int i = 0; i < p->n; ++i) {
    HIP_TRY(hipEventSynchronize(p->ev[2 * i + 1]));
    float t = 0.f;
    HIP_TRY(hipEventElapsedTime(&t, p->ev[2 * i], p->ev[2 * i + 1]));
    ms[p->phase[i]] += (double)t;
    counts[p->phase[i]] += 1;
  }
  p->n = 0;
  return AFQMC_OK;
}

void prof_free(ProfRec* p) {
  if (!p->ev) return;
  for (int i = 0; i < 2 * p->cap; ++i) (void)hipEventDestroy(p->ev[i]);
  delete[] p->ev;
  delete[] p->phase;
  p->ev = nullptr;
  p->phase = nullptr;
  p->cap = p->n = p->on = 0;
}

extern "C" int afqmc_prof_enable(afqmc_ctx* c, int32_t on) {
  if (!c) return afqmc_set_error(AFQMC_ERR_ARG, "null context");
  return prof_enable(&c->prof, on);
}

extern "C" int afqmc_prof_phases(afqmc_ctx* c, uint32_t mask) {
  if (!c) return afqmc_set_error(AFQMC_ERR_ARG, "null context");
  c->prof.mask = mask;
  return AFQMC_OK;
}

extern "C" int afqmc_prof_collect(afqmc_ctx* c, double* ms, int64_t* counts) {
  if (!c || !ms || !counts) return afqmc_set_error(AFQMC_ERR_ARG, "null argument");
  return prof_collect(&c->prof, ms, counts);
}

extern "C" int afqmc_overlap(afqmc_ctx* c, int32_t W, const void* walkers_up,
                             const void* walkers_dn, void* overlaps) {
  int rc = check_w(c, W);
  if (rc) return rc;
  if ((rc = check_walkers(c, walkers_up, walkers_dn))) return rc;
  if (!overlaps) return afqmc_set_error(AFQMC_ERR_ARG, "null overlaps");
  if ((rc = set_device(c))) return rc;
  if ((rc = launch_measure(c, W, static_cast<const cplx*>(walkers_up),
                           static_cast<const cplx*>(walkers_dn), false)))
    return rc;
  return launch_overlap_combine(c, W, static_cast<cplx*>(overlaps));
}

extern "C" int afqmc_force_bias(afqmc_ctx* c, int32_t W, const void* walkers_up,
                                const void* walkers_dn, void* force_bias) {
  int rc = check_w(c, W);
  if (rc) return rc;
  if ((rc = check_walkers(c, walkers_up, walkers_dn))) return rc;
  if (!force_bias) return afqmc_set_error(AFQMC_ERR_ARG, "null force_bias");
  if ((rc = set_device(c))) return rc;
  c->green_src = nullptr;
  // the same contraction the step uses: float64, or split-bf16 in mixed mode
  if ((rc = launch_measure(c, W, static_cast<const cplx*>(walkers_up),
                           static_cast<const cplx*>(walkers_dn), !c->mixed, c->mixed)))
    return rc;
  if ((rc = force_bias_gemm(c, W, step_fb_mode(c)))) return rc;
  return launch_fb_out(c, W, static_cast<cplx*>(force_bias));
}

extern "C" int afqmc_energy(afqmc_ctx* c, int32_t W, const void* walkers_up,
                            const void* walkers_dn, void* energies) {
  int rc = check_w(c, W);
  if (rc) return rc;
  if ((rc = check_walkers(c, walkers_up, walkers_dn))) return rc;
  if (!energies) return afqmc_set_error(AFQMC_ERR_ARG, "null energies");
  if ((rc = set_device(c))) return rc;
  c->green_src = nullptr;
  // Green's function in float64 (one-body term, c_g), plus split-bf16 planes for the
  // exchange contraction of the mixed mode
  if ((rc = launch_measure(c, W, static_cast<const cplx*>(walkers_up),
                           static_cast<const cplx*>(walkers_dn), true, c->mixed)))
    return rc;
  PhaseTimer t(c, AFQMC_PH_ENERGY);
  // c_g = force bias, always contracted in float64
  if ((rc = force_bias_gemm(c, W, FB_F64))) return rc;
  return launch_energy(c, W, static_cast<cplx*>(energies));
}

extern "C" int afqmc_orthonormalize(afqmc_ctx* c, int32_t W, void* walkers_up,
                                    void* walkers_dn) {
  int rc = check_w(c, W);
  if (rc) return rc;
  if ((rc = check_walkers(c, walkers_up, walkers_dn))) return rc;
  if ((rc = set_device(c))) return rc;
  c->green_src = nullptr;
  PhaseTimer t(c, AFQMC_PH_QR);
  return launch_qr(c, W, static_cast<cplx*>(walkers_up), static_cast<cplx*>(walkers_dn));
}

extern "C" int afqmc_sr(afqmc_ctx* c, int32_t W, void* walkers_up, void* walkers_dn,
                        double* weights, double zeta, int32_t* idx_out, int32_t flags) {
  int rc = check_w(c, W);
  if (rc) return rc;
  if ((rc = check_walkers(c, walkers_up, walkers_dn))) return rc;
  if (!weights) return afqmc_set_error(AFQMC_ERR_ARG, "null weights");
  if (!(zeta >= 0.0 && zeta < 1.0))
    return afqmc_set_error(AFQMC_ERR_ARG, "zeta must lie in [0, 1)");
  if ((rc = set_device(c))) return rc;
  c->green_src = nullptr;
  PhaseTimer t(c, AFQMC_PH_SR);
  if (flags & ~AFQMC_SR_GLOBAL_ORDER) return afqmc_set_error(AFQMC_ERR_ARG, "unknown sr flags");
  return launch_sr(c, W, static_cast<cplx*>(walkers_up), static_cast<cplx*>(walkers_dn), weights,
                   zeta, idx_out, (flags & AFQMC_SR_GLOBAL_ORDER) != 0);
}

extern "C" int afqmc_sr_key(afqmc_ctx* c, int32_t W, void* walkers_up, void* walkers_dn,
                            double* weights, uint64_t* key, int32_t* idx_out, int32_t flags) {
  int rc = check_w(c, W);
  if (rc) return rc;
  if ((rc = check_walkers(c, walkers_up, walkers_dn))) return rc;
  if (!weights || !key) return afqmc_set_error(AFQMC_ERR_ARG, "null weights or key");
  if (flags & ~AFQMC_SR_GLOBAL_ORDER) return afqmc_set_error(AFQMC_ERR_ARG, "unknown sr flags");
  if ((rc = set_device(c))) return rc;
  c->green_src = nullptr;
  PhaseTimer t(c, AFQMC_PH_SR);
  return launch_sr(c, W, static_cast<cplx*>(walkers_up), static_cast<cplx*>(walkers_dn), weights,
                   0.0, idx_out, (flags & AFQMC_SR_GLOBAL_ORDER) != 0, key);
}

extern "C" int afqmc_estimate_mix(afqmc_ctx* c, double* scalars, const double* stats,
                                  int32_t nblocks) {
  if (!c) return afqmc_set_error(AFQMC_ERR_ARG, "null context");
  if (!scalars || !stats || nblocks < 1)
    return afqmc_set_error(AFQMC_ERR_ARG, "null scalars/stats or nblocks < 1");
  int rc = set_device(c);
  if (rc) return rc;
  return launch_estimate_mix(c, nblocks, AFQMC_BLOCK_STATS_LEN, stats, scalars);
}

extern "C" int afqmc_fields_normal_key(afqmc_ctx* c, uint64_t* key, double* fields,
                                       int64_t count) {
  if (!c) return afqmc_set_error(AFQMC_ERR_ARG, "null context");
  if (!key || !fields || count < 0) return afqmc_set_error(AFQMC_ERR_ARG, "bad key or fields");
  if (count == 0) return AFQMC_OK;
  int rc = set_device(c);
  if (rc) return rc;
  PhaseTimer t(c, AFQMC_PH_RNG);
  return launch_fields_normal_key(c, key, fields, count);
}

extern "C" int afqmc_fields_normal(afqmc_ctx* c, uint64_t seed, uint64_t counter,
                                   double* fields, int64_t count) {
  if (!c) return afqmc_set_error(AFQMC_ERR_ARG, "null context");
  if (!fields || count < 0) return afqmc_set_error(AFQMC_ERR_ARG, "bad fields buffer");
  if (count == 0) return AFQMC_OK;
  int rc = set_device(c);
  if (rc) return rc;
  PhaseTimer t(c, AFQMC_PH_RNG);
  return launch_fields_normal(c, seed, counter, fields, count);
}
