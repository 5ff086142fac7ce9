// Shared device helpers and the context layout of the gfx950 AFQMC step library.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <mutex>

#include "../../include/cs_afqmc_amd.h"

typedef double double4_t __attribute__((ext_vector_type(4)));
typedef float float4_t __attribute__((ext_vector_type(4)));

#define AFQMC_MAX_NORB 256
#define AFQMC_MAX_NOCC 64

// ---------------------------------------------------------------- complex float64
struct cplx {
  double re, im;
};

__host__ __device__ inline cplx cmake(double r, double i) { return cplx{r, i}; }
__host__ __device__ inline cplx cadd(cplx a, cplx b) { return cplx{a.re + b.re, a.im + b.im}; }
__host__ __device__ inline cplx csub(cplx a, cplx b) { return cplx{a.re - b.re, a.im - b.im}; }
__host__ __device__ inline cplx cmul(cplx a, cplx b) {
  return cplx{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
__host__ __device__ inline cplx cscale(cplx a, double s) { return cplx{a.re * s, a.im * s}; }
__host__ __device__ inline cplx cconj(cplx a) { return cplx{a.re, -a.im}; }
// a * conj(b)
__host__ __device__ inline cplx cmulc(cplx a, cplx b) {
  return cplx{a.re * b.re + a.im * b.im, a.im * b.re - a.re * b.im};
}
// Smith's algorithm for a / b (robust against overflow of |b|^2)
__host__ __device__ inline cplx cdiv(cplx a, cplx b) {
  if (fabs(b.re) >= fabs(b.im)) {
    double r = b.im / b.re, d = b.re + b.im * r;
    return cplx{(a.re + a.im * r) / d, (a.im - a.re * r) / d};
  } else {
    double r = b.re / b.im, d = b.re * r + b.im;
    return cplx{(a.re * r + a.im) / d, (a.im * r - a.re) / d};
  }
}
__host__ __device__ inline double cabs1(cplx a) { return fabs(a.re) + fabs(a.im); }

// ---------------------------------------------------------------- bf16 split
// x ~= hi + lo with hi = bf16(x), lo = bf16(x - hi), both rounded to nearest even
// (gfx950's v_cvt_pk_bf16_f32 rounds to nearest even)
typedef float f32x2_t __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2_t __attribute__((ext_vector_type(2)));
__device__ inline uint16_t f32_to_bf16_rne(float f) {
  return __builtin_bit_cast(uint16_t, (__bf16)f);
}
// two floats -> one dword of bf16 (a in the low half)
__device__ inline uint32_t pack_bf16x2(float a, float b) {
  return __builtin_bit_cast(uint32_t, __builtin_convertvector(f32x2_t{a, b}, bf16x2_t));
}
// the floats a packed dword stands for
__device__ inline float bf16_lo_val(uint32_t u) { return __uint_as_float(u << 16); }
__device__ inline float bf16_hi_val(uint32_t u) { return __uint_as_float(u & 0xFFFF0000u); }
// x ~= hi + lo with hi = bf16(float(x)) and lo = bf16(float(x) - hi) (the difference is
// exact in float32).  The empty asm keeps the double -> float rounding a step of its own:
// folded into one double -> bf16 conversion it becomes a round-to-odd sequence of about
// ten instructions.
__device__ inline uint32_t split_bf16_pair(double x) {
  float f = (float)x;
  __asm__("" : "+v"(f));
  const uint16_t h = f32_to_bf16_rne(f);
  const uint16_t l = f32_to_bf16_rne(f - __uint_as_float((uint32_t)h << 16));
  return (uint32_t)h | ((uint32_t)l << 16);
}
__device__ inline void split_bf16(double x, uint16_t* hi, uint16_t* lo) {
  const uint32_t hl = split_bf16_pair(x);
  *hi = (uint16_t)hl;
  *lo = (uint16_t)(hl >> 16);
}

// ---------------------------------------------------------------- walker weights
struct WeightsArgs {
  int W, nspin;
  double dt, sqdt, h0_prop;
  const cplx* dets;
  const cplx* terms;
  double* weights;
  cplx* overlaps;
  double* scalars;
};

__device__ inline cplx cexp_(cplx a) {
  const double e = exp(a.re);
  double s, c;
  sincos(a.im, &s, &c);
  return cplx{e * c, e * s};
}

// Importance function, phaseless projection, clipping and the weight cap of walker w
// (propagation.py:133-160) given its new overlap and the population-control shift pcs;
// writes weights[w] and overlaps[w] and returns the new weight.
__device__ inline double walker_weight_update(const WeightsArgs& a, double pcs, int w,
                                              cplx onew) {
  const cplx oold = a.overlaps[w];
  const cplx st = a.terms[2 * w], ft = a.terms[2 * w + 1];
  const cplx e = cplx{-a.sqdt * st.re + ft.re + a.dt * (pcs + a.h0_prop), -a.sqdt * st.im + ft.im};
  const cplx imp = cdiv(cmul(cexp_(e), onew), oold);
  const cplx rb = cdiv(cmul(cexp_(cplx{-a.sqdt * st.re, -a.sqdt * st.im}), onew), oold);
  const double theta = atan2(rb.im, rb.re);
  double ipl = hypot(imp.re, imp.im) * cos(theta);
  if (isnan(ipl)) ipl = 0.0;
  if (ipl < 1.0e-3) ipl = 0.0;
  if (ipl > 100.0) ipl = 0.0;
  double wt = ipl * a.weights[w];
  if (wt > 100.0) wt = 0.0;
  a.weights[w] = wt;
  a.overlaps[w] = onew;
  return wt;
}

// ---------------------------------------------------------------- GEMM tile order
// Block -> (tile_m, tile_n) for a tiles_m x tiles_n grid.  Blocks are dealt round-robin
// to the 8 XCDs (block b runs on XCD b % 8), each with its own L2, so every XCD gets a
// contiguous run of tiles.  Within a run the tile index walks the dimension whose
// operand panels are shared: rows fastest when B (the N side) is the larger operand, so
// an XCD reads each B panel once; columns fastest when A is larger (m_major = true).
// The walked operand is then re-read for every panel of the other one; when it does not
// fit an XCD's 4 MB L2, the run goes in bands of up to 4 panels of the shared operand
// (as many as fill half the L2), walked together, so the walked operand streams once
// per band instead of once per panel.
constexpr int64_t XCD_L2_BYTES = 4 << 20;
__device__ inline int xcd_band(int64_t walked_bytes, int64_t shared_panel_bytes) {
  if (walked_bytes <= XCD_L2_BYTES) return 1;
  const int64_t fit = (XCD_L2_BYTES / 2) / (shared_panel_bytes > 0 ? shared_panel_bytes : 1);
  return fit < 1 ? 1 : (fit > 4 ? 4 : (int)fit);
}
__device__ inline void xcd_tile(int bx, int tiles_m, int tiles_n, bool m_major, int band,
                                int* tm, int* tn) {
  const int nblk = tiles_m * tiles_n;
  const int q = nblk >> 3, r = nblk & 7, x = bx & 7, i = bx >> 3;
  const int b = x * q + (x < r ? x : r) + i;
  if (m_major) {  // columns fastest, in bands of `band` row tiles
    const int bi = b / (band * tiles_n), m0 = bi * band;
    const int gm = tiles_m - m0 < band ? tiles_m - m0 : band;
    const int w = b - bi * band * tiles_n;
    *tn = w / gm;
    *tm = m0 + w - (w / gm) * gm;
  } else {  // rows fastest, in bands of `band` column tiles
    const int bi = b / (band * tiles_m), n0 = bi * band;
    const int gn = tiles_n - n0 < band ? tiles_n - n0 : band;
    const int w = b - bi * band * tiles_m;
    *tm = w / gn;
    *tn = n0 + w - (w / gn) * gn;
  }
}

// ---------------------------------------------------------------- reductions
template <int NT>
__device__ inline double block_sum(double v, double* red) {
  // deterministic tree: wave shuffle, then one wave over the per-wave partials
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red[wid] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x < 64) {
    s = (lane < NT / 64) ? red[lane] : 0.0;
    for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
    if (lane == 0) red[0] = s;
  }
  __syncthreads();
  s = red[0];
  __syncthreads();
  return s;
}

// ---------------------------------------------------------------- phase timing
// Event pairs around the phases of a step (afqmc_prof_*, afqmc_group_prof_*): while
// `on`, each phase whose bit is set in `mask` (0: all) records a pair on its stream.
struct ProfRec {
  int on;
  uint32_t mask;
  int n;               // recorded event pairs
  int cap;
  hipEvent_t* ev;      // 2 * cap events
  int* phase;          // phase of each pair
};
int prof_enable(ProfRec* p, int on);
int prof_collect(ProfRec* p, double* ms, int64_t* counts);
void prof_free(ProfRec* p);

// ---------------------------------------------------------------- context
struct afqmc_ctx {
  int device;
  hipStream_t stream;
  afqmc_dims d;
  int nspin;          // 1 (restricted) or 2
  int chol_spins;     // 1, or 2 for separate alpha / beta Cholesky vectors
  int nocc[2];
  int mult;           // 2 for restricted (both spins share the walker), 1 otherwise
  int kdim[2];        // n_sigma * N
  double h0, h0_prop;

  // intermediates (device, float64)
  double* trial[2];      // (N, n_sigma)
  double* rot_h1[2];     // (n_sigma, N)
  double* rot_chol[2];   // (G, n_sigma, N)
  double* chol_t;        // (N*N, G): transposed Cholesky vectors
  double* exp_h1[2];     // (N, N)
  cplx* mf_shifts;       // (G)

  // reduced-precision copies (mixed mode, AFQMC_PREC_BF16X3)
  int mixed;                 // 1 when the contractions run split-bf16
  int f32;                   // 1 when the contractions run on float32 MFMA (AFQMC_PREC_FP32)
  int taylor_bf;             // 1 when the Taylor kernels run split-bf16 (mixed mode)
  int Gp;                    // nchol rounded up to 32 (bf16 plane leading dimension)
  int Np32;                  // norb rounded up to 32
  int vhs_nv;                // stored entries of one V plane: N*N, or N(N+1)/2 when V is
                             // kept as its upper triangle (symmetric Cholesky vectors)
  uint16_t* chol_t_bf[2];    // hi / lo planes of chol_t, (vhs_nv, Gp)
  uint16_t* rot_chol_bf[2][2];  // [spin][hi/lo], (G, n_sigma, Np32)
  uint16_t* xs_bf[2];        // [hi/lo] (2W, Gp): i sqrt(dt) x~
  uint16_t* green_bf[2][2];  // [spin][hi/lo] (W, 2, n_sigma, Np32)
  float* vhs_f32;            // chol_spins x [(W, 2, vhs_nv) + 8 floats of padding]
  int64_t vhs_f32_stride;    // floats between the spins' V blocks
  unsigned long long* sym_stat;  // (2) Cholesky symmetry check at creation
  uint16_t* emi_bf[2][2];    // [spin][hi/lo] (N, Np32): split-bf16 exp_h1 - I

  // workspace (device)
  double* green[2];      // (W, 2, n_sigma*N): split re / im planes per walker
  double* fb;            // fb_cap slices of (W, 2, G) split-K partial force biases
  int fb_cap;            // slices allocated
  int fb_parts;          // slices written by the last force-bias contraction
  double* xs;            // (W, 2, G): i sqrt(dt) * shifted fields, split
  double* vhs;           // (W, 2, N*N)
  cplx* terms;           // (W, 2): shift_term, fb_term
  cplx* dets;            // (W, 2): det(C_sigma^H w_sigma)
  cplx* energy_x;        // (sum_s g_tiles_s, W) exchange partial sums
  cplx* energy_acc;      // (W): partial energies
  cplx* walk_tmp;        // gather target for SR (W * N * n_max)
  double* sr_cum;        // (W) cumulative |w|
  int32_t* sr_idx;       // (W)
  int64_t device_bytes;
  // the Green's function in `green` belongs to these walkers (set by afqmc_propagate)
  const void* green_src;
  int green_w;

  // phase timing (afqmc_prof_*)
  ProfRec prof;
};

// Spin s's transposed Cholesky vectors, their split-bf16 planes and its V buffers: one
// set shared by both spins unless the vectors are spin resolved.
inline int chol_spin(const afqmc_ctx* c, int s) { return c->chol_spins == 2 ? s : 0; }
inline double* chol_t_of(const afqmc_ctx* c, int s) {
  return c->chol_t + (size_t)chol_spin(c, s) * c->d.norb * c->d.norb * c->d.nchol;
}
inline uint16_t* chol_t_bf_of(const afqmc_ctx* c, int h, int s) {
  return c->chol_t_bf[h] + (size_t)chol_spin(c, s) * c->vhs_nv * c->Gp;
}
inline double* vhs_of(const afqmc_ctx* c, int s) {
  return c->vhs + (size_t)chol_spin(c, s) * c->d.n_walkers * 2 * c->d.norb * c->d.norb;
}
inline float* vhs_f32_of(const afqmc_ctx* c, int s) {
  return c->vhs_f32 + (size_t)chol_spin(c, s) * c->vhs_f32_stride;
}

// phase timer: records an event pair around a phase when profiling is on
struct PhaseTimer {
  ProfRec* p;
  hipStream_t s;
  int slot;
  PhaseTimer(ProfRec* rec, hipStream_t st, int phase) : p(rec), s(st), slot(-1) {
    if (p->on && p->n < p->cap && (p->mask == 0 || ((p->mask >> phase) & 1u))) {
      slot = p->n++;
      p->phase[slot] = phase;
      (void)hipEventRecord(p->ev[2 * slot], s);
    }
  }
  PhaseTimer(afqmc_ctx* c, int phase) : PhaseTimer(&c->prof, c->stream, phase) {}
  ~PhaseTimer() {
    if (slot >= 0) (void)hipEventRecord(p->ev[2 * slot + 1], s);
  }
};

// status helpers (capi.hip)
int afqmc_set_error(int code, const char* fmt, ...);

#define HIP_TRY(expr)                                                              \
  do {                                                                             \
    hipError_t _e = (expr);                                                        \
    if (_e != hipSuccess)                                                          \
      return afqmc_set_error(AFQMC_ERR_HIP, "%s failed: %s (%s:%d)", #expr,        \
                             hipGetErrorString(_e), __FILE__, __LINE__);           \
  } while (0)

#define LAUNCH_CHECK()                                                             \
  do {                                                                             \
    hipError_t _e = hipGetLastError();                                             \
    if (_e != hipSuccess)                                                          \
      return afqmc_set_error(AFQMC_ERR_HIP, "kernel launch failed: %s (%s:%d)",   \
                             hipGetErrorString(_e), __FILE__, __LINE__);           \
  } while (0)

// Raise a kernel's dynamic-LDS limit to `bytes` when it exceeds the 64 KiB default and
// the limit already set (`current`, one static per kernel).  The limit only grows, under
// a lock, so threads launching the same kernel with different sizes never lower it.
inline std::mutex& dyn_lds_mutex() {
  static std::mutex m;
  return m;
}
inline int ensure_dyn_lds(const void* fn, size_t bytes, int* current) {
  if (bytes <= 64 * 1024) return AFQMC_OK;
  std::lock_guard<std::mutex> lock(dyn_lds_mutex());
  if ((int)bytes <= *current) return AFQMC_OK;
  HIP_TRY(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
  *current = (int)bytes;
  return AFQMC_OK;
}

// ---------------------------------------------------------------- kernel launchers
// gemm_f64.hip: C[m][n] = alpha * sum_k A[m][k] * B[n][k] (+ beta * C), row-major
int launch_gemm_nt_f64(hipStream_t s, int M, int N, int K, double alpha, const double* A,
                       int lda, const double* B, int ldb, double beta, double* C, int ldc);
// split-K form: slice z of C (C + z * c_split_stride) gets alpha * A[:, Kz] B[:, Kz]^T
int launch_gemm_nt_f64_split(hipStream_t s, int M, int N, int K, double alpha, const double* A,
                             int lda, const double* B, int ldb, double* C, int ldc, int splits,
                             int64_t c_split_stride);
int gemm_f64_splits(int M, int N, int K);
// gemm_f32.hip: the same products on float32 MFMA (float64 operands rounded as staged,
// float32 accumulation, float64 C)
int launch_gemm_nt_f32(hipStream_t s, int M, int N, int K, double alpha, const double* A,
                       int lda, const double* B, int ldb, double beta, double* C, int ldc);
int launch_gemm_nt_f32_split(hipStream_t s, int M, int N, int K, double alpha, const double* A,
                             int lda, const double* B, int ldb, double* C, int ldc, int splits,
                             int64_t c_split_stride);
// dst (cols x rows) = src (rows x cols)^T, float64
int launch_transpose_f64(hipStream_t s, int rows, int cols, const double* src, double* dst);
// gemm_bf16x3.hip: split-bf16 NT GEMM (K, lda, ldb multiples of 32 / 8), float / double out
int launch_gemm_nt_bf16x3_f32(hipStream_t s, int M, int N, int K, float alpha,
                              const uint16_t* Ah, const uint16_t* Al, int64_t lda,
                              const uint16_t* Bh, const uint16_t* Bl, int64_t ldb, float* C,
                              int64_t ldc, int splits, int64_t c_split_stride);
int launch_gemm_nt_bf16x3_f64(hipStream_t s, int M, int N, int K, float alpha,
                              const uint16_t* Ah, const uint16_t* Al, int64_t lda,
                              const uint16_t* Bh, const uint16_t* Bl, int64_t ldb, double* C,
                              int64_t ldc, int splits, int64_t c_split_stride);
int launch_split_bf16(hipStream_t s, int rows, int cols, const double* x, int64_t ld_src,
                      uint16_t* hi, uint16_t* lo, int64_t ld_dst);
int launch_sym_check(hipStream_t s, int N, int G, const double* chol_t, unsigned long long* out);
int launch_split_bf16_sym(hipStream_t s, int N, int G, const double* chol_t, uint16_t* hi,
                          uint16_t* lo, int ld_dst);
// measure.hip: green_f64 -> c->green (float64 split layout); green_bf -> c->green_bf
// With tail, the last spin's launch also applies the per-walker weight update (thread 0 of
// each walker's block); the population-control shift then comes from launch_pcs.
int launch_measure(afqmc_ctx* c, int W, const cplx* wup, const cplx* wdn, bool green_f64,
                   bool green_bf = false, const WeightsArgs* tail = nullptr);
// taylor.hip
int launch_taylor(afqmc_ctx* c, int W, cplx* wup, cplx* wdn);
// step.hip
int launch_fields(afqmc_ctx* c, int W, const double* fields, cplx* fb_out);
WeightsArgs weights_args(const afqmc_ctx* c, int W, double* weights, cplx* overlaps,
                         double* scalars);
int launch_pcs(afqmc_ctx* c, int W, const double* weights, double* scalars);
int launch_overlap_combine(afqmc_ctx* c, int W, cplx* overlaps);
int launch_fb_out(afqmc_ctx* c, int W, cplx* fb_out);
int launch_block_stats(afqmc_ctx* c, int W, const cplx* energies, const double* weights,
                       double* scalars, double* out);
int launch_cs_outliers(afqmc_ctx* c, int W, double* w1, double* w2, double z_thresh);
// energy.hip
int launch_energy(afqmc_ctx* c, int W, cplx* energies);
int energy_g_tiles(const afqmc_ctx* c, int s);
// qr.hip
int launch_qr(afqmc_ctx* c, int W, cplx* wup, cplx* wdn);
// sr.hip
int launch_sr(afqmc_ctx* c, int W, cplx* wup, cplx* wdn, double* weights, double zeta,
              int32_t* idx_out, bool global_order, uint64_t* key = nullptr);
int launch_fields_normal(afqmc_ctx* c, uint64_t seed, uint64_t counter, double* out,
                         int64_t count);
int launch_fields_normal_key(afqmc_ctx* c, uint64_t* key, double* out, int64_t count);
int launch_estimate_mix(afqmc_ctx* c, int nblocks, int stride, const double* stats,
                        double* scalars);

// ---------------------------------------------------------------- grouped launches
// Several independent problems (fragments x spins) share one grid: problem p owns the
// blocks [start[p], start[p + 1]) and reads its kernel arguments from args[p]; both
// tables live in device memory.  The kernel bodies are the single-problem ones.
__device__ inline int group_of(const int* __restrict__ start, int nprob, int b) {
  int lo = 0, hi = nprob - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (start[mid] <= b)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

struct MeasureArgs {
  int N, n, spin, ld16;
  const cplx* walkers;
  const double* trial;
  cplx* dets;
  double* green;
  uint16_t* ghi;
  uint16_t* glo;
};
struct FieldsArgs {
  int G, nparts, Gp, W;
  double sqdt;
  int64_t part_stride;
  const double* fb;
  const cplx* mf;
  const double* fields;  // step s reads fields + s * W * G
  double* xs;
  uint16_t* xs_hi;
  uint16_t* xs_lo;
  cplx* terms;
};
struct TaylorArgs {
  int N, n, K, rows_pad;
  int nv;           // stored entries of one V plane (N*N, or N(N+1)/2 packed)
  cplx* walkers;
  const void* E;    // float64 / float32 exp_h1, or the hi plane of exp_h1 - I (bf16 kernel)
  const void* E2;   // lo plane of exp_h1 - I (bf16 kernel)
  const void* vhs;  // float64 or float32 by kernel
};
struct GemmArgs {
  int M, N, K, tiles_m, tiles_n;
  float alpha;
  int64_t lda, ldb, ldc;
  const void* A;   // float64, or the bf16 hi plane
  const void* Al;  // bf16 lo plane (split-bf16 only)
  const void* B;
  const void* Bl;
  void* C;
};

// block ends: QR, overlap products, local energies
struct QrArgs {
  int m, n, use_lds;
  cplx* walkers;
};
struct OvcArgs {
  int W, nspin;
  const cplx* dets;
  cplx* out;
};
struct ExchArgs {
  int G, n, K, W, gpt, wpt, g_tiles, gc;
  const void* A;   // float64, or the bf16 hi plane (half-rotated Cholesky vectors)
  const void* Al;  // bf16 lo plane (split-bf16 only)
  const void* B;   // float64 Green's function, or its bf16 hi plane
  const void* Bl;
  cplx* xpart;
};
struct FinishArgs {
  int G, N, nspin, mult, n0, n1, nparts, gt0, gt1, W;
  int64_t part_stride;
  const double* fb;
  const cplx* xpart;
  const double* green0;
  const double* green1;
  const double* rot_h1_0;
  const double* rot_h1_1;
  double h0;
  cplx* energies;
};
// exchange kernel variants (launch_exchange_grouped)
enum ExchKind {
  EXCH_BF16X3_WN2 = 0,
  EXCH_BF16X3_WN4 = 1,
  EXCH_F64_V1 = 2,
  EXCH_F64_V2 = 3,
  EXCH_F32_V1 = 4,
  EXCH_F32_V2 = 5
};
// the exchange launch of context c, spin s, over W walkers (xpart: c's partial sums of
// that spin), and the finishing launch's arguments; used by the single and grouped forms
void exchange_plan(const afqmc_ctx* c, int s, int W, ExchArgs* a, int* kind, int* nblocks);
void energy_finish_args(const afqmc_ctx* c, int W, cplx* energies, FinishArgs* a);
int launch_exchange_grouped(hipStream_t st, int kind, const ExchArgs* args, const int* start,
                            int nprob, int nblocks);
int launch_energy_finish_grouped(hipStream_t st, const FinishArgs* args, const int* start,
                                 int nprob, int nblocks);
// LDS bytes of the QR walker copy (0: factorised in place in global memory)
size_t qr_lds_bytes(int m, int n);
int launch_qr_grouped(hipStream_t st, const QrArgs* args, const int* start, int nprob,
                      int nblocks, size_t shm);
int launch_overlap_combine_grouped(hipStream_t st, const OvcArgs* args, const int* start,
                                   int nprob, int nblocks);
constexpr int OVC_NT = 256;  // threads (walkers) per overlap-combine block
// a block's fields draw from a device-resident key {seed, counter}
struct DrawArgs {
  uint64_t* key;
  double* out;
  int64_t count;
};
int launch_fields_normal_key_grouped(hipStream_t st, const DrawArgs* args, const int* start,
                                     int nprob, int nblocks);

enum TaylorKind {
  TAYLOR_REG_F64 = 0,
  TAYLOR_STREAM_F64 = 1,
  TAYLOR_REG_BF16X3 = 2,
  TAYLOR_STREAM_BF16X3 = 3
};

struct TaylorPlan {
  int kind, nkgv, ct, nthreads, blocks_per_walker;
  size_t shm;
};
void taylor_plan(const afqmc_ctx* c, int s, TaylorPlan* pl);
void measure_plan(const afqmc_ctx* c, int s, int* t2, int* stage, size_t* shm);

int launch_measure_grouped(hipStream_t s, int t2, int stage, const MeasureArgs* args,
                           const int* start, int nprob, int nblocks, size_t shm);
int launch_fields_grouped(hipStream_t s, const FieldsArgs* args, const int* start, int nprob,
                          int nblocks, int step);
int launch_weights_grouped(hipStream_t s, const WeightsArgs* args, int nprob);
int launch_taylor_grouped(hipStream_t s, int kind, int nkgv, int ct, const TaylorArgs* args,
                          const int* start, int nprob, int nblocks, int nthreads, size_t shm);
// bm: 64 or 128 square tiles; out: 0 float64 C, 1 float32 C (split-bf16 only)
int launch_gemm_bf16x3_grouped(hipStream_t s, int bm, int out, const GemmArgs* args,
                               const int* start, int nprob, int nblocks);
int launch_gemm_f64_grouped(hipStream_t s, int bm, bool vec2, const GemmArgs* args,
                            const int* start, int nprob, int nblocks);
int launch_gemm_f32_grouped(hipStream_t s, int bm, bool vec2, const GemmArgs* args,
                            const int* start, int nprob, int nblocks);
