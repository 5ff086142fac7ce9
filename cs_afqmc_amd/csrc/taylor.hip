// Trotterised propagator on walker tiles:  w <- E (t0 + sum_{k=1}^{K} V^k t0 / k!),
// t0 = E w, with E = exp_h1 (real N x N) and V = VHS (complex N x N, per walker).
// Follows propagator._apply_trotprop_det (propagation.py:67-92): carry_k = V carry_{k-1}
// unscaled, each term divided by k!, summed, then added to t0.
//
// One workgroup per (walker, spin, 16-column chunk of the walker).  ceil(N/16) waves;
// wave r owns output rows [16 r, 16 r + 16).  The 16-column chunk of the current
// vector t lives in LDS (re / im planes, double buffered); V and E rows stream from
// global memory as the A operand of v_mfma_f64_16x16x4f64.  Columns of a walker are
// independent under left multiplication, so chunks never communicate.
#include "common.h"

namespace {

// LDS: T[buf][plane][q][j], q < rows_pad, j < 16
__device__ inline double* tptr(double* T, int rows_pad, int buf, int plane) {
  return T + (int64_t)((buf * 2 + plane) * rows_pad) * 16;
}

// acc{R,I} (+)= A x T for a real A (E):  A rows [rb, rb+16), all N columns
__device__ inline void mm_real(const double* __restrict__ A, int N, int rb, const double* tr,
                               const double* ti, int lane, double4_t& accR, double4_t& accI) {
  const int p = rb + (lane & 15);
  const int s = lane >> 4;
  const bool prow = p < N;
  const int nkg = (N + 3) >> 2;
  for (int kg = 0; kg < nkg; ++kg) {
    const int q = 4 * kg + s;
    const double a = (prow && q < N) ? A[(int64_t)p * N + q] : 0.0;
    const double br = tr[q * 16 + (lane & 15)];
    const double bi = ti[q * 16 + (lane & 15)];
    accR = __builtin_amdgcn_mfma_f64_16x16x4f64(a, br, accR, 0, 0, 0);
    accI = __builtin_amdgcn_mfma_f64_16x16x4f64(a, bi, accI, 0, 0, 0);
  }
}

// complex V (split planes) x complex T
__device__ inline void mm_cplx(const double* __restrict__ Vr, const double* __restrict__ Vi,
                               int N, int rb, const double* tr, const double* ti, int lane,
                               double4_t& accR, double4_t& accI) {
  const int p = rb + (lane & 15);
  const int s = lane >> 4;
  const bool prow = p < N;
  const int nkg = (N + 3) >> 2;
  for (int kg = 0; kg < nkg; ++kg) {
    const int q = 4 * kg + s;
    const bool ok = prow && q < N;
    const double ar = ok ? Vr[(int64_t)p * N + q] : 0.0;
    const double ai = ok ? Vi[(int64_t)p * N + q] : 0.0;
    const double br = tr[q * 16 + (lane & 15)];
    const double bi = ti[q * 16 + (lane & 15)];
    accR = __builtin_amdgcn_mfma_f64_16x16x4f64(ar, br, accR, 0, 0, 0);
    accR = __builtin_amdgcn_mfma_f64_16x16x4f64(-ai, bi, accR, 0, 0, 0);
    accI = __builtin_amdgcn_mfma_f64_16x16x4f64(ar, bi, accI, 0, 0, 0);
    accI = __builtin_amdgcn_mfma_f64_16x16x4f64(ai, br, accI, 0, 0, 0);
  }
}

// write a wave's 16x16 output tile (rows rb + lane/16 + 4r, col lane%16) into T planes
__device__ inline void put_tile(double* tr, double* ti, int rb, int lane, const double4_t& R,
                                const double4_t& I) {
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int q = rb + (lane >> 4) + 4 * r;
    tr[q * 16 + (lane & 15)] = R[r];
    ti[q * 16 + (lane & 15)] = I[r];
  }
}

__device__ __forceinline__ void taylor_body(int N, int n, int n_terms_k,
                                            cplx* __restrict__ walkers,
                                            const double* __restrict__ expE,
                                            const double* __restrict__ vhs, int rows_pad,
                                            const int w, const int chunk) {
  extern __shared__ double T[];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int rb = wave * 16;
  const int c0 = chunk * 16;
  cplx* phi = walkers + (int64_t)w * N * n;
  const double* Vr = vhs + (int64_t)w * 2 * N * N;
  const double* Vi = Vr + (int64_t)N * N;

  // load the chunk (zero padded) into buffer 0
  {
    double* tr = tptr(T, rows_pad, 0, 0);
    double* ti = tptr(T, rows_pad, 0, 1);
    for (int e = threadIdx.x; e < rows_pad * 16; e += blockDim.x) {
      const int q = e >> 4, j = e & 15;
      cplx v = cplx{0.0, 0.0};
      if (q < N && c0 + j < n) v = phi[(int64_t)q * n + c0 + j];
      tr[e] = v.re;
      ti[e] = v.im;
    }
  }
  __syncthreads();

  // t0 = E w
  double4_t t0R = {0, 0, 0, 0}, t0I = {0, 0, 0, 0};
  mm_real(expE, N, rb, tptr(T, rows_pad, 0, 0), tptr(T, rows_pad, 0, 1), lane, t0R, t0I);
  put_tile(tptr(T, rows_pad, 1, 0), tptr(T, rows_pad, 1, 1), rb, lane, t0R, t0I);
  __syncthreads();

  double4_t sR = {0, 0, 0, 0}, sI = {0, 0, 0, 0};
  int cur = 1;
  double fact = 1.0;
  for (int k = 1; k <= n_terms_k; ++k) {
    fact *= (double)k;
    double4_t cR = {0, 0, 0, 0}, cI = {0, 0, 0, 0};
    mm_cplx(Vr, Vi, N, rb, tptr(T, rows_pad, cur, 0), tptr(T, rows_pad, cur, 1), lane, cR, cI);
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      sR[r] += cR[r] / fact;
      sI[r] += cI[r] / fact;
    }
    put_tile(tptr(T, rows_pad, cur ^ 1, 0), tptr(T, rows_pad, cur ^ 1, 1), rb, lane, cR, cI);
    __syncthreads();
    cur ^= 1;
  }
  // R = t0 + sum_k terms
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    t0R[r] += sR[r];
    t0I[r] += sI[r];
  }
  put_tile(tptr(T, rows_pad, 0, 0), tptr(T, rows_pad, 0, 1), rb, lane, t0R, t0I);
  __syncthreads();
  double4_t oR = {0, 0, 0, 0}, oI = {0, 0, 0, 0};
  mm_real(expE, N, rb, tptr(T, rows_pad, 0, 0), tptr(T, rows_pad, 0, 1), lane, oR, oI);
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int p = rb + (lane >> 4) + 4 * r;
    const int j = c0 + (lane & 15);
    if (p < N && j < n) phi[(int64_t)p * n + j] = cplx{oR[r], oI[r]};
  }
}

__global__ __launch_bounds__(1024) void taylor_kernel(
    int N, int n, int n_terms_k, cplx* __restrict__ walkers, const double* __restrict__ expE,
    const double* __restrict__ vhs, int rows_pad) {
  taylor_body(N, n, n_terms_k, walkers, expE, vhs, rows_pad, blockIdx.x, blockIdx.y);
}

// grouped: local block b = walker * chunks + chunk, chunks = ceil(n / 16)
__global__ __launch_bounds__(1024) void taylor_grouped_kernel(const TaylorArgs* __restrict__ args,
                                                              const int* __restrict__ start,
                                                              int nprob) {
  const int p = group_of(start, nprob, blockIdx.x);
  const TaylorArgs a = args[p];
  const int b = blockIdx.x - start[p], chunks = (a.n + 15) / 16;
  taylor_body(a.N, a.n, a.K, a.walkers, static_cast<const double*>(a.E),
              static_cast<const double*>(a.vhs), a.rows_pad, b / chunks, b % chunks);
}

// ------------------------------------------------------------------------------------
// Register-resident form (N <= 128, n <= 32).  One workgroup per (walker, spin); wave r
// keeps rows [16 r, 16 r + 16) of A = [Vr | Vi] (N x 2Np, Np = N rounded up to 4) in
// registers for the whole series, so V is read from HBM once.  The complex product is
// one real product  [ur | ui] = [Vr | Vi] [[tr, ti], [-ti, tr]]  whose B operand lives
// in LDS as 2Np rows: rows [0, Np) = [tr | ti], rows [Np, 2Np) = [-ti | tr] (columns
// padded to CTP*16), so k-group kg reads rows 4 kg .. 4 kg + 3 of one uniform array.
// The series is evaluated in Horner form  y <- t0 + V y / k  (k = K..1), which equals
// t0 + sum_k V^k t0 / k! of propagation.py:67-92.
template <int CTP>
struct RegTile {
  static constexpr int RS = 16 * CTP + ((CTP % 2 == 0) ? 16 : 0);  // == 16 (mod 32) doubles
};

template <int NKGV, int CTP>
__device__ __forceinline__ void taylor_reg_body(int N, int n, int n_terms_k,
                                                cplx* __restrict__ walkers,
                                                const double* __restrict__ expE,
                                                const double* __restrict__ vhs, const int w) {
  constexpr int RS = RegTile<CTP>::RS;
  extern __shared__ double TT[];  // [2 Np][RS]
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int rb = wave * 16;
  const int Np = (N + 3) & ~3;
  const int nkg = Np >> 2;         // k-groups of one half (E products)
  const int nkgv = 2 * nkg;        // k-groups of [Vr | Vi]
  const int s = lane >> 4, col = lane & 15;
  const int p = rb + col;          // A-operand row of this lane
  const bool prow = p < N;
  const int n2 = 2 * n;
  cplx* phi = walkers + (int64_t)w * N * n;
  const double* Vr = vhs + (int64_t)w * 2 * N * N;
  const double* Vi = Vr + (int64_t)N * N;

  // A = [Vr | Vi] rows in registers: k-group kg, slot s -> column q = 4 kg + s
  double va[NKGV];
#pragma unroll
  for (int kg = 0; kg < NKGV; ++kg) {
    const int q = 4 * kg + s;
    double v = 0.0;
    if (kg < nkgv && prow) {
      if (q < Np) {
        if (q < N) v = Vr[(int64_t)p * N + q];
      } else if (q - Np < N) {
        v = Vi[(int64_t)p * N + q - Np];
      }
    }
    va[kg] = v;
  }

  // walker tile: [tr | ti] and [-ti | tr], zero padded
  for (int e = threadIdx.x; e < Np * RS; e += blockDim.x) {
    const int q = e / RS, c = e - (e / RS) * RS;
    double up = 0.0, lo = 0.0;
    if (q < N && c < n2) {
      const int j = c < n ? c : c - n;
      const cplx z = phi[(int64_t)q * n + j];
      up = c < n ? z.re : z.im;
      lo = c < n ? -z.im : z.re;
    }
    TT[e] = up;
    TT[Np * RS + e] = lo;
  }
  __syncthreads();

  const double* bbase = TT + s * RS + col;  // + kg * 4 RS + ct * 16

  // y rows -> both halves (rows >= Np never written: they belong to the lower half)
  auto put = [&](const double4_t* v) {
#pragma unroll
    for (int ct = 0; ct < CTP; ++ct)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = rb + s + 4 * r;
        const int cc = ct * 16 + col;
        if (row < Np) {
          TT[row * RS + cc] = v[ct][r];
          if (cc < n)
            TT[(Np + row) * RS + cc + n] = v[ct][r];
          else if (cc < n2)
            TT[(Np + row) * RS + cc - n] = -v[ct][r];
        }
      }
  };

  auto e_product = [&](double4_t* acc) {
#pragma unroll
    for (int ct = 0; ct < CTP; ++ct) acc[ct] = double4_t{0.0, 0.0, 0.0, 0.0};
    double a = (prow && s < N) ? expE[(int64_t)p * N + s] : 0.0;
    for (int kg = 0; kg < nkg; ++kg) {
      const int qn = 4 * (kg + 1) + s;
      const double an = (kg + 1 < nkg && prow && qn < N) ? expE[(int64_t)p * N + qn] : 0.0;
#pragma unroll
      for (int ct = 0; ct < CTP; ++ct)
        acc[ct] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, bbase[kg * 4 * RS + ct * 16], acc[ct],
                                                       0, 0, 0);
      a = an;
    }
  };

  double4_t t0[CTP], u[CTP];
  e_product(t0);
  __syncthreads();
  put(t0);
  __syncthreads();
  for (int k = n_terms_k; k >= 1; --k) {
#pragma unroll
    for (int ct = 0; ct < CTP; ++ct) u[ct] = double4_t{0.0, 0.0, 0.0, 0.0};
    double b[CTP];
#pragma unroll
    for (int ct = 0; ct < CTP; ++ct) b[ct] = bbase[ct * 16];
#pragma unroll
    for (int kg = 0; kg < NKGV; ++kg) {
      if (kg < nkgv) {
        // software pipeline: B of k-group kg + 1 is read from LDS under the MFMAs of kg
        double bn[CTP];
#pragma unroll
        for (int ct = 0; ct < CTP; ++ct)
          bn[ct] = (kg + 1 < nkgv) ? bbase[(kg + 1) * 4 * RS + ct * 16] : 0.0;
#pragma unroll
        for (int ct = 0; ct < CTP; ++ct)
          u[ct] = __builtin_amdgcn_mfma_f64_16x16x4f64(va[kg], b[ct], u[ct], 0, 0, 0);
#pragma unroll
        for (int ct = 0; ct < CTP; ++ct) b[ct] = bn[ct];
      }
      __builtin_amdgcn_sched_barrier(0);
    }
    __syncthreads();
    const double rk = (double)k;
#pragma unroll
    for (int ct = 0; ct < CTP; ++ct)
#pragma unroll
      for (int r = 0; r < 4; ++r) u[ct][r] = t0[ct][r] + u[ct][r] / rk;
    put(u);
    __syncthreads();
  }
  e_product(u);
#pragma unroll
  for (int ct = 0; ct < CTP; ++ct)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int row = rb + s + 4 * r;
      const int cc = ct * 16 + col;
      if (row < N && cc < n2) {
        double* dst = reinterpret_cast<double*>(phi + (int64_t)row * n + (cc < n ? cc : cc - n));
        dst[cc < n ? 0 : 1] = u[ct][r];
      }
    }
}

template <int NKGV, int CTP>
__global__ __launch_bounds__(512) void taylor_reg_kernel(
    int N, int n, int n_terms_k, cplx* __restrict__ walkers, const double* __restrict__ expE,
    const double* __restrict__ vhs) {
  taylor_reg_body<NKGV, CTP>(N, n, n_terms_k, walkers, expE, vhs, blockIdx.x);
}

template <int NKGV, int CTP>
__global__ __launch_bounds__(512) void taylor_reg_grouped_kernel(
    const TaylorArgs* __restrict__ args, const int* __restrict__ start, int nprob) {
  const int p = group_of(start, nprob, blockIdx.x);
  const TaylorArgs a = args[p];
  taylor_reg_body<NKGV, CTP>(a.N, a.n, a.K, a.walkers, static_cast<const double*>(a.E),
                             static_cast<const double*>(a.vhs), blockIdx.x - start[p]);
}

template <int NKGV, int CTP>
int launch_reg(afqmc_ctx* c, int W, int n, cplx* wk, const double* E, const double* V) {
  const int N = c->d.norb;
  const int nwaves = (N + 15) / 16;
  const int Np = (N + 3) & ~3;
  const size_t shm = (size_t)2 * Np * RegTile<CTP>::RS * sizeof(double);
  static int lds_limit = 0;
  const int rc = ensure_dyn_lds((const void*)taylor_reg_kernel<NKGV, CTP>, shm, &lds_limit);
  if (rc) return rc;
  hipLaunchKernelGGL((taylor_reg_kernel<NKGV, CTP>), dim3(W), dim3(nwaves * 64), shm, c->stream,
                     N, n, c->d.n_exp_terms - 1, wk, E, V);
  LAUNCH_CHECK();
  return AFQMC_OK;
}

// ------------------------------------------------------------------------------------
// Split-bf16 register-resident form for the mixed mode (N <= 128, n <= 32).  The complex
// product is one real block product [ur | ui] = [Vr | Vi] [[yr, yi], [-yi, yr]] on
// v_mfma_f32_16x16x32_bf16, with each operand x = hi + lo (bf16 planes) and the three
// leading products al bh + ah bl + ah bh accumulated in float32:
//   A (V rows, registers): lane l holds A[l % 16][8 (l / 16) + e] of each 32-wide k block
//   B (y, LDS [column][k], k contiguous): lane l holds B[8 (l / 16) + e][l % 16]
//   D (float4): lane l, element r holds D[4 (l / 16) + r][l % 16]
// The k axis is packed tightly: with Nh = N rounded up to 4, k in [0, Nh) is the real
// part of V (top half of the y column), k in [Nh, 2 Nh) the imaginary part (bottom half),
// and KB = ceil(2 Nh / 32) k blocks cover it.  A lane's four D rows 4 (l / 16) + r are
// consecutive k of one LDS column, so y goes to LDS as 8-byte stores.
// The one-body factor is applied as E y = y + (E - I) y with E - I in split bf16 (rows of
// Np32 = 32 ceil(N / 32) entries, zero beyond N): its entries are O(dt), so the product's
// relative error stays far below float32 rounding of E y itself.  The Horner recursion
// y <- t0 + V y / k and all additions run in float32; the walker is read and written in
// float64.
// LDS rows (one column of y) are 32 KB + 16 bf16 long, 32 bytes past a multiple of 64,
// and 16-byte chunk j of column c is stored at chunk j ^ ((c >> 2) & 1).  With that, the
// sixteen columns a wave's B operand reads (8 consecutive k each) meet no bank conflict,
// and the D-layout stores of put (4 consecutive k) cost no more than with unswizzled
// rows 16 bytes past a multiple of 32, which conflict on the reads.
template <int KB>
struct BfTile {
  static constexpr int LDK = 32 * KB + 16;  // LDS row (one column of y) in bf16 elements
  static constexpr int KH = (KB + 1) / 2;   // 32-wide k blocks of E - I: ceil(N / 32)
};
inline int bf_ldk(int kb) { return 32 * kb + 16; }  // BfTile<kb>::LDK for a runtime kb
// element offset of k (a multiple of 4) in LDS column c
__device__ inline int bf_koff(int c, int k) { return (((k >> 3) ^ ((c >> 2) & 1)) << 3) | (k & 7); }
// y's real part spans k < Nh, its imaginary part k in [Nh, 2 Nh); Nh = N rounded up to 8
// keeps both halves' 4-row groups in the same chunk halves
__host__ __device__ inline int bf_nh(int N) { return (N + 7) & ~7; }

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef unsigned int u32x4_t __attribute__((ext_vector_type(4)));
typedef unsigned int u32x2_t __attribute__((ext_vector_type(2)));

// 8 floats -> hi and lo bf16 planes (4 dwords each)
__device__ inline void split8(const float (&v)[8], u32x4_t& ph, u32x4_t& pl) {
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const uint32_t h = pack_bf16x2(v[2 * j], v[2 * j + 1]);
    ph[j] = h;
    pl[j] = pack_bf16x2(v[2 * j] - bf16_lo_val(h), v[2 * j + 1] - bf16_hi_val(h));
  }
}

template <int KB, int CTP>
__device__ __forceinline__ void taylor_bf_body(int N, int n, int n_terms_k,
                                               cplx* __restrict__ walkers,
                                               const uint16_t* __restrict__ emi_hi,
                                               const uint16_t* __restrict__ emi_lo,
                                               const float* __restrict__ vhs, int nv,
                                               const int w) {
  constexpr int LDK = BfTile<KB>::LDK;
  constexpr int KH = BfTile<KB>::KH;
  extern __shared__ uint16_t YB[];  // [2 planes][CTP * 16 columns][LDK]
  uint16_t* Yh = YB;
  uint16_t* Yl = YB + CTP * 16 * LDK;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int rb = wave * 16;
  const int np32 = 32 * KH;
  const int Nh = bf_nh(N);
  const int q4 = lane >> 4, col = lane & 15;
  const int q4s = q4 ^ ((col >> 2) & 1);  // this lane's B chunk within a 32-wide k block
  const int p = rb + col;  // A row of this lane
  const bool prow = p < N;
  const int n2 = 2 * n;
  cplx* phi = walkers + (int64_t)w * N * n;
  const float* Vr = vhs + (int64_t)w * 2 * nv;
  const float* Vi = Vr + nv;
  // V planes hold all N^2 entries row-major, or (nv = N(N+1)/2, symmetric V) the upper
  // triangle row by row: entry (a, b), a <= b, at a N - a (a - 1) / 2 + b - a
  const bool packed = nv != N * N;

  // walker -> registers (D layout), issued first so that its round trip overlaps the V
  // loads'; clamped into the walker and masked after, like the V loads below
  float4_t y[CTP], t0[CTP];
  {
    const double* wz = reinterpret_cast<const double*>(phi);
#pragma unroll
    for (int ct = 0; ct < CTP; ++ct)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = rb + 4 * q4 + r;
        const int c = ct * 16 + col;
        const bool ok = (row < N) & (c < n2);
        const int rc = min(row, N - 1);
        const int cc = c < n ? c : min(c - n, n - 1);
        const double z = wz[2 * ((int64_t)rc * n + cc) + (c < n ? 0 : 1)];
        y[ct][r] = ok ? (float)z : 0.f;
      }
  }

  // V rows in registers, split once; loads are clamped into the matrix and masked after,
  // so they issue back to back without branches
  bf16x8_t vh[KB], vl[KB];
  const int pc = prow ? p : 0;
#pragma unroll
  for (int kb = 0; kb < KB; ++kb) {
    float v[8];
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const int k = kb * 32 + 8 * q4 + e;
      const bool im = k >= Nh;
      const int kk = im ? k - Nh : k;
      const bool ok = prow && kk < N;
      const int kc = ok ? kk : 0;
      const int a = pc < kc ? pc : kc, b = pc < kc ? kc : pc;
      const int idx = packed ? a * N - (a * (a - 1)) / 2 + (b - a) : pc * N + kc;
      const float t = (im ? Vi : Vr)[idx];
      v[e] = ok ? t : 0.f;
    }
    u32x4_t ph, pl;
    split8(v, ph, pl);
    vh[kb] = __builtin_bit_cast(bf16x8_t, ph);
    vl[kb] = __builtin_bit_cast(bf16x8_t, pl);
  }

  {
    u32x4_t* z = reinterpret_cast<u32x4_t*>(YB);
    const int nz = 2 * CTP * 16 * LDK / 8;
    for (int e = threadIdx.x; e < nz; e += blockDim.x) z[e] = u32x4_t{0u, 0u, 0u, 0u};
  }
  __syncthreads();

  // y (D layout, float32) -> LDS planes: column c < n holds [yr_c ; -yi_c], column n + c
  // holds [yi_c ; yr_c] (top half at k = row, bottom half at k = Nh + row)
  auto put = [&](const float4_t* y) {
    const int row0 = rb + 4 * q4;
#pragma unroll
    for (int ct = 0; ct < CTP; ++ct) {
      const int c = ct * 16 + col;
      if (row0 < N && c < n2) {
        const uint32_t h01 = pack_bf16x2(y[ct][0], y[ct][1]);
        const uint32_t h23 = pack_bf16x2(y[ct][2], y[ct][3]);
        const uint32_t l01 = pack_bf16x2(y[ct][0] - bf16_lo_val(h01), y[ct][1] - bf16_hi_val(h01));
        const uint32_t l23 = pack_bf16x2(y[ct][2] - bf16_lo_val(h23), y[ct][3] - bf16_hi_val(h23));
        const int ot = c * LDK + bf_koff(c, row0);
        *reinterpret_cast<u32x2_t*>(Yh + ot) = u32x2_t{h01, h23};
        *reinterpret_cast<u32x2_t*>(Yl + ot) = u32x2_t{l01, l23};
        const int cb = c < n ? c + n : c - n;
        const uint32_t sg = c < n ? 0u : 0x80008000u;  // bottom half: +yr, -yi
        const int ob = cb * LDK + bf_koff(cb, Nh + row0);
        *reinterpret_cast<u32x2_t*>(Yh + ob) = u32x2_t{h01 ^ sg, h23 ^ sg};
        *reinterpret_cast<u32x2_t*>(Yl + ob) = u32x2_t{l01 ^ sg, l23 ^ sg};
      }
    }
  };
  auto bload = [&](const uint16_t* plane, int ct, int kb) {
    return __builtin_bit_cast(
        bf16x8_t, *reinterpret_cast<const u32x4_t*>(plane + (ct * 16 + col) * LDK + kb * 32 + 8 * q4s));
  };
  // acc = (E - I) y: A (rows of E - I) streams from L2; k < Np32 reads the top half of
  // each column and, past Nh, bottom-half entries that meet zero columns of E - I
  auto e_product = [&](float4_t* acc) {
#pragma unroll
    for (int ct = 0; ct < CTP; ++ct) acc[ct] = float4_t{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kb = 0; kb < KH; ++kb) {
      u32x4_t ah = {0, 0, 0, 0}, al = {0, 0, 0, 0};
      if (prow) {
        ah = *reinterpret_cast<const u32x4_t*>(emi_hi + (int64_t)p * np32 + kb * 32 + 8 * q4);
        al = *reinterpret_cast<const u32x4_t*>(emi_lo + (int64_t)p * np32 + kb * 32 + 8 * q4);
      }
      const bf16x8_t a_h = __builtin_bit_cast(bf16x8_t, ah), a_l = __builtin_bit_cast(bf16x8_t, al);
#pragma unroll
      for (int ct = 0; ct < CTP; ++ct) {
        const bf16x8_t bh = bload(Yh, ct, kb), bl = bload(Yl, ct, kb);
        acc[ct] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a_l, bh, acc[ct], 0, 0, 0);
        acc[ct] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a_h, bl, acc[ct], 0, 0, 0);
        acc[ct] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a_h, bh, acc[ct], 0, 0, 0);
      }
    }
  };

  put(y);
  __syncthreads();
  e_product(t0);
#pragma unroll
  for (int ct = 0; ct < CTP; ++ct) t0[ct] += y[ct];
  __syncthreads();
  put(t0);
  __syncthreads();
  for (int k = n_terms_k; k >= 1; --k) {
#pragma unroll
    for (int ct = 0; ct < CTP; ++ct) y[ct] = float4_t{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kb = 0; kb < KB; ++kb) {
#pragma unroll
      for (int ct = 0; ct < CTP; ++ct) {
        const bf16x8_t bh = bload(Yh, ct, kb), bl = bload(Yl, ct, kb);
        y[ct] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(vl[kb], bh, y[ct], 0, 0, 0);
        y[ct] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(vh[kb], bl, y[ct], 0, 0, 0);
        y[ct] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(vh[kb], bh, y[ct], 0, 0, 0);
      }
    }
    __syncthreads();
    const float rk = 1.0f / (float)k;
#pragma unroll
    for (int ct = 0; ct < CTP; ++ct) y[ct] = t0[ct] + y[ct] * rk;
    put(y);
    __syncthreads();
  }
  float4_t acc[CTP];
  e_product(acc);
#pragma unroll
  for (int ct = 0; ct < CTP; ++ct)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int row = rb + 4 * q4 + r;
      const int c = ct * 16 + col;
      if (row < N && c < n2) {
        double* dst = reinterpret_cast<double*>(phi + (int64_t)row * n + (c < n ? c : c - n));
        dst[c < n ? 0 : 1] = (double)(y[ct][r] + acc[ct][r]);
      }
    }
}

// Two 7- or 8-wave workgroups per CU need at most 128 VGPRs (4 waves per SIMD); the
// widest tiles (n > 24 with N > 96) would spill inside the series at that budget.
constexpr int bf_waves_per_simd(int KB, int CTP) { return (CTP == 4 && KB >= 7) ? 3 : 4; }

template <int KB, int CTP>
__global__ __launch_bounds__(512, bf_waves_per_simd(KB, CTP)) void taylor_bf_kernel(int N, int n, int n_terms_k,
                                                           cplx* __restrict__ walkers,
                                                           const uint16_t* __restrict__ emi_hi,
                                                           const uint16_t* __restrict__ emi_lo,
                                                           const float* __restrict__ vhs,
                                                           int nv) {
  taylor_bf_body<KB, CTP>(N, n, n_terms_k, walkers, emi_hi, emi_lo, vhs, nv, blockIdx.x);
}

template <int KB, int CTP>
__global__ __launch_bounds__(512, bf_waves_per_simd(KB, CTP)) void taylor_bf_grouped_kernel(
    const TaylorArgs* __restrict__ args, const int* __restrict__ start, int nprob) {
  const int pidx = group_of(start, nprob, blockIdx.x);
  const TaylorArgs a = args[pidx];
  taylor_bf_body<KB, CTP>(a.N, a.n, a.K, a.walkers, static_cast<const uint16_t*>(a.E),
                          static_cast<const uint16_t*>(a.E2), static_cast<const float*>(a.vhs),
                          a.nv, blockIdx.x - start[pidx]);
}

// k blocks of the packed [Vr | Vi] axis
inline int bf_kblocks(int N) { return (2 * bf_nh(N) + 31) / 32; }
// threads of a register-form workgroup: a wave per 16 rows, with 6 or 7 waves rounded up
// to 8 (the extra waves own rows >= N: zero work, but every SIMD then holds two waves of
// the group at its barriers, and at four waves per SIMD two such groups fill a CU either
// way); 5 waves stay 5, which lets three groups share a CU
inline int bf_threads(int N) {
  const int w = (N + 15) / 16;
  return 64 * ((w == 6 || w == 7) ? 8 : w);
}

template <int KB, int CTP>
int launch_bf(afqmc_ctx* c, int W, int n, cplx* wk, int s) {
  const int N = c->d.norb;
  if (BfTile<KB>::KH * 32 != c->Np32 || bf_kblocks(N) != KB)
    return afqmc_set_error(AFQMC_ERR_ARG, "taylor bf16x3: k blocks %d do not fit N %d", KB, N);
  const size_t shm = (size_t)2 * CTP * 16 * BfTile<KB>::LDK * sizeof(uint16_t);
  static int lds_limit = 0;
  const int rc = ensure_dyn_lds((const void*)taylor_bf_kernel<KB, CTP>, shm, &lds_limit);
  if (rc) return rc;
  hipLaunchKernelGGL((taylor_bf_kernel<KB, CTP>), dim3(W), dim3(bf_threads(N)), shm, c->stream, N,
                     n, c->d.n_exp_terms - 1, wk, c->emi_bf[s][0], c->emi_bf[s][1],
                     vhs_f32_of(c, s), c->vhs_nv);
  LAUNCH_CHECK();
  return AFQMC_OK;
}

template <int KB>
int launch_bf_ct(afqmc_ctx* c, int W, int n, cplx* wk, int s) {
  switch ((2 * n + 15) / 16) {
    case 1: return launch_bf<KB, 1>(c, W, n, wk, s);
    case 2: return launch_bf<KB, 2>(c, W, n, wk, s);
    case 3: return launch_bf<KB, 3>(c, W, n, wk, s);
    default: return launch_bf<KB, 4>(c, W, n, wk, s);
  }
}

// ------------------------------------------------------------------------------------
// Split-bf16 streaming form for the mixed mode beyond the register kernel's range
// (N <= 256, n <= 64).  One workgroup per walker of one spin; wave r owns RT 16-row
// output tiles starting at row 16 RT r, and every column.  y = [yr | yi] (N x 2 n, each
// half padded to 16 CT columns) lives in LDS as hi / lo bf16 planes stored [column][k]
// with k contiguous, so a lane's B operand (8 consecutive k of one column) is one 16-byte
// read; k is padded with zeros to 32 ceil(N / 32).  V streams from global memory as
// float32 rows, 32 k at a time with the next block in flight while the current one
// multiplies, and is split into bf16 planes in registers.  The complex product is four
// real products,
//   ur = Vr yr - Vi yi,  ui = Vr yi + Vi yr,
// each as three bf16 products (al bh + ah bl + ah bh) on v_mfma_f32_16x16x32_bf16 with
// float32 accumulation; the planes of -Vi are those of Vi with the sign bits flipped
// (round-to-nearest-even is symmetric).  E y = y + (E - I) y uses the split planes of
// E - I (rows of 32 ceil(N / 32) entries) as the register form does, and the Horner
// recursion y <- t0 + V y / k runs in float32 registers.  The body takes RT 16-row tiles
// per wave (RT = 2 shares each B operand between two tiles at 256 registers per lane);
// one tile per wave, up to 16 waves at 128 registers, is what runs: with the packed-V
// gathers it measured faster for three and four column tiles despite spilling 25-48
// registers around the Horner steps (UHF N = 200: 2.21 -> 2.04 ms per step).
typedef float f4u_t __attribute__((ext_vector_type(4), aligned(4)));  // dword-aligned rows

inline int bfs_ldk(int N) { return 32 * ((N + 31) / 32) + 16; }
inline size_t bfs_shm(int N, int ct) {
  return (size_t)2 * 32 * ct * bfs_ldk(N) * sizeof(uint16_t);
}
constexpr int bfs_rt(int) { return 1; }
inline int bfs_threads(int N, int ct) {
  const int rows = 16 * bfs_rt(ct);
  return 64 * ((N + rows - 1) / rows);
}

template <int CT, int RT>
__device__ __forceinline__ void taylor_bfs_body(int N, int n, int n_terms_k,
                                                cplx* __restrict__ walkers,
                                                const uint16_t* __restrict__ emi_hi,
                                                const uint16_t* __restrict__ emi_lo,
                                                const float* __restrict__ vhs, int nv,
                                                const int w) {
  constexpr int NC = 32 * CT;  // LDS columns: CT tiles of yr, then CT tiles of yi
  const int KBk = (N + 31) >> 5;
  const int LDK = 32 * KBk + 16;  // rows and chunk swizzle as in the register form
  const int np32 = 32 * KBk;
  extern __shared__ uint16_t YB[];  // [2 planes][NC][LDK]
  uint16_t* Yh = YB;
  uint16_t* Yl = YB + NC * LDK;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int rb = wave * 16 * RT;
  const int q4 = lane >> 4, col = lane & 15;
  const int sw = (col >> 2) & 1;  // chunk swizzle of this lane's LDS columns
  const int q4s = q4 ^ sw;
  // uniform bases with 32-bit lane offsets (saddr + voffset addressing)
  cplx* phi = walkers + (int64_t)w * N * n;
  // V planes: all N^2 entries row-major, or (nv = N(N+1)/2, symmetric V) the upper
  // triangle row by row, entry (a, b), a <= b, at a N - a (a - 1) / 2 + b - a
  const bool packed = nv != N * N;
  const float* Vr = vhs + (int64_t)w * 2 * nv;
  const float* Vi = Vr + nv;
  bool prow[RT];
  int prw[RT], voff[RT], eoff[RT];
#pragma unroll
  for (int t = 0; t < RT; ++t) {
    const int p = rb + 16 * t + col;  // A row of this lane in tile t
    prow[t] = p < N;
    prw[t] = prow[t] ? p : 0;
    // start of row p (packed: of its upper part, entry (p, p))
    voff[t] = packed ? prw[t] * N - (prw[t] * (prw[t] - 1)) / 2 : prw[t] * N;
    eoff[t] = prw[t] * np32;
  }

  {
    u32x4_t* z = reinterpret_cast<u32x4_t*>(YB);
    const int nz = 2 * NC * LDK / 8;
    for (int e = threadIdx.x; e < nz; e += blockDim.x) z[e] = u32x4_t{0u, 0u, 0u, 0u};
  }
  __syncthreads();

  // D layout (rows 4 (l / 16) + r of a tile, column l % 16) -> LDS planes: four
  // consecutive k of one column, 8-byte stores.  Rows >= N and columns >= n carry zeros
  // (their A rows / B columns are zero), which keeps the padding zero.
  auto put = [&](const float4_t (*yr)[CT], const float4_t (*yi)[CT]) {
#pragma unroll
    for (int t = 0; t < RT; ++t)
#pragma unroll
      for (int h = 0; h < 2; ++h)
#pragma unroll
        for (int ct = 0; ct < CT; ++ct) {
          const float4_t v = h ? yi[t][ct] : yr[t][ct];
          const int off = (h * 16 * CT + ct * 16 + col) * LDK + rb + 16 * t + 8 * ((q4 >> 1) ^ sw) + 4 * (q4 & 1);
          const uint32_t h01 = pack_bf16x2(v[0], v[1]);
          const uint32_t h23 = pack_bf16x2(v[2], v[3]);
          const uint32_t l01 = pack_bf16x2(v[0] - bf16_lo_val(h01), v[1] - bf16_hi_val(h01));
          const uint32_t l23 = pack_bf16x2(v[2] - bf16_lo_val(h23), v[3] - bf16_hi_val(h23));
          *reinterpret_cast<u32x2_t*>(Yh + off) = u32x2_t{h01, h23};
          *reinterpret_cast<u32x2_t*>(Yl + off) = u32x2_t{l01, l23};
        }
  };
  auto bload = [&](const uint16_t* plane, int c0, int kb) {
    return __builtin_bit_cast(
        bf16x8_t, *reinterpret_cast<const u32x4_t*>(plane + (c0 + col) * LDK + kb * 32 + 8 * q4s));
  };
  auto zero = [&](float4_t (*a)[CT]) {
#pragma unroll
    for (int t = 0; t < RT; ++t)
#pragma unroll
      for (int ct = 0; ct < CT; ++ct) a[t][ct] = float4_t{0.f, 0.f, 0.f, 0.f};
  };
  // acc = (E - I) y, a real operator on both halves
  auto e_product = [&](float4_t (*ar)[CT], float4_t (*ai)[CT]) {
    zero(ar);
    zero(ai);
#pragma unroll 1
    for (int kb = 0; kb < KBk; ++kb) {
      bf16x8_t a_h[RT], a_l[RT];
#pragma unroll
      for (int t = 0; t < RT; ++t) {
        u32x4_t ah = {0, 0, 0, 0}, al = {0, 0, 0, 0};
        if (prow[t]) {
          ah = *reinterpret_cast<const u32x4_t*>(emi_hi + (eoff[t] + kb * 32 + 8 * q4));
          al = *reinterpret_cast<const u32x4_t*>(emi_lo + (eoff[t] + kb * 32 + 8 * q4));
        }
        a_h[t] = __builtin_bit_cast(bf16x8_t, ah);
        a_l[t] = __builtin_bit_cast(bf16x8_t, al);
      }
#pragma unroll
      for (int ct = 0; ct < CT; ++ct) {
        const bf16x8_t rh = bload(Yh, ct * 16, kb), rl = bload(Yl, ct * 16, kb);
#pragma unroll
        for (int t = 0; t < RT; ++t) {
          ar[t][ct] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a_l[t], rh, ar[t][ct], 0, 0, 0);
          ar[t][ct] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a_h[t], rl, ar[t][ct], 0, 0, 0);
          ar[t][ct] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a_h[t], rh, ar[t][ct], 0, 0, 0);
        }
        const bf16x8_t ih = bload(Yh, (CT + ct) * 16, kb), il = bload(Yl, (CT + ct) * 16, kb);
#pragma unroll
        for (int t = 0; t < RT; ++t) {
          ai[t][ct] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a_l[t], ih, ai[t][ct], 0, 0, 0);
          ai[t][ct] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a_h[t], il, ai[t][ct], 0, 0, 0);
          ai[t][ct] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a_h[t], ih, ai[t][ct], 0, 0, 0);
        }
      }
    }
  };
  // 8 consecutive k of a V row, for chunks that start inside the row (others are zero).
  // A chunk on or right of the diagonal is contiguous in the stored row: two 16-byte
  // loads.  A chunk that straddles k = N reads at most 7 floats past the row end, into the
  // next row or, for the last row of the last walker, into the 8 floats of tail padding
  // of the VHS buffer, and is masked after the load.  In the packed layout a chunk left
  // of the diagonal reads entries (k, p) of rows k < p one by one; for a fixed k the 16
  // rows of a lane group are consecutive floats of one stored row.
  auto vload = [&](int kb, float (*r)[8], float (*i)[8]) {
    const int k0 = kb * 32 + 8 * q4;
#pragma unroll
    for (int t = 0; t < RT; ++t) {
      f4u_t r0 = {0.f, 0.f, 0.f, 0.f}, r1 = r0, i0 = r0, i1 = r0;
      const int p = prw[t];
      if (prow[t] && k0 < N && (!packed || k0 >= p)) {
        const int base = voff[t] + (packed ? k0 - p : k0);
        r0 = *reinterpret_cast<const f4u_t*>(Vr + base);
        r1 = *reinterpret_cast<const f4u_t*>(Vr + (base + 4));
        i0 = *reinterpret_cast<const f4u_t*>(Vi + base);
        i1 = *reinterpret_cast<const f4u_t*>(Vi + (base + 4));
      }
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        r[t][e] = r0[e];
        r[t][4 + e] = r1[e];
        i[t][e] = i0[e];
        i[t][4 + e] = i1[e];
      }
      if (packed && prow[t] && k0 < p) {
#pragma unroll
        for (int e = 0; e < 8; ++e) {
          const int k = k0 + e < N ? k0 + e : N - 1;
          const int a = k < p ? k : p, b = k < p ? p : k;
          const int idx = a * N - (a * (a - 1)) / 2 + (b - a);
          r[t][e] = Vr[idx];
          i[t][e] = Vi[idx];
        }
      }
      if (k0 + 8 > N) {
#pragma unroll
        for (int e = 0; e < 8; ++e)
          if (k0 + e >= N) r[t][e] = i[t][e] = 0.f;
      }
    }
  };
  auto v_product = [&](float4_t (*ur)[CT], float4_t (*ui)[CT]) {
    zero(ur);
    zero(ui);
    float rr[RT][8], ri[RT][8];
    vload(0, rr, ri);
#pragma unroll 1
    for (int kb = 0; kb < KBk; ++kb) {
      u32x4_t rh[RT], rl[RT], ih[RT], il[RT];
#pragma unroll
      for (int t = 0; t < RT; ++t) {
        split8(rr[t], rh[t], rl[t]);
        split8(ri[t], ih[t], il[t]);
      }
      if (kb + 1 < KBk) vload(kb + 1, rr, ri);  // next block in flight
      auto bf = [](const u32x4_t& x) { return __builtin_bit_cast(bf16x8_t, x); };
      // rows of yr: ur += Vr yr, ui += Vi yr
#pragma unroll
      for (int ct = 0; ct < CT; ++ct) {
        const bf16x8_t yh = bload(Yh, ct * 16, kb), yl = bload(Yl, ct * 16, kb);
#pragma unroll
        for (int t = 0; t < RT; ++t) {
          ur[t][ct] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bf(rl[t]), yh, ur[t][ct], 0, 0, 0);
          ur[t][ct] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bf(rh[t]), yl, ur[t][ct], 0, 0, 0);
          ur[t][ct] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bf(rh[t]), yh, ur[t][ct], 0, 0, 0);
          ui[t][ct] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bf(il[t]), yh, ui[t][ct], 0, 0, 0);
          ui[t][ct] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bf(ih[t]), yl, ui[t][ct], 0, 0, 0);
          ui[t][ct] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bf(ih[t]), yh, ui[t][ct], 0, 0, 0);
        }
      }
      // rows of yi: ur += (-Vi) yi, ui += Vr yi (Vi's planes negated in place)
      const u32x4_t sg = {0x80008000u, 0x80008000u, 0x80008000u, 0x80008000u};
#pragma unroll
      for (int t = 0; t < RT; ++t) {
        ih[t] ^= sg;
        il[t] ^= sg;
      }
#pragma unroll
      for (int ct = 0; ct < CT; ++ct) {
        const bf16x8_t yh = bload(Yh, (CT + ct) * 16, kb), yl = bload(Yl, (CT + ct) * 16, kb);
#pragma unroll
        for (int t = 0; t < RT; ++t) {
          ur[t][ct] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bf(il[t]), yh, ur[t][ct], 0, 0, 0);
          ur[t][ct] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bf(ih[t]), yl, ur[t][ct], 0, 0, 0);
          ur[t][ct] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bf(ih[t]), yh, ur[t][ct], 0, 0, 0);
          ui[t][ct] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bf(rl[t]), yh, ui[t][ct], 0, 0, 0);
          ui[t][ct] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bf(rh[t]), yl, ui[t][ct], 0, 0, 0);
          ui[t][ct] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bf(rh[t]), yh, ui[t][ct], 0, 0, 0);
        }
      }
    }
  };

  // walker (float64) -> float32 D layout -> LDS
  float4_t t0r[RT][CT], t0i[RT][CT], ur[RT][CT], ui[RT][CT];
#pragma unroll
  for (int t = 0; t < RT; ++t)
#pragma unroll
    for (int ct = 0; ct < CT; ++ct)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = rb + 16 * t + 4 * q4 + r;
        const int c = ct * 16 + col;
        float re = 0.f, im = 0.f;
        if (row < N && c < n) {
          const cplx z = phi[row * n + c];
          re = (float)z.re;
          im = (float)z.im;
        }
        t0r[t][ct][r] = re;
        t0i[t][ct][r] = im;
      }
  put(t0r, t0i);
  __syncthreads();
  e_product(ur, ui);
#pragma unroll
  for (int t = 0; t < RT; ++t)
#pragma unroll
    for (int ct = 0; ct < CT; ++ct) {
      t0r[t][ct] += ur[t][ct];
      t0i[t][ct] += ui[t][ct];
    }
  __syncthreads();
  put(t0r, t0i);
  __syncthreads();
  for (int k = n_terms_k; k >= 1; --k) {
    v_product(ur, ui);
    __syncthreads();
    const float rk = 1.0f / (float)k;
#pragma unroll
    for (int t = 0; t < RT; ++t)
#pragma unroll
      for (int ct = 0; ct < CT; ++ct) {
        ur[t][ct] = t0r[t][ct] + ur[t][ct] * rk;
        ui[t][ct] = t0i[t][ct] + ui[t][ct] * rk;
      }
    put(ur, ui);
    __syncthreads();
  }
  float4_t ar[RT][CT], ai[RT][CT];
  e_product(ar, ai);
#pragma unroll
  for (int t = 0; t < RT; ++t)
#pragma unroll
    for (int ct = 0; ct < CT; ++ct)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = rb + 16 * t + 4 * q4 + r;
        const int c = ct * 16 + col;
        if (row < N && c < n)
          phi[row * n + c] = cplx{(double)(ur[t][ct][r] + ar[t][ct][r]),
                                  (double)(ui[t][ct][r] + ai[t][ct][r])};
      }
}

template <int CT>
__global__ __launch_bounds__(1024 / bfs_rt(CT)) void taylor_bfs_kernel(
    int N, int n, int n_terms_k, cplx* __restrict__ walkers, const uint16_t* __restrict__ emi_hi,
    const uint16_t* __restrict__ emi_lo, const float* __restrict__ vhs, int nv) {
  taylor_bfs_body<CT, bfs_rt(CT)>(N, n, n_terms_k, walkers, emi_hi, emi_lo, vhs, nv,
                                  blockIdx.x);
}

template <int CT>
__global__ __launch_bounds__(1024 / bfs_rt(CT)) void taylor_bfs_grouped_kernel(
    const TaylorArgs* __restrict__ args, const int* __restrict__ start, int nprob) {
  const int p = group_of(start, nprob, blockIdx.x);
  const TaylorArgs a = args[p];
  taylor_bfs_body<CT, bfs_rt(CT)>(
      a.N, a.n, a.K, a.walkers, static_cast<const uint16_t*>(a.E),
      static_cast<const uint16_t*>(a.E2), static_cast<const float*>(a.vhs), a.nv,
      blockIdx.x - start[p]);
}

// Both spins of an unrestricted walker set that share one V (spin-summed Cholesky
// vectors) in one launch.  Block b = 16 q + 8 s + r takes spin s of walker 8 q + r:
// blocks are dealt round-robin to the 8 XCDs (block b on XCD b % 8), so the two blocks
// of a walker run on the same XCD, one dispatch round apart, and the second one's V
// rows mostly come from that XCD's L2 instead of HBM.
template <int CT>
__global__ __launch_bounds__(1024 / bfs_rt(CT)) void taylor_bfs_spins_kernel(
    int N, int n0, int n1, int n_terms_k, int W, cplx* __restrict__ w0, cplx* __restrict__ w1,
    const uint16_t* __restrict__ e0h, const uint16_t* __restrict__ e0l,
    const uint16_t* __restrict__ e1h, const uint16_t* __restrict__ e1l,
    const float* __restrict__ vhs, int nv) {
  const int b = blockIdx.x;
  const int s = (b >> 3) & 1, w = 8 * (b >> 4) + (b & 7);
  if (w >= W) return;  // (whole workgroup: the tail of the last group of 8 walkers)
  taylor_bfs_body<CT, bfs_rt(CT)>(N, s ? n1 : n0, n_terms_k, s ? w1 : w0, s ? e1h : e0h,
                                  s ? e1l : e0l, vhs, nv, w);
}

template <int CT>
int launch_bfs_spins(afqmc_ctx* c, int W, cplx* wup, cplx* wdn) {
  const int N = c->d.norb;
  if (N > 256 || c->nocc[0] > 16 * CT || c->nocc[1] > 16 * CT || 32 * ((N + 31) / 32) != c->Np32)
    return afqmc_set_error(AFQMC_ERR_ARG, "taylor bf16x3 streaming: N %d column tiles %d", N, CT);
  const size_t shm = bfs_shm(N, CT);
  static int lds_limit = 0;
  const int rc = ensure_dyn_lds((const void*)taylor_bfs_spins_kernel<CT>, shm, &lds_limit);
  if (rc) return rc;
  const int groups = (W + 7) / 8;
  hipLaunchKernelGGL((taylor_bfs_spins_kernel<CT>), dim3(16 * groups), dim3(bfs_threads(N, CT)),
                     shm, c->stream, N, c->nocc[0], c->nocc[1], c->d.n_exp_terms - 1, W, wup,
                     wdn, c->emi_bf[0][0], c->emi_bf[0][1], c->emi_bf[1][0], c->emi_bf[1][1],
                     vhs_f32_of(c, 0), c->vhs_nv);
  LAUNCH_CHECK();
  return AFQMC_OK;
}

template <int CT>
int launch_bfs(afqmc_ctx* c, int W, int n, cplx* wk, int s) {
  const int N = c->d.norb;
  if (N > 256 || n > 16 * CT || 32 * ((N + 31) / 32) != c->Np32)
    return afqmc_set_error(AFQMC_ERR_ARG, "taylor bf16x3 streaming: N %d n %d column tiles %d",
                           N, n, CT);
  const size_t shm = bfs_shm(N, CT);
  static int lds_limit = 0;
  const int rc = ensure_dyn_lds((const void*)taylor_bfs_kernel<CT>, shm, &lds_limit);
  if (rc) return rc;
  hipLaunchKernelGGL((taylor_bfs_kernel<CT>), dim3(W), dim3(bfs_threads(N, CT)), shm, c->stream,
                     N, n, c->d.n_exp_terms - 1, wk, c->emi_bf[s][0], c->emi_bf[s][1],
                     vhs_f32_of(c, s), c->vhs_nv);
  LAUNCH_CHECK();
  return AFQMC_OK;
}

template <int NKGV>
int launch_reg_ct(afqmc_ctx* c, int W, int n, cplx* wk, const double* E, const double* V) {
  const int ctp = (2 * n + 15) / 16;
  switch (ctp) {
    case 1: return launch_reg<NKGV, 1>(c, W, n, wk, E, V);
    case 2: return launch_reg<NKGV, 2>(c, W, n, wk, E, V);
    case 3: return launch_reg<NKGV, 3>(c, W, n, wk, E, V);
    default: return launch_reg<NKGV, 4>(c, W, n, wk, E, V);
  }
}

}  // namespace

// The kernel launch_taylor picks for spin s, as a plan for grouped launches.
void taylor_plan(const afqmc_ctx* c, int s, TaylorPlan* pl) {
  const int N = c->d.norb, n = c->nocc[s];
  const int Np = (N + 3) & ~3;
  const int nkgv = 2 * ((N + 3) / 4);
  const int bucket = nkgv <= 32 ? 32 : nkgv <= 48 ? 48 : nkgv <= 56 ? 56 : 64;
  const int ctp = (2 * n + 15) / 16 > 4 ? 4 : (2 * n + 15) / 16;
  pl->nthreads = ((N + 15) / 16) * 64;
  pl->blocks_per_walker = 1;
  pl->nkgv = 0;
  pl->ct = 0;
  if (c->taylor_bf) {
    if (N <= 128 && n <= 32) {
      const int kb = bf_kblocks(N);
      pl->kind = TAYLOR_REG_BF16X3;
      pl->nthreads = bf_threads(N);
      pl->nkgv = kb;
      pl->ct = ctp;
      pl->shm = (size_t)2 * ctp * 16 * bf_ldk(kb) * sizeof(uint16_t);
    } else {  // N <= AFQMC_MAX_NORB = 256, n <= AFQMC_MAX_NOCC = 64
      pl->kind = TAYLOR_STREAM_BF16X3;
      pl->ct = (n + 15) / 16;
      pl->nthreads = bfs_threads(N, pl->ct);
      pl->shm = bfs_shm(N, pl->ct);
    }
  } else if (N <= 128 && c->nocc[0] <= 32 && c->nocc[1] <= 32) {
    pl->kind = TAYLOR_REG_F64;
    pl->nkgv = bucket;
    pl->ct = ctp;
    const int rs = 16 * ctp + ((ctp % 2 == 0) ? 16 : 0);
    pl->shm = (size_t)2 * Np * rs * sizeof(double);
  } else {
    pl->kind = TAYLOR_STREAM_F64;
    pl->blocks_per_walker = (n + 15) / 16;
    pl->shm = (size_t)((N + 15) / 16) * 16 * 16 * 4 * sizeof(double);
  }
}

int launch_taylor(afqmc_ctx* c, int W, cplx* wup, cplx* wdn) {
  const int N = c->d.norb;
  if (c->taylor_bf && c->nspin == 2 && vhs_f32_of(c, 0) == vhs_f32_of(c, 1) &&
      !(N <= 128 && c->nocc[0] <= 32) && !(N <= 128 && c->nocc[1] <= 32) &&
      (c->nocc[0] + 15) / 16 == (c->nocc[1] + 15) / 16) {
    // both spins streaming with one shared V: one launch, the spins of a walker side by side
    switch ((c->nocc[0] + 15) / 16) {
      case 1: return launch_bfs_spins<1>(c, W, wup, wdn);
      case 2: return launch_bfs_spins<2>(c, W, wup, wdn);
      case 3: return launch_bfs_spins<3>(c, W, wup, wdn);
      default: return launch_bfs_spins<4>(c, W, wup, wdn);
    }
  }
  if (c->taylor_bf) {
    for (int s = 0; s < c->nspin; ++s) {
      const int n = c->nocc[s];
      cplx* wk = (s == 0) ? wup : wdn;
      int rc;
      if (N <= 128 && n <= 32) {
        switch (bf_kblocks(N)) {
          case 1: rc = launch_bf_ct<1>(c, W, n, wk, s); break;
          case 2: rc = launch_bf_ct<2>(c, W, n, wk, s); break;
          case 3: rc = launch_bf_ct<3>(c, W, n, wk, s); break;
          case 4: rc = launch_bf_ct<4>(c, W, n, wk, s); break;
          case 5: rc = launch_bf_ct<5>(c, W, n, wk, s); break;
          case 6: rc = launch_bf_ct<6>(c, W, n, wk, s); break;
          case 7: rc = launch_bf_ct<7>(c, W, n, wk, s); break;
          default: rc = launch_bf_ct<8>(c, W, n, wk, s); break;
        }
      } else {
        switch ((n + 15) / 16) {
          case 1: rc = launch_bfs<1>(c, W, n, wk, s); break;
          case 2: rc = launch_bfs<2>(c, W, n, wk, s); break;
          case 3: rc = launch_bfs<3>(c, W, n, wk, s); break;
          default: rc = launch_bfs<4>(c, W, n, wk, s); break;
        }
      }
      if (rc) return rc;
    }
    return AFQMC_OK;
  }
  if (N <= 128 && c->nocc[0] <= 32 && c->nocc[1] <= 32) {
    for (int s = 0; s < c->nspin; ++s) {
      const int n = c->nocc[s];
      cplx* wk = (s == 0) ? wup : wdn;
      // register budget: 2 ceil(N/4) k-groups of [Vr | Vi] per lane, bucketed
      const int nkgv = 2 * ((N + 3) / 4);
      const double* V = vhs_of(c, s);
      const int rc = nkgv <= 32   ? launch_reg_ct<32>(c, W, n, wk, c->exp_h1[s], V)
                     : nkgv <= 48 ? launch_reg_ct<48>(c, W, n, wk, c->exp_h1[s], V)
                     : nkgv <= 56 ? launch_reg_ct<56>(c, W, n, wk, c->exp_h1[s], V)
                                  : launch_reg_ct<64>(c, W, n, wk, c->exp_h1[s], V);
      if (rc) return rc;
    }
    return AFQMC_OK;
  }
  const int nwaves = (N + 15) / 16;
  const int rows_pad = nwaves * 16;
  const size_t shm = (size_t)rows_pad * 16 * 4 * sizeof(double);
  static int lds_limit = 0;
  const int lrc = ensure_dyn_lds((const void*)taylor_kernel, shm, &lds_limit);
  if (lrc) return lrc;
  for (int s = 0; s < c->nspin; ++s) {
    const int n = c->nocc[s];
    cplx* wk = (s == 0) ? wup : wdn;
    dim3 grid(W, (n + 15) / 16);
    hipLaunchKernelGGL(taylor_kernel, grid, dim3(nwaves * 64), shm, c->stream, N, n,
                       c->d.n_exp_terms - 1, wk, c->exp_h1[s], vhs_of(c, s), rows_pad);
    LAUNCH_CHECK();
  }
  return AFQMC_OK;
}

namespace {

template <typename KernelFn>
int launch_grouped_fn(KernelFn fn, hipStream_t st, const TaylorArgs* args, const int* start,
                      int nprob, int nblocks, int nthreads, size_t shm, int* lds_limit) {
  const int rc = ensure_dyn_lds((const void*)fn, shm, lds_limit);
  if (rc) return rc;
  hipLaunchKernelGGL(fn, dim3(nblocks), dim3(nthreads), shm, st, args, start, nprob);
  LAUNCH_CHECK();
  return AFQMC_OK;
}

template <int NKGV>
int reg_grouped(int ct, hipStream_t st, const TaylorArgs* args, const int* start, int nprob,
                int nblocks, int nthreads, size_t shm) {
  static int lim[4] = {};
  const int ci = ct < 1 ? 0 : (ct > 4 ? 3 : ct - 1);
  switch (ct) {
    case 1: return launch_grouped_fn(taylor_reg_grouped_kernel<NKGV, 1>, st, args, start, nprob, nblocks, nthreads, shm, &lim[ci]);
    case 2: return launch_grouped_fn(taylor_reg_grouped_kernel<NKGV, 2>, st, args, start, nprob, nblocks, nthreads, shm, &lim[ci]);
    case 3: return launch_grouped_fn(taylor_reg_grouped_kernel<NKGV, 3>, st, args, start, nprob, nblocks, nthreads, shm, &lim[ci]);
    case 4: return launch_grouped_fn(taylor_reg_grouped_kernel<NKGV, 4>, st, args, start, nprob, nblocks, nthreads, shm, &lim[ci]);
  }
  return afqmc_set_error(AFQMC_ERR_ARG, "grouped taylor: column tiles %d", ct);
}

}  // namespace

// Grouped Taylor launch for problems of one kernel kind and template key (nkgv bucket,
// column tiles ct) and one block size nthreads = 64 * ceil(N / 16).
int launch_taylor_grouped(hipStream_t st, int kind, int nkgv, int ct, const TaylorArgs* args,
                          const int* start, int nprob, int nblocks, int nthreads, size_t shm) {
  if (kind == TAYLOR_REG_F64) {
    switch (nkgv) {
      case 32: return reg_grouped<32>(ct, st, args, start, nprob, nblocks, nthreads, shm);
      case 48: return reg_grouped<48>(ct, st, args, start, nprob, nblocks, nthreads, shm);
      case 56: return reg_grouped<56>(ct, st, args, start, nprob, nblocks, nthreads, shm);
      case 64: return reg_grouped<64>(ct, st, args, start, nprob, nblocks, nthreads, shm);
      default: return afqmc_set_error(AFQMC_ERR_ARG, "grouped taylor: nkgv %d", nkgv);
    }
  }
  if (kind == TAYLOR_REG_BF16X3) {
    static int lim[8][4] = {};
    const int ci = ct < 1 ? 0 : (ct > 4 ? 3 : ct - 1);
#define BF_CASE(KB, C)                                                                      \
  if (nkgv == KB && ct == C)                                                                \
    return launch_grouped_fn(taylor_bf_grouped_kernel<KB, C>, st, args, start, nprob, nblocks, \
                             nthreads, shm, &lim[KB - 1][ci]);
#define BF_CASES(KB) BF_CASE(KB, 1) BF_CASE(KB, 2) BF_CASE(KB, 3) BF_CASE(KB, 4)
    BF_CASES(1) BF_CASES(2) BF_CASES(3) BF_CASES(4)
    BF_CASES(5) BF_CASES(6) BF_CASES(7) BF_CASES(8)
#undef BF_CASES
#undef BF_CASE
    return afqmc_set_error(AFQMC_ERR_ARG, "grouped taylor: k blocks %d, column tiles %d", nkgv, ct);
  }
  if (kind == TAYLOR_STREAM_BF16X3) {
    static int lim[4] = {};
    switch (ct) {
      case 1: return launch_grouped_fn(taylor_bfs_grouped_kernel<1>, st, args, start, nprob, nblocks, nthreads, shm, &lim[0]);
      case 2: return launch_grouped_fn(taylor_bfs_grouped_kernel<2>, st, args, start, nprob, nblocks, nthreads, shm, &lim[1]);
      case 3: return launch_grouped_fn(taylor_bfs_grouped_kernel<3>, st, args, start, nprob, nblocks, nthreads, shm, &lim[2]);
      case 4: return launch_grouped_fn(taylor_bfs_grouped_kernel<4>, st, args, start, nprob, nblocks, nthreads, shm, &lim[3]);
      default: return afqmc_set_error(AFQMC_ERR_ARG, "grouped taylor: column tiles %d", ct);
    }
  }
  static int lim64 = 0;
  return launch_grouped_fn(taylor_grouped_kernel, st, args, start, nprob, nblocks, nthreads, shm,
                           &lim64);
}
