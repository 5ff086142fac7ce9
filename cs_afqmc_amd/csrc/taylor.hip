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

__global__ __launch_bounds__(1024) void taylor_kernel(
    int N, int n, int n_terms_k, cplx* __restrict__ walkers, const double* __restrict__ expE,
    const double* __restrict__ vhs, int rows_pad) {
  extern __shared__ double T[];
  const int w = blockIdx.x, chunk = blockIdx.y;
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

}  // namespace

int launch_taylor(afqmc_ctx* c, int W, cplx* wup, cplx* wdn) {
  const int N = c->d.norb;
  const int nwaves = (N + 15) / 16;
  const int rows_pad = nwaves * 16;
  const size_t shm = (size_t)rows_pad * 16 * 4 * sizeof(double);
  if (shm > 64 * 1024) {
    static bool attr_set = false;
    if (!attr_set) {
      HIP_TRY(hipFuncSetAttribute((const void*)taylor_kernel,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
      attr_set = true;
    }
  }
  for (int s = 0; s < c->nspin; ++s) {
    const int n = c->nocc[s];
    cplx* wk = (s == 0) ? wup : wdn;
    dim3 grid(W, (n + 15) / 16);
    hipLaunchKernelGGL(taylor_kernel, grid, dim3(nwaves * 64), shm, c->stream, N, n,
                       c->d.n_exp_terms - 1, wk, c->exp_h1[s], c->vhs, rows_pad);
    LAUNCH_CHECK();
  }
  return AFQMC_OK;
}
