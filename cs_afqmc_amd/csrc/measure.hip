// Trial measurements per walker and spin: O = C^T w (n x n), det(O) by Gauss-Jordan
// elimination with partial pivoting (pivots as in LAPACK getrf: first max of
// |re| + |im|), O^-1, and the half-rotated Green's function G = (w O^-1)^T (n x N).
//
// Follows rhf._calc_overlap_restricted / _calc_green (wavefunctions.py:478-501) and
// uhf._calc_overlap / _calc_green (wavefunctions.py:658-692).  The trial C is real.
// One workgroup per (walker, spin); O and its inverse live in LDS as [O | I].
//
// Green output layout: green[w][0][i*N + p] = Re G[i][p], green[w][1][...] = Im,
// i.e. a row-major (2W, n*N) real matrix -- the A operand of the force-bias GEMM.
#include "common.h"

namespace {

constexpr int MEAS_NT = 256;

__global__ __launch_bounds__(MEAS_NT) void measure_kernel(
    int N, int n, const cplx* __restrict__ walkers, const double* __restrict__ trial,
    cplx* __restrict__ dets, int spin, double* __restrict__ green, int want_green) {
  extern __shared__ double smem[];
  const int n2 = 2 * n;
  cplx* M = reinterpret_cast<cplx*>(smem);  // n x 2n augmented matrix
  cplx* fcol = M + n * n2;                  // pivot column copy (n)
  __shared__ int piv_s;

  const int w = blockIdx.x, tid = threadIdx.x;
  const cplx* phi = walkers + (int64_t)w * N * n;

  // O = C^T phi (left half), identity (right half)
  for (int e = tid; e < n * n; e += MEAS_NT) {
    const int i = e / n, j = e - (e / n) * n;
    double sr = 0.0, si = 0.0;
    for (int p = 0; p < N; ++p) {
      const double a = trial[p * n + i];
      const cplx b = phi[p * n + j];
      sr = fma(a, b.re, sr);
      si = fma(a, b.im, si);
    }
    M[i * n2 + j] = cplx{sr, si};
    M[i * n2 + n + j] = cplx{i == j ? 1.0 : 0.0, 0.0};
  }
  __syncthreads();

  cplx det = cplx{1.0, 0.0};
  for (int k = 0; k < n; ++k) {
    if (tid < 64) {
      const int r = k + tid;
      double v = (r < n) ? cabs1(M[r * n2 + k]) : -1.0;
      int idx = r;
      for (int off = 32; off > 0; off >>= 1) {
        const double v2 = __shfl_xor(v, off, 64);
        const int i2 = __shfl_xor(idx, off, 64);
        if (v2 > v || (v2 == v && i2 < idx)) {
          v = v2;
          idx = i2;
        }
      }
      if (tid == 0) piv_s = idx;
    }
    __syncthreads();
    const int pr = piv_s;
    if (pr != k) {
      for (int c = tid; c < n2; c += MEAS_NT) {
        const cplx a = M[k * n2 + c];
        M[k * n2 + c] = M[pr * n2 + c];
        M[pr * n2 + c] = a;
      }
    }
    __syncthreads();
    const cplx pivot = M[k * n2 + k];
    if (tid == 0) {
      det = cmul(det, pivot);
      if (pr != k) det = cplx{-det.re, -det.im};
    }
    for (int r = tid; r < n; r += MEAS_NT) fcol[r] = M[r * n2 + k];
    __syncthreads();
    const cplx inv = cdiv(cplx{1.0, 0.0}, pivot);
    for (int c = k + 1 + tid; c < n2; c += MEAS_NT) M[k * n2 + c] = cmul(M[k * n2 + c], inv);
    __syncthreads();
    const int width = n2 - k - 1;
    for (int e = tid; e < n * width; e += MEAS_NT) {
      const int r = e / width, c = k + 1 + (e - (e / width) * width);
      if (r != k) {
        const cplx f = fcol[r], m = M[k * n2 + c];
        cplx x = M[r * n2 + c];
        x.re -= f.re * m.re - f.im * m.im;
        x.im -= f.re * m.im + f.im * m.re;
        M[r * n2 + c] = x;
      }
    }
    __syncthreads();
  }
  if (tid == 0) dets[(int64_t)w * 2 + spin] = det;
  if (!want_green) return;

  // G[i][p] = sum_j phi[p][j] * Oinv[j][i]
  double* gre = green + (int64_t)w * 2 * n * N;
  double* gim = gre + (int64_t)n * N;
  for (int e = tid; e < n * N; e += MEAS_NT) {
    const int i = e / N, p = e - (e / N) * N;
    double sr = 0.0, si = 0.0;
    for (int j = 0; j < n; ++j) {
      const cplx a = phi[p * n + j];
      const cplx b = M[j * n2 + n + i];
      sr += a.re * b.re - a.im * b.im;
      si += a.re * b.im + a.im * b.re;
    }
    gre[e] = sr;
    gim[e] = si;
  }
}

}  // namespace

int launch_measure(afqmc_ctx* c, int W, const cplx* wup, const cplx* wdn, bool want_green) {
  const int N = c->d.norb;
  for (int s = 0; s < c->nspin; ++s) {
    const int n = c->nocc[s];
    const size_t shm = (size_t)(n * 2 * n + n) * sizeof(cplx);
    const cplx* wk = (s == 0) ? wup : wdn;
    if (shm > 64 * 1024) {
      static bool attr_set = false;
      if (!attr_set) {
        HIP_TRY(hipFuncSetAttribute((const void*)measure_kernel,
                                    hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
        attr_set = true;
      }
    }
    hipLaunchKernelGGL(measure_kernel, dim3(W), dim3(MEAS_NT), shm, c->stream, N, n, wk,
                       c->trial[s], c->dets, s, c->green[s], want_green ? 1 : 0);
    LAUNCH_CHECK();
  }
  return AFQMC_OK;
}
