// Trial measurements per walker and spin: O = C^T w (n x n), det(O) by Gauss-Jordan
// elimination with partial pivoting (pivots as in LAPACK getrf: first max of
// |re| + |im|), O^-1, and the half-rotated Green's function G = (w O^-1)^T (n x N).
//
// Follows rhf._calc_overlap_restricted / _calc_green (wavefunctions.py:478-501) and
// uhf._calc_overlap / _calc_green (wavefunctions.py:658-692).  The trial C is real.
// One workgroup per (walker, spin).  Rows of the walker and of the trial are staged
// through LDS in chunks of R rows (coalesced global reads); O and its inverse live in
// LDS as the augmented matrix [O | I].
//
// Green output layout: green[w][0][i*N + p] = Re G[i][p], green[w][1][...] = Im,
// i.e. a row-major (2W, n*N) real matrix -- the A operand of the force-bias GEMM.
#include "common.h"

namespace {

constexpr int MEAS_NT = 256;
constexpr int MEAS_STAGE_BYTES = 48 * 1024;  // walker + trial row chunk
constexpr int MEAS_MAX_OUT = (AFQMC_MAX_NOCC * AFQMC_MAX_NOCC + MEAS_NT - 1) / MEAS_NT;

// rows per staged chunk: at most 48 KiB of walker + trial rows, and the whole LDS
// footprint (augmented matrix included) within 150 KiB
__host__ __device__ inline int meas_rows(int N, int n) {
  int budget = 150 * 1024 - (2 * n * n + n) * 16;
  if (budget > MEAS_STAGE_BYTES) budget = MEAS_STAGE_BYTES;
  int r = budget / (n * 24);
  if (r < 1) r = 1;
  return r < N ? r : N;
}

__global__ __launch_bounds__(MEAS_NT) void measure_kernel(
    int N, int n, const cplx* __restrict__ walkers, const double* __restrict__ trial,
    cplx* __restrict__ dets, int spin, double* __restrict__ green, int want_green) {
  extern __shared__ double smem[];
  const int n2 = 2 * n;
  const int R = meas_rows(N, n);
  cplx* M = reinterpret_cast<cplx*>(smem);  // n x 2n augmented matrix
  cplx* fcol = M + n * n2;                  // pivot column copy (n)
  cplx* ph = fcol + n;                      // staged walker rows (R x n)
  double* tr = reinterpret_cast<double*>(ph + R * n);  // staged trial rows (R x n)
  __shared__ int piv_s;

  const int w = blockIdx.x, tid = threadIdx.x;
  const cplx* phi = walkers + (int64_t)w * N * n;

  // ---- O = C^T phi, accumulated over row chunks; thread owns entries e = tid + k*NT
  double accr[MEAS_MAX_OUT], acci[MEAS_MAX_OUT];
#pragma unroll
  for (int k = 0; k < MEAS_MAX_OUT; ++k) accr[k] = acci[k] = 0.0;
  for (int p0 = 0; p0 < N; p0 += R) {
    const int rows = (N - p0 < R) ? N - p0 : R;
    for (int e = tid; e < rows * n; e += MEAS_NT) {
      ph[e] = phi[(int64_t)p0 * n + e];
      tr[e] = trial[(int64_t)p0 * n + e];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < MEAS_MAX_OUT; ++k) {
      const int e = tid + k * MEAS_NT;
      if (e < n * n) {
        const int i = e / n, j = e - (e / n) * n;
        double sr = accr[k], si = acci[k];
        for (int p = 0; p < rows; ++p) {
          const double a = tr[p * n + i];
          const cplx b = ph[p * n + j];
          sr = fma(a, b.re, sr);
          si = fma(a, b.im, si);
        }
        accr[k] = sr;
        acci[k] = si;
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int k = 0; k < MEAS_MAX_OUT; ++k) {
    const int e = tid + k * MEAS_NT;
    if (e < n * n) {
      const int i = e / n, j = e - (e / n) * n;
      M[i * n2 + j] = cplx{accr[k], acci[k]};
      M[i * n2 + n + j] = cplx{i == j ? 1.0 : 0.0, 0.0};
    }
  }
  __syncthreads();

  // ---- Gauss-Jordan with partial pivoting: [O | I] -> [I | O^-1], det from pivots
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
      for (int c = k + tid; c < n2; c += MEAS_NT) {
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

  // ---- G[i][p] = sum_j phi[p][j] * Oinv[j][i], walker rows staged again
  double* gre = green + (int64_t)w * 2 * n * N;
  double* gim = gre + (int64_t)n * N;
  for (int p0 = 0; p0 < N; p0 += R) {
    const int rows = (N - p0 < R) ? N - p0 : R;
    for (int e = tid; e < rows * n; e += MEAS_NT) ph[e] = phi[(int64_t)p0 * n + e];
    __syncthreads();
    for (int e = tid; e < n * rows; e += MEAS_NT) {
      const int i = e / rows, p = e - (e / rows) * rows;
      double sr = 0.0, si = 0.0;
      for (int j = 0; j < n; ++j) {
        const cplx a = ph[p * n + j];
        const cplx b = M[j * n2 + n + i];
        sr += a.re * b.re - a.im * b.im;
        si += a.re * b.im + a.im * b.re;
      }
      gre[(int64_t)i * N + p0 + p] = sr;
      gim[(int64_t)i * N + p0 + p] = si;
    }
    __syncthreads();
  }
}

}  // namespace

int launch_measure(afqmc_ctx* c, int W, const cplx* wup, const cplx* wdn, bool want_green) {
  const int N = c->d.norb;
  for (int s = 0; s < c->nspin; ++s) {
    const int n = c->nocc[s];
    const int R = meas_rows(N, n);
    const size_t shm = (size_t)(n * 2 * n + n + R * n) * sizeof(cplx) + (size_t)R * n * 8;
    const cplx* wk = (s == 0) ? wup : wdn;
    static int lds_limit = 0;
    const int rc = ensure_dyn_lds((const void*)measure_kernel, shm, &lds_limit);
    if (rc) return rc;
    hipLaunchKernelGGL(measure_kernel, dim3(W), dim3(MEAS_NT), shm, c->stream, N, n, wk,
                       c->trial[s], c->dets, s, c->green[s], want_green ? 1 : 0);
    LAUNCH_CHECK();
  }
  return AFQMC_OK;
}
