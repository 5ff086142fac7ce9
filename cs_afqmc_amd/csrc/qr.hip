// Walker re-orthonormalisation: reduced Householder QR, walker <- Q.
// Follows linalg_utils.qr_vmap / qr_vmap_uhf (linalg_utils.py:51-65), i.e.
// jnp.linalg.qr, whose CPU lowering is LAPACK geqrf + orgqr/ungqr.  For the sizes
// here (n <= 64 < the blocked crossover) those are the unblocked zgeqr2 + zung2r:
//   zlarfg: beta = -sign(Re alpha) * |(alpha, x)|, tau = ((beta - Re a)/beta, -Im a/beta),
//           v = [1; x / (alpha - beta)], R[i][i] = beta (real)
//   geqr2 applies H(i)^H = I - conj(tau) v v^H to the trailing columns;
//   ung2r accumulates Q = H(0) ... H(n-1) backwards, in place.
// One workgroup per (walker, spin).  The walker is copied to LDS when it fits, else
// it is factorised in place in global memory.
#include "common.h"

namespace {

constexpr int QR_NT = 256;
// dynamic LDS for the walker copy: 160 KiB per workgroup less the static arrays below
constexpr size_t QR_LDS_LIMIT = 156 * 1024;

__device__ inline double dlapy3(double x, double y, double z) {
  const double xa = fabs(x), ya = fabs(y), za = fabs(z);
  const double w = fmax(xa, fmax(ya, za));
  if (w == 0.0) return xa + ya + za;
  const double a = xa / w, b = ya / w, c = za / w;
  return w * sqrt(a * a + b * b + c * c);
}

// w_j = sum_{r >= i} conj(v_r) A[r][j] for j in (i, n); one wave per column
__device__ inline void reflect_dot(cplx* A, int m, int n, int i, cplx* wvec) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int j = i + 1 + wave; j < n; j += QR_NT / 64) {
    double sr = 0.0, si = 0.0;
    for (int r = i + lane; r < m; r += 64) {
      const cplx v = (r == i) ? cplx{1.0, 0.0} : A[r * n + i];
      const cplx a = A[r * n + j];
      // conj(v) * a
      sr += v.re * a.re + v.im * a.im;
      si += v.re * a.im - v.im * a.re;
    }
    for (int off = 32; off > 0; off >>= 1) {
      sr += __shfl_down(sr, off, 64);
      si += __shfl_down(si, off, 64);
    }
    if (lane == 0) wvec[j] = cplx{sr, si};
  }
}

// A[r][j] -= t * v_r * w_j for r >= i, j in (i, n)
__device__ inline void reflect_update(cplx* A, int m, int n, int i, cplx t, const cplx* wvec) {
  const int width = n - i - 1;
  if (width <= 0) return;
  const int total = (m - i) * width;
  for (int e = threadIdx.x; e < total; e += QR_NT) {
    const int r = i + e / width, j = i + 1 + (e - (e / width) * width);
    const cplx v = (r == i) ? cplx{1.0, 0.0} : A[r * n + i];
    const cplx tv = cmul(t, v);
    A[r * n + j] = csub(A[r * n + j], cmul(tv, wvec[j]));
  }
}

__device__ void qr_body(cplx* A, int m, int n, cplx* tau, cplx* wvec, double* red,
                        cplx* bcast) {
  // ---- zgeqr2
  for (int i = 0; i < n; ++i) {
    double part = 0.0;
    for (int r = i + 1 + threadIdx.x; r < m; r += QR_NT) {
      const cplx x = A[r * n + i];
      part += x.re * x.re + x.im * x.im;
    }
    const double xnorm = sqrt(block_sum<QR_NT>(part, red));
    if (threadIdx.x == 0) {
      const cplx alpha = A[i * n + i];
      cplx t = cplx{0.0, 0.0}, scal = cplx{1.0, 0.0};
      if (!(xnorm == 0.0 && alpha.im == 0.0)) {
        const double nrm = dlapy3(alpha.re, alpha.im, xnorm);
        const double beta = (alpha.re >= 0.0) ? -nrm : nrm;
        t = cplx{(beta - alpha.re) / beta, -alpha.im / beta};
        scal = cdiv(cplx{1.0, 0.0}, cplx{alpha.re - beta, alpha.im});
        A[i * n + i] = cplx{beta, 0.0};
      }
      tau[i] = t;
      bcast[0] = scal;
    }
    __syncthreads();
    const cplx scal = bcast[0];
    for (int r = i + 1 + threadIdx.x; r < m; r += QR_NT) A[r * n + i] = cmul(A[r * n + i], scal);
    __syncthreads();
    if (i < n - 1) {
      reflect_dot(A, m, n, i, wvec);
      __syncthreads();
      reflect_update(A, m, n, i, cconj(tau[i]), wvec);
      __syncthreads();
    }
  }
  // ---- zung2r
  for (int i = n - 1; i >= 0; --i) {
    const cplx t = tau[i];
    if (i < n - 1) {
      reflect_dot(A, m, n, i, wvec);
      __syncthreads();
      reflect_update(A, m, n, i, t, wvec);
      __syncthreads();
    }
    const cplx mt = cplx{-t.re, -t.im};
    for (int r = i + 1 + threadIdx.x; r < m; r += QR_NT) A[r * n + i] = cmul(mt, A[r * n + i]);
    for (int l = threadIdx.x; l < i; l += QR_NT) A[l * n + i] = cplx{0.0, 0.0};
    if (threadIdx.x == 0) A[i * n + i] = cplx{1.0 - t.re, -t.im};
    __syncthreads();
  }
}

__global__ __launch_bounds__(QR_NT) void qr_kernel(int m, int n, cplx* __restrict__ walkers,
                                                   int use_lds) {
  extern __shared__ double smem[];
  __shared__ cplx tau[AFQMC_MAX_NOCC];
  __shared__ cplx wvec[AFQMC_MAX_NOCC];
  __shared__ double red[QR_NT / 64];
  __shared__ cplx bcast[1];
  cplx* g = walkers + (int64_t)blockIdx.x * m * n;
  cplx* A = g;
  if (use_lds) {
    A = reinterpret_cast<cplx*>(smem);
    for (int e = threadIdx.x; e < m * n; e += QR_NT) A[e] = g[e];
    __syncthreads();
  }
  qr_body(A, m, n, tau, wvec, red, bcast);
  if (use_lds) {
    for (int e = threadIdx.x; e < m * n; e += QR_NT) g[e] = A[e];
  }
}

}  // namespace

int launch_qr(afqmc_ctx* c, int W, cplx* wup, cplx* wdn) {
  const int N = c->d.norb;
  for (int s = 0; s < c->nspin; ++s) {
    const int n = c->nocc[s];
    const size_t bytes = (size_t)N * n * sizeof(cplx);
    const int use_lds = bytes <= QR_LDS_LIMIT ? 1 : 0;
    const size_t shm = use_lds ? bytes : 0;
    static int lds_limit = 0;
    const int rc = ensure_dyn_lds((const void*)qr_kernel, shm, &lds_limit);
    if (rc) return rc;
    hipLaunchKernelGGL(qr_kernel, dim3(W), dim3(QR_NT), shm, c->stream, N, n,
                       (s == 0) ? wup : wdn, use_lds);
    LAUNCH_CHECK();
  }
  return AFQMC_OK;
}
