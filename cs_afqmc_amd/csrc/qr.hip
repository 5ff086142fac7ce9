// Walker re-orthonormalisation: reduced Householder QR, walker <- Q.
// Follows linalg_utils.qr_vmap / qr_vmap_uhf (linalg_utils.py:51-65), i.e.
// jnp.linalg.qr, whose CPU lowering is LAPACK geqrf + orgqr/ungqr.  For the sizes
// here (n <= 64 < the blocked crossover) those are the unblocked zgeqr2 + zung2r:
//   zlarfg: beta = -sign(Re alpha) * |(alpha, x)|, tau = ((beta - Re a)/beta, -Im a/beta),
//           v = [1; x / (alpha - beta)], R[i][i] = beta (real)
//   geqr2 applies H(i)^H = I - conj(tau) v v^H to the trailing columns;
//   ung2r accumulates Q = H(0) ... H(n-1) backwards, in place.
// One workgroup of four waves per (walker, spin).  The walker is copied to LDS
// (column-major) when it fits, else it is factorised in place in global memory.
#include "common.h"

extern "C" __device__ double __ockl_wfred_add_f64(double);

namespace {

constexpr int QR_NT = 256;
constexpr int QR_WAVES = QR_NT / 64;
constexpr int QR_RPL = (AFQMC_MAX_NORB + 63) / 64;  // rows per lane
// dynamic LDS for the walker copy: 160 KiB per workgroup less the static arrays below
constexpr size_t QR_LDS_LIMIT = 156 * 1024;
// column stride of the LDS copy (odd, so 16-byte accesses of consecutive rows of one
// column and of one row's consecutive columns spread over the banks)
__host__ __device__ inline int qr_ldc(int m) { return m | 1; }

// The walker A (m x n) is addressed as A[r * ldr + j * ldc]: column-major in LDS, or
// row-major in place in global memory.  Column j belongs to wave j % 4, which applies
// every reflector to it, so one barrier per reflector orders the reflector's vector
// before the other waves read it:
//   zgeqr2, step i: the owner of column i forms the reflector (norm by a wave reduction;
//     beta, tau and the scaling are computed by all its lanes alike), then after the
//     barrier each wave applies H(i)^H = I - conj(tau) v v^H to its columns j > i.
//   zung2r, step i (backwards): each wave applies H(i) to its columns j > i, then after
//     the barrier the owner of column i overwrites v with column i of Q.
// A wave's column updates of one step are ordered before its own work on the next step,
// and no wave touches another wave's columns.
__device__ __forceinline__ void qr_apply(cplx* A, int ldr, int ldc, int m, int n, int i,
                                         cplx tc, int lane, int wave) {
  // v_r for the lane's rows r = i + lane + 64 k (v_i = 1; rows past m contribute 0),
  // loads from clamped rows, masked after
  cplx v[QR_RPL];
#pragma unroll
  for (int k = 0; k < QR_RPL; ++k) {
    const int r = i + lane + 64 * k;
    const cplx x = A[(int64_t)min(r, m - 1) * ldr + (int64_t)i * ldc];
    v[k] = __builtin_unpredictable(r < m) ? (r == i ? cplx{1.0, 0.0} : x) : cplx{0.0, 0.0};
  }
  const int j0 = i + 1 + ((wave - (i + 1)) % QR_WAVES + QR_WAVES) % QR_WAVES;
  for (int j = j0; j < n; j += QR_WAVES) {
    cplx a[QR_RPL];
    double sr = 0.0, si = 0.0;
#pragma unroll
    for (int k = 0; k < QR_RPL; ++k) {
      const int r = i + lane + 64 * k;
      a[k] = A[(int64_t)min(r, m - 1) * ldr + (int64_t)j * ldc];
      // conj(v) * a (v is zero past m)
      sr += v[k].re * a[k].re + v[k].im * a[k].im;
      si += v[k].re * a[k].im - v[k].im * a[k].re;
    }
    const cplx sv = cplx{__ockl_wfred_add_f64(sr), __ockl_wfred_add_f64(si)};
    const cplx ts = cmul(tc, sv);
#pragma unroll
    for (int k = 0; k < QR_RPL; ++k) {
      const int r = i + lane + 64 * k;
      if (r < m) A[(int64_t)r * ldr + (int64_t)j * ldc] = csub(a[k], cmul(v[k], ts));
    }
  }
}

// lds: A in LDS.  In global memory a barrier (a workgroup-scope fence) also follows
// every reflector's application, so the lanes of the next step read the updated values.
__device__ void qr_body(cplx* A, int ldr, int ldc, int m, int n, cplx* tau, bool lds) {
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  auto at = [&](int r, int j) -> cplx& { return A[(int64_t)r * ldr + (int64_t)j * ldc]; };
  // ---- zgeqr2
  for (int i = 0; i < n; ++i) {
    if (wave == i % QR_WAVES) {
      // zlarfg on column i: beta = -sign(Re alpha) |(alpha, x)|, tau = ((beta - Re alpha)
      // / beta, -Im alpha / beta), x <- x / (alpha - beta)
      double part = 0.0;
      for (int r = i + 1 + lane; r < m; r += 64) {
        const cplx x = at(r, i);
        part += x.re * x.re + x.im * x.im;
      }
      const double xnorm = sqrt(__ockl_wfred_add_f64(part));
      const cplx alpha = at(i, i);
      cplx t = cplx{0.0, 0.0}, scal = cplx{1.0, 0.0};
      const bool refl = !(xnorm == 0.0 && alpha.im == 0.0);
      double beta = 0.0;
      if (refl) {
        // |(alpha, xnorm)| scaled by the largest magnitude (dlapy3)
        const double xa = fabs(alpha.re), ya = fabs(alpha.im);
        const double w = fmax(xa, fmax(ya, xnorm));
        const double iw = 1.0 / w;
        const double a = xa * iw, b = ya * iw, c = xnorm * iw;
        const double nrm = w * sqrt(a * a + b * b + c * c);
        beta = alpha.re >= 0.0 ? -nrm : nrm;
        const double ib = 1.0 / beta;
        t = cplx{(beta - alpha.re) * ib, -alpha.im * ib};
        const cplx d = cplx{alpha.re - beta, alpha.im};
        const double id = 1.0 / (d.re * d.re + d.im * d.im);
        scal = cplx{d.re * id, -d.im * id};
      }
      for (int r = i + 1 + lane; r < m; r += 64) at(r, i) = cmul(at(r, i), scal);
      if (lane == 0) {
        if (refl) at(i, i) = cplx{beta, 0.0};
        tau[i] = t;
      }
    }
    __syncthreads();
    if (i + 1 < n) qr_apply(A, ldr, ldc, m, n, i, cconj(tau[i]), lane, wave);
    if (!lds) __syncthreads();
  }
  // ---- zung2r: Q = H(0) ... H(n-1) applied to the first n columns of I, in place
  for (int i = n - 1; i >= 0; --i) {
    const cplx t = tau[i];
    if (i + 1 < n) qr_apply(A, ldr, ldc, m, n, i, t, lane, wave);
    __syncthreads();
    if (wave == i % QR_WAVES) {
      const cplx mt = cplx{-t.re, -t.im};
      for (int r = i + 1 + lane; r < m; r += 64) at(r, i) = cmul(mt, at(r, i));
      for (int l = lane; l < i; l += 64) at(l, i) = cplx{0.0, 0.0};
      if (lane == 0) at(i, i) = cplx{1.0 - t.re, -t.im};
    }
  }
  __syncthreads();
}

// one walker (walker pointer g, m x n): LDS copy when use_lds, else in place
__device__ __forceinline__ void qr_walker(int m, int n, cplx* __restrict__ g, int use_lds,
                                          double* smem, cplx* tau) {
  if (use_lds) {
    // column-major copy: element (r, j) at A[j * ldc + r]
    const int ldc = qr_ldc(m);
    cplx* A = reinterpret_cast<cplx*>(smem);
    for (int e = threadIdx.x; e < m * n; e += QR_NT) {
      const int r = e / n, j = e - r * n;
      A[j * ldc + r] = g[e];
    }
    __syncthreads();
    qr_body(A, 1, ldc, m, n, tau, true);
    for (int e = threadIdx.x; e < m * n; e += QR_NT) {
      const int r = e / n, j = e - r * n;
      g[e] = A[j * ldc + r];
    }
  } else {
    qr_body(g, n, 1, m, n, tau, false);
  }
}

__global__ __launch_bounds__(QR_NT) void qr_kernel(int m, int n, cplx* __restrict__ walkers,
                                                   int use_lds) {
  extern __shared__ double smem[];
  __shared__ cplx tau[AFQMC_MAX_NOCC];
  qr_walker(m, n, walkers + (int64_t)blockIdx.x * m * n, use_lds, smem, tau);
}

__global__ __launch_bounds__(QR_NT) void qr_grouped_kernel(const QrArgs* __restrict__ args,
                                                           const int* __restrict__ start,
                                                           int nprob) {
  extern __shared__ double smem[];
  __shared__ cplx tau[AFQMC_MAX_NOCC];
  const int p = group_of(start, nprob, blockIdx.x);
  const QrArgs a = args[p];
  qr_walker(a.m, a.n, a.walkers + (int64_t)(blockIdx.x - start[p]) * a.m * a.n, a.use_lds,
            smem, tau);
}

}  // namespace

size_t qr_lds_bytes(int m, int n) {
  const size_t bytes = (size_t)qr_ldc(m) * n * sizeof(cplx);
  return bytes <= QR_LDS_LIMIT ? bytes : 0;
}

int launch_qr_grouped(hipStream_t st, const QrArgs* args, const int* start, int nprob,
                      int nblocks, size_t shm) {
  static int lds_limit = 0;
  const int rc = ensure_dyn_lds((const void*)qr_grouped_kernel, shm, &lds_limit);
  if (rc) return rc;
  hipLaunchKernelGGL(qr_grouped_kernel, dim3(nblocks), dim3(QR_NT), shm, st, args, start, nprob);
  LAUNCH_CHECK();
  return AFQMC_OK;
}

int launch_qr(afqmc_ctx* c, int W, cplx* wup, cplx* wdn) {
  const int N = c->d.norb;
  for (int s = 0; s < c->nspin; ++s) {
    const int n = c->nocc[s];
    const size_t bytes = (size_t)qr_ldc(N) * n * sizeof(cplx);
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
