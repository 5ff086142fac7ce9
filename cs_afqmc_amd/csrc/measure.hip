// Trial measurements per walker and spin: O = C^T w (n x n), det(O) by Gauss-Jordan
// elimination with partial pivoting (pivots as in LAPACK getrf: first max of
// |re| + |im|), X = O^-1, and the half-rotated Green's function G = (w X)^T (n x N).
//
// Follows rhf._calc_overlap_restricted / _calc_green (wavefunctions.py:478-501) and
// uhf._calc_overlap / _calc_green (wavefunctions.py:658-692).  The trial C is real.
//
// One workgroup per (walker, spin).  For n <= 24 it has two waves: they split the k
// chunks of the overlap product (partials summed in a fixed order through LDS) and the
// column tiles of the Green's function, and wave 0 runs the register Gauss-Jordan while
// both waves' first Green operands are in flight; with one walker per SIMD (W = 4 x CUs)
// this halves the chains of dependent global loads each walker waits on.  Wider systems
// use one wave and the LDS elimination.  Both products run on
// v_mfma_f64_16x16x4f64 in real block form:
//   [Or | Oi]  = C^T [wr | wi]                         (n x 2n,  K = N)
//   [Gr ; Gi]  = [[Xr^T, -Xi^T], [Xi^T, Xr^T]] [wr^T ; wi^T]   (2n x N, K = 2n)
// The second product's rows are exactly the rows of the split Green layout
// green[w][c][i][p] (c = re / im), so results go straight to global memory.
// [O | I] lives in LDS (n x 2n complex).
#include <algorithm>

#include "common.h"

extern "C" __device__ unsigned long long __ockl_wfred_max_u64(unsigned long long);

namespace {

constexpr int MEAS_KC = 4;  // k-groups whose operands are loaded ahead of their MFMAs

__device__ inline double readlane_f64(double x, int l) {
  const unsigned long long u = (unsigned long long)__double_as_longlong(x);
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)u, l);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(u >> 32), l);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// Gauss-Jordan on [O | I] (n x 2n, n <= 8 T2 <= 24) held in registers: lane c keeps
// column c.  Pivot k: lane k finds the first maximum of |re| + |im| in its column
// (rows >= k), the row exchange and the pivot-row scaling are per-lane register
// operations, and each other row r is updated with the multiplier M[r][k] read from
// lane k (v_readlane).  No LDS traffic and no barriers inside the elimination.
// Returns det(O); on exit M (LDS) holds [* | O^-1].
template <int T2>
__device__ __forceinline__ cplx gauss_jordan_registers(cplx* M, int n, int lane) {
  constexpr int NR = 8 * T2;
  const int n2 = 2 * n;
  double vr[NR], vi[NR];
#pragma unroll
  for (int r = 0; r < NR; ++r) {
    cplx x = cplx{0.0, 0.0};
    if (r < n && lane < n2) x = M[r * n2 + lane];
    vr[r] = x.re;
    vi[r] = x.im;
  }
  cplx det = cplx{1.0, 0.0};
#pragma unroll
  for (int k = 0; k < NR; ++k) {
    if (k < n) {
      int best = k;
      double bv = -1.0;
#pragma unroll
      for (int r = k; r < NR; ++r) {
        const double a = fabs(vr[r]) + fabs(vi[r]);
        if (r < n && a > bv) {
          bv = a;
          best = r;
        }
      }
      const int pr = __builtin_amdgcn_readlane(best, k);
      if (pr != k) {
#pragma unroll
        for (int r = k + 1; r < NR; ++r) {
          if (r == pr) {
            const double tr = vr[r], ti = vi[r];
            vr[r] = vr[k];
            vi[r] = vi[k];
            vr[k] = tr;
            vi[k] = ti;
          }
        }
      }
      const cplx piv = cplx{readlane_f64(vr[k], k), readlane_f64(vi[k], k)};
      det = cmul(det, piv);
      if (pr != k) det = cplx{-det.re, -det.im};
      const double d2 = 1.0 / (piv.re * piv.re + piv.im * piv.im);
      const cplx inv = cplx{piv.re * d2, -piv.im * d2};
      const double mr = vr[k] * inv.re - vi[k] * inv.im;
      const double mi = vr[k] * inv.im + vi[k] * inv.re;
      vr[k] = mr;
      vi[k] = mi;
#pragma unroll
      for (int r = 0; r < NR; ++r) {
        if (r != k && r < n) {
          // x -= f m as four fused multiply-adds
          const double fr = readlane_f64(vr[r], k), fi = readlane_f64(vi[r], k);
          vr[r] = fma(-fr, mr, fma(fi, mi, vr[r]));
          vi[r] = fma(-fr, mi, fma(-fi, mr, vi[r]));
        }
      }
    }
  }
#pragma unroll
  for (int r = 0; r < NR; ++r)
    if (r < n && lane < n2) M[r * n2 + lane] = cplx{vr[r], vi[r]};
  return det;
}

// waves per walker (two for n <= 24, four beyond) and the LDS the two-wave kernel's
// cross-wave overlap reduction needs
constexpr int meas_waves(int t2) { return t2 <= 3 ? 2 : 4; }
// waves per SIMD the register budget is compiled for (n > 48 needs more than 256)
constexpr int meas_wpe(int t2) { return t2 >= 7 ? 1 : 2; }
constexpr int meas_ti(int t2) { return t2 > 3 ? 1 : (t2 + 1) / 2; }
constexpr size_t meas_red_bytes(int t2) {
  return t2 <= 3 ? (size_t)meas_ti(t2) * t2 * 4 * 64 * sizeof(double) : 0;
}
// walkers up to this size are staged in LDS by the two-wave kernel (four blocks per CU)
constexpr size_t MEAS_STAGE_MAX = 40 * 1024;

// [Or | Oi] = C^T [wr | wi] for the row tiles [tb, tb + TI) of O, into M (rows i, columns
// c < 2n, K = N).  All TI x T2 accumulators are live, so each B value feeds TI MFMAs; the
// operands of k-group chunk j + NW are loaded while chunk j runs.  With NW = 2 the waves
// take alternate chunks, wave 1's partial sums reach wave 0 through red (summed in the
// order wave 0 + wave 1) and wave 0 writes M; `aliased` adds the barrier that keeps red's
// writes after both waves' last read of a staged walker sharing that LDS.
template <int T2, int TI, int NW>
__device__ __forceinline__ void overlap_tiles(int N, int n, int tb, const double* __restrict__ trial,
                                              const double* __restrict__ wsrc, cplx* M,
                                              double* red, bool aliased, int lane, int wv) {
  const int n2 = 2 * n;
  const int s = lane >> 4, col = lane & 15;
  const int nkg = (N + 3) >> 2;
  int boff[T2];
#pragma unroll
  for (int tc = 0; tc < T2; ++tc) {
    const int c = tc * 16 + col;
    boff[tc] = c < n2 ? 2 * (c < n ? c : c - n) + (c < n ? 0 : 1) : -1;
  }
  double4_t acc[TI][T2];
#pragma unroll
  for (int ti = 0; ti < TI; ++ti)
#pragma unroll
    for (int tc = 0; tc < T2; ++tc) acc[ti][tc] = double4_t{0.0, 0.0, 0.0, 0.0};
  // two operand buffers with compile-time names (a runtime buffer index would put them
  // in scratch memory)
  double a0[MEAS_KC][TI], b0[MEAS_KC][T2], a1[MEAS_KC][TI], b1[MEAS_KC][T2];
  auto load = [&](double (&a)[MEAS_KC][TI], double (&b)[MEAS_KC][T2], int kg0) {
#pragma unroll
    for (int u = 0; u < MEAS_KC; ++u) {
      const int p = 4 * (kg0 + u) + s;
      const bool pok = p < N;
#pragma unroll
      for (int ti = 0; ti < TI; ++ti) {
        const int i = (tb + ti) * 16 + col;
        a[u][ti] = (pok && i < n) ? trial[p * n + i] : 0.0;
      }
#pragma unroll
      for (int tc = 0; tc < T2; ++tc)
        b[u][tc] = (pok && boff[tc] >= 0) ? wsrc[(int64_t)2 * p * n + boff[tc]] : 0.0;
    }
  };
  auto mma = [&](const double (&a)[MEAS_KC][TI], const double (&b)[MEAS_KC][T2]) {
#pragma unroll
    for (int u = 0; u < MEAS_KC; ++u)
#pragma unroll
      for (int ti = 0; ti < TI; ++ti)
#pragma unroll
        for (int tc = 0; tc < T2; ++tc)
          acc[ti][tc] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[u][ti], b[u][tc], acc[ti][tc], 0, 0, 0);
  };
  const int nch = (nkg + MEAS_KC - 1) / MEAS_KC;
  if (wv < nch) load(a0, b0, wv * MEAS_KC);
  for (int ch = wv; ch < nch; ch += 2 * NW) {
    if (ch + NW < nch) load(a1, b1, (ch + NW) * MEAS_KC);
    mma(a0, b0);
    if (ch + 2 * NW < nch) load(a0, b0, (ch + 2 * NW) * MEAS_KC);
    if (ch + NW < nch) mma(a1, b1);
  }
  if constexpr (NW > 1) {
    if (aliased) __syncthreads();
    if (wv == 1) {
#pragma unroll
      for (int ti = 0; ti < TI; ++ti)
#pragma unroll
        for (int tc = 0; tc < T2; ++tc)
#pragma unroll
          for (int r = 0; r < 4; ++r) red[((ti * T2 + tc) * 4 + r) * 64 + lane] = acc[ti][tc][r];
    }
    __syncthreads();
    if (wv != 0) return;
#pragma unroll
    for (int ti = 0; ti < TI; ++ti)
#pragma unroll
      for (int tc = 0; tc < T2; ++tc)
#pragma unroll
        for (int r = 0; r < 4; ++r) acc[ti][tc][r] += red[((ti * T2 + tc) * 4 + r) * 64 + lane];
  }
#pragma unroll
  for (int tc = 0; tc < T2; ++tc) {
    if (boff[tc] < 0) continue;
    const int c = tc * 16 + col;
    const int cj = c < n ? c : c - n, ci = c < n ? 0 : 1;
#pragma unroll
    for (int ti = 0; ti < TI; ++ti)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = (tb + ti) * 16 + s + 4 * r;
        if (row < n) reinterpret_cast<double*>(&M[row * n2 + cj])[ci] = acc[ti][tc][r];
      }
  }
}

// thread 0: det of this spin and, on the last spin's launch, the walker's weight
__device__ inline void finish_walker(cplx det, cplx* dets, int w, int spin,
                                     const WeightsArgs& tail) {
  dets[(int64_t)w * 2 + spin] = det;
  if (tail.weights) {
    const cplx onew = tail.nspin == 1 ? cmul(det, det) : cmul(dets[(int64_t)w * 2], det);
    walker_weight_update(tail, tail.scalars[1], w, onew);
  }
}

// n <= 24 (T2 <= 3): two waves per walker.  A walker of at most MEAS_STAGE_MAX bytes is
// staged in LDS first (one round of loads), and the region is then reused for [O | I]
// and the overlap partials.  Wave 0 runs the register Gauss-Jordan while each wave's
// first Green tile of the walker operand loads; the Green product's A operand (from X) is
// then read into registers once and the walker operand runs two tiles ahead.
template <int T2, bool STAGE>
__device__ __forceinline__ void measure_narrow_body(
    int N, int n, const cplx* __restrict__ walkers, const double* __restrict__ trial,
    cplx* __restrict__ dets, int spin, double* __restrict__ green, uint16_t* __restrict__ ghi,
    uint16_t* __restrict__ glo, int ld16, const int w, const WeightsArgs& tail) {
  constexpr int NW = 2, TI = meas_ti(T2);
  constexpr int NKG = 4 * T2;  // k groups of 2n <= 16 T2
  extern __shared__ double smem[];
  const int n2 = 2 * n;
  cplx* M = reinterpret_cast<cplx*>(smem);              // n x 2n augmented matrix
  double* red = reinterpret_cast<double*>(M + n * n2);  // wave 1's overlap partials
  // wave index as a scalar: the waves' paths below are real branches, so registers that
  // live on one wave's path only can share storage with the other's
  const int lane = threadIdx.x & 63, wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int s = lane >> 4, col = lane & 15;
  const cplx* phi = walkers + (int64_t)w * N * n;
  const double* phid = reinterpret_cast<const double*>(phi);

  if constexpr (STAGE) {
    cplx* ph = reinterpret_cast<cplx*>(smem);
    for (int e = threadIdx.x; e < N * n; e += 64 * NW) ph[e] = phi[e];
    __syncthreads();
    overlap_tiles<T2, TI, NW>(N, n, 0, trial, reinterpret_cast<const double*>(ph), M, red,
                              true, lane, wv);
  } else {
    overlap_tiles<T2, TI, NW>(N, n, 0, trial, phid, M, red, false, lane, wv);
  }
  for (int e = threadIdx.x; e < n * n; e += 64 * NW) {
    const int i = e / n, j = e - (e / n) * n;
    M[i * n2 + n + j] = cplx{i == j ? 1.0 : 0.0, 0.0};
  }
  __syncthreads();

  // ---- Gauss-Jordan with partial pivoting on wave 0: [O | I] -> [* | X], det from
  // pivots; then [Gr ; Gi] = [[Xr^T, -Xi^T], [Xi^T, Xr^T]] [wr^T ; wi^T], K = 2n,
  // columns p in tiles of 16, tiles split between the waves.
  const bool want_green = green || ghi;
  const int tp_n = (N + 15) >> 4;
  // B[k][p] = re / im of w[p][k mod n]: 32-bit offsets within a walker row (tile
  // invariant), one row base per tile, clamped loads masked after
  int boff[NKG];
#pragma unroll
  for (int kg = 0; kg < NKG; ++kg) {
    const int k = 4 * kg + s;
    boff[kg] = k < n2 ? 2 * (k < n ? k : k - n) + (k < n ? 0 : 1) : -1;
  }
  auto load_b = [&](double (&b)[NKG], int tp) {
    const int p = tp * 16 + col;
    const double* rowp = phid + (int64_t)2 * (p < N ? p : 0) * n;
#pragma unroll
    for (int kg = 0; kg < NKG; ++kg) {
      const double v = rowp[boff[kg] < 0 ? 0 : boff[kg]];
      b[kg] = (boff[kg] >= 0 && p < N) ? v : 0.0;
    }
  };
  // wave 1 loads its first tile while wave 0 eliminates
  double b[NKG];
  cplx det = cplx{1.0, 0.0};
  if (wv == 0)
    det = gauss_jordan_registers<T2>(M, n, lane);
  else if (want_green && wv < tp_n)
    load_b(b, wv);
  __syncthreads();
  if (threadIdx.x == 0) finish_walker(det, dets, w, spin, tail);
  if (!want_green) return;
  if (wv == 0) load_b(b, 0);

  // A operand of every tile, from X in LDS, read once
  double a[NKG][T2];
#pragma unroll
  for (int kg = 0; kg < NKG; ++kg)
#pragma unroll
    for (int tr = 0; tr < T2; ++tr) {
      const int k = 4 * kg + s;
      const int r = tr * 16 + col;
      double v = 0.0;
      if (k < n2 && r < n2) {
        const int kj = k < n ? k : k - n;
        const double* x = reinterpret_cast<const double*>(&M[kj * n2 + n + (r < n ? r : r - n)]);
        // X[kj][ri]: (r < n, k < n): Xr; (r < n, k >= n): -Xi; (r >= n, k < n): Xi; else Xr
        v = (r < n) == (k < n) ? x[0] : ((r < n) ? -x[1] : x[1]);
      }
      a[kg][tr] = v;
    }
  double* gw = green ? green + (int64_t)w * n2 * N : nullptr;
  uint16_t* gh = ghi ? ghi + (int64_t)w * n2 * ld16 : nullptr;
  uint16_t* gl = ghi ? glo + (int64_t)w * n2 * ld16 : nullptr;
  // tile tp; the walker operand of tile tp + NW loads into b once the MFMAs have read it
  auto tile = [&](int tp) {
    double4_t acc[T2];
#pragma unroll
    for (int tr = 0; tr < T2; ++tr) acc[tr] = double4_t{0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int kg = 0; kg < NKG; ++kg)
#pragma unroll
      for (int tr = 0; tr < T2; ++tr)
        acc[tr] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[kg][tr], b[kg], acc[tr], 0, 0, 0);
    if (tp + NW < tp_n) load_b(b, tp + NW);
    // row tr * 16 + 4 q + s of column p: one base pointer per lane, wave-uniform offsets
    const int p = tp * 16 + col;
    if (p >= N) return;
    double* gwp = gw ? gw + (int64_t)s * N + p : nullptr;
    uint16_t* ghp = gh ? gh + (int64_t)s * ld16 + p : nullptr;
    uint16_t* glp = gh ? gl + (int64_t)s * ld16 + p : nullptr;
#pragma unroll
    for (int tr = 0; tr < T2; ++tr)
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int r0 = tr * 16 + 4 * q;
        if (r0 + s < n2) {
          if (gwp) gwp[(int64_t)r0 * N] = acc[tr][q];
          if (ghp) split_bf16(acc[tr][q], &ghp[(int64_t)r0 * ld16], &glp[(int64_t)r0 * ld16]);
        }
      }
  };
  for (int tp = wv; tp < tp_n; tp += NW) tile(tp);
}

// n > 24: four waves per walker.  They take alternate row tiles of O (each over the
// whole K = N), share the Gauss-Jordan elimination on [O | I] in LDS (every wave finds
// the pivot from the same column with its own wave reduction, so no broadcast is needed;
// the row updates are split by trailing column and row group), and take alternate
// column tiles of the Green's function.
template <int T2, int TI, bool STAGE>
__device__ __forceinline__ void measure_wide_body(
    int N, int n, const cplx* __restrict__ walkers, const double* __restrict__ trial,
    cplx* __restrict__ dets, int spin, double* __restrict__ green, uint16_t* __restrict__ ghi,
    uint16_t* __restrict__ glo, int ld16, const int w, const WeightsArgs& tail) {
  constexpr int NW = 4;
  extern __shared__ double smem[];
  const int n2 = 2 * n;
  cplx* M = reinterpret_cast<cplx*>(smem);  // n x 2n augmented matrix
  cplx* fcol = M + n * n2;                  // pivot column copy (n)
  const int tid = threadIdx.x, lane = tid & 63;
  const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int s = lane >> 4, col = lane & 15;
  const cplx* phi = walkers + (int64_t)w * N * n;
  // walker operand: staged once into LDS with 16-byte loads (all in flight together)
  // when it fits, else read from global memory (L2)
  const double* phid = reinterpret_cast<const double*>(phi);
  if (STAGE) {
    cplx* ph = fcol + n;
    for (int e = tid; e < N * n; e += 64 * NW) ph[e] = phi[e];
    phid = reinterpret_cast<const double*>(ph);
  }
  __syncthreads();
  // row tiles of O, one per wave in turn
  const int ti_total = (n + 15) >> 4;
  for (int tb = wv; tb < ti_total; tb += NW)
    overlap_tiles<T2, 1, 1>(N, n, tb, trial, phid, M, nullptr, false, lane, 0);
  for (int e = tid; e < n * n; e += 64 * NW) {
    const int i = e / n, j = e - (e / n) * n;
    M[i * n2 + n + j] = cplx{i == j ? 1.0 : 0.0, 0.0};
  }
  __syncthreads();

  // ---- Gauss-Jordan with partial pivoting: [O | I] -> [* | X], det from pivots.
  cplx det = cplx{1.0, 0.0};
  for (int k = 0; k < n; ++k) {
    int pr;
    {
      // first maximum of |re| + |im| in column k (rows >= k): one 64-bit key per lane,
      // float32 magnitude in the high word, 63 - row in the low word
      const int r = k + lane;
      unsigned long long key = 0ull;
      if (r < n) {
        const float v = (float)cabs1(M[r * n2 + k]);
        key = (((unsigned long long)__float_as_uint(v)) << 32) | (unsigned)(63 - lane);
      }
      key = __ockl_wfred_max_u64(key);  // DPP wave reduction, result in every lane
      pr = k + 63 - (int)__builtin_amdgcn_readfirstlane((unsigned)(key & 0xFFFFFFFFull));
    }
    const cplx pivot = M[pr * n2 + k];
    const double d2 = 1.0 / (pivot.re * pivot.re + pivot.im * pivot.im);
    const cplx inv = cplx{pivot.re * d2, -pivot.im * d2};
    if (tid == 0) {
      det = cmul(det, pivot);
      if (pr != k) det = cplx{-det.re, -det.im};
    }
    // (column k is not written again, so the exchange below needs no barrier first)
    // pivot column as seen after the row exchange; exchange + scale the pivot row
    for (int r = tid; r < n; r += 64 * NW) fcol[r] = M[(r == pr ? k : r) * n2 + k];
    for (int c = k + 1 + tid; c < n2; c += 64 * NW) {
      double* mk = reinterpret_cast<double*>(&M[k * n2 + c]);
      double* mp = reinterpret_cast<double*>(&M[pr * n2 + c]);
      const double ar = mk[0], ai = mk[1], pr_ = mp[0], pi_ = mp[1];
      mk[0] = pr_ * inv.re - pi_ * inv.im;
      mk[1] = pr_ * inv.im + pi_ * inv.re;
      if (pr != k) {
        mp[0] = ar;
        mp[1] = ai;
      }
    }
    __syncthreads();
    // trailing columns nc = 2n - k - 1: thread -> (column k + 1 + tid % nc, row group
    // tid / nc), the rows split into as many groups as 256 threads allow
    const int nc = n2 - k - 1;
    const int groups = min(16, (64 * NW) / nc);
    const int rg = tid / nc;
    const int c = k + 1 + (tid - rg * nc);
    const int rh = (n + groups - 1) / groups;
    const int r_lo = rg * rh, r_hi = min(n, r_lo + rh);
    if (rg < groups) {
      const cplx m = M[k * n2 + c];
      for (int r0 = r_lo; r0 < r_hi; r0 += 8) {
        // eight independent loads, then the updates, then the stores
        cplx x[8], f[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          const int r = r0 + u < r_hi ? r0 + u : r_hi - 1;
          x[u] = M[r * n2 + c];
          f[u] = fcol[r];
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          const int r = r0 + u;
          if (r < r_hi && r != k) {
            cplx y = x[u];
            y.re -= f[u].re * m.re - f[u].im * m.im;
            y.im -= f[u].re * m.im + f[u].im * m.re;
            M[r * n2 + c] = y;
          }
        }
      }
    }
    __syncthreads();
  }
  if (tid == 0) finish_walker(det, dets, w, spin, tail);
  if (!(green || ghi)) return;

  // ---- [Gr ; Gi] = [[Xr^T, -Xi^T], [Xi^T, Xr^T]] [wr^T ; wi^T], K = 2n, columns p
  double* gw = green ? green + (int64_t)w * n2 * N : nullptr;
  uint16_t* gh = ghi ? ghi + (int64_t)w * n2 * ld16 : nullptr;
  uint16_t* gl = ghi ? glo + (int64_t)w * n2 * ld16 : nullptr;
  const int tp_n = (N + 15) >> 4;
    const int nkg = (n2 + 3) >> 2;
    for (int tp = wv; tp < tp_n; tp += NW) {
      const int p = tp * 16 + col;  // B column = orbital p
      double4_t acc[T2];
#pragma unroll
      for (int tr = 0; tr < T2; ++tr) acc[tr] = double4_t{0.0, 0.0, 0.0, 0.0};
      for (int kg0 = 0; kg0 < nkg; kg0 += MEAS_KC) {
        double a[MEAS_KC][T2], b[MEAS_KC];
#pragma unroll
        for (int u = 0; u < MEAS_KC; ++u) {
          const int k = 4 * (kg0 + u) + s;
          const bool kok = k < n2;
          const int kj = k < n ? k : k - n;
          b[u] = (kok && p < N) ? phid[2 * ((int64_t)p * n + kj) + (k < n ? 0 : 1)] : 0.0;
#pragma unroll
          for (int tr = 0; tr < T2; ++tr) {
            const int r = tr * 16 + col;  // A row
            double v = 0.0;
            if (kok && r < n2) {
              const cplx x = M[kj * n2 + n + (r < n ? r : r - n)];  // X[kj][ri]
              // (r < n, k < n): Xr; (r < n, k >= n): -Xi; (r >= n, k < n): Xi; else Xr
              v = (r < n) == (k < n) ? x.re : ((r < n) ? -x.im : x.im);
            }
            a[u][tr] = v;
          }
        }
#pragma unroll
        for (int u = 0; u < MEAS_KC; ++u)
#pragma unroll
          for (int tr = 0; tr < T2; ++tr)
            acc[tr] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[u][tr], b[u], acc[tr], 0, 0, 0);
      }
#pragma unroll
      for (int tr = 0; tr < T2; ++tr)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int row = tr * 16 + s + 4 * q;
          if (row < n2 && p < N) {
            if (gw) gw[(int64_t)row * N + p] = acc[tr][q];
            if (gh) split_bf16(acc[tr][q], &gh[(int64_t)row * ld16 + p], &gl[(int64_t)row * ld16 + p]);
          }
        }
    }
}

template <int T2, int TI, bool STAGE>
__global__ __launch_bounds__(64 * meas_waves(T2), meas_wpe(T2)) void measure_kernel(
    int N, int n, const cplx* __restrict__ walkers, const double* __restrict__ trial,
    cplx* __restrict__ dets, int spin, double* __restrict__ green, uint16_t* __restrict__ ghi,
    uint16_t* __restrict__ glo, int ld16, WeightsArgs tail) {
  if constexpr (T2 <= 3)
    measure_narrow_body<T2, STAGE>(N, n, walkers, trial, dets, spin, green, ghi, glo, ld16,
                                   blockIdx.x, tail);
  else
    measure_wide_body<T2, TI, STAGE>(N, n, walkers, trial, dets, spin, green, ghi, glo, ld16,
                                     blockIdx.x, tail);
}

template <int T2, int TI, bool STAGE>
__global__ __launch_bounds__(64 * meas_waves(T2), meas_wpe(T2)) void measure_grouped_kernel(const MeasureArgs* __restrict__ args,
                                                             const int* __restrict__ start,
                                                             int nprob) {
  const int p = group_of(start, nprob, blockIdx.x);
  const MeasureArgs a = args[p];
  const WeightsArgs none{};
  if constexpr (T2 <= 3)
    measure_narrow_body<T2, STAGE>(a.N, a.n, a.walkers, a.trial, a.dets, a.spin, a.green,
                                   a.ghi, a.glo, a.ld16, blockIdx.x - start[p], none);
  else
    measure_wide_body<T2, TI, STAGE>(a.N, a.n, a.walkers, a.trial, a.dets, a.spin, a.green,
                                     a.ghi, a.glo, a.ld16, blockIdx.x - start[p], none);
}

template <int T2>
int launch_measure_grouped_t(hipStream_t st, bool stage, const MeasureArgs* args,
                             const int* start, int nprob, int nblocks, size_t shm) {
  constexpr int TI = meas_ti(T2);
  const dim3 block(64 * meas_waves(T2));
  if (stage) {
    hipLaunchKernelGGL((measure_grouped_kernel<T2, TI, true>), dim3(nblocks), block, shm, st,
                       args, start, nprob);
  } else {
    static int lds_limit = 0;
    const int rc =
        ensure_dyn_lds((const void*)measure_grouped_kernel<T2, TI, false>, shm, &lds_limit);
    if (rc) return rc;
    hipLaunchKernelGGL((measure_grouped_kernel<T2, TI, false>), dim3(nblocks), block, shm, st,
                       args, start, nprob);
  }
  LAUNCH_CHECK();
  return AFQMC_OK;
}

template <int T2>
int launch_measure_t(afqmc_ctx* c, int W, int n, const cplx* wk, int s, bool g64, bool gbf,
                     const WeightsArgs& tail) {
  // O row tiles per pass: all ceil(n / 16) of them unless the accumulators would spill
  constexpr int TI = meas_ti(T2);
  const dim3 block(64 * meas_waves(T2));
  const int N = c->d.norb;
  int t2;
  bool stage;
  size_t shm;
  measure_plan(c, s, &t2, &stage, &shm);
  if (stage) {  // small walkers only: staging costs occupancy
    hipLaunchKernelGGL((measure_kernel<T2, TI, true>), dim3(W), block, shm, c->stream, N,
                       n, wk, c->trial[s], c->dets, s, g64 ? c->green[s] : nullptr,
                       gbf ? c->green_bf[s][0] : nullptr, gbf ? c->green_bf[s][1] : nullptr,
                       c->Np32, tail);
    LAUNCH_CHECK();
    return AFQMC_OK;
  }
  static int lds_limit = 0;
  const int rc = ensure_dyn_lds((const void*)measure_kernel<T2, TI, false>, shm, &lds_limit);
  if (rc) return rc;
  hipLaunchKernelGGL((measure_kernel<T2, TI, false>), dim3(W), block, shm, c->stream, c->d.norb, n, wk,
                     c->trial[s], c->dets, s, g64 ? c->green[s] : nullptr,
                     gbf ? c->green_bf[s][0] : nullptr, gbf ? c->green_bf[s][1] : nullptr, c->Np32,
                     tail);
  LAUNCH_CHECK();
  return AFQMC_OK;
}

}  // namespace

int launch_measure(afqmc_ctx* c, int W, const cplx* wup, const cplx* wdn, bool g64, bool gbf,
                   const WeightsArgs* tail) {
  for (int sp = 0; sp < c->nspin; ++sp) {
    const int n = c->nocc[sp];
    const cplx* wk = (sp == 0) ? wup : wdn;
    const WeightsArgs t = (tail && sp == c->nspin - 1) ? *tail : WeightsArgs{};
    int rc;
    switch ((2 * n + 15) / 16) {
      case 1: rc = launch_measure_t<1>(c, W, n, wk, sp, g64, gbf, t); break;
      case 2: rc = launch_measure_t<2>(c, W, n, wk, sp, g64, gbf, t); break;
      case 3: rc = launch_measure_t<3>(c, W, n, wk, sp, g64, gbf, t); break;
      case 4: rc = launch_measure_t<4>(c, W, n, wk, sp, g64, gbf, t); break;
      case 5: rc = launch_measure_t<5>(c, W, n, wk, sp, g64, gbf, t); break;
      case 6: rc = launch_measure_t<6>(c, W, n, wk, sp, g64, gbf, t); break;
      case 7: rc = launch_measure_t<7>(c, W, n, wk, sp, g64, gbf, t); break;
      default: rc = launch_measure_t<8>(c, W, n, wk, sp, g64, gbf, t); break;
    }
    if (rc) return rc;
  }
  return AFQMC_OK;
}

// Grouped form: one launch for every (fragment, spin) problem with the same column-tile
// count t2 and staging choice; shm is the largest LDS need among them.
int launch_measure_grouped(hipStream_t st, int t2, bool stage, const MeasureArgs* args,
                           const int* start, int nprob, int nblocks, size_t shm) {
  switch (t2) {
    case 1: return launch_measure_grouped_t<1>(st, stage, args, start, nprob, nblocks, shm);
    case 2: return launch_measure_grouped_t<2>(st, stage, args, start, nprob, nblocks, shm);
    case 3: return launch_measure_grouped_t<3>(st, stage, args, start, nprob, nblocks, shm);
    case 4: return launch_measure_grouped_t<4>(st, stage, args, start, nprob, nblocks, shm);
    case 5: return launch_measure_grouped_t<5>(st, stage, args, start, nprob, nblocks, shm);
    case 6: return launch_measure_grouped_t<6>(st, stage, args, start, nprob, nblocks, shm);
    case 7: return launch_measure_grouped_t<7>(st, stage, args, start, nprob, nblocks, shm);
    default: return launch_measure_grouped_t<8>(st, stage, args, start, nprob, nblocks, shm);
  }
}

// The kernel launch_measure picks for spin s, as a plan for grouped launches.
void measure_plan(const afqmc_ctx* c, int s, int* t2, bool* stage, size_t* shm) {
  const int n = c->nocc[s], N = c->d.norb;
  const int t = (2 * n + 15) / 16;
  *t2 = t > 8 ? 8 : t;
  const size_t walker = (size_t)N * n * sizeof(cplx);
  if (*t2 <= 3) {  // two-wave kernel: [O | I] and partials alias the staged walker
    const size_t m_red = (size_t)n * 2 * n * sizeof(cplx) + meas_red_bytes(*t2);
    *stage = walker <= MEAS_STAGE_MAX;
    *shm = *stage ? std::max(walker, m_red) : m_red;
    return;
  }
  const size_t base = (size_t)(n * 2 * n + n) * sizeof(cplx);
  *stage = base + walker <= 24 * 1024;  // small walkers only: staging costs occupancy
  *shm = *stage ? base + walker : base;
}
