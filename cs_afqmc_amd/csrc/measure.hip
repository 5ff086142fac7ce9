// Trial measurements per walker and spin: O = C^T w (n x n), O^-1 and det(O) by a blocked
// in-place Gauss-Jordan inversion with partial pivoting (pivot: the largest |re| + |im|
// over the remaining rows, LAPACK's izamax measure), and the half-rotated Green's
// function G = (w O^-1)^T (n x N).
//
// Follows rhf._calc_overlap_restricted / _calc_green (wavefunctions.py:478-501) and
// uhf._calc_overlap / _calc_green (wavefunctions.py:658-692).  The trial C is real.
//
// One workgroup per (walker, spin): two waves for n <= 24, four beyond.  The overlap and
// Green products run on v_mfma_f64_16x16x4f64 in real block form:
//   [Or | Oi]  = C^T [wr | wi]                         (n x 2n,  K = N)
//   [Gr ; Gi]  = [[Xr^T, -Xi^T], [Xi^T, Xr^T]] [wr^T ; wi^T]   (2n x N, K = 2n)
// The second product's rows are exactly the rows of the split Green layout
// green[w][c][i][p] (c = re / im), so results go straight to global memory.
//
// Inversion: O is overwritten in LDS (n x n complex) by a blocked in-place Gauss-Jordan
// elimination with implicit pivoting (rows stay in place), in panels of GJ_PB = 4 pivot
// columns.  The wave that owns a panel's 16-column block factors it with a lane per row
// (pivot search by a wave reduction, the panel's other columns updated in registers); the
// panel columns then hold the panel's transform, and every other column is updated by a
// complex rank-4 product on MFMA (four real MFMAs per 16 x 16 tile), each wave over its
// own column block.  With r_k the pivot row of column k, O^-1[i][j] = S[r_i][r^-1(j)] for
// the stored S; the Green product reads X through these two maps.
#include <algorithm>

#include "common.h"


namespace {

constexpr int MEAS_KC = 4;  // k-groups whose operands are loaded ahead of their MFMAs

__device__ inline double readlane_f64(double x, int l) {
  const unsigned long long u = (unsigned long long)__double_as_longlong(x);
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)u, l);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(u >> 32), l);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

__device__ inline cplx readlane_c(cplx x, int l) {
  return cplx{readlane_f64(x.re, l), readlane_f64(x.im, l)};
}

constexpr int GJ_PB = 4;  // panel width: the K of one f64 MFMA

// LDS of the inversion: B (n x ld complex), then the state below (256 ints)
__host__ __device__ constexpr int gj_ld(int n) { return n | 1; }
__host__ __device__ constexpr size_t gj_bytes(int n) {
  return (size_t)n * gj_ld(n) * sizeof(cplx) + 256 * sizeof(int);
}
struct GjState {
  int rho[64];     // rho[k]: the pivot row of column k
  int rhoinv[64];  // its inverse (filled at the end)
  int vrow[64];    // row order of the equivalent row-exchanging elimination (for the sign)
  unsigned long long used;      // rows pivoted so far
  unsigned long long pmask[2];  // pivot rows of the current panel (double-buffered)
  unsigned long long pad;
  cplx det;
};
static_assert(sizeof(GjState) <= 256 * sizeof(int), "GjState exceeds its LDS slot");

// Barrier over the workgroup's LDS traffic only: waits for this wave's LDS operations,
// not for its global loads in flight (the Green product's first operands are loaded
// before the inversion and stay outstanding across it)
__device__ inline void lds_barrier() {
  asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
}

// 1 / x for x > 0 of moderate range: hardware estimate plus two Newton steps
__device__ inline double rcp_f64(double x) {
  double y = __builtin_amdgcn_rcp(x);
  double e = fma(-x, y, 1.0);
  y = fma(y, e, y);
  e = fma(-x, y, 1.0);
  return fma(y, e, y);
}

// Panel [k0, k0 + pb) on one wave, lane = row r < n (n <= 64), rows kept in place
// (implicit pivoting).  Pivot of column k: the largest |re| + |im| over the rows not yet
// pivoted (compared to 2^-17 relative, the first row among near-equal ones).  The pivot row r_k is scaled by 1 / piv (its column-k entry becomes
// 1 / piv) and every other row r loses f_r = A[r][k] times it (its column-k entry
// becomes -f_r / piv), over the panel's columns only.  Updates the state (pivot rows,
// used rows, row order, det).
__device__ __forceinline__ void gj_panel(cplx* B, int ld, int n, int k0, int pb, int panel,
                                         GjState* st, int lane) {
  const bool live = lane < n;
  unsigned long long used = st->used;
  int vr = st->vrow[lane];
  cplx det = st->det;
  unsigned long long pm = 0ull;
  int rk[GJ_PB];
  cplx p[GJ_PB];
  // loads clamped into the matrix and masked after, so they issue back to back
#pragma unroll
  for (int j = 0; j < GJ_PB; ++j) {
    const cplx v = B[(live ? lane : 0) * ld + k0 + (j < pb ? j : 0)];
    p[j] = (live && j < pb) ? v : cplx{0.0, 0.0};
    rk[j] = 0;
  }
#pragma unroll
  for (int j = 0; j < GJ_PB; ++j) {
    if (j >= pb) continue;
    const int k = k0 + j;
    // one 32-bit key per candidate row: |re| + |im| as float32 with its low 6 bits
    // replaced by 63 - row, so an integer wave maximum picks the largest magnitude and,
    // among magnitudes equal to 2^-17 relative, the first row
    const bool cand = live && !((used >> lane) & 1ull);
    const float amag = (float)(fabs(p[j].re) + fabs(p[j].im));
    unsigned key = 0u;
    if (cand) key = (amag == amag ? (__float_as_uint(amag) & ~63u) : 0u) | (unsigned)(63 - lane);
    const unsigned kmax = __ockl_wfred_max_u32(key);
    // (a column of NaN or zeros only: the first remaining row)
    const int pr = kmax ? 63 - (int)(kmax & 63u) : __builtin_ctzll(__ballot(cand));
    used |= 1ull << pr;
    pm |= 1ull << pr;
    rk[j] = pr;
    // the row-exchanging elimination would swap positions k and q (row pr's position)
    const int q = __builtin_amdgcn_readfirstlane((int)__builtin_ctzll(__ballot(vr == pr)));
    const int vk = __builtin_amdgcn_readlane(vr, k);
    vr = lane == k ? pr : (lane == q ? vk : vr);
    const cplx piv = readlane_c(p[j], pr);
    det = cmul(det, piv);
    if (q != k) det = cplx{-det.re, -det.im};
    const double d2 = rcp_f64(piv.re * piv.re + piv.im * piv.im);
    const cplx ipiv = cplx{piv.re * d2, -piv.im * d2};
    cplx qv[GJ_PB];
#pragma unroll
    for (int jj = 0; jj < GJ_PB; ++jj) qv[jj] = jj == j ? ipiv : cmul(readlane_c(p[jj], pr), ipiv);
    const cplx f = p[j];
    const bool pivrow = lane == pr;
#pragma unroll
    for (int jj = 0; jj < GJ_PB; ++jj) {
      if (jj == j) {
        const cplx e = cmul(f, ipiv);
        p[jj] = pivrow ? ipiv : cplx{-e.re, -e.im};
      } else {
        const cplx e = cmul(f, qv[jj]);
        p[jj] = pivrow ? qv[jj] : cplx{p[jj].re - e.re, p[jj].im - e.im};
      }
    }
  }
#pragma unroll
  for (int j = 0; j < GJ_PB; ++j)
    if (live && j < pb) B[lane * ld + k0 + j] = p[j];
  st->vrow[lane] = vr;
  if (lane < pb) st->rho[k0 + lane] = lane == 0 ? rk[0] : lane == 1 ? rk[1] : lane == 2 ? rk[2] : rk[3];
  if (lane == 0) {
    st->used = used;
    st->pmask[panel & 1] = pm;
    st->det = det;
  }
}

// In-place inversion of the n x n complex matrix B (row stride ld, LDS; st right after
// it) by NW >= ceil(n / 16) waves; TT >= ceil(n / 16) row tiles.  Wave w owns column block w (all its row tiles): it
// factors the panels that fall in its block and applies every panel's transform to its
// block, so one barrier per panel orders the factorization before the other blocks'
// updates.  For panel K (pivot rows R_K) and a column x outside it:
//   x <- T x[R_K] + (x with the rows R_K cleared),  T = the panel's columns,
// four real MFMAs per 16 x 16 tile.  On return B holds S with
//   O^-1[i][j] = S[rho[i]][rhoinv[j]],
// and st->det = det(O).  Ends with a barrier.
template <int NW, int TT>
__device__ __forceinline__ void gj_invert(cplx* B, int ld, int n, GjState* st, int lane,
                                          int wv) {
  const int tid = threadIdx.x;
  if (tid < 64) st->vrow[tid] = tid;
  if (tid == 0) {
    st->used = 0ull;
    st->det = cplx{1.0, 0.0};
  }
  const int tt = (n + 15) >> 4;
  const int kk = lane >> 4, c16 = lane & 15;
  const int cb = 16 * wv + c16;  // this wave's column
  const bool mine = wv < tt;
  lds_barrier();
  for (int k0 = 0, panel = 0; k0 < n; k0 += GJ_PB, ++panel) {
    const int pb = min(GJ_PB, n - k0);
    if (wv == (k0 >> 4)) gj_panel(B, ld, n, k0, pb, panel, st, lane);
    lds_barrier();
    if (!mine) continue;
    // loads clamped into the matrix and masked after, so they issue back to back
    const unsigned long long pm = st->pmask[panel & 1];
    const bool kok = kk < pb, cok = cb < n;
    const int kc = k0 + (kok ? kk : 0), cbc = cok ? cb : 0;
    const cplx xv = B[st->rho[kc] * ld + cbc];
    const double br = (kok && cok) ? xv.re : 0.0, bi = (kok && cok) ? xv.im : 0.0;
    double4_t dr[TT], di[TT];
#pragma unroll
    for (int tr = 0; tr < TT; ++tr) {
      dr[tr] = di[tr] = double4_t{0.0, 0.0, 0.0, 0.0};
      if (tr >= tt) continue;
      const int ra = 16 * tr + c16;
      const cplx tv = B[(ra < n ? ra : 0) * ld + kc];
      const double ar = (kok && ra < n) ? tv.re : 0.0, ai = (kok && ra < n) ? tv.im : 0.0;
      cplx dv[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int row = 16 * tr + kk + 4 * q;
        dv[q] = B[(row < n ? row : 0) * ld + cbc];
      }
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int row = 16 * tr + kk + 4 * q;
        const bool ok = row < n && cok && !((pm >> row) & 1ull);
        dr[tr][q] = ok ? dv[q].re : 0.0;
        di[tr][q] = ok ? dv[q].im : 0.0;
      }
      dr[tr] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar, br, dr[tr], 0, 0, 0);
      dr[tr] = __builtin_amdgcn_mfma_f64_16x16x4f64(-ai, bi, dr[tr], 0, 0, 0);
      di[tr] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar, bi, di[tr], 0, 0, 0);
      di[tr] = __builtin_amdgcn_mfma_f64_16x16x4f64(ai, br, di[tr], 0, 0, 0);
    }
    // every read of this block precedes the first write (the wave's LDS operations stay
    // in order); the panel's own columns keep T
    if (cok && (cb < k0 || cb >= k0 + pb)) {
#pragma unroll
      for (int tr = 0; tr < TT; ++tr) {
        if (tr >= tt) continue;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int row = 16 * tr + kk + 4 * q;
          if (row < n) B[row * ld + cb] = cplx{dr[tr][q], di[tr][q]};
        }
      }
    }
  }
  lds_barrier();
  if (tid < n) st->rhoinv[st->rho[tid]] = tid;
  lds_barrier();
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

// [Or | Oi] = C^T [wr | wi] for the row tiles [tb, tb + TI) of O, into M (complex n x n,
// row stride ldm; K = N).  All TI x T2 accumulators are live, so each B value feeds TI MFMAs; the
// operands of k-group chunk j + NW are loaded while chunk j runs.  With NW = 2 the waves
// take alternate chunks, wave 1's partial sums reach wave 0 through red (summed in the
// order wave 0 + wave 1) and wave 0 writes M; `aliased` adds the barrier that keeps red's
// writes after both waves' last read of a staged walker sharing that LDS.
template <int T2, int TI, int NW>
__device__ __forceinline__ void overlap_tiles(int N, int n, int tb, const double* __restrict__ trial,
                                              const double* __restrict__ wsrc, cplx* M, int ldm,
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
  // Loads come from clamped (always valid) addresses and issue unconditionally.  Rows
  // i >= n and columns c >= 2n of the product are never stored, so only the k padding
  // p >= N needs a zero, on one operand; it is applied where the MFMA reads the operand,
  // so the wait for a chunk's loads falls just before its MFMAs.
  auto load = [&](double (&a)[MEAS_KC][TI], double (&b)[MEAS_KC][T2], int kg0) {
#pragma unroll
    for (int u = 0; u < MEAS_KC; ++u) {
      const int pc = min(4 * (kg0 + u) + s, N - 1);
#pragma unroll
      for (int ti = 0; ti < TI; ++ti) a[u][ti] = trial[pc * n + min((tb + ti) * 16 + col, n - 1)];
#pragma unroll
      for (int tc = 0; tc < T2; ++tc) b[u][tc] = wsrc[(int64_t)2 * pc * n + max(boff[tc], 0)];
    }
  };
  auto mma = [&](const double (&a)[MEAS_KC][TI], const double (&b)[MEAS_KC][T2], int kg0) {
#pragma unroll
    for (int u = 0; u < MEAS_KC; ++u) {
      const bool pok = 4 * (kg0 + u) + s < N;
#pragma unroll
      for (int ti = 0; ti < TI; ++ti) {
        const double av = __builtin_unpredictable(pok) ? a[u][ti] : 0.0;
#pragma unroll
        for (int tc = 0; tc < T2; ++tc)
          acc[ti][tc] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, b[u][tc], acc[ti][tc], 0, 0, 0);
      }
    }
  };
  const int nch = (nkg + MEAS_KC - 1) / MEAS_KC;
  if (wv < nch) load(a0, b0, wv * MEAS_KC);
  for (int ch = wv; ch < nch; ch += 2 * NW) {
    if (ch + NW < nch) load(a1, b1, (ch + NW) * MEAS_KC);
    mma(a0, b0, ch * MEAS_KC);
    if (ch + 2 * NW < nch) load(a0, b0, (ch + 2 * NW) * MEAS_KC);
    if (ch + NW < nch) mma(a1, b1, (ch + NW) * MEAS_KC);
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
        if (row < n) reinterpret_cast<double*>(&M[row * ldm + cj])[ci] = acc[ti][tc][r];
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

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef unsigned int u32x4_t __attribute__((ext_vector_type(4)));

// 8 floats -> hi and lo bf16 planes, x ~= hi + lo
__device__ inline void split8_bf(const float (&v)[8], bf16x8_t& h8, bf16x8_t& l8) {
  u32x4_t ph, pl;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const uint32_t h = pack_bf16x2(v[2 * j], v[2 * j + 1]);
    ph[j] = h;
    pl[j] = pack_bf16x2(v[2 * j] - bf16_lo_val(h), v[2 * j + 1] - bf16_hi_val(h));
  }
  h8 = __builtin_bit_cast(bf16x8_t, ph);
  l8 = __builtin_bit_cast(bf16x8_t, pl);
}

// The mixed mode's Green product, where only the split-bf16 planes are wanted (they are
// the A operand of the split-bf16 force-bias GEMM): [Gr ; Gi] = X' [wr^T ; wi^T] with
// X' = [[Xr^T, -Xi^T], [Xi^T, Xr^T]] (2n x 2n), on v_mfma_f32_16x16x32_bf16 with both
// operands split into hi + lo bf16 planes (al bh + ah bl + ah bh, float32 accumulation)
// rather than on float64 MFMA.  K = 2n runs in KB blocks of 32, zero beyond 2n in A.
//   A (lane l): row r = 16 tr + l % 16, k = 32 kb + 8 (l / 16) + e, read once from X in LDS
//   B (lane l): k = 32 kb + 8 (l / 16) + e, column p = 16 tp + l % 16, from the walker
//   D (lane l, element j): row 16 tr + 4 (l / 16) + j, column 16 tp + l % 16
// bpre holds column tile wv's walker operand, loaded by the caller (it is in flight
// during the inversion).  The waves take alternate column tiles.
template <int T2>
struct GreenBf {
  static constexpr int KB = (T2 + 1) / 2;  // 2n <= 16 T2
};
template <int T2>
__device__ __forceinline__ void green_bf_load_b(const double* __restrict__ phid, int N, int n,
                                                int tp, int lane, double (&b)[GreenBf<T2>::KB][8]) {
  const int s = lane >> 4, col = lane & 15;
  const double* rowp = phid + (int64_t)2 * min(tp * 16 + col, N - 1) * n;
#pragma unroll
  for (int kb = 0; kb < GreenBf<T2>::KB; ++kb)
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const int k = kb * 32 + 8 * s + e;  // k >= 2n meets zeros of A: any valid entry
      const int kj = min(k < n ? k : k - n, n - 1);
      b[kb][e] = rowp[2 * kj + (k < n ? 0 : 1)];
    }
}
template <int T2, int NW>
__device__ __forceinline__ void green_bf_tiles(int N, int n, const double* __restrict__ phid,
                                               const cplx* M, int ld, const GjState* st,
                                               uint16_t* __restrict__ gh,
                                               uint16_t* __restrict__ gl, int ld16, int lane,
                                               int wv, double (&bpre)[GreenBf<T2>::KB][8]) {
  constexpr int KB = GreenBf<T2>::KB;
  const int n2 = 2 * n;
  const int s = lane >> 4, col = lane & 15;
  const int tp_n = (N + 15) >> 4;
  bf16x8_t ah[T2][KB], al[T2][KB];
  {
    int xcol[T2];
#pragma unroll
    for (int tr = 0; tr < T2; ++tr) {
      const int r = tr * 16 + col;  // (rows r >= 2n are not stored)
      xcol[tr] = st->rhoinv[min(r < n ? r : r - n, n - 1)];
    }
#pragma unroll
    for (int kb = 0; kb < KB; ++kb) {
      float v[T2][8];
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        const int k = kb * 32 + 8 * s + e;
        const int xr = st->rho[min(k < n ? k : k - n, n - 1)] * ld;
#pragma unroll
        for (int tr = 0; tr < T2; ++tr) {
          const int r = tr * 16 + col;
          const cplx x = M[xr + xcol[tr]];  // X[kj][ri]
          // (r < n, k < n): Xr; (r < n, k >= n): -Xi; (r >= n, k < n): Xi; else Xr
          const double d = (r < n) == (k < n) ? x.re : ((r < n) ? -x.im : x.im);
          v[tr][e] = k < n2 ? (float)d : 0.f;
        }
      }
#pragma unroll
      for (int tr = 0; tr < T2; ++tr) split8_bf(v[tr], ah[tr][kb], al[tr][kb]);
    }
  }
  for (int tp = wv; tp < tp_n; tp += NW) {
    bf16x8_t bh[KB], bl[KB];
#pragma unroll
    for (int kb = 0; kb < KB; ++kb) {
      float f[8];
#pragma unroll
      for (int e = 0; e < 8; ++e) f[e] = (float)bpre[kb][e];
      split8_bf(f, bh[kb], bl[kb]);
    }
    if (tp + NW < tp_n) green_bf_load_b<T2>(phid, N, n, tp + NW, lane, bpre);
    float4_t acc[T2];
#pragma unroll
    for (int tr = 0; tr < T2; ++tr) {
      acc[tr] = float4_t{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int kb = 0; kb < KB; ++kb) {
        acc[tr] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[tr][kb], bh[kb], acc[tr], 0, 0, 0);
        acc[tr] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[tr][kb], bl[kb], acc[tr], 0, 0, 0);
        acc[tr] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[tr][kb], bh[kb], acc[tr], 0, 0, 0);
      }
    }
    const int p = tp * 16 + col;
    if (p >= N) continue;
#pragma unroll
    for (int tr = 0; tr < T2; ++tr)
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int r = tr * 16 + 4 * s + j;
        if (r < n2) {
          const float v = acc[tr][j];
          const uint16_t h = f32_to_bf16_rne(v);
          gh[(int64_t)r * ld16 + p] = h;
          gl[(int64_t)r * ld16 + p] = f32_to_bf16_rne(v - __uint_as_float((uint32_t)h << 16));
        }
      }
  }
}

// n <= 24 (T2 <= 3): two waves per walker.  STAGE 2: the walker is staged in LDS (one
// round of loads) and stays there for the Green product, with O and the inversion's
// state after it; with two row tiles of O each wave then forms one of them over the whole
// K, else the waves split K and wave 1's partial sums reach wave 0 through LDS.  STAGE 1
// (walker too large to keep beside O): staged for the overlap product only, the region
// then reused for O, the state and the partials, the Green product reading the walker
// from global memory.  STAGE 0: no staging.  Each wave's first Green tile of the walker
// operand loads while the waves invert O; the Green product's A operand (from X) is then
// read into registers once and the walker operand runs two tiles ahead.
template <int T2, int STAGE>
__device__ __forceinline__ void measure_narrow_body(
    int N, int n, const cplx* __restrict__ walkers, const double* __restrict__ trial,
    cplx* __restrict__ dets, int spin, double* __restrict__ green, uint16_t* __restrict__ ghi,
    uint16_t* __restrict__ glo, int ld16, const int w, const WeightsArgs& tail) {
  constexpr int NW = 2, TI = meas_ti(T2);
  constexpr int NKG = 4 * T2;  // k groups of 2n <= 16 T2
  extern __shared__ double smem[];
  const int n2 = 2 * n, ld = gj_ld(n);
  cplx* ph = reinterpret_cast<cplx*>(smem);                   // staged walker (STAGE > 0)
  cplx* M = STAGE == 2 ? ph + N * n : ph;                     // O, then its inverse: n x ld
  GjState* st = reinterpret_cast<GjState*>(M + n * ld);       // the inversion's state
  double* red = reinterpret_cast<double*>(reinterpret_cast<int*>(st) + 256);  // overlap partials
  // wave index as a scalar: the waves' paths below are real branches, so registers that
  // live on one wave's path only can share storage with the other's
  const int lane = threadIdx.x & 63, wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int s = lane >> 4, col = lane & 15;
  const cplx* phi = walkers + (int64_t)w * N * n;
  const double* phid = reinterpret_cast<const double*>(phi);

  if constexpr (STAGE == 2) {
    for (int e = threadIdx.x; e < N * n; e += 64 * NW) ph[e] = phi[e];
    __syncthreads();
    phid = reinterpret_cast<const double*>(ph);
    if constexpr (TI == 2)
      overlap_tiles<T2, 1, 1>(N, n, wv, trial, phid, M, ld, nullptr, false, lane, 0);
    else
      overlap_tiles<T2, TI, NW>(N, n, 0, trial, phid, M, ld, red, false, lane, wv);
  } else if constexpr (STAGE == 1) {
    for (int e = threadIdx.x; e < N * n; e += 64 * NW) ph[e] = phi[e];
    __syncthreads();
    overlap_tiles<T2, TI, NW>(N, n, 0, trial, reinterpret_cast<const double*>(ph), M, ld,
                              red, true, lane, wv);
  } else {
    overlap_tiles<T2, TI, NW>(N, n, 0, trial, phid, M, ld, red, false, lane, wv);
  }

  // ---- O^-1 and det(O) (blocked in-place Gauss-Jordan, both waves); then
  // [Gr ; Gi] = [[Xr^T, -Xi^T], [Xi^T, Xr^T]] [wr^T ; wi^T], K = 2n, columns p in tiles
  // of 16, tiles split between the waves.
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
  // (clamped, unmasked: columns p >= N are not stored, and A is zero for k >= 2n)
  auto load_b = [&](double (&b)[NKG], int tp) {
    const double* rowp = phid + (int64_t)2 * min(tp * 16 + col, N - 1) * n;
#pragma unroll
    for (int kg = 0; kg < NKG; ++kg) b[kg] = rowp[max(boff[kg], 0)];
  };
  // each wave's first tile loads while the waves invert O
  double b[NKG];
  double bbf[GreenBf<T2>::KB][8];
  if (green && wv < tp_n) load_b(b, wv);
  if (!green && ghi && wv < tp_n) green_bf_load_b<T2>(phid, N, n, wv, lane, bbf);
  gj_invert<NW, 2>(M, ld, n, st, lane, wv);
  if (threadIdx.x == 0) finish_walker(st->det, dets, w, spin, tail);
  if (!want_green) return;
  if (!green) {  // split-bf16 planes only (mixed mode's step)
    green_bf_tiles<T2, NW>(N, n, phid, M, ld, st, ghi + (int64_t)w * n2 * ld16,
                           glo + (int64_t)w * n2 * ld16, ld16, lane, wv, bbf);
    return;
  }

  // A operand of every tile, from X in LDS, read once: the pivot maps of the lane's rows
  // and columns first, then every element, each load from a clamped index
  int xrow[NKG], xcol[T2];
#pragma unroll
  for (int kg = 0; kg < NKG; ++kg) {
    const int k = 4 * kg + s;
    xrow[kg] = st->rho[min(k < n ? k : k - n, n - 1)] * ld;
  }
#pragma unroll
  for (int tr = 0; tr < T2; ++tr) {
    const int r = tr * 16 + col;
    xcol[tr] = st->rhoinv[min(r < n ? r : r - n, n - 1)];
  }
  double a[NKG][T2];
#pragma unroll
  for (int kg = 0; kg < NKG; ++kg)
#pragma unroll
    for (int tr = 0; tr < T2; ++tr) {
      const int k = 4 * kg + s;
      const int r = tr * 16 + col;
      const cplx x = M[xrow[kg] + xcol[tr]];
      // X[kj][ri]: (r < n, k < n): Xr; (r < n, k >= n): -Xi; (r >= n, k < n): Xi; else Xr
      const double v = (r < n) == (k < n) ? x.re : ((r < n) ? -x.im : x.im);
      a[kg][tr] = k < n2 ? v : 0.0;  // (rows r >= 2n are not stored)
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
// whole K = N), share the blocked inversion of O, and take alternate column tiles of the
// Green's function.
template <int T2, int TI, bool STAGE>
__device__ __forceinline__ void measure_wide_body(
    int N, int n, const cplx* __restrict__ walkers, const double* __restrict__ trial,
    cplx* __restrict__ dets, int spin, double* __restrict__ green, uint16_t* __restrict__ ghi,
    uint16_t* __restrict__ glo, int ld16, const int w, const WeightsArgs& tail) {
  constexpr int NW = 4;
  extern __shared__ double smem[];
  const int n2 = 2 * n, ld = gj_ld(n);
  cplx* M = reinterpret_cast<cplx*>(smem);               // O, then its inverse: n x ld
  GjState* st = reinterpret_cast<GjState*>(M + n * ld);  // the inversion's state
  const int tid = threadIdx.x, lane = tid & 63;
  const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int s = lane >> 4, col = lane & 15;
  const cplx* phi = walkers + (int64_t)w * N * n;
  // walker operand: staged once into LDS with 16-byte loads (all in flight together)
  // when it fits, else read from global memory (L2)
  const double* phid = reinterpret_cast<const double*>(phi);
  if (STAGE) {
    cplx* ph = reinterpret_cast<cplx*>(reinterpret_cast<int*>(st) + 256);
    for (int e = tid; e < N * n; e += 64 * NW) ph[e] = phi[e];
    phid = reinterpret_cast<const double*>(ph);
  }
  __syncthreads();
  // row tiles of O, one per wave in turn
  const int ti_total = (n + 15) >> 4;
  for (int tb = wv; tb < ti_total; tb += NW)
    overlap_tiles<T2, 1, 1>(N, n, tb, trial, phid, M, ld, nullptr, false, lane, 0);

  // ---- O^-1 and det(O): blocked in-place Gauss-Jordan on all four waves
  gj_invert<NW, 4>(M, ld, n, st, lane, wv);
  const cplx det = st->det;
  if (tid == 0) finish_walker(det, dets, w, spin, tail);
  if (!(green || ghi)) return;

  // ---- [Gr ; Gi] = [[Xr^T, -Xi^T], [Xi^T, Xr^T]] [wr^T ; wi^T], K = 2n, columns p
  double* gw = green ? green + (int64_t)w * n2 * N : nullptr;
  uint16_t* gh = ghi ? ghi + (int64_t)w * n2 * ld16 : nullptr;
  uint16_t* gl = ghi ? glo + (int64_t)w * n2 * ld16 : nullptr;
  const int tp_n = (N + 15) >> 4;
    const int nkg = (n2 + 3) >> 2;
    // columns of X in the lane's A rows r = tr * 16 + col (pivot map, clamped index:
    // rows r >= 2n are not stored)
    int xcol[T2];
#pragma unroll
    for (int tr = 0; tr < T2; ++tr) {
      const int r = tr * 16 + col;
      xcol[tr] = st->rhoinv[min(r < n ? r : r - n, n - 1)];
    }
    for (int tp = wv; tp < tp_n; tp += NW) {
      const int p = tp * 16 + col;  // B column = orbital p
      const int pc = min(p, N - 1);   // (columns p >= N are not stored)
      double4_t acc[T2];
#pragma unroll
      for (int tr = 0; tr < T2; ++tr) acc[tr] = double4_t{0.0, 0.0, 0.0, 0.0};
      for (int kg0 = 0; kg0 < nkg; kg0 += MEAS_KC) {
        // loads from clamped indices, unconditional; A is zero for k >= 2n
        double a[MEAS_KC][T2], b[MEAS_KC];
#pragma unroll
        for (int u = 0; u < MEAS_KC; ++u) {
          const int k = 4 * (kg0 + u) + s;
          const int kj = min(k < n ? k : k - n, n - 1);
          b[u] = phid[2 * ((int64_t)pc * n + kj) + (k < n ? 0 : 1)];
          const int xr = st->rho[kj] * ld;
#pragma unroll
          for (int tr = 0; tr < T2; ++tr) {
            const int r = tr * 16 + col;  // A row
            const cplx x = M[xr + xcol[tr]];  // X[kj][ri]
            // (r < n, k < n): Xr; (r < n, k >= n): -Xi; (r >= n, k < n): Xi; else Xr
            const double v = (r < n) == (k < n) ? x.re : ((r < n) ? -x.im : x.im);
            a[u][tr] = __builtin_unpredictable(k < n2) ? v : 0.0;
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

template <int T2, int TI, int STAGE>
__global__ __launch_bounds__(64 * meas_waves(T2), meas_wpe(T2)) void measure_kernel(
    int N, int n, const cplx* __restrict__ walkers, const double* __restrict__ trial,
    cplx* __restrict__ dets, int spin, double* __restrict__ green, uint16_t* __restrict__ ghi,
    uint16_t* __restrict__ glo, int ld16, WeightsArgs tail) {
  if constexpr (T2 <= 3)
    measure_narrow_body<T2, STAGE>(N, n, walkers, trial, dets, spin, green, ghi, glo, ld16,
                                   blockIdx.x, tail);
  else
    measure_wide_body<T2, TI, STAGE != 0>(N, n, walkers, trial, dets, spin, green, ghi, glo,
                                          ld16, blockIdx.x, tail);
}

template <int T2, int TI, int STAGE>
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
    measure_wide_body<T2, TI, STAGE != 0>(a.N, a.n, a.walkers, a.trial, a.dets, a.spin,
                                          a.green, a.ghi, a.glo, a.ld16, blockIdx.x - start[p],
                                          none);
}

template <int T2, int STAGE>
int launch_measure_grouped_s(hipStream_t st, const MeasureArgs* args, const int* start,
                             int nprob, int nblocks, size_t shm) {
  constexpr int TI = meas_ti(T2);
  static int lds_limit = 0;
  const int rc =
      ensure_dyn_lds((const void*)measure_grouped_kernel<T2, TI, STAGE>, shm, &lds_limit);
  if (rc) return rc;
  hipLaunchKernelGGL((measure_grouped_kernel<T2, TI, STAGE>), dim3(nblocks),
                     dim3(64 * meas_waves(T2)), shm, st, args, start, nprob);
  LAUNCH_CHECK();
  return AFQMC_OK;
}

template <int T2>
int launch_measure_grouped_t(hipStream_t st, int stage, const MeasureArgs* args,
                             const int* start, int nprob, int nblocks, size_t shm) {
  if constexpr (T2 <= 3)
    if (stage == 2) return launch_measure_grouped_s<T2, 2>(st, args, start, nprob, nblocks, shm);
  if (stage != 0) return launch_measure_grouped_s<T2, 1>(st, args, start, nprob, nblocks, shm);
  return launch_measure_grouped_s<T2, 0>(st, args, start, nprob, nblocks, shm);
}

template <int T2, int STAGE>
int launch_measure_s(afqmc_ctx* c, int W, int n, const cplx* wk, int s, bool g64, bool gbf,
                     const WeightsArgs& tail, size_t shm) {
  constexpr int TI = meas_ti(T2);
  static int lds_limit = 0;
  const int rc = ensure_dyn_lds((const void*)measure_kernel<T2, TI, STAGE>, shm, &lds_limit);
  if (rc) return rc;
  hipLaunchKernelGGL((measure_kernel<T2, TI, STAGE>), dim3(W), dim3(64 * meas_waves(T2)), shm,
                     c->stream, c->d.norb, n, wk, c->trial[s], c->dets, s,
                     g64 ? c->green[s] : nullptr, gbf ? c->green_bf[s][0] : nullptr,
                     gbf ? c->green_bf[s][1] : nullptr, c->Np32, tail);
  LAUNCH_CHECK();
  return AFQMC_OK;
}

template <int T2>
int launch_measure_t(afqmc_ctx* c, int W, int n, const cplx* wk, int s, bool g64, bool gbf,
                     const WeightsArgs& tail) {
  int t2, stage;
  size_t shm;
  measure_plan(c, s, &t2, &stage, &shm);
  if constexpr (T2 <= 3)
    if (stage == 2) return launch_measure_s<T2, 2>(c, W, n, wk, s, g64, gbf, tail, shm);
  if (stage != 0) return launch_measure_s<T2, 1>(c, W, n, wk, s, g64, gbf, tail, shm);
  return launch_measure_s<T2, 0>(c, W, n, wk, s, g64, gbf, tail, shm);
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
int launch_measure_grouped(hipStream_t st, int t2, int stage, const MeasureArgs* args,
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

// The kernel launch_measure picks for spin s, as a plan for grouped launches: column tiles
// t2 of the Green product, the staging mode and the LDS bytes.
void measure_plan(const afqmc_ctx* c, int s, int* t2, int* stage, size_t* shm) {
  const int n = c->nocc[s], N = c->d.norb;
  const int t = (2 * n + 15) / 16;
  *t2 = t > 8 ? 8 : t;
  const size_t walker = (size_t)N * n * sizeof(cplx);
  if (*t2 <= 3) {  // two-wave kernel
    const size_t red = meas_red_bytes(*t2);
    const size_t kept = walker + gj_bytes(n) + (meas_ti(*t2) == 2 ? 0 : red);
    const size_t m_red = gj_bytes(n) + red;
    if (kept <= MEAS_STAGE_MAX) {  // walker kept beside O for the Green product
      *stage = 2;
      *shm = kept;
    } else if (walker <= MEAS_STAGE_MAX) {  // O, state and partials alias the walker
      *stage = 1;
      *shm = std::max(walker, m_red);
    } else {
      *stage = 0;
      *shm = m_red;
    }
    return;
  }
  const size_t base = gj_bytes(n);
  *stage = base + walker <= 24 * 1024 ? 1 : 0;  // small walkers only: staging costs occupancy
  *shm = *stage ? base + walker : base;
}
