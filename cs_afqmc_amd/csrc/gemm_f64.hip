// FP64 MFMA GEMM for gfx950:  C[m][n] = alpha * sum_k A[m][k] * B[n][k] + beta * C[m][n]
// (row-major, both operands K-contiguous).  Used for the Cholesky contractions of the
// step: force bias (split Green x rot_chol^T) and VHS (i sqrt(dt) x~ x chol).
//
// v_mfma_f64_16x16x4f64 lane layout (measured on gfx950):
//   A operand: lane l holds A[l % 16][l / 16]
//   B operand: lane l holds B[l / 16][l % 16]
//   D (double4): lane l, element r holds D[l / 16 + 4 r][l % 16]
#include "common.h"

namespace {

constexpr int GEMM_BK = 16;
constexpr int GEMM_PAD = 16;  // row stride == 16 (mod 32) doubles: k and k+1 rows hit disjoint bank halves

template <int BM, int BN, int WM, int WN>
__global__ __launch_bounds__(256) void gemm_nt_f64_kernel(
    int M, int N, int K, double alpha, const double* __restrict__ A, int lda,
    const double* __restrict__ B, int ldb, double beta, double* __restrict__ C, int ldc,
    int tiles_m, int tiles_n) {
  constexpr int NT = 256;
  constexpr int BK = GEMM_BK;
  constexpr int TM = BM / WM / 16;
  constexpr int TN = BN / WN / 16;
  constexpr int A_PER_T = BM * BK / NT;  // doubles of A loaded per thread per k-tile
  constexpr int B_PER_T = BN * BK / NT;
  static_assert(WM * WN == 4, "four waves");
  static_assert(A_PER_T >= 1 && B_PER_T >= 1, "tile too small");

  __shared__ double As[2][BK][BM + GEMM_PAD];
  __shared__ double Bs[2][BK][BN + GEMM_PAD];

  // XCD-aware block -> tile map: blocks are dealt round-robin to the 8 XCDs, so give
  // each XCD a contiguous run of tiles (m fastest: the run shares B column panels).
  const int nblk = tiles_m * tiles_n;
  int b = blockIdx.x;
  if ((nblk & 7) == 0) b = (b & 7) * (nblk >> 3) + (b >> 3);
  const int tile_m = b % tiles_m, tile_n = b / tiles_m;
  const int m0 = tile_m * BM, n0 = tile_n * BN;

  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int wm = wave % WM, wn = wave / WM;

  // global load mapping: 16 consecutive threads read 16 consecutive k of one row
  const int lk = t % BK, lr = t / BK;  // lr in [0, NT/BK)
  constexpr int RSTEP = NT / BK;

  double ra[A_PER_T], rb[B_PER_T];
  auto load_tile = [&](int k0) {
    const int k = k0 + lk;
#pragma unroll
    for (int i = 0; i < A_PER_T; ++i) {
      const int r = m0 + lr + i * RSTEP;
      ra[i] = (r < M && k < K) ? A[(int64_t)r * lda + k] : 0.0;
    }
#pragma unroll
    for (int i = 0; i < B_PER_T; ++i) {
      const int r = n0 + lr + i * RSTEP;
      rb[i] = (r < N && k < K) ? B[(int64_t)r * ldb + k] : 0.0;
    }
  };
  auto store_tile = [&](int buf) {
#pragma unroll
    for (int i = 0; i < A_PER_T; ++i) As[buf][lk][lr + i * RSTEP] = ra[i];
#pragma unroll
    for (int i = 0; i < B_PER_T; ++i) Bs[buf][lk][lr + i * RSTEP] = rb[i];
  };

  double4_t acc[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) acc[i][j] = double4_t{0.0, 0.0, 0.0, 0.0};

  const int nk = (K + BK - 1) / BK;
  load_tile(0);
  store_tile(0);
  __syncthreads();
  const int am = wm * (BM / WM) + (lane & 15);
  const int bn = wn * (BN / WN) + (lane & 15);
  const int kq = lane >> 4;
  for (int kt = 0; kt < nk; ++kt) {
    const int cur = kt & 1;
    if (kt + 1 < nk) load_tile((kt + 1) * BK);
#pragma unroll
    for (int ks = 0; ks < BK / 4; ++ks) {
      double af[TM], bf[TN];
#pragma unroll
      for (int i = 0; i < TM; ++i) af[i] = As[cur][ks * 4 + kq][am + i * 16];
#pragma unroll
      for (int j = 0; j < TN; ++j) bf[j] = Bs[cur][ks * 4 + kq][bn + j * 16];
#pragma unroll
      for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(af[i], bf[j], acc[i][j], 0, 0, 0);
    }
    if (kt + 1 < nk) store_tile(cur ^ 1);
    __syncthreads();
  }

  // epilogue
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) {
      const int col = n0 + wn * (BN / WN) + j * 16 + (lane & 15);
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = m0 + wm * (BM / WM) + i * 16 + (lane >> 4) + 4 * r;
        if (row < M && col < N) {
          double* p = C + (int64_t)row * ldc + col;
          const double v = alpha * acc[i][j][r];
          *p = (beta == 0.0) ? v : v + beta * (*p);
        }
      }
    }
}

}  // namespace

int launch_gemm_nt_f64(hipStream_t s, int M, int N, int K, double alpha, const double* A,
                       int lda, const double* B, int ldb, double beta, double* C, int ldc) {
  if (M <= 0 || N <= 0) return AFQMC_OK;
  if (K <= 0) return afqmc_set_error(AFQMC_ERR_ARG, "gemm: K=%d", K);
  // large problems: 128x128 tiles; otherwise 64x64 so that the grid covers the 256 CUs
  const int64_t big_tiles = (int64_t)((M + 127) / 128) * ((N + 127) / 128);
  if (big_tiles >= 512) {
    const int tm = (M + 127) / 128, tn = (N + 127) / 128;
    hipLaunchKernelGGL((gemm_nt_f64_kernel<128, 128, 2, 2>), dim3(tm * tn), dim3(256), 0, s,
                       M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, tm, tn);
  } else {
    const int tm = (M + 63) / 64, tn = (N + 63) / 64;
    hipLaunchKernelGGL((gemm_nt_f64_kernel<64, 64, 2, 2>), dim3(tm * tn), dim3(256), 0, s, M,
                       N, K, alpha, A, lda, B, ldb, beta, C, ldc, tm, tn);
  }
  LAUNCH_CHECK();
  return AFQMC_OK;
}
