// FP64 MFMA GEMM for gfx950:  C[m][n] = alpha * sum_k A[m][k] * B[n][k] + beta * C[m][n]
// (row-major, both operands K-contiguous).  Used for the Cholesky contractions of the
// step: force bias (split Green x rot_chol^T), VHS (i sqrt(dt) x~ x chol) and the
// local-energy f_g matrices (rot_chol x Green^T).
//
// v_mfma_f64_16x16x4f64 lane layout (measured on gfx950):
//   A operand: lane l holds A[l % 16][l / 16]
//   B operand: lane l holds B[l / 16][l % 16]
//   D (double4): lane l, element r holds D[l / 16 + 4 r][l % 16]
//
// Tiles of BM x BN with BK = 16, four waves (2 x 2), operands staged k-major in LDS
// (double buffered, register prefetch of the next k-tile).  Optional split-K: grid.y
// splits the K range and each split writes alpha * partial product to its own slice
// C + split * c_split_stride (beta must be 0); consumers sum the slices in order.
#include "common.h"

namespace {

constexpr int GEMM_BK = 16;
constexpr int GEMM_PAD = 16;  // row stride == 16 (mod 32) doubles: k and k+1 rows hit disjoint bank halves
// Element (k, m) of a k-major LDS tile sits at column m ^ gemm_swz<VEC>(k).  The MFMA
// operand reads (16 consecutive m at k and k + 1 per 32 lanes) see the same banks as
// without it: k and k + 1 share the swizzle, which keeps 16-aligned groups of m
// together.  The staging stores (8 or 16 k rows x 4 or 2 m per 32 lanes) without it put
// the rows of one k parity on one bank pair, an 8-way conflict; with it the 32 lanes
// meet 32 bank pairs.
template <int VEC>
__device__ inline int gemm_swz(int k) { return ((k >> 1) & 7) << (VEC == 2 ? 2 : 1); }

template <int BM, int BN, int WM, int WN, int VEC>
__device__ __forceinline__ void gemm_nt_f64_body(
    int M, int N, int K, double alpha, const double* __restrict__ A, int lda,
    const double* __restrict__ B, int ldb, double beta, double* __restrict__ C, int ldc,
    int tiles_m, int tiles_n, int k_per_split, int64_t c_split_stride, const int bx,
    const int by) {
  constexpr int NT = 256;
  constexpr int BK = GEMM_BK;
  constexpr int TM = BM / WM / 16;
  constexpr int TN = BN / WN / 16;
  constexpr int KT = BK / VEC;             // threads along k per row
  constexpr int RSTEP = NT / KT;           // rows covered per load pass
  constexpr int A_LD = BM / RSTEP;         // vector loads of A per thread per k-tile
  constexpr int B_LD = BN / RSTEP;
  static_assert(WM * WN == 4, "four waves");
  static_assert(A_LD >= 1 && B_LD >= 1, "tile too small");

  __shared__ double As[2][BK][BM + GEMM_PAD];
  __shared__ double Bs[2][BK][BN + GEMM_PAD];

  // XCD-aware block -> tile map (contiguous tile runs per XCD; the larger operand's
  // panels are the shared ones)
  int tile_m, tile_n;
  const bool m_major = M > N;
  const int band = xcd_band((int64_t)(m_major ? N : M) * k_per_split * 8,
                            (int64_t)(m_major ? BM : BN) * k_per_split * 8);
  xcd_tile(bx, tiles_m, tiles_n, m_major, band, &tile_m, &tile_n);
  const int m0 = tile_m * BM, n0 = tile_n * BN;
  const int kbeg = by * k_per_split;
  const int kend = (kbeg + k_per_split < K) ? kbeg + k_per_split : K;
  C += (int64_t)by * c_split_stride;

  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int wm = wave % WM, wn = wave / WM;
  const int lk = (t % KT) * VEC, lr = t / KT;

  double ra[A_LD][VEC], rb[B_LD][VEC];
  auto load_tile = [&](int k0) {
    const int k = k0 + lk;
    const bool kok = k < kend;  // VEC == 2 needs even k ranges (checked at launch)
#pragma unroll
    for (int i = 0; i < A_LD; ++i) {
      const int r = m0 + lr + i * RSTEP;
      if (VEC == 2) {
        double2 v = make_double2(0.0, 0.0);
        if (r < M && kok) v = *reinterpret_cast<const double2*>(A + (int64_t)r * lda + k);
        ra[i][0] = v.x;
        ra[i][VEC - 1] = v.y;
      } else {
        ra[i][0] = (r < M && kok) ? A[(int64_t)r * lda + k] : 0.0;
      }
    }
#pragma unroll
    for (int i = 0; i < B_LD; ++i) {
      const int r = n0 + lr + i * RSTEP;
      if (VEC == 2) {
        double2 v = make_double2(0.0, 0.0);
        if (r < N && kok) v = *reinterpret_cast<const double2*>(B + (int64_t)r * ldb + k);
        rb[i][0] = v.x;
        rb[i][VEC - 1] = v.y;
      } else {
        rb[i][0] = (r < N && kok) ? B[(int64_t)r * ldb + k] : 0.0;
      }
    }
  };
  auto store_tile = [&](int buf) {
#pragma unroll
    for (int i = 0; i < A_LD; ++i)
#pragma unroll
      for (int v = 0; v < VEC; ++v)
        As[buf][lk + v][(lr + i * RSTEP) ^ gemm_swz<VEC>(lk + v)] = ra[i][v];
#pragma unroll
    for (int i = 0; i < B_LD; ++i)
#pragma unroll
      for (int v = 0; v < VEC; ++v)
        Bs[buf][lk + v][(lr + i * RSTEP) ^ gemm_swz<VEC>(lk + v)] = rb[i][v];
  };

  double4_t acc[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) acc[i][j] = double4_t{0.0, 0.0, 0.0, 0.0};

  const int nk = (kend - kbeg + BK - 1) / BK;
  if (nk > 0) {
    load_tile(kbeg);
    store_tile(0);
  }
  __syncthreads();
  const int am = wm * (BM / WM) + (lane & 15);
  const int bn = wn * (BN / WN) + (lane & 15);
  const int kq = lane >> 4;
  for (int kt = 0; kt < nk; ++kt) {
    const int cur = kt & 1;
    if (kt + 1 < nk) load_tile(kbeg + (kt + 1) * BK);
#pragma unroll
    for (int ks = 0; ks < BK / 4; ++ks) {
      double af[TM], bf[TN];
      const int sw = gemm_swz<VEC>(ks * 4 + kq);
#pragma unroll
      for (int i = 0; i < TM; ++i) af[i] = As[cur][ks * 4 + kq][(am + i * 16) ^ sw];
#pragma unroll
      for (int j = 0; j < TN; ++j) bf[j] = Bs[cur][ks * 4 + kq][(bn + j * 16) ^ sw];
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

template <int BM, int BN, int WM, int WN, int VEC>
__global__ __launch_bounds__(256) void gemm_nt_f64_kernel(
    int M, int N, int K, double alpha, const double* __restrict__ A, int lda,
    const double* __restrict__ B, int ldb, double beta, double* __restrict__ C, int ldc,
    int tiles_m, int tiles_n, int k_per_split, int64_t c_split_stride) {
  gemm_nt_f64_body<BM, BN, WM, WN, VEC>(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, tiles_m,
                                        tiles_n, k_per_split, c_split_stride, blockIdx.x,
                                        blockIdx.y);
}

// grouped: problem p owns tiles [start[p], start[p + 1]), no split-K, beta = 0
template <int BM, int BN, int VEC>
__global__ __launch_bounds__(256) void gemm_f64_grouped_kernel(const GemmArgs* __restrict__ args,
                                                               const int* __restrict__ start,
                                                               int nprob) {
  const int p = group_of(start, nprob, blockIdx.x);
  const GemmArgs a = args[p];
  gemm_nt_f64_body<BM, BN, 2, 2, VEC>(a.M, a.N, a.K, (double)a.alpha,
                                      static_cast<const double*>(a.A), (int)a.lda,
                                      static_cast<const double*>(a.B), (int)a.ldb, 0.0,
                                      static_cast<double*>(a.C), (int)a.ldc, a.tiles_m, a.tiles_n,
                                      (a.K + GEMM_BK - 1) / GEMM_BK * GEMM_BK, 0,
                                      blockIdx.x - start[p], 0);
}

template <int BM, int BN>
int launch_tiles(hipStream_t s, int M, int N, int K, double alpha, const double* A, int lda,
                 const double* B, int ldb, double beta, double* C, int ldc, int splits,
                 int k_per_split, int64_t c_split_stride, bool vec2) {
  const int tm = (M + BM - 1) / BM, tn = (N + BN - 1) / BN;
  dim3 grid(tm * tn, splits);
  if (vec2)
    hipLaunchKernelGGL((gemm_nt_f64_kernel<BM, BN, 2, 2, 2>), grid, dim3(256), 0, s, M, N, K,
                       alpha, A, lda, B, ldb, beta, C, ldc, tm, tn, k_per_split,
                       c_split_stride);
  else
    hipLaunchKernelGGL((gemm_nt_f64_kernel<BM, BN, 2, 2, 1>), grid, dim3(256), 0, s, M, N, K,
                       alpha, A, lda, B, ldb, beta, C, ldc, tm, tn, k_per_split,
                       c_split_stride);
  LAUNCH_CHECK();
  return AFQMC_OK;
}

}  // namespace

int gemm_f64_splits(int M, int N, int K) {
  // split K only when the 64x64 tile grid cannot fill the 256 CUs twice over
  const int64_t tiles = (int64_t)((M + 63) / 64) * ((N + 63) / 64);
  if (tiles >= 512) return 1;
  int sp = (int)((512 + tiles - 1) / tiles);
  const int max_sp = K / 256;  // keep >= 256 k per split
  if (sp > max_sp) sp = max_sp;
  if (sp > 8) sp = 8;
  return sp < 1 ? 1 : sp;
}

int launch_gemm_nt_f64_split(hipStream_t s, int M, int N, int K, double alpha, const double* A,
                             int lda, const double* B, int ldb, double* C, int ldc, int splits,
                             int64_t c_split_stride) {
  if (M <= 0 || N <= 0) return AFQMC_OK;
  if (K <= 0) return afqmc_set_error(AFQMC_ERR_ARG, "gemm: K=%d", K);
  if (splits < 1) splits = 1;
  int kps = (K + splits - 1) / splits;
  kps = (kps + GEMM_BK - 1) / GEMM_BK * GEMM_BK;
  const bool vec2 = ((lda | ldb | K) % 2 == 0) && ((uintptr_t)A % 16 == 0) &&
                    ((uintptr_t)B % 16 == 0);
  const int64_t big_tiles = (int64_t)((M + 127) / 128) * ((N + 127) / 128);
  if (big_tiles * splits >= 512)
    return launch_tiles<128, 128>(s, M, N, K, alpha, A, lda, B, ldb, 0.0, C, ldc, splits, kps,
                                  c_split_stride, vec2);
  return launch_tiles<64, 64>(s, M, N, K, alpha, A, lda, B, ldb, 0.0, C, ldc, splits, kps,
                              c_split_stride, vec2);
}

int launch_gemm_nt_f64(hipStream_t s, int M, int N, int K, double alpha, const double* A,
                       int lda, const double* B, int ldb, double beta, double* C, int ldc) {
  if (M <= 0 || N <= 0) return AFQMC_OK;
  if (K <= 0) return afqmc_set_error(AFQMC_ERR_ARG, "gemm: K=%d", K);
  const bool vec2 = ((lda | ldb | K) % 2 == 0) && ((uintptr_t)A % 16 == 0) &&
                    ((uintptr_t)B % 16 == 0);
  const int64_t big_tiles = (int64_t)((M + 127) / 128) * ((N + 127) / 128);
  if (big_tiles >= 512)
    return launch_tiles<128, 128>(s, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, 1, K, 0,
                                  vec2);
  return launch_tiles<64, 64>(s, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, 1, K, 0, vec2);
}

int launch_gemm_f64_grouped(hipStream_t s, int bm, bool vec2, const GemmArgs* args,
                            const int* start, int nprob, int nblocks) {
  if (bm == 128) {
    if (vec2)
      hipLaunchKernelGGL((gemm_f64_grouped_kernel<128, 128, 2>), dim3(nblocks), dim3(256), 0, s,
                         args, start, nprob);
    else
      hipLaunchKernelGGL((gemm_f64_grouped_kernel<128, 128, 1>), dim3(nblocks), dim3(256), 0, s,
                         args, start, nprob);
  } else {
    if (vec2)
      hipLaunchKernelGGL((gemm_f64_grouped_kernel<64, 64, 2>), dim3(nblocks), dim3(256), 0, s,
                         args, start, nprob);
    else
      hipLaunchKernelGGL((gemm_f64_grouped_kernel<64, 64, 1>), dim3(nblocks), dim3(256), 0, s,
                         args, start, nprob);
  }
  LAUNCH_CHECK();
  return AFQMC_OK;
}

namespace {

// dst (cols x rows) = src (rows x cols)^T through 32 x 32 LDS tiles
__global__ __launch_bounds__(256) void transpose_f64_kernel(int rows, int cols,
                                                           const double* __restrict__ src,
                                                           double* __restrict__ dst) {
  __shared__ double tile[32][33];
  const int r0 = blockIdx.y * 32, c0 = blockIdx.x * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
  for (int j = ty; j < 32; j += 8) {
    const int r = r0 + j, c = c0 + tx;
    if (r < rows && c < cols) tile[j][tx] = src[(int64_t)r * cols + c];
  }
  __syncthreads();
  for (int j = ty; j < 32; j += 8) {
    const int c = c0 + j, r = r0 + tx;
    if (r < rows && c < cols) dst[(int64_t)c * rows + r] = tile[tx][j];
  }
}

}  // namespace

int launch_transpose_f64(hipStream_t s, int rows, int cols, const double* src, double* dst) {
  if (rows <= 0 || cols <= 0) return AFQMC_OK;
  dim3 grid((cols + 31) / 32, (rows + 31) / 32);
  if (grid.y > 65535) return afqmc_set_error(AFQMC_ERR_SIZE, "transpose: %d rows", rows);
  hipLaunchKernelGGL(transpose_f64_kernel, grid, dim3(256), 0, s, rows, cols, src, dst);
  LAUNCH_CHECK();
  return AFQMC_OK;
}
