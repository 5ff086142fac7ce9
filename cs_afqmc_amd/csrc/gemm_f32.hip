// FP32-MFMA GEMM for gfx950 on float64 operands:
//   C[m][n] = alpha * sum_k fl32(A[m][k]) * fl32(B[n][k]) + beta * C[m][n]
// (row-major, both operands K-contiguous, C float64).  The contractions of the FP32
// precision mode (force bias, VHS) read the same float64 operands as the float64 path;
// each element is rounded to float32 as it is staged into LDS and the products run on
// v_mfma_f32_16x16x4f32 with float32 accumulation, so the mode needs no extra copies of
// the Cholesky vectors or the Green's function.
//
// v_mfma_f32_16x16x4f32 lane layout (as v_mfma_f64_16x16x4f64 for the operands, as the
// other float32-accumulating 16 x 16 MFMAs for the result):
//   A operand: lane l holds A[l % 16][l / 16]
//   B operand: lane l holds B[l / 16][l % 16]   (B as k x n)
//   D (float4): lane l, element r holds D[4 (l / 16) + r][l % 16]
//
// Tiles of BM x BN with BK = 16, four waves (2 x 2), float operands k-major in LDS
// (double buffered, register prefetch of the next k-tile).  Split-K as in the float64
// GEMM: grid.y splits K and split z writes alpha * its partial product to
// C + z * c_split_stride (beta must be 0); consumers add the slices in order.
#include "common.h"

namespace {

constexpr int F32_BK = 16;
constexpr int F32_PAD = 16;  // row stride == 16 (mod 32) floats: rows k, k+1 use disjoint banks
// Element (k, m) of a k-major LDS tile sits at column m ^ f32_swz<VEC>(k).  The MFMA
// operand reads (16 consecutive m at k = 4 s .. 4 s + 3 per wave) share the swizzle and
// see the same banks as without it; the staging stores (16 or 8 k rows x 4 or 8 m per
// wave) without it put four k rows on one bank, a 4-way conflict, and with it the 64
// lanes meet 64 banks.
template <int VEC>
__device__ inline int f32_swz(int k) { return ((k >> 2) & 3) << (VEC == 2 ? 3 : 2); }

template <int BM, int BN, int WM, int WN, int VEC>
__device__ __forceinline__ void gemm_nt_f32_body(
    int M, int N, int K, double alpha, const double* __restrict__ A, int lda,
    const double* __restrict__ B, int ldb, double beta, double* __restrict__ C, int ldc,
    int tiles_m, int tiles_n, int k_per_split, int64_t c_split_stride, const int bx,
    const int by) {
  constexpr int NT = 256;
  constexpr int BK = F32_BK;
  constexpr int TM = BM / WM / 16;
  constexpr int TN = BN / WN / 16;
  constexpr int KT = BK / VEC;
  constexpr int RSTEP = NT / KT;
  constexpr int A_LD = BM / RSTEP;
  constexpr int B_LD = BN / RSTEP;
  static_assert(WM * WN == 4, "four waves");
  static_assert(A_LD >= 1 && B_LD >= 1, "tile too small");

  __shared__ float As[2][BK][BM + F32_PAD];
  __shared__ float Bs[2][BK][BN + F32_PAD];

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

  float ra[A_LD][VEC], rb[B_LD][VEC];
  auto load_tile = [&](int k0) {
    const int k = k0 + lk;
    const bool kok = k < kend;  // VEC == 2 needs even k ranges (checked at launch)
#pragma unroll
    for (int i = 0; i < A_LD; ++i) {
      const int r = m0 + lr + i * RSTEP;
      if (VEC == 2) {
        double2 v = make_double2(0.0, 0.0);
        if (r < M && kok) v = *reinterpret_cast<const double2*>(A + (int64_t)r * lda + k);
        ra[i][0] = (float)v.x;
        ra[i][VEC - 1] = (float)v.y;
      } else {
        ra[i][0] = (r < M && kok) ? (float)A[(int64_t)r * lda + k] : 0.0f;
      }
    }
#pragma unroll
    for (int i = 0; i < B_LD; ++i) {
      const int r = n0 + lr + i * RSTEP;
      if (VEC == 2) {
        double2 v = make_double2(0.0, 0.0);
        if (r < N && kok) v = *reinterpret_cast<const double2*>(B + (int64_t)r * ldb + k);
        rb[i][0] = (float)v.x;
        rb[i][VEC - 1] = (float)v.y;
      } else {
        rb[i][0] = (r < N && kok) ? (float)B[(int64_t)r * ldb + k] : 0.0f;
      }
    }
  };
  auto store_tile = [&](int buf) {
#pragma unroll
    for (int i = 0; i < A_LD; ++i)
#pragma unroll
      for (int v = 0; v < VEC; ++v)
        As[buf][lk + v][(lr + i * RSTEP) ^ f32_swz<VEC>(lk + v)] = ra[i][v];
#pragma unroll
    for (int i = 0; i < B_LD; ++i)
#pragma unroll
      for (int v = 0; v < VEC; ++v)
        Bs[buf][lk + v][(lr + i * RSTEP) ^ f32_swz<VEC>(lk + v)] = rb[i][v];
  };

  float4_t acc[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) acc[i][j] = float4_t{0.f, 0.f, 0.f, 0.f};

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
      float af[TM], bf[TN];
      const int sw = f32_swz<VEC>(ks * 4 + kq);
#pragma unroll
      for (int i = 0; i < TM; ++i) af[i] = As[cur][ks * 4 + kq][(am + i * 16) ^ sw];
#pragma unroll
      for (int j = 0; j < TN; ++j) bf[j] = Bs[cur][ks * 4 + kq][(bn + j * 16) ^ sw];
#pragma unroll
      for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(af[i], bf[j], acc[i][j], 0, 0, 0);
    }
    if (kt + 1 < nk) store_tile(cur ^ 1);
    __syncthreads();
  }

#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) {
      const int col = n0 + wn * (BN / WN) + j * 16 + (lane & 15);
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = m0 + wm * (BM / WM) + i * 16 + 4 * (lane >> 4) + r;
        if (row < M && col < N) {
          double* p = C + (int64_t)row * ldc + col;
          const double v = alpha * (double)acc[i][j][r];
          *p = (beta == 0.0) ? v : v + beta * (*p);
        }
      }
    }
}

template <int BM, int BN, int VEC>
__global__ __launch_bounds__(256) void gemm_nt_f32_kernel(
    int M, int N, int K, double alpha, const double* __restrict__ A, int lda,
    const double* __restrict__ B, int ldb, double beta, double* __restrict__ C, int ldc,
    int tiles_m, int tiles_n, int k_per_split, int64_t c_split_stride) {
  gemm_nt_f32_body<BM, BN, 2, 2, VEC>(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, tiles_m,
                                      tiles_n, k_per_split, c_split_stride, blockIdx.x,
                                      blockIdx.y);
}

// grouped: problem p owns tiles [start[p], start[p + 1]), no split-K, beta = 0
template <int BM, int BN, int VEC>
__global__ __launch_bounds__(256) void gemm_f32_grouped_kernel(const GemmArgs* __restrict__ args,
                                                               const int* __restrict__ start,
                                                               int nprob) {
  const int p = group_of(start, nprob, blockIdx.x);
  const GemmArgs a = args[p];
  gemm_nt_f32_body<BM, BN, 2, 2, VEC>(a.M, a.N, a.K, (double)a.alpha,
                                      static_cast<const double*>(a.A), (int)a.lda,
                                      static_cast<const double*>(a.B), (int)a.ldb, 0.0,
                                      static_cast<double*>(a.C), (int)a.ldc, a.tiles_m, a.tiles_n,
                                      (a.K + F32_BK - 1) / F32_BK * F32_BK, 0,
                                      blockIdx.x - start[p], 0);
}

template <int BM, int BN>
int launch_f32_tiles(hipStream_t s, int M, int N, int K, double alpha, const double* A, int lda,
                     const double* B, int ldb, double beta, double* C, int ldc, int splits,
                     int k_per_split, int64_t c_split_stride, bool vec2) {
  const int tm = (M + BM - 1) / BM, tn = (N + BN - 1) / BN;
  dim3 grid(tm * tn, splits);
  if (vec2)
    hipLaunchKernelGGL((gemm_nt_f32_kernel<BM, BN, 2>), grid, dim3(256), 0, s, M, N, K, alpha,
                       A, lda, B, ldb, beta, C, ldc, tm, tn, k_per_split, c_split_stride);
  else
    hipLaunchKernelGGL((gemm_nt_f32_kernel<BM, BN, 1>), grid, dim3(256), 0, s, M, N, K, alpha,
                       A, lda, B, ldb, beta, C, ldc, tm, tn, k_per_split, c_split_stride);
  LAUNCH_CHECK();
  return AFQMC_OK;
}

bool f32_vec2(int lda, int ldb, int K, const double* A, const double* B) {
  return ((lda | ldb | K) % 2 == 0) && ((uintptr_t)A % 16 == 0) && ((uintptr_t)B % 16 == 0);
}

}  // namespace

int launch_gemm_nt_f32_split(hipStream_t s, int M, int N, int K, double alpha, const double* A,
                             int lda, const double* B, int ldb, double* C, int ldc, int splits,
                             int64_t c_split_stride) {
  if (M <= 0 || N <= 0) return AFQMC_OK;
  if (K <= 0) return afqmc_set_error(AFQMC_ERR_ARG, "gemm: K=%d", K);
  if (splits < 1) splits = 1;
  int kps = (K + splits - 1) / splits;
  kps = (kps + F32_BK - 1) / F32_BK * F32_BK;
  const bool vec2 = f32_vec2(lda, ldb, K, A, B);
  const int64_t big_tiles = (int64_t)((M + 127) / 128) * ((N + 127) / 128);
  if (big_tiles * splits >= 512)
    return launch_f32_tiles<128, 128>(s, M, N, K, alpha, A, lda, B, ldb, 0.0, C, ldc, splits,
                                      kps, c_split_stride, vec2);
  return launch_f32_tiles<64, 64>(s, M, N, K, alpha, A, lda, B, ldb, 0.0, C, ldc, splits, kps,
                                  c_split_stride, vec2);
}

int launch_gemm_nt_f32(hipStream_t s, int M, int N, int K, double alpha, const double* A,
                       int lda, const double* B, int ldb, double beta, double* C, int ldc) {
  if (M <= 0 || N <= 0) return AFQMC_OK;
  if (K <= 0) return afqmc_set_error(AFQMC_ERR_ARG, "gemm: K=%d", K);
  const bool vec2 = f32_vec2(lda, ldb, K, A, B);
  const int64_t big_tiles = (int64_t)((M + 127) / 128) * ((N + 127) / 128);
  if (big_tiles >= 512)
    return launch_f32_tiles<128, 128>(s, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, 1, K, 0,
                                      vec2);
  return launch_f32_tiles<64, 64>(s, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, 1, K, 0,
                                  vec2);
}

int launch_gemm_f32_grouped(hipStream_t s, int bm, bool vec2, const GemmArgs* args,
                            const int* start, int nprob, int nblocks) {
  if (bm == 128) {
    if (vec2)
      hipLaunchKernelGGL((gemm_f32_grouped_kernel<128, 128, 2>), dim3(nblocks), dim3(256), 0, s,
                         args, start, nprob);
    else
      hipLaunchKernelGGL((gemm_f32_grouped_kernel<128, 128, 1>), dim3(nblocks), dim3(256), 0, s,
                         args, start, nprob);
  } else {
    if (vec2)
      hipLaunchKernelGGL((gemm_f32_grouped_kernel<64, 64, 2>), dim3(nblocks), dim3(256), 0, s,
                         args, start, nprob);
    else
      hipLaunchKernelGGL((gemm_f32_grouped_kernel<64, 64, 1>), dim3(nblocks), dim3(256), 0, s,
                         args, start, nprob);
  }
  LAUNCH_CHECK();
  return AFQMC_OK;
}
