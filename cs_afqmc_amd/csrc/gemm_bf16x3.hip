// Split-bf16 ("bf16x3") GEMM for gfx950 on v_mfma_f32_16x16x32_bf16:
//   C[m][n] = alpha * sum_k A[m][k] * B[n][k]   (row-major, both operands K-contiguous)
// Each operand x is stored as two bf16 planes, x = hi + lo with hi = bf16(x) and
// lo = bf16(x - hi) (round to nearest even), so the pair carries 16 mantissa bits.  The
// product keeps the three leading terms  Ah Bh + Ah Bl + Al Bh  (the dropped Al Bl is
// ~2^-16 relative) with float32 accumulation in the MFMA.  Used for the Cholesky
// contractions of the mixed-precision step (VHS, force bias, local-energy f_g).
//
// v_mfma_f32_16x16x32_bf16 lane layout (measured on gfx950):
//   A: lane l holds A[l % 16][8 (l / 16) + e], e = 0..7      (16 bytes, K-contiguous)
//   B: lane l holds B[8 (l / 16) + e][l % 16]                  (= row l%16 of an NT B)
//   D (float4): lane l, element r holds D[4 (l / 16) + r][l % 16]
// Operand planes have leading dimensions that are multiples of 32 elements, zero padded
// in k, so K (a multiple of 32) needs no tail handling.
#include "common.h"

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef unsigned int u32x4_t __attribute__((ext_vector_type(4)));

namespace {

constexpr int BK = 32;
// LDS rows of BK bf16 (64 B), the four 16-byte chunks of row r stored in the order
// c ^ ((r >> 1) & 3).  Measured on gfx950 with 16-byte LDS reads: a wave reading 16 rows
// x 16 B at each of the four chunks (an MFMA operand) runs at the speed of a contiguous
// read with this order, and 1.6x slower with rows in plain order or with c ^ ((r >> 2) & 3);
// the tile stores (4 lanes per 64-byte row) cost the same in every order
constexpr int LDSW = BK;
__device__ inline int lds_chunk(int r, int c) { return (c ^ ((r >> 1) & 3)) << 3; }

__device__ inline uint16_t f32_to_bf16(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);  // round to nearest even (finite inputs)
  return (uint16_t)(u >> 16);
}
__device__ inline float bf16_to_f32(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// x (float64, rows x cols, ld_src) -> hi / lo bf16 planes (rows x ld_dst, zero padded)
__global__ void split_bf16_kernel(int rows, int cols, const double* __restrict__ x, int64_t ld_src,
                                  uint16_t* __restrict__ hi, uint16_t* __restrict__ lo,
                                  int64_t ld_dst) {
  const int64_t total = (int64_t)rows * ld_dst;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = e / ld_dst, c = e - r * ld_dst;
    uint16_t h = 0, l = 0;
    if (c < cols) {
      const double v = x[r * ld_src + c];
      h = f32_to_bf16((float)v);
      l = f32_to_bf16((float)(v - (double)bf16_to_f32(h)));
    }
    hi[e] = h;
    lo[e] = l;
  }
}

// Symmetry of the Cholesky vectors from their transposed copy chol_t (N*N rows, G):
// out[0] = max |L_g[p][k] - L_g[k][p]| and out[1] = max |L_g[p][k]| as the bit
// patterns of non-negative doubles (ordered like the values, so atomicMax applies).
__global__ void sym_check_kernel(int N, int G, const double* __restrict__ chol_t,
                                 unsigned long long* __restrict__ out) {
  double asym = 0.0, amax = 0.0;
  const int64_t total = (int64_t)N * N * G;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t pk = e / G, g = e - pk * G;
    const int p = (int)(pk / N), k = (int)(pk - (int64_t)p * N);
    const double a = chol_t[e], b = chol_t[((int64_t)k * N + p) * G + g];
    asym = fmax(asym, fabs(a - b));
    amax = fmax(amax, fabs(a));
  }
  for (int off = 32; off > 0; off >>= 1) {
    asym = fmax(asym, __shfl_xor(asym, off, 64));
    amax = fmax(amax, __shfl_xor(amax, off, 64));
  }
  if ((threadIdx.x & 63) == 0) {
    atomicMax(&out[0], (unsigned long long)__double_as_longlong(asym));
    atomicMax(&out[1], (unsigned long long)__double_as_longlong(amax));
  }
}

// Upper triangle (p <= k, row-major: row p holds k = p .. N-1) of symmetric Cholesky
// vectors -> hi / lo bf16 planes, one packed row per (p, k) pair, Gp columns (zero
// padded): the B operand of the VHS product when it forms the packed triangle of V.
__global__ void split_bf16_sym_kernel(int N, int G, const double* __restrict__ chol_t,
                                      uint16_t* __restrict__ hi, uint16_t* __restrict__ lo,
                                      int ld_dst) {
  const int p = blockIdx.x / N, k = blockIdx.x - (blockIdx.x / N) * N;
  if (k < p) return;
  const int64_t t = (int64_t)p * N - (int64_t)p * (p - 1) / 2 + (k - p);
  const double* src = chol_t + ((int64_t)p * N + k) * G;
  for (int g = threadIdx.x; g < ld_dst; g += blockDim.x) {
    uint16_t h = 0, l = 0;
    if (g < G) {
      const double v = src[g];
      h = f32_to_bf16((float)v);
      l = f32_to_bf16((float)(v - (double)bf16_to_f32(h)));
    }
    hi[t * ld_dst + g] = h;
    lo[t * ld_dst + g] = l;
  }
}

template <int BM, int BN, typename OutT>
__device__ __forceinline__ void gemm_nt_bf16x3_body(
    int M, int N, int K, float alpha, const uint16_t* __restrict__ Ah,
    const uint16_t* __restrict__ Al, int64_t lda, const uint16_t* __restrict__ Bh,
    const uint16_t* __restrict__ Bl, int64_t ldb, OutT* __restrict__ C, int64_t ldc, int tiles_m,
    int tiles_n, int k_per_split, int64_t c_split_stride, const int bx, const int by) {
  constexpr int TM = BM / 2 / 16, TN = BN / 2 / 16;  // 2 x 2 waves
  constexpr int A_LD = BM * 4 / 256, B_LD = BN * 4 / 256;  // 16-byte chunks per thread
  __shared__ __attribute__((aligned(16))) uint16_t sA[2][2][BM * LDSW];  // [buf][hi/lo]
  __shared__ __attribute__((aligned(16))) uint16_t sB[2][2][BN * LDSW];

  int tile_m, tile_n;
  // operand bytes: hi + lo planes, 4 per element over this split's K
  const bool m_major = M > N;
  const int band = xcd_band((int64_t)(m_major ? N : M) * k_per_split * 4,
                            (int64_t)(m_major ? BM : BN) * k_per_split * 4);
  xcd_tile(bx, tiles_m, tiles_n, m_major, band, &tile_m, &tile_n);
  const int m0 = tile_m * BM, n0 = tile_n * BN;
  const int kbeg = by * k_per_split;
  const int kend = (kbeg + k_per_split < K) ? kbeg + k_per_split : K;
  C += (int64_t)by * c_split_stride;

  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int wm = wave & 1, wn = wave >> 1;

  u32x4_t ra[2][A_LD], rb[2][B_LD];
  auto load_tile = [&](int k0) {
#pragma unroll
    for (int j = 0; j < A_LD; ++j) {
      const int q = t + 256 * j, row = q >> 2, ch = q & 3;
      const int r = m0 + row;
      u32x4_t vh = {0, 0, 0, 0}, vl = {0, 0, 0, 0};
      if (r < M) {
        vh = *reinterpret_cast<const u32x4_t*>(Ah + r * lda + k0 + ch * 8);
        vl = *reinterpret_cast<const u32x4_t*>(Al + r * lda + k0 + ch * 8);
      }
      ra[0][j] = vh;
      ra[1][j] = vl;
    }
#pragma unroll
    for (int j = 0; j < B_LD; ++j) {
      const int q = t + 256 * j, row = q >> 2, ch = q & 3;
      const int r = n0 + row;
      u32x4_t vh = {0, 0, 0, 0}, vl = {0, 0, 0, 0};
      if (r < N) {
        vh = *reinterpret_cast<const u32x4_t*>(Bh + r * ldb + k0 + ch * 8);
        vl = *reinterpret_cast<const u32x4_t*>(Bl + r * ldb + k0 + ch * 8);
      }
      rb[0][j] = vh;
      rb[1][j] = vl;
    }
  };
  auto store_tile = [&](int buf) {
#pragma unroll
    for (int j = 0; j < A_LD; ++j) {
      const int q = t + 256 * j, row = q >> 2, ch = q & 3;
#pragma unroll
      for (int h = 0; h < 2; ++h)
        *reinterpret_cast<u32x4_t*>(&sA[buf][h][row * LDSW + lds_chunk(row, ch)]) = ra[h][j];
    }
#pragma unroll
    for (int j = 0; j < B_LD; ++j) {
      const int q = t + 256 * j, row = q >> 2, ch = q & 3;
#pragma unroll
      for (int h = 0; h < 2; ++h)
        *reinterpret_cast<u32x4_t*>(&sB[buf][h][row * LDSW + lds_chunk(row, ch)]) = rb[h][j];
    }
  };

  float4_t acc[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) acc[i][j] = float4_t{0.f, 0.f, 0.f, 0.f};

  const int nk = (kend - kbeg) / BK;
  if (nk > 0) {
    load_tile(kbeg);
    store_tile(0);
  }
  __syncthreads();
  const int arow = wm * (BM / 2) + (lane & 15);
  const int brow = wn * (BN / 2) + (lane & 15);
  // k chunk lane >> 4 of rows whose bits 2-3 are those of lane (row = 16 t + lane % 16)
  const int koff = lds_chunk(lane & 15, lane >> 4);
  for (int kt = 0; kt < nk; ++kt) {
    const int cur = kt & 1;
    if (kt + 1 < nk) load_tile(kbeg + (kt + 1) * BK);
    bf16x8_t ah[TM], al[TM], bh[TN], bl[TN];
#pragma unroll
    for (int i = 0; i < TM; ++i) {
      const int o = (arow + i * 16) * LDSW + koff;
      ah[i] = __builtin_bit_cast(bf16x8_t, *reinterpret_cast<const u32x4_t*>(&sA[cur][0][o]));
      al[i] = __builtin_bit_cast(bf16x8_t, *reinterpret_cast<const u32x4_t*>(&sA[cur][1][o]));
    }
#pragma unroll
    for (int j = 0; j < TN; ++j) {
      const int o = (brow + j * 16) * LDSW + koff;
      bh[j] = __builtin_bit_cast(bf16x8_t, *reinterpret_cast<const u32x4_t*>(&sB[cur][0][o]));
      bl[j] = __builtin_bit_cast(bf16x8_t, *reinterpret_cast<const u32x4_t*>(&sB[cur][1][o]));
    }
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
      for (int j = 0; j < TN; ++j) {
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[i], bh[j], acc[i][j], 0, 0, 0);
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bl[j], acc[i][j], 0, 0, 0);
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bh[j], acc[i][j], 0, 0, 0);
      }
    if (kt + 1 < nk) store_tile(cur ^ 1);
    __syncthreads();
  }

#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) {
      const int col = n0 + wn * (BN / 2) + j * 16 + (lane & 15);
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = m0 + wm * (BM / 2) + i * 16 + 4 * (lane >> 4) + r;
        if (row < M && col < N) C[(int64_t)row * ldc + col] = (OutT)(alpha * acc[i][j][r]);
      }
    }
}

template <int BM, int BN, typename OutT>
__global__ __launch_bounds__(256) void gemm_nt_bf16x3_kernel(
    int M, int N, int K, float alpha, const uint16_t* __restrict__ Ah,
    const uint16_t* __restrict__ Al, int64_t lda, const uint16_t* __restrict__ Bh,
    const uint16_t* __restrict__ Bl, int64_t ldb, OutT* __restrict__ C, int64_t ldc, int tiles_m,
    int tiles_n, int k_per_split, int64_t c_split_stride) {
  gemm_nt_bf16x3_body<BM, BN, OutT>(M, N, K, alpha, Ah, Al, lda, Bh, Bl, ldb, C, ldc, tiles_m,
                                    tiles_n, k_per_split, c_split_stride, blockIdx.x, blockIdx.y);
}

// grouped: problem p owns tiles [start[p], start[p + 1]), no split-K
template <int BM, int BN, typename OutT>
__global__ __launch_bounds__(256) void gemm_bf16x3_grouped_kernel(const GemmArgs* __restrict__ args,
                                                                  const int* __restrict__ start,
                                                                  int nprob) {
  const int p = group_of(start, nprob, blockIdx.x);
  const GemmArgs a = args[p];
  gemm_nt_bf16x3_body<BM, BN, OutT>(
      a.M, a.N, a.K, a.alpha, static_cast<const uint16_t*>(a.A),
      static_cast<const uint16_t*>(a.Al), a.lda, static_cast<const uint16_t*>(a.B),
      static_cast<const uint16_t*>(a.Bl), a.ldb, static_cast<OutT*>(a.C), a.ldc, a.tiles_m,
      a.tiles_n, a.K, 0, blockIdx.x - start[p], 0);
}

template <int BM, int BN, typename OutT>
int launch_bf(hipStream_t s, int M, int N, int K, float alpha, const uint16_t* Ah,
              const uint16_t* Al, int64_t lda, const uint16_t* Bh, const uint16_t* Bl, int64_t ldb,
              OutT* C, int64_t ldc, int splits, int kps, int64_t c_split_stride) {
  const int tm = (M + BM - 1) / BM, tn = (N + BN - 1) / BN;
  hipLaunchKernelGGL((gemm_nt_bf16x3_kernel<BM, BN, OutT>), dim3(tm * tn, splits), dim3(256), 0,
                     s, M, N, K, alpha, Ah, Al, lda, Bh, Bl, ldb, C, ldc, tm, tn, kps,
                     c_split_stride);
  LAUNCH_CHECK();
  return AFQMC_OK;
}

template <typename OutT>
int gemm_bf16x3(hipStream_t s, int M, int N, int K, float alpha, const uint16_t* Ah,
                const uint16_t* Al, int64_t lda, const uint16_t* Bh, const uint16_t* Bl,
                int64_t ldb, OutT* C, int64_t ldc, int splits, int64_t c_split_stride) {
  if (M <= 0 || N <= 0) return AFQMC_OK;
  if (K <= 0 || K % BK || lda % 8 || ldb % 8)
    return afqmc_set_error(AFQMC_ERR_ARG, "bf16x3 gemm: K=%d lda=%lld ldb=%lld", K,
                           (long long)lda, (long long)ldb);
  if (splits < 1) splits = 1;
  int kps = (K / BK + splits - 1) / splits * BK;
  const int64_t big = (int64_t)((M + 127) / 128) * ((N + 127) / 128);
  if (big * splits >= 512)
    return launch_bf<128, 128, OutT>(s, M, N, K, alpha, Ah, Al, lda, Bh, Bl, ldb, C, ldc, splits,
                                     kps, c_split_stride);
  return launch_bf<64, 64, OutT>(s, M, N, K, alpha, Ah, Al, lda, Bh, Bl, ldb, C, ldc, splits, kps,
                                 c_split_stride);
}

}  // namespace

int launch_split_bf16(hipStream_t s, int rows, int cols, const double* x, int64_t ld_src,
                      uint16_t* hi, uint16_t* lo, int64_t ld_dst) {
  const int64_t total = (int64_t)rows * ld_dst;
  if (total <= 0) return AFQMC_OK;
  const int64_t blocks = (total + 255) / 256 < 65536 ? (total + 255) / 256 : 65536;
  hipLaunchKernelGGL(split_bf16_kernel, dim3((unsigned)blocks), dim3(256), 0, s, rows, cols, x,
                     ld_src, hi, lo, ld_dst);
  LAUNCH_CHECK();
  return AFQMC_OK;
}

int launch_sym_check(hipStream_t s, int N, int G, const double* chol_t,
                     unsigned long long* out) {
  const int64_t total = (int64_t)N * N * G;
  const int64_t blocks = (total + 255) / 256 < 4096 ? (total + 255) / 256 : 4096;
  hipLaunchKernelGGL(sym_check_kernel, dim3((unsigned)blocks), dim3(256), 0, s, N, G, chol_t, out);
  LAUNCH_CHECK();
  return AFQMC_OK;
}

int launch_split_bf16_sym(hipStream_t s, int N, int G, const double* chol_t, uint16_t* hi,
                          uint16_t* lo, int ld_dst) {
  hipLaunchKernelGGL(split_bf16_sym_kernel, dim3((unsigned)(N * N)), dim3(256), 0, s, N, G,
                     chol_t, hi, lo, ld_dst);
  LAUNCH_CHECK();
  return AFQMC_OK;
}

int launch_gemm_nt_bf16x3_f32(hipStream_t s, int M, int N, int K, float alpha,
                              const uint16_t* Ah, const uint16_t* Al, int64_t lda,
                              const uint16_t* Bh, const uint16_t* Bl, int64_t ldb, float* C,
                              int64_t ldc, int splits, int64_t c_split_stride) {
  return gemm_bf16x3<float>(s, M, N, K, alpha, Ah, Al, lda, Bh, Bl, ldb, C, ldc, splits,
                            c_split_stride);
}

int launch_gemm_nt_bf16x3_f64(hipStream_t s, int M, int N, int K, float alpha,
                              const uint16_t* Ah, const uint16_t* Al, int64_t lda,
                              const uint16_t* Bh, const uint16_t* Bl, int64_t ldb, double* C,
                              int64_t ldc, int splits, int64_t c_split_stride) {
  return gemm_bf16x3<double>(s, M, N, K, alpha, Ah, Al, lda, Bh, Bl, ldb, C, ldc, splits,
                             c_split_stride);
}

int launch_gemm_bf16x3_grouped(hipStream_t s, int bm, int out, const GemmArgs* args,
                               const int* start, int nprob, int nblocks) {
  if (bm == 128) {
    if (out)
      hipLaunchKernelGGL((gemm_bf16x3_grouped_kernel<128, 128, float>), dim3(nblocks), dim3(256),
                         0, s, args, start, nprob);
    else
      hipLaunchKernelGGL((gemm_bf16x3_grouped_kernel<128, 128, double>), dim3(nblocks),
                         dim3(256), 0, s, args, start, nprob);
  } else {
    if (out)
      hipLaunchKernelGGL((gemm_bf16x3_grouped_kernel<64, 64, float>), dim3(nblocks), dim3(256), 0,
                         s, args, start, nprob);
    else
      hipLaunchKernelGGL((gemm_bf16x3_grouped_kernel<64, 64, double>), dim3(nblocks), dim3(256),
                         0, s, args, start, nprob);
  }
  LAUNCH_CHECK();
  return AFQMC_OK;
}
