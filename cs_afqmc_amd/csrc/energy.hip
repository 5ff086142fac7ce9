// Local energy, rhf._calc_energy_restricted (wavefunctions.py:529-541) and
// uhf._calc_energy (wavefunctions.py:711-743):
//   f_g = rot_chol_s[g] G_s^T (n x n),  c_g = sum_s m_s tr f_g,  x_g = sum_s m_s tr(f_g f_g)
//   E = sum_g (c_g^2 - x_g) / 2 + sum_s m_s sum G_s * rot_h1_s + h0
// with m_s = 2 for restricted walkers (one spin standing for both) and 1 otherwise;
// for m = 2 this is 2 sum c^2 - exc + 2 sum G rot_h1 + h0, the restricted formula.
//
// c_g equals the force bias of the walker (fb_g = sum_s m_s sum_ip rot_chol_s[g,i,p]
// G_s[i,p], wavefunctions.py:503-511 / :694-709), so it comes from the float64
// force-bias GEMM.  The exchange sums x_g need every f_g: a fused kernel forms a tile
// of F[(g,i)][(w,c,j)] = sum_p rot_chol[g][i][p] green[w][c][j][p] on MFMA (float64, or
// split-bf16 with float32 accumulation in the mixed mode), holding whole (g, w) blocks
// -- g_per_tile = BM / n rows of g and w_per_tile = BN / 2n walkers -- so the tile is
// reduced to sum_g x_g per walker in LDS and F never reaches HBM.  Per-tile partial
// sums are added in tile order by the finishing kernel, so results are deterministic.
// The tile holds whole (g, w) blocks for n <= 64 occupied orbitals per spin, the
// library's limit (AFQMC_MAX_NOCC).
#include <type_traits>

#include "common.h"

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef unsigned int u32x4_t __attribute__((ext_vector_type(4)));

namespace {

constexpr int ENER_NT = 256;

// (g-tile, walker-tile) of block bx.  Blocks are dealt round-robin to the 8 XCDs (block
// b runs on XCD b % 8), each with its own 4 MB L2, and each XCD takes a contiguous run of
// the pairs in this order: chunks of gc g-tiles (as many Cholesky tiles as fill half an
// L2), within a chunk walker tiles slowest.  The blocks an XCD has in flight then share
// one or two walker tiles and cycle through Cholesky tiles that stay in its L2, so the
// Cholesky operand is fetched about once per XCD and chunk, and the Green's-function
// operand about once per chunk (instead of the whole Cholesky operand once per walker
// tile with g-tiles fastest).
__device__ inline void exchange_tile_of(int bx, int g_tiles, int w_tiles, int gc, int* gt,
                                        int* wt) {
  const int nblk = g_tiles * w_tiles;
  const int q = nblk >> 3, r = nblk & 7, x = bx & 7, i = bx >> 3;
  const int b = x * q + (x < r ? x : r) + i;
  const int per = gc * w_tiles, c = b / per, g0 = c * gc;
  const int gcc = g_tiles - g0 < gc ? g_tiles - g0 : gc;
  const int rem = b - c * per;
  *wt = rem / gcc;
  *gt = g0 + rem - (rem / gcc) * gcc;
}

// g-tiles per chunk of exchange_tile_of: Cholesky tiles of tile_bytes filling half an L2
inline int exchange_chunk(int g_tiles, int64_t tile_bytes) {
  const int64_t fit = (int64_t)(2 << 20) / (tile_bytes > 0 ? tile_bytes : 1);
  return fit < 1 ? 1 : (fit > g_tiles ? g_tiles : (int)fit);
}

// ---------------------------------------------------------------- fused exchange

// Exchange sums of the (g, w) blocks of an F tile held in LDS (row stride ldf):
// F row gl*n + i, column wl*2n + c*n + j.  x_w = sum_g sum_ij f(i,j) f(j,i) is taken over
// the pairs i <= j (weight 2 off the diagonal): each lane owns pairs and runs down the
// tile's g.  With at least as many walkers as waves each wave takes walkers wl = v, v+NWV,
// ...; with fewer, the waves split each walker's g range and their partial sums (parts,
// one entry per wave in LDS) are added in wave order.  Results go to xpart[w].
template <typename T, int NT = ENER_NT>
__device__ inline void exchange_tile(const T* __restrict__ Fs, int ldf, int n, int g_used,
                                     int w_used, int w0, cplx* __restrict__ xpart,
                                     cplx* parts) {
  constexpr int NWV = NT / 64;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int npairs = n * (n + 1) / 2;
  const int gstride = n * ldf;
  const int wpw = w_used >= NWV ? 1 : NWV / w_used;  // waves per walker
  const int items = w_used * wpw;
  const int gc = (g_used + wpw - 1) / wpw;
  for (int it = wave; it < items; it += NWV) {
    const int wl = it / wpw, part = it - wl * wpw;
    const int gb = part * gc, ge = min(g_used, gb + gc);
    const T* base = Fs + wl * 2 * n + gb * gstride;
    double xr = 0.0, xi = 0.0;
    for (int t = lane; t < npairs; t += 64) {
      // pair t = a (a + 1) / 2 + b, b <= a: (i, j) = (b, a)
      int a = (int)((sqrtf(8.0f * (float)t + 1.0f) - 1.0f) * 0.5f);
      if (a * (a + 1) / 2 > t) --a;
      if ((a + 1) * (a + 2) / 2 <= t) ++a;
      const int i = t - a * (a + 1) / 2, j = a;
      const double wgt = i == j ? 1.0 : 2.0;
      const T* pa = base + i * ldf + j;  // f(i, j): re pa[0], im pa[n]
      const T* pb = base + j * ldf + i;  // f(j, i)
      double sr = 0.0, si = 0.0;
      for (int gl = gb; gl < ge; ++gl) {
        const double ar = (double)pa[0], ai = (double)pa[n];
        const double br = (double)pb[0], bi = (double)pb[n];
        sr += ar * br - ai * bi;
        si += ar * bi + ai * br;
        pa += gstride;
        pb += gstride;
      }
      xr += wgt * sr;
      xi += wgt * si;
    }
    for (int off = 32; off > 0; off >>= 1) {
      xr += __shfl_down(xr, off, 64);
      xi += __shfl_down(xi, off, 64);
    }
    if (lane == 0) {
      if (wpw == 1)
        xpart[w0 + wl] = cplx{xr, xi};
      else
        parts[it] = cplx{xr, xi};
    }
  }
  if (wpw > 1) {  // (uniform in the workgroup)
    __syncthreads();
    if (threadIdx.x < w_used) {
      double xr = 0.0, xi = 0.0;
      for (int q = 0; q < wpw; ++q) {
        xr += parts[threadIdx.x * wpw + q].re;
        xi += parts[threadIdx.x * wpw + q].im;
      }
      xpart[w0 + threadIdx.x] = cplx{xr, xi};
    }
  }
}

// float64 MFMA tile 64 x 128 (4 waves as 2 x 2, each 32 x 64), BK = 16, operands k-major
// in LDS, double buffered.  With F32 (the FP32 precision mode) the same tile runs on
// v_mfma_f32_16x16x4f32: operands rounded to float32 as they are staged, float32
// accumulation and F tile (its D rows are 4 (l / 16) + r instead of l / 16 + 4 r); the
// exchange sums still accumulate in float64.
constexpr int EF_BM = 64, EF_BN = 128, EF_BK = 16, EF_PAD = 16;
// Element (k, m) of a k-major operand tile sits at column m ^ ef_swz<F32, VEC>(k): the
// MFMA operand reads see the same banks as without it, and the staging stores (lanes
// along k) meet distinct banks instead of 8 (float64) or 4 (float32) k rows on one
template <bool F32, int VEC>
__device__ inline int ef_swz(int k) {
  if constexpr (F32) return ((k >> 2) & 3) << (VEC == 2 ? 3 : 2);
  return ((k >> 1) & 7) << (VEC == 2 ? 2 : 1);
}
constexpr int EF_LDF = 129;  // F tile row stride (odd: transposed reads spread over banks)
constexpr int EF_OPS_BYTES = 2 * EF_BK * (EF_BM + EF_PAD) * 8 + 2 * EF_BK * (EF_BN + EF_PAD) * 8;
constexpr int EF_F_BYTES = EF_BM * EF_LDF * 8;
constexpr int EF_F_PARTS = EF_F_BYTES + (ENER_NT / 64) * 16;
constexpr int EF_LDS = EF_OPS_BYTES > EF_F_PARTS ? EF_OPS_BYTES : EF_F_PARTS;

template <int VEC, bool F32>
__device__ __forceinline__ void exchange_f64_body(
    int G, int n, int N, int W, int gpt, int wpt, int g_tiles, int gc,
    const double* __restrict__ A, const double* __restrict__ B, cplx* __restrict__ xpart,
    const int bx) {
  typedef typename std::conditional<F32, float, double>::type T;
  typedef typename std::conditional<F32, float4_t, double4_t>::type AccT;
  constexpr int BM = EF_BM, BN = EF_BN, BK = EF_BK;
  constexpr int TM = BM / 2 / 16, TN = BN / 2 / 16;
  constexpr int KT = BK / VEC, RSTEP = ENER_NT / KT, A_LD = BM / RSTEP, B_LD = BN / RSTEP;
  __shared__ __attribute__((aligned(16))) unsigned char lds[EF_LDS];
  typedef T AsT[BK][BM + EF_PAD];
  typedef T BsT[BK][BN + EF_PAD];
  AsT* As = reinterpret_cast<AsT*>(lds);
  BsT* Bs = reinterpret_cast<BsT*>(lds + 2 * sizeof(AsT));
  T* Fs = reinterpret_cast<T*>(lds);

  int gt, wt;
  exchange_tile_of(bx, g_tiles, (W + wpt - 1) / wpt, gc, &gt, &wt);
  const int g0 = gt * gpt, w0 = wt * wpt;
  const int M = G * n, NB = 2 * W * n;
  const int m0 = g0 * n, n0 = w0 * 2 * n;
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int wm = wave & 1, wn = wave >> 1;
  const int lk = (t % KT) * VEC, lr = t / KT;

  T ra[A_LD][VEC], rb[B_LD][VEC];
  auto load_tile = [&](int k0) {
    const int k = k0 + lk;
    const bool kok = k < N;  // VEC == 2 needs N even (checked at launch)
#pragma unroll
    for (int i = 0; i < A_LD; ++i) {
      const int r = m0 + lr + i * RSTEP;
      if (VEC == 2) {
        double2 v = make_double2(0.0, 0.0);
        if (r < M && kok) v = *reinterpret_cast<const double2*>(A + (int64_t)r * N + k);
        ra[i][0] = (T)v.x;
        ra[i][VEC - 1] = (T)v.y;
      } else {
        ra[i][0] = (r < M && kok) ? (T)A[(int64_t)r * N + k] : (T)0;
      }
    }
#pragma unroll
    for (int i = 0; i < B_LD; ++i) {
      const int r = n0 + lr + i * RSTEP;
      if (VEC == 2) {
        double2 v = make_double2(0.0, 0.0);
        if (r < NB && kok) v = *reinterpret_cast<const double2*>(B + (int64_t)r * N + k);
        rb[i][0] = (T)v.x;
        rb[i][VEC - 1] = (T)v.y;
      } else {
        rb[i][0] = (r < NB && kok) ? (T)B[(int64_t)r * N + k] : (T)0;
      }
    }
  };
  auto store_tile = [&](int buf) {
#pragma unroll
    for (int i = 0; i < A_LD; ++i)
#pragma unroll
      for (int v = 0; v < VEC; ++v)
        As[buf][lk + v][(lr + i * RSTEP) ^ ef_swz<F32, VEC>(lk + v)] = ra[i][v];
#pragma unroll
    for (int i = 0; i < B_LD; ++i)
#pragma unroll
      for (int v = 0; v < VEC; ++v)
        Bs[buf][lk + v][(lr + i * RSTEP) ^ ef_swz<F32, VEC>(lk + v)] = rb[i][v];
  };

  AccT acc[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) acc[i][j] = AccT{0, 0, 0, 0};

  const int nk = (N + BK - 1) / BK;
  load_tile(0);
  store_tile(0);
  __syncthreads();
  const int am = wm * (BM / 2) + (lane & 15);
  const int bn = wn * (BN / 2) + (lane & 15);
  const int kq = lane >> 4;
  for (int kt = 0; kt < nk; ++kt) {
    const int cur = kt & 1;
    if (kt + 1 < nk) load_tile((kt + 1) * BK);
#pragma unroll
    for (int ks = 0; ks < BK / 4; ++ks) {
      T af[TM], bf[TN];
      const int sw = ef_swz<F32, VEC>(ks * 4 + kq);
#pragma unroll
      for (int i = 0; i < TM; ++i) af[i] = As[cur][ks * 4 + kq][(am + i * 16) ^ sw];
#pragma unroll
      for (int j = 0; j < TN; ++j) bf[j] = Bs[cur][ks * 4 + kq][(bn + j * 16) ^ sw];
#pragma unroll
      for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j) {
          if constexpr (F32)
            acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(af[i], bf[j], acc[i][j], 0, 0, 0);
          else
            acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(af[i], bf[j], acc[i][j], 0, 0, 0);
        }
    }
    if (kt + 1 < nk) store_tile(cur ^ 1);
    __syncthreads();
  }
  // F tile -> LDS (the operand buffers are dead after the last barrier)
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) {
      const int col = wn * (BN / 2) + j * 16 + (lane & 15);
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = wm * (BM / 2) + i * 16 + (F32 ? 4 * (lane >> 4) + r : (lane >> 4) + 4 * r);
        Fs[row * EF_LDF + col] = acc[i][j][r];
      }
    }
  __syncthreads();
  const int g_used = (G - g0 < gpt) ? G - g0 : gpt;
  const int w_used = (W - w0 < wpt) ? W - w0 : wpt;
  exchange_tile<T>(Fs, EF_LDF, n, g_used, w_used, w0, xpart + (int64_t)gt * W,
                   reinterpret_cast<cplx*>(lds + EF_F_BYTES));
}

template <int VEC, bool F32>
__global__ __launch_bounds__(ENER_NT) void exchange_f64_kernel(
    int G, int n, int N, int W, int gpt, int wpt, int g_tiles, int gc,
    const double* __restrict__ A, const double* __restrict__ B, cplx* __restrict__ xpart) {
  exchange_f64_body<VEC, F32>(G, n, N, W, gpt, wpt, g_tiles, gc, A, B, xpart, blockIdx.x);
}

template <int VEC, bool F32>
__global__ __launch_bounds__(ENER_NT) void exchange_f64_grouped_kernel(
    const ExchArgs* __restrict__ args, const int* __restrict__ start, int nprob) {
  const int p = group_of(start, nprob, blockIdx.x);
  const ExchArgs a = args[p];
  exchange_f64_body<VEC, F32>(a.G, a.n, a.K, a.W, a.gpt, a.wpt, a.g_tiles, a.gc,
                              static_cast<const double*>(a.A), static_cast<const double*>(a.B),
                              a.xpart, blockIdx.x - start[p]);
}

// split-bf16 tile 128 x 64 WN on v_mfma_f32_16x16x32_bf16 (2 x WN waves, each 64 x 64),
// BK = 32; operand planes have K = Np32 (zero padded), so there is no k tail.  WN = 2
// (128 columns, 4 waves) or, where it holds more whole walkers (2n columns each) per
// tile, WN = 4 (256 columns, 8 waves): the Cholesky tile is then read once per more
// walkers.
// operand rows of 32 bf16 (64 B) with the chunk order c ^ ((r >> 1) & 3), as in the
// split-bf16 GEMM, under which the 16-row operand reads meet no LDS bank conflicts
constexpr int EB_BM = 128, EB_BK = 32, EB_LDSW = EB_BK;
__device__ inline int eb_chunk(int r, int c) { return (c ^ ((r >> 1) & 3)) << 3; }
template <int WN>
struct EbTile {
  static constexpr int NT = 128 * WN, BN = 64 * WN, LDF = BN + 1;
  static constexpr int OPS = 2 * 2 * EB_BM * EB_LDSW * 2 + 2 * 2 * BN * EB_LDSW * 2;
  static constexpr int F = EB_BM * LDF * 4;
  static constexpr int F_PARTS = F + (NT / 64) * 16;
  static constexpr int LDS = OPS > F_PARTS ? OPS : F_PARTS;
};

template <int WN>
__device__ __forceinline__ void exchange_bf16x3_body(
    int G, int n, int K, int W, int gpt, int wpt, int g_tiles, int gc,
    const uint16_t* __restrict__ Ah,
    const uint16_t* __restrict__ Al, const uint16_t* __restrict__ Bh,
    const uint16_t* __restrict__ Bl, cplx* __restrict__ xpart, const int bx) {
  typedef EbTile<WN> Tile;
  constexpr int NT = Tile::NT, BM = EB_BM, BN = Tile::BN, LDSW = EB_LDSW;
  constexpr int TM = BM / 2 / 16, TN = BN / WN / 16;
  constexpr int A_LD = BM * 4 / NT, B_LD = BN * 4 / NT;
  static_assert(A_LD >= 1 && B_LD >= 1, "tile loads");
  __shared__ __attribute__((aligned(16))) unsigned char lds[Tile::LDS];
  typedef uint16_t PlaneA[2][BM * LDSW];  // [hi/lo]
  typedef uint16_t PlaneB[2][BN * LDSW];
  PlaneA* sA = reinterpret_cast<PlaneA*>(lds);
  PlaneB* sB = reinterpret_cast<PlaneB*>(lds + 2 * sizeof(PlaneA));
  float* Fs = reinterpret_cast<float*>(lds);

  int gt, wt;
  exchange_tile_of(bx, g_tiles, (W + wpt - 1) / wpt, gc, &gt, &wt);
  const int g0 = gt * gpt, w0 = wt * wpt;
  const int M = G * n, NB = 2 * W * n;
  const int m0 = g0 * n, n0 = w0 * 2 * n;
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int wm = wave & 1, wn = wave >> 1;  // wn < WN

  u32x4_t ra[2][A_LD], rb[2][B_LD];
  auto load_tile = [&](int k0) {
#pragma unroll
    for (int j = 0; j < A_LD; ++j) {
      const int q = t + NT * j, row = q >> 2, ch = q & 3;
      const int r = m0 + row;
      u32x4_t vh = {0, 0, 0, 0}, vl = {0, 0, 0, 0};
      if (r < M) {
        vh = *reinterpret_cast<const u32x4_t*>(Ah + (int64_t)r * K + k0 + ch * 8);
        vl = *reinterpret_cast<const u32x4_t*>(Al + (int64_t)r * K + k0 + ch * 8);
      }
      ra[0][j] = vh;
      ra[1][j] = vl;
    }
#pragma unroll
    for (int j = 0; j < B_LD; ++j) {
      const int q = t + NT * j, row = q >> 2, ch = q & 3;
      const int r = n0 + row;
      u32x4_t vh = {0, 0, 0, 0}, vl = {0, 0, 0, 0};
      if (r < NB) {
        vh = *reinterpret_cast<const u32x4_t*>(Bh + (int64_t)r * K + k0 + ch * 8);
        vl = *reinterpret_cast<const u32x4_t*>(Bl + (int64_t)r * K + k0 + ch * 8);
      }
      rb[0][j] = vh;
      rb[1][j] = vl;
    }
  };
  auto store_tile = [&](int buf) {
#pragma unroll
    for (int j = 0; j < A_LD; ++j) {
      const int q = t + NT * j, row = q >> 2, ch = q & 3;
#pragma unroll
      for (int h = 0; h < 2; ++h)
        *reinterpret_cast<u32x4_t*>(&sA[buf][h][row * LDSW + eb_chunk(row, ch)]) = ra[h][j];
    }
#pragma unroll
    for (int j = 0; j < B_LD; ++j) {
      const int q = t + NT * j, row = q >> 2, ch = q & 3;
#pragma unroll
      for (int h = 0; h < 2; ++h)
        *reinterpret_cast<u32x4_t*>(&sB[buf][h][row * LDSW + eb_chunk(row, ch)]) = rb[h][j];
    }
  };

  float4_t acc[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) acc[i][j] = float4_t{0.f, 0.f, 0.f, 0.f};

  const int nk = K / EB_BK;
  load_tile(0);
  store_tile(0);
  __syncthreads();
  const int arow = wm * (BM / 2) + (lane & 15);
  const int brow = wn * (BN / WN) + (lane & 15);
  // k chunk lane >> 4 of rows whose bits 2-3 are those of lane (row = 16 t + lane % 16)
  const int koff = eb_chunk(lane & 15, lane >> 4);
  for (int kt = 0; kt < nk; ++kt) {
    const int cur = kt & 1;
    if (kt + 1 < nk) load_tile((kt + 1) * EB_BK);
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
      const int col = wn * (BN / WN) + j * 16 + (lane & 15);
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = wm * (BM / 2) + i * 16 + 4 * (lane >> 4) + r;
        Fs[row * Tile::LDF + col] = acc[i][j][r];
      }
    }
  __syncthreads();
  const int g_used = (G - g0 < gpt) ? G - g0 : gpt;
  const int w_used = (W - w0 < wpt) ? W - w0 : wpt;
  exchange_tile<float, NT>(Fs, Tile::LDF, n, g_used, w_used, w0, xpart + (int64_t)gt * W,
                           reinterpret_cast<cplx*>(lds + Tile::F));
}

template <int WN>
__global__ __launch_bounds__(128 * WN) void exchange_bf16x3_kernel(
    int G, int n, int K, int W, int gpt, int wpt, int g_tiles, int gc,
    const uint16_t* __restrict__ Ah, const uint16_t* __restrict__ Al,
    const uint16_t* __restrict__ Bh, const uint16_t* __restrict__ Bl, cplx* __restrict__ xpart) {
  exchange_bf16x3_body<WN>(G, n, K, W, gpt, wpt, g_tiles, gc, Ah, Al, Bh, Bl, xpart,
                           blockIdx.x);
}

template <int WN>
__global__ __launch_bounds__(128 * WN) void exchange_bf16x3_grouped_kernel(
    const ExchArgs* __restrict__ args, const int* __restrict__ start, int nprob) {
  const int p = group_of(start, nprob, blockIdx.x);
  const ExchArgs a = args[p];
  exchange_bf16x3_body<WN>(a.G, a.n, a.K, a.W, a.gpt, a.wpt, a.g_tiles, a.gc,
                           static_cast<const uint16_t*>(a.A), static_cast<const uint16_t*>(a.Al),
                           static_cast<const uint16_t*>(a.B), static_cast<const uint16_t*>(a.Bl),
                           a.xpart, blockIdx.x - start[p]);
}

// E_w = sum_g (c_g^2 - x_g) / 2 + sum_s m_s sum G_s rot_h1_s + h0 with c_g the summed
// force-bias slices and x_g = sum over spins and g-tiles (in order) of m_s * xpart.
__device__ __forceinline__ void energy_finish_body(
    int G, int N, int nspin, int mult, int n0, int n1, const double* __restrict__ fb, int nparts,
    int64_t part_stride, const cplx* __restrict__ xpart, int gt0, int gt1, int W,
    const double* __restrict__ green0, const double* __restrict__ green1,
    const double* __restrict__ rot_h1_0, const double* __restrict__ rot_h1_1, double h0,
    cplx* __restrict__ energies, const int bx) {
  __shared__ double red[ENER_NT / 64];
  const int w = bx;
  double c2r = 0.0, c2i = 0.0;
  for (int g = threadIdx.x; g < G; g += ENER_NT) {
    const double* p = fb + (int64_t)w * 2 * G + g;
    double re = p[0], im = p[G];
    for (int s = 1; s < nparts; ++s) {
      re += p[s * part_stride];
      im += p[s * part_stride + G];
    }
    c2r += re * re - im * im;
    c2i += 2.0 * re * im;
  }
  double xr = 0.0, xi = 0.0;
  const int gts[2] = {gt0, gt1};
  const cplx* xs = xpart;
  for (int s = 0; s < nspin; ++s) {
    for (int t = threadIdx.x; t < gts[s]; t += ENER_NT) {
      const cplx v = xs[(int64_t)t * W + w];
      xr += mult * v.re;
      xi += mult * v.im;
    }
    xs += (int64_t)gts[s] * W;
  }
  const int ns[2] = {n0, n1};
  const double* gs[2] = {green0, green1};
  const double* hs[2] = {rot_h1_0, rot_h1_1};
  double e1r = 0.0, e1i = 0.0;
  for (int s = 0; s < nspin; ++s) {
    const int kd = ns[s] * N;
    const double* gr = gs[s] + (int64_t)w * 2 * kd;
    const double* gi = gr + kd;
    for (int e = threadIdx.x; e < kd; e += ENER_NT) {
      e1r += mult * gr[e] * hs[s][e];
      e1i += mult * gi[e] * hs[s][e];
    }
  }
  const double C2r = block_sum<ENER_NT>(c2r, red);
  const double C2i = block_sum<ENER_NT>(c2i, red);
  const double Xr = block_sum<ENER_NT>(xr, red);
  const double Xi = block_sum<ENER_NT>(xi, red);
  const double E1r = block_sum<ENER_NT>(e1r, red);
  const double E1i = block_sum<ENER_NT>(e1i, red);
  if (threadIdx.x == 0)
    energies[w] = cplx{(C2r - Xr) / 2.0 + E1r + h0, (C2i - Xi) / 2.0 + E1i};
}

__global__ __launch_bounds__(ENER_NT) void energy_finish_kernel(
    int G, int N, int nspin, int mult, int n0, int n1, const double* __restrict__ fb, int nparts,
    int64_t part_stride, const cplx* __restrict__ xpart, int gt0, int gt1, int W,
    const double* __restrict__ green0, const double* __restrict__ green1,
    const double* __restrict__ rot_h1_0, const double* __restrict__ rot_h1_1, double h0,
    cplx* __restrict__ energies) {
  energy_finish_body(G, N, nspin, mult, n0, n1, fb, nparts, part_stride, xpart, gt0, gt1, W,
                     green0, green1, rot_h1_0, rot_h1_1, h0, energies, blockIdx.x);
}

__global__ __launch_bounds__(ENER_NT) void energy_finish_grouped_kernel(
    const FinishArgs* __restrict__ args, const int* __restrict__ start, int nprob) {
  const int p = group_of(start, nprob, blockIdx.x);
  const FinishArgs a = args[p];
  energy_finish_body(a.G, a.N, a.nspin, a.mult, a.n0, a.n1, a.fb, a.nparts, a.part_stride,
                     a.xpart, a.gt0, a.gt1, a.W, a.green0, a.green1, a.rot_h1_0, a.rot_h1_1,
                     a.h0, a.energies, blockIdx.x - start[p]);
}

}  // namespace

// Waves across the split-bf16 exchange tile for n occupied orbitals: 4 (256 columns)
// where that tile holds at least 1.2 times the walker columns per Cholesky row of the
// 128-column one, else 2.
int exchange_wn(int n) {
  const int c2 = (128 / (2 * n)) * 2, c4 = (256 / (2 * n)) * 2;
  return 5 * c4 >= 12 * c2 ? 4 : 2;
}

int energy_g_tiles(const afqmc_ctx* c, int s) {
  const int n = c->nocc[s];
  if (n < 1) return 0;
  const int gpt = (c->mixed ? EB_BM : EF_BM) / n;
  return (c->d.nchol + gpt - 1) / gpt;
}

void exchange_plan(const afqmc_ctx* c, int s, int W, ExchArgs* a, int* kind, int* nblocks) {
  const int N = c->d.norb, G = c->d.nchol, n = c->nocc[s];
  cplx* xp = c->energy_x;
  for (int t = 0; t < s; ++t) xp += (int64_t)energy_g_tiles(c, t) * W;
  *a = ExchArgs{};
  a->G = G;
  a->n = n;
  a->W = W;
  a->xpart = xp;
  if (c->mixed) {
    const int wn = exchange_wn(n);
    a->K = c->Np32;
    a->gpt = EB_BM / n;
    a->wpt = 64 * wn / (2 * n);
    a->g_tiles = (G + a->gpt - 1) / a->gpt;
    a->gc = exchange_chunk(a->g_tiles, (int64_t)a->gpt * n * c->Np32 * 4);
    a->A = c->rot_chol_bf[s][0];
    a->Al = c->rot_chol_bf[s][1];
    a->B = c->green_bf[s][0];
    a->Bl = c->green_bf[s][1];
    *kind = wn == 4 ? EXCH_BF16X3_WN4 : EXCH_BF16X3_WN2;
  } else {
    const bool vec2 = (N % 2 == 0);
    a->K = N;
    a->gpt = EF_BM / n;
    a->wpt = EF_BN / (2 * n);
    a->g_tiles = (G + a->gpt - 1) / a->gpt;
    a->gc = exchange_chunk(a->g_tiles, (int64_t)a->gpt * n * N * 8);
    a->A = c->rot_chol[s];
    a->B = c->green[s];
    *kind = c->f32 ? (vec2 ? EXCH_F32_V2 : EXCH_F32_V1) : (vec2 ? EXCH_F64_V2 : EXCH_F64_V1);
  }
  *nblocks = a->g_tiles * ((W + a->wpt - 1) / a->wpt);
}

void energy_finish_args(const afqmc_ctx* c, int W, cplx* energies, FinishArgs* a) {
  *a = FinishArgs{};
  a->G = c->d.nchol;
  a->N = c->d.norb;
  a->nspin = c->nspin;
  a->mult = c->mult;
  a->n0 = c->nocc[0];
  a->n1 = c->nocc[1];
  a->nparts = c->fb_parts;
  a->gt0 = energy_g_tiles(c, 0);
  a->gt1 = c->nspin > 1 ? energy_g_tiles(c, 1) : 0;
  a->W = W;
  a->part_stride = (int64_t)2 * W * c->d.nchol;
  a->fb = c->fb;
  a->xpart = c->energy_x;
  a->green0 = c->green[0];
  a->green1 = c->green[1];
  a->rot_h1_0 = c->rot_h1[0];
  a->rot_h1_1 = c->rot_h1[1];
  a->h0 = c->h0;
  a->energies = energies;
}

int launch_exchange_grouped(hipStream_t st, int kind, const ExchArgs* args, const int* start,
                            int nprob, int nblocks) {
  switch (kind) {
    case EXCH_BF16X3_WN2:
      hipLaunchKernelGGL(exchange_bf16x3_grouped_kernel<2>, dim3(nblocks), dim3(EbTile<2>::NT),
                         0, st, args, start, nprob);
      break;
    case EXCH_BF16X3_WN4:
      hipLaunchKernelGGL(exchange_bf16x3_grouped_kernel<4>, dim3(nblocks), dim3(EbTile<4>::NT),
                         0, st, args, start, nprob);
      break;
    case EXCH_F64_V1:
      hipLaunchKernelGGL((exchange_f64_grouped_kernel<1, false>), dim3(nblocks), dim3(ENER_NT),
                         0, st, args, start, nprob);
      break;
    case EXCH_F64_V2:
      hipLaunchKernelGGL((exchange_f64_grouped_kernel<2, false>), dim3(nblocks), dim3(ENER_NT),
                         0, st, args, start, nprob);
      break;
    case EXCH_F32_V1:
      hipLaunchKernelGGL((exchange_f64_grouped_kernel<1, true>), dim3(nblocks), dim3(ENER_NT),
                         0, st, args, start, nprob);
      break;
    default:
      hipLaunchKernelGGL((exchange_f64_grouped_kernel<2, true>), dim3(nblocks), dim3(ENER_NT),
                         0, st, args, start, nprob);
      break;
  }
  LAUNCH_CHECK();
  return AFQMC_OK;
}

int launch_energy_finish_grouped(hipStream_t st, const FinishArgs* args, const int* start,
                                 int nprob, int nblocks) {
  hipLaunchKernelGGL(energy_finish_grouped_kernel, dim3(nblocks), dim3(ENER_NT), 0, st, args,
                     start, nprob);
  LAUNCH_CHECK();
  return AFQMC_OK;
}

// Needs the float64 Green's function of every spin in c->green (and, in the mixed mode,
// its split-bf16 planes) and the float64 force bias in c->fb (c->fb_parts slices).
int launch_energy(afqmc_ctx* c, int W, cplx* energies) {
  static_assert(AFQMC_MAX_NOCC <= EF_BM && 2 * AFQMC_MAX_NOCC <= EF_BN,
                "a (g, w) block must fit one tile");
  static_assert(AFQMC_MAX_NOCC <= EB_BM && 2 * AFQMC_MAX_NOCC <= EbTile<2>::BN,
                "a (g, w) block must fit one tile");
  for (int s = 0; s < c->nspin; ++s) {
    ExchArgs a;
    int kind, nb;
    exchange_plan(c, s, W, &a, &kind, &nb);
    const dim3 grid(nb);
    switch (kind) {
      case EXCH_BF16X3_WN4:
      case EXCH_BF16X3_WN2: {
        const uint16_t* Ah = static_cast<const uint16_t*>(a.A);
        const uint16_t* Al = static_cast<const uint16_t*>(a.Al);
        const uint16_t* Bh = static_cast<const uint16_t*>(a.B);
        const uint16_t* Bl = static_cast<const uint16_t*>(a.Bl);
        if (kind == EXCH_BF16X3_WN4)
          hipLaunchKernelGGL(exchange_bf16x3_kernel<4>, grid, dim3(EbTile<4>::NT), 0, c->stream,
                             a.G, a.n, a.K, W, a.gpt, a.wpt, a.g_tiles, a.gc, Ah, Al, Bh, Bl,
                             a.xpart);
        else
          hipLaunchKernelGGL(exchange_bf16x3_kernel<2>, grid, dim3(EbTile<2>::NT), 0, c->stream,
                             a.G, a.n, a.K, W, a.gpt, a.wpt, a.g_tiles, a.gc, Ah, Al, Bh, Bl,
                             a.xpart);
        break;
      }
      default: {
        const double* A = static_cast<const double*>(a.A);
        const double* B = static_cast<const double*>(a.B);
        const dim3 block(ENER_NT);
        if (kind == EXCH_F32_V2)
          hipLaunchKernelGGL((exchange_f64_kernel<2, true>), grid, block, 0, c->stream, a.G,
                             a.n, a.K, W, a.gpt, a.wpt, a.g_tiles, a.gc, A, B, a.xpart);
        else if (kind == EXCH_F32_V1)
          hipLaunchKernelGGL((exchange_f64_kernel<1, true>), grid, block, 0, c->stream, a.G,
                             a.n, a.K, W, a.gpt, a.wpt, a.g_tiles, a.gc, A, B, a.xpart);
        else if (kind == EXCH_F64_V2)
          hipLaunchKernelGGL((exchange_f64_kernel<2, false>), grid, block, 0, c->stream, a.G,
                             a.n, a.K, W, a.gpt, a.wpt, a.g_tiles, a.gc, A, B, a.xpart);
        else
          hipLaunchKernelGGL((exchange_f64_kernel<1, false>), grid, block, 0, c->stream, a.G,
                             a.n, a.K, W, a.gpt, a.wpt, a.g_tiles, a.gc, A, B, a.xpart);
        break;
      }
    }
    LAUNCH_CHECK();
  }
  FinishArgs f;
  energy_finish_args(c, W, energies, &f);
  hipLaunchKernelGGL(energy_finish_kernel, dim3(W), dim3(ENER_NT), 0, c->stream, f.G, f.N,
                     f.nspin, f.mult, f.n0, f.n1, f.fb, f.nparts, f.part_stride, f.xpart, f.gt0,
                     f.gt1, W, f.green0, f.green1, f.rot_h1_0, f.rot_h1_1, f.h0, energies);
  LAUNCH_CHECK();
  return AFQMC_OK;
}
