// Stochastic reconfiguration (sr.py:24-48) and the auxiliary-field generator.
//
// SR: c = cumsum |w| (sequential, left to right, as numpy.cumsum), total = c[W-1],
// z_i = total * (i + zeta) / W (or ((i + zeta) / W) * total, the order of the
// gather/scatter variant), idx_i = searchsorted(c, z_i, side='left') clamped to
// W - 1 (jax gathers clamp out-of-range indices), walkers <- walkers[idx],
// weights <- total / W.  Indices are integers, so they are reproduced bit for bit.
//
// Fields: Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11) keyed by the 64-bit seed,
// counter = (pair index, stream counter); two 53-bit uniforms per draw feed a
// Box-Muller transform (float64).
#include <algorithm>

#include "common.h"

namespace {

// ---- Philox4x32-10
struct u32x4 {
  uint32_t x, y, z, w;
};

__device__ inline void mulhilo(uint32_t a, uint32_t b, uint32_t& hi, uint32_t& lo) {
  const uint64_t p = (uint64_t)a * (uint64_t)b;
  hi = (uint32_t)(p >> 32);
  lo = (uint32_t)p;
}

__device__ inline u32x4 philox4x32_10(u32x4 c, uint32_t k0, uint32_t k1) {
  const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
  const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    uint32_t hi0, lo0, hi1, lo1;
    mulhilo(M0, c.x, hi0, lo0);
    mulhilo(M1, c.z, hi1, lo1);
    c = u32x4{hi1 ^ c.y ^ k0, lo1, hi0 ^ c.w ^ k1, lo0};
    k0 += W0;
    k1 += W1;
  }
  return c;
}

constexpr int SR_NT = 1024;
constexpr int SR_CHUNK = 4096;  // |w| staged through LDS in chunks of this many

// One workgroup.  The running sum stays sequential (left to right, one thread) so that
// the cumulative weights -- and hence the integer indices -- are those of numpy.cumsum;
// the other threads only stage |w| into LDS ahead of it and run the searches.
// global_order selects z_i = ((i + zeta) / W) * total (the gather/scatter form of
// sr.py:66-106 on one rank) instead of z_i = total * (i + zeta) / W (sr.py:24-34).
// key (device, may be NULL): {seed, counter}; zeta is then the uniform draw of that key
// (the host's rng.uniform: Philox4x32-10 of counter (0, 0xFFFFFFFF, counter) under the
// seed, top 53 bits of the first two words) and the counter advances by one, as
// rng.split does before the draw.
__global__ __launch_bounds__(SR_NT) void sr_index_kernel(int W, double zeta, int global_order,
                                                         double* __restrict__ weights,
                                                         double* __restrict__ cum,
                                                         int32_t* __restrict__ idx,
                                                         int32_t* __restrict__ idx_out,
                                                         uint64_t* __restrict__ key) {
  __shared__ double absw[SR_CHUNK];
  __shared__ double total_s;
  __shared__ double zeta_s;
  if (key && threadIdx.x == 0) {
    const uint64_t seed = key[0], ctr = key[1];
    const u32x4 r = philox4x32_10(u32x4{0u, 0xFFFFFFFFu, (uint32_t)ctr, (uint32_t)(ctr >> 32)},
                                  (uint32_t)seed, (uint32_t)(seed >> 32));
    const uint64_t a = ((uint64_t)r.y << 32) | r.x;
    zeta_s = (double)(a >> 11) * 0x1.0p-53;
    key[1] = ctr + 1;
  }
  double run = 0.0;
  for (int base = 0; base < W; base += SR_CHUNK) {
    const int len = W - base < SR_CHUNK ? W - base : SR_CHUNK;
    for (int i = threadIdx.x; i < len; i += SR_NT) absw[i] = fabs(weights[base + i]);
    __syncthreads();
    if (threadIdx.x == 0) {
      for (int i = 0; i < len; ++i) {
        run += absw[i];
        cum[base + i] = run;
      }
      total_s = run;
    }
    __syncthreads();
  }
  const double total = total_s;
  const double avg = total / (double)W;
  if (key) zeta = zeta_s;  // written before the first barrier above
  for (int i = threadIdx.x; i < W; i += SR_NT) {
    const double z = global_order ? (((double)i + zeta) / (double)W) * total
                                  : total * ((double)i + zeta) / (double)W;
    // first j with cum[j] >= z
    int lo = 0, hi = W;
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (cum[mid] < z)
        lo = mid + 1;
      else
        hi = mid;
    }
    const int j = lo < W ? lo : W - 1;
    idx[i] = j;
    if (idx_out) idx_out[i] = j;
    weights[i] = avg;
  }
}

// dst[i] = src[idx[i]] for walkers of `elems` complex numbers each
__global__ void gather_kernel(int W, int elems, const int32_t* __restrict__ idx,
                              const cplx* __restrict__ src, cplx* __restrict__ dst) {
  const int i = blockIdx.y;
  const cplx* s = src + (int64_t)idx[i] * elems;
  cplx* d = dst + (int64_t)i * elems;
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < elems; e += gridDim.x * blockDim.x)
    d[e] = s[e];
}

// Normals 2 pair and 2 pair + 1 of the stream (seed, counter): one Philox4x32-10 block
// at counter (pair, counter) gives two 53-bit uniforms u1 in (0, 1], u2 in [0, 1), and
// Box-Muller turns them into rad cos and rad sin.
__device__ inline void normal_pair(uint64_t seed, uint64_t counter, int64_t pair,
                                   double* __restrict__ out, int64_t count) {
  const u32x4 c = u32x4{(uint32_t)pair, (uint32_t)((uint64_t)pair >> 32), (uint32_t)counter,
                        (uint32_t)(counter >> 32)};
  const u32x4 r = philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
  const uint64_t a = ((uint64_t)r.y << 32) | r.x;
  const uint64_t b = ((uint64_t)r.w << 32) | r.z;
  const double u1 = ((double)(a >> 11) + 1.0) * 0x1.0p-53;  // (0, 1]
  const double u2 = (double)(b >> 11) * 0x1.0p-53;          // [0, 1)
  const double rad = sqrt(-2.0 * log(u1));
  double s, co;
  sincos(2.0 * M_PI * u2, &s, &co);
  out[2 * pair] = rad * co;
  if (2 * pair + 1 < count) out[2 * pair + 1] = rad * s;
}

__global__ void normal_kernel(uint64_t seed, uint64_t counter, double* __restrict__ out,
                              int64_t count) {
  const int64_t pair = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (2 * pair < count) normal_pair(seed, counter, pair, out, count);
}

// The same draws keyed by a device-resident {seed, counter} (read, not modified: the
// counter advances in key_advance_kernel, stream-ordered after this launch).
__global__ void normal_key_kernel(const uint64_t* __restrict__ key, double* __restrict__ out,
                                  int64_t count) {
  const int64_t pair = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (2 * pair < count) normal_pair(key[0], key[1], pair, out, count);
}

__global__ void key_advance_kernel(uint64_t* __restrict__ key) {
  if (threadIdx.x == 0) key[1] += 1;
}

// grouped form: member p draws its count normals from its own key (blocks
// [start[p], start[p + 1]) of 256 pairs), then every key advances
__global__ __launch_bounds__(256) void normal_key_grouped_kernel(
    const DrawArgs* __restrict__ args, const int* __restrict__ start, int nprob) {
  const int p = group_of(start, nprob, blockIdx.x);
  const DrawArgs a = args[p];
  const int64_t pair = (int64_t)(blockIdx.x - start[p]) * 256 + threadIdx.x;
  if (2 * pair < a.count) normal_pair(a.key[0], a.key[1], pair, a.out, a.count);
}

__global__ void key_advance_grouped_kernel(const DrawArgs* __restrict__ args, int nprob) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < nprob) args[p].key[1] += 1;
}

}  // namespace

int launch_sr(afqmc_ctx* c, int W, cplx* wup, cplx* wdn, double* weights, double zeta,
              int32_t* idx_out, bool global_order, uint64_t* key) {
  hipLaunchKernelGGL(sr_index_kernel, dim3(1), dim3(SR_NT), 0, c->stream, W, zeta,
                     global_order ? 1 : 0, weights, c->sr_cum, c->sr_idx, idx_out, key);
  LAUNCH_CHECK();
  const int N = c->d.norb;
  for (int s = 0; s < c->nspin; ++s) {
    cplx* wk = (s == 0) ? wup : wdn;
    const int elems = N * c->nocc[s];
    const int bx = (elems + 255) / 256 < 8 ? (elems + 255) / 256 : 8;
    hipLaunchKernelGGL(gather_kernel, dim3(bx, W), dim3(256), 0, c->stream, W, elems,
                       c->sr_idx, wk, c->walk_tmp);
    LAUNCH_CHECK();
    HIP_TRY(hipMemcpyAsync(wk, c->walk_tmp, (size_t)W * elems * sizeof(cplx),
                           hipMemcpyDeviceToDevice, c->stream));
  }
  return AFQMC_OK;
}

int launch_fields_normal(afqmc_ctx* c, uint64_t seed, uint64_t counter, double* out,
                         int64_t count) {
  const int64_t pairs = (count + 1) / 2;
  hipLaunchKernelGGL(normal_kernel, dim3((unsigned)((pairs + 255) / 256)), dim3(256), 0,
                     c->stream, seed, counter, out, count);
  LAUNCH_CHECK();
  return AFQMC_OK;
}

int launch_fields_normal_key_grouped(hipStream_t st, const DrawArgs* args, const int* start,
                                     int nprob, int nblocks) {
  hipLaunchKernelGGL(normal_key_grouped_kernel, dim3(nblocks), dim3(256), 0, st, args, start,
                     nprob);
  LAUNCH_CHECK();
  hipLaunchKernelGGL(key_advance_grouped_kernel, dim3((nprob + 255) / 256), dim3(256), 0, st,
                     args, nprob);
  LAUNCH_CHECK();
  return AFQMC_OK;
}

int launch_fields_normal_key(afqmc_ctx* c, uint64_t* key, double* out, int64_t count) {
  const int64_t pairs = (count + 1) / 2;
  hipLaunchKernelGGL(normal_key_kernel, dim3((unsigned)((pairs + 255) / 256)), dim3(256), 0,
                     c->stream, key, out, count);
  LAUNCH_CHECK();
  hipLaunchKernelGGL(key_advance_kernel, dim3(1), dim3(64), 0, c->stream, key);
  LAUNCH_CHECK();
  return AFQMC_OK;
}

// ---------------------------------------------------------------- allocation-free SR
// afqmc_sr_static: the resampling of afqmc_sr for callers that cannot allocate (code
// under graph capture, tight serving loops).  It needs no context: the cumulative
// weights and the indices -- the scratch that a context's workspace holds for afqmc_sr --
// live in the file-scope array below, which the runtime places on every device the
// library runs on for the life of the process, and the walkers are gathered out of place
// into the caller's output buffers.
namespace {

constexpr int64_t SR_STATIC_MAX_WALKERS = int64_t(1) << 20;
// cumulative weights (W float64), then indices (W int32): 12 MiB
constexpr int64_t SR_STATIC_BYTES = SR_STATIC_MAX_WALKERS * (sizeof(double) + sizeof(int32_t));
static __device__ __align__(256) unsigned char scratch[SR_STATIC_BYTES];

constexpr int SR_STATIC_MAX_DEVICES = 64;
std::mutex g_scratch_mutex;
void* g_scratch_ptr[SR_STATIC_MAX_DEVICES] = {};

// The device address of `scratch` on the current device, looked up once per device with
// that device current (the array's host-side name is not a device address).
int scratch_for_current_device(void** out) {
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  if (dev < 0 || dev >= SR_STATIC_MAX_DEVICES)
    return afqmc_set_error(AFQMC_ERR_ARG, "device %d outside [0, %d)", dev,
                           SR_STATIC_MAX_DEVICES);
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return afqmc_set_error(AFQMC_ERR_HIP, "hipGetSymbolAddress returned NULL");
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return AFQMC_OK;
}

// dst[i * elems + k] = src[idx[i] * elems + k] over all W * elems entries
__global__ void gather_flat_kernel(int64_t total, int elems, const int32_t* __restrict__ idx,
                                   const cplx* __restrict__ src, cplx* __restrict__ dst) {
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = e / elems;
    dst[e] = src[(int64_t)idx[i] * elems + (e - i * elems)];
  }
}

}  // namespace

extern "C" int afqmc_sr_static(int32_t n_walkers, int32_t elems_up, int32_t elems_dn,
                               const void* walkers_up, const void* walkers_dn, void* out_up,
                               void* out_dn, double* weights, double zeta, int32_t* idx_out,
                               int32_t flags, void* stream) {
  if (n_walkers < 1 || n_walkers > SR_STATIC_MAX_WALKERS)
    return afqmc_set_error(AFQMC_ERR_SIZE,
                           "afqmc_sr_static: n_walkers=%d outside [1, %lld] (its fixed "
                           "scratch holds %lld walkers)",
                           n_walkers, (long long)SR_STATIC_MAX_WALKERS,
                           (long long)SR_STATIC_MAX_WALKERS);
  if (elems_up < 1 || !walkers_up || !out_up || !weights)
    return afqmc_set_error(AFQMC_ERR_ARG, "afqmc_sr_static: null walkers/out/weights or "
                                          "elems_up < 1");
  if (elems_dn < 0 || (elems_dn > 0 && (!walkers_dn || !out_dn)))
    return afqmc_set_error(AFQMC_ERR_ARG, "afqmc_sr_static: elems_dn > 0 needs both "
                                          "walkers_dn and out_dn");
  if (walkers_up == out_up || (elems_dn > 0 && walkers_dn == out_dn))
    return afqmc_set_error(AFQMC_ERR_ARG, "afqmc_sr_static gathers out of place: the "
                                          "output buffers must differ from the inputs");
  if (!(zeta >= 0.0 && zeta < 1.0))
    return afqmc_set_error(AFQMC_ERR_ARG, "zeta must lie in [0, 1)");
  if (flags & ~AFQMC_SR_GLOBAL_ORDER)
    return afqmc_set_error(AFQMC_ERR_ARG, "unknown sr flags");
  void* base = nullptr;
  int rc = scratch_for_current_device(&base);
  if (rc) return rc;
  hipStream_t st = static_cast<hipStream_t>(stream);
  // both scratch arrays are written in full before they are read (the running sum by the
  // index kernel's sequential pass, the indices by its searches): nothing to clear
  double* cum = static_cast<double*>(base);
  int32_t* idx = reinterpret_cast<int32_t*>(static_cast<unsigned char*>(base) +
                                            SR_STATIC_MAX_WALKERS * sizeof(double));
  hipLaunchKernelGGL(sr_index_kernel, dim3(1), dim3(SR_NT), 0, st, (int)n_walkers, zeta,
                     (flags & AFQMC_SR_GLOBAL_ORDER) ? 1 : 0, weights, cum, idx, idx_out,
                     (uint64_t*)nullptr);
  LAUNCH_CHECK();
  const int elems[2] = {elems_up, elems_dn};
  const cplx* src[2] = {static_cast<const cplx*>(walkers_up),
                        static_cast<const cplx*>(walkers_dn)};
  cplx* dst[2] = {static_cast<cplx*>(out_up), static_cast<cplx*>(out_dn)};
  for (int s = 0; s < 2; ++s) {
    if (elems[s] == 0) continue;
    const int64_t total = (int64_t)n_walkers * elems[s];
    const int64_t blocks = std::min<int64_t>((total + 255) / 256, 65536);
    hipLaunchKernelGGL(gather_flat_kernel, dim3((unsigned)blocks), dim3(256), 0, st, total,
                       elems[s], idx, src[s], dst[s]);
    LAUNCH_CHECK();
  }
  return AFQMC_OK;
}
