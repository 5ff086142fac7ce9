// Element-wise parts of the phaseless step (propagator.propagate, propagation.py:100-165):
//   fields_kernel  : field shifts, shifted fields, shift / force-bias terms
//                    (propagation.py:121-127) and the VHS GEMM operand i sqrt(dt) x~
//   weights_grouped_kernel : importance function, phaseless projection, clipping, weight
//                    cap and the population-control shift (propagation.py:133-164), one
//                    block per fragment; the single-context step applies the per-walker
//                    part in the measurement kernel and the shift in pcs_kernel
#include "common.h"

namespace {

constexpr int FIELDS_NT = 256;
constexpr int WEIGHTS_NT = 1024;

__device__ inline cplx overlap_of(const cplx* dets, int w, int nspin) {
  const cplx d0 = dets[2 * w];
  return nspin == 1 ? cmul(d0, d0) : cmul(d0, dets[2 * w + 1]);
}

// one workgroup per walker
// force bias of walker w, entry g: sum of the split-K slices in slice order
__device__ inline cplx fb_sum(const double* __restrict__ fb, int nparts, int64_t part_stride,
                              int w, int G, int g) {
  const double* p = fb + (int64_t)w * 2 * G + g;
  double re = p[0], im = p[G];
  for (int s = 1; s < nparts; ++s) {
    re += p[s * part_stride];
    im += p[s * part_stride + G];
  }
  return cplx{re, im};
}

// Four block sums in one pass, each with block_sum's tree (wave shuffle, then one wave
// over the per-wave partials: the same order and rounding); valid in thread 0 only.
template <int NT>
__device__ inline void block_sum4_first(double (&v)[4], double* red) {
  constexpr int NWV = NT / 64;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1)
#pragma unroll
    for (int c = 0; c < 4; ++c) v[c] += __shfl_down(v[c], off, 64);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) {
#pragma unroll
    for (int c = 0; c < 4; ++c) red[c * NWV + wid] = v[c];
  }
  __syncthreads();
  if (threadIdx.x < 64) {
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      double t = (lane < NWV) ? red[c * NWV + lane] : 0.0;
      for (int off = 32; off > 0; off >>= 1) t += __shfl_down(t, off, 64);
      v[c] = t;
    }
  }
}

// Fields of walker w by one workgroup of NT threads, NT = 256 or 64.  The sums st and ft
// always run in the order of 256 threads: virtual thread v adds its g = v, v + 256, ...;
// then each virtual wave of 64 is reduced by the wave tree and the four wave sums as
// (w0 + w2) + (w1 + w3) -- block_sum4_first's order.  With NT = 64 one wave carries the
// four virtual waves (no LDS, no barrier), so both forms give bit-identical results and
// the launch can take either.
template <int NT>
__device__ __forceinline__ void fields_body(
    int G, double sqdt, const double* __restrict__ fb, int nparts, int64_t part_stride,
    const cplx* __restrict__ mf, const double* __restrict__ fields, double* __restrict__ xs,
    uint16_t* __restrict__ xs_hi, uint16_t* __restrict__ xs_lo, int Gp,
    cplx* __restrict__ terms, cplx* __restrict__ fb_out, const int w) {
  constexpr int VT = FIELDS_NT;  // threads whose summation order is kept
  constexpr int R = VT / NT;     // virtual threads per thread
  static_assert(VT % NT == 0 && NT % 64 == 0, "whole waves");
  const double* x = fields + (int64_t)w * G;
  double* xr = xs + (int64_t)w * 2 * G;
  double* xi = xr + G;
  cplx st[R], ft[R];
#pragma unroll
  for (int r = 0; r < R; ++r) st[r] = ft[r] = cplx{0.0, 0.0};
  for (int g0 = 0; g0 < G; g0 += VT) {
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int g = g0 + (int)threadIdx.x + NT * r;
      if (g >= G) continue;
      const cplx f = fb_sum(fb, nparts, part_stride, w, G, g);
      const cplx m = mf[g];
      // field_shifts = -sqrt(dt) * (1j * fb - mf_shifts)
      const cplx ifb_m = cplx{-f.im - m.re, f.re - m.im};
      const cplx fs = cplx{-sqdt * ifb_m.re, -sqdt * ifb_m.im};
      const double xv = x[g];
      // shifted_fields = fields - field_shifts
      const cplx xt = cplx{xv - fs.re, -fs.im};
      st[r] = cadd(st[r], cmul(xt, m));
      // fields * field_shifts - field_shifts * field_shifts / 2
      const cplx fs2 = cmul(fs, fs);
      ft[r] = cadd(ft[r], cplx{xv * fs.re - fs2.re / 2.0, xv * fs.im - fs2.im / 2.0});
      // VHS operand: 1j * sqrt(dt) * shifted_fields (float64, or split-bf16 planes)
      if (xs_hi) {
        const int64_t o = (int64_t)w * 2 * Gp + g;
        split_bf16(-sqdt * xt.im, &xs_hi[o], &xs_lo[o]);
        split_bf16(sqdt * xt.re, &xs_hi[o + Gp], &xs_lo[o + Gp]);
      } else {
        xr[g] = -sqdt * xt.im;
        xi[g] = sqdt * xt.re;
      }
      if (fb_out) fb_out[(int64_t)w * G + g] = f;
    }
  }
  double v[4];
  if constexpr (R == 1) {
    __shared__ double red[4 * (VT / 64)];
    v[0] = st[0].re;
    v[1] = st[0].im;
    v[2] = ft[0].re;
    v[3] = ft[0].im;
    block_sum4_first<VT>(v, red);
  } else {
    static_assert(R == 4, "four virtual waves");
    double ws[R][4];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      ws[r][0] = st[r].re;
      ws[r][1] = st[r].im;
      ws[r][2] = ft[r].re;
      ws[r][3] = ft[r].im;
#pragma unroll
      for (int off = 32; off > 0; off >>= 1)
#pragma unroll
        for (int c = 0; c < 4; ++c) ws[r][c] += __shfl_down(ws[r][c], off, 64);
    }
    // block_sum4_first's second stage over the wave sums at lanes 0..3 of one wave:
    // each sum first meets zeros (x + 0), then (w0 + w2) + (w1 + w3)
#pragma unroll
    for (int c = 0; c < 4; ++c)
      v[c] = ((ws[0][c] + 0.0) + (ws[2][c] + 0.0)) + ((ws[1][c] + 0.0) + (ws[3][c] + 0.0));
  }
  if (threadIdx.x == 0) {
    terms[2 * w] = cplx{v[0], v[1]};
    terms[2 * w + 1] = cplx{v[2], v[3]};
  }
}

template <int NT>
__global__ __launch_bounds__(NT) void fields_kernel(
    int G, double sqdt, const double* __restrict__ fb, int nparts, int64_t part_stride,
    const cplx* __restrict__ mf, const double* __restrict__ fields, double* __restrict__ xs,
    uint16_t* __restrict__ xs_hi, uint16_t* __restrict__ xs_lo, int Gp,
    cplx* __restrict__ terms, cplx* __restrict__ fb_out) {
  fields_body<NT>(G, sqdt, fb, nparts, part_stride, mf, fields, xs, xs_hi, xs_lo, Gp, terms,
                  fb_out, blockIdx.x);
}

// one block per walker of every fragment; step selects the fields of this step
template <int NT>
__global__ __launch_bounds__(NT) void fields_grouped_kernel(
    const FieldsArgs* __restrict__ args, const int* __restrict__ start, int nprob, int step) {
  const int p = group_of(start, nprob, blockIdx.x);
  const FieldsArgs a = args[p];
  fields_body<NT>(a.G, a.sqdt, a.fb, a.nparts, a.part_stride, a.mf,
                  a.fields + (int64_t)step * a.W * a.G, a.xs, a.xs_hi, a.xs_lo, a.Gp, a.terms,
                  nullptr, blockIdx.x - start[p]);
}

// threads per walker of the fields kernels: one wave when there are walkers enough to
// fill the chip with waves, else four (the results do not depend on the choice)
inline int fields_nt(int64_t walkers) { return walkers >= 8192 ? 64 : FIELDS_NT; }

// single workgroup over all walkers (the weight sum needs every walker)
__device__ __forceinline__ void weights_body(
    int W, int nspin, double dt, double sqdt, double h0_prop, const cplx* __restrict__ dets,
    const cplx* __restrict__ terms, double* __restrict__ weights, cplx* __restrict__ overlaps,
    double* __restrict__ scalars) {
  __shared__ double red[WEIGHTS_NT / 64];
  const double e_estimate = scalars[0];
  const double pcs = scalars[1];
  const WeightsArgs a{W, nspin, dt, sqdt, h0_prop, dets, terms, weights, overlaps, scalars};
  double part = 0.0;
  for (int w = threadIdx.x; w < W; w += WEIGHTS_NT)
    part += walker_weight_update(a, pcs, w, overlap_of(dets, w, nspin));
  const double total = block_sum<WEIGHTS_NT>(part, red);
  if (threadIdx.x == 0) scalars[1] = e_estimate - 0.1 * (log(total / (double)W) / dt);
}

// population-control shift from weights the measurement kernel already updated
// (propagation.py:161-164): same strided sum and tree as weights_body
__global__ __launch_bounds__(WEIGHTS_NT) void pcs_kernel(int W, double dt,
                                                         const double* __restrict__ weights,
                                                         double* __restrict__ scalars) {
  __shared__ double red[WEIGHTS_NT / 64];
  double part = 0.0;
  for (int w = threadIdx.x; w < W; w += WEIGHTS_NT) part += weights[w];
  const double total = block_sum<WEIGHTS_NT>(part, red);
  if (threadIdx.x == 0) scalars[1] = scalars[0] - 0.1 * (log(total / (double)W) / dt);
}

// one block per fragment
__global__ __launch_bounds__(WEIGHTS_NT) void weights_grouped_kernel(
    const WeightsArgs* __restrict__ args) {
  const WeightsArgs a = args[blockIdx.x];
  weights_body(a.W, a.nspin, a.dt, a.sqdt, a.h0_prop, a.dets, a.terms, a.weights, a.overlaps,
               a.scalars);
}

// sampling.py:79-91: clamp, weighted block energy, population-control mix
__global__ __launch_bounds__(WEIGHTS_NT) void block_stats_kernel(
    int W, double dt, const cplx* __restrict__ energies, const double* __restrict__ weights,
    double* __restrict__ scalars, double* __restrict__ out) {
  __shared__ double red[WEIGHTS_NT / 64];
  const double e_est = scalars[0];
  const double cap = sqrt(2.0 / dt);
  double sw = 0.0, sew = 0.0, nz = 0.0, ncl = 0.0;
  for (int w = threadIdx.x; w < W; w += WEIGHTS_NT) {
    double e = energies[w].re;
    if (fabs(e - e_est) > cap) {
      e = e_est;
      ncl += 1.0;
    }
    const double wt = weights[w];
    sw += wt;
    sew += e * wt;
    nz += (wt == 0.0) ? 1.0 : 0.0;
  }
  const double bw = block_sum<WEIGHTS_NT>(sw, red);
  const double bew = block_sum<WEIGHTS_NT>(sew, red);
  const double bz = block_sum<WEIGHTS_NT>(nz, red);
  const double bc = block_sum<WEIGHTS_NT>(ncl, red);
  if (threadIdx.x == 0) {
    const double be = bew / bw;
    out[0] = be;
    out[1] = bw;
    out[2] = bz;
    out[3] = bc;
    scalars[1] = 0.9 * scalars[1] + 0.1 * be;
  }
}

// The driver's estimate update after a sampler call (driver.py:308,
// lno_afqmc/ccsd_pt2/run_afqmc_nompi.py:89): e_estimate <- 0.9 e_estimate + 0.1 E with
// E = sum_b E_b w_b / sum_b w_b over the call's nblocks rows of block statistics
// (rows of `stride` doubles: block energy, block weight, ...), summed in row order;
// then the next call's start, pop_control_ene_shift <- e_estimate (sampling.py:329).
__global__ void estimate_mix_kernel(int nblocks, int stride, const double* __restrict__ stats,
                                    double* __restrict__ scalars) {
  if (threadIdx.x != 0) return;
  double sew = 0.0, sw = 0.0;
  for (int b = 0; b < nblocks; ++b) {
    sew += stats[(int64_t)b * stride] * stats[(int64_t)b * stride + 1];
    sw += stats[(int64_t)b * stride + 1];
  }
  const double e = 0.9 * scalars[0] + 0.1 * (sew / sw);
  scalars[0] = e;
  scalars[1] = e;
}

// corr_sample.py:551-585 (corr_otler_rm): z-scores of the weights of two correlated
// populations, z = |w - mean| / std (population std); a walker whose larger z-score
// exceeds z_thresh gets weight 0 in both populations.  Two-pass mean / variance with
// the deterministic block reduction.
__global__ __launch_bounds__(WEIGHTS_NT) void cs_outlier_kernel(int W, double z_thresh,
                                                                double* __restrict__ w1,
                                                                double* __restrict__ w2) {
  __shared__ double red[WEIGHTS_NT / 64];
  double s1 = 0.0, s2 = 0.0;
  for (int i = threadIdx.x; i < W; i += WEIGHTS_NT) {
    s1 += w1[i];
    s2 += w2[i];
  }
  const double m1 = block_sum<WEIGHTS_NT>(s1, red) / (double)W;
  const double m2 = block_sum<WEIGHTS_NT>(s2, red) / (double)W;
  double v1 = 0.0, v2 = 0.0;
  for (int i = threadIdx.x; i < W; i += WEIGHTS_NT) {
    v1 += (w1[i] - m1) * (w1[i] - m1);
    v2 += (w2[i] - m2) * (w2[i] - m2);
  }
  const double sd1 = sqrt(block_sum<WEIGHTS_NT>(v1, red) / (double)W);
  const double sd2 = sqrt(block_sum<WEIGHTS_NT>(v2, red) / (double)W);
  for (int i = threadIdx.x; i < W; i += WEIGHTS_NT) {
    const double z1 = fabs(w1[i] - m1) / sd1, z2 = fabs(w2[i] - m2) / sd2;
    const double z = z1 > z2 ? z1 : z2;  // a NaN z1 picks z2, as jnp.where does
    if (z > z_thresh) {
      w1[i] = 0.0;
      w2[i] = 0.0;
    }
  }
}

__global__ void overlap_combine_kernel(int W, int nspin, const cplx* __restrict__ dets,
                                       cplx* __restrict__ overlaps) {
  const int w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w < W) overlaps[w] = overlap_of(dets, w, nspin);
}

__global__ __launch_bounds__(OVC_NT) void overlap_combine_grouped_kernel(
    const OvcArgs* __restrict__ args, const int* __restrict__ start, int nprob) {
  const int p = group_of(start, nprob, blockIdx.x);
  const OvcArgs a = args[p];
  const int w = (blockIdx.x - start[p]) * OVC_NT + threadIdx.x;
  if (w < a.W) a.out[w] = overlap_of(a.dets, w, a.nspin);
}

__global__ void fb_out_kernel(int W, int G, const double* __restrict__ fb, int nparts,
                              int64_t part_stride, cplx* __restrict__ out) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e < (int64_t)W * G) {
    const int w = (int)(e / G), g = (int)(e - (int64_t)w * G);
    out[e] = fb_sum(fb, nparts, part_stride, w, G, g);
  }
}

}  // namespace

int launch_fields(afqmc_ctx* c, int W, const double* fields, cplx* fb_out) {
  const auto kern = fields_nt(W) == 64 ? fields_kernel<64> : fields_kernel<FIELDS_NT>;
  hipLaunchKernelGGL(kern, dim3(W), dim3(fields_nt(W)), 0, c->stream, c->d.nchol,
                     sqrt(c->d.dt), c->fb, c->fb_parts, (int64_t)2 * W * c->d.nchol, c->mf_shifts,
                     fields, c->xs, c->mixed ? c->xs_bf[0] : nullptr,
                     c->mixed ? c->xs_bf[1] : nullptr, c->Gp, c->terms, fb_out);
  LAUNCH_CHECK();
  return AFQMC_OK;
}

WeightsArgs weights_args(const afqmc_ctx* c, int W, double* weights, cplx* overlaps,
                         double* scalars) {
  return WeightsArgs{W, c->nspin, c->d.dt, sqrt(c->d.dt), c->h0_prop, c->dets, c->terms,
                     weights, overlaps, scalars};
}

int launch_pcs(afqmc_ctx* c, int W, const double* weights, double* scalars) {
  hipLaunchKernelGGL(pcs_kernel, dim3(1), dim3(WEIGHTS_NT), 0, c->stream, W, c->d.dt, weights,
                     scalars);
  LAUNCH_CHECK();
  return AFQMC_OK;
}

int launch_block_stats(afqmc_ctx* c, int W, const cplx* energies, const double* weights,
                       double* scalars, double* out) {
  hipLaunchKernelGGL(block_stats_kernel, dim3(1), dim3(WEIGHTS_NT), 0, c->stream, W, c->d.dt,
                     energies, weights, scalars, out);
  LAUNCH_CHECK();
  return AFQMC_OK;
}

int launch_estimate_mix(afqmc_ctx* c, int nblocks, int stride, const double* stats,
                        double* scalars) {
  hipLaunchKernelGGL(estimate_mix_kernel, dim3(1), dim3(64), 0, c->stream, nblocks, stride, stats,
                     scalars);
  LAUNCH_CHECK();
  return AFQMC_OK;
}

int launch_overlap_combine_grouped(hipStream_t st, const OvcArgs* args, const int* start,
                                   int nprob, int nblocks) {
  hipLaunchKernelGGL(overlap_combine_grouped_kernel, dim3(nblocks), dim3(OVC_NT), 0, st, args,
                     start, nprob);
  LAUNCH_CHECK();
  return AFQMC_OK;
}

int launch_overlap_combine(afqmc_ctx* c, int W, cplx* overlaps) {
  hipLaunchKernelGGL(overlap_combine_kernel, dim3((W + 255) / 256), dim3(256), 0, c->stream, W,
                     c->nspin, c->dets, overlaps);
  LAUNCH_CHECK();
  return AFQMC_OK;
}

int launch_fb_out(afqmc_ctx* c, int W, cplx* fb_out) {
  const int64_t total = (int64_t)W * c->d.nchol;
  hipLaunchKernelGGL(fb_out_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0,
                     c->stream, W, c->d.nchol, c->fb, c->fb_parts, (int64_t)2 * W * c->d.nchol,
                     fb_out);
  LAUNCH_CHECK();
  return AFQMC_OK;
}

int launch_fields_grouped(hipStream_t st, const FieldsArgs* args, const int* start, int nprob,
                          int nblocks, int step) {
  const auto kern = fields_nt(nblocks) == 64 ? fields_grouped_kernel<64>
                                               : fields_grouped_kernel<FIELDS_NT>;
  hipLaunchKernelGGL(kern, dim3(nblocks), dim3(fields_nt(nblocks)), 0, st, args, start, nprob,
                     step);
  LAUNCH_CHECK();
  return AFQMC_OK;
}

int launch_weights_grouped(hipStream_t st, const WeightsArgs* args, int nprob) {
  hipLaunchKernelGGL(weights_grouped_kernel, dim3(nprob), dim3(WEIGHTS_NT), 0, st, args);
  LAUNCH_CHECK();
  return AFQMC_OK;
}

int launch_cs_outliers(afqmc_ctx* c, int W, double* w1, double* w2, double z_thresh) {
  hipLaunchKernelGGL(cs_outlier_kernel, dim3(1), dim3(WEIGHTS_NT), 0, c->stream, W, z_thresh, w1,
                     w2);
  LAUNCH_CHECK();
  return AFQMC_OK;
}
