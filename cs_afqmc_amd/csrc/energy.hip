// Local energy, rhf._calc_energy_restricted (wavefunctions.py:529-541) and
// uhf._calc_energy (wavefunctions.py:711-743):
//   f_g = rot_chol_s[g] G_s^T (n x n),  c_g = sum_s m_s tr f_g,  x_g = sum_s m_s tr(f_g f_g)
//   E = sum_g (c_g^2 - x_g) / 2 + sum_s m_s sum G_s * rot_h1_s + h0
// with m_s = 2 for restricted walkers (one spin standing for both) and 1 otherwise;
// for m = 2 this is 2 sum c^2 - exc + 2 sum G rot_h1 + h0, the restricted formula.
//
// The f_g for a chunk of walkers come from one FP64 MFMA GEMM per spin:
//   F[(g,i)][(w,c,k)] = sum_p rot_chol[g][i][p] * green[w][c][k][p]
// (rows of the split Green buffer are (w, re/im, k) with stride N), then a reduction
// kernel forms c_g, x_g and the energy per walker.
#include "common.h"

namespace {

constexpr int ENER_NT = 256;

// one workgroup per walker of the chunk; wave v handles g = v, v+4, ...
__global__ __launch_bounds__(ENER_NT) void energy_reduce_kernel(
    int G, int N, int nspin, int mult, int n0, int n1, int wc, const double* __restrict__ F0,
    const double* __restrict__ F1, const double* __restrict__ green0,
    const double* __restrict__ green1, const double* __restrict__ rot_h1_0,
    const double* __restrict__ rot_h1_1, double h0, int w_base, cplx* __restrict__ energies) {
  __shared__ double red[ENER_NT / 64];
  const int wl = blockIdx.x;  // walker within chunk
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int ns[2] = {n0, n1};
  const double* Fs[2] = {F0, F1};
  const double* gs[2] = {green0, green1};
  const double* hs[2] = {rot_h1_0, rot_h1_1};

  // two-body part
  double e2r = 0.0, e2i = 0.0;  // accumulated in lane 0 of each wave
  for (int g = wave; g < G; g += ENER_NT / 64) {
    double cr = 0.0, ci = 0.0, xr = 0.0, xi = 0.0;
    for (int s = 0; s < nspin; ++s) {
      const int n = ns[s];
      const int64_t ldc = (int64_t)wc * 2 * n;
      const double* base = Fs[s] + (int64_t)g * n * ldc + (int64_t)wl * 2 * n;
      double tr = 0.0, ti = 0.0, ar = 0.0, ai = 0.0;
      for (int e = lane; e < n * n; e += 64) {
        const int i = e / n, j = e - (e / n) * n;
        const double fr = base[i * ldc + j], fi = base[i * ldc + n + j];
        const double tr2 = base[j * ldc + i], ti2 = base[j * ldc + n + i];
        ar += fr * tr2 - fi * ti2;
        ai += fr * ti2 + fi * tr2;
        if (i == j) {
          tr += fr;
          ti += fi;
        }
      }
      cr += mult * tr;
      ci += mult * ti;
      xr += mult * ar;
      xi += mult * ai;
    }
    for (int off = 32; off > 0; off >>= 1) {
      cr += __shfl_down(cr, off, 64);
      ci += __shfl_down(ci, off, 64);
      xr += __shfl_down(xr, off, 64);
      xi += __shfl_down(xi, off, 64);
    }
    if (lane == 0) {
      e2r += (cr * cr - ci * ci - xr) / 2.0;
      e2i += (2.0 * cr * ci - xi) / 2.0;
    }
  }
  // one-body part: sum_s m_s sum_{k,p} G_s[k][p] rot_h1_s[k][p]
  double e1r = 0.0, e1i = 0.0;
  for (int s = 0; s < nspin; ++s) {
    const int kd = ns[s] * N;
    const double* gr = gs[s] + (int64_t)(w_base + wl) * 2 * kd;
    const double* gi = gr + kd;
    for (int e = threadIdx.x; e < kd; e += ENER_NT) {
      e1r += mult * gr[e] * hs[s][e];
      e1i += mult * gi[e] * hs[s][e];
    }
  }
  const double E2r = block_sum<ENER_NT>(e2r, red);
  const double E2i = block_sum<ENER_NT>(e2i, red);
  const double E1r = block_sum<ENER_NT>(e1r, red);
  const double E1i = block_sum<ENER_NT>(e1i, red);
  if (threadIdx.x == 0) energies[w_base + wl] = cplx{E2r + E1r + h0, E2i + E1i};
}

}  // namespace

int launch_energy(afqmc_ctx* c, int W, cplx* energies) {
  const int N = c->d.norb, G = c->d.nchol;
  const int wc_max = c->energy_wchunk;
  double* Fs[2] = {c->energy_f, nullptr};
  if (c->nspin == 2) {
    const int64_t f0 = (int64_t)G * c->nocc[0] * wc_max * 2 * c->nocc[0];
    Fs[1] = c->energy_f + f0;
  }
  for (int w0 = 0; w0 < W; w0 += wc_max) {
    const int wc = (W - w0 < wc_max) ? (W - w0) : wc_max;
    for (int s = 0; s < c->nspin; ++s) {
      const int n = c->nocc[s];
      // F[(g,i)][(w,c,k)], M = G*n, Ncols = wc*2*n, K = N
      int rc = launch_gemm_nt_f64(c->stream, G * n, wc * 2 * n, N, 1.0, c->rot_chol[s], N,
                                  c->green[s] + (int64_t)w0 * 2 * n * N, N, 0.0, Fs[s],
                                  wc * 2 * n);
      if (rc) return rc;
    }
    hipLaunchKernelGGL(energy_reduce_kernel, dim3(wc), dim3(ENER_NT), 0, c->stream, G, N,
                       c->nspin, c->mult, c->nocc[0], c->nocc[1], wc, Fs[0], Fs[1],
                       c->green[0], c->green[1], c->rot_h1[0], c->rot_h1[1], c->h0, w0,
                       energies);
    LAUNCH_CHECK();
  }
  return AFQMC_OK;
}
