// Grouped step over many fragments: independent AFQMC problems of uneven size (norb,
// nocc, nchol, walkers), one afqmc_ctx each, advanced together.  Each phase of the step
// is one launch over all fragments (or a few, one per kernel template key): problem p
// owns a range of blocks and reads its arguments from a device table, so small
// fragments fill the GPU together instead of one after another.
//
// Per step:  [Green of all walkers]  force-bias GEMMs  fields  VHS GEMMs  Taylor
//            new overlaps + Green  weights
// Tables are rebuilt when the member set changes and uploaded once; a step only passes
// the step index (which selects each fragment's fields in its bound fields block).
#include <algorithm>
#include <cstring>
#include <map>
#include <tuple>
#include <vector>

#include "common.h"

namespace {

struct Member {
  afqmc_ctx* c;
  cplx* wup;
  cplx* wdn;
  double* weights;
  cplx* overlaps;
  const double* fields;
  int n_steps;
  double* scalars;
};

struct Launch {
  int k0, k1, k2, k3;  // kernel key (meaning per table)
  size_t shm;
  int nprob, nblocks;
  size_t args_off, start_off;  // byte offsets into the device table
};

}  // namespace

struct afqmc_group {
  int device = 0;
  hipStream_t stream = nullptr;
  int max_frag = 0;
  std::vector<Member> mem;
  bool dirty = true;
  bool mixed = false;
  bool f32 = false;           // FP32 mode: the float64 operands on float32 MFMA
  int fb_bm = 0, vhs_bm = 0;  // GEMM tile edge (0: heuristic)
  int cg_bm = 0;              // tile edge of the energy's float64 c_g product (0: heuristic)
  bool f64_vec2 = true;
  // built tables
  std::vector<Launch> meas, taylor;
  Launch fb{}, vhs{}, fields{}, weights{};
  unsigned char* dev_tab = nullptr;
  size_t dev_cap = 0;
  std::vector<unsigned char> host_tab;
  unsigned char* pinned = nullptr;  // staging of the table upload
  size_t pinned_cap = 0;
  hipEvent_t upload_done = nullptr;  // the last upload has read `pinned`
  // side streams (default priority) for the kernel instances of one phase, which then
  // run side by side instead of one after the other; created on first use
  static constexpr int kSide = 3;
  hipStream_t side[kSide] = {};
  hipEvent_t fork = nullptr, join[kSide] = {};
  bool side_ready = false;
  // block ends (afqmc_group_orthonormalize / _overlap / _energy): tables built with the
  // step's; en_ptrs are the energies outputs the finishing table was built for
  Launch qr{}, ovc{}, fb64{}, fin{};
  std::vector<Launch> meas_ov, meas_en, exch;
  bool fb64_vec2 = true;
  std::vector<void*> en_ptrs;
  // block fields draws (afqmc_group_fields_normal_key), keys the table was built for
  Launch draw{};
  std::vector<void*> key_ptrs;
  ProfRec prof{};                    // afqmc_group_prof_*
};

namespace {

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

template <typename T>
void put_table(std::vector<unsigned char>& tab, Launch& L, const std::vector<T>& args,
               const std::vector<int>& start) {
  L.args_off = align256(tab.size());
  tab.resize(L.args_off + args.size() * sizeof(T));
  std::memcpy(tab.data() + L.args_off, args.data(), args.size() * sizeof(T));
  L.start_off = align256(tab.size());
  tab.resize(L.start_off + start.size() * sizeof(int));
  std::memcpy(tab.data() + L.start_off, start.data(), start.size() * sizeof(int));
}

int heuristic_bm(int64_t tiles128) { return tiles128 >= 512 ? 128 : 64; }

int build(afqmc_group* g) {
  g->host_tab.clear();
  g->f64_vec2 = true;
  g->meas.clear();
  g->taylor.clear();
  std::vector<unsigned char>& tab = g->host_tab;

  // ---- measurement (Green / overlaps): key (t2, stage)
  {
    std::map<std::pair<int, int>, std::vector<std::pair<MeasureArgs, int>>> by_key;
    std::map<std::pair<int, int>, size_t> shm;
    for (const Member& m : g->mem) {
      afqmc_ctx* c = m.c;
      for (int s = 0; s < c->nspin; ++s) {
        int t2, stage;
        size_t sh;
        measure_plan(c, s, &t2, &stage, &sh);
        MeasureArgs a{};
        a.N = c->d.norb;
        a.n = c->nocc[s];
        a.spin = s;
        a.ld16 = c->Np32;
        a.walkers = s == 0 ? m.wup : m.wdn;
        a.trial = c->trial[s];
        a.dets = c->dets;
        a.green = g->mixed ? nullptr : c->green[s];
        a.ghi = g->mixed ? c->green_bf[s][0] : nullptr;
        a.glo = g->mixed ? c->green_bf[s][1] : nullptr;
        const auto key = std::make_pair(t2, stage);
        by_key[key].push_back({a, c->d.n_walkers});
        shm[key] = std::max(shm[key], sh);
      }
    }
    for (auto& kv : by_key) {
      Launch L{};
      L.k0 = kv.first.first;
      L.k1 = kv.first.second;
      L.shm = shm[kv.first];
      std::vector<MeasureArgs> args;
      std::vector<int> start;
      int b = 0;
      for (auto& pa : kv.second) {
        args.push_back(pa.first);
        start.push_back(b);
        b += pa.second;
      }
      L.nprob = (int)args.size();
      L.nblocks = b;
      put_table(tab, L, args, start);
      g->meas.push_back(L);
    }
  }

  // ---- force-bias GEMMs: one problem per (fragment, spin); part s of fb gets spin s
  {
    std::vector<GemmArgs> args;
    std::vector<int> start;
    int64_t t128 = 0;
    for (const Member& m : g->mem)
      for (int s = 0; s < m.c->nspin; ++s)
        t128 += (int64_t)((2 * m.c->d.n_walkers + 127) / 128) * ((m.c->d.nchol + 127) / 128);
    const int bm = g->fb_bm ? g->fb_bm : heuristic_bm(t128);
    int b = 0;
    for (const Member& m : g->mem) {
      afqmc_ctx* c = m.c;
      const int W = c->d.n_walkers, G = c->d.nchol;
      for (int s = 0; s < c->nspin; ++s) {
        GemmArgs a{};
        a.M = 2 * W;
        a.N = G;
        a.alpha = (float)c->mult;
        a.ldc = G;
        a.C = c->fb + (int64_t)s * 2 * W * G;
        if (g->mixed) {
          a.K = c->nocc[s] * c->Np32;
          a.A = c->green_bf[s][0];
          a.Al = c->green_bf[s][1];
          a.B = c->rot_chol_bf[s][0];
          a.Bl = c->rot_chol_bf[s][1];
        } else {
          a.K = c->kdim[s];
          a.A = c->green[s];
          a.B = c->rot_chol[s];
          if (a.K % 2) g->f64_vec2 = false;
        }
        a.lda = a.ldb = a.K;
        a.tiles_m = (a.M + bm - 1) / bm;
        a.tiles_n = (a.N + bm - 1) / bm;
        args.push_back(a);
        start.push_back(b);
        b += a.tiles_m * a.tiles_n;
      }
    }
    g->fb = Launch{};
    g->fb.k0 = bm;
    g->fb.nprob = (int)args.size();
    g->fb.nblocks = b;
    put_table(tab, g->fb, args, start);
  }

  // ---- fields: one block per walker
  {
    std::vector<FieldsArgs> args;
    std::vector<int> start;
    int b = 0;
    for (const Member& m : g->mem) {
      afqmc_ctx* c = m.c;
      FieldsArgs a{};
      a.G = c->d.nchol;
      a.nparts = c->nspin;
      a.Gp = c->Gp;
      a.W = c->d.n_walkers;
      a.sqdt = sqrt(c->d.dt);
      a.part_stride = (int64_t)2 * a.W * a.G;
      a.fb = c->fb;
      a.mf = c->mf_shifts;
      a.fields = m.fields;
      a.xs = c->xs;
      a.xs_hi = g->mixed ? c->xs_bf[0] : nullptr;
      a.xs_lo = g->mixed ? c->xs_bf[1] : nullptr;
      a.terms = c->terms;
      args.push_back(a);
      start.push_back(b);
      b += a.W;
    }
    g->fields = Launch{};
    g->fields.nprob = (int)args.size();
    g->fields.nblocks = b;
    put_table(tab, g->fields, args, start);
  }

  // ---- VHS GEMMs: one problem per fragment (per spin with spin-resolved vectors)
  {
    std::vector<GemmArgs> args;
    std::vector<int> start;
    int64_t t128 = 0;
    for (const Member& m : g->mem)
      t128 += (int64_t)m.c->chol_spins * ((2 * m.c->d.n_walkers + 127) / 128) *
              (((int64_t)m.c->d.norb * m.c->d.norb + 127) / 128);
    const int bm = g->vhs_bm ? g->vhs_bm : heuristic_bm(t128);
    int b = 0;
    for (const Member& m : g->mem) {
      afqmc_ctx* c = m.c;
      const int W = c->d.n_walkers, N = c->d.norb;
      for (int cs = 0; cs < c->chol_spins; ++cs) {
        GemmArgs a{};
        a.M = 2 * W;
        a.N = g->mixed ? c->vhs_nv : N * N;
        a.alpha = 1.0f;
        a.ldc = a.N;
        if (g->mixed) {
          a.K = c->Gp;
          a.A = c->xs_bf[0];
          a.Al = c->xs_bf[1];
          a.B = chol_t_bf_of(c, 0, cs);
          a.Bl = chol_t_bf_of(c, 1, cs);
          a.C = vhs_f32_of(c, cs);
        } else {
          a.K = c->d.nchol;
          a.A = c->xs;
          a.B = chol_t_of(c, cs);
          a.C = vhs_of(c, cs);
          if (a.K % 2) g->f64_vec2 = false;
        }
        a.lda = a.ldb = a.K;
        a.tiles_m = (a.M + bm - 1) / bm;
        a.tiles_n = (a.N + bm - 1) / bm;
        args.push_back(a);
        start.push_back(b);
        b += a.tiles_m * a.tiles_n;
      }
    }
    g->vhs = Launch{};
    g->vhs.k0 = bm;
    g->vhs.nprob = (int)args.size();
    g->vhs.nblocks = b;
    put_table(tab, g->vhs, args, start);
  }

  // ---- Taylor: key (kind, nkgv, ct, block size)
  {
    std::map<std::tuple<int, int, int, int>, std::vector<std::pair<TaylorArgs, int>>> by_key;
    std::map<std::tuple<int, int, int, int>, size_t> shm;
    for (const Member& m : g->mem) {
      afqmc_ctx* c = m.c;
      for (int s = 0; s < c->nspin; ++s) {
        TaylorPlan pl;
        taylor_plan(c, s, &pl);
        TaylorArgs a{};
        a.N = c->d.norb;
        a.n = c->nocc[s];
        a.K = c->d.n_exp_terms - 1;
        a.rows_pad = ((a.N + 15) / 16) * 16;
        a.nv = g->mixed ? c->vhs_nv : a.N * a.N;
        a.walkers = s == 0 ? m.wup : m.wdn;
        a.E = c->exp_h1[s];
        if (pl.kind == TAYLOR_REG_BF16X3 || pl.kind == TAYLOR_STREAM_BF16X3) {
          a.E = c->emi_bf[s][0];
          a.E2 = c->emi_bf[s][1];
        }
        a.vhs = g->mixed ? (const void*)vhs_f32_of(c, s) : (const void*)vhs_of(c, s);
        const auto key = std::make_tuple(pl.kind, pl.nkgv, pl.ct, pl.nthreads);
        by_key[key].push_back({a, c->d.n_walkers * pl.blocks_per_walker});
        shm[key] = std::max(shm[key], pl.shm);
      }
    }
    for (auto& kv : by_key) {
      Launch L{};
      L.k0 = std::get<0>(kv.first);
      L.k1 = std::get<1>(kv.first);
      L.k2 = std::get<2>(kv.first);
      L.k3 = std::get<3>(kv.first);
      L.shm = shm[kv.first];
      std::vector<TaylorArgs> args;
      std::vector<int> start;
      int b = 0;
      for (auto& pa : kv.second) {
        args.push_back(pa.first);
        start.push_back(b);
        b += pa.second;
      }
      L.nprob = (int)args.size();
      L.nblocks = b;
      put_table(tab, L, args, start);
      g->taylor.push_back(L);
    }
  }
  // largest instances first (work ~ walkers x k blocks x column tiles x threads; for the
  // measurement ~ walkers x T2^2): they start on the group's stream, the smaller ones
  // run beside them
  std::stable_sort(g->taylor.begin(), g->taylor.end(), [](const Launch& a, const Launch& b) {
    return (int64_t)a.nblocks * a.k1 * a.k2 * a.k3 > (int64_t)b.nblocks * b.k1 * b.k2 * b.k3;
  });
  std::stable_sort(g->meas.begin(), g->meas.end(), [](const Launch& a, const Launch& b) {
    return (int64_t)a.nblocks * a.k0 * a.k0 > (int64_t)b.nblocks * b.k0 * b.k0;
  });

  // ---- weights: one block per fragment
  {
    std::vector<WeightsArgs> args;
    std::vector<int> start;
    for (const Member& m : g->mem) {
      afqmc_ctx* c = m.c;
      WeightsArgs a{};
      a.W = c->d.n_walkers;
      a.nspin = c->nspin;
      a.dt = c->d.dt;
      a.sqdt = sqrt(c->d.dt);
      a.h0_prop = c->h0_prop;
      a.dets = c->dets;
      a.terms = c->terms;
      a.weights = m.weights;
      a.overlaps = m.overlaps;
      a.scalars = m.scalars;
      start.push_back((int)args.size());
      args.push_back(a);
    }
    g->weights = Launch{};
    g->weights.nprob = (int)args.size();
    g->weights.nblocks = (int)args.size();
    put_table(tab, g->weights, args, start);
  }

  // ---- block ends
  {
    // QR: one block per walker and spin
    std::vector<QrArgs> qa;
    std::vector<int> qs;
    size_t qshm = 0;
    int b = 0;
    for (const Member& m : g->mem) {
      afqmc_ctx* c = m.c;
      for (int s = 0; s < c->nspin; ++s) {
        const size_t lb = qr_lds_bytes(c->d.norb, c->nocc[s]);
        QrArgs a{};
        a.m = c->d.norb;
        a.n = c->nocc[s];
        a.use_lds = lb > 0 ? 1 : 0;
        a.walkers = s == 0 ? m.wup : m.wdn;
        qshm = std::max(qshm, lb);
        qa.push_back(a);
        qs.push_back(b);
        b += c->d.n_walkers;
      }
    }
    g->qr = Launch{};
    g->qr.shm = qshm;
    g->qr.nprob = (int)qa.size();
    g->qr.nblocks = b;
    put_table(tab, g->qr, qa, qs);
  }
  {
    // overlaps (determinants only) and Green's functions for the energy (float64, plus
    // the split planes in the mixed mode): the step's measurement grouping
    for (int which = 0; which < 2; ++which) {
      std::vector<Launch>& out = which == 0 ? g->meas_ov : g->meas_en;
      out.clear();
      std::map<std::pair<int, int>, std::vector<std::pair<MeasureArgs, int>>> by_key;
      std::map<std::pair<int, int>, size_t> shm;
      for (const Member& m : g->mem) {
        afqmc_ctx* c = m.c;
        for (int s = 0; s < c->nspin; ++s) {
          int t2, stage;
          size_t sh;
          measure_plan(c, s, &t2, &stage, &sh);
          MeasureArgs a{};
          a.N = c->d.norb;
          a.n = c->nocc[s];
          a.spin = s;
          a.ld16 = c->Np32;
          a.walkers = s == 0 ? m.wup : m.wdn;
          a.trial = c->trial[s];
          a.dets = c->dets;
          a.green = which == 1 ? c->green[s] : nullptr;
          a.ghi = (which == 1 && c->mixed) ? c->green_bf[s][0] : nullptr;
          a.glo = (which == 1 && c->mixed) ? c->green_bf[s][1] : nullptr;
          const auto key = std::make_pair(t2, stage);
          by_key[key].push_back({a, c->d.n_walkers});
          shm[key] = std::max(shm[key], sh);
        }
      }
      for (auto& kv : by_key) {
        Launch L{};
        L.k0 = kv.first.first;
        L.k1 = kv.first.second;
        L.shm = shm[kv.first];
        std::vector<MeasureArgs> args;
        std::vector<int> start;
        int b = 0;
        for (auto& pa : kv.second) {
          args.push_back(pa.first);
          start.push_back(b);
          b += pa.second;
        }
        L.nprob = (int)args.size();
        L.nblocks = b;
        put_table(tab, L, args, start);
        out.push_back(L);
      }
      std::stable_sort(out.begin(), out.end(), [](const Launch& x, const Launch& y) {
        return (int64_t)x.nblocks * x.k0 * x.k0 > (int64_t)y.nblocks * y.k0 * y.k0;
      });
    }
  }
  {
    // overlap = product of the spins' determinants
    std::vector<OvcArgs> oa;
    std::vector<int> os;
    int b = 0;
    for (const Member& m : g->mem) {
      OvcArgs a{};
      a.W = m.c->d.n_walkers;
      a.nspin = m.c->nspin;
      a.dets = m.c->dets;
      a.out = m.overlaps;
      oa.push_back(a);
      os.push_back(b);
      b += (a.W + OVC_NT - 1) / OVC_NT;
    }
    g->ovc = Launch{};
    g->ovc.nprob = (int)oa.size();
    g->ovc.nblocks = b;
    put_table(tab, g->ovc, oa, os);
  }
  {
    // c_g of the energy: float64 force-bias products, part s of fb from spin s
    std::vector<GemmArgs> args;
    std::vector<int> start;
    int64_t t128 = 0;
    for (const Member& m : g->mem)
      for (int s = 0; s < m.c->nspin; ++s)
        t128 += (int64_t)((2 * m.c->d.n_walkers + 127) / 128) * ((m.c->d.nchol + 127) / 128);
    const int bm = g->cg_bm ? g->cg_bm : heuristic_bm(t128);
    g->fb64_vec2 = true;
    int b = 0;
    for (const Member& m : g->mem) {
      afqmc_ctx* c = m.c;
      const int W = c->d.n_walkers, G = c->d.nchol;
      for (int s = 0; s < c->nspin; ++s) {
        GemmArgs a{};
        a.M = 2 * W;
        a.N = G;
        a.K = c->kdim[s];
        a.alpha = (float)c->mult;
        a.A = c->green[s];
        a.B = c->rot_chol[s];
        a.C = c->fb + (int64_t)s * 2 * W * G;
        a.lda = a.ldb = a.K;
        a.ldc = G;
        if (a.K % 2) g->fb64_vec2 = false;
        a.tiles_m = (a.M + bm - 1) / bm;
        a.tiles_n = (a.N + bm - 1) / bm;
        args.push_back(a);
        start.push_back(b);
        b += a.tiles_m * a.tiles_n;
      }
    }
    g->fb64 = Launch{};
    g->fb64.k0 = bm;
    g->fb64.nprob = (int)args.size();
    g->fb64.nblocks = b;
    put_table(tab, g->fb64, args, start);
  }
  {
    // exchange sums, one launch per kernel variant
    g->exch.clear();
    std::map<int, std::vector<std::pair<ExchArgs, int>>> by_kind;
    for (const Member& m : g->mem)
      for (int s = 0; s < m.c->nspin; ++s) {
        ExchArgs a;
        int kind, nb;
        exchange_plan(m.c, s, m.c->d.n_walkers, &a, &kind, &nb);
        by_kind[kind].push_back({a, nb});
      }
    for (auto& kv : by_kind) {
      Launch L{};
      L.k0 = kv.first;
      std::vector<ExchArgs> args;
      std::vector<int> start;
      int b = 0;
      for (auto& pa : kv.second) {
        args.push_back(pa.first);
        start.push_back(b);
        b += pa.second;
      }
      L.nprob = (int)args.size();
      L.nblocks = b;
      put_table(tab, L, args, start);
      g->exch.push_back(L);
    }
    std::stable_sort(g->exch.begin(), g->exch.end(),
                     [](const Launch& x, const Launch& y) { return x.nblocks > y.nblocks; });
  }
  {
    // fields blocks: member i draws n_steps x W x nchol normals from its key
    std::vector<DrawArgs> da;
    std::vector<int> ds;
    int b = 0;
    for (size_t i = 0; i < g->mem.size(); ++i) {
      const Member& m = g->mem[i];
      DrawArgs a{};
      a.key = i < g->key_ptrs.size() ? static_cast<uint64_t*>(g->key_ptrs[i]) : nullptr;
      a.out = const_cast<double*>(m.fields);
      a.count = (int64_t)m.n_steps * m.c->d.n_walkers * m.c->d.nchol;
      da.push_back(a);
      ds.push_back(b);
      b += (int)(((a.count + 1) / 2 + 255) / 256);
    }
    g->draw = Launch{};
    g->draw.nprob = (int)da.size();
    g->draw.nblocks = b;
    put_table(tab, g->draw, da, ds);
  }
  {
    // energies: one block per walker (the force-bias parts are the spins' products)
    std::vector<FinishArgs> fa;
    std::vector<int> fs;
    int b = 0;
    for (size_t i = 0; i < g->mem.size(); ++i) {
      afqmc_ctx* c = g->mem[i].c;
      FinishArgs a;
      energy_finish_args(c, c->d.n_walkers,
                         i < g->en_ptrs.size() ? static_cast<cplx*>(g->en_ptrs[i]) : nullptr,
                         &a);
      a.nparts = c->nspin;
      fa.push_back(a);
      fs.push_back(b);
      b += c->d.n_walkers;
    }
    g->fin = Launch{};
    g->fin.nprob = (int)fa.size();
    g->fin.nblocks = b;
    put_table(tab, g->fin, fa, fs);
  }

  // Upload without draining the stream: launches already queued read the old table
  // before this stream-ordered copy overwrites it; the pinned staging buffer is reused
  // once the previous copy has read it.
  if (!g->upload_done) HIP_TRY(hipEventCreateWithFlags(&g->upload_done, hipEventDisableTiming));
  if (tab.size() > g->dev_cap) {
    HIP_TRY(hipStreamSynchronize(g->stream));
    if (g->dev_tab) HIP_TRY(hipFree(g->dev_tab));
    g->dev_tab = nullptr;
    g->dev_cap = 0;
    HIP_TRY(hipMalloc(&g->dev_tab, 2 * tab.size()));
    g->dev_cap = 2 * tab.size();
  }
  HIP_TRY(hipEventSynchronize(g->upload_done));
  if (tab.size() > g->pinned_cap) {
    if (g->pinned) HIP_TRY(hipHostFree(g->pinned));
    g->pinned = nullptr;
    g->pinned_cap = 0;
    HIP_TRY(hipHostMalloc(reinterpret_cast<void**>(&g->pinned), 2 * tab.size(),
                          hipHostMallocDefault));
    g->pinned_cap = 2 * tab.size();
  }
  std::memcpy(g->pinned, tab.data(), tab.size());
  HIP_TRY(hipMemcpyAsync(g->dev_tab, g->pinned, tab.size(), hipMemcpyHostToDevice, g->stream));
  HIP_TRY(hipEventRecord(g->upload_done, g->stream));
  g->dirty = false;
  return AFQMC_OK;
}

template <typename T>
const T* targs(const afqmc_group* g, const Launch& L) {
  return reinterpret_cast<const T*>(g->dev_tab + L.args_off);
}
const int* tstart(const afqmc_group* g, const Launch& L) {
  return reinterpret_cast<const int*>(g->dev_tab + L.start_off);
}

int ensure_side(afqmc_group* g) {
  if (g->side_ready) return AFQMC_OK;
  for (int i = 0; i < afqmc_group::kSide; ++i) {
    HIP_TRY(hipStreamCreateWithFlags(&g->side[i], hipStreamNonBlocking));
    HIP_TRY(hipEventCreateWithFlags(&g->join[i], hipEventDisableTiming));
  }
  HIP_TRY(hipEventCreateWithFlags(&g->fork, hipEventDisableTiming));
  g->side_ready = true;
  return AFQMC_OK;
}

// launch(L, stream) for every launch of one phase (ordered by decreasing work): the
// first on the group's stream, the others round robin on side streams that wait for the
// work queued so far and are joined back, so the instances overlap (their tails and the
// small instances fill the chip while the large ones run).  Works under graph capture:
// the side streams join the capture through the fork event.
template <typename F>
int run_spread(afqmc_group* g, const std::vector<Launch>& ls, F&& launch) {
  const int n = (int)ls.size();
  if (n <= 1) return n == 1 ? launch(ls[0], g->stream) : AFQMC_OK;
  int rc = ensure_side(g);
  if (rc) return rc;
  const int used = std::min(n - 1, (int)afqmc_group::kSide);
  HIP_TRY(hipEventRecord(g->fork, g->stream));
  for (int i = 0; i < used; ++i) HIP_TRY(hipStreamWaitEvent(g->side[i], g->fork, 0));
  for (int i = 0; i < n && !rc; ++i)
    rc = launch(ls[i], i == 0 ? g->stream : g->side[(i - 1) % used]);
  for (int i = 0; i < used; ++i) {
    HIP_TRY(hipEventRecord(g->join[i], g->side[i]));
    HIP_TRY(hipStreamWaitEvent(g->stream, g->join[i], 0));
  }
  return rc;
}

int run_measure(afqmc_group* g) {
  return run_spread(g, g->meas, [&](const Launch& L, hipStream_t st) {
    return launch_measure_grouped(st, L.k0, L.k1, targs<MeasureArgs>(g, L), tstart(g, L),
                                  L.nprob, L.nblocks, L.shm);
  });
}

int run_gemm(afqmc_group* g, const Launch& L, int out_f32) {
  if (g->mixed)
    return launch_gemm_bf16x3_grouped(g->stream, L.k0, out_f32, targs<GemmArgs>(g, L),
                                      tstart(g, L), L.nprob, L.nblocks);
  if (g->f32)
    return launch_gemm_f32_grouped(g->stream, L.k0, g->f64_vec2, targs<GemmArgs>(g, L),
                                   tstart(g, L), L.nprob, L.nblocks);
  return launch_gemm_f64_grouped(g->stream, L.k0, g->f64_vec2, targs<GemmArgs>(g, L),
                                 tstart(g, L), L.nprob, L.nblocks);
}

int set_group_device(const afqmc_group* g) {
  HIP_TRY(hipSetDevice(g->device));
  return AFQMC_OK;
}

}  // namespace

extern "C" int afqmc_group_create(afqmc_group** out, int32_t max_fragments, void* stream) {
  if (!out) return afqmc_set_error(AFQMC_ERR_ARG, "null argument");
  *out = nullptr;
  if (max_fragments < 1) return afqmc_set_error(AFQMC_ERR_ARG, "max_fragments=%d", max_fragments);
  afqmc_group* g = new afqmc_group();
  g->max_frag = max_fragments;
  g->stream = static_cast<hipStream_t>(stream);
  if (hipGetDevice(&g->device) != hipSuccess) {
    delete g;
    return afqmc_set_error(AFQMC_ERR_HIP, "hipGetDevice failed");
  }
  *out = g;
  return AFQMC_OK;
}

extern "C" int afqmc_group_destroy(afqmc_group* g) {
  if (!g) return AFQMC_OK;
  (void)hipSetDevice(g->device);
  if (g->dev_tab || g->pinned) (void)hipStreamSynchronize(g->stream);
  if (g->dev_tab) (void)hipFree(g->dev_tab);
  if (g->pinned) (void)hipHostFree(g->pinned);
  if (g->upload_done) (void)hipEventDestroy(g->upload_done);
  if (g->side_ready) {
    for (int i = 0; i < afqmc_group::kSide; ++i) {
      (void)hipStreamSynchronize(g->side[i]);
      (void)hipStreamDestroy(g->side[i]);
      (void)hipEventDestroy(g->join[i]);
    }
    (void)hipEventDestroy(g->fork);
  }
  prof_free(&g->prof);
  delete g;
  return AFQMC_OK;
}

extern "C" int afqmc_group_set_stream(afqmc_group* g, void* stream) {
  if (!g) return afqmc_set_error(AFQMC_ERR_ARG, "null group");
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (st != g->stream && g->dev_tab) {
    // launches queued on the old stream may still read the table, so it is not replaced
    // before they finish.  A stream that is capturing a graph runs nothing now and may
    // not be synchronised against: the captured launches read the table as it stands
    // (a rebuild would synchronise, which fails the capture), and ordering the graph's
    // replays after earlier work is the caller's, as for any graph.
    int rc = set_group_device(g);
    if (rc) return rc;
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    HIP_TRY(hipStreamIsCapturing(st, &cs));
    if (cs != hipStreamCaptureStatusActive) HIP_TRY(hipStreamSynchronize(g->stream));
  }
  g->stream = st;
  return AFQMC_OK;
}

extern "C" int afqmc_group_clear(afqmc_group* g) {
  if (!g) return afqmc_set_error(AFQMC_ERR_ARG, "null group");
  g->mem.clear();
  g->dirty = true;
  return AFQMC_OK;
}

extern "C" int afqmc_group_add(afqmc_group* g, afqmc_ctx* c, void* walkers_up, void* walkers_dn,
                               double* weights, void* overlaps, const double* fields_block,
                               int32_t n_steps, double* scalars) {
  if (!g || !c) return afqmc_set_error(AFQMC_ERR_ARG, "null group or context");
  if ((int)g->mem.size() >= g->max_frag)
    return afqmc_set_error(AFQMC_ERR_SIZE, "group holds at most %d fragments", g->max_frag);
  if (!walkers_up || (c->nspin == 2 && !walkers_dn) || !weights || !overlaps || !fields_block ||
      !scalars || n_steps < 1)
    return afqmc_set_error(AFQMC_ERR_ARG, "null buffer or n_steps < 1");
  const bool mixed = c->mixed != 0;
  const bool f32 = c->f32 != 0;
  if (!g->mem.empty() && (mixed != g->mixed || f32 != g->f32))
    return afqmc_set_error(AFQMC_ERR_ARG, "all fragments of a group share one precision mode");
  if (c->device != g->device)
    return afqmc_set_error(AFQMC_ERR_ARG, "context on device %d, group on device %d", c->device,
                           g->device);
  for (const Member& m : g->mem)
    if (m.c == c) return afqmc_set_error(AFQMC_ERR_ARG, "context already in the group");
  if (g->mem.empty()) {
    g->mixed = mixed;
    g->f32 = f32;
  }
  Member m{c, static_cast<cplx*>(walkers_up), static_cast<cplx*>(walkers_dn), weights,
           static_cast<cplx*>(overlaps), fields_block, n_steps, scalars};
  g->mem.push_back(m);
  g->dirty = true;
  return AFQMC_OK;
}

extern "C" int afqmc_group_size(const afqmc_group* g) { return g ? (int)g->mem.size() : 0; }

extern "C" int afqmc_group_propagate(afqmc_group* g, int32_t step, int32_t flags) {
  if (!g) return afqmc_set_error(AFQMC_ERR_ARG, "null group");
  if (g->mem.empty()) return AFQMC_OK;
  int rc = set_group_device(g);
  if (rc) return rc;
  for (const Member& m : g->mem)
    if (step < 0 || step >= m.n_steps)
      return afqmc_set_error(AFQMC_ERR_ARG, "step %d outside the bound fields block [0, %d)",
                             step, m.n_steps);
  if (g->dirty && (rc = build(g))) return rc;
  bool reuse = (flags & AFQMC_STEP_REUSE_GREEN) != 0;
  for (const Member& m : g->mem)
    if (m.c->green_src != m.wup || m.c->green_w != m.c->d.n_walkers) reuse = false;
  for (const Member& m : g->mem) m.c->green_src = nullptr;
  if (!reuse) {
    PhaseTimer t(&g->prof, g->stream, AFQMC_PH_GREEN);
    if ((rc = run_measure(g))) return rc;
  }
  {
    PhaseTimer t(&g->prof, g->stream, AFQMC_PH_FB_GEMM);
    if ((rc = run_gemm(g, g->fb, 0))) return rc;
  }
  for (const Member& m : g->mem) m.c->fb_parts = m.c->nspin;
  {
    PhaseTimer t(&g->prof, g->stream, AFQMC_PH_FIELDS);
    if ((rc = launch_fields_grouped(g->stream, targs<FieldsArgs>(g, g->fields),
                                    tstart(g, g->fields), g->fields.nprob, g->fields.nblocks,
                                    step)))
      return rc;
  }
  {
    PhaseTimer t(&g->prof, g->stream, AFQMC_PH_VHS_GEMM);
    if ((rc = run_gemm(g, g->vhs, 1))) return rc;
  }
  {
    PhaseTimer t(&g->prof, g->stream, AFQMC_PH_TAYLOR);
    rc = run_spread(g, g->taylor, [&](const Launch& L, hipStream_t st) {
      return launch_taylor_grouped(st, L.k0, L.k1, L.k2, targs<TaylorArgs>(g, L), tstart(g, L),
                                   L.nprob, L.nblocks, L.k3, L.shm);
    });
    if (rc) return rc;
  }
  {
    PhaseTimer t(&g->prof, g->stream, AFQMC_PH_OVERLAP);
    if ((rc = run_measure(g))) return rc;
  }
  for (const Member& m : g->mem) {
    m.c->green_src = m.wup;
    m.c->green_w = m.c->d.n_walkers;
  }
  PhaseTimer t(&g->prof, g->stream, AFQMC_PH_WEIGHTS);
  return launch_weights_grouped(g->stream, targs<WeightsArgs>(g, g->weights), g->weights.nprob);
}

// ---- block ends of every member (sampling.py:74-91 around each fragment's block):
// the per-context afqmc_orthonormalize, afqmc_overlap and afqmc_energy as grouped
// launches, each fragment's results exactly those of its own context's kernels

namespace {
int group_ready(afqmc_group* g) {
  int rc = set_group_device(g);
  if (rc) return rc;
  if (g->dirty && (rc = build(g))) return rc;
  for (const Member& m : g->mem) m.c->green_src = nullptr;
  return AFQMC_OK;
}
}  // namespace

extern "C" int afqmc_group_orthonormalize(afqmc_group* g) {
  if (!g) return afqmc_set_error(AFQMC_ERR_ARG, "null group");
  if (g->mem.empty()) return AFQMC_OK;
  int rc = group_ready(g);
  if (rc) return rc;
  PhaseTimer t(&g->prof, g->stream, AFQMC_PH_QR);
  return launch_qr_grouped(g->stream, targs<QrArgs>(g, g->qr), tstart(g, g->qr), g->qr.nprob,
                           g->qr.nblocks, g->qr.shm);
}

extern "C" int afqmc_group_overlap(afqmc_group* g) {
  if (!g) return afqmc_set_error(AFQMC_ERR_ARG, "null group");
  if (g->mem.empty()) return AFQMC_OK;
  int rc = group_ready(g);
  if (rc) return rc;
  rc = run_spread(g, g->meas_ov, [&](const Launch& L, hipStream_t st) {
    return launch_measure_grouped(st, L.k0, L.k1, targs<MeasureArgs>(g, L), tstart(g, L),
                                  L.nprob, L.nblocks, L.shm);
  });
  if (rc) return rc;
  return launch_overlap_combine_grouped(g->stream, targs<OvcArgs>(g, g->ovc), tstart(g, g->ovc),
                                        g->ovc.nprob, g->ovc.nblocks);
}

extern "C" int afqmc_group_energy(afqmc_group* g, void* const* energies) {
  if (!g) return afqmc_set_error(AFQMC_ERR_ARG, "null group");
  if (g->mem.empty()) return AFQMC_OK;
  if (!energies) return afqmc_set_error(AFQMC_ERR_ARG, "null energies");
  const size_t nm = g->mem.size();
  for (size_t i = 0; i < nm; ++i)
    if (!energies[i]) return afqmc_set_error(AFQMC_ERR_ARG, "null energies of member %d", (int)i);
  if (g->en_ptrs.size() != nm || !std::equal(g->en_ptrs.begin(), g->en_ptrs.end(), energies)) {
    g->en_ptrs.assign(energies, energies + nm);
    g->dirty = true;
  }
  int rc = group_ready(g);
  if (rc) return rc;
  rc = run_spread(g, g->meas_en, [&](const Launch& L, hipStream_t st) {
    return launch_measure_grouped(st, L.k0, L.k1, targs<MeasureArgs>(g, L), tstart(g, L),
                                  L.nprob, L.nblocks, L.shm);
  });
  if (rc) return rc;
  PhaseTimer t(&g->prof, g->stream, AFQMC_PH_ENERGY);
  // c_g = force bias, contracted in float64 as the single context's energy does
  if ((rc = launch_gemm_f64_grouped(g->stream, g->fb64.k0, g->fb64_vec2,
                                    targs<GemmArgs>(g, g->fb64), tstart(g, g->fb64),
                                    g->fb64.nprob, g->fb64.nblocks)))
    return rc;
  for (const Member& m : g->mem) m.c->fb_parts = m.c->nspin;
  rc = run_spread(g, g->exch, [&](const Launch& L, hipStream_t st) {
    return launch_exchange_grouped(st, L.k0, targs<ExchArgs>(g, L), tstart(g, L), L.nprob,
                                   L.nblocks);
  });
  if (rc) return rc;
  return launch_energy_finish_grouped(g->stream, targs<FinishArgs>(g, g->fin),
                                      tstart(g, g->fin), g->fin.nprob, g->fin.nblocks);
}

extern "C" int afqmc_group_fields_normal_key(afqmc_group* g, void* const* keys) {
  if (!g) return afqmc_set_error(AFQMC_ERR_ARG, "null group");
  if (g->mem.empty()) return AFQMC_OK;
  if (!keys) return afqmc_set_error(AFQMC_ERR_ARG, "null keys");
  const size_t nm = g->mem.size();
  for (size_t i = 0; i < nm; ++i)
    if (!keys[i]) return afqmc_set_error(AFQMC_ERR_ARG, "null key of member %d", (int)i);
  if (g->key_ptrs.size() != nm || !std::equal(g->key_ptrs.begin(), g->key_ptrs.end(), keys)) {
    g->key_ptrs.assign(keys, keys + nm);
    g->dirty = true;
  }
  int rc = set_group_device(g);
  if (rc) return rc;
  if (g->dirty && (rc = build(g))) return rc;
  PhaseTimer t(&g->prof, g->stream, AFQMC_PH_RNG);
  return launch_fields_normal_key_grouped(g->stream, targs<DrawArgs>(g, g->draw),
                                          tstart(g, g->draw), g->draw.nprob, g->draw.nblocks);
}

extern "C" int afqmc_group_prof_enable(afqmc_group* g, int32_t on) {
  if (!g) return afqmc_set_error(AFQMC_ERR_ARG, "null group");
  return prof_enable(&g->prof, on);
}

extern "C" int afqmc_group_prof_phases(afqmc_group* g, uint32_t mask) {
  if (!g) return afqmc_set_error(AFQMC_ERR_ARG, "null group");
  g->prof.mask = mask;
  return AFQMC_OK;
}

extern "C" int afqmc_group_prof_collect(afqmc_group* g, double* ms, int64_t* counts) {
  if (!g || !ms || !counts) return afqmc_set_error(AFQMC_ERR_ARG, "null argument");
  return prof_collect(&g->prof, ms, counts);
}

// Times both GEMM tile edges (64, 128) for every grouped GEMM of the current member set --
// the step's force-bias and VHS products and the energy's float64 c_g product -- and keeps
// the faster edge for each.  Runs the GEMMs on the bound buffers: their outputs (force
// bias, VHS) are workspace that the next step or energy call rewrites before reading, and
// the float64 Green's functions and rotated vectors the c_g product reads are allocated
// with the context.  The k order of every output is the same for both edges, so the choice
// does not change results.  Synchronises.
extern "C" int afqmc_group_tune(afqmc_group* g, double* ms_out) {
  if (!g) return afqmc_set_error(AFQMC_ERR_ARG, "null group");
  if (g->mem.empty()) return AFQMC_OK;
  int rc = set_group_device(g);
  if (rc) return rc;
  hipEvent_t e0, e1;
  HIP_TRY(hipEventCreate(&e0));
  HIP_TRY(hipEventCreate(&e1));
  float best[3][2];  // [fb / vhs / c_g][64 / 128]
  auto run = [&](int which) {
    if (which == 2)
      return launch_gemm_f64_grouped(g->stream, g->fb64.k0, g->fb64_vec2,
                                     targs<GemmArgs>(g, g->fb64), tstart(g, g->fb64),
                                     g->fb64.nprob, g->fb64.nblocks);
    return run_gemm(g, which ? g->vhs : g->fb, which);
  };
  for (int bi = 0; bi < 2; ++bi) {
    g->fb_bm = g->vhs_bm = g->cg_bm = bi ? 128 : 64;
    if ((rc = build(g))) break;
    for (int which = 0; which < 3 && !rc; ++which) {
      rc = run(which);  // warm-up
      if (rc) break;
      (void)hipEventRecord(e0, g->stream);
      for (int r = 0; r < 3 && !rc; ++r) rc = run(which);
      (void)hipEventRecord(e1, g->stream);
      (void)hipEventSynchronize(e1);
      float ms = 0.f;
      (void)hipEventElapsedTime(&ms, e0, e1);
      best[which][bi] = ms / 3.0f;
    }
    if (rc) break;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc) return rc;
  g->fb_bm = best[0][1] < best[0][0] ? 128 : 64;
  g->vhs_bm = best[1][1] < best[1][0] ? 128 : 64;
  g->cg_bm = best[2][1] < best[2][0] ? 128 : 64;
  if (ms_out)
    for (int w = 0; w < 3; ++w) {
      ms_out[2 * w] = best[w][0];
      ms_out[2 * w + 1] = best[w][1];
    }
  g->dirty = true;
  return AFQMC_OK;
}
