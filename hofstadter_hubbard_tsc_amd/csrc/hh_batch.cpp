// Batched H.psi and batched Lanczos over independent superblocks (different
// flux, doping or couplings), the reference's model_wave
// (code/mat_vec.f90:3476-3620) and lanczos_memo_new (code/lanczos.f90:4-176)
// applied to every member at once: one grouped launch per H.psi stage for all
// members, the members' vector work in segmented launches, one host round
// trip for all members' scalars.
#include <algorithm>
#include <cmath>
#include <memory>
#include <vector>

#include "hh_runtime.hpp"

using namespace hh;

namespace {

constexpr int BATCH_MAX = 32;   // 8 buffer slots per member in a uint8 slot index

// member i's task list appended to `dst` with its buffer slots moved to 8*i+
void append_tasks(TaskList& dst, const TaskList& src, int member) {
  const int t0 = (int)dst.terms.size();
  for (Term tm : src.terms) {
    tm.a_buf = (uint8_t)(8 * member + tm.a_buf);
    tm.b_buf = (uint8_t)(8 * member + tm.b_buf);
    dst.terms.push_back(tm);
  }
  for (Tile tl : src.tiles) {
    tl.t0 += t0;
    tl.t1 += t0;
    tl.c_buf = 8 * member + tl.c_buf;
    dst.tiles.push_back(tl);
  }
  dst.flops_alg += src.flops_alg;
  dst.flops_exec += src.flops_exec;
}

void count_matvec(hh_ctx* ctx, const hh_batch* b) {
  ctx->stats.matvecs += 1;
  ctx->stats.flops_alg += b->flops_alg;
  ctx->stats.stage1_flops_alg += b->stage1_flops;
  ctx->stats.stage2_flops_alg += b->stage2_flops;
  ctx->stats.flops_exec += b->flops_exec;
  ctx->stats.stage1_flops_exec += b->stage1_exec;
  ctx->stats.stage2_flops_exec += b->stage2_exec;
}

// Segment layout, tables and fixed buffers of the batched Lanczos, made on
// the first call: member i's elements at off[i] (a multiple of vec_span()),
// the H.psi input Q and output W, and the buffer table of the grouped H.psi
// Q -> W.  The table is written again on every call, so that it names the
// members' current blocks and intermediates whenever the graph replays.
void lz_layout(hh_ctx* ctx, hh_batch* b) {
  const int nm = (int)b->sbs.size();
  if (b->off.empty()) {
    const int64_t span = vec_span();
    std::vector<int32_t> bseg, sfirst;
    std::vector<int64_t> send;
    b->off.resize(nm);
    int64_t o = 0;
    for (int i = 0; i < nm; ++i) {
      const int64_t n = b->sbs[i]->plan.sb.dim;
      need(n > 0, "empty superblock in batch");
      const int64_t nblk = (n + span - 1) / span;
      need((int64_t)bseg.size() + nblk < (int64_t)1 << 31, "batch too large for the segmented vector kernels");
      b->off[i] = o;
      sfirst.push_back((int32_t)bseg.size());
      for (int64_t k = 0; k < nblk; ++k) bseg.push_back(i);
      send.push_back(o + n);
      o += nblk * span;
    }
    sfirst.push_back((int32_t)bseg.size());
    b->npad = o;
    b->seg_bseg.alloc(sizeof(int32_t) * bseg.size());
    b->seg_send.alloc(sizeof(int64_t) * send.size());
    b->seg_sfirst.alloc(sizeof(int32_t) * sfirst.size());
    check_hip(hipMemcpyAsync(b->seg_bseg.p, bseg.data(), sizeof(int32_t) * bseg.size(), hipMemcpyHostToDevice,
                             ctx->stream), "h2d");
    check_hip(hipMemcpyAsync(b->seg_send.p, send.data(), sizeof(int64_t) * send.size(), hipMemcpyHostToDevice,
                             ctx->stream), "h2d");
    check_hip(hipMemcpyAsync(b->seg_sfirst.p, sfirst.data(), sizeof(int32_t) * sfirst.size(),
                             hipMemcpyHostToDevice, ctx->stream), "h2d");
    b->seg.nseg = nm;
    b->seg.nb = (int)bseg.size();
    b->seg.bseg = (const int32_t*)b->seg_bseg.p;
    b->seg.send = (const int64_t*)b->seg_send.p;
    b->seg.sfirst = (const int32_t*)b->seg_sfirst.p;
    b->Q.alloc(sizeof(double2) * (size_t)b->npad);
    b->W.alloc(sizeof(double2) * (size_t)b->npad);
    b->lz_table.alloc(sizeof(double2*) * 8 * nm);
    b->active.alloc(sizeof(int32_t) * nm);
    b->seg.active = (const int32_t*)b->active.p;
    check_hip(hipStreamSynchronize(ctx->stream), "batch layout");
  }
  std::vector<double2*> tab(8 * nm);
  for (int i = 0; i < nm; ++i) {
    const hh_superblock* sb = b->sbs[i];
    const BufPtrs p = bufs_of(b->Q.c() + b->off[i], b->W.c() + b->off[i], sb->sys, sb->env, sb->X.c(),
                              sb->merged.c(), nullptr, nullptr);
    for (int k = 0; k < 8; ++k) tab[8 * i + k] = p.p[k];
  }
  check_hip(hipMemcpyAsync(b->lz_table.p, tab.data(), sizeof(double2*) * tab.size(), hipMemcpyHostToDevice,
                           ctx->stream), "h2d table");
  check_hip(hipStreamSynchronize(ctx->stream), "batch table");
}

// The grouped FP64 H.psi W = H Q of every member still running (the tiles of
// a member whose flag in `active` is 0 are skipped): the two stages'
// launches, replayed from a HIP graph captured on first use (the graph reads
// its buffer pointers and the flags from device memory when it runs), or
// launched one by one.
void lz_hpsi(hh_ctx* ctx, hh_batch* b) {
  const BufTable t{(double2* const*)b->lz_table.p, (const int32_t*)b->active.p};
  hipStream_t s = ctx->stream;
  if (!b->use_graph) {
    launch_tiles(b->st1, t, s, PREC_FP64, 1, nullptr);
    launch_tiles(b->st2, t, s, PREC_FP64, 2, nullptr);
    return;
  }
  if (!b->graph_exec) {
    check_hip(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), "begin capture");
    try {
      launch_tiles(b->st1, t, s, PREC_FP64, 1, nullptr);
      launch_tiles(b->st2, t, s, PREC_FP64, 2, nullptr);
    } catch (...) {
      hipGraph_t g = nullptr;
      (void)hipStreamEndCapture(s, &g);
      if (g) (void)hipGraphDestroy(g);
      throw;
    }
    check_hip(hipStreamEndCapture(s, &b->graph), "end capture");
    check_hip(hipGraphInstantiate(&b->graph_exec, b->graph, nullptr, nullptr, 0), "graph instantiate");
  }
  check_hip(hipGraphLaunch(b->graph_exec, s), "graph launch");
}

// Per-member state of the lock-step Lanczos, following lanczos() of the
// single-superblock path step for step.
struct Member {
  std::vector<double> alpha, beta;
  double ritz = 0.0, b_last = 0.0, ritz_change = 0.0;
  int last = 1;
  bool done = false, converged = false;
};

}  // namespace

extern "C" {

int hh_batch_create(hh_ctx* ctx, int nsb, hh_superblock* const* sbs, hh_batch** out) {
  if (!out) return 1;
  *out = nullptr;
  return guard(ctx, [&] {
    need(nsb >= 1 && nsb <= BATCH_MAX, "a batch holds 1 to 32 superblocks");
    need(sbs != nullptr, "null superblock list");
    auto b = std::make_unique<hh_batch>();
    TaskList s1, s2;
    for (int i = 0; i < nsb; ++i) {
      need(sbs[i] != nullptr, "null superblock in batch");
      b->sbs.push_back(sbs[i]);
      append_tasks(s1, sbs[i]->plan.stage1, i);
      append_tasks(s2, sbs[i]->plan.stage2, i);
      b->total_dim += sbs[i]->plan.sb.dim;
    }
    const LaunchConfig cfg = ctx->forced_cfg ? ctx->forced : LaunchConfig();
    b->st1.upload(s1, ctx->stream, nullptr, 0, cfg);
    b->st2.upload(s2, ctx->stream, nullptr, 0, cfg);
    b->stage1_flops = s1.flops_alg;
    b->stage2_flops = s2.flops_alg;
    b->flops_alg = s1.flops_alg + s2.flops_alg;
    b->flops_exec = s1.flops_exec + s2.flops_exec;
    b->stage1_exec = s1.flops_exec;
    b->stage2_exec = s2.flops_exec;
    b->table.alloc(sizeof(double2*) * 8 * nsb);
    b->host_table.assign(8 * nsb, nullptr);
    if (ctx->autotune && !ctx->forced_cfg) autotune_batch(ctx, b.get(), s1, s2);
    *out = b.release();
  });
}

int hh_batch_destroy(hh_batch* b) {
  delete b;
  return 0;
}

int hh_batch_info(const hh_batch* b, int64_t* total_dim, double* flops_alg) {
  if (!b) return 1;
  if (total_dim) *total_dim = b->total_dim;
  if (flops_alg) *flops_alg = b->flops_alg;
  return 0;
}

int hh_batch_set_graph(hh_batch* b, int on) {
  if (!b) return 1;
  b->use_graph = on != 0;
  return 0;
}

int hh_batch_model_wave(hh_ctx* ctx, hh_batch* b, const double* const* psi_in, double* const* psi_out) {
  return guard(ctx, [&] {
    need(b && psi_in && psi_out, "null argument");
    need(ctx->precision != PREC_FP32_STORE, "the batched H.psi runs in precision 0 or 1");
    const int n = (int)b->sbs.size();
    for (int i = 0; i < n; ++i) {
      need(psi_in[i] && psi_out[i] && psi_in[i] != psi_out[i], "batch needs distinct input and output vectors");
      const BufPtrs p = bufs_of((double2*)psi_in[i], (double2*)psi_out[i], b->sbs[i]->sys, b->sbs[i]->env,
                                b->sbs[i]->X.c(), b->sbs[i]->merged.c(), nullptr, nullptr);
      for (int k = 0; k < 8; ++k) b->host_table[8 * i + k] = p.p[k];
    }
    check_hip(hipMemcpyAsync(b->table.p, b->host_table.data(), sizeof(double2*) * b->host_table.size(),
                             hipMemcpyHostToDevice, ctx->stream),
              "h2d table");
    const BufTable t{(double2* const*)b->table.p};
    const double t0 = ctx->profile ? wall() : 0.0;
    if (ctx->profile) check_hip(hipEventRecord(ctx->ev[0], ctx->stream), "event");
    launch_tiles(b->st1, t, ctx->stream, ctx->precision, 1, ctx->profile ? ctx->ev[1] : nullptr);
    if (ctx->profile) check_hip(hipEventRecord(ctx->ev[2], ctx->stream), "event");
    launch_tiles(b->st2, t, ctx->stream, ctx->precision, 2, ctx->profile ? ctx->ev[3] : nullptr);
    if (ctx->profile) {
      check_hip(hipEventRecord(ctx->ev[4], ctx->stream), "event");
      check_hip(hipEventSynchronize(ctx->ev[4]), "event sync");
      float ms[4] = {0.f, 0.f, 0.f, 0.f};
      for (int i = 0; i < 4; ++i) check_hip(hipEventElapsedTime(&ms[i], ctx->ev[i], ctx->ev[i + 1]), "elapsed");
      ctx->stats.stage1_ms += ms[0];
      ctx->stats.reduce_ms += ms[1] + ms[3];
      ctx->stats.stage2_ms += ms[2];
      ctx->stats.matvec_ms += (wall() - t0) * 1e3;
    }
    count_matvec(ctx, b);
    check_hip(hipStreamSynchronize(ctx->stream), "batch model_wave");
  });
}

int hh_batch_lanczos(hh_ctx* ctx, hh_batch* b, double* const* psi, double lanzero, int max_iter, double* e0,
                     int32_t* iters, int32_t* converged) {
  return guard(ctx, [&] {
    need(b && psi && e0 && iters, "null argument");
    need(max_iter >= 1, "max_iter must be >= 1");
    const int nm = (int)b->sbs.size();
    for (int i = 0; i < nm; ++i) need(psi[i] != nullptr, "null start vector");
    lz_layout(ctx, b);
    hipStream_t s = ctx->stream;
    const int64_t np = b->npad;
    const SegDev& g = b->seg;
    b->V.alloc(sizeof(double2) * (size_t)np * (size_t)max_iter);
    const int ldc = max_iter + 1;
    // device scalars: alpha, |w|^2, beta, sw, cs, |w'|^2, factor (nm each),
    // then the Gram-Schmidt / Ritz coefficients (nm x ldc)
    b->scal.alloc(sizeof(double) * (size_t)nm * (7 + ldc));
    b->partial.alloc(sizeof(double) * (size_t)g.nb * (size_t)std::max(max_iter, 1));
    double* sc = (double*)b->scal.p;
    double *d_alpha = sc, *d_wsq = sc + nm, *d_beta = sc + 2 * nm, *d_sw = sc + 3 * nm, *d_cs = sc + 4 * nm;
    double *d_nw = sc + 5 * nm, *d_fac = sc + 6 * nm, *d_c = sc + 7 * nm;
    double* part = (double*)b->partial.p;
    double2* V = b->V.c();
    double2* Q = b->Q.c();
    double2* W = b->W.c();
    auto vec = [&](int j) { return V + (size_t)j * np; };
    auto fetch = [&](const double* src, int count) {
      std::vector<double> h(count);
      check_hip(hipMemcpyAsync(h.data(), src, sizeof(double) * count, hipMemcpyDeviceToHost, s), "d2h");
      check_hip(hipStreamSynchronize(s), "sync");
      return h;
    };
    auto put = [&](double* dst, const std::vector<double>& h) {
      check_hip(hipMemcpyAsync(dst, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice, s), "h2d");
    };
    auto finite = [](double x) {
      if (!std::isfinite(x)) throw HHError(-4, "non-finite value in batched Lanczos: H.psi produced NaN or Inf");
      return x;
    };
    // with profiling on, matvec_ms gains the device time of each grouped H.psi
    auto hpsi = [&]() {
      if (ctx->profile) check_hip(hipEventRecord(ctx->ev[0], s), "event");
      lz_hpsi(ctx, b);
      if (ctx->profile) {
        check_hip(hipEventRecord(ctx->ev[4], s), "event");
        check_hip(hipEventSynchronize(ctx->ev[4]), "event sync");
        float ms = 0.f;
        check_hip(hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[4]), "elapsed");
        ctx->stats.matvec_ms += ms;
      }
      count_matvec(ctx, b);
    };
    // running members; a stopped member's H.psi tiles, multi-dots and
    // updates are skipped from the next launch on (its vectors are zero).
    // Every member runs at the start.
    std::vector<Member> m(nm);
    std::vector<int32_t> run(nm, 1);
    auto put_run = [&]() {
      for (int i = 0; i < nm; ++i) run[i] = m[i].done ? 0 : 1;
      check_hip(hipMemcpyAsync(b->active.p, run.data(), sizeof(int32_t) * nm, hipMemcpyHostToDevice, s), "h2d");
    };
    put_run();
    // v1 = psi / |psi| per member, also into Q
    for (int i = 0; i < nm; ++i)
      check_hip(hipMemcpyAsync(vec(0) + b->off[i], psi[i], sizeof(double2) * (size_t)b->sbs[i]->plan.sb.dim,
                               hipMemcpyDeviceToDevice, s), "d2d");
    dev_seg_mdots(vec(0), 0, 1, vec(0), g, d_nw, part, s);
    std::vector<double> fac(nm);
    {
      const std::vector<double> n2 = fetch(d_nw, nm);
      for (int i = 0; i < nm; ++i) {
        const double nr = std::sqrt(n2[i]);
        need(std::isfinite(nr) && nr > 0.0, "zero or non-finite start vector");
        fac[i] = 1.0 / nr;
      }
    }
    put(d_fac, fac);
    dev_seg_scale2(vec(0), Q, vec(0), d_fac, g, s);
    std::vector<double> zeros(nm, 0.0);
    put(d_beta, zeros);
    // alpha_j = Re<v_j, H v_j> and the three-term update (with |w|^2) for the
    // members still running (the others carry zero vectors)
    auto alpha_and_update = [&](int j) {
      dev_seg_mdots(vec(j - 1), 0, 1, W, g, d_alpha, part, s);
      dev_seg_lupd_norm(W, vec(j - 1), d_alpha, j >= 2 ? vec(j - 2) : nullptr, d_beta, g, d_wsq, part, s);
      std::vector<double> h(2 * nm);
      check_hip(hipMemcpyAsync(h.data(), d_alpha, sizeof(double) * 2 * nm, hipMemcpyDeviceToHost, s), "d2h");
      check_hip(hipStreamSynchronize(s), "sync");
      return h;
    };
    hpsi();
    std::vector<double> r = alpha_and_update(1);
    for (int i = 0; i < nm; ++i) {
      m[i].alpha.push_back(finite(r[i]));
      m[i].ritz = m[i].alpha[0];
      m[i].b_last = std::sqrt(std::max(r[nm + i], 0.0));
    }
    int L = 1;   // stored Lanczos vectors (the running members' count)
    while (L < max_iter) {
      // beta, the stopping test on it, and the Gram-Schmidt scalars
      std::vector<double> sw(nm, 0.0), cs(nm, 0.0);
      for (int i = 0; i < nm; ++i) {
        Member& x = m[i];
        if (x.done) continue;
        const double bb = std::sqrt(finite(r[nm + i]));
        x.beta.push_back(bb);
        x.b_last = bb;
        if (std::fabs(bb) <= lanzero) {
          x.converged = x.done = true;
          continue;
        }
        sw[i] = 1.0 / bb;
        cs[i] = -1.0 / bb;
      }
      if (std::all_of(m.begin(), m.end(), [](const Member& x) { return x.done; })) break;
      put_run();
      // classical Gram-Schmidt against every stored vector (real
      // coefficients), w <- (w - sum_j Re<v_j, w> v_j) / beta; a finished
      // member's w becomes zero
      put(d_sw, sw);
      put(d_cs, cs);
      dev_seg_mdots(V, np, L, W, g, d_c, part, s);
      dev_seg_combine_norm(W, d_sw, V, np, L, d_c, L, d_cs, g, d_nw, part, s);
      const std::vector<double> nw2 = fetch(d_nw, nm);
      for (int i = 0; i < nm; ++i) {
        const double nw = std::sqrt(std::max(nw2[i], 0.0));
        fac[i] = (!m[i].done && nw > 0.0) ? 1.0 / nw : 0.0;
      }
      put(d_fac, fac);
      dev_seg_scale2(vec(L), Q, W, d_fac, g, s);
      ++L;
      std::vector<double> bprev(nm, 0.0);
      for (int i = 0; i < nm; ++i)
        if (!m[i].done) {
          ++m[i].last;
          bprev[i] = m[i].beta[m[i].last - 2];
        }
      put(d_beta, bprev);
      hpsi();
      r = alpha_and_update(L);
      for (int i = 0; i < nm; ++i) {
        Member& x = m[i];
        if (x.done) continue;
        x.alpha.push_back(finite(r[i]));
        x.b_last = std::sqrt(std::max(r[nm + i], 0.0));
        std::vector<double> d(x.alpha), e(x.beta);
        e.resize(d.size());
        tridiag_ql(d, e, nullptr);
        x.ritz_change = std::fabs(d[0] - x.ritz);
        x.ritz = d[0];
        if (x.ritz_change <= lanzero) x.converged = true;
        if (x.converged || x.last == max_iter) x.done = true;
      }
      if (std::all_of(m.begin(), m.end(), [](const Member& x) { return x.done; })) break;
    }
    // Ritz vectors from each member's own Lanczos vectors (coefficients past
    // a member's count are zero), normalised, copied out
    std::vector<double> coef((size_t)nm * ldc, 0.0), ones(nm, 1.0);
    std::vector<double> resid(nm);
    for (int i = 0; i < nm; ++i) {
      Member& x = m[i];
      std::vector<double> d(x.alpha), e(x.beta), z;
      e.resize(d.size());
      tridiag_ql(d, e, &z);
      for (int k = 0; k < x.last; ++k) coef[(size_t)i * ldc + k] = z[k];
      e0[i] = d[0];
      iters[i] = x.last;
      if (converged) converged[i] = x.converged ? 1 : 0;
      resid[i] = x.b_last * std::fabs(z[x.last - 1]);
    }
    put(d_c, coef);
    put(d_sw, zeros);
    put(d_cs, ones);
    dev_seg_combine_norm(W, d_sw, V, np, L, d_c, ldc, d_cs, g, d_nw, part, s);
    const std::vector<double> nw2 = fetch(d_nw, nm);
    for (int i = 0; i < nm; ++i) {
      need(std::isfinite(nw2[i]) && nw2[i] > 0.0, "batched Lanczos produced a zero Ritz vector");
      fac[i] = 1.0 / std::sqrt(nw2[i]);
    }
    put(d_fac, fac);
    dev_seg_scale2(W, nullptr, W, d_fac, g, s);
    for (int i = 0; i < nm; ++i)
      check_hip(hipMemcpyAsync(psi[i], W + b->off[i], sizeof(double2) * (size_t)b->sbs[i]->plan.sb.dim,
                               hipMemcpyDeviceToDevice, s), "d2d");
    check_hip(hipStreamSynchronize(s), "batch lanczos");
    ctx->last_converged = std::all_of(m.begin(), m.end(), [](const Member& x) { return x.converged; }) ? 1 : 0;
    ctx->last_residual = *std::max_element(resid.begin(), resid.end());
  });
}

}  // extern "C"
