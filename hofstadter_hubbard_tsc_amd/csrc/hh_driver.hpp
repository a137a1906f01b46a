// The sweep driver shared by the runtime (hh_runtime.cpp: steps, checkpoint,
// C ABI) and the measurement chains (hh_measure.cpp): all blocks,
// truncations and the carried state resident on the device, one step at a
// time (warm_up_point / left_to_right / right_to_left, code/warm_up.f90:
// 1284-1849).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "../../include/hhdmrg.h"
#include "hh_core.hpp"
#include "hh_engine.hpp"
#include "hh_runtime.hpp"
#include "hh_solver.hpp"

namespace hh {

// the step's pieces (hh_runtime.cpp)
hh_trunc* truncate(hh_ctx* ctx, hh_superblock* sb, int side, const double2* psi, int kept_min, int kept_max,
                   double trun_err, const double2* psi2 = nullptr);
hh_trunc* trunc_identity(hh_ctx* ctx, hh_superblock* sb, int side);
hh_block* grow(hh_ctx* ctx, hh_superblock* sb, int side, const hh_trunc* t);
void wave_trans(hh_ctx* ctx, int dir, hh_superblock* osb, const double2* psi_old, const hh_trunc* st,
                const hh_trunc* et, hh_superblock* nsb, double2* psi_new);
hh_superblock* bind(hh_ctx* ctx, const hh_block* sys, const hh_block* env, int N);
hh_block* site_block(hh_ctx* ctx, int side);
// one FP64 task list on the context's stream, waited for
void launch_on(hh_ctx* ctx, const TaskList& tl, const BufPtrs& b);

// a block of `length` sites with sector table `sec` and its operator arena
hh_block* new_block(hh_ctx* ctx, int side, int length, const std::vector<Sector>& sec, const std::vector<int>& sites);
// a sweep schedule the driver can run (from the caller or from a checkpoint)
void check_schedule(const hh_schedule& s);

struct Driver;
// the driver's whole state to / from one binary file (hh_checkpoint.cpp)
void save_driver(Driver& d, const char* path);
void load_driver(Driver& d, const char* path);

// ----------------------------------------------------------------------------
// sweep driver: Get_Ground_State (code/quantum.f90:179-288)
// ----------------------------------------------------------------------------
struct Driver {
  hh_ctx* ctx;
  hh_schedule S;
  std::map<int, std::unique_ptr<hh_block>> sysb, envb;
  std::map<int, std::unique_ptr<hh_trunc>> syst, envt;
  std::vector<hh_step_record> recs;
  DevMem psi, psi2;
  int64_t psi_dim = 0;
  uint64_t step = 0;
  // levels > 1 (lanczos_memo_new with levels, Get_density_matrix_*_new): the
  // carried vectors of the levels above the first and their transform
  // scratch, the current step's level energies and every record's
  int levels = 1;
  bool level_carry = false;
  std::vector<std::unique_ptr<DevMem>> lpsi;
  int64_t lpsi_dim = 0;
  int solved_levels = 1;
  std::vector<double> cur_e;
  std::vector<std::vector<double>> lev_e;
  // previous step (for the wave transformation)
  std::unique_ptr<hh_superblock> prev_sb;

  static constexpr double TRUN_ERR = 1.0000000180025095e-18;   // float32(1e-18) widened

  // the stored block or truncation of `len` sites (a windowed fork holds
  // only some of them)
  template <typename T>
  static T* held(const std::map<int, std::unique_ptr<T>>& m, int len, const char* what) {
    auto it = m.find(len);
    if (it == m.end())
      throw HHError(1, std::string("this driver holds no ") + what + " of " + std::to_string(len) +
                           " sites (outside the window of a windowed fork, or not built yet)");
    return it->second.get();
  }

  int kmin(int kmax) const { return std::min(3 * kmax / 4, 3600); }

  // wall time per phase of the current step: bind, wave, lanczos, trunc, grow
  double ph[5] = {0, 0, 0, 0, 0};
  double ph_t = 0.0;
  void lap_start() {
    for (double& x : ph) x = 0.0;
    ph_t = wall();
  }
  void lap(int i) {
    check_hip(hipStreamSynchronize(ctx->stream), "phase sync");
    const double t = wall();
    ph[i] += t - ph_t;
    ph_t = t;
  }

  // learned preconditioner: network, captured training triples, mode
  std::unique_ptr<PcNet> net;
  std::vector<std::unique_ptr<PcSample>> samples;
  int capture_left = 0;
  int pc_mode = PC_LEARNED;
  double fall_factor = DAV_FALL_FACTOR;
  int64_t min_dim = DAV_MIN_DIM;
  int dav_kmax = DAV_KMAX;
  DevMem cap_u, cap_r;

  // the step's ground-state solve: the plain form (Lanczos with FP64 H.psi),
  // the mixed-precision Lanczos rounds, or the chosen Davidson form
  SolveOut solve(hh_superblock* sb, double2* p, double lanzero) {
    const int solver = sb->plan.sb.dim < min_dim ? 0 : S.solver;
    if (solver == 2) {
      bool fb = false;
      const bool have_net = net && net->trained;
      const int mode = (pc_mode == PC_LEARNED && !have_net) ? PC_DIAGONAL : pc_mode;
      const float* w = mode == PC_LEARNED ? (const float*)net->w.p : nullptr;
      if (capture_left <= 0) {
        SolveOut so = davidson(ctx, sb, p, lanzero, S.max_iter, mode, w, &fb, AnchorHook(), fall_factor, dav_kmax);
        so.fallback = fb;
        return so;
      }
      // record the restart points of this solve (unit vector, exact residual,
      // Rayleigh quotient); the last CAP_PER_SOLVE + 1 are kept in a ring
      const int64_t n = sb->plan.sb.dim;
      constexpr int CAP_PER_SOLVE = 3;
      struct Anc {
        DevMem u, r;
        double theta = 0.0;
      };
      std::vector<Anc> ring(CAP_PER_SOLVE + 1);
      int count = 0;
      AnchorHook hook = [&](const double2* u, const double2* r, double th) {
        Anc& a = ring[count % ring.size()];
        a.u.alloc(sizeof(double2) * (size_t)n);
        a.r.alloc(sizeof(double2) * (size_t)n);
        dev_scale_copy(a.u.c(), u, 1.0, n, ctx->stream);
        dev_scale_copy(a.r.c(), r, 1.0, n, ctx->stream);
        a.theta = th;
        ++count;
      };
      SolveOut so = davidson(ctx, sb, p, lanzero, S.max_iter, mode, w, &fb, hook, fall_factor, dav_kmax);
      so.fallback = fb;
      // the last restart of a converged solve is the returned vector itself
      // (its correction is zero): the ones before it are the triples
      const int last = (so.converged && !fb) ? count - 1 : count;
      const int first = std::max(0, last - CAP_PER_SOLVE);
      for (int i = first; i < last && capture_left > 0; ++i) {
        Anc& a = ring[i % ring.size()];
        samples.push_back(make_sample(ctx, sb, a.u.c(), a.r.c(), a.theta, p));
        --capture_left;
      }
      return so;
    }
    if (solver == 1) return mixed_lanczos(ctx, sb, p, lanzero, S.max_iter, 4);
    if (capture_left <= 0) return lanczos(ctx, sb, p, lanzero, S.max_iter);
    // plain solve that also records a training triple of the preconditioner:
    // the unit start vector, its FP64 residual and Rayleigh quotient
    const int64_t n = sb->plan.sb.dim;
    double theta0 = 0.0;
    LanczosOpts o;
    o.on_first = [&](const double2* v1, const double2* r1, double a1) {
      cap_u.alloc(sizeof(double2) * (size_t)n);
      cap_r.alloc(sizeof(double2) * (size_t)n);
      dev_scale_copy(cap_u.c(), v1, 1.0, n, ctx->stream);
      dev_scale_copy(cap_r.c(), r1, 1.0, n, ctx->stream);
      theta0 = a1;
    };
    SolveOut so = lanczos(ctx, sb, p, lanzero, S.max_iter, o);
    samples.push_back(make_sample(ctx, sb, cap_u.c(), cap_r.c(), theta0, p));
    --capture_left;
    return so;
  }

  // the start vectors of the levels above the first: seeded afresh at every
  // step (wave_initialize of levels 2.. after the wave transformation,
  // code/warm_up.f90:1616-1620), or the vectors already held when the step
  // carries the state of the same superblock.  With level_carry the previous
  // step's upper levels go through the same wave transformation as level 1
  // (dir 0 LR / 1 RL with its truncations ta, tb) instead of being seeded.
  void level_starts(int64_t n, bool carry, hh_superblock* sb = nullptr, int dir = 0, const hh_trunc* ta = nullptr,
                    const hh_trunc* tb = nullptr) {
    const bool held = carry && lpsi_dim == n && (int)lpsi.size() == levels - 1;
    if (level_carry && sb && ta && prev_sb && lpsi_dim == prev_sb->plan.sb.dim && (int)lpsi.size() == levels - 1) {
      // psi2 is free here: it holds the previous step's level 1
      for (auto& v : lpsi) {
        wave_trans(ctx, dir, prev_sb.get(), v->c(), ta, tb, sb, psi_buf(n, psi2));
        std::swap(v->p, psi2.p);
        std::swap(v->bytes, psi2.bytes);
      }
      lpsi_dim = n;
      return;
    }
    lpsi.resize(levels - 1);
    if (held) return;
    for (int l = 1; l < levels; ++l) {
      if (!lpsi[l - 1]) lpsi[l - 1].reset(new DevMem());
      dev_start_vector(psi_buf(n, *lpsi[l - 1]), n, S.seed + step + 7919ull * (uint64_t)l, ctx->stream);
    }
  }

  // every level's solve (levels > 1): level 1 into p, level l into lpsi[l-1],
  // each projected off the levels below; the first level's outcome is the
  // step's (energy, iterations, convergence, residual) with the H.psi count
  // and time of all levels, all energies go to cur_e
  SolveOut solve_levels(hh_superblock* sb, double2* p, double lanzero) {
    need(S.solver == 0, "several levels are solved by the plain Lanczos only (solver 0)");
    const int64_t n = sb->plan.sb.dim;
    // a level needs a start vector outside the 2 l real directions of the
    // levels below it, so a superblock of dimension n holds at most n levels;
    // the rest report NaN
    const int nl = (int)std::min<int64_t>(levels, n);
    DevMem defl;
    defl.alloc(sizeof(double2) * (size_t)n * (size_t)(2 * (levels - 1)));
    cur_e.assign(levels, std::nan(""));
    SolveOut first;
    for (int l = 0; l < nl; ++l) {
      double2* v = l == 0 ? p : lpsi[l - 1]->c();
      LanczosOpts opt;
      opt.deflate = defl.c();
      opt.ndeflate = 2 * l;
      const SolveOut so = lanczos(ctx, sb, v, lanzero, S.max_iter, opt);
      if (l == 0) {
        first = so;
      } else {
        // the step's H.psi count and time cover every level
        first.mv_seconds += so.mv_seconds;
        first.hpsi_fp64 += so.hpsi_fp64;
        first.hpsi_fp32 += so.hpsi_fp32;
      }
      cur_e[l] = so.e0;
      if (l + 1 < nl) {
        // this level and i times it join the deflation set (v is unit)
        dev_scale_copy(defl.c() + (size_t)(2 * l) * n, v, 1.0, n, ctx->stream);
        dev_cscale_copy(defl.c() + (size_t)(2 * l + 1) * n, v, 0.0, 1.0, n, ctx->stream);
      }
    }
    lpsi_dim = n;
    solved_levels = nl;
    return first;
  }
  SolveOut solve_step(hh_superblock* sb, double2* p, double lanzero) {
    if (levels == 1) return solve(sb, p, lanzero);
    return solve_levels(sb, p, lanzero);
  }
  // the second level's vector for the mixed density matrix (null with one level)
  const double2* level2() { return levels > 1 && solved_levels > 1 ? lpsi[0]->c() : nullptr; }

  void record(int phase, hh_superblock* sb, const SolveOut& so, const hh_trunc* t, double secs, int64_t mv0,
              double fl0) {
    const double E = so.e0, mvs = so.mv_seconds;
    hh_step_record r{};
    r.phase = phase;
    r.sys_len = sb->sys->d.length;
    r.env_len = sb->env->d.length;
    r.iters = so.iters;
    r.converged = so.converged ? 1 : 0;
    r.ritz_change = so.ritz_change;
    r.residual = so.residual;
    r.hpsi_fp64 = so.hpsi_fp64;
    r.hpsi_fp32 = so.hpsi_fp32;
    r.solver = S.solver;
    r.fallback = so.fallback ? 1 : 0;
    r.num_state = t ? t->num_state : 0;
    r.dim = sb->plan.sb.dim;
    r.energy = E;
    r.trunc_err = t ? t->err : 0.0;
    r.entropy = t ? t->entropy : 0.0;
    r.seconds = secs;
    r.matvec_seconds = mvs;
    r.matvec_flops = ctx->stats.flops_alg - fl0;
    r.bind_seconds = ph[0];
    r.wave_seconds = ph[1];
    r.lanczos_seconds = ph[2];
    r.trunc_seconds = ph[3];
    r.grow_seconds = ph[4];
    (void)mv0;
    recs.push_back(r);
    lev_e.push_back(levels > 1 ? cur_e : std::vector<double>{E});
    if (verbose()) {
      std::fprintf(stderr,
                   "# step phase=%d sys=%d env=%d dim=%lld iters=%d%s%s kept=%d E=%.12f res=%.1e %.2fs "
                   "(bind %.3f wave %.3f lanczos %.3f [H.psi %.3f] trunc %.3f grow %.3f)\n",
                   r.phase, r.sys_len, r.env_len, (long long)r.dim, r.iters, r.converged ? "" : "(cap)",
                   r.fallback ? "(fallback)" : "", r.num_state,
                   r.energy, r.residual, r.seconds,
                   r.bind_seconds, r.wave_seconds, r.lanczos_seconds, r.matvec_seconds, r.trunc_seconds,
                   r.grow_seconds);
      std::fflush(stderr);
    }
  }

  double2* psi_buf(int64_t n, DevMem& m) {
    m.alloc(sizeof(double2) * (size_t)std::max<int64_t>(n, 1));
    return m.c();
  }

  void warm_up() {
    const ModelParams& m = ctx->model;
    const int kmax = S.kept_max_warm > 0 ? S.kept_max_warm : S.kept_max;
    const double lanzero = S.tol * S.warm_tol_factor;
    const double terr = TRUN_ERR * 100;
    sysb[1].reset(site_block(ctx, 0));
    envb[1].reset(site_block(ctx, 1));
    int label = 1;
    while (label + 1 <= m.nsite() / 2) {
      const double t0 = wall();
      const int64_t mv0 = ctx->stats.matvecs;
      const double fl0 = ctx->stats.flops_alg;
      const int N = electrons_for_length(m, 2 * label + 2);
      lap_start();
      std::unique_ptr<hh_superblock> sb(bind(ctx, sysb[label].get(), envb[label].get(), N));
      lap(0);
      const int64_t n = sb->plan.sb.dim;
      double2* p = psi_buf(n, psi);
      dev_start_vector(p, n, S.seed + step, ctx->stream);
      if (levels > 1) level_starts(n, false);
      ++step;
      lap(1);
      const SolveOut so = solve_step(sb.get(), p, lanzero);
      lap(2);
      std::unique_ptr<hh_trunc> ts, te;
      if (sb->plan.sbs.total_dim() <= kmax) ts.reset(trunc_identity(ctx, sb.get(), 0));
      else ts.reset(truncate(ctx, sb.get(), 0, p, kmin(kmax), kmax, terr, level2()));
      if (sb->plan.ebs.total_dim() <= kmax) te.reset(trunc_identity(ctx, sb.get(), 1));
      else te.reset(truncate(ctx, sb.get(), 1, p, kmin(kmax), kmax, terr, level2()));
      lap(3);
      std::unique_ptr<hh_block> ns(grow(ctx, sb.get(), 0, ts.get()));
      std::unique_ptr<hh_block> ne(grow(ctx, sb.get(), 1, te.get()));
      lap(4);
      record(0, sb.get(), so, ts.get(), wall() - t0, mv0, fl0);
      sysb[label + 1] = std::move(ns);
      envb[label + 1] = std::move(ne);
      syst[label + 1] = std::move(ts);
      envt[label + 1] = std::move(te);
      ++label;
    }
  }

  // one finite-sweep step; returns the bound superblock (kept for the next
  // step's wave transformation)
  void sweep_lr(int bg, int end, double lanzero, double terr, int kmax, bool fresh, int phase) {
    const ModelParams& m = ctx->model;
    const int ns_ = m.nsite();
    int label = bg;
    bool first = true;
    while (label + 1 <= end) {
      const double t0 = wall();
      const int64_t mv0 = ctx->stats.matvecs;
      const double fl0 = ctx->stats.flops_alg;
      hh_block* sb_sys = held(sysb, label, "sys block");
      hh_block* sb_env = held(envb, ns_ - label - 2, "env block");
      lap_start();
      std::unique_ptr<hh_superblock> sb(bind(ctx, sb_sys, sb_env, m.n_elec));
      lap(0);
      const int64_t n = sb->plan.sb.dim;
      double2* p;
      if (first && !prev_sb) {
        const bool carry = !fresh && psi_dim == n;
        if (!carry) {
          p = psi_buf(n, psi);
          dev_start_vector(p, n, S.seed + step, ctx->stream);
        } else {
          p = psi.c();
        }
        if (levels > 1) level_starts(n, carry);
      } else {
        const hh_trunc* ta = held(syst, label, "sys truncation");
        const hh_trunc* tb = held(envt, sb_env->d.length + 1, "env truncation");
        double2* np = psi_buf(n, psi2);
        wave_trans(ctx, 0, prev_sb.get(), psi.c(), ta, tb, sb.get(), np);
        std::swap(psi.p, psi2.p);
        std::swap(psi.bytes, psi2.bytes);
        p = psi.c();
        if (levels > 1) level_starts(n, false, sb.get(), 0, ta, tb);
      }
      ++step;
      lap(1);
      const SolveOut so = solve_step(sb.get(), p, lanzero);
      lap(2);
      psi_dim = n;
      std::unique_ptr<hh_trunc> ts(truncate(ctx, sb.get(), 0, p, kmin(kmax), kmax, terr, level2()));
      lap(3);
      std::unique_ptr<hh_block> nb(grow(ctx, sb.get(), 0, ts.get()));
      lap(4);
      record(phase, sb.get(), so, ts.get(), wall() - t0, mv0, fl0);
      // the previous superblock references blocks that stay stored
      prev_sb = std::move(sb);
      sysb[label + 1] = std::move(nb);
      syst[label + 1] = std::move(ts);
      first = false;
      ++label;
    }
  }

  void sweep_rl(int bg, int end, double lanzero, double terr, int kmax, int phase) {
    const ModelParams& m = ctx->model;
    const int ns_ = m.nsite();
    int label = bg;
    bool first = true;
    while (label - 1 >= end) {
      const double t0 = wall();
      const int64_t mv0 = ctx->stats.matvecs;
      const double fl0 = ctx->stats.flops_alg;
      hh_block* sb_sys = held(sysb, label - 1, "sys block");
      hh_block* sb_env = held(envb, ns_ - label - 1, "env block");
      lap_start();
      std::unique_ptr<hh_superblock> sb(bind(ctx, sb_sys, sb_env, m.n_elec));
      lap(0);
      const int64_t n = sb->plan.sb.dim;
      double2* p;
      if (first && !prev_sb) {
        const bool carry = psi_dim == n;
        if (!carry) {
          p = psi_buf(n, psi);
          dev_start_vector(p, n, S.seed + step, ctx->stream);
        } else {
          p = psi.c();
        }
        if (levels > 1) level_starts(n, carry);
      } else {
        const hh_trunc* ta = held(syst, sb_sys->d.length + 1, "sys truncation");
        const hh_trunc* tb = held(envt, ns_ - label - 1, "env truncation");
        double2* np = psi_buf(n, psi2);
        wave_trans(ctx, 1, prev_sb.get(), psi.c(), ta, tb, sb.get(), np);
        std::swap(psi.p, psi2.p);
        std::swap(psi.bytes, psi2.bytes);
        p = psi.c();
        if (levels > 1) level_starts(n, false, sb.get(), 1, ta, tb);
      }
      ++step;
      lap(1);
      const SolveOut so = solve_step(sb.get(), p, lanzero);
      lap(2);
      psi_dim = n;
      std::unique_ptr<hh_trunc> te(truncate(ctx, sb.get(), 1, p, kmin(kmax), kmax, terr, level2()));
      lap(3);
      std::unique_ptr<hh_block> nb(grow(ctx, sb.get(), 1, te.get()));
      lap(4);
      record(phase, sb.get(), so, te.get(), wall() - t0, mv0, fl0);
      prev_sb = std::move(sb);
      envb[ns_ - label] = std::move(nb);
      envt[ns_ - label] = std::move(te);
      first = false;
      --label;
    }
  }

  void run() {
    const int n = ctx->model.nsite();
    const int kw = S.kept_max_warm > 0 ? S.kept_max_warm : S.kept_max;
    warm_up();
    sweep_lr(n / 2 - 1, n - 2, S.tol * S.first_tol_factor, TRUN_ERR * 15, kw, true, 1);
    prev_sb.reset();
    sweep_rl(n - 2, 1, S.tol * S.first_tol_factor, TRUN_ERR * 15, kw, 1);
    for (int k = 0; k < S.n_sweeps; ++k) {
      prev_sb.reset();
      sweep_lr(1, n - 2, S.tol, TRUN_ERR, S.kept_max, false, k + 2);
      prev_sb.reset();
      sweep_rl(n - 2, 1, S.tol, TRUN_ERR, S.kept_max, k + 2);
    }
    prev_sb.reset();
  }
};

}  // namespace hh

// the most levels a driver targets (the reference's runs use 1 or 2)
constexpr int MAX_LEVELS = 8;

struct hh_driver {
  hh::Driver d;
  size_t taken = 0;
  size_t last_first = 0;   // first record of the last warmup / sweep call
};
