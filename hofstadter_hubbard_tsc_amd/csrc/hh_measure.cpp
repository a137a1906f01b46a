// Measurements on the sweep driver's stored blocks and truncations
// (code/measure.f90, code/correlation.f90): site operators of a block carried
// to its current length through the truncations, applied to the ground state
// as grouped GEMM task lists on the device (density profile, correlations
// across the cut and within a block, every-pair matrices, singlet pairs).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <atomic>
#include <cmath>
#include <complex>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/hhdmrg.h"
#include "hh_core.hpp"
#include "hh_driver.hpp"
#include "hh_engine.hpp"
#include "hh_runtime.hpp"

using namespace hh;

namespace {

// Tracked rank-0 operators of one side's block, carried to its current
// length through the stored truncations (Get_sys_oper_dia / Get_env_oper_dia,
// and Get_sys2b for two-site products): operator q in sector-diagonal layout
// of `sec` at q * osz; lbl[q] = {a, 0} for the density of block site a, {a, b}
// for the product n_a n_b.
struct Tracked {
  DevMem ops;
  std::vector<Sector> sec;
  std::vector<std::pair<int, int>> lbl;
  int64_t osz = 0;
  int find(int a, int b) const {
    for (size_t q = 0; q < lbl.size(); ++q)
      if (lbl[q].first == a && lbl[q].second == b) return (int)q;
    return -1;
  }
};

// row = 0: the density of every site (what = 1: its double occupancy
// n_up n_down).  row = r >= 1: the density of site r and every product
// n_r n_j (j != r) of block sites; the densities n_j with j < r are carried
// only until site r joins.
void track_chain(hh_ctx* ctx, Driver& d, int side, int len, int row, Tracked& T, int what = 0) {
  const ModelParams& m = ctx->model;
  const auto& blocks = side == 0 ? d.sysb : d.envb;
  const auto& truns = side == 0 ? d.syst : d.envt;
  // length 1: the site block itself; its density has reduced elements sqrt(2S+1) N
  T.sec = blocks.at(1)->d.sec;
  T.lbl = {{1, 0}};
  {
    std::vector<double2> h(T.sec.size());
    for (size_t i = 0; i < T.sec.size(); ++i) {
      const double f = what == 1 ? (T.sec[i].N == 2 ? 1.0 : 0.0) : T.sec[i].N;
      h[i] = make_double2(std::sqrt((double)T.sec[i].S2 + 1) * f, 0.0);
    }
    T.ops.alloc(sizeof(double2) * h.size());
    check_hip(hipMemcpyAsync(T.ops.p, h.data(), sizeof(double2) * h.size(), hipMemcpyHostToDevice, ctx->stream),
              "h2d");
  }
  DevMem nxt, X;
  for (int l = 2; l <= len; ++l) {
    need(blocks.count(l) && truns.count(l), "the driver holds no truncation for an intermediate length");
    const hh_trunc* t = truns.at(l).get();
    std::vector<TrackSrc> srcs;
    std::vector<std::pair<int, int>> nl;
    if (row == 0 || l < row) {
      for (size_t q = 0; q < T.lbl.size(); ++q) {
        srcs.push_back({TrackSrc::CARRY, (int)q});
        nl.push_back(T.lbl[q]);
      }
      srcs.push_back({TrackSrc::SITE, row == 0 ? what : 0});
      nl.push_back({l, 0});
    } else if (l == row) {
      srcs.push_back({TrackSrc::SITE, 0});
      nl.push_back({row, 0});
      for (size_t q = 0; q < T.lbl.size(); ++q) {
        srcs.push_back({TrackSrc::PRODUCT, (int)q});
        nl.push_back({T.lbl[q].first, row});
      }
    } else {
      for (size_t q = 0; q < T.lbl.size(); ++q) {
        srcs.push_back({TrackSrc::CARRY, (int)q});
        nl.push_back(T.lbl[q]);
      }
      srcs.push_back({TrackSrc::PRODUCT, T.find(row, 0)});
      nl.push_back({row, l});
    }
    TrackPlan P = compile_track_ops(m, T.sec, t->kept, srcs);
    nxt.alloc(sizeof(double2) * (size_t)std::max<int64_t>(P.new_size * (int64_t)srcs.size(), 1));
    X.alloc(sizeof(double2) * (size_t)std::max<int64_t>(P.x_size, 1));
    BufPtrs b{};
    b.p[B_SYS] = T.ops.c();
    b.p[B_AUX] = t->evec.c();
    b.p[B_X] = X.c();
    b.p[B_OUT] = nxt.c();
    DevTasks s1, s2;
    s1.upload(P.stage1, ctx->stream);
    s2.upload(P.stage2, ctx->stream);
    launch_tiles(s1, b, ctx->stream, PREC_FP64);
    launch_tiles(s2, b, ctx->stream, PREC_FP64);
    check_hip(hipStreamSynchronize(ctx->stream), "operator chain");
    std::swap(T.ops.p, nxt.p);
    std::swap(T.ops.bytes, nxt.bytes);
    T.sec.clear();
    for (auto& k : t->kept) T.sec.push_back({P.eb.sec[k.a].N, P.eb.sec[k.a].S2, k.keep});
    T.lbl = nl;
  }
  T.osz = diag_op_size(T.sec);
}

// Re <x_j, w> for j < nv into h (host), x_j = X + j*stride
void mdots_to_host(hh_ctx* ctx, const double2* X, int64_t stride, int nv, const double2* w, int64_t n, double* h) {
  DevMem part, res;
  part.alloc(sizeof(double) * (size_t)vec_blocks(n) * nv);
  res.alloc(sizeof(double) * nv);
  dev_mdots(X, stride, nv, w, n, (double*)res.p, (double*)part.p, ctx->stream);
  check_hip(hipMemcpyAsync(h, res.p, sizeof(double) * nv, hipMemcpyDeviceToHost, ctx->stream), "d2h");
  check_hip(hipStreamSynchronize(ctx->stream), "measurement");
}

// <n_i> at every lattice position for the ground state psi of the superblock
// (sys block sys_len, env block nsite - sys_len - 2) of the driver's stored
// blocks: measure_operator_dia (code/measure.f90:3-61) with the operator
// chain of Get_sys_oper_dia / Get_env_oper_dia.  out has nsite entries,
// indexed by lattice position - 1.  what = 1 gives the double occupancy
// <n_i,up n_i,down> instead (the same chain with st_double as the site
// operator).
void density_profile(hh_ctx* ctx, Driver& d, int sys_len, const double2* psi, double* out, int what = 0) {
  const ModelParams& m = ctx->model;
  const int nsite = m.nsite();
  const int env_len = nsite - sys_len - 2;
  need(sys_len >= 1 && env_len >= 1, "superblock lengths out of range");
  need(d.sysb.count(sys_len) && d.envb.count(env_len), "the driver holds no block of that length");
  std::unique_ptr<hh_superblock> sb(bind(ctx, d.sysb.at(sys_len).get(), d.envb.at(env_len).get(), m.n_elec));
  const SuperBasis& S = sb->plan.sb;
  for (int side = 0; side < 2; ++side) {
    const int len = side == 0 ? sys_len : env_len;
    Tracked T;
    track_chain(ctx, d, side, len, 0, T, what);
    const EnlBasis& eb = side == 0 ? sb->plan.sbs : sb->plan.ebs;
    // block sites: <O_k> = Re <O_k, rho'>
    DevMem rho, Y;
    rho.alloc(sizeof(double2) * (size_t)std::max<int64_t>(T.osz, 1));
    Y.alloc(sizeof(double2) * (size_t)std::max<int64_t>(S.dim, 1));
    BufPtrs b{};
    b.p[B_IN] = (double2*)psi;
    b.p[B_OUT] = rho.c();
    launch_on(ctx, compile_block_rho(S, side, eb, T.sec), b);
    std::vector<double> h(len + 1);
    mdots_to_host(ctx, T.ops.c(), T.osz, len, rho.c(), T.osz, h.data());
    // centre site: <n> = Re <psi, n_site psi>
    dev_zero(Y.c(), S.dim, ctx->stream);
    b.p[B_OUT] = Y.c();
    launch_on(ctx, compile_site_density(S, side, eb, what), b);
    mdots_to_host(ctx, psi, 0, 1, Y.c(), S.dim, &h[len]);
    for (int k = 1; k <= len; ++k) out[(side == 0 ? k : nsite + 1 - k) - 1] = h[k - 1];
    out[(side == 0 ? sys_len + 1 : nsite - env_len) - 1] = h[len];
  }
}

// Env-side entries of a correlation row: with Y = (A (x) 1) psi for a sys-side
// operator A, <A n_j> = Re <O_j, R> for the env block sites, R = sum Y^T
// conj(psi) over the env block part, and Re <psi, n_site Y> for the env centre
// site.  row points at nsite values indexed by lattice position - 1.
void env_entries(hh_ctx* ctx, hh_superblock* sb, const Tracked& TE, int env_len, const double2* psi,
                 const double2* Y, double* row) {
  const int nsite = ctx->model.nsite();
  const SuperBasis& S = sb->plan.sb;
  DevMem R, W;
  R.alloc(sizeof(double2) * (size_t)std::max<int64_t>(TE.osz, 1));
  W.alloc(sizeof(double2) * (size_t)std::max<int64_t>(S.dim, 1));
  std::vector<double> h(env_len + 1);
  BufPtrs r{};
  r.p[B_IN] = (double2*)psi;
  r.p[B_X] = (double2*)Y;
  r.p[B_OUT] = R.c();
  launch_on(ctx, compile_block_rho(S, 1, sb->plan.ebs, TE.sec, B_X, B_IN), r);
  mdots_to_host(ctx, TE.ops.c(), TE.osz, env_len, R.c(), TE.osz, h.data());
  dev_zero(W.c(), S.dim, ctx->stream);
  BufPtrs w{};
  w.p[B_IN] = (double2*)Y;
  w.p[B_OUT] = W.c();
  launch_on(ctx, compile_site_density(S, 1, sb->plan.ebs), w);
  mdots_to_host(ctx, psi, 0, 1, W.c(), S.dim, &h[env_len]);
  for (int j = 1; j <= env_len; ++j) row[nsite - j] = h[j - 1];   // env block site j: position nsite + 1 - j
  row[nsite - env_len - 1] = h[env_len];                          // env centre: position nsite - env_len
}

// Y = (A (x) 1) psi for tracked sys operator q of T, or for the sys centre
// site's density when q < 0
void apply_sys(hh_ctx* ctx, hh_superblock* sb, const Tracked& T, int q, const double2* psi, double2* Y) {
  const SuperBasis& S = sb->plan.sb;
  BufPtrs b{};
  b.p[B_IN] = (double2*)psi;
  b.p[B_AUX] = T.ops.c();
  b.p[B_OUT] = Y;
  dev_zero(Y, S.dim, ctx->stream);
  launch_on(ctx, compile_apply_sys(S, sb->plan.sbs, T.sec, q < 0 ? 0 : (int64_t)q * T.osz, q < 0), b);
}

// <n_i n_j> for i on the sys side (block sites and the sys centre site) and j
// on the env side, from the same chains.  out is nsite x nsite (row-major,
// lattice positions - 1), symmetric; pairs on the same side are NaN.
void cross_density(hh_ctx* ctx, Driver& d, int sys_len, const double2* psi, double* out) {
  const ModelParams& m = ctx->model;
  const int nsite = m.nsite();
  const int env_len = nsite - sys_len - 2;
  need(sys_len >= 1 && env_len >= 1, "superblock lengths out of range");
  need(d.sysb.count(sys_len) && d.envb.count(env_len), "the driver holds no block of that length");
  std::unique_ptr<hh_superblock> sb(bind(ctx, d.sysb.at(sys_len).get(), d.envb.at(env_len).get(), m.n_elec));
  Tracked TS, TE;
  track_chain(ctx, d, 0, sys_len, 0, TS);
  track_chain(ctx, d, 1, env_len, 0, TE);
  for (int i = 0; i < nsite * nsite; ++i) out[i] = std::numeric_limits<double>::quiet_NaN();
  DevMem Y;
  Y.alloc(sizeof(double2) * (size_t)std::max<int64_t>(sb->plan.sb.dim, 1));
  for (int k = 1; k <= sys_len + 1; ++k) {   // k = sys_len + 1: the sys centre site (position k)
    apply_sys(ctx, sb.get(), TS, k <= sys_len ? k - 1 : -1, psi, Y.c());
    double* row = out + (int64_t)(k - 1) * nsite;
    env_entries(ctx, sb.get(), TE, env_len, psi, Y.c(), row);
    for (int j = sys_len + 2; j <= nsite; ++j) out[(int64_t)(j - 1) * nsite + (k - 1)] = row[j - 1];
  }
}

// Operators of one kind (S: rank 1; c^dag: rank 1/2, fermionic) of every
// site of one side's block, carried to its current length through the stored
// truncations (Get_sys_oper_ndia / update_trun_ndia): operator k-1 is site k,
// layout `lay` over the sectors `sec`, lay.size elements apart.
struct TrackedRank {
  DevMem ops;
  std::vector<Sector> sec;
  OpLayout lay;
  int n = 0;
};

// stage 1 and stage 2 of a chain step, run to completion on the context stream
void run_plan(hh_ctx* ctx, const TrackPlan& P, const BufPtrs& b) {
  DevTasks s1, s2;
  s1.upload(P.stage1, ctx->stream);
  s2.upload(P.stage2, ctx->stream);
  launch_tiles(s1, b, ctx->stream, PREC_FP64);
  launch_tiles(s2, b, ctx->stream, PREC_FP64);
  check_hip(hipStreamSynchronize(ctx->stream), "operator chain");
}

// the sectors of the block that truncation t makes from a block with sectors prev
std::vector<Sector> kept_sectors(const EnlBasis& eb, const hh_trunc* t) {
  std::vector<Sector> out;
  for (auto& k : t->kept) out.push_back({eb.sec[k.a].N, eb.sec[k.a].S2, k.keep});
  return out;
}

// T := the site operator of `kind` on the one-site block (site multiplets
// have dimension 1, so each layout sub-block is one element)
void rank_start(hh_ctx* ctx, const std::vector<Sector>& site_sec, int kind, TrackedRank& T) {
  T.sec = site_sec;
  T.lay = make_layout(T.sec, op_k2(kind), op_dN(kind));
  T.n = 1;
  std::vector<double2> h((size_t)std::max<int64_t>(T.lay.size, 1), make_double2(0.0, 0.0));
  for (auto& sub : T.lay.subs) {
    const cplx v = site_val(ctx->model.kind, kind, sub.bra, sub.ket);
    h[sub.off] = make_double2(v.real(), v.imag());
  }
  T.ops.alloc(sizeof(double2) * h.size());
  check_hip(hipMemcpyAsync(T.ops.p, h.data(), sizeof(double2) * h.size(), hipMemcpyHostToDevice, ctx->stream),
            "h2d");
}

// one growth step of T through truncation t (compile_track_rank): the
// operators `carry` (all when empty and keep_all) and, when add_site, the
// new site's
void rank_step(hh_ctx* ctx, const hh_trunc* t, int kind, int side, TrackedRank& T, bool keep_all = true,
               const std::vector<int>& carry = {}, bool add_site = true) {
  std::vector<int> c = carry;
  if (keep_all) {
    c.resize(T.n);
    for (int q = 0; q < T.n; ++q) c[q] = q;
  }
  TrackPlan P = compile_track_rank(ctx->model, T.sec, t->kept, c, add_site, kind, side);
  const int nout = (int)c.size() + (add_site ? 1 : 0);
  DevMem nxt, X;
  nxt.alloc(sizeof(double2) * (size_t)std::max<int64_t>(P.new_size * nout, 1));
  X.alloc(sizeof(double2) * (size_t)std::max<int64_t>(P.x_size, 1));
  BufPtrs b{};
  b.p[B_SYS] = T.ops.c();
  b.p[B_AUX] = t->evec.c();
  b.p[B_X] = X.c();
  b.p[B_OUT] = nxt.c();
  run_plan(ctx, P, b);
  std::swap(T.ops.p, nxt.p);
  std::swap(T.ops.bytes, nxt.bytes);
  T.sec = kept_sectors(P.eb, t);
  T.lay = make_layout(T.sec, op_k2(kind), op_dN(kind));
  T.n = nout;
}

// only = 0: the operator of every site; only = k >= 1: the operator of site
// k alone (nothing is carried before k joins)
void track_rank_chain(hh_ctx* ctx, Driver& d, int side, int len, int kind, TrackedRank& T, int only = 0) {
  const auto& blocks = side == 0 ? d.sysb : d.envb;
  const auto& truns = side == 0 ? d.syst : d.envt;
  rank_start(ctx, blocks.at(1)->d.sec, kind, T);
  if (only > 1) T.n = 0;
  for (int l = 2; l <= len; ++l) {
    need(blocks.count(l) && truns.count(l), "the driver holds no truncation for an intermediate length");
    const hh_trunc* t = truns.at(l).get();
    if (only == 0) rank_step(ctx, t, kind, side, T);
    else if (l < only) rank_step(ctx, t, kind, side, T, false, {}, false);
    else if (l == only) rank_step(ctx, t, kind, side, T, false, {}, true);
    else rank_step(ctx, t, kind, side, T, false, {0}, false);
  }
}

// A task list repeated n times: copy j writes its outputs (slot B_OUT) j *
// out_stride further on and reads the operator in slot op_buf j * op_stride
// further on.  Used for one measurement term against a run of carried
// operators of one layout, whose plans differ only in the operator's offset.
TaskList replicate(const TaskList& src, int n, int64_t out_stride, uint8_t op_buf, int64_t op_stride) {
  TaskList dst;
  for (const Term& tm : src.terms) need(tm.a_buf != B_OUT && tm.b_buf != B_OUT, "replicated term reads its output");
  for (int j = 0; j < n; ++j) {
    const int t0 = (int)dst.terms.size();
    for (Term tm : src.terms) {
      if (tm.a_buf == op_buf) tm.a_off += (int64_t)j * op_stride;
      if (tm.b_buf == op_buf) tm.b_off += (int64_t)j * op_stride;
      dst.terms.push_back(tm);
    }
    for (Tile tl : src.tiles) {
      tl.t0 += t0;
      tl.t1 += t0;
      if (tl.c_buf == B_OUT) tl.c_off += (int64_t)j * out_stride;
      dst.tiles.push_back(tl);
    }
  }
  dst.flops_alg = src.flops_alg * n;
  dst.flops_exec = src.flops_exec * n;
  return dst;
}

// Y_j = term_j psi for the n single-term plans term_j = P with its operator
// in slot op_buf moved by j * op_stride: stage 1 once (it reads only the
// other operator and psi), stage 2 repeated; Y_j at Y + j * P.sb.dim.
void run_term_family(hh_ctx* ctx, const MatvecPlan& P, int n, uint8_t op_buf, int64_t op_stride, BufPtrs b,
                     double2* Y) {
  DevMem X;
  X.alloc(sizeof(double2) * (size_t)std::max<int64_t>(P.x_size, 1));
  b.p[B_X] = X.c();
  b.p[B_OUT] = Y;
  for (const Term& tm : P.stage1.terms) need(tm.a_buf != op_buf && tm.b_buf != op_buf, "stage 1 reads the operator");
  DevTasks s1, s2;
  s1.upload(P.stage1, ctx->stream);
  s2.upload(n == 1 ? P.stage2 : replicate(P.stage2, n, P.sb.dim, op_buf, op_stride), ctx->stream);
  launch_tiles(s1, b, ctx->stream, PREC_FP64);
  launch_tiles(s2, b, ctx->stream, PREC_FP64);
  check_hip(hipStreamSynchronize(ctx->stream), "measurement");
}

// <psi| term |psi> for every sys-side i (or only i = only_i) and env-side j
// (centre sites included): the two-point term amp [A_i (x) B_j]^0 of
// operator kind `kind` (env operator read as the annihilator partner when
// `tilde`), compiled by compile_matvec's single-term mode and applied to psi.
// The env block partners of one i share one plan (run_term_family); the
// expectations of all of them come from one multi-dot.  val(i, j, v)
// receives the complex expectation; i = sys_len + 1 and j = env_len + 1 are
// the centre sites.
template <typename Sink>
void cross_terms(hh_ctx* ctx, Driver& d, int sys_len, const double2* psi, int kind, bool tilde, cplx amp,
                 Sink val, int only_i = 0) {
  const ModelParams& m = ctx->model;
  const int nsite = m.nsite();
  const int env_len = nsite - sys_len - 2;
  need(sys_len >= 1 && env_len >= 1, "superblock lengths out of range");
  need(d.sysb.count(sys_len) && d.envb.count(env_len), "the driver holds no block of that length");
  // with only_i the sys chain carries that one site's operator (none for the centre)
  TrackedRank TS, TE;
  if (only_i == 0) track_rank_chain(ctx, d, 0, sys_len, kind, TS);
  else if (only_i <= sys_len) track_rank_chain(ctx, d, 0, sys_len, kind, TS, only_i);
  else TS.sec = d.sysb.at(sys_len)->d.sec;
  track_rank_chain(ctx, d, 1, env_len, kind, TE);
  BlockDesc sd, ed;
  sd.side = 0;
  sd.length = sys_len;
  sd.sec = TS.sec;
  ed.side = 1;
  ed.length = env_len;
  ed.sec = TE.sec;
  const Geometry g = make_geometry(m, sys_len, env_len);
  DevMem Y, ipsi;
  std::vector<double> re(env_len + 1), im(env_len + 1);
  const double t_start = wall();
  double t_compile = 0.0;
  for (int i = 1; i <= sys_len + 1; ++i) {
    if (only_i && i != only_i) continue;
    auto term = [&](bool env_block) {
      MeasureTerm t;
      t.amp = amp;
      t.s.kind = t.e.kind = kind;
      t.e.tilde = tilde;
      if (i <= sys_len) {
        t.s.block = true;
        t.s.buf = B_SYS;
        t.s.base = only_i ? 0 : (int64_t)(i - 1) * TS.lay.size;
        t.s.lay = &TS.lay;
      }
      if (env_block) {
        t.e.block = true;
        t.e.buf = B_ENV;
        t.e.base = 0;
        t.e.lay = &TE.lay;
      }
      return t;
    };
    const MeasureTerm tb = term(true), tc = term(false);
    const double t0 = wall();
    const MatvecPlan Pb = compile_matvec(m, g, sd, ed, m.n_elec, &tb);
    const MatvecPlan Pc = compile_matvec(m, g, sd, ed, m.n_elec, &tc);
    t_compile += wall() - t0;
    const int64_t n = Pb.sb.dim;
    Y.alloc(sizeof(double2) * (size_t)std::max<int64_t>(n * (env_len + 1), 1));
    if (!ipsi.p) {   // i psi: Re <y, i psi> = Im <psi, y>
      ipsi.alloc(sizeof(double2) * (size_t)std::max<int64_t>(n, 1));
      dev_cscale_copy(ipsi.c(), psi, 0.0, 1.0, n, ctx->stream);
    }
    BufPtrs b{};
    b.p[B_IN] = (double2*)psi;
    b.p[B_SYS] = TS.ops.c();
    b.p[B_ENV] = TE.ops.c();
    run_term_family(ctx, Pb, env_len, B_ENV, TE.lay.size, b, Y.c());          // env block sites 1 .. env_len
    run_term_family(ctx, Pc, 1, B_ENV, 0, b, Y.c() + (int64_t)env_len * n);   // env centre site
    mdots_to_host(ctx, Y.c(), n, env_len + 1, psi, n, re.data());
    mdots_to_host(ctx, Y.c(), n, env_len + 1, ipsi.c(), n, im.data());
    for (int j = 1; j <= env_len + 1; ++j) val(i, j, cplx(re[j - 1], im[j - 1]));
    if (verbose() && !only_i && (i % 8 == 0 || i == sys_len + 1)) {
      std::fprintf(stderr, "# cross terms: sys site %d/%d done, %.2f s (%.2f s compiling)\n", i, sys_len + 1,
                   wall() - t_start, t_compile);
      std::fflush(stderr);
    }
  }
}

// <S_i . S_j> for i on the sys side (block sites and the sys centre site) and
// j on the env side: the pairs of Get_density_cor with oper1 = oper2 = the
// spin (code/correlation.f90:43-195; sys_block_env_block_cor_ndia and the
// block/site variants, :1666-2656).  Each pair is the exchange-bond term
// -sqrt(3) [S_i (x) S_j]^0.  out as cross_density.
void spin_cross(hh_ctx* ctx, Driver& d, int sys_len, const double2* psi, double* out) {
  const int nsite = ctx->model.nsite();
  const int env_len = nsite - sys_len - 2;
  for (int i = 0; i < nsite * nsite; ++i) out[i] = std::numeric_limits<double>::quiet_NaN();
  cross_terms(ctx, d, sys_len, psi, OP_S, false, -std::sqrt(3.0), [&](int i, int j, cplx v) {
    const int pj = j <= env_len ? nsite + 1 - j : nsite - env_len;   // sys site i sits at position i
    out[(int64_t)(i - 1) * nsite + (pj - 1)] = v.real();
    out[(int64_t)(pj - 1) * nsite + (i - 1)] = v.real();
  });
}

// sum_sigma <c^dag_i,sigma c_j,sigma> for i on the sys side and j on the env
// side (Get_density_cor with st_elec_up / st_elec_down, flag 'F'): the hopping
// bond's term sqrt(2) [c^dag_i (x) c~_j]^0 without its h.c.  out is nsite x
// nsite complex (re, im interleaved, row-major by lattice position - 1); the
// mirrored entry holds the conjugate, <c^dag_j c_i>; pairs on one side NaN.
void hopping_cross(hh_ctx* ctx, Driver& d, int sys_len, const double2* psi, double* out) {
  const int nsite = ctx->model.nsite();
  const int env_len = nsite - sys_len - 2;
  for (int i = 0; i < 2 * nsite * nsite; ++i) out[i] = std::numeric_limits<double>::quiet_NaN();
  cross_terms(ctx, d, sys_len, psi, OP_CDAG, true, std::sqrt(2.0), [&](int i, int j, cplx v) {
    const int pj = j <= env_len ? nsite + 1 - j : nsite - env_len;
    const int64_t a = (int64_t)(i - 1) * nsite + (pj - 1), b = (int64_t)(pj - 1) * nsite + (i - 1);
    out[2 * a] = v.real();
    out[2 * a + 1] = v.imag();
    out[2 * b] = v.real();
    out[2 * b + 1] = -v.imag();
  });
}

// Singlet pair creators P_12 = [c^dag_1 x c^dag_2]^0 of bonds (s1 < s2,
// block-site labels) of one side's block, carried to its current length
// (Get_sys_two / Get_SCOP_Operator_Env): operator q is bonds[q], layout `lay`
// = make_layout(sec, 0, 2), lay.size elements apart.
struct TrackedPairs {
  DevMem ops;
  std::vector<Sector> sec;
  OpLayout lay;
  std::vector<std::pair<int, int>> bonds;
};

void track_pair_chain(hh_ctx* ctx, Driver& d, int side, int len, const std::vector<std::pair<int, int>>& want,
                      TrackedPairs& T) {
  const ModelParams& m = ctx->model;
  const auto& blocks = side == 0 ? d.sysb : d.envb;
  const auto& truns = side == 0 ? d.syst : d.envt;
  int last = 0;
  for (auto& b : want) {
    need(b.first >= 1 && b.first < b.second && b.second <= len, "pair bond outside the block");
    last = std::max(last, b.second);
  }
  // c^dag of every site, one step behind the pairs
  TrackedRank C;
  rank_start(ctx, blocks.at(1)->d.sec, OP_CDAG, C);
  T.sec = C.sec;
  T.lay = make_layout(T.sec, 0, 2);
  T.bonds.clear();
  for (int l = 2; l <= len; ++l) {
    need(blocks.count(l) && truns.count(l), "the driver holds no truncation for an intermediate length");
    const hh_trunc* t = truns.at(l).get();
    std::vector<PairSrc> srcs;
    std::vector<std::pair<int, int>> nb;
    for (size_t q = 0; q < T.bonds.size(); ++q) {
      srcs.push_back({PairSrc::CARRY, (int)q});
      nb.push_back(T.bonds[q]);
    }
    for (auto& b : want)
      if (b.second == l) {
        srcs.push_back({PairSrc::CREATE, b.first - 1});
        nb.push_back(b);
      }
    TrackPlan P = compile_track_pairs(m, T.sec, t->kept, srcs, side);
    DevMem nxt, X;
    nxt.alloc(sizeof(double2) * (size_t)std::max<int64_t>(P.new_size * (int64_t)srcs.size(), 1));
    X.alloc(sizeof(double2) * (size_t)std::max<int64_t>(P.x_size, 1));
    BufPtrs b{};
    b.p[B_SYS] = T.ops.p ? T.ops.c() : nullptr;
    b.p[B_ENV] = C.ops.c();
    b.p[B_AUX] = t->evec.c();
    b.p[B_X] = X.c();
    b.p[B_OUT] = nxt.c();
    run_plan(ctx, P, b);
    std::swap(T.ops.p, nxt.p);
    std::swap(T.ops.bytes, nxt.bytes);
    T.sec = kept_sectors(P.eb, t);
    T.lay = make_layout(T.sec, 0, 2);
    T.bonds = nb;
    if (l < last) rank_step(ctx, t, OP_CDAG, side, C);   // c^dag for the bonds still to be formed
  }
}

// <Delta^dag_a Delta_b> for bonds a in the sys block and b in the env block,
// Delta^dag_ij = (c^dag_i,up c^dag_j,down - c^dag_i,down c^dag_j,up)/sqrt(2):
// the block-block pairs of Get_mSCOP_Operator_cor0 (code/correlation.f90:
// 198-327, sys_block_env_block_cor_ndia).  Each pair is one rank-0 two-point
// term [P_a (x) P_b^dag]^0, the env annihilator read as the adjoint of its
// carried creator.  bonds are nb pairs of lattice positions; out is nb x nb
// complex (re, im), entry (b, a) the conjugate; other entries NaN.
void pair_cross(hh_ctx* ctx, Driver& d, int sys_len, const double2* psi, int nb, const int32_t* bonds,
                double* out) {
  const ModelParams& m = ctx->model;
  const int nsite = m.nsite();
  const int env_len = nsite - sys_len - 2;
  need(sys_len >= 1 && env_len >= 1, "superblock lengths out of range");
  need(d.sysb.count(sys_len) && d.envb.count(env_len), "the driver holds no block of that length");
  need(nb >= 0, "negative bond count");
  for (int64_t i = 0; i < 2 * (int64_t)nb * nb; ++i) out[i] = std::numeric_limits<double>::quiet_NaN();
  std::vector<std::pair<int, int>> sw, ew;   // block-site labels
  std::vector<int> sidx(nb, -1), eidx(nb, -1);
  for (int q = 0; q < nb; ++q) {
    int u = std::min(bonds[2 * q], bonds[2 * q + 1]), v = std::max(bonds[2 * q], bonds[2 * q + 1]);
    need(u >= 1 && v <= nsite && u != v, "bond sites out of range");
    if (v <= sys_len) {
      sidx[q] = (int)sw.size();
      sw.push_back({u, v});
    } else if (u >= nsite + 1 - env_len) {   // env block: label nsite + 1 - position
      eidx[q] = (int)ew.size();
      ew.push_back({nsite + 1 - v, nsite + 1 - u});
    }
  }
  if (sw.empty() || ew.empty()) return;
  TrackedPairs TS, TE;
  track_pair_chain(ctx, d, 0, sys_len, sw, TS);
  track_pair_chain(ctx, d, 1, env_len, ew, TE);
  auto slot = [](const TrackedPairs& T, std::pair<int, int> b) {
    for (size_t q = 0; q < T.bonds.size(); ++q)
      if (T.bonds[q] == b) return (int64_t)q;
    throw HHError(1, "pair operator missing");
  };
  BlockDesc sd, ed;
  sd.side = 0;
  sd.length = sys_len;
  sd.sec = TS.sec;
  ed.side = 1;
  ed.length = env_len;
  ed.sec = TE.sec;
  const Geometry g = make_geometry(m, sys_len, env_len);
  DevMem Y, X, ipsi;
  for (int a = 0; a < nb; ++a) {
    if (sidx[a] < 0) continue;
    for (int b = 0; b < nb; ++b) {
      if (eidx[b] < 0) continue;
      MeasureTerm t;
      t.amp = 1.0;
      t.s.block = t.e.block = true;
      t.s.kind = t.e.kind = OP_PAIR;
      t.s.buf = B_SYS;
      t.s.base = slot(TS, sw[sidx[a]]) * TS.lay.size;
      t.s.lay = &TS.lay;
      t.e.buf = B_ENV;
      t.e.tilde = true;
      t.e.base = slot(TE, ew[eidx[b]]) * TE.lay.size;
      t.e.lay = &TE.lay;
      MatvecPlan P = compile_matvec(m, g, sd, ed, m.n_elec, &t);
      const int64_t n = P.sb.dim;
      Y.alloc(sizeof(double2) * (size_t)std::max<int64_t>(n, 1));
      X.alloc(sizeof(double2) * (size_t)std::max<int64_t>(P.x_size, 1));
      if (!ipsi.p) {
        ipsi.alloc(sizeof(double2) * (size_t)std::max<int64_t>(n, 1));
        dev_cscale_copy(ipsi.c(), psi, 0.0, 1.0, n, ctx->stream);
      }
      BufPtrs bp{};
      bp.p[B_IN] = (double2*)psi;
      bp.p[B_OUT] = Y.c();
      bp.p[B_SYS] = TS.ops.c();
      bp.p[B_ENV] = TE.ops.c();
      bp.p[B_X] = X.c();
      DevTasks s1, s2;
      s1.upload(P.stage1, ctx->stream);
      s2.upload(P.stage2, ctx->stream);
      launch_tiles(s1, bp, ctx->stream, PREC_FP64);
      launch_tiles(s2, bp, ctx->stream, PREC_FP64);
      double re = 0.0, im = 0.0;
      mdots_to_host(ctx, psi, 0, 1, Y.c(), n, &re);
      mdots_to_host(ctx, ipsi.c(), 0, 1, Y.c(), n, &im);
      const int64_t ab = (int64_t)a * nb + b, ba = (int64_t)b * nb + a;
      out[2 * ab] = re;
      out[2 * ab + 1] = im;
      out[2 * ba] = re;
      out[2 * ba + 1] = -im;
    }
  }
}

// <Delta^dag_a Delta_b> for pairs of bonds a = (s1, s2), b = (s3, s4) inside
// the sys block with s2 < s3 (Get_sys_two1, code/correlation.f90:370-428,
// measured by measure_sys_block, :611-673 of measure.f90): four chains
// stepped together through the stored truncations -- c^dag of every site,
// the pair creators P_a (formed when s2 joins), W = P_a c~_3 (formed when s3
// joins) and the scalars Z = -[W x c~_4]^0 = P_a Delta_b (formed when s4
// joins) -- and Z read off the block's reduced density matrix.  out[q] is
// the complex value of quads[q] = {s1, s2, s3, s4} (sys-block labels).
void pair_same_block(hh_ctx* ctx, Driver& d, int sys_len, const double2* psi, const std::vector<std::array<int, 4>>& quads,
                     cplx* out) {
  const ModelParams& m = ctx->model;
  const int nsite = m.nsite();
  const int env_len = nsite - sys_len - 2;
  need(sys_len >= 1 && env_len >= 1, "superblock lengths out of range");
  need(d.sysb.count(sys_len) && d.envb.count(env_len), "the driver holds no block of that length");
  for (int l = 2; l <= sys_len; ++l) need(d.syst.count(l), "the driver holds no truncation for an intermediate length");
  for (auto& q : quads)
    need(q[0] >= 1 && q[0] < q[1] && q[1] < q[2] && q[2] < q[3] && q[3] <= sys_len,
         "bonds must lie in the sys block with the first bond before the second");
  const int nq = (int)quads.size();
  TrackedRank C;   // c^dag of every site
  rank_start(ctx, d.sysb.at(1)->d.sec, OP_CDAG, C);
  // P, W, Z: per quad, slot index in each chain (-1 while not formed)
  DevMem Pops, Wops, Zops, X;
  std::vector<int> pq, wq, zq;   // quad of each slot
  std::vector<Sector> sec = C.sec;
  auto run = [&](const TrackPlan& P, const BufPtrs& b) { run_plan(ctx, P, b); };
  auto slot_of = [](const std::vector<int>& v, int q) {
    for (size_t i = 0; i < v.size(); ++i)
      if (v[i] == q) return (int)i;
    return -1;
  };
  for (int l = 2; l <= sys_len; ++l) {
    const hh_trunc* t = d.syst.at(l).get();
    // Z: carry, form from W
    std::vector<TrackSrc> zs;
    std::vector<int> nz;
    for (size_t i = 0; i < zq.size(); ++i) {
      zs.push_back({TrackSrc::CARRY, (int)i});
      nz.push_back(zq[i]);
    }
    for (int q = 0; q < nq; ++q)
      if (quads[q][3] == l) {
        zs.push_back({TrackSrc::QUAD_PRODUCT, slot_of(wq, q)});
        nz.push_back(q);
      }
    // W: carry, form from P
    std::vector<HoleSrc> ws;
    std::vector<int> nw;
    for (size_t i = 0; i < wq.size(); ++i)
      if (quads[wq[i]][3] > l) {   // still needed for a later Z
        ws.push_back({HoleSrc::CARRY, (int)i});
        nw.push_back(wq[i]);
      }
    for (int q = 0; q < nq; ++q)
      if (quads[q][2] == l) {
        ws.push_back({HoleSrc::PAIR_HOLE, slot_of(pq, q)});
        nw.push_back(q);
      }
    // P: carry, form from c^dag
    std::vector<PairSrc> ps;
    std::vector<int> np;
    for (size_t i = 0; i < pq.size(); ++i)
      if (quads[pq[i]][2] > l) {
        ps.push_back({PairSrc::CARRY, (int)i});
        np.push_back(pq[i]);
      }
    for (int q = 0; q < nq; ++q)
      if (quads[q][1] == l) {
        ps.push_back({PairSrc::CREATE, quads[q][0] - 1});
        np.push_back(q);
      }
    DevMem nZ, nW, nP;
    {
      TrackPlan Pl = compile_track_ops(m, sec, t->kept, zs);
      nZ.alloc(sizeof(double2) * (size_t)std::max<int64_t>(Pl.new_size * (int64_t)zs.size(), 1));
      X.alloc(sizeof(double2) * (size_t)std::max<int64_t>(Pl.x_size, 1));
      BufPtrs b{};
      b.p[B_SYS] = Zops.p ? Zops.c() : nullptr;
      b.p[B_ENV] = Wops.p ? Wops.c() : nullptr;
      b.p[B_AUX] = t->evec.c();
      b.p[B_X] = X.c();
      b.p[B_OUT] = nZ.c();
      run(Pl, b);
    }
    {
      TrackPlan Pl = compile_track_hole(m, sec, t->kept, ws);
      nW.alloc(sizeof(double2) * (size_t)std::max<int64_t>(Pl.new_size * (int64_t)ws.size(), 1));
      X.alloc(sizeof(double2) * (size_t)std::max<int64_t>(Pl.x_size, 1));
      BufPtrs b{};
      b.p[B_SYS] = Wops.p ? Wops.c() : nullptr;
      b.p[B_ENV] = Pops.p ? Pops.c() : nullptr;
      b.p[B_AUX] = t->evec.c();
      b.p[B_X] = X.c();
      b.p[B_OUT] = nW.c();
      run(Pl, b);
    }
    {
      TrackPlan Pl = compile_track_pairs(m, sec, t->kept, ps, 0);
      nP.alloc(sizeof(double2) * (size_t)std::max<int64_t>(Pl.new_size * (int64_t)ps.size(), 1));
      X.alloc(sizeof(double2) * (size_t)std::max<int64_t>(Pl.x_size, 1));
      BufPtrs b{};
      b.p[B_SYS] = Pops.p ? Pops.c() : nullptr;
      b.p[B_ENV] = C.ops.c();
      b.p[B_AUX] = t->evec.c();
      b.p[B_X] = X.c();
      b.p[B_OUT] = nP.c();
      run(Pl, b);
    }
    rank_step(ctx, t, OP_CDAG, 0, C);
    std::swap(Zops.p, nZ.p);
    std::swap(Zops.bytes, nZ.bytes);
    std::swap(Wops.p, nW.p);
    std::swap(Wops.bytes, nW.bytes);
    std::swap(Pops.p, nP.p);
    std::swap(Pops.bytes, nP.bytes);
    zq = nz;
    wq = nw;
    pq = np;
    sec = C.sec;
  }
  need((int)zq.size() == nq, "four-site operator missing");
  std::unique_ptr<hh_superblock> sb(bind(ctx, d.sysb.at(sys_len).get(), d.envb.at(env_len).get(), m.n_elec));
  const SuperBasis& S = sb->plan.sb;
  const int64_t osz = diag_op_size(sec);
  DevMem rho, irho;
  rho.alloc(sizeof(double2) * (size_t)std::max<int64_t>(osz, 1));
  irho.alloc(sizeof(double2) * (size_t)std::max<int64_t>(osz, 1));
  BufPtrs b{};
  b.p[B_IN] = (double2*)psi;
  b.p[B_OUT] = rho.c();
  launch_on(ctx, compile_block_rho(S, 0, sb->plan.sbs, sec), b);
  dev_cscale_copy(irho.c(), rho.c(), 0.0, 1.0, osz, ctx->stream);
  std::vector<double> hr(nq), hi(nq);
  mdots_to_host(ctx, Zops.c(), osz, nq, rho.c(), osz, hr.data());
  mdots_to_host(ctx, Zops.c(), osz, nq, irho.c(), osz, hi.data());
  for (int i = 0; i < nq; ++i) out[zq[i]] = cplx(hr[i], hi[i]);
}

// One row of Get_density_cor for the spin (kind OP_S: <S_r . S_j>) or the
// one-body correlation (OP_CDAG: sum_sigma <c^dag_r c_j>), for a row site r
// on the sys side (1 .. sys_len + 1) and every other site j.  Block pairs
// use the scalar products formed when their later site joins (Get_sys2b:
// S_a . S_b, or sum_sigma c^dag_a c_b with a < b) and carried as rank-0
// operators; a block site with the sys centre site is the within-block bond
// term (sys_block_site_cor_ndia); the env side is the cross term for i = r.
// out has nsite complex values, out[r-1] = NaN.
void corr_row(hh_ctx* ctx, Driver& d, int sys_len, int r, const double2* psi, int kind, cplx* out) {
  const ModelParams& m = ctx->model;
  const int nsite = m.nsite();
  const int env_len = nsite - sys_len - 2;
  need(sys_len >= 1 && env_len >= 1, "superblock lengths out of range");
  need(r >= 1 && r <= sys_len + 1, "the row site must lie on the sys side of the cut (1 .. sys_len + 1)");
  need(d.sysb.count(sys_len) && d.envb.count(env_len), "the driver holds no block of that length");
  for (int l = 2; l <= sys_len; ++l) need(d.syst.count(l), "the driver holds no truncation for an intermediate length");
  const double nan = std::numeric_limits<double>::quiet_NaN();
  for (int j = 0; j < nsite; ++j) out[j] = cplx(nan, nan);
  const bool spin = kind == OP_S;
  const bool centre = r == sys_len + 1;
  const cplx amp = spin ? cplx(-std::sqrt(3.0)) : cplx(std::sqrt(2.0));
  // the operator of every site (C) and the products with site r (T), stepped together
  TrackedRank C;
  rank_start(ctx, d.sysb.at(1)->d.sec, kind, C);
  Tracked T;
  T.sec = C.sec;
  DevMem nxt, X;
  const TrackSrc::What prod = spin ? TrackSrc::SPIN_PRODUCT : TrackSrc::HOP_PRODUCT;
  for (int l = 2; l <= sys_len; ++l) {
    const hh_trunc* t = d.syst.at(l).get();
    if (!centre) {
      std::vector<TrackSrc> srcs;
      std::vector<std::pair<int, int>> nl;   // (earlier site, later site)
      for (size_t q = 0; q < T.lbl.size(); ++q) {
        srcs.push_back({TrackSrc::CARRY, (int)q});
        nl.push_back(T.lbl[q]);
      }
      if (l == r)
        for (int j = 1; j < r; ++j) {
          srcs.push_back({prod, j - 1});
          nl.push_back({j, r});
        }
      if (l > r) {   // C holds only the row site's operator from step r on
        srcs.push_back({prod, 0});
        nl.push_back({r, l});
      }
      TrackPlan P = compile_track_ops(m, T.sec, t->kept, srcs);
      nxt.alloc(sizeof(double2) * (size_t)std::max<int64_t>(P.new_size * (int64_t)srcs.size(), 1));
      X.alloc(sizeof(double2) * (size_t)std::max<int64_t>(P.x_size, 1));
      BufPtrs b{};
      b.p[B_SYS] = T.ops.p ? T.ops.c() : nullptr;
      b.p[B_ENV] = C.ops.c();
      b.p[B_AUX] = t->evec.c();
      b.p[B_X] = X.c();
      b.p[B_OUT] = nxt.c();
      run_plan(ctx, P, b);
      std::swap(T.ops.p, nxt.p);
      std::swap(T.ops.bytes, nxt.bytes);
      T.lbl = nl;
    }
    if (centre || l < r) rank_step(ctx, t, kind, 0, C);   // every site's operator
    else if (l == r) rank_step(ctx, t, kind, 0, C, false, {}, true);
    else rank_step(ctx, t, kind, 0, C, false, {0}, false);
    T.sec = C.sec;
  }
  T.osz = diag_op_size(T.sec);
  std::unique_ptr<hh_superblock> sb(bind(ctx, d.sysb.at(sys_len).get(), d.envb.at(env_len).get(), m.n_elec));
  const SuperBasis& S = sb->plan.sb;
  DevMem ipsi;
  ipsi.alloc(sizeof(double2) * (size_t)std::max<int64_t>(S.dim, 1));
  dev_cscale_copy(ipsi.c(), psi, 0.0, 1.0, S.dim, ctx->stream);
  // pairs inside the sys block: <Q> = Re <Q, rho'> + i Re <Q, i rho'>
  if (!centre && !T.lbl.empty()) {
    DevMem rho, irho;
    rho.alloc(sizeof(double2) * (size_t)std::max<int64_t>(T.osz, 1));
    irho.alloc(sizeof(double2) * (size_t)std::max<int64_t>(T.osz, 1));
    BufPtrs b{};
    b.p[B_IN] = (double2*)psi;
    b.p[B_OUT] = rho.c();
    launch_on(ctx, compile_block_rho(S, 0, sb->plan.sbs, T.sec), b);
    dev_cscale_copy(irho.c(), rho.c(), 0.0, 1.0, T.osz, ctx->stream);
    const int nq = (int)T.lbl.size();
    std::vector<double> hr(nq), hi(nq);
    mdots_to_host(ctx, T.ops.c(), T.osz, nq, rho.c(), T.osz, hr.data());
    mdots_to_host(ctx, T.ops.c(), T.osz, nq, irho.c(), T.osz, hi.data());
    for (int q = 0; q < nq; ++q) {
      const auto& lb = T.lbl[q];
      // the product is O_a O_b (a < b): the row entry of j < r is its conjugate
      const cplx v(hr[q], hi[q]);
      if (lb.first == r) out[lb.second - 1] = v;
      else out[lb.first - 1] = std::conj(v);
    }
  }
  // block site with the sys centre site: the within-block bond term, block
  // operator first (O_j O_centre)
  {
    BlockDesc sd, ed;
    sd.side = 0;
    sd.length = sys_len;
    sd.sec = C.sec;
    ed.side = 1;
    ed.length = env_len;
    ed.sec = sb->env->d.sec;
    // one plan for the first carried operator; with the centre as row site
    // the other block sites follow by moving the operator offset
    MeasureTerm t;
    t.within = true;
    t.amp = amp;
    t.s.block = true;
    t.s.kind = t.e.kind = kind;
    t.e.tilde = !spin;
    t.s.buf = B_SYS;
    t.s.base = 0;
    t.s.lay = &C.lay;
    const MatvecPlan P = compile_matvec(m, make_geometry(m, sys_len, env_len), sd, ed, m.n_elec, &t);
    const int nj = centre ? sys_len : 1;   // otherwise C holds only the row site's operator
    const int64_t n = P.sb.dim;
    DevMem Y;
    Y.alloc(sizeof(double2) * (size_t)std::max<int64_t>(n * nj, 1));
    BufPtrs b{};
    b.p[B_IN] = (double2*)psi;
    b.p[B_SYS] = C.ops.c();
    run_term_family(ctx, P, nj, B_SYS, C.lay.size, b, Y.c());
    std::vector<double> re(nj), im(nj);
    mdots_to_host(ctx, Y.c(), n, nj, psi, n, re.data());
    mdots_to_host(ctx, Y.c(), n, nj, ipsi.c(), n, im.data());
    if (centre)
      for (int j = 1; j <= sys_len; ++j) out[j - 1] = std::conj(cplx(re[j - 1], im[j - 1]));
    else
      out[sys_len] = cplx(re[0], im[0]);
  }
  // env side
  cross_terms(ctx, d, sys_len, psi, kind, !spin, amp, [&](int, int j, cplx v) {
    out[(j <= env_len ? nsite + 1 - j : nsite - env_len) - 1] = v;
  }, r);
}

void spin_row(hh_ctx* ctx, Driver& d, int sys_len, int r, const double2* psi, double* out) {
  const int nsite = ctx->model.nsite();
  std::vector<cplx> v(nsite);
  corr_row(ctx, d, sys_len, r, psi, OP_S, v.data());
  for (int j = 0; j < nsite; ++j) out[j] = v[j].real();
}

// One row of Get_density_cor (code/correlation.f90:43-195): <n_r n_j> for a
// site r on the sys side (block or centre) and every other site j.  Pairs of
// block sites use the tracked products n_r n_j (Get_sys2b / measure_sys_block);
// a block site with the sys centre site uses the transition matrix of
// (n_centre psi, psi) (sys_block_site_cor_ndia); the env side uses
// env_entries.  out has nsite values, out[r-1] = NaN.
void density_row(hh_ctx* ctx, Driver& d, int sys_len, int r, const double2* psi, double* out) {
  const ModelParams& m = ctx->model;
  const int nsite = m.nsite();
  const int env_len = nsite - sys_len - 2;
  need(sys_len >= 1 && env_len >= 1, "superblock lengths out of range");
  need(r >= 1 && r <= sys_len + 1, "the row site must lie on the sys side of the cut (1 .. sys_len + 1)");
  need(d.sysb.count(sys_len) && d.envb.count(env_len), "the driver holds no block of that length");
  std::unique_ptr<hh_superblock> sb(bind(ctx, d.sysb.at(sys_len).get(), d.envb.at(env_len).get(), m.n_elec));
  const SuperBasis& S = sb->plan.sb;
  for (int j = 0; j < nsite; ++j) out[j] = std::numeric_limits<double>::quiet_NaN();
  Tracked TS, TE;
  const bool centre = r == sys_len + 1;
  track_chain(ctx, d, 0, sys_len, centre ? 0 : r, TS);
  track_chain(ctx, d, 1, env_len, 0, TE);
  DevMem Y, rho;
  Y.alloc(sizeof(double2) * (size_t)std::max<int64_t>(S.dim, 1));
  rho.alloc(sizeof(double2) * (size_t)std::max<int64_t>(TS.osz, 1));
  const int nops = (int)TS.lbl.size();
  std::vector<double> h(nops);
  BufPtrs b{};
  b.p[B_IN] = (double2*)psi;
  b.p[B_OUT] = rho.c();
  if (!centre) {
    // pairs inside the sys block: <n_r n_j> = Re <P_rj, rho'(psi, psi)>
    launch_on(ctx, compile_block_rho(S, 0, sb->plan.sbs, TS.sec), b);
    mdots_to_host(ctx, TS.ops.c(), TS.osz, nops, rho.c(), TS.osz, h.data());
    for (int q = 0; q < nops; ++q) {
      const auto& l = TS.lbl[q];
      if (l.second != 0) out[(l.first == r ? l.second : l.first) - 1] = h[q];
    }
  }
  // the sys centre site: Y = n_centre psi; <n_c n_j> = Re <O_j, rho'(Y, psi)>
  apply_sys(ctx, sb.get(), TS, -1, psi, Y.c());
  b.p[B_X] = Y.c();
  launch_on(ctx, compile_block_rho(S, 0, sb->plan.sbs, TS.sec, B_X, B_IN), b);
  mdots_to_host(ctx, TS.ops.c(), TS.osz, nops, rho.c(), TS.osz, h.data());
  if (centre) {
    for (int q = 0; q < nops; ++q) out[TS.lbl[q].first - 1] = h[q];
  } else {
    out[sys_len] = h[TS.find(r, 0)];
    apply_sys(ctx, sb.get(), TS, TS.find(r, 0), psi, Y.c());
  }
  env_entries(ctx, sb.get(), TE, env_len, psi, Y.c(), out);
}

// One chain step for a long source list: compile_track_ops in chunks whose
// stage-1 scratch stays within TRACK_X_ELEMS complex elements, every chunk
// writing its operators at its place in `out`.  The chunks are compiled on up
// to TRACK_THREADS host threads (the compilers are pure functions of their
// inputs), then uploaded and launched in order on the context stream; the
// chunks share one scratch buffer, and the task lists of a chunk stay alive
// until the next upload has synchronised the stream.  Returns the new sectors.
constexpr int64_t TRACK_X_ELEMS = int64_t(1) << 27;   // 2 GiB of complex128
constexpr int TRACK_THREADS = 16;
std::vector<Sector> track_ops_step(hh_ctx* ctx, const std::vector<Sector>& sec, const hh_trunc* t,
                                   const std::vector<TrackSrc>& srcs, int side, const double2* carried,
                                   const double2* rank_ops, DevMem& out, double* compile_seconds = nullptr) {
  const ModelParams& m = ctx->model;
  const double tc = wall();
  const TrackPlan P1 = compile_track_ops(m, sec, t->kept, {srcs.at(0)}, side);
  const int64_t per = std::max<int64_t>(P1.x_size, 1), nsrc = (int64_t)srcs.size();
  const int64_t chunk = std::max<int64_t>(1, std::min<int64_t>(nsrc, TRACK_X_ELEMS / per));
  const int64_t nchunks = (nsrc + chunk - 1) / chunk;
  std::vector<TrackPlan> plans((size_t)nchunks);
  {
    std::atomic<int64_t> next{0};
    std::atomic<bool> failed{false};
    std::string why;
    std::mutex mu;
    auto work = [&]() {
      for (int64_t c; (c = next++) < nchunks && !failed;) {
        try {
          const std::vector<TrackSrc> part(srcs.begin() + c * chunk, srcs.begin() + std::min(nsrc, (c + 1) * chunk));
          plans[(size_t)c] = compile_track_ops(m, sec, t->kept, part, side);
        } catch (const std::exception& e) {
          std::lock_guard<std::mutex> lk(mu);
          if (!failed) why = e.what();
          failed = true;
        }
      }
    };
    const int nth = (int)std::min<int64_t>(nchunks, TRACK_THREADS);
    std::vector<std::thread> th;
    for (int k = 1; k < nth; ++k) th.emplace_back(work);
    work();
    for (auto& x : th) x.join();
    need(!failed, "operator chain compilation: " + why);
  }
  if (compile_seconds) *compile_seconds += wall() - tc;
  out.alloc(sizeof(double2) * (size_t)std::max<int64_t>(P1.new_size * nsrc, 1));
  DevMem X;
  X.alloc(sizeof(double2) * (size_t)(per * chunk));
  std::unique_ptr<DevTasks> run1, run2;   // the chunk in flight
  for (int64_t c = 0; c < nchunks; ++c) {
    auto s1 = std::make_unique<DevTasks>(), s2 = std::make_unique<DevTasks>();
    s1->upload(plans[(size_t)c].stage1, ctx->stream);   // synchronises the stream: the previous chunk is done
    s2->upload(plans[(size_t)c].stage2, ctx->stream);
    run1 = std::move(s1);
    run2 = std::move(s2);
    BufPtrs b{};
    b.p[B_SYS] = (double2*)carried;
    b.p[B_ENV] = (double2*)rank_ops;
    b.p[B_AUX] = t->evec.c();
    b.p[B_X] = X.c();
    b.p[B_OUT] = out.c() + c * chunk * P1.new_size;
    launch_tiles(*run1, b, ctx->stream, PREC_FP64);
    launch_tiles(*run2, b, ctx->stream, PREC_FP64);
  }
  check_hip(hipStreamSynchronize(ctx->stream), "operator chain");
  return kept_sectors(P1.eb, t);
}

// Every pair a < b of block sites of one side, formed when b joins and
// carried to length len (Get_sys_oper_cor_ndia1 / Get_env_oper_cor_ndia1,
// code/correlation.f90:851-974, :1097-1216).  Density (kind OP_N): T holds
// the densities {a, 0} and the products n_a n_b {a, b}.  Spin and hopping: T
// holds the products S_a . S_b or sum_sigma c^dag_a c_b {a, b}, C the
// rank-k operator of every site.
void block_pair_chain(hh_ctx* ctx, Driver& d, int side, int len, int kind, Tracked& T, TrackedRank& C) {
  const auto& blocks = side == 0 ? d.sysb : d.envb;
  const auto& truns = side == 0 ? d.syst : d.envt;
  const bool dens = kind == OP_N;
  T.lbl.clear();
  if (dens) {
    T.sec = blocks.at(1)->d.sec;
    T.lbl = {{1, 0}};
    std::vector<double2> h(T.sec.size());
    for (size_t i = 0; i < T.sec.size(); ++i) h[i] = make_double2(std::sqrt((double)T.sec[i].S2 + 1) * T.sec[i].N, 0.0);
    T.ops.alloc(sizeof(double2) * h.size());
    check_hip(hipMemcpyAsync(T.ops.p, h.data(), sizeof(double2) * h.size(), hipMemcpyHostToDevice, ctx->stream),
              "h2d");
  } else {
    rank_start(ctx, blocks.at(1)->d.sec, kind, C);
    T.sec = C.sec;
  }
  DevMem nxt;
  const TrackSrc::What prod = kind == OP_S ? TrackSrc::SPIN_PRODUCT : TrackSrc::HOP_PRODUCT;
  const double t_start = wall();
  double t_compile = 0.0;
  for (int l = 2; l <= len; ++l) {
    need(blocks.count(l) && truns.count(l), "the driver holds no truncation for an intermediate length");
    const hh_trunc* t = truns.at(l).get();
    std::vector<TrackSrc> srcs;
    std::vector<std::pair<int, int>> nl;
    for (size_t q = 0; q < T.lbl.size(); ++q) {
      srcs.push_back({TrackSrc::CARRY, (int)q});
      nl.push_back(T.lbl[q]);
    }
    if (dens) {
      srcs.push_back({TrackSrc::SITE, 0});
      nl.push_back({l, 0});
      for (size_t q = 0; q < T.lbl.size(); ++q)
        if (T.lbl[q].second == 0) {
          srcs.push_back({TrackSrc::PRODUCT, (int)q});
          nl.push_back({T.lbl[q].first, l});
        }
    } else {
      for (int a = 1; a < l; ++a) {   // C holds the operators of sites 1 .. l-1 in order
        srcs.push_back({prod, a - 1});
        nl.push_back({a, l});
      }
    }
    T.sec = track_ops_step(ctx, T.sec, t, srcs, side, T.ops.p ? T.ops.c() : nullptr, dens ? nullptr : C.ops.c(),
                           nxt, &t_compile);
    std::swap(T.ops.p, nxt.p);
    std::swap(T.ops.bytes, nxt.bytes);
    T.lbl = nl;
    if (!dens) rank_step(ctx, t, kind, side, C);
    if (verbose() && (l % 8 == 0 || l == len)) {
      std::fprintf(stderr, "# pair chain side=%d length=%d/%d: %zu operators, %.2f s (%.2f s compiling)\n", side, l,
                   len, T.lbl.size(), wall() - t_start, t_compile);
      std::fflush(stderr);
    }
  }
  T.osz = diag_op_size(T.sec);
}

// Get_oper_cor_ndia1 (code/correlation.f90:2891-3107) for oper = oper1 = n
// (kind OP_N), the spin (OP_S) or st_elec_up / st_elec_down (OP_CDAG): the
// correlation of every pair of sites of the superblock and the on-site
// values, out[(p-1) * nsite + (q-1)] for lattice positions p, q.
//   * across the cut: cross_density / cross_terms;
//   * pairs of block sites of one side: the carried products of
//     block_pair_chain read off that block's reduced density matrix;
//   * a block site with its side's centre site: the transition matrix of
//     (n_centre psi, psi) for the density, the within-block bond term
//     (single-term compile_matvec, one plan per side) for spin and hopping;
//   * on site: <n^2> = <n> + 2 <n_up n_down>, <S.S> = 3/4 <n - 2 n_up n_down>,
//     sum_sigma <c^dag c> = <n> (st_oper x st_oper1^dag of one site).
// Spin and density are real and symmetric; hopping is Hermitian,
// out[p][q] = sum_sigma <c^dag_p c_q>.
void corr_all(hh_ctx* ctx, Driver& d, int sys_len, const double2* psi, int kind, cplx* out) {
  const ModelParams& m = ctx->model;
  const int nsite = m.nsite();
  const int env_len = nsite - sys_len - 2;
  need(sys_len >= 1 && env_len >= 1, "superblock lengths out of range");
  need(d.sysb.count(sys_len) && d.envb.count(env_len), "the driver holds no block of that length");
  const double nan = std::numeric_limits<double>::quiet_NaN();
  for (int64_t i = 0; i < (int64_t)nsite * nsite; ++i) out[i] = cplx(nan, nan);
  auto put = [&](int p, int q, cplx v) {   // lattice positions; the mirrored entry holds the conjugate
    out[(int64_t)(p - 1) * nsite + (q - 1)] = v;
    out[(int64_t)(q - 1) * nsite + (p - 1)] = std::conj(v);
  };
  const bool dens = kind == OP_N, spin = kind == OP_S;
  const cplx amp = spin ? cplx(-std::sqrt(3.0)) : cplx(std::sqrt(2.0));
  // across the cut
  if (dens) {
    std::vector<double> c((size_t)nsite * nsite);
    cross_density(ctx, d, sys_len, psi, c.data());
    for (int p = 1; p <= sys_len + 1; ++p)
      for (int q = sys_len + 2; q <= nsite; ++q) put(p, q, c[(size_t)(p - 1) * nsite + (q - 1)]);
  } else {
    cross_terms(ctx, d, sys_len, psi, kind, !spin, amp, [&](int i, int j, cplx v) {
      put(i, j <= env_len ? nsite + 1 - j : nsite - env_len, v);
    });
  }
  // on site
  {
    std::vector<double> n(nsite), dd(nsite, 0.0);
    density_profile(ctx, d, sys_len, psi, n.data(), 0);
    if (m.kind == HUBBARD && kind != OP_CDAG) density_profile(ctx, d, sys_len, psi, dd.data(), 1);
    for (int p = 0; p < nsite; ++p)
      out[(int64_t)p * nsite + p] = dens ? n[p] + 2.0 * dd[p] : spin ? 0.75 * (n[p] - 2.0 * dd[p]) : n[p];
  }
  std::unique_ptr<hh_superblock> sb(bind(ctx, d.sysb.at(sys_len).get(), d.envb.at(env_len).get(), m.n_elec));
  const SuperBasis& S = sb->plan.sb;
  DevMem ipsi, Y;
  ipsi.alloc(sizeof(double2) * (size_t)std::max<int64_t>(S.dim, 1));
  dev_cscale_copy(ipsi.c(), psi, 0.0, 1.0, S.dim, ctx->stream);
  for (int side = 0; side < 2; ++side) {
    const int len = side == 0 ? sys_len : env_len;
    auto pos = [&](int k) { return side == 0 ? k : nsite + 1 - k; };   // k = len + 1: the centre site
    const EnlBasis& eb = side == 0 ? sb->plan.sbs : sb->plan.ebs;
    Tracked T;
    TrackedRank C;
    block_pair_chain(ctx, d, side, len, kind, T, C);
    // pairs of block sites: <Q> = Re <Q, rho'> + i Re <Q, i rho'>
    const int nq = (int)T.lbl.size();
    DevMem rho, irho;
    rho.alloc(sizeof(double2) * (size_t)std::max<int64_t>(T.osz, 1));
    BufPtrs b{};
    b.p[B_IN] = (double2*)psi;
    b.p[B_OUT] = rho.c();
    launch_on(ctx, compile_block_rho(S, side, eb, T.sec), b);
    std::vector<double> hr(nq), hi(nq, 0.0);
    if (nq > 0) {
      mdots_to_host(ctx, T.ops.c(), T.osz, nq, rho.c(), T.osz, hr.data());
      if (!dens && !spin) {
        irho.alloc(sizeof(double2) * (size_t)std::max<int64_t>(T.osz, 1));
        dev_cscale_copy(irho.c(), rho.c(), 0.0, 1.0, T.osz, ctx->stream);
        mdots_to_host(ctx, T.ops.c(), T.osz, nq, irho.c(), T.osz, hi.data());
      }
    }
    for (int q = 0; q < nq; ++q)
      if (T.lbl[q].second != 0) put(pos(T.lbl[q].first), pos(T.lbl[q].second), cplx(hr[q], hi[q]));
    // a block site with this side's centre site
    Y.alloc(sizeof(double2) * (size_t)std::max<int64_t>(S.dim * std::max(len, 1), 1));
    if (dens) {
      // Y = n_centre psi; <n_a n_centre> = Re <O_a, rho'(Y, psi)>
      dev_zero(Y.c(), S.dim, ctx->stream);
      BufPtrs w{};
      w.p[B_IN] = (double2*)psi;
      w.p[B_OUT] = Y.c();
      launch_on(ctx, compile_site_density(S, side, eb), w);
      BufPtrs r{};
      r.p[B_IN] = (double2*)psi;
      r.p[B_X] = Y.c();
      r.p[B_OUT] = rho.c();
      launch_on(ctx, compile_block_rho(S, side, eb, T.sec, B_X, B_IN), r);
      mdots_to_host(ctx, T.ops.c(), T.osz, nq, rho.c(), T.osz, hr.data());
      for (int q = 0; q < nq; ++q)
        if (T.lbl[q].second == 0) put(pos(T.lbl[q].first), pos(len + 1), hr[q]);
    } else {
      // amp [O_a (x) O~_centre]^0 inside the enlarged block: one plan for
      // a = 1, the others by moving the operator offset
      BlockDesc sd, ed;
      sd.side = 0;
      sd.length = sys_len;
      sd.sec = side == 0 ? C.sec : sb->sys->d.sec;
      ed.side = 1;
      ed.length = env_len;
      ed.sec = side == 1 ? C.sec : sb->env->d.sec;
      MeasureTerm t;
      t.within = true;
      t.side = side;
      t.amp = amp;
      t.s.block = true;
      t.s.kind = t.e.kind = kind;
      t.e.tilde = !spin;
      t.s.buf = B_SYS;
      t.s.base = 0;
      t.s.lay = &C.lay;
      const MatvecPlan P = compile_matvec(m, make_geometry(m, sys_len, env_len), sd, ed, m.n_elec, &t);
      need(P.sb.dim == S.dim, "measurement superblock differs from the bound one");
      BufPtrs w{};
      w.p[B_IN] = (double2*)psi;
      w.p[B_SYS] = C.ops.c();
      run_term_family(ctx, P, len, B_SYS, C.lay.size, w, Y.c());
      std::vector<double> yr(len), yi(len);
      mdots_to_host(ctx, Y.c(), S.dim, len, psi, S.dim, yr.data());
      mdots_to_host(ctx, Y.c(), S.dim, len, ipsi.c(), S.dim, yi.data());
      for (int a = 1; a <= len; ++a) put(pos(a), pos(len + 1), cplx(yr[a - 1], yi[a - 1]));
    }
  }
}

void corr_all_real(hh_ctx* ctx, Driver& d, int sys_len, const double2* psi, int kind, double* out) {
  const int nsite = ctx->model.nsite();
  std::vector<cplx> v((size_t)nsite * nsite);
  corr_all(ctx, d, sys_len, psi, kind, v.data());
  for (size_t i = 0; i < v.size(); ++i) out[i] = v[i].real();
}

}  // namespace

extern "C" {

int hh_driver_density(hh_driver* h, int sys_len, const double* psi, double* n_out) {
  if (!h) return 1;
  return guard(h->d.ctx, [&] {
    need(psi && n_out, "null argument");
    density_profile(h->d.ctx, h->d, sys_len, (const double2*)psi, n_out);
  });
}

int hh_driver_density_cross(hh_driver* h, int sys_len, const double* psi, double* nn_out) {
  if (!h) return 1;
  return guard(h->d.ctx, [&] {
    need(psi && nn_out, "null argument");
    cross_density(h->d.ctx, h->d, sys_len, (const double2*)psi, nn_out);
  });
}

int hh_driver_density_row(hh_driver* h, int sys_len, int site, const double* psi, double* nn_out) {
  if (!h) return 1;
  return guard(h->d.ctx, [&] {
    need(psi && nn_out, "null argument");
    density_row(h->d.ctx, h->d, sys_len, site, (const double2*)psi, nn_out);
  });
}

int hh_driver_spin_cross(hh_driver* h, int sys_len, const double* psi, double* ss_out) {
  if (!h) return 1;
  return guard(h->d.ctx, [&] {
    need(psi && ss_out, "null argument");
    spin_cross(h->d.ctx, h->d, sys_len, (const double2*)psi, ss_out);
  });
}

int hh_driver_hopping_cross(hh_driver* h, int sys_len, const double* psi, double* g_out) {
  if (!h) return 1;
  return guard(h->d.ctx, [&] {
    need(psi && g_out, "null argument");
    hopping_cross(h->d.ctx, h->d, sys_len, (const double2*)psi, g_out);
  });
}

int hh_driver_pair_cross(hh_driver* h, int sys_len, const double* psi, int nbond, const int32_t* bonds,
                         double* p_out) {
  if (!h) return 1;
  return guard(h->d.ctx, [&] {
    need(psi && p_out && (nbond == 0 || bonds), "null argument");
    pair_cross(h->d.ctx, h->d, sys_len, (const double2*)psi, nbond, bonds, p_out);
  });
}

int hh_driver_spin_row(hh_driver* h, int sys_len, int site, const double* psi, double* ss_out) {
  if (!h) return 1;
  return guard(h->d.ctx, [&] {
    need(psi && ss_out, "null argument");
    spin_row(h->d.ctx, h->d, sys_len, site, (const double2*)psi, ss_out);
  });
}

int hh_driver_hopping_row(hh_driver* h, int sys_len, int site, const double* psi, double* g_out) {
  if (!h) return 1;
  return guard(h->d.ctx, [&] {
    need(psi && g_out, "null argument");
    corr_row(h->d.ctx, h->d, sys_len, site, (const double2*)psi, OP_CDAG, (cplx*)g_out);
  });
}

int hh_driver_density_all(hh_driver* h, int sys_len, const double* psi, double* nn_out) {
  if (!h) return 1;
  return guard(h->d.ctx, [&] {
    need(psi && nn_out, "null argument");
    corr_all_real(h->d.ctx, h->d, sys_len, (const double2*)psi, OP_N, nn_out);
  });
}

int hh_driver_spin_all(hh_driver* h, int sys_len, const double* psi, double* ss_out) {
  if (!h) return 1;
  return guard(h->d.ctx, [&] {
    need(psi && ss_out, "null argument");
    corr_all_real(h->d.ctx, h->d, sys_len, (const double2*)psi, OP_S, ss_out);
  });
}

int hh_driver_hopping_all(hh_driver* h, int sys_len, const double* psi, double* g_out) {
  if (!h) return 1;
  return guard(h->d.ctx, [&] {
    need(psi && g_out, "null argument");
    corr_all(h->d.ctx, h->d, sys_len, (const double2*)psi, OP_CDAG, (cplx*)g_out);
  });
}

int hh_driver_pair_block(hh_driver* h, int sys_len, const double* psi, int nquad, const int32_t* quads,
                         double* p_out) {
  if (!h) return 1;
  return guard(h->d.ctx, [&] {
    need(psi && p_out && (nquad == 0 || quads), "null argument");
    need(nquad >= 0, "negative count");
    std::vector<std::array<int, 4>> q(nquad);
    for (int i = 0; i < nquad; ++i) {
      int a0 = quads[4 * i], a1 = quads[4 * i + 1], b0 = quads[4 * i + 2], b1 = quads[4 * i + 3];
      q[i] = {std::min(a0, a1), std::max(a0, a1), std::min(b0, b1), std::max(b0, b1)};
    }
    pair_same_block(h->d.ctx, h->d, sys_len, (const double2*)psi, q, (cplx*)p_out);
  });
}


}  // extern "C"
