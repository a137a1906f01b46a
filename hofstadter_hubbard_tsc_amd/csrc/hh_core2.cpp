// Task compilers for block growth + renormalisation, density matrices and
// the wave transformation; counter-based start vectors.
#include <cmath>
#include <map>
#include <stdexcept>
#include <tuple>

#include "hh_core.hpp"

namespace hh {

namespace {
inline double nsign(int N) { return (N & 1) ? -1.0 : 1.0; }

inline Term mk(cplx c, uint8_t abuf, int64_t aoff, int ald, uint8_t aop, uint8_t bbuf, int64_t boff, int bld,
               uint8_t bop, int K) {
  Term t;
  t.cr = c.real();
  t.ci = c.imag();
  t.a_off = aoff;
  t.b_off = boff;
  t.a_ld = ald;
  t.b_ld = bld;
  t.K = K;
  t.a_buf = abuf;
  t.b_buf = bbuf;
  t.a_op = aop;
  t.b_op = bop;
  return t;
}

struct Region {
  int buf;
  int64_t off;
  int ld, M, N;
  bool operator<(const Region& o) const { return std::tie(buf, off, ld, M, N) < std::tie(o.buf, o.off, o.ld, o.M, o.N); }
};
struct Collector {
  std::map<Region, std::vector<Term>> m;
  std::vector<Region> order;
  void ensure(const Region& r) {
    if (m.find(r) == m.end()) {
      m[r];
      order.push_back(r);
    }
  }
  void add(const Region& r, const Term& t) {
    ensure(r);
    m[r].push_back(t);
  }
  void emit(TaskList& tl) {
    for (auto& r : order) {
      tl.add_tile(r.buf, r.off, r.ld, r.M, r.N, 0);
      for (auto& t : m[r]) {
        tl.add_term(t);
        tl.flops_exec += t.K > 0 ? 8.0 * r.M * r.N * t.K : 8.0 * r.M * r.N;
      }
    }
    tl.flops_alg = tl.flops_exec;
  }
};
}  // namespace

// ============================================================================
// Get_Hamiltonian_Sys/Env (code/get_ham.f90:4-659) with update_block_*,
// update_site_* (code/update.f90:419-1381) and update_trun_dia/ndia
// (code/update.f90:1387-1605): X = O_enlarged * U, O_new = U^dag X.
// ============================================================================
GrowPlan compile_grow(const ModelParams& m, const Geometry& g, const BlockDesc& blk,
                      const std::vector<KeptSector>& kept) {
  GrowPlan P;
  const auto st = site_sectors(m.kind);
  const int nst = (int)st.size();
  const int nk = n_op_kinds(m.kind);
  P.eb = enlarge_basis(blk.sec, st);
  const EnlBasis& eb = P.eb;
  const int side = blk.side;
  const int L = blk.length;

  BlockDesc& out = P.out;
  out.side = side;
  out.length = L + 1;
  for (auto& k : kept) out.sec.push_back({eb.sec[k.a].N, eb.sec[k.a].S2, k.keep});
  std::vector<int> bs = boundary_sites(m, L + 1);
  for (int k : bs) {
    bool old = false;
    for (int s : blk.sites) old |= (s == k);
    if (old || k == L + 1) out.sites.push_back(k);
  }
  layout_block(out, nk);

  std::map<int, int> kidx;   // enlarged sector -> new sector
  for (int i = 0; i < (int)kept.size(); ++i) kidx[kept[i].a] = i;
  auto pos_of = [&](int k) { return side == 0 ? g.sys_pos(k) : g.env_pos(k); };
  const int pn = pos_of(L + 1);

  // merged operators of the block sites bonded to the new site
  std::vector<std::pair<int, cplx>> ph, pj, pv;
  for (size_t a = 0; a < blk.sites.size(); ++a) {
    Coupling c;
    bool ok = side == 0 ? g.coupling(pos_of(blk.sites[a]), pn, c) : g.coupling(pn, pos_of(blk.sites[a]), c);
    if (!ok) continue;
    if (c.hop != 0.0) ph.push_back({(int)a, side == 0 ? c.hop : std::conj(c.hop)});
    if (c.J != 0.0) pj.push_back({(int)a, c.J});
    if (c.V != 0.0) pv.push_back({(int)a, c.V});
  }
  struct Mrg { int kind; int64_t off; double amp; bool tilde; };
  std::vector<Mrg> mrg;
  auto add_merge = [&](int kind, const std::vector<std::pair<int, cplx>>& parts, double amp, bool tilde) {
    MergeSpec ms;
    ms.side = 0;   // merged ops of this block (slot B_SYS)
    ms.kind = kind;
    ms.parts = parts;
    ms.off = P.merged_size;
    P.merged_size += blk.lay[kind].size;
    P.merges.push_back(ms);
    mrg.push_back({kind, ms.off, amp, tilde});
  };
  if (!ph.empty()) add_merge(OP_CDAG, ph, std::sqrt(2.0), side == 1);
  if (nk > 1 && !pj.empty()) add_merge(OP_S, pj, -std::sqrt(3.0), false);
  if (nk > 1 && !pv.empty()) add_merge(OP_N, pv, 1.0, false);
  for (auto& ms : P.merges) {
    const OpLayout& lay = blk.lay[ms.kind];
    for (auto& sub : lay.subs) {
      P.merge_tasks.add_tile(B_MRG, ms.off + sub.off, sub.rows, sub.rows, sub.cols, 0);
      for (auto& p : ms.parts)
        P.merge_tasks.add_term(mk(p.second, B_SYS, blk.op_off[p.first][ms.kind] + sub.off, sub.rows, OP_NN, 0, 0, 0,
                                  OP_NN, 0));
    }
  }

  Collector s1, s2;
  // ---------------- X_H = H_enl * U ----------------
  std::vector<int64_t> xh(kept.size());
  for (int i = 0; i < (int)kept.size(); ++i) {
    const KeptSector& ks = kept[i];
    const int a = ks.a, J2 = eb.sec[a].S2, da = eb.sec[a].dim, kp = ks.keep;
    xh[i] = P.x_size;
    P.x_size += (int64_t)da * kp;
    auto rows = [&](int pos, int d) { return Region{B_X, xh[i] + pos, da, d, kp}; };
    for (auto& ss : eb.subs[a]) s1.ensure(rows(ss.pos, ss.dim));
    for (auto& ss : eb.subs[a]) {
      const int b = ss.b, s = ss.s, mb = ss.dim;
      // embedded block H (update_block_dia)
      s1.add(rows(ss.pos, mb), mk(std::sqrt((double)(J2 + 1) / (eb.bl[b].S2 + 1)), B_SYS, blk.h_off[b], mb, OP_NN,
                                  B_AUX, ks.u_off + ss.pos, ks.u_ld, OP_NN, mb));
      // on-site U of the new site (update_site_dia with st_double)
      if (m.kind == HUBBARD && m.U != 0.0 && s == 2)
        s1.add(rows(ss.pos, mb), mk(m.U * std::sqrt((double)J2 + 1), B_AUX, ks.u_off + ss.pos, ks.u_ld, OP_NN, 0, 0,
                                    0, OP_NN, 0));
    }
    // bonds block site <-> new site (Get_hamiltonian_ndia / _dia)
    for (auto& mg : mrg) {
      const OpLayout& lay = blk.lay[mg.kind];
      const int k2 = op_k2(mg.kind);
      for (auto& ss : eb.subs[a]) {
        const int b = ss.b, s = ss.s;
        for (int si = 0; si < (int)lay.subs.size(); ++si) {
          const SubBlk& sub = lay.subs[si];
          int bfrom, bto;
          double sg = 1.0;
          if (!mg.tilde) {
            if (sub.ket != b) continue;
            bfrom = b;
            bto = sub.bra;
          } else {
            if (sub.bra != b) continue;
            bfrom = b;
            bto = sub.ket;
            sg = -phase2(1 + eb.bl[bfrom].S2 - eb.bl[bto].S2);
          }
          for (int sp = 0; sp < nst; ++sp) {
            cplx sv;
            if (mg.kind == OP_CDAG) sv = side == 0 ? site_tilde(m.kind, st, sp, s) : site_val(m.kind, OP_CDAG, sp, s);
            else sv = site_val(m.kind, mg.kind, sp, s);
            if (sv == 0.0) continue;
            int pos_out = -1, d_out = 0;
            for (auto& s2x : eb.subs[a])
              if (s2x.b == bto && s2x.s == sp) { pos_out = s2x.pos; d_out = s2x.dim; }
            if (pos_out < 0) continue;
            double c = x12(eb.bl[bto].S2, eb.bl[bfrom].S2, st[sp].S2, st[s].S2, J2, k2) * sg;
            if (c == 0.0) continue;
            if (op_fermion(mg.kind)) c *= side == 0 ? nsign(eb.bl[bfrom].N) : nsign(st[s].N);
            if (side == 1 && (k2 & 1)) c = -c;
            cplx coef = mg.amp * c * sv * std::sqrt((double)J2 + 1);
            const uint8_t fop = mg.tilde ? OP_C : OP_NN;
            s1.add(rows(pos_out, d_out), mk(coef, B_MRG, mg.off + sub.off, sub.rows, fop, B_AUX,
                                             ks.u_off + ss.pos, ks.u_ld, OP_NN, ss.dim));
            if (mg.kind == OP_CDAG) {   // h.c. of the hopping bond
              s1.add(rows(ss.pos, ss.dim), mk(std::conj(coef), B_MRG, mg.off + sub.off, sub.rows,
                                              mg.tilde ? OP_NN : OP_C, B_AUX, ks.u_off + pos_out, ks.u_ld, OP_NN,
                                              d_out));
            }
          }
        }
      }
    }
  }
  // ---------------- X_O = O_enl * U for every kept operator ----------------
  for (size_t si = 0; si < out.sites.size(); ++si) {
    const int k = out.sites[si];
    int old_idx = -1;
    for (size_t q = 0; q < blk.sites.size(); ++q)
      if (blk.sites[q] == k) old_idx = (int)q;
    for (int kind = 0; kind < nk; ++kind) {
      const OpLayout& nl = out.lay[kind];
      const int k2 = op_k2(kind);
      const bool ferm = op_fermion(kind);
      for (auto& nsb : nl.subs) {
        const KeptSector& kb = kept[nsb.bra];
        const KeptSector& kk = kept[nsb.ket];
        const int ap = kb.a, a = kk.a;
        const int J2p = eb.sec[ap].S2, J2 = eb.sec[a].S2, dap = eb.sec[ap].dim, kp = kk.keep;
        const int64_t xo = P.x_size;
        P.x_size += (int64_t)dap * kp;
        auto rows = [&](int pos, int d) { return Region{B_X, xo + pos, dap, d, kp}; };
        for (auto& ss : eb.subs[ap]) s1.ensure(rows(ss.pos, ss.dim));
        if (old_idx >= 0) {
          // old boundary operator embedded (update_block_ndia_*)
          const OpLayout& ol = blk.lay[kind];
          for (auto& osb : ol.subs)
            for (auto& ss : eb.subs[a]) {
              if (ss.b != osb.ket) continue;
              int pos_out = -1, d_out = 0;
              for (auto& t : eb.subs[ap])
                if (t.b == osb.bra && t.s == ss.s) { pos_out = t.pos; d_out = t.dim; }
              if (pos_out < 0) continue;
              double c = e1(eb.bl[osb.bra].S2, eb.bl[osb.ket].S2, st[ss.s].S2, J2p, J2, k2);
              if (c == 0.0) continue;
              if (side == 1 && ferm) c *= nsign(st[ss.s].N);
              s1.add(rows(pos_out, d_out), mk(c, B_SYS, blk.op_off[old_idx][kind] + osb.off, osb.rows, OP_NN, B_AUX,
                                              kk.u_off + ss.pos, kk.u_ld, OP_NN, ss.dim));
            }
        } else {
          // new-site operator embedded (update_site_ndia_*)
          for (auto& ss : eb.subs[a])
            for (int sp = 0; sp < nst; ++sp) {
              cplx v = site_val(m.kind, kind, sp, ss.s);
              if (v == 0.0) continue;
              int pos_out = -1, d_out = 0;
              for (auto& t : eb.subs[ap])
                if (t.b == ss.b && t.s == sp) { pos_out = t.pos; d_out = t.dim; }
              if (pos_out < 0) continue;
              double c = e2(st[sp].S2, st[ss.s].S2, eb.bl[ss.b].S2, J2p, J2, k2);
              if (c == 0.0) continue;
              if (side == 0 && ferm) c *= nsign(eb.bl[ss.b].N);
              s1.add(rows(pos_out, d_out), mk(c * v, B_AUX, kk.u_off + ss.pos, kk.u_ld, OP_NN, 0, 0, 0, OP_NN, 0));
            }
        }
        // stage 2: O_new[bra <- ket] = U_bra^dag X
        Region r{B_OUT, out.op_off[si][kind] + nsb.off, nsb.rows, nsb.rows, nsb.cols};
        s2.add(r, mk(1.0, B_AUX, kb.u_off, kb.u_ld, OP_C, B_X, xo, dap, OP_NN, dap));
      }
    }
  }
  for (int i = 0; i < (int)kept.size(); ++i) {
    const KeptSector& ks = kept[i];
    const int da = eb.sec[ks.a].dim;
    Region r{B_OUT, out.h_off[i], ks.keep, ks.keep, ks.keep};
    s2.add(r, mk(1.0, B_AUX, ks.u_off, ks.u_ld, OP_C, B_X, xh[i], da, OP_NN, da));
  }
  s1.emit(P.stage1);
  s2.emit(P.stage2);
  return P;
}

// ============================================================================
// Get_density_matrix_sys/env (code/get_den.f90:4-99)
// ============================================================================
TaskList compile_density(const SuperBasis& sb, int side, std::vector<int>& rho_sector, std::vector<int64_t>& rho_off,
                         std::vector<int>& rho_dim) {
  TaskList tl;
  int64_t off = 0;
  rho_sector.clear();
  rho_off.clear();
  rho_dim.clear();
  for (int i = 0; i < (int)sb.pairs.size(); ++i) {
    const int r = sb.rows[i], c = sb.cols[i];
    const int d = side == 0 ? r : c;
    rho_sector.push_back(side == 0 ? sb.pairs[i].first : sb.pairs[i].second);
    rho_off.push_back(off);
    rho_dim.push_back(d);
    tl.add_tile(B_OUT, off, d, d, d, 0);
    if (side == 0) tl.add_term(mk(1.0, B_IN, sb.off[i], r, OP_NN, B_IN, sb.off[i], r, OP_C, c));
    else tl.add_term(mk(1.0, B_IN, sb.off[i], r, OP_T, B_IN, sb.off[i], r, OP_R, r));
    tl.flops_alg += 8.0 * d * d * (side == 0 ? c : r);
    off += (int64_t)d * d;
  }
  tl.flops_exec = tl.flops_alg;
  return tl;
}

// ============================================================================
// site densities: operator chain of Get_sys_oper_dia / Get_env_oper_dia and
// the expectation of measure_operator_dia (code/measure.f90:3-61)
// ============================================================================
int64_t diag_op_size(const std::vector<Sector>& sec) {
  int64_t n = 0;
  for (auto& s : sec) n += (int64_t)s.dim * s.dim;
  return n;
}

std::vector<int64_t> diag_op_offsets(const std::vector<Sector>& sec) {
  std::vector<int64_t> o;
  int64_t n = 0;
  for (auto& s : sec) {
    o.push_back(n);
    n += (int64_t)s.dim * s.dim;
  }
  return o;
}

TrackPlan compile_track_ops(const ModelParams& m, const std::vector<Sector>& prev,
                             const std::vector<KeptSector>& kept, const std::vector<TrackSrc>& srcs, int side) {
  TrackPlan P;
  const auto st = site_sectors(m.kind);
  P.eb = enlarge_basis(prev, st);
  const EnlBasis& eb = P.eb;
  std::vector<Sector> nsec;
  for (auto& k : kept) nsec.push_back({eb.sec[k.a].N, eb.sec[k.a].S2, k.keep});
  const std::vector<int64_t> ooff = diag_op_offsets(prev), noff = diag_op_offsets(nsec);
  P.old_size = diag_op_size(prev);
  P.new_size = diag_op_size(nsec);
  const OpLayout sl = make_layout(prev, 2, 0), cl = make_layout(prev, 1, 1);
  const int nst = (int)st.size();
  Collector s1, s2;
  for (size_t q = 0; q < srcs.size(); ++q) {
    const TrackSrc& src = srcs[q];
    for (int i = 0; i < (int)kept.size(); ++i) {
      const KeptSector& ks = kept[i];
      const int a = ks.a, J2 = eb.sec[a].S2, da = eb.sec[a].dim, kp = ks.keep;
      const int64_t xo = P.x_size;
      P.x_size += (int64_t)da * kp;
      auto rows = [&](int pos, int d) { return Region{B_X, xo + pos, da, d, kp}; };
      for (auto& ss : eb.subs[a]) s1.ensure(rows(ss.pos, ss.dim));
      if (src.what == TrackSrc::SPIN_PRODUCT || src.what == TrackSrc::HOP_PRODUCT ||
          src.what == TrackSrc::QUAD_PRODUCT) {
        // -sqrt(3) [S_k(block) x S(site)]^0, sqrt(2) [c^dag_k(block) x
        // c~(site)]^0 or -[W_k(block) x c~(site)]^0 on (b', s') <- (b, s),
        // same J (the bond terms of growth)
        const bool hop = src.what != TrackSrc::SPIN_PRODUCT;
        const double amp = src.what == TrackSrc::SPIN_PRODUCT ? -std::sqrt(3.0)
                           : src.what == TrackSrc::HOP_PRODUCT ? std::sqrt(2.0) : -1.0;
        const OpLayout& ly = hop ? cl : sl;
        const int k2 = hop ? 1 : 2;
        for (auto& ssb : ly.subs)
          for (auto& ss : eb.subs[a]) {
            if (ss.b != ssb.ket) continue;
            for (int sp = 0; sp < nst; ++sp) {
              const cplx sv = hop ? site_tilde(m.kind, st, sp, ss.s) : site_val(m.kind, OP_S, sp, ss.s);
              if (sv == 0.0) continue;
              int pos_out = -1, d_out = 0;
              for (auto& t : eb.subs[a])
                if (t.b == ssb.bra && t.s == sp) { pos_out = t.pos; d_out = t.dim; }
              if (pos_out < 0) continue;
              double c = x12(eb.bl[ssb.bra].S2, eb.bl[ssb.ket].S2, st[sp].S2, st[ss.s].S2, J2, k2);
              if (c == 0.0) continue;
              // c~ of a sys site passes the block (ket state); c^dag of the
              // block passes an env site after c~ has acted (N_s - 1)
              if (hop) c *= side == 0 ? nsign(eb.bl[ss.b].N) : -nsign(st[ss.s].N);
              const cplx coef = amp * c * sv * std::sqrt((double)J2 + 1);
              s1.add(rows(pos_out, d_out), mk(coef, B_ENV, (int64_t)src.k * ly.size + ssb.off, ssb.rows, OP_NN,
                                              B_AUX, ks.u_off + ss.pos, ks.u_ld, OP_NN, ss.dim));
            }
          }
      }
      for (auto& ss : eb.subs[a]) {
        if (src.what == TrackSrc::SPIN_PRODUCT || src.what == TrackSrc::HOP_PRODUCT ||
            src.what == TrackSrc::QUAD_PRODUCT)
          break;
        const int mb = ss.dim;
        const double ns = (double)st[ss.s].N;
        if (src.what == TrackSrc::SITE) {
          // density (or double occupancy) of the new site: sqrt(2J+1) f_s
          // (update_site_dia)
          const double f = src.k == 1 ? (st[ss.s].N == 2 ? 1.0 : 0.0) : ns;
          if (f != 0.0)
            s1.add(rows(ss.pos, mb), mk(f * std::sqrt((double)J2 + 1), B_AUX, ks.u_off + ss.pos, ks.u_ld, OP_NN, 0,
                                        0, 0, OP_NN, 0));
          continue;
        }
        // rank-0 block operator: sqrt((2J+1)/(2S_b+1)) O_b on every (b, s)
        // (update_block_dia); a product with the new site's density scales
        // the (b, s) block by n_s
        double c = std::sqrt((double)(J2 + 1) / (eb.bl[ss.b].S2 + 1));
        if (src.what == TrackSrc::PRODUCT) c *= ns;
        if (c == 0.0) continue;
        s1.add(rows(ss.pos, mb), mk(c, B_SYS, (int64_t)src.k * P.old_size + ooff[ss.b], mb, OP_NN, B_AUX,
                                    ks.u_off + ss.pos, ks.u_ld, OP_NN, mb));
      }
      // O_new[i] = U_i^dag X
      Region r{B_OUT, (int64_t)q * P.new_size + noff[i], kp, kp, kp};
      s2.add(r, mk(1.0, B_AUX, ks.u_off, ks.u_ld, OP_C, B_X, xo, da, OP_NN, da));
    }
  }
  s1.emit(P.stage1);
  s2.emit(P.stage2);
  return P;
}

TrackPlan compile_track_rank(const ModelParams& m, const std::vector<Sector>& prev,
                             const std::vector<KeptSector>& kept, int nold, int kind, int side) {
  std::vector<int> all(nold);
  for (int q = 0; q < nold; ++q) all[q] = q;
  return compile_track_rank(m, prev, kept, all, true, kind, side);
}

TrackPlan compile_track_rank(const ModelParams& m, const std::vector<Sector>& prev,
                             const std::vector<KeptSector>& kept, const std::vector<int>& carry, bool add_site,
                             int kind, int side) {
  TrackPlan P;
  const auto st = site_sectors(m.kind);
  const int nst = (int)st.size();
  const int k2 = op_k2(kind), dN = op_dN(kind);
  const bool ferm = op_fermion(kind);
  P.eb = enlarge_basis(prev, st);
  const EnlBasis& eb = P.eb;
  std::vector<Sector> nsec;
  for (auto& k : kept) nsec.push_back({eb.sec[k.a].N, eb.sec[k.a].S2, k.keep});
  const OpLayout ol = make_layout(prev, k2, dN), nl = make_layout(nsec, k2, dN);
  P.old_size = ol.size;
  P.new_size = nl.size;
  Collector s1, s2;
  const int nout = (int)carry.size() + (add_site ? 1 : 0);
  for (int q = 0; q < nout; ++q) {
    const int src = q < (int)carry.size() ? carry[q] : -1;   // -1: the new site's operator
    for (auto& nsb : nl.subs) {
      const KeptSector& kb = kept[nsb.bra];
      const KeptSector& kk = kept[nsb.ket];
      const int ap = kb.a, a = kk.a;
      const int J2p = eb.sec[ap].S2, J2 = eb.sec[a].S2, dap = eb.sec[ap].dim, kp = kk.keep;
      const int64_t xo = P.x_size;
      P.x_size += (int64_t)dap * kp;
      auto rows = [&](int pos, int d) { return Region{B_X, xo + pos, dap, d, kp}; };
      for (auto& ss : eb.subs[ap]) s1.ensure(rows(ss.pos, ss.dim));
      if (src >= 0) {
        // carried block operator: e1 recoupling on every (b', s) <- (b, s);
        // an env block follows its site in the fermion order
        for (auto& osb : ol.subs)
          for (auto& ss : eb.subs[a]) {
            if (ss.b != osb.ket) continue;
            int pos_out = -1, d_out = 0;
            for (auto& t : eb.subs[ap])
              if (t.b == osb.bra && t.s == ss.s) { pos_out = t.pos; d_out = t.dim; }
            if (pos_out < 0) continue;
            double c = e1(eb.bl[osb.bra].S2, eb.bl[osb.ket].S2, st[ss.s].S2, J2p, J2, k2);
            if (c == 0.0) continue;
            if (side == 1 && ferm) c *= nsign(st[ss.s].N);
            s1.add(rows(pos_out, d_out), mk(c, B_SYS, (int64_t)src * P.old_size + osb.off, osb.rows, OP_NN, B_AUX,
                                            kk.u_off + ss.pos, kk.u_ld, OP_NN, ss.dim));
          }
      } else {
        // operator of the new site: e2 recoupling on every (b, s') <- (b, s);
        // a sys site follows its block in the fermion order
        for (auto& ss : eb.subs[a])
          for (int sp = 0; sp < nst; ++sp) {
            const cplx v = site_val(m.kind, kind, sp, ss.s);
            if (v == 0.0) continue;
            int pos_out = -1, d_out = 0;
            for (auto& t : eb.subs[ap])
              if (t.b == ss.b && t.s == sp) { pos_out = t.pos; d_out = t.dim; }
            if (pos_out < 0) continue;
            double c = e2(st[sp].S2, st[ss.s].S2, eb.bl[ss.b].S2, J2p, J2, k2);
            if (c == 0.0) continue;
            if (side == 0 && ferm) c *= nsign(eb.bl[ss.b].N);
            s1.add(rows(pos_out, d_out), mk(c * v, B_AUX, kk.u_off + ss.pos, kk.u_ld, OP_NN, 0, 0, 0, OP_NN, 0));
          }
      }
      // O_new[bra <- ket] = U_bra^dag X
      Region r{B_OUT, (int64_t)q * P.new_size + nsb.off, nsb.rows, nsb.rows, nsb.cols};
      s2.add(r, mk(1.0, B_AUX, kb.u_off, kb.u_ld, OP_C, B_X, xo, dap, OP_NN, dap));
    }
  }
  s1.emit(P.stage1);
  s2.emit(P.stage2);
  return P;
}

TrackPlan compile_track_pairs(const ModelParams& m, const std::vector<Sector>& prev,
                              const std::vector<KeptSector>& kept, const std::vector<PairSrc>& srcs, int side) {
  TrackPlan P;
  const auto st = site_sectors(m.kind);
  const int nst = (int)st.size();
  P.eb = enlarge_basis(prev, st);
  const EnlBasis& eb = P.eb;
  std::vector<Sector> nsec;
  for (auto& k : kept) nsec.push_back({eb.sec[k.a].N, eb.sec[k.a].S2, k.keep});
  const OpLayout ol = make_layout(prev, 0, 2), cl = make_layout(prev, 1, 1), nl = make_layout(nsec, 0, 2);
  P.old_size = ol.size;
  P.new_size = nl.size;
  Collector s1, s2;
  for (size_t q = 0; q < srcs.size(); ++q) {
    const PairSrc& src = srcs[q];
    for (auto& nsb : nl.subs) {
      const KeptSector& kb = kept[nsb.bra];
      const KeptSector& kk = kept[nsb.ket];
      const int ap = kb.a, a = kk.a;
      const int J2p = eb.sec[ap].S2, J2 = eb.sec[a].S2, dap = eb.sec[ap].dim, kp = kk.keep;
      const int64_t xo = P.x_size;
      P.x_size += (int64_t)dap * kp;
      auto rows = [&](int pos, int d) { return Region{B_X, xo + pos, dap, d, kp}; };
      for (auto& ss : eb.subs[ap]) s1.ensure(rows(ss.pos, ss.dim));
      if (src.what == PairSrc::CARRY) {
        // carried rank-0 operator: e1 recoupling on every (b', s) <- (b, s)
        for (auto& osb : ol.subs)
          for (auto& ss : eb.subs[a]) {
            if (ss.b != osb.ket) continue;
            int pos_out = -1, d_out = 0;
            for (auto& t : eb.subs[ap])
              if (t.b == osb.bra && t.s == ss.s) { pos_out = t.pos; d_out = t.dim; }
            if (pos_out < 0) continue;
            const double c = e1(eb.bl[osb.bra].S2, eb.bl[osb.ket].S2, st[ss.s].S2, J2p, J2, 0);
            if (c == 0.0) continue;
            s1.add(rows(pos_out, d_out), mk(c, B_SYS, (int64_t)src.k * P.old_size + osb.off, osb.rows, OP_NN, B_AUX,
                                            kk.u_off + ss.pos, kk.u_ld, OP_NN, ss.dim));
          }
      } else if (J2p == J2) {
        // [c^dag(block) x c^dag(new site)]^0 on (b', s') <- (b, s), the
        // scalar product of Get_hamiltonian_ndia's bond terms with the
        // site's creator in place of its annihilator
        for (auto& csb : cl.subs)
          for (auto& ss : eb.subs[a]) {
            if (ss.b != csb.ket) continue;
            for (int sp = 0; sp < nst; ++sp) {
              const cplx sv = site_val(m.kind, OP_CDAG, sp, ss.s);
              if (sv == 0.0) continue;
              int pos_out = -1, d_out = 0;
              for (auto& t : eb.subs[ap])
                if (t.b == csb.bra && t.s == sp) { pos_out = t.pos; d_out = t.dim; }
              if (pos_out < 0) continue;
              double c = x12(eb.bl[csb.bra].S2, eb.bl[csb.ket].S2, st[sp].S2, st[ss.s].S2, J2, 1);
              if (c == 0.0) continue;
              c *= side == 0 ? nsign(eb.bl[ss.b].N) : nsign(st[ss.s].N);
              if (side == 1) c = -c;
              const cplx coef = c * sv * std::sqrt((double)J2 + 1);
              s1.add(rows(pos_out, d_out), mk(coef, B_ENV, (int64_t)src.k * cl.size + csb.off, csb.rows, OP_NN,
                                              B_AUX, kk.u_off + ss.pos, kk.u_ld, OP_NN, ss.dim));
            }
          }
      }
      // P_new[bra <- ket] = U_bra^dag X
      Region r{B_OUT, (int64_t)q * P.new_size + nsb.off, nsb.rows, nsb.rows, nsb.cols};
      s2.add(r, mk(1.0, B_AUX, kb.u_off, kb.u_ld, OP_C, B_X, xo, dap, OP_NN, dap));
    }
  }
  s1.emit(P.stage1);
  s2.emit(P.stage2);
  return P;
}

TrackPlan compile_track_hole(const ModelParams& m, const std::vector<Sector>& prev,
                             const std::vector<KeptSector>& kept, const std::vector<HoleSrc>& srcs) {
  TrackPlan P;
  const auto st = site_sectors(m.kind);
  const int nst = (int)st.size();
  P.eb = enlarge_basis(prev, st);
  const EnlBasis& eb = P.eb;
  std::vector<Sector> nsec;
  for (auto& k : kept) nsec.push_back({eb.sec[k.a].N, eb.sec[k.a].S2, k.keep});
  const OpLayout ol = make_layout(prev, 1, 1), pl = make_layout(prev, 0, 2), nl = make_layout(nsec, 1, 1);
  P.old_size = ol.size;
  P.new_size = nl.size;
  Collector s1, s2;
  for (size_t q = 0; q < srcs.size(); ++q) {
    const HoleSrc& src = srcs[q];
    for (auto& nsb : nl.subs) {
      const KeptSector& kb = kept[nsb.bra];
      const KeptSector& kk = kept[nsb.ket];
      const int ap = kb.a, a = kk.a;
      const int J2p = eb.sec[ap].S2, J2 = eb.sec[a].S2, dap = eb.sec[ap].dim, kp = kk.keep;
      const int64_t xo = P.x_size;
      P.x_size += (int64_t)dap * kp;
      auto rows = [&](int pos, int d) { return Region{B_X, xo + pos, dap, d, kp}; };
      for (auto& ss : eb.subs[ap]) s1.ensure(rows(ss.pos, ss.dim));
      if (src.what == HoleSrc::CARRY) {
        // carried rank-1/2 operator of the sys block: e1 on (b', s) <- (b, s)
        for (auto& osb : ol.subs)
          for (auto& ss : eb.subs[a]) {
            if (ss.b != osb.ket) continue;
            int pos_out = -1, d_out = 0;
            for (auto& t : eb.subs[ap])
              if (t.b == osb.bra && t.s == ss.s) { pos_out = t.pos; d_out = t.dim; }
            if (pos_out < 0) continue;
            const double c = e1(eb.bl[osb.bra].S2, eb.bl[osb.ket].S2, st[ss.s].S2, J2p, J2, 1);
            if (c == 0.0) continue;
            s1.add(rows(pos_out, d_out), mk(c, B_SYS, (int64_t)src.k * P.old_size + osb.off, osb.rows, OP_NN, B_AUX,
                                            kk.u_off + ss.pos, kk.u_ld, OP_NN, ss.dim));
          }
      } else {
        // P(block, rank 0) c~(site) on (b', s') <- (b, s): P acts as
        // P_b'b / sqrt(2S_b+1), c~ as the e2-recoupled site element with the
        // string (-1)^N_b of a site operator passing the sys block
        for (auto& psb : pl.subs)
          for (auto& ss : eb.subs[a]) {
            if (ss.b != psb.ket) continue;
            for (int sp = 0; sp < nst; ++sp) {
              const cplx sv = site_tilde(m.kind, st, sp, ss.s);
              if (sv == 0.0) continue;
              int pos_out = -1, d_out = 0;
              for (auto& t : eb.subs[ap])
                if (t.b == psb.bra && t.s == sp) { pos_out = t.pos; d_out = t.dim; }
              if (pos_out < 0) continue;
              double c = e2(st[sp].S2, st[ss.s].S2, eb.bl[ss.b].S2, J2p, J2, 1);
              if (c == 0.0) continue;
              c *= nsign(eb.bl[ss.b].N) / std::sqrt((double)eb.bl[ss.b].S2 + 1);
              s1.add(rows(pos_out, d_out), mk(c * sv, B_ENV, (int64_t)src.k * pl.size + psb.off, psb.rows, OP_NN,
                                              B_AUX, kk.u_off + ss.pos, kk.u_ld, OP_NN, ss.dim));
            }
          }
      }
      // W_new[bra <- ket] = U_bra^dag X
      Region r{B_OUT, (int64_t)q * P.new_size + nsb.off, nsb.rows, nsb.rows, nsb.cols};
      s2.add(r, mk(1.0, B_AUX, kb.u_off, kb.u_ld, OP_C, B_X, xo, dap, OP_NN, dap));
    }
  }
  s1.emit(P.stage1);
  s2.emit(P.stage2);
  return P;
}

TrackPlan compile_track_spin(const ModelParams& m, const std::vector<Sector>& prev,
                             const std::vector<KeptSector>& kept, int nold) {
  return compile_track_rank(m, prev, kept, nold, OP_S, 0);
}

TrackPlan compile_track_step(const ModelParams& m, const std::vector<Sector>& prev,
                             const std::vector<KeptSector>& kept, int nold) {
  std::vector<TrackSrc> srcs;
  for (int k = 0; k < nold; ++k) srcs.push_back({TrackSrc::CARRY, k});
  srcs.push_back({TrackSrc::SITE, 0});
  return compile_track_ops(m, prev, kept, srcs);
}

TaskList compile_block_rho(const SuperBasis& sb, int side, const EnlBasis& eb, const std::vector<Sector>& blk,
                           uint8_t first, uint8_t second) {
  const std::vector<int64_t> off = diag_op_offsets(blk);
  Collector c;
  for (int i = 0; i < (int)sb.pairs.size(); ++i) {
    const int a = side == 0 ? sb.pairs[i].first : sb.pairs[i].second;
    const int r = sb.rows[i], cc = sb.cols[i];
    for (auto& ss : eb.subs[a]) {
      const int d = ss.dim;
      const Region reg{B_OUT, off[ss.b], d, d, d};
      const double w = 1.0 / std::sqrt((double)blk[ss.b].S2 + 1);
      if (side == 0)   // rows of the block part: x_sub psi_sub^dag
        c.add(reg, mk(w, first, sb.off[i] + ss.pos, r, OP_NN, second, sb.off[i] + ss.pos, r, OP_C, cc));
      else   // columns: x_sub^T conj(psi_sub)
        c.add(reg, mk(w, first, sb.off[i] + (int64_t)ss.pos * r, r, OP_T, second, sb.off[i] + (int64_t)ss.pos * r,
                      r, OP_R, r));
    }
  }
  // sectors that no psi block touches still get a (zero) tile
  for (size_t b = 0; b < blk.size(); ++b) c.ensure(Region{B_OUT, off[b], blk[b].dim, blk[b].dim, blk[b].dim});
  TaskList tl;
  c.emit(tl);
  return tl;
}

TaskList compile_apply_sys(const SuperBasis& sb, const EnlBasis& eb, const std::vector<Sector>& blk, int64_t op_off,
                           bool site) {
  const std::vector<int64_t> off = diag_op_offsets(blk);
  TaskList tl;
  for (int i = 0; i < (int)sb.pairs.size(); ++i) {
    const int a = sb.pairs[i].first;
    const int r = sb.rows[i], c = sb.cols[i];
    for (auto& ss : eb.subs[a]) {
      if (site) {   // rows of site multiplets with N = 0 stay as the caller cleared them
        const int n = eb.st[ss.s].N;
        if (n == 0) continue;
        tl.add_tile(B_OUT, sb.off[i] + ss.pos, r, ss.dim, c, 0);
        tl.add_term(mk((double)n, B_IN, sb.off[i] + ss.pos, r, OP_NN, 0, 0, 0, OP_NN, 0));
      } else {
        tl.add_tile(B_OUT, sb.off[i] + ss.pos, r, ss.dim, c, 0);
        tl.add_term(mk(1.0 / std::sqrt((double)blk[ss.b].S2 + 1), B_AUX, op_off + off[ss.b], ss.dim, OP_NN, B_IN,
                       sb.off[i] + ss.pos, r, OP_NN, ss.dim));
      }
    }
  }
  return tl;
}

TaskList compile_site_density(const SuperBasis& sb, int side, const EnlBasis& eb, int what) {
  TaskList tl;
  for (int i = 0; i < (int)sb.pairs.size(); ++i) {
    const int a = side == 0 ? sb.pairs[i].first : sb.pairs[i].second;
    const int r = sb.rows[i], c = sb.cols[i];
    for (auto& ss : eb.subs[a]) {
      const int n = what == 1 ? (eb.st[ss.s].N == 2 ? 1 : 0) : eb.st[ss.s].N;
      if (n == 0) continue;
      if (side == 0) {
        tl.add_tile(B_OUT, sb.off[i] + ss.pos, r, ss.dim, c, 0);
        tl.add_term(mk((double)n, B_IN, sb.off[i] + ss.pos, r, OP_NN, 0, 0, 0, OP_NN, 0));
      } else {
        tl.add_tile(B_OUT, sb.off[i] + (int64_t)ss.pos * r, r, r, ss.dim, 0);
        tl.add_term(mk((double)n, B_IN, sb.off[i] + (int64_t)ss.pos * r, r, OP_NN, 0, 0, 0, OP_NN, 0));
      }
    }
  }
  return tl;
}

// ============================================================================
// wave_trans_LR / wave_trans_RL (code/warm_up.f90:1855-2170)
// ============================================================================
TransPlan compile_wave_trans_lr(const SuperBasis& old_sb, const EnlBasis& old_sys, const EnlBasis& old_env,
                                const std::vector<KeptSector>& sys_trun, const std::vector<KeptSector>& env_trun_of_E,
                                const EnlBasis& new_sys, const EnlBasis& /*new_env*/, const SuperBasis& new_sb) {
  (void)old_sys;
  TransPlan P;
  Collector s1, s2;
  std::map<int, int> strun;
  for (int i = 0; i < (int)sys_trun.size(); ++i) strun[sys_trun[i].a] = i;
  std::map<std::pair<int, int>, int> npair;
  for (int i = 0; i < (int)new_sb.pairs.size(); ++i) npair[new_sb.pairs[i]] = i;
  // every output row range exists (zero when nothing lands)
  for (int i = 0; i < (int)new_sb.pairs.size(); ++i)
    for (auto& ss : new_sys.subs[new_sb.pairs[i].first])
      s2.ensure(Region{B_OUT, new_sb.off[i] + ss.pos, new_sb.rows[i], ss.dim, new_sb.cols[i]});
  for (int i = 0; i < (int)old_sb.pairs.size(); ++i) {
    const int x = old_sb.pairs[i].first, y = old_sb.pairs[i].second;
    auto it = strun.find(x);
    if (it == strun.end()) continue;
    const int ia = it->second;
    const KeptSector& us = sys_trun[ia];
    const int J2 = old_env.sec[y].S2;
    for (auto& es : old_env.subs[y]) {
      const int E = es.b, e = es.s, p = es.pos, d = es.dim;
      const KeptSector& ue = env_trun_of_E[E];
      const int E2 = old_env.bl[E].S2, e2 = old_env.st[e].S2;
      // new sys enlarged sector holding (ia, e) at total spin j_E
      int A = -1, pa = 0, da = 0;
      for (int a = 0; a < (int)new_sys.sec.size(); ++a) {
        if (new_sys.sec[a].S2 != E2) continue;
        for (auto& ss : new_sys.subs[a])
          if (ss.b == ia && ss.s == e) { A = a; pa = ss.pos; da = ss.dim; }
      }
      if (A < 0) continue;
      auto ip = npair.find({A, ue.a});
      if (ip == npair.end()) continue;
      const int i2 = ip->second;
      const int64_t xo = P.x_size;
      P.x_size += (int64_t)us.keep * d;
      s1.add(Region{B_X, xo, us.keep, us.keep, d},
             mk(1.0, B_AUX, us.u_off, us.u_ld, OP_C, B_IN, old_sb.off[i] + (int64_t)p * old_sb.rows[i],
                old_sb.rows[i], OP_NN, old_sb.rows[i]));
      const double sgn = phase2(3 * E2 + 3 * e2 + J2);
      s2.add(Region{B_OUT, new_sb.off[i2] + pa, new_sb.rows[i2], da, new_sb.cols[i2]},
             mk(sgn, B_X, xo, us.keep, OP_NN, B_AUX2, ue.u_off, ue.u_ld, OP_T, d));
    }
  }
  s1.emit(P.stage1);
  s2.emit(P.stage2);
  return P;
}

TransPlan compile_wave_trans_rl(const SuperBasis& old_sb, const EnlBasis& old_sys, const EnlBasis& old_env,
                                const std::vector<KeptSector>& sys_trun_of_A, const std::vector<KeptSector>& env_trun,
                                const EnlBasis& /*new_sys*/, const EnlBasis& new_env, const SuperBasis& new_sb) {
  TransPlan P;
  Collector s1, s2;
  std::map<int, int> etrun;
  for (int i = 0; i < (int)env_trun.size(); ++i) etrun[env_trun[i].a] = i;
  std::map<std::pair<int, int>, int> npair;
  for (int i = 0; i < (int)new_sb.pairs.size(); ++i) npair[new_sb.pairs[i]] = i;
  for (int i = 0; i < (int)new_sb.pairs.size(); ++i)
    for (auto& ss : new_env.subs[new_sb.pairs[i].second])
      s2.ensure(Region{B_OUT, new_sb.off[i] + (int64_t)ss.pos * new_sb.rows[i], new_sb.rows[i], new_sb.rows[i], ss.dim});
  for (int i = 0; i < (int)old_sb.pairs.size(); ++i) {
    const int x = old_sb.pairs[i].first, y = old_sb.pairs[i].second;
    auto it = etrun.find(y);
    if (it == etrun.end()) continue;
    const int jf = it->second;
    const KeptSector& ue = env_trun[jf];
    const int J2 = old_env.sec[y].S2;
    (void)old_env;
    for (auto& ss : old_sys.subs[x]) {
      const int A = ss.b, s = ss.s, p = ss.pos, d = ss.dim;
      const KeptSector& us = sys_trun_of_A[A];
      const int A2 = old_sys.bl[A].S2, s2v = old_sys.st[s].S2;
      int B = -1, pb = 0, db = 0;
      for (int b = 0; b < (int)new_env.sec.size(); ++b) {
        if (new_env.sec[b].S2 != A2) continue;
        for (auto& es : new_env.subs[b])
          if (es.b == jf && es.s == s) { B = b; pb = es.pos; db = es.dim; }
      }
      if (B < 0) continue;
      auto ip = npair.find({us.a, B});
      if (ip == npair.end()) continue;
      const int i2 = ip->second;
      const int64_t xo = P.x_size;
      P.x_size += (int64_t)d * ue.keep;
      s1.add(Region{B_X, xo, d, d, ue.keep},
             mk(1.0, B_IN, old_sb.off[i] + p, old_sb.rows[i], OP_NN, B_AUX2, ue.u_off, ue.u_ld, OP_R, old_sb.cols[i]));
      const double sgn = phase2(s2v + J2 - A2);
      s2.add(Region{B_OUT, new_sb.off[i2] + (int64_t)pb * new_sb.rows[i2], new_sb.rows[i2], new_sb.rows[i2], db},
             mk(sgn, B_AUX, us.u_off, us.u_ld, OP_NN, B_X, xo, d, OP_NN, d));
    }
  }
  s1.emit(P.stage1);
  s2.emit(P.stage2);
  return P;
}

// ============================================================================
// start vectors: splitmix64(seed*2^32 + 2g (+1)) -> uniform [0,1), normalised
// (the oracle's start_vector draws the same stream)
// ============================================================================
static inline uint64_t splitmix64(uint64_t x) {
  uint64_t z = x + 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
void start_vector(int64_t dim, uint64_t seed, std::vector<cplx>& out) {
  out.resize(dim);
  const uint64_t base = (seed & 0xFFFFFFFFull) << 32;
  const double scale = 1.0 / 9007199254740992.0;
  double nrm = 0.0;
  for (int64_t g = 0; g < dim; ++g) {
    double re = (double)(splitmix64(base + 2 * (uint64_t)g) >> 11) * scale;
    double im = (double)(splitmix64(base + 2 * (uint64_t)g + 1) >> 11) * scale;
    out[g] = cplx(re, im);
    nrm += re * re + im * im;
  }
  nrm = std::sqrt(nrm);
  for (auto& v : out) v /= nrm;
}

}  // namespace hh
