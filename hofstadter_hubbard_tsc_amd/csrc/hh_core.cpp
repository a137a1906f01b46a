// Host-side SU(2) algebra, lattice geometry, block layouts and task compilers.
// Every function cites the reference routine it restates; the lowering to
// grouped GEMM tiles is this library's own design.
#include "hh_core.hpp"

#include <algorithm>
#include <cmath>
#include <functional>
#include <set>
#include <stdexcept>
#include <tuple>

namespace hh {

// ============================================================================
// SU(2)
// ============================================================================
namespace {
struct FactTable {
  long double f[512];
  FactTable() {
    f[0] = 1.0L;
    for (int i = 1; i < 512; ++i) f[i] = f[i - 1] * (long double)i;
  }
};
const FactTable& facts() {
  static FactTable t;
  return t;
}
inline bool tri(int a, int b, int c) { return std::abs(a - b) <= c && c <= a + b && ((a + b + c) & 1) == 0; }
}  // namespace

double w6j(int j1, int j2, int j3, int l1, int l2, int l3) {
  // Racah formula, doubled arguments (W6JS, code/update.f90:1665-1716)
  if (!(tri(j1, j2, j3) && tri(j1, l2, l3) && tri(l1, j2, l3) && tri(l1, l2, j3))) return 0.0;
  const long double* F = facts().f;
  const int s1 = (j1 + j2 + j3) / 2, s2 = (j1 + l2 + l3) / 2, s3 = (l1 + j2 + l3) / 2, s4 = (l1 + l2 + j3) / 2;
  const int p1 = (j1 + j2 + l1 + l2) / 2, p2 = (j2 + j3 + l2 + l3) / 2, p3 = (j3 + j1 + l3 + l1) / 2;
  if (p1 + 1 >= 511 || p2 + 1 >= 511 || p3 + 1 >= 511) throw std::runtime_error("w6j: spin too large");
  auto delta = [&](int a, int b, int c) {
    return F[(a + b - c) / 2] * F[(a - b + c) / 2] * F[(-a + b + c) / 2] / F[(a + b + c) / 2 + 1];
  };
  long double sum = 0.0L;
  const int tmin = std::max(std::max(s1, s2), std::max(s3, s4));
  const int tmax = std::min(p1, std::min(p2, p3));
  for (int t = tmin; t <= tmax; ++t) {
    long double den = F[t - s1] * F[t - s2] * F[t - s3] * F[t - s4] * F[p1 - t] * F[p2 - t] * F[p3 - t];
    long double term = F[t + 1] / den;
    sum += (t & 1) ? -term : term;
  }
  long double theta = delta(j1, j2, j3) * delta(j1, l2, l3) * delta(l1, j2, l3) * delta(l1, l2, j3);
  return (double)(sum * std::sqrt(theta));
}

double e1(int b2p, int b2, int s2, int J2p, int J2, int k2) {
  return phase2(b2p + s2 + J2 + k2) * std::sqrt((double)(J2 + 1) * (J2p + 1)) * w6j(b2p, J2p, s2, J2, b2, k2);
}
double e2(int s2p, int s2, int b2, int J2p, int J2, int k2) {
  return phase2(b2 + s2 + J2p + k2) * std::sqrt((double)(J2 + 1) * (J2p + 1)) * w6j(s2p, J2p, b2, J2, s2, k2);
}
double x12(int b2p, int b2, int s2p, int s2, int J2, int k2) {
  return phase2(b2 + k2 + s2p + J2) / std::sqrt((double)(k2 + 1)) * w6j(b2p, b2, k2, s2, s2p, J2);
}

// ============================================================================
// lattice (Get_Lattice / dmrg_setup / get_phase restated: the Landau gauge,
// or get_phase's own gauge-1 pattern)
// ============================================================================
namespace {
void lattice_bonds(const ModelParams& m, int ncol, std::map<std::pair<int, int>, Coupling>& out) {
  const int Ly = m.Ly;
  const double tw = 2.0 * M_PI * m.flux;
  const bool tj = m.kind == TJ;
  auto put = [&](int a, int b, cplx amp) {
    int u = std::min(a, b), v = std::max(a, b);
    if (a > b) amp = std::conj(amp);
    Coupling c{amp, tj ? m.J : 0.0, tj ? -0.25 * m.J : 0.0};
    out[{u, v}] = c;
  };
  if (m.gauge == GAUGE_REFERENCE) {
    // get_phase with gauge = 1 (code/quantum.f90:299-326): the factor of the
    // bond from (x, y) to (x, y+1) is -(-1)^x, to (x+1, y) exp(i theta), to
    // (x+1, y-1) (-1)^x; the neighbours are those of dmrg_setup's triangular
    // branch (code/warm_up.f90:469-505)
    for (int x = 1; x <= ncol; ++x) {
      const double alt = (x % 2 == 0) ? 1.0 : -1.0;   // (-1)^x
      for (int y = 1; y <= Ly; ++y) {
        const int u = y + (x - 1) * Ly;
        if (y < Ly) put(u, u + 1, cplx(-m.t * -alt, 0.0));
        else if (Ly > 2 && m.periodic_y) put(u, 1 + (x - 1) * Ly, cplx(-m.t * -alt, 0.0));
        if (x < ncol) {
          put(u, u + Ly, -m.t * std::exp(cplx(0.0, m.theta)));
          if (y > 1) put(u, u + Ly - 1, cplx(-m.t * alt, 0.0));
          else if (Ly > 2 && m.periodic_y) put(u, u + 2 * Ly - 1, cplx(-m.t * alt, 0.0));
        }
      }
    }
    return;
  }
  for (int x = 1; x <= ncol; ++x) {
    const cplx hy = -m.t * std::exp(cplx(0.0, tw * x));
    for (int y = 1; y <= Ly; ++y) {
      const int u = y + (x - 1) * Ly;
      if (y < Ly) {
        put(u, u + 1, hy);
      } else if (Ly > 2 && m.periodic_y) {
        put(1 + (x - 1) * Ly, u, std::conj(hy));
      }
      if (x < ncol) {
        put(u, u + Ly, cplx(-m.t, 0.0));
        if (m.lattice == TRIANGULAR) {
          // (x, y) <- (x+1, y-1): the same gauge's straight-line phase, x + 1/2 along the bond
          const cplx hd = -m.t * std::exp(cplx(0.0, -tw * (x + 0.5)));
          if (y > 1) put(u, u + Ly - 1, hd);
          else if (Ly > 2 && m.periodic_y) put(u, u + 2 * Ly - 1, hd);
        }
      }
    }
  }
}
}  // namespace

bool Geometry::coupling(int pu, int pv, Coupling& out) const {
  auto it = bonds.find({std::min(pu, pv), std::max(pu, pv)});
  if (it == bonds.end()) return false;
  out = it->second;
  if (pu > pv) out.hop = std::conj(out.hop);
  return true;
}

Geometry make_geometry(const ModelParams& m, int ns, int ne) {
  Geometry g;
  g.ns = ns;
  g.ne = ne;
  const int nn = ns + ne + 2;
  g.ncol = (nn - 1) / m.Ly + 1;
  if (g.ncol == 1) g.ncol = 2;
  g.totsite = m.Ly * g.ncol;
  lattice_bonds(m, g.ncol, g.bonds);
  return g;
}

std::vector<int> boundary_sites(const ModelParams& m, int length) {
  // sites with a full-lattice neighbour beyond the block (dmrg_setup0,
  // code/warm_up.f90:311-341); the env uses the same table by inversion.
  std::map<std::pair<int, int>, Coupling> b;
  lattice_bonds(m, m.Lx, b);
  std::vector<int> out;
  for (int k = 1; k <= length; ++k) {
    bool any = false;
    for (auto& kv : b) {
      int u = kv.first.first, v = kv.first.second;
      if ((u == k && v > length) || (v == k && u > length)) { any = true; break; }
    }
    if (any) out.push_back(k);
  }
  return out;
}

int electrons_for_length(const ModelParams& m, int nsites) {
  // warm-up electron count, code/warm_up.f90:1386-1403
  int n = (int)(1.0 * nsites / m.nsite() * m.n_elec + 0.001);
  n -= n % 2;
  return std::min(n, m.n_elec);
}

// ============================================================================
// bases (Get_basis, Get_super_basis)
// ============================================================================
std::vector<Sector> site_sectors(int kind) {
  if (kind == HUBBARD) return {{0, 0, 1}, {1, 1, 1}, {2, 0, 1}};
  return {{0, 0, 1}, {1, 1, 1}};
}

int EnlBasis::find(int N, int S2) const {
  for (size_t a = 0; a < sec.size(); ++a)
    if (sec[a].N == N && sec[a].S2 == S2) return (int)a;
  return -1;
}
int EnlBasis::total_dim() const {
  int d = 0;
  for (auto& s : sec) d += s.dim;
  return d;
}

EnlBasis enlarge_basis(const std::vector<Sector>& bl, const std::vector<Sector>& st) {
  // Get_basis (code/update.f90:11-112) + QuickSort_Integer (:240-273)
  std::set<std::pair<int, int>> pre;
  for (auto& b : bl)
    for (auto& s : st)
      for (int S = std::abs(b.S2 - s.S2); S <= b.S2 + s.S2; S += 2) pre.insert({b.N + s.N, S});
  std::vector<std::pair<int, int>> q(pre.begin(), pre.end());
  std::sort(q.begin(), q.end(), [](const std::pair<int, int>& x, const std::pair<int, int>& y) {
    if (x.second != y.second) return x.second > y.second;
    return x.first > y.first;
  });
  EnlBasis eb;
  eb.bl = bl;
  eb.st = st;
  for (auto& ns : q) {
    std::vector<SubSector> lst;
    int pos = 0;
    for (int b = 0; b < (int)bl.size(); ++b)
      for (int s = 0; s < (int)st.size(); ++s)
        if (bl[b].N + st[s].N == ns.first && std::abs(bl[b].S2 - st[s].S2) <= ns.second &&
            ns.second <= bl[b].S2 + st[s].S2) {
          lst.push_back({b, s, pos, bl[b].dim});
          pos += bl[b].dim;
        }
    eb.sec.push_back({ns.first, ns.second, pos});
    eb.subs.push_back(lst);
  }
  return eb;
}

SuperBasis super_basis(const EnlBasis& s, const EnlBasis& e, int N) {
  // Get_super_basis (code/update.f90:118-171), total S = 0, sys-major
  SuperBasis sb;
  for (int x = 0; x < (int)s.sec.size(); ++x)
    for (int y = 0; y < (int)e.sec.size(); ++y)
      if (e.sec[y].N == N - s.sec[x].N && e.sec[y].S2 == s.sec[x].S2) {
        if (s.sec[x].dim == 0 || e.sec[y].dim == 0) continue;
        sb.pairs.push_back({x, y});
        sb.off.push_back(sb.dim);
        sb.rows.push_back(s.sec[x].dim);
        sb.cols.push_back(e.sec[y].dim);
        sb.dim += (int64_t)s.sec[x].dim * e.sec[y].dim;
      }
  return sb;
}

// ============================================================================
// layouts
// ============================================================================
OpLayout make_layout(const std::vector<Sector>& sec, int k2, int dN) {
  OpLayout L;
  L.k2 = k2;
  L.dN = dN;
  for (int ket = 0; ket < (int)sec.size(); ++ket)
    for (int bra = 0; bra < (int)sec.size(); ++bra) {
      if (sec[bra].N != sec[ket].N + dN) continue;
      if (!tri(sec[ket].S2, k2, sec[bra].S2)) continue;
      SubBlk sb{bra, ket, sec[bra].dim, sec[ket].dim, L.size};
      L.index[{bra, ket}] = (int)L.subs.size();
      L.subs.push_back(sb);
      L.size += (int64_t)sec[bra].dim * sec[ket].dim;
    }
  return L;
}

int n_op_kinds(int model_kind) { return model_kind == HUBBARD ? 1 : 3; }

void layout_block(BlockDesc& b, int nkinds) {
  b.h_off.clear();
  b.size = 0;
  for (auto& s : b.sec) {
    b.h_off.push_back(b.size);
    b.size += (int64_t)s.dim * s.dim;
  }
  for (int k = 0; k < 3; ++k) b.lay[k] = make_layout(b.sec, op_k2(k), op_dN(k));
  b.op_off.assign(b.sites.size(), std::vector<int64_t>(3, -1));
  for (size_t i = 0; i < b.sites.size(); ++i)
    for (int k = 0; k < nkinds; ++k) {
      b.op_off[i][k] = b.size;
      b.size += b.lay[k].size;
    }
}

// ============================================================================
// site operators (Get_site_operator_su2, code/warm_up.f90:818-991)
// ============================================================================
// reduced element <s'||O||s> of a site operator; sector indices of site_sectors
cplx site_val(int model_kind, int kind, int sp, int s) {
  const double r2 = std::sqrt(2.0);
  if (kind == OP_CDAG) {
    if (sp == 1 && s == 0) return -r2;
    if (model_kind == HUBBARD && sp == 2 && s == 1) return r2;
    return 0.0;
  }
  if (kind == OP_S) return (sp == 1 && s == 1) ? std::sqrt(1.5) : 0.0;
  if (kind == OP_N) {
    if (sp != s) return 0.0;
    if (s == 1) return r2;
    if (s == 2) return 2.0;
    return 0.0;
  }
  return 0.0;
}
// annihilator partner of the site c^dag: <j||a~||j'> = -(-1)^(1/2+j'-j) conj(<j'||c^dag||j>)
cplx site_tilde(int model_kind, const std::vector<Sector>& st, int sp, int s) {
  cplx c = site_val(model_kind, OP_CDAG, s, sp);
  if (c == 0.0) return 0.0;
  return -phase2(1 + st[s].S2 - st[sp].S2) * std::conj(c);
}

namespace {
inline double nsign(int N) { return (N & 1) ? -1.0 : 1.0; }

// (block sector, site sector) -> all enlarged (sector, pos, dim)
struct SubIndex {
  std::map<std::pair<int, int>, std::vector<std::tuple<int, int, int>>> m;
  explicit SubIndex(const EnlBasis& eb) {
    for (int a = 0; a < (int)eb.sec.size(); ++a)
      for (auto& ss : eb.subs[a]) m[{ss.b, ss.s}].push_back(std::make_tuple(a, ss.pos, ss.dim));
  }
  const std::vector<std::tuple<int, int, int>>& get(int b, int s) const {
    static const std::vector<std::tuple<int, int, int>> empty;
    auto it = m.find({b, s});
    return it == m.end() ? empty : it->second;
  }
  // position of (b, s) inside enlarged sector a, -1 if absent
  int pos_in(int b, int s, int a) const {
    for (auto& t : get(b, s))
      if (std::get<0>(t) == a) return std::get<1>(t);
    return -1;
  }
};

inline Term mk_term(cplx c, uint8_t abuf, int64_t aoff, int ald, uint8_t aop, uint8_t bbuf, int64_t boff, int bld,
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
inline uint8_t adj_op(uint8_t op) {
  switch (op) {
    case OP_NN: return OP_C;
    case OP_C: return OP_NN;
    case OP_T: return OP_R;
    default: return OP_T;
  }
}

// Tile keyed collection: output region -> terms, emitted in a fixed order.
struct TileKey {
  int buf;
  int64_t off;
  int ld, M, N;
  bool operator<(const TileKey& o) const {
    return std::tie(buf, off, ld, M, N) < std::tie(o.buf, o.off, o.ld, o.M, o.N);
  }
};
struct TileCollector {
  std::map<TileKey, std::vector<Term>> m;
  std::vector<TileKey> order;
  void ensure(const TileKey& k) {
    if (m.find(k) == m.end()) {
      m[k];
      order.push_back(k);
    }
  }
  void add(const TileKey& k, const Term& t) {
    ensure(k);
    m[k].push_back(t);
  }
  void emit(TaskList& tl, int accumulate) {
    for (auto& k : order) {
      tl.add_tile(k.buf, k.off, k.ld, k.M, k.N, accumulate);
      for (auto& t : m[k]) tl.add_term(t);
    }
  }
};

// merged-operator bookkeeping
struct MergeBuilder {
  std::vector<MergeSpec>& specs;
  int64_t& size;
  MergeBuilder(std::vector<MergeSpec>& s, int64_t& z) : specs(s), size(z) {}
  int add(int side, int kind, const std::vector<std::pair<int, cplx>>& parts, const OpLayout& lay) {
    MergeSpec m;
    m.side = side;
    m.kind = kind;
    m.parts = parts;
    m.off = size;
    size += lay.size;
    specs.push_back(m);
    return (int)specs.size() - 1;
  }
};

void compile_merges(const std::vector<MergeSpec>& specs, const BlockDesc* blocks[2], TaskList& tl,
                    uint8_t side_buf[2]) {
  for (auto& ms : specs) {
    const BlockDesc& b = *blocks[ms.side];
    const OpLayout& lay = b.lay[ms.kind];
    for (auto& sub : lay.subs) {
      tl.add_tile(B_MRG, ms.off + sub.off, sub.rows, sub.rows, sub.cols, 0);
      for (auto& p : ms.parts) {
        int64_t src = b.op_off[p.first][ms.kind] + sub.off;
        tl.add_term(mk_term(p.second, side_buf[ms.side], src, sub.rows, OP_NN, 0, 0, 0, OP_NN, 0));
      }
      tl.flops_alg += 8.0 * sub.rows * sub.cols * ms.parts.size();
    }
  }
}

}  // namespace

// ============================================================================
// matvec compiler: hmerge (code/mat_vec.f90:3623-3852) + model_wave (:3476-3620)
// ============================================================================
namespace {
// A piece of a two-body term: an operator on the sys/env block (matrix
// sub-blocks) or on a centre site (scalar reduced elements).
struct Piece {
  bool block = false;
  int side = 0;          // 0 sys, 1 env
  int kind = OP_CDAG;
  bool tilde = false;    // annihilator partner stored as s*conj(C)^T of a c^dag layout
  uint8_t buf = B_SYS;
  int64_t base = 0;      // arena offset of the (merged or stored) operator
  const OpLayout* lay = nullptr;
  int k2 = 1;
  bool fermion = true;
  // site piece: value(s', s)
  std::function<cplx(int, int)> site;
};

// one transition of a piece from a ket sector: (bra, factor-source)
struct Trans {
  int to;               // bra block sector or bra site sector
  int sub = -1;         // layout sub index (block piece)
  cplx val = 0.0;       // site piece value
  double s = 1.0;       // tilde sign
};

std::vector<Trans> transitions(const Piece& p, int from, const std::vector<Sector>& sec, int nsite_sec) {
  std::vector<Trans> out;
  if (p.block) {
    if (!p.tilde) {
      for (int i = 0; i < (int)p.lay->subs.size(); ++i)
        if (p.lay->subs[i].ket == from) out.push_back({p.lay->subs[i].bra, i, 0.0, 1.0});
    } else {
      // a~[e' <- e] = s(e',e) conj(C[e <- e'])^T : C sub with bra = from
      for (int i = 0; i < (int)p.lay->subs.size(); ++i)
        if (p.lay->subs[i].bra == from) {
          int ep = p.lay->subs[i].ket;
          // fermionic c^dag: the annihilator partner's sign; a bosonic
          // rank-0 operator: its plain adjoint
          double s = p.fermion ? -phase2(1 + sec[from].S2 - sec[ep].S2) : 1.0;
          out.push_back({ep, i, 0.0, s});
        }
    }
  } else {
    for (int sp = 0; sp < nsite_sec; ++sp) {
      cplx v = p.site(sp, from);
      if (v != 0.0) out.push_back({sp, -1, v, 1.0});
    }
  }
  return out;
}
}  // namespace

MatvecPlan compile_matvec(const ModelParams& m, const Geometry& g, const BlockDesc& sys, const BlockDesc& env,
                          int N, const MeasureTerm* only) {
  MatvecPlan P;
  const auto st = site_sectors(m.kind);
  const int nst = (int)st.size();
  P.sbs = enlarge_basis(sys.sec, st);
  P.ebs = enlarge_basis(env.sec, st);
  P.sb = super_basis(P.sbs, P.ebs, N);
  const EnlBasis& SB = P.sbs;
  const EnlBasis& EB = P.ebs;
  const SuperBasis& sb = P.sb;
  SubIndex sidx(SB), eidx(EB);
  std::map<int, int> sys_pair, env_pair;   // enlarged sector -> super index
  for (int i = 0; i < (int)sb.pairs.size(); ++i) {
    sys_pair[sb.pairs[i].first] = i;
    env_pair[sb.pairs[i].second] = i;
  }
  MergeBuilder mb(P.merges, P.merged_size);

  auto site_fn = [&](int kind) { return [&, kind](int sp, int s) { return site_val(m.kind, kind, sp, s); }; };
  auto tilde_fn = [&]() { return [&](int sp, int s) { return site_tilde(m.kind, st, sp, s); }; };

  TileCollector out, xs;
  std::map<std::tuple<int, int, int, int, int>, int64_t> xoff;   // (i, rowsub pos, piece id, sub, adj) -> X offset
  int piece_id_counter = 0;

  // psi tile helpers
  auto in_tile_off = [&](int i, int rp, int cq) { return sb.off[i] + rp + (int64_t)cq * sb.rows[i]; };

  // ---------- cross terms: sys piece P x env piece Q, coupled to a scalar --------
  auto cross = [&](const Piece& Ps, const Piece& Qe, cplx amp, bool herm_partner) {
    const int k2 = Ps.k2;
    const int pid = piece_id_counter++;
    for (int i = 0; i < (int)sb.pairs.size(); ++i) {
      const int al = sb.pairs[i].first, be = sb.pairs[i].second;
      const int J2 = SB.sec[al].S2;
      const int Nal = SB.sec[al].N;
      for (auto& rs : SB.subs[al]) {
        const int b = rs.b, s = rs.s;
        // sys transitions
        std::vector<Trans> tp = transitions(Ps, Ps.block ? b : s, SB.bl, nst);
        for (auto& tr : tp) {
          const int bh = Ps.block ? tr.to : b;
          const int sh = Ps.block ? s : tr.to;
          for (auto& tgt : sidx.get(bh, sh)) {
            const int alp = std::get<0>(tgt);
            auto itp = sys_pair.find(alp);
            if (itp == sys_pair.end()) continue;
            const int ip = itp->second;
            const int bep = sb.pairs[ip].second;
            const int J2p = SB.sec[alp].S2;
            const int rpos_out = std::get<1>(tgt), rdim_out = std::get<2>(tgt);
            double tf;
            if (Ps.block) tf = e1(SB.bl[bh].S2, SB.bl[b].S2, st[s].S2, J2p, J2, k2);
            else tf = e2(st[sh].S2, st[s].S2, SB.bl[b].S2, J2p, J2, k2) * ((Ps.fermion) ? nsign(SB.bl[b].N) : 1.0);
            if (tf == 0.0) continue;
            for (auto& cs : EB.subs[be]) {
              const int e = cs.b, tau = cs.s;
              std::vector<Trans> tq = transitions(Qe, Qe.block ? e : tau, EB.bl, nst);
              for (auto& tq1 : tq) {
                const int eh = Qe.block ? tq1.to : e;
                const int th = Qe.block ? tau : tq1.to;
                const int cpos_out = eidx.pos_in(eh, th, bep);
                if (cpos_out < 0) continue;
                const int cdim_out = EB.bl[eh].dim;
                double uf;
                if (Qe.block)
                  uf = e1(EB.bl[eh].S2, EB.bl[e].S2, st[tau].S2, J2p, J2, k2) *
                       ((Qe.fermion) ? nsign(st[tau].N) : 1.0) * tq1.s;
                else
                  uf = e2(st[th].S2, st[tau].S2, EB.bl[e].S2, J2p, J2, k2);
                if (uf == 0.0) continue;
                cplx coef = amp * tf * uf / std::sqrt((double)(k2 + 1) * (J2 + 1) * (J2p + 1));
                if (Ps.fermion) coef *= nsign(Nal);
                if (!Ps.block) coef *= tr.val;
                if (!Qe.block) coef *= tq1.val;
                const int mb = SB.bl[b].dim, me = EB.bl[e].dim;
                const int mbo = rdim_out, meo = cdim_out;
                TileKey ko{B_OUT, in_tile_off(ip, rpos_out, cpos_out), sb.rows[ip], mbo, meo};
                TileKey ki{B_OUT, in_tile_off(i, rs.pos, cs.pos), sb.rows[i], mb, me};
                const int64_t tin = in_tile_off(i, rs.pos, cs.pos);
                const int64_t tout = in_tile_off(ip, rpos_out, cpos_out);
                // operand descriptors
                int64_t a_off = 0, b_off = 0;
                int a_ld = 0, b_ld = 0;
                uint8_t b_op = OP_NN;
                if (Ps.block) {
                  const SubBlk& sub = Ps.lay->subs[tr.sub];
                  a_off = Ps.base + sub.off;
                  a_ld = sub.rows;
                }
                if (Qe.block) {
                  const SubBlk& sub = Qe.lay->subs[tq1.sub];
                  b_off = Qe.base + sub.off;
                  b_ld = sub.rows;
                  b_op = Qe.tilde ? OP_R : OP_T;
                }
                if (Ps.block && Qe.block) {
                  // two stages: X = A psi_i[rows, all cols]; out += c X[:, cols] op(B)
                  for (int adj = 0; adj < (herm_partner ? 2 : 1); ++adj) {
                    const int xi = adj ? ip : i;
                    const int xr = adj ? rpos_out : rs.pos;
                    const int xm = adj ? mb : mbo;      // X rows
                    const int xk = adj ? mbo : mb;
                    auto key = std::make_tuple(xi, xr, pid, tr.sub, adj);
                    auto itx = xoff.find(key);
                    int64_t xo;
                    if (itx == xoff.end()) {
                      xo = P.x_size;
                      P.x_size += (int64_t)xm * sb.cols[xi];
                      xoff[key] = xo;
                      TileKey kx{B_X, xo, xm, xm, sb.cols[xi]};
                      xs.add(kx, mk_term(1.0, Ps.buf, a_off, a_ld, adj ? OP_C : OP_NN, B_IN,
                                         sb.off[xi] + xr, sb.rows[xi], OP_NN, xk));
                      P.stage1.flops_exec += 8.0 * xm * sb.cols[xi] * xk;
                    } else {
                      xo = itx->second;
                    }
                    // reference association (mat_vec.f90:1749-1770): first GEMM
                    // counted on stage 1, second on stage 2
                    if (!adj) {
                      out.add(ko, mk_term(coef, B_X, xo + (int64_t)cs.pos * xm, xm, OP_NN, Qe.buf, b_off, b_ld, b_op, me));
                      P.stage1.flops_alg += 8.0 * (double)mbo * me * mb;
                      P.stage2.flops_alg += 8.0 * (double)mbo * meo * me;
                    } else {
                      out.add(ki, mk_term(std::conj(coef), B_X, xo + (int64_t)cpos_out * xm, xm, OP_NN, Qe.buf, b_off,
                                          b_ld, adj_op(b_op), meo));
                      P.stage1.flops_alg += 8.0 * (double)mb * meo * mbo;
                      P.stage2.flops_alg += 8.0 * (double)mb * me * meo;
                    }
                  }
                } else if (Ps.block) {
                  out.add(ko, mk_term(coef, Ps.buf, a_off, a_ld, OP_NN, B_IN, tin, sb.rows[i], OP_NN, mb));
                  P.stage2.flops_alg += 8.0 * mbo * me * mb;
                  if (herm_partner) {
                    out.add(ki, mk_term(std::conj(coef), Ps.buf, a_off, a_ld, OP_C, B_IN, tout, sb.rows[ip], OP_NN, mbo));
                    P.stage2.flops_alg += 8.0 * mb * me * mbo;
                  }
                } else if (Qe.block) {
                  out.add(ko, mk_term(coef, B_IN, tin, sb.rows[i], OP_NN, Qe.buf, b_off, b_ld, b_op, me));
                  P.stage2.flops_alg += 8.0 * mb * meo * me;
                  if (herm_partner) {
                    out.add(ki, mk_term(std::conj(coef), B_IN, tout, sb.rows[ip], OP_NN, Qe.buf, b_off, b_ld,
                                        adj_op(b_op), meo));
                    P.stage2.flops_alg += 8.0 * mb * me * meo;
                  }
                } else {
                  out.add(ko, mk_term(coef, B_IN, tin, sb.rows[i], OP_NN, 0, 0, 0, OP_NN, 0));
                  P.stage2.flops_alg += 8.0 * mb * me;
                  if (herm_partner) {
                    out.add(ki, mk_term(std::conj(coef), B_IN, tout, sb.rows[ip], OP_NN, 0, 0, 0, OP_NN, 0));
                    P.stage2.flops_alg += 8.0 * mb * me;
                  }
                }
              }
            }
          }
        }
      }
    }
  };

  // ---------- within-side scalar [block op x site op]^0 --------------------------
  auto within = [&](int side, const Piece& Pb, const Piece& Ps, cplx amp, bool herm_partner) {
    const int k2 = Pb.k2;
    const EnlBasis& EBs = side == 0 ? SB : EB;
    const SubIndex& ix = side == 0 ? sidx : eidx;
    for (int i = 0; i < (int)sb.pairs.size(); ++i) {
      const int al = side == 0 ? sb.pairs[i].first : sb.pairs[i].second;
      const int J2 = EBs.sec[al].S2;
      for (auto& rs : EBs.subs[al]) {
        const int b = rs.b, s = rs.s;
        for (auto& tb : transitions(Pb, b, EBs.bl, nst)) {
          for (auto& ts : transitions(Ps, s, EBs.bl, nst)) {
            const int pos_out = ix.pos_in(tb.to, ts.to, al);
            if (pos_out < 0) continue;
            double c = x12(EBs.bl[tb.to].S2, EBs.bl[b].S2, st[ts.to].S2, st[s].S2, J2, k2) * tb.s;
            if (c == 0.0) continue;
            if (Pb.fermion) c *= side == 0 ? nsign(EBs.bl[b].N) : nsign(st[s].N);
            if (side == 1 && (k2 & 1)) c = -c;
            cplx coef = amp * c * ts.val;
            const SubBlk& sub = Pb.lay->subs[tb.sub];
            const int64_t boff = Pb.base + sub.off;
            const int mi = EBs.bl[b].dim, mo = EBs.bl[tb.to].dim;
            if (side == 0) {
              for (auto& cs : EB.subs[sb.pairs[i].second]) {
                const int me = EB.bl[cs.b].dim;
                TileKey ko{B_OUT, in_tile_off(i, pos_out, cs.pos), sb.rows[i], mo, me};
                TileKey ki{B_OUT, in_tile_off(i, rs.pos, cs.pos), sb.rows[i], mi, me};
                out.add(ko, mk_term(coef, Pb.buf, boff, sub.rows, Pb.tilde ? OP_C : OP_NN, B_IN,
                                    in_tile_off(i, rs.pos, cs.pos), sb.rows[i], OP_NN, mi));
                P.stage2.flops_alg += 8.0 * mo * me * mi;
                if (herm_partner) {
                  out.add(ki, mk_term(std::conj(coef), Pb.buf, boff, sub.rows, Pb.tilde ? OP_NN : OP_C, B_IN,
                                      in_tile_off(i, pos_out, cs.pos), sb.rows[i], OP_NN, mo));
                  P.stage2.flops_alg += 8.0 * mi * me * mo;
                }
              }
            } else {
              for (auto& r2 : SB.subs[sb.pairs[i].first]) {
                const int ms = SB.bl[r2.b].dim;
                TileKey ko{B_OUT, in_tile_off(i, r2.pos, pos_out), sb.rows[i], ms, mo};
                TileKey ki{B_OUT, in_tile_off(i, r2.pos, rs.pos), sb.rows[i], ms, mi};
                const uint8_t bop = Pb.tilde ? OP_R : OP_T;
                out.add(ko, mk_term(coef, B_IN, in_tile_off(i, r2.pos, rs.pos), sb.rows[i], OP_NN, Pb.buf, boff,
                                    sub.rows, bop, mi));
                P.stage2.flops_alg += 8.0 * ms * mo * mi;
                if (herm_partner) {
                  out.add(ki, mk_term(std::conj(coef), B_IN, in_tile_off(i, r2.pos, pos_out), sb.rows[i], OP_NN,
                                      Pb.buf, boff, sub.rows, adj_op(bop), mo));
                  P.stage2.flops_alg += 8.0 * ms * mi * mo;
                }
              }
            }
          }
        }
      }
    }
  };

  // ---------- every output tile exists (zero-filled when no term lands) ---------
  for (int i = 0; i < (int)sb.pairs.size(); ++i)
    for (auto& rs : SB.subs[sb.pairs[i].first])
      for (auto& cs : EB.subs[sb.pairs[i].second])
        out.ensure(TileKey{B_OUT, in_tile_off(i, rs.pos, cs.pos), sb.rows[i], SB.bl[rs.b].dim, EB.bl[cs.b].dim});

  // ---------- couplings ----------------------------------------------------------
  const int psc = g.sys_centre(), pec = g.env_centre();
  auto sys_block_piece = [&](int kind, int64_t base, uint8_t buf) {
    Piece p;
    p.block = true; p.side = 0; p.kind = kind; p.tilde = false; p.buf = buf; p.base = base;
    p.lay = &sys.lay[kind]; p.k2 = op_k2(kind); p.fermion = op_fermion(kind);
    return p;
  };
  auto env_block_piece = [&](int kind, int64_t base, uint8_t buf, bool tilde) {
    Piece p;
    p.block = true; p.side = 1; p.kind = kind; p.tilde = tilde; p.buf = buf; p.base = base;
    p.lay = &env.lay[kind]; p.k2 = op_k2(kind); p.fermion = op_fermion(kind);
    return p;
  };
  auto site_piece = [&](int side, int kind, bool tilde) {
    Piece p;
    p.block = false; p.side = side; p.kind = kind; p.k2 = op_k2(kind); p.fermion = op_fermion(kind);
    if (tilde) p.site = tilde_fn();
    else p.site = site_fn(kind);
    return p;
  };
  if (only) {
    auto piece = [&](const MeasurePiece& q, int side) {
      if (!q.block) return site_piece(side, q.kind, q.tilde);
      Piece p;
      p.block = true; p.side = side; p.kind = q.kind; p.tilde = q.tilde; p.buf = q.buf; p.base = q.base;
      p.lay = q.lay; p.k2 = op_k2(q.kind); p.fermion = op_fermion(q.kind);
      return p;
    };
    if (only->within) within(only->side, piece(only->s, only->side), piece(only->e, only->side), only->amp, false);
    else cross(piece(only->s, 0), piece(only->e, 1), only->amp, false);
  } else {
    // ---------- block Hamiltonians (block_vec_dia / vec_block_dia) -----------------
    for (int i = 0; i < (int)sb.pairs.size(); ++i) {
      const int al = sb.pairs[i].first, be = sb.pairs[i].second;
      for (auto& rs : SB.subs[al])
        for (auto& cs : EB.subs[be]) {
          const int mb = SB.bl[rs.b].dim, me = EB.bl[cs.b].dim;
          TileKey k{B_OUT, in_tile_off(i, rs.pos, cs.pos), sb.rows[i], mb, me};
          const int64_t tin = in_tile_off(i, rs.pos, cs.pos);
          out.add(k, mk_term(1.0 / std::sqrt((double)SB.bl[rs.b].S2 + 1), B_SYS, sys.h_off[rs.b], mb, OP_NN, B_IN, tin,
                             sb.rows[i], OP_NN, mb));
          out.add(k, mk_term(1.0 / std::sqrt((double)EB.bl[cs.b].S2 + 1), B_IN, tin, sb.rows[i], OP_NN, B_ENV,
                             env.h_off[cs.b], me, OP_T, me));
          P.stage2.flops_alg += 8.0 * mb * me * (mb + me);
          // on-site U of the two centre sites (site_vec_site_dia with st_double)
          if (m.kind == HUBBARD && m.U != 0.0) {
            double u = 0.0;
            if (rs.s == 2) u += m.U;
            if (cs.s == 2) u += m.U;
            if (u != 0.0) {
              out.add(k, mk_term(u, B_IN, tin, sb.rows[i], OP_NN, 0, 0, 0, OP_NN, 0));
              P.stage2.flops_alg += 8.0 * mb * me;
            }
          }
        }
    }
    // bonds of every stored boundary site against the centre sites / env sites
    struct Bond { int si; int ej; Coupling c; };
    std::vector<std::pair<int, Coupling>> sb_sc, sb_ec, eb_sc, eb_ec;   // (site idx, coupling)
    std::vector<Bond> bb;
    for (size_t a = 0; a < sys.sites.size(); ++a) {
      Coupling c;
      const int p = g.sys_pos(sys.sites[a]);
      if (g.coupling(p, psc, c)) sb_sc.push_back({(int)a, c});
      if (g.coupling(p, pec, c)) sb_ec.push_back({(int)a, c});
      for (size_t e = 0; e < env.sites.size(); ++e)
        if (g.coupling(p, g.env_pos(env.sites[e]), c)) bb.push_back({(int)a, (int)e, c});
    }
    for (size_t e = 0; e < env.sites.size(); ++e) {
      Coupling c;
      const int p = g.env_pos(env.sites[e]);
      if (g.coupling(psc, p, c)) eb_sc.push_back({(int)e, c});
      if (g.coupling(pec, p, c)) eb_ec.push_back({(int)e, c});
    }
    Coupling cc;
    const bool has_ss = g.coupling(psc, pec, cc);

    const double r2 = std::sqrt(2.0), r3 = std::sqrt(3.0);
    const int nk = n_op_kinds(m.kind);
    // helper to build a merge and return its piece
    auto merge_sys = [&](int kind, const std::vector<std::pair<int, cplx>>& parts) {
      int id = mb.add(0, kind, parts, sys.lay[kind]);
      return sys_block_piece(kind, P.merges[id].off, B_MRG);
    };
    auto merge_env = [&](int kind, const std::vector<std::pair<int, cplx>>& parts, bool tilde) {
      int id = mb.add(1, kind, parts, env.lay[kind]);
      return env_block_piece(kind, P.merges[id].off, B_MRG, tilde);
    };

    // sys block <-> sys centre (block_site_vec_ndia / _dia)
    if (!sb_sc.empty()) {
      std::vector<std::pair<int, cplx>> ph, pj, pv;
      for (auto& x : sb_sc) {
        if (x.second.hop != 0.0) ph.push_back({x.first, x.second.hop});
        if (x.second.J != 0.0) pj.push_back({x.first, x.second.J});
        if (x.second.V != 0.0) pv.push_back({x.first, x.second.V});
      }
      if (!ph.empty()) within(0, merge_sys(OP_CDAG, ph), site_piece(0, OP_CDAG, true), r2, true);
      if (nk > 1 && !pj.empty()) within(0, merge_sys(OP_S, pj), site_piece(0, OP_S, false), -r3, false);
      if (nk > 1 && !pv.empty()) within(0, merge_sys(OP_N, pv), site_piece(0, OP_N, false), 1.0, false);
    }
    // env centre <-> env block (vec_block_site_ndia / _dia); fermion order centre first
    if (!eb_ec.empty()) {
      std::vector<std::pair<int, cplx>> ph, pj, pv;
      for (auto& x : eb_ec) {
        if (x.second.hop != 0.0) ph.push_back({x.first, std::conj(x.second.hop)});
        if (x.second.J != 0.0) pj.push_back({x.first, x.second.J});
        if (x.second.V != 0.0) pv.push_back({x.first, x.second.V});
      }
      if (!ph.empty()) within(1, merge_env(OP_CDAG, ph, true), site_piece(1, OP_CDAG, false), r2, true);
      if (nk > 1 && !pj.empty()) within(1, merge_env(OP_S, pj, false), site_piece(1, OP_S, false), -r3, false);
      if (nk > 1 && !pv.empty()) within(1, merge_env(OP_N, pv, false), site_piece(1, OP_N, false), 1.0, false);
    }
    // sys centre <-> env centre (site_vec_site_ndia / _dia)
    if (has_ss) {
      if (cc.hop != 0.0) cross(site_piece(0, OP_CDAG, false), site_piece(1, OP_CDAG, true), r2 * cc.hop, true);
      if (nk > 1 && cc.J != 0.0) cross(site_piece(0, OP_S, false), site_piece(1, OP_S, false), -r3 * cc.J, false);
      if (nk > 1 && cc.V != 0.0) cross(site_piece(0, OP_N, false), site_piece(1, OP_N, false), cc.V, false);
    }
    // sys block <-> env centre (block_vec_site_ndia / _dia)
    if (!sb_ec.empty()) {
      std::vector<std::pair<int, cplx>> ph, pj, pv;
      for (auto& x : sb_ec) {
        if (x.second.hop != 0.0) ph.push_back({x.first, x.second.hop});
        if (x.second.J != 0.0) pj.push_back({x.first, x.second.J});
        if (x.second.V != 0.0) pv.push_back({x.first, x.second.V});
      }
      if (!ph.empty()) cross(merge_sys(OP_CDAG, ph), site_piece(1, OP_CDAG, true), r2, true);
      if (nk > 1 && !pj.empty()) cross(merge_sys(OP_S, pj), site_piece(1, OP_S, false), -r3, false);
      if (nk > 1 && !pv.empty()) cross(merge_sys(OP_N, pv), site_piece(1, OP_N, false), 1.0, false);
    }
    // sys centre <-> env block (site_vec_block_ndia / _dia)
    if (!eb_sc.empty()) {
      std::vector<std::pair<int, cplx>> ph, pj, pv;
      for (auto& x : eb_sc) {
        if (x.second.hop != 0.0) ph.push_back({x.first, std::conj(x.second.hop)});
        if (x.second.J != 0.0) pj.push_back({x.first, x.second.J});
        if (x.second.V != 0.0) pv.push_back({x.first, x.second.V});
      }
      if (!ph.empty()) cross(site_piece(0, OP_CDAG, false), merge_env(OP_CDAG, ph, true), r2, true);
      if (nk > 1 && !pj.empty()) cross(site_piece(0, OP_S, false), merge_env(OP_S, pj, false), -r3, false);
      if (nk > 1 && !pv.empty()) cross(site_piece(0, OP_N, false), merge_env(OP_N, pv, false), 1.0, false);
    }
    // block <-> block (block_vec_block_ndia / _dia): one term per boundary site
    // on the side with more cross-bonded sites, partners merged on the other
    // side (hmerge's inb / coup_bb rule, code/mat_vec.f90:3739-3771)
    if (!bb.empty()) {
      std::set<int> ssites, esites;
      for (auto& x : bb) { ssites.insert(x.si); esites.insert(x.ej); }
      const bool per_env = esites.size() < ssites.size();
      if (!per_env) {
        for (int si : ssites) {
          std::vector<std::pair<int, cplx>> ph, pj, pv;
          for (auto& x : bb)
            if (x.si == si) {
              if (x.c.hop != 0.0) ph.push_back({x.ej, std::conj(x.c.hop)});
              if (x.c.J != 0.0) pj.push_back({x.ej, x.c.J});
              if (x.c.V != 0.0) pv.push_back({x.ej, x.c.V});
            }
          if (!ph.empty())
            cross(sys_block_piece(OP_CDAG, sys.op_off[si][OP_CDAG], B_SYS), merge_env(OP_CDAG, ph, true), r2, true);
          if (nk > 1 && !pj.empty())
            cross(sys_block_piece(OP_S, sys.op_off[si][OP_S], B_SYS), merge_env(OP_S, pj, false), -r3, false);
          if (nk > 1 && !pv.empty())
            cross(sys_block_piece(OP_N, sys.op_off[si][OP_N], B_SYS), merge_env(OP_N, pv, false), 1.0, false);
        }
      } else {
        for (int ej : esites) {
          std::vector<std::pair<int, cplx>> ph, pj, pv;
          for (auto& x : bb)
            if (x.ej == ej) {
              if (x.c.hop != 0.0) ph.push_back({x.si, x.c.hop});
              if (x.c.J != 0.0) pj.push_back({x.si, x.c.J});
              if (x.c.V != 0.0) pv.push_back({x.si, x.c.V});
            }
          if (!ph.empty())
            cross(merge_sys(OP_CDAG, ph), env_block_piece(OP_CDAG, env.op_off[ej][OP_CDAG], B_ENV, true), r2, true);
          if (nk > 1 && !pj.empty())
            cross(merge_sys(OP_S, pj), env_block_piece(OP_S, env.op_off[ej][OP_S], B_ENV, false), -r3, false);
          if (nk > 1 && !pv.empty())
            cross(merge_sys(OP_N, pv), env_block_piece(OP_N, env.op_off[ej][OP_N], B_ENV, false), 1.0, false);
        }
      }
    }
  }

  xs.emit(P.stage1, 0);
  out.emit(P.stage2, 0);
  // executed flops of stage 2 single-stage terms equal their algorithmic count
  double fe = 0.0;
  for (size_t ti = 0; ti < P.stage2.tiles.size(); ++ti) {
    const Tile& T = P.stage2.tiles[ti];
    for (int k = T.t0; k < T.t1; ++k) fe += P.stage2.terms[k].K > 0 ? 8.0 * T.M * T.N * P.stage2.terms[k].K : 8.0 * T.M * T.N;
  }
  P.stage2.flops_exec = fe;
  // compulsory bytes: psi in + out and every operator sub-block read once
  {
    std::set<std::pair<int, int64_t>> seen;
    double elems = 0.0;
    auto visit = [&](const TaskList& tl) {
      for (const Tile& T : tl.tiles)
        for (int k = T.t0; k < T.t1; ++k) {
          const Term& t = tl.terms[k];
          if (t.K == 0) continue;
          if (t.a_buf == B_SYS || t.a_buf == B_ENV || t.a_buf == B_MRG)
            if (seen.insert({t.a_buf, t.a_off}).second) elems += (double)T.M * t.K;
          if (t.b_buf == B_SYS || t.b_buf == B_ENV || t.b_buf == B_MRG)
            if (seen.insert({t.b_buf, t.b_off}).second) elems += (double)t.K * T.N;
        }
    };
    visit(P.stage1);
    visit(P.stage2);
    P.bytes_alg = 16.0 * (2.0 * (double)sb.dim + elems);
  }
  const BlockDesc* blocks[2] = {&sys, &env};
  uint8_t bufs[2] = {B_SYS, B_ENV};
  compile_merges(P.merges, blocks, P.merge_tasks, bufs);
  return P;
}

}  // namespace hh
