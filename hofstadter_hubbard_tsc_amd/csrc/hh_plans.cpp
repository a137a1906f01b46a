// Host-only plans behind the C ABI (include/hhdmrg.h): the task lists of a
// superblock H.psi or of a block growth compiled from sector tables alone,
// with no device, for callers that run the tasks themselves (hh_run_tasks)
// or inspect the layouts.
#include <string>
#include <vector>

#include "../../include/hhdmrg.h"
#include "hh_core.hpp"
#include "hh_driver.hpp"
#include "hh_runtime.hpp"

using namespace hh;


struct hh_plan {
  BlockDesc blk[3];
  MatvecPlan mv;
  GrowPlan gp;
  bool is_grow = false;
  std::string err;
};

namespace {
BlockDesc make_desc(const ModelParams& m, int side, int len, int nsec, const int32_t* N, const int32_t* S2,
                    const int32_t* dim, int nsites, const int32_t* sites) {
  BlockDesc d;
  d.side = side;
  d.length = len;
  for (int i = 0; i < nsec; ++i) d.sec.push_back({N[i], S2[i], dim[i]});
  for (int i = 0; i < nsites; ++i) d.sites.push_back(sites[i]);
  layout_block(d, n_op_kinds(m.kind));
  return d;
}
}  // namespace

extern "C" {

int hh_plan_matvec(const hh_model_params* mp, int sys_len, int sys_nsec, const int32_t* sys_N, const int32_t* sys_S2,
                   const int32_t* sys_dim, int sys_nsites, const int32_t* sys_sites, int env_len, int env_nsec,
                   const int32_t* env_N, const int32_t* env_S2, const int32_t* env_dim, int env_nsites,
                   const int32_t* env_sites, int N, hh_plan** out) {
  try {
    need(mp != nullptr, "null model");
    check_gauge(mp);
    ModelParams m = to_model(mp);
    auto* p = new hh_plan();
    p->blk[0] = make_desc(m, 0, sys_len, sys_nsec, sys_N, sys_S2, sys_dim, sys_nsites, sys_sites);
    p->blk[1] = make_desc(m, 1, env_len, env_nsec, env_N, env_S2, env_dim, env_nsites, env_sites);
    Geometry g = make_geometry(m, sys_len, env_len);
    p->mv = compile_matvec(m, g, p->blk[0], p->blk[1], N);
    *out = p;
    return 0;
  } catch (const std::exception& e) {
    create_error() = e.what();
    return 1;
  }
}

int hh_plan_grow(const hh_model_params* mp, int side, int length, int nsec, const int32_t* Nq, const int32_t* S2q,
                 const int32_t* dimq, int nsites, const int32_t* sites, int ns, int ne, int nkept,
                 const int32_t* kept_a, const int32_t* kept_n, const int64_t* u_off, const int32_t* u_ld,
                 hh_plan** out) {
  try {
    need(mp != nullptr, "null model");
    check_gauge(mp);
    ModelParams m = to_model(mp);
    auto* p = new hh_plan();
    p->is_grow = true;
    p->blk[side] = make_desc(m, side, length, nsec, Nq, S2q, dimq, nsites, sites);
    Geometry g = make_geometry(m, ns, ne);
    std::vector<KeptSector> kept;
    for (int i = 0; i < nkept; ++i) kept.push_back({kept_a[i], kept_n[i], u_off[i], u_ld[i]});
    p->gp = compile_grow(m, g, p->blk[side], kept);
    p->blk[2] = p->gp.out;
    *out = p;
    return 0;
  } catch (const std::exception& e) {
    create_error() = e.what();
    return 1;
  }
}

int hh_plan_destroy(hh_plan* p) {
  delete p;
  return 0;
}

int hh_plan_tasks(const hh_plan* p, int stage, const void** tiles, int32_t* ntiles, const void** terms,
                  int32_t* nterms, int64_t* scratch) {
  if (!p || stage < 0 || stage > 2) return 1;
  const TaskList* tl;
  int64_t sc;
  if (p->is_grow) {
    tl = stage == 0 ? &p->gp.merge_tasks : stage == 1 ? &p->gp.stage1 : &p->gp.stage2;
    sc = stage == 0 ? p->gp.merged_size : p->gp.x_size;
  } else {
    tl = stage == 0 ? &p->mv.merge_tasks : stage == 1 ? &p->mv.stage1 : &p->mv.stage2;
    sc = stage == 0 ? p->mv.merged_size : p->mv.x_size;
  }
  if (tiles) *tiles = tl->tiles.data();
  if (ntiles) *ntiles = (int32_t)tl->tiles.size();
  if (terms) *terms = tl->terms.data();
  if (nterms) *nterms = (int32_t)tl->terms.size();
  if (scratch) *scratch = sc;
  return 0;
}

int hh_plan_block(const hh_plan* p, int which, int64_t* arena, int32_t* nsec, int32_t* N, int32_t* S2, int32_t* dim,
                  int64_t* h_off, int32_t* nsites, int32_t* sites) {
  if (!p || which < 0 || which > 2) return 1;
  const BlockDesc& d = p->blk[which];
  if (arena) *arena = d.size;
  if (nsec) *nsec = (int32_t)d.sec.size();
  for (size_t i = 0; i < d.sec.size(); ++i) {
    if (N) N[i] = d.sec[i].N;
    if (S2) S2[i] = d.sec[i].S2;
    if (dim) dim[i] = d.sec[i].dim;
    if (h_off) h_off[i] = d.h_off[i];
  }
  if (nsites) *nsites = (int32_t)d.sites.size();
  if (sites)
    for (size_t i = 0; i < d.sites.size(); ++i) sites[i] = d.sites[i];
  return 0;
}

int hh_plan_layout(const hh_plan* p, int which, int kind, int32_t* nsub, int32_t* bra, int32_t* ket, int64_t* off) {
  if (!p || which < 0 || which > 2 || kind < 0 || kind > 2) return 1;
  const OpLayout& L = p->blk[which].lay[kind];
  if (nsub) *nsub = (int32_t)L.subs.size();
  for (size_t i = 0; i < L.subs.size(); ++i) {
    if (bra) bra[i] = L.subs[i].bra;
    if (ket) ket[i] = L.subs[i].ket;
    if (off) off[i] = L.subs[i].off;
  }
  return 0;
}

int hh_plan_op_offset(const hh_plan* p, int which, int site_index, int kind, int64_t* off) {
  if (!p || which < 0 || which > 2) return 1;
  const BlockDesc& d = p->blk[which];
  if (site_index < 0 || site_index >= (int)d.sites.size() || kind < 0 || kind > 2) return 1;
  *off = d.op_off[site_index][kind];
  return *off < 0 ? 1 : 0;
}

int hh_plan_enlarged(const hh_plan* p, int side, int32_t* nsec, int32_t* N, int32_t* S2, int32_t* dim, int32_t* nsub,
                     int32_t* subs) {
  if (!p || p->is_grow || (side != 0 && side != 1) || !nsec) return 1;
  const EnlBasis& e = side == 0 ? p->mv.sbs : p->mv.ebs;
  *nsec = (int32_t)e.sec.size();
  int32_t k = 0;
  for (size_t a = 0; a < e.sec.size(); ++a) {
    if (N) N[a] = e.sec[a].N;
    if (S2) S2[a] = e.sec[a].S2;
    if (dim) dim[a] = e.sec[a].dim;
    for (const SubSector& q : e.subs[a]) {
      if (subs) {
        int32_t* r = subs + 5 * (size_t)k;
        r[0] = (int32_t)a;
        r[1] = q.b;
        r[2] = q.s;
        r[3] = q.pos;
        r[4] = q.dim;
      }
      ++k;
    }
  }
  if (nsub) *nsub = k;
  return 0;
}

int hh_plan_super(const hh_plan* p, int64_t* dim, int32_t* npairs, int32_t* sys_sec, int32_t* env_sec, int32_t* rows,
                  int32_t* cols, int64_t* off) {
  if (!p || p->is_grow) return 1;
  const SuperBasis& S = p->mv.sb;
  if (dim) *dim = S.dim;
  if (npairs) *npairs = (int32_t)S.pairs.size();
  for (size_t i = 0; i < S.pairs.size(); ++i) {
    if (sys_sec) sys_sec[i] = S.pairs[i].first;
    if (env_sec) env_sec[i] = S.pairs[i].second;
    if (rows) rows[i] = S.rows[i];
    if (cols) cols[i] = S.cols[i];
    if (off) off[i] = S.off[i];
  }
  return 0;
}

}  // extern "C"
