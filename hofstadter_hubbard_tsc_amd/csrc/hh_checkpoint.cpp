// Checkpoint / restart of the sweep driver in this library's own binary
// layout (hh_driver_save / hh_driver_load).
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "../../include/hhdmrg.h"
#include "hh_core.hpp"
#include "hh_driver.hpp"
#include "hh_engine.hpp"
#include "hh_runtime.hpp"

namespace hh {

// ----------------------------------------------------------------------------
// checkpoint / restart of a driver: what model_to_disk, truns_to_disk and
// wave_to_disk (code/space_info.f90:1406-1762) keep for restart.dat's resume
// (code/quantum.f90:43-47, :207-253) -- every stored block with its
// operators, every truncation, the carried wave function and the step
// counter -- in one binary file of this library's own layout.
// ----------------------------------------------------------------------------
namespace {
constexpr char CK_MAGIC[8] = {'H', 'H', 'D', 'M', 'R', 'G', 'C', 'K'};
// version 3 adds the levels record (the levels setting and the vectors of the
// upper levels); version 2 files (one level) are still read
constexpr uint32_t CK_VERSION = 3;

struct CkOut {
  FILE* f;
  void raw(const void* p, size_t n) { need(std::fwrite(p, 1, n, f) == n, "checkpoint write failed"); }
  template <typename V> void put(V v) { raw(&v, sizeof(V)); }
  void sectors(const std::vector<Sector>& v) {
    put<uint32_t>((uint32_t)v.size());
    for (auto& q : v) {
      put<int32_t>(q.N);
      put<int32_t>(q.S2);
      put<int32_t>(q.dim);
    }
  }
  void device(hh_ctx* ctx, const void* dev, int64_t elems) {
    put<int64_t>(elems);
    if (elems <= 0) return;
    std::vector<double2> h((size_t)elems);
    check_hip(hipMemcpyAsync(h.data(), dev, sizeof(double2) * h.size(), hipMemcpyDeviceToHost, ctx->stream), "d2h");
    check_hip(hipStreamSynchronize(ctx->stream), "d2h");
    raw(h.data(), sizeof(double2) * h.size());
  }
};
struct CkIn {
  FILE* f;
  void raw(void* p, size_t n) { need(std::fread(p, 1, n, f) == n, "checkpoint file is truncated"); }
  template <typename V> V get() {
    V v;
    raw(&v, sizeof(V));
    return v;
  }
  std::vector<Sector> sectors() {
    const uint32_t n = get<uint32_t>();
    need(n < (1u << 24), "checkpoint file is corrupt (sector count)");
    std::vector<Sector> v(n);
    for (auto& q : v) {
      q.N = get<int32_t>();
      q.S2 = get<int32_t>();
      q.dim = get<int32_t>();
      need(q.dim >= 0 && q.S2 >= 0, "checkpoint file is corrupt (sector)");
    }
    return v;
  }
  // elements into a device buffer of at least `expect` elements (expect < 0: any)
  int64_t device(hh_ctx* ctx, DevMem& dst, int64_t expect) {
    const int64_t n = get<int64_t>();
    need(n >= 0 && (expect < 0 || n == expect), "checkpoint file is corrupt (array length)");
    dst.alloc(sizeof(double2) * (size_t)std::max<int64_t>(n, 1));
    if (n > 0) {
      std::vector<double2> h((size_t)n);
      raw(h.data(), sizeof(double2) * h.size());
      check_hip(hipMemcpyAsync(dst.p, h.data(), sizeof(double2) * h.size(), hipMemcpyHostToDevice, ctx->stream),
                "h2d");
      check_hip(hipStreamSynchronize(ctx->stream), "h2d");
    }
    return n;
  }
};

void put_model(CkOut& o, const ModelParams& m) {
  o.put<int32_t>(m.kind);
  o.put<int32_t>(m.Lx);
  o.put<int32_t>(m.Ly);
  o.put<int32_t>(m.lattice);
  o.put<int32_t>(m.periodic_y ? 1 : 0);
  o.put<int32_t>(m.n_elec);
  o.put<double>(m.t);
  o.put<double>(m.U);
  o.put<double>(m.J);
  o.put<double>(m.flux);
  o.put<int32_t>(m.gauge);
  o.put<double>(m.theta);
}

}  // namespace

void save_driver(Driver& d, const char* path) {
  hh_ctx* ctx = d.ctx;
  std::unique_ptr<FILE, int (*)(FILE*)> f(std::fopen(path, "wb"), &std::fclose);
  need(f != nullptr, "cannot open the checkpoint file for writing");
  CkOut o{f.get()};
  o.raw(CK_MAGIC, sizeof(CK_MAGIC));
  o.put<uint32_t>(CK_VERSION);
  put_model(o, ctx->model);
  o.put<hh_schedule>(d.S);
  o.put<uint64_t>(d.step);
  o.put<int64_t>(d.psi_dim);
  o.device(ctx, d.psi.p, d.psi.p ? d.psi_dim : 0);
  // levels record: the setting, then the upper levels' vectors when they
  // belong to the carried superblock
  o.put<int32_t>(d.levels);
  const bool lv = d.lpsi_dim == d.psi_dim && d.psi_dim > 0 && (int)d.lpsi.size() == d.levels - 1;
  o.put<int32_t>(lv ? (int32_t)d.lpsi.size() : 0);
  if (lv)
    for (const auto& v : d.lpsi) o.device(ctx, v->p, d.psi_dim);
  for (int side = 0; side < 2; ++side) {
    const auto& blocks = side == 0 ? d.sysb : d.envb;
    o.put<uint32_t>((uint32_t)blocks.size());
    for (auto& kv : blocks) {
      const BlockDesc& b = kv.second->d;
      o.put<int32_t>(kv.first);
      o.put<int32_t>(b.length);
      o.sectors(b.sec);
      o.put<uint32_t>((uint32_t)b.sites.size());
      for (int k : b.sites) o.put<int32_t>(k);
      o.device(ctx, kv.second->arena.p, b.size);
    }
    const auto& truns = side == 0 ? d.syst : d.envt;
    o.put<uint32_t>((uint32_t)truns.size());
    for (auto& kv : truns) {
      const hh_trunc* t = kv.second.get();
      o.put<int32_t>(kv.first);
      o.sectors(t->eb.bl);
      o.put<uint32_t>((uint32_t)t->kept.size());
      for (auto& k : t->kept) {
        o.put<int32_t>(k.a);
        o.put<int32_t>(k.keep);
        o.put<int64_t>(k.u_off);
        o.put<int32_t>(k.u_ld);
      }
      o.put<uint32_t>((uint32_t)t->kept_per_rho.size());
      for (int k : t->kept_per_rho) o.put<int32_t>(k);
      o.put<int32_t>(t->num_state);
      o.put<double>(t->err);
      o.put<double>(t->entropy);
      o.device(ctx, t->evec.p, t->evec.p ? (int64_t)(t->evec.bytes / sizeof(double2)) : 0);
    }
  }
  need(std::fflush(f.get()) == 0, "checkpoint write failed");
}

void load_driver(Driver& d, const char* path) {
  hh_ctx* ctx = d.ctx;
  std::unique_ptr<FILE, int (*)(FILE*)> f(std::fopen(path, "rb"), &std::fclose);
  need(f != nullptr, "cannot open the checkpoint file");
  CkIn in{f.get()};
  char magic[8];
  in.raw(magic, sizeof(magic));
  need(std::memcmp(magic, CK_MAGIC, sizeof(magic)) == 0, "not a checkpoint of this library");
  const uint32_t version = in.get<uint32_t>();
  need(version == 2 || version == CK_VERSION, "checkpoint version not supported");
  const ModelParams& m = ctx->model;
  const int kind = in.get<int32_t>(), Lx = in.get<int32_t>(), Ly = in.get<int32_t>(), lattice = in.get<int32_t>();
  const int periodic = in.get<int32_t>(), n_elec = in.get<int32_t>();
  const double t = in.get<double>(), U = in.get<double>(), J = in.get<double>(), flux = in.get<double>();
  const int gauge = in.get<int32_t>();
  const double theta = in.get<double>();
  need(kind == m.kind && Lx == m.Lx && Ly == m.Ly && lattice == m.lattice && periodic == (m.periodic_y ? 1 : 0) &&
           n_elec == m.n_elec && t == m.t && U == m.U && J == m.J && flux == m.flux && gauge == m.gauge &&
           theta == m.theta,
       "the checkpoint was written for another model than the context's");
  Driver fresh;   // build the whole state aside; the driver is replaced only on success
  fresh.ctx = ctx;
  fresh.S = in.get<hh_schedule>();
  check_schedule(fresh.S);
  fresh.step = in.get<uint64_t>();
  fresh.psi_dim = in.get<int64_t>();
  in.device(ctx, fresh.psi, fresh.psi_dim > 0 ? fresh.psi_dim : 0);
  if (version >= 3) {
    fresh.levels = in.get<int32_t>();
    need(fresh.levels >= 1 && fresh.levels <= MAX_LEVELS && 2 * (fresh.levels - 1) < fresh.S.max_iter &&
             (fresh.levels == 1 || fresh.S.solver == 0),
         "checkpoint file is corrupt (levels)");
    const int nl = in.get<int32_t>();
    need(nl == 0 || (nl == fresh.levels - 1 && fresh.psi_dim > 0), "checkpoint file is corrupt (level vectors)");
    for (int l = 0; l < nl; ++l) {
      fresh.lpsi.emplace_back(new DevMem());
      in.device(ctx, *fresh.lpsi.back(), fresh.psi_dim);
    }
    fresh.lpsi_dim = nl > 0 ? fresh.psi_dim : 0;
  }
  const auto st = site_sectors(m.kind);
  for (int side = 0; side < 2; ++side) {
    auto& blocks = side == 0 ? fresh.sysb : fresh.envb;
    const uint32_t nb = in.get<uint32_t>();
    for (uint32_t i = 0; i < nb; ++i) {
      const int key = in.get<int32_t>(), length = in.get<int32_t>();
      need(key == length && length >= 1 && length <= m.nsite(), "checkpoint file is corrupt (block length)");
      std::vector<Sector> sec = in.sectors();
      const uint32_t ns = in.get<uint32_t>();
      need(ns <= (uint32_t)m.nsite(), "checkpoint file is corrupt (sites)");
      std::vector<int> sites(ns);
      for (auto& k : sites) k = in.get<int32_t>();
      std::unique_ptr<hh_block> b(new_block(ctx, side, length, sec, sites));
      in.device(ctx, b->arena, b->d.size);
      blocks[key] = std::move(b);
    }
    auto& truns = side == 0 ? fresh.syst : fresh.envt;
    const uint32_t nt = in.get<uint32_t>();
    for (uint32_t i = 0; i < nt; ++i) {
      const int key = in.get<int32_t>();
      need(key >= 1 && key <= m.nsite(), "checkpoint file is corrupt (truncation length)");
      std::unique_ptr<hh_trunc> tr(new hh_trunc());
      tr->side = side;
      tr->eb = enlarge_basis(in.sectors(), st);
      const uint32_t nk = in.get<uint32_t>();
      need(nk <= tr->eb.sec.size(), "checkpoint file is corrupt (kept sectors)");
      tr->kept.resize(nk);
      int64_t need_elems = 0;
      for (auto& k : tr->kept) {
        k.a = in.get<int32_t>();
        k.keep = in.get<int32_t>();
        k.u_off = in.get<int64_t>();
        k.u_ld = in.get<int32_t>();
        need(k.a >= 0 && k.a < (int)tr->eb.sec.size() && k.keep >= 0 && k.keep <= tr->eb.sec[k.a].dim &&
                 k.u_ld >= tr->eb.sec[k.a].dim && k.u_off >= 0,
             "checkpoint file is corrupt (kept sector)");
        need_elems = std::max<int64_t>(need_elems, k.u_off + (int64_t)k.u_ld * k.keep);
      }
      const uint32_t nr = in.get<uint32_t>();
      need(nr < (1u << 24), "checkpoint file is corrupt (rho blocks)");
      tr->kept_per_rho.resize(nr);
      for (auto& k : tr->kept_per_rho) k = in.get<int32_t>();
      tr->num_state = in.get<int32_t>();
      tr->err = in.get<double>();
      tr->entropy = in.get<double>();
      const int64_t n = in.device(ctx, tr->evec, -1);
      need(n >= need_elems, "checkpoint file is corrupt (eigenvector arena)");
      truns[key] = std::move(tr);
    }
  }
  need(std::fgetc(f.get()) == EOF, "checkpoint file is corrupt (data after the last record)");
  d.S = fresh.S;
  d.step = fresh.step;
  d.psi_dim = fresh.psi_dim;
  std::swap(d.psi.p, fresh.psi.p);
  std::swap(d.psi.bytes, fresh.psi.bytes);
  d.sysb.swap(fresh.sysb);
  d.envb.swap(fresh.envb);
  d.syst.swap(fresh.syst);
  d.envt.swap(fresh.envt);
  d.prev_sb.reset();   // the next step starts from the carried wave function
  d.levels = fresh.levels;
  d.lpsi.swap(fresh.lpsi);
  d.lpsi_dim = fresh.lpsi_dim;
}
}  // namespace hh
