// The reference's own file formats behind the C ABI (include/hhdmrg.h):
// the wavefunction, truncation and basis files of code/space_info.f90
// (wave_to_disk / wave_from_disk, truns_to_disk / truns_from_disk via
// block_to_disk / block_from_disk, basis_to_disk / basis_from_disk), written
// and read on the host around device-resident psi vectors and truncations.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <initializer_list>
#include <memory>
#include <string>
#include <vector>

#include "../../include/hhdmrg.h"
#include "hh_core.hpp"
#include "hh_engine.hpp"
#include "hh_runtime.hpp"

using namespace hh;

extern "C" {

// ---- wave_to_disk / wave_from_disk (code/space_info.f90:1406-1468) --------
// The reference's wavefunction file: Fortran sequential unformatted records,
// each framed by its byte length as a 4-byte integer before and after.
// Records: sys_len, env_len, num, dim (default integers; the lengths are
// those of the enlarged blocks), then per super sector in super-basis order
// an empty record, (N_sys, 2S_sys), (N_env, 2S_env), (sys_dim, env_dim) and
// the sys_dim x env_dim complex*16 block, column-major.
namespace {
struct FortranOut {
  FILE* f;
  const char* what;   // "wave file" or "truncation file", for the messages
  void record(const void* p, size_t n) {
    need(n < (size_t(1) << 31), std::string(what) + ": record too large for 4-byte record markers");
    const int32_t m = (int32_t)n;
    need(std::fwrite(&m, 4, 1, f) == 1 && (n == 0 || std::fwrite(p, 1, n, f) == n) && std::fwrite(&m, 4, 1, f) == 1,
         std::string("cannot write the ") + what);
  }
  void ints(std::initializer_list<int64_t> v) {
    std::vector<int32_t> a;
    for (int64_t x : v) a.push_back((int32_t)x);
    record(a.data(), 4 * a.size());
  }
};
struct FortranIn {
  FILE* f;
  const char* what;   // "wave file" or "truncation file", for the messages
  // one record whose length must be `expect` bytes; the leading marker is
  // checked before anything is allocated, so a damaged marker cannot ask for
  // a large buffer
  std::vector<char> record(size_t expect) {
    int32_t m = 0, m2 = 0;
    need(std::fread(&m, 4, 1, f) == 1 && m >= 0, std::string(what) + ": truncated or not a Fortran unformatted file");
    need((size_t)m == expect, std::string(what) + ": unexpected record length");
    std::vector<char> b((size_t)m);
    need((m == 0 || std::fread(b.data(), 1, (size_t)m, f) == (size_t)m) && std::fread(&m2, 4, 1, f) == 1 && m2 == m,
         std::string(what) + ": damaged record");
    return b;
  }
  std::vector<int32_t> ints(size_t n) {
    std::vector<char> b = record(4 * n);
    std::vector<int32_t> v(n);
    std::memcpy(v.data(), b.data(), 4 * n);
    return v;
  }
};
struct FileCloser {
  FILE* f;
  ~FileCloser() {
    if (f) std::fclose(f);
  }
};
}  // namespace

int hh_wave_to_disk(hh_ctx* ctx, const hh_superblock* sb, const double* psi, const char* path) {
  return guard(ctx, [&] {
    need(sb && psi && path, "null argument");
    const SuperBasis& S = sb->plan.sb;
    std::vector<double2> h((size_t)std::max<int64_t>(S.dim, 1));
    check_hip(hipMemcpyAsync(h.data(), psi, sizeof(double2) * (size_t)S.dim, hipMemcpyDeviceToHost, ctx->stream),
              "d2h");
    check_hip(hipStreamSynchronize(ctx->stream), "d2h");
    need(S.dim < (int64_t(1) << 31), "superblock too large for the reference's default-integer dim");
    FileCloser fc{std::fopen(path, "wb")};
    need(fc.f != nullptr, std::string("cannot open ") + path);
    FortranOut o{fc.f, "wave file"};
    o.ints({sb->sys->d.length + 1});
    o.ints({sb->env->d.length + 1});
    o.ints({(int64_t)S.pairs.size()});
    o.ints({S.dim});
    for (size_t i = 0; i < S.pairs.size(); ++i) {
      const Sector& a = sb->plan.sbs.sec[S.pairs[i].first];
      const Sector& b = sb->plan.ebs.sec[S.pairs[i].second];
      o.record(nullptr, 0);
      o.ints({a.N, a.S2});
      o.ints({b.N, b.S2});
      o.ints({S.rows[i], S.cols[i]});
      o.record(h.data() + S.off[i], sizeof(double2) * (size_t)S.rows[i] * S.cols[i]);
    }
    need(std::fflush(fc.f) == 0, "cannot write the wave file");
  });
}

int hh_wave_from_disk(hh_ctx* ctx, const hh_superblock* sb, const char* path, double* psi) {
  return guard(ctx, [&] {
    need(sb && psi && path, "null argument");
    const SuperBasis& S = sb->plan.sb;
    FileCloser fc{std::fopen(path, "rb")};
    need(fc.f != nullptr, std::string("cannot open ") + path);
    FortranIn in{fc.f, "wave file"};
    need(in.ints(1)[0] == sb->sys->d.length + 1 && in.ints(1)[0] == sb->env->d.length + 1,
         "wave file: block lengths differ from the superblock's");
    need(in.ints(1)[0] == (int32_t)S.pairs.size() && in.ints(1)[0] == (int32_t)S.dim,
         "wave file: sector count or dimension differs from the superblock's");
    std::vector<double2> h((size_t)std::max<int64_t>(S.dim, 1));
    for (size_t i = 0; i < S.pairs.size(); ++i) {
      const Sector& a = sb->plan.sbs.sec[S.pairs[i].first];
      const Sector& b = sb->plan.ebs.sec[S.pairs[i].second];
      in.record(0);   // the empty record before each sector
      const auto qa = in.ints(2), qb = in.ints(2), dm = in.ints(2);
      need(qa[0] == a.N && qa[1] == a.S2 && qb[0] == b.N && qb[1] == b.S2 && dm[0] == S.rows[i] &&
               dm[1] == S.cols[i],
           "wave file: super sector " + std::to_string(i + 1) + " differs from the superblock's");
      const std::vector<char> v = in.record(sizeof(double2) * (size_t)S.rows[i] * S.cols[i]);
      std::memcpy(h.data() + S.off[i], v.data(), v.size());
    }
    check_hip(hipMemcpyAsync(psi, h.data(), sizeof(double2) * (size_t)S.dim, hipMemcpyHostToDevice, ctx->stream),
              "h2d");
    check_hip(hipStreamSynchronize(ctx->stream), "h2d");
  });
}

// truns_to_disk / truns_from_disk (code/space_info.f90:1621-1661), i.e.
// block_to_disk / block_from_disk (:1470-1521) of the truncation operator
// Get_truncation_operator builds (code/get_den.f90:362-421): records len, num,
// dim, up_dif, down_dif, then per kept sector an empty record, (N, 2S,
// down_dif = 0), (row_dim, sdim) and the row_dim x sdim complex*16 block of
// the kept eigenvectors (the last sdim columns of the ascending solve).  The
// reference's environment density matrix is psi^dag psi
// (Get_density_matrix_env, code/get_den.f90:84-91), the complex conjugate of
// this library's psi^T conj(psi), so environment truncations are written and
// read conjugated.
int hh_trunc_to_disk(hh_ctx* ctx, const hh_trunc* t, int length, const char* path) {
  return guard(ctx, [&] {
    need(t && path, "null argument");
    need(length >= 1, "length must be the enlarged block's length (>= 1)");
    std::vector<double2> h((size_t)std::max<size_t>(t->evec.bytes / sizeof(double2), 1));
    check_hip(hipMemcpyAsync(h.data(), t->evec.p, t->evec.bytes, hipMemcpyDeviceToHost, ctx->stream), "d2h");
    check_hip(hipStreamSynchronize(ctx->stream), "d2h");
    int64_t dim = 0;
    for (const KeptSector& k : t->kept) dim += k.keep;
    FileCloser fc{std::fopen(path, "wb")};
    need(fc.f != nullptr, std::string("cannot open ") + path);
    FortranOut o{fc.f, "truncation file"};
    o.ints({length});
    o.ints({(int64_t)t->kept.size()});
    o.ints({dim});
    o.ints({0});
    o.ints({0});
    std::vector<double2> blk;
    for (const KeptSector& k : t->kept) {
      const Sector& a = t->eb.sec[k.a];
      o.record(nullptr, 0);
      o.ints({a.N, a.S2, 0});
      o.ints({a.dim, k.keep});
      blk.resize((size_t)a.dim * k.keep);
      for (int c = 0; c < k.keep; ++c)
        std::memcpy(blk.data() + (size_t)c * a.dim, h.data() + k.u_off + (int64_t)c * k.u_ld, sizeof(double2) * a.dim);
      if (t->side == 1)
        for (auto& x : blk) x.y = -x.y;
      o.record(blk.data(), sizeof(double2) * blk.size());
    }
    need(std::fflush(fc.f) == 0, "cannot write the truncation file");
  });
}

int hh_trunc_from_disk(hh_ctx* ctx, const hh_block* prev, int side, const char* path, int32_t* length,
                       hh_trunc** out) {
  if (out) *out = nullptr;
  return guard(ctx, [&] {
    need(prev && path && out, "null argument");
    need(side == 0 || side == 1, "side must be 0 (sys) or 1 (env)");
    FileCloser fc{std::fopen(path, "rb")};
    need(fc.f != nullptr, std::string("cannot open ") + path);
    FortranIn in{fc.f, "truncation file"};
    const int32_t len = in.ints(1)[0], num = in.ints(1)[0], dim = in.ints(1)[0];
    need(in.ints(1)[0] == 0 && in.ints(1)[0] == 0, "truncation file: the operator must be a scalar with no N shift");
    need(len == prev->d.length + 1, "truncation file: its length is not the enlarged length of this block");
    need(num >= 0, "truncation file: negative sector count");
    std::unique_ptr<hh_trunc> t(new hh_trunc());
    t->side = side;
    t->eb = enlarge_basis(prev->d.sec, site_sectors(ctx->model.kind));
    std::vector<double2> arena;
    std::vector<bool> seen(t->eb.sec.size(), false);
    int64_t total = 0;
    for (int32_t i = 0; i < num; ++i) {
      in.record(0);   // the empty record before each sector
      const auto q = in.ints(3), d = in.ints(2);
      need(q[2] == 0, "truncation file: sector with a spin change");
      int a = -1;
      for (size_t j = 0; j < t->eb.sec.size(); ++j)
        if (t->eb.sec[j].N == q[0] && t->eb.sec[j].S2 == q[1]) a = (int)j;
      need(a >= 0 && !seen[a], "truncation file: sector (" + std::to_string(q[0]) + ", " + std::to_string(q[1]) +
                                   ") is not an enlarged sector of this block, or repeats");
      seen[a] = true;
      need(d[0] == t->eb.sec[a].dim && d[1] >= 1 && d[1] <= d[0], "truncation file: block dimensions do not fit");
      const std::vector<char> v = in.record(sizeof(double2) * (size_t)d[0] * d[1]);
      KeptSector k;
      k.a = a;
      k.keep = d[1];
      k.u_off = (int64_t)arena.size();
      k.u_ld = d[0];
      arena.resize(arena.size() + (size_t)d[0] * d[1]);
      std::memcpy(arena.data() + k.u_off, v.data(), v.size());
      if (side == 1)
        for (size_t e = 0; e < (size_t)d[0] * d[1]; ++e) arena[k.u_off + e].y = -arena[k.u_off + e].y;
      t->kept.push_back(k);
      total += d[1];
    }
    need(total == dim, "truncation file: dim differs from the sum of the kept columns");
    t->num_state = (int)total;
    t->evec.alloc(sizeof(double2) * std::max<size_t>(arena.size(), 1));
    if (!arena.empty()) {
      check_hip(hipMemcpyAsync(t->evec.p, arena.data(), sizeof(double2) * arena.size(), hipMemcpyHostToDevice,
                               ctx->stream), "h2d");
      check_hip(hipStreamSynchronize(ctx->stream), "h2d");
    }
    if (length) *length = len;
    *out = t.release();
  });
}

// basis_to_disk / basis_from_disk (code/space_info.f90:1665-1762): Get_basis
// (code/update.f90:11-112) of a block and one site as formatted text, with
// the reference's edit descriptors: (A4,I6) for len / num / dim / idx / pos,
// (A11,I4) new_num_up, (A13,I4) new_num_down, (A10,I4,2X,A12,I4) for the
// block and site quantum numbers of a subsector, and an empty list-directed
// record before each sector.  num_up / num_down are N and 2S; idx ends equal
// to the subsector count and pos is the 0-based offset in the sector.
namespace {
struct BasisLine {   // one formatted record: a label and one or two integers
  std::string label, label2;
  int width = 6;
  int64_t a = 0, b = 0;
  bool two = false;
};
std::vector<BasisLine> basis_lines(const EnlBasis& eb, int length) {
  std::vector<BasisLine> L;
  auto one = [&](const char* lab, int64_t v, int w) {
    BasisLine x;
    x.label = lab;
    x.a = v;
    x.width = w;
    L.push_back(x);
  };
  auto two = [&](const char* l1, int64_t v1, const char* l2, int64_t v2) {
    BasisLine x;
    x.label = l1;
    x.label2 = l2;
    x.a = v1;
    x.b = v2;
    x.width = 4;
    x.two = true;
    L.push_back(x);
  };
  one("len=", length, 6);
  one("num=", (int64_t)eb.sec.size(), 6);
  one("dim=", eb.total_dim(), 6);
  for (size_t x = 0; x < eb.sec.size(); ++x) {
    L.push_back(BasisLine());   // the empty record
    one("idx=", (int64_t)eb.subs[x].size(), 6);
    one("num=", (int64_t)eb.subs[x].size(), 6);
    one("dim=", eb.sec[x].dim, 6);
    one("new_num_up=", eb.sec[x].N, 4);
    one("new_num_down=", eb.sec[x].S2, 4);
    for (const SubSector& q : eb.subs[x]) {
      one("pos=", q.pos, 6);
      one("dim=", q.dim, 6);
      two("bl_num_up=", eb.bl[q.b].N, "bl_num_down=", eb.bl[q.b].S2);
      two("st_num_up=", eb.st[q.s].N, "st_num_down=", eb.st[q.s].S2);
    }
  }
  return L;
}
// Fortran Iw output: right-justified in w columns, or w asterisks when the
// value does not fit
std::string fortran_int(int64_t v, int w) {
  std::string t = std::to_string(v);
  if ((int)t.size() > w) return std::string(w, '*');
  return std::string(w - t.size(), ' ') + t;
}
}  // namespace

int hh_basis_to_disk(hh_ctx* ctx, const hh_block* blk, const char* path) {
  return guard(ctx, [&] {
    need(blk && path, "null argument");
    const EnlBasis eb = enlarge_basis(blk->d.sec, site_sectors(ctx->model.kind));
    const std::vector<BasisLine> L = basis_lines(eb, blk->d.length + 1);
    std::string out;
    for (const BasisLine& x : L) {
      if (x.label.empty()) {
        out += "\n";
        continue;
      }
      const std::string v = fortran_int(x.a, x.width);
      need(v[0] != '*', "basis file: " + x.label + " " + std::to_string(x.a) +
                            " does not fit the reference's I" + std::to_string(x.width) + " field");
      out += x.label + v;
      if (x.two) {
        const std::string v2 = fortran_int(x.b, x.width);
        need(v2[0] != '*', "basis file: " + x.label2 + " does not fit the reference's I4 field");
        out += "  " + x.label2 + v2;
      }
      out += "\n";
    }
    FileCloser fc{std::fopen(path, "wb")};
    need(fc.f != nullptr, std::string("cannot open ") + path);
    need(std::fwrite(out.data(), 1, out.size(), fc.f) == out.size() && std::fflush(fc.f) == 0,
         "cannot write the basis file");
  });
}

int hh_basis_from_disk(hh_ctx* ctx, const hh_block* blk, const char* path, int32_t* length) {
  return guard(ctx, [&] {
    need(blk && path, "null argument");
    FileCloser fc{std::fopen(path, "rb")};
    need(fc.f != nullptr, std::string("cannot open ") + path);
    const EnlBasis eb = enlarge_basis(blk->d.sec, site_sectors(ctx->model.kind));
    const std::vector<BasisLine> L = basis_lines(eb, blk->d.length + 1);
    int lineno = 0;
    auto next_line = [&]() {
      std::string t;
      int c;
      while ((c = std::fgetc(fc.f)) != EOF && c != '\n') t.push_back((char)c);
      need(c != EOF || !t.empty(), "basis file: ends early (line " + std::to_string(lineno + 1) + ")");
      if (!t.empty() && t.back() == '\r') t.pop_back();
      ++lineno;
      return t;
    };
    // an Iw input field: blanks are ignored, an empty field reads as 0
    auto field = [&](const std::string& t, size_t at, int w) {
      std::string f = at < t.size() ? t.substr(at, (size_t)w) : std::string();
      std::string d;
      for (char ch : f)
        if (ch != ' ') d.push_back(ch);
      if (d.empty()) return (int64_t)0;
      need(d.find_first_not_of("+-0123456789") == std::string::npos,
           "basis file: line " + std::to_string(lineno) + " holds no integer where one is expected");
      return (int64_t)std::stoll(d);
    };
    for (size_t i = 0; i < L.size(); ++i) {
      const BasisLine& x = L[i];
      const std::string t = next_line();
      if (x.label.empty()) continue;   // list-directed read of nothing: skips the record
      const size_t l1 = x.label.size();
      int64_t a = field(t, l1, x.width), b = 0;
      if (x.two) b = field(t, l1 + (size_t)x.width + 2 + x.label2.size(), x.width);
      if (i == 0) {
        if (length) *length = (int32_t)a;
        need(a == x.a, "basis file: len differs from the enlarged length of this block");
        continue;
      }
      need(a == x.a && b == x.b, "basis file: line " + std::to_string(lineno) + " (" + x.label +
                                     ") differs from Get_basis of this block and one site");
    }
  });
}

}  // extern "C"
