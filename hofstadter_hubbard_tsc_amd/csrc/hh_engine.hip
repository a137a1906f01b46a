// Grouped complex GEMM over tile lists and Lanczos vector kernels for gfx950.
//
// One wave64 owns a 32x32 macro tile of one output region and walks every
// term that lands on it, or a contiguous share of those terms whose partial
// sum a second kernel adds in a fixed order (output-stationary, no atomics,
// deterministic).  Complex products run as four real 16x16x4 FP64 (or FP32)
// MFMAs per k-step; the term coefficient and the operand's conjugation are
// applied to the registers between load and MFMA.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <mutex>
#include <string>
#include <stdexcept>
#include <type_traits>

#include <cmath>

#include "hh_engine.hpp"

// edge work items run the narrow pipelines when more than this share of a
// launch's MFMA work would fall on empty 16 x 16 sub-tiles
#ifndef HH_EDGE_SHARE
#define HH_EDGE_SHARE 0.2
#endif

namespace hh {

void check_hip(hipError_t e, const char* what) {
  if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

// ----------------------------------------------------------------------------
// task upload: 32x32 macro tiles, one per wave; long term lists are split
// over several waves and reduced in a fixed order
// ----------------------------------------------------------------------------
static constexpr int WA = 2, WB = 2;   // 16x16 MFMA sub-tiles per wave: rows, columns
static constexpr int MTR = 16 * WA;     // macro tile rows (one wave)
static constexpr int MTC = 16 * WB;     // macro tile columns
// A wide macro tile (LaunchConfig::wide) has WBW sub-tile columns: with
// wide 1 a region whose last 32-column chunk would leave at most 16 columns
// over takes them into one 32 x 48 tile, so its rows' operand panels are not
// read again for a nearly empty tile; with wide 2 every chunk is 48 wide
// (each operand panel of the rows is read for 48 output columns, not 32).
static constexpr int WBW = 3;
static constexpr int MTCW = 16 * WBW;
// partial sums of a split macro tile are stored column-major in strips of
// MTR x 16 elements: a 32-wide tile's partial takes two strips, a 48-wide
// one three, so plain tiles need no room for the wide shape
static constexpr int STRIP = MTR * 16;
static constexpr int WAVES = 4;        // independent waves per workgroup
// Complex products by the 3M scheme (3 real MFMAs per complex product) in
// every shape class, FP64 and FP32; HH_CPLX3 = 0 builds the four-product form
// instead.
#ifndef HH_CPLX3
#define HH_CPLX3 1
#endif
// Term lists are cut so that the launch holds about LaunchConfig::target_items
// equal work items (enough that longest-first dispatch leaves a short tail);
// macro tiles are listed in bands of LaunchConfig::band macro rows, column by
// column, and runs of XCHUNK consecutive workgroups are kept on one XCD:
// neighbouring items then share operand panels in that XCD's L2.
static constexpr int XCHUNK = 8;

void DevTasks::release() {
  if (!borrowed) {
    if (tiles) (void)hipFree(tiles);
    if (terms) (void)hipFree(terms);
    if (work) (void)hipFree(work);
    if (red) (void)hipFree(red);
    if (partial) (void)hipFree(partial);
  }
  borrowed = false;
  tiles = nullptr;
  terms = nullptr;
  work = nullptr;
  red = nullptr;
  partial = nullptr;
  ntiles = nterms = nwork = nred = nslots = 0;
}

void DevTasks::upload(const TaskList& tl, hipStream_t s, void* place, size_t cap, const LaunchConfig& config) {
  release();
  cfg = config;
  if (cfg.target_items < 1 || cfg.max_groups < 1 || cfg.band < 1 || cfg.edge < -1 || cfg.edge > 1 || cfg.wide < 0 ||
      cfg.wide > 2)
    throw std::runtime_error("invalid launch configuration");
  if (place && (sizeof(Tile) * tl.tiles.size() + sizeof(Term) * tl.terms.size() > cap))
    throw HHCapacity("task list tables exceed the fixed scratch");
  ntiles = (int)tl.tiles.size();
  nterms = (int)tl.terms.size();
  flops_alg = tl.flops_alg;
  flops_exec = tl.flops_exec;
  // the kernel addresses an operand sub-block with 32-bit element offsets
  // from its base: reject sub-blocks that would not fit
  for (const Tile& t : tl.tiles)
    for (int k = t.t0; k < t.t1; ++k) {
      const Term& tm = tl.terms[k];
      if (tm.K <= 0) continue;
      const bool a_rowfast = (tm.a_op == OP_NN || tm.a_op == OP_R), b_kfast = (tm.b_op == OP_NN || tm.b_op == OP_R);
      const int64_t a_end = a_rowfast ? (int64_t)(t.M - 1) + (int64_t)(tm.K - 1) * tm.a_ld
                                      : (int64_t)(tm.K - 1) + (int64_t)(t.M - 1) * tm.a_ld;
      const int64_t b_end = b_kfast ? (int64_t)(tm.K - 1) + (int64_t)(t.N - 1) * tm.b_ld
                                    : (int64_t)(t.N - 1) + (int64_t)(tm.K - 1) * tm.b_ld;
      if (a_end >= (int64_t(1) << 32) || b_end >= (int64_t(1) << 32))
        throw std::runtime_error("operand sub-block exceeds 2^32 elements");
    }
  // column chunks of a region: 32 wide, the last one up to 48 wide when
  // wide tiles are on and it would otherwise leave at most 16 columns over
  auto chunks = [&](int N) {
    std::vector<std::pair<int, int>> c;   // (first column, width)
    const int step = cfg.wide == 2 ? MTCW : MTC;
    for (int c0 = 0; c0 < N; c0 += step) c.push_back({c0, std::min(step, N - c0)});
    if (cfg.wide == 1 && c.size() >= 2 && c.back().second <= 16) {
      const int w = c.back().second;
      c.pop_back();
      c.back().second += w;
    }
    return c;
  };
  // work per macro tile = sum of max(K, 1) over its terms
  double total = 0.0;
  std::vector<double> tw(ntiles, 0.0);
  for (int i = 0; i < ntiles; ++i) {
    const Tile& t = tl.tiles[i];
    double c = 0.0;
    for (int k = t.t0; k < t.t1; ++k) c += (double)std::max(tl.terms[k].K, 1);
    tw[i] = c;
    total += c * (double)((t.M + MTR - 1) / MTR) * (double)chunks(t.N).size();
  }
  const double target = std::max(64.0, total / cfg.target_items);
  // a work item with its K-work and the width of its column chunk (32, or
  // up to 48 for a wide macro tile)
  struct Item {
    double work;
    WorkItem it;
    int width;
  };
  std::vector<Item> items;
  std::vector<ReduceItem> reds;
  int slot = 0;
  for (int i = 0; i < ntiles; ++i) {
    const Tile& t = tl.tiles[i];
    // contiguous term groups of roughly equal work
    std::vector<int> cuts{t.t0};
    const int groups = std::max(1, std::min(cfg.max_groups, (int)std::ceil(tw[i] / target)));
    if (groups > 1) {
      double acc = 0.0;
      const double per = tw[i] / groups;
      for (int k = t.t0; k < t.t1; ++k) {
        acc += (double)std::max(tl.terms[k].K, 1);
        if (acc >= per * (double)cuts.size() && k + 1 < t.t1 && (int)cuts.size() < groups) cuts.push_back(k + 1);
      }
    }
    cuts.push_back(t.t1);
    const int g = (int)cuts.size() - 1;
    // macro tiles in bands of cfg.band macro rows, column by column inside a band
    const int nr = (t.M + MTR - 1) / MTR;
    const std::vector<std::pair<int, int>> cols = chunks(t.N);
    const int nc = (int)cols.size();
    const int band = cfg.band;
    for (int rb = 0; rb < nr; rb += band)
      for (int ci = 0; ci < nc; ++ci)
        for (int ri = rb; ri < std::min(nr, rb + band); ++ri) {
          const int r = ri * MTR, c = cols[ci].first;
          const int width = cols[ci].second > MTC ? MTCW : MTC, strips = width / 16;
          if (g == 1) {
            items.push_back({tw[i], WorkItem{i, r, c, t.t0, t.t1, -1}, cols[ci].second});
          } else {
            for (int q = 0; q < g; ++q) {
              double wq = 0.0;
              for (int k = cuts[q]; k < cuts[q + 1]; ++k) wq += (double)std::max(tl.terms[k].K, 1);
              items.push_back({wq, WorkItem{i, r, c, cuts[q], cuts[q + 1], slot + q * strips}, cols[ci].second});
            }
            reds.push_back(ReduceItem{i, r, c, slot, g, width});
            slot += g * strips;
          }
        }
  }
  // share of the MFMA work that falls on empty 16x16 sub-tiles of edge tiles
  {
    double all = 0.0, empty = 0.0;
    for (auto& it : items) {
      const Tile& t = tl.tiles[it.it.tile];
      const int width = it.width > MTC ? MTCW : MTC;
      const int mv = std::min(MTR, t.M - it.it.r0), nv = it.width;
      const int full = (MTR / 16) * (width / 16), used = ((mv + 15) / 16) * ((nv + 15) / 16);
      all += it.work * full;
      empty += it.work * (full - used);
    }
    edge_skip = cfg.edge < 0 ? (all > 0.0 && empty > HH_EDGE_SHARE * all) : cfg.edge == 1;
  }
  // longest work items first (LPT) inside each shape class
  // classes: 0 full 32 x 32, 1 at most 16 rows, 2 at most 16 columns, 3
  // both (the last three only with edge_skip); 4 wide 32 x 48, 5 wide with at
  // most 16 rows (with edge_skip)
  auto cls = [&](const Item& x) {
    const Tile& t = tl.tiles[x.it.tile];
    if (x.width > MTC) return 4 + ((edge_skip && t.M - x.it.r0 <= 16) ? 1 : 0);
    if (!edge_skip) return 0;
    return (t.M - x.it.r0 <= 16 ? 1 : 0) | (x.width <= 16 ? 2 : 0);
  };
  std::stable_sort(items.begin(), items.end(),
                   [&](const Item& a, const Item& b) {
                     const int ca = cls(a), cb = cls(b);
                     if (ca != cb) return ca < cb;
                     return a.work > b.work;
                   });
  std::vector<WorkItem> w;
  w.reserve(items.size());
  for (int c = 0; c < NCLASS + 1; ++c) cls_off[c] = 0;
  for (auto& it : items) {
    w.push_back(it.it);
    cls_off[cls(it) + 1] += 1;
  }
  for (int c = 1; c < NCLASS + 1; ++c) cls_off[c] += cls_off[c - 1];
  nwork = (int)w.size();
  nred = (int)reds.size();
  nslots = slot;
  if (place) {
    // carve the tables out of the given memory, 256-byte aligned; nothing is
    // allocated and release() leaves the memory alone
    auto up = [](size_t x) { return (x + 255) & ~size_t(255); };
    const size_t o_tiles = 0, o_terms = o_tiles + up(sizeof(Tile) * ntiles);
    const size_t o_work = o_terms + up(sizeof(Term) * nterms);
    const size_t o_red = o_work + up(sizeof(WorkItem) * nwork);
    const size_t o_part = o_red + up(sizeof(ReduceItem) * nred);
    const size_t end = o_part + sizeof(double2) * (size_t)nslots * STRIP;
    if (end > cap) {
      nwork = nred = nslots = ntiles = nterms = 0;
      throw HHCapacity("task list needs " + std::to_string(end) + " bytes of device tables and partial sums, the "
                       "fixed scratch holds " + std::to_string(cap));
    }
    char* base = (char*)place;
    borrowed = true;
    tiles = ntiles ? (Tile*)(base + o_tiles) : nullptr;
    terms = nterms ? (Term*)(base + o_terms) : nullptr;
    work = nwork ? (WorkItem*)(base + o_work) : nullptr;
    red = nred ? (ReduceItem*)(base + o_red) : nullptr;
    partial = nred ? (double2*)(base + o_part) : nullptr;
    if (ntiles)
      check_hip(hipMemcpyAsync(tiles, tl.tiles.data(), sizeof(Tile) * ntiles, hipMemcpyHostToDevice, s), "copy");
    if (nterms)
      check_hip(hipMemcpyAsync(terms, tl.terms.data(), sizeof(Term) * nterms, hipMemcpyHostToDevice, s), "copy");
    if (nwork)
      check_hip(hipMemcpyAsync(work, w.data(), sizeof(WorkItem) * nwork, hipMemcpyHostToDevice, s), "copy");
    if (nred)
      check_hip(hipMemcpyAsync(red, reds.data(), sizeof(ReduceItem) * nred, hipMemcpyHostToDevice, s), "copy");
    check_hip(hipStreamSynchronize(s), "task upload");
    return;
  }
  if (ntiles) {
    check_hip(hipMalloc(&tiles, sizeof(Tile) * ntiles), "hipMalloc tiles");
    check_hip(hipMemcpyAsync(tiles, tl.tiles.data(), sizeof(Tile) * ntiles, hipMemcpyHostToDevice, s), "copy tiles");
  }
  if (nterms) {
    check_hip(hipMalloc(&terms, sizeof(Term) * nterms), "hipMalloc terms");
    check_hip(hipMemcpyAsync(terms, tl.terms.data(), sizeof(Term) * nterms, hipMemcpyHostToDevice, s), "copy terms");
  }
  if (nwork) {
    check_hip(hipMalloc(&work, sizeof(WorkItem) * nwork), "hipMalloc work");
    check_hip(hipMemcpyAsync(work, w.data(), sizeof(WorkItem) * nwork, hipMemcpyHostToDevice, s), "copy work");
  }
  if (nred) {
    check_hip(hipMalloc(&red, sizeof(ReduceItem) * nred), "hipMalloc red");
    check_hip(hipMemcpyAsync(red, reds.data(), sizeof(ReduceItem) * nred, hipMemcpyHostToDevice, s), "copy red");
    check_hip(hipMalloc(&partial, sizeof(double2) * (size_t)nslots * STRIP), "hipMalloc partial");
  }
  check_hip(hipStreamSynchronize(s), "task upload");
}

// ----------------------------------------------------------------------------
// fixed scratch of the allocation-free entry point
// ----------------------------------------------------------------------------
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

namespace {
constexpr int MAX_DEVICES = 64;
std::mutex scratch_mu;
void* scratch_ptr[MAX_DEVICES] = {};
}  // namespace

void* noalloc_scratch_init(int device) {
  if (device < 0 || device >= MAX_DEVICES) throw std::runtime_error("device ordinal out of range for the scratch");
  std::lock_guard<std::mutex> g(scratch_mu);
  if (!scratch_ptr[device]) {
    // the array's name is a host-side stand-in; its device address on the
    // current device comes from the runtime
    void* p = nullptr;
    check_hip(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
    if (!p) throw std::runtime_error("hipGetSymbolAddress(scratch) returned no address");
    scratch_ptr[device] = p;
  }
  return scratch_ptr[device];
}

void* noalloc_scratch(int device) {
  if (device < 0 || device >= MAX_DEVICES) return nullptr;
  std::lock_guard<std::mutex> g(scratch_mu);
  return scratch_ptr[device];
}

// ----------------------------------------------------------------------------
// MFMA helpers
// ----------------------------------------------------------------------------
typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

template <typename T> struct Acc;
// Operand layouts of both 16x16x4 MFMAs: lane l holds A[l%16][l/16] and
// B[l/16][l%16].  Output element v of lane l sits in column l%16 and row
// 4*(l/16)+v for f32 but row (l/16)+4*v for f64 (checked on gfx950).
template <> struct Acc<double> {
  typedef d4 V;
  static __device__ __forceinline__ V mfma(double a, double b, V c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lk, int v) { return lk + 4 * v; }
};
template <> struct Acc<float> {
  typedef f4 V;
  static __device__ __forceinline__ V mfma(float a, float b, V c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lk, int v) { return 4 * lk + v; }
};

// operand storage of a launch: complex128 buffers (BufPtrs, BufTable) or
// complex64 buffers (BufPtrs32, the FP32-storage form)
template <typename Bufs> struct StoreOf { typedef double2 type; };
template <> struct StoreOf<BufPtrs32> { typedef float2 type; };

__device__ __forceinline__ double2 to_d2(double2 v) { return v; }
__device__ __forceinline__ double2 to_d2(float2 v) { return make_double2((double)v.x, (double)v.y); }
__device__ __forceinline__ void put(double2* p, double re, double im) { *p = make_double2(re, im); }
__device__ __forceinline__ void put(float2* p, double re, double im) { *p = make_float2((float)re, (float)im); }

template <typename S>
__device__ __forceinline__ double2 ld_op(const S* __restrict__ p, int64_t idx, bool conj) {
  double2 v = to_d2(p[idx]);
  if (conj) v.y = -v.y;
  return v;
}

// ----------------------------------------------------------------------------
// grouped tile kernel: every wave owns one 32x32 macro tile and streams the
// MFMA operands straight from HBM / L2 (lane l loads A[l%16][k] and
// B[k][l%16] for its k = l/16 of each k-step), so waves never wait on each
// other: no LDS, no barriers.  Each shape class (NA x NB of the 2 x 2
// sub-tiles of 16 x 16 hold data) is its own instantiation with a single
// pipeline, so no wave carries the code or registers of another shape.
// CPLX3 selects three real MFMAs per complex product (FP64) over four.
// ----------------------------------------------------------------------------
template <typename T, int TAG, typename Bufs, int NA, int NB, bool CPLX3, int WBT>
__global__ __launch_bounds__(256, 2) void k_tiles(const Tile* __restrict__ tiles, const Term* __restrict__ terms,
                                               const WorkItem* __restrict__ work, int nwork, Bufs bufs,
                                               double2* __restrict__ partial) {
  typedef typename Acc<T>::V V;
  typedef typename StoreOf<Bufs>::type S;   // operand element (complex128 or complex64)
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  int wg = blockIdx.x;
  {
    // workgroups reach the 8 XCDs round-robin; give each XCD runs of XCHUNK
    // consecutive workgroups (a bijection on the full blocks of 8*XCHUNK)
    const int full = ((int)gridDim.x / (8 * XCHUNK)) * 8 * XCHUNK;
    if (wg < full) {
      const int x = wg & 7, j = wg >> 3;
      wg = (j / XCHUNK) * 8 * XCHUNK + x * XCHUNK + (j % XCHUNK);
    }
  }
  // wave-uniform: lets the compiler keep the work item and terms in SGPRs
  const int wi = __builtin_amdgcn_readfirstlane(wg * WAVES + wave);
  if (wi >= nwork) return;
  const WorkItem w = work[wi];
  const Tile tile = tiles[w.tile];
  if constexpr (std::is_same<Bufs, BufTable>::value) {
    // a finished batch member's tiles are skipped (wave-uniform)
    if (bufs.active && bufs.active[tile.c_buf >> 3] == 0) return;
  }
  const int r0 = w.r0, c0 = w.c0;
  const int Mv = min(16 * WA, tile.M - r0), Nv = min(16 * WBT, tile.N - c0);
  const int li = lane & 15, lk = lane >> 4;

  V accR[WA][WBT], accI[WA][WBT];
#pragma unroll
  for (int a = 0; a < WA; ++a)
#pragma unroll
    for (int b = 0; b < WBT; ++b) {
      accR[a][b] = V{0, 0, 0, 0};
      accI[a][b] = V{0, 0, 0, 0};
    }
  // elementwise (K = 0) terms, straight into the accumulators (before the
  // GEMM terms in the four-product form, after them in the 3M form)
  auto elementwise = [&]() {
    for (int ti = w.t0; ti < w.t1; ++ti) {
      const Term tm = terms[ti];
      if (tm.K != 0) continue;
      const S* __restrict__ A = bufs.p[tm.a_buf];
      const bool cj = tm.a_op == OP_R || tm.a_op == OP_C;
#pragma unroll
      for (int a = 0; a < WA; ++a)
#pragma unroll
        for (int b = 0; b < WBT; ++b)
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            const int r = a * 16 + Acc<T>::row(lk, v), c = b * 16 + li;
            if (r < Mv && c < Nv) {
              double2 x = ld_op(A, tm.a_off + (int64_t)(r0 + r) + (int64_t)(c0 + c) * tm.a_ld, cj);
              accR[a][b][v] += (T)(tm.cr * x.x - tm.ci * x.y);
              accI[a][b][v] += (T)(tm.cr * x.y + tm.ci * x.x);
            }
          }
    }
  };
  // Complex products by the 3M scheme: with P1 = sum ar br, P2 = sum ai bi
  // and P3 = sum (ar + ai)(br + bi), Re = P1 - P2 and Im = P3 - P1 - P2.
  // During the GEMM terms accR holds P1, accP holds P2 and accI holds P3, all
  // starting from zero.
  constexpr bool M3 = CPLX3;
  V accP[M3 ? WA : 1][M3 ? WBT : 1];
  if constexpr (M3) {
#pragma unroll
    for (int a = 0; a < WA; ++a)
#pragma unroll
      for (int b = 0; b < WBT; ++b) accP[a][b] = V{0, 0, 0, 0};
  } else {
    elementwise();
  }
  // GEMM terms (K > 0): one software pipeline over the whole term range.  The
  // operands of k-step kb are converted to MFMA inputs, the loads of the next
  // k-step go out into the other register buffer (the first k-step of the next
  // term when kb is the last one of this term), then the MFMAs of kb run while
  // those loads fly.  Every load is unconditional and in bounds: rows / columns
  // past the region edge read the last valid row / column (never stored), k
  // past K reads k = K-1 and is zeroed in the A operand at use.
  struct OpAddr {
    const S* Ab;
    const S* Bb;
    uint32_t pa[WA], pb[WBT];   // k-independent part of each lane's element offset
    int a_ks, b_ks, K;
    double sa, sb, cr, ci;
  };
  auto make_op = [&](const Term& tm) {
    OpAddr o;
    const bool a_rowfast = (tm.a_op == OP_NN || tm.a_op == OP_R);
    const bool b_kfast = (tm.b_op == OP_NN || tm.b_op == OP_R);
    o.Ab = bufs.p[tm.a_buf] + tm.a_off;
    o.Bb = bufs.p[tm.b_buf] + tm.b_off;
    const int a_rs = a_rowfast ? 1 : tm.a_ld, b_cs = b_kfast ? tm.b_ld : 1;
    o.a_ks = a_rowfast ? tm.a_ld : 1;
    o.b_ks = b_kfast ? 1 : tm.b_ld;
#pragma unroll
    for (int a = 0; a < WA; ++a) o.pa[a] = (uint32_t)(min(r0 + 16 * a + li, r0 + Mv - 1) * a_rs);
#pragma unroll
    for (int b = 0; b < WBT; ++b) o.pb[b] = (uint32_t)(min(c0 + 16 * b + li, c0 + Nv - 1) * b_cs);
    o.K = tm.K;
    o.sa = (tm.a_op == OP_C || tm.a_op == OP_R) ? -1.0 : 1.0;
    o.sb = (tm.b_op == OP_C || tm.b_op == OP_R) ? -1.0 : 1.0;
    o.cr = tm.cr;
    o.ci = tm.ci;
    return o;
  };
  // NA x NB of the WA x WBT sub-tiles hold data (the kernels of the narrow
  // shape classes run work items with Mv <= 16 and / or Nv <= 16: fewer loads
  // and MFMAs, and one pipeline per kernel)
  auto load = [&](S (&xa)[WA], S (&xb)[WBT], const OpAddr& o, int k0) {
    const uint32_t kc = (uint32_t)min(k0 + lk, o.K - 1);
#pragma unroll
    for (int a = 0; a < NA; ++a) xa[a] = o.Ab[o.pa[a] + kc * (uint32_t)o.a_ks];
#pragma unroll
    for (int b = 0; b < NB; ++b) xb[b] = o.Bb[o.pb[b] + kc * (uint32_t)o.b_ks];
  };
  auto next_gemm = [&](int ti) {
    while (ti < w.t1 && terms[ti].K == 0) ++ti;
    return ti;
  };
  auto run = [&]() {
    int ti = next_gemm(w.t0);
    if (ti >= w.t1) return;
    OpAddr c = make_op(terms[ti]);
    int tn = next_gemm(ti + 1);
    Term dn = terms[tn < w.t1 ? tn : ti];   // descriptor of the following GEMM term
    S xa0[WA], xb0[WBT], xa1[WA], xb1[WBT];
    load(xa0, xb0, c, 0);
    int kb = 0;
    bool done = false;
    auto step = [&](const S (&xa)[WA], const S (&xb)[WBT], S (&ya)[WA], S (&yb)[WBT]) {
      // only the last k-step of a term can hold k >= K
      const T kf = (kb + 4 <= c.K || kb + lk < c.K) ? T(1) : T(0);
      const T cr = (T)c.cr, ci = (T)c.ci, sa = (T)c.sa, sb = (T)c.sb;
      T ar[WA], ai[WA], br[WBT], bi[WBT];
#pragma unroll
      for (int a = 0; a < NA; ++a) {
        const T xr = kf * (T)xa[a].x, xi = kf * sa * (T)xa[a].y;
        ar[a] = cr * xr - ci * xi;
        ai[a] = cr * xi + ci * xr;
      }
#pragma unroll
      for (int b = 0; b < NB; ++b) {
        br[b] = (T)xb[b].x;
        bi[b] = sb * (T)xb[b].y;
      }
      const bool last = kb + 4 >= c.K;
      // both arms issue the same loads, so the wait counts stay path-independent
      if (last && tn < w.t1) load(ya, yb, make_op(dn), 0);
      else load(ya, yb, c, kb + 4);
      __builtin_amdgcn_sched_barrier(0);
      if constexpr (M3) {
        T pa[WA], pb[WBT];
#pragma unroll
        for (int a = 0; a < NA; ++a) pa[a] = ar[a] + ai[a];
#pragma unroll
        for (int b = 0; b < NB; ++b) pb[b] = br[b] + bi[b];
#pragma unroll
        for (int a = 0; a < NA; ++a)
#pragma unroll
          for (int b = 0; b < NB; ++b) {
            accR[a][b] = Acc<T>::mfma(ar[a], br[b], accR[a][b]);
            accP[a][b] = Acc<T>::mfma(ai[a], bi[b], accP[a][b]);
            accI[a][b] = Acc<T>::mfma(pa[a], pb[b], accI[a][b]);
          }
      } else {
#pragma unroll
        for (int a = 0; a < NA; ++a)
#pragma unroll
          for (int b = 0; b < NB; ++b) {
            accR[a][b] = Acc<T>::mfma(ar[a], br[b], accR[a][b]);
            accR[a][b] = Acc<T>::mfma(-ai[a], bi[b], accR[a][b]);
            accI[a][b] = Acc<T>::mfma(ar[a], bi[b], accI[a][b]);
            accI[a][b] = Acc<T>::mfma(ai[a], br[b], accI[a][b]);
          }
      }
      if (!last) {
        kb += 4;
      } else if (tn >= w.t1) {
        done = true;
      } else {
        c = make_op(dn);
        ti = tn;
        tn = next_gemm(ti + 1);
        dn = terms[tn < w.t1 ? tn : ti];
        kb = 0;
      }
    };
    if constexpr (M3) {
      // one step per iteration; the next step's operands move into the
      // current buffer (the accumulators then keep one home across the loop)
      while (true) {
        step(xa0, xb0, xa1, xb1);
        if (done) break;
#pragma unroll
        for (int a = 0; a < NA; ++a) xa0[a] = xa1[a];
#pragma unroll
        for (int b = 0; b < NB; ++b) xb0[b] = xb1[b];
      }
    } else {
      while (true) {
        step(xa0, xb0, xa1, xb1);
        if (done) break;
        step(xa1, xb1, xa0, xb0);
        if (done) break;
      }
    }
  };
  static_assert(NA >= 1 && NA <= WA && NB >= 1 && NB <= WBT, "sub-tile counts of a shape class");
  run();
  if constexpr (M3) {
#pragma unroll
    for (int a = 0; a < WA; ++a)
#pragma unroll
      for (int b = 0; b < WBT; ++b) {
        accI[a][b] = accI[a][b] - accR[a][b] - accP[a][b];
        accR[a][b] = accR[a][b] - accP[a][b];
      }
    elementwise();
  }
  // write the macro tile once (or its partial sum)
  S* __restrict__ C = bufs.p[tile.c_buf];
#pragma unroll
  for (int a = 0; a < WA; ++a)
#pragma unroll
    for (int b = 0; b < WBT; ++b)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int r = a * 16 + Acc<T>::row(lk, v), c = b * 16 + li;
        if (r < Mv && c < Nv) {
          double re = (double)accR[a][b][v], im = (double)accI[a][b][v];
          if (w.slot >= 0) {
            partial[(int64_t)w.slot * STRIP + r + c * MTR] = make_double2(re, im);
          } else {
            const int64_t o = tile.c_off + (int64_t)(r0 + r) + (int64_t)(c0 + c) * tile.c_ld;
            if (tile.accumulate) {
              const double2 old = to_d2(C[o]);
              re += old.x;
              im += old.y;
            }
            put(C + o, re, im);
          }
        }
      }
}

// C (+)= sum over the slots of a split macro tile, slot order fixed
template <typename Bufs>
__global__ __launch_bounds__(256) void k_reduce(const Tile* __restrict__ tiles, const ReduceItem* __restrict__ red,
                                                Bufs bufs, const double2* __restrict__ partial) {
  const ReduceItem ri = red[blockIdx.x];
  const Tile tile = tiles[ri.tile];
  if constexpr (std::is_same<Bufs, BufTable>::value) {
    if (bufs.active && bufs.active[tile.c_buf >> 3] == 0) return;
  }
  const int width = ri.width;   // 32, or 48 for a wide macro tile
  const int strips = width / 16; // partial-sum strips per slot
  const int Mv = min(MTR, tile.M - ri.r0), Nv = min(width, tile.N - ri.c0);
  typename StoreOf<Bufs>::type* __restrict__ C = bufs.p[tile.c_buf];
  for (int e = threadIdx.x; e < MTR * width; e += 256) {
    const int r = e % MTR, c = e / MTR;
    if (r >= Mv || c >= Nv) continue;
    double re = 0.0, im = 0.0;
    for (int q = 0; q < ri.nslot; ++q) {
      const double2 x = partial[(int64_t)(ri.slot0 + q * strips) * STRIP + e];
      re += x.x;
      im += x.y;
    }
    const int64_t o = tile.c_off + (int64_t)(ri.r0 + r) + (int64_t)(ri.c0 + c) * tile.c_ld;
    if (tile.accumulate) {
      const double2 old = to_d2(C[o]);
      re += old.x;
      im += old.y;
    }
    put(C + o, re, im);
  }
}

template <typename T, typename Bufs, int NA, int NB, bool CPLX3, int WBT>
static void launch_ts(const DevTasks& t, const Bufs& b, hipStream_t s, int tag, int c) {
  const int n = t.cls_off[c + 1] - t.cls_off[c];
  if (n <= 0) return;
  const WorkItem* w = t.work + t.cls_off[c];
  const dim3 g((n + WAVES - 1) / WAVES), blk(64 * WAVES);
#define HH_LAUNCH(TG) \
  hipLaunchKernelGGL((k_tiles<T, TG, Bufs, NA, NB, CPLX3, WBT>), g, blk, 0, s, t.tiles, t.terms, w, n, b, t.partial)
  if (tag == 1) HH_LAUNCH(1);
  else if (tag == 2) HH_LAUNCH(2);
  else if (tag == 3) HH_LAUNCH(3);
  else if (tag == 4) HH_LAUNCH(4);
  else HH_LAUNCH(0);
#undef HH_LAUNCH
}
template <typename T, typename Bufs>
static void launch_t(const DevTasks& t, const Bufs& b, hipStream_t s, int tag) {
  constexpr bool m3 = HH_CPLX3 != 0;
  launch_ts<T, Bufs, 2, 2, m3, WB>(t, b, s, tag, 0);
  launch_ts<T, Bufs, 1, 2, m3, WB>(t, b, s, tag, 1);
  launch_ts<T, Bufs, 2, 1, m3, WB>(t, b, s, tag, 2);
  launch_ts<T, Bufs, 1, 1, m3, WB>(t, b, s, tag, 3);
  if (t.cls_off[NCLASS] > t.cls_off[4]) {   // wide macro tiles
    launch_ts<T, Bufs, 2, WBW, m3, WBW>(t, b, s, tag, 4);
    launch_ts<T, Bufs, 1, WBW, m3, WBW>(t, b, s, tag, 5);
  }
}

template <typename Bufs>
static void launch_any(const DevTasks& t, const Bufs& b, hipStream_t s, int precision, int tag, hipEvent_t mid) {
  if (t.nwork > 0) {
    // complex64 storage always computes in FP32
    if constexpr (std::is_same<Bufs, BufPtrs32>::value) {
      launch_t<float>(t, b, s, tag);
    } else {
      if (precision == PREC_FP32) launch_t<float>(t, b, s, tag);
      else launch_t<double>(t, b, s, tag);
    }
    check_hip(hipGetLastError(), "k_tiles launch");
  }
  if (mid) check_hip(hipEventRecord(mid, s), "event");
  if (t.nwork == 0) return;
  if (t.nred) {
    hipLaunchKernelGGL(k_reduce<Bufs>, dim3(t.nred), dim3(256), 0, s, t.tiles, t.red, b, t.partial);
    check_hip(hipGetLastError(), "k_reduce launch");
  }
}

void launch_tiles(const DevTasks& t, const BufPtrs& b, hipStream_t s, int precision, int tag, hipEvent_t mid) {
  launch_any(t, b, s, precision, tag, mid);
}
void launch_tiles(const DevTasks& t, const BufTable& b, hipStream_t s, int precision, int tag, hipEvent_t mid) {
  launch_any(t, b, s, precision, tag, mid);
}
void launch_tiles(const DevTasks& t, const BufPtrs32& b, hipStream_t s, int tag, hipEvent_t mid) {
  launch_any(t, b, s, PREC_FP32, tag, mid);
}

// ----------------------------------------------------------------------------
// vector kernels
// ----------------------------------------------------------------------------
static constexpr int VB = 256;
static constexpr int NPART = 512;   // partial sums per vector (fixed => deterministic)

int rdot_partials() { return NPART; }

__global__ void k_zero(double2* p, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = make_double2(0.0, 0.0);
}
__global__ void k_scale(double2* p, double a, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    double2 v = p[i];
    p[i] = make_double2(v.x * a, v.y * a);
  }
}
__global__ void k_copy(double2* d, const double2* s, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    d[i] = s[i];
}
__global__ void k_lupd(double2* w, const double2* v1, double a, const double2* v2, double b, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    double2 x = w[i], y = v1[i];
    x.x -= a * y.x;
    x.y -= a * y.y;
    if (v2) {
      double2 z = v2[i];
      x.x -= b * z.x;
      x.y -= b * z.y;
    }
    w[i] = x;
  }
}

// partial[j*NPART + blk] = sum over the fixed slice blk of Re(conj(V_j) w)
__global__ __launch_bounds__(VB) void k_rdot_part(const double2* V, int64_t stride, const double2* w, int64_t n,
                                                  double* partial) {
  __shared__ double red[VB];
  const int j = blockIdx.y;
  const double2* v = V + (int64_t)j * stride;
  const int64_t chunk = (n + NPART - 1) / NPART;
  const int64_t lo = blockIdx.x * chunk, hi = min(n, lo + chunk);
  double acc = 0.0;
  for (int64_t i = lo + threadIdx.x; i < hi; i += VB) {
    double2 a = v[i], b = w[i];
    acc += a.x * b.x + a.y * b.y;
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s = VB / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[j * NPART + blockIdx.x] = red[0];
}
__global__ __launch_bounds__(VB) void k_rdot_final(const double* partial, double* out) {
  __shared__ double red[VB];
  const int j = blockIdx.x;
  double acc = 0.0;
  for (int i = threadIdx.x; i < NPART; i += VB) acc += partial[j * NPART + i];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s = VB / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[j] = red[0];
}
__global__ void k_combine(double2* w, double beta, const double2* V, int64_t stride, int nv, const double* c,
                          int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    double2 x = w[i];
    double re = beta * x.x, im = beta * x.y;
    for (int j = 0; j < nv; ++j) {
      double2 v = V[(int64_t)j * stride + i];
      re += c[j] * v.x;
      im += c[j] * v.y;
    }
    w[i] = make_double2(re, im);
  }
}

__device__ __forceinline__ uint64_t splitmix64(uint64_t x) {
  uint64_t z = x + 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
__global__ void k_start(double2* p, int64_t n, uint64_t base) {
  const double scale = 1.0 / 9007199254740992.0;
  for (int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; g < n; g += (int64_t)gridDim.x * blockDim.x) {
    double re = (double)(splitmix64(base + 2 * (uint64_t)g) >> 11) * scale;
    double im = (double)(splitmix64(base + 2 * (uint64_t)g + 1) >> 11) * scale;
    p[g] = make_double2(re, im);
  }
}

// ----------------------------------------------------------------------------
// fused Lanczos kernels.  A vector of length n is cut into blocks of
// VB*EPT consecutive elements (thread t of block b owns elements
// b*VB*EPT + t + VB*q, q < EPT); every reduction sums a thread's elements in
// q order, the 64 lanes of a wave by xor-shuffle, the 4 waves in order, and
// the blocks in order in a second kernel: deterministic for a given n.
// ----------------------------------------------------------------------------
static constexpr int EPT = 8;
static constexpr int MDOT_J = 64;   // vectors per pass of the multi-dot

int vec_blocks(int64_t n) { return (int)std::max<int64_t>(1, (n + VB * EPT - 1) / (VB * EPT)); }

__device__ __forceinline__ double wave_sum(double x) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
  return x;
}

// partial[j * nblk + b] = sum over block b of Re(conj(V_j) w), j < nv.
// w is read once into registers; V_j streams past it.
__global__ __launch_bounds__(VB) void k_mdot_part(const double2* __restrict__ V, int64_t stride, int nv,
                                                  const double2* __restrict__ w, int64_t n,
                                                  double* __restrict__ partial) {
  __shared__ double red[MDOT_J][VB / 64];
  const int64_t base = (int64_t)blockIdx.x * VB * EPT + threadIdx.x;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  double2 x[EPT];
#pragma unroll
  for (int q = 0; q < EPT; ++q) {
    const int64_t i = base + (int64_t)VB * q;
    x[q] = i < n ? w[i] : make_double2(0.0, 0.0);
  }
  for (int j0 = 0; j0 < nv; j0 += MDOT_J) {
    const int jn = min(MDOT_J, nv - j0);
    for (int jj = 0; jj < jn; ++jj) {
      const double2* __restrict__ v = V + (int64_t)(j0 + jj) * stride;
      double acc = 0.0;
#pragma unroll
      for (int q = 0; q < EPT; ++q) {
        const int64_t i = base + (int64_t)VB * q;
        if (i < n) {
          const double2 a = v[i];
          acc += a.x * x[q].x + a.y * x[q].y;
        }
      }
      acc = wave_sum(acc);
      if (lane == 0) red[jj][wv] = acc;
    }
    __syncthreads();
    if ((int)threadIdx.x < jn) {
      double t = 0.0;
#pragma unroll
      for (int k = 0; k < VB / 64; ++k) t += red[threadIdx.x][k];
      partial[(int64_t)(j0 + threadIdx.x) * gridDim.x + blockIdx.x] = t;
    }
    __syncthreads();
  }
}

// out[j] = sum over b of partial[j * nblk + b] (one block per j)
__global__ __launch_bounds__(VB) void k_block_sum(const double* __restrict__ partial, int nblk,
                                                  double* __restrict__ out) {
  __shared__ double red[VB / 64];
  const int j = blockIdx.x;
  double acc = 0.0;
  for (int i = threadIdx.x; i < nblk; i += VB) acc += partial[(int64_t)j * nblk + i];
  acc = wave_sum(acc);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int k = 0; k < VB / 64; ++k) t += red[k];
    out[j] = t;
  }
}

// block sum of |y|^2 over this thread's elements, written to partial[blockIdx.x]
__device__ __forceinline__ void block_norm(double acc, double* __restrict__ partial) {
  __shared__ double red[VB / 64];
  acc = wave_sum(acc);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int k = 0; k < VB / 64; ++k) t += red[k];
    partial[blockIdx.x] = t;
  }
}

// w = w - a v1 - b v2 (v2 may be null); partial |w|^2 per block.  With a_dev
// set, a is read from device memory (a scalar the stream wrote earlier)
// instead of the argument.
__global__ __launch_bounds__(VB) void k_lupd_norm(double2* __restrict__ w, const double2* __restrict__ v1,
                                                  double a_arg, const double* __restrict__ a_dev,
                                                  const double2* __restrict__ v2, double b, int64_t n,
                                                  double* __restrict__ partial) {
  const double a = a_dev ? *a_dev : a_arg;
  const int64_t base = (int64_t)blockIdx.x * VB * EPT + threadIdx.x;
  double acc = 0.0;
#pragma unroll
  for (int q = 0; q < EPT; ++q) {
    const int64_t i = base + (int64_t)VB * q;
    if (i < n) {
      double2 x = w[i];
      const double2 y = v1[i];
      x.x -= a * y.x;
      x.y -= a * y.y;
      if (v2) {
        const double2 z = v2[i];
        x.x -= b * z.x;
        x.y -= b * z.y;
      }
      w[i] = x;
      acc += x.x * x.x + x.y * x.y;
    }
  }
  block_norm(acc, partial);
}

// w = sw * w + cs * sum_j c_j V_j (c on the device); partial |w|^2 per block
__global__ __launch_bounds__(VB) void k_combine_norm(double2* __restrict__ w, double sw,
                                                     const double2* __restrict__ V, int64_t stride, int nv,
                                                     const double* __restrict__ c, double cs, int64_t n,
                                                     double* __restrict__ partial) {
  const int64_t base = (int64_t)blockIdx.x * VB * EPT + threadIdx.x;
  double2 x[EPT];
#pragma unroll
  for (int q = 0; q < EPT; ++q) {
    const int64_t i = base + (int64_t)VB * q;
    x[q] = make_double2(0.0, 0.0);
    if (i < n && sw != 0.0) {
      const double2 y = w[i];
      x[q] = make_double2(sw * y.x, sw * y.y);
    }
  }
  for (int j = 0; j < nv; ++j) {
    const double cj = cs * c[j];
    const double2* __restrict__ v = V + (int64_t)j * stride;
#pragma unroll
    for (int q = 0; q < EPT; ++q) {
      const int64_t i = base + (int64_t)VB * q;
      if (i < n) {
        const double2 y = v[i];
        x[q].x += cj * y.x;
        x[q].y += cj * y.y;
      }
    }
  }
  double acc = 0.0;
#pragma unroll
  for (int q = 0; q < EPT; ++q) {
    const int64_t i = base + (int64_t)VB * q;
    if (i < n) {
      w[i] = x[q];
      acc += x[q].x * x[q].x + x[q].y * x[q].y;
    }
  }
  block_norm(acc, partial);
}

// w = sum_j (cv_j V_j + cw_j W_j) (coefficients on the device); partial |w|^2
__global__ __launch_bounds__(VB) void k_combine2_norm(double2* __restrict__ w, const double2* __restrict__ V,
                                                      const double2* __restrict__ W, int64_t stride, int nv,
                                                      const double* __restrict__ cv, const double* __restrict__ cw,
                                                      int64_t n, double* __restrict__ partial) {
  const int64_t base = (int64_t)blockIdx.x * VB * EPT + threadIdx.x;
  double2 x[EPT];
#pragma unroll
  for (int q = 0; q < EPT; ++q) x[q] = make_double2(0.0, 0.0);
  for (int j = 0; j < nv; ++j) {
    const double a = cv[j], b = cw[j];
    const double2* __restrict__ v = V + (int64_t)j * stride;
    const double2* __restrict__ u = W + (int64_t)j * stride;
#pragma unroll
    for (int q = 0; q < EPT; ++q) {
      const int64_t i = base + (int64_t)VB * q;
      if (i < n) {
        const double2 y = v[i], z = u[i];
        x[q].x += a * y.x + b * z.x;
        x[q].y += a * y.y + b * z.y;
      }
    }
  }
  double acc = 0.0;
#pragma unroll
  for (int q = 0; q < EPT; ++q) {
    const int64_t i = base + (int64_t)VB * q;
    if (i < n) {
      w[i] = x[q];
      acc += x[q].x * x[q].x + x[q].y * x[q].y;
    }
  }
  block_norm(acc, partial);
}

__global__ __launch_bounds__(VB) void k_scale_copy(double2* __restrict__ d, const double2* __restrict__ src, double a,
                                                   int64_t n) {
  const int64_t base = (int64_t)blockIdx.x * VB * EPT + threadIdx.x;
#pragma unroll
  for (int q = 0; q < EPT; ++q) {
    const int64_t i = base + (int64_t)VB * q;
    if (i < n) {
      const double2 y = src[i];
      d[i] = make_double2(a * y.x, a * y.y);
    }
  }
}

static inline int grid_for(int64_t n) { return (int)std::min<int64_t>((n + VB - 1) / VB, 8192); }

void dev_zero(double2* p, int64_t n, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_zero, dim3(grid_for(n)), dim3(VB), 0, s, p, n);
}
void dev_scale(double2* p, double a, int64_t n, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_scale, dim3(grid_for(n)), dim3(VB), 0, s, p, a, n);
}
void dev_copy(double2* d, const double2* src, int64_t n, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_copy, dim3(grid_for(n)), dim3(VB), 0, s, d, src, n);
}
void dev_lanczos_update(double2* w, const double2* v1, double a, const double2* v2, double b, int64_t n,
                        hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_lupd, dim3(grid_for(n)), dim3(VB), 0, s, w, v1, a, v2, b, n);
}
void dev_rdots(const double2* V, int64_t stride, int nv, const double2* w, int64_t n, double* out_dev,
               double* partial_dev, hipStream_t s) {
  if (nv <= 0) return;
  hipLaunchKernelGGL(k_rdot_part, dim3(NPART, nv), dim3(VB), 0, s, V, stride, w, n, partial_dev);
  hipLaunchKernelGGL(k_rdot_final, dim3(nv), dim3(VB), 0, s, partial_dev, out_dev);
}
void dev_combine(double2* w, double beta, const double2* V, int64_t stride, int nv, const double* c_dev, int64_t n,
                 hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_combine, dim3(grid_for(n)), dim3(VB), 0, s, w, beta, V, stride, nv, c_dev, n);
}
void dev_mdots(const double2* V, int64_t stride, int nv, const double2* w, int64_t n, double* out_dev,
               double* partial_dev, hipStream_t s) {
  if (nv <= 0) return;
  const int nb = vec_blocks(n);
  hipLaunchKernelGGL(k_mdot_part, dim3(nb), dim3(VB), 0, s, V, stride, nv, w, n, partial_dev);
  hipLaunchKernelGGL(k_block_sum, dim3(nv), dim3(VB), 0, s, partial_dev, nb, out_dev);
}
void dev_lanczos_update_norm(double2* w, const double2* v1, double a, const double2* v2, double b, int64_t n,
                             double* out_dev, double* partial_dev, hipStream_t s, const double* a_dev) {
  const int nb = vec_blocks(n);
  hipLaunchKernelGGL(k_lupd_norm, dim3(nb), dim3(VB), 0, s, w, v1, a, a_dev, v2, b, n, partial_dev);
  hipLaunchKernelGGL(k_block_sum, dim3(1), dim3(VB), 0, s, partial_dev, nb, out_dev);
}
void dev_combine_norm(double2* w, double sw, const double2* V, int64_t stride, int nv, const double* c_dev, double cs,
                      int64_t n, double* out_dev, double* partial_dev, hipStream_t s) {
  const int nb = vec_blocks(n);
  hipLaunchKernelGGL(k_combine_norm, dim3(nb), dim3(VB), 0, s, w, sw, V, stride, nv, c_dev, cs, n, partial_dev);
  hipLaunchKernelGGL(k_block_sum, dim3(1), dim3(VB), 0, s, partial_dev, nb, out_dev);
}
void dev_combine2_norm(double2* w, const double2* V, const double2* W, int64_t stride, int nv, const double* cv_dev,
                       const double* cw_dev, int64_t n, double* out_dev, double* partial_dev, hipStream_t s) {
  const int nb = vec_blocks(n);
  hipLaunchKernelGGL(k_combine2_norm, dim3(nb), dim3(VB), 0, s, w, V, W, stride, nv, cv_dev, cw_dev, n, partial_dev);
  hipLaunchKernelGGL(k_block_sum, dim3(1), dim3(VB), 0, s, partial_dev, nb, out_dev);
}
void dev_scale_copy(double2* d, const double2* src, double a, int64_t n, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_scale_copy, dim3(vec_blocks(n)), dim3(VB), 0, s, d, src, a, n);
}
// ----------------------------------------------------------------------------
// segmented vector kernels: several independent vectors stored one after the
// other, each starting at a multiple of VB * EPT elements, so that every
// block lies inside one segment.  bseg[b] is the segment of block b, send[s]
// the end (exclusive) of segment s's elements; padding past it is never read
// or written.  Within a segment, threads, blocks and the reduction order are
// those of the unsegmented kernels above, so a segment's sums equal those of
// the same vector alone bit for bit.
// ----------------------------------------------------------------------------
int vec_span() { return VB * EPT; }

__global__ __launch_bounds__(VB) void k_smdot_part(const double2* __restrict__ V, int64_t stride, int nv,
                                                   const double2* __restrict__ w, const int32_t* __restrict__ bseg,
                                                   const int64_t* __restrict__ send,
                                                   const int32_t* __restrict__ active, double* __restrict__ partial) {
  __shared__ double red[MDOT_J][VB / 64];
  if (active && active[bseg[blockIdx.x]] == 0) {   // block-uniform
    for (int j = threadIdx.x; j < nv; j += VB) partial[(int64_t)j * gridDim.x + blockIdx.x] = 0.0;
    return;
  }
  const int64_t n = send[bseg[blockIdx.x]];
  const int64_t base = (int64_t)blockIdx.x * VB * EPT + threadIdx.x;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  double2 x[EPT];
#pragma unroll
  for (int q = 0; q < EPT; ++q) {
    const int64_t i = base + (int64_t)VB * q;
    x[q] = i < n ? w[i] : make_double2(0.0, 0.0);
  }
  for (int j0 = 0; j0 < nv; j0 += MDOT_J) {
    const int jn = min(MDOT_J, nv - j0);
    for (int jj = 0; jj < jn; ++jj) {
      const double2* __restrict__ v = V + (int64_t)(j0 + jj) * stride;
      double acc = 0.0;
#pragma unroll
      for (int q = 0; q < EPT; ++q) {
        const int64_t i = base + (int64_t)VB * q;
        if (i < n) {
          const double2 a = v[i];
          acc += a.x * x[q].x + a.y * x[q].y;
        }
      }
      acc = wave_sum(acc);
      if (lane == 0) red[jj][wv] = acc;
    }
    __syncthreads();
    if ((int)threadIdx.x < jn) {
      double t = 0.0;
#pragma unroll
      for (int k = 0; k < VB / 64; ++k) t += red[threadIdx.x][k];
      partial[(int64_t)(j0 + threadIdx.x) * gridDim.x + blockIdx.x] = t;
    }
    __syncthreads();
  }
}

// out[s * nv + j] = sum over the blocks of segment s of partial[j * nb + b],
// in the order of k_block_sum (one workgroup per (s, j))
__global__ __launch_bounds__(VB) void k_seg_sum(const double* __restrict__ partial, int nb, int nv,
                                                const int32_t* __restrict__ sfirst, double* __restrict__ out) {
  __shared__ double red[VB / 64];
  const int s = blockIdx.x / nv, j = blockIdx.x % nv;
  const int b0 = sfirst[s], nblk = sfirst[s + 1] - b0;
  double acc = 0.0;
  for (int i = threadIdx.x; i < nblk; i += VB) acc += partial[(int64_t)j * nb + b0 + i];
  acc = wave_sum(acc);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int k = 0; k < VB / 64; ++k) t += red[k];
    out[(int64_t)s * nv + j] = t;
  }
}

// w = w - a_s v1 - b_s v2 (v2 may be null); partial |w|^2 per block
__global__ __launch_bounds__(VB) void k_slupd_norm(double2* __restrict__ w, const double2* __restrict__ v1,
                                                   const double* __restrict__ a, const double2* __restrict__ v2,
                                                   const double* __restrict__ b, const int32_t* __restrict__ bseg,
                                                   const int64_t* __restrict__ send,
                                                   const int32_t* __restrict__ active, double* __restrict__ partial) {
  const int sg = bseg[blockIdx.x];
  if (active && active[sg] == 0) {   // block-uniform
    if (threadIdx.x == 0) partial[blockIdx.x] = 0.0;
    return;
  }
  const int64_t n = send[sg];
  const double as = a[sg], bs = v2 ? b[sg] : 0.0;
  const int64_t base = (int64_t)blockIdx.x * VB * EPT + threadIdx.x;
  double acc = 0.0;
#pragma unroll
  for (int q = 0; q < EPT; ++q) {
    const int64_t i = base + (int64_t)VB * q;
    if (i < n) {
      double2 x = w[i];
      const double2 y = v1[i];
      x.x -= as * y.x;
      x.y -= as * y.y;
      if (v2) {
        const double2 z = v2[i];
        x.x -= bs * z.x;
        x.y -= bs * z.y;
      }
      w[i] = x;
      acc += x.x * x.x + x.y * x.y;
    }
  }
  block_norm(acc, partial);
}

// w = sw_s w + cs_s sum_j c[s * ldc + j] V_j; partial |w|^2 per block
__global__ __launch_bounds__(VB) void k_scombine_norm(double2* __restrict__ w, const double* __restrict__ sw,
                                                      const double2* __restrict__ V, int64_t stride, int nv,
                                                      const double* __restrict__ c, int ldc,
                                                      const double* __restrict__ cs,
                                                      const int32_t* __restrict__ bseg,
                                                      const int64_t* __restrict__ send, double* __restrict__ partial) {
  const int sg = bseg[blockIdx.x];
  const int64_t n = send[sg];
  const double sws = sw[sg], css = cs[sg];
  const double* __restrict__ cc = c + (int64_t)sg * ldc;
  const int64_t base = (int64_t)blockIdx.x * VB * EPT + threadIdx.x;
  double2 x[EPT];
#pragma unroll
  for (int q = 0; q < EPT; ++q) {
    const int64_t i = base + (int64_t)VB * q;
    x[q] = make_double2(0.0, 0.0);
    if (i < n && sws != 0.0) {
      const double2 y = w[i];
      x[q] = make_double2(sws * y.x, sws * y.y);
    }
  }
  for (int j = 0; j < nv; ++j) {
    const double cj = css * cc[j];
    if (cj == 0.0) continue;   // block-uniform; 0 * v adds nothing (and leaves unused vectors unread)
    const double2* __restrict__ v = V + (int64_t)j * stride;
#pragma unroll
    for (int q = 0; q < EPT; ++q) {
      const int64_t i = base + (int64_t)VB * q;
      if (i < n) {
        const double2 y = v[i];
        x[q].x += cj * y.x;
        x[q].y += cj * y.y;
      }
    }
  }
  double acc = 0.0;
#pragma unroll
  for (int q = 0; q < EPT; ++q) {
    const int64_t i = base + (int64_t)VB * q;
    if (i < n) {
      w[i] = x[q];
      acc += x[q].x * x[q].x + x[q].y * x[q].y;
    }
  }
  block_norm(acc, partial);
}

// d1 = a_s src and, when d2 is set, d2 = the same
__global__ __launch_bounds__(VB) void k_sscale2(double2* __restrict__ d1, double2* __restrict__ d2,
                                                const double2* __restrict__ src, const double* __restrict__ a,
                                                const int32_t* __restrict__ bseg, const int64_t* __restrict__ send) {
  const int sg = bseg[blockIdx.x];
  const int64_t n = send[sg];
  const double as = a[sg];
  const int64_t base = (int64_t)blockIdx.x * VB * EPT + threadIdx.x;
#pragma unroll
  for (int q = 0; q < EPT; ++q) {
    const int64_t i = base + (int64_t)VB * q;
    if (i < n) {
      const double2 y = src[i];
      const double2 r = make_double2(as * y.x, as * y.y);
      d1[i] = r;
      if (d2) d2[i] = r;
    }
  }
}

void dev_seg_mdots(const double2* V, int64_t stride, int nv, const double2* w, const SegDev& g, double* out_dev,
                   double* partial_dev, hipStream_t s) {
  if (nv <= 0) return;
  hipLaunchKernelGGL(k_smdot_part, dim3(g.nb), dim3(VB), 0, s, V, stride, nv, w, g.bseg, g.send, g.active,
                     partial_dev);
  hipLaunchKernelGGL(k_seg_sum, dim3(g.nseg * nv), dim3(VB), 0, s, partial_dev, g.nb, nv, g.sfirst, out_dev);
}
void dev_seg_lupd_norm(double2* w, const double2* v1, const double* a_dev, const double2* v2, const double* b_dev,
                       const SegDev& g, double* out_dev, double* partial_dev, hipStream_t s) {
  hipLaunchKernelGGL(k_slupd_norm, dim3(g.nb), dim3(VB), 0, s, w, v1, a_dev, v2, b_dev, g.bseg, g.send, g.active,
                     partial_dev);
  hipLaunchKernelGGL(k_seg_sum, dim3(g.nseg), dim3(VB), 0, s, partial_dev, g.nb, 1, g.sfirst, out_dev);
}
void dev_seg_combine_norm(double2* w, const double* sw_dev, const double2* V, int64_t stride, int nv,
                          const double* c_dev, int ldc, const double* cs_dev, const SegDev& g, double* out_dev,
                          double* partial_dev, hipStream_t s) {
  hipLaunchKernelGGL(k_scombine_norm, dim3(g.nb), dim3(VB), 0, s, w, sw_dev, V, stride, nv, c_dev, ldc, cs_dev,
                     g.bseg, g.send, partial_dev);
  hipLaunchKernelGGL(k_seg_sum, dim3(g.nseg), dim3(VB), 0, s, partial_dev, g.nb, 1, g.sfirst, out_dev);
}
void dev_seg_scale2(double2* d1, double2* d2, const double2* src, const double* a_dev, const SegDev& g,
                    hipStream_t s) {
  hipLaunchKernelGGL(k_sscale2, dim3(g.nb), dim3(VB), 0, s, d1, d2, src, a_dev, g.bseg, g.send);
}

// d = (ar + i ai) src
__global__ __launch_bounds__(VB) void k_cscale_copy(double2* __restrict__ d, const double2* __restrict__ src,
                                                    double ar, double ai, int64_t n) {
  const int64_t base = (int64_t)blockIdx.x * VB * EPT + threadIdx.x;
#pragma unroll
  for (int q = 0; q < EPT; ++q) {
    const int64_t i = base + (int64_t)VB * q;
    if (i < n) {
      const double2 y = src[i];
      d[i] = make_double2(ar * y.x - ai * y.y, ar * y.y + ai * y.x);
    }
  }
}
void dev_cscale_copy(double2* d, const double2* src, double ar, double ai, int64_t n, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_cscale_copy, dim3(vec_blocks(n)), dim3(VB), 0, s, d, src, ar, ai, n);
}
__global__ __launch_bounds__(VB) void k_to_f32(float2* __restrict__ d, const double2* __restrict__ src, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const double2 y = src[i];
    d[i] = make_float2((float)y.x, (float)y.y);
  }
}
__global__ __launch_bounds__(VB) void k_to_f64(double2* __restrict__ d, const float2* __restrict__ src, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const float2 y = src[i];
    d[i] = make_double2((double)y.x, (double)y.y);
  }
}
__global__ __launch_bounds__(VB) void k_conj(double2* __restrict__ p, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    p[i].y = -p[i].y;
}
void dev_conj(double2* p, int64_t n, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_conj, dim3(grid_for(n)), dim3(VB), 0, s, p, n);
}
void dev_to_f32(float2* d, const double2* src, int64_t n, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_to_f32, dim3(grid_for(n)), dim3(VB), 0, s, d, src, n);
}
void dev_to_f64(double2* d, const float2* src, int64_t n, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_to_f64, dim3(grid_for(n)), dim3(VB), 0, s, d, src, n);
}

void dev_start_vector(double2* p, int64_t n, uint64_t seed, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_start, dim3(grid_for(n)), dim3(VB), 0, s, p, n, (seed & 0xFFFFFFFFull) << 32);
}

}  // namespace hh
