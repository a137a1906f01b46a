// Autotuning of the grouped H.psi launches, keyed on the sector-size
// histogram of the task list.
//
// How a task list is cut into work items (LaunchConfig: work items per
// launch, term groups per macro tile, shape classes for edge items, band
// width of the macro-tile order) sets how evenly the waves are loaded, how
// many partial sums are written and reduced, and how well neighbouring work
// items share operand panels in L2.  The best cut depends on the shapes of
// the GEMMs: a few large sectors (the DMRG steps at m = 2000) or many small
// ones (the batched synthetic superblocks).  The tuner times each stage's
// task list under a small set of configurations on the superblock's own
// buffers and keeps the fastest; the choice is cached under a key made from
// the flop-weighted histogram of the list's (M, N, K) shapes in factors of
// two, so later superblocks of the same shape statistics (the next sweep
// steps, the next batch of a scan) reuse it without timing.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <map>
#include <vector>

#include "hh_runtime.hpp"

namespace hh {

namespace {

int lg2(int64_t x) {
  int k = 0;
  while (((int64_t)1 << k) < x) ++k;
  return k;
}

void mix(uint64_t& h, uint64_t v) {
  h ^= v + 0x9E3779B97F4A7C15ull + (h << 6) + (h >> 2);
  h *= 0x100000001B3ull;
}

// the candidates: the fixed configuration first, then the grid
std::vector<LaunchConfig> candidates() {
  std::vector<LaunchConfig> c{LaunchConfig()};
  for (int edge : {0, 1})
    for (int ti : {32768, 65536, 131072})
      for (int mg : {16, 32, 64}) {
        LaunchConfig x;
        x.edge = edge;
        x.target_items = ti;
        x.max_groups = mg;
        c.push_back(x);
      }
  for (int band : {2, 8}) {
    LaunchConfig x;
    x.band = band;
    c.push_back(x);
  }
  // wide 32 x 48 tiles for regions that end in a short column chunk
  for (int wide : {1, 2})
    for (int edge : {0, 1})
      for (int ti : {32768, 65536}) {
        LaunchConfig x;
        x.wide = wide;
        x.edge = edge;
        x.target_items = ti;
        c.push_back(x);
      }
  return c;
}

// average device time (ms) of one launch of tl cut by cfg
template <typename Bufs>
double time_config(hh_ctx* ctx, const TaskList& tl, const LaunchConfig& cfg, const Bufs& bufs, int tag,
                   DevTasks& scratch, hipEvent_t e0, hipEvent_t e1) {
  hipStream_t s = ctx->stream;
  scratch.upload(tl, s, nullptr, 0, cfg);
  launch_tiles(scratch, bufs, s, PREC_FP64, tag, nullptr);   // first launch: code and caches warm
  constexpr int REPS = 3;
  check_hip(hipEventRecord(e0, s), "event");
  for (int r = 0; r < REPS; ++r) launch_tiles(scratch, bufs, s, PREC_FP64, tag, nullptr);
  check_hip(hipEventRecord(e1, s), "event");
  check_hip(hipEventSynchronize(e1), "event sync");
  float ms = 0.f;
  check_hip(hipEventElapsedTime(&ms, e0, e1), "elapsed");
  return ms / REPS;
}

// The configuration of one stage's task list: from the cache, or timed.
// A candidate replaces the fixed configuration only when it is at least 2 %
// faster, so that timing noise does not move the choice.
template <typename Bufs>
LaunchConfig choose(hh_ctx* ctx, const TaskList& tl, const Bufs& bufs, int tag, uint64_t salt) {
  const uint64_t key = shape_key(tl) ^ salt;
  auto it = ctx->tune_cache.find(key);
  if (it != ctx->tune_cache.end()) {
    ++ctx->tune_hits;
    return it->second;
  }
  const double t0 = wall();
  hipEvent_t e0 = nullptr, e1 = nullptr;
  check_hip(hipEventCreate(&e0), "event");
  check_hip(hipEventCreate(&e1), "event");
  LaunchConfig best;
  try {
    DevTasks scratch;
    const std::vector<LaunchConfig> c = candidates();
    const double base = time_config(ctx, tl, c[0], bufs, tag, scratch, e0, e1);
    double best_ms = base;
    for (size_t i = 1; i < c.size(); ++i) {
      const double ms = time_config(ctx, tl, c[i], bufs, tag, scratch, e0, e1);
      if (ms < best_ms) {
        best_ms = ms;
        best = c[i];
      }
    }
    if (best_ms > 0.98 * base) best = LaunchConfig();
    if (verbose())
      std::fprintf(stderr, "[hhdmrg] autotune stage %d: %.3f ms fixed, %.3f ms best (items %d, groups %d, edge %d, "
                   "band %d, wide %d)\n", tag, base, best_ms, best.target_items, best.max_groups, best.edge, best.band,
                   best.wide);
  } catch (...) {
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    throw;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  ctx->tune_cache[key] = best;
  ++ctx->tune_runs;
  ctx->tune_seconds += wall() - t0;
  return best;
}

}  // namespace

uint64_t shape_key(const TaskList& tl) {
  // flop-weighted histogram of (log2 M, log2 N, log2 K) of the GEMM terms,
  // each bucket's share rounded to 1/8 of the total, and log2 of the total
  std::map<uint32_t, double> h;
  double tot = 0.0;
  int64_t elementwise = 0;
  for (const Tile& t : tl.tiles)
    for (int k = t.t0; k < t.t1; ++k) {
      const Term& tm = tl.terms[k];
      if (tm.K <= 0) {
        ++elementwise;
        continue;
      }
      const double f = 8.0 * (double)t.M * (double)t.N * (double)tm.K;
      h[(uint32_t)(lg2(t.M) << 16 | lg2(t.N) << 8 | lg2(tm.K))] += f;
      tot += f;
    }
  uint64_t key = 0xCBF29CE484222325ull;
  mix(key, (uint64_t)lg2((int64_t)std::max(tot, 1.0)));
  mix(key, (uint64_t)lg2(elementwise + 1));
  for (const auto& kv : h) {
    const int q = (int)std::lround(8.0 * kv.second / tot);
    if (q == 0) continue;
    mix(key, kv.first);
    mix(key, (uint64_t)q);
  }
  return key;
}

void autotune_superblock(hh_ctx* ctx, hh_superblock* sb) {
  const int64_t n = sb->plan.sb.dim;
  DevMem in, out;
  in.alloc(sizeof(double2) * (size_t)std::max<int64_t>(n, 1));
  out.alloc(sizeof(double2) * (size_t)std::max<int64_t>(n, 1));
  check_hip(hipMemsetAsync(in.p, 0, sizeof(double2) * (size_t)std::max<int64_t>(n, 1), ctx->stream), "memset");
  const BufPtrs b = bufs_of(in.c(), out.c(), sb->sys, sb->env, sb->X.c(), sb->merged.c(), nullptr, nullptr);
  // stage 1 first: it leaves X = A . 0 = 0 for the stage-2 timings
  const LaunchConfig c1 = choose(ctx, sb->plan.stage1, b, 1, 0x1111);
  if (!(c1 == sb->st1.cfg)) sb->st1.upload(sb->plan.stage1, ctx->stream, nullptr, 0, c1);
  const LaunchConfig c2 = choose(ctx, sb->plan.stage2, b, 2, 0x2222);
  if (!(c2 == sb->st2.cfg)) sb->st2.upload(sb->plan.stage2, ctx->stream, nullptr, 0, c2);
  check_hip(hipStreamSynchronize(ctx->stream), "autotune");
}

void autotune_batch(hh_ctx* ctx, hh_batch* b, const TaskList& s1, const TaskList& s2) {
  const int nm = (int)b->sbs.size();
  std::vector<DevMem> in(nm), out(nm);
  for (int i = 0; i < nm; ++i) {
    const size_t bytes = sizeof(double2) * (size_t)std::max<int64_t>(b->sbs[i]->plan.sb.dim, 1);
    in[i].alloc(bytes);
    out[i].alloc(bytes);
    check_hip(hipMemsetAsync(in[i].p, 0, bytes, ctx->stream), "memset");
    const hh_superblock* sb = b->sbs[i];
    const BufPtrs p = bufs_of(in[i].c(), out[i].c(), sb->sys, sb->env, sb->X.c(), sb->merged.c(), nullptr, nullptr);
    for (int k = 0; k < 8; ++k) b->host_table[8 * i + k] = p.p[k];
  }
  check_hip(hipMemcpyAsync(b->table.p, b->host_table.data(), sizeof(double2*) * b->host_table.size(),
                           hipMemcpyHostToDevice, ctx->stream), "h2d table");
  const BufTable t{(double2* const*)b->table.p};
  const LaunchConfig c1 = choose(ctx, s1, t, 1, 0x3333);
  if (!(c1 == b->st1.cfg)) b->st1.upload(s1, ctx->stream, nullptr, 0, c1);
  const LaunchConfig c2 = choose(ctx, s2, t, 2, 0x4444);
  if (!(c2 == b->st2.cfg)) b->st2.upload(s2, ctx->stream, nullptr, 0, c2);
  check_hip(hipStreamSynchronize(ctx->stream), "autotune");
}

}  // namespace hh

using namespace hh;

extern "C" {

int hh_set_autotune(hh_ctx* ctx, int mode) {
  if (!ctx) return 1;
  return guard(ctx, [&] {
    need(mode == 0 || mode == 1, "autotune mode must be 0 or 1");
    ctx->autotune = mode;
  });
}

int hh_set_launch_config(hh_ctx* ctx, int target_items, int max_groups, int edge, int band, int wide) {
  if (!ctx) return 1;
  return guard(ctx, [&] {
    if (target_items == 0) {
      ctx->forced_cfg = false;
      return;
    }
    LaunchConfig c;
    c.target_items = target_items;
    c.max_groups = max_groups;
    c.edge = edge;
    c.band = band;
    c.wide = wide;
    need(target_items >= 1 && max_groups >= 1 && max_groups <= 1024 && band >= 1 && edge >= -1 && edge <= 1 &&
             wide >= 0 && wide <= 2,
         "launch configuration out of range");
    ctx->forced = c;
    ctx->forced_cfg = true;
  });
}

int hh_autotune_stats(const hh_ctx* ctx, int32_t* keys, int32_t* tunes, int32_t* hits, double* seconds) {
  if (!ctx) return 1;
  if (keys) *keys = (int32_t)ctx->tune_cache.size();
  if (tunes) *tunes = ctx->tune_runs;
  if (hits) *hits = ctx->tune_hits;
  if (seconds) *seconds = ctx->tune_seconds;
  return 0;
}

}  // extern "C"
