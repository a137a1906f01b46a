// Device engine: grouped complex GEMM over tile lists (FP64 / FP32 MFMA on
// gfx950) and the deterministic vector kernels of the Lanczos solver.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "hh_core.hpp"

namespace hh {

struct BufPtrs { double2* p[8]; };
// Buffer pointers held in device memory (up to 256 slots): lets one launch
// address the buffers of many superblocks.  With `active` set, the tiles of
// member k (buffer slots 8k .. 8k+7) run only while active[k] != 0.
struct BufTable {
  double2* const* p;
  const int32_t* active = nullptr;
};
// complex64 buffer slots: the FP32-storage form of a task list (operands,
// intermediates and outputs all single precision, FP32 MFMA)
struct BufPtrs32 { float2* p[8]; };
// One wave's share of the work: a 32x32 macro tile of output region `tile`
// and the term range [t0, t1); slot >= 0 sends the partial sum to the
// partial buffer (tiles whose term list is split across several waves),
// starting at strip `slot` (a strip is 32 x 16 complex elements).
struct WorkItem { int32_t tile, r0, c0, t0, t1, slot; };
// Deterministic reduction of split macro tiles: C = sum of nslot partials,
// the q-th starting at strip slot0 + q * width / 16.
struct ReduceItem { int32_t tile, r0, c0, slot0, nslot, width; };

// How a task list is cut into work items: about target_items work items of
// equal K-work per launch, at most max_groups term groups per macro tile,
// shape classes for edge items (edge 1 always, 0 never, -1 when more than a
// fifth of the MFMA work would fall on empty 16 x 16 sub-tiles), macro tiles
// listed in bands of `band` macro rows, and the column chunks: wide 0 all
// 32 wide, 1 a region's last chunk of at most 16 columns merged into the one
// before it (a 32 x 48 tile), 2 all 48 wide.
struct LaunchConfig {
  int target_items = 65536;
  int max_groups = 32;
  int edge = -1;
  int band = 4;
  int wide = 0;
  bool operator==(const LaunchConfig& o) const {
    return target_items == o.target_items && max_groups == o.max_groups && edge == o.edge && band == o.band &&
           wide == o.wide;
  }
};
// shape classes of the work items (each its own kernel instantiation)
constexpr int NCLASS = 6;

// Tile list resident on the device, launched as one grouped kernel.
struct DevTasks {
  Tile* tiles = nullptr;
  Term* terms = nullptr;
  WorkItem* work = nullptr;
  ReduceItem* red = nullptr;
  double2* partial = nullptr;
  int ntiles = 0, nterms = 0, nwork = 0, nred = 0, nslots = 0;
  double flops_alg = 0.0, flops_exec = 0.0;
  // work items grouped by shape class, each class run by its own kernel:
  // 0 full 32 x 32, 1 at most 16 rows, 2 at most 16 columns, 3 both, 4 wide
  // 32 x 48, 5 wide with at most 16 rows; items
  // of class c are work[cls_off[c] .. cls_off[c + 1]).  Edge items are
  // classed only when more than a set share of the MFMA work would fall on
  // empty 16 x 16 sub-tiles; otherwise every item is class 0.
  int cls_off[NCLASS + 1] = {0, 0, 0, 0, 0, 0, 0};
  bool edge_skip = false;
  // borrowed: the arrays live in caller-provided device memory (never freed here)
  bool borrowed = false;
  LaunchConfig cfg;   // the configuration of the last upload
  // Device arrays from hipMalloc, or -- when `place` is given -- carved out of
  // the `cap` bytes at `place` (throws HHCapacity when they do not fit).
  void upload(const TaskList& tl, hipStream_t s, void* place = nullptr, size_t cap = 0,
              const LaunchConfig& config = LaunchConfig());
  void release();
  ~DevTasks() { release(); }
};

// A task list whose device tables do not fit the memory given to upload().
struct HHCapacity : std::runtime_error {
  using std::runtime_error::runtime_error;
};

// Fixed device scratch of the allocation-free entry point
// (hh_run_tasks_noalloc): a file-scope device array of this many bytes in
// the engine's code object, present on every GPU the library runs on until
// the process ends.
constexpr size_t NOALLOC_SCRATCH_BYTES = size_t(64) << 20;
// Address of that array on `device`, which must be the current device: taken
// with hipGetSymbolAddress on the first call for the device and kept.
void* noalloc_scratch_init(int device);
// The address found by noalloc_scratch_init for `device` (nullptr before it).
void* noalloc_scratch(int device);

// Arithmetic of an H.psi: FP64 MFMA on complex128 operands (the plain form),
// FP32 MFMA on complex128 operands, or FP32 MFMA on complex64 copies of the
// operands, intermediates and vectors (FP32 storage).
enum Precision { PREC_FP64 = 0, PREC_FP32 = 1, PREC_FP32_STORE = 2 };

// C tiles = sum of terms; one wave per 32x32 macro tile and term range.
// tag names the launch site so profiles separate them: 0 generic, 1 H.psi
// stage 1 (A.psi products), 2 H.psi stage 2 (all output tiles), 3 and 4 the
// same two inside a profiled region.  `mid`, when given, is recorded between
// the tile kernel and the split-sum reduction.
void launch_tiles(const DevTasks& t, const BufPtrs& b, hipStream_t s, int precision = PREC_FP64, int tag = 0,
                  hipEvent_t mid = nullptr);
void launch_tiles(const DevTasks& t, const BufTable& b, hipStream_t s, int precision = PREC_FP64, int tag = 0,
                  hipEvent_t mid = nullptr);
void launch_tiles(const DevTasks& t, const BufPtrs32& b, hipStream_t s, int tag = 0, hipEvent_t mid = nullptr);
// p = conj(p), n elements
void dev_conj(double2* p, int64_t n, hipStream_t s);
// complex128 <-> complex64 copies of n elements
void dev_to_f32(float2* d, const double2* src, int64_t n, hipStream_t s);
void dev_to_f64(double2* d, const float2* src, int64_t n, hipStream_t s);

// vector kernels (complex128 vectors of length n)
void dev_zero(double2* p, int64_t n, hipStream_t s);
void dev_scale(double2* p, double a, int64_t n, hipStream_t s);
void dev_copy(double2* dst, const double2* src, int64_t n, hipStream_t s);
// w = w - a*v1 - b*v2 (v2 may be null)
void dev_lanczos_update(double2* w, const double2* v1, double a, const double2* v2, double b, int64_t n, hipStream_t s);
// out[j] = Re <V_j, w>, j < nv, deterministic two-pass reduction; V_j = V + j*stride
void dev_rdots(const double2* V, int64_t stride, int nv, const double2* w, int64_t n, double* out_dev,
               double* partial_dev, hipStream_t s);
// w = beta*w + sum_j c_j V_j  (c on the device)
void dev_combine(double2* w, double beta, const double2* V, int64_t stride, int nv, const double* c_dev, int64_t n,
                 hipStream_t s);
int rdot_partials();   // partial-sum slots per vector

// Fused Lanczos kernels with deterministic reductions (block order fixed by n).
int vec_blocks(int64_t n);   // partial-sum slots per vector of length n
// out[j] = Re <V_j, w>, j < nv; w is read once
void dev_mdots(const double2* V, int64_t stride, int nv, const double2* w, int64_t n, double* out_dev,
               double* partial_dev, hipStream_t s);
// w = w - a*v1 - b*v2 (v2 may be null); out[0] = |w|^2; a is read from a_dev
// on the device when a_dev is set
void dev_lanczos_update_norm(double2* w, const double2* v1, double a, const double2* v2, double b, int64_t n,
                             double* out_dev, double* partial_dev, hipStream_t s, const double* a_dev = nullptr);
// w = sw*w + cs * sum_j c_j V_j (c on the device); out[0] = |w|^2
void dev_combine_norm(double2* w, double sw, const double2* V, int64_t stride, int nv, const double* c_dev, double cs,
                      int64_t n, double* out_dev, double* partial_dev, hipStream_t s);
// w = sum_j (cv_j V_j + cw_j W_j) over two strided vector sets (c on the
// device); out[0] = |w|^2
void dev_combine2_norm(double2* w, const double2* V, const double2* W, int64_t stride, int nv, const double* cv_dev,
                       const double* cw_dev, int64_t n, double* out_dev, double* partial_dev, hipStream_t s);
// d = a * src
void dev_scale_copy(double2* d, const double2* src, double a, int64_t n, hipStream_t s);

// Segmented vectors: nseg vectors stored one after the other, segment s
// starting at a multiple of vec_span() elements and holding send[s] - start
// elements; every block of the launch lies inside one segment (bseg), and
// segment s owns blocks sfirst[s] .. sfirst[s + 1] - 1.  Per-segment sums
// equal the unsegmented kernels' sums of the same vector bit for bit.
int vec_span();
struct SegDev {
  int nseg = 0, nb = 0;
  const int32_t* bseg = nullptr;
  const int64_t* send = nullptr;
  const int32_t* sfirst = nullptr;
  // with `active` set, the multi-dot and the three-term update skip the
  // blocks of segments whose flag is 0 (their sums come out as 0)
  const int32_t* active = nullptr;
};
// out[s * nv + j] = Re<V_j, w> over segment s
void dev_seg_mdots(const double2* V, int64_t stride, int nv, const double2* w, const SegDev& g, double* out_dev,
                   double* partial_dev, hipStream_t s);
// w -= a_s v1 + b_s v2 per segment (v2 may be null); out[s] = |w_s|^2
void dev_seg_lupd_norm(double2* w, const double2* v1, const double* a_dev, const double2* v2, const double* b_dev,
                       const SegDev& g, double* out_dev, double* partial_dev, hipStream_t s);
// w = sw_s w + cs_s sum_j c[s * ldc + j] V_j per segment; out[s] = |w_s|^2
void dev_seg_combine_norm(double2* w, const double* sw_dev, const double2* V, int64_t stride, int nv,
                          const double* c_dev, int ldc, const double* cs_dev, const SegDev& g, double* out_dev,
                          double* partial_dev, hipStream_t s);
// d1 = a_s src per segment, and d2 = the same when d2 is not null
void dev_seg_scale2(double2* d1, double2* d2, const double2* src, const double* a_dev, const SegDev& g,
                    hipStream_t s);
// d = (ar + i ai) src
void dev_cscale_copy(double2* d, const double2* src, double ar, double ai, int64_t n, hipStream_t s);

void check_hip(hipError_t e, const char* what);
// counter-based start vector generated on the device (same stream as the host)
void dev_start_vector(double2* p, int64_t n, uint64_t seed, hipStream_t s);

}  // namespace hh
