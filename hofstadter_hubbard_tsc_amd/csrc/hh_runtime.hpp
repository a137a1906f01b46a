// Runtime objects shared by the library's translation units: the context,
// device-resident blocks, bound superblocks and truncations behind the C ABI
// (include/hhdmrg.h), device memory ownership, errors, and the H.psi and
// ground-state solvers that more than one file calls.
#pragma once
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>

#include <chrono>
#include <cstdlib>
#include <functional>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/hhdmrg.h"
#include "hh_core.hpp"
#include "hh_engine.hpp"
#include "hh_heev.hpp"

namespace hh {

struct HHError : std::runtime_error {
  int code;
  HHError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};
inline void need(bool ok, const std::string& msg) {
  if (!ok) throw HHError(1, msg);
}

// device allocation owned by one object
struct DevMem {
  void* p = nullptr;
  size_t bytes = 0;
  DevMem() = default;
  DevMem(const DevMem&) = delete;
  DevMem& operator=(const DevMem&) = delete;
  ~DevMem() { reset(); }
  void reset() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  void alloc(size_t b) {
    if (b <= bytes && p) return;
    reset();
    if (b == 0) b = 16;
    check_hip(hipMalloc(&p, b), "hipMalloc");
    bytes = b;
  }
  double2* c() const { return (double2*)p; }
  float2* f() const { return (float2*)p; }
};

// HHDMRG_VERBOSE=1: progress lines on stderr (sweep steps, truncation phases,
// measurement chains)
inline bool verbose() {
  static const bool v = std::getenv("HHDMRG_VERBOSE") && std::getenv("HHDMRG_VERBOSE")[0] == '1';
  return v;
}

inline double wall() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}


// Outcome of one ground-state solve.  converged: the reference's stopping
// rule held (|change of the lowest Ritz value| <= lanzero, or beta <=
// lanzero, code/lanczos.f90:77-80, :102) before the vector budget ran out;
// ritz_change is the last change of the lowest Ritz value; residual is
// |H psi - E psi| of the returned unit vector (for Lanczos the estimate
// beta_k |z_k| of the Krylov relation, exact up to the loss of orthogonality).
struct SolveOut {
  double e0 = 0.0;
  int iters = 0;
  double mv_seconds = 0.0;
  bool converged = false;
  double ritz_change = 0.0;
  double residual = 0.0;
  int hpsi_fp64 = 0, hpsi_fp32 = 0;
  bool fallback = false;   // the chosen form handed the solve to the mixed Lanczos
};

// Options of one Lanczos run: the arithmetic of its first H.psi and of the
// others (-1: the context's precision), and a test called with alpha_1 =
// <v_1, H v_1> and beta_1 = |H v_1 - alpha_1 v_1| right after the first H.psi
// that may end the run there (psi then comes back normalised, E = alpha_1).
struct LanczosOpts {
  int first_prec = -1, rest_prec = -1;
  std::function<bool(double, double)> after_first;
  // called once after the first H.psi with v_1 (the unit start vector),
  // H v_1 - alpha_1 v_1 and alpha_1 (device vectors valid during the call)
  std::function<void(const double2*, const double2*, double)> on_first;
  // levels > 1 (Gram_Schmidt_Lanczos, code/space_info.f90:1222-1279): the
  // start vector and every new Lanczos vector are projected off `ndeflate`
  // unit vectors at stride n, orthonormal in the real inner product (each
  // lower level u and i u, so that the projection is complex)
  const double2* deflate = nullptr;
  int ndeflate = 0;
};

}  // namespace hh

// ============================================================================
// objects (the opaque handles of include/hhdmrg.h)
// ============================================================================
using hh::BlockDesc;
using hh::DevTasks;
using hh::EnlBasis;
using hh::Geometry;
using hh::KeptSector;
using hh::MatvecPlan;
using hh::ModelParams;
using hh::PREC_FP64;

struct hh_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  rocblas_handle rb = nullptr;
  // side streams (default priority) for independent per-sector eigensolver
  // calls, each with its own rocBLAS handle; joined back with events
  static constexpr int NSIDE = 7;
  hipStream_t side[NSIDE] = {};
  rocblas_handle rbs[NSIDE] = {};
  hipEvent_t fork = nullptr, join[NSIDE] = {};
  ModelParams model;
  bool model_set = false;
  int precision = PREC_FP64;
  std::string err;
  // Lanczos workspace
  hh::DevMem V, dots, partial, scratch_w;
  // complex64 copies of the H.psi input and output (FP32-storage form)
  hh::DevMem in32, out32;
  // profiling
  bool profile = false;
  hipEvent_t ev[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
  hh_stats stats{};
  std::vector<hh::DevMem*> user_allocs;
  // outcome of the last ground-state solve (hh_lanczos_status)
  int last_converged = 0;
  double last_ritz_change = 0.0, last_residual = 0.0;
  // autotuning of the H.psi launches (hh_tune.cpp): on/off, the chosen
  // configurations by task-list shape key, counters
  int autotune = 0;
  // a launch cut forced on every task list uploaded from now on (tests,
  // experiments); overrides the autotuner when set
  bool forced_cfg = false;
  hh::LaunchConfig forced;
  std::map<uint64_t, hh::LaunchConfig> tune_cache;
  int tune_runs = 0, tune_hits = 0;
  double tune_seconds = 0.0;
  // eigensolver of the truncation: 0 the hand-written one (hh_heev.hip), 1
  // rocSOLVER zheevd per sector; truncations that the hand-written solver
  // hands back (a sector above HEEV_MAXN, vectors it cannot orthonormalise)
  // are counted and redone with zheevd
  int eig_solver = 0;
  bool eig_warm = false;   // the library eigensolver's set-up has run
  hh::HeevWork heev;
  int64_t eig_handbacks = 0;
};

struct hh_block {
  BlockDesc d;
  hh::DevMem arena;
  // complex64 copy of the arena for the FP32-storage H.psi, made on first use
  // (a block's content never changes after it is built or uploaded; an
  // upload drops the copy)
  mutable hh::DevMem arena32;
  mutable bool have32 = false;
};

struct hh_superblock {
  const hh_block* sys = nullptr;
  const hh_block* env = nullptr;
  int N = 0;
  Geometry g;
  MatvecPlan plan;
  DevTasks merge, st1, st2;
  hh::DevMem merged, X;
  // complex64 merged operators and stage-1 intermediate (FP32-storage form)
  hh::DevMem merged32, X32;
  bool have32 = false;
};

// Several superblocks' H.psi task lists merged into one list per stage; the
// buffer slots of member i are 8*i + (B_IN .. B_AUX2) in a device table.
struct hh_batch {
  std::vector<hh_superblock*> sbs;
  DevTasks st1, st2;
  hh::DevMem table;
  std::vector<double2*> host_table;
  double flops_alg = 0.0, stage1_flops = 0.0, stage2_flops = 0.0, flops_exec = 0.0;
  double stage1_exec = 0.0, stage2_exec = 0.0;
  int64_t total_dim = 0;

  // batched Lanczos: every member's vectors as one segmented vector (member
  // i at off[i], a multiple of hh::vec_span(), npad elements in all), the
  // Lanczos vectors V (stride npad), the H.psi input Q and output W at fixed
  // addresses, the segment tables, and the grouped H.psi Q -> W captured as
  // a HIP graph (its buffer table lz_table holds Q, W and the members'
  // blocks and intermediates)
  std::vector<int64_t> off;
  int64_t npad = 0;
  hh::DevMem V, Q, W, lz_table, seg_bseg, seg_send, seg_sfirst, scal, partial;
  hh::DevMem active;   // int32 per member: 0 once the member's solve has stopped
  hh::SegDev seg;
  bool use_graph = true;
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  ~hh_batch() {
    if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
    if (graph) (void)hipGraphDestroy(graph);
  }
};

struct hh_trunc {
  int side = 0;
  EnlBasis eb;
  std::vector<KeptSector> kept;
  std::vector<int> kept_per_rho;   // per rho block (super-basis order)
  hh::DevMem evec;
  int num_state = 0;
  double err = 0.0, entropy = 0.0;
};


namespace hh {

// C-ABI guard: exceptions -> status codes, message kept in the context
template <typename F>
int guard(hh_ctx* ctx, F&& f) {
  try {
    f();
    return 0;
  } catch (const HHError& e) {
    if (ctx) ctx->err = e.what();
    return e.code;
  } catch (const std::exception& e) {
    if (ctx) ctx->err = e.what();
    return -1;
  }
}

// symmetric tridiagonal eigenproblem (implicit-shift QL): ascending
// eigenvalues in d (off-diagonal e), and the eigenvector of the lowest in *z0
void tridiag_ql(std::vector<double>& d, std::vector<double> e, std::vector<double>* z0);

// buffer-slot pointers of one superblock's H.psi
BufPtrs bufs_of(double2* in, double2* out, const hh_block* sys, const hh_block* env, double2* X, double2* mrg,
                double2* aux, double2* aux2);

// H.psi of a bound superblock in the context's precision
void matvec(hh_ctx* ctx, hh_superblock* sb, const double2* in, double2* out);
// H.psi in the FP32-storage form (complex64 copies of every operand)
void matvec32(hh_ctx* ctx, hh_superblock* sb, const double2* in, double2* out);
void ensure32(hh_ctx* ctx, const hh_block* b);
// Lanczos workspace of nvec + 1 vectors of length n and its reduction scratch
void ensure_ws(hh_ctx* ctx, int64_t n, int nvec);
// lanczos_memo_new (code/lanczos.f90:4-176), levels = 1
SolveOut lanczos(hh_ctx* ctx, hh_superblock* sb, double2* psi, double lanzero, int max_iter,
                 const LanczosOpts& opt = LanczosOpts());
// mixed-precision Lanczos rounds (FP32-storage H.psi, FP64 restarts)
SolveOut mixed_lanczos(hh_ctx* ctx, hh_superblock* sb, double2* psi, double lanzero, int max_iter, int max_rounds);

// Autotuning of the grouped H.psi launches (hh_tune.cpp): the shape key of a
// task list (its flop-weighted (M, N, K) histogram in factors of two), and
// the re-cut of a bound superblock's or a batch's stage lists by the
// fastest of a set of launch configurations, cached by shape key.
uint64_t shape_key(const TaskList& tl);
void autotune_superblock(hh_ctx* ctx, hh_superblock* sb);
void autotune_batch(hh_ctx* ctx, hh_batch* b, const TaskList& s1, const TaskList& s2);

// The C model struct as the task compilers take it, and the check of its
// gauge fields (hh_runtime.cpp)
ModelParams to_model(const hh_model_params* p);
void check_gauge(const hh_model_params* p);
// message of a failed create call (the object it would have made does not exist)
std::string& create_error();

}  // namespace hh
