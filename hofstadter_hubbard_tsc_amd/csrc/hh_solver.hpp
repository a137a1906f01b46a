// The chosen ground-state form: Davidson with FP32-storage H.psi, FP64
// restarts and the learned per-sector preconditioner, plus the training of
// that preconditioner on triples captured from plain solves.
#pragma once
#include <functional>
#include <memory>
#include <vector>

#include "hh_precond.hpp"
#include "hh_runtime.hpp"

namespace hh {

// Preconditioner of a Davidson solve.
enum PrecondMode { PC_LEARNED = 0, PC_DIAGONAL = 1, PC_NONE = 2 };

// The network's parameters and Adam state (FP32, MlpLayout::SIZE each).
struct PcNet {
  DevMem w, g, m, v, loss;
  int adam_t = 0;
  bool trained = false;
  void init(hipStream_t s, uint64_t seed);
};

// One training triple: a superblock's sector table and diagonal estimate,
// the FP64 residual r of a start vector u with Rayleigh quotient theta, and
// the exact correction delta = psi* / Re<u, psi*> - u toward the ground
// state psi* the plain solver found.
struct PcSample {
  PcDev P;
  DevMem r, delta;
  double theta = 0.0, dn = 0.0;
  int n_mid = 0;
};

// The diagonal estimate and sector table of a bound superblock on the device.
void precond_setup(hh_ctx* ctx, const hh_superblock* sb, PcDev& P);

// Davidson ground state of sb from the start vector psi (overwritten with
// the result).  max_hpsi bounds the H.psi applications of both precisions.
// A round of FP32 expansions ends by the reference's stopping rule
// (code/lanczos.f90:77-80: the lowest Ritz value moved by at most lanzero)
// or at the subspace limit, and restarts from its Ritz vector with an FP64
// H.psi; the solve has converged when that exact Rayleigh quotient moved by
// at most lanzero since the previous restart, or when it confirms to lanzero
// the Ritz value of a round that ended by the stopping rule.  net is the trained
// network (PC_LEARNED) or null.  *fell_back is set when the FP64 residual
// stopped falling and the mixed-precision Lanczos finished the solve.
// Called at every FP64 restart with the unit vector u, its exact residual
// H u - theta u and theta (device vectors valid during the call).
typedef std::function<void(const double2*, const double2*, double)> AnchorHook;
// Largest Davidson subspace before a restart (default; hh_driver_set_subspace
// sets it per driver, up to DAV_KMAX_LIMIT).
constexpr int DAV_KMAX = 32;
constexpr int DAV_KMAX_LIMIT = 128;
// fall_factor: the FP64 residual must fall below fall_factor times its value
// at the previous restart, or the mixed Lanczos takes over.
constexpr double DAV_FALL_FACTOR = 0.9;
// Superblocks below this dimension are solved by the plain Lanczos under
// solvers 1 and 2: there the H.psi is a small part of the step and the
// Davidson's extra host round trips per expansion cost more than its fewer
// H.psi save.
constexpr int64_t DAV_MIN_DIM = 262144;
SolveOut davidson(hh_ctx* ctx, hh_superblock* sb, double2* psi, double lanzero, int max_hpsi, int mode,
                  const float* net, bool* fell_back, const AnchorHook& on_anchor = AnchorHook(),
                  double fall_factor = DAV_FALL_FACTOR, int kmax_limit = DAV_KMAX);

// Record one training triple from a finished plain solve: u0 is the unit
// start vector, r0 = H u0 - theta0 u0 its FP64 residual, psi the ground state.
std::unique_ptr<PcSample> make_sample(hh_ctx* ctx, const hh_superblock* sb, const double2* u0, const double2* r0,
                                      double theta0, const double2* psi);

// Train the network on the samples (epochs passes in a fixed order, Adam);
// mean loss per epoch in *losses.
void pc_train(hh_ctx* ctx, PcNet& net, const std::vector<std::unique_ptr<PcSample>>& samples, int epochs, double lr,
              std::vector<double>* losses);

}  // namespace hh
