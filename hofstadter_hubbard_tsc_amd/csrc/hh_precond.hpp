// Learned per-sector Davidson preconditioner: device kernels and their host
// entry points.
//
// The superblock vector psi is a list of sector blocks p = (sys sector x,
// env sector y), rows x cols each.  For element i = (a, b) of block p the
// preconditioner is
//
//     t_i = alpha_p * r_i / (max(d_i - theta, 0) + sigma_p),
//
// with r the residual, theta the current Ritz value and d an estimate of
// the diagonal of H: the diagonal of the sys and env block Hamiltonians in
// the enlarged bases (d_i = hs[a] + he[b]).  The per-sector pair
// (sigma_p, alpha_p) comes from a small MLP evaluated on features of block p
// (diagonal statistics, residual share, quantum numbers):
//     8 -> 256 (tanh) -> 256 (tanh) -> 2,
//     sigma_p = PC_SIGMA0 * scale * exp(PC_BOUND * tanh(o_0)),
//     alpha_p = exp(PC_BOUND * tanh(o_1)),
// a bounded correction (a factor within [1/4, 4]) of the fixed diagonal form
// sigma = PC_SIGMA0 * scale, alpha = 1, which the network gives at o = 0: a
// sector unlike the training data cannot push the preconditioner further
// than that from the fixed form.  scale is the spread of d.
// The network is trained on (diagonal, residual, exact correction) triples:
// the loss of one triple is 1 - Re<t, delta>^2 / (|t|^2 |delta|^2), the part
// of the exact correction delta that one Davidson step along t misses (it
// does not depend on the overall scale of t).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

namespace hh {

constexpr int PC_NF = 8;      // features per sector
constexpr int PC_H = 256;     // hidden width (two hidden layers)
constexpr int PC_NO = 2;      // outputs per sector: sigma, alpha (pre-activation)
constexpr int PC_CHUNK = 32768;   // elements per reduction workgroup
// shift of the fixed diagonal form, in units of the spread of d
constexpr double PC_SIGMA0 = 0.1;
// largest log-factor by which the network moves sigma_p and alpha_p (ln 4)
constexpr double PC_BOUND = 1.3862943611198906;

// parameter offsets in the flat FP32 parameter vector
struct MlpLayout {
  static constexpr int W1 = 0, B1 = W1 + PC_H * PC_NF, W2 = B1 + PC_H, B2 = W2 + PC_H * PC_H, W3 = B2 + PC_H,
                       B3 = W3 + PC_NO * PC_H, SIZE = B3 + PC_NO;
};

// One superblock's sector table on the device.
struct PairTable {
  int np = 0;
  int64_t n = 0;
  // per sector: element offset, rows, cols, offset into hs / he, 2S and N of
  // the sys sector
  std::vector<int64_t> off;
  std::vector<int32_t> rows, cols, soff, eoff, s2, nq;
};

// Device buffers of a preconditioner evaluation for one superblock.
struct PcDev {
  int np = 0;
  int64_t n = 0;
  int64_t* off = nullptr;
  int32_t* rows = nullptr;
  int32_t* cols = nullptr;
  int32_t* s2 = nullptr;
  int32_t* nq = nullptr;
  int32_t* pid = nullptr;        // sector of every element
  double* d = nullptr;           // diagonal estimate per element
  double* stats = nullptr;       // per sector: sum d, sum d^2, min d, sum |r|^2
  float* feat = nullptr;         // np x PC_NF
  float* h1 = nullptr;           // np x PC_H
  float* h2 = nullptr;           // np x PC_H
  float* out = nullptr;          // np x PC_NO
  double* par = nullptr;         // per sector: sigma, alpha
  double* gout = nullptr;        // per sector: the four loss sums (k_loss_sums)
  float* dout = nullptr;         // np x PC_NO
  float* dh2 = nullptr;          // np x PC_H
  float* dh1 = nullptr;          // np x PC_H
  double* glob = nullptr;        // global scalars: scale of d, |r|^2
  // sectors cut into chunks of at most PC_CHUNK elements (one workgroup
  // each); chunks of sector p are cfirst[p] .. cfirst[p + 1] - 1
  int nchunk = 0;
  int32_t* ch_pair = nullptr;
  int64_t* ch_off = nullptr;     // first element (global index)
  int32_t* ch_cnt = nullptr;
  int32_t* cfirst = nullptr;     // np + 1
  double* cstats = nullptr;      // per chunk: 4 partial sums
  void release();
  ~PcDev() { release(); }
};

// Upload the sector table and build d from the enlarged-basis diagonals
// hs / he (host vectors); `n_mid` centres the N feature.
void pc_setup(PcDev& P, const PairTable& t, const std::vector<double>& hs, const std::vector<double>& he, int n_mid,
              hipStream_t s);
// Features of every sector from d, r and theta (after pc_setup).
void pc_features(PcDev& P, const double2* r, double theta, int n_mid, hipStream_t s);
// MLP forward: feat -> out (parameters at w, FP32), then sigma / alpha.
void pc_forward(PcDev& P, const float* w, hipStream_t s);
// Fixed per-sector parameters (no network): sigma = sigma0 * scale, alpha = 1.
void pc_fixed(PcDev& P, double sigma0, hipStream_t s);
// t = alpha_p r / (max(d - theta, 0) + sigma_p)
void pc_apply(const PcDev& P, const double2* r, double theta, double2* t, hipStream_t s);
// Training: loss and dL/d(sigma_p, alpha_p) for target delta (|delta|^2 =
// dn), then dL/d(out) -> parameter gradient g (FP32, size MlpLayout::SIZE).
// The loss of the sample is written to loss_dev[0].
void pc_backward(PcDev& P, const double2* r, const double2* delta, double theta, double dn, const float* w, float* g,
                 double* loss_dev, hipStream_t s);
// Adam step on n parameters (bias-corrected with step number t >= 1).
void pc_adam(float* w, const float* g, float* m, float* v, int n, float lr, int t, hipStream_t s);
// host -> device block-H diagonal gather: out[i] = Re(arena[idx[i]])
void dev_gather_re(double* out, const double2* arena, const int64_t* idx, int n, hipStream_t s);

}  // namespace hh
