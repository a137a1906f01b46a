// Runtime objects behind the C ABI (include/hhdmrg.h): context, device-resident
// blocks, bound superblocks, truncations, the Lanczos solver, the sweep
// driver's steps (hh_driver.hpp) and its checkpoint.  All arithmetic runs in
// the HIP engine (hh_engine.hip) and the hand-written density-matrix
// eigensolver (hh_heev.hip; rocSOLVER zheevd on request); the host only
// compiles task lists, applies the integer kept-count rule and steers the
// sweep.  Measurements live in hh_measure.cpp, the reference's file formats in
// hh_files.cpp, the Davidson form in hh_davidson.cpp, batches in hh_batch.cpp.
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <array>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <climits>
#include <cstring>
#include <functional>
#include <initializer_list>
#include <limits>
#include <map>
#include <memory>
#include <mutex>
#include <numeric>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "../../include/hhdmrg.h"
#include "hh_core.hpp"
#include "hh_driver.hpp"
#include "hh_engine.hpp"
#include "hh_runtime.hpp"
#include "hh_solver.hpp"

using namespace hh;

namespace hh {

// ----------------------------------------------------------------------------
// symmetric tridiagonal eigenproblem (DSTEV stand-in): implicit-shift QL
// (the textbook tql2 method); ascending eigenvalues, optional first vector.
// ----------------------------------------------------------------------------
void tridiag_ql(std::vector<double>& d, std::vector<double> e, std::vector<double>* z0) {
  const int n = (int)d.size();
  std::vector<double> Z;
  if (z0) {
    Z.assign((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i) Z[(size_t)i * n + i] = 1.0;
  }
  e.resize(n);
  e[n - 1] = 0.0;
  for (int l = 0; l < n; ++l) {
    int iter = 0;
    while (true) {
      int m;
      for (m = l; m < n - 1; ++m) {
        const double dd = std::fabs(d[m]) + std::fabs(d[m + 1]);
        if (std::fabs(e[m]) <= 1e-300 || std::fabs(e[m]) <= std::numeric_limits<double>::epsilon() * dd) break;
      }
      if (m == l) break;
      if (++iter > 200) throw HHError(-3, "tridiagonal QL did not converge");
      double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
      double r = std::hypot(g, 1.0);
      g = d[m] - d[l] + e[l] / (g + (g >= 0 ? std::fabs(r) : -std::fabs(r)));
      double s = 1.0, c = 1.0, p = 0.0;
      int i;
      for (i = m - 1; i >= l; --i) {
        double f = s * e[i], b = c * e[i];
        r = std::hypot(f, g);
        e[i + 1] = r;
        if (r == 0.0) {
          d[i + 1] -= p;
          e[m] = 0.0;
          break;
        }
        s = f / r;
        c = g / r;
        g = d[i + 1] - p;
        r = (d[i] - g) * s + 2.0 * c * b;
        p = s * r;
        d[i + 1] = g + p;
        g = c * r - b;
        if (z0)
          for (int k = 0; k < n; ++k) {
            double& zi1 = Z[(size_t)(i + 1) * n + k];
            double& zi = Z[(size_t)i * n + k];
            const double t = zi1;
            zi1 = s * zi + c * t;
            zi = c * zi - s * t;
          }
      }
      if (r == 0.0 && i >= l) continue;
      d[l] -= p;
      e[l] = g;
      e[m] = 0.0;
    }
  }
  // sort ascending (with vectors)
  std::vector<int> idx(n);
  std::iota(idx.begin(), idx.end(), 0);
  std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return d[a] < d[b]; });
  std::vector<double> ds(n);
  for (int i = 0; i < n; ++i) ds[i] = d[idx[i]];
  if (z0) {
    z0->assign(n, 0.0);
    for (int k = 0; k < n; ++k) (*z0)[k] = Z[(size_t)idx[0] * n + k];
  }
  d = ds;
}

BufPtrs bufs_of(double2* in, double2* out, const hh_block* sys, const hh_block* env, double2* X, double2* mrg,
                double2* aux, double2* aux2) {
  BufPtrs b;
  b.p[B_IN] = in;
  b.p[B_OUT] = out;
  b.p[B_SYS] = sys ? sys->arena.c() : nullptr;
  b.p[B_ENV] = env ? env->arena.c() : nullptr;
  b.p[B_X] = X;
  b.p[B_MRG] = mrg;
  b.p[B_AUX] = aux;
  b.p[B_AUX2] = aux2;
  return b;
}

}  // namespace hh

namespace hh {

hh_block* new_block(hh_ctx* ctx, int side, int length, const std::vector<Sector>& sec, const std::vector<int>& sites) {
  auto* b = new hh_block();
  b->d.side = side;
  b->d.length = length;
  b->d.sec = sec;
  b->d.sites = sites;
  layout_block(b->d, n_op_kinds(ctx->model.kind));
  b->arena.alloc(sizeof(double2) * (size_t)std::max<int64_t>(b->d.size, 1));
  return b;
}

}  // namespace hh

namespace hh {

void ensure32(hh_ctx* ctx, const hh_block* b) {
  if (b->have32) return;
  b->arena32.alloc(sizeof(float2) * (size_t)std::max<int64_t>(b->d.size, 1));
  dev_to_f32(b->arena32.f(), b->arena.c(), b->d.size, ctx->stream);
  b->have32 = true;
}

// H.psi in the FP32-storage form: complex64 copies of the blocks, merged
// operators, stage-1 intermediate and vectors; FP32 MFMA, FP32 accumulation.
// in and out stay complex128 (converted on the way in and out).
void matvec32(hh_ctx* ctx, hh_superblock* sb, const double2* in, double2* out) {
  const int64_t n = sb->plan.sb.dim;
  ensure32(ctx, sb->sys);
  ensure32(ctx, sb->env);
  if (!sb->have32) {
    sb->merged32.alloc(sizeof(float2) * (size_t)std::max<int64_t>(sb->plan.merged_size, 1));
    sb->X32.alloc(sizeof(float2) * (size_t)std::max<int64_t>(sb->plan.x_size, 1));
    dev_to_f32(sb->merged32.f(), sb->merged.c(), sb->plan.merged_size, ctx->stream);
    sb->have32 = true;
  }
  ctx->in32.alloc(sizeof(float2) * (size_t)std::max<int64_t>(n, 1));
  ctx->out32.alloc(sizeof(float2) * (size_t)std::max<int64_t>(n, 1));
  BufPtrs32 b{};
  b.p[B_IN] = ctx->in32.f();
  b.p[B_OUT] = ctx->out32.f();
  b.p[B_SYS] = sb->sys->arena32.f();
  b.p[B_ENV] = sb->env->arena32.f();
  b.p[B_X] = sb->X32.f();
  b.p[B_MRG] = sb->merged32.f();
  const double t0 = ctx->profile ? wall() : 0.0;
  if (ctx->profile) check_hip(hipEventRecord(ctx->ev[0], ctx->stream), "event");
  dev_to_f32(ctx->in32.f(), in, n, ctx->stream);
  launch_tiles(sb->st1, b, ctx->stream, ctx->profile ? 3 : 1, ctx->profile ? ctx->ev[1] : nullptr);
  if (ctx->profile) check_hip(hipEventRecord(ctx->ev[2], ctx->stream), "event");
  launch_tiles(sb->st2, b, ctx->stream, ctx->profile ? 4 : 2, ctx->profile ? ctx->ev[3] : nullptr);
  dev_to_f64(out, ctx->out32.f(), n, ctx->stream);
  if (ctx->profile) {
    check_hip(hipEventRecord(ctx->ev[4], ctx->stream), "event");
    check_hip(hipEventSynchronize(ctx->ev[4]), "event sync");
    float ms[4] = {0.f, 0.f, 0.f, 0.f};
    for (int i = 0; i < 4; ++i) check_hip(hipEventElapsedTime(&ms[i], ctx->ev[i], ctx->ev[i + 1]), "elapsed");
    ctx->stats.stage1_ms += ms[0];
    ctx->stats.reduce_ms += ms[1] + ms[3];
    ctx->stats.stage2_ms += ms[2];
    ctx->stats.stage2_ms_fp32 += ms[2];
    ctx->stats.matvec_ms += (wall() - t0) * 1e3;
  }
  ctx->stats.matvecs += 1;
  ctx->stats.matvecs_fp32 += 1;
  ctx->stats.stage2_flops_alg_fp32 += sb->plan.stage2.flops_alg;
  ctx->stats.flops_alg += sb->plan.stage1.flops_alg + sb->plan.stage2.flops_alg;
  ctx->stats.stage1_flops_alg += sb->plan.stage1.flops_alg;
  ctx->stats.stage2_flops_alg += sb->plan.stage2.flops_alg;
  ctx->stats.flops_exec += sb->plan.stage1.flops_exec + sb->plan.stage2.flops_exec;
  ctx->stats.stage1_flops_exec += sb->plan.stage1.flops_exec;
  ctx->stats.stage2_flops_exec += sb->plan.stage2.flops_exec;
  ctx->stats.bytes_alg += 0.5 * sb->plan.bytes_alg;
}

void matvec(hh_ctx* ctx, hh_superblock* sb, const double2* in, double2* out) {
  if (ctx->precision == PREC_FP32_STORE) {
    matvec32(ctx, sb, in, out);
    return;
  }
  BufPtrs b = bufs_of((double2*)in, out, sb->sys, sb->env, sb->X.c(), sb->merged.c(), nullptr, nullptr);
  if (!ctx->profile) {
    launch_tiles(sb->st1, b, ctx->stream, ctx->precision, 1);
    launch_tiles(sb->st2, b, ctx->stream, ctx->precision, 2);
  } else {
    const double t0 = wall();
    check_hip(hipEventRecord(ctx->ev[0], ctx->stream), "event");
    launch_tiles(sb->st1, b, ctx->stream, ctx->precision, 3, ctx->ev[1]);
    check_hip(hipEventRecord(ctx->ev[2], ctx->stream), "event");
    launch_tiles(sb->st2, b, ctx->stream, ctx->precision, 4, ctx->ev[3]);
    check_hip(hipEventRecord(ctx->ev[4], ctx->stream), "event");
    check_hip(hipEventSynchronize(ctx->ev[4]), "event sync");
    float ms[4] = {0.f, 0.f, 0.f, 0.f};
    for (int i = 0; i < 4; ++i) check_hip(hipEventElapsedTime(&ms[i], ctx->ev[i], ctx->ev[i + 1]), "elapsed");
    ctx->stats.stage1_ms += ms[0];
    ctx->stats.reduce_ms += ms[1] + ms[3];
    ctx->stats.stage2_ms += ms[2];
    ctx->stats.matvec_ms += (wall() - t0) * 1e3;
  }
  ctx->stats.matvecs += 1;
  ctx->stats.flops_alg += sb->plan.stage1.flops_alg + sb->plan.stage2.flops_alg;
  ctx->stats.stage1_flops_alg += sb->plan.stage1.flops_alg;
  ctx->stats.stage2_flops_alg += sb->plan.stage2.flops_alg;
  ctx->stats.flops_exec += sb->plan.stage1.flops_exec + sb->plan.stage2.flops_exec;
  ctx->stats.stage1_flops_exec += sb->plan.stage1.flops_exec;
  ctx->stats.stage2_flops_exec += sb->plan.stage2.flops_exec;
  ctx->stats.bytes_alg += sb->plan.bytes_alg;
}

double rdot(hh_ctx* ctx, const double2* a, const double2* b, int64_t n) {
  double r = 0.0;
  dev_rdots(a, 0, 1, b, n, (double*)ctx->dots.p, (double*)ctx->partial.p, ctx->stream);
  check_hip(hipMemcpyAsync(&r, ctx->dots.p, sizeof(double), hipMemcpyDeviceToHost, ctx->stream), "d2h");
  check_hip(hipStreamSynchronize(ctx->stream), "sync");
  return r;
}

void ensure_ws(hh_ctx* ctx, int64_t n, int nvec) {
  ctx->V.alloc(sizeof(double2) * (size_t)n * (size_t)(nvec + 1));
  ctx->dots.alloc(sizeof(double) * (nvec + 2));
  ctx->partial.alloc(sizeof(double) * (size_t)std::max(rdot_partials(), vec_blocks(n)) * (nvec + 2));
}

// Lanczos, lanczos_memo_new (code/lanczos.f90:4-176), levels = 1.
SolveOut lanczos(hh_ctx* ctx, hh_superblock* sb, double2* psi, double lanzero, int max_iter,
                 const LanczosOpts& opt) {
  const int64_t n = sb->plan.sb.dim;
  need(n > 0, "empty superblock");
  need(max_iter >= 1, "max_iter must be >= 1");
  ensure_ws(ctx, n, max_iter + 1);
  double2* V = ctx->V.c();
  double2* w = V + (size_t)max_iter * n;   // last slot is the work vector
  auto vec = [&](int j) { return V + (size_t)j * n; };
  double mv = 0.0;
  int n64 = 0, n32 = 0, calls = 0;
  auto hpsi = [&](const double2* x, double2* y) {
    const double t = wall();
    const int saved = ctx->precision;
    const int prec = calls == 0 ? opt.first_prec : opt.rest_prec;
    if (prec >= 0) ctx->precision = prec;
    (ctx->precision == PREC_FP64 ? n64 : n32) += 1;
    ++calls;
    try {
      matvec(ctx, sb, x, y);
    } catch (...) {
      ctx->precision = saved;
      throw;
    }
    ctx->precision = saved;
    check_hip(hipStreamSynchronize(ctx->stream), "matvec");
    mv += wall() - t;
  };
  double* dots = (double*)ctx->dots.p;
  double* part = (double*)ctx->partial.p;
  auto fetch = [&](int count) {   // first `count` device scalars -> host
    std::vector<double> h(count);
    check_hip(hipMemcpyAsync(h.data(), dots, sizeof(double) * count, hipMemcpyDeviceToHost, ctx->stream), "d2h");
    check_hip(hipStreamSynchronize(ctx->stream), "sync");
    return h;
  };
  // v1 = psi / |psi|
  dev_mdots(psi, 0, 1, psi, n, dots, part, ctx->stream);
  const double nrm0 = std::sqrt(fetch(1)[0]);
  need(std::isfinite(nrm0) && nrm0 > 0.0, "zero or non-finite start vector");
  dev_scale_copy(vec(0), psi, 1.0 / nrm0, n, ctx->stream);
  const int nd = opt.ndeflate;
  need(nd == 0 || (opt.deflate && nd + 1 < max_iter + 2), "bad deflation set");
  // x -= sum_j Re<d_j, x> d_j over the deflation set, |x|^2 -> *sq
  auto deflate = [&](double2* x, double* sq) {
    dev_mdots(opt.deflate, n, nd, x, n, dots + 1, part, ctx->stream);
    dev_combine_norm(x, 1.0, opt.deflate, n, nd, dots + 1, -1.0, n, sq, part, ctx->stream);
  };
  if (nd > 0) {
    deflate(vec(0), dots);
    const double nv = std::sqrt(fetch(1)[0]);
    need(std::isfinite(nv) && nv > 1e-12, "start vector lies in the span of the lower levels");
    dev_scale(vec(0), 1.0 / nv, n, ctx->stream);
  }
  hpsi(vec(0), w);
  // a non-finite H.psi (NaN or Inf in a block or in psi) stops the solve with an error
  auto finite = [](double x) {
    if (!std::isfinite(x)) throw HHError(-4, "non-finite value in Lanczos: H.psi produced NaN or Inf");
    return x;
  };
  // alpha_j = Re<v_j, H v_j> stays on the device and the three-term update
  //   w -= alpha_j v_j + beta_{j-1} v_{j-1}   (|w|^2 in the same pass)
  // reads it there, so alpha_j and |w|^2 come back in one read.  When the
  // stopping test then ends the solve, that update was spare work on the
  // scratch vector w only.
  double* wsq = dots + max_iter + 2;   // past the Gram-Schmidt coefficients
  std::vector<double> alpha, beta;
  auto alpha_and_update = [&](int j) {
    dev_mdots(vec(j - 1), 0, 1, w, n, dots, part, ctx->stream);
    dev_lanczos_update_norm(w, vec(j - 1), 0.0, j >= 2 ? vec(j - 2) : nullptr, j >= 2 ? beta[j - 2] : 0.0, n, wsq,
                            part, ctx->stream, dots);
    // levels > 1: Lanczos on the projected operator, beta the norm after
    // the projection (alpha_j in dots[0] stays)
    if (nd > 0) deflate(w, wsq);
    double h[2];
    check_hip(hipMemcpyAsync(&h[0], dots, sizeof(double), hipMemcpyDeviceToHost, ctx->stream), "d2h");
    check_hip(hipMemcpyAsync(&h[1], wsq, sizeof(double), hipMemcpyDeviceToHost, ctx->stream), "d2h");
    check_hip(hipStreamSynchronize(ctx->stream), "sync");
    return std::make_pair(h[0], h[1]);
  };
  SolveOut out;
  auto r = alpha_and_update(1);
  alpha.push_back(finite(r.first));
  double ritz = alpha[0];
  int last = 1;
  double b_last = std::sqrt(std::max(r.second, 0.0));   // norm of the Krylov remainder
  if (opt.on_first) opt.on_first(vec(0), w, alpha[0]);
  if (opt.after_first && opt.after_first(alpha[0], b_last)) {
    dev_scale_copy(psi, vec(0), 1.0, n, ctx->stream);
    check_hip(hipStreamSynchronize(ctx->stream), "lanczos");
    out.e0 = alpha[0];
    out.iters = 1;
    out.mv_seconds = mv;
    out.converged = true;
    out.residual = b_last;
    out.hpsi_fp64 = n64;
    out.hpsi_fp32 = n32;
    ctx->last_converged = 1;
    ctx->last_ritz_change = 0.0;
    ctx->last_residual = b_last;
    return out;
  }
  while (last < max_iter) {
    const double b = std::sqrt(finite(r.second));
    beta.push_back(b);
    b_last = b;
    if (std::fabs(b) <= lanzero) {
      out.converged = true;
      break;
    }
    // normalise, then classical Gram-Schmidt against every stored vector with
    // real coefficients (Gram_Schmidt_process, code/space_info.f90:1176-1214;
    // the vectors are unit norm):  w <- (w - sum_j Re<v_j,w> v_j) / b
    dev_mdots(V, n, last, w, n, dots + 1, part, ctx->stream);
    dev_combine_norm(w, 1.0 / b, V, n, last, dots + 1, -1.0 / b, n, dots, part, ctx->stream);
    const double nw = std::sqrt(fetch(1)[0]);
    dev_scale_copy(vec(last), w, 1.0 / nw, n, ctx->stream);
    ++last;
    hpsi(vec(last - 1), w);
    r = alpha_and_update(last);
    alpha.push_back(finite(r.first));
    b_last = std::sqrt(std::max(r.second, 0.0));
    std::vector<double> d(alpha), e(beta);
    e.resize(d.size());
    tridiag_ql(d, e, nullptr);
    out.ritz_change = std::fabs(d[0] - ritz);
    if (out.ritz_change <= lanzero) out.converged = true;
    if (out.converged || last == max_iter) {
      ritz = d[0];
      break;
    }
    ritz = d[0];
  }
  // Ritz vector from the stored Lanczos vectors
  std::vector<double> d(alpha), e(beta), z;
  e.resize(d.size());
  tridiag_ql(d, e, &z);
  check_hip(hipMemcpyAsync(dots + 1, z.data(), sizeof(double) * last, hipMemcpyHostToDevice, ctx->stream), "h2d");
  dev_combine_norm(psi, 0.0, V, n, last, dots + 1, 1.0, n, dots, part, ctx->stream);
  const double np = std::sqrt(fetch(1)[0]);
  dev_scale(psi, 1.0 / np, n, ctx->stream);
  check_hip(hipStreamSynchronize(ctx->stream), "lanczos");
  out.e0 = d[0];
  out.iters = last;
  out.mv_seconds = mv;
  out.residual = b_last * std::fabs(z[last - 1]);
  out.hpsi_fp64 = n64;
  out.hpsi_fp32 = n32;
  ctx->last_converged = out.converged ? 1 : 0;
  ctx->last_ritz_change = out.ritz_change;
  ctx->last_residual = out.residual;
  return out;
}

// Mixed-precision ground state: Lanczos rounds whose first H.psi is FP64 and
// whose other H.psi run in the FP32-storage form.  The FP64 product of the
// round's start vector u gives its exact Rayleigh quotient alpha_1 and the
// exact residual r = H u - alpha_1 u; the round's Krylov space is spanned by
// u, r and the FP32 products, so its Ritz vector corrects u along the exact
// residual (iterative refinement: each round shrinks the error of u by about
// the relative FP32 error of H over the spectral gap).  Rounds restart from
// the Ritz vector until two successive FP64 Rayleigh quotients agree to
// lanzero (the reference's stopping rule applied to exact energies); the
// returned energy is that FP64 Rayleigh quotient of the returned vector and
// the residual its exact residual.
SolveOut mixed_lanczos(hh_ctx* ctx, hh_superblock* sb, double2* psi, double lanzero, int max_iter, int max_rounds) {
  SolveOut tot;
  double prev = std::numeric_limits<double>::quiet_NaN();
  bool stop = false;
  double a1 = 0.0, b1 = 0.0;
  LanczosOpts o;
  o.first_prec = PREC_FP64;
  o.rest_prec = PREC_FP32_STORE;
  o.after_first = [&](double a, double b) {
    a1 = a;
    b1 = b;
    stop = std::isfinite(prev) && (std::fabs(a - prev) <= lanzero || b <= lanzero);
    return stop;
  };
  for (int round = 0; round < max_rounds; ++round) {
    const SolveOut so = lanczos(ctx, sb, psi, lanzero, max_iter, o);
    tot.iters += so.iters;
    tot.mv_seconds += so.mv_seconds;
    tot.hpsi_fp64 += so.hpsi_fp64;
    tot.hpsi_fp32 += so.hpsi_fp32;
    tot.ritz_change = std::isfinite(prev) ? std::fabs(a1 - prev) : so.ritz_change;
    tot.e0 = a1;          // exact Rayleigh quotient of this round's start vector
    tot.residual = b1;    // and its exact residual
    if (stop) {
      tot.converged = true;
      break;
    }
    prev = a1;
  }
  if (!tot.converged) {
    // out of rounds: one last FP64 product fixes the energy and residual of
    // the vector that is returned
    LanczosOpts last;
    last.first_prec = PREC_FP64;
    last.after_first = [&](double a, double b) {
      a1 = a;
      b1 = b;
      return true;
    };
    const SolveOut so = lanczos(ctx, sb, psi, lanzero, max_iter, last);
    tot.iters += so.iters;
    tot.mv_seconds += so.mv_seconds;
    tot.hpsi_fp64 += so.hpsi_fp64;
    tot.ritz_change = std::fabs(a1 - prev);
    tot.e0 = a1;
    tot.residual = b1;
  }
  ctx->last_converged = tot.converged ? 1 : 0;
  ctx->last_ritz_change = tot.ritz_change;
  ctx->last_residual = tot.residual;
  return tot;
}

}  // namespace hh

namespace hh {

// Eigendecomposition of every sector matrix (ZHEEV, code/get_den.f90:242),
// ascending eigenvalues and eigenvectors in place.  The sectors are
// independent.  A divide-and-conquer zheevd is a long chain of small kernels
// (about 20-30 us per column) with host synchronisations inside, so the
// sectors are spread over the context stream and its side streams (largest
// first, each to the stream with the fewest columns so far) and every stream
// is driven by its own host thread.  Sectors are not batched: a strided-
// batched zheevd of sectors padded to one size returns NaN eigenvectors
// (with info = 0) for every padded member once the size reaches the
// divide-and-conquer range, whatever the pad block, because the padded
// matrix is block diagonal.
// rocBLAS / rocSOLVER set up kernels and workspaces lazily on first use.
// The library eigensolver runs on several handles from several host
// threads, so that first use happens here, before its first call on the
// context, once per handle, on one thread: every handle solves small
// diagonal problems of the sizes that select each of zheevd's code paths.
void warm_eigensolver(hh_ctx* c) {
  rocblas_initialize();
  const int sizes[3] = {8, 100, 600};
  const int nmax = 600;
  DevMem a, w, e, inf;
  a.alloc(sizeof(double2) * nmax * nmax);
  w.alloc(sizeof(double) * nmax);
  e.alloc(sizeof(double) * nmax);
  inf.alloc(sizeof(rocblas_int));
  std::vector<double2> h((size_t)nmax * nmax);
  rocblas_handle hs[hh_ctx::NSIDE + 1] = {c->rb};
  hipStream_t ss[hh_ctx::NSIDE + 1] = {c->stream};
  for (int k = 0; k < hh_ctx::NSIDE; ++k) {
    hs[k + 1] = c->rbs[k];
    ss[k + 1] = c->side[k];
  }
  for (int k = 0; k <= hh_ctx::NSIDE; ++k)
    for (int n : sizes) {
      for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) h[i + (size_t)j * n] = make_double2(i == j ? 1.0 + i : 0.0, 0.0);
      check_hip(hipMemcpyAsync(a.p, h.data(), sizeof(double2) * n * n, hipMemcpyHostToDevice, ss[k]), "h2d");
      if (rocsolver_zheevd(hs[k], rocblas_evect_original, rocblas_fill_upper, n, (rocblas_double_complex*)a.p, n,
                           (double*)w.p, (double*)e.p, (rocblas_int*)inf.p) != rocblas_status_success)
        throw HHError(-2, "rocsolver_zheevd warm-up failed");
      check_hip(hipStreamSynchronize(ss[k]), "eigensolver warm-up");
    }
}

template <typename OnFail>
void heev_sectors(hh_ctx* ctx, double2* evec, const std::vector<int64_t>& roff, const std::vector<int>& rdim,
                  double* evals, double* work, int64_t work_stride, rocblas_int* info, OnFail on_fail) {
  if (!ctx->eig_warm) {   // the library's lazy set-up, once, before any side thread uses a handle
    warm_eigensolver(ctx);
    ctx->eig_warm = true;
  }
  const size_t ns = rdim.size();
  std::vector<int64_t> eo(ns, 0);
  for (size_t i = 1; i < ns; ++i) eo[i] = eo[i - 1] + rdim[i - 1];
  std::vector<int> idx(ns);
  std::iota(idx.begin(), idx.end(), 0);
  std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return rdim[a] > rdim[b]; });
  constexpr int NS = hh_ctx::NSIDE + 1;
  rocblas_handle h[NS] = {ctx->rb};
  for (int k = 1; k < NS; ++k) h[k] = ctx->rbs[k - 1];
  std::vector<int> mine[NS];
  double load[NS] = {0.0};
  for (int i : idx) {
    int k = 0;
    for (int q = 1; q < NS; ++q)
      if (load[q] < load[k]) k = q;
    load[k] += 64.0 + rdim[i];
    mine[k].push_back(i);
  }
  // the density matrices were written on the context stream
  check_hip(hipEventRecord(ctx->fork, ctx->stream), "event");
  for (int k = 1; k < NS; ++k) check_hip(hipStreamWaitEvent(ctx->side[k - 1], ctx->fork, 0), "wait");
  std::atomic<bool> ok{true};
  auto run = [&](int k) {
    if (hipSetDevice(ctx->device) != hipSuccess) {
      ok = false;
      return;
    }
    for (int i : mine[k])
      if (rocsolver_zheevd(h[k], rocblas_evect_original, rocblas_fill_upper, rdim[i],
                           (rocblas_double_complex*)(evec + roff[i]), rdim[i], evals + eo[i],
                           work + (int64_t)k * work_stride, info + i) != rocblas_status_success)
        ok = false;
  };
  // HHDMRG_EIG_THREADS=0 issues every stream's calls from this thread (the
  // streams still overlap where zheevd does not synchronise); used under
  // tools that intercept kernel launches and are not safe with concurrent
  // launching threads
  static const bool threads = !(std::getenv("HHDMRG_EIG_THREADS") && std::getenv("HHDMRG_EIG_THREADS")[0] == '0');
  if (threads) {
    std::vector<std::thread> th;
    for (int k = 1; k < NS; ++k)
      if (!mine[k].empty()) th.emplace_back(run, k);
    run(0);
    for (auto& t : th) t.join();
  } else {
    for (int k = 0; k < NS; ++k) run(k);
  }
  for (int k = 1; k < NS; ++k) {
    check_hip(hipEventRecord(ctx->join[k - 1], ctx->side[k - 1]), "event");
    check_hip(hipStreamWaitEvent(ctx->stream, ctx->join[k - 1], 0), "wait");
  }
  if (!ok) {
    check_hip(hipStreamSynchronize(ctx->stream), "eig");
    on_fail();
    throw HHError(-2, "rocsolver_zheevd failed");
  }
}

// Get_density_matrix_sys_new / _env_new with levels >= 2 (code/get_den.f90:
// 103-177): rho = (0.1 rho(psi_1) + 1.9 rho(psi_2)) / 2, with rho(psi_1)
// already in `rho` (`dt` is the density task list).  0.1 and 1.9 are
// default-real literals there, hence the float constants.
void mix_levels(hh_ctx* ctx, const DevTasks& dt, double2* rho, const double2* psi2, int64_t total) {
  DevMem tmp;
  tmp.alloc(sizeof(double2) * (size_t)std::max<int64_t>(total, 1));
  launch_tiles(dt, bufs_of((double2*)psi2, tmp.c(), nullptr, nullptr, nullptr, nullptr, nullptr, nullptr),
               ctx->stream, PREC_FP64);
  dev_scale(rho, 0.5 * (double)0.1f, total, ctx->stream);
  dev_lanczos_update(rho, tmp.c(), -0.5 * (double)1.9f, nullptr, 0.0, total, ctx->stream);
  check_hip(hipStreamSynchronize(ctx->stream), "density matrix of two levels");
}

// Reduced density matrix + Get_truncation_operator (code/get_den.f90:184-428).
// The eigendecomposition (ZHEEV per sector, get_den.f90:242) runs in the
// hand-written eigensolver (hh_heev.hip: every eigenvalue, then the
// eigenvectors of the kept ones only) or, with eig_solver 1 or for a
// truncation the hand-written solver hands back, rocSOLVER zheevd per sector.
hh_trunc* truncate(hh_ctx* ctx, hh_superblock* sb, int side, const double2* psi, int kept_min, int kept_max,
                   double trun_err, const double2* psi2) {
  const SuperBasis& S = sb->plan.sb;
  const double t0 = wall();
  std::vector<int> rsec, rdim;
  std::vector<int64_t> roff;
  TaskList tl = compile_density(S, side, rsec, roff, rdim);
  int64_t total = 0;
  for (int d : rdim) total += (int64_t)d * d;
  std::unique_ptr<hh_trunc> t(new hh_trunc());
  t->side = side;
  t->eb = side == 0 ? sb->plan.sbs : sb->plan.ebs;
  t->evec.alloc(sizeof(double2) * (size_t)std::max<int64_t>(total, 1));
  DevTasks dt;
  dt.upload(tl, ctx->stream);
  launch_tiles(dt, bufs_of((double2*)psi, t->evec.c(), nullptr, nullptr, nullptr, nullptr, nullptr, nullptr),
               ctx->stream, PREC_FP64);
  if (psi2) mix_levels(ctx, dt, t->evec.c(), psi2, total);
  int64_t nev = 0;
  int dmax = 1;
  for (int d : rdim) {
    nev += d;
    dmax = std::max(dmax, d);
  }
  const bool hand = ctx->eig_solver == 0 && dmax <= HEEV_MAXN;
  if (ctx->eig_solver == 0 && !hand) ++ctx->eig_handbacks;
  DevMem evals, work, info;
  evals.alloc(sizeof(double) * (size_t)std::max<int64_t>(nev, 1));
  double t1 = 0.0;
  if (verbose()) {
    check_hip(hipStreamSynchronize(ctx->stream), "density matrix");
    t1 = wall();
  }
  std::vector<rocblas_int> inf;
  if (hand) {
    ctx->heev.timing = verbose();
    ctx->heev.values(ctx->stream, t->evec.c(), roff, rdim, (double*)evals.p);
  } else {
    work.alloc(sizeof(double) * (size_t)dmax * (hh_ctx::NSIDE + 1));
    info.alloc(sizeof(rocblas_int) * rdim.size() + 16);
    heev_sectors(ctx, t->evec.c(), roff, rdim, (double*)evals.p, (double*)work.p, dmax, (rocblas_int*)info.p, []() {});
    inf.resize(rdim.size());
    check_hip(hipMemcpyAsync(inf.data(), info.p, sizeof(rocblas_int) * rdim.size(), hipMemcpyDeviceToHost,
                             ctx->stream), "d2h");
  }
  std::vector<double> ev(nev);
  check_hip(hipMemcpyAsync(ev.data(), evals.p, sizeof(double) * nev, hipMemcpyDeviceToHost, ctx->stream), "d2h");
  check_hip(hipStreamSynchronize(ctx->stream), "eig");
  const double t2 = verbose() ? wall() : 0.0;
  for (auto x : inf)
    if (x != 0) throw HHError(-2, "rocsolver_zheevd did not converge");
  for (double x : ev)
    if (!std::isfinite(x)) throw HHError(-2, "the density-matrix eigensolver returned a non-finite eigenvalue");
  // global descending sort, ties by position in the concatenated ascending
  // per-sector spectra (deterministic stand-in for QuickSort_Double)
  std::vector<int64_t> order(nev);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int64_t a, int64_t b) {
    if (ev[a] != ev[b]) return ev[a] > ev[b];
    return a < b;
  });
  std::vector<double> sv(nev);
  double tot = 0.0;
  for (int64_t i = 0; i < nev; ++i) {
    sv[i] = ev[order[i]];
    tot += sv[i];
  }
  for (auto& x : sv) x /= tot;
  int64_t num;
  double err;
  if (nev <= kept_min) {
    num = nev;
    err = 0.0;
  } else {
    err = 0.0;
    for (int64_t i = kept_min; i < nev; ++i) err += sv[i];
    num = kept_min;
    while (err > trun_err) {
      ++num;
      err -= sv[num - 1];
      if (err <= trun_err || num >= kept_max || num >= nev) break;
    }
  }
  double ent = 0.0;
  for (double x : sv)
    if (x >= 1.0e-15) ent -= x * std::log(x);
  std::vector<int> keep(rdim.size(), 0);
  {
    std::vector<int> owner(nev);
    int64_t o = 0;
    for (size_t i = 0; i < rdim.size(); ++i)
      for (int k = 0; k < rdim[i]; ++k) owner[o++] = (int)i;
    for (int64_t i = 0; i < num; ++i) keep[owner[order[i]]]++;
  }
  if (hand && !ctx->heev.vectors(ctx->stream, t->evec.c(), keep, (double*)evals.p)) {
    // the kept vectors could not be orthonormalised: redo this truncation
    // (density matrix included) with zheevd
    ++ctx->eig_handbacks;
    if (verbose()) {
      std::fprintf(stderr, "# truncation side=%d: the hand-written eigensolver hands back (%s)\n", side,
                   ctx->heev.fail_info.c_str());
      std::fflush(stderr);
    }
    const int saved = ctx->eig_solver;
    ctx->eig_solver = 1;
    try {
      hh_trunc* r = truncate(ctx, sb, side, psi, kept_min, kept_max, trun_err, psi2);
      ctx->eig_solver = saved;
      return r;
    } catch (...) {
      ctx->eig_solver = saved;
      throw;
    }
  }
  if (verbose()) {
    if (hand) check_hip(hipStreamSynchronize(ctx->stream), "eigenvectors");
    std::fprintf(stderr, "# truncation side=%d: density matrix %.4f s, %s %.4f s + vectors %.4f s (%zu sectors, "
                 "largest %d, %lld columns, %lld kept%s)\n", side, t1 - t0, hand ? "eigenvalues" : "zheevd", t2 - t1,
                 wall() - t2, rdim.size(), dmax, (long long)nev, (long long)num,
                 hand ? (", " + std::to_string(ctx->heev.ns_iterations) + " orthonormalisation steps, " +
                         std::to_string(ctx->heev.unresolved) + " completed, largest cluster " +
                         std::to_string(ctx->heev.max_cluster) + "; ms: reduction " +
                         std::to_string(ctx->heev.ms[0]) + " bisection " + std::to_string(ctx->heev.ms[1]) +
                         " inverse iteration " + std::to_string(ctx->heev.ms[2]) + " orthonormalisation " +
                         std::to_string(ctx->heev.ms[3]) + " back-transformation " +
                         std::to_string(ctx->heev.ms[4])).c_str() : "");
    std::fflush(stderr);
  }
  for (size_t i = 0; i < rdim.size(); ++i) {
    if (keep[i] == 0) continue;
    KeptSector ks;
    ks.a = rsec[i];
    ks.keep = keep[i];
    ks.u_off = roff[i] + (int64_t)(rdim[i] - keep[i]) * rdim[i];
    ks.u_ld = rdim[i];
    t->kept.push_back(ks);
  }
  t->kept_per_rho = keep;
  t->num_state = (int)num;
  t->err = err;
  t->entropy = ent;
  return t.release();
}

hh_trunc* trunc_identity(hh_ctx* ctx, hh_superblock* sb, int side) {
  auto* t = new hh_trunc();
  t->side = side;
  t->eb = side == 0 ? sb->plan.sbs : sb->plan.ebs;
  int64_t total = 0;
  for (auto& s : t->eb.sec) total += (int64_t)s.dim * s.dim;
  std::vector<double2> h((size_t)std::max<int64_t>(total, 1), make_double2(0.0, 0.0));
  int64_t o = 0;
  for (int a = 0; a < (int)t->eb.sec.size(); ++a) {
    const int d = t->eb.sec[a].dim;
    for (int i = 0; i < d; ++i) h[o + (int64_t)i * d + i] = make_double2(1.0, 0.0);
    t->kept.push_back({a, d, o, d});
    t->num_state += d;
    o += (int64_t)d * d;
  }
  t->evec.alloc(sizeof(double2) * h.size());
  check_hip(hipMemcpyAsync(t->evec.p, h.data(), sizeof(double2) * h.size(), hipMemcpyHostToDevice, ctx->stream), "h2d");
  check_hip(hipStreamSynchronize(ctx->stream), "sync");
  return t;
}

// The truncation's enlarged basis is the one it is applied to: same sectors
// (N, 2S, dim) in the same order.  Guards the kernels against a truncation
// of another block (its offsets and leading dimensions index that basis).
bool same_sectors(const EnlBasis& a, const EnlBasis& b) {
  if (a.sec.size() != b.sec.size()) return false;
  for (size_t i = 0; i < a.sec.size(); ++i)
    if (a.sec[i].N != b.sec[i].N || a.sec[i].S2 != b.sec[i].S2 || a.sec[i].dim != b.sec[i].dim) return false;
  return true;
}

hh_block* grow(hh_ctx* ctx, hh_superblock* sb, int side, const hh_trunc* t) {
  need(t->side == side, "truncation side mismatch");
  need(same_sectors(t->eb, side == 0 ? sb->plan.sbs : sb->plan.ebs),
       "the truncation was made for another enlarged block than this superblock's");
  const hh_block* blk = side == 0 ? sb->sys : sb->env;
  GrowPlan gp = compile_grow(ctx->model, sb->g, blk->d, t->kept);
  auto* nb = new hh_block();
  nb->d = gp.out;
  nb->arena.alloc(sizeof(double2) * (size_t)std::max<int64_t>(nb->d.size, 1));
  DevMem mrg, X;
  mrg.alloc(sizeof(double2) * (size_t)std::max<int64_t>(gp.merged_size, 1));
  X.alloc(sizeof(double2) * (size_t)std::max<int64_t>(gp.x_size, 1));
  DevTasks a, b, c;
  a.upload(gp.merge_tasks, ctx->stream);
  b.upload(gp.stage1, ctx->stream);
  c.upload(gp.stage2, ctx->stream);
  BufPtrs bp = bufs_of(nullptr, nb->arena.c(), blk, nullptr, X.c(), mrg.c(), t->evec.c(), nullptr);
  launch_tiles(a, bp, ctx->stream, PREC_FP64);
  launch_tiles(b, bp, ctx->stream, PREC_FP64);
  launch_tiles(c, bp, ctx->stream, PREC_FP64);
  check_hip(hipStreamSynchronize(ctx->stream), "grow");
  return nb;
}

void wave_trans(hh_ctx* ctx, int dir, hh_superblock* osb, const double2* psi_old, const hh_trunc* st,
                const hh_trunc* et, hh_superblock* nsb, double2* psi_new) {
  need(st->side == 0 && et->side == 1, "wave_trans needs a sys and an env truncation");
  // LR: st truncates the old sys enlarged block, et built the old env block
  // from the new env enlarged block; RL: st built the old sys block from the
  // new sys enlarged block, et truncates the old env enlarged block
  const EnlBasis& sb_basis = dir == 0 ? osb->plan.sbs : nsb->plan.sbs;
  const EnlBasis& eb_basis = dir == 0 ? nsb->plan.ebs : osb->plan.ebs;
  need(same_sectors(st->eb, sb_basis) && same_sectors(et->eb, eb_basis),
       "wave_trans: a truncation does not belong to these superblocks");
  TransPlan tp = dir == 0 ? compile_wave_trans_lr(osb->plan.sb, osb->plan.sbs, osb->plan.ebs, st->kept, et->kept,
                                                  nsb->plan.sbs, nsb->plan.ebs, nsb->plan.sb)
                          : compile_wave_trans_rl(osb->plan.sb, osb->plan.sbs, osb->plan.ebs, st->kept, et->kept,
                                                  nsb->plan.sbs, nsb->plan.ebs, nsb->plan.sb);
  DevMem X;
  X.alloc(sizeof(double2) * (size_t)std::max<int64_t>(tp.x_size, 1));
  DevTasks a, b;
  a.upload(tp.stage1, ctx->stream);
  b.upload(tp.stage2, ctx->stream);
  BufPtrs bp = bufs_of((double2*)psi_old, psi_new, nullptr, nullptr, X.c(), nullptr, st->evec.c(), et->evec.c());
  launch_tiles(a, bp, ctx->stream, PREC_FP64);
  launch_tiles(b, bp, ctx->stream, PREC_FP64);
  ensure_ws(ctx, 1, 1);
  const double n2 = rdot(ctx, psi_new, psi_new, nsb->plan.sb.dim);
  if (n2 > 0.0) dev_scale(psi_new, 1.0 / std::sqrt(n2), nsb->plan.sb.dim, ctx->stream);
  check_hip(hipStreamSynchronize(ctx->stream), "wave_trans");
}

hh_superblock* bind(hh_ctx* ctx, const hh_block* sys, const hh_block* env, int N) {
  need(sys->d.side == 0 && env->d.side == 1, "bind expects a sys block and an env block");
  auto* s = new hh_superblock();
  s->sys = sys;
  s->env = env;
  s->N = N;
  s->g = make_geometry(ctx->model, sys->d.length, env->d.length);
  s->plan = compile_matvec(ctx->model, s->g, sys->d, env->d, N);
  s->merged.alloc(sizeof(double2) * (size_t)std::max<int64_t>(s->plan.merged_size, 1));
  s->X.alloc(sizeof(double2) * (size_t)std::max<int64_t>(s->plan.x_size, 1));
  s->merge.upload(s->plan.merge_tasks, ctx->stream);
  const LaunchConfig cfg = ctx->forced_cfg ? ctx->forced : LaunchConfig();
  s->st1.upload(s->plan.stage1, ctx->stream, nullptr, 0, cfg);
  s->st2.upload(s->plan.stage2, ctx->stream, nullptr, 0, cfg);
  launch_tiles(s->merge, bufs_of(nullptr, nullptr, sys, env, nullptr, s->merged.c(), nullptr, nullptr), ctx->stream,
               PREC_FP64);
  check_hip(hipStreamSynchronize(ctx->stream), "bind");
  if (ctx->autotune && !ctx->forced_cfg) {
    try {
      autotune_superblock(ctx, s);
    } catch (...) {
      delete s;
      throw;
    }
  }
  return s;
}

// the C model struct as the task compilers take it
ModelParams to_model(const hh_model_params* p) {
  ModelParams m;
  m.kind = p->kind;
  m.Lx = p->Lx;
  m.Ly = p->Ly;
  m.lattice = p->lattice;
  m.periodic_y = p->periodic_y;
  m.n_elec = p->n_elec;
  m.t = p->t;
  m.U = p->U;
  m.J = p->J;
  m.flux = p->flux;
  m.gauge = p->gauge;
  m.theta = p->theta;
  return m;
}

// the gauge fields of a C model struct checked
void check_gauge(const hh_model_params* p) {
  need(p->gauge == GAUGE_LANDAU || p->gauge == GAUGE_REFERENCE, "gauge must be 0 (Landau) or 1 (get_phase)");
  need(p->reserved == 0, "reserved field of hh_model_params must be 0");
  if (p->gauge == GAUGE_REFERENCE) {
    need(p->lattice == TRIANGULAR, "gauge 1 (get_phase) is defined on the triangular lattice only");
    need(p->flux == 0.0, "gauge 1 takes its flux from theta; flux must be 0");
    need(std::isfinite(p->theta), "theta must be finite");
  }
}

void launch_on(hh_ctx* ctx, const TaskList& tl, const BufPtrs& b) {
  DevTasks dt;
  dt.upload(tl, ctx->stream);
  launch_tiles(dt, b, ctx->stream, PREC_FP64);
  check_hip(hipStreamSynchronize(ctx->stream), "measurement");
}

hh_block* site_block(hh_ctx* ctx, int side) {
  const ModelParams& m = ctx->model;
  auto st = site_sectors(m.kind);
  std::vector<int> sites = boundary_sites(m, 1);
  hh_block* b = new_block(ctx, side, 1, st, sites);
  std::vector<double2> h((size_t)b->d.size, make_double2(0.0, 0.0));
  if (m.kind == HUBBARD) h[b->d.h_off[2]] = make_double2(m.U, 0.0);
  const int nk = n_op_kinds(m.kind);
  for (size_t si = 0; si < b->d.sites.size(); ++si)
    for (int k = 0; k < nk; ++k)
      for (auto& sub : b->d.lay[k].subs) {
        cplx v = site_val(m.kind, k, sub.bra, sub.ket);
        h[b->d.op_off[si][k] + sub.off] = make_double2(v.real(), v.imag());
      }
  check_hip(hipMemcpyAsync(b->arena.p, h.data(), sizeof(double2) * h.size(), hipMemcpyHostToDevice, ctx->stream),
            "h2d");
  check_hip(hipStreamSynchronize(ctx->stream), "sync");
  return b;
}

// A sweep schedule the driver can run (from the caller or from a checkpoint).
void check_schedule(const hh_schedule& s) {
  need(s.kept_max >= 1 && s.kept_max_warm >= 0 && s.n_sweeps >= 0, "bad schedule: kept counts / sweeps");
  need(s.max_iter >= 2 && s.max_iter <= 1000, "bad schedule: max_iter must be in [2, 1000]");
  need(std::isfinite(s.tol) && s.tol >= 0.0 && std::isfinite(s.warm_tol_factor) && s.warm_tol_factor > 0.0 &&
           std::isfinite(s.first_tol_factor) && s.first_tol_factor > 0.0,
       "bad schedule: tolerances");
  need(s.precision == PREC_FP64, "the driver's H.psi precision must be 0; the solver field selects the mixed forms");
  need(s.solver >= 0 && s.solver <= 2, "bad schedule: solver must be 0, 1 or 2");
}


// message of a failed hh_ctx_create (the context no longer exists)
std::string& create_error() {
  static thread_local std::string s;
  return s;
}


}  // namespace

// ============================================================================
// C ABI
// ============================================================================
extern "C" {


int hh_ctx_create(int device, hh_ctx** out) {
  if (!out) return 1;
  auto* c = new hh_ctx();
  c->device = device;
  int rc = guard(c, [&] {
    int n = 0;
    check_hip(hipGetDeviceCount(&n), "hipGetDeviceCount");
    if (device < 0 || device >= n) throw HHError(-1, "no such HIP device");
    check_hip(hipSetDevice(device), "hipSetDevice");
    // device address of the allocation-free entry point's fixed scratch,
    // once per device, with the device current
    (void)noalloc_scratch_init(device);
    check_hip(hipStreamCreate(&c->stream), "hipStreamCreate");
    if (rocblas_create_handle(&c->rb) != rocblas_status_success) throw HHError(-2, "rocblas_create_handle");
    rocblas_set_stream(c->rb, c->stream);
    for (auto& e : c->ev) check_hip(hipEventCreate(&e), "hipEventCreate");
    check_hip(hipEventCreateWithFlags(&c->fork, hipEventDisableTiming), "hipEventCreate");
    for (int k = 0; k < hh_ctx::NSIDE; ++k) {
      check_hip(hipStreamCreateWithFlags(&c->side[k], hipStreamNonBlocking), "hipStreamCreate");
      if (rocblas_create_handle(&c->rbs[k]) != rocblas_status_success) throw HHError(-2, "rocblas_create_handle");
      rocblas_set_stream(c->rbs[k], c->side[k]);
      check_hip(hipEventCreateWithFlags(&c->join[k], hipEventDisableTiming), "hipEventCreate");
    }
  });
  if (rc != 0) {
    create_error() = c->err;
    hh_ctx_destroy(c);
    *out = nullptr;
    return rc;
  }
  *out = c;
  return 0;
}

int hh_ctx_destroy(hh_ctx* ctx) {
  if (!ctx) return 0;
  for (auto* m : ctx->user_allocs) delete m;
  for (auto& e : ctx->ev)
    if (e) (void)hipEventDestroy(e);
  if (ctx->rb) rocblas_destroy_handle(ctx->rb);
  for (int k = 0; k < hh_ctx::NSIDE; ++k) {
    if (ctx->rbs[k]) rocblas_destroy_handle(ctx->rbs[k]);
    if (ctx->side[k]) (void)hipStreamDestroy(ctx->side[k]);
    if (ctx->join[k]) (void)hipEventDestroy(ctx->join[k]);
  }
  if (ctx->fork) (void)hipEventDestroy(ctx->fork);
  ctx->V.reset();
  ctx->dots.reset();
  ctx->partial.reset();
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return 0;
}

const char* hh_last_error(const hh_ctx* ctx) { return ctx ? ctx->err.c_str() : create_error().c_str(); }

int hh_boundary_sites(hh_ctx* ctx, int length, int32_t* n, int32_t* sites) {
  return guard(ctx, [&] {
    need(ctx->model_set, "set the model first");
    need(n != nullptr, "null count");
    need(length >= 1 && length < ctx->model.nsite(), "block length out of range");
    const std::vector<int> b = boundary_sites(ctx->model, length);
    *n = (int32_t)b.size();
    if (sites)
      for (size_t i = 0; i < b.size(); ++i) sites[i] = b[i];
  });
}

int hh_set_model(hh_ctx* ctx, const hh_model_params* p) {
  return guard(ctx, [&] {
    need(p != nullptr, "null model");
    need(p->kind == HUBBARD || p->kind == TJ, "model kind must be 0 (Hubbard) or 1 (t-J)");
    need(p->Lx >= 2 && p->Ly >= 1, "lattice must have Lx >= 2, Ly >= 1");
    need(p->n_elec >= 0 && p->n_elec <= 2 * p->Lx * p->Ly && p->n_elec % 2 == 0, "n_elec must be even (S = 0)");
    need(p->lattice == SQUARE || p->lattice == TRIANGULAR, "lattice must be 0 (square) or 1 (triangular)");
    check_gauge(p);
    ctx->model = to_model(p);
    ctx->model_set = true;
  });
}

int hh_dev_alloc(hh_ctx* ctx, int64_t bytes, void** dev) {
  return guard(ctx, [&] {
    need(dev && bytes >= 0, "bad allocation request");
    check_hip(hipMalloc(dev, (size_t)std::max<int64_t>(bytes, 16)), "hipMalloc");
  });
}
int hh_dev_free(hh_ctx* ctx, void* dev) {
  return guard(ctx, [&] { check_hip(hipFree(dev), "hipFree"); });
}
int hh_memcpy_h2d(hh_ctx* ctx, void* dev, const void* host, int64_t bytes) {
  return guard(ctx, [&] {
    check_hip(hipMemcpyAsync(dev, host, bytes, hipMemcpyHostToDevice, ctx->stream), "h2d");
    check_hip(hipStreamSynchronize(ctx->stream), "sync");
  });
}
int hh_memcpy_d2h(hh_ctx* ctx, void* host, const void* dev, int64_t bytes) {
  return guard(ctx, [&] {
    check_hip(hipMemcpyAsync(host, dev, bytes, hipMemcpyDeviceToHost, ctx->stream), "d2h");
    check_hip(hipStreamSynchronize(ctx->stream), "sync");
  });
}

int hh_block_site(hh_ctx* ctx, int side, hh_block** out) {
  return guard(ctx, [&] {
    need(ctx->model_set, "hh_set_model first");
    need(side == 0 || side == 1, "side must be 0 or 1");
    *out = site_block(ctx, side);
  });
}

int hh_block_create(hh_ctx* ctx, int side, int length, int nsec, const int32_t* N, const int32_t* S2,
                    const int32_t* dim, int nsites, const int32_t* sites, hh_block** out) {
  return guard(ctx, [&] {
    need(ctx->model_set, "hh_set_model first");
    need(side == 0 || side == 1, "side must be 0 or 1");
    need(nsec >= 0 && length >= 1, "bad block shape");
    std::vector<Sector> sec;
    for (int i = 0; i < nsec; ++i) {
      need(dim[i] >= 0 && S2[i] >= 0 && (N[i] - S2[i]) % 2 == 0, "sector (N, 2S) parity / dim error");
      sec.push_back({N[i], S2[i], dim[i]});
    }
    std::vector<int> st;
    for (int i = 0; i < nsites; ++i) {
      need(sites[i] >= 1 && sites[i] <= length, "boundary site outside the block");
      st.push_back(sites[i]);
    }
    *out = new_block(ctx, side, length, sec, st);
  });
}

int hh_block_destroy(hh_block* b) {
  delete b;
  return 0;
}

int hh_block_info(const hh_block* b, int64_t* arena, int32_t* nsec, int32_t* nsites, int32_t* length) {
  if (!b) return 1;
  if (arena) *arena = b->d.size;
  if (nsec) *nsec = (int32_t)b->d.sec.size();
  if (nsites) *nsites = (int32_t)b->d.sites.size();
  if (length) *length = b->d.length;
  return 0;
}
int hh_block_sectors(const hh_block* b, int32_t* N, int32_t* S2, int32_t* dim, int64_t* h_off) {
  if (!b) return 1;
  for (size_t i = 0; i < b->d.sec.size(); ++i) {
    if (N) N[i] = b->d.sec[i].N;
    if (S2) S2[i] = b->d.sec[i].S2;
    if (dim) dim[i] = b->d.sec[i].dim;
    if (h_off) h_off[i] = b->d.h_off[i];
  }
  return 0;
}
int hh_block_sites(const hh_block* b, int32_t* sites) {
  if (!b) return 1;
  for (size_t i = 0; i < b->d.sites.size(); ++i) sites[i] = b->d.sites[i];
  return 0;
}
int hh_block_layout(const hh_block* b, int kind, int32_t* nsub, int32_t* bra, int32_t* ket, int64_t* off) {
  if (!b || kind < 0 || kind > 2) return 1;
  const OpLayout& L = b->d.lay[kind];
  if (nsub) *nsub = (int32_t)L.subs.size();
  for (size_t i = 0; i < L.subs.size(); ++i) {
    if (bra) bra[i] = L.subs[i].bra;
    if (ket) ket[i] = L.subs[i].ket;
    if (off) off[i] = L.subs[i].off;
  }
  return 0;
}
int hh_block_op_offset(const hh_block* b, int site_index, int kind, int64_t* off) {
  if (!b || site_index < 0 || site_index >= (int)b->d.sites.size() || kind < 0 || kind > 2) return 1;
  *off = b->d.op_off[site_index][kind];
  return *off < 0 ? 1 : 0;
}
int hh_block_upload(hh_ctx* ctx, hh_block* b, const double* src) {
  return guard(ctx, [&] {
    need(b && src, "null argument");
    b->have32 = false;   // the complex64 copy, if any, is stale now
    check_hip(hipMemcpyAsync(b->arena.p, src, sizeof(double2) * b->d.size, hipMemcpyDefault, ctx->stream), "upload");
    check_hip(hipStreamSynchronize(ctx->stream), "sync");
  });
}
int hh_block_download(hh_ctx* ctx, const hh_block* b, double* dst) {
  return guard(ctx, [&] {
    need(b && dst, "null argument");
    check_hip(hipMemcpyAsync(dst, b->arena.p, sizeof(double2) * b->d.size, hipMemcpyDefault, ctx->stream),
              "download");
    check_hip(hipStreamSynchronize(ctx->stream), "sync");
  });
}

int hh_superblock_bind(hh_ctx* ctx, const hh_block* sys, const hh_block* env, int N, hh_superblock** out) {
  return guard(ctx, [&] {
    need(ctx->model_set, "hh_set_model first");
    need(sys && env && out, "null argument");
    *out = bind(ctx, sys, env, N);
  });
}
int hh_superblock_destroy(hh_superblock* sb) {
  delete sb;
  return 0;
}
int hh_superblock_info(const hh_superblock* sb, int64_t* dim, int32_t* npairs, double* flops_alg,
                       double* flops_exec) {
  if (!sb) return 1;
  if (dim) *dim = sb->plan.sb.dim;
  if (npairs) *npairs = (int32_t)sb->plan.sb.pairs.size();
  if (flops_alg) *flops_alg = sb->plan.stage1.flops_alg + sb->plan.stage2.flops_alg;
  if (flops_exec) *flops_exec = sb->plan.stage1.flops_exec + sb->plan.stage2.flops_exec;
  return 0;
}
int hh_superblock_pairs(const hh_superblock* sb, int32_t* sys_sec, int32_t* env_sec, int32_t* rows, int32_t* cols,
                        int64_t* off) {
  if (!sb) return 1;
  const SuperBasis& S = sb->plan.sb;
  for (size_t i = 0; i < S.pairs.size(); ++i) {
    if (sys_sec) sys_sec[i] = S.pairs[i].first;
    if (env_sec) env_sec[i] = S.pairs[i].second;
    if (rows) rows[i] = S.rows[i];
    if (cols) cols[i] = S.cols[i];
    if (off) off[i] = S.off[i];
  }
  return 0;
}
int hh_superblock_enlarged(const hh_superblock* sb, int side, int32_t* nsec, int32_t* N, int32_t* S2, int32_t* dim) {
  if (!sb) return 1;
  const EnlBasis& E = side == 0 ? sb->plan.sbs : sb->plan.ebs;
  if (nsec) *nsec = (int32_t)E.sec.size();
  for (size_t i = 0; i < E.sec.size(); ++i) {
    if (N) N[i] = E.sec[i].N;
    if (S2) S2[i] = E.sec[i].S2;
    if (dim) dim[i] = E.sec[i].dim;
  }
  return 0;
}

int hh_model_wave(hh_ctx* ctx, hh_superblock* sb, const double* psi_in, double* psi_out) {
  return guard(ctx, [&] {
    need(psi_in && psi_out && psi_in != psi_out, "model_wave needs distinct input and output vectors");
    matvec(ctx, sb, (const double2*)psi_in, (double2*)psi_out);
    check_hip(hipStreamSynchronize(ctx->stream), "model_wave");
  });
}

int hh_lanczos(hh_ctx* ctx, hh_superblock* sb, double* psi, double lanzero, int max_iter, double* e0, int* iters) {
  return guard(ctx, [&] {
    need(psi && e0 && iters, "null argument");
    const SolveOut so = lanczos(ctx, sb, (double2*)psi, lanzero, max_iter);
    *e0 = so.e0;
    *iters = so.iters;
  });
}

int hh_lanczos_levels(hh_ctx* ctx, hh_superblock* sb, double* const* psi, int levels, double lanzero, int max_iter,
                      double* e, int32_t* iters) {
  return guard(ctx, [&] {
    need(sb && psi && e && iters, "null argument");
    need(levels >= 1 && 2 * (levels - 1) + 2 < max_iter + 2, "levels must be >= 1 and below max_iter / 2");
    const int64_t n = sb->plan.sb.dim;
    for (int l = 0; l < levels; ++l) need(psi[l] != nullptr, "null level vector");
    DevMem defl;
    if (levels > 1) defl.alloc(sizeof(double2) * (size_t)n * (size_t)(2 * (levels - 1)));
    for (int l = 0; l < levels; ++l) {
      LanczosOpts opt;
      opt.deflate = defl.c();
      opt.ndeflate = 2 * l;
      const SolveOut so = lanczos(ctx, sb, (double2*)psi[l], lanzero, max_iter, opt);
      e[l] = so.e0;
      iters[l] = so.iters;
      if (l + 1 < levels) {
        // this level and i times it join the deflation set (psi[l] is unit)
        dev_scale_copy(defl.c() + (size_t)(2 * l) * n, (const double2*)psi[l], 1.0, n, ctx->stream);
        dev_cscale_copy(defl.c() + (size_t)(2 * l + 1) * n, (const double2*)psi[l], 0.0, 1.0, n, ctx->stream);
      }
    }
    check_hip(hipStreamSynchronize(ctx->stream), "lanczos levels");
  });
}

int hh_lanczos_status(const hh_ctx* ctx, int32_t* converged, double* ritz_change, double* residual) {
  if (!ctx) return 1;
  if (converged) *converged = ctx->last_converged;
  if (ritz_change) *ritz_change = ctx->last_ritz_change;
  if (residual) *residual = ctx->last_residual;
  return 0;
}

int hh_density_matrix_layout(const hh_superblock* sb, int side, int32_t* nblk, int64_t* size, int32_t* sector,
                             int32_t* dim, int64_t* off) {
  if (!sb || !nblk || !size || (side != 0 && side != 1)) return 1;
  std::vector<int> rsec, rdim;
  std::vector<int64_t> roff;
  (void)compile_density(sb->plan.sb, side, rsec, roff, rdim);
  *nblk = (int32_t)rsec.size();
  *size = 0;
  for (size_t i = 0; i < rsec.size(); ++i) {
    *size += (int64_t)rdim[i] * rdim[i];
    if (sector) sector[i] = rsec[i];
    if (dim) dim[i] = rdim[i];
    if (off) off[i] = roff[i];
  }
  return 0;
}

int hh_density_matrix(hh_ctx* ctx, hh_superblock* sb, int side, const double* psi, double* rho) {
  return guard(ctx, [&] {
    need(sb && psi && rho, "null argument");
    need(side == 0 || side == 1, "side must be 0 (sys) or 1 (env)");
    std::vector<int> rsec, rdim;
    std::vector<int64_t> roff;
    TaskList tl = compile_density(sb->plan.sb, side, rsec, roff, rdim);
    BufPtrs b{};
    b.p[B_IN] = (double2*)psi;
    b.p[B_OUT] = (double2*)rho;
    launch_on(ctx, tl, b);
  });
}

int hh_density_matrix_levels(hh_ctx* ctx, hh_superblock* sb, int side, const double* psi1, const double* psi2,
                             double* rho) {
  return guard(ctx, [&] {
    need(sb && psi1 && psi2 && rho, "null argument");
    need(side == 0 || side == 1, "side must be 0 (sys) or 1 (env)");
    std::vector<int> rsec, rdim;
    std::vector<int64_t> roff;
    TaskList tl = compile_density(sb->plan.sb, side, rsec, roff, rdim);
    int64_t total = 0;
    for (int d : rdim) total += (int64_t)d * d;
    DevTasks dt;
    dt.upload(tl, ctx->stream);
    launch_tiles(dt, bufs_of((double2*)psi1, (double2*)rho, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr),
                 ctx->stream, PREC_FP64);
    mix_levels(ctx, dt, (double2*)rho, (const double2*)psi2, total);
  });
}

int hh_truncate_levels(hh_ctx* ctx, hh_superblock* sb, int side, const double* psi1, const double* psi2,
                       int kept_min, int kept_max, double trun_err, hh_trunc** out, int* num_state, double* err,
                       double* entropy) {
  return guard(ctx, [&] {
    need(side == 0 || side == 1, "side must be 0 or 1");
    need(kept_min >= 0 && kept_max >= 1, "bad kept range");
    need(sb && psi1 && psi2 && out, "null argument");
    hh_trunc* t = truncate(ctx, sb, side, (const double2*)psi1, kept_min, kept_max, trun_err, (const double2*)psi2);
    *out = t;
    if (num_state) *num_state = t->num_state;
    if (err) *err = t->err;
    if (entropy) *entropy = t->entropy;
  });
}

int hh_truncate(hh_ctx* ctx, hh_superblock* sb, int side, const double* psi, int kept_min, int kept_max,
                double trun_err, hh_trunc** out, int* num_state, double* err, double* entropy) {
  return guard(ctx, [&] {
    need(side == 0 || side == 1, "side must be 0 or 1");
    need(kept_min >= 0 && kept_max >= 1, "bad kept range");
    hh_trunc* t = truncate(ctx, sb, side, (const double2*)psi, kept_min, kept_max, trun_err);
    *out = t;
    if (num_state) *num_state = t->num_state;
    if (err) *err = t->err;
    if (entropy) *entropy = t->entropy;
  });
}
int hh_trunc_identity(hh_ctx* ctx, hh_superblock* sb, int side, hh_trunc** out) {
  return guard(ctx, [&] { *out = trunc_identity(ctx, sb, side); });
}
int hh_trunc_destroy(hh_trunc* t) {
  delete t;
  return 0;
}
int hh_trunc_kept(const hh_trunc* t, int32_t* nsec, int32_t* sector, int32_t* keep) {
  if (!t) return 1;
  if (nsec) *nsec = (int32_t)t->kept.size();
  for (size_t i = 0; i < t->kept.size(); ++i) {
    if (sector) sector[i] = t->kept[i].a;
    if (keep) keep[i] = t->kept[i].keep;
  }
  return 0;
}

int hh_renormalize(hh_ctx* ctx, const hh_trunc* t, int nsub, const int32_t* bra, const int32_t* ket,
                   const int64_t* off, const double* op_dev, int32_t* out_nsub, int64_t* out_size, int32_t* out_bra,
                   int32_t* out_ket, int64_t* out_off, double* out_dev) {
  return guard(ctx, [&] {
    need(t && nsub >= 0 && (nsub == 0 || (bra && ket && off)), "null argument");
    const auto& sec = t->eb.sec;
    std::vector<int> kidx(sec.size(), -1);
    for (size_t j = 0; j < t->kept.size(); ++j) kidx[t->kept[j].a] = (int)j;
    // output sub-blocks: (bra, ket) pairs with both sectors kept, in input order
    TaskList s1, s2;
    int64_t xo = 0, oo = 0;
    std::vector<int32_t> ob, ok;
    std::vector<int64_t> oof;
    for (int i = 0; i < nsub; ++i) {
      need(bra[i] >= 0 && bra[i] < (int)sec.size() && ket[i] >= 0 && ket[i] < (int)sec.size() && off[i] >= 0,
           "sub-block sector out of range");
      const int jb = kidx[bra[i]], jk = kidx[ket[i]];
      if (jb < 0 || jk < 0) continue;
      const KeptSector& kb = t->kept[jb];
      const KeptSector& kk = t->kept[jk];
      const int db = sec[bra[i]].dim, dk = sec[ket[i]].dim;
      // X = O U_ket (db x keep_ket), then out = U_bra^dag X (keep_bra x keep_ket)
      s1.add_tile(B_X, xo, db, db, kk.keep);
      s1.add_term(Term{1.0, 0.0, off[i], kk.u_off, db, kk.u_ld, dk, B_IN, B_AUX, OP_NN, OP_NN});
      s2.add_tile(B_OUT, oo, kb.keep, kb.keep, kk.keep);
      s2.add_term(Term{1.0, 0.0, kb.u_off, xo, kb.u_ld, db, db, B_AUX, B_X, OP_C, OP_NN});
      ob.push_back(jb);
      ok.push_back(jk);
      oof.push_back(oo);
      xo += (int64_t)db * kk.keep;
      oo += (int64_t)kb.keep * kk.keep;
    }
    if (out_nsub) *out_nsub = (int32_t)ob.size();
    if (out_size) *out_size = oo;
    for (size_t i = 0; i < ob.size(); ++i) {
      if (out_bra) out_bra[i] = ob[i];
      if (out_ket) out_ket[i] = ok[i];
      if (out_off) out_off[i] = oof[i];
    }
    if (!out_dev || ob.empty()) return;
    need(op_dev != nullptr, "null operator");
    DevMem X;
    X.alloc(sizeof(double2) * (size_t)std::max<int64_t>(xo, 1));
    DevTasks a, b;
    a.upload(s1, ctx->stream);
    b.upload(s2, ctx->stream);
    BufPtrs bp = bufs_of((double2*)op_dev, (double2*)out_dev, nullptr, nullptr, X.c(), nullptr, t->evec.c(), nullptr);
    launch_tiles(a, bp, ctx->stream, PREC_FP64);
    launch_tiles(b, bp, ctx->stream, PREC_FP64);
    check_hip(hipStreamSynchronize(ctx->stream), "renormalize");
  });
}

int hh_block_conjugate(hh_ctx* ctx, hh_block* b) {
  return guard(ctx, [&] {
    need(b != nullptr, "null block");
    dev_conj(b->arena.c(), b->d.size, ctx->stream);
    b->have32 = false;
    check_hip(hipStreamSynchronize(ctx->stream), "conjugate");
  });
}

int hh_trunc_layout(const hh_trunc* t, int64_t* size, int64_t* u_off, int32_t* u_ld) {
  if (!t) return 1;
  if (size) *size = (int64_t)(t->evec.bytes / sizeof(double2));
  for (size_t i = 0; i < t->kept.size(); ++i) {
    if (u_off) u_off[i] = t->kept[i].u_off;
    if (u_ld) u_ld[i] = t->kept[i].u_ld;
  }
  return 0;
}

int hh_trunc_download(hh_ctx* ctx, const hh_trunc* t, double* dst) {
  return guard(ctx, [&] {
    need(t && dst, "null argument");
    check_hip(hipMemcpyAsync(dst, t->evec.p, t->evec.bytes, hipMemcpyDefault, ctx->stream), "download");
    check_hip(hipStreamSynchronize(ctx->stream), "sync");
  });
}

int hh_grow_block(hh_ctx* ctx, hh_superblock* sb, int side, const hh_trunc* t, hh_block** out) {
  return guard(ctx, [&] { *out = grow(ctx, sb, side, t); });
}

int hh_wave_trans(hh_ctx* ctx, int dir, hh_superblock* old_sb, const double* psi_old, const hh_trunc* sys_t,
                  const hh_trunc* env_t, hh_superblock* new_sb, double* psi_new) {
  return guard(ctx, [&] {
    need(dir == 0 || dir == 1, "dir must be 0 (LR) or 1 (RL)");
    wave_trans(ctx, dir, old_sb, (const double2*)psi_old, sys_t, env_t, new_sb, (double2*)psi_new);
  });
}

int hh_dmrg_run(hh_ctx* ctx, const hh_schedule* s, hh_step_record* rec, int max_rec, int* nrec) {
  return guard(ctx, [&] {
    need(ctx->model_set, "hh_set_model first");
    need(s != nullptr, "null schedule");
    check_schedule(*s);
    need(ctx->model.nsite() >= 4, "lattice too small for two-site DMRG");
    Driver d;
    d.ctx = ctx;
    d.S = *s;
    ctx->precision = s->precision;
    d.run();
    const int n = (int)std::min<size_t>(d.recs.size(), (size_t)std::max(max_rec, 0));
    if (rec) std::memcpy(rec, d.recs.data(), sizeof(hh_step_record) * n);
    if (nrec) *nrec = (int)d.recs.size();
  });
}

}  // extern "C"



extern "C" {

int hh_driver_create(hh_ctx* ctx, const hh_schedule* s, hh_driver** out) {
  return guard(ctx, [&] {
    need(ctx->model_set, "hh_set_model first");
    need(s != nullptr, "null schedule");
    check_schedule(*s);
    auto* h = new hh_driver();
    h->d.ctx = ctx;
    h->d.S = *s;
    *out = h;
  });
}

int hh_driver_destroy(hh_driver* h) {
  delete h;
  return 0;
}

static void take_records(hh_driver* h, hh_step_record* rec, int max_rec, int* nrec) {
  const size_t n = h->d.recs.size() - h->taken;
  const size_t c = std::min<size_t>(n, (size_t)std::max(max_rec, 0));
  if (rec && c) std::memcpy(rec, h->d.recs.data() + h->taken, sizeof(hh_step_record) * c);
  if (nrec) *nrec = (int)n;
  h->last_first = h->taken;
  h->taken = h->d.recs.size();
}

int hh_driver_warmup(hh_driver* h, hh_step_record* rec, int max_rec, int* nrec) {
  return guard(h->d.ctx, [&] {
    h->d.ctx->precision = h->d.S.precision;
    h->d.warm_up();
    take_records(h, rec, max_rec, nrec);
  });
}

int hh_driver_sweep(hh_driver* h, int dir, int bg, int end, double lanzero, double trun_err, int kept_max, int fresh,
                    int phase, hh_step_record* rec, int max_rec, int* nrec) {
  return guard(h->d.ctx, [&] {
    const int n = h->d.ctx->model.nsite();
    need(dir == 0 || dir == 1, "dir must be 0 (LR) or 1 (RL)");
    need(bg >= 1 && bg <= n - 1 && end >= 1 && end <= n - 1, "sweep bounds outside the lattice");
    if (dir == 0) h->d.sweep_lr(bg, end, lanzero, trun_err, kept_max, fresh != 0, phase);
    else h->d.sweep_rl(bg, end, lanzero, trun_err, kept_max, phase);
    take_records(h, rec, max_rec, nrec);
  });
}

int hh_driver_reset_transform(hh_driver* h) {
  h->d.prev_sb.reset();
  return 0;
}

int hh_driver_capture(hh_driver* h, int max_samples) {
  if (!h) return 1;
  return guard(h->d.ctx, [&] {
    need(max_samples >= 0 && max_samples <= 64, "max_samples must be in [0, 64]");
    h->d.capture_left = max_samples;
  });
}

int hh_driver_train_precond(hh_driver* h, int epochs, double lr, double* losses, int32_t* nsamples) {
  if (!h) return 1;
  return guard(h->d.ctx, [&] {
    Driver& d = h->d;
    need(std::isfinite(lr) && lr > 0.0, "learning rate must be positive");
    if (!d.net) {
      d.net.reset(new PcNet());
      d.net->init(d.ctx->stream, d.S.seed);
    }
    std::vector<double> ls;
    pc_train(d.ctx, *d.net, d.samples, epochs, lr, &ls);
    if (losses)
      for (size_t i = 0; i < ls.size(); ++i) losses[i] = ls[i];
    if (nsamples) *nsamples = (int32_t)d.samples.size();
  });
}

int hh_driver_set_precond(hh_driver* h, int mode) {
  if (!h) return 1;
  return guard(h->d.ctx, [&] {
    need(mode == PC_LEARNED || mode == PC_DIAGONAL || mode == PC_NONE, "mode must be 0, 1 or 2");
    h->d.pc_mode = mode;
  });
}

int hh_driver_set_fallback(hh_driver* h, double factor) {
  if (!h) return 1;
  return guard(h->d.ctx, [&] {
    need(std::isfinite(factor) && factor > 0.0, "the fallback factor must be positive");
    h->d.fall_factor = factor;
  });
}

// dst's network becomes a device copy of src's (parameters and Adam state)
static void copy_net(hh_ctx* ctx, Driver& dst, const Driver& src) {
  dst.net.reset(new PcNet());
  for (DevMem* m : {&dst.net->w, &dst.net->g, &dst.net->m, &dst.net->v}) m->alloc(sizeof(float) * MlpLayout::SIZE);
  dst.net->loss.alloc(sizeof(double) * 4);
  for (int k = 0; k < 3; ++k) {
    const DevMem& f = k == 0 ? src.net->w : k == 1 ? src.net->m : src.net->v;
    DevMem& t = k == 0 ? dst.net->w : k == 1 ? dst.net->m : dst.net->v;
    check_hip(hipMemcpyAsync(t.p, f.p, f.bytes, hipMemcpyDeviceToDevice, ctx->stream), "d2d");
  }
  dst.net->adam_t = src.net->adam_t;
  dst.net->trained = src.net->trained;
}

int hh_driver_set_min_dim(hh_driver* h, int64_t min_dim) {
  if (!h) return 1;
  return guard(h->d.ctx, [&] {
    need(min_dim >= 0, "min_dim must be >= 0");
    h->d.min_dim = min_dim;
  });
}

int hh_driver_set_subspace(hh_driver* h, int kmax) {
  if (!h) return 1;
  return guard(h->d.ctx, [&] {
    need(kmax >= 2 && kmax <= DAV_KMAX_LIMIT, "the Davidson subspace size must be in [2, 128]");
    h->d.dav_kmax = kmax;
  });
}

int hh_driver_set_levels(hh_driver* h, int levels) {
  if (!h) return 1;
  return guard(h->d.ctx, [&] {
    need(levels >= 1 && levels <= MAX_LEVELS, "levels must be in [1, 8]");
    need(2 * (levels - 1) < h->d.S.max_iter, "levels must stay below max_iter / 2");
    need(levels == 1 || h->d.S.solver == 0, "several levels are solved by the plain Lanczos only (solver 0)");
    h->d.levels = levels;
  });
}

int hh_driver_set_level_start(hh_driver* h, int carry) {
  if (!h) return 1;
  return guard(h->d.ctx, [&] {
    need(carry == 0 || carry == 1, "carry must be 0 (seeded, the reference's) or 1 (wave-transformed)");
    h->d.level_carry = carry != 0;
  });
}

int hh_driver_levels(const hh_driver* h, int32_t* levels) {
  if (!h || !levels) return 1;
  *levels = h->d.levels;
  return 0;
}

int hh_driver_level_energies(const hh_driver* h, int max_rec, double* e, int* nrec) {
  if (!h) return 1;
  return guard(h->d.ctx, [&] {
    const Driver& d = h->d;
    const size_t first = h->last_first, n = h->taken - first;
    const size_t c = std::min<size_t>(n, (size_t)std::max(max_rec, 0));
    need(c == 0 || e != nullptr, "null argument");
    for (size_t i = 0; i < c; ++i) {
      const std::vector<double>& v = d.lev_e[first + i];
      for (int l = 0; l < d.levels; ++l) e[i * d.levels + l] = l < (int)v.size() ? v[l] : std::nan("");
    }
    if (nrec) *nrec = (int)n;
  });
}

int hh_driver_copy_precond(hh_driver* dst, const hh_driver* src) {
  if (!dst || !src) return 1;
  return guard(dst->d.ctx, [&] {
    need(dst->d.ctx == src->d.ctx, "both drivers must belong to one context");
    need(src->d.net != nullptr, "the source driver has no network (train it first)");
    if (dst == src) return;
    copy_net(dst->d.ctx, dst->d, src->d);
    check_hip(hipStreamSynchronize(dst->d.ctx->stream), "copy_precond");
  });
}

namespace {
// a copy of the driver holding the blocks and truncations whose length lies in
// [lo[side], hi[side]] on each side (all of them with the full windows)
int fork_driver(const hh_driver* src, const int lo[2], const int hi[2], hh_driver** out) {
  if (!src || !out) return 1;
  hh_ctx* ctx = src->d.ctx;
  return guard(ctx, [&] {
    for (int side = 0; side < 2; ++side) need(lo[side] <= hi[side], "empty length window");
    if (src->d.prev_sb) {
      const int ls = src->d.prev_sb->sys->d.length, le = src->d.prev_sb->env->d.length;
      need(ls >= lo[0] && ls <= hi[0] && le >= lo[1] && le <= hi[1],
           "the length window leaves out a block of the driver's current superblock");
    }
    std::unique_ptr<hh_driver> h(new hh_driver());
    Driver& d = h->d;
    const Driver& s = src->d;
    d.ctx = ctx;
    d.S = s.S;
    d.step = s.step;
    d.psi_dim = s.psi_dim;
    auto copy_dev = [&](DevMem& dst, const DevMem& from) {
      if (!from.p) return;
      dst.alloc(from.bytes);
      check_hip(hipMemcpyAsync(dst.p, from.p, from.bytes, hipMemcpyDeviceToDevice, ctx->stream), "d2d");
    };
    for (int side = 0; side < 2; ++side) {
      const auto& sb = side == 0 ? s.sysb : s.envb;
      auto& db = side == 0 ? d.sysb : d.envb;
      for (const auto& kv : sb) {
        if (kv.first < lo[side] || kv.first > hi[side]) continue;
        std::unique_ptr<hh_block> b(new hh_block());
        b->d = kv.second->d;
        copy_dev(b->arena, kv.second->arena);
        db[kv.first] = std::move(b);
      }
      const auto& st = side == 0 ? s.syst : s.envt;
      auto& dt = side == 0 ? d.syst : d.envt;
      for (const auto& kv : st) {
        if (kv.first < lo[side] || kv.first > hi[side]) continue;
        std::unique_ptr<hh_trunc> t(new hh_trunc());
        const hh_trunc& o = *kv.second;
        t->side = o.side;
        t->eb = o.eb;
        t->kept = o.kept;
        t->kept_per_rho = o.kept_per_rho;
        t->num_state = o.num_state;
        t->err = o.err;
        t->entropy = o.entropy;
        copy_dev(t->evec, o.evec);
        dt[kv.first] = std::move(t);
      }
    }
    copy_dev(d.psi, s.psi);
    d.pc_mode = s.pc_mode;
    d.fall_factor = s.fall_factor;
    d.min_dim = s.min_dim;
    d.dav_kmax = s.dav_kmax;
    d.levels = s.levels;
    d.level_carry = s.level_carry;
    for (const auto& v : s.lpsi) {
      d.lpsi.emplace_back(new DevMem());
      copy_dev(*d.lpsi.back(), *v);
    }
    d.lpsi_dim = s.lpsi_dim;
    d.solved_levels = s.solved_levels;
    if (s.net) copy_net(ctx, d, s);
    if (s.prev_sb) {
      // the previous step's superblock, bound again to the copied blocks
      d.prev_sb.reset(bind(ctx, d.sysb.at(s.prev_sb->sys->d.length).get(), d.envb.at(s.prev_sb->env->d.length).get(),
                           s.prev_sb->N));
    }
    check_hip(hipStreamSynchronize(ctx->stream), "fork");
    *out = h.release();
  });
}
}  // namespace

int hh_driver_fork(const hh_driver* src, hh_driver** out) {
  const int lo[2] = {INT_MIN, INT_MIN}, hi[2] = {INT_MAX, INT_MAX};
  return fork_driver(src, lo, hi, out);
}

int hh_driver_fork_window(const hh_driver* src, int sys_lo, int sys_hi, int env_lo, int env_hi, hh_driver** out) {
  const int lo[2] = {sys_lo, env_lo}, hi[2] = {sys_hi, env_hi};
  return fork_driver(src, lo, hi, out);
}

int hh_driver_set_solver(hh_driver* h, int solver) {
  if (!h) return 1;
  return guard(h->d.ctx, [&] {
    need(solver >= 0 && solver <= 2, "solver must be 0, 1 or 2");
    need(solver == 0 || h->d.levels == 1, "several levels are solved by the plain Lanczos only (solver 0)");
    h->d.S.solver = solver;
  });
}

int hh_driver_wave(const hh_driver* h, int32_t* sys_len, int32_t* env_len, int64_t* dim, double* psi) {
  if (!h) return 1;
  return guard(h->d.ctx, [&] {
    const hh_superblock* sb = h->d.prev_sb.get();
    need(sb != nullptr && h->d.psi_dim == sb->plan.sb.dim, "no sweep step since the driver was created or reset");
    if (sys_len) *sys_len = sb->sys->d.length;
    if (env_len) *env_len = sb->env->d.length;
    if (dim) *dim = sb->plan.sb.dim;
    if (psi) {
      hh_ctx* ctx = h->d.ctx;
      check_hip(hipMemcpyAsync(psi, h->d.psi.p, sizeof(double2) * (size_t)sb->plan.sb.dim, hipMemcpyDeviceToDevice,
                               ctx->stream), "d2d");
      check_hip(hipStreamSynchronize(ctx->stream), "sync");
    }
  });
}

int hh_driver_block_count(const hh_driver* h, int side, int32_t* count) {
  if (!h || !count) return 1;
  *count = (int32_t)(side == 0 ? h->d.sysb.size() : h->d.envb.size());
  return 0;
}

int hh_driver_block(const hh_driver* h, int side, int length, const hh_block** out) {
  if (!h || !out) return 1;
  const auto& m = side == 0 ? h->d.sysb : h->d.envb;
  auto it = m.find(length);
  if (it == m.end()) return 1;
  *out = it->second.get();
  return 0;
}

int hh_driver_trunc(const hh_driver* h, int side, int length, const hh_trunc** out) {
  if (!h || !out) return 1;
  const auto& m = side == 0 ? h->d.syst : h->d.envt;
  auto it = m.find(length);
  if (it == m.end()) return 1;
  *out = it->second.get();
  return 0;
}

int hh_driver_save(hh_driver* h, const char* path) {
  if (!h) return 1;
  return guard(h->d.ctx, [&] {
    need(path != nullptr, "null path");
    save_driver(h->d, path);
  });
}

int hh_driver_load(hh_driver* h, const char* path) {
  if (!h) return 1;
  return guard(h->d.ctx, [&] {
    need(path != nullptr, "null path");
    load_driver(h->d, path);
  });
}

int hh_eigh_sectors(hh_ctx* ctx, int nsec, const int32_t* dim, const int32_t* keep, double* blocks_dev,
                    double* evals_dev, int solver) {
  return guard(ctx, [&] {
    need(nsec >= 1 && dim && keep && blocks_dev && evals_dev, "null or empty argument");
    need(solver == 0 || solver == 1, "solver must be 0 (hand-written) or 1 (zheevd)");
    std::vector<int64_t> off(nsec);
    std::vector<int> d(nsec), kp(nsec);
    int64_t o = 0;
    int dmax = 1;
    for (int i = 0; i < nsec; ++i) {
      need(dim[i] >= 1 && keep[i] >= 0 && keep[i] <= dim[i], "bad block dimension or kept count");
      off[i] = o;
      d[i] = dim[i];
      kp[i] = keep[i];
      o += (int64_t)dim[i] * dim[i];
      dmax = std::max(dmax, dim[i]);
    }
    double2* A = (double2*)blocks_dev;
    if (solver == 0) {
      if (dmax > HEEV_MAXN) throw HHError(3, "a block is larger than the hand-written eigensolver takes");
      ctx->heev.values(ctx->stream, A, off, d, evals_dev);
      if (!ctx->heev.vectors(ctx->stream, A, kp, evals_dev))
        throw HHError(3, "the hand-written eigensolver handed the problem back: " + ctx->heev.fail_info);
    } else {
      DevMem work, info;
      work.alloc(sizeof(double) * (size_t)dmax * (hh_ctx::NSIDE + 1));
      info.alloc(sizeof(rocblas_int) * nsec + 16);
      heev_sectors(ctx, A, off, d, evals_dev, (double*)work.p, dmax, (rocblas_int*)info.p, []() {});
      std::vector<rocblas_int> inf(nsec);
      check_hip(hipMemcpyAsync(inf.data(), info.p, sizeof(rocblas_int) * nsec, hipMemcpyDeviceToHost, ctx->stream),
                "d2h");
      check_hip(hipStreamSynchronize(ctx->stream), "eig");
      for (auto x : inf)
        if (x != 0) throw HHError(-2, "rocsolver_zheevd did not converge");
    }
    check_hip(hipStreamSynchronize(ctx->stream), "eig");
  });
}

int hh_set_eigensolver(hh_ctx* ctx, int solver) {
  return guard(ctx, [&] {
    need(solver == 0 || solver == 1, "solver must be 0 (hand-written) or 1 (zheevd)");
    ctx->eig_solver = solver;
  });
}

int hh_eigensolver_status(const hh_ctx* ctx, int32_t* solver, int64_t* handbacks, int32_t* ortho_steps,
                          double* ortho_error) {
  if (!ctx) return 1;
  if (solver) *solver = ctx->eig_solver;
  if (handbacks) *handbacks = ctx->eig_handbacks;
  if (ortho_steps) *ortho_steps = ctx->heev.ns_iterations;
  if (ortho_error) *ortho_error = ctx->heev.ortho_error;
  return 0;
}

int hh_set_precision(hh_ctx* ctx, int precision) {
  return guard(ctx, [&] {
    need(precision == PREC_FP64 || precision == PREC_FP32 || precision == PREC_FP32_STORE,
         "precision must be 0 (FP64), 1 (FP32 MFMA) or 2 (FP32 storage)");
    ctx->precision = precision;
  });
}

int hh_set_profiling(hh_ctx* ctx, int profile) {
  ctx->profile = profile != 0;
  return 0;
}
int hh_get_stats(hh_ctx* ctx, hh_stats* out) {
  if (!ctx || !out) return 1;
  *out = ctx->stats;
  return 0;
}
int hh_reset_stats(hh_ctx* ctx) {
  if (!ctx) return 1;
  ctx->stats = hh_stats{};
  return 0;
}

int hh_superblock_tasks(const hh_superblock* sb, int stage, const void** tiles, int32_t* ntiles, const void** terms,
                        int32_t* nterms, int64_t* scratch) {
  if (!sb) return 1;
  const TaskList* tl = stage == 0 ? &sb->plan.merge_tasks : stage == 1 ? &sb->plan.stage1 : &sb->plan.stage2;
  if (tiles) *tiles = tl->tiles.data();
  if (ntiles) *ntiles = (int32_t)tl->tiles.size();
  if (terms) *terms = tl->terms.data();
  if (nterms) *nterms = (int32_t)tl->terms.size();
  if (scratch) *scratch = stage == 0 ? sb->plan.merged_size : sb->plan.x_size;
  return 0;
}

}  // extern "C"

namespace {
// A caller's task list checked against its buffer slots (hh_run_tasks).
TaskList checked_tasks(const void* tiles, int32_t ntiles, const void* terms, int32_t nterms, void* const* bufs_dev,
                       int precision) {
  need(ntiles >= 0 && nterms >= 0 && bufs_dev, "bad task list");
  need(ntiles == 0 || tiles, "null tile array");
  need(nterms == 0 || terms, "null term array");
  need(precision == PREC_FP64 || precision == PREC_FP32, "precision must be 0 or 1");
  TaskList tl;
  tl.tiles.assign((const Tile*)tiles, (const Tile*)tiles + ntiles);
  tl.terms.assign((const Term*)terms, (const Term*)terms + nterms);
  for (const Tile& t : tl.tiles) {
    need(t.t0 >= 0 && t.t1 <= nterms && t.t0 <= t.t1 && t.M >= 0 && t.N >= 0, "tile out of range");
    need(t.c_buf >= 0 && t.c_buf < 8 && bufs_dev[t.c_buf], "tile buffer slot");
    for (int k = t.t0; k < t.t1; ++k) {
      const Term& m = tl.terms[k];
      need(m.a_buf < 8 && bufs_dev[m.a_buf] && m.a_op <= 3 && m.b_op <= 3 && m.K >= 0, "term operand");
      if (m.K > 0) need(m.b_buf < 8 && bufs_dev[m.b_buf], "term operand");
    }
  }
  return tl;
}
}  // namespace

extern "C" {

int hh_run_tasks(hh_ctx* ctx, const void* tiles, int32_t ntiles, const void* terms, int32_t nterms,
                 void* const* bufs_dev, int precision) {
  return guard(ctx, [&] {
    const TaskList tl = checked_tasks(tiles, ntiles, terms, nterms, bufs_dev, precision);
    DevTasks dt;
    dt.upload(tl, ctx->stream, nullptr, 0, ctx->forced_cfg ? ctx->forced : LaunchConfig());
    BufPtrs b;
    for (int i = 0; i < 8; ++i) b.p[i] = (double2*)bufs_dev[i];
    launch_tiles(dt, b, ctx->stream, precision);
    check_hip(hipStreamSynchronize(ctx->stream), "hh_run_tasks");
  });
}

// hh_run_tasks with the task tables, work items and split-sum partials in
// the fixed device scratch (its address on this context's device was taken
// at hh_ctx_create): no device memory is allocated or freed.  Every launch
// rewrites the parts of the scratch it reads, so nothing needs clearing
// between calls; calls that overlap on one GPU share the array and race.
int hh_run_tasks_noalloc(hh_ctx* ctx, const void* tiles, int32_t ntiles, const void* terms, int32_t nterms,
                         void* const* bufs_dev, int precision) {
  return guard(ctx, [&] {
    const TaskList tl = checked_tasks(tiles, ntiles, terms, nterms, bufs_dev, precision);
    void* scr = noalloc_scratch(ctx->device);
    need(scr != nullptr, "no fixed scratch on this device");
    DevTasks dt;
    try {
      dt.upload(tl, ctx->stream, scr, NOALLOC_SCRATCH_BYTES, ctx->forced_cfg ? ctx->forced : LaunchConfig());
    } catch (const HHCapacity& e) {
      throw HHError(1, e.what());
    }
    BufPtrs b;
    for (int i = 0; i < 8; ++i) b.p[i] = (double2*)bufs_dev[i];
    launch_tiles(dt, b, ctx->stream, precision);
    check_hip(hipStreamSynchronize(ctx->stream), "hh_run_tasks_noalloc");
  });
}

int hh_precond_eval(hh_ctx* ctx, int np, const int32_t* rows, const int32_t* cols, const int32_t* s2,
                    const int32_t* nq, int n_mid, const double* hs, const double* he, const double* r,
                    const double* delta, double theta, const float* w, float* feat, float* out, double* par,
                    double* loss, float* grad) {
  return guard(ctx, [&] {
    need(np >= 1 && rows && cols && s2 && nq && hs && he && r && delta && w, "null or empty argument");
    PairTable t;
    t.np = np;
    int64_t n = 0;
    int32_t so = 0, eo = 0;
    for (int p = 0; p < np; ++p) {
      need(rows[p] >= 1 && cols[p] >= 1, "empty sector");
      t.off.push_back(n);
      t.rows.push_back(rows[p]);
      t.cols.push_back(cols[p]);
      t.soff.push_back(so);
      t.eoff.push_back(eo);
      t.s2.push_back(s2[p]);
      t.nq.push_back(nq[p]);
      n += (int64_t)rows[p] * cols[p];
      so += rows[p];
      eo += cols[p];
    }
    t.n = n;
    std::vector<double> vhs(hs, hs + so), vhe(he, he + eo);
    PcDev P;
    pc_setup(P, t, vhs, vhe, n_mid, ctx->stream);
    DevMem dr, dd, dw, dg, dl;
    dr.alloc(sizeof(double2) * (size_t)n);
    dd.alloc(sizeof(double2) * (size_t)n);
    dw.alloc(sizeof(float) * MlpLayout::SIZE);
    dg.alloc(sizeof(float) * MlpLayout::SIZE);
    dl.alloc(sizeof(double) * 4);
    double dn = 0.0;
    for (int64_t i = 0; i < 2 * n; ++i) dn += delta[i] * delta[i];
    need(dn > 0.0, "zero target");
    check_hip(hipMemcpyAsync(dr.p, r, sizeof(double2) * n, hipMemcpyHostToDevice, ctx->stream), "h2d");
    check_hip(hipMemcpyAsync(dd.p, delta, sizeof(double2) * n, hipMemcpyHostToDevice, ctx->stream), "h2d");
    check_hip(hipMemcpyAsync(dw.p, w, sizeof(float) * MlpLayout::SIZE, hipMemcpyHostToDevice, ctx->stream), "h2d");
    pc_features(P, dr.c(), theta, n_mid, ctx->stream);
    pc_forward(P, (const float*)dw.p, ctx->stream);
    pc_backward(P, dr.c(), dd.c(), theta, dn, (const float*)dw.p, (float*)dg.p, (double*)dl.p, ctx->stream);
    auto down = [&](void* dst, const void* src, size_t b) {
      if (dst) check_hip(hipMemcpyAsync(dst, src, b, hipMemcpyDeviceToHost, ctx->stream), "d2h");
    };
    down(feat, P.feat, sizeof(float) * (size_t)np * PC_NF);
    down(out, P.out, sizeof(float) * (size_t)np * PC_NO);
    down(par, P.par, sizeof(double) * 2 * (size_t)np);
    down(loss, dl.p, sizeof(double));
    down(grad, dg.p, sizeof(float) * MlpLayout::SIZE);
    check_hip(hipStreamSynchronize(ctx->stream), "precond eval");
  });
}

int hh_precond_size(void) { return MlpLayout::SIZE; }

int hh_start_vector(hh_ctx* ctx, double* psi, int64_t n, uint64_t seed) {
  return guard(ctx, [&] {
    dev_start_vector((double2*)psi, n, seed, ctx->stream);
    check_hip(hipStreamSynchronize(ctx->stream), "start");
    // normalise (the host-side stream is normalised the same way)
    ensure_ws(ctx, 1, 1);
    const double n2 = rdot(ctx, (double2*)psi, (double2*)psi, n);
    dev_scale((double2*)psi, 1.0 / std::sqrt(n2), n, ctx->stream);
    check_hip(hipStreamSynchronize(ctx->stream), "start");
  });
}

}  // extern "C"
