// Davidson ground-state solver with the learned per-sector preconditioner
// (the chosen form of the DMRG inner loop) and the training of that
// preconditioner.  The reference solves every step with lanczos_memo_new
// (code/lanczos.f90:4-176); this solver replaces it with
//   * subspace expansion vectors t = P(r) from the preconditioner of
//     hh_precond.hpp (learned, fixed diagonal or none),
//   * H.psi of the expansion vectors in the FP32-storage form,
//   * restarts whose first vector is the Ritz vector with an FP64 H.psi
//     (an exact Rayleigh quotient and residual: iterative refinement of the
//     FP32 products), and
//   * the reference's stopping rule applied to successive FP64 Rayleigh
//     quotients: |E_k - E_{k-1}| <= lanzero.
// Inner products are real parts, as the reference's wave_product
// (code/space_info.f90:1018-1034), so the subspace matrix is real symmetric.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <limits>
#include <vector>

#include "hh_solver.hpp"

namespace hh {

namespace {

// Lowest eigenpair of a real symmetric k x k matrix (row-major), cyclic
// Jacobi rotations.
void lowest_eig(const std::vector<double>& A0, int k, double* lam, std::vector<double>& vec) {
  std::vector<double> A(A0), V((size_t)k * k, 0.0);
  for (int i = 0; i < k; ++i) V[(size_t)i * k + i] = 1.0;
  for (int sweep = 0; sweep < 60; ++sweep) {
    double off = 0.0, diag = 0.0;
    for (int i = 0; i < k; ++i) {
      diag += A[(size_t)i * k + i] * A[(size_t)i * k + i];
      for (int j = i + 1; j < k; ++j) off += A[(size_t)i * k + j] * A[(size_t)i * k + j];
    }
    if (off <= 1e-30 * std::max(diag, 1e-300)) break;
    for (int p = 0; p < k; ++p)
      for (int q = p + 1; q < k; ++q) {
        const double apq = A[(size_t)p * k + q];
        if (std::fabs(apq) < 1e-300) continue;
        const double app = A[(size_t)p * k + p], aqq = A[(size_t)q * k + q];
        const double th = (aqq - app) / (2.0 * apq);
        const double t = (th >= 0 ? 1.0 : -1.0) / (std::fabs(th) + std::sqrt(th * th + 1.0));
        const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
        for (int r = 0; r < k; ++r) {
          const double arp = A[(size_t)r * k + p], arq = A[(size_t)r * k + q];
          A[(size_t)r * k + p] = c * arp - s * arq;
          A[(size_t)r * k + q] = s * arp + c * arq;
        }
        for (int r = 0; r < k; ++r) {
          const double apr = A[(size_t)p * k + r], aqr = A[(size_t)q * k + r];
          A[(size_t)p * k + r] = c * apr - s * aqr;
          A[(size_t)q * k + r] = s * apr + c * aqr;
        }
        for (int r = 0; r < k; ++r) {
          const double vrp = V[(size_t)r * k + p], vrq = V[(size_t)r * k + q];
          V[(size_t)r * k + p] = c * vrp - s * vrq;
          V[(size_t)r * k + q] = s * vrp + c * vrq;
        }
      }
  }
  int m = 0;
  for (int i = 1; i < k; ++i)
    if (A[(size_t)i * k + i] < A[(size_t)m * k + m]) m = i;
  *lam = A[(size_t)m * k + m];
  vec.assign(k, 0.0);
  for (int r = 0; r < k; ++r) vec[r] = V[(size_t)r * k + m];
}

// diagonal estimate of an enlarged block's Hamiltonian in its enlarged basis
// (sector order, subsector spos + block state): the block Hamiltonian's own
// diagonal energies H_bb(a, a) / sqrt(2S_b + 1) (scalar operators are stored
// times sqrt(2S + 1)) plus the site's on-site U; bond terms inside the
// enlarged block are left out (the hopping has no diagonal part)
std::vector<double> enlarged_diag(hh_ctx* ctx, const hh_block* blk, const EnlBasis& eb) {
  const BlockDesc& b = blk->d;
  std::vector<int64_t> idx;
  std::vector<int64_t> first(b.sec.size(), 0);
  for (size_t q = 0; q < b.sec.size(); ++q) {
    first[q] = (int64_t)idx.size();
    for (int a = 0; a < b.sec[q].dim; ++a) idx.push_back(b.h_off[q] + (int64_t)a * (b.sec[q].dim + 1));
  }
  std::vector<double> hv(idx.size(), 0.0);
  if (!idx.empty()) {
    DevMem di, dv;
    di.alloc(sizeof(int64_t) * idx.size());
    dv.alloc(sizeof(double) * idx.size());
    check_hip(hipMemcpyAsync(di.p, idx.data(), sizeof(int64_t) * idx.size(), hipMemcpyHostToDevice, ctx->stream),
              "h2d");
    dev_gather_re((double*)dv.p, blk->arena.c(), (const int64_t*)di.p, (int)idx.size(), ctx->stream);
    check_hip(hipMemcpyAsync(hv.data(), dv.p, sizeof(double) * idx.size(), hipMemcpyDeviceToHost, ctx->stream), "d2h");
    check_hip(hipStreamSynchronize(ctx->stream), "diag");
  }
  std::vector<double> out((size_t)eb.total_dim(), 0.0);
  int64_t off = 0;
  for (size_t x = 0; x < eb.sec.size(); ++x) {
    for (const SubSector& q : eb.subs[x]) {
      const Sector& bs = b.sec[q.b];
      const Sector& st = eb.st[q.s];
      const double site = (ctx->model.kind == HUBBARD && st.N == 2) ? ctx->model.U : 0.0;
      const double scale = 1.0 / std::sqrt((double)bs.S2 + 1.0);
      // site multiplets have dimension 1: the subsector holds the block states
      for (int a = 0; a < q.dim; ++a) out[off + q.pos + a] = hv[first[q.b] + (a % bs.dim)] * scale + site;
    }
    off += eb.sec[x].dim;
  }
  return out;
}

PairTable pair_table(const hh_superblock* sb) {
  const SuperBasis& S = sb->plan.sb;
  PairTable t;
  t.np = (int)S.pairs.size();
  t.n = S.dim;
  std::vector<int32_t> so(sb->plan.sbs.sec.size() + 1, 0), eo(sb->plan.ebs.sec.size() + 1, 0);
  for (size_t x = 0; x < sb->plan.sbs.sec.size(); ++x) so[x + 1] = so[x] + sb->plan.sbs.sec[x].dim;
  for (size_t y = 0; y < sb->plan.ebs.sec.size(); ++y) eo[y + 1] = eo[y] + sb->plan.ebs.sec[y].dim;
  for (int p = 0; p < t.np; ++p) {
    const int x = S.pairs[p].first, y = S.pairs[p].second;
    t.off.push_back(S.off[p]);
    t.rows.push_back(S.rows[p]);
    t.cols.push_back(S.cols[p]);
    t.soff.push_back(so[x]);
    t.eoff.push_back(eo[y]);
    t.s2.push_back(sb->plan.sbs.sec[x].S2);
    t.nq.push_back(sb->plan.sbs.sec[x].N);
  }
  return t;
}

uint64_t mix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

}  // namespace

void PcNet::init(hipStream_t s, uint64_t seed) {
  const int n = MlpLayout::SIZE;
  std::vector<float> h(n, 0.f);
  auto uni = [&](uint64_t i) { return (double)(mix64(seed * 0x100000001B3ull + i) >> 11) * (1.0 / 9007199254740992.0); };
  // Glorot-uniform hidden layers; a small output layer and zero output bias,
  // so the untrained network starts at the fixed diagonal preconditioner
  // (o = 0: sigma = PC_SIGMA0 * scale, alpha = 1)
  const double a1 = std::sqrt(6.0 / (PC_NF + PC_H)), a2 = std::sqrt(6.0 / (PC_H + PC_H)), a3 = 0.01;
  for (int i = 0; i < PC_H * PC_NF; ++i) h[MlpLayout::W1 + i] = (float)(a1 * (2.0 * uni(i) - 1.0));
  for (int i = 0; i < PC_H * PC_H; ++i) h[MlpLayout::W2 + i] = (float)(a2 * (2.0 * uni(100000 + i) - 1.0));
  for (int i = 0; i < PC_NO * PC_H; ++i) h[MlpLayout::W3 + i] = (float)(a3 * (2.0 * uni(200000 + i) - 1.0));
  h[MlpLayout::B3] = 0.f;
  h[MlpLayout::B3 + 1] = 0.f;
  for (DevMem* b : {&w, &g, &m, &v}) b->alloc(sizeof(float) * n);
  loss.alloc(sizeof(double) * 4);
  check_hip(hipMemcpyAsync(w.p, h.data(), sizeof(float) * n, hipMemcpyHostToDevice, s), "h2d net");
  std::vector<float> z(n, 0.f);
  check_hip(hipMemcpyAsync(m.p, z.data(), sizeof(float) * n, hipMemcpyHostToDevice, s), "h2d");
  check_hip(hipMemcpyAsync(v.p, z.data(), sizeof(float) * n, hipMemcpyHostToDevice, s), "h2d");
  check_hip(hipStreamSynchronize(s), "net init");
  adam_t = 0;
  trained = false;
}

void precond_setup(hh_ctx* ctx, const hh_superblock* sb, PcDev& P) {
  const PairTable t = pair_table(sb);
  const std::vector<double> hs = enlarged_diag(ctx, sb->sys, sb->plan.sbs);
  const std::vector<double> he = enlarged_diag(ctx, sb->env, sb->plan.ebs);
  pc_setup(P, t, hs, he, ctx->model.n_elec / 2, ctx->stream);
}

SolveOut davidson(hh_ctx* ctx, hh_superblock* sb, double2* psi, double lanzero, int max_hpsi, int mode,
                  const float* net, bool* fell_back, const AnchorHook& on_anchor, double fall_factor,
                  int kmax_limit) {
  const int64_t n = sb->plan.sb.dim;
  need(n > 0, "empty superblock");
  need(max_hpsi >= 2, "max_iter must be >= 2");
  need(mode != PC_LEARNED || net != nullptr, "the learned preconditioner has no trained network");
  if (fell_back) *fell_back = false;
  need(kmax_limit >= 2 && kmax_limit <= DAV_KMAX_LIMIT, "Davidson subspace size out of range");
  const int kmax = std::min<int>(kmax_limit, std::max<int>(2, max_hpsi));
  ensure_ws(ctx, n, 2 * kmax + 1);   // kmax basis vectors, kmax products, residual, expansion
  double2* base = ctx->V.c();
  auto Vs = [&](int j) { return base + (size_t)j * n; };
  auto Ws = [&](int j) { return base + (size_t)(kmax + j) * n; };
  double2* r = Vs(2 * kmax);
  double2* t = Vs(2 * kmax + 1);
  double* dots = (double*)ctx->dots.p;
  double* part = (double*)ctx->partial.p;
  DevMem coefm;   // subspace coefficients on the device: basis part, product part
  coefm.alloc(sizeof(double) * (2 * kmax + 4));
  double* cv = (double*)coefm.p;
  double* cw = cv + kmax + 1;
  PcDev P;
  if (mode != PC_NONE) precond_setup(ctx, sb, P);
  const int n_mid = ctx->model.n_elec / 2;
  SolveOut out;
  double mv = 0.0;
  auto hpsi = [&](const double2* x, double2* y, int prec) {
    const double t0 = wall();
    const int saved = ctx->precision;
    ctx->precision = prec;
    try {
      matvec(ctx, sb, x, y);
    } catch (...) {
      ctx->precision = saved;
      throw;
    }
    ctx->precision = saved;
    check_hip(hipStreamSynchronize(ctx->stream), "matvec");
    mv += wall() - t0;
    (prec == PREC_FP64 ? out.hpsi_fp64 : out.hpsi_fp32) += 1;
  };
  auto fetch = [&](const double* src, int count) {
    std::vector<double> h(count);
    check_hip(hipMemcpyAsync(h.data(), src, sizeof(double) * count, hipMemcpyDeviceToHost, ctx->stream), "d2h");
    check_hip(hipStreamSynchronize(ctx->stream), "sync");
    return h;
  };
  auto fetch2 = [&](const double* a0, const double* a1) {
    double h[2];
    check_hip(hipMemcpyAsync(&h[0], a0, sizeof(double), hipMemcpyDeviceToHost, ctx->stream), "d2h");
    check_hip(hipMemcpyAsync(&h[1], a1, sizeof(double), hipMemcpyDeviceToHost, ctx->stream), "d2h");
    check_hip(hipStreamSynchronize(ctx->stream), "sync");
    return std::make_pair(h[0], h[1]);
  };
  auto put = [&](double* dst, const std::vector<double>& h) {
    check_hip(hipMemcpyAsync(dst, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice, ctx->stream), "h2d");
  };
  auto finite = [](double x) {
    if (!std::isfinite(x)) throw HHError(-4, "non-finite value in Davidson: H.psi produced NaN or Inf");
    return x;
  };
  // the FP64 anchor: u (in V_0) -> W_0 = H u exactly, T = [theta], r exact
  std::vector<double> T((size_t)kmax * kmax, 0.0), c{1.0};
  int k = 1;
  double theta = 0.0, rho = 0.0;
  auto anchor = [&]() {
    hpsi(Vs(0), Ws(0), PREC_FP64);
    dev_mdots(Vs(0), 0, 1, Ws(0), n, dots, part, ctx->stream);
    theta = finite(fetch(dots, 1)[0]);
    std::fill(T.begin(), T.end(), 0.0);
    T[0] = theta;
    k = 1;
    c.assign(1, 1.0);
    // r = W_0 - theta V_0
    dev_scale_copy(r, Ws(0), 1.0, n, ctx->stream);
    dev_lanczos_update_norm(r, Vs(0), theta, nullptr, 0.0, n, dots, part, ctx->stream);
    rho = std::sqrt(std::max(fetch(dots, 1)[0], 0.0));
    if (on_anchor) on_anchor(Vs(0), r, theta);
  };
  // start vector
  dev_mdots(psi, 0, 1, psi, n, dots, part, ctx->stream);
  const double nrm0 = std::sqrt(fetch(dots, 1)[0]);
  need(std::isfinite(nrm0) && nrm0 > 0.0, "zero or non-finite start vector");
  dev_scale_copy(Vs(0), psi, 1.0 / nrm0, n, ctx->stream);
  anchor();
  double anchor_e = theta, anchor_rho = rho;
  bool r_ready = true;   // r holds the residual of the current Ritz vector
  int used = 1;
  while (true) {
    if (!r_ready) {
      // r = sum_j c_j (W_j - theta V_j), one pass over both sets
      std::vector<double> a(k), w(c);
      for (int j = 0; j < k; ++j) a[j] = -theta * c[j];
      put(cv, a);
      put(cw, w);
      dev_combine2_norm(r, Vs(0), Ws(0), n, k, cv, cw, n, dots, part, ctx->stream);
      r_ready = true;
    }
    // expansion vector: the preconditioned residual, written to V_k
    double2* tk = Vs(k);
    if (mode == PC_NONE) {
      dev_scale_copy(tk, r, 1.0, n, ctx->stream);
    } else {
      pc_features(P, r, theta, n_mid, ctx->stream);
      if (mode == PC_LEARNED) pc_forward(P, net, ctx->stream);
      else pc_fixed(P, PC_SIGMA0, ctx->stream);
      pc_apply(P, r, theta, tk, ctx->stream);
    }
    // classical Gram-Schmidt with real coefficients against the basis
    // (Gram_Schmidt_process, code/space_info.f90:1176-1214); one multi-dot
    // gives the coefficients and |t|^2 (V_k = t), and a second pass runs only
    // when the first removed more than half of the norm (DGKS criterion)
    double nt = 0.0, nt0 = 0.0;
    for (int pass = 0; pass < 2; ++pass) {
      dev_mdots(Vs(0), n, k + 1, tk, n, cv, part, ctx->stream);
      dev_combine_norm(tk, 1.0, Vs(0), n, k, cv, -1.0, n, dots, part, ctx->stream);
      const auto pr = fetch2(cv + k, dots);
      if (pass == 0) nt0 = std::sqrt(std::max(pr.first, 0.0));
      nt = std::sqrt(std::max(pr.second, 0.0));
      if (nt * nt >= 0.5 * pr.first) break;
    }
    const bool stuck = !(nt > 1e-10 * nt0) || !std::isfinite(nt);
    double dtheta = 0.0;
    if (!stuck) {
      dev_scale(tk, 1.0 / nt, n, ctx->stream);
      hpsi(tk, Ws(k), PREC_FP32_STORE);
      ++used;
      dev_mdots(Vs(0), n, k + 1, Ws(k), n, dots, part, ctx->stream);
      const std::vector<double> col = fetch(dots, k + 1);
      for (int j = 0; j <= k; ++j) {
        T[(size_t)j * kmax + k] = finite(col[j]);
        T[(size_t)k * kmax + j] = col[j];
      }
      ++k;
      std::vector<double> Tk((size_t)k * k);
      for (int i = 0; i < k; ++i)
        for (int j = 0; j < k; ++j) Tk[(size_t)i * k + j] = T[(size_t)i * kmax + j];
      double lam;
      lowest_eig(Tk, k, &lam, c);
      dtheta = std::fabs(lam - theta);
      theta = lam;
      r_ready = false;
    }
    const bool out_of_budget = used >= max_hpsi;
    if (stuck || dtheta <= lanzero || k == kmax || out_of_budget) {
      // the round's last Ritz value (FP32 products) and whether the round
      // ended by the stopping rule itself
      const double theta_round = theta;
      const bool stagnated = !stuck && dtheta <= lanzero;
      // restart from the Ritz vector with an FP64 product
      if (k > 1) {
        put(cv, c);
        dev_combine_norm(t, 0.0, Vs(0), n, k, cv, 1.0, n, dots, part, ctx->stream);
        const double nu = std::sqrt(fetch(dots, 1)[0]);
        dev_scale_copy(Vs(0), t, 1.0 / nu, n, ctx->stream);
      }
      anchor();
      ++used;
      r_ready = true;
      const double de = std::fabs(theta - anchor_e);
      out.ritz_change = de;
      if (verbose())
        std::fprintf(stderr, "#   davidson restart: %d FP32 products so far, round ended by %s; round Ritz %.12f, "
                     "exact %.12f (moved %.2e since the last restart), residual %.3e (was %.3e)\n",
                     out.hpsi_fp32, stuck ? "stagnation of t" : (stagnated ? "the stopping rule" : "a full subspace or the budget"),
                     theta_round, theta, de, rho, anchor_rho);
      // converged: the exact Rayleigh quotient moved by at most lanzero since
      // the previous restart, or the round ended by the stopping rule and the
      // exact Rayleigh quotient of its Ritz vector confirms the round's last
      // Ritz value to lanzero (the FP32 products did not move it)
      const bool confirmed = stagnated && std::fabs(theta - theta_round) <= lanzero;
      if (de <= lanzero || rho <= lanzero || confirmed) {
        out.converged = true;
        break;
      }
      if (used >= max_hpsi) break;
      if (rho > fall_factor * anchor_rho && mode != PC_NONE) {
        // the FP64 residual did not fall: finish with the mixed-precision
        // Lanczos from the current vector
        if (fell_back) *fell_back = true;
        dev_scale_copy(psi, Vs(0), 1.0, n, ctx->stream);
        check_hip(hipStreamSynchronize(ctx->stream), "davidson");
        SolveOut so = mixed_lanczos(ctx, sb, psi, lanzero, std::max(2, max_hpsi - used), 4);
        so.hpsi_fp64 += out.hpsi_fp64;
        so.hpsi_fp32 += out.hpsi_fp32;
        so.iters += used;
        so.mv_seconds += mv;
        return so;
      }
      anchor_e = theta;
      anchor_rho = rho;
    }
  }
  dev_scale_copy(psi, Vs(0), 1.0, n, ctx->stream);
  check_hip(hipStreamSynchronize(ctx->stream), "davidson");
  out.e0 = theta;
  out.iters = used;
  out.residual = rho;
  out.mv_seconds = mv;
  ctx->last_converged = out.converged ? 1 : 0;
  ctx->last_ritz_change = out.ritz_change;
  ctx->last_residual = out.residual;
  return out;
}

std::unique_ptr<PcSample> make_sample(hh_ctx* ctx, const hh_superblock* sb, const double2* u0, const double2* r0,
                                      double theta0, const double2* psi) {
  const int64_t n = sb->plan.sb.dim;
  std::unique_ptr<PcSample> s(new PcSample());
  precond_setup(ctx, sb, s->P);
  s->r.alloc(sizeof(double2) * (size_t)n);
  s->delta.alloc(sizeof(double2) * (size_t)n);
  dev_scale_copy(s->r.c(), r0, 1.0, n, ctx->stream);
  ensure_ws(ctx, 1, 1);
  double* dots = (double*)ctx->dots.p;
  double* part = (double*)ctx->partial.p;
  auto one = [&]() {
    double h = 0.0;
    check_hip(hipMemcpyAsync(&h, dots, sizeof(double), hipMemcpyDeviceToHost, ctx->stream), "d2h");
    check_hip(hipStreamSynchronize(ctx->stream), "sync");
    return h;
  };
  // the partial-sum buffer must hold one vector's blocks
  ctx->partial.alloc(sizeof(double) * (size_t)std::max(rdot_partials(), vec_blocks(n)) * 4);
  part = (double*)ctx->partial.p;
  dev_mdots(u0, 0, 1, psi, n, dots, part, ctx->stream);
  const double ov = one();
  need(std::fabs(ov) > 1e-8, "start vector orthogonal to the ground state: no training triple");
  // delta = psi / Re<u0, psi> - u0
  dev_scale_copy(s->delta.c(), psi, 1.0 / ov, n, ctx->stream);
  dev_lanczos_update(s->delta.c(), u0, 1.0, nullptr, 0.0, n, ctx->stream);
  dev_mdots(s->delta.c(), 0, 1, s->delta.c(), n, dots, part, ctx->stream);
  s->dn = one();
  s->theta = theta0;
  s->n_mid = ctx->model.n_elec / 2;
  need(s->dn > 0.0 && std::isfinite(s->dn), "degenerate training triple");
  return s;
}

void pc_train(hh_ctx* ctx, PcNet& net, const std::vector<std::unique_ptr<PcSample>>& samples, int epochs, double lr,
              std::vector<double>* losses) {
  need(!samples.empty(), "no training triples captured");
  need(epochs >= 1, "epochs must be >= 1");
  const int n = MlpLayout::SIZE;
  for (int e = 0; e < epochs; ++e) {
    double sum = 0.0;
    for (const auto& sp : samples) {
      PcSample& s = *sp;
      pc_features(s.P, s.r.c(), s.theta, s.n_mid, ctx->stream);
      pc_forward(s.P, (const float*)net.w.p, ctx->stream);
      pc_backward(s.P, s.r.c(), s.delta.c(), s.theta, s.dn, (const float*)net.w.p, (float*)net.g.p,
                  (double*)net.loss.p, ctx->stream);
      pc_adam((float*)net.w.p, (const float*)net.g.p, (float*)net.m.p, (float*)net.v.p, n, (float)lr, ++net.adam_t,
              ctx->stream);
      double l = 0.0;
      check_hip(hipMemcpyAsync(&l, net.loss.p, sizeof(double), hipMemcpyDeviceToHost, ctx->stream), "d2h");
      check_hip(hipStreamSynchronize(ctx->stream), "train");
      sum += l;
    }
    if (losses) losses->push_back(sum / (double)samples.size());
  }
  net.trained = true;
}

}  // namespace hh
