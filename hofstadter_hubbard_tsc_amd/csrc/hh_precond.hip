// Learned per-sector Davidson preconditioner: diagonal estimate, sector
// features, the preconditioner itself, the training loss gradient, and the
// forward / backward / Adam kernels of the small MLP (FP32).  Every
// reduction runs in a fixed order, so results are bit-reproducible.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <stdexcept>
#include <string>

#include "hh_engine.hpp"
#include "hh_precond.hpp"

namespace hh {

static constexpr int PB = 256;   // threads per workgroup

template <typename T>
static void dalloc(T*& p, size_t n) {
  check_hip(hipMalloc((void**)&p, sizeof(T) * std::max<size_t>(n, 1)), "hipMalloc precond");
}
template <typename T>
static void dfree(T*& p) {
  if (p) (void)hipFree(p);
  p = nullptr;
}

void PcDev::release() {
  dfree(off);
  dfree(rows);
  dfree(cols);
  dfree(s2);
  dfree(nq);
  dfree(pid);
  dfree(d);
  dfree(stats);
  dfree(feat);
  dfree(h1);
  dfree(h2);
  dfree(out);
  dfree(par);
  dfree(gout);
  dfree(dout);
  dfree(dh2);
  dfree(dh1);
  dfree(glob);
  dfree(ch_pair);
  dfree(ch_off);
  dfree(ch_cnt);
  dfree(cfirst);
  dfree(cstats);
  np = 0;
  n = 0;
  nchunk = 0;
}

// deterministic workgroup sum of one double per thread (all threads get it)
__device__ double block_sum(double x, double* sh) {
  sh[threadIdx.x] = x;
  __syncthreads();
  for (int s = PB / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
    __syncthreads();
  }
  const double r = sh[0];
  __syncthreads();
  return r;
}
__device__ double block_min(double x, double* sh) {
  sh[threadIdx.x] = x;
  __syncthreads();
  for (int s = PB / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sh[threadIdx.x] = fmin(sh[threadIdx.x], sh[threadIdx.x + s]);
    __syncthreads();
  }
  const double r = sh[0];
  __syncthreads();
  return r;
}

// d[i] = hs[soff_p + a] + he[eoff_p + b] and pid[i] = p for element (a, b)
// of sector p; one workgroup per chunk
__global__ __launch_bounds__(PB) void k_build_d(const int32_t* __restrict__ ch_pair, const int64_t* __restrict__ ch_off,
                                                const int32_t* __restrict__ ch_cnt, const int64_t* __restrict__ off,
                                                const int32_t* __restrict__ rows, const int32_t* __restrict__ soff,
                                                const int32_t* __restrict__ eoff, const double* __restrict__ hs,
                                                const double* __restrict__ he, double* __restrict__ d,
                                                int32_t* __restrict__ pid) {
  const int c = blockIdx.x;
  const int p = ch_pair[c];
  const int R = rows[p];
  const int64_t o0 = ch_off[c], e0 = o0 - off[p];
  for (int e = threadIdx.x; e < ch_cnt[c]; e += PB) {
    const int64_t le = e0 + e;
    const int a = (int)(le % R), b = (int)(le / R);
    d[o0 + e] = hs[soff[p] + a] + he[eoff[p] + b];
    pid[o0 + e] = p;
  }
}

// per chunk: sum d, sum d^2, min d, sum |r|^2
__global__ __launch_bounds__(PB) void k_chunk_stats(const int64_t* __restrict__ ch_off,
                                                    const int32_t* __restrict__ ch_cnt, const double* __restrict__ d,
                                                    const double2* __restrict__ r, double* __restrict__ cstats) {
  __shared__ double sh[PB];
  const int c = blockIdx.x;
  const int64_t o = ch_off[c];
  const int cnt = ch_cnt[c];
  double s1 = 0.0, s2 = 0.0, mn = 1e300, rr = 0.0;
  for (int e = threadIdx.x; e < cnt; e += PB) {
    const double x = d[o + e];
    s1 += x;
    s2 += x * x;
    mn = fmin(mn, x);
    const double2 y = r[o + e];
    rr += y.x * y.x + y.y * y.y;
  }
  s1 = block_sum(s1, sh);
  s2 = block_sum(s2, sh);
  mn = block_min(mn, sh);
  rr = block_sum(rr, sh);
  if (threadIdx.x == 0) {
    cstats[4 * c + 0] = s1;
    cstats[4 * c + 1] = s2;
    cstats[4 * c + 2] = mn;
    cstats[4 * c + 3] = rr;
  }
}

// per sector: its chunks' sums in chunk order (slot 2 is a minimum when
// mode_min is set)
__global__ void k_chunk_reduce(int np, const int32_t* __restrict__ cfirst, const double* __restrict__ cstats,
                               double* __restrict__ stats, int mode_min) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np) return;
  double a[4] = {0.0, 0.0, mode_min ? 1e300 : 0.0, 0.0};
  for (int c = cfirst[p]; c < cfirst[p + 1]; ++c) {
    a[0] += cstats[4 * c + 0];
    a[1] += cstats[4 * c + 1];
    a[2] = mode_min ? fmin(a[2], cstats[4 * c + 2]) : a[2] + cstats[4 * c + 2];
    a[3] += cstats[4 * c + 3];
  }
  for (int q = 0; q < 4; ++q) stats[4 * p + q] = a[q];
}

// one workgroup: global scale of d and |r|^2 from the sector sums, then the
// features of every sector
__global__ __launch_bounds__(PB) void k_features(int np, const int32_t* __restrict__ rows,
                                                 const int32_t* __restrict__ cols, const int32_t* __restrict__ s2,
                                                 const int32_t* __restrict__ nq, const double* __restrict__ stats,
                                                 double theta, int n_mid, float* __restrict__ feat,
                                                 double* __restrict__ glob) {
  __shared__ double sh[PB];
  double c = 0.0, s1 = 0.0, sq = 0.0, rr = 0.0;
  for (int p = threadIdx.x; p < np; p += PB) {
    c += (double)rows[p] * cols[p];
    s1 += stats[4 * p + 0];
    sq += stats[4 * p + 1];
    rr += stats[4 * p + 3];
  }
  c = block_sum(c, sh);
  s1 = block_sum(s1, sh);
  sq = block_sum(sq, sh);
  rr = block_sum(rr, sh);
  const double mean = s1 / fmax(c, 1.0);
  double sd = sqrt(fmax(sq / fmax(c, 1.0) - mean * mean, 0.0));
  sd = fmax(sd, fmax(1e-3 * fabs(mean), 1e-12));
  if (threadIdx.x == 0) {
    glob[0] = sd;
    glob[1] = rr;
  }
  const double lr = 0.5 * log10(fmax(rr, 1e-300));
  for (int p = threadIdx.x; p < np; p += PB) {
    const double cnt = (double)rows[p] * cols[p];
    const double m = stats[4 * p + 0] / cnt;
    const double v = fmax(stats[4 * p + 1] / cnt - m * m, 0.0);
    float* f = feat + (size_t)p * PC_NF;
    f[0] = (float)((m - theta) / sd);
    f[1] = (float)((stats[4 * p + 2] - theta) / sd);
    f[2] = (float)(sqrt(v) / sd);
    f[3] = (float)(log(cnt) / 10.0);
    f[4] = (float)(log(stats[4 * p + 3] / fmax(rr, 1e-300) + 1e-30) / 10.0);
    f[5] = (float)(lr / 4.0);
    f[6] = (float)(s2[p] / 8.0);
    f[7] = (float)((nq[p] - n_mid) / 8.0);
  }
}

// Y[s, o] = act(sum_i X[s, i] W[o, i] + b[o]); act 1 = tanh, 0 = identity
__global__ __launch_bounds__(PB) void k_dense(const float* __restrict__ X, const float* __restrict__ W,
                                              const float* __restrict__ b, float* __restrict__ Y, int B, int in, int out,
                                              int act) {
  const int64_t g = blockIdx.x * (int64_t)PB + threadIdx.x;
  if (g >= (int64_t)B * out) return;
  const int s = (int)(g / out), o = (int)(g % out);
  float acc = b[o];
  const float* x = X + (size_t)s * in;
  const float* w = W + (size_t)o * in;
  for (int i = 0; i < in; ++i) acc += x[i] * w[i];
  Y[g] = act ? tanhf(acc) : acc;
}

// sigma_p = PC_SIGMA0 scale exp(PC_BOUND tanh(o_0)), alpha_p = exp(PC_BOUND tanh(o_1))
__global__ void k_head(int np, const float* __restrict__ out, const double* __restrict__ glob, double* __restrict__ par) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np) return;
  const double o0 = out[2 * p], o1 = out[2 * p + 1];
  par[2 * p] = PC_SIGMA0 * glob[0] * exp(PC_BOUND * tanh(o0));
  par[2 * p + 1] = exp(PC_BOUND * tanh(o1));
}

__global__ void k_fixed(int np, double sigma0, const double* __restrict__ glob, double* __restrict__ par) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np) return;
  par[2 * p] = sigma0 * glob[0];
  par[2 * p + 1] = 1.0;
}

__global__ __launch_bounds__(PB) void k_apply(int64_t n, const int32_t* __restrict__ pid, const double* __restrict__ d,
                                              const double* __restrict__ par, const double2* __restrict__ r,
                                              double theta, double2* __restrict__ t) {
  for (int64_t i = blockIdx.x * (int64_t)PB + threadIdx.x; i < n; i += (int64_t)gridDim.x * PB) {
    const int p = pid[i];
    const double den = fmax(d[i] - theta, 0.0) + par[2 * p];
    const double f = par[2 * p + 1] / den;
    const double2 y = r[i];
    t[i] = make_double2(f * y.x, f * y.y);
  }
}

// per chunk of sector p, with t = alpha r / den, den = max(d - theta, 0) +
// sigma: S1 = sum Re(conj(delta) t) / den, S2 = sum |t|^2 / den,
// S3 = sum Re(conj(delta) t), S4 = sum |t|^2
__global__ __launch_bounds__(PB) void k_loss_sums(const int32_t* __restrict__ ch_pair,
                                                  const int64_t* __restrict__ ch_off,
                                                  const int32_t* __restrict__ ch_cnt, const double* __restrict__ d,
                                                  const double* __restrict__ par, const double2* __restrict__ r,
                                                  const double2* __restrict__ delta, double theta,
                                                  double* __restrict__ cstats) {
  __shared__ double sh[PB];
  const int c = blockIdx.x;
  const int p = ch_pair[c];
  const int64_t o = ch_off[c];
  const int cnt = ch_cnt[c];
  const double sig = par[2 * p], al = par[2 * p + 1];
  double s1 = 0.0, s2 = 0.0, s3 = 0.0, s4 = 0.0;
  for (int e = threadIdx.x; e < cnt; e += PB) {
    const double den = fmax(d[o + e] - theta, 0.0) + sig;
    const double2 y = r[o + e], z = delta[o + e];
    const double tx = al * y.x / den, ty = al * y.y / den;
    const double dt = z.x * tx + z.y * ty, tt = tx * tx + ty * ty;
    s1 += dt / den;
    s2 += tt / den;
    s3 += dt;
    s4 += tt;
  }
  s1 = block_sum(s1, sh);
  s2 = block_sum(s2, sh);
  s3 = block_sum(s3, sh);
  s4 = block_sum(s4, sh);
  if (threadIdx.x == 0) {
    cstats[4 * c + 0] = s1;
    cstats[4 * c + 1] = s2;
    cstats[4 * c + 2] = s3;
    cstats[4 * c + 3] = s4;
  }
}

// one workgroup: the loss L = 1 - A^2 / (B |delta|^2) with A = Re<t, delta>,
// B = |t|^2 (the part of delta a Davidson step along t can capture is A^2 /
// (B |delta|^2), whatever the scale of t), and dL/d(out) per sector
__global__ __launch_bounds__(PB) void k_loss_head(int np, const double* __restrict__ gout,
                                                  const double* __restrict__ par, const float* __restrict__ out,
                                                  double dn,
                                                  float* __restrict__ dout, double* __restrict__ loss) {
  __shared__ double sh[PB];
  double A = 0.0, B = 0.0;
  for (int p = threadIdx.x; p < np; p += PB) {
    A += gout[4 * p + 2];
    B += gout[4 * p + 3];
  }
  A = block_sum(A, sh);
  B = block_sum(B, sh);
  B = fmax(B, 1e-300);
  if (threadIdx.x == 0) loss[0] = 1.0 - A * A / (B * dn);
  const double ka = 2.0 * A / (B * dn), kb = 2.0 * A * A / (B * B * dn);
  for (int p = threadIdx.x; p < np; p += PB) {
    const double sg = par[2 * p], al = par[2 * p + 1];
    const double gs = ka * gout[4 * p + 0] - kb * gout[4 * p + 1];
    const double ga = (-ka * gout[4 * p + 2] + kb * gout[4 * p + 3]) / al;
    const double t0 = tanh((double)out[2 * p]), t1 = tanh((double)out[2 * p + 1]);
    // d sigma / d o_0 = sigma PC_BOUND (1 - tanh^2), likewise alpha
    dout[2 * p] = (float)(gs * sg * PC_BOUND * (1.0 - t0 * t0));
    dout[2 * p + 1] = (float)(ga * al * PC_BOUND * (1.0 - t1 * t1));
  }
}

// dX[s, i] = (sum_o dY[s, o] W[o, i]) * (1 - H[s, i]^2)  (tanh layer below)
__global__ __launch_bounds__(PB) void k_dense_back_in(const float* __restrict__ dY, const float* __restrict__ W,
                                                      const float* __restrict__ H, float* __restrict__ dX, int B,
                                                      int in, int out) {
  const int64_t g = blockIdx.x * (int64_t)PB + threadIdx.x;
  if (g >= (int64_t)B * in) return;
  const int s = (int)(g / in), i = (int)(g % in);
  float acc = 0.f;
  for (int o = 0; o < out; ++o) acc += dY[(size_t)s * out + o] * W[(size_t)o * in + i];
  const float h = H[g];
  dX[g] = acc * (1.f - h * h);
}

// dW[o, i] = sum_s dY[s, o] X[s, i];  db[o] = sum_s dY[s, o]
__global__ __launch_bounds__(PB) void k_dense_back_w(const float* __restrict__ dY, const float* __restrict__ X,
                                                     float* __restrict__ dW, float* __restrict__ db, int B, int in,
                                                     int out) {
  const int64_t g = blockIdx.x * (int64_t)PB + threadIdx.x;
  if (g < (int64_t)out * in) {
    const int o = (int)(g / in), i = (int)(g % in);
    float acc = 0.f;
    for (int s = 0; s < B; ++s) acc += dY[(size_t)s * out + o] * X[(size_t)s * in + i];
    dW[g] = acc;
  }
  if (g < out) {
    float acc = 0.f;
    for (int s = 0; s < B; ++s) acc += dY[(size_t)s * out + g];
    db[g] = acc;
  }
}

__global__ __launch_bounds__(PB) void k_adam(float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m,
                                             float* __restrict__ v, int n, float lr, float c1, float c2) {
  const int i = blockIdx.x * PB + threadIdx.x;
  if (i >= n) return;
  const float b1 = 0.9f, b2 = 0.999f;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + 1e-8f);
}

__global__ __launch_bounds__(PB) void k_gather_re(double* __restrict__ out, const double2* __restrict__ a,
                                                  const int64_t* __restrict__ idx, int n) {
  const int i = blockIdx.x * PB + threadIdx.x;
  if (i < n) out[i] = a[idx[i]].x;
}

static inline unsigned nblk(int64_t n) { return (unsigned)std::max<int64_t>(1, (n + PB - 1) / PB); }

void pc_setup(PcDev& P, const PairTable& t, const std::vector<double>& hs, const std::vector<double>& he, int n_mid,
              hipStream_t s) {
  (void)n_mid;
  P.release();
  P.np = t.np;
  P.n = t.n;
  const int np = t.np;
  dalloc(P.off, np);
  dalloc(P.rows, np);
  dalloc(P.cols, np);
  dalloc(P.s2, np);
  dalloc(P.nq, np);
  dalloc(P.pid, t.n);
  dalloc(P.d, t.n);
  dalloc(P.stats, 4 * (size_t)np);
  dalloc(P.feat, (size_t)np * PC_NF);
  dalloc(P.h1, (size_t)np * PC_H);
  dalloc(P.h2, (size_t)np * PC_H);
  dalloc(P.out, (size_t)np * PC_NO);
  dalloc(P.par, 2 * (size_t)np);
  dalloc(P.gout, 4 * (size_t)np);
  dalloc(P.dout, (size_t)np * PC_NO);
  dalloc(P.dh2, (size_t)np * PC_H);
  dalloc(P.dh1, (size_t)np * PC_H);
  dalloc(P.glob, 4);
  // chunk table
  std::vector<int32_t> cpair, ccnt, cfirst(np + 1, 0);
  std::vector<int64_t> coff;
  for (int p = 0; p < np; ++p) {
    cfirst[p] = (int32_t)cpair.size();
    const int64_t cnt = (int64_t)t.rows[p] * t.cols[p];
    for (int64_t e = 0; e < cnt; e += PC_CHUNK) {
      cpair.push_back(p);
      coff.push_back(t.off[p] + e);
      ccnt.push_back((int32_t)std::min<int64_t>(PC_CHUNK, cnt - e));
    }
  }
  cfirst[np] = (int32_t)cpair.size();
  P.nchunk = (int)cpair.size();
  dalloc(P.ch_pair, P.nchunk);
  dalloc(P.ch_off, P.nchunk);
  dalloc(P.ch_cnt, P.nchunk);
  dalloc(P.cfirst, np + 1);
  dalloc(P.cstats, 4 * (size_t)P.nchunk);
  int32_t *soff = nullptr, *eoff = nullptr;
  double *dhs = nullptr, *dhe = nullptr;
  dalloc(soff, np);
  dalloc(eoff, np);
  dalloc(dhs, hs.size());
  dalloc(dhe, he.size());
  auto up = [&](void* dst, const void* src, size_t b) {
    if (b) check_hip(hipMemcpyAsync(dst, src, b, hipMemcpyHostToDevice, s), "h2d precond");
  };
  up(P.off, t.off.data(), sizeof(int64_t) * np);
  up(P.rows, t.rows.data(), sizeof(int32_t) * np);
  up(P.cols, t.cols.data(), sizeof(int32_t) * np);
  up(P.s2, t.s2.data(), sizeof(int32_t) * np);
  up(P.nq, t.nq.data(), sizeof(int32_t) * np);
  up(soff, t.soff.data(), sizeof(int32_t) * np);
  up(eoff, t.eoff.data(), sizeof(int32_t) * np);
  up(dhs, hs.data(), sizeof(double) * hs.size());
  up(dhe, he.data(), sizeof(double) * he.size());
  up(P.ch_pair, cpair.data(), sizeof(int32_t) * P.nchunk);
  up(P.ch_off, coff.data(), sizeof(int64_t) * P.nchunk);
  up(P.ch_cnt, ccnt.data(), sizeof(int32_t) * P.nchunk);
  up(P.cfirst, cfirst.data(), sizeof(int32_t) * (np + 1));
  if (P.nchunk)
    hipLaunchKernelGGL(k_build_d, dim3(P.nchunk), dim3(PB), 0, s, P.ch_pair, P.ch_off, P.ch_cnt, P.off, P.rows, soff,
                       eoff, dhs, dhe, P.d, P.pid);
  check_hip(hipGetLastError(), "k_build_d");
  check_hip(hipStreamSynchronize(s), "precond setup");
  dfree(soff);
  dfree(eoff);
  dfree(dhs);
  dfree(dhe);
}

void pc_features(PcDev& P, const double2* r, double theta, int n_mid, hipStream_t s) {
  if (!P.np) return;
  hipLaunchKernelGGL(k_chunk_stats, dim3(P.nchunk), dim3(PB), 0, s, P.ch_off, P.ch_cnt, P.d, r, P.cstats);
  hipLaunchKernelGGL(k_chunk_reduce, dim3(nblk(P.np)), dim3(PB), 0, s, P.np, P.cfirst, P.cstats, P.stats, 1);
  hipLaunchKernelGGL(k_features, dim3(1), dim3(PB), 0, s, P.np, P.rows, P.cols, P.s2, P.nq, P.stats, theta, n_mid,
                     P.feat, P.glob);
  check_hip(hipGetLastError(), "precond features");
}

void pc_forward(PcDev& P, const float* w, hipStream_t s) {
  if (!P.np) return;
  const int B = P.np;
  hipLaunchKernelGGL(k_dense, dim3(nblk((int64_t)B * PC_H)), dim3(PB), 0, s, P.feat, w + MlpLayout::W1,
                     w + MlpLayout::B1, P.h1, B, PC_NF, PC_H, 1);
  hipLaunchKernelGGL(k_dense, dim3(nblk((int64_t)B * PC_H)), dim3(PB), 0, s, P.h1, w + MlpLayout::W2,
                     w + MlpLayout::B2, P.h2, B, PC_H, PC_H, 1);
  hipLaunchKernelGGL(k_dense, dim3(nblk((int64_t)B * PC_NO)), dim3(PB), 0, s, P.h2, w + MlpLayout::W3,
                     w + MlpLayout::B3, P.out, B, PC_H, PC_NO, 0);
  hipLaunchKernelGGL(k_head, dim3(nblk(B)), dim3(PB), 0, s, B, P.out, P.glob, P.par);
  check_hip(hipGetLastError(), "precond forward");
}

void pc_fixed(PcDev& P, double sigma0, hipStream_t s) {
  if (!P.np) return;
  hipLaunchKernelGGL(k_fixed, dim3(nblk(P.np)), dim3(PB), 0, s, P.np, sigma0, P.glob, P.par);
  check_hip(hipGetLastError(), "precond fixed");
}

void pc_apply(const PcDev& P, const double2* r, double theta, double2* t, hipStream_t s) {
  if (P.n <= 0) return;
  const unsigned g = (unsigned)std::min<int64_t>(nblk(P.n), 8192);
  hipLaunchKernelGGL(k_apply, dim3(g), dim3(PB), 0, s, P.n, P.pid, P.d, P.par, r, theta, t);
  check_hip(hipGetLastError(), "precond apply");
}

void pc_backward(PcDev& P, const double2* r, const double2* delta, double theta, double dn, const float* w, float* g,
                 double* loss_dev, hipStream_t s) {
  if (!P.np) return;
  const int B = P.np;
  hipLaunchKernelGGL(k_loss_sums, dim3(P.nchunk), dim3(PB), 0, s, P.ch_pair, P.ch_off, P.ch_cnt, P.d, P.par, r, delta,
                     theta, P.cstats);
  hipLaunchKernelGGL(k_chunk_reduce, dim3(nblk(B)), dim3(PB), 0, s, B, P.cfirst, P.cstats, P.gout, 0);
  hipLaunchKernelGGL(k_loss_head, dim3(1), dim3(PB), 0, s, B, P.gout, P.par, P.out, dn, P.dout, loss_dev);
  // layer 3 (linear): dW3, db3; dH2 through tanh
  hipLaunchKernelGGL(k_dense_back_w, dim3(nblk((int64_t)PC_NO * PC_H)), dim3(PB), 0, s, P.dout, P.h2,
                     g + MlpLayout::W3, g + MlpLayout::B3, B, PC_H, PC_NO);
  hipLaunchKernelGGL(k_dense_back_in, dim3(nblk((int64_t)B * PC_H)), dim3(PB), 0, s, P.dout, w + MlpLayout::W3, P.h2,
                     P.dh2, B, PC_H, PC_NO);
  // layer 2
  hipLaunchKernelGGL(k_dense_back_w, dim3(nblk((int64_t)PC_H * PC_H)), dim3(PB), 0, s, P.dh2, P.h1,
                     g + MlpLayout::W2, g + MlpLayout::B2, B, PC_H, PC_H);
  hipLaunchKernelGGL(k_dense_back_in, dim3(nblk((int64_t)B * PC_H)), dim3(PB), 0, s, P.dh2, w + MlpLayout::W2, P.h1,
                     P.dh1, B, PC_H, PC_H);
  // layer 1
  hipLaunchKernelGGL(k_dense_back_w, dim3(nblk((int64_t)PC_H * PC_NF)), dim3(PB), 0, s, P.dh1, P.feat,
                     g + MlpLayout::W1, g + MlpLayout::B1, B, PC_NF, PC_H);
  check_hip(hipGetLastError(), "precond backward");
}

void pc_adam(float* w, const float* g, float* m, float* v, int n, float lr, int t, hipStream_t s) {
  const float c1 = 1.f - std::pow(0.9f, (float)t), c2 = 1.f - std::pow(0.999f, (float)t);
  hipLaunchKernelGGL(k_adam, dim3(nblk(n)), dim3(PB), 0, s, w, g, m, v, n, lr, c1, c2);
  check_hip(hipGetLastError(), "adam");
}

void dev_gather_re(double* out, const double2* arena, const int64_t* idx, int n, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_gather_re, dim3(nblk(n)), dim3(PB), 0, s, out, arena, idx, n);
  check_hip(hipGetLastError(), "gather");
}

}  // namespace hh
