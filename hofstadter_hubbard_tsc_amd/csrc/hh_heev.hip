// Hand-written eigensolver of the density-matrix sectors (the per-sector
// ZHEEV of Get_truncation_operator, code/get_den.f90:242); the algorithm is
// described in hh_heev.hpp.  The reference's ZHEEV call is one LAPACK call per
// sector; a library eigensolver on the GPU issues about twenty small kernels
// per column of every sector.  Here every sector of a truncation advances
// together: one launch per Householder column (all sectors), one bisection
// launch, one inverse-iteration launch, a few orthonormalisation launches and
// one back-transformation launch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

#include "hh_engine.hpp"
#include "hh_heev.hpp"

namespace hh {

void HeevBuf::ensure(size_t b) {
  if (b <= bytes && p) return;
  if (p) (void)hipFree(p);
  p = nullptr;
  bytes = 0;
  if (b == 0) b = 16;
  check_hip(hipMalloc(&p, b), "hipMalloc eigensolver");
  bytes = b;
}

HeevBuf::~HeevBuf() {
  if (p) (void)hipFree(p);
}

namespace {

constexpr int HT = 256;   // threads per workgroup
constexpr int CW = 8;     // trailing columns per workgroup in a Householder step
constexpr int GT = 32;    // Gram / Newton-Schulz tile

__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// a conj(b)
__device__ __forceinline__ double2 cmulc(double2 a, double2 b) {
  return make_double2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
// conj(a) b
__device__ __forceinline__ double2 cjmul(double2 a, double2 b) {
  return make_double2(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}

// butterfly sum over the 64 lanes: every lane ends with the same bits
__device__ __forceinline__ double wsum(double x) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o);
  return x;
}
__device__ __forceinline__ double wmax(double x) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) x = fmax(x, __shfl_xor(x, o));
  return x;
}

// Hermitian blocks from their upper triangle: A(r, c) = conj(A(c, r)) for
// r > c, real diagonal.  One 32 x 32 tile pair per workgroup (bi >= bj).
__global__ __launch_bounds__(HT) void k_herm_fix(const HeevSec* __restrict__ secs, const int4* __restrict__ tiles,
                                                 double2* __restrict__ A) {
  const int4 t = tiles[blockIdx.x];   // (sector, bi, bj, -)
  const HeevSec S = secs[t.x];
  const int n = S.n;
  double2* M = A + S.a;
  __shared__ double2 sm[32][33];
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  for (int yy = ty; yy < 32; yy += 8) {
    const int r = t.z * 32 + tx, c = t.y * 32 + yy;   // upper tile: rows of block bj, columns of block bi
    if (r < n && c < n) sm[yy][tx] = M[r + (int64_t)c * n];
  }
  __syncthreads();
  for (int yy = ty; yy < 32; yy += 8) {
    const int r = t.y * 32 + tx, c = t.z * 32 + yy;   // lower tile: rows of block bi, columns of block bj
    if (r < n && c < n) {
      const double2 v = sm[tx][yy];                   // A(c, r)
      if (r > c) M[r + (int64_t)c * n] = make_double2(v.x, -v.y);
      else if (r == c) M[r + (int64_t)c * n] = make_double2(v.x, 0.0);
    }
  }
}

// One Householder column k of every sector (zhetd2's lower form: reflector
// H_k = I - tau_k v_k v_k^H with v_k(k + 1) = 1 maps A(k + 1 : n, k) to
// beta_k e_1, and the trailing block becomes H_k^H A H_k through the rank-2
// update A -= v w^H + w v^H with p = tau A v, w = p - tau (p^H v) v / 2).
// The update of reflector k - 1 is applied here, fused with the product p
// of reflector k: each workgroup owns CW trailing columns of one sector; it
// applies reflector k - 1 to column k (read by every workgroup of the sector,
// written by none in this step) to build reflector k in LDS, then updates
// its own columns once and forms their entries of p_k and their share of
// p_k^H v_k.  The four waves split the rows (64-row chunks dealt round
// robin, RCH chunks per pass so that their loads are in flight together)
// and each takes all of the workgroup's columns; the column sums meet in LDS
// in a fixed order.  The step is a chain of dependent memory round trips, so
// every phase issues its loads at once: the workgroup record carries the
// sector's fields, the column phase is unrolled over at most HEEV_MAXN / HT
// rows per thread.  v, p and the share sums alternate between two halves of
// their buffers (step k reads half (k - 1) & 1, writes half k & 1).  The
// workgroup holding column k writes d_k, e_k = beta_k, tau_k and v_k, and
// stores reflector k - 1 into the strictly lower part of column k - 1.
constexpr int RCH = 3;
constexpr int CROWS = HEEV_MAXN / HT;   // column-phase rows per thread
__global__ __launch_bounds__(HT) void k_hetrd_step(int k, const HeevWg* __restrict__ wgs, double2* __restrict__ A,
                                                   double* __restrict__ dd, double* __restrict__ ee,
                                                   double2* __restrict__ tau, double2* __restrict__ vb,
                                                   double2* __restrict__ pb, double2* __restrict__ part, int64_t ncol,
                                                   int64_t npart) {
  const HeevWg g = wgs[blockIdx.x];
  const int n = g.n;
  if (k >= n) return;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int c_lo = g.q * CW, c_hi = min(n, c_lo + CW);
  const int cur = k & 1, prv = cur ^ 1;
  double2* part_cur = part + (int64_t)cur * npart + g.ch0;
  if (c_hi <= k) {   // every column of this chunk is finished
    if (tid == 0) part_cur[g.q] = make_double2(0.0, 0.0);
    return;
  }
  __shared__ double2 vn[HEEV_MAXN];
  __shared__ double red[4];
  __shared__ double2 redc[4];
  __shared__ double2 redp[4][CW];
  double2* M = A + g.a;
  const int64_t ld = n;
  const double2* vo = vb + (int64_t)prv * ncol + g.c;
  const double2* po = pb + (int64_t)prv * ncol + g.c;
  const bool upd = k >= 1;
  // loads of the step issued together: the chunks' shares, column k, and
  // v, p of reflector k - 1 over the column's rows
  double2 sh = make_double2(0.0, 0.0);
  double2 ca[CROWS], cv[CROWS], cp[CROWS];
  double2 t_prev = make_double2(0.0, 0.0), vk = make_double2(0.0, 0.0), pk = make_double2(0.0, 0.0);
  if (upd) {
    const double2* pp = part + (int64_t)prv * npart + g.ch0;
    for (int q = tid; q < g.nch; q += HT) {
      const double2 x = pp[q];
      sh.x += x.x;
      sh.y += x.y;
    }
    t_prev = tau[g.c + k - 1];
    vk = vo[k];
    pk = po[k];
  }
#pragma unroll
  for (int i = 0; i < CROWS; ++i) {
    const int r = k + tid + HT * i;
    ca[i] = cv[i] = cp[i] = make_double2(0.0, 0.0);
    if (r < n) {
      ca[i] = M[r + (int64_t)k * ld];
      if (upd) {
        cv[i] = vo[r];
        cp[i] = po[r];
      }
    }
  }
  double2 alpha = make_double2(0.0, 0.0);   // w = p + alpha v
  if (upd) {
    const double sr = wsum(sh.x), si = wsum(sh.y);
    if (lane == 0) redc[wv] = make_double2(sr, si);
    __syncthreads();
    const double2 sm = make_double2((redc[0].x + redc[1].x) + (redc[2].x + redc[3].x),
                                    (redc[0].y + redc[1].y) + (redc[2].y + redc[3].y));
    const double2 ts = cmul(t_prev, sm);
    alpha = make_double2(-0.5 * ts.x, -0.5 * ts.y);
  }
  // column k after reflector k - 1, rows k .. n - 1
  const double2 wk = cadd(pk, cmul(alpha, vk));
  double xs = 0.0;
#pragma unroll
  for (int i = 0; i < CROWS; ++i) {
    const int r = k + tid + HT * i;
    if (r < n) {
      double2 a = ca[i];
      if (upd) {
        const double2 w = cadd(cp[i], cmul(alpha, cv[i]));
        a = csub(csub(a, cmulc(cv[i], wk)), cmulc(w, vk));
      }
      vn[r] = a;
      if (r >= k + 2) xs += a.x * a.x + a.y * a.y;
    }
  }
  xs = wsum(xs);
  if (lane == 0) red[wv] = xs;
  __syncthreads();
  const double xn2 = (red[0] + red[1]) + (red[2] + red[3]);
  const bool refl = k <= n - 2;
  const double2 akk = vn[k];
  double2 tk = make_double2(0.0, 0.0), sc = make_double2(0.0, 0.0);
  double beta = 0.0;
  if (refl) {
    const double2 ah = vn[k + 1];
    if (xn2 == 0.0 && ah.y == 0.0) {
      beta = ah.x;   // nothing to annihilate: H_k = I
    } else {
      const double nrm = sqrt(ah.x * ah.x + ah.y * ah.y + xn2);
      beta = ah.x >= 0.0 ? -nrm : nrm;
      tk = make_double2((beta - ah.x) / beta, -ah.y / beta);
      const double2 den = make_double2(ah.x - beta, ah.y);   // v = x / (alpha - beta)
      const double m2 = den.x * den.x + den.y * den.y;
      sc = make_double2(den.x / m2, -den.y / m2);
    }
  }
  __syncthreads();   // vn[k], vn[k + 1] read everywhere before vn becomes v_k
  if (refl)
    for (int r = k + 1 + tid; r < n; r += HT) vn[r] = r == k + 1 ? make_double2(1.0, 0.0) : cmul(vn[r], sc);
  __syncthreads();
  if (g.q == k / CW) {
    if (tid == 0) {
      dd[g.c + k] = akk.x;
      if (refl) {
        ee[g.c + k] = beta;
        tau[g.c + k] = tk;
      }
    }
    if (refl) {
      double2* vc = vb + (int64_t)cur * ncol + g.c;
      for (int r = k + 1 + tid; r < n; r += HT) vc[r] = vn[r];
    }
    if (upd)
      for (int r = k + tid; r < n; r += HT) M[r + (int64_t)(k - 1) * ld] = vo[r];
  }
  if (!refl) {
    if (tid == 0) part_cur[g.q] = make_double2(0.0, 0.0);
    return;
  }
  // own trailing columns c0 .. c_hi - 1 (at most CW), rows k + 1 .. n - 1
  const int c0 = max(c_lo, k + 1), nc = max(c_hi - c0, 0);
  double2 vcl[CW], wcl[CW], acc[CW];
#pragma unroll
  for (int j = 0; j < CW; ++j) {
    acc[j] = make_double2(0.0, 0.0);
    vcl[j] = make_double2(0.0, 0.0);
    wcl[j] = make_double2(0.0, 0.0);
    if (upd && j < nc) {
      vcl[j] = vo[c0 + j];
      wcl[j] = cadd(po[c0 + j], cmul(alpha, vcl[j]));
    }
  }
  for (int rb = k + 1 + 64 * wv; rb < n; rb += HT * RCH) {
    double2 a[RCH][CW], v[RCH], w[RCH];
#pragma unroll
    for (int h = 0; h < RCH; ++h) {
      const int r = rb + HT * h + lane;
      v[h] = w[h] = make_double2(0.0, 0.0);
      if (r < n) {
        if (upd) {
          v[h] = vo[r];
          w[h] = po[r];
        }
#pragma unroll
        for (int j = 0; j < CW; ++j)
          if (j < nc) a[h][j] = M[r + (int64_t)(c0 + j) * ld];
      }
    }
#pragma unroll
    for (int h = 0; h < RCH; ++h) {
      const int r = rb + HT * h + lane;
      if (r < n) {
        w[h] = cadd(w[h], cmul(alpha, v[h]));
        const double2 x = vn[r];
#pragma unroll
        for (int j = 0; j < CW; ++j) {
          if (j < nc) {
            double2 b = a[h][j];
            if (upd) {
              b = csub(csub(b, cmulc(v[h], wcl[j])), cmulc(w[h], vcl[j]));
              M[r + (int64_t)(c0 + j) * ld] = b;
            }
            const double2 t = cjmul(b, x);   // (A^H v)_c = sum_r conj(A(r, c)) v(r)
            acc[j].x += t.x;
            acc[j].y += t.y;
          }
        }
      }
    }
  }
#pragma unroll
  for (int j = 0; j < CW; ++j) {
    const double sr = wsum(acc[j].x), si = wsum(acc[j].y);
    if (lane == 0) redp[wv][j] = make_double2(sr, si);
  }
  __syncthreads();
  if (tid < 64) {
    double2* pc = pb + (int64_t)cur * ncol + g.c;
    double2 t = make_double2(0.0, 0.0);
    if (tid < nc) {
      const double2 sm = make_double2((redp[0][tid].x + redp[1][tid].x) + (redp[2][tid].x + redp[3][tid].x),
                                      (redp[0][tid].y + redp[1][tid].y) + (redp[2][tid].y + redp[3][tid].y));
      const double2 p = cmul(tk, sm);
      pc[c0 + tid] = p;
      t = cjmul(p, vn[c0 + tid]);
    }
    const double cx = wsum(t.x), cy = wsum(t.y);
    if (tid == 0) part_cur[g.q] = make_double2(cx, cy);
  }
}

// number of eigenvalues of the tridiagonal (d, e) below x (Sturm sequence of
// the LDL^T pivots; pivots smaller than pivmin are replaced by -pivmin)
__device__ int sturm_count(const double* __restrict__ d, const double* __restrict__ e, int n, double x,
                           double pivmin) {
  double q = d[0] - x;
  if (fabs(q) <= pivmin) q = -pivmin;
  int c = q < 0.0 ? 1 : 0;
  for (int j = 1; j < n; ++j) {
    q = (d[j] - x) - e[j - 1] * (e[j - 1] / q);
    if (fabs(q) <= pivmin) q = -pivmin;
    c += q < 0.0 ? 1 : 0;
  }
  return c;
}

// Eigenvalue i (ascending) of each sector's tridiagonal by multisection:
// a group of MS lanes per eigenvalue evaluates the Sturm count at MS interior
// points of its bracket at once and keeps the sub-interval where the count
// passes i (the group's ballot bits), from the Gershgorin interval until the
// bracket is a few ulps wide or below 1e-24 of the spectrum's scale.
constexpr int MS = 8;
__global__ __launch_bounds__(HT) void k_bisect(const HeevSec* __restrict__ secs, const int32_t* __restrict__ owner,
                                               int total, const double* __restrict__ d, const double* __restrict__ e,
                                               double* __restrict__ w) {
  const int gt = blockIdx.x * HT + threadIdx.x;
  const int gi = gt / MS, j = gt % MS;
  if (gi >= total) return;   // whole groups leave together (HT is a multiple of MS)
  const HeevSec S = secs[owner[gi]];
  const int n = S.n, i = (int)(gi - S.c);
  const double* dd = d + S.c;
  const double* ee = e + S.c;
  double gl = dd[0], gu = dd[0], emax2 = 0.0;
  for (int q = 0; q < n; ++q) {
    const double r = (q > 0 ? fabs(ee[q - 1]) : 0.0) + (q < n - 1 ? fabs(ee[q]) : 0.0);
    gl = fmin(gl, dd[q] - r);
    gu = fmax(gu, dd[q] + r);
    if (q < n - 1) emax2 = fmax(emax2, ee[q] * ee[q]);
  }
  const double pivmin = DBL_MIN * fmax(1.0, emax2);
  const double scale = fmax(fabs(gl), fabs(gu));
  const double pad = 4.0 * DBL_EPSILON * scale * n + 4.0 * pivmin;
  double lo = gl - pad, hi = gu + pad;
  const double atol = fmax(1e-24 * scale, 4.0 * pivmin);
  const int gbase = (threadIdx.x & 63) & ~(MS - 1);
  for (int it = 0; it < 128; ++it) {
    if (hi - lo <= fmax(2.0 * DBL_EPSILON * fmax(fabs(lo), fabs(hi)), atol)) break;
    const double step = (hi - lo) / (MS + 1);
    const double x = lo + (double)(j + 1) * step;
    const bool below = sturm_count(dd, ee, n, x, pivmin) > i;   // lambda_i < x
    const unsigned bits = (unsigned)((__ballot(below) >> gbase) & ((1ull << MS) - 1));
    const int jf = bits ? __builtin_ctz(bits) : MS;   // first point above lambda_i
    const double nlo = jf > 0 ? lo + (double)jf * step : lo;
    const double nhi = jf < MS ? lo + (double)(jf + 1) * step : hi;
    if (nlo == lo && nhi == hi) break;   // the bracket no longer shrinks
    lo = nlo;
    hi = nhi;
  }
  if (j == 0) w[gi] = 0.5 * (lo + hi);
}

// counter-based uniform numbers in [-1, 1) (splitmix64)
__device__ __forceinline__ double urand(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  x ^= x >> 31;
  return (double)(x >> 11) * (2.0 / 9007199254740992.0) - 1.0;
}

// Inverse iteration for the kept eigenvalues, one workgroup of one wave per
// cluster (a run of kept eigenvalues of one sector whose neighbours lie
// within 1e-14 of the sector's largest |eigenvalue|: independent solves for
// two eigenvalues that close could give nearly parallel vectors; most
// clusters hold one eigenvalue, and vectors of eigenvalues further apart
// overlap by at most about 1e-2, which the orthonormalisation removes).  Lane 0 factors T - sigma I = P L U by Gaussian elimination
// with partial pivoting (U has two superdiagonals, L one multiplier per
// column) at the cluster's first eigenvalue sigma, pivots below eps |T|
// raised to that size, and runs the O(n) solves; the 64 lanes share the rest.
// Each vector of the cluster starts random and takes three solves, each
// followed by Gram-Schmidt against the cluster's earlier vectors and unit
// 2-norm, so an exactly degenerate cluster (a null space, a repeated
// eigenvalue) still gets independent vectors.  The right-hand side is scaled
// by the pivot floor before each solve (one tiny pivot then gives an O(1)
// solution); a solution entry above 1e150 (several tiny pivots in a chain)
// rescales what is done.  Scratch and vectors are laid out [row][vector]
// within a sector.
__global__ __launch_bounds__(64) void k_stein(const HeevSec* __restrict__ secs, const int4* __restrict__ clus,
                                              const double* __restrict__ d, const double* __restrict__ e,
                                              const double* __restrict__ w, double* __restrict__ scr,
                                              double* __restrict__ Z) {
  const int4 o = clus[blockIdx.x];   // (sector, first vector, vectors, -)
  const HeevSec S = secs[o.x];
  const int n = S.n, kp = S.keep, q0 = o.y, m = o.z;
  const int lane = threadIdx.x;
  const double* dd = d + S.c;
  const double* ee = e + S.c;
  const int64_t g0 = S.c + n - kp + q0;   // global index of the cluster's first eigenvalue
  const double lam = w[g0];
  double tn = 0.0;
  for (int i = lane; i < n; i += 64)
    tn = fmax(tn, fabs(dd[i]) + (i > 0 ? fabs(ee[i - 1]) : 0.0) + (i < n - 1 ? fabs(ee[i]) : 0.0));
  tn = wmax(tn);
  const double ptol = fmax(DBL_EPSILON * tn, 1e-290);
  const int64_t st = kp, plane = (int64_t)n * kp;
  double* U0 = scr + S.w + q0;
  double* U1 = U0 + plane;
  double* U2 = U1 + plane;
  double* L = U2 + plane;
  double* P = L + plane;
  if (lane == 0) {
    // the working row (w0, w1) at columns (i, i + 1) against row i + 1
    double w0 = dd[0] - lam, w1 = n > 1 ? ee[0] : 0.0;
    for (int i = 0; i < n - 1; ++i) {
      const double sub = ee[i], dn = dd[i + 1] - lam, s2 = i < n - 2 ? ee[i + 1] : 0.0;
      if (fabs(w0) >= fabs(sub)) {
        const double l = w0 != 0.0 ? sub / w0 : 0.0;
        U0[i * st] = w0;
        U1[i * st] = w1;
        U2[i * st] = 0.0;
        L[i * st] = l;
        P[i * st] = 0.0;
        w0 = dn - l * w1;
        w1 = s2;
      } else {
        const double l = w0 / sub;
        U0[i * st] = sub;
        U1[i * st] = dn;
        U2[i * st] = s2;
        L[i * st] = l;
        P[i * st] = 1.0;
        w0 = w1 - l * dn;
        w1 = -l * s2;
      }
    }
    U0[(n - 1) * st] = w0;
    U1[(n - 1) * st] = 0.0;
    U2[(n - 1) * st] = 0.0;
    for (int i = 0; i < n; ++i) {
      const double u = U0[i * st];
      if (fabs(u) < ptol) U0[i * st] = u < 0.0 ? -ptol : ptol;
    }
  }
  for (int j = 0; j < m; ++j) {
    double* y = Z + S.z + q0 + j;
    for (int i = lane; i < n; i += 64) y[i * st] = urand((uint64_t)(g0 + j) * 0x100000001B3ull + (uint64_t)i);
    for (int it = 0; it < 3; ++it) {
      __syncthreads();   // lane 0 reads what every lane wrote
      if (lane == 0) {
        double yw = y[0] * ptol;
        for (int i = 0; i < n - 1; ++i) {
          const double yr = y[(i + 1) * st] * ptol, l = L[i * st];
          if (P[i * st] == 0.0) {
            y[i * st] = yw;
            yw = yr - l * yw;
          } else {
            y[i * st] = yr;
            yw = yw - l * yr;
          }
        }
        y[(n - 1) * st] = yw;
        double x1 = 0.0, x2 = 0.0;
        for (int i = n - 1; i >= 0; --i) {
          double x = (y[i * st] - U1[i * st] * x1 - U2[i * st] * x2) / U0[i * st];
          if (fabs(x) > 1e150) {
            const double f = 1.0 / fabs(x);
            for (int jj = i + 1; jj < n; ++jj) y[jj * st] *= f;
            x *= f;
            x1 *= f;
            x2 *= f;
          }
          y[i * st] = x;
          x2 = x1;
          x1 = x;
        }
      }
      __syncthreads();   // every lane reads what lane 0 wrote
      // Gram-Schmidt against the cluster's earlier vectors (two passes), then
      // unit norm
      for (int pass = 0; pass < (j > 0 ? 2 : 0); ++pass)
        for (int p = 0; p < j; ++p) {
          const double* zp = Z + S.z + q0 + p;
          double c = 0.0;
          for (int i = lane; i < n; i += 64) c += zp[i * st] * y[i * st];
          c = wsum(c);
          for (int i = lane; i < n; i += 64) y[i * st] -= c * zp[i * st];
        }
      double nn = 0.0;
      for (int i = lane; i < n; i += 64) nn += y[i * st] * y[i * st];
      nn = wsum(nn);
      const double sc = nn > 0.0 ? 1.0 / sqrt(nn) : 0.0;
      for (int i = lane; i < n; i += 64) y[i * st] *= sc;
    }
  }
}

// The kept eigenvalues of a sector that lie within 1e-13 of its scale above
// the sector's smallest eigenvalue (the numerical null space of a
// rank-deficient density matrix, kept when the kept count exceeds its rank)
// cannot be told apart by inverse iteration; their kept vectors are any
// orthonormal basis of the complement of the sector's resolved kept vectors,
// which is their eigenspace to that precision.  The first u columns of the
// sector's Z take random numbers here (32 x 32 tiles: sector, row block,
// column block), are projected twice off the resolved columns u .. keep - 1
// (k_xgram, k_xproj) and are made orthonormal with the rest by Newton-Schulz.
__global__ __launch_bounds__(HT) void k_fill(const HeevSec* __restrict__ secs, const int4* __restrict__ tiles,
                                             const int32_t* __restrict__ nu, double* __restrict__ Z) {
  const int4 t = tiles[blockIdx.x];   // (sector, row block, column block, -)
  const HeevSec S = secs[t.x];
  const int u = nu[t.x], kp = S.keep;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  const int b = t.z * GT + tx;
  for (int rr = ty; rr < GT; rr += 8) {
    const int i = t.y * GT + rr;
    if (i < S.n && b < u) Z[S.z + (int64_t)i * kp + b] = urand(((uint64_t)(S.c + i) << 20) ^ (0x5A5Aull + b));
  }
}

// C = Z(:, u:keep)^T Z(:, 0:u) of one sector, one 32 x 32 tile per workgroup
// (sector, resolved-column block, completion-column block); C is keep-u x u
// at the sector's Gram offset.
__global__ __launch_bounds__(HT) void k_xgram(const HeevSec* __restrict__ secs, const int4* __restrict__ tiles,
                                              const int32_t* __restrict__ nu, const double* __restrict__ Z,
                                              double* __restrict__ G) {
  const int4 t = tiles[blockIdx.x];
  const HeevSec S = secs[t.x];
  const int n = S.n, kp = S.keep, u = nu[t.x], nr = kp - u;
  const double* Zs = Z + S.z;
  __shared__ double za[GT][GT + 1], zb[GT][GT + 1];
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  for (int i0 = 0; i0 < n; i0 += GT) {
    for (int rr = ty; rr < GT; rr += 8) {
      const int i = i0 + rr, a = t.y * GT + tx, b = t.z * GT + tx;
      za[rr][tx] = (i < n && a < nr) ? Zs[(int64_t)i * kp + u + a] : 0.0;
      zb[rr][tx] = (i < n && b < u) ? Zs[(int64_t)i * kp + b] : 0.0;
    }
    __syncthreads();
#pragma unroll 8
    for (int rr = 0; rr < GT; ++rr) {
      const double x = za[rr][tx];
#pragma unroll
      for (int j = 0; j < 4; ++j) acc[j] += x * zb[rr][ty + 8 * j];
    }
    __syncthreads();
  }
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int a = t.y * GT + tx, b = t.z * GT + ty + 8 * j;
    if (a < nr && b < u) G[S.g + (int64_t)a * u + b] = acc[j];
  }
}

// Z(:, 0:u) -= Z(:, u:keep) C, one 32-row x 32-column tile per workgroup
// (sector, row block, completion-column block); the tile reads only the
// resolved columns and writes only completion columns.
__global__ __launch_bounds__(HT) void k_xproj(const HeevSec* __restrict__ secs, const int4* __restrict__ tiles,
                                              const int32_t* __restrict__ nu, const double* __restrict__ G,
                                              double* __restrict__ Z) {
  const int4 t = tiles[blockIdx.x];
  const HeevSec S = secs[t.x];
  const int n = S.n, kp = S.keep, u = nu[t.x], nr = kp - u;
  double* Zs = Z + S.z;
  __shared__ double zs[GT][GT + 1], cs[GT][GT + 1];
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  const int b = t.z * GT + tx;
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  for (int a0 = 0; a0 < nr; a0 += GT) {
    for (int rr = ty; rr < GT; rr += 8) {
      const int i = t.y * GT + rr, a = a0 + tx, ca = a0 + rr;
      zs[rr][tx] = (i < n && a < nr) ? Zs[(int64_t)i * kp + u + a] : 0.0;
      cs[rr][tx] = (ca < nr && b < u) ? G[S.g + (int64_t)ca * u + b] : 0.0;
    }
    __syncthreads();
#pragma unroll 8
    for (int aa = 0; aa < GT; ++aa) {
      const double c = cs[aa][tx];
#pragma unroll
      for (int j = 0; j < 4; ++j) acc[j] += zs[ty + 8 * j][aa] * c;
    }
    __syncthreads();
  }
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int i = t.y * GT + ty + 8 * j;
    if (i < n && b < u) Zs[(int64_t)i * kp + b] -= acc[j];
  }
}

// unit 2-norm for the completion columns 0 .. u - 1 of each sector: one
// thread per column (sector, column) of the list
__global__ __launch_bounds__(HT) void k_colnorm(const HeevSec* __restrict__ secs, const int2* __restrict__ cols,
                                                int count, double* __restrict__ Z) {
  const int g = blockIdx.x * HT + threadIdx.x;
  if (g >= count) return;
  const int2 o = cols[g];
  const HeevSec S = secs[o.x];
  double* z = Z + S.z + o.y;
  const int64_t st = S.keep;
  double nn = 0.0;
  for (int i = 0; i < S.n; ++i) nn += z[i * st] * z[i * st];
  const double sc = nn > 0.0 ? 1.0 / sqrt(nn) : 0.0;
  for (int i = 0; i < S.n; ++i) z[i * st] *= sc;
}

// G = Z^T Z of a sector's kept vectors, one 32 x 32 tile per workgroup;
// err[t] = max |G - I| over the tile (NaN if any entry is), rsum[32 t + r] =
// sum of |G(a, b)| over the tile's columns for its row a = 32 ta + r.
__global__ __launch_bounds__(HT) void k_gram(const HeevSec* __restrict__ secs, const int4* __restrict__ tiles,
                                             const double* __restrict__ Z, double* __restrict__ G,
                                             double* __restrict__ err, double* __restrict__ rsum) {
  const int4 t = tiles[blockIdx.x];   // (sector, ta, tb, -)
  const HeevSec S = secs[t.x];
  const int n = S.n, kp = S.keep;
  const double* Zs = Z + S.z;
  __shared__ double za[GT][GT + 1], zb[GT][GT + 1];
  __shared__ double r1[4], rs[8][GT];
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;   // a = ta*32 + tx, b = tb*32 + ty + 8 j
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  for (int i0 = 0; i0 < n; i0 += GT) {
    for (int rr = ty; rr < GT; rr += 8) {
      const int i = i0 + rr, a = t.y * GT + tx, b = t.z * GT + tx;
      za[rr][tx] = (i < n && a < kp) ? Zs[(int64_t)i * kp + a] : 0.0;
      zb[rr][tx] = (i < n && b < kp) ? Zs[(int64_t)i * kp + b] : 0.0;
    }
    __syncthreads();
#pragma unroll 8
    for (int rr = 0; rr < GT; ++rr) {
      const double x = za[rr][tx];
#pragma unroll
      for (int j = 0; j < 4; ++j) acc[j] += x * zb[rr][ty + 8 * j];
    }
    __syncthreads();
  }
  double me = 0.0, ra = 0.0, nan = 0.0;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int a = t.y * GT + tx, b = t.z * GT + ty + 8 * j;
    if (a < kp && b < kp) {
      G[S.g + (int64_t)a * kp + b] = acc[j];
      me = fmax(me, fabs(acc[j] - (a == b ? 1.0 : 0.0)));
      ra += fabs(acc[j]);
      nan += acc[j] - acc[j];   // NaN if any entry is NaN or infinite
    }
  }
  if (nan != nan) me = nan;
  rs[ty][tx] = ra;
  const int wv = threadIdx.x >> 6;
  const double mw = wmax(me);
  const double nw = wsum(nan);
  if ((threadIdx.x & 63) == 0) r1[wv] = nw != nw ? nw : mw;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = fmax(fmax(r1[0], r1[1]), fmax(r1[2], r1[3]));
    for (int w = 0; w < 4; ++w)
      if (r1[w] != r1[w]) m = r1[w];
    err[blockIdx.x] = m;
  }
  if (ty == 0) {
    double r = 0.0;
    for (int y = 0; y < 8; ++y) r += rs[y][tx];
    rsum[(int64_t)blockIdx.x * GT + tx] = r;
  }
}

// One Newton-Schulz step on each unfinished sector: Z' = s Z (3 I - s^2 G) / 2
// with the sector's scale s (so that s Z has singular values in (0, 1]);
// a finished sector (s < 0) is copied.  One 32-row x 32-column tile of Z'
// per workgroup.
__global__ __launch_bounds__(HT) void k_ns(const HeevSec* __restrict__ secs, const int4* __restrict__ tiles,
                                           const double* __restrict__ Z, const double* __restrict__ G,
                                           const double* __restrict__ scl, double* __restrict__ Z2) {
  const int4 t = tiles[blockIdx.x];   // (sector, row block, column block, -)
  const HeevSec S = secs[t.x];
  const int n = S.n, kp = S.keep;
  const double s = scl[t.x];
  const double* Zs = Z + S.z;
  double* Z2s = Z2 + S.z;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;   // column b = tb*32 + tx, rows i = ti*32 + ty + 8 j
  const int b = t.z * GT + tx;
  if (s < 0.0) {
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int i = t.y * GT + ty + 8 * j;
      if (i < n && b < kp) Z2s[(int64_t)i * kp + b] = Zs[(int64_t)i * kp + b];
    }
    return;
  }
  __shared__ double zs[GT][GT + 1], gs[GT][GT + 1];
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  for (int a0 = 0; a0 < kp; a0 += GT) {
    for (int rr = ty; rr < GT; rr += 8) {
      const int i = t.y * GT + rr, a = a0 + tx, ga = a0 + rr;
      zs[rr][tx] = (i < n && a < kp) ? Zs[(int64_t)i * kp + a] : 0.0;
      gs[rr][tx] = (ga < kp && b < kp) ? G[S.g + (int64_t)ga * kp + b] : 0.0;
    }
    __syncthreads();
#pragma unroll 8
    for (int aa = 0; aa < GT; ++aa) {
      const double g = gs[aa][tx];
#pragma unroll
      for (int j = 0; j < 4; ++j) acc[j] += zs[ty + 8 * j][aa] * g;
    }
    __syncthreads();
  }
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int i = t.y * GT + ty + 8 * j;
    if (i < n && b < kp) Z2s[(int64_t)i * kp + b] = s * (1.5 * Zs[(int64_t)i * kp + b] - 0.5 * s * s * acc[j]);
  }
}

// Back-transformation U = Q Z = H_0 H_1 ... H_{n-2} Z in blocks of NB
// reflectors: the block j0 .. j0 + nb - 1 is I - V T V^H (V the block's
// reflectors, T upper triangular: T(i, i) = tau_i, T(0:i, i) = -tau_i
// T(0:i, 0:i) V(:, 0:i)^H v_i), applied from the last block to the first as
// U -= V (T (V^H U)).
constexpr int NB = 32;

// reflector j of a sector (zero at rows <= j; 1 at row j + 1)
__device__ __forceinline__ double2 refl_at(const double2* M, int n, int j, int r) {
  return (j <= n - 2 && r > j) ? M[r + (int64_t)j * n] : make_double2(0.0, 0.0);
}

// U = Z (real, [row][vector]) as complex, column-major n x keep
__global__ __launch_bounds__(HT) void k_z2u(const HeevSec* __restrict__ secs, int nsec, int64_t total,
                                            const int64_t* __restrict__ zoff, const double* __restrict__ Z,
                                            double2* __restrict__ U) {
  for (int64_t x = (int64_t)blockIdx.x * HT + threadIdx.x; x < total; x += (int64_t)gridDim.x * HT) {
    int lo = 0, hi = nsec - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (zoff[mid] <= x) lo = mid;
      else hi = mid - 1;
    }
    const HeevSec S = secs[lo];
    const int64_t e = x - zoff[lo];
    const int q = (int)(e / S.n), r = (int)(e - (int64_t)q * S.n);
    U[S.z + e] = make_double2(Z[S.z + (int64_t)r * S.keep + q], 0.0);
  }
}

// T of every block of every sector: one workgroup per (sector, block)
__global__ __launch_bounds__(HT) void k_wy_t(const HeevSec* __restrict__ secs, const int2* __restrict__ list,
                                             const double2* __restrict__ A, const double2* __restrict__ tau,
                                             double2* __restrict__ T) {
  const int2 o = list[blockIdx.x];   // (sector, block)
  const HeevSec S = secs[o.x];
  const int n = S.n, j0 = NB * o.y, nb = min(NB, n - 1 - j0);
  const double2* M = A + S.a;
  __shared__ double2 sv[NB][NB + 1];   // V^H V
  __shared__ double2 st[NB][NB + 1];   // T
  __shared__ double2 vs[NB][NB + 1];   // (column, row) chunk of V
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;   // S(a = tx, b = ty + 8 j)
  double2 acc[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) acc[j] = make_double2(0.0, 0.0);
  for (int r0 = j0 + 1; r0 < n; r0 += NB) {
    for (int cc = ty; cc < NB; cc += 8) {
      const int r = r0 + tx;
      vs[cc][tx] = r < n ? refl_at(M, n, j0 + cc, r) : make_double2(0.0, 0.0);
    }
    __syncthreads();
#pragma unroll 8
    for (int rr = 0; rr < NB; ++rr) {
      const double2 va = vs[tx][rr];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const double2 t = cjmul(va, vs[ty + 8 * j][rr]);
        acc[j].x += t.x;
        acc[j].y += t.y;
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    sv[tx][ty + 8 * j] = acc[j];
    st[tx][ty + 8 * j] = make_double2(0.0, 0.0);
  }
  __syncthreads();
  for (int i = 0; i < nb; ++i) {
    // T(0:i, i) = -tau_i T(0:i, 0:i) S(0:i, i); one thread per row a < i
    const double2 ti = tau[S.c + j0 + i];
    double2 val = make_double2(0.0, 0.0);
    if (threadIdx.x < i) {
      const int a = threadIdx.x;
      for (int b = a; b < i; ++b) {
        const double2 t = cmul(st[a][b], sv[b][i]);
        val.x += t.x;
        val.y += t.y;
      }
      val = cmul(make_double2(-ti.x, -ti.y), val);
    }
    __syncthreads();
    if (threadIdx.x < i) st[threadIdx.x][i] = val;
    if (threadIdx.x == 0) st[i][i] = ti;
    __syncthreads();
  }
  double2* To = T + S.cb + (int64_t)o.y * NB * NB;
  for (int e = threadIdx.x; e < NB * NB; e += HT) To[e] = st[e % NB][e / NB];   // column-major
}

// W = V^H U for block m in 256-row pieces: one workgroup per (sector,
// 32-column tile of U, row block); the pieces are summed in row-block order
// by k_wy_u (a row block with no row below the block's reflectors writes
// zeros)
__global__ __launch_bounds__(HT) void k_wy_w(int m, const HeevSec* __restrict__ secs, const int4* __restrict__ list,
                                             const double2* __restrict__ A, const double2* __restrict__ U,
                                             double2* __restrict__ W) {
  const int4 o = list[blockIdx.x];   // (sector, column tile, row block, -)
  const HeevSec S = secs[o.x];
  const int n = S.n, kp = S.keep, j0 = NB * m;
  if (j0 > n - 2) return;
  const int rlo = max(j0 + 1, o.z * HT), rhi = min(n, (o.z + 1) * HT);
  const double2* M = A + S.a;
  const double2* Us = U + S.z;
  __shared__ double2 vs[NB][NB + 1], us[NB][NB + 1];
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;   // W(a = tx, col = ty + 8 j)
  double2 acc[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) acc[j] = make_double2(0.0, 0.0);
  for (int r0 = rlo; r0 < rhi; r0 += NB) {
    for (int cc = ty; cc < NB; cc += 8) {
      const int r = r0 + tx, col = o.y * NB + cc;
      vs[cc][tx] = r < rhi ? refl_at(M, n, j0 + cc, r) : make_double2(0.0, 0.0);
      us[cc][tx] = (r < rhi && col < kp) ? Us[r + (int64_t)col * n] : make_double2(0.0, 0.0);
    }
    __syncthreads();
#pragma unroll 8
    for (int rr = 0; rr < NB; ++rr) {
      const double2 v = vs[tx][rr];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const double2 t = cjmul(v, us[ty + 8 * j][rr]);
        acc[j].x += t.x;
        acc[j].y += t.y;
      }
    }
    __syncthreads();
  }
  double2* Ws = W + S.wb;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int col = o.y * NB + ty + 8 * j;
    if (col < kp) Ws[((int64_t)o.z * kp + col) * NB + tx] = acc[j];
  }
}

// U(rows, cols) -= V (T W) for block m: one workgroup per (sector, column
// tile, 256-row block); T W of the tile formed in LDS
__global__ __launch_bounds__(HT) void k_wy_u(int m, const HeevSec* __restrict__ secs, const int4* __restrict__ list,
                                             const double2* __restrict__ A, const double2* __restrict__ T,
                                             const double2* __restrict__ W, double2* __restrict__ U) {
  const int4 o = list[blockIdx.x];   // (sector, column tile, row block, -)
  const HeevSec S = secs[o.x];
  const int n = S.n, kp = S.keep, j0 = NB * m;
  if (j0 > n - 2) return;
  const int rlo = max(j0 + 1, o.z * HT), rhi = min(n, (o.z + 1) * HT);
  if (rlo >= rhi) return;
  const double2* M = A + S.a;
  const double2* Tm = T + S.cb + (int64_t)m * NB * NB;
  const double2* Ws = W + S.wb;
  __shared__ double2 tw[NB][NB + 1];   // (T W)(a, col)
  __shared__ double2 ws[NB][NB + 1];
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  const int nrb = (n + HT - 1) / HT;
  for (int e = threadIdx.x; e < NB * NB; e += HT) {
    const int a = e % NB, c = e / NB, col = o.y * NB + c;
    double2 sm = make_double2(0.0, 0.0);
    if (col < kp)
      for (int rb = 0; rb < nrb; ++rb) {
        const double2 x = Ws[((int64_t)rb * kp + col) * NB + a];
        sm.x += x.x;
        sm.y += x.y;
      }
    ws[a][c] = sm;
  }
  __syncthreads();
  for (int e = threadIdx.x; e < NB * NB; e += HT) {
    const int a = e % NB, c = e / NB;
    double2 sm = make_double2(0.0, 0.0);
    for (int b = a; b < NB; ++b) {
      const double2 t = cmul(Tm[a + (int64_t)b * NB], ws[b][c]);
      sm.x += t.x;
      sm.y += t.y;
    }
    tw[a][c] = sm;
  }
  __syncthreads();
  // thread: row r = rlo + threadIdx.x, all 32 columns of the tile
  const int r = rlo + threadIdx.x;
  if (r < rhi) {
    double2 v[NB];
#pragma unroll
    for (int a = 0; a < NB; ++a) v[a] = refl_at(M, n, j0 + a, r);
    double2* Us = U + S.z;
    for (int c = 0; c < NB; ++c) {
      const int col = o.y * NB + c;
      if (col >= kp) break;
      double2 sm = make_double2(0.0, 0.0);
#pragma unroll
      for (int a = 0; a < NB; ++a) {
        const double2 t = cmul(v[a], tw[a][c]);
        sm.x += t.x;
        sm.y += t.y;
      }
      double2* up = Us + r + (int64_t)col * n;
      *up = csub(*up, sm);
    }
  }
  (void)tx;
  (void)ty;
}

// every block: the kept vectors as its last keep columns, zero elsewhere
__global__ __launch_bounds__(HT) void k_place(const HeevSec* __restrict__ secs, int nsec, int64_t total,
                                              double2* __restrict__ A, const double2* __restrict__ U) {
  for (int64_t x = (int64_t)blockIdx.x * HT + threadIdx.x; x < total; x += (int64_t)gridDim.x * HT) {
    int lo = 0, hi = nsec - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (secs[mid].a <= x) lo = mid;
      else hi = mid - 1;
    }
    const HeevSec S = secs[lo];
    const int64_t e = x - S.a;
    const int n = S.n;
    const int c = (int)(e / n), r = (int)(e - (int64_t)c * n);
    const int c0 = n - S.keep;
    A[x] = c >= c0 ? U[S.z + (int64_t)(c - c0) * n + r] : make_double2(0.0, 0.0);
  }
}

double now_ms() {
  return 1e3 * std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

template <typename T>
void upload(HeevBuf& b, const std::vector<T>& h, hipStream_t s) {
  b.ensure(sizeof(T) * std::max<size_t>(h.size(), 1));
  if (!h.empty())
    check_hip(hipMemcpyAsync(b.p, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice, s), "eigensolver tables");
}

}  // namespace

void HeevWork::values(hipStream_t s, double2* A, const std::vector<int64_t>& off, const std::vector<int>& dim,
                      double* evals) {
  const int ns = (int)dim.size();
  sec_.assign(ns, HeevSec{});
  nblk_ = 0;
  if (ns == 0) return;
  off_ = off;
  dim_ = dim;
  h_wgr_.clear();
  h_tile_.clear();
  h_owner_.clear();
  int64_t col = 0, blk = 0, cbo = 0;
  int ch = 0;
  nmax_ = 0;
  h_rg_.clear();
  for (int i = 0; i < ns; ++i) {
    const int n = dim[i];
    if (n < 1 || n > HEEV_MAXN) throw std::runtime_error("eigensolver: sector dimension out of range");
    if (off[i] != blk) throw std::runtime_error("eigensolver: sector blocks must be contiguous");
    HeevSec& S = sec_[i];
    S.a = off[i];
    S.c = col;
    S.n = n;
    S.ch0 = ch;
    S.nch = (n + CW - 1) / CW;
    for (int q = 0; q < S.nch; ++q) h_wgr_.push_back(HeevWg{S.a, S.c, n, S.ch0, S.nch, q});
    const int T = (n + 31) / 32;
    for (int bi = 0; bi < T; ++bi)
      for (int bj = 0; bj <= bi; ++bj) {
        h_tile_.push_back(i);
        h_tile_.push_back(bi);
        h_tile_.push_back(bj);
        h_tile_.push_back(0);
      }
    for (int j = 0; j < n; ++j) h_owner_.push_back(i);
    S.cb = cbo;   // T factors of the sector's blocks of NB reflectors
    const int nbk = n >= 2 ? (n - 2) / NB + 1 : 0;
    for (int m = 0; m < nbk; ++m) {
      h_rg_.push_back(i);
      h_rg_.push_back(m);
    }
    cbo += (int64_t)nbk * NB * NB;
    col += n;
    ch += S.nch;
    blk += (int64_t)n * n;
    nmax_ = std::max(nmax_, n);
  }
  ncol_ = col;
  nblk_ = blk;
  npart_ = ch;
  nwg_ = (int)h_wgr_.size();
  ntile_ = (int)(h_tile_.size() / 4);
  upload(secs_, sec_, s);
  upload(wgs_, h_wgr_, s);
  upload(tiles_, h_tile_, s);
  upload(owner_, h_owner_, s);
  upload(rg_, h_rg_, s);
  cg_.ensure(sizeof(double2) * (size_t)std::max<int64_t>(cbo, 1));
  d_.ensure(sizeof(double) * ncol_);
  e_.ensure(sizeof(double) * ncol_);
  tau_.ensure(sizeof(double2) * ncol_);
  vb_.ensure(sizeof(double2) * 2 * ncol_);
  pb_.ensure(sizeof(double2) * 2 * ncol_);
  part_.ensure(sizeof(double2) * 2 * (size_t)npart_);
  const HeevSec* S = secs_.as<HeevSec>();
  double t0 = 0.0;
  if (timing) {
    check_hip(hipStreamSynchronize(s), "sync");
    t0 = now_ms();
  }
  hipLaunchKernelGGL(k_herm_fix, dim3(ntile_), dim3(HT), 0, s, S, tiles_.as<int4>(), A);
  for (int k = 0; k < nmax_; ++k)
    hipLaunchKernelGGL(k_hetrd_step, dim3(nwg_), dim3(HT), 0, s, k, wgs_.as<HeevWg>(), A, d_.as<double>(),
                       e_.as<double>(), tau_.as<double2>(), vb_.as<double2>(), pb_.as<double2>(),
                       part_.as<double2>(), ncol_, (int64_t)npart_);
  if (!h_rg_.empty())
    hipLaunchKernelGGL(k_wy_t, dim3((unsigned)(h_rg_.size() / 2)), dim3(HT), 0, s, S, rg_.as<int2>(),
                       (const double2*)A, (const double2*)tau_.as<double2>(), cg_.as<double2>());
  if (timing) {
    check_hip(hipStreamSynchronize(s), "sync");
    ms[0] = now_ms() - t0;
    t0 = now_ms();
  }
  hipLaunchKernelGGL(k_bisect, dim3((unsigned)((ncol_ * MS + HT - 1) / HT)), dim3(HT), 0, s, S,
                     owner_.as<int32_t>(), (int)ncol_, d_.as<double>(), e_.as<double>(), evals);
  check_hip(hipGetLastError(), "eigensolver launch");
  if (timing) {
    check_hip(hipStreamSynchronize(s), "sync");
    ms[1] = now_ms() - t0;
  }
}

bool HeevWork::vectors(hipStream_t s, double2* A, const std::vector<int>& keep, const double* evals) {
  const int ns = (int)sec_.size();
  if ((int)keep.size() != ns) throw std::runtime_error("eigensolver: keep has the wrong length");
  h_vown_.clear();
  h_gtile_.clear();
  max_cluster = 0;
  unresolved = 0;
  fail_info.clear();
  h_ztile_.clear();
  h_wyw_.clear();
  h_wyu_.clear();
  h_zoff_.assign(ns, 0);
  int64_t z = 0, w = 0, g = 0, wb = 0;
  int mblk = 0;
  for (int i = 0; i < ns; ++i) {
    HeevSec& S = sec_[i];
    const int n = S.n, kp = keep[i];
    if (kp < 0 || kp > n) throw std::runtime_error("eigensolver: kept count out of range");
    S.keep = kp;
    S.z = z;
    S.w = w;
    S.g = g;
    S.wb = wb;   // NB x keep partial products per 256-row block
    h_zoff_[i] = z;
    z += (int64_t)n * kp;
    w += 5 * (int64_t)n * kp;
    g += (int64_t)kp * kp;
    wb += (int64_t)NB * kp * ((n + HT - 1) / HT);
    for (int q = 0; q < kp; ++q) {
      h_vown_.push_back(i);
      h_vown_.push_back(q);
    }
    if (kp > 0 && n >= 2) {
      mblk = std::max(mblk, (n - 2) / NB + 1);
      const int Tc = (kp + NB - 1) / NB, Tr = (n + HT - 1) / HT;
      for (int tc = 0; tc < Tc; ++tc) {
        h_wyw_.push_back(i);
        h_wyw_.push_back(tc);
        for (int tr = 0; tr < Tr; ++tr) {
          h_wyu_.push_back(i);
          h_wyu_.push_back(tc);
          h_wyu_.push_back(tr);
          h_wyu_.push_back(0);
        }
      }
    }
    const int Tk = (kp + GT - 1) / GT, Tn = (n + GT - 1) / GT;
    for (int ta = 0; ta < Tk; ++ta)
      for (int tb = 0; tb < Tk; ++tb) {
        h_gtile_.push_back(i);
        h_gtile_.push_back(ta);
        h_gtile_.push_back(tb);
        h_gtile_.push_back(0);
      }
    for (int ti = 0; ti < Tn; ++ti)
      for (int tb = 0; tb < Tk; ++tb) {
        h_ztile_.push_back(i);
        h_ztile_.push_back(ti);
        h_ztile_.push_back(tb);
        h_ztile_.push_back(0);
      }
  }
  const int nv = (int)(h_vown_.size() / 2);
  // clusters of kept eigenvalues (host copy of the eigenvalues)
  std::vector<double> ev((size_t)std::max<int64_t>(ncol_, 1));
  if (nv > 0) {
    check_hip(hipMemcpyAsync(ev.data(), evals, sizeof(double) * ncol_, hipMemcpyDeviceToHost, s), "d2h");
    check_hip(hipStreamSynchronize(s), "eigenvalues");
  }
  h_clus_.clear();
  h_nu_.assign(ns, 0);
  h_ftile_.clear();
  h_xtile_.clear();
  h_ucol_.clear();
  for (int i = 0; i < ns && nv > 0; ++i) {
    const HeevSec& S = sec_[i];
    const int n = S.n, kp = S.keep;
    if (kp == 0) continue;
    const double* lam = ev.data() + S.c;
    const double scale = std::max(std::fabs(lam[0]), std::fabs(lam[n - 1]));
    const double ctol = 1e-14 * scale;
    // unresolved: the lowest kept eigenvalues within 1e-13 of the scale above
    // the sector's smallest one (all discarded ones lie below them)
    int u = 0;
    while (u < kp && lam[n - kp + u] <= lam[0] + 1e-13 * scale) ++u;
    h_nu_[i] = u;
    if (u > 0) {
      const int Tn = (n + GT - 1) / GT, Tu = (u + GT - 1) / GT, Tr = (kp - u + GT - 1) / GT;
      for (int ti = 0; ti < Tn; ++ti)
        for (int tb = 0; tb < Tu; ++tb) {
          h_ftile_.push_back(i);
          h_ftile_.push_back(ti);
          h_ftile_.push_back(tb);
          h_ftile_.push_back(0);
        }
      for (int ta = 0; ta < Tr; ++ta)
        for (int tb = 0; tb < Tu; ++tb) {
          h_xtile_.push_back(i);
          h_xtile_.push_back(ta);
          h_xtile_.push_back(tb);
          h_xtile_.push_back(0);
        }
      unresolved += u;
      for (int b = 0; b < u; ++b) {
        h_ucol_.push_back(i);
        h_ucol_.push_back(b);
      }
    }
    int q = u;
    while (q < kp) {
      int mm = 1;
      while (q + mm < kp && lam[n - kp + q + mm] - lam[n - kp + q + mm - 1] <= ctol) ++mm;
      h_clus_.push_back(i);
      h_clus_.push_back(q);
      h_clus_.push_back(mm);
      h_clus_.push_back(0);
      max_cluster = std::max(max_cluster, mm);
      q += mm;
    }
  }
  const int nclus = (int)(h_clus_.size() / 4);
  const int ngt = (int)(h_gtile_.size() / 4), nzt = (int)(h_ztile_.size() / 4);

  upload(secs_, sec_, s);
  upload(vown_, h_clus_, s);
  upload(nu_, h_nu_, s);
  upload(ftile_, h_ftile_, s);
  upload(xtile_, h_xtile_, s);
  upload(ucol_, h_ucol_, s);
  const int nft = (int)(h_ftile_.size() / 4), nxt = (int)(h_xtile_.size() / 4);
  upload(gtile_, h_gtile_, s);
  upload(ztile_, h_ztile_, s);
  upload(bt_, h_wyw_, s);
  upload(wyu_, h_wyu_, s);
  upload(zoff_, h_zoff_, s);
  const HeevSec* S = secs_.as<HeevSec>();
  U_.ensure(sizeof(double2) * std::max<int64_t>(z, 1));
  ns_iterations = 0;
  ortho_error = 0.0;
  ms[2] = ms[3] = ms[4] = 0.0;
  double t0_bt = 0.0;
  if (nv > 0) {
    scratch_.ensure(sizeof(double) * std::max<int64_t>(w, 1));
    Z_.ensure(sizeof(double) * std::max<int64_t>(z, 1));
    Z2_.ensure(sizeof(double) * std::max<int64_t>(z, 1));
    G_.ensure(sizeof(double) * std::max<int64_t>(g, 1));
    gerr_.ensure(sizeof(double) * (size_t)std::max(ngt, 1) * (1 + GT));
    HeevBuf* zc = &Z_;
    HeevBuf* zn = &Z2_;
    double t0 = 0.0;
    if (timing) t0 = now_ms();
    if (nclus)
      hipLaunchKernelGGL(k_stein, dim3(nclus), dim3(64), 0, s, S, vown_.as<int4>(), d_.as<double>(), e_.as<double>(),
                         evals, scratch_.as<double>(), zc->as<double>());
    if (nft) {
      hipLaunchKernelGGL(k_fill, dim3(nft), dim3(HT), 0, s, S, ftile_.as<int4>(), nu_.as<int32_t>(), zc->as<double>());
      for (int pass = 0; pass < 2 && nxt; ++pass) {
        hipLaunchKernelGGL(k_xgram, dim3(nxt), dim3(HT), 0, s, S, xtile_.as<int4>(), nu_.as<int32_t>(),
                           (const double*)zc->as<double>(), G_.as<double>());
        hipLaunchKernelGGL(k_xproj, dim3(nft), dim3(HT), 0, s, S, ftile_.as<int4>(), nu_.as<int32_t>(),
                           (const double*)G_.as<double>(), zc->as<double>());
      }
      const int nuc = (int)(h_ucol_.size() / 2);
      hipLaunchKernelGGL(k_colnorm, dim3((nuc + HT - 1) / HT), dim3(HT), 0, s, S, ucol_.as<int2>(), nuc,
                         zc->as<double>());
    }
    if (timing) {
      check_hip(hipStreamSynchronize(s), "sync");
      ms[2] = now_ms() - t0;
      t0 = now_ms();
    }
    std::vector<double> err((size_t)ngt * (1 + GT)), scl(ns, -1.0);
    std::vector<std::vector<double>> rows(ns);
    scl_.ensure(sizeof(double) * ns);
    constexpr double TOL = 1e-12;
    constexpr int MAXIT = 100;
    for (int it = 0;; ++it) {
      hipLaunchKernelGGL(k_gram, dim3(ngt), dim3(HT), 0, s, S, gtile_.as<int4>(), zc->as<double>(), G_.as<double>(),
                         gerr_.as<double>(), gerr_.as<double>() + ngt);
      check_hip(hipMemcpyAsync(err.data(), gerr_.p, sizeof(double) * err.size(), hipMemcpyDeviceToHost, s), "d2h");
      check_hip(hipStreamSynchronize(s), "eigensolver orthonormalisation");
      std::vector<double> emax(ns, 0.0), rmax(ns, 0.0);
      for (int i = 0; i < ns; ++i) rows[i].assign(keep[i], 0.0);
      for (int t = 0; t < ngt; ++t) {
        const int i = h_gtile_[4 * t], ta = h_gtile_[4 * t + 1];
        if (!(err[t] <= emax[i])) emax[i] = err[t];   // NaN wins
        for (int r = 0; r < GT && ta * GT + r < keep[i]; ++r) rows[i][ta * GT + r] += err[ngt + (size_t)t * GT + r];
      }
      for (int i = 0; i < ns; ++i)
        for (double x : rows[i]) rmax[i] = std::max(rmax[i], x);
      bool all = true;
      ortho_error = 0.0;
      for (int i = 0; i < ns; ++i) {
        if (keep[i] == 0) continue;
        if (!(emax[i] == emax[i]) || !std::isfinite(rmax[i])) {
          fail_info = "non-finite Gram matrix in sector " + std::to_string(i) + " (n " + std::to_string(sec_[i].n) +
                      ", kept " + std::to_string(keep[i]) + ") at orthonormalisation step " + std::to_string(it);
          return false;
        }
        ortho_error = std::max(ortho_error, emax[i]);
        if (emax[i] <= TOL) {
          scl[i] = -1.0;
        } else {
          all = false;
          // the step converges while the singular values of s Z stay below
          // sqrt(3); |G|_2 is at most the largest absolute row sum of G
          scl[i] = rmax[i] <= 2.5 ? 1.0 : 1.0 / std::sqrt(rmax[i]);
        }
      }
      ns_iterations = it;
      if (all) break;
      if (it >= MAXIT) {
        int worst = 0;
        for (int i = 0; i < ns; ++i)
          if (keep[i] && emax[i] > emax[worst]) worst = i;
        fail_info = "max |Z^T Z - I| = " + std::to_string(emax[worst]) + " after " + std::to_string(MAXIT) +
                    " orthonormalisation steps in sector " + std::to_string(worst) + " (n " +
                    std::to_string(sec_[worst].n) + ", kept " + std::to_string(keep[worst]) + ")";
        return false;
      }
      check_hip(hipMemcpyAsync(scl_.p, scl.data(), sizeof(double) * ns, hipMemcpyHostToDevice, s), "h2d");
      check_hip(hipStreamSynchronize(s), "h2d");
      hipLaunchKernelGGL(k_ns, dim3(nzt), dim3(HT), 0, s, S, ztile_.as<int4>(), zc->as<double>(), G_.as<double>(),
                         scl_.as<double>(), zn->as<double>());
      std::swap(zc, zn);
    }
    if (timing) {
      ms[3] = now_ms() - t0;
      t0_bt = now_ms();
    }
    // U = Q Z by blocks of NB reflectors, last block first
    Wb_.ensure(sizeof(double2) * (size_t)std::max<int64_t>(wb, 1));
    const unsigned zg = (unsigned)std::min<int64_t>((z + HT - 1) / HT, 16384);
    hipLaunchKernelGGL(k_z2u, dim3(std::max(zg, 1u)), dim3(HT), 0, s, S, ns, z, (const int64_t*)zoff_.as<int64_t>(),
                       (const double*)zc->as<double>(), U_.as<double2>());
    const int nww = (int)(h_wyw_.size() / 2), nwu = (int)(h_wyu_.size() / 4);
    for (int m = mblk - 1; m >= 0 && nww > 0; --m) {
      hipLaunchKernelGGL(k_wy_w, dim3(nwu), dim3(HT), 0, s, m, S, wyu_.as<int4>(), (const double2*)A,
                         (const double2*)U_.as<double2>(), Wb_.as<double2>());
      hipLaunchKernelGGL(k_wy_u, dim3(nwu), dim3(HT), 0, s, m, S, wyu_.as<int4>(), (const double2*)A,
                         (const double2*)cg_.as<double2>(), (const double2*)Wb_.as<double2>(), U_.as<double2>());
    }
  }
  const unsigned grid = (unsigned)std::min<int64_t>((nblk_ + HT - 1) / HT, 16384);
  hipLaunchKernelGGL(k_place, dim3(std::max(grid, 1u)), dim3(HT), 0, s, S, ns, nblk_, A,
                     (const double2*)U_.as<double2>());
  check_hip(hipGetLastError(), "eigensolver launch");
  check_hip(hipStreamSynchronize(s), "eigensolver");
  if (timing && nv > 0) ms[4] = now_ms() - t0_bt;
  return true;
}

}  // namespace hh
