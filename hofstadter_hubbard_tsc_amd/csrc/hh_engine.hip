// Grouped complex GEMM over tile lists and Lanczos vector kernels for gfx950.
//
// One workgroup (4 wave64) owns a 64x64 macro tile of one output region and
// walks every term that lands on it (output-stationary, no atomics, fixed
// summation order => deterministic).  Complex products run as four real
// 16x16x4 MFMAs per k-step on LDS-staged operand tiles; the term coefficient
// is folded into the left operand while it is staged.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <stdexcept>

#include "hh_engine.hpp"

namespace hh {

void check_hip(hipError_t e, const char* what) {
  if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

// ----------------------------------------------------------------------------
// task upload
// ----------------------------------------------------------------------------
static constexpr int BM = 64, BN = 64, KC = 16, LPAD = 16;

void DevTasks::release() {
  if (tiles) (void)hipFree(tiles);
  if (terms) (void)hipFree(terms);
  if (work) (void)hipFree(work);
  tiles = nullptr;
  terms = nullptr;
  work = nullptr;
  ntiles = nterms = nwork = 0;
}

void DevTasks::upload(const TaskList& tl, hipStream_t s) {
  release();
  ntiles = (int)tl.tiles.size();
  nterms = (int)tl.terms.size();
  flops_alg = tl.flops_alg;
  flops_exec = tl.flops_exec;
  std::vector<WorkItem> w;
  // largest tiles first: long work items start early (simple LPT ordering)
  std::vector<std::pair<double, int>> cost;
  for (int i = 0; i < ntiles; ++i) {
    const Tile& t = tl.tiles[i];
    double c = 1.0;
    for (int k = t.t0; k < t.t1; ++k) c += (double)std::max(tl.terms[k].K, 1);
    cost.push_back({-c * std::min(t.M, BM) * std::min(t.N, BN), i});
  }
  std::stable_sort(cost.begin(), cost.end());
  for (auto& ci : cost) {
    const Tile& t = tl.tiles[ci.second];
    for (int r = 0; r < t.M; r += BM)
      for (int c = 0; c < t.N; c += BN) w.push_back({ci.second, r, c, 0});
  }
  nwork = (int)w.size();
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
  check_hip(hipStreamSynchronize(s), "task upload");
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

__device__ __forceinline__ double2 ld_op(const double2* __restrict__ p, int64_t idx, bool conj) {
  double2 v = p[idx];
  if (conj) v.y = -v.y;
  return v;
}

// ----------------------------------------------------------------------------
// grouped tile kernel
// ----------------------------------------------------------------------------
template <typename T, int TAG>
__global__ __launch_bounds__(256) void k_tiles(const Tile* __restrict__ tiles, const Term* __restrict__ terms,
                                               const WorkItem* __restrict__ work, BufPtrs bufs) {
  typedef typename Acc<T>::V V;
  __shared__ T Lr[KC][BM + LPAD], Li[KC][BM + LPAD];
  __shared__ T Rr[KC][BN + LPAD], Ri[KC][BN + LPAD];

  const WorkItem w = work[blockIdx.x];
  const Tile tile = tiles[w.tile];
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32;
  const int r0 = w.r0, c0 = w.c0;
  const int Mv = min(BM, tile.M - r0), Nv = min(BN, tile.N - c0);
  const int li = lane & 15, lk = lane >> 4;

  V accR[2][2], accI[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      accR[a][b] = V{0, 0, 0, 0};
      accI[a][b] = V{0, 0, 0, 0};
    }
  // elementwise (K == 0) contributions accumulate in double
  double exR[2][2][4], exI[2][2][4];
  bool has_ex = false;
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int v = 0; v < 4; ++v) exR[a][b][v] = exI[a][b][v] = 0.0;

  for (int ti = tile.t0; ti < tile.t1; ++ti) {
    const Term tm = terms[ti];
    const double2* __restrict__ A = bufs.p[tm.a_buf];
    if (tm.K == 0) {
      has_ex = true;
      const bool cj = tm.a_op == OP_R || tm.a_op == OP_C;
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            const int r = wr + a * 16 + Acc<T>::row(lk, v), c = wc + b * 16 + li;
            if (r < Mv && c < Nv) {
              double2 x = ld_op(A, tm.a_off + (int64_t)(r0 + r) + (int64_t)(c0 + c) * tm.a_ld, cj);
              exR[a][b][v] += tm.cr * x.x - tm.ci * x.y;
              exI[a][b][v] += tm.cr * x.y + tm.ci * x.x;
            }
          }
      continue;
    }
    const double2* __restrict__ B = bufs.p[tm.b_buf];
    const bool a_rowfast = (tm.a_op == OP_NN || tm.a_op == OP_R);
    const bool a_conj = (tm.a_op == OP_C || tm.a_op == OP_R);
    const bool b_kfast = (tm.b_op == OP_NN || tm.b_op == OP_R);
    const bool b_conj = (tm.b_op == OP_C || tm.b_op == OP_R);
    for (int kb = 0; kb < tm.K; kb += KC) {
      // stage op(A) * coef  (BM x KC)
#pragma unroll
      for (int j = 0; j < (BM * KC) / 256; ++j) {
        const int idx = tid + 256 * j;
        int r, k;
        if (a_rowfast) { r = idx % BM; k = idx / BM; }
        else { k = idx % KC; r = idx / KC; }
        double2 x = make_double2(0.0, 0.0);
        if (r < Mv && kb + k < tm.K) {
          const int64_t o = a_rowfast ? tm.a_off + (int64_t)(r0 + r) + (int64_t)(kb + k) * tm.a_ld
                                      : tm.a_off + (int64_t)(kb + k) + (int64_t)(r0 + r) * tm.a_ld;
          x = ld_op(A, o, a_conj);
        }
        Lr[k][r] = (T)(tm.cr * x.x - tm.ci * x.y);
        Li[k][r] = (T)(tm.cr * x.y + tm.ci * x.x);
      }
      // stage op(B)  (KC x BN)
#pragma unroll
      for (int j = 0; j < (BN * KC) / 256; ++j) {
        const int idx = tid + 256 * j;
        int k, c;
        if (b_kfast) { k = idx % KC; c = idx / KC; }
        else { c = idx % BN; k = idx / BN; }
        double2 x = make_double2(0.0, 0.0);
        if (c < Nv && kb + k < tm.K) {
          const int64_t o = b_kfast ? tm.b_off + (int64_t)(kb + k) + (int64_t)(c0 + c) * tm.b_ld
                                    : tm.b_off + (int64_t)(c0 + c) + (int64_t)(kb + k) * tm.b_ld;
          x = ld_op(B, o, b_conj);
        }
        Rr[k][c] = (T)x.x;
        Ri[k][c] = (T)x.y;
      }
      __syncthreads();
#pragma unroll
      for (int kk = 0; kk < KC; kk += 4) {
        T ar[2], ai[2], br[2], bi[2];
#pragma unroll
        for (int a = 0; a < 2; ++a) {
          ar[a] = Lr[kk + lk][wr + a * 16 + li];
          ai[a] = Li[kk + lk][wr + a * 16 + li];
        }
#pragma unroll
        for (int b = 0; b < 2; ++b) {
          br[b] = Rr[kk + lk][wc + b * 16 + li];
          bi[b] = Ri[kk + lk][wc + b * 16 + li];
        }
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
          for (int b = 0; b < 2; ++b) {
            accR[a][b] = Acc<T>::mfma(ar[a], br[b], accR[a][b]);
            accR[a][b] = Acc<T>::mfma(-ai[a], bi[b], accR[a][b]);
            accI[a][b] = Acc<T>::mfma(ar[a], bi[b], accI[a][b]);
            accI[a][b] = Acc<T>::mfma(ai[a], br[b], accI[a][b]);
          }
      }
      __syncthreads();
    }
  }
  // write the macro tile once
  double2* __restrict__ C = bufs.p[tile.c_buf];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int r = wr + a * 16 + Acc<T>::row(lk, v), c = wc + b * 16 + li;
        if (r < Mv && c < Nv) {
          double re = (double)accR[a][b][v], im = (double)accI[a][b][v];
          if (has_ex) {
            re += exR[a][b][v];
            im += exI[a][b][v];
          }
          const int64_t o = tile.c_off + (int64_t)(r0 + r) + (int64_t)(c0 + c) * tile.c_ld;
          if (tile.accumulate) {
            double2 old = C[o];
            re += old.x;
            im += old.y;
          }
          C[o] = make_double2(re, im);
        }
      }
}

template <typename T>
static void launch_t(const DevTasks& t, const BufPtrs& b, hipStream_t s, int tag) {
  const dim3 g(t.nwork), blk(256);
  if (tag == 1) hipLaunchKernelGGL((k_tiles<T, 1>), g, blk, 0, s, t.tiles, t.terms, t.work, b);
  else if (tag == 2) hipLaunchKernelGGL((k_tiles<T, 2>), g, blk, 0, s, t.tiles, t.terms, t.work, b);
  else hipLaunchKernelGGL((k_tiles<T, 0>), g, blk, 0, s, t.tiles, t.terms, t.work, b);
}

void launch_tiles(const DevTasks& t, const BufPtrs& b, hipStream_t s, int precision, int tag) {
  if (t.nwork == 0) return;
  if (precision == PREC_FP32) launch_t<float>(t, b, s, tag);
  else launch_t<double>(t, b, s, tag);
  check_hip(hipGetLastError(), "k_tiles launch");
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
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
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
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
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
void dev_start_vector(double2* p, int64_t n, uint64_t seed, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_start, dim3(grid_for(n)), dim3(VB), 0, s, p, n, (seed & 0xFFFFFFFFull) << 32);
}

}  // namespace hh
