// Dense LU solver of the iBEM system on gfx950 (replaces Eigen FullPivLU, systemSolverElastic3D.h:55).
//
// Blocked right-looking LU with partial (row) pivoting, column-major, in place:
//   for each 256-wide panel: recursive panel factorisation (leaves of 32 columns factored by one
//   cooperative kernel whose workgroups split the panel rows and meet at grid barriers for the pivot
//   search and row swap), row interchanges on the other columns, recursive unit-lower TRSM, and the
//   trailing-matrix update C -= L21 U12 on MFMA (v_mfma_f64_16x16x4f64, or f32 for mixed precision).
// The trailing update holds 2n^3/3 of the flops; everything else is memory- or latency-bound.
#include <hip/hip_cooperative_groups.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>

#include "ibem3d_kernels.h"

namespace cg = cooperative_groups;

namespace ibem {

// ================================================================================ MFMA GEMM ====
typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

__device__ inline v4d mfma16(double a, double b, v4d c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
__device__ inline v4f mfma16(float a, float b, v4f c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

// Accumulator layout of the 16x16x4 MFMAs (measured on gfx950): register v of lane l holds
//   f64: D[(l / 16) + 4 v][l % 16]        f32: D[4 (l / 16) + v][l % 16]
template <typename T>
struct Vec4;
template <>
struct Vec4<double> {
  typedef v4d type;
  static __device__ inline int row(int lk, int v) { return lk + 4 * v; }
};
template <>
struct Vec4<float> {
  typedef v4f type;
  static __device__ inline int row(int lk, int v) { return 4 * lk + v; }
};

constexpr int kBM = 128, kBN = 128;

template <typename T>
struct GemmCfg;
#ifndef IBEM_GEMM_BK_F64
#define IBEM_GEMM_BK_F64 16
#endif
#ifndef IBEM_GEMM_BK_F32
#define IBEM_GEMM_BK_F32 32
#endif
#ifndef IBEM_GEMM_MINB
#define IBEM_GEMM_MINB 2
#endif
template <>
struct GemmCfg<double> {
  static constexpr int BK = IBEM_GEMM_BK_F64;
};
template <>
struct GemmCfg<float> {
  static constexpr int BK = IBEM_GEMM_BK_F32;
};

// C[M x N] -= A[M x K] B[K x N], column-major. Workgroup tile 128 x 128 (256 threads, four waves
// in a 2 x 2 grid, each wave a 64 x 64 block of 4 x 4 MFMA 16x16 tiles). K advances in tiles of BK:
// the next tile is fetched into registers while the MFMAs consume the current one, and LDS is
// double-buffered so each K tile costs one barrier. LDS layouts: A as [k][m] (row stride 144,
// conflict-free fragment reads for the two half-waves), B as [n][k] (stride BK + 1).
// MFMA operand layout (16x16x4): lane l supplies A[l % 16][l / 16] and B[l / 16][l % 16];
// accumulator rows per Vec4<T>::row.
template <typename T, int Role>
__global__ __launch_bounds__(256, IBEM_GEMM_MINB) void k_gemm_sub(long M, long N, long K, const T* __restrict__ A,
                                                  long lda, const T* __restrict__ B, long ldb,
                                                  T* __restrict__ C, long ldc) {
  typedef typename Vec4<T>::type V4;
  constexpr int BK = GemmCfg<T>::BK;
  constexpr int LDA_S = kBM + 16;  // [k][m]
  constexpr int LDB_S = BK + 1;    // [n][k]
  constexpr int EA = kBM * BK / 256;  // A elements per thread per tile
  constexpr int EB = kBN * BK / 256;
  __shared__ T sA[2][BK * LDA_S];
  __shared__ T sB[2][kBN * LDB_S];
  const long m0 = (long)blockIdx.x * kBM, n0 = (long)blockIdx.y * kBN;
  const int tid = threadIdx.x;
  const int wave = tid / 64, lane = tid % 64;
  const int wm = (wave & 1) * 64, wn = (wave >> 1) * 64;
  const int lr = lane % 16, lk = lane / 16;
  V4 acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = V4{0, 0, 0, 0};
  T ra[EA], rb[EB];
  // per-thread fixed coordinates: A element e is (m = tid % 128, k = tid / 128 + (256 / 128) e),
  // B element e is (k = tid % BK, n = tid / BK + (256 / BK) e)
  const int am = tid % kBM, ak = tid / kBM;
  const int bk = tid % BK, bn = tid / BK;
  const bool a_in = (m0 + am) < M;
  const T* pa = A + (long)ak * lda + m0 + am;
  const T* pb = B + (n0 + bn) * ldb + bk;
  auto fetch = [&](long k0) {
#pragma unroll
    for (int e = 0; e < EA; ++e) {
      const long gk = k0 + ak + (256 / kBM) * e;
      ra[e] = (a_in && gk < K) ? pa[(k0 + (256 / kBM) * e) * lda] : T(0);
    }
#pragma unroll
    for (int e = 0; e < EB; ++e) {
      const long gn = n0 + bn + (256 / BK) * e;
      const long gk = k0 + bk;
      rb[e] = (gn < N && gk < K) ? pb[(long)(256 / BK) * e * ldb + k0] : T(0);
    }
  };
  auto stash = [&](int buf) {
#pragma unroll
    for (int e = 0; e < EA; ++e) {
      const int idx = tid + 256 * e;
      sA[buf][(idx / kBM) * LDA_S + idx % kBM] = ra[e];
    }
#pragma unroll
    for (int e = 0; e < EB; ++e) {
      const int idx = tid + 256 * e;
      sB[buf][(idx / BK) * LDB_S + idx % BK] = rb[e];
    }
  };
  const long ntiles = (K + BK - 1) / BK;
  fetch(0);
  stash(0);
  __syncthreads();
  for (long t = 0; t < ntiles; ++t) {
    const int buf = (int)(t & 1);
    if (t + 1 < ntiles) fetch((t + 1) * BK);
    const T* a_s = sA[buf];
    const T* b_s = sB[buf];
#pragma unroll
    for (int kk = 0; kk < BK; kk += 4) {
      T a[4], b[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) a[i] = a_s[(kk + lk) * LDA_S + wm + i * 16 + lr];
#pragma unroll
      for (int j = 0; j < 4; ++j) b[j] = b_s[(wn + j * 16 + lr) * LDB_S + kk + lk];
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = mfma16(a[i], b[j], acc[i][j]);
    }
    if (t + 1 < ntiles) stash(buf ^ 1);
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const long nn = n0 + wn + j * 16 + lr;
      if (nn >= N) continue;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const long m = m0 + wm + i * 16 + Vec4<T>::row(lk, v);
        if (m < M) C[nn * ldc + m] -= acc[i][j][v];
      }
    }
}

// Role 0: panel / TRSM updates; Role 1: the trailing-matrix update of the blocked LU (a separate
// kernel symbol so profiles separate the dominant launches).
template <typename T, int Role = 0>
static void gemm_sub(long M, long N, long K, const T* A, long lda, const T* B, long ldb, T* C,
                     long ldc, hipStream_t st) {
  if (M <= 0 || N <= 0 || K <= 0) return;
  dim3 grid((unsigned)((M + kBM - 1) / kBM), (unsigned)((N + kBN - 1) / kBN));
  hipLaunchKernelGGL((k_gemm_sub<T, Role>), grid, dim3(256), 0, st, M, N, K, A, lda, B, ldb, C, ldc);
}

void gemm_f64_sub(long M, long N, long K, const double* A, long lda, const double* B, long ldb,
                  double* C, long ldc, hipStream_t st) {
  gemm_sub<double>(M, N, K, A, lda, B, ldb, C, ldc, st);
}
void gemm_f32_sub(long M, long N, long K, const float* A, long lda, const float* B, long ldb,
                  float* C, long ldc, hipStream_t st) {
  gemm_sub<float>(M, N, K, A, lda, B, ldb, C, ldc, st);
}

// ================================================================ panel (cooperative kernel) ====
constexpr int kPanelW = 32;      // columns factored by one cooperative launch
constexpr int kPanelThreads = 256;
constexpr int kPanelBlocksDefault = 32;  // fewer workgroups make each grid barrier cheaper

// Factor A[j0:n, j0:j0+w) with partial pivoting. Workgroup g owns panel rows
// [j0 + g chunk, j0 + (g+1) chunk). Per column: (a) scale the previous column and update the
// remaining panel columns of the own rows, then the local |max| of the current column;
// grid barrier; (b) workgroup 0 reduces the partial maxima, swaps the two rows across the panel
// width and records the pivot; grid barrier.
template <typename T>
__global__ __launch_bounds__(kPanelThreads) void k_panel(T* __restrict__ A, long lda, long n,
                                                         long j0, int w, int* __restrict__ ipiv,
                                                         double* __restrict__ part,
                                                         int* __restrict__ info) {
  cg::grid_group grid = cg::this_grid();
  __shared__ double sv[kPanelThreads];
  __shared__ long si[kPanelThreads];
  const int G = gridDim.x, g = blockIdx.x, tid = threadIdx.x;
  const long rows = n - j0;
  const long chunk = (rows + G - 1) / G;
  const long rb = j0 + (long)g * chunk;
  const long re = rb + chunk < n ? rb + chunk : n;
  const long jend = j0 + w;
  for (int jj = 0; jj <= w; ++jj) {
    const long col = j0 + jj;
    if (jj > 0) {
      const long pc = col - 1;
      const T piv = A[pc * lda + pc];
      const long r0 = rb > pc + 1 ? rb : pc + 1;
      for (long r = r0 + tid; r < re; r += kPanelThreads) {
        const T l = A[pc * lda + r] / piv;
        A[pc * lda + r] = l;
        for (long c = col; c < jend; ++c) A[c * lda + r] -= l * A[c * lda + pc];
      }
    }
    if (jj == w) break;
    double best = -1.0;
    long bi = -1;
    const long r0 = rb > col ? rb : col;
    for (long r = r0 + tid; r < re; r += kPanelThreads) {
      const double v = fabs((double)A[col * lda + r]);
      if (v > best) {
        best = v;
        bi = r;
      }
    }
    sv[tid] = best;
    si[tid] = bi;
    __syncthreads();
    for (int s = kPanelThreads / 2; s > 0; s >>= 1) {
      if (tid < s) {
        const double v2 = sv[tid + s];
        const long i2 = si[tid + s];
        if (v2 > sv[tid] || (v2 == sv[tid] && i2 >= 0 && (si[tid] < 0 || i2 < si[tid]))) {
          sv[tid] = v2;
          si[tid] = i2;
        }
      }
      __syncthreads();
    }
    if (tid == 0) {
      part[2 * g] = sv[0];
      part[2 * g + 1] = (double)si[0];
    }
    grid.sync();
    if (g == 0) {
      double bv = -1.0;
      long bidx = -1;
      for (int q = tid; q < G; q += kPanelThreads) {
        const double v = part[2 * q];
        const long ix = (long)part[2 * q + 1];
        if (ix >= 0 && (v > bv || (v == bv && ix < bidx))) {
          bv = v;
          bidx = ix;
        }
      }
      sv[tid] = bv;
      si[tid] = bidx;
      __syncthreads();
      for (int s = kPanelThreads / 2; s > 0; s >>= 1) {
        if (tid < s) {
          const double v2 = sv[tid + s];
          const long i2 = si[tid + s];
          if (i2 >= 0 && (v2 > sv[tid] || (v2 == sv[tid] && (si[tid] < 0 || i2 < si[tid])))) {
            sv[tid] = v2;
            si[tid] = i2;
          }
        }
        __syncthreads();
      }
      long p = si[0];
      if (p < 0 || sv[0] == 0.0) {
        if (tid == 0 && *info == 0) *info = (int)(col + 1);
        p = col;
      }
      if (tid == 0) ipiv[col] = (int)p;
      if (p != col) {
        for (long c = j0 + tid; c < jend; c += kPanelThreads) {
          const T t = A[c * lda + col];
          A[c * lda + col] = A[c * lda + p];
          A[c * lda + p] = t;
        }
      }
    }
    grid.sync();
  }
}

// ============================================================================ row swaps ==========
// Apply the interchanges ipiv[p0..p1) in order to columns [c0, c0 + nc).
template <typename T>
__global__ __launch_bounds__(256) void k_laswp(T* __restrict__ A, long lda, long c0, long nc,
                                               const int* __restrict__ ipiv, long p0, long p1) {
  const long c = c0 + (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= c0 + nc) return;
  T* col = A + c * lda;
  for (long i = p0; i < p1; ++i) {
    const long p = ipiv[i];
    if (p != i) {
      const T t = col[i];
      col[i] = col[p];
      col[p] = t;
    }
  }
}

template <typename T>
static void laswp(T* A, long lda, long c0, long nc, const int* ipiv, long p0, long p1,
                  hipStream_t st) {
  if (nc <= 0 || p1 <= p0) return;
  hipLaunchKernelGGL(k_laswp<T>, dim3((unsigned)((nc + 255) / 256)), dim3(256), 0, st, A, lda, c0, nc,
                     ipiv, p0, p1);
}

// ============================================================ small unit-lower triangular solve ==
// B[m x N] <- L^{-1} B, L unit lower m x m (m <= 32); one thread per column of B.
template <typename T>
__global__ __launch_bounds__(256) void k_trsm_small(const T* __restrict__ L, long ldl, int m,
                                                    T* __restrict__ B, long ldb, long N) {
  __shared__ T sL[32][33];
  for (int u = threadIdx.x; u < m * m; u += blockDim.x) {
    const int i = u % m, j = u / m;
    sL[i][j] = L[(long)j * ldl + i];
  }
  __syncthreads();
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= N) return;
  T x[32];
  T* col = B + c * ldb;
#pragma unroll
  for (int i = 0; i < 32; ++i) x[i] = i < m ? col[i] : T(0);
#pragma unroll
  for (int i = 1; i < 32; ++i) {
    T v = x[i];
#pragma unroll
    for (int j = 0; j < i; ++j) v -= sL[i][j] * x[j];
    if (i < m) x[i] = v;
  }
#pragma unroll
  for (int i = 0; i < 32; ++i)
    if (i < m) col[i] = x[i];
}

template <typename T>
static void trsm_lower_unit(const T* L, long ldl, long m, T* B, long ldb, long N, hipStream_t st) {
  if (m <= 0 || N <= 0) return;
  if (m <= 32) {
    hipLaunchKernelGGL(k_trsm_small<T>, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, st, L, ldl,
                       (int)m, B, ldb, N);
    return;
  }
  const long m1 = (m / 2 + 31) / 32 * 32;
  trsm_lower_unit(L, ldl, m1, B, ldb, N, st);
  gemm_sub<T>(m - m1, N, m1, L + m1, ldl, B, ldb, B + m1, ldb, st);
  trsm_lower_unit(L + m1 * ldl + m1, ldl, m - m1, B + m1, ldb, N, st);
}

// ================================================================================ LU driver =====
template <typename T>
struct LuCtx {
  T* A;
  long n, lda;
  LuWork* W;
  hipStream_t st;
  void (*tick)(void*, const char*, int);
  void* tctx;
  int err;
};

template <typename T>
static void panel_leaf(LuCtx<T>& c, long j0, int w) {
  void* args[] = {(void*)&c.A, (void*)&c.lda, (void*)&c.n, (void*)&j0, (void*)&w,
                  (void*)&c.W->d_ipiv, (void*)&c.W->d_part, (void*)&c.W->d_info};
  long rows = c.n - j0;
  int G = c.W->coop_blocks;
  long need = (rows + 63) / 64;  // at least 64 rows per workgroup
  if (need < G) G = (int)std::max<long>(1, need);
  hipError_t e = hipLaunchCooperativeKernel((const void*)k_panel<T>, dim3(G), dim3(kPanelThreads),
                                            args, 0, c.st);
  if (e != hipSuccess) c.err = 1;
}

template <typename T>
static void panel_rec(LuCtx<T>& c, long j0, long w) {
  if (c.err) return;
  if (w <= kPanelW) {
    panel_leaf(c, j0, (int)w);
    return;
  }
  const long w1 = (w / 2 + kPanelW - 1) / kPanelW * kPanelW;
  const long lda = c.lda;
  panel_rec(c, j0, w1);
  laswp<T>(c.A, lda, j0 + w1, w - w1, c.W->d_ipiv, j0, j0 + w1, c.st);
  trsm_lower_unit<T>(c.A + j0 * lda + j0, lda, w1, c.A + (j0 + w1) * lda + j0, lda, w - w1, c.st);
  gemm_sub<T>(c.n - j0 - w1, w - w1, w1, c.A + j0 * lda + j0 + w1, lda, c.A + (j0 + w1) * lda + j0,
              lda, c.A + (j0 + w1) * lda + j0 + w1, lda, c.st);
  panel_rec(c, j0 + w1, w - w1);
  laswp<T>(c.A, lda, j0, w1, c.W->d_ipiv, j0 + w1, j0 + w, c.st);
}

constexpr long kLuNB = 256;

template <typename T>
static int lu_factor(T* A, long n, long lda, LuWork& W, hipStream_t st,
                     void (*tick)(void*, const char*, int), void* tctx) {
  LuCtx<T> c{A, n, lda, &W, st, tick, tctx, 0};
  if (W.coop_blocks <= 0) {
    int dev = 0, cus = 0, occ = 0;
    hipGetDevice(&dev);
    hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, (const void*)k_panel<T>, kPanelThreads, 0);
    W.coop_blocks = std::max(1, std::min(occ, 2) * cus);
    if (const char* ev = getenv("IBEM3D_PANEL_BLOCKS")) {
      const int v = atoi(ev);
      if (v > 0) W.coop_blocks = std::min(v, W.coop_blocks);
    } else {
      W.coop_blocks = std::min(W.coop_blocks, kPanelBlocksDefault);
    }
  }
  hipMemsetAsync(W.d_info, 0, sizeof(int), st);
  for (long k = 0; k < n; k += kLuNB) {
    const long kb = std::min(kLuNB, n - k);
    if (tick) tick(tctx, "lu.panel", 0);
    panel_rec(c, k, kb);
    if (tick) tick(tctx, "lu.panel", 1);
    if (c.err) return -1;
    if (tick) tick(tctx, "lu.swap_trsm", 0);
    laswp<T>(A, lda, 0, k, W.d_ipiv, k, k + kb, st);
    if (k + kb < n) {
      laswp<T>(A, lda, k + kb, n - k - kb, W.d_ipiv, k, k + kb, st);
      trsm_lower_unit<T>(A + k * lda + k, lda, kb, A + (k + kb) * lda + k, lda, n - k - kb, st);
    }
    if (tick) tick(tctx, "lu.swap_trsm", 1);
    if (k + kb < n) {
      if (tick) tick(tctx, "lu.gemm", 0);
      gemm_sub<T, 1>(n - k - kb, n - k - kb, kb, A + k * lda + k + kb, lda, A + (k + kb) * lda + k,
                     lda, A + (k + kb) * lda + k + kb, lda, st);
      if (tick) tick(tctx, "lu.gemm", 1);
    }
  }
  return 0;
}

int lu_factor_f64(double* A, long n, long lda, LuWork& W, hipStream_t st,
                  void (*tick)(void*, const char*, int), void* tctx) {
  return lu_factor<double>(A, n, lda, W, st, tick, tctx);
}
int lu_factor_f32(float* A, long n, long lda, LuWork& W, hipStream_t st,
                  void (*tick)(void*, const char*, int), void* tctx) {
  return lu_factor<float>(A, n, lda, W, st, tick, tctx);
}

// ======================================================================= triangular solves ======
constexpr int kTB = 256;

template <typename T>
__global__ void k_apply_pivots(T* x, long n, const int* __restrict__ ipiv) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  for (long i = 0; i < n; ++i) {
    const long p = ipiv[i];
    if (p != i) {
      const T t = x[i];
      x[i] = x[p];
      x[p] = t;
    }
  }
}

// x[k0:k0+kb] <- L_kk^{-1} x (unit lower) or U_kk^{-1} x (upper), one workgroup.
template <typename T, bool Upper>
__global__ __launch_bounds__(kTB) void k_trsv_block(const T* __restrict__ LU, long lda, long k0,
                                                    int kb, T* __restrict__ x) {
  __shared__ T sx[kTB];
  const int t = threadIdx.x;
  if (t < kb) sx[t] = x[k0 + t];
  __syncthreads();
  if (!Upper) {
    for (int i = 0; i < kb; ++i) {
      if (t > i && t < kb) sx[t] -= LU[(k0 + i) * lda + k0 + t] * sx[i];
      __syncthreads();
    }
  } else {
    for (int i = kb - 1; i >= 0; --i) {
      if (t == i) sx[i] /= LU[(k0 + i) * lda + k0 + i];
      __syncthreads();
      if (t < i) sx[t] -= LU[(k0 + i) * lda + k0 + t] * sx[i];
      __syncthreads();
    }
  }
  if (t < kb) x[k0 + t] = sx[t];
}

// y[0:m] -= A[0:m, 0:k] x[0:k] (column-major), k <= kTB; one thread per row
template <typename T>
__global__ __launch_bounds__(256) void k_gemv_sub_small(long m, int k, const T* __restrict__ A,
                                                        long lda, const T* __restrict__ x,
                                                        T* __restrict__ y) {
  __shared__ T sx[kTB];
  for (int j = threadIdx.x; j < k; j += blockDim.x) sx[j] = x[j];
  __syncthreads();
  const long r = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= m) return;
  T acc = 0;
  for (int j = 0; j < k; ++j) acc += A[(long)j * lda + r] * sx[j];
  y[r] -= acc;
}

template <typename T>
static void lu_solve(const T* LU, long n, long lda, const int* ipiv, T* x, hipStream_t st) {
  hipLaunchKernelGGL(k_apply_pivots<T>, dim3(1), dim3(1), 0, st, x, n, ipiv);
  for (long k = 0; k < n; k += kTB) {
    const int kb = (int)std::min<long>(kTB, n - k);
    hipLaunchKernelGGL((k_trsv_block<T, false>), dim3(1), dim3(kTB), 0, st, LU, lda, k, kb, x);
    const long m = n - k - kb;
    if (m > 0)
      hipLaunchKernelGGL(k_gemv_sub_small<T>, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, st, m,
                         kb, LU + k * lda + k + kb, lda, x + k, x + k + kb);
  }
  const long last = ((n - 1) / kTB) * kTB;
  for (long k = last; k >= 0; k -= kTB) {
    const int kb = (int)std::min<long>(kTB, n - k);
    hipLaunchKernelGGL((k_trsv_block<T, true>), dim3(1), dim3(kTB), 0, st, LU, lda, k, kb, x);
    if (k > 0)
      hipLaunchKernelGGL(k_gemv_sub_small<T>, dim3((unsigned)((k + 255) / 256)), dim3(256), 0, st, k,
                         kb, LU + k * lda, lda, x + k, x);
  }
}

void lu_solve_f64(const double* LU, long n, long lda, const int* ipiv, double* x, hipStream_t st) {
  lu_solve<double>(LU, n, lda, ipiv, x, st);
}
void lu_solve_f32(const float* LU, long n, long lda, const int* ipiv, float* x, hipStream_t st) {
  lu_solve<float>(LU, n, lda, ipiv, x, st);
}

// ========================================================================= equilibration ========
__global__ __launch_bounds__(256) void k_rowmax(const double* __restrict__ A, long n, long lda,
                                                double* __restrict__ r) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double m = 0.0;
  for (long j = 0; j < n; ++j) m = fmax(m, fabs(A[j * lda + i]));
  r[i] = m > 0.0 ? 1.0 / m : 1.0;
}

__global__ __launch_bounds__(256) void k_colscale(double* __restrict__ A, long n, long lda,
                                                  const double* __restrict__ r,
                                                  double* __restrict__ c) {
  __shared__ double red[256];
  const long j = blockIdx.x;
  double* col = A + j * lda;
  double m = 0.0;
  for (long i = threadIdx.x; i < n; i += blockDim.x) m = fmax(m, fabs(col[i] * r[i]));
  red[threadIdx.x] = m;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + s]);
    __syncthreads();
  }
  const double cj = red[0] > 0.0 ? 1.0 / red[0] : 1.0;
  if (threadIdx.x == 0) c[j] = cj;
  for (long i = threadIdx.x; i < n; i += blockDim.x) col[i] = col[i] * r[i] * cj;
}

void equilibrate_f64(double* A, long n, long lda, double* r, double* c, hipStream_t st) {
  hipLaunchKernelGGL(k_rowmax, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, A, n, lda, r);
  hipLaunchKernelGGL(k_colscale, dim3((unsigned)n), dim3(256), 0, st, A, n, lda, r, c);
}

// ==================================================================================== misc =======
// y = alpha A x + beta y, column chunks of 1024 reduced through a partial buffer per row chunk.
__global__ __launch_bounds__(256) void k_gemv(long m, long n, double alpha,
                                              const double* __restrict__ A, long lda,
                                              const double* __restrict__ x, double beta,
                                              double* __restrict__ y) {
  const long r = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= m) return;
  double acc0 = 0.0, acc1 = 0.0, acc2 = 0.0, acc3 = 0.0;
  long j = 0;
  for (; j + 4 <= n; j += 4) {
    acc0 += A[j * lda + r] * x[j];
    acc1 += A[(j + 1) * lda + r] * x[j + 1];
    acc2 += A[(j + 2) * lda + r] * x[j + 2];
    acc3 += A[(j + 3) * lda + r] * x[j + 3];
  }
  for (; j < n; ++j) acc0 += A[j * lda + r] * x[j];
  const double v = alpha * ((acc0 + acc1) + (acc2 + acc3));
  y[r] = beta == 0.0 ? v : v + beta * y[r];
}

void gemv_f64(long m, long n, double alpha, const double* A, long lda, const double* x, double beta,
              double* y, hipStream_t st) {
  hipLaunchKernelGGL(k_gemv, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, st, m, n, alpha, A, lda,
                     x, beta, y);
}

__global__ __launch_bounds__(256) void k_f64_to_f32(const double* __restrict__ A, long n, long lda,
                                                    float* __restrict__ B, long ldb) {
  const long j = blockIdx.y;
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) B[j * ldb + i] = (float)A[j * lda + i];
}

void convert_f64_to_f32(const double* A, long n, long lda, float* B, long ldb, hipStream_t st) {
  hipLaunchKernelGGL(k_f64_to_f32, dim3((unsigned)((n + 255) / 256), (unsigned)n), dim3(256), 0, st, A,
                     n, lda, B, ldb);
}

}  // namespace ibem
