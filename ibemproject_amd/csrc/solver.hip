// Dense LU solver of the iBEM system on gfx950 (replaces Eigen FullPivLU, systemSolverElastic3D.h:55).
//
// Blocked right-looking LU with row pivoting, column-major, in place:
//   for each IBEM_LU_NB-wide (1024) panel: recursive panel factorisation (leaves of 32 columns
//   factored with tournament pivoting in two launches), row interchanges on the other columns,
//   unit-lower TRSM, and the trailing-matrix update C -= L21 U12 on MFMA
//   (v_mfma_f64_16x16x4f64, or FP32 / split or single FP16 / BF16 products for mixed precision).
// The trailing update holds 2n^3/3 of the flops; everything else is memory- or latency-bound.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <vector>

#include "ibem3d_kernels.h"

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

// Tile order of the split GEMM (1D grid). Workgroups are dispatched to the 8 XCDs round-robin by
// linear id and each XCD has its own L2; tiles are enumerated in groups of kGroupM M-tiles (M
// fastest, then N) and every XCD takes one contiguous eighth of that enumeration, so the
// workgroups resident on an XCD cover a compact block of tiles and share operand rows and columns
// in its L2 instead of re-reading them from memory.
#ifndef IBEM_GROUP_M
#define IBEM_GROUP_M 8
#endif
constexpr int kXcds = 8, kGroupM = IBEM_GROUP_M;
__device__ inline bool tile_of(long L, long mt, long nt, long& tm, long& tn) {
  const long T = mt * nt;
  const long per = (T + kXcds - 1) / kXcds;
  const long t = (L % kXcds) * per + L / kXcds;
  if (t >= T) return false;
  const long grp = t / (kGroupM * nt);
  const long first = grp * kGroupM;
  const long gsz = (mt - first) < kGroupM ? (mt - first) : kGroupM;
  const long in = t - grp * kGroupM * nt;
  tm = first + in % gsz;
  tn = in / gsz;
  return true;
}
static inline unsigned tile_grid(long M, long N) {
  const long T = ((M + kBM - 1) / kBM) * ((N + kBN - 1) / kBN);
  return (unsigned)(((T + kXcds - 1) / kXcds) * kXcds);
}

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
        if (m < M) {
          if constexpr (Role == 2)
            C[nn * ldc + m] = acc[i][j][v];
          else
            C[nn * ldc + m] -= acc[i][j][v];
        }
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

// C[M x N] -= A[M x K] B[K x N] for the panel's narrowest updates (K = 32 or 64, N <= 64, M tall):
// one thread per row of C holds its row of A in registers, B is staged in LDS (read as broadcasts),
// four columns of C per pass. Light workgroups (no MFMA tiles) so that they start quickly on the
// few CUs the panel gets while the bulk update runs.
constexpr int kSkinnyMax = 64;
template <typename T, int K>
__global__ __launch_bounds__(256, sizeof(T) == 4 ? 4 : 2) void k_gemm_skinny(
    long M, int N, const T* __restrict__ A, long lda, const T* __restrict__ B, long ldb,
    T* __restrict__ C, long ldc) {
  __shared__ __align__(16) T sB[kSkinnyMax][K];  // sB[j][k] = B(k, j)
  const int t = threadIdx.x;
  for (int u = t; u < K * N; u += 256) {
    const int k = u % K, j = u / K;
    sB[j][k] = B[(long)j * ldb + k];
  }
  __syncthreads();
  const long r = (long)blockIdx.x * 256 + t;
  if (r >= M) return;
  T a[K];
#pragma unroll
  for (int k = 0; k < K; ++k) a[k] = A[(long)k * lda + r];
  // four columns at a time: their C loads are issued together, ahead of the four dot products
  // (one column at a time, every column would wait a full memory latency for its C)
  int j = 0;
#pragma unroll 1
  for (; j + 4 <= N; j += 4) {
    T cv[4];
#pragma unroll
    for (int jj = 0; jj < 4; ++jj) cv[jj] = C[(long)(j + jj) * ldc + r];
#pragma unroll
    for (int jj = 0; jj < 4; ++jj) {
      const T* b = sB[j + jj];
      T p[4] = {T(0), T(0), T(0), T(0)};
#pragma unroll
      for (int k = 0; k < K; k += 4)
#pragma unroll
        for (int q = 0; q < 4; ++q) p[q] = fma(a[k + q], b[k + q], p[q]);
      C[(long)(j + jj) * ldc + r] = cv[jj] - ((p[0] + p[1]) + (p[2] + p[3]));
    }
  }
#pragma unroll 1
  for (; j < N; ++j) {
    const T* b = sB[j];
    T p[4] = {T(0), T(0), T(0), T(0)};
#pragma unroll
    for (int k = 0; k < K; k += 4)
#pragma unroll
      for (int q = 0; q < 4; ++q) p[q] = fma(a[k + q], b[k + q], p[q]);
    C[(long)j * ldc + r] -= (p[0] + p[1]) + (p[2] + p[3]);
  }
}
template <typename T>
static void gemm_skinny(long M, long N, long K, const T* A, long lda, const T* B, long ldb, T* C,
                        long ldc, hipStream_t st) {
  if (M <= 0 || N <= 0) return;
  const dim3 g((unsigned)((M + 255) / 256));
  if (K == 32)
    hipLaunchKernelGGL((k_gemm_skinny<T, 32>), g, dim3(256), 0, st, M, (int)N, A, lda, B, ldb, C, ldc);
  else
    hipLaunchKernelGGL((k_gemm_skinny<T, 64>), g, dim3(256), 0, st, M, (int)N, A, lda, B, ldb, C, ldc);
}

void gemm_f64_sub(long M, long N, long K, const double* A, long lda, const double* B, long ldb,
                  double* C, long ldc, hipStream_t st) {
  gemm_sub<double>(M, N, K, A, lda, B, ldb, C, ldc, st);
}
void gemm_f32_sub(long M, long N, long K, const float* A, long lda, const float* B, long ldb,
                  float* C, long ldc, hipStream_t st) {
  gemm_sub<float>(M, N, K, A, lda, B, ldb, C, ldc, st);
}

// ================================================= FP32 update on FP16 MFMA (three-term split) ==
// C -= A B in FP32 arithmetic accuracy from half-precision products. Each operand row (of A) and
// column (of B) is scaled by a power of two that puts its largest magnitude in [2^14, 2^15), and
// split x = hi + lo with hi = fp16(x), lo = fp16(x - hi) (22 significant bits). Then
//   A B = S_A^-1 (Ahi Bhi + Ahi Blo + Alo Bhi) S_B^-1 + O(2^-22 |A| |B|),
// with every half x half product exact in the MFMA's FP32 accumulator, and the dropped Alo Blo term
// below FP32 rounding of C for K = 256. Scaling keeps operands out of the FP16 subnormal range down
// to 2^-28 of their row / column maximum. Three v_mfma_f32_16x16x32_f16 (2.5 PFLOP/s dense)
// replace 8 v_mfma_f32_16x16x4f32 (157.3 TFLOP/s) per 16x16x32 block.
typedef _Float16 half8 __attribute__((ext_vector_type(8)));
#ifndef IBEM_LU_NB
#define IBEM_LU_NB 1024
#endif
constexpr int kSplitK = IBEM_LU_NB;     // operand depth (the LU block size)
constexpr int kSplitPadK = IBEM_LU_NB;  // row stride of the split operands, in halves

__device__ inline float pow2_scale(float amax) {
  // 2^(15 - e) with amax in [2^(e-1), 2^e); 1 for amax == 0
  if (!(amax > 0.0f)) return 1.0f;
  int e;
  frexpf(amax, &e);
  return ldexpf(1.0f, 15 - e);
}

__device__ inline void split2(float x, _Float16& hi, _Float16& lo) {
  hi = (_Float16)x;
  lo = (_Float16)(x - (float)hi);
}

// Product types of the split-operand GEMMs (kProd*, ibem3d_kernels.h). BF16 operands are stored
// as their bit patterns in the same 16-bit arrays.
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
// product types with a lo part (three products per multiply-add: hi hi + hi lo + lo hi)
template <int PROD>
constexpr bool has_lo() { return PROD == kProdF16x3 || PROD == kProdBF16x3; }
template <int PROD>
__device__ inline void split_to(float x, _Float16& hi, _Float16& lo) {
  if constexpr (PROD == kProdBF16) {
    hi = __builtin_bit_cast(_Float16, (__bf16)x);
  } else if constexpr (PROD == kProdBF16x3) {
    const __bf16 h = (__bf16)x;
    hi = __builtin_bit_cast(_Float16, h);
    lo = __builtin_bit_cast(_Float16, (__bf16)(x - (float)h));
  } else if constexpr (PROD == kProdF16) {
    hi = (_Float16)x;
  } else {
    split2(x, hi, lo);
  }
}
template <int PROD>
__device__ inline v4f mma16(half8 a, half8 b, v4f c) {
  if constexpr (PROD == kProdBF16 || PROD == kProdBF16x3)
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a),
                                                   __builtin_bit_cast(bf16x8, b), c, 0, 0, 0);
  else
    return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
}

// A operand: rows r < M of the column-major M x K block at A (lda). Output row-major [M][K] hi / lo
// (k contiguous, row stride kSplitPadK) and the row's inverse scale; k in [K, kpad) written as zeros,
// k >= kpad not written (kpad: a multiple of 64, the depth the GEMM reads). Workgroup = kSplitRW
// rows, one row per lane of each wavefront so that every column read is 256 contiguous bytes: a
// pass for the row maxima, then the transpose and split in chunks of kSplitChunk columns through LDS
// (the second read of A mostly hits L2).
constexpr int kSplitRW = 64;      // rows per workgroup
constexpr int kSplitChunk = 128;  // k per LDS round
template <int PROD>
__global__ __launch_bounds__(256) void k_split_rows(long M, int K, int kpad, const float* __restrict__ A,
                                                    long lda, _Float16* __restrict__ hi,
                                                    _Float16* __restrict__ lo,
                                                    float* __restrict__ inv_scale) {
  __shared__ float tile[kSplitRW][kSplitChunk + 1];
  __shared__ float red[4][kSplitRW];
  __shared__ float ssc[kSplitRW];
  const long r0 = (long)blockIdx.x * kSplitRW;
  const int t = threadIdx.x, rl = t & 63, g = t >> 6;
  const long r = r0 + rl;
  // eight loads in flight per thread (one at a time, each pass would wait a full memory latency)
  float m = 0.0f;
  if (r < M) {
    int k = g;
    for (; k + 28 < K; k += 32) {
      float v[8];
#pragma unroll
      for (int q = 0; q < 8; ++q) v[q] = A[(long)(k + 4 * q) * lda + r];
#pragma unroll
      for (int q = 0; q < 8; ++q) m = fmaxf(m, fabsf(v[q]));
    }
    for (; k < K; k += 4) m = fmaxf(m, fabsf(A[(long)k * lda + r]));
  }
  red[g][rl] = m;
  __syncthreads();
  if (g == 0) {
    const float mm = fmaxf(fmaxf(red[0][rl], red[1][rl]), fmaxf(red[2][rl], red[3][rl]));
    const float sc = pow2_scale(mm);
    ssc[rl] = sc;
    if (r < M) inv_scale[r] = 1.0f / sc;
  }
  for (int k0 = 0; k0 < kpad; k0 += kSplitChunk) {
    __syncthreads();  // ssc ready / the previous chunk has been written out
#pragma unroll
    for (int kk = g; kk < kSplitChunk; kk += 4) {
      const int k = k0 + kk;
      tile[rl][kk] = (r < M && k < K) ? A[(long)k * lda + r] : 0.0f;
    }
    __syncthreads();
    // 16-byte stores: 8 consecutive k of one row per pass
    for (int u = t; u < kSplitRW * (kSplitChunk / 8); u += 256) {
      const int rr = u / (kSplitChunk / 8), k8 = (u % (kSplitChunk / 8)) * 8;
      const long row = r0 + rr;
      if (row >= M || k0 + k8 >= kpad) continue;
      half8 h, l;
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        _Float16 x, y = 0;
        split_to<PROD>(tile[rr][k8 + q] * ssc[rr], x, y);
        h[q] = x;
        l[q] = y;
      }
      *(half8*)(hi + row * kSplitPadK + k0 + k8) = h;
      if constexpr (has_lo<PROD>()) *(half8*)(lo + row * kSplitPadK + k0 + k8) = l;
    }
  }
}

// B operand: columns c < N of the column-major K x N block at B (ldb), output [N][K] hi / lo (k
// contiguous, row stride kSplitPadK; written for k < kpad) and the column's inverse scale. One
// wavefront per column. With a single product (PROD != kProdF16x3) only hi is written.
template <int PROD>
__global__ __launch_bounds__(256) void k_split_cols(long N, int K, int kpad, const float* __restrict__ B,
                                                    long ldb, _Float16* __restrict__ hi,
                                                    _Float16* __restrict__ lo,
                                                    float* __restrict__ inv_scale) {
  constexpr int VPL = kSplitPadK / 64;  // values per lane (a multiple of 4)
  const long c = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (c >= N) return;
  const float* col = B + c * ldb;
  float v[VPL];
  float m = 0.0f;
#pragma unroll
  for (int q = 0; q < VPL; ++q) {
    const int k = lane * VPL + q;
    v[q] = k < K ? col[k] : 0.0f;
    m = fmaxf(m, fabsf(v[q]));
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o));
  const float sc = pow2_scale(m);
  if (lane == 0) inv_scale[c] = 1.0f / sc;
  typedef _Float16 half4 __attribute__((ext_vector_type(4)));
  if (lane * VPL >= kpad) return;
#pragma unroll
  for (int q4 = 0; q4 < VPL; q4 += 4) {
    half4 h, l;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      _Float16 a, b = 0;
      split_to<PROD>(v[q4 + q] * sc, a, b);
      h[q] = a;
      l[q] = b;
    }
    *(half4*)(hi + c * kSplitPadK + lane * VPL + q4) = h;
    if constexpr (has_lo<PROD>()) *(half4*)(lo + c * kSplitPadK + lane * VPL + q4) = l;
  }
}

// C[M x N] (column-major, ldc) -= diag(ia) (Ah Bh^T + Ah Bl^T + Al Bh^T) diag(ib), with A* [M][K]
// and B* [N][K] (k contiguous, row stride kSplitPadK). Workgroup tile 128 x 128, four waves in a
// 2 x 2 grid of 64 x 64 (4 x 4 blocks of v_mfma_f32_16x16x32_f16). K advances BK at a time: the
// next chunk is fetched into registers while the MFMAs consume the current one from LDS; the three
// products run as three passes over the 16 accumulators, so consecutive MFMAs are independent.
// LDS rows are padded by 8 halves so the 16-byte fragment reads of a half-wave hit distinct banks.
// With EPI = 1 the accumulators go through LDS and C is updated with coalesced 16-byte accesses.
// Fragment layout (16x16x32): lane l supplies A[l % 16][8 (l / 16) .. +7] and
// B[8 (l / 16) .. +7][l % 16]; accumulator register v of lane l is D[4 (l / 16) + v][l % 16].
#ifndef IBEM_F16_BK
#define IBEM_F16_BK 64
#endif
#ifndef IBEM_F16_EPI
#define IBEM_F16_EPI 1
#endif
#ifndef IBEM_F16_PF
#define IBEM_F16_PF 1
#endif
#ifndef IBEM_F16_PAD
#define IBEM_F16_PAD 8
#endif
template <int BK, int EPI, int PF, int NCH = kSplitPadK / BK, int PROD = kProdF16x3>
__global__ __launch_bounds__(256, 2) void k_gemm_f16x3(long M, long N, int K,
                                                       const _Float16* __restrict__ Ah,
                                                       const _Float16* __restrict__ Al,
                                                       const _Float16* __restrict__ Bh,
                                                       const _Float16* __restrict__ Bl,
                                                       const float* __restrict__ ia,
                                                       const float* __restrict__ ib,
                                                       float* __restrict__ C, long ldc) {
  (void)K;  // operands are zero-padded to NCH * BK, so every tile runs all NCH chunks
  constexpr int LS = BK + IBEM_F16_PAD;  // LDS row stride (halves)
  constexpr int VPR = BK / 8;            // 16-byte vectors per row and array
  constexpr int NV = kBM * VPR / 256;    // vectors per thread and array
  constexpr int CS = kBM + 4;            // epilogue tile column stride (floats)
  constexpr int SMEM = (EPI && kBN * CS * 2 > 4 * kBM * LS) ? kBN * CS * 2 : 4 * kBM * LS;
  __shared__ __attribute__((aligned(16))) _Float16 smem[SMEM];
  _Float16* sAh = smem;
  _Float16* sAl = smem + kBM * LS;
  _Float16* sBh = smem + 2 * kBM * LS;
  _Float16* sBl = smem + 3 * kBM * LS;
  long tm, tn;
  if (!tile_of(blockIdx.x, (M + kBM - 1) / kBM, (N + kBN - 1) / kBN, tm, tn)) return;
  const long m0 = tm * kBM, n0 = tn * kBN;
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int wm = (wave & 1) * 64, wn = (wave >> 1) * 64;
  const int lr = lane & 15, lk = lane >> 4;
  v4f acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = v4f{0, 0, 0, 0};
  // register ring: chunk c lives in slot c % PF, fetched PF chunks ahead of its use
  half8 rAh[PF][NV], rAl[PF][NV], rBh[PF][NV], rBl[PF][NV];
  auto fetch = [&](int k0, int sl) {
#pragma unroll
    for (int e = 0; e < NV; ++e) {
      const int v = tid + 256 * e, row = v / VPR, part = v % VPR;
      const long ga = m0 + row, gb = n0 + row;
      const long ko = k0 + part * 8;
      const half8 z = {0, 0, 0, 0, 0, 0, 0, 0};
      rAh[sl][e] = ga < M ? *(const half8*)(Ah + ga * kSplitPadK + ko) : z;
      rBh[sl][e] = gb < N ? *(const half8*)(Bh + gb * kSplitPadK + ko) : z;
      if constexpr (has_lo<PROD>()) {
        rAl[sl][e] = ga < M ? *(const half8*)(Al + ga * kSplitPadK + ko) : z;
        rBl[sl][e] = gb < N ? *(const half8*)(Bl + gb * kSplitPadK + ko) : z;
      }
    }
  };
  auto stash = [&](int sl) {
#pragma unroll
    for (int e = 0; e < NV; ++e) {
      const int v = tid + 256 * e, row = v / VPR, part = v % VPR;
      *(half8*)(sAh + row * LS + part * 8) = rAh[sl][e];
      *(half8*)(sBh + row * LS + part * 8) = rBh[sl][e];
      if constexpr (has_lo<PROD>()) {
        *(half8*)(sAl + row * LS + part * 8) = rAl[sl][e];
        *(half8*)(sBl + row * LS + part * 8) = rBl[sl][e];
      }
    }
  };
#pragma unroll
  for (int c = 0; c < PF && c < NCH; ++c) fetch(c * BK, c);
#pragma unroll
  for (int ch = 0; ch < NCH; ++ch) {
    __syncthreads();
    stash(ch % PF);
    __syncthreads();
    if (ch + PF < NCH) fetch((ch + PF) * BK, ch % PF);
#pragma unroll
    for (int kk = 0; kk < BK; kk += 32) {
      half8 ah[4], al[4], bh[4], bl[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int off = (wm + i * 16 + lr) * LS + kk + lk * 8;
        ah[i] = *(const half8*)(sAh + off);
        if constexpr (has_lo<PROD>()) al[i] = *(const half8*)(sAl + off);
      }
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int off = (wn + j * 16 + lr) * LS + kk + lk * 8;
        bh[j] = *(const half8*)(sBh + off);
        if constexpr (has_lo<PROD>()) bl[j] = *(const half8*)(sBl + off);
      }
      if constexpr (has_lo<PROD>()) {
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
          for (int j = 0; j < 4; ++j) acc[i][j] = mma16<PROD>(al[i], bh[j], acc[i][j]);
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
          for (int j = 0; j < 4; ++j) acc[i][j] = mma16<PROD>(ah[i], bl[j], acc[i][j]);
      }
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = mma16<PROD>(ah[i], bh[j], acc[i][j]);
    }
  }
  if constexpr (EPI == 0) {
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const long nn = n0 + wn + j * 16 + lr;
      if (nn >= N) continue;
      const float sb = ib[nn];
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const long m = m0 + wm + i * 16 + 4 * lk + v;
          if (m < M) C[nn * ldc + m] -= acc[i][j][v] * ia[m] * sb;
        }
    }
  } else {
    // tile 128 x 128 floats, column-major with stride CS, over the operand LDS
    float* sc = (float*)smem;
    __syncthreads();
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j)
#pragma unroll
        for (int v = 0; v < 4; ++v) sc[(wn + j * 16 + lr) * CS + wm + i * 16 + 4 * lk + v] = acc[i][j][v];
    __syncthreads();
    // thread: 4 consecutive rows of one column per pass; 32 threads per column
    const bool full = (m0 + kBM <= M) && ((ldc & 3) == 0) && ((((size_t)C) & 15) == 0);
    for (int u = tid; u < kBN * (kBM / 4); u += 256) {
      const int cl = u / (kBM / 4), r4 = (u % (kBM / 4)) * 4;
      const long nn = n0 + cl;
      if (nn >= N) continue;
      const float sb = ib[nn];
      const long m = m0 + r4;
      float* cp = C + nn * ldc + m;
      if (full) {
        float4 c = *(float4*)cp;
        const float4 a = *(const float4*)(ia + m);
        c.x -= sc[cl * CS + r4] * a.x * sb;
        c.y -= sc[cl * CS + r4 + 1] * a.y * sb;
        c.z -= sc[cl * CS + r4 + 2] * a.z * sb;
        c.w -= sc[cl * CS + r4 + 3] * a.w * sb;
        *(float4*)cp = c;
      } else {
#pragma unroll
        for (int q = 0; q < 4; ++q)
          if (m + q < M) cp[q] -= sc[cl * CS + r4 + q] * ia[m + q] * sb;
      }
    }
  }
}

// Large trailing updates: the same product with workgroup tile 256 x 256, eight waves in a 2 x 4
// grid of 128 x 64 (8 x 4 blocks of v_mfma_f32_16x16x32_f16), one workgroup per CU. Operands move
// from memory to LDS by direct loads (global_load_lds_dwordx4: no register staging, no LDS store
// instructions), double-buffered in stages of 32 k (4 arrays x 256 rows x 32 halves x 2 stages =
// 128 KB of LDS). Per MAC this reads a third less LDS than the 128 x 128 tile and half the operand
// bytes from L2. A direct load writes lane l's 16 bytes at LDS offset 16 l of the instruction's
// 1 KB block, so the lane -> (row, chunk) assignment sets the layout: 4 lanes per row (64
// contiguous bytes from memory), logical chunk c of row r at physical chunk c ^ ((r / 4) % 4), so
// the 16 rows a half-wave fragment read touches fall in distinct banks. Rows past M (N) load the
// last valid row instead of zeros; they only feed accumulators that are never stored. Every
// accumulator sees the three products in the same order and K steps as k_gemm_f16x3, so both
// kernels give identical results.
typedef __attribute__((address_space(3))) void lds_void_t;
constexpr int kBigT = 256;                // tile rows and columns
constexpr int kBigK = 32;                 // k per LDS stage
constexpr int kBigArr = kBigT * kBigK;    // halves per operand array and stage
constexpr int kBigWaves = 8;
constexpr int kBigQ = 16 / kBigWaves;     // 16-row blocks each wave loads per array and stage
template <int PROD>
__global__ __launch_bounds__(64 * kBigWaves, 1) void k_gemm_f16x3_256(
    long M, long N, const _Float16* __restrict__ Ah, const _Float16* __restrict__ Al,
    const _Float16* __restrict__ Bh, const _Float16* __restrict__ Bl, const float* __restrict__ ia,
    const float* __restrict__ ib, float* __restrict__ C, long ldc) {
  __shared__ __attribute__((aligned(16))) _Float16 smem[2 * 4 * kBigArr];
  const long mt = (M + kBigT - 1) / kBigT, nt = (N + kBigT - 1) / kBigT;
  // persistent: workgroup b takes tiles b, b + G, b + 2G, ... (G = grid size, a multiple of the
  // XCD count, so every tile of a workgroup maps to its own XCD's block of tiles)
  // tile_of maps L to XCD L % 8's contiguous range only for L below the padded tile count
  const long Lend = ((mt * nt + kXcds - 1) / kXcds) * kXcds;
  long L = blockIdx.x, tm, tn;
  if (L >= Lend || !tile_of(L, mt, nt, tm, tn)) return;
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int wm = (wave & 1) * 128, wn = (wave >> 1) * 64;
  const int lr = lane & 15, lk = lane >> 4;
  const char* ah0 = (const char*)Ah;
  const char* al0 = (const char*)Al;
  const char* bh0 = (const char*)Bh;
  const char* bl0 = (const char*)Bl;
  // byte offsets (from the array base, at k = 0) of the 16 bytes this lane loads per block q
  unsigned offA[kBigQ], offB[kBigQ];
  auto offsets = [&](long m0, long n0) {
#pragma unroll
    for (int q = 0; q < kBigQ; ++q) {
      const int row = wave * (kBigQ * 16) + q * 16 + (lane >> 2);
      const int lc = (lane & 3) ^ ((row >> 2) & 3);
      const long ga = m0 + row < M ? m0 + row : M - 1, gb = n0 + row < N ? n0 + row : N - 1;
      offA[q] = (unsigned)(ga * kSplitPadK + lc * 8) * 2u;
      offB[q] = (unsigned)(gb * kSplitPadK + lc * 8) * 2u;
    }
  };
  auto issue = [&](int st, int k0) {
    const unsigned kb = (unsigned)k0 * 2u;
    _Float16* base = smem + st * 4 * kBigArr + wave * kBigQ * 512;
#pragma unroll
    for (int q = 0; q < kBigQ; ++q) {
      __builtin_amdgcn_global_load_lds((const void*)(ah0 + (offA[q] + kb)),
                                       (lds_void_t*)(base + 0 * kBigArr + q * 512), 16, 0, 0);
      if constexpr (has_lo<PROD>())
        __builtin_amdgcn_global_load_lds((const void*)(al0 + (offA[q] + kb)),
                                         (lds_void_t*)(base + 1 * kBigArr + q * 512), 16, 0, 0);
      __builtin_amdgcn_global_load_lds((const void*)(bh0 + (offB[q] + kb)),
                                       (lds_void_t*)(base + 2 * kBigArr + q * 512), 16, 0, 0);
      if constexpr (has_lo<PROD>())
        __builtin_amdgcn_global_load_lds((const void*)(bl0 + (offB[q] + kb)),
                                         (lds_void_t*)(base + 3 * kBigArr + q * 512), 16, 0, 0);
    }
  };
  const int off = lr * kBigK + ((lk ^ (lr >> 2)) * 8);  // this lane's fragment within a 16-row block
  constexpr int NCH = kSplitPadK / kBigK;
  // s_waitcnt vmcnt(0) (expcnt and lgkmcnt left at their maxima): this wave's direct loads landed
  constexpr int kWaitVm0 = 0x0F70;
  const bool al16 = ((ldc & 3) == 0) && ((((size_t)C) & 15) == 0);
  offsets(tm * kBigT, tn * kBigT);
  issue(0, 0);
  for (;;) {
    const long m0 = tm * kBigT, n0 = tn * kBigT;
    __builtin_amdgcn_s_waitcnt(kWaitVm0);
    __syncthreads();
    v4f acc[8][4];
#pragma unroll
    for (int i = 0; i < 8; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) acc[i][j] = v4f{0, 0, 0, 0};
    for (int kc = 0; kc < NCH; ++kc) {
      if (kc + 1 < NCH) issue((kc + 1) & 1, (kc + 1) * kBigK);  // the other stage, freed by the barrier
      const _Float16* S = smem + (kc & 1) * 4 * kBigArr;
      half8 bh[4], bl[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        bh[j] = *(const half8*)(S + 2 * kBigArr + (wn + 16 * j) * kBigK + off);
        if constexpr (has_lo<PROD>()) bl[j] = *(const half8*)(S + 3 * kBigArr + (wn + 16 * j) * kBigK + off);
      }
#pragma unroll
      for (int i = 0; i < 8; ++i) {
        const half8 ah = *(const half8*)(S + (wm + 16 * i) * kBigK + off);
        if constexpr (has_lo<PROD>()) {
          const half8 al = *(const half8*)(S + kBigArr + (wm + 16 * i) * kBigK + off);
#pragma unroll
          for (int j = 0; j < 4; ++j) acc[i][j] = mma16<PROD>(al, bh[j], acc[i][j]);
#pragma unroll
          for (int j = 0; j < 4; ++j) acc[i][j] = mma16<PROD>(ah, bl[j], acc[i][j]);
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = mma16<PROD>(ah, bh[j], acc[i][j]);
      }
      __builtin_amdgcn_s_waitcnt(kWaitVm0);
      __syncthreads();
    }
    // both stages are free (barrier above): the next tile's first stage loads under this epilogue
    L += gridDim.x;
    long tm2 = 0, tn2 = 0;
    const bool more = L < Lend && tile_of(L, mt, nt, tm2, tn2);
    if (more) {
      offsets(tm2 * kBigT, tn2 * kBigT);
      issue(0, 0);
    }
    // accumulator register v of lane l in block (i, j): row m0 + wm + 16 i + 4 (l / 16) + v, column
    // n0 + wn + 16 j + l % 16 -- four consecutive rows of one column, one 16-byte access
    if (al16 && m0 + kBigT <= M && n0 + kBigT <= N) {
      // whole tile: the next column block's C loads are in flight while the current one is updated
      float4 a4[8];
      float sbv[4];
#pragma unroll
      for (int i = 0; i < 8; ++i) a4[i] = *(const float4*)(ia + m0 + wm + 16 * i + 4 * lk);
#pragma unroll
      for (int j = 0; j < 4; ++j) sbv[j] = ib[n0 + wn + 16 * j + lr];
      float* cb = C + (n0 + wn + lr) * ldc + m0 + wm + 4 * lk;
      float4 cur[8], nxt[8];
#pragma unroll
      for (int i = 0; i < 8; ++i) cur[i] = *(const float4*)(cb + 16 * i);
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        if (j + 1 < 4) {
#pragma unroll
          for (int i = 0; i < 8; ++i) nxt[i] = *(const float4*)(cb + (long)(16 * (j + 1)) * ldc + 16 * i);
        }
#pragma unroll
        for (int i = 0; i < 8; ++i) {
          float4 c = cur[i];
          c.x -= acc[i][j][0] * a4[i].x * sbv[j];
          c.y -= acc[i][j][1] * a4[i].y * sbv[j];
          c.z -= acc[i][j][2] * a4[i].z * sbv[j];
          c.w -= acc[i][j][3] * a4[i].w * sbv[j];
          *(float4*)(cb + (long)(16 * j) * ldc + 16 * i) = c;
        }
        if (j + 1 < 4) {
#pragma unroll
          for (int i = 0; i < 8; ++i) cur[i] = nxt[i];
        }
      }
    } else {
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const long nn = n0 + wn + 16 * j + lr;
        if (nn >= N) continue;
        const float sb = ib[nn];
#pragma unroll
        for (int i = 0; i < 8; ++i) {
          const long m = m0 + wm + 16 * i + 4 * lk;
          float* cp = C + nn * ldc + m;
          if (al16 && m + 3 < M) {
            float4 c = *(float4*)cp;
            const float4 a = *(const float4*)(ia + m);
            c.x -= acc[i][j][0] * a.x * sb;
            c.y -= acc[i][j][1] * a.y * sb;
            c.z -= acc[i][j][2] * a.z * sb;
            c.w -= acc[i][j][3] * a.w * sb;
            *(float4*)cp = c;
          } else {
#pragma unroll
            for (int v = 0; v < 4; ++v)
              if (m + v < M) cp[v] -= acc[i][j][v] * ia[m + v] * sb;
          }
        }
      }
    }
    if (!more) break;
    // the offsets are recomputed rather than kept live through the epilogue (register pressure)
    asm volatile("" : "+s"(tm2), "+s"(tn2));
    tm = tm2;
    tn = tn2;
    offsets(tm * kBigT, tn * kBigT);
  }
}

// Workspace in doubles for split operands of up to n rows / columns
size_t f16x3_workspace_doubles(long n) {
  return ((size_t)4 * n * kSplitPadK * sizeof(_Float16) + 2 * n * sizeof(float)) / sizeof(double) + 16;
}
// The FP16-mode LU's workspace: the bulk trailing updates' split operands and the panel's
size_t lu_f16_workspace_doubles(long n) { return 2 * f16x3_workspace_doubles(n); }

// Split operands in the workspace (capacity cap rows / columns)
struct F16x3Ws {
  _Float16 *Ah, *Al, *Bh, *Bl;
  float *ia, *ib;
};
static F16x3Ws f16x3_ws(double* ws, long cap) {
  F16x3Ws w;
  w.Ah = (_Float16*)ws;
  w.Al = w.Ah + (size_t)cap * kSplitPadK;
  w.Bh = w.Al + (size_t)cap * kSplitPadK;
  w.Bl = w.Bh + (size_t)cap * kSplitPadK;
  w.ia = (float*)(w.Bl + (size_t)cap * kSplitPadK);
  w.ib = w.ia + cap;
  return w;
}
// rows of A (M x K block) into w.Ah / w.Al / w.ia; columns of B (K x N block) into the split-B
// slots [c0, c0 + N)
// prod: kProdF16x3 (hi and lo), kProdF16 or kProdBF16 (hi only)
static void f16x3_split_rows(long M, int K, const float* A, long lda, const F16x3Ws& w, hipStream_t st,
                             int kpad = kSplitPadK, int prod = kProdF16x3) {
  const dim3 g((unsigned)((M + kSplitRW - 1) / kSplitRW)), b(256);
  if (prod == kProdF16)
    hipLaunchKernelGGL(k_split_rows<kProdF16>, g, b, 0, st, M, K, kpad, A, lda, w.Ah, w.Al, w.ia);
  else if (prod == kProdBF16)
    hipLaunchKernelGGL(k_split_rows<kProdBF16>, g, b, 0, st, M, K, kpad, A, lda, w.Ah, w.Al, w.ia);
  else if (prod == kProdBF16x3)
    hipLaunchKernelGGL(k_split_rows<kProdBF16x3>, g, b, 0, st, M, K, kpad, A, lda, w.Ah, w.Al, w.ia);
  else
    hipLaunchKernelGGL(k_split_rows<kProdF16x3>, g, b, 0, st, M, K, kpad, A, lda, w.Ah, w.Al, w.ia);
}
static void f16x3_split_cols(long N, int K, const float* B, long ldb, const F16x3Ws& w, long c0,
                             hipStream_t st, int kpad = kSplitPadK, int prod = kProdF16x3) {
  if (N <= 0) return;
  const dim3 g((unsigned)((N + 3) / 4)), b(256);
  _Float16 *h = w.Bh + c0 * kSplitPadK, *l = w.Bl + c0 * kSplitPadK;
  if (prod == kProdF16)
    hipLaunchKernelGGL(k_split_cols<kProdF16>, g, b, 0, st, N, K, kpad, B, ldb, h, l, w.ib + c0);
  else if (prod == kProdBF16)
    hipLaunchKernelGGL(k_split_cols<kProdBF16>, g, b, 0, st, N, K, kpad, B, ldb, h, l, w.ib + c0);
  else if (prod == kProdBF16x3)
    hipLaunchKernelGGL(k_split_cols<kProdBF16x3>, g, b, 0, st, N, K, kpad, B, ldb, h, l, w.ib + c0);
  else
    hipLaunchKernelGGL(k_split_cols<kProdF16x3>, g, b, 0, st, N, K, kpad, B, ldb, h, l, w.ib + c0);
}
static void f16x3_split(long M, long N, int K, const float* A, long lda, const float* B, long ldb,
                        const F16x3Ws& w, hipStream_t st, int prod = kProdF16x3) {
  f16x3_split_rows(M, K, A, lda, w, st, kSplitPadK, prod);
  f16x3_split_cols(N, K, B, ldb, w, 0, st, kSplitPadK, prod);
}
// CUs of the current device (cached per device; 256 if the query fails)
static int device_cus() {
  static int cus[64] = {0};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0) dev = 0;
  int* slot = dev < 64 ? &cus[dev] : nullptr;
  if (slot && *slot > 0) return *slot;
  int n = 0;
  if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n <= 0) n = 256;
  if (slot) *slot = n;
  return n;
}
// Workgroups of the persistent 256 x 256 kernel: one per CU the stream may use (cus, 0 = all)
static unsigned big_gemm_grid(long tiles, int cus) {
  long g = (long)((cus > 0 ? cus : device_cus()) / kXcds) * kXcds;
  const long padded = ((tiles + kXcds - 1) / kXcds) * kXcds;
  if (g < kXcds) g = kXcds;
  return (unsigned)(g < padded ? g : padded);
}
#ifndef IBEM_BIG_MIN_TILES
#define IBEM_BIG_MIN_TILES 512
#endif
// Runtime tunables (process-wide; set through ibem3d_set_tuning before the work they shape is
// launched or captured): the 256 x 256 tile count from which trailing updates take the persistent
// kernel, and the share (percent) of the panel_cus schedule the bulk update leaves to the panel.
static long g_big_min_tiles = IBEM_BIG_MIN_TILES;
static long g_panel_cu_percent = 100;
static long g_inner_iters = 6;     // inner GMRES iterations on the 16-bit A per outer direction
static long g_inner_tol_exp = 3;   // inner GMRES stops at relative residual 10^-exp
static long g_rows_pointwise = 0;  // 1: per-Gauss-point boundary->inclusion rows on planar elements too
static long g_bj_nopiv = 1;        // block-Jacobi boundary block: 1 no interchanges, 0 pivoting, 2 fallback
static long g_bj_overlap = 1;      // 1: its triangular solves beside the FP16 product
static long g_stress_beside = 1;   // 1: stresses beside each GMRES cycle's true residual
static long g_bj_fp32 = 1;         // 1: the interchange-free boundary factor in FP32
static long g_res_update = 1;      // 1: a cycle's residual from its own products when it converges
int set_tuning(int key, long value) {
  if (key == kTuneBigMinTiles && value >= 1) {
    g_big_min_tiles = value;
    return 0;
  }
  if (key == kTunePanelCuPercent && value >= 0 && value <= 400) {
    g_panel_cu_percent = value;
    return 0;
  }
  if (key == kTuneInnerIters && value >= 1 && value <= kInnerItersMax) {
    g_inner_iters = value;
    return 0;
  }
  if (key == kTuneInnerTolExp && value >= 1 && value <= 8) {
    g_inner_tol_exp = value;
    return 0;
  }
  if (key == kTuneRowsPointwise && (value == 0 || value == 1)) {
    g_rows_pointwise = value;
    return 0;
  }
  if (key == kTuneBjNopiv && value >= 0 && value <= 2) {
    g_bj_nopiv = value;
    return 0;
  }
  if (key == kTuneBjOverlap && (value == 0 || value == 1)) {
    g_bj_overlap = value;
    return 0;
  }
  if (key == kTuneStressBeside && (value == 0 || value == 1)) {
    g_stress_beside = value;
    return 0;
  }
  if (key == kTuneBjFp32 && (value == 0 || value == 1)) {
    g_bj_fp32 = value;
    return 0;
  }
  if (key == kTuneResidualUpdate && (value == 0 || value == 1)) {
    g_res_update = value;
    return 0;
  }
  return -1;
}
long get_tuning(int key) {
  switch (key) {
    case kTuneBigMinTiles: return g_big_min_tiles;
    case kTunePanelCuPercent: return g_panel_cu_percent;
    case kTuneInnerIters: return g_inner_iters;
    case kTuneInnerTolExp: return g_inner_tol_exp;
    case kTuneRowsPointwise: return g_rows_pointwise;
    case kTuneBjNopiv: return g_bj_nopiv;
    case kTuneBjOverlap: return g_bj_overlap;
    case kTuneStressBeside: return g_stress_beside;
    case kTuneBjFp32: return g_bj_fp32;
    case kTuneResidualUpdate: return g_res_update;
    default: return -1;
  }
}
// C[M x N] -= split A times the split B columns [c0, c0 + N); the persistent 256 x 256 kernel runs
// on at most cus CUs (0: all)
template <int PROD>
static void f16x3_gemm_t(long M, long N, int K, const F16x3Ws& w, float* C, long ldc, hipStream_t st,
                         long c0, int cus) {
  // 256 x 256 tiles once there are at least two per CU (one workgroup per CU), else 128 x 128
  const long big = ((M + kBigT - 1) / kBigT) * ((N + kBigT - 1) / kBigT);
  if (big >= g_big_min_tiles) {  // operands are zero-padded to kSplitPadK: K is implied
    hipLaunchKernelGGL(k_gemm_f16x3_256<PROD>, dim3(big_gemm_grid(big, cus)),
                       dim3(64 * kBigWaves), 0, st, M, N, w.Ah, w.Al, w.Bh + c0 * kSplitPadK,
                       w.Bl + c0 * kSplitPadK, w.ia, w.ib + c0, C, ldc);
    return;
  }
  constexpr int NCH = kSplitPadK / IBEM_F16_BK;
  hipLaunchKernelGGL((k_gemm_f16x3<IBEM_F16_BK, IBEM_F16_EPI, IBEM_F16_PF, NCH, PROD>), dim3(tile_grid(M, N)),
                     dim3(256), 0, st, M, N, K, w.Ah, w.Al, w.Bh + c0 * kSplitPadK,
                     w.Bl + c0 * kSplitPadK, w.ia, w.ib + c0, C, ldc);
}
static void f16x3_gemm(long M, long N, int K, const F16x3Ws& w, float* C, long ldc, hipStream_t st,
                       long c0 = 0, int cus = 0, int prod = kProdF16x3) {
  if (prod == kProdF16)
    f16x3_gemm_t<kProdF16>(M, N, K, w, C, ldc, st, c0, cus);
  else if (prod == kProdBF16)
    f16x3_gemm_t<kProdBF16>(M, N, K, w, C, ldc, st, c0, cus);
  else if (prod == kProdBF16x3)
    f16x3_gemm_t<kProdBF16x3>(M, N, K, w, C, ldc, st, c0, cus);
  else
    f16x3_gemm_t<kProdF16x3>(M, N, K, w, C, ldc, st, c0, cus);
}

// C[M x N] -= split A times split B for a depth K <= 512 (operands split with kpad = K rounded up to
// 64): the 128 x 128 kernel reads only that depth
static void f16x3_gemm_short(long M, long N, int K, const F16x3Ws& w, float* C, long ldc, hipStream_t st) {
  const dim3 g(tile_grid(M, N)), b(256);
  if (K <= 256)
    hipLaunchKernelGGL((k_gemm_f16x3<64, 1, 1, 4>), g, b, 0, st, M, N, K, w.Ah, w.Al, w.Bh, w.Bl, w.ia, w.ib, C, ldc);
  else
    hipLaunchKernelGGL((k_gemm_f16x3<64, 1, 1, 8>), g, b, 0, st, M, N, K, w.Ah, w.Al, w.Bh, w.Bl, w.ia, w.ib, C, ldc);
}

// C[M x N] -= A[M x K] B[K x N] (column-major FP32, K <= 1024) through the split FP16 products.
void gemm_f32_sub_f16x3(long M, long N, int K, const float* A, long lda, const float* B, long ldb,
                        float* C, long ldc, double* ws, hipStream_t st, int prod) {
  if (M <= 0 || N <= 0 || K <= 0) return;
  const F16x3Ws w = f16x3_ws(ws, M > N ? M : N);
  f16x3_split(M, N, K, A, lda, B, ldb, w, st, prod);
  f16x3_gemm(M, N, K, w, C, ldc, st, 0, 0, prod);
}

// ======================================================= panel leaf: tournament pivoting (CALU) ==
// A leaf factors w <= 32 columns A[j0:n, j0:j0+w) in two launches, with no grid barrier:
//  k_ts_tree  every workgroup takes 256 panel rows (thread = row, the row's w values in registers)
//             and runs partial-pivoting elimination on them to pick w candidate pivot rows, whose
//             ORIGINAL values it writes out. The last workgroup of each group of 8 to finish
//             (atomic counter) merges the group's 8 x w candidates the same way, and so on up a
//             tree; the last one standing (the root) holds the w pivot rows in order with their
//             L11\U11 factors. It turns the pivot order into LAPACK-style sequential interchanges
//             (ipiv) and saves the original contents of the top rows that move down.
//  k_ts_apply every panel row: the top w rows receive L11\U11; every other row (with its new
//             content where an interchange moved one there) becomes its row of L21 = A21 U11^-1.
// Tournament pivoting (communication-avoiding LU, Grigori, Demmel and Xiang 2011) replaces the
// column-by-column global pivot search; it is backward stable in practice like partial pivoting.
// Pivot ties follow a fixed rule over a fixed tree shape, so results are deterministic.
constexpr int kPanelW = 32;   // leaf width
#ifndef IBEM_TS_ROWS
#define IBEM_TS_ROWS 256
#endif
constexpr int kTsRows = IBEM_TS_ROWS;       // rows per workgroup (one per thread) at every tree level
constexpr int kTsWaves = kTsRows / 64;      // wavefronts per tree node
constexpr int kTsFan = kTsRows / kPanelW;   // candidate sets merged per tree node

static long ts_nodes(long n) { return 2 * ((n + kTsRows - 1) / kTsRows) + 16; }
size_t lu_workspace_doubles(long n) {
  const long nodes = ts_nodes(n);
  return (size_t)nodes * kPanelW * kPanelW + 2 * kPanelW * kPanelW +
         ((size_t)nodes * kPanelW + (size_t)nodes + kPanelW + 8) / 2 + 1;
}
struct TsWs {
  double* cand;   // nodes x 32 x 32 candidate rows (original values)
  double* lu11;   // 32 x 32, root's L11\U11 rows in pivot order
  double* saved;  // 32 x 32, original contents of the top rows that move down
  int* cidx;      // nodes x 32 candidate row indices (-1: none)
  int* cnt;       // nodes arrival counters (zero between launches)
  int* dest;      // 32 positions receiving a saved row
  int* ndest;     // 1
};
static TsWs ts_ws(double* base, long n) {
  const long nodes = ts_nodes(n);
  TsWs w;
  w.cand = base;
  w.lu11 = base + (size_t)nodes * kPanelW * kPanelW;
  w.saved = w.lu11 + kPanelW * kPanelW;
  w.cidx = (int*)(w.saved + kPanelW * kPanelW);
  w.cnt = w.cidx + (size_t)nodes * kPanelW;
  w.dest = w.cnt + nodes;
  w.ndest = w.dest + kPanelW;
  return w;
}
// Pivot keys: 32 bits, larger wins, 0 = no candidate. A candidate's key is its magnitude's bits
// shifted to leave room for the +1 that keeps a zero value distinguishable from "no row": FP32 bits
// 30..1, FP64 bits 62..32 (exponent and 20 mantissa bits); both are monotonic in the magnitude.
// Magnitudes that agree in these bits tie, and ties go to the lowest lane, then the lowest wave:
// a fixed rule over a fixed tree shape, so results are deterministic.
__device__ inline unsigned ts_key(float v) { return (__float_as_uint(fabsf(v)) >> 1) + 1u; }
__device__ inline unsigned ts_key(double v) {
  return (unsigned)((unsigned long long)__double_as_longlong(fabs(v)) >> 32) + 1u;
}
// Columns [c, kPanelW) of this thread's row into its LDS slot, one element per store instruction.
// Written with the compiler's own stores, the row is gathered into aligned 16-byte register tuples,
// which ties a[] to a tuple and makes every later partial update of it copy the whole array. The
// caller waits for the stores (lgkmcnt) before its barrier.
template <typename T, int C>
__device__ __forceinline__ void ts_store_tail(T* dst, const T (&a)[kPanelW]) {
  const unsigned base = (unsigned)(size_t)dst;
#pragma unroll
  for (int k = C; k < kPanelW; ++k) {
    if constexpr (sizeof(T) == 4)
      asm volatile("ds_write_b32 %0, %1 offset:%2" ::"v"(base), "v"(a[k]), "i"(4 * k) : "memory");
    else
      asm volatile("ds_write_b64 %0, %1 offset:%2" ::"v"(base), "v"(a[k]), "i"(8 * k) : "memory");
  }
}
template <int CTRL>
__device__ inline unsigned dpp_max_u32(unsigned k) {
  return max(k, (unsigned)__builtin_amdgcn_update_dpp(0, (int)k, CTRL, 0xF, 0xF, false));
}
// maximum over the 64 lanes (all lanes active), in every lane: within each row of 16 lanes by DPP
// (quad permutations, half-row and row mirrors), then across the four rows by readlane
__device__ inline unsigned wave_max_u32(unsigned k) {
  k = dpp_max_u32<0xB1>(k);   // quad_perm [1,0,3,2]
  k = dpp_max_u32<0x4E>(k);   // quad_perm [2,3,0,1]
  k = dpp_max_u32<0x141>(k);  // row_half_mirror
  k = dpp_max_u32<0x140>(k);  // row_mirror
  const unsigned k0 = (unsigned)__builtin_amdgcn_readlane((int)k, 0);
  const unsigned k1 = (unsigned)__builtin_amdgcn_readlane((int)k, 16);
  const unsigned k2 = (unsigned)__builtin_amdgcn_readlane((int)k, 32);
  const unsigned k3 = (unsigned)__builtin_amdgcn_readlane((int)k, 48);
  return max(max(k0, k1), max(k2, k3));
}

// Partial-pivoting elimination of the workgroup's rows (thread = row), choosing up to w pivot rows.
// a: the row's values, updated in place (multipliers left of the pivot column, reduced values right
// of it); mystep: the step at which this row was chosen, or -1. Returns the number of steps that
// found a row. Per step: every unchosen row stores the columns the step can need in its own LDS
// slot (independent of the search, so the stores overlap it); a wave maximum of the pivot keys,
// packed with the lowest winning row index into one 64-bit word per wave; one barrier; every
// thread takes the largest of the four words (one LDS round trip) and eliminates with the winner's
// slot. The words alternate by step parity, so one barrier per step suffices. The update is
// branch-free: rows that do not take part use a zero multiplier (all values are finite), which
// leaves them unchanged.
constexpr int kTsStride = kPanelW + 1;  // LDS row slot stride (elements): consecutive rows in distinct banks
template <typename T, int C>
__device__ __forceinline__ void ts_step(T (&a)[kPanelW], bool valid, int w, unsigned long long* skey,
                                        T* sp, int& mystep, int& nsel) {
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
  constexpr int par = C & 1;
  const bool cand = valid && mystep < 0 && C < w;  // steps past w find no candidate
  const unsigned key = cand ? ts_key(a[C]) : 0u;
  const unsigned wmax = wave_max_u32(key);
  const unsigned long long hit = __ballot(key == wmax);
  const int wl = __ffsll((long long)hit) - 1;
  // larger key wins; equal keys go to the lower row index (lowest lane, then lowest wave)
  if (lane == 0) skey[par * kTsWaves + wv] = ((unsigned long long)wmax << 32) | (unsigned)(kTsRows - 1 - (wv * 64 + wl));
  // one slot per row: a chosen row stops writing, so the slot a step reads after its barrier (the
  // winner's) is never overwritten while other threads still read it
  if (valid && mystep < 0) ts_store_tail<T, C>(sp + t * kTsStride, a);
  asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
  __syncthreads();
  unsigned long long best = skey[par * kTsWaves];
#pragma unroll
  for (int q = 1; q < kTsWaves; ++q) {
    const unsigned long long k2 = skey[par * kTsWaves + q];
    best = k2 > best ? k2 : best;
  }
  // no branch on `found`: a column without candidates gives l = 0 everywhere (the slot read then
  // holds some row's finite values)
  const bool found = (best >> 32) != 0ull;
  const int win = kTsRows - 1 - (int)(best & 0xffffffffull);
  nsel = found ? C + 1 : nsel;
  if (found && t == win) mystep = C;
  const T* prow = sp + win * kTsStride;
  T pr[kPanelW];
#pragma unroll
  for (int k = C; k < kPanelW; ++k) pr[k] = prow[k];
  const bool act = found && valid && mystep < 0;
  const T piv = pr[C];
  const T l = (act && piv != T(0)) ? a[C] / piv : T(0);
  a[C] = act ? l : a[C];
#pragma unroll
  for (int k = C + 1; k < kPanelW; ++k) a[k] = fma(-l, pr[k], a[k]);
}
template <typename T, int C = 0>
__device__ __forceinline__ void ts_steps(T (&a)[kPanelW], bool valid, int w, unsigned long long* skey,
                                         T* sp, int& mystep, int& nsel) {
  if constexpr (C < kPanelW) {
    ts_step<T, C>(a, valid, w, skey, sp, mystep, nsel);
    ts_steps<T, C + 1>(a, valid, w, skey, sp, mystep, nsel);
  }
}
template <typename T>
__device__ __forceinline__ int ts_select(T (&a)[kPanelW], bool valid, int w, unsigned long long* skey,
                                         T* sp, int& mystep) {
  mystep = -1;
  int nsel = 0;
  ts_steps<T>(a, valid, w, skey, sp, mystep, nsel);
  return nsel;
}

// position -> original row held there, over the <= 2w positions the interchanges touch
__device__ inline int ts_who(const int* pos, const int* who, int np, int p) {
  for (int q = 0; q < np; ++q)
    if (pos[q] == p) return who[q];
  return p;
}
__device__ inline void ts_set(int* pos, int* who, int& np, int p, int v) {
  for (int q = 0; q < np; ++q)
    if (pos[q] == p) {
      who[q] = v;
      return;
    }
  pos[np] = p;
  who[np] = v;
  ++np;
}

template <typename T>
__global__ __launch_bounds__(kTsRows) void k_ts_tree(const T* __restrict__ A, long lda, long n,
                                                     long j0, int w, int P, TsWs ws,
                                                     int* __restrict__ ipiv, int* __restrict__ info) {
  __shared__ __align__(16) unsigned long long skey[2 * kTsWaves];
  __shared__ __align__(16) T sp[kTsRows * kTsStride];
  __shared__ int s_last;
  __shared__ int s_win[kPanelW];
  __shared__ double s_diag[kPanelW];
  __shared__ int s_pos[2 * kPanelW], s_who[2 * kPanelW];
  __shared__ int s_nd, s_nsel;
  __shared__ int s_src[kPanelW];
  const int t = threadIdx.x;
  int node = blockIdx.x, count = P;
  long off = 0;  // first node slot of the current tree level
  const long r = j0 + (long)node * kTsRows + t;
  bool valid = r < n;
  int gidx = valid ? (int)r : -1;
  T a[kPanelW];
#pragma unroll
  for (int c = 0; c < kPanelW; ++c) a[c] = (valid && c < w) ? A[(j0 + c) * lda + r] : T(0);
  // where this row's original values live: -1 = in A (first level), else a candidate slot
  long src = -1;
  int mystep = -1, nsel = 0;
  for (;;) {
    nsel = ts_select<T>(a, valid, w, skey, sp, mystep);
    if (count == 1) break;  // this workgroup holds the root set
    const long slot = off + node;
    if (mystep >= 0) {
      // the row's original values: all loads first, then the stores (no serialised round trips)
      double v[kPanelW];
#pragma unroll
      for (int c = 0; c < kPanelW; ++c)
        v[c] = c < w ? (src < 0 ? (double)A[(j0 + c) * lda + gidx] : ws.cand[src * kPanelW + c]) : 0.0;
      double* dst = ws.cand + (slot * kPanelW + mystep) * kPanelW;
#pragma unroll
      for (int c = 0; c < kPanelW; ++c) dst[c] = v[c];
      ws.cidx[slot * kPanelW + mystep] = gidx;
    }
    if (t >= nsel && t < kPanelW) ws.cidx[slot * kPanelW + t] = -1;
    // release: every wavefront waits for its own candidate stores and writes them back to the
    // device-coherent level before the barrier, so thread 0's arrival count below is ordered after
    // all of them (the per-XCD L2s are not coherent with each other)
    __threadfence();
    __syncthreads();
    const int group = node / kTsFan;
    const int gsize = min(kTsFan, count - group * kTsFan);
    const long next_off = off + count;
    if (t == 0) {
      const int old = atomicAdd(&ws.cnt[next_off + group], 1);
      s_last = old == gsize - 1;
    }
    __syncthreads();
    if (!s_last) return;
    __threadfence();  // acquire, in every wavefront: the group's candidates
    if (t == 0) ws.cnt[next_off + group] = 0;  // ready for the next leaf (a later launch)
    __syncthreads();
    const int child = group * kTsFan + t / kPanelW, k = t % kPanelW;
    valid = (t / kPanelW) < gsize && k < w;
    gidx = valid ? ws.cidx[(off + child) * kPanelW + k] : -1;
    valid = valid && gidx >= 0;
    src = (off + child) * kPanelW + k;
#pragma unroll
    for (int c = 0; c < kPanelW; ++c) a[c] = (valid && c < w) ? (T)ws.cand[src * kPanelW + c] : T(0);
    node = group;
    off = next_off;
    count = (count + kTsFan - 1) / kTsFan;
  }
  // ---- root: the pivot rows in order with their L11\U11 ----
  if (mystep >= 0) {
#pragma unroll
    for (int c = 0; c < kPanelW; ++c)
      if (c < w) ws.lu11[mystep * kPanelW + c] = (double)a[c];
    s_win[mystep] = gidx;
    double d = 0.0;
#pragma unroll
    for (int c = 0; c < kPanelW; ++c)
      if (c == mystep) d = (double)a[c];
    s_diag[mystep] = d;
  }
  if (t == 0) s_nsel = nsel;
  __syncthreads();
  if (t < 64) {
    // wave 0: lane q keeps slot q of the table (position -> original row) of the <= 2w positions
    // the interchanges touch; lookups are ballots, so each step costs a few wave instructions
    const int q = t, ns = s_nsel;
    const unsigned long long zero_piv =
        __ballot(q < w && (q >= ns || s_diag[q < kPanelW ? q : 0] == 0.0));
    if (zero_piv && q == 0 && *info == 0) *info = (int)(j0 + __ffsll((long long)zero_piv));
    int pos = -1, who = -1, np = 0;
    for (int i = 0; i < w; ++i) {
      const int top = (int)(j0 + i);
      int p = top;
      if (i < ns) {
        const int rw = s_win[i];
        const unsigned long long mw = __ballot(q < np && who == rw);
        p = mw ? __shfl(pos, __ffsll((long long)mw) - 1) : rw;
      }
      if (q == 0) ipiv[top] = p;
      if (p != top) {
        const unsigned long long mt = __ballot(q < np && pos == top);
        const unsigned long long mp = __ballot(q < np && pos == p);
        const int w1 = mt ? __shfl(who, __ffsll((long long)mt) - 1) : top;
        const int w2 = mp ? __shfl(who, __ffsll((long long)mp) - 1) : p;
        int lt = mt ? __ffsll((long long)mt) - 1 : np;
        if (!mt) ++np;
        int lp = mp ? __ffsll((long long)mp) - 1 : np;
        if (!mp) ++np;
        if (q == lt) {
          pos = top;
          who = w2;
        }
        if (q == lp) {
          pos = p;
          who = w1;
        }
      }
    }
    // positions below the leaf's top rows that now hold a (displaced) top row
    const bool moved = q < np && pos >= j0 + w && who != pos;
    const unsigned long long mm = __ballot(moved);
    const int e = __popcll(mm & ((1ull << q) - 1ull));
    if (moved) {
      ws.dest[e] = pos;
      s_src[e] = who;
    }
    if (q == 0) {
      *ws.ndest = __popcll(mm);
      s_nd = __popcll(mm);
    }
  }
  __syncthreads();
  for (int u = t; u < s_nd * kPanelW; u += kTsRows) {
    const int e = u / kPanelW, c = u % kPanelW;
    if (c < w) ws.saved[e * kPanelW + c] = (double)A[(j0 + c) * lda + s_src[e]];
  }
}

template <typename T>
__global__ __launch_bounds__(kTsRows) void k_ts_apply(T* __restrict__ A, long lda, long n, long j0,
                                                      int w, TsWs ws) {
  __shared__ T sLU[kPanelW][kPanelW + 1];            // root rows in pivot order: L11\U11
  __shared__ __align__(16) T sUt[kPanelW][kPanelW];  // sUt[c][k] = U11(k, c), k < c (else 0)
  __shared__ T sR[kPanelW];                          // 1 / U11(c, c) (1 where the pivot is zero)
  __shared__ int sd[kPanelW];
  __shared__ int snd;
  const int t = threadIdx.x;
  for (int u = t; u < kPanelW * kPanelW; u += kTsRows) {
    const int i = u / kPanelW, c = u % kPanelW;
    const T v = (i < w && c < w) ? (T)ws.lu11[i * kPanelW + c] : T(0);
    sLU[i][c] = v;
    sUt[c][i] = i < c ? v : T(0);
  }
  if (t < kPanelW) {
    const T d = t < w ? (T)ws.lu11[t * kPanelW + t] : T(0);
    sR[t] = d != T(0) ? T(1) / d : T(1);
    sd[t] = ws.dest[t];
  }
  if (t == 0) snd = *ws.ndest;
  __syncthreads();
  const long r = j0 + (long)blockIdx.x * kTsRows + t;
  if (r >= n) return;
  if (r < j0 + w) {
    const int i = (int)(r - j0);
#pragma unroll
    for (int c = 0; c < kPanelW; ++c)
      if (c < w) A[(j0 + c) * lda + r] = sLU[i][c];
    return;
  }
  int e = -1;
  for (int q = 0; q < snd; ++q)
    if (sd[q] == r) e = q;
  // columns past w read column w - 1 (inside the panel) and are zeroed: no per-column branches
  T* col = A + j0 * lda + r;
  T a[kPanelW];
  if (e >= 0) {
#pragma unroll
    for (int c = 0; c < kPanelW; ++c) {
      const T v = (T)ws.saved[e * kPanelW + min(c, w - 1)];
      a[c] = c < w ? v : T(0);
    }
  } else {
#pragma unroll
    for (int c = 0; c < kPanelW; ++c) {
      const T v = col[(long)min(c, w - 1) * lda];
      a[c] = c < w ? v : T(0);
    }
  }
  // the row of L21 = a U11^-1: l_c = (a_c - sum_{k<c} l_k U11(k, c)) / U11(c, c), with two partial
  // sums to halve the dependent chain (entries past w stay zero). Column c's LDS row address passes
  // through an empty asm that consumes l_{c-2}: otherwise every U read is hoisted to the top and
  // the 496 values spill.
  typedef __attribute__((address_space(3))) const T LdsT;
#pragma unroll
  for (int c = 0; c < kPanelW; ++c) {
    LdsT* ut = (LdsT*)&sUt[c][0];
    if (c >= 2) asm volatile("" : "+v"(ut) : "v"(a[c - 2]));
    T x0 = a[c], x1 = T(0);
#pragma unroll
    for (int k = 0; k + 1 < c; k += 2) {
      x0 = fma(-a[k], ut[k], x0);
      x1 = fma(-a[k + 1], ut[k + 1], x1);
    }
    if (c & 1) x0 = fma(-a[c - 1], ut[c - 1], x0);
    a[c] = (x0 + x1) * sR[c];
  }
  if (w == kPanelW) {
#pragma unroll
    for (int c = 0; c < kPanelW; ++c) col[(long)c * lda] = a[c];
  } else {
#pragma unroll
    for (int c = 0; c < kPanelW; ++c)
      if (c < w) col[(long)c * lda] = a[c];
  }
}

// ============================================================================ row swaps ==========
// Apply the interchanges ipiv[p0..p1) in order to columns [c0, c0 + nc). One thread per column
// walking them in order would be bound by its dependent memory round trips (a 1,024-interchange
// step took ~0.9 ms per call that way on wide trailing matrices). Every workgroup composes
// ipiv[p0, p0 + m) into moves A[dst] <- A[src] over the positions they touch, then moves its
// columns, kLaswpCG at a time: all reads, a barrier, all writes.
//
// The composition runs in parallel, from ipiv[p0 + i] >= p0 + i (step i swaps row p0 + i with a row
// at or below it, so row p0 + i is final after step i). With c(j) the original row whose entry sits
// at row p0 + j just before step j:
//   c(j) = c(prev(j)) for prev(j) = the last step j' < j with ipiv[p0 + j'] = p0 + j, else p0 + j,
// so c(j) = p0 + root(j) along the prev chain (pointer jumping, log2 m rounds). Then
//   row p0 + i ends with c(i) if ipiv[p0 + i] = p0 + i, else with what row q = ipiv[p0 + i] held
//   just before step i: c(j) for the previous step j < i with the same pivot row q, else q itself;
//   a row q >= p0 + m ends with c(j) of the last step j with pivot row q.
// "Previous step with the same pivot row" comes from one bitonic sort of (pivot row, step) keys.
// About 70 barriers at m = 1024, where composing the interchanges one after another in LDS was a
// chain of 1024 dependent read-write rounds (~60 us per workgroup).
constexpr int kLaswpCG = 4;    // columns per read/write round
// Move list of ipiv[p0, p0 + m), m <= IBEM_LU_NB (shared by the workgroup): returns the number of
// moves A[mdst[q]] <- A[msrc[q]]. Ends with a barrier.
struct LaswpLds {
  unsigned long long key[IBEM_LU_NB];  // (pivot row << 32 | step), sorted
  int prv[IBEM_LU_NB];                 // prev(j), then root(j) by pointer jumping
  int pocc[IBEM_LU_NB];                // previous step with the same pivot row, or -1
  int mdst[2 * IBEM_LU_NB], msrc[2 * IBEM_LU_NB];
  int nmv;
};
constexpr int kLaswpPer = IBEM_LU_NB / 256;  // steps per thread
static_assert(IBEM_LU_NB % 256 == 0 && (IBEM_LU_NB & (IBEM_LU_NB - 1)) == 0, "sort and step layout");
__device__ int laswp_moves(const int* __restrict__ ipiv, long p0, int m, LaswpLds& S) {
  const int t = threadIdx.x;
  const int b0 = (int)p0;
  int m2 = 1;
  while (m2 < m) m2 <<= 1;
  for (int i = t; i < m2; i += 256) {
    S.key[i] = i < m ? ((unsigned long long)(unsigned)ipiv[p0 + i] << 32) | (unsigned)i : ~0ull;
    if (i < m) S.prv[i] = -1;
  }
  if (t == 0) S.nmv = 0;
  __syncthreads();
  for (int i = t; i < m; i += 256) {
    const int p = (int)(S.key[i] >> 32);
    if (p > b0 + i && p < b0 + m) atomicMax(&S.prv[p - b0], i);
  }
  // ascending bitonic sort of the keys (m2 - m padding keys sort last)
  for (int k = 2; k <= m2; k <<= 1)
    for (int j = k >> 1; j > 0; j >>= 1) {
      __syncthreads();
      for (int u = t; u < (m2 >> 1); u += 256) {
        const int a = ((u & ~(j - 1)) << 1) | (u & (j - 1)), b = a + j;
        const unsigned long long ka = S.key[a], kb = S.key[b];
        if ((ka > kb) == ((a & k) == 0)) {
          S.key[a] = kb;
          S.key[b] = ka;
        }
      }
    }
  __syncthreads();
  // roots of the prev chains (prev(j) < j): prv[j] <- root(j)
  for (int j = t; j < m; j += 256)
    if (S.prv[j] < 0) S.prv[j] = j;
  for (int span = 1; span < m; span <<= 1) {
    __syncthreads();
    int v[kLaswpPer];
#pragma unroll
    for (int u = 0; u < kLaswpPer; ++u) {
      const int j = t + 256 * u;
      v[u] = j < m ? S.prv[S.prv[j]] : 0;
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < kLaswpPer; ++u) {
      const int j = t + 256 * u;
      if (j < m) S.prv[j] = v[u];
    }
  }
  __syncthreads();
  for (int s = t; s < m; s += 256) {
    const unsigned long long kk = S.key[s];
    const int q = (int)(kk >> 32), i = (int)(kk & 0xffffffffull);
    const bool first = s == 0 || (int)(S.key[s - 1] >> 32) != q;
    const bool last = s == m - 1 || (int)(S.key[s + 1] >> 32) != q;
    S.pocc[i] = first ? -1 : (int)(S.key[s - 1] & 0xffffffffull);
    if (last && q >= b0 + m) {  // a row below the range ends with its last step's entry
      const int w = atomicAdd(&S.nmv, 1);
      S.mdst[w] = q;
      S.msrc[w] = b0 + S.prv[i];
    }
  }
  __syncthreads();
  for (int i = t; i < m; i += 256) {
    const int q = ipiv[p0 + i];
    int src;
    if (q == b0 + i) {
      src = b0 + S.prv[i];
    } else {
      const int j = S.pocc[i];
      src = j >= 0 ? b0 + S.prv[j] : q;
    }
    if (src != b0 + i) {
      const int w = atomicAdd(&S.nmv, 1);
      S.mdst[w] = b0 + i;
      S.msrc[w] = src;
    }
  }
  __syncthreads();
  return S.nmv;
}
// The moves applied to columns [g, min(g + kLaswpCG, ce)): all reads, a barrier, all writes, a
// barrier (the next group reads the same rows).
template <typename T>
__device__ __forceinline__ void laswp_apply_group(T* __restrict__ A, long lda, long g, long ce, int nm,
                                                  const LaswpLds& S) {
  const int t = threadIdx.x;
  T v[kLaswpCG][8];
#pragma unroll
  for (int j = 0; j < kLaswpCG; ++j) {
    const T* col = A + min(g + j, ce - 1) * lda;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const int q = t + 256 * k;
      v[j][k] = q < nm ? col[S.msrc[q]] : T(0);
    }
  }
  __builtin_amdgcn_s_waitcnt(0);  // every read of the group has returned before the barrier
  __syncthreads();
#pragma unroll
  for (int j = 0; j < kLaswpCG; ++j) {
    if (g + j < ce) {
      T* col = A + (g + j) * lda;
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        const int q = t + 256 * k;
        if (q < nm) col[S.mdst[q]] = v[j][k];
      }
    }
  }
  __syncthreads();
}
template <typename T>
__global__ __launch_bounds__(256) void k_laswp_mv(T* __restrict__ A, long lda, long c0, long nc,
                                                  const int* __restrict__ ipiv, long p0, int m,
                                                  int cpb) {
  static_assert(2 * IBEM_LU_NB <= 8 * 256, "moves per column <= 8 per thread");
  __shared__ LaswpLds S;
  const int nm = laswp_moves(ipiv, p0, m, S);
  const long cb = c0 + (long)blockIdx.x * cpb, ce = min(cb + cpb, c0 + nc);
  for (long g = cb; g < ce; g += kLaswpCG) laswp_apply_group(A, lda, g, ce, nm, S);
}

template <typename T>
static void laswp(T* A, long lda, long c0, long nc, const int* ipiv, long p0, long p1,
                         hipStream_t st) {
  if (nc <= 0 || p1 <= p0) return;
  // ~1024 workgroups on wide ranges (four per CU: each group's read round trip overlaps the others')
  const long cpb = std::max<long>(kLaswpCG, (nc + 1023) / 1024 / kLaswpCG * kLaswpCG);
  hipLaunchKernelGGL(k_laswp_mv<T>, dim3((unsigned)((nc + cpb - 1) / cpb)), dim3(256), 0, st, A, lda,
                     c0, nc, ipiv, p0, (int)(p1 - p0), (int)cpb);
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

// B[m x N] <- L^{-1} B for a unit-lower L with m <= 64 Q, for the few, tall-ish right-hand sides
// inside a panel (N up to a few hundred): one wavefront per column of B, lane l holding rows
// l + 64 q in registers. Step i broadcasts x_i from its lane (readlane) and every lane updates its
// rows below i with column i of L, which the workgroup stages through LDS in chunks of CW columns
// (only rows below the current 64-row block are staged and updated). Every index is a compile-time
// constant, so the whole solve is straight-line code.

// a lane's value, uniform across the wavefront (the lane index is uniform)
template <typename T>
__device__ __forceinline__ T readlane_t(T v, int lane) {
  if constexpr (sizeof(T) == 4) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
  } else {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffll), lane);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
  }
}

// Column col (one per wavefront) of B, rows [0, m): x <- L^-1 x with unit-lower L (m <= 64 Q).
// All 256 threads call it (it stages L through sL with barriers).
template <typename T, int Q, int CW>
__device__ __forceinline__ void trsm_wave_cols(const T* __restrict__ L, long ldl, int m, T* __restrict__ B,
                                               long ldb, long N, long col, T (*sL)[64 * Q]) {
  const int lane = threadIdx.x & 63;
  const bool has = col < N;
  T* bcol = B + (has ? col : 0) * ldb;
  T x[Q];
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    const int r = lane + 64 * q;
    x[q] = (has && r < m) ? bcol[r] : T(0);
  }
#pragma unroll
  for (int qb = 0; qb < Q; ++qb) {
    if (64 * qb < m) {  // uniform; a condition rather than a break keeps the loop unrolled
      const int rows = 64 * (Q - qb);  // staged rows [64 qb, 64 Q)
#pragma unroll
      for (int h = 0; h < 64 / CW; ++h) {
        const int i0 = 64 * qb + h * CW;
        __syncthreads();  // the previous chunk has been used
#pragma unroll 2
        for (int u = threadIdx.x; u < CW * rows; u += 256) {
          const int c = u / rows, rr = u % rows, r = 64 * qb + rr, i = i0 + c;
          sL[c][rr] = (r > i && r < m) ? L[(long)i * ldl + r] : T(0);
        }
        __syncthreads();
        // steps past m are harmless: x_i = 0 there and the staged column is zero
#pragma unroll
        for (int c = 0; c < CW; ++c) {
          const T xi = readlane_t(x[qb], h * CW + c);
#pragma unroll
          for (int q = qb; q < Q; ++q) x[q] = fma(-sL[c][64 * (q - qb) + lane], xi, x[q]);
        }
      }
    }
  }
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    const int r = lane + 64 * q;
    if (has && r < m) bcol[r] = x[q];
  }
}
template <typename T>
constexpr int trsm_cw() { return sizeof(T) == 4 ? 32 : 16; }  // staged columns per chunk (32 KB at Q = 4)
template <typename T, int Q>
__global__ __launch_bounds__(256) void k_trsm_wave(const T* __restrict__ L, long ldl, int m,
                                                   T* __restrict__ B, long ldb, long N) {
  constexpr int CW = trsm_cw<T>();
  __shared__ T sL[CW][64 * Q];  // sL[c][r - 64 qb] = L(r, i0 + c) for r > i0 + c
  trsm_wave_cols<T, Q, CW>(L, ldl, m, B, ldb, N, (long)blockIdx.x * 4 + (threadIdx.x >> 6), sL);
}
// Inside a panel: the m interchanges ipiv[p0, p0 + m) applied to the N columns at B - p0 rows (of
// A), then their top m rows solved with the unit-lower L: the interchange and TRSM launches of a
// recursion node in one. Workgroup = 4 columns, moved together, then one wavefront per column
// solves it (reading back only what this workgroup wrote).
template <typename T, int Q>
__global__ __launch_bounds__(256) void k_laswp_trsm(T* __restrict__ A, long lda, long c0, long N,
                                                    const int* __restrict__ ipiv, long p0, int m,
                                                    const T* __restrict__ L) {
  constexpr int CW = trsm_cw<T>();
  __shared__ LaswpLds S;
  __shared__ T sL[CW][64 * Q];
  const int nm = laswp_moves(ipiv, p0, m, S);
  const long g = c0 + (long)blockIdx.x * kLaswpCG;
  laswp_apply_group(A, lda, g, c0 + N, nm, S);
  trsm_wave_cols<T, Q, CW>(L, lda, m, A + c0 * lda + p0, lda, N, (long)blockIdx.x * 4 + (threadIdx.x >> 6), sL);
}
static_assert(kLaswpCG == 4, "the fused interchange + TRSM kernel moves one column per wavefront");

template <typename T>
static void trsm_lower_unit(const T* L, long ldl, long m, T* B, long ldb, long N, hipStream_t st) {
  if (m <= 0 || N <= 0) return;
  if (N <= 4096 && m <= 256) {  // few right-hand sides (inside a panel): one wavefront per column
    const dim3 g((unsigned)((N + 3) / 4));
    if (m <= 64)
      hipLaunchKernelGGL((k_trsm_wave<T, 1>), g, dim3(256), 0, st, L, ldl, (int)m, B, ldb, N);
    else if (m <= 128)
      hipLaunchKernelGGL((k_trsm_wave<T, 2>), g, dim3(256), 0, st, L, ldl, (int)m, B, ldb, N);
    else
      hipLaunchKernelGGL((k_trsm_wave<T, 4>), g, dim3(256), 0, st, L, ldl, (int)m, B, ldb, N);
    return;
  }
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
  LuTick tick;
  void* tctx;
  int err;
  const F16x3Ws* pw;  // split operands of the panel's own updates (FP16 mode), else null
};

template <typename T>
static void panel_leaf(LuCtx<T>& c, long j0, int w) {
  const long rows = c.n - j0;
  const int P = (int)((rows + kTsRows - 1) / kTsRows);
  const TsWs ws = ts_ws(c.W->d_ws, c.n);
  hipLaunchKernelGGL(k_ts_tree<T>, dim3(P), dim3(kTsRows), 0, c.st, c.A, c.lda, c.n, j0, w, P, ws,
                     c.W->d_ipiv, c.W->d_info);
  hipLaunchKernelGGL(k_ts_apply<T>, dim3(P), dim3(kTsRows), 0, c.st, c.A, c.lda, c.n, j0, w, ws);
}

constexpr long kPanelSplitK = 256;      // panel updates at least this deep use the split products
constexpr long kPanelSplitRows = 2048;  // ... when at least this tall
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
  if (w1 <= 256) {  // the node's interchanges and U12 in one launch
    const dim3 g((unsigned)((w - w1 + 3) / 4));
    const T* L11 = c.A + j0 * lda + j0;
    if (w1 <= 64)
      hipLaunchKernelGGL((k_laswp_trsm<T, 1>), g, dim3(256), 0, c.st, c.A, lda, j0 + w1, w - w1, c.W->d_ipiv, j0, (int)w1, L11);
    else if (w1 <= 128)
      hipLaunchKernelGGL((k_laswp_trsm<T, 2>), g, dim3(256), 0, c.st, c.A, lda, j0 + w1, w - w1, c.W->d_ipiv, j0, (int)w1, L11);
    else
      hipLaunchKernelGGL((k_laswp_trsm<T, 4>), g, dim3(256), 0, c.st, c.A, lda, j0 + w1, w - w1, c.W->d_ipiv, j0, (int)w1, L11);
  } else {
    laswp<T>(c.A, lda, j0 + w1, w - w1, c.W->d_ipiv, j0, j0 + w1, c.st);
    trsm_lower_unit<T>(c.A + j0 * lda + j0, lda, w1, c.A + (j0 + w1) * lda + j0, lda, w - w1, c.st);
  }
  const long mr = c.n - j0 - w1;
  if (sizeof(T) == 4 && c.pw && w1 >= kPanelSplitK && mr >= kPanelSplitRows) {
    // the deep updates of the panel (three quarters of its flops) on the split FP16 products: the
    // panel gets few CUs while the bulk update runs, and these do 5 x the work per CU of FP32 MFMA
    const int kp = (int)((w1 + 63) / 64 * 64);
    f16x3_split_rows(mr, (int)w1, (const float*)(c.A + j0 * lda + j0 + w1), lda, *c.pw, c.st, kp);
    f16x3_split_cols(w - w1, (int)w1, (const float*)(c.A + (j0 + w1) * lda + j0), lda, *c.pw, 0, c.st, kp);
    f16x3_gemm_short(mr, w - w1, (int)w1, *c.pw, (float*)(c.A + (j0 + w1) * lda + j0 + w1), lda, c.st);
  } else if ((w1 == 32 || w1 == 64) && w - w1 <= kSkinnyMax) {
    gemm_skinny<T>(mr, w - w1, w1, c.A + j0 * lda + j0 + w1, lda, c.A + (j0 + w1) * lda + j0, lda,
                   c.A + (j0 + w1) * lda + j0 + w1, lda, c.st);
  } else {
    gemm_sub<T>(mr, w - w1, w1, c.A + j0 * lda + j0 + w1, lda, c.A + (j0 + w1) * lda + j0, lda,
                c.A + (j0 + w1) * lda + j0 + w1, lda, c.st);
  }
  panel_rec(c, j0 + w1, w - w1);
  laswp<T>(c.A, lda, j0, w1, c.W->d_ipiv, j0 + w1, j0 + w, c.st);
}

// ================================================================ LU without interchanges =====
// For a matrix whose equilibrated diagonal dominates (the block-Jacobi preconditioner's boundary
// block in its natural column order: every row's own unknown on the diagonal), the LU needs no
// pivoting: a panel's leaf is one launch, with no candidate tree and no interchanges.

// lane k's value of v to every lane (readlane: a scalar broadcast, no LDS round trip)
__device__ __forceinline__ double lane_bcast(double v, int k) {
  const unsigned long long b = __double_as_longlong(v);
  const unsigned lo = __builtin_amdgcn_readlane((unsigned)b, k), hi = __builtin_amdgcn_readlane((unsigned)(b >> 32), k);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
__device__ __forceinline__ float lane_bcast(float v, int k) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), k));
}

// Leaf of 32 columns at j0. Every workgroup: wavefront 0 loads the 32 x 32 diagonal block (row r in lane r's registers) and factors it, the pivot row
// broadcast by readlane; the factored block L11\U11 and U11's reciprocal diagonal go to LDS; then
// each thread below the block solves its row of L21 against U11, unrolled. The original block must
// stay in A until every workgroup has read it: each counts its arrival (cnt[0]) once its wavefront 0
// holds the block, and the last to arrive writes the factored block back from LDS, the identity
// interchanges, info = column + 1 of the first pivot whose magnitude is below piv_tol (or not a
// number), and resets the count for the next leaf.
template <typename T>
__global__ __launch_bounds__(256) void k_leaf_nopiv(T* __restrict__ A, long lda, long n, long j0,
                                                   int* __restrict__ ipiv, int* __restrict__ info,
                                                   int* __restrict__ cnt, double piv_tol) {
  constexpr int W = kPanelW, w = kPanelW;
  __shared__ T F[W][W + 1];
  __shared__ T rdiag[W];
  __shared__ int last;
  const int t = threadIdx.x;
  if (t < 64) {
    T a[W];
    const int r = t < W ? t : W - 1;
#pragma unroll
    for (int c = 0; c < W; ++c) a[c] = A[(j0 + c) * lda + j0 + r];
    if (t == 0) {
      __atomic_thread_fence(__ATOMIC_SEQ_CST);  // the block's loads are done (their values are used)
      last = (__hip_atomic_fetch_add(cnt, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) ==
              (int)gridDim.x - 1);
    }
#pragma unroll
    for (int k = 0; k < W; ++k) {
      {
        const T piv = lane_bcast(a[k], k);
        const bool below = t > k && t < w;
        const T l = a[k] / piv;
        if (below) a[k] = l;
#pragma unroll
        for (int c = k + 1; c < W; ++c) {
          const T ukc = lane_bcast(a[c], k);
          if (below) a[c] -= l * ukc;
        }
      }
    }
    if (t < W) {
#pragma unroll
      for (int c = 0; c < W; ++c) F[t][c] = a[c];
      rdiag[t] = T(1) / a[t];
    }
  }
  __syncthreads();
  if (last && t < w) {
    for (int c = 0; c < w; ++c) A[(j0 + c) * lda + j0 + t] = F[t][c];
    ipiv[j0 + t] = (int)(j0 + t);
    if (!(fabs((double)F[t][t]) >= piv_tol)) atomicCAS(info, 0, (int)(j0 + t + 1));
    if (t == 0) *cnt = 0;
  }
  const long i = j0 + w + (long)blockIdx.x * 256 + t;
  if (i >= n) return;
  T x[W];
#pragma unroll
  for (int c = 0; c < W; ++c) x[c] = A[(j0 + c) * lda + i];
#pragma unroll
  for (int c = 0; c < W; ++c) {
    T v = x[c];
#pragma unroll
    for (int k = 0; k < c; ++k) v -= x[k] * F[k][c];
    x[c] = v * rdiag[c];
  }
#pragma unroll
  for (int c = 0; c < W; ++c)
    A[(j0 + c) * lda + i] = x[c];
}

// The last, partial leaf (w < 32; once per factorisation): the diagonal block factored in LDS, row r
// by thread r with a barrier per step, the same arrival count and write-back as k_leaf_nopiv.
template <typename T>
__global__ __launch_bounds__(256) void k_leaf_nopiv_part(T* __restrict__ A, long lda, long n, long j0, int w,
                                                        int* __restrict__ ipiv, int* __restrict__ info,
                                                        int* __restrict__ cnt, double piv_tol) {
  constexpr int W = kPanelW;
  __shared__ T F[W][W + 1];
  __shared__ int last;
  const int t = threadIdx.x;
  for (int u = t; u < w * w; u += 256) F[u % w][u / w] = A[(j0 + u / w) * lda + j0 + u % w];
  __syncthreads();
  if (t == 0) {
    __atomic_thread_fence(__ATOMIC_SEQ_CST);  // every thread's loads of the block are done
    last = (__hip_atomic_fetch_add(cnt, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == (int)gridDim.x - 1);
  }
  for (int k = 0; k < w; ++k) {
    if (t > k && t < w) {
      const T l = F[t][k] / F[k][k];
      F[t][k] = l;
      for (int c = k + 1; c < w; ++c) F[t][c] -= l * F[k][c];
    }
    __syncthreads();
  }
  if (last && t < w) {
    for (int c = 0; c < w; ++c) A[(j0 + c) * lda + j0 + t] = F[t][c];
    ipiv[j0 + t] = (int)(j0 + t);
    if (!(fabs((double)F[t][t]) >= piv_tol)) atomicCAS(info, 0, (int)(j0 + t + 1));
    if (t == 0) *cnt = 0;
  }
  const long i = j0 + w + (long)blockIdx.x * 256 + t;
  if (i >= n) return;
  for (int c = 0; c < w; ++c) {  // x U11 = a, column by column in place
    T v = A[(j0 + c) * lda + i];
    for (int k = 0; k < c; ++k) v -= A[(j0 + k) * lda + i] * F[k][c];
    A[(j0 + c) * lda + i] = v / F[c][c];
  }
}

template <typename T>
static void panel_rec_nopiv(LuCtx<T>& c, long j0, long w, double piv_tol) {
  const long lda = c.lda;
  if (w <= kPanelW) {
    const long below = c.n - j0 - w;
    const unsigned P = (unsigned)std::max<long>(1, (below + 255) / 256);
    if (w == kPanelW)
      hipLaunchKernelGGL(k_leaf_nopiv<T>, dim3(P), dim3(256), 0, c.st, c.A, lda, c.n, j0, c.W->d_ipiv,
                         c.W->d_info, c.W->d_info + 1, piv_tol);
    else
      hipLaunchKernelGGL(k_leaf_nopiv_part<T>, dim3(P), dim3(256), 0, c.st, c.A, lda, c.n, j0, (int)w,
                         c.W->d_ipiv, c.W->d_info, c.W->d_info + 1, piv_tol);
    return;
  }
  const long w1 = (w / 2 + kPanelW - 1) / kPanelW * kPanelW;
  panel_rec_nopiv(c, j0, w1, piv_tol);
  trsm_lower_unit<T>(c.A + j0 * lda + j0, lda, w1, c.A + (j0 + w1) * lda + j0, lda, w - w1, c.st);
  const long mr = c.n - j0 - w1;
  if ((w1 == 32 || w1 == 64) && w - w1 <= kSkinnyMax)
    gemm_skinny<T>(mr, w - w1, w1, c.A + j0 * lda + j0 + w1, lda, c.A + (j0 + w1) * lda + j0, lda,
                   c.A + (j0 + w1) * lda + j0 + w1, lda, c.st);
  else
    gemm_sub<T>(mr, w - w1, w1, c.A + j0 * lda + j0 + w1, lda, c.A + (j0 + w1) * lda + j0, lda,
                c.A + (j0 + w1) * lda + j0 + w1, lda, c.st);
  panel_rec_nopiv(c, j0 + w1, w - w1, piv_tol);
}

constexpr long kLuNB = IBEM_LU_NB;
// CUs the bulk trailing update leaves idle for the concurrent panel, by trailing size m. The panel
// (latency-bound, cost ~ a + b m) and the bulk update (~ m^2) overlap; under a full-GPU update the
// panel's own kernels (tree, internal FP32 updates) run 1.5-3x slower than alone. Early steps are
// update-bound and keep most CUs on the update; later steps are panel-bound and give it more.
// Breakpoints from a fit of measured panel / update times on a 64,614-unknown system.
#ifndef IBEM_PANEL_M1
#define IBEM_PANEL_M1 41000
#define IBEM_PANEL_M2 29000
#define IBEM_PANEL_M3 20000
#endif
static int panel_cus_base(long m) {
  if (m > IBEM_PANEL_M1) return 32;
  if (m > IBEM_PANEL_M2) return 64;
  if (m > IBEM_PANEL_M3) return 96;
  return 128;
}
// the schedule scaled by the tunable, in whole XCDs (8 CUs), below the device's CU count
static int panel_cus(long m) {
  const long c = (panel_cus_base(m) * g_panel_cu_percent / 100 + 4) / 8 * 8;
  return (int)std::min<long>(c, device_cus() - 8);
}
constexpr long kNarrowCols = 4096;  // columns up to which U12 uses the wavefront TRSM, not the block inverses

// Trailing update of step k restricted to columns [c0, c0 + nc) of the trailing matrix (relative to
// column k + kb): C -= L21 U12(:, cols).
template <typename T>
static void trailing_update(T* A, long n, long lda, long k, long kb, long c0, long nc,
                            const F16x3Ws* w, hipStream_t st, int cus = 0, int prod = kProdF16x3) {
  const long m = n - k - kb;
  if (m <= 0 || nc <= 0) return;
  T* C = A + (k + kb + c0) * lda + k + kb;
  if (w)
    f16x3_gemm(m, nc, (int)kb, *w, (float*)C, lda, st, c0, cus, prod);
  else
    gemm_sub<T, 1>(m, nc, kb, A + k * lda + k + kb, lda, A + (k + kb + c0) * lda + k, lda, C, lda, st);
}

constexpr int kTB = 128;  // diagonal-block size of the triangular inverses

// X = inverse of the diagonal block b of unit-lower L (Upper = false) or of U (Upper = true),
// stored column-major (X[j * kTB + t] = X(t, j)) at inv + b * kTB * kTB. Thread j = column j.
// The 128 x 128 triangle is staged in LDS first when it fits beside X (FP32: 2 x 66 KB); in FP64
// each step reads its row of the triangle from memory.
template <typename T, bool Upper>
__global__ __launch_bounds__(kTB) void k_trinv(const T* __restrict__ LU, long lda, long n,
                                               T* __restrict__ inv) {
  constexpr bool kStage = sizeof(T) == 4;
  __shared__ T X[kTB][kTB + 1];
  __shared__ T Ls[kStage ? kTB : 1][kStage ? kTB + 1 : 1];  // Ls[t][k] = triangle (t, k)
  __shared__ T srow[kTB];
  __shared__ T sdiag;
  const long k0 = (long)blockIdx.x * kTB;
  const int kb = (int)min((long)kTB, n - k0);
  const int j = threadIdx.x;
  if constexpr (kStage) {
    // column j of the block is contiguous: thread j walks it (rows t), all threads in step
    for (int c = 0; c < kb; ++c)
      if (j < kb) Ls[j][c] = LU[(k0 + c) * lda + k0 + j];
    __syncthreads();
  }
  for (int tt = 0; tt < kb; ++tt) {
    const int t = Upper ? kb - 1 - tt : tt;
    const T* row;
    if constexpr (kStage) {
      row = Ls[t];
      if (Upper && j == 0) sdiag = Ls[t][t];
      __syncthreads();
    } else {
      // row t of the triangle: L(t, 0..t-1) or U(t, t..kb-1)
      if (!Upper ? (j < t) : (j >= t && j < kb)) srow[j] = LU[(k0 + j) * lda + k0 + t];
      __syncthreads();
      if (Upper && j == 0) sdiag = srow[t];
      __syncthreads();
      row = srow;
    }
    // The sum runs over the whole range on one side of t, the same for every thread (X is zero
    // outside its triangle: rows computed so far hold zeros there), with eight independent partial
    // sums so that the LDS reads pipeline instead of forming one dependent chain.
    const int ka = Upper ? t + 1 : 0, kz = Upper ? kb : t;
    T p[8] = {T(0), T(0), T(0), T(0), T(0), T(0), T(0), T(0)};
    int k = ka;
    for (; k + 8 <= kz; k += 8) {
#pragma unroll
      for (int q = 0; q < 8; ++q) p[q] += row[k + q] * X[k + q][j];
    }
    for (; k < kz; ++k) p[0] += row[k] * X[k][j];
    const T acc = ((p[0] + p[1]) + (p[2] + p[3])) + ((p[4] + p[5]) + (p[6] + p[7]));
    if (j < kb) {
      T v = T(0);
      if (!Upper) {
        if (j < t)
          v = -acc;
        else if (j == t)
          v = T(1);
      } else {
        if (j >= t) v = sdiag != T(0) ? ((j == t ? T(1) : T(0)) - acc) / sdiag : T(0);
      }
      X[t][j] = v;
    }
    __syncthreads();
  }
  T* out = inv + (size_t)blockIdx.x * kTB * kTB;
  for (int u = j; u < kTB * kTB; u += kTB) {
    const int c = u / kTB, t = u % kTB;
    out[u] = (c < kb && t < kb) ? X[t][c] : T(0);
  }
}

// U12 = L11^-1 A12 for the block row of one step (L11 unit lower, kb x kb; A12 kb x N), through
// the inverses D_i of L11's 128 x 128 diagonal blocks (k_trinv):
//   U_i = D_i (A_i - sum_{k<i} L_ik U_k),
// as full-height GEMMs (128-row tiles) instead of a recursion down to 32-column solves and
// skinny updates. U is assembled in W.d_trsm (kb x N) and copied back over A12.
template <typename T>
static void trsm_by_inverse(const T* L, long ldl, long kb, T* B, long ldb, long N, LuWork& W,
                            hipStream_t st) {
  if (kb <= 0 || N <= 0) return;
  const long nbk = (kb + kTB - 1) / kTB;
  T* D = (T*)W.d_trinv;  // nbk x (128 x 128)
  T* U = (T*)W.d_trsm;   // kb x N, leading dimension kb
  hipLaunchKernelGGL((k_trinv<T, false>), dim3((unsigned)nbk), dim3(kTB), 0, st, L, ldl, kb, D);
  for (long i = 0; i < nbk; ++i) {
    const long r0 = i * kTB, rb = std::min<long>(kTB, kb - r0);
    if (i > 0)
      gemm_sub<T, 0>(rb, N, r0, L + r0, ldl, U, kb, B + r0, ldb, st);  // A_i -= L_i,<i U_<i
    gemm_sub<T, 2>(rb, N, rb, D + (size_t)i * kTB * kTB, kTB, B + r0, ldb, U + r0, kb, st);
  }
  hipMemcpy2DAsync(B, sizeof(T) * ldb, U, sizeof(T) * kb, sizeof(T) * kb, N, hipMemcpyDeviceToDevice, st);
}

// Blocked right-looking LU with look-ahead: after the row interchanges and U12 of step k, the
// trailing update first covers the next panel's columns; that panel is then factored on a second
// stream while the rest of the trailing update runs on the main stream. The two never touch the
// same columns: the panel works inside columns [k + kb, k + 2 kb), the rest of the update writes
// columns >= k + 2 kb and reads L21 (columns [k, k + kb)) and U12 (rows [k, k + kb)).
int lu_streams(LuWork& W) {
  if (!W.st2) {
    if (hipStreamCreateWithFlags(&W.st2, hipStreamNonBlocking) != hipSuccess) return -1;
    if (hipEventCreateWithFlags(&W.ev_a, hipEventDisableTiming) != hipSuccess) return -1;
    if (hipEventCreateWithFlags(&W.ev_b, hipEventDisableTiming) != hipSuccess) return -1;
    if (hipEventCreateWithFlags(&W.ev_r, hipEventDisableTiming) != hipSuccess) return -1;
    if (hipEventCreateWithFlags(&W.ev_cols, hipEventDisableTiming) != hipSuccess) return -1;
  }
  return 0;
}

template <typename T>
static int lu_factor(T* A, long n, long lda, LuWork& W, hipStream_t st, LuTick tick, void* tctx,
                     double* f16ws = nullptr, bool cols_pending = false, int prod = kProdF16x3) {
  if (lu_streams(W)) return -1;
  // f16ws holds two split workspaces: the bulk trailing updates' and the panel's own
  const F16x3Ws wspanel = f16ws ? f16x3_ws(f16ws + f16x3_workspace_doubles(n), n) : F16x3Ws{};
  const F16x3Ws* pw = f16ws ? &wspanel : nullptr;
  LuCtx<T> c{A, n, lda, &W, st, tick, tctx, 0, pw};
  LuCtx<T> c2{A, n, lda, &W, W.st2, tick, tctx, 0, pw};
  fill_zero(W.d_info, sizeof(int), st);
  const F16x3Ws wsp = f16ws ? f16x3_ws(f16ws, n) : F16x3Ws{};
  const F16x3Ws* w = f16ws ? &wsp : nullptr;
  if (tick) tick(tctx, "lu.panel", 0, st);
  panel_rec(c, 0, std::min(kLuNB, n));
  if (tick) tick(tctx, "lu.panel", 1, st);
  // the columns the first panel did not need (copied on the second stream)
  if (cols_pending && hipStreamWaitEvent(st, W.ev_cols, 0) != hipSuccess) return -1;
  for (long k = 0; k < n; k += kLuNB) {
    const long kb = std::min(kLuNB, n - k);
    if (k > 0) hipStreamWaitEvent(st, W.ev_b, 0);  // panel k, factored on the second stream
    if (k + kb >= n) break;
    const long m = n - k - kb;            // trailing rows and columns
    const long kb2 = std::min(kLuNB, m);  // width of the next panel
    const T* L11 = A + k * lda + k;
    // (1) the critical path: the next panel's columns get this step's interchanges, their U12 rows
    // and their trailing update, in the latency-bound forms. The interchanges go to the trailing
    // columns only; L's earlier block columns keep their order (the solve applies the
    // interchanges to the vector instead).
    if (tick) tick(tctx, "lu.swap_trsm", 0, st);
    laswp<T>(A, lda, k + kb, kb2, W.d_ipiv, k, k + kb, st);
    trsm_lower_unit<T>(L11, lda, kb, A + (k + kb) * lda + k, lda, kb2, st);
    if (tick) tick(tctx, "lu.swap_trsm", 1, st);
    if (w) {
      if (tick) tick(tctx, "lu.split", 0, st);
      f16x3_split_rows(m, (int)kb, (const float*)(A + k * lda + k + kb), lda, *w, st, kSplitPadK, prod);
      f16x3_split_cols(kb2, (int)kb, (const float*)(A + (k + kb) * lda + k), lda, *w, 0, st, kSplitPadK, prod);
      if (tick) tick(tctx, "lu.split", 1, st);
    }
    if (tick) tick(tctx, "lu.gemm", 0, st);
    trailing_update<T>(A, n, lda, k, kb, 0, kb2, w, st, 0, prod);
    if (tick) tick(tctx, "lu.gemm", 1, st);
    hipEventRecord(W.ev_a, st);
    // (2) the remaining columns (U12 through the block inverses when wide) while the panel runs; enqueued
    // before the panel's many small launches, so that the host's time spent launching them never
    // leaves the main stream waiting. The bulk update leaves panel_cus(m) CUs idle for the panel.
    const long nr = m - kb2;
    if (nr > 0) {
      T* B2 = A + (k + kb + kb2) * lda + k;
      if (tick) tick(tctx, "lu.swap_trsm_rest", 0, st);
      laswp<T>(A, lda, k + kb + kb2, nr, W.d_ipiv, k, k + kb, st);
      if (sizeof(T) == 4 && nr > kNarrowCols)
        trsm_by_inverse<T>(L11, lda, kb, B2, lda, nr, W, st);
      else
        trsm_lower_unit<T>(L11, lda, kb, B2, lda, nr, st);
      if (tick) tick(tctx, "lu.swap_trsm_rest", 1, st);
      if (w) {
        if (tick) tick(tctx, "lu.split", 0, st);
        f16x3_split_cols(nr, (int)kb, (const float*)B2, lda, *w, kb2, st, kSplitPadK, prod);
        if (tick) tick(tctx, "lu.split", 1, st);
      }
      if (tick) tick(tctx, "lu.gemm", 0, st);
      trailing_update<T>(A, n, lda, k, kb, kb2, nr, w, st, device_cus() - panel_cus(m), prod);
      if (tick) tick(tctx, "lu.gemm", 1, st);
    }
    hipStreamWaitEvent(W.st2, W.ev_a, 0);
    if (tick) tick(tctx, "lu.panel", 0, W.st2);
    panel_rec(c2, k + kb, kb2);
    if (tick) tick(tctx, "lu.panel", 1, W.st2);
    hipEventRecord(W.ev_b, W.st2);
  }
  return 0;
}

int lu_factor_f64(double* A, long n, long lda, LuWork& W, hipStream_t st, LuTick tick, void* tctx) {
  return lu_factor<double>(A, n, lda, W, st, tick, tctx);
}

// The blocked right-looking LU of lu_factor with its look-ahead, without interchanges (panels by
// panel_rec_nopiv); W.d_info reports a pivot below piv_tol.
#ifndef IBEM_NOPIV_NB
#define IBEM_NOPIV_NB 256
#endif
constexpr long kNopivNB = IBEM_NOPIV_NB;  // panel width: the panel's own updates scale with it
template <typename T>
static int lu_factor_nopiv(T* A, long n, long lda, LuWork& W, hipStream_t st, double piv_tol) {
  if (lu_streams(W)) return -1;
  LuCtx<T> c{A, n, lda, &W, st, nullptr, nullptr, 0, nullptr};
  LuCtx<T> c2{A, n, lda, &W, W.st2, nullptr, nullptr, 0, nullptr};
  fill_zero(W.d_info, 2 * sizeof(int), st);  // info and the leaves' arrival count
  panel_rec_nopiv(c, 0, std::min(kNopivNB, n), piv_tol);
  for (long k = 0; k < n; k += kNopivNB) {
    const long kb = std::min(kNopivNB, n - k);
    if (k > 0) hipStreamWaitEvent(st, W.ev_b, 0);  // panel k, factored on the second stream
    if (k + kb >= n) break;
    const long m = n - k - kb;
    const long kb2 = std::min(kNopivNB, m);
    const T* L11 = A + k * lda + k;
    trsm_lower_unit<T>(L11, lda, kb, A + (k + kb) * lda + k, lda, kb2, st);
    trailing_update<T>(A, n, lda, k, kb, 0, kb2, nullptr, st);
    hipEventRecord(W.ev_a, st);
    const long nr = m - kb2;
    if (nr > 0) {
      trsm_lower_unit<T>(L11, lda, kb, A + (k + kb + kb2) * lda + k, lda, nr, st);
      trailing_update<T>(A, n, lda, k, kb, kb2, nr, nullptr, st);
    }
    hipStreamWaitEvent(W.st2, W.ev_a, 0);
    panel_rec_nopiv(c2, k + kb, kb2, piv_tol);
    hipEventRecord(W.ev_b, W.st2);
  }
  return 0;
}
int lu_factor_f64_nopiv(double* A, long n, long lda, LuWork& W, hipStream_t st, double piv_tol) {
  return lu_factor_nopiv<double>(A, n, lda, W, st, piv_tol);
}
int lu_factor_f32_nopiv(float* A, long n, long lda, LuWork& W, hipStream_t st, double piv_tol) {
  return lu_factor_nopiv<float>(A, n, lda, W, st, piv_tol);
}
int lu_factor_f32(float* A, long n, long lda, LuWork& W, hipStream_t st, LuTick tick, void* tctx,
                  double* f16ws, bool cols_pending, int prod) {
  return lu_factor<float>(A, n, lda, W, st, tick, tctx, f16ws, cols_pending && n > kLuNB, prod);
}

// ======================================================================= triangular solves ======
// After the factorisation: the row permutation as a gather list (built on the host from ipiv) and
// the inverses of the 128 x 128 diagonal blocks of L and U (k_trinv, one workgroup per block). A
// solve is then one launch per block column and direction: every workgroup subtracts its rows'
// share of the just-solved block (a coalesced tall GEMV), and workgroup 0, whose rows are the next
// diagonal block, finishes that block at once with the precomputed inverse. The factors are read
// once per solve (n^2 elements), so the solve runs at HBM speed rather than at one workgroup's.

size_t lu_trinv_doubles() { return (size_t)IBEM_LU_NB * kTB; }
size_t lu_trsm_doubles(long n) { return (size_t)IBEM_LU_NB * n; }

size_t lu_inverse_doubles(long n) {
  const long nb = (n + kTB - 1) / kTB;
  return (size_t)nb * kTB * kTB * 2;
}

// Solve steps of kTS = 2 kTB rows, run by workgroups of kTT = 1024 threads: every launch reduces
// all rows past the just-solved step with its 256 columns (a coalesced tall GEMV, 4 threads per row),
// and workgroup 0, whose rows are the next step, finishes that step at once: two kTB-blocks, each
// through its precomputed inverse, the second after the first's contribution (8 threads per row).
// Short per-thread chains keep each dependent launch near its memory latency.
constexpr int kTS = 2 * kTB;
constexpr int kTT = 1024;
static_assert(IBEM_LU_NB % kTS == 0, "an LU step spans whole solve steps");
static_assert(kTT == 4 * kTS && kTT == 8 * kTB, "thread layout of the solve steps");

// Factor accessors of the solves: entry (r, c) of the stored factor and the scale that multiplies
// the stored values of its kTB x kTB tile (every sum the solves form runs inside one tile, so the
// scale is applied once per partial sum).
template <typename T>
struct FacPlain {  // the factor as the LU left it (FP32 or FP64)
  const T* p;
  long ld;
  __device__ T at(long r, long c) const { return p[c * ld + r]; }
  __device__ T scale(long, long) const { return T(1); }
};
// FP16 copy of the FP32 factor: the values of the kTB x kTB tile (rows [kTB i, kTB i + kTB),
// columns [kTB j, ...)) scaled by a power of two, whose inverse is isc[j * ntr + i]
struct Fac16 {
  const unsigned short* p;
  long ld;
  const float* isc;
  long ntr;
  __device__ float at(long r, long c) const { return (float)__builtin_bit_cast(_Float16, p[c * ld + r]); }
  __device__ float scale(long r, long c) const { return isc[(c / kTB) * ntr + r / kTB]; }
};

// Row sums of a rows x cols block (rows, cols <= kTB) at (r0, c0) of M, a factor accessor, times xs
// (LDS): thread (g, i), i = t % kTB, g = t / kTB, sums columns [16 g, 16 g + 16) of row i, and the
// 8 partials are added in a fixed order. Returns row t's sum to threads t < rows. All kTT threads
// call it (it holds barriers); pp: kTT entries of LDS.
template <typename T, typename FA>
__device__ T block_mv(const FA& M, long r0, long c0, int rows, int cols, const T* xs, T* pp) {
  const int t = threadIdx.x, i = t % kTB, g = t / kTB;
  T acc0 = T(0), acc1 = T(0);
  if (i < rows) {
    const long r = r0 + i, cb = c0 + 16 * g;
    if (16 * g + 16 <= cols) {
#pragma unroll
      for (int c = 0; c < 16; c += 2) {
        acc0 += M.at(r, cb + c) * xs[16 * g + c];
        acc1 += M.at(r, cb + c + 1) * xs[16 * g + c + 1];
      }
    } else {
      for (int c = 0; 16 * g + c < cols; ++c) acc0 += M.at(r, cb + c) * xs[16 * g + c];
    }
    if (16 * g < cols) {
      const T sc = M.scale(r, cb);
      acc0 *= sc;
      acc1 *= sc;
    }
  }
  pp[t] = acc0 + acc1;
  __syncthreads();
  T r = T(0);
  if (t < rows)
    r = ((pp[t] + pp[kTB + t]) + (pp[2 * kTB + t] + pp[3 * kTB + t])) +
        ((pp[4 * kTB + t] + pp[5 * kTB + t]) + (pp[6 * kTB + t] + pp[7 * kTB + t]));
  __syncthreads();
  return r;
}

// y[r0, r0 + kn) <- the solution of the kn x kn diagonal block at r0 (kn <= kTS) for the reduced
// right-hand side sy (LDS). Lower (forward sweep, unit L): x0 = D0 sy0, x1 = D1 (sy1 - L10 x0).
// Upper (backward sweep): x1 = D1 sy1, x0 = D0 (sy0 - U01 x1). D = inverses of the kTB-blocks
// (inv + block * kTB^2, column-major); sz: kTS + kTB entries of LDS; pp: kTT entries of LDS.
template <typename T, bool Upper, typename FA>
__device__ void trsv_finish(const FA& LU, const T* __restrict__ inv, long r0,
                            int kn, const T* sy, T* sz, T* pp, T* __restrict__ y) {
  const int t = threadIdx.x;
  const int n0 = min(kTB, kn), n1 = kn - n0;
  const T* D0 = inv + (size_t)(r0 / kTB) * kTB * kTB;
  const T* D1 = D0 + (size_t)kTB * kTB;
  // first the block that does not depend on the other: 0 for Lower, 1 for Upper
  const int fa = Upper && n1 > 0 ? n0 : 0, na = Upper && n1 > 0 ? n1 : n0;
  const T xa = block_mv(FacPlain<T>{Upper && n1 > 0 ? D1 : D0, kTB}, 0, 0, na, na, sy + fa, pp);
  if (t < na) {
    sz[fa + t] = xa;
    y[r0 + fa + t] = xa;
  }
  if (n1 == 0) return;
  __syncthreads();
  // the other block, reduced by the first block's solution through the off-diagonal block of LU:
  // column (r0 + fa + c), row (r0 + fb + t)
  const int fb = Upper ? 0 : n0, nb = Upper ? n0 : n1;
  const T w = block_mv(LU, r0 + fb, r0 + fa, nb, na, sz + fa, pp);
  if (t < nb) sz[kTS + t] = sy[fb + t] - w;
  __syncthreads();
  const T xb = block_mv(FacPlain<T>{Upper ? D0 : D1, kTB}, 0, 0, nb, nb, sz + kTS, pp);
  if (t < nb) y[r0 + fb + t] = xb;
}

// Forward sweep. Step [k0, k0 + kTS) of y is solved; the rows past it are reduced by its columns
// of L, and workgroup 0 (the next step's rows) finishes the next step unless it starts an LU step
// (solve_next = 0: k_trsv_fwd_step does it after the step's interchanges).
template <typename T, typename FA>
__global__ __launch_bounds__(1024) void k_trsv_fwd(FA LU, long n, long k0,
                                                   const T* __restrict__ inv, T* __restrict__ y,
                                                   int solve_next) {
  __shared__ T sx[kTS];
  __shared__ T sy[kTS];
  __shared__ T sz[kTS + kTB];
  __shared__ T pp[kTT];
  const int t = threadIdx.x, rl = t % kTS, g = t / kTS;
  if (t < kTS) sx[t] = y[k0 + t];  // k0 + kTS < n
  __syncthreads();
  const long r0 = k0 + kTS + (long)blockIdx.x * kTS;
  const long r = r0 + rl;
  T p0 = T(0), p1 = T(0), p2 = T(0), p3 = T(0), sc = T(1);
  if (r < n) {
    const long cb = k0 + 64 * g;
    const T* xs = sx + 64 * g;
#pragma unroll
    for (int c = 0; c < 64; c += 4) {
      p0 += LU.at(r, cb + c) * xs[c];
      p1 += LU.at(r, cb + c + 1) * xs[c + 1];
      p2 += LU.at(r, cb + c + 2) * xs[c + 2];
      p3 += LU.at(r, cb + c + 3) * xs[c + 3];
    }
    sc = LU.scale(r, cb);
  }
  pp[t] = ((p0 + p1) + (p2 + p3)) * sc;
  __syncthreads();
  const bool own = blockIdx.x == 0 && solve_next;
  if (t < kTS && r < n) {
    const T v = y[r] - ((pp[t] + pp[kTS + t]) + (pp[2 * kTS + t] + pp[3 * kTS + t]));
    if (own)
      sy[t] = v;
    else
      y[r] = v;
  }
  if (own) {
    __syncthreads();
    trsv_finish<T, false>(LU, inv, r0, (int)min((long)kTS, n - r0), sy, sz, pp, y);
  }
}

// Forward sweep at the first solve step of an LU step: the step's interchanges applied to y as moves
// y[dst] <- y[src] (all reads, then all writes), then the solve step at r0 finished. One workgroup;
// an LU step has at most 2 IBEM_LU_NB moves.
template <typename T, typename FA>
__global__ __launch_bounds__(1024) void k_trsv_fwd_step(const int* __restrict__ mv,
                                                        const int* __restrict__ nmvp, long r0,
                                                        long n, FA LU,
                                                        const T* __restrict__ inv, T* __restrict__ y) {
  __shared__ T sv[2 * IBEM_LU_NB];
  __shared__ T sy[kTS];
  __shared__ T sz[kTS + kTB];
  __shared__ T pp[kTT];
  const int t = threadIdx.x;
  const int nmv = *nmvp;
  for (int q = t; q < nmv; q += kTT) sv[q] = y[mv[2 * q + 1]];
  __syncthreads();
  for (int q = t; q < nmv; q += kTT) y[mv[2 * q]] = sv[q];
  __syncthreads();
  const int kn = (int)min((long)kTS, n - r0);
  if (t < kn) sy[t] = y[r0 + t];
  __syncthreads();
  trsv_finish<T, false>(LU, inv, r0, kn, sy, sz, pp, y);
}

// Backward sweep. Step [k0, k0 + kb) of y is solved; the rows above it are reduced by its columns
// of U, and workgroup 0 (rows [k0 - kTS, k0)) finishes that step. With k0 < 0 the launch only
// finishes the last step [lastk, n).
template <typename T, typename FA>
__global__ __launch_bounds__(1024) void k_trsv_bwd(FA LU, long n, long k0,
                                                   int kb, long lastk, const T* __restrict__ inv,
                                                   T* __restrict__ y) {
  __shared__ T sx[kTS];
  __shared__ T sy[kTS];
  __shared__ T sz[kTS + kTB];
  __shared__ T pp[kTT];
  const int t = threadIdx.x;
  if (k0 < 0) {
    const int kn = (int)(n - lastk);
    if (t < kn) sy[t] = y[lastk + t];
    __syncthreads();
    trsv_finish<T, true>(LU, inv, lastk, kn, sy, sz, pp, y);
    return;
  }
  const int rl = t % kTS, g = t / kTS;
  if (t < kTS) sx[t] = t < kb ? y[k0 + t] : T(0);
  __syncthreads();
  const long r0 = k0 - (long)(blockIdx.x + 1) * kTS;  // >= 0: k0 is a multiple of kTS
  const long r = r0 + rl;
  T p0 = T(0), p1 = T(0), p2 = T(0), p3 = T(0), sc = T(1);
  {
    const long cb = k0 + 64 * g;
    const T* xs = sx + 64 * g;
    if (kb == kTS) {
#pragma unroll
      for (int c = 0; c < 64; c += 4) {
        p0 += LU.at(r, cb + c) * xs[c];
        p1 += LU.at(r, cb + c + 1) * xs[c + 1];
        p2 += LU.at(r, cb + c + 2) * xs[c + 2];
        p3 += LU.at(r, cb + c + 3) * xs[c + 3];
      }
    } else {  // the last step: only its kb < kTS columns exist
      for (int c = 0; 64 * g + c < kb && c < 64; ++c) p0 += LU.at(r, cb + c) * xs[c];
    }
    if (64 * g < kb) sc = LU.scale(r, cb);
  }
  pp[t] = ((p0 + p1) + (p2 + p3)) * sc;
  __syncthreads();
  T v = T(0);
  if (t < kTS) v = y[r] - ((pp[t] + pp[kTS + t]) + (pp[2 * kTS + t] + pp[3 * kTS + t]));
  if (blockIdx.x == 0) {
    if (t < kTS) sy[t] = v;
    __syncthreads();
    trsv_finish<T, true>(LU, inv, r0, kTS, sy, sz, pp, y);
  } else if (t < kTS) {
    y[r] = v;
  }
}

// Each LU step's interchanges composed into a move list on the device (laswp_moves, the composition
// the interchange kernels use): step j's (dst, src) pairs at perm + j * kMvCap ints, its count at
// perm[steps * kMvCap + j]. The solve applies them to the vector as all reads, then all writes, so
// their order (atomic slot claims) does not matter. No host round trip: the preparation can be
// captured in a graph.
constexpr long kMvCap = 4 * IBEM_LU_NB;  // ints per step: at most 2 IBEM_LU_NB moves of two ints
static long lu_steps(long n) { return (n + IBEM_LU_NB - 1) / IBEM_LU_NB; }
size_t lu_moves_ints(long n) { return (size_t)lu_steps(n) * (kMvCap + 1); }

__global__ __launch_bounds__(256) void k_step_moves(const int* __restrict__ ipiv, long n, long steps,
                                                    int* __restrict__ perm) {
  __shared__ LaswpLds S;
  const long j = blockIdx.x, p0 = j * IBEM_LU_NB;
  const int m = (int)min((long)IBEM_LU_NB, n - p0);
  const int nm = laswp_moves(ipiv, p0, m, S);
  int* out = perm + j * kMvCap;
  for (int q = threadIdx.x; q < nm; q += 256) {
    out[2 * q] = S.mdst[q];
    out[2 * q + 1] = S.msrc[q];
  }
  if (threadIdx.x == 0) perm[steps * kMvCap + j] = nm;
}

template <typename T>
static int lu_prepare(const T* LU, long n, long lda, LuWork& W, hipStream_t st) {
  const long nb = (n + kTB - 1) / kTB;
  T* inv = (T*)W.d_inv;
  hipLaunchKernelGGL((k_trinv<T, false>), dim3((unsigned)nb), dim3(kTB), 0, st, LU, lda, n, inv);
  hipLaunchKernelGGL((k_trinv<T, true>), dim3((unsigned)nb), dim3(kTB), 0, st, LU, lda, n,
                     inv + (size_t)nb * kTB * kTB);
  hipLaunchKernelGGL(k_step_moves, dim3((unsigned)lu_steps(n)), dim3(256), 0, st, W.d_ipiv, n, lu_steps(n),
                     W.d_perm);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

template <typename T, typename FA>
static void lu_solve(const FA& LU, long n, const LuWork& W, T* x, hipStream_t st) {
  const long nb = (n + kTB - 1) / kTB;
  const long ns = (n + kTS - 1) / kTS;
  const T* invL = (const T*)W.d_inv;
  const T* invU = invL + (size_t)nb * kTB * kTB;
  T* y = (T*)W.d_tmp;
  hipMemcpyAsync(y, x, sizeof(T) * n, hipMemcpyDeviceToDevice, st);
  for (long s = 0; s < ns; ++s) {
    const long k0 = s * kTS;
    if (k0 % IBEM_LU_NB == 0) {  // first solve step of an LU step: its interchanges, then the step
      const long j = k0 / IBEM_LU_NB;
      const int* mv = (const int*)W.d_perm;
      hipLaunchKernelGGL((k_trsv_fwd_step<T, FA>), dim3(1), dim3(kTT), 0, st, mv + j * kMvCap,
                         mv + lu_steps(n) * kMvCap + j, k0, n, LU, invL, y);
    }
    if (s + 1 < ns) {
      const long rows = n - k0 - kTS;
      const int solve_next = ((s + 1) * kTS) % IBEM_LU_NB != 0;
      hipLaunchKernelGGL((k_trsv_fwd<T, FA>), dim3((unsigned)((rows + kTS - 1) / kTS)), dim3(kTT), 0, st, LU,
                         n, k0, invL, y, solve_next);
    }
  }
  const long lastk = (ns - 1) * kTS;
  hipLaunchKernelGGL((k_trsv_bwd<T, FA>), dim3(1), dim3(kTT), 0, st, LU, n, -1L, 0, lastk, invU, y);
  for (long s = ns - 1; s >= 1; --s) {
    const long k0 = s * kTS;
    const int kb = (int)std::min<long>(kTS, n - k0);
    hipLaunchKernelGGL((k_trsv_bwd<T, FA>), dim3((unsigned)s), dim3(kTT), 0, st, LU, n, k0, kb, lastk,
                       invU, y);
  }
  hipMemcpyAsync(x, y, sizeof(T) * n, hipMemcpyDeviceToDevice, st);
}

int lu_prepare_solve_f64(const double* LU, long n, long lda, LuWork& W, hipStream_t st) {
  return lu_prepare<double>(LU, n, lda, W, st);
}
int lu_prepare_solve_f32(const float* LU, long n, long lda, LuWork& W, hipStream_t st) {
  return lu_prepare<float>(LU, n, lda, W, st);
}
void lu_solve_f64(const double* LU, long n, long lda, const LuWork& W, double* x, hipStream_t st) {
  lu_solve<double>(FacPlain<double>{LU, lda}, n, W, x, st);
}
void lu_solve_f32(const float* LU, long n, long lda, const LuWork& W, float* x, hipStream_t st) {
  lu_solve<float>(FacPlain<float>{LU, lda}, n, W, x, st);
}

// ============================================== FP16 copy of the factor (compact storage) =======
// Every kTB x kTB tile of the FP32 factor is scaled by the power of two that brings its largest
// magnitude into [2^14, 2^15) and stored in FP16; the tile's inverse scale goes to isc. One
// workgroup per tile: the maximum, then the conversion (the second read hits the cache).
long factor16_tile_rows(long n) { return (n + kTB - 1) / kTB; }
size_t factor16_scales(long n) { return (size_t)factor16_tile_rows(n) * factor16_tile_rows(n); }

__global__ __launch_bounds__(256) void k_factor16(const float* __restrict__ LU, long n, long lda,
                                                  unsigned short* __restrict__ F, long ldf,
                                                  float* __restrict__ isc) {
  __shared__ float red[4];
  const long ntr = (n + kTB - 1) / kTB;
  const long ti = blockIdx.x, tj = blockIdx.y;
  const long r0 = ti * kTB, c0 = tj * kTB;
  const int rows = (int)min((long)kTB, n - r0), cols = (int)min((long)kTB, n - c0);
  const int t = threadIdx.x, rl = t % kTB, cg = t / kTB;  // two columns per pass
  float m = 0.0f;
  if (rl < rows)
    for (int c = cg; c < cols; c += 2) m = fmaxf(m, fabsf(LU[(c0 + c) * lda + r0 + rl]));
  for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o));
  if ((t & 63) == 0) red[t >> 6] = m;
  __syncthreads();
  const float mm = fmaxf(fmaxf(red[0], red[1]), fmaxf(red[2], red[3]));
  const float sc = pow2_scale(mm);
  if (t == 0) isc[tj * ntr + ti] = 1.0f / sc;
  if (rl < rows)
    for (int c = cg; c < cols; c += 2)
      F[(c0 + c) * ldf + r0 + rl] =
          __builtin_bit_cast(unsigned short, (_Float16)(LU[(c0 + c) * lda + r0 + rl] * sc));
}

void factor_to_16(const float* LU, long n, long lda, unsigned short* F, long ldf, float* isc, hipStream_t st) {
  const long ntr = factor16_tile_rows(n);
  hipLaunchKernelGGL(k_factor16, dim3((unsigned)ntr, (unsigned)ntr), dim3(256), 0, st, LU, n, lda, F, ldf, isc);
}

void lu_solve_16(const unsigned short* F, long n, long ldf, const float* isc, const LuWork& W, float* x,
                 hipStream_t st) {
  lu_solve<float>(Fac16{F, ldf, isc, factor16_tile_rows(n)}, n, W, x, st);
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

// Row maxima in two passes: workgroup (256 rows, one of kRowChunks column chunks) -> partial
// maxima, then r_i = 1 / max. Every column is read by many workgroups at once (coalesced, HBM rate).
constexpr int kRowChunks = 32;
__global__ __launch_bounds__(256) void k_rowmax_part(const double* __restrict__ A, long n, long lda,
                                                     double* __restrict__ part) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const long per = (n + kRowChunks - 1) / kRowChunks;
  const long j0 = blockIdx.y * per, j1 = j0 + per < n ? j0 + per : n;
  double m0 = 0.0, m1 = 0.0, m2 = 0.0, m3 = 0.0;
  long j = j0;
  for (; j + 4 <= j1; j += 4) {
    m0 = fmax(m0, fabs(A[j * lda + i]));
    m1 = fmax(m1, fabs(A[(j + 1) * lda + i]));
    m2 = fmax(m2, fabs(A[(j + 2) * lda + i]));
    m3 = fmax(m3, fabs(A[(j + 3) * lda + i]));
  }
  for (; j < j1; ++j) m0 = fmax(m0, fabs(A[j * lda + i]));
  part[blockIdx.y * n + i] = fmax(fmax(m0, m1), fmax(m2, m3));
}
__global__ __launch_bounds__(256) void k_rowmax_done(const double* __restrict__ part, long n,
                                                     double* __restrict__ r) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double m = 0.0;
  for (int q = 0; q < kRowChunks; ++q) m = fmax(m, part[q * n + i]);
  r[i] = m > 0.0 ? 1.0 / m : 1.0;
}
// Column j: c_j = 1 / max_i |r_i A_ij|, out[:, j] = R A C (in TO precision); A is only read.
// One workgroup per column, kColThreads threads (one workgroup per CU): the column is read for its
// maximum, then again for the scaled copy; with few columns in flight the second read finds them in
// the device's last-level cache.
#ifndef IBEM_COL_THREADS
#define IBEM_COL_THREADS 1024
#endif
constexpr int kColThreads = IBEM_COL_THREADS;
// perm (optional): output column j is A's column perm[j].
template <typename TO>
__global__ __launch_bounds__(kColThreads) void k_colscale_copy(const double* __restrict__ A, long n, long lda,
                                                               const double* __restrict__ r,
                                                               double* __restrict__ c, TO* __restrict__ out,
                                                               long ldo, long j0, const int* __restrict__ perm) {
  __shared__ double red[kColThreads / 64];
  const long j = j0 + blockIdx.x;
  const double* col = A + (perm ? (long)perm[j] : j) * lda;
  double m = 0.0;
  for (long i = threadIdx.x; i < n; i += kColThreads) m = fmax(m, fabs(col[i] * r[i]));
  for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
  __syncthreads();
  double mm = red[0];
#pragma unroll
  for (int q = 1; q < kColThreads / 64; ++q) mm = fmax(mm, red[q]);
  const double cj = mm > 0.0 ? 1.0 / mm : 1.0;
  if (threadIdx.x == 0) c[j] = cj;
  TO* o = out + j * ldo;
  for (long i = threadIdx.x; i < n; i += kColThreads) o[i] = (TO)(col[i] * r[i] * cj);
}

size_t equilibrate_workspace_doubles(long n) { return (size_t)kRowChunks * n; }

// Row maxima, then the scaled copy. With st_rest, only columns [0, cols_first) are copied on st; the
// rest are copied on st_rest, whose completion ev_rest records. Returns nonzero when the events
// that order the two streams could not be recorded or waited on.
template <typename TO>
static int equilibrate_copy(const double* A, long n, long lda, double* r, double* c, TO* out, long ldo,
                            double* ws, hipStream_t st, hipStream_t st_rest = nullptr,
                            hipEvent_t ev_r = nullptr, hipEvent_t ev_rest = nullptr, long cols_first = 0,
                            const int* perm = nullptr) {
  hipLaunchKernelGGL(k_rowmax_part, dim3((unsigned)((n + 255) / 256), kRowChunks), dim3(256), 0, st, A,
                     n, lda, ws);
  hipLaunchKernelGGL(k_rowmax_done, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, ws, n, r);
  const long first = st_rest && cols_first > 0 && cols_first < n ? cols_first : n;
  hipLaunchKernelGGL(k_colscale_copy<TO>, dim3((unsigned)first), dim3(kColThreads), 0, st, A, n, lda, r, c,
                     out, ldo, 0L, perm);
  if (first < n) {
    if (hipEventRecord(ev_r, st) != hipSuccess) return -1;  // the row scales
    if (hipStreamWaitEvent(st_rest, ev_r, 0) != hipSuccess) return -1;
    hipLaunchKernelGGL(k_colscale_copy<TO>, dim3((unsigned)(n - first)), dim3(kColThreads), 0, st_rest, A, n,
                       lda, r, c, out, ldo, first, perm);
    if (hipEventRecord(ev_rest, st_rest) != hipSuccess) return -1;
  }
  return 0;
}
void equilibrate_copy_f64(const double* A, long n, long lda, double* r, double* c, double* out, long ldo,
                          double* ws, hipStream_t st, const int* perm) {
  (void)equilibrate_copy<double>(A, n, lda, r, c, out, ldo, ws, st, nullptr, nullptr, nullptr, 0, perm);
}
int equilibrate_copy_f32(const double* A, long n, long lda, double* r, double* c, float* out, long ldo,
                         double* ws, hipStream_t st, LuWork* W, const int* perm) {
  if (!W) return equilibrate_copy<float>(A, n, lda, r, c, out, ldo, ws, st, nullptr, nullptr, nullptr, 0, perm);
  if (perm) return -1;
  if (lu_streams(*W)) return -1;
  // the columns past the first panel are copied beside it
  return equilibrate_copy<float>(A, n, lda, r, c, out, ldo, ws, st, W->st2, W->ev_r, W->ev_cols, kLuNB);
}

// ---- in place (compact storage: the factor overwrites the FP64 matrix) ----
// Column j: c_j = 1 / max_i |r_i A_ij|, then A[:, j] <- R A C, each element read and written by the
// same thread
__global__ __launch_bounds__(kColThreads) void k_colscale_inplace(double* A, long n, long lda,
                                                                  const double* __restrict__ r,
                                                                  double* __restrict__ c) {
  __shared__ double red[kColThreads / 64];
  const long j = blockIdx.x;
  double* col = A + j * lda;
  double m = 0.0;
  for (long i = threadIdx.x; i < n; i += kColThreads) m = fmax(m, fabs(col[i] * r[i]));
  for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
  __syncthreads();
  double mm = red[0];
#pragma unroll
  for (int q = 1; q < kColThreads / 64; ++q) mm = fmax(mm, red[q]);
  const double cj = mm > 0.0 ? 1.0 / mm : 1.0;
  if (threadIdx.x == 0) c[j] = cj;
  for (long i = threadIdx.x; i < n; i += kColThreads) col[i] = col[i] * r[i] * cj;
}

void equilibrate_in_place_f64(double* A, long n, long lda, double* r, double* c, double* ws, hipStream_t st) {
  hipLaunchKernelGGL(k_rowmax_part, dim3((unsigned)((n + 255) / 256), kRowChunks), dim3(256), 0, st, A,
                     n, lda, ws);
  hipLaunchKernelGGL(k_rowmax_done, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, ws, n, r);
  hipLaunchKernelGGL(k_colscale_inplace, dim3((unsigned)n), dim3(kColThreads), 0, st, A, n, lda, r, c);
}

// FP64 A (n x n, leading dimension lda >= n) -> FP32 R A C with leading dimension lda32 (>= n) in the
// same memory. FP32 column j occupies bytes [4 lda32 j, 4 lda32 (j + 1)), which lie inside FP64
// columns < j: columns are converted in increasing blocks [j0, j1) with 4 lda32 j1 <= 8 lda j0, so
// that a launch writes only over FP64 columns that earlier launches have read, and never over the
// columns it reads itself. The first columns (whose FP32 place overlaps themselves) go through tmp
// first. Returns nonzero when tmp holds fewer than two columns.
int equilibrate_to_f32_in_place(double* A, long n, long lda, double* r, double* c, long lda32, float* tmp,
                                size_t tmp_floats, double* ws, hipStream_t st) {
  if (lda < n || lda32 < n || lda32 >= 2 * lda) return -1;
  const long S = std::min<long>(n, (long)(tmp_floats / (size_t)lda32));
  if (S < 2 && S < n) return -1;
  hipLaunchKernelGGL(k_rowmax_part, dim3((unsigned)((n + 255) / 256), kRowChunks), dim3(256), 0, st, A,
                     n, lda, ws);
  hipLaunchKernelGGL(k_rowmax_done, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, ws, n, r);
  hipLaunchKernelGGL(k_colscale_copy<float>, dim3((unsigned)S), dim3(kColThreads), 0, st, A, n, lda, r, c,
                     tmp, lda32, 0L, nullptr);
  float* out = (float*)A;
  for (long j0 = S; j0 < n;) {
    const long j1 = std::min<long>(n, (2 * lda * j0) / lda32);  // 4 lda32 j1 <= 8 lda j0
    if (j1 <= j0) return -1;
    hipLaunchKernelGGL(k_colscale_copy<float>, dim3((unsigned)(j1 - j0)), dim3(kColThreads), 0, st, A, n,
                       lda, r, c, out, lda32, j0, nullptr);
    j0 = j1;
  }
  if (hipMemcpyAsync(out, tmp, sizeof(float) * (size_t)S * lda32, hipMemcpyDeviceToDevice, st) != hipSuccess)
    return -1;
  return 0;
}

void equilibrate_f64(double* A, long n, long lda, double* r, double* c, hipStream_t st) {
  hipLaunchKernelGGL(k_rowmax, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, A, n, lda, r);
  hipLaunchKernelGGL(k_colscale, dim3((unsigned)n), dim3(256), 0, st, A, n, lda, r, c);
}

// ==================================================================================== misc =======
// Zero fill by a kernel (stream-ordered, and a kernel node when captured in a graph, where memset
// nodes are avoided): 16-byte stores over the aligned part, then the tail in 4-byte words.
__global__ __launch_bounds__(256) void k_zero(unsigned* __restrict__ p, size_t words) {
  const size_t q = words / 4;
  uint4* p4 = (uint4*)p;
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < q; i += (size_t)gridDim.x * 256)
    p4[i] = make_uint4(0u, 0u, 0u, 0u);
  const size_t t = 4 * q + (size_t)blockIdx.x * 256 + threadIdx.x;
  if (t < words) p[t] = 0u;
}
void fill_zero(void* p, size_t bytes, hipStream_t st) {
  if (!bytes) return;
  const size_t words = bytes / 4;  // every buffer here holds 4- or 8-byte elements
  const size_t blocks = std::min<size_t>((words / 4 + 255) / 256 + 1, 65536);
  hipLaunchKernelGGL(k_zero, dim3((unsigned)blocks), dim3(256), 0, st, (unsigned*)p, words);
}
// y = alpha A x + beta y. The columns are split into kGemvChunks ranges so that enough loads are in
// flight to stream A at HBM speed (one row per thread over all n columns keeps too few); the range
// partial sums are added in a fixed order by a second kernel.
constexpr int kGemvChunks = 64;
size_t gemv_workspace_doubles(long m) { return (size_t)kGemvChunks * m; }

// thread = rows 2 u and 2 u + 1 (one 16-byte load per column when lda is even and A 16-byte aligned,
// else one row per thread); each row keeps eight partial sums over its columns j = q (mod 8) of the
// chunk, added in a fixed order. A is streamed with nontemporal loads (read once per product; keeping
// it out of the caches measured 3-5 % faster).
typedef double dvec2 __attribute__((ext_vector_type(2)));
__global__ __launch_bounds__(256) void k_gemv_part(long m, long n, const double* __restrict__ A,
                                                   long lda, const double* __restrict__ x,
                                                   double* __restrict__ part) {
  const long cw = (n + gridDim.y - 1) / gridDim.y;  // columns per chunk
  const long j0 = (long)blockIdx.y * cw, j1 = min(n, j0 + cw);
  const bool pairs = (lda & 1) == 0 && (((size_t)A) & 15) == 0;
  if (pairs) {
    const long r = 2 * ((long)blockIdx.x * blockDim.x + threadIdx.x);
    if (r >= m) return;
    if (r + 1 < m) {
      double a0[8] = {0, 0, 0, 0, 0, 0, 0, 0}, a1[8] = {0, 0, 0, 0, 0, 0, 0, 0};
      long j = j0;
      for (; j + 8 <= j1; j += 8) {
        dvec2 v[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) v[q] = __builtin_nontemporal_load((const dvec2*)(A + (j + q) * lda + r));
#pragma unroll
        for (int q = 0; q < 8; ++q) {
          const double xq = x[j + q];
          a0[q] += v[q].x * xq;
          a1[q] += v[q].y * xq;
        }
      }
      for (; j < j1; ++j) {
        const double2 v = *(const double2*)(A + j * lda + r);
        a0[0] += v.x * x[j];
        a1[0] += v.y * x[j];
      }
      part[(long)blockIdx.y * m + r] = ((a0[0] + a0[1]) + (a0[2] + a0[3])) + ((a0[4] + a0[5]) + (a0[6] + a0[7]));
      part[(long)blockIdx.y * m + r + 1] = ((a1[0] + a1[1]) + (a1[2] + a1[3])) + ((a1[4] + a1[5]) + (a1[6] + a1[7]));
      return;
    }
    // the last row of an odd m: the single-row path below
  }
  const long r = pairs ? 2 * ((long)blockIdx.x * blockDim.x + threadIdx.x)
                       : (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= m) return;
  double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  long j = j0;
  for (; j + 8 <= j1; j += 8) {
#pragma unroll
    for (int q = 0; q < 8; ++q) acc[q] += A[(j + q) * lda + r] * x[j + q];
  }
  for (; j < j1; ++j) acc[0] += A[j * lda + r] * x[j];
  part[(long)blockIdx.y * m + r] = ((acc[0] + acc[1]) + (acc[2] + acc[3])) + ((acc[4] + acc[5]) + (acc[6] + acc[7]));
}

__global__ __launch_bounds__(256) void k_gemv_done(long m, int chunks, double alpha,
                                                   const double* __restrict__ part, double beta,
                                                   double* __restrict__ y) {
  const long r = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= m) return;
  double v = 0.0;
  for (int c = 0; c < chunks; ++c) v += part[(long)c * m + r];
  v *= alpha;
  y[r] = beta == 0.0 ? v : v + beta * y[r];
}

__global__ __launch_bounds__(256) void k_gemv_done_n(long m, int chunks, double alpha,
                                                     const double* __restrict__ part, double beta,
                                                     double* __restrict__ y) {
  const long r = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= m) return;
  double v = 0.0;
  for (int c = 0; c < chunks; ++c) v += part[(long)c * m + r];
  v *= alpha;
  y[r] = beta == 0.0 ? v : v + beta * y[r];
}

void gemv_f64(long m, long n, double alpha, const double* A, long lda, const double* x, double beta,
              double* y, double* ws, hipStream_t st, int chunks) {
  const bool pairs = (lda & 1) == 0 && (((size_t)A) & 15) == 0;
  const long rows_per_wg = pairs ? 512 : 256;
  chunks = std::max(1, std::min(chunks, kGemvChunks));
  hipLaunchKernelGGL(k_gemv_part, dim3((unsigned)((m + rows_per_wg - 1) / rows_per_wg), (unsigned)chunks),
                     dim3(256), 0, st, m, n, A, lda, x, ws);
  hipLaunchKernelGGL(k_gemv_done, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, st, m, chunks, alpha, ws,
                     beta, y);
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

// ---- 16-bit copy of A for the inner matrix-vector products of block-Jacobi GMRES -------------------
// A16[j * ld16 + i] = fp16(A[j * lda + i] 2^-e), one power-of-two exponent e per tile of kM16R rows x
// kM16C columns (256 x 32: a thread of the build holds 32 values), chosen so the tile's largest magnitude lands in [2^13, 2^14) (an FP16 significand of
// 11 bits for everything within 2^27 of it); scale[ct * row_tiles + rt] = 2^e. Rows m <= i < ld16 hold
// zeros, so the product reads whole 16-byte row groups.
constexpr int kM16R = 256, kM16C = 32;
long m16_ld(long n) { return (n + kM16R - 1) / kM16R * kM16R; }
size_t m16_scale_count(long n) { return (size_t)(m16_ld(n) / kM16R) * (size_t)((n + kM16C - 1) / kM16C); }

// one workgroup per tile: thread t holds rows 2 (t % 128) and 2 (t % 128) + 1 of the 32 columns
// c = t / 128 (mod 2) (16-byte nontemporal loads of A, read once), the tile maximum by wavefront and LDS
// reductions, then the FP16 values: a full tile goes through LDS, eight columns at a time, so that
// every thread stores eight rows of one column (16 bytes); an edge tile stores its pairs directly
__global__ __launch_bounds__(256) void k_m16_build(const double* __restrict__ A, long m, long n, long lda,
                                                   _Float16* __restrict__ A16, long ld16,
                                                   float* __restrict__ scale) {
  __shared__ double wmax[4];
  __shared__ unsigned tp[4][kM16R];
  // grid: x = column tile, y = row tile (consecutive workgroups take neighbouring 32-column blocks of
  // the same rows: measured 50.2 ms against 54.7 with x = row tile at 160,374 unknowns)
  const long rt = blockIdx.y, ct = blockIdx.x;
  const long i = rt * kM16R + 2 * (threadIdx.x & 127);
  const int c0 = threadIdx.x >> 7;  // this thread's columns: c0, c0 + 2, ...
  const long j0 = ct * kM16C;
  const int nc = (int)min((long)kM16C, n - j0);
  const bool pair = ((lda & 1) == 0) && ((((size_t)A) & 15) == 0) && i + 1 < m;
  const bool full = nc == kM16C && rt * kM16R + kM16R <= m && pair;  // uniform: lda, A, whole tile
  dvec2 v[kM16C / 2];
  double mx = 0.0;
#pragma unroll
  for (int u = 0; u < kM16C / 2; ++u) {
    const int c = c0 + 2 * u;
    dvec2 w = {0.0, 0.0};
    if (c < nc) {
      const double* p = A + (j0 + c) * lda + i;
      if (pair) {
        w = __builtin_nontemporal_load((const dvec2*)p);
      } else {
        if (i < m) w.x = p[0];
        if (i + 1 < m) w.y = p[1];
      }
    }
    v[u] = w;
    mx = fmax(mx, fmax(fabs(w.x), fabs(w.y)));
  }
  for (int o = 32; o > 0; o >>= 1) mx = fmax(mx, __shfl_xor(mx, o));
  if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = mx;
  __syncthreads();
  mx = fmax(fmax(wmax[0], wmax[1]), fmax(wmax[2], wmax[3]));
  int ex = 0;
  if (mx > 0) frexp(mx, &ex);  // mx = f 2^ex, f in [0.5, 1)
  const int e = mx > 0 ? ex - 14 : 0;
  const double down = ldexp(1.0, -e);
  if (threadIdx.x == 0) scale[ct * gridDim.y + rt] = ldexpf(1.0f, e);
  typedef _Float16 half2v __attribute__((ext_vector_type(2)));
  unsigned pk[kM16C / 2];  // the two FP16 values of column c0 + 2 u
#pragma unroll
  for (int u = 0; u < kM16C / 2; ++u) {
    half2v h;
    h.x = (_Float16)(float)(v[u].x * down);
    h.y = (_Float16)(float)(v[u].y * down);
    pk[u] = __builtin_bit_cast(unsigned, h);
  }
  if (full) {
    // group g: columns c0 + 2 (4 g + q), q < 4, of both halves; tp[q][128 c0 + t % 128] holds the pair
    // of rows 2 (t % 128) .. + 1; thread (q, half cc, octet o) then stores rows 8 o .. 8 o + 7
    const int q = threadIdx.x >> 6, cc = (threadIdx.x >> 5) & 1, oct = threadIdx.x & 31;
#pragma unroll
    for (int g = 0; g < kM16C / 8; ++g) {
      __syncthreads();  // the previous group's reads of tp are done
#pragma unroll
      for (int k = 0; k < 4; ++k) tp[k][(threadIdx.x & 127) + 128 * c0] = pk[4 * g + k];
      __syncthreads();
      const unsigned* src = &tp[q][128 * cc + 4 * oct];
      // nontemporal: the copy is read back only after all of it is written (46.4 ms with these
      // stores and the grid above, 50.2 with ordinary stores)
      typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
      const u32x4 w4 = {src[0], src[1], src[2], src[3]};
      __builtin_nontemporal_store(w4, (u32x4*)(A16 + (j0 + cc + 2 * (4 * g + q)) * ld16 + rt * kM16R + 8 * oct));
    }
  } else {
#pragma unroll
    for (int u = 0; u < kM16C / 2; ++u) {
      const int c = c0 + 2 * u;
      if (c < nc) *(unsigned*)(A16 + (j0 + c) * ld16 + i) = pk[u];
    }
  }
}

void m16_build(const double* A, long n, long lda, unsigned short* A16, long ld16, float* scale, hipStream_t st) {
  hipLaunchKernelGGL(k_m16_build, dim3((unsigned)((n + kM16C - 1) / kM16C), (unsigned)(ld16 / kM16R)), dim3(256), 0,
                     st, A, n, n, lda, (_Float16*)A16, ld16, scale);
}

// part[(chunk0 + c) * m + r] = sum over chunk c's columns of A16[r, j] 2^e x_j, the column tiles
// [t0, t1) split into gridDim.y chunks: a thread holds 8 consecutive rows (one 16-byte nontemporal
// load per column), FP32 sums within a 32-column tile, FP64 across tiles; the partial sums of all
// kM16Chunks chunks are added in a fixed order by k_gemv_done_n. (Measured at n = 160,374: 16 chunks
// with nontemporal loads 6.8 TB/s; 32 chunks with cached loads 6.1; 16 rows per thread 4.2.)
constexpr int kM16Chunks = 16;
__global__ __launch_bounds__(256) void k_gemv16_part(long m, long n, const _Float16* __restrict__ A16, long ld16,
                                                     const float* __restrict__ scale, long row_tiles,
                                                     const double* __restrict__ x, double* __restrict__ part,
                                                     long t0, long t1, int chunk0) {
  const long r0 = ((long)blockIdx.x * 256 + threadIdx.x) * 8;
  if (r0 >= m) return;
  const long tpc = (t1 - t0 + gridDim.y - 1) / gridDim.y;  // column tiles per chunk
  const long ta = t0 + (long)blockIdx.y * tpc, tz = min(t1, ta + tpc);
  const long rt = r0 / kM16R;
  double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (long ct = ta; ct < tz; ++ct) {
    const long j0 = ct * kM16C, j1 = min(n, j0 + kM16C);
    float at[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    long j = j0;
    for (; j + 8 <= j1; j += 8) {
      half8 h[8];
#pragma unroll
      for (int q = 0; q < 8; ++q) h[q] = __builtin_nontemporal_load((const half8*)(A16 + (j + q) * ld16 + r0));
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const float xq = (float)x[j + q];
#pragma unroll
        for (int k = 0; k < 8; ++k) at[k] = fmaf((float)h[q][k], xq, at[k]);
      }
    }
    for (; j < j1; ++j) {
      const half8 h = *(const half8*)(A16 + j * ld16 + r0);
      const float xq = (float)x[j];
#pragma unroll
      for (int k = 0; k < 8; ++k) at[k] = fmaf((float)h[k], xq, at[k]);
    }
    const double sc = (double)scale[ct * row_tiles + rt];
#pragma unroll
    for (int k = 0; k < 8; ++k) acc[k] += sc * (double)at[k];
  }
#pragma unroll
  for (int k = 0; k < 8; ++k)
    if (r0 + k < m) part[(long)(chunk0 + blockIdx.y) * m + r0 + k] = acc[k];
}

long m16_tiles(long n) { return (n + kM16C - 1) / kM16C; }

void gemv_m16_part(long n, const unsigned short* A16, long ld16, const float* scale, const double* x, double* ws,
                   long t0, long t1, int chunk0, int chunks, hipStream_t st) {
  static_assert(kM16Chunks <= kGemvChunks, "the product's partial sums share the FP64 product's workspace");
  hipLaunchKernelGGL(k_gemv16_part, dim3((unsigned)((n + 2047) / 2048), (unsigned)chunks), dim3(256), 0, st, n, n,
                     (const _Float16*)A16, ld16, scale, ld16 / kM16R, x, ws, t0, t1, chunk0);
}

void gemv_m16_finish(long n, double alpha, const double* ws, double beta, double* y, hipStream_t st, int chunks) {
  hipLaunchKernelGGL(k_gemv_done_n, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, n, chunks, alpha, ws,
                     beta, y);
}

void gemv_m16(long n, double alpha, const unsigned short* A16, long ld16, const float* scale, const double* x,
              double beta, double* y, double* ws, hipStream_t st) {
  gemv_m16_part(n, A16, ld16, scale, x, ws, 0, m16_tiles(n), 0, kM16Chunks, st);
  gemv_m16_finish(n, alpha, ws, beta, y, st, kM16Chunks);
}

}  // namespace ibem
