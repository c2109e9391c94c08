// Stresses at interior points (postProcessorElastic3D.cpp:224-587), FP64, gfx950.
//   s = C0 : ( sum_h D_4x/5x/6x . E_h  - eigenstrain polynomial if the point is inside a sphere )
//       + sum_e int_e ( Dsu u + Dst t ) dS          (C1 = 0 variant of the strain kernel)
// Inclusion term at a point outside sphere h: outside a sphere the eight potential tables are
// T_f = C_f0(a) d^beta r + C_f1(a) d^beta r^-1 at x - c_h, and D(ij | c) is linear in the tables, so
//   sum_c D(ij | c) w_c X_c = sum_beta d^beta r Vr_ij(beta) + d^beta r^-1 Vt_ij(beta)
// with per-sphere moments V (k_stress_mom: the transpose of D applied to w_c X_c, folded with C_fb;
// 144 per strain pair: d r of orders 4-6 and d r^-1 of orders 2-6). k_stress_incl evaluates that
// sum with lane = point over a chunk of spheres (the moments are wave-uniform, read through the
// scalar cache) and marks the spheres that contain the point; those few go through the tables in
// k_stress as before.
// k_stress: one workgroup per point (8 waves). Boundary term: wave v takes elements v, v+8, ...;
// lane = Gauss point; each lane builds the Kelvin tables it needs (d r of orders 3 and 4, d r^-1 of
// orders 1 and 2) in registers with generated straight-line code. Inclusion term: the chunk sums of
// k_stress_incl, plus the containing spheres four at a time, their eight potential tables built in
// LDS, then thread (ij, c) accumulating D(ij | c) X_c.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "ibem3d_kernels.h"
#include "ibem3d_math.h"

namespace ibem {

constexpr int kPostThreads = 512;  // 8 waves: the boundary term's element loop; waves 0-5 = strain pairs
constexpr int kPostTab = 84;  // |beta| <= 6 for the D_4x/5x/6x tables
constexpr int kPostSph = kPostThreads / kPostTab;  // containing spheres per round of the inclusion term (6)
constexpr int kMomR = 64, kMomT = 80, kMom = kMomR + kMomT;  // moments per strain pair
constexpr int kInclWaves = 4096;  // target waves of k_stress_incl (points x sphere chunks)

// Chunking of the spheres for k_stress_incl: at most 64 spheres per chunk (one mask bit each), and
// about kInclWaves waves over the points' 64-lane blocks.
static void incl_chunks(int num, int nump, int& chunks, int& per) {
  if (num <= 0 || nump <= 0) {
    chunks = per = 0;
    return;
  }
  const int pblocks = (nump + 63) / 64;
  long c = (kInclWaves + pblocks - 1) / pblocks;
  c = std::max<long>(c, (num + 63) / 64);
  c = std::min<long>(c, num);
  per = (int)((num + c - 1) / c);
  chunks = (num + per - 1) / per;
}

size_t stress_workspace_doubles(int num, int nump) {
  int chunks, per;
  incl_chunks(num, nump, chunks, per);
  // moments, chunk sums [point][ij][chunk], inside masks [chunk][point] (one 64-bit word each)
  return 6ul * kMom * std::max(num, 0) + 7ul * chunks * std::max(nump, 0);
}

IBEM_HD double sel8(const double* v, int f) {
  return f < 4 ? (f < 2 ? (f == 0 ? v[0] : v[1]) : (f == 2 ? v[2] : v[3]))
               : (f < 6 ? (f == 4 ? v[4] : v[5]) : (f == 6 ? v[6] : v[7]));
}

// Per-sphere moments: thread (ij, b) runs the transpose of D(ij | c) for c = 0..59 in order, weighted
// by w_c X_c (w_c = mapw, x2 for merged quadratic unknowns) and by the outside coefficient C_fb of
// each table entry; b = 0 collects the d r weights, b = 1 the d r^-1 weights.
__global__ __launch_bounds__(64) void k_stress_mom(const double* __restrict__ rad, Mat M,
                                                   const double* __restrict__ E60, double* __restrict__ V) {
  __shared__ double sv[12][kPostTab];
  const int h = blockIdx.x, tid = threadIdx.x;
  for (int u = tid; u < 12 * kPostTab; u += 64) (&sv[0][0])[u] = 0.0;
  __syncthreads();
  if (tid < 12) {
    const int ij = tid / 2, b = tid % 2;
    double C[8][4];
    pot_coefs(rad[h], false, C);
    double Cb[8];
#pragma unroll
    for (int f = 0; f < 8; ++f) Cb[f] = b ? C[f][1] : C[f][0];
    double* row = sv[tid];
    auto sink = [&](int f, int t, double v) { row[t] += v * sel8(Cb, f); };
    for (int c = 0; c < 60; ++c) {
      const Slot Cc = slot60(c);
      const double wc = mapw(Cc.i, Cc.j) * ((Cc.order == 2 && Cc.p != Cc.q) ? 2.0 : 1.0);
      Dten_adj(sink, M, Cc.w, pair_i(ij), pair_j(ij), Cc.i, Cc.j, wc * E60[60L * h + c]);
    }
  }
  __syncthreads();
  // V[h][ij][s]: s < 64 -> d r weight at tidx 20 + s, else d r^-1 weight at tidx s - 60
  for (int u = tid; u < 6 * kMom; u += 64) {
    const int ij = u / kMom, sl = u % kMom;
    V[(size_t)h * 6 * kMom + u] = sl < kMomR ? sv[2 * ij][20 + sl] : sv[2 * ij + 1][sl - 60];
  }
}

// Inclusion strain at 64 points (lane = point) from the spheres of chunk blockIdx.y that do not
// contain the point; the chunk's sum to part[(p * 6 + ij) * chunks + chunk], the containing spheres
// as bits of mask[chunk * nump + p].
__global__ __launch_bounds__(64) void k_stress_incl(const double* __restrict__ cen,
                                                    const double* __restrict__ rad, int num, int per,
                                                    int chunks, const double* __restrict__ pts, int nump,
                                                    const double* __restrict__ V, double* __restrict__ part,
                                                    unsigned long long* __restrict__ mask) {
  const int p = blockIdx.x * 64 + threadIdx.x;
  const int pc = p < nump ? p : nump - 1;
  const int k = blockIdx.y;
  const int h0 = k * per, h1 = min(num, h0 + per);
  const double x0 = pts[3 * pc], x1 = pts[3 * pc + 1], x2 = pts[3 * pc + 2];
  double e[6] = {0, 0, 0, 0, 0, 0};
  unsigned long long m = 0;
  for (int h = h0; h < h1; ++h) {
    const double x[3] = {x0 - cen[3 * h], x1 - cen[3 * h + 1], x2 - cen[3 * h + 2]};
    const bool inside = is_inside(x, rad[h]);
    double y0[7], y1[7], y2[7], PR[7], PT[7];
    kelvin_point<6>(x, y0, y1, y2, PR, PT);
    double eh[6] = {0, 0, 0, 0, 0, 0};
    incl_stress_moments(y0, y1, y2, PR, PT, V + (size_t)h * 6 * kMom, eh);
#pragma unroll
    for (int ij = 0; ij < 6; ++ij) e[ij] += inside ? 0.0 : eh[ij];
    m |= inside ? (1ull << (h - h0)) : 0ull;
  }
  if (p < nump) {
#pragma unroll
    for (int ij = 0; ij < 6; ++ij) part[((size_t)p * 6 + ij) * chunks + k] = e[ij];
    mask[(size_t)k * nump + p] = m;
  }
}

__global__ __launch_bounds__(kPostThreads) void k_stress(
    const double* __restrict__ qp, const int* __restrict__ conn, int NE,
    const double* __restrict__ cen, const double* __restrict__ rad, int num,
    const double* __restrict__ pts, Mat M, const double* __restrict__ U,
    const double* __restrict__ T, const double* __restrict__ E60, int chunks, int per, int nump,
    const double* __restrict__ part, const unsigned long long* __restrict__ mask, double* __restrict__ s) {
  // LDS: the inclusion term's potential tables, then the reduction buffers after them
  __shared__ double smem[kPostSph * 8 * kPostTab + kPostThreads * 6];
  double* ptab = smem;  // kPostSph x 8 x kPostTab
  double (*red)[6] = reinterpret_cast<double (*)[6]>(smem + kPostSph * 8 * kPostTab);  // kPostThreads x 6
  const int p = blockIdx.x;
  const int tid = threadIdx.x, wave = tid / 64, lane = tid % 64;
  const double x0 = pts[3 * p], x1 = pts[3 * p + 1], x2 = pts[3 * p + 2];
  double acc[6] = {0, 0, 0, 0, 0, 0};
  // ---- boundary term (post_stress_BEM :224-365) ----
  for (int e = wave; e < NE; e += kPostThreads / 64) {
    const double* Q = qp + ((size_t)e * 64 + lane) * kQpStride;
    const double x[3] = {Q[0] - x0, Q[1] - x1, Q[2] - x2};
    // the lane's Kelvin tables in registers (generated straight-line code, kelvin_stress)
    double y0[7], y1[7], y2[7], PR[7], PT[5], tab[45];
    kelvin_point(x, y0, y1, y2, PR, PT);
    kelvin_stress(y0, y1, y2, PR, PT, tab);
    double u[3] = {0, 0, 0}, tr[3] = {0, 0, 0};
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int nd = conn[4 * e + q] - 1;
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        u[k] += Q[6 + q] * U[3 * nd + k];
        tr[k] += Q[6 + q] * T[12 * e + 3 * q + k];
      }
    }
    const int e0[3] = {0, 0, 0};
    const double w = Q[10];
    // every table index below is a compile-time constant once these loops unroll
#pragma unroll
    for (int l = 0; l < 6; ++l)
#pragma unroll
      for (int n = 0; n < 3; ++n) {
        double su, st;
        strain_kernel_n(tab, tab + 35, M, M.lam0, M.mu0, Q + 3, pair_i(l), pair_j(l), e0, 1, n, su, st);
        acc[l] += w * (su * u[n] + st * tr[n]);
      }
  }
  // ---- inclusion term (postprocessor :367-522) ----
  // spheres the point lies outside of: the chunk sums of k_stress_incl, chunk k at lane k % 64 of
  // wave ij; the containing spheres (mask bits, in sphere order): kPostSph at a time, thread
  // (slot, u) builds entry u of the slot's eight potential tables, then thread (ij, c) accumulates
  // D(ij | c) X_c over the round's spheres in order
  const int ij = tid / 64, c = tid % 64;  // thread -> (strain pair, unknown)
  const int slot = tid / kPostTab, u = tid % kPostTab;
  double ein = 0.0;
  if (ij < 6)
    for (int k = c; k < chunks; k += 64) ein += part[((size_t)p * 6 + ij) * chunks + k];
  __shared__ int shs[kPostSph];
  __shared__ int snh;
  int inside_h = -1;
  int ck = 0;                                            // thread 0's cursor over the masks
  unsigned long long cm = chunks > 0 ? mask[p] : 0ull;
  for (;;) {
    __syncthreads();
    if (tid == 0) {
      int nh = 0;
      while (nh < kPostSph && ck < chunks) {
        if (cm == 0) {
          ++ck;
          cm = ck < chunks ? mask[(size_t)ck * nump + p] : 0ull;
          continue;
        }
        shs[nh++] = ck * per + (__ffsll((long long)cm) - 1);
        cm &= cm - 1;
      }
      snh = nh;
    }
    __syncthreads();
    const int nh = snh;
    if (nh == 0) break;
    inside_h = shs[nh - 1];
    if (slot < nh) {
      const int h = shs[slot];
      const double xh[3] = {x0 - cen[3 * h], x1 - cen[3 * h + 1], x2 - cen[3 * h + 2]};
      RadPoint P;
      rad_point(P, xh);
      double C[8][4];
      pot_coefs(rad[h], true, C);
      double v[8];
      pot_entry(P, true, C, u, v);
      for (int f = 0; f < 8; ++f) ptab[(slot * 8 + f) * kPostTab + u] = v[f];
    }
    __syncthreads();
    if (c < 60 && ij < 6) {
      // strain e_ij += D(i,j | m,n | w) map_mn (x2 for merged quadratic unknowns) X_c
      const Slot Cc = slot60(c);
      const double f = mapw(Cc.i, Cc.j) * ((Cc.order == 2 && Cc.p != Cc.q) ? 2.0 : 1.0);
      for (int k = 0; k < nh; ++k) {
        const double d = Dten(ptab + k * 8 * kPostTab, kPostTab, M, Cc.w, pair_i(ij), pair_j(ij), Cc.i,
                              Cc.j, 0, 0, 0);
        ein += d * f * E60[60L * shs[k] + c];
      }
    }
    if (nh < kPostSph) break;
  }
  // reduce the inclusion strain over c for each ij (one wave per ij), then the boundary term
  red[tid][0] = ein;
  __syncthreads();
  for (int sft = 32; sft > 0; sft >>= 1) {
    if (lane < sft) red[tid][0] += red[tid + sft][0];
    __syncthreads();
  }
  __shared__ double estr[6];
  if (lane == 0 && wave < 6) estr[wave] = red[tid][0];
  __syncthreads();
  // boundary term: reduce each wave's 64 lanes, then thread 0 adds the six wave sums
  for (int l = 0; l < 6; ++l) red[tid][l] = acc[l];
  __syncthreads();
  for (int sft = 32; sft > 0; sft >>= 1) {
    if (lane < sft)
      for (int l = 0; l < 6; ++l) red[tid][l] += red[tid + sft][l];
    __syncthreads();
  }
  __shared__ double sbnd[6];
  if (tid == 0)
    for (int l = 0; l < 6; ++l) {
      double v = 0.0;
      for (int wv = 0; wv < kPostThreads / 64; ++wv) v += red[64 * wv][l];
      sbnd[l] = v;
    }
  __syncthreads();
  if (tid == 0) {
    double e6[6];
    for (int l = 0; l < 6; ++l) e6[l] = estr[l];
    if (inside_h >= 0) {  // last containing sphere wins (:383-390)
      const double* E = E60 + 60L * inside_h;
      const double d0 = x0 - cen[3 * inside_h], d1 = x1 - cen[3 * inside_h + 1],
                   d2 = x2 - cen[3 * inside_h + 2];
      for (int t = 0; t < 6; ++t) {
        double v = E[t] + E[6 + 3 * t] * d0 + E[6 + 3 * t + 1] * d1 + E[6 + 3 * t + 2] * d2;
        for (int kl = 0; kl < 6; ++kl) {
          const int k = pair_i(kl), l = pair_j(kl);
          v += mapw(k, l) * E[24 + 6 * t + kl] * sel3(d0, d1, d2, k) * sel3(d0, d1, d2, l);
        }
        e6[t] -= v;
      }
    }
    for (int t = 0; t < 6; ++t) {
      const int i = pair_i(t), j = pair_j(t);
      double v = 0.0;
      for (int mn = 0; mn < 6; ++mn) v += C0(M, i, j, pair_i(mn), pair_j(mn)) * mapw(pair_i(mn), pair_j(mn)) * e6[mn];
      s[6L * p + t] = v + sbnd[t];
    }
  }
}

// The stress kernel's inputs from the device solution x: U[i] for the 3 NN node dofs (unknown:
// x[node_col[i]], known: uval[i]) and T per element slot (unknown: x[slot_col], known: tval, -2:
// 0), the values getUTVector assembles (BEMBuilderElastic3D.cpp:2053-2117).
__global__ __launch_bounds__(256) void k_stress_inputs(const double* __restrict__ x,
                                                       const int* __restrict__ node_col,
                                                       const double* __restrict__ uval,
                                                       const int* __restrict__ slot_col,
                                                       const double* __restrict__ tval, int NN, int NE,
                                                       double* __restrict__ U, double* __restrict__ T) {
  const long i = (long)blockIdx.x * 256 + threadIdx.x;
  if (i < 3L * NN) {
    const int c = node_col[i];
    U[i] = c >= 0 ? x[c] : uval[i];
  } else if (i < 3L * NN + 12L * NE) {
    const long k = i - 3L * NN;
    const int c = slot_col[k];
    T[k] = c == -1 ? tval[k] : (c >= 0 ? x[c] : 0.0);
  }
}

void launch_stress_inputs(const double* x, const int* node_col, const double* uval, const int* slot_col,
                          const double* tval, int NN, int NE, double* U, double* T, hipStream_t st) {
  const long m = 3L * NN + 12L * NE;
  hipLaunchKernelGGL(k_stress_inputs, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, st, x, node_col,
                     uval, slot_col, tval, NN, NE, U, T);
}

void launch_stress(const double* qp, const int* conn, const double* nodes, int NN, int NE,
                   const double* cen, const double* rad, int num, const double* pts, int nump,
                   const Mat& M, const double* U, const double* T, const double* E60, double* ws,
                   double* s, hipStream_t st) {
  (void)nodes;
  (void)NN;
  if (nump <= 0) return;
  int chunks, per;
  incl_chunks(num, nump, chunks, per);
  double* V = ws;
  double* part = V + 6L * kMom * std::max(num, 0);
  auto* mask = reinterpret_cast<unsigned long long*>(part + 6L * chunks * nump);
  if (chunks > 0) {
    hipLaunchKernelGGL(k_stress_mom, dim3(num), dim3(64), 0, st, rad, M, E60, V);
    hipLaunchKernelGGL(k_stress_incl, dim3((nump + 63) / 64, chunks), dim3(64), 0, st, cen, rad, num, per,
                       chunks, pts, nump, V, part, mask);
  }
  hipLaunchKernelGGL(k_stress, dim3(nump), dim3(kPostThreads), 0, st, qp, conn, NE, cen, rad, num, pts,
                     M, U, T, E60, chunks, per, nump, part, mask, s);
}

}  // namespace ibem
