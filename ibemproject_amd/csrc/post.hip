// Stresses at interior points (postProcessorElastic3D.cpp:224-587), FP64, gfx950.
//   s = C0 : ( sum_h D_4x/5x/6x . E_h  - eigenstrain polynomial if the point is inside a sphere )
//       + sum_e int_e ( Dsu u + Dst t ) dS          (C1 = 0 variant of the strain kernel)
// One workgroup per point (6 waves). Boundary term: wave v takes elements v, v+6, ...; lane = Gauss
// point; each lane builds the Kelvin tables it needs (d r of orders 3 and 4, d r^-1 of orders 1
// and 2) in registers with generated straight-line code. Inclusion term: four spheres at a time, their eight potential tables built in LDS, then
// thread (ij, c) accumulates D(ij | c) X_c.
#include <hip/hip_runtime.h>

#include "ibem3d_kernels.h"
#include "ibem3d_math.h"

namespace ibem {

constexpr int kPostThreads = 384;
constexpr int kPostTab = 84;  // |beta| <= 6 for the D_4x/5x/6x tables
constexpr int kPostSph = kPostThreads / kPostTab;  // spheres per round of the inclusion term (4)

__global__ __launch_bounds__(kPostThreads) void k_stress(
    const double* __restrict__ qp, const int* __restrict__ conn, int NE,
    const double* __restrict__ cen, const double* __restrict__ rad, int num,
    const double* __restrict__ pts, Mat M, const double* __restrict__ U,
    const double* __restrict__ T, const double* __restrict__ E60, double* __restrict__ s) {
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
  // kPostSph spheres per round: thread (slot, u) builds entry u of the slot's eight potential
  // tables, then thread (ij, c) accumulates D(ij | c) X_c over the round's spheres in order
  const int ij = tid / 64, c = tid % 64;  // thread -> (strain pair, unknown)
  const int slot = tid / kPostTab, u = tid % kPostTab;
  double ein = 0.0;
  int inside_h = -1;
  for (int h0 = 0; h0 < num; h0 += kPostSph) {
    const int nh = min(kPostSph, num - h0);
    for (int k = 0; k < nh; ++k) {
      const int h = h0 + k;
      const double xh[3] = {x0 - cen[3 * h], x1 - cen[3 * h + 1], x2 - cen[3 * h + 2]};
      const double a = rad[h];
      if (xh[0] * xh[0] + xh[1] * xh[1] + xh[2] * xh[2] - a * a <= 0) inside_h = h;
    }
    __syncthreads();
    if (slot < nh) {
      const int h = h0 + slot;
      const double xh[3] = {x0 - cen[3 * h], x1 - cen[3 * h + 1], x2 - cen[3 * h + 2]};
      const double a = rad[h];
      RadPoint P;
      rad_point(P, xh);
      const bool inside = is_inside(xh, a);
      double C[8][4];
      pot_coefs(a, inside, C);
      double v[8];
      pot_entry(P, inside, C, u, v);
      for (int f = 0; f < 8; ++f) ptab[(slot * 8 + f) * kPostTab + u] = v[f];
    }
    __syncthreads();
    if (c < 60) {
      // strain e_ij += D(i,j | m,n | w) map_mn (x2 for merged quadratic unknowns) X_c
      const Slot Cc = slot60(c);
      const double f = mapw(Cc.i, Cc.j) * ((Cc.order == 2 && Cc.p != Cc.q) ? 2.0 : 1.0);
      for (int k = 0; k < nh; ++k) {
        const double d = Dten(ptab + k * 8 * kPostTab, kPostTab, M, Cc.w, pair_i(ij), pair_j(ij), Cc.i,
                              Cc.j, 0, 0, 0);
        ein += d * f * E60[60L * (h0 + k) + c];
      }
    }
  }
  // reduce the inclusion strain over c for each ij (one wave per ij), then the boundary term
  red[tid][0] = ein;
  __syncthreads();
  for (int sft = 32; sft > 0; sft >>= 1) {
    if (lane < sft) red[tid][0] += red[tid + sft][0];
    __syncthreads();
  }
  __shared__ double estr[6];
  if (lane == 0) estr[wave] = red[tid][0];
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
                   const Mat& M, const double* U, const double* T, const double* E60, double* s,
                   hipStream_t st) {
  (void)nodes;
  (void)NN;
  if (nump <= 0) return;
  hipLaunchKernelGGL(k_stress, dim3(nump), dim3(kPostThreads), 0, st, qp, conn, NE, cen, rad, num, pts,
                     M, U, T, E60, s);
}

}  // namespace ibem
