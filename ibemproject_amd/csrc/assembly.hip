// Assembly kernels of the elastic-3D sphere system, FP64, gfx950.
//
// The system is built directly in its BC-folded, column-major, 60-unknown layout (n = 3NN + 60num):
//   rows    [0, 3NN)           boundary collocation rows (3 per node)
//   rows    [3NN, n)           inclusion rows (60 per sphere)
//   columns [0, 3NN)           displacement-unknown node dofs and traction-unknown keys (column map)
//   columns [3NN, n)           eigenstrain unknowns (60 per sphere)
// Every kernel owns the rows it writes, so accumulation is deterministic and needs no atomics:
//   k_incl   inclusion rows x eigenstrain columns  (integratorElastic3D.cpp:341-682)
//   k_field  boundary rows  x eigenstrain columns  (integratorElastic3D.cpp:135-340)
//   k_bem    boundary rows  x boundary columns + b (BEMBuilderElastic3D.cpp:1355-1509, ApplyBC)
//   k_rows   inclusion rows x boundary columns + b (BEMBuilderElastic3D.cpp:1511-1867, ApplyBC)
#include <hip/hip_runtime.h>

#include "ibem3d_kernels.h"
#include "ibem3d_math.h"

namespace ibem {

// ---- element quadrature geometry: 64 Gauss points x (g[3], nv[3], N[4], w*jac) -----------------
__global__ __launch_bounds__(64) void k_qpoints(const double* __restrict__ nodes,
                                                const int* __restrict__ conn, int NE,
                                                double* __restrict__ qp) {
  const int e = blockIdx.x;
  const int g = threadIdx.x;
  if (e >= NE) return;
  double el[12];
  for (int q = 0; q < 4; ++q) {
    const int nd = conn[4 * e + q] - 1;
    for (int c = 0; c < 3; ++c) el[3 * q + c] = nodes[3 * nd + c];
  }
  const int a = g / 8, b = g % 8;  // xi index outer, eta inner (BEMBuilderElastic3D.cpp:1247-1252)
  QPoint Q;
  qpoint(el, gauss8_x(a), gauss8_x(b), Q);
  double* o = qp + ((size_t)e * 64 + g) * kQpStride;
  o[0] = Q.g[0];
  o[1] = Q.g[1];
  o[2] = Q.g[2];
  o[3] = Q.nv[0];
  o[4] = Q.nv[1];
  o[5] = Q.nv[2];
  o[6] = Q.N[0];
  o[7] = Q.N[1];
  o[8] = Q.N[2];
  o[9] = Q.N[3];
  o[10] = gauss8_w(a) * gauss8_w(b) * Q.jac;
}

// ---- K1: inclusion-inclusion blocks -----------------------------------------------------------
// Workgroup = kInclTile consecutive target spheres (rows) x one source sphere h (columns). Per pair:
// the four radial basis tables (d^beta of r, r^-1 outside; of rho .. rho^4 inside, |beta| <= 8) and
// the potential coefficients, then the weighted tables Psi^w,g / Phi^w,g for the ten weights, all in
// LDS; then thread = row of the tile: its table positions once (incl_row), and the 60 columns
// unrolled with the column decode resolved at compile time and the row's conditions as selects
// (incl_row_entries), one coalesced column segment per column (rows of consecutive targets are
// contiguous). The weighted tables replace the basis tables
// in the same LDS (each thread holds its kInclPer weighted values in registers across a barrier), so
// that 38 KB per workgroup lets four workgroups share a CU.
constexpr int kInclThreads = 256;
constexpr int kInclU = kInclTile * kInclW;                              // weighted values per workgroup
constexpr int kInclPer = (kInclU + kInclThreads - 1) / kInclThreads;  // per thread
static_assert(4 * kTabN + kRadF <= kInclW, "a pair's basis tables and radial factors fit in its weighted-table slot");
static_assert(kInclTile * 64 <= kInclThreads, "one wave per target sphere in the entry phase");

__global__ __launch_bounds__(kInclThreads, 4) void k_incl(const double* __restrict__ cen,
                                                       const double* __restrict__ rad, int num, long lda,
                                                       long row0, Mat M, int rowmask,
                                                       const int* __restrict__ wdesc,
                                                       double* __restrict__ A) {
  // pair ps: basis tables [4][kTabN] at sU + ps kInclW, later its weighted tables [PS | PH] there
  __shared__ double sU[kInclU];
  __shared__ double sC[kInclTile][8][4];
  const int h = blockIdx.y;
  const int s0 = blockIdx.x * kInclTile;
  const double ch0 = cen[3 * h], ch1 = cen[3 * h + 1], ch2 = cen[3 * h + 2];
  const double a = rad[h];
  const int t = threadIdx.x;
  if (t < kInclTile && s0 + t < num) {
    const int s = s0 + t;
    const double x[3] = {cen[3 * s] - ch0, cen[3 * s + 1] - ch1, cen[3 * s + 2] - ch2};
    double C[8][4];
    pot_coefs(a, is_inside(x, a), C);
#pragma unroll
    for (int f = 0; f < 8; ++f)
#pragma unroll
      for (int b = 0; b < 4; ++b) sC[t][f][b] = C[f][b];
    // the pair's radial factors and coordinate powers, after its basis tables (outside pairs)
    rad_factors(x, sU + t * kInclW + 4 * kTabN);
  }
  __syncthreads();
  for (int u = t; u < kInclTile * kTabN; u += kInclThreads) {
    const int ps = u / kTabN, tt = u % kTabN;
    const int s = s0 + ps;
    if (s >= num) continue;
    const double x[3] = {cen[3 * s] - ch0, cen[3 * s + 1] - ch1, cen[3 * s + 2] - ch2};
    int b0, b1, b2;
    tdecode(tt, b0, b1, b2);
    double B0, B1, B2 = 0.0, B3 = 0.0;
    if (is_inside(x, a)) {
      RadPoint P;
      rad_point(P, x);
      rad_basis(P, true, b0, b1, b2, B0, B1, B2, B3);
    } else {
      rad_basis_out(sU + ps * kInclW + 4 * kTabN, b0, b1, b2, B0, B1);
    }
    double* basis = sU + ps * kInclW;
    basis[tt] = B0;
    basis[kTabN + tt] = B1;
    basis[2 * kTabN + tt] = B2;
    basis[3 * kTabN + tt] = B3;
  }
  __syncthreads();
  double wv[kInclPer];
#pragma unroll
  for (int k = 0; k < kInclPer; ++k) {
    const int u = t + kInclThreads * k;
    const int ps = u / kInclW, v = u % kInclW;
    wv[k] = (u < kInclU && s0 + ps < num) ? incl_weighted_val_d(sU + ps * kInclW, sC[ps], wdesc[v]) : 0.0;
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < kInclPer; ++k) {
    const int u = t + kInclThreads * k;
    if (u < kInclU) sU[u] = wv[k];
  }
  __syncthreads();
  const int ps = t / 64, r = t % 64;
  const int s = s0 + ps;
  if (r >= 60 || s >= num) return;
  const InclRow R = incl_row(r);
  if (!((rowmask >> R.order) & 1)) return;
  const long row = row0 + 60L * s + r;
  double* Acol = A + (row0 + 60L * h) * lda + row;
  const double* PS = sU + ps * kInclW;
  incl_row_entries(incl_row_pack(R), PS, PS + kWgt * kPS, M, s == h, Acol, lda);
}

// ---- K3: boundary rows x eigenstrain columns ----------------------------------------------------
// Thread = (node j, component i); row 3j+i; workgroup = 64 nodes x one sphere. The node's three
// threads first fill its table of radial values (field_tab_fill, 55 multi-indices) in LDS; then
// each writes its row's 60 columns from the table.
__global__ __launch_bounds__(192) void k_field(const double* __restrict__ nodes, int NN,
                                               const double* __restrict__ cen,
                                               const double* __restrict__ rad, long lda, long row0,
                                               Mat M, double* __restrict__ A) {
  __shared__ double stab[64][kFieldTab];
  const int K = blockIdx.y;
  const int jl = threadIdx.x / 3, i = threadIdx.x % 3;
  const int j = blockIdx.x * 64 + jl;
  const int jj = j < NN ? j : NN - 1;  // rows past NN fill a table nobody reads
  const double x[3] = {nodes[3 * jj] - cen[3 * K], nodes[3 * jj + 1] - cen[3 * K + 1],
                       nodes[3 * jj + 2] - cen[3 * K + 2]};
  const double a = rad[K];
  RadPoint P;
  rad_point(P, x);
  const bool inside = is_inside(x, a);
  double C[8][4];
  pot_coefs(a, inside, C);
  if (inside) {
    for (int u = i; u < kFieldMi; u += 3) field_tab_fill(P, inside, C, u, stab[jl]);
  } else if (i == 0) {
    // outside (every node of a valid RVE): one thread per node forms all 55 multi-indices' d r and
    // d r^-1 in straight-line code and writes the tables' entries (no lane-dependent loops)
    double y0[7], y1[7], y2[7], PR[7], PT[6];
    kelvin_point<5>(x, y0, y1, y2, PR, PT);
    double* tab = stab[jl];
    auto out = [&](int b0, int b1, int b2, double dr, double dri) {
      const int o = b0 + b1 + b2, idx = mi_index(b0, b1, b2);
#pragma unroll
      for (int f = 0; f < 8; ++f)
        if (field_tab_order(f) == o) tab[field_tab_off(f) + idx] = C[f][0] * dr + C[f][1] * dri;
    };
    field_basis(y0, y1, y2, PR, PT, out);
  }
  __syncthreads();
  if (j >= NN) return;
  const long row = 3L * j + i;
  double* Acol = A + (row0 + 60L * K) * lda + row;
  for (int c = 0; c < 60; ++c) Acol[(long)c * lda] = field_entry_tab(stab[jl], M, i, c);
}

// ---- K4: boundary rows x boundary columns, b, rigid-body diagonal -------------------------------
// Row = (collocation node i, component ii). For every element the row's Kelvin U*, T* are integrated
// against the four shape functions (8 x 8 Gauss, or 2 x 8 x 8 on collapsed triangles when node i
// belongs to the element, skipping T for node i's own shape function: BEMBuilderElastic3D.cpp:
// 1169-1353), and the 12 + 12 values are folded through the column map: H columns of
// displacement-unknown dofs and -G columns of traction-unknown keys into A, known values into b
// (ApplyBC :1960-2016). The diagonal block is minus the row sum of H over all node columns
// (:1478-1501). Two kernels per round of kBemChunk elements, so that the quadrature is spread over
// (node block, element block) workgroups rather than over the 3 NN rows alone:
//   k_bem_quad  workgroup = 64 nodes x 3 components x kBemEPB elements: the 24 values of every
//               (row, element) pair into a workspace, [element][value][row];
//   k_bem_fold  thread = (row, component): the values in element order through the column map
//               (the order of a single pass over all elements, so results do not depend on the
//               chunking); k_bem_finish adds the three components' known-value sums to b.
constexpr int kBemEPB = 4;  // elements per quadrature workgroup

__global__ __launch_bounds__(192) void k_bem_quad(const double* __restrict__ qp,
                                                  const int* __restrict__ conn,
                                                  const double* __restrict__ nodes, int NN, int e0,
                                                  int ne, double E0, double nu0,
                                                  double* __restrict__ W) {
  __shared__ double sq[64 * kQpStride];
  const long R = 3L * NN;
  const int i = blockIdx.x * 64 + threadIdx.x / 3;
  const int ii = threadIdx.x % 3;
  const bool active = i < NN;
  const int ic = active ? i : 0;
  const double xi0 = nodes[3 * ic], xi1 = nodes[3 * ic + 1], xi2 = nodes[3 * ic + 2];
  const long row = 3L * i + ii;
  const KelvinC K = kelvin_consts(E0, nu0);
  const int el0 = blockIdx.y * kBemEPB, el1 = min(el0 + kBemEPB, ne);
  for (int el = el0; el < el1; ++el) {
    const int e = e0 + el;
    __syncthreads();
    for (int u = threadIdx.x; u < 64 * kQpStride; u += blockDim.x)
      sq[u] = qp[(size_t)e * 64 * kQpStride + u];
    __syncthreads();
    if (!active) continue;
    int sc[4];
    int lpi = -1;
    for (int q = 0; q < 4; ++q) {
      sc[q] = conn[4 * e + q];
      if (sc[q] == i + 1) lpi = q;
    }
    double aU[4][3], aT[4][3];
    for (int q = 0; q < 4; ++q)
      for (int c = 0; c < 3; ++c) aU[q][c] = aT[q][c] = 0.0;
    if (lpi < 0) {
      for (int g = 0; g < 64; ++g) {
        const double* Q = sq + g * kQpStride;
        const double d0 = Q[0] - xi0, d1 = Q[1] - xi1, d2 = Q[2] - xi2;
        const double ir = 1.0 / sqrt(d0 * d0 + d1 * d1 + d2 * d2);
        const double dxr[3] = {d0 * ir, d1 * ir, d2 * ir};
        double up[3], tp[3];
        kelvin_ut_row(ii, dxr, ir, K, Q + 3, up, tp);
        const double w = Q[10];
        for (int q = 0; q < 4; ++q) {
          const double nw = Q[6 + q] * w;
          for (int c = 0; c < 3; ++c) {
            aU[q][c] += up[c] * nw;
            aT[q][c] += tp[c] * nw;
          }
        }
      }
    } else {
      double elx[12];
      for (int q = 0; q < 4; ++q)
        for (int c = 0; c < 3; ++c) elx[3 * q + c] = nodes[3 * (sc[q] - 1) + c];
      for (int ntr = 0; ntr < 2; ++ntr)
        for (int ga = 0; ga < 8; ++ga)
          for (int gb = 0; gb < 8; ++gb) {
            double xs, es, jb;
            duffy(ntr, lpi, gauss8_x(ga), gauss8_x(gb), xs, es, jb);
            QPoint Q;
            qpoint(elx, xs, es, Q);
            const double d0 = Q.g[0] - xi0, d1 = Q.g[1] - xi1, d2 = Q.g[2] - xi2;
            const double ir = 1.0 / sqrt(d0 * d0 + d1 * d1 + d2 * d2);
            const double dxr[3] = {d0 * ir, d1 * ir, d2 * ir};
            double up[3], tp[3];
            kelvin_ut_row(ii, dxr, ir, K, Q.nv, up, tp);
            const double w = gauss8_w(ga) * gauss8_w(gb) * Q.jac * jb;
            for (int q = 0; q < 4; ++q) {
              const double nw = Q.N[q] * w;
              for (int c = 0; c < 3; ++c) {
                aU[q][c] += up[c] * nw;
                if (q != lpi) aT[q][c] += tp[c] * nw;
              }
            }
          }
    }
    double* out = W + (size_t)el * 24 * R + row;
    for (int q = 0; q < 4; ++q)
      for (int c = 0; c < 3; ++c) {
        out[(3 * q + c) * R] = aU[q][c];
        out[(12 + 3 * q + c) * R] = aT[q][c];
      }
  }
}

// Thread = (row, component c): the columns of component c only (node dof 3 nd + c, traction keys
// of component c), so the three threads of a row never write the same entry. rs and bacc carry the
// row sum of H and the known-value terms of (row, c) across rounds ([c][R] each); the last round
// adds the rigid-body diagonal and leaves the three known-value sums for k_bem_finish.
__global__ __launch_bounds__(256) void k_bem_fold(const int* __restrict__ conn, int NN, int e0, int ne,
                                                  long lda, const int* __restrict__ node_col,
                                                  const int* __restrict__ slot_col,
                                                  const double* __restrict__ uval,
                                                  const double* __restrict__ tval,
                                                  const double* __restrict__ W,
                                                  double* __restrict__ carry, int first, int last,
                                                  double* __restrict__ A, int write_A) {
  const long R = 3L * NN;
  const long row = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const int c = blockIdx.y;
  if (row >= R) return;
  double rs = first ? 0.0 : carry[c * R + row];
  double bacc = first ? 0.0 : carry[(3 + c) * R + row];
  for (int el = 0; el < ne; ++el) {
    const int e = e0 + el;
    const double* in = W + (size_t)el * 24 * R + row;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int nd = conn[4 * e + q] - 1;
      const int dof = 3 * nd + c;
      const double hv = in[(12 + 3 * q + c) * R];
      rs += hv;
      const int col = node_col[dof];
      if (col >= 0) {
        if (write_A) A[(long)col * lda + row] += hv;
      } else {
        bacc -= hv * uval[dof];
      }
      const int slot = 12 * e + 3 * q + c;
      const double gv = in[(3 * q + c) * R];
      const int kc = slot_col[slot];
      if (kc >= 0) {
        if (write_A) A[(long)kc * lda + row] -= gv;
      } else if (kc == -1) {
        bacc += gv * tval[slot];
      }
    }
  }
  if (last) {
    const int dof = 3 * (int)(row / 3) + c;
    const int col = node_col[dof];
    if (col >= 0) {
      if (write_A) A[(long)col * lda + row] -= rs;
    } else {
      bacc += rs * uval[dof];
    }
  }
  carry[c * R + row] = rs;
  carry[(3 + c) * R + row] = bacc;
}

// ---- the boundary phase as a gather (whole-mesh workspace) --------------------------------------
// After k_bem_quad over every element (W[e][v][row]):
//   k_bem_rowsum  thread = (row, component c, element part p): over the part's elements in order,
//                 rs = the sum of H over every node dof of component c (the rigid-body diagonal,
//                 BEMBuilderElastic3D.cpp:1478-1501) and bk = the known-value terms of b (ApplyBC
//                 :1960-2016): -H u for known displacements, +G t for known tractions
//   k_bem_rowfin  thread = row: the parts added in order; the own dof's diagonal term goes to b when
//                 its displacement is known; b[row] += the three components' sums in order; rs kept
//   k_bem_gather  thread = (row, column): the column's H (+) and G (-) values in element order,
//                 minus rs_c of the row when the column is the row's own node dof c; written once.
constexpr int kBemParts = 8;
size_t bem_gather_workspace_doubles(int NN, int NE) {
  const size_t R = 3 * (size_t)NN;
  return (size_t)NE * 24 * R + (size_t)kBemParts * 6 * R + 3 * R;
}

__global__ __launch_bounds__(256) void k_bem_rowsum(const int* __restrict__ conn, int NN, int NE,
                                                    const int* __restrict__ node_col,
                                                    const int* __restrict__ slot_col,
                                                    const double* __restrict__ uval,
                                                    const double* __restrict__ tval,
                                                    const double* __restrict__ W, double* __restrict__ part) {
  const long R = 3L * NN;
  const long row = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const int c = blockIdx.y, p = blockIdx.z;
  if (row >= R) return;
  const int e0 = (int)((long)NE * p / kBemParts), e1 = (int)((long)NE * (p + 1) / kBemParts);
  double rs = 0.0, bk = 0.0;
  for (int e = e0; e < e1; ++e) {
    const double* in = W + (size_t)e * 24 * R + row;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int dof = 3 * (conn[4 * e + q] - 1) + c;
      const double hv = in[(12 + 3 * q + c) * R];
      rs += hv;
      if (node_col[dof] < 0) bk -= hv * uval[dof];
      const int slot = 12 * e + 3 * q + c;
      if (slot_col[slot] == -1) bk += in[(3 * q + c) * R] * tval[slot];
    }
  }
  part[((size_t)p * 6 + c) * R + row] = rs;
  part[((size_t)p * 6 + 3 + c) * R + row] = bk;
}

__global__ __launch_bounds__(256) void k_bem_rowfin(int NN, const int* __restrict__ node_col,
                                                    const double* __restrict__ uval,
                                                    const double* __restrict__ part, double* __restrict__ rs,
                                                    double* __restrict__ b) {
  const long R = 3L * NN;
  const long row = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= R) return;
  double bsum[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    double r = 0.0, k = 0.0;
    for (int p = 0; p < kBemParts; ++p) {
      r += part[((size_t)p * 6 + c) * R + row];
      k += part[((size_t)p * 6 + 3 + c) * R + row];
    }
    const int dof = 3 * (int)(row / 3) + c;
    if (node_col[dof] < 0) k += r * uval[dof];
    rs[c * R + row] = r;
    bsum[c] = k;
  }
  b[row] += (bsum[0] + bsum[1]) + bsum[2];
}

constexpr int kBemGatherCols = 8;  // columns per workgroup of the gather
__global__ __launch_bounds__(256) void k_bem_gather(int NN, long lda, const int* __restrict__ cptr,
                                                    const int* __restrict__ cent, const int* __restrict__ cdof,
                                                    const double* __restrict__ W, const double* __restrict__ rs,
                                                    double* __restrict__ A) {
  const long R = 3L * NN;
  const long row = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= R) return;
  const int i = (int)(row / 3);
  const long j0 = (long)blockIdx.y * kBemGatherCols, j1 = min(R, j0 + kBemGatherCols);
  for (long j = j0; j < j1; ++j) {
    double v = 0.0;
    for (int k = cptr[j]; k < cptr[j + 1]; ++k) {
      const int x = cent[k];
      const double w = W[(size_t)(x & 0x7fffffff) * R + row];
      v = x < 0 ? v - w : v + w;
    }
    const int d = cdof[j];
    if (d >= 0 && d / 3 == i) v -= rs[(d % 3) * R + row];
    A[j * lda + row] = v;
  }
}

void launch_bem_gather(const double* qp, const int* conn, const double* nodes, int NN, int NE, long lda,
                       double E0, double nu0, const int* node_col, const int* slot_col, const double* uval,
                       const double* tval, double* A, double* b, double* ws, const BemGather& G,
                       hipStream_t st, bool write_A) {
  const long R = 3L * NN;
  double* W = ws;
  double* part = ws + (size_t)NE * 24 * R;
  double* rs = part + (size_t)kBemParts * 6 * R;
  hipLaunchKernelGGL(k_bem_quad, dim3((NN + 63) / 64, (NE + kBemEPB - 1) / kBemEPB), dim3(192), 0, st, qp, conn,
                     nodes, NN, 0, NE, E0, nu0, W);
  const unsigned rb = (unsigned)((R + 255) / 256);
  hipLaunchKernelGGL(k_bem_rowsum, dim3(rb, 3, kBemParts), dim3(256), 0, st, conn, NN, NE, node_col, slot_col, uval,
                     tval, W, part);
  hipLaunchKernelGGL(k_bem_rowfin, dim3(rb), dim3(256), 0, st, NN, node_col, uval, part, rs, b);
  if (write_A)
    hipLaunchKernelGGL(k_bem_gather, dim3(rb, (unsigned)((R + kBemGatherCols - 1) / kBemGatherCols)), dim3(256), 0,
                       st, NN, lda, G.cptr, G.cent, G.cdof, W, rs, A);
}

// b[row] += the three components' known-value sums, in component order
__global__ __launch_bounds__(256) void k_bem_finish(int NN, const double* __restrict__ carry,
                                                    double* __restrict__ b) {
  const long R = 3L * NN;
  const long row = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= R) return;
  b[row] += (carry[3 * R + row] + carry[4 * R + row]) + carry[5 * R + row];
}

// ---- K2: inclusion rows x boundary columns + b -------------------------------------------------
// Workgroup = one sphere s (its 60 rows), 256 threads = 4 wavefronts; loops over every element.
// Per element: (1) the element's 64 Gauss points are staged in LDS; (2) lane = Gauss point and
// wavefront = entry group build the 64 Kelvin derivative tables at x = point - centre
// (BEMBuilderElastic3D.cpp:800-803) in LDS (kelvin_group, generated straight-line code);
// (3) lane = row, wavefront w integrates Gauss points 16w .. 16w+15: the row's strain kernels
// (row_gp, the contraction of get_for_stress_BEM{,1,2}) times shape function x weight x jacobian,
// for the three components n and the four nodes, H and G parts; (4) the four partial sums meet in
// LDS and are folded through the column map as in ApplyBC: -fac * H into displacement-unknown
// columns or b, fac * G into traction-key columns (negated) or b, fac = 0.5 for second-order rows
// (BEMBuilderElastic3D.cpp:1593-1842, 1960-2016). Each sphere's rows are written by one workgroup
// in element order: deterministic, no atomics.
__global__ __launch_bounds__(256) void k_rows(const double* __restrict__ qp,
                                              const int* __restrict__ conn,
                                              const double* __restrict__ cen, int NE, long lda,
                                              long row0, Mat M, int rowmask,
                                              const int* __restrict__ node_col,
                                              const int* __restrict__ slot_col,
                                              const double* __restrict__ uval,
                                              const double* __restrict__ tval,
                                              double* __restrict__ A, double* __restrict__ b,
                                              int write_A) {
  constexpr int QN = 64 * kQpStride;             // quadrature data of one element
  constexpr int QPT = (QN + 255) / 256;           // per thread
  __shared__ double sq[2][QN];                    // current / next element
  __shared__ double stab[64 * kKStride];          // tables; then the wavefront partial sums
  __shared__ int scol[2][24];                     // the element's 12 H and 12 G target columns
  __shared__ double sknown[2][24];                // known displacement / traction values
  __shared__ double sb[4][64];
  const int s = blockIdx.x;
  const int t = threadIdx.x, wv = t >> 6, lane = t & 63;
  const Slot R = slot60(lane < 60 ? lane : 59);
  const bool active = lane < 60 && ((rowmask >> R.order) & 1);
  const RowTerms RT = row_terms(M, R);
  const double fac = R.order == 2 ? 0.5 : 1.0;
  const double c0 = cen[3 * s], c1 = cen[3 * s + 1], c2 = cen[3 * s + 2];
  const long row = row0 + 60L * s + lane;
  double bacc = 0.0;
  // element e's column map (H: node dofs 3 q + n; G: slots 3 q + n) into buffer e & 1
  auto load_cols = [&](int e) {
    if (t < 24) {
      const int par = e & 1;
      const int k = t < 12 ? t : t - 12, q = k / 3, nc = k % 3;
      if (t < 12) {
        const int dof = 3 * (conn[4 * e + q] - 1) + nc;
        const int col = node_col[dof];
        scol[par][t] = col;
        sknown[par][t] = col >= 0 ? 0.0 : uval[dof];
      } else {
        const int slot = 12 * e + 3 * q + nc;
        const int kc = slot_col[slot];
        scol[par][t] = kc;
        sknown[par][t] = kc == -1 ? tval[slot] : 0.0;
      }
    }
  };
  for (int u = t; u < QN; u += 256) sq[0][u] = qp[u];
  double qn[QPT];
  for (int e = 0; e < NE; ++e) {
    const int cur = e & 1;
    load_cols(e);
    if (e + 1 < NE) {
#pragma unroll
      for (int k = 0; k < QPT; ++k) {
        const int u = t + 256 * k;
        qn[k] = u < QN ? qp[(size_t)(e + 1) * QN + u] : 0.0;
      }
    }
    __syncthreads();  // sq[cur] complete; the previous element's fold is done with stab
    {
      const double* Q = sq[cur] + lane * kQpStride;
      const double x[3] = {Q[0] - c0, Q[1] - c1, Q[2] - c2};
      double y0[7], y1[7], y2[7], PR[7], PT[5];
      kelvin_point(x, y0, y1, y2, PR, PT);
      double* tab = stab + lane * kKStride;
      switch (wv) {
        case 0: kelvin_group<0>(y0, y1, y2, PR, PT, tab); break;
        case 1: kelvin_group<1>(y0, y1, y2, PR, PT, tab); break;
        case 2: kelvin_group<2>(y0, y1, y2, PR, PT, tab); break;
        default: kelvin_group<3>(y0, y1, y2, PR, PT, tab); break;
      }
    }
    __syncthreads();
    double aH[12], aG[12];  // index n * 4 + q
#pragma unroll
    for (int o = 0; o < 12; ++o) aH[o] = aG[o] = 0.0;
    if (active) {
      for (int gi = 0; gi < 16; ++gi) {
        const int g = 16 * wv + gi;
        const double* Q = sq[cur] + g * kQpStride;
        double su[3], st[3];
        row_gp2(RT, stab + g * kKStride, Q + 3, su, st);
        const double w = Q[10];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const double nw = Q[6 + q] * w;
#pragma unroll
          for (int nn = 0; nn < 3; ++nn) {
            aH[nn * 4 + q] += su[nn] * nw;
            aG[nn * 4 + q] += st[nn] * nw;
          }
        }
      }
    }
    if (e + 1 < NE) {
#pragma unroll
      for (int k = 0; k < QPT; ++k) {
        const int u = t + 256 * k;
        if (u < QN) sq[cur ^ 1][u] = qn[k];
      }
    }
    __syncthreads();
#pragma unroll
    for (int o = 0; o < 12; ++o) {
      stab[(wv * 24 + o) * 64 + lane] = aH[o];
      stab[(wv * 24 + 12 + o) * 64 + lane] = aG[o];
    }
    __syncthreads();
    if (active) {
      // all reads of A first, then the updates (no store/load round trip between outputs)
      double v[6], old[6];
      int col[6];
#pragma unroll
      for (int oo = 0; oo < 6; ++oo) {
        const int o = wv * 6 + oo;
        v[oo] = stab[o * 64 + lane] + stab[(24 + o) * 64 + lane] + stab[(48 + o) * 64 + lane] +
                stab[(72 + o) * 64 + lane];
        const int og = o < 12 ? o : o - 12;
        const int k = 3 * (og % 4) + og / 4;  // 3 q + n
        col[oo] = scol[cur][(o < 12 ? 0 : 12) + k];
        old[oo] = col[oo] >= 0 && write_A ? A[(long)col[oo] * lda + row] : 0.0;
      }
#pragma unroll
      for (int oo = 0; oo < 6; ++oo) {
        const int o = wv * 6 + oo;
        const int og = o < 12 ? o : o - 12;
        const int k = 3 * (og % 4) + og / 4;
        if (o < 12) {
          const double hv = -fac * v[oo];
          if (col[oo] >= 0) {
            if (write_A) A[(long)col[oo] * lda + row] = old[oo] + hv;
          } else {
            bacc -= hv * sknown[cur][k];
          }
        } else {
          const double gv = fac * v[oo];
          if (col[oo] >= 0) {
            if (write_A) A[(long)col[oo] * lda + row] = old[oo] - gv;
          } else {
            bacc += gv * sknown[cur][12 + k];  // zero unless the traction is known
          }
        }
      }
    }
  }
  sb[wv][lane] = active ? bacc : 0.0;
  __syncthreads();
  if (wv == 0 && active) b[row] += sb[0][lane] + sb[1][lane] + sb[2][lane] + sb[3][lane];
}

// ---- K2, moment form (every element planar) ------------------------------------------------------
// Normals within kPlanarTol of each other count as one: replacing them by the first point's changes
// an entry by at most ~kPlanarTol of its magnitude.
constexpr double kPlanarTol = 1e-14;

// Barrier that orders LDS only: global loads issued before it stay in flight.
__device__ __forceinline__ void barrier_lds() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// eknown[e][k]: the known displacement (k < 12, node dof 3 q + n) or traction (k >= 12, slot
// 3 q + n) that column k of element e folds into b; zero where the value is an unknown.
__global__ __launch_bounds__(256) void k_elem_known(const int* __restrict__ conn, int NE,
                                                    const int* __restrict__ node_col,
                                                    const int* __restrict__ slot_col,
                                                    const double* __restrict__ uval,
                                                    const double* __restrict__ tval,
                                                    double* __restrict__ eknown) {
  const long i = (long)blockIdx.x * 256 + threadIdx.x;
  if (i >= 24L * NE) return;
  const int e = (int)(i / 24), k = (int)(i % 24);
  const int kk = k < 12 ? k : k - 12, q = kk / 3, nc = kk % 3;
  double v;
  if (k < 12) {
    const int dof = 3 * (conn[4 * e + q] - 1) + nc;
    v = node_col[dof] >= 0 ? 0.0 : uval[dof];
  } else {
    const int slot = 12 * e + 3 * q + nc;
    v = slot_col[slot] == -1 ? tval[slot] : 0.0;
  }
  eknown[i] = v;
}

// On a planar element the unit normal is the same at every Gauss point, so a row's strain kernels
// su_n, st_n are linear in the Kelvin table with coefficients that do not depend on the point, and
//   sum_g w_g N_q(g) su_n(tab_g) = su_n(P_q),   P_q[slot] = sum_g w_g N_q(g) tab_g[slot].
// Per element: (1) the 64 tables as in k_rows; (2) thread = (used slot, half of the Gauss points)
// forms the four moments P_q of its slot over 32 points (w N_q from cw, wave-uniform addresses, so
// scalar loads); (3) lane = row, wavefront = shape function q: one contraction of the moment table
// (the two halves summed as they are read) gives the row's H and G entries of node q, folded
// through the column map as in k_rows. The contraction runs 4 times per element and row instead of
// 64; the Gauss points, weights and tables are the same as k_rows', only the order of the sums
// differs.
// The fold adds into columns that the elements around a node share. A column the previous element
// wrote takes that element's value from LDS (sfwd); every other column was last written two or
// more elements ago, so its A entries are read at the top of the element, before (1), and arrive
// while (1) and (2) run; a column's first element writes without reading. The next element's
// quadrature points and column entries are loaded into registers before (1) as well. Barriers: one
// full barrier after (1), one ordering LDS only after (2).
constexpr int kTabStride = 109;  // per Gauss point: slots 10 .. 118 of the Kelvin table
constexpr int kTabBase = 10;     // table slot of the first stored value
constexpr int kSqN = 64 * 3 + 3; // staged per element: the 64 points, then the element normal
__global__ __launch_bounds__(256, 2) void k_rows_m(const double* __restrict__ qp,
                                                const double* __restrict__ cw,
                                                const int* __restrict__ ecol,
                                                const double* __restrict__ eknown,
                                                const double* __restrict__ cen, int NE, long lda,
                                                long row0, Mat M, int rowmask,
                                                double* __restrict__ A, double* __restrict__ b,
                                                int write_A, double* __restrict__ sink) {
  __shared__ double sq[2][kSqN + 1];                // current / next element
  __shared__ double stab[64 * kTabStride];          // the 64 Gauss points' tables
  __shared__ double smom[2][4][kTabStride];         // moments: [half of the points][q][slot]
  __shared__ double sfwd[24][64];                   // the previous element's new values [column][row]
  __shared__ int scol[2][24];
  __shared__ double sknown[2][24];
  double* sb = stab;  // after the last element: per-wave right-hand-side sums (4 x 64)
  const int s = blockIdx.x;
  const int t = threadIdx.x, lane = t & 63;
  const int wv = __builtin_amdgcn_readfirstlane(t >> 6);
  // step (3): lane = row, wave = shape function
  const Slot R = slot60(lane < 60 ? lane : 59);
  const bool active = lane < 60 && ((rowmask >> R.order) & 1);
  // the row's table positions packed (8 registers across the loop), unpacked in step (3)
  RowTermsP RP = row_terms_pack(row_terms(M, R), R.order);
  const double fac = R.order == 2 ? 0.5 : 1.0;
  const double c0 = cen[3 * s], c1 = cen[3 * s + 1], c2 = cen[3 * s + 2];
  const long row = row0 + 60L * s + lane;
  // step (2): used slot mB, Gauss points 32 half .. 32 half + 31
  const int half = wv >> 1;
  const int mB = (wv & 1) * 64 + lane;
  const bool actB = mB < kKUsed;
  const int slotB = kelvin_used_slot(actB ? mB : 0) - kTabBase;
  // staged quadrature value u of element e: the points' coordinates, then the normal
  const int tq = t < kSqN ? t : kSqN - 1;
  const long qoff = tq < 192 ? (long)(tq / 3) * kQpStride + tq % 3 : 3 + (tq - 192);
  const int tc = t < 24 ? t : 23;
  double bacc = 0.0;
  if (NE <= 0) return;
  if (t < kSqN) sq[0][t] = qp[qoff];
  if (t < 24) {
    scol[0][t] = ecol[t];
    sknown[0][t] = eknown[t];
  }
  __syncthreads();
  const int q = wv;
  double qn = 0.0, kn = 0.0;
  int cn = 0;
  for (int e = 0; e < NE; ++e) {
    const int cur = e & 1;
    if (e + 1 < NE) {  // element e + 1 into registers (stored to LDS in step 2), no branches
      qn = qp[(size_t)(e + 1) * 64 * kQpStride + qoff];
      cn = ecol[24 * (e + 1) + tc];
      kn = eknown[24 * (e + 1) + tc];
    }
    // (0) this element's columns of node q; reads of the columns written two or more elements ago
    // (issued after the prefetch, which step 2 waits for: the counter is in order)
    int col[6];
    bool rd[6];
    double ld[6];
#pragma unroll
    for (int nn = 0; nn < 3; ++nn) {
      col[nn] = scol[cur][3 * q + nn];
      col[3 + nn] = scol[cur][12 + 3 * q + nn];
    }
#pragma unroll
    for (int o = 0; o < 6; ++o) {
      rd[o] = active && write_A && col[o] >= 0 && !(col[o] & (kRowsColFirst | kRowsColFwd));
      ld[o] = A[rd[o] ? (long)(col[o] & kRowsColMask) * lda + row : row0];
    }
    {  // (1) tables
      const double* G = sq[cur] + 3 * lane;
      const double x[3] = {G[0] - c0, G[1] - c1, G[2] - c2};
      double y0[7], y1[7], y2[7], PR[7], PT[5];
      kelvin_point(x, y0, y1, y2, PR, PT);
      double* tab = stab + lane * kTabStride - kTabBase;
      switch (wv) {
        case 0: kelvin_group<0>(y0, y1, y2, PR, PT, tab); break;
        case 1: kelvin_group<1>(y0, y1, y2, PR, PT, tab); break;
        case 2: kelvin_group<2>(y0, y1, y2, PR, PT, tab); break;
        default: kelvin_group<3>(y0, y1, y2, PR, PT, tab); break;
      }
    }
    __syncthreads();  // tables complete; the previous element's forwarded values visible
    double fw[6];
#pragma unroll
    for (int o = 0; o < 6; ++o)
      fw[o] = col[o] >= 0 && (col[o] & kRowsColFwd) ? sfwd[(col[o] >> kRowsColFwdShift) & 31][lane] : 0.0;
    if (actB) {  // (2) moments; weights of four points per step (two 64-byte scalar loads, one wait)
      const double* We = cw + ((size_t)e * 64 + 32 * half) * 4;
      const double* Te = stab + 32 * half * kTabStride + slotB;
      double p0 = 0.0, p1 = 0.0, p2 = 0.0, p3 = 0.0;
      for (int gi = 0; gi < 32; gi += 4) {
        double w[16], v[4];
#pragma unroll
        for (int k = 0; k < 16; ++k) w[k] = We[4 * gi + k];
#pragma unroll
        for (int k = 0; k < 4; ++k) v[k] = Te[(gi + k) * kTabStride];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          p0 += v[k] * w[4 * k];
          p1 += v[k] * w[4 * k + 1];
          p2 += v[k] * w[4 * k + 2];
          p3 += v[k] * w[4 * k + 3];
        }
      }
      smom[half][0][slotB] = p0;
      smom[half][1][slotB] = p1;
      smom[half][2][slotB] = p2;
      smom[half][3][slotB] = p3;
    }
    if (e + 1 < NE) {
      if (t < kSqN) sq[cur ^ 1][t] = qn;
      if (t < 24) {
        scol[cur ^ 1][t] = cn;
        sknown[cur ^ 1][t] = kn;
      }
    }
    barrier_lds();
    {  // (3) contraction and fold: straight-line for every lane (lanes past row 59 and rows of an
       // order not asked for compute row 59's values and store them to the sink)
      const double* nv = sq[cur] + 192;  // the element's normal (equal at every point)
#pragma unroll
      for (int i = 0; i < 7; ++i) asm volatile("" : "+v"(RP.po[i]));  // unpack here, not before the loop
      asm volatile("" : "+v"(RP.kl));
      const RowTerms RT = row_terms_unpack(RP, M);
      double su[3], st[3];
      row_gp2_t(RT, TabSum2{smom[0][q] - kTabBase, smom[1][q] - kTabBase}, nv, su, st);
      double val[6];
#pragma unroll
      for (int nn = 0; nn < 3; ++nn) {
        const double hv = -fac * su[nn], gv = fac * st[nn];
        val[nn] = (rd[nn] ? ld[nn] : fw[nn]) + hv;
        val[3 + nn] = (rd[3 + nn] ? ld[3 + nn] : fw[3 + nn]) - gv;
        const double kh = active && col[nn] < 0 ? sknown[cur][3 * q + nn] : 0.0;
        const double kg = active && col[3 + nn] < 0 ? sknown[cur][12 + 3 * q + nn] : 0.0;
        bacc += gv * kg - hv * kh;  // known values fold into b (as in k_rows)
        sfwd[3 * q + nn][lane] = val[nn];
        sfwd[12 + 3 * q + nn][lane] = val[3 + nn];
      }
#pragma unroll
      for (int o = 0; o < 6; ++o) {
        const bool wr = active && write_A && col[o] >= 0;
        double* dst = wr ? A + (long)(col[o] & kRowsColMask) * lda + row : sink + t;
        *dst = val[o];
      }
    }
  }
  __syncthreads();  // the last element's tables were read before the last contraction
  sb[wv * 64 + lane] = active ? bacc : 0.0;
  __syncthreads();
  if (wv == 0 && active) b[row] += sb[lane] + sb[64 + lane] + sb[128 + lane] + sb[192 + lane];
}

// flag[e] = 1 when the element's 64 unit normals agree with the first point's to kPlanarTol in every
// component (a planar element: its normals differ by rounding only), else 0;
// cw[e][g][q] = w_g jac_g N_q(g), the moment form's weights.
__global__ __launch_bounds__(64) void k_planar(const double* __restrict__ qp, int NE,
                                               int* __restrict__ flag, double* __restrict__ cw) {
  const int e = blockIdx.x;
  const int g = threadIdx.x;
  if (e >= NE) return;
  const double* Q0 = qp + (size_t)e * 64 * kQpStride;
  const double* Q = Q0 + g * kQpStride;
  const bool same = fabs(Q[3] - Q0[3]) <= kPlanarTol && fabs(Q[4] - Q0[4]) <= kPlanarTol &&
                    fabs(Q[5] - Q0[5]) <= kPlanarTol;
  const bool all = __all(same);
  if (g == 0) flag[e] = all ? 1 : 0;
  double* o = cw + ((size_t)e * 64 + g) * 4;
#pragma unroll
  for (int q = 0; q < 4; ++q) o[q] = Q[10] * Q[6 + q];
}

// ---- launchers ----------------------------------------------------------------------------------
void launch_qpoints(const double* nodes, const int* conn, int NE, double* qp, hipStream_t st) {
  hipLaunchKernelGGL(k_qpoints, dim3(NE), dim3(64), 0, st, nodes, conn, NE, qp);
}

// descriptors of the kInclW weighted values of a pair (incl_weighted_desc), computed once per context
__global__ void k_incl_desc(int* __restrict__ desc) {
  const int u = blockIdx.x * 256 + threadIdx.x;
  if (u < kInclW) desc[u] = incl_weighted_desc(u);
}

void launch_incl(const double* cen, const double* rad, int num, long lda, long row0, const Mat& M,
                 int rowmask, const int* wdesc, double* A, hipStream_t st) {
  dim3 grid((num + kInclTile - 1) / kInclTile, num);
  hipLaunchKernelGGL(k_incl, grid, dim3(kInclThreads), 0, st, cen, rad, num, lda, row0, M, rowmask, wdesc, A);
}

void launch_incl_desc(int* desc, hipStream_t st) {
  hipLaunchKernelGGL(k_incl_desc, dim3((kInclW + 255) / 256), dim3(256), 0, st, desc);
}

void launch_field(const double* nodes, int NN, const double* cen, const double* rad, int num, long lda,
                  long row0, const Mat& M, double* A, hipStream_t st) {
  dim3 grid((NN + 63) / 64, num);
  hipLaunchKernelGGL(k_field, grid, dim3(192), 0, st, nodes, NN, cen, rad, lda, row0, M, A);
}

size_t bem_workspace_doubles(int NN, int NE) {
  return ((size_t)std::min(NE, kBemChunk) * 24 + 6) * 3 * (size_t)NN;
}

void launch_bem(const double* qp, const int* conn, const double* nodes, int NN, int NE, long lda,
                double E0, double nu0, const int* node_col, const int* slot_col, const double* uval,
                const double* tval, double* A, double* b, double* ws, hipStream_t st, bool write_A) {
  const long R = 3L * NN;
  double* W = ws;
  double* carry = ws + (size_t)std::min(NE, kBemChunk) * 24 * R;
  for (int e0 = 0; e0 < NE; e0 += kBemChunk) {
    const int ne = std::min(kBemChunk, NE - e0);
    hipLaunchKernelGGL(k_bem_quad, dim3((NN + 63) / 64, (ne + kBemEPB - 1) / kBemEPB), dim3(192), 0, st,
                       qp, conn, nodes, NN, e0, ne, E0, nu0, W);
    hipLaunchKernelGGL(k_bem_fold, dim3((unsigned)((R + 255) / 256), 3), dim3(256), 0, st, conn, NN, e0,
                       ne, lda, node_col, slot_col, uval, tval, W, carry, (int)(e0 == 0),
                       (int)(e0 + ne >= NE), A, (int)write_A);
  }
  hipLaunchKernelGGL(k_bem_finish, dim3((unsigned)((R + 255) / 256)), dim3(256), 0, st, NN, carry, b);
}

void launch_rows(const double* qp, const int* conn, const double* cen, int num, int NE, long lda,
                 long row0, const Mat& M, int rowmask, const int* node_col, const int* slot_col,
                 const double* uval, const double* tval, double* A, double* b, hipStream_t st,
                 bool write_A, const RowsMoment* mf) {
  if (mf) {
    hipLaunchKernelGGL(k_elem_known, dim3((unsigned)((24L * NE + 255) / 256)), dim3(256), 0, st, conn, NE,
                       node_col, slot_col, uval, tval, mf->eknown);
    hipLaunchKernelGGL(k_rows_m, dim3(num), dim3(256), 0, st, qp, mf->cw, mf->ecol, mf->eknown, cen, NE, lda,
                       row0, M, rowmask, A, b, (int)write_A, mf->sink);
  } else
    hipLaunchKernelGGL(k_rows, dim3(num), dim3(256), 0, st, qp, conn, cen, NE, lda, row0, M, rowmask,
                       node_col, slot_col, uval, tval, A, b, (int)write_A);
}

void launch_planar(const double* qp, int NE, int* flag, double* cw, hipStream_t st) {
  if (NE > 0) hipLaunchKernelGGL(k_planar, dim3(NE), dim3(64), 0, st, qp, NE, flag, cw);
}

}  // namespace ibem
