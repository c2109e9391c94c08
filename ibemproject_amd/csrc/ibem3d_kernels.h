// Kernel launchers shared between the C-ABI layer and the kernel translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "ibem3d_math.h"

namespace ibem {

constexpr int kQpStride = 11;  // per Gauss point: g[3], nv[3], N[4], w*jac
constexpr int kInclTile = 4;   // target spheres per k_incl workgroup

// ---- assembly (assembly.hip) ----
// A is column-major with leading dimension lda (>= n): entry (row, col) at A[col * lda + row].
void launch_qpoints(const double* nodes, const int* conn, int NE, double* qp, hipStream_t st);
// wdesc: the kInclW pair-independent descriptors of the weighted tables (launch_incl_desc)
void launch_incl(const double* cen, const double* rad, int num, long lda, long row0, const Mat& M,
                 int rowmask, const int* wdesc, double* A, hipStream_t st);
void launch_incl_desc(int* desc, hipStream_t st);  // kInclW descriptors
void launch_field(const double* nodes, int NN, const double* cen, const double* rad, int num, long lda,
                  long row0, const Mat& M, double* A, hipStream_t st);
// k_bem's workspace: kBemChunk elements x 24 values x 3 NN rows, plus 6 x 3 NN carried sums
constexpr int kBemChunk = 256;
size_t bem_workspace_doubles(int NN, int NE);
void launch_bem(const double* qp, const int* conn, const double* nodes, int NN, int NE, long lda,
                double E0, double nu0, const int* node_col, const int* slot_col, const double* uval,
                const double* tval, double* A, double* b, double* ws, hipStream_t st,
                bool write_A = true);
// The boundary phase as a gather over a whole-mesh workspace (24 NE 3NN doubles of quadrature values
// plus row sums, bem_gather_workspace_doubles): every boundary column's contributions listed in
// element order (cptr: 3NN + 1 offsets into cent; an entry is e * 24 + value index, with bit 31 set
// for the G values that are subtracted), cdof[j] = the node dof of displacement column j (-1 for a
// traction key). Each entry of A's boundary block is written once (no accumulation into A), and b
// gets the known-value terms. write_A = false: b only.
struct BemGather {
  const int* cptr;
  const int* cent;
  const int* cdof;
};
size_t bem_gather_workspace_doubles(int NN, int NE);
void launch_bem_gather(const double* qp, const int* conn, const double* nodes, int NN, int NE, long lda,
                       double E0, double nu0, const int* node_col, const int* slot_col, const double* uval,
                       const double* tval, double* A, double* b, double* ws, const BemGather& G,
                       hipStream_t st, bool write_A = true);
void launch_rows(const double* qp, const int* conn, const double* cen, int num, int NE, long lda,
                 long row0, const Mat& M, int rowmask, const int* node_col, const int* slot_col,
                 const double* uval, const double* tval, double* A, double* b, hipStream_t st,
                 bool write_A = true, const struct RowsMoment* mf = nullptr);
// The moment form of the boundary->inclusion rows (every element planar): cw[e][g][q] =
// w_g jac_g N_q(g); ecol[e][k] = the BC-folded column (bits 0..23) of element e's H (k < 12, node
// dof 3 q + n) or G (k >= 12, slot 3 q + n) entries, -1 / -2 where the value is known, with
// kRowsColFirst set on the first element (in element order) that touches the column, and
// kRowsColFwd with the column's index k' in element e - 1 (bits 24..28) where element e - 1 touches
// it too; eknown: NE x 24 workspace; sink: 256 doubles that absorb the stores of entries not
// written (never read).
struct RowsMoment {
  const double* cw;
  const int* ecol;
  double* eknown;
  double* sink;
};
constexpr int kRowsColFirst = 1 << 30;
constexpr int kRowsColFwd = 1 << 29;
constexpr int kRowsColFwdShift = 24;
constexpr int kRowsColMask = (1 << 24) - 1;
// flag[e] = 1 when element e's 64 quadrature-point normals agree to rounding (planar element).
void launch_planar(const double* qp, int NE, int* flag, double* cw, hipStream_t st);
// (write_A = false: b only, accumulated in the same order as with the matrix entries)

// ---- postprocess (post.hip) ----
// U (3 NN node dofs) and T (12 NE element slots) of the stress kernel from the device solution x
void launch_stress_inputs(const double* x, const int* node_col, const double* uval, const int* slot_col,
                          const double* tval, int NN, int NE, double* U, double* T, hipStream_t st);
// workspace of launch_stress (doubles): per-sphere moments, per-chunk inclusion strains, masks
size_t stress_workspace_doubles(int num, int nump);
void launch_stress(const double* qp, const int* conn, const double* nodes, int NN, int NE,
                   const double* cen, const double* rad, int num, const double* pts, int nump,
                   const Mat& M, const double* U, const double* T, const double* E60, double* ws,
                   double* s, hipStream_t st);

// ---- dense solver (solver.hip) ----
struct Timer;  // optional per-launch timing hook (api.cpp)

// C[M x N] -= A[M x K] * B[K x N], column-major, FP64 MFMA.
void gemm_f64_sub(long M, long N, long K, const double* A, long lda, const double* B, long ldb,
                  double* C, long ldc, hipStream_t st);
// same in FP32 (FP32 MFMA)
void gemm_f32_sub(long M, long N, long K, const float* A, long lda, const float* B, long ldb,
                  float* C, long ldc, hipStream_t st);
// Product types of the split-operand GEMMs: three FP16 MFMA products of hi/lo splits (FP32-level
// accuracy), one FP16 product, one BF16 product, or three BF16 products of hi/lo BF16 splits (16
// significant bits per operand), each of operands whose rows (A) and columns (B) are scaled by
// powers of two that bring their largest magnitude into [2^14, 2^15).
enum { kProdF16x3 = 0, kProdF16 = 1, kProdBF16 = 2, kProdBF16x3 = 3 };
// same in FP32 arithmetic from 16-bit MFMA products (K <= 1024); ws holds
// f16x3_workspace_doubles(max(M, N)) doubles
size_t f16x3_workspace_doubles(long n);
void gemm_f32_sub_f16x3(long M, long N, int K, const float* A, long lda, const float* B, long ldb,
                        float* C, long ldc, double* ws, hipStream_t st, int prod = kProdF16x3);

// Blocked right-looking LU with row pivoting (tournament pivoting inside 32-column leaves), in
// place; ipiv[k] = row swapped with row k (absolute, 0-based, applied in order). The block column
// of L computed at a step keeps the row order of the end of that step: later steps' interchanges
// are not applied to it (nothing in the factorisation reads it again), and the solve applies each
// step's interchanges to the vector instead. d_info receives the first column + 1 whose pivot is
// exactly zero. Returns nonzero on a launch error.
// timing callback: (context, phase name, 0 = begin / 1 = end, stream the phase runs on)
typedef void (*LuTick)(void*, const char*, int, hipStream_t);
struct LuWork {
  hipStream_t st2 = nullptr;             // second stream: look-ahead panels
  hipEvent_t ev_a = nullptr, ev_b = nullptr;
  hipEvent_t ev_r = nullptr, ev_cols = nullptr;  // equilibration: row scales ready / last columns copied
  int* d_ipiv = nullptr;   // n
  double* d_ws = nullptr;  // lu_workspace_doubles(n), zero-filled once at allocation
  int* d_info = nullptr;   // 2 ints: info, and the no-pivot LU leaves' arrival count
  int* d_perm = nullptr;   // lu_moves_ints(n): each step's interchanges as (dst, src) moves + counts
  double* d_inv = nullptr; // lu_inverse_doubles(n): inverses of the diagonal blocks of L and U
  double* d_tmp = nullptr; // n: solve scratch
  double* d_trinv = nullptr; // lu_trsm_workspace: inverses of one block row's diagonal blocks
  double* d_trsm = nullptr;  // lu_trsm_workspace: U12 of one step (kb x n)
};
size_t lu_workspace_doubles(long n);
size_t lu_inverse_doubles(long n);
size_t lu_moves_ints(long n);
// doubles of d_trinv and d_trsm for an n x n factorisation
size_t lu_trinv_doubles();
size_t lu_trsm_doubles(long n);
int lu_factor_f64(double* A, long n, long lda, LuWork& W, hipStream_t st, LuTick tick, void* tctx);
// LU without row interchanges (identity ipiv, so the solves below apply unchanged); W.d_info = column
// + 1 of the first pivot with magnitude below piv_tol
int lu_factor_f64_nopiv(double* A, long n, long lda, LuWork& W, hipStream_t st, double piv_tol);
int lu_factor_f32_nopiv(float* A, long n, long lda, LuWork& W, hipStream_t st, double piv_tol);
// f16ws != nullptr: the trailing updates run on 16-bit MFMA products of type prod (kProd*) and the
// panels' deep updates on split FP16 products (workspace of lu_f16_workspace_doubles(n) doubles);
// otherwise both run on FP32 MFMA
size_t lu_f16_workspace_doubles(long n);
// cols_pending: columns past the first panel are still being copied (ev_cols, see
// equilibrate_copy_f32); the factorisation waits for them after its first panel
int lu_factor_f32(float* A, long n, long lda, LuWork& W, hipStream_t st, LuTick tick, void* tctx, double* f16ws,
                  bool cols_pending = false, int prod = kProdF16x3);
// the LU's second stream and events (created once)
int lu_streams(LuWork& W);
// After a factorisation: each step's interchanges as a move list and the diagonal-block inverses,
// all on the device (no synchronisation). Returns nonzero on a launch error.
int lu_prepare_solve_f64(const double* LU, long n, long lda, LuWork& W, hipStream_t st);
int lu_prepare_solve_f32(const float* LU, long n, long lda, LuWork& W, hipStream_t st);
// Solve with the factors: x (in/out, n) <- A^{-1} x
void lu_solve_f64(const double* LU, long n, long lda, const LuWork& W, double* x, hipStream_t st);
void lu_solve_f32(const float* LU, long n, long lda, const LuWork& W, float* x, hipStream_t st);

// row / column equilibration: r_i = 1 / max_j |A_ij|, c_j = 1 / max_i |r_i A_ij|; A <- R A C
void equilibrate_f64(double* A, long n, long lda, double* r, double* c, hipStream_t st);
// y = alpha A x + beta y (FP64, column-major); ws holds gemv_workspace_doubles(m) doubles; the
// columns are split into chunks (at most 64) whose partial sums are added in a fixed order
size_t gemv_workspace_doubles(long m);
void gemv_f64(long m, long n, double alpha, const double* A, long lda, const double* x, double beta,
              double* y, double* ws, hipStream_t st, int chunks = 64);
// 16-bit copy of the n x n FP64 A (per 256 x 32 tile power-of-two scaled FP16, leading dimension
// m16_ld(n), m16_scale_count(n) float scales) and y = alpha A16 x + beta y with it (FP32 sums within a
// tile, FP64 across; ws as for gemv_f64)
long m16_ld(long n);
size_t m16_scale_count(long n);
void m16_build(const double* A, long n, long lda, unsigned short* A16, long ld16, float* scale, hipStream_t st);
void gemv_m16(long n, double alpha, const unsigned short* A16, long ld16, const float* scale, const double* x,
              double beta, double* y, double* ws, hipStream_t st);
// The same product in parts: the partial sums of column tiles [t0, t1) (32 columns each, m16_tiles(n)
// in all) into chunks [chunk0, chunk0 + chunks) of ws (at most 16); once chunks [0, total) hold
// partial sums, gemv_m16_finish adds them in a fixed order, y = alpha sum + beta y.
long m16_tiles(long n);
void gemv_m16_part(long n, const unsigned short* A16, long ld16, const float* scale, const double* x, double* ws,
                   long t0, long t1, int chunk0, int chunks, hipStream_t st);
void gemv_m16_finish(long n, double alpha, const double* ws, double beta, double* y, hipStream_t st, int total);
// out = R A C with the same scaling, A left unchanged (FP64 or FP32 output); ws holds
// equilibrate_workspace_doubles(n) doubles. With W, the FP32 copy writes its first LU panel's columns
// on st and the rest on W's second stream, recording W.ev_cols: the next lu_factor_f32 on W must
// then be called with cols_pending. Returns nonzero on a HIP error.
size_t equilibrate_workspace_doubles(long n);
// perm (optional): output column j is A's column perm[j] (c then belongs to the output's columns)
void equilibrate_copy_f64(const double* A, long n, long lda, double* r, double* c, double* out, long ldo,
                          double* ws, hipStream_t st, const int* perm = nullptr);
int equilibrate_copy_f32(const double* A, long n, long lda, double* r, double* c, float* out, long ldo,
                         double* ws, hipStream_t st, LuWork* W = nullptr, const int* perm = nullptr);
// Compact storage (the factor takes the place of the FP64 system matrix): A <- R A C in place, or
// FP64 A (leading dimension n) -> FP32 R A C in the same memory with leading dimension lda32
// (n <= lda32 < 2n), the first columns staged through tmp (at least two columns of lda32 floats).
// Return nonzero on a HIP error or too small a tmp.
void equilibrate_in_place_f64(double* A, long n, long lda, double* r, double* c, double* ws, hipStream_t st);
int equilibrate_to_f32_in_place(double* A, long n, long lda, double* r, double* c, long lda32, float* tmp,
                                size_t tmp_floats, double* ws, hipStream_t st);
// FP16 copy of the FP32 factor: each 128 x 128 tile scaled by a power of two (largest magnitude in
// [2^14, 2^15)), column-major with leading dimension ldf; isc holds factor16_scales(n) inverse
// scales. lu_solve_16 solves with it (FP32 vectors), using the inverses of the diagonal blocks that
// lu_prepare_solve_f32 computed from the FP32 factor.
size_t factor16_scales(long n);
void factor_to_16(const float* LU, long n, long lda, unsigned short* F, long ldf, float* isc, hipStream_t st);
void lu_solve_16(const unsigned short* F, long n, long ldf, const float* isc, const LuWork& W, float* x,
                 hipStream_t st);
// Runtime tunables of the LU (process-wide): the 256 x 256 tile count from which trailing updates
// take the persistent kernel, and the percentage of the panel CU schedule the bulk update leaves
// to the concurrent panel. set_tuning returns nonzero for an unknown key or a value out of range.
// kTuneInnerIters / kTuneInnerTolExp: the inner GMRES of IBEM3D_SOLVE_BJ_GMRES16 (iterations per outer
// direction, at most kInnerItersMax; stopping tolerance 10^-exp).
// kTuneRowsPointwise: 1 makes the boundary->inclusion rows use the per-Gauss-point contraction even
// where every element is planar (default 0: the moment form there).
// kTuneBjNopiv: the block-Jacobi boundary block's factorisation: 1 (default) without interchanges
// in its natural column order where the column map allows, 0 always with partial pivoting, 2 with a
// pivot threshold no pivot meets (the partial-pivoting fallback runs every time).
// kTuneBjOverlap: 1 (default) runs the boundary block's triangular solves of each inner iteration on
// the second stream beside the FP16 product's sphere columns; 0 runs them before the product.
// kTuneStressBeside: 1 (default) computes the stresses of each GMRES cycle's new x on the second
// stream beside its residual check, kept when that residual ends the solve; 0 only on request.
// kTuneBjFp32: 1 (default) factors the block-Jacobi boundary block in FP32 when it needs no row
// interchanges (the partial-pivoting fallback stays FP64); 0 in FP64.
// kTuneResidualUpdate: 1 (default) checks each GMRES cycle's new x with r - (A Z) y, from the FP64
// products of the cycle (b - A x by a product only when that misses the tolerance); 0 always by a
// product.
enum { kTuneBigMinTiles = 1, kTunePanelCuPercent = 2, kTuneInnerIters = 3, kTuneInnerTolExp = 4,
       kTuneRowsPointwise = 5, kTuneBjNopiv = 6, kTuneBjOverlap = 7, kTuneStressBeside = 8,
       kTuneBjFp32 = 9, kTuneResidualUpdate = 10 };
constexpr int kInnerItersMax = 16;
int set_tuning(int key, long value);
long get_tuning(int key);
// Zero bytes (a multiple of 4; p 16-byte aligned) with a kernel instead of a memset (capturable)
void fill_zero(void* p, size_t bytes, hipStream_t st);
// FP64 -> FP32 copy of a column-major matrix
void convert_f64_to_f32(const double* A, long n, long lda, float* B, long ldb, hipStream_t st);

}  // namespace ibem
