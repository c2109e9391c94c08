// C-ABI of the MI355X iBEM elastic-3D engine (include/ibem3d.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "../../include/ibem3d.h"
#include "batch_internal.h"
#include "ibem3d_kernels.h"

using namespace ibem;

namespace {

std::string g_last_error;

struct TimeEntry {
  std::string name;
  hipEvent_t a, b;
};

}  // namespace

struct ibem3d_ctx {
  int dev = 0;
  hipStream_t st = nullptr;
  int NN = 0, NE = 0, num = 0, nump = 0;
  long n = 0;
  double E0 = 0, nu0 = 0, E1 = 0, nu1 = 0;
  Mat M;
  std::vector<double> h_bc, h_tbc;
  std::vector<int> node_col, slot_col;
  // device inputs
  double *d_nodes = nullptr, *d_cen = nullptr, *d_rad = nullptr, *d_pts = nullptr;
  double *d_qp = nullptr, *d_uval = nullptr, *d_tval = nullptr;
  int *d_conn = nullptr, *d_node_col = nullptr, *d_slot_col = nullptr;
  // system
  double *d_A = nullptr, *d_b = nullptr, *d_x = nullptr;
  long lda = 0;  // A's leading dimension: n rounded up to 64 doubles, so every column starts on a
                 // 512-byte boundary (unaligned columns cost the products ~10 % more line fetches)
  // solver
  double *d_r = nullptr, *d_c = nullptr, *d_work = nullptr, *d_res = nullptr;
  double* d_Af = nullptr;  // FP64 factor copy
  float* d_A32 = nullptr;
  long lda32 = 0;  // leading dimension of d_A32: a multiple of 64 for aligned 16-byte accesses
  double* d_f16ws = nullptr;  // split operands of the FP16-MFMA trailing update
  double* d_eqws = nullptr;   // equilibration partial row maxima
  float* d_v32 = nullptr;
  // GMRES-IR: Krylov basis V (kGmres + 1 columns), preconditioned directions Z (kGmres columns),
  // dot-product partial sums, small coefficient vector
  double *d_gV = nullptr, *d_gZ = nullptr, *d_gpart = nullptr, *d_gh = nullptr;
  double* d_gW = nullptr;  // the GMRES cycle's FP64 products A z_j, before orthogonalisation
  double* d_bemws = nullptr;  // boundary-boundary quadrature values (whole mesh, or one round of elements)
  // the boundary phase as a gather (launch_bem_gather) when its whole-mesh workspace fits: every
  // boundary column's contributions in element order (CSR), each column's node dof
  int *d_bg_ptr = nullptr, *d_bg_ent = nullptr, *d_bg_dof = nullptr;
  bool bem_gather = false;
  double* d_gemvws = nullptr; // residual GEMV partial sums
  int prepared_mode = -1;      // mode (with IBEM3D_SOLVE_COMPACT) whose factors are in place (-1: none)
  // compact storage: 16-bit copy of the FP32 factor (leading dimension ldf) and its tile scales
  unsigned short* d_F16 = nullptr;
  float* d_F16s = nullptr;
  long ldf = 0;
  bool a_factored = false;     // d_A holds a factor in place of the assembled system
  // block-Jacobi preconditioner (IBEM3D_SOLVE_BJ_GMRES): the equilibrated boundary block R A_bb C
  // (nb x nb, nb = 3 NN) with its LU workspace, and the inverse of every sphere's 60 x 60 diagonal block
  double *d_Abb = nullptr, *d_rb = nullptr, *d_cb = nullptr, *d_Binv = nullptr, *d_tb = nullptr;
  double* d_tb32 = nullptr;  // an FP32 boundary factor's right-hand side (nb floats)
  bool bj_f32 = false;       // the boundary factor in d_Abb is FP32 (leading dimension nb)
  // the boundary block's natural column order (column of row k = node dof k's own unknown), when
  // the column map makes it a permutation; the block is then factored without interchanges
  int* d_bperm = nullptr;
  std::vector<int> h_bperm;
  bool bj_nopiv = false;       // the current boundary factor is of the permuted block
  LuWork Wb;
  // after a block-Jacobi solve the next assembly factors the boundary block and inverts the sphere
  // blocks on a second stream, beside the assembly kernels that do not write them (st_aux, events
  // in / out); pending: launched and not yet joined by the main stream; ready: the preconditioner
  // matches the assembled A
  hipStream_t st_aux = nullptr;
  hipEvent_t ev_aux_in = nullptr, ev_aux_in2 = nullptr, ev_aux_out = nullptr;
  hipEvent_t ev_ov_in = nullptr, ev_ov_out = nullptr;  // inner iterations: boundary solve beside the product
  bool bj_pending = false, bj_ready = false;
  // IBEM3D_SOLVE_BJ_GMRES16: the 16-bit copy of A (leading dimension ld16) with its tile scales, and
  // the inner GMRES's basis, directions and product vector
  unsigned short* d_A16 = nullptr;
  float* d_A16s = nullptr;
  long ld16 = 0;
  double *d_iV = nullptr, *d_iZ = nullptr, *d_iw = nullptr;
  long inner_iters = 0;        // inner iterations of the last solve
  bool m16_skipped = false;    // the last BJ_GMRES16 solve ran as BJ_GMRES (no room for the copy)
  // initial guess of the eigenstrain unknowns (60 num) for the block-Jacobi GMRES, if set
  double* d_x0s = nullptr;
  bool has_x0 = false;
  bool bj_fallback = false;    // the last block-Jacobi solve fell back to the mixed LU
  double warm_rel = -1.0;      // relative residual of the last warm start (-1: none)
  double warm_alpha = 0.0;     // the scale the last warm start applied to the guess
  LuWork W;
  unsigned phases = 0;
  bool solved = false;
  // every boundary element planar: the moment form of k_rows (weights, element columns, known values)
  double* d_cw = nullptr;
  int* d_ecol = nullptr;
  int* d_wdesc = nullptr;      // k_incl's weighted-table descriptors (kInclW)
  double* d_eknown = nullptr;
  double* d_sink = nullptr;
  RowsMoment rows_mf{};
  double rel_res = -1.0;
  int iters = 0;
  std::string err;
  std::vector<TimeEntry> times;
  std::map<std::string, hipEvent_t> open;
  std::vector<hipEvent_t> pool;             // recorded-and-read events, ready for reuse
  std::vector<std::string> folded_order;    // per-name totals of intervals already folded
  std::map<std::string, double> folded_ms;
  long timing_errors = 0;
  // stress inputs and output (kept between calls)
  double *d_stU = nullptr, *d_stT = nullptr, *d_sts = nullptr, *d_stws = nullptr;
  // stresses computed on the second stream beside a solve's final residual check (ev_st_out marks
  // their end): stress_pending while the main stream has not waited for them, stress_spec while they
  // belong to the current solution
  bool stress_pending = false, stress_spec = false;
  hipEvent_t ev_st_in = nullptr, ev_st_out = nullptr;
  // host copies for the volume averages
  std::vector<double> h_qp, h_rad;
  std::vector<int> h_conn;
};

namespace {

int fail(ibem3d_ctx* c, int code, const std::string& msg) {
  if (c) c->err = msg;
  g_last_error = msg;
  return code;
}

#define HIPC(call)                                                                          \
  do {                                                                                      \
    hipError_t _e = (call);                                                                 \
    if (_e != hipSuccess)                                                                   \
      return fail(ctx, IBEM3D_EHIP, std::string(#call) + ": " + hipGetErrorString(_e));    \
  } while (0)

// Phase timing with HIP events on the stream each phase runs on. Events come from a per-context
// pool; once many intervals are pending, the finished ones are folded into per-name totals and
// their events returned to the pool, so a long-lived context holds a bounded number of events.
constexpr size_t kFoldAt = 4096;

hipEvent_t take_event(ibem3d_ctx* c);
void fold_times(ibem3d_ctx* c);

void tick(void* p, const char* name, int end, hipStream_t s) {
  ibem3d_ctx* c = (ibem3d_ctx*)p;
  if (!s) s = c->st;
  if (!end) {
    hipEvent_t a = take_event(c);
    if (!a || hipEventRecord(a, s) != hipSuccess) {
      c->timing_errors++;
      return;
    }
    c->open[name] = a;
  } else {
    auto it = c->open.find(name);
    if (it == c->open.end()) return;
    hipEvent_t b = take_event(c);
    if (!b || hipEventRecord(b, s) != hipSuccess) {
      c->timing_errors++;
      c->pool.push_back(it->second);
      c->open.erase(it);
      return;
    }
    c->times.push_back({name, it->second, b});
    c->open.erase(it);
    if (c->times.size() >= kFoldAt) fold_times(c);
  }
}

hipEvent_t take_event(ibem3d_ctx* c) {
  if (!c->pool.empty()) {
    hipEvent_t e = c->pool.back();
    c->pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

// add the finished intervals to the per-name totals and return their events to the pool
void fold_times(ibem3d_ctx* c) {
  hipStreamSynchronize(c->st);
  if (c->W.st2) hipStreamSynchronize(c->W.st2);
  for (const auto& t : c->times) {
    float v = 0;
    if (hipEventElapsedTime(&v, t.a, t.b) != hipSuccess) c->timing_errors++;
    if (!c->folded_ms.count(t.name)) c->folded_order.push_back(t.name);
    c->folded_ms[t.name] += v;
    c->pool.push_back(t.a);
    c->pool.push_back(t.b);
  }
  c->times.clear();
}

struct Scope {
  ibem3d_ctx* c;
  const char* n;
  hipStream_t s;
  Scope(ibem3d_ctx* c_, const char* n_, hipStream_t s_ = nullptr) : c(c_), n(n_), s(s_) { tick(c, n, 0, s); }
  ~Scope() { tick(c, n, 1, s); }
};

template <typename T>
int dalloc(ibem3d_ctx* ctx, T** p, size_t count) {
  if (count == 0) count = 1;
  HIPC(hipMalloc((void**)p, count * sizeof(T)));
  return 0;
}

template <typename T>
int upload(ibem3d_ctx* ctx, T** p, const T* h, size_t count) {
  if (int rc = dalloc(ctx, p, count)) return rc;
  if (h && count) HIPC(hipMemcpyAsync(*p, h, count * sizeof(T), hipMemcpyHostToDevice, ctx->st));
  return 0;
}

// Column map of the BC fold (BEMBuilderElastic3D::ApplyBC, BEMBuilderElastic3D.cpp:1960-2016):
// displacement-unknown dofs in dof order, then traction-unknown (node, dof) keys in first-encounter
// order over (element, local dof); later slots with the same key share the column (Find :1894).
int build_column_map(ibem3d_ctx* ctx, const ibem3d_problem* p) {
  const int NN = p->NN, NE = p->NE;
  ctx->node_col.assign(3 * NN, -1);
  ctx->slot_col.assign(12 * NE, -1);
  int t = 0;
  for (int j = 0; j < 3 * NN; ++j) {
    const double f = p->bc[2 * j];
    if (f == 1.0) ctx->node_col[j] = t++;
  }
  std::vector<int> key(3 * NN, -1);
  for (int e = 0; e < NE; ++e)
    for (int i = 0; i < 12; ++i) {
      const double f = p->tbc[24 * e + i];
      const int nd = p->conn[4 * e + i / 3] - 1;
      if (nd < 0 || nd >= NN) return fail(ctx, IBEM3D_EINVAL, "element node id out of range");
      const int k = 3 * nd + i % 3;
      if (f == 1.0) {
        ctx->slot_col[12 * e + i] = -1;
      } else if (f == 0.0) {
        if (key[k] < 0) key[k] = t++;
        ctx->slot_col[12 * e + i] = key[k];
      } else {
        ctx->slot_col[12 * e + i] = -2;  // neither branch of the reference: ignored
      }
    }
  if (t != 3 * NN)
    return fail(ctx, IBEM3D_EINVAL,
                "boundary conditions: displacement unknowns + traction keys = " + std::to_string(t) +
                    " != 3 NN = " + std::to_string(3 * NN));
  // the natural column of row k (node dof k): its displacement unknown, else its traction key; kept
  // when that is a permutation of the boundary columns
  ctx->h_bperm.assign(3 * NN, -1);
  std::vector<char> hit(3 * NN, 0);
  bool perm_ok = true;
  for (int k = 0; k < 3 * NN && perm_ok; ++k) {
    const int c = ctx->node_col[k] >= 0 ? ctx->node_col[k] : key[k];
    if (c < 0 || hit[c]) perm_ok = false;
    else hit[c] = 1, ctx->h_bperm[k] = c;
  }
  if (!perm_ok) ctx->h_bperm.clear();
  return 0;
}

bool finite_all(const double* p, size_t n) {
  for (size_t i = 0; i < n; ++i)
    if (!std::isfinite(p[i])) return false;
  return true;
}

double dnrm2_host(const std::vector<double>& v) {
  double s = 0.0;
  for (double x : v) s += x * x;
  return std::sqrt(s);
}

}  // namespace

extern "C" {

const char* ibem3d_last_error(const ibem3d_ctx* ctx) {
  return ctx ? ctx->err.c_str() : g_last_error.c_str();
}

long ibem3d_size(const ibem3d_ctx* ctx) { return ctx ? ctx->n : 0; }

int ibem3d_set_tuning(int key, long value) {
  if (set_tuning(key, value)) {
    g_last_error = "set_tuning: unknown key or value out of range";
    return IBEM3D_EINVAL;
  }
  return 0;
}
long ibem3d_get_tuning(int key) { return get_tuning(key); }

static bool fits(size_t bytes);
static void zero_boundary_columns(ibem3d_ctx* ctx, hipStream_t st);

ibem3d_ctx* ibem3d_create(const ibem3d_problem* p, int device) {
  if (!p || p->NN <= 0 || p->NE <= 0 || p->num < 0 || p->nump < 0 || !p->nodes || !p->conn ||
      !p->bc || !p->tbc || (p->num > 0 && (!p->centres || !p->radii)) ||
      (p->nump > 0 && !p->points)) {
    g_last_error = "ibem3d_create: invalid problem description";
    return nullptr;
  }
  if (!finite_all(p->nodes, 3 * (size_t)p->NN) || !finite_all(p->bc, 6 * (size_t)p->NN) ||
      !finite_all(p->tbc, 24 * (size_t)p->NE) ||
      (p->num && (!finite_all(p->centres, 3 * (size_t)p->num) || !finite_all(p->radii, p->num))) ||
      (p->nump && !finite_all(p->points, 3 * (size_t)p->nump)) || !std::isfinite(p->E0) ||
      !std::isfinite(p->nu0) || !std::isfinite(p->E1) || !std::isfinite(p->nu1)) {
    g_last_error = "ibem3d_create: non-finite input";
    return nullptr;
  }
  for (int h = 0; h < p->num; ++h)
    if (!(p->radii[h] > 0)) {
      g_last_error = "ibem3d_create: sphere radius must be positive";
      return nullptr;
    }
  ibem3d_ctx* ctx = new ibem3d_ctx();
  auto bail = [&](const std::string& m) -> ibem3d_ctx* {
    g_last_error = m.empty() ? ctx->err : m;
    ibem3d_destroy(ctx);
    return nullptr;
  };
  ctx->dev = device;
  if (hipSetDevice(device) != hipSuccess) return bail("ibem3d_create: hipSetDevice failed");
  if (hipStreamCreateWithFlags(&ctx->st, hipStreamNonBlocking) != hipSuccess)
    return bail("ibem3d_create: stream creation failed");
  ctx->NN = p->NN;
  ctx->NE = p->NE;
  ctx->num = p->num;
  ctx->nump = p->nump;
  ctx->n = 3L * p->NN + 60L * p->num;
  ctx->E0 = p->E0;
  ctx->nu0 = p->nu0;
  ctx->E1 = p->E1;
  ctx->nu1 = p->nu1;
  ctx->M = make_mat(p->E0, p->nu0, p->E1, p->nu1);
  ctx->h_bc.assign(p->bc, p->bc + 6 * (size_t)p->NN);
  ctx->h_tbc.assign(p->tbc, p->tbc + 24 * (size_t)p->NE);
  if (build_column_map(ctx, p)) return bail("");
  std::vector<double> uval(3 * p->NN), tval(12 * (size_t)p->NE);
  for (int j = 0; j < 3 * p->NN; ++j) uval[j] = p->bc[2 * j + 1];
  for (int e = 0; e < p->NE; ++e)
    for (int i = 0; i < 12; ++i) tval[12 * e + i] = p->tbc[24 * e + 12 + i];
  const long n = ctx->n;
  ctx->lda = (n + 63) / 64 * 64;
  if (upload(ctx, &ctx->d_nodes, p->nodes, 3 * (size_t)p->NN) ||
      upload(ctx, &ctx->d_conn, p->conn, 4 * (size_t)p->NE) ||
      upload(ctx, &ctx->d_cen, p->centres, 3 * (size_t)p->num) ||
      upload(ctx, &ctx->d_rad, p->radii, (size_t)p->num) ||
      upload(ctx, &ctx->d_pts, p->points, 3 * (size_t)p->nump) ||
      upload(ctx, &ctx->d_uval, uval.data(), uval.size()) ||
      upload(ctx, &ctx->d_tval, tval.data(), tval.size()) ||
      upload(ctx, &ctx->d_node_col, ctx->node_col.data(), ctx->node_col.size()) ||
      upload(ctx, &ctx->d_slot_col, ctx->slot_col.data(), ctx->slot_col.size()) ||
      (!ctx->h_bperm.empty() && upload(ctx, &ctx->d_bperm, ctx->h_bperm.data(), ctx->h_bperm.size())) ||
      dalloc(ctx, &ctx->d_qp, (size_t)p->NE * 64 * kQpStride) ||
      dalloc(ctx, &ctx->d_A, (size_t)ctx->lda * n) || dalloc(ctx, &ctx->d_b, (size_t)n) ||
      dalloc(ctx, &ctx->d_x, (size_t)n))
    return bail("");
  fill_zero(ctx->d_b, sizeof(double) * n, ctx->st);  // b is accumulated by the assembly kernels
  launch_qpoints(ctx->d_nodes, ctx->d_conn, p->NE, ctx->d_qp, ctx->st);
  {
    // boundary columns' contributions in element order: H of displacement-unknown node dofs (added),
    // G of traction-unknown slots (subtracted), as ApplyBC folds them (BEMBuilderElastic3D.cpp:
    // 1960-2016); the gather form needs the whole mesh's quadrature values at once
    const long R = 3L * p->NN;
    std::vector<int> cnt(R + 1, 0), dof(R, -1);
    for (int e = 0; e < p->NE; ++e)
      for (int k = 0; k < 12; ++k) {
        const int d = 3 * (p->conn[4 * e + k / 3] - 1) + k % 3;
        if (ctx->node_col[d] >= 0) ++cnt[ctx->node_col[d] + 1];
        if (ctx->slot_col[12 * e + k] >= 0) ++cnt[ctx->slot_col[12 * e + k] + 1];
      }
    for (long j = 0; j < R; ++j) cnt[j + 1] += cnt[j];
    std::vector<int> ent(cnt[R]), pos(cnt.begin(), cnt.end() - 1);
    for (int e = 0; e < p->NE; ++e)
      for (int k = 0; k < 12; ++k) {
        const int d = 3 * (p->conn[4 * e + k / 3] - 1) + k % 3;
        if (ctx->node_col[d] >= 0) ent[pos[ctx->node_col[d]]++] = 24 * e + 12 + k;
        if (ctx->slot_col[12 * e + k] >= 0)
          ent[pos[ctx->slot_col[12 * e + k]]++] = (int)((24u * e + k) | 0x80000000u);
      }
    for (long d = 0; d < R; ++d)
      if (ctx->node_col[d] >= 0) dof[ctx->node_col[d]] = (int)d;
    ctx->bem_gather = 24L * p->NE < (1L << 31) &&
                      fits(sizeof(double) * bem_gather_workspace_doubles(p->NN, p->NE));
    if (ctx->bem_gather &&
        (upload(ctx, &ctx->d_bg_ptr, cnt.data(), cnt.size()) || upload(ctx, &ctx->d_bg_ent, ent.data(), ent.size()) ||
         upload(ctx, &ctx->d_bg_dof, dof.data(), dof.size()) ||
         dalloc(ctx, &ctx->d_bemws, bem_gather_workspace_doubles(p->NN, p->NE))))
      return bail("");
  }
  if (dalloc(ctx, &ctx->d_wdesc, kInclW)) return bail("");
  launch_incl_desc(ctx->d_wdesc, ctx->st);
  {
    // planar elements (one normal at every quadrature point) let k_rows integrate the Kelvin tables
    // before the row contraction
    int* d_flag = nullptr;
    std::vector<int> flag(p->NE, 0);
    if (hipMalloc((void**)&d_flag, sizeof(int) * flag.size()) != hipSuccess)
      return bail("ibem3d_create: hipMalloc (element flags)");
    if (dalloc(ctx, &ctx->d_cw, (size_t)p->NE * 64 * 4)) {
      hipFree(d_flag);
      return bail("");
    }
    launch_planar(ctx->d_qp, p->NE, d_flag, ctx->d_cw, ctx->st);
    const bool ok = hipGetLastError() == hipSuccess &&
                    hipMemcpyAsync(flag.data(), d_flag, sizeof(int) * p->NE, hipMemcpyDeviceToHost,
                                   ctx->st) == hipSuccess &&
                    hipStreamSynchronize(ctx->st) == hipSuccess;
    hipFree(d_flag);
    if (!ok) return bail("ibem3d_create: quadrature setup failed");
    if (!std::all_of(flag.begin(), flag.end(), [](int f) { return f == 1; }) ||
        3L * p->NN > kRowsColMask) {
      hipFree(ctx->d_cw);
      ctx->d_cw = nullptr;
    } else {
      // element columns in element order: the first toucher of every column marked, and the
      // columns the previous element touched too marked with their index there
      std::vector<int> ecol(24 * (size_t)p->NE);
      std::vector<char> seen(3 * (size_t)p->NN, 0);
      auto raw = [&](int e, int k) {
        const int kk = k < 12 ? k : k - 12, q = kk / 3, nc = kk % 3;
        return k < 12 ? ctx->node_col[3 * (p->conn[4 * e + q] - 1) + nc] : ctx->slot_col[12 * e + 3 * q + nc];
      };
      for (int e = 0; e < p->NE; ++e)
        for (int k = 0; k < 24; ++k) {
          const int c = raw(e, k);
          int code = c;
          if (c >= 0) {
            if (!seen[c]) {
              code |= kRowsColFirst;
            } else if (e > 0) {
              for (int k2 = 0; k2 < 24; ++k2)
                if (raw(e - 1, k2) == c) {
                  code |= kRowsColFwd | (k2 << kRowsColFwdShift);
                  break;
                }
            }
            seen[c] = 1;
          }
          ecol[24 * (size_t)e + k] = code;
        }
      if (upload(ctx, &ctx->d_ecol, ecol.data(), ecol.size()) ||
          dalloc(ctx, &ctx->d_eknown, 24 * (size_t)p->NE) || dalloc(ctx, &ctx->d_sink, 256))
        return bail("");
      ctx->rows_mf = RowsMoment{ctx->d_cw, ctx->d_ecol, ctx->d_eknown, ctx->d_sink};
    }
  }
  zero_boundary_columns(ctx, ctx->st);
  return ctx;
}

void ibem3d_destroy(ibem3d_ctx* ctx) {
  if (!ctx) return;
  hipSetDevice(ctx->dev);
  if (ctx->st) hipStreamSynchronize(ctx->st);
  if (ctx->st_aux) hipStreamSynchronize(ctx->st_aux);
  ibem3d_reset_timings(ctx);
  for (hipEvent_t e : ctx->pool) hipEventDestroy(e);
  ctx->pool.clear();
  void* ptrs[] = {ctx->d_stU,   ctx->d_stT,      ctx->d_sts,      ctx->d_stws,     ctx->d_cw,   ctx->d_ecol, ctx->d_eknown,
                  ctx->d_sink,  ctx->d_wdesc,    ctx->d_bperm,
                  ctx->d_nodes, ctx->d_cen,      ctx->d_rad,      ctx->d_pts,  ctx->d_qp,
                  ctx->d_uval,  ctx->d_tval,     ctx->d_conn,     ctx->d_node_col, ctx->d_slot_col,
                  ctx->d_A,     ctx->d_b,        ctx->d_x,        ctx->d_r,    ctx->d_c,
                  ctx->d_work,  ctx->d_res,      ctx->d_Af,       ctx->d_A32,  ctx->d_v32, ctx->d_f16ws, ctx->d_eqws,
                  ctx->d_F16,   ctx->d_F16s,     ctx->d_Abb,      ctx->d_rb,   ctx->d_cb,
                  ctx->d_Binv,  ctx->d_tb,       ctx->d_tb32,     ctx->d_x0s,      ctx->Wb.d_ipiv, ctx->Wb.d_ws,
                  ctx->d_A16,   ctx->d_A16s,     ctx->d_iV,       ctx->d_iZ,   ctx->d_iw,
                  ctx->Wb.d_info, ctx->Wb.d_perm, ctx->Wb.d_inv,  ctx->Wb.d_tmp, ctx->Wb.d_trinv,
                  ctx->Wb.d_trsm, ctx->d_bg_ptr, ctx->d_bg_ent, ctx->d_bg_dof,
                  ctx->d_gV,    ctx->d_gZ,       ctx->d_gW,       ctx->d_gpart, ctx->d_gh,  ctx->d_bemws,
                  ctx->d_gemvws,
                  ctx->W.d_ipiv, ctx->W.d_ws,  ctx->W.d_info, ctx->W.d_perm, ctx->W.d_inv, ctx->W.d_tmp,
                  ctx->W.d_trinv, ctx->W.d_trsm};
  for (void* q : ptrs)
    if (q) hipFree(q);
  if (ctx->st_aux) hipStreamDestroy(ctx->st_aux);
  for (hipEvent_t e : {ctx->ev_aux_in, ctx->ev_aux_in2, ctx->ev_aux_out, ctx->ev_ov_in, ctx->ev_ov_out,
                       ctx->ev_st_in, ctx->ev_st_out})
    if (e) hipEventDestroy(e);
  if (ctx->W.st2) hipStreamDestroy(ctx->W.st2);
  if (ctx->Wb.st2) hipStreamDestroy(ctx->Wb.st2);
  for (hipEvent_t e : {ctx->Wb.ev_a, ctx->Wb.ev_b, ctx->Wb.ev_r, ctx->Wb.ev_cols, ctx->W.ev_r, ctx->W.ev_cols})
    if (e) hipEventDestroy(e);
  if (ctx->W.ev_a) hipEventDestroy(ctx->W.ev_a);
  if (ctx->W.ev_b) hipEventDestroy(ctx->W.ev_b);
  if (ctx->st) hipStreamDestroy(ctx->st);
  delete ctx;
}

#define CTX_CHECK()                                                           \
  do {                                                                        \
    if (!ctx) {                                                               \
      g_last_error = "null context";                                          \
      return IBEM3D_EINVAL;                                                   \
    }                                                                         \
    if (hipSetDevice(ctx->dev) != hipSuccess) return fail(ctx, IBEM3D_EHIP, "hipSetDevice"); \
  } while (0)

// k_rows' moment form when every element is planar and the pointwise form is not asked for
static const RowsMoment* rows_moment_form(const ibem3d_ctx* ctx) {
  return ctx->d_cw && get_tuning(kTuneRowsPointwise) == 0 ? &ctx->rows_mf : nullptr;
}

// A's boundary columns before an assembly. The gather form of the boundary phase writes every entry
// of the boundary rows and the moment form of the rows phase every entry of the sphere rows (a
// column's first element writes it), so nothing needs clearing then; the per-point rows form
// accumulates into its entries, which are then cleared before it runs (k_zero_rows_mask).
static void zero_boundary_columns(ibem3d_ctx* ctx, hipStream_t st) {
  if (!ctx->bem_gather) fill_zero(ctx->d_A, sizeof(double) * (size_t)ctx->lda * 3 * ctx->NN, st);
}

// the boundary phase (A's boundary block and b's known-value terms; b only without write_A)
static int bem_enqueue(ibem3d_ctx* ctx, hipStream_t st, bool write_A) {
  if (ctx->bem_gather) {
    launch_bem_gather(ctx->d_qp, ctx->d_conn, ctx->d_nodes, ctx->NN, ctx->NE, ctx->lda, ctx->E0, ctx->nu0,
                      ctx->d_node_col, ctx->d_slot_col, ctx->d_uval, ctx->d_tval, ctx->d_A, ctx->d_b, ctx->d_bemws,
                      BemGather{ctx->d_bg_ptr, ctx->d_bg_ent, ctx->d_bg_dof}, st, write_A);
    return 0;
  }
  if (!ctx->d_bemws)
    if (int rc = dalloc(ctx, &ctx->d_bemws, bem_workspace_doubles(ctx->NN, ctx->NE))) return rc;
  launch_bem(ctx->d_qp, ctx->d_conn, ctx->d_nodes, ctx->NN, ctx->NE, ctx->lda, ctx->E0, ctx->nu0, ctx->d_node_col,
             ctx->d_slot_col, ctx->d_uval, ctx->d_tval, ctx->d_A, ctx->d_b, ctx->d_bemws, st, write_A);
  return 0;
}

// A[row, j] = 0 for the sphere rows (row0 + 60 h + r) of the orders in rowmask and columns j < ncols
__global__ __launch_bounds__(256) void k_zero_rows_mask(double* __restrict__ A, long lda, long row0, long nrows,
                                                        int rowmask) {
  const long r = (long)blockIdx.x * 256 + threadIdx.x;
  if (r >= nrows) return;
  const int loc = (int)(r % 60), order = loc < 6 ? 0 : (loc < 24 ? 1 : 2);
  if ((rowmask >> order) & 1) A[(long)blockIdx.y * lda + row0 + r] = 0.0;
}

static int after_launch(ibem3d_ctx* ctx, const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(ctx, IBEM3D_EHIP, std::string(what) + ": " + hipGetErrorString(e));
  return 0;
}

// The boundary phases accumulate into A and b: a phase that already ran since the last reset would
// count its entries twice, so it is refused.
static int phase_guard(ibem3d_ctx* ctx, unsigned bits, const char* what) {
  if (ctx->phases & bits)
    return fail(ctx, IBEM3D_ESTATE, std::string(what) + ": phase already ran since the last reset");
  return 0;
}

int ibem3d_add_bem(ibem3d_ctx* ctx) {
  CTX_CHECK();
  if (int rc = phase_guard(ctx, 1u, "add_bem")) return rc;
  {
    Scope s(ctx, "assemble.bem");
    if (int rc = bem_enqueue(ctx, ctx->st, true)) return rc;
  }
  ctx->phases |= 1u;
  ctx->solved = false;
  return after_launch(ctx, "k_bem");
}

int ibem3d_add_field(ibem3d_ctx* ctx) {
  CTX_CHECK();
  if (int rc = phase_guard(ctx, 2u, "add_field")) return rc;
  if (ctx->num > 0) {
    Scope s(ctx, "assemble.field");
    launch_field(ctx->d_nodes, ctx->NN, ctx->d_cen, ctx->d_rad, ctx->num, ctx->lda, 3L * ctx->NN,
                 ctx->M, ctx->d_A, ctx->st);
  }
  ctx->phases |= 2u;
  ctx->solved = false;
  return after_launch(ctx, "k_field");
}

int ibem3d_add_incl(ibem3d_ctx* ctx, int order) {
  CTX_CHECK();
  if (order < -1 || order > 2) return fail(ctx, IBEM3D_EINVAL, "add_incl: order must be -1, 0, 1, 2");
  const int mask = order < 0 ? 7 : (1 << order);
  if (int rc = phase_guard(ctx, (unsigned)mask << 2, "add_incl")) return rc;
  if (ctx->num > 0) {
    Scope s(ctx, "assemble.incl");
    launch_incl(ctx->d_cen, ctx->d_rad, ctx->num, ctx->lda, 3L * ctx->NN, ctx->M, mask, ctx->d_wdesc, ctx->d_A,
                ctx->st);
  }
  ctx->phases |= (unsigned)mask << 2;
  ctx->solved = false;
  return after_launch(ctx, "k_incl");
}

int ibem3d_add_bem_rows(ibem3d_ctx* ctx, int order) {
  CTX_CHECK();
  if (order < -1 || order > 2) return fail(ctx, IBEM3D_EINVAL, "add_bem_rows: order must be -1, 0, 1, 2");
  const int mask = order < 0 ? 7 : (1 << order);
  if (int rc = phase_guard(ctx, (unsigned)mask << 5, "add_bem_rows")) return rc;
  if (ctx->num > 0) {
    Scope s(ctx, "assemble.rows");
    if (!rows_moment_form(ctx) && ctx->bem_gather) {  // the per-point form accumulates: clear first
      const long nr = 60L * ctx->num;
      hipLaunchKernelGGL(k_zero_rows_mask, dim3((unsigned)((nr + 255) / 256), (unsigned)(3 * ctx->NN)), dim3(256), 0,
                         ctx->st, ctx->d_A, ctx->lda, 3L * ctx->NN, nr, mask);
    }
    launch_rows(ctx->d_qp, ctx->d_conn, ctx->d_cen, ctx->num, ctx->NE, ctx->lda, 3L * ctx->NN, ctx->M,
                mask, ctx->d_node_col, ctx->d_slot_col, ctx->d_uval, ctx->d_tval, ctx->d_A, ctx->d_b,
                ctx->st, true, rows_moment_form(ctx));
  }
  ctx->phases |= (unsigned)mask << 5;
  ctx->solved = false;
  return after_launch(ctx, "k_rows");
}

int ibem3d_apply_bc(ibem3d_ctx* ctx) {
  CTX_CHECK();
  if (ctx->phases != 0xFFu)
    return fail(ctx, IBEM3D_ESTATE, "apply_bc: every assembly phase must run first (mask " +
                                        std::to_string(ctx->phases) + ")");
  HIPC(hipStreamSynchronize(ctx->st));
  return 0;
}

// the block-Jacobi preconditioner's setup (defined with the solver below)
static int bj_alloc(ibem3d_ctx* ctx);
static int bj_enqueue_boundary(ibem3d_ctx* ctx, hipStream_t st, bool allow_nopiv = true);
static void bj_enqueue_blocks(ibem3d_ctx* ctx, hipStream_t st);
static int join_aux(ibem3d_ctx* ctx);

// the second stream and the events that order it against the main stream
static int ensure_aux(ibem3d_ctx* ctx) {
  if (ctx->st_aux) return 0;
  HIPC(hipStreamCreateWithFlags(&ctx->st_aux, hipStreamNonBlocking));
  for (hipEvent_t* e : {&ctx->ev_aux_in, &ctx->ev_aux_in2, &ctx->ev_aux_out, &ctx->ev_ov_in, &ctx->ev_ov_out,
                        &ctx->ev_st_in, &ctx->ev_st_out})
    HIPC(hipEventCreateWithFlags(e, hipEventDisableTiming));
  return 0;
}

int ibem3d_assemble(ibem3d_ctx* ctx) {
  CTX_CHECK();
  int rc;
  // After a block-Jacobi solve the next one needs the same preconditioner of the new A: its boundary
  // block (written only by the boundary phase) is factored, and its sphere blocks (written only by
  // the inclusion phase) inverted, on a second stream while the other phases run. Every phase of the
  // main stream is enqueued first: the setup is some two thousand small launches, and the host can
  // wait in a launch while the second stream's queue drains, which must not hold back the assembly.
  const bool bj_next = ctx->prepared_mode == IBEM3D_SOLVE_BJ_GMRES || ctx->prepared_mode == IBEM3D_SOLVE_BJ_GMRES16;
  if (bj_next) {
    if ((rc = join_aux(ctx))) return rc;
    if ((rc = ensure_aux(ctx))) return rc;
    if ((rc = bj_alloc(ctx))) return rc;
  }
  if ((rc = ibem3d_add_bem(ctx))) return rc;
  if (bj_next) HIPC(hipEventRecord(ctx->ev_aux_in, ctx->st));  // boundary block complete
  if ((rc = ibem3d_add_field(ctx))) return rc;
  // all three eigenstrain orders in one pass each: every row belongs to one order, so the entries
  // are the ones the per-order calls write, and the per-pair tables are built once instead of three
  // times
  if ((rc = ibem3d_add_incl(ctx, -1))) return rc;
  if (bj_next) HIPC(hipEventRecord(ctx->ev_aux_in2, ctx->st));  // sphere blocks complete
  if ((rc = ibem3d_add_bem_rows(ctx, -1))) return rc;
  if (bj_next) {
    HIPC(hipStreamWaitEvent(ctx->st_aux, ctx->ev_aux_in, 0));
    {
      Scope sc(ctx, "assemble.bj_boundary_aux", ctx->st_aux);
      if ((rc = bj_enqueue_boundary(ctx, ctx->st_aux))) return rc;
    }
    HIPC(hipStreamWaitEvent(ctx->st_aux, ctx->ev_aux_in2, 0));
    {
      Scope sc(ctx, "assemble.bj_blocks_aux", ctx->st_aux);
      bj_enqueue_blocks(ctx, ctx->st_aux);
    }
    HIPC(hipEventRecord(ctx->ev_aux_out, ctx->st_aux));
    ctx->bj_pending = true;
    if ((rc = after_launch(ctx, "block-Jacobi setup beside the assembly"))) return rc;
  }
  return ibem3d_apply_bc(ctx);
}

// the main stream waits for stresses still being computed on the second stream (they read the
// solution and the known values); with invalidate, they no longer count as the current solution's
static int join_stress(ibem3d_ctx* ctx, bool invalidate) {
  if (ctx->stress_pending) {
    HIPC(hipStreamWaitEvent(ctx->st, ctx->ev_st_out, 0));
    ctx->stress_pending = false;
  }
  if (invalidate) ctx->stress_spec = false;
  return 0;
}

int ibem3d_reset(ibem3d_ctx* ctx) {
  CTX_CHECK();
  if (int rc = join_stress(ctx, true)) return rc;
  if (int rc = join_aux(ctx)) return rc;  // the second stream may still read A's boundary block
  ctx->bj_ready = false;
  const long n = ctx->n;
  zero_boundary_columns(ctx, ctx->st);
  fill_zero(ctx->d_b, sizeof(double) * n, ctx->st);
  if (int rc = after_launch(ctx, "reset")) return rc;
  ctx->phases = 0;
  ctx->solved = false;
  if (ctx->a_factored) {
    // a factor in A's memory is overwritten by the next assembly
    if (ctx->prepared_mode == (IBEM3D_SOLVE_FP64_LU | IBEM3D_SOLVE_COMPACT)) ctx->prepared_mode = -1;
    ctx->a_factored = false;
  }
  return 0;
}

// Assemble the system again over A's memory (compact storage: A was overwritten by a factor), with
// the solve state kept. The kernels are deterministic, so A and b are the ones assembled before.
static int reassemble(ibem3d_ctx* ctx) {
  const bool solved = ctx->solved;
  const int prepared = ctx->prepared_mode;
  if (int rc = ibem3d_reset(ctx)) return rc;
  if (int rc = ibem3d_assemble(ctx)) return rc;
  ctx->solved = solved;
  ctx->prepared_mode = prepared == (IBEM3D_SOLVE_FP64_LU | IBEM3D_SOLVE_COMPACT) ? -1 : prepared;
  return 0;
}

int ibem3d_get_system(ibem3d_ctx* ctx, double* A, double* b) {
  CTX_CHECK();
  if (ctx->a_factored)
    return fail(ctx, IBEM3D_ESTATE, "get_system: A holds a factor (compact solve); assemble again");
  if (ctx->phases != 0xFFu)
    return fail(ctx, IBEM3D_ESTATE, "get_system: every assembly phase must run first (mask " +
                                        std::to_string(ctx->phases) + ")");
  const long n = ctx->n;
  if (A)
    HIPC(hipMemcpy2DAsync(A, sizeof(double) * n, ctx->d_A, sizeof(double) * ctx->lda, sizeof(double) * n, n,
                          hipMemcpyDeviceToHost, ctx->st));
  if (b) HIPC(hipMemcpyAsync(b, ctx->d_b, sizeof(double) * n, hipMemcpyDeviceToHost, ctx->st));
  HIPC(hipStreamSynchronize(ctx->st));
  return 0;
}

// ---- solve ------------------------------------------------------------------------------------
// the device workspaces of an LU of order n (FP64-sized, so FP32 fits too)
static int alloc_luwork(ibem3d_ctx* ctx, LuWork& W, long n) {
  if (W.d_ipiv) return 0;
  if (int rc = dalloc(ctx, &W.d_ipiv, (size_t)n)) return rc;
  if (int rc = dalloc(ctx, &W.d_ws, lu_workspace_doubles(n))) return rc;
  fill_zero(W.d_ws, sizeof(double) * lu_workspace_doubles(n), ctx->st);
  if (int rc = dalloc(ctx, &W.d_info, 2)) return rc;  // info, and the no-pivot leaves' arrival count
  if (int rc = dalloc(ctx, &W.d_perm, lu_moves_ints(n))) return rc;
  if (int rc = dalloc(ctx, &W.d_inv, lu_inverse_doubles(n))) return rc;
  if (int rc = dalloc(ctx, &W.d_tmp, (size_t)n)) return rc;
  if (int rc = dalloc(ctx, &W.d_trinv, lu_trinv_doubles())) return rc;
  if (int rc = dalloc(ctx, &W.d_trsm, lu_trsm_doubles(n))) return rc;
  return 0;
}

static int ensure_solver_ws(ibem3d_ctx* ctx) {
  const long n = ctx->n;
  if (!ctx->W.d_ipiv) {
    if (int rc = alloc_luwork(ctx, ctx->W, n)) return rc;
    if (!ctx->d_eqws)
      if (int rc = dalloc(ctx, &ctx->d_eqws, equilibrate_workspace_doubles(n))) return rc;
    if (int rc = dalloc(ctx, &ctx->d_r, (size_t)n)) return rc;
    if (int rc = dalloc(ctx, &ctx->d_c, (size_t)n)) return rc;
    if (int rc = dalloc(ctx, &ctx->d_work, (size_t)n)) return rc;
    if (int rc = dalloc(ctx, &ctx->d_res, (size_t)n)) return rc;
    if (int rc = dalloc(ctx, &ctx->d_gemvws, gemv_workspace_doubles(n))) return rc;
  }
  return 0;
}

__global__ void k_scale_vec(double* y, const double* x, const double* s, long n) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = x[i] * s[i];
}
// y[perm[i]] = x[i] s[i]: the solution of the column-permuted boundary block back in column order
__global__ void k_scale_scatter(double* y, const double* x, const double* s, const int* perm, long n) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[perm[i]] = x[i] * s[i];
}
__global__ void k_scale_to_f32(float* y, const double* x, const double* s, long n) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = (float)(x[i] * s[i]);
}
__global__ void k_axpy_scaled_f32(double* x, const float* d, const double* s, long n) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] += (double)d[i] * s[i];
}

static dim3 vgrid(long n) { return dim3((unsigned)((n + 255) / 256)); }

// ---- GMRES-IR building blocks (FP64; reductions in a fixed order, so results are reproducible) ----
constexpr int kGmres = 100;     // Krylov vectors per cycle
constexpr int kDotBlocks = 128; // partial sums per dot product

__global__ void k_scale_from_f32(double* y, const float* x, const double* s, long n) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = (double)x[i] * s[i];
}
__global__ void k_scale_by(double* y, const double* x, double a, long n) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = a * x[i];
}
// y <- a x + c y
__global__ void k_axpby(double* y, const double* x, double a, double c, long n) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = a * x[i] + c * y[i];
}
// part[i * kDotBlocks + b] = block b's share of V_i . w, i < gridDim.y
__global__ __launch_bounds__(256) void k_dot_part(const double* V, long ldv, const double* w, long n,
                                                  double* part) {
  __shared__ double red[256];
  const double* v = V + (long)blockIdx.y * ldv;
  double acc = 0.0;
  for (long t = (long)blockIdx.x * 256 + threadIdx.x; t < n; t += (long)kDotBlocks * 256)
    acc += v[t] * w[t];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.y * kDotBlocks + blockIdx.x] = red[0];
}
// y += sign * sum_{i < k} c_i V_i
__global__ void k_combo(double* y, const double* V, long ldv, int k, const double* c, double sign,
                        long n) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  double acc = 0.0;
  for (int i = 0; i < k; ++i) acc += c[i] * V[(long)i * ldv + t];
  y[t] += sign * acc;
}

// h[0..k) = V_i . w for the first k columns of V (device partial sums, host total in a fixed order)
static int dots(ibem3d_ctx* ctx, const double* V, int k, const double* w, double* h) {
  const long n = ctx->n;
  hipLaunchKernelGGL(k_dot_part, dim3(kDotBlocks, (unsigned)k), dim3(256), 0, ctx->st, V, n, w, n,
                     ctx->d_gpart);
  std::vector<double> part((size_t)k * kDotBlocks);
  HIPC(hipMemcpyAsync(part.data(), ctx->d_gpart, sizeof(double) * part.size(), hipMemcpyDeviceToHost,
                      ctx->st));
  HIPC(hipStreamSynchronize(ctx->st));
  for (int i = 0; i < k; ++i) {
    double acc = 0.0;
    for (int b = 0; b < kDotBlocks; ++b) acc += part[(size_t)i * kDotBlocks + b];
    h[i] = acc;
  }
  return 0;
}

// ---- block-Jacobi preconditioner -------------------------------------------------------------
// Inverse of every sphere's 60 x 60 diagonal block of A (rows and columns 3 NN + 60 h ...): Gauss-
// Jordan elimination with partial pivoting on [B | I] in LDS, one workgroup per sphere. The output is
// column-major, inv + 3600 h. A zero pivot leaves the block's inverse as NaN (GMRES then stops and
// the solve falls back to the LU).
constexpr int kBlk = 60;
__global__ __launch_bounds__(256) void k_block_inv(const double* __restrict__ A, long n, long row0,
                                                   double* __restrict__ inv) {
  __shared__ double M[kBlk][2 * kBlk + 1];
  __shared__ int piv;
  const int h = blockIdx.x, t = threadIdx.x;
  const long base = row0 + (long)kBlk * h;
  for (int u = t; u < kBlk * kBlk; u += 256) {
    const int c = u / kBlk, r = u % kBlk;  // column-major reads: consecutive threads, consecutive rows
    M[r][c] = A[(base + c) * n + base + r];
    M[r][kBlk + c] = r == c ? 1.0 : 0.0;
  }
  __syncthreads();
  for (int k = 0; k < kBlk; ++k) {
    if (t == 0) {
      int p = k;
      double best = fabs(M[k][k]);
      for (int i = k + 1; i < kBlk; ++i)
        if (fabs(M[i][k]) > best) {
          best = fabs(M[i][k]);
          p = i;
        }
      piv = p;
    }
    __syncthreads();
    const int p = piv;
    if (p != k)
      for (int c = t; c < 2 * kBlk; c += 256) {
        const double a = M[k][c];
        M[k][c] = M[p][c];
        M[p][c] = a;
      }
    __syncthreads();
    const double d = 1.0 / M[k][k];
    __syncthreads();
    for (int c = t; c < 2 * kBlk; c += 256) M[k][c] *= d;
    __syncthreads();
    for (int u = t; u < kBlk * 2 * kBlk; u += 256) {
      const int i = u / (2 * kBlk), c = u % (2 * kBlk);
      if (i != k && c != k) M[i][c] -= M[i][k] * M[k][c];
    }
    __syncthreads();
    // the column k multipliers M[i][k] are read by the whole pass above: they are zeroed only now
    for (int i = t; i < kBlk; i += 256)
      if (i != k) M[i][k] = 0.0;
    __syncthreads();
  }
  double* out = inv + (size_t)kBlk * kBlk * h;
  for (int u = t; u < kBlk * kBlk; u += 256) {
    const int c = u / kBlk, r = u % kBlk;
    out[u] = M[r][kBlk + c];
  }
}

// z_h = inv_h v_h for every sphere block h (rows 3 NN + 60 h ...)
__global__ __launch_bounds__(64) void k_block_mv(const double* __restrict__ inv, const double* __restrict__ v,
                                                 double* __restrict__ z, long row0) {
  __shared__ double sv[kBlk];
  const int h = blockIdx.x, t = threadIdx.x;
  const long base = row0 + (long)kBlk * h;
  if (t < kBlk) sv[t] = v[base + t];
  __syncthreads();
  if (t >= kBlk) return;
  const double* B = inv + (size_t)kBlk * kBlk * h;
  double acc0 = 0.0, acc1 = 0.0;
  for (int c = 0; c < kBlk; c += 2) {
    acc0 += B[c * kBlk + t] * sv[c];
    acc1 += B[(c + 1) * kBlk + t] * sv[c + 1];
  }
  z[base + t] = acc0 + acc1;
}

// z = M^-1 v with M = blockdiag(A_bb, A_11, ..., A_num,num): the boundary block through its
// equilibrated LU, C_b (R_b A_bb C_b)^-1 R_b, the sphere blocks through their inverses
// x_b = C t for the boundary block's solution t (scattered back when its columns were permuted)
static void bj_unscale(ibem3d_ctx* ctx, double* xb, const double* t, hipStream_t st) {
  const long nb = 3L * ctx->NN;
  if (ctx->bj_nopiv)
    hipLaunchKernelGGL(k_scale_scatter, vgrid(nb), dim3(256), 0, st, xb, t, ctx->d_cb, ctx->d_bperm, nb);
  else
    hipLaunchKernelGGL(k_scale_vec, vgrid(nb), dim3(256), 0, st, xb, t, ctx->d_cb, nb);
}

__global__ void k_to_f32(float* __restrict__ y, const double* __restrict__ x, long n) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = (float)x[i];
}
__global__ void k_from_f32(double* __restrict__ y, const float* __restrict__ x, long n) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = (double)x[i];
}

// d_tb <- (R A_bb C)^-1 d_tb on stream st (an FP32 factor solves with FP32 vectors)
static void bj_boundary_solve(ibem3d_ctx* ctx, hipStream_t st) {
  const long nb = 3L * ctx->NN;
  if (ctx->bj_f32) {
    float* t = reinterpret_cast<float*>(ctx->d_tb32);
    hipLaunchKernelGGL(k_to_f32, vgrid(nb), dim3(256), 0, st, t, ctx->d_tb, nb);
    lu_solve_f32(reinterpret_cast<const float*>(ctx->d_Abb), nb, nb, ctx->Wb, t, st);
    hipLaunchKernelGGL(k_from_f32, vgrid(nb), dim3(256), 0, st, ctx->d_tb, t, nb);
  } else {
    lu_solve_f64(ctx->d_Abb, nb, nb, ctx->Wb, ctx->d_tb, st);
  }
}

static void bj_apply(ibem3d_ctx* ctx, const double* v, double* z) {
  const long nb = 3L * ctx->NN;
  hipLaunchKernelGGL(k_scale_vec, vgrid(nb), dim3(256), 0, ctx->st, ctx->d_tb, v, ctx->d_rb, nb);
  bj_boundary_solve(ctx, ctx->st);
  bj_unscale(ctx, z, ctx->d_tb, ctx->st);
  if (ctx->num > 0)
    hipLaunchKernelGGL(k_block_mv, dim3(ctx->num), dim3(64), 0, ctx->st, ctx->d_Binv, v, z, nb);
}

// the preconditioner from the assembled A: boundary block copied, equilibrated and factored; every
// sphere block inverted
static int bj_alloc(ibem3d_ctx* ctx) {
  const long nb = 3L * ctx->NN;
  if (!ctx->d_Abb) {
    if (int rc = dalloc(ctx, &ctx->d_Abb, (size_t)nb * nb)) return rc;
    if (int rc = dalloc(ctx, &ctx->d_rb, (size_t)nb)) return rc;
    if (int rc = dalloc(ctx, &ctx->d_cb, (size_t)nb)) return rc;
    if (int rc = dalloc(ctx, &ctx->d_tb, (size_t)nb)) return rc;
    if (int rc = dalloc(ctx, &ctx->d_tb32, (size_t)nb)) return rc;
    if (int rc = dalloc(ctx, &ctx->d_Binv, (size_t)kBlk * kBlk * std::max(ctx->num, 1))) return rc;
    if (int rc = alloc_luwork(ctx, ctx->Wb, nb)) return rc;
    if (!ctx->d_eqws)
      if (int rc = dalloc(ctx, &ctx->d_eqws, equilibrate_workspace_doubles(ctx->n))) return rc;
  }
  return 0;
}

// the boundary block's copy, equilibration and LU on stream st (reads A's first 3 NN rows and columns)
// Pivots of the equilibrated, column-permuted boundary block below this fall back to partial pivoting
constexpr double kBjPivTol = 1e-6;

static int bj_enqueue_boundary(ibem3d_ctx* ctx, hipStream_t st, bool allow_nopiv) {
  const long n = ctx->n, nb = 3L * ctx->NN;
  const long mode = get_tuning(kTuneBjNopiv);
  ctx->bj_nopiv = allow_nopiv && ctx->d_bperm && mode != 0;
  ctx->bj_f32 = ctx->bj_nopiv && get_tuning(kTuneBjFp32) != 0;
  if (ctx->bj_f32) {
    float* F = reinterpret_cast<float*>(ctx->d_Abb);
    if (equilibrate_copy_f32(ctx->d_A, nb, ctx->lda, ctx->d_rb, ctx->d_cb, F, nb, ctx->d_eqws, st, nullptr,
                             ctx->d_bperm) ||
        lu_factor_f32_nopiv(F, nb, nb, ctx->Wb, st, mode == 2 ? 1e300 : kBjPivTol))
      return fail(ctx, IBEM3D_EHIP, "boundary block LU launch failed");
    if (lu_prepare_solve_f32(F, nb, nb, ctx->Wb, st))
      return fail(ctx, IBEM3D_EHIP, "boundary block solve preparation failed");
    return 0;
  }
  if (ctx->bj_nopiv) {
    // rows and their own unknowns on the diagonal: no interchanges (info reports a small pivot)
    equilibrate_copy_f64(ctx->d_A, nb, ctx->lda, ctx->d_rb, ctx->d_cb, ctx->d_Abb, nb, ctx->d_eqws, st,
                         ctx->d_bperm);
    if (lu_factor_f64_nopiv(ctx->d_Abb, nb, nb, ctx->Wb, st, mode == 2 ? 1e300 : kBjPivTol))
      return fail(ctx, IBEM3D_EHIP, "boundary block LU launch failed");
  } else {
    equilibrate_copy_f64(ctx->d_A, nb, ctx->lda, ctx->d_rb, ctx->d_cb, ctx->d_Abb, nb, ctx->d_eqws, st);
    if (lu_factor_f64(ctx->d_Abb, nb, nb, ctx->Wb, st, nullptr, nullptr))
      return fail(ctx, IBEM3D_EHIP, "boundary block LU launch failed");
  }
  if (lu_prepare_solve_f64(ctx->d_Abb, nb, nb, ctx->Wb, st))
    return fail(ctx, IBEM3D_EHIP, "boundary block solve preparation failed");
  return 0;
}

// every sphere block's inverse on stream st (reads the sphere blocks of A's diagonal)
static void bj_enqueue_blocks(ibem3d_ctx* ctx, hipStream_t st) {
  if (ctx->num > 0)
    hipLaunchKernelGGL(k_block_inv, dim3(ctx->num), dim3(256), 0, st, ctx->d_A, ctx->lda, 3L * ctx->NN,
                       ctx->d_Binv);
}

static int bj_check(ibem3d_ctx* ctx) {
  int info = 0;
  HIPC(hipMemcpyAsync(&info, ctx->Wb.d_info, sizeof(int), hipMemcpyDeviceToHost, ctx->st));
  HIPC(hipStreamSynchronize(ctx->st));
  if (info && ctx->bj_nopiv) {
    // a small pivot without interchanges: the boundary block again, with partial pivoting
    if (int rc = bj_enqueue_boundary(ctx, ctx->st, false)) return rc;
    HIPC(hipMemcpyAsync(&info, ctx->Wb.d_info, sizeof(int), hipMemcpyDeviceToHost, ctx->st));
    HIPC(hipStreamSynchronize(ctx->st));
  }
  if (info) return fail(ctx, IBEM3D_ESOLVE, "boundary block LU: zero pivot");
  ctx->bj_ready = true;
  return after_launch(ctx, "block-Jacobi setup");
}

// the main stream takes over the setup launched beside the assembly
static int join_aux(ibem3d_ctx* ctx) {
  if (!ctx->bj_pending) return 0;
  ctx->bj_pending = false;
  HIPC(hipStreamWaitEvent(ctx->st, ctx->ev_aux_out, 0));
  return bj_check(ctx);
}

static int bj_setup(ibem3d_ctx* ctx) {
  if (int rc = bj_alloc(ctx)) return rc;
  {
    Scope s(ctx, "solve.bj_setup");
    if (int rc = bj_enqueue_boundary(ctx, ctx->st)) return rc;
    bj_enqueue_blocks(ctx, ctx->st);
  }
  return bj_check(ctx);
}

// v <- (LU)^-1 v with the factors of the prepared mixed mode: the FP32 factor, or its 16-bit copy
// (compact storage)
static void precond_solve(ibem3d_ctx* ctx, float* v) {
  if (ctx->prepared_mode & IBEM3D_SOLVE_COMPACT)
    lu_solve_16(ctx->d_F16, ctx->n, ctx->ldf, ctx->d_F16s, ctx->W, v, ctx->st);
  else
    lu_solve_f32(ctx->d_A32, ctx->n, ctx->lda32, ctx->W, v, ctx->st);
}

// z ~ A^-1 v for the outer GMRES of IBEM3D_SOLVE_BJ_GMRES16: GMRES on A16 z = v from z = 0, right-
// preconditioned by block Jacobi, CGS2 Arnoldi, at most kTuneInnerIters iterations, stopping at the
// relative residual 10^-kTuneInnerTolExp (the 16-bit matrix holds A to about 2^-12, so a tighter
// inner solve buys the outer iteration nothing)
// z = M^-1 v, then w = A16 z. With IBEM3D_TUNE_BJ_OVERLAP the boundary block's triangular solves
// (a latency-bound chain of small launches) run on the second stream while the product covers the
// column tiles past the boundary block (from the sphere blocks' part of z); the tiles holding
// boundary columns follow once that part of z is solved. The chunk sums are added in a fixed order.
static int bj_apply_m16(ibem3d_ctx* ctx, const double* v, double* z, double* w) {
  const long n = ctx->n, nb = 3L * ctx->NN;
  const long tiles = m16_tiles(n), tb = (nb + 31) / 32;  // tiles [0, tb) hold the boundary columns
  if (!get_tuning(kTuneBjOverlap) || tb >= tiles) {
    bj_apply(ctx, v, z);
    Scope s(ctx, "solve.gemv_m16");
    gemv_m16(n, 1.0, ctx->d_A16, ctx->ld16, ctx->d_A16s, z, 0.0, w, ctx->d_gemvws, ctx->st);
    return 0;
  }
  if (int rc = ensure_aux(ctx)) return rc;
  // the main stream's launches first: the host then spends the boundary solve's ~90 launches while
  // the product runs, not before it
  hipLaunchKernelGGL(k_scale_vec, vgrid(nb), dim3(256), 0, ctx->st, ctx->d_tb, v, ctx->d_rb, nb);
  HIPC(hipEventRecord(ctx->ev_ov_in, ctx->st));
  if (ctx->num > 0)
    hipLaunchKernelGGL(k_block_mv, dim3(ctx->num), dim3(64), 0, ctx->st, ctx->d_Binv, v, z, nb);
  // chunks of the boundary tiles and of the rest: 6 + 3 instead of the plain product's 16, so that
  // the boundary solve's 1024-thread workgroups find room beside the product's (measured on the
  // 2,500-sphere RVE: 12 + 4 0.716 s per step, 6 + 3 0.701, 5 + 3 0.697 with a slower product, 8 + 2
  // 0.724, 7 + 3 0.728)
  constexpr int kFirst = 3, kRest = 6;
  {
    Scope s(ctx, "solve.gemv_m16");
    gemv_m16_part(n, ctx->d_A16, ctx->ld16, ctx->d_A16s, z, ctx->d_gemvws, tb, tiles, kFirst, kRest, ctx->st);
  }
  HIPC(hipStreamWaitEvent(ctx->st_aux, ctx->ev_ov_in, 0));
  bj_boundary_solve(ctx, ctx->st_aux);
  bj_unscale(ctx, z, ctx->d_tb, ctx->st_aux);
  HIPC(hipEventRecord(ctx->ev_ov_out, ctx->st_aux));
  HIPC(hipStreamWaitEvent(ctx->st, ctx->ev_ov_out, 0));
  {
    Scope s(ctx, "solve.gemv_m16");
    gemv_m16_part(n, ctx->d_A16, ctx->ld16, ctx->d_A16s, z, ctx->d_gemvws, 0, tb, 0, kFirst, ctx->st);
    gemv_m16_finish(n, 1.0, ctx->d_gemvws, 0.0, w, ctx->st, kFirst + kRest);
  }
  return 0;
}

static int inner_m16(ibem3d_ctx* ctx, const double* v, double* z) {
  const long n = ctx->n;
  const int kmax = (int)get_tuning(kTuneInnerIters);
  const double tol = std::pow(10.0, -(double)get_tuning(kTuneInnerTolExp));
  double* V = ctx->d_iV;
  double* Z = ctx->d_iZ;
  double* w = ctx->d_iw;
  double beta = 0.0;
  if (int rc = dots(ctx, v, 1, v, &beta)) return rc;
  beta = std::sqrt(beta);
  fill_zero(z, sizeof(double) * n, ctx->st);
  if (!(beta > 0)) return 0;
  hipLaunchKernelGGL(k_scale_by, vgrid(n), dim3(256), 0, ctx->st, V, v, 1.0 / beta, n);
  std::vector<double> H((size_t)(kmax + 1) * kmax, 0.0), cs(kmax), sn(kmax), g(kmax + 1, 0.0);
  g[0] = beta;
  int k = 0;
  for (int j = 0; j < kmax; ++j) {
    double* vj = V + (size_t)j * n;
    double* zj = Z + (size_t)j * n;
    if (int rc = bj_apply_m16(ctx, vj, zj, w)) return rc;
    double* Hj = H.data() + (size_t)j * (kmax + 1);
    std::vector<double> h(j + 1);
    for (int pass = 0; pass < 2; ++pass) {
      if (int rc = dots(ctx, V, j + 1, w, h.data())) return rc;
      HIPC(hipMemcpyAsync(ctx->d_gh, h.data(), sizeof(double) * (j + 1), hipMemcpyHostToDevice, ctx->st));
      hipLaunchKernelGGL(k_combo, vgrid(n), dim3(256), 0, ctx->st, w, V, n, j + 1, ctx->d_gh, -1.0, n);
      for (int i = 0; i <= j; ++i) Hj[i] += h[i];
    }
    double hn = 0.0;
    if (int rc = dots(ctx, w, 1, w, &hn)) return rc;
    hn = std::sqrt(hn);
    Hj[j + 1] = hn;
    if (hn > 0 && j + 1 < kmax)
      hipLaunchKernelGGL(k_scale_by, vgrid(n), dim3(256), 0, ctx->st, V + (size_t)(j + 1) * n, w, 1.0 / hn, n);
    for (int i = 0; i < j; ++i) {
      const double t = cs[i] * Hj[i] + sn[i] * Hj[i + 1];
      Hj[i + 1] = -sn[i] * Hj[i] + cs[i] * Hj[i + 1];
      Hj[i] = t;
    }
    const double den = std::hypot(Hj[j], Hj[j + 1]);
    cs[j] = den > 0 ? Hj[j] / den : 1.0;
    sn[j] = den > 0 ? Hj[j + 1] / den : 0.0;
    Hj[j] = den;
    Hj[j + 1] = 0.0;
    g[j + 1] = -sn[j] * g[j];
    g[j] = cs[j] * g[j];
    k = j + 1;
    ctx->inner_iters += 1;
    if (!(std::fabs(g[j + 1]) > tol * beta)) break;  // also stops on a breakdown (NaN)
  }
  std::vector<double> y(k);
  for (int i = k - 1; i >= 0; --i) {
    double acc = g[i];
    for (int c = i + 1; c < k; ++c) acc -= H[(size_t)c * (kmax + 1) + i] * y[c];
    const double d = H[(size_t)i * (kmax + 1) + i];
    y[i] = d != 0 ? acc / d : 0.0;
  }
  HIPC(hipMemcpyAsync(ctx->d_gh, y.data(), sizeof(double) * k, hipMemcpyHostToDevice, ctx->st));
  hipLaunchKernelGGL(k_combo, vgrid(n), dim3(256), 0, ctx->st, z, Z, n, k, ctx->d_gh, 1.0, n);
  return after_launch(ctx, "inner GMRES");
}

static void free_m16(ibem3d_ctx* ctx) {
  for (void* q : {(void*)ctx->d_A16, (void*)ctx->d_A16s, (void*)ctx->d_iV, (void*)ctx->d_iZ, (void*)ctx->d_iw})
    if (q) hipFree(q);
  ctx->d_A16 = nullptr;
  ctx->d_A16s = nullptr;
  ctx->d_iV = ctx->d_iZ = ctx->d_iw = nullptr;
}

// ||d_res|| / ||b||
static int residual_norm(ibem3d_ctx* ctx, double* rel) {
  const long n = ctx->n;
  std::vector<double> r(n), b(n);
  HIPC(hipMemcpyAsync(r.data(), ctx->d_res, sizeof(double) * n, hipMemcpyDeviceToHost, ctx->st));
  HIPC(hipMemcpyAsync(b.data(), ctx->d_b, sizeof(double) * n, hipMemcpyDeviceToHost, ctx->st));
  HIPC(hipStreamSynchronize(ctx->st));
  const double nb = dnrm2_host(b);
  *rel = nb > 0 ? dnrm2_host(r) / nb : dnrm2_host(r);
  return 0;
}

// relative residual ||b - A x|| / ||b|| with the FP64 system matrix (b - A x left in d_res)
static int residual(ibem3d_ctx* ctx, double* rel) {
  const long n = ctx->n;
  HIPC(hipMemcpyAsync(ctx->d_res, ctx->d_b, sizeof(double) * n, hipMemcpyDeviceToDevice, ctx->st));
  {
    Scope s(ctx, "solve.residual");
    gemv_f64(n, n, -1.0, ctx->d_A, ctx->lda, ctx->d_x, 1.0, ctx->d_res, ctx->d_gemvws, ctx->st);
  }
  return residual_norm(ctx, rel);
}

// The stresses of the x just formed, on the second stream while the main stream checks its residual
// (from the cycle's products, or by a product): kept for ibem3d_stress when that residual ends the
// solve, dropped when another cycle changes x (IBEM3D_TUNE_STRESS_BESIDE).
static int stress_alloc(ibem3d_ctx* ctx);
static void enqueue_stress(ibem3d_ctx* ctx, hipStream_t st);
static int stress_beside(ibem3d_ctx* ctx) {
  if (ctx->nump == 0 || !get_tuning(kTuneStressBeside)) return 0;
  if (int rc = ensure_aux(ctx)) return rc;
  if (int rc = stress_alloc(ctx)) return rc;
  HIPC(hipEventRecord(ctx->ev_st_in, ctx->st));
  HIPC(hipStreamWaitEvent(ctx->st_aux, ctx->ev_st_in, 0));
  {
    Scope sc(ctx, "post.stress", ctx->st_aux);
    enqueue_stress(ctx, ctx->st_aux);
  }
  HIPC(hipEventRecord(ctx->ev_st_out, ctx->st_aux));
  ctx->stress_pending = ctx->stress_spec = true;
  return after_launch(ctx, "k_stress beside the residual");
}

// GMRES-IR (Carson and Higham): FP64 GMRES on A d = r, right-preconditioned by the FP32 factors of
// the equilibrated system, M^-1 v = C LU32^-1 (R v); restarted every kGmres vectors, each cycle ending
// with the FP64 residual b - A x of its new x: r - (A Z) y from the cycle's own products when that
// meets tol (kTuneResidualUpdate), else by a product. Arnoldi by classical Gram-Schmidt applied twice;
// the small Hessenberg least-squares problem by Givens rotations on the host. iters counts
// preconditioner applications (LU32 solves). A cycle that does not converge costs one FP64
// matrix-vector product more than the same number of classic refinement steps, so the solve uses it
// when classic refinement stalls.
// warm: start from the current x, whose residual b - A x is in d_res with relative norm rel0.
static int gmres_ir(ibem3d_ctx* ctx, double tol, bool warm, double rel0, int max_cycles = 10) {
  const long n = ctx->n;
  if (!ctx->d_gV) {
    if (int rc = dalloc(ctx, &ctx->d_gV, (size_t)(kGmres + 1) * n)) return rc;
    if (int rc = dalloc(ctx, &ctx->d_gZ, (size_t)kGmres * n)) return rc;
    if (int rc = dalloc(ctx, &ctx->d_gW, (size_t)kGmres * n)) return rc;
    if (int rc = dalloc(ctx, &ctx->d_gpart, (size_t)(kGmres + 1) * kDotBlocks)) return rc;
    if (int rc = dalloc(ctx, &ctx->d_gh, (size_t)kGmres + 1)) return rc;
  }
  double* V = ctx->d_gV;
  double* Z = ctx->d_gZ;
  double* w = ctx->d_work;
  const bool upd = get_tuning(kTuneResidualUpdate) != 0;
  double bnorm = 0.0;
  if (int rc = dots(ctx, ctx->d_b, 1, ctx->d_b, &bnorm)) return rc;
  bnorm = std::sqrt(bnorm);
  double rel = bnorm > 0 ? 1.0 : 0.0;
  if (warm) {
    rel = rel0;
  } else {
    fill_zero(ctx->d_x, sizeof(double) * n, ctx->st);
    HIPC(hipMemcpyAsync(ctx->d_res, ctx->d_b, sizeof(double) * n, hipMemcpyDeviceToDevice, ctx->st));
  }
  double beta = rel * bnorm;
  for (int cyc = 0; cyc < max_cycles && rel > tol; ++cyc) {
    // V_0 = r / ||r||
    hipLaunchKernelGGL(k_scale_by, vgrid(n), dim3(256), 0, ctx->st, V, ctx->d_res, 1.0 / beta, n);
    std::vector<double> H((size_t)(kGmres + 1) * kGmres, 0.0), cs(kGmres), sn(kGmres),
        g(kGmres + 1, 0.0);
    g[0] = beta;
    int k = 0;
    for (int j = 0; j < kGmres; ++j) {
      double* vj = V + (size_t)j * n;
      double* zj = Z + (size_t)j * n;
      if (ctx->prepared_mode == IBEM3D_SOLVE_BJ_GMRES16 && !ctx->m16_skipped) {
        Scope s(ctx, "solve.inner_m16");
        if (int rc = inner_m16(ctx, vj, zj)) return rc;
      } else if (ctx->prepared_mode == IBEM3D_SOLVE_BJ_GMRES || ctx->prepared_mode == IBEM3D_SOLVE_BJ_GMRES16) {
        Scope s(ctx, "solve.precond_bj");
        bj_apply(ctx, vj, zj);
      } else {
        Scope s(ctx, "solve.trsv_f32");
        hipLaunchKernelGGL(k_scale_to_f32, vgrid(n), dim3(256), 0, ctx->st, ctx->d_v32, vj, ctx->d_r, n);
        precond_solve(ctx, ctx->d_v32);
        hipLaunchKernelGGL(k_scale_from_f32, vgrid(n), dim3(256), 0, ctx->st, zj, ctx->d_v32, ctx->d_c, n);
      }
      {
        Scope s(ctx, "solve.residual");
        gemv_f64(n, n, 1.0, ctx->d_A, ctx->lda, zj, 0.0, w, ctx->d_gemvws, ctx->st);
      }
      if (upd)
        HIPC(hipMemcpyAsync(ctx->d_gW + (size_t)j * n, w, sizeof(double) * n, hipMemcpyDeviceToDevice,
                            ctx->st));
      if (int rc = after_launch(ctx, "gmres")) return rc;
      double* Hj = H.data() + (size_t)j * (kGmres + 1);  // column j
      {
        Scope s(ctx, "solve.gmres");
        std::vector<double> h(j + 1);
        for (int pass = 0; pass < 2; ++pass) {
          if (int rc = dots(ctx, V, j + 1, w, h.data())) return rc;
          HIPC(hipMemcpyAsync(ctx->d_gh, h.data(), sizeof(double) * (j + 1), hipMemcpyHostToDevice,
                              ctx->st));
          hipLaunchKernelGGL(k_combo, vgrid(n), dim3(256), 0, ctx->st, w, V, n, j + 1, ctx->d_gh, -1.0, n);
          for (int i = 0; i <= j; ++i) Hj[i] += h[i];
        }
        double hn = 0.0;
        if (int rc = dots(ctx, w, 1, w, &hn)) return rc;
        hn = std::sqrt(hn);
        Hj[j + 1] = hn;
        if (hn > 0)
          hipLaunchKernelGGL(k_scale_by, vgrid(n), dim3(256), 0, ctx->st, V + (size_t)(j + 1) * n, w,
                             1.0 / hn, n);
      }
      // rotate column j into upper-triangular form and update the residual norm estimate
      for (int i = 0; i < j; ++i) {
        const double t = cs[i] * Hj[i] + sn[i] * Hj[i + 1];
        Hj[i + 1] = -sn[i] * Hj[i] + cs[i] * Hj[i + 1];
        Hj[i] = t;
      }
      const double den = std::hypot(Hj[j], Hj[j + 1]);
      cs[j] = den > 0 ? Hj[j] / den : 1.0;
      sn[j] = den > 0 ? Hj[j + 1] / den : 0.0;
      Hj[j] = den;
      Hj[j + 1] = 0.0;
      g[j + 1] = -sn[j] * g[j];
      g[j] = cs[j] * g[j];
      k = j + 1;
      ctx->iters += 1;
      if (!(std::fabs(g[j + 1]) > 0.5 * tol * bnorm)) break;  // also stops on a breakdown (NaN)
    }
    // y = R^-1 g over the k columns; x += Z y
    std::vector<double> y(k);
    for (int i = k - 1; i >= 0; --i) {
      double acc = g[i];
      for (int c = i + 1; c < k; ++c) acc -= H[(size_t)c * (kGmres + 1) + i] * y[c];
      const double d = H[(size_t)i * (kGmres + 1) + i];
      y[i] = d != 0 ? acc / d : 0.0;
    }
    HIPC(hipMemcpyAsync(ctx->d_gh, y.data(), sizeof(double) * k, hipMemcpyHostToDevice, ctx->st));
    if (int rc = join_stress(ctx, true)) return rc;  // the previous cycle's stresses read x
    hipLaunchKernelGGL(k_combo, vgrid(n), dim3(256), 0, ctx->st, ctx->d_x, Z, n, k, ctx->d_gh, 1.0, n);
    if (int rc = stress_beside(ctx)) return rc;
    bool met = false;
    if (upd) {
      // b - A x = r - (A Z) y with the cycle's own FP64 products: the new x's residual without
      // another pass over A (it differs from b - A x computed by a product by rounding alone)
      hipLaunchKernelGGL(k_combo, vgrid(n), dim3(256), 0, ctx->st, ctx->d_res, ctx->d_gW, n, k, ctx->d_gh,
                         -1.0, n);
      if (int rc = residual_norm(ctx, &rel)) return rc;
      met = rel <= tol;
    }
    if (!met)
      if (int rc = residual(ctx, &rel)) return rc;  // r = b - A x in d_res
    ctx->rel_res = rel;
    if (!(rel == rel)) return fail(ctx, IBEM3D_ESOLVE, "refinement diverged (NaN)");
    beta = rel * bnorm;
  }
  ctx->solved = true;
  if (rel > tol)
    return fail(ctx, IBEM3D_ESOLVE, "refinement stopped at relative residual " + std::to_string(rel));
  return 0;
}

// Free device memory is checked before the second n x n copy: without room for it the factor takes
// the place of A (IBEM3D_SOLVE_COMPACT). kReserve covers the workspaces allocated beside it.
constexpr size_t kReserve = (size_t)4 << 30;
static bool fits(size_t bytes) {
  size_t fr = 0, tot = 0;
  if (hipMemGetInfo(&fr, &tot) != hipSuccess) return true;  // hipMalloc decides
  return bytes + kReserve <= fr;
}

static int check_info(ibem3d_ctx* ctx, const char* what) {
  int info = 0;
  HIPC(hipMemcpyAsync(&info, ctx->W.d_info, sizeof(int), hipMemcpyDeviceToHost, ctx->st));
  HIPC(hipStreamSynchronize(ctx->st));
  if (info) return fail(ctx, IBEM3D_ESOLVE, std::string(what) + ": zero pivot at column " + std::to_string(info - 1));
  return 0;
}

// classic refinement x <- x + C LU^-1 (R r) from x = 0; GMRES-IR takes over from the current x when
// an iteration gains less than a factor of 10 (from zero when it diverges)
static int classic_ir(ibem3d_ctx* ctx, double tol) {
  const long n = ctx->n;
  fill_zero(ctx->d_x, sizeof(double) * n, ctx->st);
  HIPC(hipMemcpyAsync(ctx->d_res, ctx->d_b, sizeof(double) * n, hipMemcpyDeviceToDevice, ctx->st));
  double rel = 1.0;
  const int max_it = 30;
  for (int it = 0; it < max_it; ++it) {
    {
      Scope s(ctx, "solve.trsv_f32");
      hipLaunchKernelGGL(k_scale_to_f32, vgrid(n), dim3(256), 0, ctx->st, ctx->d_v32, ctx->d_res,
                         ctx->d_r, n);
      precond_solve(ctx, ctx->d_v32);
      hipLaunchKernelGGL(k_axpy_scaled_f32, vgrid(n), dim3(256), 0, ctx->st, ctx->d_x, ctx->d_v32,
                         ctx->d_c, n);
    }
    if (int rc = after_launch(ctx, "mixed refinement")) return rc;
    double rel_new = 0.0;
    if (int rc = residual(ctx, &rel_new)) return rc;
    ctx->iters = it + 1;
    ctx->rel_res = rel_new;
    if (!(rel_new == rel_new)) return fail(ctx, IBEM3D_ESOLVE, "refinement diverged (NaN)");
    if (rel_new <= tol) {
      ctx->solved = true;
      return 0;
    }
    if (rel_new > 0.1 * rel) {
      if (rel_new > rel) return gmres_ir(ctx, tol, false, 1.0);  // diverging: restart from zero
      return gmres_ir(ctx, tol, true, rel_new);
    }
    rel = rel_new;
  }
  ctx->solved = true;
  return fail(ctx, IBEM3D_ESOLVE, "refinement stopped at relative residual " + std::to_string(rel));
}

static int solve_fp64(ibem3d_ctx* ctx, bool reuse, bool compact) {
  const long n = ctx->n;
  if (!reuse) {
    ctx->prepared_mode = -1;
    double* F = ctx->d_A;
    if (compact) {
      // R A C in place; A is gone until the next assembly
      Scope s(ctx, "solve.equilibrate");
      equilibrate_in_place_f64(ctx->d_A, n, ctx->lda, ctx->d_r, ctx->d_c, ctx->d_eqws, ctx->st);
      ctx->a_factored = true;
    } else {
      // keep A for the residual; factor an equilibrated copy R A C
      if (!ctx->d_Af)
        if (int rc = dalloc(ctx, &ctx->d_Af, (size_t)ctx->lda * n)) return rc;
      Scope s(ctx, "solve.equilibrate");
      equilibrate_copy_f64(ctx->d_A, n, ctx->lda, ctx->d_r, ctx->d_c, ctx->d_Af, ctx->lda, ctx->d_eqws, ctx->st);
      F = ctx->d_Af;
    }
    {
      Scope s(ctx, "solve.lu_f64");
      if (lu_factor_f64(F, n, ctx->lda, ctx->W, ctx->st, tick, ctx))
        return fail(ctx, IBEM3D_EHIP, "LU launch failed");
    }
    if (int rc = check_info(ctx, "LU")) return rc;
    {
      Scope s(ctx, "solve.prepare");
      if (lu_prepare_solve_f64(F, n, ctx->lda, ctx->W, ctx->st))
        return fail(ctx, IBEM3D_EHIP, "solve preparation failed");
    }
    ctx->prepared_mode = IBEM3D_SOLVE_FP64_LU | (compact ? IBEM3D_SOLVE_COMPACT : 0);
  }
  const double* F = (ctx->prepared_mode & IBEM3D_SOLVE_COMPACT) ? ctx->d_A : ctx->d_Af;
  {
    Scope s(ctx, "solve.trsv_f64");
    hipLaunchKernelGGL(k_scale_vec, vgrid(n), dim3(256), 0, ctx->st, ctx->d_x, ctx->d_b, ctx->d_r, n);
    lu_solve_f64(F, n, ctx->lda, ctx->W, ctx->d_x, ctx->st);
    hipLaunchKernelGGL(k_scale_vec, vgrid(n), dim3(256), 0, ctx->st, ctx->d_x, ctx->d_x, ctx->d_c, n);
  }
  if (int rc = after_launch(ctx, "fp64 solve")) return rc;
  ctx->solved = true;
  if (ctx->a_factored) {
    // compact: A holds the factor; ibem3d_residual assembles it again for the residual
    HIPC(hipStreamSynchronize(ctx->st));
    ctx->rel_res = -1.0;
    return 0;
  }
  return residual(ctx, &ctx->rel_res);
}

static int solve_mixed(ibem3d_ctx* ctx, int mode, double tol, bool gmres_first, bool reuse, bool compact) {
  const long n = ctx->n;
  if (!(tol > 0)) tol = 1e-10;
  const int prod = mode == IBEM3D_SOLVE_MIXED_FP16_1   ? kProdF16
                   : mode == IBEM3D_SOLVE_MIXED_BF16   ? kProdBF16
                   : mode == IBEM3D_SOLVE_MIXED_BF16X3 ? kProdBF16x3
                                                       : kProdF16x3;
  const bool split = mode != IBEM3D_SOLVE_MIXED_FP32;
  ctx->lda32 = (n + 63) / 64 * 64;
  if (!ctx->d_v32)
    if (int rc = dalloc(ctx, &ctx->d_v32, (size_t)n)) return rc;
  if (!reuse) {
    ctx->prepared_mode = -1;
    // the split workspace also stages the first columns of the in-place FP32 conversion
    if ((split || compact) && !ctx->d_f16ws)
      if (int rc = dalloc(ctx, &ctx->d_f16ws, lu_f16_workspace_doubles(n))) return rc;
    float* F32 = nullptr;
    if (compact) {
      F32 = (float*)ctx->d_A;
      Scope s(ctx, "solve.equilibrate");
      const size_t tmp_floats = lu_f16_workspace_doubles(n) * 2;
      ctx->a_factored = true;
      if (equilibrate_to_f32_in_place(ctx->d_A, n, ctx->lda, ctx->d_r, ctx->d_c, ctx->lda32, (float*)ctx->d_f16ws,
                                      tmp_floats, ctx->d_eqws, ctx->st))
        return fail(ctx, IBEM3D_EHIP, "in-place FP32 conversion failed");
    } else {
      if (!ctx->d_A32)
        if (int rc = dalloc(ctx, &ctx->d_A32, (size_t)ctx->lda32 * n)) return rc;
      F32 = ctx->d_A32;
      // R A C straight from A into the FP32 factor storage (A itself stays for the residual); columns
      // past the first panel are copied on the LU's second stream while that panel factors
      Scope s(ctx, "solve.equilibrate");
      if (equilibrate_copy_f32(ctx->d_A, n, ctx->lda, ctx->d_r, ctx->d_c, ctx->d_A32, ctx->lda32, ctx->d_eqws,
                               ctx->st, &ctx->W))
        return fail(ctx, IBEM3D_EHIP, "equilibration launch failed");
    }
    {
      Scope s(ctx, "solve.lu_f32");
      if (lu_factor_f32(F32, n, ctx->lda32, ctx->W, ctx->st, tick, ctx, split ? ctx->d_f16ws : nullptr,
                        !compact, prod))
        return fail(ctx, IBEM3D_EHIP, "LU launch failed");
    }
    if (int rc = check_info(ctx, "LU(fp32)")) return rc;
    {
      Scope s(ctx, "solve.prepare");
      if (lu_prepare_solve_f32(F32, n, ctx->lda32, ctx->W, ctx->st))
        return fail(ctx, IBEM3D_EHIP, "solve preparation failed");
    }
    if (compact) {
      // the factor in FP16 (per-tile scaled) beside A, then A assembled again over the FP32 factor;
      // FP16 storage for every product type: with the tile scaling its 11-bit significand is the
      // more accurate 16-bit format (BF16 storage left the BF16-product solve of a 160,374-unknown
      // RVE without GMRES convergence)
      ctx->ldf = ctx->lda32;
      if (!ctx->d_F16) {
        if (int rc = dalloc(ctx, &ctx->d_F16, (size_t)ctx->ldf * n)) return rc;
        if (int rc = dalloc(ctx, &ctx->d_F16s, factor16_scales(n))) return rc;
      }
      {
        Scope s(ctx, "solve.factor16");
        factor_to_16(F32, n, ctx->lda32, ctx->d_F16, ctx->ldf, ctx->d_F16s, ctx->st);
      }
      if (int rc = after_launch(ctx, "factor16")) return rc;
      ctx->prepared_mode = mode | IBEM3D_SOLVE_COMPACT;
      {
        Scope s(ctx, "solve.reassemble");
        if (int rc = reassemble(ctx)) return rc;
      }
    } else {
      ctx->prepared_mode = mode;
    }
  }
  // a 16-bit factor leaves a contraction too weak for classic refinement: GMRES-IR from the start
  if (gmres_first || (ctx->prepared_mode & IBEM3D_SOLVE_COMPACT)) return gmres_ir(ctx, tol, false, 1.0);
  return classic_ir(ctx, tol);
}

// FP64 GMRES right-preconditioned by block Jacobi (no n x n factor). With an initial guess of the
// eigenstrain unknowns, the boundary unknowns start from the boundary block's solve against it,
// x_b = A_bb^-1 (b_b - A_bs x_s). When GMRES does not reach tol the mixed single-FP16 LU solves the
// system instead (the fallback, compact storage when the copy does not fit).
constexpr int kBjCycles = 3;  // GMRES cycles (of kGmres iterations) before the LU fallback
// mode: IBEM3D_SOLVE_BJ_GMRES, or IBEM3D_SOLVE_BJ_GMRES16 (the 16-bit copy of A built here when it fits)
static int solve_bj(ibem3d_ctx* ctx, double tol, bool reuse, int mode) {
  const long n = ctx->n, nb = 3L * ctx->NN;
  if (!(tol > 0)) tol = 1e-10;
  ctx->bj_fallback = false;
  if (!reuse || ctx->prepared_mode != mode) {
    ctx->prepared_mode = -1;
    ctx->m16_skipped = false;
    // the FP16 copy reads A only: it is built while the second stream may still be finishing the
    // block-Jacobi setup launched beside the assembly (which reads A too)
    if (mode == IBEM3D_SOLVE_BJ_GMRES16) {
      const long ld16 = m16_ld(n);
      const size_t need = ctx->d_A16 ? 0
                                     : 2 * (size_t)ld16 * n + 4 * m16_scale_count(n) +
                                           8 * (size_t)(2 * kInnerItersMax + 2) * n;
      if (need > 0 && !fits(need)) {
        // the LU modes' factor storage is of no use once this mode is prepared: its memory first
        for (void** q : {(void**)&ctx->d_Af, (void**)&ctx->d_A32, (void**)&ctx->d_F16, (void**)&ctx->d_F16s})
          if (*q) {
            hipFree(*q);
            *q = nullptr;
          }
      }
      if (need == 0 || fits(need)) {
        if (!ctx->d_A16) {
          if (int rc = dalloc(ctx, &ctx->d_A16, (size_t)ld16 * n)) return rc;
          if (int rc = dalloc(ctx, &ctx->d_A16s, m16_scale_count(n))) return rc;
          if (int rc = dalloc(ctx, &ctx->d_iV, (size_t)(kInnerItersMax + 1) * n)) return rc;
          if (int rc = dalloc(ctx, &ctx->d_iZ, (size_t)kInnerItersMax * n)) return rc;
          if (int rc = dalloc(ctx, &ctx->d_iw, (size_t)n)) return rc;
          ctx->ld16 = ld16;
        }
        {
          Scope s(ctx, "solve.m16_build");
          m16_build(ctx->d_A, n, ctx->lda, ctx->d_A16, ctx->ld16, ctx->d_A16s, ctx->st);
        }
        if (int rc = after_launch(ctx, "16-bit copy of A")) return rc;
      } else {
        ctx->m16_skipped = true;
      }
    }
    if (int rc = join_aux(ctx)) return rc;
    if (!ctx->bj_ready)  // else built beside the assembly of this A
      if (int rc = bj_setup(ctx)) return rc;
    ctx->prepared_mode = mode;
  } else if (int rc = join_aux(ctx)) {
    return rc;
  }
  int rc;
  if (ctx->has_x0 && ctx->num > 0) {
    Scope s(ctx, "solve.warm_start");
    // x = [0; x_s], r = b - A x, then x_b = A_bb^-1 r_b
    fill_zero(ctx->d_x, sizeof(double) * nb, ctx->st);
    HIPC(hipMemcpyAsync(ctx->d_x + nb, ctx->d_x0s, sizeof(double) * (n - nb), hipMemcpyDeviceToDevice, ctx->st));
    HIPC(hipMemcpyAsync(ctx->d_res, ctx->d_b, sizeof(double) * n, hipMemcpyDeviceToDevice, ctx->st));
    gemv_f64(n, n, -1.0, ctx->d_A, ctx->lda, ctx->d_x, 1.0, ctx->d_res, ctx->d_gemvws, ctx->st);
    hipLaunchKernelGGL(k_scale_vec, vgrid(nb), dim3(256), 0, ctx->st, ctx->d_tb, ctx->d_res, ctx->d_rb, nb);
    bj_boundary_solve(ctx, ctx->st);
    bj_unscale(ctx, ctx->d_x, ctx->d_tb, ctx->st);
    double rel0 = 0.0;
    if ((rc = residual(ctx, &rel0))) return rc;
    // x0 <- alpha x0 with alpha minimising ||b - alpha A x0|| (A x0 = b - r0): the start is never
    // worse than x = 0, whatever the quality of the guess
    double bb = 0, br = 0, rr = 0;
    if ((rc = dots(ctx, ctx->d_b, 1, ctx->d_b, &bb)) || (rc = dots(ctx, ctx->d_b, 1, ctx->d_res, &br)) ||
        (rc = dots(ctx, ctx->d_res, 1, ctx->d_res, &rr)))
      return rc;
    const double den = bb - 2 * br + rr;
    const double alpha = den > 0 ? std::min(1.0, std::max(0.0, (bb - br) / den)) : 0.0;
    hipLaunchKernelGGL(k_scale_by, vgrid(n), dim3(256), 0, ctx->st, ctx->d_x, ctx->d_x, alpha, n);
    hipLaunchKernelGGL(k_axpby, vgrid(n), dim3(256), 0, ctx->st, ctx->d_res, ctx->d_b, 1.0 - alpha, alpha, n);
    rel0 = bb > 0 ? std::sqrt(std::max(0.0, (1 - alpha) * (1 - alpha) * bb + 2 * alpha * (1 - alpha) * br +
                                              alpha * alpha * rr) / bb)
                  : 0.0;
    ctx->rel_res = rel0;
    ctx->warm_rel = rel0;
    ctx->warm_alpha = alpha;
    rc = rel0 <= tol ? 0 : gmres_ir(ctx, tol, true, rel0, kBjCycles);
  } else {
    ctx->warm_rel = -1.0;
    rc = gmres_ir(ctx, tol, false, 1.0, kBjCycles);
  }
  if (rc == IBEM3D_ESOLVE) {
    // the fallback: the mixed LU (it reports its own refinement steps)
    ctx->bj_fallback = true;
    if (int rc2 = join_stress(ctx, true)) return rc2;
    const int bj_iters = ctx->iters;
    ctx->iters = 0;
    ctx->solved = false;
    free_m16(ctx);  // room for the LU
    const bool compact = !fits(sizeof(float) * (size_t)((n + 63) / 64 * 64) * n);
    rc = solve_mixed(ctx, IBEM3D_SOLVE_MIXED_FP16_1, tol, false, false, compact);
    ctx->iters += bj_iters;
  }
  ctx->solved = rc == 0 || ctx->solved;
  return rc;
}

int ibem3d_solve(ibem3d_ctx* ctx, int mode, double tol) {
  CTX_CHECK();
  if (int rc = join_stress(ctx, true)) return rc;
  const bool gmres_first = (mode & IBEM3D_SOLVE_GMRES) != 0;
  const bool reuse = (mode & IBEM3D_SOLVE_REUSE) != 0;
  bool compact = (mode & IBEM3D_SOLVE_COMPACT) != 0;
  mode &= ~(IBEM3D_SOLVE_GMRES | IBEM3D_SOLVE_REUSE | IBEM3D_SOLVE_COMPACT);
  if (mode < IBEM3D_SOLVE_FP64_LU || mode > IBEM3D_SOLVE_MIXED_BF16X3)
    return fail(ctx, IBEM3D_EINVAL, "solve: unknown mode");
  const bool bj = mode == IBEM3D_SOLVE_BJ_GMRES || mode == IBEM3D_SOLVE_BJ_GMRES16;
  const long n = ctx->n;
  if (reuse) {
    if (ctx->prepared_mode < 0 || (ctx->prepared_mode & ~IBEM3D_SOLVE_COMPACT) != mode)
      return fail(ctx, IBEM3D_ESTATE, "solve: no factors of this mode to reuse");
    compact = (ctx->prepared_mode & IBEM3D_SOLVE_COMPACT) != 0;
  } else if (!compact && !bj) {
    // the second n x n copy (FP64 factor, or FP32 factor with its 64-aligned columns)
    const size_t need = mode == IBEM3D_SOLVE_FP64_LU ? (ctx->d_Af ? 0 : sizeof(double) * (size_t)ctx->lda * n)
                                                     : (ctx->d_A32 ? 0 : sizeof(float) * (size_t)((n + 63) / 64 * 64) * n);
    compact = need > 0 && !fits(need);
  }
  if (ctx->phases != 0xFFu) return fail(ctx, IBEM3D_ESTATE, "solve: system not assembled");
  if (!bj)  // the block-Jacobi modes join the second stream themselves, after the FP16 copy's build
    if (int rc = join_aux(ctx)) return rc;
  if (ctx->a_factored && !(reuse && mode == IBEM3D_SOLVE_FP64_LU))
    return fail(ctx, IBEM3D_ESTATE, "solve: A holds a factor (compact solve); assemble again");
  if (int rc = ensure_solver_ws(ctx)) return rc;
  ctx->iters = 0;
  ctx->inner_iters = 0;
  ctx->rel_res = -1.0;
  if (mode != IBEM3D_SOLVE_BJ_GMRES16 && ctx->d_A16) {
    // the 16-bit copy of A gives its memory back to the other modes
    if (ctx->prepared_mode == IBEM3D_SOLVE_BJ_GMRES16) ctx->prepared_mode = -1;
    free_m16(ctx);
  }
  if (mode == IBEM3D_SOLVE_FP64_LU) return solve_fp64(ctx, reuse, compact);
  if (bj) return solve_bj(ctx, tol, reuse, mode);
  return solve_mixed(ctx, mode, tol, gmres_first, reuse, compact);
}

// Initial guess of the eigenstrain unknowns (60 num, the device's 60 layout) for the next block-
// Jacobi GMRES solves; NULL clears it.
int ibem3d_set_initial_guess(ibem3d_ctx* ctx, const double* x_eig) {
  CTX_CHECK();
  if (int rc = join_stress(ctx, true)) return rc;
  if (!x_eig) {
    ctx->has_x0 = false;
    return 0;
  }
  for (long i = 0; i < 60L * ctx->num; ++i)
    if (!std::isfinite(x_eig[i])) return fail(ctx, IBEM3D_EINVAL, "set_initial_guess: non-finite value");
  if (!ctx->d_x0s)
    if (int rc = dalloc(ctx, &ctx->d_x0s, (size_t)std::max(60L * ctx->num, 1L))) return rc;
  HIPC(hipMemcpyAsync(ctx->d_x0s, x_eig, sizeof(double) * 60L * ctx->num, hipMemcpyHostToDevice, ctx->st));
  HIPC(hipStreamSynchronize(ctx->st));
  ctx->has_x0 = true;
  return 0;
}

int ibem3d_solve_fell_back(const ibem3d_ctx* ctx) { return ctx && ctx->bj_fallback ? 1 : 0; }

int ibem3d_bj_nopiv(const ibem3d_ctx* ctx) { return ctx && ctx->bj_nopiv ? (ctx->bj_f32 ? 2 : 1) : 0; }

int ibem3d_rows_moment_form(const ibem3d_ctx* ctx) { return ctx && rows_moment_form(ctx) ? 1 : 0; }

long ibem3d_inner_iterations(const ibem3d_ctx* ctx) {
  if (!ctx) return -1;
  return ctx->prepared_mode == IBEM3D_SOLVE_BJ_GMRES16 && ctx->m16_skipped ? -1 : ctx->inner_iters;
}

int ibem3d_warm_start_info(const ibem3d_ctx* ctx, double* rel0, double* alpha) {
  if (!ctx) return IBEM3D_EINVAL;
  if (rel0) *rel0 = ctx->warm_rel;
  if (alpha) *alpha = ctx->warm_alpha;
  return 0;
}

// Relative residual of the current solution against the assembled FP64 system; after a compact
// FP64 solve (A overwritten by its factor) the system is assembled again first.
int ibem3d_residual(ibem3d_ctx* ctx, double* rel) {
  CTX_CHECK();
  if (!ctx->solved) return fail(ctx, IBEM3D_ESTATE, "residual: no solution");
  if (ctx->a_factored) {
    Scope s(ctx, "solve.reassemble");
    if (int rc = reassemble(ctx)) return rc;
  }
  if (int rc = ensure_solver_ws(ctx)) return rc;
  double r = 0.0;
  if (int rc = residual(ctx, &r)) return rc;
  ctx->rel_res = r;
  if (rel) *rel = r;
  return 0;
}

int ibem3d_solve_info(const ibem3d_ctx* ctx, double* rel, int* iters) {
  if (!ctx) return IBEM3D_EINVAL;
  if (rel) *rel = ctx->rel_res;
  if (iters) *iters = ctx->iters;
  return 0;
}

// New known boundary values with the same known/unknown pattern: the folded system matrix does not
// depend on the values, only b does, so the factors stay valid.
static int upload_known_values(ibem3d_ctx* ctx, const double* bc_values, const double* tbc_values) {
  if (!bc_values || !tbc_values) return fail(ctx, IBEM3D_EINVAL, "known values: null values");
  const int NN = ctx->NN, NE = ctx->NE;
  std::vector<double> uval(3L * NN), tval(12L * NE);
  for (long j = 0; j < 3L * NN; ++j) {
    if (!std::isfinite(bc_values[j])) return fail(ctx, IBEM3D_EINVAL, "known values: non-finite value");
    ctx->h_bc[2 * j + 1] = bc_values[j];
    uval[j] = bc_values[j];
  }
  for (long e = 0; e < NE; ++e)
    for (int i = 0; i < 12; ++i) {
      if (!std::isfinite(tbc_values[12 * e + i]))
        return fail(ctx, IBEM3D_EINVAL, "known values: non-finite value");
      ctx->h_tbc[24 * e + 12 + i] = tbc_values[12 * e + i];
      tval[12 * e + i] = tbc_values[12 * e + i];
    }
  HIPC(hipMemcpyAsync(ctx->d_uval, uval.data(), sizeof(double) * uval.size(), hipMemcpyHostToDevice,
                      ctx->st));
  HIPC(hipMemcpyAsync(ctx->d_tval, tval.data(), sizeof(double) * tval.size(), hipMemcpyHostToDevice,
                      ctx->st));
  HIPC(hipStreamSynchronize(ctx->st));
  return 0;
}

int ibem3d_set_known_values(ibem3d_ctx* ctx, const double* bc_values, const double* tbc_values) {
  CTX_CHECK();
  if (int rc = join_stress(ctx, true)) return rc;
  if (int rc = upload_known_values(ctx, bc_values, tbc_values)) return rc;
  return ibem3d_reset(ctx);
}

// b only (BEMBuilderElastic3D::ApplyBC's known-value folds, BEMBuilderElastic3D.cpp:1973-1990, of the
// boundary and boundary->inclusion phases), in the order of the full assembly; A, or the factor
// that occupies its memory, is not touched.
int ibem3d_update_known_values(ibem3d_ctx* ctx, const double* bc_values, const double* tbc_values) {
  CTX_CHECK();
  if (int rc = join_stress(ctx, true)) return rc;
  if (ctx->phases != 0xFFu)
    return fail(ctx, IBEM3D_ESTATE, "update_known_values: the system must have been assembled");
  if (int rc = upload_known_values(ctx, bc_values, tbc_values)) return rc;
  fill_zero(ctx->d_b, sizeof(double) * ctx->n, ctx->st);
  {
    Scope s(ctx, "assemble.rhs");
    if (int rc = bem_enqueue(ctx, ctx->st, false)) return rc;
    if (ctx->num > 0)
      launch_rows(ctx->d_qp, ctx->d_conn, ctx->d_cen, ctx->num, ctx->NE, ctx->lda, 3L * ctx->NN, ctx->M, 7,
                  ctx->d_node_col, ctx->d_slot_col, ctx->d_uval, ctx->d_tval, ctx->d_A, ctx->d_b,
                  ctx->st, false, rows_moment_form(ctx));
  }
  ctx->solved = false;
  return after_launch(ctx, "update_known_values");
}

// Host copies of the quadrature table, connectivity and radii (downloaded once) for the averages
static int host_geometry(ibem3d_ctx* ctx) {
  if (!ctx->h_qp.empty()) return 0;
  ctx->h_qp.resize((size_t)ctx->NE * 64 * kQpStride);
  ctx->h_conn.resize(4L * ctx->NE);
  ctx->h_rad.resize(ctx->num);
  HIPC(hipMemcpyAsync(ctx->h_qp.data(), ctx->d_qp, sizeof(double) * ctx->h_qp.size(), hipMemcpyDeviceToHost,
                      ctx->st));
  HIPC(hipMemcpyAsync(ctx->h_conn.data(), ctx->d_conn, sizeof(int) * ctx->h_conn.size(),
                      hipMemcpyDeviceToHost, ctx->st));
  if (ctx->num > 0)
    HIPC(hipMemcpyAsync(ctx->h_rad.data(), ctx->d_rad, sizeof(double) * ctx->num, hipMemcpyDeviceToHost,
                        ctx->st));
  HIPC(hipStreamSynchronize(ctx->st));
  return 0;
}

// getUTVector (BEMBuilderElastic3D.cpp:2053-2117) of a host solution x in the 78 layout; U or T may
// be null
static void ut_from_x(const ibem3d_ctx* ctx, const double* x, double* U, double* T) {
  const int NN = ctx->NN, NE = ctx->NE;
  if (U) {
    for (int i = 0; i < 3 * NN; ++i) U[i] = ctx->node_col[i] >= 0 ? x[ctx->node_col[i]] : ctx->h_bc[2 * i + 1];
    for (int h = 0; h < ctx->num; ++h) {
      const double* s = x + 3L * NN + 60L * h;
      double* d = U + 3L * NN + 78L * h;
      for (int u = 0; u < 24; ++u) d[u] = s[u];
      for (int t = 0; t < 6; ++t)
        for (int k = 0; k < 3; ++k)
          for (int l = 0; l < 3; ++l) d[24 + 9 * t + 3 * k + l] = s[24 + 6 * t + pair_t(k, l)];
    }
  }
  if (T) {
    long m = 0;
    for (int e = 0; e < NE; ++e)
      for (int i = 0; i < 12; ++i) {
        const int c = ctx->slot_col[12 * e + i];
        if (c == -1)
          T[m++] = ctx->h_tbc[24 * e + 12 + i];
        else if (c >= 0)
          T[m++] = x[c];
      }
  }
}

// Volume averages of stress and strain from the boundary solution (divergence theorem; perfect
// bonding, no body force) on the quadrature table, and the enclosed volume V = (1/3) int x . n dS.
static void fields_from_ut(const ibem3d_ctx* ctx, const double* U, const double* T, double* sigma,
                           double* eps, double* volume) {
  const double* qp = ctx->h_qp.data();
  const int* conn = ctx->h_conn.data();
  double S[3][3] = {}, E[3][3] = {}, V = 0.0;
  for (int e = 0; e < ctx->NE; ++e) {
    for (int g = 0; g < 64; ++g) {
      const double* Q = qp + ((size_t)e * 64 + g) * kQpStride;  // g[3], nv[3], N[4], w*jac
      double t[3] = {0, 0, 0}, u[3] = {0, 0, 0};
      for (int q = 0; q < 4; ++q) {
        const int nd = conn[4 * e + q] - 1;
        for (int i = 0; i < 3; ++i) {
          if (T) t[i] += Q[6 + q] * T[12L * e + 3 * q + i];
          if (U) u[i] += Q[6 + q] * U[3L * nd + i];
        }
      }
      const double w = Q[10];
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
          S[i][j] += w * t[i] * Q[j];
          E[i][j] += w * 0.5 * (u[i] * Q[3 + j] + u[j] * Q[3 + i]);
        }
      V += w * (Q[0] * Q[3] + Q[1] * Q[4] + Q[2] * Q[5]) / 3.0;
    }
  }
  static const int pi[6] = {0, 0, 0, 1, 1, 2}, pj[6] = {0, 1, 2, 1, 2, 2};
  for (int c = 0; c < 6; ++c) {
    if (sigma) sigma[c] = 0.5 * (S[pi[c]][pj[c]] + S[pj[c]][pi[c]]) / V;
    if (eps) eps[c] = E[pi[c]][pj[c]] / V;
  }
  if (volume) *volume = V;
}

// <e*_ij> = (1/V) sum_spheres V_s (E_ij + (a_s^2 / 5) sum_k E_ij,kk) from U (78 layout)
static void eps_star_from_u(const ibem3d_ctx* ctx, const double* U, double V, double* eps_star) {
  double acc[6] = {0, 0, 0, 0, 0, 0};
  for (int h = 0; h < ctx->num; ++h) {
    const double a = ctx->h_rad[h], Vs = 4.0 / 3.0 * kPi * a * a * a;
    const double* E = U + 3L * ctx->NN + 78L * h;  // E_t, E_t,k, E_t,kl (78 layout)
    for (int t = 0; t < 6; ++t) {
      const double tr = E[24 + 9 * t + 0] + E[24 + 9 * t + 4] + E[24 + 9 * t + 8];
      acc[t] += Vs * (E[t] + a * a / 5.0 * tr);
    }
  }
  for (int t = 0; t < 6; ++t) eps_star[t] = acc[t] / V;
}

static int host_solution(ibem3d_ctx* ctx, std::vector<double>& x) {
  x.resize(ctx->n);
  HIPC(hipMemcpyAsync(x.data(), ctx->d_x, sizeof(double) * ctx->n, hipMemcpyDeviceToHost, ctx->st));
  HIPC(hipStreamSynchronize(ctx->st));
  return 0;
}

int ibem3d_average_eigenstrain(ibem3d_ctx* ctx, double* eps_star) {
  CTX_CHECK();
  if (!ctx->solved) return fail(ctx, IBEM3D_ESTATE, "average_eigenstrain: no solution");
  if (!eps_star) return fail(ctx, IBEM3D_EINVAL, "average_eigenstrain: null output");
  if (int rc = host_geometry(ctx)) return rc;
  std::vector<double> x, U(3L * ctx->NN + 78L * ctx->num);
  if (int rc = host_solution(ctx, x)) return rc;
  ut_from_x(ctx, x.data(), U.data(), nullptr);
  double V = 0.0;
  fields_from_ut(ctx, nullptr, nullptr, nullptr, nullptr, &V);
  if (!(V > 0)) return fail(ctx, IBEM3D_EINVAL, "average_eigenstrain: non-positive enclosed volume");
  eps_star_from_u(ctx, U.data(), V, eps_star);
  return 0;
}

int ibem3d_average_fields(ibem3d_ctx* ctx, double* sigma, double* eps, double* volume) {
  CTX_CHECK();
  if (!ctx->solved) return fail(ctx, IBEM3D_ESTATE, "average_fields: no solution");
  if (int rc = host_geometry(ctx)) return rc;
  std::vector<double> x, U(3L * ctx->NN + 78L * ctx->num), T(12L * ctx->NE);
  if (int rc = host_solution(ctx, x)) return rc;
  ut_from_x(ctx, x.data(), U.data(), T.data());
  double V = 0.0;
  fields_from_ut(ctx, U.data(), T.data(), sigma, eps, &V);
  if (!(V > 0)) return fail(ctx, IBEM3D_EINVAL, "average_fields: non-positive enclosed volume");
  if (volume) *volume = V;
  return 0;
}

// getUTVector (BEMBuilderElastic3D.cpp:2053-2117) in the 78 layout
int ibem3d_get_ut(ibem3d_ctx* ctx, double* U, double* T, double* X60) {
  CTX_CHECK();
  if (!ctx->solved) return fail(ctx, IBEM3D_ESTATE, "get_ut: no solution");
  std::vector<double> x;
  if (int rc = host_solution(ctx, x)) return rc;
  if (X60) std::memcpy(X60, x.data(), sizeof(double) * ctx->n);
  ut_from_x(ctx, x.data(), U, T);
  return 0;
}

// ---- stresses at the postprocess points on the device -----------------------------------------
static int stress_alloc(ibem3d_ctx* ctx) {
  if (ctx->d_stU) return 0;
  if (int rc = dalloc(ctx, &ctx->d_stU, 3L * ctx->NN)) return rc;
  if (int rc = dalloc(ctx, &ctx->d_stT, 12L * ctx->NE)) return rc;
  if (int rc = dalloc(ctx, &ctx->d_stws, (long)stress_workspace_doubles(ctx->num, ctx->nump))) return rc;
  return dalloc(ctx, &ctx->d_sts, 6L * ctx->nump);
}

// the stresses of the current solution d_x into d_sts, on stream st
static void enqueue_stress(ibem3d_ctx* ctx, hipStream_t st) {
  const int NN = ctx->NN, NE = ctx->NE;
  launch_stress_inputs(ctx->d_x, ctx->d_node_col, ctx->d_uval, ctx->d_slot_col, ctx->d_tval, NN, NE,
                       ctx->d_stU, ctx->d_stT, st);
  launch_stress(ctx->d_qp, ctx->d_conn, ctx->d_nodes, NN, NE, ctx->d_cen, ctx->d_rad, ctx->num, ctx->d_pts,
                ctx->nump, ctx->M, ctx->d_stU, ctx->d_stT, ctx->d_x + 3L * NN, ctx->d_stws, ctx->d_sts, st);
}

int ibem3d_stress(ibem3d_ctx* ctx, double* s) {
  CTX_CHECK();
  if (!ctx->solved) return fail(ctx, IBEM3D_ESTATE, "stress: no solution");
  if (ctx->nump == 0) return 0;
  if (int rc = join_stress(ctx, false)) return rc;
  if (!ctx->stress_spec) {
    // U and T from the device solution (no host round trip), then the kernel
    if (int rc = stress_alloc(ctx)) return rc;
    {
      Scope sc(ctx, "post.stress");
      enqueue_stress(ctx, ctx->st);
    }
    if (int rc = after_launch(ctx, "k_stress")) return rc;
  }
  HIPC(hipMemcpyAsync(s, ctx->d_sts, sizeof(double) * 6 * ctx->nump, hipMemcpyDeviceToHost, ctx->st));
  HIPC(hipStreamSynchronize(ctx->st));
  return 0;
}

int ibem3d_timings(const ibem3d_ctx* ctx, char* names, int cap, double* ms, int max_entries) {
  if (!ctx) return IBEM3D_EINVAL;
  fold_times(const_cast<ibem3d_ctx*>(ctx));
  int cnt = 0, pos = 0;
  for (const auto& nm : ctx->folded_order) {
    if (cnt >= max_entries) break;
    if (names) {
      if (pos + (int)nm.size() + 1 > cap) break;
      std::memcpy(names + pos, nm.c_str(), nm.size() + 1);
      pos += (int)nm.size() + 1;
    }
    if (ms) ms[cnt] = ctx->folded_ms.at(nm);
    ++cnt;
  }
  return cnt;
}

void ibem3d_reset_timings(ibem3d_ctx* ctx) {
  if (!ctx) return;
  fold_times(ctx);
  for (auto& kv : ctx->open) ctx->pool.push_back(kv.second);
  ctx->open.clear();
  ctx->folded_order.clear();
  ctx->folded_ms.clear();
}

}  // extern "C"

// ================================================================ batched solves (batch.cpp) ====
// relative residual on the device: part[0..B) = block sums of r^2, part[B..2B) = of b^2 (fixed
// order), then out = sqrt(sum r^2 / sum b^2)
__global__ __launch_bounds__(256) void k_sumsq_part(const double* __restrict__ v, long n,
                                                    double* __restrict__ part) {
  __shared__ double red[256];
  double acc = 0.0;
  for (long t = (long)blockIdx.x * 256 + threadIdx.x; t < n; t += (long)kDotBlocks * 256) acc += v[t] * v[t];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}
__global__ void k_rel_finish(const double* __restrict__ part, double* __restrict__ out) {
  if (threadIdx.x != 0) return;
  double r = 0.0, b = 0.0;
  for (int i = 0; i < kDotBlocks; ++i) {
    r += part[i];
    b += part[kDotBlocks + i];
  }
  *out = b > 0.0 ? sqrt(r / b) : sqrt(r);
}

namespace ibem_batch {

hipStream_t stream(ibem3d_ctx* ctx) { return ctx->st; }
long size(ibem3d_ctx* ctx) { return ctx->n; }
size_t dot_blocks() { return kDotBlocks; }

int prepare(ibem3d_ctx* ctx, int mode) {
  CTX_CHECK();
  if (mode < IBEM3D_SOLVE_MIXED_FP32 || mode > IBEM3D_SOLVE_MIXED_BF16)
    return fail(ctx, IBEM3D_EINVAL, "batch: mixed modes only");
  const long n = ctx->n;
  if (int rc = ensure_solver_ws(ctx)) return rc;
  ctx->lda32 = (n + 63) / 64 * 64;
  if (!ctx->d_A32)
    if (int rc = dalloc(ctx, &ctx->d_A32, (size_t)ctx->lda32 * n)) return rc;
  if (!ctx->d_v32)
    if (int rc = dalloc(ctx, &ctx->d_v32, (size_t)n)) return rc;
  if (mode != IBEM3D_SOLVE_MIXED_FP32 && !ctx->d_f16ws)
    if (int rc = dalloc(ctx, &ctx->d_f16ws, lu_f16_workspace_doubles(n))) return rc;
  if (!ctx->d_bemws)
    if (int rc = dalloc(ctx, &ctx->d_bemws, bem_workspace_doubles(ctx->NN, ctx->NE))) return rc;
  if (lu_streams(ctx->W)) return fail(ctx, IBEM3D_EHIP, "batch: LU streams");
  if (int rc = host_geometry(ctx)) return rc;
  HIPC(hipStreamSynchronize(ctx->st));
  return 0;
}

int enqueue(ibem3d_ctx* ctx, int mode, int K, int nloads, const double* d_uvals, double* d_xs,
            double* d_rel, double* d_part) {
  const long n = ctx->n, NN = ctx->NN;
  hipStream_t st = ctx->st;
  const int prod = mode == IBEM3D_SOLVE_MIXED_FP16_1   ? kProdF16
                   : mode == IBEM3D_SOLVE_MIXED_BF16   ? kProdBF16
                   : mode == IBEM3D_SOLVE_MIXED_BF16X3 ? kProdBF16x3
                                                       : kProdF16x3;
  const bool split = mode != IBEM3D_SOLVE_MIXED_FP32;
  HIPC(hipMemcpyAsync(ctx->d_uval, d_uvals, sizeof(double) * 3 * NN, hipMemcpyDeviceToDevice, st));
  // assembly (every phase, all orders at once)
  zero_boundary_columns(ctx, st);
  fill_zero(ctx->d_b, sizeof(double) * n, st);
  if (int rc = bem_enqueue(ctx, st, true)) return rc;
  if (ctx->num > 0 && !rows_moment_form(ctx) && ctx->bem_gather) {  // the per-point rows accumulate
    const long nr = 60L * ctx->num;
    hipLaunchKernelGGL(k_zero_rows_mask, dim3((unsigned)((nr + 255) / 256), (unsigned)(3 * NN)), dim3(256), 0, st,
                       ctx->d_A, ctx->lda, 3L * NN, nr, 7);
  }
  if (ctx->num > 0) {
    launch_field(ctx->d_nodes, ctx->NN, ctx->d_cen, ctx->d_rad, ctx->num, ctx->lda, 3 * NN, ctx->M, ctx->d_A, st);
    launch_incl(ctx->d_cen, ctx->d_rad, ctx->num, ctx->lda, 3 * NN, ctx->M, 7, ctx->d_wdesc, ctx->d_A, st);
    launch_rows(ctx->d_qp, ctx->d_conn, ctx->d_cen, ctx->num, ctx->NE, ctx->lda, 3 * NN, ctx->M, 7, ctx->d_node_col,
                ctx->d_slot_col, ctx->d_uval, ctx->d_tval, ctx->d_A, ctx->d_b, st, true, rows_moment_form(ctx));
  }
  // mixed LU of R A C
  if (equilibrate_copy_f32(ctx->d_A, n, ctx->lda, ctx->d_r, ctx->d_c, ctx->d_A32, ctx->lda32, ctx->d_eqws, st, &ctx->W))
    return fail(ctx, IBEM3D_EHIP, "batch: equilibration");
  if (lu_factor_f32(ctx->d_A32, n, ctx->lda32, ctx->W, st, nullptr, nullptr, split ? ctx->d_f16ws : nullptr, true,
                    prod))
    return fail(ctx, IBEM3D_EHIP, "batch: LU");
  if (lu_prepare_solve_f32(ctx->d_A32, n, ctx->lda32, ctx->W, st)) return fail(ctx, IBEM3D_EHIP, "batch: prepare");
  for (int t = 0; t < nloads; ++t) {
    if (t > 0) {
      HIPC(hipMemcpyAsync(ctx->d_uval, d_uvals + (size_t)t * 3 * NN, sizeof(double) * 3 * NN,
                          hipMemcpyDeviceToDevice, st));
      fill_zero(ctx->d_b, sizeof(double) * n, st);
      if (int rc = bem_enqueue(ctx, st, false)) return rc;
      if (ctx->num > 0)
        launch_rows(ctx->d_qp, ctx->d_conn, ctx->d_cen, ctx->num, ctx->NE, ctx->lda, 3 * NN, ctx->M, 7,
                    ctx->d_node_col, ctx->d_slot_col, ctx->d_uval, ctx->d_tval, ctx->d_A, ctx->d_b, st, false,
                    rows_moment_form(ctx));
    }
    // K steps of classic refinement from x = 0: x <- x + C LU^-1 R (b - A x)
    fill_zero(ctx->d_x, sizeof(double) * n, st);
    HIPC(hipMemcpyAsync(ctx->d_res, ctx->d_b, sizeof(double) * n, hipMemcpyDeviceToDevice, st));
    for (int it = 0; it < K; ++it) {
      hipLaunchKernelGGL(k_scale_to_f32, vgrid(n), dim3(256), 0, st, ctx->d_v32, ctx->d_res, ctx->d_r, n);
      lu_solve_f32(ctx->d_A32, n, ctx->lda32, ctx->W, ctx->d_v32, st);
      hipLaunchKernelGGL(k_axpy_scaled_f32, vgrid(n), dim3(256), 0, st, ctx->d_x, ctx->d_v32, ctx->d_c, n);
      HIPC(hipMemcpyAsync(ctx->d_res, ctx->d_b, sizeof(double) * n, hipMemcpyDeviceToDevice, st));
      gemv_f64(n, n, -1.0, ctx->d_A, ctx->lda, ctx->d_x, 1.0, ctx->d_res, ctx->d_gemvws, st);
    }
    HIPC(hipMemcpyAsync(d_xs + (size_t)t * n, ctx->d_x, sizeof(double) * n, hipMemcpyDeviceToDevice, st));
    hipLaunchKernelGGL(k_sumsq_part, dim3(kDotBlocks), dim3(256), 0, st, ctx->d_res, n, d_part);
    hipLaunchKernelGGL(k_sumsq_part, dim3(kDotBlocks), dim3(256), 0, st, ctx->d_b, n, d_part + kDotBlocks);
    hipLaunchKernelGGL(k_rel_finish, dim3(1), dim3(64), 0, st, d_part, d_rel + t);
  }
  return after_launch(ctx, "batch enqueue");
}

void mark_solved(ibem3d_ctx* ctx, int mode) {
  ctx->phases = 0xFFu;
  ctx->a_factored = false;
  ctx->prepared_mode = mode;
  ctx->solved = true;
}

int averages(ibem3d_ctx* ctx, const double* x, const double* uvals, double* sigma, double* eps,
             double* eps_star, double* V) {
  CTX_CHECK();
  if (int rc = host_geometry(ctx)) return rc;
  for (long j = 0; j < 3L * ctx->NN; ++j) ctx->h_bc[2 * j + 1] = uvals[j];
  std::vector<double> U(3L * ctx->NN + 78L * ctx->num), T(12L * ctx->NE);
  ut_from_x(ctx, x, U.data(), T.data());
  double vol = 0.0;
  fields_from_ut(ctx, U.data(), T.data(), sigma, eps, &vol);
  if (!(vol > 0)) return fail(ctx, IBEM3D_EINVAL, "averages: non-positive enclosed volume");
  eps_star_from_u(ctx, U.data(), vol, eps_star);
  if (V) *V = vol;
  return 0;
}

}  // namespace ibem_batch
