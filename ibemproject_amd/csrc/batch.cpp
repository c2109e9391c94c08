// Batched RVE stiffness (include/ibem3d.h, ibem3d_batch_*): many independent RVEs, each with
// displacements prescribed on its whole boundary, solved for the six unit macro strains on one
// factorisation each. RVEs are grouped into memory waves; a wave's solves (assembly, mixed LU,
// six right-hand sides with fixed-count refinement) are captured from one root stream as parallel
// branches of a single graph and run by one graph launch. Loads whose refinement did not reach the
// tolerance inside the graph are refined further on the host (classic refinement, then GMRES-IR,
// on the same factors). No counterpart in the reference, which solves one RVE per process
// (configElastic3D.cpp:27-30).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <thread>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/ibem3d.h"
#include "batch_internal.h"

namespace {

std::string g_batch_error;

struct Rve {
  int NN, NE, num;
  double E0, nu0, E1, nu1;
  std::vector<double> nodes, bc, tbc, centres, radii;
  std::vector<int> conn;
  long n;
};

double now() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// displacement values of the unit macro strain t (pair order 00 01 02 11 12 22, e_ij = e_ji = 1):
// u = e_t . x at every node
void unit_load(const Rve& r, int t, double* u) {
  static const int pi[6] = {0, 0, 0, 1, 1, 2}, pj[6] = {0, 1, 2, 1, 2, 2};
  const int i = pi[t], j = pj[t];
  for (int a = 0; a < r.NN; ++a) {
    const double* x = r.nodes.data() + 3L * a;
    double* ua = u + 3L * a;
    ua[0] = ua[1] = ua[2] = 0.0;
    ua[i] += x[j];
    if (i != j) ua[j] += x[i];
  }
}

// s = C0 : e (isotropic, E and nu), components in pair order
void isotropic_stress(double E, double nu, const double* e, double* s) {
  const double lam = E * nu / ((1 + nu) * (1 - 2 * nu)), mu = E / (2 * (1 + nu));
  const double tr = e[0] + e[3] + e[5];
  static const bool diag[6] = {true, false, false, true, false, true};
  for (int t = 0; t < 6; ++t) s[t] = (diag[t] ? lam * tr : 0.0) + 2 * mu * e[t];
}

// device bytes a context of n unknowns holds for a batched mixed solve (A, the FP32 factor, the split
// workspace, solver workspaces, the six solutions), used to size the memory waves
double rve_bytes(long n, int mode) {
  const double lda = (double)((n + 63) / 64 * 64);  // A and the FP32 factor: 64-aligned columns
  double b = 8.0 * lda * n + 4.0 * lda * n;
  if (mode != IBEM3D_SOLVE_MIXED_FP32) b += 2.0 * (4.0 * n * 1024 * 2 + 8.0 * n);
  b += 8.0 * 1024 * n + 2.0 * 8 * 128 * n + 8.0 * 32 * n * 2 + 8.0 * 6 * n + 64.0 * n;
  return b + 64.0 * (1 << 20);  // boundary quadrature workspace and small buffers
}

}  // namespace

struct ibem3d_batch {
  int dev = 0;
  std::vector<Rve> rves;
};

extern "C" {

const char* ibem3d_batch_last_error(void) { return g_batch_error.c_str(); }

ibem3d_batch* ibem3d_batch_create(const ibem3d_problem* probs, int count, int device) {
  if (!probs || count <= 0) {
    g_batch_error = "batch_create: no problems";
    return nullptr;
  }
  ibem3d_batch* b = new ibem3d_batch();
  b->dev = device;
  b->rves.resize(count);
  for (int k = 0; k < count; ++k) {
    const ibem3d_problem& p = probs[k];
    if (p.NN <= 0 || p.NE <= 0 || p.num < 0 || !p.nodes || !p.conn || !p.bc || !p.tbc ||
        (p.num > 0 && (!p.centres || !p.radii))) {
      g_batch_error = "batch_create: invalid problem " + std::to_string(k);
      delete b;
      return nullptr;
    }
    // the stiffness loads prescribe every boundary displacement: all displacements known, all
    // tractions unknown
    for (long j = 0; j < 3L * p.NN; ++j)
      if (p.bc[2 * j] != 0.0) {
        g_batch_error = "batch_create: problem " + std::to_string(k) + " has an unknown displacement";
        delete b;
        return nullptr;
      }
    for (long e = 0; e < p.NE; ++e)
      for (int i = 0; i < 12; ++i)
        if (p.tbc[24 * e + i] != 0.0) {
          g_batch_error = "batch_create: problem " + std::to_string(k) + " has a known traction";
          delete b;
          return nullptr;
        }
    Rve& r = b->rves[k];
    r.NN = p.NN;
    r.NE = p.NE;
    r.num = p.num;
    r.E0 = p.E0;
    r.nu0 = p.nu0;
    r.E1 = p.E1;
    r.nu1 = p.nu1;
    r.nodes.assign(p.nodes, p.nodes + 3L * p.NN);
    r.conn.assign(p.conn, p.conn + 4L * p.NE);
    r.bc.assign(p.bc, p.bc + 6L * p.NN);
    r.tbc.assign(p.tbc, p.tbc + 24L * p.NE);
    r.centres.assign(p.centres, p.centres + 3L * p.num);
    r.radii.assign(p.radii, p.radii + p.num);
    r.n = 3L * p.NN + 60L * p.num;
  }
  return b;
}

void ibem3d_batch_destroy(ibem3d_batch* b) { delete b; }

int ibem3d_batch_stiffness(ibem3d_batch* b, const ibem3d_batch_options* opt, double tol, double* C,
                           double* C_bnd, double* rel, int* host_refined, double* times, double* eig) {
  if (!b || !opt || !C) {
    g_batch_error = "batch_stiffness: null argument";
    return IBEM3D_EINVAL;
  }
  const int mode = opt->mode;
  if (mode < IBEM3D_SOLVE_MIXED_FP32 || mode > IBEM3D_SOLVE_MIXED_BF16 || opt->refine_steps < 1 ||
      opt->wave_max < 1 || !(opt->mem_budget_gb > 0)) {
    g_batch_error = "batch_stiffness: invalid options";
    return IBEM3D_EINVAL;
  }
  if (!(tol > 0)) tol = 1e-10;
  if (hipSetDevice(b->dev) != hipSuccess) {
    g_batch_error = "batch_stiffness: hipSetDevice";
    return IBEM3D_EHIP;
  }
  const int count = (int)b->rves.size();
  const double budget = opt->mem_budget_gb * 1e9;
  double t_capture = 0, t_inst = 0, t_run = 0, t_host = 0, t_setup = 0;
  int waves = 0;
  hipStream_t root = nullptr;
  if (hipStreamCreateWithFlags(&root, hipStreamNonBlocking) != hipSuccess) {
    g_batch_error = "batch_stiffness: stream";
    return IBEM3D_EHIP;
  }
  int rc = 0;
  const size_t nblk = ibem_batch::dot_blocks();
  std::vector<long> eig_off(count + 1, 0);  // each RVE's eigenstrain block: 6 loads x 60 num
  for (int k = 0; k < count; ++k) eig_off[k + 1] = eig_off[k] + 6L * 60 * b->rves[k].num;
  for (int k0 = 0; k0 < count && rc == 0;) {
    // the wave: consecutive RVEs while they fit the budget (at least one)
    int k1 = k0;
    double used = 0;
    while (k1 < count && k1 - k0 < opt->wave_max) {
      const double need = rve_bytes(b->rves[k1].n, mode);
      if (k1 > k0 && used + need > budget) break;
      used += need;
      ++k1;
    }
    ++waves;
    const double ts = now();
    const int W = k1 - k0;
    std::vector<ibem3d_ctx*> ctx(W, nullptr);
    std::vector<double*> d_u(W, nullptr), d_x(W, nullptr), d_rel(W, nullptr), d_part(W, nullptr);
    std::vector<hipEvent_t> ev(W, nullptr);
    hipEvent_t e0 = nullptr;
    auto cleanup = [&]() {
      for (int w = 0; w < W; ++w) {
        if (ctx[w]) ibem3d_destroy(ctx[w]);
        hipFree(d_u[w]);
        hipFree(d_x[w]);
        hipFree(d_rel[w]);
        hipFree(d_part[w]);
        if (ev[w]) hipEventDestroy(ev[w]);
      }
      if (e0) hipEventDestroy(e0);
    };
    for (int w = 0; w < W && rc == 0; ++w) {
      const Rve& r = b->rves[k0 + w];
      ibem3d_problem p{r.NN, r.NE, r.num, 0, r.E0, r.nu0, r.E1, r.nu1, r.nodes.data(), r.conn.data(),
                       r.bc.data(), r.tbc.data(), r.centres.data(), r.radii.data(), nullptr};
      ctx[w] = ibem3d_create(&p, b->dev);
      if (!ctx[w]) {
        g_batch_error = std::string("batch: create: ") + ibem3d_last_error(nullptr);
        rc = IBEM3D_EHIP;
        break;
      }
      if ((rc = ibem_batch::prepare(ctx[w], mode))) {
        g_batch_error = std::string("batch: prepare: ") + ibem3d_last_error(ctx[w]);
        break;
      }
      std::vector<double> u(6 * 3L * r.NN);
      for (int t = 0; t < 6; ++t) unit_load(r, t, u.data() + 3L * r.NN * t);
      if (hipMalloc((void**)&d_u[w], sizeof(double) * u.size()) != hipSuccess ||
          hipMalloc((void**)&d_x[w], sizeof(double) * 6 * r.n) != hipSuccess ||
          hipMalloc((void**)&d_rel[w], sizeof(double) * 6) != hipSuccess ||
          hipMalloc((void**)&d_part[w], sizeof(double) * 2 * nblk) != hipSuccess ||
          hipMemcpy(d_u[w], u.data(), sizeof(double) * u.size(), hipMemcpyHostToDevice) != hipSuccess ||
          hipEventCreateWithFlags(&ev[w], hipEventDisableTiming) != hipSuccess) {
        g_batch_error = "batch: device buffers";
        rc = IBEM3D_EHIP;
      }
    }
    if (rc == 0 && hipEventCreateWithFlags(&e0, hipEventDisableTiming) != hipSuccess) rc = IBEM3D_EHIP;
    if (rc) {
      cleanup();
      break;
    }
    const double tc = now();
    t_setup += tc - ts;
    const int K = opt->refine_steps;
    if (opt->use_graph) {
      // one graph: fork from the root stream into every context's streams, join back
      hipGraph_t graph = nullptr;
      hipGraphExec_t exec = nullptr;
      bool ok = hipStreamBeginCapture(root, hipStreamCaptureModeRelaxed) == hipSuccess &&
                hipEventRecord(e0, root) == hipSuccess;
      for (int w = 0; w < W && ok; ++w) {
        hipStream_t s = ibem_batch::stream(ctx[w]);
        ok = hipStreamWaitEvent(s, e0, 0) == hipSuccess &&
             ibem_batch::enqueue(ctx[w], mode, K, 6, d_u[w], d_x[w], d_rel[w], d_part[w]) == 0 &&
             hipEventRecord(ev[w], s) == hipSuccess && hipStreamWaitEvent(root, ev[w], 0) == hipSuccess;
      }
      hipError_t ec = hipStreamEndCapture(root, &graph);
      ok = ok && ec == hipSuccess && graph;
      const double ti = now();
      t_capture += ti - tc;
      ok = ok && hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) == hipSuccess;
      const double tl = now();
      t_inst += tl - ti;
      ok = ok && hipGraphLaunch(exec, root) == hipSuccess && hipStreamSynchronize(root) == hipSuccess;
      t_run += now() - tl;
      if (exec) hipGraphExecDestroy(exec);
      if (graph) hipGraphDestroy(graph);
      if (!ok) {
        g_batch_error = std::string("batch: graph capture / launch failed: ") + hipGetErrorString(hipGetLastError());
        rc = IBEM3D_EHIP;
      }
    } else {
      // the same work launched directly, every context on its own streams; with host_threads > 1
      // the contexts are enqueued from that many host threads (the launch rate of one thread
      // bounds the small RVEs' latency-bound solves)
      std::atomic<int> next{0}, bad{0};
      auto worker = [&]() {
        if (hipSetDevice(b->dev) != hipSuccess) {
          bad = 1;
          return;
        }
        for (int w; (w = next++) < W;) {
          if (ibem_batch::enqueue(ctx[w], mode, K, 6, d_u[w], d_x[w], d_rel[w], d_part[w]) != 0) bad = 1;
        }
      };
      const int nt = std::max(1, std::min(opt->host_threads, W));
      if (nt == 1) {
        worker();
      } else {
        std::vector<std::thread> th;
        for (int i = 0; i < nt; ++i) th.emplace_back(worker);
        for (auto& t : th) t.join();
      }
      bool ok = !bad;
      for (int w = 0; w < W; ++w) ok = hipStreamSynchronize(ibem_batch::stream(ctx[w])) == hipSuccess && ok;
      t_run += now() - tc;
      if (!ok) {
        g_batch_error = "batch: launch failed";
        rc = IBEM3D_EHIP;
      }
    }
    // results: averages per load; loads above the tolerance refined on the host
    const double th = now();
    for (int w = 0; w < W && rc == 0; ++w) {
      const int k = k0 + w;
      const Rve& r = b->rves[k];
      ibem_batch::mark_solved(ctx[w], mode);
      std::vector<double> xs(6 * r.n), rl(6), u(3L * r.NN), tz(12L * r.NE, 0.0);
      if (hipMemcpy(xs.data(), d_x[w], sizeof(double) * xs.size(), hipMemcpyDeviceToHost) != hipSuccess ||
          hipMemcpy(rl.data(), d_rel[w], sizeof(double) * 6, hipMemcpyDeviceToHost) != hipSuccess) {
        rc = IBEM3D_EHIP;
        break;
      }
      for (int t = 0; t < 6 && rc == 0; ++t) {
        double* x = xs.data() + (size_t)t * r.n;
        int refined = 0;
        unit_load(r, t, u.data());
        if (!(rl[t] <= tol)) {
          if ((rc = ibem3d_update_known_values(ctx[w], u.data(), tz.data())) ||
              (rc = ibem3d_solve(ctx[w], mode | IBEM3D_SOLVE_REUSE, tol)) ||
              (rc = ibem3d_get_ut(ctx[w], nullptr, nullptr, x)) ||
              (rc = ibem3d_solve_info(ctx[w], &rl[t], nullptr))) {
            g_batch_error = std::string("batch: host refinement: ") + ibem3d_last_error(ctx[w]);
            break;
          }
          refined = 1;
        }
        double sig[6], eps[6], est[6], d[6], V = 0;
        if ((rc = ibem_batch::averages(ctx[w], x, u.data(), sig, eps, est, &V))) break;
        for (int c = 0; c < 6; ++c) d[c] = eps[c] - est[c];
        double s_inc[6];
        isotropic_stress(r.E0, r.nu0, d, s_inc);
        for (int c = 0; c < 6; ++c) {
          C[36L * k + 6 * c + t] = s_inc[c];
          if (C_bnd) C_bnd[36L * k + 6 * c + t] = sig[c];
        }
        if (rel) rel[6L * k + t] = rl[t];
        if (host_refined) host_refined[6L * k + t] = refined;
        if (eig)
          std::memcpy(eig + eig_off[k] + 60L * r.num * t, x + 3L * r.NN, sizeof(double) * 60 * r.num);
      }
    }
    t_host += now() - th;
    cleanup();
    k0 = k1;
  }
  hipStreamDestroy(root);
  if (times) {
    times[0] = t_setup;
    times[1] = t_capture;
    times[2] = t_inst;
    times[3] = t_run;
    times[4] = t_host;
    times[5] = waves;
  }
  return rc;
}

}  // extern "C"
