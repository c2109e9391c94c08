// ibem3d_run -- native driver for one elastic-3D spherical-inclusion problem directory.
//
// Does what the reference's ibemRunner does for calculation type 1 (ibemRunner.cpp:37-120,
// serviceWrapper.h:53-77): preprocess (read position.txt, BEM.txt, postprocess.txt), iBEM system
// (addBEM, addFieldToBEM, the three inclusion orders, the three boundary orders:
// iBEMSystemService.h:58-67), system solver (ApplyBC, solve, getUTVector: systemSolverService.h:
// 41-45), postprocess (getFlux -> stress.txt: postProcessorElastic3D.cpp:525-587) -- on one GPU
// through the C ABI.
//
//   ibem3d_run [dir] [--out DIR] [--mode fp64|mixed|mixed16|fp16_1|bf16|bj|bj16] [--tol T] [--device D] [--full] [--timings]
//
// Writes DIR/stress.txt in the reference's format; --full also writes stress_full.txt (%.17g).
// Exit status 0 on success, 1 on a usage or input error, 2 on an engine error.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/ibem3d.h"

namespace {

int usage() {
  std::fprintf(stderr,
               "usage: ibem3d_run [dir] [--out DIR] [--mode fp64|mixed|mixed16|fp16_1|bf16|bj|bj16] [--tol T] [--device D] "
               "[--full] [--timings]\n");
  return 1;
}

bool read_calc_type(const std::string& dir, int& id) {
  FILE* f = std::fopen((dir + "/calculationType.txt").c_str(), "r");
  if (!f) return false;
  const bool ok = std::fscanf(f, "%d", &id) == 1;
  std::fclose(f);
  return ok;
}

}  // namespace

int main(int argc, char** argv) {
  std::string dir = ".", out;
  int mode = IBEM3D_SOLVE_MIXED_FP32, device = 0;
  double tol = 1e-10;
  bool full = false, timings = false;
  for (int i = 1; i < argc; ++i) {
    const std::string a = argv[i];
    auto next = [&](const char* what) -> const char* {
      if (i + 1 >= argc) {
        std::fprintf(stderr, "ibem3d_run: %s needs a value\n", what);
        std::exit(usage());
      }
      return argv[++i];
    };
    if (a == "--out") {
      out = next("--out");
    } else if (a == "--mode") {
      const std::string m = next("--mode");
      if (m == "fp64")
        mode = IBEM3D_SOLVE_FP64_LU;
      else if (m == "mixed")
        mode = IBEM3D_SOLVE_MIXED_FP32;
      else if (m == "mixed16")
        mode = IBEM3D_SOLVE_MIXED_FP16;
      else if (m == "fp16_1")
        mode = IBEM3D_SOLVE_MIXED_FP16_1;
      else if (m == "bf16")
        mode = IBEM3D_SOLVE_MIXED_BF16;
      else if (m == "bj")
        mode = IBEM3D_SOLVE_BJ_GMRES;
      else if (m == "bj16")
        mode = IBEM3D_SOLVE_BJ_GMRES16;
      else
        return usage();
    } else if (a == "--tol") {
      tol = std::atof(next("--tol"));
    } else if (a == "--device") {
      device = std::atoi(next("--device"));
    } else if (a == "--full") {
      full = true;
    } else if (a == "--timings") {
      timings = true;
    } else if (!a.empty() && a[0] == '-') {
      return usage();
    } else {
      dir = a;
    }
  }
  if (out.empty()) out = dir;

  int id = 1;
  if (read_calc_type(dir, id) && id != 1) {
    std::fprintf(stderr, "ibem3d_run: calculationType %d is not the elastic-3D sphere path (1)\n", id);
    return 1;
  }
  const auto t0 = std::chrono::steady_clock::now();
  ibem3d_inputs* in = ibem3d_read_inputs(dir.c_str());
  if (!in) {
    std::fprintf(stderr, "ibem3d_run: %s\n", ibem3d_io_error());
    return 1;
  }
  const ibem3d_problem* p = ibem3d_inputs_problem(in);
  std::fprintf(stderr, "Title: %s | NN %d NE %d spheres %d points %d\n", ibem3d_inputs_title(in), p->NN,
               p->NE, p->num, p->nump);
  ibem3d_ctx* ctx = ibem3d_create(p, device);
  if (!ctx) {
    std::fprintf(stderr, "ibem3d_run: %s\n", ibem3d_last_error(nullptr));
    ibem3d_free_inputs(in);
    return 2;
  }
  const int nump = p->nump;
  const long n = ibem3d_size(ctx);
  std::vector<double> s(6 * (size_t)nump);
  int rc = ibem3d_add_bem(ctx);
  if (!rc) rc = ibem3d_add_field(ctx);
  for (int o = 0; o < 3 && !rc; ++o) rc = ibem3d_add_incl(ctx, o);
  for (int o = 0; o < 3 && !rc; ++o) rc = ibem3d_add_bem_rows(ctx, o);
  if (!rc) rc = ibem3d_apply_bc(ctx);
  if (!rc) rc = ibem3d_solve(ctx, mode, tol);
  if (!rc && nump > 0) rc = ibem3d_stress(ctx, s.data());
  if (rc) {
    std::fprintf(stderr, "ibem3d_run: %s\n", ibem3d_last_error(ctx));
    ibem3d_destroy(ctx);
    ibem3d_free_inputs(in);
    return 2;
  }
  double rel = 0.0;
  int its = 0;
  ibem3d_solve_info(ctx, &rel, &its);
  const auto t1 = std::chrono::steady_clock::now();
  std::fprintf(stderr, "n = %ld, relative residual %.3e (%d refinement iterations), %.3f s\n", n, rel,
               its, std::chrono::duration<double>(t1 - t0).count());
  if (timings) {
    std::vector<char> names(8192);
    std::vector<double> ms(128);
    const int k = ibem3d_timings(ctx, names.data(), (int)names.size(), ms.data(), (int)ms.size());
    const char* q = names.data();
    for (int i = 0; i < k; ++i) {
      std::fprintf(stderr, "  %-22s %10.3f ms\n", q, ms[i]);
      q += std::strlen(q) + 1;
    }
  }
  int wrc = ibem3d_write_stress((out + "/stress.txt").c_str(), s.data(), nump, 0);
  if (!wrc && full) wrc = ibem3d_write_stress((out + "/stress_full.txt").c_str(), s.data(), nump, 1);
  if (wrc) std::fprintf(stderr, "ibem3d_run: %s\n", ibem3d_io_error());
  ibem3d_destroy(ctx);
  ibem3d_free_inputs(in);
  return wrc ? 1 : 0;
}
