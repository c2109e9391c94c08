// Reference-side plugins that route iBEMProject's elastic-3D spherical-inclusion phases to the
// MI355X engine (libibem3d.so, include/ibem3d.h).
//
// This header is meant to be compiled INSIDE the reference's source tree (it includes the
// reference's own interface headers: Config.h, BEMBuilder.h, integrator.h, systemSolver.h,
// postProcessor.h, configElastic3D.h). It replaces, one for one:
//   configElastic3D        (configElastic3D.h:41-146)      -> configElastic3DGPU
//   BEMBuilderElastic3D    (BEMBuilder.h:27-35)             -> BEMBuilderElastic3DGPU
//   integratorElastic3D    (integrator.h:28-34)             -> integratorElastic3DGPU
//   systemSolverElastic3D  (systemSolver.h:28-31)           -> systemSolverElastic3DGPU
//   postProcessorElastic3D (postProcessor.h:26-30)          -> postProcessorElastic3DGPU
// and is registered by name in classFinder.h / inputGenerator.cpp.
//
// Because serviceWrapper re-creates the plugin objects for every service (serviceWrapper.h:40-46),
// the engine context lives on the config object. The config reads the same three text files but
// does not allocate the dense HMAT / GMAT / A the reference allocates at configElastic3D.cpp:95-101
// (33 GB each at 1000 spheres); the system lives only on the GPU.
//
// Errors: the reference's plugins return void and report nothing; these throw std::runtime_error
// with the engine's message, which the reference's main() turns into a terminate with the message.
#pragma once

#include <cstdlib>
#include <stdexcept>
#include <string>
#include <vector>

#include "BEMBuilder.h"
#include "Config.h"
#include "configElastic3D.h"
#include "ibem3d.h"
#include "integrator.h"
#include "postProcessor.h"
#include "systemSolver.h"

class configElastic3DGPU : public configElastic3D {
 public:
  ibem3d_inputs* inputs = nullptr;
  ibem3d_ctx* ctx = nullptr;
  int device = 0;
  int solveMode = IBEM3D_SOLVE_MIXED_FP32;  // IBEM3D_SOLVE_FP64_LU for the plain FP64 LU
  double tolerance = 1e-10;                 // relative residual of the refined solve

  configElastic3DGPU() {
    if (const char* m = std::getenv("IBEM3D_SOLVE_MODE")) solveMode = std::atoi(m);
    if (const char* d = std::getenv("IBEM3D_DEVICE")) device = std::atoi(d);
  }
  ~configElastic3DGPU() {
    if (ctx) ibem3d_destroy(ctx);
    if (inputs) ibem3d_free_inputs(inputs);
  }
  configElastic3DGPU(const configElastic3DGPU&) = delete;
  configElastic3DGPU& operator=(const configElastic3DGPU&) = delete;

  // Config::setUp (Config.h:36-42) reads the five inputs in order; here one native reader does all
  // of them and the engine context is created on the GPU.
  void setUp() override {
    inputs = ibem3d_read_inputs(".");
    if (!inputs) throw std::runtime_error(ibem3d_io_error());
    const ibem3d_problem* p = ibem3d_inputs_problem(inputs);
    NN = p->NN;
    NE = p->NE;
    num = p->num;
    nump = p->nump;
    EE = p->E0;
    XNU = p->nu0;
    eigen_point = MatrixXd(num, 3);
    radius = VectorXd(num);
    for (int k = 0; k < num; ++k) {
      for (int j = 0; j < 3; ++j) eigen_point(k, j) = p->centres[3 * k + j];
      radius(k) = p->radii[k];
    }
    x_p = MatrixXd(nump, 3);
    for (int k = 0; k < nump; ++k)
      for (int j = 0; j < 3; ++j) x_p(k, j) = p->points[3 * k + j];
    U = VectorXd::Zero(3 * NN + 78 * num);
    T = VectorXd::Zero(12 * NE);
    ctx = ibem3d_create(p, device);
    if (!ctx) throw std::runtime_error(ibem3d_last_error(nullptr));
  }

  string getbembuilder() override { return "BEMBuilderElastic3DGPU"; }
  string getintegrator() override { return "integratorElastic3DGPU"; }
  string getsystemsolver() override { return "systemSolverElastic3DGPU"; }
  string getpostprocessor() override { return "postProcessorElastic3DGPU"; }

  static configElastic3DGPU& of(Config& c) { return dynamic_cast<configElastic3DGPU&>(c); }
  void check(int rc) const {
    if (rc != IBEM3D_OK) throw std::runtime_error(ibem3d_last_error(ctx));
  }
};

// BEMBuilderElastic3D.cpp:1355 (addBEM), :1511 / :1629 / :1743 (flux-equivalent rows),
// :1920 (ApplyBC), :2053 (getUTVector)
class BEMBuilderElastic3DGPU : public BEMBuilder {
 public:
  void addBEM(Config& c) override { auto& g = configElastic3DGPU::of(c); g.check(ibem3d_add_bem(g.ctx)); }
  void addFluxEquivalentBEM(Config& c) override {
    auto& g = configElastic3DGPU::of(c);
    g.check(ibem3d_add_bem_rows(g.ctx, 0));
  }
  void addFluxEquivalentFirstOrderBEM(Config& c) override {
    auto& g = configElastic3DGPU::of(c);
    g.check(ibem3d_add_bem_rows(g.ctx, 1));
  }
  void addFluxEquivalentSecondOrderBEM(Config& c) override {
    auto& g = configElastic3DGPU::of(c);
    g.check(ibem3d_add_bem_rows(g.ctx, 2));
  }
  void ApplyBC(Config& c) override { auto& g = configElastic3DGPU::of(c); g.check(ibem3d_apply_bc(g.ctx)); }
  void getUTVector(Config& c) override {
    auto& g = configElastic3DGPU::of(c);
    g.check(ibem3d_get_ut(g.ctx, g.U.data(), g.T.data(), nullptr));
  }
};

// integratorElastic3D.cpp:135 (addFieldToBEM), :341 / :455 / :568 (inclusion rows)
class integratorElastic3DGPU : public integrator {
 public:
  void addFieldToBEM(Config& c) override { auto& g = configElastic3DGPU::of(c); g.check(ibem3d_add_field(g.ctx)); }
  void addFluxEquivalentInclusion(Config& c) override {
    auto& g = configElastic3DGPU::of(c);
    g.check(ibem3d_add_incl(g.ctx, 0));
  }
  void addFluxEquivalentFirstOrderInclusion(Config& c) override {
    auto& g = configElastic3DGPU::of(c);
    g.check(ibem3d_add_incl(g.ctx, 1));
  }
  void addFluxEquivalentSecondOrderInclusion(Config& c) override {
    auto& g = configElastic3DGPU::of(c);
    g.check(ibem3d_add_incl(g.ctx, 2));
  }
};

// systemSolverElastic3D.h:44-60
class systemSolverElastic3DGPU : public systemSolver {
 public:
  void solve(Config& c) override {
    auto& g = configElastic3DGPU::of(c);
    g.check(ibem3d_solve(g.ctx, g.solveMode, g.tolerance));
  }
};

// postProcessorElastic3D.cpp:525-587: stresses at the points, config.s_pt, stress.txt
class postProcessorElastic3DGPU : public postProcessor {
 public:
  void getFlux(Config& c) override {
    auto& g = configElastic3DGPU::of(c);
    std::vector<double> s(6 * (size_t)g.nump);
    if (g.nump > 0) g.check(ibem3d_stress(g.ctx, s.data()));
    g.s_pt = new double*[g.nump];
    for (int k = 0; k < g.nump; ++k) {
      g.s_pt[k] = new double[6];
      for (int j = 0; j < 6; ++j) g.s_pt[k][j] = s[6 * k + j];
    }
    if (ibem3d_write_stress("stress.txt", s.data(), g.nump, 0) != IBEM3D_OK)
      throw std::runtime_error(ibem3d_io_error());
  }
};
