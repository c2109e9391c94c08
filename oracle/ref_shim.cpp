// Test infrastructure only: a C entry layer around the reference iBEM sources, compiled from where
// they lie (see oracle/Makefile). Used in the build container to write golden vectors and to pin the
// oracle, and (as a prebuilt library) as the "reference" CPU baseline. Never part of the product.
//
// It calls the reference's own functions and plugin methods; nothing here restates their math.
//   D_40..D_62, Green_0k, IGreen_*   integratorElastic3DHelper.cpp:2394-2624
//   Integ3                            BEMBuilderElastic3D.cpp:1169-1353
//   Integ3_stress{,1,2}               BEMBuilderElastic3D.cpp:982-1167, 674-778, 324-427
//   plugin phases                     iBEMSystemService.h:58-67, systemSolverService.h:41-45
#include "configElastic3D.h"
#include "BEMBuilderElastic3D.h"
#include "integratorElastic3D.h"
#include "integratorElastic3DHelper.h"
#include "postProcessorElastic3D.h"
#include "systemSolverElastic3D.h"

#include <omp.h>
#include <unistd.h>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

// free functions defined (non-static) in BEMBuilderElastic3D.cpp
void Integ3_stress(double** Elcor, double** DDsu, double** DDst, double* point);
void Integ3_stress1(double** Elcor, double** DDsuw, double** DDstw, double* point);
void Integ3_stress2(double** Elcor, double** DDsuwz, double** DDstwz, double* point);

using namespace Elastic3DSpherical;

namespace {

struct Rows {
  std::vector<std::vector<double>> store;
  std::vector<double*> ptr;
  Rows(int n, int m) : store(n, std::vector<double>(m, 0.0)), ptr(n) {
    for (int i = 0; i < n; ++i) ptr[i] = store[i].data();
  }
  double** p() { return ptr.data(); }
};

void dump(const std::string& path, const double* p, size_t n) {
  FILE* f = fopen(path.c_str(), "wb");
  if (!f) return;
  fwrite(p, sizeof(double), n, f);
  fclose(f);
}

}  // namespace

extern "C" {

void ref_set_materials(double E0, double v0, double E1, double v1) {
  mu_0 = E0 / (2.0 * (1.0 + v0));
  nu0 = v0;
  mu_1 = E1 / (2.0 * (1.0 + v1));
  nu1 = v1;
}

// kind in {40,50,60,41,51,61,42,52,62}; idx holds the reference's integer arguments in its own order
// (i,j,k,l[,p[,q]][,r[,z]]); xp = field point, xo = sphere centre, a = radius.
double ref_D(int kind, const int* g, const double* xp, const double* xo, double a) {
  VectorXd rad(1);
  rad(0) = a;
  MatrixXd o(1, 3), p(1, 3);
  for (int c = 0; c < 3; ++c) {
    o(0, c) = xo[c];
    p(0, c) = xp[c];
  }
  switch (kind) {
    case 40: return D_40(g[0], g[1], g[2], g[3], 0, 0, rad, o, p);
    case 50: return D_50(g[0], g[1], g[2], g[3], g[4], 0, 0, rad, o, p);
    case 60: return D_60(g[0], g[1], g[2], g[3], g[4], g[5], 0, 0, rad, o, p);
    case 41: return D_41(g[0], g[1], g[2], g[3], g[4], 0, 0, rad, o, p);
    case 51: return D_51(g[0], g[1], g[2], g[3], g[4], g[5], 0, 0, rad, o, p);
    case 61: return D_61(g[0], g[1], g[2], g[3], g[4], g[5], g[6], 0, 0, rad, o, p);
    case 42: return D_42(g[0], g[1], g[2], g[3], g[4], g[5], 0, 0, rad, o, p);
    case 52: return D_52(g[0], g[1], g[2], g[3], g[4], g[5], g[6], 0, 0, rad, o, p);
    case 62: return D_62(g[0], g[1], g[2], g[3], g[4], g[5], g[6], g[7], 0, 0, rad, o, p);
  }
  return 0.0;
}

double ref_green(int order, const int* g, const double* xin) {
  double x[3] = {xin[0], xin[1], xin[2]};
  switch (order) {
    case 0: return Green(x, g[0], g[1]);
    case 1: return Green_01(x, g[0], g[1], g[2]);
    case 2: return Green_02(x, g[0], g[1], g[2], g[3]);
    case 3: return Green_03(x, g[0], g[1], g[2], g[3], g[4]);
    case 4: return Green_04(x, g[0], g[1], g[2], g[3], g[4], g[5]);
  }
  return 0.0;
}

// kind 1 -> IGreen_01(i,j,l), 11 -> IGreen_11(i,j,p,l), 21 -> IGreen_21(i,j,p,q,l)
double ref_igreen(int kind, const int* g, const double* xin, double a) {
  double x[3] = {xin[0], xin[1], xin[2]};
  switch (kind) {
    case 1: return IGreen_01(x, a, g[0], g[1], g[2]);
    case 11: return IGreen_11(x, a, g[0], g[1], g[2], g[3]);
    case 21: return IGreen_21(x, a, g[0], g[1], g[2], g[3], g[4]);
  }
  return 0.0;
}

// Element x collocation-node Kelvin quadrature of one element (4 nodes, 1-based ids in sctrb)
// against all nn nodes. dUe, dTe: row-major (3*nn) x 12.
void ref_integ3(double E, double nu, const int* sctrb_in, const double* nodes_rowmajor, int nn,
                double* dUe, double* dTe) {
  BEMBuilderElastic3D b;
  int sctrb[4];
  for (int j = 0; j < 4; ++j) sctrb[j] = sctrb_in[j];
  MatrixXd xP(nn, 3);
  for (int i = 0; i < nn; ++i)
    for (int c = 0; c < 3; ++c) xP(i, c) = nodes_rowmajor[3 * i + c];
  Rows el(4, 3), U(3 * nn, 12), T(3 * nn, 12);
  for (int j = 0; j < 4; ++j)
    for (int c = 0; c < 3; ++c) el.p()[j][c] = xP(sctrb[j] - 1, c);
  b.Integ3(E, nu, sctrb, el.p(), U.p(), T.p(), xP, nn);
  for (int i = 0; i < 3 * nn; ++i)
    for (int k = 0; k < 12; ++k) {
      dUe[12 * i + k] = U.p()[i][k];
      dTe[12 * i + k] = T.p()[i][k];
    }
}

// Element x sphere-centre quadrature, order 0/1/2. DDsu/DDst: row-major rows x 12,
// rows = 6 / 18 / 54.
void ref_integ3_stress(int order, const double* elcor_rowmajor, const double* point_in, double* DDsu,
                       double* DDst) {
  int rows = order == 0 ? 6 : (order == 1 ? 18 : 54);
  Rows el(4, 3), su(rows, 12), st(rows, 12);
  for (int j = 0; j < 4; ++j)
    for (int c = 0; c < 3; ++c) el.p()[j][c] = elcor_rowmajor[3 * j + c];
  double point[3] = {point_in[0], point_in[1], point_in[2]};
  if (order == 0) Integ3_stress(el.p(), su.p(), st.p(), point);
  if (order == 1) Integ3_stress1(el.p(), su.p(), st.p(), point);
  if (order == 2) Integ3_stress2(el.p(), su.p(), st.p(), point);
  for (int i = 0; i < rows; ++i)
    for (int k = 0; k < 12; ++k) {
      DDsu[12 * i + k] = su.p()[i][k];
      DDst[12 * i + k] = st.p()[i][k];
    }
}

// Full reference run on the txt inputs in `dir` (position.txt, BEM.txt, postprocess.txt), in the
// order iBEMSystemService::run -> systemSolverService::run -> postprocessService (getFlux).
// Binary dumps (raw little-endian float64, Eigen column-major) go to `out`:
//   H_<phase>.f64 / G_<phase>.f64 after each assembly phase when (flags & 1),
//   H.f64, G.f64 after assembly, A.f64, b.f64 after ApplyBC, X.f64, U.f64, T.f64, stress.f64 (nump x 6
//   row-major), dims.txt, times.txt (seconds per phase). stress.txt is written into `out` by getFlux.
// stop_after: 1 = assembly, 2 = +ApplyBC, 3 = +solve, 4 = +postprocess.
int ref_run(const char* dir, const char* out, int flags, int stop_after) {
  std::string d(dir), o(out);
  configElastic3D cfg;
  cfg.particleInput = d + "/position.txt";
  cfg.bemInput = d + "/BEM.txt";
  cfg.postProcessInput = d + "/postprocess.txt";
  cfg.setUp();
  BEMBuilderElastic3D bem;
  integratorElastic3D integ;
  systemSolverElastic3D solver;
  postProcessorElastic3D post;
  BEMBuilder& B = bem;
  integrator& I = integ;
  systemSolver& S = solver;
  postProcessor& P = post;

  const int n = 3 * cfg.NN + 78 * cfg.num;
  FILE* dims = fopen((o + "/dims.txt").c_str(), "w");
  if (dims) {
    fprintf(dims, "NN %d\nNE %d\nnum %d\nnump %d\nn %d\n", cfg.NN, cfg.NE, cfg.num, cfg.nump, n);
    fclose(dims);
  }
  FILE* tf = fopen((o + "/times.txt").c_str(), "w");
  auto phase = [&](const char* name, auto&& fn) {
    double t0 = omp_get_wtime();
    fn();
    double t1 = omp_get_wtime();
    if (tf) fprintf(tf, "%s %.6f\n", name, t1 - t0);
    if ((flags & 1) && cfg.HMAT.size() > 0) {
      dump(o + "/H_" + name + ".f64", cfg.HMAT.data(), cfg.HMAT.size());
      dump(o + "/G_" + name + ".f64", cfg.GMAT.data(), cfg.GMAT.size());
    }
  };
  phase("addBEM", [&] { B.addBEM(cfg); });
  phase("addFieldToBEM", [&] { I.addFieldToBEM(cfg); });
  phase("incl0", [&] { I.addFluxEquivalentInclusion(cfg); });
  phase("incl1", [&] { I.addFluxEquivalentFirstOrderInclusion(cfg); });
  phase("incl2", [&] { I.addFluxEquivalentSecondOrderInclusion(cfg); });
  phase("bem0", [&] { B.addFluxEquivalentBEM(cfg); });
  phase("bem1", [&] { B.addFluxEquivalentFirstOrderBEM(cfg); });
  phase("bem2", [&] { B.addFluxEquivalentSecondOrderBEM(cfg); });
  dump(o + "/H.f64", cfg.HMAT.data(), cfg.HMAT.size());
  dump(o + "/G.f64", cfg.GMAT.data(), cfg.GMAT.size());
  if (stop_after >= 2) {
    flags &= ~1;
    phase("ApplyBC", [&] { B.ApplyBC(cfg); });
    dump(o + "/A.f64", cfg.A.data(), cfg.A.size());
    dump(o + "/b.f64", cfg.b.data(), cfg.b.size());
  }
  if (stop_after >= 3) {
    phase("solve", [&] { S.solve(cfg); });
    phase("getUTVector", [&] { B.getUTVector(cfg); });
    dump(o + "/X.f64", cfg.X.data(), cfg.X.size());
    dump(o + "/U.f64", cfg.U.data(), cfg.U.size());
    dump(o + "/T.f64", cfg.T.data(), cfg.T.size());
  }
  if (stop_after >= 4) {
    char cwd[4096];
    if (!getcwd(cwd, sizeof cwd)) cwd[0] = 0;
    if (chdir(o.c_str()) != 0) return 2;
    phase("getFlux", [&] { P.getFlux(cfg); });
    if (cwd[0] && chdir(cwd) != 0) return 3;
    std::vector<double> s(6 * cfg.nump);
    for (int k = 0; k < cfg.nump; ++k)
      for (int c = 0; c < 6; ++c) s[6 * k + c] = cfg.s_pt[k][c];
    dump(o + "/stress.f64", s.data(), s.size());
  }
  if (tf) fclose(tf);
  return 0;
}

// Eigen full-pivot LU solve (the reference solver, systemSolverElastic3D.h:55) of a column-major
// n x n system; returns seconds spent in the factor+solve.
double ref_fullpiv_solve(int n, const double* A, const double* b, double* x) {
  MatrixXd M = Map<const MatrixXd>(A, n, n);
  VectorXd r = Map<const VectorXd>(b, n);
  double t0 = omp_get_wtime();
  VectorXd s = M.fullPivLu().solve(r);
  double t1 = omp_get_wtime();
  for (int i = 0; i < n; ++i) x[i] = s(i);
  return t1 - t0;
}

}  // extern "C"
