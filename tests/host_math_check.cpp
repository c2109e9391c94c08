// Test infrastructure: compiles the kernels' per-entry math (ibem3d_math.h) for the host so the CPU
// test suite can check it against the oracle without a GPU. Not part of the product library.
#include "../ibemproject_amd/csrc/ibem3d_math.h"

#include <cmath>
#include <vector>

using namespace ibem;

extern "C" {

// 60 x 60 inclusion block (column-major) for offset x = target centre - source centre, radius a.
void hc_incl_block(double E0, double nu0, double E1, double nu1, const double* x, double a,
                   int self, double* out) {
  Mat M = make_mat(E0, nu0, E1, nu1);
  RadPoint P;
  rad_point(P, x);
  const bool inside = is_inside(x, a);
  double C[8][4];
  pot_coefs(a, inside, C);
  std::vector<double> tab(8 * kTabN);
  for (int t = 0; t < kTabN; ++t) {
    double v[8];
    pot_entry(P, inside, C, t, v);
    for (int f = 0; f < 8; ++f) tab[f * kTabN + t] = v[f];
  }
  for (int c = 0; c < 60; ++c)
    for (int r = 0; r < 60; ++r) out[c * 60 + r] = incl_entry(tab.data(), kTabN, M, self != 0, r, c);
}

// 3 x 60 node block (column-major) for x = node - sphere centre
void hc_field_block(double E0, double nu0, double E1, double nu1, const double* x, double a,
                    double* out) {
  Mat M = make_mat(E0, nu0, E1, nu1);
  RadPoint P;
  rad_point(P, x);
  const bool inside = is_inside(x, a);
  double C[8][4];
  pot_coefs(a, inside, C);
  for (int c = 0; c < 60; ++c)
    for (int i = 0; i < 3; ++i) out[c * 3 + i] = field_entry(P, inside, C, M, i, c);
}

// The same block through the per-(node, sphere) radial table the GPU kernel uses
void hc_field_block_tab(double E0, double nu0, double E1, double nu1, const double* x, double a,
                        double* out) {
  Mat M = make_mat(E0, nu0, E1, nu1);
  RadPoint P;
  rad_point(P, x);
  const bool inside = is_inside(x, a);
  double C[8][4];
  pot_coefs(a, inside, C);
  double tab[kFieldTab];
  for (int u = 0; u < kFieldMi; ++u) field_tab_fill(P, inside, C, u, tab);
  for (int c = 0; c < 60; ++c)
    for (int i = 0; i < 3; ++i) out[c * 3 + i] = field_entry_tab(tab, M, i, c);
}

// Element x sphere-centre block in the 60-row layout: H part (60 x 12, = -fac DDsu) and G part
// (60 x 12, = fac DDst), row-major; post != 0 selects the stress (postprocessor) variant with
// order-0 rows only (6 x 12).
void hc_bem_rows(double E0, double nu0, double E1, double nu1, const double* el, const double* point,
                 int post, double* Hout, double* Gout) {
  Mat M = make_mat(E0, nu0, E1, nu1);
  const int rows = post ? 6 : 60;
  for (int q = 0; q < rows * 12; ++q) Hout[q] = Gout[q] = 0.0;
  std::vector<double> tabS(kTabN), tabT(kTabN);
  for (int a = 0; a < 8; ++a)
    for (int b = 0; b < 8; ++b) {
      QPoint Q;
      qpoint(el, gauss8_x(a), gauss8_x(b), Q);
      const double W = gauss8_w(a) * gauss8_w(b) * Q.jac;
      double x[3] = {Q.g[0] - point[0], Q.g[1] - point[1], Q.g[2] - point[2]};
      RadPoint P;
      rad_point(P, x);
      for (int t = 0; t < tidx(0, 0, 6) + 1; ++t) kelvin_entry(P, t, tabS[t], tabT[t]);
      for (int r = 0; r < rows; ++r) {
        Slot R = slot60(r);
        double su[3], st[3];
        const double dlam = post ? M.lam0 : M.dlam, dmu = post ? M.mu0 : M.dmu;
        strain_kernel(tabS.data(), tabT.data(), M, dlam, dmu, Q.nv, R.i, R.j, R.e,
                      R.order == 1 ? -1 : 1, su, st);
        const double fac = R.order == 2 ? 0.5 : 1.0;
        for (int qn = 0; qn < 4; ++qn)
          for (int n = 0; n < 3; ++n) {
            Hout[r * 12 + 3 * qn + n] += (post ? 1.0 : -fac) * su[n] * Q.N[qn] * W;
            Gout[r * 12 + 3 * qn + n] += fac * st[n] * Q.N[qn] * W;
          }
      }
    }
}

}  // extern "C"

extern "C" {
// Per-Gauss-point strain rows two ways at offset x with unit normal nv: the generated tables +
// row_gp (what k_rows runs) and the per-entry tables + strain_kernel_n. out_new/out_old: 60 x 6
// row-major (su0 su1 su2 st0 st1 st2); tabs: the 119-entry generated table.
void hc_row_gp(double E0, double nu0, double E1, double nu1, const double* x, const double* nv,
               double* out_new, double* out_old, double* tabs) {
  Mat M = make_mat(E0, nu0, E1, nu1);
  double y0[7], y1[7], y2[7], PR[7], PT[5];
  kelvin_point(x, y0, y1, y2, PR, PT);
  double tab[kKStride];
  for (int i = 0; i < kKStride; ++i) tab[i] = 0.0;
  kelvin_group<0>(y0, y1, y2, PR, PT, tab);
  kelvin_group<1>(y0, y1, y2, PR, PT, tab);
  kelvin_group<2>(y0, y1, y2, PR, PT, tab);
  kelvin_group<3>(y0, y1, y2, PR, PT, tab);
  for (int i = 0; i < kKStride; ++i) tabs[i] = tab[i];
  RadPoint P;
  rad_point(P, x);
  double tS[84], tT[35];
  for (int t = 0; t < 84; ++t) {
    double sv, tv;
    kelvin_entry(P, t, sv, tv);
    tS[t] = sv;
    if (t < 35) tT[t] = tv;
  }
  for (int r = 0; r < 60; ++r) {
    const Slot R = slot60(r);
    const RowTerms RT = row_terms(M, R);
    row_gp2(RT, tab, nv, out_new + 6 * r, out_new + 6 * r + 3);
    double chk[6];
    row_gp(RT, tab, nv, chk, chk + 3);
    for (int q = 0; q < 6; ++q)
      if (chk[q] != out_new[6 * r + q] && std::abs(chk[q] - out_new[6 * r + q]) > 1e-13 * (1e-300 + std::abs(chk[q])) * 1e3)
        out_new[6 * r + q] = 1e300;  // row_gp and row_gp2 disagree
    const int sg = R.order == 1 ? -1 : 1;
    for (int n = 0; n < 3; ++n)
      strain_kernel_n(tS, tT, M, M.dlam, M.dmu, nv, R.i, R.j, R.e, sg, n, out_old[6 * r + n],
                      out_old[6 * r + 3 + n]);
  }
}
}

extern "C" {
// The same 60 x 60 inclusion block through the per-pair weighted tables (what k_incl runs).
void hc_incl_block_w(double E0, double nu0, double E1, double nu1, const double* x, double a,
                     int self, double* out) {
  Mat M = make_mat(E0, nu0, E1, nu1);
  RadPoint P;
  rad_point(P, x);
  const bool inside = is_inside(x, a);
  double C[8][4];
  pot_coefs(a, inside, C);
  std::vector<double> basis(4 * kTabN), PS(kWgt * kPS), PH(kWgt * kPH);
  for (int t = 0; t < kTabN; ++t) {
    int b0, b1, b2;
    tdecode(t, b0, b1, b2);
    double B0, B1, B2, B3;
    rad_basis(P, inside, b0, b1, b2, B0, B1, B2, B3);
    basis[t] = B0;
    basis[kTabN + t] = B1;
    basis[2 * kTabN + t] = B2;
    basis[3 * kTabN + t] = B3;
  }
  for (int u = 0; u < kInclW; ++u) incl_weighted(basis.data(), C, u, PS.data(), PH.data());
  if (self >= 2) {  // the weighted tables through the pair-independent descriptors (k_incl)
    for (int u = 0; u < kInclW; ++u) {
      const double v = incl_weighted_val_d(basis.data(), C, incl_weighted_desc(u));
      const double ref = u < kWgt * kPS ? PS[u] : PH[u - kWgt * kPS];
      if (!(v == ref)) out[0] = 1e300;  // descriptor path disagrees
      if (u < kWgt * kPS) PS[u] = v; else PH[u - kWgt * kPS] = v;
    }
  }
  for (int r = 0; r < 60; ++r) {
    const InclRow R = incl_row(r);
    if (self >= 2) {  // the unrolled, branch-free row (what k_incl runs); self = 3 for the self block
      incl_row_entries(incl_row_pack(R), PS.data(), PH.data(), M, self == 3, out + r, 60);
      continue;
    }
    for (int c = 0; c < 60; ++c) out[c * 60 + r] = incl_entry_w(R, PS.data(), PH.data(), M, self != 0, c);
  }
}
}

extern "C" {
// The eight radial functions' coefficients (pot_coefs) for radius a, inside or outside: C[8][4].
void hc_pot_coefs(double a, int inside, double* out) {
  double C[8][4];
  pot_coefs(a, inside != 0, C);
  for (int i = 0; i < 8; ++i)
    for (int j = 0; j < 4; ++j) out[4 * i + j] = C[i][j];
}
}

extern "C" {
// One element (corner nodes el[12]) against one sphere centre c[3], the boundary->inclusion strain
// rows integrated two ways: pointwise (k_rows: contraction at every Gauss point, then the weighted
// sum) and in moment form (k_rows_m: the weighted sums of the Kelvin tables over the two halves of
// the points, then one contraction per shape function, with the normal of the first point).
// out_pw / out_mom: 60 x 24 row-major, su_n summed against N_q at n * 4 + q, st_n at 12 + n * 4 + q.
void hc_rows_element(double E0, double nu0, double E1, double nu1, const double* c, const double* el,
                     double* out_pw, double* out_mom) {
  Mat M = make_mat(E0, nu0, E1, nu1);
  std::vector<double> tabs(64 * kKStride, 0.0);
  QPoint Q[64];
  double w[64];
  for (int g = 0; g < 64; ++g) {
    const int a = g / 8, b = g % 8;
    qpoint(el, gauss8_x(a), gauss8_x(b), Q[g]);
    w[g] = gauss8_w(a) * gauss8_w(b) * Q[g].jac;
    const double x[3] = {Q[g].g[0] - c[0], Q[g].g[1] - c[1], Q[g].g[2] - c[2]};
    double y0[7], y1[7], y2[7], PR[7], PT[5];
    kelvin_point(x, y0, y1, y2, PR, PT);
    double* tab = tabs.data() + g * kKStride;
    kelvin_group<0>(y0, y1, y2, PR, PT, tab);
    kelvin_group<1>(y0, y1, y2, PR, PT, tab);
    kelvin_group<2>(y0, y1, y2, PR, PT, tab);
    kelvin_group<3>(y0, y1, y2, PR, PT, tab);
  }
  std::vector<double> mom(2 * 4 * kKStride, 0.0);
  for (int half = 0; half < 2; ++half)
    for (int m = 0; m < kKUsed; ++m) {
      const int slot = kelvin_used_slot(m);
      for (int q = 0; q < 4; ++q) {
        double p = 0.0;
        for (int gi = 0; gi < 32; ++gi) {
          const int g = 32 * half + gi;
          p += tabs[g * kKStride + slot] * (w[g] * Q[g].N[q]);
        }
        mom[(half * 4 + q) * kKStride + slot] = p;
      }
    }
  for (int r = 0; r < 60; ++r) {
    const RowTerms RT = row_terms(M, slot60(r));
    double aH[12] = {0}, aG[12] = {0};
    for (int g = 0; g < 64; ++g) {
      double su[3], st[3];
      row_gp2(RT, tabs.data() + g * kKStride, Q[g].nv, su, st);
      for (int q = 0; q < 4; ++q)
        for (int n = 0; n < 3; ++n) {
          aH[n * 4 + q] += su[n] * (Q[g].N[q] * w[g]);
          aG[n * 4 + q] += st[n] * (Q[g].N[q] * w[g]);
        }
    }
    for (int o = 0; o < 12; ++o) {
      out_pw[24 * r + o] = aH[o];
      out_pw[24 * r + 12 + o] = aG[o];
    }
    for (int q = 0; q < 4; ++q) {
      double su[3], st[3];
      row_gp2_t(RT, TabSum2{mom.data() + q * kKStride, mom.data() + (4 + q) * kKStride}, Q[0].nv, su, st);
      for (int n = 0; n < 3; ++n) {
        out_mom[24 * r + n * 4 + q] = su[n];
        out_mom[24 * r + 12 + n * 4 + q] = st[n];
      }
    }
  }
}
}
