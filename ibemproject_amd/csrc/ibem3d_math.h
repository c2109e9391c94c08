// Closed-form math of the elastic-3D sphere path, shared by the HIP kernels (device) and the host
// validation build (host). FP64 throughout.
//
// Sphere potentials. For a ball B of radius a centred at 0 and x = field point - centre,
//   Phi^w(x) = int_B w(y) / |x - y| dy,  Psi^w(x) = int_B w(y) |x - y| dy,  w in {1, y_p, y_p y_q},
// are the functions whose derivatives the reference hand-expands (integratorElastic3DHelper.cpp:41-750:
// Phi_0k, Phi_1k, Phi_2k, Psi_0k, Psi_1k, Psi_2k). Every one of them is a derivative of a radial
// function f(rho), rho = |x|^2:
//   Phi^1 = F0,  Phi^p = d_p F1,  Phi^pq = d_pq F2b + delta_pq F2a,
//   Psi^1 = G0,  Psi^p = d_p G1,  Psi^pq = d_pq G2b + delta_pq G2a,
// with, outside the ball (the reference's lambda > 0 branch), V = 4 pi a^3 / 3,
//   F0 = V r^-1, F1 = -(V a^2/5) r^-1, F2a = (V a^2/5) r^-1, F2b = (4 pi a^7/105) r^-1,
//   G0 = V r + (V a^2/5) r^-1, G1 = -(V a^2/5) r - (4 pi a^7/105) r^-1,
//   G2a = (V a^2/5) r + (4 pi a^7/105) r^-1, G2b = (4 pi a^7/105) r + (4 pi a^9/945) r^-1
// (multipole expansion of |x-y|^{+-1} against the ball's moments), and inside (lambda = 0) by the
// polynomials
//   F0 = -2pi/3 rho, F1 = pi a^2/3 rho - pi/10 rho^2, F2a = -F1, F2b = pi a^2/20 rho^2 - pi/84 rho^3,
//   G0 = -pi/15 rho^2, G1 = pi a^2/30 rho^2 - pi/210 rho^3, G2a = -G1,
//   G2b = pi a^2/420 rho^3 - pi/3024 rho^4,
// exact for first and higher derivatives of Phi and third and higher of Psi, which is all the path
// uses. (The same functions as the reference's V_n / Phi_.1 / Psi_.3 formulas, :169-543, with I_n
// constant inside. Derivation: both kernels expanded in Legendre polynomials of the angle between x
// and y, the angles integrated in closed form, the radius shell by shell split at |y| = |x|;
// tests/test_potentials.py carries it out in sympy and checks pot_coefs against it.)
//
// Radial derivatives. d^beta rho^s (beta a multi-index, k = |beta|) is
//   sum_j f^(k-j)(rho) 2^(k-j) M_j(beta, x),  f = rho^s,
//   M_j = sum_{j0+j1+j2 = j} prod_c beta_c! / (j_c! 2^j_c (beta_c - 2 j_c)!) x_c^(beta_c - 2 j_c),
// i.e. every way of pairing j index slots into Kronecker deltas. Tables hold d^beta for |beta| <= 8
// in degree-major order (tidx).
#pragma once

#include <math.h>

#if defined(__HIPCC__)
#define IBEM_HD __host__ __device__ inline
#else
#define IBEM_HD inline
#endif
// Instruction-scheduling fence in device code (keeps the scheduler from interleaving unrolled
// blocks, which bounds their live ranges); nothing in the host build of these functions.
#if defined(__HIP_DEVICE_COMPILE__)
#define IBEM_SCHED_FENCE() __builtin_amdgcn_sched_barrier(0)
#else
#define IBEM_SCHED_FENCE() ((void)0)
#endif

namespace ibem {

constexpr double kPi = 3.141592653589793;  // GlobalConstant.h:26
constexpr int kTabOrder = 8;
constexpr int kTabN = 165;  // entries of degree <= 8 in 3 variables

// pair (i,j), i<=j, in the order 00 01 02 11 12 22 (configElastic3D.cpp:312-318)
IBEM_HD int pair_i(int t) { return t < 3 ? 0 : (t < 5 ? 1 : 2); }
IBEM_HD int pair_j(int t) { return t < 3 ? t : (t < 5 ? t - 2 : 2); }
IBEM_HD int pair_t(int i, int j) {
  if (i > j) { int s = i; i = j; j = s; }
  return i == 0 ? j : (i == 1 ? 2 + j : 5);
}
// weight of the unknown E_mn, m<=n, standing for both E_mn and E_nm (GlobalConstant.h:28)
IBEM_HD double mapw(int m, int n) { return m == n ? 1.0 : 2.0; }

// packed index of the multi-index (b0,b1,b2), degree-major, b0 then b1 descending
IBEM_HD int tidx(int b0, int b1, int b2) {
  int d = b0 + b1 + b2, e = d - b0;
  return d * (d + 1) * (d + 2) / 6 + e * (e + 1) / 2 + (e - b1);
}
IBEM_HD void tdecode(int t, int& b0, int& b1, int& b2) {
  int d = 0;
  while ((d + 1) * (d + 2) * (d + 3) / 6 <= t) ++d;
  int r = t - d * (d + 1) * (d + 2) / 6;
  int e = 0;
  while ((e + 1) * (e + 2) / 2 <= r) ++e;
  int f = r - e * (e + 1) / 2;
  b0 = d - e;
  b1 = e - f;
  b2 = d - b0 - b1;
}

// beta! / (j! 2^j (beta-2j)!) for beta <= 8, j <= 4
IBEM_HD double pair_count(int b, int j) {
  // rows b = 0..8, cols j = 0..4
  static constexpr double t[9][5] = {{1, 0, 0, 0, 0},    {1, 0, 0, 0, 0},     {1, 1, 0, 0, 0},
                          {1, 3, 0, 0, 0},    {1, 6, 3, 0, 0},     {1, 10, 15, 0, 0},
                          {1, 15, 45, 15, 0}, {1, 21, 105, 105, 0}, {1, 28, 210, 420, 105}};
  return t[b][j];
}

// Geometry of one evaluation point for the radial tables.
struct RadPoint {
  double x[3];
  double rho;  // |x|^2
  double rs;   // sqrt(rho)
};

IBEM_HD void rad_point(RadPoint& P, const double* x) {
  P.x[0] = x[0];
  P.x[1] = x[1];
  P.x[2] = x[2];
  P.rho = x[0] * x[0] + x[1] * x[1] + x[2] * x[2];
  P.rs = sqrt(P.rho);
}

IBEM_HD double ipow(double x, int e) {
  double r = 1.0;
  for (int q = 0; q < e; ++q) r *= x;
  return r;
}

// Pair sums M_j(beta, x), j = 0..4, returned in five registers (no private arrays: every index
// below is resolved by selects, so nothing spills to scratch on the GPU).
struct PairSums {
  double m0, m1, m2, m3, m4;
  IBEM_HD double get(int j) const {
    return j == 0 ? m0 : (j == 1 ? m1 : (j == 2 ? m2 : (j == 3 ? m3 : m4)));
  }
};

IBEM_HD PairSums pair_sums(const RadPoint& P, int b0, int b1, int b2) {
  PairSums S{0.0, 0.0, 0.0, 0.0, 0.0};
  const double x0s = P.x[0] * P.x[0], x1s = P.x[1] * P.x[1], x2s = P.x[2] * P.x[2];
  double p0 = ipow(P.x[0], b0 & 1);
  for (int j0 = b0 / 2; j0 >= 0; --j0, p0 *= x0s) {
    const double c0 = pair_count(b0, j0) * p0;
    double p1 = ipow(P.x[1], b1 & 1);
    for (int j1 = b1 / 2; j1 >= 0; --j1, p1 *= x1s) {
      const double c1 = c0 * pair_count(b1, j1) * p1;
      double p2 = ipow(P.x[2], b2 & 1);
      for (int j2 = b2 / 2; j2 >= 0; --j2, p2 *= x2s) {
        const double v = c1 * pair_count(b2, j2) * p2;
        const int j = j0 + j1 + j2;
        S.m0 += j == 0 ? v : 0.0;
        S.m1 += j == 1 ? v : 0.0;
        S.m2 += j == 2 ? v : 0.0;
        S.m3 += j == 3 ? v : 0.0;
        S.m4 += j == 4 ? v : 0.0;
      }
    }
  }
  return S;
}

// d^beta of r and 1/r (outside) from the pair sums: out0 = d^beta r, out1 = d^beta r^-1
IBEM_HD void rad_half(const RadPoint& P, int k, const PairSums& M, double& out0, double& out1) {
  // f = rho^s; f^(m) = s (s-1) ... (s-m+1) rho^(s-m)
  const double inv2 = 2.0 / P.rho;
  double sp = 0.0, sm = 0.0;
  for (int j = 0; j <= k / 2; ++j) {
    const int m = k - j;
    double ffp = 1.0, ffm = 1.0, pw = 1.0;
    for (int q = 0; q < m; ++q) {
      ffp *= (0.5 - q);
      ffm *= (-0.5 - q);
      pw *= inv2;
    }
    const double mj = M.get(j);
    sp += ffp * pw * mj;
    sm += ffm * pw * mj;
  }
  out0 = sp * P.rs;
  out1 = sm / P.rs;
}

// d^beta of rho^s (inside, s = 1..4) from the pair sums
IBEM_HD double rad_poly1(const RadPoint& P, int k, const PairSums& M, int s) {
  double acc = 0.0;
  for (int j = 0; j <= k / 2; ++j) {
    const int m = k - j;
    if (m > s) continue;
    double ff = 1.0;
    for (int q = 0; q < m; ++q) ff *= (double)(s - q);
    acc += ff * ipow(P.rho, s - m) * (double)(1 << m) * M.get(j);
  }
  return acc;
}

// basis values at multi-index beta: outside (d r, d r^-1, 0, 0), inside (d rho .. d rho^4)
IBEM_HD void rad_basis(const RadPoint& P, bool inside, int b0, int b1, int b2, double& B0, double& B1,
                       double& B2, double& B3) {
  const int k = b0 + b1 + b2;
  const PairSums M = pair_sums(P, b0, b1, b2);
  if (inside) {
    B0 = rad_poly1(P, k, M, 1);
    B1 = rad_poly1(P, k, M, 2);
    B2 = rad_poly1(P, k, M, 3);
    B3 = rad_poly1(P, k, M, 4);
  } else {
    rad_half(P, k, M, B0, B1);
    B2 = B3 = 0.0;
  }
}

// The outside basis (d^beta r, d^beta r^-1) from per-point factors F (kRadF doubles): F[m] =
// 2^m f^(m)(rho) for f = rho^(1/2) and F[9 + m] for f = rho^(-1/2) (m <= 8), F[18 + 9 c + p] = x_c^p
// (p <= 8); the same pair sums as rad_basis with each term's radial factor applied as it is summed.
constexpr int kRadF = 45;
IBEM_HD void rad_factors(const double* x, double* F) {
  const double rho = x[0] * x[0] + x[1] * x[1] + x[2] * x[2];
  const double r = sqrt(rho), inv2 = 2.0 / rho;
  F[0] = r;
  F[9] = 1.0 / r;
  for (int m = 1; m <= 8; ++m) {
    F[m] = F[m - 1] * (0.5 - (m - 1)) * inv2;
    F[9 + m] = F[9 + m - 1] * (-0.5 - (m - 1)) * inv2;
  }
  for (int c = 0; c < 3; ++c) {
    F[18 + 9 * c] = 1.0;
    for (int p = 1; p <= 8; ++p) F[18 + 9 * c + p] = F[18 + 9 * c + p - 1] * x[c];
  }
}
IBEM_HD void rad_basis_out(const double* F, int b0, int b1, int b2, double& B0, double& B1) {
  const int k = b0 + b1 + b2;
  double sp = 0.0, sm = 0.0;
  for (int j0 = 0; 2 * j0 <= b0; ++j0) {
    const double c0 = pair_count(b0, j0) * F[18 + b0 - 2 * j0];
    for (int j1 = 0; 2 * j1 <= b1; ++j1) {
      const double c1 = c0 * pair_count(b1, j1) * F[27 + b1 - 2 * j1];
      for (int j2 = 0; 2 * j2 <= b2; ++j2) {
        const double v = c1 * pair_count(b2, j2) * F[36 + b2 - 2 * j2];
        const int m = k - (j0 + j1 + j2);
        sp += v * F[m];
        sm += v * F[9 + m];
      }
    }
  }
  B0 = sp;
  B1 = sm;
}

// Potential table slots
enum { kF0 = 0, kF1 = 1, kF2a = 2, kF2b = 3, kG0 = 4, kG1 = 5, kG2a = 6, kG2b = 7 };

// Coefficients of the eight radial functions on their basis:
//   outside: basis (r, r^-1); inside: basis (rho, rho^2, rho^3, rho^4).
IBEM_HD void pot_coefs(double a, bool inside, double C[8][4]) {
  for (int i = 0; i < 8; ++i)
    for (int j = 0; j < 4; ++j) C[i][j] = 0.0;
  const double a2 = a * a;
  if (!inside) {
    const double V = 4.0 * kPi * a * a2 / 3.0;
    const double c2 = V * a2 / 5.0;                            // V a^2/5
    const double c7 = 4.0 * kPi * a2 * a2 * a2 * a / 105.0;    // 4 pi a^7/105
    const double c9 = 4.0 * kPi * a2 * a2 * a2 * a2 * a / 945.0;  // 4 pi a^9/945
    C[kF0][1] = V;
    C[kF1][1] = -c2;
    C[kF2a][1] = c2;
    C[kF2b][1] = c7;
    C[kG0][0] = V;
    C[kG0][1] = c2;
    C[kG1][0] = -c2;
    C[kG1][1] = -c7;
    C[kG2a][0] = c2;
    C[kG2a][1] = c7;
    C[kG2b][0] = c7;
    C[kG2b][1] = c9;
  } else {
    C[kF0][0] = -2.0 * kPi / 3.0;
    C[kF1][0] = kPi * a2 / 3.0;
    C[kF1][1] = -kPi / 10.0;
    C[kF2a][0] = -kPi * a2 / 3.0;
    C[kF2a][1] = kPi / 10.0;
    C[kF2b][1] = kPi * a2 / 20.0;
    C[kF2b][2] = -kPi / 84.0;
    C[kG0][1] = -kPi / 15.0;
    C[kG1][1] = kPi * a2 / 30.0;
    C[kG1][2] = -kPi / 210.0;
    C[kG2a][1] = -kPi * a2 / 30.0;
    C[kG2a][2] = kPi / 210.0;
    C[kG2b][2] = kPi * a2 / 420.0;
    C[kG2b][3] = -kPi / 3024.0;
  }
}

// The reference's inside test: lambda = max(|x|^2 - a^2, 0) == 0 (integratorElastic3DHelper.cpp:136)
IBEM_HD bool is_inside(const double* x, double a) {
  return !((x[0] * x[0] + x[1] * x[1] + x[2] * x[2] - a * a) > 0);
}

// All eight potential tables at table index t (one entry of each table).
IBEM_HD void pot_entry(const RadPoint& P, bool inside, const double C[8][4], int t, double* out8) {
  int b0, b1, b2;
  tdecode(t, b0, b1, b2);
  double B0, B1, B2, B3;
  rad_basis(P, inside, b0, b1, b2, B0, B1, B2, B3);
  for (int f = 0; f < 8; ++f) out8[f] = C[f][0] * B0 + C[f][1] * B1 + C[f][2] * B2 + C[f][3] * B3;
}

// Derivative of Phi^w (psi = 0) or Psi^w (psi = 1) with multi-index (c0,c1,c2) from the tables
// laid out as tab[f * stride + tidx]. w: 0 -> 1, 1+p -> y_p, 4+pair_t(p,q) -> y_p y_q.
IBEM_HD double pot(const double* tab, int stride, int psi, int w, int c0, int c1, int c2) {
  const double* T = tab + (psi ? 4 : 0) * stride;
  if (w == 0) return T[tidx(c0, c1, c2)];
  if (w < 4) {
    const int p = w - 1;
    return T[stride + tidx(c0 + (p == 0), c1 + (p == 1), c2 + (p == 2))];
  }
  const int p = pair_i(w - 4), q = pair_j(w - 4);
  double v = T[3 * stride + tidx(c0 + (p == 0) + (q == 0), c1 + (p == 1) + (q == 1),
                                 c2 + (p == 2) + (q == 2))];
  if (p == q) v += T[2 * stride + tidx(c0, c1, c2)];
  return v;
}

// Material constants (configElastic3D.cpp:54, :119; integratorElastic3DHelper.cpp:2376-2389)
struct Mat {
  double mu0, nu0, lam0, mu1, nu1, lam1;
  double dlam, dmu;  // lam0 - lam1, mu0 - mu1
  double cD;         // 1 / (8 pi (1 - nu0))
  double cg1, cg2;   // 1/(4 pi mu0), 1/(16 pi mu0 (1 - nu0))
};

IBEM_HD Mat make_mat(double E0, double nu0, double E1, double nu1) {
  Mat m;
  m.mu0 = E0 / (2.0 * (1.0 + nu0));
  m.nu0 = nu0;
  m.lam0 = 2.0 * m.mu0 * nu0 / (1.0 - 2.0 * nu0);
  m.mu1 = E1 / (2.0 * (1.0 + nu1));
  m.nu1 = nu1;
  m.lam1 = 2.0 * m.mu1 * nu1 / (1.0 - 2.0 * nu1);
  m.dlam = m.lam0 - m.lam1;
  m.dmu = m.mu0 - m.mu1;
  m.cD = 1.0 / (8.0 * kPi * (1.0 - nu0));
  m.cg1 = 1.0 / (4.0 * kPi * m.mu0);
  m.cg2 = 1.0 / (16.0 * kPi * m.mu0 * (1.0 - nu0));
  return m;
}

IBEM_HD double C0(const Mat& M, int i, int j, int k, int l) {
  return M.lam0 * (i == j) * (k == l) + M.mu0 * ((i == k) * (j == l) + (i == l) * (j == k));
}

// Eshelby-type tensor D(k,l | m,n | w | extra) of the reference (D_40..D_62,
// integratorElastic3DHelper.cpp:2394-2554) with strain indices (k,l), eigenstrain (m,n), weight w
// and extra derivative counts e[3]:
//   cD [ Psi^w,mnkl - 2 nu0 d_mn Phi^w,kl - (1-nu0)(d_kn Phi^w,ml + d_ln Phi^w,mk
//        + d_km Phi^w,nl + d_lm Phi^w,nk) ]      (every derivative also carries e)
IBEM_HD double Dten(const double* tab, int stride, const Mat& M, int w, int k, int l, int m, int n,
                    int e0, int e1, int e2) {
  // multi-index arithmetic by comparisons and values passed by value: no private array whose
  // address escapes, so nothing goes to scratch memory
#define IBEM_C3(a, b) e0 + ((a) == 0) + ((b) == 0), e1 + ((a) == 1) + ((b) == 1), e2 + ((a) == 2) + ((b) == 2)
  double v = pot(tab, stride, 1, w, e0 + (m == 0) + (n == 0) + (k == 0) + (l == 0),
                 e1 + (m == 1) + (n == 1) + (k == 1) + (l == 1),
                 e2 + (m == 2) + (n == 2) + (k == 2) + (l == 2));
  if (m == n) v -= 2.0 * M.nu0 * pot(tab, stride, 0, w, IBEM_C3(k, l));
  double s = 0.0;
  if (k == n) s += pot(tab, stride, 0, w, IBEM_C3(m, l));
  if (l == n) s += pot(tab, stride, 0, w, IBEM_C3(m, k));
  if (k == m) s += pot(tab, stride, 0, w, IBEM_C3(n, l));
  if (l == m) s += pot(tab, stride, 0, w, IBEM_C3(n, k));
#undef IBEM_C3
  v -= (1.0 - M.nu0) * s;
  return M.cD * v;
}

// sum_k D(k,k | m,n | w | e)
IBEM_HD double Dtrace(const double* tab, int stride, const Mat& M, int w, int m, int n, int e0,
                      int e1, int e2) {
  double v = 0.0;
  for (int k = 0; k < 3; ++k) v += Dten(tab, stride, M, w, k, k, m, n, e0, e1, e2);
  return v;
}

// Transpose of pot(): the table entry pot() reads (table f = 0..7, index t) and its weight, passed
// to sink(f, t, s) -- so that sum over calls of sink's s T_f[t] equals the forward value.
template <class Sink>
IBEM_HD void pot_adj(Sink& sink, int psi, int w, int c0, int c1, int c2, double s) {
  const int f0 = psi ? 4 : 0;
  if (w == 0) {
    sink(f0, tidx(c0, c1, c2), s);
  } else if (w < 4) {
    const int p = w - 1;
    sink(f0 + 1, tidx(c0 + (p == 0), c1 + (p == 1), c2 + (p == 2)), s);
  } else {
    const int p = pair_i(w - 4), q = pair_j(w - 4);
    sink(f0 + 3, tidx(c0 + (p == 0) + (q == 0), c1 + (p == 1) + (q == 1), c2 + (p == 2) + (q == 2)), s);
    if (p == q) sink(f0 + 2, tidx(c0, c1, c2), s);
  }
}

// Transpose of Dten() with e = 0: s times each table entry's weight in D(k,l | m,n | w) to sink
template <class Sink>
IBEM_HD void Dten_adj(Sink& sink, const Mat& M, int w, int k, int l, int m, int n, double s) {
#define IBEM_C3(a, b) ((a) == 0) + ((b) == 0), ((a) == 1) + ((b) == 1), ((a) == 2) + ((b) == 2)
  const double t = M.cD * s, u = -(1.0 - M.nu0) * t;
  pot_adj(sink, 1, w, (m == 0) + (n == 0) + (k == 0) + (l == 0), (m == 1) + (n == 1) + (k == 1) + (l == 1),
          (m == 2) + (n == 2) + (k == 2) + (l == 2), t);
  if (m == n) pot_adj(sink, 0, w, IBEM_C3(k, l), -2.0 * M.nu0 * t);
  if (k == n) pot_adj(sink, 0, w, IBEM_C3(m, l), u);
  if (l == n) pot_adj(sink, 0, w, IBEM_C3(m, k), u);
  if (k == m) pot_adj(sink, 0, w, IBEM_C3(n, l), u);
  if (l == m) pot_adj(sink, 0, w, IBEM_C3(n, k), u);
#undef IBEM_C3
}

// Row / column decode of the 60-unknown sphere block:
//   0..5   E_ij            (t = pair)
//   6..23  E_ij,k          6 + 3 t + k
//   24..59 E_ij,kl, k<=l   24 + 6 t + pair(k,l)
struct Slot {
  int order;  // 0, 1, 2
  int i, j;   // strain pair
  int e[3];   // extra derivative counts (rows) / weight indices (columns)
  int p, q;   // extra indices (p only for order 1)
  int w;      // weight id for columns
};
IBEM_HD Slot slot60(int r) {
  // components set by comparisons (an indexed store into e[] would put the struct in scratch)
  Slot s;
  s.p = s.q = -1;
  if (r < 6) {
    s.order = 0;
    s.i = pair_i(r);
    s.j = pair_j(r);
    s.w = 0;
  } else if (r < 24) {
    s.order = 1;
    const int t = (r - 6) / 3;
    s.i = pair_i(t);
    s.j = pair_j(t);
    s.p = (r - 6) % 3;
    s.w = 1 + s.p;
  } else {
    s.order = 2;
    const int t = (r - 24) / 6, kl = (r - 24) % 6;
    s.i = pair_i(t);
    s.j = pair_j(t);
    s.p = pair_i(kl);
    s.q = pair_j(kl);
    s.w = 4 + kl;
  }
  s.e[0] = (s.p == 0) + (s.q == 0);
  s.e[1] = (s.p == 1) + (s.q == 1);
  s.e[2] = (s.p == 2) + (s.q == 2);
  return s;
}

// One entry of the inclusion-inclusion block (target sphere s rows, source sphere h columns) in
// the 60-unknown layout. The reference builds it in three passes (integratorElastic3D.cpp:341-682):
//   row order 0: A = map_mn ( -sum_kl (C0-C1)_ijkl D_40/50/60 + d_sh C0_ijmn [col order 0] )
//   row order 1: A = map_mn ( -sum_kl (C0-C1)_ijkl D_41/51/61 + d_sh C0_ijmn d_pr [col order 1] )
//   row order 2: A = map_mn ( -0.5 sum_kl (C0-C1)_ijkl D_42/52/62 + d_sh C0_ijmn d_pr d_qz )
// Merging the 78-layout columns (p,q) and (q,p) of a quadratic unknown doubles the D part for
// p < q; keeping only rows r <= z of the symmetric second-order rows leaves the identity intact.
IBEM_HD double incl_entry(const double* tab, int stride, const Mat& M, bool self, int r, int c) {
  const Slot R = slot60(r);
  const Slot Cc = slot60(c);
  const int m = Cc.i, n = Cc.j, i = R.i, j = R.j;
  const double mp = mapw(m, n);
  double sum = 2.0 * M.dmu * Dten(tab, stride, M, Cc.w, i, j, m, n, R.e[0], R.e[1], R.e[2]);
  if (i == j) sum += M.dlam * Dtrace(tab, stride, M, Cc.w, m, n, R.e[0], R.e[1], R.e[2]);
  const double fac = (R.order == 2 ? 0.5 : 1.0) * ((Cc.order == 2 && Cc.p != Cc.q) ? 2.0 : 1.0);
  double v = -fac * sum;
  if (self && R.order == Cc.order && R.p == Cc.p && R.q == Cc.q) v += C0(M, i, j, m, n);
  return mp * v;
}

// ---- node x sphere columns (addFieldToBEM, integratorElastic3D.cpp:135-340) -------------------
// H(3 node + i, sphere column) = map sum_{jj,ll} C0(jj,ll,m,n) IGreen_w(i,jj,ll) (x 2 for merged
// quadratic columns), IGreen_w(i,j,l) = cg1 (d_ij Phi^w,l + d_il Phi^w,j)/2 - cg2 Psi^w,ijl
// (integratorElastic3DHelper.cpp:2604-2624). Potentials evaluated directly (no table).
IBEM_HD double pot_direct(const RadPoint& P, bool inside, const double C[8][4], int psi, int w,
                          int c0, int c1, int c2) {
  auto radial = [&](int f, int b0, int b1, int b2) {
    double B0, B1, B2, B3;
    rad_basis(P, inside, b0, b1, b2, B0, B1, B2, B3);
    double r = 0.0;
#pragma unroll
    for (int q = 0; q < 8; ++q)
      if (q == f) r = C[q][0] * B0 + C[q][1] * B1 + C[q][2] * B2 + C[q][3] * B3;
    return r;
  };
  const int base = psi ? 4 : 0;
  if (w == 0) return radial(base + 0, c0, c1, c2);
  if (w < 4) {
    const int p = w - 1;
    return radial(base + 1, c0 + (p == 0), c1 + (p == 1), c2 + (p == 2));
  }
  const int p = pair_i(w - 4), q = pair_j(w - 4);
  double v = radial(base + 3, c0 + (p == 0) + (q == 0), c1 + (p == 1) + (q == 1),
                    c2 + (p == 2) + (q == 2));
  if (p == q) v += radial(base + 2, c0, c1, c2);
  return v;
}

IBEM_HD double field_entry(const RadPoint& P, bool inside, const double C[8][4], const Mat& M,
                           int i, int col) {
  const Slot Cc = slot60(col);
  const int m = Cc.i, n = Cc.j, w = Cc.w;
  auto phi1 = [&](int a) { return pot_direct(P, inside, C, 0, w, a == 0, a == 1, a == 2); };
  auto psi3 = [&](int a, int b, int d) {
    return pot_direct(P, inside, C, 1, w, (a == 0) + (b == 0) + (d == 0), (a == 1) + (b == 1) + (d == 1),
                      (a == 2) + (b == 2) + (d == 2));
  };
  // sum_{jj,ll} C0(jj,ll,m,n) IG(i,jj,ll) = lam0 d_mn sum_k IG(i,k,k) + mu0 (IG(i,m,n) + IG(i,n,m))
  double v = 0.0;
  if (m == n) {
    double tr = M.cg1 * phi1(i);
    for (int k = 0; k < 3; ++k) tr -= M.cg2 * psi3(i, k, k);
    v += M.lam0 * tr;
  }
  double s = -2.0 * M.cg2 * psi3(i, m, n);
  if (i == m) s += M.cg1 * phi1(n);
  if (i == n) s += M.cg1 * phi1(m);
  v += M.mu0 * s;
  const double fac = (Cc.order == 2 && Cc.p != Cc.q) ? 2.0 : 1.0;
  return mapw(m, n) * fac * v;
}

// The same block from a per-(node, sphere) table of the radial values it uses. field_entry asks
// pot_direct for radial(f, b) = sum_k C[f][k] B_k(b) at 78 distinct (function, multi-index) pairs
// (f0, f2 at |b| = 1; f1 at 2; f3, f4, f6 at 3; f5 at 4; f7 at 5) and recomputes each many times
// (~1,400 basis evaluations per block against 55 distinct multi-indices). The table holds each
// value once, computed by the same expression; field_entry_tab combines them as field_entry does.
constexpr int kFieldTab = 78;
constexpr int kFieldMi = 55;  // multi-indices of orders 1..5
IBEM_HD int field_tab_order(int f) {
  return f == 0 || f == 2 ? 1 : f == 1 ? 2 : f == 5 ? 4 : f == 7 ? 5 : 3;
}
IBEM_HD int field_tab_off(int f) {
  return f == 0 ? 0 : f == 1 ? 3 : f == 2 ? 9 : f == 3 ? 12 : f == 4 ? 22 : f == 5 ? 32 : f == 6 ? 47 : 57;
}
// position of multi-index b among those of its order: b1 + b2 = k, index k (k + 1) / 2 + b2
IBEM_HD int mi_index(int b0, int b1, int b2) {
  (void)b0;
  const int k = b1 + b2;
  return k * (k + 1) / 2 + b2;
}
// multi-index number u (orders 1..5, in order) -> order and b
IBEM_HD void mi_decode(int u, int& o, int& b0, int& b1, int& b2) {
  o = u < 3 ? 1 : u < 9 ? 2 : u < 19 ? 3 : u < 34 ? 4 : 5;
  const int idx = u - (o == 1 ? 0 : o == 2 ? 3 : o == 3 ? 9 : o == 4 ? 19 : 34);
  int k = 0;
  while ((k + 1) * (k + 2) / 2 <= idx) ++k;
  b2 = idx - k * (k + 1) / 2;
  b1 = k - b2;
  b0 = o - k;
}
// table entries of multi-index number u
IBEM_HD void field_tab_fill(const RadPoint& P, bool inside, const double C[8][4], int u, double* tab) {
  int o, b0, b1, b2;
  mi_decode(u, o, b0, b1, b2);
  double B0, B1, B2, B3;
  rad_basis(P, inside, b0, b1, b2, B0, B1, B2, B3);
  const int idx = mi_index(b0, b1, b2);
#pragma unroll
  for (int f = 0; f < 8; ++f)
    if (field_tab_order(f) == o)
      tab[field_tab_off(f) + idx] = C[f][0] * B0 + C[f][1] * B1 + C[f][2] * B2 + C[f][3] * B3;
}
IBEM_HD double pot_tab(const double* tab, int psi, int w, int c0, int c1, int c2) {
  auto T = [&](int f, int b0, int b1, int b2) { return tab[field_tab_off(f) + mi_index(b0, b1, b2)]; };
  const int base = psi ? 4 : 0;
  if (w == 0) return T(base + 0, c0, c1, c2);
  if (w < 4) {
    const int p = w - 1;
    return T(base + 1, c0 + (p == 0), c1 + (p == 1), c2 + (p == 2));
  }
  const int p = pair_i(w - 4), q = pair_j(w - 4);
  double v = T(base + 3, c0 + (p == 0) + (q == 0), c1 + (p == 1) + (q == 1), c2 + (p == 2) + (q == 2));
  if (p == q) v += T(base + 2, c0, c1, c2);
  return v;
}
IBEM_HD double field_entry_tab(const double* tab, const Mat& M, int i, int col) {
  const Slot Cc = slot60(col);
  const int m = Cc.i, n = Cc.j, w = Cc.w;
  auto phi1 = [&](int a) { return pot_tab(tab, 0, w, a == 0, a == 1, a == 2); };
  auto psi3 = [&](int a, int b, int d) {
    return pot_tab(tab, 1, w, (a == 0) + (b == 0) + (d == 0), (a == 1) + (b == 1) + (d == 1),
                   (a == 2) + (b == 2) + (d == 2));
  };
  double v = 0.0;
  if (m == n) {
    double tr = M.cg1 * phi1(i);
    for (int k = 0; k < 3; ++k) tr -= M.cg2 * psi3(i, k, k);
    v += M.lam0 * tr;
  }
  double s = -2.0 * M.cg2 * psi3(i, m, n);
  if (i == m) s += M.cg1 * phi1(n);
  if (i == n) s += M.cg1 * phi1(m);
  v += M.mu0 * s;
  const double fac = (Cc.order == 2 && Cc.p != Cc.q) ? 2.0 : 1.0;
  return mapw(m, n) * fac * v;
}

// ---- boundary element geometry (BEMElasticHelper3D.cpp) ---------------------------------------
// 8-point Gauss-Legendre table as printed in BEMBuilderElastic3D.cpp:1203-1222
IBEM_HD double gauss8_x(int a) {
  static constexpr double g[8] = {0.960289856497536, -0.960289856497536, 0.796666477413627, -0.796666477413627,
                       0.525532409916329, -0.525532409916329, 0.183434642495650, -0.183434642495650};
  return g[a];
}
IBEM_HD double gauss8_w(int a) {
  static constexpr double w[8] = {0.101228536290376, 0.101228536290376, 0.222381034453374, 0.222381034453374,
                       0.313706645877887, 0.313706645877887, 0.362683783378362, 0.362683783378362};
  return w[a];
}

// Geometry of one quadrature point of a 4-node bilinear element: position, unit normal
// (d/dxi x d/deta), shape values, Jacobian (Serendip_func :191, Normal_Jac :218, Cartesian :264).
struct QPoint {
  double g[3], nv[3], N[4], jac;
};
IBEM_HD void qpoint(const double* el, double xi, double eta, QPoint& Q) {
  Q.N[0] = 0.25 * (1 - xi) * (1 - eta);
  Q.N[1] = 0.25 * (1 + xi) * (1 - eta);
  Q.N[2] = 0.25 * (1 + xi) * (1 + eta);
  Q.N[3] = 0.25 * (1 - xi) * (1 + eta);
  const double dx[4] = {-0.25 * (1 - eta), 0.25 * (1 - eta), 0.25 * (1 + eta), -0.25 * (1 + eta)};
  const double de[4] = {-0.25 * (1 - xi), -0.25 * (1 + xi), 0.25 * (1 + xi), 0.25 * (1 - xi)};
  double vx[3] = {0, 0, 0}, ve[3] = {0, 0, 0};
  for (int c = 0; c < 3; ++c) {
    Q.g[c] = 0.0;
    for (int q = 0; q < 4; ++q) {
      vx[c] += dx[q] * el[3 * q + c];
      ve[c] += de[q] * el[3 * q + c];
      Q.g[c] += Q.N[q] * el[3 * q + c];
    }
  }
  double n0 = vx[1] * ve[2] - ve[1] * vx[2];
  double n1 = vx[2] * ve[0] - vx[0] * ve[2];
  double n2 = vx[0] * ve[1] - vx[1] * ve[0];
  Q.jac = sqrt(n0 * n0 + n1 * n1 + n2 * n2);
  Q.nv[0] = n0 / Q.jac;
  Q.nv[1] = n1 / Q.jac;
  Q.nv[2] = n2 / Q.jac;
}

// Collapsed-triangle (Duffy) map of the quad at singular corner lpi, triangle ntr
// (Trans_Tri :29-100, Triangle_Coord :102-188): returns (xi, eta) and the map's Jacobian.
IBEM_HD void duffy(int ntr, int lpi, double xb, double eb, double& xi, double& eta, double& jb) {
  static constexpr int corners[2][4][3] = {{{1, 2, 0}, {2, 3, 1}, {0, 1, 2}, {0, 1, 3}},
                                {{2, 3, 0}, {3, 0, 1}, {3, 0, 2}, {1, 2, 3}}};
  static constexpr double XC[4] = {-1, 1, 1, -1}, EC[4] = {-1, -1, 1, 1};
  const double Nb[3] = {0.25 * (1 + xb) * (1 - eb), 0.25 * (1 + xb) * (1 + eb), 0.5 * (1 - xb)};
  const double dNx[3] = {0.25 * (1 - eb), 0.25 * (1 + eb), -0.5};
  const double dNe[3] = {-0.25 * (1 + xb), 0.25 * (1 + xb), 0.0};
  xi = eta = 0.0;
  double J00 = 0, J01 = 0, J10 = 0, J11 = 0;
  for (int q = 0; q < 3; ++q) {
    const int c = corners[ntr][lpi][q];
    xi += Nb[q] * XC[c];
    eta += Nb[q] * EC[c];
    J00 += dNx[q] * XC[c];
    J01 += dNx[q] * EC[c];
    J10 += dNe[q] * XC[c];
    J11 += dNe[q] * EC[c];
  }
  jb = J00 * J11 - J01 * J10;
}

// Kelvin displacement / traction kernels U*, T* (UK :276-294, TK :297-321)
IBEM_HD void kelvin_ut(const double* dxr, double r, double E, double nu, const double* nv,
                       double UP[3][3], double TP[3][3]) {
  const double G = E / (2 * (1 + nu));
  const double conr = 1.0 / (16 * kPi * G * (1 - nu)) / r;
  const double c1 = 3 - 4 * nu, c3 = 1 - 2 * nu;
  const double Conr = 1.0 / (8 * kPi * (1 - nu)) / (r * r);
  const double cth = nv[0] * dxr[0] + nv[1] * dxr[1] + nv[2] * dxr[2];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      UP[i][j] = conr * (c1 * (i == j) + dxr[i] * dxr[j]);
      TP[i][j] = (i == j) ? -Conr * (c3 + 3 * dxr[i] * dxr[i]) * cth
                          : -Conr * (3 * dxr[i] * dxr[j] * cth - c3 * (nv[j] * dxr[i] - nv[i] * dxr[j]));
    }
}

// ---- inclusion-inclusion block from per-pair weighted tables -----------------------------------
// The same entries as incl_entry, regrouped so that a row's table positions are computed once:
//   PS_w[g] = Psi^w,g  (|g| <= 6)   and   PH_w[g] = Phi^w,g  (|g| <= 4)
// for the ten weights w, built per pair from the radial basis tables; then, for row (i,j,e) and
// column (m,n,w),
//   D(i,j) = cD [ PS_w(e+i+j+m+n) - 2 nu d_mn PH_w(e+i+j)
//                 - (1-nu) (d_in PH_w(e+m+j) + d_jn PH_w(e+m+i) + d_im PH_w(e+n+j) + d_jm PH_w(e+n+i)) ]
//   sum_k D(k,k) = cD [ sum_k PS_w(e+m+n+2k) - 2 nu d_mn sum_k PH_w(e+2k) - 4 (1-nu) PH_w(e+m+n) ]
constexpr int kPS = 84, kPH = 35, kWgt = 10;
constexpr int kInclW = kWgt * (kPS + kPH);  // weighted entries per pair

// Basis table value of radial function f at multi-index (b0,b1,b2): basis is [4][kTabN].
IBEM_HD double fbasis_t(const double* basis, const double (*C)[4], int f, int t) {
  return C[f][0] * basis[t] + C[f][1] * basis[kTabN + t] + C[f][2] * basis[2 * kTabN + t] +
         C[f][3] * basis[3 * kTabN + t];
}
IBEM_HD double fbasis(const double* basis, const double (*C)[4], int f, int b0, int b1, int b2) {
  return fbasis_t(basis, C, f, tidx(b0, b1, b2));
}

// One weighted entry u (0 <= u < kInclW): PS[u] for u < kWgt kPS, else PH[u - kWgt kPS].
IBEM_HD double incl_weighted_val(const double* basis, const double (*C)[4], int u) {
  const bool psi = u < kWgt * kPS;
  const int v = psi ? u : u - kWgt * kPS;
  const int per = psi ? kPS : kPH;
  const int w = v / per, g = v % per;
  int c0, c1, c2;
  tdecode(g, c0, c1, c2);
  const int base = psi ? 4 : 0;
  double val;
  if (w == 0) {
    val = fbasis(basis, C, base, c0, c1, c2);
  } else if (w < 4) {
    const int p = w - 1;
    val = fbasis(basis, C, base + 1, c0 + (p == 0), c1 + (p == 1), c2 + (p == 2));
  } else {
    const int p = pair_i(w - 4), q = pair_j(w - 4);
    val = fbasis(basis, C, base + 3, c0 + (p == 0) + (q == 0), c1 + (p == 1) + (q == 1),
                 c2 + (p == 2) + (q == 2));
    if (p == q) val += fbasis(basis, C, base + 2, c0, c1, c2);
  }
  return val;
}
// The same value through a descriptor of u that does not depend on the pair (incl_weighted_desc):
// basis position t1 of function f1, and for p == q second-order weights the second term t2, f2.
IBEM_HD int incl_weighted_desc(int u) {
  const bool psi = u < kWgt * kPS;
  const int v = psi ? u : u - kWgt * kPS;
  const int per = psi ? kPS : kPH;
  const int w = v / per, g = v % per;
  int c0, c1, c2;
  tdecode(g, c0, c1, c2);
  const int base = psi ? 4 : 0;
  if (w == 0) return tidx(c0, c1, c2) | base << 8;
  if (w < 4) {
    const int p = w - 1;
    return tidx(c0 + (p == 0), c1 + (p == 1), c2 + (p == 2)) | (base + 1) << 8;
  }
  const int p = pair_i(w - 4), q = pair_j(w - 4);
  const int d = tidx(c0 + (p == 0) + (q == 0), c1 + (p == 1) + (q == 1), c2 + (p == 2) + (q == 2)) | (base + 3) << 8;
  return p == q ? d | 1 << 11 | tidx(c0, c1, c2) << 12 | (base + 2) << 20 : d;
}
IBEM_HD double incl_weighted_val_d(const double* basis, const double (*C)[4], int d) {
  const double val = fbasis_t(basis, C, (d >> 8) & 7, d & 255);
  const bool two = (d >> 11) & 1;
  const double val2 = fbasis_t(basis, C, two ? (d >> 20) & 7 : 0, two ? (d >> 12) & 255 : 0);
  return two ? val + val2 : val;
}

IBEM_HD void incl_weighted(const double* basis, const double (*C)[4], int u, double* PS, double* PH) {
  const double val = incl_weighted_val(basis, C, u);
  if (u < kWgt * kPS)
    PS[u] = val;
  else
    PH[u - kWgt * kPS] = val;
}

// Table positions of one row (i, j, e), computed once and reused for all 60 columns.
struct InclRow {
  int iS[6];    // e + i + j + mn          (PS), mn in pair order
  int iT[18];   // e + mn + 2k             (PS), 3 mn + k
  int iXi[3];   // e + x + i               (PH)
  int iXj[3];   // e + x + j               (PH)
  int imn[6];   // e + mn                  (PH)
  int i2k[3];   // e + 2k                  (PH)
  int iij;      // e + i + j               (PH)
  int i, j, order, p, q;
  double fac;   // 0.5 for second-order rows
};
IBEM_HD InclRow incl_row(int r) {
  const Slot R = slot60(r);
  InclRow o;
  o.i = R.i;
  o.j = R.j;
  o.order = R.order;
  o.p = R.p;
  o.q = R.q;
  o.fac = R.order == 2 ? 0.5 : 1.0;
  const int e0 = R.e[0], e1 = R.e[1], e2 = R.e[2];
  const int i = R.i, j = R.j;
  const int a0 = e0 + (i == 0) + (j == 0), a1 = e1 + (i == 1) + (j == 1), a2 = e2 + (i == 2) + (j == 2);
  o.iij = tidx(a0, a1, a2);
#pragma unroll
  for (int u = 0; u < 6; ++u) {
    const int m = pair_i(u), n = pair_j(u);
    const int d0 = (m == 0) + (n == 0), d1 = (m == 1) + (n == 1), d2 = (m == 2) + (n == 2);
    o.iS[u] = tidx(a0 + d0, a1 + d1, a2 + d2);
    o.imn[u] = tidx(e0 + d0, e1 + d1, e2 + d2);
#pragma unroll
    for (int k = 0; k < 3; ++k)
      o.iT[3 * u + k] = tidx(e0 + d0 + 2 * (k == 0), e1 + d1 + 2 * (k == 1), e2 + d2 + 2 * (k == 2));
  }
#pragma unroll
  for (int x = 0; x < 3; ++x) {
    o.iXi[x] = tidx(e0 + (x == 0) + (i == 0), e1 + (x == 1) + (i == 1), e2 + (x == 2) + (i == 2));
    o.iXj[x] = tidx(e0 + (x == 0) + (j == 0), e1 + (x == 1) + (j == 1), e2 + (x == 2) + (j == 2));
    o.i2k[x] = tidx(e0 + 2 * (x == 0), e1 + 2 * (x == 1), e2 + 2 * (x == 2));
  }
  return o;
}

// select a[k] (k < 6 / k < 3) by comparisons, so a runtime k never indexes a private array
IBEM_HD int sel6i(const int* a, int k) {
  return k == 0 ? a[0] : (k == 1 ? a[1] : (k == 2 ? a[2] : (k == 3 ? a[3] : (k == 4 ? a[4] : a[5]))));
}
IBEM_HD int sel3i(const int* a, int k) { return k == 0 ? a[0] : (k == 1 ? a[1] : a[2]); }

// Table positions of one row for one eigenstrain pair (m,n) = pair mn, shared by its ten weights.
struct InclCol {
  int m, n, iS, iT0, iT1, iT2, imn, iXjm, iXim, iXjn, iXin;
};
IBEM_HD InclCol incl_col(const InclRow& R, int mn) {
  InclCol K;
  K.m = pair_i(mn);
  K.n = pair_j(mn);
  K.iS = sel6i(R.iS, mn);
  K.imn = sel6i(R.imn, mn);
  K.iT0 = mn == 0 ? R.iT[0] : (mn == 1 ? R.iT[3] : (mn == 2 ? R.iT[6] : (mn == 3 ? R.iT[9] : (mn == 4 ? R.iT[12] : R.iT[15]))));
  K.iT1 = mn == 0 ? R.iT[1] : (mn == 1 ? R.iT[4] : (mn == 2 ? R.iT[7] : (mn == 3 ? R.iT[10] : (mn == 4 ? R.iT[13] : R.iT[16]))));
  K.iT2 = mn == 0 ? R.iT[2] : (mn == 1 ? R.iT[5] : (mn == 2 ? R.iT[8] : (mn == 3 ? R.iT[11] : (mn == 4 ? R.iT[14] : R.iT[17]))));
  K.iXjm = sel3i(R.iXj, K.m);
  K.iXim = sel3i(R.iXi, K.m);
  K.iXjn = sel3i(R.iXj, K.n);
  K.iXin = sel3i(R.iXi, K.n);
  return K;
}

// column of the 60-unknown layout for eigenstrain pair mn and weight w
IBEM_HD int incl_colidx(int mn, int w) {
  return w == 0 ? mn : (w < 4 ? 6 + 3 * mn + (w - 1) : 24 + 6 * mn + (w - 4));
}

// Entry (row R, pair mn via K, weight w); equals incl_entry(r, incl_colidx(mn, w)).
IBEM_HD double incl_entry_mw(const InclRow& R, const InclCol& K, const double* PS, const double* PH,
                             const Mat& M, bool self, int w) {
  const int m = K.m, n = K.n, i = R.i, j = R.j;
  const double* ps = PS + w * kPS;
  const double* ph = PH + w * kPH;
  const double nu = M.nu0;
  double D = ps[K.iS];
  if (m == n) D -= 2.0 * nu * ph[R.iij];
  double t = 0.0;
  if (i == n) t += ph[K.iXjm];
  if (j == n) t += ph[K.iXim];
  if (i == m) t += ph[K.iXjn];
  if (j == m) t += ph[K.iXin];
  D -= (1.0 - nu) * t;
  double sum = 2.0 * M.dmu * M.cD * D;
  if (i == j) {
    double tr = ps[K.iT0] + ps[K.iT1] + ps[K.iT2];
    if (m == n) tr -= 2.0 * nu * (ph[R.i2k[0]] + ph[R.i2k[1]] + ph[R.i2k[2]]);
    tr -= 4.0 * (1.0 - nu) * ph[K.imn];
    sum += M.dlam * M.cD * tr;
  }
  // column slot of weight w: order 0 (w = 0), 1 (w = 1..3, p = w - 1), 2 (w = 4.., p<=q pair w - 4)
  const int corder = w == 0 ? 0 : (w < 4 ? 1 : 2);
  const int cp = w == 0 ? -1 : (w < 4 ? w - 1 : pair_i(w - 4));
  const int cq = w < 4 ? -1 : pair_j(w - 4);
  const double fac = R.fac * ((corder == 2 && cp != cq) ? 2.0 : 1.0);
  double v = -fac * sum;
  if (self && R.order == corder && R.p == cp && R.q == cq) v += C0(M, i, j, m, n);
  return mapw(m, n) * v;
}

// InclRow's 40 table positions (each < 128) packed four to an int, and (i, j, order, p, q).
struct InclRowP {
  int po[10];  // iS[6] iT[18] iXi[3] iXj[3] imn[6] i2k[3] iij
  int meta;    // i | j << 2 | order << 4 | (p + 1) << 6 | (q + 1) << 8
};

IBEM_HD InclRowP incl_row_pack(const InclRow& R) {
  int o[40];
  for (int u = 0; u < 6; ++u) {
    o[u] = R.iS[u];
    o[30 + u] = R.imn[u];
  }
  for (int u = 0; u < 18; ++u) o[6 + u] = R.iT[u];
  for (int x = 0; x < 3; ++x) {
    o[24 + x] = R.iXi[x];
    o[27 + x] = R.iXj[x];
    o[36 + x] = R.i2k[x];
  }
  o[39] = R.iij;
  InclRowP p;
  for (int k = 0; k < 10; ++k) p.po[k] = o[4 * k] | o[4 * k + 1] << 8 | o[4 * k + 2] << 16 | o[4 * k + 3] << 24;
  p.meta = R.i | R.j << 2 | R.order << 4 | (R.p + 1) << 6 | (R.q + 1) << 8;
  return p;
}

// All 60 entries of row P of an inclusion block, written to Acol[c * lda] (c = the 60-layout
// column): the same values and the same order of operations as incl_entry_mw, with the column loop
// unrolled so that (m, n, w) are constants, and every term read from the tables and then kept or
// dropped by a select, so that the row-dependent conditions (i == n, ...) cost no branches.
IBEM_HD void incl_row_entries(const InclRowP& P, const double* PS, const double* PH, const Mat& M,
                              bool self, double* Acol, long lda) {
  auto at = [&](int k) { return (P.po[k >> 2] >> (8 * (k & 3))) & 255; };
  const int i = P.meta & 3, j = (P.meta >> 2) & 3, order = (P.meta >> 4) & 3;
  const int rp = ((P.meta >> 6) & 3) - 1, rq = ((P.meta >> 8) & 3) - 1;
  const double rfac = order == 2 ? 0.5 : 1.0;
  const double nu = M.nu0;
  const bool dij = i == j;
#pragma unroll
  for (int mn = 0; mn < 6; ++mn) {
    const int m = pair_i(mn), n = pair_j(mn);
    const bool din = i == n, djn = j == n, dim = i == m, djm = j == m;
    const double c0 = C0(M, i, j, m, n);
    const int oS = at(mn), oT0 = at(6 + 3 * mn), oT1 = at(7 + 3 * mn), oT2 = at(8 + 3 * mn);
    const int oXjm = at(27 + m), oXim = at(24 + m), oXjn = at(27 + n), oXin = at(24 + n);
    const int omn = at(30 + mn), oij = at(39);
#pragma unroll
    for (int w = 0; w < kWgt; ++w) {
      const double* ps = PS + w * kPS;
      const double* ph = PH + w * kPH;
      double D = ps[oS];
      if (m == n) D -= 2.0 * nu * ph[oij];
      const double a0 = ph[oXjm], a1 = ph[oXim], a2 = ph[oXjn], a3 = ph[oXin];
      double t = 0.0;
      t += din ? a0 : 0.0;
      t += djn ? a1 : 0.0;
      t += dim ? a2 : 0.0;
      t += djm ? a3 : 0.0;
      D -= (1.0 - nu) * t;
      double sum = 2.0 * M.dmu * M.cD * D;
      double tr = ps[oT0] + ps[oT1] + ps[oT2];
      if (m == n) tr -= 2.0 * nu * (ph[at(36)] + ph[at(37)] + ph[at(38)]);
      tr -= 4.0 * (1.0 - nu) * ph[omn];
      sum += dij ? M.dlam * M.cD * tr : 0.0;
      // the column's order and (p, q): p = w - 1 for first order, the pair w - 4 for second
      const int corder = w == 0 ? 0 : (w < 4 ? 1 : 2);
      const int cp = w == 0 ? -1 : (w < 4 ? w - 1 : (w - 4 < 3 ? 0 : (w - 4 < 5 ? 1 : 2)));
      const int cq = w < 4 ? -1 : (w - 4 < 3 ? w - 4 : (w - 4 < 5 ? w - 6 : 2));
      const double fac = rfac * ((corder == 2 && cp != cq) ? 2.0 : 1.0);
      double v = -fac * sum;
      v += (self && order == corder && rp == cp && rq == cq) ? c0 : 0.0;
      Acol[(long)incl_colidx(mn, w) * lda] = mapw(m, n) * v;
    }
    IBEM_SCHED_FENCE();  // one pair's ten entries at a time: bounded live ranges
  }
}

// Entry (row R, column c); equals incl_entry.
IBEM_HD double incl_entry_w(const InclRow& R, const double* PS, const double* PH, const Mat& M,
                            bool self, int c) {
  const Slot Cc = slot60(c);
  const InclCol K = incl_col(R, pair_t(Cc.i, Cc.j));
  return incl_entry_mw(R, K, PS, PH, M, self, Cc.w);
}

// ---- boundary -> inclusion strain kernels (get_for_stress_BEM{,1,2}, BEMBuilderElastic3D.cpp) ---
// Kelvin tensor G_ij,beta = d_ij cg1 d^beta r^-1 - cg2 d^(beta+e_i+e_j) r. With the harmonic and
// biharmonic identities lap r^-1 = 0, lap r = 2 r^-1, the reference's contractions reduce to
//   TG1[n] = kap T(n), TG[i][j] = kap T(ij), sum_i G_mi,ni = kap T(mn), trace TG = 0,
// kap = cg1 - 2 cg2, T = d r^-1, S = d r (all with the row's extra derivatives), so
//   Dsu[kl|n] = sg sum_m nv_m [ 2 dlam mu0 kap d_kl T(mn) + 2 dmu lam0 kap d_mn T(kl)
//               + dmu mu0 ( cg1 (d_mk T(nl) + d_ml T(nk) + d_nk T(ml) + d_nl T(mk)) - 4 cg2 S(mknl) ) ]
//   Dst[kl|n] = -sg ( dlam kap d_kl T(n) + dmu ( cg1 (d_nk T(l) + d_nl T(k)) - 2 cg2 S(nkl) ) )
// with sg = -1 for first-order rows (BEMBuilderElastic3D.cpp:654-659) and +1 otherwise.
// tabS / tabT: tables of d r and d r^-1 (degree-major, tidx) at x = quadrature point - sphere
// centre, the reference's x (BEMBuilderElastic3D.cpp:800-803).
IBEM_HD void strain_kernel_n(const double* tabS, const double* tabT, const Mat& M, double dlam,
                             double dmu, const double* nv, int k, int l, const int* e, int sg, int n,
                             double& su_out, double& st_out) {
  const double kap = M.cg1 - 2.0 * M.cg2;
  auto T = [&](int a, int b) {
    int c[3] = {e[0], e[1], e[2]};
    if (a >= 0) c[a]++;
    if (b >= 0) c[b]++;
    return tabT[tidx(c[0], c[1], c[2])];
  };
  auto S4 = [&](int a, int b, int d, int f) {
    int c[3] = {e[0], e[1], e[2]};
    c[a]++;
    c[b]++;
    c[d]++;
    if (f >= 0) c[f]++;
    return tabS[tidx(c[0], c[1], c[2])];
  };
  double su = 0.0;
#pragma unroll
  for (int m = 0; m < 3; ++m) {
    double t = 0.0;
    if (k == l) t += 2.0 * dlam * M.mu0 * kap * T(m, n);
    if (m == n) t += 2.0 * dmu * M.lam0 * kap * T(k, l);
    double g = -4.0 * M.cg2 * S4(m, k, n, l);
    if (m == k) g += M.cg1 * T(n, l);
    if (m == l) g += M.cg1 * T(n, k);
    if (n == k) g += M.cg1 * T(m, l);
    if (n == l) g += M.cg1 * T(m, k);
    t += dmu * M.mu0 * g;
    su += t * nv[m];
  }
  double st = -2.0 * M.cg2 * S4(n, k, l, -1);
  if (n == k) st += M.cg1 * T(l, -1);
  if (n == l) st += M.cg1 * T(k, -1);
  st = dmu * st;
  if (k == l) st += dlam * kap * T(n, -1);
  su_out = sg * su;
  st_out = -sg * st;
}

IBEM_HD void strain_kernel(const double* tabS, const double* tabT, const Mat& M, double dlam,
                           double dmu, const double* nv, int k, int l, const int* e, int sg,
                           double* su3, double* st3) {
  for (int n = 0; n < 3; ++n) strain_kernel_n(tabS, tabT, M, dlam, dmu, nv, k, l, e, sg, n, su3[n], st3[n]);
}

// Material constants of the Kelvin kernels: cu = 1 / (16 pi G (1 - nu)), ct = 1 / (8 pi (1 - nu)),
// c1 = 3 - 4 nu, c3 = 1 - 2 nu
struct KelvinC {
  double cu, ct, c1, c3;
};
IBEM_HD KelvinC kelvin_consts(double E, double nu) {
  const double G = E / (2 * (1 + nu));
  return KelvinC{1.0 / (16 * kPi * G * (1 - nu)), 1.0 / (8 * kPi * (1 - nu)), 3 - 4 * nu, 1 - 2 * nu};
}

// Kelvin kernel row ii of U* and T* at unit direction dxr and inverse distance ir (UK/TK,
// BEMElasticHelper3D.cpp:276-321): U*_ij = cu / r (c1 d_ij + r_i r_j),
// T*_ij = -ct / r^2 ( (c3 d_ij + 3 r_i r_j) dr/dn - c3 (n_j r_i - n_i r_j) )
IBEM_HD void kelvin_ut_row(int ii, const double* dxr, double ir, const KelvinC& K, const double* nv,
                           double* up, double* tp) {
  const double conr = K.cu * ir, Conr = K.ct * ir * ir;
  const double di = ii == 0 ? dxr[0] : (ii == 1 ? dxr[1] : dxr[2]);
  const double ni = ii == 0 ? nv[0] : (ii == 1 ? nv[1] : nv[2]);
  const double cth = nv[0] * dxr[0] + nv[1] * dxr[1] + nv[2] * dxr[2];
  for (int j = 0; j < 3; ++j) {
    up[j] = conr * (K.c1 * (ii == j) + di * dxr[j]);
    tp[j] = (ii == j) ? -Conr * (K.c3 + 3 * di * di) * cth
                      : -Conr * (3 * di * dxr[j] * cth - K.c3 * (nv[j] * di - ni * dxr[j]));
  }
}

// ---- per-Gauss-point Kelvin tables for the boundary->inclusion rows ----------------------------
// Layout of one point's table (kKStride doubles): d^beta r at tidx(beta) for 3 <= |beta| <= 6,
// d^beta r^-1 at 84 + tidx(beta) for 1 <= |beta| <= 4 (other slots unused).
constexpr int kKStride = 119;

#include "kelvin_tables.inc"

// Inputs of kelvin_group at x: y_c[p] = (2 x_c)^p, PR[m] = d^m/drho^m rho^(1/2) (m <= 6),
// PT[m] = d^m/drho^m rho^(-1/2) (m <= MT).
template <int MT = 4>
IBEM_HD void kelvin_point(const double* x, double* y0, double* y1, double* y2, double* PR,
                          double* PT) {
  const double rho = x[0] * x[0] + x[1] * x[1] + x[2] * x[2];
  const double inv = 1.0 / rho, r = sqrt(rho);
  y0[0] = y1[0] = y2[0] = 1.0;
#pragma unroll
  for (int p = 1; p <= 6; ++p) {
    y0[p] = y0[p - 1] * (2.0 * x[0]);
    y1[p] = y1[p - 1] * (2.0 * x[1]);
    y2[p] = y2[p - 1] * (2.0 * x[2]);
  }
  PR[0] = r;
#pragma unroll
  for (int m = 1; m <= 6; ++m) PR[m] = PR[m - 1] * (0.5 - (m - 1)) * inv;
  PT[0] = 1.0 / r;
#pragma unroll
  for (int m = 1; m <= MT; ++m) PT[m] = PT[m - 1] * (-0.5 - (m - 1)) * inv;
}

// select t[i] (i = 0..2 / 0..5) without an indexed private array
IBEM_HD double sel3(double t0, double t1, double t2, int i) { return i == 0 ? t0 : (i == 1 ? t1 : t2); }
IBEM_HD double sel6(const double* t, int i) {
  return i == 0 ? t[0] : (i == 1 ? t[1] : (i == 2 ? t[2] : (i == 3 ? t[3] : (i == 4 ? t[4] : t[5]))));
}

// Per-row constants of one inclusion row for the strain kernels at a Gauss point (the same
// contraction as strain_kernel_n, regrouped so the three components n share their table reads):
//   V_x = sum_m nv_m T(e+m+x),  W_x = sum_m nv_m S(e+m+x+k+l),
//   su_n = sg [ A1 V_n + A2 nv_n T(e+k+l) + A3 W_n
//               + A4 (nv_k T(e+n+l) + nv_l T(e+n+k) + d_nk V_l + d_nl V_k) ]
//   st_n = -sg [ B1 S(e+n+k+l) + B2 (d_nk T(e+l) + d_nl T(e+k)) + B3 T(e+n) ]
// with A1 = d_kl 2 dlam mu0 kap, A2 = 2 dmu lam0 kap, A3 = -4 dmu mu0 cg2, A4 = dmu mu0 cg1,
// B1 = -2 dmu cg2, B2 = dmu cg1, B3 = d_kl dlam kap, kap = cg1 - 2 cg2.
struct RowTerms {
  int oT[6];   // T(e + m + n), (m,n) in pair order 00 01 02 11 12 22
  int oS[6];   // S(e + m + n + k + l)
  int oSn[3];  // S(e + n + k + l)
  int oTn[3];  // T(e + n)
  int oTnk[3], oTnl[3];  // T(e + n + k), T(e + n + l)
  int oTkl, oTk, oTl;    // T(e + k + l), T(e + k), T(e + l)
  int k, l;
  double dk[3], dl[3];   // [n == k], [n == l]
  double A1, A2, A3, A4, B1, B2, B3, sg;
};

IBEM_HD RowTerms row_terms(const Mat& M, const Slot& R) {
  RowTerms t;
  t.k = R.i;
  t.l = R.j;
  const double kap = M.cg1 - 2.0 * M.cg2;
  const bool kl = R.i == R.j;
  t.A1 = kl ? 2.0 * M.dlam * M.mu0 * kap : 0.0;
  t.A2 = 2.0 * M.dmu * M.lam0 * kap;
  t.A3 = -4.0 * M.dmu * M.mu0 * M.cg2;
  t.A4 = M.dmu * M.mu0 * M.cg1;
  t.B1 = -2.0 * M.dmu * M.cg2;
  t.B2 = M.dmu * M.cg1;
  t.B3 = kl ? M.dlam * kap : 0.0;
  t.sg = R.order == 1 ? -1.0 : 1.0;
  const int e0 = R.e[0], e1 = R.e[1], e2 = R.e[2];
  const int k0 = (R.i == 0) + (R.j == 0), k1 = (R.i == 1) + (R.j == 1), k2 = (R.i == 2) + (R.j == 2);
  for (int u = 0; u < 6; ++u) {
    const int m = pair_i(u), n = pair_j(u);
    const int a0 = e0 + (m == 0) + (n == 0), a1 = e1 + (m == 1) + (n == 1), a2 = e2 + (m == 2) + (n == 2);
    t.oT[u] = 84 + tidx(a0, a1, a2);
    t.oS[u] = tidx(a0 + k0, a1 + k1, a2 + k2);
  }
  for (int n = 0; n < 3; ++n) {
    const int a0 = e0 + (n == 0), a1 = e1 + (n == 1), a2 = e2 + (n == 2);
    t.oSn[n] = tidx(a0 + k0, a1 + k1, a2 + k2);
    t.oTn[n] = 84 + tidx(a0, a1, a2);
    t.oTnk[n] = 84 + tidx(a0 + (R.i == 0), a1 + (R.i == 1), a2 + (R.i == 2));
    t.oTnl[n] = 84 + tidx(a0 + (R.j == 0), a1 + (R.j == 1), a2 + (R.j == 2));
    t.dk[n] = n == R.i ? 1.0 : 0.0;
    t.dl[n] = n == R.j ? 1.0 : 0.0;
  }
  t.oTkl = 84 + tidx(e0 + k0, e1 + k1, e2 + k2);
  t.oTk = 84 + tidx(e0 + (R.i == 0), e1 + (R.i == 1), e2 + (R.i == 2));
  t.oTl = 84 + tidx(e0 + (R.j == 0), e1 + (R.j == 1), e2 + (R.j == 2));
  return t;
}

// RowTerms with its 27 table positions (each < 128) packed four to an int and the row's constants
// reduced to k, l and the order: 8 ints, unpacked to the same RowTerms where they are used.
struct RowTermsP {
  int po[7];
  int kl;  // k | l << 2 | order << 4
};

IBEM_HD RowTermsP row_terms_pack(const RowTerms& t, int order) {
  int o[28] = {0};
  for (int u = 0; u < 6; ++u) {
    o[u] = t.oT[u];
    o[6 + u] = t.oS[u];
  }
  for (int n = 0; n < 3; ++n) {
    o[12 + n] = t.oSn[n];
    o[15 + n] = t.oTn[n];
    o[18 + n] = t.oTnk[n];
    o[21 + n] = t.oTnl[n];
  }
  o[24] = t.oTkl;
  o[25] = t.oTk;
  o[26] = t.oTl;
  RowTermsP p;
  for (int i = 0; i < 7; ++i)
    p.po[i] = o[4 * i] | o[4 * i + 1] << 8 | o[4 * i + 2] << 16 | o[4 * i + 3] << 24;
  p.kl = t.k | t.l << 2 | order << 4;
  return p;
}

IBEM_HD RowTerms row_terms_unpack(const RowTermsP& p, const Mat& M) {
  RowTerms t;
  auto at = [&](int i) { return (p.po[i >> 2] >> (8 * (i & 3))) & 255; };
#pragma unroll
  for (int u = 0; u < 6; ++u) {
    t.oT[u] = at(u);
    t.oS[u] = at(6 + u);
  }
#pragma unroll
  for (int n = 0; n < 3; ++n) {
    t.oSn[n] = at(12 + n);
    t.oTn[n] = at(15 + n);
    t.oTnk[n] = at(18 + n);
    t.oTnl[n] = at(21 + n);
  }
  t.oTkl = at(24);
  t.oTk = at(25);
  t.oTl = at(26);
  t.k = p.kl & 3;
  t.l = (p.kl >> 2) & 3;
  const int order = p.kl >> 4;
  const double kap = M.cg1 - 2.0 * M.cg2;
  const bool kl = t.k == t.l;
  t.A1 = kl ? 2.0 * M.dlam * M.mu0 * kap : 0.0;
  t.A2 = 2.0 * M.dmu * M.lam0 * kap;
  t.A3 = -4.0 * M.dmu * M.mu0 * M.cg2;
  t.A4 = M.dmu * M.mu0 * M.cg1;
  t.B1 = -2.0 * M.dmu * M.cg2;
  t.B2 = M.dmu * M.cg1;
  t.B3 = kl ? M.dlam * kap : 0.0;
  t.sg = order == 1 ? -1.0 : 1.0;
#pragma unroll
  for (int n = 0; n < 3; ++n) {
    t.dk[n] = n == t.k ? 1.0 : 0.0;
    t.dl[n] = n == t.l ? 1.0 : 0.0;
  }
  return t;
}

// row_gp with every table value read at a precomputed position (no selects): the k- and l-shifted
// entries come from the table, and the d_nk, d_nl terms are multiplied in. Same values as row_gp.
// tab(o) returns the table value at slot o (one table, or the sum of partial tables).
template <class Tab>
IBEM_HD void row_gp2_t(const RowTerms& R, const Tab& tab, const double* nv, double* su, double* st) {
  double T6[6], S6[6], Tnk[3], Tnl[3], Sn[3], Tn[3];
#pragma unroll
  for (int u = 0; u < 6; ++u) {
    T6[u] = tab(R.oT[u]);
    S6[u] = tab(R.oS[u]);
  }
#pragma unroll
  for (int n = 0; n < 3; ++n) {
    Tnk[n] = tab(R.oTnk[n]);
    Tnl[n] = tab(R.oTnl[n]);
    Sn[n] = tab(R.oSn[n]);
    Tn[n] = tab(R.oTn[n]);
  }
  const double Tkl = tab(R.oTkl), Tk = tab(R.oTk), Tl = tab(R.oTl);
  const double V0 = nv[0] * T6[0] + nv[1] * T6[1] + nv[2] * T6[2];
  const double V1 = nv[0] * T6[1] + nv[1] * T6[3] + nv[2] * T6[4];
  const double V2 = nv[0] * T6[2] + nv[1] * T6[4] + nv[2] * T6[5];
  const double W0 = nv[0] * S6[0] + nv[1] * S6[1] + nv[2] * S6[2];
  const double W1 = nv[0] * S6[1] + nv[1] * S6[3] + nv[2] * S6[4];
  const double W2 = nv[0] * S6[2] + nv[1] * S6[4] + nv[2] * S6[5];
  // V_k = sum_m nv_m T(e+m+k), V_l likewise; nv_k, nv_l as dot products with the indicators
  const double Vk = nv[0] * Tnk[0] + nv[1] * Tnk[1] + nv[2] * Tnk[2];
  const double Vl = nv[0] * Tnl[0] + nv[1] * Tnl[1] + nv[2] * Tnl[2];
  const double nvk = nv[0] * R.dk[0] + nv[1] * R.dk[1] + nv[2] * R.dk[2];
  const double nvl = nv[0] * R.dl[0] + nv[1] * R.dl[1] + nv[2] * R.dl[2];
  const double V[3] = {V0, V1, V2}, W[3] = {W0, W1, W2};
#pragma unroll
  for (int n = 0; n < 3; ++n) {
    const double g = nvk * Tnl[n] + nvl * Tnk[n] + R.dk[n] * Vl + R.dl[n] * Vk;
    su[n] = R.sg * (R.A1 * V[n] + R.A2 * nv[n] * Tkl + R.A3 * W[n] + R.A4 * g);
    st[n] = -R.sg * (R.B1 * Sn[n] + R.B3 * Tn[n] + R.B2 * (R.dk[n] * Tl + R.dl[n] * Tk));
  }
}

struct TabOne {
  const double* p;
  IBEM_HD double operator()(int o) const { return p[o]; }
};
struct TabSum2 {
  const double* p;
  const double* q;
  IBEM_HD double operator()(int o) const { return p[o] + q[o]; }
};

IBEM_HD void row_gp2(const RowTerms& R, const double* tab, const double* nv, double* su, double* st) {
  row_gp2_t(R, TabOne{tab}, nv, su, st);
}

// Slots of the Kelvin table that the strain kernels read: d^beta r for 3 <= |beta| <= 6 at
// 10 .. 83, d^beta r^-1 for 1 <= |beta| <= 4 at 85 .. 118.
constexpr int kKUsed = 108;
IBEM_HD int kelvin_used_slot(int m) { return m < 74 ? 10 + m : 11 + m; }

// su[n], st[n] (n = 0..2) of one row at one Gauss point with unit normal nv and table tab
IBEM_HD void row_gp(const RowTerms& R, const double* tab, const double* nv, double* su, double* st) {
  double T6[6], S6[6];
#pragma unroll
  for (int u = 0; u < 6; ++u) {
    T6[u] = tab[R.oT[u]];
    S6[u] = tab[R.oS[u]];
  }
  const double Sn0 = tab[R.oSn[0]], Sn1 = tab[R.oSn[1]], Sn2 = tab[R.oSn[2]];
  const double Tn0 = tab[R.oTn[0]], Tn1 = tab[R.oTn[1]], Tn2 = tab[R.oTn[2]];
  // symmetric 3x3 from pair order: (0,0)=0 (0,1)=1 (0,2)=2 (1,1)=3 (1,2)=4 (2,2)=5
  const double V0 = nv[0] * T6[0] + nv[1] * T6[1] + nv[2] * T6[2];
  const double V1 = nv[0] * T6[1] + nv[1] * T6[3] + nv[2] * T6[4];
  const double V2 = nv[0] * T6[2] + nv[1] * T6[4] + nv[2] * T6[5];
  const double W0 = nv[0] * S6[0] + nv[1] * S6[1] + nv[2] * S6[2];
  const double W1 = nv[0] * S6[1] + nv[1] * S6[3] + nv[2] * S6[4];
  const double W2 = nv[0] * S6[2] + nv[1] * S6[4] + nv[2] * S6[5];
  const int k = R.k, l = R.l;
  const double Tkl = sel6(T6, pair_t(k, l));
  const double nvk = sel3(nv[0], nv[1], nv[2], k), nvl = sel3(nv[0], nv[1], nv[2], l);
  const double Vk = sel3(V0, V1, V2, k), Vl = sel3(V0, V1, V2, l);
  const double Tk = sel3(Tn0, Tn1, Tn2, k), Tl = sel3(Tn0, Tn1, Tn2, l);
#pragma unroll
  for (int n = 0; n < 3; ++n) {
    const double Vn = n == 0 ? V0 : (n == 1 ? V1 : V2);
    const double Wn = n == 0 ? W0 : (n == 1 ? W1 : W2);
    const double Snn = n == 0 ? Sn0 : (n == 1 ? Sn1 : Sn2);
    const double Tnn = n == 0 ? Tn0 : (n == 1 ? Tn1 : Tn2);
    const double Tnl = sel6(T6, pair_t(n, l)), Tnk = sel6(T6, pair_t(n, k));
    double g = nvk * Tnl + nvl * Tnk;
    if (n == k) g += Vl;
    if (n == l) g += Vk;
    su[n] = R.sg * (R.A1 * Vn + R.A2 * nv[n] * Tkl + R.A3 * Wn + R.A4 * g);
    double h = R.B1 * Snn + R.B3 * Tnn;
    if (n == k) h += R.B2 * Tl;
    if (n == l) h += R.B2 * Tk;
    st[n] = -R.sg * h;
  }
}

// One entry of the Kelvin tables: s = d^beta r, tt = d^beta r^-1 at table index t.
IBEM_HD void kelvin_entry(const RadPoint& P, int t, double& s, double& tt) {
  int b0, b1, b2;
  tdecode(t, b0, b1, b2);
  const PairSums M = pair_sums(P, b0, b1, b2);
  rad_half(P, b0 + b1 + b2, M, s, tt);
}

}  // namespace ibem
