/* CPU oracle for the iBEM elastic-3D spherical-inclusion path -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of the reference algorithm, used as the checker for the HIP path.
 * Paths in citations are relative to the reference root.
 *
 * How the closed forms are evaluated. The reference hand-expands every derivative of the sphere
 * potentials (integratorElastic3DHelper.cpp:41-750) and of the Kelvin tensor (:754-2600). Here each
 * potential is built from the same base quantities -- I_n(lambda) (:41), V_n = (I_n - x.x I_{n+1})/2
 * (:169), V_n,p = -x_p I_{n+1} (:182), the weighted Phi_.1 first derivatives (:279, :316, :355) and
 * the weighted Psi_.3 third derivatives (:429, :464, :529) -- as a truncated 3-variable Taylor
 * series ("jet") about the evaluation point. Higher derivatives (Phi_.2..Phi_.6, Psi_.4..Psi_.7,
 * R_0k, fR_0k) are then read off the jets exactly, instead of being expanded by hand. The interior
 * rule of the reference (lambda = 0 => every lambda-derivative term vanishes, :51-54) becomes
 * "I_n is a constant jet".
 */
#include "ibem_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define PI 3.141592653589793 /* GlobalConstant.h:26 */
static const int KD[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
/* weight of the symmetric pair (m,n), m<=n: 1 on the diagonal, 2 off it (GlobalConstant.h:28) */
static const int MAPW[3][3] = {{1, 2, 2}, {0, 1, 2}, {0, 0, 1}};
/* position of (i,j), i<=j, in the order 00 01 02 11 12 22 (configElastic3D.cpp:312-318) */
static const int TIJ[3][3] = {{0, 1, 2}, {1, 3, 4}, {2, 4, 5}};

/* ===================================================================================== jets ==== */
#define JK 6  /* truncation order */
#define NJ 84 /* number of monomials of degree <= 6 in 3 variables */

typedef struct {
  double c[NJ];
} jet;

static int g_mono[NJ][3];
static int g_deg[NJ];
static int g_idx[JK + 1][JK + 1][JK + 1];
static int g_nmul;
static short g_mul_a[1024], g_mul_b[1024], g_mul_c[1024];
static int g_init = 0;

static void jets_init(void) {
  if (g_init) return;
  int n = 0;
  for (int d = 0; d <= JK; ++d)
    for (int a = d; a >= 0; --a)
      for (int b = d - a; b >= 0; --b) {
        int c = d - a - b;
        g_mono[n][0] = a;
        g_mono[n][1] = b;
        g_mono[n][2] = c;
        g_deg[n] = d;
        g_idx[a][b][c] = n;
        ++n;
      }
  g_nmul = 0;
  for (int i = 0; i < NJ; ++i)
    for (int j = 0; j < NJ; ++j)
      if (g_deg[i] + g_deg[j] <= JK) {
        g_mul_a[g_nmul] = (short)i;
        g_mul_b[g_nmul] = (short)j;
        g_mul_c[g_nmul] = (short)g_idx[g_mono[i][0] + g_mono[j][0]][g_mono[i][1] + g_mono[j][1]]
                                       [g_mono[i][2] + g_mono[j][2]];
        ++g_nmul;
      }
  g_init = 1;
}

__attribute__((constructor)) static void oracle_ctor(void) { jets_init(); }

static void jzero(jet* r) { memset(r, 0, sizeof(jet)); }
static void jconst(jet* r, double v) {
  jzero(r);
  r->c[0] = v;
}
/* coordinate x_p about x0 */
static void jcoord(jet* r, const double* x0, int p) {
  jzero(r);
  r->c[0] = x0[p];
  r->c[g_idx[p == 0][p == 1][p == 2]] = 1.0;
}
static void jaxpy(jet* r, double a, const jet* x) { /* r += a x */
  for (int i = 0; i < NJ; ++i) r->c[i] += a * x->c[i];
}
static void jscale(jet* r, double a) {
  for (int i = 0; i < NJ; ++i) r->c[i] *= a;
}
static void jmul(jet* r, const jet* x, const jet* y) {
  jet t;
  jzero(&t);
  for (int k = 0; k < g_nmul; ++k) t.c[g_mul_c[k]] += x->c[g_mul_a[k]] * y->c[g_mul_b[k]];
  *r = t;
}
/* r = d/dx_p x (valid order drops by one) */
static void jderiv(jet* r, const jet* x, int p) {
  jet t;
  jzero(&t);
  for (int i = 0; i < NJ; ++i) {
    int a = g_mono[i][0] + (p == 0), b = g_mono[i][1] + (p == 1), c = g_mono[i][2] + (p == 2);
    if (a + b + c > JK) continue;
    t.c[i] = (double)(p == 0 ? a : (p == 1 ? b : c)) * x->c[g_idx[a][b][c]];
  }
  *r = t;
}
/* f(rho(x)) with f = rho^s, rho = |x|^2, as a jet about x0 */
static void jpow_rho(jet* r, const double* x0, double s) {
  double rho0 = x0[0] * x0[0] + x0[1] * x0[1] + x0[2] * x0[2];
  jet delta, dk;
  jzero(&delta); /* rho(x0+h) - rho0 = 2 x0.h + h.h */
  for (int p = 0; p < 3; ++p) {
    delta.c[g_idx[p == 0][p == 1][p == 2]] = 2.0 * x0[p];
    delta.c[g_idx[2 * (p == 0)][2 * (p == 1)][2 * (p == 2)]] = 1.0;
  }
  jconst(r, pow(rho0, s));
  jconst(&dk, 1.0);
  double coef = 1.0; /* s (s-1) ... (s-k+1) / k! */
  for (int k = 1; k <= JK; ++k) {
    jmul(&dk, &dk, &delta);
    coef *= (s - (k - 1)) / (double)k;
    jaxpy(r, coef * pow(rho0, s - k), &dk);
  }
}
/* derivative of order |list| read off a jet: d^alpha f = alpha! c_alpha */
static double jd(const jet* J, const int* list, int nl) {
  int al[3] = {0, 0, 0};
  for (int i = 0; i < nl; ++i) al[list[i]]++;
  double f = 1.0;
  for (int q = 0; q < 3; ++q)
    for (int k = 2; k <= al[q]; ++k) f *= k;
  return f * J->c[g_idx[al[0]][al[1]][al[2]]];
}

/* ============================================================== sphere potentials (:41-750) ==== */
/* weight index w: 0 -> 1, 1+p -> x_p, 4+TIJ[p][q] -> x_p x_q */
typedef struct {
  double x[3], a;
  int inside;
  jet X[3];           /* coordinate jets */
  jet phi[10][3];     /* Phi^w,i      (Phi_01 :279, Phi_11 :316, Phi_21 :355) */
  jet psi[10][3][3][3]; /* Psi^w,ijk  (Psi_03 :429, Psi_13 :464, Psi_23 :529) */
} sph_ctx;

static void sph_build(sph_ctx* S, const double* x, double a) {
  jets_init();
  S->x[0] = x[0];
  S->x[1] = x[1];
  S->x[2] = x[2];
  S->a = a;
  double xx = x[0] * x[0] + x[1] * x[1] + x[2] * x[2];
  /* lambda = max(|x|^2 - a^2, 0) (:136-140) */
  S->inside = !((xx - a * a) > 0);
  for (int p = 0; p < 3; ++p) jcoord(&S->X[p], x, p);
  /* I_m, m = 0..4 (:41-45): 4 pi a^3 / ((2m+1) (a^2+lambda)^(m+1/2)) */
  jet I[5];
  for (int m = 0; m < 5; ++m) {
    double cm = 4.0 * PI * a * a * a / (2.0 * m + 1.0);
    if (S->inside)
      jconst(&I[m], cm / pow(a * a, m + 0.5));
    else {
      jpow_rho(&I[m], x, -(m + 0.5));
      jscale(&I[m], cm);
    }
  }
  jet rho;
  jzero(&rho);
  for (int p = 0; p < 3; ++p) {
    jet t;
    jmul(&t, &S->X[p], &S->X[p]);
    jaxpy(&rho, 1.0, &t);
  }
  /* V_n (:169-180), V_n,p = -x_p I_{n+1} (:182-193), V_n,pq (:195-205) for n = 0..3 */
  jet V[4], Vp[4][3], Vpq[4][3][3];
  for (int n = 0; n < 4; ++n) {
    jet t;
    jmul(&t, &rho, &I[n + 1]);
    V[n] = I[n];
    jaxpy(&V[n], -1.0, &t);
    jscale(&V[n], 0.5);
    for (int p = 0; p < 3; ++p) {
      jmul(&Vp[n][p], &S->X[p], &I[n + 1]);
      jscale(&Vp[n][p], -1.0);
    }
    for (int p = 0; p < 3; ++p)
      for (int q = 0; q < 3; ++q) jderiv(&Vpq[n][p][q], &Vp[n][p], q);
  }
  const double a2 = a * a, a4 = a2 * a2;
  /* W0 = V0 - a^2 V1 (Psi_03), W1 = V1 - a^2 V2 (Psi_13), W2 = V2 - a^2 V3 and
   * W21 = V1 - 2 a^2 V2 + a^4 V3 (Psi_23), with their first and second derivatives */
  jet W[4], Wp[4][3], Wpq[4][3][3];
  const double cw[4][4] = {{1, -a2, 0, 0}, {0, 1, -a2, 0}, {0, 0, 1, -a2}, {0, 1, -2 * a2, a4}};
  for (int w = 0; w < 4; ++w) {
    jzero(&W[w]);
    for (int n = 0; n < 4; ++n)
      if (cw[w][n] != 0) jaxpy(&W[w], cw[w][n], &V[n]);
    for (int p = 0; p < 3; ++p) {
      jzero(&Wp[w][p]);
      for (int n = 0; n < 4; ++n)
        if (cw[w][n] != 0) jaxpy(&Wp[w][p], cw[w][n], &Vp[n][p]);
      for (int q = 0; q < 3; ++q) {
        jzero(&Wpq[w][p][q]);
        for (int n = 0; n < 4; ++n)
          if (cw[w][n] != 0) jaxpy(&Wpq[w][p][q], cw[w][n], &Vpq[n][p][q]);
      }
    }
  }
  jet t1, t2;
  /* ---- Phi^w,i ---- */
  for (int i = 0; i < 3; ++i) {
    S->phi[0][i] = Vp[0][i]; /* Phi_01 = V_n1(0,i) (:279-283) */
    for (int n = 0; n < 3; ++n) { /* Phi_11 = a^2 (d_in V_1 + x_n V_1,i) (:316-320) */
      jet* r = &S->phi[1 + n][i];
      jmul(r, &S->X[n], &Vp[1][i]);
      if (i == n) jaxpy(r, 1.0, &V[1]);
      jscale(r, a2);
    }
    for (int m = 0; m < 3; ++m)
      for (int n = m; n < 3; ++n) {
        /* Phi_21 = a^2 ( a^2 (d_mi x_n + d_ni x_m) V_2 + a^2 x_m x_n V_2,i
         *               - d_mn x_i (V_1 - a^2 V_2) )                       (:355-359) */
        jet* r = &S->phi[4 + TIJ[m][n]][i];
        jzero(r);
        if (m == i) {
          jmul(&t1, &S->X[n], &V[2]);
          jaxpy(r, a2, &t1);
        }
        if (n == i) {
          jmul(&t1, &S->X[m], &V[2]);
          jaxpy(r, a2, &t1);
        }
        jmul(&t1, &S->X[m], &S->X[n]);
        jmul(&t2, &t1, &Vp[2][i]);
        jaxpy(r, a2, &t2);
        if (m == n) {
          jet v = V[1];
          jaxpy(&v, -a2, &V[2]);
          jmul(&t1, &S->X[i], &v);
          jaxpy(r, -1.0, &t1);
        }
        jscale(r, a2);
      }
  }
  /* ---- Psi^w,ijk ---- */
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      for (int k = j; k < 3; ++k) {
        /* Psi_03 = d_ij W0,k + d_ik W0,j + x_i W0,jk (:429-433) */
        {
          jet* r = &S->psi[0][i][j][k];
          jmul(r, &S->X[i], &Wpq[0][j][k]);
          if (i == j) jaxpy(r, 1.0, &Wp[0][k]);
          if (i == k) jaxpy(r, 1.0, &Wp[0][j]);
        }
        /* Psi_13 (:464-471) */
        for (int n = 0; n < 3; ++n) {
          jet* r = &S->psi[1 + n][i][j][k];
          jzero(r);
          double c0 = (double)(KD[i][n] * KD[j][k] + KD[n][j] * KD[i][k] + KD[i][j] * KD[n][k]);
          if (c0 != 0) jaxpy(r, c0, &W[1]);
          if (KD[i][n]) {
            jmul(&t1, &S->X[j], &Wp[1][k]);
            jaxpy(r, 1.0, &t1);
          }
          if (KD[n][j]) {
            jmul(&t1, &S->X[i], &Wp[1][k]);
            jaxpy(r, 1.0, &t1);
          }
          if (KD[i][j]) {
            jmul(&t1, &S->X[n], &Wp[1][k]);
            jaxpy(r, 1.0, &t1);
          }
          if (KD[n][k]) {
            jmul(&t1, &S->X[i], &Wp[1][j]);
            jaxpy(r, 1.0, &t1);
          }
          if (KD[i][k]) {
            jmul(&t1, &S->X[n], &Wp[1][j]);
            jaxpy(r, 1.0, &t1);
          }
          jmul(&t1, &S->X[n], &S->X[i]);
          jmul(&t2, &t1, &Wpq[1][j][k]);
          jaxpy(r, 1.0, &t2);
          jscale(r, a2);
        }
        /* Psi_23 (:529-543) */
        for (int m = 0; m < 3; ++m)
          for (int n = m; n < 3; ++n) {
            jet* r = &S->psi[4 + TIJ[m][n]][i][j][k];
            jet acc;
            jzero(&acc);
            const int dmi = KD[m][i], dni = KD[n][i], dmj = KD[m][j], dnj = KD[n][j];
            const int dmk = KD[m][k], dnk = KD[n][k], dij = KD[i][j], dik = KD[i][k];
            const int djk = KD[j][k];
            /* scalar-delta x single-coordinate terms multiplying W2 */
            double cx[3] = {0, 0, 0}; /* coefficient of x_c W2 */
            cx[k] += (double)(dmi * dnj + dni * dmj);
            cx[j] += (double)(dmi * dnk + dni * dmk);
            cx[n] += (double)(dmi * djk);
            cx[m] += (double)(dni * djk);
            cx[i] += (double)(dmj * dnk + dnj * dmk);
            cx[n] += (double)(dmj * dik + dij * dmk);
            cx[m] += (double)(dnj * dik + dij * dnk);
            for (int c = 0; c < 3; ++c)
              if (cx[c] != 0) {
                jmul(&t1, &S->X[c], &W[2]);
                jaxpy(&acc, cx[c], &t1);
              }
            /* (d_mi x_n + d_ni x_m) x_j W2,k */
            if (dmi) {
              jmul(&t1, &S->X[n], &S->X[j]);
              jmul(&t2, &t1, &Wp[2][k]);
              jaxpy(&acc, 1.0, &t2);
            }
            if (dni) {
              jmul(&t1, &S->X[m], &S->X[j]);
              jmul(&t2, &t1, &Wp[2][k]);
              jaxpy(&acc, 1.0, &t2);
            }
            /* (d_mj x_n x_i + d_nj x_m x_i + d_ij x_m x_n) W2,k */
            if (dmj) {
              jmul(&t1, &S->X[n], &S->X[i]);
              jmul(&t2, &t1, &Wp[2][k]);
              jaxpy(&acc, 1.0, &t2);
            }
            if (dnj) {
              jmul(&t1, &S->X[m], &S->X[i]);
              jmul(&t2, &t1, &Wp[2][k]);
              jaxpy(&acc, 1.0, &t2);
            }
            if (dij) {
              jmul(&t1, &S->X[m], &S->X[n]);
              jmul(&t2, &t1, &Wp[2][k]);
              jaxpy(&acc, 1.0, &t2);
            }
            /* (d_mk x_n x_i + d_nk x_m x_i + d_ik x_m x_n) W2,j */
            if (dmk) {
              jmul(&t1, &S->X[n], &S->X[i]);
              jmul(&t2, &t1, &Wp[2][j]);
              jaxpy(&acc, 1.0, &t2);
            }
            if (dnk) {
              jmul(&t1, &S->X[m], &S->X[i]);
              jmul(&t2, &t1, &Wp[2][j]);
              jaxpy(&acc, 1.0, &t2);
            }
            if (dik) {
              jmul(&t1, &S->X[m], &S->X[n]);
              jmul(&t2, &t1, &Wp[2][j]);
              jaxpy(&acc, 1.0, &t2);
            }
            /* x_m x_n x_i W2,jk */
            jmul(&t1, &S->X[m], &S->X[n]);
            jmul(&t2, &t1, &S->X[i]);
            jmul(&t1, &t2, &Wpq[2][j][k]);
            jaxpy(&acc, 1.0, &t1);
            jscale(&acc, a4);
            if (m == n) {
              /* a^2 d_mn ( -d_ij x_k W21 - (d_ik x_j + d_jk x_i) W21 - x_i x_j W21,k ) */
              jet b;
              jzero(&b);
              double cy[3] = {0, 0, 0};
              cy[k] -= dij;
              cy[j] -= dik;
              cy[i] -= djk;
              for (int c = 0; c < 3; ++c)
                if (cy[c] != 0) {
                  jmul(&t1, &S->X[c], &W[3]);
                  jaxpy(&b, cy[c], &t1);
                }
              jmul(&t1, &S->X[i], &S->X[j]);
              jmul(&t2, &t1, &Wp[3][k]);
              jaxpy(&b, -1.0, &t2);
              jaxpy(&acc, a2, &b);
            }
            *r = acc;
          }
      }
  for (int w = 0; w < 10; ++w)
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j)
        for (int k = 0; k < j; ++k) S->psi[w][i][j][k] = S->psi[w][i][k][j];
}

static sph_ctx* sph_new(const double* x, double a) {
  sph_ctx* S = (sph_ctx*)malloc(sizeof(sph_ctx));
  sph_build(S, x, a);
  return S;
}

/* Psi^w with derivative indices (b0,b1,b2 | extra...) and Phi^w with (b0 | extra...) */
static double PSI(const sph_ctx* S, int w, int b0, int b1, int b2, const int* ex, int ne) {
  return jd(&S->psi[w][b0][b1][b2], ex, ne);
}
static double PHI(const sph_ctx* S, int w, int b0, const int* ex, int ne) {
  return jd(&S->phi[w][b0], ex, ne);
}

/* D_4x/D_5x/D_6x (:2394-2554), generic over the weight w and the trailing derivative list
 * (r) or (r,z): D = 1/(8 pi (1-nu0)) [ Psi^w,(k l i | j ext) - 2 nu0 d_kl Phi^w,(i | j ext)
 *   - (1-nu0) ( Phi^w,(k|j ext) d_il + Phi^w,(k|i ext) d_jl + Phi^w,(l|j ext) d_ik
 *               + Phi^w,(l|i ext) d_jk ) ]                                                 */
static double Dgen(const sph_ctx* S, double nu0, int w, int i, int j, int k, int l, const int* ext,
                   int ne) {
  int ej[3] = {j, ne > 0 ? ext[0] : 0, ne > 1 ? ext[1] : 0};
  int ei[3] = {i, ne > 0 ? ext[0] : 0, ne > 1 ? ext[1] : 0};
  double v = PSI(S, w, k, l, i, ej, 1 + ne);
  if (k == l) v -= 2.0 * nu0 * PHI(S, w, i, ej, 1 + ne);
  double s = 0.0;
  if (i == l) s += PHI(S, w, k, ej, 1 + ne);
  if (j == l) s += PHI(S, w, k, ei, 1 + ne);
  if (i == k) s += PHI(S, w, l, ej, 1 + ne);
  if (j == k) s += PHI(S, w, l, ei, 1 + ne);
  v -= (1.0 - nu0) * s;
  return v / (8.0 * PI * (1.0 - nu0));
}

static int weight_of(int kind, const int* g, const int** ext, int* ne) {
  switch (kind) {
    case 40: *ext = g + 4; *ne = 0; return 0;
    case 50: *ext = g + 5; *ne = 0; return 1 + g[4];
    case 60: *ext = g + 6; *ne = 0; return 4 + TIJ[g[4]][g[5]];
    case 41: *ext = g + 4; *ne = 1; return 0;
    case 51: *ext = g + 5; *ne = 1; return 1 + g[4];
    case 61: *ext = g + 6; *ne = 1; return 4 + TIJ[g[4]][g[5]];
    case 42: *ext = g + 4; *ne = 2; return 0;
    case 52: *ext = g + 5; *ne = 2; return 1 + g[4];
    case 62: *ext = g + 6; *ne = 2; return 4 + TIJ[g[4]][g[5]];
  }
  return -1;
}

double orc_D(const orc_material* m, int kind, const int* g, const double* x, double a) {
  const int* ext;
  int ne;
  int w = weight_of(kind, g, &ext, &ne);
  if (w < 0) return NAN;
  sph_ctx* S = sph_new(x, a);
  double v = Dgen(S, m->nu0, w, g[0], g[1], g[2], g[3], ext, ne);
  free(S);
  return v;
}

static double mu_of(double E, double nu) { return E / (2.0 * (1.0 + nu)); }

/* IGreen_01/11/21 (:2604-2624) on a built context */
static double igreen_ctx(const sph_ctx* S, double mu0, double nu0, int kind, const int* g) {
  int w, i = g[0], j = g[1], l;
  if (kind == 1) {
    w = 0;
    l = g[2];
  } else if (kind == 11) {
    w = 1 + g[2];
    l = g[3];
  } else {
    w = 4 + TIJ[g[2]][g[3]];
    l = g[4];
  }
  double v = 0.0;
  if (i == j) v += 0.5 * PHI(S, w, l, NULL, 0);
  if (i == l) v += 0.5 * PHI(S, w, j, NULL, 0);
  v /= (4.0 * PI * mu0);
  v -= PSI(S, w, i, j, l, NULL, 0) / (16.0 * PI * mu0 * (1.0 - nu0));
  return v;
}

double orc_igreen(const orc_material* m, int kind, const int* g, const double* x, double a) {
  sph_ctx* S = sph_new(x, a);
  double v = igreen_ctx(S, mu_of(m->E0, m->nu0), m->nu0, kind, g);
  free(S);
  return v;
}

/* ============================================================ Kelvin tensor (:754-2600) ======== */
typedef struct {
  jet inv;  /* 1/R  (fR with sign = 0, :952) */
  jet R;    /* R    (:754) */
  double cg1, cg2; /* 1/(4 pi mu0), 1/(16 pi mu0 (1-nu0)) */
} kel_ctx;

static void kel_build(kel_ctx* K, const double* x, double mu0, double nu0) {
  jets_init();
  jpow_rho(&K->inv, x, -0.5);
  jpow_rho(&K->R, x, 0.5);
  K->cg1 = 1.0 / (4.0 * PI * mu0);
  K->cg2 = 1.0 / (16.0 * PI * mu0 * (1.0 - nu0));
}

/* Green_0k: G_ij,(list) = d_ij/(4 pi mu0) (1/R),(list) - R,(ij list)/(16 pi mu0 (1-nu0)) */
static double kgreen(const kel_ctx* K, int i, int j, const int* list, int nl) {
  int full[8];
  full[0] = i;
  full[1] = j;
  for (int q = 0; q < nl; ++q) full[2 + q] = list[q];
  double v = -K->cg2 * jd(&K->R, full, nl + 2);
  if (i == j) v += K->cg1 * jd(&K->inv, list, nl);
  return v;
}

double orc_green(const orc_material* m, int order, const int* g, const double* x) {
  kel_ctx K;
  kel_build(&K, x, mu_of(m->E0, m->nu0), m->nu0);
  return kgreen(&K, g[0], g[1], g + 2, order);
}

/* ================================================== BEM element helpers (BEMElasticHelper3D) === */
/* 8-point Gauss-Legendre table exactly as printed in BEMBuilderElastic3D.cpp:1203-1222 */
static const double GP8[8] = {0.960289856497536, -0.960289856497536, 0.796666477413627,
                              -0.796666477413627, 0.525532409916329, -0.525532409916329,
                              0.183434642495650, -0.183434642495650};
static const double GW8[8] = {0.101228536290376, 0.101228536290376, 0.222381034453374,
                              0.222381034453374, 0.313706645877887, 0.313706645877887,
                              0.362683783378362, 0.362683783378362};

/* bilinear shape functions (Serendip_func :191-216) */
static void shape4(double xi, double eta, double* N) {
  N[0] = 0.25 * (1 - xi) * (1 - eta);
  N[1] = 0.25 * (1 + xi) * (1 - eta);
  N[2] = 0.25 * (1 + xi) * (1 + eta);
  N[3] = 0.25 * (1 - xi) * (1 + eta);
}
/* Jacobian and unit normal d/dxi x d/deta (Normal_Jac :218-262) */
static double normal_jac(double xi, double eta, const double* el, double* nv) {
  double dx[4] = {-0.25 * (1 - eta), 0.25 * (1 - eta), 0.25 * (1 + eta), -0.25 * (1 + eta)};
  double de[4] = {-0.25 * (1 - xi), -0.25 * (1 + xi), 0.25 * (1 + xi), 0.25 * (1 - xi)};
  double vx[3] = {0, 0, 0}, ve[3] = {0, 0, 0};
  for (int c = 0; c < 3; ++c)
    for (int q = 0; q < 4; ++q) {
      vx[c] += dx[q] * el[3 * q + c];
      ve[c] += de[q] * el[3 * q + c];
    }
  nv[0] = vx[1] * ve[2] - ve[1] * vx[2];
  nv[1] = vx[2] * ve[0] - vx[0] * ve[2];
  nv[2] = vx[0] * ve[1] - vx[1] * ve[0];
  double J = sqrt(nv[0] * nv[0] + nv[1] * nv[1] + nv[2] * nv[2]);
  nv[0] /= J;
  nv[1] /= J;
  nv[2] /= J;
  return J;
}
/* Duffy split of the quad at corner l_pi into two collapsed triangles (Trans_Tri :29-100,
 * Triangle_Coord :102-188): corner lists (ntr, l_pi) -> three quad corners, the last one the
 * singular corner. */
static const int TRI_CORNERS[2][4][3] = {
    {{1, 2, 0}, {2, 3, 1}, {0, 1, 2}, {0, 1, 3}},
    {{2, 3, 0}, {3, 0, 1}, {3, 0, 2}, {1, 2, 3}},
};
static void trans_tri(int ntr, int lpi, double xb, double eb, double* xi, double* eta,
                      double* jacb) {
  static const int XC[4] = {-1, 1, 1, -1}, EC[4] = {-1, -1, 1, 1};
  double Nb[3] = {0.25 * (1 + xb) * (1 - eb), 0.25 * (1 + xb) * (1 + eb), 0.5 * (1 - xb)};
  double dNx[3] = {0.25 * (1 - eb), 0.25 * (1 + eb), -0.5};
  double dNe[3] = {-0.25 * (1 + xb), 0.25 * (1 + xb), 0.0};
  double cx[3], ce[3];
  for (int q = 0; q < 3; ++q) {
    int c = TRI_CORNERS[ntr][lpi][q];
    cx[q] = XC[c];
    ce[q] = EC[c];
  }
  *xi = 0;
  *eta = 0;
  double J[2][2] = {{0, 0}, {0, 0}};
  for (int q = 0; q < 3; ++q) {
    *xi += Nb[q] * cx[q];
    *eta += Nb[q] * ce[q];
    J[0][0] += dNx[q] * cx[q];
    J[0][1] += dNx[q] * ce[q];
    J[1][0] += dNe[q] * cx[q];
    J[1][1] += dNe[q] * ce[q];
  }
  *jacb = J[0][0] * J[1][1] - J[0][1] * J[1][0];
}
/* Kelvin displacement and traction kernels (UK :276-294, TK :297-321) */
static void kelvin_ut(const double* dxr, double r, double E, double nu, const double* nv,
                      double UP[3][3], double TP[3][3]) {
  double G = E / (2 * (1 + nu));
  double conr = 1.0 / (16 * PI * G * (1 - nu)) / r;
  double c1 = 3 - 4 * nu;
  double c3 = 1 - 2 * nu;
  double Conr = 1.0 / (8 * PI * (1 - nu)) / (r * r);
  double cth = nv[0] * dxr[0] + nv[1] * dxr[1] + nv[2] * dxr[2];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      UP[i][j] = conr * (c1 * KD[i][j] + dxr[i] * dxr[j]);
      if (i == j)
        TP[i][j] = -Conr * (c3 + 3 * dxr[i] * dxr[i]) * cth;
      else
        TP[i][j] = -Conr * (3 * dxr[i] * dxr[j] * cth - c3 * (nv[j] * dxr[i] - nv[i] * dxr[j]));
    }
}

static void gp_accum(const double* el, double xi, double eta, double wjac_extra, const double* xP,
                     double E, double nu, int skip_node, double* dU, double* dT) {
  double N[4], nv[3], g[3] = {0, 0, 0};
  shape4(xi, eta, N);
  double J = normal_jac(xi, eta, el, nv) * wjac_extra;
  for (int q = 0; q < 4; ++q)
    for (int c = 0; c < 3; ++c) g[c] += N[q] * el[3 * q + c];
  double d[3] = {g[0] - xP[0], g[1] - xP[1], g[2] - xP[2]};
  double r = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
  double dxr[3] = {d[0] / r, d[1] / r, d[2] / r};
  double UP[3][3], TP[3][3];
  kelvin_ut(dxr, r, E, nu, nv, UP, TP);
  for (int ii = 0; ii < 3; ++ii)
    for (int jj = 0; jj < 3; ++jj)
      for (int n = 0; n < 4; ++n) {
        dU[12 * ii + 3 * n + jj] += N[n] * UP[ii][jj] * J;
        if (n != skip_node) dT[12 * ii + 3 * n + jj] += N[n] * TP[ii][jj] * J;
      }
}

void orc_integ3(const orc_material* m, const int* sctrb, const double* nodes, int nn, double* dUe,
                double* dTe) {
  double el[12];
  for (int q = 0; q < 4; ++q)
    for (int c = 0; c < 3; ++c) el[3 * q + c] = nodes[3 * (sctrb[q] - 1) + c];
  for (int i = 0; i < nn; ++i) {
    int lpi = -1;
    for (int q = 0; q < 4; ++q)
      if (sctrb[q] == i + 1) lpi = q;
    double* dU = dUe + 36 * i; /* rows 3i..3i+2 of a row-major (3nn x 12) block */
    double* dT = dTe + 36 * i;
    if (lpi < 0) {
      for (int a = 0; a < 8; ++a)
        for (int b = 0; b < 8; ++b)
          gp_accum(el, GP8[a], GP8[b], GW8[a] * GW8[b], nodes + 3 * i, m->E0, m->nu0, -1, dU, dT);
    } else {
      for (int ntr = 0; ntr < 2; ++ntr)
        for (int a = 0; a < 8; ++a)
          for (int b = 0; b < 8; ++b) {
            double xi, eta, jb;
            trans_tri(ntr, lpi, GP8[a], GP8[b], &xi, &eta, &jb);
            gp_accum(el, xi, eta, GW8[a] * GW8[b] * jb, nodes + 3 * i, m->E0, m->nu0, lpi, dU, dT);
          }
    }
  }
}

/* ===================================== boundary -> inclusion strain rows (Integ3_stress*) ====== */
/* get_for_stress_BEM{,1,2} (BEMBuilderElastic3D.cpp:780-980, 429-672, 37-322) and the postprocessor
 * variant get_for_stress_BEM_post (postProcessorElastic3D.cpp:32-114), order 0 only, with
 * (dlam, dmu) = (lam0 - lam1, mu0 - mu1), or (lam0, mu0) for the postprocessor.
 * Output Dsu/Dst: rows x 3, rows = 6 | 18 | 54 ordered (k<=l)[, w[, z]]. */
static void stress_kernel(const kel_ctx* K, int order, double lam0, double mu0, double dlam,
                          double dmu, const double* nv, double* Dsu, double* Dst) {
  int rows = order == 0 ? 6 : (order == 1 ? 18 : 54);
  int nw = order == 0 ? 1 : (order == 1 ? 3 : 9);
  memset(Dsu, 0, sizeof(double) * rows * 3);
  memset(Dst, 0, sizeof(double) * rows * 3);
  for (int wz = 0; wz < nw; ++wz) {
    int ext[2] = {order == 2 ? wz / 3 : wz, wz % 3};
    int ne = order;
    /* TG1[n] = sum_i G_ni,(i ext); TG[i][j] = sum_l G_li,(l j ext);
     * sgm[m][n] = sum_i lam0 d_mn TG[i][i] + (2 mu0 G_mi,(n i ext) | mu0 (G_mi,(n i ext) + G_ni,(m i ext))) */
    double TG1[3], TG[3][3], sgm[3][3];
    int lst[4];
    for (int n = 0; n < 3; ++n) {
      TG1[n] = 0;
      for (int i = 0; i < 3; ++i) {
        lst[0] = i;
        lst[1] = ext[0];
        lst[2] = ext[1];
        TG1[n] += kgreen(K, n, i, lst, 1 + ne);
      }
    }
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        TG[i][j] = 0;
        for (int l = 0; l < 3; ++l) {
          lst[0] = l;
          lst[1] = j;
          lst[2] = ext[0];
          lst[3] = ext[1];
          TG[i][j] += kgreen(K, l, i, lst, 2 + ne);
        }
      }
    double trTG = TG[0][0] + TG[1][1] + TG[2][2];
    for (int m = 0; m < 3; ++m)
      for (int n = 0; n < 3; ++n) {
        double s = 0;
        for (int i = 0; i < 3; ++i) {
          lst[0] = n;
          lst[1] = i;
          lst[2] = ext[0];
          lst[3] = ext[1];
          double gm = kgreen(K, m, i, lst, 2 + ne);
          if (order == 0) {
            s += 2.0 * mu0 * gm; /* BEMBuilderElastic3D.cpp:952 */
          } else {
            lst[0] = m;
            double gn = kgreen(K, n, i, lst, 2 + ne);
            s += mu0 * gm + mu0 * gn; /* :639, :286 */
          }
        }
        /* the reference adds lam0 d_mn TG[i][i] inside the i loop: lam0 d_mn trace(TG) */
        sgm[m][n] = s + (m == n ? lam0 * trTG : 0.0);
      }
    for (int k = 0; k < 3; ++k)
      for (int l = k; l < 3; ++l) {
        int t = order == 0 ? TIJ[k][l] : (order == 1 ? 3 * TIJ[k][l] + wz : 9 * TIJ[k][l] + wz);
        for (int n = 0; n < 3; ++n) {
          double su = 0;
          for (int m = 0; m < 3; ++m) {
            int a1[4] = {n, l, ext[0], ext[1]}, a2[4] = {n, k, ext[0], ext[1]};
            int a3[4] = {m, l, ext[0], ext[1]}, a4[4] = {m, k, ext[0], ext[1]};
            double gsum = kgreen(K, m, k, a1, 2 + ne) + kgreen(K, m, l, a2, 2 + ne) +
                          kgreen(K, n, k, a3, 2 + ne) + kgreen(K, n, l, a4, 2 + ne);
            double term = (k == l ? dlam * sgm[m][n] : 0.0) +
                          dmu * ((m == n ? lam0 * (TG[k][l] + TG[l][k]) : 0.0) + mu0 * gsum);
            su += term * nv[m];
          }
          int b1[3] = {l, ext[0], ext[1]}, b2[3] = {k, ext[0], ext[1]};
          double st = (k == l ? dlam * TG1[n] : 0.0) + dmu * kgreen(K, n, k, b1, 1 + ne) +
                      dmu * kgreen(K, n, l, b2, 1 + ne);
          /* signs: order 0 Dsu += / Dst -=; order 1 Dsu -= / Dst +=; order 2 as order 0 */
          double sg = order == 1 ? -1.0 : 1.0;
          Dsu[3 * t + n] += sg * su;
          Dst[3 * t + n] -= sg * st;
        }
      }
  }
}

void orc_integ3_stress(const orc_material* m, int order, int post, const double* el,
                       const double* point, double* DDsu, double* DDst) {
  double mu0 = mu_of(m->E0, m->nu0), mu1 = mu_of(m->E1, m->nu1);
  double lam0 = 2.0 * mu0 * m->nu0 / (1 - 2 * m->nu0);
  double lam1 = 2.0 * mu1 * m->nu1 / (1 - 2 * m->nu1);
  double dlam = post ? lam0 : lam0 - lam1;
  double dmu = post ? mu0 : mu0 - mu1;
  int rows = order == 0 ? 6 : (order == 1 ? 18 : 54);
  double Dsu[54 * 3], Dst[54 * 3];
  for (int a = 0; a < 8; ++a)
    for (int b = 0; b < 8; ++b) {
      double N[4], nv[3], g[3] = {0, 0, 0};
      shape4(GP8[a], GP8[b], N);
      double J = normal_jac(GP8[a], GP8[b], el, nv);
      double W = GW8[a] * GW8[b];
      for (int q = 0; q < 4; ++q)
        for (int c = 0; c < 3; ++c) g[c] += N[q] * el[3 * q + c];
      double x[3] = {g[0] - point[0], g[1] - point[1], g[2] - point[2]};
      kel_ctx K;
      kel_build(&K, x, mu0, m->nu0);
      stress_kernel(&K, order, lam0, mu0, dlam, dmu, nv, Dsu, Dst);
      for (int r = 0; r < rows; ++r)
        for (int q = 0; q < 4; ++q)
          for (int c = 0; c < 3; ++c) {
            DDsu[12 * r + 3 * q + c] += Dsu[3 * r + c] * N[q] * W * J;
            DDst[12 * r + 3 * q + c] += Dst[3 * r + c] * N[q] * W * J;
          }
    }
}

/* ============================================================== system assembly, 78 layout ==== */
/* per-sphere unknown offsets (configElastic3D.cpp:308-343) */
static int i78_ij(int i, int j) { return TIJ[i][j]; }
static int i78_ijk(int i, int j, int k) { return 6 + 3 * TIJ[i][j] + k; }
static int i78_ijkl(int i, int j, int k, int l) { return 24 + 9 * TIJ[i][j] + 3 * k + l; }

static double C_iso(double lam, double mu, int i, int j, int k, int l) {
  /* Constitutive / Constitutive1 (integratorElastic3DHelper.cpp:2376-2389) */
  return lam * KD[i][j] * KD[k][l] + mu * (KD[i][k] * KD[j][l] + KD[i][l] * KD[j][k]);
}

/* inclusion-inclusion block for target s, source h into H (rows/cols 3NN + 78 .) */
static void incl_block(const orc_problem* p, int s, int h, int orders, double* H, int n78) {
  const orc_material* m = &p->mat;
  double mu0 = mu_of(m->E0, m->nu0), mu1 = mu_of(m->E1, m->nu1);
  double lam0 = 2.0 * mu0 * m->nu0 / (1 - 2 * m->nu0);
  double lam1 = 2.0 * mu1 * m->nu1 / (1 - 2 * m->nu1);
  double x[3];
  for (int c = 0; c < 3; ++c) x[c] = p->centres[3 * s + c] - p->centres[3 * h + c];
  sph_ctx* S = sph_new(x, p->radii[h]);
  const int base_r = 3 * p->NN + 78 * s, base_c = 3 * p->NN + 78 * h;
  const double sym = (s == h) ? 1.0 : 0.0;
#define HH(r, c) H[(size_t)(base_c + (c)) * n78 + (base_r + (r))]
  /* sum_kl (C0 - C1)_ijkl D(k,l,m,n,...) */
#define DCON(w, ext, ne, res)                                                             \
  do {                                                                                     \
    double _a = 0;                                                                         \
    for (int kk = 0; kk < 3; ++kk)                                                         \
      for (int ll = 0; ll < 3; ++ll) {                                                     \
        double dc = C_iso(lam0, mu0, i, j, kk, ll) - C_iso(lam1, mu1, i, j, kk, ll);       \
        if (dc != 0) _a += dc * Dgen(S, m->nu0, (w), kk, ll, mm, nn, (ext), (ne));         \
      }                                                                                    \
    res = _a;                                                                              \
  } while (0)
  for (int i = 0; i < 3; ++i)
    for (int j = i; j < 3; ++j)
      for (int mm = 0; mm < 3; ++mm)
        for (int nn = mm; nn < 3; ++nn) {
          const double mp = MAPW[mm][nn];
          double c0 = C_iso(lam0, mu0, i, j, mm, nn);
          if (orders & 1) { /* integratorElastic3D.cpp:377-427 */
            double d;
            DCON(0, NULL, 0, d);
            HH(i78_ij(i, j), i78_ij(mm, nn)) += mp * (-d + sym * c0);
            for (int pp = 0; pp < 3; ++pp) {
              DCON(1 + pp, NULL, 0, d);
              HH(i78_ij(i, j), i78_ijk(mm, nn, pp)) += mp * (-d);
            }
            for (int pp = 0; pp < 3; ++pp)
              for (int qq = 0; qq < 3; ++qq) {
                DCON(4 + TIJ[pp][qq], NULL, 0, d);
                HH(i78_ij(i, j), i78_ijkl(mm, nn, pp, qq)) += mp * (-d);
              }
          }
          if (orders & 2) { /* :484-545 */
            for (int r = 0; r < 3; ++r) {
              int ex[1] = {r};
              double d;
              DCON(0, ex, 1, d);
              HH(i78_ijk(i, j, r), i78_ij(mm, nn)) += mp * (-d);
              for (int pp = 0; pp < 3; ++pp) {
                DCON(1 + pp, ex, 1, d);
                HH(i78_ijk(i, j, r), i78_ijk(mm, nn, pp)) += mp * (-d + sym * c0 * KD[pp][r]);
              }
              for (int pp = 0; pp < 3; ++pp)
                for (int qq = 0; qq < 3; ++qq) {
                  DCON(4 + TIJ[pp][qq], ex, 1, d);
                  HH(i78_ijk(i, j, r), i78_ijkl(mm, nn, pp, qq)) += mp * (-d);
                }
            }
          }
          if (orders & 4) { /* :598-660, rows carry -0.5 */
            for (int r = 0; r < 3; ++r)
              for (int z = 0; z < 3; ++z) {
                int ex[2] = {r, z};
                double d;
                DCON(0, ex, 2, d);
                HH(i78_ijkl(i, j, r, z), i78_ij(mm, nn)) += mp * (-0.5 * d);
                for (int pp = 0; pp < 3; ++pp) {
                  DCON(1 + pp, ex, 2, d);
                  HH(i78_ijkl(i, j, r, z), i78_ijk(mm, nn, pp)) += mp * (-0.5 * d);
                }
                for (int pp = 0; pp < 3; ++pp)
                  for (int qq = 0; qq < 3; ++qq) {
                    DCON(4 + TIJ[pp][qq], ex, 2, d);
                    HH(i78_ijkl(i, j, r, z), i78_ijkl(mm, nn, pp, qq)) +=
                        mp * (-0.5 * d + sym * c0 * KD[pp][r] * KD[qq][z]);
                  }
              }
          }
        }
#undef DCON
#undef HH
  free(S);
}

int orc_assemble78(const orc_problem* p, int phases, double* H, double* G) {
  jets_init();
  const int NN = p->NN, NE = p->NE, num = p->num;
  const int n78 = 3 * NN + 78 * num;
  const size_t ng = (size_t)12 * NE;
  memset(H, 0, sizeof(double) * (size_t)n78 * n78);
  memset(G, 0, sizeof(double) * (size_t)n78 * ng);
  const orc_material* m = &p->mat;
  const double mu0 = mu_of(m->E0, m->nu0);
  const double lam0 = 2.0 * mu0 * m->nu0 / (1 - 2 * m->nu0);
#define H_(r, c) H[(size_t)(c) * n78 + (r)]
#define G_(r, c) G[(size_t)(c) * n78 + (r)]
  if (phases & ORC_ADD_BEM) { /* BEMBuilderElastic3D.cpp:1355-1509 */
    double* dU = (double*)calloc((size_t)36 * NN, sizeof(double));
    double* dT = (double*)calloc((size_t)36 * NN, sizeof(double));
    for (int e = 0; e < NE; ++e) {
      memset(dU, 0, sizeof(double) * 36 * NN);
      memset(dT, 0, sizeof(double) * 36 * NN);
      const int* sc = p->conn + 4 * e;
      orc_integ3(m, sc, p->nodes, NN, dU, dT);
      for (int l = 0; l < 3 * NN; ++l)
        for (int k = 0; k < 12; ++k) G_(l, 12 * e + k) = dU[12 * l + k];
      for (int q = 0; q < 4; ++q)
        for (int kk = 0; kk < 3; ++kk)
          for (int l = 0; l < 3 * NN; ++l) H_(l, 3 * (sc[q] - 1) + kk) += dT[12 * l + 3 * q + kk];
    }
    free(dU);
    free(dT);
    /* rigid-body diagonal (:1478-1501) */
    double* diag = (double*)calloc((size_t)9 * NN, sizeof(double));
    for (int i = 0; i < NN; ++i)
      for (int j = 0; j < 3; ++j)
        for (int k = 0; k < 3 * NN; ++k) diag[3 * k + j] += H_(k, 3 * i + j);
    for (int q = 0; q < NN; ++q)
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) H_(3 * q + a, 3 * q + b) = -diag[3 * (3 * q + a) + b];
    free(diag);
  }
  if (phases & ORC_ADD_FIELD) { /* integratorElastic3D.cpp:135-340 */
#pragma omp parallel for schedule(dynamic)
    for (int K = 0; K < num; ++K)
      for (int jn = 0; jn < NN; ++jn) {
        double x[3];
        for (int c = 0; c < 3; ++c) x[c] = p->nodes[3 * jn + c] - p->centres[3 * K + c];
        sph_ctx* S = sph_new(x, p->radii[K]);
        for (int i = 0; i < 3; ++i)
          for (int mm = 0; mm < 3; ++mm)
            for (int nn = mm; nn < 3; ++nn) {
              const double mp = MAPW[mm][nn];
              /* HMAT = -DD = + sum map C0(jj,ll,m,n) IGreen(x, a, i, jj, [p, [q,]] ll) */
              double s0 = 0, s1[3] = {0, 0, 0}, s2[9] = {0};
              for (int jj = 0; jj < 3; ++jj)
                for (int ll = 0; ll < 3; ++ll) {
                  double c0 = C_iso(lam0, mu0, jj, ll, mm, nn);
                  if (c0 == 0) continue;
                  int g1[3] = {i, jj, ll};
                  s0 += mp * c0 * igreen_ctx(S, mu0, m->nu0, 1, g1);
                  for (int pp = 0; pp < 3; ++pp) {
                    int g2[4] = {i, jj, pp, ll};
                    s1[pp] += mp * c0 * igreen_ctx(S, mu0, m->nu0, 11, g2);
                    for (int qq = 0; qq < 3; ++qq) {
                      int g3[5] = {i, jj, pp, qq, ll};
                      s2[3 * pp + qq] += mp * c0 * igreen_ctx(S, mu0, m->nu0, 21, g3);
                    }
                  }
                }
              const int c0 = 3 * NN + 78 * K;
              H_(3 * jn + i, c0 + i78_ij(mm, nn)) = s0;
              for (int pp = 0; pp < 3; ++pp) H_(3 * jn + i, c0 + i78_ijk(mm, nn, pp)) = s1[pp];
              for (int pp = 0; pp < 3; ++pp)
                for (int qq = 0; qq < 3; ++qq)
                  H_(3 * jn + i, c0 + i78_ijkl(mm, nn, pp, qq)) = s2[3 * pp + qq];
            }
        free(S);
      }
  }
  int io = ((phases & ORC_INCL0) ? 1 : 0) | ((phases & ORC_INCL1) ? 2 : 0) |
           ((phases & ORC_INCL2) ? 4 : 0);
  if (io) {
#pragma omp parallel for schedule(dynamic)
    for (int s = 0; s < num; ++s)
      for (int h = 0; h < num; ++h) incl_block(p, s, h, io, H, n78);
  }
  for (int order = 0; order < 3; ++order) { /* BEMBuilderElastic3D.cpp:1511-1867 */
    int bit = order == 0 ? ORC_BEM0 : (order == 1 ? ORC_BEM1 : ORC_BEM2);
    if (!(phases & bit)) continue;
    int rows = order == 0 ? 6 : (order == 1 ? 18 : 54);
    int off = order == 0 ? 0 : (order == 1 ? 6 : 24);
    double fac = order == 2 ? 0.5 : 1.0;
#pragma omp parallel for schedule(dynamic)
    for (int e = 0; e < NE; ++e) {
      double DDsu[54 * 12], DDst[54 * 12], el[12];
      const int* sc = p->conn + 4 * e;
      for (int q = 0; q < 4; ++q)
        for (int c = 0; c < 3; ++c) el[3 * q + c] = p->nodes[3 * (sc[q] - 1) + c];
      for (int K = 0; K < num; ++K) {
        memset(DDsu, 0, sizeof DDsu);
        memset(DDst, 0, sizeof DDst);
        orc_integ3_stress(m, order, 0, el, p->centres + 3 * K, DDsu, DDst);
        const int r0 = 3 * NN + 78 * K + off;
#pragma omp critical
        {
          for (int q = 0; q < 4; ++q)
            for (int k = 0; k < 3; ++k)
              for (int l = 0; l < rows; ++l)
                H_(r0 + l, 3 * (sc[q] - 1) + k) -= fac * DDsu[12 * l + 3 * q + k];
        }
        for (int l = 0; l < rows; ++l)
          for (int k = 0; k < 12; ++k) G_(r0 + l, 12 * e + k) = fac * DDst[12 * l + k];
      }
    }
  }
#undef H_
#undef G_
  return 0;
}

/* ============================================================================ BC fold ========== */
int orc_column_map(const orc_problem* p, int* node_col, int* slot_col) {
  const int NN = p->NN, NE = p->NE;
  int t = 0;
  for (int j = 0; j < 3 * NN; ++j) node_col[j] = (p->bc[2 * j] == 1.0) ? t++ : -1;
  /* traction-unknown keys (node, dof) in first-encounter order over (e, i)
   * (BEMBuilderElastic3D.cpp:1984-2016; Find :1894-1918 compares node + 0.1 (dof+1)) */
  int* key_col = (int*)malloc(sizeof(int) * 3 * NN);
  for (int j = 0; j < 3 * NN; ++j) key_col[j] = -1;
  for (int e = 0; e < NE; ++e)
    for (int i = 0; i < 12; ++i) {
      double f = p->tbc[24 * e + i];
      int key = 3 * (p->conn[4 * e + i / 3] - 1) + i % 3;
      if (f == 1.0)
        slot_col[12 * e + i] = -1;
      else if (f == 0.0) {
        if (key_col[key] < 0) key_col[key] = t++;
        slot_col[12 * e + i] = key_col[key];
      } else
        slot_col[12 * e + i] = -2; /* neither branch of the reference: ignored */
    }
  free(key_col);
  return t;
}

int orc_apply_bc(const orc_problem* p, const double* H, const double* G, double* A, double* b) {
  const int NN = p->NN, NE = p->NE, n = 3 * NN + 78 * p->num;
  int* nc = (int*)malloc(sizeof(int) * 3 * NN);
  int* sc = (int*)malloc(sizeof(int) * 12 * NE);
  int t = orc_column_map(p, nc, sc);
  memset(A, 0, sizeof(double) * (size_t)n * n);
  memset(b, 0, sizeof(double) * n);
  if (t != 3 * NN) {
    free(nc);
    free(sc);
    return -1;
  }
  for (int j = 0; j < 3 * NN; ++j) {
    if (nc[j] >= 0)
      memcpy(A + (size_t)nc[j] * n, H + (size_t)j * n, sizeof(double) * n);
    else
      for (int i = 0; i < n; ++i) b[i] -= H[(size_t)j * n + i] * p->bc[2 * j + 1];
  }
  for (int e = 0; e < NE; ++e)
    for (int i = 0; i < 12; ++i) {
      const double* g = G + (size_t)(12 * e + i) * n;
      if (sc[12 * e + i] == -1) {
        double tv = p->tbc[24 * e + 12 + i];
        for (int r = 0; r < n; ++r) b[r] += tv * g[r];
      } else if (sc[12 * e + i] >= 0) {
        double* a = A + (size_t)sc[12 * e + i] * n;
        for (int r = 0; r < n; ++r) a[r] -= g[r];
      }
    }
  memcpy(A + (size_t)3 * NN * n, H + (size_t)3 * NN * n, sizeof(double) * (size_t)n * 78 * p->num);
  free(nc);
  free(sc);
  return 0;
}

int orc_get_ut(const orc_problem* p, const double* X, double* U, double* T) {
  const int NN = p->NN, NE = p->NE, n = 3 * NN + 78 * p->num;
  int* nc = (int*)malloc(sizeof(int) * 3 * NN);
  int* sc = (int*)malloc(sizeof(int) * 12 * NE);
  orc_column_map(p, nc, sc);
  for (int i = 0; i < 3 * NN; ++i) U[i] = nc[i] >= 0 ? X[nc[i]] : p->bc[2 * i + 1];
  for (int i = 3 * NN; i < n; ++i) U[i] = X[i];
  int mI = 0;
  for (int e = 0; e < NE; ++e)
    for (int i = 0; i < 12; ++i) {
      int c = sc[12 * e + i];
      if (c == -1)
        T[mI++] = p->tbc[24 * e + 12 + i];
      else if (c >= 0)
        T[mI++] = X[c];
    }
  free(nc);
  free(sc);
  return 0;
}

/* ============================================================================ layouts ========== */
static const int KL_I[6] = {0, 0, 0, 1, 1, 2}, KL_J[6] = {0, 1, 2, 1, 2, 2};

/* 60-layout offset of each 78-layout offset (rows: -1 for duplicated (r>z) rows) */
static void map_78_60(int* row_of78, int* col_a, int* col_b) {
  for (int u = 0; u < 24; ++u) {
    row_of78[u] = u;
    col_a[u] = u;
    col_b[u] = -1;
  }
  for (int t = 0; t < 6; ++t)
    for (int k = 0; k < 3; ++k)
      for (int l = 0; l < 3; ++l) {
        int u = 24 + 9 * t + 3 * k + l;
        row_of78[u] = (k <= l) ? 24 + 6 * t + TIJ[k][l] : -1;
      }
  /* for each 60 column: its one or two 78 columns */
  for (int t = 0; t < 6; ++t)
    for (int kl = 0; kl < 6; ++kl) {
      int k = KL_I[kl], l = KL_J[kl];
      col_a[24 + 6 * t + kl] = 24 + 9 * t + 3 * k + l;
      col_b[24 + 6 * t + kl] = (k < l) ? 24 + 9 * t + 3 * l + k : -1;
    }
}

int orc_reduce_78_to_60(const orc_problem* p, const double* A78, const double* b78, double* A60,
                        double* b60) {
  const int NN = p->NN, num = p->num;
  const int n78 = 3 * NN + 78 * num, n60 = 3 * NN + 60 * num;
  int row_of78[78], col_a[60], col_b[60];
  map_78_60(row_of78, col_a, col_b);
  int* r60 = (int*)malloc(sizeof(int) * n78);
  for (int r = 0; r < 3 * NN; ++r) r60[r] = r;
  for (int h = 0; h < num; ++h)
    for (int u = 0; u < 78; ++u)
      r60[3 * NN + 78 * h + u] = row_of78[u] < 0 ? -1 : 3 * NN + 60 * h + row_of78[u];
  for (int c = 0; c < n60; ++c) {
    int ca, cb;
    if (c < 3 * NN) {
      ca = c;
      cb = -1;
    } else {
      int h = (c - 3 * NN) / 60, v = (c - 3 * NN) % 60;
      ca = 3 * NN + 78 * h + col_a[v];
      cb = col_b[v] < 0 ? -1 : 3 * NN + 78 * h + col_b[v];
    }
    for (int r = 0; r < n78; ++r) {
      if (r60[r] < 0) continue;
      double v = A78[(size_t)ca * n78 + r];
      if (cb >= 0) v += A78[(size_t)cb * n78 + r];
      A60[(size_t)c * n60 + r60[r]] = v;
    }
  }
  for (int r = 0; r < n78; ++r)
    if (r60[r] >= 0) b60[r60[r]] = b78[r];
  free(r60);
  return 0;
}

int orc_expand_60_to_78(const orc_problem* p, const double* x60, double* x78) {
  const int NN = p->NN;
  for (int r = 0; r < 3 * NN; ++r) x78[r] = x60[r];
  for (int h = 0; h < p->num; ++h) {
    const double* s = x60 + 3 * NN + 60 * h;
    double* d = x78 + 3 * NN + 78 * h;
    for (int u = 0; u < 24; ++u) d[u] = s[u];
    for (int t = 0; t < 6; ++t)
      for (int k = 0; k < 3; ++k)
        for (int l = 0; l < 3; ++l)
          d[24 + 9 * t + 3 * k + l] = s[24 + 6 * t + TIJ[k < l ? k : l][k < l ? l : k]];
  }
  return 0;
}

/* ============================================================================ solvers ========== */
int orc_fullpiv_solve(int n, double* A, const double* b, double* x) {
  /* FullPivLU::compute (FullPivLU.h:425-497): at step k pick the entry of largest magnitude in
   * the trailing corner, swap its row and column to k, scale the column, rank-1 update. */
  int* rp = (int*)malloc(sizeof(int) * n);
  int* cq = (int*)malloc(sizeof(int) * n);
  for (int i = 0; i < n; ++i) {
    rp[i] = i;
    cq[i] = i;
  }
#define AA(i, j) A[(size_t)(j) * n + (i)]
  int rc = 0;
  for (int k = 0; k < n; ++k) {
    double best = -1;
    int bi = k, bj = k;
    for (int j = k; j < n; ++j)
      for (int i = k; i < n; ++i) {
        double v = fabs(AA(i, j));
        if (v > best) {
          best = v;
          bi = i;
          bj = j;
        }
      }
    if (best == 0.0) {
      rc = -1;
      break;
    }
    if (bi != k) {
      for (int j = 0; j < n; ++j) {
        double t = AA(k, j);
        AA(k, j) = AA(bi, j);
        AA(bi, j) = t;
      }
      int t = rp[k];
      rp[k] = rp[bi];
      rp[bi] = t;
    }
    if (bj != k) {
      for (int i = 0; i < n; ++i) {
        double t = AA(i, k);
        AA(i, k) = AA(i, bj);
        AA(i, bj) = t;
      }
      int t = cq[k];
      cq[k] = cq[bj];
      cq[bj] = t;
    }
    double piv = AA(k, k);
    for (int i = k + 1; i < n; ++i) AA(i, k) /= piv;
#pragma omp parallel for if (n - k > 256)
    for (int j = k + 1; j < n; ++j) {
      double f = AA(k, j);
      if (f != 0)
        for (int i = k + 1; i < n; ++i) AA(i, j) -= AA(i, k) * f;
    }
  }
  if (rc == 0) {
    double* y = (double*)malloc(sizeof(double) * n);
    for (int i = 0; i < n; ++i) y[i] = b[rp[i]];
    for (int j = 0; j < n; ++j)
      for (int i = j + 1; i < n; ++i) y[i] -= AA(i, j) * y[j];
    for (int j = n - 1; j >= 0; --j) {
      y[j] /= AA(j, j);
      for (int i = 0; i < j; ++i) y[i] -= AA(i, j) * y[j];
    }
    for (int i = 0; i < n; ++i) x[cq[i]] = y[i];
    free(y);
  }
#undef AA
  free(rp);
  free(cq);
  return rc;
}

int orc_partialpiv_solve(int n, double* A, const double* b, double* x) {
#define AA(i, j) A[(size_t)(j) * n + (i)]
  double* y = (double*)malloc(sizeof(double) * n);
  for (int i = 0; i < n; ++i) y[i] = b[i];
  for (int k = 0; k < n; ++k) {
    int bi = k;
    double best = fabs(AA(k, k));
    for (int i = k + 1; i < n; ++i)
      if (fabs(AA(i, k)) > best) {
        best = fabs(AA(i, k));
        bi = i;
      }
    if (best == 0.0) {
      free(y);
      return -1;
    }
    if (bi != k) {
      for (int j = 0; j < n; ++j) {
        double t = AA(k, j);
        AA(k, j) = AA(bi, j);
        AA(bi, j) = t;
      }
      double t = y[k];
      y[k] = y[bi];
      y[bi] = t;
    }
    double piv = AA(k, k);
    for (int i = k + 1; i < n; ++i) AA(i, k) /= piv;
#pragma omp parallel for if (n - k > 256)
    for (int j = k + 1; j < n; ++j) {
      double f = AA(k, j);
      if (f != 0)
        for (int i = k + 1; i < n; ++i) AA(i, j) -= AA(i, k) * f;
    }
  }
  for (int j = 0; j < n; ++j)
    for (int i = j + 1; i < n; ++i) y[i] -= AA(i, j) * y[j];
  for (int j = n - 1; j >= 0; --j) {
    y[j] /= AA(j, j);
    for (int i = 0; i < j; ++i) y[i] -= AA(i, j) * y[j];
  }
  for (int i = 0; i < n; ++i) x[i] = y[i];
  free(y);
#undef AA
  return 0;
}

/* ========================================================================= postprocess ========== */
int orc_stress(const orc_problem* p, const double* U, const double* T, double* s) {
  jets_init();
  const int NN = p->NN, NE = p->NE, num = p->num, nump = p->nump;
  const orc_material* m = &p->mat;
  const double mu0 = mu_of(m->E0, m->nu0);
  const double lam0 = 2.0 * mu0 * m->nu0 / (1 - 2 * m->nu0);
#pragma omp parallel for schedule(dynamic)
  for (int KK = 0; KK < nump; ++KK) {
    const double* xp = p->points + 3 * KK;
    /* boundary term post_stress_BEM (postProcessorElastic3D.cpp:224-365) */
    double sb[6] = {0, 0, 0, 0, 0, 0};
    for (int e = 0; e < NE; ++e) {
      double DDsu[6 * 12] = {0}, DDst[6 * 12] = {0}, el[12];
      const int* sc = p->conn + 4 * e;
      for (int q = 0; q < 4; ++q)
        for (int c = 0; c < 3; ++c) el[3 * q + c] = p->nodes[3 * (sc[q] - 1) + c];
      orc_integ3_stress(m, 0, 1, el, xp, DDsu, DDst);
      for (int l = 0; l < 6; ++l)
        for (int q = 0; q < 4; ++q)
          for (int k = 0; k < 3; ++k) {
            sb[l] += DDsu[12 * l + 3 * q + k] * U[3 * (sc[q] - 1) + k];
            sb[l] += DDst[12 * l + 3 * q + k] * T[12 * e + 3 * q + k];
          }
    }
    /* inclusion term and interior eigenstrain (:367-522); last containing sphere wins */
    int inside = -1;
    for (int h = 0; h < num; ++h) {
      double d2 = 0;
      for (int c = 0; c < 3; ++c) {
        double d = xp[c] - p->centres[3 * h + c];
        d2 += d * d;
      }
      if (d2 - p->radii[h] * p->radii[h] <= 0) inside = h;
    }
    double ep[6] = {0, 0, 0, 0, 0, 0};
    for (int h = 0; h < num; ++h) {
      double x[3];
      for (int c = 0; c < 3; ++c) x[c] = xp[c] - p->centres[3 * h + c];
      sph_ctx* S = sph_new(x, p->radii[h]);
      const double* E = U + 3 * NN + 78 * h;
      for (int i = 0; i < 3; ++i)
        for (int j = i; j < 3; ++j) {
          double acc = 0;
          for (int mm = 0; mm < 3; ++mm)
            for (int nn = mm; nn < 3; ++nn) {
              double mp = MAPW[mm][nn];
              acc += Dgen(S, m->nu0, 0, i, j, mm, nn, NULL, 0) * mp * E[i78_ij(mm, nn)];
              for (int pp = 0; pp < 3; ++pp)
                acc += Dgen(S, m->nu0, 1 + pp, i, j, mm, nn, NULL, 0) * mp * E[i78_ijk(mm, nn, pp)];
              for (int pp = 0; pp < 3; ++pp)
                for (int qq = 0; qq < 3; ++qq)
                  acc += Dgen(S, m->nu0, 4 + TIJ[pp][qq], i, j, mm, nn, NULL, 0) * mp *
                         E[i78_ijkl(mm, nn, pp, qq)];
            }
          ep[TIJ[i][j]] += acc;
        }
      free(S);
    }
    if (inside >= 0) {
      const double* E = U + 3 * NN + 78 * inside;
      double d[3];
      for (int c = 0; c < 3; ++c) d[c] = xp[c] - p->centres[3 * inside + c];
      for (int i = 0; i < 3; ++i)
        for (int j = i; j < 3; ++j) {
          double e0 = E[i78_ij(i, j)], e1 = 0, e2 = 0;
          for (int k = 0; k < 3; ++k) e1 += E[i78_ijk(i, j, k)] * d[k];
          for (int k = 0; k < 3; ++k)
            for (int l = k; l < 3; ++l) e2 += MAPW[k][l] * E[i78_ijkl(i, j, k, l)] * d[k] * d[l];
          ep[TIJ[i][j]] -= e0 + e1 + e2;
        }
    }
    for (int i = 0; i < 3; ++i)
      for (int j = i; j < 3; ++j) {
        double v = 0;
        for (int mm = 0; mm < 3; ++mm)
          for (int nn = mm; nn < 3; ++nn)
            v += C_iso(lam0, mu0, i, j, mm, nn) * MAPW[mm][nn] * ep[TIJ[mm][nn]];
        s[6 * KK + TIJ[i][j]] = v + sb[TIJ[i][j]];
      }
  }
  return 0;
}
