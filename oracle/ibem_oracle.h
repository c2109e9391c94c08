/* CPU oracle for the iBEM elastic-3D spherical-inclusion path.
 *
 * TEST INFRASTRUCTURE ONLY. This library restates the reference's algorithm in plain C so that the
 * HIP path can be checked against it. Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it, and only as the checker; the product never calls it.
 *
 * Every function cites the reference file:line it follows (paths relative to the reference root).
 * Matrices are column-major (Eigen's default storage in the reference), indices 0-based except the
 * element connectivity, which stays one-based as in BEM.txt.
 */
#ifndef IBEM_ORACLE_H
#define IBEM_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  double E0, nu0, E1, nu1; /* matrix and particle moduli (BEM.txt line 2-3, position.txt line 2-3) */
} orc_material;

/* ---- closed forms (integratorElastic3DHelper.cpp) ---------------------------------------------- */
/* kind 40,50,60,41,51,61,42,52,62 with the reference's argument order, D_40:2394 ... D_62:2540.
 * x = field point - sphere centre, a = sphere radius. */
double orc_D(const orc_material* m, int kind, const int* idx, const double* x, double a);
/* Green_0k (order 0..4), integratorElastic3DHelper.cpp:2559-2593 */
double orc_green(const orc_material* m, int order, const int* idx, const double* x);
/* IGreen_01/11/21 (kind 1, 11, 21), integratorElastic3DHelper.cpp:2604-2624 */
double orc_igreen(const orc_material* m, int kind, const int* idx, const double* x, double a);

/* ---- boundary quadrature (BEMBuilderElastic3D.cpp) --------------------------------------------- */
/* Integ3 (BEMBuilderElastic3D.cpp:1169-1353): element with one-based nodes sctrb against nn
 * collocation nodes (row-major nn x 3). dUe, dTe: row-major (3 nn) x 12, accumulated (+=). */
void orc_integ3(const orc_material* m, const int* sctrb, const double* nodes, int nn, double* dUe,
                double* dTe);
/* Integ3_stress / Integ3_stress1 / Integ3_stress2 (BEMBuilderElastic3D.cpp:982-1167, 674-778,
 * 324-427). elcor row-major 4 x 3; rows 6 / 18 / 54; DDsu, DDst row-major rows x 12 (+=).
 * post = 1 selects the postprocessor variant (postProcessorElastic3D.cpp:32-220: C1 = 0). */
void orc_integ3_stress(const orc_material* m, int order, int post, const double* elcor,
                       const double* point, double* DDsu, double* DDst);

/* ---- whole-system assembly in the reference's 78-unknown layout -------------------------------- */
typedef struct {
  int NN, NE, num, nump;
  orc_material mat;
  const double* nodes;   /* NN x 3 row-major */
  const int* conn;       /* NE x 4 row-major, one-based */
  const double* bc;      /* 3NN x 2 row-major (flag, value) */
  const double* tbc;     /* NE x 24 row-major */
  const double* centres; /* num x 3 row-major */
  const double* radii;   /* num */
  const double* points;  /* nump x 3 row-major */
} orc_problem;

/* phase bit mask for orc_assemble78 */
enum {
  ORC_ADD_BEM = 1,        /* BEMBuilderElastic3D::addBEM :1355-1509 */
  ORC_ADD_FIELD = 2,      /* integratorElastic3D::addFieldToBEM :135-340 */
  ORC_INCL0 = 4,          /* integratorElastic3D::addFluxEquivalentInclusion :341-454 */
  ORC_INCL1 = 8,          /* ...FirstOrderInclusion :455-567 */
  ORC_INCL2 = 16,         /* ...SecondOrderInclusion :568-682 */
  ORC_BEM0 = 32,          /* BEMBuilderElastic3D::addFluxEquivalentBEM :1511-1627 */
  ORC_BEM1 = 64,          /* ...FirstOrderBEM :1629-1741 */
  ORC_BEM2 = 128,         /* ...SecondOrderBEM :1743-1867 */
  ORC_ALL = 255
};
/* H: n78 x n78, G: n78 x 12NE, column-major, both zeroed here. n78 = 3NN + 78 num. */
int orc_assemble78(const orc_problem* p, int phases, double* H, double* G);

/* ApplyBC (BEMBuilderElastic3D.cpp:1920-2051): A (n78 x n78), b (n78), colmap (3NN ints, the A
 * column of each traction-unknown key or displacement-unknown dof, see orc_column_map). */
int orc_apply_bc(const orc_problem* p, const double* H, const double* G, double* A, double* b);

/* Column map of the BC fold, shared by the 78 and 60 layouts (BEMBuilderElastic3D.cpp:1960-2016):
 * node_col[3NN]   A column of node dof j if its displacement is unknown, else -1;
 * slot_col[12NE]  A column of element traction dof (e, i) if the traction is unknown, else -1
 *                 (several slots share a column when they have the same (node, dof) key).
 * Returns the number of BEM columns (must equal 3NN for a square system), or -1. */
int orc_column_map(const orc_problem* p, int* node_col, int* slot_col);

/* ---- 78 <-> 60 unknown layouts ----------------------------------------------------------------- */
/* The quadratic eigenstrain block E_ij,kl is symmetric in (k,l): rows (r,z) with r > z duplicate
 * rows (z,r) and columns (p,q),(q,p) merge. 60 = 6 + 18 + 36 per sphere. */
int orc_reduce_78_to_60(const orc_problem* p, const double* A78, const double* b78, double* A60,
                        double* b60);
int orc_expand_60_to_78(const orc_problem* p, const double* x60, double* x78);

/* ---- solvers ----------------------------------------------------------------------------------- */
/* Complete-pivot LU solve (Eigen FullPivLU, FullPivLU.h:425-520 as called at
 * systemSolverElastic3D.h:55). A is overwritten. Returns 0 or -1 when singular. */
int orc_fullpiv_solve(int n, double* A, const double* b, double* x);
/* Partial-pivot LU solve (row pivoting), the GPU's algorithm in its unblocked form. */
int orc_partialpiv_solve(int n, double* A, const double* b, double* x);

/* getUTVector (BEMBuilderElastic3D.cpp:2053-2117): X (n78) -> U (n78), T (12NE) */
int orc_get_ut(const orc_problem* p, const double* X78, double* U, double* T);

/* getFlux stresses (postProcessorElastic3D.cpp:224-587): U (>= 3NN + 78num), T (12NE);
 * s: nump x 6 row-major (s00 s01 s02 s11 s12 s22). */
int orc_stress(const orc_problem* p, const double* U, const double* T, double* s);

#ifdef __cplusplus
}
#endif
#endif
