/*
 * vforacle.h -- CPU restatement (FP64, plain C) of VPFHF's u- and V-subproblem
 * operators.  TEST INFRASTRUCTURE ONLY: the oracle is the checker.  Only the
 * test suite, the smoke entry and the bench's cpu_baseline leg may load it;
 * the product library never links or calls it.
 *
 * Every function names the reference code it restates (paths relative to the
 * bourdin/VPFHF tree).  Loop orders follow the reference so that summation
 * order (and therefore rounding) matches it as closely as a restatement can.
 *
 * Layout (DMDA natural order, single rank; VFCommon.c:321-376):
 *   node(i,j,k) = i + nx*(j + ny*k),  u-dof = 3*node + c,  V-dof = node
 *   cell(ei,ej,ek) = ei + (nx-1)*(ej + (ny-1)*ek)
 *
 * Matrices are stored in a fixed 27-neighbour ELL layout ("vfo_ell"):
 *   row r = dof*node + c, slot s = ((dk+1)*3 + (dj+1))*3 + (di+1), then *dof + c2
 *   col[r*W + slot] = column index or -1 for a neighbour outside the grid.
 * vfo_ell_to_csr() converts to a sorted CSR identical to the AIJ matrix that
 * DMCreateMatrix(daVect/daScal) preallocates (box stencil, width 1).
 */
#ifndef VFORACLE_H
#define VFORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* BC types, faces and vertices: VFCartFE.h:93-127 */
enum { VFO_NONE = 0, VFO_ZERO = 1, VFO_ONE = 2, VFO_FIXED = 3 };
enum { VFO_X0 = 0, VFO_X1, VFO_Y0, VFO_Y1, VFO_Z0, VFO_Z1 };
enum { VFO_X0Y0Z0 = 0, VFO_X1Y0Z0, VFO_X0Y1Z0, VFO_X1Y1Z0,
       VFO_X0Y0Z1, VFO_X1Y0Z1, VFO_X0Y1Z1, VFO_X1Y1Z1 };

/* VFBC (VFCartFE.h:129-136) without the unparsed edges */
typedef struct {
  int    face[6];
  double faceValue[6];
  int    vertex[8];
  double vertexValue[8];
} vfo_bc;

/* VFCartFEElement3D (VFCartFE.h:60-73) */
typedef struct {
  double lx, ly, lz;
  double weight[27];
  double phi[2][2][2][27];
  double dphi[2][2][2][3][27];
} vfo_elem;

/* One problem: grid + properties + BCs.  Properties follow VFMatProp/VFProp
 * (VFCommon.h:69-90); E is per element (E_elem) or uniform (E). */
typedef struct {
  int    nx, ny, nz;            /* nodes per direction */
  double lx, ly, lz;            /* domain lengths */
  double E, nu, alpha, beta, Gc;
  double eta, epsilon, irrevtol;
  int    atnum;                 /* 1 or 2 */
  int    hasCrackPressure;      /* -pressurize */
  const double *E_elem;         /* optional per-element E (N_e) */
  vfo_bc bcU[3];
  vfo_bc bcV;
  int    unilateral;            /* 0 UNILATERAL_NONE, 1 UNILATERAL_NOCOMPRESSION
                                   (VFCommon.h:120-123, -unilateral) */
} vfo_problem;

/* Nodal fields (all length N_n except U, BCU: 3*N_n) */
typedef struct {
  const double *V, *VIrrev, *theta, *thetaRef, *pressure;
} vfo_fields;

typedef struct {
  int64_t nrows;
  int     width;      /* 81 (u) or 27 (V) */
  int64_t *col;       /* nrows*width, -1 = empty */
  double  *val;       /* nrows*width */
} vfo_ell;

/* ---- basis: VFCartFE.c:90-116, 198-269 ---- */
void vfo_elem1d(double lx, double w[3], double phi[2][3], double dphi[2][3]);
void vfo_elem3d_init(vfo_elem *e, double lx, double ly, double lz);

/* ---- material: VFCommon.c:583-586 ---- */
void vfo_lame(double E, double nu, double *lambda, double *mu);
double vfo_atcv(int atnum);  /* VFCommon.c:482-492 */

/* ---- element-local kernels (VFMech.c) ----
 * u8/v8/... are the 8 nodal values in local order (k*2+j)*2+i; u24 is
 * interleaved 3*((k*2+j)*2+i)+c.                                          */
void vfo_bilinear_u(double *Ke576, const double v8[8], double lambda, double mu,
                    double eta, const vfo_elem *e);                 /* VFMech.c:227-290 */
void vfo_grad_u_thermoporo(double *b24, const double v8[8], const double dth8[8],
                           const double p8[8], double lambda, double mu,
                           double alpha, double beta, double eta,
                           const vfo_elem *e);                     /* VFMech.c:386-447 */
void vfo_grad_u_crackpressure(double *b24, const double v8[8], const double p8[8],
                              const vfo_elem *e);                  /* VFMech.c:533-589 */
void vfo_elastic_energy_density(double W27[27], const double u24[24],
                                const double dth8[8], double lambda, double mu,
                                double alpha, const vfo_elem *e);  /* VFMech.c:32-106 */
/* UNILATERAL_NOCOMPRESSION variants: the spherical part is not degraded
 * where tr(eps(u)) < 0 (UNILATERAL_THRES = 0, VFMech.c:6) */
void vfo_bilinear_u_nocomp(double *Ke576, const double u24[24], const double v8[8],
                           double lambda, double mu, double eta,
                           const vfo_elem *e);                     /* VFMech.c:299-376 */
void vfo_grad_u_thermoporo_nocomp(double *b24, const double u24[24], const double v8[8],
                                  const double dth8[8], const double p8[8], double lambda,
                                  double mu, double alpha, double beta, double eta,
                                  const vfo_elem *e);              /* VFMech.c:456-524 */
void vfo_energy_density_sd_nocomp(double S27[27], double D27[27], const double u24[24],
                                  const double dth8[8], double lambda, double mu,
                                  double alpha, const vfo_elem *e); /* VFMech.c:115-218 */
void vfo_bilinear_v_coupling_nocomp(double *K64, const double u24[24], const double dth8[8],
                                    double lambda, double mu, double alpha,
                                    const vfo_elem *e);            /* VFMech.c:2125-2153 */
double vfo_elastic_energy_nocomp_local(const double u24[24], const double v8[8],
                                       const double dth8[8], const double p8[8], double lambda,
                                       double mu, double alpha, double beta, double eta,
                                       const vfo_elem *e);         /* VFMech.c:819-908 */
void vfo_bilinear_v_at(double *K64, int atnum, double Gc, double atCv,
                       double epsilon, const vfo_elem *e);        /* VFMech.c:2040-2084 */
void vfo_bilinear_v_coupling(double *K64, const double u24[24], const double dth8[8],
                             double lambda, double mu, double alpha,
                             const vfo_elem *e);                   /* VFMech.c:2093-2115 */
void vfo_residual_v_at(double b8[8], int atnum, double Gc, double atCv,
                       double epsilon, const vfo_elem *e);        /* VFMech.c:2350-2384 */
void vfo_residual_v_crackpressure(double b8[8], const double u24[24],
                                  const double p8[8], const vfo_elem *e); /* VFMech.c:2294-2341 */
double vfo_elastic_energy_local(const double u24[24], const double v8[8],
                                const double dth8[8], const double p8[8],
                                double lambda, double mu, double alpha,
                                double beta, double eta,
                                const vfo_elem *e);                /* VFMech.c:725-809 */
double vfo_pressure_work_local(const double u24[24], const double v8[8],
                               const double p8[8], const vfo_elem *e); /* VFMech.c:920-974 */
double vfo_surface_energy_local(const double v8[8], int atnum, double Gc,
                                double atCv, double epsilon,
                                const vfo_elem *e);                /* VFMech.c:2394-2465 */
double vfo_crack_volume_local(const double u24[24], const double v8[8],
                              const vfo_elem *e, double *cellavg); /* VFPermfield.c:15-71 */

/* ---- penny cracks: VFCracks.c:101-227 ---- */
double vfo_distance_to_penny(const double x[3], const double center[3], double r,
                             double theta_deg, double phi_deg);
/* builds V for one crack (V is overwritten, VFCracks.c:152/201 set V=1 first) */
void vfo_penny_build_v(double *V, const vfo_problem *P, const double center[3],
                       double r, double theta_deg, double phi_deg, double thickness);

/* ---- geometry ---- */
int64_t vfo_nnodes(const vfo_problem *P);
int64_t vfo_ncells(const vfo_problem *P);
void    vfo_coords(const vfo_problem *P, double *X, double *Y, double *Z); /* VFCommon.c:395-406 */

/* ---- BC kit: VFCartFE.c:279-1454 ---- */
void vfo_bc_init(vfo_bc *bc);
/* mask[dof*node+c] = 1 iff Dirichlet (union of faces and vertices) */
void vfo_bc_mask(const vfo_problem *P, const vfo_bc *bc, int dof, uint8_t *mask);
void vfo_vec_set_from_bc(const vfo_problem *P, const vfo_bc *bc, int dof, double *BCvec);   /* :279-483 */
void vfo_vec_apply_dirichlet(const vfo_problem *P, const vfo_bc *bc, int dof,
                             double *vec, const double *BCvec);                           /* :493-738 */
void vfo_residual_apply_dirichlet(const vfo_problem *P, const vfo_bc *bc, int dof,
                                  double *F, const double *U, const double *BCvec);       /* :747-999 */
/* BCfiles parser: "-U_X0_BC_0 ZERO" / "-V_X1_BC ONE" / "-U_X0_0 val" lines */
int  vfo_bc_parse_options(const char *text, const char *prefix, int dof, vfo_bc *bc);

/* ---- global operators ---- */
vfo_ell *vfo_ell_alloc(int64_t nrows, int width);
void     vfo_ell_free(vfo_ell *A);
void     vfo_ell_matvec(const vfo_ell *A, const double *x, double *y);
/* CSR export: returns nnz; rowptr (nrows+1), colidx, vals preallocated by caller
 * if non-NULL (call once with NULLs to size). */
int64_t  vfo_ell_to_csr(const vfo_ell *A, int64_t *rowptr, int64_t *colidx, double *vals);

/* VF_UIJacobian + MatApplyDirichletBC (rows zeroed, diag 1): VFMech.c:1897-2029 */
vfo_ell *vfo_u_jacobian(const vfo_problem *P, const vfo_fields *F);
/* the same at the state U (NOCOMPRESSION: the matrix depends on U) */
vfo_ell *vfo_u_jacobian_at(const vfo_problem *P, const vfo_fields *F, const double *U);
/* VF_UResidual: VFMech.c:1456-1887 (no in-situ) */
void vfo_u_residual(const vfo_problem *P, const vfo_fields *F, const double *U,
                    const double *BCU, double *res);
/* the element RHS alone: sum_e (thermo-poro + crack pressure) b_e, no BC */
void vfo_u_rhs(const vfo_problem *P, const vfo_fields *F, double *b);
/* VF_VIJacobian + MatApplyDirichletBCRowCol + VF_IrrevApplyEQMat: VFMech.c:2883-3029 */
vfo_ell *vfo_v_jacobian(const vfo_problem *P, const vfo_fields *F, const double *U);
/* VF_VResidual: VFMech.c:2687-2879 */
void vfo_v_residual(const vfo_problem *P, const vfo_fields *F, const double *U,
                    const double *V, double *res);
/* element RHS of V alone (AT + crack pressure), no BC */
void vfo_v_rhs(const vfo_problem *P, const vfo_fields *F, const double *U, double *b);

/* VF_UEnergy3D (VFMech.c:1111-1427) and VF_VEnergy3D (VFMech.c:2474-2526) */
void vfo_u_energy(const vfo_problem *P, const vfo_fields *F, const double *U,
                  double *elastic, double *pressureWork);
double vfo_v_energy(const vfo_problem *P, const double *V);
/* VolumetricCrackOpening (VFPermfield.c:75-150); cellfield optional (N_e) */
double vfo_crack_volume(const vfo_problem *P, const double *U, const double *V,
                        double *cellfield);

/* per-cell energy contributions, 3 per cell interleaved: elastic
 * (VF_ElasticEnergy3D_local, VFMech.c:725-806), crack pressure work
 * (VF_PressureWork3D_local, VFMech.c:920-974, 0 without -pressurize) and
 * surface energy
 * (VF_AT1/AT2SurfaceEnergy3D_local, VFMech.c:2394-2465): the element terms
 * whose sums VF_UEnergy3D / VF_VEnergy3D report */
void vfo_energy_fields(const vfo_problem *P, const vfo_fields *F, const double *U, const double *V,
                       double *cell3);

/* CellToNodeInterpolation (VFPermfield.c:1156-1327): node value = sum over the
 * adjacent cells of cell value * cell volume / sum of those volumes; dof 1 or
 * 3 values per cell / node, interleaved */
void vfo_cell_to_node(const vfo_problem *P, const double *cell, double *node, int dof);

/* ---- solvers (semantics of VFCommon.c:1005-1088, 1132-1157, 1192-1231) ---- */
/* Jacobi PCG on an ELL matrix; returns iterations, <0 on breakdown */
int vfo_pcg(const vfo_ell *A, const double *b, double *x, double rtol,
            double atol, int maxit, double *relres);
/* u-solve: one Newton step of SNESNEWTONLS on the linear u-problem; with
 * NOCOMPRESSION, Newton iterations (Jacobian at the iterate, backtracking on
 * ||F||) until ||F|| <= rtol ||F(U0)||.  Returns CG iterations.
 * U must satisfy the BCs on entry (VFTimeStepPrepare). */
int vfo_u_solve(const vfo_problem *P, const vfo_fields *F, double *U,
                const double *BCU, double rtol, int maxit, double *relres);
/* V-solve: bound-constrained minimisation (SNESVINEWTONRSLS semantics):
 * lb <= V <= ub with Dirichlet/irreversibility rows fixed.  ub==NULL -> 1. */
int vfo_v_solve(const vfo_problem *P, const vfo_fields *F, const double *U,
                double *V, const double *ub, double tol, int maxit);

int vfo_num_threads(void);
void vfo_set_num_threads(int n);

#ifdef __cplusplus
}
#endif
#endif
