/*
 * vfg.h -- C-ABI of the MI355X-native VPFHF u-/V-subproblem operators.
 *
 * Plain pointers and sizes only.  Every entry point returns 0 (VFG_OK) or a
 * negative VFG_E_* code, never aborts and never prints; vfg_last_error()
 * gives the message of the last failure on a handle (or of a failed
 * vfg_create when called with NULL).
 *
 * Layout is the DMDA natural order of the reference (VFCommon.c:321-376):
 *   node(i,j,k) = i + nx*(j + ny*k); displacement dof 3*node + c (interleaved);
 *   scalar fields (V, VIrrev, theta, thetaRef, pressure) one value per node;
 *   cell(ei,ej,ek) = ei + (nx-1)*(ej + (ny-1)*ek) for per-element arrays.
 * A PETSc Vec array of daVect/daScal therefore passes through unchanged.
 *
 * `loc` says where the array pointers of a call live: VFG_HOST (copied
 * through the handle's staging buffers) or VFG_DEVICE (HIP device pointers,
 * used in place on the handle's stream).  No pointer is retained after a call
 * returns; vfg_set_fields copies its inputs.
 *
 * Threading: one handle = one HIP stream; a handle is not thread-safe;
 * distinct handles are independent.
 */
#ifndef VFG_H
#define VFG_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define VFG_HOST 0
#define VFG_DEVICE 1

enum {
  VFG_OK = 0,
  VFG_E_ARG = -1,      /* invalid argument / shape */
  VFG_E_HIP = -2,      /* HIP runtime error */
  VFG_E_NOMEM = -3,    /* device allocation failed */
  VFG_E_BREAKDOWN = -4,/* CG breakdown (non-positive curvature) */
  VFG_E_MAXIT = -5,    /* iteration limit reached before tolerance */
  VFG_E_NAN = -6,      /* NaN/Inf detected in a solve */
  VFG_E_BUDGET = -7,   /* request exceeds the admission budget */
  VFG_E_NODEV = -8     /* no GPU / library not usable */
};

/* BC types, face and vertex order of VFCartFE.h:93-127 */
enum { VFG_NONE = 0, VFG_ZERO = 1, VFG_ONE = 2, VFG_FIXED = 3 };
enum { VFG_X0 = 0, VFG_X1, VFG_Y0, VFG_Y1, VFG_Z0, VFG_Z1 };
enum { VFG_X0Y0Z0 = 0, VFG_X1Y0Z0, VFG_X0Y1Z0, VFG_X1Y1Z0,
       VFG_X0Y0Z1, VFG_X1Y0Z1, VFG_X0Y1Z1, VFG_X1Y1Z1 };

/* grid of nodes, as "-n nx,ny,nz -l lx,ly,lz" (VFCommon.c:277-315) */
typedef struct {
  int    nx, ny, nz;
  double lx, ly, lz;
} vfg_grid;

/* VFMatProp + VFProp subset on the u/V path (VFCommon.h:69-90).  E is the
 * uniform Young's modulus, overridden per element by vfg_set_fields(E_elem). */
typedef struct {
  double E, nu, alpha, beta, Gc;
  double eta, epsilon, irrevtol;
  int    atnum;            /* 1: AT1, 2: AT2 */
  int    hasCrackPressure; /* -pressurize */
  int    unilateral;       /* VFG_UNILATERAL_* (-unilateral, VFCommon.h:120-123) */
} vfg_props;
/* UNILATERAL_NOCOMPRESSION: where tr(eps(U)) < 0 the spherical part of the
 * u-operator and of the thermo-poro load is not degraded, and the crack is
 * not driven by compressive volumetric energy (VFMech.c:115-218, 299-376,
 * 456-524, 819-908, 2125-2153).  The u-problem is then nonlinear: its
 * Jacobian depends on the state set by vfg_u_set_state, and vfg_u_solve runs
 * Newton iterations. */
enum { VFG_UNILATERAL_NONE = 0, VFG_UNILATERAL_NOCOMPRESSION = 1 };

/* VFBC (VFCartFE.h:129-136); edges are not parsed by the reference and are
 * absent here.  vertexValue is read for FIXED vertices. */
typedef struct {
  int    face[6];
  double faceValue[6];
  int    vertex[8];
  double vertexValue[8];
} vfg_bc;

typedef struct vfg_handle vfg_handle;

/* ------------------------------------------------------------ lifecycle */
/* Replaces VFGeometryInitialize/VFPropGet/VFMatPropGet for one problem
 * (VFCommon.c:277-588).  Fields start as V = VIrrev = 1, theta = thetaRef =
 * 0, pressure = 0, uniform E (VFFieldsInitialize, VFCommon.c:629-671, with the
 * thermal/pressure inputs supplied explicitly by the caller). */
int vfg_create(const vfg_grid *grid, const vfg_props *props, vfg_handle **out);
int vfg_destroy(vfg_handle *h);
const char *vfg_last_error(const vfg_handle *h);
int vfg_set_props(vfg_handle *h, const vfg_props *props);
/* VFBCSetFromOptions result for "U" (3 components) and "V" (1):
 * VFCartFE.c:1506-1574 */
int vfg_set_bc(vfg_handle *h, const vfg_bc u_bc[3], const vfg_bc *v_bc);
/* Nodal fields (N_n) and per-element E (N_e).  NULL keeps the current value. */
int vfg_set_fields(vfg_handle *h, const double *V, const double *VIrrev, const double *theta,
                   const double *thetaRef, const double *pressure, const double *E_elem,
                   int loc);
int vfg_get_fields(vfg_handle *h, double *V, double *VIrrev, int loc);
int64_t vfg_num_nodes(const vfg_handle *h);
int64_t vfg_num_cells(const vfg_handle *h);
int vfg_synchronize(vfg_handle *h);

/* ------------------------------------------------------------- u operator */
/* MatShell MATOP_MULT of the assembled u-Jacobian: VF_UIJacobian
 * (VFMech.c:1897-2029) then MatApplyDirichletBC (VFCartFE.c:1147-1296):
 * y = K x on free rows, y_i = x_i on Dirichlet rows (columns kept). */
int vfg_u_apply(vfg_handle *h, const double *x, double *y, int loc);
/* VF_UIJacobian(U) under UNILATERAL_NOCOMPRESSION (VFMech.c:1942-2023): sets
 * the state U (3*N_n) whose tension/compression pattern the u-Jacobian
 * (vfg_u_apply, vfg_u_diag and their FP32 twins) uses; the pattern of U = 0
 * (all tensile) until called.  No effect for UNILATERAL_NONE. */
int vfg_u_set_state(vfg_handle *h, const double *U, int loc);
int vfg_u_apply_f32(vfg_handle *h, const float *x, float *y, int loc);
/* How the FP64 u-apply (vfg_u_apply and the FP64 u-solve) holds the integrand
 * weights W_g s_g E_e of VF_BilinearFormU3D_local (VFMech.c:240-252): the
 * recomputing form forms s_g = v_g^2 + eta from the nodal V in every apply;
 * the stored form keeps the 27 products per element (216 B per element of
 * device memory, allocated on first use), refreshed on the first apply after
 * V, E or the properties change -- the role VF_UIJacobian's assembly plays
 * for the reference's MatMult (VFMech.c:1897-2029).  Both give the same
 * operator (to rounding).  Default: VFG_U_FORM_RECOMPUTE. */
enum { VFG_U_FORM_RECOMPUTE = 0, VFG_U_FORM_STORED = 1 };
int vfg_u_set_form(vfg_handle *h, int form);
/* PCApply of a PCSHELL around the u-MatShell: z = M r with M one FP32
 * geometric-multigrid V-cycle (the VFG_PC_MG preconditioner of vfg_u_solve;
 * FP64 in and out) on the free rows and z = r on the Dirichlet rows, the
 * matrix-free stand-in for the PCHYPRE/GAMG preconditioner the reference sets
 * on the u-KSP (VFCommon.c:1041-1047).  Each call refreshes the hierarchy's
 * coefficients from the current fields. */
int vfg_u_pc_apply(vfg_handle *h, const double *r, double *z, int loc);
/* MATOP_GET_DIAGONAL of the same matrix (1 on Dirichlet rows) */
int vfg_u_diag(vfg_handle *h, double *d, int loc);
/* element load vector sum_e b_e of VF_GradientUThermoPoro3D_local
 * (VFMech.c:386-447) + VF_GradientUCrackPressure3D_local (VFMech.c:533-589);
 * no boundary treatment */
int vfg_u_rhs(vfg_handle *h, double *b, int loc);
/* SNES function callback VF_UResidual (VFMech.c:1456-1887):
 * F = K U - b, then ResidualApplyDirichletBC (VFCartFE.c:747-999) */
int vfg_u_residual(vfg_handle *h, const double *U, const double *BCU, double *F, int loc);

/* ------------------------------------------------------------- V operator */
/* The V-Jacobian callback VF_VIJacobian (VFMech.c:2883-3029) depends on U
 * only through the coupling term; vfg_v_set_u refreshes it (and the V load
 * vector) from U.  Must be called before the V calls below. */
int vfg_v_set_u(vfg_handle *h, const double *U, int loc);
/* MATOP_MULT of J_V after MatApplyDirichletBCRowCol (VFCartFE.c:1305-1454)
 * and, for AT2, VF_IrrevApplyEQMat (VFMech.c:2637-2682): rows and columns
 * of fixed nodes replaced by identity. */
int vfg_v_apply(vfg_handle *h, const double *x, double *y, int loc);
int vfg_v_diag(vfg_handle *h, double *d, int loc);
/* element load vector: VF_ResidualVAT*3D_local (VFMech.c:2350-2384) +
 * VF_ResidualVCrackPressure3D_local (VFMech.c:2294-2341); no BC */
int vfg_v_rhs(vfg_handle *h, double *b, int loc);
/* SNES function callback VF_VResidual (VFMech.c:2687-2879) at the U of the
 * last vfg_v_set_u */
int vfg_v_residual(vfg_handle *h, const double *V, double *F, int loc);

/* ----------------------------------------------------------------- BC kit */
/* which: 0 = U (3 dof/node), 1 = V (1 dof/node) */
int vfg_bc_mask(vfg_handle *h, int which, uint8_t *mask, int loc);            /* VFCartFE.c:1147-1296 */
int vfg_vec_set_from_bc(vfg_handle *h, int which, double *B, int loc);        /* VFCartFE.c:279-483 */
int vfg_vec_apply_dirichlet(vfg_handle *h, int which, double *vec, const double *B,
                            int loc);                                         /* VFCartFE.c:493-738 */

/* ------------------------------------------------------------ diagnostics */
/* out[0] elastic energy, out[1] in-situ work (always 0: out of scope),
 * out[2] pressure work (VF_UEnergy3D, VFMech.c:1111-1427), out[3] surface
 * energy (VF_VEnergy3D, VFMech.c:2474-2526). */
int vfg_energies(vfg_handle *h, const double *U, const double *V, double out[4], int loc);
/* vfg_energies without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * U and V are device pointers, out is a host array; the reduction partials
 * live in a static 8 MiB device array per GPU, one group per workgroup of
 * 16 x 16 element columns by 8 layers (2 on grids of fewer than 64 such
 * columns), so grids are limited to 262,144 such blocks, 536,870,912 cells
 * (VFG_E_ARG beyond).  Calls that overlap on one GPU share
 * that array and race; the call synchronizes the handle's stream.  Results
 * are bitwise those of vfg_energies. */
int vfg_energies_static(vfg_handle *h, const double *U, const double *V, double out[4]);
/* VolumetricCrackOpening (VFPermfield.c:75-150); cellfield (N_e) optional */
int vfg_crack_volume(vfg_handle *h, const double *U, const double *V, double *vol,
                     double *cellfield, int loc);
/* per-cell energy terms, 3 per cell interleaved (cell3[3e + c]): c = 0 the
 * element's elastic energy (VF_ElasticEnergy3D_local, VFMech.c:725-806), 1 its
 * crack pressure work (VF_PressureWork3D_local, VFMech.c:920-974), 2 its
 * surface energy (VF_AT1/AT2SurfaceEnergy3D_local, VFMech.c:2394-2465); the
 * terms VF_UEnergy3D and VF_VEnergy3D sum (VFMech.c:1111-1427, 2474-2526).
 * Feeds vfg_cell_to_node with dof 3 for nodal energy fields. */
int vfg_energy_fields(vfg_handle *h, const double *U, const double *V, double *cell3, int loc);
/* CellToNodeInterpolation (VFPermfield.c:1156-1327): node (dof*N_n) = the
 * volume-weighted mean of the adjacent cells of cell (dof*N_e), components
 * interleaved, 1 <= dof <= 6.  Overwrites node; every reference caller zeroes
 * it first (VecSet(...,0.) at VFPermfield.c:144, 563, 1398, 1922 and
 * VFFlow_SNESStandardFEM.c:228, 340), so the results agree. */
int vfg_cell_to_node(vfg_handle *h, const double *cell, double *node, int dof, int loc);

/* ---------------------------------------------------------- crack geometry */
/* VFPennyCrackBuildVAT1/AT2 (VFCracks.c:136-227) for one crack into V */
int vfg_penny_build_v(vfg_handle *h, double *V, const double center[3], double r,
                      double theta_deg, double phi_deg, double thickness, int loc);

/* ----------------------------------------------------------------- solvers */
/* VFG_PC_MG (u-solve, default): geometric multigrid V-cycle in FP32, the
 * matrix-free counterpart of the reference's PCHYPRE / GAMG preconditioners
 * (VFCommon.c:1041-1047); any grid (coarse levels take ceil(n/2) elements per
 * direction down to 4). */
enum { VFG_PC_JACOBI = 0, VFG_PC_CHEBYSHEV = 1, VFG_PC_MG = 2 };
enum { VFG_PREC_FP64 = 0, VFG_PREC_MIXED = 1 };
typedef struct {
  double rtol;      /* relative residual tolerance (default 1e-10) */
  double atol;      /* absolute residual tolerance (default 0) */
  int    maxit;     /* iteration cap (default 100000) */
  int    pc;        /* VFG_PC_* (default VFG_PC_MG) */
  int    precision; /* VFG_PREC_* : mixed = FP32 inner solves inside FP64
                       iterative refinement */
  int    cheb_degree;
} vfg_solve_opts;
typedef struct {
  int    iterations;  /* inner iterations (CG steps) */
  int    outer;       /* refinement / active-set iterations */
  double relres;      /* final FP64 relative residual */
  int    status;      /* 0 or VFG_E_* */
} vfg_solve_info;
void vfg_solve_opts_default(vfg_solve_opts *o);
/* VF_StepU (VFMech.c:1436-1451): one Newton step of SNESNEWTONLS with KSPCG
 * on the linear u-problem.  U must carry its Dirichlet values on entry
 * (VFTimeStepPrepare, VFCommon.c:1132-1157).  UNILATERAL_NOCOMPRESSION:
 * Newton iterations (Jacobian at each iterate, FP64 CG under the selected
 * preconditioner, backtracking on ||F||) until ||F|| <= rtol ||F(U0)||;
 * info->outer counts them, precision is ignored. */
int vfg_u_solve(vfg_handle *h, double *U, const double *BCU, const vfg_solve_opts *o,
                vfg_solve_info *info, int loc);
/* VF_StepV (VFMech.c:3057-3082) with SNESVINEWTONRSLS semantics: bounds
 * 0 <= V <= ub (ub = NULL -> 1), Dirichlet and AT2-irreversible nodes held. */
int vfg_v_solve(vfg_handle *h, const double *U, double *V, const double *ub,
                const vfg_solve_opts *o, vfg_solve_info *info, int loc);

/* -------------------------------------------------------------- ensembles */
/* Independent problems of different sizes solved together.  The reference
 * solves one problem per run (VF_StepU, VFMech.c:1436-1451); an ensemble
 * runs that u-solve for every member at once: one persistent launch applies
 * every member's operator, CG iterations of all members are captured in a
 * HIP graph, and each member keeps its own scalars, so its result does not
 * depend on the other members.  Members must outlive the batch and be
 * UNILATERAL_NONE problems (VFG_E_ARG otherwise). */
typedef struct vfg_batch vfg_batch;
int vfg_batch_create(vfg_handle *const *members, int n, vfg_batch **out);
int vfg_batch_destroy(vfg_batch *b);
const char *vfg_batch_last_error(const vfg_batch *b);
int vfg_batch_size(const vfg_batch *b);
/* y[p] = A_p x[p] for every member (device pointers) in one persistent launch */
int vfg_batch_u_apply(vfg_batch *b, const double *const *x, double *const *y);
int vfg_batch_u_apply_f32(vfg_batch *b, const float *const *x, float *const *y);
/* times nrep persistent applies on internal vectors (HIP events), total ms */
int vfg_batch_time_u_apply(vfg_batch *b, int fp32, int nrep, double *ms);
/* picks the persistent apply's schedule (0: 16 x 16 element tiles on the
 * register march, 1: 8 x 8 tiles on the register march, 2: 16 x 16 tiles on
 * the LDS-ring march), z-chunk length (0.5 / 1 / 2 x the whole-round choice)
 * and grid size (one workgroup per item, one or two rounds of resident
 * workgroups) by timing every combination; best ms per apply */
int vfg_batch_autotune(vfg_batch *b, int fp32, double *best_ms);
typedef struct {
  int    schedule;  /* 0, 1, 2 as above */
  double oz_scale;  /* z-chunk length relative to the whole-round choice */
  int    grid;      /* 0: one workgroup per item; else the persistent grid */
  double ms;        /* measured time of one apply */
} vfg_tune_point;
/* the points the last vfg_batch_autotune timed (returns their number; copies
 * at most cap) and the configuration in use */
int vfg_batch_tune_points(const vfg_batch *b, vfg_tune_point *out, int cap);
int vfg_batch_tuning(const vfg_batch *b, vfg_tune_point *pick);
/* graph capture of `chunk` CG iterations per replay (enable = 0: plain launches) */
int vfg_batch_set_graph(vfg_batch *b, int enable, int chunk);
/* VF_StepU for every member; info has one entry per member */
int vfg_batch_u_solve(vfg_batch *b, double *const *U, const double *const *BCU, const vfg_solve_opts *o,
                      vfg_solve_info *info, int loc);

/* ---------------------------------------------------------------- serving */
/* Continuous batching of alternate-minimisation requests.  The reference runs
 * one problem per process and its loop (VFCommon.c:1192-1231, test1.c:53-69)
 * alternates VF_StepU and VF_StepV until ||V - Vold||_inf < altmintol; the
 * engine keeps up to max_slots such loops resident on one GPU.  Each call of
 * vfg_engine_step advances every resident request: one graph replay runs
 * `graph_chunk` CG iterations of every request currently in a u- or V-CG
 * phase (one persistent apply per operator over all of them), and between
 * replays each request takes its own refinement / active-set / line-search
 * decisions from its own device scalars, so its result is bitwise the same
 * alone and in any mix.  A request leaves after am_maxit iterations or when
 * ||V - Vold||_inf < altmintol; requests join at the next step.
 * Submitted arrays are device pointers owned by the caller (U, BCU of
 * 3*nn entries, V, VIrrev of nn), which must stay valid until the request
 * is reported by vfg_engine_poll; the handle keeps the request's fields. */
typedef struct vfg_engine vfg_engine;
typedef struct {
  int    max_slots;   /* requests resident at once (64) */
  int    precision;   /* VFG_PREC_MIXED (default): FP32 inner u-CG in FP64 refinement */
  double u_rtol;      /* u-solve relative residual (1e-10) */
  double v_rtol;      /* V-solve relative projected gradient (1e-8) */
  int    maxit;       /* CG iteration cap of an inner solve (100000) */
  int    graph_chunk; /* CG iterations per replay (1) */
  int    use_graph;   /* 1: captured hipGraph replays; 0: plain launches */
  int    pc;          /* u preconditioner of the mixed engine: VFG_PC_MG (default; every
                         resident request's FP32 V-cycle, batched level by level inside the
                         replays, under the FP32 CG of the mixed engine or the FP64 CG of the
                         FP64 engine) or VFG_PC_JACOBI */
} vfg_engine_opts;
typedef struct {
  int64_t tag;
  int     status;         /* 0 or the VFG_E_* of the first failing solve */
  int     am_iterations;  /* alternate-minimisation iterations done */
  int     u_its, v_outer, v_its;
  double  errV;           /* last ||V - Vold||_inf */
} vfg_engine_done;
void vfg_engine_opts_default(vfg_engine_opts *o);
int  vfg_engine_create(const vfg_engine_opts *o, vfg_engine **out);
int  vfg_engine_destroy(vfg_engine *e);
const char *vfg_engine_last_error(const vfg_engine *e);
int  vfg_engine_free_slots(const vfg_engine *e);
/* VIrrev is the AT1 upper bound (ignored for AT2); returns VFG_E_BUDGET when
 * no slot is free, VFG_E_ARG for a UNILATERAL_NOCOMPRESSION problem */
int  vfg_engine_submit(vfg_engine *e, vfg_handle *h, double *U, const double *BCU, double *V,
                       const double *VIrrev, int am_maxit, double altmintol, int64_t tag);
/* one scheduling step; returns the number of resident requests (>= 0) */
int  vfg_engine_step(vfg_engine *e);
/* finished requests since the last poll (at most max); returns the count */
int  vfg_engine_poll(vfg_engine *e, vfg_engine_done *out, int max);
int  vfg_engine_synchronize(vfg_engine *e);
int  vfg_engine_stats(const vfg_engine *e, int64_t *ticks, int64_t *replays);
/* device memory held by a handle (fields, coefficients, solver vectors) */
int64_t vfg_device_bytes(const vfg_handle *h);

/* ------------------------------------------------- device vector helpers */
/* The PETSc Vec operations the host loop of the reference performs on the
 * fields (VecCopy, VecPointwiseMin, VecAXPY+VecNorm(NORM_INFINITY), VecScale,
 * VecSet; VFCommon.c:827-857, 1141, 1209-1215), on device arrays of length n. */
int vfg_vec_copy(vfg_handle *h, const double *x, double *y, int64_t n);
int vfg_vec_pointwise_min(vfg_handle *h, const double *a, const double *b, double *out, int64_t n);
int vfg_vec_maxabs_diff(vfg_handle *h, const double *a, const double *b, int64_t n, double *out);
int vfg_vec_scale(vfg_handle *h, double *x, double s, int64_t n);
int vfg_vec_set(vfg_handle *h, double *x, double v, int64_t n);

/* ------------------------------------------- device memory for callers */
int vfg_device_alloc(size_t bytes, void **ptr);
int vfg_device_free(void *ptr);
/* kind: 0 host->device, 1 device->host, 2 device->device (synchronous) */
int vfg_memcpy(void *dst, const void *src, size_t bytes, int kind);

/* --------------------------------------------------------------- timing */
/* Runs `nrep` u-applies (device pointers; pairs (x[i % nbuf], y[i % nbuf]))
 * on the handle's stream between two HIP events; returns total ms. */
int vfg_time_u_apply(vfg_handle *h, const double *const *x, double *const *y, int nbuf,
                     int nrep, int fp32, double *ms);
/* The same for the masked V-Jacobian apply (needs vfg_v_set_u first). */
int vfg_time_v_apply(vfg_handle *h, const double *const *x, double *const *y, int nbuf,
                     int nrep, double *ms);

#ifdef __cplusplus
}
#endif
#endif
