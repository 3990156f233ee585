/*
 * vfhost.h -- C host layer above the vfg C-ABI: the parts of the reference's
 * context/driver code that sit on the u/V path, with every field resident on
 * the GPU.
 *
 *   vfh_options_*          VFCtxGet / VFPropGet / VFMatPropGet / VFGeometryInitialize
 *                          option subset and defaults (VFCommon.c:90-588),
 *                          VFBCSetFromOptions (VFCartFE.c:1506-1574),
 *                          VFPennyCrackGet (VFCracks.c:14-31)
 *   vfh_fields_initialize  V / VIrrev from penny cracks (VFCommon.c:822-857)
 *   vfh_timestep_prepare   VFTimeStepPrepare (VFCommon.c:1132-1157)
 *   vfh_altmin             the alternate-minimisation loop of
 *                          VFFractureTimeStep (VFCommon.c:1192-1231) and
 *                          test1.c:53-69: u-solve, V-solve, ||V - Vold||_inf
 *   vfh_energies           VF_UEnergy3D + VF_VEnergy3D + VolumetricCrackOpening
 *
 * Errors: negative VFG_E_* codes; vfh_last_error() gives the message.
 */
#ifndef VFHOST_H
#define VFHOST_H
#include "vfg.h"

#ifdef __cplusplus
extern "C" {
#endif

#define VFH_MAX_PENNY 64

typedef struct {
  double center[3], r, theta, phi, thickness;
} vfh_penny;

typedef struct {
  vfg_grid  grid;          /* -n nx,ny,nz -l lx,ly,lz */
  vfg_props props;         /* -E -nu -alpha -beta -Gc -epsilon -eta -irrevtol -atnum -pressurize */
  vfg_bc    bcU[3], bcV;   /* -U_<ENT>_BC_<c> TYPE, -U_<FACE>_<c> val, -V_<ENT>_BC TYPE */
  double    altmintol;     /* -altmintol (1e-4) */
  int       altminmaxit;   /* -altminmaxit (10000) */
  double    u_rtol;        /* -U_ksp_rtol (1e-8, the KSP default of VFSolversInitialize) */
  double    v_rtol;        /* -V_snes_rtol (1e-8) */
  int       npc;           /* -npc */
  vfh_penny pc[VFH_MAX_PENNY]; /* -pc<i>_center -pc<i>_r -pc<i>_theta -pc<i>_phi -pc<i>_thickness */
  int       precision;     /* VFG_PREC_FP64 (default) or VFG_PREC_MIXED (-vfg_precision mixed) */
  int       pc_type;       /* VFG_PC_MG (default; -U_pc_type hypre|gamg|mg, -vfg_pc mg),
                              VFG_PC_JACOBI (-vfg_pc jacobi, -U_pc_type jacobi) or
                              VFG_PC_CHEBYSHEV (-vfg_pc chebyshev, mixed precision) */
} vfh_options;

typedef struct vfh_ctx vfh_ctx;

typedef struct {
  int    iterations;   /* alternate-minimisation iterations done */
  int    u_its;        /* CG iterations of all u-solves */
  int    v_outer;      /* active-set Newton steps of all V-solves */
  int    v_its;        /* CG iterations of all V-solves */
  double errV;         /* last ||V - Vold||_inf */
  int    status;       /* 0 or VFG_E_* of the first failing solve */
} vfh_altmin_info;

void vfh_options_default(vfh_options *o);
/* PETSc-style option text ("-key value" tokens, '#' comments, case-insensitive
 * keys); may be called repeatedly (options files, then command line). */
int  vfh_options_parse(vfh_options *o, const char *text);
const char *vfh_last_error(const vfh_ctx *ctx);

int  vfh_create(const vfh_options *o, vfh_ctx **out);
int  vfh_destroy(vfh_ctx *ctx);
vfg_handle *vfh_handle(vfh_ctx *ctx);
/* device arrays owned by the context: U, BCU (3 N_n), V, VIrrev (N_n) */
int  vfh_device_fields(vfh_ctx *ctx, double **U, double **BCU, double **V, double **VIrrev);

/* material / loading inputs (host arrays, NULL keeps): per-element E,
 * temperature, reference temperature, pressure */
int  vfh_set_inputs(vfh_ctx *ctx, const double *E_elem, const double *theta, const double *thetaRef,
                    const double *pressure);
/* V = VIrrev = min over the penny cracks of their V profiles (or 1) */
int  vfh_fields_initialize(vfh_ctx *ctx);
/* BCU = scale * VecSetFromBC(bcU)  (test1.c:49-50 scales by the time value) */
int  vfh_set_bcu(vfh_ctx *ctx, double scale);
int  vfh_timestep_prepare(vfh_ctx *ctx);
/* maxit <= 0: run until ||V - Vold|| < altmintol or altminmaxit (the
 * do-while of test1.c:53-69); maxit > 0: exactly min(maxit, altminmaxit)
 * iterations unless converged earlier when stop_on_tol is set */
int  vfh_altmin(vfh_ctx *ctx, int maxit, int stop_on_tol, vfh_altmin_info *info);
/* out: elastic, in-situ work (0), pressure work, surface energy, crack volume */
int  vfh_energies(vfh_ctx *ctx, double out[5]);
/* fields->VolCrackOpening, the nodal crack-opening field of
 * VolumetricCrackOpening (VFPermfield.c:75-146): host array of N_n */
int  vfh_crack_opening_field(vfh_ctx *ctx, double *node);
/* per-cell (elastic, pressure work, surface) energy terms, host array of
 * 3 N_e interleaved (element terms of VFMech.c:1111-1427, 2474-2526) */
int  vfh_energy_fields(vfh_ctx *ctx, double *cell3);
int  vfh_get_state(vfh_ctx *ctx, double *U, double *V, double *VIrrev);
/* the options the context was created with */
const vfh_options *vfh_options_of(const vfh_ctx *ctx);

/* ------------------------------------------------------------ field output
 * PETSc binary Vec records as VecView writes them (int32 1211214, int32
 * length, doubles; big-endian), nvec vectors of len[v] values (vec[v] = NULL:
 * zeros); append != 0 adds to an existing file */
int  vfh_petsc_vec_write(const char *file, int nvec, const double *const *vec, const int64_t *len, int append);
/* FieldsBinaryWrite (VFCommon.c:1353-1375): <prefix>.<step, 5 digits>.bin
 * with U, velocity, V, pmult, theta, pressure, VolCrackOpening,
 * VolLeakOffRate; the flow fields as the reference initialises them (0),
 * theta and pressure from the caller (NULL: 0) */
int  vfh_fields_binary_write(vfh_ctx *ctx, const char *prefix, int step, const double *theta,
                             const double *pressure);
/* the Coordinates Vec of <prefix>.bin (VFCommon.c:383-424); the DMView record
 * the reference writes before it depends on the PETSc version and is omitted */
int  vfh_coordinates_write(vfh_ctx *ctx, const char *file);
/* VTK XML structured grid (.vts, ASCII) with named nodal and cell arrays */
int  vfh_vts_write(const char *file, int nx, int ny, int nz, double lx, double ly, double lz, int npt,
                   const char *const *pnames, const double *const *pdata, int ncell,
                   const char *const *cnames, const double *const *cdata);
/* FieldsVTKWrite (VFCommon.c:1415-1460) for this path's fields:
 * <prefix>_nodal.<step>.vts (Displacement_0..2, Fracture, Temperature,
 * Pressure, Volumetric Crack Opening) and <prefix>_cell.<step>.vts (elastic
 * energy, pressure work, surface energy of each cell) */
int  vfh_fields_vts_write(vfh_ctx *ctx, const char *prefix, int step, const double *theta,
                          const double *pressure);
int  vfh_set_state(vfh_ctx *ctx, const double *U, const double *V, const double *VIrrev);

/* ---------------------------------------------------------------- serving
 * A request loop over the vfg engine: requests (each a complete problem given
 * by the reference's option text plus its input fields) arrive at their
 * arrival times, are admitted first come first served while an engine slot is
 * free and the HBM they will hold fits the budget, are set up on the GPU
 * (VFGeometryInitialize ... VFTimeStepPrepare, as a test driver does) and run
 * their alternate-minimisation iterations in the engine's continuous batch;
 * their fields stay in HBM from admission to completion.  Latency is
 * completion - arrival. */
typedef struct {
  const char   *options;                /* reference option text of the request */
  const double *E_elem, *theta, *thetaRef, *pressure;  /* host inputs (NULL: defaults) */
  double        arrival;                /* seconds after the start of the run */
  int           am_maxit;               /* iterations (fewer once ||V - Vold|| < altmintol) */
  double       *U_out, *V_out;          /* optional host copies of the final U and V */
  double       *opening_out;            /* optional nodal crack opening (N_n) */
  double       *energy_out;             /* optional per-cell energy terms (3 N_e) */
} vfh_request;

typedef struct {
  int     status;                       /* 0, VFG_E_BUDGET (larger than the budget) or a solver code */
  int     am_iterations, u_its, v_outer, v_its;
  double  errV;
  double  t_admit, t_done, latency;     /* seconds since the start; completion - arrival */
  double  energies[5];                  /* elastic, in-situ (0), pressure work, surface, crack volume */
  int64_t device_bytes;                 /* HBM the request held */
} vfh_request_result;

typedef struct {
  vfg_engine_opts engine;
  double hbm_budget;                    /* bytes resident requests may hold (200e9) */
  int    realtime;                      /* 1: honour arrival times; 0: all arrive at t = 0 */
} vfh_serve_opts;

typedef struct {
  double  wall, solves, solves_per_s;   /* a solve is one u- or one V-solve */
  double  p50, p99, mean_latency, max_latency;
  int     completed, rejected, max_resident;
  double  peak_bytes;                   /* largest admitted total (estimate) */
  int64_t ticks, replays;
} vfh_serve_stats;

void    vfh_serve_opts_default(vfh_serve_opts *o);
/* HBM one request of nx*ny*nz nodes holds while resident (admission estimate) */
int64_t vfh_request_bytes(int nx, int ny, int nz);
int     vfh_serve(const vfh_serve_opts *o, const vfh_request *req, int nreq, vfh_request_result *res,
                  vfh_serve_stats *stats);

#ifdef __cplusplus
}
#endif
#endif
