/*
 * vfhost.c -- host layer of the u/V path in C over the vfg C-ABI.
 *
 * Restates, for one problem on one GPU, the reference's option handling
 * (VFCtxGet/VFPropGet/VFMatPropGet/VFGeometryInitialize, VFCommon.c:90-588;
 * VFBCSetFromOptions, VFCartFE.c:1506-1574; VFPennyCrackGet, VFCracks.c:14-31),
 * the crack part of VFFieldsInitialize (VFCommon.c:822-857), VFTimeStepPrepare
 * (VFCommon.c:1132-1157) and the alternate-minimisation loop of
 * VFFractureTimeStep (VFCommon.c:1192-1231).  All fields stay on the device;
 * the host only moves scalars.
 */
#define _POSIX_C_SOURCE 200809L
#include "../../include/vfhost.h"
#include <ctype.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include <time.h>

struct vfh_ctx {
  vfh_options o;
  vfg_handle *h;
  int64_t nn;
  double *U, *BCU, *V, *VIrrev, *Vold, *tmp;
  char err[512];
};

static char g_err[512];

static int fail(vfh_ctx *c, int code, const char *msg) {
  snprintf(c ? c->err : g_err, 512, "%s", msg);
  if (c) snprintf(g_err, sizeof g_err, "%s", msg);
  return code;
}
static int vchk(vfh_ctx *c, int rc) {
  if (rc) fail(c, rc, vfg_last_error(c->h));
  return rc;
}
#define CHK(c, call)                  \
  do {                                \
    int rc_ = (call);                 \
    if (rc_) return vchk((c), rc_);   \
  } while (0)

const char *vfh_last_error(const vfh_ctx *c) { return c ? c->err : g_err; }

/* defaults of VFGeometryInitialize (VFCommon.c:297-313), VFPropGet
 * (VFCommon.c:469-491), VFMatPropGet (VFCommon.c:513-520), VFCtxGet
 * (VFCommon.c:171-178) and VFSolversInitialize (VFCommon.c:1041) */
void vfh_options_default(vfh_options *o) {
  memset(o, 0, sizeof *o);
  o->grid.nx = o->grid.ny = o->grid.nz = 11;
  o->grid.lx = o->grid.ly = o->grid.lz = 1.;
  o->props.E = 1.; o->props.nu = 0.; o->props.alpha = 1e-5; o->props.beta = 1e-4; o->props.Gc = 1.;
  o->props.eta = 1e-5; o->props.epsilon = 2e-1; o->props.irrevtol = 5e-2; o->props.atnum = 2;
  o->props.hasCrackPressure = 0;
  o->altmintol = 1e-4;
  o->altminmaxit = 10000;
  o->u_rtol = 1e-8;
  o->v_rtol = 1e-8;
  o->precision = VFG_PREC_FP64;
  o->pc_type = VFG_PC_MG;
}

static int ieq(const char *a, const char *b) {
  for (; *a && *b; a++, b++)
    if (tolower((unsigned char)*a) != tolower((unsigned char)*b)) return 0;
  return *a == *b;
}

static int parse_reals(const char *v, double *out, int maxn) {
  int n = 0;
  while (v && *v && n < maxn) {
    char *end;
    out[n] = strtod(v, &end);
    if (end == v) return -1;
    n++;
    v = end;
    if (*v == ',') v++;
  }
  return n;
}

static const char *FACE[6] = {"X0", "X1", "Y0", "Y1", "Z0", "Z1"};
static const char *VERT[8] = {"X0Y0Z0", "X1Y0Z0", "X0Y1Z0", "X1Y1Z0", "X0Y0Z1", "X1Y0Z1", "X0Y1Z1", "X1Y1Z1"};
static const char *BCT[4] = {"NONE", "ZERO", "ONE", "FIXED"};

/* -U_<ENT>_BC_<c> TYPE | -U_<FACE>_<c> value | -V_<ENT>_BC TYPE | -V_<FACE> value */
static int parse_bc(vfh_options *o, const char *key, const char *val) {
  char want[96];
  for (int which = 0; which < 2; which++) {
    int dof = which == 0 ? 3 : 1;
    const char *pre = which == 0 ? "U" : "V";
    vfg_bc *bc = which == 0 ? o->bcU : &o->bcV;
    for (int c = 0; c < dof; c++)
      for (int e = 0; e < 14; e++) {
        const char *ent = e < 6 ? FACE[e] : VERT[e - 6];
        if (dof == 1) snprintf(want, sizeof want, "-%s_%s_BC", pre, ent);
        else snprintf(want, sizeof want, "-%s_%s_BC_%d", pre, ent, c);
        if (ieq(key, want)) {
          int t = -1;
          for (int q = 0; q < 4; q++)
            if (val && ieq(val, BCT[q])) t = q;
          if (t < 0) return -1;
          if (e < 6) bc[c].face[e] = t; else bc[c].vertex[e - 6] = t;
          return 1;
        }
        if (e < 6) {
          if (dof == 1) snprintf(want, sizeof want, "-%s_%s", pre, ent);
          else snprintf(want, sizeof want, "-%s_%s_%d", pre, ent, c);
          if (ieq(key, want)) {
            if (!val) return -1;
            bc[c].faceValue[e] = atof(val);
            return 1;
          }
        }
      }
  }
  return 0;
}

int vfh_options_parse(vfh_options *o, const char *text) {
  if (!o || !text) return fail(NULL, VFG_E_ARG, "vfh_options_parse: NULL argument");
  char *buf = strdup(text);
  char *tok[4096];
  int nt = 0;
  for (char *p = buf; *p;) {
    while (*p && isspace((unsigned char)*p)) p++;
    if (!*p) break;
    if (*p == '#') {
      while (*p && *p != '\n') p++;
      continue;
    }
    if (nt < 4096) tok[nt++] = p;
    while (*p && !isspace((unsigned char)*p)) p++;
    if (*p) *p++ = 0;
  }
  int rc = 0;
  for (int i = 0; i < nt && !rc; i++) {
    const char *k = tok[i];
    if (k[0] != '-' || !(isalpha((unsigned char)k[1]) || k[1] == '_')) continue;
    const char *v = NULL;
    if (i + 1 < nt && !(tok[i + 1][0] == '-' && (isalpha((unsigned char)tok[i + 1][1]) || tok[i + 1][1] == '_')))
      v = tok[++i];
    double r[3];
    int pi;
    char sub[64];
    if (ieq(k, "-n")) {
      if (parse_reals(v, r, 3) != 3) rc = -1;
      else { o->grid.nx = (int)r[0]; o->grid.ny = (int)r[1]; o->grid.nz = (int)r[2]; }
    } else if (ieq(k, "-l")) {
      if (parse_reals(v, r, 3) != 3) rc = -1;
      else { o->grid.lx = r[0]; o->grid.ly = r[1]; o->grid.lz = r[2]; }
    } else if (ieq(k, "-E") && v) o->props.E = atof(v);
    else if (ieq(k, "-nu") && v) o->props.nu = atof(v);
    else if (ieq(k, "-alpha") && v) o->props.alpha = atof(v);
    else if (ieq(k, "-beta") && v) o->props.beta = atof(v);
    else if (ieq(k, "-Gc") && v) o->props.Gc = atof(v);
    else if (ieq(k, "-epsilon") && v) o->props.epsilon = atof(v);
    else if (ieq(k, "-eta") && v) o->props.eta = atof(v);
    else if (ieq(k, "-irrevtol") && v) o->props.irrevtol = atof(v);
    else if (ieq(k, "-atnum") && v) o->props.atnum = atoi(v);
    else if (ieq(k, "-unilateral") && v) { /* PetscOptionsEnum, VFCommon.c:176-177 */
      if (ieq(v, "NOCOMPRESSION") || ieq(v, "UNILATERAL_NOCOMPRESSION") || ieq(v, "1"))
        o->props.unilateral = VFG_UNILATERAL_NOCOMPRESSION;
      else if (ieq(v, "NONE") || ieq(v, "UNILATERAL_NONE") || ieq(v, "0"))
        o->props.unilateral = VFG_UNILATERAL_NONE;
      else
        rc = -1;
    }
    else if (ieq(k, "-pressurize")) o->props.hasCrackPressure = !(v && (ieq(v, "0") || ieq(v, "false")));
    else if (ieq(k, "-altmintol") && v) o->altmintol = atof(v);
    else if (ieq(k, "-altminmaxit") && v) o->altminmaxit = atoi(v);
    else if (ieq(k, "-U_ksp_rtol") && v) o->u_rtol = atof(v);
    else if (ieq(k, "-V_snes_rtol") && v) o->v_rtol = atof(v);
    else if (ieq(k, "-npc") && v) {
      o->npc = atoi(v);
      if (o->npc < 0 || o->npc > VFH_MAX_PENNY) rc = -1;
    } else if (ieq(k, "-vfg_precision") && v) o->precision = ieq(v, "mixed") ? VFG_PREC_MIXED : VFG_PREC_FP64;
    else if (ieq(k, "-vfg_pc") && v)
      o->pc_type = ieq(v, "chebyshev") ? VFG_PC_CHEBYSHEV : ieq(v, "mg") ? VFG_PC_MG : VFG_PC_JACOBI;
    /* the reference's multigrid choices for the u-KSP (VFCommon.c:1041-1047,
     * scripts/test16-*.sh) map to the matrix-free multigrid */
    else if (ieq(k, "-U_pc_type") && v)
      o->pc_type = (ieq(v, "hypre") || ieq(v, "gamg") || ieq(v, "mg") || ieq(v, "ml")) ? VFG_PC_MG : VFG_PC_JACOBI;
    else if (sscanf(k, "-pc%d_%63s", &pi, sub) == 2 && pi >= 0 && pi < VFH_MAX_PENNY) {
      vfh_penny *c = &o->pc[pi];
      if (ieq(sub, "center")) { if (parse_reals(v, c->center, 3) != 3) rc = -1; }
      else if (ieq(sub, "r") && v) c->r = atof(v);
      else if (ieq(sub, "theta") && v) c->theta = atof(v);
      else if (ieq(sub, "phi") && v) c->phi = atof(v);
      else if (ieq(sub, "thickness") && v) c->thickness = atof(v);
    } else if (k[1] == 'U' || k[1] == 'V' || k[1] == 'u' || k[1] == 'v') {
      if (parse_bc(o, k, v) < 0) rc = -1;
    }
    if (rc) snprintf(g_err, sizeof g_err, "vfh_options_parse: bad value for %s", k);
  }
  free(buf);
  return rc ? VFG_E_ARG : VFG_OK;
}

int vfh_create(const vfh_options *o, vfh_ctx **out) {
  if (!o || !out) return fail(NULL, VFG_E_ARG, "vfh_create: NULL argument");
  *out = NULL;
  vfh_ctx *c = calloc(1, sizeof *c);
  c->o = *o;
  int rc = vfg_create(&o->grid, &o->props, &c->h);
  if (rc) {
    fail(NULL, rc, vfg_last_error(NULL));
    free(c);
    return rc;
  }
  c->nn = vfg_num_nodes(c->h);
  double **bufs[6] = {&c->U, &c->BCU, &c->V, &c->VIrrev, &c->Vold, &c->tmp};
  size_t sz[6] = {24, 24, 8, 8, 8, 8};
  for (int i = 0; i < 6; i++) {
    if ((rc = vfg_device_alloc(sz[i] * (size_t)c->nn, (void **)bufs[i]))) {
      fail(c, rc, vfg_last_error(NULL));
      snprintf(g_err, sizeof g_err, "%s", c->err);
      vfh_destroy(c);
      return rc;
    }
  }
  if ((rc = vfg_set_bc(c->h, o->bcU, &o->bcV)) || (rc = vfg_vec_set(c->h, c->U, 0., 3 * c->nn)) ||
      (rc = vfg_vec_set(c->h, c->BCU, 0., 3 * c->nn)) || (rc = vfg_vec_set(c->h, c->V, 1., c->nn)) ||
      (rc = vfg_vec_set(c->h, c->VIrrev, 1., c->nn)) || (rc = vfg_synchronize(c->h))) {
    fail(c, rc, vfg_last_error(c->h));
    snprintf(g_err, sizeof g_err, "%s", c->err);
    vfh_destroy(c);
    return rc;
  }
  *out = c;
  return VFG_OK;
}

int vfh_destroy(vfh_ctx *c) {
  if (!c) return VFG_OK;
  double *bufs[6] = {c->U, c->BCU, c->V, c->VIrrev, c->Vold, c->tmp};
  for (int i = 0; i < 6; i++)
    if (bufs[i]) vfg_device_free(bufs[i]);
  if (c->h) vfg_destroy(c->h);
  free(c);
  return VFG_OK;
}

vfg_handle *vfh_handle(vfh_ctx *c) { return c ? c->h : NULL; }

int vfh_device_fields(vfh_ctx *c, double **U, double **BCU, double **V, double **VIrrev) {
  if (!c) return fail(NULL, VFG_E_ARG, "NULL context");
  if (U) *U = c->U;
  if (BCU) *BCU = c->BCU;
  if (V) *V = c->V;
  if (VIrrev) *VIrrev = c->VIrrev;
  return VFG_OK;
}

int vfh_set_inputs(vfh_ctx *c, const double *E_elem, const double *theta, const double *thetaRef,
                   const double *pressure) {
  if (!c) return fail(NULL, VFG_E_ARG, "NULL context");
  CHK(c, vfg_set_fields(c->h, NULL, NULL, theta, thetaRef, pressure, E_elem, VFG_HOST));
  return VFG_OK;
}

/* VecSet(V,1); VecSet(VIrrev,1); per crack: Build V; VIrrev = min(VIrrev, V);
 * then V = VIrrev (VFCommon.c:826-857) */
int vfh_fields_initialize(vfh_ctx *c) {
  if (!c) return fail(NULL, VFG_E_ARG, "NULL context");
  CHK(c, vfg_vec_set(c->h, c->V, 1., c->nn));
  CHK(c, vfg_vec_set(c->h, c->VIrrev, 1., c->nn));
  for (int i = 0; i < c->o.npc; i++) {
    const vfh_penny *p = &c->o.pc[i];
    CHK(c, vfg_penny_build_v(c->h, c->V, p->center, p->r, p->theta, p->phi, p->thickness, VFG_DEVICE));
    CHK(c, vfg_vec_pointwise_min(c->h, c->VIrrev, c->V, c->VIrrev, c->nn));
  }
  CHK(c, vfg_vec_copy(c->h, c->VIrrev, c->V, c->nn));
  CHK(c, vfg_set_fields(c->h, c->V, c->VIrrev, NULL, NULL, NULL, NULL, VFG_DEVICE));
  return VFG_OK;
}

int vfh_set_bcu(vfh_ctx *c, double scale) {
  if (!c) return fail(NULL, VFG_E_ARG, "NULL context");
  CHK(c, vfg_vec_set(c->h, c->BCU, 0., 3 * c->nn));
  CHK(c, vfg_vec_set_from_bc(c->h, 0, c->BCU, VFG_DEVICE));
  CHK(c, vfg_vec_scale(c->h, c->BCU, scale, 3 * c->nn));
  return VFG_OK;
}

/* VFTimeStepPrepare: VIrrev = V; U|D = BCU; V|D from its own BC values; the
 * AT1 upper bound is VIrrev (VFCommon.c:1141-1154) */
int vfh_timestep_prepare(vfh_ctx *c) {
  if (!c) return fail(NULL, VFG_E_ARG, "NULL context");
  CHK(c, vfg_vec_copy(c->h, c->V, c->VIrrev, c->nn));
  CHK(c, vfg_vec_apply_dirichlet(c->h, 0, c->U, c->BCU, VFG_DEVICE));
  CHK(c, vfg_vec_apply_dirichlet(c->h, 1, c->V, c->V, VFG_DEVICE));
  CHK(c, vfg_set_fields(c->h, c->V, c->VIrrev, NULL, NULL, NULL, NULL, VFG_DEVICE));
  return VFG_OK;
}

int vfh_altmin(vfh_ctx *c, int maxit, int stop_on_tol, vfh_altmin_info *info) {
  if (!c) return fail(NULL, VFG_E_ARG, "NULL context");
  vfh_altmin_info inf;
  memset(&inf, 0, sizeof inf);
  vfg_solve_opts uo, vo;
  vfg_solve_opts_default(&uo);
  vfg_solve_opts_default(&vo);
  uo.rtol = c->o.u_rtol;
  uo.precision = c->o.precision;
  uo.pc = c->o.pc_type;
  vo.rtol = c->o.v_rtol;
  int rc = VFG_OK;
  double errV = 1e10;
  for (;;) {
    vfg_solve_info si = {0, 0, 0., 0};
    rc = vfg_u_solve(c->h, c->U, c->BCU, &uo, &si, VFG_DEVICE);
    inf.u_its += si.iterations;
    if (rc) break;
    if ((rc = vfg_vec_copy(c->h, c->V, c->Vold, c->nn))) break;
    vfg_solve_info sv = {0, 0, 0., 0};
    rc = vfg_v_solve(c->h, c->U, c->V, c->o.props.atnum == 1 ? c->VIrrev : NULL, &vo, &sv, VFG_DEVICE);
    inf.v_outer += sv.outer;
    inf.v_its += sv.iterations;
    if (rc) break;
    if ((rc = vfg_set_fields(c->h, c->V, NULL, NULL, NULL, NULL, NULL, VFG_DEVICE))) break;
    if ((rc = vfg_vec_maxabs_diff(c->h, c->V, c->Vold, c->nn, &errV))) break;
    inf.iterations++;
    if (maxit > 0) {
      if (inf.iterations >= maxit || inf.iterations > c->o.altminmaxit) break;
      if (stop_on_tol && errV < c->o.altmintol) break;
    } else if (!(errV >= c->o.altmintol && inf.iterations <= c->o.altminmaxit)) {
      break;  /* do { ... altminit++ } while (errV >= altmintol && altminit <= altminmaxit) */
    }
  }
  inf.errV = errV;
  inf.status = rc;
  if (info) *info = inf;
  if (rc) return vchk(c, rc);
  return VFG_OK;
}

int vfh_energies(vfh_ctx *c, double out[5]) {
  if (!c || !out) return fail(c, VFG_E_ARG, "vfh_energies: NULL argument");
  double e[4], vol = 0.;
  CHK(c, vfg_energies(c->h, c->U, c->V, e, VFG_DEVICE));
  CHK(c, vfg_crack_volume(c->h, c->U, c->V, &vol, NULL, VFG_DEVICE));
  out[0] = e[0]; out[1] = e[1]; out[2] = e[2]; out[3] = e[3]; out[4] = vol;
  return VFG_OK;
}

int vfh_crack_opening_field(vfh_ctx *c, double *node) {
  if (!c || !node) return fail(c, VFG_E_ARG, "vfh_crack_opening_field: NULL argument");
  const int64_t ne = (int64_t)(c->o.grid.nx - 1) * (c->o.grid.ny - 1) * (c->o.grid.nz - 1);
  double *d = NULL, vol = 0.;
  CHK(c, vfg_device_alloc(8 * (size_t)(ne + c->nn), (void **)&d));
  int rc = vfg_crack_volume(c->h, c->U, c->V, &vol, d, VFG_DEVICE);
  if (!rc) rc = vfg_cell_to_node(c->h, d, d + ne, 1, VFG_DEVICE);
  if (!rc) rc = vfg_synchronize(c->h);
  if (!rc) rc = vfg_memcpy(node, d + ne, 8 * (size_t)c->nn, 1);
  vfg_device_free(d);
  return vchk(c, rc);
}

int vfh_energy_fields(vfh_ctx *c, double *cell3) {
  if (!c || !cell3) return fail(c, VFG_E_ARG, "vfh_energy_fields: NULL argument");
  const int64_t ne = (int64_t)(c->o.grid.nx - 1) * (c->o.grid.ny - 1) * (c->o.grid.nz - 1);
  double *d = NULL;
  CHK(c, vfg_device_alloc(24 * (size_t)ne, (void **)&d));
  int rc = vfg_energy_fields(c->h, c->U, c->V, d, VFG_DEVICE);
  if (!rc) rc = vfg_memcpy(cell3, d, 24 * (size_t)ne, 1);
  vfg_device_free(d);
  return vchk(c, rc);
}

const vfh_options *vfh_options_of(const vfh_ctx *c) { return c ? &c->o : NULL; }

int vfh_get_state(vfh_ctx *c, double *U, double *V, double *VIrrev) {
  if (!c) return fail(NULL, VFG_E_ARG, "NULL context");
  CHK(c, vfg_synchronize(c->h));
  if (U) CHK(c, vfg_memcpy(U, c->U, 24 * (size_t)c->nn, 1));
  if (V) CHK(c, vfg_memcpy(V, c->V, 8 * (size_t)c->nn, 1));
  if (VIrrev) CHK(c, vfg_memcpy(VIrrev, c->VIrrev, 8 * (size_t)c->nn, 1));
  return VFG_OK;
}

int vfh_set_state(vfh_ctx *c, const double *U, const double *V, const double *VIrrev) {
  if (!c) return fail(NULL, VFG_E_ARG, "NULL context");
  CHK(c, vfg_synchronize(c->h));
  if (U) CHK(c, vfg_memcpy(c->U, U, 24 * (size_t)c->nn, 0));
  if (V) CHK(c, vfg_memcpy(c->V, V, 8 * (size_t)c->nn, 0));
  if (VIrrev) CHK(c, vfg_memcpy(c->VIrrev, VIrrev, 8 * (size_t)c->nn, 0));
  CHK(c, vfg_set_fields(c->h, V ? c->V : NULL, VIrrev ? c->VIrrev : NULL, NULL, NULL, NULL, NULL, VFG_DEVICE));
  return VFG_OK;
}

/* ------------------------------------------------------------------ serving */
static double now_s(void) {
  struct timespec t;
  clock_gettime(CLOCK_MONOTONIC, &t);
  return (double)t.tv_sec + 1e-9 * (double)t.tv_nsec;
}

void vfh_serve_opts_default(vfh_serve_opts *o) {
  if (!o) return;
  vfg_engine_opts_default(&o->engine);
  o->hbm_budget = 200e9;
  o->realtime = 1;
}

/* context arrays 80 B/node; handle fields, staging and solver vectors
 * 446 B/node; the multigrid hierarchy of the u-solve 70 B/node (FP32
 * level-0 vectors, coarse levels, transfer and set-up work); E 8 B/cell; the
 * V coupling table, 27 doubles per element slot of every 16 x 16 workgroup
 * tile and element layer; reduction scratch */
int64_t vfh_request_bytes(int nx, int ny, int nz) {
  const int64_t nn = (int64_t)nx * ny * nz, ne = (int64_t)(nx - 1) * (ny - 1) * (nz - 1);
  const int64_t vq = (int64_t)((nx + 14) / 15) * ((ny + 14) / 15) * (nz - 1) * 27 * 256 * 8;
  return 600 * nn + 8 * ne + vq + ((int64_t)5 << 20);
}

static int request_setup(const vfh_request *r, const vfh_options *o, vfh_ctx **out) {
  vfh_ctx *c;
  int rc = vfh_create(o, &c);
  if (rc) return rc;
  if ((rc = vfh_set_inputs(c, r->E_elem, r->theta, r->thetaRef, r->pressure)) || (rc = vfh_fields_initialize(c)) ||
      (rc = vfh_set_bcu(c, 1.0)) || (rc = vfh_timestep_prepare(c)) || (rc = vfg_synchronize(c->h))) {
    snprintf(g_err, sizeof g_err, "%s", c->err);
    vfh_destroy(c);
    return rc;
  }
  *out = c;
  return VFG_OK;
}

static int cmp_dbl(const void *a, const void *b) {
  double x = *(const double *)a, y = *(const double *)b;
  return x < y ? -1 : x > y;
}
static double pct(const double *sorted, int n, double q) { /* nearest rank */
  if (n < 1) return 0.;
  int k = (int)ceil(q * n) - 1;
  if (k < 0) k = 0;
  if (k >= n) k = n - 1;
  return sorted[k];
}

int vfh_serve(const vfh_serve_opts *so, const vfh_request *req, int nreq, vfh_request_result *res,
              vfh_serve_stats *stats) {
  if (!req || nreq < 1 || !res) return fail(NULL, VFG_E_ARG, "vfh_serve: bad arguments");
  vfh_serve_opts d;
  if (!so) {
    vfh_serve_opts_default(&d);
    so = &d;
  }
  memset(res, 0, sizeof(*res) * (size_t)nreq);
  int *order = malloc(sizeof(int) * (size_t)nreq);
  vfh_ctx **ctx = calloc((size_t)nreq, sizeof *ctx);
  int64_t *need = calloc((size_t)nreq, sizeof *need);
  vfh_options *opt = malloc(sizeof(vfh_options) * (size_t)nreq);
  double *lat = malloc(sizeof(double) * (size_t)nreq);
  if (!order || !ctx || !need || !opt || !lat) {
    free(order); free(ctx); free(need); free(opt); free(lat);
    return fail(NULL, VFG_E_NOMEM, "vfh_serve: host allocation failed");
  }
  int rc = VFG_OK;
  for (int i = 0; i < nreq && !rc; i++) {
    order[i] = i;
    vfh_options_default(&opt[i]);
    if (!req[i].options || (rc = vfh_options_parse(&opt[i], req[i].options))) {
      if (!rc) rc = fail(NULL, VFG_E_ARG, "vfh_serve: request without options");
      break;
    }
    if (req[i].am_maxit < 1) rc = fail(NULL, VFG_E_ARG, "vfh_serve: am_maxit < 1");
    need[i] = vfh_request_bytes(opt[i].grid.nx, opt[i].grid.ny, opt[i].grid.nz);
  }
  /* arrival order (stable insertion sort) */
  for (int i = 1; i < nreq; i++) {
    int k = order[i], j = i - 1;
    while (j >= 0 && req[order[j]].arrival > req[k].arrival) {
      order[j + 1] = order[j];
      j--;
    }
    order[j + 1] = k;
  }
  vfg_engine *e = NULL;
  if (!rc && (rc = vfg_engine_create(&so->engine, &e))) fail(NULL, rc, "vfh_serve: vfg_engine_create failed");
  int next = 0, head = 0, done = 0, resident = 0, max_res = 0, rejected = 0;
  double used = 0., peak = 0.;
  vfg_engine_done dn[64];
  const double t0 = now_s();
  while (!rc && done < nreq) {
    const double t = now_s() - t0;
    while (next < nreq && (!so->realtime || req[order[next]].arrival <= t)) next++;
    while (head < next) { /* first come, first served */
      const int i = order[head];
      if ((double)need[i] > so->hbm_budget) {
        res[i].status = VFG_E_BUDGET;
        res[i].t_admit = res[i].t_done = now_s() - t0;
        head++;
        done++;
        rejected++;
        continue;
      }
      if (vfg_engine_free_slots(e) < 1 || used + (double)need[i] > so->hbm_budget) break;
      if ((rc = request_setup(&req[i], &opt[i], &ctx[i]))) break;
      vfh_ctx *c = ctx[i];
      if ((rc = vfg_engine_submit(e, c->h, c->U, c->BCU, c->V, c->VIrrev, req[i].am_maxit, opt[i].altmintol, i))) {
        fail(NULL, rc, vfg_engine_last_error(e));
        break;
      }
      res[i].t_admit = now_s() - t0;
      used += (double)need[i];
      if (used > peak) peak = used;
      head++;
      if (++resident > max_res) max_res = resident;
    }
    if (rc) break;
    if (resident == 0) {
      if (next < nreq && head == next && so->realtime) { /* idle until the next arrival */
        double wait = req[order[next]].arrival - (now_s() - t0);
        if (wait > 0.) {
          struct timespec ts;
          ts.tv_sec = (time_t)wait;
          ts.tv_nsec = (long)((wait - (double)ts.tv_sec) * 1e9);
          nanosleep(&ts, NULL);
        }
      }
      continue;
    }
    int r = vfg_engine_step(e);
    if (r < 0) {
      rc = fail(NULL, r, vfg_engine_last_error(e));
      break;
    }
    int nd;
    while (!rc && (nd = vfg_engine_poll(e, dn, 64)) > 0) {
      for (int k = 0; k < nd && !rc; k++) {
        const int i = (int)dn[k].tag;
        vfh_request_result *q = &res[i];
        vfh_ctx *c = ctx[i];
        q->status = dn[k].status;
        q->am_iterations = dn[k].am_iterations;
        q->u_its = dn[k].u_its;
        q->v_outer = dn[k].v_outer;
        q->v_its = dn[k].v_its;
        q->errV = dn[k].errV;
        q->device_bytes = vfg_device_bytes(c->h) + 80 * c->nn;
        if ((rc = vfh_energies(c, q->energies)) ||
            ((req[i].U_out || req[i].V_out) && (rc = vfh_get_state(c, req[i].U_out, req[i].V_out, NULL))) ||
            (req[i].opening_out && (rc = vfh_crack_opening_field(c, req[i].opening_out))) ||
            (req[i].energy_out && (rc = vfh_energy_fields(c, req[i].energy_out)))) {
          snprintf(g_err, sizeof g_err, "%s", c->err);
          break;
        }
        q->t_done = now_s() - t0;
        q->latency = q->t_done - (so->realtime ? req[i].arrival : 0.);
        vfh_destroy(c);
        ctx[i] = NULL;
        used -= (double)need[i];
        resident--;
        done++;
      }
    }
  }
  const double wall = now_s() - t0;
  if (stats) {
    memset(stats, 0, sizeof *stats);
    int n = 0;
    double solves = 0., sum = 0.;
    for (int i = 0; i < nreq; i++) {
      if (res[i].status == VFG_E_BUDGET) continue;
      lat[n++] = res[i].latency;
      sum += res[i].latency;
      solves += 2. * res[i].am_iterations;
    }
    qsort(lat, (size_t)n, sizeof(double), cmp_dbl);
    stats->wall = wall;
    stats->solves = solves;
    stats->solves_per_s = wall > 0. ? solves / wall : 0.;
    stats->p50 = pct(lat, n, .50);
    stats->p99 = pct(lat, n, .99);
    stats->mean_latency = n ? sum / n : 0.;
    stats->max_latency = n ? lat[n - 1] : 0.;
    stats->completed = n;
    stats->rejected = rejected;
    stats->max_resident = max_res;
    stats->peak_bytes = peak;
    if (e) vfg_engine_stats(e, &stats->ticks, &stats->replays);
  }
  for (int i = 0; i < nreq; i++)
    if (ctx[i]) vfh_destroy(ctx[i]);
  if (e) vfg_engine_destroy(e);
  free(order); free(ctx); free(need); free(opt); free(lat);
  return rc;
}
