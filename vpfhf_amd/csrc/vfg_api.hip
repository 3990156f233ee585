// vfg_api.hip -- the vfg C-ABI: handles, field management, operator entry
// points, diagnostics and device memory helpers.  Solvers live in
// vfg_solve.hip.  Nothing here computes on the host: every numerical result
// comes from the HIP kernels; the library refuses to run without a GPU.
#include <hip/hip_runtime.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>
#include <string>
#include <vector>
#include "vfg_handle.h"

using namespace vfg;

static thread_local char g_err[512];

int vfg_fail(vfg_handle *h, int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  if (h) h->err = buf;
  snprintf(g_err, sizeof g_err, "%s", buf);
  return code;
}

#define HIPCHK(h, call)                                                                        \
  do {                                                                                         \
    hipError_t e_ = (call);                                                                    \
    if (e_ != hipSuccess) return vfg_fail(h, VFG_E_HIP, "%s: %s", #call, hipGetErrorString(e_)); \
  } while (0)

int vfg_alloc(vfg_handle *h, void **p, size_t bytes) {
  if (*p) return VFG_OK;
  hipError_t e = hipMalloc(p, bytes ? bytes : 8);
  if (e != hipSuccess) {
    *p = nullptr;
    return vfg_fail(h, VFG_E_NOMEM, "hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
  }
  h->allocs.push_back(*p);
  h->bytes += bytes;
  return VFG_OK;
}

static void bcmask_from(const vfg_bc *bc, int dof, BcMask *m) {
  for (int c = 0; c < dof; c++) {
    uint8_t f = 0, v = 0;
    for (int q = 0; q < 6; q++)
      if (bc[c].face[q] != VFG_NONE) f |= (uint8_t)(1u << q);
    for (int q = 0; q < 8; q++)
      if (bc[c].vertex[q] != VFG_NONE) v |= (uint8_t)(1u << q);
    m->face[c] = f;
    m->vert[c] = v;
  }
}

static void set_props_(vfg_handle *h, const vfg_props *pr) {
  DevProblem &P = h->P;
  h->props = *pr;
  P.E = pr->E; P.nu = pr->nu; P.alpha = pr->alpha; P.beta = pr->beta; P.Gc = pr->Gc;
  P.eta = pr->eta; P.eps = pr->epsilon; P.irrevtol = pr->irrevtol;
  P.atnum = pr->atnum; P.hasCrackPressure = pr->hasCrackPressure;
  P.unilateral = pr->unilateral;
  // VFMatPropGet (VFCommon.c:583-586) for E = 1
  P.lam1 = 1. * pr->nu / (1. + pr->nu) / (1. - 2. * pr->nu);
  P.mu1 = 1. / (1. + pr->nu) * .5;
  P.atcv = pr->atnum == 1 ? 2. / 3. : .5;  // VFCommon.c:482-492
}

static int check_props(const vfg_props *pr) {
  if (!pr) return 0;
  if (pr->atnum != 1 && pr->atnum != 2) return 0;
  if (pr->unilateral != VFG_UNILATERAL_NONE && pr->unilateral != VFG_UNILATERAL_NOCOMPRESSION) return 0;
  if (!(pr->nu > -1. && pr->nu < .5)) return 0;
  if (!(pr->epsilon > 0.)) return 0;
  return 1;
}

// copy an input array to device memory (dst) from host or device
int vfg_in(vfg_handle *h, void *dst, const void *src, size_t bytes, int loc) {
  HIPCHK(h, hipMemcpyAsync(dst, src, bytes, loc == VFG_HOST ? hipMemcpyHostToDevice : hipMemcpyDeviceToDevice,
                           h->stream));
  return VFG_OK;
}


// --------------------------------------------------------------- operators
// Host-pointer calls go through two staging vectors of 3*N_n doubles.
int vfg_scratch(vfg_handle *h, int64_t ndoubles) {
  if (h->stage_n >= ndoubles) return VFG_OK;
  if (h->stage) {
    hipFree(h->stage);
    for (auto &p : h->allocs)
      if (p == h->stage) p = nullptr;
    h->stage = nullptr;
    h->bytes -= 8 * (size_t)h->stage_n;
    h->stage_n = 0;
  }
  int rc = vfg_alloc(h, (void **)&h->stage, 8 * ndoubles);
  if (!rc) h->stage_n = ndoubles;
  return rc;
}


int vfg_u_prepare(vfg_handle *h) {
  DevProblem &P = h->P;
  if (h->u_form != VFG_U_FORM_STORED || P.unilateral || P.uq) return VFG_OK;
  int rc = vfg_alloc(h, (void **)&h->uq_buf, 8 * (size_t)uq_doubles(P.G));
  if (rc) return rc;
  launch_u_qsetup(P, h->uq_buf, h->stream);
  P.uq = h->uq_buf;
  return VFG_OK;
}

extern "C" {

const char *vfg_last_error(const vfg_handle *h) { return h ? h->err.c_str() : g_err; }

int vfg_create(const vfg_grid *g, const vfg_props *pr, vfg_handle **out) {
  if (!out) return vfg_fail(nullptr, VFG_E_ARG, "vfg_create: out is NULL");
  *out = nullptr;
  if (!g || g->nx < 2 || g->ny < 2 || g->nz < 2 || !(g->lx > 0) || !(g->ly > 0) || !(g->lz > 0))
    return vfg_fail(nullptr, VFG_E_ARG, "vfg_create: grid needs >= 2 nodes and positive lengths per direction");
  if (!check_props(pr)) return vfg_fail(nullptr, VFG_E_ARG, "vfg_create: invalid properties (atnum, nu, epsilon)");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
    return vfg_fail(nullptr, VFG_E_NODEV, "vfg_create: no HIP device available");
  vfg_handle *h = new vfg_handle();
  DevProblem &P = h->P;
  memset(&P, 0, sizeof P);
  h->grid = *g;
  P.G.nx = g->nx; P.G.ny = g->ny; P.G.nz = g->nz;
  P.G.nex = g->nx - 1; P.G.ney = g->ny - 1; P.G.nez = g->nz - 1;
  P.G.hx = g->lx / (g->nx - 1.); P.G.hy = g->ly / (g->ny - 1.); P.G.hz = g->lz / (g->nz - 1.);
  P.nn = (int64_t)g->nx * g->ny * g->nz;
  P.ne = (int64_t)P.G.nex * P.G.ney * P.G.nez;
  set_props_(h, pr);
  if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) {
    delete h;
    return vfg_fail(nullptr, VFG_E_HIP, "vfg_create: hipStreamCreate failed");
  }
  int rc;
  if ((rc = vfg_alloc(h, (void **)&P.V, 8 * P.nn)) || (rc = vfg_alloc(h, (void **)&P.VIrrev, 8 * P.nn)) ||
      (rc = vfg_alloc(h, (void **)&P.dth, 8 * P.nn)) || (rc = vfg_alloc(h, (void **)&P.p, 8 * P.nn)) ||
      (rc = vfg_alloc(h, (void **)&P.V32, 4 * P.nn)) || (rc = vfg_alloc(h, (void **)&P.zero3, 24 * P.nn)) ||
      (rc = vfg_alloc(h, (void **)&P.vfix, P.nn)) || (rc = vfg_alloc(h, (void **)&P.bV, 8 * P.nn)) ||
      (rc = vfg_alloc(h, (void **)&h->part, 8 * 4 * 65536)) || (rc = vfg_alloc(h, (void **)&h->scal, 8 * 64)) ||
      (rc = vfg_alloc(h, (void **)&P.cmask, 4 * P.ne))) {
    snprintf(g_err, sizeof g_err, "%s", h->err.c_str());
    vfg_destroy(h);
    return rc;
  }
  launch_fill(P.V, 1., P.nn, h->stream);
  launch_fill(P.VIrrev, 1., P.nn, h->stream);
  launch_fill(P.dth, 0., P.nn, h->stream);
  launch_fill(P.p, 0., P.nn, h->stream);
  launch_fill(P.zero3, 0., 3 * P.nn, h->stream);
  launch_copy_d2f(P.V, P.V32, P.nn, h->stream);
  hipMemsetAsync(P.cmask, 0, 4 * P.ne, h->stream);  // the state U = 0: every point tensile
  for (int c = 0; c < 3; c++) memset(&P.hbcU[c], 0, sizeof(vfg_bc));
  memset(&P.hbcV, 0, sizeof(vfg_bc));
  launch_v_fixmask(P, h->stream);
  if (hipStreamSynchronize(h->stream) != hipSuccess || hipGetLastError() != hipSuccess) {
    vfg_destroy(h);
    return vfg_fail(nullptr, VFG_E_HIP, "vfg_create: kernel launch failed (is the library built for this GPU?)");
  }
  *out = h;
  return VFG_OK;
}

int vfg_destroy(vfg_handle *h) {
  if (!h) return VFG_OK;
  if (h->stream) hipStreamSynchronize(h->stream);
  for (void *p : h->allocs) hipFree(p);
  if (h->stream) hipStreamDestroy(h->stream);
  delete h->mg;
  delete h;
  return VFG_OK;
}

int vfg_set_props(vfg_handle *h, const vfg_props *pr) {
  if (!h) return vfg_fail(nullptr, VFG_E_ARG, "NULL handle");
  if (!check_props(pr)) return vfg_fail(h, VFG_E_ARG, "vfg_set_props: invalid properties");
  set_props_(h, pr);
  vfg_u_invalidate(h);
  launch_v_fixmask(h->P, h->stream);
  if (h->mg) h->mg->setups = 0;  // the multigrid smoother bounds depend on nu
  return VFG_OK;
}

int vfg_set_bc(vfg_handle *h, const vfg_bc u_bc[3], const vfg_bc *v_bc) {
  if (!h) return vfg_fail(nullptr, VFG_E_ARG, "NULL handle");
  for (int c = 0; c < 3 && u_bc; c++)
    for (int q = 0; q < 14; q++) {
      int t = q < 6 ? u_bc[c].face[q] : u_bc[c].vertex[q - 6];
      if (t < VFG_NONE || t > VFG_FIXED) return vfg_fail(h, VFG_E_ARG, "vfg_set_bc: bad BC type %d", t);
    }
  if (u_bc) {
    for (int c = 0; c < 3; c++) h->P.hbcU[c] = u_bc[c];
    bcmask_from(u_bc, 3, &h->P.bcU);
    if (h->mg) h->mg->setups = 0;  // and on the Dirichlet rows
  }
  if (v_bc) {
    h->P.hbcV = *v_bc;
    bcmask_from(v_bc, 1, &h->P.bcV);
  }
  launch_v_fixmask(h->P, h->stream);
  return VFG_OK;
}

int64_t vfg_num_nodes(const vfg_handle *h) { return h ? h->P.nn : -1; }
int64_t vfg_num_cells(const vfg_handle *h) { return h ? h->P.ne : -1; }

int vfg_synchronize(vfg_handle *h) {
  if (!h) return vfg_fail(nullptr, VFG_E_ARG, "NULL handle");
  HIPCHK(h, hipStreamSynchronize(h->stream));
  HIPCHK(h, hipGetLastError());
  return VFG_OK;
}

int vfg_set_fields(vfg_handle *h, const double *V, const double *VIrrev, const double *theta,
                   const double *thetaRef, const double *pressure, const double *E_elem, int loc) {
  if (!h) return vfg_fail(nullptr, VFG_E_ARG, "NULL handle");
  DevProblem &P = h->P;
  int rc;
  if ((theta == nullptr) != (thetaRef == nullptr))
    return vfg_fail(h, VFG_E_ARG, "vfg_set_fields: theta and thetaRef must be given together");
  if (V && (rc = vfg_in(h, P.V, V, 8 * P.nn, loc))) return rc;
  if (VIrrev && (rc = vfg_in(h, P.VIrrev, VIrrev, 8 * P.nn, loc))) return rc;
  if (pressure && (rc = vfg_in(h, P.p, pressure, 8 * P.nn, loc))) return rc;
  if (theta) {
    // dtheta = theta - thetaRef, the difference every reference kernel forms
    if ((rc = vfg_scratch(h, 2 * P.nn))) return rc;
    if ((rc = vfg_in(h, h->stage, theta, 8 * P.nn, loc))) return rc;
    if ((rc = vfg_in(h, h->stage + P.nn, thetaRef, 8 * P.nn, loc))) return rc;
    launch_sub(h->stage, h->stage + P.nn, P.dth, P.nn, h->stream);
  }
  if (E_elem) {
    if ((rc = vfg_alloc(h, (void **)&P.E_elem, 8 * P.ne))) return rc;
    if ((rc = vfg_in(h, P.E_elem, E_elem, 8 * P.ne, loc))) return rc;
  }
  if (V) launch_copy_d2f(P.V, P.V32, P.nn, h->stream);
  if (V || E_elem) vfg_u_invalidate(h);
  if (VIrrev) launch_v_fixmask(P, h->stream);
  if (loc == VFG_HOST) HIPCHK(h, hipStreamSynchronize(h->stream));
  return VFG_OK;
}

int vfg_get_fields(vfg_handle *h, double *V, double *VIrrev, int loc) {
  if (!h) return vfg_fail(nullptr, VFG_E_ARG, "NULL handle");
  hipMemcpyKind k = loc == VFG_HOST ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice;
  if (V) HIPCHK(h, hipMemcpyAsync(V, h->P.V, 8 * h->P.nn, k, h->stream));
  if (VIrrev) HIPCHK(h, hipMemcpyAsync(VIrrev, h->P.VIrrev, 8 * h->P.nn, k, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return VFG_OK;
}

struct IO {
  vfg_handle *h;
  int loc;
  const void *in;
  void *out;
  size_t inb, outb;
  void *din, *dout;
  int rc;
  IO(vfg_handle *h_, int loc_, const void *in_, size_t inb_, void *out_, size_t outb_)
      : h(h_), loc(loc_), in(in_), out(out_), inb(inb_), outb(outb_), din(nullptr), dout(nullptr), rc(0) {
    if (loc == VFG_DEVICE) {
      din = (void *)in;
      dout = out;
      return;
    }
    int64_t need = (int64_t)((inb + 7) / 8 + (outb + 7) / 8);
    if ((rc = vfg_scratch(h, need > 0 ? need : 1))) return;
    din = h->stage;
    dout = (char *)h->stage + ((inb + 7) / 8) * 8;
    if (in && inb) rc = vfg_in(h, din, in, inb, VFG_HOST);
  }
  int finish() {
    if (rc) return rc;
    if (loc == VFG_HOST && out && outb) {
      if (hipMemcpyAsync(out, dout, outb, hipMemcpyDeviceToHost, h->stream) != hipSuccess)
        return vfg_fail(h, VFG_E_HIP, "device-to-host copy failed");
    }
    if (hipStreamSynchronize(h->stream) != hipSuccess || hipGetLastError() != hipSuccess)
      return vfg_fail(h, VFG_E_HIP, "kernel execution failed");
    return VFG_OK;
  }
};

#define NEED_H(h) \
  if (!h) return vfg_fail(nullptr, VFG_E_ARG, "NULL handle")

int vfg_u_set_state(vfg_handle *h, const double *U, int loc) {
  NEED_H(h);
  if (!U) return vfg_fail(h, VFG_E_ARG, "vfg_u_set_state: NULL array");
  if (!h->P.unilateral) return VFG_OK;
  IO io(h, loc, U, 24 * h->P.nn, nullptr, 0);
  if (io.rc) return io.rc;
  launch_u_cmask(h->P, (const double *)io.din, h->stream);
  return io.finish();
}

int vfg_u_set_form(vfg_handle *h, int form) {
  NEED_H(h);
  if (form != VFG_U_FORM_RECOMPUTE && form != VFG_U_FORM_STORED)
    return vfg_fail(h, VFG_E_ARG, "vfg_u_set_form: unknown form %d", form);
  h->u_form = form;
  vfg_u_invalidate(h);
  return VFG_OK;
}

int vfg_u_apply(vfg_handle *h, const double *x, double *y, int loc) {
  NEED_H(h);
  if (!x || !y) return vfg_fail(h, VFG_E_ARG, "vfg_u_apply: NULL array");
  int prc = vfg_u_prepare(h);
  if (prc) return prc;
  IO io(h, loc, x, 24 * h->P.nn, y, 24 * h->P.nn);
  if (io.rc) return io.rc;
  launch_u_apply(h->P, (const double *)io.din, (double *)io.dout, h->stream);
  return io.finish();
}

int vfg_u_apply_f32(vfg_handle *h, const float *x, float *y, int loc) {
  NEED_H(h);
  if (!x || !y) return vfg_fail(h, VFG_E_ARG, "vfg_u_apply_f32: NULL array");
  IO io(h, loc, x, 12 * h->P.nn, y, 12 * h->P.nn);
  if (io.rc) return io.rc;
  launch_u_apply_f32(h->P, (const float *)io.din, (float *)io.dout, h->stream);
  return io.finish();
}

int vfg_u_diag(vfg_handle *h, double *d, int loc) {
  NEED_H(h);
  if (!d) return vfg_fail(h, VFG_E_ARG, "vfg_u_diag: NULL array");
  IO io(h, loc, nullptr, 0, d, 24 * h->P.nn);
  if (io.rc) return io.rc;
  launch_u_diag(h->P, (double *)io.dout, h->stream);
  return io.finish();
}

int vfg_u_rhs(vfg_handle *h, double *b, int loc) {
  NEED_H(h);
  if (!b) return vfg_fail(h, VFG_E_ARG, "vfg_u_rhs: NULL array");
  IO io(h, loc, nullptr, 0, b, 24 * h->P.nn);
  if (io.rc) return io.rc;
  launch_u_rhs(h->P, nullptr, nullptr, (double *)io.dout, 0, h->stream);
  return io.finish();
}

int vfg_u_residual(vfg_handle *h, const double *U, const double *BCU, double *F, int loc) {
  NEED_H(h);
  if (!U || !BCU || !F) return vfg_fail(h, VFG_E_ARG, "vfg_u_residual: NULL array");
  const int64_t n3 = 3 * h->P.nn;
  if (loc == VFG_DEVICE) {
    launch_u_rhs(h->P, U, BCU, F, 1, h->stream);
    return vfg_synchronize(h);
  }
  int rc;
  if ((rc = vfg_scratch(h, 3 * n3))) return rc;
  double *dU = h->stage, *dB = h->stage + n3, *dF = h->stage + 2 * n3;
  if ((rc = vfg_in(h, dU, U, 8 * n3, VFG_HOST)) || (rc = vfg_in(h, dB, BCU, 8 * n3, VFG_HOST))) return rc;
  launch_u_rhs(h->P, dU, dB, dF, 1, h->stream);
  HIPCHK(h, hipMemcpyAsync(F, dF, 8 * n3, hipMemcpyDeviceToHost, h->stream));
  return vfg_synchronize(h);
}

int vfg_v_set_u(vfg_handle *h, const double *U, int loc) {
  NEED_H(h);
  if (!U) return vfg_fail(h, VFG_E_ARG, "vfg_v_set_u: NULL U");
  DevProblem &P = h->P;
  int rc;
  if ((rc = vfg_alloc(h, (void **)&P.Vq, 8 * (size_t)vq_doubles(P.G)))) return rc;
  IO io(h, loc, U, 24 * P.nn, nullptr, 0);
  if (io.rc) return io.rc;
  launch_v_setup(P, (const double *)io.din, h->stream);
  h->v_ready = true;
  return io.finish();
}

#define NEED_V(h) \
  if (!h->v_ready) return vfg_fail(h, VFG_E_ARG, "V operator used before vfg_v_set_u")

int vfg_v_apply(vfg_handle *h, const double *x, double *y, int loc) {
  NEED_H(h);
  NEED_V(h);
  if (!x || !y) return vfg_fail(h, VFG_E_ARG, "vfg_v_apply: NULL array");
  IO io(h, loc, x, 8 * h->P.nn, y, 8 * h->P.nn);
  if (io.rc) return io.rc;
  launch_v_apply(h->P, (const double *)io.din, (double *)io.dout, 0, h->stream);
  return io.finish();
}

int vfg_v_diag(vfg_handle *h, double *d, int loc) {
  NEED_H(h);
  NEED_V(h);
  if (!d) return vfg_fail(h, VFG_E_ARG, "vfg_v_diag: NULL array");
  IO io(h, loc, nullptr, 0, d, 8 * h->P.nn);
  if (io.rc) return io.rc;
  launch_v_diag(h->P, (double *)io.dout, h->stream);
  return io.finish();
}

int vfg_v_rhs(vfg_handle *h, double *b, int loc) {
  NEED_H(h);
  NEED_V(h);
  if (!b) return vfg_fail(h, VFG_E_ARG, "vfg_v_rhs: NULL array");
  HIPCHK(h, hipMemcpyAsync(b, h->P.bV, 8 * h->P.nn,
                           loc == VFG_HOST ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice, h->stream));
  return vfg_synchronize(h);
}

int vfg_v_residual(vfg_handle *h, const double *V, double *F, int loc) {
  NEED_H(h);
  NEED_V(h);
  if (!V || !F) return vfg_fail(h, VFG_E_ARG, "vfg_v_residual: NULL array");
  IO io(h, loc, V, 8 * h->P.nn, F, 8 * h->P.nn);
  if (io.rc) return io.rc;
  launch_v_apply(h->P, (const double *)io.din, (double *)io.dout, 1, h->stream);
  return io.finish();
}

// ------------------------------------------------------------------ BC kit
int vfg_bc_mask(vfg_handle *h, int which, uint8_t *mask, int loc) {
  NEED_H(h);
  if (!mask || (which != 0 && which != 1)) return vfg_fail(h, VFG_E_ARG, "vfg_bc_mask: bad arguments");
  size_t nb = (which == 0 ? 3 : 1) * h->P.nn;
  IO io(h, loc, nullptr, 0, mask, nb);
  if (io.rc) return io.rc;
  launch_bc_mask(h->P, which, (uint8_t *)io.dout, h->stream);
  return io.finish();
}

int vfg_vec_set_from_bc(vfg_handle *h, int which, double *B, int loc) {
  NEED_H(h);
  if (!B || (which != 0 && which != 1)) return vfg_fail(h, VFG_E_ARG, "vfg_vec_set_from_bc: bad arguments");
  size_t nb = 8 * (which == 0 ? 3 : 1) * h->P.nn;
  IO io(h, loc, B, nb, nullptr, 0);
  if (io.rc) return io.rc;
  launch_vec_set_from_bc(h->P, which, (double *)io.din, h->stream);
  if (loc == VFG_HOST) HIPCHK(h, hipMemcpyAsync(B, io.din, nb, hipMemcpyDeviceToHost, h->stream));
  return io.finish();
}

int vfg_vec_apply_dirichlet(vfg_handle *h, int which, double *vec, const double *B, int loc) {
  NEED_H(h);
  if (!vec || !B || (which != 0 && which != 1)) return vfg_fail(h, VFG_E_ARG, "vfg_vec_apply_dirichlet: bad arguments");
  size_t nb = 8 * (which == 0 ? 3 : 1) * h->P.nn;
  if (loc == VFG_DEVICE) {
    launch_vec_apply_dirichlet(h->P, which, vec, B, h->stream);
    return vfg_synchronize(h);
  }
  int rc;
  if ((rc = vfg_scratch(h, (int64_t)(2 * nb / 8)))) return rc;
  // the same host array for vec and B keeps the reference's in-place semantics
  const bool alias = (const void *)vec == (const void *)B;
  double *dv = h->stage, *dB = alias ? h->stage : h->stage + nb / 8;
  if ((rc = vfg_in(h, dv, vec, nb, VFG_HOST)) || (!alias && (rc = vfg_in(h, dB, B, nb, VFG_HOST)))) return rc;
  launch_vec_apply_dirichlet(h->P, which, dv, dB, h->stream);
  HIPCHK(h, hipMemcpyAsync(vec, dv, nb, hipMemcpyDeviceToHost, h->stream));
  return vfg_synchronize(h);
}

// ------------------------------------------------------------- diagnostics
static int energies_(vfg_handle *h, const double *U, const double *V, double *cellfield, double *efield,
                     double out[4], int loc) {
  DevProblem &P = h->P;
  int rc;
  int nb = energy_blocks(P);
  if (nb > 65536) {
    if ((rc = vfg_alloc(h, (void **)&h->part_big, 8 * 4 * (size_t)nb))) return rc;
  }
  double *part = nb > 65536 ? h->part_big : h->part;
  const double *dU = U, *dV = V;
  double *dcell = cellfield, *deng = efield;
  if (loc == VFG_HOST) {
    if ((rc = vfg_scratch(h, 4 * P.nn + 4 * P.ne))) return rc;
    double *sU = h->stage, *sV = h->stage + 3 * P.nn;
    if ((rc = vfg_in(h, sU, U, 24 * P.nn, VFG_HOST)) || (rc = vfg_in(h, sV, V, 8 * P.nn, VFG_HOST))) return rc;
    dU = sU; dV = sV;
    dcell = cellfield ? h->stage + 4 * P.nn : nullptr;
    deng = efield ? h->stage + 4 * P.nn + P.ne : nullptr;
  }
  launch_energies(P, dU, dV, part, dcell, deng, h->scal, h->stream);
  HIPCHK(h, hipMemcpyAsync(out, h->scal, 32, hipMemcpyDeviceToHost, h->stream));
  if (loc == VFG_HOST && cellfield)
    HIPCHK(h, hipMemcpyAsync(cellfield, dcell, 8 * P.ne, hipMemcpyDeviceToHost, h->stream));
  if (loc == VFG_HOST && efield)
    HIPCHK(h, hipMemcpyAsync(efield, deng, 24 * P.ne, hipMemcpyDeviceToHost, h->stream));
  return vfg_synchronize(h);
}

int vfg_energies(vfg_handle *h, const double *U, const double *V, double out[4], int loc) {
  NEED_H(h);
  if (!U || !V || !out) return vfg_fail(h, VFG_E_ARG, "vfg_energies: NULL array");
  double r[4];
  int rc = energies_(h, U, V, nullptr, nullptr, r, loc);
  if (rc) return rc;
  out[0] = r[0];
  out[1] = 0.;
  out[2] = r[1];
  out[3] = r[2];
  return VFG_OK;
}

int vfg_crack_volume(vfg_handle *h, const double *U, const double *V, double *vol, double *cellfield, int loc) {
  NEED_H(h);
  if (!U || !V || !vol) return vfg_fail(h, VFG_E_ARG, "vfg_crack_volume: NULL array");
  double r[4];
  int rc = energies_(h, U, V, cellfield, nullptr, r, loc);
  if (rc) return rc;
  *vol = r[3];
  return VFG_OK;
}

int vfg_energy_fields(vfg_handle *h, const double *U, const double *V, double *cell3, int loc) {
  NEED_H(h);
  if (!U || !V || !cell3) return vfg_fail(h, VFG_E_ARG, "vfg_energy_fields: NULL array");
  double r[4];
  return energies_(h, U, V, nullptr, cell3, r, loc);
}

int vfg_cell_to_node(vfg_handle *h, const double *cell, double *node, int dof, int loc) {
  NEED_H(h);
  if (!cell || !node) return vfg_fail(h, VFG_E_ARG, "vfg_cell_to_node: NULL array");
  if (dof < 1 || dof > 6) return vfg_fail(h, VFG_E_ARG, "vfg_cell_to_node: dof must be 1..6");
  IO io(h, loc, cell, 8 * (size_t)dof * h->P.ne, node, 8 * (size_t)dof * h->P.nn);
  if (io.rc) return io.rc;
  launch_cell_to_node(h->P, h->grid.lx, h->grid.ly, h->grid.lz, (const double *)io.din, (double *)io.dout, dof,
                      h->stream);
  return io.finish();
}

int vfg_penny_build_v(vfg_handle *h, double *V, const double center[3], double r, double theta_deg,
                      double phi_deg, double thickness, int loc) {
  NEED_H(h);
  if (!V || !center || r < 0.) return vfg_fail(h, VFG_E_ARG, "vfg_penny_build_v: bad arguments");
  IO io(h, loc, nullptr, 0, V, 8 * h->P.nn);
  if (io.rc) return io.rc;
  launch_penny(h->P, (double *)io.dout, center, r, theta_deg, phi_deg, thickness, h->stream);
  return io.finish();
}

// ------------------------------------------------------ device vector helpers
int64_t vfg_device_bytes(const vfg_handle *h) { return h ? (int64_t)h->bytes : -1; }

int vfg_vec_copy(vfg_handle *h, const double *x, double *y, int64_t n) {
  NEED_H(h);
  if (!x || !y || n < 0) return vfg_fail(h, VFG_E_ARG, "vfg_vec_copy: bad arguments");
  HIPCHK(h, hipMemcpyAsync(y, x, 8 * n, hipMemcpyDeviceToDevice, h->stream));
  return VFG_OK;
}
int vfg_vec_pointwise_min(vfg_handle *h, const double *a, const double *b, double *out, int64_t n) {
  NEED_H(h);
  if (!a || !b || !out || n < 0) return vfg_fail(h, VFG_E_ARG, "vfg_vec_pointwise_min: bad arguments");
  launch_vmin(a, b, out, n, h->stream);
  return VFG_OK;
}
int vfg_vec_maxabs_diff(vfg_handle *h, const double *a, const double *b, int64_t n, double *out) {
  NEED_H(h);
  if (!a || !b || !out || n < 0) return vfg_fail(h, VFG_E_ARG, "vfg_vec_maxabs_diff: bad arguments");
  launch_maxabs_diff(a, b, n, h->part, h->scal + 32, h->stream);
  HIPCHK(h, hipMemcpyAsync(out, h->scal + 32, 8, hipMemcpyDeviceToHost, h->stream));
  return vfg_synchronize(h);
}
int vfg_vec_scale(vfg_handle *h, double *x, double v, int64_t n) {
  NEED_H(h);
  if (!x || n < 0) return vfg_fail(h, VFG_E_ARG, "vfg_vec_scale: bad arguments");
  launch_scale(x, v, n, h->stream);
  return VFG_OK;
}
int vfg_vec_set(vfg_handle *h, double *x, double v, int64_t n) {
  NEED_H(h);
  if (!x || n < 0) return vfg_fail(h, VFG_E_ARG, "vfg_vec_set: bad arguments");
  launch_fill(x, v, n, h->stream);
  return VFG_OK;
}

// ------------------------------------------------------------ memory helpers
int vfg_device_alloc(size_t bytes, void **ptr) {
  if (!ptr) return vfg_fail(nullptr, VFG_E_ARG, "vfg_device_alloc: NULL ptr");
  hipError_t e = hipMalloc(ptr, bytes ? bytes : 8);
  if (e != hipSuccess) return vfg_fail(nullptr, VFG_E_NOMEM, "hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
  return VFG_OK;
}
int vfg_device_free(void *ptr) {
  if (ptr && hipFree(ptr) != hipSuccess) return vfg_fail(nullptr, VFG_E_HIP, "hipFree failed");
  return VFG_OK;
}
int vfg_memcpy(void *dst, const void *src, size_t bytes, int kind) {
  hipMemcpyKind k = kind == 0 ? hipMemcpyHostToDevice : kind == 1 ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice;
  if (kind < 0 || kind > 2) return vfg_fail(nullptr, VFG_E_ARG, "vfg_memcpy: bad kind");
  hipError_t e = hipMemcpy(dst, src, bytes, k);
  if (e != hipSuccess) return vfg_fail(nullptr, VFG_E_HIP, "hipMemcpy: %s", hipGetErrorString(e));
  return VFG_OK;
}

int vfg_time_u_apply(vfg_handle *h, const double *const *x, double *const *y, int nbuf, int nrep, int fp32,
                     double *ms) {
  NEED_H(h);
  if (!x || !y || nbuf < 1 || nrep < 1 || !ms) return vfg_fail(h, VFG_E_ARG, "vfg_time_u_apply: bad arguments");
  if (!fp32) {
    int rc = vfg_u_prepare(h);
    if (rc) return rc;
  }
  hipEvent_t e0, e1;
  HIPCHK(h, hipEventCreate(&e0));
  HIPCHK(h, hipEventCreate(&e1));
  HIPCHK(h, hipEventRecord(e0, h->stream));
  for (int i = 0; i < nrep; i++) {
    if (fp32)
      launch_u_apply_f32(h->P, (const float *)x[i % nbuf], (float *)y[i % nbuf], h->stream);
    else
      launch_u_apply(h->P, x[i % nbuf], y[i % nbuf], h->stream);
  }
  HIPCHK(h, hipEventRecord(e1, h->stream));
  HIPCHK(h, hipEventSynchronize(e1));
  float t = 0.f;
  HIPCHK(h, hipEventElapsedTime(&t, e0, e1));
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  *ms = t;
  HIPCHK(h, hipGetLastError());
  return VFG_OK;
}

int vfg_time_v_apply(vfg_handle *h, const double *const *x, double *const *y, int nbuf, int nrep, double *ms) {
  NEED_H(h);
  if (!x || !y || nbuf < 1 || nrep < 1 || !ms) return vfg_fail(h, VFG_E_ARG, "vfg_time_v_apply: bad arguments");
  if (!h->v_ready) return vfg_fail(h, VFG_E_ARG, "vfg_time_v_apply: V operator used before vfg_v_set_u");
  hipEvent_t e0, e1;
  HIPCHK(h, hipEventCreate(&e0));
  HIPCHK(h, hipEventCreate(&e1));
  HIPCHK(h, hipEventRecord(e0, h->stream));
  for (int i = 0; i < nrep; i++) launch_v_apply(h->P, x[i % nbuf], y[i % nbuf], 0, h->stream);
  HIPCHK(h, hipEventRecord(e1, h->stream));
  HIPCHK(h, hipEventSynchronize(e1));
  float t = 0.f;
  HIPCHK(h, hipEventElapsedTime(&t, e0, e1));
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  *ms = t;
  HIPCHK(h, hipGetLastError());
  return VFG_OK;
}

}  // extern "C"
