// vfg_solve.hip -- GPU solvers behind vfg_u_solve / vfg_v_solve.
//
// u: VF_StepU (VFMech.c:1436-1451) runs one Newton step of SNESNEWTONLS on a
//    problem that is linear in U, with KSPCG on the Jacobian whose Dirichlet
//    rows are identity (VFCommon.c:1018-1048).  With U carrying its Dirichlet
//    values the Newton system is K_II d_I = -F_I, solved here by
//    preconditioned CG started from U (r = -F(U)), which is the same iterate
//    sequence without the separate correction vector.
// V: VF_StepV (VFMech.c:3057-3082) with SNESVINEWTONRSLS (VFCommon.c:1052-1086):
//    minimise the quadratic V-energy over lb <= V <= ub with Dirichlet and
//    (AT2) irreversible nodes held.  Reduced-space active-set Newton: the
//    active set holds nodes at a bound whose gradient points outward; the
//    Newton step solves the reduced system on the inactive set by PCG; the
//    projected step is accepted under an Armijo test.
// All vectors, scalars and reductions stay on the device; the host reads a
// few scalars every `check` iterations to test convergence.
#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>
#include <functional>
#include <utility>
#include "vfg_handle.h"

namespace vfg {

enum { S_RZ = 0, S_PQ, S_ALPHA, S_BETA, S_RR, S_RR0, S_FLAG, S_TMP0, S_TMP1, S_TMP2, S_TOL2, S_DONE, S_ITS };
constexpr int RB = 256;

__device__ __forceinline__ double bsum(double v) {
  __shared__ double sh[RB / 64];
  return block_sum<RB>(v, sh);
}

// z = dinv * r; p = z (init) ; partial sums of r.z and r.r
__global__ __launch_bounds__(RB) void k_pcg_init(const double *__restrict__ r, double *__restrict__ z,
                                                 double *__restrict__ p, const double *__restrict__ dinv,
                                                 int64_t n, double *part) {
  double rz = 0., rr = 0.;
  for (int64_t i = (int64_t)blockIdx.x * RB + threadIdx.x; i < n; i += (int64_t)gridDim.x * RB) {
    double ri = r[i], zi = dinv[i] * ri;
    z[i] = zi;
    p[i] = zi;
    rz += ri * zi;
    rr += ri * ri;
  }
  rz = bsum(rz);
  rr = bsum(rr);
  if (threadIdx.x == 0) { part[2 * blockIdx.x] = rz; part[2 * blockIdx.x + 1] = rr; }
}

// op 0: rz, rr, rr0 from init; op 1: alpha = rz / (p.q); op 2: beta
__global__ __launch_bounds__(RB) void k_pcg_scalar(const double *part, int nb, int op, double *sc) {
  double a = 0., b = 0.;
  const int w = op == 1 ? 1 : 2;
  for (int i = threadIdx.x; i < nb; i += RB) {
    a += part[(int64_t)i * w];
    if (w == 2) b += part[(int64_t)i * w + 1];
  }
  a = bsum(a);
  b = bsum(b);
  if (threadIdx.x != 0) return;
  if (op == 0) {
    sc[S_RZ] = a; sc[S_RR] = b; sc[S_RR0] = b; sc[S_FLAG] = 0.;
    sc[S_TOL2] = 0.; sc[S_DONE] = 0.; sc[S_ITS] = 0.;
  } else if (sc[S_DONE] != 0.) {
    return;  // converged: the rest of the iteration batch does nothing
  } else if (op == 1) {
    sc[S_PQ] = a;
    if (!(a > 0.) || !isfinite(a)) { sc[S_ALPHA] = 0.; sc[S_FLAG] = isfinite(a) ? 1. : 2.; }
    else sc[S_ALPHA] = sc[S_RZ] / a;
  } else {
    double rz = sc[S_RZ];
    sc[S_BETA] = rz != 0. ? a / rz : 0.;
    sc[S_RZ] = a;
    sc[S_RR] = b;
    sc[S_ITS] += 1.;
    if (!isfinite(a) || !isfinite(b)) sc[S_FLAG] = 2.;
    else if (sc[S_TOL2] > 0. && b <= sc[S_TOL2]) sc[S_DONE] = 1.;  // only solvers whose kernels all check S_DONE set S_TOL2
  }
}

__global__ __launch_bounds__(RB) void k_dotp(const double *__restrict__ a, const double *__restrict__ b, int64_t n,
                                             double *part) {
  double s = 0.;
  for (int64_t i = (int64_t)blockIdx.x * RB + threadIdx.x; i < n; i += (int64_t)gridDim.x * RB) s += a[i] * b[i];
  s = bsum(s);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

__global__ __launch_bounds__(RB) void k_pcg_xr(double *__restrict__ x, double *__restrict__ r,
                                               double *__restrict__ z, const double *__restrict__ p,
                                               const double *__restrict__ q, const double *__restrict__ dinv,
                                               int64_t n, const double *sc, double *part) {
  const double al = sc[S_ALPHA];
  double rz = 0., rr = 0.;
  for (int64_t i = (int64_t)blockIdx.x * RB + threadIdx.x; i < n; i += (int64_t)gridDim.x * RB) {
    x[i] += al * p[i];
    double ri = r[i] - al * q[i];
    r[i] = ri;
    double zi = dinv[i] * ri;
    z[i] = zi;
    rz += ri * zi;
    rr += ri * ri;
  }
  rz = bsum(rz);
  rr = bsum(rr);
  if (threadIdx.x == 0) { part[2 * blockIdx.x] = rz; part[2 * blockIdx.x + 1] = rr; }
}

__global__ __launch_bounds__(RB) void k_pcg_p(double *__restrict__ p, const double *__restrict__ z, int64_t n,
                                              const double *sc) {
  const double be = sc[S_BETA];
  for (int64_t i = (int64_t)blockIdx.x * RB + threadIdx.x; i < n; i += (int64_t)gridDim.x * RB)
    p[i] = z[i] + be * p[i];
}

// p = z + beta p with z the FP32 preconditioner output (exact in FP64)
__global__ __launch_bounds__(RB) void k_pcg_p32(double *__restrict__ p, const float *__restrict__ z32, int64_t n,
                                                const double *sc) {
  const double be = sc[S_BETA];
  for (int64_t i = (int64_t)blockIdx.x * RB + threadIdx.x; i < n; i += (int64_t)gridDim.x * RB)
    p[i] = (double)z32[i] + be * p[i];
}

__global__ void k_recip(const double *d, double *dinv, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dinv[i] = d[i] != 0. ? 1. / d[i] : 1.;
}
__global__ void k_neg(const double *a, double *b, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = -a[i];
}

static unsigned nb_(int64_t n) { return (unsigned)((n + 255) / 256); }

typedef std::function<void(const double *, double *)> ApplyFn;
// apply with the x.y partials written to h->part: (x, y) -> their number, 0
// when the operator has no such kernel (nothing launched then)
typedef std::function<int(const double *, double *)> DotApplyFn;
// fused apply: (p_old, p, q) -> number of p.q partials written to h->part
typedef std::function<int(const double *, double *, double *)> FusedFn;
// CG update over a subset of the nodes: (p, q) -> number of (r.z, r.r)
// partial pairs written to h->part
typedef std::function<int(const double *, const double *)> XrFn;


// Preconditioned CG with the direction update and the p.q product fused into
// the operator apply (launch_v_apply_cg_list): per iteration
// one apply, one vector update of x, r, z and two scalar steps.  p alternates
// between two arrays.  Same convergence test and reporting as pcg().
static int pcg_fused(vfg_handle *h, int64_t n, const FusedFn &A, const double *dinv, double *x, double *r,
                     double *z, double *p, double *p2, double *q, double rtol, double atol, int maxit, int check,
                     vfg_solve_info *info, const XrFn *xr = nullptr, double rn0_known = -1., int first = 0);

// Preconditioned CG on the device.  On entry x is the initial guess and r the
// initial residual b - A x.  Converged when ||r|| <= max(rtol*||r0||, atol).
static int pcg(vfg_handle *h, int64_t n, const ApplyFn &A, const double *dinv, double *x, double *r, double *z,
               double *p, double *q, double rtol, double atol, int maxit, int check, vfg_solve_info *info,
               const DotApplyFn *Ad = nullptr) {
  hipStream_t s = h->stream;
  double *sc = h->scal, *part = h->part;
  const int nb = red_blocks(n);
  hipLaunchKernelGGL(k_pcg_init, dim3(nb), dim3(RB), 0, s, r, z, p, dinv, n, part);
  hipLaunchKernelGGL(k_pcg_scalar, dim3(1), dim3(RB), 0, s, (const double *)part, nb, 0, sc);
  double hs[8];
  if (hipMemcpyAsync(hs, sc, 64, hipMemcpyDeviceToHost, s) != hipSuccess || hipStreamSynchronize(s) != hipSuccess)
    return vfg_fail(h, VFG_E_HIP, "pcg: initial reduction failed");
  const double rn0 = sqrt(hs[S_RR0]);
  double rn = rn0;
  int it = 0, status = VFG_OK;
  const double tol = fmax(rtol * rn0, atol);
  if (!isfinite(rn0)) status = VFG_E_NAN;
  double rate = -1.;  // observed residual reduction per iteration
  while (status == VFG_OK && rn > tol && it < maxit) {
    int todo = check < maxit - it ? check : maxit - it;
    // run only the iterations the observed rate says are left (at most
    // check, at least check / 4: the Jacobi-PCG residual is not monotone, and
    // one lucky drop must not shrink the batches to single iterations, each
    // ending in a host synchronisation)
    if (rate > 0. && rate < 1.) {
      const double need = ceil(log(tol / rn) / log(rate));
      const int floor_ = check / 4 > 1 ? check / 4 : 1;
      if (need >= 1. && need < todo) todo = (int)need > floor_ ? (int)need : (floor_ < todo ? floor_ : todo);
    }
    const double rn_before = rn;
    for (int k = 0; k < todo; k++) {
      int na = Ad ? (*Ad)(p, q) : 0;  // q = A p with the p.q partials
      if (!na) {
        A(p, q);
        hipLaunchKernelGGL(k_dotp, dim3(nb), dim3(RB), 0, s, (const double *)p, (const double *)q, n, part);
        na = nb;
      }
      hipLaunchKernelGGL(k_pcg_scalar, dim3(1), dim3(RB), 0, s, (const double *)part, na, 1, sc);
      hipLaunchKernelGGL(k_pcg_xr, dim3(nb), dim3(RB), 0, s, x, r, z, (const double *)p, (const double *)q, dinv, n,
                         (const double *)sc, part);
      hipLaunchKernelGGL(k_pcg_scalar, dim3(1), dim3(RB), 0, s, (const double *)part, nb, 2, sc);
      hipLaunchKernelGGL(k_pcg_p, dim3(nb), dim3(RB), 0, s, p, (const double *)z, n, (const double *)sc);
    }
    it += todo;
    if (hipMemcpyAsync(hs, sc, 64, hipMemcpyDeviceToHost, s) != hipSuccess || hipStreamSynchronize(s) != hipSuccess)
      return vfg_fail(h, VFG_E_HIP, "pcg: iteration failed");
    rn = sqrt(hs[S_RR]);
    if (rn_before > 0. && rn > 0.) rate = pow(rn / rn_before, 1. / todo);
    if (hs[S_FLAG] == 2. || !isfinite(rn)) status = VFG_E_NAN;
    else if (hs[S_FLAG] == 1.) status = rn <= tol ? VFG_OK : VFG_E_BREAKDOWN;
    if (hs[S_FLAG] != 0.) break;
  }
  if (status == VFG_OK && rn > tol) status = VFG_E_MAXIT;
  if (info) {
    info->iterations += it;
    info->relres = rn0 > 0. ? rn / rn0 : 0.;
    info->status = status;
  }
  return status;
}

__global__ void k_set_scalar(double *sc, int i, double v) { sc[i] = v; }
// two device counts next to a scalar, so that one copy reads all three
__global__ void k_pack_counts(const int *a, const int *b, double *out) {
  out[0] = (double)*a;
  out[1] = (double)*b;
}

static int pcg_fused(vfg_handle *h, int64_t n, const FusedFn &A, const double *dinv, double *x, double *r,
                     double *z, double *p, double *p2, double *q, double rtol, double atol, int maxit, int check,
                     vfg_solve_info *info, const XrFn *xr, double rn0_known, int first) {
  hipStream_t s = h->stream;
  double *sc = h->scal, *part = h->part;
  const int nb = red_blocks(n);
  hipLaunchKernelGGL(k_pcg_init, dim3(nb), dim3(RB), 0, s, r, z, p, dinv, n, part);
  hipLaunchKernelGGL(k_pcg_scalar, dim3(1), dim3(RB), 0, s, (const double *)part, nb, 0, sc);
  hipLaunchKernelGGL(k_set_scalar, dim3(1), dim3(1), 0, s, sc, (int)S_BETA, 0.);  // first direction p = z
  double hs[13];
  // ||r0|| from the caller when it has just read it back; else read here
  if (rn0_known < 0. &&
      (hipMemcpyAsync(hs, sc, 64, hipMemcpyDeviceToHost, s) != hipSuccess || hipStreamSynchronize(s) != hipSuccess))
    return vfg_fail(h, VFG_E_HIP, "pcg: initial reduction failed");
  const double rn0 = rn0_known < 0. ? sqrt(hs[S_RR0]) : rn0_known;
  double rn = rn0;
  int it = 0, status = VFG_OK;
  const double tol = fmax(rtol * rn0, atol);
  if (!isfinite(rn0)) status = VFG_E_NAN;
  // with xr (the listed kernels), the iteration that reaches tol marks the
  // solve done on the device and the rest of its batch exits at once
  if (xr) hipLaunchKernelGGL(k_set_scalar, dim3(1), dim3(1), 0, s, sc, (int)S_TOL2, tol * tol);
  double *pa = p, *pb = p2;  // previous / current direction
  bool first_batch = true;
  while (status == VFG_OK && rn > tol && it < maxit) {
    int todo = check < maxit - it ? check : maxit - it;
    // first > 0: the caller's estimate of the iterations needed (launches
    // after the device-side stop cost a launch each, for nothing)
    if (first_batch && first > 0 && first < todo) todo = first;
    first_batch = false;
    for (int k = 0; k < todo; k++) {
      const int na = A(pa, pb, q);  // pb = z + beta pa, q = A pb, part = p.q per workgroup
      hipLaunchKernelGGL(k_pcg_scalar, dim3(1), dim3(RB), 0, s, (const double *)part, na, 1, sc);
      int nx = nb;
      if (xr) nx = (*xr)(pb, q);
      else
        hipLaunchKernelGGL(k_pcg_xr, dim3(nb), dim3(RB), 0, s, x, r, z, (const double *)pb, (const double *)q, dinv,
                           n, (const double *)sc, part);
      hipLaunchKernelGGL(k_pcg_scalar, dim3(1), dim3(RB), 0, s, (const double *)part, nx, 2, sc);
      double *t = pa;
      pa = pb;
      pb = t;
    }
    if (hipMemcpyAsync(hs, sc, 8 * 13, hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
      return vfg_fail(h, VFG_E_HIP, "pcg: iteration failed");
    it = (int)hs[S_ITS];  // iterations that ran (with xr the batch stops at convergence)
    rn = sqrt(hs[S_RR]);
    if (hs[S_FLAG] == 2. || !isfinite(rn)) status = VFG_E_NAN;
    else if (hs[S_FLAG] == 1.) status = rn <= tol ? VFG_OK : VFG_E_BREAKDOWN;
    if (hs[S_FLAG] != 0. || hs[S_DONE] != 0.) break;
  }
  if (status == VFG_OK && rn > tol) status = VFG_E_MAXIT;
  if (info) {
    info->iterations += it;
    info->relres = rn0 > 0. ? rn / rn0 : 0.;
    info->status = status;
  }
  return status;
}

static int work(vfg_handle *h, int nvec, int64_t n) {
  for (int i = 0; i < nvec; i++) {
    if (h->w[i]) continue;
    int rc = vfg_alloc(h, (void **)&h->w[i], 8 * (size_t)n);
    if (rc) return rc;
  }
  return VFG_OK;
}
}  // namespace vfg
int vfg_work(vfg_handle *h, int n64, int n32) {
  const int64_t n = 3 * h->P.nn;
  int rc = vfg::work(h, n64, n);
  for (int i = 0; i < n32 && !rc; i++) rc = vfg_alloc(h, (void **)&h->wf[i], 4 * (size_t)n);
  return rc;
}
namespace vfg {

// ---------------------------------------------------------------- V helpers
__global__ void k_v_clamp(double *V, const uint8_t *fix, const double *ub, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || fix[i]) return;
  double hi = ub ? ub[i] : 1., v = V[i];
  V[i] = v < 0. ? 0. : v > hi ? hi : v;
}

// active set + projected-gradient partial norm; r = -g on the inactive set
__global__ __launch_bounds__(RB) void k_v_active(const double *V, const double *g, const uint8_t *fix,
                                                 const double *ub, uint8_t *act, double *r, int64_t n,
                                                 double *part) {
  double s = 0.;
  for (int64_t i = (int64_t)blockIdx.x * RB + threadIdx.x; i < n; i += (int64_t)gridDim.x * RB) {
    double hi = ub ? ub[i] : 1., v = V[i], gi = g[i];
    bool a = fix[i] || (v <= 0. && gi > 0.) || (v >= hi && gi < 0.);
    act[i] = a ? 1 : 0;
    r[i] = a ? 0. : -gi;
    if (!a) s += gi * gi;
  }
  s = bsum(s);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// Vn = P(V + lam d) on free nodes; D = Vn - V
__global__ void k_v_step(const double *V, const double *d, const uint8_t *act, const uint8_t *fix,
                         const double *ub, double lam, double *Vn, double *D, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double v = V[i], x = v;
  if (!act[i]) {
    double hi = ub ? ub[i] : 1.;
    x = v + lam * d[i];
    if (!fix[i]) x = x < 0. ? 0. : x > hi ? hi : x;
  }
  Vn[i] = x;
  D[i] = x - v;
}

__global__ void k_zero_fixed(double *g, const uint8_t *fix, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && fix[i]) g[i] = 0.;
}

static unsigned nbv(int64_t n) { return (unsigned)((n + 255) / 256); }
void launch_v_clamp(double *V, const uint8_t *fix, const double *ub, int64_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_v_clamp, dim3(nbv(n)), dim3(256), 0, s, V, fix, ub, n);
}
void launch_v_active(const double *V, const double *g, const uint8_t *fix, const double *ub, uint8_t *act,
                     double *r, int64_t n, double *part, hipStream_t s) {
  hipLaunchKernelGGL(k_v_active, dim3(red_blocks(n)), dim3(RB), 0, s, V, g, fix, ub, act, r, n, part);
}
void launch_v_step(const double *V, const double *d, const uint8_t *act, const uint8_t *fix, const double *ub,
                   double lam, double *Vn, double *D, int64_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_v_step, dim3(nbv(n)), dim3(256), 0, s, V, d, act, fix, ub, lam, Vn, D, n);
}
void launch_zero_fixed(double *g, const uint8_t *fix, int64_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_zero_fixed, dim3(nbv(n)), dim3(256), 0, s, g, fix, n);
}

static double dot_host(vfg_handle *h, const double *a, const double *b, int64_t n, int *rc) {
  launch_dot(a, b, n, h->part, h->scal + 16, h->stream);
  double v = 0.;
  if (hipMemcpyAsync(&v, h->scal + 16, 8, hipMemcpyDeviceToHost, h->stream) != hipSuccess ||
      hipStreamSynchronize(h->stream) != hipSuccess)
    *rc = vfg_fail(h, VFG_E_HIP, "dot failed");
  return v;
}


// ----------------------------------------------------------- FP32 inner PCG
// Vectors in FP32, all inner products and CG scalars in FP64.
__global__ __launch_bounds__(RB) void k_pcg32_init(const float *__restrict__ r, const float *__restrict__ z,
                                                   float *__restrict__ p, int64_t n, double *part) {
  double rz = 0., rr = 0.;
  for (int64_t i = (int64_t)blockIdx.x * RB + threadIdx.x; i < n; i += (int64_t)gridDim.x * RB) {
    double ri = r[i], zi = z[i];
    p[i] = z[i];
    rz += ri * zi;
    rr += ri * ri;
  }
  rz = bsum(rz);
  rr = bsum(rr);
  if (threadIdx.x == 0) { part[2 * blockIdx.x] = rz; part[2 * blockIdx.x + 1] = rr; }
}
__global__ __launch_bounds__(RB) void k_dotp32(const float *__restrict__ a, const float *__restrict__ b, int64_t n,
                                               double *part) {
  double s = 0.;
  for (int64_t i = (int64_t)blockIdx.x * RB + threadIdx.x; i < n; i += (int64_t)gridDim.x * RB)
    s += (double)a[i] * (double)b[i];
  s = bsum(s);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}
// x += alpha p ; r -= alpha q   (z is refreshed by the preconditioner)
__global__ __launch_bounds__(RB) void k_pcg32_xr(float *__restrict__ x, float *__restrict__ r,
                                                 const float *__restrict__ p, const float *__restrict__ q,
                                                 int64_t n, const double *sc) {
  const double al = sc[S_ALPHA];
  for (int64_t i = (int64_t)blockIdx.x * RB + threadIdx.x; i < n; i += (int64_t)gridDim.x * RB) {
    x[i] = (float)((double)x[i] + al * p[i]);
    r[i] = (float)((double)r[i] - al * q[i]);
  }
}
// partial sums of r.z and r.r
__global__ __launch_bounds__(RB) void k_rz32(const float *__restrict__ r, const float *__restrict__ z, int64_t n,
                                             double *part) {
  double rz = 0., rr = 0.;
  for (int64_t i = (int64_t)blockIdx.x * RB + threadIdx.x; i < n; i += (int64_t)gridDim.x * RB) {
    double ri = r[i];
    rz += ri * z[i];
    rr += ri * ri;
  }
  rz = bsum(rz);
  rr = bsum(rr);
  if (threadIdx.x == 0) { part[2 * blockIdx.x] = rz; part[2 * blockIdx.x + 1] = rr; }
}
__global__ __launch_bounds__(RB) void k_pcg32_p(float *__restrict__ p, const float *__restrict__ z, int64_t n,
                                                const double *sc) {
  const double be = sc[S_BETA];
  for (int64_t i = (int64_t)blockIdx.x * RB + threadIdx.x; i < n; i += (int64_t)gridDim.x * RB)
    p[i] = (float)((double)z[i] + be * p[i]);
}
__global__ void k_jacobi32(const float *__restrict__ dinv, const float *__restrict__ r, float *__restrict__ z,
                           int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) z[i] = dinv[i] * r[i];
}
// Chebyshev step: res = r - Az ; d = c1 d + c2 D^-1 res ; z += d
__global__ void k_cheb32(const float *__restrict__ r, const float *__restrict__ Az, const float *__restrict__ dinv,
                         float *__restrict__ d, float *__restrict__ z, float c1, float c2, int first, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float res = first ? r[i] : r[i] - Az[i];
  float di = (first ? 0.f : c1 * d[i]) + c2 * dinv[i] * res;
  d[i] = di;
  z[i] = first ? di : z[i] + di;
}
// scale r64 by 1/s into FP32 ; x64 += s * d32
__global__ void k_r64to32(const double *r, float *r32, double inv, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) r32[i] = (float)(r[i] * inv);
}
__global__ void k_fill32(float *x, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = 0.f;
}
__global__ void k_axpy3264(double *x, const float *d, double s, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] += s * (double)d[i];
}

struct Cheb {
  int degree;
  double lmin, lmax;
};

static int mg_apply(vfg_handle *h, const float *r, float *z, const double *r64 = nullptr, double *part = nullptr);

// z = M^-1 r  (Jacobi, Chebyshev(degree) polynomial in D^-1 A, or, with
// degree < 0, one multigrid V-cycle)
static void precond32(vfg_handle *h, const std::function<void(const float *, float *)> &A, const float *dinv,
                      const float *r, float *z, float *d, float *t, int64_t n, const Cheb &cb) {
  hipStream_t s = h->stream;
  if (cb.degree < 0) {
    mg_apply(h, r, z);
    return;
  }
  if (cb.degree <= 1) {
    hipLaunchKernelGGL(k_jacobi32, dim3(nb_(n)), dim3(256), 0, s, dinv, r, z, n);
    return;
  }
  const double th = .5 * (cb.lmax + cb.lmin), de = .5 * (cb.lmax - cb.lmin);
  const double s1 = th / de;
  double rho_old = 1. / s1;
  hipLaunchKernelGGL(k_cheb32, dim3(nb_(n)), dim3(256), 0, s, r, (const float *)nullptr, dinv, d, z, 0.f,
                     (float)(1. / th), 1, n);
  for (int k = 1; k < cb.degree; k++) {
    A(z, t);
    double rho = 1. / (2. * s1 - rho_old);
    hipLaunchKernelGGL(k_cheb32, dim3(nb_(n)), dim3(256), 0, s, r, (const float *)t, dinv, d, z,
                       (float)(rho * rho_old), (float)(2. * rho / de), 0, n);
    rho_old = rho;
  }
}

// Largest eigenvalue of the Lanczos tridiagonal built from CG coefficients
// (T_kk = 1/a_k + b_{k-1}/a_{k-1}, T_k,k+1 = sqrt(b_k)/a_k), by bisection with
// Sturm counts; an estimate of lambda_max(M^-1 A) from below.
static double lanczos_lmax(const double *ab, int m) {
  if (m < 1) return 0.;
  double d[64], e[64], hi = 0.;
  for (int k = 0; k < m; k++) {
    double a = ab[2 * k], ap = k ? ab[2 * (k - 1)] : 0., bp = k ? ab[2 * (k - 1) + 1] : 0.;
    d[k] = (a > 0 ? 1. / a : 0.) + (k && ap > 0 ? bp / ap : 0.);
    e[k] = (a > 0 ? sqrt(fmax(ab[2 * k + 1], 0.)) / a : 0.);
  }
  for (int k = 0; k < m; k++) hi = fmax(hi, d[k] + (k ? e[k - 1] : 0.) + (k < m - 1 ? e[k] : 0.));
  double lo = 0.;
  for (int it = 0; it < 100; it++) {
    double x = .5 * (lo + hi), q = 1.;
    int cnt = 0;  // eigenvalues below x
    for (int k = 0; k < m; k++) {
      q = d[k] - x - (k ? e[k - 1] * e[k - 1] / q : 0.);
      if (q == 0.) q = 1e-300;
      if (q < 0.) cnt++;
    }
    if (cnt >= m) hi = x; else lo = x;
  }
  return hi;
}

// Inner PCG in FP32 on A32 d = r (d0 = 0); returns iterations; lanczos
// estimates of the extreme eigenvalues of M^-1 A are accumulated from the CG
// coefficients when est != nullptr (used to size the Chebyshev interval).
// alpha, beta and the breakdown flag of CG iteration k, kept for the host
__global__ void k_ab_save(const double *sc, double *hist, int k) {
  hist[3 * k] = sc[S_ALPHA];
  hist[3 * k + 1] = sc[S_BETA];
  hist[3 * k + 2] = sc[S_FLAG];
}

static int pcg32(vfg_handle *h, int64_t n, const std::function<void(const float *, float *)> &A, const float *dinv,
                 float *x, float *r, float *z, float *p, float *q, float *w1, float *w2, const Cheb &cb,
                 double rtol, int maxit, double *lmax_est, int *status) {
  hipStream_t s = h->stream;
  double *sc = h->scal, *part = h->part;
  const int nb = red_blocks(n);
  precond32(h, A, dinv, r, z, w1, w2, n, cb);
  hipLaunchKernelGGL(k_pcg32_init, dim3(nb), dim3(RB), 0, s, (const float *)r, (const float *)z, p, n, part);
  hipLaunchKernelGGL(k_pcg_scalar, dim3(1), dim3(RB), 0, s, (const double *)part, nb, 0, sc);
  double hs[8];
  if (hipMemcpyAsync(hs, sc, 64, hipMemcpyDeviceToHost, s) != hipSuccess || hipStreamSynchronize(s) != hipSuccess) {
    *status = VFG_E_HIP;
    return 0;
  }
  const double rn0 = sqrt(hs[S_RR0]);
  double rn = rn0, ab[128];
  int it = 0, n_ab = 0;
  *status = VFG_OK;
  if (lmax_est) {
    // eigenvalue estimate: a fixed number of steps (at most 64), the
    // coefficients recorded on the device and read once
    if (!h->abh && vfg_alloc(h, (void **)&h->abh, 8 * 3 * 64)) {
      *status = VFG_E_NOMEM;
      return 0;
    }
    const int m = maxit < 64 ? maxit : 64;
    for (int k = 0; k < m; k++) {
      A(p, q);
      hipLaunchKernelGGL(k_dotp32, dim3(nb), dim3(RB), 0, s, (const float *)p, (const float *)q, n, part);
      hipLaunchKernelGGL(k_pcg_scalar, dim3(1), dim3(RB), 0, s, (const double *)part, nb, 1, sc);
      hipLaunchKernelGGL(k_pcg32_xr, dim3(nb), dim3(RB), 0, s, x, r, (const float *)p, (const float *)q, n,
                         (const double *)sc);
      precond32(h, A, dinv, r, z, w1, w2, n, cb);
      hipLaunchKernelGGL(k_rz32, dim3(nb), dim3(RB), 0, s, (const float *)r, (const float *)z, n, part);
      hipLaunchKernelGGL(k_pcg_scalar, dim3(1), dim3(RB), 0, s, (const double *)part, nb, 2, sc);
      hipLaunchKernelGGL(k_pcg32_p, dim3(nb), dim3(RB), 0, s, p, (const float *)z, n, (const double *)sc);
      hipLaunchKernelGGL(k_ab_save, dim3(1), dim3(1), 0, s, (const double *)sc, h->abh, k);
    }
    double hist[3 * 64];
    if (hipMemcpyAsync(hist, h->abh, 8 * 3 * m, hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess) {
      *status = VFG_E_HIP;
      return 0;
    }
    for (int k = 0; k < m; k++) {
      if (hist[3 * k + 2] == 2.) { *status = VFG_E_NAN; break; }
      if (hist[3 * k + 2] != 0.) break;  // breakdown: the steps before it stand
      ab[2 * n_ab] = hist[3 * k];
      ab[2 * n_ab + 1] = hist[3 * k + 1];
      n_ab++;
    }
    *lmax_est = lanczos_lmax(ab, n_ab);
    return m;
  }
  const int check = 10;
  double rate = -1.;  // observed residual reduction per iteration
  while (rn > rtol * rn0 && it < maxit) {
    int todo = check < maxit - it ? check : maxit - it;
    // run only the iterations the observed rate says are left (at most check,
    // at least check / 4, as in pcg)
    if (rate > 0. && rate < 1.) {
      const double need = ceil(log(rtol * rn0 / rn) / log(rate));
      const int floor_ = check / 4 > 1 ? check / 4 : 1;
      if (need >= 1. && need < todo) todo = (int)need > floor_ ? (int)need : (floor_ < todo ? floor_ : todo);
    }
    const double rn_before = rn;
    for (int k = 0; k < todo; k++) {
      A(p, q);
      hipLaunchKernelGGL(k_dotp32, dim3(nb), dim3(RB), 0, s, (const float *)p, (const float *)q, n, part);
      hipLaunchKernelGGL(k_pcg_scalar, dim3(1), dim3(RB), 0, s, (const double *)part, nb, 1, sc);
      hipLaunchKernelGGL(k_pcg32_xr, dim3(nb), dim3(RB), 0, s, x, r, (const float *)p, (const float *)q, n,
                         (const double *)sc);
      precond32(h, A, dinv, r, z, w1, w2, n, cb);
      hipLaunchKernelGGL(k_rz32, dim3(nb), dim3(RB), 0, s, (const float *)r, (const float *)z, n, part);
      hipLaunchKernelGGL(k_pcg_scalar, dim3(1), dim3(RB), 0, s, (const double *)part, nb, 2, sc);
      hipLaunchKernelGGL(k_pcg32_p, dim3(nb), dim3(RB), 0, s, p, (const float *)z, n, (const double *)sc);
    }
    it += todo;
    if (hipMemcpyAsync(hs, sc, 64, hipMemcpyDeviceToHost, s) != hipSuccess || hipStreamSynchronize(s) != hipSuccess) {
      *status = VFG_E_HIP;
      return it;
    }
    rn = sqrt(hs[S_RR]);
    if (rn_before > 0. && rn > 0.) rate = pow(rn / rn_before, 1. / todo);
    if (hs[S_FLAG] == 2. || !isfinite(rn)) { *status = VFG_E_NAN; break; }
    if (hs[S_FLAG] == 1.) { *status = rn <= rtol * rn0 ? VFG_OK : VFG_E_BREAKDOWN; break; }
  }
  return it;
}


// ---------------------------------------------------------------- multigrid
// V-cycle: Chebyshev(MG_SMOOTH)-Jacobi pre- and post-smoothing on
// [lmax/11, lmax] (lmax = 1.1 x the Lanczos estimate of D^-1 A), restriction
// of the residual, recursion, prolongation; the coarsest level gets
// Chebyshev(MG_COARSE_DEG) on [lmax/30, lmax].  All in FP32 with the level
// operators of vfg_mg.hip.  Each level has ceil(n/2) elements per direction
// where n > MG_COARSEST, else n; coarsening stops when no direction changes
// (the coarsest level has at most 8 elements per direction, or the fine
// level's count: at 128^3 five levels down to 8^3).
int mg_build(vfg_handle *h) {
  if (h->mg) return VFG_OK;
  MgHier *m = new MgHier();
  int rc = VFG_OK;
  MgLevel &l0 = m->lv[0];
  l0.P = h->P;
  l0.P.uq = nullptr;  // level 0 runs the FP32 apply only
  l0.n = 3 * h->P.nn;
  rc = vfg_alloc(h, (void **)&l0.dinv, 4 * (size_t)l0.n);
  if (!rc) rc = vfg_alloc(h, (void **)&l0.y, 4 * (size_t)l0.n);
  if (!rc) rc = vfg_alloc(h, (void **)&l0.d, 4 * (size_t)l0.n);
  int L = 1;
  Grid G = h->P.G;
  auto halve = [](int n) { return n > MG_COARSEST ? (n + 1) / 2 : n; };
  while (!rc && L < MgHier::MAXL) {
    Grid Gc;
    Gc.nex = halve(G.nex); Gc.ney = halve(G.ney); Gc.nez = halve(G.nez);
    if (Gc.nex == G.nex && Gc.ney == G.ney && Gc.nez == G.nez) break;
    Gc.nx = Gc.nex + 1; Gc.ny = Gc.ney + 1; Gc.nz = Gc.nez + 1;
    Gc.hx = G.hx * G.nex / Gc.nex; Gc.hy = G.hy * G.ney / Gc.ney; Gc.hz = G.hz * G.nez / Gc.nez;
    MgLevel &lc = m->lv[L];
    DevProblem &Pc = lc.P;
    memset(&Pc, 0, sizeof Pc);
    Pc.G = Gc;
    Pc.nn = (int64_t)Gc.nx * Gc.ny * Gc.nz;
    Pc.ne = (int64_t)Gc.nex * Gc.ney * Gc.nez;
    lc.n = 3 * Pc.nn;
    if ((rc = vfg_alloc(h, (void **)&Pc.E_elem, 8 * (size_t)Pc.ne)) ||
        (rc = vfg_alloc(h, (void **)&Pc.V, 8 * (size_t)Pc.nn)) ||
        (rc = vfg_alloc(h, (void **)&Pc.V32, 4 * (size_t)Pc.nn)))
      break;
    float **fv[5] = {&lc.dinv, &lc.b, &lc.x, &lc.y, &lc.d};
    for (int i = 0; i < 5 && !rc; i++) rc = vfg_alloc(h, (void **)fv[i], 4 * (size_t)lc.n);
    if (rc) break;
    launch_fill(Pc.V, 1., Pc.nn, h->stream);
    launch_mg_fill32(Pc.V32, 1.f, Pc.nn, h->stream);
    G = Gc;
    L++;
  }
  if (!rc && L > 1) rc = vfg_alloc(h, (void **)&m->dtmp, 8 * (size_t)m->lv[1].n);
  if (rc) {
    delete m;
    return rc;
  }
  m->L = L;
  h->mg = m;
  return VFG_OK;
}

// Refreshes the hierarchy from the current problem (V, E, properties, BCs):
// coarse coefficients, FP32 inverse diagonals and the smoother intervals.
// dinv64: the problem's FP64 inverse diagonal.
int mg_setup(vfg_handle *h, const double *dinv64) {
  int rc = mg_build(h);
  if (rc) return rc;
  MgHier &m = *h->mg;
  hipStream_t s = h->stream;
  const int64_t nf = 3 * h->P.nn;
  for (int i = 0; i < 7; i++)
    if (!h->wf[i] && (rc = vfg_alloc(h, (void **)&h->wf[i], 4 * (size_t)nf))) return rc;
  if (m.L > 1 && (rc = work(h, 8, nf))) return rc;  // w[7]: per-element coefficients of the problem
  m.lv[0].P = h->P;
  launch_copy_d2f(dinv64, m.lv[0].dinv, nf, s);
  for (int l = 1; l < m.L; l++) {
    MgLevel &lc = m.lv[l], &lf = m.lv[l - 1];
    DevProblem &Pc = lc.P;
    Pc.E = 1.; Pc.nu = h->P.nu; Pc.lam1 = h->P.lam1; Pc.mu1 = h->P.mu1; Pc.eta = 0.;
    Pc.bcU = h->P.bcU;
    for (int c = 0; c < 3; c++) Pc.hbcU[c] = h->P.hbcU[c];
    if (l == 1) {
      launch_mg_elem_coef(h->P, h->w[7], s);
      launch_mg_coarsen(h->P.G, h->w[7], Pc.G, Pc.E_elem, s);
    } else {
      launch_mg_coarsen(lf.P.G, lf.P.E_elem, Pc.G, Pc.E_elem, s);
    }
    launch_u_diag(Pc, m.dtmp, s);
    launch_mg_recip32(m.dtmp, lc.dinv, lc.n, s);
  }
  // lambda_max(D^-1 A) per level from 10 Jacobi-PCG steps (Lanczos).  The
  // top of the Jacobi-scaled spectrum is set by the element geometry and the
  // Poisson ratio, not by V or E (measured: 2.55, 2.50, 2.30, 1.79 on the
  // levels of a cracked 32^3 problem against 2.5 for a uniform one), so the
  // estimates of the first set-up are reused and refreshed every 16th.
  if (m.setups++ % 16 != 0) return VFG_OK;
  float **w = h->wf;
  for (int l = 0; l < m.L; l++) {
    MgLevel &lv = m.lv[l];
    launch_mg_seed(w[1], lv.n, s);
    launch_mg_fill32(w[0], 0.f, lv.n, s);
    std::function<void(const float *, float *)> A = [&](const float *in, float *out) {
      launch_u_apply_f32(lv.P, in, out, s);
    };
    double lmax = 0.;
    int st = 0;
    pcg32(h, lv.n, A, lv.dinv, w[0], w[1], w[2], w[3], w[4], w[5], w[6], Cheb{1, 0., 0.}, 0., 10, &lmax, &st);
    if (st == VFG_E_HIP || st == VFG_E_NAN || !(lmax > 0.))
      return vfg_fail(h, st ? st : VFG_E_NAN, "multigrid set-up: eigenvalue estimate failed on level %d", l);
    lv.lmax = 1.1 * lmax;
  }
  return VFG_OK;
}

// Chebyshev coefficients of mg_cheb_entry's step i on [lo, hi]
static void cheb_step_coefs(int k, double lo, double hi, float *c1, float *c2) {
  const double th = .5 * (hi + lo), de = .5 * (hi - lo), sg = th / de;
  double rho = 1. / sg;
  for (int i = 0; i < k; i++) {
    c1[i] = 0.f;
    c2[i] = (float)(1. / th);
    if (i > 0) {
      const double rn = 1. / (2. * sg - rho);
      c1[i] = (float)(rn * rho);
      c2[i] = (float)(2. * rn / de);
      rho = rn;
    }
  }
}

static void mg_smooth(vfg_handle *h, MgLevel &L, int k, double lo, double hi, bool xzero) {
  hipStream_t s = h->stream;
  float c1[16], c2[16];
  cheb_step_coefs(k, lo, hi, c1, c2);
  for (int i = 0; i < k; i++) {
    const float *y = nullptr;
    if (i > 0 || !xzero) {
      launch_u_apply_f32(L.P, L.x, L.y, s);
      y = L.y;
    }
    launch_mg_cheb(L.b, y, L.dinv, L.d, L.x, c1[i], c2[i], i == 0, i == 0 && xzero, L.n, s);
  }
}

// r64/part (level 0 only): the post-smoothing also reduces r64 . x into part;
// returns the number of partials (0: not fused, the caller reduces)
static int mg_vcycle(vfg_handle *h, int l, const double *r64 = nullptr, double *part = nullptr) {
  MgHier &m = *h->mg;
  MgLevel &L = m.lv[l];
  hipStream_t s = h->stream;
  if (l == m.L - 1) {
    // a coarse level of at most 8^3 elements: the whole solve in one launch
    if (l > 0 && mg_coarse_fits(L.P)) {
      MgChebCoefs cc;
      static_assert(MG_COARSE_DEG <= 8, "MgChebCoefs holds 8 steps");
      cc.k = MG_COARSE_DEG;
      cheb_step_coefs(MG_COARSE_DEG, L.lmax / 30., L.lmax, cc.c1, cc.c2);
      launch_mg_coarse_cheb(L.P, L.b, L.dinv, L.x, cc, s);
      return 0;
    }
    mg_smooth(h, L, MG_COARSE_DEG, L.lmax / 30., L.lmax, true);
    return 0;
  }
  MgLevel &C = m.lv[l + 1];
  static_assert(MG_SMOOTH == 1, "the fused smoothing kernels do one Chebyshev step");
  // one Chebyshev-Jacobi step from x = 0 fused with its apply, restriction of
  // b - A x, coarse correction, and one step fused with its apply: the values
  // of mg_smooth(1) + apply + restrict ... + mg_smooth(1) in three passes less.
  // The level iterates in d (A d in y) and leaves its result in x.
  float c1[1], c2[1];
  cheb_step_coefs(1, L.lmax / 11., L.lmax, c1, c2);
  launch_mg_presmooth(L.P, L.b, L.dinv, c2[0], L.d, L.y, s);
  // L.x is written only by the post-smoothing: it is the restriction's work space
  float *tmp = mg_restrict_tmp_floats(L.P.G, C.P.G) <= L.n ? L.x : nullptr;
  launch_mg_restrict(L.P.G, C.P.G, L.P.bcU, L.b, L.y, C.b, s, tmp);
  mg_vcycle(h, l + 1);
  launch_mg_prolong(L.P.G, C.P.G, L.P.bcU, C.x, L.d, s);
  return launch_mg_postsmooth(L.P, L.b, L.dinv, c2[0], L.d, L.x, r64, part, s);
}

// z = V-cycle(r) on the problem's grid (r: FP32, zero on Dirichlet rows);
// with r64, also the partials of r64 . z in part (returns their number, 0:
// none written)
static int mg_apply(vfg_handle *h, const float *r, float *z, const double *r64, double *part) {
  MgLevel &L = h->mg->lv[0];
  L.b = const_cast<float *>(r);
  L.x = z;
  return mg_vcycle(h, 0, r64, part);
}

// x += alpha p ; r -= alpha q ; r32 = r; with part, one partial of r.r per block
__global__ __launch_bounds__(RB) void k_mg_xr(double *__restrict__ x, double *__restrict__ r,
                                              const double *__restrict__ p, const double *__restrict__ q,
                                              float *__restrict__ r32, int64_t n, const double *sc, double *part) {
  const double al = sc[S_ALPHA];
  double rr = 0.;
  for (int64_t i = (int64_t)blockIdx.x * RB + threadIdx.x; i < n; i += (int64_t)gridDim.x * RB) {
    x[i] += al * p[i];
    const double ri = r[i] - al * q[i];
    r[i] = ri;
    r32[i] = (float)ri;
    rr += ri * ri;
  }
  if (part) {
    rr = bsum(rr);
    if (threadIdx.x == 0) part[blockIdx.x] = rr;
  }
}
// op 2 of k_pcg_scalar (beta, rz, rr) from separate partials: rz from prz[0, nrz), rr from prr[0, nrr)
__global__ __launch_bounds__(RB) void k_pcg_beta_split(const double *prz, int nrz, const double *prr, int nrr,
                                                       double *sc) {
  double a = 0., b = 0.;
  for (int i = threadIdx.x; i < nrz; i += RB) a += prz[i];
  for (int i = threadIdx.x; i < nrr; i += RB) b += prr[i];
  a = bsum(a);
  b = bsum(b);
  if (threadIdx.x != 0 || sc[S_DONE] != 0.) return;
  const double rz = sc[S_RZ];
  sc[S_BETA] = rz != 0. ? a / rz : 0.;
  sc[S_RZ] = a;
  sc[S_RR] = b;
  sc[S_ITS] += 1.;
  if (!isfinite(a) || !isfinite(b)) sc[S_FLAG] = 2.;
  else if (sc[S_TOL2] > 0. && b <= sc[S_TOL2]) sc[S_DONE] = 1.;
}
// partial sums of r.z and r.r with z = z32 (and p = z when p != nullptr)
__global__ __launch_bounds__(RB) void k_mg_rz(const double *__restrict__ r, const float *__restrict__ z32,
                                              double *__restrict__ p, int64_t n, double *part) {
  double rz = 0., rr = 0.;
  for (int64_t i = (int64_t)blockIdx.x * RB + threadIdx.x; i < n; i += (int64_t)gridDim.x * RB) {
    const double ri = r[i], zi = (double)z32[i];
    if (p) p[i] = zi;
    rz += ri * zi;
    rr += ri * ri;
  }
  rz = bsum(rz);
  rr = bsum(rr);
  if (threadIdx.x == 0) { part[2 * blockIdx.x] = rz; part[2 * blockIdx.x + 1] = rr; }
}

// FP64 PCG with the FP32 V-cycle as preconditioner; same contract as pcg().
static int pcg_mg(vfg_handle *h, int64_t n, const ApplyFn &A, double *x, double *r, double *p, double *q,
                  float *r32, float *z32, double rtol, double atol, int maxit, int check, vfg_solve_info *info,
                  const DotApplyFn *Ad = nullptr) {
  hipStream_t s = h->stream;
  double *sc = h->scal, *part = h->part;
  const int nb = red_blocks(n);
  launch_copy_d2f(r, r32, n, s);
  mg_apply(h, r32, z32);
  hipLaunchKernelGGL(k_mg_rz, dim3(nb), dim3(RB), 0, s, (const double *)r, (const float *)z32, p, n, part);
  hipLaunchKernelGGL(k_pcg_scalar, dim3(1), dim3(RB), 0, s, (const double *)part, nb, 0, sc);
  double hs[8];
  if (hipMemcpyAsync(hs, sc, 64, hipMemcpyDeviceToHost, s) != hipSuccess || hipStreamSynchronize(s) != hipSuccess)
    return vfg_fail(h, VFG_E_HIP, "pcg: initial reduction failed");
  const double rn0 = sqrt(hs[S_RR0]);
  double rn = rn0;
  int it = 0, status = VFG_OK;
  const double tol = fmax(rtol * rn0, atol);
  if (!isfinite(rn0)) status = VFG_E_NAN;
  double rate = -1.;  // observed residual reduction per iteration
  while (status == VFG_OK && rn > tol && it < maxit) {
    int todo = check < maxit - it ? check : maxit - it;
    // multigrid convergence is steady: run only the iterations the observed
    // rate says are left (at most check), so few run past convergence
    if (rate > 0. && rate < 1.) {
      const double need = ceil(log(tol / rn) / log(rate));
      if (need >= 1. && need < todo) todo = (int)need;
    }
    const double rn_before = rn;
    for (int k = 0; k < todo; k++) {
      int na = Ad ? (*Ad)(p, q) : 0;  // q = A p with the p.q partials
      if (!na) {
        A(p, q);
        hipLaunchKernelGGL(k_dotp, dim3(nb), dim3(RB), 0, s, (const double *)p, (const double *)q, n, part);
        na = nb;
      }
      hipLaunchKernelGGL(k_pcg_scalar, dim3(1), dim3(RB), 0, s, (const double *)part, na, 1, sc);
      // r.r partials from the update, r.z partials from the V-cycle's last
      // kernel (in two regions of part past the apply's partials)
      double *prr = part + 65536, *prz = part + 2 * 65536;
      hipLaunchKernelGGL(k_mg_xr, dim3(nb), dim3(RB), 0, s, x, r, (const double *)p, (const double *)q, r32, n,
                         (const double *)sc, prr);
      const int nrz = mg_apply(h, r32, z32, r, prz);  // at most 2 x 65536 partials
      if (nrz > 0) {
        hipLaunchKernelGGL(k_pcg_beta_split, dim3(1), dim3(RB), 0, s, (const double *)prz, nrz, (const double *)prr,
                           nb, sc);
      } else {
        hipLaunchKernelGGL(k_mg_rz, dim3(nb), dim3(RB), 0, s, (const double *)r, (const float *)z32,
                           (double *)nullptr, n, part);
        hipLaunchKernelGGL(k_pcg_scalar, dim3(1), dim3(RB), 0, s, (const double *)part, nb, 2, sc);
      }
      hipLaunchKernelGGL(k_pcg_p32, dim3(nb), dim3(RB), 0, s, p, (const float *)z32, n, (const double *)sc);
    }
    it += todo;
    if (hipMemcpyAsync(hs, sc, 64, hipMemcpyDeviceToHost, s) != hipSuccess || hipStreamSynchronize(s) != hipSuccess)
      return vfg_fail(h, VFG_E_HIP, "pcg: iteration failed");
    rn = sqrt(hs[S_RR]);
    if (rn_before > 0. && rn > 0.) rate = pow(rn / rn_before, 1. / todo);
    if (hs[S_FLAG] == 2. || !isfinite(rn)) status = VFG_E_NAN;
    else if (hs[S_FLAG] == 1.) status = rn <= tol ? VFG_OK : VFG_E_BREAKDOWN;
    if (hs[S_FLAG] != 0.) break;
  }
  if (status == VFG_OK && rn > tol) status = VFG_E_MAXIT;
  if (info) {
    info->iterations += it;
    info->relres = rn0 > 0. ? rn / rn0 : 0.;
    info->status = status;
  }
  return status;
}

// x = x0 + t d
__global__ void k_lin(double *__restrict__ x, const double *__restrict__ x0, const double *__restrict__ d, double t,
                      int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = x0[i] + t * d[i];
}

// UNILATERAL_NOCOMPRESSION: SNESNEWTONLS on F(U) = K(U) U - b(U)
// (VF_UResidual, VFMech.c:1456-1887) with the Jacobian K(U) of VF_UIJacobian
// (VFMech.c:1897-2029).  Each step solves J d = -F by PCG to
// min(0.1, tol / (2 ||F||)) relative (inexact Newton: the last step lands
// below tol), then halves d until ||F|| decreases (the line search's role).
// On entry r = -F(x); x carries its Dirichlet values.
static int u_solve_newton(vfg_handle *h, double *x, double *r, double *z, double *p, double *q, double *dinv,
                          const double *B, const vfg_solve_opts *o, vfg_solve_info *inf) {
  DevProblem &P = h->P;
  hipStream_t s = h->stream;
  const int64_t n = 3 * P.nn;
  int rc = work(h, 10, n);
  if (rc) return rc;
  double *d = h->w[8], *x0 = h->w[9];
  int e2 = 0;
  double rn = sqrt(dot_host(h, r, r, n, &e2));
  if (e2) return e2;
  const double rn0 = rn, tol = fmax(o->rtol * rn0, o->atol);
  ApplyFn A = [&](const double *in, double *out) { launch_u_apply(P, in, out, s); };
  int newton = 0, status = VFG_OK;
  for (; newton < 50; newton++) {
    if (!isfinite(rn)) { status = VFG_E_NAN; break; }
    if (rn <= tol || rn == 0.) break;
    launch_u_cmask(P, x, s);  // the Jacobian at the iterate
    launch_u_diag(P, q, s);
    hipLaunchKernelGGL(k_recip, dim3(nb_(n)), dim3(256), 0, s, (const double *)q, dinv, n);
    launch_fill(d, 0., n, s);
    const double inner = fmax(1e-14, fmin(0.1, .5 * tol / rn));
    vfg_solve_info li = {0, 0, 0., 0};
    int prc;
    if (o->pc == VFG_PC_MG) {
      // the compression state, and with it the operator, changes from step to
      // step: fresh smoother intervals every step (a stale lambda_max lets the
      // Chebyshev smoother amplify modes above its interval)
      if (h->mg) h->mg->setups = 0;
      if ((rc = mg_setup(h, dinv))) return rc;
      prc = pcg_mg(h, n, A, d, r, p, q, h->wf[0], h->wf[1], inner, 0., o->maxit, 4, &li);
    } else {
      prc = pcg(h, n, A, dinv, d, r, z, p, q, inner, 0., o->maxit, 20, &li);
    }
    inf->iterations += li.iterations;
    if (prc == VFG_E_HIP || prc == VFG_E_NAN) return prc;
    hipMemcpyAsync(x0, x, 8 * n, hipMemcpyDeviceToDevice, s);
    double t = 1.;
    for (int ls = 0; ls < 20; ls++) {
      hipLaunchKernelGGL(k_lin, dim3(nb_(n)), dim3(256), 0, s, x, (const double *)x0, (const double *)d, t, n);
      launch_u_rhs(P, x, B, r, 1, s);
      hipLaunchKernelGGL(k_neg, dim3(nb_(n)), dim3(256), 0, s, (const double *)r, r, n);
      const double rt = sqrt(dot_host(h, r, r, n, &e2));
      if (e2) return e2;
      if (rt < rn || ls == 19) {
        rn = rt;
        break;
      }
      t *= .5;
    }
  }
  inf->outer = newton;
  inf->relres = rn0 > 0. ? rn / rn0 : 0.;
  if (status == VFG_OK && rn > tol) status = VFG_E_MAXIT;
  return status;
}

// FP64 iterative refinement around the FP32 inner PCG (precision form):
// r64 = b - A64 x in FP64; solve A32 d = r64/|r64| in FP32 to inner_rtol;
// x += |r64| d in FP64; repeat until |r64| <= rtol |r64_0|.
// FP64 refinement around the FP32 inner PCG.  The residual of each step is
// -F(x) = bt - A x: one FP64 apply and a subtraction, with bt (the load vector
// with the Dirichlet targets, formed once) in bt and A x in y64.
static int u_solve_mixed(vfg_handle *h, double *x, double *r64, double *t64, const double *dinv64, const double *B,
                         double *bt, double *y64, const vfg_solve_opts *o, vfg_solve_info *inf) {
  DevProblem &P = h->P;
  hipStream_t s = h->stream;
  const int64_t n = 3 * P.nn;
  for (int i = 0; i < 7; i++)
    if (!h->wf[i]) {
      int rc = vfg_alloc(h, (void **)&h->wf[i], 4 * (size_t)n);
      if (rc) return rc;
    }
  float *d = h->wf[0], *r = h->wf[1], *z = h->wf[2], *p = h->wf[3], *q = h->wf[4], *w1 = h->wf[5], *w2 = h->wf[6];
  float *dinv = (float *)t64;  // FP32 reciprocal diagonal lives in the first half of t64
  launch_copy_d2f(dinv64, dinv, n, s);
  std::function<void(const float *, float *)> A32 = [&](const float *in, float *out) {
    launch_u_apply_f32(P, in, out, s);
  };
  Cheb cb = {1, 0., 0.};
  // inner tolerance: 1e-5 under the V-cycle (two outer steps reach 1e-10;
  // at 256^3 33 inner iterations against 42 with three steps at 1e-4, 37 at
  // 1e-6), 1e-4 under the Jacobi and Chebyshev preconditioners
  const double inner_rtol = o->pc == VFG_PC_MG ? 1e-5 : 1e-4;
  int rc = VFG_OK;
  if (o->pc == VFG_PC_MG) {
    if ((rc = mg_setup(h, dinv64))) return rc;
    cb.degree = -1;
  }
  launch_u_rhs(P, x, B, bt, 2, s);
  double rn0 = -1., rn_prev = 0.;
  for (int outer = 0; outer < 60; outer++) {
    int e2 = 0;
    double rn = sqrt(dot_host(h, r64, r64, n, &e2));
    if (e2) return e2;
    if (!isfinite(rn)) return VFG_E_NAN;
    if (rn0 < 0.) rn0 = rn;
    inf->relres = rn0 > 0. ? rn / rn0 : 0.;
    // Converged at the tolerance, or at the rounding floor of the FP64
    // residual: an outer step whose inner solve reduced its system by
    // inner_rtol no longer halves the true residual (a solve that starts
    // close to its solution, as in later alternate-minimisation iterations,
    // can ask for rtol * rn0 below that floor; the reference's SNES stops
    // there on its step-size test).
    if (rn <= fmax(o->rtol * rn0, o->atol) || rn == 0. || (outer > 0 && rn > 0.5 * rn_prev)) {
      inf->outer = outer;
      return VFG_OK;
    }
    rn_prev = rn;
    hipLaunchKernelGGL(k_r64to32, dim3(nb_(n)), dim3(256), 0, s, (const double *)r64, r, 1. / rn, n);
    hipLaunchKernelGGL(k_fill32, dim3(nb_(n)), dim3(256), 0, s, d, n);
    if (o->pc == VFG_PC_CHEBYSHEV && cb.degree == 1) {
      // lambda_max(D^-1 A) from 12 Jacobi-PCG steps on this residual, then restart
      double lmax = 0.;
      int st = 0;
      pcg32(h, n, A32, dinv, d, r, z, p, q, w1, w2, cb, 0., 12, &lmax, &st);
      inf->iterations += 12;
      if (st == VFG_E_HIP) return st;
      cb.degree = o->cheb_degree > 1 ? o->cheb_degree : 4;
      cb.lmax = 1.1 * lmax;
      cb.lmin = cb.lmax / 30.;
      hipLaunchKernelGGL(k_r64to32, dim3(nb_(n)), dim3(256), 0, s, (const double *)r64, r, 1. / rn, n);
      hipLaunchKernelGGL(k_fill32, dim3(nb_(n)), dim3(256), 0, s, d, n);
    }
    int st = 0;
    int its = pcg32(h, n, A32, dinv, d, r, z, p, q, w1, w2, cb, inner_rtol, o->maxit, nullptr, &st);
    inf->iterations += its * (cb.degree > 1 ? cb.degree : 1);  // a V-cycle step counts once
    if (st == VFG_E_HIP || st == VFG_E_NAN) return st;
    hipLaunchKernelGGL(k_axpy3264, dim3(nb_(n)), dim3(256), 0, s, x, (const float *)d, rn, n);
    launch_u_apply(P, x, y64, s);  // -F(x) = bt - A x in FP64
    launch_sub(bt, y64, r64, n, s);
    inf->outer = outer + 1;
  }
  return rc ? rc : VFG_E_MAXIT;
}

}  // namespace vfg

using namespace vfg;

extern "C" {

void vfg_solve_opts_default(vfg_solve_opts *o) {
  if (!o) return;
  o->rtol = 1e-10;
  o->atol = 0.;
  o->maxit = 100000;
  o->pc = VFG_PC_MG;  // the reference preconditions its u-CG with multigrid (PCHYPRE)
  o->precision = VFG_PREC_FP64;
  o->cheb_degree = 4;
}

// on the u-Dirichlet rows: dst = 0 (src null) or dst = src
__global__ void k_u_dirichlet_rows(Grid G, BcMask bc, const double *src, double *dst) {
  const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= (int64_t)G.nx * G.ny * G.nz) return;
  const int i = (int)(n % G.nx), j = (int)((n / G.nx) % G.ny), k = (int)(n / ((int64_t)G.nx * G.ny));
  for (int c = 0; c < 3; c++)
    if (on_dirichlet(G, bc.face[c], bc.vert[c], i, j, k)) dst[3 * n + c] = src ? src[3 * n + c] : 0.;
}

// z = M r with M the V-cycle of VFG_PC_MG for the current coefficients (the
// hierarchy's coefficients, diagonals and smoother bounds are refreshed
// first) on the free rows, and z = r on the Dirichlet rows (the inverse of the
// operator's identity rows): M = diag(I, M_free) is symmetric positive
// definite when the V-cycle is
int vfg_u_pc_apply(vfg_handle *h, const double *r, double *z, int loc) {
  if (!h) return vfg_fail(nullptr, VFG_E_ARG, "NULL handle");
  if (!r || !z) return vfg_fail(h, VFG_E_ARG, "vfg_u_pc_apply: NULL array");
  const int64_t n = 3 * h->P.nn;
  hipStream_t s = h->stream;
  int rc;
  if ((rc = vfg_work(h, 7, 2))) return rc;
  double *rd = h->w[1], *rin = h->w[2], *q = h->w[4], *dinv = h->w[5];
  if ((rc = vfg_in(h, rin, r, 8 * n, loc))) return rc;
  launch_u_diag(h->P, q, s);
  hipLaunchKernelGGL(k_recip, dim3(nb_(n)), dim3(256), 0, s, (const double *)q, dinv, n);
  if ((rc = mg_setup(h, dinv))) return rc;
  if (hipMemcpyAsync(rd, rin, 8 * n, hipMemcpyDeviceToDevice, s) != hipSuccess)
    return vfg_fail(h, VFG_E_HIP, "vfg_u_pc_apply: copy failed");
  const unsigned nbn = (unsigned)((h->P.nn + 255) / 256);
  hipLaunchKernelGGL(k_u_dirichlet_rows, dim3(nbn), dim3(256), 0, s, h->P.G, h->P.bcU, (const double *)nullptr, rd);
  launch_copy_d2f(rd, h->wf[0], n, s);
  mg_apply(h, h->wf[0], h->wf[1]);
  launch_copy_f2d(h->wf[1], rd, n, s);
  hipLaunchKernelGGL(k_u_dirichlet_rows, dim3(nbn), dim3(256), 0, s, h->P.G, h->P.bcU, (const double *)rin, rd);
  hipError_t e = hipMemcpyAsync(z, rd, 8 * n, loc == VFG_HOST ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice, s);
  if (e != hipSuccess || hipStreamSynchronize(s) != hipSuccess)
    return vfg_fail(h, VFG_E_HIP, "vfg_u_pc_apply: kernel execution failed");
  return VFG_OK;
}

int vfg_u_solve(vfg_handle *h, double *U, const double *BCU, const vfg_solve_opts *o, vfg_solve_info *info, int loc) {
  if (!h) return vfg_fail(nullptr, VFG_E_ARG, "NULL handle");
  if (!U || !BCU) return vfg_fail(h, VFG_E_ARG, "vfg_u_solve: NULL array");
  vfg_solve_opts d;
  vfg_solve_opts_default(&d);
  if (!o) o = &d;
  vfg_solve_info inf = {0, 0, 0., 0};
  const int64_t n = 3 * h->P.nn;
  hipStream_t s = h->stream;
  int rc;
  if ((rc = work(h, 7, n))) return rc;
  double *x = h->w[0], *r = h->w[1], *z = h->w[2], *p = h->w[3], *q = h->w[4], *dinv = h->w[5], *B = h->w[6];
  if ((rc = vfg_in(h, x, U, 8 * n, loc)) || (rc = vfg_in(h, B, BCU, 8 * n, loc))) return rc;
  if (o->precision != VFG_PREC_MIXED && (rc = vfg_u_prepare(h))) return rc;
  // VFTimeStepPrepare: U carries its Dirichlet values
  launch_vec_apply_dirichlet(h->P, 0, x, B, s);
  launch_u_rhs(h->P, x, B, r, 1, s);  // F(U)
  hipLaunchKernelGGL(k_neg, dim3(nb_(n)), dim3(256), 0, s, (const double *)r, r, n);
  launch_u_diag(h->P, q, s);
  hipLaunchKernelGGL(k_recip, dim3(nb_(n)), dim3(256), 0, s, (const double *)q, dinv, n);
  const DevProblem &P = h->P;
  if (P.unilateral) {
    rc = u_solve_newton(h, x, r, z, p, q, dinv, B, o, &inf);
  } else if (o->precision == VFG_PREC_MIXED) {
    rc = u_solve_mixed(h, x, r, q, dinv, B, z, p, o, &inf);
  } else if (o->pc == VFG_PC_CHEBYSHEV) {
    rc = vfg_fail(h, VFG_E_ARG, "vfg_u_solve: the Chebyshev preconditioner is provided for the mixed-precision solve");
  } else if (o->pc == VFG_PC_MG) {
    ApplyFn A = [&](const double *in, double *out) { launch_u_apply(P, in, out, s); };
    DotApplyFn Ad = [&](const double *in, double *out) { return launch_u_apply_dot(P, in, out, h->part, s); };
    rc = mg_setup(h, dinv);
    if (!rc) rc = pcg_mg(h, n, A, x, r, p, q, h->wf[0], h->wf[1], o->rtol, o->atol, o->maxit, 4, &inf, &Ad);
    inf.outer = 1;
  } else {
    // the u-operator is issue-bound at one wave per SIMD: fusing the direction
    // update into its staging, as the V-solve does, measured slower than the
    // separate vector kernels (282 us against 215 + 33 us at 128^3)
    ApplyFn A = [&](const double *in, double *out) { launch_u_apply(P, in, out, s); };
    DotApplyFn Ad = [&](const double *in, double *out) { return launch_u_apply_dot(P, in, out, h->part, s); };
    rc = pcg(h, n, A, dinv, x, r, z, p, q, o->rtol, o->atol, o->maxit, 20, &inf, &Ad);
    inf.outer = 1;
  }
  if (info) *info = inf;
  if (rc == VFG_E_HIP) return rc;
  hipError_t e = hipMemcpyAsync(U, x, 8 * n, loc == VFG_HOST ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice, s);
  if (e != hipSuccess || hipStreamSynchronize(s) != hipSuccess) return vfg_fail(h, VFG_E_HIP, "vfg_u_solve: copy-out failed");
  if (rc) return vfg_fail(h, rc, "vfg_u_solve: CG stopped with status %d after %d iterations (relres %.3e)", rc,
                          inf.iterations, inf.relres);
  return VFG_OK;
}

int vfg_v_solve(vfg_handle *h, const double *U, double *V, const double *ub, const vfg_solve_opts *o,
                vfg_solve_info *info, int loc) {
  if (!h) return vfg_fail(nullptr, VFG_E_ARG, "NULL handle");
  if (!U || !V) return vfg_fail(h, VFG_E_ARG, "vfg_v_solve: NULL array");
  vfg_solve_opts d;
  vfg_solve_opts_default(&d);
  if (!o) o = &d;
  vfg_solve_info inf = {0, 0, 0., 0};
  DevProblem &P = h->P;
  const int64_t n = P.nn;
  hipStream_t s = h->stream;
  int rc;
  if ((rc = work(h, 13, 3 * n))) return rc;
  if ((rc = vfg_alloc(h, (void **)&h->act, n))) return rc;
  const int vcap = v_chunk_cap(P.G);
  if ((rc = vfg_alloc(h, (void **)&h->vflag, 4 * (size_t)vcap)) ||
      (rc = vfg_alloc(h, (void **)&h->vlist, 4 * (size_t)(vcap + 2))) ||
      (rc = vfg_alloc(h, (void **)&h->vflag2, 4 * (size_t)vcap)) ||
      (rc = vfg_alloc(h, (void **)&h->vlist2, 4 * (size_t)(vcap + 2))))
    return rc;
  double *x = h->w[0], *r = h->w[1], *z = h->w[2], *p = h->w[3], *q = h->w[4], *dinv = h->w[5];
  double *Vd = h->w[6], *g = h->w[7], *Vn = h->w[8], *D = h->w[9], *KD = h->w[10], *ubd = nullptr;
  double *Ud = h->w[11];
  uint8_t *act = h->act;
  if ((rc = vfg_in(h, Ud, U, 24 * n, loc)) || (rc = vfg_in(h, Vd, V, 8 * n, loc))) return rc;
  if (ub) {
    ubd = KD + n;  // KD is 3n long: second third holds the bound
    if ((rc = vfg_in(h, ubd, ub, 8 * n, loc))) return rc;
  }
  // VF_VIJacobian refresh: coupling from U; Dirichlet values on V
  if ((rc = vfg_alloc(h, (void **)&P.Vq, 8 * (size_t)vq_doubles(P.G)))) return rc;
  launch_v_setup(P, Ud, s);
  h->v_ready = true;
  launch_fill(KD, 0., n, s);
  launch_v_fixmask(P, s);
  launch_vec_apply_dirichlet(P, 1, Vd, Vd, s);
  hipLaunchKernelGGL(k_v_clamp, dim3(nb_(n)), dim3(256), 0, s, Vd, (const uint8_t *)P.vfix, (const double *)ubd, n);
  // inexact Newton: the reduced system is solved to eta_k = 0.1 |g_k| / |g_0|
  // (between the final tolerance and 0.1), tight only as the iteration converges
  const double inner_floor = o->rtol < 1e-12 ? o->rtol : 1e-12;
  double pg0 = -1.;
  int it;
  const int nb = red_blocks(n);
  int status = VFG_OK;
  int nstep = -1;  // chunks around the last step's free nodes (-1: no step yet)
  int last_cg = 0;  // CG iterations of the last Newton step
  for (it = 0; it < (o->maxit < 200 ? o->maxit : 200); it++) {
    // F(V): free rows = K V - b.  A step changes V on the free nodes only, so
    // F changes only on the nodes that share an element with one: after the
    // first step F is re-formed on the chunks around them (the dilated list
    // of the step's free nodes) and kept elsewhere
    if (nstep < 0) launch_v_apply(P, Vd, g, 1, s);
    else launch_v_apply_list(P, Vd, g, 1, h->vlist2, nstep, s);
    hipLaunchKernelGGL(k_v_active, dim3(nb), dim3(RB), 0, s, (const double *)Vd, (const double *)g,
                       (const uint8_t *)P.vfix, (const double *)ubd, act, r, n, h->part);
    launch_dot(r, r, n, h->part + 4096, h->scal + 16, s);  // ||g_I||^2 (same value, fixed order)
    // With irreversibility most nodes sit at their upper bound: the reduced
    // system lives on a band around the cracks.  The chunks of node planes
    // holding a free node are listed, and CG runs over those only.
    launch_v_chunk_list(P, act, h->vflag, h->vlist, s);
    launch_v_dilated_list(P, h->vflag, h->vflag2, h->vlist2, s);
    // ||g_I||^2 and the two list lengths in one read
    hipLaunchKernelGGL(k_pack_counts, dim3(1), dim3(1), 0, s, (const int *)(h->vlist + vcap),
                       (const int *)(h->vlist2 + vcap), h->scal + 17);
    double got[3];
    if (hipMemcpyAsync(got, h->scal + 16, 24, hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
      return vfg_fail(h, VFG_E_HIP, "vfg_v_solve: reduction failed");
    const double pg2 = got[0];
    const int nlist = (int)got[1], ndil = (int)got[2];
    double pg = sqrt(pg2);
    if (!isfinite(pg)) { status = VFG_E_NAN; break; }
    if (pg0 < 0.) pg0 = pg > 0. ? pg : 1.;
    if (pg <= o->rtol * pg0 || pg <= 1e-14) break;
    // reduced Newton system on the inactive set.  The unlisted chunks hold
    // active nodes only, where r = z = 0; they are never written, so p, q
    // start at zero everywhere and stay zero there.  The diagonal is formed
    // in q on the listed chunks (1 / 0 -> 1 elsewhere); the first CG apply
    // overwrites it there.
    const int *vlist = h->vlist;
    launch_fill(q, 0., n, s);
    launch_v_diag_list(P, q, act, vlist, nlist, s);
    hipLaunchKernelGGL(k_recip, dim3(nb_(n)), dim3(256), 0, s, (const double *)q, dinv, n);
    launch_fill(x, 0., n, s);
    launch_fill(h->w[12], 0., n, s);
    FusedFn A = [&](const double *pold, double *pn, double *qq) {
      return launch_v_apply_cg_list(P, pold, pn, z, h->scal + S_BETA, qq, h->part, act, vlist, nlist, s,
                                    h->scal + S_DONE);
    };
    XrFn XR = [&](const double *pc, const double *qc) {
      return launch_v_xr_list(P.G, vlist, nlist, x, r, z, pc, qc, dinv, h->scal + S_ALPHA, h->part, s,
                              h->scal + S_DONE);
    };
    const double inner_rtol = fmax(inner_floor, fmin(0.1, 0.1 * pg / pg0));
    // r = -g on the free nodes, so ||r0|| = pg; the first batch holds two
    // iterations more than the previous Newton step's solve took
    const int its0 = inf.iterations;
    int prc = pcg_fused(h, n, A, dinv, x, r, z, p, h->w[12], q, inner_rtol, 0., 20000, 20, &inf, &XR, pg,
                        last_cg > 0 ? last_cg + 2 : 0);
    last_cg = inf.iterations - its0;
    if (prc == VFG_E_HIP || prc == VFG_E_NAN) { status = prc; break; }
    // projected step with Armijo backtracking on f(V) = 1/2 V.KV - b.V
    hipLaunchKernelGGL(k_zero_fixed, dim3(nb_(n)), dim3(256), 0, s, g, (const uint8_t *)P.vfix, n);
    double lam = 1.;
    for (int ls = 0; ls < 40; ls++) {
      hipLaunchKernelGGL(k_v_step, dim3(nb_(n)), dim3(256), 0, s, (const double *)Vd, (const double *)x,
                         (const uint8_t *)act, (const uint8_t *)P.vfix, (const double *)ubd, lam, Vn, D, n);
      // D vanishes off the free nodes: K D is needed on the listed chunks only
      // (KD is zeroed when the solve starts, so it is finite everywhere)
      launch_v_apply_list(P, D, KD, 2, vlist, nlist, s);
      // g.D and D.KD, read back together
      launch_dot(g, D, n, h->part, h->scal + 16, s);
      launch_dot(D, KD, n, h->part, h->scal + 17, s);
      double dots[2];
      if (hipMemcpyAsync(dots, h->scal + 16, 16, hipMemcpyDeviceToHost, s) != hipSuccess ||
          hipStreamSynchronize(s) != hipSuccess)
        return vfg_fail(h, VFG_E_HIP, "vfg_v_solve: line-search reduction failed");
      const double gd = dots[0], dkd = dots[1];
      if (gd + .5 * dkd <= 1e-4 * gd || ls == 39) break;
      lam *= .5;
    }
    hipMemcpyAsync(Vd, Vn, 8 * n, hipMemcpyDeviceToDevice, s);
    nstep = ndil;
  }
  inf.outer = it;
  if (status == VFG_OK && it >= (o->maxit < 200 ? o->maxit : 200)) status = VFG_E_MAXIT;
  inf.status = status;
  if (info) *info = inf;
  hipError_t e = hipMemcpyAsync(V, Vd, 8 * n, loc == VFG_HOST ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice, s);
  if (e != hipSuccess || hipStreamSynchronize(s) != hipSuccess) return vfg_fail(h, VFG_E_HIP, "vfg_v_solve: copy-out failed");
  if (status) return vfg_fail(h, status, "vfg_v_solve: stopped with status %d after %d Newton steps", status, it);
  return VFG_OK;
}

}  // extern "C"
