// vfg_engine.hip -- serving engine: continuous batching of alternate-
// minimisation requests (u-solve, then the bound-constrained V-solve, repeated;
// VFCommon.c:1192-1231, test1.c:53-69).
//
// Every resident request keeps its own handle (fields, coefficients, solver
// vectors) in HBM across all its iterations.  One replay of a captured hipGraph
// advances K CG iterations of every request that is in a CG phase: the FP32
// u-CG inside FP64 iterative refinement (or the plain FP64 u-CG) and the FP64
// V-CG on the inactive set.  Each CG iteration is one persistent apply launch
// per operator over a work list of (request, tile, z-chunk) items, chunked
// vector updates over all requests' vectors and one per-request scalar step;
// list lengths live on the device, so requests join and leave by table
// updates and the graph is captured once.  Between replays the host advances
// every request's state machine (refinement step, active-set Newton step,
// projected line search, ||V - Vold||_inf) from that request's own device
// scalars only, so a request's result is bitwise the same alone and in any
// mix.  The V-solve steps are those of vfg_v_solve (vfg_solve.hip).
#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>
#include <string>
#include <vector>
#include "vfg_batch_apply.h"
#include "vfg_bmg.h"
#include "vfg_handle.h"

namespace vfg {

// per-slot device scalars
enum {
  E_RZ = 0, E_PQ, E_ALPHA, E_BETA, E_RR, E_RR0, E_ACTIVE, E_STATUS, E_ITS, E_RTOL2, E_MAXIT,
  E_RN, E_PG2, E_GD, E_DKD, E_ERRV,
  E_MG,    // 1: this CG is preconditioned by the slot's multigrid V-cycle
  E_MGON,  // the V-cycle runs for the slot in this iteration (E_MG and E_ACTIVE)
  E_NS
};
constexpr int ECH = 4096;  // vector entries per chunk
constexpr int ERB = 256;

// vector slots of a request (its handle's solver work vectors, or U)
enum { QX = 0, QR, QZ, QP, QQ, QD, QX32, QR32, QZ32, QP32, QQ32, QD32, QN };
struct EVec {
  void *v[QN];
};
struct EChunk {
  int p;          // slot
  int pi;         // partial-sum index
  int64_t start;  // offset in the slot's vectors
  int len, pad;
};
// device counters: items of the u32 / u64 / V lists, chunks of the FP32 / FP64 lists
enum { C_U32 = 0, C_U64, C_V, C_F32, C_F64, C_N = 8 };

__device__ __forceinline__ double esum(double v) {
  __shared__ double sh[ERB / 64];
  return block_sum<ERB>(v, sh);
}

// ---------------------------------------------------------------- graph kernels
template <typename T>
__device__ __forceinline__ double chunk_dot(const EVec &t, int sa, int sb, const EChunk &c) {
  const T *a = (const T *)t.v[sa] + c.start, *b = (const T *)t.v[sb] + c.start;
  double s = 0.;
  for (int i = threadIdx.x; i < c.len; i += ERB) s += (double)a[i] * (double)b[i];
  return s;
}

// part[pi] = p.q over each chunk of the requests in CG
__global__ __launch_bounds__(ERB) void k_e_dot(const EVec *__restrict__ vt, const EChunk *__restrict__ c32,
                                               const EChunk *__restrict__ c64, const int *__restrict__ cnt,
                                               const double *__restrict__ sc, double *part) {
  const int n32 = cnt[C_F32], n = n32 + cnt[C_F64];
  for (int c = blockIdx.x; c < n; c += gridDim.x) {
    const bool f = c < n32;
    const EChunk k = f ? c32[c] : c64[c - n32];
    if (sc[(int64_t)k.p * E_NS + E_ACTIVE] == 0.) continue;  // uniform per block
    const EVec t = vt[k.p];
    double s = f ? chunk_dot<float>(t, QP32, QQ32, k) : chunk_dot<double>(t, QP, QQ, k);
    s = esum(s);
    if (threadIdx.x == 0) part[k.pi] = s;
  }
}

template <typename T>
__device__ __forceinline__ void chunk_xr(const EVec &t, int o, const EChunk &c, double al, double &rz, double &rr,
                                         bool r32 = false) {
  T *x = (T *)t.v[o + QX] + c.start, *r = (T *)t.v[o + QR] + c.start, *z = (T *)t.v[o + QZ] + c.start;
  const T *p = (const T *)t.v[o + QP] + c.start, *q = (const T *)t.v[o + QQ] + c.start;
  const T *dinv = (const T *)t.v[o + QD] + c.start;
  float *rf = (float *)t.v[QR32] + c.start;  // FP32 copy of r for the V-cycle (an FP64 multigrid CG)
  for (int k = threadIdx.x; k < c.len; k += ERB) {
    x[k] = (T)((double)x[k] + al * (double)p[k]);
    T ri = (T)((double)r[k] - al * (double)q[k]);
    r[k] = ri;
    if (r32) rf[k] = (float)ri;
    T zi = (T)(dinv[k] * ri);
    z[k] = zi;
    rz += (double)ri * (double)zi;
    rr += (double)ri * (double)ri;
  }
}

// x += alpha p; r -= alpha q; z = dinv r; part[2 pi] = r.z, part[2 pi + 1] = r.r
__global__ __launch_bounds__(ERB) void k_e_xr(const EVec *__restrict__ vt, const EChunk *__restrict__ c32,
                                              const EChunk *__restrict__ c64, const int *__restrict__ cnt,
                                              const double *__restrict__ sc, double *part) {
  const int n32 = cnt[C_F32], n = n32 + cnt[C_F64];
  for (int c = blockIdx.x; c < n; c += gridDim.x) {
    const bool f = c < n32;
    const EChunk k = f ? c32[c] : c64[c - n32];
    const double *S = sc + (int64_t)k.p * E_NS;
    if (S[E_ACTIVE] == 0.) continue;
    const EVec t = vt[k.p];
    double rz = 0., rr = 0.;
    if (f) chunk_xr<float>(t, QX32, k, S[E_ALPHA], rz, rr);
    else chunk_xr<double>(t, QX, k, S[E_ALPHA], rz, rr, S[E_MGON] != 0.);
    rz = esum(rz);
    rr = esum(rr);
    if (threadIdx.x == 0) {
      part[2 * k.pi] = rz;
      part[2 * k.pi + 1] = rr;
    }
  }
}

template <typename T>
__device__ __forceinline__ void chunk_p(const EVec &t, int o, const EChunk &c, double be) {
  T *p = (T *)t.v[o + QP] + c.start;
  const T *z = (const T *)t.v[o + QZ] + c.start;
  for (int k = threadIdx.x; k < c.len; k += ERB) p[k] = (T)((double)z[k] + be * (double)p[k]);
}

// p = z + beta p
__global__ __launch_bounds__(ERB) void k_e_p(const EVec *__restrict__ vt, const EChunk *__restrict__ c32,
                                             const EChunk *__restrict__ c64, const int *__restrict__ cnt,
                                             const double *__restrict__ sc) {
  const int n32 = cnt[C_F32], n = n32 + cnt[C_F64];
  for (int c = blockIdx.x; c < n; c += gridDim.x) {
    const bool f = c < n32;
    const EChunk k = f ? c32[c] : c64[c - n32];
    const double *S = sc + (int64_t)k.p * E_NS;
    if (S[E_ACTIVE] == 0.) continue;
    if (f) chunk_p<float>(vt[k.p], QX32, k, S[E_BETA]);
    else chunk_p<double>(vt[k.p], QX, k, S[E_BETA]);
  }
}

// part[2 pi] = r.z over the chunks of the slots whose z came from the
// V-cycle this iteration (k_e_xr wrote the Jacobi value there): FP32 chunks
// (the refinement's inner CG), and FP64 chunks (an FP64 CG under the FP32
// V-cycle), where z is also copied to the FP64 z the direction update reads
__global__ __launch_bounds__(ERB) void k_e_rz(const EVec *__restrict__ vt, const EChunk *__restrict__ c32,
                                              const EChunk *__restrict__ c64, const int *__restrict__ cnt,
                                              const double *__restrict__ sc, double *part) {
  const int n32 = cnt[C_F32], n = n32 + cnt[C_F64];
  for (int c = blockIdx.x; c < n; c += gridDim.x) {
    const bool f = c < n32;
    const EChunk k = f ? c32[c] : c64[c - n32];
    if (sc[(int64_t)k.p * E_NS + E_MGON] == 0.) continue;
    double rz;
    if (f) {
      rz = chunk_dot<float>(vt[k.p], QR32, QZ32, k);
    } else {
      const EVec t = vt[k.p];
      const double *r = (const double *)t.v[QR] + k.start;
      const float *z32 = (const float *)t.v[QZ32] + k.start;
      double *z = (double *)t.v[QZ] + k.start;
      rz = 0.;
      for (int i = threadIdx.x; i < k.len; i += ERB) {
        const double zi = (double)z32[i];
        z[i] = zi;
        rz += r[i] * zi;
      }
    }
    rz = esum(rz);
    if (threadIdx.x == 0) part[2 * k.pi] = rz;
  }
}

// per-slot CG scalars from the chunk partials [first, first + nch), fixed order.
// op 1: alpha = rz / p.q (breakdown if p.q <= 0); op 2: beta, convergence
__global__ __launch_bounds__(ERB) void k_e_scalar(const double *part, const int *first, const int *nch, int op,
                                                  double *sc) {
  const int p = blockIdx.x;
  const int m = nch[p];
  if (m == 0) return;
  const int w = op == 1 ? 1 : 2;
  double a = 0., b = 0.;
  for (int i = threadIdx.x; i < m; i += ERB) {
    a += part[(int64_t)(first[p] + i) * w];
    if (w == 2) b += part[(int64_t)(first[p] + i) * w + 1];
  }
  a = esum(a);
  b = esum(b);
  if (threadIdx.x != 0) return;
  double *S = sc + (int64_t)p * E_NS;
  if (S[E_ACTIVE] == 0.) {
    S[E_MGON] = 0.;
    return;
  }
  if (op == 1) {
    S[E_PQ] = a;
    if (!(a > 0.) || !isfinite(a)) {
      S[E_ALPHA] = 0.;
      S[E_ACTIVE] = 0.;
      S[E_STATUS] = isfinite(a) ? (double)VFG_E_BREAKDOWN : (double)VFG_E_NAN;
    } else {
      S[E_ALPHA] = S[E_RZ] / a;
    }
    S[E_MGON] = S[E_ACTIVE] != 0. && S[E_MG] != 0. ? 1. : 0.;
  } else {
    const double rz = S[E_RZ];
    S[E_BETA] = rz != 0. ? a / rz : 0.;
    S[E_RZ] = a;
    S[E_RR] = b;
    S[E_ITS] += 1.;
    if (!isfinite(a) || !isfinite(b)) {
      S[E_ACTIVE] = 0.;
      S[E_STATUS] = (double)VFG_E_NAN;
    } else if (b <= S[E_RTOL2] * S[E_RR0]) {
      S[E_ACTIVE] = 0.;
    } else if (S[E_ITS] >= S[E_MAXIT]) {
      S[E_ACTIVE] = 0.;
      S[E_STATUS] = (double)VFG_E_MAXIT;
    }
  }
}

// ------------------------------------------------------- per-request kernels
// z = dinv r (dinv == nullptr: z is given, the V-cycle's), p = z, partials of r.z and r.r
template <typename T>
__global__ __launch_bounds__(ERB) void k_e_init(const T *r, const T *dinv, T *z, T *p, int64_t n, double *part) {
  double rz = 0., rr = 0.;
  for (int64_t i = (int64_t)blockIdx.x * ERB + threadIdx.x; i < n; i += (int64_t)gridDim.x * ERB) {
    double ri = r[i];
    T zi = dinv ? (T)(dinv[i] * r[i]) : z[i];
    if (dinv) z[i] = zi;
    p[i] = zi;
    rz += ri * (double)zi;
    rr += ri * ri;
  }
  rz = esum(rz);
  rr = esum(rr);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = rz;
    part[2 * blockIdx.x + 1] = rr;
  }
}
__global__ __launch_bounds__(ERB) void k_e_init_fin(const double *part, int nb, double *S, double rtol,
                                                    double maxit, double mg) {
  double a = 0., b = 0.;
  for (int i = threadIdx.x; i < nb; i += ERB) {
    a += part[2 * i];
    b += part[2 * i + 1];
  }
  a = esum(a);
  b = esum(b);
  if (threadIdx.x != 0) return;
  S[E_RZ] = a;
  S[E_RR] = b;
  S[E_RR0] = b;
  S[E_ITS] = 0.;
  S[E_ALPHA] = 0.;
  S[E_BETA] = 0.;
  S[E_RTOL2] = rtol * rtol;
  S[E_MAXIT] = maxit;
  S[E_STATUS] = isfinite(b) ? 0. : (double)VFG_E_NAN;
  S[E_ACTIVE] = (b > 0. && isfinite(b)) ? 1. : 0.;
  S[E_MG] = mg;
  S[E_MGON] = mg != 0. && S[E_ACTIVE] != 0. ? 1. : 0.;
}
// a slot leaves its CG: no V-cycle for it until its next multigrid CG
__global__ void k_e_mgoff(double *S) {
  S[E_MG] = 0.;
  S[E_MGON] = 0.;
}
__global__ void k_e_flag(double *f, int p, double v) { f[p] = v; }
__global__ void k_e_neg(double *r, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) r[i] = -r[i];
}
// D = 1 / diag (1 where the diagonal is 0); D32 = FP32 copy
__global__ void k_e_prep(double *D, float *D32, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double d = D[i] != 0. ? 1. / D[i] : 1.;
  D[i] = d;
  if (D32) D32[i] = (float)d;
}
// R32 = R * s ; X32 = 0
__global__ void k_e_r32(const double *R, double s, float *R32, float *X32, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  R32[i] = (float)(R[i] * s);
  X32[i] = 0.f;
}
// Z = (double) Z32
__global__ void k_e_f2d(const float *Z32, double *Z, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) Z[i] = (double)Z32[i];
}
// U += s * X32
__global__ void k_e_axpy(double *U, const float *X32, double s, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) U[i] += s * (double)X32[i];
}

static unsigned enb(int64_t n) { return (unsigned)((n + 255) / 256); }

}  // namespace vfg

using namespace vfg;

// ------------------------------------------------------------------ host side
namespace {

enum { ST_FREE = 0, ST_U_BEGIN, ST_U_CHECK, ST_U_CG, ST_V_BEGIN, ST_V_GRAD, ST_V_GCHECK, ST_V_CG, ST_V_LS, ST_V_END,
       ST_AM_CHECK };
enum { K_NONE = 0, K_U32, K_U64, K_V64 };

struct ESlot {
  int st = ST_FREE;
  int kind = K_NONE;  // CG kind while in a CG phase
  vfg_handle *h = nullptr;
  double *U = nullptr, *V = nullptr;
  const double *BCU = nullptr, *ub = nullptr;
  int am_maxit = 1;
  double altmintol = 1e-4;
  int64_t tag = 0;
  int am = 0, outer = 0, newton = 0, ls = 0;
  double rn0 = -1., rs = 0., pg0 = -1., lam = 1.;
  int u_its = 0, v_its = 0, v_outer = 0, status = 0;
  double errV = 0.;
  bool vop_made = false;
};

}  // namespace

struct vfg_engine {
  vfg_engine_opts o;
  hipStream_t s = nullptr;
  std::string err;
  int ns = 0;
  std::vector<ESlot> slot;
  std::vector<vfg_engine_done> done;
  double *d_sc = nullptr, *h_sc = nullptr;  // h_sc pinned
  EVec *d_vt = nullptr;
  std::vector<EVec> h_vt;
  void *d_uop = nullptr, *d_vop = nullptr;
  std::vector<char> h_uop, h_vop;
  size_t uop_sz = 0, vop_sz = 0;
  BItem *d_it[3] = {nullptr, nullptr, nullptr};
  int cap_it[3] = {0, 0, 0};
  std::vector<BItem> h_it[3];
  EChunk *d_ch[2] = {nullptr, nullptr};
  int cap_ch[2] = {0, 0};
  std::vector<EChunk> h_ch[2];
  int *d_first = nullptr, *d_nch = nullptr, *d_cnt = nullptr;
  std::vector<int> h_first, h_nch;
  int h_cnt[C_N] = {0};
  double *d_part = nullptr;
  int cap_part = 0;
  bool dirty_lists = true, dirty_vt = true, dirty_uop = true, dirty_vop = true;
  hipGraphExec_t g = nullptr, g_mg = nullptr;  // replay graphs without / with the batched V-cycle
  int grid_apply = 0, grid_vapply = 0, grid_vec = 0;
  int64_t ticks = 0, replays = 0;
  // multigrid of the mixed engine: one member per slot; d_one: flags of an
  // eager V-cycle of a single slot
  bool mg = false, dirty_mg = false;
  BatchMG *bmg = nullptr;
  double *d_one = nullptr;
};

static int efail(vfg_engine *e, int code, const std::string &msg) {
  if (e) e->err = msg;
  return code;
}
#define ECHK(e, call)                                                          \
  do {                                                                         \
    hipError_t er_ = (call);                                                   \
    if (er_ != hipSuccess) return efail(e, VFG_E_HIP, hipGetErrorString(er_)); \
  } while (0)

static void drop_graphs(vfg_engine *e) {
  if (e->g) hipGraphExecDestroy(e->g);
  if (e->g_mg) hipGraphExecDestroy(e->g_mg);
  e->g = e->g_mg = nullptr;
}

static int egrow(vfg_engine *e, void **p, int *cap, int need, size_t elem) {
  if (need <= *cap) return VFG_OK;
  int nc = *cap ? *cap : 1024;
  while (nc < need) nc *= 2;
  if (*p) ECHK(e, hipFree(*p));
  *p = nullptr;
  if (hipMalloc(p, elem * (size_t)nc) != hipSuccess) {
    *p = nullptr;
    *cap = 0;
    return efail(e, VFG_E_NOMEM, "vfg_engine: device allocation failed");
  }
  *cap = nc;
  drop_graphs(e);  // the captured graphs hold the old address
  return VFG_OK;
}

static double *scal(vfg_engine *e, int p, int f) { return e->d_sc + (size_t)p * E_NS + f; }
static double hs(const vfg_engine *e, int p, int f) { return e->h_sc[(size_t)p * E_NS + f]; }

// CG init of one slot: z = dinv r (dinv == nullptr: z from the V-cycle), p = z,
// scalars (ACTIVE = 1 unless r = 0)
template <typename T>
static void cg_init(vfg_engine *e, int p, const T *r, const T *dinv, T *z, T *pv, int64_t n, double rtol, int maxit) {
  vfg_handle *h = e->slot[p].h;
  const int nb = red_blocks(n);
  hipLaunchKernelGGL(k_e_init<T>, dim3(nb), dim3(ERB), 0, e->s, r, dinv, z, pv, n, h->part);
  hipLaunchKernelGGL(k_e_init_fin, dim3(1), dim3(ERB), 0, e->s, (const double *)h->part, nb, scal(e, p, 0), rtol,
                     (double)maxit, dinv ? 0. : 1.);
}

// the multigrid work lists follow the resident requests; a list that moved
// invalidates the captured replay graph
static int mg_lists(vfg_engine *e) {
  if (!e->mg || !e->dirty_mg) return VFG_OK;
  bool grew = false;
  int rc = bmg_rebuild(e->bmg, e->s, &grew);
  if (rc) return efail(e, rc, "vfg_engine: multigrid work lists failed");
  if (grew) drop_graphs(e);
  e->dirty_mg = false;
  return VFG_OK;
}

static void set_vt_u(vfg_engine *e, int p) {
  ESlot &q = e->slot[p];
  EVec &t = e->h_vt[p];
  t.v[QX] = q.U;
  for (int i = 1; i < 6; i++) t.v[QX + i] = q.h->w[i];
  for (int i = 0; i < 6; i++) t.v[QX32 + i] = q.h->wf[i];
  e->dirty_vt = true;
}
static void set_vt_v(vfg_engine *e, int p) {
  ESlot &q = e->slot[p];
  EVec &t = e->h_vt[p];
  for (int i = 0; i < 6; i++) t.v[QX + i] = q.h->w[i];
  e->dirty_vt = true;
}

static void enter_cg(vfg_engine *e, int p, int kind) {
  e->slot[p].kind = kind;
  e->dirty_lists = true;
}
static void leave_cg(vfg_engine *e, int p) {
  e->slot[p].kind = K_NONE;
  e->dirty_lists = true;
  hipLaunchKernelGGL(k_e_mgoff, dim3(1), dim3(1), 0, e->s, scal(e, p, 0));
}

// line-search candidate: Vn = P(V + lam x), D = Vn - V, KD = K D, g.D, D.KD
static void v_ls_step(vfg_engine *e, int p) {
  ESlot &q = e->slot[p];
  vfg_handle *h = q.h;
  DevProblem &P = h->P;
  const int64_t n = P.nn;
  launch_v_step(q.V, h->w[0], h->act, P.vfix, q.ub, q.lam, h->w[8], h->w[9], n, e->s);
  launch_v_apply(P, h->w[9], h->w[10], 2, e->s);
  launch_dot(h->w[7], h->w[9], n, h->part, scal(e, p, E_GD), e->s);
  launch_dot(h->w[9], h->w[10], n, h->part, scal(e, p, E_DKD), e->s);
}

static void finish(vfg_engine *e, int p) {
  ESlot &q = e->slot[p];
  vfg_engine_done d;
  d.tag = q.tag;
  d.status = q.status;
  d.am_iterations = q.am;
  d.u_its = q.u_its;
  d.v_outer = q.v_outer;
  d.v_its = q.v_its;
  d.errV = q.errV;
  e->done.push_back(d);
  if (q.kind != K_NONE) leave_cg(e, p);
  if (e->mg) {
    bmg_set_member(e->bmg, p, nullptr, nullptr, nullptr);
    e->dirty_mg = true;
  }
  q = ESlot();
  e->dirty_lists = true;
}

// Advances one request's alternate-minimisation state machine as far as it
// can go without a device result it has not read yet.
static int advance(vfg_engine *e, int p) {
  ESlot &q = e->slot[p];
  vfg_handle *h = q.h;
  DevProblem &P = h->P;
  hipStream_t s = e->s;
  const int64_t n3 = 3 * P.nn, nn = P.nn;
  const bool mixed = e->o.precision == VFG_PREC_MIXED;
  for (;;) {
    switch (q.st) {
      case ST_U_BEGIN: {  // VF_StepU: U carries its Dirichlet values; r = -F(U)
        set_vt_u(e, p);
        launch_vec_apply_dirichlet(P, 0, q.U, q.BCU, s);
        launch_u_rhs(P, q.U, q.BCU, h->w[QR], 1, s);
        hipLaunchKernelGGL(k_e_neg, dim3(enb(n3)), dim3(256), 0, s, h->w[QR], n3);
        launch_u_diag(P, h->w[QD], s);
        hipLaunchKernelGGL(k_e_prep, dim3(enb(n3)), dim3(256), 0, s, h->w[QD], mixed ? h->wf[QD32 - QX32] : nullptr,
                           n3);
        if (e->mg) {
          // coefficients, diagonals and smoother bounds of the slot's hierarchy
          // from the current V; the set-up's eigenvalue estimate borrows the
          // FP32 work vectors, so the FP32 inverse diagonal is written again
          int rc = mg_lists(e);  // a request admitted since the last step: its levels' records
          if (rc) return rc;
          if ((rc = bmg_setup_member(e->bmg, p, h->w[QD], s))) return efail(e, rc, "vfg_engine: multigrid set-up failed");
          launch_copy_d2f(h->w[QD], h->wf[QD32 - QX32], n3, s);
        }
        q.outer = 0;
        q.rn0 = -1.;
        if (mixed) {
          launch_dot(h->w[QR], h->w[QR], n3, h->part, scal(e, p, E_RN), s);
          q.st = ST_U_CHECK;
        } else if (e->mg) {  // FP64 CG under the slot's FP32 V-cycle: z0 = V-cycle(r0) of this slot alone
          int rc = mg_lists(e);
          if (rc) return rc;
          launch_copy_d2f(h->w[QR], h->wf[QR32 - QX32], n3, s);
          hipLaunchKernelGGL(k_e_flag, dim3(1), dim3(1), 0, s, e->d_one, p, 1.);
          bmg_vcycle(e->bmg, e->d_one, 1, s);
          hipLaunchKernelGGL(k_e_flag, dim3(1), dim3(1), 0, s, e->d_one, p, 0.);
          hipLaunchKernelGGL(k_e_f2d, dim3(enb(n3)), dim3(256), 0, s, (const float *)h->wf[QZ32 - QX32], h->w[QZ], n3);
          cg_init<double>(e, p, h->w[QR], (const double *)nullptr, h->w[QZ], h->w[QP], n3, e->o.u_rtol, e->o.maxit);
          enter_cg(e, p, K_U64);
          q.st = ST_U_CG;
        } else {
          cg_init<double>(e, p, h->w[QR], h->w[QD], h->w[QZ], h->w[QP], n3, e->o.u_rtol, e->o.maxit);
          enter_cg(e, p, K_U64);
          q.st = ST_U_CG;
        }
        return VFG_OK;
      }
      case ST_U_CHECK: {  // refinement: |r| against the tolerance, else an FP32 inner solve
        const double rn = sqrt(hs(e, p, E_RN));
        if (!isfinite(rn)) {
          q.status = VFG_E_NAN;
          finish(e, p);
          return VFG_OK;
        }
        if (q.rn0 < 0.) q.rn0 = rn;
        // converged at the tolerance or at the FP64 residual's rounding floor
        // (an outer step no longer halves it: the rule of u_solve_mixed)
        const bool conv = rn <= e->o.u_rtol * q.rn0 || rn == 0. || (q.outer > 0 && rn > 0.5 * q.rs);
        if (conv || q.outer >= 60) {
          if (!conv && !q.status) q.status = VFG_E_MAXIT;
          q.st = ST_V_BEGIN;
          break;
        }
        q.outer++;
        q.rs = rn;
        hipLaunchKernelGGL(k_e_r32, dim3(enb(n3)), dim3(256), 0, s, (const double *)h->w[QR], 1. / rn, h->wf[1],
                           h->wf[0], n3);
        if (e->mg) {  // z0 = V-cycle(r0) of this slot alone
          int rc = mg_lists(e);
          if (rc) return rc;
          hipLaunchKernelGGL(k_e_flag, dim3(1), dim3(1), 0, s, e->d_one, p, 1.);
          bmg_vcycle(e->bmg, e->d_one, 1, s);
          hipLaunchKernelGGL(k_e_flag, dim3(1), dim3(1), 0, s, e->d_one, p, 0.);
          // 1e-4: with 1e-5 the served stream's u-CG iterations per request rose
          // 133 -> 147 and its solves/s fell 3% (the batched solves gain from it)
          cg_init<float>(e, p, h->wf[1], (const float *)nullptr, h->wf[2], h->wf[3], n3, 1e-4, e->o.maxit);
        } else {
          cg_init<float>(e, p, h->wf[1], h->wf[5], h->wf[2], h->wf[3], n3, 1e-4, e->o.maxit);
        }
        enter_cg(e, p, K_U32);
        q.st = ST_U_CG;
        return VFG_OK;
      }
      case ST_U_CG: {
        if (hs(e, p, E_ACTIVE) != 0.) return VFG_OK;  // still iterating
        q.u_its += (int)hs(e, p, E_ITS);
        const int st = (int)hs(e, p, E_STATUS);
        leave_cg(e, p);
        if (st == VFG_E_NAN) {
          q.status = st;
          finish(e, p);
          return VFG_OK;
        }
        if (mixed) {  // x += |r| d in FP64, r = -F(x)
          hipLaunchKernelGGL(k_e_axpy, dim3(enb(n3)), dim3(256), 0, s, q.U, (const float *)h->wf[0], q.rs, n3);
          launch_u_rhs(P, q.U, q.BCU, h->w[QR], 1, s);
          hipLaunchKernelGGL(k_e_neg, dim3(enb(n3)), dim3(256), 0, s, h->w[QR], n3);
          launch_dot(h->w[QR], h->w[QR], n3, h->part, scal(e, p, E_RN), s);
          q.st = ST_U_CHECK;
          return VFG_OK;
        }
        if (st && !q.status) q.status = st;
        q.st = ST_V_BEGIN;
        break;
      }
      case ST_V_BEGIN: {  // VF_StepV set-up (vfg_v_solve): coupling from U, Dirichlet V, projection
        int rc;
        if ((rc = vfg_alloc(h, (void **)&P.Vq, 8 * (size_t)vq_doubles(P.G))) || (rc = vfg_alloc(h, (void **)&h->act, nn)))
          return efail(e, rc, h->err);
        if (!q.vop_made) {
          make_vop(P, h->w[QP], h->w[QQ], h->act, &e->h_vop[(size_t)p * e->vop_sz]);
          q.vop_made = true;
          e->dirty_vop = true;
        }
        set_vt_v(e, p);
        launch_v_setup(P, q.U, s);
        h->v_ready = true;
        launch_v_fixmask(P, s);
        ECHK(e, hipMemcpyAsync(h->w[11], q.V, 8 * nn, hipMemcpyDeviceToDevice, s));  // Vold
        launch_vec_apply_dirichlet(P, 1, q.V, q.V, s);
        launch_v_clamp(q.V, P.vfix, q.ub, nn, s);
        q.newton = 0;
        q.pg0 = -1.;
        q.st = ST_V_GRAD;
        break;
      }
      case ST_V_GRAD: {  // g = F(V); active set; r = -g on the inactive set
        launch_v_apply(P, q.V, h->w[7], 1, s);
        launch_v_active(q.V, h->w[7], P.vfix, q.ub, h->act, h->w[QR], nn, h->part, s);
        launch_dot(h->w[QR], h->w[QR], nn, h->part, scal(e, p, E_PG2), s);
        q.st = ST_V_GCHECK;
        return VFG_OK;
      }
      case ST_V_GCHECK: {
        const double pg = sqrt(hs(e, p, E_PG2));
        if (!isfinite(pg)) {
          q.status = VFG_E_NAN;
          finish(e, p);
          return VFG_OK;
        }
        if (q.pg0 < 0.) q.pg0 = pg > 0. ? pg : 1.;
        const bool conv = pg <= e->o.v_rtol * q.pg0 || pg <= 1e-14;
        if (conv || q.newton >= 200) {
          if (!conv && !q.status) q.status = VFG_E_MAXIT;
          q.st = ST_V_END;
          break;
        }
        launch_v_diag(P, h->w[QD], s, h->act);
        hipLaunchKernelGGL(k_e_prep, dim3(enb(nn)), dim3(256), 0, s, h->w[QD], (float *)nullptr, nn);
        launch_fill(h->w[QX], 0., nn, s);
        // inexact Newton forcing term, as vfg_v_solve
        const double eta = fmax(fmin(e->o.v_rtol, 1e-12), fmin(0.1, 0.1 * pg / q.pg0));
        cg_init<double>(e, p, h->w[QR], h->w[QD], h->w[QZ], h->w[QP], nn, eta, 20000);
        enter_cg(e, p, K_V64);
        q.st = ST_V_CG;
        return VFG_OK;
      }
      case ST_V_CG: {
        if (hs(e, p, E_ACTIVE) != 0.) return VFG_OK;
        q.v_its += (int)hs(e, p, E_ITS);
        leave_cg(e, p);
        if ((int)hs(e, p, E_STATUS) == VFG_E_NAN) {
          q.status = VFG_E_NAN;
          finish(e, p);
          return VFG_OK;
        }
        launch_zero_fixed(h->w[7], P.vfix, nn, s);
        q.lam = 1.;
        q.ls = 0;
        v_ls_step(e, p);
        q.st = ST_V_LS;
        return VFG_OK;
      }
      case ST_V_LS: {  // Armijo on f(V) = 1/2 V.KV - b.V along the projected path
        const double gd = hs(e, p, E_GD), dkd = hs(e, p, E_DKD);
        if (gd + .5 * dkd <= 1e-4 * gd || q.ls == 39) {
          ECHK(e, hipMemcpyAsync(q.V, h->w[8], 8 * nn, hipMemcpyDeviceToDevice, s));
          q.newton++;
          q.v_outer++;
          q.st = ST_V_GRAD;
          break;
        }
        q.lam *= .5;
        q.ls++;
        v_ls_step(e, p);
        return VFG_OK;
      }
      case ST_V_END: {  // the handle's V for the next u-solve; ||V - Vold||_inf
        ECHK(e, hipMemcpyAsync(P.V, q.V, 8 * nn, hipMemcpyDeviceToDevice, s));
        launch_copy_d2f(P.V, P.V32, nn, s);
        vfg_u_invalidate(q.h);
        launch_maxabs_diff(q.V, h->w[11], nn, h->part, scal(e, p, E_ERRV), s);
        q.st = ST_AM_CHECK;
        return VFG_OK;
      }
      case ST_AM_CHECK: {
        q.errV = hs(e, p, E_ERRV);
        q.am++;
        if (q.am >= q.am_maxit || q.errV < q.altmintol || q.status) {
          finish(e, p);
          return VFG_OK;
        }
        q.st = ST_U_BEGIN;
        break;
      }
      default:
        return VFG_OK;
    }
  }
}

// work lists of the requests in a CG phase
static int rebuild_lists(vfg_engine *e) {
  for (int k = 0; k < 3; k++) e->h_it[k].clear();
  for (int k = 0; k < 2; k++) e->h_ch[k].clear();
  const int slots = batch_slots();
  int pi = 0;
  for (int p = 0; p < e->ns; p++) e->h_first[p] = e->h_nch[p] = 0;
  for (int pass = 0; pass < 2; pass++)  // FP32 chunks first, then FP64: partial indices follow
    for (int p = 0; p < e->ns; p++) {
      ESlot &q = e->slot[p];
      if (q.st == ST_FREE || q.kind == K_NONE) continue;
      const bool f32 = q.kind == K_U32;
      if ((pass == 0) != f32) continue;
      const int64_t len = q.kind == K_V64 ? q.h->P.nn : 3 * q.h->P.nn;
      e->h_first[p] = pi;
      for (int64_t st = 0; st < len; st += ECH)
        e->h_ch[f32 ? 0 : 1].push_back({p, pi++, st, (int)(len - st < ECH ? len - st : ECH), 0});
      e->h_nch[p] = pi - e->h_first[p];
    }
  for (int p = 0; p < e->ns; p++) {
    ESlot &q = e->slot[p];
    if (q.st == ST_FREE || q.kind == K_NONE) continue;
    const Grid &G = q.h->P.G;
    const int tx = (G.nx + 14) / 15, ty = (G.ny + 14) / 15;
    const int k = q.kind == K_U32 ? 0 : q.kind == K_U64 ? 1 : 2;
    const int oz = k == 2 ? v_oz(G) : choose_oz_pub(G, tx * ty, slots);
    const int tz = (G.nz + oz - 1) / oz;
    for (int z = 0; z < tz; z++)
      for (int y = 0; y < ty; y++)
        for (int x = 0; x < tx; x++) e->h_it[k].push_back({p, x, y, z});
  }
  int rc;
  for (int k = 0; k < 3; k++)
    if ((rc = egrow(e, (void **)&e->d_it[k], &e->cap_it[k], (int)e->h_it[k].size(), sizeof(BItem)))) return rc;
  for (int k = 0; k < 2; k++)
    if ((rc = egrow(e, (void **)&e->d_ch[k], &e->cap_ch[k], (int)e->h_ch[k].size(), sizeof(EChunk)))) return rc;
  if ((rc = egrow(e, (void **)&e->d_part, &e->cap_part, 2 * pi, sizeof(double)))) return rc;
  e->h_cnt[C_U32] = (int)e->h_it[0].size();
  e->h_cnt[C_U64] = (int)e->h_it[1].size();
  e->h_cnt[C_V] = (int)e->h_it[2].size();
  e->h_cnt[C_F32] = (int)e->h_ch[0].size();
  e->h_cnt[C_F64] = (int)e->h_ch[1].size();
  for (int k = 0; k < 3; k++)
    if (!e->h_it[k].empty())
      ECHK(e, hipMemcpyAsync(e->d_it[k], e->h_it[k].data(), sizeof(BItem) * e->h_it[k].size(),
                             hipMemcpyHostToDevice, e->s));
  for (int k = 0; k < 2; k++)
    if (!e->h_ch[k].empty())
      ECHK(e, hipMemcpyAsync(e->d_ch[k], e->h_ch[k].data(), sizeof(EChunk) * e->h_ch[k].size(),
                             hipMemcpyHostToDevice, e->s));
  ECHK(e, hipMemcpyAsync(e->d_first, e->h_first.data(), sizeof(int) * e->ns, hipMemcpyHostToDevice, e->s));
  ECHK(e, hipMemcpyAsync(e->d_nch, e->h_nch.data(), sizeof(int) * e->ns, hipMemcpyHostToDevice, e->s));
  ECHK(e, hipMemcpyAsync(e->d_cnt, e->h_cnt, sizeof(int) * C_N, hipMemcpyHostToDevice, e->s));
  e->dirty_lists = false;
  return VFG_OK;
}

// one CG iteration of every request in a CG phase; with_mg: some request is in
// a multigrid u-CG, so the batched V-cycle runs (for the flagged slots)
static void e_iter(vfg_engine *e, bool with_mg) {
  hipStream_t s = e->s;
  const double *fl = e->d_sc + E_ACTIVE;
  if (e->o.precision == VFG_PREC_MIXED)
    launch_engine_u_apply(1, 0, e->d_uop, e->d_it[0], 0, e->d_cnt + C_U32, e->grid_apply, fl, E_NS, s);
  else
    launch_engine_u_apply(0, 0, e->d_uop, e->d_it[1], 0, e->d_cnt + C_U64, e->grid_apply, fl, E_NS, s);
  launch_engine_v_apply(e->d_vop, e->d_it[2], e->d_cnt + C_V, e->grid_vapply, fl, E_NS, s);
  hipLaunchKernelGGL(k_e_dot, dim3(e->grid_vec), dim3(ERB), 0, s, (const EVec *)e->d_vt, (const EChunk *)e->d_ch[0],
                     (const EChunk *)e->d_ch[1], (const int *)e->d_cnt, (const double *)e->d_sc, e->d_part);
  hipLaunchKernelGGL(k_e_scalar, dim3(e->ns), dim3(ERB), 0, s, (const double *)e->d_part, (const int *)e->d_first,
                     (const int *)e->d_nch, 1, e->d_sc);
  hipLaunchKernelGGL(k_e_xr, dim3(e->grid_vec), dim3(ERB), 0, s, (const EVec *)e->d_vt, (const EChunk *)e->d_ch[0],
                     (const EChunk *)e->d_ch[1], (const int *)e->d_cnt, (const double *)e->d_sc, e->d_part);
  if (with_mg) {  // z = V-cycle(r) of the multigrid slots, then their r.z
    bmg_vcycle(e->bmg, e->d_sc + E_MGON, E_NS, s);
    hipLaunchKernelGGL(k_e_rz, dim3(e->grid_vec), dim3(ERB), 0, s, (const EVec *)e->d_vt, (const EChunk *)e->d_ch[0],
                       (const EChunk *)e->d_ch[1], (const int *)e->d_cnt, (const double *)e->d_sc, e->d_part);
  }
  hipLaunchKernelGGL(k_e_scalar, dim3(e->ns), dim3(ERB), 0, s, (const double *)e->d_part, (const int *)e->d_first,
                     (const int *)e->d_nch, 2, e->d_sc);
  hipLaunchKernelGGL(k_e_p, dim3(e->grid_vec), dim3(ERB), 0, s, (const EVec *)e->d_vt, (const EChunk *)e->d_ch[0],
                     (const EChunk *)e->d_ch[1], (const int *)e->d_cnt, (const double *)e->d_sc);
}

static int replay(vfg_engine *e) {
  const int K = e->o.graph_chunk;
  bool with_mg = false;
  if (e->mg)
    for (const ESlot &q : e->slot) with_mg |= q.st != ST_FREE && (q.kind == K_U32 || q.kind == K_U64);
  if (!e->o.use_graph) {
    for (int k = 0; k < K; k++) e_iter(e, with_mg);
    ECHK(e, hipGetLastError());
    e->replays++;
    return VFG_OK;
  }
  hipGraphExec_t &gx = with_mg ? e->g_mg : e->g;
  if (!gx) {
    hipGraph_t gr;
    ECHK(e, hipStreamBeginCapture(e->s, hipStreamCaptureModeThreadLocal));
    for (int k = 0; k < K; k++) e_iter(e, with_mg);
    ECHK(e, hipStreamEndCapture(e->s, &gr));
    hipError_t er = hipGraphInstantiate(&gx, gr, nullptr, nullptr, 0);
    hipGraphDestroy(gr);
    if (er != hipSuccess) {
      gx = nullptr;
      return efail(e, VFG_E_HIP, "vfg_engine: hipGraphInstantiate failed");
    }
  }
  ECHK(e, hipGraphLaunch(gx, e->s));
  e->replays++;
  return VFG_OK;
}

extern "C" {

void vfg_engine_opts_default(vfg_engine_opts *o) {
  if (!o) return;
  o->max_slots = 64;
  o->precision = VFG_PREC_MIXED;
  o->u_rtol = 1e-10;
  o->v_rtol = 1e-8;
  o->maxit = 100000;
  o->graph_chunk = 1;  // measured best: 456 solves/s at 1 against 430 at 4 and 388 at 16 (200-request stream)
  o->use_graph = 1;
  o->pc = VFG_PC_MG;
}

const char *vfg_engine_last_error(const vfg_engine *e) { return e ? e->err.c_str() : "NULL engine"; }

int vfg_engine_destroy(vfg_engine *e) {
  if (!e) return VFG_OK;
  if (e->s) hipStreamSynchronize(e->s);
  drop_graphs(e);
  bmg_destroy(e->bmg);
  if (e->d_one) hipFree(e->d_one);
  void *ps[] = {e->d_sc, e->d_vt, e->d_uop, e->d_vop, e->d_it[0], e->d_it[1], e->d_it[2], e->d_ch[0], e->d_ch[1],
                e->d_first, e->d_nch, e->d_cnt, e->d_part};
  for (void *p : ps)
    if (p) hipFree(p);
  if (e->h_sc) hipHostFree(e->h_sc);
  if (e->s) hipStreamDestroy(e->s);
  delete e;
  return VFG_OK;
}

int vfg_engine_create(const vfg_engine_opts *o, vfg_engine **out) {
  if (!out) return VFG_E_ARG;
  *out = nullptr;
  vfg_engine_opts d;
  vfg_engine_opts_default(&d);
  if (!o) o = &d;
  if (o->max_slots < 1 || o->max_slots > 4096 || o->graph_chunk < 1 || !(o->u_rtol > 0.) || !(o->v_rtol > 0.) ||
      o->maxit < 1 || (o->precision != VFG_PREC_MIXED && o->precision != VFG_PREC_FP64) ||
      (o->pc != VFG_PC_MG && o->pc != VFG_PC_JACOBI))
    return VFG_E_ARG;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return VFG_E_NODEV;
  vfg_engine *e = new vfg_engine();
  e->o = *o;
  e->ns = o->max_slots;
  e->slot.assign(e->ns, ESlot());
  e->h_vt.assign(e->ns, EVec());
  e->h_first.assign(e->ns, 0);
  e->h_nch.assign(e->ns, 0);
  e->uop_sz = batch_op_bytes(o->precision == VFG_PREC_MIXED ? 1 : 0);
  e->vop_sz = vop_bytes();
  e->h_uop.assign(e->uop_sz * e->ns, 0);
  e->h_vop.assign(e->vop_sz * e->ns, 0);
  hipDeviceProp_t prop;
  int cus = 256;
  if (hipGetDeviceProperties(&prop, dev) == hipSuccess) cus = prop.multiProcessorCount;
  e->grid_apply = batch_slots();
  e->grid_vapply = 3 * cus;
  e->grid_vec = 4 * cus;
  bool ok = hipStreamCreateWithFlags(&e->s, hipStreamNonBlocking) == hipSuccess &&
            hipMalloc((void **)&e->d_sc, 8 * E_NS * (size_t)e->ns) == hipSuccess &&
            hipHostMalloc((void **)&e->h_sc, 8 * E_NS * (size_t)e->ns, hipHostMallocDefault) == hipSuccess &&
            hipMalloc((void **)&e->d_vt, sizeof(EVec) * e->ns) == hipSuccess &&
            hipMalloc(&e->d_uop, e->uop_sz * e->ns) == hipSuccess &&
            hipMalloc(&e->d_vop, e->vop_sz * e->ns) == hipSuccess &&
            hipMalloc((void **)&e->d_first, sizeof(int) * e->ns) == hipSuccess &&
            hipMalloc((void **)&e->d_nch, sizeof(int) * e->ns) == hipSuccess &&
            hipMalloc((void **)&e->d_cnt, sizeof(int) * C_N) == hipSuccess;
  e->mg = o->pc == VFG_PC_MG;
  if (ok && e->mg)
    ok = bmg_create_slots(e->ns, &e->bmg) == VFG_OK && hipMalloc((void **)&e->d_one, 8 * (size_t)e->ns) == hipSuccess &&
         hipMemset(e->d_one, 0, 8 * (size_t)e->ns) == hipSuccess;
  if (!ok) {
    vfg_engine_destroy(e);
    return VFG_E_NOMEM;
  }
  memset(e->h_sc, 0, 8 * E_NS * (size_t)e->ns);
  if (hipMemcpy(e->d_sc, e->h_sc, 8 * E_NS * (size_t)e->ns, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemset(e->d_vt, 0, sizeof(EVec) * e->ns) != hipSuccess ||
      hipMemset(e->d_uop, 0, e->uop_sz * e->ns) != hipSuccess ||
      hipMemset(e->d_vop, 0, e->vop_sz * e->ns) != hipSuccess) {
    vfg_engine_destroy(e);
    return VFG_E_HIP;
  }
  int rc = rebuild_lists(e);  // empty lists, device counters = 0
  if (!rc && hipStreamSynchronize(e->s) != hipSuccess) rc = VFG_E_HIP;
  if (rc) {
    vfg_engine_destroy(e);
    return rc;
  }
  *out = e;
  return VFG_OK;
}

int vfg_engine_free_slots(const vfg_engine *e) {
  if (!e) return VFG_E_ARG;
  int n = 0;
  for (const ESlot &q : e->slot) n += q.st == ST_FREE;
  return n;
}

int vfg_engine_submit(vfg_engine *e, vfg_handle *h, double *U, const double *BCU, double *V, const double *VIrrev,
                      int am_maxit, double altmintol, int64_t tag) {
  if (!e) return VFG_E_ARG;
  if (!h || !U || !BCU || !V || !VIrrev || am_maxit < 1) return efail(e, VFG_E_ARG, "vfg_engine_submit: bad arguments");
  if (h->P.unilateral)
    return efail(e, VFG_E_ARG, "vfg_engine_submit: UNILATERAL_NOCOMPRESSION requests run on the single-problem path");
  int p = -1;
  for (int i = 0; i < e->ns && p < 0; i++)
    if (e->slot[i].st == ST_FREE) p = i;
  if (p < 0) return efail(e, VFG_E_BUDGET, "vfg_engine_submit: no free slot");
  int rc = vfg_work(h, 12, 6);
  if (rc) return efail(e, rc, h->err);
  // the handle's own stream may still hold set-up work
  ECHK(e, hipStreamSynchronize(h->stream));
  ESlot &q = e->slot[p];
  q = ESlot();
  q.h = h;
  q.U = U;
  q.BCU = BCU;
  q.V = V;
  q.ub = h->P.atnum == 1 ? VIrrev : nullptr;
  q.am_maxit = am_maxit;
  q.altmintol = altmintol;
  q.tag = tag;
  q.st = ST_U_BEGIN;
  const int mixed = e->o.precision == VFG_PREC_MIXED;
  const Grid &G = h->P.G;
  const int oz = choose_oz_pub(G, ((G.nx + 14) / 15) * ((G.ny + 14) / 15), batch_slots());
  make_uop(h->P, mixed ? (const void *)h->wf[QP32 - QX32] : (const void *)h->w[QP],
           mixed ? (void *)h->wf[QQ32 - QX32] : (void *)h->w[QQ], mixed, oz, &e->h_uop[(size_t)p * e->uop_sz]);
  e->dirty_uop = true;
  set_vt_u(e, p);
  if (e->mg) {  // the request's hierarchy (built once per handle) joins the batched V-cycle
    if ((rc = bmg_set_member(e->bmg, p, h, h->wf[QR32 - QX32], h->wf[QZ32 - QX32]))) {
      q = ESlot();
      return efail(e, rc, "vfg_engine_submit: multigrid hierarchy failed");
    }
    e->dirty_mg = true;
  }
  return VFG_OK;
}

int vfg_engine_step(vfg_engine *e) {
  if (!e) return VFG_E_ARG;
  int used = 0;
  for (const ESlot &q : e->slot) used += q.st != ST_FREE;
  if (!used) return 0;
  ECHK(e, hipMemcpyAsync(e->h_sc, e->d_sc, 8 * E_NS * (size_t)e->ns, hipMemcpyDeviceToHost, e->s));
  ECHK(e, hipStreamSynchronize(e->s));
  for (int p = 0; p < e->ns; p++) {
    if (e->slot[p].st == ST_FREE) continue;
    int rc = advance(e, p);
    if (rc) return rc;
  }
  if (e->dirty_vt) {
    ECHK(e, hipMemcpyAsync(e->d_vt, e->h_vt.data(), sizeof(EVec) * e->ns, hipMemcpyHostToDevice, e->s));
    e->dirty_vt = false;
  }
  if (e->dirty_uop) {
    ECHK(e, hipMemcpyAsync(e->d_uop, e->h_uop.data(), e->uop_sz * e->ns, hipMemcpyHostToDevice, e->s));
    e->dirty_uop = false;
  }
  if (e->dirty_vop) {
    ECHK(e, hipMemcpyAsync(e->d_vop, e->h_vop.data(), e->vop_sz * e->ns, hipMemcpyHostToDevice, e->s));
    e->dirty_vop = false;
  }
  int rc;
  if (e->dirty_lists && (rc = rebuild_lists(e))) return rc;
  if ((rc = mg_lists(e))) return rc;
  bool any_cg = false;
  used = 0;
  for (const ESlot &q : e->slot) {
    used += q.st != ST_FREE;
    any_cg |= q.st != ST_FREE && q.kind != K_NONE;
  }
  if (any_cg && (rc = replay(e))) return rc;
  ECHK(e, hipGetLastError());
  e->ticks++;
  return used;
}

int vfg_engine_poll(vfg_engine *e, vfg_engine_done *out, int max) {
  if (!e || (!out && max > 0) || max < 0) return VFG_E_ARG;
  int n = (int)e->done.size() < max ? (int)e->done.size() : max;
  for (int i = 0; i < n; i++) out[i] = e->done[i];
  e->done.erase(e->done.begin(), e->done.begin() + n);
  return n;
}

int vfg_engine_synchronize(vfg_engine *e) {
  if (!e) return VFG_E_ARG;
  ECHK(e, hipStreamSynchronize(e->s));
  return VFG_OK;
}

int vfg_engine_stats(const vfg_engine *e, int64_t *ticks, int64_t *replays) {
  if (!e) return VFG_E_ARG;
  if (ticks) *ticks = e->ticks;
  if (replays) *replays = e->replays;
  return VFG_OK;
}

}  // extern "C"
