// vfg_batch.hip -- ensembles of independent problems of different sizes.
//
// Apply: one persistent launch walks a work list of (member, tile x, tile y,
// z-chunk) items with the marching node-gather of the single-problem kernel;
// every member keeps its own grid, coefficients and Dirichlet table.
//
// Solve: preconditioned CG on all members at once over concatenated vectors.
// Vector work is cut into chunks that never straddle two members; inner
// products are per-chunk partial sums reduced per member in a fixed order, and
// every member has its own FP64 scalars.  A member that has converged (or
// broken down) is frozen on the device (alpha = beta = 0), so the iterate
// sequence of each member is independent of the others and of how many
// iterations the ensemble runs: a member's result is bitwise the same alone
// and in any mix.  Chunks of iterations are captured once in a hipGraph and
// replayed until every member is frozen.
//
// Mixed precision: FP64 iterative refinement around the FP32 inner CG, the
// refinement residual computed by the FP64 operator of each member.
#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>
#include <string>
#include <vector>
#include "vfg_handle.h"

namespace vfg {

struct BChunk {
  int p;
  int64_t start;
  int len;
};
constexpr int CH = 4096;  // vector elements per chunk
constexpr int BRB = 256;

enum { B_RZ = 0, B_PQ, B_ALPHA, B_BETA, B_RR, B_RR0, B_ACTIVE, B_STATUS, B_SCALE, B_ITS, B_NS = 16 };

__device__ __forceinline__ double bbsum(double v) {
  __shared__ double sh[BRB / 64];
  return block_sum<BRB>(v, sh);
}

template <typename T>
__global__ __launch_bounds__(BRB) void k_binit(const T *__restrict__ r, T *__restrict__ z, T *__restrict__ p,
                                               const T *__restrict__ dinv, const BChunk *__restrict__ ch,
                                               double *part) {
  const BChunk c = ch[blockIdx.x];
  double rz = 0., rr = 0.;
  for (int i = threadIdx.x; i < c.len; i += BRB) {
    int64_t k = c.start + i;
    double ri = r[k];
    T zi = (T)(dinv[k] * r[k]);
    z[k] = zi;
    p[k] = zi;
    rz += ri * (double)zi;
    rr += ri * ri;
  }
  rz = bbsum(rz);
  rr = bbsum(rr);
  if (threadIdx.x == 0) { part[2 * blockIdx.x] = rz; part[2 * blockIdx.x + 1] = rr; }
}

template <typename T>
__global__ __launch_bounds__(BRB) void k_bdot(const T *__restrict__ a, const T *__restrict__ b,
                                              const BChunk *__restrict__ ch, double *part) {
  const BChunk c = ch[blockIdx.x];
  double s = 0.;
  for (int i = threadIdx.x; i < c.len; i += BRB) s += (double)a[c.start + i] * (double)b[c.start + i];
  s = bbsum(s);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// per-member scalars; block p reduces its chunks [first, first+n)
// op 0: init (rz, rr, rr0, active); op 1: alpha; op 2: beta + convergence
__global__ __launch_bounds__(BRB) void k_bscalar(const double *part, const int *first, const int *nch, int op,
                                                 double *sc, double rtol, int maxit) {
  const int p = blockIdx.x;
  const int w = op == 1 ? 1 : 2;
  double a = 0., b = 0.;
  for (int i = threadIdx.x; i < nch[p]; i += BRB) {
    a += part[(int64_t)(first[p] + i) * w];
    if (w == 2) b += part[(int64_t)(first[p] + i) * w + 1];
  }
  a = bbsum(a);
  b = bbsum(b);
  if (threadIdx.x != 0) return;
  double *S = sc + (int64_t)p * B_NS;
  if (op == 0) {
    S[B_RZ] = a; S[B_RR] = b; S[B_RR0] = b; S[B_ITS] = 0.; S[B_STATUS] = 0.;
    S[B_ACTIVE] = (b > 0. && isfinite(b)) ? 1. : 0.;
    if (!isfinite(b)) S[B_STATUS] = (double)VFG_E_NAN;
    S[B_ALPHA] = 0.; S[B_BETA] = 0.;
  } else if (op == 1) {
    S[B_PQ] = a;
    if (S[B_ACTIVE] == 0.) { S[B_ALPHA] = 0.; return; }
    if (!(a > 0.) || !isfinite(a)) {
      S[B_ALPHA] = 0.; S[B_ACTIVE] = 0.;
      S[B_STATUS] = isfinite(a) ? (double)VFG_E_BREAKDOWN : (double)VFG_E_NAN;
    } else {
      S[B_ALPHA] = S[B_RZ] / a;
    }
  } else {
    if (S[B_ACTIVE] == 0.) { S[B_BETA] = 0.; return; }
    double rz = S[B_RZ];
    S[B_BETA] = rz != 0. ? a / rz : 0.;
    S[B_RZ] = a;
    S[B_RR] = b;
    S[B_ITS] += 1.;
    if (!isfinite(a) || !isfinite(b)) { S[B_ACTIVE] = 0.; S[B_STATUS] = (double)VFG_E_NAN; S[B_BETA] = 0.; }
    else if (b <= rtol * rtol * S[B_RR0]) { S[B_ACTIVE] = 0.; S[B_BETA] = 0.; }
    else if (S[B_ITS] >= maxit) { S[B_ACTIVE] = 0.; S[B_STATUS] = (double)VFG_E_MAXIT; S[B_BETA] = 0.; }
  }
}

template <typename T>
__global__ __launch_bounds__(BRB) void k_bxr(T *__restrict__ x, T *__restrict__ r, T *__restrict__ z,
                                             const T *__restrict__ p, const T *__restrict__ q,
                                             const T *__restrict__ dinv, const BChunk *__restrict__ ch,
                                             const double *sc, double *part) {
  const BChunk c = ch[blockIdx.x];
  const double al = sc[(int64_t)c.p * B_NS + B_ALPHA];
  double rz = 0., rr = 0.;
  for (int i = threadIdx.x; i < c.len; i += BRB) {
    int64_t k = c.start + i;
    x[k] = (T)((double)x[k] + al * (double)p[k]);
    T ri = (T)((double)r[k] - al * (double)q[k]);
    r[k] = ri;
    T zi = (T)(dinv[k] * ri);
    z[k] = zi;
    rz += (double)ri * (double)zi;
    rr += (double)ri * (double)ri;
  }
  rz = bbsum(rz);
  rr = bbsum(rr);
  if (threadIdx.x == 0) { part[2 * blockIdx.x] = rz; part[2 * blockIdx.x + 1] = rr; }
}

template <typename T>
__global__ __launch_bounds__(BRB) void k_bp(T *__restrict__ p, const T *__restrict__ z, const BChunk *__restrict__ ch,
                                            const double *sc) {
  const BChunk c = ch[blockIdx.x];
  const double be = sc[(int64_t)c.p * B_NS + B_BETA];
  for (int i = threadIdx.x; i < c.len; i += BRB) {
    int64_t k = c.start + i;
    p[k] = (T)((double)z[k] + be * (double)p[k]);
  }
}

// r32 = r64 * scale_p ; d32 = 0
__global__ __launch_bounds__(BRB) void k_b64to32(const double *r64, float *r32, float *d32,
                                                 const BChunk *__restrict__ ch, const double *scale) {
  const BChunk c = ch[blockIdx.x];
  const double s = scale[c.p];
  for (int i = threadIdx.x; i < c.len; i += BRB) {
    int64_t k = c.start + i;
    r32[k] = (float)(r64[k] * s);
    d32[k] = 0.f;
  }
}
// x64 += inv_scale_p * d32
__global__ __launch_bounds__(BRB) void k_baxpy(double *x, const float *d, const BChunk *__restrict__ ch,
                                               const double *inv_scale) {
  const BChunk c = ch[blockIdx.x];
  const double s = inv_scale[c.p];
  for (int i = threadIdx.x; i < c.len; i += BRB) x[c.start + i] += s * (double)d[c.start + i];
}
__global__ void k_bprep(double *R, double *D, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  R[i] = -R[i];
  D[i] = D[i] != 0. ? 1. / D[i] : 1.;
}
__global__ void k_bneg(double *R, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) R[i] = -R[i];
}
__global__ void k_bd2f(const double *D, float *F, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) F[i] = (float)D[i];
}
__global__ __launch_bounds__(BRB) void k_bmember_sum(const double *part, const int *first, const int *nch,
                                                     double *out) {
  const int p = blockIdx.x;
  double a = 0.;
  for (int i = threadIdx.x; i < nch[p]; i += BRB) a += part[first[p] + i];
  a = bbsum(a);
  if (threadIdx.x == 0) out[p] = a;
}
static unsigned bnb(int64_t n) { return (unsigned)((n + 255) / 256); }
void batch_prepare_vectors(double *R, double *D, int64_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_bprep, dim3(bnb(n)), dim3(256), 0, s, R, D, n);
}
void batch_dinv_to_f32(const double *D, float *D32, int64_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_bd2f, dim3(bnb(n)), dim3(256), 0, s, D, D32, n);
}
void batch_member_sums(const double *part, const int *first, const int *nch, int np, double *out, hipStream_t s) {
  hipLaunchKernelGGL(k_bmember_sum, dim3(np), dim3(BRB), 0, s, part, first, nch, out);
}
void batch_negate(double *R, int64_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_bneg, dim3(bnb(n)), dim3(256), 0, s, R, n);
}

}  // namespace vfg

// ------------------------------------------------------------------ host side
#include "vfg_batch_apply.h"

using namespace vfg;

struct vfg_batch {
  std::vector<vfg_handle *> m;
  int np = 0;
  hipStream_t stream = nullptr;
  std::string err;
  std::vector<int64_t> off;  // member offsets (doubles) of 3*nn vectors
  int64_t n3 = 0;
  std::vector<BItem> items;
  std::vector<int> oz;
  BItem *d_items = nullptr;
  int nitems = 0;
  BChunk *d_ch = nullptr;
  int nch = 0;
  int *d_first = nullptr, *d_nch = nullptr;
  double *d_part = nullptr, *d_sc = nullptr, *d_scale = nullptr, *d_iscale = nullptr;
  double *X = nullptr, *R = nullptr, *Z = nullptr, *Pv = nullptr, *Q = nullptr, *Dinv = nullptr, *B = nullptr;
  float *X32 = nullptr, *R32 = nullptr, *Z32 = nullptr, *P32 = nullptr, *Q32 = nullptr, *Dinv32 = nullptr;
  void *d_ops = nullptr, *d_ops32 = nullptr, *d_opsx = nullptr;
  int grid = 0;  // persistent grid size (0: one block per item)
  int graph_on = 1, graph_chunk = 10;
  hipGraphExec_t gexec = nullptr, gexec32 = nullptr;
  double g_rtol[2] = {-1., -1.};  // rtol / maxit baked into each captured graph
  int g_maxit[2] = {-1, -1};
  std::vector<void *> allocs;
};

static int bfail(vfg_batch *b, int code, const char *msg) {
  if (b) b->err = msg;
  return code;
}
static int balloc(vfg_batch *b, void **p, size_t bytes) {
  if (hipMalloc(p, bytes ? bytes : 8) != hipSuccess) {
    *p = nullptr;
    return bfail(b, VFG_E_NOMEM, "vfg_batch: device allocation failed");
  }
  b->allocs.push_back(*p);
  return VFG_OK;
}
#define BCHK(b, call)                                                   \
  do {                                                                  \
    hipError_t e_ = (call);                                             \
    if (e_ != hipSuccess) return bfail(b, VFG_E_HIP, hipGetErrorString(e_)); \
  } while (0)

static void build_items(vfg_batch *b, int slots, double oz_scale) {
  b->items.clear();
  b->oz.assign(b->np, 1);
  for (int p = 0; p < b->np; p++) {
    const Grid &G = b->m[p]->P.G;
    int tx = (G.nx + 14) / 15, ty = (G.ny + 14) / 15;
    int oz = choose_oz_pub(G, tx * ty, slots);
    oz = (int)(oz * oz_scale + .5);
    if (oz < 2) oz = 2;
    if (oz > G.nz) oz = G.nz;
    b->oz[p] = oz;
    int tz = (G.nz + oz - 1) / oz;
    for (int z = 0; z < tz; z++)
      for (int y = 0; y < ty; y++)
        for (int x = 0; x < tx; x++) b->items.push_back({p, x, y, z});
  }
  b->nitems = (int)b->items.size();
}

extern "C" {

const char *vfg_batch_last_error(const vfg_batch *b) { return b ? b->err.c_str() : "NULL batch"; }

int vfg_batch_create(vfg_handle *const *members, int n, vfg_batch **out) {
  if (!out || !members || n < 1) return VFG_E_ARG;
  *out = nullptr;
  vfg_batch *b = new vfg_batch();
  b->np = n;
  for (int i = 0; i < n; i++) {
    if (!members[i]) { delete b; return VFG_E_ARG; }
    b->m.push_back(members[i]);
  }
  if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) { delete b; return VFG_E_HIP; }
  // layout of concatenated vectors and reduction chunks
  std::vector<BChunk> ch;
  std::vector<int> first(n), nch(n);
  int64_t o = 0;
  for (int p = 0; p < n; p++) {
    b->off.push_back(o);
    int64_t len = 3 * members[p]->P.nn;
    first[p] = (int)ch.size();
    for (int64_t s = 0; s < len; s += CH) ch.push_back({p, o + s, (int)(len - s < CH ? len - s : CH)});
    nch[p] = (int)ch.size() - first[p];
    o += len;
  }
  b->n3 = o;
  b->nch = (int)ch.size();
  int rc = 0;
  const size_t n3 = (size_t)b->n3;
  if ((rc = balloc(b, (void **)&b->d_ch, sizeof(BChunk) * ch.size())) ||
      (rc = balloc(b, (void **)&b->d_first, sizeof(int) * n)) || (rc = balloc(b, (void **)&b->d_nch, sizeof(int) * n)) ||
      (rc = balloc(b, (void **)&b->d_part, 16 * ch.size())) || (rc = balloc(b, (void **)&b->d_sc, 8 * B_NS * n)) ||
      (rc = balloc(b, (void **)&b->d_scale, 8 * n)) || (rc = balloc(b, (void **)&b->d_iscale, 8 * n)) ||
      (rc = balloc(b, (void **)&b->X, 8 * n3)) || (rc = balloc(b, (void **)&b->R, 8 * n3)) ||
      (rc = balloc(b, (void **)&b->Z, 8 * n3)) || (rc = balloc(b, (void **)&b->Pv, 8 * n3)) ||
      (rc = balloc(b, (void **)&b->Q, 8 * n3)) || (rc = balloc(b, (void **)&b->Dinv, 8 * n3)) ||
      (rc = balloc(b, (void **)&b->B, 8 * n3)) || (rc = balloc(b, (void **)&b->X32, 4 * n3)) ||
      (rc = balloc(b, (void **)&b->R32, 4 * n3)) || (rc = balloc(b, (void **)&b->Z32, 4 * n3)) ||
      (rc = balloc(b, (void **)&b->P32, 4 * n3)) || (rc = balloc(b, (void **)&b->Q32, 4 * n3)) ||
      (rc = balloc(b, (void **)&b->Dinv32, 4 * n3)) ||
      (rc = balloc(b, &b->d_ops, batch_op_bytes(0) * n)) || (rc = balloc(b, &b->d_ops32, batch_op_bytes(1) * n)) ||
      (rc = balloc(b, &b->d_opsx, batch_op_bytes(0) * n))) {
    vfg_batch_destroy(b);
    return rc;
  }
  hipMemcpy(b->d_ch, ch.data(), sizeof(BChunk) * ch.size(), hipMemcpyHostToDevice);
  hipMemcpy(b->d_first, first.data(), sizeof(int) * n, hipMemcpyHostToDevice);
  hipMemcpy(b->d_nch, nch.data(), sizeof(int) * n, hipMemcpyHostToDevice);
  int slots = batch_slots();
  build_items(b, slots, 1.0);
  if ((rc = balloc(b, (void **)&b->d_items, sizeof(BItem) * b->items.size()))) {
    vfg_batch_destroy(b);
    return rc;
  }
  hipMemcpy(b->d_items, b->items.data(), sizeof(BItem) * b->items.size(), hipMemcpyHostToDevice);
  b->grid = 0;
  *out = b;
  return VFG_OK;
}

int vfg_batch_destroy(vfg_batch *b) {
  if (!b) return VFG_OK;
  if (b->stream) hipStreamSynchronize(b->stream);
  if (b->gexec) hipGraphExecDestroy(b->gexec);
  if (b->gexec32) hipGraphExecDestroy(b->gexec32);
  for (void *p : b->allocs) hipFree(p);
  if (b->stream) hipStreamDestroy(b->stream);
  delete b;
  return VFG_OK;
}

int vfg_batch_size(const vfg_batch *b) { return b ? b->np : -1; }

// persistent apply over arbitrary per-member device vectors
int vfg_batch_u_apply(vfg_batch *b, const double *const *x, double *const *y) {
  if (!b || !x || !y) return VFG_E_ARG;
  std::vector<const void *> xs(x, x + b->np), ys(y, y + b->np);
  fill_ops(b->m, xs.data(), (void *const *)ys.data(), 0, b->oz, b->d_opsx, b->stream);
  launch_batch_apply(0, b->d_opsx, b->d_items, b->nitems, b->grid, b->stream);
  BCHK(b, hipStreamSynchronize(b->stream));
  BCHK(b, hipGetLastError());
  return VFG_OK;
}

int vfg_batch_u_apply_f32(vfg_batch *b, const float *const *x, float *const *y) {
  if (!b || !x || !y) return VFG_E_ARG;
  std::vector<const void *> xs(x, x + b->np), ys(y, y + b->np);
  fill_ops(b->m, xs.data(), (void *const *)ys.data(), 1, b->oz, b->d_ops32, b->stream);
  launch_batch_apply(1, b->d_ops32, b->d_items, b->nitems, b->grid, b->stream);
  BCHK(b, hipStreamSynchronize(b->stream));
  BCHK(b, hipGetLastError());
  return VFG_OK;
}

int vfg_batch_time_u_apply(vfg_batch *b, int fp32, int nrep, double *ms) {
  if (!b || nrep < 1 || !ms) return VFG_E_ARG;
  std::vector<const void *> xs(b->np);
  std::vector<void *> ys(b->np);
  for (int p = 0; p < b->np; p++) {
    xs[p] = fp32 ? (const void *)(b->P32 + b->off[p]) : (const void *)(b->Pv + b->off[p]);
    ys[p] = fp32 ? (void *)(b->Q32 + b->off[p]) : (void *)(b->Q + b->off[p]);
  }
  void *ops = fp32 ? b->d_ops32 : b->d_opsx;
  fill_ops(b->m, xs.data(), ys.data(), fp32, b->oz, ops, b->stream);
  hipEvent_t e0, e1;
  hipEventCreate(&e0);
  hipEventCreate(&e1);
  hipEventRecord(e0, b->stream);
  for (int i = 0; i < nrep; i++) launch_batch_apply(fp32, ops, b->d_items, b->nitems, b->grid, b->stream);
  hipEventRecord(e1, b->stream);
  hipEventSynchronize(e1);
  float t = 0.f;
  hipEventElapsedTime(&t, e0, e1);
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  *ms = t;
  BCHK(b, hipGetLastError());
  return VFG_OK;
}

// Autotuner: times the persistent apply for a few (z-chunk scale, grid) pairs
// and keeps the fastest; returns the chosen grid (0 = one block per item).
int vfg_batch_autotune(vfg_batch *b, int fp32, double *best_ms) {
  if (!b) return VFG_E_ARG;
  const int slots = batch_slots();
  const double scales[3] = {0.5, 1.0, 2.0};
  const int grids[3] = {0, slots, 2 * slots};
  double best = 1e300;
  double bscale = 1.0;
  int bgrid = 0;
  for (double sc : scales) {
    build_items(b, slots, sc);
    hipFree(b->d_items);
    for (auto &p : b->allocs)
      if (p == b->d_items) p = nullptr;
    if (balloc(b, (void **)&b->d_items, sizeof(BItem) * b->items.size())) return VFG_E_NOMEM;
    hipMemcpy(b->d_items, b->items.data(), sizeof(BItem) * b->items.size(), hipMemcpyHostToDevice);
    for (int g : grids) {
      b->grid = g;
      double ms = 0.;
      vfg_batch_time_u_apply(b, fp32, 2, &ms);
      vfg_batch_time_u_apply(b, fp32, 5, &ms);
      if (ms < best) { best = ms; bscale = sc; bgrid = g; }
    }
  }
  build_items(b, slots, bscale);
  hipFree(b->d_items);
  for (auto &p : b->allocs)
    if (p == b->d_items) p = nullptr;
  if (balloc(b, (void **)&b->d_items, sizeof(BItem) * b->items.size())) return VFG_E_NOMEM;
  hipMemcpy(b->d_items, b->items.data(), sizeof(BItem) * b->items.size(), hipMemcpyHostToDevice);
  b->grid = bgrid;
  if (b->gexec) { hipGraphExecDestroy(b->gexec); b->gexec = nullptr; }
  if (b->gexec32) { hipGraphExecDestroy(b->gexec32); b->gexec32 = nullptr; }
  if (best_ms) *best_ms = best / 5.;
  return VFG_OK;
}

int vfg_batch_set_graph(vfg_batch *b, int enable, int chunk) {
  if (!b || chunk < 1) return VFG_E_ARG;
  b->graph_on = enable;
  b->graph_chunk = chunk;
  if (b->gexec) { hipGraphExecDestroy(b->gexec); b->gexec = nullptr; }
  if (b->gexec32) { hipGraphExecDestroy(b->gexec32); b->gexec32 = nullptr; }
  return VFG_OK;
}

}  // extern "C"

// one CG iteration of the whole ensemble on vectors of type T
template <typename T>
static void cg_iter(vfg_batch *b, T *x, T *r, T *z, T *p, T *q, const T *dinv, void *ops, int fp32, double rtol,
                    int maxit) {
  hipStream_t s = b->stream;
  launch_batch_apply(fp32, ops, b->d_items, b->nitems, b->grid, s);
  hipLaunchKernelGGL(k_bdot<T>, dim3(b->nch), dim3(BRB), 0, s, (const T *)p, (const T *)q, (const BChunk *)b->d_ch,
                     b->d_part);
  hipLaunchKernelGGL(k_bscalar, dim3(b->np), dim3(BRB), 0, s, (const double *)b->d_part, (const int *)b->d_first,
                     (const int *)b->d_nch, 1, b->d_sc, rtol, maxit);
  hipLaunchKernelGGL(k_bxr<T>, dim3(b->nch), dim3(BRB), 0, s, x, r, z, (const T *)p, (const T *)q, dinv,
                     (const BChunk *)b->d_ch, (const double *)b->d_sc, b->d_part);
  hipLaunchKernelGGL(k_bscalar, dim3(b->np), dim3(BRB), 0, s, (const double *)b->d_part, (const int *)b->d_first,
                     (const int *)b->d_nch, 2, b->d_sc, rtol, maxit);
  hipLaunchKernelGGL(k_bp<T>, dim3(b->nch), dim3(BRB), 0, s, p, (const T *)z, (const BChunk *)b->d_ch,
                     (const double *)b->d_sc);
}

// runs CG on all members to rtol; the iterate state (x, r) is prepared by the caller
template <typename T>
static int cg_run(vfg_batch *b, T *x, T *r, T *z, T *p, T *q, const T *dinv, void *ops, int fp32, double rtol,
                  int maxit, hipGraphExec_t *gexec, std::vector<double> &S) {
  hipStream_t s = b->stream;
  hipLaunchKernelGGL(k_binit<T>, dim3(b->nch), dim3(BRB), 0, s, (const T *)r, z, p, dinv, (const BChunk *)b->d_ch,
                     b->d_part);
  hipLaunchKernelGGL(k_bscalar, dim3(b->np), dim3(BRB), 0, s, (const double *)b->d_part, (const int *)b->d_first,
                     (const int *)b->d_nch, 0, b->d_sc, rtol, maxit);
  const int K = b->graph_chunk;
  const int gi = fp32 ? 1 : 0;
  if (b->graph_on && *gexec && (b->g_rtol[gi] != rtol || b->g_maxit[gi] != maxit)) {
    hipGraphExecDestroy(*gexec);
    *gexec = nullptr;
  }
  if (b->graph_on && !*gexec) {
    b->g_rtol[gi] = rtol;
    b->g_maxit[gi] = maxit;
    hipGraph_t g;
    BCHK(b, hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    for (int k = 0; k < K; k++) cg_iter<T>(b, x, r, z, p, q, dinv, ops, fp32, rtol, maxit);
    BCHK(b, hipStreamEndCapture(s, &g));
    hipError_t e = hipGraphInstantiate(gexec, g, nullptr, nullptr, 0);
    hipGraphDestroy(g);
    if (e != hipSuccess) return bfail(b, VFG_E_HIP, "hipGraphInstantiate failed");
  }
  S.assign((size_t)b->np * B_NS, 0.);
  for (int done = 0; done < maxit + K; done += K) {
    if (b->graph_on) BCHK(b, hipGraphLaunch(*gexec, s));
    else
      for (int k = 0; k < K; k++) cg_iter<T>(b, x, r, z, p, q, dinv, ops, fp32, rtol, maxit);
    BCHK(b, hipMemcpyAsync(S.data(), b->d_sc, 8 * S.size(), hipMemcpyDeviceToHost, s));
    BCHK(b, hipStreamSynchronize(s));
    bool any = false;
    for (int m = 0; m < b->np; m++) any |= S[(size_t)m * B_NS + B_ACTIVE] != 0.;
    if (!any) break;
  }
  return VFG_OK;
}

extern "C" int vfg_batch_u_solve(vfg_batch *b, double *const *U, const double *const *BCU, const vfg_solve_opts *o,
                                 vfg_solve_info *info, int loc) {
  if (!b || !U || !BCU) return VFG_E_ARG;
  vfg_solve_opts d;
  vfg_solve_opts_default(&d);
  if (!o) o = &d;
  hipStream_t s = b->stream;
  // per member: x = U with Dirichlet values, r = -F(x), dinv = 1/diag
  for (int p = 0; p < b->np; p++) {
    vfg_handle *h = b->m[p];
    const int64_t n = 3 * h->P.nn;
    double *x = b->X + b->off[p], *r = b->R + b->off[p], *B = b->B + b->off[p], *dv = b->Dinv + b->off[p];
    hipMemcpyKind k = loc == VFG_HOST ? hipMemcpyHostToDevice : hipMemcpyDeviceToDevice;
    BCHK(b, hipMemcpyAsync(x, U[p], 8 * n, k, s));
    BCHK(b, hipMemcpyAsync(B, BCU[p], 8 * n, k, s));
    launch_vec_apply_dirichlet(h->P, 0, x, B, s);
    launch_u_rhs(h->P, x, B, r, 1, s);
    launch_u_diag(h->P, dv, s);
  }
  batch_prepare_vectors(b->R, b->Dinv, b->n3, s);  // r = -F, dinv = 1/diag
  std::vector<const void *> xs(b->np), xs32(b->np);
  std::vector<void *> ys(b->np), ys32(b->np);
  for (int p = 0; p < b->np; p++) {
    xs[p] = b->Pv + b->off[p];
    ys[p] = b->Q + b->off[p];
    xs32[p] = b->P32 + b->off[p];
    ys32[p] = b->Q32 + b->off[p];
  }
  std::vector<int> status(b->np, 0), iters(b->np, 0), outer(b->np, 0);
  std::vector<double> relres(b->np, 0.), S;
  int rc = VFG_OK;
  if (o->precision != VFG_PREC_MIXED) {
    fill_ops(b->m, xs.data(), ys.data(), 0, b->oz, b->d_ops, s);
    rc = cg_run<double>(b, b->X, b->R, b->Z, b->Pv, b->Q, b->Dinv, b->d_ops, 0, o->rtol, o->maxit, &b->gexec, S);
    for (int p = 0; p < b->np && !rc; p++) {
      const double *Sp = &S[(size_t)p * B_NS];
      iters[p] = (int)Sp[B_ITS];
      status[p] = (int)Sp[B_STATUS];
      relres[p] = Sp[B_RR0] > 0 ? sqrt(Sp[B_RR] / Sp[B_RR0]) : 0.;
      outer[p] = 1;
    }
  } else {
    // FP64 refinement around the FP32 inner CG
    fill_ops(b->m, xs32.data(), ys32.data(), 1, b->oz, b->d_ops32, s);
    batch_dinv_to_f32(b->Dinv, b->Dinv32, b->n3, s);
    std::vector<double> rn0(b->np, -1.), scale(b->np), iscale(b->np), nrm(b->np);
    std::vector<int> done(b->np, 0);
    const double inner_rtol = 1e-4;
    for (int it = 0; it < 60 && !rc; it++) {
      // per-member FP64 residual norms
      hipLaunchKernelGGL(k_bdot<double>, dim3(b->nch), dim3(BRB), 0, s, (const double *)b->R, (const double *)b->R,
                         (const BChunk *)b->d_ch, b->d_part);
      batch_member_sums(b->d_part, b->d_first, b->d_nch, b->np, b->d_scale, s);
      BCHK(b, hipMemcpyAsync(nrm.data(), b->d_scale, 8 * b->np, hipMemcpyDeviceToHost, s));
      BCHK(b, hipStreamSynchronize(s));
      bool any = false;
      for (int p = 0; p < b->np; p++) {
        double rn = sqrt(nrm[p]);
        if (rn0[p] < 0.) rn0[p] = rn;
        relres[p] = rn0[p] > 0. ? rn / rn0[p] : 0.;
        if (!isfinite(rn)) { status[p] = VFG_E_NAN; done[p] = 1; }
        if (!done[p] && (rn <= fmax(o->rtol * rn0[p], o->atol) || rn == 0.)) done[p] = 1;
        if (!done[p]) { outer[p] = it + 1; any = true; }
        scale[p] = done[p] ? 0. : 1. / rn;
        iscale[p] = done[p] ? 0. : rn;
      }
      if (!any) break;
      BCHK(b, hipMemcpyAsync(b->d_scale, scale.data(), 8 * b->np, hipMemcpyHostToDevice, s));
      BCHK(b, hipMemcpyAsync(b->d_iscale, iscale.data(), 8 * b->np, hipMemcpyHostToDevice, s));
      hipLaunchKernelGGL(k_b64to32, dim3(b->nch), dim3(BRB), 0, s, (const double *)b->R, b->R32, b->X32,
                         (const BChunk *)b->d_ch, (const double *)b->d_scale);
      rc = cg_run<float>(b, b->X32, b->R32, b->Z32, b->P32, b->Q32, b->Dinv32, b->d_ops32, 1, inner_rtol, o->maxit,
                         &b->gexec32, S);
      for (int p = 0; p < b->np; p++) {
        iters[p] += (int)S[(size_t)p * B_NS + B_ITS];
        int st = (int)S[(size_t)p * B_NS + B_STATUS];
        if (st == VFG_E_NAN) { status[p] = st; done[p] = 1; }
      }
      hipLaunchKernelGGL(k_baxpy, dim3(b->nch), dim3(BRB), 0, s, b->X, (const float *)b->X32, (const BChunk *)b->d_ch,
                         (const double *)b->d_iscale);
      for (int p = 0; p < b->np; p++) {
        if (done[p]) continue;
        vfg_handle *h = b->m[p];
        launch_u_rhs(h->P, b->X + b->off[p], b->B + b->off[p], b->R + b->off[p], 1, s);
      }
      batch_negate(b->R, b->n3, s);
      if (it == 59)
        for (int p = 0; p < b->np; p++)
          if (!done[p]) status[p] = VFG_E_MAXIT;
    }
  }
  if (rc) return rc;
  for (int p = 0; p < b->np; p++) {
    hipMemcpyKind k = loc == VFG_HOST ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice;
    BCHK(b, hipMemcpyAsync(U[p], b->X + b->off[p], 8 * 3 * b->m[p]->P.nn, k, s));
    if (info) {
      info[p].iterations = iters[p];
      info[p].outer = outer[p];
      info[p].relres = relres[p];
      info[p].status = status[p];
    }
  }
  BCHK(b, hipStreamSynchronize(s));
  BCHK(b, hipGetLastError());
  for (int p = 0; p < b->np; p++)
    if (status[p]) return bfail(b, status[p], "vfg_batch_u_solve: a member did not converge (see info[].status)");
  return VFG_OK;
}
