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
  int p;          // member
  int64_t start;  // offset inside the member's vectors
  int len;
};
// per-member vector table: slots 0..6 FP64 (X,R,Z,P,Q,Dinv,B), 7..12 FP32
// (X,R,Z,P,Q,Dinv); the vectors belong to the member handle
enum { VX = 0, VR, VZ, VP, VQ, VD, VB, VX32, VR32, VZ32, VP32, VQ32, VD32, NV };
struct BVec {
  void *v[NV];
};
template <typename T>
__device__ __forceinline__ T *vec(const BVec *t, const BChunk &c, int slot) {
  return (T *)t[c.p].v[slot] + c.start;
}
constexpr int CH = 4096;  // vector elements per chunk
constexpr int BRB = 256;

enum { B_RZ = 0, B_PQ, B_ALPHA, B_BETA, B_RR, B_RR0, B_ACTIVE, B_STATUS, B_SCALE, B_ITS, B_NS = 16 };

__device__ __forceinline__ double bbsum(double v) {
  __shared__ double sh[BRB / 64];
  return block_sum<BRB>(v, sh);
}

// slot offset o = 0 (FP64 set) or VX32 (FP32 set)
template <typename T>
__global__ __launch_bounds__(BRB) void k_binit(const BVec *vt, int o, const BChunk *__restrict__ ch, double *part) {
  const BChunk c = ch[blockIdx.x];
  const T *r = vec<T>(vt, c, o + VR), *dinv = vec<T>(vt, c, o + VD);
  T *z = vec<T>(vt, c, o + VZ), *p = vec<T>(vt, c, o + VP);
  double rz = 0., rr = 0.;
  for (int k = threadIdx.x; k < c.len; k += BRB) {
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

// sc (optional): per-member scalars; chunks of frozen members are skipped
template <typename T>
__global__ __launch_bounds__(BRB) void k_bdot(const BVec *vt, int sa, int sb, const BChunk *__restrict__ ch,
                                              const double *sc, double *part) {
  const BChunk c = ch[blockIdx.x];
  if (sc && sc[(int64_t)c.p * B_NS + B_ACTIVE] == 0.) return;  // frozen member
  const T *a = vec<T>(vt, c, sa), *b = vec<T>(vt, c, sb);
  double s = 0.;
  for (int i = threadIdx.x; i < c.len; i += BRB) s += (double)a[i] * (double)b[i];
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
__global__ __launch_bounds__(BRB) void k_bxr(const BVec *vt, int o, const BChunk *__restrict__ ch, const double *sc,
                                             double *part) {
  const BChunk c = ch[blockIdx.x];
  if (sc[(int64_t)c.p * B_NS + B_ACTIVE] == 0.) return;  // frozen member
  T *x = vec<T>(vt, c, o + VX), *r = vec<T>(vt, c, o + VR), *z = vec<T>(vt, c, o + VZ);
  const T *p = vec<T>(vt, c, o + VP), *q = vec<T>(vt, c, o + VQ), *dinv = vec<T>(vt, c, o + VD);
  const double al = sc[(int64_t)c.p * B_NS + B_ALPHA];
  double rz = 0., rr = 0.;
  for (int k = threadIdx.x; k < c.len; k += BRB) {
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

// multigrid CG: x += alpha p, r -= alpha q (z comes from the V-cycle)
template <typename T>
__global__ __launch_bounds__(BRB) void k_bxr_mg(const BVec *vt, int o, const BChunk *__restrict__ ch,
                                                const double *sc) {
  const BChunk c = ch[blockIdx.x];
  if (sc[(int64_t)c.p * B_NS + B_ACTIVE] == 0.) return;  // frozen member
  T *x = vec<T>(vt, c, o + VX), *r = vec<T>(vt, c, o + VR);
  const T *p = vec<T>(vt, c, o + VP), *q = vec<T>(vt, c, o + VQ);
  const double al = sc[(int64_t)c.p * B_NS + B_ALPHA];
  for (int k = threadIdx.x; k < c.len; k += BRB) {
    x[k] = (T)((double)x[k] + al * (double)p[k]);
    r[k] = (T)((double)r[k] - al * (double)q[k]);
  }
}
// multigrid CG: partial sums of r.z and r.r (and p = z at the start); sc
// (optional) skips frozen members
template <typename T>
__global__ __launch_bounds__(BRB) void k_brz(const BVec *vt, int o, const BChunk *__restrict__ ch, const double *sc,
                                             int setp, double *part) {
  const BChunk c = ch[blockIdx.x];
  if (sc && sc[(int64_t)c.p * B_NS + B_ACTIVE] == 0.) return;
  const T *r = vec<T>(vt, c, o + VR), *z = vec<T>(vt, c, o + VZ);
  T *p = vec<T>(vt, c, o + VP);
  double rz = 0., rr = 0.;
  for (int k = threadIdx.x; k < c.len; k += BRB) {
    const double ri = r[k];
    rz += ri * (double)z[k];
    rr += ri * ri;
    if (setp) p[k] = z[k];
  }
  rz = bbsum(rz);
  rr = bbsum(rr);
  if (threadIdx.x == 0) { part[2 * blockIdx.x] = rz; part[2 * blockIdx.x + 1] = rr; }
}

template <typename T>
__global__ __launch_bounds__(BRB) void k_bp(const BVec *vt, int o, const BChunk *__restrict__ ch, const double *sc) {
  const BChunk c = ch[blockIdx.x];
  if (sc[(int64_t)c.p * B_NS + B_ACTIVE] == 0.) return;  // frozen member
  T *p = vec<T>(vt, c, o + VP);
  const T *z = vec<T>(vt, c, o + VZ);
  const double be = sc[(int64_t)c.p * B_NS + B_BETA];
  for (int k = threadIdx.x; k < c.len; k += BRB) p[k] = (T)((double)z[k] + be * (double)p[k]);
}

// FP64 CG under the FP32 V-cycle (the single-problem default, batched):
// x += alpha p, r -= alpha q in FP64, r32 = r for the V-cycle
__global__ __launch_bounds__(BRB) void k_bxr_mg64(const BVec *vt, const BChunk *__restrict__ ch, const double *sc) {
  const BChunk c = ch[blockIdx.x];
  if (sc[(int64_t)c.p * B_NS + B_ACTIVE] == 0.) return;  // frozen member
  double *x = vec<double>(vt, c, VX), *r = vec<double>(vt, c, VR);
  const double *p = vec<double>(vt, c, VP), *q = vec<double>(vt, c, VQ);
  float *r32 = vec<float>(vt, c, VR32);
  const double al = sc[(int64_t)c.p * B_NS + B_ALPHA];
  for (int k = threadIdx.x; k < c.len; k += BRB) {
    x[k] += al * p[k];
    const double ri = r[k] - al * q[k];
    r[k] = ri;
    r32[k] = (float)ri;
  }
}
// r32 = r (before the first V-cycle)
__global__ __launch_bounds__(BRB) void k_br32(const BVec *vt, const BChunk *__restrict__ ch) {
  const BChunk c = ch[blockIdx.x];
  const double *r = vec<double>(vt, c, VR);
  float *r32 = vec<float>(vt, c, VR32);
  for (int k = threadIdx.x; k < c.len; k += BRB) r32[k] = (float)r[k];
}
// partial sums of r.z and r.r with z the FP32 V-cycle output (and p = z at the start)
__global__ __launch_bounds__(BRB) void k_brz64(const BVec *vt, const BChunk *__restrict__ ch, const double *sc, int setp,
                                               double *part) {
  const BChunk c = ch[blockIdx.x];
  if (sc && sc[(int64_t)c.p * B_NS + B_ACTIVE] == 0.) return;
  const double *r = vec<double>(vt, c, VR);
  const float *z = vec<float>(vt, c, VZ32);
  double *p = vec<double>(vt, c, VP);
  double rz = 0., rr = 0.;
  for (int k = threadIdx.x; k < c.len; k += BRB) {
    const double ri = r[k], zi = (double)z[k];
    rz += ri * zi;
    rr += ri * ri;
    if (setp) p[k] = zi;
  }
  rz = bbsum(rz);
  rr = bbsum(rr);
  if (threadIdx.x == 0) { part[2 * blockIdx.x] = rz; part[2 * blockIdx.x + 1] = rr; }
}
// p = z + beta p with z the FP32 V-cycle output
__global__ __launch_bounds__(BRB) void k_bp64(const BVec *vt, const BChunk *__restrict__ ch, const double *sc) {
  const BChunk c = ch[blockIdx.x];
  if (sc[(int64_t)c.p * B_NS + B_ACTIVE] == 0.) return;  // frozen member
  double *p = vec<double>(vt, c, VP);
  const float *z = vec<float>(vt, c, VZ32);
  const double be = sc[(int64_t)c.p * B_NS + B_BETA];
  for (int k = threadIdx.x; k < c.len; k += BRB) p[k] = (double)z[k] + be * p[k];
}

// R32 = R * scale_p ; X32 = 0
__global__ __launch_bounds__(BRB) void k_b64to32(const BVec *vt, const BChunk *__restrict__ ch, const double *scale) {
  const BChunk c = ch[blockIdx.x];
  const double *r64 = vec<double>(vt, c, VR);
  float *r32 = vec<float>(vt, c, VR32), *d32 = vec<float>(vt, c, VX32);
  const double s = scale[c.p];
  for (int k = threadIdx.x; k < c.len; k += BRB) {
    r32[k] = (float)(r64[k] * s);
    d32[k] = 0.f;
  }
}
// X += inv_scale_p * X32
__global__ __launch_bounds__(BRB) void k_baxpy(const BVec *vt, const BChunk *__restrict__ ch, const double *inv_scale) {
  const BChunk c = ch[blockIdx.x];
  double *x = vec<double>(vt, c, VX);
  const float *d = vec<float>(vt, c, VX32);
  const double s = inv_scale[c.p];
  for (int k = threadIdx.x; k < c.len; k += BRB) x[k] += s * (double)d[k];
}
// R = -R ; Dinv = 1/Dinv ; Dinv32 = Dinv
__global__ __launch_bounds__(BRB) void k_bprep(const BVec *vt, const BChunk *__restrict__ ch) {
  const BChunk c = ch[blockIdx.x];
  double *R = vec<double>(vt, c, VR), *D = vec<double>(vt, c, VD);
  float *D32 = vec<float>(vt, c, VD32);
  for (int k = threadIdx.x; k < c.len; k += BRB) {
    R[k] = -R[k];
    double di = D[k] != 0. ? 1. / D[k] : 1.;
    D[k] = di;
    D32[k] = (float)di;
  }
}
__global__ __launch_bounds__(BRB) void k_bneg(const BVec *vt, const BChunk *__restrict__ ch) {
  const BChunk c = ch[blockIdx.x];
  double *R = vec<double>(vt, c, VR);
  for (int k = threadIdx.x; k < c.len; k += BRB) R[k] = -R[k];
}
__global__ void k_bset_active(double *sc) { sc[(int64_t)blockIdx.x * B_NS + B_ACTIVE] = 1.; }

__global__ __launch_bounds__(BRB) void k_bmember_sum(const double *part, const int *first, const int *nch,
                                                     double *out) {
  const int p = blockIdx.x;
  double a = 0.;
  for (int i = threadIdx.x; i < nch[p]; i += BRB) a += part[first[p] + i];
  a = bbsum(a);
  if (threadIdx.x == 0) out[p] = a;
}
void batch_member_sums(const double *part, const int *first, const int *nch, int np, double *out, hipStream_t s) {
  hipLaunchKernelGGL(k_bmember_sum, dim3(np), dim3(BRB), 0, s, part, first, nch, out);
}

}  // namespace vfg

// ------------------------------------------------------------------ host side
#include "vfg_batch_apply.h"
#include "vfg_bmg.h"

using namespace vfg;

struct vfg_batch {
  std::vector<vfg_handle *> m;
  int np = 0;
  hipStream_t stream = nullptr;
  std::string err;
  std::vector<BItem> items;
  std::vector<int> oz;
  BItem *d_items = nullptr;
  int nitems = 0;
  int *d_nitems = nullptr;  // nitems on the device (the flagged persistent apply reads it)
  BChunk *d_ch = nullptr;
  int nch = 0;
  BVec *d_vt = nullptr;  // per-member vector table (member work buffers)
  std::vector<BVec> vt;
  int *d_first = nullptr, *d_nch = nullptr;
  double *d_part = nullptr, *d_sc = nullptr, *d_scale = nullptr, *d_iscale = nullptr;
  void *d_ops = nullptr, *d_ops32 = nullptr, *d_opsx = nullptr;
  int grid = 0;  // persistent grid size (0: one block per item)
  int sched = 0;  // persistent-apply schedule (batch_tile, BATCH_SCHEDS)
  double oz_scale = 1.;
  std::vector<vfg_tune_point> tune;  // the last autotuning's measurements
  int graph_on = 1, graph_chunk = 10;
  hipGraphExec_t gexec = nullptr, gexec32 = nullptr;
  double g_rtol[2] = {-1., -1.};  // rtol / maxit / preconditioner baked into each captured graph
  int g_maxit[2] = {-1, -1}, g_mg[2] = {-1, -1};
  BatchMG *bmg = nullptr;  // members' multigrid (built at the first VFG_PC_MG solve)
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
static void bfree(vfg_batch *b, void *p) {
  if (!p) return;
  hipFree(p);
  for (auto &q : b->allocs)
    if (q == p) q = nullptr;
}
#define BCHK(b, call)                                                   \
  do {                                                                  \
    hipError_t e_ = (call);                                             \
    if (e_ != hipSuccess) return bfail(b, VFG_E_HIP, hipGetErrorString(e_)); \
  } while (0)

static void drop_graphs(vfg_batch *b) {
  if (b->gexec) { hipGraphExecDestroy(b->gexec); b->gexec = nullptr; }
  if (b->gexec32) { hipGraphExecDestroy(b->gexec32); b->gexec32 = nullptr; }
}

static int build_items(vfg_batch *b, int slots, double oz_scale) {
  b->items.clear();
  b->oz.assign(b->np, 1);
  b->oz_scale = oz_scale;
  const int ow = batch_tile(b->sched) - 1;  // owned node columns per tile
  for (int p = 0; p < b->np; p++) {
    const Grid &G = b->m[p]->P.G;
    int tx = (G.nx + ow - 1) / ow, ty = (G.ny + ow - 1) / ow;
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
  bfree(b, b->d_items);
  b->d_items = nullptr;
  int rc = balloc(b, (void **)&b->d_items, sizeof(BItem) * b->items.size());
  if (rc) return rc;
  BCHK(b, hipMemcpy(b->d_items, b->items.data(), sizeof(BItem) * b->items.size(), hipMemcpyHostToDevice));
  if (!b->d_nitems && (rc = balloc(b, (void **)&b->d_nitems, sizeof(int)))) return rc;
  BCHK(b, hipMemcpy(b->d_nitems, &b->nitems, sizeof(int), hipMemcpyHostToDevice));
  drop_graphs(b);  // captured graphs hold the old item list
  return VFG_OK;
}

extern "C" {

const char *vfg_batch_last_error(const vfg_batch *b) { return b ? b->err.c_str() : "NULL batch"; }

int vfg_batch_create(vfg_handle *const *members, int n, vfg_batch **out) {
  if (!out || !members || n < 1) return VFG_E_ARG;
  *out = nullptr;
  for (int i = 0; i < n; i++)
    if (!members[i] || members[i]->P.unilateral) return VFG_E_ARG;  // linear u-problems only
  vfg_batch *b = new vfg_batch();
  b->np = n;
  b->m.assign(members, members + n);
  if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) {
    delete b;
    return VFG_E_HIP;
  }
  int rc = VFG_OK;
  // member work vectors and the pointer table
  b->vt.resize(n);
  for (int p = 0; p < n && !rc; p++) {
    vfg_handle *h = members[p];
    if ((rc = vfg_work(h, 7, 6))) {
      bfail(b, rc, h->err.c_str());
      break;
    }
    for (int i = 0; i < 7; i++) b->vt[p].v[VX + i] = h->w[i];
    for (int i = 0; i < 6; i++) b->vt[p].v[VX32 + i] = h->wf[i];
  }
  // reduction chunks: never straddle two members
  std::vector<BChunk> ch;
  std::vector<int> first(n), nch(n);
  for (int p = 0; p < n; p++) {
    int64_t len = 3 * members[p]->P.nn;
    first[p] = (int)ch.size();
    for (int64_t s = 0; s < len; s += CH) ch.push_back({p, s, (int)(len - s < CH ? len - s : CH)});
    nch[p] = (int)ch.size() - first[p];
  }
  b->nch = (int)ch.size();
  if (rc || (rc = balloc(b, (void **)&b->d_ch, sizeof(BChunk) * ch.size())) ||
      (rc = balloc(b, (void **)&b->d_vt, sizeof(BVec) * n)) ||
      (rc = balloc(b, (void **)&b->d_first, sizeof(int) * n)) || (rc = balloc(b, (void **)&b->d_nch, sizeof(int) * n)) ||
      (rc = balloc(b, (void **)&b->d_part, 16 * ch.size())) || (rc = balloc(b, (void **)&b->d_sc, 8 * B_NS * n)) ||
      (rc = balloc(b, (void **)&b->d_scale, 8 * n)) || (rc = balloc(b, (void **)&b->d_iscale, 8 * n)) ||
      (rc = balloc(b, &b->d_ops, batch_op_bytes(0) * n)) || (rc = balloc(b, &b->d_ops32, batch_op_bytes(1) * n)) ||
      (rc = balloc(b, &b->d_opsx, batch_op_bytes(0) * n)) || (rc = build_items(b, batch_slots(), 1.0))) {
    vfg_batch_destroy(b);
    return rc;
  }
  if (hipMemcpy(b->d_ch, ch.data(), sizeof(BChunk) * ch.size(), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(b->d_vt, b->vt.data(), sizeof(BVec) * n, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(b->d_first, first.data(), sizeof(int) * n, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(b->d_nch, nch.data(), sizeof(int) * n, hipMemcpyHostToDevice) != hipSuccess) {
    vfg_batch_destroy(b);
    return VFG_E_HIP;
  }
  b->grid = 0;
  *out = b;
  return VFG_OK;
}

int vfg_batch_destroy(vfg_batch *b) {
  if (!b) return VFG_OK;
  if (b->stream) hipStreamSynchronize(b->stream);
  drop_graphs(b);
  bmg_destroy(b->bmg);
  for (void *p : b->allocs)
    if (p) hipFree(p);
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
  launch_batch_apply(0, b->sched, b->d_opsx, b->d_items, b->nitems, b->grid, b->stream);
  BCHK(b, hipStreamSynchronize(b->stream));
  BCHK(b, hipGetLastError());
  return VFG_OK;
}

int vfg_batch_u_apply_f32(vfg_batch *b, const float *const *x, float *const *y) {
  if (!b || !x || !y) return VFG_E_ARG;
  std::vector<const void *> xs(x, x + b->np), ys(y, y + b->np);
  fill_ops(b->m, xs.data(), (void *const *)ys.data(), 1, b->oz, b->d_ops32, b->stream);
  launch_batch_apply(1, b->sched, b->d_ops32, b->d_items, b->nitems, b->grid, b->stream);
  BCHK(b, hipStreamSynchronize(b->stream));
  BCHK(b, hipGetLastError());
  return VFG_OK;
}

// times nrep persistent applies p -> q over the members' CG direction vectors
int vfg_batch_time_u_apply(vfg_batch *b, int fp32, int nrep, double *ms) {
  if (!b || nrep < 1 || !ms) return VFG_E_ARG;
  std::vector<const void *> xs(b->np);
  std::vector<void *> ys(b->np);
  for (int p = 0; p < b->np; p++) {
    xs[p] = b->vt[p].v[fp32 ? VP32 : VP];
    ys[p] = b->vt[p].v[fp32 ? VQ32 : VQ];
  }
  void *ops = fp32 ? b->d_ops32 : b->d_opsx;
  fill_ops(b->m, xs.data(), ys.data(), fp32, b->oz, ops, b->stream);
  hipEvent_t e0, e1;
  BCHK(b, hipEventCreate(&e0));
  BCHK(b, hipEventCreate(&e1));
  hipEventRecord(e0, b->stream);
  for (int i = 0; i < nrep; i++) launch_batch_apply(fp32, b->sched, ops, b->d_items, b->nitems, b->grid, b->stream);
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

// Autotuner: times the persistent apply for every (schedule, z-chunk scale,
// grid) point -- schedules: 16 x 16 element tiles on the register march, 8 x 8
// tiles on the register march, 16 x 16 tiles on the LDS-ring march; z-chunk
// lengths 0.5, 1 and 2 times the whole-round choice; one workgroup per item,
// or a persistent grid of one or two rounds of resident workgroups -- keeps the
// fastest and records every point (vfg_batch_tune_points); best_ms receives
// the time of one apply.
int vfg_batch_autotune(vfg_batch *b, int fp32, double *best_ms) {
  if (!b) return VFG_E_ARG;
  const double scales[3] = {0.5, 1.0, 2.0};
  double best = 1e300, bscale = 1.0;
  int bgrid = 0, bsched = 0, rc;
  b->tune.clear();
  for (int sc = 0; sc < BATCH_SCHEDS; sc++) {
    b->sched = sc;
    const int slots = batch_slots_s(fp32, sc);
    const int grids[3] = {0, slots, 2 * slots};
    for (double z : scales) {
      if ((rc = build_items(b, slots, z))) return rc;
      for (int g : grids) {
        b->grid = g;
        double ms = 0.;
        if ((rc = vfg_batch_time_u_apply(b, fp32, 2, &ms)) || (rc = vfg_batch_time_u_apply(b, fp32, 5, &ms)))
          return rc;
        b->tune.push_back({sc, z, g, ms / 5.});
        if (ms < best) { best = ms; bscale = z; bgrid = g; bsched = sc; }
      }
    }
  }
  b->sched = bsched;
  if ((rc = build_items(b, batch_slots_s(fp32, bsched), bscale))) return rc;
  b->grid = bgrid;
  drop_graphs(b);
  if (best_ms) *best_ms = best / 5.;
  return VFG_OK;
}

int vfg_batch_tune_points(const vfg_batch *b, vfg_tune_point *out, int cap) {
  if (!b) return VFG_E_ARG;
  const int n = (int)b->tune.size();
  for (int i = 0; i < n && i < cap && out; i++) out[i] = b->tune[i];
  return n;
}

int vfg_batch_tuning(const vfg_batch *b, vfg_tune_point *pick) {
  if (!b || !pick) return VFG_E_ARG;
  pick->schedule = b->sched;
  pick->oz_scale = b->oz_scale;
  pick->grid = b->grid;
  pick->ms = 0.;
  for (const vfg_tune_point &t : b->tune)
    if (t.schedule == b->sched && t.oz_scale == b->oz_scale && t.grid == b->grid) pick->ms = t.ms;
  return VFG_OK;
}

int vfg_batch_set_graph(vfg_batch *b, int enable, int chunk) {
  if (!b || chunk < 1) return VFG_E_ARG;
  b->graph_on = enable;
  b->graph_chunk = chunk;
  drop_graphs(b);
  return VFG_OK;
}

}  // extern "C"

// one CG iteration of the whole ensemble; o = 0 (FP64 vectors) or VX32;
// mg: the FP32 preconditioner is the members' V-cycle instead of Jacobi
template <typename T>
static void cg_iter(vfg_batch *b, int o, void *ops, int fp32, double rtol, int maxit, bool mg) {
  hipStream_t s = b->stream;
  const BChunk *ch = b->d_ch;
  // frozen members are skipped by the apply and the vector kernels
  const int g = b->grid > 0 && b->grid < b->nitems ? b->grid : b->nitems;
  launch_engine_u_apply(fp32, b->sched, ops, b->d_items, 0, b->d_nitems, g, b->d_sc + B_ACTIVE, B_NS, s);
  hipLaunchKernelGGL(k_bdot<T>, dim3(b->nch), dim3(BRB), 0, s, (const BVec *)b->d_vt, o + VP, o + VQ, ch,
                     (const double *)b->d_sc, b->d_part);
  hipLaunchKernelGGL(k_bscalar, dim3(b->np), dim3(BRB), 0, s, (const double *)b->d_part, (const int *)b->d_first,
                     (const int *)b->d_nch, 1, b->d_sc, rtol, maxit);
  constexpr bool F64 = sizeof(T) == 8;  // FP64 CG: the V-cycle works on the FP32 copies r32, z32
  if (mg) {
    if constexpr (F64)
      hipLaunchKernelGGL(k_bxr_mg64, dim3(b->nch), dim3(BRB), 0, s, (const BVec *)b->d_vt, ch, (const double *)b->d_sc);
    else
      hipLaunchKernelGGL(k_bxr_mg<T>, dim3(b->nch), dim3(BRB), 0, s, (const BVec *)b->d_vt, o, ch,
                         (const double *)b->d_sc);
    bmg_vcycle(b->bmg, b->d_sc + B_ACTIVE, B_NS, s);
    if constexpr (F64)
      hipLaunchKernelGGL(k_brz64, dim3(b->nch), dim3(BRB), 0, s, (const BVec *)b->d_vt, ch, (const double *)b->d_sc, 0,
                         b->d_part);
    else
      hipLaunchKernelGGL(k_brz<T>, dim3(b->nch), dim3(BRB), 0, s, (const BVec *)b->d_vt, o, ch, (const double *)b->d_sc,
                         0, b->d_part);
  } else {
    hipLaunchKernelGGL(k_bxr<T>, dim3(b->nch), dim3(BRB), 0, s, (const BVec *)b->d_vt, o, ch, (const double *)b->d_sc,
                       b->d_part);
  }
  hipLaunchKernelGGL(k_bscalar, dim3(b->np), dim3(BRB), 0, s, (const double *)b->d_part, (const int *)b->d_first,
                     (const int *)b->d_nch, 2, b->d_sc, rtol, maxit);
  if (F64 && mg)
    hipLaunchKernelGGL(k_bp64, dim3(b->nch), dim3(BRB), 0, s, (const BVec *)b->d_vt, ch, (const double *)b->d_sc);
  else
    hipLaunchKernelGGL(k_bp<T>, dim3(b->nch), dim3(BRB), 0, s, (const BVec *)b->d_vt, o, ch, (const double *)b->d_sc);
}

// runs CG on all members to rtol; the iterate state (x, r) is prepared by the caller
template <typename T>
static int cg_run(vfg_batch *b, int o, void *ops, int fp32, double rtol, int maxit, hipGraphExec_t *gexec,
                  std::vector<double> &S, bool mg) {
  hipStream_t s = b->stream;
  if (mg) {
    // every member starts: the V-cycle on all of them (flags = 1)
    if constexpr (sizeof(T) == 8)
      hipLaunchKernelGGL(k_br32, dim3(b->nch), dim3(BRB), 0, s, (const BVec *)b->d_vt, (const BChunk *)b->d_ch);
    hipLaunchKernelGGL(k_bset_active, dim3(b->np), dim3(1), 0, s, b->d_sc);
    bmg_vcycle(b->bmg, b->d_sc + B_ACTIVE, B_NS, s);
    if constexpr (sizeof(T) == 8)
      hipLaunchKernelGGL(k_brz64, dim3(b->nch), dim3(BRB), 0, s, (const BVec *)b->d_vt, (const BChunk *)b->d_ch,
                         (const double *)nullptr, 1, b->d_part);
    else
      hipLaunchKernelGGL(k_brz<T>, dim3(b->nch), dim3(BRB), 0, s, (const BVec *)b->d_vt, o, (const BChunk *)b->d_ch,
                         (const double *)nullptr, 1, b->d_part);
  } else {
    hipLaunchKernelGGL(k_binit<T>, dim3(b->nch), dim3(BRB), 0, s, (const BVec *)b->d_vt, o, (const BChunk *)b->d_ch,
                       b->d_part);
  }
  hipLaunchKernelGGL(k_bscalar, dim3(b->np), dim3(BRB), 0, s, (const double *)b->d_part, (const int *)b->d_first,
                     (const int *)b->d_nch, 0, b->d_sc, rtol, maxit);
  const int K = b->graph_chunk;
  const int gi = fp32 ? 1 : 0;
  if (b->graph_on && *gexec && (b->g_rtol[gi] != rtol || b->g_maxit[gi] != maxit || b->g_mg[gi] != (int)mg)) {
    hipGraphExecDestroy(*gexec);
    *gexec = nullptr;
  }
  if (b->graph_on && !*gexec) {
    b->g_rtol[gi] = rtol;
    b->g_maxit[gi] = maxit;
    b->g_mg[gi] = (int)mg;
    hipGraph_t g;
    BCHK(b, hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    for (int k = 0; k < K; k++) cg_iter<T>(b, o, ops, fp32, rtol, maxit, mg);
    BCHK(b, hipStreamEndCapture(s, &g));
    hipError_t e = hipGraphInstantiate(gexec, g, nullptr, nullptr, 0);
    hipGraphDestroy(g);
    if (e != hipSuccess) {
      *gexec = nullptr;
      return bfail(b, VFG_E_HIP, "hipGraphInstantiate failed");
    }
  }
  S.assign((size_t)b->np * B_NS, 0.);
  for (int done = 0; done < maxit + K; done += K) {
    if (b->graph_on) BCHK(b, hipGraphLaunch(*gexec, s));
    else
      for (int k = 0; k < K; k++) cg_iter<T>(b, o, ops, fp32, rtol, maxit, mg);
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
  const int np = b->np;
  const dim3 gch(b->nch), blk(BRB);
  const BVec *vt = b->d_vt;
  const BChunk *ch = b->d_ch;
  auto V64 = [&](int p, int slot) { return (double *)b->vt[p].v[slot]; };
  std::vector<const void *> xs(np), xs32(np);
  std::vector<void *> ys(np), ys32(np);
  for (int p = 0; p < np; p++) {
    xs[p] = b->vt[p].v[VP];
    ys[p] = b->vt[p].v[VQ];
    xs32[p] = b->vt[p].v[VP32];
    ys32[p] = b->vt[p].v[VQ32];
  }
  // per member: x = U with Dirichlet values, r = F(x), dinv = diag
  hipMemcpyKind kin = loc == VFG_HOST ? hipMemcpyHostToDevice : hipMemcpyDeviceToDevice;
  for (int p = 0; p < np; p++) {
    vfg_handle *h = b->m[p];
    const int64_t n = 3 * h->P.nn;
    BCHK(b, hipMemcpyAsync(V64(p, VX), U[p], 8 * n, kin, s));
    BCHK(b, hipMemcpyAsync(V64(p, VB), BCU[p], 8 * n, kin, s));
    launch_vec_apply_dirichlet(h->P, 0, V64(p, VX), V64(p, VB), s);
    launch_u_rhs(h->P, V64(p, VX), V64(p, VB), V64(p, VR), 1, s);
    launch_u_diag(h->P, V64(p, VD), s);
  }
  hipLaunchKernelGGL(k_bprep, gch, blk, 0, s, vt, ch);  // r = -F, dinv = 1/diag (+ FP32 copy)
  std::vector<int> status(np, 0), iters(np, 0), outer(np, 0);
  // multigrid: every member's FP32 V-cycle, batched level by level; under the
  // FP64 CG (the single-problem default) or the FP32 CG of the refinement
  const bool mg = o->pc == VFG_PC_MG;
  std::vector<double> relres(np, 0.), S;
  int rc = VFG_OK;
  auto mg_setup_all = [&]() -> int {
    if (!b->bmg) {
      std::vector<float *> r32(np), z32(np);
      for (int p = 0; p < np; p++) {
        r32[p] = (float *)b->vt[p].v[VR32];
        z32[p] = (float *)b->vt[p].v[VZ32];
      }
      BCHK(b, hipStreamSynchronize(s));
      int e = bmg_create(b->m, r32.data(), z32.data(), &b->bmg);
      if (e) return bfail(b, e, "batched multigrid set-up failed");
      drop_graphs(b);
    }
    std::vector<const double *> dinv(np);
    for (int p = 0; p < np; p++) dinv[p] = (const double *)b->vt[p].v[VD];
    int e = bmg_setup(b->bmg, dinv.data(), s);
    return e ? bfail(b, e, "batched multigrid set-up failed") : VFG_OK;
  };
  if (o->precision != VFG_PREC_MIXED) {
    fill_ops(b->m, xs.data(), ys.data(), 0, b->oz, b->d_ops, s);
    if (mg && (rc = mg_setup_all())) return rc;
    rc = cg_run<double>(b, 0, b->d_ops, 0, o->rtol, o->maxit, &b->gexec, S, mg);
    for (int p = 0; p < np && !rc; p++) {
      const double *Sp = &S[(size_t)p * B_NS];
      iters[p] = (int)Sp[B_ITS];
      status[p] = (int)Sp[B_STATUS];
      relres[p] = Sp[B_RR0] > 0 ? sqrt(Sp[B_RR] / Sp[B_RR0]) : 0.;
      outer[p] = 1;
    }
  } else {
    // FP64 refinement around the FP32 inner CG
    fill_ops(b->m, xs32.data(), ys32.data(), 1, b->oz, b->d_ops32, s);
    if (mg && (rc = mg_setup_all())) return rc;
    std::vector<double> rn0(np, -1.), rprev(np, -1.), scale(np), iscale(np), nrm(np);
    std::vector<int> done(np, 0);
    const double inner_rtol = mg ? 1e-5 : 1e-4;  // as the single-problem refinement
    const int max_outer = 60;
    for (int it = 0; it < max_outer && !rc; it++) {
      // per-member FP64 residual norms
      hipLaunchKernelGGL(k_bdot<double>, gch, blk, 0, s, vt, (int)VR, (int)VR, ch, (const double *)nullptr,
                         b->d_part);
      batch_member_sums(b->d_part, b->d_first, b->d_nch, np, b->d_scale, s);
      BCHK(b, hipMemcpyAsync(nrm.data(), b->d_scale, 8 * np, hipMemcpyDeviceToHost, s));
      BCHK(b, hipStreamSynchronize(s));
      bool any = false;
      for (int p = 0; p < np; p++) {
        double rn = sqrt(nrm[p]);
        if (rn0[p] < 0.) rn0[p] = rn;
        if (!done[p]) relres[p] = rn0[p] > 0. ? rn / rn0[p] : 0.;
        if (!done[p] && !isfinite(rn)) { status[p] = VFG_E_NAN; done[p] = 1; }
        // converged at the tolerance, or at the FP64 residual's rounding floor
        // (an outer step no longer halves it: the rule of the single-problem solve)
        if (!done[p] && (rn <= fmax(o->rtol * rn0[p], o->atol) || rn == 0. || (rprev[p] > 0. && rn > .5 * rprev[p])))
          done[p] = 1;
        rprev[p] = rn;
        if (!done[p]) { outer[p] = it + 1; any = true; }
        scale[p] = done[p] ? 0. : 1. / rn;
        iscale[p] = done[p] ? 0. : rn;
      }
      if (!any) break;
      if (it == max_outer - 1) {
        for (int p = 0; p < np; p++)
          if (!done[p]) status[p] = VFG_E_MAXIT;
        break;
      }
      BCHK(b, hipMemcpyAsync(b->d_scale, scale.data(), 8 * np, hipMemcpyHostToDevice, s));
      BCHK(b, hipMemcpyAsync(b->d_iscale, iscale.data(), 8 * np, hipMemcpyHostToDevice, s));
      hipLaunchKernelGGL(k_b64to32, gch, blk, 0, s, vt, ch, (const double *)b->d_scale);
      rc = cg_run<float>(b, VX32, b->d_ops32, 1, inner_rtol, o->maxit, &b->gexec32, S, mg);
      for (int p = 0; p < np && !rc; p++) {
        if (done[p]) continue;
        iters[p] += (int)S[(size_t)p * B_NS + B_ITS];
        if ((int)S[(size_t)p * B_NS + B_STATUS] == VFG_E_NAN) { status[p] = VFG_E_NAN; done[p] = 1; }
      }
      hipLaunchKernelGGL(k_baxpy, gch, blk, 0, s, vt, ch, (const double *)b->d_iscale);
      for (int p = 0; p < np; p++) {
        if (done[p]) continue;
        vfg_handle *h = b->m[p];
        launch_u_rhs(h->P, V64(p, VX), V64(p, VB), V64(p, VR), 1, s);
      }
      hipLaunchKernelGGL(k_bneg, gch, blk, 0, s, vt, ch);
    }
  }
  if (rc) return rc;
  hipMemcpyKind kout = loc == VFG_HOST ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice;
  for (int p = 0; p < np; p++) {
    BCHK(b, hipMemcpyAsync(U[p], V64(p, VX), 8 * 3 * b->m[p]->P.nn, kout, s));
    if (info) {
      info[p].iterations = iters[p];
      info[p].outer = outer[p];
      info[p].relres = relres[p];
      info[p].status = status[p];
    }
  }
  BCHK(b, hipStreamSynchronize(s));
  BCHK(b, hipGetLastError());
  for (int p = 0; p < np; p++)
    if (status[p]) return bfail(b, status[p], "vfg_batch_u_solve: a member did not converge (see info[].status)");
  return VFG_OK;
}
