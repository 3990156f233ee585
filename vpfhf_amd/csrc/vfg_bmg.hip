// vfg_bmg.hip -- the multigrid V-cycle of every member of an ensemble at once.
//
// Each member keeps its own hierarchy (MgHier of its handle: levels,
// coefficients, diagonals, smoother bounds).  The V-cycle runs level by level
// over all members: on level l one persistent launch applies the level-l
// operator of every member that has that level, one launch performs a
// Chebyshev step of all of them (each member with its own coefficients),
// one launch restricts and one prolongs, each member on its own pair of
// grids.  A member's coarsest level gets Chebyshev(MG_COARSE_DEG) while the
// others smooth and descend.  Members that are frozen (converged) are skipped
// by every launch.  No launch mixes data of two members, so a member's result
// does not depend on the others.
#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>
#include <vector>
#include "vfg_handle.h"
#include "vfg_mg.h"
#include "vfg_batch_apply.h"
#include "vfg_bmg.h"

namespace vfg {

constexpr int BCH = 4096;  // entries per Chebyshev chunk
constexpr int BTR = 256;   // threads per transfer chunk: one fine node each (prolongation), four per coarse dof (restriction)


// Launch grids are list capacities; dn holds the live chunk count, so a
// captured graph stays valid while lists are rebuilt in place.
__global__ __launch_bounds__(256) void k_bmg_cheb(const BmgVec *__restrict__ lvt, int level,
                                                  const BmgChunk *__restrict__ ch, const int *__restrict__ dn,
                                                  const BmgCoef *__restrict__ coef, int step, int use_y, int xzero,
                                                  const double *__restrict__ flags, int fstride) {
  if ((int)blockIdx.x >= *dn) return;
  const BmgChunk c = ch[blockIdx.x];
  if (flags[(int64_t)c.p * fstride] == 0.) return;
  const BmgVec &v = lvt[c.p * BML + level];
  const BmgCoef k = coef[((c.p * BML + level) * 2 + v.coarsest) * BMK + step];
  for (int t = threadIdx.x; t < c.len; t += 256)
    mg_cheb_entry(v.b, use_y ? v.y : nullptr, v.dinv, v.d, v.x, k.c1, k.c2, step == 0, xzero, c.start + t);
}

// chunks of BTR / 4 coarse dofs (level + 1) of members that descend, four
// lanes per dof: lane kk sums the fine planes kk, kk + 4, ... of the gather
// and the four are added by shuffles in lane order (as k_mg_restrict4), so a
// thread's serial gather is a quarter as long
__global__ __launch_bounds__(256) void k_bmg_restrict(const BmgVec *__restrict__ lvt, int level,
                                                      const BmgChunk *__restrict__ ch, const int *__restrict__ dn,
                                                      const double *__restrict__ flags, int fstride) {
  if ((int)blockIdx.x >= *dn) return;
  const BmgChunk c = ch[blockIdx.x];
  if (flags[(int64_t)c.p * fstride] == 0.) return;
  const BmgVec &f = lvt[c.p * BML + level];
  const BmgVec &g = lvt[c.p * BML + level + 1];
  const int t = (int)threadIdx.x >> 2, kk = (int)threadIdx.x & 3;
  const bool ok = t < c.len;
  const int64_t dof = c.start + (ok ? t : 0), N = dof / 3;
  const float v = ok ? mg_restrict_dof_plane(f.G, g.G, f.bc, f.b, f.y, N, (int)(dof - 3 * N), kk) : 0.f;
  const float v1 = __shfl_down(v, 1, 4), v2 = __shfl_down(v, 2, 4), v3 = __shfl_down(v, 3, 4);
  if (ok && kk == 0) g.b[dof] = ((v + v1) + v2) + v3;
}

// chunks over the fine nodes (level) of members that descend: the coarse
// correction is added to the level's iterate d (the post-smoothing leaves x)
__global__ __launch_bounds__(256) void k_bmg_prolong(const BmgVec *__restrict__ lvt, int level,
                                                     const BmgChunk *__restrict__ ch, const int *__restrict__ dn,
                                                     const double *__restrict__ flags, int fstride) {
  if ((int)blockIdx.x >= *dn) return;
  const BmgChunk c = ch[blockIdx.x];
  if (flags[(int64_t)c.p * fstride] == 0.) return;
  const BmgVec &f = lvt[c.p * BML + level];
  const BmgVec &g = lvt[c.p * BML + level + 1];
  for (int t = threadIdx.x; t < c.len; t += 256) mg_prolong_node(f.G, g.G, f.bc, g.x, f.d, c.start + t);
}

// ------------------------------------------------------------------ host side
// A device array that grows (never shrinks) with a host mirror; growth
// changes its address, which the owner of a captured graph must learn.
template <typename V>
struct Grow {
  V *d = nullptr;
  int cap = 0;
  std::vector<V> h;
};
struct BmgList {
  Grow<BItem> items;
  Grow<BmgChunk> ch;
};
struct BmgPlan {
  void *ops = nullptr;  // one FP32 apply record per member (level operator, x -> y)
  void *ops_pre = nullptr, *ops_post = nullptr;  // fused smoothing records per member
  BmgList M, D, C;      // members with this level / that descend from it / whose coarsest it is
  Grow<BmgChunk> rch;   // restriction: coarse dofs of the descending members
  Grow<BmgChunk> pch;   // prolongation: fine nodes of the descending members
  // live counts: M items, M chunks, D items, D chunks, C items, C chunks, rch, pch
  int *dcnt = nullptr;
  int hcnt[8] = {0};
  int cfit = 0;  // every member whose coarsest level this is fits the one-launch solve
};
struct BatchMG {
  std::vector<vfg_handle *> m;        // per member (nullptr: empty slot)
  std::vector<float *> r32, z32;
  int np = 0, slots = 0;
  int Lcap = 0;                       // levels whose lists have room (the V-cycle's launches)
  std::vector<int> L;                 // levels per member (0: empty)
  std::vector<std::vector<int>> oz;   // [p][l]
  BmgPlan lv[BML];
  BmgVec *d_lvt = nullptr;
  BmgCoef *d_coef = nullptr;
  std::vector<BmgVec> h_lvt;          // staging of the tables (kept: uploads are asynchronous)
  std::vector<BmgCoef> h_coef;
  std::vector<std::vector<char>> h_ops, h_pre, h_post;
  std::vector<void *> allocs;         // fixed allocations
  std::vector<void *> grown;          // growable buffers (freed with the object)
};

static int galloc(BatchMG *g, void **p, size_t bytes) {
  if (hipMalloc(p, bytes ? bytes : 8) != hipSuccess) {
    *p = nullptr;
    return VFG_E_NOMEM;
  }
  g->allocs.push_back(*p);
  return VFG_OK;
}

// sizes the device array for its host mirror; *grew is set when it moved
template <typename V>
static int grow(BatchMG *g, Grow<V> &a, bool *grew) {
  const int need = (int)a.h.size();
  if (need <= a.cap) return VFG_OK;
  const int nc = a.cap == 0 ? need : (need > 2 * a.cap ? need : 2 * a.cap);
  if (a.d) {
    for (auto &q : g->grown)
      if (q == (void *)a.d) q = nullptr;
    hipFree(a.d);
  }
  a.d = nullptr;
  a.cap = 0;
  if (hipMalloc((void **)&a.d, sizeof(V) * (size_t)nc) != hipSuccess) {
    a.d = nullptr;
    return VFG_E_NOMEM;
  }
  g->grown.push_back(a.d);
  a.cap = nc;
  *grew = true;
  return VFG_OK;
}
template <typename V>
static int push(Grow<V> &a, hipStream_t s) {
  if (a.h.empty()) return VFG_OK;
  return hipMemcpyAsync(a.d, a.h.data(), sizeof(V) * a.h.size(), hipMemcpyHostToDevice, s) == hipSuccess
             ? VFG_OK
             : VFG_E_HIP;
}

static const Grid &level_grid(const BatchMG *g, int p, int l) { return g->m[p]->mg->lv[l].P.G; }

// apply items and Chebyshev chunks of the members selected by sel(p)
template <typename Sel>
static void fill_list(BatchMG *g, int l, Sel sel, BmgList &out) {
  out.items.h.clear();
  out.ch.h.clear();
  for (int p = 0; p < g->np; p++) {
    if (g->L[p] <= l || !sel(p)) continue;
    const Grid &G = level_grid(g, p, l);
    const int tx = (G.nx + 14) / 15, ty = (G.ny + 14) / 15, OZ = g->oz[p][l];
    for (int z = 0; z < (G.nz + OZ - 1) / OZ; z++)
      for (int y = 0; y < ty; y++)
        for (int x = 0; x < tx; x++) out.items.h.push_back({p, x, y, z});
    const int64_t n = 3 * (int64_t)G.nx * G.ny * G.nz;
    for (int64_t s = 0; s < n; s += BCH) out.ch.h.push_back({p, s, (int)(n - s < BCH ? n - s : BCH)});
  }
}

int bmg_set_member(BatchMG *g, int p, vfg_handle *h, float *r32, float *z32) {
  if (p < 0 || p >= g->np) return VFG_E_ARG;
  g->m[p] = h;
  g->r32[p] = r32;
  g->z32[p] = z32;
  g->L[p] = 0;
  if (!h) return VFG_OK;
  int rc = mg_build(h);
  if (rc) return rc;
  g->L[p] = h->mg->L;
  for (int l = 0; l < g->L[p]; l++) {
    const Grid &G = level_grid(g, p, l);
    int oz = choose_oz_pub(G, ((G.nx + 14) / 15) * ((G.ny + 14) / 15), g->slots);
    g->oz[p][l] = oz < 2 ? 2 : oz > G.nz ? G.nz : oz;
  }
  return VFG_OK;
}

int bmg_rebuild(BatchMG *g, hipStream_t s, bool *grew) {
  int rc = VFG_OK;
  int Lmax = 0;
  for (int p = 0; p < g->np; p++) Lmax = g->L[p] > Lmax ? g->L[p] : Lmax;
  for (int l = 0; l < BML && !rc; l++) {
    BmgPlan &P = g->lv[l];
    fill_list(g, l, [](int) { return true; }, P.M);
    fill_list(g, l, [&](int p) { return g->L[p] - 1 > l; }, P.D);
    fill_list(g, l, [&](int p) { return g->L[p] - 1 == l; }, P.C);
    P.rch.h.clear();
    P.pch.h.clear();
    for (int p = 0; p < g->np; p++) {
      if (g->L[p] - 1 <= l) continue;
      const Grid &Gf = level_grid(g, p, l), &Gc = level_grid(g, p, l + 1);
      const int64_t nc = 3 * (int64_t)Gc.nx * Gc.ny * Gc.nz, nf = (int64_t)Gf.nx * Gf.ny * Gf.nz;
      constexpr int RC = BTR / 4;  // four lanes per coarse dof
      for (int64_t st = 0; st < nc; st += RC) P.rch.h.push_back({p, st, (int)(nc - st < RC ? nc - st : RC)});
      for (int64_t st = 0; st < nf; st += BTR) P.pch.h.push_back({p, st, (int)(nf - st < BTR ? nf - st : BTR)});
    }
    if (l >= Lmax && !P.dcnt) continue;  // nothing on this level, ever
    if (!P.dcnt && (rc = galloc(g, (void **)&P.dcnt, 8 * sizeof(int)))) break;
    if (!P.ops && (rc = galloc(g, &P.ops, batch_op_bytes(1) * g->np))) break;
    if (!P.ops_pre && (rc = galloc(g, &P.ops_pre, mg_smooth_op_bytes(0) * g->np))) break;
    if (!P.ops_post && (rc = galloc(g, &P.ops_post, mg_smooth_op_bytes(1) * g->np))) break;
    BmgList *ls[3] = {&P.M, &P.D, &P.C};
    for (int k = 0; k < 3 && !rc; k++)
      if ((rc = grow(g, ls[k]->items, grew)) || (rc = grow(g, ls[k]->ch, grew))) break;
    if (rc || (rc = grow(g, P.rch, grew)) || (rc = grow(g, P.pch, grew))) break;
    for (int k = 0; k < 3; k++) {
      P.hcnt[2 * k] = (int)ls[k]->items.h.size();
      P.hcnt[2 * k + 1] = (int)ls[k]->ch.h.size();
    }
    P.hcnt[6] = (int)P.rch.h.size();
    P.hcnt[7] = (int)P.pch.h.size();
    int cfit = P.C.ch.h.empty() ? 0 : 1;
    for (int p = 0; p < g->np && cfit; p++)
      if (g->m[p] && g->L[p] - 1 == l && !mg_coarse_fits(g->m[p]->mg->lv[l].P)) cfit = 0;
    if (cfit != P.cfit) *grew = true;  // the V-cycle's launches change
    P.cfit = cfit;
    for (int k = 0; k < 3 && !rc; k++)
      if ((rc = push(ls[k]->items, s)) || (rc = push(ls[k]->ch, s))) break;
    if (rc || (rc = push(P.rch, s)) || (rc = push(P.pch, s))) break;
    if (hipMemcpyAsync(P.dcnt, P.hcnt, sizeof P.hcnt, hipMemcpyHostToDevice, s) != hipSuccess) rc = VFG_E_HIP;
  }
  int Lcap = 0;
  for (int l = 0; l < BML; l++)
    if (g->lv[l].M.ch.cap > 0) Lcap = l + 1;
  if (Lcap != g->Lcap) *grew = true;
  g->Lcap = Lcap;
  // the counts above are read from their host arrays when the copies run
  if (!rc && hipStreamSynchronize(s) != hipSuccess) rc = VFG_E_HIP;
  return rc;
}

static int bmg_alloc(int np, BatchMG **out) {
  *out = nullptr;
  BatchMG *g = new BatchMG();
  g->np = np;
  g->m.assign(np, nullptr);
  g->r32.assign(np, nullptr);
  g->z32.assign(np, nullptr);
  g->slots = batch_slots();
  g->L.assign(np, 0);
  g->oz.assign(np, std::vector<int>(BML, 2));
  g->h_lvt.assign((size_t)np * BML, BmgVec{});
  g->h_coef.assign((size_t)np * BML * 2 * BMK, BmgCoef{0.f, 0.f});
  g->h_ops.assign(BML, std::vector<char>(batch_op_bytes(1) * np, 0));
  g->h_pre.assign(BML, std::vector<char>(mg_smooth_op_bytes(0) * np, 0));
  g->h_post.assign(BML, std::vector<char>(mg_smooth_op_bytes(1) * np, 0));
  int rc = galloc(g, (void **)&g->d_lvt, sizeof(BmgVec) * np * BML);
  if (!rc) rc = galloc(g, (void **)&g->d_coef, sizeof(BmgCoef) * np * BML * 2 * BMK);
  if (rc) {
    bmg_destroy(g);
    return rc;
  }
  *out = g;
  return VFG_OK;
}

int bmg_create_slots(int np, BatchMG **out) { return bmg_alloc(np, out); }

int bmg_create(const std::vector<vfg_handle *> &m, float *const *r32, float *const *z32, BatchMG **out) {
  BatchMG *g;
  int rc = bmg_alloc((int)m.size(), &g);
  if (rc) return rc;
  for (int p = 0; p < g->np && !rc; p++) rc = bmg_set_member(g, p, m[p], r32[p], z32[p]);
  bool grew = false;
  hipStream_t s = m.empty() ? nullptr : m[0]->stream;
  if (!rc) rc = bmg_rebuild(g, s, &grew);
  if (rc) {
    bmg_destroy(g);
    return rc;
  }
  *out = g;
  return VFG_OK;
}

void bmg_destroy(BatchMG *g) {
  if (!g) return;
  for (void *p : g->allocs)
    if (p) hipFree(p);
  for (void *p : g->grown)
    if (p) hipFree(p);
  delete g;
}

// Chebyshev coefficients of one phase: step 0 c2 = 1/theta; step i c1, c2 of
// the three-term recurrence on [lo, hi] (as the single-problem smoother)
static void cheb_coefs(double lo, double hi, int k, BmgCoef *out) {
  const double th = .5 * (hi + lo), de = .5 * (hi - lo), sg = th / de;
  double rho = 1. / sg;
  out[0] = {0.f, (float)(1. / th)};
  for (int i = 1; i < k; i++) {
    const double rn = 1. / (2. * sg - rho);
    out[i] = {(float)(rn * rho), (float)(2. * rn / de)};
    rho = rn;
  }
}

// member p's level vectors, smoother coefficients and level operators into the staging tables
static void stage_member(BatchMG *g, int p) {
  MgHier &H = *g->m[p]->mg;
  for (int l = 0; l < g->L[p]; l++) {
    MgLevel &V = H.lv[l];
    BmgVec &v = g->h_lvt[(size_t)p * BML + l];
    v.b = l == 0 ? g->r32[p] : V.b;
    v.x = l == 0 ? g->z32[p] : V.x;
    v.y = V.y;
    v.d = V.d;
    v.dinv = V.dinv;
    v.G = V.P.G;
    v.bc = V.P.bcU;
    v.coarsest = l == g->L[p] - 1;
    BmgCoef *cf = &g->h_coef[((size_t)p * BML + l) * 2 * BMK];
    cheb_coefs(V.lmax / 11., V.lmax, MG_SMOOTH, cf);           // smoother
    cheb_coefs(V.lmax / 30., V.lmax, MG_COARSE_DEG, cf + BMK);  // coarsest level
    make_uop(V.P, v.x, v.y, 1, g->oz[p][l], g->h_ops[l].data() + batch_op_bytes(1) * p);
    // descending levels: pre-smoothing from 0 into d (A d into y), post-smoothing d -> x
    static_assert(MG_SMOOTH == 1, "the fused smoothing records do one Chebyshev step");
    make_mg_smooth_op(V.P, 0, v.b, v.dinv, cf[0].c2, nullptr, v.d, v.y, g->oz[p][l],
                      g->h_pre[l].data() + mg_smooth_op_bytes(0) * p);
    make_mg_smooth_op(V.P, 1, v.b, v.dinv, cf[0].c2, v.d, v.x, nullptr, g->oz[p][l],
                      g->h_post[l].data() + mg_smooth_op_bytes(1) * p);
  }
}
static int upload_member(BatchMG *g, int p, hipStream_t s) {
  const size_t ob = batch_op_bytes(1);
  if (hipMemcpyAsync(g->d_lvt + (size_t)p * BML, &g->h_lvt[(size_t)p * BML], sizeof(BmgVec) * BML,
                     hipMemcpyHostToDevice, s) != hipSuccess ||
      hipMemcpyAsync(g->d_coef + (size_t)p * BML * 2 * BMK, &g->h_coef[(size_t)p * BML * 2 * BMK],
                     sizeof(BmgCoef) * BML * 2 * BMK, hipMemcpyHostToDevice, s) != hipSuccess)
    return VFG_E_HIP;
  const size_t pb = mg_smooth_op_bytes(0), qb = mg_smooth_op_bytes(1);
  for (int l = 0; l < g->L[p]; l++)
    if (!g->lv[l].ops || !g->lv[l].ops_pre || !g->lv[l].ops_post ||
        hipMemcpyAsync((char *)g->lv[l].ops + ob * p, g->h_ops[l].data() + ob * p, ob, hipMemcpyHostToDevice, s) !=
            hipSuccess ||
        hipMemcpyAsync((char *)g->lv[l].ops_pre + pb * p, g->h_pre[l].data() + pb * p, pb, hipMemcpyHostToDevice,
                       s) != hipSuccess ||
        hipMemcpyAsync((char *)g->lv[l].ops_post + qb * p, g->h_post[l].data() + qb * p, qb, hipMemcpyHostToDevice,
                       s) != hipSuccess)
      return VFG_E_HIP;
  return VFG_OK;
}

int bmg_setup(BatchMG *g, const double *const *dinv64, hipStream_t s) {
  if (hipStreamSynchronize(s) != hipSuccess) return VFG_E_HIP;  // dinv64 is written on the batch stream
  int rc;
  for (int p = 0; p < g->np; p++)
    if (g->m[p] && (rc = mg_setup(g->m[p], dinv64[p]))) return rc;
  for (int p = 0; p < g->np; p++)
    if (g->m[p] && hipStreamSynchronize(g->m[p]->stream) != hipSuccess) return VFG_E_HIP;
  for (int p = 0; p < g->np; p++)
    if (g->m[p]) {
      stage_member(g, p);
      if ((rc = upload_member(g, p, s))) return rc;
    }
  return hipStreamSynchronize(s) == hipSuccess ? VFG_OK : VFG_E_HIP;
}

int bmg_setup_member(BatchMG *g, int p, const double *dinv64, hipStream_t s) {
  if (p < 0 || p >= g->np || !g->m[p]) return VFG_E_ARG;
  if (hipStreamSynchronize(s) != hipSuccess) return VFG_E_HIP;  // dinv64 and the fields are written on s
  int rc = mg_setup(g->m[p], dinv64);
  if (rc) return rc;
  if (hipStreamSynchronize(g->m[p]->stream) != hipSuccess) return VFG_E_HIP;
  stage_member(g, p);
  if ((rc = upload_member(g, p, s))) return rc;
  return hipStreamSynchronize(s) == hipSuccess ? VFG_OK : VFG_E_HIP;  // the staging is reused
}

static void apply(BatchMG *g, int l, const BmgList &L, int which, const double *flags, int fstride, hipStream_t s) {
  if (!L.items.cap) return;
  const int grid = L.items.cap < g->slots ? L.items.cap : g->slots;
  launch_engine_u_apply(1, 0, g->lv[l].ops, L.items.d, 0, g->lv[l].dcnt + 2 * which, grid, flags, fstride, s);
}
static void cheb(BatchMG *g, int l, const BmgList &L, int which, int step, int use_y, int xzero, const double *flags,
                 int fstride, hipStream_t s) {
  if (!L.ch.cap) return;
  hipLaunchKernelGGL(k_bmg_cheb, dim3(L.ch.cap), dim3(256), 0, s, (const BmgVec *)g->d_lvt, l,
                     (const BmgChunk *)L.ch.d, (const int *)(g->lv[l].dcnt + 2 * which + 1),
                     (const BmgCoef *)g->d_coef, step, use_y, xzero, flags, fstride);
}

static void smooth(BatchMG *g, int l, const BmgList &L, int post, const double *flags, int fstride, hipStream_t s) {
  if (!L.items.cap) return;
  const int grid = L.items.cap < g->slots ? L.items.cap : g->slots;
  launch_batch_mg_smooth(post, post ? g->lv[l].ops_post : g->lv[l].ops_pre, L.items.d, g->lv[l].dcnt + 2, grid, flags,
                         fstride, s);
}

// Descending members: one Chebyshev-Jacobi step from 0 fused with its apply
// (d, A d), restriction of b - A d, the coarse correction added to d, one step
// fused with its apply (d -> x) -- the single-problem V-cycle's kernels over
// the work list.  Members whose coarsest level this is: the one-launch
// Chebyshev solve, or Chebyshev steps with separate applies.
void bmg_vcycle(BatchMG *g, const double *flags, int fstride, hipStream_t s) {
  for (int l = 0; l < g->Lcap; l++) {
    BmgPlan &P = g->lv[l];
    if (P.D.ch.cap) smooth(g, l, P.D, 0, flags, fstride, s);
    if (P.cfit) {
      launch_bmg_coarse_cheb(P.ops, (const BmgVec *)g->d_lvt, l, (const BmgChunk *)P.C.ch.d,
                             (const int *)(P.dcnt + 5), P.C.ch.cap, (const BmgCoef *)g->d_coef, flags, fstride, s);
    } else {
      for (int i = 0; i < BMK && P.C.ch.cap; i++) {
        if (i > 0) apply(g, l, P.C, 2, flags, fstride, s);
        cheb(g, l, P.C, 2, i, i > 0, i == 0, flags, fstride, s);
      }
    }
    if (P.D.ch.cap && P.rch.cap)
      hipLaunchKernelGGL(k_bmg_restrict, dim3(P.rch.cap), dim3(256), 0, s, (const BmgVec *)g->d_lvt, l,
                         (const BmgChunk *)P.rch.d, (const int *)(P.dcnt + 6), flags, fstride);
  }
  for (int l = g->Lcap - 2; l >= 0; l--) {
    BmgPlan &P = g->lv[l];
    if (!P.D.ch.cap) continue;
    if (P.pch.cap)
      hipLaunchKernelGGL(k_bmg_prolong, dim3(P.pch.cap), dim3(256), 0, s, (const BmgVec *)g->d_lvt, l,
                         (const BmgChunk *)P.pch.d, (const int *)(P.dcnt + 7), flags, fstride);
    smooth(g, l, P.D, 1, flags, fstride, s);
  }
}

}  // namespace vfg
