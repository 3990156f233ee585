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

constexpr int BML = MgHier::MAXL;
constexpr int BMK = MG_COARSE_DEG > MG_SMOOTH ? MG_COARSE_DEG : MG_SMOOTH;
constexpr int BCH = 4096;  // entries per Chebyshev chunk
constexpr int BTR = 256;   // entries per transfer chunk: one gather per thread

struct BmgVec {
  float *b, *x, *y, *d;
  const float *dinv;
  Grid G;
  BcMask bc;
  int coarsest;  // 1 on the member's coarsest level
};
struct BmgChunk {
  int p;
  int64_t start;
  int len;
};
struct BmgCoef {
  float c1, c2;
};

__global__ __launch_bounds__(256) void k_bmg_cheb(const BmgVec *__restrict__ lvt, int level,
                                                  const BmgChunk *__restrict__ ch, const BmgCoef *__restrict__ coef,
                                                  int step, int use_y, int xzero, const double *__restrict__ flags,
                                                  int fstride) {
  const BmgChunk c = ch[blockIdx.x];
  if (flags[(int64_t)c.p * fstride] == 0.) return;
  const BmgVec &v = lvt[c.p * BML + level];
  const BmgCoef k = coef[((c.p * BML + level) * 2 + v.coarsest) * BMK + step];
  for (int t = threadIdx.x; t < c.len; t += 256)
    mg_cheb_entry(v.b, use_y ? v.y : nullptr, v.dinv, v.d, v.x, k.c1, k.c2, step == 0, xzero, c.start + t);
}

// chunks over the coarse dofs (level + 1) of members that descend
__global__ __launch_bounds__(256) void k_bmg_restrict(const BmgVec *__restrict__ lvt, int level,
                                                      const BmgChunk *__restrict__ ch, const double *__restrict__ flags,
                                                      int fstride) {
  const BmgChunk c = ch[blockIdx.x];
  if (flags[(int64_t)c.p * fstride] == 0.) return;
  const BmgVec &f = lvt[c.p * BML + level];
  const BmgVec &g = lvt[c.p * BML + level + 1];
  for (int t = threadIdx.x; t < c.len; t += 256) {
    const int64_t dof = c.start + t, N = dof / 3;
    g.b[dof] = mg_restrict_dof(f.G, g.G, f.bc, f.b, f.y, N, (int)(dof - 3 * N));
  }
}

// chunks over the fine nodes (level) of members that descend
__global__ __launch_bounds__(256) void k_bmg_prolong(const BmgVec *__restrict__ lvt, int level,
                                                     const BmgChunk *__restrict__ ch, const double *__restrict__ flags,
                                                     int fstride) {
  const BmgChunk c = ch[blockIdx.x];
  if (flags[(int64_t)c.p * fstride] == 0.) return;
  const BmgVec &f = lvt[c.p * BML + level];
  const BmgVec &g = lvt[c.p * BML + level + 1];
  for (int t = threadIdx.x; t < c.len; t += 256) mg_prolong_node(f.G, g.G, f.bc, g.x, f.x, c.start + t);
}

// ------------------------------------------------------------------ host side
struct BmgList {
  BItem *items = nullptr;
  int nitems = 0;
  int *dcount = nullptr;
  BmgChunk *ch = nullptr;
  int nch = 0;
};
struct BmgPlan {
  void *ops = nullptr;  // one FP32 apply record per member (level operator, x -> y)
  BmgList M, D, C;      // members with this level / that descend from it / whose coarsest it is
  BmgChunk *rch = nullptr;  // restriction: coarse dofs of the descending members
  int nrch = 0;
  BmgChunk *pch = nullptr;  // prolongation: fine nodes of the descending members
  int npch = 0;
};
struct BatchMG {
  std::vector<vfg_handle *> m;
  std::vector<float *> r32, z32;
  int np = 0, Lmax = 0, slots = 0;
  std::vector<int> L;                 // levels per member
  std::vector<std::vector<int>> oz;   // [p][l]
  BmgPlan lv[BML];
  BmgVec *d_lvt = nullptr;
  BmgCoef *d_coef = nullptr;
  std::vector<void *> allocs;
};

static int galloc(BatchMG *g, void **p, size_t bytes) {
  if (hipMalloc(p, bytes ? bytes : 8) != hipSuccess) {
    *p = nullptr;
    return VFG_E_NOMEM;
  }
  g->allocs.push_back(*p);
  return VFG_OK;
}

template <typename V>
static int upload(BatchMG *g, V **dst, const std::vector<V> &src) {
  int rc = galloc(g, (void **)dst, sizeof(V) * src.size());
  if (rc) return rc;
  if (!src.empty() && hipMemcpy(*dst, src.data(), sizeof(V) * src.size(), hipMemcpyHostToDevice) != hipSuccess)
    return VFG_E_HIP;
  return VFG_OK;
}

static const Grid &level_grid(const BatchMG *g, int p, int l) { return g->m[p]->mg->lv[l].P.G; }

// apply items and Chebyshev chunks of the members selected by sel(p)
template <typename Sel>
static int build_list(BatchMG *g, int l, Sel sel, BmgList &out) {
  std::vector<BItem> items;
  std::vector<BmgChunk> ch;
  for (int p = 0; p < g->np; p++) {
    if (g->L[p] <= l || !sel(p)) continue;
    const Grid &G = level_grid(g, p, l);
    const int tx = (G.nx + 14) / 15, ty = (G.ny + 14) / 15, OZ = g->oz[p][l];
    for (int z = 0; z < (G.nz + OZ - 1) / OZ; z++)
      for (int y = 0; y < ty; y++)
        for (int x = 0; x < tx; x++) items.push_back({p, x, y, z});
    const int64_t n = 3 * (int64_t)G.nx * G.ny * G.nz;
    for (int64_t s = 0; s < n; s += BCH) ch.push_back({p, s, (int)(n - s < BCH ? n - s : BCH)});
  }
  out.nitems = (int)items.size();
  out.nch = (int)ch.size();
  int rc;
  if ((rc = upload(g, &out.items, items)) || (rc = upload(g, &out.ch, ch))) return rc;
  std::vector<int> cnt(1, out.nitems);
  return upload(g, &out.dcount, cnt);
}

int bmg_create(const std::vector<vfg_handle *> &m, float *const *r32, float *const *z32, BatchMG **out) {
  *out = nullptr;
  BatchMG *g = new BatchMG();
  g->m = m;
  g->np = (int)m.size();
  g->r32.assign(r32, r32 + g->np);
  g->z32.assign(z32, z32 + g->np);
  g->slots = batch_slots();
  int rc = VFG_OK;
  g->L.assign(g->np, 1);
  g->oz.assign(g->np, std::vector<int>(BML, 2));
  for (int p = 0; p < g->np && !rc; p++) {
    if ((rc = mg_build(m[p]))) break;
    g->L[p] = m[p]->mg->L;
    g->Lmax = g->L[p] > g->Lmax ? g->L[p] : g->Lmax;
    for (int l = 0; l < g->L[p]; l++) {
      const Grid &G = level_grid(g, p, l);
      int oz = choose_oz_pub(G, ((G.nx + 14) / 15) * ((G.ny + 14) / 15), g->slots);
      g->oz[p][l] = oz < 2 ? 2 : oz > G.nz ? G.nz : oz;
    }
  }
  for (int l = 0; l < g->Lmax && !rc; l++) {
    BmgPlan &P = g->lv[l];
    if ((rc = galloc(g, &P.ops, batch_op_bytes(1) * g->np))) break;
    if ((rc = build_list(g, l, [](int) { return true; }, P.M)) ||
        (rc = build_list(g, l, [&](int p) { return g->L[p] - 1 > l; }, P.D)) ||
        (rc = build_list(g, l, [&](int p) { return g->L[p] - 1 == l; }, P.C)))
      break;
    std::vector<BmgChunk> rch, pch;
    for (int p = 0; p < g->np; p++) {
      if (g->L[p] - 1 <= l) continue;
      const Grid &Gf = level_grid(g, p, l), &Gc = level_grid(g, p, l + 1);
      const int64_t nc = 3 * (int64_t)Gc.nx * Gc.ny * Gc.nz, nf = (int64_t)Gf.nx * Gf.ny * Gf.nz;
      for (int64_t s = 0; s < nc; s += BTR) rch.push_back({p, s, (int)(nc - s < BTR ? nc - s : BTR)});
      for (int64_t s = 0; s < nf; s += BTR) pch.push_back({p, s, (int)(nf - s < BTR ? nf - s : BTR)});
    }
    P.nrch = (int)rch.size();
    P.npch = (int)pch.size();
    if ((rc = upload(g, &P.rch, rch)) || (rc = upload(g, &P.pch, pch))) break;
  }
  if (!rc) rc = galloc(g, (void **)&g->d_lvt, sizeof(BmgVec) * g->np * BML);
  if (!rc) rc = galloc(g, (void **)&g->d_coef, sizeof(BmgCoef) * g->np * BML * 2 * BMK);
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

int bmg_setup(BatchMG *g, const double *const *dinv64, hipStream_t s) {
  if (hipStreamSynchronize(s) != hipSuccess) return VFG_E_HIP;  // dinv64 is written on the batch stream
  int rc;
  for (int p = 0; p < g->np; p++)
    if ((rc = mg_setup(g->m[p], dinv64[p]))) return rc;
  for (int p = 0; p < g->np; p++)
    if (hipStreamSynchronize(g->m[p]->stream) != hipSuccess) return VFG_E_HIP;
  std::vector<BmgVec> lvt((size_t)g->np * BML);
  std::vector<BmgCoef> coef((size_t)g->np * BML * 2 * BMK, BmgCoef{0.f, 0.f});
  std::vector<std::vector<char>> ops(g->Lmax, std::vector<char>(batch_op_bytes(1) * g->np, 0));
  for (int p = 0; p < g->np; p++) {
    MgHier &H = *g->m[p]->mg;
    for (int l = 0; l < g->L[p]; l++) {
      MgLevel &V = H.lv[l];
      BmgVec &v = lvt[(size_t)p * BML + l];
      v.b = l == 0 ? g->r32[p] : V.b;
      v.x = l == 0 ? g->z32[p] : V.x;
      v.y = V.y;
      v.d = V.d;
      v.dinv = V.dinv;
      v.G = V.P.G;
      v.bc = V.P.bcU;
      v.coarsest = l == g->L[p] - 1;
      BmgCoef *cf = &coef[((size_t)p * BML + l) * 2 * BMK];
      cheb_coefs(V.lmax / 11., V.lmax, MG_SMOOTH, cf);           // smoother
      cheb_coefs(V.lmax / 30., V.lmax, MG_COARSE_DEG, cf + BMK);  // coarsest level
      make_uop(V.P, v.x, v.y, 1, g->oz[p][l], ops[l].data() + batch_op_bytes(1) * p);
    }
  }
  if (hipMemcpyAsync(g->d_lvt, lvt.data(), sizeof(BmgVec) * lvt.size(), hipMemcpyHostToDevice, s) != hipSuccess ||
      hipMemcpyAsync(g->d_coef, coef.data(), sizeof(BmgCoef) * coef.size(), hipMemcpyHostToDevice, s) != hipSuccess)
    return VFG_E_HIP;
  for (int l = 0; l < g->Lmax; l++)
    if (hipMemcpyAsync(g->lv[l].ops, ops[l].data(), ops[l].size(), hipMemcpyHostToDevice, s) != hipSuccess)
      return VFG_E_HIP;
  return hipStreamSynchronize(s) == hipSuccess ? VFG_OK : VFG_E_HIP;
}

static void apply(BatchMG *g, int l, const BmgList &L, const double *flags, int fstride, hipStream_t s) {
  if (!L.nitems) return;
  const int grid = L.nitems < g->slots ? L.nitems : g->slots;
  launch_engine_u_apply(1, g->lv[l].ops, L.items, L.dcount, grid, flags, fstride, s);
}
static void cheb(BatchMG *g, int l, const BmgList &L, int step, int use_y, int xzero, const double *flags,
                 int fstride, hipStream_t s) {
  if (!L.nch) return;
  hipLaunchKernelGGL(k_bmg_cheb, dim3(L.nch), dim3(256), 0, s, (const BmgVec *)g->d_lvt, l, (const BmgChunk *)L.ch,
                     (const BmgCoef *)g->d_coef, step, use_y, xzero, flags, fstride);
}

void bmg_vcycle(BatchMG *g, const double *flags, int fstride, hipStream_t s) {
  for (int l = 0; l < g->Lmax; l++) {
    BmgPlan &P = g->lv[l];
    // pre-smoothing (descending members) and the coarsest-level solve, from x = 0
    for (int i = 0; i < BMK; i++) {
      const BmgList &L = i < MG_SMOOTH ? P.M : P.C;
      if (!L.nch) break;
      if (i > 0) apply(g, l, L, flags, fstride, s);
      cheb(g, l, L, i, i > 0, i == 0, flags, fstride, s);
    }
    if (P.D.nch) {
      apply(g, l, P.D, flags, fstride, s);  // residual b - A x, restricted
      hipLaunchKernelGGL(k_bmg_restrict, dim3(P.nrch), dim3(256), 0, s, (const BmgVec *)g->d_lvt, l,
                         (const BmgChunk *)P.rch, flags, fstride);
    }
  }
  for (int l = g->Lmax - 2; l >= 0; l--) {
    BmgPlan &P = g->lv[l];
    if (!P.D.nch) continue;
    hipLaunchKernelGGL(k_bmg_prolong, dim3(P.npch), dim3(256), 0, s, (const BmgVec *)g->d_lvt, l,
                       (const BmgChunk *)P.pch, flags, fstride);
    for (int i = 0; i < MG_SMOOTH; i++) {  // post-smoothing
      apply(g, l, P.D, flags, fstride, s);
      cheb(g, l, P.D, i, 1, 0, flags, fstride, s);
    }
  }
}

}  // namespace vfg
