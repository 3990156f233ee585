// vfg_bmg.h -- batched multigrid V-cycle over the members of an ensemble (internal).
#pragma once
#include <vector>
#include <hip/hip_runtime.h>
#include "vfg_handle.h"

namespace vfg {

constexpr int BML = MgHier::MAXL;
constexpr int BMK = MG_COARSE_DEG > MG_SMOOTH ? MG_COARSE_DEG : MG_SMOOTH;

// device tables of the batched V-cycle: a member's level vectors, work-list
// chunks (member, first entry, count) and Chebyshev step coefficients
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

// the coarsest-level Chebyshev(MG_COARSE_DEG) solve, from x = 0, of every
// member listed in ch[0, *dn) whose chunk starts its vector, one workgroup
// each with the level in LDS (levels that satisfy mg_coarse_fits); ops: the
// level's FP32 apply records, one per member
void launch_bmg_coarse_cheb(const void *ops, const BmgVec *lvt, int level, const BmgChunk *ch, const int *dn,
                            int grid, const BmgCoef *coef, const double *flags, int fstride, hipStream_t s);

struct BatchMG;
// level plans of the members' hierarchies; r32[p] / z32[p]: the member's FP32
// residual (V-cycle input) and preconditioned residual (output)
int bmg_create(const std::vector<vfg_handle *> &m, float *const *r32, float *const *z32, BatchMG **out);
// np member slots, all empty (the serving engine: members come and go)
int bmg_create_slots(int np, BatchMG **out);
void bmg_destroy(BatchMG *g);
// sets (h != nullptr) or clears member p; the work lists change at the next bmg_rebuild
int bmg_set_member(BatchMG *g, int p, vfg_handle *h, float *r32, float *z32);
// rebuilds the work lists of the current members on stream s (synchronised);
// *grew is set when a device list moved or the level count changed, i.e.
// when a graph that captured bmg_vcycle must be captured again
int bmg_rebuild(BatchMG *g, hipStream_t s, bool *grew);
// per solve: every member's coefficients, diagonals and smoother bounds
// (dinv64[p]: the member's FP64 inverse diagonal), then the device tables
int bmg_setup(BatchMG *g, const double *const *dinv64, hipStream_t s);
// the same for member p alone
int bmg_setup_member(BatchMG *g, int p, const double *dinv64, hipStream_t s);
// z = V-cycle(r) for every member with flags[p * fstride] != 0
void bmg_vcycle(BatchMG *g, const double *flags, int fstride, hipStream_t s);

}  // namespace vfg
