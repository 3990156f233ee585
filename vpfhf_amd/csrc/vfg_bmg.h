// vfg_bmg.h -- batched multigrid V-cycle over the members of an ensemble (internal).
#pragma once
#include <vector>
#include <hip/hip_runtime.h>

struct vfg_handle;

namespace vfg {

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
