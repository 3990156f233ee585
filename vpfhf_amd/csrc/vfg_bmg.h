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
void bmg_destroy(BatchMG *g);
// per solve: every member's coefficients, diagonals and smoother bounds
// (dinv64[p]: the member's FP64 inverse diagonal), then the device tables
int bmg_setup(BatchMG *g, const double *const *dinv64, hipStream_t s);
// z = V-cycle(r) for every member with flags[p * fstride] != 0
void bmg_vcycle(BatchMG *g, const double *flags, int fstride, hipStream_t s);

}  // namespace vfg
