// vfg_batch_apply.h -- persistent multi-problem apply (internal; kernels in vfg_u.hip).
#pragma once
#include <vector>
#include "vfg_kernels.h"

struct vfg_handle;

namespace vfg {

size_t batch_op_bytes(int fp32);
int batch_slots();
int choose_oz_pub(const Grid &G, int tiles_xy, int slots);
// writes one {UApplyOp, OZ} record per member (x[p] -> y[p]) into d_ops
void fill_ops(const std::vector<vfg_handle *> &m, const void *const *x, void *const *y, int fp32,
              const std::vector<int> &oz, void *d_ops, hipStream_t s);
// grid == 0: one workgroup per item; otherwise a persistent grid striding the items
void launch_batch_apply(int fp32, const void *d_ops, const BItem *items, int nitems, int grid, hipStream_t s);

void batch_member_sums(const double *part, const int *first, const int *nch, int np, double *out, hipStream_t s);

}  // namespace vfg
