// vfg_batch_apply.h -- persistent multi-problem apply (internal; kernels in vfg_u.hip).
#pragma once
#include <vector>
#include "vfg_kernels.h"

struct vfg_handle;

namespace vfg {

size_t batch_op_bytes(int fp32);
// persistent-apply schedules: 0: 16 x 16 tiles, register march; 1: 8 x 8
// tiles, register march; 2: 16 x 16 tiles, LDS-ring march
constexpr int BATCH_SCHEDS = 3;
int batch_tile(int sched);             // element tile width (owned node columns + 1)
int batch_slots_s(int fp32, int sched);  // resident workgroups of that apply on the device
int batch_slots();                     // schedule 0, FP64
int choose_oz_pub(const Grid &G, int tiles_xy, int slots);
// writes one {UApplyOp, OZ} record per member (x[p] -> y[p]) into d_ops
void fill_ops(const std::vector<vfg_handle *> &m, const void *const *x, void *const *y, int fp32,
              const std::vector<int> &oz, void *d_ops, hipStream_t s);
// grid == 0: one workgroup per item; otherwise a persistent grid striding the items
void launch_batch_apply(int fp32, int sched, const void *d_ops, const BItem *items, int nitems, int grid,
                        hipStream_t s);

// persistent applies over (device-counted, when dcount is given) work lists;
// items of members whose flags[p * fstride] is 0 are skipped
void launch_engine_u_apply(int fp32, int sched, const void *d_ops, const BItem *items, int nitems,
                           const int *dcount, int grid, const double *flags, int fstride, hipStream_t s);
void make_uop(const DevProblem &P, const void *x, void *y, int fp32, int OZ, void *out);  // host record
// multigrid smoothing fused with the level apply, one record per member
// (launch_mg_presmooth / launch_mg_postsmooth of every listed member at once;
// 16 x 16 tiles, register march): post = 0: x = c2 D^-1 b into xo, A x into
// ro; post = 1: xo = xin + c2 D^-1 (b - A xin)
size_t mg_smooth_op_bytes(int post);
void make_mg_smooth_op(const DevProblem &P, int post, const float *b, const float *dinv, float c2, const float *xin,
                       float *xo, float *ro, int OZ, void *out);  // host record
void launch_batch_mg_smooth(int post, const void *d_ops, const BItem *items, const int *dcount, int grid,
                            const double *flags, int fstride, hipStream_t s);
size_t vop_bytes();
int v_oz(const Grid &G);
void make_vop(const DevProblem &P, const double *x, double *y, const uint8_t *mask, void *out);
void launch_engine_v_apply(const void *d_ops, const BItem *items, const int *dcount, int grid, const double *flags,
                           int fstride, hipStream_t s);

void batch_member_sums(const double *part, const int *first, const int *nch, int np, double *out, hipStream_t s);

}  // namespace vfg
