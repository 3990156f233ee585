// vfg_handle.h -- the opaque vfg_handle (internal).
#pragma once
#include <string>
#include <vector>
#include "vfg_kernels.h"

namespace vfg {
// Multigrid hierarchy of the u-operator (allocated once per handle, its
// coefficients refreshed at every u-solve).  Level 0 is the problem itself;
// level l+1 halves the element counts of level l.
struct MgLevel {
  DevProblem P;                     // level operator (level >= 1: E_elem = coefficient, V = 1, eta = 0)
  int64_t n = 0;                    // 3 * nodes
  float *dinv = nullptr;            // FP32 inverse diagonal
  float *b = nullptr, *x = nullptr; // rhs / iterate (level 0: the caller's arrays)
  float *y = nullptr, *d = nullptr; // A x, Chebyshev direction
  double lmax = 0.;                 // upper bound of the D^-1 A spectrum used by the smoother
};
struct MgHier {
  static constexpr int MAXL = 12;
  int L = 0;
  MgLevel lv[MAXL];
  int setups = 0;          // smoother bounds are re-estimated on every 16th set-up
  double *dtmp = nullptr;  // FP64 diagonal of level 1 (largest coarse level)
};
}  // namespace vfg

struct vfg_handle {
  vfg::DevProblem P;
  vfg_grid grid;
  vfg_props props;
  hipStream_t stream = nullptr;
  std::string err;
  std::vector<void *> allocs;
  size_t bytes = 0;
  double *stage = nullptr;  // staging for host-pointer calls
  int64_t stage_n = 0;
  double *part = nullptr;      // reduction partials (4 x 65536)
  double *part_big = nullptr;  // partials for very large grids
  double *scal = nullptr;      // device scalars (64)
  double *abh = nullptr;       // CG coefficient history of an eigenvalue estimate (3 x 64)
  uint32_t *vflag = nullptr;   // V-solve: node-plane chunks holding free nodes (v_chunk_cap flags)
  int *vlist = nullptr;        //          and their list, count and chunk length (v_chunk_cap + 2)
  uint32_t *vflag2 = nullptr;  // the same dilated by one tile and one plane,
  int *vlist2 = nullptr;       //          and its list
  bool v_ready = false;
  // solver work space (allocated on first use)
  double *w[16] = {nullptr};
  float *wf[8] = {nullptr};
  uint8_t *act = nullptr;
  vfg::MgHier *mg = nullptr;  // u-solve multigrid (VFG_PC_MG)
  // FP64 u-apply form (VFG_U_FORM_*); the stored form keeps its integrand
  // weights in uq_buf, valid while P.uq is set (cleared when V, E or eta change)
  int u_form = 0;  // VFG_U_FORM_RECOMPUTE
  double *uq_buf = nullptr;
};

namespace vfg {
// multigrid (vfg_solve.hip): V-cycle smoothing degree, coarsest-level degree,
// smallest coarse element count per direction
constexpr int MG_SMOOTH = 1, MG_COARSE_DEG = 6, MG_COARSEST = 8;
int mg_build(vfg_handle *h);                      // hierarchy shape and buffers (once)
int mg_setup(vfg_handle *h, const double *dinv64);  // coefficients, diagonals, bounds (per solve)
}  // namespace vfg

int vfg_fail(vfg_handle *h, int code, const char *fmt, ...);
// the stored form's weights for the current coefficients (no-op for the
// recomputing form and for NOCOMPRESSION problems); u_invalidate drops them
int vfg_u_prepare(vfg_handle *h);
inline void vfg_u_invalidate(vfg_handle *h) { h->P.uq = nullptr; }
int vfg_alloc(vfg_handle *h, void **p, size_t bytes);
int vfg_scratch(vfg_handle *h, int64_t ndoubles);
// solver work vectors of 3*nn entries: w[0..n64) FP64 and wf[0..n32) FP32
int vfg_work(vfg_handle *h, int n64, int n32);
int vfg_in(vfg_handle *h, void *dst, const void *src, size_t bytes, int loc);
