// vfg_kernels.h -- device problem description and kernel launchers (internal).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "vfg_internal.h"
#include "../../include/vfg.h"

namespace vfg {

// one work item of the persistent multi-problem apply
struct BItem {
  int p, bx, by, bz;
};

// Everything a kernel needs about one problem.  Device pointers are owned by
// the handle (vfg_api.hip); host copies of the BC tables travel by value.
struct DevProblem {
  Grid G;
  int64_t nn, ne;
  // properties (VFMatProp/VFProp); lam1/mu1 are the Lame constants for E = 1
  double E, nu, lam1, mu1, alpha, beta, Gc, eta, eps, irrevtol, atcv;
  int atnum, hasCrackPressure;
  int unilateral;   // VFG_UNILATERAL_*
  uint32_t *cmask;  // per element: Gauss points where tr(eps(U_J)) < 0 (Jacobian state)
  BcMask bcU, bcV;
  vfg_bc hbcU[3], hbcV;
  // fields
  double *V, *VIrrev, *dth, *p, *E_elem;  // E_elem may be null (uniform E)
  float *V32;                             // FP32 copy of V for the FP32 apply
  double *zero3;                          // 3*nn zeros
  // V-subproblem state, refreshed by vfg_v_set_u
  double *uq;      // FP64 u-apply, stored form: W_g s_g E_e vol per element (uq_doubles); null: recomputed
  double *Vq;      // W_g * (c_AT/eps [AT2] + 2 W^el_g), tile-blocked (vq_doubles)
  double *bV;      // V load vector (nn)
  uint8_t *vfix;   // fixed V rows/cols (Dirichlet + AT2 irreversibility)
};

// u operator (vfg_u.hip)
// size of the stored integrand weights: node-owned tiles x element layers x 27 x 256
int64_t uq_doubles(const Grid &G);
void launch_u_qsetup(const DevProblem &P, double *uq, hipStream_t s);
void launch_u_apply(const DevProblem &P, const double *x, double *y, hipStream_t s);
void launch_u_apply_f32(const DevProblem &P, const float *x, float *y, hipStream_t s);
// FP32 multigrid level: one Chebyshev-Jacobi step from x = 0 fused with its
// apply (x = c2 D^-1 b, y = A x: the restriction forms b - y on the free
// rows), and one step from x fused with its apply (xo = x + c2 D^-1 (b - A x));
// xo must not alias x.  With r64 the post-smoothing also writes one partial of
// r64 . xo per workgroup to part and returns their number (else 0 partials).
void launch_mg_presmooth(const DevProblem &P, const float *b, const float *dinv, float c2, float *x, float *y,
                         hipStream_t s);
int launch_mg_postsmooth(const DevProblem &P, const float *b, const float *dinv, float c2, const float *x, float *xo,
                         const double *r64, double *part, hipStream_t s);
// FP64 y = A x (no NOCOMPRESSION) with one x.y partial per workgroup in part;
// returns the number of partials (0, and nothing launched, for NOCOMPRESSION)
int launch_u_apply_dot(const DevProblem &P, const double *x, double *y, double *part, hipStream_t s);
// Chebyshev(k)-Jacobi from x = 0 on a small FP32 level in one launch (k <= 8;
// step i: coefficients c1[i], c2[i] of mg_cheb_entry); mg_coarse_fits tells
// whether the level qualifies (at most 8^3 elements, no NOCOMPRESSION)
struct MgChebCoefs {
  float c1[8], c2[8];
  int k;
};
bool mg_coarse_fits(const DevProblem &P);
void launch_mg_coarse_cheb(const DevProblem &P, const float *b, const float *dinv, float *x, const MgChebCoefs &cc,
                           hipStream_t s);
void launch_u_diag(const DevProblem &P, double *d, hipStream_t s);
// cmask from the state U (UNILATERAL_NOCOMPRESSION)
void launch_u_cmask(const DevProblem &P, const double *U, hipStream_t s);
// mode 0: out = element load vector b; mode 1: out = VF_UResidual(U);
// mode 2 (UNILATERAL_NONE only): b on the free rows and, on the Dirichlet
// rows, the value ResidualApplyDirichletBC subtracts from U there -- so that
// VF_UResidual(U) = A U - out with A the MatShell operator (identity rows)
void launch_u_rhs(const DevProblem &P, const double *U, const double *BCU, double *out, int mode,
                  hipStream_t s);

// V operator (vfg_v.hip)
void launch_v_setup(const DevProblem &P, const double *U, hipStream_t s);  // Vq, bV
// Vq size: tile-blocked [tile][element layer][27][256] (vfg_v.hip)
int64_t vq_doubles(const Grid &G);
void launch_v_fixmask(const DevProblem &P, hipStream_t s);                 // vfix
// mode 0: masked Jacobian apply; mode 1: residual K x - bV with BC rows
// mode 2: plain K x (no masks, no BC).  `mask` overrides vfix (reduced systems).
void launch_v_apply(const DevProblem &P, const double *x, double *y, int mode, hipStream_t s,
                    const uint8_t *mask = nullptr);
void launch_v_diag(const DevProblem &P, double *d, hipStream_t s, const uint8_t *mask = nullptr);
// Chunks of the V kernels: node planes [zc L, zc L + L) of a 15 x 15 node
// tile, id tile + txy zc.  launch_v_chunk_list lists, in increasing id, the
// chunks holding a node with mask == 0 into list[0, count), with count in
// list[cap] and L in list[cap + 1] (cap = v_chunk_cap, flag holds cap words).
int v_chunk_cap(const Grid &G);
void launch_v_chunk_list(const DevProblem &P, const uint8_t *mask, uint32_t *flag, int *list, hipStream_t s);
// the same list for the flags of launch_v_chunk_list dilated by one tile and
// one node plane (flag2: v_chunk_cap words): the chunks whose nodes share an
// element with a listed node, where K x changes when x changes on the listed
// nodes only
void launch_v_dilated_list(const DevProblem &P, const uint32_t *flag, uint32_t *flag2, int *list2, hipStream_t s);
// CG-fused masked V apply (mask: fixed or active nodes) over the listed
// chunks: p = z + beta p_old staged, q = K p, p and q written on the chunks'
// owned nodes (p, q must be zero on the other chunks), one partial of p.q per
// chunk; returns the number of partials (= count)
int launch_v_apply_cg_list(const DevProblem &P, const double *pold, double *p, const double *z, const double *beta,
                           double *q, double *part, const uint8_t *mask, const int *list, int count, hipStream_t s,
                           const double *done = nullptr);
// launch_v_apply and launch_v_diag over the listed chunks' owned nodes (the
// other entries of y, d are left as they are)
void launch_v_apply_list(const DevProblem &P, const double *x, double *y, int mode, const int *list, int count,
                         hipStream_t s, const uint8_t *mask = nullptr);
void launch_v_diag_list(const DevProblem &P, double *d, const uint8_t *mask, const int *list, int count,
                        hipStream_t s);
// the CG update x += a p, r -= a q, z = dinv r over the listed chunks, with
// (r.z, r.r) partial pairs; returns their number (= count)
int launch_v_xr_list(const Grid &G, const int *list, int count, double *x, double *r, double *z, const double *p,
                     const double *q, const double *dinv, const double *alpha, double *part, hipStream_t s,
                     const double *done = nullptr);
// (both: the launch does nothing once *done != 0, a device-side stop of the CG)

// misc (vfg_misc.hip)
void launch_bc_mask(const DevProblem &P, int which, uint8_t *mask, hipStream_t s);
void launch_vec_set_from_bc(const DevProblem &P, int which, double *B, hipStream_t s);
void launch_vec_apply_dirichlet(const DevProblem &P, int which, double *vec, const double *B,
                                hipStream_t s);
// partial sums per block: out[b*4 + {elastic, pressure work, surface, crack volume}]
int energy_blocks(const DevProblem &P);
void launch_energies(const DevProblem &P, const double *U, const double *V, double *partials,
                     double *cellfield, double *efield, double *result4, hipStream_t s);
// geometric multigrid pieces (vfg_mg.hip)
void launch_mg_elem_coef(const DevProblem &P, double *out, hipStream_t s);
void launch_mg_coarsen(const Grid &Gf, const double *cf, const Grid &Gc, double *cc, hipStream_t s);
// tmp (optional, mg_restrict_tmp_floats entries): work space of the
// separable restriction between nested large levels
void launch_mg_restrict(const Grid &Gf, const Grid &Gc, const BcMask &bc, const float *b, const float *y,
                        float *bcoarse, hipStream_t s, float *tmp = nullptr);
int64_t mg_restrict_tmp_floats(const Grid &Gf, const Grid &Gc);
void launch_mg_prolong(const Grid &Gf, const Grid &Gc, const BcMask &bc, const float *xc, float *xf,
                       hipStream_t s);
void launch_mg_cheb(const float *b, const float *y, const float *dinv, float *d, float *x, float c1, float c2,
                    int dfirst, int xzero, int64_t n, hipStream_t s);
void launch_mg_recip32(const double *d, float *dinv, int64_t n, hipStream_t s);
void launch_mg_fill32(float *x, float v, int64_t n, hipStream_t s);
void launch_mg_seed(float *x, int64_t n, hipStream_t s);
void launch_cell_to_node(const DevProblem &P, double lx, double ly, double lz, const double *cell, double *node,
                         int dof, hipStream_t s);
void launch_penny(const DevProblem &P, double *V, const double c[3], double r, double theta,
                  double phi, double thick, hipStream_t s);

// vector kernels for the solvers (vfg_misc.hip); all reductions deterministic
int red_blocks(int64_t n);
void launch_dot(const double *a, const double *b, int64_t n, double *partials, double *out,
                hipStream_t s);
void launch_fill(double *x, double v, int64_t n, hipStream_t s);
void launch_sub(const double *a, const double *b, double *out, int64_t n, hipStream_t s);
void launch_vmin(const double *a, const double *b, double *o, int64_t n, hipStream_t s);
void launch_scale(double *x, double v, int64_t n, hipStream_t s);
void launch_maxabs_diff(const double *a, const double *b, int64_t n, double *part, double *out, hipStream_t s);
void launch_copy_d2f(const double *x, float *y, int64_t n, hipStream_t s);
void launch_copy_f2d(const float *x, double *y, int64_t n, hipStream_t s);

// V-solve pieces (vfg_solve.hip): projection onto [0, ub] off the fixed nodes;
// active set + r = -g on the inactive set; projected step Vn = P(V + lam d),
// D = Vn - V; g = 0 on fixed nodes
void launch_v_clamp(double *V, const uint8_t *fix, const double *ub, int64_t n, hipStream_t s);
void launch_v_active(const double *V, const double *g, const uint8_t *fix, const double *ub, uint8_t *act,
                     double *r, int64_t n, double *part, hipStream_t s);
void launch_v_step(const double *V, const double *d, const uint8_t *act, const uint8_t *fix, const double *ub,
                   double lam, double *Vn, double *D, int64_t n, hipStream_t s);
void launch_zero_fixed(double *g, const uint8_t *fix, int64_t n, hipStream_t s);

}  // namespace vfg
