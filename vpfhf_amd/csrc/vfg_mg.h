// vfg_mg.h -- grid-transfer device code shared by the single-problem and the
// batched multigrid (internal).
#pragma once
#include "vfg_internal.h"

namespace vfg {

// Coarse grids cover the same box with ceil(n/2) elements per direction
// (directions already at the minimum keep theirs), so the grids are nested
// only where n is even.  Everything between two levels is built from the 1-D
// map xi(i) = i * nc / nf of fine node i onto the coarse node axis:
//   prolongation: trilinear interpolation at xi, weight 1 - |xi - I| for the
//     two coarse nodes I bracketing xi;
//   restriction: its transpose (the same weights, gathered per coarse node);
//   coefficients: the overlap-weighted mean of the fine elements under each
//     coarse element.
__device__ __forceinline__ double mg_xi(int i, int nf, int nc) { return (double)i * (double)nc / (double)nf; }
__device__ __forceinline__ float mg_w(double xi, int I) {
  const double d = fabs(xi - (double)I);
  return d < 1. ? (float)(1. - d) : 0.f;
}

// ------------------------------------------------- restriction / prolongation
// fine nodes i of one axis that coarse node I gathers from, with weights
__device__ __forceinline__ int mg_gather(int I, int nf, int nc, int *i0, float *w) {
  const int lo = max(0, (int)floor((double)(I - 1) * nf / nc));
  const int hi = min(nf, (int)ceil((double)(I + 1) * nf / nc));
  int m = 0;
  for (int i = lo; i <= hi && m < 6; i++) {
    const float wi = mg_w(mg_xi(i, nf, nc), I);
    if (wi > 0.f) {
      if (m == 0) *i0 = i;
      w[m++] = wi;
    }
  }
  return m;
}

// coarse rhs entry (node N, component c) = P^T (b - y) over the free fine rows
// (y = nullptr: b already holds the residual); zero on coarse Dirichlet rows
__device__ __forceinline__ float mg_restrict_dof(const Grid &Gf, const Grid &Gc, const BcMask &bc,
                                                 const float *__restrict__ b, const float *__restrict__ y,
                                                 int64_t N, int c) {
  const int I = (int)(N % Gc.nx), J = (int)((N / Gc.nx) % Gc.ny), K = (int)(N / ((int64_t)Gc.nx * Gc.ny));
  if (on_dirichlet(Gc, bc.face[c], bc.vert[c], I, J, K)) return 0.f;
  int i0 = 0, j0 = 0, k0 = 0;
  float wx[6], wy[6], wz[6];
  const int mx = mg_gather(I, Gf.nex, Gc.nex, &i0, wx);
  const int my = mg_gather(J, Gf.ney, Gc.ney, &j0, wy);
  const int mz = mg_gather(K, Gf.nez, Gc.nez, &k0, wz);
  float acc = 0.f;
  for (int kk = 0; kk < mz; kk++)
    for (int jj = 0; jj < my; jj++) {
      float row = 0.f;
      for (int ii = 0; ii < mx; ii++) {
        const int i = i0 + ii, j = j0 + jj, k = k0 + kk;
        if (on_dirichlet(Gf, bc.face[c], bc.vert[c], i, j, k)) continue;
        const int64_t n = 3 * node_id(Gf, i, j, k) + c;
        row += wx[ii] * (y ? b[n] - y[n] : b[n]);
      }
      acc += wy[jj] * wz[kk] * row;
    }
  return acc;
}

// The same entry split over the fine planes: lane kk of four sums planes
// k0 + kk, k0 + kk + 4, ...; the four lanes' values add up to the entry.
__device__ __forceinline__ float mg_restrict_dof_plane(const Grid &Gf, const Grid &Gc, const BcMask &bc,
                                                       const float *__restrict__ b, const float *__restrict__ y,
                                                       int64_t N, int c, int kk) {
  const int I = (int)(N % Gc.nx), J = (int)((N / Gc.nx) % Gc.ny), K = (int)(N / ((int64_t)Gc.nx * Gc.ny));
  if (on_dirichlet(Gc, bc.face[c], bc.vert[c], I, J, K)) return 0.f;
  int i0 = 0, j0 = 0, k0 = 0;
  float wx[6], wy[6], wz[6];
  const int mx = mg_gather(I, Gf.nex, Gc.nex, &i0, wx);
  const int my = mg_gather(J, Gf.ney, Gc.ney, &j0, wy);
  const int mz = mg_gather(K, Gf.nez, Gc.nez, &k0, wz);
  float acc = 0.f;
  for (int kp = kk; kp < mz; kp += 4)
    for (int jj = 0; jj < my; jj++) {
      float row = 0.f;
      for (int ii = 0; ii < mx; ii++) {
        const int i = i0 + ii, j = j0 + jj, k = k0 + kp;
        if (on_dirichlet(Gf, bc.face[c], bc.vert[c], i, j, k)) continue;
        const int64_t n = 3 * node_id(Gf, i, j, k) + c;
        row += wx[ii] * (y ? b[n] - y[n] : b[n]);
      }
      acc += wy[jj] * wz[kp] * row;
    }
  return acc;
}

// x_fine(node n) += P x_coarse on the free fine rows
__device__ __forceinline__ void mg_prolong_node(const Grid &Gf, const Grid &Gc, const BcMask &bc,
                                                const float *__restrict__ xc, float *__restrict__ xf, int64_t n) {
  const int i = (int)(n % Gf.nx), j = (int)((n / Gf.nx) % Gf.ny), k = (int)(n / ((int64_t)Gf.nx * Gf.ny));
  const double xi = mg_xi(i, Gf.nex, Gc.nex), yj = mg_xi(j, Gf.ney, Gc.ney), zk = mg_xi(k, Gf.nez, Gc.nez);
  const int I0 = min((int)floor(xi), Gc.nex - 1), J0 = min((int)floor(yj), Gc.ney - 1),
            K0 = min((int)floor(zk), Gc.nez - 1);
  const float wx[2] = {mg_w(xi, I0), mg_w(xi, I0 + 1)}, wy[2] = {mg_w(yj, J0), mg_w(yj, J0 + 1)},
              wz[2] = {mg_w(zk, K0), mg_w(zk, K0 + 1)};
  float acc[3] = {0.f, 0.f, 0.f};
  for (int c3 = 0; c3 < 2; c3++)
    for (int c2 = 0; c2 < 2; c2++)
      for (int c1 = 0; c1 < 2; c1++) {
        const float w = wx[c1] * wy[c2] * wz[c3];
        if (w == 0.f) continue;
        const int64_t N = node_id(Gc, I0 + c1, J0 + c2, K0 + c3);
        for (int c = 0; c < 3; c++) acc[c] += w * xc[3 * N + c];
      }
  for (int c = 0; c < 3; c++)
    if (!on_dirichlet(Gf, bc.face[c], bc.vert[c], i, j, k)) xf[3 * n + c] += acc[c];
}

// one Chebyshev step on the Jacobi-scaled system, entry i:
//   res = b - y (use_y) or b;  d = c1 d + c2 D^-1 res (dfirst: c2 D^-1 res);
//   x = x + d (xzero: x = d)
__device__ __forceinline__ void mg_cheb_entry(const float *__restrict__ b, const float *__restrict__ y,
                                              const float *__restrict__ dinv, float *__restrict__ d,
                                              float *__restrict__ x, float c1, float c2, int dfirst, int xzero,
                                              int64_t i) {
  const float res = y ? b[i] - y[i] : b[i];
  const float di = (dfirst ? 0.f : c1 * d[i]) + c2 * dinv[i] * res;
  d[i] = di;
  x[i] = xzero ? di : x[i] + di;
}

}  // namespace vfg
