// vfg_mg.hip -- geometric multigrid pieces for the u-solve preconditioner.
//
// The reference preconditions its u-CG with hypre BoomerAMG (PCHYPRE,
// VFCommon.c:1041-1047) on the assembled Jacobian; its large runs use PETSc
// GAMG with Chebyshev smoothers (-U_pc_type gamg, scripts/test16-*.sh).  On
// the structured grid of this path the same role is played matrix-free by
// geometric multigrid: 2:1 coarsening of the element grid, coarse operators
// by rediscretisation, trilinear prolongation and its transpose.
//
//   level 1 coefficient of coarse element C: the mean over its 8 fine children
//     of E_e * sbar_e, sbar_e = sum_g w_g (v_g^2 + eta) (weights summing to 1);
//   level l+1 from level l: the mean over the 8 children;
//   coarse operator: the Q1 elasticity operator with s = 1 and E = coefficient
//     (the u-apply with V = 1, eta = 0), so every level runs the same kernels.
// Dirichlet rows are excluded from restriction and prolongation on every
// level (their masks come from the same face/vertex table on each grid).
#include "vfg_internal.h"
#include "vfg_kernels.h"

namespace vfg {

// ------------------------------------------------------------ coefficients
__global__ void k_mg_coef0(Grid G, const double *__restrict__ V, const double *__restrict__ E_elem, double E,
                           double eta, Grid Gc, double *__restrict__ coef) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= (int64_t)Gc.nex * Gc.ney * Gc.nez) return;
  const int ci = (int)(c % Gc.nex), cj = (int)((c / Gc.nex) % Gc.ney), ck = (int)(c / ((int64_t)Gc.nex * Gc.ney));
  double acc = 0.;
  for (int ch = 0; ch < 8; ch++) {
    const int ei = 2 * ci + (ch & 1), ej = 2 * cj + ((ch >> 1) & 1), ek = 2 * ck + (ch >> 2);
    double v[8];
    for (int a = 0; a < 8; a++) v[a] = V[node_id(G, ei + (a & 1), ej + ((a >> 1) & 1), ek + (a >> 2))];
    double s = 0.;
    for (int gz = 0; gz < 3; gz++)
      for (int gy = 0; gy < 3; gy++)
        for (int gx = 0; gx < 3; gx++) {
          double vg = 0.;
          for (int a = 0; a < 8; a++)
            vg += ((a & 1) ? gP1(gx) : gP0(gx)) * (((a >> 1) & 1) ? gP1(gy) : gP0(gy)) *
                  ((a >> 2) ? gP1(gz) : gP0(gz)) * v[a];
          s += gW(gx) * gW(gy) * gW(gz) * (vg * vg + eta);
        }
    acc += (E_elem ? E_elem[cell_id(G, ei, ej, ek)] : E) * s;
  }
  coef[c] = acc * .125;
}

__global__ void k_mg_coarsen(Grid Gf, const double *__restrict__ cf, Grid Gc, double *__restrict__ cc) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= (int64_t)Gc.nex * Gc.ney * Gc.nez) return;
  const int ci = (int)(c % Gc.nex), cj = (int)((c / Gc.nex) % Gc.ney), ck = (int)(c / ((int64_t)Gc.nex * Gc.ney));
  double acc = 0.;
  for (int ch = 0; ch < 8; ch++)
    acc += cf[cell_id(Gf, 2 * ci + (ch & 1), 2 * cj + ((ch >> 1) & 1), 2 * ck + (ch >> 2))];
  cc[c] = acc * .125;
}

// ------------------------------------------------- restriction / prolongation
// coarse rhs = P^T (b - y) on the free fine rows; 27 fine nodes per coarse node
// with weights (1/2, 1, 1/2) per direction; zero on coarse Dirichlet rows
__global__ __launch_bounds__(256) void k_mg_restrict(Grid Gf, Grid Gc, BcMask bc, const float *__restrict__ b,
                                                     const float *__restrict__ y, float *__restrict__ bcoarse) {
  const int64_t N = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (N >= (int64_t)Gc.nx * Gc.ny * Gc.nz) return;
  const int I = (int)(N % Gc.nx), J = (int)((N / Gc.nx) % Gc.ny), K = (int)(N / ((int64_t)Gc.nx * Gc.ny));
  float acc[3] = {0.f, 0.f, 0.f};
  for (int dk = -1; dk <= 1; dk++) {
    const int k = 2 * K + dk;
    if (k < 0 || k >= Gf.nz) continue;
    for (int dj = -1; dj <= 1; dj++) {
      const int j = 2 * J + dj;
      if (j < 0 || j >= Gf.ny) continue;
      for (int di = -1; di <= 1; di++) {
        const int i = 2 * I + di;
        if (i < 0 || i >= Gf.nx) continue;
        const float w = (di ? .5f : 1.f) * (dj ? .5f : 1.f) * (dk ? .5f : 1.f);
        const int64_t n = node_id(Gf, i, j, k);
        for (int c = 0; c < 3; c++) {
          if (on_dirichlet(Gf, bc.face[c], bc.vert[c], i, j, k)) continue;
          acc[c] += w * (b[3 * n + c] - y[3 * n + c]);
        }
      }
    }
  }
  for (int c = 0; c < 3; c++)
    bcoarse[3 * N + c] = on_dirichlet(Gc, bc.face[c], bc.vert[c], I, J, K) ? 0.f : acc[c];
}

// x_fine += P x_coarse on the free fine rows (trilinear interpolation)
__global__ __launch_bounds__(256) void k_mg_prolong(Grid Gf, Grid Gc, BcMask bc, const float *__restrict__ xc,
                                                    float *__restrict__ xf) {
  const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= (int64_t)Gf.nx * Gf.ny * Gf.nz) return;
  const int i = (int)(n % Gf.nx), j = (int)((n / Gf.nx) % Gf.ny), k = (int)(n / ((int64_t)Gf.nx * Gf.ny));
  const int i0 = i >> 1, j0 = j >> 1, k0 = k >> 1;
  const int ni = (i & 1) ? 2 : 1, nj = (j & 1) ? 2 : 1, nk = (k & 1) ? 2 : 1;
  const float wi = (i & 1) ? .5f : 1.f, wj = (j & 1) ? .5f : 1.f, wk = (k & 1) ? .5f : 1.f;
  float acc[3] = {0.f, 0.f, 0.f};
  for (int c3 = 0; c3 < nk; c3++)
    for (int c2 = 0; c2 < nj; c2++)
      for (int c1 = 0; c1 < ni; c1++) {
        const int64_t N = node_id(Gc, i0 + c1, j0 + c2, k0 + c3);
        const float w = wi * wj * wk;
        for (int c = 0; c < 3; c++) acc[c] += w * xc[3 * N + c];
      }
  for (int c = 0; c < 3; c++)
    if (!on_dirichlet(Gf, bc.face[c], bc.vert[c], i, j, k)) xf[3 * n + c] += acc[c];
}

// ----------------------------------------------------------------- smoother
// one Chebyshev step on the Jacobi-scaled system:
//   res = b - y (y = A x; y == nullptr: x = 0, res = b)
//   d = c1 d + c2 D^-1 res   (dfirst: d = c2 D^-1 res)
//   x = x + d                (xzero: x = d)
__global__ void k_mg_cheb(const float *__restrict__ b, const float *__restrict__ y, const float *__restrict__ dinv,
                          float *__restrict__ d, float *__restrict__ x, float c1, float c2, int dfirst, int xzero,
                          int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float res = y ? b[i] - y[i] : b[i];
  const float di = (dfirst ? 0.f : c1 * d[i]) + c2 * dinv[i] * res;
  d[i] = di;
  x[i] = xzero ? di : x[i] + di;
}

__global__ void k_mg_recip32(const double *__restrict__ d, float *__restrict__ dinv, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dinv[i] = d[i] != 0. ? (float)(1. / d[i]) : 1.f;
}

__global__ void k_mg_fill32(float *x, float v, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = v;
}

// deterministic start vector in [-1, 1) for the eigenvalue estimates
__global__ void k_mg_seed(float *x, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint64_t z = (uint64_t)i * 0x9E3779B97F4A7C15ull + 0x632BE59BD9B4E019ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  x[i] = (float)((double)(z >> 11) * (2.0 / 9007199254740992.0) - 1.0);
}

static unsigned nbm(int64_t n) { return (unsigned)((n + 255) / 256); }

void launch_mg_coef0(const DevProblem &Pf, const Grid &Gc, double *coef, hipStream_t s) {
  const int64_t ne = (int64_t)Gc.nex * Gc.ney * Gc.nez;
  hipLaunchKernelGGL(k_mg_coef0, dim3(nbm(ne)), dim3(256), 0, s, Pf.G, (const double *)Pf.V,
                     (const double *)Pf.E_elem, Pf.E, Pf.eta, Gc, coef);
}
void launch_mg_coarsen(const Grid &Gf, const double *cf, const Grid &Gc, double *cc, hipStream_t s) {
  const int64_t ne = (int64_t)Gc.nex * Gc.ney * Gc.nez;
  hipLaunchKernelGGL(k_mg_coarsen, dim3(nbm(ne)), dim3(256), 0, s, Gf, cf, Gc, cc);
}
void launch_mg_restrict(const Grid &Gf, const Grid &Gc, const BcMask &bc, const float *b, const float *y,
                        float *bcoarse, hipStream_t s) {
  const int64_t nn = (int64_t)Gc.nx * Gc.ny * Gc.nz;
  hipLaunchKernelGGL(k_mg_restrict, dim3(nbm(nn)), dim3(256), 0, s, Gf, Gc, bc, b, y, bcoarse);
}
void launch_mg_prolong(const Grid &Gf, const Grid &Gc, const BcMask &bc, const float *xc, float *xf,
                       hipStream_t s) {
  const int64_t nn = (int64_t)Gf.nx * Gf.ny * Gf.nz;
  hipLaunchKernelGGL(k_mg_prolong, dim3(nbm(nn)), dim3(256), 0, s, Gf, Gc, bc, xc, xf);
}
void launch_mg_cheb(const float *b, const float *y, const float *dinv, float *d, float *x, float c1, float c2,
                    int dfirst, int xzero, int64_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_mg_cheb, dim3(nbm(n)), dim3(256), 0, s, b, y, dinv, d, x, c1, c2, dfirst, xzero, n);
}
void launch_mg_recip32(const double *d, float *dinv, int64_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_mg_recip32, dim3(nbm(n)), dim3(256), 0, s, d, dinv, n);
}
void launch_mg_fill32(float *x, float v, int64_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_mg_fill32, dim3(nbm(n)), dim3(256), 0, s, x, v, n);
}
void launch_mg_seed(float *x, int64_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_mg_seed, dim3(nbm(n)), dim3(256), 0, s, x, n);
}

}  // namespace vfg
