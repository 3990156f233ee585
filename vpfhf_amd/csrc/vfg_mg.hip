// vfg_mg.hip -- geometric multigrid pieces for the u-solve preconditioner.
//
// The reference preconditions its u-CG with hypre BoomerAMG (PCHYPRE,
// VFCommon.c:1041-1047) on the assembled Jacobian; its large runs use PETSc
// GAMG with Chebyshev smoothers (-U_pc_type gamg, scripts/test16-*.sh).  On
// the structured grid of this path the same role is played matrix-free by
// geometric multigrid: about 2:1 coarsening of the element grid, coarse operators
// by rediscretisation, trilinear prolongation and its transpose.
//
//   level 1 coefficient of coarse element C: the mean of E_e * sbar_e over the
//     fine elements under C (weighted by overlap), sbar_e = sum_g w_g (v_g^2 +
//     eta) (weights summing to 1); level l+1 from level l likewise;
//   coarse operator: the Q1 elasticity operator with s = 1 and E = coefficient
//     (the u-apply with V = 1, eta = 0), so every level runs the same kernels.
// Dirichlet rows are excluded from restriction and prolongation on every
// level (their masks come from the same face/vertex table on each grid).
#include "vfg_internal.h"
#include "vfg_kernels.h"
#include "vfg_mg.h"

namespace vfg {

// ------------------------------------------------------------ coefficients
// per fine element: E_e * sbar_e, sbar_e = sum_g w_g c_g with c_g = v_g^2 + eta,
// or, at the compressed points of a UNILATERAL_NOCOMPRESSION Jacobian (cmask),
// the factor that gives the isotropic coarse operator the point's P-wave
// modulus: c_g = (kappa (1 + eta) + 4/3 mu s_g) / (lam + 2 mu), kappa =
// lam + 2/3 mu (the spherical part there is not degraded)
__global__ void k_mg_elem_coef(Grid G, const double *__restrict__ V, const double *__restrict__ E_elem, double E,
                               double eta, const uint32_t *__restrict__ cmask, double lam1, double mu1,
                               double *__restrict__ out) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (int64_t)G.nex * G.ney * G.nez) return;
  const int ei = (int)(e % G.nex), ej = (int)((e / G.nex) % G.ney), ek = (int)(e / ((int64_t)G.nex * G.ney));
  double v[8];
  for (int a = 0; a < 8; a++) v[a] = V[node_id(G, ei + (a & 1), ej + ((a >> 1) & 1), ek + (a >> 2))];
  const uint32_t cm = cmask ? cmask[e] : 0u;
  const double kap1 = lam1 + 2. * mu1 / 3., pw = 1. / (lam1 + 2. * mu1);
  double s = 0.;
  for (int gz = 0; gz < 3; gz++)
    for (int gy = 0; gy < 3; gy++)
      for (int gx = 0; gx < 3; gx++) {
        double vg = 0.;
        for (int a = 0; a < 8; a++)
          vg += ((a & 1) ? gP1(gx) : gP0(gx)) * (((a >> 1) & 1) ? gP1(gy) : gP0(gy)) * ((a >> 2) ? gP1(gz) : gP0(gz)) *
                v[a];
        const double sg = vg * vg + eta;
        const bool comp = (cm >> ((gz * 3 + gy) * 3 + gx)) & 1u;
        s += gW(gx) * gW(gy) * gW(gz) * (comp ? (kap1 * (1. + eta) + 4. * mu1 * sg / 3.) * pw : sg);
      }
  out[e] = (E_elem ? E_elem[e] : E) * s;
}

// fine elements e overlapping coarse element C along one axis, with overlaps
__device__ __forceinline__ int mg_overlap(int C, int nf, int nc, int *e0, double *ov) {
  const double a = (double)C * nf / nc, b = (double)(C + 1) * nf / nc;
  const int lo = (int)floor(a);
  int m = 0;
  for (int e = lo; e < nf && (double)e < b && m < 4; e++) {
    const double o = fmin((double)(e + 1), b) - fmax((double)e, a);
    if (o > 0.) {
      if (m == 0) *e0 = e;
      ov[m++] = o;
    }
  }
  return m;
}

__global__ void k_mg_coarsen(Grid Gf, const double *__restrict__ cf, Grid Gc, double *__restrict__ cc) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= (int64_t)Gc.nex * Gc.ney * Gc.nez) return;
  const int ci = (int)(c % Gc.nex), cj = (int)((c / Gc.nex) % Gc.ney), ck = (int)(c / ((int64_t)Gc.nex * Gc.ney));
  int x0 = 0, y0 = 0, z0 = 0;
  double ox[4], oy[4], oz[4];
  const int mx = mg_overlap(ci, Gf.nex, Gc.nex, &x0, ox);
  const int my = mg_overlap(cj, Gf.ney, Gc.ney, &y0, oy);
  const int mz = mg_overlap(ck, Gf.nez, Gc.nez, &z0, oz);
  double acc = 0., wsum = 0.;
  for (int k = 0; k < mz; k++)
    for (int j = 0; j < my; j++)
      for (int i = 0; i < mx; i++) {
        const double w = ox[i] * oy[j] * oz[k];
        acc += w * cf[cell_id(Gf, x0 + i, y0 + j, z0 + k)];
        wsum += w;
      }
  cc[c] = acc / wsum;
}

// one thread per coarse (node, component)
__global__ __launch_bounds__(256) void k_mg_restrict(Grid Gf, Grid Gc, BcMask bc, const float *__restrict__ b,
                                                     const float *__restrict__ y, float *__restrict__ bcoarse) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 3 * (int64_t)Gc.nx * Gc.ny * Gc.nz) return;
  const int64_t N = t / 3;
  bcoarse[t] = mg_restrict_dof(Gf, Gc, bc, b, y, N, (int)(t - 3 * N));
}

// small coarse grids: four lanes per (node, component), each over every
// fourth fine plane (one plane each under ceil(n/2) coarsening), summed in
// lane order
__global__ __launch_bounds__(256) void k_mg_restrict4(Grid Gf, Grid Gc, BcMask bc, const float *__restrict__ b,
                                                      const float *__restrict__ y, float *__restrict__ bcoarse) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x, t = g >> 2;
  const int kk = (int)(g & 3);
  const bool ok = t < 3 * (int64_t)Gc.nx * Gc.ny * Gc.nz;
  const int64_t N = ok ? t / 3 : 0;
  float v = ok ? mg_restrict_dof_plane(Gf, Gc, bc, b, y, N, (int)(t - 3 * N), kk) : 0.f;
  const float v1 = __shfl_down(v, 1, 4), v2 = __shfl_down(v, 2, 4), v3 = __shfl_down(v, 3, 4);
  if (ok && kk == 0) bcoarse[t] = ((v + v1) + v2) + v3;
}

// Nested levels (nf = 2 nc in every direction): the same transfers with the
// weights of the 1-D map xi = i / 2 written out -- fine node 2I gets coarse
// node I with weight 1, fine node 2I + 1 gets I and I + 1 with 1/2 each --
// and one thread per node doing all three components (the general kernels
// form the weights from the map in FP64 per entry).
__device__ __forceinline__ float mg_w2(int d) { return d == 0 ? 1.f : .5f; }
// grid (ceil(nx * ny / 256), nz): node (i, j) of plane blockIdx.y from a
// 32-bit index (no 64-bit division per thread)
__global__ __launch_bounds__(256) void k_mg_restrict2(Grid Gf, Grid Gc, BcMask bc, const float *__restrict__ b,
                                                      const float *__restrict__ y, float *__restrict__ bcoarse) {
  const int t = blockIdx.x * 256 + threadIdx.x;
  if (t >= Gc.nx * Gc.ny) return;
  const int J = t / Gc.nx, I = t - J * Gc.nx, K = blockIdx.y;
  float acc[3] = {0.f, 0.f, 0.f};
  for (int dk = -1; dk <= 1; dk++) {
    const int k = 2 * K + dk;
    if (k < 0 || k >= Gf.nz) continue;
    for (int dj = -1; dj <= 1; dj++) {
      const int j = 2 * J + dj;
      if (j < 0 || j >= Gf.ny) continue;
      float row[3] = {0.f, 0.f, 0.f};
      for (int di = -1; di <= 1; di++) {
        const int i = 2 * I + di;
        if (i < 0 || i >= Gf.nx) continue;
        const int64_t n = 3 * node_id(Gf, i, j, k);
#pragma unroll
        for (int c = 0; c < 3; c++)
          if (!on_dirichlet(Gf, bc.face[c], bc.vert[c], i, j, k))
            row[c] += mg_w2(di) * (y ? b[n + c] - y[n + c] : b[n + c]);
      }
      const float w = mg_w2(dj) * mg_w2(dk);
#pragma unroll
      for (int c = 0; c < 3; c++) acc[c] += w * row[c];
    }
  }
  const int64_t N = node_id(Gc, I, J, K);
#pragma unroll
  for (int c = 0; c < 3; c++) bcoarse[3 * N + c] = on_dirichlet(Gc, bc.face[c], bc.vert[c], I, J, K) ? 0.f : acc[c];
}
__global__ __launch_bounds__(256) void k_mg_prolong2(Grid Gf, Grid Gc, BcMask bc, const float *__restrict__ xc,
                                                     float *__restrict__ xf) {
  const int t = blockIdx.x * 256 + threadIdx.x;
  if (t >= Gf.nx * Gf.ny) return;
  const int j = t / Gf.nx, i = t - j * Gf.nx, k = blockIdx.y;
  const int I0 = i >> 1, J0 = j >> 1, K0 = k >> 1;
  const int mi = i & 1, mj = j & 1, mk = k & 1;  // odd: between I0 and I0 + 1
  const float wx = mi ? .5f : 1.f, wy = mj ? .5f : 1.f, wz = mk ? .5f : 1.f;
  float acc[3] = {0.f, 0.f, 0.f};
  for (int c3 = 0; c3 <= mk; c3++)
    for (int c2 = 0; c2 <= mj; c2++)
      for (int c1 = 0; c1 <= mi; c1++) {
        const float w = wx * wy * wz;
        const int64_t N = 3 * node_id(Gc, I0 + c1, J0 + c2, K0 + c3);
#pragma unroll
        for (int c = 0; c < 3; c++) acc[c] += w * xc[N + c];
      }
  const int64_t n = node_id(Gf, i, j, k);
#pragma unroll
  for (int c = 0; c < 3; c++)
    if (!on_dirichlet(Gf, bc.face[c], bc.vert[c], i, j, k)) xf[3 * n + c] += acc[c];
}
// Nested restriction in three separable passes (large levels): x, then y,
// then z, each a contiguous sweep.  t1 holds (Gc.nx, Gf.ny, Gf.nz) nodes, t2
// (Gc.nx, Gc.ny, Gf.nz); the fine Dirichlet rows are zeroed in the x pass,
// the coarse ones in the z pass.
__global__ __launch_bounds__(256) void k_mg_rsx(Grid Gf, Grid Gc, BcMask bc, const float *__restrict__ b,
                                                const float *__restrict__ y, float *__restrict__ t1) {
  const int t = blockIdx.x * 256 + threadIdx.x;
  if (t >= Gc.nx * Gf.ny) return;
  const int j = t / Gc.nx, I = t - j * Gc.nx, k = blockIdx.y;
  float acc[3] = {0.f, 0.f, 0.f};
  for (int di = -1; di <= 1; di++) {
    const int i = 2 * I + di;
    if (i < 0 || i >= Gf.nx) continue;
    const int64_t n = 3 * node_id(Gf, i, j, k);
#pragma unroll
    for (int c = 0; c < 3; c++)
      if (!on_dirichlet(Gf, bc.face[c], bc.vert[c], i, j, k)) acc[c] += mg_w2(di) * (y ? b[n + c] - y[n + c] : b[n + c]);
  }
  const int64_t o = 3 * ((int64_t)I + (int64_t)Gc.nx * (j + (int64_t)Gf.ny * k));
#pragma unroll
  for (int c = 0; c < 3; c++) t1[o + c] = acc[c];
}
__global__ __launch_bounds__(256) void k_mg_rsy(Grid Gf, Grid Gc, const float *__restrict__ t1, float *__restrict__ t2) {
  const int t = blockIdx.x * 256 + threadIdx.x;
  if (t >= Gc.nx * Gc.ny) return;
  const int J = t / Gc.nx, I = t - J * Gc.nx, k = blockIdx.y;
  float acc[3] = {0.f, 0.f, 0.f};
  for (int dj = -1; dj <= 1; dj++) {
    const int j = 2 * J + dj;
    if (j < 0 || j >= Gf.ny) continue;
    const int64_t n = 3 * ((int64_t)I + (int64_t)Gc.nx * (j + (int64_t)Gf.ny * k));
#pragma unroll
    for (int c = 0; c < 3; c++) acc[c] += mg_w2(dj) * t1[n + c];
  }
  const int64_t o = 3 * ((int64_t)I + (int64_t)Gc.nx * (J + (int64_t)Gc.ny * k));
#pragma unroll
  for (int c = 0; c < 3; c++) t2[o + c] = acc[c];
}
__global__ __launch_bounds__(256) void k_mg_rsz(Grid Gf, Grid Gc, BcMask bc, const float *__restrict__ t2,
                                                float *__restrict__ bcoarse) {
  const int t = blockIdx.x * 256 + threadIdx.x;
  if (t >= Gc.nx * Gc.ny) return;
  const int J = t / Gc.nx, I = t - J * Gc.nx, K = blockIdx.y;
  float acc[3] = {0.f, 0.f, 0.f};
  for (int dk = -1; dk <= 1; dk++) {
    const int k = 2 * K + dk;
    if (k < 0 || k >= Gf.nz) continue;
    const int64_t n = 3 * ((int64_t)I + (int64_t)Gc.nx * (J + (int64_t)Gc.ny * k));
#pragma unroll
    for (int c = 0; c < 3; c++) acc[c] += mg_w2(dk) * t2[n + c];
  }
  const int64_t N = node_id(Gc, I, J, K);
#pragma unroll
  for (int c = 0; c < 3; c++) bcoarse[3 * N + c] = on_dirichlet(Gc, bc.face[c], bc.vert[c], I, J, K) ? 0.f : acc[c];
}
int64_t mg_restrict_tmp_floats(const Grid &Gf, const Grid &Gc) {
  return 3 * ((int64_t)Gc.nx * Gf.ny * Gf.nz + (int64_t)Gc.nx * Gc.ny * Gf.nz);
}
static bool mg_nested(const Grid &Gf, const Grid &Gc) {
  return Gf.nex == 2 * Gc.nex && Gf.ney == 2 * Gc.ney && Gf.nez == 2 * Gc.nez;
}

__global__ __launch_bounds__(256) void k_mg_prolong(Grid Gf, Grid Gc, BcMask bc, const float *__restrict__ xc,
                                                    float *__restrict__ xf) {
  const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= (int64_t)Gf.nx * Gf.ny * Gf.nz) return;
  mg_prolong_node(Gf, Gc, bc, xc, xf, n);
}

// ----------------------------------------------------------------- smoother
__global__ void k_mg_cheb(const float *__restrict__ b, const float *__restrict__ y, const float *__restrict__ dinv,
                          float *__restrict__ d, float *__restrict__ x, float c1, float c2, int dfirst, int xzero,
                          int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) mg_cheb_entry(b, y, dinv, d, x, c1, c2, dfirst, xzero, i);
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

void launch_mg_elem_coef(const DevProblem &P, double *out, hipStream_t s) {
  hipLaunchKernelGGL(k_mg_elem_coef, dim3(nbm(P.ne)), dim3(256), 0, s, P.G, (const double *)P.V,
                     (const double *)P.E_elem, P.E, P.eta, (const uint32_t *)(P.unilateral ? P.cmask : nullptr),
                     P.lam1, P.mu1, out);
}
void launch_mg_coarsen(const Grid &Gf, const double *cf, const Grid &Gc, double *cc, hipStream_t s) {
  const int64_t ne = (int64_t)Gc.nex * Gc.ney * Gc.nez;
  hipLaunchKernelGGL(k_mg_coarsen, dim3(nbm(ne)), dim3(256), 0, s, Gf, cf, Gc, cc);
}
void launch_mg_restrict(const Grid &Gf, const Grid &Gc, const BcMask &bc, const float *b, const float *y,
                        float *bcoarse, hipStream_t s, float *tmp) {
  const int64_t n3 = 3 * (int64_t)Gc.nx * Gc.ny * Gc.nz;
  // few coarse entries: latency-bound, split each over its fine planes
  if (n3 <= (1 << 16)) hipLaunchKernelGGL(k_mg_restrict4, dim3(nbm(4 * n3)), dim3(256), 0, s, Gf, Gc, bc, b, y, bcoarse);
  else if (mg_nested(Gf, Gc) && tmp) {
    float *t1 = tmp, *t2 = tmp + 3 * (int64_t)Gc.nx * Gf.ny * Gf.nz;
    hipLaunchKernelGGL(k_mg_rsx, dim3(nbm((int64_t)Gc.nx * Gf.ny), Gf.nz), dim3(256), 0, s, Gf, Gc, bc, b, y, t1);
    hipLaunchKernelGGL(k_mg_rsy, dim3(nbm((int64_t)Gc.nx * Gc.ny), Gf.nz), dim3(256), 0, s, Gf, Gc, (const float *)t1, t2);
    hipLaunchKernelGGL(k_mg_rsz, dim3(nbm((int64_t)Gc.nx * Gc.ny), Gc.nz), dim3(256), 0, s, Gf, Gc, bc, (const float *)t2,
                       bcoarse);
  } else if (mg_nested(Gf, Gc))
    hipLaunchKernelGGL(k_mg_restrict2, dim3(nbm((int64_t)Gc.nx * Gc.ny), Gc.nz), dim3(256), 0, s, Gf, Gc, bc, b, y,
                       bcoarse);
  else hipLaunchKernelGGL(k_mg_restrict, dim3(nbm(n3)), dim3(256), 0, s, Gf, Gc, bc, b, y, bcoarse);
}
void launch_mg_prolong(const Grid &Gf, const Grid &Gc, const BcMask &bc, const float *xc, float *xf,
                       hipStream_t s) {
  const int64_t nn = (int64_t)Gf.nx * Gf.ny * Gf.nz;
  if (mg_nested(Gf, Gc))
    hipLaunchKernelGGL(k_mg_prolong2, dim3(nbm((int64_t)Gf.nx * Gf.ny), Gf.nz), dim3(256), 0, s, Gf, Gc, bc, xc, xf);
  else hipLaunchKernelGGL(k_mg_prolong, dim3(nbm(nn)), dim3(256), 0, s, Gf, Gc, bc, xc, xf);
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
