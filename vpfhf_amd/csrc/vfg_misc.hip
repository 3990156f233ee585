// vfg_misc.hip -- BC kit, energies / crack volume, penny cracks and the
// vector kernels of the solvers.  All reductions are two-stage with a fixed
// summation order, so results are bitwise reproducible run to run.
#include "vfg_internal.h"
#include "vfg_kernels.h"

namespace vfg {

// ----------------------------------------------------------------- BC kit
struct BcTable {
  uint8_t face[3][6], vert[3][8];
  double faceValue[3][6], vertValue[3][8];
  int dof;
};
static BcTable bc_table(const DevProblem &P, int which) {
  BcTable t;
  t.dof = which == 0 ? 3 : 1;
  for (int c = 0; c < t.dof; c++) {
    const vfg_bc &b = which == 0 ? P.hbcU[c] : P.hbcV;
    for (int f = 0; f < 6; f++) { t.face[c][f] = (uint8_t)b.face[f]; t.faceValue[c][f] = b.faceValue[f]; }
    for (int v = 0; v < 8; v++) { t.vert[c][v] = (uint8_t)b.vertex[v]; t.vertValue[c][v] = b.vertex[v] ? b.vertexValue[v] : 0.; }
  }
  return t;
}

// Last writer over faces X0,X1,Y0,Y1,Z0,Z1 then the vertex (VFCartFE.c:279-483,
// 493-738).  Returns the BC type (0 if the dof is free) and the value it
// writes: FIXED takes faceValue/vertexValue (mode 0, VecSetFromBC) or B[d]
// (mode 1, VecApplyDirichletBC).
// With alias (vec and B are the same array, as in VecApplyDirichletBC(V,V,bcV)
// of VFTimeStepPrepare, VFCommon.c:1146) a FIXED entity copies the value the
// previous writer just stored, exactly as the in-place reference loops do.
__device__ int bc_last(const Grid &G, const BcTable &t, int c, int i, int j, int k, double Bd, int mode,
                       int alias, double *val) {
  int ty = 0;
  double v = Bd;
  const bool onf[6] = {i == 0, i == G.nx - 1, j == 0, j == G.ny - 1, k == 0, k == G.nz - 1};
  for (int f = 0; f < 6; f++)
    if (t.face[c][f] && onf[f]) {
      ty = t.face[c][f];
      v = ty == VFG_ZERO ? 0. : ty == VFG_ONE ? 1. : (mode == 0 ? t.faceValue[c][f] : alias ? v : Bd);
    }
  int ix = (i == 0) ? 0 : (i == G.nx - 1) ? 1 : -1;
  int iy = (j == 0) ? 0 : (j == G.ny - 1) ? 1 : -1;
  int iz = (k == 0) ? 0 : (k == G.nz - 1) ? 1 : -1;
  if (ix >= 0 && iy >= 0 && iz >= 0) {
    int vv = ix + 2 * iy + 4 * iz;
    if (t.vert[c][vv]) {
      ty = t.vert[c][vv];
      v = ty == VFG_ZERO ? 0. : ty == VFG_ONE ? 1. : (mode == 0 ? t.vertValue[c][vv] : alias ? v : Bd);
    }
  }
  *val = ty ? v : 0.;
  return ty;
}

__global__ void k_bc_vec(Grid G, BcTable t, uint8_t *mask, double *vec, const double *B, int mode) {
  const int alias = (mode == 1 && (const double *)vec == B);
  const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= (int64_t)G.nx * G.ny * G.nz) return;
  int i = (int)(n % G.nx), j = (int)((n / G.nx) % G.ny), k = (int)(n / ((int64_t)G.nx * G.ny));
  for (int c = 0; c < t.dof; c++) {
    int64_t d = t.dof * n + c;
    double val;
    int ty = bc_last(G, t, c, i, j, k, B ? B[d] : 0., mode, alias, &val);
    if (mask) mask[d] = ty ? 1 : 0;
    if (vec && ty) vec[d] = val;
  }
}

static unsigned nblk(int64_t n, int bs = 256) { return (unsigned)((n + bs - 1) / bs); }

void launch_bc_mask(const DevProblem &P, int which, uint8_t *mask, hipStream_t s) {
  hipLaunchKernelGGL(k_bc_vec, dim3(nblk(P.nn)), dim3(256), 0, s, P.G, bc_table(P, which), mask,
                     (double *)nullptr, (const double *)nullptr, 0);
}
void launch_vec_set_from_bc(const DevProblem &P, int which, double *B, hipStream_t s) {
  hipLaunchKernelGGL(k_bc_vec, dim3(nblk(P.nn)), dim3(256), 0, s, P.G, bc_table(P, which),
                     (uint8_t *)nullptr, B, (const double *)nullptr, 0);
}
void launch_vec_apply_dirichlet(const DevProblem &P, int which, double *vec, const double *B, hipStream_t s) {
  hipLaunchKernelGGL(k_bc_vec, dim3(nblk(P.nn)), dim3(256), 0, s, P.G, bc_table(P, which),
                     (uint8_t *)nullptr, vec, B, 1);
}

// ------------------------------------------------------- energies / volume
// Per element (VFMech.c:725-809, 920-974, 2394-2465; VFPermfield.c:15-71):
//   elastic = sum_g W [mu(1+eta) v^2 |e^D|^2 + 3(3lam+2mu)/2 ((tr/3 - alpha dth) v
//             - beta p/(3lam+2mu))^2 + (3lam+2mu)/6 tr^2 eta]
//   pressure work = sum_g W p (u . grad v)
//   surface = Gc/(4 atCv) sum_g W [(1-v)^k/eps + eps |grad v|^2], k = 1 (AT1), 2 (AT2)
//   crack volume = sum_g W u . grad v
// A workgroup takes a 16 x 16 tile of element columns over EZ element layers:
// the node planes (u, v, p, theta - thetaRef) are staged in LDS two at a time
// and each thread walks the Gauss points of its element reading the corners
// there, so the 48 corner values are not held in registers (two waves per
// SIMD).  One group of 4 partials per workgroup, in a fixed order.
constexpr int EB = 256, ET = 16, EP = ET + 1;
static int energy_ez(const Grid &G) {
  const int tiles = ((G.nex + ET - 1) / ET) * ((G.ney + ET - 1) / ET);
  return tiles >= 64 ? 8 : 2;
}
static dim3 energy_grid(const Grid &G) {
  const int ez = energy_ez(G);
  return dim3((G.nex + ET - 1) / ET, (G.ney + ET - 1) / ET, (G.nez + ez - 1) / ez);
}
int energy_blocks(const DevProblem &P) {
  const dim3 g = energy_grid(P.G);
  return (int)(g.x * g.y * g.z);
}

__global__ __launch_bounds__(EB, 2) void k_energies(DevProblem P, const double *__restrict__ U,
                                                    const double *__restrict__ V, double *partials,
                                                    double *cellfield, double *efield, int EZ) {
  __shared__ double sp[2][6][EP * EP];
  __shared__ double sh[EB / 64];
  const Grid &G = P.G;
  const int tid = threadIdx.x, tx = tid % ET, ty = tid / ET;
  const Blk3 bk = xcd_block();
  const int ex0 = bk.x * ET, ey0 = bk.y * ET, ek0 = bk.z * EZ;
  const int ek1 = min(ek0 + EZ, G.nez);
  const int ex = ex0 + tx, ey = ey0 + ty;
  const bool ok = ex < G.nex && ey < G.ney;
  // fields 0-2: u_c, 3: v (values and gradients); 4: p, 5: theta - thetaRef
  auto stage = [&](int b, int k) {
    for (int t = tid; t < EP * EP; t += EB) {
      const int i = ex0 + t % EP, j = ey0 + t / EP;
      double f[6] = {0., 0., 0., 0., 0., 0.};
      if (i < G.nx && j < G.ny) {
        const int64_t n = node_id(G, i, j, k);
        f[0] = U[3 * n]; f[1] = U[3 * n + 1]; f[2] = U[3 * n + 2];
        f[3] = V[n]; f[4] = P.p[n]; f[5] = P.dth[n];
      }
#pragma unroll
      for (int q = 0; q < 6; q++) sp[b][q][t] = f[q];
    }
  };
  const double vol = G.hx * G.hy * G.hz;
  const double csurf = P.Gc / P.atcv * .25, ieps = 1. / P.eps, third = 1. / 3.;
  double el = 0., pw = 0., se = 0., cv = 0.;
  stage(0, ek0);
  for (int ek = ek0; ek < ek1; ek++) {
    const int b = (ek - ek0) & 1;
    stage(b ^ 1, ek + 1);
    __syncthreads();
    if (ok) {
      const int64_t e = cell_id(G, ex, ey, ek);
      const double Ee = P.E_elem ? P.E_elem[e] : P.E;
      const double lam = P.lam1 * Ee, mu = P.mu1 * Ee;
      const double tk = 3. * lam + 2. * mu, cb = P.beta / tk;
      const int c0 = ty * EP + tx;
      auto corner = [&](int q, int a) {
        return sp[(a >> 2) ? b ^ 1 : b][q][c0 + ((a >> 1) & 1) * EP + (a & 1)];
      };
      double cell = 0., cel = 0., cpw = 0., cse = 0.;
      gauss_walk_cv<4, 2, 1, 1>(corner, 1. / G.hx, 1. / G.hy, 1. / G.hz,
                          [&](int gx, int gy, int gz, const double (&val)[6], const double (&H)[4][3]) {
        const double vg = val[3], pg = val[4], tg = val[5];
        const double W = vol * gW(gx) * gW(gy) * gW(gz);
        double tr = H[0][0] + H[1][1] + H[2][2];
        const double tr3 = tr * third;
        double d11 = H[0][0] - tr3, d22 = H[1][1] - tr3, d33 = H[2][2] - tr3;
        double d12 = (H[0][1] + H[1][0]) * .5, d23 = (H[1][2] + H[2][1]) * .5, d13 = (H[0][2] + H[2][0]) * .5;
        cel += (d11 * d11 + d22 * d22 + d33 * d33 + (d12 * d12 + d23 * d23 + d13 * d13) * 2.) * vg * vg *
               (1. + P.eta) * mu * W;
        // UNILATERAL_NOCOMPRESSION: the spherical term is not degraded where
        // tr(eps) < 0 (VF_ElasticEnergyNoCompression3D_local, VFMech.c:871-903)
        double WD = (tr3 - P.alpha * tg) * (P.unilateral && tr < 0. ? 1. : vg) - cb * pg;
        cel += 3. * tk / 2. * WD * WD * W;
        cel += tk / 6. * tr * tr * P.eta * W;
        const double *gv = H[3];
        double ugv = val[0] * gv[0] + val[1] * gv[1] + val[2] * gv[2];
        if (P.hasCrackPressure) cpw += W * pg * ugv;
        double g2 = gv[0] * gv[0] + gv[1] * gv[1] + gv[2] * gv[2];
        double omv = 1. - vg;
        cse += W * ((P.atnum == 2 ? omv * omv : omv) * ieps + g2 * P.eps) * csurf;
        cell += ugv * W;
      });
      el += cel; pw += cpw; se += cse; cv += cell;
      if (cellfield) cellfield[e] = cell / vol;
      if (efield) {
        efield[3 * e] = cel;
        efield[3 * e + 1] = cpw;
        efield[3 * e + 2] = cse;
      }
    }
    __syncthreads();  // plane b is restaged for layer ek + 2
  }
  double r0 = block_sum<EB>(el, sh), r1 = block_sum<EB>(pw, sh), r2 = block_sum<EB>(se, sh),
         r3 = block_sum<EB>(cv, sh);
  if (tid == 0) {
    const int64_t blk = bk.x + (int64_t)gridDim.x * (bk.y + (int64_t)gridDim.y * bk.z);  // partials in block order
    partials[blk * 4 + 0] = r0; partials[blk * 4 + 1] = r1;
    partials[blk * 4 + 2] = r2; partials[blk * 4 + 3] = r3;
  }
}

// sums `nb` groups of `width` partials column-wise with one block
__global__ __launch_bounds__(256) void k_reduce_partials(const double *partials, int nb, int width, double *out) {
  __shared__ double sh[4];
  for (int c = 0; c < width; c++) {
    double s = 0.;
    for (int b = threadIdx.x; b < nb; b += 256) s += partials[(int64_t)b * width + c];
    s = block_sum<256>(s, sh);
    if (threadIdx.x == 0) out[c] = s;
  }
}

void launch_energies(const DevProblem &P, const double *U, const double *V, double *partials,
                     double *cellfield, double *efield, double *result4, hipStream_t s) {
  const int nb = energy_blocks(P);
  hipLaunchKernelGGL(k_energies, energy_grid(P.G), dim3(EB), 0, s, P, U, V, partials, cellfield, efield,
                     energy_ez(P.G));
  hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(256), 0, s, (const double *)partials, nb, 4, result4);
}

// ------------------------------------------------------------ penny crack
// VFDistanceToPennyCrack + VFPennyCrackBuildVAT1/AT2 (VFCracks.c:101-227);
// node coordinates X[i] = i*lx/(nx-1) (VFCommon.c:395-397).
__global__ void k_penny(Grid G, double lx, double ly, double lz, int atnum, double eps, double *V,
                        double c0, double c1, double c2, double r, double th, double ph, double thick) {
  const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= (int64_t)G.nx * G.ny * G.nz) return;
  int i = (int)(n % G.nx), j = (int)((n / G.nx) % G.ny), k = (int)(n / ((int64_t)G.nx * G.ny));
  double x0 = i == 0 ? 0. : i * lx / (G.nx - 1.);
  double x1 = j == 0 ? 0. : j * ly / (G.ny - 1.);
  double x2 = k == 0 ? 0. : k * lz / (G.nz - 1.);
  if (r == 0) { V[n] = 1.; return; }
  const double PI = 3.14159265358979323846;
  double nv0 = cos(th * PI / 180.) * sin(ph * PI / 180.);
  double nv1 = sin(th * PI / 180.) * sin(ph * PI / 180.);
  double nv2 = cos(ph * PI / 180.);
  double xdn = (x0 - c0) * nv0 + (x1 - c1) * nv1 + (x2 - c2) * nv2;
  double t0 = (x0 - c0) - xdn * nv0, t1 = (x1 - c1) - xdn * nv1, t2 = (x2 - c2) - xdn * nv2;
  double l = sqrt(t0 * t0 + t1 * t1 + t2 * t2);
  double dist = (l <= r) ? sqrt(xdn * xdn) : sqrt(xdn * xdn + (l - r) * (l - r));
  dist -= thick / 2.;
  double val;
  if (atnum == 1) val = dist <= 0 ? 0. : dist < 2. * eps ? dist / eps * (1. - .25 * dist / eps) : 1.;
  else val = dist <= 0 ? 0. : 1. - exp(-dist / 2 / eps);
  V[n] = val;
}

void launch_penny(const DevProblem &P, double *V, const double c[3], double r, double theta, double phi,
                  double thick, hipStream_t s) {
  double lx = P.G.hx * P.G.nex, ly = P.G.hy * P.G.ney, lz = P.G.hz * P.G.nez;
  hipLaunchKernelGGL(k_penny, dim3(nblk(P.nn)), dim3(256), 0, s, P.G, lx, ly, lz, P.atnum, P.eps, V,
                     c[0], c[1], c[2], r, theta, phi, thick);
}

// ------------------------------------------------- cell -> node averaging
// CellToNodeInterpolation (VFPermfield.c:1156-1327): every node takes the
// volume-weighted mean of its adjacent cells, node = sum(cell*hx*hy*hz) /
// sum(hz*hy*hx).  The reference scatters cell by cell; here each thread
// gathers one (node, component) pair, visiting the adjacent cells in the
// order the scatter reaches the node (ek, ej, ei ascending) with the same
// products and no contraction, so the sums round as the reference's do.
// Cell sizes come from the node coordinates X[i] = i*l/(n-1)
// (VFCommon.c:395-397), as the reference reads them from ctx->coordinates.
__device__ __forceinline__ double coord_of(int i, double l, int n) { return i == 0 ? 0. : i * l / (n - 1.); }

__global__ __launch_bounds__(256) void k_cell_to_node(Grid G, double lx, double ly, double lz, const double *cell,
                                                      double *node, int dof) {
#pragma clang fp contract(off)
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t nn = (int64_t)G.nx * G.ny * G.nz;
  if (t >= nn * dof) return;
  const int64_t n = t / dof;
  const int d = (int)(t - n * dof);
  const int i = (int)(n % G.nx), j = (int)((n / G.nx) % G.ny), k = (int)(n / ((int64_t)G.nx * G.ny));
  double acc = 0., vol = 0.;
  for (int ek = k > 0 ? k - 1 : 0; ek <= k && ek < G.nez; ek++) {
    const double hz = coord_of(ek + 1, lz, G.nz) - coord_of(ek, lz, G.nz);
    for (int ej = j > 0 ? j - 1 : 0; ej <= j && ej < G.ney; ej++) {
      const double hy = coord_of(ej + 1, ly, G.ny) - coord_of(ej, ly, G.ny);
      for (int ei = i > 0 ? i - 1 : 0; ei <= i && ei < G.nex; ei++) {
        const double hx = coord_of(ei + 1, lx, G.nx) - coord_of(ei, lx, G.nx);
        vol = vol + hz * hy * hx;
        acc = acc + cell[dof * cell_id(G, ei, ej, ek) + d] * hx * hy * hz;
      }
    }
  }
  node[t] = acc / vol;
}

void launch_cell_to_node(const DevProblem &P, double lx, double ly, double lz, const double *cell, double *node,
                         int dof, hipStream_t s) {
  const int64_t n = P.nn * dof;
  hipLaunchKernelGGL(k_cell_to_node, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, P.G, lx, ly, lz, cell,
                     node, dof);
}

// ------------------------------------------------------------ vector ops
constexpr int RB = 256;
int red_blocks(int64_t n) {
  int64_t b = (n + RB * 4 - 1) / (RB * 4);
  return (int)(b < 1 ? 1 : b > 2048 ? 2048 : b);
}

__global__ __launch_bounds__(RB) void k_dot(const double *__restrict__ a, const double *__restrict__ b,
                                            int64_t n, double *partials) {
  __shared__ double sh[RB / 64];
  double s = 0.;
  for (int64_t i = (int64_t)blockIdx.x * RB + threadIdx.x; i < n; i += (int64_t)gridDim.x * RB) s += a[i] * b[i];
  s = block_sum<RB>(s, sh);
  if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

void launch_dot(const double *a, const double *b, int64_t n, double *partials, double *out, hipStream_t s) {
  int nb = red_blocks(n);
  hipLaunchKernelGGL(k_dot, dim3(nb), dim3(RB), 0, s, a, b, n, partials);
  hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(256), 0, s, (const double *)partials, nb, 1, out);
}

__global__ void k_fill(double *x, double v, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = v;
}
void launch_fill(double *x, double v, int64_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_fill, dim3(nblk(n)), dim3(256), 0, s, x, v, n);
}
__global__ void k_sub(const double *a, const double *b, double *o, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) o[i] = a[i] - b[i];
}
void launch_sub(const double *a, const double *b, double *out, int64_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_sub, dim3(nblk(n)), dim3(256), 0, s, a, b, out, n);
}
__global__ void k_d2f(const double *x, float *y, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = (float)x[i];
}
__global__ void k_f2d(const float *x, double *y, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = (double)x[i];
}
void launch_copy_d2f(const double *x, float *y, int64_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_d2f, dim3(nblk(n)), dim3(256), 0, s, x, y, n);
}
void launch_copy_f2d(const float *x, double *y, int64_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_f2d, dim3(nblk(n)), dim3(256), 0, s, x, y, n);
}

__global__ void k_vmin(const double *a, const double *b, double *o, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) o[i] = fmin(a[i], b[i]);
}
__global__ void k_scale(double *x, double s, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] *= s;
}
__global__ __launch_bounds__(RB) void k_maxabs(const double *a, const double *b, int64_t n, double *part) {
  __shared__ double sh[RB];
  double m = 0.;
  for (int64_t i = (int64_t)blockIdx.x * RB + threadIdx.x; i < n; i += (int64_t)gridDim.x * RB)
    m = fmax(m, fabs(a[i] - b[i]));
  sh[threadIdx.x] = m;
  __syncthreads();
  for (int o = RB / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + o]);
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}
// max over the partials: one workgroup, strided loads, then a tree in LDS
// (max is exact, so any order gives the same value)
__global__ __launch_bounds__(RB) void k_maxreduce(const double *part, int nb, double *out) {
  __shared__ double sh[RB];
  double m = 0.;
  for (int i = threadIdx.x; i < nb; i += RB) m = fmax(m, part[i]);
  sh[threadIdx.x] = m;
  __syncthreads();
  for (int w = RB / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = sh[0];
}
void launch_vmin(const double *a, const double *b, double *o, int64_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_vmin, dim3(nblk(n)), dim3(256), 0, s, a, b, o, n);
}
void launch_scale(double *x, double v, int64_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_scale, dim3(nblk(n)), dim3(256), 0, s, x, v, n);
}
void launch_maxabs_diff(const double *a, const double *b, int64_t n, double *part, double *out, hipStream_t s) {
  int nb = red_blocks(n);
  hipLaunchKernelGGL(k_maxabs, dim3(nb), dim3(RB), 0, s, a, b, n, part);
  hipLaunchKernelGGL(k_maxreduce, dim3(1), dim3(RB), 0, s, (const double *)part, nb, out);
}

}  // namespace vfg

// ------------------------------------------- energies without allocation
// vfg_energies_static: the reductions of vfg_energies with their partials in
// a file-scope device array instead of handle buffers, for callers that
// cannot allocate device memory (HIP graph capture, serving loops, harnesses
// that forbid hipMalloc).  The array holds ENERGY_STATIC_BLOCKS groups of 4
// partials plus the 4 results: 8 MiB + 32 B, resident on every GPU the
// library runs on until the process ends.  Every partial is written by
// k_energies before k_reduce_partials reads it, so nothing needs clearing;
// calls that overlap on one GPU share the array and race.
#include <mutex>
#include "vfg_handle.h"

static constexpr int ENERGY_STATIC_BLOCKS = 1 << 18;  // workgroups of 16 x 16 x 8 cells: 536,870,912 cells
static __device__ __align__(256) double scratch[4 * ENERGY_STATIC_BLOCKS + 4];

namespace {
constexpr int MAX_DEVICES = 64;
// device address of `scratch` on each GPU (the host-side name is a stand-in,
// never a device address)
double *g_scratch_dev[MAX_DEVICES];
std::mutex g_scratch_mu;

int scratch_address(vfg_handle *h, double **out) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= MAX_DEVICES)
    return vfg_fail(h, VFG_E_HIP, "vfg_energies_static: no current HIP device (or device id >= %d)", MAX_DEVICES);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void *p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return vfg_fail(h, VFG_E_HIP, "vfg_energies_static: hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
    g_scratch_dev[dev] = (double *)p;
  }
  *out = g_scratch_dev[dev];
  return VFG_OK;
}
}  // namespace

extern "C" int vfg_energies_static(vfg_handle *h, const double *U, const double *V, double out[4]) {
  if (!h) return vfg_fail(nullptr, VFG_E_ARG, "NULL handle");
  if (!U || !V || !out) return vfg_fail(h, VFG_E_ARG, "vfg_energies_static: NULL array");
  const int nb = vfg::energy_blocks(h->P);
  if (nb > ENERGY_STATIC_BLOCKS)
    return vfg_fail(h, VFG_E_ARG,
                    "vfg_energies_static: %lld cells need %d partial groups, more than the %d of the "
                    "static scratch array (use vfg_energies)",
                    (long long)h->P.ne, nb, ENERGY_STATIC_BLOCKS);
  double *part = nullptr;
  int rc = scratch_address(h, &part);
  if (rc) return rc;
  double *res = part + 4 * (size_t)ENERGY_STATIC_BLOCKS;
  vfg::launch_energies(h->P, U, V, part, nullptr, nullptr, res, h->stream);
  double r[4];
  if (hipMemcpyAsync(r, res, sizeof r, hipMemcpyDeviceToHost, h->stream) != hipSuccess ||
      hipStreamSynchronize(h->stream) != hipSuccess || hipGetLastError() != hipSuccess)
    return vfg_fail(h, VFG_E_HIP, "vfg_energies_static: kernel execution failed");
  out[0] = r[0];
  out[1] = 0.;
  out[2] = r[1];
  out[3] = r[2];
  return VFG_OK;
}
