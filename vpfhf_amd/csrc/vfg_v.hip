// vfg_v.hip -- V-subproblem (AT1/AT2 phase field) operators on gfx950.
//
// Element matrix (VFMech.c:2040-2115):
//   K[a][b] = c sum_g W_g (phi_a phi_b / eps [AT2 only] + eps grad phi_a . grad phi_b)
//           + sum_g W_g phi_a phi_b 2 W^el_g,            c = Gc / atCv / 2,
//   W^el_g = 1/2 sigma(e~):e~,  e~ = e(u) - alpha dtheta I  (VFMech.c:32-106).
// Split as K = ks * L + M_q with the constant Q1 stiffness L (8x8, exact for
// the 27-point rule) and the point-weighted mass M_q, q_g = W_g (c/eps [AT2] +
// 2 W^el_g) precomputed once per U (vfg_v_set_u) and stored [27][N_e].
#include "vfg_internal.h"
#include "vfg_kernels.h"

namespace vfg {

struct VConst {
  double L[64];     // sum_g W_g grad phi_a . grad phi_b (element of size h)
  double ks;        // c * eps
  double crhs;      // Gc/atCv/eps * (.5 AT2 | .25 AT1), VFMech.c:2350-2384
};

static VConst make_vconst(const DevProblem &P) {
  VConst C;
  const double h[3] = {P.G.hx, P.G.hy, P.G.hz};
  for (int a = 0; a < 8; a++)
    for (int b = 0; b < 8; b++) {
      double s = 0.;
      for (int g = 0; g < 27; g++) {
        int gx = g % 3, gy = (g / 3) % 3, gz = g / 9;
        const int ga[3] = {gx, gy, gz};
        double W = h[0] * h[1] * h[2] * gW(gx) * gW(gy) * gW(gz);
        double da[3], db[3];
        for (int d = 0; d < 3; d++) {
          double pa = 1., pb = 1.;
          for (int e = 0; e < 3; e++) {
            int ia = (a >> e) & 1, ib = (b >> e) & 1;
            if (e == d) {
              pa *= (ia ? 1. : -1.) / h[e];
              pb *= (ib ? 1. : -1.) / h[e];
            } else {
              pa *= ia ? gP1(ga[e]) : gP0(ga[e]);
              pb *= ib ? gP1(ga[e]) : gP0(ga[e]);
            }
          }
          da[d] = pa; db[d] = pb;
        }
        s += W * (da[0] * db[0] + da[1] * db[1] + da[2] * db[2]);
      }
      C.L[a * 8 + b] = s;
    }
  double c = P.Gc / P.atcv * .5;
  C.ks = c * P.eps;
  C.crhs = P.atnum == 2 ? P.Gc / P.atcv / P.eps * .5 : P.Gc / P.atcv / P.eps * .25;
  return C;
}

// -------------------------------------------------------------- setup: Vq
// One block per (tile, element layer), one thread per tile slot: element
// (15 bx - 1 + tx, 15 by - 1 + ty) of the V operator's tile-blocked layout
// (halo elements are computed by every tile that holds them), so each
// coefficient is written as one aligned, contiguous 512-byte segment per
// wavefront.  Gauss-point values by sum factorisation (gauss_walk).
__global__ __launch_bounds__(256) void k_v_setup(DevProblem P, const double *__restrict__ U) {
  const Grid &G = P.G;
  const int tid = threadIdx.x;
  const Blk3 b = xcd_block();
  const int ei = b.x * 15 - 1 + tid % 16, ej = b.y * 15 - 1 + tid / 16, ek = b.z;
  const int tile = b.x + gridDim.x * b.y;
  double *out = P.Vq + tiled_index(tile, ek, G.nez, 27, tid, 0, 256);
  if (ei < 0 || ei >= G.nex || ej < 0 || ej >= G.ney) {
    for (int g = 0; g < 27; g++) out[g * 256] = 0.;
    return;
  }
  const int64_t e = cell_id(G, ei, ej, ek);
  double fld[4][8];  // u_c (values and gradients), theta - thetaRef (values)
  for (int a = 0; a < 8; a++) {
    int64_t n = node_id(G, ei + (a & 1), ej + ((a >> 1) & 1), ek + (a >> 2));
    for (int c = 0; c < 3; c++) fld[c][a] = U[3 * n + c];
    fld[3][a] = P.dth[n];
  }
  const double Ee = P.E_elem ? P.E_elem[e] : P.E;
  const double lam = P.lam1 * Ee, mu = P.mu1 * Ee;
  const double vol = G.hx * G.hy * G.hz;
  const double cm = P.atnum == 2 ? P.Gc / P.atcv * .5 / P.eps : 0.;
  gauss_walk<3, 1>(fld, 1. / G.hx, 1. / G.hy, 1. / G.hz,
                   [&](int gx, int gy, int gz, const double (&val)[4], const double (&H)[3][3]) {
    const double t = val[3];
    double e11 = H[0][0] - P.alpha * t, e22 = H[1][1] - P.alpha * t, e33 = H[2][2] - P.alpha * t;
    double e12 = (H[1][0] + H[0][1]) * .5, e23 = (H[2][1] + H[1][2]) * .5, e13 = (H[2][0] + H[0][2]) * .5;
    double s11 = (lam + 2. * mu) * e11 + lam * e22 + lam * e33;
    double s22 = lam * e11 + (lam + 2. * mu) * e22 + lam * e33;
    double s33 = lam * e11 + lam * e22 + (lam + 2. * mu) * e33;
    double W;
    if (P.unilateral) {
      // ElasticEnergyDensitySphericalDeviatoricNoCompression3D_local
      // (VFMech.c:183-208): 1/2 sigma(eps - e0) : eps, less kappa tr^2 / 2
      // where tr(eps) < 0 (S + D of VFMech.c:2125-2153)
      const double tr = (H[0][0] + H[1][1]) + H[2][2];
      W = (s11 * H[0][0] + s22 * H[1][1] + s33 * H[2][2]) * .5 + 2. * mu * (e12 * e12 + e23 * e23 + e13 * e13);
      if (tr < 0.) W -= (lam + 2. * mu / 3.) * tr * tr * .5;
    } else {
      W = (s11 * e11 + s22 * e22 + s33 * e33) * .5 + 2. * mu * (e12 * e12 + e23 * e23 + e13 * e13);
    }
    const double Wg = vol * gW(gx) * gW(gy) * gW(gz);
    out[((gz * 3 + gy) * 3 + gx) * 256] = Wg * (cm + 2. * W);
  });
}

int64_t vq_doubles(const Grid &G) {
  return (int64_t)((G.nx + 14) / 15) * ((G.ny + 14) / 15) * G.nez * 27 * 256;
}

__global__ void k_v_fixmask(DevProblem P) {
  const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= P.nn) return;
  const Grid &G = P.G;
  int i = (int)(n % G.nx), j = (int)((n / G.nx) % G.ny), k = (int)(n / ((int64_t)G.nx * G.ny));
  bool f = on_dirichlet(G, P.bcV.face[0], P.bcV.vert[0], i, j, k);
  if (P.atnum == 2 && P.VIrrev[n] <= P.irrevtol) f = true;  // VF_IrrevApplyEQMat
  P.vfix[n] = f ? 1 : 0;
}

// ------------------------------------------------------------------- apply
// the 27 coupling weights of one element, read one layer ahead by the march
struct VQ {
  double q[27];
};
__device__ __forceinline__ VQ load_vq(const double *__restrict__ Vq, int tile, int slot, int ek, int nez) {
  VQ a;
  const double *b = Vq + tiled_index(tile, ek, nez, 27, slot, 0, 256);
#pragma unroll
  for (int g = 0; g < 27; g++) a.q[g] = b[g * 256];
  return a;
}

// y_a = ks sum_b L[a][b] x_b + sum_g q_g phi_a(g) x(g); the weights are read
// one Gauss plane (9 values) at a time, the next plane's loads in flight while
// the current one is used (keeps the register count at four waves per SIMD)
__device__ __forceinline__ void v_elem_apply_qp(const double (&x)[8], const double *__restrict__ qb,
                                                const VConst &C, double (&r)[8]) {
  double qa[9], qn[9];
#pragma unroll
  for (int g = 0; g < 9; g++) qa[g] = qb[g * 256];
#pragma unroll
  for (int a = 0; a < 8; a++) {
    double s = 0.;
#pragma unroll
    for (int b = 0; b < 8; b++) s += C.L[a * 8 + b] * x[b];
    r[a] = C.ks * s;
  }
  double ax[2][2][3];  // [k][j][gx]
#pragma unroll
  for (int k = 0; k < 2; k++)
#pragma unroll
    for (int j = 0; j < 2; j++)
#pragma unroll
      for (int gx = 0; gx < 3; gx++) ax[k][j][gx] = interp<double>(gx, x[(k * 2 + j) * 2], x[(k * 2 + j) * 2 + 1]);
  double bz[2][2][3];
#pragma unroll
  for (int k = 0; k < 2; k++)
#pragma unroll
    for (int j = 0; j < 2; j++)
#pragma unroll
      for (int gx = 0; gx < 3; gx++) bz[k][j][gx] = 0.;
#pragma unroll
  for (int gz = 0; gz < 3; gz++) {
    if (gz < 2)
#pragma unroll
      for (int g = 0; g < 9; g++) qn[g] = qb[((gz + 1) * 9 + g) * 256];
#pragma unroll
    for (int gy = 0; gy < 3; gy++)
#pragma unroll
      for (int gx = 0; gx < 3; gx++) {
        double a0 = interp<double>(gy, ax[0][0][gx], ax[0][1][gx]);
        double a1 = interp<double>(gy, ax[1][0][gx], ax[1][1][gx]);
        double xv = interp<double>(gz, a0, a1);
        double f = qa[gy * 3 + gx] * xv;
        double f0 = gP0(gz) * f, f1 = gP1(gz) * f;
        bz[0][0][gx] += gP0(gy) * f0; bz[0][1][gx] += gP1(gy) * f0;
        bz[1][0][gx] += gP0(gy) * f1; bz[1][1][gx] += gP1(gy) * f1;
      }
#pragma unroll
    for (int g = 0; g < 9; g++) qa[g] = qn[g];
  }
#pragma unroll
  for (int k = 0; k < 2; k++)
#pragma unroll
    for (int j = 0; j < 2; j++) {
      double s0 = 0., s1 = 0.;
#pragma unroll
      for (int gx = 0; gx < 3; gx++) { s0 += gP0(gx) * bz[k][j][gx]; s1 += gP1(gx) * bz[k][j][gx]; }
      r[(k * 2 + j) * 2] += s0;
      r[(k * 2 + j) * 2 + 1] += s1;
    }
}

struct VApplyOp {
  static constexpr int NF = 1, NC = 1;
  static constexpr bool TILED = true;  // element(tile, slot, ek, ...): q is tile-blocked
  const double *x;
  double *y;
  const double *q, *bV, *VIrrev;
  const uint8_t *vfix;
  Grid G;
  int64_t ne;
  VConst C;
  BcMask bc;
  uint8_t ftype[6], vtype[8];
  int mode, atnum;
  double irrevtol;
  __device__ double load(int, int64_t n) const { return (mode == 0 && vfix[n]) ? 0. : x[n]; }
  // mode 0: masked Jacobian; 1: VF_VResidual; 2: plain K x
  // the coupling weights are read inside the layer: prefetching them a layer
  // ahead costs a wave of occupancy and is slower
  __device__ void element(int tile, int slot, int ek, const double (&xin)[2][4][1], double (&r)[8][1]) const {
    double xe[8], re[8];
#pragma unroll
    for (int a = 0; a < 8; a++) xe[a] = xin[a >> 2][a & 3][0];
    v_elem_apply_qp(xe, q + tiled_index(tile, ek, G.nez, 27, slot, 0, 256), C, re);
#pragma unroll
    for (int a = 0; a < 8; a++) r[a][0] = re[a];
  }
  __device__ void store(int i, int j, int k, const double (&acc)[1], const double (&)[1]) const {
    int64_t n = node_id(G, i, j, k);
    if (mode == 0) {
      y[n] = vfix[n] ? x[n] : acc[0];
      return;
    }
    if (mode == 2) {
      y[n] = acc[0];
      return;
    }
    // VF_VResidual: F = K V - b; AT2 irreversible rows -> 0 (VF_IrrevApplyEQVec);
    // then ResidualApplyDirichletBC with V as its own BC vector.
    double F = acc[0] - bV[n];
    if (atnum == 2 && VIrrev[n] <= irrevtol) F = 0.;
    if (on_dirichlet(G, bc.face[0], bc.vert[0], i, j, k)) {
      int t = 0;
      const bool onf[6] = {i == 0, i == G.nx - 1, j == 0, j == G.ny - 1, k == 0, k == G.nz - 1};
      for (int f = 0; f < 6; f++)
        if (ftype[f] && onf[f]) t = ftype[f];
      int ix = (i == 0) ? 0 : (i == G.nx - 1) ? 1 : -1;
      int iy = (j == 0) ? 0 : (j == G.ny - 1) ? 1 : -1;
      int iz = (k == 0) ? 0 : (k == G.nz - 1) ? 1 : -1;
      if (ix >= 0 && iy >= 0 && iz >= 0 && vtype[ix + 2 * iy + 4 * iz]) t = vtype[ix + 2 * iy + 4 * iz];
      double v = x[n];
      F = t == 2 ? v - 1. : t == 3 ? 0. : v;
    }
    y[n] = F;
  }
};

// CG-fused masked apply (mode 0): stages p = z + beta p_old off the mask, writes
// p and q = K p, one partial of p.q per workgroup; p_old and p are different
// arrays (neighbouring workgroups still read p_old at their halo nodes)
struct VApplyCgOp : VApplyOp {
  const double *z;
  const double *beta;
  double *pw;
  double *part;
  __device__ double load(int, int64_t n) const { return vfix[n] ? 0. : z[n] + *beta * x[n]; }
  __device__ double store(int i, int j, int k, const double (&acc)[1], const double (&own)[1]) const {
    const int64_t n = node_id(G, i, j, k);
    const double yv = vfix[n] ? own[0] : acc[0];
    y[n] = yv;
    pw[n] = own[0];
    return own[0] * yv;
  }
  __device__ void partial(int slot, double v) const { part[slot] = v; }
};

__global__ __launch_bounds__(256, 3) void k_v_apply(VApplyOp op, int OZ) { march<double, 16, 16>(op.G, op, OZ, xcd_block()); }

// The same over a list of workgroups.  The reduced Newton systems of the
// V-solve live on the free nodes, a thin band around the cracks: a workgroup
// owning no free node has p = q = 0 throughout and is left out.  The list
// (k_v_list_chunks) holds chunk ids tile + txy * zc, a chunk being the node
// planes [zc L, zc L + L) of a 15 x 15 node tile; L sits in list[cap + 1].
// Partials are written compactly, one per list entry.
struct VChunk {
  int bx, by, bz, L;
};
__device__ __forceinline__ VChunk chunk_at(const Grid &G, const int *__restrict__ list, int cap, int w) {
  const int gx = (G.nx + 14) / 15, txy = gx * ((G.ny + 14) / 15);
  const int c = list[w], t = c % txy;
  return VChunk{t % gx, t / gx, c / txy, list[cap + 1]};
}
__global__ __launch_bounds__(256, 3) void k_v_apply_cg_list(VApplyCgOp op, const int *__restrict__ list, int cap,
                                                           const double *__restrict__ done) {
  if (done && *done != 0.) return;  // the CG has converged
  const VChunk q = chunk_at(op.G, list, cap, blockIdx.x);
  march<double, 16, 16>(op.G, op, q.L, q.bx, q.by, q.bz, blockIdx.x);
}
// the plain apply (any mode) over the listed chunks
__global__ __launch_bounds__(256, 3) void k_v_apply_list(VApplyOp op, const int *__restrict__ list, int cap) {
  const VChunk q = chunk_at(op.G, list, cap, blockIdx.x);
  march<double, 16, 16>(op.G, op, q.L, q.bx, q.by, q.bz);
}

__global__ void k_fill_u32(uint32_t *a, uint32_t v, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = v;
}

// flag[tile + txy k] = 1 where node plane k of a 15 x 15 node tile holds a
// node with mask == 0 (flag cleared beforehand)
__global__ void k_v_mark_planes(Grid G, const uint8_t *__restrict__ mask, uint32_t *__restrict__ flag) {
  const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= (int64_t)G.nx * G.ny * G.nz || mask[n]) return;
  const int i = (int)(n % G.nx), j = (int)((n / G.nx) % G.ny), k = (int)(n / ((int64_t)G.nx * G.ny));
  const int gx = (G.nx + 14) / 15, txy = gx * ((G.ny + 14) / 15);
  flag[i / 15 + gx * (j / 15) + (int64_t)txy * k] = 1u;  // same value from every writer
}

// out[tile + txy k] = 1 where a flag is set on one of the 3 x 3 tiles around
// the tile, on planes k - 1 .. k + 1: the (tile, plane) pairs holding a node
// that shares an element with a node of a flagged pair (a node at a tile edge
// meets the next tile's nodes, a node plane the planes next to it)
__global__ void k_v_dilate_planes(const uint32_t *__restrict__ in, uint32_t *__restrict__ out, int gx, int gy,
                                  int nz) {
  const int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t txy = (int64_t)gx * gy;
  if (f >= txy * nz) return;
  const int t = (int)(f % txy), k = (int)(f / txy), tx = t % gx, ty = t / gx;
  uint32_t any = 0u;
  for (int dk = -1; dk <= 1; dk++) {
    const int kk = k + dk;
    if (kk < 0 || kk >= nz) continue;
    for (int dy = -1; dy <= 1; dy++) {
      const int yy = ty + dy;
      if (yy < 0 || yy >= gy) continue;
      for (int dx = -1; dx <= 1; dx++) {
        const int xx = tx + dx;
        if (xx >= 0 && xx < gx) any |= in[xx + (int64_t)gx * yy + txy * kk];
      }
    }
  }
  out[f] = any ? 1u : 0u;
}

constexpr int VLT = 1024;  // threads of the chunk-list kernel

__device__ __forceinline__ int block_isum(int v, int *sh) {
  __syncthreads();
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int o = VLT / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
    __syncthreads();
  }
  return sh[0];
}

// One workgroup of VLT threads.  Chunk length L = ceil(flagged planes /
// target), clamped to [lmin, lmax]: a thin band gives many one-plane chunks
// (latency, not work, bounds the march of a few chunks), a wide one long
// chunks (each chunk re-reads one halo plane).  The flagged chunks are listed
// in increasing id; list[cap] = their count, list[cap + 1] = L.
__global__ __launch_bounds__(VLT) void k_v_list_chunks(const uint32_t *__restrict__ flag, int txy, int nz, int lmin,
                                                       int lmax, int target, int *__restrict__ list, int cap) {
  __shared__ int sh[VLT];
  const int tid = threadIdx.x;
  const int64_t nf = (int64_t)txy * nz;
  int f = 0;
  for (int64_t i = tid; i < nf; i += VLT) f += flag[i] != 0u;
  const int F = block_isum(f, sh);
  int L = (F + target - 1) / target;
  L = L < lmin ? lmin : L > lmax ? lmax : L;
  if (L < 1) L = 1;
  const int nzc = (nz + L - 1) / L;
  const int64_t nch = (int64_t)txy * nzc, per = (nch + VLT - 1) / VLT;
  const int64_t c0 = tid * per < nch ? tid * per : nch, c1 = c0 + per < nch ? c0 + per : nch;
  auto hit = [&](int64_t c) {
    const int t = (int)(c % txy), k0 = (int)(c / txy) * L, k1 = k0 + L < nz ? k0 + L : nz;
    for (int k = k0; k < k1; k++)
      if (flag[t + (int64_t)txy * k]) return true;
    return false;
  };
  int cnt = 0;
  for (int64_t c = c0; c < c1; c++) cnt += hit(c);
  // exclusive scan of the per-thread counts
  __syncthreads();
  sh[tid] = cnt;
  __syncthreads();
  for (int o = 1; o < VLT; o <<= 1) {
    const int v = tid >= o ? sh[tid - o] : 0;
    __syncthreads();
    sh[tid] += v;
    __syncthreads();
  }
  int off = sh[tid] - cnt;
  for (int64_t c = c0; c < c1; c++)
    if (hit(c)) list[off++] = (int)c;
  if (tid == VLT - 1) {
    list[cap] = sh[VLT - 1];
    list[cap + 1] = L;
  }
}

// x += alpha p, r -= alpha q, z = dinv r and the partials of r.z, r.r over
// the owned nodes of the listed chunks (p = q = r = 0 elsewhere)
__global__ __launch_bounds__(256) void k_v_xr_list(Grid G, const int *__restrict__ list, int cap,
                                                   double *__restrict__ x, double *__restrict__ r,
                                                   double *__restrict__ z, const double *__restrict__ p,
                                                   const double *__restrict__ q, const double *__restrict__ dinv,
                                                   const double *__restrict__ alpha, double *__restrict__ part,
                                                   const double *__restrict__ done) {
  if (done && *done != 0.) return;  // the CG has converged
  __shared__ double sh[4];
  const int gx = (G.nx + 14) / 15, txy = gx * ((G.ny + 14) / 15);
  const int c = list[blockIdx.x], L = list[cap + 1], t = c % txy;
  const int tx = threadIdx.x % 16, ty = threadIdx.x / 16;
  const int i = (t % gx) * 15 + tx, j = (t / gx) * 15 + ty, k0 = (c / txy) * L;
  const int k1 = k0 + L < G.nz ? k0 + L : G.nz;
  const double al = *alpha;
  double rz = 0., rr = 0.;
  if (tx < 15 && ty < 15 && i < G.nx && j < G.ny)
    for (int k = k0; k < k1; k++) {
      const int64_t n = node_id(G, i, j, k);
      x[n] += al * p[n];
      const double ri = r[n] - al * q[n];
      r[n] = ri;
      const double zi = dinv[n] * ri;
      z[n] = zi;
      rz += ri * zi;
      rr += ri * ri;
    }
  rz = block_sum<256>(rz, sh);
  rr = block_sum<256>(rr, sh);
  if (threadIdx.x == 0) { part[2 * blockIdx.x] = rz; part[2 * blockIdx.x + 1] = rr; }
}

// ---------------------------------------------------------------- diagonal
struct VDiagOp {
  static constexpr int NF = 1, NC = 1;
  static constexpr bool TILED = true;
  double *d;
  const double *q;
  const uint8_t *vfix;
  Grid G;
  int64_t ne;
  VConst C;
  __device__ double load(int, int64_t) const { return 0.; }
  __device__ VQ aux(int tile, int slot, int ek) const { return load_vq(q, tile, slot, ek, G.nez); }
  // d_a = ks L[a][a] + sum_g q_g phi_a(g)^2, phi_a^2 = P_i(gx)^2 P_j(gy)^2 P_k(gz)^2:
  // contracted direction by direction (x, then y, then z)
  __device__ void element(int, int, int, const double (&)[2][4][1], const VQ &Q, double (&r)[8][1]) const {
    double bx[3][3][2];  // [gz][gy][i]
#pragma unroll
    for (int gz = 0; gz < 3; gz++)
#pragma unroll
      for (int gy = 0; gy < 3; gy++) {
        double s0 = 0., s1 = 0.;
#pragma unroll
        for (int gx = 0; gx < 3; gx++) {
          const double qv = Q.q[(gz * 3 + gy) * 3 + gx];
          s0 += gP0(gx) * gP0(gx) * qv;
          s1 += gP1(gx) * gP1(gx) * qv;
        }
        bx[gz][gy][0] = s0;
        bx[gz][gy][1] = s1;
      }
    double by[3][2][2];  // [gz][j][i]
#pragma unroll
    for (int gz = 0; gz < 3; gz++)
#pragma unroll
      for (int i = 0; i < 2; i++) {
        double s0 = 0., s1 = 0.;
#pragma unroll
        for (int gy = 0; gy < 3; gy++) {
          s0 += gP0(gy) * gP0(gy) * bx[gz][gy][i];
          s1 += gP1(gy) * gP1(gy) * bx[gz][gy][i];
        }
        by[gz][0][i] = s0;
        by[gz][1][i] = s1;
      }
#pragma unroll
    for (int a = 0; a < 8; a++) {
      const int i = a & 1, j = (a >> 1) & 1, k = a >> 2;
      double s = 0.;
#pragma unroll
      for (int gz = 0; gz < 3; gz++) s += (k ? gP1(gz) * gP1(gz) : gP0(gz) * gP0(gz)) * by[gz][j][i];
      r[a][0] = C.ks * C.L[a * 9] + s;
    }
  }
  __device__ void store(int i, int j, int k, const double (&acc)[1], const double (&)[1]) const {
    int64_t n = node_id(G, i, j, k);
    d[n] = vfix[n] ? 1. : acc[0];
  }
};

__global__ __launch_bounds__(256) void k_v_diag(VDiagOp op, int OZ) { march<double, 16, 16>(op.G, op, OZ, xcd_block()); }
__global__ __launch_bounds__(256) void k_v_diag_list(VDiagOp op, const int *__restrict__ list, int cap) {
  const VChunk q = chunk_at(op.G, list, cap, blockIdx.x);
  march<double, 16, 16>(op.G, op, q.L, q.bx, q.by, q.bz);
}

// ------------------------------------------------------------ load vector
// b_a = crhs sum_g W phi_a + [pressurize] sum_g W p (u . grad phi_a)
struct VRhsOp {
  static constexpr int NF = 4, NC = 1;
  const double *U, *p;
  double *b;
  Grid G;
  double crhs;
  int crack;
  __device__ double load(int f, int64_t n) const { return f < 3 ? U[3 * n + f] : p[n]; }
  // AT part: c_rhs sum_g W_g phi_a(g) = c_rhs vol / 8 for every corner (the
  // 3-point rule integrates phi_a exactly).  Crack-pressure part: the flux
  // F(g) = W_g p_g u_g projected on grad phi_a, sum-factorised: the x-flux
  // summed over gx per (gy, gz), the y-flux over gy per (gx, gz), the z-flux
  // over gz per (gx, gy), each then projected on the two other axes.
  __device__ void element(int, int, int, const double (&X)[2][4][4], double (&r)[8][1]) const {
    const double vol = G.hx * G.hy * G.hz;
    const double at = crhs * vol * .125;
    if (!crack) {
#pragma unroll
      for (int a = 0; a < 8; a++) r[a][0] += at;
      return;
    }
    double fld[4][8];
#pragma unroll
    for (int a = 0; a < 8; a++)
#pragma unroll
      for (int f = 0; f < 4; f++) fld[f][a] = X[a >> 2][a & 3][f];
    double Rx[2][2] = {{0., 0.}, {0., 0.}}, Ry[2][2] = {{0., 0.}, {0., 0.}}, Sz[3][3], Sy[3], Sx = 0.;
#pragma unroll
    for (int q = 0; q < 3; q++) {
      Sy[q] = 0.;
#pragma unroll
      for (int w = 0; w < 3; w++) Sz[q][w] = 0.;
    }
    gauss_walk<0, 4>(fld, 0., 0., 0., [&](int gx, int gy, int gz, const double (&val)[4], const double (&)[1][3]) {
      const double wp = vol * gW(gx) * gW(gy) * gW(gz) * val[3];
      Sx += wp * val[0];
      Sy[gx] += wp * val[1];
      Sz[gy][gx] += wp * val[2];
      if (gx == 2) {  // end of a row: the x-flux of (gy, gz) onto phi_j(y) phi_k(z)
#pragma unroll
        for (int k = 0; k < 2; k++)
#pragma unroll
          for (int j = 0; j < 2; j++) Rx[k][j] += (k ? gP1(gz) : gP0(gz)) * (j ? gP1(gy) : gP0(gy)) * Sx;
        Sx = 0.;
      }
      if (gx == 2 && gy == 2) {  // end of a layer: the y-flux of (gx, gz) onto phi_i(x) phi_k(z)
#pragma unroll
        for (int q = 0; q < 3; q++) {
#pragma unroll
          for (int k = 0; k < 2; k++)
#pragma unroll
            for (int i = 0; i < 2; i++) Ry[k][i] += (k ? gP1(gz) : gP0(gz)) * (i ? gP1(q) : gP0(q)) * Sy[q];
          Sy[q] = 0.;
        }
      }
    });
    double Rz[2][2];  // [j][i]
#pragma unroll
    for (int j = 0; j < 2; j++)
#pragma unroll
      for (int i = 0; i < 2; i++) {
        double acc = 0.;
#pragma unroll
        for (int gy = 0; gy < 3; gy++)
#pragma unroll
          for (int gx = 0; gx < 3; gx++) acc += (j ? gP1(gy) : gP0(gy)) * (i ? gP1(gx) : gP0(gx)) * Sz[gy][gx];
        Rz[j][i] = acc;
      }
    const double ihx = 1. / G.hx, ihy = 1. / G.hy, ihz = 1. / G.hz;
#pragma unroll
    for (int a = 0; a < 8; a++) {
      const int i = a & 1, j = (a >> 1) & 1, k = a >> 2;
      const double vx = (i ? ihx : -ihx) * Rx[k][j], vy = (j ? ihy : -ihy) * Ry[k][i],
                   vz = (k ? ihz : -ihz) * Rz[j][i];
      r[a][0] += at + (vx + vy + vz);
    }
  }
  __device__ void store(int i, int j, int k, const double (&acc)[1], const double (&)[4]) const {
    b[node_id(G, i, j, k)] = acc[0];
  }
};

__global__ __launch_bounds__(256) void k_v_rhs(VRhsOp op, int OZ) { march<double, 16, 16>(op.G, op, OZ, xcd_block()); }

// ------------------------------------------------------------ launchers
static dim3 mgrid(const Grid &G, int OZ) {
  return dim3((G.nx + 14) / 15, (G.ny + 14) / 15, (G.nz + OZ - 1) / OZ);
}
int choose_oz(const Grid &G, int tiles_xy, int slots);
// z-chunk length of the V kernels (three workgroups per CU): whole rounds of
// workgroups, as for the u kernels; a partial last round costs a full one.
static int oz_for(const Grid &G) { return choose_oz(G, ((G.nx + 14) / 15) * ((G.ny + 14) / 15), 768); }

void launch_v_setup(const DevProblem &P, const double *U, hipStream_t s) {
  hipLaunchKernelGGL(k_v_setup, dim3((P.G.nx + 14) / 15, (P.G.ny + 14) / 15, P.G.nez), dim3(256), 0, s, P, U);
  VRhsOp op;
  op.U = U; op.p = P.p; op.b = P.bV; op.G = P.G;
  op.crhs = make_vconst(P).crhs;
  op.crack = P.hasCrackPressure;
  int OZ = oz_for(P.G);
  hipLaunchKernelGGL(k_v_rhs, mgrid(P.G, OZ), dim3(256), 0, s, op, OZ);
}

void launch_v_fixmask(const DevProblem &P, hipStream_t s) {
  hipLaunchKernelGGL(k_v_fixmask, dim3((unsigned)((P.nn + 255) / 256)), dim3(256), 0, s, P);
}

static VApplyOp make_vapply_op(const DevProblem &P, const double *x, double *y, int mode, const uint8_t *mask) {
  VApplyOp op;
  op.x = x; op.y = y; op.q = P.Vq; op.bV = P.bV; op.VIrrev = P.VIrrev; op.vfix = mask ? mask : P.vfix;
  op.G = P.G; op.ne = P.ne; op.C = make_vconst(P); op.bc = P.bcV;
  for (int f = 0; f < 6; f++) op.ftype[f] = (uint8_t)P.hbcV.face[f];
  for (int v = 0; v < 8; v++) op.vtype[v] = (uint8_t)P.hbcV.vertex[v];
  op.mode = mode; op.atnum = P.atnum; op.irrevtol = P.irrevtol;
  return op;
}

void launch_v_apply(const DevProblem &P, const double *x, double *y, int mode, hipStream_t s,
                    const uint8_t *mask) {
  const VApplyOp op = make_vapply_op(P, x, y, mode, mask);
  int OZ = oz_for(P.G);
  hipLaunchKernelGGL(k_v_apply, mgrid(P.G, OZ), dim3(256), 0, s, op, OZ);
}

void launch_v_apply_list(const DevProblem &P, const double *x, double *y, int mode, const int *list, int count,
                         hipStream_t s, const uint8_t *mask) {
  if (count <= 0) return;
  const VApplyOp op = make_vapply_op(P, x, y, mode, mask);
  hipLaunchKernelGGL(k_v_apply_list, dim3((unsigned)count), dim3(256), 0, s, op, list, v_chunk_cap(P.G));
}

// ------------------------------------------------ persistent multi-problem apply
// Masked Jacobian applies of several problems in one launch (serving engine):
// a work list of (member, tile x, tile y, z-chunk); the item count lives on the
// device and members whose flags[p * fstride] is 0 are skipped.
struct BVApply {
  VApplyOp op;
  int OZ;
};
__global__ __launch_bounds__(256, 3) void k_batch_v_apply(const BVApply *__restrict__ probs,
                                                       const BItem *__restrict__ items,
                                                       const int *__restrict__ dcount,
                                                       const double *__restrict__ flags, int fstride) {
  const int n = *dcount;
  for (int it = blockIdx.x; it < n; it += gridDim.x) {
    const BItem w = items[it];
    if (flags[(int64_t)w.p * fstride] == 0.) continue;  // uniform per block
    const VApplyOp &op = probs[w.p].op;  // read in place: the record is large
    const int OZ = probs[w.p].OZ;
    __syncthreads();  // LDS of the previous item is free
    march<double, 16, 16>(op.G, op, OZ, w.bx, w.by, w.bz);
  }
}
size_t vop_bytes() { return sizeof(BVApply); }
int v_oz(const Grid &G) { return oz_for(G); }
void make_vop(const DevProblem &P, const double *x, double *y, const uint8_t *mask, void *out) {
  BVApply *b = (BVApply *)out;
  VApplyOp &op = b->op;
  op.x = x; op.y = y; op.q = P.Vq; op.bV = P.bV; op.VIrrev = P.VIrrev; op.vfix = mask ? mask : P.vfix;
  op.G = P.G; op.ne = P.ne; op.C = make_vconst(P); op.bc = P.bcV;
  for (int f = 0; f < 6; f++) op.ftype[f] = (uint8_t)P.hbcV.face[f];
  for (int v = 0; v < 8; v++) op.vtype[v] = (uint8_t)P.hbcV.vertex[v];
  op.mode = 0; op.atnum = P.atnum; op.irrevtol = P.irrevtol;
  b->OZ = oz_for(P.G);
}
void launch_engine_v_apply(const void *d_ops, const BItem *items, const int *dcount, int grid, const double *flags,
                           int fstride, hipStream_t s) {
  hipLaunchKernelGGL(k_batch_v_apply, dim3(grid), dim3(256), 0, s, (const BVApply *)d_ops, items, dcount, flags,
                     fstride);
}

static VApplyCgOp make_vcg_op(const DevProblem &P, const double *pold, double *p, const double *z,
                              const double *beta, double *q, double *part, const uint8_t *mask) {
  VApplyCgOp op;
  op.x = pold; op.y = q; op.q = P.Vq; op.bV = P.bV; op.VIrrev = P.VIrrev; op.vfix = mask ? mask : P.vfix;
  op.G = P.G; op.ne = P.ne; op.C = make_vconst(P); op.bc = P.bcV;
  for (int f = 0; f < 6; f++) op.ftype[f] = (uint8_t)P.hbcV.face[f];
  for (int v = 0; v < 8; v++) op.vtype[v] = (uint8_t)P.hbcV.vertex[v];
  op.mode = 0; op.atnum = P.atnum; op.irrevtol = P.irrevtol;
  op.z = z; op.beta = beta; op.pw = p; op.part = part;
  return op;
}

int v_chunk_cap(const Grid &G) { return ((G.nx + 14) / 15) * ((G.ny + 14) / 15) * G.nz; }

void launch_v_chunk_list(const DevProblem &P, const uint8_t *mask, uint32_t *flag, int *list, hipStream_t s) {
  const int cap = v_chunk_cap(P.G), txy = cap / P.G.nz;
  // at most 32768 + txy chunks, whatever the band: partials fit the reduction buffers
  const int lmin = (cap + 32767) / 32768;
  hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)((cap + 255) / 256)), dim3(256), 0, s, flag, 0u, (int64_t)cap);
  hipLaunchKernelGGL(k_v_mark_planes, dim3((unsigned)((P.nn + 255) / 256)), dim3(256), 0, s, P.G, mask, flag);
  hipLaunchKernelGGL(k_v_list_chunks, dim3(1), dim3(VLT), 0, s, (const uint32_t *)flag, txy, P.G.nz, lmin, P.G.nz,
                     1024, list, cap);
}

void launch_v_dilated_list(const DevProblem &P, const uint32_t *flag, uint32_t *flag2, int *list2, hipStream_t s) {
  const int cap = v_chunk_cap(P.G), gx = (P.G.nx + 14) / 15, gy = (P.G.ny + 14) / 15, txy = gx * gy;
  const int lmin = (cap + 32767) / 32768;
  hipLaunchKernelGGL(k_v_dilate_planes, dim3((unsigned)((cap + 255) / 256)), dim3(256), 0, s, flag, flag2, gx, gy,
                     P.G.nz);
  hipLaunchKernelGGL(k_v_list_chunks, dim3(1), dim3(VLT), 0, s, (const uint32_t *)flag2, txy, P.G.nz, lmin, P.G.nz,
                     1024, list2, cap);
}

int launch_v_apply_cg_list(const DevProblem &P, const double *pold, double *p, const double *z, const double *beta,
                           double *q, double *part, const uint8_t *mask, const int *list, int count, hipStream_t s,
                           const double *done) {
  if (count <= 0) return 0;
  const VApplyCgOp op = make_vcg_op(P, pold, p, z, beta, q, part, mask);
  hipLaunchKernelGGL(k_v_apply_cg_list, dim3((unsigned)count), dim3(256), 0, s, op, list, v_chunk_cap(P.G), done);
  return count;
}

int launch_v_xr_list(const Grid &G, const int *list, int count, double *x, double *r, double *z, const double *p,
                     const double *q, const double *dinv, const double *alpha, double *part, hipStream_t s,
                     const double *done) {
  if (count <= 0) return 0;
  hipLaunchKernelGGL(k_v_xr_list, dim3((unsigned)count), dim3(256), 0, s, G, list, v_chunk_cap(G), x, r, z, p, q,
                     dinv, alpha, part, done);
  return count;
}

void launch_v_diag(const DevProblem &P, double *d, hipStream_t s, const uint8_t *mask) {
  VDiagOp op;
  op.d = d; op.q = P.Vq; op.vfix = mask ? mask : P.vfix; op.G = P.G; op.ne = P.ne; op.C = make_vconst(P);
  int OZ = oz_for(P.G);
  hipLaunchKernelGGL(k_v_diag, mgrid(P.G, OZ), dim3(256), 0, s, op, OZ);
}

void launch_v_diag_list(const DevProblem &P, double *d, const uint8_t *mask, const int *list, int count,
                        hipStream_t s) {
  if (count <= 0) return;
  VDiagOp op;
  op.d = d; op.q = P.Vq; op.vfix = mask ? mask : P.vfix; op.G = P.G; op.ne = P.ne; op.C = make_vconst(P);
  hipLaunchKernelGGL(k_v_diag_list, dim3((unsigned)count), dim3(256), 0, s, op, list, v_chunk_cap(P.G));
}

}  // namespace vfg
