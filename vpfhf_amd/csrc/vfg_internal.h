// vfg_internal.h -- device-side building blocks shared by the vfg kernels.
//
// Element basis: trilinear Q1 on an axis-aligned box, 3-point Gauss-Legendre
// per direction (27 points), exactly the rule of VFCartFEElement3DInit
// (VFCartFE.c:90-116, 216-269).  The 27-point tables are never stored: every
// quantity is evaluated by sum factorisation with the 1-D tables below.
//
// Node accumulation ("gather") without atomics: a workgroup owns a tile of
// (TX-1) x (TY-1) node columns and marches upward through a z-chunk of node
// planes.  One thread computes one element of the (TX x TY) element columns
// that touch the owned nodes (a one-element halo is recomputed), element
// contributions are combined across the x neighbour with a lane shuffle and
// across the y neighbour through LDS, and the plane below the current element
// layer is complete after each step.  Every node sums its 8 element
// contributions in a fixed order, so results are bitwise reproducible.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <type_traits>
#include <utility>

namespace vfg {

// 1-D Gauss tables on [0,1]: P0[g] = phi_0(xi_g), P1[g] = phi_1(xi_g) and the
// weights w[g] (times the element length), VFCartFE.c:100-113.
#define VFG_SQ35 0.77459666924148337704  // sqrt(3/5)
__host__ __device__ constexpr double gP0(int g) { return g == 0 ? (1. + VFG_SQ35) * .5 : g == 1 ? .5 : (1. - VFG_SQ35) * .5; }
__host__ __device__ constexpr double gP1(int g) { return g == 0 ? (1. - VFG_SQ35) * .5 : g == 1 ? .5 : (1. + VFG_SQ35) * .5; }
__host__ __device__ constexpr double gW(int g) { return g == 1 ? 8. / 18. : 5. / 18.; }

struct Grid {
  int nx, ny, nz;     // nodes
  int nex, ney, nez;  // elements
  double hx, hy, hz;  // uniform element size
};

__device__ __forceinline__ int64_t node_id(const Grid &G, int i, int j, int k) {
  return (int64_t)i + (int64_t)G.nx * ((int64_t)j + (int64_t)G.ny * k);
}
__device__ __forceinline__ int64_t cell_id(const Grid &G, int i, int j, int k) {
  return (int64_t)i + (int64_t)G.nex * ((int64_t)j + (int64_t)G.ney * k);
}

// Workgroups go to the eight XCDs round-robin by linear id, and each XCD has
// its own L2.  xcd_block renumbers a 3-D grid so that XCD k runs one
// contiguous range of blocks in (x, y, z) order: tiles that are neighbours in
// x and y, whose halo node lines overlap, then meet in the same L2.
struct Blk3 {
  int x, y, z;
};
__device__ __forceinline__ Blk3 xcd_block() {
  const unsigned gx = gridDim.x, gy = gridDim.y, n = gx * gy * gridDim.z;
  const unsigned id = blockIdx.x + gx * (blockIdx.y + gy * blockIdx.z);
  const unsigned per = n / 8, rem = n % 8, x = id % 8, loc = id / 8;
  const unsigned l = (x < rem ? x * (per + 1) : rem * (per + 1) + (x - rem) * per) + loc;
  return Blk3{(int)(l % gx), (int)((l / gx) % gy), (int)(l / (gx * gy))};
}

// tr(eps(u)) at Gauss point (gx, gy, gz) of an element from its 8 nodal
// displacements (local node a = (k*2+j)*2+i), accumulated per component over
// the nodes and summed as (d1u1 + d2u2) + d3u3: every kernel that branches on
// its sign (UNILATERAL_NOCOMPRESSION) evaluates it this way.
__device__ __forceinline__ double tr_eps_at(const double (&u)[8][3], int gx, int gy, int gz, double ihx,
                                            double ihy, double ihz) {
  double h0 = 0., h1 = 0., h2 = 0.;
  for (int a = 0; a < 8; a++) {
    const int i = a & 1, j = (a >> 1) & 1, k = a >> 2;
    const double px = i ? gP1(gx) : gP0(gx), py = j ? gP1(gy) : gP0(gy), pz = k ? gP1(gz) : gP0(gz);
    h0 += (i ? ihx : -ihx) * py * pz * u[a][0];
    h1 += (j ? ihy : -ihy) * px * pz * u[a][1];
    h2 += (k ? ihz : -ihz) * px * py * u[a][2];
  }
  return (h0 + h1) + h2;
}
// bit g = (gz*3+gy)*3+gx set where tr(eps(u)) < 0 (UNILATERAL_THRES = 0, VFMech.c:6)
__device__ __forceinline__ uint32_t compression_mask(const double (&u)[8][3], double ihx, double ihy, double ihz) {
  uint32_t m = 0u;
  for (int g = 0; g < 27; g++)
    if (tr_eps_at(u, g % 3, (g / 3) % 3, g / 9, ihx, ihy, ihz) < 0.) m |= 1u << g;
  return m;
}

// Dirichlet row set of one component, computed from node indices
// (MatApplyDirichletBC, VFCartFE.c:1147-1296): bit f of `face` set iff face f
// is not NONE; bit v of `vert` set iff vertex v is not NONE.
struct BcMask {
  uint8_t face[3];
  uint8_t vert[3];
};

__device__ __forceinline__ bool on_dirichlet(const Grid &G, uint8_t face, uint8_t vert, int i, int j,
                                             int k) {
  bool m = ((face & 1) && i == 0) || ((face & 2) && i == G.nx - 1) || ((face & 4) && j == 0) ||
           ((face & 8) && j == G.ny - 1) || ((face & 16) && k == 0) || ((face & 32) && k == G.nz - 1);
  if (vert) {
    int ix = (i == 0) ? 0 : (i == G.nx - 1) ? 1 : -1;
    int iy = (j == 0) ? 0 : (j == G.ny - 1) ? 1 : -1;
    int iz = (k == 0) ? 0 : (k == G.nz - 1) ? 1 : -1;
    if (ix >= 0 && iy >= 0 && iz >= 0 && ((vert >> (ix + 2 * iy + 4 * iz)) & 1)) m = true;
  }
  return m;
}

// wave64 / block reductions in FP64 with a fixed order (deterministic)
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
template <int NT>
__device__ __forceinline__ double block_sum(double v, double *sh /* NT/64 */) {
  v = wave_sum(v);
  const int w = threadIdx.x / 64, l = threadIdx.x % 64;
  if (l == 0) sh[w] = v;
  __syncthreads();
  double s = 0.;
  if (threadIdx.x == 0)
    for (int i = 0; i < NT / 64; i++) s += sh[i];
  __syncthreads();
  return s;  // valid in thread 0
}

// ---------------------------------------------------------------- marching
// Op interface:
//   static constexpr int NF  -- staged nodal scalar fields
//   static constexpr int NC  -- accumulated components per node
//   __device__ T load(int f, int64_t node)  -- value of staged field f at a node
//   __device__ void element(int ex,int ey,int ek, const T (&xin)[2][4][NF], T (&r)[8][NC])
//   __device__ void store(int i,int j,int k, const T (&acc)[NC], const T (&own)[NF])
// Optional per-element coefficients read one layer ahead (hides their latency):
//   __device__ A aux(int ex,int ey,int ek); element(ex,ey,ek, xin, aux, r)
// with A any trivially copyable type (a value, or a struct of values).
// Corners of an element are a = (k*2+j)*2+i; xin[k][j*2+i][f].
template <class Op, class = void>
struct has_aux : std::false_type {};
template <class Op>
struct has_aux<Op, std::void_t<decltype(&Op::aux)>> : std::true_type {};
// Ops that keep per-element data in a tile-blocked layout (one 16x16 slot
// block per workgroup tile and element layer; halo elements stored in every
// tile that computes them) declare TILED and take (tile, slot, ek) instead of
// (ex, ey, ek): a wavefront then reads each coefficient as one aligned,
// contiguous 512-byte segment.
template <class Op, class = void>
struct is_tiled : std::false_type {};
template <class Op>
struct is_tiled<Op, std::void_t<decltype(Op::TILED)>> : std::true_type {};
// tile-blocked layout: [tile][element layer][coefficient][TX*TY slots]
__host__ __device__ inline int64_t tiled_index(int tile, int ek, int nez, int ncoef, int slot, int coef, int nslot) {
  return (((int64_t)tile * nez + ek) * ncoef + coef) * nslot + slot;
}

// Ops whose store also reads data of the owned node that is not staged define
//   __device__ O own_load(int i, int j, int k)      (O trivially copyable)
//   store(i, j, k, acc, own, O o)
// the register march loads it one node plane ahead, beside the staged plane,
// so its latency is not exposed at the store.
template <class Op, class = void>
struct has_own : std::false_type {};
template <class Op>
struct has_own<Op, std::void_t<decltype(&Op::own_load)>> : std::true_type {};
template <class Op, class = void>
struct own_of {
  using type = int;
};
template <class Op>
struct own_of<Op, std::void_t<decltype(&Op::own_load)>> {
  using type = decltype(std::declval<const Op &>().own_load(0, 0, 0));
};

// Ops that also reduce over their owned nodes define
//   __device__ double store(...)      (returns the node's contribution)
//   __device__ void partial(int slot, double sum)
// the march sums the contributions of each thread, reduces them over the
// workgroup in a fixed order and hands the sum to partial(rslot, .).
template <class Op, class = void>
struct has_red : std::false_type {};
template <class Op>
struct has_red<Op, std::void_t<decltype(&Op::partial)>> : std::true_type {};

template <class Op, class = void>
struct aux_of {
  using type = int;
};
template <class Op>
struct aux_of<Op, std::void_t<decltype(&Op::aux)>> {
  using type = decltype(std::declval<const Op &>().aux(0, 0, 0));
};

template <typename T, int TX, int TY, class Op>
__device__ __forceinline__ void march(const Grid &G, const Op &op, int OZ, int bx, int by, int bz, int rslot = 0) {
  constexpr int NF = Op::NF, NC = Op::NC;
  constexpr int NT = TX * TY, OX = TX - 1, OY = TY - 1, SX = TX + 1, SY = TY + 1;
  constexpr int NPF = (SX * SY + NT - 1) / NT;  // staged nodes per thread
  constexpr bool AUX = has_aux<Op>::value;
  constexpr bool TILED = is_tiled<Op>::value;
  constexpr bool RED = has_red<Op>::value;
  __shared__ T s_in[NF][SY * SX];
  __shared__ T s_red[2 * NC][NT];
  const int tid = threadIdx.x, tx = tid % TX, ty = tid / TX;
  const int ox0 = bx * OX, oy0 = by * OY, oz0 = bz * OZ;
  const int ex = ox0 - 1 + tx, ey = oy0 - 1 + ty;
  const bool exy_ok = ex >= 0 && ex < G.nex && ey >= 0 && ey < G.ney;
  const bool owner = tx < OX && ty < OY && (ox0 + tx) < G.nx && (oy0 + ty) < G.ny;
  const int64_t plane = (int64_t)G.nx * G.ny;
  const int tile = bx + ((G.nx + OX - 1) / OX) * by;  // workgroup tile, for tiled ops

  // staged nodes of this thread: in-plane offsets, fixed for the whole march
  int64_t soff[NPF];
  bool sok[NPF];
#pragma unroll
  for (int q = 0; q < NPF; q++) {
    const int n = tid + q * NT;
    const int lx = n % SX, ly = n / SX;
    const int gx = ox0 - 1 + lx, gy = oy0 - 1 + ly;
    sok[q] = n < SX * SY && gx >= 0 && gx < G.nx && gy >= 0 && gy < G.ny;
    soff[q] = sok[q] ? (int64_t)gx + (int64_t)G.nx * gy : 0;
  }
  // The next node plane is loaded into registers while the current element
  // layer computes, and written to LDS after the next barrier.
  T pf[NPF][NF];
  auto fetch = [&](int iz) {
    const bool zok = iz >= 0 && iz < G.nz;  // uniform
#pragma unroll
    for (int q = 0; q < NPF; q++) {
      const bool ok = zok && sok[q];
      const int64_t nd = soff[q] + plane * iz;
#pragma unroll
      for (int f = 0; f < NF; f++) pf[q][f] = ok ? op.load(f, nd) : T(0);
    }
  };
  auto commit = [&]() {
#pragma unroll
    for (int q = 0; q < NPF; q++) {
      const int n = tid + q * NT;
      if (n < SX * SY)
#pragma unroll
        for (int f = 0; f < NF; f++) s_in[f][n] = pf[q][f];
    }
  };
  using AuxT = typename aux_of<Op>::type;
  auto aux_at = [&](int ek) -> AuxT {
    AuxT a{};
    if constexpr (AUX) {
      if (exy_ok && ek >= 0 && ek < G.nez) {
        if constexpr (TILED) a = op.aux(tile, tid, ek);
        else a = op.aux(ex, ey, ek);
      }
    }
    return a;
  };
  auto read_plane = [&](T (&xp)[4][NF]) {
#pragma unroll
    for (int q = 0; q < 4; q++)
#pragma unroll
      for (int f = 0; f < NF; f++) xp[q][f] = s_in[f][(ty + (q >> 1)) * SX + tx + (q & 1)];
  };
  T carry[NC];
#pragma unroll
  for (int c = 0; c < NC; c++) carry[c] = T(0);
  double racc = 0.;  // owned-node contributions (reducing ops)
  using OwnT = typename own_of<Op>::type;
  auto own_at = [&](int k) -> OwnT {
    OwnT o{};
    if constexpr (has_own<Op>::value)
      if (owner && k >= oz0 && k >= 0 && k < G.nz) o = op.own_load(ox0 + tx, oy0 + ty, k);
    return o;
  };

  // one element layer: bottom plane xb (node plane ek), top plane xt (ek + 1)
  auto layer = [&](int ek, const T (&xb)[4][NF], T (&xt)[4][NF], AuxT &aux_next, OwnT &own_next) {
    const AuxT aux_cur = aux_next;
    const OwnT own_cur = own_next;
    __syncthreads();
    commit();       // plane ek + 1
    fetch(ek + 2);  // in flight during this layer
    aux_next = aux_at(ek + 1);
    own_next = own_at(ek + 1);
    __syncthreads();
    read_plane(xt);
    T r[8][NC];
#pragma unroll
    for (int a = 0; a < 8; a++)
#pragma unroll
      for (int c = 0; c < NC; c++) r[a][c] = T(0);
    if (exy_ok && ek >= 0 && ek < G.nez) {
      T xin[2][4][NF];
#pragma unroll
      for (int q = 0; q < 4; q++)
#pragma unroll
        for (int f = 0; f < NF; f++) {
          xin[0][q][f] = xb[q][f];
          xin[1][q][f] = xt[q][f];
        }
      const int e0 = TILED ? tile : ex, e1 = TILED ? tid : ey;
      if constexpr (AUX) op.element(e0, e1, ek, xin, aux_cur, r);
      else op.element(e0, e1, ek, xin, r);
    }
    // x-combine: node (ex+1) gets corner i=1 of this element and corner i=0
    // of the element to the right (next lane inside the 16-lane row).
    T cx[2][2][NC];
#pragma unroll
    for (int kz = 0; kz < 2; kz++)
#pragma unroll
      for (int jy = 0; jy < 2; jy++)
#pragma unroll
        for (int c = 0; c < NC; c++) {
          T right = __shfl_down(r[(kz * 2 + jy) * 2 + 0][c], 1, TX);
          cx[kz][jy][c] = r[(kz * 2 + jy) * 2 + 1][c] + right;
        }
    // y-combine through LDS: node (ey+1) gets row jy=1 of this element row and
    // row jy=0 of the element row above.
#pragma unroll
    for (int kz = 0; kz < 2; kz++)
#pragma unroll
      for (int c = 0; c < NC; c++) s_red[kz * NC + c][tid] = cx[kz][0][c];
    __syncthreads();
    if (owner) {
      T bot[NC];
#pragma unroll
      for (int c = 0; c < NC; c++) {
        bot[c] = carry[c] + (cx[0][1][c] + s_red[c][tid + TX]);
        carry[c] = cx[1][1][c] + s_red[NC + c][tid + TX];
      }
      if (ek >= oz0 && ek >= 0) {
        if constexpr (has_own<Op>::value) {
          if constexpr (RED) racc += op.store(ox0 + tx, oy0 + ty, ek, bot, xb[3], own_cur);
          else op.store(ox0 + tx, oy0 + ty, ek, bot, xb[3], own_cur);
        } else {
          if constexpr (RED) racc += op.store(ox0 + tx, oy0 + ty, ek, bot, xb[3]);
          else op.store(ox0 + tx, oy0 + ty, ek, bot, xb[3]);
        }
      }
    }
  };

  T xa[4][NF], xb2[4][NF];  // ping-pong planes: no register copies between layers
  fetch(oz0 - 1);
  commit();
  fetch(oz0);
  AuxT aux_next = aux_at(oz0 - 1);
  OwnT own_next = own_at(oz0 - 1);
  __syncthreads();
  read_plane(xa);
  const int ek_end = min(oz0 + OZ - 1, G.nz - 1);
  for (int ek = oz0 - 1; ek <= ek_end; ek += 2) {
    layer(ek, xa, xb2, aux_next, own_next);
    if (ek + 1 > ek_end) break;
    layer(ek + 1, xb2, xa, aux_next, own_next);
  }
  if constexpr (RED) {
    __shared__ double s_rs[NT / 64];
    const double v = block_sum<NT>(racc, s_rs);
    if (tid == 0) op.partial(rslot, v);
  }
}

// march with the node planes kept in LDS.  Three planes rotate through LDS
// (bottom and top of the current element layer, and the one being committed
// for the next layer), so the element computation can re-read its corners
// instead of keeping them in registers:
//   __device__ void element_lds(int ex, int ey, int ek, const XS &xs, A aux, T (&r)[8][NC])
// with xs(k, q, f) the value of staged field f at corner (k, q = j*2+i).  The
// next plane is prefetched through registers as in march; the combination of
// the element results (x by lane shuffles, y through LDS, z carried) and the
// stores are those of march.  Two barriers per layer.  No tiled or reducing
// ops.  PF: the next plane is loaded one layer ahead into registers (held
// during the element computation); without PF it is loaded at the start of
// its layer and its latency is left to the other resident workgroup.
template <typename T, int TX, int TY, bool PF, class Op>
__device__ __forceinline__ void march_lds(const Grid &G, const Op &op, int OZ, int bx, int by, int bz,
                                          int rslot = 0) {
  constexpr int NF = Op::NF, NC = Op::NC;
  constexpr int NT = TX * TY, OX = TX - 1, OY = TY - 1, SX = TX + 1, SY = TY + 1, PL = SX * SY;
  constexpr int NPF = (PL + NT - 1) / NT;
  constexpr bool RED = has_red<Op>::value;
  static_assert(!is_tiled<Op>::value, "march_lds: untiled ops only");
  double racc = 0.;  // owned-node contributions (reducing ops)
  __shared__ T s_in[3][NF][PL];
  __shared__ T s_red[2 * NC][NT];
  const int tid = threadIdx.x, tx = tid % TX, ty = tid / TX;
  const int ox0 = bx * OX, oy0 = by * OY, oz0 = bz * OZ;
  const int ex = ox0 - 1 + tx, ey = oy0 - 1 + ty;
  const bool exy_ok = ex >= 0 && ex < G.nex && ey >= 0 && ey < G.ney;
  const bool owner = tx < OX && ty < OY && (ox0 + tx) < G.nx && (oy0 + ty) < G.ny;
  const int64_t plane = (int64_t)G.nx * G.ny;
  int64_t soff[NPF];
  bool sok[NPF];
#pragma unroll
  for (int q = 0; q < NPF; q++) {
    const int n = tid + q * NT;
    const int lx = n % SX, ly = n / SX;
    const int gx = ox0 - 1 + lx, gy = oy0 - 1 + ly;
    sok[q] = n < PL && gx >= 0 && gx < G.nx && gy >= 0 && gy < G.ny;
    soff[q] = sok[q] ? (int64_t)gx + (int64_t)G.nx * gy : 0;
  }
  T pf[NPF][NF];
  auto fetch = [&](int iz) {
    const bool zok = iz >= 0 && iz < G.nz;  // uniform
#pragma unroll
    for (int q = 0; q < NPF; q++) {
      const bool ok = zok && sok[q];
      const int64_t nd = soff[q] + plane * iz;
#pragma unroll
      for (int f = 0; f < NF; f++) pf[q][f] = ok ? op.load(f, nd) : T(0);
    }
  };
  auto commit = [&](int b) {
#pragma unroll
    for (int q = 0; q < NPF; q++) {
      const int n = tid + q * NT;
      if (n < PL)
#pragma unroll
        for (int f = 0; f < NF; f++) s_in[b][f][n] = pf[q][f];
    }
  };
  using AuxT = typename aux_of<Op>::type;
  auto aux_at = [&](int ek) -> AuxT {
    AuxT a{};
    if constexpr (has_aux<Op>::value)
      if (exy_ok && ek >= 0 && ek < G.nez) a = op.aux(ex, ey, ek);
    return a;
  };
  const int c0 = ty * SX + tx;  // corner (j, i) = (0, 0) of this thread's element in a plane
  T carry[NC];
#pragma unroll
  for (int c = 0; c < NC; c++) carry[c] = T(0);

  // planes rotate through the buffers: bottom bb, top bt, next bn
  int bb = 0, bt = 1, bn = 2;
  fetch(oz0 - 1);
  commit(bb);
  if constexpr (PF) fetch(oz0);
  AuxT aux_next = aux_at(oz0 - 1);
  const int ek_end = min(oz0 + OZ - 1, G.nz - 1);
  for (int ek = oz0 - 1; ek <= ek_end; ek++) {
    const AuxT aux_cur = aux_next;
    // buffer bt last held a plane read two layers ago; the barriers since
    // then order this write after those reads
    if constexpr (PF) {
      commit(bt);     // plane ek + 1
      fetch(ek + 2);  // in flight during this layer
    } else {
      fetch(ek + 1);
      commit(bt);
    }
    aux_next = aux_at(ek + 1);
    __syncthreads();
    T r[8][NC];
#pragma unroll
    for (int a = 0; a < 8; a++)
#pragma unroll
      for (int c = 0; c < NC; c++) r[a][c] = T(0);
    if (exy_ok && ek >= 0 && ek < G.nez) {
      const T *pb = &s_in[bb][0][c0], *pt = &s_in[bt][0][c0];
      auto xs = [pb, pt](int k, int q, int f) -> T {
        return (k ? pt : pb)[f * PL + (q >> 1) * SX + (q & 1)];
      };
      if constexpr (has_aux<Op>::value) op.element_lds(ex, ey, ek, xs, aux_cur, r);
      else op.element_lds(ex, ey, ek, xs, r);
    }
    T cx[2][2][NC];
#pragma unroll
    for (int kz = 0; kz < 2; kz++)
#pragma unroll
      for (int jy = 0; jy < 2; jy++)
#pragma unroll
        for (int c = 0; c < NC; c++) {
          T right = __shfl_down(r[(kz * 2 + jy) * 2 + 0][c], 1, TX);
          cx[kz][jy][c] = r[(kz * 2 + jy) * 2 + 1][c] + right;
        }
#pragma unroll
    for (int kz = 0; kz < 2; kz++)
#pragma unroll
      for (int c = 0; c < NC; c++) s_red[kz * NC + c][tid] = cx[kz][0][c];
    __syncthreads();
    if (owner) {
      T bot[NC], own[NF];
#pragma unroll
      for (int c = 0; c < NC; c++) {
        bot[c] = carry[c] + (cx[0][1][c] + s_red[c][tid + TX]);
        carry[c] = cx[1][1][c] + s_red[NC + c][tid + TX];
      }
      if (ek >= oz0 && ek >= 0) {
#pragma unroll
        for (int f = 0; f < NF; f++) own[f] = s_in[bb][f][c0 + SX + 1];
        if constexpr (RED) racc += op.store(ox0 + tx, oy0 + ty, ek, bot, own);
        else op.store(ox0 + tx, oy0 + ty, ek, bot, own);
      }
    }
    const int b = bb;
    bb = bt;
    bt = bn;
    bn = b;
  }
  if constexpr (RED) {
    __shared__ double s_rs[NT / 64];
    const double v = block_sum<NT>(racc, s_rs);
    if (tid == 0) op.partial(rslot, v);
  }
}

// ------------------------------------------------- Gauss-point walk
// Values (and, for the first NG fields, gradients) of trilinear nodal fields
// at the 27 Gauss points of one element, by sum factorisation one axis at a
// time: the 8 corner values f[a], a = (k*2+j)*2+i, are interpolated along z
// for each Gauss layer, then along y, then along x, with each derivative
// taken at the stage of its axis (d/dz of the corner pairs, constant in z).
// f(gx, gy, gz, val[NG+NV], grad[NG][3]) is called for every point in the
// order gz, gy, gx.  About 11 operations per point and gradient field and 3
// per value-only field, against 32 and 8 for the direct 8-corner sums.
// The corner values come from cv(q, a) (field q, corner a), read where they
// are used: from LDS, the array need not stay in registers for the walk.
// UZ, UY: unroll factors of the Gauss-layer and Gauss-row loops (1 keeps one
// layer's or row's values live, for walks with many fields).  With UZ = 1 the
// corners are re-read in every layer (a compiler barrier keeps the reads from
// being hoisted out of the loop, which would hold them all in registers).
template <int NG, int NV, int UZ = 3, int UY = 3, typename CV, typename F>
__device__ __forceinline__ void gauss_walk_cv(const CV &cv, double ihx, double ihy, double ihz, const F &f) {
  constexpr int NF = NG + NV;
  double dz[NG > 0 ? NG : 1][4];  // d/dz at the 4 (x, y) corners
#pragma unroll
  for (int q = 0; q < NG; q++)
#pragma unroll
    for (int c = 0; c < 4; c++) dz[q][c] = (cv(q, c + 4) - cv(q, c)) * ihz;
#pragma unroll UZ
  for (int gz = 0; gz < 3; gz++) {
    if constexpr (UZ == 1) __asm__ __volatile__("" ::: "memory");
    const double pz = gP1(gz);
    double z[NF][4];
#pragma unroll
    for (int q = 0; q < NF; q++)
#pragma unroll
      for (int c = 0; c < 4; c++) {
        const double f0 = cv(q, c);
        z[q][c] = f0 + pz * (cv(q, c + 4) - f0);
      }
#pragma unroll UY
    for (int gy = 0; gy < 3; gy++) {
      const double py = gP1(gy);
      double y[NF][2], dyv[NG > 0 ? NG : 1][2], dzy[NG > 0 ? NG : 1][2];
#pragma unroll
      for (int q = 0; q < NF; q++)
#pragma unroll
        for (int i = 0; i < 2; i++) {
          const double d = z[q][i + 2] - z[q][i];
          y[q][i] = z[q][i] + py * d;
          if (q < NG) {
            dyv[q][i] = d * ihy;
            dzy[q][i] = dz[q][i] + py * (dz[q][i + 2] - dz[q][i]);
          }
        }
#pragma unroll
      for (int gx = 0; gx < 3; gx++) {
        const double px = gP1(gx);
        double val[NF], grad[NG > 0 ? NG : 1][3];
#pragma unroll
        for (int q = 0; q < NF; q++) {
          const double d = y[q][1] - y[q][0];
          val[q] = y[q][0] + px * d;
          if (q < NG) {
            grad[q][0] = d * ihx;
            grad[q][1] = dyv[q][0] + px * (dyv[q][1] - dyv[q][0]);
            grad[q][2] = dzy[q][0] + px * (dzy[q][1] - dzy[q][0]);
          }
        }
        f(gx, gy, gz, val, grad);
      }
    }
  }
}
template <int NG, int NV, typename F>
__device__ __forceinline__ void gauss_walk(const double (&fld)[NG + NV][8], double ihx, double ihy, double ihz,
                                           const F &f) {
  gauss_walk_cv<NG, NV>([&](int q, int a) { return fld[q][a]; }, ihx, ihy, ihz, f);
}

template <typename T>
__device__ __forceinline__ T interp(int g, T f0, T f1) {
  return T(gP0(g)) * f0 + T(gP1(g)) * f1;
}

// the marches at a renumbered block (xcd_block)
template <typename T, int TX, int TY, class Op>
__device__ __forceinline__ void march(const Grid &G, const Op &op, int OZ, Blk3 b, int rslot = 0) {
  march<T, TX, TY>(G, op, OZ, b.x, b.y, b.z, rslot);
}
template <typename T, int TX, int TY, bool PF, class Op>
__device__ __forceinline__ void march_lds(const Grid &G, const Op &op, int OZ, Blk3 b, int rslot = 0) {
  march_lds<T, TX, TY, PF>(G, op, OZ, b.x, b.y, b.z, rslot);
}

}  // namespace vfg
