// vfg_u.hip -- u-subproblem operators on gfx950.
//
// Element bilinear form (VF_BilinearFormU3D_local, VFMech.c:227-290):
//   Ke[(a,c1),(b,c2)] = sum_g W_g s_g [lam d_c1 phi_a d_c2 phi_b
//                                      + mu d_c2 phi_a d_c1 phi_b
//                                      + delta_c1c2 mu grad phi_a . grad phi_b]
//   s_g = v_g^2 + eta.
// Applied matrix-free: (Ke x)_{a,c} = sum_g sum_d sigma_cd(g) d_d phi_a(g) with
//   sigma = W s (lam tr(H) I + mu (H + H^T)),  H_cd = d_d u_c  (displacement
//   gradient of x).  All 27-point quantities are obtained by sum
// factorisation with the 1-D Gauss tables; d_x u_c is constant along x, so it
// takes 9 distinct values per component.
#include "vfg_internal.h"
#include "vfg_kernels.h"
#include "vfg_batch_apply.h"
#include "vfg_handle.h"
#include "vfg_mg.h"
#include "vfg_bmg.h"

namespace vfg {

// Apply of one element: xin holds (x0,x1,x2,v) at the 8 corners.
// cE = E_e * hx*hy*hz; lam1/mu1 are the Lame constants for E = 1.
// The gradients are carried pre-scaled by mu (H' = mu H), so that
//   sigma / (W s cE) = lr tr(H') I + H' + H'^T,  lr = lam / mu,
// and each stress entry is accumulated with one FMA by its weight W s cE.
// UNI (UNILATERAL_NOCOMPRESSION, VFMech.c:299-376): at the Gauss points set in
// cm the spherical part keeps kappa (1 + eta) instead of kappa s, i.e.
//   sigma / (W cE) = (kr - 2 s / 3) tr(H') I + s (H' + H'^T),
//   kr = kappa (1 + eta) / mu = (lr + 2/3)(1 + eta).
template <typename T, bool UNI = false>
__device__ __forceinline__ void u_elem_apply(const T (&X)[2][4][4], T cE, T lam1, T mu1, T eta,
                                             T ihx, T ihy, T ihz, T (&r)[8][3], uint32_t cm = 0u) {
  const T lr = lam1 / mu1;
  const T kr = (lr + T(2. / 3.)) * (T(1) + eta);
  const T mx = mu1 * ihx, my = mu1 * ihy, mz = mu1 * ihz;
  // differences along each direction, scaled by mu/h
  T Dx[3][2][2], Dy[3][2][2], Dz[3][2][2];  // Dx[c][k][j], Dy[c][k][i], Dz[c][j][i]
#pragma unroll
  for (int c = 0; c < 3; c++)
#pragma unroll
    for (int p = 0; p < 2; p++)
#pragma unroll
      for (int q = 0; q < 2; q++) {
        Dx[c][p][q] = (X[p][q * 2 + 1][c] - X[p][q * 2 + 0][c]) * mx;
        Dy[c][p][q] = (X[p][2 + q][c] - X[p][q][c]) * my;
        Dz[c][p][q] = (X[1][p * 2 + q][c] - X[0][p * 2 + q][c]) * mz;
      }
  // mu d_z u_c at (gx,gy): constant in z
  T Gz[3][3][3];  // [c][gy][gx]
#pragma unroll
  for (int c = 0; c < 3; c++)
#pragma unroll
    for (int gx = 0; gx < 3; gx++) {
      T a0 = interp<T>(gx, Dz[c][0][0], Dz[c][0][1]);
      T a1 = interp<T>(gx, Dz[c][1][0], Dz[c][1][1]);
#pragma unroll
      for (int gy = 0; gy < 3; gy++) Gz[c][gy][gx] = interp<T>(gy, a0, a1);
    }
  // v interpolated in x and y on both z-faces
  T vxy[2][3][3];  // [k][gy][gx]
#pragma unroll
  for (int k = 0; k < 2; k++)
#pragma unroll
    for (int gx = 0; gx < 3; gx++) {
      T a0 = interp<T>(gx, X[k][0][3], X[k][1][3]);
      T a1 = interp<T>(gx, X[k][2][3], X[k][3][3]);
#pragma unroll
      for (int gy = 0; gy < 3; gy++) vxy[k][gy][gx] = interp<T>(gy, a0, a1);
    }

  T Rx[3][2][2], Ry[3][2][2], Sz[3][3][3];  // Rx[c][k][j], Ry[c][k][i], Sz[c][gy][gx]
#pragma unroll
  for (int c = 0; c < 3; c++)
#pragma unroll
    for (int p = 0; p < 2; p++)
#pragma unroll
      for (int q = 0; q < 2; q++) Rx[c][p][q] = Ry[c][p][q] = T(0);
#pragma unroll
  for (int c = 0; c < 3; c++)
#pragma unroll
    for (int gy = 0; gy < 3; gy++)
#pragma unroll
      for (int gx = 0; gx < 3; gx++) Sz[c][gy][gx] = T(0);

#pragma unroll
  for (int gz = 0; gz < 3; gz++) {
    T Gx[3][3], Gy[3][3];  // Gx[c][gy], Gy[c][gx]
#pragma unroll
    for (int c = 0; c < 3; c++) {
      T bx0 = interp<T>(gz, Dx[c][0][0], Dx[c][1][0]);
      T bx1 = interp<T>(gz, Dx[c][0][1], Dx[c][1][1]);
      T by0 = interp<T>(gz, Dy[c][0][0], Dy[c][1][0]);
      T by1 = interp<T>(gz, Dy[c][0][1], Dy[c][1][1]);
#pragma unroll
      for (int g = 0; g < 3; g++) {
        Gx[c][g] = interp<T>(g, bx0, bx1);
        Gy[c][g] = interp<T>(g, by0, by1);
      }
    }
    T Sx[3][3], Sy[3][3];  // Sx[c][gy] = sum_gx sigma_c0, Sy[c][gx] = sum_gy sigma_c1
#pragma unroll
    for (int c = 0; c < 3; c++)
#pragma unroll
      for (int g = 0; g < 3; g++) Sx[c][g] = Sy[c][g] = T(0);
    const T wz = cE * T(gW(gz));
#pragma unroll
    for (int gy = 0; gy < 3; gy++) {
      const T wzy = wz * T(gW(gy));
#pragma unroll
      for (int gx = 0; gx < 3; gx++) {
        const T v = interp<T>(gz, vxy[0][gy][gx], vxy[1][gy][gx]);
        const T wq = wzy * T(gW(gx));
        const T cq = wq * (v * v + eta);
        const T H00 = Gx[0][gy], H01 = Gy[0][gx], H02 = Gz[0][gy][gx];
        const T H10 = Gx[1][gy], H11 = Gy[1][gx], H12 = Gz[1][gy][gx];
        const T H20 = Gx[2][gy], H21 = Gy[2][gx], H22 = Gz[2][gy][gx];
        const T t01 = H01 + H10, t02 = H02 + H20, t12 = H12 + H21;
        if constexpr (UNI) {
          const bool comp = (cm >> ((gz * 3 + gy) * 3 + gx)) & 1u;
          const T lw = comp ? wq * (kr - T(2. / 3.) * (v * v + eta)) : cq * lr;
          const T L = lw * (H00 + H11 + H22);
          const T c2 = cq + cq;
          Sx[0][gy] += L + c2 * H00; Sx[1][gy] += cq * t01; Sx[2][gy] += cq * t02;
          Sy[0][gx] += cq * t01; Sy[1][gx] += L + c2 * H11; Sy[2][gx] += cq * t12;
          Sz[0][gy][gx] += cq * t02; Sz[1][gy][gx] += cq * t12; Sz[2][gy][gx] += L + c2 * H22;
        } else {
          const T lt = lr * (H00 + H11 + H22);
          const T t00 = lt + T(2) * H00, t11 = lt + T(2) * H11, t22 = lt + T(2) * H22;
          Sx[0][gy] += cq * t00; Sx[1][gy] += cq * t01; Sx[2][gy] += cq * t02;
          Sy[0][gx] += cq * t01; Sy[1][gx] += cq * t11; Sy[2][gx] += cq * t12;
          Sz[0][gy][gx] += cq * t02; Sz[1][gy][gx] += cq * t12; Sz[2][gy][gx] += cq * t22;
        }
      }
    }
    const T pz0 = T(gP0(gz)), pz1 = T(gP1(gz));
#pragma unroll
    for (int c = 0; c < 3; c++) {
      T ax0 = T(0), ax1 = T(0), ay0 = T(0), ay1 = T(0);
#pragma unroll
      for (int g = 0; g < 3; g++) {
        ax0 += T(gP0(g)) * Sx[c][g]; ax1 += T(gP1(g)) * Sx[c][g];
        ay0 += T(gP0(g)) * Sy[c][g]; ay1 += T(gP1(g)) * Sy[c][g];
      }
      Rx[c][0][0] += pz0 * ax0; Rx[c][0][1] += pz0 * ax1;
      Rx[c][1][0] += pz1 * ax0; Rx[c][1][1] += pz1 * ax1;
      Ry[c][0][0] += pz0 * ay0; Ry[c][0][1] += pz0 * ay1;
      Ry[c][1][0] += pz1 * ay0; Ry[c][1][1] += pz1 * ay1;
    }
  }
  // project Sz onto the bilinear basis in (x,y); scale each direction by 1/h once
#pragma unroll
  for (int c = 0; c < 3; c++) {
    T b[3][2];  // [gy][i]
#pragma unroll
    for (int gy = 0; gy < 3; gy++) {
      T a0 = T(0), a1 = T(0);
#pragma unroll
      for (int gx = 0; gx < 3; gx++) { a0 += T(gP0(gx)) * Sz[c][gy][gx]; a1 += T(gP1(gx)) * Sz[c][gy][gx]; }
      b[gy][0] = a0; b[gy][1] = a1;
    }
    T Rz[2][2];  // [j][i]
#pragma unroll
    for (int i = 0; i < 2; i++) {
      T a0 = T(0), a1 = T(0);
#pragma unroll
      for (int gy = 0; gy < 3; gy++) { a0 += T(gP0(gy)) * b[gy][i]; a1 += T(gP1(gy)) * b[gy][i]; }
      Rz[0][i] = a0 * ihz; Rz[1][i] = a1 * ihz;
    }
    T sx[2][2], sy[2][2];
#pragma unroll
    for (int p = 0; p < 2; p++)
#pragma unroll
      for (int q = 0; q < 2; q++) { sx[p][q] = Rx[c][p][q] * ihx; sy[p][q] = Ry[c][p][q] * ihy; }
#pragma unroll
    for (int k = 0; k < 2; k++)
#pragma unroll
      for (int j = 0; j < 2; j++)
#pragma unroll
        for (int i = 0; i < 2; i++) {
          const T vx = i ? sx[k][j] : -sx[k][j];
          const T vy = j ? sy[k][i] : -sy[k][i];
          const T vz = k ? Rz[j][i] : -Rz[j][i];
          r[(k * 2 + j) * 2 + i][c] = vx + vy + vz;
        }
  }
}

// The same element apply for march_lds: the corners are re-read from LDS in
// every Gauss layer gz (xs(k, q, f)), so that only one layer's
// interpolants are live and the FP64 kernel fits 256 registers (two waves
// per SIMD).  Per layer: the corners interpolated to the layer (Uz, Vz) and
// their z-differences (Dz); x-derivatives Gx[c][gy] for the layer; then
// gx-outer, gy-inner over the 9 points, with interpolation written as
// a + p1 (b - a) (one FMA per point and value).  Accumulation and the final
// projections are those of u_elem_apply.  UNILATERAL_NONE only: the
// NOCOMPRESSION apply keeps the register march.
// QS != nullptr (stored coefficients): the 27 products W_g s_g E_e vol of the
// element are read from QS[g * QSTRIDE], g = (gz * 3 + gy) * 3 + gx, instead
// of being formed from the nodal v (then xs needs no field 3).
template <typename T, typename XS, typename OUT, int QSTRIDE>
__device__ __forceinline__ void u_elem_lds_impl(const XS &xs, T cE, T lam1, T mu1, T eta, T ihx, T ihy, T ihz,
                                                const OUT &out, const T *__restrict__ QS) {
  constexpr bool STORED = QSTRIDE > 0;
  const T lr = lam1 / mu1;
  const T mx = mu1 * ihx, my = mu1 * ihy, mz = mu1 * ihz;
  // Rx[c][k][j], Ry[c][k][i]; Sz[c][gy][i]: the z-stress summed over the Gauss
  // layers and projected onto x at once (18 accumulators rather than 27: the
  // element then fits two waves per SIMD without spilling)
  T Rx[3][2][2], Ry[3][2][2], Sz[3][3][2];
#pragma unroll
  for (int c = 0; c < 3; c++) {
#pragma unroll
    for (int p = 0; p < 2; p++)
#pragma unroll
      for (int q = 0; q < 2; q++) Rx[c][p][q] = Ry[c][p][q] = T(0);
#pragma unroll
    for (int p = 0; p < 3; p++)
#pragma unroll
      for (int q = 0; q < 2; q++) Sz[c][p][q] = T(0);
  }
#pragma unroll
  for (int gz = 0; gz < 3; gz++) {
    const T pz0 = T(gP0(gz)), pz1 = T(gP1(gz));
    // compiler barrier: the corners are read again from LDS for this layer
    // rather than kept live across layers
    asm volatile("" ::: "memory");
    T Dz[4][3], Vz[4], Gx[3][3], by0[3], byd[3];
    T Ql[3][3];  // stored coefficients of the layer [gx][gy]
    if constexpr (STORED) {
#pragma unroll
      for (int gx = 0; gx < 3; gx++)
#pragma unroll
        for (int gy = 0; gy < 3; gy++) Ql[gx][gy] = QS[((gz * 3 + gy) * 3 + gx) * QSTRIDE];
    }
    {
      T Uz[4][3];
#pragma unroll
      for (int q = 0; q < 4; q++) {
#pragma unroll
        for (int c = 0; c < 3; c++) {
          const T a = xs(0, q, c), d = xs(1, q, c) - a;
          Uz[q][c] = a + pz1 * d;
          Dz[q][c] = d * mz;
        }
        if constexpr (!STORED) {
          const T a = xs(0, q, 3);
          Vz[q] = a + pz1 * (xs(1, q, 3) - a);
        }
      }
#pragma unroll
      for (int c = 0; c < 3; c++) {
        const T bx0 = (Uz[1][c] - Uz[0][c]) * mx, bxd = (Uz[3][c] - Uz[2][c]) * mx - bx0;
#pragma unroll
        for (int gy = 0; gy < 3; gy++) Gx[c][gy] = bx0 + T(gP1(gy)) * bxd;
        by0[c] = (Uz[2][c] - Uz[0][c]) * my;
        byd[c] = (Uz[3][c] - Uz[1][c]) * my - by0[c];
      }
    }
    T Dzx0[3], Dzx2[3];
#pragma unroll
    for (int c = 0; c < 3; c++) {
      Dzx0[c] = Dz[1][c] - Dz[0][c];
      Dzx2[c] = Dz[3][c] - Dz[2][c];
    }
    T vdx0 = T(0), vdx2 = T(0);
    if constexpr (!STORED) { vdx0 = Vz[1] - Vz[0]; vdx2 = Vz[3] - Vz[2]; }
    T Sx[3][3], ay[3][2];  // Sx[c][gy] = sum_gx sigma_c0; ay[c][i] = y-projection of sum_gy sigma_c1
#pragma unroll
    for (int c = 0; c < 3; c++) {
#pragma unroll
      for (int g = 0; g < 3; g++) Sx[c][g] = T(0);
      ay[c][0] = ay[c][1] = T(0);
    }
    const T wz = cE * T(gW(gz));
#pragma unroll
    for (int gx = 0; gx < 3; gx++) {
      const T px = T(gP1(gx));
      T Gy[3], az0[3], azd[3], Sy[3];
#pragma unroll
      for (int c = 0; c < 3; c++) {
        Gy[c] = by0[c] + px * byd[c];
        az0[c] = Dz[0][c] + px * Dzx0[c];
        azd[c] = Dz[2][c] + px * Dzx2[c] - az0[c];
        Sy[c] = T(0);
      }
      T av0 = T(0), avd = T(0), w0 = T(0), w1 = T(0);
      if constexpr (!STORED) {
        av0 = Vz[0] + px * vdx0;
        avd = Vz[2] + px * vdx2 - av0;
        const T wzx = wz * T(gW(gx));
        w0 = wzx * T(gW(0));
        w1 = wzx * T(gW(1));
      }
#pragma unroll
      for (int gy = 0; gy < 3; gy++) {
        T cq;
        if constexpr (STORED) cq = Ql[gx][gy];
        else {
          const T v = av0 + T(gP1(gy)) * avd;
          cq = (gy == 1 ? w1 : w0) * (v * v + eta);
        }
        const T H00 = Gx[0][gy], H01 = Gy[0], H02 = az0[0] + T(gP1(gy)) * azd[0];
        const T H10 = Gx[1][gy], H11 = Gy[1], H12 = az0[1] + T(gP1(gy)) * azd[1];
        const T H20 = Gx[2][gy], H21 = Gy[2], H22 = az0[2] + T(gP1(gy)) * azd[2];
        const T t01 = H01 + H10, t02 = H02 + H20, t12 = H12 + H21;
        const T lt = lr * (H00 + H11 + H22);
        const T t00 = lt + T(2) * H00, t11 = lt + T(2) * H11, t22 = lt + T(2) * H22;
        Sx[0][gy] += cq * t00; Sx[1][gy] += cq * t01; Sx[2][gy] += cq * t02;
        Sy[0] += cq * t01; Sy[1] += cq * t11; Sy[2] += cq * t12;
        const T z0 = cq * t02, z1 = cq * t12, z2 = cq * t22;
        Sz[0][gy][0] += T(gP0(gx)) * z0; Sz[0][gy][1] += T(gP1(gx)) * z0;
        Sz[1][gy][0] += T(gP0(gx)) * z1; Sz[1][gy][1] += T(gP1(gx)) * z1;
        Sz[2][gy][0] += T(gP0(gx)) * z2; Sz[2][gy][1] += T(gP1(gx)) * z2;
      }
#pragma unroll
      for (int c = 0; c < 3; c++) {
        ay[c][0] += T(gP0(gx)) * Sy[c];
        ay[c][1] += T(gP1(gx)) * Sy[c];
      }
    }
#pragma unroll
    for (int c = 0; c < 3; c++) {
      T ax0 = T(0), ax1 = T(0);
#pragma unroll
      for (int g = 0; g < 3; g++) { ax0 += T(gP0(g)) * Sx[c][g]; ax1 += T(gP1(g)) * Sx[c][g]; }
      Rx[c][0][0] += pz0 * ax0; Rx[c][0][1] += pz0 * ax1;
      Rx[c][1][0] += pz1 * ax0; Rx[c][1][1] += pz1 * ax1;
      Ry[c][0][0] += pz0 * ay[c][0]; Ry[c][0][1] += pz0 * ay[c][1];
      Ry[c][1][0] += pz1 * ay[c][0]; Ry[c][1][1] += pz1 * ay[c][1];
    }
  }
#pragma unroll
  for (int c = 0; c < 3; c++) {
    const T (&b)[3][2] = Sz[c];  // [gy][i]
    T Rz[2][2];  // [j][i]
#pragma unroll
    for (int i = 0; i < 2; i++) {
      T a0 = T(0), a1 = T(0);
#pragma unroll
      for (int gy = 0; gy < 3; gy++) { a0 += T(gP0(gy)) * b[gy][i]; a1 += T(gP1(gy)) * b[gy][i]; }
      Rz[0][i] = a0 * ihz; Rz[1][i] = a1 * ihz;
    }
#pragma unroll
    for (int k = 0; k < 2; k++)
#pragma unroll
      for (int j = 0; j < 2; j++)
#pragma unroll
        for (int i = 0; i < 2; i++) {
          const T vx = (i ? Rx[c][k][j] : -Rx[c][k][j]) * ihx;
          const T vy = (j ? Ry[c][k][i] : -Ry[c][k][i]) * ihy;
          const T vz = k ? Rz[j][i] : -Rz[j][i];
          out((k * 2 + j) * 2 + i, c, vx + vy + vz);
        }
  }
}

// recomputing form: s_g = v_g^2 + eta from the staged nodal v (field 3)
template <typename T, typename XS, typename OUT>
__device__ __forceinline__ void u_elem_apply_lds_out(const XS &xs, T cE, T lam1, T mu1, T eta, T ihx, T ihy, T ihz,
                                                     const OUT &out) {
  u_elem_lds_impl<T, XS, OUT, 0>(xs, cE, lam1, mu1, eta, ihx, ihy, ihz, out, nullptr);
}
// stored form: the 27 weights W_g s_g E_e vol at QS[g * QSTRIDE]
template <int QSTRIDE, typename T, typename XS, typename OUT>
__device__ __forceinline__ void u_elem_apply_lds_stored(const XS &xs, T lam1, T mu1, T ihx, T ihy, T ihz,
                                                        const OUT &out, const T *__restrict__ QS) {
  static_assert(QSTRIDE > 0, "stored weights need their stride");
  u_elem_lds_impl<T, XS, OUT, QSTRIDE>(xs, T(0), lam1, mu1, T(0), ihx, ihy, ihz, out, QS);
}
// the same with the 24 outputs in registers
template <typename T, typename XS>
__device__ __forceinline__ void u_elem_apply_lds(const XS &xs, T cE, T lam1, T mu1, T eta, T ihx, T ihy, T ihz,
                                                 T (&r)[8][3]) {
  u_elem_apply_lds_out<T>(xs, cE, lam1, mu1, eta, ihx, ihy, ihz, [&](int a, int c, T v) { r[a][c] = v; });
}

template <typename T>
struct UAuxM {
  T E;
  uint32_t m;  // compressed Gauss points of the Jacobian state
};

template <typename T, bool UNI = false>
struct UApplyOp {
  static constexpr int NF = 4, NC = 3;
  const T *x;
  T *y;
  const T *v;
  const double *E_elem;
  const uint32_t *cmask;  // UNI only
  Grid G;
  BcMask bc;
  T E, lam1, mu1, eta, vol, ihx, ihy, ihz;
  __device__ T load(int f, int64_t n) const { return f < 3 ? x[3 * n + f] : v[n]; }
  __device__ auto aux(int ex, int ey, int ek) const {
    const int64_t c = cell_id(G, ex, ey, ek);
    const T Ee = E_elem ? T(E_elem[c]) : E;
    if constexpr (UNI) return UAuxM<T>{Ee, cmask[c]};
    else return Ee;
  }
  template <typename A>
  __device__ void element(int, int, int, const T (&xin)[2][4][4], A a, T (&r)[8][3]) const {
    if constexpr (UNI) u_elem_apply<T, true>(xin, a.E * vol, lam1, mu1, eta, ihx, ihy, ihz, r, a.m);
    else u_elem_apply<T>(xin, a * vol, lam1, mu1, eta, ihx, ihy, ihz, r);
  }
  template <typename XS, typename A>
  __device__ void element_lds(int, int, int, const XS &xs, A a, T (&r)[8][3]) const {
    static_assert(!UNI, "the NOCOMPRESSION apply uses the register march");
    u_elem_apply_lds<T>(xs, a * vol, lam1, mu1, eta, ihx, ihy, ihz, r);
  }
  // the element's outputs handed to out(corner, component, value) as they are formed
  template <typename XS, typename A, typename OUT>
  __device__ void element_out(int, int, int, const XS &xs, A a, const OUT &out) const {
    static_assert(!UNI, "the NOCOMPRESSION apply uses the register march");
    u_elem_apply_lds_out<T>(xs, a * vol, lam1, mu1, eta, ihx, ihy, ihz, out);
  }
  __device__ void store(int i, int j, int k, const T (&acc)[3], const T (&own)[4]) const {
    int64_t n = node_id(G, i, j, k);
#pragma unroll
    for (int c = 0; c < 3; c++)
      y[3 * n + c] = on_dirichlet(G, bc.face[c], bc.vert[c], i, j, k) ? own[c] : acc[c];
  }
};

// FP32: at most 256 registers, two waves per SIMD
template <typename T, int TX, int TY, bool UNI = false>
__global__ __launch_bounds__(TX *TY, sizeof(T) == 4 ? 2 : 1) void k_u_apply(UApplyOp<T, UNI> op, int OZ) {
  march<T, TX, TY>(op.G, op, OZ, xcd_block());
}

// FP32 multigrid apply fused with its Chebyshev(1)-Jacobi step (mg_cheb_entry
// with one step, vfg_solve.hip's V-cycle):
//   pre  (from x = 0): x = c2 D^-1 b staged on the fly; stores x and A x (the
//        restriction forms b - A x on the free rows);
//   post: xo = x + c2 D^-1 (b - A x), A x = x on Dirichlet rows, into a
//        buffer other than x (neighbouring tiles still stage x); b and D^-1
//        of the owned node are loaded a node plane ahead (own_load).  The
//        rz form also reduces r . xo (r: the FP64 residual the V-cycle
//        preconditions) over the owned rows, one partial per workgroup.
struct UMgPreOp : UApplyOp<float, false> {
  const float *b, *dinv;
  float *xo, *ro;
  float c2;
  __device__ float load(int f, int64_t n) const { return f == 3 ? v[n] : c2 * dinv[3 * n + f] * b[3 * n + f]; }
  __device__ void store(int i, int j, int k, const float (&acc)[3], const float (&own)[4]) const {
    const int64_t n = node_id(G, i, j, k);
#pragma unroll
    for (int c = 0; c < 3; c++) {
      xo[3 * n + c] = own[c];
      ro[3 * n + c] = acc[c];
    }
  }
};
struct MgOwn {
  float b[3], di[3];
};
struct MgOwnR {
  float b[3], di[3];
  double r[3];
};
template <bool RZ>
struct UMgPostOp : UApplyOp<float, false> {
  const float *b, *dinv;
  float *xo;
  float c2;
  const double *r64;  // RZ
  double *part;       // RZ
  __device__ auto own_load(int i, int j, int k) const {
    const int64_t n = node_id(G, i, j, k);
    std::conditional_t<RZ, MgOwnR, MgOwn> o;
#pragma unroll
    for (int c = 0; c < 3; c++) {
      o.b[c] = b[3 * n + c];
      o.di[c] = dinv[3 * n + c];
      if constexpr (RZ) o.r[c] = r64[3 * n + c];
    }
    return o;
  }
  template <typename O>
  __device__ auto store(int i, int j, int k, const float (&acc)[3], const float (&own)[4], const O &o) const {
    const int64_t n = node_id(G, i, j, k);
    double rz = 0.;
#pragma unroll
    for (int c = 0; c < 3; c++) {
      const float yv = on_dirichlet(G, bc.face[c], bc.vert[c], i, j, k) ? own[c] : acc[c];
      const float z = own[c] + c2 * o.di[c] * (o.b[c] - yv);
      xo[3 * n + c] = z;
      if constexpr (RZ) rz += o.r[c] * (double)z;
    }
    if constexpr (RZ) return rz;
  }
};
struct UMgPostRzOp : UMgPostOp<true> {
  __device__ void partial(int slot, double v) const { part[slot] = v; }
};
template <int TX, int TY>
__global__ __launch_bounds__(TX *TY, 2) void k_u_mg_pre(UMgPreOp op, int OZ) {
  march<float, TX, TY>(op.G, op, OZ, xcd_block());
}
template <int TX, int TY>
__global__ __launch_bounds__(TX *TY, 2) void k_u_mg_post(UMgPostOp<false> op, int OZ) {
  march<float, TX, TY>(op.G, op, OZ, xcd_block());
}
template <int TX, int TY>
__global__ __launch_bounds__(TX *TY, 2) void k_u_mg_post_rz(UMgPostRzOp op, int OZ) {
  march<float, TX, TY>(op.G, op, OZ, xcd_block(), blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z));
}

// FP64 apply with the corners in LDS: at most 256 registers, two waves per
// SIMD; the next plane is loaded at the start of its layer (the other
// resident workgroup covers the latency; a register prefetch measured slower)
__global__ __launch_bounds__(256, 2) void k_u_apply_lds(UApplyOp<double, false> op, int OZ) {
  const Blk3 b = xcd_block();
  march_lds<double, 16, 16, false>(op.G, op, OZ, b.x, b.y, b.z);
}

// the same apply, also reducing p.q (p = the input, q = the output, every
// row) to one partial per workgroup: the CG's p.q without another pass
struct UApplyDotOp : UApplyOp<double, false> {
  double *part;
  __device__ double store(int i, int j, int k, const double (&acc)[3], const double (&own)[4]) const {
    const int64_t n = node_id(G, i, j, k);
    double d = 0.;
#pragma unroll
    for (int c = 0; c < 3; c++) {
      const double yv = on_dirichlet(G, bc.face[c], bc.vert[c], i, j, k) ? own[c] : acc[c];
      y[3 * n + c] = yv;
      d += own[c] * yv;
    }
    return d;
  }
  __device__ void partial(int slot, double v) const { part[slot] = v; }
};
__global__ __launch_bounds__(256, 2) void k_u_apply_lds_dot(UApplyDotOp op, int OZ) {
  const Blk3 b = xcd_block();
  march_lds<double, 16, 16, false>(op.G, op, OZ, b.x, b.y, b.z,
                                   blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z));
}

// ------------------------------------------------ stored-form FP64 apply
// The stored form (variant B): q_g = W_g (v_g^2 + eta) E_e vol for the 27
// Gauss points of every element, the integrand weights of
// VF_BilinearFormU3D_local (VFMech.c:240-252), computed once per coefficient
// state (k_u_qsetup_lds) so that the apply stages only x.
// The stored form on the node-owned march (march_lds): 16 x 16 element slots
// per tile, (15 bx - 1 + tx, 15 by - 1 + ty), halo elements stored in every
// tile that computes them: [tile][element layer][27][256], the layout of the
// V operator's weights.
template <bool DOT>
struct UApplyLdsqOp {
  static constexpr int NF = 3, NC = 3;
  const double *x;
  double *y, *part;
  const double *uq;
  Grid G;
  BcMask bc;
  double lam1, mu1, ihx, ihy, ihz;
  __device__ double load(int f, int64_t n) const { return x[3 * n + f]; }
  template <typename XS>
  __device__ void element_lds(int, int, int ek, const XS &xs, double (&r)[8][3]) const {
    const int tile = blockIdx.x + gridDim.x * blockIdx.y;
    const double *q = uq + ((int64_t)tile * G.nez + ek) * 27 * 256 + threadIdx.x;
    u_elem_apply_lds_stored<256, double>(xs, lam1, mu1, ihx, ihy, ihz, [&](int a, int c, double v) { r[a][c] = v; },
                                         q);
  }
  __device__ auto store(int i, int j, int k, const double (&acc)[3], const double (&own)[3]) const {
    const int64_t n = node_id(G, i, j, k);
    double d = 0.;
#pragma unroll
    for (int c = 0; c < 3; c++) {
      const double yv = on_dirichlet(G, bc.face[c], bc.vert[c], i, j, k) ? own[c] : acc[c];
      y[3 * n + c] = yv;
      d += own[c] * yv;
    }
    if constexpr (DOT) return d;
  }
};
struct UApplyLdsqDotOp : UApplyLdsqOp<true> {
  __device__ void partial(int slot, double v) const { part[slot] = v; }
};
__global__ __launch_bounds__(256, 2) void k_u_apply_ldsq(UApplyLdsqOp<false> op, int OZ) {
  march_lds<double, 16, 16, false>(op.G, op, OZ, blockIdx.x, blockIdx.y, blockIdx.z);
}
__global__ __launch_bounds__(256, 2) void k_u_apply_ldsq_dot(UApplyLdsqDotOp op, int OZ) {
  march_lds<double, 16, 16, false>(op.G, op, OZ, blockIdx.x, blockIdx.y, blockIdx.z,
                                   blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z));
}
// q of the node-owned tiles: block (bx, by, ek), thread = slot
__global__ __launch_bounds__(256) void k_u_qsetup_lds(Grid G, const double *__restrict__ V,
                                                      const double *__restrict__ E_elem, double E, double eta,
                                                      double *__restrict__ uq) {
  const int tid = threadIdx.x;
  const Blk3 b = xcd_block();
  const int ex = b.x * 15 - 1 + tid % 16, ey = b.y * 15 - 1 + tid / 16, ek = b.z;
  double *q = uq + ((int64_t)(b.x + gridDim.x * b.y) * G.nez + ek) * 27 * 256 + tid;
  if (ex < 0 || ex >= G.nex || ey < 0 || ey >= G.ney) {
    for (int g = 0; g < 27; g++) q[g * 256] = 0.;
    return;
  }
  double v[8];
  for (int a = 0; a < 8; a++) v[a] = V[node_id(G, ex + (a & 1), ey + ((a >> 1) & 1), ek + (a >> 2))];
  const double cE = (E_elem ? E_elem[cell_id(G, ex, ey, ek)] : E) * (G.hx * G.hy * G.hz);
  for (int gz = 0; gz < 3; gz++)
    for (int gy = 0; gy < 3; gy++)
      for (int gx = 0; gx < 3; gx++) {
        double vg = 0.;
        for (int a = 0; a < 8; a++)
          vg += ((a & 1) ? gP1(gx) : gP0(gx)) * (((a >> 1) & 1) ? gP1(gy) : gP0(gy)) * ((a >> 2) ? gP1(gz) : gP0(gz)) *
                v[a];
        q[((gz * 3 + gy) * 3 + gx) * 256] = cE * gW(gz) * gW(gx) * gW(gy) * (vg * vg + eta);
      }
}


int64_t uq_doubles(const Grid &G) { return (int64_t)((G.nx + 14) / 15) * ((G.ny + 14) / 15) * G.nez * 27 * 256; }
void launch_u_qsetup(const DevProblem &P, double *uq, hipStream_t s) {
  const dim3 g((P.G.nx + 14) / 15, (P.G.ny + 14) / 15, P.G.nez);
  hipLaunchKernelGGL(k_u_qsetup_lds, g, dim3(256), 0, s, P.G, (const double *)P.V, (const double *)P.E_elem, P.E,
                     P.eta, uq);
}

// ---------------------------------------------------------------- diagonal
// Ke[(a,c),(a,c)] = sum_g W s [(lam+mu) (d_c phi_a)^2 + mu |grad phi_a|^2]
struct UDiagOp {
  static constexpr int NF = 1, NC = 3;
  const double *v;
  double *d;
  const double *E_elem;
  const uint32_t *cmask;  // UNILATERAL_NOCOMPRESSION state (nullptr: none)
  Grid G;
  BcMask bc;
  double E, lam1, mu1, eta, vol;
  __device__ double load(int, int64_t n) const { return v[n]; }
  // diag_{a,c} = sum_g A_g (d_c N_a)^2 + M_g |grad N_a|^2 with M = W s mu and
  // A = M + the spherical weight.  (d_x N_a)^2 = P_j(y)^2 P_k(z)^2 / hx^2 does
  // not depend on the node's x-index i, so the x-term needs the sums of A and M
  // over gx on the 3 x 3 (gy, gz) lines only, contracted with P_j^2 P_k^2; the
  // y- and z-terms alike.
  __device__ void element(int ex, int ey, int ek, const double (&xin)[2][4][1], double (&r)[8][3]) const {
    double Ee = E_elem ? E_elem[cell_id(G, ex, ey, ek)] : E;
    double cE = Ee * vol;
    const uint32_t cm = cmask ? cmask[cell_id(G, ex, ey, ek)] : 0u;
    const double kap1 = (lam1 + 2. * mu1 / 3.) * (1. + eta);
    const double i2[3] = {1. / (G.hx * G.hx), 1. / (G.hy * G.hy), 1. / (G.hz * G.hz)};
    double vx[2][2][3];  // v along x on the (j, k) lines
#pragma unroll
    for (int k = 0; k < 2; k++)
#pragma unroll
      for (int j = 0; j < 2; j++)
#pragma unroll
        for (int gx = 0; gx < 3; gx++) vx[k][j][gx] = gP0(gx) * xin[k][j * 2][0] + gP1(gx) * xin[k][j * 2 + 1][0];
    // [A, M] summed over gx ([gz][gy]), over gy ([gz][gx]), over gz ([gy][gx])
    double Sx[2][3][3] = {}, Sy[2][3][3] = {}, Sz[2][3][3] = {};
#pragma unroll
    for (int gz = 0; gz < 3; gz++)
#pragma unroll
      for (int gy = 0; gy < 3; gy++)
#pragma unroll
        for (int gx = 0; gx < 3; gx++) {
          const double vv = gP0(gz) * (gP0(gy) * vx[0][0][gx] + gP1(gy) * vx[0][1][gx]) +
                            gP1(gz) * (gP0(gy) * vx[1][0][gx] + gP1(gy) * vx[1][1][gx]);
          const double w = cE * gW(gx) * gW(gy) * gW(gz), s = vv * vv + eta;
          // spherical weight: lam s, or kappa (1 + eta) - 2 mu s / 3 at a compressed point
          const double lw = ((cm >> ((gz * 3 + gy) * 3 + gx)) & 1u) ? w * (kap1 - 2. * mu1 * s / 3.) : w * s * lam1;
          const double mw = w * s * mu1, aw = lw + mw;
          Sx[0][gz][gy] += aw;
          Sx[1][gz][gy] += mw;
          Sy[0][gz][gx] += aw;
          Sy[1][gz][gx] += mw;
          Sz[0][gy][gx] += aw;
          Sz[1][gy][gx] += mw;
        }
    // T[q][b][a] = sum over the two remaining directions of S[q] P_a^2 P_b^2
    double Tx[2][2][2], Ty[2][2][2], Tz[2][2][2];
#pragma unroll
    for (int q = 0; q < 2; q++)
#pragma unroll
      for (int b = 0; b < 2; b++)
#pragma unroll
        for (int a = 0; a < 2; a++) {
          double tx = 0., ty = 0., tz = 0.;
#pragma unroll
          for (int u = 0; u < 3; u++) {
            const double pb = b ? gP1(u) : gP0(u);
#pragma unroll
            for (int t = 0; t < 3; t++) {
              const double pa = a ? gP1(t) : gP0(t), w2 = pa * pa * pb * pb;
              tx += Sx[q][u][t] * w2;
              ty += Sy[q][u][t] * w2;
              tz += Sz[q][u][t] * w2;
            }
          }
          Tx[q][b][a] = tx;  // [k][j]
          Ty[q][b][a] = ty;  // [k][i]
          Tz[q][b][a] = tz;  // [j][i]
        }
#pragma unroll
    for (int a = 0; a < 8; a++) {
      const int i = a & 1, j = (a >> 1) & 1, k = a >> 2;
      const double g2 = i2[0] * Tx[1][k][j] + i2[1] * Ty[1][k][i] + i2[2] * Tz[1][j][i];
      r[a][0] += i2[0] * Tx[0][k][j] + g2;
      r[a][1] += i2[1] * Ty[0][k][i] + g2;
      r[a][2] += i2[2] * Tz[0][j][i] + g2;
    }
  }
  __device__ void store(int i, int j, int k, const double (&acc)[3], const double (&)[1]) const {
    int64_t n = node_id(G, i, j, k);
    for (int c = 0; c < 3; c++) d[3 * n + c] = on_dirichlet(G, bc.face[c], bc.vert[c], i, j, k) ? 1. : acc[c];
  }
};

__global__ __launch_bounds__(256) void k_u_diag(UDiagOp op, int OZ) { march<double, 16, 16>(op.G, op, OZ, xcd_block()); }

// ------------------------------------------------------- load vector / residual
// b_{a,c} = sum_g W d_c phi_a [(3lam+2mu) alpha dth s + beta p v]          (thermo-poro)
//         + sum_g W p (d_c v) phi_a                       (crack pressure, optional)
// mode 0: out = b (no BC); mode 1: out = K U - b with ResidualApplyDirichletBC
// Load terms of the residual for one element: at every Gauss point the
// damaged thermal / poroelastic stress tp (times grad N_a) and the crack
// pressure term cp grad(v) (times N_a), both scaled by sgn, added into r.
// The fields are interpolated along x first, and the projection is contracted
// over gx, then gy, then added per gz.
template <bool UNI>
__device__ __forceinline__ void rhs_loads(const double (&X)[2][4][6], double coefalpha, double beta, double eta,
                                          double vol, bool crack, const double (&ih)[3], uint32_t cm, double sgn,
                                          double (&r)[8][3]) {
  double f1[3][2][2][3], dvx[2][2];  // [v, th, p][k][j][gx] interpolated along x; dv/dx on the (j, k) lines
#pragma unroll
  for (int k = 0; k < 2; k++)
#pragma unroll
    for (int j = 0; j < 2; j++) {
      const double *a0 = X[k][j * 2], *a1 = X[k][j * 2 + 1];
#pragma unroll
      for (int f = 0; f < 3; f++)
#pragma unroll
        for (int gx = 0; gx < 3; gx++) f1[f][k][j][gx] = gP0(gx) * a0[3 + f] + gP1(gx) * a1[3 + f];
      dvx[k][j] = (a1[3] - a0[3]) * ih[0];
    }
#pragma unroll
  for (int gz = 0; gz < 3; gz++) {
    double SA[3], SAi[2][3], SB[3][2][3];  // sums over gx: of tp; of tp P_i; of B_c P_i  ([..][gy])
#pragma unroll
    for (int gy = 0; gy < 3; gy++) {
      const double wjk[2][2] = {{gP0(gy) * gP0(gz), gP1(gy) * gP0(gz)}, {gP0(gy) * gP1(gz), gP1(gy) * gP1(gz)}};
      const double gvx = wjk[0][0] * dvx[0][0] + wjk[0][1] * dvx[0][1] + wjk[1][0] * dvx[1][0] + wjk[1][1] * dvx[1][1];
      SA[gy] = 0.;
#pragma unroll
      for (int i = 0; i < 2; i++) {
        SAi[i][gy] = 0.;
#pragma unroll
        for (int c = 0; c < 3; c++) SB[c][i][gy] = 0.;
      }
#pragma unroll
      for (int gx = 0; gx < 3; gx++) {
        double vv = 0., th = 0., pr = 0.;
#pragma unroll
        for (int k = 0; k < 2; k++)
#pragma unroll
          for (int j = 0; j < 2; j++) {
            vv += wjk[k][j] * f1[0][k][j][gx];
            th += wjk[k][j] * f1[1][k][j][gx];
            pr += wjk[k][j] * f1[2][k][j][gx];
          }
        const double gvy = ih[1] * (gP0(gz) * (f1[0][0][1][gx] - f1[0][0][0][gx]) +
                                    gP1(gz) * (f1[0][1][1][gx] - f1[0][1][0][gx]));
        const double gvz = ih[2] * (gP0(gy) * (f1[0][1][0][gx] - f1[0][0][0][gx]) +
                                    gP1(gy) * (f1[0][1][1][gx] - f1[0][0][1][gx]));
        const double W = vol * gW(gx) * gW(gy) * gW(gz);
        const bool comp = UNI && ((cm >> ((gz * 3 + gy) * 3 + gx)) & 1u);  // VFMech.c:505-513
        const double tp = sgn * (comp ? W * (coefalpha * th * (1. + eta) + beta * pr)
                                      : W * (coefalpha * th * (vv * vv + eta) + beta * pr * vv));
        const double cp = crack ? sgn * W * pr : 0.;
        const double B[3] = {cp * gvx, cp * gvy, cp * gvz};
        SA[gy] += tp;
        SAi[0][gy] += tp * gP0(gx);
        SAi[1][gy] += tp * gP1(gx);
#pragma unroll
        for (int c = 0; c < 3; c++) {
          SB[c][0][gy] += B[c] * gP0(gx);
          SB[c][1][gy] += B[c] * gP1(gx);
        }
      }
    }
    // corner a = (i, j, k): x-term sx P_j P_k, y-term P_i sy P_k, z-term
    // P_i P_j sz (s = -+1/h), N-term P_i P_j P_k
#pragma unroll
    for (int k = 0; k < 2; k++) {
      const double pz = k ? gP1(gz) : gP0(gz), sz = k ? ih[2] : -ih[2];
#pragma unroll
      for (int j = 0; j < 2; j++) {
        const double sy = j ? ih[1] : -ih[1];
        double TAx = 0., TAy[2] = {0., 0.}, TAz[2] = {0., 0.}, TB[3][2] = {{0., 0.}, {0., 0.}, {0., 0.}};
#pragma unroll
        for (int gy = 0; gy < 3; gy++) {
          const double py = j ? gP1(gy) : gP0(gy);
          TAx += SA[gy] * py;
#pragma unroll
          for (int i = 0; i < 2; i++) {
            TAy[i] += SAi[i][gy];
            TAz[i] += SAi[i][gy] * py;
#pragma unroll
            for (int c = 0; c < 3; c++) TB[c][i] += SB[c][i][gy] * py;
          }
        }
#pragma unroll
        for (int i = 0; i < 2; i++) {
          const double sx = i ? ih[0] : -ih[0];
          const int a = (k * 2 + j) * 2 + i;
          r[a][0] += pz * (sx * TAx + TB[0][i]);
          r[a][1] += pz * (sy * TAy[i] + TB[1][i]);
          r[a][2] += sz * TAz[i] + pz * TB[2][i];
        }
      }
    }
  }
}

template <bool UNI>
struct URhsOp {
  static constexpr int NF = 6, NC = 3;  // u0,u1,u2,v,dth,p
  const double *U, *v, *dth, *p, *BCU;
  double *out;
  const double *E_elem;
  Grid G;
  BcMask bc;
  uint8_t ftype[3][6];   // face types for the residual BC values
  uint8_t vtype[3][8];
  double E, lam1, mu1, eta, vol, alpha, beta;
  int crack, mode;
  __device__ double load(int f, int64_t n) const {
    return f < 3 ? U[3 * n + f] : f == 3 ? v[n] : f == 4 ? dth[n] : p[n];
  }
  __device__ void element(int ex, int ey, int ek, const double (&X)[2][4][6], double (&r)[8][3]) const {
    double Ee = E_elem ? E_elem[cell_id(G, ex, ey, ek)] : E;
    double lam = lam1 * Ee, mu = mu1 * Ee;
    const double ih[3] = {1. / G.hx, 1. / G.hy, 1. / G.hz};
    // UNILATERAL_NOCOMPRESSION: the branch of every Gauss point follows U itself
    // (VFMech.c:1566-1615: VF_UResidual evaluates both terms at the iterate)
    uint32_t cm = 0u;
    if constexpr (UNI) {
      double u[8][3];
      for (int a = 0; a < 8; a++)
        for (int c = 0; c < 3; c++) u[a][c] = X[a >> 2][a & 3][c];
      cm = compression_mask(u, ih[0], ih[1], ih[2]);
    }
    if (mode == 1) {
      double Xa[2][4][4];
      for (int k = 0; k < 2; k++)
        for (int q = 0; q < 4; q++)
          for (int f = 0; f < 4; f++) Xa[k][q][f] = X[k][q][f];
      u_elem_apply<double, UNI>(Xa, Ee * vol, lam1, mu1, eta, ih[0], ih[1], ih[2], r, cm);
    }
    const double coefalpha = (3. * lam + 2. * mu) * alpha;
    rhs_loads<UNI>(X, coefalpha, beta, eta, vol, crack != 0, ih, cm, mode == 1 ? -1. : 1., r);
  }
  // value written by ResidualApplyDirichletBC: traversal order X0..Z1 then
  // vertices, later writers win (VFCartFE.c:747-999)
  __device__ void store(int i, int j, int k, const double (&acc)[3], const double (&own)[6]) const {
    int64_t n = node_id(G, i, j, k);
    for (int c = 0; c < 3; c++) {
      double val = acc[c];
      if (mode == 1 && on_dirichlet(G, bc.face[c], bc.vert[c], i, j, k)) {
        int t = 0;
        const bool onf[6] = {i == 0, i == G.nx - 1, j == 0, j == G.ny - 1, k == 0, k == G.nz - 1};
        for (int f = 0; f < 6; f++)
          if (ftype[c][f] && onf[f]) t = ftype[c][f];
        int ix = (i == 0) ? 0 : (i == G.nx - 1) ? 1 : -1;
        int iy = (j == 0) ? 0 : (j == G.ny - 1) ? 1 : -1;
        int iz = (k == 0) ? 0 : (k == G.nz - 1) ? 1 : -1;
        if (ix >= 0 && iy >= 0 && iz >= 0 && vtype[c][ix + 2 * iy + 4 * iz]) t = vtype[c][ix + 2 * iy + 4 * iz];
        double u = own[c];
        val = t == 3 ? u - BCU[3 * n + c] : t == 2 ? u - 1. : u;
      }
      out[3 * n + c] = val;
    }
  }
};

// Load vector of one element under UNILATERAL_NONE, sum-factorised over two
// Gauss-point walks (values of v, dth, p; then v with its gradient and p,
// kept apart so that each walk's accumulators fit the registers of two waves
// per SIMD):
//   thermo-poro (VFMech.c:386-447): b[a][c] += sum_g W d_c phi_a T,
//     T = (3 lam + 2 mu) alpha dth (v^2 + eta) + beta p v  -- the x-, y- and
//     z-derivative projections of one scalar: T summed over gx per (gy, gz),
//     over gy per (gx, gz), over gz per (gx, gy);
//   crack pressure (VFMech.c:533-589): b[a][c] += sum_g W p (d_c v) phi_a --
//     a value projection, summed over gx, then gy, then gz.
// Every contribution is multiplied by sgn (-1 for the residual).
template <typename CV>
__device__ __forceinline__ void u_loads_elem(const CV &cv, double coefalpha, double beta, double eta,
                                             double vol, bool crack, double ihx, double ihy, double ihz, double sgn,
                                             double (&r)[8][3]) {
  // thermo-poro: values of v, dth, p (fields 0-2)
  {
    double Tx = 0., Ty[3] = {0., 0., 0.}, Tz[3][3], Rx[2][2] = {{0., 0.}, {0., 0.}}, Ry[2][2] = {{0., 0.}, {0., 0.}};
#pragma unroll
    for (int q = 0; q < 3; q++)
#pragma unroll
      for (int w = 0; w < 3; w++) Tz[q][w] = 0.;
    gauss_walk_cv<0, 3>(cv, ihx, ihy, ihz, [&](int gx, int gy, int gz, const double (&val)[3], const double (&)[1][3]) {
      const double W = vol * gW(gx) * gW(gy) * gW(gz);
      const double v = val[0];
      const double T = sgn * (W * (coefalpha * val[1] * (v * v + eta) + beta * val[2] * v));
      Tx += T;
      Ty[gx] += T;
      Tz[gy][gx] += T;
      if (gx == 2) {  // end of a row (gy, gz)
#pragma unroll
        for (int k = 0; k < 2; k++)
#pragma unroll
          for (int j = 0; j < 2; j++) Rx[k][j] += (k ? gP1(gz) : gP0(gz)) * (j ? gP1(gy) : gP0(gy)) * Tx;
        Tx = 0.;
      }
      if (gx == 2 && gy == 2) {  // end of a layer gz
#pragma unroll
        for (int q = 0; q < 3; q++) {
#pragma unroll
          for (int k = 0; k < 2; k++)
#pragma unroll
            for (int i = 0; i < 2; i++) Ry[k][i] += (k ? gP1(gz) : gP0(gz)) * (i ? gP1(q) : gP0(q)) * Ty[q];
          Ty[q] = 0.;
        }
      }
    });
#pragma unroll
    for (int a = 0; a < 8; a++) {
      const int i = a & 1, j = (a >> 1) & 1, k = a >> 2;
      double rz = 0.;
#pragma unroll
      for (int gy = 0; gy < 3; gy++)
#pragma unroll
        for (int gx = 0; gx < 3; gx++) rz += (j ? gP1(gy) : gP0(gy)) * (i ? gP1(gx) : gP0(gx)) * Tz[gy][gx];
      r[a][0] += (i ? ihx : -ihx) * Rx[k][j];
      r[a][1] += (j ? ihy : -ihy) * Ry[k][i];
      r[a][2] += (k ? ihz : -ihz) * rz;
    }
  }
  if (!crack) return;
  // crack pressure: v with its gradient (field 0) and p (field 2); the value
  // projection summed over gx (Vx), gy (Vy), then gz straight into r
  double Vx[3][2], Vy[3][2][2];
#pragma unroll
  for (int c = 0; c < 3; c++)
#pragma unroll
    for (int i = 0; i < 2; i++) {
      Vx[c][i] = 0.;
#pragma unroll
      for (int j = 0; j < 2; j++) Vy[c][j][i] = 0.;
    }
  gauss_walk_cv<1, 1>([&](int q, int a) { return cv(q ? 2 : 0, a); }, ihx, ihy, ihz,
                      [&](int gx, int gy, int gz, const double (&val)[2], const double (&gv)[1][3]) {
    const double cp = sgn * (vol * gW(gx) * gW(gy) * gW(gz)) * val[1];
#pragma unroll
    for (int c = 0; c < 3; c++) {
      const double Gc = cp * gv[0][c];
      Vx[c][0] += gP0(gx) * Gc;
      Vx[c][1] += gP1(gx) * Gc;
    }
    if (gx == 2) {  // end of a row (gy, gz)
#pragma unroll
      for (int c = 0; c < 3; c++)
#pragma unroll
        for (int i = 0; i < 2; i++) {
          Vy[c][0][i] += gP0(gy) * Vx[c][i];
          Vy[c][1][i] += gP1(gy) * Vx[c][i];
          Vx[c][i] = 0.;
        }
    }
    if (gx == 2 && gy == 2) {  // end of a layer gz: into the corners (j, i) of both node planes
#pragma unroll
      for (int c = 0; c < 3; c++)
#pragma unroll
        for (int j = 0; j < 2; j++)
#pragma unroll
          for (int i = 0; i < 2; i++) {
            r[(0 * 2 + j) * 2 + i][c] += gP0(gz) * Vy[c][j][i];
            r[(1 * 2 + j) * 2 + i][c] += gP1(gz) * Vy[c][j][i];
            Vy[c][j][i] = 0.;
          }
    }
  });
}

// The load vector alone (UNILATERAL_NONE: the loads do not depend on U):
// mode 0 stores b; mode 1 finishes the residual F = K U - b over an output
// that already holds K U (the FP64 apply), and writes the Dirichlet rows'
// ResidualApplyDirichletBC values from U (read there only); mode 2 stores b
// with the Dirichlet rows' targets instead, so that F(U) = A U - b for the
// MatShell operator A.
struct ULoadsOp {
  static constexpr int NF = 3, NC = 3;  // v, dth, p
  const double *U, *v, *dth, *p, *BCU;
  double *out;
  const double *E_elem;
  Grid G;
  BcMask bc;
  uint8_t ftype[3][6];
  uint8_t vtype[3][8];
  double E, lam1, mu1, eta, vol, alpha, beta;
  int crack, mode;
  __device__ double load(int f, int64_t n) const { return f == 0 ? v[n] : f == 1 ? dth[n] : p[n]; }
  __device__ double aux(int ex, int ey, int ek) const { return E_elem ? E_elem[cell_id(G, ex, ey, ek)] : E; }
  __device__ void element(int, int, int, const double (&X)[2][4][3], double Ee, double (&r)[8][3]) const {
    double fld[3][8];
#pragma unroll
    for (int a = 0; a < 8; a++)
#pragma unroll
      for (int f = 0; f < 3; f++) fld[f][a] = X[a >> 2][a & 3][f];
    const double coefalpha = (3. * lam1 + 2. * mu1) * Ee * alpha;
    u_loads_elem([&](int q, int a) { return fld[q][a]; }, coefalpha, beta, eta, vol, crack != 0, 1. / G.hx,
                 1. / G.hy, 1. / G.hz, mode == 1 ? -1. : 1., r);
  }
  template <typename XS>
  __device__ void element_lds(int, int, int, const XS &xs, double Ee, double (&r)[8][3]) const {
    const double coefalpha = (3. * lam1 + 2. * mu1) * Ee * alpha;
    u_loads_elem([&](int q, int a) { return xs(a >> 2, a & 3, q); }, coefalpha, beta, eta, vol, crack != 0, 1. / G.hx, 1. / G.hy, 1. / G.hz, mode == 1 ? -1. : 1., r);
  }
  __device__ void store(int i, int j, int k, const double (&acc)[3], const double (&)[3]) const {
    const int64_t n = node_id(G, i, j, k);
    for (int c = 0; c < 3; c++) {
      double val = acc[c];
      if (mode == 2 && on_dirichlet(G, bc.face[c], bc.vert[c], i, j, k)) {
        // the value the residual's Dirichlet row subtracts from U
        val = dirichlet_value(i, j, k, c, n);
      } else if (mode == 1) {
        if (on_dirichlet(G, bc.face[c], bc.vert[c], i, j, k)) val = U[3 * n + c] - dirichlet_value(i, j, k, c, n);
        else val = out[3 * n + c] + acc[c];  // K U - b
      }
      out[3 * n + c] = val;
    }
  }
  // ResidualApplyDirichletBC's target on a Dirichlet row: BCU (type 3), 1 (type 2), else 0
  __device__ double dirichlet_value(int i, int j, int k, int c, int64_t n) const {
    int t = 0;
    const bool onf[6] = {i == 0, i == G.nx - 1, j == 0, j == G.ny - 1, k == 0, k == G.nz - 1};
    for (int f = 0; f < 6; f++)
      if (ftype[c][f] && onf[f]) t = ftype[c][f];
    const int ix = (i == 0) ? 0 : (i == G.nx - 1) ? 1 : -1;
    const int iy = (j == 0) ? 0 : (j == G.ny - 1) ? 1 : -1;
    const int iz = (k == 0) ? 0 : (k == G.nz - 1) ? 1 : -1;
    if (ix >= 0 && iy >= 0 && iz >= 0 && vtype[c][ix + 2 * iy + 4 * iz]) t = vtype[c][ix + 2 * iy + 4 * iz];
    return t == 3 ? BCU[3 * n + c] : t == 2 ? 1. : 0.;
  }
};
__global__ __launch_bounds__(256, 2) void k_u_loads(ULoadsOp op, int OZ) {
  march_lds<double, 16, 16, false>(op.G, op, OZ, xcd_block());
}

template <bool UNI>
__global__ __launch_bounds__(256) void k_u_rhs(URhsOp<UNI> op, int OZ) {
  march<double, 16, 16>(op.G, op, OZ, xcd_block());
}

// ------------------------------------------------------------ host launchers
static dim3 march_grid(const Grid &G, int TX, int TY, int OZ) {
  return dim3((G.nx + TX - 2) / (TX - 1), (G.ny + TY - 2) / (TY - 1), (G.nz + OZ - 1) / OZ);
}

// z-chunk length of the marching tiles: every chunk recomputes one element
// layer below it, so long chunks waste less; the grid should still fill every
// workgroup slot of the device.  Minimise rounds * (layers per workgroup),
// rounds = ceil(workgroups / resident slots).
int choose_oz(const Grid &G, int tiles_xy, int slots) {
  if (slots < 1) slots = 256;
  int best = G.nz;
  double best_cost = 1e300;
  for (int oz = 2; oz <= G.nz; oz++) {
    int chunks = (G.nz + oz - 1) / oz;
    if (oz > 2 && (G.nz + oz - 2) / (oz - 1) == chunks) continue;  // same chunk count, shorter is free
    int64_t wg = (int64_t)tiles_xy * chunks;
    double rounds = (double)((wg + slots - 1) / slots);
    double cost = rounds * (oz + 1);
    if (cost < best_cost - 1e-9) { best_cost = cost; best = oz; }
  }
  return best;
}

// resident workgroups of a kernel on the whole device (occupancy x CUs)
template <typename K>
static int device_slots(K kernel, int threads) {
  static int cus = 0;
  if (!cus) {
    int dev = 0;
    hipGetDevice(&dev);
    hipDeviceProp_t pr;
    cus = hipGetDeviceProperties(&pr, dev) == hipSuccess ? pr.multiProcessorCount : 256;
  }
  int per = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kernel, threads, 0) != hipSuccess || per < 1) per = 1;
  return per * cus;
}

// FP32 grids of 16-64 nodes in x (coarse multigrid levels) march 8 x 8 tiles
// of 7 x 7 owned nodes: more, shorter workgroups, and no row of two 15-node
// tiles whose second owns 1-2 nodes (16-17 nodes in x), which runs the
// 16 x 16 apply at 34-37 us against 14 us (five times the memory waits)
constexpr int SMALL_NX_LO = 16, SMALL_NX_HI = 64;
template <typename T, bool UNI>
static void launch_u_apply_u(const DevProblem &P, const T *x, T *y, const T *v, hipStream_t s) {
  UApplyOp<T, UNI> op;
  op.x = x; op.y = y; op.v = v;
  op.E_elem = P.E_elem;
  op.cmask = P.cmask;
  op.G = P.G; op.bc = P.bcU;
  op.E = T(P.E); op.lam1 = T(P.lam1); op.mu1 = T(P.mu1); op.eta = T(P.eta);
  op.vol = T(P.G.hx * P.G.hy * P.G.hz);
  op.ihx = T(1. / P.G.hx); op.ihy = T(1. / P.G.hy); op.ihz = T(1. / P.G.hz);
  int tiles = ((P.G.nx + 14) / 15) * ((P.G.ny + 14) / 15);
  if constexpr (sizeof(T) == 8 && !UNI) {
    if (P.uq) {
      UApplyLdsqOp<false> q;
      q.x = (const double *)x; q.y = (double *)y; q.part = nullptr; q.uq = P.uq;
      q.G = P.G; q.bc = P.bcU; q.lam1 = P.lam1; q.mu1 = P.mu1;
      q.ihx = 1. / P.G.hx; q.ihy = 1. / P.G.hy; q.ihz = 1. / P.G.hz;
      static int lqslots = device_slots(k_u_apply_ldsq, 256);
      const int OZ = choose_oz(P.G, tiles, lqslots);
      hipLaunchKernelGGL(k_u_apply_ldsq, march_grid(P.G, 16, 16, OZ), dim3(256), 0, s, q, OZ);
      return;
    }
    static int slots = device_slots(k_u_apply_lds, 256);
    int OZ = choose_oz(P.G, tiles, slots);
    hipLaunchKernelGGL(k_u_apply_lds, march_grid(P.G, 16, 16, OZ), dim3(256), 0, s, op, OZ);
  } else {
    if constexpr (sizeof(T) == 4 && !UNI) {
      if (P.G.nx >= SMALL_NX_LO && P.G.nx <= SMALL_NX_HI) {
        static int slots = device_slots(k_u_apply<T, 8, 8, false>, 64);
        int OZ = choose_oz(P.G, ((P.G.nx + 6) / 7) * ((P.G.ny + 6) / 7), slots);
        hipLaunchKernelGGL((k_u_apply<T, 8, 8, false>), march_grid(P.G, 8, 8, OZ), dim3(64), 0, s, op, OZ);
        return;
      }
    }
    static int slots = device_slots(k_u_apply<T, 16, 16, UNI>, 256);
    int OZ = choose_oz(P.G, tiles, slots);
    hipLaunchKernelGGL((k_u_apply<T, 16, 16, UNI>), march_grid(P.G, 16, 16, OZ), dim3(256), 0, s, op, OZ);
  }
}

template <typename T>
static void launch_u_apply_t(const DevProblem &P, const T *x, T *y, const T *v, hipStream_t s) {
  if (P.unilateral) launch_u_apply_u<T, true>(P, x, y, v, s);
  else launch_u_apply_u<T, false>(P, x, y, v, s);
}

// per element: the Gauss points where tr(eps(U)) < 0
__global__ void k_u_cmask(Grid G, const double *__restrict__ U, uint32_t *__restrict__ cmask) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (int64_t)G.nex * G.ney * G.nez) return;
  const int ei = (int)(e % G.nex), ej = (int)((e / G.nex) % G.ney), ek = (int)(e / ((int64_t)G.nex * G.ney));
  double u[8][3];
  for (int a = 0; a < 8; a++) {
    const int64_t n = node_id(G, ei + (a & 1), ej + ((a >> 1) & 1), ek + (a >> 2));
    for (int c = 0; c < 3; c++) u[a][c] = U[3 * n + c];
  }
  cmask[e] = compression_mask(u, 1. / G.hx, 1. / G.hy, 1. / G.hz);
}

void launch_u_cmask(const DevProblem &P, const double *U, hipStream_t s) {
  hipLaunchKernelGGL(k_u_cmask, dim3((unsigned)((P.ne + 255) / 256)), dim3(256), 0, s, P.G, U, P.cmask);
}

// Chebyshev-Jacobi solve of a small FP32 level from x = 0 in one workgroup
// (the coarsest multigrid level, at most MGC_NE elements): x, d and the
// operator output stay in LDS.  An apply is an element pass (one element per
// thread, its 24 outputs to LDS) and a node pass that adds each node's eight
// element contributions in the order of the marching apply (x-pairs, then
// y-pairs, the lower element layer before the upper), so it gives the values
// of launch_u_apply_f32; the step is mg_cheb_entry, as in the multi-launch
// smoother.
// one apply record of the batched / engine applies
template <typename T>
struct BApply {
  UApplyOp<T> op;
  int OZ;
};

constexpr int MGC_T = 512, MGC_NE = 512, MGC_NN = 729;

__device__ __forceinline__ void mg_coarse_cheb_body(const UApplyOp<float, false> &op, const float *__restrict__ b,
                                                    const float *__restrict__ dinv, float *__restrict__ xout,
                                                    const MgChebCoefs &cc) {
  __shared__ float sx[3 * MGC_NN], sd[3 * MGC_NN], sy[3 * MGC_NN];
  __shared__ float sr[MGC_NE * 24];
  const Grid &G = op.G;
  const int tid = threadIdx.x;
  const int nn = G.nx * G.ny * G.nz, ne = G.nex * G.ney * G.nez;
  auto R = [&](int ex, int ey, int ez, int corner, int c) -> float {
    if (ex < 0 || ex >= G.nex || ey < 0 || ey >= G.ney || ez < 0 || ez >= G.nez) return 0.f;
    return sr[(ex + G.nex * (ey + G.ney * ez)) * 24 + corner * 3 + c];
  };
  // corner (i, J, K) of element (ex, ey, ez) plus corner (0, J, K) of its +x neighbour
  auto cx = [&](int ex, int ey, int ez, int J, int K, int c) {
    return R(ex, ey, ez, (K * 2 + J) * 2 + 1, c) + R(ex + 1, ey, ez, (K * 2 + J) * 2, c);
  };
  for (int i = 0; i < cc.k; i++) {
    if (i > 0) {
      for (int e = tid; e < ne; e += MGC_T) {
        const int ex = e % G.nex, ey = (e / G.nex) % G.ney, ez = e / (G.nex * G.ney);
        float xin[2][4][4];
#pragma unroll
        for (int kz = 0; kz < 2; kz++)
#pragma unroll
          for (int q = 0; q < 4; q++) {
            const int64_t n = node_id(G, ex + (q & 1), ey + (q >> 1), ez + kz);
#pragma unroll
            for (int f = 0; f < 3; f++) xin[kz][q][f] = sx[3 * n + f];
            xin[kz][q][3] = op.v[n];
          }
        float r[8][3];
        op.element(ex, ey, ez, xin, op.aux(ex, ey, ez), r);
#pragma unroll
        for (int a = 0; a < 8; a++)
#pragma unroll
          for (int c = 0; c < 3; c++) sr[e * 24 + a * 3 + c] = r[a][c];
      }
      __syncthreads();
      for (int n = tid; n < nn; n += MGC_T) {
        const int ni = n % G.nx, nj = (n / G.nx) % G.ny, nk = n / (G.nx * G.ny);
        const int ex = ni - 1, ey = nj - 1;
#pragma unroll
        for (int c = 0; c < 3; c++) {
          // element layer nk - 1 (corners K = 1), then layer nk (K = 0); in
          // each, element row ey (corners J = 1) before row ey + 1 (J = 0)
          const float carry = cx(ex, ey, nk - 1, 1, 1, c) + cx(ex, ey + 1, nk - 1, 0, 1, c);
          const float acc = carry + (cx(ex, ey, nk, 1, 0, c) + cx(ex, ey + 1, nk, 0, 0, c));
          sy[3 * n + c] = on_dirichlet(G, op.bc.face[c], op.bc.vert[c], ni, nj, nk) ? sx[3 * n + c] : acc;
        }
      }
      __syncthreads();
    }
    const float c1 = cc.c1[i], c2 = cc.c2[i];
    for (int t = tid; t < 3 * nn; t += MGC_T)
      mg_cheb_entry(b, i > 0 ? sy : nullptr, dinv, sd, sx, c1, c2, i == 0, i == 0, t);
    __syncthreads();
  }
  for (int t = tid; t < 3 * nn; t += MGC_T) xout[t] = sx[t];
}

__global__ __launch_bounds__(MGC_T) void k_mg_coarse_cheb(UApplyOp<float, false> op, const float *__restrict__ b,
                                                          const float *__restrict__ dinv, float *__restrict__ xout,
                                                          MgChebCoefs cc) {
  mg_coarse_cheb_body(op, b, dinv, xout, cc);
}

// the batched form: one workgroup per listed member (its first chunk), the
// member's level operator, vectors and coarsest-level coefficients from the
// engine's tables
__global__ __launch_bounds__(MGC_T) void k_bmg_coarse_cheb(const BApply<float> *__restrict__ ops,
                                                           const BmgVec *__restrict__ lvt, int level,
                                                           const BmgChunk *__restrict__ ch, const int *__restrict__ dn,
                                                           const BmgCoef *__restrict__ coef,
                                                           const double *__restrict__ flags, int fstride) {
  if ((int)blockIdx.x >= *dn) return;
  const BmgChunk c = ch[blockIdx.x];
  if (c.start != 0 || flags[(int64_t)c.p * fstride] == 0.) return;  // uniform in the workgroup
  const BmgVec &v = lvt[c.p * BML + level];
  MgChebCoefs cc;
  cc.k = MG_COARSE_DEG;
#pragma unroll
  for (int i = 0; i < MG_COARSE_DEG; i++) {
    const BmgCoef k = coef[((c.p * BML + level) * 2 + 1) * BMK + i];
    cc.c1[i] = k.c1;
    cc.c2[i] = k.c2;
  }
  mg_coarse_cheb_body(ops[c.p].op, v.b, v.dinv, v.x, cc);
}

bool mg_coarse_fits(const DevProblem &P) {
  return !P.unilateral && P.ne <= MGC_NE && P.nn <= MGC_NN;
}

void launch_mg_coarse_cheb(const DevProblem &P, const float *b, const float *dinv, float *x, const MgChebCoefs &cc,
                           hipStream_t s) {
  UApplyOp<float, false> op;
  op.x = nullptr; op.y = nullptr; op.v = P.V32;
  op.E_elem = P.E_elem;
  op.cmask = nullptr;
  op.G = P.G; op.bc = P.bcU;
  op.E = float(P.E); op.lam1 = float(P.lam1); op.mu1 = float(P.mu1); op.eta = float(P.eta);
  op.vol = float(P.G.hx * P.G.hy * P.G.hz);
  op.ihx = float(1. / P.G.hx); op.ihy = float(1. / P.G.hy); op.ihz = float(1. / P.G.hz);
  hipLaunchKernelGGL(k_mg_coarse_cheb, dim3(1), dim3(MGC_T), 0, s, op, b, dinv, x, cc);
}

void launch_bmg_coarse_cheb(const void *ops, const BmgVec *lvt, int level, const BmgChunk *ch, const int *dn,
                            int grid, const BmgCoef *coef, const double *flags, int fstride, hipStream_t s) {
  if (grid <= 0) return;
  hipLaunchKernelGGL(k_bmg_coarse_cheb, dim3(grid), dim3(MGC_T), 0, s, (const BApply<float> *)ops, lvt, level, ch,
                     dn, coef, flags, fstride);
}

void launch_u_apply(const DevProblem &P, const double *x, double *y, hipStream_t s) {
  launch_u_apply_t<double>(P, x, y, P.V, s);
}

int launch_u_apply_dot(const DevProblem &P, const double *x, double *y, double *part, hipStream_t s) {
  if (P.unilateral) return 0;
  UApplyDotOp op;
  op.x = x; op.y = y; op.v = P.V;
  op.E_elem = P.E_elem;
  op.cmask = nullptr;
  op.G = P.G; op.bc = P.bcU;
  op.E = P.E; op.lam1 = P.lam1; op.mu1 = P.mu1; op.eta = P.eta;
  op.vol = P.G.hx * P.G.hy * P.G.hz;
  op.ihx = 1. / P.G.hx; op.ihy = 1. / P.G.hy; op.ihz = 1. / P.G.hz;
  op.part = part;
  if (P.uq) {
    UApplyLdsqDotOp q;
    q.x = x; q.y = y; q.part = part; q.uq = P.uq;
    q.G = P.G; q.bc = P.bcU; q.lam1 = P.lam1; q.mu1 = P.mu1;
    q.ihx = 1. / P.G.hx; q.ihy = 1. / P.G.hy; q.ihz = 1. / P.G.hz;
    static int lqslots = device_slots(k_u_apply_ldsq_dot, 256);
    const int tiles = ((P.G.nx + 14) / 15) * ((P.G.ny + 14) / 15);
    const int OZ = choose_oz(P.G, tiles, lqslots);
    const dim3 g = march_grid(P.G, 16, 16, OZ);
    hipLaunchKernelGGL(k_u_apply_ldsq_dot, g, dim3(256), 0, s, q, OZ);
    return (int)(g.x * g.y * g.z);
  }
  static int slots = device_slots(k_u_apply_lds_dot, 256);
  const int tiles = ((P.G.nx + 14) / 15) * ((P.G.ny + 14) / 15);
  const int OZ = choose_oz(P.G, tiles, slots);
  const dim3 g = march_grid(P.G, 16, 16, OZ);
  hipLaunchKernelGGL(k_u_apply_lds_dot, g, dim3(256), 0, s, op, OZ);
  return (int)(g.x * g.y * g.z);
}
void launch_u_apply_f32(const DevProblem &P, const float *x, float *y, hipStream_t s) {
  launch_u_apply_t<float>(P, x, y, P.V32, s);
}

template <class Op>
static void mg_op_fill(Op &op, const DevProblem &P, const float *b, const float *dinv, float c2) {
  op.y = nullptr; op.v = P.V32;
  op.E_elem = P.E_elem;
  op.cmask = nullptr;
  op.G = P.G; op.bc = P.bcU;
  op.E = float(P.E); op.lam1 = float(P.lam1); op.mu1 = float(P.mu1); op.eta = float(P.eta);
  op.vol = float(P.G.hx * P.G.hy * P.G.hz);
  op.ihx = float(1. / P.G.hx); op.ihy = float(1. / P.G.hy); op.ihz = float(1. / P.G.hz);
  op.b = b; op.dinv = dinv; op.c2 = c2;
}
// launches kernel 8x8 or 16x16 by the grid's x extent; returns the workgroup
// count, or 0 without launching when it would exceed max_wg
template <class Op, class K8, class K16>
static int mg_launch(const DevProblem &P, const Op &op, K8 k8, K16 k16, int &slots8, int &slots16, hipStream_t s,
                     int64_t max_wg = INT32_MAX) {
  const bool small = P.G.nx >= SMALL_NX_LO && P.G.nx <= SMALL_NX_HI;
  if (small && !slots8) slots8 = device_slots(k8, 64);
  if (!small && !slots16) slots16 = device_slots(k16, 256);
  const int T = small ? 8 : 16;
  const int OZ = choose_oz(P.G, ((P.G.nx + T - 2) / (T - 1)) * ((P.G.ny + T - 2) / (T - 1)), small ? slots8 : slots16);
  const dim3 g = march_grid(P.G, T, T, OZ);
  const int64_t nwg = (int64_t)g.x * g.y * g.z;
  if (nwg > max_wg) return 0;
  if (small) hipLaunchKernelGGL(k8, g, dim3(64), 0, s, op, OZ);
  else hipLaunchKernelGGL(k16, g, dim3(256), 0, s, op, OZ);
  return (int)nwg;
}
void launch_mg_presmooth(const DevProblem &P, const float *b, const float *dinv, float c2, float *x, float *r,
                         hipStream_t s) {
  UMgPreOp op;
  mg_op_fill(op, P, b, dinv, c2);
  op.x = nullptr; op.xo = x; op.ro = r;
  static int s8 = 0, s16 = 0;
  mg_launch(P, op, k_u_mg_pre<8, 8>, k_u_mg_pre<16, 16>, s8, s16, s);
}
int launch_mg_postsmooth(const DevProblem &P, const float *b, const float *dinv, float c2, const float *x, float *xo,
                         const double *r64, double *part, hipStream_t s) {
  if (r64) {
    UMgPostRzOp op;
    mg_op_fill(op, P, b, dinv, c2);
    op.x = x; op.xo = xo; op.r64 = r64; op.part = part;
    static int s8 = 0, s16 = 0;
    const int n = mg_launch(P, op, k_u_mg_post_rz<8, 8>, k_u_mg_post_rz<16, 16>, s8, s16, s, 2 * 65536);
    if (n) return n;  // else too many workgroups for the partials: the plain form
  }
  UMgPostOp<false> op;
  mg_op_fill(op, P, b, dinv, c2);
  op.x = x; op.xo = xo; op.r64 = nullptr; op.part = nullptr;
  static int s8 = 0, s16 = 0;
  return mg_launch(P, op, k_u_mg_post<8, 8>, k_u_mg_post<16, 16>, s8, s16, s);
}

void launch_u_diag(const DevProblem &P, double *d, hipStream_t s) {
  UDiagOp op;
  op.v = P.V; op.d = d; op.E_elem = P.E_elem; op.G = P.G; op.bc = P.bcU;
  op.cmask = P.unilateral ? P.cmask : nullptr;
  op.E = P.E; op.lam1 = P.lam1; op.mu1 = P.mu1; op.eta = P.eta;
  op.vol = P.G.hx * P.G.hy * P.G.hz;
  int tiles = ((P.G.nx + 14) / 15) * ((P.G.ny + 14) / 15);
  static int slots = device_slots(k_u_diag, 256);
  int OZ = choose_oz(P.G, tiles, slots);
  hipLaunchKernelGGL(k_u_diag, march_grid(P.G, 16, 16, OZ), dim3(256), 0, s, op, OZ);
}

template <bool UNI>
static void launch_u_rhs_t(const DevProblem &P, const double *U, const double *BCU, double *out, int mode,
                           hipStream_t s) {
  URhsOp<UNI> op;
  op.U = U ? U : P.zero3; op.v = P.V; op.dth = P.dth; op.p = P.p; op.BCU = BCU;
  op.out = out; op.E_elem = P.E_elem; op.G = P.G; op.bc = P.bcU;
  for (int c = 0; c < 3; c++) {
    for (int f = 0; f < 6; f++) op.ftype[c][f] = (uint8_t)P.hbcU[c].face[f];
    for (int v = 0; v < 8; v++) op.vtype[c][v] = (uint8_t)P.hbcU[c].vertex[v];
  }
  op.E = P.E; op.lam1 = P.lam1; op.mu1 = P.mu1; op.eta = P.eta;
  op.vol = P.G.hx * P.G.hy * P.G.hz; op.alpha = P.alpha; op.beta = P.beta;
  op.crack = P.hasCrackPressure; op.mode = mode;
  int tiles = ((P.G.nx + 14) / 15) * ((P.G.ny + 14) / 15);
  static int slots = device_slots(k_u_rhs<UNI>, 256);
  int OZ = choose_oz(P.G, tiles, slots);
  hipLaunchKernelGGL(k_u_rhs<UNI>, march_grid(P.G, 16, 16, OZ), dim3(256), 0, s, op, OZ);
}

void launch_u_rhs(const DevProblem &P, const double *U, const double *BCU, double *out, int mode,
                  hipStream_t s) {
  if (P.unilateral) {  // (mode 2 is not used with NOCOMPRESSION: its load depends on U)
    launch_u_rhs_t<true>(P, U, BCU, out, mode, s);
    return;
  }
  // small grids (ensemble members, served requests): the apply and the loads
  // in one launch, which is latency-bound there
  if (mode != 2 && P.nn < (1 << 19)) {
    launch_u_rhs_t<false>(P, U, BCU, out, mode, s);
    return;
  }
  // UNILATERAL_NONE: K U by the FP64 apply, then the loads
  if (mode == 1) launch_u_apply(P, U ? U : P.zero3, out, s);
  ULoadsOp op;
  op.U = U ? U : P.zero3; op.v = P.V; op.dth = P.dth; op.p = P.p; op.BCU = BCU;
  op.out = out; op.E_elem = P.E_elem; op.G = P.G; op.bc = P.bcU;
  for (int c = 0; c < 3; c++) {
    for (int f = 0; f < 6; f++) op.ftype[c][f] = (uint8_t)P.hbcU[c].face[f];
    for (int v = 0; v < 8; v++) op.vtype[c][v] = (uint8_t)P.hbcU[c].vertex[v];
  }
  op.E = P.E; op.lam1 = P.lam1; op.mu1 = P.mu1; op.eta = P.eta;
  op.vol = P.G.hx * P.G.hy * P.G.hz; op.alpha = P.alpha; op.beta = P.beta;
  op.crack = P.hasCrackPressure; op.mode = mode;
  const int tiles = ((P.G.nx + 14) / 15) * ((P.G.ny + 14) / 15);
  static int slots = device_slots(k_u_loads, 256);
  const int OZ = choose_oz(P.G, tiles, slots);
  hipLaunchKernelGGL(k_u_loads, march_grid(P.G, 16, 16, OZ), dim3(256), 0, s, op, OZ);
}

}  // namespace vfg

// ------------------------------------------------- persistent multi-problem apply
namespace vfg {


// Persistent walk over a work list.  dcount (optional) holds the item count on
// the device, so a captured graph follows a list that changes between replays;
// flags (optional) skip items of members whose flags[p * fstride] is 0.
// Schedules (the autotuner's choice): TX x TY element tiles of (TX-1) x (TY-1)
// owned node columns, on the register march or the LDS-ring march.
template <typename T, int TX, int TY, bool LDS>
__global__ __launch_bounds__(TX *TY, (sizeof(T) == 4 || LDS) ? 2 : 1) void k_batch_u_apply(
    const BApply<T> *__restrict__ probs, const BItem *__restrict__ items, int nitems, const int *__restrict__ dcount,
    const double *__restrict__ flags, int fstride) {
  const int n = dcount ? *dcount : nitems;
  for (int it = blockIdx.x; it < n; it += gridDim.x) {
    const BItem w = items[it];
    if (flags && flags[(int64_t)w.p * fstride] == 0.) continue;  // uniform per block
    const UApplyOp<T> op = probs[w.p].op;
    const int OZ = probs[w.p].OZ;
    __syncthreads();  // LDS of the previous item is free
    if constexpr (LDS) march_lds<T, TX, TY, false>(op.G, op, OZ, w.bx, w.by, w.bz);
    else march<T, TX, TY>(op.G, op, OZ, w.bx, w.by, w.bz);
  }
}

// schedule -> tile width and threads: 0: 16 x 16 register march, 1: 8 x 8
// register march, 2: 16 x 16 LDS-ring march
int batch_tile(int sched) { return sched == 1 ? 8 : 16; }
template <typename T>
static int batch_slots_t(int sched) {
  if (sched == 1) return device_slots(k_batch_u_apply<T, 8, 8, false>, 64);
  if (sched == 2) return device_slots(k_batch_u_apply<T, 16, 16, true>, 256);
  return device_slots(k_batch_u_apply<T, 16, 16, false>, 256);
}
int batch_slots_s(int fp32, int sched) {
  static int cache[2][BATCH_SCHEDS] = {{0}};
  int &c = cache[fp32 ? 1 : 0][sched];
  if (!c) c = fp32 ? batch_slots_t<float>(sched) : batch_slots_t<double>(sched);
  return c;
}

size_t batch_op_bytes(int fp32) { return fp32 ? sizeof(BApply<float>) : sizeof(BApply<double>); }

int choose_oz_pub(const Grid &G, int tiles_xy, int slots) { return choose_oz(G, tiles_xy, slots); }

int batch_slots() { return batch_slots_s(0, 0); }

template <typename T>
static void fill_ops_t(const std::vector<vfg_handle *> &m, const void *const *x, void *const *y,
                       const std::vector<int> &oz, void *d_ops, hipStream_t s) {
  std::vector<BApply<T>> ops(m.size());
  for (size_t p = 0; p < m.size(); p++) {
    const DevProblem &P = m[p]->P;
    UApplyOp<T> &op = ops[p].op;
    op.x = (const T *)x[p];
    op.y = (T *)y[p];
    op.v = sizeof(T) == 4 ? (const T *)(const void *)P.V32 : (const T *)(const void *)P.V;
    op.E_elem = P.E_elem;
    op.G = P.G;
    op.bc = P.bcU;
    op.E = T(P.E); op.lam1 = T(P.lam1); op.mu1 = T(P.mu1); op.eta = T(P.eta);
    op.vol = T(P.G.hx * P.G.hy * P.G.hz);
    op.ihx = T(1. / P.G.hx); op.ihy = T(1. / P.G.hy); op.ihz = T(1. / P.G.hz);
    ops[p].OZ = oz[p];
  }
  hipStreamSynchronize(s);  // no kernel of the previous call still reads d_ops
  hipMemcpy(d_ops, ops.data(), sizeof(BApply<T>) * ops.size(), hipMemcpyHostToDevice);
}

void fill_ops(const std::vector<vfg_handle *> &m, const void *const *x, void *const *y, int fp32,
              const std::vector<int> &oz, void *d_ops, hipStream_t s) {
  if (fp32) fill_ops_t<float>(m, x, y, oz, d_ops, s);
  else fill_ops_t<double>(m, x, y, oz, d_ops, s);
}

void launch_batch_apply(int fp32, int sched, const void *d_ops, const BItem *items, int nitems, int grid,
                        hipStream_t s) {
  const int g = grid > 0 && grid < nitems ? grid : nitems;
  launch_engine_u_apply(fp32, sched, d_ops, items, nitems, nullptr, g, nullptr, 0, s);
}

template <typename T>
static void launch_batch_t(int sched, const void *d_ops, const BItem *items, int nitems, const int *dcount, int grid,
                           const double *flags, int fstride, hipStream_t s) {
  const BApply<T> *ops = (const BApply<T> *)d_ops;
  if (sched == 1)
    hipLaunchKernelGGL((k_batch_u_apply<T, 8, 8, false>), dim3(grid), dim3(64), 0, s, ops, items, nitems, dcount,
                       flags, fstride);
  else if (sched == 2)
    hipLaunchKernelGGL((k_batch_u_apply<T, 16, 16, true>), dim3(grid), dim3(256), 0, s, ops, items, nitems, dcount,
                       flags, fstride);
  else
    hipLaunchKernelGGL((k_batch_u_apply<T, 16, 16, false>), dim3(grid), dim3(256), 0, s, ops, items, nitems, dcount,
                       flags, fstride);
}
void launch_engine_u_apply(int fp32, int sched, const void *d_ops, const BItem *items, int nitems,
                           const int *dcount, int grid, const double *flags, int fstride, hipStream_t s) {
  if (fp32) launch_batch_t<float>(sched, d_ops, items, nitems, dcount, grid, flags, fstride, s);
  else launch_batch_t<double>(sched, d_ops, items, nitems, dcount, grid, flags, fstride, s);
}

// multigrid smoothing records and their persistent launch
template <class OpT>
struct BSmooth {
  OpT op;
  int OZ;
};
using BMgPre = BSmooth<UMgPreOp>;
using BMgPost = BSmooth<UMgPostOp<false>>;
template <class Rec>
__global__ __launch_bounds__(256, 2) void k_batch_mg_smooth(const Rec *__restrict__ recs, const BItem *__restrict__ items,
                                                            const int *__restrict__ dcount,
                                                            const double *__restrict__ flags, int fstride) {
  const int n = *dcount;
  for (int it = blockIdx.x; it < n; it += gridDim.x) {
    const BItem w = items[it];
    if (flags && flags[(int64_t)w.p * fstride] == 0.) continue;  // uniform per block
    const auto op = recs[w.p].op;
    const int OZ = recs[w.p].OZ;
    __syncthreads();  // LDS of the previous item is free
    march<float, 16, 16>(op.G, op, OZ, w.bx, w.by, w.bz);
  }
}
size_t mg_smooth_op_bytes(int post) { return post ? sizeof(BMgPost) : sizeof(BMgPre); }
void make_mg_smooth_op(const DevProblem &P, int post, const float *b, const float *dinv, float c2, const float *xin,
                       float *xo, float *ro, int OZ, void *out) {
  if (post) {
    BMgPost r;
    mg_op_fill(r.op, P, b, dinv, c2);
    r.op.x = xin; r.op.xo = xo; r.op.r64 = nullptr; r.op.part = nullptr;
    r.OZ = OZ;
    *(decltype(r) *)out = r;
  } else {
    BMgPre r;
    mg_op_fill(r.op, P, b, dinv, c2);
    r.op.x = nullptr; r.op.xo = xo; r.op.ro = ro;
    r.OZ = OZ;
    *(decltype(r) *)out = r;
  }
}
void launch_batch_mg_smooth(int post, const void *d_ops, const BItem *items, const int *dcount, int grid,
                            const double *flags, int fstride, hipStream_t s) {
  if (post)
    hipLaunchKernelGGL(k_batch_mg_smooth<BMgPost>, dim3(grid), dim3(256), 0, s, (const BMgPost *)d_ops, items, dcount,
                       flags, fstride);
  else
    hipLaunchKernelGGL(k_batch_mg_smooth<BMgPre>, dim3(grid), dim3(256), 0, s, (const BMgPre *)d_ops, items, dcount,
                       flags, fstride);
}

// one member's operator entry for a persistent list (host memory `out`)
void make_uop(const DevProblem &P, const void *x, void *y, int fp32, int OZ, void *out) {
  auto fill = [&](auto *b, auto one) {
    using T = decltype(one);
    UApplyOp<T> &op = b->op;
    op.x = (const T *)x;
    op.y = (T *)y;
    op.v = sizeof(T) == 4 ? (const T *)(const void *)P.V32 : (const T *)(const void *)P.V;
    op.E_elem = P.E_elem;
    op.G = P.G;
    op.bc = P.bcU;
    op.E = T(P.E); op.lam1 = T(P.lam1); op.mu1 = T(P.mu1); op.eta = T(P.eta);
    op.vol = T(P.G.hx * P.G.hy * P.G.hz);
    op.ihx = T(1. / P.G.hx); op.ihy = T(1. / P.G.hy); op.ihz = T(1. / P.G.hz);
    b->OZ = OZ;
  };
  if (fp32) fill((BApply<float> *)out, 0.f);
  else fill((BApply<double> *)out, 0.);
}

}  // namespace vfg
