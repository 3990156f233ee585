/*
 * vforacle.c -- CPU restatement (FP64) of VPFHF's u-/V-subproblem operators.
 * TEST INFRASTRUCTURE ONLY (the checker); never linked by the product library.
 * See vforacle.h for layout conventions.  Each function cites the reference
 * file:line it follows; loop orders and the order of floating-point operations
 * mirror the reference where it is observable.
 */
#include "vforacle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <stdio.h>
#include <ctype.h>

/* ------------------------------------------------------------------ basis */

/* VFCartFEElement1DInit: VFCartFE.c:90-116 */
void vfo_elem1d(double lx, double w[3], double phi[2][3], double dphi[2][3])
{
  w[0] = 5. * lx / 18.;
  w[1] = 8. * lx / 18.;
  w[2] = 5. * lx / 18.;
  phi[0][0] = (1. + sqrt(3. / 5.)) * .5; phi[0][1] = .5; phi[0][2] = (1. - sqrt(3. / 5.)) * .5;
  phi[1][0] = (1. - sqrt(3. / 5.)) * .5; phi[1][1] = .5; phi[1][2] = (1. + sqrt(3. / 5.)) * .5;
  for (int g = 0; g < 3; g++) { dphi[0][g] = -1. / lx; dphi[1][g] = 1. / lx; }
}

/* VFCartFEElement3DInit: VFCartFE.c:216-269 (g = gi + 3*(gj + 3*gk)) */
void vfo_elem3d_init(vfo_elem *e, double lx, double ly, double lz)
{
  double wx[3], wy[3], wz[3], px[2][3], py[2][3], pz[2][3], dx[2][3], dy[2][3], dz[2][3];
  vfo_elem1d(lx, wx, px, dx);
  vfo_elem1d(ly, wy, py, dy);
  vfo_elem1d(lz, wz, pz, dz);
  e->lx = lx; e->ly = ly; e->lz = lz;
  int g = 0;
  for (int gk = 0; gk < 3; gk++)
    for (int gj = 0; gj < 3; gj++)
      for (int gi = 0; gi < 3; gi++, g++) e->weight[g] = wz[gk] * wy[gj] * wx[gi];
  g = 0;
  for (int gk = 0; gk < 3; gk++)
    for (int gj = 0; gj < 3; gj++)
      for (int gi = 0; gi < 3; gi++, g++)
        for (int k = 0; k < 2; k++)
          for (int j = 0; j < 2; j++)
            for (int i = 0; i < 2; i++) {
              e->phi[k][j][i][g]     = pz[k][gk] * py[j][gj] * px[i][gi];
              e->dphi[k][j][i][0][g] = pz[k][gk] * py[j][gj] * dx[i][gi];
              e->dphi[k][j][i][1][g] = pz[k][gk] * dy[j][gj] * px[i][gi];
              e->dphi[k][j][i][2][g] = dz[k][gk] * py[j][gj] * px[i][gi];
            }
}

/* VFMatPropGet: VFCommon.c:583-586 */
void vfo_lame(double E, double nu, double *lambda, double *mu)
{
  *lambda = E * nu / (1. + nu) / (1. - 2. * nu);
  *mu     = E / (1. + nu) * .5;
}

/* VFPropGet: VFCommon.c:482-492 */
double vfo_atcv(int atnum) { return atnum == 1 ? 2. / 3. : .5; }

#define PHI(e, a, g) ((e)->phi[(a) >> 2][((a) >> 1) & 1][(a) & 1][g])
#define DPHI(e, a, d, g) ((e)->dphi[(a) >> 2][((a) >> 1) & 1][(a) & 1][d][g])

/* ------------------------------------------------------- element kernels */

/* VF_BilinearFormU3D_local: VFMech.c:227-290 */
void vfo_bilinear_u(double *Ke, const double v8[8], double lambda, double mu,
                    double eta, const vfo_elem *e)
{
  double s[27];
  for (int g = 0; g < 27; g++) s[g] = 0.;
  for (int a = 0; a < 8; a++)
    for (int g = 0; g < 27; g++) s[g] += v8[a] * PHI(e, a, g);
  for (int g = 0; g < 27; g++) s[g] = s[g] * s[g] + eta;
  for (int l = 0; l < 576; l++) Ke[l] = 0.;
  int l = 0;
  for (int a = 0; a < 8; a++)
    for (int c1 = 0; c1 < 3; c1++)
      for (int b = 0; b < 8; b++)
        for (int c2 = 0; c2 < 3; c2++, l++)
          for (int g = 0; g < 27; g++) {
            double mg = lambda * s[g] * DPHI(e, a, c1, g) * DPHI(e, b, c2, g);
            mg += mu * s[g] * DPHI(e, a, c2, g) * DPHI(e, b, c1, g);
            if (c1 == c2) {
              mg += mu * s[g] *
                    (DPHI(e, a, 0, g) * DPHI(e, b, 0, g) +
                     DPHI(e, a, 1, g) * DPHI(e, b, 1, g) +
                     DPHI(e, a, 2, g) * DPHI(e, b, 2, g));
            }
            Ke[l] += e->weight[g] * mg;
          }
}

/* VF_GradientUThermoPoro3D_local: VFMech.c:386-447 (accumulates into b24) */
void vfo_grad_u_thermoporo(double *b24, const double v8[8], const double dth8[8],
                           const double p8[8], double lambda, double mu,
                           double alpha, double beta, double eta, const vfo_elem *e)
{
  double coefalpha = (3. * lambda + 2. * mu) * alpha;
  double th[27], pr[27], vv[27];
  for (int g = 0; g < 27; g++) { th[g] = 0; pr[g] = 0; vv[g] = 0.; }
  for (int g = 0; g < 27; g++)
    for (int a = 0; a < 8; a++) {
      th[g] += PHI(e, a, g) * dth8[a];
      pr[g] += PHI(e, a, g) * p8[a];
      vv[g] += PHI(e, a, g) * v8[a];
    }
  for (int g = 0; g < 27; g++)
    for (int a = 0, l = 0; a < 8; a++)
      for (int c = 0; c < 3; c++, l++)
        b24[l] += e->weight[g] * DPHI(e, a, c, g) *
                  (coefalpha * th[g] * (vv[g] * vv[g] + eta) + beta * pr[g] * vv[g]);
}

/* VF_GradientUCrackPressure3D_local: VFMech.c:533-589 (accumulates) */
void vfo_grad_u_crackpressure(double *b24, const double v8[8], const double p8[8],
                              const vfo_elem *e)
{
  double pr[27], gv[3][27];
  for (int g = 0; g < 27; g++) { pr[g] = 0; gv[0][g] = gv[1][g] = gv[2][g] = 0.; }
  for (int a = 0; a < 8; a++)
    for (int g = 0; g < 27; g++) {
      pr[g] += PHI(e, a, g) * p8[a];
      for (int c = 0; c < 3; c++) gv[c][g] += DPHI(e, a, c, g) * v8[a];
    }
  for (int a = 0, l = 0; a < 8; a++)
    for (int c = 0; c < 3; c++, l++)
      for (int g = 0; g < 27; g++) b24[l] += e->weight[g] * pr[g] * gv[c][g] * PHI(e, a, g);
}

/* ElasticEnergyDensity3D_local: VFMech.c:32-106 (thermal strain folded into
 * the diagonal strains, no Biot term) */
void vfo_elastic_energy_density(double W[27], const double u24[24], const double dth8[8],
                                double lambda, double mu, double alpha, const vfo_elem *e)
{
  double e11[27], e22[27], e33[27], e12[27], e23[27], e13[27];
  for (int g = 0; g < 27; g++) {
    e11[g] = e22[g] = e33[g] = e12[g] = e23[g] = e13[g] = 0; W[g] = 0.;
  }
  for (int a = 0; a < 8; a++)
    for (int g = 0; g < 27; g++) {
      const double *u = u24 + 3 * a;
      e11[g] += DPHI(e, a, 0, g) * u[0] - alpha * PHI(e, a, g) * dth8[a];
      e22[g] += DPHI(e, a, 1, g) * u[1] - alpha * PHI(e, a, g) * dth8[a];
      e33[g] += DPHI(e, a, 2, g) * u[2] - alpha * PHI(e, a, g) * dth8[a];
      e12[g] += (DPHI(e, a, 0, g) * u[1] + DPHI(e, a, 1, g) * u[0]) * .5;
      e23[g] += (DPHI(e, a, 1, g) * u[2] + DPHI(e, a, 2, g) * u[1]) * .5;
      e13[g] += (DPHI(e, a, 0, g) * u[2] + DPHI(e, a, 2, g) * u[0]) * .5;
    }
  for (int g = 0; g < 27; g++) {
    double s11 = (lambda + 2. * mu) * e11[g] + lambda * e22[g] + lambda * e33[g];
    double s22 = lambda * e11[g] + (lambda + 2. * mu) * e22[g] + lambda * e33[g];
    double s33 = lambda * e11[g] + lambda * e22[g] + (lambda + 2. * mu) * e33[g];
    double s12 = 2. * mu * e12[g];
    double s23 = 2. * mu * e23[g];
    double s13 = 2. * mu * e13[g];
    W[g] = (s11 * e11[g] + s22 * e22[g] + s33 * e33[g]) * .5 + s12 * e12[g] + s23 * e23[g] +
           s13 * e13[g];
  }
}

/* tr(eps(u)) at the 27 Gauss points, accumulated node by node as the
 * reference does (VFMech.c:320-326) */
static void trace_eps(double tr[27], const double u24[24], const vfo_elem *e)
{
  for (int g = 0; g < 27; g++) tr[g] = 0.;
  for (int a = 0; a < 8; a++)
    for (int g = 0; g < 27; g++)
      tr[g] += DPHI(e, a, 0, g) * u24[3 * a] + DPHI(e, a, 1, g) * u24[3 * a + 1] +
               DPHI(e, a, 2, g) * u24[3 * a + 2];
}

/* VF_BilinearFormUNoCompression3D_local: VFMech.c:299-376 */
void vfo_bilinear_u_nocomp(double *Ke, const double u24[24], const double v8[8], double lambda,
                           double mu, double eta, const vfo_elem *e)
{
  const double kappa = lambda + 2. * mu / 3.;
  double s[27], tr[27];
  trace_eps(tr, u24, e);
  for (int g = 0; g < 27; g++) s[g] = 0.;
  for (int a = 0; a < 8; a++)
    for (int g = 0; g < 27; g++) s[g] += v8[a] * PHI(e, a, g);
  for (int g = 0; g < 27; g++) s[g] = s[g] * s[g] + eta;
  for (int l = 0; l < 576; l++) Ke[l] = 0.;
  int l = 0;
  for (int a = 0; a < 8; a++)
    for (int c1 = 0; c1 < 3; c1++)
      for (int b = 0; b < 8; b++)
        for (int c2 = 0; c2 < 3; c2++, l++)
          for (int g = 0; g < 27; g++) {
            const double d12 = DPHI(e, a, c1, g) * DPHI(e, b, c2, g);
            double mg = tr[g] < 0. ? kappa * (1. + eta) * d12 : kappa * s[g] * d12;
            mg -= 2. * mu * s[g] * d12 / 3.;
            mg += mu * s[g] * DPHI(e, a, c2, g) * DPHI(e, b, c1, g);
            if (c1 == c2)
              mg += mu * s[g] *
                    (DPHI(e, a, 0, g) * DPHI(e, b, 0, g) + DPHI(e, a, 1, g) * DPHI(e, b, 1, g) +
                     DPHI(e, a, 2, g) * DPHI(e, b, 2, g));
            Ke[l] += e->weight[g] * mg;
          }
}

/* VF_GradientUThermoPoroNoCompression3D_local: VFMech.c:456-524 */
void vfo_grad_u_thermoporo_nocomp(double *b24, const double u24[24], const double v8[8],
                                  const double dth8[8], const double p8[8], double lambda,
                                  double mu, double alpha, double beta, double eta,
                                  const vfo_elem *e)
{
  const double coefalpha = (3. * lambda + 2. * mu) * alpha;
  double th[27], pr[27], vv[27], tr[27];
  trace_eps(tr, u24, e);
  for (int g = 0; g < 27; g++) { th[g] = 0; pr[g] = 0; vv[g] = 0.; }
  for (int a = 0; a < 8; a++)
    for (int g = 0; g < 27; g++) {
      th[g] += PHI(e, a, g) * dth8[a];
      pr[g] += PHI(e, a, g) * p8[a];
      vv[g] += PHI(e, a, g) * v8[a];
    }
  for (int g = 0; g < 27; g++) {
    const double f = tr[g] < 0. ? coefalpha * th[g] * (1. + eta) + beta * pr[g]
                                : coefalpha * th[g] * (vv[g] * vv[g] + eta) + beta * pr[g] * vv[g];
    for (int a = 0, l = 0; a < 8; a++)
      for (int c = 0; c < 3; c++, l++) b24[l] += e->weight[g] * DPHI(e, a, c, g) * f;
  }
}

/* ElasticEnergyDensitySphericalDeviatoricNoCompression3D_local:
 * VFMech.c:115-218.  Stress from the strain less the thermal strain, energy
 * density 1/2 sigma : eps with the total strain; S = kappa tr^2 / 2 where
 * tr(eps) >= 0 (else 0), D = W - kappa tr^2 / 2. */
void vfo_energy_density_sd_nocomp(double S[27], double D[27], const double u24[24],
                                  const double dth8[8], double lambda, double mu, double alpha,
                                  const vfo_elem *e)
{
  const double kappa = lambda + 2. * mu / 3.;
  double e11[27], e22[27], e33[27], e12[27], e23[27], e13[27], e0[27];
  for (int g = 0; g < 27; g++) e11[g] = e22[g] = e33[g] = e12[g] = e23[g] = e13[g] = e0[g] = 0.;
  for (int a = 0; a < 8; a++)
    for (int g = 0; g < 27; g++) {
      const double *u = u24 + 3 * a;
      e11[g] += DPHI(e, a, 0, g) * u[0];
      e22[g] += DPHI(e, a, 1, g) * u[1];
      e33[g] += DPHI(e, a, 2, g) * u[2];
      e12[g] += (DPHI(e, a, 0, g) * u[1] + DPHI(e, a, 1, g) * u[0]) * .5;
      e23[g] += (DPHI(e, a, 1, g) * u[2] + DPHI(e, a, 2, g) * u[1]) * .5;
      e13[g] += (DPHI(e, a, 0, g) * u[2] + DPHI(e, a, 2, g) * u[0]) * .5;
      e0[g] += alpha * PHI(e, a, g) * dth8[a];
    }
  for (int g = 0; g < 27; g++) {
    const double a11 = e11[g] - e0[g], a22 = e22[g] - e0[g], a33 = e33[g] - e0[g];
    const double s11 = (lambda + 2. * mu) * a11 + lambda * a22 + lambda * a33;
    const double s22 = lambda * a11 + (lambda + 2. * mu) * a22 + lambda * a33;
    const double s33 = lambda * a11 + lambda * a22 + (lambda + 2. * mu) * a33;
    const double W = (s11 * e11[g] + s22 * e22[g] + s33 * e33[g]) * .5 + 2. * mu * e12[g] * e12[g] +
                     2. * mu * e23[g] * e23[g] + 2. * mu * e13[g] * e13[g];
    const double tr = e11[g] + e22[g] + e33[g];
    S[g] = tr < 0. ? 0. : kappa * tr * tr * .5;
    D[g] = W - kappa * tr * tr * .5;
  }
}

/* VF_BilinearFormVCouplingNoCompression3D_local: VFMech.c:2125-2153 */
void vfo_bilinear_v_coupling_nocomp(double *K, const double u24[24], const double dth8[8],
                                    double lambda, double mu, double alpha, const vfo_elem *e)
{
  double S[27], D[27];
  vfo_energy_density_sd_nocomp(S, D, u24, dth8, lambda, mu, alpha, e);
  for (int a = 0, l = 0; a < 8; a++)
    for (int b = 0; b < 8; b++, l++)
      for (int g = 0; g < 27; g++)
        K[l] += e->weight[g] * PHI(e, a, g) * PHI(e, b, g) * (S[g] + D[g]) * 2.;
}

/* VF_BilinearFormVAT13D_local / VF_BilinearFormVAT23D_local: VFMech.c:2040-2084 */
void vfo_bilinear_v_at(double *K, int atnum, double Gc, double atCv, double eps,
                       const vfo_elem *e)
{
  double coef = Gc / atCv * .5;
  for (int a = 0, l = 0; a < 8; a++)
    for (int b = 0; b < 8; b++, l++)
      for (int g = 0; g < 27; g++) {
        double gg = DPHI(e, a, 0, g) * DPHI(e, b, 0, g) + DPHI(e, a, 1, g) * DPHI(e, b, 1, g) +
                    DPHI(e, a, 2, g) * DPHI(e, b, 2, g);
        if (atnum == 2)
          K[l] += coef * e->weight[g] * (PHI(e, a, g) * PHI(e, b, g) / eps + gg * eps);
        else
          K[l] += coef * e->weight[g] * gg * eps;
      }
}

/* VF_BilinearFormVCoupling3D_local: VFMech.c:2093-2115 */
void vfo_bilinear_v_coupling(double *K, const double u24[24], const double dth8[8],
                             double lambda, double mu, double alpha, const vfo_elem *e)
{
  double W[27];
  vfo_elastic_energy_density(W, u24, dth8, lambda, mu, alpha, e);
  for (int a = 0, l = 0; a < 8; a++)
    for (int b = 0; b < 8; b++, l++)
      for (int g = 0; g < 27; g++)
        K[l] += e->weight[g] * PHI(e, a, g) * PHI(e, b, g) * W[g] * 2.;
}

/* VF_ResidualVAT13D_local / VF_ResidualVAT23D_local: VFMech.c:2350-2384 */
void vfo_residual_v_at(double b8[8], int atnum, double Gc, double atCv, double eps,
                       const vfo_elem *e)
{
  double coef = (atnum == 2) ? Gc / atCv / eps * .5 : Gc / atCv / eps * .25;
  for (int a = 0; a < 8; a++)
    for (int g = 0; g < 27; g++) b8[a] += e->weight[g] * PHI(e, a, g) * coef;
}

/* VF_ResidualVCrackPressure3D_local: VFMech.c:2294-2341 */
void vfo_residual_v_crackpressure(double b8[8], const double u24[24], const double p8[8],
                                  const vfo_elem *e)
{
  double pr[27], ue[3][27];
  for (int g = 0; g < 27; g++) { pr[g] = 0; ue[0][g] = ue[1][g] = ue[2][g] = 0.; }
  for (int a = 0; a < 8; a++)
    for (int g = 0; g < 27; g++) {
      pr[g] += PHI(e, a, g) * p8[a];
      for (int c = 0; c < 3; c++) ue[c][g] += PHI(e, a, g) * u24[3 * a + c];
    }
  for (int a = 0; a < 8; a++)
    for (int c = 0; c < 3; c++)
      for (int g = 0; g < 27; g++) b8[a] += e->weight[g] * pr[g] * ue[c][g] * DPHI(e, a, c, g);
}

/* VF_ElasticEnergy3D_local: VFMech.c:725-809 */
double vfo_elastic_energy_local(const double u24[24], const double v8[8], const double dth8[8],
                                const double p8[8], double lambda, double mu, double alpha,
                                double beta, double eta, const vfo_elem *e)
{
  double D11[27], D22[27], D33[27], D12[27], D23[27], D13[27], ve[27], pe[27], te[27];
  double threekappa = 3. * lambda + 2. * mu;
  double coefbeta = beta / threekappa;
  for (int g = 0; g < 27; g++) {
    D11[g] = D22[g] = D33[g] = D12[g] = D23[g] = D13[g] = 0.;
    ve[g] = pe[g] = te[g] = 0.;
  }
  for (int a = 0; a < 8; a++)
    for (int g = 0; g < 27; g++) {
      const double *u = u24 + 3 * a;
      D11[g] += DPHI(e, a, 0, g) * u[0];
      D22[g] += DPHI(e, a, 1, g) * u[1];
      D33[g] += DPHI(e, a, 2, g) * u[2];
      D12[g] += (DPHI(e, a, 0, g) * u[1] + DPHI(e, a, 1, g) * u[0]) * .5;
      D23[g] += (DPHI(e, a, 1, g) * u[2] + DPHI(e, a, 2, g) * u[1]) * .5;
      D13[g] += (DPHI(e, a, 0, g) * u[2] + DPHI(e, a, 2, g) * u[0]) * .5;
      ve[g] += PHI(e, a, g) * v8[a];
      pe[g] += PHI(e, a, g) * p8[a];
      te[g] += PHI(e, a, g) * dth8[a];
    }
  double E = 0.;
  for (int g = 0; g < 27; g++) {
    double tr = D11[g] + D22[g] + D33[g];
    D11[g] -= tr / 3.;
    D22[g] -= tr / 3.;
    D33[g] -= tr / 3.;
    E += (D11[g] * D11[g] + D22[g] * D22[g] + D33[g] * D33[g] +
          (D12[g] * D12[g] + D23[g] * D23[g] + D13[g] * D13[g]) * 2.) *
         ve[g] * ve[g] * (1. + eta) * mu * e->weight[g];
    double WD = (tr / 3. - alpha * te[g]) * ve[g] - coefbeta * pe[g];
    E += 3. * threekappa / 2. * WD * WD * e->weight[g];
    E += threekappa / 6. * tr * tr * eta * e->weight[g];
  }
  return E;
}

/* VF_ElasticEnergyNoCompression3D_local: VFMech.c:819-908 (the spherical
 * term uses s = 1 where tr(eps) < 0, s = v elsewhere) */
double vfo_elastic_energy_nocomp_local(const double u24[24], const double v8[8],
                                       const double dth8[8], const double p8[8], double lambda,
                                       double mu, double alpha, double beta, double eta,
                                       const vfo_elem *e)
{
  double D11[27], D22[27], D33[27], D12[27], D23[27], D13[27], ve[27], pe[27], te[27];
  double threekappa = 3. * lambda + 2. * mu;
  double coefbeta = beta / threekappa;
  for (int g = 0; g < 27; g++) {
    D11[g] = D22[g] = D33[g] = D12[g] = D23[g] = D13[g] = 0.;
    ve[g] = pe[g] = te[g] = 0.;
  }
  for (int a = 0; a < 8; a++)
    for (int g = 0; g < 27; g++) {
      const double *u = u24 + 3 * a;
      D11[g] += DPHI(e, a, 0, g) * u[0];
      D22[g] += DPHI(e, a, 1, g) * u[1];
      D33[g] += DPHI(e, a, 2, g) * u[2];
      D12[g] += (DPHI(e, a, 0, g) * u[1] + DPHI(e, a, 1, g) * u[0]) * .5;
      D23[g] += (DPHI(e, a, 1, g) * u[2] + DPHI(e, a, 2, g) * u[1]) * .5;
      D13[g] += (DPHI(e, a, 0, g) * u[2] + DPHI(e, a, 2, g) * u[0]) * .5;
      ve[g] += PHI(e, a, g) * v8[a];
      pe[g] += PHI(e, a, g) * p8[a];
      te[g] += PHI(e, a, g) * dth8[a];
    }
  double E = 0.;
  for (int g = 0; g < 27; g++) {
    double tr = D11[g] + D22[g] + D33[g];
    double s = tr < 0. ? 1. : ve[g];
    D11[g] -= tr / 3.;
    D22[g] -= tr / 3.;
    D33[g] -= tr / 3.;
    E += (D11[g] * D11[g] + D22[g] * D22[g] + D33[g] * D33[g] +
          (D12[g] * D12[g] + D23[g] * D23[g] + D13[g] * D13[g]) * 2.) *
         ve[g] * ve[g] * (1. + eta) * mu * e->weight[g];
    double WD = (tr / 3. - alpha * te[g]) * s - coefbeta * pe[g];
    E += 3. * threekappa / 2. * WD * WD * e->weight[g];
    E += threekappa / 6. * tr * tr * eta * e->weight[g];
  }
  return E;
}

/* VF_PressureWork3D_local: VFMech.c:920-974.  The reference accumulates the
 * per-point terms straight into the running global sum; so does this. */
static void pressure_work_acc(double *acc, const double u24[24], const double v8[8],
                              const double p8[8], const vfo_elem *e)
{
  double pr[27], gv[3][27], ue[3][27];
  for (int g = 0; g < 27; g++) {
    pr[g] = 0;
    for (int c = 0; c < 3; c++) gv[c][g] = ue[c][g] = 0.;
  }
  for (int a = 0; a < 8; a++)
    for (int g = 0; g < 27; g++) {
      pr[g] += PHI(e, a, g) * p8[a];
      for (int c = 0; c < 3; c++) {
        gv[c][g] += DPHI(e, a, c, g) * v8[a];
        ue[c][g] += PHI(e, a, g) * u24[3 * a + c];
      }
    }
  for (int g = 0; g < 27; g++)
    *acc += e->weight[g] * pr[g] *
            (ue[0][g] * gv[0][g] + ue[1][g] * gv[1][g] + ue[2][g] * gv[2][g]);
}

double vfo_pressure_work_local(const double u24[24], const double v8[8], const double p8[8],
                               const vfo_elem *e)
{
  double acc = 0.;
  pressure_work_acc(&acc, u24, v8, p8, e);
  return acc;
}

/* VF_AT1SurfaceEnergy3D_local / VF_AT2SurfaceEnergy3D_local: VFMech.c:2394-2465
 * (accumulates straight into the running sum, as the reference does) */
static void surface_energy_acc(double *acc, const double v8[8], int atnum, double Gc,
                               double atCv, double eps, const vfo_elem *e)
{
  double coef = Gc / atCv * .25;
  double ve[27], g0[27], g1[27], g2[27];
  for (int g = 0; g < 27; g++) ve[g] = g0[g] = g1[g] = g2[g] = 0.;
  for (int a = 0; a < 8; a++)
    for (int g = 0; g < 27; g++) {
      ve[g] += v8[a] * PHI(e, a, g);
      g0[g] += v8[a] * DPHI(e, a, 0, g);
      g1[g] += v8[a] * DPHI(e, a, 1, g);
      g2[g] += v8[a] * DPHI(e, a, 2, g);
    }
  for (int g = 0; g < 27; g++) {
    double grad2 = g0[g] * g0[g] + g1[g] * g1[g] + g2[g] * g2[g];
    if (atnum == 2)
      *acc += e->weight[g] * ((1. - ve[g]) * (1. - ve[g]) / eps + grad2 * eps) * coef;
    else
      *acc += e->weight[g] * ((1. - ve[g]) / eps + grad2 * eps) * coef;
  }
}

double vfo_surface_energy_local(const double v8[8], int atnum, double Gc, double atCv,
                                double eps, const vfo_elem *e)
{
  double acc = 0.;
  surface_energy_acc(&acc, v8, atnum, Gc, atCv, eps, e);
  return acc;
}

/* VolumetricCrackOpening3D_local: VFPermfield.c:15-71 */
double vfo_crack_volume_local(const double u24[24], const double v8[8], const vfo_elem *e,
                              double *cellavg)
{
  double ud[27], vd[27], wd[27], dx[27], dy[27], dz[27];
  for (int g = 0; g < 27; g++) ud[g] = vd[g] = wd[g] = dx[g] = dy[g] = dz[g] = 0.;
  for (int g = 0; g < 27; g++)
    for (int a = 0; a < 8; a++) {
      ud[g] += u24[3 * a + 0] * PHI(e, a, g);
      vd[g] += u24[3 * a + 1] * PHI(e, a, g);
      wd[g] += u24[3 * a + 2] * PHI(e, a, g);
      dx[g] += v8[a] * DPHI(e, a, 0, g);
      dy[g] += v8[a] * DPHI(e, a, 1, g);
      dz[g] += v8[a] * DPHI(e, a, 2, g);
    }
  double vol = 0., cell = 0., evol = 0.;
  for (int g = 0; g < 27; g++) {
    cell += (ud[g] * dx[g] + vd[g] * dy[g] + wd[g] * dz[g]) * e->weight[g];
    evol += e->weight[g];
    vol += (ud[g] * dx[g] + vd[g] * dy[g] + wd[g] * dz[g]) * e->weight[g];
  }
  if (cellavg) *cellavg = cell / evol;
  return vol;
}

/* ------------------------------------------------------------ penny cracks */

/* VFDistanceToPennyCrack: VFCracks.c:101-126 */
double vfo_distance_to_penny(const double x[3], const double c[3], double r,
                             double theta_deg, double phi_deg)
{
  const double PI = 3.14159265358979323846;
  double n[3], tau[3];
  n[0] = cos(theta_deg * PI / 180.) * sin(phi_deg * PI / 180.);
  n[1] = sin(theta_deg * PI / 180.) * sin(phi_deg * PI / 180.);
  n[2] = cos(phi_deg * PI / 180.);
  double xdotn = (x[0] - c[0]) * n[0] + (x[1] - c[1]) * n[1] + (x[2] - c[2]) * n[2];
  tau[0] = (x[0] - c[0]) - xdotn * n[0];
  tau[1] = (x[1] - c[1]) - xdotn * n[1];
  tau[2] = (x[2] - c[2]) - xdotn * n[2];
  double l = sqrt(tau[0] * tau[0] + tau[1] * tau[1] + tau[2] * tau[2]);
  if (l <= r) return sqrt(xdotn * xdotn);
  return sqrt(pow(xdotn, 2) + pow(l - r, 2));
}

/* VFPennyCrackBuildVAT1 / VFPennyCrackBuildVAT2: VFCracks.c:136-227 */
void vfo_penny_build_v(double *V, const vfo_problem *P, const double center[3], double r,
                       double theta_deg, double phi_deg, double thickness)
{
  int64_t nn = vfo_nnodes(P);
  double *X = malloc(sizeof(double) * P->nx), *Y = malloc(sizeof(double) * P->ny),
         *Z = malloc(sizeof(double) * P->nz);
  vfo_coords(P, X, Y, Z);
  for (int64_t n = 0; n < nn; n++) V[n] = 1.0;
  double eps = P->epsilon;
  for (int k = 0; k < P->nz; k++)
    for (int j = 0; j < P->ny; j++)
      for (int i = 0; i < P->nx; i++) {
        int64_t n = i + (int64_t)P->nx * (j + (int64_t)P->ny * k);
        double x[3] = {X[i], Y[j], Z[k]};
        if (r == 0) { V[n] = 1.; continue; }
        double dist = vfo_distance_to_penny(x, center, r, theta_deg, phi_deg);
        dist -= thickness / 2.;
        if (P->atnum == 1) {
          if (dist <= 0) V[n] = 0.;
          else if (dist < 2. * eps) V[n] = dist / eps * (1. - .25 * dist / eps);
          else V[n] = 1.;
        } else {
          if (dist <= 0) V[n] = 0.;
          else V[n] = 1. - exp(-dist / 2 / eps);
        }
      }
  free(X); free(Y); free(Z);
}

/* --------------------------------------------------------------- geometry */

int64_t vfo_nnodes(const vfo_problem *P) { return (int64_t)P->nx * P->ny * P->nz; }
int64_t vfo_ncells(const vfo_problem *P)
{
  return (int64_t)(P->nx - 1) * (P->ny - 1) * (P->nz - 1);
}

/* VFGeometryInitialize: VFCommon.c:391-406 */
void vfo_coords(const vfo_problem *P, double *X, double *Y, double *Z)
{
  X[0] = 0.; Y[0] = 0.; Z[0] = 0.;
  for (int k = 1; k < P->nz; k++) Z[k] = k * P->lz / (P->nz - 1.);
  for (int j = 1; j < P->ny; j++) Y[j] = j * P->ly / (P->ny - 1.);
  for (int i = 1; i < P->nx; i++) X[i] = i * P->lx / (P->nx - 1.);
}

static inline int64_t NODE(const vfo_problem *P, int i, int j, int k)
{
  return i + (int64_t)P->nx * (j + (int64_t)P->ny * k);
}

/* element material: per element E if given (VFMatProp per layer otherwise) */
static void elem_lame(const vfo_problem *P, int64_t cell, double *lambda, double *mu)
{
  double E = P->E_elem ? P->E_elem[cell] : P->E;
  vfo_lame(E, P->nu, lambda, mu);
}

/* ----------------------------------------------------------------- BC kit */

void vfo_bc_init(vfo_bc *bc) { memset(bc, 0, sizeof(*bc)); }

/* dof of (i,j,k) lies on the entity?  faces X0..Z1, then vertices */
static int on_face(const vfo_problem *P, int f, int i, int j, int k)
{
  switch (f) {
    case VFO_X0: return i == 0;
    case VFO_X1: return i == P->nx - 1;
    case VFO_Y0: return j == 0;
    case VFO_Y1: return j == P->ny - 1;
    case VFO_Z0: return k == 0;
    default:     return k == P->nz - 1;
  }
}
static int on_vertex(const vfo_problem *P, int v, int i, int j, int k)
{
  int ix = v & 1, iy = (v >> 1) & 1, iz = (v >> 2) & 1;
  return i == (ix ? P->nx - 1 : 0) && j == (iy ? P->ny - 1 : 0) && k == (iz ? P->nz - 1 : 0);
}
/* vertex application order of VecSetFromBC & co: VFCartFE.c:421-451 */
static const int VORDER[8] = {VFO_X0Y0Z0, VFO_X0Y0Z1, VFO_X0Y1Z0, VFO_X0Y1Z1,
                              VFO_X1Y0Z0, VFO_X1Y0Z1, VFO_X1Y1Z0, VFO_X1Y1Z1};

/* MatApplyDirichletBC row set (VFCartFE.c:1147-1296): union of entities */
void vfo_bc_mask(const vfo_problem *P, const vfo_bc *bc, int dof, uint8_t *mask)
{
  for (int k = 0; k < P->nz; k++)
    for (int j = 0; j < P->ny; j++)
      for (int i = 0; i < P->nx; i++)
        for (int c = 0; c < dof; c++) {
          int m = 0;
          for (int f = 0; f < 6; f++)
            if (bc[c].face[f] != VFO_NONE && on_face(P, f, i, j, k)) m = 1;
          for (int v = 0; v < 8; v++)
            if (bc[c].vertex[v] != VFO_NONE && on_vertex(P, v, i, j, k)) m = 1;
          mask[dof * NODE(P, i, j, k) + c] = (uint8_t)m;
        }
}

/* VecSetFromBC: VFCartFE.c:279-483 -- faces X0,X1,Y0,Y1,Z0,Z1 then vertices */
void vfo_vec_set_from_bc(const vfo_problem *P, const vfo_bc *bc, int dof, double *B)
{
  for (int c = 0; c < dof; c++) {
    for (int f = 0; f < 6; f++) {
      if (bc[c].face[f] == VFO_NONE) continue;
      double val = bc[c].face[f] == VFO_ZERO ? 0. : bc[c].face[f] == VFO_ONE ? 1. : bc[c].faceValue[f];
      for (int k = 0; k < P->nz; k++)
        for (int j = 0; j < P->ny; j++)
          for (int i = 0; i < P->nx; i++)
            if (on_face(P, f, i, j, k)) B[dof * NODE(P, i, j, k) + c] = val;
    }
    for (int q = 0; q < 8; q++) {
      int v = VORDER[q];
      int ix = v & 1, iy = (v >> 1) & 1, iz = (v >> 2) & 1;
      int64_t n = NODE(P, ix ? P->nx - 1 : 0, iy ? P->ny - 1 : 0, iz ? P->nz - 1 : 0);
      if (bc[c].vertex[v] == VFO_FIXED) B[dof * n + c] = bc[c].vertexValue[v];
      if (bc[c].vertex[v] == VFO_ONE) B[dof * n + c] = 1.;
      if (bc[c].vertex[v] == VFO_ZERO) B[dof * n + c] = 0.;
    }
  }
}

/* VecApplyDirichletBC (VFCartFE.c:493-738) and ResidualApplyDirichletBC
 * (VFCartFE.c:747-999) share one traversal: mode 0 writes the BC value into
 * vec, mode 1 writes U - value into F. */
static void dirichlet_traverse(const vfo_problem *P, const vfo_bc *bc, int dof, double *out,
                               const double *U, const double *B, int mode)
{
  for (int c = 0; c < dof; c++) {
    for (int f = 0; f < 6; f++) {
      int t = bc[c].face[f];
      if (t == VFO_NONE) continue;
      for (int k = 0; k < P->nz; k++)
        for (int j = 0; j < P->ny; j++)
          for (int i = 0; i < P->nx; i++) {
            if (!on_face(P, f, i, j, k)) continue;
            int64_t d = dof * NODE(P, i, j, k) + c;
            double val = t == VFO_FIXED ? B[d] : t == VFO_ONE ? 1. : 0.;
            out[d] = mode == 0 ? val : (t == VFO_ZERO ? U[d] : U[d] - val);
          }
    }
    for (int q = 0; q < 8; q++) {
      int v = VORDER[q], t = bc[c].vertex[v];
      if (t == VFO_NONE) continue;
      int ix = v & 1, iy = (v >> 1) & 1, iz = (v >> 2) & 1;
      int64_t d = dof * NODE(P, ix ? P->nx - 1 : 0, iy ? P->ny - 1 : 0, iz ? P->nz - 1 : 0) + c;
      double val = t == VFO_FIXED ? B[d] : t == VFO_ONE ? 1. : 0.;
      out[d] = mode == 0 ? val : (t == VFO_ZERO ? U[d] : U[d] - val);
    }
  }
}

void vfo_vec_apply_dirichlet(const vfo_problem *P, const vfo_bc *bc, int dof, double *vec,
                             const double *B)
{
  dirichlet_traverse(P, bc, dof, vec, NULL, B, 0);
}

void vfo_residual_apply_dirichlet(const vfo_problem *P, const vfo_bc *bc, int dof, double *F,
                                  const double *U, const double *B)
{
  dirichlet_traverse(P, bc, dof, F, U, B, 1);
}

/* Options-file parser for the BC subset of VFBCSetFromOptions
 * (VFCartFE.c:1506-1574): "-<prefix>_<FACE|VERTEX>_BC[_c] TYPE" and
 * "-<prefix>_<FACE>[_c] value".  Edges are not parsed (as in the reference). */
static const char *FNAME[6] = {"X0", "X1", "Y0", "Y1", "Z0", "Z1"};
static const char *VNAME[8] = {"X0Y0Z0", "X1Y0Z0", "X0Y1Z0", "X1Y1Z0",
                               "X0Y0Z1", "X1Y0Z1", "X0Y1Z1", "X1Y1Z1"};
static const char *TNAME[4] = {"NONE", "ZERO", "ONE", "FIXED"};

static int streq_ci(const char *a, const char *b)
{
  for (; *a && *b; a++, b++)
    if (tolower((unsigned char)*a) != tolower((unsigned char)*b)) return 0;
  return *a == *b;
}

int vfo_bc_parse_options(const char *text, const char *prefix, int dof, vfo_bc *bc)
{
  int nset = 0;
  const char *p = text;
  char key[256], val[256];
  while (*p) {
    while (*p && isspace((unsigned char)*p)) p++;
    if (!*p) break;
    if (*p == '#') { while (*p && *p != '\n') p++; continue; }
    int n = 0;
    while (*p && !isspace((unsigned char)*p) && n < 255) key[n++] = *p++;
    key[n] = 0;
    while (*p && (*p == ' ' || *p == '\t')) p++;
    n = 0;
    /* a value follows unless the next token is another option ("-" + letter);
       negative numbers ("-1.", "-.5") are values */
    int is_key = *p == '-' && (isalpha((unsigned char)p[1]) || p[1] == '_');
    if (*p && *p != '\n' && !is_key) {
      while (*p && !isspace((unsigned char)*p) && n < 255) val[n++] = *p++;
    }
    val[n] = 0;
    if (key[0] != '-') continue;
    for (int c = 0; c < dof; c++) {
      char want[128];
      for (int f = 0; f < 14; f++) {
        const char *ent = f < 6 ? FNAME[f] : VNAME[f - 6];
        if (dof == 1) snprintf(want, sizeof want, "-%s_%s_BC", prefix, ent);
        else snprintf(want, sizeof want, "-%s_%s_BC_%d", prefix, ent, c);
        if (streq_ci(key, want)) {
          int t = -1;
          for (int q = 0; q < 4; q++)
            if (streq_ci(val, TNAME[q])) t = q;
          if (t < 0) return -1;
          if (f < 6) bc[c].face[f] = t; else bc[c].vertex[f - 6] = t;
          nset++;
        }
        if (f < 6) {
          if (dof == 1) snprintf(want, sizeof want, "-%s_%s", prefix, ent);
          else snprintf(want, sizeof want, "-%s_%s_%d", prefix, ent, c);
          if (streq_ci(key, want)) { bc[c].faceValue[f] = atof(val); nset++; }
        }
      }
    }
  }
  return nset;
}

/* ------------------------------------------------------------ ELL matrices */

vfo_ell *vfo_ell_alloc(int64_t nrows, int width)
{
  vfo_ell *A = calloc(1, sizeof *A);
  A->nrows = nrows;
  A->width = width;
  A->col = malloc(sizeof(int64_t) * nrows * width);
  A->val = calloc((size_t)nrows * width, sizeof(double));
  return A;
}

void vfo_ell_free(vfo_ell *A)
{
  if (!A) return;
  free(A->col); free(A->val); free(A);
}

/* Row-parallel ELL product; each row sums its slots in increasing column
 * order, as PETSc's AIJ MatMult does. */
void vfo_ell_matvec(const vfo_ell *A, const double *x, double *y)
{
  const int W = A->width;
#pragma omp parallel for schedule(static)
  for (int64_t r = 0; r < A->nrows; r++) {
    double s = 0.;
    const int64_t *c = A->col + r * W;
    const double *v = A->val + r * W;
    for (int q = 0; q < W; q++)
      if (c[q] >= 0) s += v[q] * x[c[q]];
    y[r] = s;
  }
}

int64_t vfo_ell_to_csr(const vfo_ell *A, int64_t *rowptr, int64_t *colidx, double *vals)
{
  int64_t nnz = 0;
  for (int64_t r = 0; r < A->nrows; r++) {
    if (rowptr) rowptr[r] = nnz;
    for (int q = 0; q < A->width; q++) {
      int64_t c = A->col[r * A->width + q];
      if (c < 0) continue;
      if (colidx) colidx[nnz] = c;
      if (vals) vals[nnz] = A->val[r * A->width + q];
      nnz++;
    }
  }
  if (rowptr) rowptr[A->nrows] = nnz;
  return nnz;
}

/* DMCreateMatrix pattern: box stencil of width 1 */
static vfo_ell *pattern(const vfo_problem *P, int dof)
{
  int64_t nn = vfo_nnodes(P);
  int W = 27 * dof;
  vfo_ell *A = vfo_ell_alloc(nn * dof, W);
  for (int k = 0; k < P->nz; k++)
    for (int j = 0; j < P->ny; j++)
      for (int i = 0; i < P->nx; i++)
        for (int c = 0; c < dof; c++) {
          int64_t r = dof * NODE(P, i, j, k) + c;
          for (int dk = -1; dk <= 1; dk++)
            for (int dj = -1; dj <= 1; dj++)
              for (int di = -1; di <= 1; di++)
                for (int c2 = 0; c2 < dof; c2++) {
                  int s = (((dk + 1) * 3 + (dj + 1)) * 3 + (di + 1)) * dof + c2;
                  int ii = i + di, jj = j + dj, kk = k + dk;
                  int ok = ii >= 0 && ii < P->nx && jj >= 0 && jj < P->ny && kk >= 0 && kk < P->nz;
                  A->col[r * W + s] = ok ? dof * NODE(P, ii, jj, kk) + c2 : -1;
                }
        }
  return A;
}

/* MatSetValuesStencil(ADD) of one element block into the ELL matrix */
static void add_block(vfo_ell *A, const vfo_problem *P, int dof, int ei, int ej, int ek,
                      const double *Kl)
{
  int W = A->width, n = 8 * dof;
  for (int r = 0; r < n; r++) {
    int a = r / dof, c1 = r % dof;
    int ia = a & 1, ja = (a >> 1) & 1, ka = a >> 2;
    int64_t row = dof * NODE(P, ei + ia, ej + ja, ek + ka) + c1;
    for (int q = 0; q < n; q++) {
      int b = q / dof, c2 = q % dof;
      int ib = b & 1, jb = (b >> 1) & 1, kb = b >> 2;
      int s = (((kb - ka + 1) * 3 + (jb - ja + 1)) * 3 + (ib - ia + 1)) * dof + c2;
      A->val[row * W + s] += Kl[r * n + q];
    }
  }
}

/* MatZeroRowsStencil(diag=1) and MatZeroRowsColumnsStencil(diag=1) */
static void zero_rows(vfo_ell *A, const uint8_t *mask, int cols_too)
{
  int W = A->width;
  for (int64_t r = 0; r < A->nrows; r++) {
    if (mask[r]) {
      for (int q = 0; q < W; q++) {
        int64_t c = A->col[r * W + q];
        A->val[r * W + q] = (c == r) ? 1. : 0.;
      }
    } else if (cols_too) {
      for (int q = 0; q < W; q++) {
        int64_t c = A->col[r * W + q];
        if (c >= 0 && mask[c]) A->val[r * W + q] = 0.;
      }
    }
  }
}

/* ------------------------------------------------------- global operators */

typedef struct {
  double u24[24], v8[8], dth8[8], p8[8];
} elemdata;

static void gather(const vfo_problem *P, const vfo_fields *F, const double *U, int ei, int ej,
                   int ek, elemdata *d)
{
  for (int a = 0; a < 8; a++) {
    int64_t n = NODE(P, ei + (a & 1), ej + ((a >> 1) & 1), ek + (a >> 2));
    if (U)
      for (int c = 0; c < 3; c++) d->u24[3 * a + c] = U[3 * n + c];
    if (F) {
      d->v8[a] = F->V ? F->V[n] : 1.;
      d->dth8[a] = (F->theta && F->thetaRef) ? F->theta[n] - F->thetaRef[n] : 0.;
      d->p8[a] = F->pressure ? F->pressure[n] : 0.;
    }
  }
}

#define ELEM_LOOP(P, X, Y, Z)                                                     \
  for (int ek = 0; ek < (P)->nz - 1; ek++)                                        \
    for (int ej = 0; ej < (P)->ny - 1; ej++)                                      \
      for (int ei = 0; ei < (P)->nx - 1; ei++)

#define ELEM_SETUP(P, X, Y, Z, e, cell)                                           \
  int64_t cell = ei + (int64_t)((P)->nx - 1) * (ej + (int64_t)((P)->ny - 1) * ek); \
  vfo_elem3d_init(&(e), X[ei + 1] - X[ei], Y[ej + 1] - Y[ej], Z[ek + 1] - Z[ek]);

typedef struct { double *X, *Y, *Z; } coords_t;
static coords_t mkcoords(const vfo_problem *P)
{
  coords_t c = {malloc(sizeof(double) * P->nx), malloc(sizeof(double) * P->ny),
                malloc(sizeof(double) * P->nz)};
  vfo_coords(P, c.X, c.Y, c.Z);
  return c;
}
static void freecoords(coords_t c) { free(c.X); free(c.Y); free(c.Z); }

/* element stiffness at the state u24 (ignored for UNILATERAL_NONE) */
static void u_elem_matrix(const vfo_problem *P, const elemdata *d, double lambda, double mu,
                          const vfo_elem *e, double *Ke)
{
  if (P->unilateral)
    vfo_bilinear_u_nocomp(Ke, d->u24, d->v8, lambda, mu, P->eta, e);
  else
    vfo_bilinear_u(Ke, d->v8, lambda, mu, P->eta, e);
}

/* VF_UIJacobian (VFMech.c:1897-2029) + MatApplyDirichletBC (VFCartFE.c:1147-1296) */
vfo_ell *vfo_u_jacobian(const vfo_problem *P, const vfo_fields *F)
{
  return vfo_u_jacobian_at(P, F, NULL);
}

vfo_ell *vfo_u_jacobian_at(const vfo_problem *P, const vfo_fields *F, const double *U)
{
  vfo_ell *A = pattern(P, 3);
  coords_t C = mkcoords(P);
  vfo_elem e;
  elemdata d;
  memset(&d, 0, sizeof d);
  double Ke[576];
  ELEM_LOOP(P, C.X, C.Y, C.Z)
  {
    ELEM_SETUP(P, C.X, C.Y, C.Z, e, cell);
    double lambda, mu;
    elem_lame(P, cell, &lambda, &mu);
    gather(P, F, U, ei, ej, ek, &d);
    u_elem_matrix(P, &d, lambda, mu, &e, Ke);
    add_block(A, P, 3, ei, ej, ek, Ke);
  }
  uint8_t *mask = malloc(3 * vfo_nnodes(P));
  vfo_bc_mask(P, P->bcU, 3, mask);
  zero_rows(A, mask, 0);
  free(mask);
  freecoords(C);
  return A;
}

/* element RHS b_e = thermo-poro (+ crack pressure) for element (ei,ej,ek) */
static void u_elem_rhs(const vfo_problem *P, const elemdata *d, double lambda, double mu,
                       const vfo_elem *e, double *b24)
{
  for (int l = 0; l < 24; l++) b24[l] = 0.;
  if (P->unilateral)
    vfo_grad_u_thermoporo_nocomp(b24, d->u24, d->v8, d->dth8, d->p8, lambda, mu, P->alpha, P->beta,
                                 P->eta, e);
  else
    vfo_grad_u_thermoporo(b24, d->v8, d->dth8, d->p8, lambda, mu, P->alpha, P->beta, P->eta, e);
  if (P->hasCrackPressure) vfo_grad_u_crackpressure(b24, d->v8, d->p8, e);
}

/* VF_UResidual: VFMech.c:1456-1887 (in-situ branch out of scope) */
void vfo_u_residual(const vfo_problem *P, const vfo_fields *F, const double *U,
                    const double *BCU, double *res)
{
  int64_t nd = 3 * vfo_nnodes(P);
  for (int64_t q = 0; q < nd; q++) res[q] = 0.;
  coords_t C = mkcoords(P);
  vfo_elem e;
  elemdata d;
  double Ke[576], b24[24];
  ELEM_LOOP(P, C.X, C.Y, C.Z)
  {
    ELEM_SETUP(P, C.X, C.Y, C.Z, e, cell);
    double lambda, mu;
    elem_lame(P, cell, &lambda, &mu);
    gather(P, F, U, ei, ej, ek, &d);
    u_elem_matrix(P, &d, lambda, mu, &e, Ke);
    for (int r = 0, l = 0; r < 24; r++) {
      int a = r / 3;
      int64_t row = 3 * NODE(P, ei + (a & 1), ej + ((a >> 1) & 1), ek + (a >> 2)) + r % 3;
      for (int q = 0; q < 24; q++, l++) res[row] += Ke[l] * d.u24[q];
    }
    u_elem_rhs(P, &d, lambda, mu, &e, b24);
    for (int r = 0; r < 24; r++) {
      int a = r / 3;
      int64_t row = 3 * NODE(P, ei + (a & 1), ej + ((a >> 1) & 1), ek + (a >> 2)) + r % 3;
      res[row] -= b24[r];
    }
  }
  vfo_residual_apply_dirichlet(P, P->bcU, 3, res, U, BCU);
  freecoords(C);
}

void vfo_u_rhs(const vfo_problem *P, const vfo_fields *F, double *b)
{
  int64_t nd = 3 * vfo_nnodes(P);
  for (int64_t q = 0; q < nd; q++) b[q] = 0.;
  coords_t C = mkcoords(P);
  vfo_elem e;
  elemdata d;
  memset(&d, 0, sizeof d);  /* the load at U = 0 (NOCOMPRESSION: every point tensile) */
  double b24[24];
  ELEM_LOOP(P, C.X, C.Y, C.Z)
  {
    ELEM_SETUP(P, C.X, C.Y, C.Z, e, cell);
    double lambda, mu;
    elem_lame(P, cell, &lambda, &mu);
    gather(P, F, NULL, ei, ej, ek, &d);
    u_elem_rhs(P, &d, lambda, mu, &e, b24);
    for (int r = 0; r < 24; r++) {
      int a = r / 3;
      b[3 * NODE(P, ei + (a & 1), ej + ((a >> 1) & 1), ek + (a >> 2)) + r % 3] += b24[r];
    }
  }
  freecoords(C);
}

/* V element matrix: AT part + elastic coupling (VFMech.c:2787-2797) */
static void v_elem_matrix(const vfo_problem *P, const elemdata *d, double lambda, double mu,
                          const vfo_elem *e, double *K64)
{
  for (int l = 0; l < 64; l++) K64[l] = 0.;
  vfo_bilinear_v_at(K64, P->atnum, P->Gc, vfo_atcv(P->atnum), P->epsilon, e);
  if (P->unilateral)
    vfo_bilinear_v_coupling_nocomp(K64, d->u24, d->dth8, lambda, mu, P->alpha, e);
  else
    vfo_bilinear_v_coupling(K64, d->u24, d->dth8, lambda, mu, P->alpha, e);
}

static uint8_t *v_fixed_mask(const vfo_problem *P, const vfo_fields *F)
{
  int64_t nn = vfo_nnodes(P);
  uint8_t *mask = malloc(nn);
  vfo_bc_mask(P, &P->bcV, 1, mask);
  if (P->atnum == 2 && F->VIrrev)
    for (int64_t n = 0; n < nn; n++)
      if (F->VIrrev[n] <= P->irrevtol) mask[n] = 1;
  return mask;
}

/* VF_VIJacobian (VFMech.c:2883-3029) + MatApplyDirichletBCRowCol
 * (VFCartFE.c:1305-1454) + VF_IrrevApplyEQMat (VFMech.c:2637-2682) */
vfo_ell *vfo_v_jacobian(const vfo_problem *P, const vfo_fields *F, const double *U)
{
  vfo_ell *A = pattern(P, 1);
  coords_t C = mkcoords(P);
  vfo_elem e;
  elemdata d;
  double K64[64];
  ELEM_LOOP(P, C.X, C.Y, C.Z)
  {
    ELEM_SETUP(P, C.X, C.Y, C.Z, e, cell);
    double lambda, mu;
    elem_lame(P, cell, &lambda, &mu);
    gather(P, F, U, ei, ej, ek, &d);
    v_elem_matrix(P, &d, lambda, mu, &e, K64);
    add_block(A, P, 1, ei, ej, ek, K64);
  }
  uint8_t *mask = v_fixed_mask(P, F);
  zero_rows(A, mask, 1);
  free(mask);
  freecoords(C);
  return A;
}

static void v_elem_rhs(const vfo_problem *P, const elemdata *d, const vfo_elem *e, double *b8)
{
  for (int a = 0; a < 8; a++) b8[a] = 0.;
  vfo_residual_v_at(b8, P->atnum, P->Gc, vfo_atcv(P->atnum), P->epsilon, e);
  if (P->hasCrackPressure) vfo_residual_v_crackpressure(b8, d->u24, d->p8, e);
}

/* VF_VResidual: VFMech.c:2687-2879.  The thermo-poro term of the reference
 * (VFMech.c:2796) is overwritten before use (VFMech.c:2809-2818), so it does
 * not appear here. */
void vfo_v_residual(const vfo_problem *P, const vfo_fields *F, const double *U, const double *V,
                    double *res)
{
  int64_t nn = vfo_nnodes(P);
  for (int64_t q = 0; q < nn; q++) res[q] = 0.;
  coords_t C = mkcoords(P);
  vfo_elem e;
  elemdata d;
  double K64[64], rl[8];
  ELEM_LOOP(P, C.X, C.Y, C.Z)
  {
    ELEM_SETUP(P, C.X, C.Y, C.Z, e, cell);
    double lambda, mu;
    elem_lame(P, cell, &lambda, &mu);
    gather(P, F, U, ei, ej, ek, &d);
    v_elem_matrix(P, &d, lambda, mu, &e, K64);
    double v8[8];
    for (int a = 0; a < 8; a++)
      v8[a] = V[NODE(P, ei + (a & 1), ej + ((a >> 1) & 1), ek + (a >> 2))];
    for (int m = 0; m < 8; m++) {
      rl[m] = 0.;
      for (int l = 0; l < 8; l++) rl[m] -= K64[m * 8 + l] * v8[l];
    }
    /* VF_ResidualVAT*_local and the crack-pressure term accumulate onto rl */
    vfo_residual_v_at(rl, P->atnum, P->Gc, vfo_atcv(P->atnum), P->epsilon, &e);
    if (P->hasCrackPressure) vfo_residual_v_crackpressure(rl, d.u24, d.p8, &e);
    for (int a = 0; a < 8; a++)
      res[NODE(P, ei + (a & 1), ej + ((a >> 1) & 1), ek + (a >> 2))] -= rl[a];
  }
  if (P->atnum == 2 && F->VIrrev) /* VF_IrrevApplyEQVec: VFMech.c:2597-2628 */
    for (int64_t n = 0; n < nn; n++)
      if (F->VIrrev[n] <= P->irrevtol) res[n] = 0.;
  vfo_residual_apply_dirichlet(P, &P->bcV, 1, res, V, V);
  freecoords(C);
}

void vfo_v_rhs(const vfo_problem *P, const vfo_fields *F, const double *U, double *b)
{
  int64_t nn = vfo_nnodes(P);
  for (int64_t q = 0; q < nn; q++) b[q] = 0.;
  coords_t C = mkcoords(P);
  vfo_elem e;
  elemdata d;
  double b8[8];
  ELEM_LOOP(P, C.X, C.Y, C.Z)
  {
    ELEM_SETUP(P, C.X, C.Y, C.Z, e, cell);
    (void)cell;
    gather(P, F, U, ei, ej, ek, &d);
    v_elem_rhs(P, &d, &e, b8);
    for (int a = 0; a < 8; a++) b[NODE(P, ei + (a & 1), ej + ((a >> 1) & 1), ek + (a >> 2))] += b8[a];
  }
  freecoords(C);
}

/* VF_UEnergy3D: VFMech.c:1111-1427 (UNILATERAL_NONE, no in-situ) */
void vfo_u_energy(const vfo_problem *P, const vfo_fields *F, const double *U, double *elastic,
                  double *pressureWork)
{
  coords_t C = mkcoords(P);
  vfo_elem e;
  elemdata d;
  double myE = 0., myP = 0.;
  ELEM_LOOP(P, C.X, C.Y, C.Z)
  {
    ELEM_SETUP(P, C.X, C.Y, C.Z, e, cell);
    double lambda, mu;
    elem_lame(P, cell, &lambda, &mu);
    gather(P, F, U, ei, ej, ek, &d);
    myE += (P->unilateral ? vfo_elastic_energy_nocomp_local : vfo_elastic_energy_local)(
        d.u24, d.v8, d.dth8, d.p8, lambda, mu, P->alpha, P->beta, P->eta, &e);
    if (P->hasCrackPressure) pressure_work_acc(&myP, d.u24, d.v8, d.p8, &e);
  }
  *elastic = myE;
  *pressureWork = myP;
  freecoords(C);
}

/* VF_VEnergy3D: VFMech.c:2474-2526 */
double vfo_v_energy(const vfo_problem *P, const double *V)
{
  coords_t C = mkcoords(P);
  vfo_elem e;
  double acc = 0.;
  ELEM_LOOP(P, C.X, C.Y, C.Z)
  {
    ELEM_SETUP(P, C.X, C.Y, C.Z, e, cell);
    (void)cell;
    double v8[8];
    for (int a = 0; a < 8; a++) v8[a] = V[NODE(P, ei + (a & 1), ej + ((a >> 1) & 1), ek + (a >> 2))];
    surface_energy_acc(&acc, v8, P->atnum, P->Gc, vfo_atcv(P->atnum), P->epsilon, &e);
  }
  freecoords(C);
  return acc;
}

/* VolumetricCrackOpening: VFPermfield.c:75-150 */
double vfo_crack_volume(const vfo_problem *P, const double *U, const double *V, double *cellfield)
{
  coords_t C = mkcoords(P);
  vfo_elem e;
  elemdata d;
  double total = 0.;
  ELEM_LOOP(P, C.X, C.Y, C.Z)
  {
    ELEM_SETUP(P, C.X, C.Y, C.Z, e, cell);
    gather(P, NULL, U, ei, ej, ek, &d);
    double v8[8], avg;
    for (int a = 0; a < 8; a++) v8[a] = V[NODE(P, ei + (a & 1), ej + ((a >> 1) & 1), ek + (a >> 2))];
    total += vfo_crack_volume_local(d.u24, v8, &e, &avg);
    if (cellfield) cellfield[cell] = avg;
  }
  freecoords(C);
  return total;
}

/* element terms of VF_UEnergy3D (VFMech.c:1111-1427) and VF_VEnergy3D
 * (VFMech.c:2474-2526), one triple per cell, all three from the V argument */
void vfo_energy_fields(const vfo_problem *P, const vfo_fields *F, const double *U, const double *V,
                       double *cell3)
{
  coords_t C = mkcoords(P);
  vfo_elem e;
  elemdata d;
  ELEM_LOOP(P, C.X, C.Y, C.Z)
  {
    ELEM_SETUP(P, C.X, C.Y, C.Z, e, cell);
    double lambda, mu, pw = 0., se = 0., v8[8];
    elem_lame(P, cell, &lambda, &mu);
    gather(P, F, U, ei, ej, ek, &d);
    for (int a = 0; a < 8; a++) v8[a] = V[NODE(P, ei + (a & 1), ej + ((a >> 1) & 1), ek + (a >> 2))];
    cell3[3 * cell] = (P->unilateral ? vfo_elastic_energy_nocomp_local : vfo_elastic_energy_local)(
        d.u24, v8, d.dth8, d.p8, lambda, mu, P->alpha, P->beta, P->eta, &e);
    if (P->hasCrackPressure) pressure_work_acc(&pw, d.u24, v8, d.p8, &e);
    cell3[3 * cell + 1] = pw;
    surface_energy_acc(&se, v8, P->atnum, P->Gc, vfo_atcv(P->atnum), P->epsilon, &e);
    cell3[3 * cell + 2] = se;
  }
  freecoords(C);
}

/* VFPermfield.c:1156-1327: scatter cell * volume and volume to the 8 nodes of
 * every cell (in cell order), then divide node by node */
void vfo_cell_to_node(const vfo_problem *P, const double *cell, double *node, int dof)
{
  coords_t C = mkcoords(P);
  const int64_t nn = vfo_nnodes(P);
  double *vol = (double *)calloc((size_t)nn, sizeof(double));
  memset(node, 0, sizeof(double) * (size_t)dof * (size_t)nn);
  for (int ek = 0; ek < P->nz - 1; ek++)
    for (int ej = 0; ej < P->ny - 1; ej++)
      for (int ei = 0; ei < P->nx - 1; ei++) {
        const double hx = C.X[ei + 1] - C.X[ei], hy = C.Y[ej + 1] - C.Y[ej], hz = C.Z[ek + 1] - C.Z[ek];
        const int64_t c = (int64_t)ei + (int64_t)(P->nx - 1) * ((int64_t)ej + (int64_t)(P->ny - 1) * ek);
        for (int k = 0; k < 2; k++)
          for (int j = 0; j < 2; j++)
            for (int i = 0; i < 2; i++) {
              const int64_t n = NODE(P, ei + i, ej + j, ek + k);
              vol[n] += hz * hy * hx;
              for (int d = 0; d < dof; d++) node[dof * n + d] += cell[dof * c + d] * hx * hy * hz;
            }
      }
  for (int64_t n = 0; n < nn; n++)
    for (int d = 0; d < dof; d++) node[dof * n + d] = node[dof * n + d] / vol[n];
  free(vol);
  freecoords(C);
}

/* ----------------------------------------------------------------- solvers */

static double dot(const double *a, const double *b, int64_t n)
{
  double s = 0.;
  for (int64_t i = 0; i < n; i++) s += a[i] * b[i];
  return s;
}

/* Jacobi-preconditioned CG (KSPCG semantics, unpreconditioned-norm test):
 * converged when ||r|| <= max(rtol*||b||, atol). */
int vfo_pcg(const vfo_ell *A, const double *b, double *x, double rtol, double atol, int maxit,
            double *relres)
{
  int64_t n = A->nrows;
  double *r = malloc(sizeof(double) * n), *z = malloc(sizeof(double) * n),
         *p = malloc(sizeof(double) * n), *q = malloc(sizeof(double) * n),
         *dinv = malloc(sizeof(double) * n);
  for (int64_t i = 0; i < n; i++) {
    double dg = 0.;
    for (int s = 0; s < A->width; s++)
      if (A->col[i * A->width + s] == i) dg = A->val[i * A->width + s];
    dinv[i] = dg != 0. ? 1. / dg : 1.;
  }
  vfo_ell_matvec(A, x, q);
  for (int64_t i = 0; i < n; i++) r[i] = b[i] - q[i];
  double bn = sqrt(dot(b, b, n));
  if (bn == 0.) bn = 1.;
  double rn = sqrt(dot(r, r, n));
  int it = 0;
  double rz = 0.;
  for (int64_t i = 0; i < n; i++) z[i] = dinv[i] * r[i];
  rz = dot(r, z, n);
  for (int64_t i = 0; i < n; i++) p[i] = z[i];
  while (rn > rtol * bn && rn > atol && it < maxit) {
    vfo_ell_matvec(A, p, q);
    double pq = dot(p, q, n);
    if (pq <= 0.) { it = -1 - it; break; }
    double alpha = rz / pq;
    for (int64_t i = 0; i < n; i++) { x[i] += alpha * p[i]; r[i] -= alpha * q[i]; }
    for (int64_t i = 0; i < n; i++) z[i] = dinv[i] * r[i];
    double rz2 = dot(r, z, n);
    double beta = rz2 / rz;
    rz = rz2;
    for (int64_t i = 0; i < n; i++) p[i] = z[i] + beta * p[i];
    rn = sqrt(dot(r, r, n));
    it++;
  }
  if (relres) *relres = rn / bn;
  free(r); free(z); free(p); free(q); free(dinv);
  return it;
}

/* One Newton step of SNESNEWTONLS on the (linear) u-problem:
 * J d = -F(U), J = K with Dirichlet rows replaced by identity.  The
 * Dirichlet part of d is -F_D; the interior part solves K_II d_I =
 * -F_I - K_ID d_D (the identity rows decouple). */
static double norm2(const double *a, int64_t n)
{
  double s = 0.;
  for (int64_t i = 0; i < n; i++) s += a[i] * a[i];
  return sqrt(s);
}

/* UNILATERAL_NOCOMPRESSION: SNESNEWTONLS on F(U) = K(U) U - b(U) (VFMech.c
 * :1456-1887) with the Jacobian K(U) of VF_UIJacobian (VFMech.c:1897-2029);
 * each step solved by the linear step below, then halved until ||F||
 * decreases (the role of the line search). */
static int u_solve_linear(const vfo_problem *P, const vfo_fields *F, double *U, const double *BCU,
                          double rtol, int maxit, double *relres);
static int u_solve_newton(const vfo_problem *P, const vfo_fields *F, double *U, const double *BCU,
                          double rtol, int maxit, double *relres)
{
  int64_t nd = 3 * vfo_nnodes(P);
  double *Fr = malloc(sizeof(double) * nd), *U0 = malloc(sizeof(double) * nd),
         *D = malloc(sizeof(double) * nd);
  vfo_u_residual(P, F, U, BCU, Fr);
  const double rn0 = norm2(Fr, nd);
  double rn = rn0, rr;
  int its = 0;
  for (int newton = 0; newton < 50 && rn > rtol * rn0 && rn > 0.; newton++) {
    memcpy(U0, U, sizeof(double) * nd);
    its += u_solve_linear(P, F, U, BCU, 1e-13, maxit, &rr);  /* U = U0 + D */
    for (int64_t i = 0; i < nd; i++) D[i] = U[i] - U0[i];
    double t = 1.;
    for (int ls = 0; ls < 30; ls++) {
      for (int64_t i = 0; i < nd; i++) U[i] = U0[i] + t * D[i];
      vfo_u_residual(P, F, U, BCU, Fr);
      double rt = norm2(Fr, nd);
      if (rt < rn || ls == 29) {
        rn = rt;
        break;
      }
      t *= .5;
    }
  }
  if (relres) *relres = rn0 > 0. ? rn / rn0 : 0.;
  free(Fr);
  free(U0);
  free(D);
  return its;
}

int vfo_u_solve(const vfo_problem *P, const vfo_fields *F, double *U, const double *BCU,
                double rtol, int maxit, double *relres)
{
  if (P->unilateral) return u_solve_newton(P, F, U, BCU, rtol, maxit, relres);
  return u_solve_linear(P, F, U, BCU, rtol, maxit, relres);
}

/* one Newton step: J d = -F(U) with J = K(U) (identity Dirichlet rows) */
static int u_solve_linear(const vfo_problem *P, const vfo_fields *F, double *U, const double *BCU,
                          double rtol, int maxit, double *relres)
{
  int64_t nd = 3 * vfo_nnodes(P);
  vfo_ell *J = vfo_u_jacobian_at(P, F, U);
  double *Fr = malloc(sizeof(double) * nd), *d = calloc(nd, sizeof(double)),
         *rhs = malloc(sizeof(double) * nd), *t = malloc(sizeof(double) * nd);
  uint8_t *mask = malloc(nd);
  vfo_bc_mask(P, P->bcU, 3, mask);
  vfo_u_residual(P, F, U, BCU, Fr);
  for (int64_t i = 0; i < nd; i++) d[i] = mask[i] ? -Fr[i] : 0.;
  vfo_ell_matvec(J, d, t);
  for (int64_t i = 0; i < nd; i++) rhs[i] = mask[i] ? 0. : -Fr[i] - t[i];
  double *di = calloc(nd, sizeof(double));
  int it = vfo_pcg(J, rhs, di, rtol, 0., maxit, relres);
  for (int64_t i = 0; i < nd; i++) U[i] += mask[i] ? d[i] : di[i];
  free(Fr); free(d); free(rhs); free(t); free(mask); free(di);
  vfo_ell_free(J);
  return it;
}

/* f(V) = 1/2 V.K V - b.V over the free set (fixed entries held) */
static double qp_obj(const vfo_ell *K, const double *b, const double *V, double *tmp, int64_t n)
{
  vfo_ell_matvec(K, V, tmp);
  double f = 0.;
  for (int64_t i = 0; i < n; i++) f += .5 * V[i] * tmp[i] - b[i] * V[i];
  return f;
}

/* Bound-constrained V-solve with the semantics of SNESVINEWTONRSLS on
 * VF_VResidual/VF_VIJacobian (VFCommon.c:1052-1086, 1149-1154): minimise the
 * quadratic V-energy over lb <= V <= ub with Dirichlet and (AT2) irreversible
 * nodes held.  Reduced-space active-set Newton: the active set holds nodes at
 * a bound whose gradient pushes outward; the Newton step solves the reduced
 * system on the inactive set by CG; the projected step is accepted under an
 * Armijo test on the (strictly convex) objective.  Returns outer iterations
 * or a negative value if maxit was reached. */
int vfo_v_solve(const vfo_problem *P, const vfo_fields *F, const double *U, double *V,
                const double *ub, double tol, int maxit)
{
  int64_t n = vfo_nnodes(P);
  uint8_t *fixed = v_fixed_mask(P, F);
  /* Dirichlet values (VecApplyDirichletBC on V); irreversible nodes keep V */
  vfo_vec_apply_dirichlet(P, &P->bcV, 1, V, V);
  /* full (unmasked) K and b */
  vfo_problem Q = *P;
  vfo_bc_init(&Q.bcV);
  vfo_fields F2 = *F;
  F2.VIrrev = NULL;
  vfo_ell *K = vfo_v_jacobian(&Q, &F2, U);
  double *b = malloc(sizeof(double) * n), *g = malloc(sizeof(double) * n),
         *rhs = malloc(sizeof(double) * n), *dlt = malloc(sizeof(double) * n),
         *Vn = malloc(sizeof(double) * n), *tmp = malloc(sizeof(double) * n);
  uint8_t *act = malloc(n);
  vfo_v_rhs(P, F, U, b);
  for (int64_t i = 0; i < n; i++)
    if (!fixed[i]) {
      double hi = ub ? ub[i] : 1.;
      if (V[i] < 0.) V[i] = 0.;
      if (V[i] > hi) V[i] = hi;
    }
  double g0 = -1.;
  int it;
  for (it = 0; it < maxit; it++) {
    vfo_ell_matvec(K, V, g);
    double pg2 = 0.;
    for (int64_t i = 0; i < n; i++) {
      g[i] -= b[i];
      double hi = ub ? ub[i] : 1.;
      act[i] = fixed[i] || (V[i] <= 0. && g[i] > 0.) || (V[i] >= hi && g[i] < 0.);
      if (!act[i]) pg2 += g[i] * g[i];
    }
    double pg = sqrt(pg2);
    if (g0 < 0.) g0 = pg > 0. ? pg : 1.;
    if (pg <= tol * g0 || pg <= 1e-14) break;
    /* reduced Newton system on the inactive set */
    vfo_ell *R = vfo_ell_alloc(n, K->width);
    memcpy(R->col, K->col, sizeof(int64_t) * n * K->width);
    memcpy(R->val, K->val, sizeof(double) * n * K->width);
    zero_rows(R, act, 1);
    for (int64_t i = 0; i < n; i++) { rhs[i] = act[i] ? 0. : -g[i]; dlt[i] = 0.; }
    vfo_pcg(R, rhs, dlt, 1e-12, 0., 20000, NULL);
    vfo_ell_free(R);
    double f0 = qp_obj(K, b, V, tmp, n), lam = 1.;
    for (int ls = 0; ls < 60; ls++) {
      for (int64_t i = 0; i < n; i++) {
        double hi = ub ? ub[i] : 1.;
        double x = V[i] + (act[i] ? 0. : lam * dlt[i]);
        if (!fixed[i]) { if (x < 0.) x = 0.; if (x > hi) x = hi; }
        Vn[i] = x;
      }
      double gd = 0.;
      for (int64_t i = 0; i < n; i++) gd += g[i] * (Vn[i] - V[i]);
      double f1 = qp_obj(K, b, Vn, tmp, n);
      if (f1 <= f0 + 1e-4 * gd || ls == 59) break;
      lam *= .5;
    }
    memcpy(V, Vn, sizeof(double) * n);
  }
  free(b); free(g); free(rhs); free(dlt); free(Vn); free(tmp); free(act); free(fixed);
  vfo_ell_free(K);
  return it < maxit ? it : -it;
}

/* threads the OpenMP loops of the oracle use (for the CPU-baseline report) */
#include <omp.h>
int vfo_num_threads(void) { return omp_get_max_threads(); }
void vfo_set_num_threads(int n) { omp_set_num_threads(n > 0 ? n : 1); }
