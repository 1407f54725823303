/*
 * afem_oracle.c -- CPU FP64 oracle for the P1 assembly + PCG path.
 *
 * TEST INFRASTRUCTURE, NOT PRODUCT CODE.  See afem_oracle.h for the contract
 * and the reference file:line of each function.
 */
#define _POSIX_C_SOURCE 199309L
#include "afem_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif
#include <time.h>

/* ---- small vector helpers (Arcane Real3 semantics) ------------------------ */

static void sub3(const double* a, const double* b, double* r)
{
  r[0] = a[0] - b[0];
  r[1] = a[1] - b[1];
  r[2] = a[2] - b[2];
}

/* math::cross */
static void cross3(const double* a, const double* b, double* r)
{
  r[0] = a[1] * b[2] - a[2] * b[1];
  r[1] = a[2] * b[0] - a[0] * b[2];
  r[2] = a[0] * b[1] - a[1] * b[0];
}

static double dot3(const double* a, const double* b)
{
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

/* ---- geometry -------------------------------------------------------------- */

/* femutils/ArcaneFemFunctionsGpu.h:76-86 */
double afo_area_tria3(const double* n0, const double* n1, const double* n2)
{
  double a[3], b[3], v[3];
  sub3(n1, n0, a);
  sub3(n2, n0, b);
  cross3(a, b, v);
  return sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]) / 2.0;
}

/* femutils/ArcaneFemFunctionsGpu.h:113-128 */
double afo_volume_tetra4(const double* n0, const double* n1, const double* n2, const double* n3)
{
  double v0[3], v1[3], v2[3], c[3];
  sub3(n1, n0, v0);
  sub3(n2, n0, v1);
  sub3(n3, n0, v2);
  cross3(v1, v2, c);
  return fabs(dot3(v0, c)) / 6.0;
}

/* femutils/ArcaneFemFunctionsGpu.h:241-281 */
void afo_grad_tria3(const double* X, double dx[3], double dy[3])
{
  const double* n0 = X;
  const double* n1 = X + 3;
  const double* n2 = X + 6;
  double A2 = ((n1[0] - n0[0]) * (n2[1] - n0[1]) - (n2[0] - n0[0]) * (n1[1] - n0[1]));
  dx[0] = (n1[1] - n2[1]) / A2;
  dx[1] = (n2[1] - n0[1]) / A2;
  dx[2] = (n0[1] - n1[1]) / A2;
  dy[0] = (n2[0] - n1[0]) / A2;
  dy[1] = (n0[0] - n2[0]) / A2;
  dy[2] = (n1[0] - n0[0]) / A2;
}

/* femutils/ArcaneFemFunctionsGpu.h:414-535 (x = 0, y = 1, z = 2) */
void afo_grad_tetra4(const double* X, double dx[4], double dy[4], double dz[4])
{
  const double* n0 = X;
  const double* n1 = X + 3;
  const double* n2 = X + 6;
  const double* n3 = X + 9;
  double v0[3], v1[3], v2[3], c[3];
  sub3(n1, n0, v0);
  sub3(n2, n0, v1);
  sub3(n3, n0, v2);
  cross3(v1, v2, c);
  double V6 = fabs(dot3(v0, c));
  enum { X_ = 0, Y_ = 1, Z_ = 2 };
  dx[0] = (n1[Y_] * (n3[Z_] - n2[Z_]) + n2[Y_] * (n1[Z_] - n3[Z_]) + n3[Y_] * (n2[Z_] - n1[Z_])) / V6;
  dx[1] = (n0[Y_] * (n2[Z_] - n3[Z_]) + n2[Y_] * (n3[Z_] - n0[Z_]) + n3[Y_] * (n0[Z_] - n2[Z_])) / V6;
  dx[2] = (n0[Y_] * (n3[Z_] - n1[Z_]) + n1[Y_] * (n0[Z_] - n3[Z_]) + n3[Y_] * (n1[Z_] - n0[Z_])) / V6;
  dx[3] = (n0[Y_] * (n1[Z_] - n2[Z_]) + n1[Y_] * (n2[Z_] - n0[Z_]) + n2[Y_] * (n0[Z_] - n1[Z_])) / V6;

  dy[0] = (n1[Z_] * (n3[X_] - n2[X_]) + n2[Z_] * (n1[X_] - n3[X_]) + n3[Z_] * (n2[X_] - n1[X_])) / V6;
  dy[1] = (n0[Z_] * (n2[X_] - n3[X_]) + n2[Z_] * (n3[X_] - n0[X_]) + n3[Z_] * (n0[X_] - n2[X_])) / V6;
  dy[2] = (n0[Z_] * (n3[X_] - n1[X_]) + n1[Z_] * (n0[X_] - n3[X_]) + n3[Z_] * (n1[X_] - n0[X_])) / V6;
  dy[3] = (n0[Z_] * (n1[X_] - n2[X_]) + n1[Z_] * (n2[X_] - n0[X_]) + n2[Z_] * (n0[X_] - n1[X_])) / V6;

  dz[0] = (n1[X_] * (n3[Y_] - n2[Y_]) + n2[X_] * (n1[Y_] - n3[Y_]) + n3[X_] * (n2[Y_] - n1[Y_])) / V6;
  dz[1] = (n0[X_] * (n2[Y_] - n3[Y_]) + n2[X_] * (n3[Y_] - n0[Y_]) + n3[X_] * (n0[Y_] - n2[Y_])) / V6;
  dz[2] = (n0[X_] * (n3[Y_] - n1[Y_]) + n1[X_] * (n0[Y_] - n3[Y_]) + n3[X_] * (n1[Y_] - n0[Y_])) / V6;
  dz[3] = (n0[X_] * (n1[Y_] - n2[Y_]) + n1[X_] * (n2[Y_] - n0[Y_]) + n2[X_] * (n0[Y_] - n1[Y_])) / V6;
}

/* ---- element matrices ------------------------------------------------------- */

/* modules/poisson/ElementMatrix.h:41-49; modules/fourier/ElementMatrix.h:38-47:
 *   area * lambda * (dxU ^ dxU) + area * lambda * (dyU ^ dyU) */
void afo_elem_laplace_tria3(const double* X, double lambda, double K[9])
{
  double area = afo_area_tria3(X, X + 3, X + 6);
  double dx[3], dy[3];
  afo_grad_tria3(X, dx, dy);
  double s = area * lambda;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      K[i * 3 + j] = s * (dx[i] * dx[j]) + s * (dy[i] * dy[j]);
}

/* modules/poisson/ElementMatrix.h:93-102; modules/fourier/ElementMatrix.h:85-95 */
void afo_elem_laplace_tetra4(const double* X, double lambda, double K[16])
{
  double vol = afo_volume_tetra4(X, X + 3, X + 6, X + 9);
  double dx[4], dy[4], dz[4];
  afo_grad_tetra4(X, dx, dy, dz);
  double s = vol * lambda;
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j)
      K[i * 4 + j] = s * (dx[i] * dx[j]) + s * (dy[i] * dy[j]) + s * (dz[i] * dz[j]);
}

/* Heat element lambda K + M / dt, modules/heat/ElementMatrix.h:28-39 (tri:
 * (1/12.) * massMatrix * area / dt) and :80-91 (tet: (1/20.) * ... * volume / dt);
 * massMatrix (femutils/FemUtils.h:341-355) is 2 on the diagonal, 1 elsewhere. */
void afo_elem_heat_tria3(const double* X, double lambda, double dt, double K[9])
{
  afo_elem_laplace_tria3(X, lambda, K);
  double area = afo_area_tria3(X, X + 3, X + 6);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      K[i * 3 + j] = K[i * 3 + j] + (1.0 / 12.0) * (i == j ? 2.0 : 1.0) * area / dt;
}

void afo_elem_heat_tetra4(const double* X, double lambda, double dt, double K[16])
{
  afo_elem_laplace_tetra4(X, lambda, K);
  double vol = afo_volume_tetra4(X, X + 3, X + 6, X + 9);
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j)
      K[i * 4 + j] = K[i * 4 + j] + (1.0 / 20.0) * (i == j ? 2.0 : 1.0) * vol / dt;
}

/* ---- Q1 elements (Quad4, Hexa8) ------------------------------------------- */

/* 1/sqrt(3): the 2-point Gauss abscissa of every Q1 integral in the reference
 * (modules/poisson/ElementMatrixHexQuad.h:67-69), weights 1. */
#define AFO_GP 0.57735026918962576451

/* Shape functions and reference gradients of femutils/ShapeFunctions.h:104-129
 * (Quad4, nodes counter-clockwise from (-1,-1)) and :291-350 (Hexa8, the
 * bottom face counter-clockwise then the top face). */
static void q1_shape(int dim, double xi, double eta, double zeta, double N[8], double dxi[8], double deta[8],
                     double dzeta[8])
{
  static const double sx[8] = { -1, 1, 1, -1, -1, 1, 1, -1 };
  static const double sy[8] = { -1, -1, 1, 1, -1, -1, 1, 1 };
  static const double sz[8] = { -1, -1, -1, -1, 1, 1, 1, 1 };
  if (dim == 2) {
    for (int a = 0; a < 4; ++a) {
      N[a] = 0.25 * (1.0 + sx[a] * xi) * (1.0 + sy[a] * eta);
      dxi[a] = 0.25 * sx[a] * (1.0 + sy[a] * eta);
      deta[a] = 0.25 * sy[a] * (1.0 + sx[a] * xi);
      dzeta[a] = 0.0;
    }
    return;
  }
  for (int a = 0; a < 8; ++a) {
    N[a] = 0.125 * (1.0 + sx[a] * xi) * (1.0 + sy[a] * eta) * (1.0 + sz[a] * zeta);
    dxi[a] = 0.125 * sx[a] * (1.0 + sy[a] * eta) * (1.0 + sz[a] * zeta);
    deta[a] = 0.125 * sy[a] * (1.0 + sx[a] * xi) * (1.0 + sz[a] * zeta);
    dzeta[a] = 0.125 * sz[a] * (1.0 + sx[a] * xi) * (1.0 + sy[a] * eta);
  }
}

/* Physical gradients and det J at one point: J[i][j] = sum_a dN_a/dxi_i x_a[j],
 * grad N = J^-1 (dN/dxi, dN/deta[, dN/dzeta]) (femutils/ArcaneFemFunctionsGpu.h:
 * 296-350 for Quad4, :553-587 for Hexa8).  Returns det J (signed). */
double afo_q1_gradients(int dim, const double* X, double xi, double eta, double zeta, double N[8], double gx[8],
                        double gy[8], double gz[8])
{
  double dxi[8], deta[8], dzeta[8];
  q1_shape(dim, xi, eta, zeta, N, dxi, deta, dzeta);
  const int n = dim == 2 ? 4 : 8;
  if (dim == 2) {
    double j00 = 0, j01 = 0, j10 = 0, j11 = 0;
    for (int a = 0; a < n; ++a) {
      j00 += dxi[a] * X[3 * a + 0];
      j01 += dxi[a] * X[3 * a + 1];
      j10 += deta[a] * X[3 * a + 0];
      j11 += deta[a] * X[3 * a + 1];
    }
    const double det = j00 * j11 - j01 * j10;
    const double i00 = j11 / det, i01 = -j01 / det, i10 = -j10 / det, i11 = j00 / det;
    for (int a = 0; a < n; ++a) {
      gx[a] = i00 * dxi[a] + i01 * deta[a];
      gy[a] = i10 * dxi[a] + i11 * deta[a];
      gz[a] = 0.0;
    }
    return det;
  }
  double J[3][3] = { { 0 } };
  for (int a = 0; a < n; ++a)
    for (int d = 0; d < 3; ++d) {
      J[0][d] += dxi[a] * X[3 * a + d];
      J[1][d] += deta[a] * X[3 * a + d];
      J[2][d] += dzeta[a] * X[3 * a + d];
    }
  const double det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) -
                     J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                     J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
  double I[3][3];
  I[0][0] = (J[1][1] * J[2][2] - J[1][2] * J[2][1]) / det;
  I[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) / det;
  I[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) / det;
  I[1][0] = (J[1][2] * J[2][0] - J[1][0] * J[2][2]) / det;
  I[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) / det;
  I[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) / det;
  I[2][0] = (J[1][0] * J[2][1] - J[1][1] * J[2][0]) / det;
  I[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) / det;
  I[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) / det;
  for (int a = 0; a < n; ++a) {
    gx[a] = I[0][0] * dxi[a] + I[0][1] * deta[a] + I[0][2] * dzeta[a];
    gy[a] = I[1][0] * dxi[a] + I[1][1] * deta[a] + I[1][2] * dzeta[a];
    gz[a] = I[2][0] * dxi[a] + I[2][1] * deta[a] + I[2][2] * dzeta[a];
  }
  return det;
}

/* modules/poisson/ElementMatrixHexQuad.h:33-64 (Quad4), :159-195 (Hexa8):
 * sum over the 2x2(x2) Gauss points of (dxU^dxU + dyU^dyU [+ dzU^dzU]) detJ,
 * times the cell's lambda. */
void afo_elem_laplace_q1(int dim, const double* X, double lambda, double* K)
{
  const int n = dim == 2 ? 4 : 8, ng = dim == 2 ? 1 : 2;
  const double gp[2] = { -AFO_GP, AFO_GP };
  for (int k = 0; k < n * n; ++k)
    K[k] = 0.0;
  for (int ix = 0; ix < 2; ++ix)
    for (int iy = 0; iy < 2; ++iy)
      for (int iz = 0; iz < ng; ++iz) {
        double N[8], gx[8], gy[8], gz[8];
        const double w = afo_q1_gradients(dim, X, gp[ix], gp[iy], dim == 2 ? 0.0 : gp[iz], N, gx, gy, gz);
        for (int a = 0; a < n; ++a)
          for (int b = 0; b < n; ++b)
            K[a * n + b] += (gx[a] * gx[b]) * w + (gy[a] * gy[b]) * w + (gz[a] * gz[b]) * w;
      }
  for (int k = 0; k < n * n; ++k)
    K[k] *= lambda;
}

/* BoundaryConditions2D::applyConstantSourceToRhsQuad4 and
 * BoundaryConditions3D::applyConstantSourceToRhsHexa8
 * (femutils/ArcaneFemFunctionsGpu.cc:393-470, :875-960): every node of the
 * cell receives qdot N_a detJ at each Gauss point. */
void afo_source_q1(int dim, const double* X, double qdot, double* b)
{
  const int n = dim == 2 ? 4 : 8, ng = dim == 2 ? 1 : 2;
  const double gp[2] = { -AFO_GP, AFO_GP };
  for (int a = 0; a < n; ++a)
    b[a] = 0.0;
  for (int ix = 0; ix < 2; ++ix)
    for (int iy = 0; iy < 2; ++iy)
      for (int iz = 0; iz < ng; ++iz) {
        double N[8], gx[8], gy[8], gz[8];
        const double w = afo_q1_gradients(dim, X, gp[ix], gp[iy], dim == 2 ? 0.0 : gp[iz], N, gx, gy, gz);
        for (int a = 0; a < n; ++a)
          b[a] += qdot * N[a] * w;
      }
}

/* modules/heat/ElementMatrixHexQuad.h:33-70 (Quad4), :91-130 (Hexa8): at each
 * Gauss point (dxU^dxU) iw lambda + (dyU^dyU) iw lambda [+ dz] + (N^N) iw / dt
 * with iw = det J (the Q1 consistent mass matrix, no doubled diagonal). */
void afo_elem_heat_q1(int dim, const double* X, double lambda, double dt, double* K)
{
  const int n = dim == 2 ? 4 : 8, ng = dim == 2 ? 1 : 2;
  const double gp[2] = { -AFO_GP, AFO_GP };
  for (int k = 0; k < n * n; ++k)
    K[k] = 0.0;
  for (int ix = 0; ix < 2; ++ix)
    for (int iy = 0; iy < 2; ++iy)
      for (int iz = 0; iz < ng; ++iz) {
        double N[8], gx[8], gy[8], gz[8];
        const double iw = afo_q1_gradients(dim, X, gp[ix], gp[iy], dim == 2 ? 0.0 : gp[iz], N, gx, gy, gz);
        for (int a = 0; a < n; ++a)
          for (int b = 0; b < n; ++b)
            K[a * n + b] += (gx[a] * gx[b]) * iw * lambda + (gy[a] * gy[b]) * iw * lambda +
                            (gz[a] * gz[b]) * iw * lambda + (N[a] * N[b]) * iw / dt;
      }
}

/* modules/elasticity/ElementMatrix.h:41-58 (tri3, DoFs interleaved ux0 uy0 ux1
 * ...): (lambda + 2 mu)(dxUx^dxUx + dyUy^dyUy) area + lambda (dyUy^dxUx +
 * dxUx^dyUy) area + mu ((dxUy + dyUx)^(dyUx + dxUy)) area, with the vectors
 * restated as written there. */
void afo_elem_elasticity_tria3(const double* X, double lam, double mu, double K[36])
{
  double dx[3], dy[3];
  afo_grad_tria3(X, dx, dy);
  const double area = afo_area_tria3(X, X + 3, X + 6);
  double dxUx[6] = { 0 }, dyUx[6] = { 0 }, dxUy[6] = { 0 }, dyUy[6] = { 0 };
  for (int a = 0; a < 3; ++a) {
    dxUx[2 * a] = dx[a];
    dyUx[2 * a] = dy[a];
    dxUy[2 * a + 1] = dx[a];
    dyUy[2 * a + 1] = dy[a];
  }
  for (int i = 0; i < 6; ++i)
    for (int j = 0; j < 6; ++j) {
      const double normal = (lam + 2 * mu) * (dxUx[i] * dxUx[j] + dyUy[i] * dyUy[j]) * area;
      const double comp = lam * (dyUy[i] * dxUx[j] + dxUx[i] * dyUy[j]) * area;
      const double shear = mu * ((dxUy[i] + dyUx[i]) * (dyUx[j] + dxUy[j])) * area;
      K[i * 6 + j] = normal + comp + shear;
    }
}

/* modules/elasticity/ElementMatrixHexQuad.h:50-70 (Quad4, 8 DoFs), :137-187
 * (Hexa8, 24 DoFs), summed over the Gauss points with iw = det J: the
 * displacement-gradient vectors restated as written there (DoFs interleaved
 * per node). */
void afo_elem_elasticity_q1(int dim, const double* X, double lam, double mu, double* K)
{
  const int n = dim == 2 ? 4 : 8, nd = n * dim, ng = dim == 2 ? 1 : 2;
  const double gp[2] = { -AFO_GP, AFO_GP };
  for (int k = 0; k < nd * nd; ++k)
    K[k] = 0.0;
  for (int ix = 0; ix < 2; ++ix)
    for (int iy = 0; iy < 2; ++iy)
      for (int iz = 0; iz < ng; ++iz) {
        double N[8], g[3][8];
        const double iw = afo_q1_gradients(dim, X, gp[ix], gp[iy], dim == 2 ? 0.0 : gp[iz], N, g[0], g[1], g[2]);
        /* D[d][c][i]: derivative along d of displacement component c, DoF i */
        double D[3][3][24];
        memset(D, 0, sizeof D);
        for (int a = 0; a < n; ++a)
          for (int d = 0; d < dim; ++d)
            for (int c = 0; c < dim; ++c)
              D[d][c][a * dim + c] = g[d][a];
        for (int i = 0; i < nd; ++i)
          for (int j = 0; j < nd; ++j) {
            double normal = 0.0, comp = 0.0, shear = 0.0;
            for (int c = 0; c < dim; ++c)
              normal += D[c][c][i] * D[c][c][j];
            for (int c = 0; c < dim; ++c)
              for (int e = 0; e < dim; ++e)
                if (e != c)
                  comp += D[c][c][i] * D[e][e][j];
            for (int c = 0; c < dim; ++c)
              for (int e = c + 1; e < dim; ++e)
                shear += (D[e][c][i] + D[c][e][i]) * (D[e][c][j] + D[c][e][j]);
            K[i * nd + j] += (lam + 2 * mu) * normal * iw + lam * comp * iw + mu * shear * iw;
          }
      }
}

/* modules/elasticity/FemModule.cc:171-172 */
void afo_lame(double E, double nu, double* lame_lambda, double* lame_mu)
{
  *lame_mu = (E / (2 * (1 + nu)));
  *lame_lambda = E * nu / ((1 + nu) * (1 - 2 * nu));
}

/* modules/elasticity/ElementMatrix.h:151-189.  The 12-vectors are the
 * reference's dxUx, dyUx, ... ; '^' is the outer product
 * (femutils/FemUtils.h:258-281) and each sum of matrices is evaluated left to
 * right, then scaled by the Lame coefficient and by the volume. */
void afo_elem_elasticity_tetra4(const double* X, double lam, double mu, double K[144])
{
  double dxu[4], dyu[4], dzu[4];
  afo_grad_tetra4(X, dxu, dyu, dzu);
  double volume = afo_volume_tetra4(X, X + 3, X + 6, X + 9);

  double dxUx[12] = { 0 }, dyUx[12] = { 0 }, dzUx[12] = { 0 };
  double dxUy[12] = { 0 }, dyUy[12] = { 0 }, dzUy[12] = { 0 };
  double dxUz[12] = { 0 }, dyUz[12] = { 0 }, dzUz[12] = { 0 };
  for (int a = 0; a < 4; ++a) {
    dxUx[3 * a] = dxu[a];
    dyUx[3 * a] = dyu[a];
    dzUx[3 * a] = dzu[a];
    dxUy[3 * a + 1] = dxu[a];
    dyUy[3 * a + 1] = dyu[a];
    dzUy[3 * a + 1] = dzu[a];
    dxUz[3 * a + 2] = dxu[a];
    dyUz[3 * a + 2] = dyu[a];
    dzUz[3 * a + 2] = dzu[a];
  }
  double s1[12], s2[12], s3[12], t1[12], t2[12], t3[12];
  for (int p = 0; p < 12; ++p) {
    s1[p] = dxUy[p] + dyUx[p]; /* (dxUy + dyUx) */
    t1[p] = dyUx[p] + dxUy[p]; /* (dyUx + dxUy) */
    s2[p] = dzUy[p] + dyUz[p];
    t2[p] = dyUz[p] + dzUy[p];
    s3[p] = dxUz[p] + dzUx[p];
    t3[p] = dzUx[p] + dxUz[p];
  }
  for (int p = 0; p < 12; ++p) {
    for (int q = 0; q < 12; ++q) {
      double normal = (lam + 2 * mu) * ((dxUx[p] * dxUx[q] + dyUy[p] * dyUy[q]) + dzUz[p] * dzUz[q]);
      normal = normal * volume;
      double comp = dyUy[p] * dxUx[q] + dxUx[p] * dyUy[q];
      comp = comp + dzUz[p] * dxUx[q];
      comp = comp + dxUx[p] * dzUz[q];
      comp = comp + dyUy[p] * dzUz[q];
      comp = comp + dzUz[p] * dyUy[q];
      comp = (lam * comp) * volume;
      double shear = (s1[p] * t1[q] + s2[p] * t2[q]) + s3[p] * t3[q];
      shear = (mu * shear) * volume;
      K[p * 12 + q] = (normal + comp) + shear;
    }
  }
}

/* ---- connectivity -------------------------------------------------------------- */

int afo_node_cells(int32_t nb_node, int32_t nb_cell, int32_t npc, const int32_t* cells,
                   int32_t* offsets, int32_t* items)
{
  const int st = AFO_CELL_STRIDE(npc);
  memset(offsets, 0, sizeof(int32_t) * ((size_t)nb_node + 1));
  for (int64_t c = 0; c < nb_cell; ++c)
    for (int k = 0; k < npc; ++k) {
      int32_t n = cells[c * st + k];
      if (n < 0 || n >= nb_node)
        return -1;
      offsets[n + 1]++;
    }
  for (int32_t i = 0; i < nb_node; ++i)
    offsets[i + 1] += offsets[i];
  int32_t* fill = (int32_t*)malloc(sizeof(int32_t) * (size_t)(nb_node > 0 ? nb_node : 1));
  memcpy(fill, offsets, sizeof(int32_t) * (size_t)nb_node);
  /* cells visited in increasing id order: each list comes out sorted by cell */
  for (int64_t c = 0; c < nb_cell; ++c)
    for (int k = 0; k < npc; ++k) {
      int32_t n = cells[c * st + k];
      items[fill[n]++] = (int32_t)(c * st + k);
    }
  free(fill);
  return 0;
}

/* ---- sparsity ------------------------------------------------------------------- */

static int cmp_i32(const void* a, const void* b)
{
  int32_t x = *(const int32_t*)a, y = *(const int32_t*)b;
  return (x > y) - (x < y);
}

int64_t afo_pattern(int32_t nb_node, int32_t nb_cell, int32_t npc, const int32_t* cells,
                    int32_t* row_index, int32_t* columns, int64_t capacity)
{
  int32_t* off = (int32_t*)malloc(sizeof(int32_t) * ((size_t)nb_node + 1));
  int32_t* items = (int32_t*)malloc(sizeof(int32_t) * ((size_t)nb_cell * npc + 1));
  if (afo_node_cells(nb_node, nb_cell, npc, cells, off, items) != 0) {
    free(off);
    free(items);
    return -1;
  }
  int32_t maxdeg = 0;
  for (int32_t i = 0; i < nb_node; ++i)
    if (off[i + 1] - off[i] > maxdeg)
      maxdeg = off[i + 1] - off[i];
  int32_t* buf = (int32_t*)malloc(sizeof(int32_t) * ((size_t)maxdeg * npc + 1));
  int64_t nnz = 0;
  int overflow = 0;
  row_index[0] = 0;
  for (int32_t i = 0; i < nb_node; ++i) {
    int32_t m = 0;
    for (int32_t e = off[i]; e < off[i + 1]; ++e) {
      int32_t c = items[e] / AFO_CELL_STRIDE(npc);
      for (int k = 0; k < npc; ++k)
        buf[m++] = cells[(int64_t)c * AFO_CELL_STRIDE(npc) + k];
    }
    if (m == 0)
      buf[m++] = i; /* isolated node keeps its diagonal */
    qsort(buf, (size_t)m, sizeof(int32_t), cmp_i32);
    int32_t u = 0;
    for (int32_t k = 0; k < m; ++k)
      if (k == 0 || buf[k] != buf[k - 1])
        buf[u++] = buf[k];
    if (!overflow && nnz + u <= capacity)
      memcpy(columns + nnz, buf, sizeof(int32_t) * (size_t)u);
    else
      overflow = 1;
    nnz += u;
    row_index[i + 1] = (int32_t)nnz;
  }
  free(buf);
  free(off);
  free(items);
  return overflow ? -1 : nnz;
}

/* ---- assembly ------------------------------------------------------------------- */

/* Linear search of femutils/BSRFormat.h:281-289: returns the position x of
 * column col inside row r, or -1. */
static int32_t find_in_row(const int32_t* row_index, const int32_t* columns, int32_t r, int32_t col)
{
  for (int32_t k = row_index[r]; k < row_index[r + 1]; ++k)
    if (columns[k] == col)
      return k - row_index[r];
  return -1;
}

int afo_assemble(int dim, int32_t nb_node, const double* coords,
                 int32_t nb_cell, int32_t npc, const int32_t* cells,
                 const int32_t* row_index, const int32_t* columns,
                 int element, int nb_dof, int order_per_block,
                 const double* cell_lambda, double lame_lambda, double lame_mu,
                 double* values)
{
  const int nb = nb_dof;
  const int nb2 = nb * nb;
  const int64_t nnz = row_index[nb_node];
  const int q1 = AFO_IS_Q1(dim, npc);
  if (element == AFO_ELASTICITY && nb_dof != dim)
    return -2;
  memset(values, 0, sizeof(double) * (size_t)(nnz * nb2));
  double X[24];
  double K[576];
  const int ke = npc * nb; /* element matrix order */
  for (int64_t c = 0; c < nb_cell; ++c) {
    const int32_t* cn = cells + c * AFO_CELL_STRIDE(npc);
    for (int k = 0; k < npc; ++k)
      memcpy(X + 3 * k, coords + 3 * (int64_t)cn[k], 3 * sizeof(double));
    double lambda = cell_lambda ? cell_lambda[c] : 1.0;
    if (q1 && element == AFO_ELASTICITY)
      afo_elem_elasticity_q1(dim, X, lame_lambda, lame_mu, K);
    else if (q1 && element == AFO_HEAT) /* lame_lambda carries dt */
      afo_elem_heat_q1(dim, X, lambda, lame_lambda, K);
    else if (q1)
      afo_elem_laplace_q1(dim, X, lambda, K);
    else if (element == AFO_LAPLACE) {
      if (npc == 3)
        afo_elem_laplace_tria3(X, lambda, K);
      else
        afo_elem_laplace_tetra4(X, lambda, K);
    }
    else if (element == AFO_HEAT) { /* lame_lambda carries dt */
      if (npc == 3)
        afo_elem_heat_tria3(X, lambda, lame_lambda, K);
      else
        afo_elem_heat_tetra4(X, lambda, lame_lambda, K);
    }
    else if (npc == 3) {
      afo_elem_elasticity_tria3(X, lame_lambda, lame_mu, K);
    }
    else {
      afo_elem_elasticity_tetra4(X, lame_lambda, lame_mu, K);
    }
    for (int a = 0; a < npc; ++a) {
      int32_t r = cn[a];
      for (int b = 0; b < npc; ++b) {
        int32_t x = find_in_row(row_index, columns, r, cn[b]);
        if (x < 0)
          return -1;
        int64_t blk = (int64_t)row_index[r] + x;
        int32_t nzr = row_index[r + 1] - row_index[r];
        for (int i = 0; i < nb; ++i)
          for (int j = 0; j < nb; ++j) {
            double v = K[(nb * a + i) * ke + nb * b + j];
            int64_t idx;
            if (order_per_block)
              idx = blk * nb2 + (i * nb + j);
            else
              idx = (int64_t)row_index[r] * nb2 + nb * (x + (int64_t)i * nzr) + j;
            values[idx] += v;
          }
      }
    }
  }
  return 0;
}

/* The same cell-wise Laplace assembly and constant source on `nthreads` host
 * threads: cells split among the threads, every entry added with an atomic
 * update, as the reference's atomic scatter does on its device
 * (femutils/BSRFormat.h:257-401, doAtomic Add).  Timing baseline only; the
 * summation order differs from afo_assemble in the last bits. */
int afo_assemble_laplace_threads(int32_t nb_node, const double* coords, int32_t nb_cell, int32_t npc,
                                 const int32_t* cells, const int32_t* row_index, const int32_t* columns,
                                 double f, double* values, double* rhs, int nthreads)
{
  const int64_t nnz = row_index[nb_node];
  int bad = 0;
#pragma omp parallel num_threads(nthreads)
  {
#pragma omp for schedule(static)
    for (int64_t k = 0; k < nnz; ++k)
      values[k] = 0.0;
#pragma omp for schedule(static)
    for (int64_t k = 0; k < nb_node; ++k)
      rhs[k] = 0.0;
#pragma omp for schedule(static) reduction(| : bad)
    for (int64_t c = 0; c < nb_cell; ++c) {
      const int32_t* cn = cells + c * 4;
      double X[12], K[16];
      for (int k = 0; k < npc; ++k)
        memcpy(X + 3 * k, coords + 3 * (int64_t)cn[k], 3 * sizeof(double));
      double m;
      if (npc == 3) {
        afo_elem_laplace_tria3(X, 1.0, K);
        m = afo_area_tria3(X, X + 3, X + 6);
      }
      else {
        afo_elem_laplace_tetra4(X, 1.0, K);
        m = afo_volume_tetra4(X, X + 3, X + 6, X + 9);
      }
      for (int a = 0; a < npc; ++a) {
        const int32_t r = cn[a];
        for (int b = 0; b < npc; ++b) {
          const int32_t x = find_in_row(row_index, columns, r, cn[b]);
          if (x < 0) {
            bad = 1;
            continue;
          }
#pragma omp atomic
          values[(int64_t)row_index[r] + x] += K[a * npc + b];
        }
#pragma omp atomic
        rhs[r] += f * m / npc;
      }
    }
  }
  return bad ? -1 : 0;
}

int afo_max_threads(void)
{
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/* femutils/BSRFormat.cc:110-172 */
void afo_bsr_to_csr(int32_t nb_node, int nb, const int32_t* row_index, const int32_t* columns,
                    int32_t* csr_row, int32_t* csr_col)
{
  csr_row[0] = 0;
  int64_t off = 1;
  for (int32_t i = 0; i < nb_node; ++i)
    for (int j = 0; j < nb; ++j) {
      int32_t len = (row_index[i + 1] - row_index[i]) * nb;
      csr_row[off] = csr_row[off - 1] + len;
      ++off;
    }
  int64_t o = 0;
  for (int32_t i = 0; i < nb_node; ++i)
    for (int j = 0; j < nb; ++j)
      for (int32_t b = row_index[i]; b < row_index[i + 1]; ++b)
        for (int k = 0; k < nb; ++k)
          csr_col[o++] = columns[b] * nb + k;
}

/* BSRMatrix::findValueIndex (femutils/BSRFormat.cc:79-106) over the
 * reference's row array of nb_row starts (the last row ends at nb_col); for
 * nb_block = 1 this is CsrDoFLinearSystemImpl::indexValue
 * (femutils/internal/CsrDoFLinearSystemImpl.h:46-53).  Returns the value
 * index, or -1 when (row, col) is not in the pattern (where the reference
 * throws or indexes values[-1]). */
int64_t afo_find_value_index(int32_t nb_row, int nb_block, int order_per_block, const int32_t* rows,
                             int64_t nb_col, const int32_t* columns, const int32_t* nnz_per_row, int32_t row,
                             int32_t col)
{
  const int32_t block_row = row / nb_block;
  const int32_t block_col = col / nb_block;
  int64_t block_start = rows[block_row];
  const int64_t block_end = (block_row == nb_row - 1) ? nb_col : rows[block_row + 1];
  const int row_offset = row % nb_block;
  const int col_offset = col % nb_block;
  const int64_t nb_block_sq = (int64_t)nb_block * nb_block;
  const int64_t block_start_in_value = block_start * nb_block_sq;
  int64_t col_index = 0;
  while (block_start < block_end) {
    if (columns[block_start] == block_col) {
      if (!order_per_block)
        return block_start_in_value + (int64_t)nb_block * col_index +
               (int64_t)row_offset * nb_block * nnz_per_row[block_row] + col_offset;
      return block_start * nb_block_sq + (int64_t)row_offset * nb_block + col_offset;
    }
    ++block_start;
    ++col_index;
  }
  return -1;
}

/* ---- right-hand side ------------------------------------------------------------ */

/* femutils/ArcaneFemFunctions.h:1601-1611 (tet) and :2531-2541 (tri) */
void afo_rhs_constant_source(int dim, const double* coords, int32_t nb_cell, int32_t npc,
                             const int32_t* cells, const double* cell_scale,
                             double f, int nb_dof, int dof_index, double* rhs)
{
  if (AFO_IS_Q1(dim, npc)) {
    for (int64_t c = 0; c < nb_cell; ++c) {
      const int32_t* cn = cells + c * AFO_CELL_STRIDE(npc);
      double X[24], b[8];
      for (int k = 0; k < npc; ++k)
        memcpy(X + 3 * k, coords + 3 * (int64_t)cn[k], 3 * sizeof(double));
      afo_source_q1(dim, X, cell_scale ? f * cell_scale[c] : f, b);
      for (int k = 0; k < npc; ++k)
        rhs[(int64_t)cn[k] * nb_dof + dof_index] += b[k];
    }
    return;
  }
  for (int64_t c = 0; c < nb_cell; ++c) {
    const int32_t* cn = cells + c * 4;
    double m;
    if (npc == 3)
      m = afo_area_tria3(coords + 3 * (int64_t)cn[0], coords + 3 * (int64_t)cn[1], coords + 3 * (int64_t)cn[2]);
    else
      m = afo_volume_tetra4(coords + 3 * (int64_t)cn[0], coords + 3 * (int64_t)cn[1],
                            coords + 3 * (int64_t)cn[2], coords + 3 * (int64_t)cn[3]);
    double q = cell_scale ? f * cell_scale[c] : f;
    for (int k = 0; k < npc; ++k)
      rhs[(int64_t)cn[k] * nb_dof + dof_index] += q * m / npc;
  }
}

/* BoundaryConditions2D::integrateNodalFieldToRhsTria3 (femutils/ArcaneFemFunctions.h:2558-2590)
 * and BoundaryConditions3D::integrateNodalFieldToRhsTetra4 (:1628-1660) on the field
 * multiplied by scale first (the heat module does m_node_temperature_old.mult(1 / dt),
 * modules/heat/FemModule.cc:343): per cell, node i receives sum_j c_ij f_j with
 * c = area/6 | area/12 (tri) or volume/10 | volume/20 (tet) on | off the diagonal. */
void afo_rhs_mass_field(int dim, const double* coords, int32_t nb_cell, int32_t npc, const int32_t* cells,
                        const double* field, double scale, int nb_dof, int dof_index, double* rhs)
{
  if (AFO_IS_Q1(dim, npc)) {
    /* integrateNodalFieldToRhsQuad4 / Hexa8 (femutils/ArcaneFemFunctions.h:2609-2660,
     * :1680-1740): f_gp = sum_a N_a f_a, node i += f_gp N_i det J at each Gauss point */
    const int ng = dim == 2 ? 1 : 2;
    const double gp[2] = { -AFO_GP, AFO_GP };
    for (int64_t c = 0; c < nb_cell; ++c) {
      const int32_t* cn = cells + c * AFO_CELL_STRIDE(npc);
      double X[24], fv[8], acc[8] = { 0 };
      for (int k = 0; k < npc; ++k) {
        memcpy(X + 3 * k, coords + 3 * (int64_t)cn[k], 3 * sizeof(double));
        fv[k] = field[cn[k]] * scale;
      }
      for (int ix = 0; ix < 2; ++ix)
        for (int iy = 0; iy < 2; ++iy)
          for (int iz = 0; iz < ng; ++iz) {
            double N[8], gx[8], gy[8], gz[8];
            const double iw = afo_q1_gradients(dim, X, gp[ix], gp[iy], dim == 2 ? 0.0 : gp[iz], N, gx, gy, gz);
            double fg = 0.0;
            for (int a = 0; a < npc; ++a)
              fg += N[a] * fv[a];
            for (int i = 0; i < npc; ++i)
              acc[i] += fg * N[i] * iw;
          }
      for (int i = 0; i < npc; ++i)
        rhs[(int64_t)cn[i] * nb_dof + dof_index] += acc[i];
    }
    return;
  }
  for (int64_t c = 0; c < nb_cell; ++c) {
    const int32_t* cn = cells + c * 4;
    double m, cd, co;
    if (npc == 3) {
      m = afo_area_tria3(coords + 3 * (int64_t)cn[0], coords + 3 * (int64_t)cn[1], coords + 3 * (int64_t)cn[2]);
      cd = m / 6.0;
      co = m / 12.0;
    }
    else {
      m = afo_volume_tetra4(coords + 3 * (int64_t)cn[0], coords + 3 * (int64_t)cn[1], coords + 3 * (int64_t)cn[2],
                            coords + 3 * (int64_t)cn[3]);
      cd = m / 10.0;
      co = m / 20.0;
    }
    double fv[4];
    for (int j = 0; j < npc; ++j)
      fv[j] = field[cn[j]] * scale;
    for (int i = 0; i < npc; ++i) {
      double acc = 0.0;
      for (int j = 0; j < npc; ++j)
        acc += (i == j ? cd : co) * fv[j];
      rhs[(int64_t)cn[i] * nb_dof + dof_index] += acc;
    }
  }
}

/* modules/elasticity/BodyForce.h:149-161 */
void afo_rhs_body_force(const double* coords, int32_t nb_cell, const int32_t* cells,
                        const double f[3], double* rhs)
{
  for (int64_t c = 0; c < nb_cell; ++c) {
    const int32_t* cn = cells + c * 4;
    double volume = afo_volume_tetra4(coords + 3 * (int64_t)cn[0], coords + 3 * (int64_t)cn[1],
                                      coords + 3 * (int64_t)cn[2], coords + 3 * (int64_t)cn[3]);
    for (int k = 0; k < 4; ++k) {
      rhs[(int64_t)cn[k] * 3 + 0] += f[0] * volume / 4;
      rhs[(int64_t)cn[k] * 3 + 1] += f[1] * volume / 4;
      rhs[(int64_t)cn[k] * 3 + 2] += f[2] * volume / 4;
    }
  }
}

/* BoundaryConditions2D::applyNeumannToRhsTria3 (femutils/ArcaneFemFunctions.h:2715-2759)
 * and BoundaryConditions3D::applyNeumannToRhsTetra4 (:1793-1840), face by face in order.
 * Measures and normals as computeLengthEdge2 (:375-381), computeNormalEdge2 (:392-405),
 * computeAreaTria3 (:180-188) and computeNormalTriangle (:454-480), which swap the first
 * two nodes of a face that is not oriented outward; here that is a face whose normal
 * points towards the vertex of its cell that is not on the face. */
int afo_rhs_neumann(int dim, const double* coords, int32_t nb_cell, const int32_t* cells, int32_t nb_face,
                    const int32_t* faces, int nb_value, const double* values, int nb_dof, int dof_index,
                    double* rhs)
{
  const int k = dim, npc = dim + 1;
  const int scalar = nb_value == 1;
  for (int32_t f = 0; f < nb_face; ++f) {
    const int32_t* fn = faces + (int64_t)f * k;
    int32_t opp = -1;
    for (int32_t c = 0; c < nb_cell && opp < 0; ++c) {
      const int32_t* cn = cells + 4 * (int64_t)c;
      int hits = 0;
      for (int a = 0; a < k; ++a)
        for (int s = 0; s < npc; ++s)
          hits += cn[s] == fn[a];
      if (hits != k)
        continue;
      for (int s = 0; s < npc; ++s) {
        int on = 0;
        for (int a = 0; a < k; ++a)
          on |= cn[s] == fn[a];
        if (!on)
          opp = cn[s];
      }
    }
    if (opp < 0)
      return -1;
    const double* xo = coords + 3 * (int64_t)opp;
    const double* v0 = coords + 3 * (int64_t)fn[0];
    const double* v1 = coords + 3 * (int64_t)fn[1];
    double rhs_value;
    if (dim == 2) {
      const double ex = v1[0] - v0[0], ey = v1[1] - v0[1], ez = v1[2] - v0[2];
      const double length = sqrt(ex * ex + ey * ey + ez * ez);
      const double* a = v0;
      const double* b = v1;
      double dx = b[0] - a[0], dy = b[1] - a[1];
      double norm_n = sqrt(dx * dx + dy * dy);
      double nx = dy / norm_n, ny = -dx / norm_n;
      if (nx * (xo[0] - v0[0]) + ny * (xo[1] - v0[1]) > 0.0) {
        a = v1;
        b = v0;
        dx = b[0] - a[0];
        dy = b[1] - a[1];
        norm_n = sqrt(dx * dx + dy * dy);
        nx = dy / norm_n;
        ny = -dx / norm_n;
      }
      rhs_value = scalar ? values[0] * length / 2.0 : (nx * values[0] + ny * values[1]) * length / 2.0;
    }
    else {
      const double* v2 = coords + 3 * (int64_t)fn[2];
      double e1[3], e2[3], c[3];
      for (int d = 0; d < 3; ++d) {
        e1[d] = v1[d] - v0[d];
        e2[d] = v2[d] - v0[d];
      }
      c[0] = e1[1] * e2[2] - e1[2] * e2[1];
      c[1] = e1[2] * e2[0] - e1[0] * e2[2];
      c[2] = e1[0] * e2[1] - e1[1] * e2[0];
      const double area = sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]) / 2.0;
      if (c[0] * (xo[0] - v0[0]) + c[1] * (xo[1] - v0[1]) + c[2] * (xo[2] - v0[2]) > 0.0) {
        for (int d = 0; d < 3; ++d) {
          e1[d] = v0[d] - v1[d];
          e2[d] = v2[d] - v1[d];
        }
        c[0] = e1[1] * e2[2] - e1[2] * e2[1];
        c[1] = e1[2] * e2[0] - e1[0] * e2[2];
        c[2] = e1[0] * e2[1] - e1[1] * e2[0];
      }
      const double norm = sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
      const double nx = c[0] / norm, ny = c[1] / norm, nz = c[2] / norm;
      rhs_value = scalar ? values[0] * area / 3.0
                         : (nx * values[0] + ny * values[1] + nz * values[2]) * area / 3.0;
    }
    for (int a = 0; a < k; ++a)
      rhs[(int64_t)fn[a] * nb_dof + dof_index] += rhs_value;
  }
  return 0;
}

/* |F| of a face of k nodes: computeLengthEdge2 (femutils/ArcaneFemFunctions.h:
 * 375-381), computeAreaTria3 (:180-188) or computeAreaQuad4 (:202-212, the two
 * triangles n2 n1 n0 and n0 n3 n2). */
double afo_face_measure(int k, const double* coords, const int32_t* fn)
{
  const double* v0 = coords + 3 * (int64_t)fn[0];
  const double* v1 = coords + 3 * (int64_t)fn[1];
  double e1[3], e2[3], c[3];
  if (k == 2) {
    sub3(v1, v0, e1);
    return sqrt(dot3(e1, e1));
  }
  const double* v2 = coords + 3 * (int64_t)fn[2];
  if (k == 3) {
    sub3(v1, v0, e1);
    sub3(v2, v0, e2);
    cross3(e1, e2, c);
    return sqrt(dot3(c, c)) / 2.0;
  }
  const double* v3 = coords + 3 * (int64_t)fn[3];
  double d[3];
  sub3(v2, v1, e1);
  sub3(v0, v1, e2);
  cross3(e1, e2, c);
  sub3(v0, v3, e1);
  sub3(v2, v3, e2);
  cross3(e1, e2, d);
  return 0.5 * (sqrt(dot3(c, c)) + sqrt(dot3(d, d)));
}

/* rhs[node of F] += g |F| / k for every face F (k nodes): the convection RHS
 * (h Text) * measure / num_nodes of modules/heat/FemModule.cc:333-337, the
 * computeFacLengthOrArea rule of femutils/ArcaneFemFunctions.h:415-440. */
void afo_rhs_face_uniform(int k, const double* coords, int32_t nb_face, const int32_t* faces, double g,
                          int nb_dof, int dof_index, double* rhs)
{
  for (int32_t f = 0; f < nb_face; ++f) {
    const int32_t* fn = faces + (int64_t)f * k;
    const double v = g * afo_face_measure(k, coords, fn) / k;
    for (int a = 0; a < k; ++a)
      rhs[(int64_t)fn[a] * nb_dof + dof_index] += v;
  }
}

/* modules/heat/FemModule.cc:302-331: K_e = h * factor * massMatrix(U, U) * measure
 * with U all ones (FemUtils.h:584-596 doubles the diagonal), measure
 * computeLengthEdge2 (ArcaneFemFunctions.h:375-381) or computeAreaTria3
 * (:180-188), added entry by entry with matrixAddValue. */
int afo_convection_lhs(int k, const double* coords, int32_t nb_face, const int32_t* faces, double h,
                       const int32_t* row_index, const int32_t* columns, double* values)
{
  const double factor = k == 2 ? 1 / 6. : k == 3 ? 1 / 12. : 1 / 20.;
  for (int32_t f = 0; f < nb_face; ++f) {
    const int32_t* fn = faces + (int64_t)f * k;
    const double measure = afo_face_measure(k, coords, fn);
    for (int a = 0; a < k; ++a)
      for (int b = 0; b < k; ++b) {
        const double m = a == b ? 2.0 : 1.0;
        const int32_t x = find_in_row(row_index, columns, fn[a], fn[b]);
        if (x < 0)
          return -1;
        values[row_index[fn[a]] + x] += h * factor * m * measure;
      }
  }
  return 0;
}

/* Neumann condition on the boundary faces of a Quad4 mesh (edges, 2 nodes) or
 * a Hexa8 mesh (quadrangles, 4 nodes), femutils/ArcaneFemFunctionsGpu.cc:
 * 741-850 and :1146-1320: Gauss points on the face (2 on an edge, 2x2 on a
 * quadrangle), rhs[node j] += g N_j w |J| or (g . n) N_j w |J|.  Edges take
 * the outward normal (computeNormalFace, ArcaneFemFunctionsGpu.h:158-178);
 * quadrangles the normal t1 x t2 of the face's node order when
 * outward == 0, turned away from the cell when outward == 1.
 * Returns -1 when a face is not a face of any cell. */
int afo_rhs_neumann_q1(int dim, const double* coords, int32_t nb_cell, int32_t npc, const int32_t* cells,
                       int32_t nb_face, const int32_t* faces, int nb_value, const double* values, int nb_dof,
                       int dof_index, int outward, double* rhs)
{
  const int k = dim == 2 ? 2 : 4, st = AFO_CELL_STRIDE(npc);
  const int scalar = nb_value == 1;
  const double gp[2] = { -AFO_GP, AFO_GP };
  for (int32_t f = 0; f < nb_face; ++f) {
    const int32_t* fn = faces + (int64_t)f * k;
    /* a node of the face's cell that is not on the face (the inner side) */
    int32_t inner = -1;
    for (int32_t c = 0; c < nb_cell && inner < 0; ++c) {
      const int32_t* cn = cells + st * (int64_t)c;
      int hits = 0;
      for (int a = 0; a < k; ++a)
        for (int s2 = 0; s2 < npc; ++s2)
          hits += cn[s2] == fn[a];
      if (hits != k)
        continue;
      for (int s2 = 0; s2 < npc && inner < 0; ++s2) {
        int on = 0;
        for (int a = 0; a < k; ++a)
          on |= cn[s2] == fn[a];
        if (!on)
          inner = cn[s2];
      }
    }
    if (inner < 0)
      return -1;
    const double* xi_ = coords + 3 * (int64_t)inner;
    if (dim == 2) {
      const double* a = coords + 3 * (int64_t)fn[0];
      const double* b = coords + 3 * (int64_t)fn[1];
      const double length = sqrt((b[0] - a[0]) * (b[0] - a[0]) + (b[1] - a[1]) * (b[1] - a[1]));
      double nx = (b[1] - a[1]) / length, ny = (a[0] - b[0]) / length;
      if (nx * (xi_[0] - a[0]) + ny * (xi_[1] - a[1]) > 0.0) {
        nx = -nx;
        ny = -ny;
      }
      for (int i = 0; i < 2; ++i) {
        const double N[2] = { 0.5 * (1.0 - gp[i]), 0.5 * (1.0 + gp[i]) };
        const double iw = 1.0 * length * 0.5;
        for (int j = 0; j < 2; ++j)
          rhs[(int64_t)fn[j] * nb_dof + dof_index] +=
            (scalar ? values[0] : nx * values[0] + ny * values[1]) * N[j] * iw;
      }
      continue;
    }
    double X[12];
    for (int a = 0; a < 4; ++a)
      memcpy(X + 3 * a, coords + 3 * (int64_t)fn[a], 3 * sizeof(double));
    for (int ix = 0; ix < 2; ++ix)
      for (int iy = 0; iy < 2; ++iy) {
        double N[8], dxi[8], deta[8], dz[8];
        q1_shape(2, gp[ix], gp[iy], 0.0, N, dxi, deta, dz);
        double t1[3] = { 0, 0, 0 }, t2[3] = { 0, 0, 0 }, n[3];
        for (int a = 0; a < 4; ++a)
          for (int d = 0; d < 3; ++d) {
            t1[d] += dxi[a] * X[3 * a + d];
            t2[d] += deta[a] * X[3 * a + d];
          }
        cross3(t1, t2, n);
        const double det = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
        double u[3] = { n[0] / det, n[1] / det, n[2] / det };
        if (outward && u[0] * (xi_[0] - X[0]) + u[1] * (xi_[1] - X[1]) + u[2] * (xi_[2] - X[2]) > 0.0) {
          u[0] = -u[0];
          u[1] = -u[1];
          u[2] = -u[2];
        }
        const double iw = 1.0 * 1.0 * det;
        for (int j = 0; j < 4; ++j)
          rhs[(int64_t)fn[j] * nb_dof + dof_index] +=
            (scalar ? values[0] : u[0] * values[0] + u[1] * values[1] + u[2] * values[2]) * N[j] * iw;
      }
  }
  return 0;
}

/* Manufactured solutions of modules/poisson/ManufacturedSolutions.h:91-142:
 * which = 0 sine (sin(pi x) sin(pi y), source 2 pi^2 u), 1 poly
 * (x(1-x) y(1-y), source 2(x(1-x) + y(1-y))), 2 quadratic (x^2 + y^2,
 * source -4). */
double afo_manufactured_exact(int which, const double* x)
{
  const double Pi = 3.141592653589793238462643383279502884;
  if (which == 0)
    return sin(Pi * x[0]) * sin(Pi * x[1]);
  if (which == 1)
    return x[0] * (1.0 - x[0]) * x[1] * (1.0 - x[1]);
  return x[0] * x[0] + x[1] * x[1];
}

double afo_manufactured_source(int which, const double* x)
{
  const double Pi = 3.141592653589793238462643383279502884;
  if (which == 0)
    return 2.0 * Pi * Pi * afo_manufactured_exact(0, x);
  if (which == 1)
    return 2.0 * (x[0] * (1.0 - x[0]) + x[1] * (1.0 - x[1]));
  return -4.0;
}

/* 3-point Gauss rule of the Hexa8 manufactured-solution integrals
 * (modules/poisson/ManufacturedSolutions.cc:29-40): abscissae 0 and
 * +-sqrt(3/5), weights 8/9 and 5/9. */
static const double gp3[3] = { -0.77459666924148337704, 0.0, 0.77459666924148337704 };
static const double gw3[3] = { 5.0 / 9.0, 8.0 / 9.0, 5.0 / 9.0 };

/* modules/poisson/ManufacturedSolutions.cc:66-240 (_applyManufacturedSource):
 * tri3 source(barycentre) |T| / 3 per node; Quad4 (2x2 Gauss, weight 1) /
 * Hexa8 (3x3x3 Gauss, xi slowest) sum over the Gauss points of
 * N_i source(x_g) det J w_g (x_g = sum N_a x_a).  Returns -1 for other
 * cells (the reference has no branch for them). */
int afo_rhs_manufactured(int dim, const double* coords, int32_t nb_cell, int32_t npc, const int32_t* cells,
                         int which, double* rhs)
{
  const int st = AFO_CELL_STRIDE(npc);
  if (!(dim == 2 && npc == 3) && !AFO_IS_Q1(dim, npc))
    return -1;
  const double gp[2] = { -AFO_GP, AFO_GP };
  for (int64_t c = 0; c < nb_cell; ++c) {
    const int32_t* cn = cells + st * c;
    double X[24];
    for (int k = 0; k < npc; ++k)
      memcpy(X + 3 * k, coords + 3 * (int64_t)cn[k], 3 * sizeof(double));
    if (npc == 3) {
      const double area = afo_area_tria3(X, X + 3, X + 6);
      double xb[3];
      for (int d = 0; d < 3; ++d)
        xb[d] = (X[d] + X[3 + d] + X[6 + d]) / 3.0;
      const double src = afo_manufactured_source(which, xb);
      for (int k = 0; k < 3; ++k)
        rhs[cn[k]] += src * area / 3;
      continue;
    }
    const int np = dim == 2 ? 2 : 3;
    for (int ix = 0; ix < np; ++ix)
      for (int iy = 0; iy < np; ++iy)
        for (int iz = 0; iz < (dim == 2 ? 1 : 3); ++iz) {
          double N[8], gx[8], gy[8], gz[8];
          const double* p = dim == 2 ? gp : gp3;
          const double w = dim == 2 ? 1.0 : gw3[ix] * gw3[iy] * gw3[iz];
          const double det = afo_q1_gradients(dim, X, p[ix], p[iy], dim == 2 ? 0.0 : p[iz], N, gx, gy, gz);
          double xg[3] = { 0, 0, 0 };
          for (int a = 0; a < npc; ++a)
            for (int d = 0; d < 3; ++d)
              xg[d] += N[a] * X[3 * a + d];
          const double src = afo_manufactured_source(which, xg);
          for (int a = 0; a < npc; ++a)
            rhs[cn[a]] += N[a] * src * (det * w);
        }
  }
  return 0;
}

/* _computeManufacturedL2Error (ManufacturedSolutions.cc:270-420): tri3 at the
 * barycentre (uh the mean of the nodal values) times |T|; Quad4 at the 2x2,
 * Hexa8 at the 3x3x3 Gauss points (the rules of the source above); returns
 * sqrt of the sum. */
double afo_manufactured_l2_error(int dim, const double* coords, int32_t nb_cell, int32_t npc, const int32_t* cells,
                                 int which, const double* u)
{
  const int st = AFO_CELL_STRIDE(npc);
  const double gp[2] = { -AFO_GP, AFO_GP };
  double e2 = 0.0;
  for (int64_t c = 0; c < nb_cell; ++c) {
    const int32_t* cn = cells + st * c;
    double X[24];
    for (int k = 0; k < npc; ++k)
      memcpy(X + 3 * k, coords + 3 * (int64_t)cn[k], 3 * sizeof(double));
    if (npc == 3) {
      const double area = afo_area_tria3(X, X + 3, X + 6);
      double xb[3];
      for (int d = 0; d < 3; ++d)
        xb[d] = (X[d] + X[3 + d] + X[6 + d]) / 3.0;
      const double err = (u[cn[0]] + u[cn[1]] + u[cn[2]]) / 3.0 - afo_manufactured_exact(which, xb);
      e2 += err * err * area;
      continue;
    }
    const int np = dim == 2 ? 2 : 3;
    for (int ix = 0; ix < np; ++ix)
      for (int iy = 0; iy < np; ++iy)
        for (int iz = 0; iz < (dim == 2 ? 1 : 3); ++iz) {
          double N[8], gx[8], gy[8], gz[8];
          const double* p = dim == 2 ? gp : gp3;
          const double w = dim == 2 ? 1.0 : gw3[ix] * gw3[iy] * gw3[iz];
          const double det = afo_q1_gradients(dim, X, p[ix], p[iy], dim == 2 ? 0.0 : p[iz], N, gx, gy, gz);
          double xg[3] = { 0, 0, 0 }, uh = 0.0;
          for (int a = 0; a < npc; ++a) {
            uh += N[a] * u[cn[a]];
            for (int d = 0; d < 3; ++d)
              xg[d] += N[a] * X[3 * a + d];
          }
          const double err = uh - afo_manufactured_exact(which, xg);
          e2 += err * err * det * w;
        }
  }
  return sqrt(e2);
}

/* ---- transformations ------------------------------------------------------------ */

int afo_apply_transformations(int32_t nb_row, const int32_t* row_index, const int32_t* columns,
                              double* values, double* rhs,
                              const uint8_t* forced_info, const double* forced_value,
                              const uint8_t* elim_info, const double* elim_value)
{
  /* _fillRowColumnEliminationInfos (CsrDoFLinearSystemImpl.cc:187-230): the
   * (row, col) values touching a row-column eliminated DoF, captured before the
   * matrix is modified.  Kept in CSR order, which is the (row, col) order of
   * the reference's ordered map when columns are sorted. */
  int has_rc = 0;
  if (elim_info)
    for (int32_t i = 0; i < nb_row; ++i)
      if (elim_info[i] == AFO_ELIMINATE_ROW_COLUMN) {
        has_rc = 1;
        break;
      }
  double* saved = NULL;
  if (has_rc) {
    int64_t nnz = row_index[nb_row];
    saved = (double*)malloc(sizeof(double) * (size_t)(nnz > 0 ? nnz : 1));
    memcpy(saved, values, sizeof(double) * (size_t)nnz);
  }
  if (elim_info) {
    /* _applyRowEliminationOnMatrix (:126-152) */
    for (int32_t i = 0; i < nb_row; ++i)
      if (elim_info[i] == AFO_ELIMINATE_ROW)
        for (int32_t k = row_index[i]; k < row_index[i + 1]; ++k)
          values[k] = (columns[k] == i) ? 1.0 : 0.0;
    /* _applyRowColumnEliminationOnMatrix (:88-121); note the reference's
     * 'column_index > 0' test, which leaves column 0 untouched. */
    if (has_rc)
      for (int32_t i = 0; i < nb_row; ++i) {
        int row_elim = elim_info[i] == AFO_ELIMINATE_ROW || elim_info[i] == AFO_ELIMINATE_ROW_COLUMN;
        for (int32_t k = row_index[i]; k < row_index[i + 1]; ++k) {
          int32_t col = columns[k];
          if (col > 0) {
            int col_elim = elim_info[col] == AFO_ELIMINATE_ROW || elim_info[col] == AFO_ELIMINATE_ROW_COLUMN;
            if (row_elim || col_elim)
              values[k] = (col == i) ? 1.0 : 0.0;
          }
        }
      }
  }
  /* _applyForcedValuesToLhs (:50-72) */
  if (forced_info)
    for (int32_t i = 0; i < nb_row; ++i)
      if (forced_info[i]) {
        int32_t x = find_in_row(row_index, columns, i, i);
        if (x < 0) {
          free(saved);
          return -1;
        }
        values[row_index[i] + x] = forced_value[i];
      }
  /* applyRHSTransformation: _applyRowColumnEliminationToRHS
   * (DoFLinearSystemImplBase.cc:55-88) then _applyRowOrRowColumnEliminationOnRHS
   * (CsrDoFLinearSystemImpl.cc:157-182). */
  if (has_rc)
    for (int32_t i = 0; i < nb_row; ++i)
      for (int32_t k = row_index[i]; k < row_index[i + 1]; ++k) {
        int32_t col = columns[k];
        if (col < 0 || col == i)
          continue;
        if (!(elim_info[i] == AFO_ELIMINATE_ROW_COLUMN || elim_info[col] == AFO_ELIMINATE_ROW_COLUMN))
          continue;
        if (elim_info[i] == AFO_ELIMINATE_ROW_COLUMN)
          rhs[col] = rhs[col] - saved[k] * elim_value[i];
      }
  if (elim_info)
    for (int32_t i = 0; i < nb_row; ++i)
      if (elim_info[i] == AFO_ELIMINATE_ROW || elim_info[i] == AFO_ELIMINATE_ROW_COLUMN)
        rhs[i] = elim_value[i];
  free(saved);
  return 0;
}

/* ---- sparse kernels and PCG ------------------------------------------------------- */

void afo_spmv(int32_t nb_row, const int32_t* row_index, const int32_t* columns,
              const double* values, const double* x, double* y)
{
  for (int32_t i = 0; i < nb_row; ++i) {
    double s = 0.0;
    for (int32_t k = row_index[i]; k < row_index[i + 1]; ++k)
      s += values[k] * x[columns[k]];
    y[i] = s;
  }
}

static double dot_n(int32_t n, const double* a, const double* b)
{
  double s = 0.0;
  for (int32_t i = 0; i < n; ++i)
    s += a[i] * b[i];
  return s;
}

int afo_pcg_jacobi(int32_t n, const int32_t* row_index, const int32_t* columns,
                   const double* values, const double* b, double* x,
                   double rtol, int max_iter, double* true_relres)
{
  double* r = (double*)malloc(sizeof(double) * (size_t)n);
  double* z = (double*)malloc(sizeof(double) * (size_t)n);
  double* p = (double*)malloc(sizeof(double) * (size_t)n);
  double* q = (double*)malloc(sizeof(double) * (size_t)n);
  double* dinv = (double*)malloc(sizeof(double) * (size_t)n);
  for (int32_t i = 0; i < n; ++i) {
    double d = 0.0;
    for (int32_t k = row_index[i]; k < row_index[i + 1]; ++k)
      if (columns[k] == i)
        d = values[k];
    dinv[i] = (d != 0.0) ? 1.0 / d : 1.0;
  }
  afo_spmv(n, row_index, columns, values, x, q);
  for (int32_t i = 0; i < n; ++i) {
    r[i] = b[i] - q[i];
    z[i] = dinv[i] * r[i];
    p[i] = z[i];
  }
  double bnorm = sqrt(dot_n(n, b, b));
  double rho = dot_n(n, r, z);
  int it = 0;
  int converged = (bnorm == 0.0) || (sqrt(dot_n(n, r, r)) <= rtol * bnorm);
  while (!converged && it < max_iter) {
    afo_spmv(n, row_index, columns, values, p, q);
    double alpha = rho / dot_n(n, p, q);
    for (int32_t i = 0; i < n; ++i) {
      x[i] += alpha * p[i];
      r[i] -= alpha * q[i];
      z[i] = dinv[i] * r[i];
    }
    ++it;
    double rho_new = dot_n(n, r, z);
    double rnorm = sqrt(dot_n(n, r, r));
    if (rnorm <= rtol * bnorm) {
      converged = 1;
      break;
    }
    double beta = rho_new / rho;
    rho = rho_new;
    for (int32_t i = 0; i < n; ++i)
      p[i] = z[i] + beta * p[i];
  }
  afo_spmv(n, row_index, columns, values, x, q);
  double rs = 0.0;
  for (int32_t i = 0; i < n; ++i)
    rs += (b[i] - q[i]) * (b[i] - q[i]);
  if (true_relres)
    *true_relres = bnorm > 0 ? sqrt(rs) / bnorm : sqrt(rs);
  free(r);
  free(z);
  free(p);
  free(q);
  free(dinv);
  return converged ? it : -1;
}

/* The same Jacobi-PCG on `nthreads` OpenMP threads (rows split statically;
 * the dot products are OpenMP reductions, so their summation order and the
 * iteration count can differ from the serial version in the last bits).
 * Timing baseline only: the CPU counterpart of the PETSc cg + jacobi solve
 * (femutils/PetscDoFLinearSystemFactory.axl:5-10) on all host cores. */
int afo_pcg_jacobi_threads(int32_t n, const int32_t* row_index, const int32_t* columns, const double* values,
                           const double* b, double* x, double rtol, int max_iter, int nthreads,
                           double* true_relres)
{
  double* r = (double*)malloc(sizeof(double) * (size_t)n);
  double* z = (double*)malloc(sizeof(double) * (size_t)n);
  double* p = (double*)malloc(sizeof(double) * (size_t)n);
  double* q = (double*)malloc(sizeof(double) * (size_t)n);
  double* dinv = (double*)malloc(sizeof(double) * (size_t)n);
  double bb = 0.0, rho = 0.0, rr = 0.0;
#pragma omp parallel for num_threads(nthreads) schedule(static) reduction(+ : bb, rho, rr)
  for (int32_t i = 0; i < n; ++i) {
    double d = 0.0, ax = 0.0;
    for (int32_t k = row_index[i]; k < row_index[i + 1]; ++k) {
      if (columns[k] == i)
        d = values[k];
      ax += values[k] * x[columns[k]];
    }
    dinv[i] = (d != 0.0) ? 1.0 / d : 1.0;
    r[i] = b[i] - ax;
    z[i] = dinv[i] * r[i];
    p[i] = z[i];
    bb += b[i] * b[i];
    rho += r[i] * z[i];
    rr += r[i] * r[i];
  }
  const double bnorm = sqrt(bb);
  int it = 0;
  int converged = (bnorm == 0.0) || (sqrt(rr) <= rtol * bnorm);
  while (!converged && it < max_iter) {
    double pq = 0.0;
#pragma omp parallel for num_threads(nthreads) schedule(static) reduction(+ : pq)
    for (int32_t i = 0; i < n; ++i) {
      double s = 0.0;
      for (int32_t k = row_index[i]; k < row_index[i + 1]; ++k)
        s += values[k] * p[columns[k]];
      q[i] = s;
      pq += p[i] * s;
    }
    const double alpha = rho / pq;
    double rho_new = 0.0, rn = 0.0;
#pragma omp parallel for num_threads(nthreads) schedule(static) reduction(+ : rho_new, rn)
    for (int32_t i = 0; i < n; ++i) {
      x[i] += alpha * p[i];
      r[i] -= alpha * q[i];
      z[i] = dinv[i] * r[i];
      rho_new += r[i] * z[i];
      rn += r[i] * r[i];
    }
    ++it;
    if (sqrt(rn) <= rtol * bnorm) {
      converged = 1;
      break;
    }
    const double beta = rho_new / rho;
    rho = rho_new;
#pragma omp parallel for num_threads(nthreads) schedule(static)
    for (int32_t i = 0; i < n; ++i)
      p[i] = z[i] + beta * p[i];
  }
  double rs = 0.0;
#pragma omp parallel for num_threads(nthreads) schedule(static) reduction(+ : rs)
  for (int32_t i = 0; i < n; ++i) {
    double s = 0.0;
    for (int32_t k = row_index[i]; k < row_index[i + 1]; ++k)
      s += values[k] * x[columns[k]];
    rs += (b[i] - s) * (b[i] - s);
  }
  if (true_relres)
    *true_relres = bnorm > 0 ? sqrt(rs) / bnorm : sqrt(rs);
  free(r);
  free(z);
  free(p);
  free(q);
  free(dinv);
  return converged ? it : -1;
}

/* ---- comparison --------------------------------------------------------------- */

int64_t afo_count_differences(int64_t n, const double* ref, const double* v,
                              double epsilon, double min_value)
{
  int64_t nb = 0;
  for (int64_t i = 0; i < n; ++i) {
    double a = ref[i], b = v[i];
    if (fabs(a) < min_value && fabs(b) < min_value)
      continue;
    double d = a - b;
    if (d == 0.0)
      continue;
    double s = fabs(a) + fabs(b);
    if (!(fabs(d) / s < epsilon))
      ++nb;
  }
  return nb;
}

double afo_wtime(void)
{
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}
