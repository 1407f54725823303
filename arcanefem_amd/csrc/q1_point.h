// q1_point.h -- device helpers shared by the Q1 kernels: the isoparametric
// Quad4 / Hexa8 point (shape functions, physical gradients, det J), the
// manufactured solutions of modules/poisson/ManufacturedSolutions.h:91-142
// and their quadrature, and the per-cell squared L2 error.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace afem {

constexpr double kGauss = 0.57735026918962576451; // 1/sqrt(3), weight 1

// Shape functions, reference gradients, physical gradients and det J at the
// Gauss point g (ixi-major: g = ixi*2 + ieta in 2D, ixi*4 + ieta*2 + izeta in
// 3D, the reference's loop order) of a cell with node coordinates x.
template <int DIM>
struct Q1Point
{
  static constexpr int N = DIM == 2 ? 4 : 8;
  double Nv[N], gx[N], gy[N], gz[N];
  double det;

  // point g of the 2-point Gauss rule per direction (xi slowest)
  __device__ __forceinline__ void eval(const double (&x)[N][3], int g)
  {
    const double xi = (DIM == 2 ? (g >> 1) : (g >> 2)) ? kGauss : -kGauss;
    const double eta = (DIM == 2 ? (g & 1) : ((g >> 1) & 1)) ? kGauss : -kGauss;
    const double zeta = (DIM == 3 && (g & 1)) ? kGauss : -kGauss;
    eval_at(x, xi, eta, zeta);
  }

  __device__ __forceinline__ void eval_at(const double (&x)[N][3], double xi, double eta, double zeta)
  {
    double dxi[N], deta[N], dzeta[N];
#pragma unroll
    for (int a = 0; a < N; ++a) {
      // corner signs of node a: counter-clockwise in the (xi, eta) plane, then zeta
      const double sx = ((a & 3) == 1 || (a & 3) == 2) ? 1.0 : -1.0;
      const double sy = ((a & 3) >= 2) ? 1.0 : -1.0;
      if (DIM == 2) {
        Nv[a] = 0.25 * (1.0 + sx * xi) * (1.0 + sy * eta);
        dxi[a] = 0.25 * sx * (1.0 + sy * eta);
        deta[a] = 0.25 * sy * (1.0 + sx * xi);
        dzeta[a] = 0.0;
      }
      else {
        const double sz = a >= 4 ? 1.0 : -1.0;
        Nv[a] = 0.125 * (1.0 + sx * xi) * (1.0 + sy * eta) * (1.0 + sz * zeta);
        dxi[a] = 0.125 * sx * (1.0 + sy * eta) * (1.0 + sz * zeta);
        deta[a] = 0.125 * sy * (1.0 + sx * xi) * (1.0 + sz * zeta);
        dzeta[a] = 0.125 * sz * (1.0 + sx * xi) * (1.0 + sy * eta);
      }
    }
    if (DIM == 2) {
      double j00 = 0, j01 = 0, j10 = 0, j11 = 0;
#pragma unroll
      for (int a = 0; a < N; ++a) {
        j00 += dxi[a] * x[a][0];
        j01 += dxi[a] * x[a][1];
        j10 += deta[a] * x[a][0];
        j11 += deta[a] * x[a][1];
      }
      det = j00 * j11 - j01 * j10;
      const double i00 = j11 / det, i01 = -j01 / det, i10 = -j10 / det, i11 = j00 / det;
#pragma unroll
      for (int a = 0; a < N; ++a) {
        gx[a] = i00 * dxi[a] + i01 * deta[a];
        gy[a] = i10 * dxi[a] + i11 * deta[a];
        gz[a] = 0.0;
      }
      return;
    }
    double J[3][3] = { { 0, 0, 0 }, { 0, 0, 0 }, { 0, 0, 0 } };
#pragma unroll
    for (int a = 0; a < N; ++a)
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        J[0][d] += dxi[a] * x[a][d];
        J[1][d] += deta[a] * x[a][d];
        J[2][d] += dzeta[a] * x[a][d];
      }
    const double c00 = J[1][1] * J[2][2] - J[1][2] * J[2][1];
    const double c01 = J[1][2] * J[2][0] - J[1][0] * J[2][2];
    const double c02 = J[1][0] * J[2][1] - J[1][1] * J[2][0];
    det = J[0][0] * c00 + J[0][1] * c01 + J[0][2] * c02;
    const double r = 1.0 / det;
    const double I00 = c00 * r, I01 = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) * r,
                 I02 = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) * r;
    const double I10 = c01 * r, I11 = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) * r,
                 I12 = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) * r;
    const double I20 = c02 * r, I21 = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) * r,
                 I22 = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * r;
#pragma unroll
    for (int a = 0; a < N; ++a) {
      gx[a] = I00 * dxi[a] + I01 * deta[a] + I02 * dzeta[a];
      gy[a] = I10 * dxi[a] + I11 * deta[a] + I12 * dzeta[a];
      gz[a] = I20 * dxi[a] + I21 * deta[a] + I22 * dzeta[a];
    }
  }
};

// Manufactured sources (modules/poisson/ManufacturedSolutions.h:91-142):
// 0 sine 2 pi^2 sin(pi x) sin(pi y), 1 poly 2 (x(1-x) + y(1-y)), 2 quadratic -4.
__device__ __forceinline__ double manufactured_source(int which, double x, double y)
{
  constexpr double Pi = 3.141592653589793238462643383279502884;
  if (which == 0)
    return 2.0 * Pi * Pi * (sin(Pi * x) * sin(Pi * y));
  if (which == 1)
    return 2.0 * (x * (1.0 - x) + y * (1.0 - y));
  return -4.0;
}

// Quadrature of the manufactured-solution integrals
// (modules/poisson/ManufacturedSolutions.cc:22-40): Quad4 the 2x2 Gauss rule
// with weight 1, Hexa8 the 3x3x3 rule (abscissae 0, +-sqrt(3/5), weights 8/9,
// 5/9), xi slowest.  point() evaluates q at point g and returns its weight.
template <int DIM>
struct ManufacturedRule
{
  static constexpr int NG = DIM == 2 ? 4 : 27;
  __device__ __forceinline__ static double point(int g, Q1Point<DIM>& q, const double (&x)[Q1Point<DIM>::N][3])
  {
    if constexpr (DIM == 2) {
      q.eval(x, g);
      return 1.0;
    }
    else {
      constexpr double p = 0.77459666924148337704;
      const int ix = g / 9, iy = (g / 3) % 3, iz = g % 3;
      const double xi = ix == 0 ? -p : (ix == 1 ? 0.0 : p);
      const double eta = iy == 0 ? -p : (iy == 1 ? 0.0 : p);
      const double zeta = iz == 0 ? -p : (iz == 1 ? 0.0 : p);
      const double wx = ix == 1 ? 8.0 / 9.0 : 5.0 / 9.0, wy = iy == 1 ? 8.0 / 9.0 : 5.0 / 9.0,
                   wz = iz == 1 ? 8.0 / 9.0 : 5.0 / 9.0;
      q.eval_at(x, xi, eta, zeta);
      return wx * wy * wz;
    }
  }
};

__device__ __forceinline__ double manufactured_exact(int which, double x, double y)
{
  constexpr double Pi = 3.141592653589793238462643383279502884;
  if (which == 0)
    return sin(Pi * x) * sin(Pi * y);
  if (which == 1)
    return x * (1.0 - x) * y * (1.0 - y);
  return x * x + y * y;
}

// The squared L2 error of the nodal solution u on cell c: tri3 (mean of u -
// exact(barycentre))^2 |T|, Quad4 / Hexa8 sum_g (u_h(x_g) - exact(x_g))^2
// det J_g w_g (_computeManufacturedL2Error, modules/poisson/
// ManufacturedSolutions.cc:270-420).  DIM = 0 selects tri3.
template <int DIM>
__device__ __forceinline__ double manufactured_cell_error_sq(const double* __restrict__ coords,
                                                             const int32_t* __restrict__ cn, int which,
                                                             const double* __restrict__ u)
{
  double e2 = 0.0;
  if constexpr (DIM == 0) {
    double x[3][3];
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int d = 0; d < 3; ++d)
        x[a][d] = coords[4 * (int64_t)cn[a] + d];
    const double e1[3] = { x[1][0] - x[0][0], x[1][1] - x[0][1], x[1][2] - x[0][2] };
    const double f1[3] = { x[2][0] - x[0][0], x[2][1] - x[0][1], x[2][2] - x[0][2] };
    const double v0 = e1[1] * f1[2] - e1[2] * f1[1], v1 = e1[2] * f1[0] - e1[0] * f1[2],
                 v2 = e1[0] * f1[1] - e1[1] * f1[0];
    const double area = sqrt(v0 * v0 + v1 * v1 + v2 * v2) / 2.0;
    const double bx = (x[0][0] + x[1][0] + x[2][0]) / 3.0, by = (x[0][1] + x[1][1] + x[2][1]) / 3.0;
    const double err = (u[cn[0]] + u[cn[1]] + u[cn[2]]) / 3.0 - manufactured_exact(which, bx, by);
    e2 = err * err * area;
  }
  else {
    constexpr int N = Q1Point<DIM>::N;
    double x[N][3], un[N];
#pragma unroll
    for (int a = 0; a < N; ++a) {
      const int32_t n = cn[a];
      x[a][0] = coords[4 * (int64_t)n];
      x[a][1] = coords[4 * (int64_t)n + 1];
      x[a][2] = coords[4 * (int64_t)n + 2];
      un[a] = u[n];
    }
    for (int g = 0; g < ManufacturedRule<DIM>::NG; ++g) {
      Q1Point<DIM> q;
      const double w = ManufacturedRule<DIM>::point(g, q, x);
      double xg = 0.0, yg = 0.0, uh = 0.0;
#pragma unroll
      for (int a = 0; a < N; ++a) {
        uh += q.Nv[a] * un[a];
        xg += q.Nv[a] * x[a][0];
        yg += q.Nv[a] * x[a][1];
      }
      const double err = uh - manufactured_exact(which, xg, yg);
      e2 += err * err * q.det * w;
    }
  }
  return e2;
}

} // namespace afem
