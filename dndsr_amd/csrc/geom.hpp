// geom.hpp -- linear element maps and Gauss-Legendre rules of the setup kernels.
//
// Face-point geometry (point, unit normal, w * (JDet + tiny)) is tabulated once at
// setup by facePoint().  Evaluating it inside the face kernel from the face's node
// coordinates instead halves that kernel's HBM bytes but measured slower on gfx950
// (more VGPRs cost a wave slot per SIMD, and the kernel is latency-bound), so the
// tables stay.
#pragma once
#include "internal.hpp"

namespace dndsr
{
// Gauss-Legendre tables as the reference stores them (15 digits),
// src/Geom/Quadratures/Line.hpp:25-42.
constexpr real kGLp[4][4] = {{0, 0, 0, 0},
                             {-0.577350269189626, 0.577350269189626, 0, 0},
                             {-0.774596669241483, 0, 0.774596669241483, 0},
                             {-0.861136311594053, -0.339981043584856, 0.339981043584856, 0.861136311594053}};
constexpr real kGLw[4][4] = {{2, 0, 0, 0},
                             {1, 1, 0, 0},
                             {0.555555555555555, 0.888888888888889, 0.555555555555555, 0},
                             {0.347854845137454, 0.652145154862546, 0.652145154862546, 0.347854845137454}};

// points per direction for an integration order (QuadratureHub.hpp:39-96)
__host__ __device__ inline int glN(int order) { return order <= 1 ? 1 : order <= 3 ? 2 : order <= 5 ? 3 : 4; }
__host__ __device__ inline int paramDim(int t) { return t == E_LINE2 ? 1 : t == E_QUAD4 ? 2 : 3; }
__host__ __device__ inline int nodeCount(int t) { return t == E_LINE2 ? 2 : t == E_QUAD4 ? 4 : 8; }
__host__ __device__ inline int ruleSize(int t, int order)
{
    int n = glN(order), pd = paramDim(t);
    return pd == 1 ? n : pd == 2 ? n * n : n * n * n;
}
// tensor Gauss-Legendre point iG of an n-point-per-direction rule (first parametric
// index fastest, QuadratureHub.hpp:318-373)
__device__ inline void rulePointN(int pd, int n, int iG, real *xi, real &w)
{
    int i = iG % n, j = (iG / n) % n, k = iG / (n * n);
    xi[0] = kGLp[n - 1][i], xi[1] = 0, xi[2] = 0;
    w = kGLw[n - 1][i];
    if (pd >= 2)
        xi[1] = kGLp[n - 1][j], w = kGLw[n - 1][i] * kGLw[n - 1][j];
    if (pd == 3)
        xi[2] = kGLp[n - 1][k], w = kGLw[n - 1][i] * kGLw[n - 1][j] * kGLw[n - 1][k];
}
__device__ inline void rulePoint(int t, int order, int iG, real *xi, real &w)
{
    rulePointN(paramDim(t), glN(order), iG, xi, w);
}

// Linear element map: x(xi) and J = dx/dxi (CGNS node order)
__device__ inline void mapLinear(int t, const real (*X)[3], const real *xi, real *x, real (*J)[3])
{
    for (int r = 0; r < 3; r++)
    {
        x[r] = 0;
        for (int c = 0; c < 3; c++)
            J[r][c] = 0;
    }
    if (t == E_LINE2)
    {
        real N0 = 0.5 * (1 - xi[0]), N1 = 0.5 * (1 + xi[0]);
        for (int r = 0; r < 3; r++)
            x[r] = N0 * X[0][r] + N1 * X[1][r], J[r][0] = -0.5 * X[0][r] + 0.5 * X[1][r];
    }
    else if (t == E_QUAD4)
    {
        const real q[4][2] = {{-1, -1}, {1, -1}, {1, 1}, {-1, 1}};
        for (int a = 0; a < 4; a++)
        {
            real N = 0.25 * (1 + q[a][0] * xi[0]) * (1 + q[a][1] * xi[1]);
            real d0 = 0.25 * q[a][0] * (1 + q[a][1] * xi[1]);
            real d1 = 0.25 * q[a][1] * (1 + q[a][0] * xi[0]);
            for (int r = 0; r < 3; r++)
                x[r] += N * X[a][r], J[r][0] += d0 * X[a][r], J[r][1] += d1 * X[a][r];
        }
    }
    else
    {
        const real h[8][3] = {{-1, -1, -1}, {1, -1, -1}, {1, 1, -1}, {-1, 1, -1}, {-1, -1, 1}, {1, -1, 1}, {1, 1, 1}, {-1, 1, 1}};
        for (int a = 0; a < 8; a++)
        {
            real fx = 1 + h[a][0] * xi[0], fy = 1 + h[a][1] * xi[1], fz = 1 + h[a][2] * xi[2];
            real N = 0.125 * fx * fy * fz;
            real d0 = 0.125 * h[a][0] * fy * fz, d1 = 0.125 * h[a][1] * fx * fz, d2 = 0.125 * h[a][2] * fx * fy;
            for (int r = 0; r < 3; r++)
                x[r] += N * X[a][r], J[r][0] += d0 * X[a][r], J[r][1] += d1 * X[a][r], J[r][2] += d2 * X[a][r];
        }
    }
}
__device__ inline real cellDet(int dim, const real (*J)[3])
{
    if (dim == 2)
        return J[0][0] * J[1][1] - J[0][1] * J[1][0];
    return J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) - J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
           J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
}
__device__ inline void faceNormal(int dim, const real (*J)[3], real *n)
{
    if (dim == 2)
    {
        n[0] = J[1][0], n[1] = -J[0][0], n[2] = 0;
        return;
    }
    n[0] = J[1][0] * J[2][1] - J[2][0] * J[1][1];
    n[1] = J[2][0] * J[0][1] - J[0][0] * J[2][1];
    n[2] = J[0][0] * J[1][1] - J[1][0] * J[0][1];
}

// Face point iG of a Line2 (dim 2) or Quad4 (dim 3) face with n1d Gauss points per
// direction (FiniteVolume.cpp:353-575): physical point x, unit normal nu, the
// quadrature weight times (JDet + tiny) wj, JDet and the bare weight w.
__device__ __forceinline__ void facePoint(int dim, int n1d, int iG, const real (*X)[3], real *x, real *nu, real &wj,
                                          real &jd, real &w)
{
    const int t = dim == 2 ? E_LINE2 : E_QUAD4;
    real xi[3], J[3][3], n[3];
    rulePointN(paramDim(t), n1d, iG, xi, w);
    mapLinear(t, X, xi, x, J);
    faceNormal(dim, J, n);
    jd = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    wj = (jd + kVerySmall) * w;
    const real rj = 1.0 / jd;
    for (int r = 0; r < 3; r++)
        nu[r] = n[r] * rj;
}
} // namespace dndsr
