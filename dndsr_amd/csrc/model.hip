// model.hip -- the scalar advection-diffusion model of the reference's CFV test physics
// (ModelEvaluator, src/CFV/ModelEvaluator.hpp:18-123 and ModelEvaluator.cpp:6-115), on the
// device: one thread per face evaluates the face-point states and gradients of both sides from
// the reconstruction, the upwind advective flux and the GRP viscous flux, and writes the face's
// integrated flux; a per-cell gather in cell2face order forms the residual (no atomics).
#include "internal.hpp"

namespace dndsr
{

// one thread per face; nVars <= 5 independent scalars advected with a = (ax, ay)
template <int NDOF>
__global__ void __launch_bounds__(128) kModelRHSFace(DevMesh m, int nVars, real ax, real ay, real sigma,
                                                     const real *__restrict__ u, const real *__restrict__ uRec,
                                                     real *__restrict__ fflux)
{
    constexpr int M = NDOF - 1;
    const idx f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= m.F)
        return;
    const size_t N = m.N, F = m.F;
    const idx cL = m.f2c[f], cR = m.f2c[F + f];
    const bool bnd = cR < 0;
    const idx cs[2] = {cL, bnd ? cL : cR};
    real cen[2][3], sc[2], mom[2][M > 0 ? M : 1];
#pragma unroll
    for (int s = 0; s < 2; s++)
    {
#pragma unroll
        for (int r = 0; r < 3; r++)
            cen[s][r] = m.cent[(size_t)r * N + cs[s]];
        sc[s] = m.scale[cs[s]];
#pragma unroll
        for (int j = 0; j < M; j++)
            mom[s][j] = m.moment[(size_t)j * N + cs[s]];
    }
    real sh[3] = {0, 0, 0};
    real minVol = m.vol[cL];
    if (!bnd)
    {
        minVol = fmin(minVol, m.vol[cR]);
#pragma unroll
        for (int r = 0; r < 3; r++)
            sh[r] = m.fshift[(size_t)r * F + f];
    }
    const real distGRP = minVol / m.farea[f] * 2;
    real acc[5] = {0, 0, 0, 0, 0};
    for (int g = 0; g < m.G; g++)
    {
        real x[3], n[3];
#pragma unroll
        for (int r = 0; r < 3; r++)
            x[r] = m.fpts[((size_t)g * 3 + r) * F + f], n[r] = m.fnorm[((size_t)g * 3 + r) * F + f];
        // basis values and first derivatives at the point, from each side (R in its own frame)
        real phi[2][M > 0 ? M : 1], dphi[2][2][M > 0 ? M : 1];
#pragma unroll
        for (int s = 0; s < 2; s++)
        {
            real xs[3];
#pragma unroll
            for (int r = 0; r < 3; r++)
                xs[r] = (x[r] + (s ? sh[r] : 0.0)) - cen[s][r];
            xs[0] /= sc[s], xs[1] /= sc[s];
#pragma unroll
            for (int j = 0; j < M; j++)
            {
                phi[s][j] = basisDiff(2, 0, j + 1, xs, sc[s]) - mom[s][j];
                dphi[s][0][j] = basisDiff(2, 1, j + 1, xs, sc[s]);
                dphi[s][1][j] = basisDiff(2, 2, j + 1, xs, sc[s]);
            }
        }
        const real an = n[0] * ax + n[1] * ay;
        const real wj = m.fwj[(size_t)g * F + f];
        for (int v = 0; v < nVars; v++)
        {
            real UL = u[(size_t)v * N + cL], UR = bnd ? 0.0 : u[(size_t)v * N + cR];
            real gL0 = 0, gL1 = 0, gR0 = 0, gR1 = 0, rL = 0, rR = 0;
            if (uRec)
            {
#pragma unroll
                for (int j = 0; j < M; j++)
                {
                    const real xl = uRec[((size_t)v * M + j) * N + cL];
                    rL += phi[0][j] * xl, gL0 += dphi[0][0][j] * xl, gL1 += dphi[0][1][j] * xl;
                    if (!bnd)
                    {
                        const real xr = uRec[((size_t)v * M + j) * N + cR];
                        rR += phi[1][j] * xr, gR0 += dphi[1][0][j] * xr, gR1 += dphi[1][1][j] * xr;
                    }
                }
            }
            UL += rL;
            if (bnd) // generateBoundaryValue is zero; GradU_R = GradU_L
                gR0 = gL0, gR1 = gL1;
            else
                UR += rR;
            const real gm0 = (gR0 + gL0) * 0.5 + (1.0 / distGRP) * (n[0] * (UR - UL));
            const real gm1 = (gR1 + gL1) * 0.5 + (1.0 / distGRP) * (n[1] * (UR - UL));
            real finc = -(an * UR + an * UL) * 0.5 + 0.5 * fabs(an) * (UR - UL);
            finc += (gm0 * n[0] + gm1 * n[1]) * sigma;
            acc[v] += finc * wj;
        }
    }
    for (int v = 0; v < nVars; v++)
        fflux[(size_t)v * F + f] = acc[v];
}

__global__ void __launch_bounds__(256) kModelGather(DevMesh m, int nVars, const real *__restrict__ fflux,
                                                    real *__restrict__ rhs)
{
    const idx i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m.N)
        return;
    const size_t N = m.N, F = m.F;
    const real vol = m.vol[i];
    real acc[5] = {0, 0, 0, 0, 0};
    for (int k = 0; k < m.K; k++)
    {
        const idx f = m.c2f[(size_t)k * N + i];
        if (f < 0)
            continue;
        const bool back = m.f2c[f] == i;
        for (int v = 0; v < nVars; v++)
        {
            const real fv = fflux[(size_t)v * F + f];
            acc[v] += (back ? fv : -fv) / vol;
        }
    }
    for (int v = 0; v < nVars; v++)
        rhs[(size_t)v * N + i] = acc[v];
}

void launchModelRHS(Ctx &c, int nVars, real ax, real ay, real sigma, const real *u, const real *uRec, real *rhs)
{
    DevMesh &m = c.m;
    if (m.dim != 2)
        throw std::runtime_error("ModelEvaluator is two-dimensional");
    if (nVars < 1 || nVars > 5)
        throw std::runtime_error("ModelEvaluator: nVars must be 1..5");
    KernelTimer kt(c, K_RHS_FACE);
    const int nb = gridFor(m.F, 128);
    const int P = uRec ? m.P : 0;
    switch (P)
    {
    case 0: kModelRHSFace<1><<<nb, 128, 0, c.stream>>>(m, nVars, ax, ay, sigma, u, nullptr, c.faceFlux); break;
    case 1: kModelRHSFace<3><<<nb, 128, 0, c.stream>>>(m, nVars, ax, ay, sigma, u, uRec, c.faceFlux); break;
    case 2: kModelRHSFace<6><<<nb, 128, 0, c.stream>>>(m, nVars, ax, ay, sigma, u, uRec, c.faceFlux); break;
    default: kModelRHSFace<10><<<nb, 128, 0, c.stream>>>(m, nVars, ax, ay, sigma, u, uRec, c.faceFlux); break;
    }
    DNDSR_HIP_CHECK(hipGetLastError());
    KernelTimer kt2(c, K_RHS_GATHER);
    kModelGather<<<gridFor(m.N, 256), 256, 0, c.stream>>>(m, nVars, c.faceFlux, rhs);
    DNDSR_HIP_CHECK(hipGetLastError());
}

} // namespace dndsr
