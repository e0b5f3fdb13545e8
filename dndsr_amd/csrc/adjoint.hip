// adjoint.hip -- discrete adjoint of the compact-FV residual: vector-Jacobian
// products of EvaluateRHS and of the VR Jacobi sweep.
//
// The reference has no adjoint; these kernels differentiate exactly the discrete
// operators of vr.hip / euler.hip:
//   EvaluateRHS:  rhs_i = sum_{f in i} s_if / V_i sum_g (-w_g F(U_L(x_g), U_R(x_g), uL, uR, n_g))
//                 with U_side(x) = u_side + phi_side(x) . uRec_side and deltaLambdaCell frozen;
//   VR sweep:     y_i = c0 x_i + c1 sum_k [A^-1B_{i,k} x_j + A^-1b_{i,k} (u_j - u_i)].
// Flux derivatives come from the templated Riemann solvers evaluated on dual
// numbers (one pass per input direction), so every branch of the flux is
// differentiated as executed.  Reconstruction is not stored: the caller passes the
// same uRec the forward pass used (the Python wrapper recomputes it).
#include "dual.hpp"
#include "gas.hpp"
#include "internal.hpp"

namespace dndsr
{
// zero-mean basis of cell c at the point x (+ shift, the periodic translation into
// the cell's frame); same arithmetic as the face kernel
template <int DIM, int NDOF>
__device__ inline void cellBasisAt(const DevMesh &m, idx c, const real *x, const real *shift, real *phi)
{
    constexpr int M = NDOF - 1;
    const size_t N = m.N;
    const real isc = m.iscale[c];
    real xs[3], mom[M > 0 ? M : 1];
#pragma unroll
    for (int r = 0; r < 3; r++)
    {
        real p = x[r] + (shift ? shift[r] : 0.0);
        xs[r] = p - m.cent[(size_t)r * N + c];
        if (DIM == 3 || r < 2)
            xs[r] *= isc;
    }
#pragma unroll
    for (int j = 0; j < M; j++)
        mom[j] = m.moment[(size_t)j * N + c];
    basisRow0<DIM, NDOF>(xs, mom, phi);
}

// One thread per (face, point), point-major: t = g F + f.  Writes the 20 sensitivities
// dF.Fbar / d(UL, UR, uLmean, uRmean) of the point to pbuf[20][G F].
template <int DIM, int NDOF, int RS, bool BATCH>
__global__ void __launch_bounds__(128) kRHSPointAdj(DevMesh m, GasParams gp, const real *__restrict__ u,
                                                    const real *__restrict__ uRec, const real *__restrict__ dlc,
                                                    const real *__restrict__ rbar, real *__restrict__ pbuf)
{
    constexpr int M = NDOF - 1;
    const size_t N = m.N, F = m.F, GF = (size_t)m.G * F;
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= GF)
        return;
    const idx f = (idx)(t % F);
    const int g = (int)(t / F);
    const idx cL = m.f2c[f], cR = m.f2c[F + f];
    real x[3], n[3], sh[3];
#pragma unroll
    for (int r = 0; r < 3; r++)
    {
        x[r] = m.fpts[((size_t)g * 3 + r) * F + f];
        n[r] = m.fnorm[((size_t)g * 3 + r) * F + f];
        sh[r] = m.fshift[(size_t)r * F + f];
    }
    const real wj = m.fwj[(size_t)g * F + f];
    real UL[5], UR[5], uLm[5], uRm[5], Fbar[5];
#pragma unroll
    for (int v = 0; v < 5; v++)
    {
        uLm[v] = UL[v] = u[(size_t)v * N + cL];
        uRm[v] = UR[v] = u[(size_t)v * N + cR];
        Fbar[v] = -wj * (rbar[(size_t)v * N + cL] / m.vol[cL] - rbar[(size_t)v * N + cR] / m.vol[cR]);
    }
    if (uRec)
    {
        real phiL[M > 0 ? M : 1], phiR[M > 0 ? M : 1];
        cellBasisAt<DIM, NDOF>(m, cL, x, nullptr, phiL);
        cellBasisAt<DIM, NDOF>(m, cR, x, sh, phiR);
#pragma unroll
        for (int v = 0; v < 5; v++)
        {
            real sL = 0, sR = 0;
#pragma unroll
            for (int j = 0; j < M; j++)
            {
                sL += phiL[j] * uRec[(size_t)(v * M + j) * N + cL];
                sR += phiR[j] * uRec[(size_t)(v * M + j) * N + cR];
            }
            UL[v] += sL, UR[v] += sR;
        }
    }
    const real dLambda = dlc ? fmax(dlc[cL], dlc[cR]) : 0.0;
    real nm[3];
    if constexpr (BATCH)
    {
#pragma unroll
        for (int r = 0; r < 3; r++)
            nm[r] = m.fmeanN[(size_t)r * F + f];
    }
#pragma unroll 1
    for (int d = 0; d < 20; d++)
    {
        const int which = d / 5, comp = d % 5;
        Dual DL[5], DR[5], DLm[5], DRm[5], DF[5];
#pragma unroll
        for (int v = 0; v < 5; v++)
        {
            DL[v] = Dual(UL[v], which == 0 && v == comp ? 1.0 : 0.0);
            DR[v] = Dual(UR[v], which == 1 && v == comp ? 1.0 : 0.0);
            DLm[v] = Dual(uLm[v], which == 2 && v == comp ? 1.0 : 0.0);
            DRm[v] = Dual(uRm[v], which == 3 && v == comp ? 1.0 : 0.0);
        }
        if constexpr (BATCH)
        {
            RoeBatchFaceT<Dual> bf;
            roeBatchSetup<RS>(DLm, DRm, nm, gp, dLambda, bf);
            roeBatchPoint<RS>(DL, DR, n, gp, bf, DF);
        }
        else
        {
            Dual l0, l1, l4;
            riemannFlux<RS>(DL, DR, DLm, DRm, n, gp, dLambda, DF, l0, l1, l4);
        }
        real s = 0;
#pragma unroll
        for (int v = 0; v < 5; v++)
            s += Fbar[v] * DF[v].d;
        pbuf[(size_t)d * GF + t] = s;
    }
}

// One thread per cell: collect the point sensitivities of the cell's faces into
// uBar (N x 5) and uRecBar (5 M x N), evaluating the cell's basis at each point.
template <int DIM, int NDOF>
__global__ void __launch_bounds__(128) kRHSGatherAdj(DevMesh m, const real *__restrict__ pbuf,
                                                     real *__restrict__ ubar, real *__restrict__ urecbar)
{
    constexpr int M = NDOF - 1;
    const idx i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m.N)
        return;
    const size_t N = m.N, F = m.F, GF = (size_t)m.G * F;
    real au[5] = {0, 0, 0, 0, 0}, ar[5][M > 0 ? M : 1];
#pragma unroll
    for (int v = 0; v < 5; v++)
#pragma unroll
        for (int j = 0; j < M; j++)
            ar[v][j] = 0;
    for (int k = 0; k < m.K; k++)
    {
        const idx f = m.c2f[(size_t)k * N + i];
        if (f < 0)
            continue;
        const bool back = m.f2c[f] == i;
        const int rowU = back ? 0 : 5, rowM = back ? 10 : 15;
        real sh[3];
#pragma unroll
        for (int r = 0; r < 3; r++)
            sh[r] = m.fshift[(size_t)r * F + f];
        for (int g = 0; g < m.G; g++)
        {
            const size_t t = (size_t)g * F + f;
            real gU[5];
#pragma unroll
            for (int v = 0; v < 5; v++)
            {
                gU[v] = pbuf[(size_t)(rowU + v) * GF + t];
                au[v] += gU[v] + pbuf[(size_t)(rowM + v) * GF + t];
            }
            if (urecbar)
            {
                real x[3], phi[M > 0 ? M : 1];
#pragma unroll
                for (int r = 0; r < 3; r++)
                    x[r] = m.fpts[((size_t)g * 3 + r) * F + f];
                cellBasisAt<DIM, NDOF>(m, i, x, back ? nullptr : sh, phi);
#pragma unroll
                for (int v = 0; v < 5; v++)
#pragma unroll
                    for (int j = 0; j < M; j++)
                        ar[v][j] += phi[j] * gU[v];
            }
        }
    }
#pragma unroll
    for (int v = 0; v < 5; v++)
        ubar[(size_t)v * N + i] = au[v];
    if (urecbar)
#pragma unroll
        for (int v = 0; v < 5; v++)
#pragma unroll
            for (int j = 0; j < M; j++)
                urecbar[(size_t)(v * M + j) * N + i] = ar[v][j];
}

template <int DIM, int NDOF, int RS, bool BATCH>
static void rhsAdjLaunch(Ctx &c, const GasParams &gp, const real *u, const real *uRec, const real *dlc,
                         const real *rbar, real *ubar, real *urecbar)
{
    const size_t GF = (size_t)c.m.G * c.m.F;
    kRHSPointAdj<DIM, NDOF, RS, BATCH><<<(unsigned)((GF + 127) / 128), 128, 0, c.stream>>>(c.m, gp, u, uRec, dlc,
                                                                                          rbar, c.adjPoint);
    DNDSR_HIP_CHECK(hipGetLastError());
    kRHSGatherAdj<DIM, NDOF><<<gridFor(c.m.N, 128), 128, 0, c.stream>>>(c.m, c.adjPoint, ubar, urecbar);
    DNDSR_HIP_CHECK(hipGetLastError());
}

template <int DIM, int NDOF>
static void rhsAdjRS(Ctx &c, const GasParams &gp, const real *u, const real *uRec, const real *dlc, const real *rbar,
                     real *ubar, real *urecbar)
{
    const bool batch = gp.rsMeanValueEig != 0 && gp.rsType >= RS_Roe_M1 && gp.rsType <= RS_Roe_M5;
#define DNDSR_ADJ_CASE(RSV, B)                                                                    \
    case RSV:                                                                                     \
        rhsAdjLaunch<DIM, NDOF, RSV, B>(c, gp, u, uRec, dlc, rbar, ubar, urecbar);               \
        break;
    if (batch)
    {
        switch (gp.rsType)
        {
            DNDSR_ADJ_CASE(RS_Roe_M1, true)
            DNDSR_ADJ_CASE(RS_Roe_M2, true)
            DNDSR_ADJ_CASE(RS_Roe_M3, true)
            DNDSR_ADJ_CASE(RS_Roe_M4, true)
            DNDSR_ADJ_CASE(RS_Roe_M5, true)
        default: throw std::runtime_error("invalid batched Riemann solver");
        }
    }
    else
    {
        switch (gp.rsType)
        {
            DNDSR_ADJ_CASE(RS_Roe, false)
            DNDSR_ADJ_CASE(RS_HLLC, false)
            DNDSR_ADJ_CASE(RS_HLLEP, false)
            DNDSR_ADJ_CASE(RS_HLLEP_V1, false)
            DNDSR_ADJ_CASE(RS_Roe_M1, false)
            DNDSR_ADJ_CASE(RS_Roe_M2, false)
            DNDSR_ADJ_CASE(RS_Roe_M3, false)
            DNDSR_ADJ_CASE(RS_Roe_M4, false)
            DNDSR_ADJ_CASE(RS_Roe_M5, false)
            DNDSR_ADJ_CASE(RS_Roe_M6, false)
            DNDSR_ADJ_CASE(RS_Roe_M7, false)
            DNDSR_ADJ_CASE(RS_Roe_M8, false)
        default: throw std::runtime_error("the rs type is invalid");
        }
    }
#undef DNDSR_ADJ_CASE
}

void launchEvalRHSVJP(Ctx &c, const GasParams &gp, const real *u, const real *uRec, const real *dlc,
                      const real *rbar, real *ubar, real *urecbar)
{
    DevMesh &m = c.m;
    KernelTimer kt(c, K_ADJOINT);
    if (m.dim == 2)
    {
        if (!uRec || m.P == 1)
            rhsAdjRS<2, 3>(c, gp, u, uRec, dlc, rbar, ubar, urecbar);
        else if (m.P == 2)
            rhsAdjRS<2, 6>(c, gp, u, uRec, dlc, rbar, ubar, urecbar);
        else
            rhsAdjRS<2, 10>(c, gp, u, uRec, dlc, rbar, ubar, urecbar);
    }
    else
    {
        if (!uRec || m.P == 1)
            rhsAdjRS<3, 4>(c, gp, u, uRec, dlc, rbar, ubar, urecbar);
        else if (m.P == 2)
            rhsAdjRS<3, 10>(c, gp, u, uRec, dlc, rbar, ubar, urecbar);
        else
            throw std::runtime_error("the adjoint of EvaluateRHS supports maxOrder 1 and 2");
    }
}

// ---------------------------------------------------------------------------
// Transposed VR sweep.  Forward: y_i = c0 x_i + c1 sum_k [B_{i,k} x_{j(i,k)} + b_{i,k}(u_j - u_i)]
// (c0, c1) = (1 - w, w) Jacobi, (1, -1) recordInc, (0, -1) recordInc with uRecIsZero,
// (1, -1) without the b term for IterDiff.  Gather form for cell j: its neighbours
// i across slots kj see j across their slot ki = c2s[kj][j], so
//   xbar_j = c0 ybar_j + c1 sum_kj B_{i,ki}^T ybar_i
//   ubar_j = c1 sum_kj b_{i,ki} . ybar_i - c1 sum_k b_{j,k} . ybar_j     (per variable)
// One thread per cell; the neighbour's matrix block is read column-wise.
// ---------------------------------------------------------------------------
template <int M, int NV>
__global__ void __launch_bounds__(128) kVRIterT(DevMesh m, const real *__restrict__ ybar, real *__restrict__ xbar,
                                                real *__restrict__ ubar, real c0, real c1, int useB, int useb)
{
    const idx jc = blockIdx.x * blockDim.x + threadIdx.x;
    if (jc >= m.N)
        return;
    const size_t N = m.N;
    real ax[NV][M];
    real au[NV];
#pragma unroll
    for (int v = 0; v < NV; v++)
    {
        au[v] = 0;
#pragma unroll
        for (int r = 0; r < M; r++)
            ax[v][r] = c0 * ybar[(size_t)(v * M + r) * N + jc];
    }
    for (int kj = 0; kj < m.K; kj++)
    {
        const idx i = m.c2c[(size_t)kj * N + jc];
        if (i < 0)
            continue;
        const int ki = m.c2s[(size_t)kj * N + jc];
        const real *B = m.AinvB + (size_t)(1 + ki) * M * M * N + i;
        const real *b = m.Ainvb + (size_t)ki * M * N + i;
        const real *bj = m.Ainvb + (size_t)kj * M * N + jc;
        real yi[M];
#pragma unroll
        for (int v = 0; v < NV; v++)
        {
#pragma unroll
            for (int q = 0; q < M; q++)
                yi[q] = ybar[(size_t)(v * M + q) * N + i];
            if (useB)
#pragma unroll
                for (int r = 0; r < M; r++)
                {
                    real s = 0;
#pragma unroll
                    for (int q = 0; q < M; q++)
                        s += B[(size_t)(q * M + r) * N] * yi[q];
                    ax[v][r] += c1 * s;
                }
            if (useb)
            {
                real s = 0, sj = 0;
#pragma unroll
                for (int q = 0; q < M; q++)
                {
                    s += b[(size_t)q * N] * yi[q];
                    sj += bj[(size_t)q * N] * ybar[(size_t)(v * M + q) * N + jc];
                }
                au[v] += c1 * (s - sj);
            }
        }
    }
    if (xbar)
#pragma unroll
        for (int v = 0; v < NV; v++)
#pragma unroll
            for (int r = 0; r < M; r++)
                xbar[(size_t)(v * M + r) * N + jc] = ax[v][r];
    if (ubar)
#pragma unroll
        for (int v = 0; v < NV; v++)
            ubar[(size_t)v * N + jc] = au[v];
}

template <int M>
static void vrTLaunch(Ctx &c, int nVars, const real *ybar, real *xbar, real *ubar, real c0, real c1, int useB,
                      int useb)
{
    const int g = gridFor(c.m.N, 128);
    if (nVars == 5)
        kVRIterT<M, 5><<<g, 128, 0, c.stream>>>(c.m, ybar, xbar, ubar, c0, c1, useB, useb);
    else if (nVars == 1)
        kVRIterT<M, 1><<<g, 128, 0, c.stream>>>(c.m, ybar, xbar, ubar, c0, c1, useB, useb);
    else
        throw std::runtime_error("nVars must be 1 or 5");
    DNDSR_HIP_CHECK(hipGetLastError());
}

void launchVRIterVJP(Ctx &c, int nVars, int mode, real relax, const real *ybar, real *xbar, real *ubar)
{
    // mode 0 Jacobi (relax), 1 recordInc, 2 recordInc + uRecIsZero, 3 IterDiff
    real c0 = 1, c1 = -1;
    int useB = 1, useb = 1;
    if (mode == 0)
        c0 = 1 - relax, c1 = relax;
    else if (mode == 2)
        c0 = 0, useB = 0;
    else if (mode == 3)
        useb = 0;
    KernelTimer kt(c, K_ADJOINT);
    switch (c.m.M)
    {
    case 2: vrTLaunch<2>(c, nVars, ybar, xbar, ubar, c0, c1, useB, useb); break;
    case 3: vrTLaunch<3>(c, nVars, ybar, xbar, ubar, c0, c1, useB, useb); break;
    case 5: vrTLaunch<5>(c, nVars, ybar, xbar, ubar, c0, c1, useB, useb); break;
    case 9: vrTLaunch<9>(c, nVars, ybar, xbar, ubar, c0, c1, useB, useb); break;
    case 19: vrTLaunch<19>(c, nVars, ybar, xbar, ubar, c0, c1, useB, useb); break;
    default: throw std::runtime_error("unsupported number of reconstruction modes");
    }
}

} // namespace dndsr
