// adjoint.hip -- discrete adjoint of the compact-FV residual: vector-Jacobian
// products of EvaluateRHS and of the VR Jacobi sweep.
//
// The reference has no adjoint; these kernels differentiate exactly the discrete
// operators of vr.hip / euler.hip:
//   EvaluateRHS:  rhs_i = sum_{f in i} s_if / V_i sum_g w_g (-F(U_L, U_R, uL, uR, n_g) + Fv(U_L, U_R, G_L, G_R, n_g))
//                 with U_side(x) = u_side + phi_side(x) . uRec_side, G_side(x) = D1phi_side(x) . uRec_side,
//                 deltaLambdaCell frozen; a boundary face takes U_R = ghost(U_L), u_R = ghost(u_L) (at
//                 the mean normal), G_R = G_L;
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

// zero-mean basis and its gradient rows of cell c at x (+ shift), as the viscous pass
// evaluates them (scaled coordinates divided by the cell scale)
template <int DIM, int NDOF>
__device__ inline void cellBasisGradAt(const DevMesh &m, idx c, const real *x, const real *shift, real *phi,
                                       real (*dphi)[NDOF > 1 ? NDOF - 1 : 1])
{
    constexpr int M = NDOF - 1;
    const size_t N = m.N;
    const real s = m.scale[c];
    real xs[3];
#pragma unroll
    for (int r = 0; r < 3; r++)
        xs[r] = (x[r] + (shift ? shift[r] : 0.0)) - m.cent[(size_t)r * N + c];
    xs[0] /= s, xs[1] /= s;
    if (DIM == 3)
        xs[2] /= s;
#pragma unroll
    for (int j = 0; j < M; j++)
    {
        phi[j] = basisDiff(DIM, 0, j + 1, xs, s) - m.moment[(size_t)j * N + c];
#pragma unroll
        for (int d = 0; d < DIM; d++)
            dphi[d][j] = basisDiff(DIM, 1 + d, j + 1, xs, s);
    }
}

// value and gradient of cell c's reconstruction at x (+ shift), as the viscous pass
template <int DIM, int NDOF>
__device__ inline void cellValGradAt(const DevMesh &m, idx c, const real *x, const real *shift, const real *u,
                                     const real *uRec, real *U, real (*G)[5])
{
    constexpr int M = NDOF - 1;
    const size_t N = m.N;
#pragma unroll
    for (int v = 0; v < 5; v++)
    {
        U[v] = u[(size_t)v * N + c];
#pragma unroll
        for (int d = 0; d < 3; d++)
            G[d][v] = 0;
    }
    if (!uRec)
        return;
    real phi[M > 0 ? M : 1], dphi[DIM][M > 0 ? M : 1];
    cellBasisGradAt<DIM, NDOF>(m, c, x, shift, phi, dphi);
#pragma unroll
    for (int v = 0; v < 5; v++)
    {
        real a = 0, g[DIM];
#pragma unroll
        for (int d = 0; d < DIM; d++)
            g[d] = 0;
#pragma unroll
        for (int j = 0; j < M; j++)
        {
            const real r = uRec[(size_t)(v * M + j) * N + c];
            a += phi[j] * r;
#pragma unroll
            for (int d = 0; d < DIM; d++)
                g[d] += dphi[d][j] * r;
        }
        U[v] += a;
#pragma unroll
        for (int d = 0; d < DIM; d++)
            G[d][v] = g[d];
    }
}

// Point-sensitivity rows (of the buffer pbuf[row][G F]): 0-4 d/dU_L, 5-9 d/dU_R, 10-14
// d/du_L (cell mean), 15-19 d/du_R, and with viscous fluxes 20-34 d/dG_L[d][v]
// (row 20 + 5 d + v), 35-49 d/dG_R.
constexpr int kAdjRowsInv = 20, kAdjRowsVis = 50;

// One thread per (face, point), point-major: t = g F + f.  Writes the 20 sensitivities
// dF.Fbar / d(UL, UR, uLmean, uRmean) of the point to pbuf[20][G F].
template <int DIM, int NDOF, int RS, bool BATCH>
__global__ void __launch_bounds__(128) kRHSPointAdj(DevMesh m, GasParams gp, const real *__restrict__ u,
                                                    const real *__restrict__ uRec, const real *__restrict__ dlc,
                                                    const real *__restrict__ fbar, real *__restrict__ pbuf)
{
    constexpr int M = NDOF - 1;
    const size_t N = m.N, F = m.F, GF = (size_t)m.G * F;
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= GF)
        return;
    const idx f = (idx)(t % F);
    const int g = (int)(t / F);
    const idx cL = m.f2c[f], cR = m.f2c[F + f];
    if (cR < 0) // boundary faces: kRHSBndPointAdj
        return;
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
        Fbar[v] = -wj * fbar[(size_t)v * F + f];
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

// Viscous part of an interior face point: adds d(Fv . Fvbar)/dU_L,R to rows 0-9 and
// writes the gradient rows 20-49.  Fvbar = -Fbar (the face flux gains +Fv w).
template <int DIM, int NDOF>
__global__ void __launch_bounds__(128) kRHSPointAdjVis(DevMesh m, GasParams gp, const real *__restrict__ u,
                                                       const real *__restrict__ uRec, const real *__restrict__ fbar,
                                                       real *__restrict__ pbuf)
{
    const size_t N = m.N, F = m.F, GF = (size_t)m.G * F;
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= GF)
        return;
    const idx f = (idx)(t % F);
    const int g = (int)(t / F);
    const idx cL = m.f2c[f], cR = m.f2c[F + f];
    if (cR < 0)
        return;
    real x[3], n[3], sh[3], dB[3];
#pragma unroll
    for (int r = 0; r < 3; r++)
    {
        x[r] = m.fpts[((size_t)g * 3 + r) * F + f];
        n[r] = m.fnorm[((size_t)g * 3 + r) * F + f];
        sh[r] = m.fshift[(size_t)r * F + f];
        dB[r] = (m.bary[(size_t)r * N + cR] - sh[r]) - m.bary[(size_t)r * N + cL];
    }
    const real d0 = fmin(m.vol[cL], m.vol[cR]) / m.farea[f] * 2;
    const real wj = m.fwj[(size_t)g * F + f];
    real Fvb[5];
#pragma unroll
    for (int v = 0; v < 5; v++)
        Fvb[v] = wj * fbar[(size_t)v * F + f];
    real UL[5], UR[5], GL[3][5], GR[3][5];
    cellValGradAt<DIM, NDOF>(m, cL, x, nullptr, u, uRec, UL, GL);
    cellValGradAt<DIM, NDOF>(m, cR, x, sh, u, uRec, UR, GR);
#pragma unroll 1
    for (int d = 0; d < 40; d++)
    {
        Dual DL[5], DR[5], DGL[3][5], DGR[3][5], DF[5];
#pragma unroll
        for (int v = 0; v < 5; v++)
        {
            DL[v] = Dual(UL[v], d == v ? 1.0 : 0.0);
            DR[v] = Dual(UR[v], d == 5 + v ? 1.0 : 0.0);
#pragma unroll
            for (int q = 0; q < 3; q++)
            {
                DGL[q][v] = Dual(GL[q][v], d == 10 + 5 * q + v ? 1.0 : 0.0);
                DGR[q][v] = Dual(GR[q][v], d == 25 + 5 * q + v ? 1.0 : 0.0);
            }
        }
        viscousPointFlux(gp, DL, DR, DGL, DGR, n, dB, d0, false, false, DF);
        real sv = 0;
#pragma unroll
        for (int v = 0; v < 5; v++)
            sv += Fvb[v] * DF[v].d;
        const int row = d < 10 ? d : 20 + (d - 10);
        if (d < 10)
            pbuf[(size_t)row * GF + t] += sv;
        else
            pbuf[(size_t)row * GF + t] = sv;
    }
}

// One thread per (boundary face, point): the sensitivities of the inviscid flux against
// the ghost states (and of the viscous flux, vis) to the interior point value, cell mean
// and gradient; the front-side rows are zero.
template <int DIM, int NDOF, int RS, bool BATCH>
__global__ void __launch_bounds__(128) kRHSBndPointAdj(DevMesh m, GasParams gp, const real *__restrict__ u,
                                                       const real *__restrict__ uRec, const real *__restrict__ dlc,
                                                       const real *__restrict__ fbar, real *__restrict__ pbuf, int vis)
{
    const size_t N = m.N, F = m.F, GF = (size_t)m.G * F;
    const size_t tb = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (tb >= (size_t)m.nBnd * m.G)
        return;
    const idx f = m.bndFace[tb % m.nBnd];
    const int g = (int)(tb / m.nBnd);
    const size_t t = (size_t)g * F + f;
    const idx cL = m.f2c[f];
    const int zone = m.fzone[f], type = m.bcType[zone];
    const real *far = m.bcVal + (size_t)zone * 5;
    real x[3], n[3], nm[3], dB[3];
#pragma unroll
    for (int r = 0; r < 3; r++)
    {
        x[r] = m.fpts[((size_t)g * 3 + r) * F + f];
        n[r] = m.fnorm[((size_t)g * 3 + r) * F + f];
        nm[r] = m.fmeanN[(size_t)r * F + f];
        dB[r] = m.fcent[(size_t)r * F + f] - m.bary[(size_t)r * N + cL];
    }
    const real d0 = m.vol[cL] / m.farea[f] * 2;
    const real wj = m.fwj[(size_t)g * F + f];
    const real dLambda = dlc ? dlc[cL] : 0.0;
    real Fbar[5], uLm[5];
#pragma unroll
    for (int v = 0; v < 5; v++)
    {
        Fbar[v] = -wj * fbar[(size_t)v * F + f];
        uLm[v] = u[(size_t)v * N + cL];
    }
    real UL[5], GL[3][5];
    cellValGradAt<DIM, NDOF>(m, cL, x, nullptr, u, uRec, UL, GL);
    const bool visFace =
        vis && !(type == BC_WALL_INVIS || type == BC_SYM || type == BC_FAR || type == BC_OUT);
    const int nDir = vis ? 25 : 10;
#pragma unroll 1
    for (int d = 0; d < nDir; d++)
    {
        Dual DL[5], DLm[5], DR[5], DRm[5], DF[5];
#pragma unroll
        for (int v = 0; v < 5; v++)
        {
            DL[v] = Dual(UL[v], d == v ? 1.0 : 0.0);
            DLm[v] = Dual(uLm[v], d == 5 + v ? 1.0 : 0.0);
        }
        bcGhost(type, DLm, nm, far, gp.gamma, DRm);
        bcGhost(type, DL, n, far, gp.gamma, DR);
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
        real sv = 0;
#pragma unroll
        for (int v = 0; v < 5; v++)
            sv += Fbar[v] * DF[v].d;
        if (visFace)
        {
            Dual DGL[3][5], DV[5];
#pragma unroll
            for (int q = 0; q < 3; q++)
#pragma unroll
                for (int v = 0; v < 5; v++)
                    DGL[q][v] = Dual(GL[q][v], d == 10 + 5 * q + v ? 1.0 : 0.0);
            viscousPointFlux(gp, DL, DR, DGL, DGL, n, dB, d0, true, type == BC_WALL, DV);
#pragma unroll
            for (int v = 0; v < 5; v++)
                sv -= Fbar[v] * DV[v].d;
        }
        const int row = d < 5 ? d : (d < 10 ? 10 + (d - 5) : 20 + (d - 10));
        pbuf[(size_t)row * GF + t] = sv;
    }
    // the front side has no state of its own
#pragma unroll
    for (int v = 0; v < 5; v++)
    {
        pbuf[(size_t)(5 + v) * GF + t] = 0;
        pbuf[(size_t)(15 + v) * GF + t] = 0;
    }
    if (vis)
        for (int q = 0; q < 15; q++)
            pbuf[(size_t)(35 + q) * GF + t] = 0;
}

// One thread per cell: collect the point sensitivities of the cell's faces into
// uBar (N x 5) and uRecBar (5 M x N), evaluating the cell's basis at each point.
template <int DIM, int NDOF>
__global__ void __launch_bounds__(128) kRHSGatherAdj(DevMesh m, const real *__restrict__ pbuf,
                                                     real *__restrict__ ubar, real *__restrict__ urecbar, int vis)
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
        const int rowU = back ? 0 : 5, rowM = back ? 10 : 15, rowG = back ? 20 : 35;
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
                if (vis)
                {
                    real phi2[M > 0 ? M : 1], dphi[DIM][M > 0 ? M : 1];
                    cellBasisGradAt<DIM, NDOF>(m, i, x, back ? nullptr : sh, phi2, dphi);
#pragma unroll
                    for (int d = 0; d < DIM; d++)
#pragma unroll
                        for (int v = 0; v < 5; v++)
                        {
                            const real gg = pbuf[(size_t)(rowG + 5 * d + v) * GF + t];
#pragma unroll
                            for (int j = 0; j < M; j++)
                                ar[v][j] += dphi[d][j] * gg;
                        }
                }
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

// adjoint of the face flux from the adjoint of the residual (the gather rhs_L += ff / V_L,
// rhs_R -= ff / V_R): fbar_f = rbar_L / V_L - rbar_R / V_R
__global__ void __launch_bounds__(256) kFaceBar(DevMesh m, const real *__restrict__ rbar, real *__restrict__ fbar)
{
    const idx f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= m.F)
        return;
    const size_t N = m.N, F = m.F;
    const idx cL = m.f2c[f], cR = m.f2c[F + f];
#pragma unroll
    for (int v = 0; v < 5; v++)
        fbar[(size_t)v * F + f] = cR >= 0 ? rbar[(size_t)v * N + cL] / m.vol[cL] - rbar[(size_t)v * N + cR] / m.vol[cR]
                                          : rbar[(size_t)v * N + cL] / m.vol[cL];
}

template <int DIM, int NDOF, int RS, bool BATCH>
static void rhsAdjLaunch(Ctx &c, const GasParams &gp, const real *u, const real *uRec, const real *dlc,
                         const real *rbar, real *ubar, real *urecbar)
{
    const size_t GF = (size_t)c.m.G * c.m.F;
    const int vis = gp.viscous ? 1 : 0;
    kRHSPointAdj<DIM, NDOF, RS, BATCH><<<(unsigned)((GF + 127) / 128), 128, 0, c.stream>>>(c.m, gp, u, uRec, dlc,
                                                                                          rbar, c.adjPoint);
    // (rbar here is the face-flux adjoint [5][F], formed by launchEvalRHSVJP)
    DNDSR_HIP_CHECK(hipGetLastError());
    if (vis)
    {
        kRHSPointAdjVis<DIM, NDOF><<<(unsigned)((GF + 127) / 128), 128, 0, c.stream>>>(c.m, gp, u, uRec, rbar,
                                                                                      c.adjPoint);
        DNDSR_HIP_CHECK(hipGetLastError());
    }
    if (c.m.nBnd > 0)
    {
        const size_t nb = (size_t)c.m.nBnd * c.m.G;
        kRHSBndPointAdj<DIM, NDOF, RS, BATCH><<<(unsigned)((nb + 127) / 128), 128, 0, c.stream>>>(
            c.m, gp, u, uRec, dlc, rbar, c.adjPoint, vis);
        DNDSR_HIP_CHECK(hipGetLastError());
    }
    kRHSGatherAdj<DIM, NDOF><<<gridFor(c.m.N, 128), 128, 0, c.stream>>>(c.m, c.adjPoint, ubar, urecbar,
                                                                        vis && uRec != nullptr);
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
                      const real *rbar, real *ubar, real *urecbar, const real *faceBar)
{
    DevMesh &m = c.m;
    KernelTimer kt(c, K_ADJOINT);
    if (!faceBar)
    {
        kFaceBar<<<gridFor(m.F, 256), 256, 0, c.stream>>>(m, rbar, c.adjFaceBar);
        DNDSR_HIP_CHECK(hipGetLastError());
        faceBar = c.adjFaceBar;
    }
    rbar = faceBar; // the point kernels read the face-flux adjoint
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
            rhsAdjRS<3, 20>(c, gp, u, uRec, dlc, rbar, ubar, urecbar);
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

// Transposed boundary term of the VR sweep (forward: uRecNew_i += scale A_i^-1 BCC_i,
// BCC_i = sum_{boundary faces, g} phi(x_g) (w0^2 (u_B(x_g) - u_i)) w_g JDet_g, u_B the ghost
// of the compressed point state u_i + phi . uRec_i, or the given point value).  One
// thread per boundary cell; BCCbar = scale A^-T ybar, then the derivatives of BCC to the
// cell's mean and coefficients from dual-number passes (point values given: to the mean
// only), added to ubar / xbar.
template <int DIM, int NDOF, int NV>
__global__ void __launch_bounds__(64) kVRBoundaryT(DevMesh m, const real *__restrict__ uRec,
                                                   const real *__restrict__ u, const real *__restrict__ ybar,
                                                   real scale, real *__restrict__ xbar, real *__restrict__ ubar)
{
    constexpr int M = NDOF - 1;
    const idx t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= m.nBndCell)
        return;
    const idx i = m.bndCell[t];
    const size_t N = m.N, F = m.F;
    real Bb[M][NV];
#pragma unroll
    for (int j = 0; j < M; j++)
#pragma unroll
        for (int v = 0; v < NV; v++)
        {
            real a = 0;
#pragma unroll
            for (int r = 0; r < M; r++)
                a += m.AinvB[((size_t)r * M + j) * N + i] * ybar[((size_t)v * M + r) * N + i];
            Bb[j][v] = scale * a;
        }
    real mom[M], cen[3], um[NV];
#pragma unroll
    for (int j = 0; j < M; j++)
        mom[j] = m.moment[(size_t)j * N + i];
#pragma unroll
    for (int r = 0; r < 3; r++)
        cen[r] = m.cent[(size_t)r * N + i];
    const real isc = m.iscale[i];
#pragma unroll
    for (int v = 0; v < NV; v++)
        um[v] = u[(size_t)v * N + i];
    const bool euler = NV == 5 && m.bcType != nullptr;
    const int nDir = euler ? NV + NV * M : NV;
#pragma unroll 1
    for (int d = 0; d < nDir; d++)
    {
        real total = 0;
        for (int k = 0; k < m.K; k++)
        {
            const idx f = m.c2f[(size_t)k * N + i];
            if (f < 0 || m.c2c[(size_t)k * N + i] >= 0)
                continue;
            const int zone = m.fzone[f];
            const real w0 = m.faceW[f];
            const real w2 = w0 * w0;
            for (int g = 0; g < m.G; g++)
            {
                real xs[3], phi[M];
#pragma unroll
                for (int r = 0; r < 3; r++)
                {
                    xs[r] = m.fpts[((size_t)g * 3 + r) * F + f] - cen[r];
                    if (DIM == 3 || r < 2)
                        xs[r] *= isc;
                }
                basisRow0<DIM, NDOF>(xs, mom, phi);
                const real fac = m.fwj[(size_t)g * F + f];
                real incd[NV]; // d inc / d (direction d)
                if (euler)
                {
                    Dual umD[5], uBL[5], ULf[5], uB[5];
                    real n[3];
#pragma unroll
                    for (int v = 0; v < 5; v++)
                    {
                        umD[v] = Dual(um[v], d == v ? 1.0 : 0.0);
                        Dual r = 0;
#pragma unroll
                        for (int j = 0; j < M; j++)
                            r += phi[j] * Dual(uRec[((size_t)v * M + j) * N + i], d == NV + v * M + j ? 1.0 : 0.0);
                        uBL[v] = r;
                    }
#pragma unroll
                    for (int r = 0; r < 3; r++)
                        n[r] = m.fnorm[((size_t)g * 3 + r) * F + f];
                    compressRecPart(umD, uBL, ULf);
                    bcGhost(m.bcType[zone], ULf, n, m.bcVal + (size_t)zone * 5, m.bcGamma, uB);
#pragma unroll
                    for (int v = 0; v < NV; v++)
                        incd[v] = (uB[v] - umD[v]).d;
                }
                else
#pragma unroll
                    for (int v = 0; v < NV; v++)
                        incd[v] = d == v ? -1.0 : 0.0; // given point values: inc = value - u_i
#pragma unroll
                for (int j = 0; j < M; j++)
#pragma unroll
                    for (int v = 0; v < NV; v++)
                        total += Bb[j][v] * (((phi[j] * incd[v]) * w2) * fac);
            }
        }
        if (d < NV)
        {
            if (ubar)
                ubar[(size_t)d * N + i] += total;
        }
        else if (xbar)
            xbar[(size_t)(d - NV) * N + i] += total;
    }
}

template <int DIM, int NDOF>
static void vrBndTLaunch(Ctx &c, int nVars, const real *uRec, const real *u, const real *ybar, real scale, real *xbar,
                         real *ubar)
{
    const int g = gridFor(c.m.nBndCell, 64);
    if (nVars == 5)
        kVRBoundaryT<DIM, NDOF, 5><<<g, 64, 0, c.stream>>>(c.m, uRec, u, ybar, scale, xbar, ubar);
    else
        kVRBoundaryT<DIM, NDOF, 1><<<g, 64, 0, c.stream>>>(c.m, uRec, u, ybar, scale, xbar, ubar);
    DNDSR_HIP_CHECK(hipGetLastError());
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

// ---------------------------------------------------------------------------
// Transpose of the second-order substitute (kVR2nd): uRec_i = G_i(u) with
//   Gauss rule:     uRec_i[v][d] = 1/(V_i isc_i) sum_k c_ik[d] (u_k[v] - u_i[v]),
//                   c_ik = lThis/(lThis + lThat) A n sgn (boundary: A n / 2, u_k the ghost)
//   least squares:  uRec_i[v][d] = 1/isc_i (S_i^-1 sum_k dc_ik du_ik[v])[d], S_i = sum_k dc_ik dc_ik^T
// which is linear in u but for the boundary ghosts.  Pass 1 forms per cell the weight of its
// differences, q_i[v] = recBar_i[v] / (V_i isc_i) or S_i^-T recBar_i[v] / isc_i; pass 2 gathers
// per cell j: sum over its faces of c_ik . q_i from the neighbour i's side, minus its own
// c_jk . q_j, plus the ghosts' Jacobian (dual numbers) transposed.
// ---------------------------------------------------------------------------
template <int DIM>
__device__ inline void vr2ndFaceWeight(const DevMesh &m, idx i, int k, int method, bool plain, real *w)
{
    const size_t N = m.N, F = m.F;
    const idx f = m.c2f[(size_t)k * N + i];
    const idx o = m.c2c[(size_t)k * N + i];
    const bool back = m.f2c[f] == i;
    real bi[3], xf[3], n[3], ob[3] = {0, 0, 0};
    for (int r = 0; r < 3; r++)
    {
        bi[r] = m.bary[(size_t)r * N + i];
        xf[r] = m.fcent[(size_t)r * F + f] + (back ? 0.0 : m.fshift[(size_t)r * F + f]);
        n[r] = m.fmeanN[(size_t)r * F + f];
        if (o >= 0)
            ob[r] = m.bary[(size_t)r * N + o] + (back ? -m.fshift[(size_t)r * F + f] : m.fshift[(size_t)r * F + f]);
    }
    const real A = m.farea[f];
    if (method == 1)
    {
        if (o < 0)
        {
            for (int d = 0; d < DIM; d++)
                w[d] = 0.5 * A * n[d];
            return;
        }
        real lThis = 0, lThat = 0;
        for (int r = 0; r < 3; r++)
            lThis += (xf[r] - bi[r]) * (xf[r] - bi[r]), lThat += (ob[r] - xf[r]) * (ob[r] - xf[r]);
        lThis = plain ? 1.0 : sqrt(lThis);
        lThat = plain ? 1.0 : sqrt(lThat);
        const real sgn = back ? 1.0 : -1.0;
        for (int d = 0; d < DIM; d++)
            w[d] = lThis / (lThis + lThat + kVerySmall) * A * (n[d] * sgn);
        return;
    }
    if (o >= 0)
        for (int d = 0; d < DIM; d++)
            w[d] = ob[d] - bi[d];
    else
    {
        real pr = 0;
        for (int r = 0; r < 3; r++)
            pr += (bi[r] - m.fcent[(size_t)r * F + f]) * n[r];
        for (int d = 0; d < DIM; d++)
            w[d] = pr * -2.0 * n[d];
    }
}

template <int DIM, int NV>
__global__ void __launch_bounds__(64) kVR2ndTq(DevMesh m, const real *__restrict__ ybar, real *__restrict__ q)
{
    const idx i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m.N)
        return;
    const size_t N = m.N;
    const int method = m.s.subs2ndOrder;
    const bool plain = m.s.subs2ndOrderGGScheme == 0;
    const real isc = m.iscale[i];
    real gb[NV][DIM];
    for (int v = 0; v < NV; v++)
        for (int d = 0; d < DIM; d++)
            gb[v][d] = ybar[((size_t)v * DIM + d) * N + i] / isc;
    if (method == 1)
    {
        const real V = m.vol[i];
        for (int v = 0; v < NV; v++)
            for (int d = 0; d < DIM; d++)
                q[((size_t)v * DIM + d) * N + i] = gb[v][d] / V;
        return;
    }
    real S[DIM][DIM];
    for (int d = 0; d < DIM; d++)
        for (int e = 0; e < DIM; e++)
            S[d][e] = 0;
    for (int k = 0; k < m.K; k++)
    {
        if (m.c2f[(size_t)k * N + i] < 0)
            continue;
        real w[3];
        vr2ndFaceWeight<DIM>(m, i, k, method, plain, w);
        for (int d = 0; d < DIM; d++)
            for (int e = 0; e < DIM; e++)
                S[d][e] += w[d] * w[e];
    }
    real Si[DIM][DIM];
    if constexpr (DIM == 2)
    {
        const real det = S[0][0] * S[1][1] - S[0][1] * S[1][0];
        Si[0][0] = S[1][1] / det, Si[0][1] = -S[0][1] / det, Si[1][0] = -S[1][0] / det, Si[1][1] = S[0][0] / det;
    }
    else
    {
        real C[3][3];
        C[0][0] = S[1][1] * S[2][2] - S[1][2] * S[2][1];
        C[0][1] = S[0][2] * S[2][1] - S[0][1] * S[2][2];
        C[0][2] = S[0][1] * S[1][2] - S[0][2] * S[1][1];
        C[1][0] = S[1][2] * S[2][0] - S[1][0] * S[2][2];
        C[1][1] = S[0][0] * S[2][2] - S[0][2] * S[2][0];
        C[1][2] = S[0][2] * S[1][0] - S[0][0] * S[1][2];
        C[2][0] = S[1][0] * S[2][1] - S[1][1] * S[2][0];
        C[2][1] = S[0][1] * S[2][0] - S[0][0] * S[2][1];
        C[2][2] = S[0][0] * S[1][1] - S[0][1] * S[1][0];
        const real det = S[0][0] * C[0][0] + S[0][1] * C[1][0] + S[0][2] * C[2][0];
        for (int d = 0; d < 3; d++)
            for (int e = 0; e < 3; e++)
                Si[d][e] = C[d][e] / det;
    }
    for (int v = 0; v < NV; v++)
        for (int e = 0; e < DIM; e++)
        {
            real a = 0;
            for (int d = 0; d < DIM; d++)
                a += Si[d][e] * gb[v][d];
            q[((size_t)v * DIM + e) * N + i] = a;
        }
}

template <int DIM, int NV>
__global__ void __launch_bounds__(64) kVR2ndTGather(DevMesh m, const real *__restrict__ u,
                                                    const real *__restrict__ q, real *__restrict__ ubar)
{
    const idx j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= m.N)
        return;
    const size_t N = m.N, F = m.F;
    const int method = m.s.subs2ndOrder;
    const bool plain = m.s.subs2ndOrderGGScheme == 0;
    real qj[NV][DIM], acc[NV];
    for (int v = 0; v < NV; v++)
    {
        acc[v] = 0;
        for (int d = 0; d < DIM; d++)
            qj[v][d] = q[((size_t)v * DIM + d) * N + j];
    }
    for (int k = 0; k < m.K; k++)
    {
        const idx f = m.c2f[(size_t)k * N + j];
        if (f < 0)
            continue;
        const idx o = m.c2c[(size_t)k * N + j];
        real w[3];
        vr2ndFaceWeight<DIM>(m, j, k, method, plain, w);
        real s[NV]; // this face's own term: -c_jk . q_j
        for (int v = 0; v < NV; v++)
        {
            s[v] = 0;
            for (int d = 0; d < DIM; d++)
                s[v] += w[d] * qj[v][d];
            acc[v] -= s[v];
        }
        if (o >= 0)
        {
            // u_j as the neighbour of o, across o's slot of this face
            const int ko = m.c2s[(size_t)k * N + j];
            real wo[3];
            vr2ndFaceWeight<DIM>(m, o, ko, method, plain, wo);
            for (int v = 0; v < NV; v++)
                for (int d = 0; d < DIM; d++)
                    acc[v] += wo[d] * q[((size_t)v * DIM + d) * N + o];
        }
        else if constexpr (NV == 5)
        {
            // the ghost of u_j at the face centre (mean normal): its Jacobian transposed
            const int zone = m.fzone[f];
            real n[3];
            for (int r = 0; r < 3; r++)
                n[r] = m.fmeanN[(size_t)r * F + f];
            for (int cc = 0; cc < 5; cc++)
            {
                Dual U[5], G[5];
                for (int v = 0; v < 5; v++)
                    U[v] = Dual(u[(size_t)v * N + j], v == cc ? 1.0 : 0.0);
                bcGhost(m.bcType[zone], U, n, m.bcVal + (size_t)zone * 5, m.bcGamma, G);
                real a = 0;
                for (int v = 0; v < 5; v++)
                    a += G[v].d * s[v];
                acc[cc] += a;
            }
        }
    }
    for (int v = 0; v < NV; v++)
        ubar[(size_t)v * N + j] = acc[v];
}

void launchVRIterVJP(Ctx &c, int nVars, int mode, real relax, const real *uRec, const real *u, const real *ybar,
                     real *xbar, real *ubar)
{
    if (usesSubs2nd(c.m))
    {
        // every mode writes G(u): nothing flows back to uRec; IterDiff (mode 3) is zero
        KernelTimer kt(c, K_ADJOINT);
        const DevMesh &m = c.m;
        const size_t nR = (size_t)nVars * m.M * m.N;
        if (xbar)
            DNDSR_HIP_CHECK(hipMemsetAsync(xbar, 0, nR * sizeof(real), c.stream));
        if (mode == 3 || !ubar)
            return;
        if (nVars != 5 && nVars != 1)
            throw std::runtime_error("nVars must be 1 or 5");
        if (m.nBnd > 0 && !(nVars == 5 && m.bcType != nullptr && u != nullptr))
            throw std::runtime_error("the transposed second-order substitute on a mesh with boundary faces needs the "
                                     "Euler boundary table (nVars 5) and the base state u");
        if (!c.adj2ndQ)
            c.adj2ndQ = devAlloc<real>(c, (size_t)5 * 3 * m.N);
        const int g = gridFor(m.N, 64);
        if (m.dim == 2)
        {
            if (nVars == 5)
                kVR2ndTq<2, 5><<<g, 64, 0, c.stream>>>(m, ybar, c.adj2ndQ),
                    kVR2ndTGather<2, 5><<<g, 64, 0, c.stream>>>(m, u, c.adj2ndQ, ubar);
            else
                kVR2ndTq<2, 1><<<g, 64, 0, c.stream>>>(m, ybar, c.adj2ndQ),
                    kVR2ndTGather<2, 1><<<g, 64, 0, c.stream>>>(m, u, c.adj2ndQ, ubar);
        }
        else
        {
            if (nVars == 5)
                kVR2ndTq<3, 5><<<g, 64, 0, c.stream>>>(m, ybar, c.adj2ndQ),
                    kVR2ndTGather<3, 5><<<g, 64, 0, c.stream>>>(m, u, c.adj2ndQ, ubar);
            else
                kVR2ndTq<3, 1><<<g, 64, 0, c.stream>>>(m, ybar, c.adj2ndQ),
                    kVR2ndTGather<3, 1><<<g, 64, 0, c.stream>>>(m, u, c.adj2ndQ, ubar);
        }
        DNDSR_HIP_CHECK(hipGetLastError());
        return;
    }
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
    // boundary term (Jacobi: relax A^-1 BCC; recordInc: -A^-1 BCC; none with uRecIsZero)
    const DevMesh &m = c.m;
    if (m.nBnd == 0 || mode == 2)
        return;
    if (mode == 3)
        throw std::runtime_error("the transposed VR Jacobian (IterDiff) covers meshes without boundary faces");
    const bool euler = nVars == 5 && m.bcType != nullptr;
    if (!euler && !(m.dirVal && m.dirNV == nVars))
        throw std::runtime_error("the mesh has boundary faces: set the boundary conditions first");
    if (!u || (euler && !uRec))
        throw std::runtime_error("the transposed VR sweep on a mesh with boundary faces needs the base state uRec, u");
    const real scale = mode == 0 ? relax : -1.0;
    if (m.dim == 2)
    {
        if (m.P == 1)
            vrBndTLaunch<2, 3>(c, nVars, uRec, u, ybar, scale, xbar, ubar);
        else if (m.P == 2)
            vrBndTLaunch<2, 6>(c, nVars, uRec, u, ybar, scale, xbar, ubar);
        else
            vrBndTLaunch<2, 10>(c, nVars, uRec, u, ybar, scale, xbar, ubar);
    }
    else
    {
        if (m.P == 1)
            vrBndTLaunch<3, 4>(c, nVars, uRec, u, ybar, scale, xbar, ubar);
        else if (m.P == 2)
            vrBndTLaunch<3, 10>(c, nVars, uRec, u, ybar, scale, xbar, ubar);
        else
            vrBndTLaunch<3, 20>(c, nVars, uRec, u, ybar, scale, xbar, ubar);
    }
}

} // namespace dndsr
