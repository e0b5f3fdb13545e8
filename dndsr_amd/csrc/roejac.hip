// roejac.hip -- the Roe-Jacobian block LU-SGS of EulerEvaluatorSettings::useRoeJacobian:
// the face eigenvalues EvaluateRHS keeps for it, LUSGSMatrixInit with 5x5 diagonal blocks and
// their inverses, and the block multicolour UpdateSGS sweep (FP64 and FP32).
#include "gas.hpp"
#include "internal.hpp"

namespace dndsr
{

// ---------------------------------------------------------------------------
// Face eigenvalues (EulerEvaluator_EvaluateRHS.hxx:527-544): with useRoeJacobian the
// reference integrates the Riemann solver's entropy-fixed lam0 / lam123 / lam4 over every
// face and divides by its area, lambdaFace0/123/4 = sum_g w_g JDet_g lam(g) / A_f.  One
// thread per face: the same point states as the flux passes (means plus D0phi uRec, the
// ghost of the reconstructed state on a boundary face), the same Roe average, dLambda and
// batched eigenvalues.  lamRoe [3][F].
// ---------------------------------------------------------------------------
template <int DIM, int NDOF>
__device__ inline void sideState(const DevMesh &m, idx c, const real *u, const real *uRec, const real *x,
                                 const real *shift, real *U)
{
    constexpr int M = NDOF - 1;
    const size_t N = m.N;
#pragma unroll
    for (int v = 0; v < 5; v++)
        U[v] = u[(size_t)v * N + c];
    if (!uRec || M == 0)
        return;
    real xs[3], mom[M > 0 ? M : 1], phi[M > 0 ? M : 1];
    const real isc = m.iscale[c];
#pragma unroll
    for (int r = 0; r < 3; r++)
    {
        xs[r] = (x[r] + shift[r]) - m.cent[(size_t)r * N + c];
        if (DIM == 3 || r < 2)
            xs[r] *= isc;
    }
#pragma unroll
    for (int j = 0; j < M; j++)
        mom[j] = m.moment[(size_t)j * N + c];
    basisRow0<DIM, NDOF>(xs, mom, phi);
#pragma unroll
    for (int v = 0; v < 5; v++)
    {
        real s = 0;
#pragma unroll
        for (int j = 0; j < M; j++)
            s += phi[j] * uRec[((size_t)v * M + j) * N + c];
        U[v] += s;
    }
}

template <int DIM, int NDOF, int G, int RS, bool BATCH>
__global__ void __launch_bounds__(128) kFaceRoeLambda(DevMesh m, GasParams gp, const real *__restrict__ u,
                                                      const real *__restrict__ uRec, const real *__restrict__ dlc,
                                                      real *__restrict__ lamRoe)
{
    const idx f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= m.F)
        return;
    const size_t N = m.N, F = m.F;
    const idx cL = m.f2c[f], cR = m.f2c[F + f];
    const bool bnd = cR < 0;
    const int type = bnd ? m.bcType[m.fzone[f]] : 0;
    const real *far = bnd ? m.bcVal + (size_t)m.fzone[f] * 5 : nullptr;
    real ULm[5], URm[5], nm[3], sh[3], zero[3] = {0, 0, 0};
#pragma unroll
    for (int r = 0; r < 3; r++)
        nm[r] = m.fmeanN[(size_t)r * F + f], sh[r] = m.fshift[(size_t)r * F + f];
#pragma unroll
    for (int v = 0; v < 5; v++)
        ULm[v] = u[(size_t)v * N + cL];
    if (bnd)
        bcGhost(type, ULm, nm, far, gp.gamma, URm);
    else
#pragma unroll
        for (int v = 0; v < 5; v++)
            URm[v] = u[(size_t)v * N + cR];
    real dLambda = 0;
    if (dlc)
        dLambda = bnd ? dlc[cL] : fmax(dlc[cL], dlc[cR]);
    RoeBatchFace bf;
    if constexpr (BATCH)
        roeBatchSetup<RS>(ULm, URm, nm, gp, dLambda, bf);
    else
        roeAverage(ULm, URm, gp.gamma, bf.r);
    real le[3] = {0, 0, 0};
    for (int g = 0; g < G; g++)
    {
        real x[3], n[3], UL[5], UR[5];
#pragma unroll
        for (int r = 0; r < 3; r++)
            x[r] = m.fpts[((size_t)g * 3 + r) * F + f], n[r] = m.fnorm[((size_t)g * 3 + r) * F + f];
        sideState<DIM, NDOF>(m, cL, u, uRec, x, zero, UL);
        if (bnd)
            bcGhost(type, UL, n, far, gp.gamma, UR);
        else
            sideState<DIM, NDOF>(m, cR, u, uRec, x, sh, UR);
        real l0, l1, l4;
        if constexpr (BATCH)
            l0 = bf.lam0, l1 = bf.lam123, l4 = bf.lam4;
        else
        {
            real Fl[5];
            riemannFluxR<RS>(UL, UR, bf.r, n, gp, dLambda, Fl, l0, l1, l4);
        }
        const real wj = m.fwj[(size_t)g * F + f];
        le[0] += l0 * wj, le[1] += l1 * wj, le[2] += l4 * wj;
    }
    const real ia = 1.0 / m.farea[f];
#pragma unroll
    for (int k = 0; k < 3; k++)
        lamRoe[(size_t)k * F + f] = le[k] * ia;
}

template <int DIM, int NDOF, int G>
static void faceRoeLambdaRS(Ctx &c, const GasParams &gp, const real *u, const real *uRec, const real *dlc, real *lr)
{
    const bool batch = gp.rsMeanValueEig != 0 && gp.rsType >= RS_Roe_M1 && gp.rsType <= RS_Roe_M5;
    const int nb = gridFor(c.m.F, 128);
#define DNDSR_LAM_CASE(RSV, B)                                                                          \
    case RSV:                                                                                          \
        kFaceRoeLambda<DIM, NDOF, G, RSV, B><<<nb, 128, 0, c.stream>>>(c.m, gp, u, uRec, dlc, lr);     \
        break;
    if (batch)
        switch (gp.rsType)
        {
            DNDSR_LAM_CASE(RS_Roe_M1, true)
            DNDSR_LAM_CASE(RS_Roe_M2, true)
            DNDSR_LAM_CASE(RS_Roe_M3, true)
            DNDSR_LAM_CASE(RS_Roe_M4, true)
            DNDSR_LAM_CASE(RS_Roe_M5, true)
        default: throw std::runtime_error("invalid batched Riemann solver");
        }
    else
        switch (gp.rsType)
        {
            DNDSR_LAM_CASE(RS_Roe, false)
            DNDSR_LAM_CASE(RS_HLLC, false)
            DNDSR_LAM_CASE(RS_HLLEP, false)
            DNDSR_LAM_CASE(RS_HLLEP_V1, false)
            DNDSR_LAM_CASE(RS_Roe_M1, false)
            DNDSR_LAM_CASE(RS_Roe_M2, false)
            DNDSR_LAM_CASE(RS_Roe_M3, false)
            DNDSR_LAM_CASE(RS_Roe_M4, false)
            DNDSR_LAM_CASE(RS_Roe_M5, false)
            DNDSR_LAM_CASE(RS_Roe_M6, false)
            DNDSR_LAM_CASE(RS_Roe_M7, false)
            DNDSR_LAM_CASE(RS_Roe_M8, false)
        default: throw std::runtime_error("the rs type is invalid");
        }
#undef DNDSR_LAM_CASE
    DNDSR_HIP_CHECK(hipGetLastError());
}

template <int DIM, int NDOF>
static void faceRoeLambdaG(Ctx &c, const GasParams &gp, const real *u, const real *uRec, const real *dlc, real *lr)
{
    constexpr int G1 = DIM == 2 ? 2 : 4, G2 = DIM == 2 ? 3 : 9;
    if (c.m.G == G1)
        faceRoeLambdaRS<DIM, NDOF, G1>(c, gp, u, uRec, dlc, lr);
    else if (c.m.G == G2)
        faceRoeLambdaRS<DIM, NDOF, G2>(c, gp, u, uRec, dlc, lr);
    else if (DIM == 3 && c.m.G == 6)
        faceRoeLambdaRS<DIM, NDOF, 6>(c, gp, u, uRec, dlc, lr);
    else
        throw std::runtime_error("unsupported face rule for the face eigenvalues");
}

void launchFaceRoeLambda(Ctx &c, const GasParams &gp, const real *u, const real *uRec, const real *dlc, real *lamRoe)
{
    KernelTimer kt(c, K_RHS_FACE);
    const int p = c.m.s.maxOrder;
    if (c.m.dim == 2)
    {
        if (p == 1)
            faceRoeLambdaG<2, 3>(c, gp, u, uRec, dlc, lamRoe);
        else if (p == 2)
            faceRoeLambdaG<2, 6>(c, gp, u, uRec, dlc, lamRoe);
        else
            faceRoeLambdaG<2, 10>(c, gp, u, uRec, dlc, lamRoe);
    }
    else
    {
        if (p == 1)
            faceRoeLambdaG<3, 4>(c, gp, u, uRec, dlc, lamRoe);
        else if (p == 2)
            faceRoeLambdaG<3, 10>(c, gp, u, uRec, dlc, lamRoe);
        else
            faceRoeLambdaG<3, 20>(c, gp, u, uRec, dlc, lamRoe);
    }
}

// ---------------------------------------------------------------------------
// Block pieces.  RoeFluxIncFDiff (Gas.hpp:1209-1243) in precision T with the normal in T,
// applied to dU: |A|_Roe dU for eigenvalues lam0 / lam123 / lam4.
// ---------------------------------------------------------------------------
template <typename T>
__device__ inline void roeIncT(const T *dU, const T *n, const RoeAveT<T> &r, T lam0, T lam123, T lam4, T gamma,
                               T *incF)
{
    const T vRN = r.veloRoe[0] * n[0] + r.veloRoe[1] * n[1] + r.veloRoe[2] * n[2];
    const T dUn = dU[1] * n[0] + dU[2] * n[1] + dU[3] * n[2];
    T a23[3];
#pragma unroll
    for (int d = 0; d < 3; d++)
        a23[d] = dU[1 + d] - dU[0] * r.veloRoe[d];
    const T a23n = a23[0] * n[0] + a23[1] * n[1] + a23[2] * n[2];
#pragma unroll
    for (int d = 0; d < 3; d++)
        a23[d] -= n[d] * a23n;
    const T dU4b = dU[4] - (a23[0] * r.veloRoe[0] + a23[1] * r.veloRoe[1] + a23[2] * r.veloRoe[2]);
    T al1 = r.g1a2 * (dU[0] * (r.HRoe - vRN * vRN) + vRN * dUn - dU4b);
    T al0 = (dU[0] * (vRN + r.aRoe) - dUn - r.aRoe * al1) * r.inv2a;
    T al4 = dU[0] - (al0 + al1);
    al0 *= lam0, al1 *= lam123, al4 *= lam4;
#pragma unroll
    for (int d = 0; d < 3; d++)
        a23[d] *= lam123;
    incF[0] = al0 + al1 + al4;
    incF[4] = (r.HRoe - vRN * r.aRoe) * al0 + T(0.5) * r.vsqrRoe * al1 + (r.HRoe + vRN * r.aRoe) * al4 +
              (a23[0] * r.veloRoe[0] + a23[1] * r.veloRoe[1] + a23[2] * r.veloRoe[2]);
#pragma unroll
    for (int d = 0; d < 3; d++)
        incF[1 + d] = (r.veloRoe[d] - r.aRoe * n[d]) * al0 + (r.veloRoe[d] + r.aRoe * n[d]) * al4 +
                      r.veloRoe[d] * al1 + a23[d];
}

// the face's eigenvalues seen from a cell: lam0 and lam4 swap on the front side
// (EulerEvaluator.hxx:96, 617)
template <typename T>
__device__ inline void faceLambdas(const T *lam, const T *lamRoe, size_t F, idx f, bool front, T &l0, T &l123, T &l4)
{
    if (lamRoe)
        l0 = lamRoe[f], l123 = lamRoe[F + f], l4 = lamRoe[2 * F + f];
    else
        l0 = l123 = l4 = lam[3 * F + f];
    if (front)
    {
        T t = l0;
        l0 = l4, l4 = t;
    }
}

// JacobianDiagBlock::GetInvert (EulerJacobian.hpp:153-190): J^-1 = (D^-1 J)^-1 D^-1 with the
// scaled block inverted by Gauss-Jordan elimination with complete pivoting.  Returns false for a
// zero diagonal or a singular block.
__device__ inline bool invertBlock5(const real *J, real *Jinv)
{
    real P[5][5], X[5][5], dinv[5];
    int cp[5] = {0, 1, 2, 3, 4};
    bool ok = true;
#pragma unroll
    for (int i = 0; i < 5; i++)
    {
        ok = ok && J[i * 5 + i] != 0;
        dinv[i] = 1.0 / J[i * 5 + i];
    }
#pragma unroll
    for (int i = 0; i < 5; i++)
#pragma unroll
        for (int j = 0; j < 5; j++)
            P[i][j] = dinv[i] * J[i * 5 + j], X[i][j] = i == j ? 1.0 : 0.0;
    for (int k = 0; k < 5; k++)
    {
        int pr = k, pc = k;
        real best = -1;
        for (int i = k; i < 5; i++)
            for (int j = k; j < 5; j++)
                if (fabs(P[i][j]) > best)
                    best = fabs(P[i][j]), pr = i, pc = j;
        ok = ok && best > 0;
        for (int j = 0; j < 5; j++)
        {
            real t = P[k][j];
            P[k][j] = P[pr][j], P[pr][j] = t;
            t = X[k][j];
            X[k][j] = X[pr][j], X[pr][j] = t;
        }
        for (int i = 0; i < 5; i++)
        {
            real t = P[i][k];
            P[i][k] = P[i][pc], P[i][pc] = t;
        }
        int t = cp[k];
        cp[k] = cp[pc], cp[pc] = t;
        const real piv = 1.0 / P[k][k];
        for (int j = 0; j < 5; j++)
            P[k][j] *= piv, X[k][j] *= piv;
        for (int i = 0; i < 5; i++)
        {
            if (i == k)
                continue;
            const real fct = P[i][k];
            if (fct == 0)
                continue;
            for (int j = 0; j < 5; j++)
                P[i][j] -= fct * P[k][j], X[i][j] -= fct * X[k][j];
        }
    }
    for (int k = 0; k < 5; k++)
        for (int j = 0; j < 5; j++)
            Jinv[cp[k] * 5 + j] = X[k][j] * dinv[j];
    return ok;
}

// ---------------------------------------------------------------------------
// LUSGSMatrixInit with useRoeJacobian (EulerEvaluator.hxx:37-159, ignoreSourceTerm), one
// thread per cell: per face 0.5 alpha A_f / V_i (|A|_Roe(u_i, u_j) + lambdaVis I); a boundary
// face adds 0.5 alpha A_f / V_i (A(u_g) - |A|_Roe(u_g, u_i) - lambdaVis I) dU_g/dU_i with the
// ghost u_g at the face centre and its linearisation (velocity reversal for a no-slip wall,
// normal-momentum mirror for a slip wall or symmetry plane, zero otherwise: generateBoundaryValue
// with linMode 1, EulerEvaluator_EvaluateDt.hxx:1762-1835); then 1/dTau + 1/dt on the diagonal.
// jd and jinv are [25][N] (row-major 5x5 per cell, cell index fastest).
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(128) kLUSGSInitBlock(DevMesh m, real gamma, const real *__restrict__ dTau, real dt,
                                                       real alphaDiag, const real *__restrict__ u,
                                                       const real *__restrict__ lam, const real *__restrict__ lamRoe,
                                                       real *__restrict__ jd, real *__restrict__ jinv,
                                                       int *__restrict__ bad)
{
    const idx i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m.N)
        return;
    const size_t N = m.N, F = m.F;
    real ui[5], J[25];
#pragma unroll
    for (int v = 0; v < 5; v++)
        ui[v] = u[(size_t)v * N + i];
#pragma unroll
    for (int q = 0; q < 25; q++)
        J[q] = 0;
    const real voli = m.vol[i];
    for (int k = 0; k < m.K; k++)
    {
        const idx f = m.c2f[(size_t)k * N + i];
        if (f < 0)
            continue;
        const idx o = m.c2c[(size_t)k * N + i];
        const bool front = (m.c2flag[(size_t)k * N + i] & 1) != 0;
        real n[3], l0, l123, l4;
#pragma unroll
        for (int r = 0; r < 3; r++)
            n[r] = m.fmeanN[(size_t)r * F + f] * (front ? -1.0 : 1.0);
        faceLambdas(lam, lamRoe, F, f, front, l0, l123, l4);
        const real lvis = lam[2 * F + f];
        const real coef = (0.5 * alphaDiag) * m.farea[f] / voli;
        const bool bnd = o < 0;
        int type = 0;
        real uj[5];
        if (bnd)
        {
            const int zone = m.fzone[f];
            type = m.bcType[zone];
            bcGhost(type, ui, n, m.bcVal + (size_t)zone * 5, gamma, uj);
        }
        else
#pragma unroll
            for (int v = 0; v < 5; v++)
                uj[v] = u[(size_t)v * N + o];
        RoeAve ra;
        roeAverage(ui, uj, gamma, ra);
        // the isothermal wall linearises as the adiabatic one (the linearised ghost of
        // generateBV_ViscousWall returns before the density reset, EvaluateDt.hxx:2265-2266)
        const int typeLin = type == BC_WALL_ISO ? BC_WALL : type;
        if (bnd && (typeLin == BC_WALL || typeLin == BC_WALL_INVIS || typeLin == BC_SYM))
        {
            // (A(u_g) - |A| - lambdaVis) dU_g/dU_i, column c of dU_g/dU_i being the ghost map of e_c
            real JBC[25];
            for (int c = 0; c < 5; c++)
            {
                real e[5] = {0, 0, 0, 0, 0}, g[5];
                e[c] = 1;
                bcGhost(typeLin, e, n, m.bcVal, gamma, g);
#pragma unroll
                for (int r = 0; r < 5; r++)
                    JBC[r * 5 + c] = g[r];
            }
            real IJ[25];
            for (int c = 0; c < 5; c++)
            {
                real e[5] = {0, 0, 0, 0, 0}, dF[5], inc[5];
                e[c] = 1;
                fluxJacDu<real>(uj, n, e, lvis, gamma, dF);
                roeIncT<real>(e, n, ra, l0, l123, l4, gamma, inc);
#pragma unroll
                for (int r = 0; r < 5; r++)
                    IJ[r * 5 + c] = dF[r] - inc[r];
            }
            for (int r = 0; r < 5; r++)
                for (int c = 0; c < 5; c++)
                {
                    real acc = 0;
#pragma unroll
                    for (int q = 0; q < 5; q++)
                        acc += IJ[r * 5 + q] * JBC[q * 5 + c];
                    J[r * 5 + c] += coef * acc;
                }
        }
        for (int c = 0; c < 5; c++)
        {
            real e[5] = {0, 0, 0, 0, 0}, inc[5];
            e[c] = 1;
            roeIncT<real>(e, n, ra, l0, l123, l4, gamma, inc);
            inc[c] += lvis;
#pragma unroll
            for (int r = 0; r < 5; r++)
                J[r * 5 + c] += coef * inc[r];
        }
    }
    const real dg = 1.0 / dTau[i] + 1.0 / dt;
#pragma unroll
    for (int k = 0; k < 5; k++)
        J[k * 5 + k] += dg;
    real Ji[25];
    if (!invertBlock5(J, Ji))
        atomicOr(bad, 1);
#pragma unroll
    for (int q = 0; q < 25; q++)
    {
        if (jd)
            jd[(size_t)q * N + i] = J[q];
        jinv[(size_t)q * N + i] = Ji[q];
    }
}

void launchLUSGSInitBlock(Ctx &c, const GasParams &gp, const real *dTau, real dt, real alphaDiag, const real *u,
                          const real *lam, const real *lamRoe, real *jd, real *jinv, int *bad)
{
    KernelTimer kt(c, K_LUSGS_INIT);
    kLUSGSInitBlock<<<gridFor(c.m.N, 128), 128, 0, c.stream>>>(c.m, gp.gamma, dTau, dt, alphaDiag, u, lam, lamRoe, jd,
                                                               jinv, bad);
    DNDSR_HIP_CHECK(hipGetLastError());
}

// ---------------------------------------------------------------------------
// UpdateSGS with useRoeJacobian (EulerEvaluator.hxx:550-693), multicolour: the neighbour term
// is fluxJacobian0(u_j, u_i, n, dU_j) with the Roe dissipation and incFsign -1,
// A(u_j) dU_j - |A|_Roe(u_j, u_i) dU_j - lambdaVis dU_j, and the update x_i = JDiag_i^-1 buf
// (MatVecLeftInvert, EulerJacobian.hpp:215).  Neighbour order and the |dx| partial sums as in
// the scalar sweep.  T = double or float (every array in T).
// ---------------------------------------------------------------------------
template <typename T>
struct BlkGeom;
template <>
struct BlkGeom<real>
{
    static __device__ const real *area(const DevMesh &m) { return m.farea; }
    static __device__ const real *norm(const DevMesh &m) { return m.fmeanN; }
    static __device__ const real *vol(const DevMesh &m) { return m.vol; }
};
template <>
struct BlkGeom<float>
{
    static __device__ const float *area(const DevMesh &m) { return m.fareaF; }
    static __device__ const float *norm(const DevMesh &m) { return m.fmeanNF; }
    static __device__ const float *vol(const DevMesh &m) { return m.volF; }
};

template <typename T>
__global__ void __launch_bounds__(256) kSGSColourBlock(DevMesh m, real gammaD, T alphaDiag, const T *__restrict__ rhs,
                                                       const T *__restrict__ u, const T *__restrict__ uInc,
                                                       T *__restrict__ uIncNew, const T *__restrict__ jinv,
                                                       const T *__restrict__ lam, const T *__restrict__ lamRoe,
                                                       int forward, int gsUpdate, int uIncIsZero, idx begin, idx end,
                                                       real *__restrict__ part, int fCol, int lCol)
{
    __shared__ real sm[5][256];
    const idx t = begin + blockIdx.x * blockDim.x + threadIdx.x;
    real loc[5] = {0, 0, 0, 0, 0};
    if (t < end)
    {
        const size_t N = m.N, F = m.F;
        const idx i = m.colorCells[t];
        const T gamma = (T)gammaD;
        const T *__restrict__ farea = BlkGeom<T>::area(m);
        const T *__restrict__ fnorm = BlkGeom<T>::norm(m);
        T buf[5], ui[5];
#pragma unroll
        for (int v = 0; v < 5; v++)
            buf[v] = rhs[v * N + i], ui[v] = u[v * N + i];
        const T voli = BlkGeom<T>::vol(m)[i];
        for (int k = 0; k < m.K; k++)
        {
            const idx f = m.c2f[(size_t)k * N + i];
            const idx o = m.c2c[(size_t)k * N + i];
            const uint8_t fl = m.c2flag[(size_t)k * N + i];
            if (f < 0 || o < 0 || o == i)
                continue;
            const bool earlier = forward ? (fl & 2) != 0 : (fl & 2) == 0;
            const bool useNew = gsUpdate && earlier;
            if (uIncIsZero && !useNew)
                continue;
            // a skipped colour pass (fCol) keeps its increments in uInc, the colour the previous
            // sweep skipped (lCol) in uIncNew until its own pass
            const bool readNew = sweepReadNew(fl, fCol, lCol, m.nColor, useNew);
            const T *du = readNew ? uIncNew : uInc;
            const bool front = (fl & 1) != 0;
            T U[5], dU[5], n[3], dF[5], inc[5], l0, l123, l4;
#pragma unroll
            for (int v = 0; v < 5; v++)
                U[v] = u[v * N + o], dU[v] = du[v * N + o];
#pragma unroll
            for (int r = 0; r < 3; r++)
                n[r] = fnorm[r * F + f] * (front ? T(-1) : T(1));
            faceLambdas(lam, lamRoe, F, f, front, l0, l123, l4);
            RoeAveT<T> ra;
            roeAverage(U, ui, (real)gamma, ra);
            fluxJacDu<T>(U, n, dU, lam[2 * F + f], gamma, dF);
            roeIncT<T>(dU, n, ra, l0, l123, l4, gamma, inc);
            const T coef = (T(0.5) * alphaDiag) * farea[f] / voli;
#pragma unroll
            for (int v = 0; v < 5; v++)
                buf[v] -= coef * (dF[v] - inc[v]);
        }
#pragma unroll
        for (int r = 0; r < 5; r++)
        {
            T nw = 0;
#pragma unroll
            for (int q = 0; q < 5; q++)
                nw += jinv[(size_t)(r * 5 + q) * N + i] * buf[q];
            const T old = uIncNew[r * N + i];
            uIncNew[r * N + i] = nw;
            loc[r] = (real)fabs(nw - old);
        }
    }
#pragma unroll
    for (int v = 0; v < 5; v++)
        sm[v][threadIdx.x] = loc[v];
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1)
    {
        if (threadIdx.x < s)
#pragma unroll
            for (int v = 0; v < 5; v++)
                sm[v][threadIdx.x] += sm[v][threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0)
#pragma unroll
        for (int v = 0; v < 5; v++)
            part[(size_t)blockIdx.x * 5 + v] = sm[v][0];
}

template <typename T>
static void sgsSweepBlock(Ctx &c, const GasParams &gp, T alphaDiag, const T *rhs, const T *u, const T *uInc,
                          T *uIncNew, const T *jinv, const T *lam, const T *lamRoe, int forward, int gsUpdate,
                          int uIncIsZero, real *sumIncDev, int fCol = -1, int lCol = -1)
{
    DevMesh &m = c.m;
    const int nc = m.nColor;
    size_t off = 0;
    for (int s = 0; s < nc; s++)
    {
        const int col = forward ? s : nc - 1 - s;
        if (col == fCol)
            continue;
        const idx b = m.colorPtrHost[col], e = m.colorPtrHost[col + 1];
        if (e <= b)
            continue;
        const int nb = gridFor(e - b, 256);
        KernelTimer kt(c, K_SGS);
        kSGSColourBlock<T><<<nb, 256, 0, c.stream>>>(m, gp.gamma, alphaDiag, rhs, u, uInc, uIncNew, jinv, lam, lamRoe,
                                                    forward, gsUpdate, uIncIsZero, b, e, c.red + off * 5, fCol, lCol);
        DNDSR_HIP_CHECK(hipGetLastError());
        off += nb;
    }
    launchReduceSum5(c, (int)off, sumIncDev);
}

// A sweep of the transpose of the block matrix kSGSColourBlock sweeps: for cell i,
// y_i = JDiag_i^-T (rhs_i - sum_j O_ji^T y_j) with O_ji = (alpha/2) A_f / V_j
// (A(u_i) n_ji - |A|_Roe(u_i, u_j; n_ji) - lambdaVis I) and the face eigenvalues seen from j,
// the 5x5 block assembled column by column from the same increment routines.
__global__ void __launch_bounds__(256) kSGSColourBlockT(DevMesh m, real gamma, real alphaDiag,
                                                        const real *__restrict__ rhs, const real *__restrict__ u,
                                                        const real *__restrict__ uInc, real *__restrict__ uIncNew,
                                                        const real *__restrict__ jinv, const real *__restrict__ lam,
                                                        const real *__restrict__ lamRoe, int forward, int gsUpdate,
                                                        int uIncIsZero, idx begin, idx end, real *__restrict__ part,
                                                        int fCol, int lCol)
{
    __shared__ real sm[5][256];
    const idx t = begin + blockIdx.x * blockDim.x + threadIdx.x;
    real loc[5] = {0, 0, 0, 0, 0};
    if (t < end)
    {
        const size_t N = m.N, F = m.F;
        const idx i = m.colorCells[t];
        real buf[5], ui[5];
#pragma unroll
        for (int v = 0; v < 5; v++)
            buf[v] = rhs[v * N + i], ui[v] = u[v * N + i];
        for (int k = 0; k < m.K; k++)
        {
            const idx f = m.c2f[(size_t)k * N + i];
            const idx o = m.c2c[(size_t)k * N + i];
            const uint8_t fl = m.c2flag[(size_t)k * N + i];
            if (f < 0 || o < 0 || o == i)
                continue;
            const bool earlier = forward ? (fl & 2) != 0 : (fl & 2) == 0;
            const bool useNew = gsUpdate && earlier;
            if (uIncIsZero && !useNew)
                continue;
            const bool readNew = sweepReadNew(fl, fCol, lCol, m.nColor, useNew);
            const real *y = readNew ? uIncNew : uInc;
            const bool frontJ = (fl & 1) == 0; // the neighbour is the front cell when i is the back
            real U[5], yj[5], n[3], l0, l123, l4;
#pragma unroll
            for (int v = 0; v < 5; v++)
                U[v] = u[v * N + o], yj[v] = y[v * N + o];
#pragma unroll
            for (int r = 0; r < 3; r++)
                n[r] = m.fmeanN[r * F + f] * (frontJ ? real(-1) : real(1));
            faceLambdas(lam, lamRoe, F, f, frontJ, l0, l123, l4);
            RoeAveT<real> ra;
            roeAverage(ui, U, gamma, ra);
            const real cji = (real(0.5) * alphaDiag) * m.farea[f] / m.vol[o];
#pragma unroll
            for (int c = 0; c < 5; c++)
            {
                real e[5] = {0, 0, 0, 0, 0}, dF[5], inc[5];
                e[c] = 1;
                fluxJacDu<real>(ui, n, e, lam[2 * F + f], gamma, dF);
                roeIncT<real>(e, n, ra, l0, l123, l4, gamma, inc);
                real acc = 0;
#pragma unroll
                for (int r = 0; r < 5; r++)
                    acc += (dF[r] - inc[r]) * yj[r];
                buf[c] -= cji * acc;
            }
        }
#pragma unroll
        for (int r = 0; r < 5; r++)
        {
            real nw = 0;
#pragma unroll
            for (int q = 0; q < 5; q++)
                nw += jinv[(size_t)(q * 5 + r) * N + i] * buf[q];
            const real old = uIncNew[r * N + i];
            uIncNew[r * N + i] = nw;
            loc[r] = fabs(nw - old);
        }
    }
#pragma unroll
    for (int v = 0; v < 5; v++)
        sm[v][threadIdx.x] = loc[v];
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1)
    {
        if (threadIdx.x < s)
#pragma unroll
            for (int v = 0; v < 5; v++)
                sm[v][threadIdx.x] += sm[v][threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0)
#pragma unroll
        for (int v = 0; v < 5; v++)
            part[(size_t)blockIdx.x * 5 + v] = sm[v][0];
}

void launchSGSSweepBlockT64(Ctx &c, const GasParams &gp, real alphaDiag, const real *rhs, const real *u,
                            const real *uInc, real *uIncNew, const real *jinv, const real *lam, const real *lamRoe,
                            int forward, int gsUpdate, int uIncIsZero, real *sumIncDev, int fCol, int lCol)
{
    DevMesh &m = c.m;
    const int nc = m.nColor;
    size_t off = 0;
    for (int s = 0; s < nc; s++)
    {
        const int col = forward ? s : nc - 1 - s;
        if (col == fCol)
            continue;
        const idx b = m.colorPtrHost[col], e = m.colorPtrHost[col + 1];
        if (e <= b)
            continue;
        const int nb = gridFor(e - b, 256);
        KernelTimer kt(c, K_ADJOINT);
        kSGSColourBlockT<<<nb, 256, 0, c.stream>>>(m, gp.gamma, alphaDiag, rhs, u, uInc, uIncNew, jinv, lam, lamRoe,
                                                   forward, gsUpdate, uIncIsZero, b, e, c.red + off * 5, fCol, lCol);
        DNDSR_HIP_CHECK(hipGetLastError());
        off += nb;
    }
    launchReduceSum5(c, (int)off, sumIncDev);
}

// the increments of the cells of colour col from one array into the other
template <typename T>
__global__ void __launch_bounds__(256) kCopyColourInc(DevMesh m, idx lo, idx hi, const T *__restrict__ from,
                                                      T *__restrict__ to)
{
    const idx p = lo + blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= hi)
        return;
    const size_t N = m.N;
    const idx i = m.colorCells[p];
#pragma unroll
    for (int v = 0; v < 5; v++)
        to[(size_t)v * N + i] = from[(size_t)v * N + i];
}

// (1 + sgsIter) symmetric block sweeps from the increments in a (b: the other array), the
// result in *result.  With Gauss-Seidel updates a sweep after the first skips its first
// colour, whose pass would repeat the previous sweep's last one (same inputs, same values);
// the sweep reads that colour's increments from the array it reads, and the next sweep the
// colour from the array it writes until that colour's own pass; one copy after the last
// sweep puts the skipped colour's increments into the result.
template <typename T>
static void sgsBlockSymmetric(Ctx &c, const GasParams &gp, T alphaDiag, const T *rhs, const T *u, T *a, T *b,
                              const T *jinv, const T *lam, const T *lamRoe, int sgsIter, int gsUpdate,
                              real *sumIncDev, T **result)
{
    DevMesh &m = c.m;
    const int nC = m.nColor;
    const bool canSkip = gsUpdate && nC >= 2;
    int prevSkip = -1;
    for (int s = 0; s <= sgsIter; s++)
        for (int dir = 1; dir >= 0; dir--)
        {
            const int first = dir ? 0 : nC - 1, last = dir ? nC - 1 : 0;
            const int fCol = canSkip && !(s == 0 && dir == 1) ? first : -1;
            const int lCol = prevSkip == last ? last : -1;
            sgsSweepBlock<T>(c, gp, alphaDiag, rhs, u, a, b, jinv, lam, lamRoe, dir, gsUpdate, 0, sumIncDev, fCol,
                             lCol);
            prevSkip = fCol;
            std::swap(a, b);
        }
    if (prevSkip >= 0)
    {
        const idx lo = m.colorPtrHost[prevSkip], hi = m.colorPtrHost[prevSkip + 1];
        if (hi > lo)
        {
            KernelTimer kt(c, K_SGS);
            kCopyColourInc<T><<<gridFor(hi - lo, 256), 256, 0, c.stream>>>(m, lo, hi, b, a);
            DNDSR_HIP_CHECK(hipGetLastError());
        }
    }
    *result = a;
}
void launchCopyColourInc64(Ctx &c, int col, const real *from, real *to)
{
    const idx lo = c.m.colorPtrHost[col], hi = c.m.colorPtrHost[col + 1];
    if (hi <= lo)
        return;
    kCopyColourInc<real><<<gridFor(hi - lo, 256), 256, 0, c.stream>>>(c.m, lo, hi, from, to);
    DNDSR_HIP_CHECK(hipGetLastError());
}
void launchSGSBlockSymmetric64(Ctx &c, const GasParams &gp, real alphaDiag, const real *rhs, const real *u, real *a,
                               real *b, const real *jinv, const real *lam, const real *lamRoe, int sgsIter,
                               int gsUpdate, real *sumIncDev, real **result)
{
    sgsBlockSymmetric<real>(c, gp, alphaDiag, rhs, u, a, b, jinv, lam, lamRoe, sgsIter, gsUpdate, sumIncDev, result);
}
void launchSGSBlockSymmetric32(Ctx &c, const GasParams &gp, float alphaDiag, const float *rhs, const float *u,
                               float *a, float *b, const float *jinv, const float *lam, const float *lamRoe,
                               int sgsIter, int gsUpdate, real *sumIncDev, float **result)
{
    sgsBlockSymmetric<float>(c, gp, alphaDiag, rhs, u, a, b, jinv, lam, lamRoe, sgsIter, gsUpdate, sumIncDev, result);
}

void launchSGSSweepBlock64(Ctx &c, const GasParams &gp, real alphaDiag, const real *rhs, const real *u,
                           const real *uInc, real *uIncNew, const real *jinv, const real *lam, const real *lamRoe,
                           int forward, int gsUpdate, int uIncIsZero, real *sumIncDev)
{
    sgsSweepBlock<real>(c, gp, alphaDiag, rhs, u, uInc, uIncNew, jinv, lam, lamRoe, forward, gsUpdate, uIncIsZero,
                        sumIncDev);
}
void launchSGSSweepBlock32(Ctx &c, const GasParams &gp, float alphaDiag, const float *rhs, const float *u,
                           const float *uInc, float *uIncNew, const float *jinv, const float *lam,
                           const float *lamRoe, int forward, int gsUpdate, int uIncIsZero, real *sumIncDev)
{
    sgsSweepBlock<float>(c, gp, alphaDiag, rhs, u, uInc, uIncNew, jinv, lam, lamRoe, forward, gsUpdate, uIncIsZero,
                         sumIncDev);
}

} // namespace dndsr
