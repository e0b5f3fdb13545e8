// gas.hpp -- ideal-gas thermodynamics and approximate Riemann solvers, device side.
//
// Restates src/Euler/Gas.hpp for 3 velocity components (models NS / NS_3D keep
// (rho, rho u, rho v, rho w, E) also in 2-D, src/Euler/Euler.hpp).  Grid
// velocity is zero on every configuration of this path, so the vg terms drop.
// Each solver is a straight-line function of (UL, UR, ULm, URm, n) so the
// compiler keeps everything in VGPRs; the solver is selected by a template
// argument, never by a runtime branch inside the face loop.
#pragma once
#include "dev_common.hpp"

namespace dndsr
{
enum RSKind : int
{
    RS_Roe = 1,      // Gas::Roe (eigScheme 0, Harten-Yee H2)
    RS_HLLC = 2,     // Gas::HLLC
    RS_HLLEP = 3,    // Gas::HLLEP
    RS_Roe_M1 = 11,  // Roe + cLLF
    RS_Roe_M2 = 12,  // Lax-Friedrichs
    RS_Roe_M3 = 13,
    RS_Roe_M4 = 14,
    RS_Roe_M5 = 15,
    RS_Roe_M6 = 16,
    RS_Roe_M7 = 17,
    RS_Roe_M8 = 18,
    RS_HLLEP_V1 = 21,
};

__host__ __device__ inline real sq(real x) { return x * x; }
__host__ __device__ inline real dot3(const real *a, const real *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
__host__ __device__ inline real signTol(real a, real tol) { return a > tol ? 1 : (a < -tol ? -1 : 0); }
__host__ __device__ inline real signP(real a) { return a >= 0 ? 1 : -1; }
__host__ __device__ inline real max3(real a, real b, real c) { return fmax(a, fmax(b, c)); }

// IdealGasThermal (src/DNDS/IdealGasPhysics.hpp:37-44).  Divisions by rho are
// taken as products with one reciprocal (FP64 division costs ~10 VALU ops on
// CDNA; the product differs from the quotient by at most an ulp).
__host__ __device__ inline void thermalR(real E, real rho, real rinv, real vSqr, real gamma, real &p, real &asqr,
                                         real &H)
{
    p = (gamma - 1) * (E - rho * 0.5 * vSqr);
    asqr = gamma * p * rinv;
    H = (E + p) * rinv;
}
__host__ __device__ inline void thermal(real E, real rho, real vSqr, real gamma, real &p, real &asqr, real &H)
{
    thermalR(E, rho, 1.0 / rho, vSqr, gamma, p, asqr, H);
}

// Face-normal physical flux (GasInviscidFlux_XY, Gas.hpp:390-397)
__host__ __device__ inline void physFlux(const real *U, const real *velo, const real *n, real p, real *F)
{
    real vn = dot3(velo, n);
#pragma unroll
    for (int i = 0; i < 5; i++)
        F[i] = U[i] * vn;
    F[1] += p * n[0];
    F[2] += p * n[1];
    F[3] += p * n[2];
    F[4] += vn * p;
}

// Roe average of two mean states (ComputeRoePreamble, Gas.hpp:229-259)
struct RoeAve
{
    real veloLm[3], veloRm[3], aLm, aRm, pLm, pRm;
    real veloRoe[3], vsqrRoe, HRoe, asqrRoe, rhoRoe, aRoe;
    real g1a2;  // (gamma - 1) / asqrRoe
    real inv2a; // 1 / (2 aRoe)
};
__host__ __device__ inline void roeAverage(const real *ULm, const real *URm, real gamma, RoeAve &r)
{
    real HLm, HRm, aLm2, aRm2;
    const real rinvL = 1.0 / ULm[0], rinvR = 1.0 / URm[0];
#pragma unroll
    for (int d = 0; d < 3; d++)
        r.veloLm[d] = ULm[1 + d] * rinvL, r.veloRm[d] = URm[1 + d] * rinvR;
    thermalR(ULm[4], ULm[0], rinvL, dot3(r.veloLm, r.veloLm), gamma, r.pLm, aLm2, HLm);
    thermalR(URm[4], URm[0], rinvR, dot3(r.veloRm, r.veloRm), gamma, r.pRm, aRm2, HRm);
    r.aLm = sqrt(aLm2);
    r.aRm = sqrt(aRm2);
    real sL = sqrt(ULm[0]), sR = sqrt(URm[0]);
    const real sInv = 1.0 / (sL + sR);
#pragma unroll
    for (int d = 0; d < 3; d++)
        r.veloRoe[d] = (sL * r.veloLm[d] + sR * r.veloRm[d]) * sInv;
    r.vsqrRoe = dot3(r.veloRoe, r.veloRoe);
    r.HRoe = (sL * HLm + sR * HRm) * sInv;
    r.asqrRoe = (gamma - 1) * (r.HRoe - 0.5 * r.vsqrRoe);
    r.rhoRoe = sL * sR;
    r.aRoe = sqrt(r.asqrRoe); // NaN when asqrRoe <= 0; the NaN flag of the caller catches it
    r.g1a2 = (gamma - 1) / r.asqrRoe;
    r.inv2a = 0.5 / r.aRoe;
}

// Entropy fixes (Roe_EntropyFixer, Gas.hpp:878-1007)
template <int EIG>
__host__ __device__ inline void roeFix(real aL, real aR, real aAve, real uL, real uR, real uAve, real VAve,
                                       real dLambda, real fixScale, real incFScale, real &lam0, real &lam123, real &lam4)
{
    const real sHY = 0.05 * fixScale, sLD = 0.2 * fixScale, sHF = 0.25 * fixScale;
    if constexpr (EIG == 0)
    {
        auto Flim = [=](real v, real lam, real lamL, real lamR) {
            real th = max3(0., lam - lamL, lamR - lam) * sHY;
            return v < th ? (sq(v) / th + th) * 0.5 : v;
        };
        lam0 = Flim(lam0, uAve - aAve, uL - aL, uR - aR);
        lam123 = Flim(lam123 * incFScale, uAve, uL, uR);
        lam4 = Flim(lam4, uAve + aAve, uL + aL, uR + aR);
    }
    else if constexpr (EIG == 1)
    {
        real uLm = signTol(uL, aL * kSmall) * fmax(fabs(uL) * incFScale, aL * sLD);
        real uRm = signTol(uR, aR * kSmall) * fmax(fabs(uR) * incFScale, aR * sLD);
        lam0 = fmax(fabs(uLm - aL), fabs(uRm - aR));
        lam123 = fmax(fabs(uLm), fabs(uRm));
        lam4 = fmax(fabs(uLm + aL), fabs(uRm + aR));
    }
    else if constexpr (EIG == 3)
    {
        real aStar = fmin(fabs(uAve) / sLD, aAve);
        lam0 = fabs(uAve * incFScale - aStar);
        lam4 = fabs(uAve * incFScale + aStar);
    }
    else if constexpr (EIG == 4)
    {
        real uStar = signTol(uAve, aAve * kSmall) * fmax(aAve * sLD, fabs(uAve) * incFScale);
        lam0 = fabs(uStar - aAve);
        lam123 = fabs(uStar);
        lam4 = fabs(uStar + aAve);
    }
    else if constexpr (EIG == 5)
    {
        real vL = uL * incFScale, vR = uR * incFScale;
        real aLm = fmin(aL, fabs(vL) / sLD), aRm = fmin(aR, fabs(vR) / sLD);
        lam0 = fmax(fabs(vL - aLm), fabs(vR - aRm));
        lam123 = fmax(fabs(vL), fabs(vR));
        lam4 = fmax(fabs(vL + aLm), fabs(vR + aRm));
    }
    else if constexpr (EIG == 6 || EIG == 8)
    {
        lam0 = fmax(fabs(uAve * incFScale - aAve), dLambda * sHF);
        lam4 = fmax(fabs(uAve * incFScale + aAve), dLambda * sHF);
        lam123 = fmax(fabs(uAve * incFScale), dLambda * sHF);
    }
    else if constexpr (EIG == 7)
    {
        lam0 *= fabs(uAve * incFScale - aAve);
        lam4 *= fabs(uAve * incFScale + aAve);
        lam123 *= fabs(uAve * incFScale);
    }
    if constexpr (EIG == 7 || EIG == 8)
    {
        real th = sHY * (VAve + aAve), th2 = sq(th);
        if (lam0 < th)
            lam0 = (sq(lam0) + th2) / (2 * th);
        if (lam4 < th)
            lam4 = (sq(lam4) + th2) / (2 * th);
        if (lam123 < th)
            lam123 = (sq(lam123) + th2) / (2 * th);
    }
}

// Roe-eigenvector decomposition of a jump, |A| dU (Gas.hpp:1091-1125, 1209-1243).
// vRN is the Roe normal velocity used in the eigenvectors; n the point normal.
__host__ __device__ inline void roeDissipation(const real *dU, const real *n, const RoeAve &r, real vRN,
                                               real lam0, real lam123, real lam4, real gamma, real *incF)
{
    real dUn = dU[1] * n[0] + dU[2] * n[1] + dU[3] * n[2];
    real a23[3];
#pragma unroll
    for (int d = 0; d < 3; d++)
        a23[d] = dU[1 + d] - dU[0] * r.veloRoe[d];
    real a23n = dot3(a23, n);
#pragma unroll
    for (int d = 0; d < 3; d++)
        a23[d] -= n[d] * a23n;
    real dU4b = dU[4] - dot3(a23, r.veloRoe);
    real al1 = r.g1a2 * (dU[0] * (r.HRoe - vRN * vRN) + vRN * dUn - dU4b);
    real al0 = (dU[0] * (vRN + r.aRoe) - dUn - r.aRoe * al1) * r.inv2a;
    real al4 = dU[0] - (al0 + al1);
    al0 *= lam0, al1 *= lam123, al4 *= lam4;
#pragma unroll
    for (int d = 0; d < 3; d++)
        a23[d] *= lam123;
    incF[0] = al0 + al1 + al4;
    incF[4] = (r.HRoe - vRN * r.aRoe) * al0 + 0.5 * r.vsqrRoe * al1 + (r.HRoe + vRN * r.aRoe) * al4 + dot3(a23, r.veloRoe);
#pragma unroll
    for (int d = 0; d < 3; d++)
        incF[1 + d] = (r.veloRoe[d] - r.aRoe * n[d]) * al0 + (r.veloRoe[d] + r.aRoe * n[d]) * al4 + r.veloRoe[d] * al1 + a23[d];
}

// Point states -> physical fluxes, pressures, velocities
__host__ __device__ inline void sideFlux(const real *U, const real *n, real gamma, real *F, real &p, real &a)
{
    const real rinv = 1.0 / U[0];
    real v[3] = {U[1] * rinv, U[2] * rinv, U[3] * rinv};
    real a2, H;
    thermalR(U[4], U[0], rinv, dot3(v, v), gamma, p, a2, H);
    a = sqrt(a2);
    physFlux(U, v, n, p, F);
}

// Full Riemann flux at one point with its own Roe average of the mean states.
// Returns lam0/123/4 (after entropy fix) like the reference.
// Point part of the Riemann flux, given the face's Roe average of the mean states
// (the average depends only on ULm / URm, so a face computes it once).
template <int RS>
__host__ __device__ inline void riemannFluxR(const real *UL, const real *UR, const RoeAve &r, const real *n,
                                             const GasParams &gp, real dLambda, real *F, real &lam0, real &lam123,
                                             real &lam4)
{
    const real gamma = gp.gamma;
    real FL[5], FR[5], pL, pR, aL, aR;
    sideFlux(UL, n, gamma, FL, pL, aL);
    sideFlux(UR, n, gamma, FR, pR, aR);
    real vRN = dot3(r.veloRoe, n);
    lam0 = fabs(vRN - r.aRoe);
    lam123 = fabs(vRN);
    lam4 = fabs(vRN + r.aRoe);
    real vLm0 = dot3(r.veloLm, n), vRm0 = dot3(r.veloRm, n);
    real dU[5];
#pragma unroll
    for (int i = 0; i < 5; i++)
        dU[i] = UR[i] - UL[i];
    if constexpr (RS == RS_Roe_M2)
    {
        lam0 = lam123 = lam4 = fmax(fabs(vLm0) + r.aLm, fabs(vRm0) + r.aRm);
#pragma unroll
        for (int i = 0; i < 5; i++)
            F[i] = (FL[i] + FR[i]) * 0.5 - 0.5 * lam0 * dU[i];
    }
    else if constexpr (RS == RS_Roe || (RS >= RS_Roe_M1 && RS <= RS_Roe_M8))
    {
        constexpr int EIG = RS == RS_Roe ? 0 : RS - 10;
        roeFix<EIG>(r.aLm, r.aRm, r.aRoe, vLm0, vRm0, vRN, sqrt(r.vsqrRoe), dLambda, gp.fixScale, gp.incFScale,
                    lam0, lam123, lam4);
        real incF[5];
        roeDissipation(dU, n, r, vRN, lam0, lam123, lam4, gamma, incF);
#pragma unroll
        for (int i = 0; i < 5; i++)
            F[i] = (FL[i] + FR[i]) * 0.5 - 0.5 * incF[i];
    }
    else if constexpr (RS == RS_HLLEP || RS == RS_HLLEP_V1)
    {
        // HLLEPFlux_IdealGas (Gas.hpp:601-714)
        real dUn = dU[1] * n[0] + dU[2] * n[1] + dU[3] * n[2];
        real a23[3];
#pragma unroll
        for (int d = 0; d < 3; d++)
            a23[d] = dU[1 + d] - dU[0] * r.veloRoe[d];
        real a23n = dot3(a23, n);
#pragma unroll
        for (int d = 0; d < 3; d++)
            a23[d] -= n[d] * a23n;
        real dU4b = dU[4] - dot3(a23, r.veloRoe);
        real al1 = r.g1a2 * (dU[0] * (r.HRoe - vRN * vRN) + vRN * dUn - dU4b);
        real SL = fmin(vRN - r.aRoe, vLm0 - aL);
        real SR = fmax(vRN + r.aRoe, vRm0 + aR);
        real dfix = r.aRoe / (r.aRoe + fmax(fabs(vRN), dLambda * gp.fixScale));
        real SP = fmax(SR, 0.0), SM = fmin(SL, 0.0);
        if constexpr (RS == RS_HLLEP)
        {
            real rev[5] = {1, r.veloRoe[0], r.veloRoe[1], r.veloRoe[2], r.vsqrRoe * 0.5};
            real div = SP - SM;
            div += signP(div) * kVerySmall;
            const real dinv = 1.0 / div, spsm = SP * SM * dinv, dal = dfix * al1;
#pragma unroll
            for (int i = 0; i < 5; i++)
                F[i] = (SP * FL[i] - SM * FR[i]) * dinv + spsm * (dU[i] - rev[i] * dal);
        }
        else
        {
            const real dinv = 1.0 / (SP - SM);
            lam123 = fabs(((SP + SM) * vRN - 2.0 * (1.0 - dfix) * (SP * SM)) * dinv);
            lam4 = fabs(((SP + SM) * (vRN + r.aRoe) - 2.0 * (SP * SM)) * dinv);
            lam0 = fabs(((SP + SM) * (vRN - r.aRoe) - 2.0 * (SP * SM)) * dinv);
            real incF[5];
            roeDissipation(dU, n, r, vRN, lam0, lam123, lam4, gamma, incF);
#pragma unroll
            for (int i = 0; i < 5; i++)
                F[i] = (FL[i] + FR[i]) * 0.5 - 0.5 * incF[i];
        }
    }
    else if constexpr (RS == RS_HLLC)
    {
        // HLLCFlux_IdealGas_HartenYee (Gas.hpp:733-843)
        auto q = [&](real p, real pS) { return pS <= p ? 1.0 : sqrt(1 + (gamma + 1) / 2 / gamma * (pS / p - 1)); };
        real pS = fmax(0.0, 0.5 * (r.pLm + r.pRm) - 0.5 * (vLm0 - vRm0) * r.rhoRoe * r.aRoe);
        real SL = vLm0 - r.aLm * q(r.pLm, pS);
        real SR = vRm0 + r.aRm * q(r.pRm, pS);
        if (0 <= SL)
        {
#pragma unroll
            for (int i = 0; i < 5; i++)
                F[i] = FL[i];
        }
        else if (SR <= 0)
        {
#pragma unroll
            for (int i = 0; i < 5; i++)
                F[i] = FR[i];
        }
        else
        {
            real SS = 0;
            real div = UL[0] * (SL - vLm0) - UR[0] * (SR - vRm0);
            if (fabs(div) > kVerySmall)
                SS = (pR - pL + (UL[1] * n[0] + UL[2] * n[1] + UL[3] * n[2]) * (SL - vLm0) -
                      (UR[1] * n[0] + UR[2] * n[1] + UR[3] * n[2]) * (SR - vRm0)) /
                     div;
            real DS[5] = {0, n[0], n[1], n[2], SS};
            bool left = SS >= 0;
            const real *U = left ? UL : UR;
            const real *Fs = left ? FL : FR;
            real S = left ? SL : SR, v0 = left ? vLm0 : vRm0, p = left ? pL : pR;
            real dv = S - SS;
            if (fabs(dv) < kVerySmall)
            {
#pragma unroll
                for (int i = 0; i < 5; i++)
                    F[i] = Fs[i];
            }
            else
            {
#pragma unroll
                for (int i = 0; i < 5; i++)
                    F[i] = ((U[i] * S - Fs[i]) * SS + DS[i] * ((p + U[0] * (S - v0) * (SS - v0)) * S)) / dv;
            }
        }
    }
}

template <int RS>
__host__ __device__ inline void riemannFlux(const real *UL, const real *UR, const real *ULm, const real *URm,
                                            const real *n, const GasParams &gp, real dLambda, real *F,
                                            real &lam0, real &lam123, real &lam4)
{
    RoeAve r;
    roeAverage(ULm, URm, gp.gamma, r);
    riemannFluxR<RS>(UL, UR, r, n, gp, dLambda, F, lam0, lam123, lam4);
}

// Batched Roe flux: eigenvalues from the cell-mean Roe average on the mean face
// normal, shared across the face points (RoeFlux_IdealGas_HartenYee_Batch,
// Gas.hpp:1272-1407).  `r`, `vRNm` and the fixed lam are computed once per face.
struct RoeBatchFace
{
    RoeAve r;
    real vRN; // Roe velocity on the mean normal
    real lam0, lam123, lam4;
};
template <int RS>
__host__ __device__ inline void roeBatchSetup(const real *ULm, const real *URm, const real *nm, const GasParams &gp,
                                              real dLambda, RoeBatchFace &b)
{
    roeAverage(ULm, URm, gp.gamma, b.r);
    b.vRN = dot3(b.r.veloRoe, nm);
    b.lam0 = fabs(b.vRN - b.r.aRoe);
    b.lam123 = fabs(b.vRN);
    b.lam4 = fabs(b.vRN + b.r.aRoe);
    real vLm0 = dot3(b.r.veloLm, nm), vRm0 = dot3(b.r.veloRm, nm);
    if constexpr (RS == RS_Roe_M2)
        b.lam0 = b.lam123 = b.lam4 = fmax(fabs(vLm0) + b.r.aLm, fabs(vRm0) + b.r.aRm);
    else
        roeFix<RS - 10>(b.r.aLm, b.r.aRm, b.r.aRoe, vLm0, vRm0, b.vRN, sqrt(b.r.vsqrRoe), dLambda, gp.fixScale,
                        gp.incFScale, b.lam0, b.lam123, b.lam4);
}
template <int RS>
__host__ __device__ inline void roeBatchPoint(const real *UL, const real *UR, const real *n, const GasParams &gp,
                                              const RoeBatchFace &b, real *F)
{
    real FL[5], FR[5], pL, pR, aL, aR, dU[5];
    sideFlux(UL, n, gp.gamma, FL, pL, aL);
    sideFlux(UR, n, gp.gamma, FR, pR, aR);
#pragma unroll
    for (int i = 0; i < 5; i++)
        dU[i] = UR[i] - UL[i];
    if constexpr (RS == RS_Roe_M2)
    {
#pragma unroll
        for (int i = 0; i < 5; i++)
            F[i] = (FL[i] + FR[i]) * 0.5 - 0.5 * b.lam0 * dU[i];
    }
    else
    {
        real incF[5];
        roeDissipation(dU, n, b.r, b.vRN, b.lam0, b.lam123, b.lam4, gp.gamma, incF);
#pragma unroll
        for (int i = 0; i < 5; i++)
            F[i] = (FL[i] + FR[i]) * 0.5 - 0.5 * incF[i];
    }
}

// Scalar-Jacobian flux increment dF = A(U) dU . n - lambda dU
// (fluxJacobian0_Right_Times_du, EulerEvaluator.hpp:1100-1158; IdealGasUIncrement
// Gas.hpp:462-471; GasInviscidFluxFacialIncrement Gas.hpp:495-513)
template <typename T>
__host__ __device__ inline void fluxJacDu(const T *U, const T *n, const T *dU, T lambdaMain, T gamma, T *dF)
{
    T rinv = T(1) / U[0];
    T v0 = U[1] * rinv, v1 = U[2] * rinv, v2 = U[3] * rinv;
    T p = (gamma - T(1)) * (U[4] - U[0] * T(0.5) * (v0 * v0 + v1 * v1 + v2 * v2));
    T dv0 = (dU[1] - U[1] * rinv * dU[0]) * rinv;
    T dv1 = (dU[2] - U[2] * rinv * dU[0]) * rinv;
    T dv2 = (dU[3] - U[3] * rinv * dU[0]) * rinv;
    T dp = (gamma - T(1)) * (dU[4] - T(0.5) * ((dU[1] * v0 + dU[2] * v1 + dU[3] * v2) + (U[1] * dv0 + U[2] * dv1 + U[3] * dv2)));
    T vn = v0 * n[0] + v1 * n[1] + v2 * n[2];
    T dvn = dv0 * n[0] + dv1 * n[1] + dv2 * n[2];
    dF[0] = dU[1] * n[0] + dU[2] * n[1] + dU[3] * n[2];
    dF[1] = dU[1] * vn + U[1] * dvn + n[0] * dp;
    dF[2] = dU[2] * vn + U[2] * dvn + n[1] * dp;
    dF[3] = dU[3] * vn + U[3] * dvn + n[2] * dp;
    dF[4] = (dU[4] + dp) * vn + (U[4] + p) * dvn;
#pragma unroll
    for (int k = 0; k < 5; k++)
        dF[k] += -lambdaMain * dU[k];
}

} // namespace dndsr
