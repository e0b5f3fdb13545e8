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

#include <type_traits>

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

template <typename T>
__host__ __device__ inline T sq(T x) { return x * x; }
template <typename A, typename B>
__host__ __device__ inline A dot3(const A *a, const B *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
template <typename T>
__host__ __device__ inline real signTol(T a, T tol) { return a > tol ? 1 : (a < -tol ? -1 : 0); }
template <typename T>
__host__ __device__ inline real signP(T a) { return a >= 0 ? 1 : -1; }
template <typename T>
__host__ __device__ inline T max3(T a, T b, T c) { return fmax(a, fmax(b, c)); }

// Reciprocal for the flux arithmetic.  FP64 on the device: the hardware estimate
// (v_rcp_f64) refined by two Newton steps r <- r + r (1 - x r), each squaring the relative
// error, so the result is the correctly rounded 1/x to within an ulp -- six VALU operations
// against the ten of the IEEE division sequence.  FP32 and the host take the quotient.
template <typename T>
__host__ __device__ inline T rcpN(T x)
{
#if defined(__HIP_DEVICE_COMPILE__)
    if constexpr (sizeof(T) == 8)
    {
        double r = __builtin_amdgcn_rcp(x);
        double e = fma(-x, r, 1.0);
        r = fma(r, e, r);
        e = fma(-x, r, 1.0);
        return fma(r, e, r);
    }
#endif
    return T(1.0) / x;
}

// IdealGasThermal (src/DNDS/IdealGasPhysics.hpp:37-44).  Divisions by rho are
// taken as products with one reciprocal (FP64 division costs ~10 VALU ops on
// CDNA; the product differs from the quotient by at most an ulp).
template <typename T>
__host__ __device__ inline void thermalR(T E, T rho, T rinv, T vSqr, real gamma, T &p, T &asqr, T &H)
{
    p = (gamma - 1) * (E - rho * 0.5 * vSqr);
    asqr = gamma * p * rinv;
    H = (E + p) * rinv;
}
template <typename T>
__host__ __device__ inline void thermal(T E, T rho, T vSqr, real gamma, T &p, T &asqr, T &H)
{
    thermalR(E, rho, T(1.0) / rho, vSqr, gamma, p, asqr, H);
}

// Face-normal physical flux (GasInviscidFlux_XY, Gas.hpp:390-397)
template <typename T>
__host__ __device__ inline void physFlux(const T *U, const T *velo, const real *n, T p, T *F)
{
    T vn = dot3(velo, n);
#pragma unroll
    for (int i = 0; i < 5; i++)
        F[i] = U[i] * vn;
    F[1] += p * n[0];
    F[2] += p * n[1];
    F[3] += p * n[2];
    F[4] += vn * p;
}

// Roe average of two mean states (ComputeRoePreamble, Gas.hpp:229-259)
template <typename T>
struct RoeAveT
{
    T veloLm[3], veloRm[3], aLm, aRm, pLm, pRm;
    T veloRoe[3], vsqrRoe, HRoe, asqrRoe, rhoRoe, aRoe;
    T g1a2;  // (gamma - 1) / asqrRoe
    T inv2a; // 1 / (2 aRoe)
};
template <typename T>
__host__ __device__ inline void roeAverage(const T *ULm, const T *URm, real gamma, RoeAveT<T> &r)
{
    T HLm, HRm, aLm2, aRm2;
    const T rinvL = rcpN(ULm[0]), rinvR = rcpN(URm[0]);
#pragma unroll
    for (int d = 0; d < 3; d++)
        r.veloLm[d] = ULm[1 + d] * rinvL, r.veloRm[d] = URm[1 + d] * rinvR;
    thermalR(ULm[4], ULm[0], rinvL, dot3(r.veloLm, r.veloLm), gamma, r.pLm, aLm2, HLm);
    thermalR(URm[4], URm[0], rinvR, dot3(r.veloRm, r.veloRm), gamma, r.pRm, aRm2, HRm);
    r.aLm = sqrt(aLm2);
    r.aRm = sqrt(aRm2);
    T sL = sqrt(ULm[0]), sR = sqrt(URm[0]);
    const T sInv = rcpN(sL + sR);
#pragma unroll
    for (int d = 0; d < 3; d++)
        r.veloRoe[d] = (sL * r.veloLm[d] + sR * r.veloRm[d]) * sInv;
    r.vsqrRoe = dot3(r.veloRoe, r.veloRoe);
    r.HRoe = (sL * HLm + sR * HRm) * sInv;
    r.asqrRoe = (gamma - 1) * (r.HRoe - 0.5 * r.vsqrRoe);
    r.rhoRoe = sL * sR;
    r.aRoe = sqrt(r.asqrRoe); // NaN when asqrRoe <= 0; the NaN flag of the caller catches it
    r.g1a2 = (gamma - 1) * rcpN(r.asqrRoe);
    r.inv2a = 0.5 * rcpN(r.aRoe);
}

// Entropy fixes (Roe_EntropyFixer, Gas.hpp:878-1007)
template <int EIG, typename T>
__host__ __device__ inline void roeFix(T aL, T aR, T aAve, T uL, T uR, T uAve, T VAve, real dLambda, real fixScale,
                                       real incFScale, T &lam0, T &lam123, T &lam4)
{
    const real sHY = 0.05 * fixScale, sLD = 0.2 * fixScale, sHF = 0.25 * fixScale;
    if constexpr (EIG == 0)
    {
        auto Flim = [=](T v, T lam, T lamL, T lamR) {
            T th = max3(T(0.), lam - lamL, lamR - lam) * sHY;
            return v < th ? (sq(v) / th + th) * 0.5 : v;
        };
        lam0 = Flim(lam0, uAve - aAve, uL - aL, uR - aR);
        lam123 = Flim(lam123 * incFScale, uAve, uL, uR);
        lam4 = Flim(lam4, uAve + aAve, uL + aL, uR + aR);
    }
    else if constexpr (EIG == 1)
    {
        T uLm = signTol(uL, aL * kSmall) * fmax(fabs(uL) * incFScale, aL * sLD);
        T uRm = signTol(uR, aR * kSmall) * fmax(fabs(uR) * incFScale, aR * sLD);
        lam0 = fmax(fabs(uLm - aL), fabs(uRm - aR));
        lam123 = fmax(fabs(uLm), fabs(uRm));
        lam4 = fmax(fabs(uLm + aL), fabs(uRm + aR));
    }
    else if constexpr (EIG == 3)
    {
        T aStar = fmin(fabs(uAve) / sLD, aAve);
        lam0 = fabs(uAve * incFScale - aStar);
        lam4 = fabs(uAve * incFScale + aStar);
    }
    else if constexpr (EIG == 4)
    {
        T uStar = signTol(uAve, aAve * kSmall) * fmax(aAve * sLD, fabs(uAve) * incFScale);
        lam0 = fabs(uStar - aAve);
        lam123 = fabs(uStar);
        lam4 = fabs(uStar + aAve);
    }
    else if constexpr (EIG == 5)
    {
        T vL = uL * incFScale, vR = uR * incFScale;
        T aLm = fmin(aL, fabs(vL) / sLD), aRm = fmin(aR, fabs(vR) / sLD);
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
        T th = sHY * (VAve + aAve), th2 = sq(th);
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
template <typename T>
__host__ __device__ inline void roeDissipation(const T *dU, const real *n, const RoeAveT<T> &r, T vRN, T lam0, T lam123,
                                               T lam4, real gamma, T *incF)
{
    T dUn = dU[1] * n[0] + dU[2] * n[1] + dU[3] * n[2];
    T a23[3];
#pragma unroll
    for (int d = 0; d < 3; d++)
        a23[d] = dU[1 + d] - dU[0] * r.veloRoe[d];
    T a23n = dot3(a23, n);
#pragma unroll
    for (int d = 0; d < 3; d++)
        a23[d] -= n[d] * a23n;
    T dU4b = dU[4] - dot3(a23, r.veloRoe);
    T al1 = r.g1a2 * (dU[0] * (r.HRoe - vRN * vRN) + vRN * dUn - dU4b);
    T al0 = (dU[0] * (vRN + r.aRoe) - dUn - r.aRoe * al1) * r.inv2a;
    T al4 = dU[0] - (al0 + al1);
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
template <typename T>
__host__ __device__ inline void sideFlux(const T *U, const real *n, real gamma, T *F, T &p, T &a)
{
    const T rinv = rcpN(U[0]);
    T v[3] = {U[1] * rinv, U[2] * rinv, U[3] * rinv};
    T a2, H;
    thermalR(U[4], U[0], rinv, dot3(v, v), gamma, p, a2, H);
    a = sqrt(a2);
    physFlux(U, v, n, p, F);
}

// Full Riemann flux at one point with its own Roe average of the mean states.
// Returns lam0/123/4 (after entropy fix) like the reference.
// Point part of the Riemann flux, given the face's Roe average of the mean states
// (the average depends only on ULm / URm, so a face computes it once).
template <int RS, typename T>
__host__ __device__ inline void riemannFluxR(const T *UL, const T *UR, const RoeAveT<T> &r, const real *n,
                                             const GasParams &gp, real dLambda, T *F, T &lam0, T &lam123, T &lam4)
{
    const real gamma = gp.gamma;
    T FL[5], FR[5], pL, pR, aL, aR;
    sideFlux(UL, n, gamma, FL, pL, aL);
    sideFlux(UR, n, gamma, FR, pR, aR);
    T vRN = dot3(r.veloRoe, n);
    lam0 = fabs(vRN - r.aRoe);
    lam123 = fabs(vRN);
    lam4 = fabs(vRN + r.aRoe);
    T vLm0 = dot3(r.veloLm, n), vRm0 = dot3(r.veloRm, n);
    T dU[5];
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
        T incF[5];
        roeDissipation(dU, n, r, vRN, lam0, lam123, lam4, gamma, incF);
#pragma unroll
        for (int i = 0; i < 5; i++)
            F[i] = (FL[i] + FR[i]) * 0.5 - 0.5 * incF[i];
    }
    else if constexpr (RS == RS_HLLEP || RS == RS_HLLEP_V1)
    {
        // HLLEPFlux_IdealGas (Gas.hpp:601-714)
        T dUn = dU[1] * n[0] + dU[2] * n[1] + dU[3] * n[2];
        T a23[3];
#pragma unroll
        for (int d = 0; d < 3; d++)
            a23[d] = dU[1 + d] - dU[0] * r.veloRoe[d];
        T a23n = dot3(a23, n);
#pragma unroll
        for (int d = 0; d < 3; d++)
            a23[d] -= n[d] * a23n;
        T dU4b = dU[4] - dot3(a23, r.veloRoe);
        T al1 = r.g1a2 * (dU[0] * (r.HRoe - vRN * vRN) + vRN * dUn - dU4b);
        T SL = fmin(vRN - r.aRoe, vLm0 - aL);
        T SR = fmax(vRN + r.aRoe, vRm0 + aR);
        T dfix = r.aRoe * rcpN(r.aRoe + fmax(fabs(vRN), dLambda * gp.fixScale));
        T SP = fmax(SR, T(0.0)), SM = fmin(SL, T(0.0));
        if constexpr (RS == RS_HLLEP)
        {
            T rev[5] = {T(1), r.veloRoe[0], r.veloRoe[1], r.veloRoe[2], r.vsqrRoe * 0.5};
            T div = SP - SM;
            div += signP(div) * kVerySmall;
            const T dinv = rcpN(div), spsm = SP * SM * dinv, dal = dfix * al1;
#pragma unroll
            for (int i = 0; i < 5; i++)
                F[i] = (SP * FL[i] - SM * FR[i]) * dinv + spsm * (dU[i] - rev[i] * dal);
        }
        else
        {
            const T dinv = rcpN(SP - SM);
            lam123 = fabs(((SP + SM) * vRN - 2.0 * (1.0 - dfix) * (SP * SM)) * dinv);
            lam4 = fabs(((SP + SM) * (vRN + r.aRoe) - 2.0 * (SP * SM)) * dinv);
            lam0 = fabs(((SP + SM) * (vRN - r.aRoe) - 2.0 * (SP * SM)) * dinv);
            T incF[5];
            roeDissipation(dU, n, r, vRN, lam0, lam123, lam4, gamma, incF);
#pragma unroll
            for (int i = 0; i < 5; i++)
                F[i] = (FL[i] + FR[i]) * 0.5 - 0.5 * incF[i];
        }
    }
    else if constexpr (RS == RS_HLLC)
    {
        // HLLCFlux_IdealGas_HartenYee (Gas.hpp:733-843)
        auto q = [&](T p, T pS) { return pS <= p ? T(1.0) : sqrt(T(1) + (gamma + 1) / 2 / gamma * (pS / p - T(1))); };
        T pS = fmax(T(0.0), 0.5 * (r.pLm + r.pRm) - 0.5 * (vLm0 - vRm0) * r.rhoRoe * r.aRoe);
        T SL = vLm0 - r.aLm * q(r.pLm, pS);
        T SR = vRm0 + r.aRm * q(r.pRm, pS);
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
            T SS = 0;
            T div = UL[0] * (SL - vLm0) - UR[0] * (SR - vRm0);
            if (fabs(div) > kVerySmall)
                SS = (pR - pL + (UL[1] * n[0] + UL[2] * n[1] + UL[3] * n[2]) * (SL - vLm0) -
                      (UR[1] * n[0] + UR[2] * n[1] + UR[3] * n[2]) * (SR - vRm0)) /
                     div;
            T DS[5] = {T(0), T(n[0]), T(n[1]), T(n[2]), SS};
            bool left = SS >= 0;
            const T *U = left ? UL : UR;
            const T *Fs = left ? FL : FR;
            T S = left ? SL : SR, v0 = left ? vLm0 : vRm0, p = left ? pL : pR;
            T dv = S - SS;
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

template <int RS, typename T>
__host__ __device__ inline void riemannFlux(const T *UL, const T *UR, const T *ULm, const T *URm, const real *n,
                                            const GasParams &gp, real dLambda, T *F, T &lam0, T &lam123, T &lam4)
{
    RoeAveT<T> r;
    roeAverage(ULm, URm, gp.gamma, r);
    riemannFluxR<RS>(UL, UR, r, n, gp, dLambda, F, lam0, lam123, lam4);
}

// Batched Roe flux: eigenvalues from the cell-mean Roe average on the mean face
// normal, shared across the face points (RoeFlux_IdealGas_HartenYee_Batch,
// Gas.hpp:1272-1407).  `r`, `vRNm` and the fixed lam are computed once per face.
template <typename T>
struct RoeBatchFaceT
{
    RoeAveT<T> r;
    T vRN; // Roe velocity on the mean normal
    T lam0, lam123, lam4;
};
template <int RS, typename T>
__host__ __device__ inline void roeBatchSetup(const T *ULm, const T *URm, const real *nm, const GasParams &gp,
                                              real dLambda, RoeBatchFaceT<T> &b)
{
    roeAverage(ULm, URm, gp.gamma, b.r);
    b.vRN = dot3(b.r.veloRoe, nm);
    b.lam0 = fabs(b.vRN - b.r.aRoe);
    b.lam123 = fabs(b.vRN);
    b.lam4 = fabs(b.vRN + b.r.aRoe);
    T vLm0 = dot3(b.r.veloLm, nm), vRm0 = dot3(b.r.veloRm, nm);
    if constexpr (RS == RS_Roe_M2)
        b.lam0 = b.lam123 = b.lam4 = fmax(fabs(vLm0) + b.r.aLm, fabs(vRm0) + b.r.aRm);
    else
        roeFix<RS - 10>(b.r.aLm, b.r.aRm, b.r.aRoe, vLm0, vRm0, b.vRN, sqrt(b.r.vsqrRoe), dLambda, gp.fixScale,
                        gp.incFScale, b.lam0, b.lam123, b.lam4);
}
template <int RS, typename T>
__host__ __device__ inline void roeBatchPoint(const T *UL, const T *UR, const real *n, const GasParams &gp,
                                              const RoeBatchFaceT<T> &b, T *F)
{
    T FL[5], FR[5], pL, pR, aL, aR, dU[5];
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
        T incF[5];
        roeDissipation(dU, n, b.r, b.vRN, b.lam0, b.lam123, b.lam4, gp.gamma, incF);
#pragma unroll
        for (int i = 0; i < 5; i++)
            F[i] = (FL[i] + FR[i]) * 0.5 - 0.5 * incF[i];
    }
}

using RoeAve = RoeAveT<real>;
using RoeBatchFace = RoeBatchFaceT<real>;

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

// ---------------------------------------------------------------------------
// Boundary ghost states (generateBoundaryValue, EulerEvaluator_EvaluateDt.hxx:1743-1838),
// static frame: generateBV_FarField :1870-1950 (BCFar, BCOutP without anchors),
// generateBV_InviscidWall :2199-2215 (BCWallInvis, BCSym), generateBV_ViscousWall
// :2241-2270 (BCWall: velocity reversed; BCWallIsothermal: velocity reversed, density
// p / (Rgas T_wall)), generateBV_Outflow :2346-2353, generateBV_Inflow :2377-2390,
// generateBV_TotalConditionInflow :2417-2455 (BCInPsTs).  Codes are EulerBCType
// (src/Euler/EulerBC.hpp:35-48).  The zone values of the two gas-dependent types are stored
// prepared (dndsr_set_bc_gas): BCWallIsothermal (Rgas T_wall, ...), BCInPsTs (p0, Rgas T0,
// unit flow direction).
// ---------------------------------------------------------------------------
constexpr int BC_FAR = 1, BC_WALL = 2, BC_WALL_INVIS = 3, BC_WALL_ISO = 4, BC_OUT = 5, BC_OUTP = 6, BC_IN = 7,
              BC_IN_PSTS = 8, BC_SYM = 9;

// IdealGasThermalConservative2Primitive / Primitive2Conservative (Gas.hpp:279-317)
template <class T>
__host__ __device__ inline void cons2prim(const T *U, T *prim, real gamma)
{
    for (int k = 0; k < 5; k++)
        prim[k] = U[k] / U[0];
    const T vSqr = sq(U[1] / U[0]) + sq(U[2] / U[0]) + sq(U[3] / U[0]);
    prim[0] = U[0];
    prim[4] = (gamma - 1) * (U[4] - U[0] * 0.5 * vSqr);
}
template <class T>
__host__ __device__ inline void prim2cons(const T *prim, T *U, real gamma)
{
    for (int k = 0; k < 5; k++)
        U[k] = prim[k] * prim[0];
    const T vSqr = sq(prim[1]) + sq(prim[2]) + sq(prim[3]);
    U[0] = prim[0];
    U[4] = prim[4] / (gamma - 1) + (prim[0] * 0.5 * vSqr);
}

// the two gas-dependent ghosts (zone values prepared by dndsr_set_bc_gas):
// BCWallIsothermal far = (Rgas T_wall, ...), BCInPsTs far = (p0, Rgas T0, unit direction)
template <class T>
__host__ __device__ inline void bcGhostGasT(int type, const T *UL, const real *far, real gamma, T *UR)
{
    if (type == BC_WALL_ISO) // the mirrored state at the wall temperature
    {
        T prim[5];
        for (int k = 0; k < 5; k++)
            UR[k] = UL[k];
        for (int d = 0; d < 3; d++)
            UR[1 + d] *= -1;
        cons2prim(UR, prim, gamma);
        prim[0] = prim[4] / far[0];
        prim2cons(prim, UR, gamma);
        return;
    }
    // BC_IN_PSTS: static state of the total condition at the interior speed
    const real g1 = gamma / (gamma - 1), A = 2 * g1 * far[1]; // A = 2 Cp T0
    T vSqr = (sq(UL[1]) + sq(UL[2]) + sq(UL[3])) / sq(UL[0]);
    vSqr = fmin(vSqr, 0.95 * A);
    const T r = 1.0 - vSqr / A; // T / T0
    using ::pow;                 // the library pow for doubles; a dual number finds its own by ADL
    const T ps = far[0] * pow(r, g1);
    const T vm = sqrt(vSqr);
    T prim[5];
    prim[0] = ps / (far[1] * r);
    for (int d = 0; d < 3; d++)
        prim[1 + d] = far[2 + d] * vm;
    prim[4] = ps;
    prim2cons(prim, UR, gamma);
}
// out of line in the FP64 kernels, so that the rarely taken branch does not add its
// registers to the face and reconstruction kernels that inline bcGhost
__host__ __device__ __noinline__ inline void bcGhostGas(int type, const real *UL, const real *far, real gamma,
                                                        real *UR)
{
    bcGhostGasT<real>(type, UL, far, gamma, UR);
}

// GAS = false: a caller that knows the table has no gas-dependent zone drops that branch
template <class T, bool GAS = true>
__host__ __device__ inline void bcGhost(int type, const T *UL, const real *n, const real *far, real gamma, T *UR)
{
    if constexpr (GAS)
        if (type == BC_WALL_ISO || type == BC_IN_PSTS)
        {
            if constexpr (std::is_same<T, real>::value)
                bcGhostGas(type, UL, far, gamma, UR);
            else
                bcGhostGasT<T>(type, UL, far, gamma, UR);
            return;
        }
    switch (type)
    {
    case BC_FAR:
    case BC_OUTP:
    {
        const T un = (UL[1] * n[0] + UL[2] * n[1] + UL[3] * n[2]) / UL[0];
        T p, asqr, H;
        thermal(UL[4], UL[0], sq(UL[1] / UL[0]) + sq(UL[2] / UL[0]) + sq(UL[3] / UL[0]), gamma, p, asqr, H);
        const T a = sqrt(asqr);
        if (un - a > 0) // supersonic outflow
            for (int k = 0; k < 5; k++)
                UR[k] = UL[k];
        else if (un > 0) // subsonic outflow: exterior pressure
        {
            real fp[5];
            T lp[5];
            cons2prim(far, fp, gamma);
            cons2prim(UL, lp, gamma);
            lp[4] = fp[4];
            prim2cons(lp, UR, gamma);
        }
        else if (un + a > 0) // subsonic inflow: interior pressure
        {
            real fpr[5];
            T fp[5], lp[5];
            cons2prim(far, fpr, gamma);
            for (int k = 0; k < 5; k++)
                fp[k] = fpr[k];
            cons2prim(UL, lp, gamma);
            fp[4] = lp[4];
            prim2cons(fp, UR, gamma);
        }
        else // supersonic inflow
            for (int k = 0; k < 5; k++)
                UR[k] = far[k];
        return;
    }
    case BC_WALL_INVIS:
    case BC_SYM:
    {
        for (int k = 0; k < 5; k++)
            UR[k] = UL[k];
        const T vn = UR[1] * n[0] + UR[2] * n[1] + UR[3] * n[2];
        for (int d = 0; d < 3; d++)
            UR[1 + d] -= 2 * vn * n[d];
        return;
    }
    case BC_WALL:
        for (int k = 0; k < 5; k++)
            UR[k] = UL[k];
        for (int d = 0; d < 3; d++)
            UR[1 + d] *= -1;
        return;
    case BC_IN:
        for (int k = 0; k < 5; k++)
            UR[k] = far[k];
        return;
    default: // BC_OUT
        for (int k = 0; k < 5; k++)
            UR[k] = UL[k];
        return;
    }
}

// CompressRecPart (EulerEvaluator.hpp:1541-1631): the reconstructed point state, or the
// cell mean where it would carry negative density or internal energy
template <class T>
__host__ __device__ inline void compressRecPart(const T *umean, const T *inc, T *ret)
{
    for (int k = 0; k < 5; k++)
        ret[k] = umean[k] + inc[k];
    if (ret[0] < 0)
        for (int k = 0; k < 5; k++)
            ret[k] = umean[k];
    const T eK = (sq(ret[1]) + sq(ret[2]) + sq(ret[3])) * 0.5 / (kVerySmall + ret[0]);
    if (ret[4] - eK < 0)
        for (int k = 0; k < 5; k++)
            ret[k] = umean[k];
}

// ---- laminar viscous terms
// CpGas = Rgas gamma / (gamma - 1), src/Euler/EulerEvaluatorSettings.hpp:297
__host__ __device__ inline real cpGas(const GasParams &gp) { return gp.Rgas * gp.gamma / (gp.gamma - 1); }

// muEff, src/Euler/EulerEvaluator.hpp:751-776 (laminar part): constant mu, Sutherland's
// law at temperature T, or constant kinematic viscosity (mu = nu rho)
template <class S>
__host__ __device__ inline S muEff(const GasParams &gp, const S *U, S T)
{
    if (gp.muModel == 1)
    {
        const S TRel = T / gp.TRef;
        return gp.muGas * TRel * sqrt(TRel) * (gp.TRef + gp.CSutherland) / (T + gp.CSutherland);
    }
    if (gp.muModel == 2)
        return gp.muGas * U[0];
    return S(gp.muGas);
}

// T = p / ((gamma - 1) / gamma Cp rho), EulerEvaluator_EvaluateDt.hxx:969, :1178
template <class S>
__host__ __device__ inline S gasTemperature(const GasParams &gp, S p, S rho)
{
    return p / ((gp.gamma - 1) / gp.gamma * cpGas(gp) * rho);
}

// lamVis of EvaluateDt, EulerEvaluator_EvaluateDt.hxx:964-979: mu / rho max(4/3, gamma / Pr)
// at the face mean state; 0 for an inviscid run
__host__ __device__ inline real lamVisFace(const GasParams &gp, const real *uMean, real pMean)
{
    if (!gp.viscous)
        return 0.0;
    const real mu = muEff(gp, uMean, gasTemperature(gp, pMean, uMean[0]));
    return mu / uMean[0] * fmax(4. / 3., gp.gamma / gp.prGas);
}

// GradientCons2Prim_IdealGas, Gas.hpp:1678-1698: G[d][k] = d U_k / d x_d (conservative)
// -> Gp[d][k] = d (rho, u, v, w, p)_k / d x_d
template <class S>
__host__ __device__ inline void gradCons2Prim(const S *U, const S (*G)[5], S (*Gp)[5], real gamma)
{
    const S velo[3] = {U[1] / U[0], U[2] / U[0], U[3] / U[0]};
    const S r2 = 1.0 / sq(U[0]);
    for (int d = 0; d < 3; d++)
    {
        Gp[d][0] = G[d][0];
        for (int j = 0; j < 3; j++)
            Gp[d][1 + j] = r2 * (U[0] * G[d][1 + j] - G[d][0] * U[1 + j]);
        S a = 0, b = 0;
        for (int j = 0; j < 3; j++)
            a += G[d][1 + j] * velo[j], b += Gp[d][1 + j] * U[1 + j];
        Gp[d][4] = (gamma - 1) * (G[d][4] - 0.5 * (a + b));
    }
}

// ViscousFlux_IdealGas, Gas.hpp:1595-1655, laminar (no QCR): stress
// tau = mu (grad u + grad u^T) + lambda mu div u I with lambda = -2/3, heat flux k grad T,
// adiabatic walls drop the normal part of grad T.  Gp[d][1 + j] = d u_j / d x_d.
template <class Q>
__host__ __device__ inline void viscousFlux(const Q *U, const Q (*Gp)[5], const real *n, bool adiabatic, real gamma,
                                            Q mu, Q k, real Cp, Q *F)
{
    const real lambda = -2. / 3.;
    const Q velo[3] = {U[1] / U[0], U[2] / U[0], U[3] / U[0]};
    const Q vSqr = velo[0] * velo[0] + velo[1] * velo[1] + velo[2] * velo[2];
    const Q p = (gamma - 1) * (U[4] - U[0] * 0.5 * vSqr);
    const Q c = gamma / ((gamma - 1) * Cp * U[0] * U[0]);
    Q GradT[3];
    for (int d = 0; d < 3; d++)
        GradT[d] = c * (U[0] * Gp[d][4] - p * Gp[d][0]);
    if (adiabatic)
    {
        const Q tn = dot3(GradT, n);
        for (int d = 0; d < 3; d++)
            GradT[d] -= tn * n[d];
    }
    const Q tr = Gp[0][1] + Gp[1][2] + Gp[2][3];
    Q S[3][3];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
            S[i][j] = (Gp[i][1 + j] + Gp[j][1 + i]) * mu + (i == j ? lambda * mu * tr : Q(0.0));
    F[0] = 0;
    Q e = 0;
    for (int i = 0; i < 3; i++)
    {
        F[1 + i] = S[i][0] * n[0] + S[i][1] * n[1] + S[i][2] * n[2];
        e += (S[i][0] * velo[0] + S[i][1] * velo[1] + S[i][2] * velo[2] + k * GradT[i]) * n[i];
    }
    F[4] = e;
}

// Viscous flux at one face point from both sides' values and gradients (the GRP mean
// gradient and distance, EvaluateRHS.hxx:346-424; fluxFace EvaluateDt.hxx:1163-1215):
// dB = barycentre offset (the other cell's, or 2 x the face centre's on a boundary face,
// which the bnd flag doubles), d0 = 2 minVol / area, adiabatic for no-slip walls.
template <class T>
__host__ __device__ inline void viscousPointFlux(const GasParams &gp, const T *UL, const T *UR, const T (*GL)[5],
                                                 const T (*GR)[5], const real *n, const real *dB, real d0, bool bnd,
                                                 bool adiabatic, T *Fv)
{
    const real distBaryPerp = fabs(dot3(dB, n)) * (bnd ? 2.0 : 1.0);
    const real distGRP = fmax(fmin(distBaryPerp, d0 * 2), d0 * 0.25);
    T Um[5], Gm[3][5], Gp[3][5];
    for (int v = 0; v < 5; v++)
        Um[v] = 0.5 * (UL[v] + UR[v]);
    for (int d = 0; d < 3; d++)
        for (int v = 0; v < 5; v++)
            Gm[d][v] = (GR[d][v] + GL[d][v]) * 0.5 + (1.0 / distGRP) * (n[d] * (UR[v] - UL[v]));
    gradCons2Prim(Um, Gm, Gp, gp.gamma);
    T pM, a2M, HM;
    thermal(Um[4], Um[0], sq(Um[1] / Um[0]) + sq(Um[2] / Um[0]) + sq(Um[3] / Um[0]), gp.gamma, pM, a2M, HM);
    const T mu = muEff(gp, Um, gasTemperature(gp, pM, Um[0]));
    const real Cp = cpGas(gp);
    viscousFlux(Um, Gp, n, adiabatic, gp.gamma, mu, Cp * mu / gp.prGas, Cp, Fv);
}

} // namespace dndsr
