// setup.hip -- FiniteVolume metrics and VariationalReconstruction setup on the device.
//
// Restates, one thread per cell or face:
//   ConstructMetrics        src/CFV/VariationalReconstruction.cpp:7-44 (FiniteVolume.cpp:23-631)
//   ConstructBaseAndWeight  src/CFV/VariationalReconstruction.cpp:53-427
//   ConstructRecCoeff       src/CFV/VariationalReconstruction.cpp:507-825
// with the defaults of the path (isotropic cell-aligned h-box basis scaling,
// BaryDiff functional scale, Factorial / HQM_OPT direction weights, GWNone /
// HQM_SD geometric weights, no Green-Gauss blending, svdTolerance 0).
#include "geom.hpp"

namespace dndsr
{
// ---------------------------------------------------------------------------
// cell metrics: volume, quadrature points, w*JDet, barycentre, centre, scale, moments
// ---------------------------------------------------------------------------
template <int DIM, int NDOF>
__global__ void __launch_bounds__(128) kCellMetrics(DevMesh m)
{
    idx i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m.N)
        return;
    const int t = m.ctype ? (int)m.ctype[i] : m.cellType;
    const int nn = nodeCount(t);
    real X[8][3];
    for (int a = 0; a < nn; a++)
    {
        idx nd = m.cell2node[(size_t)a * m.N + i];
        for (int r = 0; r < 3; r++)
            X[a][r] = m.coords[(size_t)r * m.nNode + nd];
    }
    // volume with the highest rule (FiniteVolume.cpp:23-82)
    real v = 0;
    const int nMax = ruleSize(t, 6);
    for (int g = 0; g < nMax; g++)
    {
        real xi[3], w, x[3], J[3][3];
        rulePoint(t, 6, g, xi, w);
        mapLinear(t, X, xi, x, J);
        v += (1 * cellDet(DIM, J)) * w;
    }
    v += kVerySmall;
    m.vol[i] = v;
    // integration-order points, w*JDet and barycentre (FiniteVolume.cpp:84-216)
    real b[3] = {0, 0, 0};
    const int gc = ruleSize(t, m.s.intOrder);
    for (int g = gc; g < m.Gc; g++) // a mixed mesh's smaller rules are padded with zero points and weights
    {
        m.cwj[(size_t)g * m.N + i] = 0;
        for (int r = 0; r < 3; r++)
            m.cpts[((size_t)g * 3 + r) * m.N + i] = 0;
    }
    for (int g = 0; g < gc; g++)
    {
        real xi[3], w, x[3], J[3][3];
        rulePoint(t, m.s.intOrder, g, xi, w);
        mapLinear(t, X, xi, x, J);
        real jd = cellDet(DIM, J);
        m.cwj[(size_t)g * m.N + i] = (jd + kVerySmall) * w;
        for (int r = 0; r < 3; r++)
        {
            m.cpts[((size_t)g * 3 + r) * m.N + i] = x[r];
            b[r] += (x[r] * jd) * w;
        }
    }
    for (int r = 0; r < 3; r++)
        b[r] /= v, m.bary[(size_t)r * m.N + i] = b[r];
    // parametric centre (order-1 rule) (FiniteVolume.cpp:218-241)
    real cc[3];
    {
        real xi[3], J[3][3];
        centreParam(t, xi);
        mapLinear(t, X, xi, cc, J);
        for (int r = 0; r < 3; r++)
            m.cent[(size_t)r * m.N + i] = cc[r];
    }
    // aligned h-box, isotropic scale (FiniteVolume.cpp:243-267, VariationalReconstruction.hpp:331-339)
    real h[3] = {0, 0, 0};
    for (int a = 0; a < nn; a++)
        for (int r = 0; r < 3; r++)
            h[r] = fmax(h[r], fabs(X[a][r] - b[r]));
    real s = DIM == 2 ? fmax(h[0], h[1]) : fmax(h[0], fmax(h[1], h[2]));
    m.scale[i] = s;
    m.iscale[i] = 1.0 / s;
    // basis moments with the highest rule (VariationalReconstruction.cpp:86-108)
    real mm[NDOF];
    for (int k = 0; k < NDOF; k++)
        mm[k] = 0;
    for (int g = 0; g < nMax; g++)
    {
        real xi[3], w, x[3], J[3][3];
        rulePoint(t, 6, g, xi, w);
        mapLinear(t, X, xi, x, J);
        real jd = cellDet(DIM, J);
        real xs[3] = {(x[0] - cc[0]) / s, (x[1] - cc[1]) / s, DIM == 2 ? (x[2] - cc[2]) : (x[2] - cc[2]) / s};
        for (int k = 0; k < NDOF; k++)
        {
            int px, py, pz;
            monoExp(DIM, k, px, py, pz);
            mm[k] += (ipow(px, xs[0]) * ipow(py, xs[1]) * ipow(pz, xs[2]) * jd) * w;
        }
    }
    for (int k = 1; k < NDOF; k++)
        m.moment[(size_t)(k - 1) * m.N + i] = mm[k] / mm[0];
}

// ---------------------------------------------------------------------------
// face metrics: area, points, unit normals, w*JDet, mean normal, centre
// (FiniteVolume.cpp:353-575)
// ---------------------------------------------------------------------------
template <int DIM>
__global__ void __launch_bounds__(128) kFaceMetrics(DevMesh m)
{
    idx f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= m.F)
        return;
    const int t = m.faceType;
    const int nn = nodeCount(t);
    real X[4][3];
    for (int a = 0; a < nn; a++)
    {
        idx nd = m.face2node[(size_t)a * m.F + f];
        for (int r = 0; r < 3; r++)
            X[a][r] = m.coords[(size_t)r * m.nNode + nd];
    }
    real area = 0, mn[3] = {0, 0, 0};
    for (int g = 0; g < m.G; g++)
    {
        real x[3], nu[3], wj, jd, w;
        facePoint(t, m.s.intOrder, g, X, x, nu, wj, jd, w);
        area += (1 * jd) * w;
        m.fwj[(size_t)g * m.F + f] = wj;
        for (int r = 0; r < 3; r++)
        {
            m.fpts[((size_t)g * 3 + r) * m.F + f] = x[r];
            m.fnorm[((size_t)g * 3 + r) * m.F + f] = nu[r];
            mn[r] += (nu[r] * jd) * w;
        }
    }
    m.farea[f] = area + kVerySmall;
    real l = sqrt(mn[0] * mn[0] + mn[1] * mn[1] + mn[2] * mn[2]);
    real xi[3], x[3], J[3][3];
    centreParam(t, xi);
    mapLinear(t, X, xi, x, J);
    for (int r = 0; r < 3; r++)
        m.fmeanN[(size_t)r * m.F + f] = mn[r] / l, m.fcent[(size_t)r * m.F + f] = x[r];
}

// cell smooth scale (FiniteVolume.cpp:581-631): V / sum(A), then nIter rounds of
// min(self, mean(self, neighbours)), finally divided by the initial value.
__global__ void kSmoothInit(DevMesh m, real *sc, real *cur)
{
    idx i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m.N)
        return;
    real fs = 0;
    for (int k = 0; k < m.K; k++)
    {
        idx f = m.c2f[(size_t)k * m.N + i];
        if (f >= 0)
            fs += m.farea[f];
    }
    sc[i] = m.vol[i] / (fs + kVerySmall);
    cur[i] = sc[i];
}
__global__ void kSmoothIter(DevMesh m, const real *cur, real *nxt)
{
    idx i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m.N)
        return;
    real nAdj = 0, sum = 0;
    for (int k = 0; k < m.K; k++)
    {
        idx f = m.c2f[(size_t)k * m.N + i];
        if (f < 0)
            continue;
        idx o = m.f2c[f] == i ? m.f2c[(size_t)m.F + f] : m.f2c[f];
        if (o >= 0)
            nAdj += 1., sum += cur[o];
    }
    sum += nAdj * 1.0 * cur[i];
    sum /= nAdj * (1 + 1.0);
    nxt[i] = fmin(cur[i], sum);
}
__global__ void kSmoothFinal(DevMesh m, const real *sc, const real *cur)
{
    idx i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m.N)
        return;
    m.smooth[i] = cur[i] / sc[i];
}

// face functional scale and weights (VariationalReconstruction.cpp:239-427)
template <int DIM>
__global__ void kFaceWeight(DevMesh m)
{
    idx f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= m.F)
        return;
    idx cL = m.f2c[f], cR = m.f2c[(size_t)m.F + f];
    real bd[3];
    if (cR >= 0)
        for (int r = 0; r < 3; r++)
            bd[r] = (m.bary[(size_t)r * m.N + cR] - m.fshift[(size_t)r * m.F + f]) - m.bary[(size_t)r * m.N + cL];
    else
    {
        real vCB[3], n[3];
        for (int r = 0; r < 3; r++)
            vCB[r] = m.fcent[(size_t)r * m.F + f] - m.bary[(size_t)r * m.N + cL], n[r] = m.fmeanN[(size_t)r * m.F + f];
        real pr = vCB[0] * n[0] + vCB[1] * n[1] + vCB[2] * n[2];
        for (int r = 0; r < 3; r++)
            bd[r] = 2 * pr * n[r];
    }
    real l2 = 0;
    for (int r = 0; r < DIM; r++)
        l2 += bd[r] * bd[r];
    m.faceL[f] = sqrt(l2) * m.s.scaleMultiplier;
    real wg = 1;
    if (m.s.geomWeightScheme == 1)
        wg = m.s.geomWeightBias +
             pow(pow(m.farea[f], 1. / real(DIM - 1)) / sqrt(bd[0] * bd[0] + bd[1] * bd[1] + bd[2] * bd[2]),
                 m.s.geomWeightPower * 0.5);
    const real fac[4] = {1, 1, 2, 6};
    for (int p = 0; p <= m.P; p++)
    {
        real wd;
        if (m.s.dirWeightScheme == 0)
            wd = 1. / fac[p];
        else
        {
            // HQM_OPT (VariationalReconstruction.cpp:340-356)
            if (m.P == 1)
                wd = 1;
            else if (m.P == 2)
                wd = p == 0 ? 1 : (p == 1 ? 0.4643 : 0.1559);
            else
                wd = p == 0 ? 1 : (p == 1 ? .5295 : .2117);
        }
        // boundary faces: id2faceDircWeight(zone, order) * bcWeight, VariationalReconstruction.cpp:438-441
        if (cR < 0)
            wd *= (m.zoneW ? m.zoneW[(size_t)m.fzone[f] * (m.P + 1) + p] : 1.0) * m.s.bcWeight;
        m.faceW[(size_t)p * m.F + f] = wd * wg;
    }
}

// ---------------------------------------------------------------------------
// Reconstruction coefficients: one thread per cell.
// D-matrices (NDIFF x M) of basis derivatives are built per quadrature point;
// the functional (FFaceFunctional + AccumulateDiffOrderContributions,
// VariationalReconstruction.hpp:32-93, 459-505) sums the order blocks from the
// highest derivative order down, each weighted by w_o^2 * faceL^(2o) and the
// multinomial counts diffNCombs (BaseFunction.hpp:216-220).
// ---------------------------------------------------------------------------
template <int DIM, int P>
struct VRDims
{
    static constexpr int NDOF = nDof(DIM, P);
    static constexpr int M = NDOF - 1;
    static constexpr int NDIFF = NDOF;
};

__host__ __device__ inline int orderOfRow(int dim, int k)
{
    if (dim == 2)
        return k == 0 ? 0 : (k < 3 ? 1 : (k < 6 ? 2 : 3));
    return k == 0 ? 0 : (k < 4 ? 1 : (k < 10 ? 2 : 3));
}
__host__ __device__ inline int nComb(int dim, int k)
{
    const int c2[10] = {1, 1, 1, 1, 2, 1, 1, 3, 3, 1};
    const int c3[20] = {1, 1, 1, 1, 1, 1, 1, 2, 2, 2, 1, 1, 1, 3, 3, 3, 3, 3, 3, 6};
    return dim == 2 ? c2[k] : c3[k];
}

template <int DIM, int P>
__device__ inline void diffMatrix(const DevMesh &m, idx c, const real *p, int nDiff, real *D /*[nDiff][M]*/)
{
    using Dm = VRDims<DIM, P>;
    real s = m.scale[c];
    real xs[3];
    for (int r = 0; r < 3; r++)
        xs[r] = (p[r] - m.cent[(size_t)r * m.N + c]);
    xs[0] /= s, xs[1] /= s;
    if (DIM == 3)
        xs[2] /= s;
    for (int id = 0; id < nDiff; id++)
        for (int j = 0; j < Dm::M; j++)
        {
            real v = basisDiff(DIM, id, j + 1, xs, s);
            if (id == 0)
                v -= m.moment[(size_t)j * m.N + c];
            D[id * Dm::M + j] = v;
        }
}

// Conj[M x nJ] += fac * sum_o w_o^2 L^(2o) sum_{k in o} DI[k][i] DJ[k][j] comb_k
template <int DIM, int P>
__device__ inline void functionalAcc(const DevMesh &m, idx f, const real *DI, const real *DJ, int nJ, int nDiff,
                                     real fac, real *out, size_t strideOut, bool toGlobal, real *outLocal)
{
    using Dm = VRDims<DIM, P>;
    int maxO = orderOfRow(DIM, nDiff - 1);
    real fL = m.faceL[f];
    real fL2 = fL * fL;
    real sc[4];
    for (int o = 0; o <= maxO; o++)
    {
        real w = m.faceW[(size_t)o * m.F + f];
        real pw = o == 0 ? 1.0 : (o == 1 ? fL2 : (o == 2 ? fL2 * fL2 : fL2 * fL2 * fL2));
        sc[o] = w * w * pw;
    }
    for (int i = 0; i < Dm::M; i++)
        for (int j = 0; j < nJ; j++)
        {
            real tot = 0;
            for (int o = maxO; o >= 0; o--)
            {
                real r = 0;
                for (int k = 0; k < nDiff; k++)
                    if (orderOfRow(DIM, k) == o)
                        r += DI[k * Dm::M + i] * DJ[k * nJ + j] * nComb(DIM, k);
                tot += r * sc[o];
            }
            if (toGlobal)
                out[((size_t)i * nJ + j) * strideOut] += tot * fac;
            else
                outLocal[i * nJ + j] += tot * fac;
        }
}

template <int DIM, int P>
__global__ void __launch_bounds__(64) kVRCoeff(DevMesh m, real *Ascr, real *Ainv)
{
    using Dm = VRDims<DIM, P>;
    constexpr int M = Dm::M, NDIFF = Dm::NDIFF;
    idx i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m.N)
        return;
    const size_t N = m.N;
    real DI[NDIFF * M], DJ[NDIFF * M];
    for (int q = 0; q < M * M; q++)
        Ascr[q * N + i] = 0;
    // A
    for (int k = 0; k < m.K; k++)
    {
        idx f = m.c2f[(size_t)k * N + i];
        if (f < 0)
            continue;
        int side = m.f2c[f] == i ? 0 : 1;
        for (int g = 0; g < m.G; g++)
        {
            real p[3];
            for (int r = 0; r < 3; r++)
                p[r] = m.fpts[((size_t)g * 3 + r) * m.F + f] + (side ? m.fshift[(size_t)r * m.F + f] : 0.0);
            diffMatrix<DIM, P>(m, i, p, NDIFF, DI);
            functionalAcc<DIM, P>(m, f, DI, DI, M, NDIFF, m.fwj[(size_t)g * m.F + f], Ascr + i, N, true, nullptr);
        }
    }
    // B_f into AinvB block 1+k, b_f into Ainvb slot k
    for (int k = 0; k < m.K; k++)
    {
        idx f = m.c2f[(size_t)k * N + i];
        real *B = m.AinvB + ((size_t)(1 + k) * M * M) * N + i;
        real *bv = m.Ainvb + ((size_t)k * M) * N + i;
        for (int q = 0; q < M * M; q++)
            B[q * N] = 0;
        for (int q = 0; q < M; q++)
            bv[q * N] = 0;
        if (f < 0)
            continue;
        int side = m.f2c[f] == i ? 0 : 1;
        idx o = side == 0 ? m.f2c[(size_t)m.F + f] : m.f2c[f];
        for (int g = 0; g < m.G; g++)
        {
            real pI[3], pJ[3];
            for (int r = 0; r < 3; r++)
            {
                real x = m.fpts[((size_t)g * 3 + r) * m.F + f], sh = m.fshift[(size_t)r * m.F + f];
                pI[r] = x + (side ? sh : 0.0);
                pJ[r] = x + (side ? 0.0 : sh);
            }
            real fac = m.fwj[(size_t)g * m.F + f];
            diffMatrix<DIM, P>(m, i, pI, NDIFF, DI);
            if (o >= 0)
            {
                diffMatrix<DIM, P>(m, o, pJ, NDIFF, DJ);
                functionalAcc<DIM, P>(m, f, DI, DJ, M, NDIFF, fac, B, N, true, nullptr);
            }
            real one = 1.0;
            functionalAcc<DIM, P>(m, f, DI, &one, 1, 1, fac, bv, N, true, nullptr);
        }
        if (o == i) // a face coinciding with itself through periodicity
            for (int q = 0; q < M * M; q++)
                Ascr[q * N + i] -= B[q * N], B[q * N] = 0;
    }
    // Gauss-Jordan inverse with partial pivoting (A is SPD; the reference uses an
    // SVD pseudo-inverse, HardEigen.cpp:13-23, which coincides for full rank)
    real *A = Ascr + i;
    real *R = Ainv + i;
    for (int r = 0; r < M; r++)
        for (int c2 = 0; c2 < M; c2++)
            R[(r * M + c2) * N] = r == c2 ? 1.0 : 0.0;
    for (int c2 = 0; c2 < M; c2++)
    {
        int piv = c2;
        real best = fabs(A[(c2 * M + c2) * N]);
        for (int r = c2 + 1; r < M; r++)
        {
            real v = fabs(A[(r * M + c2) * N]);
            if (v > best)
                best = v, piv = r;
        }
        if (piv != c2)
            for (int k = 0; k < M; k++)
            {
                real t0 = A[(c2 * M + k) * N];
                A[(c2 * M + k) * N] = A[(piv * M + k) * N];
                A[(piv * M + k) * N] = t0;
                real t1 = R[(c2 * M + k) * N];
                R[(c2 * M + k) * N] = R[(piv * M + k) * N];
                R[(piv * M + k) * N] = t1;
            }
        real d = 1.0 / A[(c2 * M + c2) * N];
        for (int k = 0; k < M; k++)
            A[(c2 * M + k) * N] *= d, R[(c2 * M + k) * N] *= d;
        for (int r = 0; r < M; r++)
        {
            if (r == c2)
                continue;
            real fr = A[(r * M + c2) * N];
            if (fr == 0)
                continue;
            for (int k = 0; k < M; k++)
                A[(r * M + k) * N] -= fr * A[(c2 * M + k) * N], R[(r * M + k) * N] -= fr * R[(c2 * M + k) * N];
        }
    }
    // blocks: A^-1, A^-1 B_f, A^-1 b_f
    for (int q = 0; q < M * M; q++)
        m.AinvB[(size_t)q * N + i] = R[q * N];
    real tmp[M];
    for (int k = 0; k < m.K; k++)
    {
        idx f = m.c2f[(size_t)k * N + i];
        real *B = m.AinvB + ((size_t)(1 + k) * M * M) * N + i;
        real *bv = m.Ainvb + ((size_t)k * M) * N + i;
        bool bnd = f < 0 || (m.f2c[f] == i ? m.f2c[(size_t)m.F + f] : m.f2c[f]) < 0;
        if (bnd)
        {
            for (int q = 0; q < M * M; q++)
                B[q * N] = 0;
            for (int q = 0; q < M; q++)
                bv[q * N] = 0;
            continue;
        }
        for (int c2 = 0; c2 < M; c2++)
        {
            for (int r = 0; r < M; r++)
                tmp[r] = B[(r * M + c2) * N];
            for (int r = 0; r < M; r++)
            {
                real acc = 0;
                for (int k2 = 0; k2 < M; k2++)
                    acc += R[(r * M + k2) * N] * tmp[k2];
                B[(r * M + c2) * N] = acc;
            }
        }
        for (int r = 0; r < M; r++)
            tmp[r] = bv[r * N];
        for (int r = 0; r < M; r++)
        {
            real acc = 0;
            for (int k2 = 0; k2 < M; k2++)
                acc += R[(r * M + k2) * N] * tmp[k2];
            bv[r * N] = acc;
        }
    }
}

template <int DIM, int P>
static void runSetup(Ctx &c)
{
    DevMesh &m = c.m;
    constexpr int NDOF = nDof(DIM, P);
    hipStream_t st = c.stream;
    kCellMetrics<DIM, NDOF><<<gridFor(m.N, 128), 128, 0, st>>>(m);
    DNDSR_HIP_CHECK(hipGetLastError());
    kFaceMetrics<DIM><<<gridFor(m.F, 128), 128, 0, st>>>(m);
    DNDSR_HIP_CHECK(hipGetLastError());
    real *sc, *a, *b;
    DNDSR_HIP_CHECK(hipMalloc(&sc, sizeof(real) * m.N));
    DNDSR_HIP_CHECK(hipMalloc(&a, sizeof(real) * m.N));
    DNDSR_HIP_CHECK(hipMalloc(&b, sizeof(real) * m.N));
    kSmoothInit<<<gridFor(m.N, 256), 256, 0, st>>>(m, sc, a);
    for (int it = 1; it <= m.s.nIterCellSmoothScale; it++)
    {
        kSmoothIter<<<gridFor(m.N, 256), 256, 0, st>>>(m, a, b);
        std::swap(a, b);
    }
    kSmoothFinal<<<gridFor(m.N, 256), 256, 0, st>>>(m, sc, a);
    DNDSR_HIP_CHECK(hipGetLastError());
    kFaceWeight<DIM><<<gridFor(m.F, 256), 256, 0, st>>>(m);
    DNDSR_HIP_CHECK(hipGetLastError());
    constexpr int M = NDOF - 1;
    real *Ascr, *Ainv;
    DNDSR_HIP_CHECK(hipMalloc(&Ascr, sizeof(real) * (size_t)M * M * m.N));
    DNDSR_HIP_CHECK(hipMalloc(&Ainv, sizeof(real) * (size_t)M * M * m.N));
    kVRCoeff<DIM, P><<<gridFor(m.N, 64), 64, 0, st>>>(m, Ascr, Ainv);
    DNDSR_HIP_CHECK(hipGetLastError());
    DNDSR_HIP_CHECK(hipStreamSynchronize(st));
    DNDSR_HIP_CHECK(hipFree(Ascr));
    DNDSR_HIP_CHECK(hipFree(Ainv));
    DNDSR_HIP_CHECK(hipFree(sc));
    DNDSR_HIP_CHECK(hipFree(a));
    DNDSR_HIP_CHECK(hipFree(b));
}

void setupMetricsAndVR(Ctx &c)
{
    const DevMesh &m = c.m;
    if (m.dim == 2)
    {
        if (m.P == 1)
            runSetup<2, 1>(c);
        else if (m.P == 2)
            runSetup<2, 2>(c);
        else if (m.P == 3)
            runSetup<2, 3>(c);
        else
            throw std::runtime_error("maxOrder must be 1..3");
    }
    else
    {
        if (m.P == 1)
            runSetup<3, 1>(c);
        else if (m.P == 2)
            runSetup<3, 2>(c);
        else if (m.P == 3)
            runSetup<3, 3>(c);
        else
            throw std::runtime_error("maxOrder must be 1..3");
    }
}

} // namespace dndsr
