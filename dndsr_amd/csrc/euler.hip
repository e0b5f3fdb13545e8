// euler.hip -- EulerEvaluator hot path on the device: EvaluateDt, EvaluateRHS,
// LUSGSMatrixInit, the multicolour symmetric Gauss-Seidel sweeps and the
// SSP-RK3 stage combination.
#include <type_traits>
#include "gas.hpp"
#include "internal.hpp"

namespace dndsr
{
// ---------------------------------------------------------------------------
// EvaluateDt (EulerEvaluator_EvaluateDt.hxx:840-1048), inviscid.
// Face pass: spectral radius lambdaConvection = max(a_L + |v_L.n|, a_R + |v_R.n|)
// and deltaLambdaFace; cell pass: dt = CFL V smooth / (sum lambda A),
// deltaLambdaCell = max over faces; then a two-stage min for dtMin.
// lambdaFace layout [4][F]: lambdaFace, lambdaFaceC, lambdaFaceVis, lambdaFace0(=123=4).
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) kDtFace(DevMesh m, GasParams gp, const real *__restrict__ u,
                                               real *__restrict__ lam, real *__restrict__ dLamF)
{
    const real gamma = gp.gamma;
    idx f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= m.F)
        return;
    const size_t N = m.N, F = m.F;
    idx cL = m.f2c[f], cR = m.f2c[F + f];
    idx cR2 = cR >= 0 ? cR : cL;
    real UL[5], UR[5], n[3];
#pragma unroll
    for (int v = 0; v < 5; v++)
        UL[v] = u[v * N + cL], UR[v] = u[v * N + cR2];
#pragma unroll
    for (int r = 0; r < 3; r++)
        n[r] = m.fmeanN[r * F + f];
    real vL[3] = {UL[1] / UL[0], UL[2] / UL[0], UL[3] / UL[0]};
    real vR[3] = {UR[1] / UR[0], UR[2] / UR[0], UR[3] / UR[0]};
    real pL, aL2, HL, pR, aR2, HR;
    thermal(UL[4], UL[0], dot3(vL, vL), gamma, pL, aL2, HL);
    thermal(UR[4], UR[0], dot3(vR, vR), gamma, pR, aR2, HR);
    real vnL = dot3(vL, n), vnR = dot3(vR, n);
    real vs[3] = {vL[0] + vR[0], vL[1] + vR[1], vL[2] + vR[2]};
    real vnMean = 0.5 * dot3(vs, n);
    real aL = sqrt(aL2), aR = sqrt(aR2);
    real lc = fmax(aL + fabs(vnL), aR + fabs(vnR));
    real visc = 0.0; // inviscid: lamVis = 0
    if (gp.viscous)
    {
        // lamVis at the mean of the two cell means (a boundary face has UR = UL),
        // times A (1 / V_L + 1 / V_R), EvaluateDt.hxx:964-992
        real uMean[5];
#pragma unroll
        for (int k = 0; k < 5; k++)
            uMean[k] = (UL[k] + UR[k]) * 0.5;
        const real lamVis = lamVisFace(gp, uMean, (pL + pR) * 0.5);
        visc = lamVis * m.farea[f] * (1. / m.vol[cL] + 1. / m.vol[cR2]);
    }
    lam[f] = lc + visc;
    lam[F + f] = fabs(vnMean) + visc;
    lam[2 * F + f] = visc;
    lam[3 * F + f] = lc;
    real dv[3] = {vR[0] - vL[0], vR[1] - vL[1], vR[2] - vL[2]};
    real dvn = dot3(dv, n);
    dLamF[f] = max3(fabs(dvn + aR - aL), fabs(dvn - aR + aL), fabs(aL - aR));
}

__global__ void __launch_bounds__(256) kDtCell(DevMesh m, real CFL, real maxDt, const real *__restrict__ lam,
                                               const real *__restrict__ dLamF, real *__restrict__ dt,
                                               real *__restrict__ dlc, real *__restrict__ partMin)
{
    __shared__ real sm[256];
    idx i = blockIdx.x * blockDim.x + threadIdx.x;
    real d = 3e200;
    if (i < m.N)
    {
        real lc = 0, dl = 0;
        for (int k = 0; k < m.K; k++)
        {
            idx f = m.c2f[(size_t)k * m.N + i];
            if (f < 0)
                continue;
            lc += lam[f] * m.farea[f];
            dl = fmax(dl, dLamF[f]);
        }
        d = fmin(CFL * m.vol[i] * m.smooth[i] / (lc + 1e-100), maxDt);
        if (dt)
            dt[i] = d;
        if (dlc)
            dlc[i] = dl;
    }
    sm[threadIdx.x] = d;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1)
    {
        if (threadIdx.x < s)
            sm[threadIdx.x] = fmin(sm[threadIdx.x], sm[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0)
        partMin[blockIdx.x] = sm[0];
}

// single-block reductions over nPart partials (min or sum over `width` interleaved values)
__global__ void __launch_bounds__(256) kReduceMin(const real *part, int nPart, real *out)
{
    __shared__ real sm[256];
    real d = 3e200;
    for (int q = threadIdx.x; q < nPart; q += 256)
        d = fmin(d, part[q]);
    sm[threadIdx.x] = d;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1)
    {
        if (threadIdx.x < s)
            sm[threadIdx.x] = fmin(sm[threadIdx.x], sm[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0)
        out[0] = sm[0];
}
__global__ void kFillFromScalar(idx n, const real *s, real *out)
{
    idx i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        out[i] = s[0];
}

void launchEvalDt(Ctx &c, const GasParams &gp, const real *u, real CFL, real maxDt, int useLocalDt, real *dt,
                  real *lambdaFace, real *deltaLambdaCell, real *dtMinDev)
{
    DevMesh &m = c.m;
    {
        KernelTimer kt(c, K_DT_FACE);
        kDtFace<<<gridFor(m.F, 256), 256, 0, c.stream>>>(m, gp, u, lambdaFace, c.faceLamTmp);
        DNDSR_HIP_CHECK(hipGetLastError());
    }
    int nb = gridFor(m.N, 256);
    {
        KernelTimer kt(c, K_DT_CELL);
        kDtCell<<<nb, 256, 0, c.stream>>>(m, CFL, maxDt, lambdaFace, c.faceLamTmp, dt, deltaLambdaCell, c.red);
        DNDSR_HIP_CHECK(hipGetLastError());
    }
    kReduceMin<<<1, 256, 0, c.stream>>>(c.red, nb, dtMinDev);
    DNDSR_HIP_CHECK(hipGetLastError());
    if (!useLocalDt && dt)
    {
        kFillFromScalar<<<nb, 256, 0, c.stream>>>(m.N, dtMinDev, dt);
        DNDSR_HIP_CHECK(hipGetLastError());
    }
}

// ---------------------------------------------------------------------------
// EvaluateRHS (EulerEvaluator_EvaluateRHS.hxx:171-662), inviscid, interior and
// periodic faces.  Face pass: one thread per face evaluates the reconstructed
// states U = u + D0phi(x_g) uRec on both sides at every face quadrature point
// (basis recomputed from the cell centre, scale and moments), the Riemann flux,
// and integrates -F w JDet.  Gather pass (the reference's face-buffer + cell
// gather form, :640-662): rhs_i = sum over cell2face of +-F_f / V_i.  No atomics.
// ---------------------------------------------------------------------------
// Face tile kernel: 64 consecutive faces per block, one wavefront per face
// quadrature point (lane = face, wave = point).  Everything a point needs from
// the two cells -- means, reconstruction coefficients, centre, 1/scale, basis
// moments -- and the face's periodic shift is staged once into LDS rows of 64
// doubles by waves 1..G-1 (coalesced row loads; all loads of a lane are issued
// before the first wait), while wave 0 forms the face constants from the means
// (Roe average, batched-Roe eigenvalues, dLambda) and parks them in LDS.  Each
// lane's own point geometry is loaded into registers before the barrier.  After
// it every lane evaluates the basis on both sides, reconstructs, evaluates the
// flux and the G point fluxes of a face are summed in point order through LDS.
constexpr int kTileF = 64;
constexpr int kRoeRows = 20; // RoeAve fields
constexpr int kFaceRows = kRoeRows + 5; // + batch vRN, lam0, lam123, lam4, and dLambda

__device__ inline void roeStore(real *lds, int row0, int lane, const RoeAve &r)
{
    const real v[kRoeRows] = {r.veloLm[0], r.veloLm[1], r.veloLm[2], r.veloRm[0], r.veloRm[1], r.veloRm[2],
                              r.aLm, r.aRm, r.pLm, r.pRm, r.veloRoe[0], r.veloRoe[1], r.veloRoe[2],
                              r.vsqrRoe, r.HRoe, r.asqrRoe, r.rhoRoe, r.aRoe, r.g1a2, r.inv2a};
#pragma unroll
    for (int k = 0; k < kRoeRows; k++)
        lds[(row0 + k) * kTileF + lane] = v[k];
}
__device__ inline void roeLoad(const real *lds, int row0, int lane, RoeAve &r)
{
    real v[kRoeRows];
#pragma unroll
    for (int k = 0; k < kRoeRows; k++)
        v[k] = lds[(row0 + k) * kTileF + lane];
    r.veloLm[0] = v[0], r.veloLm[1] = v[1], r.veloLm[2] = v[2];
    r.veloRm[0] = v[3], r.veloRm[1] = v[4], r.veloRm[2] = v[5];
    r.aLm = v[6], r.aRm = v[7], r.pLm = v[8], r.pRm = v[9];
    r.veloRoe[0] = v[10], r.veloRoe[1] = v[11], r.veloRoe[2] = v[12];
    r.vsqrRoe = v[13], r.HRoe = v[14], r.asqrRoe = v[15], r.rhoRoe = v[16], r.aRoe = v[17];
    r.g1a2 = v[18], r.inv2a = v[19];
}

// LDS row map of the face tile (rows of kTileF doubles)
template <int NDOF>
struct RHSRows
{
    static constexpr int M = NDOF - 1;
    static constexpr int MEAN = 0;                 // [side][5]
    static constexpr int REC = 10;                 // [side][5][M]
    static constexpr int MET = REC + 10 * M;       // [side][cent 3, 1/scale, moment M]
    static constexpr int METW = 4 + M;
    static constexpr int SHIFT = MET + 2 * METW;   // [3]
    static constexpr int STAGED = SHIFT + 3;       // rows loaded from global memory
    static constexpr int FC = STAGED;              // face constants
    static constexpr int TOTAL = FC + kFaceRows;
};

// Staging of the face tile.  The row -> (array, row of the array, side) map is
// resolved at compile time: one dispatch on the staging wave's index, then every
// row is a constant, so a row costs a few scalar address operations and one
// buffer load.  Loads go out in chunks of CH rows that are in flight together.
template <int NDOF>
struct RHSStageCtx
{
    const DevMesh *m;
    const real *u, *uRec;
    real *lds;
    int lane;
    uint32_t fo, oL, oR; // byte offsets of the face and its two cells
};

template <int NDOF, int Q>
__device__ __forceinline__ real rhsLoadRow(const RHSStageCtx<NDOF> &c)
{
    using R = RHSRows<NDOF>;
    constexpr int M = R::M;
    const size_t N = c.m->N, F = c.m->F;
    const uint64_t rowN = (uint64_t)N * 8;
    if constexpr (Q < R::REC)
        return bld(rowRsrc(c.u + (size_t)(Q % 5) * N, rowN), Q < 5 ? c.oL : c.oR);
    else if constexpr (Q < R::MET)
    {
        constexpr int r = Q - R::REC, side = r / (5 * M), vj = r % (5 * M);
        // without coefficients (first-order residual) the rows read as 0
        return c.uRec ? bld(rowRsrc(c.uRec + (size_t)vj * N, rowN), side ? c.oR : c.oL)
                      : bld(rowRsrc(c.u, rowN), kNoRow);
    }
    else if constexpr (Q < R::SHIFT)
    {
        constexpr int r = Q - R::MET, side = r / R::METW, k = r % R::METW;
        const real *base = k < 3 ? c.m->cent + (size_t)k * N
                                 : (k == 3 ? c.m->iscale : c.m->moment + (size_t)(k - 4) * N);
        return bld(rowRsrc(base, rowN), side ? c.oR : c.oL);
    }
    else
        return bld(rowRsrc(c.m->fshift + (size_t)(Q - R::SHIFT) * F, (uint64_t)F * 8), c.fo);
}

template <int NDOF, int W, int NS, int T0, int I, int CH>
__device__ __forceinline__ void rhsLoadSeq(const RHSStageCtx<NDOF> &c, real *val)
{
    if constexpr (I < CH)
    {
        constexpr int Q = W + (T0 + I) * NS;
        if constexpr (Q < RHSRows<NDOF>::STAGED)
            val[I] = rhsLoadRow<NDOF, Q>(c);
        rhsLoadSeq<NDOF, W, NS, T0, I + 1, CH>(c, val);
    }
}
template <int NDOF, int W, int NS, int T0, int I, int CH>
__device__ __forceinline__ void rhsStoreSeq(const RHSStageCtx<NDOF> &c, const real *val)
{
    if constexpr (I < CH)
    {
        constexpr int Q = W + (T0 + I) * NS;
        if constexpr (Q < RHSRows<NDOF>::STAGED)
            c.lds[Q * kTileF + c.lane] = val[I];
        rhsStoreSeq<NDOF, W, NS, T0, I + 1, CH>(c, val);
    }
}
template <int NDOF, int W, int NS, int T0>
__device__ __forceinline__ void rhsStageChunks(const RHSStageCtx<NDOF> &c)
{
    constexpr int PER = (RHSRows<NDOF>::STAGED + NS - 1) / NS, CH = 12;
    if constexpr (T0 < PER)
    {
        real val[CH];
        rhsLoadSeq<NDOF, W, NS, T0, 0, CH>(c, val);
        rhsStoreSeq<NDOF, W, NS, T0, 0, CH>(c, val);
        rhsStageChunks<NDOF, W, NS, T0 + CH>(c);
    }
}
template <int NDOF, int NS, int W>
__device__ __forceinline__ void rhsStageDispatch(int w, const RHSStageCtx<NDOF> &c)
{
    if constexpr (W < NS)
    {
        if (w == W)
            rhsStageChunks<NDOF, W, NS, 0>(c);
        else
            rhsStageDispatch<NDOF, NS, W + 1>(w, c);
    }
}

template <int DIM, int NDOF, int G, int RS, bool BATCH>
__global__ void __launch_bounds__(kTileF *G) __attribute__((amdgpu_waves_per_eu(5)))
kRHSFace(DevMesh m, GasParams gp, const real *__restrict__ u, const real *__restrict__ uRec,
         const real *__restrict__ dlc, real *__restrict__ fflux)
{
    using R = RHSRows<NDOF>;
    constexpr int M = R::M;
    // waves that stage rows: all but wave 0 when there are enough waves
    constexpr int S0 = G >= 4 ? 1 : 0, NS = G - S0;
    extern __shared__ real lds[];
    const int lane = threadIdx.x % kTileF, g = waveIndex(kTileF);
    const size_t N = m.N, F = m.F;
    const uint64_t rowN = (uint64_t)N * 8, rowF = (uint64_t)F * 8;
    const int tile = xcdTile(blockIdx.x, gridDim.x);
    const idx f = tile * kTileF + lane;
    const bool valid = f < m.F;
    const uint32_t fo = (uint32_t)(valid ? f : 0) * 8u;
    const idx cL = bldi(rowRsrc(m.f2c, (uint64_t)F * 4), fo / 2);
    const idx cR = bldi(rowRsrc(m.f2c + F, (uint64_t)F * 4), fo / 2);
    const uint32_t oL = (uint32_t)cL * 8u, oR = (uint32_t)cR * 8u;

    // this lane's point geometry (registers, used after the barrier; read once, streamed)
    real x[3], n[3];
#pragma unroll
    for (int r = 0; r < 3; r++)
    {
        x[r] = bldnt(rowRsrc(m.fpts + ((size_t)g * 3 + r) * F, rowF), fo);
        n[r] = bldnt(rowRsrc(m.fnorm + ((size_t)g * 3 + r) * F, rowF), fo);
    }
    const real wj = bldnt(rowRsrc(m.fwj + (size_t)g * F, rowF), fo);

    if (g >= S0)
    {
        const RHSStageCtx<NDOF> sc{&m, u, uRec, lds, lane, fo, oL, oR};
        rhsStageDispatch<NDOF, NS, 0>(g - S0, sc);
    }
    if (g == 0)
    {
        // face constants from the cell means
        real uLm[5], uRm[5];
        real dLambda = 0;
#pragma unroll
        for (int v = 0; v < 5; v++)
        {
            const Rsrc rr = rowRsrc(u + (size_t)v * N, rowN);
            uLm[v] = bld(rr, oL), uRm[v] = bld(rr, oR);
        }
        if (dlc)
        {
            const Rsrc rr = rowRsrc(dlc, rowN);
            dLambda = fmax(bld(rr, oL), bld(rr, oR));
        }
        if (!valid)
        {
#pragma unroll
            for (int v = 0; v < 5; v++)
                uLm[v] = uRm[v] = v == 0 ? 1.0 : (v == 4 ? 2.5 : 0.0);
        }
        if constexpr (BATCH)
        {
            RoeBatchFace bf;
            real nm[3];
#pragma unroll
            for (int r = 0; r < 3; r++)
                nm[r] = bld(rowRsrc(m.fmeanN + (size_t)r * F, rowF), fo);
            roeBatchSetup<RS>(uLm, uRm, nm, gp, dLambda, bf);
            roeStore(lds, R::FC, lane, bf.r);
            lds[(R::FC + kRoeRows + 0) * kTileF + lane] = bf.vRN;
            lds[(R::FC + kRoeRows + 1) * kTileF + lane] = bf.lam0;
            lds[(R::FC + kRoeRows + 2) * kTileF + lane] = bf.lam123;
            lds[(R::FC + kRoeRows + 3) * kTileF + lane] = bf.lam4;
        }
        else
        {
            RoeAve ra;
            roeAverage(uLm, uRm, gp.gamma, ra);
            roeStore(lds, R::FC, lane, ra);
        }
        lds[(R::FC + kRoeRows + 4) * kTileF + lane] = dLambda;
    }
    __syncthreads();
    real Fp[5] = {0, 0, 0, 0, 0};
    if (valid)
    {
        real UL[5], UR[5];
#pragma unroll
        for (int v = 0; v < 5; v++)
            UL[v] = lds[(R::MEAN + v) * kTileF + lane], UR[v] = lds[(R::MEAN + 5 + v) * kTileF + lane];
        if (uRec)
        {
#pragma unroll
            for (int side = 0; side < 2; side++)
            {
                const real *met = lds + (size_t)(R::MET + side * R::METW) * kTileF + lane;
                const real isc = met[3 * kTileF];
                real xs[3], mom[M > 0 ? M : 1], phi[M > 0 ? M : 1];
#pragma unroll
                for (int r = 0; r < 3; r++)
                {
                    real p = x[r] + (side ? lds[(R::SHIFT + r) * kTileF + lane] : 0.0);
                    xs[r] = p - met[r * kTileF];
                    if (DIM == 3 || r < 2)
                        xs[r] *= isc;
                }
#pragma unroll
                for (int j = 0; j < M; j++)
                    mom[j] = met[(4 + j) * kTileF];
                basisRow0<DIM, NDOF>(xs, mom, phi);
                const real *rec = lds + (size_t)(R::REC + side * 5 * M) * kTileF + lane;
                real *U = side ? UR : UL;
#pragma unroll
                for (int v = 0; v < 5; v++)
                {
                    real s = 0;
#pragma unroll
                    for (int j = 0; j < M; j++)
                        s += phi[j] * rec[(v * M + j) * kTileF];
                    U[v] += s;
                }
            }
        }
        const real dLambda = lds[(R::FC + kRoeRows + 4) * kTileF + lane];
        RoeAve ra;
        roeLoad(lds, R::FC, lane, ra);
        real Fl[5];
        if constexpr (BATCH)
        {
            RoeBatchFace bf;
            bf.r = ra;
            bf.vRN = lds[(R::FC + kRoeRows + 0) * kTileF + lane];
            bf.lam0 = lds[(R::FC + kRoeRows + 1) * kTileF + lane];
            bf.lam123 = lds[(R::FC + kRoeRows + 2) * kTileF + lane];
            bf.lam4 = lds[(R::FC + kRoeRows + 3) * kTileF + lane];
            roeBatchPoint<RS>(UL, UR, n, gp, bf, Fl);
        }
        else
        {
            real l0, l1, l4;
            riemannFluxR<RS>(UL, UR, ra, n, gp, dLambda, Fl, l0, l1, l4);
        }
#pragma unroll
        for (int v = 0; v < 5; v++)
            Fp[v] = (-Fl[v]) * wj;
    }
    __syncthreads(); // all lanes are done with the staged data
#pragma unroll
    for (int v = 0; v < 5; v++)
        lds[(g * 5 + v) * kTileF + lane] = Fp[v];
    __syncthreads();
    for (int t = threadIdx.x; t < 5 * kTileF; t += blockDim.x)
    {
        const int v = t / kTileF, l = t % kTileF;
        const idx fout = tile * kTileF + l;
        if (fout >= m.F)
            continue;
        real acc = 0;
#pragma unroll
        for (int q = 0; q < G; q++)
            acc += lds[(q * 5 + v) * kTileF + l];
        fflux[(size_t)v * F + fout] = acc;
    }
}

__global__ void __launch_bounds__(256) kRHSGather(DevMesh m, const real *__restrict__ fflux, real *__restrict__ rhs,
                                                  StageUpdate st, int *__restrict__ nonFinite)
{
    idx i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m.N)
        return;
    const size_t N = m.N, F = m.F;
    real vol = m.vol[i];
    real acc[5] = {0, 0, 0, 0, 0};
    for (int k = 0; k < m.K; k++)
    {
        idx f = m.c2f[(size_t)k * N + i];
        if (f < 0)
            continue;
        bool back = m.f2c[f] == i;
#pragma unroll
        for (int v = 0; v < 5; v++)
        {
            real fv = fflux[v * F + f];
            acc[v] += (back ? fv : -fv) / vol;
        }
    }
    // the reference asserts on non-finite fluxes (EulerEvaluator_EvaluateRHS.hxx:387-398, 782-788);
    // here the first one raises a flag the host reads at its next synchronisation
    if (nonFinite && !(isfinite(acc[0]) && isfinite(acc[1]) && isfinite(acc[2]) && isfinite(acc[3]) && isfinite(acc[4])))
        atomicOr(nonFinite, 1);
    if (rhs)
    {
#pragma unroll
        for (int v = 0; v < 5; v++)
            rhs[v * N + i] = acc[v];
    }
    if (st.u) // same arithmetic as kAxpby(a, u, b, y, cdt, rhs, dt)
    {
        const real dt = st.dtDev ? st.dtDev[0] : 1.0;
#pragma unroll
        for (int v = 0; v < 5; v++)
        {
            real xi = st.u[v * N + i];
            if (st.save)
                st.save[v * N + i] = xi;
            real r = st.a * xi;
            if (st.y)
                r += st.b * st.y[v * N + i];
            r += st.cdt * (acc[v] * dt);
            st.u[v * N + i] = r;
        }
    }
}

typedef __attribute__((address_space(3))) void *LdsPtr;

// ---------------------------------------------------------------------------
// Multi-point, side-sequential face tile kernel: NW = ceil(G / PPW) wavefronts per
// 64-face tile, each lane evaluating PPW quadrature points of its face (lane = face),
// so one LDS read of a staged row (reconstruction coefficient, centre, moment, mean)
// feeds PPW point evaluations and the independent point computations give each wave
// instruction-level parallelism.  Rows are staged straight from global memory into
// LDS (buffer_load ... lds: a row of 64 doubles is two dword loads, lane l carrying
// half of face l/2; no VGPRs, no LDS store instructions, all rows of a wave in
// flight at once).  The two sides of the tile are staged one after the other into
// the same LDS rows: the left cells' mean,
// coefficients and metrics land, every wave reconstructs its points' left states,
// then the right cells' rows (and the periodic shift) replace them and the right
// states follow.  The LDS of a tile is one side's rows (5 + 5 M + 4 + M + 3), half
// of the two-sided kernel's, so twice as many tiles share a compute unit and hide
// each other's staging latency.  Every wave forms the face constants (Roe average of
// the means, batched-Roe eigenvalues, dLambda) itself from the means it read, so
// they take no LDS.  The point fluxes are parked in the staging rows once both
// sides are reconstructed and summed in point order.
// ---------------------------------------------------------------------------
template <int NDOF>
struct RHSSideRows
{
    static constexpr int M = NDOF - 1;
    static constexpr int MEAN = 0;            // [5]
    static constexpr int REC = 5;             // [5][M]
    static constexpr int MET = REC + 5 * M;   // centre 3, 1/scale, moment M
    static constexpr int SHIFT = MET + 4 + M; // [3], right side only
    static constexpr int STAGED = SHIFT + 3;
};

template <int NDOF, int Q>
__device__ __forceinline__ void rhsSideRow(const DevMesh &m, const real *u, const real *uRec, real *lds,
                                           const uint32_t *oc, const uint32_t *of)
{
    using R = RHSSideRows<NDOF>;
    const size_t N = m.N, F = m.F;
    const uint64_t rowN = (uint64_t)N * 8;
    real *dst = lds + Q * kTileF;
    const real *base;
    uint64_t bytes = rowN;
    const uint32_t *o = oc;
    if constexpr (Q < R::REC)
        base = u + (size_t)Q * N;
    else if constexpr (Q < R::MET)
    {
        if (!uRec) // without coefficients these rows are never read
            return;
        base = uRec + (size_t)(Q - R::REC) * N;
    }
    else if constexpr (Q < R::SHIFT)
    {
        constexpr int k = Q - R::MET;
        base = k < 3 ? m.cent + (size_t)k * N : (k == 3 ? m.iscale : m.moment + (size_t)(k - 4) * N);
    }
    else
        base = m.fshift + (size_t)(Q - R::SHIFT) * F, bytes = (uint64_t)F * 8, o = of;
    const Rsrc rs = rowRsrc(base, bytes);
#pragma unroll
    for (int h = 0; h < 2; h++)
        __builtin_amdgcn_raw_ptr_buffer_load_lds(rs, (LdsPtr)(dst + 32 * h), 4, o[h], 0, 0, 0);
}
template <int NDOF, int W, int NS, int T, int NR>
__device__ __forceinline__ void rhsSideSeq(const DevMesh &m, const real *u, const real *uRec, real *lds,
                                           const uint32_t *oc, const uint32_t *of)
{
    constexpr int Q = W + T * NS;
    if constexpr (Q < NR)
    {
        rhsSideRow<NDOF, Q>(m, u, uRec, lds, oc, of);
        rhsSideSeq<NDOF, W, NS, T + 1, NR>(m, u, uRec, lds, oc, of);
    }
}
template <int NDOF, int NS, int W, int NR>
__device__ __forceinline__ void rhsSideDispatch(int w, const DevMesh &m, const real *u, const real *uRec, real *lds,
                                                const uint32_t *oc, const uint32_t *of)
{
    if constexpr (W < NS)
    {
        if (w == W)
            rhsSideSeq<NDOF, W, NS, 0, NR>(m, u, uRec, lds, oc, of);
        else
            rhsSideDispatch<NDOF, NS, W + 1, NR>(w, m, u, uRec, lds, oc, of);
    }
}

// one side's reconstructed states at the lane's PPW points from the staged rows
template <int DIM, int NDOF, int PPW, bool SH>
__device__ __forceinline__ void rhsSideStates(const real *lds, int lane, const real (&x)[PPW][3], bool haveRec,
                                              real (&Um)[5], real (&U)[PPW][5])
{
    using R = RHSSideRows<NDOF>;
    constexpr int M = R::M;
#pragma unroll
    for (int v = 0; v < 5; v++)
    {
        Um[v] = lds[(R::MEAN + v) * kTileF + lane];
#pragma unroll
        for (int p = 0; p < PPW; p++)
            U[p][v] = Um[v];
    }
    if (!haveRec)
        return;
    const real *met = lds + R::MET * kTileF + lane;
    const real isc = met[3 * kTileF];
    real cen[3], mom[M > 0 ? M : 1], phi[PPW][M > 0 ? M : 1];
#pragma unroll
    for (int r = 0; r < 3; r++)
    {
        cen[r] = met[r * kTileF];
        if constexpr (SH)
            cen[r] -= lds[(R::SHIFT + r) * kTileF + lane]; // x + shift - cen
    }
#pragma unroll
    for (int j = 0; j < M; j++)
        mom[j] = met[(4 + j) * kTileF];
#pragma unroll
    for (int p = 0; p < PPW; p++)
    {
        real xs[3];
#pragma unroll
        for (int r = 0; r < 3; r++)
        {
            xs[r] = x[p][r] - cen[r];
            if (DIM == 3 || r < 2)
                xs[r] *= isc;
        }
        basisRow0<DIM, NDOF>(xs, mom, phi[p]);
    }
    const real *rec = lds + R::REC * kTileF + lane;
#pragma unroll
    for (int v = 0; v < 5; v++)
    {
        real sacc[PPW];
#pragma unroll
        for (int p = 0; p < PPW; p++)
            sacc[p] = 0;
#pragma unroll
        for (int j = 0; j < M; j++)
        {
            const real c = rec[(v * M + j) * kTileF];
#pragma unroll
            for (int p = 0; p < PPW; p++)
                sacc[p] += phi[p][j] * c;
        }
#pragma unroll
        for (int p = 0; p < PPW; p++)
            U[p][v] += sacc[p];
    }
}

template <int DIM, int NDOF, int G, int RS, bool BATCH, int PPW, int WPE>
__global__ void __launch_bounds__(kTileF *((G + PPW - 1) / PPW)) __attribute__((amdgpu_waves_per_eu(WPE)))
kRHSFaceS(DevMesh m, GasParams gp, const real *__restrict__ u, const real *__restrict__ uRec,
          const real *__restrict__ dlc, real *__restrict__ fflux)
{
    using R = RHSSideRows<NDOF>;
    constexpr int NW = (G + PPW - 1) / PPW;
    extern __shared__ real lds[];
    const int lane = threadIdx.x % kTileF, g = waveIndex(kTileF);
    const size_t N = m.N, F = m.F;
    const uint64_t rowN = (uint64_t)N * 8, rowF = (uint64_t)F * 8;
    const int tile = xcdTile(blockIdx.x, gridDim.x);
    const idx f = tile * kTileF + lane;
    const bool valid = f < m.F;
    const uint32_t fo = (uint32_t)(valid ? f : 0) * 8u;
    const idx cL = bldi(rowRsrc(m.f2c, (uint64_t)F * 4), fo / 2);
    const idx cR = bldi(rowRsrc(m.f2c + F, (uint64_t)F * 4), fo / 2);
    const uint32_t oL = (uint32_t)cL * 8u, oR = (uint32_t)cR * 8u;
    // byte offsets of the direct-to-LDS loads: lane l carries dword (l & 1) of rows (l >> 1)
    // and 32 + (l >> 1) of the tile, so a row of 64 doubles is two dword loads
    uint32_t ocL[2], ocR[2], of[2];
    {
        const int hl = lane >> 1;
        const uint32_t par = (uint32_t)(lane & 1) * 4u;
#pragma unroll
        for (int h = 0; h < 2; h++)
        {
            ocL[h] = (uint32_t)__shfl((int)oL, 32 * h + hl) + par;
            ocR[h] = (uint32_t)__shfl((int)oR, 32 * h + hl) + par;
            const idx f2 = tile * kTileF + 32 * h + hl;
            of[h] = (uint32_t)(f2 < m.F ? f2 : 0) * 8u + par;
        }
    }
    rhsSideDispatch<NDOF, NW, 0, R::SHIFT>(g, m, u, uRec, lds, ocL, of);
    // the lane's point coordinates (streamed once)
    real x[PPW][3];
#pragma unroll
    for (int p = 0; p < PPW; p++)
    {
        const int q = g * PPW + p < G ? g * PPW + p : G - 1; // the last wave may hold fewer points
#pragma unroll
        for (int r = 0; r < 3; r++)
            x[p][r] = bldnt(rowRsrc(m.fpts + ((size_t)q * 3 + r) * F, rowF), fo);
    }
    real dLambda = 0;
    if (dlc)
    {
        const Rsrc rr = rowRsrc(dlc, rowN);
        dLambda = fmax(bld(rr, oL), bld(rr, oR));
    }
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    real uLm[5], uRm[5], UL[PPW][5], UR[PPW][5];
    rhsSideStates<DIM, NDOF, PPW, false>(lds, lane, x, uRec != nullptr, uLm, UL);
    // the left mean waits in LDS rows past the staging and flux rows (the same for every wave),
    // so that it holds no registers through the right side
    constexpr int KEEP = R::STAGED > 5 * G ? R::STAGED : 5 * G;
    if (g == 0)
    {
#pragma unroll
        for (int v = 0; v < 5; v++)
            lds[(KEEP + v) * kTileF + lane] = uLm[v];
    }
    __syncthreads(); // the left rows are consumed
    rhsSideDispatch<NDOF, NW, 0, R::STAGED>(g, m, u, uRec, lds, ocR, of);
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    rhsSideStates<DIM, NDOF, PPW, true>(lds, lane, x, uRec != nullptr, uRm, UR);
    // the lane's point normals, loaded after the right states so that they do not hold registers
    // through them; their latency overlaps the barrier (the weights are applied in the final sum)
    real n[PPW][3];
#pragma unroll
    for (int p = 0; p < PPW; p++)
    {
        const int q = g * PPW + p < G ? g * PPW + p : G - 1;
#pragma unroll
        for (int r = 0; r < 3; r++)
            n[p][r] = bldnt(rowRsrc(m.fnorm + ((size_t)q * 3 + r) * F, rowF), fo);
    }
    real nm[3];
    if constexpr (BATCH)
    {
#pragma unroll
        for (int r = 0; r < 3; r++)
            nm[r] = bld(rowRsrc(m.fmeanN + (size_t)r * F, rowF), fo);
    }
    __syncthreads(); // the right rows are consumed: the staging rows take the point fluxes
#pragma unroll
    for (int v = 0; v < 5; v++)
        uLm[v] = lds[(KEEP + v) * kTileF + lane];
    if (!valid)
    {
#pragma unroll
        for (int v = 0; v < 5; v++)
            uLm[v] = uRm[v] = v == 0 ? 1.0 : (v == 4 ? 2.5 : 0.0);
    }
    RoeBatchFace bf;
    if constexpr (BATCH)
        roeBatchSetup<RS>(uLm, uRm, nm, gp, dLambda, bf);
    else
        roeAverage(uLm, uRm, gp.gamma, bf.r);
#pragma unroll
    for (int p = 0; p < PPW; p++)
    {
        const int q = g * PPW + p;
        if (q >= G)
            continue;
        real Fl[5];
        if constexpr (BATCH)
            roeBatchPoint<RS>(UL[p], UR[p], n[p], gp, bf, Fl);
        else
        {
            real l0, l1, l4;
            riemannFluxR<RS>(UL[p], UR[p], bf.r, n[p], gp, dLambda, Fl, l0, l1, l4);
        }
#pragma unroll
        for (int v = 0; v < 5; v++)
            lds[(q * 5 + v) * kTileF + lane] = valid ? -Fl[v] : 0.0;
    }
    __syncthreads();
    for (int t = threadIdx.x; t < 5 * kTileF; t += blockDim.x)
    {
        const int v = t / kTileF, l = t % kTileF;
        const idx fout = tile * kTileF + l;
        if (fout >= m.F)
            continue;
        real acc = 0;
#pragma unroll
        for (int q = 0; q < G; q++)
            acc += lds[(q * 5 + v) * kTileF + l] * m.fwj[(size_t)q * F + fout];
        fflux[(size_t)v * F + fout] = acc;
    }
}

// points per lane of the face tile kernel for a face rule of G points (1: one
// wavefront per point)
template <int DIM, int G>
constexpr int rhsPointsPerLane()
{
    return DIM == 3 && G == 9 ? 3 : 1;
}

// Boundary faces (EvaluateRHS.hxx:194-207, 336-345): one thread per boundary face,
// run after the tile kernel and overwriting the tile kernel's flux of that face.  The
// Roe average takes the cell mean and its ghost at the face centre (mean normal); every
// point takes the reconstructed state and its ghost at the point normal; dLambda is the
// owner cell's.  Point fluxes are summed in point order as in the tile kernels.
template <int DIM, int NDOF, int G, int RS, bool BATCH>
__global__ void __launch_bounds__(64) kRHSBoundary(DevMesh m, GasParams gp, const real *__restrict__ u,
                                                   const real *__restrict__ uRec, const real *__restrict__ dlc,
                                                   real *__restrict__ fflux)
{
    constexpr int M = NDOF - 1;
    const idx t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= m.nBnd)
        return;
    const idx f = m.bndFace[t];
    const size_t N = m.N, F = m.F;
    const idx cL = m.f2c[f];
    const int zone = m.fzone[f];
    const int type = m.bcType[zone];
    const real *far = m.bcVal + (size_t)zone * 5;
    real ULm[5], URm[5], nm[3];
#pragma unroll
    for (int v = 0; v < 5; v++)
        ULm[v] = u[(size_t)v * N + cL];
#pragma unroll
    for (int r = 0; r < 3; r++)
        nm[r] = m.fmeanN[(size_t)r * F + f];
    bcGhost(type, ULm, nm, far, gp.gamma, URm);
    const real dLambda = dlc ? dlc[cL] : 0.0;
    RoeBatchFace bf;
    RoeAve ra;
    if constexpr (BATCH)
        roeBatchSetup<RS>(ULm, URm, nm, gp, dLambda, bf);
    else
        roeAverage(ULm, URm, gp.gamma, ra);
    real mom[M > 0 ? M : 1], cen[3], isc = 0;
    if (uRec)
    {
#pragma unroll
        for (int j = 0; j < M; j++)
            mom[j] = m.moment[(size_t)j * N + cL];
#pragma unroll
        for (int r = 0; r < 3; r++)
            cen[r] = m.cent[(size_t)r * N + cL];
        isc = m.iscale[cL];
    }
    real acc[5] = {0, 0, 0, 0, 0};
    for (int g = 0; g < G; g++)
    {
        real UL[5], UR[5], n[3];
#pragma unroll
        for (int r = 0; r < 3; r++)
            n[r] = m.fnorm[((size_t)g * 3 + r) * F + f];
#pragma unroll
        for (int v = 0; v < 5; v++)
            UL[v] = ULm[v];
        if (uRec)
        {
            real xs[3], phi[M > 0 ? M : 1];
#pragma unroll
            for (int r = 0; r < 3; r++)
            {
                real p = m.fpts[((size_t)g * 3 + r) * F + f] + 0.0;
                xs[r] = p - cen[r];
                if (DIM == 3 || r < 2)
                    xs[r] *= isc;
            }
            basisRow0<DIM, NDOF>(xs, mom, phi);
#pragma unroll
            for (int v = 0; v < 5; v++)
            {
                real sacc = 0;
#pragma unroll
                for (int j = 0; j < M; j++)
                    sacc += phi[j] * uRec[((size_t)v * M + j) * N + cL];
                UL[v] += sacc;
            }
        }
        bcGhost(type, UL, n, far, gp.gamma, UR);
        real Fl[5];
        if constexpr (BATCH)
            roeBatchPoint<RS>(UL, UR, n, gp, bf, Fl);
        else
        {
            real l0, l1, l4;
            riemannFluxR<RS>(UL, UR, ra, n, gp, dLambda, Fl, l0, l1, l4);
        }
        const real wj = m.fwj[(size_t)g * F + f];
#pragma unroll
        for (int v = 0; v < 5; v++)
            acc[v] += (-Fl[v]) * wj;
    }
#pragma unroll
    for (int v = 0; v < 5; v++)
        fflux[(size_t)v * F + f] = acc[v];
}

// ---------------------------------------------------------------------------
// Laminar viscous face fluxes (EvaluateRHS.hxx:257-424 with fluxFace,
// EvaluateDt.hxx:1163-1215; finc -= visFlux at :1460).  One thread per face, after
// the inviscid face and boundary passes: at every face point both sides' values and
// gradients D1phi uRec, the GRP mean gradient
//   GradUMean = (GradL + GradR) / 2 + n (UR - UL) / distGRP,
//   distGRP = max(min(distBaryPerp, 2 d0), d0 / 4), d0 = 2 minVol / area,
// its primitive form at the mean state, and the Navier-Stokes flux; the face flux
// gains sum_g F_vis w_g (JDet + tiny).  A boundary point takes the ghost state and
// GradR = GradL; slip walls, symmetry planes, far field and outflow carry no viscous
// flux (their mean gradient is zeroed, :427-435, or the flux is, :1200-1204).
// ---------------------------------------------------------------------------
// One side of a face for the viscous pass: the cell's mean, centre, scale and basis
// moments, read once per face; with a linear basis (M <= 3) also its reconstruction
// coefficients, kept in registers across the face's points.  eval() returns the
// reconstruction's value and gradient at a point p in the cell's frame.
template <int DIM, int NDOF>
struct SideRec
{
    static constexpr int M = NDOF - 1;
    static constexpr bool REG = M <= 3;
    idx c;
    real um[5], cen[3], s, mom[M];
    real rec[REG ? 5 : 1][REG ? M : 1];

    __device__ __forceinline__ void load(const DevMesh &m, idx cc, const real *__restrict__ u,
                                         const real *__restrict__ uRec)
    {
        const size_t N = m.N;
        c = cc;
#pragma unroll
        for (int v = 0; v < 5; v++)
            um[v] = u[(size_t)v * N + c];
        if (!uRec)
            return;
#pragma unroll
        for (int r = 0; r < 3; r++)
            cen[r] = m.cent[(size_t)r * N + c];
        s = m.scale[c];
#pragma unroll
        for (int j = 0; j < M; j++)
            mom[j] = m.moment[(size_t)j * N + c];
        if constexpr (REG)
#pragma unroll
            for (int v = 0; v < 5; v++)
#pragma unroll
                for (int j = 0; j < M; j++)
                    rec[v][j] = uRec[((size_t)v * M + j) * N + c];
    }

    __device__ __forceinline__ void eval(const DevMesh &m, const real *p, const real *__restrict__ uRec, real *U,
                                         real (*Gr)[5]) const
    {
#pragma unroll
        for (int v = 0; v < 5; v++)
        {
            U[v] = um[v];
#pragma unroll
            for (int d = 0; d < 3; d++)
                Gr[d][v] = 0;
        }
        if (!uRec)
            return;
        real xs[3];
#pragma unroll
        for (int r = 0; r < 3; r++)
            xs[r] = p[r] - cen[r];
        xs[0] /= s, xs[1] /= s;
        if (DIM == 3)
            xs[2] /= s;
        real phi[M], dphi[DIM][M];
#pragma unroll
        for (int j = 0; j < M; j++)
        {
            phi[j] = basisDiff(DIM, 0, j + 1, xs, s) - mom[j];
#pragma unroll
            for (int d = 0; d < DIM; d++)
                dphi[d][j] = basisDiff(DIM, 1 + d, j + 1, xs, s);
        }
        const size_t N = m.N;
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
                real r;
                if constexpr (REG)
                    r = rec[v][j];
                else
                    r = uRec[((size_t)v * M + j) * N + c];
                a += phi[j] * r;
#pragma unroll
                for (int d = 0; d < DIM; d++)
                    g[d] += dphi[d][j] * r;
            }
            U[v] += a;
#pragma unroll
            for (int d = 0; d < DIM; d++)
                Gr[d][v] = g[d];
        }
    }
};

// Linear reconstruction (p1): the basis is x_s - moment with x_s = (x - cent) / scale, so
// the side's value at x is U0 + G (x - cent) with U0 = u - sum_j moment_j uRec_j and the
// gradient G_j = uRec_j / scale the same at every point: 23 values per side instead of the
// moments, coefficients and a per-point basis derivative.
template <int DIM>
struct SideLin
{
    real U0[5], G[3][5], cen[3];

    __device__ __forceinline__ void load(const DevMesh &m, idx c, const real *__restrict__ u,
                                         const real *__restrict__ uRec)
    {
        const size_t N = m.N;
#pragma unroll
        for (int v = 0; v < 5; v++)
        {
            U0[v] = u[(size_t)v * N + c];
#pragma unroll
            for (int d = 0; d < 3; d++)
                G[d][v] = 0;
        }
#pragma unroll
        for (int r = 0; r < 3; r++)
            cen[r] = 0;
        if (!uRec)
            return;
#pragma unroll
        for (int r = 0; r < 3; r++)
            cen[r] = m.cent[(size_t)r * N + c];
        const real s = m.scale[c];
        real mom[DIM];
#pragma unroll
        for (int j = 0; j < DIM; j++)
            mom[j] = m.moment[(size_t)j * N + c];
#pragma unroll
        for (int v = 0; v < 5; v++)
#pragma unroll
            for (int j = 0; j < DIM; j++)
            {
                const real r = uRec[((size_t)v * DIM + j) * N + c];
                U0[v] -= mom[j] * r;
                G[j][v] = r / s;
            }
    }

    __device__ __forceinline__ void eval(const DevMesh &, const real *p, const real *, real *U, real (*Gr)[5]) const
    {
#pragma unroll
        for (int v = 0; v < 5; v++)
        {
            real a = U0[v];
#pragma unroll
            for (int d = 0; d < DIM; d++)
                a += G[d][v] * (p[d] - cen[d]);
            U[v] = a;
#pragma unroll
            for (int d = 0; d < 3; d++)
                Gr[d][v] = G[d][v];
        }
    }
};
template <int DIM, int NDOF>
using ViscSide = std::conditional_t<NDOF == DIM + 1, SideLin<DIM>, SideRec<DIM, NDOF>>;

template <int DIM, int NDOF, bool GAS>
__global__ void __launch_bounds__(128) kRHSViscous(DevMesh m, GasParams gp, const real *__restrict__ u,
                                                   const real *__restrict__ uRec, real *__restrict__ fflux)
{
    const idx f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= m.F)
        return;
    const size_t N = m.N, F = m.F;
    const idx cL = m.f2c[f], cR = m.f2c[F + f];
    const bool bnd = cR < 0;
    const int type = bnd ? m.bcType[m.fzone[f]] : 0;
    if (bnd && (type == BC_WALL_INVIS || type == BC_SYM || type == BC_FAR || type == BC_OUT))
        return;
    const real *far = bnd ? m.bcVal + (size_t)m.fzone[f] * 5 : nullptr;
    real dB[3], sh[3];
    real minVol = m.vol[cL];
    ViscSide<DIM, NDOF> L, R;
    L.load(m, cL, u, uRec);
    if (!bnd)
        R.load(m, cR, u, uRec);
    if (!bnd)
    {
        minVol = fmin(minVol, m.vol[cR]);
#pragma unroll
        for (int r = 0; r < 3; r++)
        {
            sh[r] = m.fshift[(size_t)r * F + f];
            dB[r] = (m.bary[(size_t)r * N + cR] - sh[r]) - m.bary[(size_t)r * N + cL];
        }
    }
    else
#pragma unroll
        for (int r = 0; r < 3; r++)
            sh[r] = 0, dB[r] = m.fcent[(size_t)r * F + f] - m.bary[(size_t)r * N + cL];
    const real d0 = minVol / m.farea[f] * 2;
    const real gamma = gp.gamma;
    real acc[5] = {0, 0, 0, 0, 0};
    for (int g = 0; g < m.G; g++)
    {
        real x[3], n[3];
#pragma unroll
        for (int r = 0; r < 3; r++)
            x[r] = m.fpts[((size_t)g * 3 + r) * F + f], n[r] = m.fnorm[((size_t)g * 3 + r) * F + f];
        real UL[5], UR[5], GL[3][5], GR[3][5];
        L.eval(m, x, uRec, UL, GL);
        if (bnd)
        {
            bcGhost<real, GAS>(type, UL, n, far, gamma, UR);
#pragma unroll
            for (int d = 0; d < 3; d++)
#pragma unroll
                for (int v = 0; v < 5; v++)
                    GR[d][v] = GL[d][v];
        }
        else
        {
            real xR[3] = {x[0] + sh[0], x[1] + sh[1], x[2] + sh[2]};
            R.eval(m, xR, uRec, UR, GR);
        }
        real Fv[5];
        viscousPointFlux(gp, UL, UR, GL, GR, n, dB, d0, bnd, bnd && type == BC_WALL, Fv);
        const real wj = m.fwj[(size_t)g * F + f];
#pragma unroll
        for (int v = 0; v < 5; v++)
            acc[v] += Fv[v] * wj;
    }
#pragma unroll
    for (int v = 0; v < 5; v++)
        fflux[(size_t)v * F + f] += acc[v];
}

template <int DIM, int NDOF, int G, int RS, bool BATCH>
static void rhsLaunch(Ctx &c, const GasParams &gp, const real *u, const real *uRec, const real *dlc, real *ff)
{
    int rows = RHSRows<NDOF>::TOTAL;
    if (rows < 5 * G)
        rows = 5 * G;
    size_t smem = (size_t)rows * kTileF * sizeof(real);
    constexpr int PPW = rhsPointsPerLane<DIM, G>();
    if constexpr (PPW > 1)
    {
        constexpr int NW = (G + PPW - 1) / PPW;
        int srows = RHSSideRows<NDOF>::STAGED;
        if (srows < 5 * G)
            srows = 5 * G;
        srows += 5; // the left mean
        const size_t ss = (size_t)srows * kTileF * sizeof(real);
        // three waves per SIMD where the registers allow it; the cubic basis needs the room of two
        auto kern = kRHSFaceS<DIM, NDOF, G, RS, BATCH, PPW, (NDOF >= 20 ? 2 : 3)>;
        if (ss > 65536)
            DNDSR_HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void *>(kern),
                                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)ss));
        kern<<<gridFor(c.m.F, kTileF), kTileF * NW, ss, c.stream>>>(c.m, gp, u, uRec, dlc, ff);
    }
    else
    {
        auto kern = kRHSFace<DIM, NDOF, G, RS, BATCH>;
        if (smem > 65536)
            DNDSR_HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void *>(kern),
                                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem));
        kern<<<gridFor(c.m.F, kTileF), kTileF * G, smem, c.stream>>>(c.m, gp, u, uRec, dlc, ff);
    }
    DNDSR_HIP_CHECK(hipGetLastError());
    if (c.m.nBnd > 0)
    {
        kRHSBoundary<DIM, NDOF, G, RS, BATCH><<<gridFor(c.m.nBnd, 64), 64, 0, c.stream>>>(c.m, gp, u, uRec, dlc, ff);
        DNDSR_HIP_CHECK(hipGetLastError());
    }
}

template <int DIM, int NDOF, int G>
static void rhsDispatchRS(Ctx &c, const GasParams &gp, const real *u, const real *uRec, const real *dlc, real *ff)
{
    bool batch = gp.rsMeanValueEig != 0 && gp.rsType >= RS_Roe_M1 && gp.rsType <= RS_Roe_M5;
#define DNDSR_RHS_CASE(RSV)                                                  \
    case RSV:                                                               \
        rhsLaunch<DIM, NDOF, G, RSV, false>(c, gp, u, uRec, dlc, ff);       \
        break;
#define DNDSR_RHS_BCASE(RSV)                                                 \
    case RSV:                                                               \
        rhsLaunch<DIM, NDOF, G, RSV, true>(c, gp, u, uRec, dlc, ff);        \
        break;
    if (batch)
    {
        switch (gp.rsType)
        {
            DNDSR_RHS_BCASE(RS_Roe_M1)
            DNDSR_RHS_BCASE(RS_Roe_M2)
            DNDSR_RHS_BCASE(RS_Roe_M3)
            DNDSR_RHS_BCASE(RS_Roe_M4)
            DNDSR_RHS_BCASE(RS_Roe_M5)
        default: throw std::runtime_error("invalid batched Riemann solver");
        }
    }
    else
    {
        switch (gp.rsType)
        {
            DNDSR_RHS_CASE(RS_Roe)
            DNDSR_RHS_CASE(RS_HLLC)
            DNDSR_RHS_CASE(RS_HLLEP)
            DNDSR_RHS_CASE(RS_HLLEP_V1)
            DNDSR_RHS_CASE(RS_Roe_M1)
            DNDSR_RHS_CASE(RS_Roe_M2)
            DNDSR_RHS_CASE(RS_Roe_M3)
            DNDSR_RHS_CASE(RS_Roe_M4)
            DNDSR_RHS_CASE(RS_Roe_M5)
            DNDSR_RHS_CASE(RS_Roe_M6)
            DNDSR_RHS_CASE(RS_Roe_M7)
            DNDSR_RHS_CASE(RS_Roe_M8)
        default: throw std::runtime_error("the rs type is invalid");
        }
    }
#undef DNDSR_RHS_CASE
#undef DNDSR_RHS_BCASE
}

// face quadrature sizes: 2-D Line rules of 2 / 3 points (intOrder 2..5), 3-D Quad rules
// of 4 / 9 points (intOrder 2..5), Tri rules of 6 points (intOrder 3..4); the tile
// kernel is specialised on the point count.
template <int DIM, int NDOF>
static void rhsDispatch(Ctx &c, const GasParams &gp, const real *u, const real *uRec, const real *dlc, real *ff)
{
    constexpr int G1 = DIM == 2 ? 2 : 4, G2 = DIM == 2 ? 3 : 9;
    if (c.m.G == G1)
        rhsDispatchRS<DIM, NDOF, G1>(c, gp, u, uRec, dlc, ff);
    else if (c.m.G == G2)
        rhsDispatchRS<DIM, NDOF, G2>(c, gp, u, uRec, dlc, ff);
    else if (DIM == 3 && c.m.G == 6)
        rhsDispatchRS<DIM, NDOF, 6>(c, gp, u, uRec, dlc, ff);
    else
        throw std::runtime_error("EvaluateRHS supports 2/3 face points in 2-D (intOrder 2..5), 4/9 on quadrilateral "
                                 "faces (intOrder 2..5) and 6 on triangular faces (intOrder 3..4)");
}

void launchEvalRHS(Ctx &c, const GasParams &gp, const real *u, const real *uRec, const real *dlc, real *rhs,
                   const StageUpdate *stage, real *faceFluxBuf, bool flagNonFinite)
{
    DevMesh &m = c.m;
    if (m.nBnd > 0 && !m.bcType)
        throw std::runtime_error("the mesh has boundary faces: set the boundary conditions first (dndsr_set_bc)");
    real *ff = faceFluxBuf ? faceFluxBuf : c.faceFlux;
    {
        KernelTimer kt(c, K_RHS_FACE); // inviscid interior, periodic and boundary faces
        if (m.dim == 2)
        {
            if (!uRec || m.P == 1)
                rhsDispatch<2, 3>(c, gp, u, uRec, dlc, ff);
            else if (m.P == 2)
                rhsDispatch<2, 6>(c, gp, u, uRec, dlc, ff);
            else
                rhsDispatch<2, 10>(c, gp, u, uRec, dlc, ff);
        }
        else
        {
            if (!uRec || m.P == 1)
                rhsDispatch<3, 4>(c, gp, u, uRec, dlc, ff);
            else if (m.P == 2)
                rhsDispatch<3, 10>(c, gp, u, uRec, dlc, ff);
            else
                rhsDispatch<3, 20>(c, gp, u, uRec, dlc, ff);
        }
    }
    if (gp.viscous)
    {
        KernelTimer kt(c, K_RHS_VISCOUS);
        if (!(gp.prGas > 0 && gp.Rgas > 0 && gp.muGas >= 0))
            throw std::runtime_error("viscous residual needs prGas > 0, Rgas > 0 and muGas >= 0");
        const int nb = gridFor(m.F, 128);
        const int P = uRec ? m.P : 1;
        // the gas-dependent ghosts only where the table has such a zone (their branch costs
        // the kernel registers)
        auto launchVis = [&](auto dimC, auto ndofC) {
            constexpr int D = decltype(dimC)::value, ND = decltype(ndofC)::value;
            if (m.gasBC)
                kRHSViscous<D, ND, true><<<nb, 128, 0, c.stream>>>(m, gp, u, uRec, ff);
            else
                kRHSViscous<D, ND, false><<<nb, 128, 0, c.stream>>>(m, gp, u, uRec, ff);
        };
        using I2 = std::integral_constant<int, 2>;
        using I3 = std::integral_constant<int, 3>;
        if (m.dim == 2)
        {
            if (P == 1)
                launchVis(I2{}, I3{});
            else if (P == 2)
                launchVis(I2{}, std::integral_constant<int, 6>{});
            else
                launchVis(I2{}, std::integral_constant<int, 10>{});
        }
        else
        {
            if (P == 1)
                launchVis(I3{}, std::integral_constant<int, 4>{});
            else if (P == 2)
                launchVis(I3{}, std::integral_constant<int, 10>{});
            else
                launchVis(I3{}, std::integral_constant<int, 20>{});
        }
        DNDSR_HIP_CHECK(hipGetLastError());
    }
    KernelTimer kt2(c, K_RHS_GATHER);
    kRHSGather<<<gridFor(m.N, 256), 256, 0, c.stream>>>(m, ff, rhs, stage ? *stage : StageUpdate{},
                                                         flagNonFinite ? c.nonFinite : nullptr);
    DNDSR_HIP_CHECK(hipGetLastError());
}

// ---------------------------------------------------------------------------
// LUSGSMatrixInit, scalar Jacobian (EulerEvaluator.hxx:37-159, ignoreSourceTerm):
// D_i = 1/dTau_i + 1/dt + sum_f 0.5 alpha A_f lambda_f / V_i
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) kLUSGSInit(DevMesh m, const real *dTau, real dt, real alphaDiag,
                                                  const real *lam, real *jd)
{
    idx i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m.N)
        return;
    real fp = 1.0 / dTau[i] + 1.0 / dt;
    for (int k = 0; k < m.K; k++)
    {
        idx f = m.c2f[(size_t)k * m.N + i];
        if (f >= 0)
            fp += (0.5 * alphaDiag) * m.farea[f] * lam[f] / m.vol[i];
    }
    jd[i] = fp;
}
void launchLUSGSInit(Ctx &c, const real *dTau, real dt, real alphaDiag, const real *lambdaFace, real *jdiag)
{
    KernelTimer kt(c, K_LUSGS_INIT);
    kLUSGSInit<<<gridFor(c.m.N, 256), 256, 0, c.stream>>>(c.m, dTau, dt, alphaDiag, lambdaFace, jdiag);
    DNDSR_HIP_CHECK(hipGetLastError());
}

// ---------------------------------------------------------------------------
// UpdateSGS (EulerEvaluator.hxx:550-693), scalar Jacobian, multicolour form.
// Cells of one colour share no face, so one launch per colour updates them all
// in parallel; a neighbour counts as already visited when its colour precedes
// (forward) or follows (backward) the current one -- the serial reference sweep
// over cells renumbered colour by colour.  Per-block |dx| partial sums are
// written to distinct slots and reduced once at the end (deterministic).
// ---------------------------------------------------------------------------
template <typename T>
struct SGSGeom;
template <>
struct SGSGeom<real>
{
    static __device__ const real *area(const DevMesh &m) { return m.farea; }
    static __device__ const real *norm(const DevMesh &m) { return m.fmeanN; }
    static __device__ const real *vol(const DevMesh &m) { return m.vol; }
};
template <>
struct SGSGeom<float>
{
    static __device__ const float *area(const DevMesh &m) { return m.fareaF; }
    static __device__ const float *norm(const DevMesh &m) { return m.fmeanNF; }
    static __device__ const float *vol(const DevMesh &m) { return m.volF; }
};

// One thread per cell of the colour.  Per face slot the face index, the neighbour
// (c2c) and the slot flags (side, colour order) are independent loads; the face's
// geometry and the neighbour's state follow in one round.
template <typename T>
__global__ void __launch_bounds__(256) kSGSColour(DevMesh m, real gammaD, T alphaDiag, const T *__restrict__ rhs,
                                                  const T *__restrict__ u, const T *__restrict__ uInc,
                                                  T *__restrict__ uIncNew, const T *__restrict__ jd,
                                                  const T *__restrict__ lam, int col, int forward, int gsUpdate,
                                                  int uIncIsZero, idx begin, idx end, real *__restrict__ part,
                                                  int fCol, int lCol)
{
    __shared__ real sm[5][256];
    idx t = begin + blockIdx.x * blockDim.x + threadIdx.x;
    real loc[5] = {0, 0, 0, 0, 0};
    if (t < end)
    {
        const size_t N = m.N, F = m.F;
        const idx i = m.colorCells[t];
        const T gamma = (T)gammaD;
        const T *__restrict__ farea = SGSGeom<T>::area(m);
        const T *__restrict__ fnorm = SGSGeom<T>::norm(m);
        T buf[5];
#pragma unroll
        for (int v = 0; v < 5; v++)
            buf[v] = rhs[v * N + i];
        const T voli = SGSGeom<T>::vol(m)[i];
        for (int k = 0; k < m.K; k++)
        {
            const idx f = m.c2f[(size_t)k * N + i];
            const idx o = m.c2c[(size_t)k * N + i];
            const uint8_t fl = m.c2flag[(size_t)k * N + i];
            if (f < 0 || o < 0 || o == i)
                continue;
            // the neighbour precedes the cell in this sweep: forward = lower colour
            const bool earlier = forward ? (fl & 2) != 0 : (fl & 2) == 0;
            const bool useNew = gsUpdate && earlier;
            if (uIncIsZero && !useNew)
                continue;
            const bool readNew = sweepReadNew(fl, fCol, lCol, m.nColor, useNew);
            const T *du = readNew ? uIncNew : uInc;
            T U[5], dU[5], n[3], dF[5];
#pragma unroll
            for (int v = 0; v < 5; v++)
                U[v] = u[v * N + o], dU[v] = du[v * N + o];
            const T sgn = (fl & 1) ? T(-1) : T(1);
#pragma unroll
            for (int r = 0; r < 3; r++)
                n[r] = fnorm[r * F + f] * sgn;
            fluxJacDu<T>(U, n, dU, lam[f], gamma, dF);
            T coef = (T(0.5) * alphaDiag) * farea[f] / voli;
#pragma unroll
            for (int v = 0; v < 5; v++)
                buf[v] -= coef * dF[v];
        }
        T d = jd[i];
#pragma unroll
        for (int v = 0; v < 5; v++)
        {
            T old = uIncNew[v * N + i];
            T nw = buf[v] / d;
            uIncNew[v * N + i] = nw;
            loc[v] = (real)fabs(nw - old);
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

// ---------------------------------------------------------------------------
// The transpose of the scalar LU-SGS matrix, swept in the same multicolour order: the
// preconditioner of the discrete-adjoint fixed point (no counterpart in the reference).
// (P^T y)_i = D_i y_i + sum_j c_ji (A(u_i) n_ji - lambda_f I)^T y_j, c_ji = (alpha/2) A_f / V_j,
// n_ji out of the neighbour j: per face slot the matrix A(u_i) n_ji - lambda_f I is built column
// by column with fluxJacDu and applied transposed to the neighbour's y_j.
// ---------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(256) kSGSColourT(DevMesh m, real gammaD, T alphaDiag, const T *__restrict__ rhs,
                                                   const T *__restrict__ u, const T *__restrict__ uInc,
                                                   T *__restrict__ uIncNew, const T *__restrict__ jd,
                                                   const T *__restrict__ lam, int forward, int gsUpdate,
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
        const T gamma = (T)gammaD;
        const T *__restrict__ farea = SGSGeom<T>::area(m);
        const T *__restrict__ fnorm = SGSGeom<T>::norm(m);
        const T *__restrict__ vol = SGSGeom<T>::vol(m);
        T buf[5], ui[5];
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
            const T *y = readNew ? uIncNew : uInc;
            T yj[5], n[3];
#pragma unroll
            for (int v = 0; v < 5; v++)
                yj[v] = y[v * N + o];
            const T sgn = (fl & 1) ? T(1) : T(-1); // out of the neighbour
#pragma unroll
            for (int r = 0; r < 3; r++)
                n[r] = fnorm[r * F + f] * sgn;
            const T cji = (T(0.5) * alphaDiag) * farea[f] / vol[o];
#pragma unroll
            for (int c = 0; c < 5; c++)
            {
                T e[5] = {0, 0, 0, 0, 0}, col[5];
                e[c] = 1;
                fluxJacDu<T>(ui, n, e, lam[f], gamma, col);
                T acc = 0;
#pragma unroll
                for (int r = 0; r < 5; r++)
                    acc += col[r] * yj[r];
                buf[c] -= cji * acc;
            }
        }
        const T d = jd[i];
#pragma unroll
        for (int v = 0; v < 5; v++)
        {
            const T old = uIncNew[v * N + i];
            const T nw = buf[v] / d;
            uIncNew[v * N + i] = nw;
            loc[v] = (real)fabs(nw - old);
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

__global__ void __launch_bounds__(256) kReduceSum5(const real *part, int nPart, real *out)
{
    __shared__ real sm[5][256];
    real a[5] = {0, 0, 0, 0, 0};
    for (int q = threadIdx.x; q < nPart; q += 256)
#pragma unroll
        for (int v = 0; v < 5; v++)
            a[v] += part[(size_t)q * 5 + v];
#pragma unroll
    for (int v = 0; v < 5; v++)
        sm[v][threadIdx.x] = a[v];
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
            out[v] = sm[v][0];
}

void launchReduceSum5(Ctx &c, int nPart, real *out)
{
    kReduceSum5<<<1, 256, 0, c.stream>>>(c.red, nPart, out);
    DNDSR_HIP_CHECK(hipGetLastError());
}

template <typename T>
static void sgsSweep(Ctx &c, const GasParams &gp, T alphaDiag, const T *rhs, const T *u, const T *uInc, T *uIncNew,
                     const T *jd, const T *lam, int forward, int gsUpdate, int uIncIsZero, real *sumIncDev,
                     int fCol = -1, int lCol = -1)
{
    DevMesh &m = c.m;
    int nc = m.nColor;
    size_t off = 0;
    for (int s = 0; s < nc; s++)
    {
        int col = forward ? s : nc - 1 - s;
        if (col == fCol)
            continue;
        idx b = m.colorPtrHost[col], e = m.colorPtrHost[col + 1];
        if (e <= b)
            continue;
        int nb = gridFor(e - b, 256);
        KernelTimer kt(c, K_SGS);
        kSGSColour<T><<<nb, 256, 0, c.stream>>>(m, gp.gamma, alphaDiag, rhs, u, uInc, uIncNew, jd, lam, col, forward,
                                               gsUpdate, uIncIsZero, b, e, c.red + off * 5, fCol, lCol);
        DNDSR_HIP_CHECK(hipGetLastError());
        off += nb;
    }
    kReduceSum5<<<1, 256, 0, c.stream>>>(c.red, (int)off, sumIncDev);
    DNDSR_HIP_CHECK(hipGetLastError());
}

void launchSGSSweepT64(Ctx &c, const GasParams &gp, real alphaDiag, const real *rhs, const real *u, const real *uInc,
                       real *uIncNew, const real *jdiag, const real *lambdaFace, int forward, int gsUpdate,
                       int uIncIsZero, real *sumIncDev, int fCol, int lCol)
{
    DevMesh &m = c.m;
    size_t off = 0;
    for (int s = 0; s < m.nColor; s++)
    {
        const int col = forward ? s : m.nColor - 1 - s;
        if (col == fCol)
            continue;
        const idx b = m.colorPtrHost[col], e = m.colorPtrHost[col + 1];
        if (e <= b)
            continue;
        const int nb = gridFor(e - b, 256);
        KernelTimer kt(c, K_ADJOINT);
        kSGSColourT<real><<<nb, 256, 0, c.stream>>>(m, gp.gamma, alphaDiag, rhs, u, uInc, uIncNew, jdiag, lambdaFace,
                                                   forward, gsUpdate, uIncIsZero, b, e, c.red + off * 5, fCol, lCol);
        DNDSR_HIP_CHECK(hipGetLastError());
        off += nb;
    }
    kReduceSum5<<<1, 256, 0, c.stream>>>(c.red, (int)off, sumIncDev);
    DNDSR_HIP_CHECK(hipGetLastError());
}

void launchSGSSweep64(Ctx &c, const GasParams &gp, real alphaDiag, const real *rhs, const real *u, const real *uInc,
                      real *uIncNew, const real *jdiag, const real *lambdaFace, int forward, int gsUpdate,
                      int uIncIsZero, real *sumIncDev)
{
    sgsSweep<real>(c, gp, alphaDiag, rhs, u, uInc, uIncNew, jdiag, lambdaFace, forward, gsUpdate, uIncIsZero,
                   sumIncDev);
}
void launchSGSSweep32(Ctx &c, const GasParams &gp, float alphaDiag, const float *rhs, const float *u,
                      const float *uInc, float *uIncNew, const float *jdiag, const float *lambdaFace, int forward,
                      int gsUpdate, int uIncIsZero, real *sumIncDev)
{
    sgsSweep<float>(c, gp, alphaDiag, rhs, u, uInc, uIncNew, jdiag, lambdaFace, forward, gsUpdate, uIncIsZero,
                    sumIncDev);
}

// the increments of the cells of one colour from one array into the other ([5][N])
template <typename T>
__global__ void __launch_bounds__(256) kCopyColourIncS(DevMesh m, idx lo, idx hi, const T *__restrict__ from,
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

// nSweep symmetric scalar sweeps from the increments in a (b: the other array); returns the
// array holding the result.  With Gauss-Seidel updates a sweep after the first skips its first
// colour, whose pass would repeat the previous sweep's last one (same inputs, same values): the
// sweep reads that colour from the array it reads, the next sweep reads it from the array it
// writes until the colour's own pass, and one copy after the last sweep completes the result.
template <typename T>
static T *sgsSymmetric(Ctx &c, const GasParams &gp, T alphaDiag, const T *rhs, const T *u, T *a, T *b, const T *jd,
                       const T *lam, int nSweep, int gsUpdate, real *sumIncDev)
{
    DevMesh &m = c.m;
    const int nC = m.nColor;
    const bool canSkip = gsUpdate && nC >= 2;
    int prevSkip = -1;
    for (int s = 0; s < nSweep; s++)
        for (int dir = 1; dir >= 0; dir--)
        {
            const int first = dir ? 0 : nC - 1, last = dir ? nC - 1 : 0;
            const int fCol = canSkip && !(s == 0 && dir == 1) ? first : -1;
            const int lCol = prevSkip == last ? last : -1;
            sgsSweep<T>(c, gp, alphaDiag, rhs, u, a, b, jd, lam, dir, gsUpdate, 0, sumIncDev, fCol, lCol);
            prevSkip = fCol;
            std::swap(a, b);
        }
    if (prevSkip >= 0)
    {
        const idx lo = m.colorPtrHost[prevSkip], hi = m.colorPtrHost[prevSkip + 1];
        if (hi > lo)
        {
            KernelTimer kt(c, K_SGS);
            kCopyColourIncS<T><<<gridFor(hi - lo, 256), 256, 0, c.stream>>>(m, lo, hi, b, a);
            DNDSR_HIP_CHECK(hipGetLastError());
        }
    }
    return a;
}
real *launchSGSSymmetric64(Ctx &c, const GasParams &gp, real alphaDiag, const real *rhs, const real *u, real *a,
                           real *b, const real *jd, const real *lam, int nSweep, int gsUpdate, real *sumIncDev)
{
    return sgsSymmetric<real>(c, gp, alphaDiag, rhs, u, a, b, jd, lam, nSweep, gsUpdate, sumIncDev);
}
float *launchSGSSymmetric32(Ctx &c, const GasParams &gp, float alphaDiag, const float *rhs, const float *u, float *a,
                            float *b, const float *jd, const float *lam, int nSweep, int gsUpdate, real *sumIncDev)
{
    return sgsSymmetric<float>(c, gp, alphaDiag, rhs, u, a, b, jd, lam, nSweep, gsUpdate, sumIncDev);
}

// out = a x + b y + c (z * dt); dt read from device memory (1 when null) so that a
// time loop never waits on the host for dtMin.
__global__ void __launch_bounds__(256) kAxpby(idx n, real a, const real *__restrict__ x, real b,
                                              const real *__restrict__ y, real cc, const real *__restrict__ z,
                                              const real *__restrict__ dtDev, real *__restrict__ out,
                                              real *__restrict__ save)
{
    idx i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    real dt = dtDev ? dtDev[0] : 1.0;
    real xi = x[i];
    if (save)
        save[i] = xi;
    real r = a * xi;
    if (y)
        r += b * y[i];
    if (z)
        r += cc * (z[i] * dt);
    out[i] = r;
}
void launchAxpby(Ctx &c, idx n, real a, const real *x, real b, const real *y, real cdt, const real *z,
                 const real *dtDev, real *out, real *save)
{
    KernelTimer kt(c, K_AXPBY);
    kAxpby<<<gridFor(n, 256), 256, 0, c.stream>>>(n, a, x, b, y, cdt, z, dtDev, out, save);
    DNDSR_HIP_CHECK(hipGetLastError());
}

// Device-to-device copy of n doubles as one kernel on the library's stream (the runtime's
// blit path splits one large copy into several dispatches).  16-byte lanes where both
// pointers allow it.
__global__ void __launch_bounds__(256) kCopy2(size_t n2, const double2 *__restrict__ src, double2 *__restrict__ dst)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n2; i += stride)
        dst[i] = src[i];
}
__global__ void __launch_bounds__(256) kCopy1(size_t n, const real *__restrict__ src, real *__restrict__ dst)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride)
        dst[i] = src[i];
}
void launchCopy(Ctx &c, size_t n, const real *src, real *dst)
{
    if (n == 0 || src == dst)
        return;
    KernelTimer kt(c, K_AXPBY);
    const bool vec = ((reinterpret_cast<uintptr_t>(src) | reinterpret_cast<uintptr_t>(dst)) & 15) == 0;
    const size_t work = vec ? n / 2 : n;
    const unsigned nb = (unsigned)std::min<size_t>((work + 255) / 256, 8192);
    if (vec)
    {
        kCopy2<<<nb, 256, 0, c.stream>>>(n / 2, reinterpret_cast<const double2 *>(src),
                                         reinterpret_cast<double2 *>(dst));
        DNDSR_HIP_CHECK(hipGetLastError());
        if (n & 1)
            kCopy1<<<1, 64, 0, c.stream>>>(1, src + n - 1, dst + n - 1);
    }
    else
        kCopy1<<<nb, 256, 0, c.stream>>>(n, src, dst);
    DNDSR_HIP_CHECK(hipGetLastError());
}

// ---------------------------------------------------------------------------
// Implicit dual-time step pieces (ImplicitEulerDualTimeStep::Step, ODE.hpp:80-98).
// Unsteady residual rhs = xLast (1/dt) + x (-1/dt) + frhs(x), in the reference's
// operation order, fused with the per-component L1 norm (EvaluateNorm with P = 1,
// not volume-weighted, EulerEvaluator.hxx:1394-1424).  The partial sums of each
// block land in distinct slots and are reduced once (deterministic).
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) kDualResidual(idx N, real rdt, const real *__restrict__ xLast,
                                                     const real *__restrict__ x, const real *__restrict__ rhsF,
                                                     real *__restrict__ rhs, real *__restrict__ part)
{
    __shared__ real sm[5][256];
    const idx i = blockIdx.x * blockDim.x + threadIdx.x;
    real loc[5] = {0, 0, 0, 0, 0};
    if (i < N)
    {
#pragma unroll
        for (int v = 0; v < 5; v++)
        {
            const size_t k = (size_t)v * N + i;
            real r = xLast[k] * rdt;
            r += x[k] * (-rdt);
            r += rhsF[k];
            rhs[k] = r;
            loc[v] = fabs(r);
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
// The BDF unsteady residual (ImplicitBDFDualTimeStep::Step, src/Solver/ODE.hpp:440-451):
// rhs = sum_j a_j x_j - x / dt + alpha rhsF with a_j = c_j / dt over xLast and up to three
// earlier states, summed in that order; L1 partial sums as kDualResidual.
struct BDFTerms
{
    const real *x[4];
    real a[4];
    int n;
};
__global__ void __launch_bounds__(256) kDualResidualBDF(idx N, BDFTerms bt, real rdt, real alpha,
                                                        const real *__restrict__ x, const real *__restrict__ rhsF,
                                                        real *__restrict__ rhs, real *__restrict__ part)
{
    __shared__ real sm[5][256];
    const idx i = blockIdx.x * blockDim.x + threadIdx.x;
    real loc[5] = {0, 0, 0, 0, 0};
    if (i < N)
    {
#pragma unroll
        for (int v = 0; v < 5; v++)
        {
            const size_t k = (size_t)v * N + i;
            real r = 0;
            for (int j = 0; j < bt.n; j++)
                r += bt.x[j][k] * bt.a[j];
            r += x[k] * (-rdt);
            r += rhsF[k] * alpha;
            rhs[k] = r;
            loc[v] = fabs(r);
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
void launchDualResidualBDF(Ctx &c, real dt, const real *const *xs, const real *coef, int nx, real alpha,
                           const real *x, const real *rhsF, real *rhs, real *resDev)
{
    KernelTimer kt(c, K_DUAL_RES);
    if (nx < 1 || nx > 4)
        throw std::runtime_error("the BDF residual takes 1..4 earlier states");
    BDFTerms bt{};
    bt.n = nx;
    for (int j = 0; j < nx; j++)
        bt.x[j] = xs[j], bt.a[j] = coef[j] / dt;
    const int nb = gridFor(c.m.N, 256);
    kDualResidualBDF<<<nb, 256, 0, c.stream>>>(c.m.N, bt, 1. / dt, alpha, x, rhsF, rhs, c.red);
    DNDSR_HIP_CHECK(hipGetLastError());
    kReduceSum5<<<1, 256, 0, c.stream>>>(c.red, nb, resDev);
    DNDSR_HIP_CHECK(hipGetLastError());
}

__global__ void __launch_bounds__(256) kScale(idx n, real s, real *__restrict__ x)
{
    const idx i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        x[i] *= s;
}
void launchScale(Ctx &c, idx n, real s, real *x)
{
    kScale<<<gridFor(n, 256), 256, 0, c.stream>>>(n, s, x);
    DNDSR_HIP_CHECK(hipGetLastError());
}

void launchDualResidual(Ctx &c, real dt, const real *xLast, const real *x, const real *rhsF, real *rhs, real *resDev)
{
    KernelTimer kt(c, K_DUAL_RES);
    const int nb = gridFor(c.m.N, 256);
    kDualResidual<<<nb, 256, 0, c.stream>>>(c.m.N, 1.0 / dt, xLast, x, rhsF, rhs, c.red);
    DNDSR_HIP_CHECK(hipGetLastError());
    kReduceSum5<<<1, 256, 0, c.stream>>>(c.red, nb, resDev);
    DNDSR_HIP_CHECK(hipGetLastError());
}

// u += du after the positivity tests of the increment (EvaluateCellRHSAlpha,
// EulerEvaluator.hxx:1964-2020, with relax = incrementPPRelax) and AddFixedIncrement /
// ---------------------------------------------------------------------------
// Packed Gauss-Seidel sweep of the device implicit loop.  Same arithmetic as
// kSGSColour, other layout: everything the sweep touches is stored in colour-major
// position order (position p holds cell colorCells[p]), so a colour pass reads and writes
// one contiguous range whatever the caller's cell order (a locality order for the face
// and cell kernels need not group cells by colour).  The state the sweep gathers from
// neighbours (u, uInc, uIncNew) is cell-major, [N][5], so one neighbour costs one or two
// cache lines instead of ten; the face data a cell needs per slot (signed mean normal,
// area, lambdaFace) and the neighbour's position are laid out per position and slot,
// [K][5][N] and [K][N], and the cell's rhs, JDiag and volume per position, all by
// kSGSSlots once per iteration.  A neighbour precedes in the sweep exactly when its
// position is lower (the colouring has no neighbours of one colour).
// ---------------------------------------------------------------------------
// the mesh-only part of the packed layout (slot normals and areas, neighbour positions, volume):
// written once per context and precision
template <typename T>
__global__ void __launch_bounds__(256) kSGSSlotsStatic(DevMesh m, T *__restrict__ slot, idx *__restrict__ nbr)
{
    const idx p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= m.N)
        return;
    const size_t N = m.N, F = m.F;
    const idx i = m.colorCells[p];
    const T *__restrict__ farea = SGSGeom<T>::area(m);
    const T *__restrict__ fnorm = SGSGeom<T>::norm(m);
    for (int k = 0; k < m.K; k++)
    {
        const idx f = m.c2f[(size_t)k * N + i];
        const idx o = m.c2c[(size_t)k * N + i];
        T q[4] = {0, 0, 0, 0};
        if (f >= 0)
        {
            const T sgn = (m.c2flag[(size_t)k * N + i] & 1) ? T(-1) : T(1);
#pragma unroll
            for (int r = 0; r < 3; r++)
                q[r] = fnorm[r * F + f] * sgn;
            q[3] = farea[f];
        }
#pragma unroll
        for (int v = 0; v < 4; v++)
            slot[((size_t)k * 5 + v) * N + p] = q[v];
        nbr[(size_t)k * N + p] = (f < 0 || o < 0 || o == i) ? idx(-1) : m.cellPos[o];
    }
    slot[((size_t)m.K * 5 + 6) * N + p] = SGSGeom<T>::vol(m)[i];
}

// the per-iteration part: lambdaFace per slot, the cell's rhs and JDiag
template <typename T>
__global__ void __launch_bounds__(256) kSGSSlots(DevMesh m, const T *__restrict__ lam, const T *__restrict__ rhs,
                                                 const T *__restrict__ jd, T *__restrict__ slot)
{
    const idx p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= m.N)
        return;
    const size_t N = m.N;
    const idx i = m.colorCells[p];
    for (int k = 0; k < m.K; k++)
    {
        const idx f = m.c2f[(size_t)k * N + i];
        slot[((size_t)k * 5 + 4) * N + p] = f >= 0 ? lam[f] : T(0);
    }
    T *cellRows = slot + (size_t)m.K * 5 * N;
#pragma unroll
    for (int v = 0; v < 5; v++)
        cellRows[(size_t)v * N + p] = rhs[(size_t)v * N + i];
    cellRows[5 * N + p] = jd[i];
}

// the packed state: per position a 128-byte row (FP64; 64 bytes FP32) holding the increment
// (5, at kPackInc) and the state U (5, at kPackU), so that a neighbour costs one cache line;
// two such arrays, read and written alternately by the sweeps.  Each half of a row is written
// whole (the increment or the state and three zeros), so no write covers part of a 32-byte
// sector.
constexpr int kPackStride = 16, kPackInc = 0, kPackU = 8;

// Skipped colour passes (Gauss-Seidel updates only).  A symmetric sweep after the first
// begins with the colour the previous sweep ended with; that pass would read the same
// neighbour increments as the previous one and write the same values, so it is skipped,
// and the positions [fLo, fHi) of that colour keep their current increments in the array
// the sweep reads (xr) instead of the one it writes (xw).  The colour the previous sweep
// skipped is this sweep's last: until its own pass, its current increments are in xw,
// positions [lLo, lHi).  A neighbour in the first range is read from xr, in the second
// from xw, any other as usual (xw when already updated in this sweep, else xr).
struct SweepSkip
{
    idx fLo = 0, fHi = 0, lLo = 0, lHi = 0;
};
__device__ __forceinline__ bool sgsReadNew(const SweepSkip &sk, idx po, bool useNew)
{
    if (po >= sk.fLo && po < sk.fHi)
        return false;
    if (po >= sk.lLo && po < sk.lHi)
        return true;
    return useNew;
}

template <typename T>
__global__ void __launch_bounds__(256) kSGSColourPacked(DevMesh m, real gammaD, T alphaDiag, const T *__restrict__ xr,
                                                        T *__restrict__ xw, const T *__restrict__ slot,
                                                        const idx *__restrict__ nbr, int forward, int gsUpdate,
                                                        int incZero, idx begin, idx end, SweepSkip sk)
{
    constexpr int S = kPackStride;
    const idx p = begin + blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= end)
        return;
    const size_t N = m.N;
    const T gamma = (T)gammaD;
    const T *__restrict__ cellRows = slot + (size_t)m.K * 5 * N;
    T buf[5];
#pragma unroll
    for (int v = 0; v < 5; v++)
        buf[v] = cellRows[(size_t)v * N + p];
    const T voli = cellRows[6 * N + p];
    for (int k = 0; k < m.K; k++)
    {
        const idx po = nbr[(size_t)k * N + p];
        if (po < 0)
            continue;
        const bool earlier = forward ? po < p : po > p;
        const bool useNew = gsUpdate && earlier;
        if (incZero && !useNew) // the first sweep starts from a zero increment
            continue;
        const T *x = (sgsReadNew(sk, po, useNew) ? xw : xr) + (size_t)po * S;
        T U[5], dU[5], n[3], dF[5];
#pragma unroll
        for (int v = 0; v < 5; v++)
            U[v] = x[kPackU + v], dU[v] = x[kPackInc + v];
#pragma unroll
        for (int r = 0; r < 3; r++)
            n[r] = slot[((size_t)k * 5 + r) * N + p];
        const T area = slot[((size_t)k * 5 + 3) * N + p], lamf = slot[((size_t)k * 5 + 4) * N + p];
        fluxJacDu<T>(U, n, dU, lamf, gamma, dF);
        T coef = (T(0.5) * alphaDiag) * area / voli;
#pragma unroll
        for (int v = 0; v < 5; v++)
            buf[v] -= coef * dF[v];
    }
    const T d = cellRows[5 * N + p];
#pragma unroll
    for (int v = 0; v < 8; v++)
        xw[(size_t)p * S + kPackInc + v] = v < 5 ? buf[v] / d : T(0);
}

// The same pass with the slot count K fixed at compile time: all K neighbour positions are
// loaded first and then all K neighbour rows, so a thread waits for two memory round trips
// instead of two per slot (a colour pass is about one round of waves on the device, so the
// chain of dependent loads per thread is its run time).  Same arithmetic in the same slot
// order as kSGSColourPacked: bit-identical results.
template <typename T, int K>
constexpr int sgsPackedWaves()
{
    return sizeof(T) == 8 ? (K <= 4 ? 4 : 3) : 1; // FP64: the rows of 4 (6) slots in 128 (168) VGPRs
}
template <typename T, int K>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(sgsPackedWaves<T, K>())))
kSGSColourPackedK(DevMesh m, real gammaD, T alphaDiag,
                                                         const T *__restrict__ xr, T *__restrict__ xw,
                                                         const T *__restrict__ slot, const idx *__restrict__ nbr,
                                                         int forward, int gsUpdate, int incZero, idx begin, idx end,
                                                         SweepSkip sk)
{
    constexpr int S = kPackStride;
    const idx p = begin + blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= end)
        return;
    const size_t N = m.N;
    const T gamma = (T)gammaD;
    const T *__restrict__ cellRows = slot + (size_t)K * 5 * N;
    idx po[K];
#pragma unroll
    for (int k = 0; k < K; k++)
        po[k] = nbr[(size_t)k * N + p];
    // the gathers are unconditional (a slot that is not used reads the cell's own row, a valid
    // address): a load under a branch makes the compiler wait for every earlier load at the join
    bool use[K];
    T U[K][5], dU[K][5];
#pragma unroll
    for (int k = 0; k < K; k++)
    {
        const bool earlier = forward ? po[k] < p : po[k] > p;
        const bool useNew = gsUpdate && earlier;
        use[k] = po[k] >= 0 && !(incZero && !useNew); // the first sweep starts from a zero increment
        const T *x = use[k] ? (sgsReadNew(sk, po[k], useNew) ? xw : xr) + (size_t)po[k] * S : xr + (size_t)p * S;
#pragma unroll
        for (int v = 0; v < 5; v++)
            U[k][v] = x[kPackU + v], dU[k][v] = x[kPackInc + v];
    }
    T buf[5];
#pragma unroll
    for (int v = 0; v < 5; v++)
        buf[v] = cellRows[(size_t)v * N + p];
    const T voli = cellRows[6 * N + p];
#pragma unroll
    for (int k = 0; k < K; k++)
    {
        if (!use[k])
            continue;
        T n[3], dF[5];
#pragma unroll
        for (int r = 0; r < 3; r++)
            n[r] = slot[((size_t)k * 5 + r) * N + p];
        const T area = slot[((size_t)k * 5 + 3) * N + p], lamf = slot[((size_t)k * 5 + 4) * N + p];
        fluxJacDu<T>(U[k], n, dU[k], lamf, gamma, dF);
        T coef = (T(0.5) * alphaDiag) * area / voli;
#pragma unroll
        for (int v = 0; v < 5; v++)
            buf[v] -= coef * dF[v];
    }
    const T d = cellRows[5 * N + p];
#pragma unroll
    for (int v = 0; v < 8; v++)
        xw[(size_t)p * S + kPackInc + v] = v < 5 ? buf[v] / d : T(0);
}

// the state into both packed arrays in colour-major position order (the increments need no
// clearing: the first sweep reads none it has not written, see incZero)
template <typename T>
__global__ void __launch_bounds__(256) kToCellMajor(idx N, const idx *__restrict__ cellAt,
                                                    const real *__restrict__ soa, T *__restrict__ x0,
                                                    T *__restrict__ x1)
{
    constexpr int S = kPackStride;
    const idx p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= N)
        return;
    const idx i = cellAt[p];
#pragma unroll
    for (int v = 0; v < 8; v++)
    {
        const T a = v < 5 ? (T)soa[(size_t)v * N + i] : T(0);
        x0[(size_t)p * S + kPackU + v] = a, x1[(size_t)p * S + kPackU + v] = a;
    }
}

// the increments of positions [lo, hi) from one packed array into the other
template <typename T>
__global__ void __launch_bounds__(256) kCopyIncRows(idx lo, idx hi, const T *__restrict__ from, T *__restrict__ to)
{
    constexpr int S = kPackStride;
    const idx p = lo + blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= hi)
        return;
#pragma unroll
    for (int v = 0; v < 8; v++)
        to[(size_t)p * S + kPackInc + v] = from[(size_t)p * S + kPackInc + v];
}

template <typename T>
void launchSGSPackedT(Ctx &c, const GasParams &gp, T alphaDiag, const T *rhs, const real *u, const T *lam,
                      const T *jd, T *xbuf, T *slot, idx *nbr, bool fresh, int sgsIter, int gsUpdate, T **result)
{
    DevMesh &m = c.m;
    const int nb = gridFor(m.N, 256);
    if (fresh)
        kSGSSlotsStatic<T><<<nb, 256, 0, c.stream>>>(m, slot, nbr);
    T *x0 = xbuf, *x1 = xbuf + (size_t)kPackStride * m.N;
    kToCellMajor<T><<<nb, 256, 0, c.stream>>>(m.N, m.colorCells, u, x0, x1);
    kSGSSlots<T><<<nb, 256, 0, c.stream>>>(m, lam, rhs, jd, slot);
    DNDSR_HIP_CHECK(hipGetLastError());
    const int nC = m.nColor;
    const bool canSkip = gsUpdate && nC >= 2;
    int prevSkip = -1; // the colour whose pass the previous sweep skipped
    for (int s = 0; s <= sgsIter; s++)
        for (int dir = 1; dir >= 0; dir--)
        {
            const int first = dir ? 0 : nC - 1, last = dir ? nC - 1 : 0;
            const bool skipFirst = canSkip && !(s == 0 && dir == 1);
            SweepSkip sk;
            if (skipFirst)
                sk.fLo = m.colorPtrHost[first], sk.fHi = m.colorPtrHost[first + 1];
            if (prevSkip == last)
                sk.lLo = m.colorPtrHost[last], sk.lHi = m.colorPtrHost[last + 1];
            for (int q = skipFirst ? 1 : 0; q < nC; q++)
            {
                const int col = dir ? q : nC - 1 - q;
                const idx b = m.colorPtrHost[col], e = m.colorPtrHost[col + 1];
                if (e <= b)
                    continue;
                KernelTimer kt(c, K_SGS);
                const int g = gridFor(e - b, 256), z = (s == 0 && dir == 1) ? 1 : 0;
                if (m.K == 3)
                    kSGSColourPackedK<T, 3><<<g, 256, 0, c.stream>>>(m, gp.gamma, alphaDiag, x0, x1, slot, nbr, dir,
                                                                     gsUpdate, z, b, e, sk);
                else if (m.K == 4)
                    kSGSColourPackedK<T, 4><<<g, 256, 0, c.stream>>>(m, gp.gamma, alphaDiag, x0, x1, slot, nbr, dir,
                                                                     gsUpdate, z, b, e, sk);
                else if (m.K == 6)
                    kSGSColourPackedK<T, 6><<<g, 256, 0, c.stream>>>(m, gp.gamma, alphaDiag, x0, x1, slot, nbr, dir,
                                                                     gsUpdate, z, b, e, sk);
                else
                    kSGSColourPacked<T><<<g, 256, 0, c.stream>>>(m, gp.gamma, alphaDiag, x0, x1, slot, nbr, dir,
                                                                 gsUpdate, z, b, e, sk);
                DNDSR_HIP_CHECK(hipGetLastError());
            }
            prevSkip = skipFirst ? first : -1;
            std::swap(x0, x1);
        }
    if (prevSkip >= 0)
    {
        // the last sweep skipped its first colour: its increments are still in the array that
        // sweep read (now x1)
        const idx b = m.colorPtrHost[prevSkip], e = m.colorPtrHost[prevSkip + 1];
        if (e > b)
        {
            KernelTimer kt(c, K_SGS);
            kCopyIncRows<T><<<gridFor(e - b, 256), 256, 0, c.stream>>>(b, e, x1, x0);
            DNDSR_HIP_CHECK(hipGetLastError());
        }
    }
    *result = x0 + kPackInc; // the last increment, stride kPackStride
}
void launchSGSPacked64(Ctx &c, const GasParams &gp, real alphaDiag, const real *rhs, const real *u, const real *lam,
                       const real *jd, real *xbuf, real *slot, idx *nbr, bool fresh, int sgsIter, int gsUpdate,
                       real **result)
{
    launchSGSPackedT<real>(c, gp, alphaDiag, rhs, u, lam, jd, xbuf, slot, nbr, fresh, sgsIter, gsUpdate, result);
}
void launchSGSPacked32(Ctx &c, const GasParams &gp, float alphaDiag, const float *rhs, const real *u,
                       const float *lam, const float *jd, float *xbuf, float *slot, idx *nbr, bool fresh, int sgsIter,
                       int gsUpdate, float **result)
{
    launchSGSPackedT<float>(c, gp, alphaDiag, rhs, u, lam, jd, xbuf, slot, nbr, fresh, sgsIter, gsUpdate, result);
}

// ---------------------------------------------------------------------------
// Force on a boundary zone (the wall force of a drag functional): the momentum part of
// the face fluxes of the zone's faces from the last EvaluateRHS, with the sign of the
// flux out of the domain, summed in boundary-face order by one workgroup.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) kZoneForce(DevMesh m, int zone, const real *__restrict__ ff,
                                                  real *__restrict__ out)
{
    __shared__ real sm[3][256];
    real a[3] = {0, 0, 0};
    for (idx t = threadIdx.x; t < m.nBnd; t += 256)
    {
        const idx f = m.bndFace[t];
        if (m.fzone[f] != zone)
            continue;
#pragma unroll
        for (int d = 0; d < 3; d++)
            a[d] -= ff[(size_t)(1 + d) * m.F + f];
    }
#pragma unroll
    for (int d = 0; d < 3; d++)
        sm[d][threadIdx.x] = a[d];
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1)
    {
        if (threadIdx.x < s)
#pragma unroll
            for (int d = 0; d < 3; d++)
                sm[d][threadIdx.x] += sm[d][threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0)
#pragma unroll
        for (int d = 0; d < 3; d++)
            out[d] = sm[d][0];
}
void launchZoneForce(Ctx &c, int zone, real *forceDev)
{
    kZoneForce<<<1, 256, 0, c.stream>>>(c.m, zone, c.faceFlux, forceDev);
    DNDSR_HIP_CHECK(hipGetLastError());
}
__global__ void __launch_bounds__(256) kZoneForceBar(DevMesh m, int zone, const real *__restrict__ fb,
                                                     real *__restrict__ fbar)
{
    const idx f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= m.F)
        return;
    const bool on = m.f2c[(size_t)m.F + f] < 0 && m.fzone[f] == zone;
    fbar[f] = 0;
#pragma unroll
    for (int d = 0; d < 3; d++)
        fbar[(size_t)(1 + d) * m.F + f] = on ? -fb[d] : 0.0;
    fbar[(size_t)4 * m.F + f] = 0;
}
void launchZoneForceBar(Ctx &c, int zone, const real *forceBarDev, real *fbar)
{
    kZoneForceBar<<<gridFor(c.m.F, 256), 256, 0, c.stream>>>(c.m, zone, forceBarDev, fbar);
    DNDSR_HIP_CHECK(hipGetLastError());
}

// Positivity of the increment, per cell (fincrement, EulerSolver.hxx:696-738):
//   alpha_i = EvaluateCellRHSAlpha (EulerEvaluator.hxx:1938-2030; DEFAULT flag, compress 2,
//             with IdealGasGetCompressionRatioPressure<dim, 0>, Gas.hpp:1750-1860),
//   u_i    += CompressInc(u_i, alpha_i du_i) (AddFixedIncrement, EulerEvaluator.hpp:1645-1830).
// Limited cells add to stat->nLim and lower stat->alphaMinBits (the bit pattern of a positive
// double orders as the value); a cell that reaches one of the reference's assertions (non-positive
// density or internal energy, no real root, an undecayed increment) sets stat->bad.
__device__ inline real ppSq3(const real *a) { return a[0] * a[0] + a[1] * a[1] + a[2] * a[2]; }

__device__ real ppCompressionRatioPressure(const real *u, const real *inc, real newRhoEi, bool &bad)
{
    const real safetyRatio = 1 - 1e-5;
    real ret[5], un[5];
#pragma unroll
    for (int k = 0; k < 5; k++)
        ret[k] = inc[k], un[k] = u[k] + inc[k];
    const real rhoEOld = u[4] - ppSq3(u + 1) / (u[0] + kVerySmall) * 0.5;
    newRhoEi = fmax(kSmall * rhoEOld, newRhoEi);
    const real c0 = 2 * u[4] * u[0] - ppSq3(u + 1) - 2 * u[0] * newRhoEi;
    const real c1 = 2 * u[4] * ret[0] + 2 * u[0] * ret[4] -
                    2 * (u[1] * ret[1] + u[2] * ret[2] + u[3] * ret[3]) - 2 * ret[0] * newRhoEi;
    real c2 = 2 * ret[4] * ret[0] - ppSq3(ret + 1);
    c2 += (c2 >= 0 ? 1.0 : -1.0) * kVerySmall;
    real deltaC = c1 * c1 - 4 * c0 * c2;
    if (deltaC <= -c0 * c0 * kSmall)
        bad = true;
    deltaC = fmax(0., deltaC);
    const real alphaL = (-sqrt(deltaC) - c1) / (2 * c2);
    real alpha;
    if (fabs(c2) < 1e-10 * c0)
        alpha = fabs(c1) < 1e-10 * c0 ? 0 : fmin(-c0 / c1, 1.);
    else
        alpha = fmin(alphaL, 1.);
    alpha = fmax(0., alpha);
    alpha *= safetyRatio;
    if (alpha < kSmall)
        alpha = 0;
#pragma unroll
    for (int k = 0; k < 5; k++)
        ret[k] *= alpha;
    const real decay = 1 - 1e-2;
    int iter;
    for (iter = 0; iter < 1000; iter++)
    {
        real w[3] = {u[1] + ret[1], u[2] + ret[2], u[3] + ret[3]};
        const real ek = 0.5 * ppSq3(w) / (u[0] + ret[0] + kVerySmall);
        if (!(ret[4] + u[4] - ek < newRhoEi))
            break;
#pragma unroll
        for (int k = 0; k < 5; k++)
            ret[k] *= decay;
        alpha *= decay;
    }
    if (iter >= 1000)
        bad = true;
    return alpha;
}

__device__ real ppCellAlpha(const real *u, const real *res, real gamma, real relax, real rhoEps, real pEps,
                            bool &limited, bool &bad)
{
    const real safetyRatio = 1 - 1e-5;
    real inc[5];
    if (!(u[0] >= rhoEps))
        bad = true;
    real alphaRho = 1;
    const real relaxedRho = rhoEps * relax + u[0] * (1 - relax);
    if (res[0] < 0)
        alphaRho = fmin(1.0, (u[0] - relaxedRho) / (-res[0] - kSmall * res[0]));
    if (!(alphaRho >= 0 && alphaRho <= 1))
        bad = true;
    real un[5];
#pragma unroll
    for (int k = 0; k < 5; k++)
        inc[k] = res[k] * alphaRho, un[k] = u[k] + inc[k];
    const real pNew = (un[4] - 0.5 * ppSq3(un + 1) / un[0]) * (gamma - 1);
    const real pOld = (u[4] - 0.5 * ppSq3(u + 1) / u[0]) * (gamma - 1);
    real relaxedP = pEps;
    if (pNew < pOld)
        relaxedP = pEps + (pOld - pEps) * (1 - relax);
    real alphaP = 1;
    if (pNew < relaxedP)
        alphaP = fmin(alphaP, ppCompressionRatioPressure(u, inc, relaxedP / (gamma - 1), bad));
    real alpha = alphaRho * alphaP;
    limited = alpha < 1;
    if (limited)
        alpha = alpha * alpha * safetyRatio; // uRecAlphaCompressPower 1, compress 2
    return alpha;
}

__device__ void ppCompressInc(const real *u, real *ret, real gamma, real rhoEps, real pEps, bool &bad)
{
    if (!(u[0] > 0))
        bad = true;
    if (u[0] + ret[0] <= rhoEps)
    {
        const real f = (rhoEps - u[0]) / (ret[0] - kVerySmall);
#pragma unroll
        for (int k = 0; k < 5; k++)
            ret[k] *= f;
    }
    const real rhoEi = u[4] - 0.5 * ppSq3(u + 1) / (u[0] + kVerySmall);
    if (!(rhoEi > 0))
        bad = true;
    real w[3] = {u[1] + ret[1], u[2] + ret[2], u[3] + ret[3]};
    real ek = 0.5 * ppSq3(w) / (u[0] + ret[0] + kVerySmall);
    if (u[4] + ret[4] - ek <= pEps)
    {
        const real newRhoEi = pEps / (gamma - 1);
        const real c0 = 2 * u[4] * u[0] - ppSq3(u + 1) - 2 * u[0] * newRhoEi;
        const real c1 = 2 * u[4] * ret[0] + 2 * u[0] * ret[4] -
                        2 * (u[1] * ret[1] + u[2] * ret[2] + u[3] * ret[3]) - 2 * ret[0] * newRhoEi;
        const real c2 = 2 * ret[4] * ret[0] - ppSq3(ret + 1);
        const real deltaC = c1 * c1 - 4 * c0 * c2;
        if (!(deltaC > 0))
            bad = true;
        const real alpha = fmax(0., fmin((-sqrt(deltaC) - c1) / (2 * c2), 1.));
#pragma unroll
        for (int k = 0; k < 5; k++)
            ret[k] *= alpha * 0.99;
        const real decay = 1 - 1e-1;
        for (int iter = 0; iter < 1000; iter++)
        {
            real v[3] = {u[1] + ret[1], u[2] + ret[2], u[3] + ret[3]};
            ek = 0.5 * ppSq3(v) / (u[0] + ret[0] + kVerySmall);
            if (!(ret[4] + u[4] - ek < newRhoEi))
                break;
#pragma unroll
            for (int k = 0; k < 5; k++)
                ret[k] *= decay;
        }
        real v[3] = {u[1] + ret[1], u[2] + ret[2], u[3] + ret[3]};
        ek = 0.5 * ppSq3(v) / (u[0] + ret[0] + kVerySmall);
        if (ret[4] + u[4] - ek < newRhoEi * 0.5)
            bad = true;
    }
}

// CELLMAJOR: du is the packed sweep's increment, rows of kPackStride in colour-major position
// order (pos = the cell's position)
template <typename T, bool CELLMAJOR = false>
__global__ void __launch_bounds__(256) kAddIncrement(idx N, real gamma, real relax, real rhoEps, real pEps,
                                                     real *__restrict__ u, const T *__restrict__ du,
                                                     PPStat *__restrict__ stat, const idx *__restrict__ pos = nullptr)
{
    const idx i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N)
        return;
    real U[5], D[5];
    const size_t pi = CELLMAJOR ? (size_t)pos[i] : (size_t)i;
#pragma unroll
    for (int v = 0; v < 5; v++)
        U[v] = u[(size_t)v * N + i], D[v] = (real)du[CELLMAJOR ? pi * kPackStride + v : (size_t)v * N + i];
    bool limited = false, bad = false;
    const real a = ppCellAlpha(U, D, gamma, relax, rhoEps, pEps, limited, bad);
#pragma unroll
    for (int v = 0; v < 5; v++)
        D[v] *= a;
    ppCompressInc(U, D, gamma, rhoEps, pEps, bad);
#pragma unroll
    for (int v = 0; v < 5; v++)
        U[v] += D[v];
    if (!(isfinite(U[0]) && isfinite(U[1]) && isfinite(U[2]) && isfinite(U[3]) && isfinite(U[4])))
        bad = true;
    if (limited)
    {
        atomicAdd(&stat->nLim, 1);
        atomicMin(&stat->alphaMinBits, (unsigned long long)__double_as_longlong(a));
    }
    if (bad)
        atomicOr(&stat->bad, 1);
#pragma unroll
    for (int v = 0; v < 5; v++)
        u[(size_t)v * N + i] = U[v];
}
__global__ void kPPStatReset(PPStat *stat)
{
    if (threadIdx.x == 0)
    {
        stat->alphaMinBits = (unsigned long long)__double_as_longlong(1.0);
        stat->nLim = 0;
        stat->bad = 0;
    }
}
void launchPPStatReset(Ctx &c, PPStat *stat)
{
    kPPStatReset<<<1, 64, 0, c.stream>>>(stat);
    DNDSR_HIP_CHECK(hipGetLastError());
}
void launchAddIncrement64(Ctx &c, real gamma, real relax, real rhoEps, real pEps, real *u, const real *du, PPStat *stat)
{
    KernelTimer kt(c, K_INCREMENT);
    kAddIncrement<real><<<gridFor(c.m.N, 256), 256, 0, c.stream>>>(c.m.N, gamma, relax, rhoEps, pEps, u, du, stat);
    DNDSR_HIP_CHECK(hipGetLastError());
}
void launchAddIncrementCellMajor(Ctx &c, real gamma, real relax, real rhoEps, real pEps, real *u, const real *du64,
                                 const float *du32, PPStat *stat)
{
    KernelTimer kt(c, K_INCREMENT);
    if (du64)
        kAddIncrement<real, true><<<gridFor(c.m.N, 256), 256, 0, c.stream>>>(c.m.N, gamma, relax, rhoEps, pEps, u,
                                                                             du64, stat, c.m.cellPos);
    else
        kAddIncrement<float, true><<<gridFor(c.m.N, 256), 256, 0, c.stream>>>(c.m.N, gamma, relax, rhoEps, pEps, u,
                                                                              du32, stat, c.m.cellPos);
    DNDSR_HIP_CHECK(hipGetLastError());
}
void launchAddIncrement32(Ctx &c, real gamma, real relax, real rhoEps, real pEps, real *u, const float *du,
                          PPStat *stat)
{
    KernelTimer kt(c, K_INCREMENT);
    kAddIncrement<float><<<gridFor(c.m.N, 256), 256, 0, c.stream>>>(c.m.N, gamma, relax, rhoEps, pEps, u, du, stat);
    DNDSR_HIP_CHECK(hipGetLastError());
}

__global__ void __launch_bounds__(256) kToF32(size_t n, const real *__restrict__ in, float *__restrict__ out)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        out[i] = (float)in[i];
}
void launchToF32(Ctx &c, size_t n, const real *in, float *out)
{
    KernelTimer kt(c, K_CONVERT);
    kToF32<<<(unsigned)((n + 255) / 256), 256, 0, c.stream>>>(n, in, out);
    DNDSR_HIP_CHECK(hipGetLastError());
}

} // namespace dndsr
