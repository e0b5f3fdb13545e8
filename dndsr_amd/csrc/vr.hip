// vr.hip -- variational-reconstruction iteration kernels.
//
// DoReconstructionIter, Jacobi branch (SORInstead = false), SOR branch, the recordInc
// modes and the second-order substitute (subs2ndOrder),
// src/CFV/VariationalReconstruction_Reconstruction.hxx:184-615, and
// DoReconstructionIterDiff, :617-670.
#include "gas.hpp"
#include "internal.hpp"

namespace dndsr
{
// Tile kernel: 64 consecutive cells per block, one wavefront per row r of the
// M x M blocks (lane = cell).  For each face slot k the neighbours' M x nVars
// coefficient blocks and means are staged once into LDS (coalesced), then every
// row-wave applies its row of A^-1 B_k, streamed from HBM exactly once.  The
// loop is software-pipelined over the face slots: the neighbour rows and the
// matrix row of slot k+1 are loaded into registers while slot k is applied, and
// the LDS stage is double-buffered, so one barrier per slot suffices and two
// slots' worth of loads are in flight per wave.
// Modes (DIFF = false):
//   0: uRecNew = (1-w) uRec + w sum_f (A^-1B_f uRec_j + A^-1b_f (u_j - u_i))
//   1: uRecNew = uRec - sum_f (...)            (recordInc: minus the block-Jacobi residual)
//   2: uRecNew = - sum_f A^-1b_f (u_j - u_i)   (recordInc with uRecIsZero)
// DIFF = true: uRecNew = x - sum_f A^-1B_f x_j (DoReconstructionIterDiff).
constexpr int kTileC = 64;

template <int M>
struct VRRows
{
    static constexpr int RW = M <= 16 ? M : (M + 1) / 2; // row-waves per block
    static constexpr int RPT = (M + RW - 1) / RW;        // rows per thread
};

// KIND: 0 = Jacobi iteration (mode 0), 3 = recordInc (mode 1), 1 = recordInc with
// uRecIsZero (mode 2), 2 = IterDiff.  The staged mean rows hold u_j - u_i.
template <int M, int NV, int KIND>
__global__ void __launch_bounds__(kTileC *VRRows<M>::RW) __attribute__((amdgpu_waves_per_eu(5)))
kVRIter(DevMesh m, const real *__restrict__ uRec, real *__restrict__ uRecNew, const real *__restrict__ u, int mode,
        real relax)
{
    constexpr bool DIFF = KIND == 2, REC = KIND != 1, RELAX = KIND == 0;
    constexpr int RW = VRRows<M>::RW, RPT = VRRows<M>::RPT;
    constexpr int NROW = (REC ? NV * M : 0) + (DIFF ? 0 : NV); // staged neighbour rows
    constexpr int R0 = REC ? NV * M : 0;                       // first mean row
    constexpr int PER = (NROW + RW - 1) / RW;                  // staged rows per lane
    __shared__ real sNb[2][PER * RW][kTileC]; // padded so that every lane stores PER rows
    const int lane = threadIdx.x % kTileC, w = waveIndex(kTileC);
    const int tile = xcdTile(blockIdx.x, gridDim.x);
    const size_t N = m.N;
    const int K = m.K;
    const idx i = tile * kTileC + lane;
    const bool valid = i < m.N;
    const uint32_t ioff = (uint32_t)i * 8u; // reads past the row (tail lanes) return 0
    const uint64_t rowB = (uint64_t)N * 8;
    auto row = [&](const real *base) { return rowRsrc(base, rowB); };
    real acc[RPT][NV];
    const real cf = RELAX ? (1 - relax) : 1.0;
#pragma unroll
    for (int t = 0; t < RPT; t++)
    {
        const int r = w + t * RW < M ? w + t * RW : M - 1;
#pragma unroll
        for (int v = 0; v < NV; v++)
            acc[t][v] = REC ? bld(row(uRec + ((size_t)v * M + r) * N), ioff) * cf : 0.0;
    }
    // own mean of the variable this lane stages in the mean rows (q = R0 + v)
    real uOwn[PER];
#pragma unroll
    for (int t = 0; t < PER; t++)
    {
        const int q = w + t * RW;
        uOwn[t] = (!DIFF && q >= R0 && q < NROW) ? bld(row(u + (size_t)(q - R0) * N), ioff) : 0.0;
    }

    // All loads are raw buffer loads with the row descriptor in SGPRs; absent
    // neighbours (o < 0) read 0 through an out-of-range offset, so no load is
    // predicated and the compiler keeps them in flight together.
    auto loadNb = [&](idx o, real *nb) {
        const uint32_t ooff = o >= 0 ? (uint32_t)o * 8u : kNoRow;
#pragma unroll
        for (int t = 0; t < PER; t++)
        {
            const int q = w + t * RW < NROW ? w + t * RW : NROW - 1;
            const real *base = q < R0 ? uRec + (size_t)q * N : u + (size_t)(q - R0) * N;
            nb[t] = bld(row(base), ooff);
        }
    };
    auto storeNb = [&](int buf, const real *nb) {
#pragma unroll
        for (int t = 0; t < PER; t++)
            sNb[buf][w + t * RW][lane] = nb[t] - uOwn[t];
    };
    // matrix rows of slot k (AinvB / Ainvb hold K slots for every cell), streamed once
    // with the non-temporal hint so that they do not evict the neighbours' rows from L2
    auto loadB = [&](int k, real (*br)[M], real *bb) {
        const real *B = m.AinvB + ((size_t)(1 + k) * M * M) * N;
        const real *bv = m.Ainvb + ((size_t)k * M) * N;
#pragma unroll
        for (int t = 0; t < RPT; t++)
        {
            const int r = w + t * RW < M ? w + t * RW : M - 1;
            if constexpr (REC)
#pragma unroll
                for (int q = 0; q < M; q++)
                    br[t][q] = bldnt(row(B + ((size_t)r * M + q) * N), ioff);
            bb[t] = DIFF ? 0.0 : bldnt(row(bv + (size_t)r * N), ioff);
        }
    };
    auto nbIndex = [&](int k) -> idx {
        return valid ? bldi(rowRsrc(m.c2c + (size_t)k * N, (uint64_t)N * 4), (uint32_t)i * 4u) : -1;
    };

    real brow[RPT][M], bb[RPT];
    auto apply = [&](int buf, idx oc) {
        if (oc < 0)
            return;
#pragma unroll
        for (int t = 0; t < RPT; t++)
        {
            const int r = w + t * RW;
            if (r >= M)
                continue;
#pragma unroll
            for (int v = 0; v < NV; v++)
            {
                real s = 0;
                if constexpr (REC)
#pragma unroll
                    for (int q = 0; q < M; q++)
                        s += brow[t][q] * sNb[buf][v * M + q][lane];
                if constexpr (DIFF)
                    acc[t][v] -= s;
                else
                {
                    s += bb[t] * sNb[buf][R0 + v][lane];
                    if constexpr (RELAX)
                        acc[t][v] += relax * s;
                    else
                        acc[t][v] -= s;
                }
            }
        }
    };

    // Slot pipeline.  Iteration k: issue the neighbour rows of slot k+1, barrier,
    // apply slot k (its matrix rows were issued during iteration k-1), issue the
    // matrix rows of slot k+1 into the registers just freed, then park the
    // neighbour rows of slot k+1 in the other LDS buffer.  The neighbour index of
    // slot k+2 is fetched one iteration ahead.  The last slot is peeled so that
    // the loop body is straight-line code (exact waits on the in-order vmcnt).
    idx o = nbIndex(0);
    idx on = K > 1 ? nbIndex(1) : -1;
    {
        real nb[PER];
        loadNb(o, nb);
        loadB(0, brow, bb);
        storeNb(0, nb);
    }
    for (int k = 0; k + 1 < K; k++)
    {
        const int buf = k & 1;
        const int k2 = k + 2 < K ? k + 2 : K - 1;
        const idx onn = nbIndex(k2);
        real nbN[PER];
        loadNb(on, nbN);
        __syncthreads(); // sNb[buf] complete; sNb[buf ^ 1] no longer read
        apply(buf, o);
        loadB(k + 1, brow, bb);
        storeNb(buf ^ 1, nbN);
        o = on;
        on = onn;
    }
    __syncthreads();
    apply((K - 1) & 1, o);
    if (!valid)
        return;
#pragma unroll
    for (int t = 0; t < RPT; t++)
    {
        const int r = w + t * RW;
        if (r < M)
#pragma unroll
            for (int v = 0; v < NV; v++)
                stu(uRecNew + ((size_t)v * M + r) * N, (uint32_t)i, acc[t][v]);
    }
}

template <int M, int KIND>
static void vrLaunch(Ctx &c, int nVars, const real *uRec, real *uRecNew, const real *u, int mode, real relax)
{
    int g = gridFor(c.m.N, kTileC), bs = kTileC * VRRows<M>::RW;
    if (nVars == 5)
        kVRIter<M, 5, KIND><<<g, bs, 0, c.stream>>>(c.m, uRec, uRecNew, u, mode, relax);
    else if (nVars == 1)
        kVRIter<M, 1, KIND><<<g, bs, 0, c.stream>>>(c.m, uRec, uRecNew, u, mode, relax);
    else
        throw std::runtime_error("nVars must be 1 or 5");
    DNDSR_HIP_CHECK(hipGetLastError());
}
template <int M>
static void vrIterM(Ctx &c, int nVars, const real *uRec, real *uRecNew, const real *u, int mode, real relax)
{
    if (mode == 2)
        vrLaunch<M, 1>(c, nVars, uRec, uRecNew, u, mode, relax);
    else if (mode == 1)
        vrLaunch<M, 3>(c, nVars, uRec, uRecNew, u, mode, relax);
    else
        vrLaunch<M, 0>(c, nVars, uRec, uRecNew, u, mode, relax);
}
template <int M>
static void vrIterDiffM(Ctx &c, int nVars, const real *x, real *y)
{
    vrLaunch<M, 2>(c, nVars, x, y, nullptr, 0, 1.0);
}


// ---------------------------------------------------------------------------
// Boundary term of the sweep (GetBoundaryRHS / GetBoundaryRHSDiff,
// VariationalReconstruction_Reconstruction.hxx:19-176, intOrderVRBC = intOrder):
// one thread per cell that owns boundary faces, run after the tile kernel (whose
// boundary slots contribute nothing).  Per boundary face of the cell
//   BCC = sum_g D0phi(x_g)^T w_0^2 (uB(x_g) - u_i) JDet w_g,   uRecNew += scale A^-1 BCC
// with uB the ghost of the compressed point state u_i + D0phi uRec (Euler table, nVars = 5,
// the solver's FBoundary, EulerSolver.hxx:163-175) or the given point value.
// MODE 0: Jacobi (scale = relax); 1: recordInc (scale = -1); 2: the Jacobian action
// (uB(uBL + D0phi x) - uB(uBL) in place of uB - u_i, scale = -1).
// ---------------------------------------------------------------------------
// BCC of boundary face f of cell i (mom, cen, isc: the cell's moments, centre and scale;
// um its mean): the own block is read through rec(v, j) and, for LIN, the Jacobian
// direction through dx(v, j).
template <int DIM, int NDOF, int NV, bool LIN, class Rec, class Dx>
__device__ inline void boundaryFaceBCC(const DevMesh &m, idx f, const real *mom, const real *cen, real isc,
                                       const real *um, Rec rec, Dx dx, real (&BCC)[NDOF - 1][NV])
{
    constexpr int M = NDOF - 1;
    const size_t F = m.F;
    const bool euler = NV == 5 && m.bcType != nullptr;
    const int zone = m.fzone[f];
    const real w0 = m.faceW[f]; // order-0 weight; faceL^0 = 1
    const real w2 = w0 * w0;
#pragma unroll
    for (int j = 0; j < M; j++)
#pragma unroll
        for (int v = 0; v < NV; v++)
            BCC[j][v] = 0;
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
        real inc[NV];
        if (euler)
        {
            real uBL[5], n[3], ULf[5], uB[5];
#pragma unroll
            for (int v = 0; v < 5; v++)
            {
                real r = 0;
#pragma unroll
                for (int j = 0; j < M; j++)
                    r += phi[j] * rec(v, j);
                uBL[v] = r;
            }
#pragma unroll
            for (int r = 0; r < 3; r++)
                n[r] = m.fnorm[((size_t)g * 3 + r) * F + f];
            const real *far = m.bcVal + (size_t)zone * 5;
            const int type = m.bcType[zone];
            compressRecPart(um, uBL, ULf);
            bcGhost(type, ULf, n, far, m.bcGamma, uB);
            if constexpr (LIN)
            {
                real in2[5], ULp[5], uBp[5];
#pragma unroll
                for (int v = 0; v < 5; v++)
                {
                    real r = 0;
#pragma unroll
                    for (int j = 0; j < M; j++)
                        r += phi[j] * dx(v, j);
                    in2[v] = uBL[v] + r;
                }
                compressRecPart(um, in2, ULp);
                bcGhost(type, ULp, n, far, m.bcGamma, uBp);
#pragma unroll
                for (int v = 0; v < NV; v++)
                    inc[v] = uBp[v] - uB[v];
            }
            else
#pragma unroll
                for (int v = 0; v < NV; v++)
                    inc[v] = uB[v] - um[v];
        }
        else
#pragma unroll
            for (int v = 0; v < NV; v++)
                inc[v] = LIN ? 0.0 : m.dirVal[((size_t)v * m.G + g) * F + f] - um[v];
        const real fac = m.fwj[(size_t)g * F + f];
#pragma unroll
        for (int j = 0; j < M; j++)
#pragma unroll
            for (int v = 0; v < NV; v++)
                BCC[j][v] += ((phi[j] * inc[v]) * w2) * fac;
    }
}

// The boundary term in two passes.  kVRBndFace: one thread per boundary face, A^-1 BCC of the
// face (BCC from the owner cell's block, A^-1 = the cell's slot-0 block) into vrBnd; kVRBndCell:
// one thread per boundary cell adds scale times its faces' terms in slot order.  (One thread per
// boundary cell doing both would run, in every wave, each slot that any lane has on the
// boundary.)
template <int DIM, int NDOF, int NV, int MODE>
__global__ void __launch_bounds__(64) kVRBndFace(DevMesh m, const real *__restrict__ uRec, const real *__restrict__ u,
                                                 const real *__restrict__ x, real *__restrict__ out)
{
    constexpr int M = NDOF - 1;
    const idx t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= m.nBnd)
        return;
    const idx f = m.bndFace[t];
    const idx i = m.f2c[f];
    const size_t N = m.N, NB = m.nBnd;
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
        um[v] = u ? u[(size_t)v * N + i] : 0.0;
    auto rec = [&](int v, int j) { return uRec[((size_t)v * M + j) * N + i]; };
    auto dx = [&](int v, int j) { return x[((size_t)v * M + j) * N + i]; };
    real BCC[M][NV];
    boundaryFaceBCC<DIM, NDOF, NV, MODE == 2>(m, f, mom, cen, isc, um, rec, dx, BCC);
#pragma unroll
    for (int v = 0; v < NV; v++)
#pragma unroll
        for (int r = 0; r < M; r++)
        {
            real acc = 0;
#pragma unroll
            for (int c2 = 0; c2 < M; c2++)
                acc += m.AinvB[((size_t)r * M + c2) * N + i] * BCC[c2][v];
            out[((size_t)v * M + r) * NB + t] = acc;
        }
}

template <int M, int NV>
__global__ void __launch_bounds__(64) kVRBndCell(DevMesh m, const real *__restrict__ bt, real *__restrict__ uRecNew,
                                                 real scale)
{
    const idx j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= m.nBndCell)
        return;
    const idx i = m.bndCell[j];
    const size_t N = m.N, NB = m.nBnd, NC = m.nBndCell;
    for (int k = 0; k < m.K; k++)
    {
        const int t = m.bndSlot[(size_t)k * NC + j];
        if (t < 0)
            continue;
#pragma unroll
        for (int v = 0; v < NV; v++)
#pragma unroll
            for (int r = 0; r < M; r++)
                uRecNew[((size_t)v * M + r) * N + i] += scale * bt[((size_t)v * M + r) * NB + t];
    }
}

// ---------------------------------------------------------------------------
// SOR form of the sweep (SORInstead = true, recordInc = false,
// VariationalReconstruction_Reconstruction.hxx:534-589), multicolour: one launch per
// colour, one thread per cell of the colour, the cell's block kept in registers and
// updated in place.  Cells of one colour share no face, so a colour's neighbours are
// not written during its launch and the sweep equals the serial SOR in colour order.
// Per cell: uRec_i *= (1 - relax); per face in cell-face order
//   interior: uRec_i += relax (A^-1 B_f uRec_j + A^-1 b_f (u_j - u_i))
//   boundary: uRec_i += relax A^-1 BCC(uRec_i as accumulated so far), as the reference
//             hands GetBoundaryRHS the array being updated (:584).
// ---------------------------------------------------------------------------
template <int DIM, int NDOF, int NV, bool BND>
__global__ void __launch_bounds__(64) kVRSOR(DevMesh m, const idx *__restrict__ cells, idx nCells,
                                             real *__restrict__ uRec, const real *__restrict__ u, real relax)
{
    constexpr int M = NDOF - 1;
    const idx t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nCells)
        return;
    const idx i = cells[t];
    const size_t N = m.N;
    real acc[NV][M], um[NV];
#pragma unroll
    for (int v = 0; v < NV; v++)
    {
        um[v] = u[(size_t)v * N + i];
#pragma unroll
        for (int r = 0; r < M; r++)
            acc[v][r] = uRec[((size_t)v * M + r) * N + i] * (1 - relax);
    }
    for (int k = 0; k < m.K; k++)
    {
        const idx f = m.c2f[(size_t)k * N + i];
        if (f < 0)
            continue;
        const idx o = m.c2c[(size_t)k * N + i];
        if (o >= 0)
        {
            // the neighbour's block once, then the matrix row by row (each matrix entry read once)
            const real *B = m.AinvB + ((size_t)(1 + k) * M * M) * N;
            const real *bv = m.Ainvb + ((size_t)k * M) * N;
            real nb[NV][M], du[NV];
#pragma unroll
            for (int v = 0; v < NV; v++)
            {
                du[v] = u[(size_t)v * N + o] - um[v];
#pragma unroll
                for (int q = 0; q < M; q++)
                    nb[v][q] = uRec[((size_t)v * M + q) * N + o];
            }
#pragma unroll
            for (int r = 0; r < M; r++)
            {
                real br[M];
#pragma unroll
                for (int q = 0; q < M; q++)
                    br[q] = B[((size_t)r * M + q) * N + i];
                const real b = bv[(size_t)r * N + i];
#pragma unroll
                for (int v = 0; v < NV; v++)
                {
                    real s = 0;
#pragma unroll
                    for (int q = 0; q < M; q++)
                        s += br[q] * nb[v][q];
                    s += b * du[v];
                    acc[v][r] += relax * s;
                }
            }
            continue;
        }
        if constexpr (!BND) // the cell lists without a boundary face
            continue;
        real mom[M], cen[3];
#pragma unroll
        for (int j = 0; j < M; j++)
            mom[j] = m.moment[(size_t)j * N + i];
#pragma unroll
        for (int r = 0; r < 3; r++)
            cen[r] = m.cent[(size_t)r * N + i];
        real BCC[M][NV];
        auto rec = [&](int v, int j) { return acc[v][j]; };
        boundaryFaceBCC<DIM, NDOF, NV, false>(m, f, mom, cen, m.iscale[i], um, rec, rec, BCC);
        real d[NV][M];
#pragma unroll
        for (int v = 0; v < NV; v++)
#pragma unroll
            for (int r = 0; r < M; r++)
            {
                real a = 0;
#pragma unroll
                for (int c2 = 0; c2 < M; c2++)
                    a += m.AinvB[((size_t)r * M + c2) * N + i] * BCC[c2][v];
                d[v][r] = a;
            }
#pragma unroll
        for (int v = 0; v < NV; v++)
#pragma unroll
            for (int r = 0; r < M; r++)
                acc[v][r] += relax * d[v][r];
    }
#pragma unroll
    for (int v = 0; v < NV; v++)
#pragma unroll
        for (int r = 0; r < M; r++)
            uRec[((size_t)v * M + r) * N + i] = acc[v][r];
}

// ---------------------------------------------------------------------------
// Second-order substitute of the sweep (maxOrder = 1, subs2ndOrder != 0,
// DoReconstruction2ndGrad + DoReconstruction2nd, VariationalReconstruction_Reconstruction.hxx:
// 184-483): one thread per cell, a gradient from the means only.
//   1 (Gauss rule): g = (1/V) sum_f (u_j - u_i) lThis / (lThis + lThat) A_f n_f, a boundary
//     face adding (uB - u_i) A_f n_f / 2 (uB: ghost of u_i at the face centre, mean normal);
//     lThis = |x_f - bary_i|, lThat = |bary_j - x_f|, both 1 for subs2ndOrderGGScheme 0
//   2 (least squares): the normal equations of min sum_f |dx_f g^T - du_f|^2, dx_f = bary_j -
//     bary_i (boundary: the mirrored barycentre, -2 ((bary_i - x_f).n) n), du_f = u_j - u_i
//     (boundary: uB - u_i)
// and uRec_i = g^T times the cell's length scale (the first derivatives of the linear modes
// of the scaled Taylor basis are 1 / scale).
// ---------------------------------------------------------------------------
template <int DIM, int NV>
__global__ void __launch_bounds__(64) kVR2nd(DevMesh m, const real *__restrict__ u, real *__restrict__ out)
{
    const idx i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m.N)
        return;
    const size_t N = m.N, F = m.F;
    const int method = m.s.subs2ndOrder;
    const bool plain = m.s.subs2ndOrderGGScheme == 0;
    real ui[NV], bi[3], g[NV][DIM], S[DIM][DIM], T[DIM][NV];
#pragma unroll
    for (int v = 0; v < NV; v++)
        ui[v] = u[(size_t)v * N + i];
#pragma unroll
    for (int r = 0; r < 3; r++)
        bi[r] = m.bary[(size_t)r * N + i];
#pragma unroll
    for (int d = 0; d < DIM; d++)
    {
#pragma unroll
        for (int v = 0; v < NV; v++)
            g[v][d] = T[d][v] = 0;
#pragma unroll
        for (int e = 0; e < DIM; e++)
            S[d][e] = 0;
    }
    for (int k = 0; k < m.K; k++)
    {
        const idx f = m.c2f[(size_t)k * N + i];
        if (f < 0)
            continue;
        const idx o = m.c2c[(size_t)k * N + i];
        const bool back = m.f2c[f] == i;
        real xf[3], n[3], uB[NV];
#pragma unroll
        for (int r = 0; r < 3; r++)
        {
            xf[r] = m.fcent[(size_t)r * F + f] + (back ? 0.0 : m.fshift[(size_t)r * F + f]);
            n[r] = m.fmeanN[(size_t)r * F + f];
        }
        const real A = m.farea[f];
#pragma unroll
        for (int v = 0; v < NV; v++)
            uB[v] = ui[v];
        if (o < 0)
        {
            if constexpr (NV == 5)
            {
                const int zone = m.fzone[f];
                bcGhost(m.bcType[zone], ui, n, m.bcVal + (size_t)zone * 5, m.bcGamma, uB);
            }
        }
        real ob[3] = {0, 0, 0};
        if (o >= 0)
#pragma unroll
            for (int r = 0; r < 3; r++)
                ob[r] = m.bary[(size_t)r * N + o] + (back ? -m.fshift[(size_t)r * F + f] : m.fshift[(size_t)r * F + f]);
        if (method == 1)
        {
            real lThis = 0;
#pragma unroll
            for (int r = 0; r < 3; r++)
                lThis += (xf[r] - bi[r]) * (xf[r] - bi[r]);
            lThis = plain ? 1.0 : sqrt(lThis);
            if (o >= 0)
            {
                real lThat = 0;
#pragma unroll
                for (int r = 0; r < 3; r++)
                    lThat += (ob[r] - xf[r]) * (ob[r] - xf[r]);
                lThat = plain ? 1.0 : sqrt(lThat);
                const real sgn = back ? 1.0 : -1.0;
#pragma unroll
                for (int v = 0; v < NV; v++)
#pragma unroll
                    for (int d = 0; d < DIM; d++)
                        g[v][d] += (u[(size_t)v * N + o] - ui[v]) * lThis / (lThis + lThat + kVerySmall) * A *
                                   (n[d] * sgn);
            }
            else
#pragma unroll
                for (int v = 0; v < NV; v++)
#pragma unroll
                    for (int d = 0; d < DIM; d++)
                        g[v][d] += (uB[v] - ui[v]) * 0.5 * A * n[d];
        }
        else
        {
            real dc[DIM], du[NV];
            if (o >= 0)
            {
#pragma unroll
                for (int d = 0; d < DIM; d++)
                    dc[d] = ob[d] - bi[d];
#pragma unroll
                for (int v = 0; v < NV; v++)
                    du[v] = u[(size_t)v * N + o] - ui[v];
            }
            else
            {
                real pr = 0;
#pragma unroll
                for (int r = 0; r < 3; r++)
                    pr += (bi[r] - m.fcent[(size_t)r * F + f]) * n[r];
#pragma unroll
                for (int d = 0; d < DIM; d++)
                    dc[d] = pr * -2.0 * n[d];
#pragma unroll
                for (int v = 0; v < NV; v++)
                    du[v] = uB[v] - ui[v];
            }
#pragma unroll
            for (int d = 0; d < DIM; d++)
            {
#pragma unroll
                for (int e = 0; e < DIM; e++)
                    S[d][e] += dc[d] * dc[e];
#pragma unroll
                for (int v = 0; v < NV; v++)
                    T[d][v] += dc[d] * du[v];
            }
        }
    }
    if (method == 1)
    {
        const real V = m.vol[i];
#pragma unroll
        for (int v = 0; v < NV; v++)
#pragma unroll
            for (int d = 0; d < DIM; d++)
                g[v][d] /= V;
    }
    else if constexpr (DIM == 2)
    {
        const real det = S[0][0] * S[1][1] - S[0][1] * S[1][0];
#pragma unroll
        for (int v = 0; v < NV; v++)
        {
            g[v][0] = (S[1][1] * T[0][v] - S[0][1] * T[1][v]) / det;
            g[v][1] = (S[0][0] * T[1][v] - S[1][0] * T[0][v]) / det;
        }
    }
    else
    {
        real C[3][3]; // adjugate of S
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
#pragma unroll
        for (int v = 0; v < NV; v++)
#pragma unroll
            for (int d = 0; d < 3; d++)
                g[v][d] = (C[d][0] * T[0][v] + C[d][1] * T[1][v] + C[d][2] * T[2][v]) / det;
    }
    const real isc = m.iscale[i];
#pragma unroll
    for (int v = 0; v < NV; v++)
#pragma unroll
        for (int d = 0; d < DIM; d++)
            out[((size_t)v * DIM + d) * N + i] = g[v][d] / isc;
}

template <int MODE, int DIM, int NDOF>
static void vrBoundaryN(Ctx &c, int nVars, const real *uRec, const real *u, const real *x, real *uRecNew, real scale)
{
    constexpr int M = NDOF - 1;
    if (c.m.nBnd == 0)
        return;
    const int gf = gridFor(c.m.nBnd, 64), gc = gridFor(c.m.nBndCell, 64);
    if (nVars == 5)
    {
        kVRBndFace<DIM, NDOF, 5, MODE><<<gf, 64, 0, c.stream>>>(c.m, uRec, u, x, c.vrBnd);
        kVRBndCell<M, 5><<<gc, 64, 0, c.stream>>>(c.m, c.vrBnd, uRecNew, scale);
    }
    else
    {
        kVRBndFace<DIM, NDOF, 1, MODE><<<gf, 64, 0, c.stream>>>(c.m, uRec, u, x, c.vrBnd);
        kVRBndCell<M, 1><<<gc, 64, 0, c.stream>>>(c.m, c.vrBnd, uRecNew, scale);
    }
    DNDSR_HIP_CHECK(hipGetLastError());
}
static void requireBoundaryValues(const DevMesh &m, int nVars)
{
    const bool euler = nVars == 5 && m.bcType != nullptr;
    if (m.nBnd > 0 && !euler && !(m.dirVal && m.dirNV == nVars))
        throw std::runtime_error("the mesh has boundary faces: set the boundary conditions (dndsr_set_bc, nVars 5) "
                                 "or boundary point values (dndsr_set_vr_dirichlet) for this nVars");
}

template <int MODE>
static void vrBoundary(Ctx &c, int nVars, const real *uRec, const real *u, const real *x, real *uRecNew, real scale)
{
    const DevMesh &m = c.m;
    if (m.nBnd == 0)
        return;
    requireBoundaryValues(m, nVars);
    const bool euler = nVars == 5 && m.bcType != nullptr;
    if (MODE == 2 && euler && (!uRec || !u))
        throw std::runtime_error("the boundary term of the VR Jacobian needs uRec and u");
    if (m.dim == 2)
    {
        if (m.P == 1)
            vrBoundaryN<MODE, 2, 3>(c, nVars, uRec, u, x, uRecNew, scale);
        else if (m.P == 2)
            vrBoundaryN<MODE, 2, 6>(c, nVars, uRec, u, x, uRecNew, scale);
        else
            vrBoundaryN<MODE, 2, 10>(c, nVars, uRec, u, x, uRecNew, scale);
    }
    else
    {
        if (m.P == 1)
            vrBoundaryN<MODE, 3, 4>(c, nVars, uRec, u, x, uRecNew, scale);
        else if (m.P == 2)
            vrBoundaryN<MODE, 3, 10>(c, nVars, uRec, u, x, uRecNew, scale);
        else
            vrBoundaryN<MODE, 3, 20>(c, nVars, uRec, u, x, uRecNew, scale);
    }
}

void launchVRIter(Ctx &c, int nVars, const real *uRec, real *uRecNew, const real *u, int mode, real relax)
{
    if (usesSubs2nd(c.m)) // :503-512, every mode
    {
        launchVR2nd(c, nVars, u, uRecNew);
        return;
    }
    KernelTimer kt(c, K_VR_ITER);
    switch (c.m.M)
    {
    case 2: vrIterM<2>(c, nVars, uRec, uRecNew, u, mode, relax); break;
    case 3: vrIterM<3>(c, nVars, uRec, uRecNew, u, mode, relax); break;
    case 5: vrIterM<5>(c, nVars, uRec, uRecNew, u, mode, relax); break;
    case 9: vrIterM<9>(c, nVars, uRec, uRecNew, u, mode, relax); break;
    case 19: vrIterM<19>(c, nVars, uRec, uRecNew, u, mode, relax); break;
    default: throw std::runtime_error("unsupported number of reconstruction modes");
    }
    // boundary faces: recordInc with uRecIsZero takes no boundary term (:577-581)
    if (mode == 0)
        vrBoundary<0>(c, nVars, uRec, u, nullptr, uRecNew, relax);
    else if (mode == 1)
        vrBoundary<1>(c, nVars, uRec, u, nullptr, uRecNew, -1.0);
}

template <int DIM, int NDOF>
static void vrSORN(Ctx &c, int nVars, real *uRec, const real *u, real relax)
{
    const DevMesh &m = c.m;
    // per colour two launches: the cells without a boundary face (no boundary branch, fewer
    // registers), then those with one; cells of one colour are independent
    for (int col = 0; col < m.nColor; col++)
    {
        const idx b = m.colorPtrHost[col], s = m.sorSplitHost[col], e = m.colorPtrHost[col + 1];
        if (s > b)
        {
            if (nVars == 5)
                kVRSOR<DIM, NDOF, 5, false><<<gridFor(s - b, 64), 64, 0, c.stream>>>(m, m.sorCells + b, s - b, uRec, u,
                                                                                     relax);
            else
                kVRSOR<DIM, NDOF, 1, false><<<gridFor(s - b, 64), 64, 0, c.stream>>>(m, m.sorCells + b, s - b, uRec, u,
                                                                                     relax);
            DNDSR_HIP_CHECK(hipGetLastError());
        }
        if (e > s)
        {
            if (nVars == 5)
                kVRSOR<DIM, NDOF, 5, true><<<gridFor(e - s, 64), 64, 0, c.stream>>>(m, m.sorCells + s, e - s, uRec, u,
                                                                                    relax);
            else
                kVRSOR<DIM, NDOF, 1, true><<<gridFor(e - s, 64), 64, 0, c.stream>>>(m, m.sorCells + s, e - s, uRec, u,
                                                                                    relax);
            DNDSR_HIP_CHECK(hipGetLastError());
        }
    }
}

void launchVRSOR(Ctx &c, int nVars, real *uRec, const real *u, real relax)
{
    if (usesSubs2nd(c.m))
    {
        launchVR2nd(c, nVars, u, uRec);
        return;
    }
    KernelTimer kt(c, K_VR_ITER);
    const DevMesh &m = c.m;
    if (nVars != 5 && nVars != 1)
        throw std::runtime_error("nVars must be 1 or 5");
    requireBoundaryValues(m, nVars);
    if (m.dim == 2)
    {
        if (m.P == 1)
            vrSORN<2, 3>(c, nVars, uRec, u, relax);
        else if (m.P == 2)
            vrSORN<2, 6>(c, nVars, uRec, u, relax);
        else
            vrSORN<2, 10>(c, nVars, uRec, u, relax);
    }
    else
    {
        if (m.P == 1)
            vrSORN<3, 4>(c, nVars, uRec, u, relax);
        else if (m.P == 2)
            vrSORN<3, 10>(c, nVars, uRec, u, relax);
        else
            vrSORN<3, 20>(c, nVars, uRec, u, relax);
    }
}

bool usesSubs2nd(const DevMesh &m) { return m.P == 1 && m.s.subs2ndOrder != 0; }

void launchVR2nd(Ctx &c, int nVars, const real *u, real *out)
{
    KernelTimer kt(c, K_VR_ITER);
    const DevMesh &m = c.m;
    if (!usesSubs2nd(m))
        throw std::runtime_error("the second-order substitute needs maxOrder 1 and subs2ndOrder 1 or 2");
    if (nVars != 5 && nVars != 1)
        throw std::runtime_error("nVars must be 1 or 5");
    if (m.nBnd > 0 && !(nVars == 5 && m.bcType != nullptr))
        throw std::runtime_error("the second-order substitute needs the Euler boundary table (dndsr_set_bc, "
                                 "nVars 5) on meshes with boundary faces");
    const int g = gridFor(m.N, 64);
    if (m.dim == 2)
    {
        if (nVars == 5)
            kVR2nd<2, 5><<<g, 64, 0, c.stream>>>(m, u, out);
        else
            kVR2nd<2, 1><<<g, 64, 0, c.stream>>>(m, u, out);
    }
    else
    {
        if (nVars == 5)
            kVR2nd<3, 5><<<g, 64, 0, c.stream>>>(m, u, out);
        else
            kVR2nd<3, 1><<<g, 64, 0, c.stream>>>(m, u, out);
    }
    DNDSR_HIP_CHECK(hipGetLastError());
}

void launchVRIterDiff(Ctx &c, int nVars, const real *uRec, const real *u, const real *x, real *y)
{
    KernelTimer kt(c, K_VR_ITER);
    if (usesSubs2nd(c.m)) // :631-638: the substitute does not depend on uRec
    {
        DNDSR_HIP_CHECK(hipMemsetAsync(y, 0, sizeof(real) * (size_t)nVars * c.m.M * c.m.N, c.stream));
        return;
    }
    switch (c.m.M)
    {
    case 2: vrIterDiffM<2>(c, nVars, x, y); break;
    case 3: vrIterDiffM<3>(c, nVars, x, y); break;
    case 5: vrIterDiffM<5>(c, nVars, x, y); break;
    case 9: vrIterDiffM<9>(c, nVars, x, y); break;
    case 19: vrIterDiffM<19>(c, nVars, x, y); break;
    default: throw std::runtime_error("unsupported number of reconstruction modes");
    }
    vrBoundary<2>(c, nVars, uRec, u, x, y, -1.0);
}

} // namespace dndsr
