// vr.hip -- variational-reconstruction iteration kernels.
//
// DoReconstructionIter, Jacobi branch (SORInstead = false) and the recordInc
// modes, src/CFV/VariationalReconstruction_Reconstruction.hxx:485-615, and
// DoReconstructionIterDiff, :617-670.
#include "internal.hpp"

namespace dndsr
{
// Tile kernel: 64 consecutive cells per block, one wavefront per row r of the
// M x M blocks (lane = cell).  For each face slot k the neighbours' M x nVars
// coefficient blocks and means are staged once into LDS (coalesced), then every
// row-wave applies its row of A^-1 B_k, streamed from HBM exactly once.
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

template <int M, int NV, bool DIFF>
__global__ void __launch_bounds__(kTileC *VRRows<M>::RW) __attribute__((amdgpu_waves_per_eu(5))) kVRIter(DevMesh m, const real *__restrict__ uRec,
                                                                  real *__restrict__ uRecNew,
                                                                  const real *__restrict__ u, int mode, real relax)
{
    constexpr int RW = VRRows<M>::RW, RPT = VRRows<M>::RPT;
    constexpr int NROW = NV * M + (DIFF ? 0 : NV);
    __shared__ real sNb[NROW][kTileC];
    const int lane = threadIdx.x % kTileC, w = threadIdx.x / kTileC;
    const int tile = xcdTile(blockIdx.x, gridDim.x);
    const size_t N = m.N;
    const idx i = tile * kTileC + lane;
    const bool valid = i < m.N;
    real acc[RPT][NV];
    real ui[NV];
    const real cf = DIFF ? 1.0 : (mode == 0 ? (1 - relax) : (mode == 1 ? 1.0 : 0.0));
#pragma unroll
    for (int t = 0; t < RPT; t++)
    {
        const int r = w + t * RW;
#pragma unroll
        for (int v = 0; v < NV; v++)
            acc[t][v] = (valid && r < M && (DIFF || mode != 2)) ? uRec[((size_t)v * M + r) * N + i] * cf : 0.0;
    }
#pragma unroll
    for (int v = 0; v < NV; v++)
        ui[v] = (!DIFF && valid) ? u[(size_t)v * N + i] : 0.0;
    const bool needRec = DIFF || mode != 2;
    for (int k = 0; k < m.K; k++)
    {
        const idx f = valid ? m.c2f[(size_t)k * N + i] : -1;
        idx o = 0;
        if (f >= 0)
        {
            const idx L = m.f2c[f], R = m.f2c[(size_t)m.F + f];
            o = L == i ? R : L;
        }
        __syncthreads(); // previous slot's LDS reads are done
        for (int q = w; q < NROW; q += RW)
        {
            real val = 0;
            if (f >= 0)
            {
                if (q < NV * M)
                    val = needRec ? uRec[(size_t)q * N + o] : 0.0;
                else
                    val = u[(size_t)(q - NV * M) * N + o];
            }
            sNb[q][lane] = val;
        }
        __syncthreads();
        if (f < 0)
            continue;
        const real *B = m.AinvB + ((size_t)(1 + k) * M * M) * N + i;
        const real *bv = m.Ainvb + ((size_t)k * M) * N + i;
#pragma unroll
        for (int t = 0; t < RPT; t++)
        {
            const int r = w + t * RW;
            if (r >= M)
                continue;
            real brow[M];
            if (needRec)
#pragma unroll
                for (int q = 0; q < M; q++)
                    brow[q] = B[((size_t)r * M + q) * N];
            const real b = DIFF ? 0.0 : bv[(size_t)r * N];
#pragma unroll
            for (int v = 0; v < NV; v++)
            {
                real s = 0;
                if (needRec)
#pragma unroll
                    for (int q = 0; q < M; q++)
                        s += brow[q] * sNb[v * M + q][lane];
                if constexpr (DIFF)
                    acc[t][v] -= s;
                else
                {
                    s += b * (sNb[NV * M + v][lane] - ui[v]);
                    if (mode == 0)
                        acc[t][v] += relax * s;
                    else
                        acc[t][v] -= s;
                }
            }
        }
    }
    if (!valid)
        return;
#pragma unroll
    for (int t = 0; t < RPT; t++)
    {
        const int r = w + t * RW;
        if (r < M)
#pragma unroll
            for (int v = 0; v < NV; v++)
                uRecNew[((size_t)v * M + r) * N + i] = acc[t][v];
    }
}

template <int M>
static void vrIterM(Ctx &c, int nVars, const real *uRec, real *uRecNew, const real *u, int mode, real relax)
{
    int g = gridFor(c.m.N, kTileC), bs = kTileC * VRRows<M>::RW;
    if (nVars == 5)
        kVRIter<M, 5, false><<<g, bs, 0, c.stream>>>(c.m, uRec, uRecNew, u, mode, relax);
    else if (nVars == 1)
        kVRIter<M, 1, false><<<g, bs, 0, c.stream>>>(c.m, uRec, uRecNew, u, mode, relax);
    else
        throw std::runtime_error("nVars must be 1 or 5");
    DNDSR_HIP_CHECK(hipGetLastError());
}
template <int M>
static void vrIterDiffM(Ctx &c, int nVars, const real *x, real *y)
{
    int g = gridFor(c.m.N, kTileC), bs = kTileC * VRRows<M>::RW;
    if (nVars == 5)
        kVRIter<M, 5, true><<<g, bs, 0, c.stream>>>(c.m, x, y, nullptr, 0, 1.0);
    else if (nVars == 1)
        kVRIter<M, 1, true><<<g, bs, 0, c.stream>>>(c.m, x, y, nullptr, 0, 1.0);
    else
        throw std::runtime_error("nVars must be 1 or 5");
    DNDSR_HIP_CHECK(hipGetLastError());
}

void launchVRIter(Ctx &c, int nVars, const real *uRec, real *uRecNew, const real *u, int mode, real relax)
{
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
}

void launchVRIterDiff(Ctx &c, int nVars, const real *x, real *y)
{
    KernelTimer kt(c, K_VR_ITER);
    switch (c.m.M)
    {
    case 2: vrIterDiffM<2>(c, nVars, x, y); break;
    case 3: vrIterDiffM<3>(c, nVars, x, y); break;
    case 5: vrIterDiffM<5>(c, nVars, x, y); break;
    case 9: vrIterDiffM<9>(c, nVars, x, y); break;
    case 19: vrIterDiffM<19>(c, nVars, x, y); break;
    default: throw std::runtime_error("unsupported number of reconstruction modes");
    }
}

} // namespace dndsr
