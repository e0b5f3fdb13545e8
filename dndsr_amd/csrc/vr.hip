// vr.hip -- variational-reconstruction iteration kernels.
//
// DoReconstructionIter, Jacobi branch (SORInstead = false) and the recordInc
// modes, src/CFV/VariationalReconstruction_Reconstruction.hxx:485-615, and
// DoReconstructionIterDiff, :617-670.  One thread per cell: the cell's
// (K x M x M) A^-1 B blocks stream in coalesced (cell index fastest), the
// neighbour's M x nVars block is held in registers while the M rows are applied.
#include "internal.hpp"

namespace dndsr
{
// mode 0: uRecNew = (1-w) uRec + w sum_f (A^-1B_f uRec_j + A^-1b_f (u_j - u_i))
// mode 1: uRecNew = uRec - sum_f (...)            (recordInc: minus the block-Jacobi residual)
// mode 2: uRecNew = - sum_f A^-1b_f (u_j - u_i)   (recordInc with uRecIsZero)
template <int M, int NV>
__global__ void __launch_bounds__(256) kVRIter(DevMesh m, const real *__restrict__ uRec, real *__restrict__ uRecNew,
                                               const real *__restrict__ u, int mode, real relax)
{
    idx i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m.N)
        return;
    const size_t N = m.N;
    real acc[NV][M];
    real cf = mode == 0 ? (1 - relax) : (mode == 1 ? 1.0 : 0.0);
#pragma unroll
    for (int v = 0; v < NV; v++)
#pragma unroll
        for (int r = 0; r < M; r++)
            acc[v][r] = mode == 2 ? 0.0 : uRec[((size_t)v * M + r) * N + i] * cf;
    real ui[NV];
#pragma unroll
    for (int v = 0; v < NV; v++)
        ui[v] = u[(size_t)v * N + i];
    for (int k = 0; k < m.K; k++)
    {
        idx f = m.c2f[(size_t)k * N + i];
        if (f < 0)
            continue;
        idx L = m.f2c[f], R = m.f2c[(size_t)m.F + f];
        idx o = L == i ? R : L;
        real du[NV];
#pragma unroll
        for (int v = 0; v < NV; v++)
            du[v] = u[(size_t)v * N + o] - ui[v];
        const real *B = m.AinvB + ((size_t)(1 + k) * M * M) * N + i;
        const real *bv = m.Ainvb + ((size_t)k * M) * N + i;
        if (mode == 2)
        {
#pragma unroll
            for (int r = 0; r < M; r++)
            {
                real b = bv[(size_t)r * N];
#pragma unroll
                for (int v = 0; v < NV; v++)
                    acc[v][r] -= b * du[v];
            }
            continue;
        }
        real ro[NV][M];
#pragma unroll
        for (int v = 0; v < NV; v++)
#pragma unroll
            for (int q = 0; q < M; q++)
                ro[v][q] = uRec[((size_t)v * M + q) * N + o];
#pragma unroll
        for (int r = 0; r < M; r++)
        {
            real brow[M];
#pragma unroll
            for (int q = 0; q < M; q++)
                brow[q] = B[((size_t)r * M + q) * N];
            real b = bv[(size_t)r * N];
#pragma unroll
            for (int v = 0; v < NV; v++)
            {
                real s = 0;
#pragma unroll
                for (int q = 0; q < M; q++)
                    s += brow[q] * ro[v][q];
                s += b * du[v];
                if (mode == 0)
                    acc[v][r] += relax * s;
                else
                    acc[v][r] -= s;
            }
        }
    }
#pragma unroll
    for (int v = 0; v < NV; v++)
#pragma unroll
        for (int r = 0; r < M; r++)
            uRecNew[((size_t)v * M + r) * N + i] = acc[v][r];
}

// uRecNew = x - sum_f A^-1B_f x_j
template <int M, int NV>
__global__ void __launch_bounds__(256) kVRIterDiff(DevMesh m, const real *__restrict__ x, real *__restrict__ y)
{
    idx i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m.N)
        return;
    const size_t N = m.N;
    real acc[NV][M];
#pragma unroll
    for (int v = 0; v < NV; v++)
#pragma unroll
        for (int r = 0; r < M; r++)
            acc[v][r] = x[((size_t)v * M + r) * N + i];
    for (int k = 0; k < m.K; k++)
    {
        idx f = m.c2f[(size_t)k * N + i];
        if (f < 0)
            continue;
        idx L = m.f2c[f], R = m.f2c[(size_t)m.F + f];
        idx o = L == i ? R : L;
        const real *B = m.AinvB + ((size_t)(1 + k) * M * M) * N + i;
        real ro[NV][M];
#pragma unroll
        for (int v = 0; v < NV; v++)
#pragma unroll
            for (int q = 0; q < M; q++)
                ro[v][q] = x[((size_t)v * M + q) * N + o];
#pragma unroll
        for (int r = 0; r < M; r++)
        {
            real brow[M];
#pragma unroll
            for (int q = 0; q < M; q++)
                brow[q] = B[((size_t)r * M + q) * N];
#pragma unroll
            for (int v = 0; v < NV; v++)
            {
                real s = 0;
#pragma unroll
                for (int q = 0; q < M; q++)
                    s += brow[q] * ro[v][q];
                acc[v][r] -= s;
            }
        }
    }
#pragma unroll
    for (int v = 0; v < NV; v++)
#pragma unroll
        for (int r = 0; r < M; r++)
            y[((size_t)v * M + r) * N + i] = acc[v][r];
}

template <int M>
static void vrIterM(Ctx &c, int nVars, const real *uRec, real *uRecNew, const real *u, int mode, real relax)
{
    int g = gridFor(c.m.N, 256);
    if (nVars == 5)
        kVRIter<M, 5><<<g, 256, 0, c.stream>>>(c.m, uRec, uRecNew, u, mode, relax);
    else if (nVars == 1)
        kVRIter<M, 1><<<g, 256, 0, c.stream>>>(c.m, uRec, uRecNew, u, mode, relax);
    else
        throw std::runtime_error("nVars must be 1 or 5");
    DNDSR_HIP_CHECK(hipGetLastError());
}
template <int M>
static void vrIterDiffM(Ctx &c, int nVars, const real *x, real *y)
{
    int g = gridFor(c.m.N, 256);
    if (nVars == 5)
        kVRIterDiff<M, 5><<<g, 256, 0, c.stream>>>(c.m, x, y);
    else if (nVars == 1)
        kVRIterDiff<M, 1><<<g, 256, 0, c.stream>>>(c.m, x, y);
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
