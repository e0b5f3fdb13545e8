// dev_common.hpp -- shared device helpers, tables and the device mesh/state layout.
//
// Device layout (structure of arrays, cell- or face-index fastest so that a
// wavefront of consecutive cells/faces reads contiguous 512 B segments):
//   u        [5][N]            conservative means
//   uRec     [nVars][M][N]     reconstruction coefficients
//   rhs      [5][N]
//   AinvB    [K+1][M][M][N]    block 0 = A^-1, block 1+f = A^-1 B_f (K = max faces per cell)
//   Ainvb    [K][M][N]
//   face points / normals [G][3][F], w*JDet [G][F]
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace dndsr
{
using real = double;
using idx = int32_t;

constexpr real kVerySmall = 1e-200; // src/DNDS/Defines.hpp:199
constexpr real kSmall = 1e-10;      // src/DNDS/Defines.hpp:201
constexpr int kMaxG = 16;           // max face/cell quadrature points handled per entity
constexpr int kMaxFC = 6;           // max faces per cell (hexahedron)

// Monomial exponent tables in the reference's derivative / basis order
// (src/Geom/BaseFunction.hpp:53-90).  Same table for basis index and
// derivative-operator index.
__host__ __device__ inline void monoExp(int dim, int k, int &px, int &py, int &pz)
{
    // 3-D order: 1 x y z xx yy zz xy yz xz xxx yyy zzz xxy xyy yyz yzz xzz xxz xyz
    const int8_t e3[20][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {2, 0, 0}, {0, 2, 0}, {0, 0, 2}, {1, 1, 0}, {0, 1, 1}, {1, 0, 1}, {3, 0, 0}, {0, 3, 0}, {0, 0, 3}, {2, 1, 0}, {1, 2, 0}, {0, 2, 1}, {0, 1, 2}, {1, 0, 2}, {2, 0, 1}, {1, 1, 1}};
    // 2-D order: 1 x y xx xy yy xxx xxy xyy yyy
    const int8_t e2[10][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {2, 0, 0}, {1, 1, 0}, {0, 2, 0}, {3, 0, 0}, {2, 1, 0}, {1, 2, 0}, {0, 3, 0}};
    if (dim == 2)
        px = e2[k][0], py = e2[k][1], pz = e2[k][2];
    else
        px = e3[k][0], py = e3[k][1], pz = e3[k][2];
}

// number of polynomial DOFs for degree p (src/Geom/BaseFunction.hpp:233-273)
__host__ __device__ constexpr int nDof(int dim, int p)
{
    return dim == 2 ? (p == 0 ? 1 : p == 1 ? 3 : p == 2 ? 6 : 10) : (p == 0 ? 1 : p == 1 ? 4 : p == 2 ? 10 : 20);
}

__host__ __device__ inline real ipow(int p, real x)
{
    return p == 0 ? 1.0 : (p == 1 ? x : (p == 2 ? x * x : (p == 3 ? x * x * x : 1e300)));
}
__host__ __device__ inline int dfact(int p, int d)
{
    // p! / (p-d)! for d <= p (dFactorials, BaseFunction.hpp:192-195)
    if (d > p)
        return 0;
    int r = 1;
    for (int q = 0; q < d; q++)
        r *= (p - q);
    return r;
}

// Derivative `id` of basis monomial `ib` at scaled point xs with isotropic scale s
// (FPolynomial3D / FPolynomialFill, BaseFunction.hpp:222-227, 423-437).
__host__ __device__ inline real basisDiff(int dim, int id, int ib, const real *xs, real s)
{
    int px, py, pz, dx, dy, dz;
    monoExp(dim, ib, px, py, pz);
    monoExp(dim, id, dx, dy, dz);
    int c = dfact(px, dx) * dfact(py, dy) * dfact(pz, dz);
    if (!c)
        return 0.0;
    real v = c * ipow(px - dx, xs[0]) * ipow(py - dy, xs[1]) * ipow(pz - dz, xs[2]);
    real sz = dim == 2 ? 1.0 : s;
    return v / (ipow(dx, s) * ipow(dy, s) * ipow(dz, sz));
}

// Row-0 basis values (j = 1..NDOF-1) minus the cell moment, i.e. the
// zero-mean D0 row used by reconstruction (VariationalReconstruction.hpp:317-369).
template <int DIM, int NDOF>
__device__ inline void basisRow0(const real *xs, const real *moment /*[NDOF-1]*/, real *phi)
{
#pragma unroll
    for (int j = 1; j < NDOF; j++)
    {
        int px, py, pz;
        monoExp(DIM, j, px, py, pz);
        phi[j - 1] = ipow(px, xs[0]) * ipow(py, xs[1]) * ipow(pz, xs[2]) - moment[j - 1];
    }
}

// Element i of a row whose base address is wave-uniform.  The 64-bit base stays
// in SGPRs and each lane supplies one 32-bit byte offset (global_load saddr form),
// so a lane needs one VGPR per distinct index instead of two per load address.
// Every device row is below 4 GiB (entities per row < 2^29 doubles).
template <class T>
__device__ __forceinline__ T ldu(const T *__restrict__ base, uint32_t i)
{
    return *reinterpret_cast<const T *>(reinterpret_cast<const char *>(base) + (uint32_t)(i * (uint32_t)sizeof(T)));
}
template <class T>
__device__ __forceinline__ void stu(T *__restrict__ base, uint32_t i, T v)
{
    *reinterpret_cast<T *>(reinterpret_cast<char *>(base) + (uint32_t)(i * (uint32_t)sizeof(T))) = v;
}
// Raw buffer access to one device row (CDNA buffer descriptor: 48-bit base, 32-bit
// size in bytes, word 3 = 0x00020000 on gfx9-family parts).  The descriptor lives in
// SGPRs, each lane supplies a 32-bit byte offset and the row's scalar offset goes in
// soffset, so streaming many rows per lane costs no VGPRs for addresses.  A read at
// an offset at or beyond the row size returns 0 (hardware range check).
using Rsrc = __amdgpu_buffer_rsrc_t;
__device__ __forceinline__ Rsrc rowRsrc(const void *base, uint64_t bytes)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(base), (short)0,
                                             (int)(uint32_t)(bytes < 0xFFFFFFFFull ? bytes : 0xFFFFFFFFull), 0x00020000);
}
__device__ __forceinline__ double bld(Rsrc r, uint32_t voff, uint32_t soff = 0)
{
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
}
// streaming read (non-temporal hint: the line is not kept in preference to reused data)
__device__ __forceinline__ double bldnt(Rsrc r, uint32_t voff, uint32_t soff = 0)
{
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 2));
}
__device__ __forceinline__ int32_t bldi(Rsrc r, uint32_t voff, uint32_t soff = 0)
{
    return (int32_t)__builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0);
}
constexpr uint32_t kNoRow = 0xFFFFFFF0u; // byte offset that reads as 0

// wavefront index inside the block, as a scalar (uniform) value
__device__ __forceinline__ int waveIndex(int waveSize = 64)
{
    return __builtin_amdgcn_readfirstlane((int)threadIdx.x / waveSize);
}

// MI355X dispatches the workgroups of a launch round-robin over its 8 XCDs, each
// with its own L2.  Remap so that XCD x processes one contiguous range of tiles:
// neighbouring tiles (and the cells / faces they share) then stay in one L2.
constexpr int kNumXCD = 8;
__device__ inline int xcdTile(int b, int nb)
{
    const int per = nb / kNumXCD, rem = nb % kNumXCD;
    const int x = b % kNumXCD, k = b / kNumXCD;
    return x * per + (x < rem ? x : rem) + k;
}

struct GasParams
{
    real gamma;
    int rsType;
    int rsMeanValueEig;
    real fixScale;
    real incFScale;
    // laminar viscosity (idealGasProperty, src/Euler/EulerEvaluatorSettings.hpp:280-330)
    int viscous;  // 0: inviscid residual and time step
    int muModel;  // 0 constant mu, 1 Sutherland, 2 constant nu
    real Rgas, muGas, prGas, TRef, CSutherland;
};

} // namespace dndsr
