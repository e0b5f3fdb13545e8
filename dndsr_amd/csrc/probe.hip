// probe.hip -- device bandwidth self-test: a read-dominated stream (every element
// read once, one partial sum per workgroup written) and a copy, timed with HIP events
// on the context's stream.  The roofline of the hot-path kernels is quoted against
// the datasheet peak and against what these two streams reach on the same device.
#include "internal.hpp"

namespace dndsr
{
typedef double dvec2 __attribute__((ext_vector_type(2))); // 16 bytes per lane

__global__ void __launch_bounds__(256) kStreamRead(const dvec2 *__restrict__ x, size_t n2, double *__restrict__ part)
{
    __shared__ double sm[256];
    double acc = 0;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n2; i += stride)
    {
        const dvec2 v = __builtin_nontemporal_load(x + i);
        acc += v.x + v.y;
    }
    sm[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1)
    {
        if (threadIdx.x < s)
            sm[threadIdx.x] += sm[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        part[blockIdx.x] = sm[0];
}

__global__ void __launch_bounds__(256) kStreamCopy(const dvec2 *__restrict__ x, dvec2 *__restrict__ y, size_t n2)
{
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n2; i += stride)
        __builtin_nontemporal_store(__builtin_nontemporal_load(x + i), y + i);
}

// mode 0: read src (dst receives the per-workgroup partial sums, at least nBlocks doubles);
// mode 1: copy src to dst.  Returns the average milliseconds of one pass over reps passes.
double streamProbe(Ctx &c, int mode, const double *src, double *dst, size_t n, int reps)
{
    int nCU = 0;
    DNDSR_HIP_CHECK(hipDeviceGetAttribute(&nCU, hipDeviceAttributeMultiprocessorCount, c.device));
    const int nb = nCU * 8; // eight 256-thread workgroups per compute unit
    const size_t n2 = n / 2;
    hipEvent_t a, b;
    DNDSR_HIP_CHECK(hipEventCreate(&a));
    DNDSR_HIP_CHECK(hipEventCreate(&b));
    auto launch = [&]() {
        if (mode == 0)
            kStreamRead<<<nb, 256, 0, c.stream>>>(reinterpret_cast<const dvec2 *>(src), n2, dst);
        else
            kStreamCopy<<<nb, 256, 0, c.stream>>>(reinterpret_cast<const dvec2 *>(src),
                                                 reinterpret_cast<dvec2 *>(dst), n2);
        DNDSR_HIP_CHECK(hipGetLastError());
    };
    launch(); // warm-up
    DNDSR_HIP_CHECK(hipEventRecord(a, c.stream));
    for (int r = 0; r < reps; r++)
        launch();
    DNDSR_HIP_CHECK(hipEventRecord(b, c.stream));
    DNDSR_HIP_CHECK(hipEventSynchronize(b));
    float ms = 0;
    DNDSR_HIP_CHECK(hipEventElapsedTime(&ms, a, b));
    DNDSR_HIP_CHECK(hipEventDestroy(a));
    DNDSR_HIP_CHECK(hipEventDestroy(b));
    return (double)ms / reps;
}

int streamProbeBlocks(Ctx &c)
{
    int nCU = 0;
    DNDSR_HIP_CHECK(hipDeviceGetAttribute(&nCU, hipDeviceAttributeMultiprocessorCount, c.device));
    return nCU * 8;
}

} // namespace dndsr
