// linsolve.hip -- LUSGSMatrixVec and the left-preconditioned GMRES of the implicit
// solver, on the device.
//
// LUSGSMatrixVec (EulerEvaluator.hxx:175-275): Ax_i = D_i x_i + sum_f (alpha/2) A_f / V_i
//   fluxJacobian0(u_j) x_j -- the matrix whose Gauss-Seidel sweeps UpdateSGS performs.
// GMRES_LeftPreconditioned::solve (src/Solver/Linear.hpp:47-131): restarted GMRES on
//   M A x = M b with classical Gram-Schmidt; M = (1 + sgsIter) symmetric LU-SGS sweeps
//   from zero (doPrecondition, EulerSolver.hxx:1605-1641), in FP64 or FP32.
// The Krylov basis, the Gram-Schmidt dots and updates and the normalisation stay on the
// device; the host reads the Hessenberg matrix once per Arnoldi cycle and solves the
// small least-squares problem (Householder QR).
#include <cmath>
#include <vector>

#include "gas.hpp"
#include "internal.hpp"

namespace dndsr
{
__global__ void __launch_bounds__(256) kLUSGSMatVec(DevMesh m, real gamma, real alphaDiag, const real *__restrict__ u,
                                                    const real *__restrict__ x, const real *__restrict__ jd,
                                                    const real *__restrict__ lam, real *__restrict__ Ax)
{
    const idx i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m.N)
        return;
    const size_t N = m.N, F = m.F;
    real buf[5] = {0, 0, 0, 0, 0};
    const real voli = m.vol[i];
    for (int k = 0; k < m.K; k++)
    {
        const idx f = m.c2f[(size_t)k * N + i];
        const idx o = m.c2c[(size_t)k * N + i];
        const uint8_t fl = m.c2flag[(size_t)k * N + i];
        if (f < 0 || o < 0)
            continue;
        real U[5], dU[5], n[3], dF[5];
#pragma unroll
        for (int v = 0; v < 5; v++)
            U[v] = u[v * N + o], dU[v] = x[v * N + o];
        const real sgn = (fl & 1) ? -1.0 : 1.0;
#pragma unroll
        for (int r = 0; r < 3; r++)
            n[r] = m.fmeanN[r * F + f] * sgn;
        fluxJacDu<real>(U, n, dU, lam[f], gamma, dF);
#pragma unroll
        for (int v = 0; v < 5; v++)
            buf[v] -= (0.5 * alphaDiag) * m.farea[f] / voli * dF[v];
    }
    const real d = jd[i];
#pragma unroll
    for (int v = 0; v < 5; v++)
        Ax[v * N + i] = d * x[v * N + i] - buf[v];
}

void launchLUSGSMatVec(Ctx &c, const GasParams &gp, real alphaDiag, const real *u, const real *x, const real *jd,
                       const real *lam, real *Ax)
{
    KernelTimer kt(c, K_LINSOLVE);
    kLUSGSMatVec<<<gridFor(c.m.N, 256), 256, 0, c.stream>>>(c.m, gp.gamma, alphaDiag, u, x, jd, lam, Ax);
    DNDSR_HIP_CHECK(hipGetLastError());
}

// dots[i] = <V_i, w> for i < nv, V_i = V + i * n (blockIdx.y = i); per-block partials then a reduce
constexpr int kDotBlocks = 256;
__global__ void __launch_bounds__(256) kDotBatch(size_t n, const real *__restrict__ V, const real *__restrict__ w,
                                                 real *__restrict__ part)
{
    __shared__ real sm[256];
    const real *vi = V + (size_t)blockIdx.y * n;
    real s = 0;
    for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (size_t)gridDim.x * blockDim.x)
        s += vi[k] * w[k];
    sm[threadIdx.x] = s;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1)
    {
        if (threadIdx.x < h)
            sm[threadIdx.x] += sm[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        part[(size_t)blockIdx.y * gridDim.x + blockIdx.x] = sm[0];
}
// out[i * stride] = sum of the nb partials of dot i
__global__ void __launch_bounds__(256) kDotReduce(const real *__restrict__ part, int nb, real *__restrict__ out,
                                                  int stride)
{
    __shared__ real sm[256];
    real s = 0;
    for (int k = threadIdx.x; k < nb; k += 256)
        s += part[(size_t)blockIdx.x * nb + k];
    sm[threadIdx.x] = s;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1)
    {
        if (threadIdx.x < h)
            sm[threadIdx.x] += sm[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        out[(size_t)blockIdx.x * stride] = sm[0];
}
// w -= sum_{i<nv} coef[i * stride] V_i   (coefficients read from device memory)
__global__ void __launch_bounds__(256) kSubtractProjections(size_t n, int nv, const real *__restrict__ V,
                                                            const real *__restrict__ coef, int stride,
                                                            real *__restrict__ w)
{
    const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n)
        return;
    real s = w[k];
    for (int i = 0; i < nv; i++)
        s -= coef[(size_t)i * stride] * V[(size_t)i * n + k];
    w[k] = s;
}
// w *= 1 / sqrt(h2); the norm sqrt(h2) goes to *hOut (a different location than h2)
__global__ void __launch_bounds__(256) kNormalize(size_t n, const real *__restrict__ h2, real *__restrict__ hOut,
                                                  real *__restrict__ w)
{
    const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const real hv = sqrt(*h2);
    if (k < n)
        w[k] *= 1.0 / hv;
    if (k == 0)
        *hOut = hv;
}
// out = a x + b y (y may be null)
__global__ void __launch_bounds__(256) kAxpbyPlain(size_t n, real a, const real *__restrict__ x, real b,
                                                   const real *__restrict__ y, real *__restrict__ out)
{
    const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n)
        out[k] = a * x[k] + (y ? b * y[k] : 0.0);
}
__global__ void __launch_bounds__(256) kToF64(size_t n, const float *__restrict__ in, real *__restrict__ out)
{
    const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n)
        out[k] = (real)in[k];
}

static unsigned blocksFor(size_t n) { return (unsigned)((n + 255) / 256); }

// Least squares min || H y - e1 beta || for the (j+1) x j upper Hessenberg H (column-major
// in h with leading dimension ld), by Householder QR; returns the numerical rank.
static int hessenbergLeastSquares(std::vector<double> H, int ld, int j, double beta, std::vector<double> &y)
{
    const int m = j + 1;
    std::vector<double> b(m, 0.0);
    b[0] = beta;
    auto A = [&](int r, int c) -> double & { return H[(size_t)c * ld + r]; };
    double amax = 0;
    for (int c = 0; c < j; c++)
        for (int r = 0; r < m; r++)
            amax = std::max(amax, std::fabs(A(r, c)));
    int rank = j;
    for (int c = 0; c < j; c++)
    {
        double nrm = 0;
        for (int r = c; r < m; r++)
            nrm += A(r, c) * A(r, c);
        nrm = std::sqrt(nrm);
        if (nrm <= 1e-14 * amax)
        {
            rank = c;
            break;
        }
        const double alpha = A(c, c) > 0 ? -nrm : nrm;
        std::vector<double> v(m, 0.0);
        for (int r = c; r < m; r++)
            v[r] = A(r, c);
        v[c] -= alpha;
        double vv = 0;
        for (int r = c; r < m; r++)
            vv += v[r] * v[r];
        if (vv == 0)
            continue;
        for (int cc = c; cc < j; cc++)
        {
            double s = 0;
            for (int r = c; r < m; r++)
                s += v[r] * A(r, cc);
            s *= 2 / vv;
            for (int r = c; r < m; r++)
                A(r, cc) -= s * v[r];
        }
        double s = 0;
        for (int r = c; r < m; r++)
            s += v[r] * b[r];
        s *= 2 / vv;
        for (int r = c; r < m; r++)
            b[r] -= s * v[r];
    }
    y.assign(j, 0.0);
    for (int c = rank - 1; c >= 0; c--)
    {
        double s = b[c];
        for (int cc = c + 1; cc < rank; cc++)
            s -= A(c, cc) * y[cc];
        y[c] = s / A(c, c);
    }
    return rank;
}

// Preconditioner M: (1 + sgsIter) symmetric sweeps of the LU-SGS matrix from a zero start,
// applied to the right-hand side r (doPrecondition with jacobiCode 0 / 1).
void applyLUSGSPrecond(Ctx &c, const GasParams &gp, const LinWork &w, const real *r, real *out)
{
    const size_t nU = 5 * (size_t)c.m.N;
    real *sumDev = c.red + 5 * ((size_t)gridFor(c.m.N, 256) + c.m.nColor + 6);
    if (w.fp32)
    {
        launchToF32(c, nU, r, w.fR);
        DNDSR_HIP_CHECK(hipMemsetAsync(w.fX, 0, nU * sizeof(float), c.stream));
        DNDSR_HIP_CHECK(hipMemsetAsync(w.fT, 0, nU * sizeof(float), c.stream));
        const float *x = launchSGSSymmetric32(c, gp, (float)w.alphaDiag, w.fR, w.fU, w.fX, w.fT, w.fJd, w.fLam,
                                              w.sgsIter + 1, w.gsUpdate, sumDev);
        kToF64<<<blocksFor(nU), 256, 0, c.stream>>>(nU, x, out);
        DNDSR_HIP_CHECK(hipGetLastError());
    }
    else
    {
        DNDSR_HIP_CHECK(hipMemsetAsync(out, 0, nU * sizeof(real), c.stream));
        DNDSR_HIP_CHECK(hipMemsetAsync(w.dT, 0, nU * sizeof(real), c.stream));
        const real *x = launchSGSSymmetric64(c, gp, w.alphaDiag, r, w.u, out, w.dT, w.jd, w.lam, w.sgsIter + 1,
                                             w.gsUpdate, sumDev);
        if (x != out) // the result may be left in the scratch array
            launchCopy(c, nU, x, out);
    }
}

// GMRES_LeftPreconditioned::solve with FStop = false (Linear.hpp:47-131): x is the initial
// guess on input and the solution on output.  Returns the number of restart cycles run.
int gmresSolve(Ctx &c, const GasParams &gp, const LinWork &w, const real *b, real *x, int kSub, int nRestart,
               double *residualOut)
{
    const size_t nU = 5 * (size_t)c.m.N;
    const int ld = kSub + 1;
    real *V = w.V, *Vt = w.Vt, *MLb = w.MLb, *h = w.h;
    auto dots = [&](const real *Vs, int nv, const real *vec, real *out, int stride) {
        KernelTimer kt(c, K_LINSOLVE);
        kDotBatch<<<dim3(kDotBlocks, nv), 256, 0, c.stream>>>(nU, Vs, vec, w.part);
        DNDSR_HIP_CHECK(hipGetLastError());
        kDotReduce<<<nv, 256, 0, c.stream>>>(w.part, kDotBlocks, out, stride);
        DNDSR_HIP_CHECK(hipGetLastError());
    };
    auto hostScalar = [&](const real *dev) {
        DNDSR_HIP_CHECK(hipMemcpyAsync(c.hostRed, dev, sizeof(real), hipMemcpyDeviceToHost, c.stream));
        DNDSR_HIP_CHECK(hipStreamSynchronize(c.stream));
        return c.hostRed[0];
    };
    auto FA = [&](const real *in, real *out) { launchLUSGSMatVec(c, gp, w.alphaDiag, w.u, in, w.jd, w.lam, out); };
    auto FML = [&](const real *in, real *out) { applyLUSGSPrecond(c, gp, w, in, out); };

    FML(b, MLb);
    dots(MLb, 1, MLb, w.scal, 1);
    const double scaleMLb = std::sqrt(hostScalar(w.scal));
    std::vector<double> hh((size_t)ld * kSub), y;
    int iRestart = 0;
    double beta = 0;
    for (iRestart = 0; iRestart <= nRestart; iRestart++)
    {
        FA(x, Vt);
        FML(Vt, V);
        {
            KernelTimer kt(c, K_LINSOLVE);
            kAxpbyPlain<<<blocksFor(nU), 256, 0, c.stream>>>(nU, 1.0, V, -1.0, MLb, V); // V0 = ML A x - ML b
            DNDSR_HIP_CHECK(hipGetLastError());
        }
        dots(V, 1, V, w.scal, 1);
        beta = std::sqrt(hostScalar(w.scal));
        if (std::fabs(beta) < kVerySmall || iRestart == nRestart)
            break;
        {
            KernelTimer kt(c, K_LINSOLVE);
            kAxpbyPlain<<<blocksFor(nU), 256, 0, c.stream>>>(nU, -1.0 / beta, V, 0.0, nullptr, V);
            DNDSR_HIP_CHECK(hipGetLastError());
        }
        DNDSR_HIP_CHECK(hipMemsetAsync(h, 0, (size_t)ld * kSub * sizeof(real), c.stream));
        for (int j = 0; j < kSub; j++) // Arnoldi, classical Gram-Schmidt
        {
            real *Vj = V + (size_t)j * nU, *Vn = V + (size_t)(j + 1) * nU;
            FA(Vj, Vt);
            FML(Vt, Vn);
            dots(V, j + 1, Vn, h + (size_t)j * ld, 1); // h(i, j) = <Vn, V_i>
            KernelTimer kt(c, K_LINSOLVE);
            kSubtractProjections<<<blocksFor(nU), 256, 0, c.stream>>>(nU, j + 1, V, h + (size_t)j * ld, 1, Vn);
            DNDSR_HIP_CHECK(hipGetLastError());
            dots(Vn, 1, Vn, w.scal, 1);
            kNormalize<<<blocksFor(nU), 256, 0, c.stream>>>(nU, w.scal, h + (size_t)j * ld + j + 1, Vn);
            DNDSR_HIP_CHECK(hipGetLastError());
        }
        DNDSR_HIP_CHECK(hipMemcpyAsync(hh.data(), h, hh.size() * sizeof(real), hipMemcpyDeviceToHost, c.stream));
        DNDSR_HIP_CHECK(hipStreamSynchronize(c.stream));
        // the Arnoldi process stops at the first negligible subdiagonal (Linear.hpp:89-93)
        int jEnd = kSub;
        for (int j = 0; j < kSub; j++)
            if (!(hh[(size_t)j * ld + j + 1] >= scaleMLb * 1e-32))
            {
                jEnd = j;
                break;
            }
        if (jEnd < 1)
            break;
        for (int j = 0; j < jEnd; j++)
            for (int r = 0; r <= j + 1; r++)
                if (!std::isfinite(hh[(size_t)j * ld + r]))
                    throw std::runtime_error("GMRES acquired a non-finite Hessenberg coefficient");
        const int rank = hessenbergLeastSquares(hh, ld, jEnd, beta, y);
        // x += V y (coefficients through the device scalar buffer)
        std::vector<double> yy(jEnd, 0.0);
        for (int k = 0; k < (int)y.size(); k++)
            yy[k] = -y[k];
        DNDSR_HIP_CHECK(hipMemcpyAsync(w.coef, yy.data(), jEnd * sizeof(real), hipMemcpyHostToDevice, c.stream));
        {
            KernelTimer kt(c, K_LINSOLVE);
            kSubtractProjections<<<blocksFor(nU), 256, 0, c.stream>>>(nU, jEnd, V, w.coef, 1, x);
            DNDSR_HIP_CHECK(hipGetLastError());
        }
        DNDSR_HIP_CHECK(hipStreamSynchronize(c.stream)); // the host buffer of the coefficients is reused
        if (rank < kSub) // a breakdown or a rank-deficient projection: do not restart (Linear.hpp:123-127)
            break;
    }
    if (residualOut)
        residualOut[0] = beta, residualOut[1] = scaleMLb;
    return iRestart;
}

} // namespace dndsr
