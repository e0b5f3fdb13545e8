// internal.hpp -- context, device mesh and launcher declarations shared by the
// HIP translation units of libdndsr_amd.so (not part of the C-ABI).
#pragma once
#include <hip/hip_runtime.h>

#include <stdexcept>
#include <string>
#include <vector>

#include "dev_common.hpp"

#define DNDSR_HIP_CHECK(x)                                                                          \
    do                                                                                              \
    {                                                                                               \
        hipError_t e_ = (x);                                                                        \
        if (e_ != hipSuccess)                                                                       \
            throw std::runtime_error(std::string(#x) + " failed: " + hipGetErrorString(e_));        \
    } while (0)

namespace dndsr
{
enum ElemType : int
{
    E_LINE2 = 1,
    E_QUAD4 = 2,
    E_HEX8 = 3,
    E_TRI3 = 4,
    E_TET4 = 5,
};

struct VRSettingsDev
{
    int maxOrder;
    int intOrder;
    int dirWeightScheme;  // 0 Factorial, 1 HQM_OPT
    int geomWeightScheme; // 0 GWNone, 1 HQM_SD
    real geomWeightBias;
    real geomWeightPower;
    real scaleMultiplier;
    real jacobiRelax;
    int nIterCellSmoothScale;
    real bcWeight;
    bool SORInstead; // solver loops sweep the reconstruction in place (multicolour SOR)
    int subs2ndOrder;         // maxOrder 1: 0 VR, 1 Gauss rule, 2 least squares
    int subs2ndOrderGGScheme; // 1 distance-weighted face value, 0 plain average
};

// Device-resident mesh, metrics and VR coefficients (all arrays SoA, see dev_common.hpp).
struct DevMesh
{
    int dim = 0, P = 0, NDOF = 0, M = 0;
    idx N = 0, F = 0, nNode = 0;
    int K = 0;      // max faces per cell
    int G = 0;      // face quadrature points
    int Gc = 0;     // cell quadrature points
    int cellType = 0, faceType = 0; // cellType: the type of every cell, or 0 on a mixed mesh
    int8_t *ctype = nullptr;         // [N] per-cell type on a mixed (Tri3 + Quad4) mesh
    VRSettingsDev s{};
    // connectivity
    idx *f2c = nullptr;     // [2][F], -1 = boundary
    idx *c2f = nullptr;     // [K][N], -1 = absent
    idx *c2c = nullptr;     // [K][N] cell across face slot k, -1 = absent
    int32_t *c2s = nullptr; // [K][N] slot of the same face in that neighbour cell
    uint8_t *c2flag = nullptr; // [K][N] bit 0: the cell is the face's front (R) side;
                               //        bit 1: the neighbour's colour precedes the cell's;
                               //        bits 2, 3: the neighbour is in the first, last colour
    float *fareaF = nullptr, *fmeanNF = nullptr, *volF = nullptr; // FP32 copies for the FP32 sweeps
    idx *cell2node = nullptr; // [nnc][N] (nnc = the most nodes of a cell type present)
    idx *face2node = nullptr; // [nnf][F]
    real *coords = nullptr;   // [3][nNode]
    // geometry
    real *fshift = nullptr;   // [3][F]
    real *fpts = nullptr;     // [G][3][F]
    real *fnorm = nullptr;    // [G][3][F]
    real *fwj = nullptr;      // [G][F]   w * (JDet + tiny)
    real *farea = nullptr;    // [F]
    real *fmeanN = nullptr;   // [3][F]
    real *fcent = nullptr;    // [3][F]
    real *faceL = nullptr;    // [F]
    real *faceW = nullptr;    // [P+1][F]
    real *vol = nullptr, *cent = nullptr, *bary = nullptr, *scale = nullptr, *smooth = nullptr;
    real *iscale = nullptr;   // [N] 1 / scale
    real *moment = nullptr;   // [M][N]
    real *cpts = nullptr;     // [Gc][3][N]
    real *cwj = nullptr;      // [Gc][N]
    // VR
    real *AinvB = nullptr;    // [(K+1)*M*M][N]
    real *Ainvb = nullptr;    // [K*M][N]
    // multicolour ordering of cells for the Gauss-Seidel sweeps
    int nColor = 0;
    std::vector<idx> colorPtrHost;
    idx *colorCells = nullptr; // [N] cells grouped by colour (the cell at each colour-major position)
    idx *cellPos = nullptr;    // [N] the colour-major position of each cell (inverse of colorCells)
    // the SOR sweep's cell lists: per colour (the ranges of colorPtrHost) the cells without a
    // boundary face first, then those with one, from sorSplitHost[colour] on
    idx *sorCells = nullptr;
    std::vector<idx> sorSplitHost;
    int32_t *color = nullptr;  // [N]
    real globalVol = 0;
    // boundary faces (face2cell[f][1] = -1) and their zones
    int nBCZone = 0;
    bool gasBC = false; // a zone of the boundary table is BCWallIsothermal or BCInPsTs
    idx nBnd = 0, nBndCell = 0;
    int32_t *fzone = nullptr;   // [F] zone of a boundary face, -1 elsewhere
    idx *bndFace = nullptr;     // [nBnd]
    idx *bndCell = nullptr;     // [nBndCell] cells with at least one boundary face
    int32_t *bndSlot = nullptr; // [K][nBndCell] index in bndFace of the face at slot k of bndCell[j], -1 elsewhere
    real *zoneW = nullptr;      // [nBCZone][P+1] boundary direction weights (setup only)
    // boundary-condition table (dndsr_set_bc) and VR boundary point values (dndsr_set_vr_dirichlet)
    int32_t *bcType = nullptr;  // [nBCZone]
    real *bcVal = nullptr;      // [nBCZone][5]
    real bcGamma = 1.4;
    int dirNV = 0;
    const real *dirVal = nullptr; // [dirNV][G][F]
};

struct Ctx
{
    int device = 0;
    hipStream_t stream = nullptr;
    DevMesh m;
    bool haveMesh = false;
    // scratch owned by the context (grown on demand, freed at destroy)
    real *faceFlux = nullptr;   // [5][F]
    real *vrBnd = nullptr;      // [nVars][M][nBnd] A^-1 BCC of each boundary face (VR sweep boundary term)
    size_t faceFluxCap = 0;
    real *red = nullptr;        // reduction partials
    size_t redCap = 0;
    real *hostRed = nullptr;    // pinned host mirror of reductions
    real *faceLamTmp = nullptr; // [2][F] (lambdaFace, deltaLambdaFace)
    real *adjPoint = nullptr;   // [20 or 50][G F] point sensitivities of the RHS adjoint
    int adjPointRows = 0;
    real *adjFaceBar = nullptr; // [5][F] face-flux adjoint
    real *adj2ndQ = nullptr;    // [5][3][N] per-cell weights of the transposed second-order substitute
    real *adjRhs = nullptr;     // [5][N] residual of a functional's forward evaluation
    size_t faceLamCap = 0;
    int *nonFinite = nullptr;   // set by the residual gather when a cell's residual is not finite
    std::vector<void *> owned;  // device allocations freed at destroy
    // optional per-kernel timing with HIP events on the launch stream
    bool timing = false;
    struct TimedRec
    {
        int id;
        hipEvent_t a, b;
    };
    std::vector<TimedRec> recs;
    std::vector<hipEvent_t> evPool;
    double kernMs[16] = {0};
    long long kernCount[16] = {0};
};

// kernel ids reported by dndsr_get_timing
enum KernelId : int
{
    K_VR_ITER = 0,
    K_RHS_FACE = 1,
    K_RHS_GATHER = 2,
    K_DT_FACE = 3,
    K_DT_CELL = 4,
    K_AXPBY = 5,
    K_SGS = 6,
    K_LUSGS_INIT = 7,
    K_DUAL_RES = 8,
    K_INCREMENT = 9,
    K_CONVERT = 10,
    K_ADJOINT = 11,
    K_LINSOLVE = 12,
    K_RHS_VISCOUS = 13,
    K_NUM = 14,
};

// Records a HIP event pair around the launches in its scope when timing is on.
struct KernelTimer
{
    Ctx &c;
    int id;
    hipEvent_t b = nullptr;
    KernelTimer(Ctx &cc, int kid) : c(cc), id(kid)
    {
        if (!c.timing)
            return;
        hipEvent_t a = take();
        b = take();
        DNDSR_HIP_CHECK(hipEventRecord(a, c.stream));
        c.recs.push_back({id, a, b});
    }
    ~KernelTimer()
    {
        if (b)
            (void)hipEventRecord(b, c.stream);
    }
    hipEvent_t take()
    {
        if (!c.evPool.empty())
        {
            hipEvent_t e = c.evPool.back();
            c.evPool.pop_back();
            return e;
        }
        hipEvent_t e;
        DNDSR_HIP_CHECK(hipEventCreate(&e));
        return e;
    }
};

// ---- launchers (setup.hip)
void setupMetricsAndVR(Ctx &c);
// ---- launchers (vr.hip)
void launchVRIter(Ctx &c, int nVars, const real *uRec, real *uRecNew, const real *u, int mode, real relax);
void launchVRSOR(Ctx &c, int nVars, real *uRec, const real *u, real relax);
bool usesSubs2nd(const DevMesh &m);
void launchVR2nd(Ctx &c, int nVars, const real *u, real *out);
void launchVRIterDiff(Ctx &c, int nVars, const real *uRec, const real *u, const real *uRecDiff, real *uRecNew);
// ---- launchers (euler.hip)
void launchEvalDt(Ctx &c, const GasParams &gp, const real *u, real CFL, real maxDt, int useLocalDt, real *dt,
                  real *lambdaFace, real *deltaLambdaCell, real *dtMinDev);
// Optional stage combination fused into the RHS gather: u = a u + b y + cdt dt RHS (y, dtDev
// optional; save receives u first).  The face pass reads u before the gather starts.
struct StageUpdate
{
    real a = 0, b = 0, cdt = 0;
    const real *y = nullptr, *dtDev = nullptr;
    real *u = nullptr, *save = nullptr;
};
// faceFluxBuf: the [5][F] face-flux scratch (NULL = the context's); flagNonFinite: raise the
// context's non-finite flag from the gather
void launchEvalRHS(Ctx &c, const GasParams &gp, const real *u, const real *uRec, const real *deltaLambdaCell,
                   real *rhs, const StageUpdate *stage = nullptr, real *faceFluxBuf = nullptr,
                   bool flagNonFinite = true);
void launchLUSGSInit(Ctx &c, const real *dTau, real dt, real alphaDiag, const real *lambdaFace, real *jdiag);
// a sweep of the transposed scalar LU-SGS matrix (the adjoint preconditioner), FP64
void launchSGSSweepT64(Ctx &c, const GasParams &gp, real alphaDiag, const real *rhs, const real *u, const real *uInc,
                       real *uIncNew, const real *jdiag, const real *lambdaFace, int forward, int gsUpdate,
                       int uIncIsZero, real *sumIncDev, int fCol = -1, int lCol = -1);
real *launchSGSSymmetric64(Ctx &c, const GasParams &gp, real alphaDiag, const real *rhs, const real *u, real *a,
                           real *b, const real *jd, const real *lam, int nSweep, int gsUpdate, real *sumIncDev);
float *launchSGSSymmetric32(Ctx &c, const GasParams &gp, float alphaDiag, const float *rhs, const float *u, float *a,
                            float *b, const float *jd, const float *lam, int nSweep, int gsUpdate, real *sumIncDev);
void launchSGSSweep64(Ctx &c, const GasParams &gp, real alphaDiag, const real *rhs, const real *u, const real *uInc,
                      real *uIncNew, const real *jdiag, const real *lambdaFace, int forward, int gsUpdate,
                      int uIncIsZero, real *sumIncDev);
void launchSGSSweep32(Ctx &c, const GasParams &gp, float alphaDiag, const float *rhs, const float *u,
                      const float *uInc, float *uIncNew, const float *jdiag, const float *lambdaFace, int forward,
                      int gsUpdate, int uIncIsZero, real *sumIncDev);
// out = a x + b y + cdt dt z (y, z, dtDev optional); save, when given, receives x first
// Roe-Jacobian block LU-SGS (useRoeJacobian): face eigenvalues lamRoe [3][F], block JDiag and its
// inverse [25][N] (bad set when a block cannot be inverted), block sweeps
void launchFaceRoeLambda(Ctx &c, const GasParams &gp, const real *u, const real *uRec, const real *dlc, real *lamRoe);
void launchLUSGSInitBlock(Ctx &c, const GasParams &gp, const real *dTau, real dt, real alphaDiag, const real *u,
                          const real *lam, const real *lamRoe, real *jd, real *jinv, int *bad);
void launchSGSBlockSymmetric64(Ctx &c, const GasParams &gp, real alphaDiag, const real *rhs, const real *u, real *a,
                               real *b, const real *jinv, const real *lam, const real *lamRoe, int sgsIter,
                               int gsUpdate, real *sumIncDev, real **result);
void launchSGSBlockSymmetric32(Ctx &c, const GasParams &gp, float alphaDiag, const float *rhs, const float *u,
                               float *a, float *b, const float *jinv, const float *lam, const float *lamRoe,
                               int sgsIter, int gsUpdate, real *sumIncDev, float **result);
void launchSGSSweepBlock64(Ctx &c, const GasParams &gp, real alphaDiag, const real *rhs, const real *u,
                           const real *uInc, real *uIncNew, const real *jinv, const real *lam, const real *lamRoe,
                           int forward, int gsUpdate, int uIncIsZero, real *sumIncDev);
void launchSGSSweepBlockT64(Ctx &c, const GasParams &gp, real alphaDiag, const real *rhs, const real *u,
                            const real *uInc, real *uIncNew, const real *jinv, const real *lam, const real *lamRoe,
                            int forward, int gsUpdate, int uIncIsZero, real *sumIncDev, int fCol = -1,
                            int lCol = -1);
void launchCopyColourInc64(Ctx &c, int col, const real *from, real *to);
void launchSGSSweepBlock32(Ctx &c, const GasParams &gp, float alphaDiag, const float *rhs, const float *u,
                           const float *uInc, float *uIncNew, const float *jinv, const float *lam,
                           const float *lamRoe, int forward, int gsUpdate, int uIncIsZero, real *sumIncDev);
// device bandwidth self-test (probe.hip): mode 0 read stream, 1 copy; average ms per pass
double streamProbe(Ctx &c, int mode, const double *src, double *dst, size_t n, int reps);
int streamProbeBlocks(Ctx &c);
// sum of nPart 5-wide partials in c.red into out[5]
void launchReduceSum5(Ctx &c, int nPart, real *out);
void launchAxpby(Ctx &c, idx n, real a, const real *x, real b, const real *y, real cdt, const real *z,
                 const real *dtDev, real *out, real *save = nullptr);
// dst[0:n) = src[0:n) on the library's stream
void launchCopy(Ctx &c, size_t n, const real *src, real *dst);
// the BDF unsteady residual: sum_j coef_j / dt xs_j - x / dt + alpha rhsF, nx in 1..4 (xs[0] = xLast)
void launchDualResidualBDF(Ctx &c, real dt, const real *const *xs, const real *coef, int nx, real alpha,
                           const real *x, const real *rhsF, real *rhs, real *resDev);
void launchScale(Ctx &c, idx n, real s, real *x);
void launchDualResidual(Ctx &c, real dt, const real *xLast, const real *x, const real *rhsF, real *rhs,
                        real *resDev);
// statistics of the positivity-preserving increment: limited cells, smallest alpha (as the bits
// of a positive double), and a flag for states the reference asserts against
struct PPStat
{
    unsigned long long alphaMinBits;
    int nLim;
    int bad;
};
void launchPPStatReset(Ctx &c, PPStat *stat);
// fincrement: u += CompressInc(u, alpha du) with alpha from EvaluateCellRHSAlpha, per cell
void launchAddIncrement64(Ctx &c, real gamma, real relax, real rhoEps, real pEps, real *u, const real *du,
                          PPStat *stat);
void launchAddIncrement32(Ctx &c, real gamma, real relax, real rhoEps, real pEps, real *u, const float *du,
                          PPStat *stat);
void launchToF32(Ctx &c, size_t n, const real *in, float *out);
// (1 + sgsIter) forward/backward multicolour sweeps from a zero increment with the state
// gathered from neighbours stored cell-major: uA <- u as [N][5], slot <- per-cell-slot face
// data; inc / tmp are [N][5] scratch; *result points at the buffer holding the increment.
// xbuf: [2][N][16] (two packed arrays of state + increment rows); slot: [(K*5 + 7) N] (per-slot
// face data, then rhs, JDiag and volume in colour-major position order); nbr: [K N]; fresh: slot
// and nbr are new (their mesh-only part is written); *result: the increment, rows of 16 in
// colour-major position order
void launchSGSPacked64(Ctx &c, const GasParams &gp, real alphaDiag, const real *rhs, const real *u, const real *lam,
                       const real *jd, real *xbuf, real *slot, idx *nbr, bool fresh, int sgsIter, int gsUpdate,
                       real **result);
void launchSGSPacked32(Ctx &c, const GasParams &gp, float alphaDiag, const float *rhs, const real *u,
                       const float *lam, const float *jd, float *xbuf, float *slot, idx *nbr, bool fresh, int sgsIter,
                       int gsUpdate, float **result);
// u += du with du cell-major ([N][5], FP64 or FP32: exactly one of du64 / du32 non-null)
// du: the packed sweep's increment (rows of 16 in colour-major position order)
void launchAddIncrementCellMajor(Ctx &c, real gamma, real relax, real rhoEps, real pEps, real *u, const real *du64,
                                 const float *du32, PPStat *stat);
// ---- launchers (model.hip): ModelEvaluator::EvaluateRHS, scalar advection-diffusion (2-D)
void launchModelRHS(Ctx &c, int nVars, real ax, real ay, real sigma, const real *u, const real *uRec, real *rhs);
// ---- linear solver (linsolve.hip)
struct LinWork
{
    // operator: LU-SGS matrix at the state u (FP64) and its FP32 copies for FP32 sweeps
    const real *u = nullptr, *jd = nullptr, *lam = nullptr;
    const float *fU = nullptr, *fJd = nullptr, *fLam = nullptr;
    real alphaDiag = 1.0;
    int sgsIter = 0, gsUpdate = 1, fp32 = 0;
    // scratch: FP32 sweep vectors, FP64 sweep vector, Krylov basis [(k+1)][5N], temporaries
    float *fR = nullptr, *fX = nullptr, *fT = nullptr;
    real *dT = nullptr, *V = nullptr, *Vt = nullptr, *MLb = nullptr, *h = nullptr, *part = nullptr,
         *scal = nullptr, *coef = nullptr;
};
void launchLUSGSMatVec(Ctx &c, const GasParams &gp, real alphaDiag, const real *u, const real *x, const real *jd,
                       const real *lam, real *Ax);
void applyLUSGSPrecond(Ctx &c, const GasParams &gp, const LinWork &w, const real *r, real *out);
int gmresSolve(Ctx &c, const GasParams &gp, const LinWork &w, const real *b, real *x, int kSub, int nRestart,
               double *residualOut);
// ---- launchers (adjoint.hip)
// faceBar (optional, [5][F]): the adjoint of the face fluxes instead of the residual's
// (rbar unused then); without it the face adjoint is formed from rbar
void launchEvalRHSVJP(Ctx &c, const GasParams &gp, const real *u, const real *uRec, const real *dlc,
                      const real *rbar, real *ubar, real *urecbar, const real *faceBar = nullptr);
// force on the boundary faces of a zone from the last EvaluateRHS's face fluxes:
// forceDev[d] = - sum_f ff[1 + d][f] (the momentum flux out of the domain)
void launchZoneForce(Ctx &c, int zone, real *forceDev);
// face-flux adjoint of that force: fbar[1 + d][f] = -forceBar[d] on the zone's faces, 0 elsewhere
void launchZoneForceBar(Ctx &c, int zone, const real *forceBarDev, real *fbar);
void launchVRIterVJP(Ctx &c, int nVars, int mode, real relax, const real *uRec, const real *u, const real *ybar,
                     real *xbar, real *ubar);

template <typename T>
T *devAlloc(Ctx &c, size_t n)
{
    void *p = nullptr;
    if (n == 0)
        n = 1;
    DNDSR_HIP_CHECK(hipMalloc(&p, n * sizeof(T)));
    c.owned.push_back(p);
    return static_cast<T *>(p);
}

inline int gridFor(idx n, int block) { return (int)((n + block - 1) / block); }

// Skipped colour passes of the symmetric Gauss-Seidel sweeps: a sweep may skip its first colour
// (fCol; its current values stay in the array the sweep reads) and read the colour the previous
// sweep skipped (lCol) from the array it writes.  Both are the first or the last colour, which
// the slot flags mark (bits 2, 3); -1 names none.
__device__ __forceinline__ bool nbrInColour(uint8_t fl, int col, int nColor)
{
    return (col == 0 && (fl & 4)) || (col >= 0 && col == nColor - 1 && (fl & 8));
}
__device__ __forceinline__ bool sweepReadNew(uint8_t fl, int fCol, int lCol, int nColor, bool useNew)
{
    if (nbrInColour(fl, fCol, nColor))
        return false;
    if (nbrInColour(fl, lCol, nColor))
        return true;
    return useNew;
}

} // namespace dndsr
