// capi.hip -- C-ABI of libdndsr_amd.so (declarations and reference citations in
// include/dndsr_amd.h).  Host-side responsibilities: mesh upload into the device
// SoA layout, multicolour ordering of the cells, buffer management and the
// device-resident SSP-RK3 loop; every numerical operation runs in the HIP kernels
// of setup.hip, vr.hip and euler.hip.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/dndsr_amd.h"
#include "gas.hpp"
#include "geom.hpp"

using namespace dndsr;

// Scratch of dndsr_eval_rhs_noalloc: the face-flux buffer [5][F] of EvaluateRHS for meshes of up
// to kNoAllocMaxFaces faces (40 MiB), a file-scope device array that every GPU the library runs on
// holds until the process ends.  In host code `scratch` only names the symbol; its device address
// (one per GPU) comes from hipGetSymbolAddress at dndsr_create and is the only form passed on.
static constexpr size_t kNoAllocMaxFaces = size_t(1) << 20;
static __device__ __align__(256) double scratch[5 * kNoAllocMaxFaces];
static constexpr int kMaxDevices = 64;
static double *g_scratchDev[kMaxDevices] = {nullptr};

struct dndsr_ctx
{
    Ctx c;
    double *scratchDev = nullptr; // this device's address of `scratch`
    hipStream_t own = nullptr;
    // device-resident time-loop workspace
    real *wXLast = nullptr, *wURecNew = nullptr, *wLam = nullptr, *wDlc = nullptr, *wDtMin = nullptr;
    bool haveWork = false;
    // implicit dual-time workspace
    real *iXLast = nullptr, *iRhsF = nullptr, *iRhs = nullptr, *iDTau = nullptr, *iJd = nullptr, *iInc = nullptr,
         *iTmp = nullptr, *iRes = nullptr;
    float *fRhs = nullptr, *fU = nullptr, *fJd = nullptr, *fLam = nullptr, *fInc = nullptr, *fTmp = nullptr;
    // cell-major sweep state and per-cell-slot face data of the packed sweeps
    real *iSlot = nullptr, *iUA = nullptr; // packed sweep: slot data; two [N][16] state + increment arrays
    float *fSlot = nullptr, *fUA = nullptr;
    idx *iNbr = nullptr, *fNbr = nullptr; // packed sweep: the neighbours' colour-major positions, [K][N]
    PPStat *iStat = nullptr;   // positivity statistics of the current iteration
    real *iURecSave = nullptr; // uRec at entry, restored when a step fails
    int64_t ppLimited = 0;     // limited cell increments of the last implicit step
    double ppAlphaMin = 1.0;   // and their smallest alpha
    bool haveImplicit = false;
    // Roe-Jacobian block LU-SGS workspace: face eigenvalues [3][F], inverse blocks [25][N]
    // (FP64 and FP32), the singular-block flag
    real *bLamRoe = nullptr, *bJinv = nullptr;
    float *bJinvF = nullptr, *bLamRoeF = nullptr;
    int *bBad = nullptr;
    // linear-solver workspace (GMRES Krylov basis for up to linK vectors + temporaries)
    LinWork lin;
    int linK = -1;
    bool linF32 = false;
};

// workspace of the device time loops (SSP-RK3 and the implicit step share it)
static void timeLoopWork(dndsr_ctx *ctx)
{
    if (ctx->haveWork)
        return;
    Ctx &c = ctx->c;
    const size_t N = c.m.N;
    ctx->wXLast = devAlloc<real>(c, 5 * N);
    ctx->wURecNew = devAlloc<real>(c, 5 * (size_t)c.m.M * N);
    ctx->wLam = devAlloc<real>(c, 4 * (size_t)c.m.F);
    ctx->wDlc = devAlloc<real>(c, N);
    ctx->wDtMin = devAlloc<real>(c, 1);
    ctx->haveWork = true;
}

// workspace for GMRES(k) and the LU-SGS preconditioner; FP32 buffers only when needed
static LinWork &linWork(dndsr_ctx *ctx, int k, bool fp32)
{
    Ctx &c = ctx->c;
    const size_t N = c.m.N, nU = 5 * N;
    LinWork &w = ctx->lin;
    if (ctx->linK < 0)
    {
        w.dT = devAlloc<real>(c, nU);
        w.Vt = devAlloc<real>(c, nU);
        w.MLb = devAlloc<real>(c, nU);
        w.part = devAlloc<real>(c, (size_t)256 * 64);
        w.scal = devAlloc<real>(c, 8);
        w.coef = devAlloc<real>(c, 64);
        ctx->linK = 0;
    }
    if (k > ctx->linK)
    {
        if (k > 60)
            throw std::runtime_error("GMRES subspace size must be at most 60");
        w.V = devAlloc<real>(c, (size_t)(k + 1) * nU);
        w.h = devAlloc<real>(c, (size_t)(k + 1) * k);
        ctx->linK = k;
    }
    if (fp32 && !ctx->linF32)
    {
        w.fR = devAlloc<float>(c, nU);
        w.fX = devAlloc<float>(c, nU);
        w.fT = devAlloc<float>(c, nU);
        ctx->linF32 = true;
    }
    return w;
}

static thread_local std::string g_err;

#define CAPI_BEGIN try {
#define CAPI_END                                      \
    return 0;                                         \
    }                                                 \
    catch (const std::exception &e)                   \
    {                                                 \
        g_err = e.what();                             \
        return 1;                                     \
    }                                                 \
    catch (...)                                       \
    {                                                 \
        g_err = "unknown error";                      \
        return 1;                                     \
    }

static void need(bool ok, const char *msg)
{
    if (!ok)
        throw std::runtime_error(msg);
}

static void requireMesh(dndsr_ctx *ctx)
{
    need(ctx != nullptr, "null context");
    need(ctx->c.haveMesh, "no mesh: call dndsr_vr_construct first");
}

template <typename T>
static void upload(Ctx &c, T *dev, const std::vector<T> &h)
{
    DNDSR_HIP_CHECK(hipMemcpyAsync(dev, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice, c.stream));
}
template <typename T>
static std::vector<T> download(Ctx &c, const T *dev, size_t n)
{
    std::vector<T> h(n);
    DNDSR_HIP_CHECK(hipMemcpyAsync(h.data(), dev, sizeof(T) * n, hipMemcpyDeviceToHost, c.stream));
    DNDSR_HIP_CHECK(hipStreamSynchronize(c.stream));
    return h;
}

static void freeAll(Ctx &c)
{
    for (void *p : c.owned)
        (void)hipFree(p);
    c.owned.clear();
    if (c.hostRed)
        (void)hipHostFree(c.hostRed);
    c.hostRed = nullptr;
}

extern "C" {

const char *dndsr_last_error(void) { return g_err.c_str(); }

int dndsr_create(dndsr_ctx **out, int device)
{
    CAPI_BEGIN
    need(out != nullptr, "null output");
    int n = 0;
    DNDSR_HIP_CHECK(hipGetDeviceCount(&n));
    need(device >= 0 && device < n, "invalid device index");
    DNDSR_HIP_CHECK(hipSetDevice(device));
    need(device < kMaxDevices, "device index beyond the library's table");
    if (!g_scratchDev[device])
    {
        void *p = nullptr;
        DNDSR_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        need(p != nullptr, "no device address for the no-allocation scratch");
        g_scratchDev[device] = static_cast<double *>(p);
    }
    auto *ctx = new dndsr_ctx;
    ctx->scratchDev = g_scratchDev[device];
    ctx->c.device = device;
    DNDSR_HIP_CHECK(hipStreamCreateWithFlags(&ctx->own, hipStreamNonBlocking));
    ctx->c.stream = ctx->own;
    DNDSR_HIP_CHECK(hipHostMalloc(reinterpret_cast<void **>(&ctx->c.hostRed), sizeof(real) * 16));
    ctx->c.nonFinite = devAlloc<int>(ctx->c, 1);
    DNDSR_HIP_CHECK(hipMemsetAsync(ctx->c.nonFinite, 0, sizeof(int), ctx->c.stream));
    DNDSR_HIP_CHECK(hipStreamSynchronize(ctx->c.stream));
    *out = ctx;
    CAPI_END
}

int dndsr_destroy(dndsr_ctx *ctx)
{
    CAPI_BEGIN
    if (!ctx)
        return 0;
    (void)hipSetDevice(ctx->c.device);
    (void)hipStreamSynchronize(ctx->c.stream);
    for (auto &r : ctx->c.recs)
        (void)hipEventDestroy(r.a), (void)hipEventDestroy(r.b);
    for (auto e : ctx->c.evPool)
        (void)hipEventDestroy(e);
    freeAll(ctx->c);
    if (ctx->own)
        (void)hipStreamDestroy(ctx->own);
    delete ctx;
    CAPI_END
}

// Reads and clears the residual's non-finite flag (the stream must be synchronised).
static void checkFinite(Ctx &c, const char *where)
{
    int nf = 0;
    DNDSR_HIP_CHECK(hipMemcpy(&nf, c.nonFinite, sizeof(int), hipMemcpyDeviceToHost));
    if (nf)
    {
        DNDSR_HIP_CHECK(hipMemset(c.nonFinite, 0, sizeof(int)));
        throw std::runtime_error(std::string("non-finite residual (EvaluateRHS) detected ") + where);
    }
}

int dndsr_sync(dndsr_ctx *ctx)
{
    CAPI_BEGIN
    need(ctx != nullptr, "null context");
    DNDSR_HIP_CHECK(hipStreamSynchronize(ctx->c.stream));
    checkFinite(ctx->c, "since the last synchronisation");
    CAPI_END
}

int dndsr_set_stream(dndsr_ctx *ctx, void *stream)
{
    CAPI_BEGIN
    need(ctx != nullptr, "null context");
    ctx->c.stream = reinterpret_cast<hipStream_t>(stream); // NULL: the device's default (null) stream
    CAPI_END
}

int dndsr_vr_construct(dndsr_ctx *ctx, const dndsr_mesh_desc *md, const dndsr_vr_settings *s)
{
    CAPI_BEGIN
    need(ctx && md && s, "null argument");
    need(!ctx->c.haveMesh, "context already holds a mesh");
    need(md->dim == 2 || md->dim == 3, "dim must be 2 or 3");
    need(s->maxOrder >= 1 && s->maxOrder <= 3, "maxOrder must be 1..3");
    need(s->intOrder >= 0 && s->intOrder <= 6, "intOrder must be 0..6");
    need(md->nCell > 1 && md->nCell < (int64_t(1) << 31) && md->nFace < (int64_t(1) << 31), "mesh size out of range");
    Ctx &c = ctx->c;
    DNDSR_HIP_CHECK(hipSetDevice(c.device));
    DevMesh &m = c.m;
    m.dim = md->dim;
    m.P = s->maxOrder;
    m.NDOF = nDof(m.dim, m.P);
    m.M = m.NDOF - 1;
    m.N = (idx)md->nCell;
    m.F = (idx)md->nFace;
    m.nNode = (idx)md->nNode;
    m.s = VRSettingsDev{s->maxOrder, s->intOrder, s->dirWeightScheme, s->geomWeightScheme, s->geomWeightBias,
                        s->geomWeightPower, s->scaleMultiplier, s->jacobiRelax, s->nIterCellSmoothScale,
                        s->bcWeight, s->SORInstead != 0, s->subs2ndOrder, s->subs2ndOrderGGScheme};
    need(s->subs2ndOrder == 0 || s->subs2ndOrder == 1 || s->subs2ndOrder == 2,
         "subs2ndOrder must be 0, 1 (Gauss rule) or 2 (least squares); 11 (GGMP) is not provided");
    need(s->subs2ndOrderGGScheme == 0 || s->subs2ndOrderGGScheme == 1, "subs2ndOrderGGScheme must be 0 or 1");
    need(md->nBCZone >= 0 && md->nBCZone < 4096, "nBCZone out of range");
    m.nBCZone = md->nBCZone;
    const size_t N = m.N, F = m.F;
    // element types: one face type per mesh; cells of one type, or (2-D) Tri3 and Quad4 mixed
    m.cellType = md->cellType[0];
    m.faceType = md->faceType[0];
    bool mixed = false;
    for (size_t i = 0; i < N; i++)
        mixed |= md->cellType[i] != m.cellType;
    for (size_t f = 0; f < F; f++)
        need(md->faceType[f] == m.faceType, "mixed face types are not supported");
    if (m.dim == 2)
    {
        for (size_t i = 0; i < N; i++)
            need(md->cellType[i] == DNDSR_QUAD4 || md->cellType[i] == DNDSR_TRI3,
                 "2-D meshes must be Quad4 and/or Tri3 cells / Line2 faces");
        need(m.faceType == DNDSR_LINE2, "2-D meshes must have Line2 faces");
    }
    else
    {
        need(!mixed, "3-D meshes take one cell type");
        need((m.cellType == DNDSR_HEX8 && m.faceType == DNDSR_QUAD4) ||
                 (m.cellType == DNDSR_TET4 && m.faceType == DNDSR_TRI3),
             "3-D meshes must be Hex8 cells / Quad4 faces or Tet4 cells / Tri3 faces");
    }
    const int nnc = mixed ? 4 : nodeCount(m.cellType), nnf = nodeCount(m.faceType);
    std::vector<int8_t> hct;
    if (mixed)
    {
        hct.resize(N);
        for (size_t i = 0; i < N; i++)
            hct[i] = (int8_t)md->cellType[i];
        m.cellType = 0;
    }
    int K = 0;
    for (size_t i = 0; i < N; i++)
    {
        K = std::max<int>(K, (int)(md->cell2facePtr[i + 1] - md->cell2facePtr[i]));
    }
    need(K <= kMaxFC, "too many faces per cell");
    m.K = K;
    need(s->intOrder >= 0 && s->intOrder <= 6, "intOrder must be 0..6");
    m.G = ruleSize(m.faceType, s->intOrder);
    m.Gc = mixed ? std::max(ruleSize(E_QUAD4, s->intOrder), ruleSize(E_TRI3, s->intOrder))
                 : ruleSize(m.cellType, s->intOrder);
    need(m.G <= kMaxG, "too many face quadrature points");

    std::vector<real> hc(3 * (size_t)m.nNode);
    for (size_t a = 0; a < (size_t)m.nNode; a++)
        for (int r = 0; r < 3; r++)
            hc[r * m.nNode + a] = md->coords[3 * a + r];
    std::vector<idx> hcn((size_t)nnc * N, 0), hfn((size_t)nnf * F), hf2c(2 * F), hc2f((size_t)K * N, -1);
    for (size_t i = 0; i < N; i++)
    {
        const int nn = nodeCount(md->cellType[i]);
        need(md->cell2nodePtr[i + 1] - md->cell2nodePtr[i] == nn, "bad cell node count");
        for (int a = 0; a < nn; a++)
        {
            int64_t nd = md->cell2node[md->cell2nodePtr[i] + a];
            need(nd >= 0 && nd < md->nNode, "cell2node out of range");
            hcn[a * N + i] = (idx)nd;
        }
    }
    std::vector<real> hsh(3 * F);
    std::vector<int32_t> hzone(F, -1);
    std::vector<idx> hbnd;
    for (size_t f = 0; f < F; f++)
    {
        need(md->face2nodePtr[f + 1] - md->face2nodePtr[f] == nnf, "bad face node count");
        for (int a = 0; a < nnf; a++)
        {
            int64_t nd = md->face2node[md->face2nodePtr[f] + a];
            need(nd >= 0 && nd < md->nNode, "face2node out of range");
            hfn[a * F + f] = (idx)nd;
        }
        int64_t L = md->face2cell[2 * f], R = md->face2cell[2 * f + 1];
        need(L >= 0 && L < (int64_t)N && R < (int64_t)N, "face2cell out of range");
        if (R < 0)
        {
            need(R == -1 && md->faceZone && md->faceZone[f] >= 0 && md->faceZone[f] < md->nBCZone,
                 "a boundary face needs a zone in 0 .. nBCZone-1");
            hzone[f] = md->faceZone[f];
            hbnd.push_back((idx)f);
        }
        hf2c[f] = (idx)L, hf2c[F + f] = (idx)R;
        for (int r = 0; r < 3; r++)
            hsh[r * F + f] = md->faceShift[3 * f + r];
    }
    for (size_t i = 0; i < N; i++)
    {
        int k = 0;
        for (int64_t q = md->cell2facePtr[i]; q < md->cell2facePtr[i + 1]; q++, k++)
        {
            int64_t f = md->cell2face[q];
            need(f >= 0 && f < (int64_t)F, "cell2face out of range");
            need(hf2c[f] == (idx)i || hf2c[F + f] == (idx)i, "cell2face / face2cell mismatch");
            hc2f[(size_t)k * N + i] = (idx)f;
        }
    }
    // neighbour across each face slot (the VR sweep's only index) and the slot of the
    // same face in that neighbour (the transposed sweep gathers through it)
    std::vector<idx> hc2c((size_t)K * N, -1);
    std::vector<int32_t> hc2s((size_t)K * N, -1);
    for (size_t i = 0; i < N; i++)
        for (int k = 0; k < K; k++)
        {
            idx f = hc2f[(size_t)k * N + i];
            if (f < 0)
                continue;
            idx o = hf2c[f] == (idx)i ? hf2c[F + f] : hf2c[f];
            hc2c[(size_t)k * N + i] = o;
            if (o < 0) // boundary face: no neighbour, no reverse slot
                continue;
            for (int k2 = 0; k2 < K; k2++)
                if (hc2f[(size_t)k2 * N + o] == f)
                    hc2s[(size_t)k * N + i] = k2;
            need(hc2s[(size_t)k * N + i] >= 0, "face missing from its neighbour's cell2face");
        }
    // greedy multicolouring in cell order (smallest colour unused by face neighbours)
    std::vector<int32_t> col(N, -1);
    int nColor = 0;
    for (size_t i = 0; i < N; i++)
    {
        uint64_t used = 0;
        for (int k = 0; k < K; k++)
        {
            idx f = hc2f[(size_t)k * N + i];
            if (f < 0)
                continue;
            idx o = hf2c[f] == (idx)i ? hf2c[F + f] : hf2c[f];
            if (o >= 0 && o != (idx)i && col[o] >= 0 && col[o] < 64)
                used |= (uint64_t(1) << col[o]);
        }
        int cc = 0;
        while (cc < 64 && (used >> cc) & 1)
            cc++;
        need(cc < 64, "colouring needs more than 64 colours");
        col[i] = cc;
        nColor = std::max(nColor, cc + 1);
    }
    m.nColor = nColor;
    m.colorPtrHost.assign(nColor + 1, 0);
    for (size_t i = 0; i < N; i++)
        m.colorPtrHost[col[i] + 1]++;
    for (int q = 0; q < nColor; q++)
        m.colorPtrHost[q + 1] += m.colorPtrHost[q];
    std::vector<idx> cells(N), fill(m.colorPtrHost.begin(), m.colorPtrHost.end() - 1);
    for (size_t i = 0; i < N; i++)
        cells[fill[col[i]]++] = (idx)i;

    // device allocations
    m.coords = devAlloc<real>(c, hc.size());
    m.cell2node = devAlloc<idx>(c, hcn.size());
    m.face2node = devAlloc<idx>(c, hfn.size());
    m.f2c = devAlloc<idx>(c, hf2c.size());
    m.c2f = devAlloc<idx>(c, hc2f.size());
    m.c2c = devAlloc<idx>(c, hc2c.size());
    m.c2s = devAlloc<int32_t>(c, hc2s.size());
    m.fshift = devAlloc<real>(c, hsh.size());
    m.color = devAlloc<int32_t>(c, N);
    m.colorCells = devAlloc<idx>(c, N);
    m.cellPos = devAlloc<idx>(c, N);
    m.sorCells = devAlloc<idx>(c, N);
    {
        std::vector<uint8_t> hfl((size_t)K * N, 0);
        for (size_t i = 0; i < N; i++)
            for (int k = 0; k < K; k++)
            {
                idx f = hc2f[(size_t)k * N + i];
                if (f < 0)
                    continue;
                idx o = hc2c[(size_t)k * N + i];
                uint8_t fl = hf2c[f] == (idx)i ? 0 : 1;
                if (o >= 0 && col[o] < col[i])
                    fl |= 2;
                if (o >= 0 && col[o] == 0)
                    fl |= 4;
                if (o >= 0 && col[o] == nColor - 1)
                    fl |= 8;
                hfl[(size_t)k * N + i] = fl;
            }
        m.c2flag = devAlloc<uint8_t>(c, hfl.size());
        upload(c, m.c2flag, hfl);
    }
    upload(c, m.coords, hc);
    upload(c, m.cell2node, hcn);
    if (mixed)
    {
        m.ctype = devAlloc<int8_t>(c, N);
        upload(c, m.ctype, hct);
    }
    upload(c, m.face2node, hfn);
    upload(c, m.f2c, hf2c);
    upload(c, m.c2f, hc2f);
    upload(c, m.c2c, hc2c);
    upload(c, m.c2s, hc2s);
    upload(c, m.fshift, hsh);
    upload(c, m.color, col);
    upload(c, m.colorCells, cells);
    {
        std::vector<idx> pos(N);
        for (size_t q = 0; q < N; q++)
            pos[cells[q]] = (idx)q;
        upload(c, m.cellPos, pos);
        // per colour: cells without a boundary face, then cells with one
        std::vector<char> hasBnd(N, 0);
        for (size_t i = 0; i < N; i++)
            for (int k = 0; k < K; k++)
                if (hc2f[(size_t)k * N + i] >= 0 && hc2c[(size_t)k * N + i] < 0)
                    hasBnd[i] = 1;
        std::vector<idx> sc(N);
        m.sorSplitHost.assign(nColor, 0);
        for (int q = 0; q < nColor; q++)
        {
            size_t w = m.colorPtrHost[q];
            for (idx t = m.colorPtrHost[q]; t < m.colorPtrHost[q + 1]; t++)
                if (!hasBnd[cells[t]])
                    sc[w++] = cells[t];
            m.sorSplitHost[q] = (idx)w;
            for (idx t = m.colorPtrHost[q]; t < m.colorPtrHost[q + 1]; t++)
                if (hasBnd[cells[t]])
                    sc[w++] = cells[t];
        }
        upload(c, m.sorCells, sc);
    }
    // boundary faces, the cells that own them, per-zone VR direction weights
    m.nBnd = (idx)hbnd.size();
    m.fzone = devAlloc<int32_t>(c, F);
    upload(c, m.fzone, hzone);
    if (m.nBnd > 0)
    {
        m.bndFace = devAlloc<idx>(c, hbnd.size());
        upload(c, m.bndFace, hbnd);
        std::vector<idx> bcells;
        for (idx f : hbnd)
            bcells.push_back(hf2c[f]);
        std::sort(bcells.begin(), bcells.end());
        bcells.erase(std::unique(bcells.begin(), bcells.end()), bcells.end());
        m.nBndCell = (idx)bcells.size();
        m.bndCell = devAlloc<idx>(c, bcells.size());
        upload(c, m.bndCell, bcells);
        // per boundary cell and face slot: the slot's face in the boundary-face list, -1 elsewhere
        std::vector<int32_t> tOf(F, -1);
        for (size_t t = 0; t < hbnd.size(); t++)
            tOf[hbnd[t]] = (int32_t)t;
        const size_t nbc = bcells.size();
        std::vector<int32_t> bs((size_t)K * nbc, -1);
        for (size_t j = 0; j < nbc; j++)
            for (int k = 0; k < K; k++)
            {
                const idx f = hc2f[(size_t)k * N + bcells[j]];
                if (f >= 0 && hc2c[(size_t)k * N + bcells[j]] < 0)
                {
                    need(tOf[f] >= 0, "a face without a neighbour cell is missing from the boundary faces");
                    bs[(size_t)k * nbc + j] = tOf[f];
                }
            }
        m.bndSlot = devAlloc<int32_t>(c, bs.size());
        upload(c, m.bndSlot, bs);
    }
    if (m.nBCZone > 0)
    {
        std::vector<real> zw((size_t)m.nBCZone * (m.P + 1), 1.0);
        if (s->zoneDirWeight)
            std::copy(s->zoneDirWeight, s->zoneDirWeight + zw.size(), zw.begin());
        m.zoneW = devAlloc<real>(c, zw.size());
        upload(c, m.zoneW, zw);
    }
    m.fpts = devAlloc<real>(c, (size_t)m.G * 3 * F);
    m.fnorm = devAlloc<real>(c, (size_t)m.G * 3 * F);
    m.fwj = devAlloc<real>(c, (size_t)m.G * F);
    m.farea = devAlloc<real>(c, F);
    m.fmeanN = devAlloc<real>(c, 3 * F);
    m.fcent = devAlloc<real>(c, 3 * F);
    m.faceL = devAlloc<real>(c, F);
    m.faceW = devAlloc<real>(c, (size_t)(m.P + 1) * F);
    m.vol = devAlloc<real>(c, N);
    m.cent = devAlloc<real>(c, 3 * N);
    m.bary = devAlloc<real>(c, 3 * N);
    m.scale = devAlloc<real>(c, N);
    m.iscale = devAlloc<real>(c, N);
    m.smooth = devAlloc<real>(c, N);
    m.moment = devAlloc<real>(c, (size_t)m.M * N);
    m.cpts = devAlloc<real>(c, (size_t)m.Gc * 3 * N);
    m.cwj = devAlloc<real>(c, (size_t)m.Gc * N);
    m.AinvB = devAlloc<real>(c, (size_t)(K + 1) * m.M * m.M * N);
    m.Ainvb = devAlloc<real>(c, (size_t)K * m.M * N);
    c.faceFlux = devAlloc<real>(c, 5 * F);
    if (m.nBnd > 0)
        c.vrBnd = devAlloc<real>(c, (size_t)5 * m.M * m.nBnd);
    c.faceLamTmp = devAlloc<real>(c, F);
    size_t nPart = (size_t)gridFor(m.N, 256) + (size_t)nColor + 8;
    c.red = devAlloc<real>(c, nPart * 5);
    DNDSR_HIP_CHECK(hipStreamSynchronize(c.stream));

    setupMetricsAndVR(c);
    // FP32 copies of the face and cell geometry read by the FP32 Gauss-Seidel sweeps
    m.fareaF = devAlloc<float>(c, F);
    m.fmeanNF = devAlloc<float>(c, 3 * F);
    m.volF = devAlloc<float>(c, N);
    launchToF32(c, F, m.farea, m.fareaF);
    launchToF32(c, 3 * F, m.fmeanN, m.fmeanNF);
    launchToF32(c, N, m.vol, m.volF);
    auto vol = download(c, m.vol, N);
    m.globalVol = 0;
    for (real v : vol)
        m.globalVol += v;
    c.haveMesh = true;
    CAPI_END
}

int dndsr_mesh_info(dndsr_ctx *ctx, int64_t *nCell, int64_t *nFace, int *M, int *maxFaceCell, int *nGFace,
                    int *nGCell, int *nColor, double *globalVol)
{
    CAPI_BEGIN
    requireMesh(ctx);
    const DevMesh &m = ctx->c.m;
    if (nCell)
        *nCell = m.N;
    if (nFace)
        *nFace = m.F;
    if (M)
        *M = m.M;
    if (maxFaceCell)
        *maxFaceCell = m.K;
    if (nGFace)
        *nGFace = m.G;
    if (nGCell)
        *nGCell = m.Gc;
    if (nColor)
        *nColor = m.nColor;
    if (globalVol)
        *globalVol = m.globalVol;
    CAPI_END
}

int dndsr_get_cell_metrics(dndsr_ctx *ctx, double *vol, double *bary, double *cent, double *scale,
                           double *smoothRatio, double *moment)
{
    CAPI_BEGIN
    requireMesh(ctx);
    Ctx &c = ctx->c;
    const DevMesh &m = c.m;
    const size_t N = m.N;
    auto v = download(c, m.vol, N), b = download(c, m.bary, 3 * N), ce = download(c, m.cent, 3 * N);
    auto sc = download(c, m.scale, N), sm = download(c, m.smooth, N), mo = download(c, m.moment, (size_t)m.M * N);
    for (size_t i = 0; i < N; i++)
    {
        if (vol)
            vol[i] = v[i];
        if (scale)
            scale[i] = sc[i];
        if (smoothRatio)
            smoothRatio[i] = sm[i];
        for (int r = 0; r < 3; r++)
        {
            if (bary)
                bary[3 * i + r] = b[r * N + i];
            if (cent)
                cent[3 * i + r] = ce[r * N + i];
        }
        if (moment)
            for (int k = 0; k < m.M; k++)
                moment[i * m.M + k] = mo[k * N + i];
    }
    CAPI_END
}

int dndsr_get_face_metrics(dndsr_ctx *ctx, double *area, double *meanNorm, double *pts, double *norms,
                           double *wJDet, double *faceL, double *faceWeight)
{
    CAPI_BEGIN
    requireMesh(ctx);
    Ctx &c = ctx->c;
    const DevMesh &m = c.m;
    const size_t F = m.F, G = m.G;
    auto a = download(c, m.farea, F), mn = download(c, m.fmeanN, 3 * F), p = download(c, m.fpts, G * 3 * F);
    auto n = download(c, m.fnorm, G * 3 * F), w = download(c, m.fwj, G * F), l = download(c, m.faceL, F);
    auto fw = download(c, m.faceW, (size_t)(m.P + 1) * F);
    for (size_t f = 0; f < F; f++)
    {
        if (area)
            area[f] = a[f];
        if (faceL)
            faceL[f] = l[f];
        for (int r = 0; r < 3; r++)
            if (meanNorm)
                meanNorm[3 * f + r] = mn[r * F + f];
        for (size_t g = 0; g < G; g++)
        {
            for (int r = 0; r < 3; r++)
            {
                if (pts)
                    pts[(f * G + g) * 3 + r] = p[(g * 3 + r) * F + f];
                if (norms)
                    norms[(f * G + g) * 3 + r] = n[(g * 3 + r) * F + f];
            }
            if (wJDet)
                wJDet[f * G + g] = w[g * F + f];
        }
        if (faceWeight)
            for (int o = 0; o <= m.P; o++)
                faceWeight[f * (m.P + 1) + o] = fw[o * F + f];
    }
    CAPI_END
}

int dndsr_get_cell_quad(dndsr_ctx *ctx, double *pts, double *wJDet)
{
    CAPI_BEGIN
    requireMesh(ctx);
    Ctx &c = ctx->c;
    const DevMesh &m = c.m;
    const size_t N = m.N, G = m.Gc;
    auto p = download(c, m.cpts, G * 3 * N), w = download(c, m.cwj, G * N);
    for (size_t i = 0; i < N; i++)
        for (size_t g = 0; g < G; g++)
        {
            for (int r = 0; r < 3; r++)
                pts[(i * G + g) * 3 + r] = p[(g * 3 + r) * N + i];
            wJDet[i * G + g] = w[g * N + i];
        }
    CAPI_END
}

int dndsr_get_vr_coeffs(dndsr_ctx *ctx, double *AinvB, double *Ainvb)
{
    CAPI_BEGIN
    requireMesh(ctx);
    Ctx &c = ctx->c;
    const DevMesh &m = c.m;
    const size_t N = m.N, K = m.K, M = m.M;
    auto B = download(c, m.AinvB, (K + 1) * M * M * N);
    auto b = download(c, m.Ainvb, K * M * N);
    for (size_t i = 0; i < N; i++)
    {
        for (size_t q = 0; q < (K + 1) * M * M; q++)
            AinvB[i * (K + 1) * M * M + q] = B[q * N + i];
        for (size_t q = 0; q < K * M; q++)
            Ainvb[i * K * M + q] = b[q * N + i];
    }
    CAPI_END
}

int dndsr_set_vr_coeffs(dndsr_ctx *ctx, const double *AinvB, const double *Ainvb)
{
    CAPI_BEGIN
    requireMesh(ctx);
    Ctx &c = ctx->c;
    const DevMesh &m = c.m;
    const size_t N = m.N, K = m.K, M = m.M;
    std::vector<real> B((K + 1) * M * M * N), b(K * M * N);
    for (size_t i = 0; i < N; i++)
    {
        for (size_t q = 0; q < (K + 1) * M * M; q++)
            B[q * N + i] = AinvB[i * (K + 1) * M * M + q];
        for (size_t q = 0; q < K * M; q++)
            b[q * N + i] = Ainvb[i * K * M + q];
    }
    upload(c, m.AinvB, B);
    upload(c, m.Ainvb, b);
    DNDSR_HIP_CHECK(hipStreamSynchronize(c.stream));
    CAPI_END
}

int dndsr_get_colors(dndsr_ctx *ctx, int32_t *color)
{
    CAPI_BEGIN
    requireMesh(ctx);
    auto h = download(ctx->c, ctx->c.m.color, ctx->c.m.N);
    std::copy(h.begin(), h.end(), color);
    CAPI_END
}

int dndsr_h2d_cellfield(dndsr_ctx *ctx, const double *hostRows, double *dev, int nComp)
{
    CAPI_BEGIN
    requireMesh(ctx);
    need(hostRows && dev && nComp > 0, "bad argument");
    const size_t N = ctx->c.m.N;
    std::vector<real> t((size_t)nComp * N);
    for (size_t i = 0; i < N; i++)
        for (int k = 0; k < nComp; k++)
            t[(size_t)k * N + i] = hostRows[i * nComp + k];
    upload(ctx->c, dev, t);
    DNDSR_HIP_CHECK(hipStreamSynchronize(ctx->c.stream));
    CAPI_END
}

int dndsr_d2h_cellfield(dndsr_ctx *ctx, const double *dev, double *hostRows, int nComp)
{
    CAPI_BEGIN
    requireMesh(ctx);
    need(hostRows && dev && nComp > 0, "bad argument");
    const size_t N = ctx->c.m.N;
    auto t = download(ctx->c, dev, (size_t)nComp * N);
    for (size_t i = 0; i < N; i++)
        for (int k = 0; k < nComp; k++)
            hostRows[i * nComp + k] = t[(size_t)k * N + i];
    CAPI_END
}

int dndsr_vr_iter(dndsr_ctx *ctx, int nVars, double *uRec, double *uRecNew, const double *u, int mode,
                  double relax)
{
    CAPI_BEGIN
    requireMesh(ctx);
    need(mode >= 0 && mode <= 2 || mode == 4, "mode must be 0, 1, 2 or 4");
    if (relax < 0)
        relax = ctx->c.m.s.jacobiRelax;
    if (mode == 4)
    {
        need(uRec && u, "null field");
        launchVRSOR(ctx->c, nVars, uRec, u, relax);
    }
    else
    {
        need(uRec && uRecNew && u, "null field");
        need(uRec != uRecNew, "uRec and uRecNew must be different arrays (Jacobi form)");
        launchVRIter(ctx->c, nVars, uRec, uRecNew, u, mode, relax);
    }
    CAPI_END
}

int dndsr_vr_iter_diff(dndsr_ctx *ctx, int nVars, const double *uRec, const double *u, const double *uRecDiff,
                       double *uRecNew)
{
    CAPI_BEGIN
    requireMesh(ctx);
    need(uRecDiff && uRecNew && uRecDiff != uRecNew, "bad fields");
    launchVRIterDiff(ctx->c, nVars, uRec, u, uRecDiff, uRecNew);
    CAPI_END
}

static void setBC(dndsr_ctx *ctx, int nZone, const int32_t *type, const double *value, double gamma, double Rgas)
{
    requireMesh(ctx);
    Ctx &c = ctx->c;
    DevMesh &m = c.m;
    need(nZone == m.nBCZone, "the zone count differs from the mesh's nBCZone");
    need(nZone > 0 && type, "no boundary zones");
    need(gamma > 1.0, "gamma must exceed 1");
    std::vector<int32_t> t(type, type + nZone);
    std::vector<real> v((size_t)nZone * 5, 0.0);
    for (int z = 0; z < nZone; z++)
    {
        need(t[z] == BC_FAR || t[z] == BC_WALL || t[z] == BC_WALL_INVIS || t[z] == BC_WALL_ISO || t[z] == BC_OUT ||
                 t[z] == BC_OUTP || t[z] == BC_IN || t[z] == BC_IN_PSTS || t[z] == BC_SYM,
             "boundary type not supported on this path (BCFar, BCWall, BCWallInvis, BCWallIsothermal, BCOut, BCOutP, "
             "BCIn, BCInPsTs, BCSym)");
        real *vz = v.data() + (size_t)z * 5;
        if (t[z] == BC_FAR || t[z] == BC_OUTP || t[z] == BC_IN)
        {
            need(value != nullptr, "far-field / inflow zones need a value");
            for (int k = 0; k < 5; k++)
                vz[k] = value[(size_t)z * 5 + k];
            need(vz[0] > 0, "boundary state with nonpositive density");
        }
        else if (t[z] == BC_WALL_ISO || t[z] == BC_IN_PSTS)
        {
            need(Rgas > 0, "BCWallIsothermal and BCInPsTs need the gas constant (dndsr_set_bc_gas with Rgas > 0)");
            need(value != nullptr, "isothermal-wall / total-condition zones need a value");
            const double *w = value + (size_t)z * 5;
            if (t[z] == BC_WALL_ISO) // (T_wall): stored Rgas T_wall
            {
                need(w[0] > 0, "isothermal wall with nonpositive temperature");
                vz[0] = Rgas * w[0];
            }
            else // (p0, T0, direction): stored (p0, Rgas T0, unit direction)
            {
                const double dn = std::sqrt(w[2] * w[2] + w[3] * w[3] + w[4] * w[4]);
                need(w[0] > 0 && w[1] > 0 && dn > 0, "total-condition inflow needs p0 > 0, T0 > 0 and a direction");
                vz[0] = w[0], vz[1] = Rgas * w[1];
                for (int d = 0; d < 3; d++)
                    vz[2 + d] = w[2 + d] / dn;
            }
        }
    }
    if (!m.bcType)
    {
        m.bcType = devAlloc<int32_t>(c, nZone);
        m.bcVal = devAlloc<real>(c, (size_t)nZone * 5);
    }
    upload(c, m.bcType, t);
    upload(c, m.bcVal, v);
    m.bcGamma = gamma;
    m.gasBC = false;
    for (int z = 0; z < nZone; z++)
        m.gasBC = m.gasBC || t[z] == BC_WALL_ISO || t[z] == BC_IN_PSTS;
}

int dndsr_set_bc(dndsr_ctx *ctx, int nZone, const int32_t *type, const double *value, double gamma)
{
    CAPI_BEGIN
    setBC(ctx, nZone, type, value, gamma, 0.0);
    CAPI_END
}

int dndsr_set_bc_gas(dndsr_ctx *ctx, int nZone, const int32_t *type, const double *value, double gamma, double Rgas)
{
    CAPI_BEGIN
    need(Rgas > 0, "Rgas must be positive");
    setBC(ctx, nZone, type, value, gamma, Rgas);
    CAPI_END
}

int dndsr_set_vr_dirichlet(dndsr_ctx *ctx, int nVars, const double *value)
{
    CAPI_BEGIN
    requireMesh(ctx);
    need(value == nullptr || (nVars >= 1 && nVars <= 5), "nVars must be 1..5");
    ctx->c.m.dirVal = value;
    ctx->c.m.dirNV = value ? nVars : 0;
    CAPI_END
}

static GasParams toGas(const dndsr_euler_settings *es)
{
    need(es != nullptr, "null euler settings");
    need(es->gamma > 1.0, "gamma must exceed 1");
    if (es->viscous)
    {
        need(es->muModel >= 0 && es->muModel <= 2, "muModel must be 0 (constant), 1 (Sutherland) or 2 (constant nu)");
        need(es->Rgas > 0 && es->prGas > 0 && es->muGas >= 0, "viscous gas needs Rgas > 0, prGas > 0, muGas >= 0");
        need(es->muModel != 1 || (es->TRef > 0 && es->CSutherland >= 0), "Sutherland's law needs TRef > 0, CSutherland >= 0");
    }
    return GasParams{es->gamma, es->rsType, es->rsMeanValueEig, es->rsFixScale, es->rsIncFScale, es->viscous ? 1 : 0,
                     es->muModel, es->Rgas, es->muGas, es->prGas, es->TRef, es->CSutherland};
}

int dndsr_eval_dt(dndsr_ctx *ctx, const dndsr_euler_settings *es, const double *u, double CFL, double maxDt,
                  int useLocalDt, double *dt, double *lambdaFace, double *deltaLambdaCell, double *dtMinDev,
                  double *dtMinHost)
{
    CAPI_BEGIN
    requireMesh(ctx);
    need(u && dt && lambdaFace, "null field");
    Ctx &c = ctx->c;
    real *dmin = dtMinDev ? dtMinDev : c.red + 5 * ((size_t)gridFor(c.m.N, 256) + c.m.nColor + 7);
    launchEvalDt(c, toGas(es), u, CFL, maxDt, useLocalDt, dt, lambdaFace, deltaLambdaCell, dmin);
    if (dtMinHost)
    {
        DNDSR_HIP_CHECK(hipMemcpyAsync(c.hostRed, dmin, sizeof(real), hipMemcpyDeviceToHost, c.stream));
        DNDSR_HIP_CHECK(hipStreamSynchronize(c.stream));
        *dtMinHost = c.hostRed[0];
    }
    CAPI_END
}

int dndsr_eval_rhs(dndsr_ctx *ctx, const dndsr_euler_settings *es, const double *u, const double *uRec,
                   const double *deltaLambdaCell, double *rhs)
{
    CAPI_BEGIN
    requireMesh(ctx);
    need(u && rhs, "null field");
    launchEvalRHS(ctx->c, toGas(es), u, uRec, deltaLambdaCell, rhs);
    CAPI_END
}

int dndsr_eval_rhs_noalloc(dndsr_ctx *ctx, const dndsr_euler_settings *es, const double *u, const double *uRec,
                           const double *deltaLambdaCell, double *rhs)
{
    CAPI_BEGIN
    requireMesh(ctx);
    need(u && rhs, "null field");
    Ctx &c = ctx->c;
    need((size_t)c.m.F <= kNoAllocMaxFaces,
         "dndsr_eval_rhs_noalloc: the mesh has more faces than its scratch holds (1048576); use dndsr_eval_rhs");
    need(ctx->scratchDev != nullptr, "no-allocation scratch not resolved");
    // every face's flux is written by the face passes before the gather reads it, so the
    // scratch needs no clearing; timing events are not taken (they would be created here)
    const bool timing = c.timing;
    c.timing = false;
    try
    {
        launchEvalRHS(c, toGas(es), u, uRec, deltaLambdaCell, rhs, nullptr, ctx->scratchDev, false);
    }
    catch (...)
    {
        c.timing = timing;
        throw;
    }
    c.timing = timing;
    CAPI_END
}

int dndsr_model_rhs(dndsr_ctx *ctx, int nVars, double ax, double ay, double sigma, const double *u,
                    const double *uRec, double *rhs)
{
    CAPI_BEGIN
    requireMesh(ctx);
    need(u && rhs, "null field");
    need(sigma >= 0, "sigma must be non-negative");
    launchModelRHS(ctx->c, nVars, ax, ay, sigma, u, uRec, rhs);
    CAPI_END
}

int dndsr_lusgs_init(dndsr_ctx *ctx, const double *dTau, double dt, double alphaDiag, const double *lambdaFace,
                     double *jdiag)
{
    CAPI_BEGIN
    requireMesh(ctx);
    need(dTau && lambdaFace && jdiag, "null field");
    launchLUSGSInit(ctx->c, dTau, dt, alphaDiag, lambdaFace, jdiag);
    CAPI_END
}

static int *blockFlag(dndsr_ctx *ctx)
{
    if (!ctx->bBad)
    {
        ctx->bBad = devAlloc<int>(ctx->c, 1);
        DNDSR_HIP_CHECK(hipMemsetAsync(ctx->bBad, 0, sizeof(int), ctx->c.stream));
    }
    return ctx->bBad;
}

// a singular JDiag block raised the flag: clear it and report
static void checkBlocks(dndsr_ctx *ctx, int flagHost)
{
    if (!flagHost)
        return;
    DNDSR_HIP_CHECK(hipMemsetAsync(ctx->bBad, 0, sizeof(int), ctx->c.stream));
    DNDSR_HIP_CHECK(hipStreamSynchronize(ctx->c.stream));
    throw std::runtime_error("a block of JDiag is singular or has a zero diagonal entry (GetInvert)");
}

int dndsr_eval_rhs_roe_lambda(dndsr_ctx *ctx, const dndsr_euler_settings *es, const double *u, const double *uRec,
                              const double *deltaLambdaCell, double *lamRoe)
{
    CAPI_BEGIN
    requireMesh(ctx);
    need(u && lamRoe, "null field");
    launchFaceRoeLambda(ctx->c, toGas(es), u, uRec, deltaLambdaCell, lamRoe);
    CAPI_END
}

int dndsr_lusgs_init_block(dndsr_ctx *ctx, const dndsr_euler_settings *es, const double *dTau, double dt,
                           double alphaDiag, const double *u, const double *lambdaFace, const double *lamRoe,
                           double *jdiag, double *jdiagInv)
{
    CAPI_BEGIN
    requireMesh(ctx);
    need(dTau && u && lambdaFace && jdiagInv, "null field");
    Ctx &c = ctx->c;
    int *bad = blockFlag(ctx);
    launchLUSGSInitBlock(c, toGas(es), dTau, dt, alphaDiag, u, lambdaFace, lamRoe, jdiag, jdiagInv, bad);
    int h = 0;
    DNDSR_HIP_CHECK(hipMemcpyAsync(&h, bad, sizeof(int), hipMemcpyDeviceToHost, c.stream));
    DNDSR_HIP_CHECK(hipStreamSynchronize(c.stream));
    checkBlocks(ctx, h);
    CAPI_END
}

static void fetchSum(Ctx &c, real *sumIncHost)
{
    if (!sumIncHost)
        return;
    real *dev = c.red + 5 * ((size_t)gridFor(c.m.N, 256) + c.m.nColor + 6);
    // the final reduce already wrote there (see callers)
    DNDSR_HIP_CHECK(hipMemcpyAsync(c.hostRed, dev, 5 * sizeof(real), hipMemcpyDeviceToHost, c.stream));
    DNDSR_HIP_CHECK(hipStreamSynchronize(c.stream));
    for (int k = 0; k < 5; k++)
        sumIncHost[k] = c.hostRed[k];
}

int dndsr_update_sgs(dndsr_ctx *ctx, const dndsr_euler_settings *es, double alphaDiag, const double *rhs,
                     const double *u, const double *uInc, double *uIncNew, const double *jdiag,
                     const double *lambdaFace, int forward, int gsUpdate, int uIncIsZero, double *sumIncHost)
{
    CAPI_BEGIN
    requireMesh(ctx);
    need(rhs && u && uInc && uIncNew && jdiag && lambdaFace, "null field");
    need(uInc != uIncNew, "uInc and uIncNew must be different arrays");
    Ctx &c = ctx->c;
    real *sumDev = c.red + 5 * ((size_t)gridFor(c.m.N, 256) + c.m.nColor + 6);
    launchSGSSweep64(c, toGas(es), alphaDiag, rhs, u, uInc, uIncNew, jdiag, lambdaFace, forward, gsUpdate, uIncIsZero,
                     sumDev);
    fetchSum(c, sumIncHost);
    CAPI_END
}

int dndsr_update_sgs_transpose(dndsr_ctx *ctx, const dndsr_euler_settings *es, double alphaDiag, const double *rhs,
                               const double *u, const double *uInc, double *uIncNew, const double *jdiag,
                               const double *lambdaFace, int forward, int gsUpdate, int uIncIsZero,
                               double *sumIncHost)
{
    CAPI_BEGIN
    requireMesh(ctx);
    need(rhs && u && uInc && uIncNew && jdiag && lambdaFace, "null field");
    need(uInc != uIncNew, "uInc and uIncNew must be different arrays");
    Ctx &c = ctx->c;
    real *sumDev = c.red + 5 * ((size_t)gridFor(c.m.N, 256) + c.m.nColor + 6);
    launchSGSSweepT64(c, toGas(es), alphaDiag, rhs, u, uInc, uIncNew, jdiag, lambdaFace, forward, gsUpdate,
                      uIncIsZero, sumDev);
    fetchSum(c, sumIncHost);
    CAPI_END
}

int dndsr_update_sgs_f32(dndsr_ctx *ctx, const dndsr_euler_settings *es, float alphaDiag, const float *rhs,
                         const float *u, const float *uInc, float *uIncNew, const float *jdiag,
                         const float *lambdaFace, int forward, int gsUpdate, int uIncIsZero, double *sumIncHost)
{
    CAPI_BEGIN
    requireMesh(ctx);
    need(rhs && u && uInc && uIncNew && jdiag && lambdaFace, "null field");
    need(uInc != uIncNew, "uInc and uIncNew must be different arrays");
    Ctx &c = ctx->c;
    real *sumDev = c.red + 5 * ((size_t)gridFor(c.m.N, 256) + c.m.nColor + 6);
    launchSGSSweep32(c, toGas(es), alphaDiag, rhs, u, uInc, uIncNew, jdiag, lambdaFace, forward, gsUpdate, uIncIsZero,
                     sumDev);
    fetchSum(c, sumIncHost);
    CAPI_END
}

int dndsr_update_sgs_block(dndsr_ctx *ctx, const dndsr_euler_settings *es, double alphaDiag, const double *rhs,
                           const double *u, const double *uInc, double *uIncNew, const double *jdiagInv,
                           const double *lambdaFace, const double *lamRoe, int forward, int gsUpdate, int uIncIsZero,
                           double *sumIncHost)
{
    CAPI_BEGIN
    requireMesh(ctx);
    need(rhs && u && uInc && uIncNew && jdiagInv && lambdaFace, "null field");
    need(uInc != uIncNew, "uInc and uIncNew must be different arrays");
    Ctx &c = ctx->c;
    real *sumDev = c.red + 5 * ((size_t)gridFor(c.m.N, 256) + c.m.nColor + 6);
    launchSGSSweepBlock64(c, toGas(es), alphaDiag, rhs, u, uInc, uIncNew, jdiagInv, lambdaFace, lamRoe, forward,
                          gsUpdate, uIncIsZero, sumDev);
    fetchSum(c, sumIncHost);
    CAPI_END
}

int dndsr_update_sgs_block_transpose(dndsr_ctx *ctx, const dndsr_euler_settings *es, double alphaDiag,
                                     const double *rhs, const double *u, const double *uInc, double *uIncNew,
                                     const double *jdiagInv, const double *lambdaFace, const double *lamRoe,
                                     int forward, int gsUpdate, int uIncIsZero, double *sumIncHost)
{
    CAPI_BEGIN
    requireMesh(ctx);
    need(rhs && u && uInc && uIncNew && jdiagInv && lambdaFace, "null field");
    need(uInc != uIncNew, "uInc and uIncNew must be different arrays");
    Ctx &c = ctx->c;
    real *sumDev = c.red + 5 * ((size_t)gridFor(c.m.N, 256) + c.m.nColor + 6);
    launchSGSSweepBlockT64(c, toGas(es), alphaDiag, rhs, u, uInc, uIncNew, jdiagInv, lambdaFace, lamRoe, forward,
                           gsUpdate, uIncIsZero, sumDev);
    fetchSum(c, sumIncHost);
    CAPI_END
}

// nSweep symmetric sweeps of the transposed LU-SGS matrix (scalar with jd, block with jinv)
// from zero increments, ping-ponging between a and b; with Gauss-Seidel updates a sweep after
// the first skips its first colour, whose pass would repeat the previous sweep's last one (same
// inputs, same values): that sweep reads the colour from the array it reads, the next sweep
// reads it from the array it writes until the colour's own pass, and one copy after the last
// sweep puts the skipped colour's values into the result.
static real *sgsTransposeSymmetric(Ctx &c, const GasParams &gp, real alphaDiag, const real *rhs, const real *u,
                                   const real *jd, const real *jinv, const real *lam, const real *lamRoe, int nSweep,
                                   int gsUpdate, real *a, real *b)
{
    DevMesh &m = c.m;
    const size_t nU = (size_t)5 * m.N;
    real *sumDev = c.red + 5 * ((size_t)gridFor(m.N, 256) + m.nColor + 6);
    DNDSR_HIP_CHECK(hipMemsetAsync(a, 0, nU * sizeof(real), c.stream));
    DNDSR_HIP_CHECK(hipMemsetAsync(b, 0, nU * sizeof(real), c.stream));
    const int nC = m.nColor;
    const bool canSkip = gsUpdate && nC >= 2;
    int prevSkip = -1;
    for (int s = 0; s < nSweep; s++)
        for (int dir = 1; dir >= 0; dir--)
        {
            const int first = dir ? 0 : nC - 1, last = dir ? nC - 1 : 0;
            const int fCol = canSkip && !(s == 0 && dir == 1) ? first : -1;
            const int lCol = prevSkip == last ? last : -1;
            if (jinv)
                launchSGSSweepBlockT64(c, gp, alphaDiag, rhs, u, a, b, jinv, lam, lamRoe, dir, gsUpdate, 0, sumDev,
                                       fCol, lCol);
            else
                launchSGSSweepT64(c, gp, alphaDiag, rhs, u, a, b, jd, lam, dir, gsUpdate, 0, sumDev, fCol, lCol);
            prevSkip = fCol;
            std::swap(a, b);
        }
    if (prevSkip >= 0)
        launchCopyColourInc64(c, prevSkip, b, a);
    return a;
}

int dndsr_sgs_transpose_symmetric(dndsr_ctx *ctx, const dndsr_euler_settings *es, double alphaDiag,
                                  const double *rhs, const double *u, const double *jdiag, const double *jdiagInv,
                                  const double *lambdaFace, const double *lamRoe, int nSweep, int gsUpdate,
                                  double *work, double *out)
{
    CAPI_BEGIN
    requireMesh(ctx);
    need(rhs && u && lambdaFace && work && out, "null field");
    need((jdiag != nullptr) != (jdiagInv != nullptr), "pass exactly one of jdiag (scalar) and jdiagInv (block)");
    need(!jdiagInv || lamRoe, "the block sweeps need the Roe face eigenvalues (lamRoe)");
    need(work != out && work != rhs && out != rhs, "work, out and rhs must be different arrays");
    need(nSweep >= 1, "nSweep must be at least 1");
    Ctx &c = ctx->c;
    real *res = sgsTransposeSymmetric(c, toGas(es), alphaDiag, rhs, u, jdiag, jdiagInv, lambdaFace, lamRoe, nSweep,
                                      gsUpdate, out, work);
    if (res != out)
        DNDSR_HIP_CHECK(hipMemcpyAsync(out, res, (size_t)5 * c.m.N * sizeof(real), hipMemcpyDeviceToDevice,
                                       c.stream));
    CAPI_END
}

int dndsr_update_sgs_block_f32(dndsr_ctx *ctx, const dndsr_euler_settings *es, float alphaDiag, const float *rhs,
                               const float *u, const float *uInc, float *uIncNew, const float *jdiagInv,
                               const float *lambdaFace, const float *lamRoe, int forward, int gsUpdate,
                               int uIncIsZero, double *sumIncHost)
{
    CAPI_BEGIN
    requireMesh(ctx);
    need(rhs && u && uInc && uIncNew && jdiagInv && lambdaFace, "null field");
    need(uInc != uIncNew, "uInc and uIncNew must be different arrays");
    Ctx &c = ctx->c;
    real *sumDev = c.red + 5 * ((size_t)gridFor(c.m.N, 256) + c.m.nColor + 6);
    launchSGSSweepBlock32(c, toGas(es), alphaDiag, rhs, u, uInc, uIncNew, jdiagInv, lambdaFace, lamRoe, forward,
                          gsUpdate, uIncIsZero, sumDev);
    fetchSum(c, sumIncHost);
    CAPI_END
}

int dndsr_stream_probe(dndsr_ctx *ctx, int mode, const double *src, double *dst, int64_t n, int reps,
                       double *msPerPass)
{
    CAPI_BEGIN
    need(ctx != nullptr, "null context");
    need(src && dst && msPerPass, "null argument");
    need(mode == 0 || mode == 1, "mode must be 0 (read) or 1 (copy)");
    need(n >= 2 && n % 2 == 0 && reps >= 1, "n must be even and positive, reps at least 1");
    need(((uintptr_t)src % 16) == 0 && ((uintptr_t)dst % 16) == 0, "src and dst must be 16-byte aligned");
    need(mode == 1 || streamProbeBlocks(ctx->c) <= 4096, "more than 4096 partial sums");
    *msPerPass = streamProbe(ctx->c, mode, src, dst, (size_t)n, reps);
    CAPI_END
}

int dndsr_set_timing(dndsr_ctx *ctx, int enable)
{
    CAPI_BEGIN
    need(ctx != nullptr, "null context");
    ctx->c.timing = enable != 0;
    CAPI_END
}

int dndsr_get_timing(dndsr_ctx *ctx, double *msTotal, int64_t *launches, int reset)
{
    CAPI_BEGIN
    need(ctx != nullptr, "null context");
    Ctx &c = ctx->c;
    DNDSR_HIP_CHECK(hipStreamSynchronize(c.stream));
    for (auto &r : c.recs)
    {
        float ms = 0;
        DNDSR_HIP_CHECK(hipEventElapsedTime(&ms, r.a, r.b));
        c.kernMs[r.id] += ms;
        c.kernCount[r.id] += 1;
        c.evPool.push_back(r.a);
        c.evPool.push_back(r.b);
    }
    c.recs.clear();
    for (int k = 0; k < K_NUM; k++)
    {
        if (msTotal)
            msTotal[k] = c.kernMs[k];
        if (launches)
            launches[k] = c.kernCount[k];
        if (reset)
            c.kernMs[k] = 0, c.kernCount[k] = 0;
    }
    CAPI_END
}

// One reconstruction sweep of the solver loops (frhs, EulerSolver.hxx:240-250, putIntoNew =
// false): SOR in place when SORInstead, else Jacobi into recNew and the two swapped.
static void solverVRSweep(Ctx &c, real *&rec, real *&recNew, const real *u)
{
    if (usesSubs2nd(c.m)) // :503-512: the substitute reads only u
        launchVR2nd(c, 5, u, rec);
    else if (c.m.s.SORInstead)
        launchVRSOR(c, 5, rec, u, c.m.s.jacobiRelax);
    else
    {
        launchVRIter(c, 5, rec, recNew, u, 0, c.m.s.jacobiRelax);
        std::swap(rec, recNew);
    }
}

int dndsr_ssprk3(dndsr_ctx *ctx, const dndsr_euler_settings *es, double *u, double *uRec, int nStep, double CFL,
                 int nRec, double dtFixed)
{
    CAPI_BEGIN
    requireMesh(ctx);
    need(u && uRec && nStep >= 0 && nRec >= 0, "bad argument");
    Ctx &c = ctx->c;
    DevMesh &m = c.m;
    GasParams gp = toGas(es);
    const size_t N = m.N;
    const size_t nU = 5 * N, nR = 5 * (size_t)m.M * N;
    timeLoopWork(ctx);
    real *rec = uRec, *recNew = ctx->wURecNew;
    // one RK stage: VR sweeps at u, RHS(u), and u = a u + b xLast + cdt dt RHS in the gather pass
    auto stage = [&](real a, real b, real cdt, const real *dtp, real *save) {
        for (int r = 0; r < nRec; r++)
            solverVRSweep(c, rec, recNew, u);
        StageUpdate st{a, b, cdt, b != 0.0 ? ctx->wXLast : nullptr, dtp, u, save};
        launchEvalRHS(c, gp, u, rec, ctx->wDlc, nullptr, &st);
    };
    for (int step = 0; step < nStep; step++)
    {
        launchEvalDt(c, gp, u, CFL, 1e100, 0, nullptr, ctx->wLam, ctx->wDlc, ctx->wDtMin); // global dt only
        const real *dtp = ctx->wDtMin;
        if (dtFixed > 0)
        {
            c.hostRed[0] = dtFixed;
            DNDSR_HIP_CHECK(hipMemcpyAsync(ctx->wDtMin, c.hostRed, sizeof(real), hipMemcpyHostToDevice, c.stream));
            DNDSR_HIP_CHECK(hipStreamSynchronize(c.stream));
        }
        stage(1.0, 0.0, 1.0, dtp, ctx->wXLast); // xLast = u, then u1 = u + dt R(u)
        stage(0.25, 0.75, 0.25, dtp, nullptr);
        stage(2. / 3., 1. / 3., 2. / 3., dtp, nullptr);
    }
    if (rec != uRec) // odd number of VR sweeps in total: bring the result home
        launchCopy(c, nR, rec, uRec);
    DNDSR_HIP_CHECK(hipStreamSynchronize(c.stream));
    checkFinite(c, "in the SSP-RK3 steps");
    CAPI_END
}

// BDF coefficients of ImplicitBDFDualTimeStep (src/Solver/ODE.hpp:590-595): row k - 1 holds
// (alpha, c_last, c_prev1, ...) of the k-step formula
static const real kBDFCoef[4][5] = {{1. / 1., 1. / 1., 0, 0, 0},
                                    {2. / 3., 4. / 3., -1. / 3., 0, 0},
                                    {6. / 11., 18. / 11., -9. / 11., 2. / 11., 0},
                                    {12. / 25., 48. / 25., -36. / 25., 16. / 25., -3. / 25.}};

static void implicitStep(dndsr_ctx *ctx, const dndsr_euler_settings *es, const dndsr_implicit_settings *is, int kCur,
                         const double *const *xPrevs, double *u, double *uRec, int *nIter, double *res,
                         double *resRel, double *history);

int dndsr_implicit_euler_step(dndsr_ctx *ctx, const dndsr_euler_settings *es, const dndsr_implicit_settings *is,
                              double *u, double *uRec, int *nIter, double *res, double *resRel, double *history)
{
    CAPI_BEGIN
    implicitStep(ctx, es, is, 0, nullptr, u, uRec, nIter, res, resRel, history);
    CAPI_END
}

int dndsr_implicit_bdf_step(dndsr_ctx *ctx, const dndsr_euler_settings *es, const dndsr_implicit_settings *is,
                            int nPrev, const double *const *xPrevs, double *u, double *uRec, int *nIter, double *res,
                            double *resRel, double *history)
{
    CAPI_BEGIN
    need(nPrev >= 0 && nPrev <= 3, "nPrev must be 0..3 (BDF1..BDF4)");
    need(nPrev == 0 || xPrevs != nullptr, "null earlier states");
    for (int j = 0; j < nPrev; j++)
        need(xPrevs[j] != nullptr, "null earlier state");
    implicitStep(ctx, es, is, nPrev + 1, xPrevs, u, uRec, nIter, res, resRel, history);
    CAPI_END
}

// kCur 0: ImplicitEulerDualTimeStep::Step; 1..4: the BDF step with kCur - 1 earlier states.
static void implicitStep(dndsr_ctx *ctx, const dndsr_euler_settings *es, const dndsr_implicit_settings *is, int kCur,
                         const double *const *xPrevs, double *u, double *uRec, int *nIter, double *res,
                         double *resRel, double *history)
{
    requireMesh(ctx);
    need(es && is && u && uRec, "null argument");
    need(is->dt > 0 && is->CFL > 0 && is->nRec >= 0 && is->sgsIter >= 0 && is->maxIter >= 1 && is->minIter >= 0,
         "bad implicit settings");
    need(is->gmresCode >= 0 && is->gmresCode <= 2 && (is->gmresCode == 0 || (is->nGmresSpace >= 1 && is->nGmresIter >= 0)),
         "bad GMRES settings");
    need(is->gsUpdate == 0 || is->gsUpdate == 1,
         "gsUpdate (jacobiCode) must be 0 (Jacobi) or 1 (Gauss-Seidel); jacobiCode 2 (ILU) is not provided");
    need(is->incrementRelax >= 0 && is->incrementRelax <= 1, "incrementRelax must lie in [0, 1]");
    const bool roeJ = es->useRoeJacobian != 0;
    need(!roeJ || is->gmresCode == 0, "useRoeJacobian runs with the LU-SGS solver (gmresCode 0)");
    Ctx &c = ctx->c;
    DevMesh &m = c.m;
    GasParams gp = toGas(es);
    const size_t N = m.N, F = m.F;
    const size_t nU = 5 * N, nR = 5 * (size_t)m.M * N;
    timeLoopWork(ctx);
    if (roeJ && !ctx->bJinv)
    {
        ctx->bLamRoe = devAlloc<real>(c, 3 * F);
        ctx->bJinv = devAlloc<real>(c, 25 * N);
    }
    if (roeJ && is->fp32Sweeps && !ctx->bJinvF)
    {
        ctx->bJinvF = devAlloc<float>(c, 25 * N);
        ctx->bLamRoeF = devAlloc<float>(c, 3 * F);
    }
    int *bBad = roeJ ? blockFlag(ctx) : nullptr;
    if (!ctx->haveImplicit)
    {
        ctx->iXLast = devAlloc<real>(c, nU);
        ctx->iRhsF = devAlloc<real>(c, nU);
        ctx->iRhs = devAlloc<real>(c, nU);
        ctx->iDTau = devAlloc<real>(c, N);
        ctx->iJd = devAlloc<real>(c, N);
        ctx->iInc = devAlloc<real>(c, nU);
        ctx->iTmp = devAlloc<real>(c, nU);
        ctx->iRes = devAlloc<real>(c, 8);
        ctx->fRhs = devAlloc<float>(c, nU);
        if (!ctx->fU)
            ctx->fU = devAlloc<float>(c, nU);
        if (!ctx->fJd)
            ctx->fJd = devAlloc<float>(c, N);
        if (!ctx->fLam)
            ctx->fLam = devAlloc<float>(c, 4 * F);
        ctx->fInc = devAlloc<float>(c, nU);
        ctx->fTmp = devAlloc<float>(c, nU);
        ctx->iStat = devAlloc<PPStat>(c, 1);
        ctx->iURecSave = devAlloc<real>(c, nR);
        ctx->haveImplicit = true;
    }
    const real gamma = es->gamma;
    const real alphaDiag = kCur ? kBDFCoef[kCur - 1][0] : 1.0;
    const real *xs[4] = {nullptr, nullptr, nullptr, nullptr};
    real coefs[4] = {0, 0, 0, 0};
    const real rhoEps = kSmall * is->refRho * 1e-1, pEps = kSmall * is->refP * 1e-1;
    launchCopy(c, nU, u, ctx->iXLast);
    launchCopy(c, nR, uRec, ctx->iURecSave);
    real *rec = uRec, *recNew = ctx->wURecNew;
    real resBase[5] = {0, 0, 0, 0, 0}, rr[5] = {0, 0, 0, 0, 0}, rel[5] = {0, 0, 0, 0, 0};
    ctx->ppLimited = 0;
    ctx->ppAlphaMin = 1.0;
    // a failed step leaves u and uRec as they were at entry
    auto fail = [&](const std::string &msg) {
        launchCopy(c, nU, ctx->iXLast, u);
        launchCopy(c, nR, ctx->iURecSave, uRec);
        DNDSR_HIP_CHECK(hipStreamSynchronize(c.stream));
        throw std::runtime_error(msg);
    };
    int it = 0;
    for (it = 1; it <= is->maxIter; it++)
    {
        launchPPStatReset(c, ctx->iStat);
        // fdtau: local pseudo time step (alphaDiag = 1)
        launchEvalDt(c, gp, u, is->CFL, 1e100, 1, ctx->iDTau, ctx->wLam, ctx->wDlc, ctx->wDtMin);
        if (kCur) // dTau *= 1 / alphaDiag (EulerSolver.hxx:693)
            launchScale(c, (idx)N, 1. / alphaDiag, ctx->iDTau);
        // frhs: nRec VR sweeps + EvaluateRHS
        for (int r = 0; r < is->nRec; r++)
            solverVRSweep(c, rec, recNew, u);
        launchEvalRHS(c, gp, u, rec, ctx->wDlc, ctx->iRhsF);
        if (roeJ) // the face eigenvalues EvaluateRHS keeps for the Roe Jacobian
            launchFaceRoeLambda(c, gp, u, rec, ctx->wDlc, ctx->bLamRoe);
        if (kCur)
        {
            xs[0] = ctx->iXLast, coefs[0] = kBDFCoef[kCur - 1][1];
            for (int j = 0; j < kCur - 1; j++)
                xs[1 + j] = xPrevs[j], coefs[1 + j] = kBDFCoef[kCur - 1][2 + j];
            launchDualResidualBDF(c, is->dt, xs, coefs, kCur, alphaDiag, u, ctx->iRhsF, ctx->iRhs, ctx->iRes);
        }
        else
            launchDualResidual(c, is->dt, ctx->iXLast, u, ctx->iRhsF, ctx->iRhs, ctx->iRes);
        // fsolve: LUSGSMatrixInit + (1 + sgsIter) symmetric sweeps from a zero increment
        if (roeJ)
            launchLUSGSInitBlock(c, gp, ctx->iDTau, is->dt, alphaDiag, u, ctx->wLam, ctx->bLamRoe, nullptr,
                                 ctx->bJinv, bBad);
        else
            launchLUSGSInit(c, ctx->iDTau, is->dt, alphaDiag, ctx->wLam, ctx->iJd);
        if (roeJ)
        {
            // block sweeps in colour order, ping-ponging the increment from zero
            real *sumDev = c.red + 5 * ((size_t)gridFor(m.N, 256) + m.nColor + 6);
            if (is->fp32Sweeps)
            {
                launchToF32(c, nU, ctx->iRhs, ctx->fRhs);
                launchToF32(c, nU, u, ctx->fU);
                launchToF32(c, 25 * N, ctx->bJinv, ctx->bJinvF);
                launchToF32(c, 4 * F, ctx->wLam, ctx->fLam);
                launchToF32(c, 3 * F, ctx->bLamRoe, ctx->bLamRoeF);
                float *a = ctx->fInc, *b = ctx->fTmp;
                DNDSR_HIP_CHECK(hipMemsetAsync(a, 0, nU * sizeof(float), c.stream));
                DNDSR_HIP_CHECK(hipMemsetAsync(b, 0, nU * sizeof(float), c.stream));
                float *res = nullptr;
                launchSGSBlockSymmetric32(c, gp, (float)alphaDiag, ctx->fRhs, ctx->fU, a, b, ctx->bJinvF, ctx->fLam,
                                          ctx->bLamRoeF, is->sgsIter, is->gsUpdate, sumDev, &res);
                launchAddIncrement32(c, gamma, is->incrementRelax, rhoEps, pEps, u, res, ctx->iStat);
            }
            else
            {
                real *a = ctx->iInc, *b = ctx->iTmp;
                DNDSR_HIP_CHECK(hipMemsetAsync(a, 0, nU * sizeof(real), c.stream));
                DNDSR_HIP_CHECK(hipMemsetAsync(b, 0, nU * sizeof(real), c.stream));
                real *res = nullptr;
                launchSGSBlockSymmetric64(c, gp, alphaDiag, ctx->iRhs, u, a, b, ctx->bJinv, ctx->wLam, ctx->bLamRoe,
                                          is->sgsIter, is->gsUpdate, sumDev, &res);
                launchAddIncrement64(c, gamma, is->incrementRelax, rhoEps, pEps, u, res, ctx->iStat);
            }
        }
        else if (is->gmresCode != 0)
        {
            // solveLinear with gmresCode 1 (GMRES from zero) or 2 (LU-SGS result as the start),
            // EulerSolver.hxx:1407-1578; the preconditioner sweeps in FP32 when fp32Sweeps
            LinWork &w = linWork(ctx, is->nGmresSpace, is->fp32Sweeps != 0);
            w.u = u, w.jd = ctx->iJd, w.lam = ctx->wLam, w.alphaDiag = alphaDiag;
            w.sgsIter = is->sgsIter, w.gsUpdate = is->gsUpdate, w.fp32 = is->fp32Sweeps != 0;
            if (w.fp32)
            {
                launchToF32(c, nU, u, ctx->fU);
                launchToF32(c, N, ctx->iJd, ctx->fJd);
                launchToF32(c, F, ctx->wLam, ctx->fLam);
                w.fU = ctx->fU, w.fJd = ctx->fJd, w.fLam = ctx->fLam;
            }
            if (is->gmresCode == 2)
                applyLUSGSPrecond(c, gp, w, ctx->iRhs, ctx->iInc);
            else
                DNDSR_HIP_CHECK(hipMemsetAsync(ctx->iInc, 0, nU * sizeof(real), c.stream));
            gmresSolve(c, gp, w, ctx->iRhs, ctx->iInc, is->nGmresSpace, is->nGmresIter, nullptr);
            launchAddIncrement64(c, gamma, is->incrementRelax, rhoEps, pEps, u, ctx->iInc, ctx->iStat);
        }
        else if (is->fp32Sweeps)
        {
            // sweeps with the neighbour state cell-major (the sweep's gathers), FP32
            const bool fresh = !ctx->fSlot;
            if (!ctx->fSlot)
                ctx->fSlot = devAlloc<float>(c, ((size_t)m.K * 5 + 7) * N), ctx->fUA = devAlloc<float>(c, 32 * N),
                ctx->fNbr = devAlloc<idx>(c, (size_t)m.K * N);
            launchToF32(c, nU, ctx->iRhs, ctx->fRhs);
            launchToF32(c, N, ctx->iJd, ctx->fJd);
            launchToF32(c, F, ctx->wLam, ctx->fLam); // lambdaFace row only
            float *inc = nullptr;
            launchSGSPacked32(c, gp, (float)alphaDiag, ctx->fRhs, u, ctx->fLam, ctx->fJd, ctx->fUA, ctx->fSlot, ctx->fNbr,
                              fresh, is->sgsIter, is->gsUpdate, &inc);
            launchAddIncrementCellMajor(c, gamma, is->incrementRelax, rhoEps, pEps, u, nullptr, inc, ctx->iStat);
        }
        else
        {
            const bool fresh = !ctx->iSlot;
            if (!ctx->iSlot)
                ctx->iSlot = devAlloc<real>(c, ((size_t)m.K * 5 + 7) * N), ctx->iUA = devAlloc<real>(c, 32 * N),
                ctx->iNbr = devAlloc<idx>(c, (size_t)m.K * N);
            real *inc = nullptr;
            launchSGSPacked64(c, gp, alphaDiag, ctx->iRhs, u, ctx->wLam, ctx->iJd, ctx->iUA, ctx->iSlot, ctx->iNbr,
                              fresh, is->sgsIter, is->gsUpdate, &inc);
            launchAddIncrementCellMajor(c, gamma, is->incrementRelax, rhoEps, pEps, u, inc, nullptr, ctx->iStat);
        }
        // fstop: residual of this iteration relative to its running maximum (density component)
        PPStat st;
        int blockBad = 0;
        DNDSR_HIP_CHECK(hipMemcpyAsync(c.hostRed, ctx->iRes, 5 * sizeof(real), hipMemcpyDeviceToHost, c.stream));
        DNDSR_HIP_CHECK(hipMemcpyAsync(&st, ctx->iStat, sizeof(PPStat), hipMemcpyDeviceToHost, c.stream));
        if (bBad)
            DNDSR_HIP_CHECK(hipMemcpyAsync(&blockBad, bBad, sizeof(int), hipMemcpyDeviceToHost, c.stream));
        DNDSR_HIP_CHECK(hipStreamSynchronize(c.stream));
        if (blockBad)
        {
            DNDSR_HIP_CHECK(hipMemsetAsync(bBad, 0, sizeof(int), c.stream));
            fail("a block of JDiag is singular or has a zero diagonal entry (iteration " + std::to_string(it) +
                 "); u and uRec are left as they were at entry");
        }
        for (int v = 0; v < 5; v++)
            if (!std::isfinite(c.hostRed[v]))
                fail("non-finite residual in the implicit step (iteration " + std::to_string(it) +
                     "); u and uRec are left as they were at entry");
        if (st.bad)
            fail("the positivity-preserving increment met a state with non-positive density or internal energy "
                 "(iteration " + std::to_string(it) + "); u and uRec are left as they were at entry");
        ctx->ppLimited += st.nLim;
        if (st.nLim)
        {
            double am;
            std::memcpy(&am, &st.alphaMinBits, sizeof(am));
            ctx->ppAlphaMin = std::min(ctx->ppAlphaMin, am);
        }
        for (int v = 0; v < 5; v++)
        {
            rr[v] = c.hostRed[v];
            resBase[v] = std::max(resBase[v], rr[v]);
            rel[v] = rr[v] / (resBase[v] + kVerySmall);
        }
        if (history)
            history[it - 1] = rel[0];
        if (it >= is->minIter && rel[0] < is->tol)
            break;
    }
    if (it > is->maxIter)
        it = is->maxIter;
    if (rec != uRec)
        launchCopy(c, nR, rec, uRec);
    if (nIter)
        *nIter = it;
    for (int v = 0; v < 5; v++)
    {
        if (res)
            res[v] = rr[v];
        if (resRel)
            resRel[v] = rel[v];
    }
}

int dndsr_implicit_pp_stats(dndsr_ctx *ctx, int64_t *nLimited, double *alphaMin)
{
    CAPI_BEGIN
    need(ctx != nullptr, "null context");
    if (nLimited)
        *nLimited = ctx->ppLimited;
    if (alphaMin)
        *alphaMin = ctx->ppAlphaMin;
    CAPI_END
}

int dndsr_lusgs_matvec(dndsr_ctx *ctx, const dndsr_euler_settings *es, double alphaDiag, const double *u,
                       const double *x, const double *jdiag, const double *lambdaFace, double *Ax)
{
    CAPI_BEGIN
    requireMesh(ctx);
    need(es && u && x && jdiag && lambdaFace && Ax, "null field");
    need(x != Ax, "x and Ax must be different arrays");
    launchLUSGSMatVec(ctx->c, toGas(es), alphaDiag, u, x, jdiag, lambdaFace, Ax);
    CAPI_END
}

int dndsr_gmres_lusgs(dndsr_ctx *ctx, const dndsr_euler_settings *es, double alphaDiag, const double *u,
                      const double *jdiag, const double *lambdaFace, const double *b, double *x, int nGmresSpace,
                      int nGmresIter, int sgsIter, int gsUpdate, int fp32Precond, double *residual)
{
    CAPI_BEGIN
    requireMesh(ctx);
    need(es && u && jdiag && lambdaFace && b && x, "null field");
    need(nGmresSpace >= 1 && nGmresIter >= 0 && sgsIter >= 0, "bad GMRES settings");
    need(gsUpdate == 0 || gsUpdate == 1, "gsUpdate must be 0 (Jacobi) or 1 (Gauss-Seidel)");
    Ctx &c = ctx->c;
    const size_t N = c.m.N, F = c.m.F, nU = 5 * N;
    LinWork &w = linWork(ctx, nGmresSpace, fp32Precond != 0);
    w.u = u, w.jd = jdiag, w.lam = lambdaFace, w.alphaDiag = alphaDiag;
    w.sgsIter = sgsIter, w.gsUpdate = gsUpdate, w.fp32 = fp32Precond != 0;
    if (w.fp32)
    {
        if (!ctx->fU)
            ctx->fU = devAlloc<float>(c, nU);
        if (!ctx->fJd)
            ctx->fJd = devAlloc<float>(c, N);
        if (!ctx->fLam)
            ctx->fLam = devAlloc<float>(c, 4 * F);
        launchToF32(c, nU, u, ctx->fU);
        launchToF32(c, N, jdiag, ctx->fJd);
        launchToF32(c, F, lambdaFace, ctx->fLam);
        w.fU = ctx->fU, w.fJd = ctx->fJd, w.fLam = ctx->fLam;
    }
    double r[2] = {0, 0};
    gmresSolve(c, toGas(es), w, b, x, nGmresSpace, nGmresIter, r);
    DNDSR_HIP_CHECK(hipStreamSynchronize(c.stream));
    if (residual)
        residual[0] = r[0], residual[1] = r[1];
    CAPI_END
}

int dndsr_eval_rhs_vjp(dndsr_ctx *ctx, const dndsr_euler_settings *es, const double *u, const double *uRec,
                       const double *deltaLambdaCell, const double *rhsBar, double *uBar, double *uRecBar)
{
    CAPI_BEGIN
    requireMesh(ctx);
    need(es && u && rhsBar && uBar, "null field");
    need(!uRecBar || uRec, "uRecBar needs uRec");
    Ctx &c = ctx->c;
    need(c.m.nBnd == 0 || c.m.bcType, "the mesh has boundary faces: set the boundary conditions first (dndsr_set_bc)");
    const int rows = es->viscous ? 50 : 20;
    if (c.adjPointRows < rows)
    {
        c.adjPoint = devAlloc<real>(c, (size_t)rows * c.m.G * c.m.F);
        c.adjPointRows = rows;
    }
    if (!c.adjFaceBar)
        c.adjFaceBar = devAlloc<real>(c, (size_t)5 * c.m.F);
    launchEvalRHSVJP(c, toGas(es), u, uRec, deltaLambdaCell, rhsBar, uBar, uRec ? uRecBar : nullptr);
    CAPI_END
}

int dndsr_zone_force(dndsr_ctx *ctx, const dndsr_euler_settings *es, const double *u, const double *uRec,
                     const double *deltaLambdaCell, int zone, double *force)
{
    CAPI_BEGIN
    requireMesh(ctx);
    need(es && u && force, "null argument");
    Ctx &c = ctx->c;
    need(zone >= 0 && zone < c.m.nBCZone, "zone out of range");
    need(c.m.bcType, "the mesh has boundary faces: set the boundary conditions first (dndsr_set_bc)");
    if (!c.adjRhs)
        c.adjRhs = devAlloc<real>(c, (size_t)5 * c.m.N);
    launchEvalRHS(c, toGas(es), u, uRec, deltaLambdaCell, c.adjRhs);
    real *dev = c.red + 5 * ((size_t)gridFor(c.m.N, 256) + c.m.nColor + 5); // three values
    launchZoneForce(c, zone, dev);
    DNDSR_HIP_CHECK(hipMemcpyAsync(c.hostRed, dev, 3 * sizeof(real), hipMemcpyDeviceToHost, c.stream));
    DNDSR_HIP_CHECK(hipStreamSynchronize(c.stream));
    for (int d = 0; d < 3; d++)
        force[d] = c.hostRed[d];
    CAPI_END
}

int dndsr_zone_force_vjp(dndsr_ctx *ctx, const dndsr_euler_settings *es, const double *u, const double *uRec,
                         const double *deltaLambdaCell, int zone, const double *forceBar, double *uBar,
                         double *uRecBar)
{
    CAPI_BEGIN
    requireMesh(ctx);
    need(es && u && forceBar && uBar, "null argument");
    need(!uRecBar || uRec, "uRecBar needs uRec");
    Ctx &c = ctx->c;
    need(zone >= 0 && zone < c.m.nBCZone, "zone out of range");
    need(c.m.bcType, "the mesh has boundary faces: set the boundary conditions first (dndsr_set_bc)");
    const int rows = es->viscous ? 50 : 20;
    if (c.adjPointRows < rows)
    {
        c.adjPoint = devAlloc<real>(c, (size_t)rows * c.m.G * c.m.F);
        c.adjPointRows = rows;
    }
    if (!c.adjFaceBar)
        c.adjFaceBar = devAlloc<real>(c, (size_t)5 * c.m.F);
    for (int d = 0; d < 3; d++)
        c.hostRed[d] = forceBar[d];
    real *dev = c.red + 5 * ((size_t)gridFor(c.m.N, 256) + c.m.nColor + 5);
    DNDSR_HIP_CHECK(hipMemcpyAsync(dev, c.hostRed, 3 * sizeof(real), hipMemcpyHostToDevice, c.stream));
    launchZoneForceBar(c, zone, dev, c.adjFaceBar);
    launchEvalRHSVJP(c, toGas(es), u, uRec, deltaLambdaCell, nullptr, uBar, uRec ? uRecBar : nullptr, c.adjFaceBar);
    DNDSR_HIP_CHECK(hipStreamSynchronize(c.stream));
    CAPI_END
}

int dndsr_vr_iter_vjp(dndsr_ctx *ctx, int nVars, int mode, double relax, const double *uRec, const double *u,
                      const double *uRecNewBar, double *uRecBar, double *uBar)
{
    CAPI_BEGIN
    requireMesh(ctx);
    need(uRecNewBar != nullptr, "null field");
    need(mode >= 0 && mode <= 3, "mode must be 0, 1, 2 or 3");
    need(mode == 3 || uBar != nullptr, "uBar is required for modes 0-2");
    Ctx &c = ctx->c;
    launchVRIterVJP(c, nVars, mode, relax < 0 ? c.m.s.jacobiRelax : relax, uRec, u, uRecNewBar, uRecBar,
                    mode == 3 ? nullptr : uBar);
    CAPI_END
}

} // extern "C"
