// dndsr_amd.hpp -- C++17 host mirror of the reference's operator API over the C-ABI.
//
// The reference exposes this path as member templates of
//   DNDS::CFV::VariationalReconstruction<dim>   (src/CFV/VariationalReconstruction.hpp:108)
//   DNDS::Euler::EulerEvaluator<model>           (src/Euler/EulerEvaluator.hpp:76)
// operating on caller-owned ArrayDOFV / ArrayRECV.  This header keeps those class and
// member names and their argument meaning, holds the fields on the device (structure of
// arrays, cell index fastest) and calls include/dndsr_amd.h.  Errors throw
// std::runtime_error, as DNDS_check_throw does (src/DNDS/Errors.hpp:93-112).
//
// Link with libdndsr_amd.so and the HIP runtime (amdhip64).
#pragma once
#include <hip/hip_runtime_api.h>

#include <array>
#include <cmath>
#include <cstdint>
#include <functional>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "dndsr_amd.h"

namespace dndsr_amd
{
using real = double;
using index = int64_t;

inline void check(int rc)
{
    if (rc != 0)
        throw std::runtime_error(std::string("dndsr_amd: ") + dndsr_last_error());
}
inline void checkHip(hipError_t e, const char *what)
{
    if (e != hipSuccess)
        throw std::runtime_error(std::string("dndsr_amd: ") + what + ": " + hipGetErrorString(e));
}

// Owned device buffer.
template <class T>
class DeviceBuffer
{
  public:
    DeviceBuffer() = default;
    explicit DeviceBuffer(size_t n) : n_(n)
    {
        if (n)
            checkHip(hipMalloc(reinterpret_cast<void **>(&p_), n * sizeof(T)), "hipMalloc");
    }
    DeviceBuffer(const DeviceBuffer &) = delete;
    DeviceBuffer &operator=(const DeviceBuffer &) = delete;
    DeviceBuffer(DeviceBuffer &&o) noexcept : p_(std::exchange(o.p_, nullptr)), n_(std::exchange(o.n_, 0)) {}
    DeviceBuffer &operator=(DeviceBuffer &&o) noexcept
    {
        std::swap(p_, o.p_);
        std::swap(n_, o.n_);
        return *this;
    }
    ~DeviceBuffer()
    {
        if (p_)
            (void)hipFree(p_);
    }
    T *data() { return p_; }
    const T *data() const { return p_; }
    size_t size() const { return n_; }
    void zero() { checkHip(hipMemset(p_, 0, n_ * sizeof(T)), "hipMemset"); }

  private:
    T *p_ = nullptr;
    size_t n_ = 0;
};

// One device context (one stream).
class Context
{
  public:
    explicit Context(int device = 0, void *stream = nullptr)
    {
        check(dndsr_create(&h_, device));
        if (stream)
            check(dndsr_set_stream(h_, stream));
    }
    Context(const Context &) = delete;
    Context &operator=(const Context &) = delete;
    ~Context()
    {
        if (h_)
            dndsr_destroy(h_);
    }
    dndsr_ctx *handle() const { return h_; }
    void sync() const { check(dndsr_sync(h_)); }

  private:
    dndsr_ctx *h_ = nullptr;
};

// Cell field: nComp values per cell (ArrayDOFV<nVars>; ArrayRECV<nVars> with nComp = nVars * M).
// Host I/O uses the reference's row layout (rows of nComp, an uRec row is the column-major
// M x nVars block).
class CellField
{
  public:
    CellField() = default;
    CellField(const Context &ctx, index nCell, int nVars, int M = 1)
        : ctx_(&ctx), nCell_(nCell), nVars_(nVars), M_(M), buf_((size_t)nCell * nVars * M)
    {
        buf_.zero();
    }
    double *data() { return buf_.data(); }
    const double *data() const { return buf_.data(); }
    index nCell() const { return nCell_; }
    int nVars() const { return nVars_; }
    int M() const { return M_; }
    int nComp() const { return nVars_ * M_; }
    void setConstant(double v)
    {
        if (v != 0.0)
        {
            std::vector<double> h((size_t)nCell_ * nComp(), v);
            fromHostRows(h);
        }
        else
            buf_.zero();
    }
    void fromHostRows(const std::vector<double> &rows)
    {
        if (rows.size() != (size_t)nCell_ * nComp())
            throw std::runtime_error("dndsr_amd: host row array has the wrong size");
        check(dndsr_h2d_cellfield(ctx_->handle(), rows.data(), buf_.data(), nComp()));
    }
    std::vector<double> toHostRows() const
    {
        std::vector<double> rows((size_t)nCell_ * nComp());
        check(dndsr_d2h_cellfield(ctx_->handle(), buf_.data(), rows.data(), nComp()));
        return rows;
    }

  private:
    const Context *ctx_ = nullptr;
    index nCell_ = 0;
    int nVars_ = 0, M_ = 1;
    DeviceBuffer<double> buf_;
};
using ArrayDOF = CellField;
using ArrayREC = CellField;

// Host-side mesh description (UnstructuredMesh's CSR adjacency, src/Geom/Mesh/Mesh.hpp).
struct Mesh
{
    int dim = 2;
    std::vector<double> coords; // [nNode][3]
    std::vector<int32_t> cellType, faceType;
    std::vector<int64_t> cell2nodePtr, cell2node, cell2facePtr, cell2face, face2nodePtr, face2node;
    std::vector<int64_t> face2cell; // [nFace][2], -1 = boundary
    std::vector<double> faceShift;  // [nFace][3]
    std::vector<int32_t> faceZone;  // [nFace]: zone of each boundary face (empty: no boundary face)
    int32_t nBCZone = 0;
    dndsr_mesh_desc desc() const
    {
        dndsr_mesh_desc d{};
        d.dim = dim;
        d.nNode = (int64_t)(coords.size() / 3);
        d.coords = coords.data();
        d.nCell = (int64_t)cellType.size();
        d.cellType = cellType.data();
        d.cell2nodePtr = cell2nodePtr.data(), d.cell2node = cell2node.data();
        d.cell2facePtr = cell2facePtr.data(), d.cell2face = cell2face.data();
        d.nFace = (int64_t)faceType.size();
        d.faceType = faceType.data();
        d.face2nodePtr = face2nodePtr.data(), d.face2node = face2node.data();
        d.face2cell = face2cell.data();
        d.faceShift = faceShift.data();
        d.faceZone = faceZone.empty() ? nullptr : faceZone.data();
        d.nBCZone = nBCZone;
        return d;
    }
};

// VRSettings subset (src/CFV/VRSettings.hpp:37-212) with the reference's defaults.
struct VRSettings
{
    int maxOrder = 1, intOrder = 5;
    int dirWeightScheme = 0, geomWeightScheme = 0; // Factorial, GWNone
    double geomWeightBias = 0, geomWeightPower = 0.5, scaleMultiplier = 1, jacobiRelax = 1;
    int nIterCellSmoothScale = 15;
    double bcWeight = 1;
    std::vector<double> zoneDirWeight; // [nBCZone][maxOrder+1]; empty = 1
    bool SORInstead = false;           // multicolour in-place sweeps (the reference defaults to true)
    int subs2ndOrder = 0, subs2ndOrderGGScheme = 1;
    dndsr_vr_settings c() const
    {
        return {maxOrder, intOrder, dirWeightScheme, geomWeightScheme, geomWeightBias, geomWeightPower,
                scaleMultiplier, jacobiRelax, nIterCellSmoothScale, bcWeight,
                zoneDirWeight.empty() ? nullptr : zoneDirWeight.data(), SORInstead ? 1 : 0, subs2ndOrder,
                subs2ndOrderGGScheme};
    }
};

template <int dim>
class VariationalReconstruction
{
    static_assert(dim == 2 || dim == 3, "dim must be 2 or 3");

  public:
    VariationalReconstruction(Context &ctx, const Mesh &mesh, const VRSettings &s) : ctx_(ctx), mesh_(mesh), s_(s)
    {
        if (mesh.dim != dim)
            throw std::runtime_error("dndsr_amd: mesh dimension does not match VariationalReconstruction<dim>");
    }
    // ConstructMetrics / ConstructBaseAndWeight / ConstructRecCoeff (VariationalReconstruction.hpp:282-289):
    // the device builds all three at ConstructRecCoeff (or at the first use).
    // id2faceDircWeight(zone, order) multiplies the weights of boundary faces, as the
    // reference's tFGetBoundaryWeight does (VariationalReconstruction.cpp:438-441).
    void ConstructMetrics() {}
    void ConstructBaseAndWeight(const std::function<double(int, int)> &id2faceDircWeight = nullptr)
    {
        if (!id2faceDircWeight)
            return;
        std::vector<double> w((size_t)mesh_.nBCZone * (s_.maxOrder + 1));
        for (int z = 0; z < mesh_.nBCZone; z++)
            for (int p = 0; p <= s_.maxOrder; p++)
                w[(size_t)z * (s_.maxOrder + 1) + p] = id2faceDircWeight(z, p);
        if (constructed_ && w != s_.zoneDirWeight)
            throw std::runtime_error("dndsr_amd: boundary weights must be set before ConstructRecCoeff");
        s_.zoneDirWeight = w;
    }
    void ConstructRecCoeff() { construct(); }
    bool Constructed() const { return constructed_; }

    index NumCell() { construct(); return nCell_; }
    index NumFace() { construct(); return nFace_; }
    int NumRecModes() { construct(); return M_; }
    int NumBCZone() const { return mesh_.nBCZone; }
    int MaxOrder() const { return s_.maxOrder; }
    Context &context() const { return ctx_; }

    void BuildUDof(ArrayDOF &u, int nVars) { construct(); u = ArrayDOF(ctx_, nCell_, nVars, 1); }
    void BuildURec(ArrayREC &uRec, int nVars) { construct(); uRec = ArrayREC(ctx_, nCell_, nVars, M_); }

    // DoReconstructionIter (VariationalReconstruction.hpp:984-992): putIntoNew = false leaves
    // the new iterate in uRec, true in uRecNew; recordInc writes -residual to uRecNew.  With
    // SORInstead (and no recordInc) uRec is swept in place and, for putIntoNew, the two arrays
    // are swapped as the reference does (:602-606).
    void DoReconstructionIter(ArrayREC &uRec, ArrayREC &uRecNew, ArrayDOF &u, bool putIntoNew = false,
                              bool recordInc = false, bool uRecIsZero = false, double relax = -1)
    {
        if (recordInc && !putIntoNew)
            throw std::runtime_error("dndsr_amd: recordInc needs putIntoNew");
        if (s_.SORInstead && !recordInc)
        {
            check(dndsr_vr_iter(ctx_.handle(), u.nVars(), uRec.data(), nullptr, u.data(), 4, relax));
            if (putIntoNew)
                std::swap(uRec, uRecNew);
            return;
        }
        const int mode = !recordInc ? 0 : (uRecIsZero ? 2 : 1);
        check(dndsr_vr_iter(ctx_.handle(), u.nVars(), uRec.data(), uRecNew.data(), u.data(), mode, relax));
        if (!putIntoNew)
            std::swap(uRec, uRecNew);
    }
    // DoReconstructionIterDiff(uRec, uRecDiff, uRecNew, u) (VariationalReconstruction.hpp:1012-1018);
    // uRec and u feed the boundary term and may be empty on meshes without boundary faces
    void DoReconstructionIterDiff(const ArrayREC &uRec, const ArrayREC &uRecDiff, ArrayREC &uRecNew, const ArrayDOF &u)
    {
        check(dndsr_vr_iter_diff(ctx_.handle(), uRecDiff.nVars(), uRec.data(), u.data(), uRecDiff.data(),
                                 uRecNew.data()));
    }

  private:
    void construct()
    {
        if (constructed_)
            return;
        const dndsr_mesh_desc d = mesh_.desc();
        const dndsr_vr_settings s = s_.c();
        check(dndsr_vr_construct(ctx_.handle(), &d, &s));
        int64_t nC, nF;
        int K, G, Gc, nCol;
        double vol;
        check(dndsr_mesh_info(ctx_.handle(), &nC, &nF, &M_, &K, &G, &Gc, &nCol, &vol));
        nCell_ = nC, nFace_ = nF;
        constructed_ = true;
    }
    Context &ctx_;
    const Mesh &mesh_;
    VRSettings s_;
    bool constructed_ = false;
    index nCell_ = 0, nFace_ = 0;
    int M_ = 0;
};

// EulerModel subset of this path (src/Euler/Euler.hpp): NS (2-D mesh, 3 momentum
// components) and NS_3D, both with 5 variables.
enum EulerModel
{
    NS = 0,
    NS_3D = 2,
};
template <EulerModel model>
constexpr int modelDim = model == NS_3D ? 3 : 2;

// EulerEvaluatorSettings subset (src/Euler/EulerEvaluatorSettings.hpp:53-80)
struct EulerEvaluatorSettings
{
    double gamma = 1.4;
    int riemannSolverType = 1; // Gas::RiemannSolverType numbering (Roe = 1, HLLC = 2, HLLEP = 3, Roe_M1.. = 11..)
    int rsMeanValueEig = 0;
    double rsFixScale = 1, rsIncFScale = 1;
    // boundary zones (BoundaryHandler, src/Euler/EulerBC.hpp): EulerBCType code per zone and
    // its 5 conservative values (far field / inflow); empty on periodic meshes
    std::vector<int32_t> bcType;
    std::vector<double> bcValue; // [nZone][5]
    // laminar viscous terms (false: the RHS_Ignore_Viscosity residual, EvaluateRHS.hxx:78)
    // and IdealGasProperty (src/Euler/EulerEvaluatorSettings.hpp:291-300)
    bool viscous = false;
    bool useRoeJacobian = false; // EulerEvaluatorSettings.hpp:54 (block LU-SGS in the implicit step)
    struct IdealGasProperty
    {
        double Rgas = 1, muGas = 1, prGas = 0.72, TRef = 273.15, CSutherland = 110.4;
        int muModel = 1; // 0 constant, 1 Sutherland, 2 constant nu
    } idealGasProperty;
    dndsr_euler_settings c() const
    {
        const IdealGasProperty &g = idealGasProperty;
        return {gamma,   riemannSolverType, rsMeanValueEig, rsFixScale, rsIncFScale, viscous ? 1 : 0,
                g.muModel, g.Rgas,          g.muGas,        g.prGas,    g.TRef,      g.CSutherland,
                useRoeJacobian ? 1 : 0};
    }
};

template <EulerModel model>
class EulerEvaluator
{
  public:
    static constexpr int dim = modelDim<model>;
    static constexpr int nVars = 5;

    EulerEvaluator(VariationalReconstruction<dim> &vfv, const EulerEvaluatorSettings &s)
        : settings(s), vfv_(vfv), es_(s.c())
    {
        if (!s.bcType.empty())
        {
            if ((int)s.bcType.size() != vfv.NumBCZone())
                throw std::runtime_error("dndsr_amd: one boundary type per mesh zone is needed");
            // boundary weights of the VR functional by boundary type (EulerEvaluator.hpp:138-152):
            // value only (Dirichlet type) except BCOut, which takes none
            const std::vector<int32_t> types = s.bcType;
            vfv.ConstructBaseAndWeight([types](int zone, int order) -> double {
                if (types[zone] == 5 /* BCOut */)
                    return 0.;
                return order ? 0. : 1.;
            });
        }
        vfv.ConstructRecCoeff();
        if (!s.bcType.empty())
            check(dndsr_set_bc_gas(vfv.context().handle(), (int)s.bcType.size(), s.bcType.data(),
                                   s.bcValue.empty() ? nullptr : s.bcValue.data(), s.gamma, s.idealGasProperty.Rgas));
        lambdaFace = DeviceBuffer<double>(4 * (size_t)vfv.NumFace());
        lambdaFace.zero();
        deltaLambdaCell = DeviceBuffer<double>((size_t)vfv.NumCell());
        deltaLambdaCell.zero();
    }

    // EvaluateDt(dt, u, uRec, CFL, dtMinall, MaxDt, UseLocaldt, t) (EulerEvaluator.hpp:359-366)
    void EvaluateDt(ArrayDOF &dt, ArrayDOF &u, ArrayREC & /*uRec*/, double CFL, double &dtMinall, double MaxDt = 1e100,
                    bool UseLocaldt = false, double /*t*/ = 0)
    {
        check(dndsr_eval_dt(vfv_.context().handle(), &es_, u.data(), CFL, MaxDt, UseLocaldt ? 1 : 0, dt.data(),
                            lambdaFace.data(), deltaLambdaCell.data(), nullptr, &dtMinall));
    }
    // EvaluateRHS(rhs, JSource, u, uRecUnlim, uRec, ...) (EulerEvaluator.hpp:399-409), inviscid
    void EvaluateRHS(ArrayDOF &rhs, ArrayDOF &u, ArrayREC &uRec, double /*t*/ = 0)
    {
        check(dndsr_eval_rhs(vfv_.context().handle(), &es_, u.data(), uRec.data(), deltaLambdaCell.data(),
                             rhs.data()));
    }
    // LUSGSMatrixInit(JDiag, JSource, dTau, dt, alphaDiag, u, uRec, jacobianCode, t) (EulerEvaluator.hpp:427-434)
    void LUSGSMatrixInit(ArrayDOF &JDiag, ArrayDOF &dTau, double dt, double alphaDiag, ArrayDOF & /*u*/,
                         ArrayREC & /*uRec*/, int jacobianCode = 0, double /*t*/ = 0)
    {
        if (jacobianCode != 0)
            throw std::runtime_error("dndsr_amd: only the scalar Jacobian (jacobianCode 0) is on this path");
        check(dndsr_lusgs_init(vfv_.context().handle(), dTau.data(), dt, alphaDiag, lambdaFace.data(), JDiag.data()));
    }
    // UpdateSGS(alphaDiag, t, rhs, u, uInc, uIncNew, JDiag, forward, gsUpdate, sumInc, uIncIsZero)
    // (EulerEvaluator.hpp:513-522), multicolour Gauss-Seidel order
    void UpdateSGS(double alphaDiag, double /*t*/, ArrayDOF &rhs, ArrayDOF &u, ArrayDOF &uInc, ArrayDOF &uIncNew,
                   ArrayDOF &JDiag, bool forward, bool gsUpdate, std::array<double, 5> &sumInc,
                   bool uIncIsZero = false)
    {
        check(dndsr_update_sgs(vfv_.context().handle(), &es_, alphaDiag, rhs.data(), u.data(), uInc.data(),
                               uIncNew.data(), JDiag.data(), lambdaFace.data(), forward ? 1 : 0, gsUpdate ? 1 : 0,
                               uIncIsZero ? 1 : 0, sumInc.data()));
    }
    // explicit SSP-RK3 steps kept on the device (ODE.hpp:1477-1520)
    void SSPRK3(ArrayDOF &u, ArrayREC &uRec, int nStep, double CFL, int nRec = 1, double dtFixed = 0)
    {
        check(dndsr_ssprk3(vfv_.context().handle(), &es_, u.data(), uRec.data(), nStep, CFL, nRec, dtFixed));
    }
    // FixUMaxFilter(u) (EulerEvaluator.hxx:1318-1323): the reference returns at once
    void FixUMaxFilter(ArrayDOF &) const {}
    // EvaluateNorm(res, rhs, P, volWise = false) (EulerEvaluator.hpp:613), P = 1 or 2
    std::vector<double> EvaluateNorm(const ArrayDOF &rhs, int P = 1) const
    {
        std::vector<double> rows = rhs.toHostRows(), res(rhs.nVars(), 0.0);
        for (size_t i = 0; i < rows.size(); i++)
        {
            const double a = rows[i] < 0 ? -rows[i] : rows[i];
            res[i % rhs.nVars()] += P == 1 ? a : a * a;
        }
        if (P == 2)
            for (auto &r : res)
                r = std::sqrt(r);
        return res;
    }

    DeviceBuffer<double> lambdaFace;      // [4][nFace] (lambdaFace, lambdaFaceC, lambdaFaceVis, lambdaFace0)
    DeviceBuffer<double> deltaLambdaCell; // [nCell]
    EulerEvaluatorSettings settings;

  private:
    VariationalReconstruction<dim> &vfv_;
    dndsr_euler_settings es_;
};

} // namespace dndsr_amd
