// dndsr_oracle.cpp -- CPU oracle for the compact-FV residual / reconstruction / LU-SGS path.
//
// TEST INFRASTRUCTURE ONLY: a plain C++17 FP64 restatement of the reference
// (CFDLAB-THU/DNDSR) algorithms that the HIP path is checked against.  It is
// loaded only by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
// leg; the product library never uses it.  Each function cites the reference
// file:line it restates.  Loops keep the reference's order (faces in index
// order, quadrature points in rule order, cell2face in element order) so that
// summation order matches the reference's serial build.
//
// Pinning (tests/test_oracle_*.py):
//   Riemann fluxes      : reference golden vectors test/cpp/Euler/test_RiemannSolvers.cpp:257-267,
//                         consistency / symmetry / supersonic-upwind closed forms.
//   metrics, VR setup,  : reference VR error goldens test/cpp/CFV/test_Reconstruction.cpp:531-545
//   VR iteration          and test_Reconstruction3D.cpp:393-409 (uniform periodic grids),
//                         constant / linear exactness, dense brute-force fixed point.
//   EvaluateDt, RHS(p0),: reference pipeline golden test/cpp/Euler/test_EulerEvaluator.cpp:231-259
//   LUSGSMatrixInit       (IV 10x10, rho-residual L1 and increment L2).
//   RHS(high order)     : conservation (sum V rhs = 0), free-stream preservation, and the
//                         p0 limit that the pipeline golden pins.
//   UpdateSGS (GS)      : dense (D+L+U) brute force on a tiny mesh, FD check of the flux
//                         Jacobian increment.
//   SSP-RK3             : restated from src/Solver/ODE.hpp:1477-1520; pinned by the RHS pins
//                         plus the closed-form RK3 combination (parity unpinned beyond that).
#include "dndsr_oracle.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

namespace orc
{
using real = double;
using idx_t = int64_t;
static constexpr real verySmallReal = 1e-200; // src/DNDS/Defines.hpp:199
static constexpr real smallReal = 1e-10;      // src/DNDS/Defines.hpp:201
static constexpr real pi = 3.14159265358979323846;

static thread_local std::string g_err;
static int g_threads = 1;     // OpenMP threads used by the parallel loops
static int g_rhsGather = 0;   // 0: serial face-order scatter (EvaluateRHS.hxx:550-562);
                              // 1: face buffer + cell gather (the OpenMP variant, :153-159, 640-662)

struct Pt
{
    real x[3]{0, 0, 0};
};

// ---------------------------------------------------------------------------
// Quadrature.  Gauss-Legendre tables exactly as the reference stores them
// (15 significant digits), src/Geom/Quadratures/Line.hpp:25-42; tensor-product
// point order src/Geom/QuadratureHub.hpp:318-373 (first parametric index fastest);
// order -> scheme map src/Geom/QuadratureHub.hpp:39-96.
// ---------------------------------------------------------------------------
static const real GL1p[1] = {0}, GL1w[1] = {2};
static const real GL2p[2] = {-0.577350269189626, 0.577350269189626}, GL2w[2] = {1, 1};
static const real GL3p[3] = {-0.774596669241483, 0, 0.774596669241483},
                  GL3w[3] = {0.555555555555555, 0.888888888888889, 0.555555555555555};
static const real GL4p[4] = {-0.861136311594053, -0.339981043584856, 0.339981043584856, 0.861136311594053},
                  GL4w[4] = {0.347854845137454, 0.652145154862546, 0.652145154862546, 0.347854845137454};

static int lineNPoint(int order)
{
    if (order <= 1)
        return 1;
    if (order <= 3)
        return 2;
    if (order <= 5)
        return 3;
    if (order == 6)
        return 4;
    throw std::runtime_error("unsupported integration order");
}
static void glTable(int n, const real *&p, const real *&w)
{
    switch (n)
    {
    case 1: p = GL1p, w = GL1w; break;
    case 2: p = GL2p, w = GL2w; break;
    case 3: p = GL3p, w = GL3w; break;
    case 4: p = GL4p, w = GL4w; break;
    default: throw std::runtime_error("bad GL size");
    }
}
struct QP
{
    real p[3];
    real w;
};

// Hammer rules on the reference triangle (0,0)-(1,0)-(0,1) and tetrahedron
// (0,0,0)-(1,0,0)-(0,1,0)-(0,0,1), digits as the reference stores them
// (src/Geom/Quadratures/Tri.hpp:25-66, Tet.hpp:27-92); rows xi, eta, (zeta,) w.
static const real kTri1[3][1] = {{1. / 3.}, {1. / 3.}, {1. / 2.}};
static const real kTri3[3][3] = {{2. / 3., 1. / 6., 1. / 6.}, {1. / 6., 2. / 3., 1. / 6.}, {1. / 6., 1. / 6., 1. / 6.}};
static const real kTri6[3][6] = {
    {0.09157621350977102, 0.8168475729804585, 0.09157621350977052, 0.445948490915965, 0.10810301816807, 0.445948490915965},
    {0.09157621350977102, 0.09157621350977052, 0.8168475729804585, 0.445948490915965, 0.445948490915965, 0.10810301816807},
    {0.054975871827661, 0.054975871827661, 0.054975871827661, 0.1116907948390057, 0.1116907948390057, 0.1116907948390057}};
static const real kTri7[3][7] = {
    {0.3333333333333335, 0.470142064105115, 0.05971587178977, 0.470142064105115, 0.1012865073234565, 0.7974269853530875,
     0.1012865073234565},
    {0.3333333333333335, 0.470142064105115, 0.470142064105115, 0.05971587178977, 0.1012865073234565, 0.1012865073234565,
     0.7974269853530875},
    {0.1125, 0.066197076394253, 0.066197076394253, 0.066197076394253, 0.0629695902724135, 0.0629695902724135,
     0.0629695902724135}};
static const real kTri12[3][12] = {
    {0.2492867451709105, 0.501426509658179, 0.2492867451709105, 0.063089014491502, 0.8738219710169954, 0.063089014491502,
     0.3103524510337845, 0.05314504984481699, 0.6365024991213986, 0.05314504984481699, 0.6365024991213986,
     0.3103524510337845},
    {0.2492867451709105, 0.2492867451709105, 0.501426509658179, 0.063089014491502, 0.063089014491502, 0.8738219710169954,
     0.05314504984481699, 0.3103524510337845, 0.05314504984481699, 0.6365024991213986, 0.3103524510337845,
     0.6365024991213986},
    {0.05839313786318975, 0.05839313786318975, 0.05839313786318975, 0.0254224531851035, 0.0254224531851035,
     0.0254224531851035, 0.04142553780918675, 0.04142553780918675, 0.04142553780918675, 0.04142553780918675,
     0.04142553780918675, 0.04142553780918675}};
static const real kTet1[4][1] = {{0.25}, {0.25}, {0.25}, {0.1666666666666667}};
static const real kTet4[4][4] = {{0.1381966011250105, 0.1381966011250105, 0.5854101966249685, 0.1381966011250105},
                                 {0.1381966011250105, 0.5854101966249684, 0.1381966011250105, 0.1381966011250105},
                                 {0.5854101966249684, 0.1381966011250105, 0.1381966011250105, 0.1381966011250105},
                                 {0.04166666666666666, 0.04166666666666666, 0.04166666666666666, 0.04166666666666666}};
static const real kTet8[4][8] = {
    {0.1069940147705369, 0.1069940147705369, 0.6790179556883893, 0.1069940147705369, 0.3280585716625127,
     0.3280585716625127, 0.01582428501246191, 0.3280585716625127},
    {0.1069940147705369, 0.6790179556883893, 0.1069940147705369, 0.1069940147705369, 0.3280585716625127,
     0.01582428501246191, 0.3280585716625127, 0.3280585716625127},
    {0.6790179556883893, 0.1069940147705369, 0.1069940147705369, 0.1069940147705369, 0.01582428501246191,
     0.3280585716625127, 0.3280585716625127, 0.3280585716625127},
    {0.01859314997209119, 0.01859314997209119, 0.01859314997209119, 0.01859314997209119, 0.02307351669457548,
     0.02307351669457548, 0.02307351669457548, 0.02307351669457548}};
static const real kTet14[4][14] = {
    {0.09273525031089125, 0.09273525031089125, 0.7217942490673263, 0.09273525031089125, 0.3108859192633008,
     0.3108859192633008, 0.06734224221009777, 0.3108859192633008, 0.04550370412564952, 0.4544962958743505,
     0.04550370412564952, 0.04550370412564952, 0.4544962958743505, 0.4544962958743505},
    {0.09273525031089125, 0.7217942490673263, 0.09273525031089125, 0.09273525031089125, 0.3108859192633008,
     0.06734224221009777, 0.3108859192633008, 0.3108859192633008, 0.4544962958743505, 0.04550370412564952,
     0.04550370412564952, 0.4544962958743505, 0.04550370412564952, 0.4544962958743505},
    {0.7217942490673263, 0.09273525031089125, 0.09273525031089125, 0.09273525031089125, 0.06734224221009777,
     0.3108859192633008, 0.3108859192633008, 0.3108859192633008, 0.4544962958743505, 0.4544962958743505,
     0.4544962958743505, 0.04550370412564952, 0.04550370412564952, 0.04550370412564952},
    {0.01224884051939367, 0.01224884051939367, 0.01224884051939367, 0.01224884051939367, 0.01878132095300264,
     0.01878132095300264, 0.01878132095300264, 0.01878132095300264, 0.0070910034628469, 0.0070910034628469,
     0.0070910034628469, 0.0070910034628469, 0.0070910034628469, 0.0070910034628469}};
static const real kTet24[4][24] = {
    {0.2146028712591523, 0.2146028712591523, 0.3561913862225431, 0.2146028712591523, 0.3223378901422754,
     0.3223378901422754, 0.0329863295731736, 0.3223378901422754, 0.04067395853461131, 0.04067395853461131,
     0.8779781243961661, 0.04067395853461131, 0.6030056647916491, 0.6030056647916491, 0.06366100187501755,
     0.2696723314583158, 0.06366100187501755, 0.06366100187501755, 0.2696723314583158, 0.06366100187501755,
     0.06366100187501755, 0.06366100187501755, 0.2696723314583158, 0.6030056647916491},
    {0.2146028712591523, 0.3561913862225431, 0.2146028712591523, 0.2146028712591523, 0.3223378901422754,
     0.0329863295731736, 0.3223378901422754, 0.3223378901422754, 0.04067395853461131, 0.8779781243961661,
     0.04067395853461131, 0.04067395853461131, 0.06366100187501755, 0.06366100187501755, 0.06366100187501755,
     0.6030056647916491, 0.2696723314583158, 0.6030056647916491, 0.06366100187501755, 0.2696723314583158,
     0.06366100187501755, 0.6030056647916491, 0.06366100187501755, 0.2696723314583158},
    {0.3561913862225431, 0.2146028712591523, 0.2146028712591523, 0.2146028712591523, 0.0329863295731736,
     0.3223378901422754, 0.3223378901422754, 0.3223378901422754, 0.8779781243961661, 0.04067395853461131,
     0.04067395853461131, 0.04067395853461131, 0.2696723314583158, 0.06366100187501755, 0.6030056647916491,
     0.06366100187501755, 0.6030056647916491, 0.06366100187501755, 0.6030056647916491, 0.06366100187501755,
     0.2696723314583158, 0.2696723314583158, 0.06366100187501755, 0.06366100187501755},
    {0.006653791709694545, 0.006653791709694545, 0.006653791709694545, 0.006653791709694545, 0.009226196923942479,
     0.009226196923942479, 0.009226196923942479, 0.009226196923942479, 0.001679535175886773, 0.001679535175886773,
     0.001679535175886773, 0.001679535175886773, 0.008035714285714285, 0.008035714285714285, 0.008035714285714285,
     0.008035714285714285, 0.008035714285714285, 0.008035714285714285, 0.008035714285714285, 0.008035714285714285,
     0.008035714285714285, 0.008035714285714285, 0.008035714285714285, 0.008035714285714285}};

// Hammer rule for an integration order (GetQuadratureScheme, src/Geom/QuadratureHub.hpp:57-108)
template <size_t R, size_t n>
static std::vector<QP> hammer(const real (&t)[R][n])
{
    std::vector<QP> r(n);
    for (size_t i = 0; i < n; i++)
    {
        r[i] = QP{{t[0][i], t[1][i], R == 4 ? t[2][i] : 0.0}, t[R - 1][i]};
    }
    return r;
}
static std::vector<QP> simplexRule(int elemType, int order)
{
    if (elemType == ORC_TRI3)
    {
        if (order <= 1) return hammer(kTri1);
        if (order == 2) return hammer(kTri3);
        if (order <= 4) return hammer(kTri6);
        if (order == 5) return hammer(kTri7);
        if (order == 6) return hammer(kTri12);
    }
    else
    {
        if (order <= 1) return hammer(kTet1);
        if (order == 2) return hammer(kTet4);
        if (order == 3) return hammer(kTet8);
        if (order <= 5) return hammer(kTet14);
        if (order == 6) return hammer(kTet24);
    }
    throw std::runtime_error("unsupported integration order");
}

static int elemParamDim(int t)
{
    switch (t)
    {
    case ORC_LINE2: return 1;
    case ORC_QUAD4: return 2;
    case ORC_TRI3: return 2;
    case ORC_TET4: return 3;
    case ORC_HEX8: return 3;
    default: throw std::runtime_error("element type not supported by the oracle");
    }
}
static std::vector<QP> quadRule(int elemType, int order)
{
    if (elemType == ORC_TRI3 || elemType == ORC_TET4)
        return simplexRule(elemType, order);
    int pd = elemParamDim(elemType);
    int n = lineNPoint(order);
    const real *p, *w;
    glTable(n, p, w);
    std::vector<QP> r;
    int tot = pd == 1 ? n : (pd == 2 ? n * n : n * n * n);
    for (int iG = 0; iG < tot; iG++)
    {
        QP q{{0, 0, 0}, 1};
        int i = iG % n, j = (iG / n) % n, k = iG / (n * n);
        q.p[0] = p[i];
        q.w = w[i];
        if (pd >= 2)
            q.p[1] = p[j], q.w = w[i] * w[j];
        if (pd == 3)
            q.p[2] = p[k], q.w = w[i] * w[j] * w[k];
        r.push_back(q);
    }
    return r;
}
static const int INT_ORDER_MAX = 6; // src/Geom/Quadratures/QuadratureBase.hpp:70

// ---------------------------------------------------------------------------
// Linear shape functions (Line2, Quad4, Hex8, Tri3, Tet4; CGNS node order;
// simplices: src/Geom/Elements/Tri3.hpp:20-50, Tet4.hpp:20-56).
// ---------------------------------------------------------------------------
static int elemNNode(int t)
{
    switch (t)
    {
    case ORC_LINE2: return 2;
    case ORC_QUAD4: return 4;
    case ORC_HEX8: return 8;
    case ORC_TRI3: return 3;
    case ORC_TET4: return 4;
    default: throw std::runtime_error("bad element");
    }
}
static void shape(int t, const real *xi, real *N, real (*dN)[3])
{
    static const real q4[4][2] = {{-1, -1}, {1, -1}, {1, 1}, {-1, 1}};
    static const real h8[8][3] = {{-1, -1, -1}, {1, -1, -1}, {1, 1, -1}, {-1, 1, -1},
                                  {-1, -1, 1}, {1, -1, 1}, {1, 1, 1}, {-1, 1, 1}};
    switch (t)
    {
    case ORC_LINE2:
        N[0] = 0.5 * (1 - xi[0]), N[1] = 0.5 * (1 + xi[0]);
        dN[0][0] = -0.5, dN[1][0] = 0.5;
        break;
    case ORC_QUAD4:
        for (int a = 0; a < 4; a++)
        {
            N[a] = 0.25 * (1 + q4[a][0] * xi[0]) * (1 + q4[a][1] * xi[1]);
            dN[a][0] = 0.25 * q4[a][0] * (1 + q4[a][1] * xi[1]);
            dN[a][1] = 0.25 * q4[a][1] * (1 + q4[a][0] * xi[0]);
        }
        break;
    case ORC_HEX8:
        for (int a = 0; a < 8; a++)
        {
            real fx = 1 + h8[a][0] * xi[0], fy = 1 + h8[a][1] * xi[1], fz = 1 + h8[a][2] * xi[2];
            N[a] = 0.125 * fx * fy * fz;
            dN[a][0] = 0.125 * h8[a][0] * fy * fz;
            dN[a][1] = 0.125 * h8[a][1] * fx * fz;
            dN[a][2] = 0.125 * h8[a][2] * fx * fy;
        }
        break;
    case ORC_TRI3:
        N[0] = -xi[1] - xi[0] + 1, N[1] = xi[0], N[2] = xi[1];
        dN[0][0] = -1, dN[0][1] = -1, dN[1][0] = 1, dN[1][1] = 0, dN[2][0] = 0, dN[2][1] = 1;
        break;
    case ORC_TET4:
        N[0] = -xi[1] - xi[0] - xi[2] + 1, N[1] = xi[0], N[2] = xi[1], N[3] = xi[2];
        for (int a = 0; a < 4; a++)
            for (int c = 0; c < 3; c++)
                dN[a][c] = a == 0 ? -1.0 : (a == c + 1 ? 1.0 : 0.0);
        break;
    default: throw std::runtime_error("bad element");
    }
}
struct GeoPt
{
    real x[3];
    real J[3][3]; // J[r][c] = d x_r / d xi_c
};
static GeoPt mapPoint(int t, const std::vector<Pt> &nodes, const real *xi)
{
    real N[8], dN[8][3];
    shape(t, xi, N, dN);
    GeoPt g{};
    int nn = elemNNode(t), pd = elemParamDim(t);
    for (int a = 0; a < nn; a++)
        for (int r = 0; r < 3; r++)
        {
            g.x[r] += N[a] * nodes[a].x[r];
            for (int c = 0; c < pd; c++)
                g.J[r][c] += dN[a][c] * nodes[a].x[r];
        }
    return g;
}
// cell Jacobian determinant (CellJacobianDet), face Jacobian determinant (FaceJacobianDet)
// and face normal (FacialJacobianToNormVec, src/Geom/Geometric.hpp:113-130)
static real cellJDet(int dim, const GeoPt &g)
{
    if (dim == 2)
        return g.J[0][0] * g.J[1][1] - g.J[0][1] * g.J[1][0];
    return g.J[0][0] * (g.J[1][1] * g.J[2][2] - g.J[1][2] * g.J[2][1]) -
           g.J[0][1] * (g.J[1][0] * g.J[2][2] - g.J[1][2] * g.J[2][0]) +
           g.J[0][2] * (g.J[1][0] * g.J[2][1] - g.J[1][1] * g.J[2][0]);
}
static void faceNormVec(int dim, const GeoPt &g, real *n)
{
    if (dim == 2)
    {
        n[0] = g.J[1][0], n[1] = -g.J[0][0], n[2] = 0;
        return;
    }
    n[0] = g.J[1][0] * g.J[2][1] - g.J[2][0] * g.J[1][1];
    n[1] = g.J[2][0] * g.J[0][1] - g.J[0][0] * g.J[2][1];
    n[2] = g.J[0][0] * g.J[1][1] - g.J[1][0] * g.J[0][1];
}
static real norm3(const real *v) { return std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]); }

// ---------------------------------------------------------------------------
// Taylor basis (src/Geom/BaseFunction.hpp:53-140, 284-438, 441-712): scaled
// monomials about the cell centre, derivative rows in diffOperatorOrderList order.
// ---------------------------------------------------------------------------
static const int ord3D[20][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {2, 0, 0}, {0, 2, 0}, {0, 0, 2}, {1, 1, 0}, {0, 1, 1}, {1, 0, 1}, {3, 0, 0}, {0, 3, 0}, {0, 0, 3}, {2, 1, 0}, {1, 2, 0}, {0, 2, 1}, {0, 1, 2}, {1, 0, 2}, {2, 0, 1}, {1, 1, 1}};
static const int ord2D[10][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {2, 0, 0}, {1, 1, 0}, {0, 2, 0}, {3, 0, 0}, {2, 1, 0}, {1, 2, 0}, {0, 3, 0}};
static const int dFact[4][4] = {{1, 0, 0, 0}, {1, 1, 0, 0}, {1, 2, 2, 0}, {1, 3, 6, 6}};
static const int nComb2D[10] = {1, 1, 1, 1, 2, 1, 1, 3, 3, 1}; // BaseFunction.hpp:216-217
static const int nComb3D[20] = {1, 1, 1, 1, 1, 1, 1, 2, 2, 2, 1, 1, 1, 3, 3, 3, 3, 3, 3, 6};
static const int factorials[10] = {1, 1, 2, 6, 24, 120, 720, 5040, 40320, 362880};

static int nDofOf(int dim, int order)
{
    static const int d2[4] = {1, 3, 6, 10}, d3[4] = {1, 4, 10, 20};
    return dim == 2 ? d2[order] : d3[order];
}
static real iPow(int p, real x)
{
    switch (p)
    {
    case 0: return 1.;
    case 1: return x;
    case 2: return x * x;
    case 3: return x * x * x;
    default: return 1e300;
    }
}

// ---------------------------------------------------------------------------
// Gas dynamics (src/Euler/Gas.hpp, src/DNDS/IdealGasPhysics.hpp:37-44)
// Physics is always 3 velocity components (model NS / NS_3D, src/Euler/Euler.hpp).
// ---------------------------------------------------------------------------
static inline real sqr(real x) { return x * x; }
static inline real signTol(real a, real tol) { return a > tol ? 1 : (a < -tol ? -1 : 0); } // Defines.hpp:543
static inline real signP(real a) { return a >= 0 ? 1 : -1; }                                // Defines.hpp:549
static inline void thermal(real E, real rho, real vSqr, real gamma, real &p, real &asqr, real &H)
{
    p = (gamma - 1) * (E - rho * 0.5 * vSqr);
    asqr = gamma * p / rho;
    H = (E + p) / rho;
}
static inline real dot3(const real *a, const real *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

// GasInviscidFlux_XY, Gas.hpp:390-397 (vg = 0 in every configuration here)
static void fluxXY(const real *U, const real *velo, const real *n, real p, real *F)
{
    real vn = dot3(velo, n);
    for (int i = 0; i < 5; i++)
        F[i] = U[i] * vn;
    for (int d = 0; d < 3; d++)
        F[1 + d] += p * n[d];
    F[4] += vn * p;
}

struct RoePre
{
    real veloLm[3], veloRm[3], asqrLm, asqrRm, pLm, pRm, HLm, HRm, vsqrLm, vsqrRm;
    real veloRoe[3], sqrtRhoLm, sqrtRhoRm, vsqrRoe, HRoe, asqrRoe, rhoRoe, aRoe;
};
// ComputeRoePreamble, Gas.hpp:229-259
static RoePre roePreamble(const real *ULm, const real *URm, real gamma)
{
    RoePre r;
    for (int d = 0; d < 3; d++)
        r.veloLm[d] = ULm[1 + d] / ULm[0], r.veloRm[d] = URm[1 + d] / URm[0];
    r.vsqrLm = dot3(r.veloLm, r.veloLm);
    r.vsqrRm = dot3(r.veloRm, r.veloRm);
    thermal(ULm[4], ULm[0], r.vsqrLm, gamma, r.pLm, r.asqrLm, r.HLm);
    thermal(URm[4], URm[0], r.vsqrRm, gamma, r.pRm, r.asqrRm, r.HRm);
    r.sqrtRhoLm = std::sqrt(ULm[0]);
    r.sqrtRhoRm = std::sqrt(URm[0]);
    for (int d = 0; d < 3; d++)
        r.veloRoe[d] = (r.sqrtRhoLm * r.veloLm[d] + r.sqrtRhoRm * r.veloRm[d]) / (r.sqrtRhoLm + r.sqrtRhoRm);
    r.vsqrRoe = dot3(r.veloRoe, r.veloRoe);
    r.HRoe = (r.sqrtRhoLm * r.HLm + r.sqrtRhoRm * r.HRm) / (r.sqrtRhoLm + r.sqrtRhoRm);
    r.asqrRoe = (gamma - 1) * (r.HRoe - 0.5 * r.vsqrRoe);
    r.rhoRoe = r.sqrtRhoLm * r.sqrtRhoRm;
    if (!(r.asqrRoe > 0))
        throw std::runtime_error("asqrRoe <= 0");
    r.aRoe = std::sqrt(r.asqrRoe);
    return r;
}

// Roe_EntropyFixer, Gas.hpp:878-1007
static void roeEntropyFix(int eig, real aL, real aR, real aAve, real uL, real uR, real uAve,
                          real VL, real VR, real VAve, real dLambda, real fixScale, real incFScale,
                          real &lam0, real &lam123, real &lam4)
{
    (void)VL, (void)VR;
    const real scaleHartenYee = 0.05 * fixScale, scaleLD = 0.2 * fixScale, scaleHFix = 0.25 * fixScale;
    const real incFScaleA = 1.0;
    if (eig == 0)
    {
        auto Flim = [=](real v, real lam, real lamL, real lamR) {
            const real thresH = std::max({0., lam - lamL, lamR - lam}) * scaleHartenYee;
            if (v < thresH)
                return (sqr(v) / thresH + thresH) * 0.5;
            return v;
        };
        lam0 = Flim(lam0 * incFScaleA, uAve - aAve, uL - aL, uR - aR);
        lam123 = Flim(lam123 * incFScale, uAve, uL, uR);
        lam4 = Flim(lam4 * incFScaleA, uAve + aAve, uL + aL, uR + aR);
    }
    else if (eig == 1)
    {
        const real aLm = aL * incFScaleA, aRm = aR * incFScaleA;
        const real uLm = signTol(uL, aLm * smallReal) * std::max(std::abs(uL) * incFScale, aLm * scaleLD);
        const real uRm = signTol(uR, aRm * smallReal) * std::max(std::abs(uR) * incFScale, aRm * scaleLD);
        lam0 = std::max(std::abs(uLm - aLm), std::abs(uRm - aRm));
        lam123 = std::max(std::abs(uLm), std::abs(uRm));
        lam4 = std::max(std::abs(uLm + aLm), std::abs(uRm + aRm));
    }
    else if (eig == 3)
    {
        const real LDthreshold = std::abs(uAve) / scaleLD;
        const real aRoeStar = std::min(LDthreshold, aAve * incFScaleA);
        lam0 = std::abs(uAve * incFScale - aRoeStar);
        lam4 = std::abs(uAve * incFScale + aRoeStar);
    }
    else if (eig == 4)
    {
        const real uStar = signTol(uAve, aAve * smallReal) * std::max(aAve * scaleLD, std::abs(uAve) * incFScale);
        lam0 = std::abs(uStar - aAve * incFScaleA);
        lam123 = std::abs(uStar);
        lam4 = std::abs(uStar + aAve * incFScaleA);
    }
    else if (eig == 5)
    {
        const real veloLm0 = uL * incFScale, veloRm0 = uR * incFScale;
        const real aLm = std::min(aL * incFScaleA, std::abs(veloLm0) / scaleLD);
        const real aRm = std::min(aR * incFScaleA, std::abs(veloRm0) / scaleLD);
        lam0 = std::max(std::abs(veloLm0 - aLm), std::abs(veloRm0 - aRm));
        lam123 = std::max(std::abs(veloLm0), std::abs(veloRm0));
        lam4 = std::max(std::abs(veloLm0 + aLm), std::abs(veloRm0 + aRm));
    }
    else if (eig == 6 || eig == 8)
    {
        lam0 = std::max(std::abs(uAve * incFScale - aAve * incFScaleA), dLambda * scaleHFix);
        lam4 = std::max(std::abs(uAve * incFScale + aAve * incFScaleA), dLambda * scaleHFix);
        lam123 = std::max(std::abs(uAve * incFScale), dLambda * scaleHFix);
    }
    else if (eig == 7)
    {
        lam0 *= std::abs(uAve * incFScale - aAve * incFScaleA);
        lam4 *= std::abs(uAve * incFScale + aAve * incFScaleA);
        lam123 *= std::abs(uAve * incFScale);
    }
    else
        throw std::runtime_error("bad eigScheme");
    if (eig == 7 || eig == 8)
    {
        const real th = scaleHartenYee * (VAve + aAve), thS = sqr(th);
        if (lam0 < th)
            lam0 = (sqr(lam0) + thS) / (2 * th);
        if (lam4 < th)
            lam4 = (sqr(lam4) + thS) / (2 * th);
        if (lam123 < th)
            lam123 = (sqr(lam123) + thS) / (2 * th);
    }
}

// Roe dissipation |A| dU with Roe eigenvectors (Gas.hpp:1091-1125, 1209-1243)
static void roeIncF(const real *incU, const real *n, const real *veloRoe, real vsqrRoe, real aRoe,
                    real asqrRoe, real HRoe, real lam0, real lam123, real lam4, real gamma, real *incF)
{
    real veloRoeN = dot3(veloRoe, n);
    real incU123N = incU[1] * n[0] + incU[2] * n[1] + incU[3] * n[2];
    real a23V[3], a23VT[3];
    for (int d = 0; d < 3; d++)
        a23V[d] = incU[1 + d] - incU[0] * veloRoe[d];
    real a23Vn = dot3(a23V, n);
    for (int d = 0; d < 3; d++)
        a23VT[d] = a23V[d] - n[d] * a23Vn;
    real incU4b = incU[4] - dot3(a23VT, veloRoe);
    real alpha1 = (gamma - 1) / asqrRoe * (incU[0] * (HRoe - veloRoeN * veloRoeN) + veloRoeN * incU123N - incU4b);
    real alpha0 = (incU[0] * (veloRoeN + aRoe) - incU123N - aRoe * alpha1) / (2 * aRoe);
    real alpha4 = incU[0] - (alpha0 + alpha1);
    alpha0 *= lam0, alpha1 *= lam123, alpha4 *= lam4;
    for (int d = 0; d < 3; d++)
        a23VT[d] *= lam123;
    incF[0] = alpha0 + alpha1 + alpha4;
    incF[4] = (HRoe - veloRoeN * aRoe) * alpha0 + 0.5 * vsqrRoe * alpha1 + (HRoe + veloRoeN * aRoe) * alpha4 + dot3(a23VT, veloRoe);
    for (int d = 0; d < 3; d++)
        incF[1 + d] = (veloRoe[d] - aRoe * n[d]) * alpha0 + (veloRoe[d] + aRoe * n[d]) * alpha4 + veloRoe[d] * alpha1 + a23VT[d];
}

// RoeFlux_IdealGas_HartenYee<dim, eigScheme>, Gas.hpp:1038-1126 (vg = 0)
static void roeFlux(int eig, const real *UL, const real *UR, const real *ULm, const real *URm,
                    const real *n, real gamma, real *F, real dLambda, real fixScale, real incFScale,
                    real &lam0, real &lam123, real &lam4)
{
    if (!(UL[0] > 0 && UR[0] > 0))
        throw std::runtime_error("rho <= 0 in Roe flux");
    real veloL[3], veloR[3];
    for (int d = 0; d < 3; d++)
        veloL[d] = UL[1 + d] / UL[0], veloR[d] = UR[1 + d] / UR[0];
    real asqrL, asqrR, pL, pR, HL, HR;
    thermal(UL[4], UL[0], dot3(veloL, veloL), gamma, pL, asqrL, HL);
    thermal(UR[4], UR[0], dot3(veloR, veloR), gamma, pR, asqrR, HR);
    RoePre rp = roePreamble(ULm, URm, gamma);
    real FL[5], FR[5];
    fluxXY(UL, veloL, n, pL, FL);
    fluxXY(UR, veloR, n, pR, FR);
    real veloRoeN = dot3(rp.veloRoe, n);
    real veloRoe0 = veloRoeN;
    lam0 = std::abs(veloRoe0 - rp.aRoe);
    lam123 = std::abs(veloRoe0);
    lam4 = std::abs(veloRoe0 + rp.aRoe);
    real veloLm0 = dot3(rp.veloLm, n), veloRm0 = dot3(rp.veloRm, n);
    if (eig == 2)
    {
        lam0 = lam123 = lam4 = std::max(std::abs(veloLm0) + std::sqrt(rp.asqrLm), std::abs(veloRm0) + std::sqrt(rp.asqrRm));
        for (int i = 0; i < 5; i++)
            F[i] = (FL[i] + FR[i]) * 0.5 - 0.5 * lam0 * (UR[i] - UL[i]);
        return;
    }
    roeEntropyFix(eig, std::sqrt(rp.asqrLm), std::sqrt(rp.asqrRm), rp.aRoe, veloLm0, veloRm0, veloRoe0,
                  norm3(rp.veloLm), norm3(rp.veloRm), norm3(rp.veloRoe), dLambda, fixScale, incFScale,
                  lam0, lam123, lam4);
    real incU[5], incF[5];
    for (int i = 0; i < 5; i++)
        incU[i] = UR[i] - UL[i];
    roeIncF(incU, n, rp.veloRoe, rp.vsqrRoe, rp.aRoe, rp.asqrRoe, rp.HRoe, lam0, lam123, lam4, gamma, incF);
    for (int i = 0; i < 5; i++)
        F[i] = (FL[i] + FR[i]) * 0.5 - 0.5 * incF[i];
}

// HLLEPFlux_IdealGas<dim, type>, Gas.hpp:601-714 (vg = 0)
static void hllepFlux(int type, const real *UL, const real *UR, const real *ULm, const real *URm,
                      const real *n, real gamma, real *F, real dLambda, real fixScale,
                      real &lam0, real &lam123, real &lam4)
{
    if (!(UL[0] > 0 && UR[0] > 0))
        throw std::runtime_error("rho <= 0 in HLLEP flux");
    real veloL[3], veloR[3];
    for (int d = 0; d < 3; d++)
        veloL[d] = UL[1 + d] / UL[0], veloR[d] = UR[1 + d] / UR[0];
    real asqrL, asqrR, pL, pR, HL, HR;
    thermal(UL[4], UL[0], dot3(veloL, veloL), gamma, pL, asqrL, HL);
    thermal(UR[4], UR[0], dot3(veloR, veloR), gamma, pR, asqrR, HR);
    RoePre rp = roePreamble(ULm, URm, gamma);
    real FL[5], FR[5];
    fluxXY(UL, veloL, n, pL, FL);
    fluxXY(UR, veloR, n, pR, FR);
    real veloRoeN = dot3(rp.veloRoe, n);
    real veloRoe0 = veloRoeN;
    lam0 = std::abs(veloRoe0 - rp.aRoe);
    lam123 = std::abs(veloRoe0);
    lam4 = std::abs(veloRoe0 + rp.aRoe);
    real veloLm0 = dot3(rp.veloLm, n), veloRm0 = dot3(rp.veloRm, n);
    real incU[5];
    for (int i = 0; i < 5; i++)
        incU[i] = UR[i] - UL[i];
    real incU123N = incU[1] * n[0] + incU[2] * n[1] + incU[3] * n[2];
    real a23V[3], a23VT[3];
    for (int d = 0; d < 3; d++)
        a23V[d] = incU[1 + d] - incU[0] * rp.veloRoe[d];
    real a23Vn = dot3(a23V, n);
    for (int d = 0; d < 3; d++)
        a23VT[d] = a23V[d] - n[d] * a23Vn;
    real incU4b = incU[4] - dot3(a23VT, rp.veloRoe);
    real alpha1 = (gamma - 1) / rp.asqrRoe * (incU[0] * (rp.HRoe - veloRoeN * veloRoeN) + veloRoeN * incU123N - incU4b);
    real alpha0 = (incU[0] * (veloRoeN + rp.aRoe) - incU123N - rp.aRoe * alpha1) / (2 * rp.aRoe);
    real alpha4 = incU[0] - (alpha0 + alpha1);
    real SL = std::min(veloRoe0 - rp.aRoe, veloLm0 - std::sqrt(asqrL));
    real SR = std::max(veloRoe0 + rp.aRoe, veloRm0 + std::sqrt(asqrR));
    real UU = std::abs(veloRoe0);
    real dfix = rp.aRoe / (rp.aRoe + std::max(UU, dLambda * fixScale));
    real SP = std::max(SR, 0.0), SM = std::min(SL, 0.0);
    if (type != 1)
    {
        real ReV1[5] = {1, rp.veloRoe[0], rp.veloRoe[1], rp.veloRoe[2], rp.vsqrRoe * 0.5};
        real div = SP - SM;
        div += signP(div) * verySmallReal;
        for (int i = 0; i < 5; i++)
            F[i] = (SP * FL[i] - SM * FR[i]) / div + (SP * SM / div) * (UR[i] - UL[i] - dfix * ReV1[i] * alpha1);
    }
    else
    {
        real delta1 = dfix, delta2 = 0.0, delta3 = 0.0;
        lam123 = ((SP + SM) * (veloRoe0)-2.0 * (1.0 - delta1) * (SP * SM)) / (SP - SM);
        lam4 = ((SP + SM) * (veloRoe0 + rp.aRoe) - 2.0 * (1.0 - delta2) * (SP * SM)) / (SP - SM);
        lam0 = ((SP + SM) * (veloRoe0 - rp.aRoe) - 2.0 * (1.0 - delta3) * (SP * SM)) / (SP - SM);
        lam123 = std::abs(lam123), lam0 = std::abs(lam0), lam4 = std::abs(lam4);
        real incF[5];
        roeIncF(incU, n, rp.veloRoe, rp.vsqrRoe, rp.aRoe, rp.asqrRoe, rp.HRoe, lam0, lam123, lam4, gamma, incF);
        for (int i = 0; i < 5; i++)
            F[i] = (FL[i] + FR[i]) * 0.5 - 0.5 * incF[i];
    }
}

// HLLCFlux_IdealGas_HartenYee, Gas.hpp:733-843 (vg = 0)
static void hllcFlux(const real *UL, const real *UR, const real *ULm, const real *URm,
                     const real *n, real gamma, real *F, real &lam0, real &lam123, real &lam4)
{
    if (!(UL[0] > 0 && UR[0] > 0))
        throw std::runtime_error("rho <= 0 in HLLC flux");
    real veloL[3], veloR[3];
    for (int d = 0; d < 3; d++)
        veloL[d] = UL[1 + d] / UL[0], veloR[d] = UR[1 + d] / UR[0];
    real asqrL, asqrR, pL, pR, HL, HR;
    thermal(UL[4], UL[0], dot3(veloL, veloL), gamma, pL, asqrL, HL);
    thermal(UR[4], UR[0], dot3(veloR, veloR), gamma, pR, asqrR, HR);
    RoePre rp = roePreamble(ULm, URm, gamma);
    real FL[5], FR[5];
    fluxXY(UL, veloL, n, pL, FL);
    fluxXY(UR, veloR, n, pR, FR);
    real veloRoe0 = dot3(rp.veloRoe, n);
    lam0 = std::abs(veloRoe0 - rp.aRoe);
    lam123 = std::abs(veloRoe0);
    lam4 = std::abs(veloRoe0 + rp.aRoe);
    real veloLm0 = dot3(rp.veloLm, n), veloRm0 = dot3(rp.veloRm, n);
    auto HLLCq = [&](real p, real pS) {
        real q = std::sqrt(1 + (gamma + 1) / 2 / gamma * (pS / p - 1));
        if (pS <= p)
            q = 1;
        return q;
    };
    real pS = 0.5 * (rp.pLm + rp.pRm) - 0.5 * (veloLm0 - veloRm0) * rp.rhoRoe * rp.aRoe;
    pS = std::max(0.0, pS);
    real SL = veloLm0 - std::sqrt(rp.asqrLm) * HLLCq(rp.pLm, pS);
    real SR = veloRm0 + std::sqrt(rp.asqrRm) * HLLCq(rp.pRm, pS);
    if (0 <= SL)
    {
        std::copy(FL, FL + 5, F);
        return;
    }
    if (SR <= 0)
    {
        std::copy(FR, FR + 5, F);
        return;
    }
    real SS = 0;
    real div = (UL[0] * (SL - veloLm0) - UR[0] * (SR - veloRm0));
    if (std::abs(div) > verySmallReal)
        SS = (pR - pL + (UL[1] * n[0] + UL[2] * n[1] + UL[3] * n[2]) * (SL - veloLm0) -
              (UR[1] * n[0] + UR[2] * n[1] + UR[3] * n[2]) * (SR - veloRm0)) /
             div;
    real DS[5] = {0, n[0], n[1], n[2], SS};
    if (SS >= 0)
    {
        real dv = SL - SS;
        if (std::abs(dv) < verySmallReal)
            std::copy(FL, FL + 5, F);
        else
            for (int i = 0; i < 5; i++)
                F[i] = ((UL[i] * SL - FL[i]) * SS + DS[i] * ((pL + UL[0] * (SL - veloLm0) * (SS - veloLm0)) * SL)) / dv;
    }
    else
    {
        real dv = SR - SS;
        if (std::abs(dv) < verySmallReal)
            std::copy(FR, FR + 5, F);
        else
            for (int i = 0; i < 5; i++)
                F[i] = ((UR[i] * SR - FR[i]) * SS + DS[i] * ((pR + UR[0] * (SR - veloRm0) * (SS - veloRm0)) * SR)) / dv;
    }
}

// InviscidFlux_IdealGas_Dispatcher, Gas.hpp:1434-1489
static void dispatchFlux(int rsType, const real *UL, const real *UR, const real *ULm, const real *URm,
                         const real *n, real gamma, real *F, real dLambda, real fixScale, real incFScale,
                         real &lam0, real &lam123, real &lam4)
{
    if (rsType == 1)
        roeFlux(0, UL, UR, ULm, URm, n, gamma, F, dLambda, fixScale, incFScale, lam0, lam123, lam4);
    else if (rsType >= 11 && rsType <= 18)
        roeFlux(rsType - 10, UL, UR, ULm, URm, n, gamma, F, dLambda, fixScale, incFScale, lam0, lam123, lam4);
    else if (rsType == 3)
        hllepFlux(0, UL, UR, ULm, URm, n, gamma, F, dLambda, fixScale, lam0, lam123, lam4);
    else if (rsType == 21)
        hllepFlux(1, UL, UR, ULm, URm, n, gamma, F, dLambda, fixScale, lam0, lam123, lam4);
    else if (rsType == 2)
        hllcFlux(UL, UR, ULm, URm, n, gamma, F, lam0, lam123, lam4);
    else
        throw std::runtime_error("the rs type is invalid");
}

// RoeFlux_IdealGas_HartenYee_Batch, Gas.hpp:1272-1407: eigenvalues from the cell-mean
// Roe average on the mean normal nm, shared by all points of the face (vg = 0).
static void roeFluxBatch(int eig, int nB, const real (*UL)[5], const real (*UR)[5], const real *ULm,
                         const real *URm, const real (*n)[3], const real *nm, real gamma, real (*F)[5],
                         real dLambda, real fixScale, real incFScale, real &lam0, real &lam123, real &lam4)
{
    for (int iB = 0; iB < nB; iB++)
        if (!(UL[iB][0] > 0 && UR[iB][0] > 0))
            throw std::runtime_error("rho <= 0 in batch Roe flux");
    real veloLm[3], veloRm[3];
    for (int d = 0; d < 3; d++)
        veloLm[d] = ULm[1 + d] / ULm[0], veloRm[d] = URm[1 + d] / URm[0];
    real asqrLm, asqrRm, pLm, pRm, HLm, HRm;
    thermal(ULm[4], ULm[0], dot3(veloLm, veloLm), gamma, pLm, asqrLm, HLm);
    thermal(URm[4], URm[0], dot3(veloRm, veloRm), gamma, pRm, asqrRm, HRm);
    real sqrtRhoLm = std::sqrt(ULm[0]), sqrtRhoRm = std::sqrt(URm[0]);
    real veloRoe[3];
    for (int d = 0; d < 3; d++)
        veloRoe[d] = (sqrtRhoLm * veloLm[d] + sqrtRhoRm * veloRm[d]) / (sqrtRhoLm + sqrtRhoRm);
    real vsqrRoe = dot3(veloRoe, veloRoe);
    real HRoe = (sqrtRhoLm * HLm + sqrtRhoRm * HRm) / (sqrtRhoLm + sqrtRhoRm);
    real asqrRoe = (gamma - 1) * (HRoe - 0.5 * vsqrRoe);
    real veloRoeN = dot3(veloRoe, nm);
    real veloRoeRN = veloRoeN;
    real veloLm0 = dot3(veloLm, nm), veloRm0 = dot3(veloRm, nm);
    if (!(asqrRoe > 0))
        throw std::runtime_error("asqrRoe <= 0");
    real aRoe = std::sqrt(asqrRoe);
    lam0 = std::abs(veloRoeRN - aRoe);
    lam123 = std::abs(veloRoeRN);
    lam4 = std::abs(veloRoeRN + aRoe);
    std::vector<std::array<real, 5>> FL(nB), FR(nB);
    for (int iB = 0; iB < nB; iB++)
    {
        real vL[3], vR[3], pL, pR, a2, H;
        for (int d = 0; d < 3; d++)
            vL[d] = UL[iB][1 + d] / UL[iB][0], vR[d] = UR[iB][1 + d] / UR[iB][0];
        thermal(UL[iB][4], UL[iB][0], dot3(vL, vL), gamma, pL, a2, H);
        thermal(UR[iB][4], UR[iB][0], dot3(vR, vR), gamma, pR, a2, H);
        fluxXY(UL[iB], vL, n[iB], pL, FL[iB].data());
        fluxXY(UR[iB], vR, n[iB], pR, FR[iB].data());
    }
    if (eig == 2)
    {
        lam0 = lam123 = lam4 = std::max(std::abs(veloLm0) + std::sqrt(asqrLm), std::abs(veloRm0) + std::sqrt(asqrRm));
        for (int iB = 0; iB < nB; iB++)
            for (int i = 0; i < 5; i++)
                F[iB][i] = (FL[iB][i] + FR[iB][i]) * 0.5 - 0.5 * lam0 * (UR[iB][i] - UL[iB][i]);
        return;
    }
    roeEntropyFix(eig, std::sqrt(asqrLm), std::sqrt(asqrRm), aRoe, veloLm0, veloRm0, veloRoeRN,
                  norm3(veloLm), norm3(veloRm), norm3(veloRoe), dLambda, fixScale, incFScale, lam0, lam123, lam4);
    for (int iB = 0; iB < nB; iB++)
    {
        // per-point decomposition: jump projected on the point normal, eigenvalues shared,
        // veloRoeN taken on the mean normal as in the reference batch code
        const real *nn = n[iB];
        real incU[5];
        for (int i = 0; i < 5; i++)
            incU[i] = UR[iB][i] - UL[iB][i];
        real incU123N = incU[1] * nn[0] + incU[2] * nn[1] + incU[3] * nn[2];
        real a23V[3], a23VT[3];
        for (int d = 0; d < 3; d++)
            a23V[d] = incU[1 + d] - veloRoe[d] * incU[0];
        real a23Vn = dot3(a23V, nn);
        for (int d = 0; d < 3; d++)
            a23VT[d] = a23V[d] - nn[d] * a23Vn;
        real incU4b = incU[4] - dot3(veloRoe, a23VT);
        real alpha1 = (gamma - 1) / asqrRoe * (incU[0] * (HRoe - veloRoeN * veloRoeN) + veloRoeN * incU123N - incU4b);
        real alpha0 = (incU[0] * (veloRoeN + aRoe) - incU123N - aRoe * alpha1) / (2 * aRoe);
        real alpha4 = incU[0] - (alpha0 + alpha1);
        alpha0 *= lam0, alpha1 *= lam123, alpha4 *= lam4;
        for (int d = 0; d < 3; d++)
            a23VT[d] *= lam123;
        real incF[5];
        incF[0] = alpha0 + alpha1 + alpha4;
        incF[4] = (HRoe - veloRoeN * aRoe) * alpha0 + 0.5 * vsqrRoe * alpha1 + (HRoe + veloRoeN * aRoe) * alpha4 + dot3(veloRoe, a23VT);
        for (int d = 0; d < 3; d++)
            incF[1 + d] = (-aRoe * nn[d] + veloRoe[d]) * alpha0 + (aRoe * nn[d] + veloRoe[d]) * alpha4 + a23VT[d] + veloRoe[d] * alpha1;
        for (int i = 0; i < 5; i++)
            F[iB][i] = (FL[iB][i] + FR[iB][i]) * 0.5 - 0.5 * incF[i];
    }
}

// ---------------------------------------------------------------------------
// Mesh + FiniteVolume metrics + VR setup
// ---------------------------------------------------------------------------
struct VR
{
    int dim{2};
    orc_vr_settings s{};
    idx_t nCell{0}, nFace{0};
    int maxOrder{1}, NDOF{3}, M{2};
    std::vector<int32_t> cellType, faceType;
    std::vector<std::vector<Pt>> cellNodes, faceNodes;
    std::vector<std::vector<idx_t>> c2f;
    std::vector<std::array<idx_t, 2>> f2c;
    std::vector<Pt> fshift;
    // cell metrics
    std::vector<real> vol, smooth;
    std::vector<Pt> bary, cent, hbox;
    std::vector<std::vector<Pt>> cellQP;
    std::vector<std::vector<real>> cellJD; // JDet + verySmall
    std::vector<std::vector<real>> cellQw; // rule weights per cell (its type's rule)
    std::vector<std::vector<real>> moment; // [nCell][NDOF]
    // face metrics
    std::vector<real> area, faceL;
    std::vector<Pt> faceCent, faceMeanNorm, faceBaryDiff;
    std::vector<std::vector<Pt>> faceQP, faceN;
    std::vector<std::vector<real>> faceJD;
    std::vector<real> faceQw;
    std::vector<std::vector<real>> faceWeight; // [nFace][maxOrder+1]
    // VR coefficients
    std::vector<real> AinvB; // [nCell][maxFC+1][M][M]
    std::vector<real> Ainvb; // [nCell][maxFC][M]
    int maxFC{0};
    real globalVol{0};
    // boundary zones (faces with f2c[f][1] < 0) and their conditions
    std::vector<int32_t> faceZone;
    int nBCZone{0};
    std::vector<int32_t> bcType;              // EulerBCType per zone (empty: no Euler BC table set)
    std::vector<std::array<real, 5>> bcValue; // per zone
    real bcGamma{1.4};
    real bcRgas{1.0};
    int dirNVars{0};
    std::vector<real> dirValue;               // [nFace][G][dirNVars] boundary point values

    bool cellIsBack(idx_t iCell, idx_t iFace) const { return f2c[iFace][0] == iCell; }
    idx_t cellFaceOther(idx_t iCell, idx_t iFace) const { return f2c[iFace][0] == iCell ? f2c[iFace][1] : f2c[iFace][0]; }
    // GetFaceQuadraturePPhysFromCell: the face point expressed in the frame of side if2c
    Pt facePointFromSide(idx_t iFace, int if2c, const Pt &p) const
    {
        Pt r = p;
        if (if2c == 1)
            for (int d = 0; d < 3; d++)
                r.x[d] += fshift[iFace].x[d];
        return r;
    }
    // GetOtherCellBaryFromCell, FiniteVolume.hpp:319-337
    Pt otherBaryFromCell(idx_t iCell, idx_t iCellOther, idx_t iFace) const
    {
        Pt r = bary[iCellOther];
        real sgn = cellIsBack(iCell, iFace) ? -1.0 : 1.0;
        for (int d = 0; d < 3; d++)
            r.x[d] += sgn * fshift[iFace].x[d];
        return r;
    }

    // FDiffBaseValue, VariationalReconstruction.hpp:317-369 (isotropic aligned h-box scaling)
    void diffBase(idx_t iCell, const Pt &p, int nDiff, int flag, std::vector<real> &D) const
    {
        D.assign((size_t)nDiff * NDOF, 0.0);
        Pt sc = hbox[iCell];
        if (dim == 2)
        {
            real m = std::max(sc.x[0], sc.x[1]);
            sc.x[0] = sc.x[1] = m;
        }
        else
        {
            real m = std::max({sc.x[0], sc.x[1], sc.x[2]});
            sc.x[0] = sc.x[1] = sc.x[2] = m;
        }
        real xs[3];
        for (int d = 0; d < 3; d++)
            xs[d] = (p.x[d] - cent[iCell].x[d]) / sc.x[d];
        const int(*ord)[3] = dim == 2 ? ord2D : ord3D;
        for (int id = 0; id < nDiff; id++)
            for (int ib = 0; ib < NDOF; ib++)
            {
                int px = ord[ib][0], py = ord[ib][1], pz = ord[ib][2];
                int dx = ord[id][0], dy = ord[id][1], dz = ord[id][2];
                int c = (dx <= px ? dFact[px][dx] : 0) * (dy <= py ? dFact[py][dy] : 0) * (dz <= pz ? dFact[pz][dz] : 0);
                real v = c ? c * iPow(px - dx, xs[0]) * iPow(py - dy, xs[1]) * iPow(pz - dz, xs[2]) : 0.;
                D[(size_t)id * NDOF + ib] = v / (iPow(dx, sc.x[0]) * iPow(dy, sc.x[1]) * iPow(dz, sc.x[2]));
            }
        if (flag == 0)
            for (int ib = 0; ib < NDOF; ib++)
                D[ib] -= moment[iCell][ib];
    }

    // FFaceFunctional + AccumulateDiffOrderContributions, VariationalReconstruction.hpp:32-93, 459-505
    // DI: cnDiffs x nI (columns = basis 1..NDOF-1), DJ: cnDiffs x nJ; Conj += ...
    void faceFunctional(idx_t iFace, const real *DI, int nI, const real *DJ, int nJ, int cnDiffs, real *Conj, real fac) const
    {
        const int *comb = dim == 2 ? nComb2D : nComb3D;
        const int start2[4] = {0, 1, 3, 6}, cnt2[4] = {1, 2, 3, 4};
        const int start3[4] = {0, 1, 4, 10}, cnt3[4] = {1, 3, 6, 10};
        const int *st = dim == 2 ? start2 : start3, *ct = dim == 2 ? cnt2 : cnt3;
        int maxO = 0;
        for (int o = 0; o < 4; o++)
            if (st[o] + ct[o] <= cnDiffs)
                maxO = o;
        real fL = faceL[iFace];
        std::vector<real> tmp((size_t)nI * nJ, 0.0);
        for (int o = maxO; o >= 0; o--)
        {
            real fL2 = fL * fL;
            real pw = o == 0 ? 1.0 : (o == 1 ? fL2 : (o == 2 ? fL2 * fL2 : fL2 * fL2 * fL2));
            real w = faceWeight[iFace][o];
            real scale = w * w * pw;
            for (int i = 0; i < nI; i++)
                for (int j = 0; j < nJ; j++)
                {
                    real r = 0;
                    for (int k = st[o]; k < st[o] + ct[o]; k++)
                        r += DI[(size_t)k * nI + i] * DJ[(size_t)k * nJ + j] * comb[k];
                    tmp[(size_t)i * nJ + j] += r * scale;
                }
        }
        for (size_t q = 0; q < tmp.size(); q++)
            Conj[q] += tmp[q] * fac;
    }
};

static Pt toPt(const double *c)
{
    Pt p;
    p.x[0] = c[0], p.x[1] = c[1], p.x[2] = c[2];
    return p;
}

// Gauss-Jordan inverse with partial pivoting (stands in for Eigen's JacobiSVD
// pseudo-inverse, HardEigen.cpp:13-23; A is SPD here so both give A^-1).
static void invertMatrix(int n, std::vector<real> A, std::vector<real> &Ainv)
{
    Ainv.assign((size_t)n * n, 0.0);
    for (int i = 0; i < n; i++)
        Ainv[(size_t)i * n + i] = 1.0;
    for (int c = 0; c < n; c++)
    {
        int piv = c;
        real best = std::abs(A[(size_t)c * n + c]);
        for (int r = c + 1; r < n; r++)
            if (std::abs(A[(size_t)r * n + c]) > best)
                best = std::abs(A[(size_t)r * n + c]), piv = r;
        if (best == 0)
            throw std::runtime_error("singular VR matrix");
        if (piv != c)
            for (int k = 0; k < n; k++)
            {
                std::swap(A[(size_t)c * n + k], A[(size_t)piv * n + k]);
                std::swap(Ainv[(size_t)c * n + k], Ainv[(size_t)piv * n + k]);
            }
        real d = 1.0 / A[(size_t)c * n + c];
        for (int k = 0; k < n; k++)
            A[(size_t)c * n + k] *= d, Ainv[(size_t)c * n + k] *= d;
        for (int r = 0; r < n; r++)
        {
            if (r == c)
                continue;
            real f = A[(size_t)r * n + c];
            if (f == 0)
                continue;
            for (int k = 0; k < n; k++)
                A[(size_t)r * n + k] -= f * A[(size_t)c * n + k], Ainv[(size_t)r * n + k] -= f * Ainv[(size_t)c * n + k];
        }
    }
}

static void buildVR(VR &v, const orc_mesh_desc *m, const orc_vr_settings *s)
{
    v.dim = m->dim;
    v.s = *s;
    v.nCell = m->nCell;
    v.nFace = m->nFace;
    v.maxOrder = s->maxOrder;
    v.NDOF = nDofOf(v.dim, s->maxOrder);
    v.M = v.NDOF - 1;
    int dim = v.dim;
    v.cellType.assign(m->cellType, m->cellType + m->nCell);
    v.faceType.assign(m->faceType, m->faceType + m->nFace);
    // one face rule per mesh (a single face type); cells may mix types of one dimension
    for (auto t : v.cellType)
        if (elemParamDim(t) != elemParamDim(v.cellType[0]))
            throw std::runtime_error("cells of different dimensions");
    for (auto t : v.faceType)
        if (t != v.faceType[0])
            throw std::runtime_error("mixed face types are not supported");
    v.cellNodes.resize(v.nCell);
    v.c2f.resize(v.nCell);
    v.maxFC = 0;
    for (idx_t i = 0; i < v.nCell; i++)
    {
        for (idx_t k = m->cell2nodePtr[i]; k < m->cell2nodePtr[i + 1]; k++)
            v.cellNodes[i].push_back(toPt(m->coords + 3 * m->cell2node[k]));
        for (idx_t k = m->cell2facePtr[i]; k < m->cell2facePtr[i + 1]; k++)
            v.c2f[i].push_back(m->cell2face[k]);
        v.maxFC = std::max(v.maxFC, (int)v.c2f[i].size());
    }
    v.faceNodes.resize(v.nFace);
    v.f2c.resize(v.nFace);
    v.fshift.resize(v.nFace);
    for (idx_t f = 0; f < v.nFace; f++)
    {
        for (idx_t k = m->face2nodePtr[f]; k < m->face2nodePtr[f + 1]; k++)
            v.faceNodes[f].push_back(toPt(m->coords + 3 * m->face2node[k]));
        v.f2c[f] = {m->face2cell[2 * f], m->face2cell[2 * f + 1]};
        v.fshift[f] = toPt(m->faceShift + 3 * f);
    }
    v.nBCZone = m->nBCZone;
    v.faceZone.assign(v.nFace, -1);
    for (idx_t f = 0; f < v.nFace; f++)
        if (v.f2c[f][1] < 0)
        {
            if (!m->faceZone || m->faceZone[f] < 0 || m->faceZone[f] >= m->nBCZone)
                throw std::runtime_error("boundary face without a valid zone");
            v.faceZone[f] = m->faceZone[f];
        }

    // ---- cells: FiniteVolume.cpp:23-326
    v.vol.resize(v.nCell);
    v.bary.resize(v.nCell);
    v.cent.resize(v.nCell);
    v.hbox.resize(v.nCell);
    v.cellQP.resize(v.nCell);
    v.cellJD.resize(v.nCell);
    v.cellQw.resize(v.nCell);
    v.globalVol = 0;
    for (idx_t i = 0; i < v.nCell; i++)
    {
        int t = v.cellType[i];
        auto qMax = quadRule(t, INT_ORDER_MAX);
        real vv = 0;
        for (auto &q : qMax)
        {
            GeoPt g = mapPoint(t, v.cellNodes[i], q.p);
            vv += (1 * cellJDet(dim, g)) * q.w;
        }
        vv += verySmallReal;
        v.vol[i] = vv;
        v.globalVol += vv;
        auto qc = quadRule(t, s->intOrder);
        v.cellQw[i].clear();
        Pt b;
        for (auto &q : qc)
        {
            GeoPt g = mapPoint(t, v.cellNodes[i], q.p);
            real jd = cellJDet(dim, g);
            v.cellJD[i].push_back(jd + verySmallReal);
            v.cellQP[i].push_back(toPt(g.x));
            v.cellQw[i].push_back(q.w);
            for (int d = 0; d < 3; d++)
                b.x[d] += (g.x[d] * jd) * q.w;
        }
        for (int d = 0; d < 3; d++)
            b.x[d] /= vv;
        v.bary[i] = b;
        auto q1 = quadRule(t, 1);
        v.cent[i] = toPt(mapPoint(t, v.cellNodes[i], q1[0].p).x);
        Pt h;
        for (auto &nd : v.cellNodes[i])
            for (int d = 0; d < 3; d++)
                h.x[d] = std::max(h.x[d], std::abs(nd.x[d] - b.x[d]));
        if (dim == 2)
            h.x[2] = 1;
        v.hbox[i] = h;
    }
    // ---- faces: FiniteVolume.cpp:353-575
    v.area.resize(v.nFace);
    v.faceCent.resize(v.nFace);
    v.faceMeanNorm.resize(v.nFace);
    v.faceQP.resize(v.nFace);
    v.faceN.resize(v.nFace);
    v.faceJD.resize(v.nFace);
    for (idx_t f = 0; f < v.nFace; f++)
    {
        int t = v.faceType[f];
        auto qf = quadRule(t, s->intOrder);
        real a = 0;
        Pt mn;
        v.faceQw.clear();
        for (auto &q : qf)
        {
            GeoPt g = mapPoint(t, v.faceNodes[f], q.p);
            real nv[3];
            faceNormVec(dim, g, nv);
            real jd = norm3(nv);
            a += (1 * jd) * q.w;
            v.faceJD[f].push_back(jd + verySmallReal);
            v.faceQP[f].push_back(toPt(g.x));
            Pt nu;
            for (int d = 0; d < 3; d++)
                nu.x[d] = nv[d] / jd;
            v.faceN[f].push_back(nu);
            for (int d = 0; d < 3; d++)
                mn.x[d] += (nu.x[d] * jd) * q.w;
            v.faceQw.push_back(q.w);
        }
        v.area[f] = a + verySmallReal;
        real nn = norm3(mn.x);
        for (int d = 0; d < 3; d++)
            mn.x[d] /= nn;
        v.faceMeanNorm[f] = mn;
        auto q1 = quadRule(t, 1);
        v.faceCent[f] = toPt(mapPoint(t, v.faceNodes[f], q1[0].p).x);
    }
    // ---- cell smooth scale, FiniteVolume.cpp:581-631
    {
        std::vector<real> sc(v.nCell), cur(v.nCell), nw(v.nCell);
        for (idx_t i = 0; i < v.nCell; i++)
        {
            real fs = 0;
            for (auto f : v.c2f[i])
                fs += v.area[f];
            sc[i] = v.vol[i] / (fs + verySmallReal);
            cur[i] = sc[i];
        }
        nw = cur;
        for (int it = 1; it <= s->nIterCellSmoothScale; it++)
        {
            for (idx_t i = 0; i < v.nCell; i++)
            {
                real nAdj = 0, sum = 0;
                for (auto f : v.c2f[i])
                {
                    idx_t o = v.cellFaceOther(i, f);
                    if (o >= 0)
                        nAdj += 1., sum += cur[o];
                }
                real w = 1;
                sum += nAdj * w * cur[i];
                sum /= nAdj * (1 + w);
                nw[i] = std::min(cur[i], sum);
            }
            cur = nw;
        }
        v.smooth.resize(v.nCell);
        for (idx_t i = 0; i < v.nCell; i++)
            v.smooth[i] = cur[i] / sc[i];
    }

    // ---- base moments, VariationalReconstruction.cpp:62-108
    v.moment.assign(v.nCell, std::vector<real>(v.NDOF, 0.0));
#pragma omp parallel num_threads(g_threads)
    {
        std::vector<real> D;
#pragma omp for schedule(static)
        for (idx_t i = 0; i < v.nCell; i++)
        {
            std::vector<real> mm(v.NDOF, 0.0);
            auto qMax = quadRule(v.cellType[i], INT_ORDER_MAX);
            for (auto &q : qMax)
            {
                GeoPt g = mapPoint(v.cellType[i], v.cellNodes[i], q.p);
                real jd = cellJDet(dim, g);
                v.diffBase(i, toPt(g.x), 1, 1, D);
                for (int k = 0; k < v.NDOF; k++)
                    mm[k] += (D[k] * jd) * q.w;
            }
            for (int k = 0; k < v.NDOF; k++)
                v.moment[i][k] = mm[k] / mm[0];
        }
    }
    // ---- face weights and scales, VariationalReconstruction.cpp:160-427
    v.faceWeight.assign(v.nFace, std::vector<real>(v.maxOrder + 1, 0.0));
    v.faceL.resize(v.nFace);
    v.faceBaryDiff.resize(v.nFace);
    for (idx_t f = 0; f < v.nFace; f++)
    {
        idx_t cL = v.f2c[f][0], cR = v.f2c[f][1];
        Pt bd;
        if (cR >= 0)
        {
            Pt ob = v.otherBaryFromCell(cL, cR, f);
            for (int d = 0; d < 3; d++)
                bd.x[d] = ob.x[d] - v.bary[cL].x[d];
        }
        else
        {
            real vCB[3];
            for (int d = 0; d < 3; d++)
                vCB[d] = v.faceCent[f].x[d] - v.bary[cL].x[d];
            real pr = dot3(vCB, v.faceMeanNorm[f].x);
            for (int d = 0; d < 3; d++)
                bd.x[d] = 2 * pr * v.faceMeanNorm[f].x[d];
        }
        v.faceBaryDiff[f] = bd;
        real l2 = 0;
        for (int d = 0; d < dim; d++)
            l2 += bd.x[d] * bd.x[d];
        v.faceL[f] = std::sqrt(l2) * s->scaleMultiplier;
        real wg = 1;
        if (s->geomWeightScheme == 1)
            wg = s->geomWeightBias + std::pow(std::pow(v.area[f], 1. / real(dim - 1)) / norm3(bd.x), s->geomWeightPower * 0.5);
        std::vector<real> wd(v.maxOrder + 1);
        if (s->dirWeightScheme == 0)
            for (int p = 0; p <= v.maxOrder; p++)
                wd[p] = 1. / factorials[p];
        else
        {
            switch (v.maxOrder)
            {
            case 1: wd[0] = wd[1] = 1; break;
            case 2: wd[0] = 1, wd[1] = 0.4643, wd[2] = 0.1559; break;
            case 3: wd[0] = 1, wd[1] = .5295, wd[2] = wd[3] = .2117; break;
            default: throw std::runtime_error("HQM_OPT needs maxOrder 1..3");
            }
        }
        // boundary faces: id2faceDircWeight(zone, order) * bcWeight, VariationalReconstruction.cpp:438-441
        if (cR < 0)
            for (int p = 0; p <= v.maxOrder; p++)
                wd[p] *= (s->zoneDirWeight ? s->zoneDirWeight[(size_t)v.faceZone[f] * (v.maxOrder + 1) + p] : 1.0) *
                         s->bcWeight;
        for (int p = 0; p <= v.maxOrder; p++)
            v.faceWeight[f][p] = wd[p] * wg;
    }

    // ---- reconstruction coefficients, VariationalReconstruction.cpp:507-825
    int M = v.M, NDOF = v.NDOF, NDIFF = v.NDOF;
    v.AinvB.assign((size_t)v.nCell * (v.maxFC + 1) * M * M, 0.0);
    v.Ainvb.assign((size_t)v.nCell * v.maxFC * M, 0.0);
    auto extract = [&](const std::vector<real> &D, std::vector<real> &out, int nDiff) {
        out.assign((size_t)nDiff * M, 0.0);
        for (int k = 0; k < nDiff; k++)
            for (int j = 0; j < M; j++)
                out[(size_t)k * M + j] = D[(size_t)k * NDOF + 1 + j];
    };
    std::string firstErr;
#pragma omp parallel for schedule(dynamic, 64) num_threads(g_threads)
    for (idx_t i = 0; i < v.nCell; i++)
    {
        std::vector<real> Dfull, DI((size_t)NDIFF * M), DJ((size_t)NDIFF * M);
        int nf = (int)v.c2f[i].size();
        std::vector<real> A((size_t)M * M, 0.0);
        for (int ic = 0; ic < nf; ic++)
        {
            idx_t f = v.c2f[i][ic];
            int if2c = v.cellIsBack(i, f) ? 0 : 1;
            for (size_t g = 0; g < v.faceQP[f].size(); g++)
            {
                v.diffBase(i, v.facePointFromSide(f, if2c, v.faceQP[f][g]), NDIFF, 0, Dfull);
                extract(Dfull, DI, NDIFF);
                v.faceFunctional(f, DI.data(), M, DI.data(), M, NDIFF, A.data(), v.faceJD[f][g] * v.faceQw[g]);
            }
        }
        std::vector<std::vector<real>> Bs(nf), bs(nf);
        for (int ic = 0; ic < nf; ic++)
        {
            idx_t f = v.c2f[i][ic];
            idx_t o = v.cellFaceOther(i, f);
            int if2c = v.cellIsBack(i, f) ? 0 : 1;
            if (o < 0)
                continue;
            std::vector<real> B((size_t)M * M, 0.0);
            for (size_t g = 0; g < v.faceQP[f].size(); g++)
            {
                v.diffBase(i, v.facePointFromSide(f, if2c, v.faceQP[f][g]), NDIFF, 0, Dfull);
                extract(Dfull, DI, NDIFF);
                v.diffBase(o, v.facePointFromSide(f, 1 - if2c, v.faceQP[f][g]), NDIFF, 0, Dfull);
                extract(Dfull, DJ, NDIFF);
                v.faceFunctional(f, DI.data(), M, DJ.data(), M, NDIFF, B.data(), v.faceJD[f][g] * v.faceQw[g]);
            }
            if (o == i)
            {
                for (size_t q = 0; q < A.size(); q++)
                    A[q] -= B[q], B[q] = 0;
            }
            Bs[ic] = B;
        }
        for (int ic = 0; ic < nf; ic++)
        {
            idx_t f = v.c2f[i][ic];
            int if2c = v.cellIsBack(i, f) ? 0 : 1;
            std::vector<real> b(M, 0.0);
            real one = 1.0;
            for (size_t g = 0; g < v.faceQP[f].size(); g++)
            {
                v.diffBase(i, v.facePointFromSide(f, if2c, v.faceQP[f][g]), 1, 0, Dfull);
                extract(Dfull, DI, 1);
                v.faceFunctional(f, DI.data(), M, &one, 1, 1, b.data(), v.faceJD[f][g] * v.faceQw[g]);
            }
            bs[ic] = b;
        }
        std::vector<real> Ainv;
        bool ok = true;
        try
        {
            invertMatrix(M, A, Ainv);
        }
        catch (const std::exception &e)
        {
            ok = false;
#pragma omp critical
            firstErr = e.what();
        }
        if (!ok)
            continue;
        real *blk = &v.AinvB[(size_t)i * (v.maxFC + 1) * M * M];
        std::copy(Ainv.begin(), Ainv.end(), blk);
        for (int ic = 0; ic < nf; ic++)
        {
            idx_t f = v.c2f[i][ic];
            if (v.cellFaceOther(i, f) < 0)
                continue;
            real *bb = blk + (size_t)(1 + ic) * M * M;
            for (int r = 0; r < M; r++)
                for (int c = 0; c < M; c++)
                {
                    real acc = 0;
                    for (int k = 0; k < M; k++)
                        acc += Ainv[(size_t)r * M + k] * Bs[ic][(size_t)k * M + c];
                    bb[(size_t)r * M + c] = acc;
                }
            real *bv = &v.Ainvb[((size_t)i * v.maxFC + ic) * M];
            for (int r = 0; r < M; r++)
            {
                real acc = 0;
                for (int k = 0; k < M; k++)
                    acc += Ainv[(size_t)r * M + k] * bs[ic][k];
                bv[r] = acc;
            }
        }
    }
    if (!firstErr.empty())
        throw std::runtime_error(firstErr);
}


// ---------------------------------------------------------------------------
// Boundary conditions
// ---------------------------------------------------------------------------
// IdealGasThermalConservative2Primitive / Primitive2Conservative, Gas.hpp:279-317
// (three velocity components: the NS models carry rho w also in 2-D)
static void cons2prim(const real *U, real *prim, real gamma)
{
    for (int k = 0; k < 5; k++)
        prim[k] = U[k] / U[0];
    real vSqr = sqr(U[1] / U[0]) + sqr(U[2] / U[0]) + sqr(U[3] / U[0]);
    real rho = U[0], E = U[4];
    prim[0] = rho;
    prim[4] = (gamma - 1) * (E - rho * 0.5 * vSqr);
}
static void prim2cons(const real *prim, real *U, real gamma)
{
    for (int k = 0; k < 5; k++)
        U[k] = prim[k] * prim[0];
    real vSqr = sqr(prim[1]) + sqr(prim[2]) + sqr(prim[3]);
    real rho = prim[0], p = prim[4];
    U[0] = rho;
    U[4] = p / (gamma - 1) + (rho * 0.5 * vSqr);
}

// generateBoundaryValue (EulerEvaluator_EvaluateDt.hxx:1743-1838) for the inviscid types,
// static frame, no CL driver / anchors: generateBV_FarField :1870-1950 (BCFar, BCOutP),
// generateBV_InviscidWall :2199-2215 (BCWallInvis, BCSym), generateBV_ViscousWall :2241-2270
// (BCWall, velocity reversed; BCWallIsothermal, velocity reversed and the density set to
// p / (Rgas T_wall), T_wall = far[0]), generateBV_Outflow :2346-2353, generateBV_Inflow
// :2377-2390, generateBV_TotalConditionInflow :2417-2455 (BCInPsTs: far = (p0, T0, direction);
// the interior speed, clamped to 0.95 of 2 Cp T0 in square, with the static state of the
// given total pressure and temperature).
static void bcGhost(int type, const real *UL, const real *n, const real *far, real gamma, real Rgas, real *UR)
{
    switch (type)
    {
    case ORC_BC_WALL_ISO:
    {
        real prim[5];
        for (int k = 0; k < 5; k++)
            UR[k] = UL[k];
        for (int d = 0; d < 3; d++)
            UR[1 + d] *= -1;
        cons2prim(UR, prim, gamma);
        const real temp = far[0];
        if (!(prim[0] > 0 && prim[4] > 0 && temp > 0))
            throw std::runtime_error("isothermal wall: non-positive density, pressure or wall temperature");
        prim[0] = prim[4] / temp / Rgas;
        prim2cons(prim, UR, gamma);
        break;
    }
    case ORC_BC_IN_PSTS:
    {
        const real Cp = gamma / (gamma - 1) * Rgas;
        real vSqr = (sqr(UL[1]) + sqr(UL[2]) + sqr(UL[3])) / sqr(UL[0]);
        const real pStag = far[0], tStag = far[1];
        vSqr = std::min(vSqr, tStag * 2 * Cp * 0.95);
        const real tStatic = tStag - 0.5 * vSqr / Cp;
        const real pStatic = pStag * std::pow(tStatic / tStag, gamma / (gamma - 1));
        const real rStatic = pStatic / (Rgas * tStatic);
        const real dn = std::sqrt(sqr(far[2]) + sqr(far[3]) + sqr(far[4]));
        real prim[5] = {rStatic, 0, 0, 0, pStatic};
        for (int d = 0; d < 3; d++)
            prim[1 + d] = far[2 + d] / dn * std::sqrt(vSqr);
        prim2cons(prim, UR, gamma);
        break;
    }
    case ORC_BC_FAR:
    case ORC_BC_OUTP:
    {
        real un = (UL[1] * n[0] + UL[2] * n[1] + UL[3] * n[2]) / UL[0];
        real p, asqr, H;
        thermal(UL[4], UL[0], sqr(UL[1] / UL[0]) + sqr(UL[2] / UL[0]) + sqr(UL[3] / UL[0]), gamma, p, asqr, H);
        if (!(asqr >= 0))
            throw std::runtime_error("far-field ghost: negative a^2");
        real a = std::sqrt(asqr);
        const real vgN = 0;
        if (un - vgN - a > 0) // supersonic outflow
            for (int k = 0; k < 5; k++)
                UR[k] = UL[k];
        else if (un - vgN > 0) // subsonic outflow: far pressure
        {
            real fp[5], lp[5];
            cons2prim(far, fp, gamma);
            cons2prim(UL, lp, gamma);
            lp[4] = fp[4];
            prim2cons(lp, UR, gamma);
        }
        else if (un - vgN + a > 0) // subsonic inflow: interior pressure
        {
            real fp[5], lp[5];
            cons2prim(far, fp, gamma);
            cons2prim(UL, lp, gamma);
            fp[4] = lp[4];
            prim2cons(fp, UR, gamma);
        }
        else // supersonic inflow
            for (int k = 0; k < 5; k++)
                UR[k] = far[k];
        break;
    }
    case ORC_BC_WALL_INVIS:
    case ORC_BC_SYM:
    {
        for (int k = 0; k < 5; k++)
            UR[k] = UL[k];
        real vn = UR[1] * n[0] + UR[2] * n[1] + UR[3] * n[2];
        for (int d = 0; d < 3; d++)
            UR[1 + d] -= 2 * vn * n[d];
        break;
    }
    case ORC_BC_WALL:
        for (int k = 0; k < 5; k++)
            UR[k] = UL[k];
        for (int d = 0; d < 3; d++)
            UR[1 + d] *= -1;
        break;
    case ORC_BC_OUT:
        for (int k = 0; k < 5; k++)
            UR[k] = UL[k];
        break;
    case ORC_BC_IN:
        for (int k = 0; k < 5; k++)
            UR[k] = far[k];
        break;
    default:
        throw std::runtime_error("boundary type not supported on this path");
    }
}

// CompressRecPart, EulerEvaluator.hpp:1541-1631: the reconstructed point state, or the
// cell mean where it would carry negative density or internal energy
static void compressRecPart(const real *umean, const real *inc, real *ret)
{
    for (int k = 0; k < 5; k++)
        ret[k] = umean[k] + inc[k];
    if (ret[0] < 0)
        for (int k = 0; k < 5; k++)
            ret[k] = umean[k];
    real eK = (sqr(ret[1]) + sqr(ret[2]) + sqr(ret[3])) * 0.5 / (verySmallReal + ret[0]);
    if (ret[4] - eK < 0)
        for (int k = 0; k < 5; k++)
            ret[k] = umean[k];
}

// Boundary term of the VR sweep for boundary face f of cell i (GetBoundaryRHS,
// VariationalReconstruction_Reconstruction.hxx:19-96, intOrderVRBC = intOrder):
//   BCC += sum_g FFaceFunctional(D0phi(x_g), uB(x_g) - u_i) JDet w_g
// with uB the ghost of the compressed point state (Euler BC table, nVars = 5; the
// solver's FBoundary, EulerSolver.hxx:163-175) or the given point value.  With
// uRecDiff the linearised form (GetBoundaryRHSDiff :98-176, FBoundaryDiff
// EulerSolver.hxx:176-193): uB(uBL + dU) - uB(uBL), no subtraction of u_i.
static void boundaryBCC(const VR &v, idx_t i, idx_t f, int nVars, const real *uRec, const real *u,
                        const real *uRecDiff, std::vector<real> &BCC)
{
    const int M = v.M;
    const size_t blkN = (size_t)M * nVars, G = v.faceQw.size();
    const int zone = v.faceZone[f];
    const bool euler = !v.bcType.empty() && nVars == 5;
    if (!euler && !(v.dirNVars == nVars && !v.dirValue.empty()))
        throw std::runtime_error("boundary values are not set for this number of variables");
    std::vector<real> D, inc(nVars), phi(M);
    for (size_t g = 0; g < G; g++)
    {
        v.diffBase(i, v.faceQP[f][g], 1, 0, D);
        for (int k = 0; k < M; k++)
            phi[k] = D[1 + k];
        real uBL[5], dU[5], uB[5];
        for (int var = 0; var < nVars; var++)
        {
            real r = 0, rd = 0;
            for (int k = 0; k < M; k++)
            {
                r += phi[k] * uRec[i * blkN + (size_t)var * M + k];
                if (uRecDiff)
                    rd += phi[k] * uRecDiff[i * blkN + (size_t)var * M + k];
            }
            if (var < 5)
                uBL[var] = r, dU[var] = rd;
            inc[var] = r;
        }
        if (euler)
        {
            const real *n = v.faceN[f][g].x;
            const real *far = v.bcValue[zone].data();
            real ULf[5];
            compressRecPart(u + i * 5, uBL, ULf);
            bcGhost(v.bcType[zone], ULf, n, far, v.bcGamma, v.bcRgas, uB);
            if (uRecDiff)
            {
                real inc2[5], ULp[5], uBp[5];
                for (int k = 0; k < 5; k++)
                    inc2[k] = uBL[k] + dU[k];
                compressRecPart(u + i * 5, inc2, ULp);
                bcGhost(v.bcType[zone], ULp, n, far, v.bcGamma, v.bcRgas, uBp);
                for (int k = 0; k < 5; k++)
                    inc[k] = uBp[k] - uB[k];
            }
            else
                for (int k = 0; k < 5; k++)
                    inc[k] = uB[k] - u[i * 5 + k];
        }
        else
            for (int var = 0; var < nVars; var++)
                inc[var] = uRecDiff ? 0.0 : v.dirValue[((size_t)f * G + g) * nVars + var] - u[i * nVars + var];
        v.faceFunctional(f, phi.data(), M, inc.data(), nVars, 1, BCC.data(), v.faceJD[f][g] * v.faceQw[g]);
    }
}

// ---------------------------------------------------------------------------
// Analytic fields (src/Euler/SpecialFields.hpp:41-75, EulerEvaluator.hxx:1039-1135,
// test/cpp/CFV/test_Reconstruction.cpp:443-447, test_Reconstruction3D.cpp:321-325)
// ---------------------------------------------------------------------------
static void evalField(int id, const double *prm, const Pt &p, int nVars, real *out)
{
    for (int v = 0; v < nVars; v++)
        out[v] = 0;
    switch (id)
    {
    case 0: // constant: prm[0..nVars)
        for (int v = 0; v < nVars; v++)
            out[v] = prm[v];
        return;
    case 1: // linear a + b.x
        out[0] = prm[0] + prm[1] * p.x[0] + prm[2] * p.x[1] + prm[3] * p.x[2];
        return;
    case 2: // sin(kx)cos(ky)[cos(kz)] with k = 2 pi / L; prm = {L, dim}
    {
        real k = 2.0 * pi / prm[0];
        out[0] = std::sin(k * p.x[0]) * std::cos(k * p.x[1]) * (prm[1] == 3 ? std::cos(k * p.x[2]) : 1.0);
        return;
    }
    case 3: // cos(kx)+cos(ky)[+cos(kz)]
    {
        real k = 2.0 * pi / prm[0];
        out[0] = std::cos(k * p.x[0]) + std::cos(k * p.x[1]) + (prm[1] == 3 ? std::cos(k * p.x[2]) : 0.0);
        return;
    }
    case 4: // quadratic x^2 + y^2 (+ z^2)
        out[0] = p.x[0] * p.x[0] + p.x[1] * p.x[1] + prm[0] * p.x[2] * p.x[2];
        return;
    case 5: // cubic x^3 + x y^2 (test_Reconstruction.cpp:504-507) (+ z^3)
        out[0] = p.x[0] * p.x[0] * p.x[0] + p.x[0] * p.x[1] * p.x[1] + prm[0] * p.x[2] * p.x[2] * p.x[2];
        return;
    case 10: // isentropic vortex on [0,10]^2: prm = {chi, gamma, ubx, uby}
    {
        if (nVars < 5)
            throw std::runtime_error("isentropic vortex needs 5 variables");
        real chi = prm[0], gamma = prm[1], ubx = prm[2], uby = prm[3];
        real xyc = 5;
        real px = p.x[0], py = p.x[1];
        real r = std::sqrt(sqr(px - xyc) + sqr(py - xyc));
        real dT = -(gamma - 1) / (8 * gamma * sqr(pi)) * sqr(chi) * std::exp(1 - sqr(r));
        real dux = chi / 2 / pi * std::exp((1 - sqr(r)) / 2) * -(py - xyc);
        real duy = chi / 2 / pi * std::exp((1 - sqr(r)) / 2) * +(px - xyc);
        real T = dT + 1, ux = dux + ubx, uy = duy + uby, S = 1;
        real rho = std::pow(T / S, 1 / (gamma - 1));
        real pr = T * rho;
        real E = 0.5 * (sqr(ux) + sqr(uy)) * rho + pr / (gamma - 1);
        out[0] = rho, out[1] = rho * ux, out[2] = rho * uy, out[3] = 0, out[4] = E;
        return;
    }
    case 11: // Taylor-Green vortex: prm = {M0, gamma}
    {
        if (nVars < 5)
            throw std::runtime_error("Taylor-Green vortex needs 5 variables");
        real M0 = prm[0], gamma = prm[1];
        real x = p.x[0], y = p.x[1], z = p.x[2];
        real ux = std::sin(x) * std::cos(y) * std::cos(z);
        real uy = -std::cos(x) * std::sin(y) * std::cos(z);
        real pr = 1. / (gamma * sqr(M0)) + 1. / 16 * ((std::cos(2 * x) + std::cos(2 * y)) * (2 + std::cos(2 * z)));
        real rho = gamma * sqr(M0) * pr;
        real E = 0.5 * (sqr(ux) + sqr(uy)) * rho + pr / (gamma - 1);
        out[0] = rho, out[1] = rho * ux, out[2] = rho * uy, out[3] = 0, out[4] = E;
        return;
    }
    default: throw std::runtime_error("unknown field id");
    }
}

// ---------------------------------------------------------------------------
// Positivity of the increment (ideal gas, nVars = 5).
// ---------------------------------------------------------------------------
static real sq3(const real *a) { return a[0] * a[0] + a[1] * a[1] + a[2] * a[2]; }
static real signPlus(real a) { return a >= 0 ? 1 : -1; } // signP, src/DNDS/Defines.hpp:549-552

// IdealGasGetCompressionRatioPressure<dim, 0> (src/Euler/Gas.hpp:1750-1860): the largest alpha in
// [0, 1] (times the safety ratio, then decayed) with rho E_internal(u + alpha inc) >= newRhoEi.
// bad is set where the reference asserts.
static real compressionRatioPressure(const real *u, const real *inc, real newRhoEi, bool &bad)
{
    static const real safetyRatio = 1 - 1e-5;
    real ret[5], un[5];
    for (int k = 0; k < 5; k++)
        ret[k] = inc[k], un[k] = u[k] + inc[k];
    real rhoEOld = u[4] - sq3(u + 1) / (u[0] + verySmallReal) * 0.5;
    newRhoEi = std::max(smallReal * rhoEOld, newRhoEi);
    real rhoENew = un[4] - sq3(un + 1) / (un[0] + verySmallReal) * 0.5;
    real alpha = (rhoEOld - newRhoEi) / std::max(-rhoENew + rhoEOld, verySmallReal);
    if (rhoENew > rhoEOld)
        alpha = 1;
    alpha = std::max(std::min(alpha, 1.), 0.); // the convex estimate, superseded below
    real c0 = 2 * u[4] * u[0] - sq3(u + 1) - 2 * u[0] * newRhoEi;
    real c1 = 2 * u[4] * ret[0] + 2 * u[0] * ret[4] - 2 * dot3(u + 1, ret + 1) - 2 * ret[0] * newRhoEi;
    real c2 = 2 * ret[4] * ret[0] - sq3(ret + 1);
    c2 += signPlus(c2) * verySmallReal;
    real deltaC = c1 * c1 - 4 * c0 * c2;
    if (deltaC <= -c0 * c0 * smallReal)
        bad = true;
    deltaC = std::max(0., deltaC);
    real alphaL = (-std::sqrt(deltaC) - c1) / (2 * c2);
    if (std::abs(c2) < 1e-10 * c0)
        alpha = std::abs(c1) < 1e-10 * c0 ? 0 : std::min(-c0 / c1, 1.);
    else
        alpha = std::min(alphaL, 1.);
    alpha = std::max(0., alpha);
    alpha *= safetyRatio;
    if (alpha < smallReal)
        alpha = 0;
    for (int k = 0; k < 5; k++)
        ret[k] *= alpha;
    const real decay = 1 - 1e-2;
    int iter;
    for (iter = 0; iter < 1000; iter++)
    {
        real uv[3] = {u[1] + ret[1], u[2] + ret[2], u[3] + ret[3]};
        real ek = 0.5 * sq3(uv) / (u[0] + ret[0] + verySmallReal);
        if (ret[4] + u[4] - ek < newRhoEi)
        {
            for (int k = 0; k < 5; k++)
                ret[k] *= decay;
            alpha *= decay;
        }
        else
            break;
    }
    if (iter >= 1000)
        bad = true;
    return alpha;
}

// EvaluateCellRHSAlpha for one cell (src/Euler/EulerEvaluator.hxx:1938-2030; flag DEFAULT,
// uRecAlphaCompressPower 1): alpha = (alphaRho alphaP)^compress * safetyRatio when below 1.
// Returns alpha; limited tells whether the cell was limited; bad where the reference asserts.
static real cellRHSAlpha(const real *u, const real *res, real gamma, real relax, real rhoEps, real pEps, int compress,
                         bool &limited, bool &bad)
{
    static const real safetyRatio = 1 - 1e-5;
    real inc[5];
    for (int k = 0; k < 5; k++)
        inc[k] = res[k];
    if (!(u[0] >= rhoEps))
        bad = true;
    real alphaRho = 1;
    real relaxedRho = rhoEps * relax + u[0] * (1 - relax);
    if (inc[0] < 0)
        alphaRho = std::min(1.0, (u[0] - relaxedRho) / (-inc[0] - smallReal * inc[0]));
    if (!(alphaRho >= 0 && alphaRho <= 1))
        bad = true;
    for (int k = 0; k < 5; k++)
        inc[k] *= alphaRho;
    real un[5];
    for (int k = 0; k < 5; k++)
        un[k] = u[k] + inc[k];
    real pNew = (un[4] - 0.5 * sq3(un + 1) / un[0]) * (gamma - 1);
    real pOld = (u[4] - 0.5 * sq3(u + 1) / u[0]) * (gamma - 1);
    real relaxedP = pEps;
    if (pNew < pOld)
        relaxedP = pEps + (pOld - pEps) * (1 - relax);
    real alphaP = 1;
    if (pNew < relaxedP)
        alphaP = std::min(alphaP, compressionRatioPressure(u, inc, relaxedP / (gamma - 1), bad));
    real alpha = alphaRho * alphaP;
    limited = alpha < 1;
    if (limited)
        alpha = std::pow(alpha, compress) * safetyRatio;
    return alpha;
}

// CompressInc (src/Euler/EulerEvaluator.hpp:1645-1760, ideal-gas part): scales the increment so
// that density stays above rhoEps and rho E_internal above pEps / (gamma - 1).
static void compressInc(const real *u, const real *inc, real gamma, real rhoEps, real pEps, real *ret, bool &bad)
{
    for (int k = 0; k < 5; k++)
        ret[k] = inc[k];
    if (!(u[0] > 0))
        bad = true;
    if (u[0] + ret[0] <= rhoEps)
    {
        real newrho = rhoEps;
        real f = (newrho - u[0]) / (ret[0] - verySmallReal);
        for (int k = 0; k < 5; k++)
            ret[k] *= f;
    }
    real ekOld = 0.5 * sq3(u + 1) / (u[0] + verySmallReal);
    real rhoEi = u[4] - ekOld;
    if (!(rhoEi > 0))
        bad = true;
    real uv[3] = {u[1] + ret[1], u[2] + ret[2], u[3] + ret[3]};
    real ek = 0.5 * sq3(uv) / (u[0] + ret[0] + verySmallReal);
    real rhoEiNew = u[4] + ret[4] - ek;
    if (rhoEiNew <= pEps)
    {
        real newRhoEi = pEps / (gamma - 1);
        real c0 = 2 * u[4] * u[0] - sq3(u + 1) - 2 * u[0] * newRhoEi;
        real c1 = 2 * u[4] * ret[0] + 2 * u[0] * ret[4] - 2 * dot3(u + 1, ret + 1) - 2 * ret[0] * newRhoEi;
        real c2 = 2 * ret[4] * ret[0] - sq3(ret + 1);
        real deltaC = c1 * c1 - 4 * c0 * c2;
        if (!(deltaC > 0))
            bad = true;
        real alphaL = (-std::sqrt(deltaC) - c1) / (2 * c2);
        real alpha = std::max(0., std::min(alphaL, 1.));
        for (int k = 0; k < 5; k++)
            ret[k] *= alpha * 0.99;
        const real decay = 1 - 1e-1;
        for (int iter = 0; iter < 1000; iter++)
        {
            real w[3] = {u[1] + ret[1], u[2] + ret[2], u[3] + ret[3]};
            ek = 0.5 * sq3(w) / (u[0] + ret[0] + verySmallReal);
            if (ret[4] + u[4] - ek < newRhoEi)
                for (int k = 0; k < 5; k++)
                    ret[k] *= decay;
            else
                break;
        }
        real w[3] = {u[1] + ret[1], u[2] + ret[2], u[3] + ret[3]};
        ek = 0.5 * sq3(w) / (u[0] + ret[0] + verySmallReal);
        if (ret[4] + u[4] - ek < newRhoEi * 0.5)
            bad = true;
    }
}

// fincrement (src/Euler/EulerSolver.hxx:696-738): alpha_i = EvaluateCellRHSAlpha(u, inc, relax,
// compress 2, DEFAULT) per cell, then AddFixedIncrement (EulerEvaluator.hpp:1795-1830):
// u_i += CompressInc(u_i, alpha_i inc_i).  Returns the number of limited cells; alphaMin gets
// the smallest alpha (1 when none); bad when a cell hits one of the reference's assertions.
static idx_t addIncrementPP(idx_t N, real gamma, real relax, real rhoEps, real pEps, real *u, const real *du,
                            real &alphaMin, bool &bad)
{
    idx_t nLim = 0;
    alphaMin = 1;
    for (idx_t i = 0; i < N; i++)
    {
        real *U = u + i * 5;
        bool lim = false;
        real a = cellRHSAlpha(U, du + i * 5, gamma, relax, rhoEps, pEps, 2, lim, bad);
        if (lim)
            nLim++, alphaMin = std::min(alphaMin, a);
        real d[5], r[5];
        for (int k = 0; k < 5; k++)
            d[k] = du[i * 5 + k] * a;
        compressInc(U, d, gamma, rhoEps, pEps, r, bad);
        for (int k = 0; k < 5; k++)
            U[k] += r[k];
    }
    return nLim;
}

} // namespace orc

using namespace orc;

struct orc_vr
{
    VR v;
};

extern "C" {

const char *orc_last_error(void) { return g_err.c_str(); }

int orc_set_threads(int nThreads, int rhsGather)
{
    g_threads = nThreads > 0 ? nThreads : 1;
    g_rhsGather = rhsGather;
    return 0;
}

#define ORC_TRY try {
#define ORC_CATCH(ret)                     \
    }                                      \
    catch (const std::exception &e)        \
    {                                      \
        g_err = e.what();                  \
        return ret;                        \
    }

orc_vr *orc_vr_create(const orc_mesh_desc *mesh, const orc_vr_settings *s)
{
    ORC_TRY
    auto *h = new orc_vr;
    buildVR(h->v, mesh, s);
    return h;
    ORC_CATCH(nullptr)
}

void orc_vr_destroy(orc_vr *h) { delete h; }

int orc_vr_info(const orc_vr *h, int64_t *nCell, int64_t *nFace, int *M, int *maxFaceCell,
                int *nGFace, int *nGCell, double *globalVol)
{
    const VR &v = h->v;
    *nCell = v.nCell, *nFace = v.nFace, *M = v.M, *maxFaceCell = v.maxFC;
    size_t gc = 0;
    for (auto &w : v.cellQw)
        gc = std::max(gc, w.size());
    *nGFace = (int)v.faceQw.size(), *nGCell = (int)gc, *globalVol = v.globalVol;
    return 0;
}

int orc_get_cell_metrics(const orc_vr *h, double *vol, double *bary, double *cent, double *scale,
                         double *smoothRatio, double *moment)
{
    const VR &v = h->v;
    for (idx_t i = 0; i < v.nCell; i++)
    {
        vol[i] = v.vol[i];
        for (int d = 0; d < 3; d++)
            bary[3 * i + d] = v.bary[i].x[d], cent[3 * i + d] = v.cent[i].x[d];
        real sc = v.dim == 2 ? std::max(v.hbox[i].x[0], v.hbox[i].x[1])
                             : std::max({v.hbox[i].x[0], v.hbox[i].x[1], v.hbox[i].x[2]});
        scale[i] = sc;
        smoothRatio[i] = v.smooth[i];
        for (int k = 0; k < v.M; k++)
            moment[(size_t)i * v.M + k] = v.moment[i][1 + k];
    }
    return 0;
}

int orc_get_face_metrics(const orc_vr *h, double *area, double *meanNorm, double *pts, double *norms,
                         double *wJDet, double *faceL, double *faceWeight)
{
    const VR &v = h->v;
    size_t G = v.faceQw.size();
    for (idx_t f = 0; f < v.nFace; f++)
    {
        area[f] = v.area[f];
        faceL[f] = v.faceL[f];
        for (int d = 0; d < 3; d++)
            meanNorm[3 * f + d] = v.faceMeanNorm[f].x[d];
        for (size_t g = 0; g < G; g++)
        {
            for (int d = 0; d < 3; d++)
                pts[(f * G + g) * 3 + d] = v.faceQP[f][g].x[d], norms[(f * G + g) * 3 + d] = v.faceN[f][g].x[d];
            wJDet[f * G + g] = v.faceJD[f][g] * v.faceQw[g];
        }
        for (int o = 0; o <= v.maxOrder; o++)
            faceWeight[f * (v.maxOrder + 1) + o] = v.faceWeight[f][o];
    }
    return 0;
}

int orc_get_cell_quad(const orc_vr *h, double *pts, double *wJDet)
{
    const VR &v = h->v;
    size_t G = 0;
    for (auto &w : v.cellQw)
        G = std::max(G, w.size());
    // a cell with a smaller rule than the largest is padded with zero points and weights
    for (idx_t i = 0; i < v.nCell; i++)
        for (size_t g = 0; g < G; g++)
        {
            const bool in = g < v.cellQw[i].size();
            for (int d = 0; d < 3; d++)
                pts[(i * G + g) * 3 + d] = in ? v.cellQP[i][g].x[d] : 0.0;
            wJDet[i * G + g] = in ? v.cellJD[i][g] * v.cellQw[i][g] : 0.0;
        }
    return 0;
}

int orc_get_vr_coeffs(const orc_vr *h, double *AinvB, double *Ainvb)
{
    const VR &v = h->v;
    std::copy(v.AinvB.begin(), v.AinvB.end(), AinvB);
    std::copy(v.Ainvb.begin(), v.Ainvb.end(), Ainvb);
    return 0;
}

// cell averages by cell quadrature: EulerEvaluator.hxx:1046-1068, test_Reconstruction.cpp:246-259
int orc_cell_average(const orc_vr *h, int fieldId, const double *params, int nVars, double *u)
{
    ORC_TRY
    const VR &v = h->v;
    std::vector<real> val(nVars);
    for (idx_t i = 0; i < v.nCell; i++)
    {
        std::vector<real> acc(nVars, 0.0);
        for (size_t g = 0; g < v.cellQw[i].size(); g++)
        {
            evalField(fieldId, params, v.cellQP[i][g], nVars, val.data());
            for (int k = 0; k < nVars; k++)
                acc[k] += (val[k] * v.cellJD[i][g]) * v.cellQw[i][g];
        }
        for (int k = 0; k < nVars; k++)
            u[i * nVars + k] = acc[k] / v.vol[i];
    }
    return 0;
    ORC_CATCH(-1)
}

// L1 error of the reconstruction at cell quadrature points / domain volume,
// test_Reconstruction.cpp:330-353
double orc_rec_error_l1(const orc_vr *h, int fieldId, const double *params, const double *u, const double *uRec)
{
    const VR &v = h->v;
    int M = v.M;
    std::vector<real> D;
    real err = 0;
    for (idx_t i = 0; i < v.nCell; i++)
    {
        real ec = 0;
        for (size_t g = 0; g < v.cellQw[i].size(); g++)
        {
            v.diffBase(i, v.cellQP[i][g], 1, 0, D);
            real r = 0;
            for (int k = 0; k < M; k++)
                r += D[1 + k] * uRec[(size_t)i * M + k];
            real ex;
            evalField(fieldId, params, v.cellQP[i][g], 1, &ex);
            ec += (std::abs(r + u[i] - ex) * v.cellJD[i][g]) * v.cellQw[i][g];
        }
        err += ec;
    }
    return err / v.globalVol;
}

// Second-order substitute of the VR sweep (maxOrder = 1, subs2ndOrder != 0):
// DoReconstruction2ndGrad + DoReconstruction2nd,
// VariationalReconstruction_Reconstruction.hxx:184-483.  Per cell a gradient g (nVars x dim)
//   subs2ndOrder 1 (Gauss rule): g = (1/V) sum_f (u_j - u_i) lThis / (lThis + lThat) A_f n_f
//     with lThis = |x_f - bary_i|, lThat = |bary_j - x_f| (both 1 with subs2ndOrderGGScheme 0),
//     n_f the mean normal out of the cell, x_f the face centre; a boundary face adds
//     (uB - u_i) A_f n_f / 2 with uB the ghost of u_i at the face centre (mean normal)
//   subs2ndOrder 2 (least squares): g = argmin sum_f |dx_f g^T - du_f|^2 with dx_f = bary_j -
//     bary_i, du_f = u_j - u_i, and at a boundary face dx_f = -2 ((bary_i - x_f).n) n,
//     du_f = uB - u_i (solved here by Householder QR, the minimum-norm solution for full rank)
// then uRec_i (first dim modes) = D1phi^-1 g^T, D1phi the first derivatives of the linear
// modes (partial-pivot elimination).
static void vr2ndCell(const VR &v, idx_t i, int nVars, const real *u, real *out)
{
    const int dim = v.dim, M = v.M;
    const int method = v.s.subs2ndOrder;
    const bool euler = !v.bcType.empty() && nVars == 5;
    const idx_t nf = (idx_t)v.c2f[i].size();
    std::vector<real> grad((size_t)nVars * dim, 0.0);    // grad[var * dim + d]
    std::vector<real> dcs((size_t)nf * dim), dus((size_t)nf * nVars);
    auto ghost = [&](idx_t f, real *uB) {
        if (!euler)
            throw std::runtime_error("the second-order substitute needs the Euler boundary table (nVars = 5) on "
                                     "meshes with boundary faces");
        const int zone = v.faceZone[f];
        bcGhost(v.bcType[zone], u + i * 5, v.faceMeanNorm[f].x, v.bcValue[zone].data(), v.bcGamma, v.bcRgas, uB);
    };
    for (idx_t ic = 0; ic < nf; ic++)
    {
        const idx_t f = v.c2f[i][ic];
        const idx_t o = v.cellFaceOther(i, f);
        const bool back = v.cellIsBack(i, f);
        Pt xf = v.faceCent[f];
        if (!back)
            for (int d = 0; d < 3; d++)
                xf.x[d] += v.fshift[f].x[d];
        const real *n = v.faceMeanNorm[f].x;
        const real area = v.area[f];
        real uB[5];
        if (o < 0)
            ghost(f, uB);
        if (method == 1)
        {
            real lThis = 0;
            for (int d = 0; d < 3; d++)
                lThis += sqr(xf.x[d] - v.bary[i].x[d]);
            lThis = v.s.subs2ndOrderGGScheme == 0 ? 1.0 : std::sqrt(lThis);
            if (o >= 0)
            {
                const Pt ob = v.otherBaryFromCell(i, o, f);
                real lThat = 0;
                for (int d = 0; d < 3; d++)
                    lThat += sqr(ob.x[d] - xf.x[d]);
                lThat = v.s.subs2ndOrderGGScheme == 0 ? 1.0 : std::sqrt(lThat);
                const real sgn = back ? 1.0 : -1.0;
                for (int var = 0; var < nVars; var++)
                    for (int d = 0; d < dim; d++)
                        grad[(size_t)var * dim + d] += (u[o * nVars + var] - u[i * nVars + var]) * lThis /
                                                       (lThis + lThat + verySmallReal) * area * (n[d] * sgn);
            }
            else
                for (int var = 0; var < nVars; var++)
                    for (int d = 0; d < dim; d++)
                        grad[(size_t)var * dim + d] += (uB[var] - u[i * nVars + var]) * 0.5 * area * n[d];
        }
        else
        {
            if (o >= 0)
            {
                const Pt ob = v.otherBaryFromCell(i, o, f);
                for (int d = 0; d < dim; d++)
                    dcs[(size_t)ic * dim + d] = ob.x[d] - v.bary[i].x[d];
                for (int var = 0; var < nVars; var++)
                    dus[(size_t)ic * nVars + var] = u[o * nVars + var] - u[i * nVars + var];
            }
            else
            {
                real pr = 0;
                for (int d = 0; d < 3; d++)
                    pr += (v.bary[i].x[d] - v.faceCent[f].x[d]) * n[d];
                for (int d = 0; d < dim; d++)
                    dcs[(size_t)ic * dim + d] = pr * -2.0 * n[d];
                for (int var = 0; var < nVars; var++)
                    dus[(size_t)ic * nVars + var] = uB[var] - u[i * nVars + var];
            }
        }
    }
    if (method == 1)
        for (auto &g : grad)
            g /= v.vol[i];
    else
    {
        // Householder QR of dcs (nf x dim), applied to dus; back substitution
        std::vector<real> A(dcs), B(dus);
        for (int k = 0; k < dim; k++)
        {
            real nrm = 0;
            for (idx_t r = k; r < nf; r++)
                nrm += sqr(A[(size_t)r * dim + k]);
            nrm = std::sqrt(nrm);
            if (nrm == 0)
                throw std::runtime_error("least-squares gradient: rank-deficient stencil");
            const real alpha = A[(size_t)k * dim + k] > 0 ? -nrm : nrm;
            std::vector<real> w(nf, 0.0);
            for (idx_t r = k; r < nf; r++)
                w[r] = A[(size_t)r * dim + k];
            w[k] -= alpha;
            real wn = 0;
            for (idx_t r = k; r < nf; r++)
                wn += w[r] * w[r];
            if (wn == 0)
                continue;
            auto reflect = [&](real *col, int stride, int ncol) {
                for (int c = 0; c < ncol; c++)
                {
                    real dot = 0;
                    for (idx_t r = k; r < nf; r++)
                        dot += w[r] * col[(size_t)r * stride + c];
                    const real fct = 2 * dot / wn;
                    for (idx_t r = k; r < nf; r++)
                        col[(size_t)r * stride + c] -= fct * w[r];
                }
            };
            reflect(A.data(), dim, dim);
            reflect(B.data(), nVars, nVars);
        }
        for (int var = 0; var < nVars; var++)
            for (int k = dim - 1; k >= 0; k--)
            {
                real acc = B[(size_t)k * nVars + var];
                for (int c = k + 1; c < dim; c++)
                    acc -= A[(size_t)k * dim + c] * grad[(size_t)var * dim + c];
                grad[(size_t)var * dim + k] = acc / A[(size_t)k * dim + k];
            }
    }
    // D1phi (dim x dim): rows d/dx_d, columns the linear modes; solve D1phi X = g^T
    std::vector<real> D;
    v.diffBase(i, v.bary[i], 1 + dim, 1, D);
    real L[3][3], R[3][5];
    for (int d = 0; d < dim; d++)
    {
        for (int e = 0; e < dim; e++)
            L[d][e] = D[(size_t)(1 + d) * v.NDOF + 1 + e];
        for (int var = 0; var < nVars; var++)
            R[d][var] = grad[(size_t)var * dim + d];
    }
    for (int k = 0; k < dim; k++)
    {
        int piv = k;
        for (int r = k + 1; r < dim; r++)
            if (std::fabs(L[r][k]) > std::fabs(L[piv][k]))
                piv = r;
        for (int c = 0; c < dim; c++)
            std::swap(L[k][c], L[piv][c]);
        for (int var = 0; var < nVars; var++)
            std::swap(R[k][var], R[piv][var]);
        for (int r = k + 1; r < dim; r++)
        {
            const real fct = L[r][k] / L[k][k];
            for (int c = k; c < dim; c++)
                L[r][c] -= fct * L[k][c];
            for (int var = 0; var < nVars; var++)
                R[r][var] -= fct * R[k][var];
        }
    }
    for (int q = 0; q < M * nVars; q++)
        out[q] = 0;
    for (int var = 0; var < nVars; var++)
        for (int k = dim - 1; k >= 0; k--)
        {
            real acc = R[k][var];
            for (int c = k + 1; c < dim; c++)
                acc -= L[k][c] * out[(size_t)var * M + c];
            out[(size_t)var * M + k] = acc / L[k][k];
        }
}

int orc_vr_2nd(const orc_vr *h, int nVars, const double *u, double *uRec)
{
    ORC_TRY
    const VR &v = h->v;
    if (v.s.maxOrder != 1 || (v.s.subs2ndOrder != 1 && v.s.subs2ndOrder != 2))
        throw std::runtime_error("the second-order substitute needs maxOrder 1 and subs2ndOrder 1 or 2");
    if (nVars > 5)
        throw std::runtime_error("nVars must be at most 5");
    const size_t blkN = (size_t)v.M * nVars;
    std::string firstErr;
#pragma omp parallel for schedule(static) num_threads(g_threads)
    for (idx_t i = 0; i < v.nCell; i++)
    {
        try
        {
            vr2ndCell(v, i, nVars, u, uRec + i * blkN);
        }
        catch (const std::exception &e)
        {
#pragma omp critical
            firstErr = e.what();
        }
    }
    if (!firstErr.empty())
        throw std::runtime_error(firstErr);
    return 0;
    ORC_CATCH(-1)
}

static bool useSubs2nd(const VR &v) { return v.s.maxOrder == 1 && v.s.subs2ndOrder != 0; }

// DoReconstructionIter (Jacobi branch, SORInstead = false),
// VariationalReconstruction_Reconstruction.hxx:485-615; boundary faces add A^-1 BCC
static void addAinvBCC(const VR &v, idx_t i, int nVars, const std::vector<real> &BCC, real scale, real *out)
{
    const int M = v.M;
    const real *Ainv = &v.AinvB[(size_t)i * (v.maxFC + 1) * M * M];
    for (int var = 0; var < nVars; var++)
        for (int r = 0; r < M; r++)
        {
            real acc = 0;
            for (int k = 0; k < M; k++)
                acc += Ainv[(size_t)r * M + k] * BCC[(size_t)k * nVars + var];
            out[(size_t)var * M + r] += scale * acc;
        }
}

int orc_vr_iter(const orc_vr *h, int nVars, const double *uRec, const double *u, double *uRecNew)
{
    ORC_TRY
    const VR &v = h->v;
    if (useSubs2nd(v)) // :503-512
        return orc_vr_2nd(h, nVars, u, uRecNew);
    int M = v.M;
    size_t blkN = (size_t)M * nVars;
    std::string firstErr;
#pragma omp parallel for schedule(static) num_threads(g_threads)
    for (idx_t i = 0; i < v.nCell; i++)
    {
        real relax = v.s.jacobiRelax;
        real *out = uRecNew + i * blkN;
        for (size_t q = 0; q < blkN; q++)
            out[q] = uRec[i * blkN + q] * (1 - relax);
        const real *blk = &v.AinvB[(size_t)i * (v.maxFC + 1) * M * M];
        for (size_t ic = 0; ic < v.c2f[i].size(); ic++)
        {
            idx_t f = v.c2f[i][ic];
            idx_t o = v.cellFaceOther(i, f);
            if (o < 0)
            {
                std::vector<real> BCC(blkN, 0.0);
                try
                {
                    boundaryBCC(v, i, f, nVars, uRec, u, nullptr, BCC);
                }
                catch (const std::exception &e)
                {
#pragma omp critical
                    firstErr = e.what();
                }
                addAinvBCC(v, i, nVars, BCC, relax, out);
                continue;
            }
            const real *B = blk + (1 + ic) * (size_t)M * M;
            const real *bv = &v.Ainvb[((size_t)i * v.maxFC + ic) * M];
            for (int var = 0; var < nVars; var++)
            {
                real du = u[o * nVars + var] - u[i * nVars + var];
                for (int r = 0; r < M; r++)
                {
                    real acc = 0;
                    for (int k = 0; k < M; k++)
                        acc += B[(size_t)r * M + k] * uRec[o * blkN + (size_t)var * M + k];
                    acc += bv[r] * du;
                    out[(size_t)var * M + r] += relax * acc;
                }
            }
        }
    }
    if (!firstErr.empty())
        throw std::runtime_error(firstErr);
    return 0;
    ORC_CATCH(-1)
}

// DoReconstructionIter, SOR branch (SORInstead = true, recordInc = false),
// VariationalReconstruction_Reconstruction.hxx:534-589: cells in scan order, each
// updated in place: uRec_i *= (1 - relax), then per face in cell2face order
//   interior: uRec_i += relax (A^-1 B_f uRec_j + A^-1 b_f (u_j - u_i)), uRec_j as it stands
//   boundary: uRec_i += relax A^-1 BCC, GetBoundaryRHS evaluated on the uRec_i accumulated so
//             far (the reference passes the array being updated, :584)
// scanOrder NULL = natural order.
int orc_vr_sor(const orc_vr *h, int nVars, double *uRec, const double *u, const int64_t *scanOrder)
{
    ORC_TRY
    const VR &v = h->v;
    if (useSubs2nd(v)) // :503-512: the substitute reads only u
        return orc_vr_2nd(h, nVars, u, uRec);
    const int M = v.M;
    const size_t blkN = (size_t)M * nVars;
    const real relax = v.s.jacobiRelax;
    std::vector<real> BCC(blkN), nb(blkN);
    for (idx_t s = 0; s < v.nCell; s++)
    {
        const idx_t i = scanOrder ? scanOrder[s] : s;
        if (i < 0 || i >= v.nCell)
            throw std::runtime_error("scanOrder entry out of range");
        real *out = uRec + i * blkN;
        for (size_t q = 0; q < blkN; q++)
            out[q] *= 1 - relax;
        const real *blk = &v.AinvB[(size_t)i * (v.maxFC + 1) * M * M];
        for (size_t ic = 0; ic < v.c2f[i].size(); ic++)
        {
            const idx_t f = v.c2f[i][ic];
            const idx_t o = v.cellFaceOther(i, f);
            if (o < 0)
            {
                std::fill(BCC.begin(), BCC.end(), 0.0);
                boundaryBCC(v, i, f, nVars, uRec, u, nullptr, BCC);
                addAinvBCC(v, i, nVars, BCC, relax, out);
                continue;
            }
            const real *B = blk + (1 + ic) * (size_t)M * M;
            const real *bv = &v.Ainvb[((size_t)i * v.maxFC + ic) * M];
            std::fill(nb.begin(), nb.end(), 0.0);
            for (int var = 0; var < nVars; var++)
            {
                const real du = u[o * nVars + var] - u[i * nVars + var];
                for (int r = 0; r < M; r++)
                {
                    real acc = 0;
                    for (int k = 0; k < M; k++)
                        acc += B[(size_t)r * M + k] * uRec[o * blkN + (size_t)var * M + k];
                    acc += bv[r] * du;
                    nb[(size_t)var * M + r] = acc;
                }
            }
            for (size_t q = 0; q < blkN; q++)
                out[q] += relax * nb[q];
        }
    }
    return 0;
    ORC_CATCH(-1)
}

// One reconstruction sweep of the solver loops (frhs, EulerSolver.hxx:240-250, putIntoNew =
// false): in place in scan order when SORInstead, else Jacobi into tmp and copied back.
static void vrSweep(const orc_vr *h, double *uRec, const double *u, const int64_t *scanOrder, std::vector<real> &tmp)
{
    if (h->v.s.SORInstead)
    {
        if (orc_vr_sor(h, 5, uRec, u, scanOrder))
            throw std::runtime_error(g_err);
        return;
    }
    if (orc_vr_iter(h, 5, uRec, u, tmp.data()))
        throw std::runtime_error(g_err);
    std::copy(tmp.begin(), tmp.end(), uRec);
}

// DoReconstructionIterDiff, VariationalReconstruction_Reconstruction.hxx:617-670
int orc_vr_iter_diff(const orc_vr *h, int nVars, const double *uRec, const double *u, const double *uRecDiff,
                     double *uRecNew)
{
    ORC_TRY
    const VR &v = h->v;
    int M = v.M;
    size_t blkN = (size_t)M * nVars;
    if (useSubs2nd(v)) // :631-638: the substitute does not depend on uRec
    {
        std::fill(uRecNew, uRecNew + (size_t)v.nCell * blkN, 0.0);
        return 0;
    }
    for (idx_t i = 0; i < v.nCell; i++)
    {
        real *out = uRecNew + i * blkN;
        for (size_t q = 0; q < blkN; q++)
            out[q] = uRecDiff[i * blkN + q];
        const real *blk = &v.AinvB[(size_t)i * (v.maxFC + 1) * M * M];
        for (size_t ic = 0; ic < v.c2f[i].size(); ic++)
        {
            idx_t f = v.c2f[i][ic];
            idx_t o = v.cellFaceOther(i, f);
            if (o < 0)
            {
                if (!uRec || !u)
                    throw std::runtime_error("the boundary term of the VR Jacobian needs uRec and u");
                std::vector<real> BCC(blkN, 0.0);
                boundaryBCC(v, i, f, nVars, uRec, u, uRecDiff, BCC);
                addAinvBCC(v, i, nVars, BCC, -1.0, out);
                continue;
            }
            const real *B = blk + (1 + ic) * (size_t)M * M;
            for (int var = 0; var < nVars; var++)
                for (int r = 0; r < M; r++)
                {
                    real acc = 0;
                    for (int k = 0; k < M; k++)
                        acc += B[(size_t)r * M + k] * uRecDiff[o * blkN + (size_t)var * M + k];
                    out[(size_t)var * M + r] -= acc;
                }
        }
    }
    return 0;
    ORC_CATCH(-1)
}

int orc_set_bc(orc_vr *h, int nZone, const int32_t *type, const double *value, double gamma, double Rgas)
{
    ORC_TRY
    VR &v = h->v;
    if (nZone != v.nBCZone)
        throw std::runtime_error("zone count differs from the mesh's");
    v.bcType.assign(type, type + nZone);
    v.bcValue.assign(nZone, {0, 0, 0, 0, 0});
    for (int z = 0; z < nZone; z++)
        for (int k = 0; k < 5; k++)
            v.bcValue[z][k] = value ? value[z * 5 + k] : 0.0;
    v.bcGamma = gamma;
    v.bcRgas = Rgas;
    return 0;
    ORC_CATCH(-1)
}

int orc_set_vr_dirichlet(orc_vr *h, int nVars, const double *value)
{
    ORC_TRY
    VR &v = h->v;
    v.dirNVars = nVars;
    v.dirValue.assign(value, value + (size_t)v.nFace * v.faceQw.size() * nVars);
    return 0;
    ORC_CATCH(-1)
}

int orc_bc_ghost(int type, const double *UL, const double *n, const double *zoneValue, double gamma, double Rgas,
                 double *UR)
{
    ORC_TRY
    bcGhost(type, UL, n, zoneValue, gamma, Rgas, UR);
    return 0;
    ORC_CATCH(-1)
}

int orc_riemann(int rsType, int dim, const double *UL, const double *UR, const double *ULm,
                const double *URm, const double *n, double gamma, double dLambda, double fixScale,
                double incFScale, double *F, double *lam)
{
    (void)dim;
    ORC_TRY
    dispatchFlux(rsType, UL, UR, ULm, URm, n, gamma, F, dLambda, fixScale, incFScale, lam[0], lam[1], lam[2]);
    return 0;
    ORC_CATCH(-1)
}


// ---------------------------------------------------------------------------
// Laminar viscous terms
// ---------------------------------------------------------------------------
static constexpr real veryLargeReal = 3e200; // src/DNDS/Defines.hpp:195

static real cpOf(const orc_euler_settings *es) { return es->Rgas * es->gamma / (es->gamma - 1); }

// muEff, EulerEvaluator.hpp:751-776 (no turbulence)
static real muEffOf(const orc_euler_settings *es, const real *U, real T)
{
    switch (es->muModel)
    {
    case 0: return es->muGas;
    case 1:
    {
        real TRel = T / es->TRef;
        return es->muGas * TRel * std::sqrt(TRel) * (es->TRef + es->CSutherland) / (T + es->CSutherland);
    }
    case 2: return es->muGas * U[0];
    default: throw std::runtime_error("no such muModel");
    }
}

// GradientCons2Prim_IdealGas, Gas.hpp:1678-1698; G[d][var] = d U_var / d x_d
static void gradCons2Prim(const real *U, const real (*G)[5], real (*Gp)[5], real gamma)
{
    real velo[3] = {U[1] / U[0], U[2] / U[0], U[3] / U[0]};
    for (int d = 0; d < 3; d++)
        for (int k = 0; k < 5; k++)
            Gp[d][k] = G[d][k];
    for (int d = 0; d < 3; d++)
        for (int j = 0; j < 3; j++)
            Gp[d][1 + j] = (1.0 / sqr(U[0])) * (U[0] * G[d][1 + j] - G[d][0] * U[1 + j]);
    for (int d = 0; d < 3; d++)
    {
        real a = 0, b = 0;
        for (int j = 0; j < 3; j++)
            a += G[d][1 + j] * velo[j], b += Gp[d][1 + j] * U[1 + j];
        Gp[d][4] = (gamma - 1) * (G[d][4] - 0.5 * (a + b));
    }
}

// ViscousFlux_IdealGas, Gas.hpp:1595-1655 (laminar: no QCR); Gp[d][1 + j] = d u_j / d x_d
static void viscousFlux(const real *U, const real (*Gp)[5], const real *n, bool adiabatic, real gamma, real mu, real k,
                        real Cp, real *F)
{
    const real lambda = -2. / 3.;
    real velo[3] = {U[1] / U[0], U[2] / U[0], U[3] / U[0]};
    real vSqr = velo[0] * velo[0] + velo[1] * velo[1] + velo[2] * velo[2];
    real p = (gamma - 1) * (U[4] - U[0] * 0.5 * vSqr);
    real GradT[3];
    for (int d = 0; d < 3; d++)
        GradT[d] = (gamma / ((gamma - 1) * Cp * U[0] * U[0])) * (U[0] * Gp[d][4] - p * Gp[d][0]);
    if (adiabatic)
    {
        real tn = dot3(GradT, n);
        for (int d = 0; d < 3; d++)
            GradT[d] -= tn * n[d];
    }
    real tr = Gp[0][1] + Gp[1][2] + Gp[2][3];
    real S[3][3];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
            S[i][j] = (Gp[i][1 + j] + Gp[j][1 + i]) * mu + (i == j ? lambda * mu * tr : 0.0);
    F[0] = 0;
    for (int i = 0; i < 3; i++)
        F[1 + i] = S[i][0] * n[0] + S[i][1] * n[1] + S[i][2] * n[2];
    real e = 0;
    for (int i = 0; i < 3; i++)
        e += (S[i][0] * velo[0] + S[i][1] * velo[1] + S[i][2] * velo[2] + k * GradT[i]) * n[i];
    F[4] = e;
}

// lamVis of EvaluateDt, EulerEvaluator_EvaluateDt.hxx:964-979 (no turbulence)
static real lamVisOf(const orc_euler_settings *es, const real *uMean, real pMean)
{
    if (!es->viscous)
        return 0.0;
    real gamma = es->gamma, Cp = cpOf(es);
    real T = pMean / ((gamma - 1) / gamma * Cp * uMean[0]);
    real muf = muEffOf(es, uMean, T);
    return muf / uMean[0] * std::max(4. / 3., gamma / es->prGas);
}

// EvaluateDt, EulerEvaluator_EvaluateDt.hxx:840-1048 (inviscid: lamVis = 0; a boundary face
// takes UR = UL, :880-899)
int orc_eval_dt(const orc_vr *h, const orc_euler_settings *es, const double *u, double CFL,
                double maxDt, int useLocalDt, double *dt, double *lambdaFace, double *deltaLambdaCell,
                double *dtMin)
{
    ORC_TRY
    const VR &v = h->v;
    real gamma = es->gamma;
    std::vector<real> lamF(v.nFace), dLamF(v.nFace);
    for (idx_t f = 0; f < v.nFace; f++)
    {
        const real *n = v.faceMeanNorm[f].x;
        idx_t cL = v.f2c[f][0], cR = v.f2c[f][1];
        const real *UL = u + cL * 5;
        const real *UR = cR >= 0 ? u + cR * 5 : UL;
        real vL[3], vR[3], pL, aL2, HL, pR, aR2, HR;
        for (int d = 0; d < 3; d++)
            vL[d] = UL[1 + d] / UL[0], vR[d] = UR[1 + d] / UR[0];
        thermal(UL[4], UL[0], dot3(vL, vL), gamma, pL, aL2, HL);
        thermal(UR[4], UR[0], dot3(vR, vR), gamma, pR, aR2, HR);
        if (!(aL2 >= 0 && aR2 >= 0))
            throw std::runtime_error("mean value violates PP");
        real vgN = 0;
        real veloNL = dot3(vL, n), veloNR = dot3(vR, n);
        real veloNMean = 0.5 * (veloNL + veloNR);
        {
            real s[3] = {vL[0] + vR[0], vL[1] + vR[1], vL[2] + vR[2]};
            veloNMean = 0.5 * dot3(s, n);
        }
        real aL = std::sqrt(aL2), aR = std::sqrt(aR2);
        real lambdaConvection = std::max(aL + std::abs(veloNL - vgN), aR + std::abs(veloNR - vgN));
        real uMean[5];
        for (int k = 0; k < 5; k++)
            uMean[k] = cR >= 0 ? (UL[k] + UR[k]) * 0.5 : UL[k];
        real lamVis = lamVisOf(es, uMean, (pL + pR) * 0.5);
        real volR = cR >= 0 ? v.vol[cR] : v.vol[cL];
        real visc = lamVis * v.area[f] * (1. / v.vol[cL] + 1. / volR);
        lamF[f] = lambdaConvection + visc;
        lambdaFace[4 * f + 0] = lamF[f];                              // lambdaFace
        lambdaFace[4 * f + 1] = std::abs(veloNMean - vgN) + visc;     // lambdaFaceC
        lambdaFace[4 * f + 2] = visc;                                 // lambdaFaceVis
        lambdaFace[4 * f + 3] = lambdaConvection;                     // lambdaFace0/123/4
        real dvn = dot3(vR, n) - dot3(vL, n);
        {
            real dv[3] = {vR[0] - vL[0], vR[1] - vL[1], vR[2] - vL[2]};
            dvn = dot3(dv, n);
        }
        dLamF[f] = std::max({std::abs(dvn + aR - aL), std::abs(dvn - aR + aL), std::abs(aL - aR)});
    }
    real dmin = 3e200;
    for (idx_t i = 0; i < v.nCell; i++)
    {
        real lc = 0;
        for (auto f : v.c2f[i])
            lc += lamF[f] * v.area[f];
        dt[i] = std::min(CFL * v.vol[i] * v.smooth[i] / (lc + 1e-100), maxDt);
        dmin = std::min(dmin, dt[i]);
        real dl = 0;
        for (auto f : v.c2f[i])
            dl = std::max(dl, dLamF[f]);
        deltaLambdaCell[i] = dl;
    }
    *dtMin = dmin;
    if (!useLocalDt)
        for (idx_t i = 0; i < v.nCell; i++)
            dt[i] = dmin;
    return 0;
    ORC_CATCH(-1)
}

// EvaluateRHS, EulerEvaluator_EvaluateRHS.hxx:55-818 with fluxFace
// EulerEvaluator_EvaluateDt.hxx:1084-1511; inviscid; interior, periodic and boundary
// faces (ghost states from the boundary-condition table); serial direct scatter (:550-562).
int orc_eval_rhs(const orc_vr *h, const orc_euler_settings *es, const double *u, const double *uRec,
                 const double *deltaLambdaCell, double *rhs)
{
    return orc_eval_rhs_lam(h, es, u, uRec, deltaLambdaCell, rhs, nullptr);
}

static int evalRHSCore(const orc_vr *h, const orc_euler_settings *es, const double *u, const double *uRec,
                       const double *deltaLambdaCell, double *rhs, double *lamRoe, double *faceFlux);

int orc_eval_rhs_lam(const orc_vr *h, const orc_euler_settings *es, const double *u, const double *uRec,
                     const double *deltaLambdaCell, double *rhs, double *lamRoe)
{
    return evalRHSCore(h, es, u, uRec, deltaLambdaCell, rhs, lamRoe, nullptr);
}

// Force on the boundary faces of one zone: the momentum part of the integrated face fluxes
// (Riemann minus viscous flux, sum_g -finc JDet w) with the sign of the flux out of the domain.
// The wall force of a drag functional; the reference's solver integrates it in its output stage.
int orc_zone_force(const orc_vr *h, const orc_euler_settings *es, const double *u, const double *uRec,
                   const double *deltaLambdaCell, int zone, double *force)
{
    const VR &v = h->v;
    std::vector<real> rhs((size_t)v.nCell * 5), ff((size_t)v.nFace * 5);
    int rc = evalRHSCore(h, es, u, uRec, deltaLambdaCell, rhs.data(), nullptr, ff.data());
    if (rc)
        return rc;
    real a[3] = {0, 0, 0};
    for (idx_t f = 0; f < v.nFace; f++)
        if (v.f2c[f][1] < 0 && v.faceZone[f] == zone)
            for (int d = 0; d < 3; d++)
                a[d] -= ff[(size_t)f * 5 + 1 + d];
    for (int d = 0; d < 3; d++)
        force[d] = a[d];
    return 0;
}

static int evalRHSCore(const orc_vr *h, const orc_euler_settings *es, const double *u, const double *uRec,
                       const double *deltaLambdaCell, double *rhs, double *lamRoe, double *faceFlux)
{
    ORC_TRY
    const VR &v = h->v;
    int M = v.M;
    size_t G = v.faceQw.size();
    std::fill(rhs, rhs + v.nCell * 5, 0.0);
    for (idx_t f = 0; f < v.nFace; f++)
        if (v.f2c[f][1] < 0 && v.bcType.empty())
            throw std::runtime_error("boundary faces need the boundary-condition table (orc_set_bc)");
    std::vector<std::array<real, 5>> fbuf(v.nFace);
    std::string firstErr;
#pragma omp parallel num_threads(g_threads)
    {
    std::vector<real> D;
    std::vector<std::array<real, 5>> ULv(G), URv(G), fincC(G);
    std::vector<std::array<real, 3>> lamv(G); // per-point lam0 / lam123 / lam4 of the Riemann solver
    std::vector<std::array<real, 3>> nv(G);
    std::vector<std::array<std::array<real, 5>, 3>> GLv(G), GRv(G); // point gradients d U / d x_d
    const bool vis = es->viscous != 0;
    const int nDiff = vis ? v.dim + 1 : 1;
#pragma omp for schedule(static)
    for (idx_t f = 0; f < v.nFace; f++)
    {
      try
      {
        idx_t cL = v.f2c[f][0], cR = v.f2c[f][1];
        const bool bnd = cR < 0;
        const int zone = bnd ? v.faceZone[f] : -1;
        const real *ULm = u + cL * 5, *URm = bnd ? nullptr : u + cR * 5;
        real URmB[5];
        if (bnd) // URMeanXy = ghost of the mean at the face centre, EvaluateRHS.hxx:194-207
        {
            bcGhost(v.bcType[zone], ULm, v.faceMeanNorm[f].x, v.bcValue[zone].data(), es->gamma, v.bcRgas, URmB);
            URm = URmB;
        }
        for (size_t g = 0; g < G; g++)
        {
            // ULxy = u[L] + (D0Bj(L) * uRec[L])^T, EvaluateRHS.hxx:235-243, 287-297
            real UL[5], UR[5];
            v.diffBase(cL, v.facePointFromSide(f, 0, v.faceQP[f][g]), nDiff, 0, D);
            for (int var = 0; var < 5; var++)
            {
                real r = 0;
                if (uRec)
                    for (int k = 0; k < M; k++)
                        r += D[1 + k] * uRec[((size_t)cL * 5 + var) * M + k];
                UL[var] = ULm[var] + r;
            }
            // GradULxy = D1phi(L) uRec[L], EvaluateRHS.hxx:257-273 (zero rows beyond the mesh dimension)
            for (int d = 0; d < 3; d++)
                for (int var = 0; var < 5; var++)
                {
                    real r = 0;
                    if (vis && uRec && d < v.dim)
                        for (int k = 0; k < M; k++)
                            r += D[(size_t)(1 + d) * v.NDOF + 1 + k] * uRec[((size_t)cL * 5 + var) * M + k];
                    GLv[g][d][var] = r;
                }
            if (bnd) // URxy = generateBoundaryValue(ULxy, ...), GradURxy = GradULxy, EvaluateRHS.hxx:336-355
            {
                bcGhost(v.bcType[zone], UL, v.faceN[f][g].x, v.bcValue[zone].data(), es->gamma, v.bcRgas, UR);
                GRv[g] = GLv[g];
            }
            else
            {
                v.diffBase(cR, v.facePointFromSide(f, 1, v.faceQP[f][g]), nDiff, 0, D);
                for (int var = 0; var < 5; var++)
                {
                    real r = 0;
                    if (uRec)
                        for (int k = 0; k < M; k++)
                            r += D[1 + k] * uRec[((size_t)cR * 5 + var) * M + k];
                    UR[var] = URm[var] + r;
                }
                for (int d = 0; d < 3; d++)
                    for (int var = 0; var < 5; var++)
                    {
                        real r = 0;
                        if (vis && uRec && d < v.dim)
                            for (int k = 0; k < M; k++)
                                r += D[(size_t)(1 + d) * v.NDOF + 1 + k] * uRec[((size_t)cR * 5 + var) * M + k];
                        GRv[g][d][var] = r;
                    }
            }
            for (int var = 0; var < 5; var++)
                ULv[g][var] = UL[var], URv[g][var] = UR[var];
            for (int d = 0; d < 3; d++)
                nv[g][d] = v.faceN[f][g].x[d];
        }
        real dLambda = 0.0; // fluxFace, EvaluateDt.hxx:1158-1160
        if (deltaLambdaCell)
            dLambda = bnd ? deltaLambdaCell[cL] : std::max(deltaLambdaCell[cL], deltaLambdaCell[cR]);
        real lam0, lam123, lam4;
        int rs = es->rsType;
        if (es->rsMeanValueEig != 0 && rs >= 11 && rs <= 15)
        {
            std::vector<std::array<real, 5>> Fb(G);
            roeFluxBatch(rs - 10, (int)G, reinterpret_cast<const real(*)[5]>(ULv.data()),
                         reinterpret_cast<const real(*)[5]>(URv.data()), ULm, URm,
                         reinterpret_cast<const real(*)[3]>(nv.data()), v.faceMeanNorm[f].x, es->gamma,
                         reinterpret_cast<real(*)[5]>(Fb.data()), dLambda, es->rsFixScale, es->rsIncFScale,
                         lam0, lam123, lam4);
            for (size_t g = 0; g < G; g++)
                fincC[g] = Fb[g], lamv[g] = {lam0, lam123, lam4};
        }
        else
            for (size_t g = 0; g < G; g++)
            {
                dispatchFlux(rs, ULv[g].data(), URv[g].data(), ULm, URm, nv[g].data(), es->gamma, fincC[g].data(),
                             dLambda, es->rsFixScale, es->rsIncFScale, lam0, lam123, lam4);
                lamv[g] = {lam0, lam123, lam4};
            }
        if (lamRoe)
        {
            // useRoeJacobian: the face's eigenvalues for the block Jacobian, integrated over the
            // face and divided by its area (EvaluateRHS.hxx:527-544)
            real le[3] = {0, 0, 0};
            for (size_t g = 0; g < G; g++)
                for (int k = 0; k < 3; k++)
                    le[k] += (lamv[g][k] * v.faceJD[f][g]) * v.faceQw[g];
            for (int k = 0; k < 3; k++)
                lamRoe[(size_t)f * 3 + k] = le[k] / v.area[f];
        }
        if (vis)
        {
            // viscous flux, EvaluateRHS.hxx:346-424 and fluxFace EvaluateDt.hxx:1163-1215, finc -= visFlux (:1460)
            const int bt = bnd ? v.bcType[zone] : 0;
            const bool noGrad = bnd && (bt == ORC_BC_WALL_INVIS || bt == ORC_BC_OUT || bt == ORC_BC_FAR || bt == ORC_BC_SYM);
            real minVol = v.vol[cL];
            Pt dB;
            if (!bnd)
            {
                minVol = std::min(minVol, v.vol[cR]);
                Pt ob = v.otherBaryFromCell(cL, cR, f);
                for (int d = 0; d < 3; d++)
                    dB.x[d] = ob.x[d] - v.bary[cL].x[d];
            }
            else
                for (int d = 0; d < 3; d++)
                    dB.x[d] = v.faceCent[f].x[d] - v.bary[cL].x[d];
            const real gamma = es->gamma, Cp = cpOf(es);
            for (size_t g = 0; g < G; g++)
            {
                const real *n = nv[g].data();
                real distBaryPerp = std::abs(dot3(dB.x, n)) * (bnd ? 2.0 : 1.0);
                real distGRP = minVol / v.area[f] * 2;
                distGRP = std::max(std::min(distBaryPerp, distGRP * 2), distGRP * 0.25);
                if (bnd && (bt == ORC_BC_WALL_INVIS || bt == ORC_BC_SYM))
                    distGRP += veryLargeReal;
                real Um[5], Gm[3][5], Gp[3][5];
                for (int k = 0; k < 5; k++)
                    Um[k] = 0.5 * (ULv[g][k] + URv[g][k]);
                for (int d = 0; d < 3; d++)
                    for (int k = 0; k < 5; k++)
                        Gm[d][k] = noGrad ? 0.0
                                          : (GRv[g][d][k] + GLv[g][d][k]) * 0.5 +
                                                (1.0 / distGRP) * (n[d] * (URv[g][k] - ULv[g][k]));
                gradCons2Prim(Um, Gm, Gp, gamma);
                real pM, a2M, HM;
                thermal(Um[4], Um[0], sqr(Um[1] / Um[0]) + sqr(Um[2] / Um[0]) + sqr(Um[3] / Um[0]), gamma, pM, a2M, HM);
                real T = pM / ((gamma - 1) / gamma * Cp * Um[0]);
                real mu = muEffOf(es, Um, T);
                real kc = Cp * 0.0 / 0.9 + Cp * mu / es->prGas;
                real Fv[5];
                viscousFlux(Um, Gp, n, bnd && bt == ORC_BC_WALL, gamma, mu, kc, Cp, Fv);
                if (bnd && (bt == ORC_BC_WALL_INVIS || bt == ORC_BC_SYM))
                    for (int k = 0; k < 5; k++)
                        Fv[k] = 0.0;
                for (int k = 0; k < 5; k++)
                    fincC[g][k] -= Fv[k];
            }
        }
        real fluxEs[5] = {0, 0, 0, 0, 0};
        for (size_t g = 0; g < G; g++)
            for (int var = 0; var < 5; var++)
                fluxEs[var] += ((-fincC[g][var]) * v.faceJD[f][g]) * v.faceQw[g];
        for (int var = 0; var < 5; var++)
            fbuf[f][var] = fluxEs[var];
      }
      catch (const std::exception &e)
      {
#pragma omp critical
        firstErr = e.what();
      }
    }
    }
    if (!firstErr.empty())
        throw std::runtime_error(firstErr);
    if (faceFlux)
        for (idx_t f = 0; f < v.nFace; f++)
            for (int var = 0; var < 5; var++)
                faceFlux[(size_t)f * 5 + var] = fbuf[f][var];
    if (!g_rhsGather)
    {
        // serial build: direct scatter in face order, EvaluateRHS.hxx:550-562
        for (idx_t f = 0; f < v.nFace; f++)
        {
            idx_t cL = v.f2c[f][0], cR = v.f2c[f][1];
            for (int var = 0; var < 5; var++)
            {
                rhs[cL * 5 + var] += fbuf[f][var] / v.vol[cL];
                if (cR >= 0)
                    rhs[cR * 5 + var] += (-fbuf[f][var]) / v.vol[cR];
            }
        }
    }
    else
    {
        // OpenMP build: per-cell gather over cell2face, EvaluateRHS.hxx:640-662
#pragma omp parallel for schedule(static) num_threads(g_threads)
        for (idx_t i = 0; i < v.nCell; i++)
            for (idx_t f : v.c2f[i])
            {
                real sgn = v.cellIsBack(i, f) ? 1.0 : -1.0;
                for (int var = 0; var < 5; var++)
                    rhs[i * 5 + var] += (fbuf[f][var] * sgn) / v.vol[i];
            }
    }
    return 0;
    ORC_CATCH(-1)
}

// LUSGSMatrixInit, scalar mode, EulerEvaluator.hxx:37-159 (ignoreSourceTerm)
int orc_lusgs_init(const orc_vr *h, const double *dTau, double dt, double alphaDiag,
                   const double *lambdaFace, double *jdiag)
{
    const VR &v = h->v;
    for (idx_t i = 0; i < v.nCell; i++)
    {
        real fp = 1.0 / dTau[i] + 1.0 / dt;
        for (auto f : v.c2f[i])
            fp += (0.5 * alphaDiag) * v.area[f] * lambdaFace[4 * f] / v.vol[i];
        jdiag[i] = fp;
    }
    return 0;
}

// fluxJacobian0_Right_Times_du (scalar / non-Roe branch), EulerEvaluator.hpp:1100-1158,
// with IdealGasUIncrement Gas.hpp:462-471 and GasInviscidFluxFacialIncrement Gas.hpp:495-513
static void fluxJacobianTimesDu(const real *U, const real *n, const real *dU, real lambdaMain, real gamma, real *dF)
{
    real velo[3] = {U[1] / U[0], U[2] / U[0], U[3] / U[0]};
    real p, H, asqr;
    thermal(U[4], U[0], dot3(velo, velo), gamma, p, asqr, H);
    real dVelo[3];
    for (int d = 0; d < 3; d++)
        dVelo[d] = (dU[1 + d] - U[1 + d] / U[0] * dU[0]) / U[0];
    real dp = (gamma - 1) * (dU[4] - 0.5 * (dot3(dU + 1, velo) + dot3(U + 1, dVelo)));
    real vn = dot3(velo, n), dvn = dot3(dVelo, n);
    dF[0] = dot3(dU + 1, n);
    for (int d = 0; d < 3; d++)
        dF[1 + d] = dU[1 + d] * vn + U[1 + d] * dvn + n[d] * dp;
    dF[4] = (dU[4] + dp) * vn + (U[4] + p) * dvn;
    // vg = 0: F -= dU * vg.n is zero
    for (int k = 0; k < 5; k++)
        dF[k] += -1 * lambdaMain * dU[k];
}

// UpdateSGS, EulerEvaluator.hxx:550-693 (scalar JDiag, non-Roe Jacobian).  scanOrder
// gives the scan position -> cell map (NULL = natural order); "already visited" is
// decided by scan position, as the reference does with renumbered cell indices.
int orc_update_sgs(const orc_vr *h, const orc_euler_settings *es, double alphaDiag, const double *rhs,
                   const double *u, const double *uInc, double *uIncNew, const double *jdiag,
                   const double *lambdaFace, int forward, int gsUpdate, int uIncIsZero,
                   const int64_t *scanOrder, double *sumInc)
{
    const VR &v = h->v;
    idx_t N = v.nCell;
    std::vector<idx_t> pos(N), ord(N);
    for (idx_t s = 0; s < N; s++)
    {
        ord[s] = scanOrder ? scanOrder[s] : s;
        pos[ord[s]] = s;
    }
    for (int k = 0; k < 5; k++)
        sumInc[k] = 0;
    for (idx_t sIt = 0; sIt < N; sIt++)
    {
        idx_t iScan = forward ? sIt : N - 1 - sIt;
        idx_t i = ord[iScan];
        real buf[5];
        for (int k = 0; k < 5; k++)
            buf[k] = rhs[i * 5 + k];
        for (auto f : v.c2f[i])
        {
            idx_t o = v.cellFaceOther(i, f);
            int atFace = v.cellIsBack(i, f) ? 0 : 1;
            if (o < 0 || o == i)
                continue;
            real n[3];
            for (int d = 0; d < 3; d++)
                n[d] = v.faceMeanNorm[f].x[d] * (atFace ? -1 : 1);
            bool gsUseNew = gsUpdate && (forward ? (pos[o] < iScan) : (pos[o] > iScan));
            if (uIncIsZero && !gsUseNew)
                continue;
            const real *duj = gsUseNew ? uIncNew + o * 5 : uInc + o * 5;
            real dF[5];
            fluxJacobianTimesDu(u + o * 5, n, duj, lambdaFace[4 * f], es->gamma, dF);
            for (int k = 0; k < 5; k++)
                buf[k] -= (0.5 * alphaDiag) * v.area[f] / v.vol[i] * dF[k];
        }
        for (int k = 0; k < 5; k++)
        {
            real old = uIncNew[i * 5 + k];
            real nw = buf[k] / jdiag[i];
            uIncNew[i * 5 + k] = nw;
            sumInc[k] += std::abs(nw - old);
        }
    }
    return 0;
}

// fluxJacobian0_Right_Times_du with the Roe term (EulerEvaluator.hpp:1100-1158, useRoeTerm = 1):
// dF = [A(U) dU unless omitF] + s RoeFluxIncFDiff(dU; Roe average of U and UOther, lam0, lam123,
// lam4) + s lambdaVis dU, with s = incFsign (RoeFluxIncFDiff Gas.hpp:1209-1243, GetRoeAverage
// Gas.hpp:1149-1180; vg = 0).
static void fluxJacobianRoeTimesDu(const real *U, const real *UOther, const real *n, const real *dU, real lambdaVis,
                                   real lam0, real lam123, real lam4, real gamma, int incFsign, bool omitF, real *dF)
{
    if (!omitF)
        fluxJacobianTimesDu(U, n, dU, 0.0, gamma, dF);
    else
        for (int k = 0; k < 5; k++)
            dF[k] = 0;
    RoePre r = roePreamble(U, UOther, gamma);
    real inc[5];
    roeIncF(dU, n, r.veloRoe, r.vsqrRoe, r.aRoe, r.asqrRoe, r.HRoe, incFsign * lam0, incFsign * lam123,
            incFsign * lam4, gamma, inc);
    for (int k = 0; k < 5; k++)
        dF[k] += inc[k];
    for (int k = 0; k < 5; k++)
        dF[k] += incFsign * lambdaVis * dU[k];
}

// JacobianDiagBlock::GetInvert in block mode (EulerJacobian.hpp:153-190): the block scaled by
// the inverse of its diagonal, inverted by full-pivot LU (Gauss-Jordan with complete pivoting,
// the textbook form of Eigen's FullPivLU::inverse), then the diagonal scaling undone:
// J^-1 = (D^-1 J)^-1 D^-1.  J and Jinv are 5x5 row-major.
static void invertBlock5(const real *J, real *Jinv)
{
    real P[5][5], X[5][5], dinv[5];
    for (int i = 0; i < 5; i++)
    {
        if (J[i * 5 + i] == 0)
            throw std::runtime_error("block JDiag has a zero diagonal entry");
        dinv[i] = 1.0 / J[i * 5 + i];
    }
    for (int i = 0; i < 5; i++)
        for (int j = 0; j < 5; j++)
            P[i][j] = dinv[i] * J[i * 5 + j], X[i][j] = i == j ? 1.0 : 0.0;
    int colPerm[5] = {0, 1, 2, 3, 4};
    for (int k = 0; k < 5; k++)
    {
        int pr = k, pc = k;
        real best = -1;
        for (int i = k; i < 5; i++)
            for (int j = k; j < 5; j++)
                if (std::abs(P[i][j]) > best)
                    best = std::abs(P[i][j]), pr = i, pc = j;
        if (!(best > 0))
            throw std::runtime_error("block JDiag is singular");
        for (int j = 0; j < 5; j++)
            std::swap(P[k][j], P[pr][j]), std::swap(X[k][j], X[pr][j]);
        for (int i = 0; i < 5; i++)
            std::swap(P[i][k], P[i][pc]);
        std::swap(colPerm[k], colPerm[pc]);
        real piv = 1.0 / P[k][k];
        for (int j = 0; j < 5; j++)
            P[k][j] *= piv, X[k][j] *= piv;
        for (int i = 0; i < 5; i++)
        {
            if (i == k)
                continue;
            real fct = P[i][k];
            if (fct == 0)
                continue;
            for (int j = 0; j < 5; j++)
                P[i][j] -= fct * P[k][j], X[i][j] -= fct * X[k][j];
        }
    }
    // P Q = I after the eliminations with column permutation Q: P^-1 = Q X
    real Pinv[5][5];
    for (int k = 0; k < 5; k++)
        for (int j = 0; j < 5; j++)
            Pinv[colPerm[k]][j] = X[k][j];
    for (int i = 0; i < 5; i++)
        for (int j = 0; j < 5; j++)
            Jinv[i * 5 + j] = Pinv[i][j] * dinv[j];
}

// the face's Roe eigenvalues seen from cell i: lambdaFace0 and lambdaFace4 swap when i is the
// front cell (EulerEvaluator.hxx:96, 617)
static void roeLambdas(const real *lambdaFace, const real *lamRoe, idx_t f, int atFace, real &l0, real &l123,
                       real &l4)
{
    if (lamRoe)
        l0 = lamRoe[(size_t)f * 3], l123 = lamRoe[(size_t)f * 3 + 1], l4 = lamRoe[(size_t)f * 3 + 2];
    else
        l0 = l123 = l4 = lambdaFace[(size_t)f * 4 + 3];
    if (atFace)
        std::swap(l0, l4);
}

// LUSGSMatrixInit with useRoeJacobian (EulerEvaluator.hxx:37-159, ignoreSourceTerm): per face
// 0.5 alpha A_f / V_i (|A|_Roe(u_i, u_j) + lambdaVis I), for a boundary face also
// 0.5 alpha A_f / V_i (A(u_g) - |A|_Roe(u_g, u_i) - lambdaVis I) dU_g/dU_i with the ghost u_g at the
// face centre and its linearisation (generateBoundaryValue linMode 1, EvaluateDt.hxx:1762-1835),
// then 1/dTau + 1/dt on the diagonal.
int orc_lusgs_init_block(const orc_vr *h, const orc_euler_settings *es, const double *dTau, double dt,
                         double alphaDiag, const double *u, const double *lambdaFace, const double *lamRoe,
                         double *jdiag, double *jinv)
{
    ORC_TRY
    const VR &v = h->v;
    const real gamma = es->gamma;
    for (idx_t i = 0; i < v.nCell; i++)
    {
        real J[25] = {0};
        const real *ui = u + (size_t)i * 5;
        for (auto f : v.c2f[i])
        {
            idx_t o = v.cellFaceOther(i, f);
            int atFace = v.cellIsBack(i, f) ? 0 : 1;
            real n[3];
            for (int d = 0; d < 3; d++)
                n[d] = v.faceMeanNorm[f].x[d] * (atFace ? -1 : 1);
            real l0, l123, l4;
            roeLambdas(lambdaFace, lamRoe, f, atFace, l0, l123, l4);
            const real lvis = lambdaFace[(size_t)f * 4 + 2];
            const real coef = (0.5 * alphaDiag) * v.area[f] / v.vol[i];
            real uj[5];
            int bt = -1;
            if (o < 0)
            {
                int zone = v.faceZone[f];
                bt = v.bcType.empty() ? -1 : v.bcType[zone];
                if (bt < 0)
                    throw std::runtime_error("boundary faces need the boundary-condition table (orc_set_bc)");
                bcGhost(bt, ui, n, v.bcValue[zone].data(), gamma, v.bcRgas, uj);
            }
            else
                for (int k = 0; k < 5; k++)
                    uj[k] = u[(size_t)o * 5 + k];
            real jacII[25];
            for (int c = 0; c < 5; c++)
            {
                real e[5] = {0, 0, 0, 0, 0}, col[5];
                e[c] = 1;
                fluxJacobianRoeTimesDu(ui, uj, n, e, lvis, l0, l123, l4, gamma, +1, true, col);
                for (int r = 0; r < 5; r++)
                    jacII[r * 5 + c] = col[r];
            }
            if (o < 0)
            {
                // dU_g / dU_i column by column; zero for far field, in- and outflow (linMode 1)
                real JBC[25] = {0};
                // the isothermal wall linearises as the adiabatic one: linMode 1 returns before the
                // density reset (generateBV_ViscousWall, EvaluateDt.hxx:2265-2266)
                const int btLin = bt == ORC_BC_WALL_ISO ? ORC_BC_WALL : bt;
                if (btLin == ORC_BC_WALL_INVIS || btLin == ORC_BC_SYM || btLin == ORC_BC_WALL)
                    for (int c = 0; c < 5; c++)
                    {
                        real e[5] = {0, 0, 0, 0, 0}, g[5];
                        e[c] = 1;
                        bcGhost(btLin, e, n, v.bcValue[v.faceZone[f]].data(), gamma, v.bcRgas, g);
                        for (int r = 0; r < 5; r++)
                            JBC[r * 5 + c] = g[r];
                    }
                real jacIJ[25];
                for (int c = 0; c < 5; c++)
                {
                    real e[5] = {0, 0, 0, 0, 0}, col[5];
                    e[c] = 1;
                    fluxJacobianRoeTimesDu(uj, ui, n, e, lvis, l0, l123, l4, gamma, -1, false, col);
                    for (int r = 0; r < 5; r++)
                        jacIJ[r * 5 + c] = col[r];
                }
                for (int r = 0; r < 5; r++)
                    for (int c = 0; c < 5; c++)
                    {
                        real acc = 0;
                        for (int k = 0; k < 5; k++)
                            acc += jacIJ[r * 5 + k] * JBC[k * 5 + c];
                        J[r * 5 + c] += coef * acc;
                    }
            }
            for (int q = 0; q < 25; q++)
                J[q] += coef * jacII[q];
        }
        for (int k = 0; k < 5; k++)
            J[k * 5 + k] += 1.0 / dTau[i] + 1.0 / dt;
        for (int q = 0; q < 25; q++)
            jdiag[(size_t)i * 25 + q] = J[q];
        invertBlock5(J, jinv + (size_t)i * 25);
    }
    return 0;
    ORC_CATCH(-1)
}

// UpdateSGS with useRoeJacobian (EulerEvaluator.hxx:550-693): the off-diagonal term of a
// neighbour j is fluxJacobian0(u_j, u_i, n, dU_j) with the Roe term and incFsign -1, the
// update x_i = JDiag_i^-1 buf (MatVecLeftInvert, EulerJacobian.hpp:215).
int orc_update_sgs_block(const orc_vr *h, const orc_euler_settings *es, double alphaDiag, const double *rhs,
                         const double *u, const double *uInc, double *uIncNew, const double *jinv,
                         const double *lambdaFace, const double *lamRoe, int forward, int gsUpdate,
                         int uIncIsZero, const int64_t *scanOrder, double *sumInc)
{
    ORC_TRY
    const VR &v = h->v;
    idx_t N = v.nCell;
    std::vector<idx_t> pos(N), ord(N);
    for (idx_t s = 0; s < N; s++)
    {
        ord[s] = scanOrder ? scanOrder[s] : s;
        pos[ord[s]] = s;
    }
    for (int k = 0; k < 5; k++)
        sumInc[k] = 0;
    for (idx_t sIt = 0; sIt < N; sIt++)
    {
        idx_t iScan = forward ? sIt : N - 1 - sIt;
        idx_t i = ord[iScan];
        real buf[5];
        for (int k = 0; k < 5; k++)
            buf[k] = rhs[i * 5 + k];
        for (auto f : v.c2f[i])
        {
            idx_t o = v.cellFaceOther(i, f);
            int atFace = v.cellIsBack(i, f) ? 0 : 1;
            if (o < 0 || o == i)
                continue;
            real n[3];
            for (int d = 0; d < 3; d++)
                n[d] = v.faceMeanNorm[f].x[d] * (atFace ? -1 : 1);
            bool gsUseNew = gsUpdate && (forward ? (pos[o] < iScan) : (pos[o] > iScan));
            if (uIncIsZero && !gsUseNew)
                continue;
            const real *duj = gsUseNew ? uIncNew + o * 5 : uInc + o * 5;
            real l0, l123, l4;
            roeLambdas(lambdaFace, lamRoe, f, atFace, l0, l123, l4);
            real dF[5];
            fluxJacobianRoeTimesDu(u + o * 5, u + i * 5, n, duj, lambdaFace[(size_t)f * 4 + 2], l0, l123, l4,
                                   es->gamma, -1, false, dF);
            for (int k = 0; k < 5; k++)
                buf[k] -= (0.5 * alphaDiag) * v.area[f] / v.vol[i] * dF[k];
        }
        const real *Ji = jinv + (size_t)i * 25;
        for (int k = 0; k < 5; k++)
        {
            real nw = 0;
            for (int c = 0; c < 5; c++)
                nw += Ji[k * 5 + c] * buf[c];
            real old = uIncNew[i * 5 + k];
            uIncNew[i * 5 + k] = nw;
            sumInc[k] += std::abs(nw - old);
        }
    }
    return 0;
    ORC_CATCH(-1)
}

// The transpose of the block LU-SGS matrix of orc_update_sgs_block (useRoeJacobian), swept in the
// same order.  Row i of that matrix holds JDiag_i and, per neighbour j, the block
// O_ij = (alpha/2) A_f / V_i (A(u_j) n_ij - |A|_Roe(u_j, u_i; n_ij) - lambdaVis I) with the face
// eigenvalues seen from i; the transposed sweep solves
// y_i = JDiag_i^-T (b_i - sum_j O_ji^T y_j), O_ji assembled column by column from
// fluxJacobianRoeTimesDu with the roles of i and j exchanged.  No counterpart in the reference:
// the block preconditioner of the discrete adjoint.
int orc_update_sgs_block_transpose(const orc_vr *h, const orc_euler_settings *es, double alphaDiag, const double *rhs,
                                   const double *u, const double *uInc, double *uIncNew, const double *jinv,
                                   const double *lambdaFace, const double *lamRoe, int forward, int gsUpdate,
                                   int uIncIsZero, const int64_t *scanOrder, double *sumInc)
{
    ORC_TRY
    const VR &v = h->v;
    idx_t N = v.nCell;
    std::vector<idx_t> pos(N), ord(N);
    for (idx_t s = 0; s < N; s++)
    {
        ord[s] = scanOrder ? scanOrder[s] : s;
        pos[ord[s]] = s;
    }
    for (int k = 0; k < 5; k++)
        sumInc[k] = 0;
    for (idx_t sIt = 0; sIt < N; sIt++)
    {
        idx_t iScan = forward ? sIt : N - 1 - sIt;
        idx_t i = ord[iScan];
        real buf[5];
        for (int k = 0; k < 5; k++)
            buf[k] = rhs[i * 5 + k];
        for (auto f : v.c2f[i])
        {
            idx_t o = v.cellFaceOther(i, f);
            if (o < 0 || o == i)
                continue;
            bool gsUseNew = gsUpdate && (forward ? (pos[o] < iScan) : (pos[o] > iScan));
            if (uIncIsZero && !gsUseNew)
                continue;
            const real *yj = gsUseNew ? uIncNew + o * 5 : uInc + o * 5;
            int atFaceJ = v.cellIsBack(i, f) ? 1 : 0; // the neighbour is the front cell when i is the back
            real n[3];                                  // out of the neighbour
            for (int d = 0; d < 3; d++)
                n[d] = v.faceMeanNorm[f].x[d] * (atFaceJ ? -1 : 1);
            real l0, l123, l4;
            roeLambdas(lambdaFace, lamRoe, f, atFaceJ, l0, l123, l4);
            real B[5][5]; // O_ji / coef, B[r][c]
            for (int c = 0; c < 5; c++)
            {
                real e[5] = {0, 0, 0, 0, 0}, col[5];
                e[c] = 1;
                fluxJacobianRoeTimesDu(u + i * 5, u + o * 5, n, e, lambdaFace[(size_t)f * 4 + 2], l0, l123, l4,
                                       es->gamma, -1, false, col);
                for (int r = 0; r < 5; r++)
                    B[r][c] = col[r];
            }
            const real cji = (0.5 * alphaDiag) * v.area[f] / v.vol[o];
            for (int c = 0; c < 5; c++)
            {
                real acc = 0;
                for (int r = 0; r < 5; r++)
                    acc += B[r][c] * yj[r];
                buf[c] -= cji * acc;
            }
        }
        const real *Ji = jinv + (size_t)i * 25;
        for (int k = 0; k < 5; k++)
        {
            real nw = 0;
            for (int c = 0; c < 5; c++)
                nw += Ji[c * 5 + k] * buf[c];
            real old = uIncNew[i * 5 + k];
            uIncNew[i * 5 + k] = nw;
            sumInc[k] += std::abs(nw - old);
        }
    }
    return 0;
    ORC_CATCH(-1)
}

// The transpose of the scalar LU-SGS matrix P of UpdateSGS (the matrix LUSGSMatrixVec applies),
// swept in the same order: (P^T y)_i = D_i y_i + sum_j c_ji (A(u_i) n_ji - lambda_f I)^T y_j with
// c_ji = (alpha/2) A_f / V_j and n_ji the normal out of j, so a sweep solves
// y_i = (b_i - sum_j c_ji (A(u_i) n_ji - lambda_f)^T y_j) / D_i.  No counterpart in the reference:
// the preconditioner of the discrete-adjoint fixed point, whose iteration it makes converge at the
// rate of the primal LU-SGS iteration.  A(u_i) n_ji is assembled column by column from
// fluxJacobianTimesDu.
int orc_update_sgs_transpose(const orc_vr *h, const orc_euler_settings *es, double alphaDiag, const double *rhs,
                             const double *u, const double *uInc, double *uIncNew, const double *jdiag,
                             const double *lambdaFace, int forward, int gsUpdate, int uIncIsZero,
                             const int64_t *scanOrder, double *sumInc)
{
    ORC_TRY
    const VR &v = h->v;
    idx_t N = v.nCell;
    std::vector<idx_t> pos(N), ord(N);
    for (idx_t s = 0; s < N; s++)
    {
        ord[s] = scanOrder ? scanOrder[s] : s;
        pos[ord[s]] = s;
    }
    for (int k = 0; k < 5; k++)
        sumInc[k] = 0;
    for (idx_t sIt = 0; sIt < N; sIt++)
    {
        idx_t iScan = forward ? sIt : N - 1 - sIt;
        idx_t i = ord[iScan];
        real buf[5];
        for (int k = 0; k < 5; k++)
            buf[k] = rhs[i * 5 + k];
        for (auto f : v.c2f[i])
        {
            idx_t o = v.cellFaceOther(i, f);
            int atFace = v.cellIsBack(i, f) ? 0 : 1;
            if (o < 0 || o == i)
                continue;
            bool gsUseNew = gsUpdate && (forward ? (pos[o] < iScan) : (pos[o] > iScan));
            if (uIncIsZero && !gsUseNew)
                continue;
            const real *yj = gsUseNew ? uIncNew + o * 5 : uInc + o * 5;
            real n[3]; // out of the neighbour j
            for (int d = 0; d < 3; d++)
                n[d] = v.faceMeanNorm[f].x[d] * (atFace ? 1 : -1);
            real B[5][5]; // B[r][c] = (A(u_i) n_ji - lambda I)[r][c]
            for (int c = 0; c < 5; c++)
            {
                real e[5] = {0, 0, 0, 0, 0}, col[5];
                e[c] = 1;
                fluxJacobianTimesDu(u + i * 5, n, e, lambdaFace[4 * f], es->gamma, col);
                for (int r = 0; r < 5; r++)
                    B[r][c] = col[r];
            }
            const real cji = (0.5 * alphaDiag) * v.area[f] / v.vol[o];
            for (int c = 0; c < 5; c++)
            {
                real acc = 0;
                for (int r = 0; r < 5; r++)
                    acc += B[r][c] * yj[r];
                buf[c] -= cji * acc;
            }
        }
        for (int k = 0; k < 5; k++)
        {
            real old = uIncNew[i * 5 + k];
            real nw = buf[k] / jdiag[i];
            uIncNew[i * 5 + k] = nw;
            sumInc[k] += std::abs(nw - old);
        }
    }
    return 0;
    ORC_CATCH(-1)
}

// LUSGSMatrixVec (EulerEvaluator.hxx:175-275), scalar Jacobian: the matrix whose
// Gauss-Seidel sweeps UpdateSGS performs, Ax_i = D_i x_i - buf_i with
// buf_i = -sum_f (alpha/2) A_f / V_i * fluxJacobian0(u_j) x_j (lambda term included).
int orc_lusgs_matvec(const orc_vr *h, const orc_euler_settings *es, double alphaDiag, const double *u,
                     const double *x, const double *jdiag, const double *lambdaFace, double *Ax)
{
    const VR &v = h->v;
    for (idx_t i = 0; i < v.nCell; i++)
    {
        real buf[5] = {0, 0, 0, 0, 0};
        for (auto f : v.c2f[i])
        {
            idx_t o = v.cellFaceOther(i, f);
            int atFace = v.cellIsBack(i, f) ? 0 : 1;
            if (o < 0)
                continue;
            real n[3];
            for (int d = 0; d < 3; d++)
                n[d] = v.faceMeanNorm[f].x[d] * (atFace ? -1 : 1);
            real dF[5];
            fluxJacobianTimesDu(u + o * 5, n, x + o * 5, lambdaFace[4 * f], es->gamma, dF);
            for (int k = 0; k < 5; k++)
                buf[k] -= (0.5 * alphaDiag) * v.area[f] / v.vol[i] * dF[k];
        }
        for (int k = 0; k < 5; k++)
            Ax[i * 5 + k] = jdiag[i] * x[i * 5 + k] - buf[k];
    }
    return 0;
}

// ModelEvaluator::EvaluateRHS (src/CFV/ModelEvaluator.cpp:6-115, default options; model
// src/CFV/ModelEvaluator.hpp:18-123): scalar advection-diffusion of nVars independent variables
// in 2-D, rhs_i = (1/V_i) sum_f +- sum_g w JDet [ -a.n (U_L + U_R)/2 + |a.n| (U_R - U_L)/2
// + sigma (GradU_mean . n) ], with GradU_mean = (GradU_L + GradU_R)/2 + n (U_R - U_L) / distGRP,
// distGRP = 2 min(V_L, V_R) / A_f.  Boundary faces take U_R = 0 (generateBoundaryValue,
// ModelEvaluator.hpp:96-112) and GradU_R = GradU_L.  u: [N][nVars], uRec: [N][nVars][M]
// (may be NULL: first order), rhs: [N][nVars].
int orc_model_rhs(const orc_vr *h, int nVars, double ax, double ay, double sigma, const double *u, const double *uRec,
                  double *rhs)
{
    ORC_TRY
    const VR &v = h->v;
    if (v.dim != 2)
        throw std::runtime_error("ModelEvaluator is two-dimensional");
    const int M = v.M;
    const size_t G = v.faceQw.size();
    std::fill(rhs, rhs + (size_t)v.nCell * nVars, 0.0);
    std::vector<real> D, UL(nVars), UR(nVars), fluxEs(nVars);
    std::vector<std::array<real, 2>> GL(nVars), GR(nVars);
    for (idx_t f = 0; f < v.nFace; f++)
    {
        const idx_t cL = v.f2c[f][0], cR = v.f2c[f][1];
        const bool bnd = cR < 0;
        std::fill(fluxEs.begin(), fluxEs.end(), 0.0);
        for (size_t g = 0; g < G; g++)
        {
            const real *n = v.faceN[f][g].x;
            auto side = [&](idx_t c, int s, std::vector<real> &U, std::vector<std::array<real, 2>> &Gr) {
                v.diffBase(c, v.facePointFromSide(f, s, v.faceQP[f][g]), 3, 0, D);
                for (int var = 0; var < nVars; var++)
                {
                    real r = 0, g0 = 0, g1 = 0;
                    if (uRec)
                        for (int k = 0; k < M; k++)
                        {
                            const real x = uRec[((size_t)c * nVars + var) * M + k];
                            r += D[1 + k] * x;
                            g0 += D[(size_t)v.NDOF + 1 + k] * x;
                            g1 += D[(size_t)2 * v.NDOF + 1 + k] * x;
                        }
                    U[var] = u[(size_t)c * nVars + var] + r;
                    Gr[var] = {g0, g1};
                }
            };
            side(cL, 0, UL, GL);
            real minVol = v.vol[cL];
            if (!bnd)
            {
                side(cR, 1, UR, GR);
                minVol = std::min(minVol, v.vol[cR]);
            }
            else
                for (int var = 0; var < nVars; var++)
                    UR[var] = 0.0, GR[var] = GL[var];
            const real distGRP = minVol / v.area[f] * 2;
            const real an = n[0] * ax + n[1] * ay;
            for (int var = 0; var < nVars; var++)
            {
                real gm0 = (GR[var][0] + GL[var][0]) * 0.5 + (1.0 / distGRP) * (n[0] * (UR[var] - UL[var]));
                real gm1 = (GR[var][1] + GL[var][1]) * 0.5 + (1.0 / distGRP) * (n[1] * (UR[var] - UL[var]));
                real finc = -(an * UR[var] + an * UL[var]) * 0.5 + 0.5 * std::abs(an) * (UR[var] - UL[var]);
                finc += (gm0 * n[0] + gm1 * n[1]) * sigma;
                finc *= v.faceJD[f][g];
                fluxEs[var] += finc * v.faceQw[g];
            }
        }
        for (int var = 0; var < nVars; var++)
        {
            rhs[(size_t)cL * nVars + var] += fluxEs[var] / v.vol[cL];
            if (!bnd)
                rhs[(size_t)cR * nVars + var] += -fluxEs[var] / v.vol[cR];
        }
    }
    return 0;
    ORC_CATCH(-1)
}

// Explicit SSP-RK3 (src/Solver/ODE.hpp:1477-1520) around frhs = nRec x Jacobi VR + EvaluateRHS,
// fdt = EvaluateDt (global dt from CFL unless dtFixed > 0); fincrement = plain add (the
// positivity compression of EulerEvaluator.hpp:1795 is the identity for these smooth states).
int orc_ssprk3(const orc_vr *h, const orc_euler_settings *es, double *u, double *uRec, int nStep,
               double CFL, int nRec, double dtFixed)
{
    ORC_TRY
    const VR &v = h->v;
    idx_t N = v.nCell;
    size_t nR = (size_t)N * 5 * v.M;
    std::vector<real> dtv(N), lam((size_t)v.nFace * 4), dlc(N), rhs(N * 5), xLast(N * 5), tmp(nR);
    auto frhs = [&](const real *x) {
        for (int r = 0; r < nRec; r++)
            vrSweep(h, uRec, x, nullptr, tmp);
        if (orc_eval_rhs(h, es, x, uRec, dlc.data(), rhs.data()))
            throw std::runtime_error(g_err);
    };
    for (int step = 0; step < nStep; step++)
    {
        real dtMin;
        if (orc_eval_dt(h, es, u, CFL, 1e100, 0, dtv.data(), lam.data(), dlc.data(), &dtMin))
            throw std::runtime_error(g_err);
        real dt = dtFixed > 0 ? dtFixed : dtMin;
        std::copy(u, u + N * 5, xLast.begin());
        frhs(u);
        for (idx_t q = 0; q < N * 5; q++)
            u[q] += rhs[q] * dt;
        frhs(u);
        for (idx_t q = 0; q < N * 5; q++)
            u[q] = u[q] * 0.25 + xLast[q] * 0.75 + 0.25 * (rhs[q] * dt);
        frhs(u);
        for (idx_t q = 0; q < N * 5; q++)
            u[q] = u[q] * (2. / 3.) + xLast[q] * (1. / 3.) + (2. / 3.) * (rhs[q] * dt);
    }
    return 0;
    ORC_CATCH(-1)
}

// One implicit-Euler dual-time step (ImplicitEulerDualTimeStep::Step, src/Solver/ODE.hpp:80-98)
// with the EulerSolver callbacks (src/Euler/EulerSolver.hxx): fdtau :674-694 (local dTau,
// alphaDiag = 1, no dTau smoothing), frhs :645-672 / frhsOuter :147-235 (Jacobi VR sweeps,
// recLinearScheme 0), fsolve :740-822 (LUSGSMatrixInit, zero increment, solveLinear
// :1407-1500 -> doPrecondition :1605-1641: forward then backward UpdateSGS, repeated
// 1 + sgsIter times), fincrement :696-738 (increment alpha = 1 when no positivity limit
// is needed, then AddFixedIncrement), fstop -> functor_fstop (EulerSolver_Init.hxx:490-530:
// L1 norm, running-maximum base, density component).
// BDF coefficients of ImplicitBDFDualTimeStep (src/Solver/ODE.hpp:590-595): row k - 1 holds
// (alpha, c_last, c_prev1, ...) of the k-step formula
static const real kBDF[4][5] = {{1. / 1., 1. / 1., 0, 0, 0},
                                {2. / 3., 4. / 3., -1. / 3., 0, 0},
                                {6. / 11., 18. / 11., -9. / 11., 2. / 11., 0},
                                {12. / 25., 48. / 25., -36. / 25., 16. / 25., -3. / 25.}};

static int implicitStep(const orc_vr *h, const orc_euler_settings *es, const orc_implicit_settings *is, int kCur,
                        const double *xPrevs, double *u, double *uRec, const int64_t *scanOrder, int *nIter,
                        double *res, double *resRel, double *history, double *ppStats);

int orc_implicit_euler_step(const orc_vr *h, const orc_euler_settings *es, const orc_implicit_settings *is,
                            double *u, double *uRec, const int64_t *scanOrder, int *nIter, double *res,
                            double *resRel, double *history, double *ppStats)
{
    return implicitStep(h, es, is, 0, nullptr, u, uRec, scanOrder, nIter, res, resRel, history, ppStats);
}

// ImplicitBDFDualTimeStep::Step (src/Solver/ODE.hpp:427-477): the k-step BDF with nPrev earlier
// states (kCurrent = nPrev + 1), xPrevs [nPrev][N][5] newest first
int orc_implicit_bdf_step(const orc_vr *h, const orc_euler_settings *es, const orc_implicit_settings *is, int nPrev,
                          const double *xPrevs, double *u, double *uRec, const int64_t *scanOrder, int *nIter,
                          double *res, double *resRel, double *history, double *ppStats)
{
    if (nPrev < 0 || nPrev > 3)
    {
        g_err = "nPrev must be 0..3 (BDF1..BDF4)";
        return -1;
    }
    return implicitStep(h, es, is, nPrev + 1, xPrevs, u, uRec, scanOrder, nIter, res, resRel, history, ppStats);
}

// kCur 0: ImplicitEulerDualTimeStep::Step (ODE.hpp:80-98); 1..4: the BDF step with kCur - 1
// earlier states.  Per iteration: fdtau (EvaluateDt, dTau / alpha), frhs, the unsteady residual
// sum_j c_j x_j / dt - x / dt + alpha frhs, fsolve (LUSGSMatrixInit and the sweeps with alphaDiag =
// alpha), fincrement, fstop.
static int implicitStep(const orc_vr *h, const orc_euler_settings *es, const orc_implicit_settings *is, int kCur,
                        const double *xPrevs, double *u, double *uRec, const int64_t *scanOrder, int *nIter,
                        double *res, double *resRel, double *history, double *ppStats)
{
    ORC_TRY
    const real alphaDiag = kCur ? kBDF[kCur - 1][0] : 1.0;
    const VR &v = h->v;
    idx_t N = v.nCell;
    size_t nU = (size_t)N * 5, nR = nU * v.M;
    std::vector<real> dTau(N), lam((size_t)v.nFace * 4), dlc(N), rhsF(nU), rhs(nU), xLast(u, u + nU), tmp(nR),
        jd(N), inc(nU), incNew(nU);
    const bool roeJ = es->useRoeJacobian != 0;
    std::vector<real> lamRoe(roeJ ? (size_t)v.nFace * 3 : 0), jdB(roeJ ? (size_t)N * 25 : 0),
        jinvB(roeJ ? (size_t)N * 25 : 0);
    const real gamma = es->gamma;
    const real rhoEps = smallReal * is->refRho * 1e-1, pEps = smallReal * is->refP * 1e-1;
    real resBase[5] = {0, 0, 0, 0, 0}, rr[5] = {0, 0, 0, 0, 0}, rel[5] = {0, 0, 0, 0, 0};
    idx_t nLimTot = 0;
    real alphaMinAll = 1;
    int it;
    for (it = 1; it <= is->maxIter; it++)
    {
        real dtMin;
        if (orc_eval_dt(h, es, u, is->CFL, 1e100, 1, dTau.data(), lam.data(), dlc.data(), &dtMin))
            throw std::runtime_error(g_err);
        if (kCur) // dTau *= 1 / alphaDiag (EulerSolver.hxx:693)
            for (idx_t i = 0; i < N; i++)
                dTau[i] *= 1. / alphaDiag;
        for (int r = 0; r < is->nRec; r++)
            vrSweep(h, uRec, u, scanOrder, tmp);
        if (orc_eval_rhs_lam(h, es, u, uRec, dlc.data(), rhsF.data(), roeJ ? lamRoe.data() : nullptr))
            throw std::runtime_error(g_err);
        if (!kCur) // crhs = xLast / dt - x / dt + frhs
            for (size_t q = 0; q < nU; q++)
            {
                real r = xLast[q] * (1.0 / is->dt);
                r += u[q] * (-1.0 / is->dt);
                r += rhsF[q];
                rhs[q] = r;
            }
        else // BDF: sum_j c_j x_j / dt - x / dt + alpha frhs (ODE.hpp:440-451)
            for (size_t q = 0; q < nU; q++)
            {
                real r = 0;
                r += xLast[q] * (kBDF[kCur - 1][1] / is->dt);
                for (int j = 0; j < kCur - 1; j++)
                    r += xPrevs[(size_t)j * nU + q] * (kBDF[kCur - 1][2 + j] / is->dt);
                r += u[q] * (-1. / is->dt);
                r += rhsF[q] * kBDF[kCur - 1][0];
                rhs[q] = r;
            }
        if (roeJ ? orc_lusgs_init_block(h, es, dTau.data(), is->dt, alphaDiag, u, lam.data(), lamRoe.data(),
                                        jdB.data(), jinvB.data())
                 : orc_lusgs_init(h, dTau.data(), is->dt, alphaDiag, lam.data(), jd.data()))
            throw std::runtime_error(g_err);
        std::fill(inc.begin(), inc.end(), 0.0);
        real sumInc[5];
        for (int s = 0; s <= is->sgsIter; s++)
            for (int dir = 1; dir >= 0; dir--)
            {
                int rc = roeJ ? orc_update_sgs_block(h, es, alphaDiag, rhs.data(), u, inc.data(), incNew.data(),
                                                     jinvB.data(), lam.data(), lamRoe.data(), dir, is->gsUpdate, 0,
                                                     scanOrder, sumInc)
                              : orc_update_sgs(h, es, alphaDiag, rhs.data(), u, inc.data(), incNew.data(), jd.data(),
                                               lam.data(), dir, is->gsUpdate, 0, scanOrder, sumInc);
                if (rc)
                    throw std::runtime_error(g_err);
                std::swap(inc, incNew);
            }
        // fincrement: positivity-limited, compressed increment
        {
            real aMin = 1;
            bool bad = false;
            idx_t nl = addIncrementPP(N, gamma, is->incrementRelax, rhoEps, pEps, u, inc.data(), aMin, bad);
            if (bad)
                throw std::runtime_error("the positivity-preserving increment failed (a state with non-positive "
                                         "density or internal energy)");
            nLimTot += nl;
            alphaMinAll = std::min(alphaMinAll, aMin);
        }
        for (int k = 0; k < 5; k++)
            rr[k] = 0;
        for (idx_t i = 0; i < N; i++)
            for (int k = 0; k < 5; k++)
                rr[k] += std::abs(rhs[i * 5 + k]);
        for (int k = 0; k < 5; k++)
        {
            resBase[k] = std::max(resBase[k], rr[k]);
            rel[k] = rr[k] / (resBase[k] + verySmallReal);
        }
        if (history)
            history[it - 1] = rel[0];
        if (it >= is->minIter && rel[0] < is->tol)
            break;
    }
    if (nIter)
        *nIter = std::min(it, is->maxIter);
    for (int k = 0; k < 5; k++)
    {
        if (res)
            res[k] = rr[k];
        if (resRel)
            resRel[k] = rel[k];
    }
    if (ppStats)
        ppStats[0] = (double)nLimTot, ppStats[1] = alphaMinAll;
    return 0;
    ORC_CATCH(-1)
}

int orc_cell_rhs_alpha(const double *u, const double *res, double gamma, double relax, double rhoEps, double pEps,
                       int compress, double *alpha, int *limited)
{
    bool lim = false, bad = false;
    *alpha = cellRHSAlpha(u, res, gamma, relax, rhoEps, pEps, compress, lim, bad);
    *limited = lim ? 1 : 0;
    return bad ? 1 : 0;
}

int orc_compress_inc(const double *u, const double *inc, double gamma, double rhoEps, double pEps, double *ret)
{
    bool bad = false;
    compressInc(u, inc, gamma, rhoEps, pEps, ret, bad);
    return bad ? 1 : 0;
}

int64_t orc_add_increment_pp(int64_t N, double gamma, double relax, double rhoEps, double pEps, double *u,
                             const double *du, double *alphaMin, int *bad)
{
    bool b = false;
    real am = 1;
    idx_t n = addIncrementPP(N, gamma, relax, rhoEps, pEps, u, du, am, b);
    if (alphaMin)
        *alphaMin = am;
    if (bad)
        *bad = b ? 1 : 0;
    return n;
}

} // extern "C"
