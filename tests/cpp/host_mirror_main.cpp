// host_mirror_main.cpp -- drives the C++17 host mirror (include/dndsr_amd.hpp) the way a
// reference caller would (test_EulerEvaluator.cpp style): build the VR, one Jacobi VR
// sweep, EvaluateDt (local), EvaluateRHS, LUSGSMatrixInit, one forward Gauss-Seidel
// sweep.  Inputs / outputs are .npy files; the GPU test compares them with the oracle.
//   usage: host_mirror_test <indir> <outdir>
#include <cstdio>

#include "../../include/dndsr_amd.hpp"
#include "npy_io.hpp"

using namespace dndsr_amd;

template <EulerModel model>
static void run(const std::string &in, const std::string &out, const Mesh &mesh, const std::vector<double> &par)
{
    constexpr int dim = modelDim<model>;
    Context ctx(0);
    VRSettings vs;
    vs.maxOrder = (int)par[1];
    vs.intOrder = (int)par[2];
    vs.SORInstead = par.size() > 5 && par[5] != 0; // in-place colour-by-colour sweep
    VariationalReconstruction<dim> vfv(ctx, mesh, vs);
    vfv.ConstructMetrics();
    vfv.ConstructBaseAndWeight();
    vfv.ConstructRecCoeff();
    EulerEvaluatorSettings es;
    es.riemannSolverType = (int)par[3];
    EulerEvaluator<model> ev(vfv, es);
    const dndsr_amd::index N = vfv.NumCell();
    const int M = vfv.NumRecModes();

    ArrayDOF u, dt, jd, rhs, inc, incNew;
    ArrayREC rec, recNew;
    vfv.BuildUDof(u, 5);
    vfv.BuildUDof(dt, 1);
    vfv.BuildUDof(jd, 1);
    vfv.BuildUDof(rhs, 5);
    vfv.BuildUDof(inc, 5);
    vfv.BuildUDof(incNew, 5);
    vfv.BuildURec(rec, 5);
    vfv.BuildURec(recNew, 5);
    u.fromHostRows(npy::load<double>(in + "/u.npy"));
    rec.fromHostRows(npy::load<double>(in + "/uRec.npy"));

    vfv.DoReconstructionIter(rec, recNew, u); // putIntoNew = false: rec holds the new iterate (either form)
    double dtMin = 0;
    ev.EvaluateDt(dt, u, rec, par[4], dtMin, 1e100, true);
    ev.EvaluateRHS(rhs, u, rec);
    ev.LUSGSMatrixInit(jd, dt, 1e100, 1.0, u, rec);
    inc.setConstant(0.0);
    std::array<double, 5> sumInc{};
    ev.UpdateSGS(1.0, 0.0, rhs, u, inc, incNew, jd, true, true, sumInc);
    std::vector<double> norm = ev.EvaluateNorm(rhs, 1);

    std::vector<int32_t> col((size_t)N);
    check(dndsr_get_colors(ctx.handle(), col.data()));
    npy::save(out + "/uRec.npy", rec.toHostRows(), {(size_t)N, 5, (size_t)M});
    npy::save(out + "/dt.npy", dt.toHostRows(), {(size_t)N});
    npy::save(out + "/rhs.npy", rhs.toHostRows(), {(size_t)N, 5});
    npy::save(out + "/jd.npy", jd.toHostRows(), {(size_t)N});
    npy::save(out + "/inc.npy", incNew.toHostRows(), {(size_t)N, 5});
    npy::save(out + "/norm.npy", norm, {5});
    npy::save(out + "/colors.npy", col, {(size_t)N});
    npy::save(out + "/dtMin.npy", std::vector<double>{dtMin}, {1});
}

int main(int argc, char **argv)
{
    if (argc != 3)
    {
        std::fprintf(stderr, "usage: %s <indir> <outdir>\n", argv[0]);
        return 2;
    }
    const std::string in = argv[1], out = argv[2];
    try
    {
        Mesh mesh;
        const std::vector<double> par = npy::load<double>(in + "/params.npy"); // dim, maxOrder, intOrder, rs, CFL[, SORInstead]
        mesh.dim = (int)par[0];
        mesh.coords = npy::load<double>(in + "/coords.npy");
        mesh.cellType = npy::load<int32_t>(in + "/cellType.npy");
        mesh.faceType = npy::load<int32_t>(in + "/faceType.npy");
        mesh.cell2nodePtr = npy::load<int64_t>(in + "/cell2nodePtr.npy");
        mesh.cell2node = npy::load<int64_t>(in + "/cell2node.npy");
        mesh.cell2facePtr = npy::load<int64_t>(in + "/cell2facePtr.npy");
        mesh.cell2face = npy::load<int64_t>(in + "/cell2face.npy");
        mesh.face2nodePtr = npy::load<int64_t>(in + "/face2nodePtr.npy");
        mesh.face2node = npy::load<int64_t>(in + "/face2node.npy");
        mesh.face2cell = npy::load<int64_t>(in + "/face2cell.npy");
        mesh.faceShift = npy::load<double>(in + "/faceShift.npy");
        if (mesh.dim == 3)
            run<NS_3D>(in, out, mesh, par);
        else
            run<NS>(in, out, mesh, par);
    }
    catch (const std::exception &e)
    {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
    std::printf("host mirror driver done\n");
    return 0;
}
