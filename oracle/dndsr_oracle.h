/*
 * dndsr_oracle.h -- C interface of the CPU oracle.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain C++17 FP64 restatement of the
 * reference's compact-finite-volume algorithms (CFDLAB-THU/DNDSR), used as the
 * checker for the HIP path.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it.  The product library never links
 * or calls anything in oracle/.
 *
 * Array conventions follow the reference's row layout (ArrayDof<m,n> rows are
 * contiguous column-major m x n blocks, src/DNDS/ArrayDerived/ArrayEigenMatrix.hpp:23,64):
 *   u      [nCell][nVars]
 *   uRec   [nCell][nVars][M]    (column-major M x nVars: element (k,v) at k + M*v)
 *   AinvB  [nCell][nFaceCell+1][M][M] (block 0 = A^-1, block 1+ic2f = A^-1 B, row-major)
 *   Ainvb  [nCell][nFaceCell][M]
 */
#ifndef DNDSR_ORACLE_H
#define DNDSR_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* element types */
enum { ORC_LINE2 = 1, ORC_QUAD4 = 2, ORC_HEX8 = 3, ORC_TRI3 = 4, ORC_TET4 = 5 };

typedef struct
{
    int dim;                 /* geometric dimension, 2 or 3 */
    int64_t nNode;
    const double *coords;    /* [nNode][3] */
    int64_t nCell;
    const int32_t *cellType; /* [nCell] */
    const int64_t *cell2nodePtr, *cell2node;
    const int64_t *cell2facePtr, *cell2face; /* faces of each cell, element-local order */
    int64_t nFace;
    const int32_t *faceType; /* [nFace] */
    const int64_t *face2nodePtr, *face2node;
    const int64_t *face2cell;  /* [nFace][2]; -1 = boundary */
    const double *faceShift;   /* [nFace][3]: translation from the back (L) cell frame to the front (R) cell frame */
    /* boundary zone (0 .. nBCZone-1) of every face with face2cell[f][1] = -1 (the mesh's
     * face zone id, UnstructuredMesh::GetFaceZone); entries of interior faces are ignored.
     * NULL when the mesh has no boundary face. */
    const int32_t *faceZone;
    int32_t nBCZone;
} orc_mesh_desc;

typedef struct
{
    int maxOrder;
    int intOrder;
    int dirWeightScheme;  /* 0 Factorial, 1 HQM_OPT */
    int geomWeightScheme; /* 0 GWNone, 1 HQM_SD */
    double geomWeightBias;
    double geomWeightPower;
    double scaleMultiplier;
    double jacobiRelax;
    int nIterCellSmoothScale;
    /* boundary faces: face weight of order p multiplied by zoneDirWeight[zone][p] * bcWeight
     * (the id2faceDircWeight callback of ConstructBaseAndWeight,
     * VariationalReconstruction.cpp:438-441); NULL = 1 for every zone and order */
    double bcWeight;
    const double *zoneDirWeight; /* [nBCZone][maxOrder+1] */
    /* the solver loops (orc_ssprk3, orc_implicit_euler_step) sweep the reconstruction in
     * place (orc_vr_sor) instead of by Jacobi */
    int SORInstead;
    /* maxOrder 1 only: 0 = VR; 1 = Gauss rule, 2 = least squares (subs2ndOrder, VRSettings.hpp:49-50)
     * replace every VR sweep by a gradient of the means */
    int subs2ndOrder;
    int subs2ndOrderGGScheme; /* 1: distance-weighted face value, 0: plain average */
} orc_vr_settings;

typedef struct
{
    double gamma;
    int rsType;         /* Gas::RiemannSolverType numbering: Roe=1, HLLC=2, HLLEP=3, HLLEP_V1=21, Roe_M1..M9=11..19 */
    int rsMeanValueEig; /* batched mean-value eigen path for Roe_M1..Roe_M5 */
    double rsFixScale;
    double rsIncFScale;
    /* laminar viscous terms (idealGasProperty, EulerEvaluatorSettings.hpp:288-326);
     * viscous = 0: Euler.  muModel 0 constant, 1 Sutherland, 2 constant nu */
    int viscous;
    int muModel;
    double Rgas, muGas, prGas, TRef, CSutherland;
    int useRoeJacobian; /* EulerEvaluatorSettings::useRoeJacobian: block JDiag with the Roe |A| */
} orc_euler_settings;

typedef struct orc_vr orc_vr;

const char *orc_last_error(void);
/* OpenMP threads for the parallel loops; rhsGather selects the reference's OpenMP
 * face-buffer + cell-gather scatter instead of the serial face-order scatter */
int orc_set_threads(int nThreads, int rhsGather);

orc_vr *orc_vr_create(const orc_mesh_desc *mesh, const orc_vr_settings *s);
void orc_vr_destroy(orc_vr *h);
int orc_vr_info(const orc_vr *h, int64_t *nCell, int64_t *nFace, int *M, int *maxFaceCell,
                int *nGFace, int *nGCell, double *globalVol);

/* metric getters (copy out) */
int orc_get_cell_metrics(const orc_vr *h, double *vol, double *bary, double *cent,
                         double *scale, double *smoothRatio, double *moment /*[nCell][M]*/);
int orc_get_face_metrics(const orc_vr *h, double *area, double *meanNorm /*[nFace][3]*/,
                         double *pts /*[nFace][G][3]*/, double *norms /*[nFace][G][3]*/,
                         double *wJDet /*[nFace][G]*/, double *faceL, double *faceWeight /*[nFace][maxOrder+1]*/);
int orc_get_cell_quad(const orc_vr *h, double *pts /*[nCell][Gc][3]*/, double *wJDet /*[nCell][Gc]*/);
int orc_get_vr_coeffs(const orc_vr *h, double *AinvB, double *Ainvb);

/* fields */
int orc_cell_average(const orc_vr *h, int fieldId, const double *params, int nVars, double *u);
double orc_rec_error_l1(const orc_vr *h, int fieldId, const double *params, const double *u,
                        const double *uRec);

/* VR iteration: one Jacobi sweep uRecNew = (1-w) uRec + w * sum(...) */
int orc_vr_iter(const orc_vr *h, int nVars, const double *uRec, const double *u, double *uRecNew);
/* second-order substitute (settings.subs2ndOrder 1 or 2, maxOrder 1): uRec from the means only;
 * orc_vr_iter / orc_vr_sor dispatch to it and orc_vr_iter_diff returns zeros, as the reference */
int orc_vr_2nd(const orc_vr *h, int nVars, const double *u, double *uRec);
/* VR iteration, SOR form: uRec swept in place, cells in scanOrder (NULL = natural order) */
int orc_vr_sor(const orc_vr *h, int nVars, double *uRec, const double *u, const int64_t *scanOrder);
/* uRecNew = uRecDiff - sum_f AinvB_f * uRecDiff_j - A^-1 BCdiff (VR Jacobian action); the
 * boundary term needs the base state uRec, u (may be NULL on meshes without boundary faces) */
int orc_vr_iter_diff(const orc_vr *h, int nVars, const double *uRec, const double *u, const double *uRecDiff,
                     double *uRecNew);

/* Boundary conditions (EulerBCType codes, src/Euler/EulerBC.hpp:35-48): per zone a type and a
 * value vector of 5 conservative variables (far-field / inflow state; unused otherwise).
 * Used by the VR sweep for nVars = 5 (with gamma), by EvaluateRHS and by EvaluateDt. */
enum { ORC_BC_FAR = 1, ORC_BC_WALL = 2, ORC_BC_WALL_INVIS = 3, ORC_BC_WALL_ISO = 4, ORC_BC_OUT = 5, ORC_BC_OUTP = 6,
       ORC_BC_IN = 7, ORC_BC_IN_PSTS = 8, ORC_BC_SYM = 9 };
int orc_set_bc(orc_vr *h, int nZone, const int32_t *type, const double *value, double gamma, double Rgas);
/* VR sweep boundary values given as data (the TFBoundary callback of the reference tests):
 * value[f][g][var] at boundary face f, quadrature point g; used when no Euler BC applies. */
int orc_set_vr_dirichlet(orc_vr *h, int nVars, const double *value);
/* ghost state of one boundary point (generateBoundaryValue, EulerEvaluator_EvaluateDt.hxx:1743-1838) */
int orc_bc_ghost(int type, const double *UL, const double *n, const double *zoneValue, double gamma, double Rgas,
                 double *UR);

/* Riemann flux at one point; F gets the flux, lam[3] = lam0, lam123, lam4 */
int orc_riemann(int rsType, int dim, const double *UL, const double *UR, const double *ULm,
                const double *URm, const double *n, double gamma, double dLambda,
                double fixScale, double incFScale, double *F, double *lam);

/* EulerEvaluator operators (nVars = 5, inviscid, periodic / interior faces) */
int orc_eval_dt(const orc_vr *h, const orc_euler_settings *es, const double *u, double CFL,
                double maxDt, int useLocalDt, double *dt, double *lambdaFace /*[nFace][4]*/,
                double *deltaLambdaCell, double *dtMin);
int orc_eval_rhs(const orc_vr *h, const orc_euler_settings *es, const double *u,
                 const double *uRec, const double *deltaLambdaCell, double *rhs);
/* EvaluateRHS that also writes the face-averaged Riemann eigenvalues lambdaFace0/123/4
 * ([nFace][3], EvaluateRHS.hxx:527-544) when lamRoe is not NULL */
int orc_eval_rhs_lam(const orc_vr *h, const orc_euler_settings *es, const double *u,
                     const double *uRec, const double *deltaLambdaCell, double *rhs, double *lamRoe);
/* force on the boundary faces of zone `zone` (momentum flux out of the domain, Riemann minus
 * viscous flux integrated over the faces), force[3] */
int orc_zone_force(const orc_vr *h, const orc_euler_settings *es, const double *u, const double *uRec,
                   const double *deltaLambdaCell, int zone, double *force);
/* LUSGSMatrixInit with useRoeJacobian (EulerEvaluator.hxx:37-159): 5x5 blocks [N][5][5]
 * (row-major) and their inverses (JacobianDiagBlock::GetInvert, EulerJacobian.hpp:153-190);
 * lamRoe [nFace][3] = lambdaFace0/123/4, or NULL for EvaluateDt's (lambdaFace[.][3]) */
int orc_lusgs_init_block(const orc_vr *h, const orc_euler_settings *es, const double *dTau, double dt,
                         double alphaDiag, const double *u, const double *lambdaFace, const double *lamRoe,
                         double *jdiag, double *jinv);
/* UpdateSGS with the block JDiag and the Roe off-diagonal terms */
int orc_update_sgs_block(const orc_vr *h, const orc_euler_settings *es, double alphaDiag, const double *rhs,
                         const double *u, const double *uInc, double *uIncNew, const double *jinv,
                         const double *lambdaFace, const double *lamRoe, int forward, int gsUpdate,
                         int uIncIsZero, const int64_t *scanOrder, double *sumInc);
int orc_lusgs_init(const orc_vr *h, const double *dTau, double dt, double alphaDiag,
                   const double *lambdaFace, double *jdiag);
int orc_update_sgs(const orc_vr *h, const orc_euler_settings *es, double alphaDiag,
                   const double *rhs, const double *u, const double *uInc, double *uIncNew,
                   const double *jdiag, const double *lambdaFace, int forward, int gsUpdate,
                   int uIncIsZero, const int64_t *scanOrder, double *sumInc);
/* the transpose of the scalar LU-SGS matrix, swept as UpdateSGS (the adjoint preconditioner) */
int orc_update_sgs_transpose(const orc_vr *h, const orc_euler_settings *es, double alphaDiag, const double *rhs,
                             const double *u, const double *uInc, double *uIncNew, const double *jdiag,
                             const double *lambdaFace, int forward, int gsUpdate, int uIncIsZero,
                             const int64_t *scanOrder, double *sumInc);
/* the transpose of the block LU-SGS matrix of orc_update_sgs_block, swept in the same order */
int orc_update_sgs_block_transpose(const orc_vr *h, const orc_euler_settings *es, double alphaDiag, const double *rhs,
                                   const double *u, const double *uInc, double *uIncNew, const double *jinv,
                                   const double *lambdaFace, const double *lamRoe, int forward, int gsUpdate,
                                   int uIncIsZero, const int64_t *scanOrder, double *sumInc);
/* LUSGSMatrixVec (scalar Jacobian): Ax = D x + sum_f (alpha/2) A_f / V fluxJacobian0(u_j) x_j */
int orc_lusgs_matvec(const orc_vr *h, const orc_euler_settings *es, double alphaDiag, const double *u,
                     const double *x, const double *jdiag, const double *lambdaFace, double *Ax);
/* ModelEvaluator::EvaluateRHS (scalar advection-diffusion, 2-D): u [N][nVars], uRec [N][nVars][M]
 * (NULL = first order), rhs [N][nVars] */
int orc_model_rhs(const orc_vr *h, int nVars, double ax, double ay, double sigma, const double *u, const double *uRec,
                  double *rhs);
/* full explicit SSP-RK3 steps: nStep steps of (EvaluateDt once, 3 x (nRec VR + RHS)); the
 * VR sweeps run in natural cell order when SORInstead */
int orc_ssprk3(const orc_vr *h, const orc_euler_settings *es, double *u, double *uRec, int nStep,
               double CFL, int nRec, double dtFixed);

/* implicit-Euler dual-time step (ImplicitEulerDualTimeStep::Step, src/Solver/ODE.hpp:80-98) */
typedef struct
{
    double dt, CFL;
    int nRec, sgsIter, gsUpdate, maxIter, minIter;
    double tol;
    int fp32Sweeps; /* ignored by the oracle (FP64 only) */
    double incrementRelax, refRho, refP;
} orc_implicit_settings;
int orc_implicit_euler_step(const orc_vr *h, const orc_euler_settings *es, const orc_implicit_settings *is,
                            double *u, double *uRec, const int64_t *scanOrder, int *nIter, double *res,
                            double *resRel, double *history, double *ppStats /* [2]: limited cells, min alpha */);

/* ImplicitBDFDualTimeStep::Step (src/Solver/ODE.hpp:427-477) with nPrev (0..3) earlier states
 * xPrevs [nPrev][N][5], newest first: the (nPrev + 1)-step BDF formula */
int orc_implicit_bdf_step(const orc_vr *h, const orc_euler_settings *es, const orc_implicit_settings *is, int nPrev,
                          const double *xPrevs, double *u, double *uRec, const int64_t *scanOrder, int *nIter,
                          double *res, double *resRel, double *history, double *ppStats);

/* positivity of the increment (ideal gas, 5 conservative variables per cell):
 * EvaluateCellRHSAlpha for one cell (returns nonzero where the reference asserts) */
int orc_cell_rhs_alpha(const double *u, const double *res, double gamma, double relax, double rhoEps, double pEps,
                       int compress, double *alpha, int *limited);
/* CompressInc of one cell's increment */
int orc_compress_inc(const double *u, const double *inc, double gamma, double rhoEps, double pEps, double *ret);
/* fincrement over N cells (row layout [N][5]): u += CompressInc(u, alpha_i du); returns the
 * number of limited cells */
int64_t orc_add_increment_pp(int64_t N, double gamma, double relax, double rhoEps, double pEps, double *u,
                             const double *du, double *alphaMin, int *bad);

#ifdef __cplusplus
}
#endif
#endif
