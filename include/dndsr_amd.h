/*
 * dndsr_amd.h -- C-ABI of the MI355X (gfx950) compact-finite-volume hot path.
 *
 * The reference (CFDLAB-THU/DNDSR) exposes this path as C++ templates:
 *   VariationalReconstruction<dim>   src/CFV/VariationalReconstruction.hpp:108
 *   EulerEvaluator<model>            src/Euler/EulerEvaluator.hpp:76
 * Each entry point below replaces one of their member functions; the cited
 * line is the declaration it stands in for.  All entry points:
 *   - return 0 on success, nonzero on error (message via dndsr_last_error());
 *     no C++ exception crosses the boundary;
 *   - take DEVICE pointers for every field array (caller-owned, never
 *     reallocated) and enqueue their kernels on the context's stream; results
 *     are ready after dndsr_sync() unless a host output is requested;
 *   - are not thread-safe per context.
 *
 * Device field layout (structure of arrays, cell index fastest):
 *   u, rhs, uInc      [5][nCell]        (rho, rho u, rho v, rho w, E)
 *   uRec              [nVars][M][nCell] (M = NDOF(maxOrder) - 1 modes)
 *   dt, deltaLambdaCell, jdiag          [nCell]
 *   lambdaFace        [4][nFace]        (lambdaFace, lambdaFaceC, lambdaFaceVis, lambdaFace0/123/4)
 * dndsr_h2d_cellfield / dndsr_d2h_cellfield convert from / to the reference's
 * row layout (ArrayDof rows of column-major m x n blocks,
 * src/DNDS/ArrayDerived/ArrayEigenMatrix.hpp:23,64): a host row of nComp values
 * per cell, with uRec's (k, var) entry at component var*M + k.
 */
#ifndef DNDSR_AMD_H
#define DNDSR_AMD_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct dndsr_ctx dndsr_ctx;

/* element type codes of dndsr_mesh_desc */
enum { DNDSR_LINE2 = 1, DNDSR_QUAD4 = 2, DNDSR_HEX8 = 3, DNDSR_TRI3 = 4, DNDSR_TET4 = 5 };

/* Mesh as handed over by the caller (host memory).  Periodic faces are stored once,
 * in the frame of face2cell[f][0]; faceShift[f] carries a face point into the frame of
 * face2cell[f][1] (UnstructuredMesh::periodicInfo, src/Geom/Mesh/Mesh.hpp:37).
 * A boundary face has face2cell[f][1] = -1, its normal pointing out of the domain, and
 * a boundary zone faceZone[f] in 0 .. nBCZone-1 (the face zone id of
 * UnstructuredMesh::GetFaceZone); faceZone may be NULL when there is no boundary face. */
typedef struct
{
    int dim;
    int64_t nNode;
    const double *coords;    /* [nNode][3] */
    int64_t nCell;
    const int32_t *cellType; /* [nCell] */
    const int64_t *cell2nodePtr, *cell2node;
    const int64_t *cell2facePtr, *cell2face;
    int64_t nFace;
    const int32_t *faceType; /* [nFace] */
    const int64_t *face2nodePtr, *face2node;
    const int64_t *face2cell; /* [nFace][2], -1 = boundary */
    const double *faceShift;  /* [nFace][3] */
    const int32_t *faceZone;  /* [nFace]: zone of each boundary face (entries of other faces ignored) */
    int32_t nBCZone;
} dndsr_mesh_desc;

/* VRSettings subset used on this path (src/CFV/VRSettings.hpp:37-212) */
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
    /* boundary faces: the face weight of derivative order p is multiplied by
     * zoneDirWeight[zone][p] * bcWeight (the id2faceDircWeight callback of
     * ConstructBaseAndWeight, VariationalReconstruction.cpp:438-441; EulerEvaluator
     * passes 1 for the value and 0 for derivatives, 0 everywhere for BCOut);
     * NULL = 1 for every zone and order */
    double bcWeight;
    const double *zoneDirWeight; /* host, [nBCZone][maxOrder+1] */
    /* SORInstead (VRSettings.hpp:43): the solver loops (dndsr_ssprk3, dndsr_implicit_euler_step)
     * sweep the reconstruction in place, multicolour SOR (dndsr_vr_iter mode 4), instead of
     * by relaxed Jacobi.  0 = Jacobi. */
    int SORInstead;
    /* subs2ndOrder (VRSettings.hpp:49-50), maxOrder 1 only: 0 = VR; 1 = Gauss rule, 2 = least
     * squares: every VR sweep (dndsr_vr_iter, any mode, and the solver loops) is replaced by a
     * gradient of the means (DoReconstruction2nd, VariationalReconstruction_Reconstruction.hxx:
     * 184-483; GGMP = 11 is rejected) and dndsr_vr_iter_diff returns zeros.
     * subs2ndOrderGGScheme: 1 = distance-weighted face value, 0 = plain average. */
    int subs2ndOrder;
    int subs2ndOrderGGScheme;
} dndsr_vr_settings;

/* EulerEvaluatorSettings subset (src/Euler/EulerEvaluatorSettings.hpp:53-80) */
typedef struct
{
    double gamma;
    int rsType;         /* Gas::RiemannSolverType: Roe=1, HLLC=2, HLLEP=3, Roe_M1..M8=11..18, HLLEP_V1=21 */
    int rsMeanValueEig; /* cell-mean eigenvalues shared by a face's points (Roe_M1..M5) */
    double rsFixScale;
    double rsIncFScale;
    /* laminar Navier-Stokes terms (idealGasProperty, EulerEvaluatorSettings.hpp:280-330;
     * models NS / NS_3D): 0 = Euler.  muModel: 0 constant mu = muGas, 1 Sutherland's law
     * mu = muGas (T/TRef)^1.5 (TRef + CSutherland) / (T + CSutherland), 2 constant
     * nu (mu = muGas rho); conductivity Cp mu / prGas with Cp = Rgas gamma / (gamma - 1). */
    int viscous;
    int muModel;
    double Rgas;
    double muGas;
    double prGas;
    double TRef;
    double CSutherland;
    /* useRoeJacobian (EulerEvaluatorSettings.hpp:54): the implicit step builds 5x5 JDiag blocks with
     * the Roe |A| and sweeps with the Roe off-diagonal terms (dndsr_lusgs_init_block,
     * dndsr_update_sgs_block); 0 = the scalar Jacobian */
    int useRoeJacobian;
} dndsr_euler_settings;

/* ---- context ---------------------------------------------------------- */
int dndsr_create(dndsr_ctx **ctx, int device);
int dndsr_destroy(dndsr_ctx *ctx);
const char *dndsr_last_error(void);
/* Waits for the context's stream.  Returns an error when an EvaluateRHS launched since the last
 * synchronisation produced a non-finite residual in some cell (the reference asserts there,
 * src/Euler/EulerEvaluator_EvaluateRHS.hxx:387-398,782-788); the flag is cleared by the call. */
int dndsr_sync(dndsr_ctx *ctx);
/* enqueue on the caller's stream (e.g. the framework's current stream) instead of the
 * context's own one, so that the library's work is ordered with the caller's; NULL selects
 * the device's default (null) stream */
int dndsr_set_stream(dndsr_ctx *ctx, void *hipStream);

/* ---- setup ------------------------------------------------------------ */
/* ConstructMetrics + ConstructBaseAndWeight + ConstructRecCoeff, computed on the device.
 * Replaces VariationalReconstruction<dim>::ConstructMetrics / ConstructBaseAndWeight /
 * ConstructRecCoeff, src/CFV/VariationalReconstruction.hpp:282-289. */
int dndsr_vr_construct(dndsr_ctx *ctx, const dndsr_mesh_desc *mesh, const dndsr_vr_settings *s);
int dndsr_mesh_info(dndsr_ctx *ctx, int64_t *nCell, int64_t *nFace, int *M, int *maxFaceCell, int *nGFace,
                    int *nGCell, int *nColor, double *globalVol);
/* host copies of metrics in row layout:
 *   vol[N], bary[N][3], cent[N][3], scale[N], smoothRatio[N], moment[N][M]        */
int dndsr_get_cell_metrics(dndsr_ctx *ctx, double *vol, double *bary, double *cent, double *scale,
                           double *smoothRatio, double *moment);
/*   area[F], meanNorm[F][3], pts[F][G][3], norms[F][G][3], wJDet[F][G], faceL[F], faceWeight[F][P+1] */
int dndsr_get_face_metrics(dndsr_ctx *ctx, double *area, double *meanNorm, double *pts, double *norms,
                           double *wJDet, double *faceL, double *faceWeight);
/*   pts[N][Gc][3], wJDet[N][Gc] (cell quadrature) */
int dndsr_get_cell_quad(dndsr_ctx *ctx, double *pts, double *wJDet);
/* matrixAAInvB / vectorAInvB in the reference's block order
 * (src/CFV/VariationalReconstruction.hpp:169-197): AinvB[N][K+1][M][M] (block 0 = A^-1),
 * Ainvb[N][K][M].  get = download; set = upload coefficients computed elsewhere
 * (e.g. by the reference's own ConstructRecCoeff). */
int dndsr_get_vr_coeffs(dndsr_ctx *ctx, double *AinvB, double *Ainvb);
int dndsr_set_vr_coeffs(dndsr_ctx *ctx, const double *AinvB, const double *Ainvb);
/* multicolour ordering used by the Gauss-Seidel sweeps: color[N] */
int dndsr_get_colors(dndsr_ctx *ctx, int32_t *color);

/* ---- layout helpers ---------------------------------------------------- */
int dndsr_h2d_cellfield(dndsr_ctx *ctx, const double *hostRows, double *dev, int nComp);
int dndsr_d2h_cellfield(dndsr_ctx *ctx, const double *dev, double *hostRows, int nComp);

/* ---- operators ---------------------------------------------------------- */
/* DoReconstructionIter, Jacobi form (SORInstead = false):
 *   mode 0: uRecNew = (1-relax) uRec + relax sum_f (A^-1 B_f uRec_j + A^-1 b_f (u_j - u_i))
 *   mode 1: recordInc (uRecNew = uRec - sum_f ...), mode 2: recordInc with uRecIsZero.
 * SOR form (SORInstead = true, putIntoNew = false; :534-589):
 *   mode 4: uRec is updated in place, colour by colour (dndsr_get_colors), each cell as
 *   uRec_i = (1-relax) uRec_i, then per face in cell-face order uRec_i += relax (...), the
 *   boundary term evaluated on the uRec_i accumulated so far; uRecNew is not used (may be
 *   NULL).  Equals the reference's serial sweep when cells are numbered colour by colour.
 * relax < 0 uses settings.jacobiRelax.
 * Replaces VariationalReconstruction<dim>::DoReconstructionIter,
 * src/CFV/VariationalReconstruction.hpp:984-992. */
int dndsr_vr_iter(dndsr_ctx *ctx, int nVars, double *uRec, double *uRecNew, const double *u, int mode,
                  double relax);
/* DoReconstructionIterDiff: uRecNew = uRecDiff - sum_f A^-1 B_f uRecDiff_j - A^-1 BCdiff,
 * src/CFV/VariationalReconstruction.hpp:1012-1018.  The boundary term (GetBoundaryRHSDiff,
 * VariationalReconstruction_Reconstruction.hxx:98-176) is the difference of the boundary
 * values at the point states of uRec + uRecDiff and of uRec, so uRec and u are needed on
 * meshes with boundary faces (they may be NULL otherwise). */
int dndsr_vr_iter_diff(dndsr_ctx *ctx, int nVars, const double *uRec, const double *u, const double *uRecDiff,
                       double *uRecNew);

/* ---- boundary conditions ----------------------------------------------------------
 * Boundary zones of the Euler path (BoundaryHandler / EulerBCType, src/Euler/EulerBC.hpp:35-48):
 * type[z] is an EulerBCType code (BCFar 1, BCWall 2, BCWallInvis 3, BCWallIsothermal 4, BCOut 5,
 * BCOutP 6, BCIn 7, BCInPsTs 8, BCSym 9; BCSpecial is rejected) and value[z][5] the zone's
 * values (host; may be NULL when no zone needs one): the conservative far-field / inflow state
 * (BCFar, BCOutP, BCIn), the wall temperature in value[z][0] (BCWallIsothermal), or total
 * pressure, total temperature and flow direction (p0, T0, dx, dy, dz) (BCInPsTs).  The ghost
 * states follow generateBoundaryValue (EulerEvaluator_EvaluateDt.hxx:1743-2455): EvaluateRHS
 * takes them at every face point and, for the Roe average, at the face centre; the VR sweep
 * with nVars = 5 takes them at the compressed point state (EulerSolver.hxx:163-175) with the
 * given gamma.  The two gas-dependent types need the gas constant: dndsr_set_bc_gas takes it
 * (idealGasProperty.Rgas, EulerEvaluatorSettings.hpp:291-300); dndsr_set_bc rejects them. */
int dndsr_set_bc(dndsr_ctx *ctx, int nZone, const int32_t *type, const double *value, double gamma);
int dndsr_set_bc_gas(dndsr_ctx *ctx, int nZone, const int32_t *type, const double *value, double gamma, double Rgas);
/* Boundary values of the VR sweep given as data, for any nVars (the TFBoundary callback of
 * DoReconstructionIter, VariationalReconstruction.hpp:893-901, evaluated by the caller):
 * value is a device array [nVars][nGFace][nFace] of which the boundary faces are read; the
 * library keeps the pointer.  Used when the Euler table does not apply (nVars != 5 or no
 * dndsr_set_bc).  NULL clears it. */
int dndsr_set_vr_dirichlet(dndsr_ctx *ctx, int nVars, const double *value);

/* EvaluateDt, src/Euler/EulerEvaluator.hpp:359-366.  dtMinHost may be NULL (then no
 * host synchronisation happens; the minimum stays in dtMinDev if non-NULL). */
int dndsr_eval_dt(dndsr_ctx *ctx, const dndsr_euler_settings *es, const double *u, double CFL, double maxDt,
                  int useLocalDt, double *dt, double *lambdaFace, double *deltaLambdaCell, double *dtMinDev,
                  double *dtMinHost);
/* EvaluateRHS, src/Euler/EulerEvaluator.hpp:399-409 (inviscid, plus the laminar viscous
 * fluxes when es->viscous; uRec may be NULL for the first-order residual; deltaLambdaCell
 * feeds HLLE+ and may be NULL = 0).  Asynchronous: a non-finite residual is reported by the
 * next dndsr_sync (a NaN or infinite input is a value, not a fault: every kernel completes). */
int dndsr_eval_rhs(dndsr_ctx *ctx, const dndsr_euler_settings *es, const double *u, const double *uRec,
                   const double *deltaLambdaCell, double *rhs);
/* EvaluateRHS that allocates nothing: no counterpart in the reference.  For callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops).  Same arguments and
 * result as dndsr_eval_rhs; its face-flux scratch is a static device array of the library holding
 * meshes of up to 1048576 faces (40 MiB per GPU, held until the process ends; larger meshes are
 * rejected with an error).  Every call on one GPU shares that array, so calls that overlap on one
 * GPU (two contexts or streams) race.  It does not raise the non-finite flag read by dndsr_sync
 * and records no timing events. */
int dndsr_eval_rhs_noalloc(dndsr_ctx *ctx, const dndsr_euler_settings *es, const double *u, const double *uRec,
                           const double *deltaLambdaCell, double *rhs);
/* ModelEvaluator::EvaluateRHS (src/CFV/ModelEvaluator.hpp:118-119; ModelEvaluator.cpp:6-115,
 * default options): the reference's scalar advection-diffusion test physics on a 2-D mesh, for
 * nVars (1..5) independent variables u [nVars][nCell] with reconstruction uRec [nVars][M][nCell]
 * (NULL = first order); advection velocity (ax, ay) with upwinding, diffusion sigma >= 0 through
 * the mean face gradient plus the jump over 2 min(V_L, V_R) / A_f.  Boundary faces see a zero
 * exterior value (the model's generateBoundaryValue).  rhs [nVars][nCell]. */
int dndsr_model_rhs(dndsr_ctx *ctx, int nVars, double ax, double ay, double sigma, const double *u,
                    const double *uRec, double *rhs);
/* LUSGSMatrixInit (scalar Jacobian), src/Euler/EulerEvaluator.hpp:427-434. */
int dndsr_lusgs_init(dndsr_ctx *ctx, const double *dTau, double dt, double alphaDiag, const double *lambdaFace,
                     double *jdiag);
/* UpdateSGS, src/Euler/EulerEvaluator.hpp:513-522, in multicolour order.  sumInc (5
 * values) is returned on the host when sumIncHost is non-NULL (synchronises). */
int dndsr_update_sgs(dndsr_ctx *ctx, const dndsr_euler_settings *es, double alphaDiag, const double *rhs,
                     const double *u, const double *uInc, double *uIncNew, const double *jdiag,
                     const double *lambdaFace, int forward, int gsUpdate, int uIncIsZero, double *sumIncHost);
/* A sweep of the transpose of the scalar LU-SGS matrix that dndsr_update_sgs sweeps, in the same
 * multicolour order: y_i = (rhs_i - sum_j c_ji (A(u_i) n_ji - lambda_f I)^T y_j) / D_i with
 * c_ji = (alphaDiag/2) A_f / V_j.  No counterpart in the reference: the preconditioner of the
 * discrete-adjoint fixed point (dndsr_amd.adjoint.SteadyAdjoint). */
int dndsr_update_sgs_transpose(dndsr_ctx *ctx, const dndsr_euler_settings *es, double alphaDiag, const double *rhs,
                               const double *u, const double *uInc, double *uIncNew, const double *jdiag,
                               const double *lambdaFace, int forward, int gsUpdate, int uIncIsZero,
                               double *sumIncHost);
/* FP32 sweep for the mixed-precision path: every array is float (same layouts). */
int dndsr_update_sgs_f32(dndsr_ctx *ctx, const dndsr_euler_settings *es, float alphaDiag, const float *rhs,
                         const float *u, const float *uInc, float *uIncNew, const float *jdiag,
                         const float *lambdaFace, int forward, int gsUpdate, int uIncIsZero, double *sumIncHost);
/* The Roe-Jacobian block LU-SGS (useRoeJacobian).
 * dndsr_eval_rhs_roe_lambda: the face eigenvalues EvaluateRHS keeps with useRoeJacobian
 * (lambdaFace0/123/4 = the Riemann solver's entropy-fixed eigenvalues integrated over the face and
 * divided by its area, src/Euler/EulerEvaluator_EvaluateRHS.hxx:527-544), lamRoe [3][nFace].
 * dndsr_lusgs_init_block: LUSGSMatrixInit with useRoeJacobian (src/Euler/EulerEvaluator.hxx:37-159;
 * EulerEvaluator.hpp:427-434): per face (alphaDiag/2) A_f / V (|A|_Roe + lambdaVis I), boundary
 * faces with the ghost state and its linearisation, 1/dTau + 1/dt on the diagonal; jdiag (may be
 * NULL) and jdiagInv are [25][nCell] (row-major 5x5 per cell), the inverse as
 * JacobianDiagBlock::GetInvert (src/Euler/EulerJacobian.hpp:153-190).  lamRoe NULL takes
 * EvaluateDt's eigenvalue (lambdaFace row 3) for all three.  Synchronises; a singular block is an
 * error.
 * dndsr_update_sgs_block(_f32): UpdateSGS with useRoeJacobian (EulerEvaluator.hxx:550-693), the
 * neighbour term A(u_j) dU_j - |A|_Roe dU_j - lambdaVis dU_j, x_i = JDiag_i^-1 buf, in multicolour
 * order as dndsr_update_sgs. */
int dndsr_eval_rhs_roe_lambda(dndsr_ctx *ctx, const dndsr_euler_settings *es, const double *u, const double *uRec,
                              const double *deltaLambdaCell, double *lamRoe);
int dndsr_lusgs_init_block(dndsr_ctx *ctx, const dndsr_euler_settings *es, const double *dTau, double dt,
                           double alphaDiag, const double *u, const double *lambdaFace, const double *lamRoe,
                           double *jdiag, double *jdiagInv);
int dndsr_update_sgs_block(dndsr_ctx *ctx, const dndsr_euler_settings *es, double alphaDiag, const double *rhs,
                           const double *u, const double *uInc, double *uIncNew, const double *jdiagInv,
                           const double *lambdaFace, const double *lamRoe, int forward, int gsUpdate, int uIncIsZero,
                           double *sumIncHost);
/* A sweep of the transpose of the block matrix dndsr_update_sgs_block sweeps, in the same
 * multicolour order: y_i = JDiag_i^-T (rhs_i - sum_j O_ji^T y_j) with
 * O_ji = (alphaDiag/2) A_f / V_j (A(u_i) n_ji - |A|_Roe - lambdaVis I) and the face eigenvalues
 * seen from j.  No counterpart in the reference: the block preconditioner of the discrete-adjoint
 * solve (dndsr_amd.adjoint.SteadyAdjoint). */
int dndsr_update_sgs_block_transpose(dndsr_ctx *ctx, const dndsr_euler_settings *es, double alphaDiag,
                                     const double *rhs, const double *u, const double *uInc, double *uIncNew,
                                     const double *jdiagInv, const double *lambdaFace, const double *lamRoe,
                                     int forward, int gsUpdate, int uIncIsZero, double *sumIncHost);
/* nSweep symmetric (forward then backward) sweeps of the transposed LU-SGS matrix from zero
 * increments into out: scalar with jdiag (dndsr_update_sgs_transpose), block with jdiagInv and
 * lamRoe (dndsr_update_sgs_block_transpose); pass exactly one of jdiag and jdiagInv.  work: a
 * second [5][N] array.  With gsUpdate a sweep after the first skips the colour pass that would
 * repeat the previous sweep's last one: the result equals that of the single sweeps.  No
 * counterpart in the reference: the preconditioner of the discrete-adjoint solve in one call,
 * without a host synchronisation per sweep. */
int dndsr_sgs_transpose_symmetric(dndsr_ctx *ctx, const dndsr_euler_settings *es, double alphaDiag,
                                  const double *rhs, const double *u, const double *jdiag, const double *jdiagInv,
                                  const double *lambdaFace, const double *lamRoe, int nSweep, int gsUpdate,
                                  double *work, double *out);
int dndsr_update_sgs_block_f32(dndsr_ctx *ctx, const dndsr_euler_settings *es, float alphaDiag, const float *rhs,
                               const float *u, const float *uInc, float *uIncNew, const float *jdiagInv,
                               const float *lambdaFace, const float *lamRoe, int forward, int gsUpdate,
                               int uIncIsZero, double *sumIncHost);
/* Explicit SSP-RK3 time loop kept on the device (ExplicitSSPRK3TimeStepAsImplicitDualTimeStep::Step,
 * src/Solver/ODE.hpp:1477-1520): per step EvaluateDt once (global dt = CFL-limited minimum, or
 * dtFixed > 0), then three stages of nRec Jacobi VR iterations + EvaluateRHS + combination.
 * u and uRec are updated in place. */
int dndsr_ssprk3(dndsr_ctx *ctx, const dndsr_euler_settings *es, double *u, double *uRec, int nStep, double CFL,
                 int nRec, double dtFixed);

/* Settings of the implicit dual-time step (EulerSolver config: timeMarchControl,
 * implicitCFLControl, convergenceControl, linearSolverControl; src/Euler/EulerSolver.hpp:171-760). */
typedef struct
{
    double dt;             /* physical time step (dtImplicit); 1e100 marches to a steady state */
    double CFL;            /* pseudo-time CFL of the local dTau (implicitCFLControl.CFL) */
    int nRec;              /* Jacobi VR sweeps per residual evaluation (nInternalRecStep) */
    int sgsIter;           /* symmetric sweeps after the first one (linearSolverControl.sgsIter) */
    int gsUpdate;          /* jacobiCode: 1 Gauss-Seidel, 0 Jacobi */
    int maxIter;           /* nTimeStepInternal */
    int minIter;           /* nTimeStepInternalMin */
    double tol;            /* rhsThresholdInternal on the relative density residual */
    int fp32Sweeps;        /* 1: LU-SGS sweeps in FP32 (residual, VR and update stay FP64) */
    double incrementRelax; /* incrementPPRelax */
    double refRho, refP;   /* reference density / pressure of the positivity epsilons (refUPrim) */
    int gmresCode;         /* linearSolverControl.gmresCode: 0 LU-SGS, 1 GMRES, 2 LU-SGS then GMRES */
    int nGmresSpace;       /* Krylov subspace size (nGmresSpace) */
    int nGmresIter;        /* restart cycles (nGmresIter) */
} dndsr_implicit_settings;

/* One implicit-Euler dual-time step, entirely on the device
 * (ImplicitEulerDualTimeStep::Step, src/Solver/ODE.hpp:80-98, with the EulerSolver
 * callbacks fdtau / frhs / fsolve / fincrement / fstop, src/Euler/EulerSolver.hxx:645-830,
 * 1206; solveLinear / doPrecondition :1407-1658).  Per inner iteration: local dTau from
 * EvaluateDt, nRec Jacobi VR sweeps, EvaluateRHS, the unsteady residual and its L1 norm,
 * LUSGSMatrixInit, (1 + sgsIter) symmetric multicolour sweeps (FP64 or FP32) and the
 * increment (with es->useRoeJacobian: the face eigenvalues, block init and block sweeps).  Stops when the density residual relative to its running maximum is below
 * tol (after minIter iterations).  u and uRec are updated in place.  Outputs (host, may
 * be NULL): nIter, res[5] (L1 residual of the last iteration), resRel[5], history[maxIter]
 * (relative density residual per iteration).  The increment is applied as the reference's
 * fincrement (EulerSolver.hxx:696-738): per cell alpha = EvaluateCellRHSAlpha (relaxation
 * incrementRelax, compress 2; src/Euler/EulerEvaluator.hxx:1938-2030), then AddFixedIncrement
 * u += CompressInc(u, alpha inc) (src/Euler/EulerEvaluator.hpp:1645-1830); the counts are read
 * with dndsr_implicit_pp_stats.  Returns an error, with u and uRec restored to their values at
 * entry, when the residual is not finite or a cell meets a state the reference asserts against
 * (non-positive density or internal energy).  gsUpdate must be 0 or 1 (jacobiCode 2, ILU, is
 * not provided). */
int dndsr_implicit_euler_step(dndsr_ctx *ctx, const dndsr_euler_settings *es, const dndsr_implicit_settings *is,
                              double *u, double *uRec, int *nIter, double *res, double *resRel, double *history);

/* One BDF dual-time step (ImplicitBDFDualTimeStep::Step, src/Solver/ODE.hpp:427-477; coefficients
 * :590-595): as dndsr_implicit_euler_step, with the unsteady residual of the (nPrev + 1)-step
 * formula, sum_j c_j x_j / dt - x / dt + alpha frhs(x) over the entry state and the nPrev (0..3)
 * earlier states xPrevs[j] (device arrays [5][nCell], newest first), dTau scaled by 1 / alpha
 * (EulerSolver.hxx:693) and alphaDiag = alpha in LUSGSMatrixInit and the sweeps.  The caller
 * keeps the history between steps (the reference's xPrevs ring) and starts with nPrev = 0 (BDF1,
 * the implicit-Euler step), as the reference does. */
int dndsr_implicit_bdf_step(dndsr_ctx *ctx, const dndsr_euler_settings *es, const dndsr_implicit_settings *is,
                            int nPrev, const double *const *xPrevs, double *u, double *uRec, int *nIter, double *res,
                            double *resRel, double *history);

/* Positivity statistics of the last dndsr_implicit_euler_step: the number of cell increments
 * the positivity limiter scaled (summed over its iterations; the PPIncrementLimiter count the
 * reference's solver logs, EulerSolver.hxx:708-718) and the smallest alpha (1 when none). */
int dndsr_implicit_pp_stats(dndsr_ctx *ctx, int64_t *nLimited, double *alphaMin);

/* LUSGSMatrixVec (scalar Jacobian), src/Euler/EulerEvaluator.hpp:449-455:
 *   Ax = D x + sum_f (alphaDiag/2) A_f / V fluxJacobian0(u_j) x_j, the matrix UpdateSGS sweeps. */
int dndsr_lusgs_matvec(dndsr_ctx *ctx, const dndsr_euler_settings *es, double alphaDiag, const double *u,
                       const double *x, const double *jdiag, const double *lambdaFace, double *Ax);
/* GMRES_LeftPreconditioned<...>::solve (src/Solver/Linear.hpp:47-131) on the LU-SGS matrix with
 * the LU-SGS preconditioner ((1 + sgsIter) symmetric sweeps from zero, FP32 when fp32Precond):
 * x holds the initial guess and receives the solution; nGmresSpace vectors, nGmresIter restart
 * cycles.  residual (may be NULL) receives the last preconditioned residual norm and |M b|. */
int dndsr_gmres_lusgs(dndsr_ctx *ctx, const dndsr_euler_settings *es, double alphaDiag, const double *u,
                      const double *jdiag, const double *lambdaFace, const double *b, double *x, int nGmresSpace,
                      int nGmresIter, int sgsIter, int gsUpdate, int fp32Precond, double *residual);

/* ---- discrete adjoint (no counterpart in the reference; checked against central finite
 * differences of the forward operators) ----------------------------------------------
 * Vector-Jacobian product of EvaluateRHS (dndsr_eval_rhs) with deltaLambdaCell held fixed:
 *   uBar = (d rhs / d u)^T rhsBar, uRecBar = (d rhs / d uRec)^T rhsBar (overwritten; uRecBar may
 *   be NULL, or must be when uRec is NULL).  Riemann solvers: all of dndsr_eval_rhs's;
 *   maxOrder 1 and 2; inviscid or viscous (es->viscous); periodic, interior and boundary
 *   faces (the ghost states differentiated as evaluated). */
int dndsr_eval_rhs_vjp(dndsr_ctx *ctx, const dndsr_euler_settings *es, const double *u, const double *uRec,
                       const double *deltaLambdaCell, const double *rhsBar, double *uBar, double *uRecBar);
/* Force on the boundary faces of one zone (drag / lift of a wall): EvaluateRHS at (u, uRec),
 * then force[d] = sum over the zone's faces of the momentum flux out of the domain
 * (Riemann flux minus viscous flux, integrated over the face).  Synchronises; force is a
 * host array of 3.  No counterpart in the reference's evaluator (its solver integrates
 * wall forces in its output stage); the wall-force functional of an adjoint gradient. */
int dndsr_zone_force(dndsr_ctx *ctx, const dndsr_euler_settings *es, const double *u, const double *uRec,
                     const double *deltaLambdaCell, int zone, double *force);
/* uBar = (d force / d u)^T forceBar, uRecBar = (d force / d uRec)^T forceBar (forceBar a host
 * array of 3; deltaLambdaCell held fixed; overwritten outputs; uRecBar may be NULL). */
int dndsr_zone_force_vjp(dndsr_ctx *ctx, const dndsr_euler_settings *es, const double *u, const double *uRec,
                         const double *deltaLambdaCell, int zone, const double *forceBar, double *uBar,
                         double *uRecBar);
/* Transpose of one VR sweep (dndsr_vr_iter modes 0-2, mode 3 = dndsr_vr_iter_diff):
 *   uRecBar = (d uRecNew / d uRec)^T uRecNewBar, uBar = (d uRecNew / d u)^T uRecNewBar
 * (uBar unused for mode 3; uRecBar may be NULL).  relax < 0 uses settings.jacobiRelax.
 * uRec, u: the state the forward sweep read; on a mesh with boundary faces the boundary
 * term depends on it (Euler ghosts: both; given point values: u), elsewhere they may be
 * NULL.  Mode 3 covers meshes without boundary faces.  With the second-order substitute
 * (subs2ndOrder 1 or 2) every mode writes a zero uRecBar (mode 3 writes nothing else) and
 * the substitute's transpose into uBar; a bounded mesh then needs nVars 5, the boundary
 * table and u. */
int dndsr_vr_iter_vjp(dndsr_ctx *ctx, int nVars, int mode, double relax, const double *uRec, const double *u,
                      const double *uRecNewBar, double *uRecBar, double *uBar);

/* ---- instrumentation (no counterpart in the reference, whose timers are the host-side
 * PerformanceTimer slots of src/DNDS/Profiling.hpp:22-49) ------------------------------
 * When enabled, every kernel launch is bracketed by a HIP event pair on its own stream.
 * dndsr_get_timing synchronises, adds the elapsed times of all launches recorded since the
 * last call per kernel id, and returns the running totals (ms) and launch counts:
 *   0 VR iteration, 1 RHS face pass, 2 RHS cell gather, 3 EvaluateDt face pass,
 *   4 EvaluateDt cell pass, 5 RK stage combination, 6 SGS colour sweep, 7 LU-SGS init,
 *   8 dual-time residual + norm, 9 increment, 10 FP64 -> FP32 conversion, 11 adjoint kernels,
 *   12 linear-solver kernels (LUSGSMatrixVec, GMRES vector operations), 13 RHS viscous face
 *   pass.  The face pass (1) covers the inviscid interior, periodic and boundary faces only.
 * reset != 0 clears the totals after reading them. */
#define DNDSR_NUM_KERNEL_IDS 14
/* Device bandwidth self-test (no counterpart in the reference; it calibrates the roofline the
 * kernels are quoted against): mode 0 streams n doubles of src once, reading 16 bytes per lane
 * (dst receives one partial sum per workgroup and must hold 4096 doubles); mode 1 copies src
 * to dst.  Both 16-byte aligned, n even.  Returns the average milliseconds of one pass over
 * reps passes, timed with HIP events on the context's stream (synchronises). */
int dndsr_stream_probe(dndsr_ctx *ctx, int mode, const double *src, double *dst, int64_t n, int reps,
                       double *msPerPass);
int dndsr_set_timing(dndsr_ctx *ctx, int enable);
int dndsr_get_timing(dndsr_ctx *ctx, double *msTotal, int64_t *launches, int reset);

#ifdef __cplusplus
}
#endif
#endif
