/*
 * fxg — MI355X (gfx950) DGSEM right-hand side and low-storage Runge-Kutta stepper for the
 * compressible Navier-Stokes equations, with the call surface of FLEXI's hot path.
 *
 * Plain C ABI: no C++ or torch types, plain pointers and sizes. Every array crossing this
 * boundary uses FLEXI's own Fortran layout (column-major, variable index fastest, ids 1-based),
 * so a FLEXI host passes its module arrays unchanged through ISO_C_BINDING.
 *
 * Error convention: every function returns 0 on success and a nonzero code otherwise; the
 * message is available from fxg_last_error(). FLEXI aborts through Abort/CollectiveStop
 * (src/globals/globals.f90:129-220); the Fortran shim maps a nonzero code onto that call.
 *
 * Threading: a context is bound to the HIP device that was current at creation and is not
 * thread-safe; FLEXI calls the path from a single host thread per rank.
 */
#ifndef FXG_H
#define FXG_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Node types as FLEXI's PP_NodeType (src/CMakeLists.txt): 1 Gauss, 2 Gauss-Lobatto. */
#define FXG_NODE_GAUSS 1
#define FXG_NODE_GAUSS_LOBATTO 2
/* Riemann solvers, numbered as src/equations/navierstokes/riemann.f90:44-57. */
#define FXG_RIEMANN_LF 1
#define FXG_RIEMANN_ROE 3
#define FXG_RIEMANN_ROEENTROPYFIX 33
/* Split-form fluxes, numbered as src/equations/navierstokes/splitflux.f90:63-69. */
#define FXG_SPLITDG_NONE 0
#define FXG_SPLITDG_PI 4
/* Arithmetic of the flux, lifting and integral kernels. U, the RK register, the metric terms
 * and the Jacobian stay FP64 in both modes. */
#define FXG_PREC_FP64 0
#define FXG_PREC_FP32_BULK 1
/* Metric terms of fxg_create_from_nodes. FXG_METRICS_OWN: the curl form with coordinates relative
 * to each element's first node, and every element takes its own boundary metric as the face
 * metric of its sides: the freestream is kept to flux round-off on curved meshes. Use it unless a
 * run must follow FLEXI's. FXG_METRICS_REFERENCE: CalcMetrics/CalcSurfMetrics in
 * metrics.f90's operation order (absolute coordinates, one-dimensional basis changes in
 * sequence, no FMA contraction), with the master's NormVec/SurfElem shared by both sides as
 * in FLEXI: the metric terms FLEXI computes, to the rounding of the basis matrices. fxg_create
 * (metrics from the host) always shares the side's NormVec/SurfElem. */
#define FXG_METRICS_OWN 0
#define FXG_METRICS_REFERENCE 1

/* Run-time parameters. Names and meanings follow FLEXI's parameter file keys:
 * N, NodeType (compile-time PP_NodeType in FLEXI), SplitDG (src/equations/navierstokes/
 * splitflux.f90:95-101), Riemann (riemann.f90:97-113), kappa/R/Pr/mu0 (idealgas/eos.f90:79-90),
 * doWeakLifting/doConservativeLifting (src/dg/lifting/lifting.f90:72-76). */
typedef struct {
  int N;
  int nodeType;
  int splitDG;               /* FXG_SPLITDG_NONE (weak form) or FXG_SPLITDG_PI */
  int riemann;               /* FXG_RIEMANN_* */
  double kappa, R, Pr, mu0;  /* constant viscosity (PP_VISC=0) */
  int doWeakLifting;         /* only 0 (strong BR1) is supported */
  int doConservativeLifting; /* only 0 (non-conservative volume integral) is supported */
  int precision;             /* FXG_PREC_* */
  int metrics;               /* FXG_METRICS_*: how fxg_create_from_nodes forms the metric terms */
} fxg_params;

/* Mesh and metric arrays as FLEXI holds them after InitMesh (src/mesh/mesh_vars.f90:26-209):
 *   SideToElem(5,nSides)  ElemToSide(2,6,nElems)              (1-based element and side ids)
 *   Metrics_fTilde/gTilde/hTilde(3,0:N,0:N,0:N,nElems)  sJ(0:N,0:N,0:N,nElems)
 *   NormVec/TangVec1/TangVec2(3,0:N,0:N,nSides)  SurfElem(0:N,0:N,nSides)
 * Sides are boundary sides 1..nBCSides (SideToElem(2,:) = -1, as setLocalSideIDs numbers them,
 * src/mesh/prepare_mesh.f90:196-222) followed by conforming inner sides: firstInnerSide =
 * nBCSides+1, lastInnerSide = nSides; no mortars, a single rank. A mesh with boundary sides needs
 * fxg_set_boundaries before its first right-hand side. The library copies the arrays; the caller
 * keeps ownership. */
typedef struct {
  int nElems;
  int nSides;
  int firstInnerSide;
  int lastInnerSide;
  const int* SideToElem;
  const int* ElemToSide;
  const double* Metrics_fTilde;
  const double* Metrics_gTilde;
  const double* Metrics_hTilde;
  const double* sJ;
  const double* NormVec;
  const double* TangVec1;
  const double* TangVec2;
  const double* SurfElem;
} fxg_mesh;

typedef struct fxg_ctx fxg_ctx;

/* Message of the last failed call on this thread. */
const char* fxg_last_error(void);
/* Library version string. */
const char* fxg_version(void);

/* Set-up: replaces InitDG/InitLifting/InitTimeDisc allocations (src/dg/dg.f90:45-111,
 * src/dg/lifting/lifting.f90:106-185, src/timedisc/timedisc_func.f90:68-128) with device
 * allocations. The DG basis (D, D_Hat, DVolSurf, L_Hat) is built from N and nodeType as
 * InitDGBasis does (src/dg/dg.f90:132-209). */
int fxg_create(const fxg_params* params, const fxg_mesh* mesh, fxg_ctx** ctx);
/* Set-up from geometry: NodeCoords(3,0:NGeo,0:NGeo,0:NGeo,nElems) on equidistant nodes; the
 * metrics of CalcMetrics/CalcSurfMetrics (src/mesh/metrics.f90:161-787, curl form) are computed
 * on the GPU as params->metrics selects. SideToElem/ElemToSide as in fxg_mesh, boundary sides
 * first. */
int fxg_create_from_nodes(const fxg_params* params, int nElems, int nSides,
                          const int* SideToElem, const int* ElemToSide, int NGeo,
                          const double* NodeCoords, fxg_ctx** ctx);
int fxg_destroy(fxg_ctx* ctx);
/* Boundary conditions of the boundary sides (InitBC, src/equations/navierstokes/idealgas/
 * getboundaryflux.f90:64-218, with FLEXI's arrays): BC(nBCSides) the boundary index of each side
 * (mesh_vars.f90), BoundaryType(nBCs,3) = (type, state, alpha) in Fortran layout, RefStatePrim(6,
 * nRefState) = (rho,u,v,w,p,T) per RefState (equation.f90:126-141) as InitBC leaves it (for type 27
 * the entries 2..4 hold the inflow direction vector). Types (getboundaryflux.f90:224-472,
 * 480-789, 855-939): 2 Dirichlet with state = RefState index (>= 1); 12 and 121 Dirichlet with a
 * per-point state given by fxg_set_boundary_data; 23 outflow at the Mach number RefStatePrim(2,
 * state); 24 pressure outflow and 25 subsonic outflow at RefStatePrim(5,state); 27 subsonic
 * inflow (total temperature, direction, total pressure of RefState `state`); 3 adiabatic wall; 4
 * isothermal wall (temperature of RefState `state`); 9 slip wall; 91 slip wall with the wall's
 * velocity-gradient treatment. Type 22 (an exact function of time) is not evaluated by the
 * library. Both stage pipelines support them; the discrete adjoint does not. Call it before the
 * first right-hand side, with no work of the context in flight. No-op for a mesh without
 * boundary sides. */
int fxg_set_boundaries(fxg_ctx* ctx, int nBCs, const int* BC, const int* BoundaryType, int nRefState,
                       const double* RefStatePrim);
/* The boundary states of the sides of types 12 and 121, BCDataPrim(6,0:N,0:N,nBCSides) =
 * (rho,u,v,w,p,T) per boundary side point in FLEXI's layout (equation_vars.f90; filled there by
 * ReadBCFlow from a State file for 12, by ExactFunc at the start time for 121, getboundaryflux.f90:
 * 64-218, 946-1040). Replaces those two fills; entries of sides of other types are not read but
 * must be physical. Call it after fxg_set_boundaries and before a right-hand side, with no work of
 * the context in flight. */
int fxg_set_boundary_data(fxg_ctx* ctx, const double* BCDataPrim);

/* Host <-> device state transfer in FLEXI layout U(5,0:N,0:N,0:N,nElems). */
int fxg_set_state(fxg_ctx* ctx, const double* U);
int fxg_get_state(fxg_ctx* ctx, double* U);
int fxg_get_ut(fxg_ctx* ctx, double* Ut);
/* Lifted gradients of the last right-hand side, gradU{x,y,z}(nVarLifting,0:N,0:N,0:N,nElems)
 * (src/dg/lifting/lifting_vars.f90) in the layout of FLEXI's build: nVarLifting = 4 for
 * PP_OPTLIFT=1, lifting variables (u,v,w,T); 5 for FLEXI's default PP_OPTLIFT=0, (rho,u,v,w,T)
 * (src/equations/navierstokes/idealgas/eos.h:108-127). After fxg_dg_time_derivative those of U;
 * after fxg_timestep_lserkw2, as in FLEXI, those of the last stage's right-hand side (at the state
 * before the last update, recomputed there on request, equal to round-off). The density
 * gradient, which the solver itself does not need, is lifted on request from the same state.
 * fxg_get_ut likewise. */
int fxg_get_gradients(fxg_ctx* ctx, int nVarLifting, double* gx, double* gy, double* gz);
/* Taylor-Green vortex analysis, AnalyzeTestcase (src/testcase/taylorgreenvortex/testcase.f90:241-479)
 * evaluated on the device: the state and the lifted velocity gradients that fxg_get_gradients
 * would return are interpolated to NAnalyze+1 Gauss-Lobatto points per direction
 * (1 <= NAnalyze <= 31) and integrated; out[15] receives the values of one TGVAnalysis line in
 * FLEXI's order (DR_S, DR_Sd+DR_p, Ekin, Ekin comp, Enstrophy, DR_u, DR_S, DR_Sd, DR_p, maximum
 * vorticity, mean temperature, uprime, mean entropy, ED_S, ED_D), constant viscosity, rho0 = 1.
 * Replaces the body of AnalyzeTestcase on a GPU host (no download of U and the gradients). */
int fxg_tgv_analyze(fxg_ctx* ctx, int NAnalyze, double* out);

/* Geometry as used on the device (for checking fxg_create_from_nodes): same layouts as fxg_mesh
 * plus Elem_xGP(3,0:N,0:N,0:N,nElems). Any pointer may be NULL. */
int fxg_get_metrics(fxg_ctx* ctx, double* Mf, double* Mg, double* Mh, double* sJ,
                    double* NormVec, double* TangVec1, double* TangVec2, double* SurfElem,
                    double* Elem_xGP);

/* DGTimeDerivative_weakForm(t) (src/dg/dg.f90:221-678): Ut = RHS(U) on the device state. */
int fxg_dg_time_derivative(fxg_ctx* ctx, double t);
/* TimeStepByLSERKW2(t) (src/timedisc/timestep.f90:82-152) for nStages stages with FLEXI's
 * coefficient arrays RKA(2:nStages), RKb(1:nStages), RKc(2:nStages) (timedisc_vars.f90:109-214),
 * passed as RKA[0..nStages-2], RKb[0..nStages-1], RKc[0..nStages-2]. U is advanced in place. */
int fxg_timestep_lserkw2(fxg_ctx* ctx, double t, double dt, int nStages, const double* RKA,
                         const double* RKb, const double* RKc);
/* CALCTIMESTEP(errType) (src/equations/navierstokes/calctimestep.f90:89-227). CFLScale and
 * DFLScale are the values after FillCFL_DFL (src/timedisc/timedisc_func.f90:463-527).
 * errType: 0 ok, 1 NaN density, 2 NaN convective dt, 3 NaN viscous dt. dtElem may be NULL. */
int fxg_calc_timestep(fxg_ctx* ctx, double CFLScale, double DFLScale, double* dt, int* errType,
                      double* dtElem);

/* Device-pointer forms for hosts that keep the state in their own device buffers (e.g. torch
 * tensors): the arrays have the FLEXI layouts above and live on the context's device; `stream`
 * is a hipStream_t (NULL = the context's stream). The call is asynchronous on that stream. It
 * uses the context's scratch arrays (face states, face fluxes, gradients), so calls on one
 * context must be serialised: with a stream other than the context's, the caller orders the
 * calls itself (e.g. synchronises before switching streams). Afterwards fxg_get_ut and
 * fxg_get_gradients recompute those of the context's own state. */
int fxg_dg_time_derivative_dev(fxg_ctx* ctx, double t, const double* U_dev, double* Ut_dev,
                               void* stream);
int fxg_timestep_lserkw2_dev(fxg_ctx* ctx, double t, double dt, int nStages, const double* RKA,
                             const double* RKb, const double* RKc, double* U_dev, double* R_dev,
                             void* stream);
/* Discrete adjoint (FP64, Gauss-Lobatto split form, N 2..7, meshes without boundary sides; no
 * counterpart in the reference, which has no adjoint). All arrays in FLEXI layout on the
 * context's device.
 * fxg_rhs_vjp_dev: out = (d Ut / d U)^T lam, the transposed Jacobian of
 *   DGTimeDerivative_weakForm at U applied to lam (asynchronous on `stream`, NULL = the context's).
 * fxg_forward_n_steps: nSteps calls of TimeStepByLSERKW2 on U_dev with the time steps dt[0..nSteps-1]
 *   (host array; the sequence a primal run chose with CALCTIMESTEP, frozen so that the adjoint
 *   differentiates the same map). With ckpt_dev (may be NULL) the state before step k is copied to
 *   ckpt_dev[k/checkpointEvery] for every k that checkpointEvery divides (ceil(nSteps/every) states).
 *   With checkpointEvery = 0 the forward call instead stores the whole trajectory in ckpt_dev
 *   (every stage's input state and its face states; layout private to the library, size from
 *   fxg_adjoint_storage): the stages write it out of place, so the forward costs what it costs
 *   without it, and the adjoint then needs no recomputation.
 * fxg_adjoint_n_steps: ubar_dev holds dJ/dU(T) on entry and dJ/dU(0) on return, for the map
 *   U(0) -> U(T) of fxg_forward_n_steps started from U0_dev. The stage states are recomputed
 *   segment by segment from the checkpoints: those of a forward call with the same
 *   checkpointEvery (ckpt_dev), or, with ckpt_dev NULL, from a forward sweep of its own. Device
 *   memory allocated for the call: checkpointEvery*nStages + 3 states (+ the checkpoints when it
 *   takes them itself). With checkpointEvery = 0, ckpt_dev must be the trajectory of a forward
 *   call with checkpointEvery = 0 and the same steps; only the reverse stages run (one state of
 *   scratch). Synchronous.
 * fxg_adjoint_storage: the number of doubles ckpt_dev needs for nSteps steps and checkpointEvery
 *   (ceil(nSteps/checkpointEvery) states, or the trajectory for 0). */
int fxg_rhs_vjp_dev(fxg_ctx* ctx, double t, const double* U_dev, const double* lam_dev, double* out_dev,
                    void* stream);
int fxg_forward_n_steps(fxg_ctx* ctx, int nSteps, const double* dt, int nStages, const double* RKA,
                        const double* RKb, const double* RKc, double* U_dev, double* ckpt_dev,
                        int checkpointEvery, void* stream);
int fxg_adjoint_n_steps(fxg_ctx* ctx, int nSteps, const double* dt, int nStages, const double* RKA,
                        const double* RKb, const double* RKc, const double* U0_dev, const double* ckpt_dev,
                        double* ubar_dev, int checkpointEvery, void* stream);
int fxg_adjoint_storage(fxg_ctx* ctx, int nSteps, int nStages, int checkpointEvery, size_t* doubles);
/* CALCTIMESTEP of a device state without any allocation: no counterpart in the reference; for
 * callers that cannot allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc). The same arithmetic as fxg_calc_timestep on U_dev (FLEXI
 * layout, the context's device, which must be current): *dt_dev = MIN(convective, viscous) time
 * step and *errType_dev = errType (0 ok, 1 NaN density, 2 NaN convective, 3 NaN viscous dt; the
 * call itself does not fail on a NaN), both in device memory the caller owns. Asynchronous on
 * `stream` (NULL = the context's). Its per-element scratch is a fixed device array of the library
 * (about 10 MB per GPU, held until the process ends) for meshes of at most 524,288 elements;
 * larger meshes are rejected. All calls on one GPU share that array: calls that overlap on one
 * GPU (other streams, other threads, other contexts) race and must be serialised by the caller. */
int fxg_calc_timestep_dev_noalloc(fxg_ctx* ctx, const double* U_dev, double CFLScale, double DFLScale,
                                  double* dt_dev, int* errType_dev, void* stream);
/* Which stage kernels the context runs: "fused" (Gauss-Lobatto split form, N 2..7: two
 * kernels per stage, face states carried between LSERK stages) or "classic" (prolongation,
 * lifting, side fluxes, integrals). No counterpart in the reference; FXG_PIPELINE=classic in
 * the environment at fxg_create time selects the classic path everywhere. */
const char* fxg_pipeline(fxg_ctx* ctx);
/* Synchronise the context's stream. */
int fxg_sync(fxg_ctx* ctx);

/* Sizes of the context (for hosts that allocate the device-pointer arrays themselves). */
int fxg_sizes(fxg_ctx* ctx, int* N, int* nElems, int* nSides);

/* Kernel timing with HIP events (no counterpart in the reference, whose timers are wall-clock
 * FLEXITIME calls around the whole time loop, src/globals/globals.f90:660). While enabled, an
 * event pair brackets every kernel launch on the launch stream; fxg_profile_read synchronises
 * and returns the summed duration and launch count of one kernel since enabling.
 * Kernel ids: 0 prolongation, 1 lifting, 2 side fluxes, 3 volume/surface integral (+RK update),
 * 4 per-element time step, 5 time-step reduction; on the fused Gauss-Lobatto split-form path
 * 6 face states of the first stage, 7 viscous face fluxes, 8 fused stage. Enabling again resets
 * the counters. */
int fxg_profile_enable(fxg_ctx* ctx, int enable);
int fxg_profile_read(fxg_ctx* ctx, int kernel, double* total_ms, int* launches);

#ifdef __cplusplus
}
#endif
#endif /* FXG_H */
