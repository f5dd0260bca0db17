/*
 * CPU oracle for the FLEXI DGSEM Navier-Stokes right-hand side and the LSERK stepper.
 *
 * TEST INFRASTRUCTURE ONLY. This library restates, in plain C++ and FP64, the reference
 * algorithm of FLEXI (flexi-framework/flexi v25.10) for the hot path, in FLEXI's own
 * column-major data layout. Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline
 * leg may load it, and only as the checker or as the timed CPU baseline. The product
 * (libfxg.so) never links or calls it.
 *
 * Every routine cites the reference file:line it follows (paths relative to the FLEXI root).
 * Parity status of each part is documented in oracle/README.md.
 */
#pragma once
#ifdef __cplusplus
extern "C" {
#endif

/* Node types, numbered as FLEXI's PP_NodeType (src/CMakeLists.txt: GAUSS=1, GAUSS-LOBATTO=2)
 * plus the helper node sets used by the metric and analysis code. */
enum { OR_NODE_G = 1, OR_NODE_GL = 2, OR_NODE_CL = 3, OR_NODE_VISU = 4 };

/* Riemann solver ids as in riemann.f90:44-57. */
enum { OR_RIEMANN_LF = 1, OR_RIEMANN_ROE = 3, OR_RIEMANN_ROEENTROPYFIX = 33 };

typedef struct {
  int N;            /* polynomial degree */
  int nodeType;     /* OR_NODE_G or OR_NODE_GL */
  int splitDG;      /* 0: weak form (VolInt_weakForm), 1: split form with the PI flux */
  int riemann;      /* OR_RIEMANN_* */
  double kappa, R, Pr, mu0;
} or_params;

/* ---- 1D basis (interpolation/basis.f90, interpolation.f90) ---- */
void or_nodes_weights(int N, int nodeType, double* x, double* w, double* wBary);
void or_derivative_matrix(int N, int nodeType, double* D);         /* D(0:N,0:N) column-major */
void or_lagrange_polys(double x, int N, const double* xGP, const double* wBary, double* L);
void or_vandermonde(int Nin, int typeIn, int Nout, int typeOut, int modal, double* Vdm);
/* InitDGBasis + InitInterpolationBasis (dg.f90:132-209, interpolation.f90:127-160):
 * out: xGP,wGP,wBary,L_Minus,L_Plus,L_HatMinus,L_HatPlus (N+1 each) and
 * D,D_T,D_Hat,D_Hat_T,DVolSurf ((N+1)^2 each, column-major). */
void or_dg_basis(int N, int nodeType, double* xGP, double* wGP, double* wBary, double* L_Minus,
                 double* L_Plus, double* L_HatMinus, double* L_HatPlus, double* D, double* D_T,
                 double* D_Hat, double* D_Hat_T, double* DVolSurf);

/* ---- index mappings (mesh/mappings.f90) ---- */
void or_mappings(int N, int* S2V2, int* S2V, int* V2S);

/* ---- metrics (mesh/metrics.f90) ---- */
void or_calc_metrics(int N, int nodeType, int NGeo, int nElems, const double* NodeCoords,
                     int nSides, const int* ElemToSide, double* Metrics_f, double* Metrics_g,
                     double* Metrics_h, double* sJ, double* Elem_xGP, double* NormVec,
                     double* TangVec1, double* TangVec2, double* SurfElem, double* Face_xGP);

/* ---- point physics ---- */
void or_cons_to_prim(const or_params* p, const double* cons, double* prim);
void or_prim_to_cons(const or_params* p, const double* prim, double* cons);
void or_riemann_point(const or_params* p, const double* U_L, const double* U_R, const double* UPrim_L,
                      const double* UPrim_R, const double* nv, const double* t1, const double* t2,
                      double* F);
void or_split_volume_flux_pi(const double* URef, const double* UPrimRef, const double* U,
                             const double* UPrim, const double* MRef, const double* M, double* F);
void or_eval_diff_flux_point(const or_params* p, const double* UPrim, const double* gx,
                             const double* gy, const double* gz, double* f, double* g, double* h);

/* ---- the DG operator ---- */
typedef struct or_ctx or_ctx;
/* Mesh arrays in FLEXI layout. SideToElem(5,nSides), ElemToSide(2,6,nElems) hold 1-based ids.
 * Metrics_*(3,0:N,0:N,0:N,nElems), sJ(0:N,0:N,0:N,nElems), NormVec/TangVec*(3,0:N,0:N,nSides),
 * SurfElem(0:N,0:N,nSides). Sides 1..nBCSides of or_set_boundaries are boundary sides (no
 * neighbour, SideToElem(2,:) = -1); all others are inner sides. */
or_ctx* or_init(const or_params* p, int nElems, int nSides, const int* SideToElem,
                const int* ElemToSide, const double* Mf, const double* Mg, const double* Mh,
                const double* sJ, const double* NormVec, const double* TangVec1,
                const double* TangVec2, const double* SurfElem);
void or_free(or_ctx* c);
/* Boundary sides 1..nBCSides (numbered first, as setLocalSideIDs does, prepare_mesh.f90:196-222):
 * BC(nBCSides) the boundary index of each, BoundaryType(nBCs,3) = (type, state, alpha) per
 * boundary (mesh_readin.f90:167-173), RefStatePrim(6,nRefState) (equation.f90:126-141) as InitBC
 * leaves it. Types 2, 4, 23, 24, 25, 27 (BCState >= 1), 3, 9, 91, 12 and 121
 * (getboundaryflux.f90). Returns 0, or a nonzero code for an invalid side, index, type or state. */
int or_set_boundaries(or_ctx* c, int nBCSides, const int* BC, int nBCs, const int* BoundaryType, int nRefState,
                      const double* RefStatePrim);
/* BCDataPrim(6,0:N,0:N,nBCSides) of the boundaries of types 12 and 121 (equation_vars.f90). */
void or_set_boundary_data(or_ctx* c, const double* BCDataPrim);
/* One boundary point (GetBoundaryState, Lifting_GetBoundaryFlux before SurfElem, GetBoundaryFlux
 * before SurfElem): inner primitive state Pm(6), frame nv/t1/t2, inner gradients gx/gy/gz(4) of
 * (u,v,w,T); ref = RefStatePrim(:,BCState) or NULL, data = BCDataPrim of the point or NULL. */
void or_boundary_point(const or_params* p, int type, const double* ref, const double* data, const double* Pm,
                       const double* nv, const double* t1, const double* t2, const double* gx, const double* gy,
                       const double* gz, double* Pb, double* Flift, double* Flux);
/* The gradients a slip wall's (9 or 91) viscous flux sees (getboundaryflux.f90:622-776). */
void or_slip_gradients(int type, const double* nv, const double* t1, const double* t2, const double* gx,
                       const double* gy, const double* gz, double* qx, double* qy, double* qz);
void or_set_threads(int n);
/* DGTimeDerivative_weakForm (dg.f90:221-678): U(5,0:N,0:N,0:N,nElems) -> Ut. */
void or_dg_time_derivative(or_ctx* c, double t, const double* U, double* Ut);
/* Lifted gradients of the last RHS call: (4,0:N,0:N,0:N,nElems) for (u,v,w,T) in x,y,z. */
void or_get_gradients(or_ctx* c, double* gx, double* gy, double* gz);
// density gradient of the last right-hand side (PP_OPTLIFT=0), (0:N,0:N,0:N,nElems) each
void or_get_density_gradient(or_ctx* c, double* gx, double* gy, double* gz);
/* Face arrays of the last RHS call (FLEXI layout). */
void or_get_faces(or_ctx* c, double* U_master, double* U_slave, double* Flux_master);
/* CALCTIMESTEP (calctimestep.f90:89-227) with CFLScale/DFLScale already scaled by FillCFL_DFL. */
double or_calc_timestep(or_ctx* c, const double* U, double CFLScale, double DFLScale, int* errType,
                        double* dtElem);
/* FillCFL_DFL (timedisc_func.f90:463-527) for CarpenterRK4-5. */
void or_fill_cfl_dfl(int N, int nodeType, double CFL_in, double DFL_in, double* CFL_out,
                     double* DFL_out);
/* Carpenter RK4-5 coefficients (timedisc_vars.f90:175-214): RKA(2:5) RKb(1:5) RKc(2:5). */
void or_rk_coefs(double* RKA, double* RKb, double* RKc);
/* TimeStepByLSERKW2 (timestep.f90:82-152). U updated in place, Ut_tmp is the register. */
void or_timestep_lserkw2(or_ctx* c, double t, double dt, double* U, double* Ut_tmp);

/* ---- TGV test case (testcase/taylorgreenvortex/testcase.f90) ---- */
void or_tgv_ic(const or_params* p, double Ma, int n, const double* x, double* U);
/* AnalyzeTestcase (testcase.f90:241-479): out[15] in the CSV column order. */
void or_tgv_analyze(or_ctx* c, const double* U, const double* gx, const double* gy,
                    const double* gz, int NAnalyze, double* out);

/* ---- single-element entry points used by FLEXI's unit-test fixtures ---- */
void or_prolong_to_face_raw(int N, int nodeType, int nSides, const int* S2E, const double* L_Minus,
                            const double* L_Plus, int nVar, const double* Uvol, double* Um, double* Us);
void or_surf_int_raw(int N, int nodeType, int nSides, const int* S2E, const double* L_HatMinus,
                     const double* L_HatPlus, int nVar, const double* Fm, const double* Fs,
                     double slave_sign, double* Ut);

#ifdef __cplusplus
}
#endif
