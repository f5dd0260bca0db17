!> ISO_C_BINDING interface to libfxg (include/fxg.h) for a FLEXI host.
!>
!> The derived types mirror fxg_params and fxg_mesh field by field; the procedures bind the
!> C entry points one to one. Arrays are passed as FLEXI holds them (column-major, variable
!> index fastest), e.g. U(1:5,0:N,0:N,0:N,1:nElems) from MOD_DG_Vars (src/dg/dg_vars.f90),
!> so no copies or transposes are needed on the host side.
MODULE MOD_FXG
USE ISO_C_BINDING
IMPLICIT NONE
PRIVATE

INTEGER(C_INT),PARAMETER,PUBLIC :: FXG_NODE_GAUSS            = 1
INTEGER(C_INT),PARAMETER,PUBLIC :: FXG_NODE_GAUSS_LOBATTO    = 2
INTEGER(C_INT),PARAMETER,PUBLIC :: FXG_RIEMANN_LF            = 1   ! riemann.f90:44-57
INTEGER(C_INT),PARAMETER,PUBLIC :: FXG_RIEMANN_ROE           = 3
INTEGER(C_INT),PARAMETER,PUBLIC :: FXG_RIEMANN_ROEENTROPYFIX = 33
INTEGER(C_INT),PARAMETER,PUBLIC :: FXG_SPLITDG_NONE          = 0   ! splitflux.f90:63-69
INTEGER(C_INT),PARAMETER,PUBLIC :: FXG_SPLITDG_PI            = 4
INTEGER(C_INT),PARAMETER,PUBLIC :: FXG_PREC_FP64             = 0
INTEGER(C_INT),PARAMETER,PUBLIC :: FXG_PREC_FP32_BULK        = 1
INTEGER(C_INT),PARAMETER,PUBLIC :: FXG_METRICS_OWN           = 0
INTEGER(C_INT),PARAMETER,PUBLIC :: FXG_METRICS_REFERENCE     = 1

TYPE,BIND(C),PUBLIC :: fxg_params
  INTEGER(C_INT) :: N, nodeType, splitDG, riemann
  REAL(C_DOUBLE) :: kappa, R, Pr, mu0
  INTEGER(C_INT) :: doWeakLifting, doConservativeLifting, precision
  INTEGER(C_INT) :: metrics
END TYPE fxg_params

TYPE,BIND(C),PUBLIC :: fxg_mesh
  INTEGER(C_INT) :: nElems, nSides, firstInnerSide, lastInnerSide
  TYPE(C_PTR)    :: SideToElem, ElemToSide
  TYPE(C_PTR)    :: Metrics_fTilde, Metrics_gTilde, Metrics_hTilde, sJ
  TYPE(C_PTR)    :: NormVec, TangVec1, TangVec2, SurfElem
END TYPE fxg_mesh

PUBLIC :: fxg_last_error, fxg_version, fxg_create, fxg_create_from_nodes, fxg_destroy, fxg_set_boundaries
PUBLIC :: fxg_set_boundary_data
PUBLIC :: fxg_set_state, fxg_get_state, fxg_get_ut, fxg_get_gradients, fxg_get_metrics
PUBLIC :: fxg_tgv_analyze
PUBLIC :: fxg_dg_time_derivative, fxg_timestep_lserkw2, fxg_calc_timestep
PUBLIC :: fxg_dg_time_derivative_dev, fxg_timestep_lserkw2_dev, fxg_pipeline, fxg_sync, fxg_sizes
PUBLIC :: fxg_calc_timestep_dev_noalloc
PUBLIC :: fxg_rhs_vjp_dev, fxg_forward_n_steps, fxg_adjoint_n_steps, fxg_adjoint_storage
PUBLIC :: fxg_profile_enable, fxg_profile_read, fxg_string

INTERFACE
  FUNCTION fxg_last_error() BIND(C,NAME='fxg_last_error')
    IMPORT :: C_PTR
    TYPE(C_PTR) :: fxg_last_error
  END FUNCTION
  FUNCTION fxg_version() BIND(C,NAME='fxg_version')
    IMPORT :: C_PTR
    TYPE(C_PTR) :: fxg_version
  END FUNCTION
  INTEGER(C_INT) FUNCTION fxg_create(params,mesh,ctx) BIND(C,NAME='fxg_create')
    IMPORT :: C_INT, C_PTR, fxg_params, fxg_mesh
    TYPE(fxg_params),INTENT(IN) :: params
    TYPE(fxg_mesh),INTENT(IN)   :: mesh
    TYPE(C_PTR),INTENT(OUT)     :: ctx
  END FUNCTION
  INTEGER(C_INT) FUNCTION fxg_create_from_nodes(params,nElems,nSides,SideToElem,ElemToSide,NGeo,NodeCoords,ctx) &
      BIND(C,NAME='fxg_create_from_nodes')
    IMPORT :: C_INT, C_DOUBLE, C_PTR, fxg_params
    TYPE(fxg_params),INTENT(IN)      :: params
    INTEGER(C_INT),VALUE             :: nElems, nSides, NGeo
    INTEGER(C_INT),INTENT(IN)        :: SideToElem(*), ElemToSide(*)
    REAL(C_DOUBLE),INTENT(IN)        :: NodeCoords(*)
    TYPE(C_PTR),INTENT(OUT)          :: ctx
  END FUNCTION
  INTEGER(C_INT) FUNCTION fxg_destroy(ctx) BIND(C,NAME='fxg_destroy')
    IMPORT :: C_INT, C_PTR
    TYPE(C_PTR),VALUE :: ctx
  END FUNCTION
  INTEGER(C_INT) FUNCTION fxg_set_boundaries(ctx,nBCs,BC,BoundaryType,nRefState,RefStatePrim) &
      BIND(C,NAME='fxg_set_boundaries')
    IMPORT :: C_INT, C_DOUBLE, C_PTR
    TYPE(C_PTR),VALUE         :: ctx
    INTEGER(C_INT),VALUE      :: nBCs, nRefState
    INTEGER(C_INT),INTENT(IN) :: BC(*), BoundaryType(*)
    REAL(C_DOUBLE),INTENT(IN) :: RefStatePrim(*)
  END FUNCTION
  ! BCDataPrim(6,0:N,0:N,nBCSides) of the boundaries of types 12 and 121 (equation_vars.f90)
  INTEGER(C_INT) FUNCTION fxg_set_boundary_data(ctx,BCDataPrim) BIND(C,NAME='fxg_set_boundary_data')
    IMPORT :: C_INT, C_DOUBLE, C_PTR
    TYPE(C_PTR),VALUE         :: ctx
    REAL(C_DOUBLE),INTENT(IN) :: BCDataPrim(*)
  END FUNCTION
  INTEGER(C_INT) FUNCTION fxg_set_state(ctx,U) BIND(C,NAME='fxg_set_state')
    IMPORT :: C_INT, C_DOUBLE, C_PTR
    TYPE(C_PTR),VALUE         :: ctx
    REAL(C_DOUBLE),INTENT(IN) :: U(*)
  END FUNCTION
  INTEGER(C_INT) FUNCTION fxg_get_state(ctx,U) BIND(C,NAME='fxg_get_state')
    IMPORT :: C_INT, C_DOUBLE, C_PTR
    TYPE(C_PTR),VALUE          :: ctx
    REAL(C_DOUBLE),INTENT(OUT) :: U(*)
  END FUNCTION
  INTEGER(C_INT) FUNCTION fxg_get_ut(ctx,Ut) BIND(C,NAME='fxg_get_ut')
    IMPORT :: C_INT, C_DOUBLE, C_PTR
    TYPE(C_PTR),VALUE          :: ctx
    REAL(C_DOUBLE),INTENT(OUT) :: Ut(*)
  END FUNCTION
  ! gradUx/y/z(nVarLifting,0:N,0:N,0:N,nElems): nVarLifting = PP_nVarLifting of the host build
  INTEGER(C_INT) FUNCTION fxg_get_gradients(ctx,nVarLifting,gx,gy,gz) BIND(C,NAME='fxg_get_gradients')
    IMPORT :: C_INT, C_DOUBLE, C_PTR
    TYPE(C_PTR),VALUE          :: ctx
    INTEGER(C_INT),VALUE       :: nVarLifting
    REAL(C_DOUBLE),INTENT(OUT) :: gx(*), gy(*), gz(*)
  END FUNCTION
  ! the 15 values of writeBuf in AnalyzeTestcase (testcase.f90:464-466), evaluated on the device
  INTEGER(C_INT) FUNCTION fxg_tgv_analyze(ctx,NAnalyze,vals) BIND(C,NAME='fxg_tgv_analyze')
    IMPORT :: C_INT, C_DOUBLE, C_PTR
    TYPE(C_PTR),VALUE          :: ctx
    INTEGER(C_INT),VALUE       :: NAnalyze
    REAL(C_DOUBLE),INTENT(OUT) :: vals(15)
  END FUNCTION
  INTEGER(C_INT) FUNCTION fxg_get_metrics(ctx,Mf,Mg,Mh,sJ,NormVec,TangVec1,TangVec2,SurfElem,Elem_xGP) &
      BIND(C,NAME='fxg_get_metrics')
    IMPORT :: C_INT, C_PTR
    TYPE(C_PTR),VALUE :: ctx, Mf, Mg, Mh, sJ, NormVec, TangVec1, TangVec2, SurfElem, Elem_xGP
  END FUNCTION
  INTEGER(C_INT) FUNCTION fxg_dg_time_derivative(ctx,t) BIND(C,NAME='fxg_dg_time_derivative')
    IMPORT :: C_INT, C_DOUBLE, C_PTR
    TYPE(C_PTR),VALUE    :: ctx
    REAL(C_DOUBLE),VALUE :: t
  END FUNCTION
  INTEGER(C_INT) FUNCTION fxg_timestep_lserkw2(ctx,t,dt,nStages,RKA,RKb,RKc) BIND(C,NAME='fxg_timestep_lserkw2')
    IMPORT :: C_INT, C_DOUBLE, C_PTR
    TYPE(C_PTR),VALUE         :: ctx
    REAL(C_DOUBLE),VALUE      :: t, dt
    INTEGER(C_INT),VALUE      :: nStages
    REAL(C_DOUBLE),INTENT(IN) :: RKA(*), RKb(*), RKc(*)
  END FUNCTION
  INTEGER(C_INT) FUNCTION fxg_calc_timestep(ctx,CFLScale,DFLScale,dt,errType,dtElem) BIND(C,NAME='fxg_calc_timestep')
    IMPORT :: C_INT, C_DOUBLE, C_PTR
    TYPE(C_PTR),VALUE          :: ctx
    REAL(C_DOUBLE),VALUE       :: CFLScale, DFLScale
    REAL(C_DOUBLE),INTENT(OUT) :: dt
    INTEGER(C_INT),INTENT(OUT) :: errType
    TYPE(C_PTR),VALUE          :: dtElem
  END FUNCTION
  INTEGER(C_INT) FUNCTION fxg_dg_time_derivative_dev(ctx,t,U_dev,Ut_dev,stream) BIND(C,NAME='fxg_dg_time_derivative_dev')
    IMPORT :: C_INT, C_DOUBLE, C_PTR
    TYPE(C_PTR),VALUE    :: ctx, U_dev, Ut_dev, stream
    REAL(C_DOUBLE),VALUE :: t
  END FUNCTION
  INTEGER(C_INT) FUNCTION fxg_calc_timestep_dev_noalloc(ctx,U_dev,CFLScale,DFLScale,dt_dev,errType_dev,stream) &
      BIND(C,NAME='fxg_calc_timestep_dev_noalloc')
    IMPORT :: C_INT, C_DOUBLE, C_PTR
    TYPE(C_PTR),VALUE    :: ctx, U_dev, dt_dev, errType_dev, stream
    REAL(C_DOUBLE),VALUE :: CFLScale, DFLScale
  END FUNCTION
  INTEGER(C_INT) FUNCTION fxg_timestep_lserkw2_dev(ctx,t,dt,nStages,RKA,RKb,RKc,U_dev,R_dev,stream) &
      BIND(C,NAME='fxg_timestep_lserkw2_dev')
    IMPORT :: C_INT, C_DOUBLE, C_PTR
    TYPE(C_PTR),VALUE         :: ctx, U_dev, R_dev, stream
    REAL(C_DOUBLE),VALUE      :: t, dt
    INTEGER(C_INT),VALUE      :: nStages
    REAL(C_DOUBLE),INTENT(IN) :: RKA(*), RKb(*), RKc(*)
  END FUNCTION
  INTEGER(C_INT) FUNCTION fxg_rhs_vjp_dev(ctx,t,U_dev,lam_dev,out_dev,stream) BIND(C,NAME='fxg_rhs_vjp_dev')
    IMPORT :: C_INT, C_DOUBLE, C_PTR
    TYPE(C_PTR),VALUE    :: ctx, U_dev, lam_dev, out_dev, stream
    REAL(C_DOUBLE),VALUE :: t
  END FUNCTION
  INTEGER(C_INT) FUNCTION fxg_forward_n_steps(ctx,nSteps,dt,nStages,RKA,RKb,RKc,U_dev,ckpt_dev,checkpointEvery,stream) &
      BIND(C,NAME='fxg_forward_n_steps')
    IMPORT :: C_INT, C_DOUBLE, C_PTR
    TYPE(C_PTR),VALUE         :: ctx, U_dev, ckpt_dev, stream
    INTEGER(C_INT),VALUE      :: nSteps, nStages, checkpointEvery
    REAL(C_DOUBLE),INTENT(IN) :: dt(*), RKA(*), RKb(*), RKc(*)
  END FUNCTION
  INTEGER(C_INT) FUNCTION fxg_adjoint_n_steps(ctx,nSteps,dt,nStages,RKA,RKb,RKc,U0_dev,ckpt_dev,ubar_dev, &
      checkpointEvery,stream) BIND(C,NAME='fxg_adjoint_n_steps')
    IMPORT :: C_INT, C_DOUBLE, C_PTR
    TYPE(C_PTR),VALUE         :: ctx, U0_dev, ckpt_dev, ubar_dev, stream
    INTEGER(C_INT),VALUE      :: nSteps, nStages, checkpointEvery
    REAL(C_DOUBLE),INTENT(IN) :: dt(*), RKA(*), RKb(*), RKc(*)
  END FUNCTION
  INTEGER(C_INT) FUNCTION fxg_adjoint_storage(ctx,nSteps,nStages,checkpointEvery,doubles) &
      BIND(C,NAME='fxg_adjoint_storage')
    IMPORT :: C_INT, C_SIZE_T, C_PTR
    TYPE(C_PTR),VALUE              :: ctx
    INTEGER(C_INT),VALUE           :: nSteps, nStages, checkpointEvery
    INTEGER(C_SIZE_T),INTENT(OUT)  :: doubles
  END FUNCTION
  FUNCTION fxg_pipeline(ctx) BIND(C,NAME='fxg_pipeline')
    IMPORT :: C_PTR
    TYPE(C_PTR),VALUE :: ctx
    TYPE(C_PTR)       :: fxg_pipeline
  END FUNCTION
  INTEGER(C_INT) FUNCTION fxg_sync(ctx) BIND(C,NAME='fxg_sync')
    IMPORT :: C_INT, C_PTR
    TYPE(C_PTR),VALUE :: ctx
  END FUNCTION
  INTEGER(C_INT) FUNCTION fxg_sizes(ctx,N,nElems,nSides) BIND(C,NAME='fxg_sizes')
    IMPORT :: C_INT, C_PTR
    TYPE(C_PTR),VALUE          :: ctx
    INTEGER(C_INT),INTENT(OUT) :: N, nElems, nSides
  END FUNCTION
  INTEGER(C_INT) FUNCTION fxg_profile_enable(ctx,enable) BIND(C,NAME='fxg_profile_enable')
    IMPORT :: C_INT, C_PTR
    TYPE(C_PTR),VALUE    :: ctx
    INTEGER(C_INT),VALUE :: enable
  END FUNCTION
  INTEGER(C_INT) FUNCTION fxg_profile_read(ctx,kernel,total_ms,launches) BIND(C,NAME='fxg_profile_read')
    IMPORT :: C_INT, C_DOUBLE, C_PTR
    TYPE(C_PTR),VALUE          :: ctx
    INTEGER(C_INT),VALUE       :: kernel
    REAL(C_DOUBLE),INTENT(OUT) :: total_ms
    INTEGER(C_INT),INTENT(OUT) :: launches
  END FUNCTION
  FUNCTION c_strlen(s) BIND(C,NAME='strlen')
    IMPORT :: C_PTR, C_SIZE_T
    TYPE(C_PTR),VALUE :: s
    INTEGER(C_SIZE_T) :: c_strlen
  END FUNCTION
END INTERFACE

CONTAINS

!> Fortran string of a NUL-terminated C string returned by the library (fxg_last_error,
!> fxg_version, fxg_pipeline).
FUNCTION fxg_string(p) RESULT(s)
TYPE(C_PTR),INTENT(IN)       :: p
CHARACTER(LEN=:),ALLOCATABLE :: s
CHARACTER(KIND=C_CHAR),POINTER :: c(:)
INTEGER :: i, n
IF (.NOT.C_ASSOCIATED(p)) THEN
  s = ''
  RETURN
END IF
n = INT(c_strlen(p))
CALL C_F_POINTER(p, c, [n])
ALLOCATE(CHARACTER(LEN=n) :: s)
DO i = 1, n
  s(i:i) = c(i)
END DO
END FUNCTION fxg_string

END MODULE MOD_FXG
