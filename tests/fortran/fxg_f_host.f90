!> A Fortran host of the fxg library through MOD_FXG, shaped like FLEXI's time loop
!> (src/timedisc/timedisc.f90:160-212): CALCTIMESTEP, then TimeStep(t) = TimeStepByLSERKW2
!> (src/timedisc/timestep.f90:82-152), then t = t + dt, nSteps times, on the GPU state.
!>
!>   fxg_f_host <mesh.h5> <N> <nSteps> <CFLScale> <DFLScale> <state_in.h5> <state_out.h5>
!>
!> Mesh and State files go through libfxg_io (bound here with ISO_C_BINDING, as a host without
!> FLEXI's HDF5 layer would); the solver through MOD_FXG only. Scheme as the C host
!> tests/c_host/fxg_c_host.c: Gauss-Lobatto, split form PI, RoeEntropyFix, BR1, kappa 1.4,
!> R 71.42857, Pr 0.72, mu0 6.25e-4. Prints "dt=<dt>" per step and "ok t=<t> rho_sum=<sum>".
PROGRAM fxg_f_host
USE ISO_C_BINDING
USE MOD_FXG
IMPLICIT NONE
INTERFACE
  INTEGER(C_INT) FUNCTION fxg_io_mesh_size(path,nElems,nSides,NGeo) BIND(C,NAME='fxg_io_mesh_size')
    IMPORT :: C_INT, C_CHAR
    CHARACTER(KIND=C_CHAR),INTENT(IN) :: path(*)
    INTEGER(C_INT),INTENT(OUT)        :: nElems, nSides, NGeo
  END FUNCTION
  INTEGER(C_INT) FUNCTION fxg_io_read_mesh(path,nElems,nSides,NGeo,SideToElem,ElemToSide,NodeCoords) &
      BIND(C,NAME='fxg_io_read_mesh')
    IMPORT :: C_INT, C_CHAR, C_DOUBLE
    CHARACTER(KIND=C_CHAR),INTENT(IN) :: path(*)
    INTEGER(C_INT),VALUE              :: nElems, nSides, NGeo
    INTEGER(C_INT),INTENT(OUT)        :: SideToElem(*), ElemToSide(*)
    REAL(C_DOUBLE),INTENT(OUT)        :: NodeCoords(*)
  END FUNCTION
  INTEGER(C_INT) FUNCTION fxg_io_read_state(path,nVar,N,nElems,U,time,nodeType,nodeTypeLen,meshFile,meshFileLen) &
      BIND(C,NAME='fxg_io_read_state')
    IMPORT :: C_INT, C_CHAR, C_DOUBLE
    CHARACTER(KIND=C_CHAR),INTENT(IN)  :: path(*)
    INTEGER(C_INT),VALUE               :: nVar, N, nElems, nodeTypeLen, meshFileLen
    REAL(C_DOUBLE),INTENT(OUT)         :: U(*), time
    CHARACTER(KIND=C_CHAR),INTENT(OUT) :: nodeType(*), meshFile(*)
  END FUNCTION
  INTEGER(C_INT) FUNCTION fxg_io_write_state(path,projectName,meshFile,nodeType,N,nElems,time,U) &
      BIND(C,NAME='fxg_io_write_state')
    IMPORT :: C_INT, C_CHAR, C_DOUBLE
    CHARACTER(KIND=C_CHAR),INTENT(IN) :: path(*), projectName(*), meshFile(*), nodeType(*)
    INTEGER(C_INT),VALUE              :: N, nElems
    REAL(C_DOUBLE),VALUE              :: time
    REAL(C_DOUBLE),INTENT(IN)         :: U(*)
  END FUNCTION
  FUNCTION fxg_io_last_error() BIND(C,NAME='fxg_io_last_error')
    IMPORT :: C_PTR
    TYPE(C_PTR) :: fxg_io_last_error
  END FUNCTION
END INTERFACE
CHARACTER(LEN=512)          :: meshFile, stateIn, stateOut, arg
CHARACTER(KIND=C_CHAR,LEN=64)  :: nodeType
CHARACTER(KIND=C_CHAR,LEN=256) :: stateMesh
INTEGER(C_INT)              :: N, nSteps, nElems, nSides, NGeo, errType, iter, i
INTEGER(C_INT),ALLOCATABLE  :: SideToElem(:,:), ElemToSide(:,:,:)
REAL(C_DOUBLE),ALLOCATABLE  :: NodeCoords(:,:,:,:,:), U(:,:,:,:,:)
REAL(C_DOUBLE)              :: CFLScale, DFLScale, t, dt, rhoSum
REAL(C_DOUBLE)              :: RKA(4), RKb(5), RKc(4)
TYPE(fxg_params)            :: prm
TYPE(C_PTR)                 :: ctx

IF (COMMAND_ARGUMENT_COUNT().NE.7) THEN
  WRITE(*,'(A)') 'usage: fxg_f_host mesh.h5 N nSteps CFLScale DFLScale state_in.h5 state_out.h5'
  STOP 2
END IF
CALL GET_COMMAND_ARGUMENT(1, meshFile)
CALL GET_COMMAND_ARGUMENT(2, arg); READ(arg,*) N
CALL GET_COMMAND_ARGUMENT(3, arg); READ(arg,*) nSteps
CALL GET_COMMAND_ARGUMENT(4, arg); READ(arg,*) CFLScale
CALL GET_COMMAND_ARGUMENT(5, arg); READ(arg,*) DFLScale
CALL GET_COMMAND_ARGUMENT(6, stateIn)
CALL GET_COMMAND_ARGUMENT(7, stateOut)

IF (fxg_io_mesh_size(TRIM(meshFile)//C_NULL_CHAR, nElems, nSides, NGeo).NE.0) CALL AbortIO('mesh size')
ALLOCATE(SideToElem(5,nSides), ElemToSide(2,6,nElems), NodeCoords(3,0:NGeo,0:NGeo,0:NGeo,nElems))
ALLOCATE(U(5,0:N,0:N,0:N,nElems))
IF (fxg_io_read_mesh(TRIM(meshFile)//C_NULL_CHAR, nElems, nSides, NGeo, SideToElem, ElemToSide, NodeCoords).NE.0) &
  CALL AbortIO('read mesh')

prm = fxg_params(N, FXG_NODE_GAUSS_LOBATTO, FXG_SPLITDG_PI, FXG_RIEMANN_ROEENTROPYFIX, 1.4_C_DOUBLE, &
                 71.42857_C_DOUBLE, 0.72_C_DOUBLE, 6.25E-4_C_DOUBLE, 0, 0, FXG_PREC_FP64, FXG_METRICS_OWN)
IF (fxg_create_from_nodes(prm, nElems, nSides, SideToElem, ElemToSide, NGeo, NodeCoords, ctx).NE.0) &
  CALL AbortFXG('create')
IF (fxg_io_read_state(TRIM(stateIn)//C_NULL_CHAR, 5, N, nElems, U, t, nodeType, LEN(nodeType), &
                      stateMesh, LEN(stateMesh)).NE.0) CALL AbortIO('read state')
IF (fxg_set_state(ctx, U).NE.0) CALL AbortFXG('set state')

! CarpenterRK4-5 (src/timedisc/timedisc_vars.f90:200-214): RKA(2:5), RKb(1:5), RKc(2:5)
RKA = (/567301805773._C_DOUBLE/1357537059087._C_DOUBLE, 2404267990393._C_DOUBLE/2016746695238._C_DOUBLE, &
        3550918686646._C_DOUBLE/2091501179385._C_DOUBLE, 1275806237668._C_DOUBLE/842570457699._C_DOUBLE/)
RKb = (/1432997174477._C_DOUBLE/9575080441755._C_DOUBLE, 5161836677717._C_DOUBLE/13612068292357._C_DOUBLE, &
        1720146321549._C_DOUBLE/2090206949498._C_DOUBLE, 3134564353537._C_DOUBLE/4481467310338._C_DOUBLE, &
        2277821191437._C_DOUBLE/14882151754819._C_DOUBLE/)
RKc = (/1432997174477._C_DOUBLE/9575080441755._C_DOUBLE, 2526269341429._C_DOUBLE/6820363962896._C_DOUBLE, &
        2006345519317._C_DOUBLE/3224310063776._C_DOUBLE, 2802321613138._C_DOUBLE/2924317926251._C_DOUBLE/)

DO iter = 1, nSteps
  ! dt = CALCTIMESTEP(errType); CALL TimeStep(t); t = t + dt  (timedisc.f90:160-212)
  IF (fxg_calc_timestep(ctx, CFLScale, DFLScale, dt, errType, C_NULL_PTR).NE.0) CALL AbortFXG('CALCTIMESTEP')
  IF (fxg_timestep_lserkw2(ctx, t, dt, 5, RKA, RKb, RKc).NE.0) CALL AbortFXG('TimeStepByLSERKW2')
  t = t + dt
  WRITE(*,'(A,ES25.17)') 'dt=', dt
END DO

IF (fxg_get_state(ctx, U).NE.0) CALL AbortFXG('get state')
i = INDEX(nodeType, C_NULL_CHAR) - 1
IF (fxg_io_write_state(TRIM(stateOut)//C_NULL_CHAR, 'f_host'//C_NULL_CHAR, TRIM(meshFile)//C_NULL_CHAR, &
                       nodeType(1:i)//C_NULL_CHAR, N, nElems, t, U).NE.0) CALL AbortIO('write state')
rhoSum = SUM(U(1,:,:,:,:))
WRITE(*,'(A,ES25.17,A,ES25.17)') 'ok t=', t, ' rho_sum=', rhoSum
IF (fxg_destroy(ctx).NE.0) CALL AbortFXG('destroy')

CONTAINS

SUBROUTINE AbortFXG(what)
CHARACTER(LEN=*),INTENT(IN) :: what
WRITE(*,'(3A)') what, ': ', fxg_string(fxg_last_error())
STOP 1
END SUBROUTINE AbortFXG

SUBROUTINE AbortIO(what)
CHARACTER(LEN=*),INTENT(IN) :: what
WRITE(*,'(3A)') what, ': ', fxg_string(fxg_io_last_error())
STOP 1
END SUBROUTINE AbortIO

END PROGRAM fxg_f_host
