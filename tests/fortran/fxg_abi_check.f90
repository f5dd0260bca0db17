!> CPU-only check of the Fortran binding: the module compiles against include/fxg.h's ABI, the
!> library links, and errors come back through the return code and fxg_last_error before any
!> device work. Every entry point the module exports is called once (with a NULL context, or
!> with parameters the library rejects), so a binding that is missing or not PUBLIC fails to
!> compile here.
PROGRAM fxg_abi_check
USE ISO_C_BINDING
USE MOD_FXG
IMPLICIT NONE
TYPE(fxg_params) :: prm
TYPE(fxg_mesh)   :: mesh
TYPE(C_PTR)      :: ctx, nul
INTEGER(C_INT)   :: rc, nbad, ival(3), ibuf(6)
INTEGER(C_SIZE_T):: ndoubles
REAL(C_DOUBLE)   :: dval, dbuf(15), rk(5)
CHARACTER(LEN=:),ALLOCATABLE :: msg

nul = C_NULL_PTR
WRITE(*,'(A,I0)') 'sizeof(fxg_params)=', C_SIZEOF(prm)
WRITE(*,'(A,I0)') 'sizeof(fxg_mesh)=', C_SIZEOF(mesh)
WRITE(*,'(2A)') 'version=', fxg_string(fxg_version())
prm = fxg_params(9, FXG_NODE_GAUSS_LOBATTO, FXG_SPLITDG_PI, FXG_RIEMANN_ROE, 1.4_C_DOUBLE, 287._C_DOUBLE, &
                 0.72_C_DOUBLE, 1.E-3_C_DOUBLE, 0, 0, FXG_PREC_FP64, FXG_METRICS_OWN)
mesh%nElems = 1; mesh%nSides = 3; mesh%firstInnerSide = 1; mesh%lastInnerSide = 3
rc = fxg_create(prm, mesh, ctx)
msg = fxg_string(fxg_last_error())
WRITE(*,'(A,I0,2A)') 'rc=', rc, ' msg=', msg
IF (rc.EQ.0 .OR. INDEX(msg,'N must be').EQ.0) STOP 1
prm%N = 3
prm%nodeType = FXG_NODE_GAUSS
rc = fxg_create(prm, mesh, ctx)
msg = fxg_string(fxg_last_error())
WRITE(*,'(A,I0,2A)') 'rc=', rc, ' msg=', msg
IF (rc.EQ.0 .OR. INDEX(msg,'Gauss-Lobatto').EQ.0) STOP 2
rc = fxg_create_from_nodes(prm, 1, 3, ibuf, ibuf, 1, dbuf, ctx)
IF (rc.EQ.0) STOP 3

! every entry point that takes a context rejects a NULL one
nbad = 0
rk = 0.5_C_DOUBLE
IF (fxg_destroy(nul).NE.0) STOP 4                      ! destroying NULL is a no-op
IF (LEN(fxg_string(fxg_pipeline(nul))).NE.0) STOP 5
IF (fxg_set_boundaries(nul, 1, ibuf, ibuf, 1, dbuf).EQ.0) nbad = nbad + 1
IF (fxg_set_boundary_data(nul, dbuf).EQ.0) nbad = nbad + 1
IF (fxg_set_state(nul, dbuf).EQ.0) nbad = nbad + 1
IF (fxg_get_state(nul, dbuf).EQ.0) nbad = nbad + 1
IF (fxg_get_ut(nul, dbuf).EQ.0) nbad = nbad + 1
IF (fxg_get_gradients(nul, 5, dbuf, dbuf, dbuf).EQ.0) nbad = nbad + 1
IF (fxg_tgv_analyze(nul, 4, dbuf).EQ.0) nbad = nbad + 1
IF (fxg_get_metrics(nul, nul, nul, nul, nul, nul, nul, nul, nul, nul).EQ.0) nbad = nbad + 1
IF (fxg_dg_time_derivative(nul, 0._C_DOUBLE).EQ.0) nbad = nbad + 1
IF (fxg_timestep_lserkw2(nul, 0._C_DOUBLE, 1._C_DOUBLE, 5, rk, rk, rk).EQ.0) nbad = nbad + 1
IF (fxg_calc_timestep(nul, 0.9_C_DOUBLE, 0.9_C_DOUBLE, dval, rc, nul).EQ.0) nbad = nbad + 1
IF (fxg_dg_time_derivative_dev(nul, 0._C_DOUBLE, nul, nul, nul).EQ.0) nbad = nbad + 1
IF (fxg_calc_timestep_dev_noalloc(nul, nul, 1._C_DOUBLE, 1._C_DOUBLE, nul, nul, nul).EQ.0) nbad = nbad + 1
IF (fxg_timestep_lserkw2_dev(nul, 0._C_DOUBLE, 1._C_DOUBLE, 5, rk, rk, rk, nul, nul, nul).EQ.0) nbad = nbad + 1
IF (fxg_rhs_vjp_dev(nul, 0._C_DOUBLE, nul, nul, nul, nul).EQ.0) nbad = nbad + 1
IF (fxg_forward_n_steps(nul, 1, rk, 5, rk, rk, rk, nul, nul, 1, nul).EQ.0) nbad = nbad + 1
IF (fxg_adjoint_n_steps(nul, 1, rk, 5, rk, rk, rk, nul, nul, nul, 1, nul).EQ.0) nbad = nbad + 1
IF (fxg_adjoint_storage(nul, 1, 5, 1, ndoubles).EQ.0) nbad = nbad + 1
IF (fxg_sync(nul).EQ.0) nbad = nbad + 1
IF (fxg_sizes(nul, ival(1), ival(2), ival(3)).EQ.0) nbad = nbad + 1
IF (fxg_profile_enable(nul, 1).EQ.0) nbad = nbad + 1
IF (fxg_profile_read(nul, 0, dval, rc).EQ.0) nbad = nbad + 1
WRITE(*,'(A,I0)') 'accepted NULL ctx: ', nbad
IF (nbad.NE.0) STOP 6
WRITE(*,'(A)') 'OK'
END PROGRAM fxg_abi_check
