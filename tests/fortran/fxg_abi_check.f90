!> CPU-only check of the Fortran binding: the module compiles against include/fxg.h's ABI, the
!> library links, and parameter errors come back through the return code and fxg_last_error
!> before any device work.
PROGRAM fxg_abi_check
USE ISO_C_BINDING
USE MOD_FXG
IMPLICIT NONE
TYPE(fxg_params) :: prm
TYPE(fxg_mesh)   :: mesh
TYPE(C_PTR)      :: ctx
INTEGER(C_INT)   :: rc
CHARACTER(LEN=:),ALLOCATABLE :: msg

WRITE(*,'(A,I0)') 'sizeof(fxg_params)=', C_SIZEOF(prm)
WRITE(*,'(A,I0)') 'sizeof(fxg_mesh)=', C_SIZEOF(mesh)
WRITE(*,'(2A)') 'version=', fxg_string(fxg_version())
prm = fxg_params(9, FXG_NODE_GAUSS_LOBATTO, FXG_SPLITDG_PI, FXG_RIEMANN_ROE, 1.4_C_DOUBLE, 287._C_DOUBLE, &
                 0.72_C_DOUBLE, 1.E-3_C_DOUBLE, 0, 0, FXG_PREC_FP64)
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
WRITE(*,'(A)') 'OK'
END PROGRAM fxg_abi_check
