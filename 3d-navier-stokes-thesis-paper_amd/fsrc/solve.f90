!=====================================================================
! Replacement for MODULE solve (PFC/solve.f90): the banded y-solves run on
! the GPU (chan_solve_helmholtz, chan_solve_p).  strip_divergence and
! subtract_pressure_gradient keep the reference's per-mode interface for
! the kept initialisation code (PFC/init_vel.f90:275-311, 509, 531); the
! whole-field device versions of that initialisation are chan_initialize_flow
! and chan_remove_divergence.
!=====================================================================
MODULE solve
  USE, INTRINSIC :: iso_c_binding
  USE runparms
  USE gridparms
  USE diff_int
  USE chanflow_f
  USE chanflow_sync
  IMPLICIT NONE
  PRIVATE
  PUBLIC :: solve_helmholtz, solve_real_helmholtz, solve_p, &
       strip_divergence, subtract_pressure_gradient

CONTAINS

  ! PFC/solve.f90:14-50
  SUBROUTINE strip_divergence(cu, cp, coef, p, q)
    COMPLEX(KIND=prec), INTENT(INOUT) :: cu(1:ny+1,3)
    COMPLEX(KIND=prec), INTENT(OUT)   :: cp(1:ny+1)
    REAL(KIND=prec), INTENT(IN)       :: coef
    INTEGER, INTENT(IN)               :: p, q
    COMPLEX(KIND=prec) :: dv(1:ny+1), topbc, botbc
    REAL(KIND=prec)    :: tmp, alpha
    tmp = 1.d0/coef
    CALL d_dy(ny, cu(1,2), ny+1, dv, dh)
    cp(1)    = 0.d0
    cp(2:ny) = tmp*(ikx_me(p)*cu(2:ny,1) + dv(2:ny) + ikz(q)*cu(2:ny,3))
    cp(ny+1) = 0.d0
    topbc = 0.d0
    botbc = 0.d0
    alpha = -k2_me(q,p)
    CALL solve_p(cp, alpha, topbc, botbc)
    CALL subtract_pressure_gradient(cu, cp, coef, p, q)
  END SUBROUTINE strip_divergence

  ! PFC/solve.f90:54-77
  SUBROUTINE subtract_pressure_gradient(crhs, cpress, coef, p, q)
    COMPLEX(KIND=prec), DIMENSION(1:ny+1,1:3), INTENT(INOUT) :: crhs
    COMPLEX(KIND=prec), DIMENSION(1:ny+1), INTENT(IN)        :: cpress
    REAL(KIND=prec), INTENT(IN)                              :: coef
    INTEGER, INTENT(IN) :: p, q
    COMPLEX(KIND=prec), DIMENSION(1:ny) :: dp
    CALL d_dy(ny+1, cpress, ny, dp, dn)
    crhs(2:ny,1)   = crhs(2:ny,1)   - coef*ikx_me(p)*cpress(2:ny)
    crhs(2:ny-1,2) = crhs(2:ny-1,2) - coef*dp(2:ny-1)
    crhs(2:ny,3)   = crhs(2:ny,3)   - coef*ikz(q)*cpress(2:ny)
  END SUBROUTINE subtract_pressure_gradient

  ! PFC/solve.f90:154-264 (Dirichlet rows; the channel never sets the
  ! Neumann flags, and the library rejects them)
  SUBROUTINE SOLVE_HELMHOLTZ(N, NRHS, F, ALPHA, DIFF, D2, TOPBC, BOTBC, TOP_NEUMANN, BOT_NEUMANN)
    INTEGER, INTENT(IN)               :: N, NRHS, TOP_NEUMANN, BOT_NEUMANN
    COMPLEX(KIND=prec), INTENT(INOUT) :: F(N,NRHS)
    REAL(KIND=prec), INTENT(IN)       :: ALPHA, DIFF, D2(3,N)
    COMPLEX(KIND=prec), INTENT(IN)    :: TOPBC(NRHS), BOTBC(NRHS)
    REAL(KIND=prec) :: al(NRHS)
    CALL chanflow_ensure_device()
    al = ALPHA
    CALL chanflow_check(chan_solve_helmholtz(chanflow_ctx, N, NRHS, F, al, DIFF, D2, TOPBC, BOTBC, &
         TOP_NEUMANN, BOT_NEUMANN), 'solve_helmholtz')
  END SUBROUTINE SOLVE_HELMHOLTZ

  ! PFC/solve.f90:81-150: the real system is the complex one with zero
  ! imaginary parts.
  SUBROUTINE SOLVE_REAL_HELMHOLTZ(N, NRHS, F, ALPHA, DIFF, D2, TOPBC, BOTBC, TOP_NEUMANN, BOT_NEUMANN)
    INTEGER, INTENT(IN)            :: N, NRHS, TOP_NEUMANN, BOT_NEUMANN
    REAL(KIND=prec), INTENT(INOUT) :: F(N,NRHS)
    REAL(KIND=prec), INTENT(IN)    :: ALPHA, DIFF, D2(3,N)
    REAL(KIND=prec), INTENT(IN)    :: TOPBC(NRHS), BOTBC(NRHS)
    COMPLEX(KIND=prec) :: cf(N,NRHS), ct(NRHS), cb(NRHS)
    cf = CMPLX(F, 0.d0, KIND=prec)
    ct = CMPLX(TOPBC, 0.d0, KIND=prec)
    cb = CMPLX(BOTBC, 0.d0, KIND=prec)
    CALL SOLVE_HELMHOLTZ(N, NRHS, cf, ALPHA, DIFF, D2, ct, cb, TOP_NEUMANN, BOT_NEUMANN)
    F = DBLE(cf)
  END SUBROUTINE SOLVE_REAL_HELMHOLTZ

  ! PFC/solve.f90:269-338.  The device solve uses zero wall data in the
  ! re-extrapolation (:335-336); non-zero TOPBC/BOTBC add TOPBC/DM(3,2) and
  ! BOTBC/DM(1,1) to the wall values.
  SUBROUTINE SOLVE_P(P, ALPHA, TOPBC, BOTBC)
    COMPLEX(KIND=prec), INTENT(INOUT) :: P(NY+1)
    REAL(KIND=prec), INTENT(IN)       :: ALPHA
    COMPLEX(KIND=prec), INTENT(IN)    :: TOPBC, BOTBC
    REAL(KIND=prec) :: al(1)
    CALL chanflow_ensure_device()
    al(1) = ALPHA
    CALL chanflow_check(chan_solve_p(chanflow_ctx, 1, P, al), 'solve_p')
    IF (BOTBC /= (0.d0, 0.d0)) P(1) = P(1) + BOTBC/DM(1,1)
    IF (TOPBC /= (0.d0, 0.d0)) P(NY+1) = P(NY+1) + TOPBC/DM(3,2)
  END SUBROUTINE SOLVE_P

END MODULE solve
