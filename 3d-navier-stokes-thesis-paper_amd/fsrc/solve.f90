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
  USE chanflow_f
  USE chanflow_sync
  IMPLICIT NONE
  PRIVATE
  PUBLIC :: solve_helmholtz, solve_real_helmholtz, solve_p, &
       strip_divergence, subtract_pressure_gradient

CONTAINS

  ! PFC/solve.f90:14-50: one (kx,kz) column, projected on the GPU
  ! (chan_strip_divergence_modes); p, q are the x- and z-mode indices.
  SUBROUTINE strip_divergence(cu, cp, coef, p, q)
    COMPLEX(KIND=prec), INTENT(INOUT) :: cu(1:ny+1,3)
    COMPLEX(KIND=prec), INTENT(OUT)   :: cp(1:ny+1)
    REAL(KIND=prec), INTENT(IN)       :: coef
    INTEGER, INTENT(IN)               :: p, q
    INTEGER(c_int) :: ip(1), iq(1)
    CALL chanflow_ensure_ops()
    ip(1) = p
    iq(1) = q
    CALL chanflow_check(chan_strip_divergence_modes(chanflow_ctx, 1, cu, cp, coef, ip, iq), 'strip_divergence')
  END SUBROUTINE strip_divergence

  ! PFC/solve.f90:54-77: one column, on the GPU
  ! (chan_subtract_pressure_gradient_modes).
  SUBROUTINE subtract_pressure_gradient(crhs, cpress, coef, p, q)
    COMPLEX(KIND=prec), DIMENSION(1:ny+1,1:3), INTENT(INOUT) :: crhs
    COMPLEX(KIND=prec), DIMENSION(1:ny+1), INTENT(IN)        :: cpress
    REAL(KIND=prec), INTENT(IN)                              :: coef
    INTEGER, INTENT(IN) :: p, q
    INTEGER(c_int) :: ip(1), iq(1)
    CALL chanflow_ensure_ops()
    ip(1) = p
    iq(1) = q
    CALL chanflow_check(chan_subtract_pressure_gradient_modes(chanflow_ctx, 1, crhs, cpress, coef, ip, iq), &
         'subtract_pressure_gradient')
  END SUBROUTINE subtract_pressure_gradient

  ! PFC/solve.f90:154-264 (Dirichlet rows; the channel never sets the
  ! Neumann flags, and the library rejects them)
  SUBROUTINE SOLVE_HELMHOLTZ(N, NRHS, F, ALPHA, DIFF, D2, TOPBC, BOTBC, TOP_NEUMANN, BOT_NEUMANN)
    INTEGER, INTENT(IN)               :: N, NRHS, TOP_NEUMANN, BOT_NEUMANN
    COMPLEX(KIND=prec), INTENT(INOUT) :: F(N,NRHS)
    REAL(KIND=prec), INTENT(IN)       :: ALPHA, DIFF, D2(3,N)
    COMPLEX(KIND=prec), INTENT(IN)    :: TOPBC(NRHS), BOTBC(NRHS)
    REAL(KIND=prec) :: al(NRHS)
    CALL chanflow_ensure_ops()
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
    CALL chanflow_ensure_ops()
    al(1) = ALPHA
    CALL chanflow_check(chan_solve_p(chanflow_ctx, 1, P, al), 'solve_p')
    IF (BOTBC /= (0.d0, 0.d0)) P(1) = P(1) + BOTBC/DM(1,1)
    IF (TOPBC /= (0.d0, 0.d0)) P(NY+1) = P(NY+1) + TOPBC/DM(3,2)
  END SUBROUTINE SOLVE_P

END MODULE solve
