!=====================================================================
! Replacement for MODULE pass1 (PFC/pass1.f90): nonlinear runs on the GPU
! (chan_nonlinear); the NL term stays on the device for rhs_solve.
!=====================================================================
MODULE pass1
  USE, INTRINSIC :: iso_c_binding
  USE runparms
  USE gridparms
  USE chanflow_f
  USE chanflow_sync
  IMPLICIT NONE
  PRIVATE
  PUBLIC :: nonlinear, advance_scalar

CONTAINS

  ! PFC/pass1.f90:26-252.  dt is adapted on the device when dt_flag == 1;
  ! a CFL abort returns the reference's message and STOPs.
  SUBROUTINE nonlinear
    REAL(c_double) :: cfl
    CALL chanflow_ensure_device()
    CALL chanflow_check(chan_nonlinear(chanflow_ctx, step, rk_step, dt, cfl), 'nonlinear')
  END SUBROUTINE nonlinear

  ! PFC/pass1.f90:254-: passive scalars are outside this library.
  SUBROUTINE advance_scalar(maxm)
    INTEGER, INTENT(IN) :: maxm
    IF (nscalar > 0) STOP 'chanflow: advance_scalar is not provided (nscalar must be 0)'
  END SUBROUTINE advance_scalar

END MODULE pass1
