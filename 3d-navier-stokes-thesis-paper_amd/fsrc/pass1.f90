!=====================================================================
! Replacement for MODULE pass1 (PFC/pass1.f90): nonlinear runs on the GPU
! (chan_nonlinear); the NL term stays on the device for rhs_solve.
!=====================================================================
MODULE pass1
  USE, INTRINSIC :: iso_c_binding
  USE runparms
  USE gridparms
  USE data
  USE chanflow_f
  USE chanflow_sync
  IMPLICIT NONE
  PRIVATE
  PUBLIC :: nonlinear, advance_scalar

CONTAINS

  ! PFC/pass1.f90:26-252.  dt is adapted on the device when dt_flag == 1;
  ! a CFL abort returns the reference's message and STOPs.  In the time loop
  ! the NL term stays on the device for rhs_solve.  The call with step = -1
  ! is compute_initial_pressure's (PFC/init_vel.f90:444-461), whose host code
  ! reads the NL term from data1 and then rebuilds data1 and data4
  ! (:465-513): the NL term comes back into data1 and the host arrays are
  ! declared the newest copy, so the first step uploads them again.
  SUBROUTINE nonlinear
    REAL(c_double) :: cfl
    CALL chanflow_ensure_device()
    CALL chanflow_check(chan_nonlinear(chanflow_ctx, step, rk_step, dt, cfl), 'nonlinear')
    IF (step == -1) THEN
       CALL chanflow_download_data1(data1)
       chanflow_on_device = .FALSE.
    END IF
  END SUBROUTINE nonlinear

  ! PFC/pass1.f90:254-: passive scalars are outside this library.
  SUBROUTINE advance_scalar(maxm)
    INTEGER, INTENT(IN) :: maxm
    IF (nscalar > 0) STOP 'chanflow: advance_scalar is not provided (nscalar must be 0)'
  END SUBROUTINE advance_scalar

END MODULE pass1
