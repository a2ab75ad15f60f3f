!=====================================================================
! Replacement for MODULE pass1_linearized (PFC/pass1_linearized.f90).
!=====================================================================
MODULE pass1_linearized
  USE, INTRINSIC :: iso_c_binding
  USE runparms
  USE gridparms
  USE data
  USE chanflow_f
  USE chanflow_sync
  IMPLICIT NONE
  PRIVATE
  PUBLIC :: nonlinear_linearized, init_baseflow

CONTAINS

  ! PFC/pass1_linearized.f90:20-131, laminar branch: u = (1 - yh^2) cos(beta),
  ! w = (1 - yh^2) sin(beta), dt = cfl_max*delta1/32, p_grad = 0.  Reading
  ! mean.hdf (HDF4) is not provided.
  SUBROUTINE init_baseflow
    INTEGER :: k, ierr
    ALLOCATE(baseflow(nx2,local_nz,ny+1,3+nscalar), STAT=ierr)
    IF (ierr /= 0) STOP 'COULD NOT ALLOCATE ARRAY FOR BASEFLOW'
    baseflow = 0.d0
    DO k = 1,ny+1
       baseflow(1:nx2,1:local_nz,k,1) = (1.d0 - yh(k)**2)*COS(beta)
       baseflow(1:nx2,1:local_nz,k,3) = (1.d0 - yh(k)**2)*SIN(beta)
    END DO
    dt = cfl_max*delta1/32.d0
    p_grad = 0.d0
    CALL chanflow_ensure_device()
    CALL chanflow_check(chan_set_baseflow(chanflow_ctx, baseflow), 'init_baseflow')
    IF (me == 0) WRITE(*,*) 'LINEARIZED SIMULATION'
    IF (me == 0) WRITE(*,*) 'CONSTANT DT = ', dt
  END SUBROUTINE init_baseflow

  ! PFC/pass1_linearized.f90:135-509
  SUBROUTINE nonlinear_linearized
    REAL(c_double) :: cfl
    CALL chanflow_ensure_device()
    CALL chanflow_check(chan_nonlinear_linearized(chanflow_ctx, step, rk_step, dt, cfl), &
         'nonlinear_linearized')
  END SUBROUTINE nonlinear_linearized

END MODULE pass1_linearized
