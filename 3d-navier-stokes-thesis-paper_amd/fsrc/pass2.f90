!=====================================================================
! Replacement for MODULE pass2 (PFC/pass2.f90): rhs_solve runs on the GPU
! (chan_rhs_solve).  The wall data cbc_top/cbc_bot written by the kept
! read_write::boundary are uploaded at rk_step 1; at the end of a step the
! host arrays are refreshed when the kept driver reads them, and with wall
! feedback on, the four planes boundary reads are refreshed every step.
!=====================================================================
MODULE pass2
  USE, INTRINSIC :: iso_c_binding
  USE runparms
  USE gridparms
  USE data
  USE chanflow_f
  USE chanflow_sync
  IMPLICIT NONE
  PRIVATE
  PUBLIC :: rhs_solve, jet_profile

CONTAINS

  ! PFC/pass2.f90:29-180.  With crossflow = 0 the wall data change only
  ! through the feedback law of read_write::boundary, so with both gains zero
  ! cbc_top/cbc_bot (zeroed by boundary every step) go to the device once.
  ! p_grad stays on the device between substeps and comes back with the
  ! state whenever the driver reads the host arrays.
  SUBROUTINE rhs_solve
    INTEGER :: c
    LOGICAL, SAVE :: wall_data_sent = .FALSE.
    IF (crossflow == 1) STOP 'chanflow: jets in crossflow are not provided (crossflow must be 0)'
    CALL chanflow_ensure_device()
    IF (rk_step == 1 .AND. (ctrlu /= 0.d0 .OR. ctrlw /= 0.d0 .OR. .NOT. wall_data_sent)) THEN
       CALL chanflow_check(chan_set_wall_bc(chanflow_ctx, cbc_top, cbc_bot), 'rhs_solve wall data')
       wall_data_sent = .TRUE.
    END IF
    CALL chanflow_check(chan_rhs_solve(chanflow_ctx, step, rk_step), 'rhs_solve')
    IF (rk_step /= rk_last) RETURN
    IF (chanflow_host_reads_step(step)) THEN
       CALL chanflow_sync_host()
    ELSE IF (ctrlu /= 0.d0 .OR. ctrlw /= 0.d0) THEN
       ! read_write::boundary reads planes 2, 3, ny-1, ny (PFC/read_write.f90:1475-1528)
       DO c = 1,3
          CALL chanflow_download_planes(c, 2, 3, data1(1,1,2,c))
          CALL chanflow_download_planes(c, ny-1, ny, data1(1,1,ny-1,c))
       END DO
    END IF
  END SUBROUTINE rhs_solve

  ! PFC/pass2.f90:423-: jets are outside this library.
  SUBROUTINE jet_profile(time, time_incr, ctop, cbot)
    REAL(KIND=prec), INTENT(IN)       :: time, time_incr
    COMPLEX(KIND=prec), INTENT(INOUT) :: ctop(nz,local_nx), cbot(nz,local_nx)
    STOP 'chanflow: jet_profile is not provided (crossflow must be 0)'
  END SUBROUTINE jet_profile

END MODULE pass2
