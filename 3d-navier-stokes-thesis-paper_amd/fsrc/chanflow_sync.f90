!=====================================================================
! chanflow_sync: moves the reference's module state (PFC/data.f90,
! PFC/gridparms.f90, PFC/runparms.f90) between the host arrays and the
! library context.  Used by the replacement modules only.
!=====================================================================
MODULE chanflow_sync
  USE, INTRINSIC :: iso_c_binding
  USE runparms
  USE gridparms
  USE data
  USE chanflow_f
  IMPLICIT NONE
  PRIVATE
  PUBLIC :: chanflow_ensure_ops, chanflow_ensure_device, chanflow_sync_host, chanflow_host_reads_step

  LOGICAL, SAVE :: ops_uploaded = .FALSE.

CONTAINS

  ! The context and the operators, for the routines that work on host arrays
  ! passed to them (transform, solve): they neither read nor change the
  ! device copy of the module state.
  SUBROUTINE chanflow_ensure_ops()
    IF (.NOT. C_ASSOCIATED(chanflow_ctx)) &
         CALL chanflow_create(nx, ny, nz, nx2, nz2, lx, lz, re, beta, fix_mass, dt_flag, &
         cfl_max, ctrlu, ctrlw, bl_grid)
    IF (.NOT. ops_uploaded) THEN
       IF (nscalar /= 0) STOP 'chanflow: passive scalars are not provided (nscalar must be 0)'
       IF (nproc /= 1) STOP 'chanflow: run with one MPI rank (the GPU holds the whole field)'
       CALL chanflow_upload_operators(d2h, d2n, dm, dh, dn, hn, iu, dyh, dy, kx, kz, &
            buffer, target_inflow)
       ops_uploaded = .TRUE.
    END IF
  END SUBROUTINE chanflow_ensure_ops

  ! Before a routine that works on the module state (nonlinear, rhs_solve):
  ! the first such call of a run, or the first after the host arrays were
  ! declared newer by setting chanflow_on_device = .FALSE., sends the
  ! operators, the state and the scalars to the device.
  SUBROUTINE chanflow_ensure_device()
    CALL chanflow_ensure_ops()
    IF (.NOT. chanflow_on_device) THEN
       CALL chanflow_upload_state(data1, data2, data4)
       CALL chanflow_set_scalars(dt, t_total, p_grad, d_um_dy)
    END IF
  END SUBROUTINE chanflow_ensure_device

  ! Copy the newest state back into data1, data2, data4, p_grad and
  ! d_um_dy.  dt already comes back from every nonlinear call, and t_total
  ! is advanced by the driver on the host (PFC/main_mpi.f90:90).
  SUBROUTINE chanflow_sync_host()
    REAL(c_double) :: dt_dev, t_dev
    IF (.NOT. chanflow_on_device) RETURN
    CALL chanflow_download_state(data1, data2, data4)
    CALL chanflow_get_scalars(dt_dev, t_dev, p_grad, d_um_dy)
  END SUBROUTINE chanflow_sync_host

  ! True when the kept driver reads the host arrays after this step
  ! (PFC/main_mpi.f90:96-101 output, snapshot, save and l2 steps; the last step).
  LOGICAL FUNCTION chanflow_host_reads_step(s)
    INTEGER, INTENT(IN) :: s
    chanflow_host_reads_step = (s == tend)
    IF (output_step > 0) chanflow_host_reads_step = chanflow_host_reads_step .OR. MOD(s, output_step) == 0
    IF (save_step > 0)   chanflow_host_reads_step = chanflow_host_reads_step .OR. MOD(s, save_step) == 0
    IF (l2_step > 0)     chanflow_host_reads_step = chanflow_host_reads_step .OR. MOD(s, l2_step) == 0
  END FUNCTION chanflow_host_reads_step

END MODULE chanflow_sync
