! Calls libchanflow.so through the iso_c_binding interface without a GPU:
! the version string and the parameter checks of chan_create, which read
! every field of chan_params (so a layout mismatch shows as a wrong message).
PROGRAM abi_check
  USE, INTRINSIC :: iso_c_binding
  USE chanflow_f
  IMPLICIT NONE
  TYPE(chan_params) :: prm
  TYPE(c_ptr)       :: ctx
  INTEGER(c_int)    :: rc
  WRITE(*,'(A,A)') 'version: ', TRIM(c_string(chan_version()))
  CALL reset()
  prm%nx = 31
  rc = chan_create(prm, ctx)
  WRITE(*,'(A,I0,A,A)') 'nx=31 rc=', rc, ' msg: ', TRIM(c_string(chan_last_error(c_null_ptr)))
  CALL reset()
  prm%nz2 = 50
  rc = chan_create(prm, ctx)
  WRITE(*,'(A,I0,A,A)') 'nz2=50 rc=', rc, ' msg: ', TRIM(c_string(chan_last_error(c_null_ptr)))
  CALL reset()
  prm%re = -1.d0
  rc = chan_create(prm, ctx)
  WRITE(*,'(A,I0,A,A)') 're=-1 rc=', rc, ' msg: ', TRIM(c_string(chan_last_error(c_null_ptr)))
  CALL reset()
  prm%bl_grid = 1
  rc = chan_create(prm, ctx)
  WRITE(*,'(A,I0,A,A)') 'bl_grid=1 rc=', rc, ' msg: ', TRIM(c_string(chan_last_error(c_null_ptr)))
CONTAINS
  SUBROUTINE reset()
    prm%nx = 32; prm%ny = 65; prm%nz = 32; prm%nx2 = 48; prm%nz2 = 48
    prm%lx = 6.28d0; prm%lz = 3.14d0; prm%re = 2500.d0; prm%beta = 0.d0
    prm%fix_mass = 1; prm%dt_flag = 0; prm%cfl_max = 0.95d0
    prm%ctrlu = 0.d0; prm%ctrlw = 0.d0; prm%bl_grid = 0
  END SUBROUTINE reset
END PROGRAM abi_check
