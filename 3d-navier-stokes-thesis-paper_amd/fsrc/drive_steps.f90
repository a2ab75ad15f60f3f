! Drives the channel step from Fortran through the iso_c_binding layer in the
! order of the reference's timestep loop (PFC/main_mpi.f90:70-87): boundary,
! then three (nonlinear, rhs_solve) substeps per step, with data1 in the
! reference's layout data1(nz, nx/2+1, ny+1, 3).
!   drive_steps <in.bin> <out.bin> <nsteps> <dt>
! The grid is the 32 x 65 x 32 channel (3/2 dealiased to 48 x 48, box
! 2 pi x 2 x pi, Re 2500, stretch 1.5); in.bin and out.bin hold data1 as raw
! complex(8) values.
PROGRAM drive_steps
  USE, INTRINSIC :: iso_c_binding
  USE chanflow_f
  IMPLICIT NONE
  TYPE(chan_params) :: prm
  COMPLEX(c_double_complex), ALLOCATABLE, TARGET :: data1(:,:,:,:)
  REAL(c_double), ALLOCATABLE :: y(:), yh(:)
  REAL(c_double), TARGET :: dt, t0, pg
  REAL(c_double) :: cfl, pgo, dto
  INTEGER :: nsteps, step, rk, u
  CHARACTER(LEN=512) :: fin, fout, arg
  REAL(c_double), PARAMETER :: pi = 3.14159265358979323846d0

  CALL GET_COMMAND_ARGUMENT(1, fin)
  CALL GET_COMMAND_ARGUMENT(2, fout)
  CALL GET_COMMAND_ARGUMENT(3, arg)
  READ(arg, *) nsteps
  CALL GET_COMMAND_ARGUMENT(4, arg)
  READ(arg, *) dt

  prm%nx = 32; prm%ny = 65; prm%nz = 32; prm%nx2 = 48; prm%nz2 = 48
  prm%lx = 2.d0 * pi; prm%lz = pi; prm%re = 2500.d0; prm%beta = 0.d0
  prm%fix_mass = 1; prm%dt_flag = 0; prm%cfl_max = 0.95d0
  prm%ctrlu = 0.d0; prm%ctrlw = 0.d0; prm%bl_grid = 0
  CALL chanflow_check(chan_create(prm, chanflow_ctx), 'chan_create')
  ALLOCATE(y(prm%ny), yh(prm%ny + 1), data1(prm%nz, prm%nx / 2 + 1, prm%ny + 1, 3))
  CALL chanflow_check(chan_build_operators(chanflow_ctx, 1.5d0, y, yh), 'chan_build_operators')

  OPEN(NEWUNIT=u, FILE=TRIM(fin), ACCESS='STREAM', FORM='UNFORMATTED', STATUS='OLD')
  READ(u) data1
  CLOSE(u)
  CALL chanflow_check(chan_set_state(chanflow_ctx, C_LOC(data1), c_null_ptr, c_null_ptr), 'chan_set_state')
  t0 = 0.d0
  pg = -2.d0 / prm%re
  CALL chanflow_check(chan_set_scalars(chanflow_ctx, C_LOC(dt), C_LOC(t0), C_LOC(pg), c_null_ptr), 'chan_set_scalars')
  CALL chanflow_check(chan_initialize_flow(chanflow_ctx, 0), 'chan_initialize_flow')

  DO step = 1, nsteps
     CALL chanflow_check(chan_boundary(chanflow_ctx), 'boundary')
     DO rk = 1, 3
        dto = dt
        CALL chanflow_check(chan_nonlinear(chanflow_ctx, step, rk, dto, cfl), 'nonlinear')
        CALL chanflow_check(chan_rhs_solve(chanflow_ctx, step, rk, pgo), 'rhs_solve')
     END DO
  END DO

  CALL chanflow_check(chan_get_state(chanflow_ctx, C_LOC(data1), c_null_ptr, c_null_ptr), 'chan_get_state')
  OPEN(NEWUNIT=u, FILE=TRIM(fout), ACCESS='STREAM', FORM='UNFORMATTED', STATUS='REPLACE')
  WRITE(u) data1
  CLOSE(u)
  WRITE(*, '(A,I0,A,ES12.4)') 'steps=', nsteps, ' cfl=', cfl
  CALL chanflow_check(chan_destroy(chanflow_ctx), 'chan_destroy')
END PROGRAM drive_steps
