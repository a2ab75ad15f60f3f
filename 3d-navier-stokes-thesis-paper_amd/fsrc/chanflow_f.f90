!=====================================================================
! chanflow_f: iso_c_binding interface to libchanflow.so (include/chanflow.h)
! and the context shared by the replacement modules transform, pass1,
! pass1_linearized, pass2 and solve.
!
! Every C entry point returns 0 on success; chanflow_check prints the
! library's message and ERROR STOPs (a nonzero exit status), as the reference
! STOPs on its own errors (e.g. PFC/pass1.f90:243).
!
! State residency: after the first nonlinear call the newest data1, data2
! and data4 live on the GPU.  chanflow_upload_state copies the host arrays
! to the device; chanflow_download_state copies them back.  The replacement
! pass2::rhs_solve downloads at the end of a step whenever the kept driver
! reads the host arrays (output, save, l2 and final steps) and, when wall
! feedback is on, downloads the planes read_write::boundary needs.
!=====================================================================
MODULE chanflow_f
  USE, INTRINSIC :: iso_c_binding
  IMPLICIT NONE
  PRIVATE
  PUBLIC :: chan_params, chan_operators, chanflow_ctx, chanflow_check, &
            chanflow_create, chanflow_upload_operators, chanflow_upload_state, &
            chanflow_download_state, chanflow_download_data1, chanflow_download_planes, &
            chanflow_on_device, &
            chanflow_set_scalars, chanflow_get_scalars, chanflow_destroy, &
            chan_version, chan_nonlinear, chan_nonlinear_linearized, chan_set_baseflow, &
            chan_rhs_solve, chan_set_wall_bc, chan_xz_fft_complex_to_real, &
            chan_xz_fft_real_to_complex, chan_compute_divergence_of_fluxes, &
            chan_small_fft_complex_to_real, chan_small_fft_real_to_complex, &
            chan_solve_helmholtz, chan_solve_p, chan_linearized_matvec, c_string, &
            chan_strip_divergence_modes, chan_subtract_pressure_gradient_modes, &
            chan_create, chan_destroy, chan_last_error, chan_set_state, chan_get_state, &
            chan_set_scalars, chan_build_operators, chan_initialize_flow, chan_boundary

  TYPE, BIND(C) :: chan_params
     INTEGER(c_int) :: nx, ny, nz, nx2, nz2
     REAL(c_double) :: lx, lz, re, beta
     INTEGER(c_int) :: fix_mass, dt_flag
     REAL(c_double) :: cfl_max, ctrlu, ctrlw
     INTEGER(c_int) :: bl_grid
  END TYPE chan_params

  TYPE, BIND(C) :: chan_operators
     TYPE(c_ptr) :: d2h, d2n, dm, dh, dn, hn, iu, dyh, dy, kx, kz, buffer, target_inflow
  END TYPE chan_operators

  TYPE(c_ptr), SAVE :: chanflow_ctx = c_null_ptr
  LOGICAL, SAVE     :: chanflow_on_device = .FALSE.

  INTERFACE
     INTEGER(c_int) FUNCTION chan_create(prm, ctx) BIND(C, name='chan_create')
       IMPORT :: c_int, c_ptr, chan_params
       TYPE(chan_params), INTENT(IN) :: prm
       TYPE(c_ptr), INTENT(OUT)      :: ctx
     END FUNCTION chan_create
     INTEGER(c_int) FUNCTION chan_destroy(ctx) BIND(C, name='chan_destroy')
       IMPORT :: c_int, c_ptr
       TYPE(c_ptr), VALUE :: ctx
     END FUNCTION chan_destroy
     TYPE(c_ptr) FUNCTION chan_last_error(ctx) BIND(C, name='chan_last_error')
       IMPORT :: c_ptr
       TYPE(c_ptr), VALUE :: ctx
     END FUNCTION chan_last_error
     TYPE(c_ptr) FUNCTION chan_version() BIND(C, name='chan_version')
       IMPORT :: c_ptr
     END FUNCTION chan_version
     INTEGER(c_int) FUNCTION chan_upload_operators(ctx, ops) BIND(C, name='chan_upload_operators')
       IMPORT :: c_int, c_ptr, chan_operators
       TYPE(c_ptr), VALUE               :: ctx
       TYPE(chan_operators), INTENT(IN) :: ops
     END FUNCTION chan_upload_operators
     INTEGER(c_int) FUNCTION chan_build_operators(ctx, stretch, y, yh) BIND(C, name='chan_build_operators')
       IMPORT :: c_int, c_ptr, c_double
       TYPE(c_ptr), VALUE     :: ctx
       REAL(c_double), VALUE  :: stretch
       REAL(c_double)         :: y(*), yh(*)
     END FUNCTION chan_build_operators
     INTEGER(c_int) FUNCTION chan_initialize_flow(ctx, linearized) BIND(C, name='chan_initialize_flow')
       IMPORT :: c_int, c_ptr
       TYPE(c_ptr), VALUE    :: ctx
       INTEGER(c_int), VALUE :: linearized
     END FUNCTION chan_initialize_flow
     INTEGER(c_int) FUNCTION chan_boundary(ctx) BIND(C, name='chan_boundary')
       IMPORT :: c_int, c_ptr
       TYPE(c_ptr), VALUE :: ctx
     END FUNCTION chan_boundary
     INTEGER(c_int) FUNCTION chan_set_state(ctx, d1, d2, d4) BIND(C, name='chan_set_state')
       IMPORT :: c_int, c_ptr
       TYPE(c_ptr), VALUE :: ctx, d1, d2, d4
     END FUNCTION chan_set_state
     INTEGER(c_int) FUNCTION chan_get_state(ctx, d1, d2, d4) BIND(C, name='chan_get_state')
       IMPORT :: c_int, c_ptr
       TYPE(c_ptr), VALUE :: ctx, d1, d2, d4
     END FUNCTION chan_get_state
     INTEGER(c_int) FUNCTION chan_get_planes(ctx, comp, k0, k1, out) BIND(C, name='chan_get_planes')
       IMPORT :: c_int, c_ptr
       TYPE(c_ptr), VALUE    :: ctx, out
       INTEGER(c_int), VALUE :: comp, k0, k1
     END FUNCTION chan_get_planes
     INTEGER(c_int) FUNCTION chan_set_scalars(ctx, dt, t_total, p_grad, d_um_dy) &
          BIND(C, name='chan_set_scalars')
       IMPORT :: c_int, c_ptr
       TYPE(c_ptr), VALUE :: ctx, dt, t_total, p_grad, d_um_dy
     END FUNCTION chan_set_scalars
     INTEGER(c_int) FUNCTION chan_get_scalars(ctx, dt, t_total, p_grad, d_um_dy) &
          BIND(C, name='chan_get_scalars')
       IMPORT :: c_int, c_ptr
       TYPE(c_ptr), VALUE :: ctx, dt, t_total, p_grad, d_um_dy
     END FUNCTION chan_get_scalars
     ! data arrays are passed by reference (assumed size), scalars by value
     INTEGER(c_int) FUNCTION chan_nonlinear(ctx, step, rk_step, dt, cfl) BIND(C, name='chan_nonlinear')
       IMPORT :: c_int, c_ptr, c_double
       TYPE(c_ptr), VALUE    :: ctx
       INTEGER(c_int), VALUE :: step, rk_step
       REAL(c_double)        :: dt, cfl
     END FUNCTION chan_nonlinear
     INTEGER(c_int) FUNCTION chan_nonlinear_linearized(ctx, step, rk_step, dt, cfl) &
          BIND(C, name='chan_nonlinear_linearized')
       IMPORT :: c_int, c_ptr, c_double
       TYPE(c_ptr), VALUE    :: ctx
       INTEGER(c_int), VALUE :: step, rk_step
       REAL(c_double)        :: dt, cfl
     END FUNCTION chan_nonlinear_linearized
     INTEGER(c_int) FUNCTION chan_set_baseflow(ctx, baseflow) BIND(C, name='chan_set_baseflow')
       IMPORT :: c_int, c_ptr, c_double
       TYPE(c_ptr), VALUE         :: ctx
       REAL(c_double), INTENT(IN) :: baseflow(*)
     END FUNCTION chan_set_baseflow
     INTEGER(c_int) FUNCTION chan_rhs_solve(ctx, step, rk_step, p_grad) BIND(C, name='chan_rhs_solve')
       IMPORT :: c_int, c_ptr, c_double
       TYPE(c_ptr), VALUE       :: ctx
       INTEGER(c_int), VALUE    :: step, rk_step
       REAL(c_double), OPTIONAL :: p_grad
     END FUNCTION chan_rhs_solve
     INTEGER(c_int) FUNCTION chan_set_wall_bc(ctx, cbc_top, cbc_bot) BIND(C, name='chan_set_wall_bc')
       IMPORT :: c_int, c_ptr, c_double_complex
       TYPE(c_ptr), VALUE                    :: ctx
       COMPLEX(c_double_complex), INTENT(IN) :: cbc_top(*), cbc_bot(*)
     END FUNCTION chan_set_wall_bc
     INTEGER(c_int) FUNCTION chan_xz_fft_complex_to_real(ctx, cdata, data, nplanes) &
          BIND(C, name='chan_xz_fft_complex_to_real')
       IMPORT :: c_int, c_ptr, c_double, c_double_complex
       TYPE(c_ptr), VALUE                    :: ctx
       COMPLEX(c_double_complex), INTENT(IN) :: cdata(*)
       REAL(c_double), INTENT(OUT)           :: data(*)
       INTEGER(c_int), VALUE                 :: nplanes
     END FUNCTION chan_xz_fft_complex_to_real
     INTEGER(c_int) FUNCTION chan_xz_fft_real_to_complex(ctx, data, cdata, nplanes) &
          BIND(C, name='chan_xz_fft_real_to_complex')
       IMPORT :: c_int, c_ptr, c_double, c_double_complex
       TYPE(c_ptr), VALUE                     :: ctx
       REAL(c_double), INTENT(IN)             :: data(*)
       COMPLEX(c_double_complex), INTENT(OUT) :: cdata(*)
       INTEGER(c_int), VALUE                  :: nplanes
     END FUNCTION chan_xz_fft_real_to_complex
     INTEGER(c_int) FUNCTION chan_compute_divergence_of_fluxes(ctx, datax, datay, dataz, cdata, nplanes) &
          BIND(C, name='chan_compute_divergence_of_fluxes')
       IMPORT :: c_int, c_ptr, c_double, c_double_complex
       TYPE(c_ptr), VALUE                     :: ctx
       REAL(c_double), INTENT(IN)             :: datax(*), datay(*), dataz(*)
       COMPLEX(c_double_complex), INTENT(OUT) :: cdata(*)
       INTEGER(c_int), VALUE                  :: nplanes
     END FUNCTION chan_compute_divergence_of_fluxes
     INTEGER(c_int) FUNCTION chan_small_fft_complex_to_real(ctx, cdata, data, nplanes) &
          BIND(C, name='chan_small_fft_complex_to_real')
       IMPORT :: c_int, c_ptr, c_double, c_double_complex
       TYPE(c_ptr), VALUE                    :: ctx
       COMPLEX(c_double_complex), INTENT(IN) :: cdata(*)
       REAL(c_double), INTENT(OUT)           :: data(*)
       INTEGER(c_int), VALUE                 :: nplanes
     END FUNCTION chan_small_fft_complex_to_real
     INTEGER(c_int) FUNCTION chan_small_fft_real_to_complex(ctx, data, cdata, nplanes) &
          BIND(C, name='chan_small_fft_real_to_complex')
       IMPORT :: c_int, c_ptr, c_double, c_double_complex
       TYPE(c_ptr), VALUE                     :: ctx
       REAL(c_double), INTENT(IN)             :: data(*)
       COMPLEX(c_double_complex), INTENT(OUT) :: cdata(*)
       INTEGER(c_int), VALUE                  :: nplanes
     END FUNCTION chan_small_fft_real_to_complex
     INTEGER(c_int) FUNCTION chan_solve_helmholtz(ctx, n, nsys, f, alpha, diff, d2, topbc, botbc, &
          top_neumann, bot_neumann) BIND(C, name='chan_solve_helmholtz')
       IMPORT :: c_int, c_ptr, c_double, c_double_complex
       TYPE(c_ptr), VALUE                       :: ctx
       INTEGER(c_int), VALUE                    :: n, nsys, top_neumann, bot_neumann
       COMPLEX(c_double_complex), INTENT(INOUT) :: f(*)
       REAL(c_double), INTENT(IN)               :: alpha(*), d2(*)
       REAL(c_double), VALUE                    :: diff
       COMPLEX(c_double_complex), INTENT(IN)    :: topbc(*), botbc(*)
     END FUNCTION chan_solve_helmholtz
     INTEGER(c_int) FUNCTION chan_solve_p(ctx, nsys, p, alpha) BIND(C, name='chan_solve_p')
       IMPORT :: c_int, c_ptr, c_double, c_double_complex
       TYPE(c_ptr), VALUE                       :: ctx
       INTEGER(c_int), VALUE                    :: nsys
       COMPLEX(c_double_complex), INTENT(INOUT) :: p(*)
       REAL(c_double), INTENT(IN)               :: alpha(*)
     END FUNCTION chan_solve_p
     INTEGER(c_int) FUNCTION chan_strip_divergence_modes(ctx, nsys, cu, cp, coef, p, q) &
          BIND(C, name='chan_strip_divergence_modes')
       IMPORT :: c_int, c_ptr, c_double, c_double_complex
       TYPE(c_ptr), VALUE                       :: ctx
       INTEGER(c_int), VALUE                    :: nsys
       COMPLEX(c_double_complex), INTENT(INOUT) :: cu(*)
       COMPLEX(c_double_complex), INTENT(OUT)   :: cp(*)
       REAL(c_double), VALUE                    :: coef
       INTEGER(c_int), INTENT(IN)               :: p(*), q(*)
     END FUNCTION chan_strip_divergence_modes
     INTEGER(c_int) FUNCTION chan_subtract_pressure_gradient_modes(ctx, nsys, crhs, cpress, coef, p, q) &
          BIND(C, name='chan_subtract_pressure_gradient_modes')
       IMPORT :: c_int, c_ptr, c_double, c_double_complex
       TYPE(c_ptr), VALUE                       :: ctx
       INTEGER(c_int), VALUE                    :: nsys
       COMPLEX(c_double_complex), INTENT(INOUT) :: crhs(*)
       COMPLEX(c_double_complex), INTENT(IN)    :: cpress(*)
       REAL(c_double), VALUE                    :: coef
       INTEGER(c_int), INTENT(IN)               :: p(*), q(*)
     END FUNCTION chan_subtract_pressure_gradient_modes
     INTEGER(c_int) FUNCTION chan_linearized_matvec(ctx, v, w, tend, t_total) &
          BIND(C, name='chan_linearized_matvec')
       IMPORT :: c_int, c_ptr, c_double
       TYPE(c_ptr), VALUE          :: ctx
       REAL(c_double), INTENT(IN)  :: v(*)
       REAL(c_double), INTENT(OUT) :: w(*)
       INTEGER(c_int), VALUE       :: tend
       REAL(c_double)              :: t_total
     END FUNCTION chan_linearized_matvec
  END INTERFACE

CONTAINS

  ! STOP with the library's message when rc /= 0.
  SUBROUTINE chanflow_check(rc, where)
    INTEGER(c_int), INTENT(IN)   :: rc
    CHARACTER(LEN=*), INTENT(IN) :: where
    IF (rc /= 0) THEN
       WRITE(*,*) 'chanflow: ', where, ': ', c_string(chan_last_error(chanflow_ctx))
       ERROR STOP 'CHANFLOW ERROR'
    END IF
  END SUBROUTINE chanflow_check

  FUNCTION c_string(p) RESULT(s)
    TYPE(c_ptr), INTENT(IN) :: p
    CHARACTER(LEN=256)      :: s
    CHARACTER(KIND=c_char), POINTER :: cs(:)
    INTEGER :: i
    s = ' '
    IF (.NOT. C_ASSOCIATED(p)) RETURN
    CALL C_F_POINTER(p, cs, [256])
    DO i = 1, 256
       IF (cs(i) == c_null_char) EXIT
       s(i:i) = cs(i)
    END DO
  END FUNCTION c_string

  ! Create the context from the run parameters (PFC/runparms.f90, PFC/gridparms.f90).
  SUBROUTINE chanflow_create(nx, ny, nz, nx2, nz2, lx, lz, re, beta, fix_mass, dt_flag, &
       cfl_max, ctrlu, ctrlw, bl_grid)
    INTEGER, INTENT(IN)         :: nx, ny, nz, nx2, nz2, fix_mass, dt_flag, bl_grid
    REAL(c_double), INTENT(IN)  :: lx, lz, re, beta, cfl_max, ctrlu, ctrlw
    TYPE(chan_params) :: prm
    IF (C_ASSOCIATED(chanflow_ctx)) RETURN
    prm%nx = nx; prm%ny = ny; prm%nz = nz; prm%nx2 = nx2; prm%nz2 = nz2
    prm%lx = lx; prm%lz = lz; prm%re = re; prm%beta = beta
    prm%fix_mass = fix_mass; prm%dt_flag = dt_flag
    prm%cfl_max = cfl_max; prm%ctrlu = ctrlu; prm%ctrlw = ctrlw; prm%bl_grid = bl_grid
    CALL chanflow_check(chan_create(prm, chanflow_ctx), 'chan_create')
    chanflow_on_device = .FALSE.
  END SUBROUTINE chanflow_create

  SUBROUTINE chanflow_destroy()
    INTEGER(c_int) :: rc
    IF (.NOT. C_ASSOCIATED(chanflow_ctx)) RETURN
    rc = chan_destroy(chanflow_ctx)
    chanflow_ctx = c_null_ptr
    chanflow_on_device = .FALSE.
  END SUBROUTINE chanflow_destroy

  ! Upload the y-operators and wavenumbers verbatim from the Fortran arrays
  ! (PFC/setup.f90:541-754, 1440-1466).  The dummies carry TARGET so that
  ! their addresses can be taken for the duration of the call.
  SUBROUTINE chanflow_upload_operators(d2h, d2n, dm, dh, dn, hn, iu, dyh, dy, kx, kz, &
       buffer, target_inflow)
    REAL(c_double), INTENT(IN), TARGET :: d2h(*), d2n(*), dm(*), dh(*), dn(*), hn(*), iu(*), &
         dyh(*), dy(*), kx(*), kz(*), buffer(*), target_inflow(*)
    TYPE(chan_operators) :: o
    o%d2h = C_LOC(d2h); o%d2n = C_LOC(d2n); o%dm = C_LOC(dm); o%dh = C_LOC(dh)
    o%dn = C_LOC(dn); o%hn = C_LOC(hn); o%iu = C_LOC(iu); o%dyh = C_LOC(dyh)
    o%dy = C_LOC(dy); o%kx = C_LOC(kx); o%kz = C_LOC(kz); o%buffer = C_LOC(buffer)
    o%target_inflow = C_LOC(target_inflow)
    CALL chanflow_check(chan_upload_operators(chanflow_ctx, o), 'chan_upload_operators')
  END SUBROUTINE chanflow_upload_operators

  ! data1 complex(nz,nx/2+1,ny+1,3), data2 complex(ny+1,3,nz,nx/2+1),
  ! data4 complex(ny+1,nz,nx/2+1) (PFC/setup.f90:447-449).
  SUBROUTINE chanflow_upload_state(data1, data2, data4)
    COMPLEX(c_double_complex), INTENT(IN), TARGET :: data1(*), data2(*), data4(*)
    CALL chanflow_check(chan_set_state(chanflow_ctx, C_LOC(data1), C_LOC(data2), C_LOC(data4)), &
         'chan_set_state')
    chanflow_on_device = .TRUE.
  END SUBROUTINE chanflow_upload_state

  ! data1 alone (the NL term between nonlinear and rhs_solve, else the velocity)
  SUBROUTINE chanflow_download_data1(data1)
    COMPLEX(c_double_complex), INTENT(INOUT), TARGET :: data1(*)
    CALL chanflow_check(chan_get_state(chanflow_ctx, C_LOC(data1), c_null_ptr, c_null_ptr), &
         'chan_get_state')
  END SUBROUTINE chanflow_download_data1

  SUBROUTINE chanflow_download_state(data1, data2, data4)
    COMPLEX(c_double_complex), INTENT(INOUT), TARGET :: data1(*), data2(*), data4(*)
    CALL chanflow_check(chan_get_state(chanflow_ctx, C_LOC(data1), C_LOC(data2), C_LOC(data4)), &
         'chan_get_state')
  END SUBROUTINE chanflow_download_state

  ! planes k0..k1 (1-based) of component comp (1..3) of the device velocity
  ! into data1(:,:,k0:k1,comp); out points at data1(1,1,k0,comp).
  SUBROUTINE chanflow_download_planes(comp, k0, k1, out)
    INTEGER, INTENT(IN) :: comp, k0, k1
    COMPLEX(c_double_complex), INTENT(INOUT), TARGET :: out(*)
    CALL chanflow_check(chan_get_planes(chanflow_ctx, comp - 1, k0 - 1, k1 - 1, C_LOC(out)), &
         'chan_get_planes')
  END SUBROUTINE chanflow_download_planes

  SUBROUTINE chanflow_set_scalars(dt, t_total, p_grad, d_um_dy)
    REAL(c_double), INTENT(IN), TARGET :: dt, t_total, p_grad, d_um_dy(*)
    CALL chanflow_check(chan_set_scalars(chanflow_ctx, C_LOC(dt), C_LOC(t_total), C_LOC(p_grad), &
         C_LOC(d_um_dy)), 'chan_set_scalars')
  END SUBROUTINE chanflow_set_scalars

  SUBROUTINE chanflow_get_scalars(dt, t_total, p_grad, d_um_dy)
    REAL(c_double), INTENT(INOUT), TARGET :: dt, t_total, p_grad, d_um_dy(*)
    CALL chanflow_check(chan_get_scalars(chanflow_ctx, C_LOC(dt), C_LOC(t_total), C_LOC(p_grad), &
         C_LOC(d_um_dy)), 'chan_get_scalars')
  END SUBROUTINE chanflow_get_scalars

END MODULE chanflow_f
