!=====================================================================
! Replacement for MODULE transform (PFC/transform.f90): same name and
! PUBLIC list; the dealiased x-z transforms run on the GPU through
! libchanflow.so (one process, nproc = 1, so local_nx = nx/2+1,
! local_nz = nz2 and local_nz_small = nz).
!
! initialize_fft creates the library context (it is the first call of the
! reference's driver that needs the grid sizes, PFC/main_mpi.f90:33);
! the operators are uploaded by pass1::nonlinear's first call, after
! define_coordinates and define_parameters have filled them.
!=====================================================================
MODULE transform
  USE, INTRINSIC :: iso_c_binding
  USE runparms
  USE gridparms
  USE chanflow_f
  USE chanflow_sync
  IMPLICIT NONE
  PRIVATE
  PUBLIC :: initialize_fft,             finalize_fft,                  &
            xz_fft_real_to_complex,     xz_transform,                  &
            xz_fft_complex_to_real,     xz_fft_complex_to_real_padded, &
            dxz_transform,              dxz_transform_padded,          &
            small_fft_complex_to_real,  small_fft_complex_to_real_pad, &
            small_fft_real_to_complex,  compute_divergence_of_fluxes,  &
            output_fft_real_to_complex, output_fft_complex_to_real,    &
            FFTW_REAL_TO_COMPLEX,       FFTW_COMPLEX_TO_REAL

  INTEGER, PARAMETER :: FFTW_REAL_TO_COMPLEX = -1, FFTW_COMPLEX_TO_REAL = 1

CONTAINS

  SUBROUTINE initialize_fft
    CALL chanflow_create(nx, ny, nz, nx2, nz2, lx, lz, re, beta, fix_mass, dt_flag, &
         cfl_max, ctrlu, ctrlw, bl_grid)
  END SUBROUTINE initialize_fft

  SUBROUTINE finalize_fft
    CALL chanflow_destroy()
  END SUBROUTINE finalize_fft

  ! PFC/transform.f90:195-209
  SUBROUTINE xz_fft_real_to_complex(data, cdata)
    REAL(KIND=prec), INTENT(IN)     :: data(1:nx2,1:local_nz)
    COMPLEX(KIND=prec), INTENT(OUT) :: cdata(1:nz,1:local_nx)
    CALL chanflow_ensure_ops()
    CALL chanflow_check(chan_xz_fft_real_to_complex(chanflow_ctx, data, cdata, 1), &
         'xz_fft_real_to_complex')
  END SUBROUTINE xz_fft_real_to_complex

  ! PFC/transform.f90:213-227
  SUBROUTINE xz_fft_complex_to_real(cdata, data)
    COMPLEX(KIND=prec), INTENT(IN) :: cdata(1:nz,1:local_nx)
    REAL(KIND=prec), INTENT(OUT)   :: data(1:nx2,1:local_nz)
    CALL chanflow_ensure_ops()
    CALL chanflow_check(chan_xz_fft_complex_to_real(chanflow_ctx, cdata, data, 1), &
         'xz_fft_complex_to_real')
  END SUBROUTINE xz_fft_complex_to_real

  ! PFC/transform.f90:413-498: the transform plus one periodic ghost cell
  ! on each side in x and z.
  SUBROUTINE xz_fft_complex_to_real_padded(cdata, data)
    COMPLEX(KIND=prec), INTENT(IN) :: cdata(1:nz,1:local_nx)
    REAL(KIND=prec), INTENT(OUT)   :: data(0:nx2+1,0:local_nz+1)
    REAL(KIND=prec) :: tmp(1:nx2,1:nz2)
    CALL xz_fft_complex_to_real(cdata, tmp)
    CALL periodic_pad(tmp, nx2, nz2, data)
  END SUBROUTINE xz_fft_complex_to_real_padded

  SUBROUTINE periodic_pad(a, n1, n2, out)
    INTEGER, INTENT(IN)          :: n1, n2
    REAL(KIND=prec), INTENT(IN)  :: a(n1,n2)
    REAL(KIND=prec), INTENT(OUT) :: out(0:n1+1,0:n2+1)
    out(1:n1,1:n2) = a
    out(1:n1,0)    = a(:,n2)
    out(1:n1,n2+1) = a(:,1)
    out(0,:)       = out(n1,:)
    out(n1+1,:)    = out(1,:)
  END SUBROUTINE periodic_pad

  ! PFC/transform.f90:502-569: CU(nx/2+1,nz) <-> U(nx2,nz2), kx fastest.
  ! The reference's forward branch leaves CU(nx/2+1,nz/2+1) unassigned
  ! before doubling the last kx column (:556-561); that element is kept
  ! the same way here.
  SUBROUTINE XZ_TRANSFORM(CU, U, ISIGN)
    COMPLEX(KIND=prec) :: CU(NX/2+1,NZ)
    REAL(KIND=prec)    :: U(NX2,NZ2)
    INTEGER            :: ISIGN
    COMPLEX(KIND=prec) :: ct(NZ,NX/2+1), corner
    IF (ISIGN == FFTW_COMPLEX_TO_REAL) THEN
       ct = TRANSPOSE(CU)
       CALL xz_fft_complex_to_real(ct, U)
    ELSE
       corner = CU(NX/2+1,NZ/2+1)
       CALL xz_fft_real_to_complex(U, ct)
       CU = TRANSPOSE(ct)
       CU(NX/2+1,NZ/2+1) = 2.d0*corner
    END IF
  END SUBROUTINE XZ_TRANSFORM

  ! PFC/transform.f90:694-804: the field and its x and z derivatives.
  SUBROUTINE dxz_transform(cdata, data, datax, dataz)
    COMPLEX(KIND=prec), INTENT(IN) :: cdata(1:nz,1:local_nx)
    REAL(KIND=prec), INTENT(OUT), DIMENSION(1:nx2,1:local_nz) :: data, datax, dataz
    COMPLEX(KIND=prec) :: cd(1:nz,1:local_nx)
    INTEGER :: x, z
    CALL xz_fft_complex_to_real(cdata, data)
    DO x = 1,local_nx
       cd(:,x) = ikx_me(x)*cdata(:,x)
    END DO
    CALL xz_fft_complex_to_real(cd, datax)
    DO x = 1,local_nx
       DO z = 1,nz
          cd(z,x) = ikz(z)*cdata(z,x)
       END DO
    END DO
    CALL xz_fft_complex_to_real(cd, dataz)
  END SUBROUTINE dxz_transform

  ! PFC/transform.f90:808-946
  SUBROUTINE dxz_transform_padded(cdata, data, datax, dataz)
    COMPLEX(KIND=prec), INTENT(IN) :: cdata(1:nz,1:local_nx)
    REAL(KIND=prec), INTENT(OUT), DIMENSION(0:nx2+1,0:local_nz+1) :: data, datax, dataz
    REAL(KIND=prec), DIMENSION(1:nx2,1:nz2) :: a, ax, az
    CALL dxz_transform(cdata, a, ax, az)
    CALL periodic_pad(a, nx2, nz2, data)
    CALL periodic_pad(ax, nx2, nz2, datax)
    CALL periodic_pad(az, nx2, nz2, dataz)
  END SUBROUTINE dxz_transform_padded

  ! PFC/transform.f90:950-1000
  SUBROUTINE small_fft_complex_to_real(cdata, data)
    COMPLEX(KIND=prec), INTENT(IN) :: cdata(1:nz,1:local_nx)
    REAL(KIND=prec), INTENT(OUT)   :: data(1:nx,1:local_nz_small)
    CALL chanflow_ensure_ops()
    CALL chanflow_check(chan_small_fft_complex_to_real(chanflow_ctx, cdata, data, 1), &
         'small_fft_complex_to_real')
  END SUBROUTINE small_fft_complex_to_real

  ! PFC/transform.f90:1004-1067
  SUBROUTINE small_fft_complex_to_real_pad(cdata, data)
    COMPLEX(KIND=prec), INTENT(IN) :: cdata(1:nz,1:local_nx)
    REAL(KIND=prec), INTENT(OUT)   :: data(0:nx+1,0:local_nz_small+1)
    REAL(KIND=prec) :: tmp(1:nx,1:nz)
    CALL small_fft_complex_to_real(cdata, tmp)
    CALL periodic_pad(tmp, nx, nz, data)
  END SUBROUTINE small_fft_complex_to_real_pad

  ! PFC/transform.f90:1071-1142
  SUBROUTINE small_fft_real_to_complex(data, cdata)
    REAL(KIND=prec), INTENT(IN)     :: data(1:nx,1:local_nz_small)
    COMPLEX(KIND=prec), INTENT(OUT) :: cdata(1:nz,1:local_nx)
    CALL chanflow_ensure_ops()
    CALL chanflow_check(chan_small_fft_real_to_complex(chanflow_ctx, data, cdata, 1), &
         'small_fft_real_to_complex')
  END SUBROUTINE small_fft_real_to_complex

  ! PFC/transform.f90:231-254
  SUBROUTINE compute_divergence_of_fluxes(datax, datay, dataz, cdata)
    REAL(KIND=prec), INTENT(IN), DIMENSION(1:nx2,1:local_nz) :: datax, datay, dataz
    COMPLEX(KIND=prec), INTENT(OUT) :: cdata(1:nz,1:local_nx)
    CALL chanflow_ensure_ops()
    CALL chanflow_check(chan_compute_divergence_of_fluxes(chanflow_ctx, datax, datay, dataz, cdata, 1), &
         'compute_divergence_of_fluxes')
  END SUBROUTINE compute_divergence_of_fluxes

  ! PFC/transform.f90:1146-1166: in(nx,nz) -> cout(nx/2+1,nz), kx fastest.
  SUBROUTINE output_fft_real_to_complex(in, cout)
    REAL(KIND=prec), DIMENSION(nx,nz)        :: in
    COMPLEX(KIND=prec), DIMENSION(nx/2+1,nz) :: cout
    COMPLEX(KIND=prec) :: ct(nz,nx/2+1)
    CALL small_fft_real_to_complex(in, ct)
    cout = TRANSPOSE(ct)
  END SUBROUTINE output_fft_real_to_complex

  ! PFC/transform.f90:1170-1190
  SUBROUTINE output_fft_complex_to_real(cin, out)
    REAL(KIND=prec), DIMENSION(nx,nz)        :: out
    COMPLEX(KIND=prec), DIMENSION(nx/2+1,nz) :: cin
    COMPLEX(KIND=prec) :: ct(nz,nx/2+1)
    ct = TRANSPOSE(cin)
    CALL small_fft_complex_to_real(ct, out)
  END SUBROUTINE output_fft_complex_to_real

END MODULE transform
