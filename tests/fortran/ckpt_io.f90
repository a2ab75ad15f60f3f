! Writes and reads velocity checkpoints with the Fortran runtime's own
! sequential unformatted I/O, record for record as SAVE_VELOCITY and
! read_velocity do (PFC/read_write.f90:21-292): a header record
! NX, NY, NZ, LX, LZ, RE, DT, T_TOTAL, then slice(1:nz,1:ny+1,1:3) =
! data1(1:nz,x,1:ny+1,1:3) for x = 1..nx/2+1.
!   ckpt_io write <file> : writes a deterministic field
!   ckpt_io read  <file> : reads a file and prints the header and checksums
PROGRAM ckpt_io
  IMPLICIT NONE
  INTEGER, PARAMETER :: prec = SELECTED_REAL_KIND(14,32)
  INTEGER, PARAMETER :: nx = 8, ny = 5, nz = 6
  COMPLEX(KIND=prec) :: data1(nz, nx/2+1, ny+1, 3), slice(nz, ny+1, 3)
  INTEGER :: i, j, k, c, x, nxt, nyt, nzt
  REAL(KIND=prec) :: lx, lz, re, dt, t_total, s1, s2
  CHARACTER(LEN=16) :: mode
  CHARACTER(LEN=256) :: fname
  CALL GET_COMMAND_ARGUMENT(1, mode)
  CALL GET_COMMAND_ARGUMENT(2, fname)
  IF (TRIM(mode) == 'write') THEN
     DO c = 1,3
        DO j = 1,ny+1
           DO x = 1,nx/2+1
              DO k = 1,nz
                 data1(k,x,j,c) = CMPLX(1000*c + 100*j + 10*x + k, -0.5d0*(k + x + j + c), KIND=prec)
              END DO
           END DO
        END DO
     END DO
     lx = 6.283185307179586d0; lz = 3.141592653589793d0; re = 2500.d0; dt = 0.01d0; t_total = 1.25d0
     OPEN (UNIT=13, FILE=TRIM(fname), FORM='UNFORMATTED')
     WRITE(13) nx, ny, nz, lx, lz, re, dt, t_total
     DO x = 1,nx/2+1
        slice(1:nz,1:ny+1,1:3) = data1(1:nz,x,1:ny+1,1:3)
        WRITE(13) slice
     END DO
     CLOSE (UNIT=13, STATUS='KEEP')
  ELSE
     OPEN (UNIT=14, FILE=TRIM(fname), STATUS='OLD', FORM='UNFORMATTED')
     READ(14) nxt, nyt, nzt, lx, lz, re, dt, t_total
     IF (nxt /= nx .OR. nyt /= ny .OR. nzt /= nz) STOP 'INCONSISTENT PARAMETERS'
     DO x = 1,nx/2+1
        READ(14) slice
        data1(1:nz,x,1:ny+1,1:3) = slice(1:nz,1:ny+1,1:3)
     END DO
     CLOSE (UNIT=14)
     s1 = 0.d0; s2 = 0.d0
     i = 0
     DO c = 1,3
        DO j = 1,ny+1
           DO x = 1,nx/2+1
              DO k = 1,nz
                 i = i + 1
                 s1 = s1 + DBLE(i)*DBLE(data1(k,x,j,c))
                 s2 = s2 + DBLE(i)*AIMAG(data1(k,x,j,c))
              END DO
           END DO
        END DO
     END DO
     WRITE(*,'(3I6)') nxt, nyt, nzt
     WRITE(*,'(5ES25.16)') lx, lz, re, dt, t_total
     WRITE(*,'(2ES25.16)') s1, s2
  END IF
END PROGRAM ckpt_io
