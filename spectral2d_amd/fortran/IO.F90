!> @file IO.F90
!> @brief IO_m with the public interface of the reference's module
!! (src/IO.F90:39-63): WriteRestart, WriteGridHDF5 and FileName.
!!
!! The physical fields are formed through Spectral_m (libs2d on the GPU);
!! the HDF5 and Xdmf files are written by one process through libs2d_h5
!! (include/s2d_h5.h), so the layouts are those of the reference's files
!! with a single hyperslab per dataset.
MODULE IO_m

   USE ISO_C_BINDING
   USE ISO_FORTRAN_ENV,ONLY: OUTPUT_UNIT
   USE Parameters_m,ONLY: IWPF, RWPF, RWPC, CWPC
   USE S2D_C
   USE Spectral_m,ONLY: ComputeVelocity, TransformC2R, TransformR2C

   IMPLICIT NONE

   !> Output type of WriteRestart (src/IO.F90:39).
   INTEGER(KIND=IWPF),PARAMETER,PUBLIC :: HDF5_OUTPUT = 1_IWPF
   !> Length of file names (src/IO.F90:42).
   INTEGER(KIND=IWPF),PARAMETER,PUBLIC :: FILE_NAME_LENGTH = 256_IWPF

   !> Number of the next restart file (REST_00000000 is the first).
   INTEGER(KIND=IWPF),PRIVATE :: restNum = 0_IWPF

   INTERFACE FileName
      MODULE PROCEDURE FileNameRoot, FileNameNumbered
   END INTERFACE FileName

   PUBLIC :: WriteRestart, WriteGridHDF5, FileName, RestartCount
   PRIVATE :: FileNameRoot, FileNameNumbered, CStr, H5Check

CONTAINS

   !> WriteRestart, src/IO.F90:98-143: the velocity by ComputeVelocity with
   !! the inverse transform and omega, psi by TransformC2R into the real
   !! views; REST_<n>.h5 and its .xmf; then omega and psi transformed back,
   !! so wC and psiC leave as FFT(IFFT(.)) of what came in, as in the
   !! reference. uC and vC are left holding the physical velocity.
   SUBROUTINE WriteRestart(nxG, nyG, nxP, nyP, i1, i2, j1, j2, rISize, cISize, &
                           kxP, kyP, rank, nadv, time, nu, outType, &
                           uC, uR, vC, vR, wC, wR, psiC, psiR)
      INTEGER(KIND=IWPF),INTENT(IN) :: nxG, nyG, nxP, nyP, i1, i2, j1, j2
      INTEGER(KIND=IWPF),INTENT(IN) :: rISize, cISize, rank, nadv, outType
      INTEGER(KIND=IWPF),DIMENSION(cISize),INTENT(IN) :: kxP
      INTEGER(KIND=IWPF),DIMENSION(nyP),INTENT(IN) :: kyP
      REAL(KIND=RWPC),INTENT(IN) :: time, nu
      COMPLEX(KIND=CWPC),DIMENSION(cISize,nyP),INTENT(INOUT) :: uC, vC, wC, psiC
      REAL(KIND=RWPC),DIMENSION(rISize,nyP),INTENT(INOUT) :: uR, vR, wR, psiR
      CHARACTER(LEN=FILE_NAME_LENGTH) :: fname

      CALL ComputeVelocity(nxP, nyP, rISize, cISize, kxP, kyP, uC, uR, vC, vR, psiC, .TRUE.)
      CALL TransformC2R(rISize, cISize, nyP, wC, wR)
      CALL TransformC2R(rISize, cISize, nyP, psiC, psiR)
      IF (rank == 0_IWPF) THEN
         WRITE(OUTPUT_UNIT,'(A,I6.6,A,ES15.8,A,I8.8)') 'Writing REST_', restNum, &
            '.h5 at t = ', time, ' and nadv = ', nadv
      END IF
      IF (outType == HDF5_OUTPUT) THEN
         CALL FileName('REST', restNum, 'h5', fname)
         CALL H5Check(s2d_h5_write_rest(CStr(fname), INT(nxG, C_INT), INT(nyG, C_INT), &
                                        INT(rISize, C_INT), INT(nadv, C_INT), &
                                        REAL(time, C_DOUBLE), REAL(nu, C_DOUBLE), &
                                        wR, psiR, uR, vR), fname)
         CALL FileName('REST', restNum, 'xmf', fname)
         CALL H5Check(s2d_h5_write_xmf(CStr(fname), INT(nxG, C_INT), INT(nyG, C_INT), &
                                       INT(restNum, C_INT)), fname)
      END IF
      restNum = restNum + 1_IWPF
      CALL TransformR2C(nxG, nyG, nyP, rISize, cISize, wR, wC)
      CALL TransformR2C(nxG, nyG, nyP, rISize, cISize, psiR, psiC)
   END SUBROUTINE WriteRestart

   !> WriteGridHDF5, src/IO.F90:158-285 (one process writes the whole grid;
   !! comm and the block bounds are accepted for the interface).
   SUBROUTINE WriteGridHDF5(comm, rank, fName, nxG, nyG, nxP, nyP, i1, i2, j1, j2)
      INTEGER(KIND=IWPF),INTENT(IN) :: comm, rank
      CHARACTER(LEN=*),INTENT(IN) :: fName
      INTEGER(KIND=IWPF),INTENT(IN) :: nxG, nyG, nxP, nyP, i1, i2, j1, j2
      IF (rank /= 0_IWPF) RETURN
      CALL H5Check(s2d_h5_write_grid(CStr(fName), INT(nxG, C_INT), INT(nyG, C_INT)), fName)
   END SUBROUTINE WriteGridHDF5

   !> Number of restart files written so far.
   INTEGER(KIND=IWPF) FUNCTION RestartCount()
      RestartCount = restNum
   END FUNCTION RestartCount

   !> root.suffix (src/IO.F90:657-666).
   SUBROUTINE FileNameRoot(root, suffix, fname)
      CHARACTER(LEN=*),INTENT(IN) :: root, suffix
      CHARACTER(LEN=FILE_NAME_LENGTH),INTENT(OUT) :: fname
      fname = TRIM(root)//'.'//TRIM(suffix)
   END SUBROUTINE FileNameRoot

   !> root_########.suffix (src/IO.F90:674-684).
   SUBROUTINE FileNameNumbered(root, num, suffix, fname)
      CHARACTER(LEN=*),INTENT(IN) :: root, suffix
      INTEGER(KIND=IWPF),INTENT(IN) :: num
      CHARACTER(LEN=FILE_NAME_LENGTH),INTENT(OUT) :: fname
      CHARACTER(LEN=8) :: digits
      WRITE(digits,'(I8.8)') num
      fname = TRIM(root)//'_'//digits//'.'//TRIM(suffix)
   END SUBROUTINE FileNameNumbered

   !> A NUL-terminated copy of a Fortran string.
   FUNCTION CStr(s) RESULT(c)
      CHARACTER(LEN=*),INTENT(IN) :: s
      CHARACTER(KIND=C_CHAR),DIMENSION(LEN_TRIM(s)+1) :: c
      INTEGER :: i
      DO i = 1, LEN_TRIM(s)
         c(i) = s(i:i)
      END DO
      c(LEN_TRIM(s)+1) = C_NULL_CHAR
   END FUNCTION CStr

   SUBROUTINE H5Check(rc, what)
      INTEGER(C_INT),INTENT(IN) :: rc
      CHARACTER(LEN=*),INTENT(IN) :: what
      IF (rc == 0_C_INT) RETURN
      WRITE(*,'(A,A,A,I0,A)') 'libs2d_h5: writing ', TRIM(what), ' failed (status ', rc, ')'
      ERROR STOP 1
   END SUBROUTINE H5Check

END MODULE IO_m
