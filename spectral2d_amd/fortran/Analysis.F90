!> @file Analysis.F90
!> @brief Analysis_m with the public interface of the reference's module
!! (src/Analysis.F90:28-33): SetupAnalysis, Analysis, ComputeSpectrum.
!!
!! The derivatives and transforms go through Spectral_m (libs2d on the
!! GPU); the shell spectrum is binned on the GPU in FP64
!! (s2d_spectrum_uv); the skewness sums run on the host over the physical
!! fields, as the reference forms them. Output files: <root>.DAT (one line
!! per call) and SPECTRUM_########.DAT (src/Analysis.F90:60-68, 244-245,
!! 423-433).
MODULE Analysis_m

   USE ISO_C_BINDING
   USE ISO_FORTRAN_ENV,ONLY: OUTPUT_UNIT
   USE Parameters_m,ONLY: IWPF, RWPF, RWPC, CWPC
   USE S2D_C
   USE Spectral_m,ONLY: s2dCtx, ComputeVelocity, DuDx, DuDy, TransformC2R
   USE IO_m,ONLY: FILE_NAME_LENGTH, FileName

   IMPLICIT NONE

   !> Name of the Eulerian statistics file.
   CHARACTER(LEN=FILE_NAME_LENGTH),PRIVATE :: eulOut = ''

   PUBLIC :: SetupAnalysis, Analysis, ComputeSpectrum
   PRIVATE :: Skewness

CONTAINS

   !> SetupAnalysis, src/Analysis.F90:40-68: <fRoot>.DAT with its header.
   SUBROUTINE SetupAnalysis(fRoot)
      CHARACTER(LEN=*),INTENT(IN) :: fRoot
      INTEGER :: u
      CALL FileName(fRoot, 'DAT', eulOut)
      OPEN(NEWUNIT=u, FILE=TRIM(eulOut), STATUS='REPLACE', ACTION='WRITE', FORM='FORMATTED')
      WRITE(u,'(A,T3,A,T17,A,T33,A,T49,A,T65,A,T81,A,T97,A)') '#', 'NADV', 'TIME', 'TKE', &
         'EPS', 'EnstDiss', 'xSkew', 'ySkew'
      CLOSE(u)
   END SUBROUTINE SetupAnalysis

   !> Analysis, src/Analysis.F90:90-242: the x and y derivative skewness
   !! factors, the spectrum (written to SPECTRUM_<nadv>.DAT) with its TKE,
   !! dissipation and enstrophy dissipation, and one line of the Eulerian
   !! file. uC, vC (and their real views) are scratch, as in the reference;
   !! wC and psiC are left unchanged.
   SUBROUTINE Analysis(rank, nadv, time, nu, nxG, nyG, nxP, nyP, &
                       rISize, cISize, kxP, kyP, kxLG, kyLG, kxMG, kyMG, &
                       uC, uR, vC, vR, wC, wR, psiC, psiR)
      INTEGER(KIND=IWPF),INTENT(IN) :: rank, nadv, nxG, nyG, nxP, nyP
      INTEGER(KIND=IWPF),INTENT(IN) :: rISize, cISize
      REAL(KIND=RWPF),INTENT(IN) :: time, nu
      INTEGER(KIND=IWPF),DIMENSION(cISize),INTENT(IN) :: kxP
      INTEGER(KIND=IWPF),DIMENSION(nyP),INTENT(IN) :: kyP
      INTEGER(KIND=IWPF),INTENT(IN) :: kxLG, kyLG, kxMG, kyMG
      COMPLEX(KIND=CWPC),DIMENSION(cISize,nyP),INTENT(INOUT) :: uC, vC, wC, psiC
      REAL(KIND=RWPC),DIMENSION(rISize,nyP),INTENT(INOUT) :: uR, vR, wR, psiR
      REAL(KIND=RWPC) :: xSkew, ySkew
      REAL(KIND=RWPF) :: tke, eps, enstDiss
      INTEGER :: u

      IF (rank == 0_IWPF) THEN
         WRITE(OUTPUT_UNIT,'(A,ES15.8,A,I8.8)') 'Writing Eulerian statistics at t = ', time, &
            ' and nadv = ', nadv
      END IF

      ! x: the velocity, then du/dx into u and dw/dx into v, both physical
      CALL ComputeVelocity(nxP, nyP, rISize, cISize, kxP, kyP, uC, uR, vC, vR, psiC, .FALSE.)
      CALL DuDx(cISize, nyP, kxP, kyP, uC, uC)
      CALL DuDx(cISize, nyP, kxP, kyP, wC, vC)
      CALL TransformC2R(rISize, cISize, nyP, uC, uR)
      CALL TransformC2R(rISize, cISize, nyP, vC, vR)
      xSkew = Skewness(uR(1:nxP,1:nyP), vR(1:nxP,1:nyP), nxG, nyG)

      ! y: dv/dy into v and dw/dy into u
      CALL ComputeVelocity(nxP, nyP, rISize, cISize, kxP, kyP, uC, uR, vC, vR, psiC, .FALSE.)
      CALL DuDy(cISize, nyP, kxP, kyP, vC, vC)
      CALL DuDy(cISize, nyP, kxP, kyP, wC, uC)
      CALL TransformC2R(rISize, cISize, nyP, vC, vR)
      CALL TransformC2R(rISize, cISize, nyP, uC, uR)
      ySkew = Skewness(vR(1:nxP,1:nyP), uR(1:nxP,1:nyP), nxG, nyG)

      ! the spectrum of the velocity
      CALL ComputeVelocity(nxP, nyP, rISize, cISize, kxP, kyP, uC, uR, vC, vR, psiC, .FALSE.)
      CALL ComputeSpectrum(rank, nu, nxG, nyG, nxP, nyP, rISize, cISize, kxP, kyP, &
                           kxLG, kyLG, kxMG, kyMG, .TRUE., nadv, uC, vC, tke, eps, enstDiss)

      IF (rank == 0_IWPF) THEN
         OPEN(NEWUNIT=u, FILE=TRIM(eulOut), STATUS='OLD', ACTION='WRITE', FORM='FORMATTED', &
              POSITION='APPEND')
         WRITE(u,'(I12.10,6ES16.8)') nadv, time, tke, eps, enstDiss, xSkew, ySkew
         CLOSE(u)
      END IF
   END SUBROUTINE Analysis

   !> ComputeSpectrum, src/Analysis.F90:267-460: shell energies of the
   !! velocity spectra uC, vC (bins of width 1 centred on the integers,
   !! weight 1/2 at kx = 0 and kx = kxMG), the TKE, eps = 2 nu sum k^2 E(k)
   !! and the enstrophy dissipation 2 nu sum k^4 E(k); with output, also
   !! SPECTRUM_<outNum>.DAT. uC and vC are not modified.
   SUBROUTINE ComputeSpectrum(rank, nu, nxG, nyG, nxP, nyP, rISize, cISize, &
                              kxP, kyP, kxLG, kyLG, kxMG, kyMG, output, &
                              outNum, uC, vC, tke, eps, enstDiss)
      INTEGER(KIND=IWPF),INTENT(IN) :: rank, nxG, nyG, nxP, nyP, rISize, cISize
      REAL(KIND=RWPF),INTENT(IN) :: nu
      INTEGER(KIND=IWPF),DIMENSION(cISize),INTENT(IN) :: kxP
      INTEGER(KIND=IWPF),DIMENSION(nyP),INTENT(IN) :: kyP
      INTEGER(KIND=IWPF),INTENT(IN) :: kxLG, kyLG, kxMG, kyMG, outNum
      LOGICAL,INTENT(IN) :: output
      COMPLEX(KIND=CWPC),DIMENSION(cISize,nyP),INTENT(INOUT) :: uC, vC
      REAL(KIND=RWPF),INTENT(OUT) :: tke, eps, enstDiss
      TYPE(s2d_grid_t) :: g
      REAL(C_DOUBLE),DIMENSION(:),ALLOCATABLE :: ek
      INTEGER(C_INT),DIMENSION(:),ALLOCATABLE :: cnt
      REAL(C_DOUBLE) :: t, e, z
      CHARACTER(LEN=FILE_NAME_LENGTH) :: fname
      INTEGER :: nb, b, u

      CALL S2D_Check(s2d_get_grid(s2dCtx, g), s2dCtx, 'ComputeSpectrum')
      nb = INT(s2d_spectrum_bins(g))
      ALLOCATE(ek(nb), cnt(nb))
      CALL S2D_Check(s2d_set_field_host(s2dCtx, S2D_FIELD_U, uC), s2dCtx, 'ComputeSpectrum')
      CALL S2D_Check(s2d_set_field_host(s2dCtx, S2D_FIELD_V, vC), s2dCtx, 'ComputeSpectrum')
      CALL S2D_Check(s2d_spectrum_uv(s2dCtx, S2D_FIELD_U, S2D_FIELD_V, REAL(nu, C_DOUBLE), &
                                     ek, cnt, t, e, z), s2dCtx, 'ComputeSpectrum')
      tke = REAL(t, RWPF)
      eps = REAL(e, RWPF)
      enstDiss = REAL(z, RWPF)
      IF (output .AND. rank == 0_IWPF) THEN
         CALL FileName('SPECTRUM', outNum, 'DAT', fname)
         OPEN(NEWUNIT=u, FILE=TRIM(fname), STATUS='REPLACE', ACTION='WRITE', FORM='FORMATTED')
         WRITE(u,'(A,T3,A,T11,A,T28,A,T45,A)') '#', 'k', 'E(k)', 'D(k)', 'DEnst(k)'
         DO b = 1, nb
            WRITE(u,'(I6.6,ES16.8,1X,ES16.8,1X,ES16.8)') b - 1, ek(b), &
               REAL(b - 1, RWPF)**2*ek(b), REAL(b - 1, RWPF)**4*ek(b)
         END DO
         CLOSE(u)
      END IF
      DEALLOCATE(ek, cnt)
   END SUBROUTINE ComputeSpectrum

   !> -2 <a b^2> / (sqrt(<a^2>) <b^2>) over the nxG x nyG physical grid.
   REAL(KIND=RWPC) FUNCTION Skewness(a, b, nxG, nyG)
      REAL(KIND=RWPC),DIMENSION(:,:),INTENT(IN) :: a, b
      INTEGER(KIND=IWPF),INTENT(IN) :: nxG, nyG
      REAL(KIND=RWPC) :: n, ab2, a2, b2
      n = REAL(nxG, RWPC)*REAL(nyG, RWPC)
      ab2 = SUM(a*b**2)/n
      a2 = SUM(a**2)/n
      b2 = SUM(b**2)/n
      Skewness = -2.0_RWPC*ab2/(SQRT(a2)*b2)
   END FUNCTION Skewness

END MODULE Analysis_m
