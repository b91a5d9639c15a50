!> @file Routines.F90
!> @brief Applies every Spectral_m routine once to fields read from a raw
!! file, through the same module interface the driver uses, and writes
!! every result to a raw file: a command-line harness for checking the
!! drop-in module routine by routine (src/Spectral.F90:83-88).
!!
!! Usage: spectral2d_routines nx nu in.bin out.bin
!!   in.bin:  w, psi, u, v (complex (cISize, nyP) each, Fortran order),
!!            for the 3/2-rule grid of nx x nx.
!!   out.bin: the records below, in order, each one field buffer of
!!            cISize*nyP complex doubles (real results in the real view):
!!    1 DuDx(w)                         2 DuDy(w)
!!    3 ComputePsi(w)                   4 Truncate(w)
!!    5,6 ComputeVelocity(psi, .FALSE.): uC, vC
!!    7 ComputeVorticity(records 5, 6)  8,9 its (destroyed) uC, vC after
!!   10,11 ComputeVelocity(psi, .TRUE.): uR, vR
!!   12 TransformC2R(w)                13 TransformR2C(record 12)
!!   14 ComputeRHS(nu, u, v, w, psi): dW  15,16,17 its uC, vC, wC after
PROGRAM Spectral2D_routines

   USE ISO_C_BINDING
   USE Parameters_m,ONLY: IWPF, RWPF, RWPC, CWPC
   USE Spectral_m

   IMPLICIT NONE

   INTEGER(KIND=IWPF) :: nx, nxG, nyG, kxLG, kyLG, kxMG, kyMG, kxLU, kyLU, kxMU, kyMU
   INTEGER(KIND=IWPF) :: rISize, cISize, nyP, u
   REAL(KIND=RWPF) :: kMax
   REAL(KIND=RWPC) :: nu
   INTEGER(KIND=IWPF),DIMENSION(:),ALLOCATABLE :: kxG, kyG
   COMPLEX(KIND=CWPC),DIMENSION(:,:),ALLOCATABLE,TARGET :: w, psi, ui, vi, a, b, c, d
   REAL(KIND=RWPC),DIMENSION(:,:),POINTER :: aR, bR, cR, pR
   CHARACTER(LEN=512) :: arg, fin, fout

   CALL GET_COMMAND_ARGUMENT(1, arg)
   READ(arg,*) nx
   CALL GET_COMMAND_ARGUMENT(2, arg)
   READ(arg,*) nu
   CALL GET_COMMAND_ARGUMENT(3, fin)
   CALL GET_COMMAND_ARGUMENT(4, fout)

   CALL GetGridSize(nx, nx, DEALIAS_3_2_RULE, nxG, nyG, kxLG, kyLG, kxMG, kyMG, &
                    kxLU, kyLU, kxMU, kyMU, kMax, rISize, cISize, kxG, kyG)
   nyP = nyG
   CALL SetupMask(cISize, nyP, kxG, kyG, kxMG, kyMG)
   CALL SetupTruncation(cISize, nyP, kMax, kxG, kyG)
   ALLOCATE(w(cISize,nyP), psi(cISize,nyP), ui(cISize,nyP), vi(cISize,nyP))
   ALLOCATE(a(cISize,nyP), b(cISize,nyP), c(cISize,nyP), d(cISize,nyP))
   CALL C_F_POINTER(C_LOC(a), aR, [rISize,nyP])
   CALL C_F_POINTER(C_LOC(b), bR, [rISize,nyP])
   CALL C_F_POINTER(C_LOC(c), cR, [rISize,nyP])
   CALL C_F_POINTER(C_LOC(psi), pR, [rISize,nyP])
   CALL SetupPlans(nxG, nyG, nyP, rISize, cISize, aR, a)

   OPEN(NEWUNIT=u, FILE=TRIM(fin), ACCESS='STREAM', FORM='UNFORMATTED', STATUS='OLD', &
        ACTION='READ')
   READ(u) w, psi, ui, vi
   CLOSE(u)
   OPEN(NEWUNIT=u, FILE=TRIM(fout), ACCESS='STREAM', FORM='UNFORMATTED', STATUS='REPLACE', &
        ACTION='WRITE')

   CALL DuDx(cISize, nyP, kxG, kyG, w, a)
   WRITE(u) a
   CALL DuDy(cISize, nyP, kxG, kyG, w, a)
   WRITE(u) a
   CALL ComputePsi(cISize, nyP, kxG, kyG, w, a)
   WRITE(u) a
   a = w
   CALL Truncate(cISize, nyP, a)
   WRITE(u) a

   CALL ComputeVelocity(nxG, nyP, rISize, cISize, kxG, kyG, a, aR, b, bR, psi, .FALSE.)
   WRITE(u) a, b
   CALL ComputeVorticity(nxG, nyP, cISize, kxG, kyG, a, b, c)
   WRITE(u) c, a, b
   CALL ComputeVelocity(nxG, nyP, rISize, cISize, kxG, kyG, a, aR, b, bR, psi, .TRUE.)
   WRITE(u) a, b

   c = w
   CALL TransformC2R(rISize, cISize, nyP, c, cR)
   WRITE(u) c
   CALL TransformR2C(nxG, nyG, nyP, rISize, cISize, cR, c)
   WRITE(u) c

   a = ui
   b = vi
   c = w
   CALL ComputeRHS(nxG, nyG, nxG, nyP, rISize, cISize, kxG, kyG, nu, a, aR, b, bR, c, cR, &
                   psi, pR, d)
   WRITE(u) d, a, b, c
   CLOSE(u)

   DEALLOCATE(w, psi, ui, vi, a, b, c, d, kxG, kyG)
   CALL SpectralFinalize()

END PROGRAM Spectral2D_routines
