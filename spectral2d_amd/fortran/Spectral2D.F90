!> @file Spectral2D.F90
!> @brief Driver program in the shape of the reference's Spectral2D_p
!! (src/Spectral2D.F90:253-453): grid set-up, module set-up, initial
!! condition, the time loop with its print cadence, finalize. The
!! configuration comes from an optional namelist file (first command-line
!! argument) instead of compile-time PARAMETERs:
!!
!!   &spectral2d nx=256, ny=256, nu=0.01, dt=1.0e-3, cfl=0.0, nsteps=1000,
!!     tEnd=0.0, ic='taylor-green', seed=1, k0=10.0, ke=0.5, printPeriod=100,
!!     deviceResident=.TRUE., out='' /
!!
!! ic is 'taylor-green' (SetTaylorGreen) or 'turbulence' (the seeded
!! replacement of SetTurbulentVelocity). cfl > 0 recomputes dt on the GPU
!! every step. tEnd > 0 ends the run at that time instead of after nsteps
!! steps (the reference's DESIRED_END_TIME, src/Spectral2D.F90:426-430: once
!! time + dt reaches tEnd, one last step of tEnd - time). out names a raw
!! file receiving the final wC (cISize x nyP complex doubles, Fortran order)
!! for checking.
PROGRAM Spectral2D_p

   USE ISO_C_BINDING
   USE ISO_FORTRAN_ENV,ONLY: OUTPUT_UNIT
   USE Parameters_m,ONLY: IWPF, RWPF, RWPC, CWPC, PI
   USE S2D_C
   USE Spectral_m,ONLY: GetGridSize, SetupMask, SetupTruncation, SetupPlans, &
                        SpectralFinalize, TransformR2C, ComputePsi, s2dCtx, &
                        DEALIAS_3_2_RULE
   USE TimeIntegration_m,ONLY: TimeIntegrationSetup, IntegrateOneStep, &
                               ComputeTimeStep, RK3_TVD_SHU, &
                               TimeIntegrationFinalize, SetDeviceResident, &
                               UploadHostFields, SyncHostFields, cflNumber

   IMPLICIT NONE

   INTEGER(KIND=IWPF) :: nx = 96, ny = 96, nsteps = 100, printPeriod = 10
   INTEGER(KIND=IWPF) :: seed = 1
   REAL(KIND=RWPC) :: nu = 0.005_RWPC, dt = 1.0e-5_RWPC, cfl = 0.0_RWPC
   REAL(KIND=RWPF) :: tEnd = 0.0_RWPF
   REAL(KIND=RWPF) :: k0 = 10.0_RWPF, ke = 0.5_RWPF
   CHARACTER(LEN=32) :: ic = 'taylor-green'
   CHARACTER(LEN=256) :: out = ''
   LOGICAL :: deviceResident = .TRUE.
   NAMELIST /spectral2d/ nx, ny, nu, dt, cfl, nsteps, tEnd, ic, seed, k0, ke, &
                         printPeriod, deviceResident, out

   INTEGER(KIND=IWPF) :: nxG, nyG, kxLG, kyLG, kxMG, kyMG, kxLU, kyLU, kxMU, kyMU
   INTEGER(KIND=IWPF) :: rISize, cISize, nxP, nyP, j1, j2, allocLocal
   REAL(KIND=RWPF) :: kMax
   INTEGER(KIND=IWPF),DIMENSION(:),ALLOCATABLE :: kxG, kyG
   COMPLEX(KIND=CWPC),DIMENSION(:,:),ALLOCATABLE,TARGET :: wC, psiC, uC, vC
   REAL(KIND=RWPC),DIMENSION(:,:),ALLOCATABLE :: wR, psiR, uR, vR
   REAL(KIND=RWPF) :: time = 0.0_RWPF
   INTEGER(KIND=IWPF) :: nadv = 0, i, j, u, ios
   LOGICAL :: exitThisStep = .FALSE.
   CHARACTER(LEN=512) :: cfg
   TYPE(s2d_diag_t) :: diag

   ! configuration
   IF (COMMAND_ARGUMENT_COUNT() >= 1) THEN
      CALL GET_COMMAND_ARGUMENT(1, cfg)
      OPEN(NEWUNIT=u, FILE=TRIM(cfg), STATUS='OLD', ACTION='READ', IOSTAT=ios)
      IF (ios /= 0) THEN
         WRITE(*,'(A,A)') 'cannot open ', TRIM(cfg)
         ERROR STOP 1
      END IF
      READ(u, NML=spectral2d)
      CLOSE(u)
   END IF

   ! grid and modules, src/Spectral2D.F90:263-307 (one process: nyP = nyG)
   CALL GetGridSize(nx, ny, DEALIAS_3_2_RULE, nxG, nyG, kxLG, kyLG, kxMG, kyMG, &
                    kxLU, kyLU, kxMU, kyMU, kMax, rISize, cISize, kxG, kyG)
   nxP = nxG
   nyP = nyG
   j1 = 1
   j2 = nyG
   allocLocal = cISize*nyP
   CALL SetupMask(cISize, nyP, kxG, kyG, kxMG, kyMG)
   CALL SetupTruncation(cISize, nyP, kMax, kxG, kyG)
   CALL TimeIntegrationSetup(RK3_TVD_SHU, rISize, cISize, nyP, allocLocal)
   ALLOCATE(wC(cISize,nyP), psiC(cISize,nyP), uC(cISize,nyP), vC(cISize,nyP))
   ALLOCATE(wR(rISize,nyP), psiR(rISize,nyP), uR(rISize,nyP), vR(rISize,nyP))
   wC = (0.0_RWPC, 0.0_RWPC)
   psiC = (0.0_RWPC, 0.0_RWPC)
   uC = (0.0_RWPC, 0.0_RWPC)
   vC = (0.0_RWPC, 0.0_RWPC)
   CALL SetupPlans(nxG, nyG, nyP, rISize, cISize, wR, wC)

   ! initial condition
   SELECT CASE (TRIM(ic))
      CASE ('taylor-green')
         ! SetTaylorGreen, src/SetIC.F90:88-100 (x_i = i dx, 1-based)
         DO j = j1, j2
            DO i = 1, nxP
               psiR(i,j-j1+1) = SIN(2.0_RWPC*PI*REAL(i, RWPC)/REAL(nxG, RWPC))* &
                                SIN(2.0_RWPC*PI*REAL(j, RWPC)/REAL(nyG, RWPC))
               wR(i,j-j1+1) = 2.0_RWPC*SIN(2.0_RWPC*PI*REAL(i, RWPC)/REAL(nxG, RWPC))* &
                              SIN(2.0_RWPC*PI*REAL(j, RWPC)/REAL(nyG, RWPC))
            END DO
         END DO
         CALL TransformR2C(nxG, nyG, nyP, rISize, cISize, psiR, psiC)
         CALL TransformR2C(nxG, nyG, nyP, rISize, cISize, wR, wC)
      CASE ('turbulence')
         CALL S2D_Check(s2d_set_turbulence(s2dCtx, INT(seed, C_LONG_LONG), &
                                           REAL(k0, C_DOUBLE), REAL(ke, C_DOUBLE)), &
                        s2dCtx, 'SetTurbulence')
         CALL S2D_Check(s2d_get_field_host(s2dCtx, S2D_FIELD_W, wC), s2dCtx, 'IC')
         CALL ComputePsi(cISize, nyP, kxG, kyG, wC, psiC)
      CASE DEFAULT
         WRITE(*,'(A,A)') 'unknown ic ', TRIM(ic)
         ERROR STOP 1
   END SELECT
   CALL UploadHostFields(wC, psiC)
   CALL SetDeviceResident(deviceResident)
   cflNumber = cfl

   ! time loop, src/Spectral2D.F90:338-432: print cadence, then the exit
   ! conditions (FIXED_STEP_COUNT with tEnd <= 0, else DESIRED_END_TIME)
   tloop: DO
      IF (tEnd <= 0.0_RWPF .AND. nadv >= nsteps) EXIT tloop
      IF (cfl > 0.0_RWPC) THEN
         CALL IntegrateOneStepCfl()
      ELSE
         CALL IntegrateOneStep(nxG, nyG, nxP, nyP, rISize, cISize, kxG, kyG, &
                               nu, dt, uC, uR, vC, vR, wC, wR, psiC, psiR)
         time = time + dt
         nadv = nadv + 1_IWPF
      END IF
      IF (MOD(nadv, printPeriod) == 0 .OR. (tEnd <= 0.0_RWPF .AND. nadv == nsteps) &
          .OR. exitThisStep) THEN
         CALL SyncHostFields(wC, psiC)
         ! spectral magnitudes, as the reference prints (src/Spectral2D.F90:358-361)
         WRITE(OUTPUT_UNIT,'(A,I8.8,A,ES15.8,A,ES15.8,A,ES15.8,A,ES15.8,A,ES15.8)') &
            'Simulation step number: ', nadv, '; Simulation time: ', time, &
            '; Max W: ', MAXVAL(ABS(wC)), '; Min W: ', MINVAL(ABS(wC)), &
            '; Max Psi: ', MAXVAL(ABS(psiC)), '; Min Psi: ', MINVAL(ABS(psiC))
      END IF
      IF (tEnd > 0.0_RWPF) THEN
         IF (exitThisStep) EXIT tloop
         ! fixed dt: the next step is the last once it would reach tEnd
         ! (with a CFL dt the clamp follows ComputeTimeStep instead)
         IF (cfl <= 0.0_RWPC .AND. time + dt >= tEnd) THEN
            exitThisStep = .TRUE.
            dt = tEnd - time
         END IF
      END IF
   END DO tloop

   CALL S2D_Check(s2d_diagnostics(s2dCtx, diag), s2dCtx, 'diagnostics')
   WRITE(OUTPUT_UNIT,'(A,ES24.16,A,ES24.16,A,ES24.16)') 'Final TKE: ', diag%tke, &
      '; enstrophy: ', diag%enstrophy, '; time: ', diag%time
   IF (LEN_TRIM(out) > 0) THEN
      CALL SyncHostFields(wC, psiC)
      OPEN(NEWUNIT=u, FILE=TRIM(out), ACCESS='STREAM', FORM='UNFORMATTED', &
           STATUS='REPLACE', ACTION='WRITE')
      WRITE(u) wC
      CLOSE(u)
   END IF

   DEALLOCATE(wC, psiC, uC, vC, wR, psiR, uR, vR, kxG, kyG)
   CALL TimeIntegrationFinalize()
   CALL SpectralFinalize()

CONTAINS

   !> One step with the CFL dt of the state at its start, computed on the
   !! device (ComputeTimeStep fills the reference's empty stub).
   SUBROUTINE IntegrateOneStepCfl()
      CALL ComputeTimeStep(dt)
      IF (tEnd > 0.0_RWPF .AND. time + dt >= tEnd) THEN
         exitThisStep = .TRUE.
         dt = tEnd - time
      END IF
      CALL IntegrateOneStep(nxG, nyG, nxP, nyP, rISize, cISize, kxG, kyG, &
                            nu, dt, uC, uR, vC, vR, wC, wR, psiC, psiR)
      time = time + dt
      nadv = nadv + 1_IWPF
   END SUBROUTINE IntegrateOneStepCfl

END PROGRAM Spectral2D_p
