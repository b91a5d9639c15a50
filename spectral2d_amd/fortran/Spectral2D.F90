!> @file Spectral2D.F90
!> @brief Driver program in the shape of the reference's Spectral2D_p
!! (src/Spectral2D.F90:253-453): grid set-up, module set-up, the grid file,
!! initial condition, the initial restart file and statistics, the time loop
!! with its print / write / statistics cadence, finalize. The configuration
!! comes from an optional namelist file (first command-line argument)
!! instead of compile-time PARAMETERs:
!!
!!   &spectral2d nx=256, ny=256, nu=0.01, dt=1.0e-3, cfl=0.0, nsteps=1000,
!!     tEnd=0.0, ic='taylor-green', seed=1, k0=10.0, ke=0.5, printPeriod=100,
!!     writePeriod=0, statPeriod=0, statRoot='EULERIAN',
!!     deviceResident=.TRUE., out='' /
!!
!! ic is 'taylor-green' (SetTaylorGreen) or 'turbulence' (the seeded
!! replacement of SetTurbulentVelocity). cfl > 0 recomputes dt on the GPU
!! after every step (ComputeTimeStep). tEnd > 0 ends the run at that time
!! instead of after nsteps steps (the reference's DESIRED_END_TIME,
!! src/Spectral2D.F90:426-430: once time + dt reaches tEnd, one last step
!! of tEnd - time). writePeriod > 0 writes GRID.h5 and the REST_########
!! files (src/IO.F90:98-143) at the start, every writePeriod steps and at
!! the end; statPeriod > 0 writes <statRoot>.DAT and the SPECTRUM files
!! (src/Analysis.F90:90-242) on the same pattern; 0 turns either off.
!! WriteRestart round-trips wC and psiC through the transforms, as in the
!! reference; with deviceResident the host copies are synchronised before
!! and uploaded after each write. out names a raw file receiving the final
!! wC (cISize x nyP complex doubles, Fortran order) for checking.
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
   USE IO_m,ONLY: FileName, WriteGridHDF5, WriteRestart, FILE_NAME_LENGTH, HDF5_OUTPUT
   USE Analysis_m,ONLY: SetupAnalysis, Analysis

   IMPLICIT NONE

   INTEGER(KIND=IWPF) :: nx = 96, ny = 96, nsteps = 100, printPeriod = 10
   INTEGER(KIND=IWPF) :: writePeriod = 0, statPeriod = 0
   INTEGER(KIND=IWPF) :: seed = 1
   REAL(KIND=RWPC) :: nu = 0.005_RWPC, dt = 1.0e-5_RWPC, cfl = 0.0_RWPC
   REAL(KIND=RWPF) :: tEnd = 0.0_RWPF
   REAL(KIND=RWPF) :: k0 = 10.0_RWPF, ke = 0.5_RWPF
   CHARACTER(LEN=32) :: ic = 'taylor-green'
   CHARACTER(LEN=64) :: statRoot = 'EULERIAN'
   CHARACTER(LEN=256) :: out = ''
   LOGICAL :: deviceResident = .TRUE.
   NAMELIST /spectral2d/ nx, ny, nu, dt, cfl, nsteps, tEnd, ic, seed, k0, ke, &
                         printPeriod, writePeriod, statPeriod, statRoot, &
                         deviceResident, out

   INTEGER(KIND=IWPF) :: nxG, nyG, kxLG, kyLG, kxMG, kyMG, kxLU, kyLU, kxMU, kyMU
   INTEGER(KIND=IWPF) :: rISize, cISize, nxP, nyP, j1, j2, allocLocal
   REAL(KIND=RWPF) :: kMax
   INTEGER(KIND=IWPF),DIMENSION(:),ALLOCATABLE :: kxG, kyG
   ! one buffer per field with a complex and a real view, as the reference's
   ! fftw_alloc buffers (src/Spectral2D.F90:287-304)
   COMPLEX(KIND=CWPC),DIMENSION(:,:),ALLOCATABLE,TARGET :: wC, psiC, uC, vC
   REAL(KIND=RWPC),DIMENSION(:,:),POINTER :: wR, psiR, uR, vR
   REAL(KIND=RWPF) :: time = 0.0_RWPF
   INTEGER(KIND=IWPF) :: nadv = 0, i, j, u, ios
   INTEGER(KIND=IWPF) :: printStep, writeStep, statStep
   LOGICAL :: exitThisStep = .FALSE., printBool = .FALSE., writeBool = .FALSE.
   LOGICAL :: statBool = .FALSE.
   CHARACTER(LEN=512) :: cfg
   CHARACTER(LEN=FILE_NAME_LENGTH) :: fname
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
   printStep = printPeriod
   writeStep = writePeriod
   statStep = statPeriod

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
   IF (writePeriod > 0) THEN
      CALL FileName('GRID', 'h5', fname)
      CALL WriteGridHDF5(0_IWPF, 0_IWPF, fname, nxG, nyG, nxP, nyP, 1_IWPF, nxP, j1, j2)
   END IF
   CALL TimeIntegrationSetup(RK3_TVD_SHU, rISize, cISize, nyP, allocLocal)
   ALLOCATE(wC(cISize,nyP), psiC(cISize,nyP), uC(cISize,nyP), vC(cISize,nyP))
   CALL C_F_POINTER(C_LOC(wC), wR, [rISize,nyP])
   CALL C_F_POINTER(C_LOC(psiC), psiR, [rISize,nyP])
   CALL C_F_POINTER(C_LOC(uC), uR, [rISize,nyP])
   CALL C_F_POINTER(C_LOC(vC), vR, [rISize,nyP])
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
   ! the initial restart file and statistics (src/Spectral2D.F90:330-340)
   IF (writePeriod > 0) CALL WriteRestartNow()
   IF (statPeriod > 0) THEN
      CALL SetupAnalysis(TRIM(statRoot))
      CALL AnalysisNow()
   END IF
   CALL UploadHostFields(wC, psiC)
   CALL SetDeviceResident(deviceResident)
   cflNumber = cfl
   IF (cfl > 0.0_RWPC) CALL ComputeTimeStep(dt)
   ! the run's length: FIXED_STEP_COUNT (tEnd <= 0) or DESIRED_END_TIME
   IF (tEnd <= 0.0_RWPF .AND. nsteps <= 0) GOTO 900
   IF (tEnd <= 0.0_RWPF .AND. nsteps == 1) exitThisStep = .TRUE.
   IF (tEnd > 0.0_RWPF .AND. time + dt >= tEnd) THEN
      exitThisStep = .TRUE.
      dt = tEnd - time
   END IF

   ! time loop, src/Spectral2D.F90:338-432: each event runs on the step
   ! where its flag was raised one step ahead, and all of them on the last
   tloop: DO
      CALL IntegrateOneStep(nxG, nyG, nxP, nyP, rISize, cISize, kxG, kyG, &
                            nu, dt, uC, uR, vC, vR, wC, wR, psiC, psiR)
      time = time + dt
      nadv = nadv + 1_IWPF
      ! the CFL dt of the new state (ComputeTimeStep, an empty stub in the
      ! reference)
      CALL ComputeTimeStep(dt)

      IF (printBool .OR. exitThisStep) THEN
         printStep = printStep + printPeriod
         printBool = .FALSE.
         CALL SyncHostFields(wC, psiC)
         ! spectral magnitudes, as the reference prints (src/Spectral2D.F90:358-361)
         WRITE(OUTPUT_UNIT,'(A,I8.8,A,ES15.8,A,ES15.8,A,ES15.8,A,ES15.8,A,ES15.8)') &
            'Simulation step number: ', nadv, '; Simulation time: ', time, &
            '; Max W: ', MAXVAL(ABS(wC)), '; Min W: ', MINVAL(ABS(wC)), &
            '; Max Psi: ', MAXVAL(ABS(psiC)), '; Min Psi: ', MINVAL(ABS(psiC))
      ELSE IF (nadv + 1_IWPF == printStep) THEN
         printBool = .TRUE.
      END IF

      IF (writePeriod > 0) THEN
         IF (writeBool .OR. exitThisStep) THEN
            writeStep = writeStep + writePeriod
            writeBool = .FALSE.
            CALL WriteRestartNow()
         ELSE IF (nadv + 1_IWPF == writeStep) THEN
            writeBool = .TRUE.
         END IF
      END IF

      IF (statPeriod > 0) THEN
         IF (statBool .OR. exitThisStep) THEN
            statStep = statStep + statPeriod
            statBool = .FALSE.
            CALL AnalysisNow()
         ELSE IF (nadv + 1_IWPF == statStep) THEN
            statBool = .TRUE.
         END IF
      END IF

      IF (exitThisStep) EXIT tloop
      IF (tEnd <= 0.0_RWPF) THEN
         IF (nadv + 1_IWPF == nsteps) exitThisStep = .TRUE.
      ELSE IF (time + dt >= tEnd) THEN
         exitThisStep = .TRUE.
         dt = tEnd - time
      END IF
   END DO tloop

900 CONTINUE
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

   wR => NULL()
   psiR => NULL()
   uR => NULL()
   vR => NULL()
   DEALLOCATE(wC, psiC, uC, vC, kxG, kyG)
   CALL TimeIntegrationFinalize()
   CALL SpectralFinalize()

CONTAINS

   !> WriteRestart on the host arrays; in device-resident mode they are
   !! synchronised first and the round-tripped state uploaded afterwards, so
   !! both modes advance what the reference advances.
   SUBROUTINE WriteRestartNow()
      CALL SyncHostFields(wC, psiC)
      CALL WriteRestart(nxG, nyG, nxP, nyP, 1_IWPF, nxP, j1, j2, rISize, cISize, kxG, kyG, &
                        0_IWPF, nadv, time, nu, HDF5_OUTPUT, uC, uR, vC, vR, wC, wR, psiC, psiR)
      IF (deviceResident) CALL UploadHostFields(wC, psiC)
   END SUBROUTINE WriteRestartNow

   SUBROUTINE AnalysisNow()
      CALL SyncHostFields(wC, psiC)
      CALL Analysis(0_IWPF, nadv, time, nu, nxG, nyG, nxP, nyP, rISize, cISize, kxG, kyG, &
                    kxLG, kyLG, kxMG, kyMG, uC, uR, vC, vR, wC, wR, psiC, psiR)
   END SUBROUTINE AnalysisNow

END PROGRAM Spectral2D_p
