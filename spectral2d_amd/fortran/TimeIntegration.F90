!> @file TimeIntegration.F90
!> @brief TimeIntegration_m with the public interface of the reference's
!! module (src/TimeIntegration.F90:67-68), running the fused TVD-RK3 step of
!! libs2d.
!!
!! Default mode: the host wC and psiC are the state, as in the reference.
!! IntegrateOneStep uploads them, steps on the GPU and downloads them, so
!! that only wC and psiC are valid afterwards (src/Spectral2D.F90:341) and
!! whatever the caller wrote into them between steps (the reference's
!! WriteRestart round trip, src/IO.F90:118-142, a restart read, SetIC) is
!! what the next step advances.
!!
!! Device-resident mode (SetDeviceResident): the state stays on the GPU
!! between steps; the host copies are refreshed by SyncHostFields only, and
!! a host change reaches the GPU only through UploadHostFields or
!! MarkHostModified (the next step then uploads).
MODULE TimeIntegration_m

   USE ISO_C_BINDING
   USE Parameters_m,ONLY: IWPC, IWPF, RWPC, RWPF, CWPC
   USE S2D_C
   USE Spectral_m,ONLY: s2dCtx

   IMPLICIT NONE

   INTEGER(KIND=IWPF),PARAMETER,PUBLIC :: RK3_TVD_SHU = 1_IWPF

   INTEGER(KIND=IWPF),PRIVATE :: scheme = 0_IWPF
   LOGICAL,PRIVATE :: hostNewer = .TRUE.
   LOGICAL,PRIVATE :: deviceResident = .FALSE.
   !> CFL number for ComputeTimeStep (<= 0: keep dt fixed, as the reference).
   REAL(KIND=RWPC),PUBLIC :: cflNumber = 0.0_RWPC

   PUBLIC :: TimeIntegrationSetup, TimeIntegrationFinalize
   PUBLIC :: IntegrateOneStep, ComputeTimeStep
   PUBLIC :: SetDeviceResident, MarkHostModified, SyncHostFields, UploadHostFields

CONTAINS

   !> TimeIntegrationSetup, src/TimeIntegration.F90:79-108. The RK scratch
   !! (dW, wS, psiS) is device memory owned by the libs2d context.
   SUBROUTINE TimeIntegrationSetup(scheme_, rISize, cISize, nyP, allocLocal)
      INTEGER(KIND=IWPF),INTENT(IN) :: scheme_, rISize, cISize, nyP, allocLocal
      SELECT CASE (scheme_)
         CASE (RK3_TVD_SHU)
            scheme = RK3_TVD_SHU
         CASE DEFAULT
            WRITE(*,'(A,I0)') 'TimeIntegrationSetup: unknown scheme ', scheme_
            ERROR STOP 1
      END SELECT
      hostNewer = .TRUE.
   END SUBROUTINE TimeIntegrationSetup

   !> IntegrateOneStep, src/TimeIntegration.F90:130-210.
   SUBROUTINE IntegrateOneStep(nxG, nyG, nxP, nyP, rISize, cISize, kxP, kyP, &
                               nu, dt, uC, uR, vC, vR, wC, wR, psiC, psiR)
      INTEGER(KIND=IWPF),INTENT(IN) :: nxG, nyG, nxP, nyP, rISize, cISize
      INTEGER(KIND=IWPF),DIMENSION(cISize),INTENT(IN) :: kxP
      INTEGER(KIND=IWPF),DIMENSION(nyP),INTENT(IN) :: kyP
      REAL(KIND=RWPC),INTENT(IN) :: nu, dt
      COMPLEX(KIND=CWPC),DIMENSION(cISize,nyP),INTENT(INOUT) :: uC, vC, wC, psiC
      REAL(KIND=RWPC),DIMENSION(rISize,nyP),INTENT(INOUT) :: uR, vR, wR, psiR
      IF (hostNewer) THEN
         CALL S2D_Check(s2d_set_field_host(s2dCtx, S2D_FIELD_W, wC), s2dCtx, 'IntegrateOneStep')
         CALL S2D_Check(s2d_set_field_host(s2dCtx, S2D_FIELD_PSI, psiC), s2dCtx, &
                        'IntegrateOneStep')
         hostNewer = .FALSE.
      ELSE IF (.NOT. deviceResident) THEN
         ! the fused step derives psi from w (psi_hat = w_hat/|k|^2), so the
         ! host w is the whole input
         CALL S2D_Check(s2d_set_field_host(s2dCtx, S2D_FIELD_W, wC), s2dCtx, 'IntegrateOneStep')
      END IF
      CALL S2D_Check(s2d_rk3_steps(s2dCtx, 1_C_INT, REAL(nu, C_DOUBLE), REAL(dt, C_DOUBLE), &
                                   0.0_C_DOUBLE), s2dCtx, 'IntegrateOneStep')
      IF (.NOT. deviceResident) THEN
         CALL S2D_Check(s2d_get_field_host(s2dCtx, S2D_FIELD_W, wC), s2dCtx, 'IntegrateOneStep')
         CALL S2D_Check(s2d_get_field_host(s2dCtx, S2D_FIELD_PSI, psiC), s2dCtx, &
                        'IntegrateOneStep')
      END IF
   END SUBROUTINE IntegrateOneStep

   !> ComputeTimeStep, src/TimeIntegration.F90:213-215 (empty in the
   !! reference). With cflNumber > 0 it returns the CFL step of the current
   !! device state in dt; otherwise dt is left unchanged.
   SUBROUTINE ComputeTimeStep(dt)
      REAL(KIND=RWPC),INTENT(INOUT),OPTIONAL :: dt
      REAL(C_DOUBLE) :: d
      IF (.NOT. PRESENT(dt)) RETURN
      IF (cflNumber <= 0.0_RWPC) RETURN
      CALL S2D_Check(s2d_cfl_dt(s2dCtx, REAL(cflNumber, C_DOUBLE), d), s2dCtx, 'ComputeTimeStep')
      dt = REAL(d, RWPC)
   END SUBROUTINE ComputeTimeStep

   !> TimeIntegrationFinalize, src/TimeIntegration.F90:218-233.
   SUBROUTINE TimeIntegrationFinalize()
      scheme = 0_IWPF
      hostNewer = .TRUE.
   END SUBROUTINE TimeIntegrationFinalize

   !> Keep the state on the GPU between steps (host copies refreshed by
   !! SyncHostFields only).
   SUBROUTINE SetDeviceResident(flag)
      LOGICAL,INTENT(IN) :: flag
      deviceResident = flag
   END SUBROUTINE SetDeviceResident

   !> The host wC/psiC were changed outside IntegrateOneStep: upload them on
   !! the next step.
   SUBROUTINE MarkHostModified()
      hostNewer = .TRUE.
   END SUBROUTINE MarkHostModified

   !> Copy wC and psiC into the device state now.
   SUBROUTINE UploadHostFields(wC, psiC)
      COMPLEX(KIND=CWPC),DIMENSION(:,:),INTENT(IN) :: wC, psiC
      CALL S2D_Check(s2d_set_field_host(s2dCtx, S2D_FIELD_W, wC), s2dCtx, 'UploadHostFields')
      CALL S2D_Check(s2d_set_field_host(s2dCtx, S2D_FIELD_PSI, psiC), s2dCtx, 'UploadHostFields')
      hostNewer = .FALSE.
   END SUBROUTINE UploadHostFields

   !> Copy the device state into wC and psiC (device-resident mode; in the
   !! default mode the host arrays already are the state).
   SUBROUTINE SyncHostFields(wC, psiC)
      COMPLEX(KIND=CWPC),DIMENSION(:,:),INTENT(INOUT) :: wC, psiC
      IF (hostNewer .OR. .NOT. deviceResident) RETURN
      CALL S2D_Check(s2d_get_field_host(s2dCtx, S2D_FIELD_W, wC), s2dCtx, 'SyncHostFields')
      CALL S2D_Check(s2d_get_field_host(s2dCtx, S2D_FIELD_PSI, psiC), s2dCtx, 'SyncHostFields')
   END SUBROUTINE SyncHostFields

END MODULE TimeIntegration_m
