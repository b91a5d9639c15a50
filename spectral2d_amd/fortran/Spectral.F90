!> @file Spectral.F90
!> @brief Spectral_m with the public interface of the reference's module
!! (src/Spectral.F90:83-88), implemented over libs2d.
!!
!! One libs2d context replaces the module singletons (FFTW plans, masks,
!! truncation table). Routines that take host arrays copy them to device
!! scratch slots, run the HIP kernel and copy the results back, so they
!! keep the reference's in/out and side-effect semantics exactly. They use
!! the slots U, V, WS, PSIS, DW and TMP only; the W and PSI slots belong to
!! the time integrator's device-resident state (TimeIntegration_m).
MODULE Spectral_m

   USE ISO_C_BINDING
   USE Parameters_m,ONLY: IWPF, RWPF, IWPC, RWPC, CWPC
   USE S2D_C

   IMPLICIT NONE

   INTEGER(KIND=IWPF),PARAMETER,PUBLIC :: NO_DEALIASING = 0_IWPF
   INTEGER(KIND=IWPF),PARAMETER,PUBLIC :: DEALIAS_3_2_RULE = 1_IWPF
   INTEGER(KIND=IWPF),PARAMETER,PUBLIC :: DEALIAS_2_3_RULE = 2_IWPF
   INTEGER(KIND=IWPF),PARAMETER,PUBLIC :: DEALIAS_GRID_SHIFT = 3_IWPF

   !> The libs2d context (shared with TimeIntegration_m).
   TYPE(C_PTR),PUBLIC :: s2dCtx = C_NULL_PTR
   !> Arithmetic of the context (S2D_FP64 unless set before SetupPlans).
   INTEGER(C_INT),PUBLIC :: s2dPrecision = S2D_FP64
   !> Grid request remembered by GetGridSize for SetupPlans.
   INTEGER(KIND=IWPF),PRIVATE :: nxReq = 0, nyReq = 0, dealias = DEALIAS_3_2_RULE

   PUBLIC :: GetGridSize
   PUBLIC :: SetupMask, SetupPlans, SetupTruncation, SpectralFinalize
   PUBLIC :: ComputeRHS, ComputePsi, ComputeVelocity, ComputeVorticity
   PUBLIC :: TransformR2C, TransformC2R
   PUBLIC :: DuDx, DuDy
   PUBLIC :: Truncate

CONTAINS

   !> GetGridSize, src/Spectral.F90:117-219 (host only): the sizes and the
   !! FFTW-order wavenumbers from libs2d (s2d_get_grid_size, s2d_wavenumbers).
   SUBROUTINE GetGridSize(nx, ny, dealias_, nxG, nyG, &
                          kxLG, kyLG, kxMG, kyMG, &
                          kxLU, kyLU, kxMU, kyMU, &
                          kMax, rISize, cISize, kxG, kyG)
      INTEGER(KIND=IWPF),INTENT(IN) :: nx, ny, dealias_
      INTEGER(KIND=IWPF),INTENT(OUT) :: nxG, nyG
      INTEGER(KIND=IWPF),INTENT(OUT) :: kxLG, kyLG, kxMG, kyMG
      INTEGER(KIND=IWPF),INTENT(OUT) :: kxLU, kyLU, kxMU, kyMU
      INTEGER(KIND=IWPF),INTENT(OUT) :: rISize, cISize
      REAL(KIND=RWPF),INTENT(OUT) :: kMax
      INTEGER(KIND=IWPF),DIMENSION(:),ALLOCATABLE,INTENT(INOUT) :: kxG, kyG
      TYPE(s2d_grid_t) :: g

      CALL S2D_Check(s2d_get_grid_size(INT(nx, C_INT), INT(ny, C_INT), INT(dealias_, C_INT), &
                                       g), C_NULL_PTR, 'GetGridSize')
      dealias = dealias_
      nxReq = nx
      nyReq = ny
      nxG = g%nxG
      nyG = g%nyG
      kxLG = g%kxLG
      kyLG = g%kyLG
      kxMG = g%kxMG
      kyMG = g%kyMG
      kxLU = g%kxLU
      kyLU = g%kyLU
      kxMU = g%kxMU
      kyMU = g%kyMU
      kMax = REAL(g%kMax, RWPF)
      rISize = g%rISize
      cISize = g%cISize
      IF (ALLOCATED(kxG)) DEALLOCATE(kxG)
      IF (ALLOCATED(kyG)) DEALLOCATE(kyG)
      ALLOCATE(kxG(cISize), kyG(nyG))
      CALL S2D_Check(s2d_wavenumbers(g, kxG, kyG), C_NULL_PTR, 'GetGridSize')
   END SUBROUTINE GetGridSize

   !> SetupMask, src/Spectral.F90:229-260: the masks are computed by
   !! s2d_create; nothing to do before the context exists.
   SUBROUTINE SetupMask(cISize, nyP, kxP, kyP, kxMG, kyMG)
      INTEGER(KIND=IWPF),INTENT(IN) :: cISize, nyP, kxMG, kyMG
      INTEGER(KIND=IWPF),DIMENSION(cISize),INTENT(IN) :: kxP
      INTEGER(KIND=IWPF),DIMENSION(nyP),INTENT(IN) :: kyP
   END SUBROUTINE SetupMask

   !> SetupTruncation, src/Spectral.F90:272-313: computed by s2d_create.
   SUBROUTINE SetupTruncation(cISize, nyP, kMax, kxP, kyP)
      INTEGER(KIND=IWPF),INTENT(IN) :: cISize, nyP
      REAL(KIND=RWPF),INTENT(IN) :: kMax
      INTEGER(KIND=IWPF),DIMENSION(cISize),INTENT(IN) :: kxP
      INTEGER(KIND=IWPF),DIMENSION(nyP),INTENT(IN) :: kyP
   END SUBROUTINE SetupTruncation

   !> SetupPlans, src/Spectral.F90:330-353: creates the libs2d context
   !! (transform plans, twiddles, masks, truncation, device slots) and zeroes
   !! dataC as the reference does after planning.
   SUBROUTINE SetupPlans(nxG, nyG, nyP, rISize, cISize, dataR, dataC)
      INTEGER(KIND=IWPF),INTENT(IN) :: nxG, nyG, nyP, rISize, cISize
      REAL(KIND=RWPC),DIMENSION(rISize,nyP),INTENT(INOUT) :: dataR
      COMPLEX(KIND=CWPC),DIMENSION(cISize,nyP),INTENT(INOUT) :: dataC
      IF (C_ASSOCIATED(s2dCtx)) CALL SpectralFinalize()
      CALL S2D_Check(s2d_create(INT(nxReq, C_INT), INT(nyReq, C_INT), &
                                INT(dealias, C_INT), s2dPrecision, s2dCtx), &
                     C_NULL_PTR, 'SetupPlans')
      dataC(:,:) = (0.0_RWPC, 0.0_RWPC)
   END SUBROUTINE SetupPlans

   !> ComputeRHS, src/Spectral.F90:396-482.
   SUBROUTINE ComputeRHS(nxG, nyG, nxP, nyP, rISize, cISize, kxP, kyP, nu, &
                         uC, uR, vC, vR, wC, wR, psiC, psiR, dW)
      INTEGER(KIND=IWPF),INTENT(IN) :: nxG, nyG, nxP, nyP, rISize, cISize
      INTEGER(KIND=IWPF),DIMENSION(cISize),INTENT(IN) :: kxP
      INTEGER(KIND=IWPF),DIMENSION(nyP),INTENT(IN) :: kyP
      REAL(KIND=RWPC),INTENT(IN) :: nu
      COMPLEX(KIND=CWPC),DIMENSION(cISize,nyP),INTENT(INOUT) :: uC, vC, wC
      REAL(KIND=RWPC),DIMENSION(rISize,nyP),INTENT(INOUT) :: uR, vR, wR
      COMPLEX(KIND=CWPC),DIMENSION(cISize,nyP),INTENT(IN) :: psiC
      REAL(KIND=RWPC),DIMENSION(rISize,nyP),INTENT(IN) :: psiR
      COMPLEX(KIND=CWPC),DIMENSION(cISize,nyP),INTENT(OUT) :: dW
      CALL Put(S2D_FIELD_WS, wC)
      CALL Put(S2D_FIELD_PSIS, psiC)
      CALL S2D_Check(s2d_compute_rhs(s2dCtx, REAL(nu, C_DOUBLE), S2D_FIELD_U, S2D_FIELD_V, &
                                     S2D_FIELD_WS, S2D_FIELD_PSIS, S2D_FIELD_DW), &
                     s2dCtx, 'ComputeRHS')
      CALL Get(S2D_FIELD_U, uC)
      CALL Get(S2D_FIELD_V, vC)
      CALL Get(S2D_FIELD_WS, wC)
      CALL Get(S2D_FIELD_DW, dW)
   END SUBROUTINE ComputeRHS

   !> ComputePsi, src/Spectral.F90:508-560.
   SUBROUTINE ComputePsi(cISize, nyP, kxP, kyP, Qw, Qp)
      INTEGER(KIND=IWPF),INTENT(IN) :: cISize, nyP
      INTEGER(KIND=IWPF),DIMENSION(cISize),INTENT(IN) :: kxP
      INTEGER(KIND=IWPF),DIMENSION(nyP),INTENT(IN) :: kyP
      COMPLEX(KIND=CWPC),DIMENSION(cISize,nyP),INTENT(IN) :: Qw
      COMPLEX(KIND=CWPC),DIMENSION(cISize,nyP),INTENT(OUT) :: Qp
      CALL Put(S2D_FIELD_WS, Qw)
      CALL S2D_Check(s2d_compute_psi(s2dCtx, S2D_FIELD_WS, S2D_FIELD_PSIS), s2dCtx, 'ComputePsi')
      CALL Get(S2D_FIELD_PSIS, Qp)
   END SUBROUTINE ComputePsi

   !> ComputeVelocity, src/Spectral.F90:576-610.
   SUBROUTINE ComputeVelocity(nxP, nyP, rISize, cISize, kxP, kyP, &
                              uC, uR, vC, vR, PsiC, trans)
      INTEGER(KIND=IWPF),INTENT(IN) :: nxP, nyP, rISize, cISize
      INTEGER(KIND=IWPF),DIMENSION(cISize),INTENT(IN) :: kxP
      INTEGER(KIND=IWPF),DIMENSION(nyP),INTENT(IN) :: kyP
      COMPLEX(KIND=CWPC),DIMENSION(cISize,nyP),INTENT(INOUT) :: uC, vC
      REAL(KIND=RWPC),DIMENSION(rISize,nyP),INTENT(INOUT) :: uR, vR
      COMPLEX(CWPC),DIMENSION(cISize,nyP),INTENT(IN) :: PsiC
      LOGICAL,INTENT(IN) :: trans
      INTEGER(C_INT) :: t
      t = 0
      IF (trans) t = 1
      CALL Put(S2D_FIELD_PSIS, PsiC)
      CALL S2D_Check(s2d_compute_velocity(s2dCtx, S2D_FIELD_U, S2D_FIELD_V, S2D_FIELD_PSIS, t), &
                     s2dCtx, 'ComputeVelocity')
      CALL Get(S2D_FIELD_U, uC)
      CALL Get(S2D_FIELD_V, vC)
   END SUBROUTINE ComputeVelocity

   !> ComputeVorticity, src/Spectral.F90:624-647 (destroys uC and vC).
   SUBROUTINE ComputeVorticity(nxP, nyP, cISize, kxP, kyP, uC, vC, wC)
      INTEGER(KIND=IWPF),INTENT(IN) :: nxP, nyP, cISize
      INTEGER(KIND=IWPF),DIMENSION(cISize),INTENT(IN) :: kxP
      INTEGER(KIND=IWPF),DIMENSION(nyP),INTENT(IN) :: kyP
      COMPLEX(KIND=CWPC),DIMENSION(cISize,nyP),INTENT(INOUT) :: uC, vC, wC
      CALL Put(S2D_FIELD_U, uC)
      CALL Put(S2D_FIELD_V, vC)
      CALL S2D_Check(s2d_compute_vorticity(s2dCtx, S2D_FIELD_U, S2D_FIELD_V, S2D_FIELD_TMP), &
                     s2dCtx, 'ComputeVorticity')
      CALL Get(S2D_FIELD_U, uC)
      CALL Get(S2D_FIELD_V, vC)
      CALL Get(S2D_FIELD_TMP, wC)
   END SUBROUTINE ComputeVorticity

   !> DuDx, src/Spectral.F90:662-693.
   SUBROUTINE DuDx(cISize, nyP, kxP, kyP, sigIn, sigOut)
      INTEGER(KIND=IWPF),INTENT(IN) :: cISize, nyP
      INTEGER(KIND=IWPF),DIMENSION(cISize),INTENT(IN) :: kxP
      INTEGER(KIND=IWPF),DIMENSION(nyP),INTENT(IN) :: kyP
      COMPLEX(KIND=CWPC),DIMENSION(cISize,nyP),INTENT(IN) :: sigIn
      COMPLEX(KIND=CWPC),DIMENSION(cISize,nyP),INTENT(OUT) :: sigOut
      CALL Put(S2D_FIELD_TMP, sigIn)
      CALL S2D_Check(s2d_dudx(s2dCtx, S2D_FIELD_TMP, S2D_FIELD_TMP), s2dCtx, 'DuDx')
      CALL Get(S2D_FIELD_TMP, sigOut)
   END SUBROUTINE DuDx

   !> DuDy, src/Spectral.F90:708-738.
   SUBROUTINE DuDy(cISize, nyP, kxP, kyP, sigIn, sigOut)
      INTEGER(KIND=IWPF),INTENT(IN) :: cISize, nyP
      INTEGER(KIND=IWPF),DIMENSION(cISize),INTENT(IN) :: kxP
      INTEGER(KIND=IWPF),DIMENSION(nyP),INTENT(IN) :: kyP
      COMPLEX(KIND=CWPC),DIMENSION(cISize,nyP),INTENT(IN) :: sigIn
      COMPLEX(KIND=CWPC),DIMENSION(cISize,nyP),INTENT(OUT) :: sigOut
      CALL Put(S2D_FIELD_TMP, sigIn)
      CALL S2D_Check(s2d_dudy(s2dCtx, S2D_FIELD_TMP, S2D_FIELD_TMP), s2dCtx, 'DuDy')
      CALL Get(S2D_FIELD_TMP, sigOut)
   END SUBROUTINE DuDy

   !> TransformR2C, src/Spectral.F90:749-763 (in place: rData and cData are
   !! the two views of one buffer).
   SUBROUTINE TransformR2C(nxG, nyG, nyP, rISize, cISize, rData, cData)
      INTEGER(KIND=IWPF),INTENT(IN) :: nxG, nyG, nyP, rISize, cISize
      REAL(KIND=RWPC),DIMENSION(rISize,nyP),INTENT(INOUT) :: rData
      COMPLEX(KIND=CWPC),DIMENSION(cISize,nyP),INTENT(INOUT) :: cData
      CALL PutR(S2D_FIELD_TMP, rData)
      CALL S2D_Check(s2d_transform_r2c(s2dCtx, S2D_FIELD_TMP), s2dCtx, 'TransformR2C')
      CALL Get(S2D_FIELD_TMP, cData)
   END SUBROUTINE TransformR2C

   !> TransformC2R, src/Spectral.F90:772-781.
   SUBROUTINE TransformC2R(rISize, cISize, nyP, cData, rData)
      INTEGER(KIND=IWPF),INTENT(IN) :: rISize, cISize, nyP
      COMPLEX(KIND=CWPC),DIMENSION(cISize,nyP),INTENT(INOUT) :: cData
      REAL(KIND=RWPC),DIMENSION(rISize,nyP),INTENT(INOUT) :: rData
      CALL Put(S2D_FIELD_TMP, cData)
      CALL S2D_Check(s2d_transform_c2r(s2dCtx, S2D_FIELD_TMP), s2dCtx, 'TransformC2R')
      CALL GetR(S2D_FIELD_TMP, rData)
   END SUBROUTINE TransformC2R

   !> Truncate, src/Spectral.F90:788-802.
   SUBROUTINE Truncate(cISize, nyP, dataC)
      INTEGER(KIND=IWPF),INTENT(IN) :: cISize, nyP
      COMPLEX(KIND=CWPC),DIMENSION(cISize,nyP),INTENT(INOUT) :: dataC
      CALL Put(S2D_FIELD_TMP, dataC)
      CALL S2D_Check(s2d_truncate(s2dCtx, S2D_FIELD_TMP), s2dCtx, 'Truncate')
      CALL Get(S2D_FIELD_TMP, dataC)
   END SUBROUTINE Truncate

   !> SpectralFinalize, src/Spectral.F90:805-813.
   SUBROUTINE SpectralFinalize()
      INTEGER(C_INT) :: rc
      IF (C_ASSOCIATED(s2dCtx)) THEN
         rc = s2d_destroy(s2dCtx)
         s2dCtx = C_NULL_PTR
      END IF
   END SUBROUTINE SpectralFinalize

   ! ---- host <-> device copies of a whole field buffer
   SUBROUTINE Put(field, a)
      INTEGER(C_INT),INTENT(IN) :: field
      COMPLEX(KIND=CWPC),DIMENSION(:,:),INTENT(IN) :: a
      CALL S2D_Check(s2d_set_field_host(s2dCtx, field, a), s2dCtx, 'set_field_host')
   END SUBROUTINE Put

   SUBROUTINE Get(field, a)
      INTEGER(C_INT),INTENT(IN) :: field
      COMPLEX(KIND=CWPC),DIMENSION(:,:),INTENT(INOUT) :: a
      CALL S2D_Check(s2d_get_field_host(s2dCtx, field, a), s2dCtx, 'get_field_host')
   END SUBROUTINE Get

   SUBROUTINE PutR(field, a)
      INTEGER(C_INT),INTENT(IN) :: field
      REAL(KIND=RWPC),DIMENSION(:,:),INTENT(IN) :: a
      CALL S2D_Check(s2d_set_field_host(s2dCtx, field, a), s2dCtx, 'set_field_host')
   END SUBROUTINE PutR

   SUBROUTINE GetR(field, a)
      INTEGER(C_INT),INTENT(IN) :: field
      REAL(KIND=RWPC),DIMENSION(:,:),INTENT(INOUT) :: a
      CALL S2D_Check(s2d_get_field_host(s2dCtx, field, a), s2dCtx, 'get_field_host')
   END SUBROUTINE GetR

END MODULE Spectral_m
