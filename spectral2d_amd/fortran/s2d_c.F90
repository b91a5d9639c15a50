!> @file s2d_c.F90
!> @brief ISO_C_BINDING interfaces to libs2d (include/s2d.h).
!!
!! Every function returns an INTEGER(C_INT) status; S2D_Check turns a
!! nonzero status into ERROR STOP with the library's message, which keeps the
!! reference's "a failure is fatal" behaviour (the reference checks no
!! return codes at all).
MODULE S2D_C

   USE ISO_C_BINDING

   IMPLICIT NONE

   INTEGER(C_INT),PARAMETER,PUBLIC :: S2D_OK = 0
   INTEGER(C_INT),PARAMETER,PUBLIC :: S2D_FP64 = 0
   INTEGER(C_INT),PARAMETER,PUBLIC :: S2D_FP32 = 1
   INTEGER(C_INT),PARAMETER,PUBLIC :: S2D_FIELD_W = 0
   INTEGER(C_INT),PARAMETER,PUBLIC :: S2D_FIELD_PSI = 1
   INTEGER(C_INT),PARAMETER,PUBLIC :: S2D_FIELD_U = 2
   INTEGER(C_INT),PARAMETER,PUBLIC :: S2D_FIELD_V = 3
   INTEGER(C_INT),PARAMETER,PUBLIC :: S2D_FIELD_WS = 4
   INTEGER(C_INT),PARAMETER,PUBLIC :: S2D_FIELD_PSIS = 5
   INTEGER(C_INT),PARAMETER,PUBLIC :: S2D_FIELD_DW = 6
   INTEGER(C_INT),PARAMETER,PUBLIC :: S2D_FIELD_TMP = 7

   !> Mirror of s2d_diag.
   TYPE,BIND(C) :: s2d_diag_t
      REAL(C_DOUBLE) :: tke, enstrophy, palinstrophy, max_abs_w, time
      INTEGER(C_LONG_LONG) :: nadv
   END TYPE

   !> Mirror of s2d_grid (GetGridSize + GetAllocSize with one process).
   TYPE,BIND(C) :: s2d_grid_t
      INTEGER(C_INT) :: nx, ny, dealias
      INTEGER(C_INT) :: nxG, nyG
      INTEGER(C_INT) :: kxLG, kyLG, kxMG, kyMG
      INTEGER(C_INT) :: kxLU, kyLU, kxMU, kyMU
      REAL(C_DOUBLE) :: kMax
      INTEGER(C_INT) :: rISize, cISize
      INTEGER(C_INT) :: nxP, nyP, j1, j2, localJ, allocLocal
   END TYPE

   INTERFACE
      FUNCTION s2d_last_error(ctx) BIND(C, name='s2d_last_error')
         IMPORT :: C_PTR
         TYPE(C_PTR),VALUE :: ctx
         TYPE(C_PTR) :: s2d_last_error
      END FUNCTION
      INTEGER(C_INT) FUNCTION s2d_create(nx, ny, dealias, precision, ctx) &
            BIND(C, name='s2d_create')
         IMPORT :: C_INT, C_PTR
         INTEGER(C_INT),VALUE :: nx, ny, dealias, precision
         TYPE(C_PTR) :: ctx
      END FUNCTION
      INTEGER(C_INT) FUNCTION s2d_destroy(ctx) BIND(C, name='s2d_destroy')
         IMPORT :: C_INT, C_PTR
         TYPE(C_PTR),VALUE :: ctx
      END FUNCTION
      INTEGER(C_INT) FUNCTION s2d_set_field_host(ctx, field, buf) &
            BIND(C, name='s2d_set_field_host')
         IMPORT :: C_INT, C_PTR
         TYPE(C_PTR),VALUE :: ctx
         INTEGER(C_INT),VALUE :: field
         TYPE(*),DIMENSION(*) :: buf
      END FUNCTION
      INTEGER(C_INT) FUNCTION s2d_get_field_host(ctx, field, buf) &
            BIND(C, name='s2d_get_field_host')
         IMPORT :: C_INT, C_PTR
         TYPE(C_PTR),VALUE :: ctx
         INTEGER(C_INT),VALUE :: field
         TYPE(*),DIMENSION(*) :: buf
      END FUNCTION
      INTEGER(C_INT) FUNCTION s2d_transform_r2c(ctx, field) BIND(C, name='s2d_transform_r2c')
         IMPORT :: C_INT, C_PTR
         TYPE(C_PTR),VALUE :: ctx
         INTEGER(C_INT),VALUE :: field
      END FUNCTION
      INTEGER(C_INT) FUNCTION s2d_transform_c2r(ctx, field) BIND(C, name='s2d_transform_c2r')
         IMPORT :: C_INT, C_PTR
         TYPE(C_PTR),VALUE :: ctx
         INTEGER(C_INT),VALUE :: field
      END FUNCTION
      INTEGER(C_INT) FUNCTION s2d_dudx(ctx, fin, fout) BIND(C, name='s2d_dudx')
         IMPORT :: C_INT, C_PTR
         TYPE(C_PTR),VALUE :: ctx
         INTEGER(C_INT),VALUE :: fin, fout
      END FUNCTION
      INTEGER(C_INT) FUNCTION s2d_dudy(ctx, fin, fout) BIND(C, name='s2d_dudy')
         IMPORT :: C_INT, C_PTR
         TYPE(C_PTR),VALUE :: ctx
         INTEGER(C_INT),VALUE :: fin, fout
      END FUNCTION
      INTEGER(C_INT) FUNCTION s2d_truncate(ctx, field) BIND(C, name='s2d_truncate')
         IMPORT :: C_INT, C_PTR
         TYPE(C_PTR),VALUE :: ctx
         INTEGER(C_INT),VALUE :: field
      END FUNCTION
      INTEGER(C_INT) FUNCTION s2d_compute_psi(ctx, w, psi) BIND(C, name='s2d_compute_psi')
         IMPORT :: C_INT, C_PTR
         TYPE(C_PTR),VALUE :: ctx
         INTEGER(C_INT),VALUE :: w, psi
      END FUNCTION
      INTEGER(C_INT) FUNCTION s2d_compute_velocity(ctx, u, v, psi, trans) &
            BIND(C, name='s2d_compute_velocity')
         IMPORT :: C_INT, C_PTR
         TYPE(C_PTR),VALUE :: ctx
         INTEGER(C_INT),VALUE :: u, v, psi, trans
      END FUNCTION
      INTEGER(C_INT) FUNCTION s2d_compute_vorticity(ctx, u, v, w) &
            BIND(C, name='s2d_compute_vorticity')
         IMPORT :: C_INT, C_PTR
         TYPE(C_PTR),VALUE :: ctx
         INTEGER(C_INT),VALUE :: u, v, w
      END FUNCTION
      INTEGER(C_INT) FUNCTION s2d_compute_rhs(ctx, nu, u, v, w, psi, dW) &
            BIND(C, name='s2d_compute_rhs')
         IMPORT :: C_INT, C_PTR, C_DOUBLE
         TYPE(C_PTR),VALUE :: ctx
         REAL(C_DOUBLE),VALUE :: nu
         INTEGER(C_INT),VALUE :: u, v, w, psi, dW
      END FUNCTION
      INTEGER(C_INT) FUNCTION s2d_integrate_one_step(ctx, nu, dt) &
            BIND(C, name='s2d_integrate_one_step')
         IMPORT :: C_INT, C_PTR, C_DOUBLE
         TYPE(C_PTR),VALUE :: ctx
         REAL(C_DOUBLE),VALUE :: nu, dt
      END FUNCTION
      INTEGER(C_INT) FUNCTION s2d_rk3_steps(ctx, nsteps, nu, dt, cfl) &
            BIND(C, name='s2d_rk3_steps')
         IMPORT :: C_INT, C_PTR, C_DOUBLE
         TYPE(C_PTR),VALUE :: ctx
         INTEGER(C_INT),VALUE :: nsteps
         REAL(C_DOUBLE),VALUE :: nu, dt, cfl
      END FUNCTION
      INTEGER(C_INT) FUNCTION s2d_cfl_dt(ctx, cfl, dt) BIND(C, name='s2d_cfl_dt')
         IMPORT :: C_INT, C_PTR, C_DOUBLE
         TYPE(C_PTR),VALUE :: ctx
         REAL(C_DOUBLE),VALUE :: cfl
         REAL(C_DOUBLE) :: dt
      END FUNCTION
      INTEGER(C_INT) FUNCTION s2d_diagnostics(ctx, d) BIND(C, name='s2d_diagnostics')
         IMPORT :: C_INT, C_PTR, s2d_diag_t
         TYPE(C_PTR),VALUE :: ctx
         TYPE(s2d_diag_t) :: d
      END FUNCTION
      INTEGER(C_INT) FUNCTION s2d_set_turbulence(ctx, seed, k0, ke) &
            BIND(C, name='s2d_set_turbulence')
         IMPORT :: C_INT, C_PTR, C_DOUBLE, C_LONG_LONG
         TYPE(C_PTR),VALUE :: ctx
         INTEGER(C_LONG_LONG),VALUE :: seed
         REAL(C_DOUBLE),VALUE :: k0, ke
      END FUNCTION
      INTEGER(C_INT) FUNCTION s2d_get_grid_size(nx, ny, dealias, g) &
            BIND(C, name='s2d_get_grid_size')
         IMPORT :: C_INT, s2d_grid_t
         INTEGER(C_INT),VALUE :: nx, ny, dealias
         TYPE(s2d_grid_t) :: g
      END FUNCTION
      INTEGER(C_INT) FUNCTION s2d_wavenumbers(g, kxG, kyG) BIND(C, name='s2d_wavenumbers')
         IMPORT :: C_INT, s2d_grid_t
         TYPE(s2d_grid_t) :: g
         INTEGER(C_INT),DIMENSION(*) :: kxG, kyG
      END FUNCTION
      INTEGER(C_INT) FUNCTION s2d_spectrum_bins(g) BIND(C, name='s2d_spectrum_bins')
         IMPORT :: C_INT, s2d_grid_t
         TYPE(s2d_grid_t) :: g
      END FUNCTION
      INTEGER(C_INT) FUNCTION s2d_spectrum_uv(ctx, u, v, nu, ek, cnt, tke, eps, ediss) &
            BIND(C, name='s2d_spectrum_uv')
         IMPORT :: C_INT, C_PTR, C_DOUBLE
         TYPE(C_PTR),VALUE :: ctx
         INTEGER(C_INT),VALUE :: u, v
         REAL(C_DOUBLE),VALUE :: nu
         REAL(C_DOUBLE),DIMENSION(*) :: ek
         INTEGER(C_INT),DIMENSION(*) :: cnt
         REAL(C_DOUBLE) :: tke, eps, ediss
      END FUNCTION
      INTEGER(C_INT) FUNCTION s2d_get_grid(ctx, g) BIND(C, name='s2d_get_grid')
         IMPORT :: C_INT, C_PTR, s2d_grid_t
         TYPE(C_PTR),VALUE :: ctx
         TYPE(s2d_grid_t) :: g
      END FUNCTION
      ! libs2d_h5 (include/s2d_h5.h): the reference's output files
      INTEGER(C_INT) FUNCTION s2d_h5_write_rest(path, nxG, nyG, rISize, nadv, time, nu, &
                                                omega, psi, u, v) &
            BIND(C, name='s2d_h5_write_rest')
         IMPORT :: C_INT, C_CHAR, C_DOUBLE
         CHARACTER(KIND=C_CHAR),DIMENSION(*) :: path
         INTEGER(C_INT),VALUE :: nxG, nyG, rISize, nadv
         REAL(C_DOUBLE),VALUE :: time, nu
         REAL(C_DOUBLE),DIMENSION(*) :: omega, psi, u, v
      END FUNCTION
      INTEGER(C_INT) FUNCTION s2d_h5_write_grid(path, nxG, nyG) BIND(C, name='s2d_h5_write_grid')
         IMPORT :: C_INT, C_CHAR
         CHARACTER(KIND=C_CHAR),DIMENSION(*) :: path
         INTEGER(C_INT),VALUE :: nxG, nyG
      END FUNCTION
      INTEGER(C_INT) FUNCTION s2d_h5_write_xmf(path, nxG, nyG, num) &
            BIND(C, name='s2d_h5_write_xmf')
         IMPORT :: C_INT, C_CHAR
         CHARACTER(KIND=C_CHAR),DIMENSION(*) :: path
         INTEGER(C_INT),VALUE :: nxG, nyG, num
      END FUNCTION
      FUNCTION c_strlen(s) BIND(C, name='strlen')
         IMPORT :: C_PTR, C_SIZE_T
         TYPE(C_PTR),VALUE :: s
         INTEGER(C_SIZE_T) :: c_strlen
      END FUNCTION
   END INTERFACE

   PUBLIC :: S2D_Check

CONTAINS

   !> Stop with the library's message when a call failed.
   SUBROUTINE S2D_Check(rc, ctx, where)
      INTEGER(C_INT),INTENT(IN) :: rc
      TYPE(C_PTR),INTENT(IN) :: ctx
      CHARACTER(LEN=*),INTENT(IN) :: where
      CHARACTER(KIND=C_CHAR),DIMENSION(:),POINTER :: msg
      CHARACTER(LEN=512) :: text
      TYPE(C_PTR) :: p
      INTEGER :: i, n
      IF (rc == S2D_OK) RETURN
      p = s2d_last_error(ctx)
      n = INT(c_strlen(p))
      CALL C_F_POINTER(p, msg, [n])
      text = ''
      DO i = 1, MIN(n, LEN(text))
         text(i:i) = msg(i)
      END DO
      WRITE(*,'(A,A,A,I0,A,A)') 'libs2d failure in ', where, ' (status ', rc, '): ', TRIM(text)
      ERROR STOP 1
   END SUBROUTINE S2D_Check

END MODULE S2D_C
