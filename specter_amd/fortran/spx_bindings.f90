!=======================================================================
! Fortran 2003 interface to libspx (include/spx.h).
!
! A SPECTER build links libspx.so and USEs this module in place of the
! FFTW backend: device arrays are TYPE(C_PTR) handles obtained from
! spx_dev_alloc / spx_state_field; field layouts are the reference's own
! (nz,ny,nx/2+1) complex and (nx,ny,nz) real arrays.  Every function returns
! 0 on success; spx_check() reproduces the reference's abort-on-error
! behaviour (print + STOP).
!=======================================================================
MODULE spx
   USE, INTRINSIC :: iso_c_binding
   IMPLICIT NONE

   INTEGER(C_INT), PARAMETER :: SPX_FP64 = 0, SPX_FP32_MIXED = 1
   INTEGER(C_INT), PARAMETER :: SPX_SOLVER_HD = 0, SPX_SOLVER_BOUSS = 1, SPX_SOLVER_ROTBOUSS = 2
   INTEGER(C_INT), PARAMETER :: SPX_SOLVER_MHD = 3
   INTEGER(C_INT), PARAMETER :: SPX_BWALL_CONDUCTING = 0, SPX_BWALL_INSULATING = 1

   TYPE, BIND(C) :: spx_grid
      INTEGER(C_INT) :: nx, ny, nz, Cz, oz
      REAL(C_DOUBLE) :: Lx, Ly, Lz
      INTEGER(C_INT) :: precision
   END TYPE spx_grid

   TYPE, BIND(C) :: spx_params
      INTEGER(C_INT) :: solver, ord
      REAL(C_DOUBLE) :: dt, nu, kappa, xmom, xtemp
      REAL(C_DOUBLE) :: v_zsta(2), v_zend(2)
      REAL(C_DOUBLE) :: omega(3)
      REAL(C_DOUBLE) :: mu, b0(3)
      INTEGER(C_INT) :: bzsta, bzend
   END TYPE spx_params

   INTERFACE
      ! fcgram_create_plan / fcgram_destroy_plan (fftp/fcgram_mod.f90:49-177)
      INTEGER(C_INT) FUNCTION spx_plan_create(g, A, Q, stream, plan) BIND(C, name='spx_plan_create')
         IMPORT :: C_INT, C_DOUBLE, C_PTR, spx_grid
         TYPE(spx_grid), INTENT(IN) :: g
         REAL(C_DOUBLE), INTENT(IN) :: A(*), Q(*)
         TYPE(C_PTR), VALUE :: stream
         TYPE(C_PTR), INTENT(OUT) :: plan
      END FUNCTION
      INTEGER(C_INT) FUNCTION spx_plan_destroy(plan) BIND(C, name='spx_plan_destroy')
         IMPORT :: C_INT, C_PTR
         TYPE(C_PTR), VALUE :: plan
      END FUNCTION
      INTEGER(C_INT) FUNCTION spx_plan_sync(plan) BIND(C, name='spx_plan_sync')
         IMPORT :: C_INT, C_PTR
         TYPE(C_PTR), VALUE :: plan
      END FUNCTION
      ! fftp3d_real_to_complex / fftp3d_complex_to_real (fftp/fftp.fpp:388-425,789-821)
      INTEGER(C_INT) FUNCTION spx_fftp3d_real_to_complex(plan, rin, cout) &
            BIND(C, name='spx_fftp3d_real_to_complex')
         IMPORT :: C_INT, C_PTR
         TYPE(C_PTR), VALUE :: plan, rin, cout
      END FUNCTION
      INTEGER(C_INT) FUNCTION spx_fftp3d_complex_to_real(plan, cin, rout) &
            BIND(C, name='spx_fftp3d_complex_to_real')
         IMPORT :: C_INT, C_PTR
         TYPE(C_PTR), VALUE :: plan, cin, rout
      END FUNCTION
      ! fftp2d_*_xy (fftp/fftp.fpp:428-524,824-919)
      INTEGER(C_INT) FUNCTION spx_fftp2d_real_to_complex_xy(plan, rin, cout) &
            BIND(C, name='spx_fftp2d_real_to_complex_xy')
         IMPORT :: C_INT, C_PTR
         TYPE(C_PTR), VALUE :: plan, rin, cout
      END FUNCTION
      INTEGER(C_INT) FUNCTION spx_fftp2d_complex_to_real_xy(plan, cin, rout) &
            BIND(C, name='spx_fftp2d_complex_to_real_xy')
         IMPORT :: C_INT, C_PTR
         TYPE(C_PTR), VALUE :: plan, cin, rout
      END FUNCTION
      ! non-periodic y: fcgram_create_plan with Cy > 0 (fftp/fcgram_mod.f90:49-144)
      INTEGER(C_INT) FUNCTION spx_plan_create_fcy(g, A, Q, Cy, oy, Ay, Qy, nmembers, stream, &
            plan) BIND(C, name='spx_plan_create_fcy')
         IMPORT :: C_INT, C_DOUBLE, C_PTR, spx_grid
         TYPE(spx_grid), INTENT(IN) :: g
         REAL(C_DOUBLE), INTENT(IN) :: A(*), Q(*), Ay(*), Qy(*)
         INTEGER(C_INT), VALUE :: Cy, oy, nmembers
         TYPE(C_PTR), VALUE :: stream
         TYPE(C_PTR), INTENT(OUT) :: plan
      END FUNCTION
      ! fftp1d_*_x, fftp2d_*_yz (fftp/fftp.fpp:527-716,922-1056)
      INTEGER(C_INT) FUNCTION spx_fftp1d_real_to_complex_x(plan, rin, cout) &
            BIND(C, name='spx_fftp1d_real_to_complex_x')
         IMPORT :: C_INT, C_PTR
         TYPE(C_PTR), VALUE :: plan, rin, cout
      END FUNCTION
      INTEGER(C_INT) FUNCTION spx_fftp2d_real_to_complex_yz(plan, c) &
            BIND(C, name='spx_fftp2d_real_to_complex_yz')
         IMPORT :: C_INT, C_PTR
         TYPE(C_PTR), VALUE :: plan, c
      END FUNCTION
      INTEGER(C_INT) FUNCTION spx_fftp2d_complex_to_real_yz(plan, c) &
            BIND(C, name='spx_fftp2d_complex_to_real_yz')
         IMPORT :: C_INT, C_PTR
         TYPE(C_PTR), VALUE :: plan, c
      END FUNCTION
      INTEGER(C_INT) FUNCTION spx_fftp1d_complex_to_real_x(plan, cin, rout) &
            BIND(C, name='spx_fftp1d_complex_to_real_x')
         IMPORT :: C_INT, C_PTR
         TYPE(C_PTR), VALUE :: plan, cin, rout
      END FUNCTION
      ! fftp1d_*_z (fftp/fftp.fpp:720-786,1060-1094)
      INTEGER(C_INT) FUNCTION spx_fftp1d_real_to_complex_z(plan, c) &
            BIND(C, name='spx_fftp1d_real_to_complex_z')
         IMPORT :: C_INT, C_PTR
         TYPE(C_PTR), VALUE :: plan, c
      END FUNCTION
      INTEGER(C_INT) FUNCTION spx_fftp1d_complex_to_real_z(plan, c) &
            BIND(C, name='spx_fftp1d_complex_to_real_z')
         IMPORT :: C_INT, C_PTR
         TYPE(C_PTR), VALUE :: plan, c
      END FUNCTION
      ! operators (pseudo/pseudospec_hd.f90)
      INTEGER(C_INT) FUNCTION spx_derivk(plan, a, b, dir) BIND(C, name='spx_derivk')
         IMPORT :: C_INT, C_PTR
         TYPE(C_PTR), VALUE :: plan, a, b
         INTEGER(C_INT), VALUE :: dir
      END FUNCTION
      INTEGER(C_INT) FUNCTION spx_laplak(plan, a, b) BIND(C, name='spx_laplak')
         IMPORT :: C_INT, C_PTR
         TYPE(C_PTR), VALUE :: plan, a, b
      END FUNCTION
      INTEGER(C_INT) FUNCTION spx_gradre(plan, a, b, c, d, e, f) BIND(C, name='spx_gradre')
         IMPORT :: C_INT, C_PTR
         TYPE(C_PTR), VALUE :: plan, a, b, c, d, e, f
      END FUNCTION
      ! MHD: Neumann tables, prodre / vector, magnetic BCs (bboundary.f90:100-189)
      INTEGER(C_INT) FUNCTION spx_plan_load_neumann(plan, ord, Q, Qn) &
            BIND(C, name='spx_plan_load_neumann')
         IMPORT :: C_INT, C_PTR, C_DOUBLE
         TYPE(C_PTR), VALUE :: plan
         INTEGER(C_INT), VALUE :: ord
         REAL(C_DOUBLE), INTENT(IN) :: Q(*), Qn(*)
      END FUNCTION
      INTEGER(C_INT) FUNCTION spx_prodre(plan, a, b, c, d, e, f) BIND(C, name='spx_prodre')
         IMPORT :: C_INT, C_PTR
         TYPE(C_PTR), VALUE :: plan, a, b, c, d, e, f
      END FUNCTION
      INTEGER(C_INT) FUNCTION spx_vector(plan, a, b, c, d, e, f, g, h, i) &
            BIND(C, name='spx_vector')
         IMPORT :: C_INT, C_PTR
         TYPE(C_PTR), VALUE :: plan, a, b, c, d, e, f, g, h, i
      END FUNCTION
      INTEGER(C_INT) FUNCTION spx_a_imposebc_and_project(plan, ax, ay, az, ph, bsta, bend) &
            BIND(C, name='spx_a_imposebc_and_project')
         IMPORT :: C_INT, C_PTR
         TYPE(C_PTR), VALUE :: plan, ax, ay, az, ph
         INTEGER(C_INT), VALUE :: bsta, bend
      END FUNCTION
      ! boundary/vboundary.f90:67-151
      INTEGER(C_INT) FUNCTION spx_v_imposebc_and_project(plan, vx, vy, vz, pr, rki, ord, vzsta, vzend) &
            BIND(C, name='spx_v_imposebc_and_project')
         IMPORT :: C_INT, C_PTR, C_DOUBLE
         TYPE(C_PTR), VALUE :: plan, vx, vy, vz, pr
         INTEGER(C_INT), VALUE :: rki, ord
         REAL(C_DOUBLE), INTENT(IN) :: vzsta(2), vzend(2)
      END FUNCTION
      ! diagnostics (pseudo/pseudospec_hd.f90:405-1235)
      INTEGER(C_INT) FUNCTION spx_energy(plan, a, b, c, kin, out) BIND(C, name='spx_energy')
         IMPORT :: C_INT, C_PTR, C_DOUBLE
         TYPE(C_PTR), VALUE :: plan, a, b, c
         INTEGER(C_INT), VALUE :: kin
         REAL(C_DOUBLE), INTENT(OUT) :: out
      END FUNCTION
      INTEGER(C_INT) FUNCTION spx_cross(plan, a, b, c, d, e, f, out) BIND(C, name='spx_cross')
         IMPORT :: C_INT, C_PTR, C_DOUBLE
         TYPE(C_PTR), VALUE :: plan, a, b, c, d, e, f
         REAL(C_DOUBLE), INTENT(OUT) :: out
      END FUNCTION
      INTEGER(C_INT) FUNCTION spx_helicity(plan, a, b, c, out) BIND(C, name='spx_helicity')
         IMPORT :: C_INT, C_PTR, C_DOUBLE
         TYPE(C_PTR), VALUE :: plan, a, b, c
         REAL(C_DOUBLE), INTENT(OUT) :: out
      END FUNCTION
      INTEGER(C_INT) FUNCTION spx_divergence(plan, a, b, c, out) BIND(C, name='spx_divergence')
         IMPORT :: C_INT, C_PTR, C_DOUBLE
         TYPE(C_PTR), VALUE :: plan, a, b, c
         REAL(C_DOUBLE), INTENT(OUT) :: out
      END FUNCTION
      INTEGER(C_INT) FUNCTION spx_bouncheck_z(plan, a, b, bot, top) BIND(C, name='spx_bouncheck_z')
         IMPORT :: C_INT, C_PTR, C_DOUBLE
         TYPE(C_PTR), VALUE :: plan, a, b
         REAL(C_DOUBLE), INTENT(OUT) :: bot, top
      END FUNCTION
      ! scalar diagnostics (pseudospec_phd.f90:116-235) and the Nusselt number
      INTEGER(C_INT) FUNCTION spx_variance(plan, a, kin, out) BIND(C, name='spx_variance')
         IMPORT :: C_INT, C_PTR, C_DOUBLE
         TYPE(C_PTR), VALUE :: plan, a
         INTEGER(C_INT), VALUE :: kin
         REAL(C_DOUBLE), INTENT(OUT) :: out
      END FUNCTION
      INTEGER(C_INT) FUNCTION spx_product(plan, a, b, out) BIND(C, name='spx_product')
         IMPORT :: C_INT, C_PTR, C_DOUBLE
         TYPE(C_PTR), VALUE :: plan, a, b
         REAL(C_DOUBLE), INTENT(OUT) :: out
      END FUNCTION
      INTEGER(C_INT) FUNCTION spx_nusselt(plan, th, out) BIND(C, name='spx_nusselt')
         IMPORT :: C_INT, C_PTR, C_DOUBLE
         TYPE(C_PTR), VALUE :: plan, th
         REAL(C_DOUBLE), INTENT(OUT) :: out(2)
      END FUNCTION
      ! device-resident solver state (specter.fpp:1002-1165 RK loop)
      INTEGER(C_INT) FUNCTION spx_state_create(plan, prm, state) BIND(C, name='spx_state_create')
         IMPORT :: C_INT, C_PTR, spx_params
         TYPE(C_PTR), VALUE :: plan
         TYPE(spx_params), INTENT(IN) :: prm
         TYPE(C_PTR), INTENT(OUT) :: state
      END FUNCTION
      INTEGER(C_INT) FUNCTION spx_state_destroy(state) BIND(C, name='spx_state_destroy')
         IMPORT :: C_INT, C_PTR
         TYPE(C_PTR), VALUE :: state
      END FUNCTION
      INTEGER(C_INT) FUNCTION spx_state_field(state, name, ptr) BIND(C, name='spx_state_field')
         IMPORT :: C_INT, C_PTR, C_CHAR
         TYPE(C_PTR), VALUE :: state
         CHARACTER(KIND=C_CHAR), INTENT(IN) :: name(*)
         TYPE(C_PTR), INTENT(OUT) :: ptr
      END FUNCTION
      INTEGER(C_INT) FUNCTION spx_step(plan, state, nsteps) BIND(C, name='spx_step')
         IMPORT :: C_INT, C_PTR
         TYPE(C_PTR), VALUE :: plan, state
         INTEGER(C_INT), VALUE :: nsteps
      END FUNCTION
      ! per-kernel timing with HIP events (benchmark.txt)
      INTEGER(C_INT) FUNCTION spx_prof_enable(plan, on) BIND(C, name='spx_prof_enable')
         IMPORT :: C_INT, C_PTR
         TYPE(C_PTR), VALUE :: plan
         INTEGER(C_INT), VALUE :: on
      END FUNCTION
      INTEGER(C_INT) FUNCTION spx_prof_resolve(plan, nk) BIND(C, name='spx_prof_resolve')
         IMPORT :: C_INT, C_PTR
         TYPE(C_PTR), VALUE :: plan
         INTEGER(C_INT), INTENT(OUT) :: nk
      END FUNCTION
      INTEGER(C_INT) FUNCTION spx_prof_get(plan, idx, name, length, launches, total_ms, &
                                           total_bytes) BIND(C, name='spx_prof_get')
         IMPORT :: C_INT, C_PTR, C_CHAR, C_SIZE_T, C_LONG, C_DOUBLE
         TYPE(C_PTR), VALUE :: plan
         INTEGER(C_INT), VALUE :: idx
         CHARACTER(KIND=C_CHAR), INTENT(OUT) :: name(*)
         INTEGER(C_SIZE_T), VALUE :: length
         INTEGER(C_LONG), INTENT(OUT) :: launches
         REAL(C_DOUBLE), INTENT(OUT) :: total_ms, total_bytes
      END FUNCTION
      ! memory helpers
      INTEGER(C_INT) FUNCTION spx_dev_alloc(bytes, ptr) BIND(C, name='spx_dev_alloc')
         IMPORT :: C_INT, C_PTR, C_SIZE_T
         INTEGER(C_SIZE_T), VALUE :: bytes
         TYPE(C_PTR), INTENT(OUT) :: ptr
      END FUNCTION
      INTEGER(C_INT) FUNCTION spx_dev_free(ptr) BIND(C, name='spx_dev_free')
         IMPORT :: C_INT, C_PTR
         TYPE(C_PTR), VALUE :: ptr
      END FUNCTION
      INTEGER(C_INT) FUNCTION spx_memcpy_h2d(plan, dst, src, bytes) BIND(C, name='spx_memcpy_h2d')
         IMPORT :: C_INT, C_PTR, C_SIZE_T
         TYPE(C_PTR), VALUE :: plan, dst, src
         INTEGER(C_SIZE_T), VALUE :: bytes
      END FUNCTION
      INTEGER(C_INT) FUNCTION spx_memcpy_d2h(plan, dst, src, bytes) BIND(C, name='spx_memcpy_d2h')
         IMPORT :: C_INT, C_PTR, C_SIZE_T
         TYPE(C_PTR), VALUE :: plan, dst, src
         INTEGER(C_SIZE_T), VALUE :: bytes
      END FUNCTION
      INTEGER(C_INT) FUNCTION spx_memcpy_d2d(plan, dst, src, bytes) BIND(C, name='spx_memcpy_d2d')
         IMPORT :: C_INT, C_PTR, C_SIZE_T
         TYPE(C_PTR), VALUE :: plan, dst, src
         INTEGER(C_SIZE_T), VALUE :: bytes
      END FUNCTION
      INTEGER(C_INT) FUNCTION spx_last_error(buf, length) BIND(C, name='spx_last_error')
         IMPORT :: C_INT, C_CHAR, C_SIZE_T
         CHARACTER(KIND=C_CHAR), INTENT(OUT) :: buf(*)
         INTEGER(C_SIZE_T), VALUE :: length
      END FUNCTION
   END INTERFACE

CONTAINS

   ! Abort on error, as the reference does (e.g. fcgram_mod.f90:114-117).
   SUBROUTINE spx_check(rc, what)
      INTEGER(C_INT), INTENT(IN) :: rc
      CHARACTER(len=*), INTENT(IN) :: what
      CHARACTER(KIND=C_CHAR) :: buf(512)
      CHARACTER(len=512) :: msg
      INTEGER(C_INT) :: n
      INTEGER :: i
      IF (rc .eq. 0) RETURN
      n = spx_last_error(buf, INT(512, C_SIZE_T))
      msg = ' '
      DO i = 1, MIN(INT(n), 512)
         msg(i:i) = buf(i)
      END DO
      PRINT *, '[ERROR] ', what, ': ', TRIM(msg)
      STOP 1
   END SUBROUTINE spx_check

END MODULE spx
