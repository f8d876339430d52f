!=======================================================================
! SPX_HD: SPECTER-style HD driver over libspx (device-resident state).
!
! Reads SPECTER's parameter.inp namelists (&status, &boxparams, &parameter,
! &velocity, &velbound; specter.fpp:201-239) plus a &grid group that stands
! in for the compile-time NX/NY/NZ/CZ/OZ/ORD of Makefile.in (and prec:
! 0 = float64, 1 = float32 nonlinear term, the compile-time GP choice).  Starts like a
! reference restart with stat >= 1 (specter.fpp:885-934): the physical-space
! fields vx, vy, vz (and optionally fx, fy, fz, pr) are read from
! idir/<name>.NNNN.out as raw float64 arrays (nx, ny, nz-Cz) in Fortran order
! (mpiio/binary_io.f90:127-151).  The RK loop (specter.fpp:1002-1165) runs on
! the GPU through libspx with the reference's restart counters
! (specter.fpp:885-891): every cstep steps the global quantities of
! hd_global.f90 are appended to balance.txt (format of pseudospec_hd.f90:993)
! and the no-slip diagnostics to noslip_diagnostic.txt at full precision;
! every tstep steps (bench = 0) the fields are written as in
! specter.fpp:1003-1053: vx, vy, vz (and wx, wy, wz when outs >= 1) as
! v/(nx ny nz) and pr as p'/(nx ny dt) on the physical planes, to
! odir/<name>.MMMM.out.  x.txt, y.txt, z.txt hold the grid
! (specter.fpp:752-765).  mult divides dt and multiplies step, tstep and
! cstep (specter.fpp:367-370); stat = 1 writes the initial vx, vy, vz as
! the 0000 outputs (specter.fpp:920-932); helicity.txt is appended with the
! balance.txt line (hdcheck with hel = 1, pseudospec_hd.f90:980-1003).
! bench = 1 (2) skips all outputs and global quantities, times the RK loop
! and appends one line to benchmark.txt in the reference's column layout
! (specter.fpp:1182-1228; with bench = 2 also the plan-creation time):
! TCPU, TOMP and TWTIME per step from the CPU clock and the wall clock
! around the loop (which runs one captured HIP graph per step), TFFT and
! TTOT per step from HIP events around every kernel of one extra profiled
! step (TFFT: the line-transform kernels, which fuse the continuation, the
! transposes, the products and the wall update; TTOT: all kernels); TTRA,
! TCOM, TCONT, TNEU and TROB are 0 (fused into those kernels, or no MPI).
!
! solver = 'BOUSS' or 'ROTBOUSS' in &grid (the reference's compile-time
! SOLVER choice) runs the Boussinesq solver, ROTBOUSS with the Coriolis
! term of &rotation (omegax, omegay, omegaz; specter.fpp:662-677): &boussinesq (gama, xmom, xtemp; the
! coefficients used are gama*xmom and gama*xtemp, specter.fpp:476-491) and
! &scalar (kappa) are read, th (and optionally fs) start from
! th.NNNN.out / fs.NNNN.out, th is written every tstep steps, and every
! cstep steps scalar.txt (pseudospec_phd.f90:237-262 format),
! scalar_constant_diagnostic.txt (<th^2> at the walls, sboundary.f90:167-210,
! at full precision) and nusselt.txt (Nu at z = 0 and z = Lz for theta = 1 - z/Lz + th; the
! reference computes no Nusselt number) are appended.
!
! solver = 'MHD' runs the MHD solver (include/mhd/mhd_rkstep2.f90): &magfield
! (mu; m0, a0, mkdn, mkup and the mparam/aparam values are read and unused),
! &magbound (bbczsta, bbczend: 'conducting' or 'vacuum'; x and y periodic) and
! &uniformb (bx0, by0, bz0) are read, the Neumann tables Q1n{oz}.dat and
! Q2n{oz}.dat come from tdir, the vector potential ax, ay, az (and
! optionally mx, my, mz and ph, ph as p'/dt) start from idir like v; every
! tstep steps ax, ay, az (bx, by, bz when outs >= 1, laplak(a) as jx, jy, jz
! when outs = 2) and ph/(nx ny dt) are written (specter.fpp:1070-1127); every
! cstep steps the mhdcheck files are appended with hel = crs = 1
! (pseudospec_mhd.f90 mhdcheck): balance.txt (eng, ens, cur), energy.txt
! (engk, engm), helicity.txt (helk, helm), cross.txt (crh, asq).
!=======================================================================
PROGRAM spx_hd
   USE, INTRINSIC :: iso_c_binding
   USE spx
   IMPLICIT NONE

   INTEGER :: nx = 64, ny = 64, nz = 89, Cz = 25, oz = 5, ord = 2, prec = 0
   CHARACTER(len=10) :: solver = 'HD'
   REAL(8) :: gama = 1.d0, xmom = 1.d0, xtemp = 1.d0
   REAL(8) :: omegax = 0.d0, omegay = 0.d0, omegaz = 0.d0
   REAL(8) :: c0 = 0.d0, s0 = 0.d0, skdn = 1.d0, skup = 2.d0, kappa = 1.d-3
   CHARACTER(len=100) :: idir = 'in/', odir = 'out/', tdir = 'tables/'
   INTEGER :: stat = 1, mult = 1, bench = 0, outs = 0
   REAL(8) :: Lx = 1.d0, Ly = 1.d0, Lz = 1.d0
   REAL(8) :: dt = 1.d-3
   INTEGER :: step = 10, tstep = 100, cstep = 5, seed = 1000
   REAL(8) :: f0 = 0.d0, u0 = 1.d0, kdn = 2.d0, kup = 4.d0, nu = 1.d-3
   CHARACTER(len=15) :: vbczsta = 'noslip', vbczend = 'noslip'
   REAL(8) :: vxzsta = 0.d0, vxzend = 0.d0, vyzsta = 0.d0, vyzend = 0.d0
   REAL(8) :: m0 = 0.d0, a0 = 1.d0, mkdn = 2.d0, mkup = 4.d0, mu = 1.d-3
   REAL(8) :: mparam0, mparam1, mparam2, mparam3, mparam4, mparam5, mparam6, mparam7
   REAL(8) :: mparam8, mparam9, aparam0, aparam1, aparam2, aparam3, aparam4, aparam5
   REAL(8) :: aparam6, aparam7, aparam8, aparam9
   REAL(8) :: bx0 = 0.d0, by0 = 0.d0, bz0 = 0.d0
   CHARACTER(len=15) :: bbcxsta = 'periodic', bbcxend = 'periodic', bbcysta = 'periodic'
   CHARACTER(len=15) :: bbcyend = 'periodic', bbczsta = 'conducting', bbczend = 'conducting'
   NAMELIST /magfield/ m0, a0, mkdn, mkup, mu, mparam0, mparam1, mparam2, mparam3, &
      mparam4, mparam5, mparam6, mparam7, mparam8, mparam9, aparam0, aparam1, aparam2, &
      aparam3, aparam4, aparam5, aparam6, aparam7, aparam8, aparam9
   NAMELIST /magbound/ bbcxsta, bbcxend, bbcysta, bbcyend, bbczsta, bbczend
   NAMELIST /uniformb/ bx0, by0, bz0
   NAMELIST /grid/ nx, ny, nz, Cz, oz, ord, prec, solver
   NAMELIST /boussinesq/ gama, xmom, xtemp
   NAMELIST /rotation/ omegax, omegay, omegaz
   NAMELIST /scalar/ c0, s0, skdn, skup, kappa
   NAMELIST /status/ idir, odir, tdir, stat, mult, bench, outs
   NAMELIST /boxparams/ Lx, Ly, Lz
   NAMELIST /parameter/ dt, step, tstep, cstep, seed
   NAMELIST /velocity/ f0, u0, kdn, kup, nu
   NAMELIST /velbound/ vbczsta, vbczend, vxzsta, vxzend, vyzsta, vyzend

   TYPE(spx_grid) :: g
   TYPE(spx_params) :: prm
   TYPE(C_PTR) :: plan, state, dreal, dtmp
   TYPE(C_PTR) :: vx, vy, vz, pr, fx, fy, fz, th, fs
   TYPE(C_PTR) :: ax, ay, az, ph, mx, my, mz, dtmp2, dtmp3
   LOGICAL :: bouss, rot, mhd
   INTEGER :: bzs = 0, bze = 0
   REAL(8) :: engk, engm, cur, helm, crh, asq
   REAL(8), ALLOCATABLE, TARGET :: Qn(:)
   REAL(8) :: sv, sg, sp, nus(2)
   REAL(8), ALLOCATABLE, TARGET :: R(:, :, :), A(:, :), Q(:, :)
   REAL(8) :: eng, ens, pot, dv, tb, tt, nb, nt
   INTEGER :: t, timec, timet, tind, ini, np, ios, nk, ik
   REAL(8) :: khe, cpu0, cpu1, cpup0, cpup1, tfft, ttot, tms, tby
   INTEGER(8) :: clk0, clk1, clkp0, clkp1, clkrate
   INTEGER(C_LONG) :: nlaunch
   CHARACTER(KIND=C_CHAR) :: kname(64)
   CHARACTER(len=64) :: kstr
   LOGICAL :: bbenchexist
   REAL(8), ALLOCATABLE, TARGET :: R2(:, :, :)
   INTEGER(C_SIZE_T) :: rbytes, cbytes
   INTEGER :: i_
   CHARACTER(len=4) :: ext
   LOGICAL :: have

   OPEN(1, file='parameter.inp', status='old', form='formatted')
   READ(1, NML=grid)
   REWIND(1)
   READ(1, NML=status)
   REWIND(1)
   READ(1, NML=boxparams)
   REWIND(1)
   READ(1, NML=parameter)
   dt = dt / REAL(mult, 8)
   step = step * mult
   tstep = tstep * mult
   cstep = cstep * mult
   REWIND(1)
   READ(1, NML=velocity)
   REWIND(1)
   READ(1, NML=velbound, IOSTAT=ios)
   rot = TRIM(solver) .eq. 'ROTBOUSS'
   bouss = TRIM(solver) .eq. 'BOUSS' .OR. rot
   IF (bouss) THEN
      REWIND(1)
      READ(1, NML=boussinesq)
      REWIND(1)
      READ(1, NML=scalar)
      xmom = xmom * gama
      xtemp = xtemp * gama
      IF (rot) THEN
         REWIND(1)
         READ(1, NML=rotation)
      ENDIF
   ENDIF
   mhd = TRIM(solver) .eq. 'MHD'
   IF (mhd) THEN
      REWIND(1)
      READ(1, NML=magfield)
      REWIND(1)
      READ(1, NML=magbound)
      REWIND(1)
      READ(1, NML=uniformb, IOSTAT=ios)
      bzs = bkind(bbczsta)
      bze = bkind(bbczend)
   ELSE IF (.NOT. bouss .AND. TRIM(solver) .ne. 'HD') THEN
      PRINT *, '[ERROR] solver must be HD, BOUSS, ROTBOUSS or MHD. Aborting...'
      STOP 1
   ENDIF
   CLOSE(1)
   IF (TRIM(vbczsta) .ne. 'noslip' .OR. TRIM(vbczend) .ne. 'noslip') THEN
      PRINT *, '[ERROR] only no-slip walls in z are supported. Aborting...'
      STOP 1
   ENDIF
   IF (stat .lt. 1) THEN
      PRINT *, '[ERROR] SPX_HD starts from binary fields (stat >= 1). Aborting...'
      STOP 1
   ENDIF
   np = nz - Cz

   ! FC-Gram tables (fcgram_mod.f90:221-244)
   ALLOCATE(A(Cz, oz), Q(oz, oz))
   CALL read_table(TRIM(tdir)//'/A'//itoa(Cz)//'-'//itoa(oz)//'.dat', A, Cz*oz)
   CALL read_table(TRIM(tdir)//'/Q'//itoa(oz)//'.dat', Q, oz*oz)

   g = spx_grid(nx, ny, nz, Cz, oz, Lx, Ly, Lz, prec)
   CALL CPU_TIME(cpup0)
   CALL SYSTEM_CLOCK(clkp0, clkrate)
   CALL spx_check(spx_plan_create(g, A, Q, C_NULL_PTR, plan), 'spx_plan_create')
   IF (mhd) THEN
      ! load_neumann_tables (fcgram_mod.f90:260-356) for the magnetic walls
      ALLOCATE(Qn(1 + oz * oz))
      CALL read_table(TRIM(tdir)//'/Q1n'//itoa(oz)//'.dat', Qn, 1 + oz * oz)
      CALL spx_check(spx_plan_load_neumann(plan, 1, Q, Qn), 'Neumann tables 1')
      CALL read_table(TRIM(tdir)//'/Q2n'//itoa(oz)//'.dat', Qn, 1 + oz * oz)
      CALL spx_check(spx_plan_load_neumann(plan, 2, Q, Qn), 'Neumann tables 2')
   ENDIF
   CALL spx_check(spx_plan_sync(plan), 'sync')
   CALL CPU_TIME(cpup1)
   CALL SYSTEM_CLOCK(clkp1)
   IF (mhd) THEN
      prm = spx_params(SPX_SOLVER_MHD, ord, dt, nu, 0.d0, 0.d0, 0.d0, &
                       (/vxzsta, vyzsta/), (/vxzend, vyzend/), (/0.d0, 0.d0, 0.d0/), &
                       mu, (/bx0, by0, bz0/), bzs, bze)
   ELSE IF (rot) THEN
      prm = spx_params(SPX_SOLVER_ROTBOUSS, ord, dt, nu, kappa, xmom, xtemp, &
                       (/vxzsta, vyzsta/), (/vxzend, vyzend/), (/omegax, omegay, omegaz/), &
                       0.d0, (/0.d0, 0.d0, 0.d0/), 0, 0)
   ELSE IF (bouss) THEN
      prm = spx_params(SPX_SOLVER_BOUSS, ord, dt, nu, kappa, xmom, xtemp, &
                       (/vxzsta, vyzsta/), (/vxzend, vyzend/), (/0.d0, 0.d0, 0.d0/), &
                       0.d0, (/0.d0, 0.d0, 0.d0/), 0, 0)
   ELSE
      prm = spx_params(SPX_SOLVER_HD, ord, dt, nu, 0.d0, 0.d0, 0.d0, &
                       (/vxzsta, vyzsta/), (/vxzend, vyzend/), (/0.d0, 0.d0, 0.d0/), &
                       0.d0, (/0.d0, 0.d0, 0.d0/), 0, 0)
   ENDIF
   CALL spx_check(spx_state_create(plan, prm, state), 'spx_state_create')
   CALL spx_check(spx_state_field(state, 'vx'//C_NULL_CHAR, vx), 'field vx')
   CALL spx_check(spx_state_field(state, 'vy'//C_NULL_CHAR, vy), 'field vy')
   CALL spx_check(spx_state_field(state, 'vz'//C_NULL_CHAR, vz), 'field vz')
   CALL spx_check(spx_state_field(state, 'pr'//C_NULL_CHAR, pr), 'field pr')
   CALL spx_check(spx_state_field(state, 'fx'//C_NULL_CHAR, fx), 'field fx')
   CALL spx_check(spx_state_field(state, 'fy'//C_NULL_CHAR, fy), 'field fy')
   CALL spx_check(spx_state_field(state, 'fz'//C_NULL_CHAR, fz), 'field fz')
   IF (bouss) THEN
      CALL spx_check(spx_state_field(state, 'th'//C_NULL_CHAR, th), 'field th')
      CALL spx_check(spx_state_field(state, 'fs'//C_NULL_CHAR, fs), 'field fs')
   ENDIF
   IF (mhd) THEN
      CALL spx_check(spx_state_field(state, 'ax'//C_NULL_CHAR, ax), 'field ax')
      CALL spx_check(spx_state_field(state, 'ay'//C_NULL_CHAR, ay), 'field ay')
      CALL spx_check(spx_state_field(state, 'az'//C_NULL_CHAR, az), 'field az')
      CALL spx_check(spx_state_field(state, 'ph'//C_NULL_CHAR, ph), 'field ph')
      CALL spx_check(spx_state_field(state, 'mx'//C_NULL_CHAR, mx), 'field mx')
      CALL spx_check(spx_state_field(state, 'my'//C_NULL_CHAR, my), 'field my')
      CALL spx_check(spx_state_field(state, 'mz'//C_NULL_CHAR, mz), 'field mz')
   ENDIF

   ALLOCATE(R(nx, ny, nz))
   rbytes = INT(nx, C_SIZE_T) * ny * nz * 8
   cbytes = INT(nz, C_SIZE_T) * ny * (nx / 2 + 1) * 16
   CALL spx_check(spx_dev_alloc(rbytes, dreal), 'spx_dev_alloc')
   CALL spx_check(spx_dev_alloc(cbytes, dtmp), 'spx_dev_alloc')
   IF (mhd) THEN
      CALL spx_check(spx_dev_alloc(cbytes, dtmp2), 'spx_dev_alloc')
      CALL spx_check(spx_dev_alloc(cbytes, dtmp3), 'spx_dev_alloc')
   ENDIF

   ! restart: physical fields -> fftp3d_real_to_complex (specter.fpp:894-899)
   WRITE(ext, '(i4.4)') stat
   CALL load_field('vx', .TRUE., vx, have)
   CALL load_field('vy', .TRUE., vy, have)
   CALL load_field('vz', .TRUE., vz, have)
   CALL load_field('fx', .FALSE., fx, have)
   CALL load_field('fy', .FALSE., fy, have)
   CALL load_field('fz', .FALSE., fz, have)
   IF (bouss) THEN
      CALL load_field('th', .TRUE., th, have)
      CALL load_field('fs', .FALSE., fs, have)
   ENDIF
   IF (mhd) THEN
      CALL load_field('ax', .TRUE., ax, have)
      CALL load_field('ay', .TRUE., ay, have)
      CALL load_field('az', .TRUE., az, have)
      CALL load_field('mx', .FALSE., mx, have)
      CALL load_field('my', .FALSE., my, have)
      CALL load_field('mz', .FALSE., mz, have)
      ! electric potential phi' = dt*phi in the mixed domain (specter.fpp:949-958)
      CALL read_raw(TRIM(idir)//'/ph.'//ext//'.out', have)
      IF (have) THEN
         R(:, :, np+1:nz) = 0.d0
         R = R * dt
         CALL spx_check(spx_memcpy_h2d(plan, dreal, C_LOC(R), rbytes), 'h2d ph')
         CALL spx_check(spx_fftp2d_real_to_complex_xy(plan, dreal, ph), 'ph r2c')
      ENDIF
   ENDIF
   ! pressure p' = dt*p in the mixed domain (specter.fpp:901-910)
   CALL read_raw(TRIM(idir)//'/pr.'//ext//'.out', have)
   IF (have) THEN
      R(:, :, np+1:nz) = 0.d0
      R = R * dt
      CALL spx_check(spx_memcpy_h2d(plan, dreal, C_LOC(R), rbytes), 'h2d pr')
      CALL spx_check(spx_fftp2d_real_to_complex_xy(plan, dreal, pr), 'pr r2c')
   ENDIF

   CALL write_grid()

   ! initial fields as the 0000 outputs (specter.fpp:920-932)
   IF (stat .eq. 1) THEN
      ext = '0000'
      CALL store_field('vx', vx)
      CALL store_field('vy', vy)
      CALL store_field('vz', vz)
      WRITE(ext, '(i4.4)') stat
   ENDIF

   ! RK loop with restart counters (specter.fpp:885-891, 1002-1165)
   ini = (stat - 1) * tstep + 1
   tind = stat
   timet = 0
   timec = MODULO(ini - 1, cstep)
   OPEN(2, file='balance.txt', position='append')
   OPEN(3, file='noslip_diagnostic.txt', position='append')
   OPEN(9, file='helicity.txt', position='append')
   IF (bouss) THEN
      OPEN(4, file='scalar.txt', position='append')
      OPEN(7, file='scalar_constant_diagnostic.txt', position='append')
      OPEN(8, file='nusselt.txt', position='append')
   ENDIF
   IF (mhd) THEN
      OPEN(5, file='energy.txt', position='append')
      OPEN(6, file='cross.txt', position='append')
   ENDIF
   IF (bench .ge. 1) THEN
      CALL spx_check(spx_plan_sync(plan), 'sync')
      CALL CPU_TIME(cpu0)
      CALL SYSTEM_CLOCK(clk0)
   ENDIF
   DO t = ini, step
      IF (timet .eq. tstep .AND. bench .eq. 0) THEN
         timet = 0
         tind = tind + 1
         WRITE(ext, '(i4.4)') tind
         IF (outs .ge. 1) THEN
            ! curlk (pseudospec_hd.f90:130-206): wx = dy vz - dz vy,
            ! wy = dz vx - dx vz, wz = dx vy - dy vx
            CALL store_curl('wx', vz, 2, vy, 3)
            CALL store_curl('wy', vx, 3, vz, 1)
            CALL store_curl('wz', vy, 1, vx, 2)
         ENDIF
         CALL store_field('vx', vx)
         CALL store_field('vy', vy)
         CALL store_field('vz', vz)
         CALL store_pressure('pr', pr)
         IF (bouss) CALL store_field('th', th)
         IF (mhd) THEN
            IF (outs .ge. 1) THEN
               CALL store_curl('bx', az, 2, ay, 3)
               CALL store_curl('by', ax, 3, az, 1)
               CALL store_curl('bz', ay, 1, ax, 2)
            ENDIF
            IF (outs .eq. 2) THEN
               CALL store_lap('jx', ax)
               CALL store_lap('jy', ay)
               CALL store_lap('jz', az)
            ENDIF
            CALL store_field('ax', ax)
            CALL store_field('ay', ay)
            CALL store_field('az', az)
            CALL store_pressure('ph', ph)
         ENDIF
      ENDIF
      IF (timec .eq. cstep .AND. bench .eq. 0) THEN
         timec = 0
         IF (mhd) THEN
            CALL mhd_check(t)
         ELSE
            CALL spx_check(spx_energy(plan, vx, vy, vz, 1, eng), 'energy')
            CALL spx_check(spx_energy(plan, vx, vy, vz, 0, ens), 'enstrophy')
            CALL spx_check(spx_helicity(plan, vx, vy, vz, khe), 'helicity')
            CALL spx_check(spx_cross(plan, vx, vy, vz, fx, fy, fz, pot), 'cross')
            WRITE(2, '(1P E13.6,2E23.16,E24.16)') (t-1)*dt, eng, ens, pot
            WRITE(9, '(1P E13.6,E24.16)') (t-1)*dt, khe
         ENDIF
         CALL spx_check(spx_divergence(plan, vx, vy, vz, dv), 'divergence')
         CALL spx_check(spx_bouncheck_z(plan, vx, vy, tb, tt), 'bouncheck')
         CALL spx_check(spx_bouncheck_z(plan, vz, C_NULL_PTR, nb, nt), 'bouncheck')
         WRITE(3, '(1P 6E24.16)') (t-1)*dt, dv, tb, tt, nb, nt
         IF (bouss) THEN
            ! pscheck (pseudospec_phd.f90:237-262): <th^2>, <|k^2 th|^2>, <th fs>
            CALL spx_check(spx_variance(plan, th, 1, sv), 'variance')
            CALL spx_check(spx_variance(plan, th, 0, sg), 'variance')
            CALL spx_check(spx_product(plan, th, fs, sp), 'product')
            WRITE(4, '(1P E13.6,2E22.14,E23.14)') (t-1)*dt, sv, sg, sp
            CALL spx_check(spx_bouncheck_z(plan, th, C_NULL_PTR, tb, tt), 'bouncheck th')
            WRITE(7, '(1P 3E24.16)') (t-1)*dt, tb, tt
            CALL spx_check(spx_nusselt(plan, th, nus), 'nusselt')
            WRITE(8, '(1P 3E24.16)') (t-1)*dt, nus(1), nus(2)
         ENDIF
      ENDIF
      CALL spx_check(spx_step(plan, state, 1), 'spx_step')
      timet = timet + 1
      timec = timec + 1
   ENDDO
   IF (bench .ge. 1) THEN
      CALL spx_check(spx_plan_sync(plan), 'sync')
      CALL CPU_TIME(cpu1)
      CALL SYSTEM_CLOCK(clk1)
      CALL write_benchmark()
   ENDIF
   CLOSE(2)
   CLOSE(3)
   CLOSE(9)
   IF (bouss) THEN
      CLOSE(4)
      CLOSE(7)
      CLOSE(8)
   ENDIF
   IF (mhd) THEN
      CLOSE(5)
      CLOSE(6)
      CALL spx_check(spx_dev_free(dtmp2), 'free')
      CALL spx_check(spx_dev_free(dtmp3), 'free')
   ENDIF

   CALL spx_check(spx_plan_sync(plan), 'sync')
   CALL spx_check(spx_dev_free(dreal), 'free')
   CALL spx_check(spx_dev_free(dtmp), 'free')
   CALL spx_check(spx_state_destroy(state), 'state destroy')
   CALL spx_check(spx_plan_destroy(plan), 'plan destroy')

CONTAINS

   ! benchmark.txt (specter.fpp:1182-1228, column layout of the FFTW build)
   SUBROUTINE write_benchmark()
      REAL(8) :: ns, wall
      ns = REAL(MAX(step - ini + 1, 1), 8)
      wall = REAL(clk1 - clk0, 8) / REAL(clkrate, 8)
      ! one extra step with HIP events around every kernel
      CALL spx_check(spx_prof_enable(plan, 1), 'prof on')
      CALL spx_check(spx_step(plan, state, 1), 'spx_step')
      CALL spx_check(spx_prof_resolve(plan, nk), 'prof resolve')
      tfft = 0.d0
      ttot = 0.d0
      DO ik = 0, nk - 1
         CALL spx_check(spx_prof_get(plan, ik, kname, INT(64, C_SIZE_T), nlaunch, tms, tby), &
                        'prof get')
         kstr = ' '
         DO i_ = 1, 64
            IF (kname(i_) .eq. C_NULL_CHAR) EXIT
            kstr(i_:i_) = kname(i_)
         ENDDO
         ttot = ttot + tms * 1.d-3
         IF (kstr(1:1) .eq. 'z' .OR. kstr(1:1) .eq. 'y' .OR. kstr(1:1) .eq. 'x' .OR. &
             kstr(1:6) .eq. 'noslip') tfft = tfft + tms * 1.d-3
      ENDDO
      CALL spx_check(spx_prof_enable(plan, 0), 'prof off')
      INQUIRE(file='benchmark.txt', exist=bbenchexist)
      OPEN(14, file='benchmark.txt', position='append')
      IF (.NOT. bbenchexist) THEN
         WRITE(14, *) '# nx ny nz nsteps nprocs nth TCPU TOMP TWTIME TFFT TTRA TCOM', &
                      'TCONT TNEU TROB TTOT'
      ENDIF
      WRITE(14, *) nx, ny, nz, (step - ini + 1), 1, 1, (cpu1 - cpu0) / ns, wall / ns, &
                   wall / ns, tfft, 0.d0, 0.d0, 0.d0, 0.d0, 0.d0, ttot
      IF (bench .eq. 2) THEN
         WRITE(14, *) 'FFTW: Create_plan = ', (cpup1 - cpup0) / ns, &
                      REAL(clkp1 - clkp0, 8) / REAL(clkrate, 8) / ns, &
                      REAL(clkp1 - clkp0, 8) / REAL(clkrate, 8) / ns
      ENDIF
      CLOSE(14)
   END SUBROUTINE write_benchmark

   FUNCTION itoa(i) RESULT(s)
      INTEGER, INTENT(IN) :: i
      CHARACTER(len=:), ALLOCATABLE :: s
      CHARACTER(len=12) :: buf
      WRITE(buf, '(I0)') i
      s = TRIM(buf)
   END FUNCTION itoa

   SUBROUTINE read_table(path, T, n)
      CHARACTER(len=*), INTENT(IN) :: path
      INTEGER, INTENT(IN) :: n
      REAL(8), INTENT(OUT) :: T(n)
      LOGICAL :: ex
      INQUIRE(FILE=path, EXIST=ex)
      IF (.NOT. ex) THEN
         PRINT *, 'Could not find table ', path
         STOP 1
      ENDIF
      OPEN(10, FILE=path, FORM='unformatted', ACCESS='stream', ACTION='read')
      READ(10) T
      CLOSE(10)
   END SUBROUTINE read_table

   SUBROUTINE read_raw(path, found)
      CHARACTER(len=*), INTENT(IN) :: path
      LOGICAL, INTENT(OUT) :: found
      INQUIRE(FILE=path, EXIST=found)
      IF (.NOT. found) RETURN
      R = 0.d0
      OPEN(11, FILE=path, FORM='unformatted', ACCESS='stream', ACTION='read')
      READ(11) R(:, :, 1:np)
      CLOSE(11)
   END SUBROUTINE read_raw

   SUBROUTINE load_field(name, required, dst, found)
      CHARACTER(len=*), INTENT(IN) :: name
      LOGICAL, INTENT(IN) :: required
      TYPE(C_PTR), INTENT(IN) :: dst
      LOGICAL, INTENT(OUT) :: found
      CALL read_raw(TRIM(idir)//'/'//name//'.'//ext//'.out', found)
      IF (.NOT. found) THEN
         IF (required) THEN
            PRINT *, '[ERROR] missing input field ', name
            STOP 1
         ENDIF
         RETURN
      ENDIF
      CALL spx_check(spx_memcpy_h2d(plan, dreal, C_LOC(R), rbytes), 'h2d '//name)
      CALL spx_check(spx_fftp3d_real_to_complex(plan, dreal, dst), 'r2c '//name)
   END SUBROUTINE load_field

   SUBROUTINE store_field(name, src)
      CHARACTER(len=*), INTENT(IN) :: name
      TYPE(C_PTR), INTENT(IN) :: src
      CALL spx_check(spx_memcpy_d2d(plan, dtmp, src, cbytes), 'd2d '//name)
      CALL spx_check(spx_fftp3d_complex_to_real(plan, dtmp, dreal), 'c2r '//name)
      CALL spx_check(spx_memcpy_d2h(plan, C_LOC(R), dreal, rbytes), 'd2h '//name)
      R = R / (REAL(nx, 8) * REAL(ny, 8) * REAL(nz, 8))
      CALL write_raw(name)
   END SUBROUTINE store_field

   ! R (physical planes) -> odir/<name>.ext.out, raw float64 (binary_io.f90:165-226)
   SUBROUTINE write_raw(name)
      CHARACTER(len=*), INTENT(IN) :: name
      OPEN(12, FILE=TRIM(odir)//'/'//name//'.'//ext//'.out', FORM='unformatted', &
           ACCESS='stream', STATUS='replace')
      WRITE(12) R(:, :, 1:np)
      CLOSE(12)
   END SUBROUTINE write_raw

   ! one vorticity component: d_dp b - d_dm a, as v/(nx ny nz) in physical space
   SUBROUTINE store_curl(name, b, dp, a, dm)
      CHARACTER(len=*), INTENT(IN) :: name
      TYPE(C_PTR), INTENT(IN) :: a, b
      INTEGER, INTENT(IN) :: dp, dm
      IF (.NOT. ALLOCATED(R2)) ALLOCATE(R2(nx, ny, nz))
      CALL spx_check(spx_derivk(plan, a, dtmp, dm), 'derivk '//name)
      CALL spx_check(spx_fftp3d_complex_to_real(plan, dtmp, dreal), 'c2r '//name)
      CALL spx_check(spx_memcpy_d2h(plan, C_LOC(R2), dreal, rbytes), 'd2h '//name)
      CALL spx_check(spx_derivk(plan, b, dtmp, dp), 'derivk '//name)
      CALL spx_check(spx_fftp3d_complex_to_real(plan, dtmp, dreal), 'c2r '//name)
      CALL spx_check(spx_memcpy_d2h(plan, C_LOC(R), dreal, rbytes), 'd2h '//name)
      R = (R - R2) / (REAL(nx, 8) * REAL(ny, 8) * REAL(nz, 8))
      CALL write_raw(name)
   END SUBROUTINE store_curl

   ! pressure p = p'/dt (or the electric potential phi = phi'/dt) from the
   ! mixed (z, ky, kx) domain (specter.fpp:1041-1053, 1114-1126)
   SUBROUTINE store_pressure(name, src)
      CHARACTER(len=*), INTENT(IN) :: name
      TYPE(C_PTR), INTENT(IN) :: src
      CALL spx_check(spx_memcpy_d2d(plan, dtmp, src, cbytes), 'd2d '//name)
      CALL spx_check(spx_fftp2d_complex_to_real_xy(plan, dtmp, dreal), 'c2r '//name)
      CALL spx_check(spx_memcpy_d2h(plan, C_LOC(R), dreal, rbytes), 'd2h '//name)
      R = R / (REAL(nx, 8) * REAL(ny, 8) * dt)
      CALL write_raw(name)
   END SUBROUTINE store_pressure

   ! laplak(a) as a/(nx ny nz) in physical space (specter.fpp:1094-1104)
   SUBROUTINE store_lap(name, src)
      CHARACTER(len=*), INTENT(IN) :: name
      TYPE(C_PTR), INTENT(IN) :: src
      CALL spx_check(spx_laplak(plan, src, dtmp), 'laplak '//name)
      CALL spx_check(spx_fftp3d_complex_to_real(plan, dtmp, dreal), 'c2r '//name)
      CALL spx_check(spx_memcpy_d2h(plan, C_LOC(R), dreal, rbytes), 'd2h '//name)
      R = R / (REAL(nx, 8) * REAL(ny, 8) * REAL(nz, 8))
      CALL write_raw(name)
   END SUBROUTINE store_lap

   ! mhdcheck with hel = crs = 1 (pseudospec_mhd.f90 mhdcheck)
   SUBROUTINE mhd_check(tt_)
      INTEGER, INTENT(IN) :: tt_
      CALL spx_check(spx_energy(plan, vx, vy, vz, 1, engk), 'energy')
      CALL spx_check(spx_energy(plan, vx, vy, vz, 0, ens), 'enstrophy')
      CALL spx_check(spx_energy(plan, ax, ay, az, 0, engm), 'magnetic energy')
      CALL spx_check(spx_energy(plan, ax, ay, az, 2, cur), 'current')
      CALL spx_check(spx_helicity(plan, vx, vy, vz, khe), 'helicity')
      CALL spx_check(spx_helicity(plan, ax, ay, az, helm), 'magnetic helicity')
      CALL spx_check(spx_energy(plan, ax, ay, az, 1, asq), 'asq')
      CALL spx_check(spx_derivk(plan, ax, dtmp, 1), 'derivk ax')
      CALL spx_check(spx_derivk(plan, ay, dtmp2, 2), 'derivk ay')
      CALL spx_check(spx_derivk(plan, az, dtmp3, 3), 'derivk az')
      CALL spx_check(spx_cross(plan, vx, vy, vz, dtmp, dtmp2, dtmp3, crh), 'cross')
      WRITE(2, '(1P E13.6,3E23.16)') (tt_-1)*dt, engk + engm, ens, cur
      WRITE(5, '(1P E13.6,2E23.16)') (tt_-1)*dt, engk, engm
      WRITE(9, '(1P E13.6,2E24.16)') (tt_-1)*dt, khe, helm
      WRITE(6, '(1P E13.6,E23.16,E24.16)') (tt_-1)*dt, crh, asq
   END SUBROUTINE mhd_check

   ! b_parsebc (bboundary.f90:20-41) for z walls: conducting 0, vacuum 1
   INTEGER FUNCTION bkind(str)
      CHARACTER(len=*), INTENT(IN) :: str
      IF (TRIM(str) .eq. 'conducting') THEN
         bkind = SPX_BWALL_CONDUCTING
      ELSE IF (TRIM(str) .eq. 'vacuum') THEN
         bkind = SPX_BWALL_INSULATING
      ELSE
         PRINT *, '[ERROR] Unknown magnetic boundary condition in z: ', TRIM(str)
         STOP 1
      ENDIF
   END FUNCTION bkind

   ! x.txt, y.txt, z.txt (specter.fpp:683-765: periodic x, y; z on the
   ! physical points of the FC-Gram grid)
   SUBROUTINE write_grid()
      INTEGER :: i
      REAL(8), PARAMETER :: pi = 3.14159265358979323846d0
      OPEN(13, file='x.txt', action='write', status='replace')
      DO i = 1, nx
         WRITE(13, FMT='(1P E23.15)') (Lx * 2.d0 * pi / nx) * (i - 1)
      ENDDO
      CLOSE(13)
      OPEN(13, file='y.txt', action='write', status='replace')
      DO i = 1, ny
         WRITE(13, FMT='(1P E23.15)') (Ly * 2.d0 * pi / ny) * (i - 1)
      ENDDO
      CLOSE(13)
      OPEN(13, file='z.txt', action='write', status='replace')
      DO i = 1, np
         WRITE(13, FMT='(1P E23.15)') (Lz / (np - 1)) * (i - 1)
      ENDDO
      CLOSE(13)
   END SUBROUTINE write_grid

END PROGRAM spx_hd
