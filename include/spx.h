/*
 * libspx -- MI355X-native implementation of SPECTER's per-RK-substep hot path.
 *
 * C ABI only: plain pointers and sizes.  Field arguments are DEVICE pointers
 * (HBM) in the reference's Fortran layouts:
 *   spectral  (nz, ny, nx/2+1) complex, z fastest      ["S"]
 *   real      (nx, ny, nz)     real,    x fastest      ["R"]
 * Fields are float64 (complex = 2 x double) in every plan.  The plan's
 * precision selects the arithmetic of the nonlinear term:
 *   SPX_FP64        everything in float64
 *   SPX_FP32_MIXED  gradre / advect (their z, y, x transforms and products)
 *                   run in float32 and write complex float (2 x float)
 *                   results; the FC continuation sums, the Poisson and
 *                   harmonic pressure solve, the wall conditions, the RK
 *                   update and every public transform stay float64.
 * All transforms are unnormalised, as in the reference (callers scale).
 * Every call returns 0 on success or an SPX_ERR_* code; spx_last_error()
 * returns the message.  A plan is bound to one HIP stream; calls on one plan
 * must come from one host thread at a time, different plans may run
 * concurrently.  Work is enqueued on the plan's stream and is asynchronous
 * unless stated otherwise.
 *
 * Each entry point names the reference routine it replaces (paths relative to
 * the SPECTER source tree, src/).
 */
#ifndef SPX_H
#define SPX_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  SPX_OK = 0,
  SPX_ERR_ARG = 1,          /* invalid argument (null pointer, bad size, bad enum)   */
  SPX_ERR_UNSUPPORTED = 2,  /* valid but not supported (grid too large for LDS, ...) */
  SPX_ERR_HIP = 3,          /* HIP runtime error                                     */
  SPX_ERR_ALLOC = 4,        /* device allocation failed                              */
  SPX_ERR_STATE = 5         /* call out of order / missing field                     */
};

enum { SPX_FP64 = 0, SPX_FP32_MIXED = 1 };
enum { SPX_SOLVER_HD = 0, SPX_SOLVER_BOUSS = 1, SPX_SOLVER_ROTBOUSS = 2, SPX_SOLVER_MHD = 3 };
enum { SPX_BWALL_CONDUCTING = 0, SPX_BWALL_INSULATING = 1 };  /* b_parsebc codes (bboundary.f90:27-31) */

typedef struct spx_plan spx_plan;
typedef struct spx_state spx_state;

/* Grid description; mirrors grid/kes module globals (pseudospec_mod.fpp:13-91)
 * and the FCPLAN constructor arguments (fftp/fcgram_mod.f90:49-144). */
typedef struct {
  int nx, ny, nz;   /* points per direction; nz INCLUDES the Cz continuation points */
  int Cz;           /* continuation points in z (0 = periodic z, no tables)          */
  int oz;           /* FC-Gram matching order d in z                                 */
  double Lx, Ly;    /* periodic box lengths in units of 2*pi (specter.fpp:683-711)   */
  double Lz;        /* wall-to-wall length of the physical z domain (:743-749)       */
  int precision;    /* SPX_FP64 or SPX_FP32_MIXED                                    */
} spx_grid;

/* Sizes of a plan (derived values). */
typedef struct {
  int nkx;                 /* nx/2+1                                   */
  int np;                  /* nz - Cz physical z points                */
  size_t n_spectral;       /* nz*ny*nkx complex elements of an S array */
  size_t n_real;           /* nx*ny*nz real elements of an R array     */
  size_t bytes_spectral;
  size_t bytes_real;
  int z_fft_len;           /* length of the z FFT run in LDS (Bluestein size for awkward nz) */
  int z_algorithm;         /* z DFT algorithm: 0 direct, 1 Bluestein, 2 Rader               */
  size_t device_bytes;     /* device memory the plan holds now (tables, basis, scratch)      */
  size_t work_bytes;       /* more device memory the plan allocates on first use of the
                              nonlinear term (gradre work arrays; 0 once allocated)          */
  size_t work_bytes_bouss; /* and on first use of advect (Boussinesq), beyond work_bytes     */
  size_t state_bytes;      /* device memory of one spx_state (all members): HD              */
  size_t state_bytes_bouss;/* ... Boussinesq                                                 */
  int z_threads_per_line;  /* launch shape of the float64 z line kernels                      */
  int z_lines_per_block;
  int z_tables_global;     /* 1: Rader / Bluestein tables read from global memory, 0: LDS    */
  int z_wave_plan;         /* compiled one-wavefront-per-line plan of the z transforms, -1: none */
  int z_wave_plan_lo;      /* ... of the float32 z transforms of a mixed plan, -1: none      */
} spx_plan_info;

/* Replaces fcgram_create_plan (fftp/fcgram_mod.f90:49-144) including
 * load_dirichlet_tables (:180-257): A is the raw A{C}-{d}.dat table (C x d,
 * Fortran order), Q the raw Q{d}.dat table (d x d, Fortran order); the plan
 * forms dir = A*Q^T in float64.  A/Q may be NULL when Cz == 0.  stream is a
 * hipStream_t (NULL = default stream).  Plans with walls also hold the
 * harmonic-correction basis of sol_project, 2 * (distinct sqrt(kx^2+ky^2))
 * * nz complex float64 values (53 MB at 256^3, 2.8 GB at 1024^3); with the
 * environment variable SPX_HARMONIC_TABLE=0 at creation, or when it cannot be
 * allocated, the correction transforms its profiles per line instead. */
int spx_plan_create(const spx_grid* g, const double* A, const double* Q, void* stream,
                    spx_plan** out);
/* Replaces fcgram_destroy_plan (fftp/fcgram_mod.f90:147-177). */
int spx_plan_destroy(spx_plan* p);
int spx_plan_get_info(const spx_plan* p, spx_plan_info* info);
/* Launch-shape overrides (no counterpart in the reference, whose FFTW plans
 * choose their own; set by the autotuner specter_amd/autotune.py), applied
 * to plans created afterwards in this process.  spx_tune_line: for line
 * FFTs of length n with csize-byte complex elements (16 float64, 8 float32),
 * lines per workgroup and threads per line (powers of two; 0 = default) and
 * the Rader / Bluestein tables in global memory (1), LDS (0) or default (-1);
 * values the engine cannot run (LDS, 512 threads, fewer threads than the
 * length needs) are ignored.  spx_tune_wall: kappa blocks per wave (9, 18)
 * and waves per SIMD (3, 4) of the wall kernel for lines of length n. */
int spx_tune_line(int n, int csize, int lines, int threads, int tables_global);
int spx_tune_wall(int n, int blocks_per_wave, int waves_per_simd);
int spx_tune_clear(void);
/* Block the host until the plan's stream is idle. */
int spx_plan_sync(spx_plan* p);

/* ---- batched plans (no counterpart in the reference; for ensembles of
 * independent runs on one grid).  A batched plan holds nmembers runs of the
 * grid g; every field argument then holds the members one after the other
 * (member b of a spectral array at element offset b*n_spectral, of a real
 * array at b*n_real), and every call processes the nactive leading members
 * (default: all).  Diagnostics write one value per active member to their
 * output arrays.  spx_plan_create is spx_plan_create_batch with nmembers = 1.
 * Results of a member do not depend on the other members. */
int spx_plan_create_batch(const spx_grid* g, const double* A, const double* Q, int nmembers,
                          void* stream, spx_plan** out);
int spx_plan_get_members(const spx_plan* p, int* nmembers, int* nactive);

/* ---- non-periodic y.  fcgram_create_plan with Cy > 0 (fcgram_mod.f90:49-144,
 * the y tables :115-126; grid specter.fpp:706-723): as spx_plan_create_batch,
 * and the forward 3D transforms continue every y line of the physical z
 * planes by FC-Gram (fftp.fpp:565-581) before the y FFT.  g->ny INCLUDES the Cy
 * continuation points (ny must stay even: the x kernels pack row pairs), g->Ly
 * is the physical y length, dy = Ly/(ny-Cy-1), Dky = 2 pi/(dy ny), and ky runs
 * in standard DFT order.  Ay is the raw A{Cy}-{oy}.dat table (Cy x oy, Fortran
 * order), Qy the raw Q{oy}.dat table.  Cy = 0 (oy = 0, tables NULL) is
 * spx_plan_create_batch.  Plans with Cy > 0 are SPX_FP64 only, and the
 * boundary-condition calls and the solvers (spx_sol_project,
 * spx_v_imposebc_and_project, spx_s_imposebc, spx_state_create) return
 * SPX_ERR_UNSUPPORTED, as the reference stops on y boundary conditions
 * (boundary_mod.fpp:104-105, :179-180). */
int spx_plan_create_fcy(const spx_grid* g, const double* A, const double* Q, int Cy, int oy,
                        const double* Ay, const double* Qy, int nmembers, void* stream,
                        spx_plan** out);
int spx_plan_set_active(spx_plan* p, int nactive);

/* ---- transforms (fftp/fftp.fpp) ------------------------------------------ */
/* fftp3d_real_to_complex (fftp.fpp:388-425): R -> S. Continuation planes of the
 * real input (k > nz-Cz) are ignored. */
int spx_fftp3d_real_to_complex(spx_plan* p, const void* in_real, void* out_spec);
/* fftp3d_complex_to_real (fftp.fpp:789-821): S -> R, all nz planes.  Like the
 * reference, the input may be overwritten (here it is preserved). */
int spx_fftp3d_complex_to_real(spx_plan* p, void* in_spec, void* out_real);
/* fftp2d_real_to_complex_xy (fftp.fpp:428-524): per-plane 2D r2c + transpose. */
int spx_fftp2d_real_to_complex_xy(spx_plan* p, const void* in_real, void* out_spec);
/* fftp2d_complex_to_real_xy (fftp.fpp:824-919): transpose + per-plane 2D c2r. */
int spx_fftp2d_complex_to_real_xy(spx_plan* p, void* in_spec, void* out_real);
/* fftp1d_real_to_complex_z (fftp.fpp:720-786): FC continuation + z FFT, in place. */
int spx_fftp1d_real_to_complex_z(spx_plan* p, void* inout_spec);
/* fftp1d_complex_to_real_z (fftp.fpp:1060-1094): z IFFT, in place. */
int spx_fftp1d_complex_to_real_z(spx_plan* p, void* inout_spec);
/* The continuation loop alone (fftp.fpp:755-773), in place. */
int spx_fc_continue(spx_plan* p, void* inout_spec);
/* The pieces of the non-periodic-y forward 3D transform (fftp.fpp:416-418) and
 * their backward counterparts, on the S layout (the reference's backward 3D
 * transform runs z IFFT + 2D xy c2r for every Cy, :817-818, which equals the
 * yz IFFT + x c2r up to rounding); with Cy = 0 the yz transforms have no y
 * continuation.
 * fftp1d_real_to_complex_x (fftp.fpp:620-716): x r2c of all nz planes, S layout. */
int spx_fftp1d_real_to_complex_x(spx_plan* p, const void* in_real, void* out_spec);
/* fftp2d_real_to_complex_yz (fftp.fpp:527-616): y continuation of the physical
 * z planes, z continuation, 2D (y, z) FFT, in place. */
int spx_fftp2d_real_to_complex_yz(spx_plan* p, void* inout_spec);
/* fftp2d_complex_to_real_yz (fftp.fpp:922-956): 2D (y, z) IFFT, in place. */
int spx_fftp2d_complex_to_real_yz(spx_plan* p, void* inout_spec);
/* fftp1d_complex_to_real_x (fftp.fpp:961-1056): x c2r of all nz planes. */
int spx_fftp1d_complex_to_real_x(spx_plan* p, const void* in_spec, void* out_real);

/* ---- spectral operators (pseudo/pseudospec_hd.f90, pseudospec_phd.f90) ---- */
int spx_derivk(spx_plan* p, const void* a, void* b, int dir);      /* :28-94 */
int spx_laplak(spx_plan* p, const void* a, void* b);               /* :97-127 */
int spx_fc_filter(spx_plan* p, void* a);                           /* :1082-1115 */
/* gradre (:209-319): (A.grad)A in advective form; d,e,f = x,y,z components
 * (complex float in SPX_FP32_MIXED plans, complex double otherwise). */
int spx_gradre(spx_plan* p, const void* a, const void* b, const void* c, void* d, void* e,
               void* f);
/* No counterpart in the reference: gradre for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops).  Same
 * arguments and results as spx_gradre, but the work arrays come from a static
 * 128 MiB device array that every GPU the library runs on holds until the
 * process ends, instead of memory the plan allocates on first use.  Grids
 * whose work set exceeds it (spx_gradre_scratch_bytes: need > limit; e.g.
 * float64 256^3) are refused with SPX_ERR_UNSUPPORTED.  Calls that overlap on
 * one GPU share that array and race.  The plan's device must be current. */
int spx_gradre_noalloc(spx_plan* p, const void* a, const void* b, const void* c, void* d,
                       void* e, void* f);
int spx_gradre_scratch_bytes(const spx_plan* p, size_t* need, size_t* limit);
/* advect (pseudospec_phd.f90:23-113): (A.grad)B for a scalar B = d; result e
 * (complex float in SPX_FP32_MIXED plans). */
int spx_advect(spx_plan* p, const void* a, const void* b, const void* c, const void* d,
               void* e);

/* ---- boundary conditions and projection (boundary/) ---------------------- */
/* sol_project (boundary_mod.fpp:197-402): bctarget = 1 (condition on the
 * normal component, the velocity BCs) with homogeneous Dirichlet walls
 * bczsta = bczend = 0; bctarget = 0 (condition on the potential, the magnetic
 * BCs) with Dirichlet (0) or Robin (2) walls, except Robin bottom with
 * Dirichlet top (laplace_z stops on it). */
int spx_sol_project(spx_plan* p, void* a, void* b, void* c, void* d, int bctarget, int bczsta,
                    int bczend);
/* v_imposebc_and_project (vboundary.f90:67-151) for no-slip walls at z=0,Lz:
 * rki is the RK stage o, ord the RK order (pressure extrapolation factor of
 * noslip_z, :195-196); v_zsta/v_zend are the wall tangential velocities (x,y),
 * may be NULL for zero. */
int spx_v_imposebc_and_project(spx_plan* p, void* vx, void* vy, void* vz, void* pr, int rki,
                               int ord, const double* v_zsta, const double* v_zend);
/* s_imposebc (sboundary.f90:67-120): scalar = 0 at both walls. */
int spx_s_imposebc(spx_plan* p, void* th);

/* ---- MHD (pseudo/pseudospec_hd.f90, pseudo/pseudospec_mhd.f90, boundary/bboundary.f90) */
/* load_neumann_tables (fcgram_mod.f90:260-356) for the z walls: Q is the raw
 * Q{d}.dat table (d x d, Fortran order), Qn the raw Q{ord}n{d}.dat stream (a
 * float64 dxp, then d x d values); ord 1 (normal derivative) or 2 (second
 * derivative). */
int spx_plan_load_neumann(spx_plan* p, int ord, const double* Q, const double* Qn);
/* curlk (pseudospec_hd.f90:130-206): c = component dir of curl from a, b. */
int spx_curlk(spx_plan* p, const void* a, const void* b, void* c, int dir);
/* prodre (pseudospec_hd.f90:322-402): (d,e,f) = (curl A) x A. */
int spx_prodre(spx_plan* p, const void* a, const void* b, const void* c, void* d, void* e,
               void* f);
/* vector (pseudospec_mhd.f90): (g,h,i) = (a,b,c) x (d,e,f). */
int spx_vector(spx_plan* p, const void* a, const void* b, const void* c, const void* d,
               const void* e, const void* f, void* g, void* h, void* i);
/* a_imposebc_and_project (bboundary.f90:100-189): wall kinds 0 conducting, 1
 * insulating (vacuum); insulating bottom with conducting top is refused, as
 * laplace_z stops on it.  ph receives the projection potential (mixed domain). */
int spx_a_imposebc_and_project(spx_plan* p, void* ax, void* ay, void* az, void* ph, int bczsta,
                               int bczend);

/* ---- global diagnostics (host-side results, synchronous; batched plans write
 * one value per active member to each output) ------------------------------ */
int spx_energy(spx_plan* p, const void* a, const void* b, const void* c, int kin,
               double* out);                                   /* pseudospec_hd.f90:405-635 */
int spx_cross(spx_plan* p, const void* a, const void* b, const void* c, const void* d,
              const void* e, const void* f, double* out);      /* :778-940, kin = 1 */
int spx_helicity(spx_plan* p, const void* a, const void* b, const void* c,
                 double* out);                                 /* :638-775 */
int spx_divergence(spx_plan* p, const void* a, const void* b, const void* c,
                   double* out);                               /* :1118-1235 */
int spx_bouncheck_z(spx_plan* p, const void* a, const void* b, double* bot,
                    double* top);                              /* boundary_mod.fpp:681-801 */

/* variance (pseudospec_phd.f90:116-176): <|a|^2> (kin=1) or <|kk2 a|^2>
 * (kin=0) over physical z; product (:178-235): <Re(a conj b)>.  pscheck
 * (:237-262) is variance(th,1), variance(th,0), product(th,fs). */
int spx_variance(spx_plan* p, const void* a, int kin, double* out);
int spx_product(spx_plan* p, const void* a, const void* b, double* out);
/* Wall and solution diagnostics the reference does not compute (defined by
 * this library): mean wall shear nu <d(a,b)/dz> as (tau_x, tau_y) at z = 0
 * then z = Lz (4 values per member); Nusselt number 1 - Lz <d th/dz> at
 * z = 0 and z = Lz for theta_total = 1 - z/Lz + th (2 values per member);
 * max |div v| over the physical points (physical units); the horizontal
 * shell spectrum of <|v|^2> over physical z, shells s - 1/2 <= |(kx,ky)|/Dkx
 * < s + 1/2, nbins values per member. */
int spx_wall_shear(spx_plan* p, const void* a, const void* b, double nu, double* out);
int spx_nusselt(spx_plan* p, const void* th, double* out);
int spx_divergence_max(spx_plan* p, const void* a, const void* b, const void* c, double* out);
int spx_energy_spectrum(spx_plan* p, const void* a, const void* b, const void* c, int nbins,
                        double* out);

/* ---- device-resident solver state (include/hd, include/bouss, specter.fpp RK loop) ---- */
typedef struct {
  int solver;        /* SPX_SOLVER_HD, SPX_SOLVER_BOUSS or SPX_SOLVER_ROTBOUSS */
  int ord;           /* RK order ORD (specter.fpp:1157)                        */
  double dt, nu;
  double kappa;      /* scalar diffusivity (Boussinesq)                        */
  double xmom, xtemp;/* gama*xmom, gama*xtemp (specter.fpp:480-491)            */
  double v_zsta[2];  /* wall velocities (vxzsta, vyzsta)                        */
  double v_zend[2];  /* (vxzend, vyzend)                                        */
  double omega[3];   /* ROTBOUSS: (omegax, omegay, omegaz), the &rotation
                        namelist (specter.fpp:662-677); ignored otherwise      */
  double mu;         /* MHD: magnetic diffusivity (&magfield)                  */
  double b0[3];      /* MHD: uniform field (bx0, by0, bz0), the &uniformb namelist */
  int bzsta, bzend;  /* MHD: magnetic wall kinds at z = 0 and z = Lz
                        (SPX_BWALL_CONDUCTING / SPX_BWALL_INSULATING; &magbound) */
} spx_params;

int spx_state_create(spx_plan* p, const spx_params* prm, spx_state** out);
int spx_state_destroy(spx_state* s);
/* Device pointer of a state field: "vx","vy","vz","pr","fx","fy","fz","th","fs";
 * MHD: "ax","ay","az" (vector potential), "ph" (its projection potential),
 * "mx","my","mz" (electromotive forcing).
 * Fields are S arrays; the caller fills them (e.g. via spx_fftp3d_real_to_complex). */
int spx_state_field(spx_state* s, const char* name, void** dev_ptr);
/* hd_rkstep1 copy (include/hd/hd_rkstep1.f90:4-6). */
int spx_rkstep1(spx_plan* p, spx_state* s);
/* one RK stage o (include/hd/hd_rkstep2.f90:3-36 / include/bouss/bouss_rkstep2.f90:3-59 /
 * include/rotbouss/rotbouss_rkstep2.f90:3-50 / include/mhd/mhd_rkstep2.f90:3-84).
 * MHD states need SPX_FP64 plans with both Neumann tables loaded
 * (spx_plan_load_neumann). */
int spx_substep(spx_plan* p, spx_state* s, int o);
/* nsteps full steps: rkstep1 + o = ord..1 (specter.fpp:1142-1161). */
int spx_step(spx_plan* p, spx_state* s, int nsteps);
/* Whether spx_step replays one captured HIP graph per RK step (default 1)
 * instead of launching every kernel; no counterpart in the reference. */
int spx_state_set_graph(spx_state* s, int on);
/* Diffusivities of one member (nu for the velocity, kappa for th); members
 * start with the values of spx_state_create's params. */
int spx_state_set_member(spx_state* s, int member, double nu, double kappa);
/* Copy member src's fields (v, pr, forcing, th, fs) and diffusivities onto
 * member dst, e.g. to keep the active members leading when a run ends. */
int spx_state_copy_member(spx_state* s, int dst, int src);

/* ---- per-kernel timing (no counterpart in the reference; the reference's
 * gtimer accumulators ffttime/tratime/conttime, fftp_mod.fpp:32-37, play this
 * role on the CPU).  When enabled, every kernel launch of the plan is bracketed
 * by HIP events on the plan's stream; spx_prof_resolve synchronises and sums
 * elapsed time and algorithmic bytes (compulsory HBM reads + writes) per kernel. */
int spx_prof_enable(spx_plan* p, int on);   /* on = 1 clears previous records */
int spx_prof_resolve(spx_plan* p, int* nkernels);
int spx_prof_get(spx_plan* p, int idx, char* name, size_t len, long* launches,
                 double* total_ms, double* total_bytes);

/* ---- memory helpers for callers without a HIP binding (e.g. Fortran) ------ */
int spx_dev_alloc(size_t bytes, void** ptr);
int spx_dev_free(void* ptr);
int spx_memcpy_h2d(spx_plan* p, void* dst, const void* src, size_t bytes);  /* synchronous */
int spx_memcpy_d2h(spx_plan* p, void* dst, const void* src, size_t bytes);  /* synchronous */
int spx_memcpy_d2d(spx_plan* p, void* dst, const void* src, size_t bytes);  /* on the plan's stream */

/* Last error message of the calling thread; returns its length. */
int spx_last_error(char* buf, size_t len);

#ifdef __cplusplus
}
#endif
#endif /* SPX_H */
