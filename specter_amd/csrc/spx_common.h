// Shared definitions of the SPECTER MI355X hot-path library (libspx).
//
// Layouts are the reference's Fortran layouts, held in HBM unchanged:
//   spectral S : (nz, ny, nkx) complex, z fastest   -- fftp.fpp:388-425 output
//   real     R : (nx, ny, nz)  real,    x fastest
//   xy-stage X : (nkx, ny, nz) complex, kx fastest  -- the reference's c1 buffer
//                (fftp.fpp:446) before the (kx,y,z)->(z,y,kx) transpose
#pragma once

#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/spx.h"

namespace spx {

template <typename T> struct cvec;
template <> struct cvec<double> { using type = double2; };
template <> struct cvec<float> { using type = float2; };

// Device-side description of one batched line DFT of length n, computed
// directly (n smooth), by Rader's algorithm (n prime, n-1 smooth) or by
// Bluestein's chirp-z (otherwise) with an M-point smooth FFT in LDS.
constexpr int kMaxStages = 16;
template <typename C>
struct LineFFT {
  int n;      // logical DFT length
  int M;      // length of the in-LDS FFT actually run
  int mode;   // 0 direct, 1 Bluestein, 2 Rader
  int nstage;
  int radix[kMaxStages];
  unsigned magic[kMaxStages];  // ceil(2^32 / Ns) of each stage (Ns = product of earlier radices)
  int tpl, lgtpl;  // threads per line (power of two) and its log2
  int nl, lgnl;    // lines per workgroup (power of two) and its log2
  int ldl;         // LDS stride between lines (complex elements)
  int tab;         // complex slots of the tables staged at the head of LDS
  const C* tw;     // W_M[m] = exp(-2 pi i m / M), m < M
  const C* chirp;  // Bluestein: w[m] = exp(-pi i m^2 / n), m < n
  const C* bfil;   // Bluestein: FFT_M(wrapped conj chirp)/M ; Rader: FFT_{n-1}(b)/(n-1)
  const int* perm; // Rader: g^q mod n (n-1 entries) then g^-p mod n (n-1 entries)
};

// Host-side owner of the tables behind a LineFFT.
struct LineFFTHost {
  int n = 0, M = 0, mode = 0, nstage = 0, tpl = 0, lgtpl = 0, nl = 0, lgnl = 0, ldl = 0;
  int wave = -1;  // compiled one-wavefront-per-line plan of the line (spx_wave.h), -1: none
  int tab = 0;  // LDS table slots: twiddles M, plus bfil M and chirp n
                // (Bluestein) or bfil M and the 2M Rader indices (Rader)
  int radix[kMaxStages] = {0};
  unsigned magic[kMaxStages] = {0};
  void* tw = nullptr;
  void* chirp = nullptr;
  void* bfil = nullptr;
  void* perm = nullptr;
  size_t bytes = 0;  // device bytes of the tables above
  // LDS bytes of a launch: tables + nbuf line buffers
  size_t lds_bytes(size_t csize, int nbuf = 1) const {
    return (size_t(tab) + size_t(nbuf) * size_t(nl) * size_t(ldl)) * csize;
  }
  LineFFTHost with_lines(int lines) const {
    LineFFTHost h = *this;
    h.nl = lines < 1 ? 1 : lines;
    h.lgnl = 0;
    while ((1 << h.lgnl) < h.nl) ++h.lgnl;
    return h;
  }
  template <typename C> LineFFT<C> dev() const {
    LineFFT<C> f;
    f.n = n; f.M = M; f.mode = mode; f.nstage = nstage;
    f.tpl = tpl; f.lgtpl = lgtpl; f.nl = nl; f.lgnl = lgnl; f.ldl = ldl; f.tab = tab;
    for (int i = 0; i < kMaxStages; ++i) {
      f.radix[i] = radix[i];
      f.magic[i] = magic[i];
    }
    f.tw = static_cast<const C*>(tw);
    f.chirp = static_cast<const C*>(chirp);
    f.bfil = static_cast<const C*>(bfil);
    f.perm = static_cast<const int*>(perm);
    return f;
  }
};

// Per-launch HIP-event timing of the plan's kernels (off unless enabled).
struct ProfRec {
  int kind;
  hipEvent_t a, b;
  double bytes;
};
struct Profiler {
  bool on = false;
  const char* suffix = "";  // appended to the names of launches (e.g. "_bnd")
  std::vector<std::string> names;
  std::vector<ProfRec> recs;
  std::vector<hipEvent_t> pool;
  size_t used = 0;
  // resolved totals per kernel name
  std::vector<long> launches;
  std::vector<double> total_ms, total_bytes;
};

// the wall update as a rank-(2+C) correction on the matrix cores (spx_wall.hip)
struct WallMfmaHost {
  bool on = false;
  int nbk = 0;   // 16-row blocks of the 2 nz real line values
  int kbw = 0;   // blocks per wave
  int nw = 0;    // waves per workgroup
  int mt = 0;    // 16-row tiles of the partial inverse DFT (2 (2d + C) rows)
  int kt = 0;    // 4-column steps of the correction (2 (2 + C) columns)
  int wpe = 0;   // waves per SIMD the kernel's register budget is sized for
  int tab = 1;   // 1: A operands from the roots of unity in LDS, 0: fragment tables
  unsigned magic = 0;  // floor(2^32 / nz) + 1 (x mod nz by multiplication)
  double* winv = nullptr;
  double* wfwd = nullptr;
  double* roots = nullptr;  // exp(2 pi i m / nz), m < nz
};

}  // namespace spx

// Opaque plan (FCPLAN + fftp plans + module globals of kes/grid, fcgram_mod.f90:34-38,
// pseudospec_mod.fpp:13-91).
struct spx_plan {
  spx_grid g;
  int nkx = 0, np = 0;
  size_t nc = 0, nr = 0;  // complex / real element counts of one member's S and R arrays
  hipStream_t stream = nullptr;
  // Batched plans hold nb independent runs ("members") of the same grid.
  // Every field argument then holds the members one after the other (member b
  // at offset b*nc or b*nr).  Calls process the nact leading members; mem is
  // the member count of the current launch view (see MemberView).  Spectral
  // and z-line kernels treat the members as extra (kx, ky) lines; the
  // xy-stage and real buffers of a multi-plane pass hold, per member, only
  // the planes the pass works on, members one after the other, so the y and
  // x passes see mem times as many planes.
  int nb = 1, nact = 1, mem = 1;
  int device = 0;           // HIP device current at creation
  bool xfused = true;       // gradre's x stage in one launch (SPX_XFUSED=0: six)
  bool fused_project = true;  // sol_project in one pass per line (SPX_FUSED_PROJECT=0: off)
  int wkx() const { return nkx * mem; }
  long wlines() const { return long(g.ny) * nkx * mem; }
  size_t wnc() const { return nc * size_t(mem); }
  size_t wnr() const { return nr * size_t(mem); }
  // The state, the public transforms, the boundary / projection / diagnostics
  // kernels and the RK update are float64 in every plan.  A mixed-precision
  // plan (SPX_FP32_MIXED) evaluates the nonlinear term (gradre / advect:
  // z, y, x transforms and products) in float32; csize / rsize are the
  // element sizes of that pipeline's buffers and of the nonlinear terms.
  int csize = 16;         // bytes per complex element of the nonlinear pipeline
  int rsize = 8;          // bytes per real element of the nonlinear pipeline
  bool fp32 = false;      // SPX_FP32_MIXED
  // FC-Gram continuation matrix dir = A*Q^T, row-major [C][d], float64 always.
  double* d_dir = nullptr;
  std::vector<double> h_dir;
  // Non-periodic y (Cy > 0): the forward 3D transforms continue every y line
  // of the physical z planes from its first and last oy physical points
  // before the y FFT; diry = Ay*Qy^T as dir.  Cy = 0: periodic y.
  int Cy = 0, oy = 0, npy = 0;
  double* d_diry = nullptr;
  std::vector<double> h_diry;
  // wavenumbers (float64), filter factors, z grid and khom(kx,ky); the kx,
  // fkx and khom tables are repeated once per member
  double *d_kx = nullptr, *d_ky = nullptr, *d_kz = nullptr;
  double *d_fkx = nullptr, *d_fky = nullptr, *d_fkz = nullptr;
  double *d_z = nullptr, *d_khom = nullptr;
  std::vector<double> h_kx, h_ky, h_kz, h_z, h_khom, h_fkx, h_fky, h_fkz;
  double Dkx = 0, Dky = 0, Dkz = 0, dz = 0;
  // float64 line FFTs (x, y, z; fz2: z with half the lines per workgroup for
  // the two-buffer kernels) and, in mixed plans, the float32 line FFTs of the
  // nonlinear pipeline.
  spx::LineFFTHost fx, fy, fz, fz2;
  spx::LineFFTHost fx32, fy32, fz32;
  // one-wavefront-per-line variants of fz / fz32 for the z transforms, when
  // the z length has a compiled wave plan (wave < 0: unused)
  spx::LineFFTHost fzw, fzw32;
  spx::WallMfmaHost wall;
  bool work_bouss = false;
  // scratch
  void* X = nullptr;        // xy-stage buffer, nc float64 complex
  // gradre / advect work arrays (allocated on first use, csize / rsize elements):
  //   XV[c]: z+y inverse transform of v_c, XZ[c]: of d_z v_c, XY[c]: y-derivative stage
  //   RV[c]: v_c in real space, RO[0..3]: nonlinear products (x,y,z, scalar)
  void* XV[4] = {nullptr};
  void* XZ[4] = {nullptr};
  void* XY[4] = {nullptr};
  void* RV[3] = {nullptr};
  void* RO[4] = {nullptr};
  void* R0 = nullptr;       // general real scratch (Boussinesq round trip)
  // MHD: real work fields (physical planes of every member), allocated on
  // first use by prodre / vector / the MHD stage; Neumann vectors neu
  // (order 1) and neu2 (order 2) of the z walls (fcgram_mod.f90:347-355)
  void* MR[12] = {nullptr};
  void* MS[6] = {nullptr};   // spectral scratch of energy(kin = 2), allocated on first use
  double neu[2][16] = {{0}};
  bool neu_loaded[2] = {false, false};
  // mixed plans with walls: float64 boundary-plane buffers of the nonlinear
  // pipeline (2d planes per component and member; see zfft_s2x_bnd)
  void* XVb = nullptr;      // z IFFT of v_c, then its y IFFT (4 components)
  void* XZb = nullptr;      // of d_z v_c
  void* XYb = nullptr;      // of d_y v_c
  void* RVb = nullptr;      // v_c in real space (3 components)
  void* ROb = nullptr;      // products (4 components)
  void* XFb = nullptr;      // forward x, y transforms of the products (4 components)
  void* d_wbnd = nullptr;   // exp(+2 pi i k / nz), then exp(+2 pi i k (np-d) / nz)
  void* d_wtop = nullptr;   // exp(+2 pi i k (np-1) / nz): the top wall of sol_project
  void* DS = nullptr;       // diagnostics scratch: spectral (nb members), allocated on use
  void* DR = nullptr;       // diagnostics scratch: real (nb members)
  void* bc = nullptr;       // wall data of sol_project, 2 complex per (kx,ky) line
  void* pw = nullptr;       // IFFT_z(d)/nz at the two walls, 2 complex per line
  // harmonic-correction basis of sol_project: for each distinct khom value u,
  // rows 2u and 2u+1 (nz complex each) hold FFT_z(cont(exp(kh (z - Lz)))) and
  // FFT_z(cont(exp(-kh z))) (kh = 0: of z and of 1); d_hidx maps a line of any
  // member to its u.  Null when it could not be allocated: the harmonic kernel
  // then transforms the two profiles of every line itself.
  void* d_htab = nullptr;
  int* d_hidx = nullptr;
  int nhtab = 0;
  double* part = nullptr;   // per-line partial sums of the diagnostics
  std::vector<double> h_part;
  std::vector<void*> owned;
  size_t dev_bytes = 0;     // device bytes of the allocations in owned
  spx::Profiler prof;
};

namespace spx {

// error handling ------------------------------------------------------------
void set_error(const std::string& msg);
int fail(int code, const std::string& msg);
#define SPX_HIP(call)                                                       \
  do {                                                                      \
    hipError_t e_ = (call);                                                 \
    if (e_ != hipSuccess)                                                   \
      return ::spx::fail(SPX_ERR_HIP, std::string(#call) + ": " +          \
                                          hipGetErrorString(e_));           \
  } while (0)
#define SPX_CHECK_LAUNCH() SPX_HIP(hipGetLastError())

// The RK stage update of one component applied while a kernel loads it
// (include/hd/hd_rkstep2.f90:11-32, same operation order as op_rk_update):
//   v <- c0 + ((nu*(-kk2 v) - ((nl*fx)*fy)*fz + f)*dt)*rmp
// Wall conditions of the wall-update kernels: comp 0 vx (kx), comp 1 vy (ky);
// scalar mode: zero at both walls (s_constant_z).
struct WallParams {
  double tau;          // pressure time-extrapolation factor
  double vb[2][2];     // [comp][pos] wall velocity
  double nxny;         // nx*ny
  int scalar;          // 1: s_constant_z (zero at walls)
};

// Magnetic wall round trip (boundary/bboundary.f90:192-344): per component
// and wall (0: z = 0, 1: z = Lz) what happens to the wall value in the mixed
// domain: 0 nothing, 1 set to zero (int_conducting_z), 2 zero + Neumann
// reconstruction from a zero first normal derivative, 3 the same from a zero
// second derivative (conducting_z), 4 zero + Robin reconstruction with khom
// (insulating_z).  neu1 / neu2: the plan's Neumann vectors (float64, d <= 16).
struct BWallParams {
  int mode[3][2];
  double neu1[16];
  double neu2[16];
};

struct RkLoadComp {
  const void* c0;       // float64 complex
  const void* nl;       // nonlinear term, float64 or float32 complex (nl32)
  const void* f;        // float64 complex forcing
  const double* nu;     // diffusivity per (member, kx)
};
struct RkLoad {
  RkLoadComp c[2];
  int nl32;
  const double *kx, *ky, *kz, *fkx, *fky, *fkz;
  double dt, rmp;
};

// Tags the timing records of the launches in a scope with a name suffix.
struct ProfTag {
  spx::Profiler& pr;
  const char* old;
  ProfTag(spx::Profiler& prof, const char* sfx) : pr(prof), old(prof.suffix) { prof.suffix = sfx; }
  ~ProfTag() { pr.suffix = old; }
};

// Sets the member count of the launch view for a scope.
struct MemberView {
  spx_plan* p;
  int old;
  MemberView(spx_plan* plan, int m) : p(plan), old(plan->mem) { plan->mem = m; }
  ~MemberView() { p->mem = old; }
};

// RAII timing scope around one launch; records nothing when profiling is off.
struct ProfScope {
  spx_plan* p;
  int rec = -1;
  ProfScope(spx_plan* plan, const char* name, double bytes);
  ~ProfScope();
};

// FFT launchers (spx_fft.hip) ------------------------------------------------
// Every launcher works on float64 buffers; those with an `lo` argument run the
// float32 kernels of a mixed plan instead when lo is true (zfft_s2x then reads
// a float64 S and writes float32 X; the others read and write float32).
// csize: bytes per complex element (16: float64 tables, 8: float32 tables)
// want_lines: lines per workgroup at most (0: as many as max_block threads
// hold); the tables plus line buffers stay within lds_budget bytes
int build_line_fft(LineFFTHost& f, int n, int csize, int want_lines, int max_block = 256,
                   size_t lds_budget = 64 * 1024);
// the one-wavefront-per-line plan of length n (f.wave < 0 if none is compiled)
int build_wave_fft(LineFFTHost& f, int n, int csize);
// mark an x / y line plan of 512 points (one wavefront per line) for the
// compiled 512-point wave plan
void attach_xy_wave(LineFFTHost& f, int csize);
void free_line_fft(LineFFTHost& f);

// z transforms in place on S (contiguous lines).
//   cont:   1 = apply the FC continuation before a forward transform
//   scale:  multiply k < np by scale after a backward transform (1.0 = none)
int zfft_inplace(spx_plan* p, void* S, int forward, int cont, double scale);
// the continuation loop alone (fftp.fpp:755-773), in place on S
int fc_continue_only(spx_plan* p, void* S);
// forward 3D last pass: X (planes k < np) -> continuation -> z FFT -> S.
// nofft = 1: transpose only (all nz planes), for fftp2d_real_to_complex_xy.
// Xb (mixed plans, float32 X): float64 boundary planes [member][q][ky][kx], q < 2d,
// that replace X's planes z = 0..d-1, np-d..np-1 and feed the continuation.
int zfft_x2s(spx_plan* p, const void* X, void* S, int nofft, bool lo = false,
             const void* Xb = nullptr);
// float64 z IFFT of S (and of (i kz) S) at the 2d boundary planes of every
// line, for ncomp float64 spectral inputs: XVb / XZb planes [comp][member][q][ky][kx].
int zfft_s2x_bnd(spx_plan* p, const void* const* S, int ncomp, void* XVb, void* XZb);
// backward 3D first pass: S * (i k_dir) (dir 0 = none) -> z IFFT -> X, planes < nzout.
// X2 != nullptr: also X2 <- z IFFT of (i kz) S from the same load (dual).
int zfft_s2x(spx_plan* p, const void* S, void* X, int dir, int nzout, int nofft,
             void* X2 = nullptr, bool lo = false);
// y transform of X planes < nplanes: Xout <- y FFT(Xin) (in place when equal);
// Y2 != nullptr: also Y2 <- y FFT of (i ky) Xin (dual, backward only).
// cont (forward, plans with Cy > 0): FC-Gram continuation of every y line
// before the FFT (rows ny-Cy.. from the first and last oy physical rows).
int yfft(spx_plan* p, void* X, int forward, int nplanes, const void* Xin = nullptr,
         void* Y2 = nullptr, bool lo = false, bool cont = false);
int xr2c(spx_plan* p, const void* R, void* X, int nplanes, bool lo = false);
// x half complex -> real of planes < nplanes; dirx = 1 multiplies by i kx first.
int xc2r(spx_plan* p, const void* X, void* R, int nplanes, int dirx = 0, bool lo = false);
// gradient products of one component: out = ((u*Dx + v*Dy) + w*Dz)*tmp with
// Dx = c2r(i kx YV), Dy = c2r(YY), Dz = c2r(XZ) (pseudospec_hd.f90:254-312 order)
// ncomp > 1: components stored back to back (YV/YY/XZ blocks of nkx*ny*nplanes*mem
// complex, out blocks of nx*ny*nplanes*mem real), one launch
int xgrad_prod(spx_plan* p, const void* YV, const void* YY, const void* XZ, const void* u,
               const void* v, const void* w, void* out, bool lo = false, int nplanes = -1,
               int ncomp = 1);
// the three components' x stage in one launch on the np planes: RV = c2r(YV)
// (stored only when RV is given) and out_c as xgrad_prod, bit for bit
int xgrad_fused(spx_plan* p, void* const* YV, void* const* YY, void* const* XZ, void* const* RV,
                void* const* out, bool lo);
// fused v_imposebc: goto_domain_w_boundaries + noslip_z + goto_3d_fourier on vx,vy
// rk: the RK update of vx, vy (c[0], c[1]) is applied on load (null: none)
// Launch-shape overrides set by an autotuner (spx_tune_*), consulted when a
// plan builds its line FFTs and wall tables; 0 / -1 fields keep the default.
struct LineTune {
  int lines = 0;     // lines per workgroup (power of two)
  int threads = 0;   // threads per line (power of two, >= the minimum for n)
  int gtab = -1;     // Rader / Bluestein tables: 1 global memory, 0 LDS
};
struct WallTune {
  int kbw = 0;       // kappa blocks per wave: 9 or 18
  int wpe = 0;       // waves per SIMD of the register budget: 3 or 4 (kbw 9)
};
void set_line_tune(int n, int csize, const LineTune& t);
bool get_line_tune(int n, int csize, LineTune* t);
void set_wall_tune(int n, const WallTune& t);
bool get_wall_tune(int n, WallTune* t);
void clear_tunes();

// wall update on the matrix cores (spx_wall.hip): tables at plan creation
// (no-op when the (d, C) pair or nz is not covered), the launch returns
// SPX_ERR_UNSUPPORTED when the plan has no tables
int build_wall_mfma(spx_plan* p);
int wall_mfma(spx_plan* p, void* vx, void* vy, const void* pr, const WallParams& wp, int ncomp,
              const RkLoad* rk);
int noslip_fused(spx_plan* p, void* vx, void* vy, const void* pr, double tau,
                 const double vzsta[2], const double vzend[2], const RkLoad* rk = nullptr);
// s_imposebc: z IFFT, 1/nz, th = 0 at both walls, continuation + z FFT
int sconst_fused(spx_plan* p, void* th);
// sol_project stages (bctarget = 1, Neumann potential)
// d <- IFFT_z(d)/nz + harmonic a; (k0, k1): laplace_z kinds of the two walls
// (0 Dirichlet, 1 Neumann, 2 Robin), (1,1) for the velocity projection
int proj_pressure(spx_plan* p, void* d, int k0 = 1, int k1 = 1);
// the magnetic wall round trip of ncomp components (a2 unused for ncomp 2)
int bwall_fused(spx_plan* p, void* a0, void* a1, void* a2, int ncomp, const BWallParams& bp);
// proj_pressure at z = 0 and z = np-1 only, from p->pw (the rest of d keeps
// the potential): all the next stage's wall update reads
int pressure_walls(spx_plan* p, void* d);
int proj_harmonic(spx_plan* p, void* a, void* b, void* c, int k0 = 1,
                  int k1 = 1);  // v -= grad(harmonic)
// op_poisson_project_bc + proj_harmonic in one pass per line (plans with the
// harmonic basis table and nz <= 1088; SPX_ERR_UNSUPPORTED otherwise)
int project_harmonic(spx_plan* p, void* a, void* b, void* c, void* d, const RkLoad* rk,
                     bool walls_only);
int build_harmonic_table(spx_plan* p);  // p->d_htab / d_hidx (kept null on allocation failure)
// z IFFT of a linear combination of derivatives, reduced per line over physical z
// mode 0: sum |u|^2 ; the combination is acc = m0(in0) (+/-) m1(in1) (+/-) m2(in2)
struct LinComb {
  const void* in[3];
  int dir[3];    // 0 none, 1 x, 2 y, 3 z (i k_dir), 4 kk2
  int sgn[3];    // +1 / -1 for terms 1,2 (term 0 always +1); 0 = term absent
};
// gradre (pseudospec_hd.f90:209-319) through the plan's current work arrays
// XV, XZ, XY, RV, RO (and the boundary-plane buffers of mixed plans); keep_rv:
// leave v in real space in RV (advect on the same v reuses it)
int gradre_impl(spx_plan* p, const void* const v[3], void* const out[3], bool keep_rv = true);

int zline_sumsq(spx_plan* p, const LinComb& lc, double* line_out);
int zline_dot(spx_plan* p, const void* a, const void* b, double* line_out);
// sum_k<np Re(IFFT(lincomb) conj IFFT(b)) per line
int zline_dot_lc(spx_plan* p, const LinComb& lc, const void* b, double* line_out);
int zline_walls(spx_plan* p, const void* a, double* line_out /*2 per line*/);
// IFFT_z(lincomb) at z = 0 and z = Lz: (re, im) bottom, (re, im) top per line
int zline_wall_values(spx_plan* p, const LinComb& lc, double* line_out /*4 per line*/);

// pointwise launchers (spx_ops.hip) -------------------------------------------
int op_derivk(spx_plan* p, const void* a, void* b, int dir);
int op_laplak(spx_plan* p, const void* a, void* b);
int op_filter(spx_plan* p, void* a);
int op_scale(spx_plan* p, void* a, size_t n, double s);
// nl: float64, or float32 with nl_lo (the nonlinear terms of a mixed plan)
// nu[c]: device table of the component's diffusivity per (member, kx), nb*nkx values
int op_rk_update(spx_plan* p, int ncomp, void* const* v, const void* const* c0, void* const* nl,
                 const void* const* f, const void* const* src, const double* src_coef,
                 const double* const* nu, double dt, double rmp, bool nl_lo = false);
int op_poisson_project(spx_plan* p, void* a, void* b, void* c, void* d);
// poisson_project plus the wall data bc of sol_project (boundary_mod.fpp:262-352)
// walls_only: d is left unwritten at z = 0 and z = np-1 and IFFT_z(d)/nz at
// those two points goes to p->pw (for pressure_walls)
// MHD pointwise kernels (spx_ops.hip)
int op_curlk(spx_plan* p, const void* a, const void* b, void* c, int dir);
int op_curl3(spx_plan* p, const void* const* in, void* const* out);
int op_cross_real(spx_plan* p, const void* const* A, const void* const* B, void* const* O,
                  long n);
int op_set_mode0(spx_plan* p, void* f0, void* f1, void* f2, const double v[3]);
int op_mhd_update(spx_plan* p, void* const* v, void* const* a, const void* const* c1,
                  const void* const* c9, const void* const* N, const void* const* L,
                  const void* const* M, const void* const* f, const void* const* m,
                  const double* nu, double mu, double dt, double rmp);
// sol_project's first pass for a condition on the potential (bctarget 0),
// wall kinds k0, k1 in {0 Dirichlet, 2 Robin}; bc -> p->bc
int op_poisson_project_pot(spx_plan* p, void* a, void* b, void* c, void* d, int k0, int k1);
int op_poisson_project_bc(spx_plan* p, void* a, void* b, void* c, void* d,
                          const RkLoad* rk = nullptr,  // rk: update of c (c[0]) on load
                          bool walls_only = false);
int op_axpy_field(spx_plan* p, void* y, const void* x, double alpha);
int op_copy_fields(spx_plan* p, int n, const void* const* src, void* const* dst);
// N_c += (2 Omega x v)_c, c = x, y, z (include/rotbouss/rotbouss_rkstep2.f90:13-19);
// nl_lo: the nonlinear terms are float32
int op_coriolis(spx_plan* p, void* const* nl, const void* const* v, const double omega[3],
                bool nl_lo);
// d <- i kx a + i ky b + i kz c (all active members)
int op_div_spec(spx_plan* p, const void* a, const void* b, const void* c, void* d);
// per member: block maxima of |r| over n values at member stride ms -> out[member][block]
int op_absmax(spx_plan* p, const void* r, long n, long ms, double* out, int* nblocks);
int op_real_div3(spx_plan* p, void* r, double nx, double ny, double nz);

}  // namespace spx
