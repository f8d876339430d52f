// C ABI of libspx: plan (FCPLAN), transforms (fftp.fpp), operators
// (pseudospec_*.f90), boundary conditions (boundary/), diagnostics and the
// device-resident RK solver state (include/hd, include/bouss, specter.fpp).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>

#include "spx_common.h"

namespace spx {

static thread_local std::string g_err;

void set_error(const std::string& msg) { g_err = msg; }
int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

ProfScope::ProfScope(spx_plan* plan, const char* name0, double bytes) : p(plan) {
  if (!p || !p->prof.on) return;
  Profiler& pr = p->prof;
  const std::string name = std::string(name0) + pr.suffix;
  int kind = -1;
  for (size_t i = 0; i < pr.names.size(); ++i)
    if (pr.names[i] == name) kind = int(i);
  if (kind < 0) {
    pr.names.push_back(name);
    kind = int(pr.names.size()) - 1;
  }
  while (pr.pool.size() < pr.used + 2) {
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return;
    pr.pool.push_back(e);
  }
  ProfRec r;
  r.kind = kind;
  r.a = pr.pool[pr.used++];
  r.b = pr.pool[pr.used++];
  r.bytes = bytes;
  if (hipEventRecord(r.a, p->stream) != hipSuccess) return;
  pr.recs.push_back(r);
  rec = int(pr.recs.size()) - 1;
}

ProfScope::~ProfScope() {
  if (rec >= 0) (void)hipEventRecord(p->prof.recs[rec].b, p->stream);
}

static int alloc_dev(spx_plan* p, void** ptr, size_t bytes) {
  const hipError_t e = hipMalloc(ptr, bytes);
  if (e != hipSuccess) {
    *ptr = nullptr;
    return fail(SPX_ERR_ALLOC,
                "hipMalloc(" + std::to_string(bytes) + ") failed: " + hipGetErrorString(e));
  }
  if (p) {
    p->owned.push_back(*ptr);
    p->dev_bytes += bytes;
  }
  return SPX_OK;
}

template <typename V>
static int upload(spx_plan* p, double** dst, const V& host) {
  int rc = alloc_dev(p, reinterpret_cast<void**>(dst), host.size() * sizeof(double));
  if (rc) return rc;
  SPX_HIP(hipMemcpy(*dst, host.data(), host.size() * sizeof(double), hipMemcpyHostToDevice));
  return SPX_OK;
}

// device bytes of the work arrays ensure_work allocates: the gradre set (with
// the boundary-plane buffers of mixed plans) and the advect additions
static void work_sizes(const spx_plan* p, size_t* hd, size_t* bouss) {
  const size_t cb = p->nc * p->nb * p->csize, rb = p->nr * p->nb * p->rsize;
  size_t h = 3 * (3 * cb + 2 * rb) + p->nr * p->nb * 8;
  if (p->fp32 && p->g.Cz > 0) {
    const size_t d2 = 2 * size_t(p->g.oz), nbm = size_t(p->nb);
    const size_t bc = size_t(p->nkx) * p->g.ny * d2 * nbm * 16;
    const size_t br = size_t(p->g.nx) * p->g.ny * d2 * nbm * 8;
    h += 16 * bc + 7 * br;
  }
  *hd = h;
  *bouss = 3 * cb + rb;
}

// device bytes of one solver state (all members): 10 float64 spectral fields,
// the 3 nonlinear terms in the pipeline's precision, the diffusivity tables;
// Boussinesq adds th, fs, the th start copy and its nonlinear term
static size_t state_size(const spx_plan* p, bool bouss) {
  const size_t c16 = p->nc * size_t(p->nb) * 16, cs = p->nc * size_t(p->nb) * p->csize;
  size_t b = 10 * c16 + 3 * cs + 2 * size_t(p->nb) * p->nkx * sizeof(double);
  if (bouss) b += 3 * c16 + cs;
  return b;
}

// gradre / advect work arrays in the nonlinear pipeline's precision
static int ensure_work(spx_plan* p, bool bouss) {
  int rc;
  const size_t cb = p->nc * p->nb * p->csize, rb = p->nr * p->nb * p->rsize;
  if (!p->XV[0]) {
    for (int c = 0; c < 3; ++c) {
      if ((rc = alloc_dev(p, &p->XV[c], cb)) || (rc = alloc_dev(p, &p->XZ[c], cb)) ||
          (rc = alloc_dev(p, &p->XY[c], cb)) || (rc = alloc_dev(p, &p->RV[c], rb)) ||
          (rc = alloc_dev(p, &p->RO[c], rb)))
        return rc;
    }
    if ((rc = alloc_dev(p, &p->R0, p->nr * p->nb * 8))) return rc;
  }
  if (bouss && !p->work_bouss) {
    if ((rc = alloc_dev(p, &p->XV[3], cb)) || (rc = alloc_dev(p, &p->XZ[3], cb)) ||
        (rc = alloc_dev(p, &p->XY[3], cb)) || (rc = alloc_dev(p, &p->RO[3], rb)))
      return rc;
    p->work_bouss = true;
  }
  if (p->fp32 && p->g.Cz > 0 && !p->XVb) {
    const size_t d2 = 2 * size_t(p->g.oz), nbm = size_t(p->nb);
    const size_t bc = size_t(p->nkx) * p->g.ny * d2 * nbm * 16;   // one component
    const size_t br = size_t(p->g.nx) * p->g.ny * d2 * nbm * 8;
    if ((rc = alloc_dev(p, &p->XVb, 4 * bc)) || (rc = alloc_dev(p, &p->XZb, 4 * bc)) ||
        (rc = alloc_dev(p, &p->XYb, 4 * bc)) || (rc = alloc_dev(p, &p->RVb, 3 * br)) ||
        (rc = alloc_dev(p, &p->ROb, 4 * br)) || (rc = alloc_dev(p, &p->XFb, 4 * bc)))
      return rc;
  }
  return SPX_OK;
}

// complex / real boundary-plane block of component c in the current view
static inline void* bnd_c(const spx_plan* p, void* base, int c) {
  return static_cast<char*>(base) +
         size_t(c) * p->mem * 2 * p->g.oz * size_t(p->nkx) * p->g.ny * 16;
}
static inline void* bnd_r(const spx_plan* p, void* base, int c) {
  return static_cast<char*>(base) +
         size_t(c) * p->mem * 2 * p->g.oz * size_t(p->g.nx) * p->g.ny * 8;
}

// 3D transforms on internal buffers -------------------------------------------
// lo: float32 buffers and kernels of a mixed plan (X holds float32 data then);
// xb: float64 boundary planes for the continuation (mixed plans)
static int r2c3(spx_plan* p, const void* in, void* out, bool lo = false,
                const void* xb = nullptr) {
  int rc;
  if ((rc = xr2c(p, in, p->X, p->np, lo))) return rc;
  if ((rc = yfft(p, p->X, 1, p->np, nullptr, nullptr, lo, p->Cy > 0))) return rc;
  return zfft_x2s(p, p->X, out, 0, lo, xb);
}

static int c2r3(spx_plan* p, const void* in, void* out, int dirn, int nplanes) {
  int rc;
  if ((rc = zfft_s2x(p, in, p->X, dirn, nplanes, 0))) return rc;
  if ((rc = yfft(p, p->X, 0, nplanes))) return rc;
  return xc2r(p, p->X, out, nplanes);
}

// ---- MHD building blocks -----------------------------------------------------
static int ensure_mhd_work(spx_plan* p) {
  if (p->MR[0]) return SPX_OK;
  const size_t rb = size_t(p->g.nx) * p->g.ny * p->np * p->nb * 8;
  for (int q = 0; q < 12; ++q)
    if (int rc = alloc_dev(p, &p->MR[q], rb)) return rc;
  return SPX_OK;
}

// real values (physical planes, members of the view packed) of A and B, then
// out = A x B through real space (pseudospec_hd.f90:383-399, pseudospec_mhd.f90
// vector).  RA / RB receive the real fields unless have_a / have_b say they
// already hold them; RO is scratch.
static int cross_spec(spx_plan* p, const void* const* A, const void* const* B, void* const* out,
                      void* const* RA, void* const* RB, void* const* RO, bool have_a,
                      bool have_b) {
  int rc;
  for (int c = 0; c < 3 && !have_a; ++c)
    if ((rc = c2r3(p, A[c], RA[c], 0, p->np))) return rc;
  for (int c = 0; c < 3 && !have_b; ++c)
    if ((rc = c2r3(p, B[c], RB[c], 0, p->np))) return rc;
  const long n = long(p->g.nx) * p->g.ny * p->np * p->mem;
  if ((rc = op_cross_real(p, RA, RB, RO, n))) return rc;
  for (int c = 0; c < 3; ++c)
    if ((rc = r2c3(p, RO[c], out[c]))) return rc;
  return SPX_OK;
}

// sol_project with the condition on the potential (bctarget 0; wall kinds
// k0, k1: 0 Dirichlet, 2 Robin; boundary_mod.fpp:197-402)
static int sol_project_pot(spx_plan* p, void* a, void* b, void* c, void* d, int k0, int k1) {
  int rc;
  if ((rc = op_poisson_project_pot(p, a, b, c, d, k0, k1))) return rc;
  if ((rc = proj_pressure(p, d, k0, k1))) return rc;
  return proj_harmonic(p, a, b, c, k0, k1);
}

// a_imposebc_and_project (bboundary.f90:100-189) for the walls' kinds
// (0 conducting, 1 insulating): az(0,0,0) = 0; conducting walls zero the
// tangential components; projection with Dirichlet / Robin conditions on the
// potential; then the wall values are reconstructed (conducting: zero second
// derivative of ax, ay and zero first derivative of az; insulating: Robin
// with khom for each component)
static int a_imposebc(spx_plan* p, void* ax, void* ay, void* az, void* ph, int bs, int be) {
  if (!p->neu_loaded[0] || !p->neu_loaded[1])
    return fail(SPX_ERR_STATE, "magnetic walls need the Neumann tables (spx_plan_load_neumann)");
  int rc;
  const double zero[3] = {0.0, 0.0, 0.0};
  if ((rc = op_set_mode0(p, nullptr, nullptr, az, zero))) return rc;
  BWallParams bp;
  std::memset(&bp, 0, sizeof(bp));
  std::memcpy(bp.neu1, p->neu[0], sizeof(bp.neu1));
  std::memcpy(bp.neu2, p->neu[1], sizeof(bp.neu2));
  const int kind[2] = {bs, be};
  if (bs == 0 || be == 0) {
    for (int w = 0; w < 2; ++w) bp.mode[0][w] = bp.mode[1][w] = kind[w] == 0 ? 1 : 0;
    if ((rc = bwall_fused(p, ax, ay, nullptr, 2, bp))) return rc;
  }
  if ((rc = sol_project_pot(p, ax, ay, az, ph, 2 * bs, 2 * be))) return rc;
  for (int w = 0; w < 2; ++w) {
    bp.mode[0][w] = bp.mode[1][w] = kind[w] == 0 ? 3 : 4;
    bp.mode[2][w] = kind[w] == 0 ? 2 : 4;
  }
  return bwall_fused(p, ax, ay, az, 3, bp);
}

static double diag_tmp(const spx_plan* p) {
  const double t = double(p->g.nx) * double(p->g.ny) * double(p->g.nz);
  return 1.0 / (t * t) / double(p->g.nz - p->g.Cz);
}

static int fetch_parts(spx_plan* p, size_t count) {
  if (p->h_part.size() < count) p->h_part.resize(count);
  SPX_HIP(hipMemcpyAsync(p->h_part.data(), p->part, count * sizeof(double),
                         hipMemcpyDeviceToHost, p->stream));
  SPX_HIP(hipStreamSynchronize(p->stream));
  return SPX_OK;
}

// weighted sum over the (kx,ky) lines of member b: weight 1 for kx = 0, 2 otherwise
static double weighted_sum(const spx_plan* p, const std::vector<double>& acc, int b = 0) {
  const long ny = p->g.ny, nl = long(p->g.ny) * p->nkx, o = long(b) * nl;
  double s0 = 0.0, s1 = 0.0;
  for (long l = 0; l < ny; ++l) s0 += acc[o + l];
  for (long l = ny; l < nl; ++l) s1 += acc[o + l];
  return s0 + 2.0 * s1;
}

static inline void* coff(void* a, size_t elems, int b) {
  return static_cast<char*>(a) + size_t(b) * elems * 16;
}
static inline const void* coff(const void* a, size_t elems, int b) {
  return static_cast<const char*>(a) + size_t(b) * elems * 16;
}
static inline void* roff(void* a, size_t elems, int b) {
  return static_cast<char*>(a) + size_t(b) * elems * 8;
}
static inline const void* roff(const void* a, size_t elems, int b) {
  return static_cast<const char*>(a) + size_t(b) * elems * 8;
}

}  // namespace spx

using namespace spx;

struct spx_state {
  spx_plan* p = nullptr;
  spx_params prm;
  void *vx = nullptr, *vy = nullptr, *vz = nullptr, *pr = nullptr;
  void *fx = nullptr, *fy = nullptr, *fz = nullptr, *th = nullptr, *fs = nullptr;
  void *C1 = nullptr, *C2 = nullptr, *C3 = nullptr, *C7 = nullptr;
  void* N[4] = {nullptr, nullptr, nullptr, nullptr};
  // MHD: vector potential, its projection potential, electromotive forcing,
  // stage start copies (C9-C11), b = curl a (C12-C14), Lorentz force and
  // electromotive force (C15-C17 at their two uses)
  void *ax = nullptr, *ay = nullptr, *az = nullptr, *ph = nullptr;
  void *mx = nullptr, *my = nullptr, *mz = nullptr;
  void* C9[3] = {nullptr, nullptr, nullptr};
  void* B[3] = {nullptr, nullptr, nullptr};
  void* L[3] = {nullptr, nullptr, nullptr};
  void* M[3] = {nullptr, nullptr, nullptr};
  std::vector<void*> owned;
  // diffusivities per (member, kx): nu for the velocity, kappa for th
  double* d_nu = nullptr;
  double* d_kappa = nullptr;
  std::vector<double> h_nu, h_kappa;
  // one full RK step (rkstep1 + ord stages) captured as a HIP graph, per
  // active-member count
  bool use_graph = true;
  hipStream_t cap = nullptr;
  std::map<int, std::pair<hipGraph_t, hipGraphExec_t>> graphs;
};

#define SPX_REQ(cond, msg) \
  do {                     \
    if (!(cond)) return fail(SPX_ERR_ARG, msg); \
  } while (0)

extern "C" {

int spx_last_error(char* buf, size_t len) {
  if (buf && len) {
    std::snprintf(buf, len, "%s", g_err.c_str());
  }
  return int(g_err.size());
}

// dir = A*Q^T (fcgram_mod.f90:250-252), textbook sum in k order, row-major [C][d]
static std::vector<double> dir_matrix(const double* A, const double* Q, int C, int d) {
  std::vector<double> dir(size_t(C) * d, 0.0);
  for (int ii = 0; ii < C; ++ii)
    for (int jj = 0; jj < d; ++jj) {
      double s = 0.0;
      for (int k = 0; k < d; ++k) s = s + A[ii + size_t(C) * k] * Q[jj + size_t(d) * k];
      dir[size_t(ii) * d + jj] = s;
    }
  return dir;
}

static int plan_create(const spx_grid* g, const double* A, const double* Q, int Cy, int oy,
                       const double* Ay, const double* Qy, int nmembers, void* stream,
                       spx_plan** out);

int spx_plan_create(const spx_grid* g, const double* A, const double* Q, void* stream,
                    spx_plan** out) {
  return plan_create(g, A, Q, 0, 0, nullptr, nullptr, 1, stream, out);
}

int spx_plan_create_batch(const spx_grid* g, const double* A, const double* Q, int nmembers,
                          void* stream, spx_plan** out) {
  return plan_create(g, A, Q, 0, 0, nullptr, nullptr, nmembers, stream, out);
}

int spx_plan_create_fcy(const spx_grid* g, const double* A, const double* Q, int Cy, int oy,
                        const double* Ay, const double* Qy, int nmembers, void* stream,
                        spx_plan** out) {
  return plan_create(g, A, Q, Cy, oy, Ay, Qy, nmembers, stream, out);
}

static int plan_create(const spx_grid* g, const double* A, const double* Q, int Cy, int oy,
                       const double* Ay, const double* Qy, int nmembers, void* stream,
                       spx_plan** out) {
  SPX_REQ(g && out, "spx_plan_create: null argument");
  *out = nullptr;
  SPX_REQ(nmembers >= 1, "nmembers must be >= 1");
  SPX_REQ(g->nx >= 2 && g->nx % 2 == 0, "nx must be even and >= 2");
  SPX_REQ(g->ny >= 2 && g->ny % 2 == 0, "ny must be even and >= 2");
  if (Cy > 0) {
    SPX_REQ(Ay && Qy, "FC-Gram tables Ay and Qy are required when Cy > 0");
    SPX_REQ(oy >= 1, "matching order oy must be >= 1 when Cy > 0");
    SPX_REQ(g->ny - Cy >= oy + 1, "too few physical y points for the matching order");
    if (g->precision != SPX_FP64)
      return fail(SPX_ERR_UNSUPPORTED,
                  "non-periodic y needs SPX_FP64: the y continuation amplifies float32 "
                  "rounding of the boundary rows by its row sums (~1e4)");
  } else {
    SPX_REQ(Cy == 0 && oy == 0, "Cy == 0 requires oy == 0 (periodic y)");
  }
  SPX_REQ(g->nz >= 2, "nz must be >= 2");
  SPX_REQ(g->Cz >= 0 && g->Cz < g->nz, "Cz out of range");
  SPX_REQ(g->Lx > 0 && g->Ly > 0 && g->Lz > 0, "box lengths must be positive");
  SPX_REQ(g->precision == SPX_FP64 || g->precision == SPX_FP32_MIXED,
          "precision must be SPX_FP64 or SPX_FP32_MIXED");
  if (g->Cz > 0) {
    SPX_REQ(A && Q, "FC-Gram tables A and Q are required when Cz > 0");
    SPX_REQ(g->oz >= 1, "matching order oz must be >= 1 when Cz > 0");
    SPX_REQ(g->nz - g->Cz >= g->oz + 1, "too few physical z points for the matching order");
  } else {
    SPX_REQ(g->oz == 0, "Cz == 0 requires oz == 0 (periodic z)");
  }
  spx_plan* p = new spx_plan();
  p->g = *g;
  p->nb = p->nact = nmembers;
  p->mem = 1;
  p->stream = static_cast<hipStream_t>(stream);
  if (const char* e = std::getenv("SPX_FUSED_PROJECT")) p->fused_project = std::atoi(e) != 0;
  if (const char* e = std::getenv("SPX_XFUSED")) p->xfused = std::atoi(e) != 0;
  if (hipGetDevice(&p->device) != hipSuccess) {
    delete p;
    return fail(SPX_ERR_HIP, "spx_plan_create: hipGetDevice failed");
  }
  const int nx = g->nx, ny = g->ny, nz = g->nz, C = g->Cz, d = g->oz;
  p->nkx = nx / 2 + 1;
  p->np = nz - C;
  p->nc = size_t(nz) * ny * p->nkx;
  p->nr = size_t(nx) * ny * nz;
  p->fp32 = g->precision == SPX_FP32_MIXED;
  p->csize = p->fp32 ? 8 : 16;
  p->rsize = p->fp32 ? 4 : 8;
  int rc = SPX_OK;
  auto bail = [&](int code) {
    spx_plan_destroy(p);
    return code;
  };
  if (C > 0) {
    p->h_dir = dir_matrix(A, Q, C, d);
    if ((rc = upload(p, &p->d_dir, p->h_dir))) return bail(rc);
  }
  p->Cy = Cy;
  p->oy = oy;
  p->npy = ny - Cy;
  if (Cy > 0) {
    p->h_diry = dir_matrix(Ay, Qy, Cy, oy);
    if ((rc = upload(p, &p->d_diry, p->h_diry))) return bail(rc);
  }
  // grids and wavenumbers (specter.fpp:683-800); non-periodic y: Ly is the
  // physical length, dy = Ly/(ny-Cy-1) and Dky = 2 pi/(dy ny) (:720-723)
  const double PI = 3.14159265358979323846;
  p->Dkx = 1.0 / g->Lx;
  p->Dky = Cy > 0 ? 2.0 * PI / (g->Ly / double(ny - Cy - 1) * ny) : 1.0 / g->Ly;
  if (C > 0) {
    p->dz = g->Lz / double(nz - C - 1);
    p->Dkz = 2.0 * PI / (p->dz * nz);
  } else {
    p->dz = g->Lz * 2.0 * PI / nz;
    p->Dkz = 1.0 / g->Lz;
  }
  p->h_kx.resize(p->nkx);
  p->h_ky.resize(ny);
  p->h_kz.resize(nz);
  p->h_z.resize(nz);
  for (int i = 0; i < p->nkx; ++i) p->h_kx[i] = double(i < nx / 2 ? i : i - nx) * p->Dkx;
  for (int j = 0; j < ny; ++j) p->h_ky[j] = double(j < ny / 2 ? j : j - ny) * p->Dky;
  for (int k = 0; k < nz; ++k) p->h_kz[k] = double(k < (nz + 1) / 2 ? k : k - nz) * p->Dkz;
  for (int k = 0; k < nz; ++k) p->h_z[k] = p->dz * double(k);
  p->h_khom.resize(size_t(p->nkx) * ny);
  for (int i = 0; i < p->nkx; ++i)
    for (int j = 0; j < ny; ++j)
      p->h_khom[size_t(j) + size_t(ny) * i] =
          std::sqrt(p->h_kx[i] * p->h_kx[i] + p->h_ky[j] * p->h_ky[j]);
  // fc_filter factors (pseudospec_hd.f90:1099-1105)
  const double alpha = 16 * std::log(10.0), pw = 2 * 50.0;
  p->h_fkx.resize(p->nkx);
  p->h_fky.resize(ny);
  p->h_fkz.resize(nz);
  for (int i = 0; i < p->nkx; ++i)
    p->h_fkx[i] = std::exp(-alpha * std::pow(2 * p->h_kx[i] / nx / p->Dkx, pw));
  for (int j = 0; j < ny; ++j)
    p->h_fky[j] = std::exp(-alpha * std::pow(2 * p->h_ky[j] / ny / p->Dky, pw));
  for (int k = 0; k < nz; ++k)
    p->h_fkz[k] = std::exp(-alpha * std::pow(2 * p->h_kz[k] / nz / p->Dkz, pw));
  // tables indexed by the (member, kx) line group repeat once per member
  auto repeat = [&](const std::vector<double>& v) {
    std::vector<double> r;
    r.reserve(v.size() * size_t(nmembers));
    for (int b = 0; b < nmembers; ++b) r.insert(r.end(), v.begin(), v.end());
    return r;
  };
  if ((rc = upload(p, &p->d_kx, repeat(p->h_kx))) || (rc = upload(p, &p->d_ky, p->h_ky)) ||
      (rc = upload(p, &p->d_kz, p->h_kz)) || (rc = upload(p, &p->d_z, p->h_z)) ||
      (rc = upload(p, &p->d_khom, repeat(p->h_khom))) ||
      (rc = upload(p, &p->d_fkx, repeat(p->h_fkx))) || (rc = upload(p, &p->d_fky, p->h_fky)) ||
      (rc = upload(p, &p->d_fkz, p->h_fkz)))
    return bail(rc);
  // line FFT plans
  if ((rc = build_line_fft(p->fx, nx, 16, 0))) return bail(rc);
  // y lines (stride nkx in memory): up to 512 threads per workgroup, so a row
  // of a workgroup's load spans 8 consecutive kx at ny = 512
  if ((rc = build_line_fft(p->fy, ny, 16, 16, 512, 80 * 1024))) return bail(rc);
  attach_xy_wave(p->fx, 16);
  attach_xy_wave(p->fy, 16);
  if ((rc = build_line_fft(p->fz, nz, 16, 16))) return bail(rc);
  p->fz2 = p->fz.with_lines(p->fz.nl / 2);
  if ((rc = build_wave_fft(p->fzw, nz, 16))) return bail(rc);
  if (C > 0) {
    // exp(+2 pi i k (np-1) / nz): the top-wall value of a z line as one DFT sum
    std::vector<double> w(size_t(nz) * 2);
    const long double TWO_PI = 6.283185307179586476925286766559005768L;
    for (int k = 0; k < nz; ++k) {
      const long r = (long(k) * (p->np - 1)) % nz;
      const long double ang = TWO_PI * (long double)r / (long double)nz;
      w[2 * size_t(k)] = double(cosl(ang));
      w[2 * size_t(k) + 1] = double(sinl(ang));
    }
    if ((rc = upload(p, reinterpret_cast<double**>(&p->d_wtop), w))) return bail(rc);
    // the wall solvers' tables; plans with a y continuation have no boundary
    // conditions (the reference stops at boundary_mod.fpp:104-105, :179-180)
    if (Cy == 0) {
      if ((rc = build_harmonic_table(p))) return bail(rc);
      if ((rc = build_wall_mfma(p))) return bail(rc);
    }
  }
  if (p->fp32) {
    if ((rc = build_line_fft(p->fx32, nx, 8, 0))) return bail(rc);
    if ((rc = build_line_fft(p->fy32, ny, 8, 16, 512, 80 * 1024))) return bail(rc);
    attach_xy_wave(p->fx32, 8);
    attach_xy_wave(p->fy32, 8);
    if ((rc = build_line_fft(p->fz32, nz, 8, 16))) return bail(rc);
    if ((rc = build_wave_fft(p->fzw32, nz, 8))) return bail(rc);
    if (C > 0) {
      if (d > 8) return bail(fail(SPX_ERR_UNSUPPORTED, "mixed plans: matching order above 8"));
      // exp(+2 pi i k z / nz) for z = 1 (first nz entries) and z = np - d
      std::vector<double> w(size_t(2) * nz * 2);
      const long double TWO_PI = 6.283185307179586476925286766559005768L;
      for (int t = 0; t < 2; ++t) {
        const long z = t == 0 ? 1 : long(p->np) - d;
        for (int k = 0; k < nz; ++k) {
          const long r = (long(k) * z) % nz;
          const long double a = TWO_PI * (long double)r / (long double)nz;
          w[2 * (size_t(t) * nz + k)] = double(cosl(a));
          w[2 * (size_t(t) * nz + k) + 1] = double(sinl(a));
        }
      }
      if ((rc = upload(p, reinterpret_cast<double**>(&p->d_wbnd), w))) return bail(rc);
    }
  }
  // scratch
  const size_t nlines = size_t(ny) * p->nkx * nmembers;
  if ((rc = alloc_dev(p, &p->X, p->nc * nmembers * 16))) return bail(rc);
  if ((rc = alloc_dev(p, &p->bc, nlines * 2 * 16))) return bail(rc);
  if ((rc = alloc_dev(p, &p->pw, nlines * 2 * 16))) return bail(rc);
  // per-line partials of the diagnostics (up to 4 per line), block maxima (64 per member)
  const size_t nparts = std::max(nlines * 4, size_t(64) * nmembers);
  if ((rc = alloc_dev(p, reinterpret_cast<void**>(&p->part), nparts * sizeof(double))))
    return bail(rc);
  *out = p;
  return SPX_OK;
}

int spx_plan_destroy(spx_plan* p) {
  if (!p) return SPX_OK;
  if (p->stream) (void)hipStreamSynchronize(p->stream);
  else (void)hipDeviceSynchronize();
  free_line_fft(p->fx);
  free_line_fft(p->fy);
  free_line_fft(p->fz);
  free_line_fft(p->fx32);
  free_line_fft(p->fy32);
  free_line_fft(p->fz32);
  free_line_fft(p->fzw);
  free_line_fft(p->fzw32);
  for (hipEvent_t e : p->prof.pool) (void)hipEventDestroy(e);
  for (void* q : p->owned) (void)hipFree(q);
  delete p;
  return SPX_OK;
}

int spx_plan_get_info(const spx_plan* p, spx_plan_info* info) {
  SPX_REQ(p && info, "spx_plan_get_info: null argument");
  info->nkx = p->nkx;
  info->np = p->np;
  info->n_spectral = p->nc;
  info->n_real = p->nr;
  info->bytes_spectral = p->nc * 16;
  info->bytes_real = p->nr * 8;
  info->z_fft_len = p->fz.M;
  info->z_algorithm = p->fz.mode;
  size_t tab = 0;
  for (const spx::LineFFTHost* f :
       {&p->fx, &p->fy, &p->fz, &p->fx32, &p->fy32, &p->fz32, &p->fzw, &p->fzw32})
    tab += f->bytes;
  info->device_bytes = p->dev_bytes + tab;
  size_t whd, wb;
  work_sizes(p, &whd, &wb);
  info->work_bytes = p->XV[0] ? 0 : whd;
  info->work_bytes_bouss = p->work_bouss ? 0 : wb;
  info->state_bytes = state_size(p, false);
  info->state_bytes_bouss = state_size(p, true);
  info->z_threads_per_line = p->fz.tpl;
  info->z_lines_per_block = p->fz.nl;
  info->z_tables_global = (p->fz.mode != 0 && p->fz.tab == p->fz.M) ? 1 : 0;
  info->z_wave_plan = p->fzw.wave;
  info->z_wave_plan_lo = p->fp32 ? p->fzw32.wave : -1;
  return SPX_OK;
}

int spx_tune_line(int n, int csize, int lines, int threads, int tables_global) {
  SPX_REQ(n >= 1 && (csize == 8 || csize == 16), "spx_tune_line: bad length or element size");
  LineTune t;
  t.lines = lines;
  t.threads = threads;
  t.gtab = tables_global;
  set_line_tune(n, csize, t);
  return SPX_OK;
}

int spx_tune_wall(int n, int blocks_per_wave, int waves_per_simd) {
  SPX_REQ(n >= 1, "spx_tune_wall: bad length");
  WallTune t;
  t.kbw = blocks_per_wave;
  t.wpe = waves_per_simd;
  set_wall_tune(n, t);
  return SPX_OK;
}

int spx_tune_clear(void) {
  clear_tunes();
  return SPX_OK;
}

int spx_plan_get_members(const spx_plan* p, int* nmembers, int* nactive) {
  SPX_REQ(p && nmembers && nactive, "spx_plan_get_members: null argument");
  *nmembers = p->nb;
  *nactive = p->nact;
  return SPX_OK;
}

int spx_plan_set_active(spx_plan* p, int nactive) {
  SPX_REQ(p, "spx_plan_set_active: null plan");
  SPX_REQ(nactive >= 1 && nactive <= p->nb, "spx_plan_set_active: nactive out of range");
  p->nact = nactive;
  return SPX_OK;
}

int spx_plan_sync(spx_plan* p) {
  SPX_REQ(p, "spx_plan_sync: null plan");
  SPX_HIP(hipStreamSynchronize(p->stream));
  return SPX_OK;
}

// ---- transforms -------------------------------------------------------------
// The 3D and 2D transforms take member-contiguous real arrays, so batched
// plans run them member by member; everything line-based runs on all active
// members at once.
int spx_fftp3d_real_to_complex(spx_plan* p, const void* in, void* out) {
  SPX_REQ(p && in && out, "spx_fftp3d_real_to_complex: null argument");
  MemberView mv(p, 1);
  for (int b = 0; b < p->nact; ++b)
    if (int rc = r2c3(p, roff(in, p->nr, b), coff(out, p->nc, b))) return rc;
  return SPX_OK;
}

int spx_fftp3d_complex_to_real(spx_plan* p, void* in, void* out) {
  SPX_REQ(p && in && out, "spx_fftp3d_complex_to_real: null argument");
  MemberView mv(p, 1);
  for (int b = 0; b < p->nact; ++b)
    if (int rc = c2r3(p, coff(in, p->nc, b), roff(out, p->nr, b), 0, p->g.nz)) return rc;
  return SPX_OK;
}

int spx_fftp2d_real_to_complex_xy(spx_plan* p, const void* in, void* out) {
  SPX_REQ(p && in && out, "spx_fftp2d_real_to_complex_xy: null argument");
  MemberView mv(p, 1);
  int rc;
  for (int b = 0; b < p->nact; ++b) {
    if ((rc = xr2c(p, roff(in, p->nr, b), p->X, p->g.nz))) return rc;
    if ((rc = yfft(p, p->X, 1, p->g.nz))) return rc;
    if ((rc = zfft_x2s(p, p->X, coff(out, p->nc, b), 1))) return rc;
  }
  return SPX_OK;
}

int spx_fftp2d_complex_to_real_xy(spx_plan* p, void* in, void* out) {
  SPX_REQ(p && in && out, "spx_fftp2d_complex_to_real_xy: null argument");
  MemberView mv(p, 1);
  int rc;
  for (int b = 0; b < p->nact; ++b) {
    if ((rc = zfft_s2x(p, coff(in, p->nc, b), p->X, 0, p->g.nz, 1))) return rc;
    if ((rc = yfft(p, p->X, 0, p->g.nz))) return rc;
    if ((rc = xc2r(p, p->X, roff(out, p->nr, b), p->g.nz))) return rc;
  }
  return SPX_OK;
}

// The transforms of non-periodic y, on the S layout.  Each works for any
// plan; with Cy = 0 the yz transforms have no y continuation.
int spx_fftp1d_real_to_complex_x(spx_plan* p, const void* in, void* out) {
  SPX_REQ(p && in && out, "spx_fftp1d_real_to_complex_x: null argument");
  MemberView mv(p, 1);
  int rc;
  for (int b = 0; b < p->nact; ++b) {
    if ((rc = xr2c(p, roff(in, p->nr, b), p->X, p->g.nz))) return rc;
    if ((rc = zfft_x2s(p, p->X, coff(out, p->nc, b), 1))) return rc;
  }
  return SPX_OK;
}

int spx_fftp2d_real_to_complex_yz(spx_plan* p, void* s) {
  SPX_REQ(p && s, "spx_fftp2d_real_to_complex_yz: null argument");
  MemberView mv(p, 1);
  int rc;
  // y continuation + y FFT of the physical z planes, then z continuation +
  // z FFT: the two continuations and the y FFT act on different axes and
  // commute with each other
  for (int b = 0; b < p->nact; ++b) {
    void* sb = coff(s, p->nc, b);
    if ((rc = zfft_s2x(p, sb, p->X, 0, p->np, 1))) return rc;
    if ((rc = yfft(p, p->X, 1, p->np, nullptr, nullptr, false, p->Cy > 0))) return rc;
    if ((rc = zfft_x2s(p, p->X, sb, 0))) return rc;
  }
  return SPX_OK;
}

int spx_fftp2d_complex_to_real_yz(spx_plan* p, void* s) {
  SPX_REQ(p && s, "spx_fftp2d_complex_to_real_yz: null argument");
  MemberView mv(p, 1);
  int rc;
  for (int b = 0; b < p->nact; ++b) {
    void* sb = coff(s, p->nc, b);
    if ((rc = zfft_s2x(p, sb, p->X, 0, p->g.nz, 0))) return rc;
    if ((rc = yfft(p, p->X, 0, p->g.nz))) return rc;
    if ((rc = zfft_x2s(p, p->X, sb, 1))) return rc;
  }
  return SPX_OK;
}

int spx_fftp1d_complex_to_real_x(spx_plan* p, const void* in, void* out) {
  SPX_REQ(p && in && out, "spx_fftp1d_complex_to_real_x: null argument");
  MemberView mv(p, 1);
  int rc;
  for (int b = 0; b < p->nact; ++b) {
    if ((rc = zfft_s2x(p, coff(in, p->nc, b), p->X, 0, p->g.nz, 1))) return rc;
    if ((rc = xc2r(p, p->X, roff(out, p->nr, b), p->g.nz))) return rc;
  }
  return SPX_OK;
}

int spx_fftp1d_real_to_complex_z(spx_plan* p, void* s) {
  SPX_REQ(p && s, "spx_fftp1d_real_to_complex_z: null argument");
  MemberView mv(p, p->nact);
  return zfft_inplace(p, s, 1, 1, 1.0);
}

int spx_fftp1d_complex_to_real_z(spx_plan* p, void* s) {
  SPX_REQ(p && s, "spx_fftp1d_complex_to_real_z: null argument");
  MemberView mv(p, p->nact);
  return zfft_inplace(p, s, 0, 0, 1.0);
}

int spx_fc_continue(spx_plan* p, void* s) {
  SPX_REQ(p && s, "spx_fc_continue: null argument");
  MemberView mv(p, p->nact);
  return fc_continue_only(p, s);
}

// ---- operators ----------------------------------------------------------------
int spx_derivk(spx_plan* p, const void* a, void* b, int dir) {
  SPX_REQ(p && a && b, "spx_derivk: null argument");
  SPX_REQ(dir >= 1 && dir <= 3, "spx_derivk: dir must be 1, 2 or 3");
  MemberView mv(p, p->nact);
  return op_derivk(p, a, b, dir);
}

int spx_laplak(spx_plan* p, const void* a, void* b) {
  SPX_REQ(p && a && b, "spx_laplak: null argument");
  MemberView mv(p, p->nact);
  return op_laplak(p, a, b);
}

int spx_fc_filter(spx_plan* p, void* a) {
  SPX_REQ(p && a, "spx_fc_filter: null argument");
  MemberView mv(p, p->nact);
  return op_filter(p, a);
}

}  // extern "C"

namespace spx {

// (v.grad) v_c for c = x,y,z in advective form (pseudospec_hd.f90:209-319).
// The x and y derivatives commute with the z transform (kx, ky are constant
// along a z line), so each component needs two z IFFTs (v and d_z v) instead
// of four; the y pass yields v and d_y v from one load; the x pass forms the
// three derivatives and accumulates the products in the reference's order.
// In mixed plans the whole pipeline runs in float32: the first z pass reads
// the float64 state and the outputs are float32.
int gradre_impl(spx_plan* p, const void* const v[3], void* const out[3], bool keep_rv) {
  int rc;
  const int np = p->np;
  const bool lo = p->fp32;
  for (int c = 0; c < 3; ++c)
    if ((rc = zfft_s2x(p, v[c], p->XV[c], 0, np, 0, p->XZ[c], lo))) return rc;
  for (int c = 0; c < 3; ++c) {
    if ((rc = yfft(p, p->XV[c], 0, np, nullptr, p->XY[c], lo))) return rc;
    if ((rc = yfft(p, p->XZ[c], 0, np, nullptr, nullptr, lo))) return rc;
  }
  rc = p->xfused ? xgrad_fused(p, p->XV, p->XY, p->XZ, keep_rv ? p->RV : nullptr, p->RO, lo)
                 : SPX_ERR_UNSUPPORTED;
  if (rc != SPX_OK && rc != SPX_ERR_UNSUPPORTED) return rc;
  if (rc == SPX_ERR_UNSUPPORTED) {
    for (int c = 0; c < 3; ++c)
      if ((rc = xc2r(p, p->XV[c], p->RV[c], np, 0, lo))) return rc;
    for (int c = 0; c < 3; ++c)
      if ((rc = xgrad_prod(p, p->XV[c], p->XY[c], p->XZ[c], p->RV[0], p->RV[1], p->RV[2],
                           p->RO[c], lo)))
        return rc;
  }
  const bool bnd = lo && p->g.Cz > 0;
  const int d2 = 2 * p->g.oz;
  if (bnd) {
    // float64 boundary planes: z IFFT at the 2d planes, y and x passes and
    // products there, forward x, y passes (three components per launch)
    ProfTag tag(p->prof, "_bnd");
    if ((rc = zfft_s2x_bnd(p, v, 3, p->XVb, p->XZb))) return rc;
    if ((rc = yfft(p, p->XVb, 0, 3 * d2, nullptr, p->XYb))) return rc;
    if ((rc = yfft(p, p->XZb, 0, 3 * d2))) return rc;
    if ((rc = xc2r(p, p->XVb, p->RVb, 3 * d2, 0))) return rc;
    if ((rc = xgrad_prod(p, p->XVb, p->XYb, p->XZb, bnd_r(p, p->RVb, 0), bnd_r(p, p->RVb, 1),
                         bnd_r(p, p->RVb, 2), p->ROb, false, d2, 3)))
      return rc;
    if ((rc = xr2c(p, p->ROb, p->XFb, 3 * d2))) return rc;
    if ((rc = yfft(p, p->XFb, 1, 3 * d2))) return rc;
  }
  for (int c = 0; c < 3; ++c)
    if ((rc = r2c3(p, p->RO[c], out[c], lo, bnd ? bnd_c(p, p->XFb, c) : nullptr))) return rc;
  return SPX_OK;
}

// (v.grad) th (pseudospec_phd.f90:23-113); have_rv: RV already holds v in
// real space (computed by gradre_impl on the same v).
static int advect_impl(spx_plan* p, const void* const v[3], const void* th, void* out,
                       bool have_rv) {
  int rc;
  const int np = p->np;
  const bool lo = p->fp32;
  const bool bnd = lo && p->g.Cz > 0;
  const int d2 = 2 * p->g.oz;
  if (!have_rv) {
    for (int c = 0; c < 3; ++c) {
      if ((rc = zfft_s2x(p, v[c], p->XV[c], 0, np, 0, nullptr, lo))) return rc;
      if ((rc = yfft(p, p->XV[c], 0, np, nullptr, nullptr, lo))) return rc;
      if ((rc = xc2r(p, p->XV[c], p->RV[c], np, 0, lo))) return rc;
    }
    if (bnd) {
      ProfTag tag(p->prof, "_bnd");
      if ((rc = zfft_s2x_bnd(p, v, 3, p->XVb, p->XZb))) return rc;
      if ((rc = yfft(p, p->XVb, 0, 3 * d2))) return rc;
      if ((rc = xc2r(p, p->XVb, p->RVb, 3 * d2, 0))) return rc;
    }
  }
  if (bnd) {
    // the scalar's float64 boundary planes (component slot 3)
    ProfTag tag(p->prof, "_bnd");
    const void* ths[1] = {th};
    void* xv3 = bnd_c(p, p->XVb, 3);
    void* xz3 = bnd_c(p, p->XZb, 3);
    void* xy3 = bnd_c(p, p->XYb, 3);
    if ((rc = zfft_s2x_bnd(p, ths, 1, xv3, xz3))) return rc;
    if ((rc = yfft(p, xv3, 0, d2, nullptr, xy3))) return rc;
    if ((rc = yfft(p, xz3, 0, d2))) return rc;
    if ((rc = xgrad_prod(p, xv3, xy3, xz3, bnd_r(p, p->RVb, 0), bnd_r(p, p->RVb, 1),
                         bnd_r(p, p->RVb, 2), bnd_r(p, p->ROb, 3), false, d2)))
      return rc;
    if ((rc = xr2c(p, bnd_r(p, p->ROb, 3), bnd_c(p, p->XFb, 3), d2))) return rc;
    if ((rc = yfft(p, bnd_c(p, p->XFb, 3), 1, d2))) return rc;
  }
  if ((rc = zfft_s2x(p, th, p->XV[3], 0, np, 0, p->XZ[3], lo))) return rc;
  if ((rc = yfft(p, p->XV[3], 0, np, nullptr, p->XY[3], lo))) return rc;
  if ((rc = yfft(p, p->XZ[3], 0, np, nullptr, nullptr, lo))) return rc;
  if ((rc = xgrad_prod(p, p->XV[3], p->XY[3], p->XZ[3], p->RV[0], p->RV[1], p->RV[2],
                       p->RO[3], lo)))
    return rc;
  return r2c3(p, p->RO[3], out, lo, bnd ? bnd_c(p, p->XFb, 3) : nullptr);
}

}  // namespace spx

extern "C" {

int spx_gradre(spx_plan* p, const void* a, const void* b, const void* c, void* d, void* e,
               void* f) {
  SPX_REQ(p && a && b && c && d && e && f, "spx_gradre: null argument");
  int rc;
  if ((rc = ensure_work(p, false))) return rc;
  MemberView mv(p, p->nact);
  const void* v[3] = {a, b, c};
  void* o[3] = {d, e, f};
  return gradre_impl(p, v, o, false);
}

int spx_advect(spx_plan* p, const void* a, const void* b, const void* c, const void* dd,
               void* e) {
  SPX_REQ(p && a && b && c && dd && e, "spx_advect: null argument");
  int rc;
  if ((rc = ensure_work(p, true))) return rc;
  MemberView mv(p, p->nact);
  const void* v[3] = {a, b, c};
  return advect_impl(p, v, dd, e, false);
}

// ---- boundary conditions --------------------------------------------------------
int spx_sol_project(spx_plan* p, void* a, void* b, void* c, void* d, int bctarget, int bczsta,
                    int bczend) {
  SPX_REQ(p && a && b && c && d, "spx_sol_project: null argument");
  if (p->g.Cz <= 0) return fail(SPX_ERR_UNSUPPORTED, "sol_project needs walls in z (Cz > 0)");
  if (p->Cy > 0) return fail(SPX_ERR_UNSUPPORTED, "non-periodic y (Cy > 0) has no boundary conditions (boundary_mod.fpp:104-105)");
  if (bctarget == 0) {
    // conditions on the potential: Dirichlet (0) or Robin (2) walls; laplace_z
    // has no Robin-bottom / Dirichlet-top case (boundary_mod.fpp:622-627)
    const bool ok0 = bczsta == 0 || bczsta == 2, ok1 = bczend == 0 || bczend == 2;
    if (!ok0 || !ok1 || (bczsta == 2 && bczend == 0))
      return fail(SPX_ERR_UNSUPPORTED, "sol_project: unsupported potential boundary kinds");
    MemberView mv(p, p->nact);
    return sol_project_pot(p, a, b, c, d, bczsta, bczend);
  }
  if (bctarget != 1 || bczsta != 0 || bczend != 0)
    return fail(SPX_ERR_UNSUPPORTED,
                "sol_project: bctarget=1 supports homogeneous Dirichlet walls only");
  MemberView mv(p, p->nact);
  int rc = project_harmonic(p, a, b, c, d, nullptr, false);
  if (rc == SPX_OK) return proj_pressure(p, d);
  if (rc != SPX_ERR_UNSUPPORTED) return rc;
  if ((rc = op_poisson_project_bc(p, a, b, c, d))) return rc;
  if ((rc = proj_pressure(p, d))) return rc;
  return proj_harmonic(p, a, b, c);
}

int spx_v_imposebc_and_project(spx_plan* p, void* vx, void* vy, void* vz, void* pr, int rki,
                               int ord, const double* v_zsta, const double* v_zend) {
  SPX_REQ(p && vx && vy && vz && pr, "spx_v_imposebc_and_project: null argument");
  SPX_REQ(rki >= 1 && ord >= 1 && rki <= ord, "spx_v_imposebc_and_project: bad RK stage");
  if (p->g.Cz <= 0) return fail(SPX_ERR_UNSUPPORTED, "no-slip walls need Cz > 0");
  if (p->Cy > 0) return fail(SPX_ERR_UNSUPPORTED, "non-periodic y (Cy > 0) has no boundary conditions (boundary_mod.fpp:104-105)");
  double tau = 1.0 / double(rki);
  if (rki != ord) tau = double(rki + 1) * tau;
  MemberView mv(p, p->nact);
  int rc;
  if ((rc = noslip_fused(p, vx, vy, pr, tau, v_zsta, v_zend))) return rc;
  return spx_sol_project(p, vx, vy, vz, pr, 1, 0, 0);
}

int spx_s_imposebc(spx_plan* p, void* th) {
  SPX_REQ(p && th, "spx_s_imposebc: null argument");
  if (p->g.Cz <= 0) return fail(SPX_ERR_UNSUPPORTED, "scalar walls need Cz > 0");
  if (p->Cy > 0) return fail(SPX_ERR_UNSUPPORTED, "non-periodic y (Cy > 0) has no boundary conditions (boundary_mod.fpp:104-105)");
  MemberView mv(p, p->nact);
  return sconst_fused(p, th);
}

int spx_plan_load_neumann(spx_plan* p, int ord, const double* Q, const double* Qn) {
  SPX_REQ(p && Q && Qn, "spx_plan_load_neumann: null argument");
  SPX_REQ(ord == 1 || ord == 2, "spx_plan_load_neumann: ord must be 1 or 2");
  if (p->g.Cz <= 0) return fail(SPX_ERR_UNSUPPORTED, "Neumann tables need walls in z (Cz > 0)");
  const int d = p->g.oz;
  if (d > 16) return fail(SPX_ERR_UNSUPPORTED, "Neumann tables: matching order above 16");
  // neu = Q(d,:) Qn^T with Qn after the leading dxp (Fortran order), the last
  // entry scaled by (dz/dxp)^ord (fcgram_mod.f90:347-355)
  const double dxp = Qn[0];
  const double* q = Qn + 1;
  double* neu = p->neu[ord - 1];
  for (int j = 0; j < d; ++j) {
    double acc = 0.0;
    for (int k = 0; k < d; ++k) acc = acc + Q[(d - 1) + size_t(d) * k] * q[j + size_t(d) * k];
    neu[j] = acc;
  }
  const double r = p->dz / dxp;
  neu[d - 1] = ord == 1 ? neu[d - 1] * r : neu[d - 1] * (r * r);
  p->neu_loaded[ord - 1] = true;
  return SPX_OK;
}

int spx_curlk(spx_plan* p, const void* a, const void* b, void* c, int dir) {
  SPX_REQ(p && a && b && c, "spx_curlk: null argument");
  SPX_REQ(dir >= 1 && dir <= 3, "spx_curlk: dir must be 1, 2 or 3");
  MemberView mv(p, p->nact);
  return op_curlk(p, a, b, c, dir);
}

int spx_prodre(spx_plan* p, const void* a, const void* b, const void* c, void* d, void* e,
               void* f) {
  SPX_REQ(p && a && b && c && d && e && f, "spx_prodre: null argument");
  int rc;
  if ((rc = ensure_mhd_work(p))) return rc;
  MemberView mv(p, p->nact);
  const void* v[3] = {a, b, c};
  void* w[3] = {d, e, f};
  if ((rc = op_curl3(p, v, w))) return rc;
  const void* wc[3] = {d, e, f};
  return cross_spec(p, wc, v, w, p->MR, p->MR + 3, p->MR + 9, false, false);
}

int spx_vector(spx_plan* p, const void* a, const void* b, const void* c, const void* d,
               const void* e, const void* f, void* g, void* h, void* i) {
  SPX_REQ(p && a && b && c && d && e && f && g && h && i, "spx_vector: null argument");
  int rc;
  if ((rc = ensure_mhd_work(p))) return rc;
  MemberView mv(p, p->nact);
  const void* A[3] = {a, b, c};
  const void* B[3] = {d, e, f};
  void* O[3] = {g, h, i};
  return cross_spec(p, A, B, O, p->MR, p->MR + 3, p->MR + 9, false, false);
}

int spx_a_imposebc_and_project(spx_plan* p, void* ax, void* ay, void* az, void* ph, int bczsta,
                               int bczend) {
  SPX_REQ(p && ax && ay && az && ph, "spx_a_imposebc_and_project: null argument");
  SPX_REQ((bczsta == 0 || bczsta == 1) && (bczend == 0 || bczend == 1),
          "spx_a_imposebc_and_project: wall kinds are 0 (conducting) or 1 (insulating)");
  if (p->g.Cz <= 0) return fail(SPX_ERR_UNSUPPORTED, "magnetic walls need Cz > 0");
  if (p->Cy > 0) return fail(SPX_ERR_UNSUPPORTED, "non-periodic y (Cy > 0) has no boundary conditions (bboundary.f90:132-137)");
  if (bczsta == 1 && bczend == 0)
    return fail(SPX_ERR_UNSUPPORTED,
                "insulating bottom with conducting top: laplace_z has no Robin-Dirichlet case");
  MemberView mv(p, p->nact);
  return a_imposebc(p, ax, ay, az, ph, bczsta, bczend);
}

// ---- diagnostics ------------------------------------------------------------------
// (the diagnostics run in a view of all active members; acc holds one value
// per line of every member)
static int sumsq_lines(spx_plan* p, const LinComb& lc, std::vector<double>& acc, double w) {
  const size_t nl = size_t(p->wlines());
  int rc;
  if ((rc = zline_sumsq(p, lc, p->part))) return rc;
  if ((rc = fetch_parts(p, nl))) return rc;
  for (size_t l = 0; l < nl; ++l) acc[l] += w * p->h_part[l];
  return SPX_OK;
}

static LinComb one(const void* a, int dir) {
  LinComb lc;
  lc.in[0] = a;
  lc.in[1] = lc.in[2] = nullptr;
  lc.dir[0] = dir;
  lc.dir[1] = lc.dir[2] = 0;
  lc.sgn[0] = 1;
  lc.sgn[1] = lc.sgn[2] = 0;
  return lc;
}

static LinComb two(const void* a, int da, const void* b, int db, int sg) {
  LinComb lc = one(a, da);
  lc.in[1] = b;
  lc.dir[1] = db;
  lc.sgn[1] = sg;
  return lc;
}

int spx_energy(spx_plan* p, const void* a, const void* b, const void* c, int kin, double* out) {
  SPX_REQ(p && a && b && c && out, "spx_energy: null argument");
  SPX_REQ(kin >= 0 && kin <= 2, "spx_energy: kin must be 0, 1 or 2");
  int rc;
  if (kin == 2 && !p->MS[0])
    for (int q = 0; q < 6; ++q)
      if ((rc = alloc_dev(p, &p->MS[q], p->nc * p->nb * 16))) return rc;
  MemberView mv(p, p->nact);
  std::vector<double> acc(size_t(p->wlines()), 0.0);
  if (kin == 1) {
    for (const void* q : {a, b, c})
      if ((rc = sumsq_lines(p, one(q, 0), acc, 1.0))) return rc;
  } else if (kin == 2) {
    // curl of curl by the reference's curlk calls (pseudospec_hd.f90:555-560)
    const void* v[3] = {a, b, c};
    const void* w[3] = {p->MS[0], p->MS[1], p->MS[2]};
    if ((rc = op_curl3(p, v, p->MS))) return rc;
    if ((rc = op_curl3(p, w, p->MS + 3))) return rc;
    for (int q = 3; q < 6; ++q)
      if ((rc = sumsq_lines(p, one(p->MS[q], 0), acc, 1.0))) return rc;
  } else {
    // curlk(b,c,1) = d_y c - d_z b ; curlk(a,c,2) = d_z a - d_x c, counted twice
    // (pseudospec_hd.f90:527,540)
    if ((rc = sumsq_lines(p, two(c, 2, b, 3, -1), acc, 1.0))) return rc;
    if ((rc = sumsq_lines(p, two(a, 3, c, 1, -1), acc, 2.0))) return rc;
  }
  for (int b = 0; b < p->nact; ++b) out[b] = weighted_sum(p, acc, b) * diag_tmp(p);
  return SPX_OK;
}

int spx_cross(spx_plan* p, const void* a, const void* b, const void* c, const void* d,
              const void* e, const void* f, double* out) {
  SPX_REQ(p && a && b && c && d && e && f && out, "spx_cross: null argument");
  MemberView mv(p, p->nact);
  const size_t nl = size_t(p->wlines());
  std::vector<double> acc(nl, 0.0);
  const void* L[3] = {a, b, c};
  const void* Rr[3] = {d, e, f};
  int rc;
  for (int t = 0; t < 3; ++t) {
    if ((rc = zline_dot(p, L[t], Rr[t], p->part))) return rc;
    if ((rc = fetch_parts(p, nl))) return rc;
    for (size_t l = 0; l < nl; ++l) acc[l] += p->h_part[l];
  }
  for (int b = 0; b < p->nact; ++b) out[b] = weighted_sum(p, acc, b) * diag_tmp(p);
  return SPX_OK;
}

int spx_helicity(spx_plan* p, const void* a, const void* b, const void* c, double* out) {
  SPX_REQ(p && a && b && c && out, "spx_helicity: null argument");
  MemberView mv(p, p->nact);
  const size_t nl = size_t(p->wlines());
  std::vector<double> acc(nl, 0.0);
  // curlk(b,c,1) = d_y c - d_z b ; curlk(a,c,2) = d_z a - d_x c ; curlk(a,b,3) = d_x b - d_y a
  const LinComb w[3] = {two(c, 2, b, 3, -1), two(a, 3, c, 1, -1), two(b, 1, a, 2, -1)};
  const void* v[3] = {a, b, c};
  int rc;
  for (int t = 0; t < 3; ++t) {
    if ((rc = zline_dot_lc(p, w[t], v[t], p->part))) return rc;
    if ((rc = fetch_parts(p, nl))) return rc;
    for (size_t l = 0; l < nl; ++l) acc[l] += p->h_part[l];
  }
  for (int m = 0; m < p->nact; ++m) out[m] = weighted_sum(p, acc, m) * diag_tmp(p);
  return SPX_OK;
}

int spx_divergence(spx_plan* p, const void* a, const void* b, const void* c, double* out) {
  SPX_REQ(p && a && b && c && out, "spx_divergence: null argument");
  LinComb lc = two(a, 1, b, 2, 1);
  lc.in[2] = c;
  lc.dir[2] = 3;
  lc.sgn[2] = 1;
  MemberView mv(p, p->nact);
  std::vector<double> acc(size_t(p->wlines()), 0.0);
  int rc = sumsq_lines(p, lc, acc, 1.0);
  if (rc) return rc;
  for (int m = 0; m < p->nact; ++m) out[m] = weighted_sum(p, acc, m) * diag_tmp(p);
  return SPX_OK;
}

int spx_bouncheck_z(spx_plan* p, const void* a, const void* b, double* bot, double* top) {
  SPX_REQ(p && a && bot && top, "spx_bouncheck_z: null argument");
  MemberView mv(p, p->nact);
  const size_t nl = size_t(p->wlines());
  std::vector<double> ab(nl, 0.0), at(nl, 0.0);
  int rc;
  for (const void* q : {a, b}) {
    if (!q) continue;
    if ((rc = zline_walls(p, q, p->part))) return rc;
    if ((rc = fetch_parts(p, 2 * nl))) return rc;
    for (size_t l = 0; l < nl; ++l) {
      ab[l] += p->h_part[2 * l];
      at[l] += p->h_part[2 * l + 1];
    }
  }
  const double t = 1.0 / double(long(p->g.nx) * p->g.ny * p->g.nz);
  for (int m = 0; m < p->nact; ++m) {
    bot[m] = weighted_sum(p, ab, m) * (t * t);
    top[m] = weighted_sum(p, at, m) * (t * t);
  }
  return SPX_OK;
}

// ---- scalar and wall diagnostics ---------------------------------------------------
int spx_variance(spx_plan* p, const void* a, int kin, double* out) {
  SPX_REQ(p && a && out, "spx_variance: null argument");
  SPX_REQ(kin == 0 || kin == 1, "spx_variance: kin must be 0 or 1");
  MemberView mv(p, p->nact);
  std::vector<double> acc(size_t(p->wlines()), 0.0);
  int rc = sumsq_lines(p, one(a, kin == 1 ? 0 : 4), acc, 1.0);
  if (rc) return rc;
  for (int m = 0; m < p->nact; ++m) out[m] = weighted_sum(p, acc, m) * diag_tmp(p);
  return SPX_OK;
}

int spx_product(spx_plan* p, const void* a, const void* b, double* out) {
  SPX_REQ(p && a && b && out, "spx_product: null argument");
  MemberView mv(p, p->nact);
  const size_t nl = size_t(p->wlines());
  int rc;
  if ((rc = zline_dot(p, a, b, p->part))) return rc;
  if ((rc = fetch_parts(p, nl))) return rc;
  std::vector<double> acc(p->h_part.begin(), p->h_part.begin() + nl);
  for (int m = 0; m < p->nact; ++m) out[m] = weighted_sum(p, acc, m) * diag_tmp(p);
  return SPX_OK;
}

// horizontal means of d a/dz at z = 0 and z = Lz (physical units), per member
static int plane_mean_dz(spx_plan* p, const void* a, std::vector<double>& bot,
                         std::vector<double>& top) {
  const size_t nl = size_t(p->g.ny) * p->nkx;
  int rc;
  if ((rc = zline_wall_values(p, one(a, 3), p->part))) return rc;
  if ((rc = fetch_parts(p, 4 * nl * p->nact))) return rc;
  const double s = 1.0 / (double(p->g.nx) * double(p->g.ny) * double(p->g.nz));
  bot.assign(p->nact, 0.0);
  top.assign(p->nact, 0.0);
  for (int m = 0; m < p->nact; ++m) {
    const size_t l0 = size_t(m) * nl;  // the (kx, ky) = (0, 0) line of member m
    bot[m] = p->h_part[4 * l0] * s;
    top[m] = p->h_part[4 * l0 + 2] * s;
  }
  return SPX_OK;
}

int spx_wall_shear(spx_plan* p, const void* a, const void* b, double nu, double* out) {
  SPX_REQ(p && a && b && out, "spx_wall_shear: null argument");
  SPX_REQ(p->g.Cz > 0, "spx_wall_shear: needs walls in z");
  MemberView mv(p, p->nact);
  std::vector<double> ab, at, bb, bt;
  int rc;
  if ((rc = plane_mean_dz(p, a, ab, at)) || (rc = plane_mean_dz(p, b, bb, bt))) return rc;
  for (int m = 0; m < p->nact; ++m) {
    out[4 * m] = nu * ab[m];
    out[4 * m + 1] = nu * bb[m];
    out[4 * m + 2] = nu * at[m];
    out[4 * m + 3] = nu * bt[m];
  }
  return SPX_OK;
}

int spx_nusselt(spx_plan* p, const void* th, double* out) {
  SPX_REQ(p && th && out, "spx_nusselt: null argument");
  SPX_REQ(p->g.Cz > 0, "spx_nusselt: needs walls in z");
  MemberView mv(p, p->nact);
  std::vector<double> b, t;
  int rc = plane_mean_dz(p, th, b, t);
  if (rc) return rc;
  for (int m = 0; m < p->nact; ++m) {
    out[2 * m] = 1.0 - p->g.Lz * b[m];
    out[2 * m + 1] = 1.0 - p->g.Lz * t[m];
  }
  return SPX_OK;
}

int spx_divergence_max(spx_plan* p, const void* a, const void* b, const void* c, double* out) {
  SPX_REQ(p && a && b && c && out, "spx_divergence_max: null argument");
  int rc;
  if (!p->DS && (rc = alloc_dev(p, &p->DS, p->nc * p->nb * 16))) return rc;
  if (!p->DR && (rc = alloc_dev(p, &p->DR, p->nr * p->nb * 8))) return rc;
  MemberView mv(p, p->nact);
  if ((rc = op_div_spec(p, a, b, c, p->DS))) return rc;
  // physical planes of each member, members back to back
  if ((rc = c2r3(p, p->DS, p->DR, 0, p->np))) return rc;
  const long per = long(p->g.nx) * p->g.ny * p->np;
  int nblocks = 0;
  if ((rc = op_absmax(p, p->DR, per, per, p->part, &nblocks))) return rc;
  if ((rc = fetch_parts(p, size_t(nblocks) * p->nact))) return rc;
  const double s = 1.0 / (double(p->g.nx) * double(p->g.ny) * double(p->g.nz));
  for (int m = 0; m < p->nact; ++m) {
    double mx = 0.0;
    for (int q = 0; q < nblocks; ++q) mx = std::max(mx, p->h_part[size_t(m) * nblocks + q]);
    out[m] = mx * s;
  }
  return SPX_OK;
}

int spx_energy_spectrum(spx_plan* p, const void* a, const void* b, const void* c, int nbins,
                        double* out) {
  SPX_REQ(p && a && b && c && out && nbins > 0, "spx_energy_spectrum: bad argument");
  MemberView mv(p, p->nact);
  std::vector<double> acc(size_t(p->wlines()), 0.0);
  int rc;
  for (const void* q : {a, b, c})
    if ((rc = sumsq_lines(p, one(q, 0), acc, 1.0))) return rc;
  const long ny = p->g.ny, nl = long(p->g.ny) * p->nkx;
  const double tmp = diag_tmp(p);
  for (int m = 0; m < p->nact; ++m) {
    double* o = out + size_t(m) * nbins;
    std::fill(o, o + nbins, 0.0);
    for (long l = 0; l < nl; ++l) {
      const long i = l / ny, j = l - i * ny;
      const double kh = std::sqrt(p->h_kx[i] * p->h_kx[i] + p->h_ky[j] * p->h_ky[j]) / p->Dkx;
      const long sh = long(std::floor(kh + 0.5));
      if (sh < nbins) o[sh] += (i == 0 ? 1.0 : 2.0) * acc[size_t(m) * nl + l] * tmp;
    }
  }
  return SPX_OK;
}

// ---- solver state ------------------------------------------------------------------
int spx_state_create(spx_plan* p, const spx_params* prm, spx_state** out) {
  SPX_REQ(p && prm && out, "spx_state_create: null argument");
  SPX_REQ(prm->solver == SPX_SOLVER_HD || prm->solver == SPX_SOLVER_BOUSS ||
              prm->solver == SPX_SOLVER_ROTBOUSS || prm->solver == SPX_SOLVER_MHD,
          "spx_state_create: unknown solver");
  SPX_REQ(prm->ord >= 1, "spx_state_create: ord must be >= 1");
  if (p->g.Cz <= 0) return fail(SPX_ERR_UNSUPPORTED, "the solvers need walls in z (Cz > 0)");
  if (p->Cy > 0) return fail(SPX_ERR_UNSUPPORTED, "non-periodic y (Cy > 0) has no boundary conditions (boundary_mod.fpp:104-105)");
  *out = nullptr;
  const bool mhd = prm->solver == SPX_SOLVER_MHD;
  if (mhd) {
    if (p->fp32) return fail(SPX_ERR_UNSUPPORTED, "the MHD solver runs in SPX_FP64 plans");
    SPX_REQ((prm->bzsta == 0 || prm->bzsta == 1) && (prm->bzend == 0 || prm->bzend == 1),
            "spx_state_create: magnetic wall kinds are 0 (conducting) or 1 (insulating)");
    if (prm->bzsta == 1 && prm->bzend == 0)
      return fail(SPX_ERR_UNSUPPORTED,
                  "insulating bottom with conducting top: laplace_z has no Robin-Dirichlet case");
    if (!p->neu_loaded[0] || !p->neu_loaded[1])
      return fail(SPX_ERR_STATE, "the MHD solver needs the Neumann tables (spx_plan_load_neumann)");
  }
  const bool bouss = prm->solver == SPX_SOLVER_BOUSS || prm->solver == SPX_SOLVER_ROTBOUSS;
  int rc = ensure_work(p, bouss);
  if (rc) return rc;
  if (mhd && (rc = ensure_mhd_work(p))) return rc;
  spx_state* s = new spx_state();
  s->p = p;
  s->prm = *prm;
  void** fields[] = {&s->vx, &s->vy, &s->vz, &s->pr, &s->fx, &s->fy, &s->fz,
                     &s->C1, &s->C2, &s->C3, &s->N[0], &s->N[1], &s->N[2]};
  void** bfields[] = {&s->th, &s->fs, &s->C7, &s->N[3]};
  // the nonlinear terms N are in the nonlinear pipeline's precision
  auto mk = [&](void** dst) -> int {
    const bool nl = dst == &s->N[0] || dst == &s->N[1] || dst == &s->N[2] || dst == &s->N[3];
    const size_t bytes = p->nc * size_t(p->nb) * (nl ? p->csize : 16);
    if (hipMalloc(dst, bytes) != hipSuccess) return fail(SPX_ERR_ALLOC, "state allocation");
    s->owned.push_back(*dst);
    if (hipMemsetAsync(*dst, 0, bytes, p->stream) != hipSuccess)
      return fail(SPX_ERR_HIP, "state clear");
    return SPX_OK;
  };
  for (void** f : fields)
    if ((rc = mk(f))) {
      spx_state_destroy(s);
      return rc;
    }
  if (bouss)
    for (void** f : bfields)
      if ((rc = mk(f))) {
        spx_state_destroy(s);
        return rc;
      }
  if (mhd) {
    void** mfields[] = {&s->ax,   &s->ay,   &s->az,   &s->ph,   &s->mx,   &s->my,
                        &s->mz,   &s->C9[0], &s->C9[1], &s->C9[2], &s->B[0], &s->B[1],
                        &s->B[2], &s->L[0], &s->L[1], &s->L[2], &s->M[0], &s->M[1],
                        &s->M[2]};
    for (void** f : mfields)
      if ((rc = mk(f))) {
        spx_state_destroy(s);
        return rc;
      }
  }
  // per-(member, kx) diffusivity tables, every member starting from prm
  s->h_nu.assign(size_t(p->nb) * p->nkx, prm->nu);
  s->h_kappa.assign(size_t(p->nb) * p->nkx, prm->kappa);
  const size_t tb = s->h_nu.size() * sizeof(double);
  for (double** t : {&s->d_nu, &s->d_kappa}) {
    if (hipMalloc(reinterpret_cast<void**>(t), tb) != hipSuccess) {
      *t = nullptr;
      spx_state_destroy(s);
      return fail(SPX_ERR_ALLOC, "state allocation");
    }
    s->owned.push_back(*t);
  }
  if (hipMemcpy(s->d_nu, s->h_nu.data(), tb, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(s->d_kappa, s->h_kappa.data(), tb, hipMemcpyHostToDevice) != hipSuccess) {
    spx_state_destroy(s);
    return fail(SPX_ERR_HIP, "state tables upload");
  }
  if (hipStreamSynchronize(p->stream) != hipSuccess) {
    spx_state_destroy(s);
    return fail(SPX_ERR_HIP, "state clear sync");
  }
  *out = s;
  return SPX_OK;
}

static void drop_graph(spx_state* s) {
  for (auto& kv : s->graphs) {
    if (kv.second.second) (void)hipGraphExecDestroy(kv.second.second);
    if (kv.second.first) (void)hipGraphDestroy(kv.second.first);
  }
  s->graphs.clear();
}

int spx_state_destroy(spx_state* s) {
  if (!s) return SPX_OK;
  if (s->p) (void)hipStreamSynchronize(s->p->stream);
  drop_graph(s);
  if (s->cap) (void)hipStreamDestroy(s->cap);
  for (void* q : s->owned) (void)hipFree(q);
  delete s;
  return SPX_OK;
}

int spx_state_field(spx_state* s, const char* name, void** ptr) {
  SPX_REQ(s && name && ptr, "spx_state_field: null argument");
  struct {
    const char* n;
    void* v;
  } tab[] = {{"vx", s->vx}, {"vy", s->vy}, {"vz", s->vz}, {"pr", s->pr}, {"fx", s->fx},
             {"fy", s->fy}, {"fz", s->fz}, {"th", s->th}, {"fs", s->fs}, {"ax", s->ax},
             {"ay", s->ay}, {"az", s->az}, {"ph", s->ph}, {"mx", s->mx}, {"my", s->my},
             {"mz", s->mz}};
  for (auto& t : tab)
    if (std::strcmp(t.n, name) == 0) {
      if (!t.v) return fail(SPX_ERR_STATE, std::string("field ") + name + " not in this solver");
      *ptr = t.v;
      return SPX_OK;
    }
  return fail(SPX_ERR_ARG, std::string("unknown field ") + name);
}

int spx_rkstep1(spx_plan* p, spx_state* s) {
  SPX_REQ(p && s && s->p == p, "spx_rkstep1: bad plan/state");
  MemberView mv(p, p->nact);
  const void* src[4] = {s->vx, s->vy, s->vz, s->th};
  void* dst[4] = {s->C1, s->C2, s->C3, s->C7};
  if (int rc = op_copy_fields(p, s->th ? 4 : 3, src, dst)) return rc;
  if (!s->ax) return SPX_OK;
  // mhd_rkstep1 also keeps the vector potential (include/mhd/mhd_rkstep1.f90)
  const void* asrc[3] = {s->ax, s->ay, s->az};
  return op_copy_fields(p, 3, asrc, s->C9);
}

// One RK stage of the MHD solver (include/mhd/mhd_rkstep2.f90:3-84).  Real
// work fields: MR[0..2] curl v, then j; MR[3..5] v; MR[6..8] b; MR[9..11] products.
static int mhd_substep(spx_plan* p, spx_state* s, int o) {
  const spx_params& q = s->prm;
  const double rmp = 1.0 / double(o);
  int rc;
  MemberView mv(p, p->nact);
  void* A[3] = {s->ax, s->ay, s->az};
  void* V[3] = {s->vx, s->vy, s->vz};
  void** R = p->MR;
  // b = curl a with the uniform field in the (0,0,0) mode (:6-16), j = curl b in a (:18-21)
  if ((rc = op_curl3(p, A, s->B))) return rc;
  const double b0[3] = {((q.b0[0] * p->g.nx) * p->g.ny) * p->g.nz,
                        ((q.b0[1] * p->g.nx) * p->g.ny) * p->g.nz,
                        ((q.b0[2] * p->g.nx) * p->g.ny) * p->g.nz};
  if ((rc = op_set_mode0(p, s->B[0], s->B[1], s->B[2], b0))) return rc;
  if ((rc = op_curl3(p, s->B, A))) return rc;
  // (curl v) x v (:28), j x b (:31), v x b (:45)
  void* W[3] = {s->N[0], s->N[1], s->N[2]};
  if ((rc = op_curl3(p, V, W))) return rc;
  if ((rc = cross_spec(p, W, V, W, R, R + 3, R + 9, false, false))) return rc;
  if ((rc = cross_spec(p, A, s->B, s->L, R, R + 6, R + 9, false, false))) return rc;
  if ((rc = cross_spec(p, V, s->B, s->M, R + 3, R + 6, R + 9, true, true))) return rc;
  // filters, laplak and the stage update (:33-76)
  const void* c1[3] = {s->C1, s->C2, s->C3};
  const void* f[3] = {s->fx, s->fy, s->fz};
  const void* m[3] = {s->mx, s->my, s->mz};
  if ((rc = op_mhd_update(p, V, A, c1, s->C9, W, s->L, s->M, f, m, s->d_nu, q.mu, q.dt, rmp)))
    return rc;
  // boundary conditions and projections (:80-81)
  if ((rc = spx_v_imposebc_and_project(p, s->vx, s->vy, s->vz, s->pr, o, q.ord, q.v_zsta,
                                       q.v_zend)))
    return rc;
  return a_imposebc(p, s->ax, s->ay, s->az, s->ph, q.bzsta, q.bzend);
}

// One RK stage.  full_pressure = false (inside a full step, for every stage
// but the last): the pressure is formed only at the two walls, which is all
// the next stage's wall update reads; the last stage of the step forms all of
// it, so the state after a step is the reference's.
static int substep_impl(spx_plan* p, spx_state* s, int o, bool full_pressure) {
  SPX_REQ(p && s && s->p == p, "spx_substep: bad plan/state");
  const spx_params& q = s->prm;
  SPX_REQ(o >= 1 && o <= q.ord, "spx_substep: stage out of range");
  if (q.solver == SPX_SOLVER_MHD) return mhd_substep(p, s, o);
  const bool rot = q.solver == SPX_SOLVER_ROTBOUSS;
  const bool bouss = q.solver == SPX_SOLVER_BOUSS || rot;
  const double rmp = 1.0 / double(o);
  int rc;
  MemberView mv(p, p->nact);
  const void* vel[3] = {s->vx, s->vy, s->vz};
  void* nlo[3] = {s->N[0], s->N[1], s->N[2]};
  if ((rc = gradre_impl(p, vel, nlo, bouss))) return rc;
  if (bouss && (rc = advect_impl(p, vel, s->th, s->N[3], true))) return rc;
  // rotating frame: 2 Omega x v joins the nonlinear term before the buoyancy
  // and the filter (include/rotbouss/rotbouss_rkstep2.f90:13-21)
  if (rot && (rc = op_coriolis(p, nlo, vel, q.omega, p->fp32))) return rc;
  void* v[4] = {s->vx, s->vy, s->vz, s->th};
  const void* c0[4] = {s->C1, s->C2, s->C3, s->C7};
  void* nl[4] = {s->N[0], s->N[1], s->N[2], s->N[3]};
  const void* f[4] = {s->fx, s->fy, s->fz, s->fs};
  const void* src[4] = {nullptr, nullptr, s->th, s->vz};
  const double coef[4] = {0.0, 0.0, q.xmom, q.xtemp};
  const double* nu[4] = {s->d_nu, s->d_nu, s->d_nu, s->d_kappa};
  // the RK update of vx, vy is applied as the wall kernel loads them; HD
  // updates vz as the projection loads it
  RkLoad rk;
  std::memset(&rk, 0, sizeof(rk));
  rk.nl32 = p->fp32 ? 1 : 0;
  rk.kx = p->d_kx;
  rk.ky = p->d_ky;
  rk.kz = p->d_kz;
  rk.fkx = p->d_fkx;
  rk.fky = p->d_fky;
  rk.fkz = p->d_fkz;
  rk.dt = q.dt;
  rk.rmp = rmp;
  for (int c = 0; c < 2; ++c) rk.c[c] = RkLoadComp{c0[c], nl[c], f[c], s->d_nu};
  double tau = 1.0 / double(o);
  if (o != q.ord) tau = double(o + 1) * tau;
  if (!bouss) {
    if ((rc = noslip_fused(p, s->vx, s->vy, s->pr, tau, q.v_zsta, q.v_zend, &rk))) return rc;
    RkLoad rz = rk;
    rz.c[0] = RkLoadComp{c0[2], nl[2], f[2], s->d_nu};
    rc = project_harmonic(p, s->vx, s->vy, s->vz, s->pr, &rz, !full_pressure);
    if (rc == SPX_OK)
      return full_pressure ? proj_pressure(p, s->pr) : pressure_walls(p, s->pr);
    if (rc != SPX_ERR_UNSUPPORTED) return rc;
    if ((rc = op_poisson_project_bc(p, s->vx, s->vy, s->vz, s->pr, &rz, !full_pressure)))
      return rc;
    if ((rc = full_pressure ? proj_pressure(p, s->pr) : pressure_walls(p, s->pr))) return rc;
    return proj_harmonic(p, s->vx, s->vy, s->vz);
  }
  // Boussinesq: vz and th are updated together first (the buoyancy and heat
  // current terms couple them through their values before this stage)
  if ((rc = op_rk_update(p, 2, v + 2, c0 + 2, nl + 2, f + 2, src + 2, coef + 2, nu + 2, q.dt, rmp,
                         p->fp32)))
    return rc;
  if ((rc = noslip_fused(p, s->vx, s->vy, s->pr, tau, q.v_zsta, q.v_zend, &rk))) return rc;
  rc = project_harmonic(p, s->vx, s->vy, s->vz, s->pr, nullptr, !full_pressure);
  if (rc == SPX_ERR_UNSUPPORTED) {
    if ((rc = op_poisson_project_bc(p, s->vx, s->vy, s->vz, s->pr, nullptr, !full_pressure)))
      return rc;
    if ((rc = full_pressure ? proj_pressure(p, s->pr) : pressure_walls(p, s->pr))) return rc;
    rc = proj_harmonic(p, s->vx, s->vy, s->vz);
  } else if (rc == SPX_OK) {
    rc = full_pressure ? proj_pressure(p, s->pr) : pressure_walls(p, s->pr);
  }
  if (rc) return rc;
  if ((rc = spx_s_imposebc(p, s->th))) return rc;
  // the rotating solver ends its stage here (rotbouss_rkstep2.f90:49-50)
  if (rot) return SPX_OK;
  if ((rc = op_filter(p, s->th))) return rc;
  // round trip of th through real space (include/bouss/bouss_rkstep2.f90:56-59)
  if ((rc = c2r3(p, s->th, p->R0, 0, p->np))) return rc;
  if ((rc = op_real_div3(p, p->R0, double(p->g.nx), double(p->g.ny), double(p->g.nz))))
    return rc;
  return r2c3(p, p->R0, s->th);
}

int spx_substep(spx_plan* p, spx_state* s, int o) { return substep_impl(p, s, o, true); }

static int step_once(spx_plan* p, spx_state* s) {
  int rc;
  if ((rc = spx_rkstep1(p, s))) return rc;
  for (int o = s->prm.ord; o >= 1; --o)
    if ((rc = substep_impl(p, s, o, o == 1))) return rc;
  return SPX_OK;
}

// Capture step_once on a private stream (a capture needs a stream other than
// the null stream); the graph is then launched on the plan's stream.
static int capture_step(spx_plan* p, spx_state* s) {
  if (!s->cap) SPX_HIP(hipStreamCreateWithFlags(&s->cap, hipStreamNonBlocking));
  hipStream_t orig = p->stream;
  p->stream = s->cap;
  hipError_t e = hipStreamBeginCapture(s->cap, hipStreamCaptureModeRelaxed);
  if (e != hipSuccess) {
    p->stream = orig;
    return fail(SPX_ERR_HIP, std::string("hipStreamBeginCapture: ") + hipGetErrorString(e));
  }
  const int rc = step_once(p, s);
  hipGraph_t g = nullptr;
  e = hipStreamEndCapture(s->cap, &g);
  p->stream = orig;
  if (rc != SPX_OK || e != hipSuccess) {
    if (g) (void)hipGraphDestroy(g);
    return rc != SPX_OK ? rc
                        : fail(SPX_ERR_HIP, std::string("hipStreamEndCapture: ") +
                                                hipGetErrorString(e));
  }
  hipGraphExec_t x = nullptr;
  e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    (void)hipGraphDestroy(g);
    return fail(SPX_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
  }
  s->graphs[p->nact] = std::make_pair(g, x);
  return SPX_OK;
}

int spx_step(spx_plan* p, spx_state* s, int nsteps) {
  SPX_REQ(p && s && s->p == p && nsteps >= 0, "spx_step: bad arguments");
  int rc;
  if (nsteps == 0) return SPX_OK;
  // per-kernel timing needs the individual launches
  if (!s->use_graph || p->prof.on) {
    for (int t = 0; t < nsteps; ++t)
      if ((rc = step_once(p, s))) return rc;
    return SPX_OK;
  }
  if (!s->graphs.count(p->nact) && (rc = capture_step(p, s))) return rc;
  hipGraphExec_t x = s->graphs[p->nact].second;
  for (int t = 0; t < nsteps; ++t) SPX_HIP(hipGraphLaunch(x, p->stream));
  return SPX_OK;
}

int spx_state_set_member(spx_state* s, int member, double nu, double kappa) {
  SPX_REQ(s && s->p, "spx_state_set_member: null state");
  spx_plan* p = s->p;
  SPX_REQ(member >= 0 && member < p->nb, "spx_state_set_member: member out of range");
  const size_t o = size_t(member) * p->nkx;
  for (int i = 0; i < p->nkx; ++i) {
    s->h_nu[o + i] = nu;
    s->h_kappa[o + i] = kappa;
  }
  const size_t tb = size_t(p->nkx) * sizeof(double);
  // the host tables stay alive with the state, so the copies may be asynchronous
  SPX_HIP(hipMemcpyAsync(s->d_nu + o, s->h_nu.data() + o, tb, hipMemcpyHostToDevice, p->stream));
  SPX_HIP(hipMemcpyAsync(s->d_kappa + o, s->h_kappa.data() + o, tb, hipMemcpyHostToDevice,
                         p->stream));
  return SPX_OK;
}

int spx_state_copy_member(spx_state* s, int dst, int src) {
  SPX_REQ(s && s->p, "spx_state_copy_member: null state");
  spx_plan* p = s->p;
  SPX_REQ(dst >= 0 && dst < p->nb && src >= 0 && src < p->nb,
          "spx_state_copy_member: member out of range");
  if (dst == src) return SPX_OK;
  const size_t b = p->nc * 16;
  for (void* f : {s->vx, s->vy, s->vz, s->pr, s->fx, s->fy, s->fz, s->th, s->fs}) {
    if (!f) continue;
    SPX_HIP(hipMemcpyAsync(coff(f, p->nc, dst), coff(static_cast<const void*>(f), p->nc, src), b,
                           hipMemcpyDeviceToDevice, p->stream));
  }
  const size_t o = size_t(p->nkx);
  return spx_state_set_member(s, dst, s->h_nu[src * o], s->h_kappa[src * o]);
}

int spx_state_set_graph(spx_state* s, int on) {
  SPX_REQ(s, "spx_state_set_graph: null state");
  s->use_graph = on != 0;
  if (!s->use_graph) {
    if (s->p) SPX_HIP(hipStreamSynchronize(s->p->stream));
    drop_graph(s);
  }
  return SPX_OK;
}

// ---- per-kernel timing --------------------------------------------------------------
int spx_prof_enable(spx_plan* p, int on) {
  SPX_REQ(p, "spx_prof_enable: null plan");
  SPX_HIP(hipStreamSynchronize(p->stream));
  Profiler& pr = p->prof;
  pr.on = on != 0;
  if (on) {
    pr.recs.clear();
    pr.used = 0;
    pr.names.clear();
  }
  return SPX_OK;
}

int spx_prof_resolve(spx_plan* p, int* nk) {
  SPX_REQ(p && nk, "spx_prof_resolve: null argument");
  SPX_HIP(hipStreamSynchronize(p->stream));
  Profiler& pr = p->prof;
  pr.launches.assign(pr.names.size(), 0);
  pr.total_ms.assign(pr.names.size(), 0.0);
  pr.total_bytes.assign(pr.names.size(), 0.0);
  for (const ProfRec& r : pr.recs) {
    float ms = 0.f;
    SPX_HIP(hipEventElapsedTime(&ms, r.a, r.b));
    pr.launches[r.kind] += 1;
    pr.total_ms[r.kind] += double(ms);
    pr.total_bytes[r.kind] += r.bytes;
  }
  *nk = int(pr.names.size());
  return SPX_OK;
}

int spx_prof_get(spx_plan* p, int idx, char* name, size_t len, long* launches, double* total_ms,
                 double* total_bytes) {
  SPX_REQ(p && name && launches && total_ms && total_bytes, "spx_prof_get: null argument");
  const Profiler& pr = p->prof;
  SPX_REQ(idx >= 0 && size_t(idx) < pr.launches.size(), "spx_prof_get: index out of range");
  std::snprintf(name, len, "%s", pr.names[idx].c_str());
  *launches = pr.launches[idx];
  *total_ms = pr.total_ms[idx];
  *total_bytes = pr.total_bytes[idx];
  return SPX_OK;
}

// ---- memory helpers ----------------------------------------------------------------
int spx_dev_alloc(size_t bytes, void** ptr) {
  SPX_REQ(ptr, "spx_dev_alloc: null argument");
  return alloc_dev(nullptr, ptr, bytes);
}

int spx_dev_free(void* ptr) {
  SPX_HIP(hipFree(ptr));
  return SPX_OK;
}

int spx_memcpy_h2d(spx_plan* p, void* dst, const void* src, size_t bytes) {
  SPX_REQ(p && dst && src, "spx_memcpy_h2d: null argument");
  SPX_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, p->stream));
  SPX_HIP(hipStreamSynchronize(p->stream));
  return SPX_OK;
}

int spx_memcpy_d2d(spx_plan* p, void* dst, const void* src, size_t bytes) {
  SPX_REQ(p && dst && src, "spx_memcpy_d2d: null argument");
  SPX_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, p->stream));
  return SPX_OK;
}

int spx_memcpy_d2h(spx_plan* p, void* dst, const void* src, size_t bytes) {
  SPX_REQ(p && dst && src, "spx_memcpy_d2h: null argument");
  SPX_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, p->stream));
  SPX_HIP(hipStreamSynchronize(p->stream));
  return SPX_OK;
}

}  // extern "C"
