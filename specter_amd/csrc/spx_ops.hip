// Pointwise spectral / real-space kernels of libspx.  Floating-point
// contraction is disabled so each kernel performs exactly the reference's
// operation sequence (pseudospec_hd.f90, boundary_mod.fpp, include/*/).
#include "spx_engine.h"

namespace spx {

using Tp = double;
using Cp = cpx<Tp>;

__device__ __forceinline__ Cp ik(double k, Cp v) { return mkc<Tp>(-(k * v.y), k * v.x); }

// index helpers for S (nz, ny, nkx) with z fastest
struct SIdx {
  int nz, ny;
  __device__ __forceinline__ void split(long e, int& k, int& j, long& i) const {
    const long line = e / nz;
    k = int(e - line * nz);
    i = line / ny;
    j = int(line - i * ny);
  }
};

// derivk (pseudospec_hd.f90:28-94)
__global__ void k_derivk(const Cp* __restrict__ a, Cp* __restrict__ b, long nc, SIdx ix, int dirn,
                         const double* __restrict__ kx, const double* __restrict__ ky,
                         const double* __restrict__ kz) {
  const long e = long(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= nc) return;
  int k, j;
  long i;
  ix.split(e, k, j, i);
  const double kk = dirn == 1 ? kx[i] : (dirn == 2 ? ky[j] : kz[k]);
  b[e] = ik(kk, a[e]);
}

// laplak (pseudospec_hd.f90:97-127): b = -kk2*a, kk2 = (kx^2+ky^2)+kz^2
__global__ void k_laplak(const Cp* __restrict__ a, Cp* __restrict__ b, long nc, SIdx ix,
                         const double* __restrict__ kx, const double* __restrict__ ky,
                         const double* __restrict__ kz) {
#pragma clang fp contract(off)
  const long e = long(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= nc) return;
  int k, j;
  long i;
  ix.split(e, k, j, i);
  const double kk2 = (kx[i] * kx[i] + ky[j] * ky[j]) + kz[k] * kz[k];
  b[e] = cscale(a[e], -kk2);
}

// fc_filter (pseudospec_hd.f90:1082-1115): ((a*fx)*fy)*fz
__global__ void k_filter(Cp* __restrict__ a, long nc, SIdx ix, const double* __restrict__ fx,
                         const double* __restrict__ fy, const double* __restrict__ fz) {
#pragma clang fp contract(off)
  const long e = long(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= nc) return;
  int k, j;
  long i;
  ix.split(e, k, j, i);
  a[e] = cscale(cscale(cscale(a[e], fx[i]), fy[j]), fz[k]);
}

__global__ void k_scale(double* __restrict__ a, long n, double s) {
  const long e = long(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e < n) a[e] *= s;
}

// real r <- ((r/nx)/ny)/nz  (include/bouss/bouss_rkstep2.f90:57)
__global__ void k_div3(double* __restrict__ r, long n, double a, double b, double c) {
#pragma clang fp contract(off)
  const long e = long(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e < n) r[e] = ((r[e] / a) / b) / c;
}

// Fused fc_filter(N) + laplak(v) + RK stage update (include/hd/hd_rkstep2.f90:11-32):
//   v = C0 + dt*(nu*(-kk2 v) - ((N*fx)*fy)*fz + F)*rmp
// with the Boussinesq source N <- N - coef*src applied first
// (include/bouss/bouss_rkstep2.f90:19-27).
struct RkComp {
  Cp* v;
  const Cp* c0;
  const Cp* nl;
  const Cp* f;
  const Cp* src;   // may be null
  double src_coef;
  double nu;
};
struct RkArgs {
  RkComp c[4];
};

__global__ void k_rk_update(RkArgs args, int ncomp, long nc, SIdx ix,
                            const double* __restrict__ kx, const double* __restrict__ ky,
                            const double* __restrict__ kz, const double* __restrict__ fx,
                            const double* __restrict__ fy, const double* __restrict__ fz,
                            double dt, double rmp) {
#pragma clang fp contract(off)
  const long e = long(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= nc) return;
  int k, j;
  long i;
  ix.split(e, k, j, i);
  const double kk2 = (kx[i] * kx[i] + ky[j] * ky[j]) + kz[k] * kz[k];
  // every component is read before any is written: the Boussinesq sources
  // couple vz and th
  Cp out[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    if (q >= ncomp) break;
    const RkComp& c = args.c[q];
    Cp nl = c.nl[e];
    if (c.src) nl = csub(nl, cscale(c.src[e], c.src_coef));
    nl = cscale(cscale(cscale(nl, fx[i]), fy[j]), fz[k]);
    const Cp lap = cscale(c.v[e], -kk2);
    const Cp t = cadd(csub(cscale(lap, c.nu), nl), c.f[e]);
    out[q] = cadd(c.c0[e], cscale(cscale(t, dt), rmp));
  }
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    if (q >= ncomp) break;
    args.c[q].v[e] = out[q];
  }
}

// poisson_inhomogeneous + first projection (boundary_mod.fpp:227-240, 405-448):
//   d = -i(kx a + ky b + kz c)/kk2, d(0,0,0) = 0 ; a -= i kx d ; b -= i ky d ; c -= i kz d
__global__ void k_poisson_project(Cp* __restrict__ a, Cp* __restrict__ b, Cp* __restrict__ c,
                                  Cp* __restrict__ d, long nc, SIdx ix,
                                  const double* __restrict__ kx, const double* __restrict__ ky,
                                  const double* __restrict__ kz) {
#pragma clang fp contract(off)
  const long e = long(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= nc) return;
  int k, j;
  long i;
  ix.split(e, k, j, i);
  const double kxi = kx[i], kyj = ky[j], kzk = kz[k];
  const double kk2 = (kxi * kxi + kyj * kyj) + kzk * kzk;
  Cp A = a[e], B = b[e], Cc = c[e];
  Cp s = cadd(cscale(A, kxi), cscale(B, kyj));
  s = cadd(s, cscale(Cc, kzk));
  Cp dd;
  if (e == 0) {
    dd = mkc<Tp>(0, 0);
  } else {
    dd = mkc<Tp>(s.y / kk2, -s.x / kk2);
  }
  d[e] = dd;
  a[e] = csub(A, ik(kxi, dd));
  b[e] = csub(B, ik(kyj, dd));
  c[e] = csub(Cc, ik(kzk, dd));
}

__global__ void k_axpy(Cp* __restrict__ y, const Cp* __restrict__ x, long nc, double alpha) {
#pragma clang fp contract(off)
  const long e = long(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e < nc) y[e] = csub(y[e], cscale(x[e], alpha));
}

// ---------------------------------------------------------------------------
static inline unsigned nblk(long n) { return unsigned((n + 255) / 256); }
static inline SIdx sidx(spx_plan* p) { return SIdx{p->g.nz, p->g.ny}; }

int op_derivk(spx_plan* p, const void* a, void* b, int dirn) {
  ProfScope ps_(p, "derivk", 2.0 * double(p->nc) * 16.0);
  hipLaunchKernelGGL(k_derivk, dim3(nblk(p->nc)), dim3(256), 0, p->stream,
                     static_cast<const Cp*>(a), static_cast<Cp*>(b), long(p->nc), sidx(p), dirn,
                     p->d_kx, p->d_ky, p->d_kz);
  SPX_CHECK_LAUNCH();
  return SPX_OK;
}

int op_laplak(spx_plan* p, const void* a, void* b) {
  ProfScope ps_(p, "laplak", 2.0 * double(p->nc) * 16.0);
  hipLaunchKernelGGL(k_laplak, dim3(nblk(p->nc)), dim3(256), 0, p->stream,
                     static_cast<const Cp*>(a), static_cast<Cp*>(b), long(p->nc), sidx(p),
                     p->d_kx, p->d_ky, p->d_kz);
  SPX_CHECK_LAUNCH();
  return SPX_OK;
}

int op_filter(spx_plan* p, void* a) {
  ProfScope ps_(p, "filter", 2.0 * double(p->nc) * 16.0);
  hipLaunchKernelGGL(k_filter, dim3(nblk(p->nc)), dim3(256), 0, p->stream, static_cast<Cp*>(a),
                     long(p->nc), sidx(p), p->d_fkx, p->d_fky, p->d_fkz);
  SPX_CHECK_LAUNCH();
  return SPX_OK;
}

int op_scale(spx_plan* p, void* a, size_t n, double s) {
  ProfScope ps_(p, "scale", 16.0 * double(n));
  hipLaunchKernelGGL(k_scale, dim3(nblk(long(n))), dim3(256), 0, p->stream,
                     static_cast<double*>(a), long(n), s);
  SPX_CHECK_LAUNCH();
  return SPX_OK;
}

int op_real_div3(spx_plan* p, void* r, double a, double b, double c) {
  const long n = long(p->g.nx) * p->g.ny * p->np;
  ProfScope ps_(p, "real_div3", 16.0 * double(n));
  hipLaunchKernelGGL(k_div3, dim3(nblk(n)), dim3(256), 0, p->stream, static_cast<double*>(r), n,
                     a, b, c);
  SPX_CHECK_LAUNCH();
  return SPX_OK;
}

int op_rk_update(spx_plan* p, int ncomp, void* const* v, const void* const* c0, void* const* nl,
                 const void* const* f, const void* const* src, const double* src_coef,
                 const double* nu, double dt, double rmp) {
  RkArgs args;
  for (int c = 0; c < 4; ++c) {
    RkComp& r = args.c[c];
    if (c < ncomp) {
      r.v = static_cast<Cp*>(v[c]);
      r.c0 = static_cast<const Cp*>(c0[c]);
      r.nl = static_cast<const Cp*>(nl[c]);
      r.f = static_cast<const Cp*>(f[c]);
      r.src = src ? static_cast<const Cp*>(src[c]) : nullptr;
      r.src_coef = src_coef ? src_coef[c] : 0.0;
      r.nu = nu[c];
    } else {
      r = RkComp{nullptr, nullptr, nullptr, nullptr, nullptr, 0.0, 0.0};
    }
  }
  ProfScope ps_(p, "rk_update", double(p->nc) * 16.0 * 5 * ncomp + (src ? double(p->nc) * 16.0 * ncomp : 0.0));
  hipLaunchKernelGGL(k_rk_update, dim3(nblk(p->nc)), dim3(256), 0, p->stream, args, ncomp,
                     long(p->nc), sidx(p), p->d_kx, p->d_ky, p->d_kz, p->d_fkx, p->d_fky,
                     p->d_fkz, dt, rmp);
  SPX_CHECK_LAUNCH();
  return SPX_OK;
}

int op_poisson_project(spx_plan* p, void* a, void* b, void* c, void* d) {
  ProfScope ps_(p, "poisson_project", 7.0 * double(p->nc) * 16.0);
  hipLaunchKernelGGL(k_poisson_project, dim3(nblk(p->nc)), dim3(256), 0, p->stream,
                     static_cast<Cp*>(a), static_cast<Cp*>(b), static_cast<Cp*>(c),
                     static_cast<Cp*>(d), long(p->nc), sidx(p), p->d_kx, p->d_ky, p->d_kz);
  SPX_CHECK_LAUNCH();
  return SPX_OK;
}

int op_axpy_field(spx_plan* p, void* y, const void* x, double alpha) {
  ProfScope ps_(p, "axpy", 3.0 * double(p->nc) * 16.0);
  hipLaunchKernelGGL(k_axpy, dim3(nblk(p->nc)), dim3(256), 0, p->stream, static_cast<Cp*>(y),
                     static_cast<const Cp*>(x), long(p->nc), alpha);
  SPX_CHECK_LAUNCH();
  return SPX_OK;
}

}  // namespace spx
