// Pointwise spectral / real-space kernels of libspx.  Floating-point
// contraction is disabled so each kernel performs exactly the reference's
// operation sequence (pseudospec_hd.f90, boundary_mod.fpp, include/*/).
//
// The kernels compute in float64 and are templated on the storage type of
// their fields; the state is float64 in every plan, and only the nonlinear
// terms read by the RK update are float32 in mixed-precision plans.  The
// kernels are bound by HBM traffic, so float64 arithmetic costs nothing.
#include <cmath>
#include <cstring>

#include "spx_engine.h"

namespace spx {

using Cp = cpx<double>;

__device__ __forceinline__ Cp ik(double k, Cp v) { return mkc<double>(-(k * v.y), k * v.x); }

template <typename TS>
__device__ __forceinline__ Cp ld(const cpx<TS>* a, long e) {
  const cpx<TS> v = a[e];
  return mkc<double>(double(v.x), double(v.y));
}
template <typename TS>
__device__ __forceinline__ void st(cpx<TS>* a, long e, Cp v) {
  a[e] = mkc<TS>(TS(v.x), TS(v.y));
}

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
template <typename TS>
__global__ void k_derivk(const cpx<TS>* __restrict__ a, cpx<TS>* __restrict__ b, long nc, SIdx ix,
                         int dirn,
                         const double* __restrict__ kx, const double* __restrict__ ky,
                         const double* __restrict__ kz) {
  const long e = long(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= nc) return;
  int k, j;
  long i;
  ix.split(e, k, j, i);
  const double kk = dirn == 1 ? kx[i] : (dirn == 2 ? ky[j] : kz[k]);
  st(b, e, ik(kk, ld(a, e)));
}

// laplak (pseudospec_hd.f90:97-127): b = -kk2*a, kk2 = (kx^2+ky^2)+kz^2
template <typename TS>
__global__ void k_laplak(const cpx<TS>* __restrict__ a, cpx<TS>* __restrict__ b, long nc, SIdx ix,
                         const double* __restrict__ kx, const double* __restrict__ ky,
                         const double* __restrict__ kz) {
#pragma clang fp contract(off)
  const long e = long(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= nc) return;
  int k, j;
  long i;
  ix.split(e, k, j, i);
  const double kk2 = (kx[i] * kx[i] + ky[j] * ky[j]) + kz[k] * kz[k];
  st(b, e, cscale(ld(a, e), -kk2));
}

// fc_filter (pseudospec_hd.f90:1082-1115): ((a*fx)*fy)*fz
template <typename TS>
__global__ void k_filter(cpx<TS>* __restrict__ a, long nc, SIdx ix, const double* __restrict__ fx,
                         const double* __restrict__ fy, const double* __restrict__ fz) {
#pragma clang fp contract(off)
  const long e = long(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= nc) return;
  int k, j;
  long i;
  ix.split(e, k, j, i);
  st(a, e, cscale(cscale(cscale(ld(a, e), fx[i]), fy[j]), fz[k]));
}

template <typename TR>
__global__ void k_scale(TR* __restrict__ a, long n, double s) {
  const long e = long(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e < n) a[e] = TR(double(a[e]) * s);
}

// real r <- ((r/nx)/ny)/nz  (include/bouss/bouss_rkstep2.f90:57)
template <typename TR>
__global__ void k_div3(TR* __restrict__ r, long n, double a, double b, double c) {
#pragma clang fp contract(off)
  const long e = long(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e < n) r[e] = TR(((double(r[e]) / a) / b) / c);
}

// Fused fc_filter(N) + laplak(v) + RK stage update (include/hd/hd_rkstep2.f90:11-32):
//   v = C0 + dt*(nu*(-kk2 v) - ((N*fx)*fy)*fz + F)*rmp
// with the Boussinesq source N <- N - coef*src applied first
// (include/bouss/bouss_rkstep2.f90:19-27).
template <typename TN>
struct RkComp {
  Cp* v;
  const Cp* c0;
  const cpx<TN>* nl;    // nonlinear term, float64 or float32
  const Cp* f;
  const Cp* src;        // may be null
  double src_coef;
  const double* nu;     // diffusivity per (member, kx) line group
};
template <typename TN>
struct RkArgs {
  RkComp<TN> c[4];
};

template <typename TN>
__global__ void k_rk_update(RkArgs<TN> args, int ncomp, long nc, SIdx ix,
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
    const RkComp<TN>& c = args.c[q];
    Cp nl = ld(c.nl, e);
    if (c.src) nl = csub(nl, cscale(ld(c.src, e), c.src_coef));
    nl = cscale(cscale(cscale(nl, fx[i]), fy[j]), fz[k]);
    const Cp lap = cscale(ld(c.v, e), -kk2);
    const Cp t = cadd(csub(cscale(lap, c.nu[i]), nl), ld(c.f, e));
    out[q] = cadd(ld(c.c0, e), cscale(cscale(t, dt), rmp));
  }
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    if (q >= ncomp) break;
    st(args.c[q].v, e, out[q]);
  }
}

// Coriolis terms 2 Omega x v added to the nonlinear terms in the reference's
// operation order (include/rotbouss/rotbouss_rkstep2.f90:13-19), before the
// buoyancy source and the filter that the RK update applies:
//   Nx += 2 (Oy vz - Oz vy),  Ny += 2 (Oz vx - Ox vz),  Nz += 2 (Ox vy - Oy vx)
template <typename TN>
__global__ void k_coriolis(cpx<TN>* __restrict__ nx, cpx<TN>* __restrict__ ny,
                           cpx<TN>* __restrict__ nz, const Cp* __restrict__ vx,
                           const Cp* __restrict__ vy, const Cp* __restrict__ vz, long nc,
                           double ox, double oy, double oz) {
#pragma clang fp contract(off)
  const long e = long(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= nc) return;
  const Cp a = vx[e], b = vy[e], c = vz[e];
  st(nx, e, cadd(ld(nx, e), cscale(csub(cscale(c, oy), cscale(b, oz)), 2.0)));
  st(ny, e, cadd(ld(ny, e), cscale(csub(cscale(a, oz), cscale(c, ox)), 2.0)));
  st(nz, e, cadd(ld(nz, e), cscale(csub(cscale(b, ox), cscale(a, oy)), 2.0)));
}

// curlk (pseudospec_hd.f90:130-206) for one component: dir 1: i ky b - i kz a,
// dir 2: i kz a - i kx b, dir 3: i kx b - i ky a
// (the products and differences are written out here, where the pragma keeps
// them unfused, so the result is the reference's derivk-then-subtract)
__device__ __forceinline__ Cp curl_dir(int dir, Cp a, Cp b, double kxi, double kyj, double kzk) {
#pragma clang fp contract(off)
  double k1, k2;
  Cp u, w;  // result = (i k1) u - (i k2) w
  if (dir == 1) {
    k1 = kyj, u = b, k2 = kzk, w = a;
  } else if (dir == 2) {
    k1 = kzk, u = a, k2 = kxi, w = b;
  } else {
    k1 = kxi, u = b, k2 = kyj, w = a;
  }
  const double t1x = -(k1 * u.y), t1y = k1 * u.x;
  const double t2x = -(k2 * w.y), t2y = k2 * w.x;
  return mkc<double>(t1x - t2x, t1y - t2y);
}

__global__ void k_curlk(const Cp* __restrict__ a, const Cp* __restrict__ b, Cp* __restrict__ c,
                        long nc, SIdx ix, int dir, const double* __restrict__ kx,
                        const double* __restrict__ ky, const double* __restrict__ kz) {
#pragma clang fp contract(off)
  const long e = long(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= nc) return;
  int k, j;
  long i;
  ix.split(e, k, j, i);
  c[e] = curl_dir(dir, a[e], b[e], kx[i], ky[j], kz[k]);
}

// the three components of curl (X, Y, Z) in one pass:
// (curlk(Y,Z,1), curlk(X,Z,2), curlk(X,Y,3)); outputs distinct from inputs
__global__ void k_curl3(const Cp* __restrict__ X, const Cp* __restrict__ Y,
                        const Cp* __restrict__ Z, Cp* __restrict__ ox, Cp* __restrict__ oy,
                        Cp* __restrict__ oz, long nc, SIdx ix, const double* __restrict__ kx,
                        const double* __restrict__ ky, const double* __restrict__ kz) {
#pragma clang fp contract(off)
  const long e = long(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= nc) return;
  int k, j;
  long i;
  ix.split(e, k, j, i);
  const double kxi = kx[i], kyj = ky[j], kzk = kz[k];
  const Cp x = X[e], y = Y[e], z = Z[e];
  ox[e] = curl_dir(1, y, z, kxi, kyj, kzk);
  oy[e] = curl_dir(2, x, z, kxi, kyj, kzk);
  oz[e] = curl_dir(3, x, y, kxi, kyj, kzk);
}

// A x B on real fields scaled by tmp (pseudospec_hd.f90:386-395): with
// A = (r1, r2, r3), B = (r4, r5, r6):
//   o1 = (r2 r6 - r5 r3) tmp, o2 = (r3 r4 - r6 r1) tmp, o3 = (r1 r5 - r4 r2) tmp
__global__ void k_cross_real(const double* __restrict__ r1, const double* __restrict__ r2,
                             const double* __restrict__ r3, const double* __restrict__ r4,
                             const double* __restrict__ r5, const double* __restrict__ r6,
                             double* __restrict__ o1, double* __restrict__ o2,
                             double* __restrict__ o3, long n, double tmp) {
#pragma clang fp contract(off)
  const long e = long(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const double a1 = r1[e], a2 = r2[e], a3 = r3[e], b1 = r4[e], b2 = r5[e], b3 = r6[e];
  o1[e] = (a2 * b3 - b2 * a3) * tmp;
  o2[e] = (a3 * b1 - b3 * a1) * tmp;
  o3[e] = (a1 * b2 - b1 * a2) * tmp;
}

// element 0 of each member of up to three fields <- v[c] (a real value; the
// (kx, ky, kz) = (0, 0, 0) mode), one thread per (member, field)
__global__ void k_set_mode0(Cp* __restrict__ f0, Cp* __restrict__ f1, Cp* __restrict__ f2,
                            double v0, double v1, double v2, long nc, int nmem) {
  const int t = threadIdx.x;
  if (t >= 3 * nmem) return;
  const int c = t % 3, m = t / 3;
  Cp* f = c == 0 ? f0 : (c == 1 ? f1 : f2);
  if (!f) return;
  f[long(m) * nc] = mkc<double>(c == 0 ? v0 : (c == 1 ? v1 : v2), 0.0);
}

// The MHD stage update (include/mhd/mhd_rkstep2.f90:36-76) per element:
//   v = C1 + dt*(nu*(-kk2 v) - F(N - L) + f)*rmp      (F: fc_filter)
//   a = C9 + dt*(-mu*j + F(M) + m)*rmp                 (a holds j = curl b)
struct MhdArgs {
  Cp* v[3];
  Cp* a[3];
  const Cp* c1[3];
  const Cp* c9[3];
  const Cp* N[3];
  const Cp* L[3];
  const Cp* M[3];
  const Cp* f[3];
  const Cp* m[3];
};
__global__ void k_mhd_update(MhdArgs g, long nc, SIdx ix, const double* __restrict__ kx,
                             const double* __restrict__ ky, const double* __restrict__ kz,
                             const double* __restrict__ fx, const double* __restrict__ fy,
                             const double* __restrict__ fz, const double* __restrict__ nu,
                             double mu, double dt, double rmp) {
#pragma clang fp contract(off)
  const long e = long(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= nc) return;
  int k, j;
  long i;
  ix.split(e, k, j, i);
  const double kk2 = (kx[i] * kx[i] + ky[j] * ky[j]) + kz[k] * kz[k];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    Cp nl = csub(g.N[c][e], g.L[c][e]);
    nl = cscale(cscale(cscale(nl, fx[i]), fy[j]), fz[k]);
    const Cp lap = cscale(g.v[c][e], -kk2);
    const Cp t = cadd(csub(cscale(lap, nu[i]), nl), g.f[c][e]);
    g.v[c][e] = cadd(g.c1[c][e], cscale(cscale(t, dt), rmp));
    const Cp em = cscale(cscale(cscale(g.M[c][e], fx[i]), fy[j]), fz[k]);
    const Cp jm = cscale(g.a[c][e], mu);
    const Cp u = cadd(cadd(mkc<double>(-jm.x, -jm.y), em), g.m[c][e]);
    g.a[c][e] = cadd(g.c9[c][e], cscale(cscale(u, dt), rmp));
  }
}

// sol_project's first pass with the condition on the potential (bctarget 0,
// boundary_mod.fpp:227-336): d = -i k.v/kk2, v -= i k d, and the wall data
// from C1 = d/nz and, for a Robin wall, C2 = i kz C1, at z = 0 and z = Lz as
// DFT sums: kind 0 (Dirichlet) bc = -C1(wall); kind 2 (Robin) bc = C2 - kh C1
// at z = 0 and -(C2 + kh C1) at z = Lz.
__global__ void __launch_bounds__(256) k_poisson_project_pot(
    Cp* __restrict__ a, Cp* __restrict__ b, Cp* __restrict__ c, Cp* __restrict__ d,
    Cp* __restrict__ bc, long nlines, int nz, int ny, const double* __restrict__ kx,
    const double* __restrict__ ky, const double* __restrict__ kz, const Cp* __restrict__ wtop,
    double tmp, int k0, int k1, const double* __restrict__ khom) {
#pragma clang fp contract(off)
  const int lane = threadIdx.x & 63;
  const long ln = long(blockIdx.x) * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (ln >= nlines) return;  // uniform per wavefront
  const long i = ln / ny;
  const int j = int(ln - i * ny);
  const double kxi = kx[i], kyj = ky[j];
  const bool robin = k0 == 2 || k1 == 2;
  Cp s0 = mkc<double>(0, 0), s1 = mkc<double>(0, 0);
  Cp r0 = mkc<double>(0, 0), r1 = mkc<double>(0, 0);
  for (int k = lane; k < nz; k += 64) {
    const long e = ln * nz + k;
    const double kzk = kz[k];
    const double kk2 = (kxi * kxi + kyj * kyj) + kzk * kzk;
    const Cp A = a[e], B = b[e], Cc = c[e];
    Cp sm = cadd(cscale(A, kxi), cscale(B, kyj));
    sm = cadd(sm, cscale(Cc, kzk));
    Cp dd;
    if (kk2 == 0.0) {  // the (0,0,0) mode
      dd = mkc<double>(0, 0);
    } else {
      dd = mkc<double>(sm.y / kk2, -sm.x / kk2);
    }
    d[e] = dd;
    a[e] = csub(A, ik(kxi, dd));
    b[e] = csub(B, ik(kyj, dd));
    c[e] = csub(Cc, ik(kzk, dd));
    const Cp c1 = cscale(dd, tmp);
    s0 = cadd(s0, c1);
    s1 = cadd(s1, cmul(c1, wtop[k]));
    if (robin) {
      const Cp c2 = ik(kzk, c1);
      r0 = cadd(r0, c2);
      r1 = cadd(r1, cmul(c2, wtop[k]));
    }
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    s0.x += __shfl_xor(s0.x, off);
    s0.y += __shfl_xor(s0.y, off);
    s1.x += __shfl_xor(s1.x, off);
    s1.y += __shfl_xor(s1.y, off);
    r0.x += __shfl_xor(r0.x, off);
    r0.y += __shfl_xor(r0.y, off);
    r1.x += __shfl_xor(r1.x, off);
    r1.y += __shfl_xor(r1.y, off);
  }
  if (lane == 0) {
    const double kh = khom[ln];
    bc[2 * ln] = k0 == 0 ? mkc<double>(-s0.x, -s0.y) : csub(r0, cscale(s0, kh));
    const Cp t = cadd(r1, cscale(s1, kh));
    bc[2 * ln + 1] = k1 == 0 ? mkc<double>(-s1.x, -s1.y) : mkc<double>(-t.x, -t.y);
  }
}

// poisson_inhomogeneous + first projection (boundary_mod.fpp:227-240, 405-448):
//   d = -i(kx a + ky b + kz c)/kk2, d(0,0,0) = 0 ; a -= i kx d ; b -= i ky d ; c -= i kz d
// d is float64 for every plan (the pressure potential is kept in float64).
template <typename TS>
__global__ void k_poisson_project(cpx<TS>* __restrict__ a, cpx<TS>* __restrict__ b,
                                  cpx<TS>* __restrict__ c, Cp* __restrict__ d, long nc, SIdx ix,
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
  const Cp A = ld(a, e), B = ld(b, e), Cc = ld(c, e);
  Cp s = cadd(cscale(A, kxi), cscale(B, kyj));
  s = cadd(s, cscale(Cc, kzk));
  Cp dd;
  if (kk2 == 0.0) {  // the (0,0,0) mode
    dd = mkc<double>(0, 0);
  } else {
    dd = mkc<double>(s.y / kk2, -s.x / kk2);
  }
  d[e] = dd;
  st(a, e, csub(A, ik(kxi, dd)));
  st(b, e, csub(B, ik(kyj, dd)));
  st(c, e, csub(Cc, ik(kzk, dd)));
}

// The same per element, one wavefront per (kx, ky) line, plus the wall data
// of sol_project (boundary_mod.fpp:262-352): bc = IFFT_z(c/nz) at z = 0 and
// z = Lz, evaluated as the two DFT sums sum_k c_k/nz and
// sum_k c_k/nz exp(+2 pi i k (np-1)/nz) instead of a full inverse transform.
template <typename TS>
__global__ void __launch_bounds__(256) k_poisson_project_bc(
    cpx<TS>* __restrict__ a, cpx<TS>* __restrict__ b, cpx<TS>* __restrict__ c,
    Cp* __restrict__ d, Cp* __restrict__ bc, long nlines, int nz, int ny,
    const double* __restrict__ kx, const double* __restrict__ ky, const double* __restrict__ kz,
    const Cp* __restrict__ wtop, double tmp, RkLoad rk, int rk_on, Cp* __restrict__ pw,
    int np) {
#pragma clang fp contract(off)
  const int lane = threadIdx.x & 63;
  const long ln = long(blockIdx.x) * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (ln >= nlines) return;  // uniform per wavefront
  const long i = ln / ny;
  const int j = int(ln - i * ny);
  const double kxi = kx[i], kyj = ky[j];
  Cp s0 = mkc<double>(0, 0), s1 = mkc<double>(0, 0);
  Cp q0 = mkc<double>(0, 0), q1 = mkc<double>(0, 0);  // IFFT_z(d) at the walls (pw)
  for (int k = lane; k < nz; k += 64) {
    const long e = ln * nz + k;
    const double kzk = kz[k];
    const double kk2 = (kxi * kxi + kyj * kyj) + kzk * kzk;
    const Cp A = ld(a, e), B = ld(b, e);
    Cp Cc = ld(c, e);
    if (rk_on) Cc = rk_on_load(rk, 0, Cc, e, i, j, k);
    Cp sm = cadd(cscale(A, kxi), cscale(B, kyj));
    sm = cadd(sm, cscale(Cc, kzk));
    Cp dd;
    if (kk2 == 0.0) {  // the (0,0,0) mode
      dd = mkc<double>(0, 0);
    } else {
      dd = mkc<double>(sm.y / kk2, -sm.x / kk2);
    }
    if (pw) {
      q0 = cadd(q0, dd);
      q1 = cadd(q1, cmul(dd, wtop[k]));
      if (k != 0 && k != np - 1) d[e] = dd;
    } else {
      d[e] = dd;
    }
    st(a, e, csub(A, ik(kxi, dd)));
    st(b, e, csub(B, ik(kyj, dd)));
    st(c, e, csub(Cc, ik(kzk, dd)));
    const Cp c1 = cscale(ld(c, e), tmp);
    s0 = cadd(s0, c1);
    s1 = cadd(s1, cmul(c1, wtop[k]));
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    s0.x += __shfl_xor(s0.x, off);
    s0.y += __shfl_xor(s0.y, off);
    s1.x += __shfl_xor(s1.x, off);
    s1.y += __shfl_xor(s1.y, off);
  }
  if (lane == 0) {
    bc[2 * ln] = s0;
    bc[2 * ln + 1] = s1;
  }
  if (pw) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      q0.x += __shfl_xor(q0.x, off);
      q0.y += __shfl_xor(q0.y, off);
      q1.x += __shfl_xor(q1.x, off);
      q1.y += __shfl_xor(q1.y, off);
    }
    if (lane == 0) {
      pw[2 * ln] = cscale(q0, tmp);
      pw[2 * ln + 1] = cscale(q1, tmp);
    }
  }
}

template <typename TS>
__global__ void k_axpy(cpx<TS>* __restrict__ y, const cpx<TS>* __restrict__ x, long nc,
                       double alpha) {
#pragma clang fp contract(off)
  const long e = long(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e < nc) st(y, e, csub(ld(y, e), cscale(ld(x, e), alpha)));
}

// dst[q] <- src[q] for q < n (the RK start copies; a kernel rather than
// memcpy nodes so that a captured step holds kernels only)
struct CopyArgs {
  const Cp* src[4];
  Cp* dst[4];
};
__global__ void k_copy_fields(CopyArgs args, int n, long nc) {
  const long e = long(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= nc) return;
#pragma unroll
  for (int q = 0; q < 4; ++q)
    if (q < n) args.dst[q][e] = args.src[q][e];
}

// d = ((i kx a) + (i ky b)) + (i kz c)  (divergence, pseudospec_hd.f90:1118-1235 order)
template <typename TS>
__global__ void k_div_spec(const cpx<TS>* __restrict__ a, const cpx<TS>* __restrict__ b,
                           const cpx<TS>* __restrict__ c, Cp* __restrict__ d, long nc, SIdx ix,
                           const double* __restrict__ kx, const double* __restrict__ ky,
                           const double* __restrict__ kz) {
#pragma clang fp contract(off)
  const long e = long(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= nc) return;
  int k, j;
  long i;
  ix.split(e, k, j, i);
  Cp s = cadd(ik(kx[i], ld(a, e)), ik(ky[j], ld(b, e)));
  d[e] = cadd(s, ik(kz[k], ld(c, e)));
}

// block maxima of |r| over [0, n) of member blockIdx.y (members ms apart)
__global__ void __launch_bounds__(256) k_absmax(const double* __restrict__ r, long n, long ms,
                                                double* __restrict__ out) {
  __shared__ double red[256];
  const double* x = r + blockIdx.y * ms;
  double m = 0.0;
  for (long e = long(blockIdx.x) * blockDim.x + threadIdx.x; e < n;
       e += long(gridDim.x) * blockDim.x)
    m = fmax(m, fabs(x[e]));
  red[threadIdx.x] = m;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (int(threadIdx.x) < s) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) out[long(blockIdx.y) * gridDim.x + blockIdx.x] = red[0];
}

// ---------------------------------------------------------------------------
static inline unsigned nblk(long n) { return unsigned((n + 255) / 256); }
static inline SIdx sidx(spx_plan* p) { return SIdx{p->g.nz, p->g.ny}; }

template <typename TS>
static inline cpx<TS>* CS(void* a) { return static_cast<cpx<TS>*>(a); }
template <typename TS>
static inline const cpx<TS>* CS(const void* a) { return static_cast<const cpx<TS>*>(a); }

int op_derivk(spx_plan* p, const void* a, void* b, int dirn) {
  ProfScope ps_(p, "derivk", 2.0 * double(p->wnc()) * 16.0);
  hipLaunchKernelGGL(k_derivk<double>, dim3(nblk(long(p->wnc()))), dim3(256), 0, p->stream, CS<double>(a),
                     CS<double>(b), long(p->wnc()), sidx(p), dirn, p->d_kx, p->d_ky, p->d_kz);
  SPX_CHECK_LAUNCH();
  return SPX_OK;
}

int op_laplak(spx_plan* p, const void* a, void* b) {
  ProfScope ps_(p, "laplak", 2.0 * double(p->wnc()) * 16.0);
  hipLaunchKernelGGL(k_laplak<double>, dim3(nblk(long(p->wnc()))), dim3(256), 0, p->stream, CS<double>(a),
                     CS<double>(b), long(p->wnc()), sidx(p), p->d_kx, p->d_ky, p->d_kz);
  SPX_CHECK_LAUNCH();
  return SPX_OK;
}

int op_filter(spx_plan* p, void* a) {
  ProfScope ps_(p, "filter", 2.0 * double(p->wnc()) * 16.0);
  hipLaunchKernelGGL(k_filter<double>, dim3(nblk(long(p->wnc()))), dim3(256), 0, p->stream, CS<double>(a),
                     long(p->wnc()), sidx(p), p->d_fkx, p->d_fky, p->d_fkz);
  SPX_CHECK_LAUNCH();
  return SPX_OK;
}

int op_scale(spx_plan* p, void* a, size_t n, double s) {
  ProfScope ps_(p, "scale", 16.0 * double(n));
  hipLaunchKernelGGL(k_scale<double>, dim3(nblk(long(n))), dim3(256), 0, p->stream,
                     static_cast<double*>(a), long(n), s);
  SPX_CHECK_LAUNCH();
  return SPX_OK;
}

int op_real_div3(spx_plan* p, void* r, double a, double b, double c) {
  const long n = long(p->g.nx) * p->g.ny * p->np * p->mem;
  ProfScope ps_(p, "real_div3", 16.0 * double(n));
  hipLaunchKernelGGL(k_div3<double>, dim3(nblk(n)), dim3(256), 0, p->stream,
                     static_cast<double*>(r), n, a, b, c);
  SPX_CHECK_LAUNCH();
  return SPX_OK;
}

template <typename TN>
static int rk_launch(spx_plan* p, int ncomp, void* const* v, const void* const* c0,
                     void* const* nl, const void* const* f, const void* const* src,
                     const double* src_coef, const double* const* nu, double dt, double rmp) {
  RkArgs<TN> args;
  for (int c = 0; c < 4; ++c) {
    RkComp<TN>& r = args.c[c];
    if (c < ncomp) {
      r.v = CS<double>(v[c]);
      r.c0 = CS<double>(c0[c]);
      r.nl = CS<TN>(static_cast<const void*>(nl[c]));
      r.f = CS<double>(f[c]);
      r.src = src ? CS<double>(src[c]) : nullptr;
      r.src_coef = src_coef ? src_coef[c] : 0.0;
      r.nu = nu[c];
    } else {
      r = RkComp<TN>{nullptr, nullptr, nullptr, nullptr, nullptr, 0.0, nullptr};
    }
  }
  ProfScope ps_(p, "rk_update",
                double(p->wnc()) * ncomp * (4.0 * 16.0 + sizeof(cpx<TN>) + (src ? 16.0 : 0.0)));
  hipLaunchKernelGGL(k_rk_update<TN>, dim3(nblk(long(p->wnc()))), dim3(256), 0, p->stream, args, ncomp,
                     long(p->wnc()), sidx(p), p->d_kx, p->d_ky, p->d_kz, p->d_fkx, p->d_fky,
                     p->d_fkz, dt, rmp);
  SPX_CHECK_LAUNCH();
  return SPX_OK;
}

int op_rk_update(spx_plan* p, int ncomp, void* const* v, const void* const* c0, void* const* nl,
                 const void* const* f, const void* const* src, const double* src_coef,
                 const double* const* nu, double dt, double rmp, bool nl_lo) {
  return nl_lo ? rk_launch<float>(p, ncomp, v, c0, nl, f, src, src_coef, nu, dt, rmp)
               : rk_launch<double>(p, ncomp, v, c0, nl, f, src, src_coef, nu, dt, rmp);
}

int op_coriolis(spx_plan* p, void* const* nl, const void* const* v, const double omega[3],
                bool nl_lo) {
  const long nc = long(p->wnc());
  ProfScope ps_(p, "coriolis", double(nc) * 3.0 * (16.0 + 2.0 * (nl_lo ? 8.0 : 16.0)));
  if (nl_lo)
    hipLaunchKernelGGL(k_coriolis<float>, dim3(nblk(nc)), dim3(256), 0, p->stream,
                       CS<float>(nl[0]), CS<float>(nl[1]), CS<float>(nl[2]), CS<double>(v[0]),
                       CS<double>(v[1]), CS<double>(v[2]), nc, omega[0], omega[1], omega[2]);
  else
    hipLaunchKernelGGL(k_coriolis<double>, dim3(nblk(nc)), dim3(256), 0, p->stream,
                       CS<double>(nl[0]), CS<double>(nl[1]), CS<double>(nl[2]),
                       CS<double>(v[0]), CS<double>(v[1]), CS<double>(v[2]), nc, omega[0],
                       omega[1], omega[2]);
  SPX_CHECK_LAUNCH();
  return SPX_OK;
}

int op_poisson_project(spx_plan* p, void* a, void* b, void* c, void* d) {
  ProfScope ps_(p, "poisson_project", 7.0 * double(p->wnc()) * 16.0);
  hipLaunchKernelGGL(k_poisson_project<double>, dim3(nblk(long(p->wnc()))), dim3(256), 0, p->stream,
                     CS<double>(a), CS<double>(b), CS<double>(c), CS<double>(d), long(p->wnc()),
                     sidx(p), p->d_kx, p->d_ky, p->d_kz);
  SPX_CHECK_LAUNCH();
  return SPX_OK;
}

int op_poisson_project_bc(spx_plan* p, void* a, void* b, void* c, void* d, const RkLoad* rk,
                          bool walls_only) {
  if (!p->d_wtop) return fail(SPX_ERR_STATE, "plan has no wall table");
  const long nlines = p->wlines();
  RkLoad r;
  std::memset(&r, 0, sizeof(r));
  if (rk) r = *rk;
  const double extra = rk ? (rk->nl32 ? 8.0 : 16.0) + 32.0 : 0.0;
  ProfScope ps_(p, rk ? "poisson_project_bc_rk" : "poisson_project_bc",
                double(p->wnc()) * (7.0 * 16.0 + extra));
  hipLaunchKernelGGL(k_poisson_project_bc<double>, dim3(unsigned((nlines + 3) / 4)), dim3(256), 0,
                     p->stream, CS<double>(a), CS<double>(b), CS<double>(c), CS<double>(d),
                     static_cast<Cp*>(p->bc), nlines, p->g.nz, p->g.ny, p->d_kx, p->d_ky,
                     p->d_kz, static_cast<const Cp*>(p->d_wtop), 1.0 / double(p->g.nz), r,
                     rk ? 1 : 0, walls_only ? static_cast<Cp*>(p->pw) : nullptr, p->np);
  SPX_CHECK_LAUNCH();
  return SPX_OK;
}

int op_curlk(spx_plan* p, const void* a, const void* b, void* c, int dir) {
  ProfScope ps_(p, "curlk", 3.0 * double(p->wnc()) * 16.0);
  hipLaunchKernelGGL(k_curlk, dim3(nblk(long(p->wnc()))), dim3(256), 0, p->stream, CS<double>(a),
                     CS<double>(b), CS<double>(c), long(p->wnc()), sidx(p), dir, p->d_kx,
                     p->d_ky, p->d_kz);
  SPX_CHECK_LAUNCH();
  return SPX_OK;
}

int op_curl3(spx_plan* p, const void* const* in, void* const* out) {
  ProfScope ps_(p, "curl3", 6.0 * double(p->wnc()) * 16.0);
  hipLaunchKernelGGL(k_curl3, dim3(nblk(long(p->wnc()))), dim3(256), 0, p->stream,
                     CS<double>(in[0]), CS<double>(in[1]), CS<double>(in[2]), CS<double>(out[0]),
                     CS<double>(out[1]), CS<double>(out[2]), long(p->wnc()), sidx(p), p->d_kx,
                     p->d_ky, p->d_kz);
  SPX_CHECK_LAUNCH();
  return SPX_OK;
}

int op_cross_real(spx_plan* p, const void* const* A, const void* const* B, void* const* O,
                  long n) {
  const double tmp = 1.0 / std::pow(double(p->g.nx) * double(p->g.ny) * double(p->g.nz), 2);
  ProfScope ps_(p, "cross_real", 9.0 * double(n) * 8.0);
  hipLaunchKernelGGL(k_cross_real, dim3(nblk(n)), dim3(256), 0, p->stream,
                     static_cast<const double*>(A[0]), static_cast<const double*>(A[1]),
                     static_cast<const double*>(A[2]), static_cast<const double*>(B[0]),
                     static_cast<const double*>(B[1]), static_cast<const double*>(B[2]),
                     static_cast<double*>(O[0]), static_cast<double*>(O[1]),
                     static_cast<double*>(O[2]), n, tmp);
  SPX_CHECK_LAUNCH();
  return SPX_OK;
}

int op_set_mode0(spx_plan* p, void* f0, void* f1, void* f2, const double v[3]) {
  if (3 * p->mem > 1024) return fail(SPX_ERR_UNSUPPORTED, "set_mode0: too many members");
  hipLaunchKernelGGL(k_set_mode0, dim3(1), dim3(unsigned(((3 * p->mem + 63) / 64) * 64)), 0,
                     p->stream, CS<double>(f0), CS<double>(f1), CS<double>(f2), v[0], v[1], v[2],
                     long(p->nc), p->mem);
  SPX_CHECK_LAUNCH();
  return SPX_OK;
}

int op_mhd_update(spx_plan* p, void* const* v, void* const* a, const void* const* c1,
                  const void* const* c9, const void* const* N, const void* const* L,
                  const void* const* M, const void* const* f, const void* const* m,
                  const double* nu, double mu, double dt, double rmp) {
  MhdArgs g;
  for (int c = 0; c < 3; ++c) {
    g.v[c] = CS<double>(v[c]);
    g.a[c] = CS<double>(a[c]);
    g.c1[c] = CS<double>(c1[c]);
    g.c9[c] = CS<double>(c9[c]);
    g.N[c] = CS<double>(N[c]);
    g.L[c] = CS<double>(L[c]);
    g.M[c] = CS<double>(M[c]);
    g.f[c] = CS<double>(f[c]);
    g.m[c] = CS<double>(m[c]);
  }
  ProfScope ps_(p, "mhd_update", 3.0 * double(p->wnc()) * 16.0 * 11.0);
  hipLaunchKernelGGL(k_mhd_update, dim3(nblk(long(p->wnc()))), dim3(256), 0, p->stream, g,
                     long(p->wnc()), sidx(p), p->d_kx, p->d_ky, p->d_kz, p->d_fkx, p->d_fky,
                     p->d_fkz, nu, mu, dt, rmp);
  SPX_CHECK_LAUNCH();
  return SPX_OK;
}

int op_poisson_project_pot(spx_plan* p, void* a, void* b, void* c, void* d, int k0, int k1) {
  if (!p->d_wtop) return fail(SPX_ERR_STATE, "plan has no wall table");
  const long nlines = p->wlines();
  ProfScope ps_(p, "poisson_project_pot", double(p->wnc()) * 7.0 * 16.0);
  hipLaunchKernelGGL(k_poisson_project_pot, dim3(unsigned((nlines + 3) / 4)), dim3(256), 0,
                     p->stream, CS<double>(a), CS<double>(b), CS<double>(c), CS<double>(d),
                     static_cast<Cp*>(p->bc), nlines, p->g.nz, p->g.ny, p->d_kx, p->d_ky,
                     p->d_kz, static_cast<const Cp*>(p->d_wtop), 1.0 / double(p->g.nz), k0, k1,
                     p->d_khom);
  SPX_CHECK_LAUNCH();
  return SPX_OK;
}

int op_copy_fields(spx_plan* p, int n, const void* const* src, void* const* dst) {
  CopyArgs args;
  for (int q = 0; q < 4; ++q) {
    args.src[q] = q < n ? CS<double>(src[q]) : nullptr;
    args.dst[q] = q < n ? CS<double>(dst[q]) : nullptr;
  }
  ProfScope ps_(p, "copy", 2.0 * n * double(p->wnc()) * 16.0);
  hipLaunchKernelGGL(k_copy_fields, dim3(nblk(long(p->wnc()))), dim3(256), 0, p->stream, args, n,
                     long(p->wnc()));
  SPX_CHECK_LAUNCH();
  return SPX_OK;
}

int op_div_spec(spx_plan* p, const void* a, const void* b, const void* c, void* d) {
  ProfScope ps_(p, "div_spec", 4.0 * double(p->wnc()) * 16.0);
  hipLaunchKernelGGL(k_div_spec<double>, dim3(nblk(long(p->wnc()))), dim3(256), 0, p->stream,
                     CS<double>(a), CS<double>(b), CS<double>(c), CS<double>(d),
                     long(p->wnc()), sidx(p), p->d_kx, p->d_ky, p->d_kz);
  SPX_CHECK_LAUNCH();
  return SPX_OK;
}

int op_absmax(spx_plan* p, const void* r, long n, long ms, double* out, int* nblocks) {
  const int nb = 64;  // blocks per member: out holds mem * 64 values
  ProfScope ps_(p, "absmax", 8.0 * double(n) * p->mem);
  hipLaunchKernelGGL(k_absmax, dim3(nb, unsigned(p->mem)), dim3(256), 0, p->stream,
                     static_cast<const double*>(r), n, ms, out);
  SPX_CHECK_LAUNCH();
  *nblocks = nb;
  return SPX_OK;
}

int op_axpy_field(spx_plan* p, void* y, const void* x, double alpha) {
  ProfScope ps_(p, "axpy", 3.0 * double(p->wnc()) * 16.0);
  hipLaunchKernelGGL(k_axpy<double>, dim3(nblk(long(p->wnc()))), dim3(256), 0, p->stream, CS<double>(y),
                     CS<double>(x), long(p->wnc()), alpha);
  SPX_CHECK_LAUNCH();
  return SPX_OK;
}

}  // namespace spx
