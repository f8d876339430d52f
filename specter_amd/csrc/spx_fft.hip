// Line-FFT kernels of libspx: x (real<->half-complex), y, z (with the FC
// continuation and the (kx,y,z)<->(z,y,kx) transposes fused in), and the
// z-line kernels of the boundary/projection stage.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "spx_engine.h"

namespace spx {

using Tp = double;  // element precision of the FP64 path
using Cp = cpx<Tp>;
using LF = LineFFT<Cp>;

// (0,1)*k*v evaluated as the reference's complex product (-k Im v, k Re v)
__device__ __forceinline__ Cp ikmul(double k, Cp v) { return mkc<Tp>(-(k * v.y), k * v.x); }

// ---------------------------------------------------------------------------
// z transforms in place on S (contiguous lines)
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_z_inplace(Cp* __restrict__ S, long nlines, LF f, int np,
                                                   int forward, int cont, double scale,
                                                   const double* __restrict__ dir, int C,
                                                   int d) {
  extern __shared__ unsigned char smem[];
  Cp* lds = reinterpret_cast<Cp*>(smem);
  const int tid = threadIdx.x, NT = blockDim.x, n = f.n, nl = f.nl, ldl = f.ldl;
  const long line0 = long(blockIdx.x) * nl;
  const int nload = (forward && cont) ? np : n;
  for (int e = tid; e < nl * nload; e += NT) {
    const int l = e / nload, k = e - l * nload;
    const long ln = line0 + l;
    Cp v = mkc<Tp>(0, 0);
    if (ln < nlines) v = S[ln * n + k];
    if (!forward) v.y = -v.y;
    lds[l * ldl + k] = v;
  }
  __syncthreads();
  if (forward && cont) {
    fc_continue_lds(lds, nl, ldl, np, C, d, dir);
    __syncthreads();
  }
  const int line = tid / f.tpl, lane = tid - line * f.tpl;
  line_dft(lds + line * ldl, f, lane);
  for (int e = tid; e < nl * n; e += NT) {
    const int l = e / n, k = e - l * n;
    const long ln = line0 + l;
    if (ln < nlines) {
      Cp v = lds[l * ldl + k];
      if (!forward) {
        v.y = -v.y;
        if (k < np && scale != 1.0) v = cscale(v, scale);
      }
      S[ln * n + k] = v;
    }
  }
}

// Continuation only (fftp.fpp:755-773), in place on S.
__global__ void k_fc_continue(Cp* __restrict__ S, long nlines, int n, int np, int C, int d,
                              const double* __restrict__ dir) {
#pragma clang fp contract(off)
  const long e = long(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= nlines * C) return;
  const long ln = e / C;
  const int ii = int(e - ln * C);
  const Cp* x = S + ln * n;
  const int top = np - d;
  const double* D0 = dir + ii * d;
  const double* D1 = dir + (C - 1 - ii) * d;
  double ar = D0[0] * x[top].x + D1[0] * x[d - 1].x;
  double ai = D0[0] * x[top].y + D1[0] * x[d - 1].y;
  for (int jj = 1; jj < d; ++jj) {
    ar = ar + D0[jj] * x[top + jj].x;
    ar = ar + D1[jj] * x[d - 1 - jj].x;
    ai = ai + D0[jj] * x[top + jj].y;
    ai = ai + D1[jj] * x[d - 1 - jj].y;
  }
  S[ln * n + np + ii] = mkc<Tp>(ar, ai);
}

// ---------------------------------------------------------------------------
// forward 3D last pass: X(kx,y,z) planes < np -> continuation -> z FFT -> S(z,y,kx)
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_z_x2s(const Cp* __restrict__ X, Cp* __restrict__ S, LF f,
                                               int nkx, int ny, int np, int nofft,
                                               const double* __restrict__ dir, int C, int d) {
  extern __shared__ unsigned char smem[];
  Cp* lds = reinterpret_cast<Cp*>(smem);
  const int tid = threadIdx.x, NT = blockDim.x, n = f.n, nl = f.nl, ldl = f.ldl;
  const int i0 = blockIdx.x * nl, j = blockIdx.y;
  const int nload = nofft ? n : np;
  const long pl = long(nkx) * ny;
  for (int e = tid; e < nl * nload; e += NT) {
    const int k = e / nl, l = e - k * nl, i = i0 + l;
    Cp v = mkc<Tp>(0, 0);
    if (i < nkx) v = X[i + long(nkx) * j + pl * k];
    lds[l * ldl + k] = v;
  }
  __syncthreads();
  if (!nofft) {
    if (C > 0) {
      fc_continue_lds(lds, nl, ldl, np, C, d, dir);
      __syncthreads();
    }
    const int line = tid / f.tpl, lane = tid - line * f.tpl;
    line_dft(lds + line * ldl, f, lane);
  }
  for (int e = tid; e < nl * n; e += NT) {
    const int l = e / n, k = e - l * n, i = i0 + l;
    if (i < nkx) S[k + long(n) * (j + long(ny) * i)] = lds[l * ldl + k];
  }
}

// ---------------------------------------------------------------------------
// backward 3D first pass: S * (i k_dir) -> z IFFT -> X planes < nzout
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_z_s2x(const Cp* __restrict__ S, Cp* __restrict__ X, LF f,
                                               int nkx, int ny, int nzout, int dirn, int nofft,
                                               const double* __restrict__ kx,
                                               const double* __restrict__ ky,
                                               const double* __restrict__ kz) {
  extern __shared__ unsigned char smem[];
  Cp* lds = reinterpret_cast<Cp*>(smem);
  const int tid = threadIdx.x, NT = blockDim.x, n = f.n, nl = f.nl, ldl = f.ldl;
  const int i0 = blockIdx.x * nl, j = blockIdx.y;
  for (int e = tid; e < nl * n; e += NT) {
    const int l = e / n, k = e - l * n, i = i0 + l;
    Cp v = mkc<Tp>(0, 0);
    if (i < nkx) {
      v = S[k + long(n) * (j + long(ny) * i)];
      if (dirn == 1) v = ikmul(kx[i], v);
      else if (dirn == 2) v = ikmul(ky[j], v);
      else if (dirn == 3) v = ikmul(kz[k], v);
    }
    v.y = -v.y;
    lds[l * ldl + k] = v;
  }
  __syncthreads();
  if (!nofft) {
    const int line = tid / f.tpl, lane = tid - line * f.tpl;
    line_dft(lds + line * ldl, f, lane);
  }
  const long pl = long(nkx) * ny;
  for (int e = tid; e < nl * nzout; e += NT) {
    const int k = e / nl, l = e - k * nl, i = i0 + l;
    if (i < nkx) {
      Cp v = lds[l * ldl + k];
      v.y = -v.y;
      X[i + long(nkx) * j + pl * k] = v;
    }
  }
}

// ---------------------------------------------------------------------------
// y transform in place on X (lines along y, stride nkx), planes < nplanes
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_y(Cp* __restrict__ X, LF f, int nkx, int ny, int forward) {
  extern __shared__ unsigned char smem[];
  Cp* lds = reinterpret_cast<Cp*>(smem);
  const int tid = threadIdx.x, NT = blockDim.x, nl = f.nl, ldl = f.ldl;
  const int i0 = blockIdx.x * nl;
  const long base = long(blockIdx.y) * nkx * ny;
  for (int e = tid; e < nl * ny; e += NT) {
    const int jy = e / nl, l = e - jy * nl, i = i0 + l;
    Cp v = mkc<Tp>(0, 0);
    if (i < nkx) v = X[base + i + long(nkx) * jy];
    if (!forward) v.y = -v.y;
    lds[l * ldl + jy] = v;
  }
  __syncthreads();
  const int line = tid / f.tpl, lane = tid - line * f.tpl;
  line_dft(lds + line * ldl, f, lane);
  for (int e = tid; e < nl * ny; e += NT) {
    const int jy = e / nl, l = e - jy * nl, i = i0 + l;
    if (i < nkx) {
      Cp v = lds[l * ldl + jy];
      if (!forward) v.y = -v.y;
      X[base + i + long(nkx) * jy] = v;
    }
  }
}

// ---------------------------------------------------------------------------
// x real -> half complex: rows 2p, 2p+1 packed as one complex line
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_xr2c(const Tp* __restrict__ R, Cp* __restrict__ X, LF f,
                                              int nkx, long npairs) {
  extern __shared__ unsigned char smem[];
  Cp* lds = reinterpret_cast<Cp*>(smem);
  const int tid = threadIdx.x, NT = blockDim.x, nx = f.n, nl = f.nl, ldl = f.ldl;
  const long p0 = long(blockIdx.x) * nl;
  for (int e = tid; e < nl * nx; e += NT) {
    const int l = e / nx, x = e - l * nx;
    const long p = p0 + l;
    Cp v = mkc<Tp>(0, 0);
    if (p < npairs) v = mkc<Tp>(R[x + long(nx) * (2 * p)], R[x + long(nx) * (2 * p + 1)]);
    lds[l * ldl + x] = v;
  }
  __syncthreads();
  const int line = tid / f.tpl, lane = tid - line * f.tpl;
  line_dft(lds + line * ldl, f, lane);
  const int two = 2 * nkx;
  for (int e = tid; e < nl * two; e += NT) {
    const int l = e / two, q = e - l * two;
    const long p = p0 + l;
    if (p >= npairs) continue;
    const int w = q >= nkx ? 1 : 0, k = q - w * nkx;
    const Cp Z = lds[l * ldl + k];
    const Cp Zc = lds[l * ldl + (k == 0 ? 0 : nx - k)];
    Cp o;
    if (w == 0) {
      o = mkc<Tp>((Z.x + Zc.x) * 0.5, (Z.y - Zc.y) * 0.5);
    } else {
      const Tp dr = Z.x - Zc.x, di = Z.y + Zc.y;
      o = mkc<Tp>(di * 0.5, -dr * 0.5);
    }
    X[k + long(nkx) * (2 * p + w)] = o;
  }
}

// ---------------------------------------------------------------------------
// x half complex -> real (FFTW c2r semantics: Im of kx=0 and kx=nx/2 ignored)
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_xc2r(const Cp* __restrict__ X, Tp* __restrict__ R, LF f,
                                              int nkx, long npairs) {
  extern __shared__ unsigned char smem[];
  Cp* lds = reinterpret_cast<Cp*>(smem);
  const int tid = threadIdx.x, NT = blockDim.x, nx = f.n, nl = f.nl, ldl = f.ldl;
  const int h = nx / 2;
  const long p0 = long(blockIdx.x) * nl;
  for (int e = tid; e < nl * nx; e += NT) {
    const int l = e / nx, m = e - l * nx;
    const long p = p0 + l;
    Cp v = mkc<Tp>(0, 0);
    if (p < npairs) {
      const bool up = m > h;
      const int k = up ? nx - m : m;
      Cp A = X[k + long(nkx) * (2 * p)];
      Cp B = X[k + long(nkx) * (2 * p + 1)];
      if (k == 0 || k == h) {
        A.y = 0;
        B.y = 0;
      }
      v = up ? mkc<Tp>(A.x + B.y, -A.y + B.x) : mkc<Tp>(A.x - B.y, A.y + B.x);
      v.y = -v.y;
    }
    lds[l * ldl + m] = v;
  }
  __syncthreads();
  const int line = tid / f.tpl, lane = tid - line * f.tpl;
  line_dft(lds + line * ldl, f, lane);
  for (int e = tid; e < nl * nx; e += NT) {
    const int l = e / nx, x = e - l * nx;
    const long p = p0 + l;
    if (p < npairs) {
      const Cp z = lds[l * ldl + x];
      R[x + long(nx) * (2 * p)] = z.x;
      R[x + long(nx) * (2 * p + 1)] = -z.y;
    }
  }
}

// ---------------------------------------------------------------------------
// goto_domain_w_boundaries + noslip_z + goto_3d_fourier fused per line
// (boundary_mod.fpp:72-194, vboundary.f90:154-211, sboundary.f90:122-165).
// comp 0: vx (kx), comp 1: vy (ky); scalar mode: wall values are zero.
// ---------------------------------------------------------------------------
struct WallParams {
  double tau;          // pressure time-extrapolation factor
  double vb[2][2];     // [comp][pos] wall velocity
  double nxny;         // nx*ny
  int scalar;          // 1: s_constant_z (zero at walls)
};

__global__ void __launch_bounds__(256) k_noslip(Cp* __restrict__ vx, Cp* __restrict__ vy,
                                                const Cp* __restrict__ pr, LF f, long nlines,
                                                int ny, int np, double invn, WallParams wp,
                                                const double* __restrict__ kx,
                                                const double* __restrict__ ky,
                                                const double* __restrict__ dir, int C, int d) {
#pragma clang fp contract(off)
  extern __shared__ unsigned char smem[];
  Cp* lds = reinterpret_cast<Cp*>(smem);
  const int tid = threadIdx.x, NT = blockDim.x, n = f.n, nl = f.nl, ldl = f.ldl;
  const int comp = blockIdx.y;
  Cp* V = comp ? vy : vx;
  const long line0 = long(blockIdx.x) * nl;
  for (int e = tid; e < nl * n; e += NT) {
    const int l = e / n, k = e - l * n;
    const long ln = line0 + l;
    Cp v = mkc<Tp>(0, 0);
    if (ln < nlines) v = V[ln * n + k];
    v.y = -v.y;
    lds[l * ldl + k] = v;
  }
  __syncthreads();
  const int line = tid / f.tpl, lane = tid - line * f.tpl;
  line_dft(lds + line * ldl, f, lane);
  for (int e = tid; e < nl * np; e += NT) {
    const int l = e / np, k = e - l * np;
    Cp v = lds[l * ldl + k];
    v.y = -v.y;
    lds[l * ldl + k] = cscale(v, invn);
  }
  __syncthreads();
  for (int e = tid; e < nl * 2; e += NT) {
    const int l = e >> 1, w = e & 1;
    const long ln = line0 + l;
    if (ln >= nlines) continue;
    const int ind = w ? np - 1 : 0;
    Cp v;
    if (wp.scalar) {
      v = mkc<Tp>(0, 0);
    } else {
      const long i = ln / ny;
      const int j = int(ln - i * ny);
      const Cp p = pr[ln * n + ind];
      v = cscale(ikmul(comp ? ky[j] : kx[i], p), wp.tau);
      if (i == 0 && j == 0) v = mkc<Tp>(wp.nxny * wp.vb[comp][w], 0.0);
    }
    lds[l * ldl + ind] = v;
  }
  __syncthreads();
  fc_continue_lds(lds, nl, ldl, np, C, d, dir);
  __syncthreads();
  line_dft(lds + line * ldl, f, lane);
  for (int e = tid; e < nl * n; e += NT) {
    const int l = e / n, k = e - l * n;
    const long ln = line0 + l;
    if (ln < nlines) V[ln * n + k] = lds[l * ldl + k];
  }
}

// ---------------------------------------------------------------------------
// sol_project (boundary_mod.fpp:197-402), bctarget = 1, bczsta = bczend = 0
// ---------------------------------------------------------------------------
// bc(1:2) <- IFFT_z(c * (1/nz)) at z = 0 and z = Lz   (:262-352)
__global__ void __launch_bounds__(256) k_proj_bc(const Cp* __restrict__ c, Cp* __restrict__ bc,
                                                 LF f, long nlines, int np, double tmp) {
  extern __shared__ unsigned char smem[];
  Cp* lds = reinterpret_cast<Cp*>(smem);
  const int tid = threadIdx.x, NT = blockDim.x, n = f.n, nl = f.nl, ldl = f.ldl;
  const long line0 = long(blockIdx.x) * nl;
  for (int e = tid; e < nl * n; e += NT) {
    const int l = e / n, k = e - l * n;
    const long ln = line0 + l;
    Cp v = mkc<Tp>(0, 0);
    if (ln < nlines) v = cscale(c[ln * n + k], tmp);
    v.y = -v.y;
    lds[l * ldl + k] = v;
  }
  __syncthreads();
  const int line = tid / f.tpl, lane = tid - line * f.tpl;
  line_dft(lds + line * ldl, f, lane);
  for (int e = tid; e < nl * 2; e += NT) {
    const int l = e >> 1, w = e & 1;
    const long ln = line0 + l;
    if (ln < nlines) bc[ln * 2 + w] = cconj(lds[l * ldl + (w ? np - 1 : 0)]);
  }
}

// laplace_z pure-Neumann coefficients (boundary_mod.fpp:531-560)
struct Coef {
  Cp c1, c2;
};
__device__ __forceinline__ Coef neumann_coef(Cp b1, Cp b2, double kh, double Lz, bool zero_mode) {
#pragma clang fp contract(off)
  Coef r;
  if (zero_mode) {
    r.c1 = b1;
    r.c2 = mkc<Tp>(0, 0);
    return r;
  }
  const double tmp = 1.0 / (kh * (1 - exp(-2 * kh * Lz)));
  const double e1 = exp(-kh * Lz);
  r.c1 = cscale(csub(b2, cscale(b1, e1)), tmp);
  r.c2 = cscale(cadd(mkc<Tp>(-b1.x, -b1.y), cscale(b2, e1)), tmp);
  return r;
}

// d <- IFFT_z(d)*tmp + a_harm(z)  (:354-381 with laplace_z :631-675)
__global__ void __launch_bounds__(256) k_proj_pressure(Cp* __restrict__ dd,
                                                       const Cp* __restrict__ bc, LF f,
                                                       long nlines, int ny, double tmp, double Lz,
                                                       const double* __restrict__ khom,
                                                       const double* __restrict__ z) {
#pragma clang fp contract(off)
  extern __shared__ unsigned char smem[];
  Cp* lds = reinterpret_cast<Cp*>(smem);
  const int tid = threadIdx.x, NT = blockDim.x, n = f.n, nl = f.nl, ldl = f.ldl;
  const long line0 = long(blockIdx.x) * nl;
  for (int e = tid; e < nl * n; e += NT) {
    const int l = e / n, k = e - l * n;
    const long ln = line0 + l;
    Cp v = mkc<Tp>(0, 0);
    if (ln < nlines) v = dd[ln * n + k];
    v.y = -v.y;
    lds[l * ldl + k] = v;
  }
  __syncthreads();
  const int line = tid / f.tpl, lane = tid - line * f.tpl;
  line_dft(lds + line * ldl, f, lane);
  for (int e = tid; e < nl * n; e += NT) {
    const int l = e / n, k = e - l * n;
    const long ln = line0 + l;
    if (ln >= nlines) continue;
    const long i = ln / ny;
    const int j = int(ln - i * ny);
    const bool zm = (i == 0 && j == 0);
    const double kh = khom[ln];
    const Coef cf = neumann_coef(bc[2 * ln], bc[2 * ln + 1], kh, Lz, zm);
    Cp a;
    if (zm) {
      a = mkc<Tp>(cf.c1.x * z[k] + cf.c2.x, 0.0);
    } else {
      const double ep = exp(kh * (z[k] - Lz));
      const double em = exp(-kh * z[k]);
      a = cadd(cscale(cf.c1, ep), cscale(cf.c2, em));
    }
    Cp v = lds[l * ldl + k];
    v.y = -v.y;
    dd[ln * n + k] = cadd(cscale(v, tmp), a);
  }
}

// harmonic correction: A = FFT_z(cont(a_harm)), B = FFT_z(cont(d a_harm/dz));
// vx -= i kx A, vy -= i ky A, vz -= B   (:383-399)
__global__ void __launch_bounds__(256) k_proj_harmonic(Cp* __restrict__ va, Cp* __restrict__ vb,
                                                       Cp* __restrict__ vc,
                                                       const Cp* __restrict__ bc, LF f,
                                                       long nlines, int ny, int np, double Lz,
                                                       const double* __restrict__ khom,
                                                       const double* __restrict__ z,
                                                       const double* __restrict__ kx,
                                                       const double* __restrict__ ky,
                                                       const double* __restrict__ dir, int C,
                                                       int d) {
#pragma clang fp contract(off)
  extern __shared__ unsigned char smem[];
  Cp* ldsA = reinterpret_cast<Cp*>(smem);
  const int tid = threadIdx.x, NT = blockDim.x, n = f.n, nl = f.nl, ldl = f.ldl;
  Cp* ldsB = ldsA + nl * ldl;
  const long line0 = long(blockIdx.x) * nl;
  for (int e = tid; e < nl * np; e += NT) {
    const int l = e / np, k = e - l * np;
    const long ln = line0 + l;
    Cp a = mkc<Tp>(0, 0), b = mkc<Tp>(0, 0);
    if (ln < nlines) {
      const long i = ln / ny;
      const int j = int(ln - i * ny);
      const bool zm = (i == 0 && j == 0);
      const double kh = khom[ln];
      const Coef cf = neumann_coef(bc[2 * ln], bc[2 * ln + 1], kh, Lz, zm);
      if (zm) {
        a = mkc<Tp>(cf.c1.x * z[k] + cf.c2.x, 0.0);
        b = mkc<Tp>(cf.c1.x, 0.0);
      } else {
        const double ep = exp(kh * (z[k] - Lz));
        const double em = exp(-kh * z[k]);
        const Cp t1 = cscale(cf.c1, ep), t2 = cscale(cf.c2, em);
        a = cadd(t1, t2);
        b = cscale(csub(t1, t2), kh);
      }
    }
    ldsA[l * ldl + k] = a;
    ldsB[l * ldl + k] = b;
  }
  __syncthreads();
  fc_continue_lds(ldsA, nl, ldl, np, C, d, dir);
  fc_continue_lds(ldsB, nl, ldl, np, C, d, dir);
  __syncthreads();
  const int line = tid / f.tpl, lane = tid - line * f.tpl;
  line_dft(ldsA + line * ldl, f, lane);
  line_dft(ldsB + line * ldl, f, lane);
  for (int e = tid; e < nl * n; e += NT) {
    const int l = e / n, k = e - l * n;
    const long ln = line0 + l;
    if (ln >= nlines) continue;
    const long i = ln / ny;
    const int j = int(ln - i * ny);
    const Cp A = ldsA[l * ldl + k], B = ldsB[l * ldl + k];
    const long o = ln * n + k;
    va[o] = csub(va[o], ikmul(kx[i], A));
    vb[o] = csub(vb[o], ikmul(ky[j], A));
    vc[o] = csub(vc[o], B);
  }
}

// ---------------------------------------------------------------------------
// diagnostics: per-line reductions over physical z in the mixed domain
// ---------------------------------------------------------------------------
__device__ __forceinline__ double line_reduce(double v, double* red, int line, int lane, int tpl) {
  red[line * tpl + lane] = v;
  __syncthreads();
  for (int s = tpl >> 1; s > 0; s >>= 1) {
    if (lane < s) red[line * tpl + lane] += red[line * tpl + lane + s];
    __syncthreads();
  }
  return red[line * tpl];
}

struct LinCombDev {
  const Cp* in[3];
  int dir[3];
  int sgn[3];
};

__device__ __forceinline__ Cp deriv_apply(Cp v, int dirn, double kxi, double kyj, double kzk) {
  if (dirn == 1) return ikmul(kxi, v);
  if (dirn == 2) return ikmul(kyj, v);
  if (dirn == 3) return ikmul(kzk, v);
  return v;
}

// mode 0: sum_k<np |IFFT(lincomb)|^2 ; mode 1: Re(IFFT(a) conj IFFT(b)) ; mode 2: walls |.|^2
__global__ void __launch_bounds__(256) k_zline_reduce(LinCombDev lc, const Cp* __restrict__ bsrc,
                                                      int mode, double* __restrict__ out, LF f,
                                                      long nlines, int ny, int np,
                                                      const double* __restrict__ kx,
                                                      const double* __restrict__ ky,
                                                      const double* __restrict__ kz) {
#pragma clang fp contract(off)
  extern __shared__ unsigned char smem[];
  Cp* ldsA = reinterpret_cast<Cp*>(smem);
  const int tid = threadIdx.x, NT = blockDim.x, n = f.n, nl = f.nl, ldl = f.ldl;
  Cp* ldsB = ldsA + nl * ldl;
  double* red = reinterpret_cast<double*>(ldsB + (mode == 1 ? nl * ldl : 0));
  const long line0 = long(blockIdx.x) * nl;
  for (int e = tid; e < nl * n; e += NT) {
    const int l = e / n, k = e - l * n;
    const long ln = line0 + l;
    Cp v = mkc<Tp>(0, 0), w = mkc<Tp>(0, 0);
    if (ln < nlines) {
      const long i = ln / ny;
      const int j = int(ln - i * ny);
      const double kxi = kx[i], kyj = ky[j], kzk = kz[k];
      v = deriv_apply(lc.in[0][ln * n + k], lc.dir[0], kxi, kyj, kzk);
      for (int t = 1; t < 3; ++t) {
        if (lc.sgn[t] == 0) continue;
        const Cp u = deriv_apply(lc.in[t][ln * n + k], lc.dir[t], kxi, kyj, kzk);
        v = lc.sgn[t] > 0 ? cadd(v, u) : csub(v, u);
      }
      if (mode == 1) w = bsrc[ln * n + k];
    }
    v.y = -v.y;
    w.y = -w.y;
    ldsA[l * ldl + k] = v;
    if (mode == 1) ldsB[l * ldl + k] = w;
  }
  __syncthreads();
  const int line = tid / f.tpl, lane = tid - line * f.tpl;
  line_dft(ldsA + line * ldl, f, lane);
  if (mode == 1) line_dft(ldsB + line * ldl, f, lane);
  double acc0 = 0.0, acc1 = 0.0;
  if (mode == 2) {
    if (lane == 0) {
      const Cp b = ldsA[line * ldl], t = ldsA[line * ldl + np - 1];
      acc0 = b.x * b.x + b.y * b.y;
      acc1 = t.x * t.x + t.y * t.y;
    }
  } else {
    for (int k = lane; k < np; k += f.tpl) {
      const Cp v = ldsA[line * ldl + k];
      if (mode == 0) {
        acc0 += v.x * v.x + v.y * v.y;
      } else {
        const Cp w = ldsB[line * ldl + k];
        // values hold conj(IFFT); Re(p conj(q)) is invariant under conj of both
        acc0 += v.x * w.x + v.y * w.y;
      }
    }
  }
  const long ln = line0 + line;
  if (mode == 2) {
    if (lane == 0 && ln < nlines) {
      out[2 * ln] = acc0;
      out[2 * ln + 1] = acc1;
    }
  } else {
    const double s = line_reduce(acc0, red, line, lane, f.tpl);
    if (lane == 0 && ln < nlines) out[ln] = s;
  }
}

// ---------------------------------------------------------------------------
// plan construction: factorisation, Bluestein size, long-double tables
// ---------------------------------------------------------------------------
static bool factor_small(int n, std::vector<int>& fac) {
  fac.clear();
  int a = 0;
  while (n % 2 == 0) {
    n /= 2;
    ++a;
  }
  // powers of two: prefer radix 8, then 4, avoid lone 2 when possible
  int q = a / 3, r = a % 3;
  if (r == 1 && q >= 1) {
    --q;
    fac.push_back(4);
    fac.push_back(4);
  } else if (r == 1) {
    fac.push_back(2);
  } else if (r == 2) {
    fac.push_back(4);
  }
  for (int i = 0; i < q; ++i) fac.push_back(8);
  for (int p : {3, 5, 7, 11, 13}) {
    while (n % p == 0) {
      n /= p;
      fac.push_back(p);
    }
  }
  return n == 1;
}

static bool smooth7(int m) {
  for (int p : {2, 3, 5, 7})
    while (m % p == 0) m /= p;
  return m == 1;
}

static int tpl_for(int M, const std::vector<int>& fac) {
  int need = 1;
  for (int R : fac) {
    const int nb = M / R;
    const int mb = std::max(1, 16 / R);
    need = std::max(need, (nb + mb - 1) / mb);
  }
  int t = 1;
  while (t < need) t <<= 1;
  return t;
}

static int dev_upload(void** dst, const void* src, size_t bytes) {
  if (hipMalloc(dst, bytes) != hipSuccess) return fail(SPX_ERR_ALLOC, "hipMalloc of FFT tables");
  if (hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice) != hipSuccess)
    return fail(SPX_ERR_HIP, "hipMemcpy of FFT tables");
  return SPX_OK;
}

int build_line_fft(LineFFTHost& f, int n, int csize, int want_lines) {
  std::vector<int> fac;
  f.n = n;
  if (factor_small(n, fac)) {
    f.blue = 0;
    f.M = n;
  } else {
    f.blue = 1;
    int M = 2 * n - 1;
    while (!smooth7(M)) ++M;
    f.M = M;
    factor_small(M, fac);
  }
  if (int(fac.size()) > kMaxStages) return fail(SPX_ERR_UNSUPPORTED, "FFT length has too many factors");
  f.nstage = int(fac.size());
  for (int i = 0; i < f.nstage; ++i) f.radix[i] = fac[i];
  f.tpl = tpl_for(f.M, fac);
  if (f.tpl > 256)
    return fail(SPX_ERR_UNSUPPORTED, "FFT length " + std::to_string(n) + " exceeds the LDS engine");
  f.ldl = f.M + ((f.M % 16 == 0) ? 1 : 0);
  const size_t lds_budget = 64 * 1024;
  int nl = std::max(1, 256 / f.tpl);
  if (want_lines > 0) nl = std::min(nl, want_lines);
  while (nl > 1 && size_t(nl) * f.ldl * csize > lds_budget) nl >>= 1;
  if (size_t(nl) * f.ldl * csize > 160 * 1024)
    return fail(SPX_ERR_UNSUPPORTED, "FFT line does not fit in LDS");
  f.nl = nl;
  // tables in long double
  const long double PI2 = 6.283185307179586476925286766559005768L;
  const int M = f.M;
  std::vector<long double> wr(M), wi(M);
  for (int m = 0; m < M; ++m) {
    wr[m] = cosl(PI2 * (long double)m / (long double)M);
    wi[m] = -sinl(PI2 * (long double)m / (long double)M);
  }
  std::vector<double> tw(2 * size_t(M));
  for (int m = 0; m < M; ++m) {
    tw[2 * m] = double(wr[m]);
    tw[2 * m + 1] = double(wi[m]);
  }
  int rc = dev_upload(&f.tw, tw.data(), tw.size() * sizeof(double));
  if (rc) return rc;
  if (f.blue) {
    // chirp w[m] = exp(-pi i (m^2 mod 2n)/n)
    const long double PI = PI2 / 2;
    std::vector<long double> cr(n), ci(n);
    std::vector<double> ch(2 * size_t(n));
    for (int m = 0; m < n; ++m) {
      const long long r = (long long)m * m % (2LL * n);
      cr[m] = cosl(PI * (long double)r / (long double)n);
      ci[m] = -sinl(PI * (long double)r / (long double)n);
      ch[2 * m] = double(cr[m]);
      ch[2 * m + 1] = double(ci[m]);
    }
    // b[m] = conj(w[|m|]) wrapped; B = DFT_M(b) / M
    std::vector<long double> br(M, 0.0L), bi(M, 0.0L);
    br[0] = cr[0];
    bi[0] = -ci[0];
    for (int m = 1; m < n; ++m) {
      br[m] = cr[m];
      bi[m] = -ci[m];
      br[M - m] = cr[m];
      bi[M - m] = -ci[m];
    }
    std::vector<double> bf(2 * size_t(M));
    for (int k = 0; k < M; ++k) {
      long double sr = 0, si = 0;
      for (int m = 0; m < M; ++m) {
        if (br[m] == 0.0L && bi[m] == 0.0L) continue;
        const int idx = int((long long)m * k % M);
        sr += br[m] * wr[idx] - bi[m] * wi[idx];
        si += br[m] * wi[idx] + bi[m] * wr[idx];
      }
      bf[2 * k] = double(sr / M);
      bf[2 * k + 1] = double(si / M);
    }
    if ((rc = dev_upload(&f.chirp, ch.data(), ch.size() * sizeof(double)))) return rc;
    if ((rc = dev_upload(&f.bfil, bf.data(), bf.size() * sizeof(double)))) return rc;
  }
  return SPX_OK;
}

void free_line_fft(LineFFTHost& f) {
  if (f.tw) (void)hipFree(f.tw);
  if (f.chirp) (void)hipFree(f.chirp);
  if (f.bfil) (void)hipFree(f.bfil);
  f.tw = f.chirp = f.bfil = nullptr;
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
static inline dim3 blk(const LineFFTHost& f) { return dim3(unsigned(f.nl * f.tpl)); }

// Raise the dynamic-LDS limit of a kernel when a launch needs more than 64 KiB.
template <typename K>
static int lds_limit(K kernel, size_t bytes) {
  if (bytes <= 65536) return SPX_OK;
  SPX_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kernel),
                              hipFuncAttributeMaxDynamicSharedMemorySize, int(bytes)));
  return SPX_OK;
}

int zfft_inplace(spx_plan* p, void* S, int forward, int cont, double scale) {
  const LineFFTHost& f = p->fz;
  const long nlines = long(p->g.ny) * p->nkx;
  const unsigned nb = unsigned((nlines + f.nl - 1) / f.nl);
  const int C = p->g.Cz, d = p->g.oz;
  ProfScope ps_(p, "z_inplace", double(p->nc) * (16.0 + 16.0));
  hipLaunchKernelGGL(k_z_inplace, dim3(nb), blk(f), f.lds_bytes(p->csize), p->stream,
                     static_cast<Cp*>(S), nlines, f.dev<Cp>(), p->np, forward,
                     (cont && C > 0) ? 1 : 0, scale, p->d_dir, C, d);
  SPX_CHECK_LAUNCH();
  return SPX_OK;
}

int fc_continue_only(spx_plan* p, void* S) {
  const long nlines = long(p->g.ny) * p->nkx;
  const int C = p->g.Cz;
  if (C <= 0) return SPX_OK;
  const long tot = nlines * C;
  ProfScope ps_(p, "fc_continue", double(nlines) * 16.0 * (2 * p->g.oz + C));
  hipLaunchKernelGGL(k_fc_continue, dim3(unsigned((tot + 255) / 256)), dim3(256), 0, p->stream,
                     static_cast<Cp*>(S), nlines, p->g.nz, p->np, C, p->g.oz, p->d_dir);
  SPX_CHECK_LAUNCH();
  return SPX_OK;
}

int zfft_x2s(spx_plan* p, const void* X, void* S, int nofft) {
  const LineFFTHost& f = p->fz;
  dim3 grid(unsigned((p->nkx + f.nl - 1) / f.nl), unsigned(p->g.ny));
  ProfScope ps_(p, "z_x2s", double(p->nkx) * p->g.ny * 16.0 * ((nofft ? p->g.nz : p->np) + p->g.nz));
  hipLaunchKernelGGL(k_z_x2s, grid, blk(f), f.lds_bytes(p->csize), p->stream,
                     static_cast<const Cp*>(X), static_cast<Cp*>(S), f.dev<Cp>(), p->nkx,
                     p->g.ny, p->np, nofft, p->d_dir, p->g.Cz, p->g.oz);
  SPX_CHECK_LAUNCH();
  return SPX_OK;
}

int zfft_s2x(spx_plan* p, const void* S, void* X, int dirn, int nzout, int nofft) {
  const LineFFTHost& f = p->fz;
  dim3 grid(unsigned((p->nkx + f.nl - 1) / f.nl), unsigned(p->g.ny));
  ProfScope ps_(p, "z_s2x", double(p->nkx) * p->g.ny * 16.0 * (p->g.nz + nzout));
  hipLaunchKernelGGL(k_z_s2x, grid, blk(f), f.lds_bytes(p->csize), p->stream,
                     static_cast<const Cp*>(S), static_cast<Cp*>(X), f.dev<Cp>(), p->nkx,
                     p->g.ny, nzout, dirn, nofft, p->d_kx, p->d_ky, p->d_kz);
  SPX_CHECK_LAUNCH();
  return SPX_OK;
}

int yfft(spx_plan* p, void* X, int forward, int nplanes) {
  const LineFFTHost& f = p->fy;
  dim3 grid(unsigned((p->nkx + f.nl - 1) / f.nl), unsigned(nplanes));
  ProfScope ps_(p, "y", double(p->nkx) * p->g.ny * nplanes * 32.0);
  hipLaunchKernelGGL(k_y, grid, blk(f), f.lds_bytes(p->csize), p->stream, static_cast<Cp*>(X),
                     f.dev<Cp>(), p->nkx, p->g.ny, forward);
  SPX_CHECK_LAUNCH();
  return SPX_OK;
}

int xr2c(spx_plan* p, const void* R, void* X, int nplanes) {
  const LineFFTHost& f = p->fx;
  const long npairs = long(p->g.ny) * nplanes / 2;
  dim3 grid(unsigned((npairs + f.nl - 1) / f.nl));
  ProfScope ps_(p, "xr2c", double(p->g.ny) * nplanes * (8.0 * p->g.nx + 16.0 * p->nkx));
  hipLaunchKernelGGL(k_xr2c, grid, blk(f), f.lds_bytes(p->csize), p->stream,
                     static_cast<const Tp*>(R), static_cast<Cp*>(X), f.dev<Cp>(), p->nkx, npairs);
  SPX_CHECK_LAUNCH();
  return SPX_OK;
}

int xc2r(spx_plan* p, const void* X, void* R, int nplanes) {
  const LineFFTHost& f = p->fx;
  const long npairs = long(p->g.ny) * nplanes / 2;
  dim3 grid(unsigned((npairs + f.nl - 1) / f.nl));
  ProfScope ps_(p, "xc2r", double(p->g.ny) * nplanes * (8.0 * p->g.nx + 16.0 * p->nkx));
  hipLaunchKernelGGL(k_xc2r, grid, blk(f), f.lds_bytes(p->csize), p->stream,
                     static_cast<const Cp*>(X), static_cast<Tp*>(R), f.dev<Cp>(), p->nkx, npairs);
  SPX_CHECK_LAUNCH();
  return SPX_OK;
}

int noslip_fused(spx_plan* p, void* vx, void* vy, const void* pr, double tau,
                 const double vzsta[2], const double vzend[2]) {
  const LineFFTHost& f = p->fz;
  const long nlines = long(p->g.ny) * p->nkx;
  WallParams wp;
  wp.tau = tau;
  wp.vb[0][0] = vzsta ? vzsta[0] : 0.0;
  wp.vb[1][0] = vzsta ? vzsta[1] : 0.0;
  wp.vb[0][1] = vzend ? vzend[0] : 0.0;
  wp.vb[1][1] = vzend ? vzend[1] : 0.0;
  wp.nxny = double(p->g.nx * p->g.ny);
  wp.scalar = 0;
  dim3 grid(unsigned((nlines + f.nl - 1) / f.nl), 2);
  ProfScope ps_(p, "noslip", 4.0 * double(p->nc) * 16.0);
  hipLaunchKernelGGL(k_noslip, grid, blk(f), f.lds_bytes(p->csize), p->stream,
                     static_cast<Cp*>(vx), static_cast<Cp*>(vy), static_cast<const Cp*>(pr),
                     f.dev<Cp>(), nlines, p->g.ny, p->np, 1.0 / double(p->g.nz), wp, p->d_kx,
                     p->d_ky, p->d_dir, p->g.Cz, p->g.oz);
  SPX_CHECK_LAUNCH();
  return SPX_OK;
}

int sconst_fused(spx_plan* p, void* th) {
  const LineFFTHost& f = p->fz;
  const long nlines = long(p->g.ny) * p->nkx;
  WallParams wp;
  std::memset(&wp, 0, sizeof(wp));
  wp.scalar = 1;
  dim3 grid(unsigned((nlines + f.nl - 1) / f.nl), 1);
  ProfScope ps_(p, "sconst", 2.0 * double(p->nc) * 16.0);
  hipLaunchKernelGGL(k_noslip, grid, blk(f), f.lds_bytes(p->csize), p->stream,
                     static_cast<Cp*>(th), static_cast<Cp*>(th), static_cast<const Cp*>(nullptr),
                     f.dev<Cp>(), nlines, p->g.ny, p->np, 1.0 / double(p->g.nz), wp, p->d_kx,
                     p->d_ky, p->d_dir, p->g.Cz, p->g.oz);
  SPX_CHECK_LAUNCH();
  return SPX_OK;
}

int proj_bc(spx_plan* p, const void* c) {
  const LineFFTHost& f = p->fz;
  const long nlines = long(p->g.ny) * p->nkx;
  ProfScope ps_(p, "proj_bc", double(p->nc) * 16.0);
  hipLaunchKernelGGL(k_proj_bc, dim3(unsigned((nlines + f.nl - 1) / f.nl)), blk(f),
                     f.lds_bytes(p->csize), p->stream, static_cast<const Cp*>(c),
                     static_cast<Cp*>(p->bc), f.dev<Cp>(), nlines, p->np,
                     1.0 / double(p->g.nz));
  SPX_CHECK_LAUNCH();
  return SPX_OK;
}

int proj_pressure(spx_plan* p, void* d) {
  const LineFFTHost& f = p->fz;
  const long nlines = long(p->g.ny) * p->nkx;
  ProfScope ps_(p, "proj_pressure", 2.0 * double(p->nc) * 16.0);
  hipLaunchKernelGGL(k_proj_pressure, dim3(unsigned((nlines + f.nl - 1) / f.nl)), blk(f),
                     f.lds_bytes(p->csize), p->stream, static_cast<Cp*>(d),
                     static_cast<const Cp*>(p->bc), f.dev<Cp>(), nlines, p->g.ny,
                     1.0 / double(p->g.nz), p->g.Lz, p->d_khom, p->d_z);
  SPX_CHECK_LAUNCH();
  return SPX_OK;
}

int proj_harmonic(spx_plan* p, void* a, void* b, void* c) {
  const LineFFTHost& f = p->fz2;
  const long nlines = long(p->g.ny) * p->nkx;
  if (int rc = lds_limit(k_proj_harmonic, 2 * f.lds_bytes(p->csize))) return rc;
  ProfScope ps_(p, "proj_harmonic", 6.0 * double(p->nc) * 16.0);
  hipLaunchKernelGGL(k_proj_harmonic, dim3(unsigned((nlines + f.nl - 1) / f.nl)), blk(f),
                     2 * f.lds_bytes(p->csize), p->stream, static_cast<Cp*>(a),
                     static_cast<Cp*>(b), static_cast<Cp*>(c), static_cast<const Cp*>(p->bc),
                     f.dev<Cp>(), nlines, p->g.ny, p->np, p->g.Lz, p->d_khom, p->d_z, p->d_kx,
                     p->d_ky, p->d_dir, p->g.Cz, p->g.oz);
  SPX_CHECK_LAUNCH();
  return SPX_OK;
}

static int zline_launch(spx_plan* p, const LinCombDev& lc, const void* b, int mode, double* out) {
  const LineFFTHost& f = mode == 1 ? p->fz2 : p->fz;
  const long nlines = long(p->g.ny) * p->nkx;
  size_t lds = f.lds_bytes(p->csize) * (mode == 1 ? 2 : 1) + size_t(f.nl) * f.tpl * sizeof(double);
  if (int rc = lds_limit(k_zline_reduce, lds)) return rc;
  ProfScope ps_(p, "zline_reduce", double(p->nc) * 16.0 * (mode == 1 ? 2 : 1));
  hipLaunchKernelGGL(k_zline_reduce, dim3(unsigned((nlines + f.nl - 1) / f.nl)), blk(f), lds,
                     p->stream, lc, static_cast<const Cp*>(b), mode, out, f.dev<Cp>(), nlines,
                     p->g.ny, p->np, p->d_kx, p->d_ky, p->d_kz);
  SPX_CHECK_LAUNCH();
  return SPX_OK;
}

int zline_sumsq(spx_plan* p, const LinComb& lc, double* line_out) {
  LinCombDev d;
  for (int t = 0; t < 3; ++t) {
    d.in[t] = static_cast<const Cp*>(lc.in[t]);
    d.dir[t] = lc.dir[t];
    d.sgn[t] = lc.sgn[t];
  }
  return zline_launch(p, d, nullptr, 0, line_out);
}

int zline_dot(spx_plan* p, const void* a, const void* b, double* line_out) {
  LinCombDev d;
  d.in[0] = static_cast<const Cp*>(a);
  d.in[1] = d.in[2] = nullptr;
  d.dir[0] = d.dir[1] = d.dir[2] = 0;
  d.sgn[0] = 1;
  d.sgn[1] = d.sgn[2] = 0;
  return zline_launch(p, d, b, 1, line_out);
}

int zline_walls(spx_plan* p, const void* a, double* line_out) {
  LinCombDev d;
  d.in[0] = static_cast<const Cp*>(a);
  d.in[1] = d.in[2] = nullptr;
  d.dir[0] = d.dir[1] = d.dir[2] = 0;
  d.sgn[0] = 1;
  d.sgn[1] = d.sgn[2] = 0;
  return zline_launch(p, d, nullptr, 2, line_out);
}

}  // namespace spx
