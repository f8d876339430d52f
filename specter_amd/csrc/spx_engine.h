// In-LDS batched line DFT engine for gfx950 (CDNA4, wave64).
//
// A workgroup owns nl lines of one length; tpl threads work on each line.  A
// mixed-radix Stockham autosort FFT (radices 2,3,4,5,7,8,11,13) runs in LDS
// with register staging: every pass reads R strided points per butterfly,
// applies the W_M twiddles of a single M-entry table, does the radix-R DFT in
// registers, synchronises, and writes the autosorted result back.  Lengths
// with a prime factor above 13 use Bluestein's chirp-z algorithm with a smooth
// M >= 2n-1, still entirely in LDS (one HBM read and one write per line).
// Backward transforms use conj(DFT(conj(x))).
#pragma once

#include "spx_common.h"

namespace spx {

template <typename T>
struct alignas(2 * sizeof(T)) cpx {
  T x, y;
};

template <typename T>
__device__ __forceinline__ cpx<T> mkc(T re, T im) {
  cpx<T> r;
  r.x = re;
  r.y = im;
  return r;
}
template <typename T>
__device__ __forceinline__ cpx<T> cadd(cpx<T> a, cpx<T> b) { return mkc<T>(a.x + b.x, a.y + b.y); }
template <typename T>
__device__ __forceinline__ cpx<T> csub(cpx<T> a, cpx<T> b) { return mkc<T>(a.x - b.x, a.y - b.y); }
template <typename T>
__device__ __forceinline__ cpx<T> cmul(cpx<T> a, cpx<T> b) {
  return mkc<T>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
template <typename T>
__device__ __forceinline__ cpx<T> cconj(cpx<T> a) { return mkc<T>(a.x, -a.y); }
// (a)(-i)
template <typename T>
__device__ __forceinline__ cpx<T> cmul_mi(cpx<T> a) { return mkc<T>(a.y, -a.x); }
template <typename T>
__device__ __forceinline__ cpx<T> cscale(cpx<T> a, T s) { return mkc<T>(a.x * s, a.y * s); }

// Radix-R DFT in registers, forward sign.  tw is the W_M table, M/R its stride.
template <typename T, int R>
struct Dft {
  static __device__ __forceinline__ void run(cpx<T>* v, const cpx<T>* __restrict__ tw, int s) {
    // odd R: pair r with R-r
    constexpr int H = (R - 1) / 2;
    cpx<T> a[H + 1], b[H + 1];
#pragma unroll
    for (int r = 1; r <= H; ++r) {
      a[r] = cadd(v[r], v[R - r]);
      b[r] = csub(v[r], v[R - r]);
    }
    cpx<T> y0 = v[0];
#pragma unroll
    for (int r = 1; r <= H; ++r) y0 = cadd(y0, a[r]);
    cpx<T> out[R];
    out[0] = y0;
#pragma unroll
    for (int q = 1; q <= H; ++q) {
      cpx<T> re = v[0];
      cpx<T> im = mkc<T>(0, 0);
#pragma unroll
      for (int r = 1; r <= H; ++r) {
        const cpx<T> w = tw[((r * q) % R) * s];  // cos - i sin
        re = cadd(re, cscale(a[r], w.x));
        im = cadd(im, cscale(b[r], -w.y));
      }
      out[q] = cadd(re, cmul_mi(im));
      out[R - q] = csub(re, cmul_mi(im));
    }
#pragma unroll
    for (int q = 0; q < R; ++q) v[q] = out[q];
  }
};

template <typename T>
struct Dft<T, 2> {
  static __device__ __forceinline__ void run(cpx<T>* v, const cpx<T>*, int) {
    cpx<T> a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
  }
};

template <typename T>
struct Dft<T, 4> {
  static __device__ __forceinline__ void run(cpx<T>* v, const cpx<T>*, int) {
    cpx<T> s02 = cadd(v[0], v[2]), d02 = csub(v[0], v[2]);
    cpx<T> s13 = cadd(v[1], v[3]), d13 = cmul_mi(csub(v[1], v[3]));
    v[0] = cadd(s02, s13);
    v[2] = csub(s02, s13);
    v[1] = cadd(d02, d13);
    v[3] = csub(d02, d13);
  }
};

template <typename T>
struct Dft<T, 8> {
  static __device__ __forceinline__ void run(cpx<T>* v, const cpx<T>*, int) {
    const T h = T(0.70710678118654752440084436210484903928L);
    cpx<T> e[4] = {v[0], v[2], v[4], v[6]};
    cpx<T> o[4] = {v[1], v[3], v[5], v[7]};
    Dft<T, 4>::run(e, nullptr, 0);
    Dft<T, 4>::run(o, nullptr, 0);
    // o1 *= exp(-i pi/4), o2 *= -i, o3 *= exp(-3 i pi/4)
    cpx<T> o1 = mkc<T>(h * (o[1].x + o[1].y), h * (o[1].y - o[1].x));
    cpx<T> o2 = cmul_mi(o[2]);
    cpx<T> o3 = mkc<T>(h * (o[3].y - o[3].x), -h * (o[3].x + o[3].y));
    v[0] = cadd(e[0], o[0]);
    v[4] = csub(e[0], o[0]);
    v[1] = cadd(e[1], o1);
    v[5] = csub(e[1], o1);
    v[2] = cadd(e[2], o2);
    v[6] = csub(e[2], o2);
    v[3] = cadd(e[3], o3);
    v[7] = csub(e[3], o3);
  }
};

// Butterflies a thread may stage in registers per pass (<= 16 points).
template <int R>
struct MaxB {
  static constexpr int value = (16 / R) > 0 ? (16 / R) : 1;
};

template <typename T, int R>
__device__ __forceinline__ void stockham_pass(cpx<T>* x, int M, int Ns, int lane, int tpl,
                                              const cpx<T>* __restrict__ tw) {
  constexpr int MB = MaxB<R>::value;
  const int nb = M / R;
  const int tstep = M / (Ns * R);
  const int s = M / R;
  cpx<T> v[MB][R];
#pragma unroll
  for (int b = 0; b < MB; ++b) {
    const int j = lane + b * tpl;
    if (j < nb) {
#pragma unroll
      for (int r = 0; r < R; ++r) v[b][r] = x[j + r * nb];
      if (Ns > 1) {
        const int jm = j % Ns;
        const int st = jm * tstep;
#pragma unroll
        for (int r = 1; r < R; ++r) v[b][r] = cmul(v[b][r], tw[r * st]);
      }
      Dft<T, R>::run(v[b], tw, s);
    }
  }
  __syncthreads();
#pragma unroll
  for (int b = 0; b < MB; ++b) {
    const int j = lane + b * tpl;
    if (j < nb) {
      const int jm = j % Ns;
      const int base = (j - jm) * R + jm;
#pragma unroll
      for (int r = 0; r < R; ++r) x[base + r * Ns] = v[b][r];
    }
  }
  __syncthreads();
}

// Forward FFT of length f.M on the line x (LDS).  Every thread of the
// workgroup must call it (it synchronises); x may differ per thread group.
template <typename T>
__device__ void fft_lds(cpx<T>* x, const LineFFT<cpx<T>>& f, int lane) {
  int Ns = 1;
  for (int st = 0; st < f.nstage; ++st) {
    const int R = f.radix[st];
    switch (R) {
      case 2: stockham_pass<T, 2>(x, f.M, Ns, lane, f.tpl, f.tw); break;
      case 3: stockham_pass<T, 3>(x, f.M, Ns, lane, f.tpl, f.tw); break;
      case 4: stockham_pass<T, 4>(x, f.M, Ns, lane, f.tpl, f.tw); break;
      case 5: stockham_pass<T, 5>(x, f.M, Ns, lane, f.tpl, f.tw); break;
      case 7: stockham_pass<T, 7>(x, f.M, Ns, lane, f.tpl, f.tw); break;
      case 8: stockham_pass<T, 8>(x, f.M, Ns, lane, f.tpl, f.tw); break;
      case 11: stockham_pass<T, 11>(x, f.M, Ns, lane, f.tpl, f.tw); break;
      case 13: stockham_pass<T, 13>(x, f.M, Ns, lane, f.tpl, f.tw); break;
      default: break;
    }
    Ns *= R;
  }
}

// Forward DFT of length f.n of the n values at x[0..n), result in x[0..n).
// The caller has synchronised after filling x.
template <typename T>
__device__ void line_dft(cpx<T>* x, const LineFFT<cpx<T>>& f, int lane) {
  if (!f.blue) {
    fft_lds(x, f, lane);
    return;
  }
  const int n = f.n, M = f.M, tpl = f.tpl;
  for (int m = lane; m < M; m += tpl)
    x[m] = (m < n) ? cmul(x[m], f.chirp[m]) : mkc<T>(0, 0);
  __syncthreads();
  fft_lds(x, f, lane);
  for (int m = lane; m < M; m += tpl) x[m] = cconj(cmul(x[m], f.bfil[m]));
  __syncthreads();
  fft_lds(x, f, lane);
  for (int m = lane; m < n; m += tpl) x[m] = cmul(cconj(x[m]), f.chirp[m]);
  __syncthreads();
}

// FC-Gram Dirichlet continuation of nl lines held in LDS (fftp.fpp:755-773):
// x[np+ii] = sum_jj dir(ii,jj) x[np-d+jj] + dir(C-1-ii,jj) x[d-1-jj],
// accumulated in the reference's order, in float64.
template <typename T>
__device__ void fc_continue_lds(cpx<T>* lds, int nl, int ldl, int np, int C, int d,
                                const double* __restrict__ dir) {
#pragma clang fp contract(off)
  const int top = np - d;
  for (int e = threadIdx.x; e < nl * C; e += blockDim.x) {
    const int l = e / C, ii = e - l * C;
    const cpx<T>* x = lds + l * ldl;
    const double* D0 = dir + ii * d;
    const double* D1 = dir + (C - 1 - ii) * d;
    const cpx<T> xt = x[top], xb = x[d - 1];
    double ar = D0[0] * double(xt.x) + D1[0] * double(xb.x);
    double ai = D0[0] * double(xt.y) + D1[0] * double(xb.y);
    for (int jj = 1; jj < d; ++jj) {
      const cpx<T> ut = x[top + jj], ub = x[d - 1 - jj];
      ar = ar + D0[jj] * double(ut.x);
      ar = ar + D1[jj] * double(ub.x);
      ai = ai + D0[jj] * double(ut.y);
      ai = ai + D1[jj] * double(ub.y);
    }
    lds[l * ldl + np + ii] = mkc<T>(T(ar), T(ai));
  }
}

}  // namespace spx
