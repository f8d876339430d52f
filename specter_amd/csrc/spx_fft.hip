// Line-FFT kernels of libspx: x (real<->half-complex), y, z (with the FC
// continuation and the (kx,y,z)<->(z,y,kx) transposes fused in), the fused
// gradient-product x pass of gradre/advect, and the z-line kernels of the
// boundary / projection / diagnostics stages.
//
// Build: the host part of this file is compiled in three parts so that the
// kernel instantiations build in parallel (SPX_PART = 1: plan tables and z
// transforms, 2: y and x transforms, 3: boundary, projection and diagnostics
// launchers); without SPX_PART everything is compiled in one unit.
//
// Precision: the transform kernels are templated on their arithmetic / buffer
// type T (double; float for the nonlinear pipeline of mixed-precision plans,
// where the first z pass reads the float64 state, TI = double).  The
// boundary, projection and diagnostics kernels compute in float64; they are
// templated on the storage type TS of the fields and instantiated for double.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <cstdint>
#include <thread>
#include <unordered_map>

#include "spx_engine.h"
#include "spx_wave.h"

namespace spx {

// (0,1)*k*v evaluated as the reference's complex product (-k Im v, k Re v)
template <typename T>
__device__ __forceinline__ cpx<T> ikmul(T k, cpx<T> v) { return mkc<T>(-(k * v.y), k * v.x); }

template <typename TO, typename TI>
__device__ __forceinline__ cpx<TO> ccast(cpx<TI> v) { return mkc<TO>(TO(v.x), TO(v.y)); }

// Twiddles at the head of dynamic LDS, then the line buffers.  Each thread
// belongs to line `line` (lane `lane` of tpl); xl is its line buffer.
#define SPX_PROLOGUE(CT, f)                                      \
  extern __shared__ unsigned char smem[];                        \
  CT* const tws = reinterpret_cast<CT*>(smem);                   \
  CT* const lds = tws + (f).tab;                                 \
  stage_twiddles(tws, (f));                                      \
  const int tid = threadIdx.x;                                   \
  const int line = tid >> (f).lgtpl, lane = tid & ((f).tpl - 1); \
  CT* const xl = lds + line * (f).ldl;                           \
  (void)xl;

constexpr int KEEP = kStage;  // line elements a thread can hold in registers (n <= KEEP*tpl)
// Workgroups hold nl lines of tpl threads: 256 threads, or one line of up to
// 512 threads for the longest lines (e.g. Bluestein length 2100 for nz = 1049).
constexpr int kMaxBlock = 512;

// Forward DFT of a z line in LDS: the compiled wave plan for radix-set codes
// RS_WAVE + w (tables: twiddles staged in LDS, filter and chirp global),
// the general engine otherwise.  Ends with a workgroup barrier either way.
template <typename T, int RS>
__device__ __forceinline__ void zline_dft(cpx<T>* x, const LineFFT<cpx<T>>& f, const cpx<T>* tw,
                                          int lane) {
  if constexpr (RS >= RS_WAVE) {
    using P = typename WavePlanOf<RS>::type;
    // float32-only plans are never built for float64 lines (wave_plan_for):
    // their float64 instances are not compiled
    if constexpr (sizeof(T) == 4 || P::f64) P::template dft<T>(x, lane, tw, f.bfil, f.chirp);
    __syncthreads();
  } else {
    line_dft<T, RS>(x, f, tw, lane);
  }
}

// Forward FFT of a direct (mode 0) line without the closing workgroup
// barrier: the caller synchronises the line (line_sync); the wave plan or
// the engine's passes.
template <typename T, int RS>
__device__ __forceinline__ void line_fft_nosync(cpx<T>* x, const LineFFT<cpx<T>>& f,
                                                const cpx<T>* tw, int lane) {
  if constexpr (RS >= RS_WAVE) {
    using P = typename WavePlanOf<RS>::type;
    static_assert(P::mode == 0, "direct plans only");
    P::template dft<T>(x, lane, tw, nullptr, nullptr);
  } else {
    fft_lds<T, RS>(x, f, tw, lane);
  }
}

// line elements a thread holds in registers in the z kernels (n <= KN * tpl)
template <int RS>
constexpr int keep_n() {
  if constexpr (RS >= RS_WAVE) return (WavePlanOf<RS>::type::n + 63) / 64;
  else return KEEP;
}

// idx = first, first + step, ... < count: wr(idx, rd(idx)).  KN > 0 (the
// compiled plans, count <= first + KN*step): every LDS read is issued before
// the first global store instead of one read-wait-store round trip each.
template <int KN, typename V, typename Rd, typename Wr>
__device__ __forceinline__ void lds_to_global(int first, int step, int count, Rd rd, Wr wr) {
  if constexpr (KN > 0) {
    V v[KN];
#pragma unroll
    for (int q = 0; q < KN; ++q) {
      const int idx = first + q * step;
      if (idx < count) v[q] = rd(idx);
    }
#pragma unroll
    for (int q = 0; q < KN; ++q) {
      const int idx = first + q * step;
      if (idx < count) wr(idx, v[q]);
    }
  } else {
    for (int idx = first; idx < count; idx += step) wr(idx, rd(idx));
  }
}

// XCD-aware tile order of a 2D grid.  Workgroups are handed to the 8 XCDs
// (each with its own L2) round-robin by linear id; the kx tiles of the
// transposing kernels are narrower than a cache line, so neighbouring tiles
// must meet in one L2.  Each XCD gets a contiguous range of logical tiles
// (a bijection on [0, gx*gy)): returns the logical (x, y) of this workgroup.
__device__ __forceinline__ void xcd_tile(int& bx, int& by) {
  constexpr int X = 8;
  const int gx = gridDim.x;
  const int n = gx * gridDim.y;
  const int pid = blockIdx.x + gx * blockIdx.y;
  const int q = n / X, r = n % X;
  const int x = pid % X, slot = pid / X;
  const int l = x * q + (x < r ? x : r) + slot;
  by = l / gx;
  bx = l - by * gx;
}

// ---------------------------------------------------------------------------
// z transforms in place on S (contiguous lines)
// ---------------------------------------------------------------------------
template <typename T, int RS>
__global__ void __launch_bounds__(kMaxBlock) k_z_inplace(cpx<T>* __restrict__ S, long nlines,
                                                   LineFFT<cpx<T>> f, int np, int forward,
                                                   int cont, double scale,
                                                   const double* __restrict__ dir, int C,
                                                   int d) {
  using Cp = cpx<T>;
  SPX_PROLOGUE(Cp, f);
  const int n = f.n;
  const long ln = long(blockIdx.x) * f.nl + line;
  const bool ok = ln < nlines;
  const int nload = (forward && cont) ? np : n;
  for (int k = lane; k < nload; k += f.tpl) {
    Cp v = ok ? S[ln * n + k] : mkc<T>(0, 0);
    if (!forward) v.y = -v.y;
    xl[k] = v;
  }
  __syncthreads();
  if (forward && cont) {
    fc_continue_lds(lds, f.nl, f.ldl, np, C, d, dir);
    __syncthreads();
  }
  zline_dft<T, RS>(xl, f, tws, lane);
  if (ok) {
    for (int k = lane; k < n; k += f.tpl) {
      Cp v = xl[k];
      if (!forward) {
        v.y = -v.y;
        if (k < np && scale != 1.0) v = cscale(v, T(scale));
      }
      S[ln * n + k] = v;
    }
  }
}

// Continuation only (fftp.fpp:755-773), in place on S, float64 arithmetic.
template <typename T>
__global__ void k_fc_continue(cpx<T>* __restrict__ S, long nlines, int n, int np, int C, int d,
                              const double* __restrict__ dir) {
#pragma clang fp contract(off)
  const long e = long(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= nlines * C) return;
  const long ln = e / C;
  const int ii = int(e - ln * C);
  const cpx<T>* x = S + ln * n;
  const int top = np - d;
  const double* D0 = dir + ii * d;
  const double* D1 = dir + (C - 1 - ii) * d;
  double ar = D0[0] * double(x[top].x) + D1[0] * double(x[d - 1].x);
  double ai = D0[0] * double(x[top].y) + D1[0] * double(x[d - 1].y);
  for (int jj = 1; jj < d; ++jj) {
    ar = ar + D0[jj] * double(x[top + jj].x);
    ar = ar + D1[jj] * double(x[d - 1 - jj].x);
    ai = ai + D0[jj] * double(x[top + jj].y);
    ai = ai + D1[jj] * double(x[d - 1 - jj].y);
  }
  S[ln * n + np + ii] = mkc<T>(T(ar), T(ai));
}

// ---------------------------------------------------------------------------
// forward 3D last pass: X(kx,y,z) planes < np -> continuation -> z FFT -> S(z,y,kx)
// blockIdx.z = member: S advances by n*ny*nkx, X by xstride elements per member.
// ---------------------------------------------------------------------------
// BND (mixed plans): the 2d boundary planes z = 0..d-1, np-d..np-1 come in
// float64 from Xb (planes [member][q][ky][kx], q < 2d) and the continuation
// is evaluated from them, so the float32 roundoff of the other planes is not
// amplified by the continuation.
template <typename T, int RS, bool BND>
__global__ void __launch_bounds__(kMaxBlock) k_z_x2s(const cpx<T>* __restrict__ X, cpx<T>* __restrict__ S,
                                               LineFFT<cpx<T>> f, int nkx, int ny, int np,
                                               int nofft, const double* __restrict__ dir, int C,
                                               int d, long xstride,
                                               const cpx<double>* __restrict__ Xb) {
  using Cp = cpx<T>;
  SPX_PROLOGUE(Cp, f);
  const int n = f.n;
  X += blockIdx.z * xstride;
  S += blockIdx.z * (long(n) * ny * nkx);
  int bx, j;
  xcd_tile(bx, j);
  const int i0 = bx * f.nl;
  const int nload = nofft ? n : np;
  const long pl = long(nkx) * ny;
  {
    const int l = tid & (f.nl - 1), i = i0 + l;
    if constexpr (RS >= RS_WAVE) {
      // compiled plans (tpl 64, n <= KN*tpl): the thread's loads in flight together
      constexpr int KN = keep_n<RS>();
      Cp v[KN];
#pragma unroll
      for (int q = 0; q < KN; ++q) {
        const int k = (tid >> f.lgnl) + q * f.tpl;
        v[q] = mkc<T>(0, 0);
        if (i < nkx && k < nload) v[q] = X[i + long(nkx) * j + pl * k];
      }
#pragma unroll
      for (int q = 0; q < KN; ++q) {
        const int k = (tid >> f.lgnl) + q * f.tpl;
        if (k < nload) lds[l * f.ldl + k] = v[q];
      }
    } else {
      for (int k = tid >> f.lgnl; k < nload; k += f.tpl) {
        Cp v = mkc<T>(0, 0);
        if (i < nkx) v = X[i + long(nkx) * j + pl * k];
        lds[l * f.ldl + k] = v;
      }
    }
  }
  __syncthreads();
  if (!nofft) {
    if constexpr (BND) {
#pragma clang fp contract(off)
      const int d2 = 2 * d;
      const long pb = long(nkx) * ny;
      const cpx<double>* xb = Xb + blockIdx.z * (pb * d2) + long(nkx) * j;
      // the tile's float64 boundary values, staged once behind the line
      // buffers (16-byte aligned; bnd_stage_bytes at launch): bl[l * d2 + q]
      cpx<double>* const bl = reinterpret_cast<cpx<double>*>(
          (reinterpret_cast<uintptr_t>(lds + f.nl * f.ldl) + 15) & ~uintptr_t(15));
      for (int e = tid; e < f.nl * d2; e += blockDim.x) {
        const int l = e / d2, q = e - l * d2, i = i0 + l;
        const int z = q < d ? q : np - d2 + q;
        const cpx<double> v = i < nkx ? xb[i + pb * q] : mkc<double>(0, 0);
        lds[l * f.ldl + z] = mkc<T>(T(v.x), T(v.y));
        bl[e] = v;
      }
      __syncthreads();
      // continuation (fftp.fpp:755-773 order): x[top+jj] is plane q = d+jj,
      // x[d-1-jj] is plane q = d-1-jj
      for (int e = tid; e < f.nl * C; e += blockDim.x) {
        const int l = e / C, ii = e - l * C, i = i0 + l;
        double ar = 0.0, ai = 0.0;
        if (i < nkx) {
          const cpx<double>* bq = bl + l * d2;
          const double* D0 = dir + ii * d;
          const double* D1 = dir + (C - 1 - ii) * d;
          const cpx<double> t0 = bq[d], b0 = bq[d - 1];
          ar = D0[0] * t0.x + D1[0] * b0.x;
          ai = D0[0] * t0.y + D1[0] * b0.y;
          for (int jj = 1; jj < d; ++jj) {
            const cpx<double> ut = bq[d + jj], ub = bq[d - 1 - jj];
            ar = ar + D0[jj] * ut.x;
            ar = ar + D1[jj] * ub.x;
            ai = ai + D0[jj] * ut.y;
            ai = ai + D1[jj] * ub.y;
          }
        }
        lds[l * f.ldl + np + ii] = mkc<T>(T(ar), T(ai));
      }
      __syncthreads();
    } else if (C > 0) {
      fc_continue_lds(lds, f.nl, f.ldl, np, C, d, dir);
      __syncthreads();
    }
    zline_dft<T, RS>(xl, f, tws, lane);
  }
  const int i = i0 + line;
  if (i < nkx) {
    Cp* so = S + long(n) * (j + long(ny) * i);
    lds_to_global<0, Cp>(lane, f.tpl, n, [&](int k) { return xl[k]; },
                         [&](int k, Cp v) { so[k] = v; });
  }
}

// ---------------------------------------------------------------------------
// float64 values of the backward z transform at the 2d boundary planes
// z = 0..d-1 and np-d..np-1 (the points the FC continuation reads), for the
// float32 nonlinear pipeline of mixed plans: a direct partial DFT of every
// line, 16 lanes per line (4 lines per wavefront), for v and (i kz) v
// together.  Twiddles: w1[k] = exp(+2 pi i k/n) and wt[k] = exp(+2 pi i k
// (np-d)/n) from long-double tables, then exp(+2 pi i k z/n) for the d
// consecutive z of each wall block by d-1 products (error ~ d ulp).
// blockIdx.y = component.  Output planes [comp][member][q][ky][kx].
// ---------------------------------------------------------------------------
struct BndSrc {
  const cpx<double>* s[4];
};

template <int D, int U>
__global__ void __launch_bounds__(256) k_z_s2x_bnd(BndSrc src, cpx<double>* __restrict__ XVb,
                                                   cpx<double>* __restrict__ XZb,
                                                   const cpx<double>* __restrict__ w1,
                                                   const cpx<double>* __restrict__ wt, int n,
                                                   int nkx, int ny, int mem,
                                                   const double* __restrict__ kz) {
  constexpr int D2 = 2 * D;
  constexpr int LPL = 16;  // lanes per line
  const int sub = threadIdx.x & (LPL - 1);
  const long nlines = long(nkx) * ny;
  const long ln = long(blockIdx.x) * (blockDim.x / LPL) + (threadIdx.x / LPL);
  const bool ok = ln < nlines * mem;
  const int comp = blockIdx.y;
  const cpx<double>* S = src.s[comp] + (ok ? ln : 0) * n;
  const long b = ok ? ln / nlines : 0, l = ok ? ln - b * nlines : 0;
  const long i = l / ny, jy = l - i * ny;
  const long o = ((long(comp) * mem + b) * D2) * nlines + jy * nkx + i;
  // v and (i kz) v from one load of the line.  The top block uses
  // exp(+2 pi i k (np-d+q)/n) = wt[k] exp(+2 pi i k q/n): both blocks share
  // the twiddle recurrence, the top one on u_k = v_k wt[k].
  cpx<double> acc[2][D2];
#pragma unroll
  for (int kind = 0; kind < 2; ++kind)
#pragma unroll
    for (int q = 0; q < D2; ++q) acc[kind][q] = mkc<double>(0, 0);
  if (ok) {
    // U points per lane in flight: their loads first (indices past the line
    // end clamped to 0 and their values zeroed), then the sums in the same
    // point order as one point at a time (HD 512^3, d = 5: U = 2 1.17 ms per
    // launch, U = 1 1.25, U = 4 1.27 at two waves per SIMD)
    for (int k0 = sub; k0 < n; k0 += U * LPL) {
      cpx<double> vv[U], ww[U], ss[U];
      double kk[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int k = k0 + u * LPL;
        const int kc = k < n ? k : 0;
        vv[u] = S[kc];
        kk[u] = kz[kc];
        ww[u] = wt[kc];
        ss[u] = w1[kc];
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const bool in = k0 + u * LPL < n;
        const cpx<double> v0 = in ? vv[u] : mkc<double>(0, 0);
        const cpx<double> v1 = ikmul(kk[u], v0);
        const cpx<double> wtk = ww[u];
        const cpx<double> u0 = cmul(v0, wtk), u1 = cmul(v1, wtk);
        const cpx<double> step = ss[u];
        cpx<double> t = mkc<double>(1.0, 0.0);
#pragma unroll
        for (int q = 0; q < D; ++q) {
          acc[0][q] = cadd(acc[0][q], cmul(v0, t));
          acc[0][D + q] = cadd(acc[0][D + q], cmul(u0, t));
          acc[1][q] = cadd(acc[1][q], cmul(v1, t));
          acc[1][D + q] = cadd(acc[1][D + q], cmul(u1, t));
          if (q + 1 < D) t = cmul(t, step);
        }
      }
    }
  }
  // reduce over the 16 lanes of the line (lanes of one line are contiguous)
#pragma unroll
  for (int kind = 0; kind < 2; ++kind) {
#pragma unroll
    for (int off = LPL / 2; off > 0; off >>= 1) {
#pragma unroll
      for (int q = 0; q < D2; ++q) {
        acc[kind][q].x += __shfl_xor(acc[kind][q].x, off);
        acc[kind][q].y += __shfl_xor(acc[kind][q].y, off);
      }
    }
    if (ok) {
      cpx<double>* out = kind ? XZb : XVb;
#pragma unroll
      for (int q = 0; q < D2; ++q)
        if (sub == q) out[o + q * nlines] = acc[kind][q];
    }
  }
}

// ---------------------------------------------------------------------------
// backward 3D first pass: S * (i k_dir) -> z IFFT -> X planes < nzout.
// DUAL: the same load also gives X2 <- z IFFT of (i kz) S.
// blockIdx.z = member: S advances by n*ny*nkx, X and X2 by nkx*ny*nzout.
// ---------------------------------------------------------------------------
template <typename T, int RS, bool DUAL, typename TI>
__global__ void __launch_bounds__(kMaxBlock) k_z_s2x(const cpx<TI>* __restrict__ S, cpx<T>* __restrict__ X,
                                               cpx<T>* __restrict__ X2, LineFFT<cpx<T>> f,
                                               int nkx, int ny, int nzout, int dirn, int nofft,
                                               const double* __restrict__ kx,
                                               const double* __restrict__ ky,
                                               const double* __restrict__ kz) {
  using Cp = cpx<T>;
  SPX_PROLOGUE(Cp, f);
  const int n = f.n;
  {
    const long xs = blockIdx.z * (long(nkx) * ny * nzout);
    S += blockIdx.z * (long(n) * ny * nkx);
    X += xs;
    if constexpr (DUAL) X2 += xs;
  }
  int bx, j;
  xcd_tile(bx, j);
  const int i0 = bx * f.nl;
  const int i = i0 + line;
  const bool ok = i < nkx;
  const long base = long(n) * (j + long(ny) * (ok ? i : 0));
  // the line (times i k_dirn).  The dual transform needs it again: float
  // kernels and the wave plans (whose occupancy LDS bounds) keep it in
  // registers, the other float64 kernels read it again (the second read hits
  // in L2 and keeps the registers free for occupancy)
  constexpr bool KEEPV = DUAL && (sizeof(T) == 4 || RS >= RS_WAVE);
  constexpr int KN = keep_n<RS>();
  auto load = [&](int k) {
    Cp v = mkc<T>(0, 0);
    if (ok) {
      v = ccast<T>(S[base + k]);
      if (dirn == 1) v = ikmul(T(kx[i]), v);
      else if (dirn == 2) v = ikmul(T(ky[j]), v);
      else if (dirn == 3) v = ikmul(T(kz[k]), v);
    }
    return v;
  };
  Cp keep[KEEPV ? KN : 1];
  if constexpr (RS >= RS_WAVE) {
    // compiled plans: the line's raw values (and kz) in flight together, then
    // the derivative factor and the LDS writes
    cpx<TI> raw[KN];
    double kzv[KN];
    const double kline = (ok && dirn == 1) ? kx[i] : ((ok && dirn == 2) ? ky[j] : 0.0);
#pragma unroll
    for (int q = 0; q < KN; ++q) {
      const int k = lane + q * f.tpl;
      raw[q] = mkc<TI>(0, 0);
      kzv[q] = kline;
      if (ok && k < n) {
        raw[q] = S[base + k];
        if (dirn == 3) kzv[q] = kz[k];
      }
    }
#pragma unroll
    for (int q = 0; q < KN; ++q) {
      const int k = lane + q * f.tpl;
      if (k < n) {
        Cp v = ccast<T>(raw[q]);
        if (ok && dirn != 0) v = ikmul(T(kzv[q]), v);
        if constexpr (KEEPV) keep[q] = v;
        xl[k] = cconj(v);
      }
    }
  } else {
#pragma unroll
    for (int q = 0; q < KN; ++q) {
      const int k = lane + q * f.tpl;
      if (k < n) {
        const Cp v = load(k);
        if constexpr (KEEPV) keep[q] = v;
        xl[k] = cconj(v);
      }
    }
  }
  __syncthreads();
  if (!nofft) zline_dft<T, RS>(xl, f, tws, lane);
  const long pl = long(nkx) * ny;
  const int l = tid & (f.nl - 1), ii = i0 + l;
  // (batched stores measured slower here: the z line kernels are VALU-bound)
  constexpr int KB = 0;
  if (ii < nkx) {
    Cp* xo = X + ii + long(nkx) * j;
    lds_to_global<KB, Cp>(
        tid >> f.lgnl, f.tpl, nzout, [&](int k) { return cconj(lds[l * f.ldl + k]); },
        [&](int k, Cp v) { xo[pl * k] = v; });
  }
  if constexpr (DUAL) {
    __syncthreads();
#pragma unroll
    for (int q = 0; q < KN; ++q) {
      const int k = lane + q * f.tpl;
      if (k < n) {
        if constexpr (KEEPV) xl[k] = cconj(ikmul(T(kz[k]), keep[q]));
        else xl[k] = cconj(ikmul(T(kz[k]), load(k)));
      }
    }
    __syncthreads();
    zline_dft<T, RS>(xl, f, tws, lane);
    if (ii < nkx) {
      Cp* xo = X2 + ii + long(nkx) * j;
      lds_to_global<KB, Cp>(
          tid >> f.lgnl, f.tpl, nzout, [&](int k) { return cconj(lds[l * f.ldl + k]); },
          [&](int k, Cp v) { xo[pl * k] = v; });
    }
  }
}

// ---------------------------------------------------------------------------
// y transform of X (lines along y, stride nkx), one plane per blockIdx.y.
// DUAL (backward): Y2 <- y IFFT of (i ky) Xin from the same load.
// Cy > 0 (forward): FC-Gram continuation of each line before the FFT
// (fftp.fpp:565-581): rows npy..ny-1 from rows 0..oy-1 and npy-oy..npy-1.
// ---------------------------------------------------------------------------
template <typename T, int RS, bool DUAL>
__global__ void __launch_bounds__(kMaxBlock) k_y(const cpx<T>* Xin, cpx<T>* Xout,
                                           cpx<T>* __restrict__ Y2, LineFFT<cpx<T>> f, int nkx,
                                           int ny, int forward, const double* __restrict__ ky,
                                           int npy, int Cy, int oy,
                                           const double* __restrict__ diry) {
  using Cp = cpx<T>;
  SPX_PROLOGUE(Cp, f);
  int bx, by;
  xcd_tile(bx, by);
  const int i0 = bx * f.nl;
  const long base = long(by) * nkx * ny;
  const int l = tid & (f.nl - 1), ii = i0 + l;
  const bool cont = forward && Cy > 0;
  const int nload = cont ? npy : ny;
  if constexpr (RS >= RS_WAVE) {
    // compiled plans (tpl 64, ny <= KN*tpl): all of a thread's loads in
    // flight at once instead of one load-wait-store round trip per element
    constexpr int KN = keep_n<RS>();
    Cp v[KN];
#pragma unroll
    for (int q = 0; q < KN; ++q) {
      const int jy = (tid >> f.lgnl) + q * f.tpl;
      v[q] = mkc<T>(0, 0);
      if (ii < nkx && jy < nload) v[q] = Xin[base + ii + long(nkx) * jy];
    }
#pragma unroll
    for (int q = 0; q < KN; ++q) {
      const int jy = (tid >> f.lgnl) + q * f.tpl;
      if (!forward) v[q].y = -v[q].y;
      if (jy < nload) lds[l * f.ldl + jy] = v[q];
    }
  } else {
    for (int jy = tid >> f.lgnl; jy < nload; jy += f.tpl) {
      Cp v = mkc<T>(0, 0);
      if (ii < nkx) v = Xin[base + ii + long(nkx) * jy];
      if (!forward) v.y = -v.y;
      lds[l * f.ldl + jy] = v;
    }
  }
  __syncthreads();
  if (cont) {
    fc_continue_lds(lds, f.nl, f.ldl, npy, Cy, oy, diry);
    __syncthreads();
  }
  Cp keep[DUAL ? KEEP : 1];
  if constexpr (DUAL) {
#pragma unroll
    for (int q = 0; q < KEEP; ++q) {
      const int k = lane + q * f.tpl;
      if (k < ny) keep[q] = xl[k];
    }
  }
  zline_dft<T, RS>(xl, f, tws, lane);
  constexpr int KB = RS >= RS_WAVE ? keep_n<RS>() : 0;
  if (ii < nkx) {
    Cp* xo = Xout + base + ii;
    lds_to_global<KB, Cp>(
        tid >> f.lgnl, f.tpl, ny,
        [&](int jy) {
          Cp v = lds[l * f.ldl + jy];
          if (!forward) v.y = -v.y;
          return v;
        },
        [&](int jy, Cp v) { xo[long(nkx) * jy] = v; });
  }
  if constexpr (DUAL) {
    __syncthreads();
#pragma unroll
    for (int q = 0; q < KEEP; ++q) {
      const int k = lane + q * f.tpl;
      if (k < ny) xl[k] = cconj(ikmul(T(ky[k]), cconj(keep[q])));
    }
    __syncthreads();
    zline_dft<T, RS>(xl, f, tws, lane);
    if (ii < nkx) {
      Cp* yo = Y2 + base + ii;
      lds_to_global<KB, Cp>(
          tid >> f.lgnl, f.tpl, ny, [&](int jy) { return cconj(lds[l * f.ldl + jy]); },
          [&](int jy, Cp v) { yo[long(nkx) * jy] = v; });
    }
  }
}

// ---------------------------------------------------------------------------
// x real -> half complex: rows 2p, 2p+1 packed as one complex line
// ---------------------------------------------------------------------------
template <typename T, int RS>
__global__ void __launch_bounds__(kMaxBlock) k_xr2c(const T* __restrict__ R, cpx<T>* __restrict__ X,
                                              LineFFT<cpx<T>> f, int nkx, long npairs) {
  using Cp = cpx<T>;
  SPX_PROLOGUE(Cp, f);
  const int nx = f.n;
  const long p = long(blockIdx.x) * f.nl + line;
  const bool ok = p < npairs;
  const T* r0 = R + long(nx) * (2 * (ok ? p : 0));
  // float64 rows: batched loads; float32 rows keep the loop (batching them
  // tripled the register count and halved the occupancy)
  if constexpr (RS >= RS_WAVE && sizeof(T) == 8) {
    constexpr int KN = keep_n<RS>();
    T a[KN], b[KN];
#pragma unroll
    for (int q = 0; q < KN; ++q) {
      const int x = lane + q * f.tpl;
      a[q] = b[q] = T(0);
      if (ok && x < nx) {
        a[q] = r0[x];
        b[q] = r0[x + nx];
      }
    }
#pragma unroll
    for (int q = 0; q < KN; ++q) {
      const int x = lane + q * f.tpl;
      if (x < nx) xl[x] = mkc<T>(a[q], b[q]);
    }
  } else {
    for (int x = lane; x < nx; x += f.tpl) xl[x] = ok ? mkc<T>(r0[x], r0[x + nx]) : mkc<T>(0, 0);
  }
  __syncthreads();
  zline_dft<T, RS>(xl, f, tws, lane);
  if (ok) {
    Cp* o0 = X + long(nkx) * (2 * p);
    const T h = T(0.5);
    for (int k = lane; k < nkx; k += f.tpl) {
      const Cp Z = xl[k];
      const Cp Zc = xl[k == 0 ? 0 : nx - k];
      o0[k] = mkc<T>((Z.x + Zc.x) * h, (Z.y - Zc.y) * h);
      o0[k + nkx] = mkc<T>((Z.y + Zc.y) * h, -(Z.x - Zc.x) * h);
    }
  }
}

// Load of an x c2r row pair into LDS: Z = A + iB from the two half spectra
// (FFTW c2r semantics: Im of kx = 0 and kx = nx/2 ignored), conjugated for the
// backward transform; dirx multiplies both by i kx first.
// KN > 0 (compiled plans, nx <= KN*tpl): every load of the thread is issued
// before the first is used.
template <typename T>
__device__ __forceinline__ cpx<T> c2r_pair_value(cpx<T> A, cpx<T> B, int m, int k, int h,
                                                 int dirx, double kxk) {
  if (dirx) {
    A = ikmul(T(kxk), A);
    B = ikmul(T(kxk), B);
  }
  if (k == 0 || k == h) {
    A.y = 0;
    B.y = 0;
  }
  cpx<T> v = m > h ? mkc<T>(A.x + B.y, -A.y + B.x) : mkc<T>(A.x - B.y, A.y + B.x);
  v.y = -v.y;
  return v;
}

template <typename T, int KN = 0, int CH = KN>
__device__ __forceinline__ void load_c2r_pair(cpx<T>* xl, const cpx<T>* __restrict__ a0, int nkx,
                                              int nx, int lane, int tpl, bool ok, int dirx,
                                              const double* __restrict__ kx) {
  const int h = nx / 2;
  if constexpr (KN > 0) {
    // CH elements in flight at a time (CH < KN bounds the registers: the
    // chunk loop stays a loop, so the compiler cannot hoist every load)
#pragma unroll(CH == KN ? KN : 1)
    for (int c0 = 0; c0 < KN; c0 += CH) {
      cpx<T> A[CH], B[CH];
      double kk[CH];
#pragma unroll
      for (int q = 0; q < CH; ++q) {
        const int m = lane + (c0 + q) * tpl;
        const int k = m > h ? nx - m : m;
        A[q] = B[q] = mkc<T>(0, 0);
        kk[q] = 0.0;
        if (ok && m < nx) {
          A[q] = a0[k];
          B[q] = a0[k + nkx];
          if (dirx) kk[q] = kx[k];
        }
      }
#pragma unroll
      for (int q = 0; q < CH; ++q) {
        const int m = lane + (c0 + q) * tpl;
        const int k = m > h ? nx - m : m;
        if (m < nx) xl[m] = ok ? c2r_pair_value(A[q], B[q], m, k, h, dirx, kk[q]) : mkc<T>(0, 0);
      }
    }
  } else {
    for (int m = lane; m < nx; m += tpl) {
      cpx<T> v = mkc<T>(0, 0);
      if (ok) {
        const bool up = m > h;
        const int k = up ? nx - m : m;
        cpx<T> A = a0[k];
        cpx<T> B = a0[k + nkx];
        if (dirx) {
          A = ikmul(T(kx[k]), A);
          B = ikmul(T(kx[k]), B);
        }
        if (k == 0 || k == h) {
          A.y = 0;
          B.y = 0;
        }
        v = up ? mkc<T>(A.x + B.y, -A.y + B.x) : mkc<T>(A.x - B.y, A.y + B.x);
        v.y = -v.y;
      }
      xl[m] = v;
    }
  }
}

template <int RS>
constexpr int batch_n() {
  if constexpr (RS >= RS_WAVE) return keep_n<RS>();
  else return 0;
}
// the fused x stage holds the velocity of its row pair in registers, so its
// loads are batched in chunks
template <typename T, int RS>
constexpr int xg_batch() {
  return batch_n<RS>();
}
template <typename T, int RS>
constexpr int xg_chunk() {
  // four elements in flight at a time: at 512^3 the x stage takes 2.2 ms
  // (float32) / 4.2 ms (float64) per RK stage against 2.9 / 5.5 with one
  // and 4.3 / 5.3 with all eight (spills)
  return batch_n<RS>() >= 4 ? 4 : batch_n<RS>();
}

template <typename T, int RS>
__global__ void __launch_bounds__(kMaxBlock) k_xc2r(const cpx<T>* __restrict__ X, T* __restrict__ R,
                                              LineFFT<cpx<T>> f, int nkx, long npairs, int dirx,
                                              const double* __restrict__ kx) {
  using Cp = cpx<T>;
  SPX_PROLOGUE(Cp, f);
  const int nx = f.n;
  const long p = long(blockIdx.x) * f.nl + line;
  const bool ok = p < npairs;
  load_c2r_pair<T, batch_n<RS>()>(xl, X + long(nkx) * (2 * (ok ? p : 0)), nkx, nx, lane, f.tpl, ok, dirx, kx);
  __syncthreads();
  zline_dft<T, RS>(xl, f, tws, lane);
  if (ok) {
    T* r0 = R + long(nx) * (2 * p);
    for (int x = lane; x < nx; x += f.tpl) {
      const Cp z = xl[x];
      r0[x] = z.x;
      r0[x + nx] = -z.y;
    }
  }
}

// ((u*Dx + v*Dy) + w*Dz)*tmp on a row pair, the three derivatives c2r'd in
// LDS one after the other and accumulated in registers
// (pseudospec_hd.f90:254-312, pseudospec_phd.f90:63-108).
template <typename T, int RS>
__global__ void __launch_bounds__(kMaxBlock) k_xgrad_prod(const cpx<T>* __restrict__ YV,
                                                    const cpx<T>* __restrict__ YY,
                                                    const cpx<T>* __restrict__ XZ,
                                                    const T* __restrict__ u,
                                                    const T* __restrict__ v,
                                                    const T* __restrict__ w, T* __restrict__ out,
                                                    LineFFT<cpx<T>> f, int nkx, long npairs,
                                                    double tmpd, const double* __restrict__ kx,
                                                    long cstride, long rstride) {
#pragma clang fp contract(off)
  using Cp = cpx<T>;
  SPX_PROLOGUE(Cp, f);
  // blockIdx.y: component (YV, YY, XZ advance by cstride, out by rstride)
  YV += blockIdx.y * cstride;
  YY += blockIdx.y * cstride;
  XZ += blockIdx.y * cstride;
  out += blockIdx.y * rstride;
  const int nx = f.n;
  const T tmp = T(tmpd);
  const long p = long(blockIdx.x) * f.nl + line;
  const bool ok = p < npairs;
  const long ro = long(nx) * (2 * (ok ? p : 0));
  const long co = long(nkx) * (2 * (ok ? p : 0));
  T accA[KEEP], accB[KEEP];
  const Cp* srcs[3] = {YV, YY, XZ};
  const T* mult[3] = {u, v, w};
  for (int r = 0; r < 3; ++r) {
    load_c2r_pair<T, batch_n<RS>()>(xl, srcs[r] + co, nkx, nx, lane, f.tpl, ok, r == 0 ? 1 : 0, kx);
    // the velocity values this thread multiplies with, loaded before the
    // transform so that their latency overlaps it
    const T* m = mult[r] + ro;
    T ma[KEEP], mb[KEEP];
#pragma unroll
    for (int q = 0; q < KEEP; ++q) {
      const int x = lane + q * f.tpl;
      const bool in = x < nx && ok;
      ma[q] = in ? m[x] : T(0);
      mb[q] = in ? m[x + nx] : T(0);
    }
    __syncthreads();
    zline_dft<T, RS>(xl, f, tws, lane);
#pragma unroll
    for (int q = 0; q < KEEP; ++q) {
      const int x = lane + q * f.tpl;
      if (x < nx && ok) {
        const Cp z = xl[x];
        const T a = z.x, b = -z.y;
        if (r == 0) {
          accA[q] = ma[q] * a;
          accB[q] = mb[q] * b;
        } else if (r == 1) {
          accA[q] = accA[q] + ma[q] * a;
          accB[q] = accB[q] + mb[q] * b;
        } else {
          accA[q] = (accA[q] + ma[q] * a) * tmp;
          accB[q] = (accB[q] + mb[q] * b) * tmp;
        }
      }
    }
    __syncthreads();
  }
  if (ok) {
#pragma unroll
    for (int q = 0; q < KEEP; ++q) {
      const int x = lane + q * f.tpl;
      if (x < nx) {
        out[ro + x] = accA[q];
        out[ro + x + nx] = accB[q];
      }
    }
  }
}

// gradre's whole x stage on a row pair (pseudospec_hd.f90:254-312): the
// velocity components u, v, w = c2r(YV_c) are formed first and kept in
// registers (and stored to rv[c] when given: advect reuses them), then for
// each component c ((u Dx v_c + v Dy v_c) + w Dz v_c) tmp from the c2r of
// i kx YV_c, YY_c and XZ_c, in k_xgrad_prod's operation order.  It saves the
// traffic of the real velocity (written once and read nine times per stage
// by xc2r x 3 + xgrad_prod x 3).  A line is owned by its tpl threads, so
// the twelve transforms of a row pair synchronise with line_sync.
template <typename T>
struct XGradArgs {
  const cpx<T>* yv[3];
  const cpx<T>* yy[3];
  const cpx<T>* xz[3];
  T* rv[3];   // may be null
  T* out[3];
};

template <typename T, int RS>
__global__ void __launch_bounds__(kMaxBlock, sizeof(T) == 4 ? 4 : 2) k_xgrad_fused(
    XGradArgs<T> a, LineFFT<cpx<T>> f, int nkx, long npairs, double tmpd,
    const double* __restrict__ kx) {
#pragma clang fp contract(off)
  using Cp = cpx<T>;
  SPX_PROLOGUE(Cp, f);
  const int nx = f.n;
  const T tmp = T(tmpd);
  const long p = long(blockIdx.x) * f.nl + line;
  const bool ok = p < npairs;
  const long ro = long(nx) * (2 * (ok ? p : 0));
  const long co = long(nkx) * (2 * (ok ? p : 0));
  T vel[3][2][KEEP];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    load_c2r_pair<T, xg_batch<T, RS>(), xg_chunk<T, RS>()>(xl, a.yv[c] + co, nkx, nx, lane, f.tpl, ok, 0, kx);
    // the first barrier also publishes the staged twiddles
    if (c == 0) __syncthreads();
    else line_sync(f.tpl);
    line_fft_nosync<T, RS>(xl, f, tws, lane);
    line_sync(f.tpl);
#pragma unroll
    for (int q = 0; q < KEEP; ++q) {
      const int x = lane + q * f.tpl;
      vel[c][0][q] = vel[c][1][q] = T(0);
      if (x < nx) {
        const Cp z = xl[x];
        vel[c][0][q] = z.x;
        vel[c][1][q] = -z.y;
        if (a.rv[c] && ok) {
          a.rv[c][ro + x] = z.x;
          a.rv[c][ro + x + nx] = -z.y;
        }
      }
    }
    line_sync(f.tpl);
  }
#pragma unroll 1
  for (int c = 0; c < 3; ++c) {
    T accA[KEEP], accB[KEEP];
    const Cp* srcs[3] = {a.yv[c], a.yy[c], a.xz[c]};
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      load_c2r_pair<T, xg_batch<T, RS>(), xg_chunk<T, RS>()>(xl, srcs[r] + co, nkx, nx, lane, f.tpl, ok, r == 0 ? 1 : 0, kx);
      line_sync(f.tpl);
      line_fft_nosync<T, RS>(xl, f, tws, lane);
      line_sync(f.tpl);
#pragma unroll
      for (int q = 0; q < KEEP; ++q) {
        const int x = lane + q * f.tpl;
        if (x < nx) {
          const Cp z = xl[x];
          const T va = z.x, vb = -z.y;
          if (r == 0) {
            accA[q] = vel[0][0][q] * va;
            accB[q] = vel[0][1][q] * vb;
          } else if (r == 1) {
            accA[q] = accA[q] + vel[1][0][q] * va;
            accB[q] = accB[q] + vel[1][1][q] * vb;
          } else {
            accA[q] = (accA[q] + vel[2][0][q] * va) * tmp;
            accB[q] = (accB[q] + vel[2][1][q] * vb) * tmp;
          }
        }
      }
      line_sync(f.tpl);
    }
    if (ok) {
#pragma unroll
      for (int q = 0; q < KEEP; ++q) {
        const int x = lane + q * f.tpl;
        if (x < nx) {
          a.out[c][ro + x] = accA[q];
          a.out[c][ro + x + nx] = accB[q];
        }
      }
    }
  }
}

// ---------------------------------------------------------------------------
// float64 boundary kernels; TS is the storage type of the fields
// ---------------------------------------------------------------------------
using Cd = cpx<double>;
using LFd = LineFFT<Cd>;

// goto_domain_w_boundaries + noslip_z + goto_3d_fourier fused per line
// (boundary_mod.fpp:72-194, vboundary.f90:154-211, sboundary.f90:122-165).
// comp 0: vx (kx), comp 1: vy (ky); scalar mode: wall values are zero.

template <typename TS, int RS>
__global__ void __launch_bounds__(kMaxBlock) k_noslip(cpx<TS>* __restrict__ vx, cpx<TS>* __restrict__ vy,
                                                const Cd* __restrict__ pr, LFd f, long nlines,
                                                int ny, int np, double invn, WallParams wp,
                                                const double* __restrict__ kx,
                                                const double* __restrict__ ky,
                                                const double* __restrict__ dir, int C, int d,
                                                RkLoad rk, int rk_on) {
#pragma clang fp contract(off)
  SPX_PROLOGUE(Cd, f);
  const int n = f.n;
  const int comp = blockIdx.y;
  cpx<TS>* V = comp ? vy : vx;
  const long ln = long(blockIdx.x) * f.nl + line;
  const bool ok = ln < nlines;
  const long li = ok ? ln / ny : 0;
  const int lj = ok ? int(ln - li * ny) : 0;
  // wall value of lanes 0 (z = 0) and 1 (z = Lz), read before the transform
  Cd val = mkc<double>(0, 0);
  if (lane < 2 && ok && !wp.scalar) {
    const int ind = lane ? np - 1 : 0;
    const Cd pv = pr[ln * n + ind];
    const double kxi = kx[li], kyj = ky[lj];
    val = cscale(ikmul(comp ? kyj : kxi, pv), wp.tau);
    if (kxi == 0.0 && kyj == 0.0) val = mkc<double>(wp.nxny * wp.vb[comp][lane], 0.0);
  }
  for (int k = lane; k < n; k += f.tpl) {
    Cd v = mkc<double>(0, 0);
    if (ok) {
      v = ccast<double>(V[ln * n + k]);
      if (rk_on) v = rk_on_load(rk, comp, v, ln * n + k, li, lj, k);
    }
    xl[k] = cconj(v);
  }
  __syncthreads();
  line_dft<double, RS>(xl, f, tws, lane);
  for (int k = lane; k < np; k += f.tpl) xl[k] = cscale(cconj(xl[k]), invn);
  __syncthreads();
  if (lane < 2 && ok) xl[lane ? np - 1 : 0] = val;
  __syncthreads();
  fc_continue_lds(lds, f.nl, f.ldl, np, C, d, dir);
  __syncthreads();
  line_dft<double, RS>(xl, f, tws, lane);
  if (ok)
    for (int k = lane; k < n; k += f.tpl) V[ln * n + k] = ccast<TS>(xl[k]);
}

// goto_domain_w_boundaries + the magnetic wall conditions + goto_3d_fourier
// fused per line (boundary_mod.fpp:72-194, bboundary.f90:192-344,
// fcgram_mod.f90:437-468, 550-568); blockIdx.y = component (a0, a1, a2).
template <typename TS, int RS>
__global__ void __launch_bounds__(kMaxBlock) k_bwall(Cd* __restrict__ a0, Cd* __restrict__ a1,
                                                     Cd* __restrict__ a2, LFd f, long nlines,
                                                     int np, double invn, BWallParams bp,
                                                     const double* __restrict__ khom,
                                                     const double* __restrict__ dir, int C,
                                                     int d) {
#pragma clang fp contract(off)
  SPX_PROLOGUE(Cd, f);
  const int n = f.n;
  const int comp = blockIdx.y;
  Cd* V = comp == 0 ? a0 : (comp == 1 ? a1 : a2);
  const long ln = long(blockIdx.x) * f.nl + line;
  const bool ok = ln < nlines;
  for (int k = lane; k < n; k += f.tpl) xl[k] = ok ? cconj(V[ln * n + k]) : mkc<double>(0, 0);
  __syncthreads();
  line_dft<double, RS>(xl, f, tws, lane);
  for (int k = lane; k < np; k += f.tpl) xl[k] = cscale(cconj(xl[k]), invn);
  __syncthreads();
  // lane 0: the wall at z = 0, lane 1: z = Lz (their stencils do not overlap
  // for np > d)
  if (lane < 2 && ok) {
    const int m = bp.mode[comp][lane];
    if (m) {
      const int ind = lane ? np - 1 : 0;
      xl[ind] = mkc<double>(0, 0);
      if (m >= 2) {
        const double* vec = m == 3 ? bp.neu2 : bp.neu1;
        Cd acc = cscale(xl[ind], vec[d - 1]);
        for (int k = 0; k < d - 1; ++k)
          acc = cadd(acc, cscale(xl[lane ? np - d + k : d - 1 - k], vec[k]));
        if (m == 4) {
          const double den = khom[ln] * vec[d - 1] + 1;
          acc = mkc<double>(acc.x / den, acc.y / den);
        }
        xl[ind] = acc;
      }
    }
  }
  __syncthreads();
  fc_continue_lds(lds, f.nl, f.ldl, np, C, d, dir);
  __syncthreads();
  line_dft<double, RS>(xl, f, tws, lane);
  if (ok)
    for (int k = lane; k < n; k += f.tpl) V[ln * n + k] = xl[k];
}

// laplace_z pure-Neumann coefficients (boundary_mod.fpp:531-560)
struct Coef {
  Cd c1, c2;
};
__device__ __forceinline__ Coef neumann_coef(Cd b1, Cd b2, double kh, double Lz, bool zero_mode) {
#pragma clang fp contract(off)
  Coef r;
  if (zero_mode) {
    r.c1 = b1;
    r.c2 = mkc<double>(0, 0);
    return r;
  }
  const double tmp = 1.0 / (kh * (1 - exp(-2 * kh * Lz)));
  const double e1 = exp(-kh * Lz);
  r.c1 = cscale(csub(b2, cscale(b1, e1)), tmp);
  r.c2 = cscale(cadd(mkc<double>(-b1.x, -b1.y), cscale(b2, e1)), tmp);
  return r;
}

// laplace_z coefficients for the wall kinds (k0, k1) (boundary_mod.fpp:497-621):
// 0 Dirichlet, 1 Neumann, 2 Robin with constant khom; (1,1), (0,0), (2,2), (0,2)
__device__ __forceinline__ Coef laplace_coef(int k0, int k1, Cd b1, Cd b2, double kh, double Lz,
                                             bool zero_mode) {
#pragma clang fp contract(off)
  if (k0 == 1) return neumann_coef(b1, b2, kh, Lz, zero_mode);
  Coef r;
  if (k0 == 0 && k1 == 0) {
    if (zero_mode) {
      const Cd t = csub(b2, b1);
      r.c1 = mkc<double>(t.x / Lz, t.y / Lz);
      r.c2 = b1;
      return r;
    }
    const double tmp = 1.0 / (1 - exp(-2 * kh * Lz));
    const double e1 = exp(-kh * Lz);
    r.c1 = cscale(csub(b2, cscale(b1, e1)), tmp);
    r.c2 = cscale(csub(b1, cscale(b2, e1)), tmp);
    return r;
  }
  if (k0 == 2) {  // (2,2)
    if (zero_mode) {
      r.c1 = b1;
      r.c2 = mkc<double>(0, 0);
      return r;
    }
    const double tmp = 1.0 / (2 * kh);
    r.c1 = cscale(b2, tmp);
    r.c2 = cscale(b1, tmp);
    return r;
  }
  // (0,2)
  if (zero_mode) {
    r.c1 = b2;
    r.c2 = b1;
    return r;
  }
  const double tmp = 1.0 / (2 * kh);
  r.c1 = cscale(b2, tmp);
  r.c2 = cscale(csub(cscale(cscale(b1, 2.0), kh), cscale(b2, exp(-kh * Lz))), tmp);
  return r;
}

// d <- IFFT_z(d)*tmp + a_harm(z)  (:354-381 with laplace_z :631-675); d is float64
// (TS = double: the kernels of this family are instantiated per type)
template <typename TS, int RS>
__global__ void __launch_bounds__(kMaxBlock) k_proj_pressure(Cd* __restrict__ dd,
                                                       const Cd* __restrict__ bc, LFd f,
                                                       long nlines, int ny, double tmp, double Lz,
                                                       const double* __restrict__ khom,
                                                       const double* __restrict__ z, int k0,
                                                       int k1) {
#pragma clang fp contract(off)
  SPX_PROLOGUE(Cd, f);
  const int n = f.n;
  const long ln = long(blockIdx.x) * f.nl + line;
  const bool ok = ln < nlines;
  if constexpr (RS >= RS_WAVE) {
    constexpr int KN = keep_n<RS>();
    Cd v[KN];
#pragma unroll
    for (int q = 0; q < KN; ++q) {
      const int k = lane + q * f.tpl;
      v[q] = mkc<double>(0, 0);
      if (ok && k < n) v[q] = cconj(dd[ln * n + k]);
    }
#pragma unroll
    for (int q = 0; q < KN; ++q) {
      const int k = lane + q * f.tpl;
      if (k < n) xl[k] = v[q];
    }
  } else {
    for (int k = lane; k < n; k += f.tpl) xl[k] = ok ? cconj(dd[ln * n + k]) : mkc<double>(0, 0);
  }
  // (kx, ky) = (0, 0) is the only line with khom = 0; coefficients formed
  // before the transform
  const double kh = ok ? khom[ln] : 1.0;
  const bool zm = kh == 0.0;
  Coef cf;
  cf.c1 = cf.c2 = mkc<double>(0, 0);
  if (ok) cf = laplace_coef(k0, k1, bc[2 * ln], bc[2 * ln + 1], kh, Lz, zm);
  __syncthreads();
  zline_dft<double, RS>(xl, f, tws, lane);
  if (!ok) return;
  for (int k = lane; k < n; k += f.tpl) {
    Cd a;
    if (zm) {
      a = mkc<double>(cf.c1.x * z[k] + cf.c2.x, 0.0);
    } else {
      const double ep = exp(kh * (z[k] - Lz));
      const double em = exp(-kh * z[k]);
      a = cadd(cscale(cf.c1, ep), cscale(cf.c2, em));
    }
    dd[ln * n + k] = cadd(cscale(cconj(xl[k]), tmp), a);
  }
}

// k_proj_pressure at z = 0 and z = np-1 only, one thread per line, from the
// wall values pw = IFFT_z(d)/nz formed by the projection kernel
static __global__ void k_pressure_walls(Cd* __restrict__ dd, const Cd* __restrict__ bc,
                                 const Cd* __restrict__ pw, long nlines, int n, int np,
                                 double Lz, const double* __restrict__ khom,
                                 const double* __restrict__ z) {
#pragma clang fp contract(off)
  const long ln = long(blockIdx.x) * blockDim.x + threadIdx.x;
  if (ln >= nlines) return;
  const double kh = khom[ln];
  const bool zm = kh == 0.0;
  const Coef cf = neumann_coef(bc[2 * ln], bc[2 * ln + 1], kh, Lz, zm);
#pragma unroll
  for (int w = 0; w < 2; ++w) {
    const int k = w ? np - 1 : 0;
    Cd a;
    if (zm) {
      a = mkc<double>(cf.c1.x * z[k] + cf.c2.x, 0.0);
    } else {
      const double ep = exp(kh * (z[k] - Lz));
      const double em = exp(-kh * z[k]);
      a = cadd(cscale(cf.c1, ep), cscale(cf.c2, em));
    }
    dd[ln * n + k] = cadd(pw[2 * ln + w], a);
  }
}

// harmonic correction: A = FFT_z(cont(a_harm)), B = FFT_z(cont(d a_harm/dz));
// vx -= i kx A, vy -= i ky A, vz -= B   (:383-399).  Two buffers per line.
template <typename TS, int RS>
__global__ void __launch_bounds__(kMaxBlock) k_proj_harmonic(cpx<TS>* __restrict__ va,
                                                       cpx<TS>* __restrict__ vb,
                                                       cpx<TS>* __restrict__ vc,
                                                       const Cd* __restrict__ bc, LFd f,
                                                       long nlines, int ny, int np, double Lz,
                                                       const double* __restrict__ khom,
                                                       const double* __restrict__ z,
                                                       const double* __restrict__ kx,
                                                       const double* __restrict__ ky,
                                                       const double* __restrict__ dir, int C,
                                                       int d, int k0, int k1) {
#pragma clang fp contract(off)
  SPX_PROLOGUE(Cd, f);
  const int n = f.n;
  Cd* const ldsB = lds + f.nl * f.ldl;
  Cd* const xb = ldsB + line * f.ldl;
  const long ln = long(blockIdx.x) * f.nl + line;
  const bool ok = ln < nlines;
  const long i = ok ? ln / ny : 0;
  const int j = ok ? int(ln - i * ny) : 0;
  const double kh = ok ? khom[ln] : 0.0;
  const bool zm = kh == 0.0;
  Coef cf;
  cf.c1 = cf.c2 = mkc<double>(0, 0);
  if (ok) cf = laplace_coef(k0, k1, bc[2 * ln], bc[2 * ln + 1], kh, Lz, zm);
  // the field values this thread corrects, read before the transforms (the
  // two line buffers bound the occupancy, not the registers)
  Cd pa[KEEP], pb[KEEP], pc[KEEP];
#pragma unroll
  for (int q = 0; q < KEEP; ++q) {
    const int k = lane + q * f.tpl;
    if (ok && k < n) {
      const long o = ln * n + k;
      pa[q] = ccast<double>(va[o]);
      pb[q] = ccast<double>(vb[o]);
      pc[q] = ccast<double>(vc[o]);
    }
  }
  for (int k = lane; k < np; k += f.tpl) {
    Cd a = mkc<double>(0, 0), b = mkc<double>(0, 0);
    if (ok) {
      if (zm) {
        a = mkc<double>(cf.c1.x * z[k] + cf.c2.x, 0.0);
        b = mkc<double>(cf.c1.x, 0.0);
      } else {
        const double ep = exp(kh * (z[k] - Lz));
        const double em = exp(-kh * z[k]);
        const Cd t1 = cscale(cf.c1, ep), t2 = cscale(cf.c2, em);
        a = cadd(t1, t2);
        b = cscale(csub(t1, t2), kh);
      }
    }
    xl[k] = a;
    xb[k] = b;
  }
  __syncthreads();
  fc_continue_lds(lds, f.nl, f.ldl, np, C, d, dir);
  fc_continue_lds(ldsB, f.nl, f.ldl, np, C, d, dir);
  __syncthreads();
  line_dft<double, RS>(xl, f, tws, lane);
  line_dft<double, RS>(xb, f, tws, lane);
  if (!ok) return;
  const double kxi = kx[i], kyj = ky[j];
#pragma unroll
  for (int q = 0; q < KEEP; ++q) {
    const int k = lane + q * f.tpl;
    if (k < n) {
      const Cd A = xl[k], B = xb[k];
      const long o = ln * n + k;
      va[o] = ccast<TS>(csub(pa[q], ikmul(kxi, A)));
      vb[o] = ccast<TS>(csub(pb[q], ikmul(kyj, A)));
      vc[o] = ccast<TS>(csub(pc[q], B));
    }
  }
}

// The two harmonic profiles of each distinct khom value u on the physical
// points, rows 2u and 2u+1 of tab (continued and transformed afterwards).
static __global__ void k_harm_profiles(Cd* __restrict__ tab, const double* __restrict__ khv, int U,
                                int n, int np, const double* __restrict__ z, double Lz) {
#pragma clang fp contract(off)
  const long e = long(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= long(U) * np) return;
  const int u = int(e / np), k = int(e - long(u) * np);
  const double kh = khv[u];
  double a1, a2;
  if (kh == 0.0) {
    a1 = z[k];
    a2 = 1.0;
  } else {
    a1 = exp(kh * (z[k] - Lz));
    a2 = exp(-kh * z[k]);
  }
  tab[(2 * long(u)) * n + k] = mkc<double>(a1, 0.0);
  tab[(2 * long(u) + 1) * n + k] = mkc<double>(a2, 0.0);
}

// Harmonic correction from the basis table, one wavefront per line:
// A = c1 P1 + c2 P2, B = kh (c1 P1 - c2 P2) (kh = 0: A = Re c1 P1 + Re c2 P2,
// B = Re c1 P2); vx -= i kx A, vy -= i ky A, vz -= B   (:383-399).
template <typename TS>
__global__ void __launch_bounds__(256) k_proj_harmonic_tab(cpx<TS>* __restrict__ va,
                                                         cpx<TS>* __restrict__ vb,
                                                         cpx<TS>* __restrict__ vc,
                                                         const Cd* __restrict__ bc,
                                                         const Cd* __restrict__ tab,
                                                         const int* __restrict__ hidx,
                                                         long nlines, int ny, int n, double Lz,
                                                         const double* __restrict__ khom,
                                                         const double* __restrict__ kx,
                                                         const double* __restrict__ ky, int k0,
                                                         int k1) {
#pragma clang fp contract(off)
  const long ln = long(blockIdx.x) * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (ln >= nlines) return;
  const long i = ln / ny;
  const int j = int(ln - i * ny);
  const double kh = khom[ln];
  const bool zm = kh == 0.0;
  const Coef cf = laplace_coef(k0, k1, bc[2 * ln], bc[2 * ln + 1], kh, Lz, zm);
  const Cd* P1 = tab + 2 * long(hidx[ln]) * n;
  const Cd* P2 = P1 + n;
  const double kxi = kx[i], kyj = ky[j];
  for (int k = lane; k < n; k += 64) {
    const Cd p1 = P1[k], p2 = P2[k];
    Cd A, B;
    if (zm) {
      A = cadd(cscale(p1, cf.c1.x), cscale(p2, cf.c2.x));
      B = cscale(p2, cf.c1.x);
    } else {
      const Cd u1 = cmul(cf.c1, p1), u2 = cmul(cf.c2, p2);
      A = cadd(u1, u2);
      B = cscale(csub(u1, u2), kh);
    }
    const long o = ln * n + k;
    va[o] = ccast<TS>(csub(ccast<double>(va[o]), ikmul(kxi, A)));
    vb[o] = ccast<TS>(csub(ccast<double>(vb[o]), ikmul(kyj, A)));
    vc[o] = ccast<TS>(csub(ccast<double>(vc[o]), B));
  }
}

// sol_project in one pass per line (boundary_mod.fpp:197-402 with
// poisson_inhomogeneous :405-448 and laplace_z :451-678), one wavefront per
// (kx, ky) line holding its KPL points per lane in registers:
//   d = -i(kx a + ky b + kz c)/kz2 (d(0,0,0) = 0), v -= i k d   (first projection)
//   bc = IFFT_z(c/nz) at z = 0 and z = Lz as two DFT sums        (wall data)
//   A = c1 P1 + c2 P2, B = kh (c1 P1 - c2 P2) from bc and the    (harmonic correction
//   basis table; v -= (i kx A, i ky A, B)                           from the table)
// The arithmetic of every value is that of k_poisson_project_bc followed by
// k_proj_harmonic_tab; the projected velocity never goes through HBM.
// vz takes the RK update on its load when rk_on; pw (when non-null) gets
// IFFT_z(d)/nz at the two walls and d is then not written there.
// RKM < 0: any nz, RK update chosen at run time (rk_on); RKM >= 0: nz >=
// 64 (KPL - 1), so that only the last of a lane's points can lie past the line
// end and the others are loaded without a branch between them; RKM 0: no RK
// update, 1: RK update with a float64 nonlinear term, 2: float32.  Register
// budget of the branch-free forms: 2 waves per SIMD up to 9 points per lane
// (HD 512^3: 2.43 ms per launch against 2.81 at one wave), one wave for 17
// (HD 1024^3: 23.1 ms against 25.8 for the general form).
template <int KPL, int RKM>
__global__ void __launch_bounds__(256, RKM >= 0 && KPL <= 9 ? 2 : 1) k_project_harmonic(
    Cd* __restrict__ a, Cd* __restrict__ b, Cd* __restrict__ c, Cd* __restrict__ d,
    Cd* __restrict__ bc, long nlines, int nz, int ny, const double* __restrict__ kx,
    const double* __restrict__ ky, const double* __restrict__ kz, const Cd* __restrict__ wtop,
    double tmp, RkLoad rk, int rk_on, Cd* __restrict__ pw, int np, const Cd* __restrict__ tab,
    const int* __restrict__ hidx, double Lz, const double* __restrict__ khom) {
#pragma clang fp contract(off)
  const int lane = threadIdx.x & 63;
  const long ln = long(blockIdx.x) * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (ln >= nlines) return;  // uniform per wavefront
  const long i = ln / ny;
  const int j = int(ln - i * ny);
  const double kxi = kx[i], kyj = ky[j];
  // the RK update's factors of the line (rk_on_load's, read once)
  double rkl2 = 0.0, rfx = 0.0, rfy = 0.0, rnu = 0.0;
  if constexpr (RKM > 0) {
    const double rx = rk.kx[i], ry = rk.ky[j];
    rkl2 = rx * rx + ry * ry;
    rfx = rk.fkx[i];
    rfy = rk.fky[j];
    rnu = rk.c[0].nu[i];
  }
  Cd s0 = mkc<double>(0, 0), s1 = mkc<double>(0, 0);
  Cd q0 = mkc<double>(0, 0), q1 = mkc<double>(0, 0);
  Cd pa[KPL], pb[KPL], pc[KPL];
#pragma unroll
  for (int q = 0; q < KPL; ++q) {
    const int k = lane + 64 * q;
    if ((RKM >= 0 && q < KPL - 1) || k < nz) {
      const long e = ln * nz + k;
      const double kzk = kz[k];
      const double kk2 = (kxi * kxi + kyj * kyj) + kzk * kzk;
      const Cd A = a[e], B = b[e];
      Cd Cc = c[e];
      if constexpr (RKM < 0) {
        if (rk_on) Cc = rk_on_load(rk, 0, Cc, e, i, j, k);
      } else if constexpr (RKM > 0) {
        Cd nl;
        if constexpr (RKM == 2) {
          const cpx<float> t = static_cast<const cpx<float>*>(rk.c[0].nl)[e];
          nl = mkc<double>(double(t.x), double(t.y));
        } else {
          nl = static_cast<const Cd*>(rk.c[0].nl)[e];
        }
        Cc = rk_apply(rk, Cc, nl, static_cast<const Cd*>(rk.c[0].f)[e],
                      static_cast<const Cd*>(rk.c[0].c0)[e], rkl2, rk.kz[k], rfx, rfy, rk.fkz[k],
                      rnu);
      }
      Cd sm = cadd(cscale(A, kxi), cscale(B, kyj));
      sm = cadd(sm, cscale(Cc, kzk));
      Cd dd;
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
      pa[q] = csub(A, ikmul(kxi, dd));
      pb[q] = csub(B, ikmul(kyj, dd));
      pc[q] = csub(Cc, ikmul(kzk, dd));
      const Cd c1 = cscale(pc[q], tmp);
      s0 = cadd(s0, c1);
      s1 = cadd(s1, cmul(c1, wtop[k]));
    }
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
  // harmonic correction (the k_proj_harmonic_tab arithmetic)
  const double kh = khom[ln];
  const bool zm = kh == 0.0;
  const Coef cf = neumann_coef(s0, s1, kh, Lz, zm);
  const Cd* P1 = tab + 2 * long(hidx[ln]) * nz;
  const Cd* P2 = P1 + nz;
#pragma unroll
  for (int q = 0; q < KPL; ++q) {
    const int k = lane + 64 * q;
    if ((RKM >= 0 && q < KPL - 1) || k < nz) {
      const Cd p1 = P1[k], p2 = P2[k];
      Cd A, B;
      if (zm) {
        A = cadd(cscale(p1, cf.c1.x), cscale(p2, cf.c2.x));
        B = cscale(p2, cf.c1.x);
      } else {
        const Cd u1 = cmul(cf.c1, p1), u2 = cmul(cf.c2, p2);
        A = cadd(u1, u2);
        B = cscale(csub(u1, u2), kh);
      }
      const long o = ln * nz + k;
      a[o] = csub(pa[q], ikmul(kxi, A));
      b[o] = csub(pb[q], ikmul(kyj, A));
      c[o] = csub(pc[q], B);
    }
  }
}

// ---------------------------------------------------------------------------
// diagnostics: per-line float64 reductions over physical z in the mixed domain
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

template <typename TS>
struct LinCombDev {
  const cpx<TS>* in[3];
  int dir[3];
  int sgn[3];
};

// dirn 1..3: (i k_dirn) v ; 4: kk2 v with kk2 = (kx^2 + ky^2) + kz^2
__device__ __forceinline__ Cd deriv_apply(Cd v, int dirn, double kxi, double kyj, double kzk) {
#pragma clang fp contract(off)
  if (dirn == 1) return ikmul(kxi, v);
  if (dirn == 2) return ikmul(kyj, v);
  if (dirn == 3) return ikmul(kzk, v);
  if (dirn == 4) return cscale(v, (kxi * kxi + kyj * kyj) + kzk * kzk);
  return v;
}

// mode 0: sum_k<np |IFFT(lincomb)|^2 ; mode 1: Re(IFFT(a) conj IFFT(b)) ; mode 2: walls |.|^2 ;
// mode 3: the wall values IFFT(lincomb) at z = 0 and z = Lz (4 doubles per line)
template <typename TS, int RS>
__global__ void __launch_bounds__(kMaxBlock) k_zline_reduce(LinCombDev<TS> lc,
                                                      const cpx<TS>* __restrict__ bsrc, int mode,
                                                      double* __restrict__ out, LFd f,
                                                      long nlines, int ny, int np,
                                                      const double* __restrict__ kx,
                                                      const double* __restrict__ ky,
                                                      const double* __restrict__ kz) {
#pragma clang fp contract(off)
  SPX_PROLOGUE(Cd, f);
  const int n = f.n;
  Cd* const ldsB = lds + f.nl * f.ldl;
  Cd* const xb = ldsB + line * f.ldl;
  double* red = reinterpret_cast<double*>(ldsB + (mode == 1 ? f.nl * f.ldl : 0));
  const long ln = long(blockIdx.x) * f.nl + line;
  const bool ok = ln < nlines;
  const long i = ok ? ln / ny : 0;
  const int j = ok ? int(ln - i * ny) : 0;
  for (int k = lane; k < n; k += f.tpl) {
    Cd v = mkc<double>(0, 0), w = mkc<double>(0, 0);
    if (ok) {
      const double kxi = kx[i], kyj = ky[j], kzk = kz[k];
      v = deriv_apply(ccast<double>(lc.in[0][ln * n + k]), lc.dir[0], kxi, kyj, kzk);
      for (int t = 1; t < 3; ++t) {
        if (lc.sgn[t] == 0) continue;
        const Cd u = deriv_apply(ccast<double>(lc.in[t][ln * n + k]), lc.dir[t], kxi, kyj, kzk);
        v = lc.sgn[t] > 0 ? cadd(v, u) : csub(v, u);
      }
      if (mode == 1) w = ccast<double>(bsrc[ln * n + k]);
    }
    xl[k] = cconj(v);
    if (mode == 1) xb[k] = cconj(w);
  }
  __syncthreads();
  line_dft<double, RS>(xl, f, tws, lane);
  if (mode == 1) line_dft<double, RS>(xb, f, tws, lane);
  if (mode == 2) {
    if (lane == 0 && ok) {
      const Cd b = xl[0], t = xl[np - 1];
      out[2 * ln] = b.x * b.x + b.y * b.y;
      out[2 * ln + 1] = t.x * t.x + t.y * t.y;
    }
    return;
  }
  if (mode == 3) {
    // the line holds conj(IFFT)
    if (lane == 0 && ok) {
      const Cd b = xl[0], t = xl[np - 1];
      out[4 * ln] = b.x;
      out[4 * ln + 1] = -b.y;
      out[4 * ln + 2] = t.x;
      out[4 * ln + 3] = -t.y;
    }
    return;
  }
  double acc = 0.0;
  for (int k = lane; k < np; k += f.tpl) {
    const Cd v = xl[k];
    if (mode == 0) {
      acc += v.x * v.x + v.y * v.y;
    } else {
      const Cd w = xb[k];
      // values hold conj(IFFT); Re(p conj(q)) is invariant under conj of both
      acc += v.x * w.x + v.y * w.y;
    }
  }
  const double s = line_reduce(acc, red, line, lane, f.tpl);
  if (lane == 0 && ok) out[ln] = s;
}

#if !defined(SPX_PART) || SPX_PART == 1
// ---------------------------------------------------------------------------
// plan construction: algorithm choice, factorisation, long-double tables
// ---------------------------------------------------------------------------
static bool factor_small(int n, std::vector<int>& fac) {
  fac.clear();
  int a = 0;
  while (n % 2 == 0) {
    n /= 2;
    ++a;
  }
  // powers of two: prefer radix 8, then 4, avoid a lone 2 when possible
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
  int b = 0;
  while (n % 3 == 0) {
    n /= 3;
    ++b;
  }
  for (int i = 0; i < b / 2; ++i) fac.push_back(9);
  if (b % 2) fac.push_back(3);
  for (int p : {5, 7, 11, 13}) {
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

static bool is_prime(int n) {
  if (n < 2) return false;
  for (int d = 2; d * d <= n; ++d)
    if (n % d == 0) return false;
  return true;
}

static long long powmod(long long b, long long e, long long m) {
  long long r = 1 % m;
  b %= m;
  while (e > 0) {
    if (e & 1) r = r * b % m;
    b = b * b % m;
    e >>= 1;
  }
  return r;
}

static int primitive_root(int n) {
  std::vector<int> pf;
  int m = n - 1;
  for (int d = 2; d * d <= m; ++d)
    if (m % d == 0) {
      pf.push_back(d);
      while (m % d == 0) m /= d;
    }
  if (m > 1) pf.push_back(m);
  for (int g = 2; g < n; ++g) {
    bool ok = true;
    for (int p : pf)
      if (powmod(g, (n - 1) / p, n) == 1) {
        ok = false;
        break;
      }
    if (ok) return g;
  }
  return -1;
}

static int tpl_for(int M, int n, const std::vector<int>& fac) {
  int need = (n + KEEP - 1) / KEEP;
  for (int R : fac) {
    const int nb = M / R;
    const int mb = std::max(1, kStage / R);
    need = std::max(need, (nb + mb - 1) / mb);
  }
  int t = 1;
  while (t < need) t <<= 1;
  return t;
}

static int dev_upload(void** dst, const void* src, size_t bytes, size_t* acc) {
  if (hipMalloc(dst, bytes) != hipSuccess) return fail(SPX_ERR_ALLOC, "hipMalloc of FFT tables");
  *acc += bytes;
  if (hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice) != hipSuccess)
    return fail(SPX_ERR_HIP, "hipMemcpy of FFT tables");
  return SPX_OK;
}

// upload a complex table in the plan's precision (csize 16: float64, 8: float32)
static int upload_table(void** dst, const std::vector<double>& v, int csize, size_t* acc) {
  if (csize == 16) return dev_upload(dst, v.data(), v.size() * sizeof(double), acc);
  std::vector<float> f(v.size());
  for (size_t i = 0; i < v.size(); ++i) f[i] = float(v[i]);
  return dev_upload(dst, f.data(), f.size() * sizeof(float), acc);
}

// naive long-double DFT of length M (used once per plan for the filters)
static void dft_ld(const std::vector<long double>& xr, const std::vector<long double>& xi,
                   const std::vector<long double>& wr, const std::vector<long double>& wi,
                   std::vector<double>& out, long double scale) {
  const int M = int(xr.size());
  out.assign(2 * size_t(M), 0.0);
  for (int k = 0; k < M; ++k) {
    long double sr = 0, si = 0;
    for (int m = 0; m < M; ++m) {
      if (xr[m] == 0.0L && xi[m] == 0.0L) continue;
      const int idx = int((long long)m * k % M);
      sr += xr[m] * wr[idx] - xi[m] * wi[idx];
      si += xr[m] * wi[idx] + xi[m] * wr[idx];
    }
    out[2 * k] = double(sr * scale);
    out[2 * k + 1] = double(si * scale);
  }
}

// autotuner overrides (process-wide, read at plan creation)
static std::mutex g_tune_mu;
static std::map<std::pair<int, int>, LineTune> g_line_tune;
static std::map<int, WallTune> g_wall_tune;

void set_line_tune(int n, int csize, const LineTune& t) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  g_line_tune[{n, csize}] = t;
}
bool get_line_tune(int n, int csize, LineTune* t) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  auto it = g_line_tune.find({n, csize});
  if (it == g_line_tune.end()) return false;
  *t = it->second;
  return true;
}
void set_wall_tune(int n, const WallTune& t) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  g_wall_tune[n] = t;
}
bool get_wall_tune(int n, WallTune* t) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  auto it = g_wall_tune.find(n);
  if (it == g_wall_tune.end()) return false;
  *t = it->second;
  return true;
}
void clear_tunes() {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  g_line_tune.clear();
  g_wall_tune.clear();
}

int build_line_fft(LineFFTHost& f, int n, int csize, int want_lines, int max_block,
                   size_t lds_budget) {
  std::vector<int> fac;
  f.n = n;
  if (factor_small(n, fac)) {
    f.mode = 0;
    f.M = n;
  } else if (is_prime(n) && factor_small(n - 1, fac)) {
    f.mode = 2;
    f.M = n - 1;
  } else {
    f.mode = 1;
    int M = 2 * n - 1;
    while (!smooth7(M)) ++M;
    f.M = M;
    factor_small(M, fac);
  }
  if (int(fac.size()) > kMaxStages) return fail(SPX_ERR_UNSUPPORTED, "FFT length has too many factors");
  f.nstage = int(fac.size());
  long long Ns = 1;
  for (int i = 0; i < f.nstage; ++i) {
    f.radix[i] = fac[i];
    f.magic[i] = Ns > 1 ? unsigned(((1ULL << 32) + Ns - 1) / Ns) : 0u;
    Ns *= fac[i];
  }
  // at least 2 threads per line: the wall kernels handle z = 0 and z = Lz on
  // lanes 0 and 1
  f.tpl = std::max(2, tpl_for(f.M, n, fac));
  LineTune tune;
  const bool tuned = get_line_tune(n, csize, &tune);
  auto pow2 = [](int v) { return v > 0 && (v & (v - 1)) == 0; };
  if (tuned && pow2(tune.threads) && tune.threads >= f.tpl && tune.threads <= kMaxBlock)
    f.tpl = tune.threads;
  if (f.tpl > kMaxBlock)
    return fail(SPX_ERR_UNSUPPORTED, "FFT length " + std::to_string(n) + " exceeds the LDS engine");
  f.lgtpl = 0;
  while ((1 << f.lgtpl) < f.tpl) ++f.lgtpl;
  const int span = f.mode == 0 ? n : (f.mode == 1 ? f.M : 2 * n);
  f.ldl = span + ((span % 16 == 0) ? 1 : 0);
  // tables staged in LDS: twiddles; Bluestein filter + chirp; Rader filter + indices
  if (f.mode == 0) f.tab = f.M;
  else if (f.mode == 1) f.tab = 2 * f.M + n;
  else f.tab = 2 * f.M + int((2 * size_t(f.M) * sizeof(int) + csize - 1) / csize);
  int nl = std::max(1, max_block / f.tpl);
  if (want_lines > 0) nl = std::min(nl, want_lines);
  while (nl > 1 && (size_t(f.tab) + size_t(nl) * f.ldl) * csize > lds_budget) nl >>= 1;
  // Rader / Bluestein: leave the filter / chirp / index tables in global
  // memory (L1/L2-resident) when staging them in LDS would cost workgroups
  // per CU, counted up to 4 per CU (beyond that registers bound the
  // occupancy of these kernels); kernels for this are built for radices <= 9
  bool small_radices = true;
  for (int R : fac) small_radices = small_radices && R <= 9;
  if (f.mode != 0 && small_radices) {
    const size_t lds_cu = 160 * 1024, lines = size_t(nl) * f.ldl;
    auto wgs = [&](size_t tab) {
      return std::min<size_t>(4, lds_cu / ((tab + lines) * size_t(csize)));
    };
    if (wgs(size_t(f.M)) > wgs(size_t(f.tab))) f.tab = f.M;
    if (tuned && tune.gtab == 1) f.tab = f.M;
    if (tuned && tune.gtab == 0 && f.tab == f.M)
      f.tab = f.mode == 1 ? 2 * f.M + n
                          : 2 * f.M + int((2 * size_t(f.M) * sizeof(int) + csize - 1) / csize);
  }
  if (tuned && pow2(tune.lines) && tune.lines * f.tpl <= kMaxBlock &&
      (size_t(f.tab) + size_t(tune.lines) * f.ldl) * csize <= 160 * 1024)
    nl = tune.lines;
  if ((size_t(f.tab) + size_t(nl) * f.ldl) * csize > 160 * 1024)
    return fail(SPX_ERR_UNSUPPORTED, "FFT line does not fit in LDS");
  f.nl = nl;
  f.lgnl = 0;
  while ((1 << f.lgnl) < f.nl) ++f.lgnl;
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
  int rc = upload_table(&f.tw, tw, csize, &f.bytes);
  if (rc) return rc;
  if (f.mode == 1) {
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
    // b[m] = conj(w[|m|]) wrapped; bfil = DFT_M(b) / M
    std::vector<long double> br(M, 0.0L), bi(M, 0.0L);
    br[0] = cr[0];
    bi[0] = -ci[0];
    for (int m = 1; m < n; ++m) {
      br[m] = cr[m];
      bi[m] = -ci[m];
      br[M - m] = cr[m];
      bi[M - m] = -ci[m];
    }
    std::vector<double> bf;
    dft_ld(br, bi, wr, wi, bf, 1.0L / M);
    if ((rc = upload_table(&f.chirp, ch, csize, &f.bytes))) return rc;
    if ((rc = upload_table(&f.bfil, bf, csize, &f.bytes))) return rc;
  } else if (f.mode == 2) {
    // Rader: a_q = x[g^q], b_m = W_n^(g^-m), X[g^-p] = x0 + (a (*) b)_p
    const int g = primitive_root(n);
    if (g < 0) return fail(SPX_ERR_UNSUPPORTED, "no primitive root for Rader length");
    std::vector<int> perm(2 * size_t(M));
    const long long ginv = powmod(g, n - 2, n);
    for (int q = 0; q < M; ++q) {
      perm[q] = int(powmod(g, q, n));
      perm[M + q] = int(powmod(ginv, q, n));
    }
    std::vector<long double> br(M), bi(M);
    for (int m = 0; m < M; ++m) {
      const long double ang = PI2 * (long double)perm[M + m] / (long double)n;
      br[m] = cosl(ang);
      bi[m] = -sinl(ang);
    }
    std::vector<double> bf;
    dft_ld(br, bi, wr, wi, bf, 1.0L / M);
    if ((rc = upload_table(&f.bfil, bf, csize, &f.bytes))) return rc;
    if ((rc = dev_upload(&f.perm, perm.data(), perm.size() * sizeof(int), &f.bytes))) return rc;
  }
  return SPX_OK;
}

int build_wave_fft(LineFFTHost& f, int n, int csize) {
  f = LineFFTHost();
  if (const char* e = std::getenv("SPX_WAVE"))
    if (e[0] == '0') return SPX_OK;
  LineFFTHost h;
  if (int rc = build_line_fft(h, n, csize, 0)) {
    free_line_fft(h);
    return rc;
  }
  // the z launch macro lists plans 0..2 (plan 3 is the x / y lines' plan)
  const int w = wave_plan_for(h.n, h.M, h.mode, csize);
  if (w < 0 || w > 2) {
    free_line_fft(h);
    return SPX_OK;
  }
  // one wavefront per line, 8 lines per workgroup; twiddles staged in LDS,
  // the Bluestein filter and chirp read from global memory
  h.wave = w;
  h.tpl = 64;
  h.lgtpl = 6;
  h.nl = 8;
  h.lgnl = 3;
  h.ldl = h.M + 1;
  h.tab = h.M;
  if (h.lds_bytes(size_t(csize)) > 160 * 1024) {
    free_line_fft(h);
    return SPX_OK;
  }
  f = h;
  return SPX_OK;
}

void attach_xy_wave(LineFFTHost& f, int csize) {
  for (const char* v : {"SPX_WAVE", "SPX_WAVE_XY"})
    if (const char* e = std::getenv(v))
      if (e[0] == '0') return;
  if (f.mode == 0 && f.tpl == 64 && wave_plan_for(f.n, f.M, 0, csize) == 3) f.wave = 3;
}

void free_line_fft(LineFFTHost& f) {
  if (f.tw) (void)hipFree(f.tw);
  if (f.chirp) (void)hipFree(f.chirp);
  if (f.bfil) (void)hipFree(f.bfil);
  if (f.perm) (void)hipFree(f.perm);
  f.tw = f.chirp = f.bfil = f.perm = nullptr;
  f.bytes = 0;
}
#endif  // SPX_PART 1 (plan tables)

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

// radix set of a plan: the smallest compiled kernel family that covers it
// (Rader and Bluestein plans never use the power-of-two family)
[[maybe_unused]] static int rset_of(const LineFFTHost& f) {
  if (f.wave >= 0) return RS_WAVE + f.wave;
  if (f.mode != 0 && f.tab == f.M) return RS_GTAB;  // radices <= 9 (build_line_fft)
  int rs = f.mode == 0 ? RS_POW2 : RS_SMOOTH;
  for (int i = 0; i < f.nstage; ++i) {
    const int R = f.radix[i];
    if (R == 11 || R == 13) rs = RS_ALL;
    else if (R != 2 && R != 4 && R != 8 && rs < RS_SMOOTH) rs = RS_SMOOTH;
  }
  return rs;
}

#define SPX_COMMA ,
// launch the instance KER<TY, RS TARGS> compiled for the plan's radix set
#define SPX_LAUNCH_RS(F, KER, TY, TARGS, GRID, BLOCK, LDS, STREAM, ...)                      \
  do {                                                                                       \
    const int rs_ = rset_of(F);                                                              \
    if (rs_ == RS_POW2) {                                                                    \
      if (int rc_ = lds_limit(KER<TY SPX_COMMA RS_POW2 TARGS>, (LDS))) return rc_;           \
      hipLaunchKernelGGL((KER<TY SPX_COMMA RS_POW2 TARGS>), GRID, BLOCK, LDS, STREAM,        \
                         __VA_ARGS__);                                                       \
    } else if (rs_ == RS_SMOOTH) {                                                           \
      if (int rc_ = lds_limit(KER<TY SPX_COMMA RS_SMOOTH TARGS>, (LDS))) return rc_;         \
      hipLaunchKernelGGL((KER<TY SPX_COMMA RS_SMOOTH TARGS>), GRID, BLOCK, LDS, STREAM,      \
                         __VA_ARGS__);                                                       \
    } else if (rs_ == RS_ALL) {                                                              \
      if (int rc_ = lds_limit(KER<TY SPX_COMMA RS_ALL TARGS>, (LDS))) return rc_;            \
      hipLaunchKernelGGL((KER<TY SPX_COMMA RS_ALL TARGS>), GRID, BLOCK, LDS, STREAM,         \
                         __VA_ARGS__);                                                       \
    } else if (rs_ == RS_GTAB) {                                                             \
      if (int rc_ = lds_limit(KER<TY SPX_COMMA RS_GTAB TARGS>, (LDS))) return rc_;           \
      hipLaunchKernelGGL((KER<TY SPX_COMMA RS_GTAB TARGS>), GRID, BLOCK, LDS, STREAM,        \
                         __VA_ARGS__);                                                       \
    } else {                                                                                 \
      return fail(SPX_ERR_STATE, "wave line plan passed to a general kernel");               \
    }                                                                                        \
  } while (0)

// x / y transform launch: the 512-point wave plan (RS_WAVE + 3) or the
// general families (spelled out: TARGS cannot be forwarded to another macro)
#define SPX_LAUNCH_XY(F, KER, TY, TARGS, GRID, BLOCK, LDS, STREAM, ...)                      \
  do {                                                                                       \
    const int rs_ = rset_of(F);                                                              \
    if (rs_ == RS_WAVE + 3) {                                                                \
      if (int rc_ = lds_limit(KER<TY SPX_COMMA RS_WAVE + 3 TARGS>, (LDS))) return rc_;       \
      hipLaunchKernelGGL((KER<TY SPX_COMMA RS_WAVE + 3 TARGS>), GRID, BLOCK, LDS, STREAM,    \
                         __VA_ARGS__);                                                       \
    } else if (rs_ == RS_POW2) {                                                             \
      if (int rc_ = lds_limit(KER<TY SPX_COMMA RS_POW2 TARGS>, (LDS))) return rc_;           \
      hipLaunchKernelGGL((KER<TY SPX_COMMA RS_POW2 TARGS>), GRID, BLOCK, LDS, STREAM,        \
                         __VA_ARGS__);                                                       \
    } else if (rs_ == RS_SMOOTH) {                                                           \
      if (int rc_ = lds_limit(KER<TY SPX_COMMA RS_SMOOTH TARGS>, (LDS))) return rc_;         \
      hipLaunchKernelGGL((KER<TY SPX_COMMA RS_SMOOTH TARGS>), GRID, BLOCK, LDS, STREAM,      \
                         __VA_ARGS__);                                                       \
    } else if (rs_ == RS_ALL) {                                                              \
      if (int rc_ = lds_limit(KER<TY SPX_COMMA RS_ALL TARGS>, (LDS))) return rc_;            \
      hipLaunchKernelGGL((KER<TY SPX_COMMA RS_ALL TARGS>), GRID, BLOCK, LDS, STREAM,         \
                         __VA_ARGS__);                                                       \
    } else if (rs_ == RS_GTAB) {                                                             \
      if (int rc_ = lds_limit(KER<TY SPX_COMMA RS_GTAB TARGS>, (LDS))) return rc_;           \
      hipLaunchKernelGGL((KER<TY SPX_COMMA RS_GTAB TARGS>), GRID, BLOCK, LDS, STREAM,        \
                         __VA_ARGS__);                                                       \
    } else {                                                                                 \
      return fail(SPX_ERR_STATE, "x/y line plan without a compiled kernel family");          \
    }                                                                                        \
  } while (0)

// z-transform launch: the wave-plan instance when the plan has one, else as
// SPX_LAUNCH_RS (spelled out: TARGS cannot be forwarded to another macro)
#define SPX_LAUNCH_Z(F, KER, TY, TARGS, GRID, BLOCK, LDS, STREAM, ...)                       \
  do {                                                                                       \
    const int rs_ = rset_of(F);                                                              \
    if (rs_ == RS_WAVE + 0) {                                                                \
      if (int rc_ = lds_limit(KER<TY SPX_COMMA RS_WAVE + 0 TARGS>, (LDS))) return rc_;       \
      hipLaunchKernelGGL((KER<TY SPX_COMMA RS_WAVE + 0 TARGS>), GRID, BLOCK, LDS, STREAM,    \
                         __VA_ARGS__);                                                       \
    } else if (rs_ == RS_WAVE + 1) {                                                         \
      if (int rc_ = lds_limit(KER<TY SPX_COMMA RS_WAVE + 1 TARGS>, (LDS))) return rc_;       \
      hipLaunchKernelGGL((KER<TY SPX_COMMA RS_WAVE + 1 TARGS>), GRID, BLOCK, LDS, STREAM,    \
                         __VA_ARGS__);                                                       \
    } else if (rs_ == RS_WAVE + 2) {                                                         \
      if (int rc_ = lds_limit(KER<TY SPX_COMMA RS_WAVE + 2 TARGS>, (LDS))) return rc_;       \
      hipLaunchKernelGGL((KER<TY SPX_COMMA RS_WAVE + 2 TARGS>), GRID, BLOCK, LDS, STREAM,    \
                         __VA_ARGS__);                                                       \
    } else if (rs_ == RS_WAVE + 3) {                                                         \
      if (int rc_ = lds_limit(KER<TY SPX_COMMA RS_WAVE + 3 TARGS>, (LDS))) return rc_;       \
      hipLaunchKernelGGL((KER<TY SPX_COMMA RS_WAVE + 3 TARGS>), GRID, BLOCK, LDS, STREAM,    \
                         __VA_ARGS__);                                                       \
    } else if (rs_ == RS_POW2) {                                                             \
      if (int rc_ = lds_limit(KER<TY SPX_COMMA RS_POW2 TARGS>, (LDS))) return rc_;           \
      hipLaunchKernelGGL((KER<TY SPX_COMMA RS_POW2 TARGS>), GRID, BLOCK, LDS, STREAM,        \
                         __VA_ARGS__);                                                       \
    } else if (rs_ == RS_SMOOTH) {                                                           \
      if (int rc_ = lds_limit(KER<TY SPX_COMMA RS_SMOOTH TARGS>, (LDS))) return rc_;         \
      hipLaunchKernelGGL((KER<TY SPX_COMMA RS_SMOOTH TARGS>), GRID, BLOCK, LDS, STREAM,      \
                         __VA_ARGS__);                                                       \
    } else if (rs_ == RS_ALL) {                                                              \
      if (int rc_ = lds_limit(KER<TY SPX_COMMA RS_ALL TARGS>, (LDS))) return rc_;            \
      hipLaunchKernelGGL((KER<TY SPX_COMMA RS_ALL TARGS>), GRID, BLOCK, LDS, STREAM,         \
                         __VA_ARGS__);                                                       \
    } else if (rs_ == RS_GTAB) {                                                             \
      if (int rc_ = lds_limit(KER<TY SPX_COMMA RS_GTAB TARGS>, (LDS))) return rc_;           \
      hipLaunchKernelGGL((KER<TY SPX_COMMA RS_GTAB TARGS>), GRID, BLOCK, LDS, STREAM,        \
                         __VA_ARGS__);                                                       \
    } else {                                                                                 \
      return fail(SPX_ERR_STATE, "z line plan without a compiled kernel family");            \
    }                                                                                        \
  } while (0)

template <typename T>
static inline cpx<T>* C_(void* p) { return static_cast<cpx<T>*>(p); }
template <typename T>
static inline const cpx<T>* C_(const void* p) { return static_cast<const cpx<T>*>(p); }

// ---- transforms ----
// lo selects the float32 kernels and tables of a mixed plan.
static inline const LineFFTHost& fxp(const spx_plan* p, bool lo) { return lo ? p->fx32 : p->fx; }
static inline const LineFFTHost& fyp(const spx_plan* p, bool lo) { return lo ? p->fy32 : p->fy; }
static inline const LineFFTHost& fzp(const spx_plan* p, bool lo) { return lo ? p->fz32 : p->fz; }
// the z-transform plan: the wave variant where one is compiled
static inline const LineFFTHost& fzt(const spx_plan* p, bool lo) {
  if (lo) return p->fzw32.wave >= 0 ? p->fzw32 : p->fz32;
  return p->fzw.wave >= 0 ? p->fzw : p->fz;
}

[[maybe_unused]] static int check_lo(const spx_plan* p, bool lo) {
  if (lo && !p->fp32) return fail(SPX_ERR_STATE, "float32 kernels need a mixed-precision plan");
  return SPX_OK;
}

#if !defined(SPX_PART) || SPX_PART == 1
int zfft_inplace(spx_plan* p, void* S, int forward, int cont, double scale) {
  const LineFFTHost& f = fzt(p, false);
  const long nlines = p->wlines();
  const unsigned nb = unsigned((nlines + f.nl - 1) / f.nl);
  const int C = p->g.Cz, d = p->g.oz;
  ProfScope ps_(p, "z_inplace", double(p->wnc()) * 2.0 * 16.0);
  SPX_LAUNCH_Z(f, k_z_inplace, double, , dim3(nb), blk(f), f.lds_bytes(16), p->stream,
                C_<double>(S), nlines, f.dev<Cd>(), p->np, forward, (cont && C > 0) ? 1 : 0,
                scale, p->d_dir, C, d);
  SPX_CHECK_LAUNCH();
  return SPX_OK;
}

int fc_continue_only(spx_plan* p, void* S) {
  const long nlines = p->wlines();
  const int C = p->g.Cz;
  if (C <= 0) return SPX_OK;
  const long tot = nlines * C;
  ProfScope ps_(p, "fc_continue", double(nlines) * 16.0 * (2 * p->g.oz + C));
  hipLaunchKernelGGL(k_fc_continue<double>, dim3(unsigned((tot + 255) / 256)), dim3(256), 0,
                     p->stream, C_<double>(S), nlines, p->g.nz, p->np, C, p->g.oz, p->d_dir);
  SPX_CHECK_LAUNCH();
  return SPX_OK;
}

template <typename T>
static int zfft_x2s_t(spx_plan* p, const void* X, void* S, int nofft, const LineFFTHost& f,
                      const void* Xb) {
  // members: grid.z; X holds np planes per member (all nz with nofft)
  dim3 grid(unsigned((p->nkx + f.nl - 1) / f.nl), unsigned(p->g.ny), unsigned(p->mem));
  const long xstride = long(p->nkx) * p->g.ny * (nofft ? p->g.nz : p->np);
  ProfScope ps_(p, "z_x2s", double(p->wkx()) * p->g.ny * sizeof(cpx<T>) *
                                ((nofft ? p->g.nz : p->np) + p->g.nz));
  if (Xb) {
    // plus the staged float64 boundary values of the tile (nl * 2d, aligned)
    const size_t bnd_stage_bytes = size_t(f.nl) * 2 * p->g.oz * sizeof(Cd) + 16;
    SPX_LAUNCH_Z(f, k_z_x2s, T, SPX_COMMA true, grid, blk(f),
                  f.lds_bytes(sizeof(cpx<T>)) + bnd_stage_bytes, p->stream, C_<T>(X), C_<T>(S),
                  f.dev<cpx<T>>(), p->nkx, p->g.ny, p->np, nofft, p->d_dir, p->g.Cz, p->g.oz,
                  xstride, C_<double>(Xb));
    SPX_CHECK_LAUNCH();
    return SPX_OK;
  }
  SPX_LAUNCH_Z(f, k_z_x2s, T, SPX_COMMA false, grid, blk(f), f.lds_bytes(sizeof(cpx<T>)),
                p->stream, C_<T>(X), C_<T>(S), f.dev<cpx<T>>(), p->nkx, p->g.ny, p->np, nofft, p->d_dir,
                p->g.Cz, p->g.oz, xstride, static_cast<const Cd*>(nullptr));
  SPX_CHECK_LAUNCH();
  return SPX_OK;
}
int zfft_x2s(spx_plan* p, const void* X, void* S, int nofft, bool lo, const void* Xb) {
  if (int rc = check_lo(p, lo)) return rc;
  if (Xb && (nofft || p->g.Cz <= 0)) return fail(SPX_ERR_ARG, "boundary planes need a continuation");
  return lo ? zfft_x2s_t<float>(p, X, S, nofft, fzt(p, true), Xb)
            : zfft_x2s_t<double>(p, X, S, nofft, fzt(p, false), Xb);
}

int zfft_s2x_bnd(spx_plan* p, const void* const* S, int ncomp, void* XVb, void* XZb) {
  if (!p->d_wbnd) return fail(SPX_ERR_STATE, "plan has no boundary-plane table");
  if (ncomp < 1 || ncomp > 4) return fail(SPX_ERR_ARG, "zfft_s2x_bnd: 1..4 components");
  BndSrc src;
  for (int c = 0; c < 4; ++c) src.s[c] = c < ncomp ? C_<double>(S[c]) : nullptr;
  const long lines = p->wlines();
  const dim3 grid(unsigned((lines + 15) / 16), unsigned(ncomp));
  ProfScope ps_(p, "z_s2x", double(ncomp) * double(p->wnc()) * 16.0);
  const int n = p->g.nz, nkx = p->nkx, ny = p->g.ny, mem = p->mem;
  const Cd* w1 = static_cast<const Cd*>(p->d_wbnd);
  const Cd* wt = w1 + n;
  Cd* v = C_<double>(XVb);
  Cd* z = C_<double>(XZb);
  switch (p->g.oz) {
#define SPX_BND_CASE(D)                                                                        \
  case D:                                                                                      \
    hipLaunchKernelGGL((k_z_s2x_bnd<D, 2>), grid, dim3(256), 0, p->stream, src, v, z, w1, wt, n, \
                       nkx, ny, mem, p->d_kz);                                                 \
    break;
    SPX_BND_CASE(1)
    SPX_BND_CASE(2)
    SPX_BND_CASE(3)
    SPX_BND_CASE(4)
    SPX_BND_CASE(5)
    SPX_BND_CASE(6)
    SPX_BND_CASE(7)
    SPX_BND_CASE(8)
#undef SPX_BND_CASE
    default:
      return fail(SPX_ERR_UNSUPPORTED, "boundary planes: matching order above 8");
  }
  SPX_CHECK_LAUNCH();
  return SPX_OK;
}

// S is float64 in both precisions; X (and X2) are T
template <typename T>
static int zfft_s2x_t(spx_plan* p, const void* S, void* X, int dirn, int nzout, int nofft,
                      void* X2, const LineFFTHost& f) {
  // members: grid.z; X holds nzout planes per member
  dim3 grid(unsigned((p->nkx + f.nl - 1) / f.nl), unsigned(p->g.ny), unsigned(p->mem));
  const double line = double(p->wkx()) * p->g.ny;
  const Cd* in = C_<double>(S);
  if (X2) {
    ProfScope ps_(p, "z_s2x_dual", line * (16.0 * p->g.nz + 2.0 * sizeof(cpx<T>) * nzout));
    SPX_LAUNCH_Z(f, k_z_s2x, T, SPX_COMMA true SPX_COMMA double, grid, blk(f),
                  f.lds_bytes(sizeof(cpx<T>)), p->stream, in, C_<T>(X), C_<T>(X2),
                  f.dev<cpx<T>>(), p->nkx, p->g.ny, nzout, dirn, nofft, p->d_kx, p->d_ky,
                  p->d_kz);
  } else {
    ProfScope ps_(p, "z_s2x", line * (16.0 * p->g.nz + sizeof(cpx<T>) * nzout));
    SPX_LAUNCH_Z(f, k_z_s2x, T, SPX_COMMA false SPX_COMMA double, grid, blk(f),
                  f.lds_bytes(sizeof(cpx<T>)), p->stream, in, C_<T>(X),
                  static_cast<cpx<T>*>(nullptr), f.dev<cpx<T>>(), p->nkx, p->g.ny, nzout, dirn,
                  nofft, p->d_kx, p->d_ky, p->d_kz);
  }
  SPX_CHECK_LAUNCH();
  return SPX_OK;
}
int zfft_s2x(spx_plan* p, const void* S, void* X, int dirn, int nzout, int nofft, void* X2,
             bool lo) {
  if (int rc = check_lo(p, lo)) return rc;
  return lo ? zfft_s2x_t<float>(p, S, X, dirn, nzout, nofft, X2, fzt(p, true))
            : zfft_s2x_t<double>(p, S, X, dirn, nzout, nofft, X2, fzt(p, false));
}
#endif  // SPX_PART 1 (z transforms)

#if !defined(SPX_PART) || SPX_PART == 2
template <typename T>
static int yfft_t(spx_plan* p, void* X, int forward, int nplanes, const void* Xin, void* Y2,
                  const LineFFTHost& f, bool cont) {
  // members are consecutive blocks of nplanes planes
  const int npl = nplanes * p->mem;
  dim3 grid(unsigned((p->nkx + f.nl - 1) / f.nl), unsigned(npl));
  const cpx<T>* in = C_<T>(Xin ? Xin : static_cast<const void*>(X));
  const double io = double(p->nkx) * p->g.ny * npl * sizeof(cpx<T>);
  if (Y2) {
    if (forward) return fail(SPX_ERR_ARG, "dual y transform is backward only");
    ProfScope ps_(p, "y_dual", io * 3.0);
    SPX_LAUNCH_XY(f, k_y, T, SPX_COMMA true, grid, blk(f), f.lds_bytes(sizeof(cpx<T>)),
                  p->stream, in, C_<T>(X), C_<T>(Y2), f.dev<cpx<T>>(), p->nkx, p->g.ny, forward,
                  p->d_ky, p->npy, 0, 0, static_cast<const double*>(nullptr));
  } else {
    ProfScope ps_(p, "y", io * 2.0);
    SPX_LAUNCH_XY(f, k_y, T, SPX_COMMA false, grid, blk(f), f.lds_bytes(sizeof(cpx<T>)),
                  p->stream, in, C_<T>(X), static_cast<cpx<T>*>(nullptr), f.dev<cpx<T>>(),
                  p->nkx, p->g.ny, forward, p->d_ky, p->npy, cont ? p->Cy : 0, p->oy,
                  p->d_diry);
  }
  SPX_CHECK_LAUNCH();
  return SPX_OK;
}
int yfft(spx_plan* p, void* X, int forward, int nplanes, const void* Xin, void* Y2, bool lo,
         bool cont) {
  if (int rc = check_lo(p, lo)) return rc;
  if (cont && (!forward || p->Cy <= 0 || Y2))
    return fail(SPX_ERR_ARG, "y continuation: forward transforms of plans with Cy > 0 only");
  return lo ? yfft_t<float>(p, X, forward, nplanes, Xin, Y2, p->fy32, cont)
            : yfft_t<double>(p, X, forward, nplanes, Xin, Y2, p->fy, cont);
}

template <typename T>
static int xr2c_t(spx_plan* p, const void* R, void* X, int nplanes, const LineFFTHost& f) {
  const long npairs = long(p->g.ny) * nplanes * p->mem / 2;
  dim3 grid(unsigned((npairs + f.nl - 1) / f.nl));
  ProfScope ps_(p, "xr2c", double(p->g.ny) * nplanes * p->mem *
                               (double(sizeof(T)) * p->g.nx + double(sizeof(cpx<T>)) * p->nkx));
  SPX_LAUNCH_XY(f, k_xr2c, T, , grid, blk(f), f.lds_bytes(sizeof(cpx<T>)), p->stream,
                static_cast<const T*>(R), C_<T>(X), f.dev<cpx<T>>(), p->nkx, npairs);
  SPX_CHECK_LAUNCH();
  return SPX_OK;
}
int xr2c(spx_plan* p, const void* R, void* X, int nplanes, bool lo) {
  if (int rc = check_lo(p, lo)) return rc;
  return lo ? xr2c_t<float>(p, R, X, nplanes, p->fx32) : xr2c_t<double>(p, R, X, nplanes, p->fx);
}

template <typename T>
static int xc2r_t(spx_plan* p, const void* X, void* R, int nplanes, int dirx,
                  const LineFFTHost& f) {
  const long npairs = long(p->g.ny) * nplanes * p->mem / 2;
  dim3 grid(unsigned((npairs + f.nl - 1) / f.nl));
  ProfScope ps_(p, "xc2r", double(p->g.ny) * nplanes * p->mem *
                               (double(sizeof(T)) * p->g.nx + double(sizeof(cpx<T>)) * p->nkx));
  SPX_LAUNCH_XY(f, k_xc2r, T, , grid, blk(f), f.lds_bytes(sizeof(cpx<T>)), p->stream,
                C_<T>(X), static_cast<T*>(R), f.dev<cpx<T>>(), p->nkx, npairs, dirx, p->d_kx);
  SPX_CHECK_LAUNCH();
  return SPX_OK;
}
int xc2r(spx_plan* p, const void* X, void* R, int nplanes, int dirx, bool lo) {
  if (int rc = check_lo(p, lo)) return rc;
  return lo ? xc2r_t<float>(p, X, R, nplanes, dirx, p->fx32)
            : xc2r_t<double>(p, X, R, nplanes, dirx, p->fx);
}

template <typename T>
static int xgrad_prod_t(spx_plan* p, const void* YV, const void* YY, const void* XZ,
                        const void* u, const void* v, const void* w, void* out,
                        const LineFFTHost& f, int nplanes, int ncomp) {
  const long npairs = long(p->g.ny) * nplanes * p->mem / 2;
  dim3 grid(unsigned((npairs + f.nl - 1) / f.nl), unsigned(ncomp));
  const long cstride = long(p->nkx) * p->g.ny * nplanes * p->mem;
  const long rstride = long(p->g.nx) * p->g.ny * nplanes * p->mem;
  const double t = double(p->g.nx) * double(p->g.ny) * double(p->g.nz);
  const double rows = double(p->g.ny) * nplanes * p->mem * ncomp;
  ProfScope ps_(p, "xgrad_prod", rows * (3.0 * sizeof(cpx<T>) * p->nkx +
                                         4.0 * sizeof(T) * p->g.nx));
  SPX_LAUNCH_XY(f, k_xgrad_prod, T, , grid, blk(f), f.lds_bytes(sizeof(cpx<T>)), p->stream,
                C_<T>(YV), C_<T>(YY), C_<T>(XZ), static_cast<const T*>(u),
                static_cast<const T*>(v), static_cast<const T*>(w), static_cast<T*>(out),
                f.dev<cpx<T>>(), p->nkx, npairs, 1.0 / (t * t), p->d_kx, cstride, rstride);
  SPX_CHECK_LAUNCH();
  return SPX_OK;
}
template <typename T>
static int xgrad_fused_t(spx_plan* p, void* const* YV, void* const* YY, void* const* XZ,
                         void* const* RV, void* const* out, const LineFFTHost& f) {
  const int nplanes = p->np;
  const long npairs = long(p->g.ny) * nplanes * p->mem / 2;
  const double t = double(p->g.nx) * double(p->g.ny) * double(p->g.nz);
  const double rows = double(p->g.ny) * nplanes * p->mem;
  XGradArgs<T> a;
  for (int c = 0; c < 3; ++c) {
    a.yv[c] = C_<T>(static_cast<const void*>(YV[c]));
    a.yy[c] = C_<T>(static_cast<const void*>(YY[c]));
    a.xz[c] = C_<T>(static_cast<const void*>(XZ[c]));
    a.rv[c] = RV ? static_cast<T*>(RV[c]) : nullptr;
    a.out[c] = static_cast<T*>(out[c]);
  }
  ProfScope ps_(p, "xgrad_fused", rows * (9.0 * sizeof(cpx<T>) * p->nkx +
                                          (RV ? 6.0 : 3.0) * sizeof(T) * p->g.nx));
  if (f.wave == 3)
    hipLaunchKernelGGL((k_xgrad_fused<T, RS_WAVE + 3>), dim3(unsigned((npairs + f.nl - 1) / f.nl)),
                       blk(f), f.lds_bytes(sizeof(cpx<T>)), p->stream, a, f.dev<cpx<T>>(),
                       p->nkx, npairs, 1.0 / (t * t), p->d_kx);
  else
    hipLaunchKernelGGL((k_xgrad_fused<T, RS_POW2>), dim3(unsigned((npairs + f.nl - 1) / f.nl)),
                       blk(f), f.lds_bytes(sizeof(cpx<T>)), p->stream, a, f.dev<cpx<T>>(),
                       p->nkx, npairs, 1.0 / (t * t), p->d_kx);
  SPX_CHECK_LAUNCH();
  return SPX_OK;
}
int xgrad_fused(spx_plan* p, void* const* YV, void* const* YY, void* const* XZ, void* const* RV,
                void* const* out, bool lo) {
  if (int rc = check_lo(p, lo)) return rc;
  // power-of-two x lines only: the other radix families spill the velocity
  // registers
  const int rs = rset_of(lo ? p->fx32 : p->fx);
  if (rs != RS_POW2 && rs != RS_WAVE + 3) return SPX_ERR_UNSUPPORTED;
  return lo ? xgrad_fused_t<float>(p, YV, YY, XZ, RV, out, p->fx32)
            : xgrad_fused_t<double>(p, YV, YY, XZ, RV, out, p->fx);
}

int xgrad_prod(spx_plan* p, const void* YV, const void* YY, const void* XZ, const void* u,
               const void* v, const void* w, void* out, bool lo, int nplanes, int ncomp) {
  if (int rc = check_lo(p, lo)) return rc;
  if (nplanes < 0) nplanes = p->np;
  return lo ? xgrad_prod_t<float>(p, YV, YY, XZ, u, v, w, out, p->fx32, nplanes, ncomp)
            : xgrad_prod_t<double>(p, YV, YY, XZ, u, v, w, out, p->fx, nplanes, ncomp);
}
#endif  // SPX_PART 2 (y and x transforms)

#if !defined(SPX_PART) || SPX_PART == 3
// ---- float64 boundary / projection / diagnostics kernels ----
static int noslip_launch(spx_plan* p, void* vx, void* vy, const void* pr, const WallParams& wp,
                         int ncomp, const char* name, const RkLoad* rk) {
  if (p->wall.on) return wall_mfma(p, vx, vy, pr, wp, ncomp, rk);
  const LineFFTHost& f = p->fz;
  const long nlines = p->wlines();
  dim3 grid(unsigned((nlines + f.nl - 1) / f.nl), unsigned(ncomp));
  RkLoad r;
  std::memset(&r, 0, sizeof(r));
  if (rk) r = *rk;
  const double extra = rk ? (rk->nl32 ? 8.0 : 16.0) + 32.0 : 0.0;  // nl, c0, f reads
  ProfScope ps_(p, rk ? "noslip_rk" : name,
                ncomp * double(p->wnc()) * (2.0 * 16.0 + extra));
  SPX_LAUNCH_RS(f, k_noslip, double, , grid, blk(f), f.lds_bytes(16), p->stream, C_<double>(vx),
                C_<double>(vy), C_<double>(pr), f.dev<Cd>(), nlines, p->g.ny, p->np,
                1.0 / double(p->g.nz), wp, p->d_kx, p->d_ky, p->d_dir, p->g.Cz, p->g.oz, r,
                rk ? 1 : 0);
  SPX_CHECK_LAUNCH();
  return SPX_OK;
}

int noslip_fused(spx_plan* p, void* vx, void* vy, const void* pr, double tau,
                 const double vzsta[2], const double vzend[2], const RkLoad* rk) {
  WallParams wp;
  wp.tau = tau;
  wp.vb[0][0] = vzsta ? vzsta[0] : 0.0;
  wp.vb[1][0] = vzsta ? vzsta[1] : 0.0;
  wp.vb[0][1] = vzend ? vzend[0] : 0.0;
  wp.vb[1][1] = vzend ? vzend[1] : 0.0;
  wp.nxny = double(p->g.nx * p->g.ny);
  wp.scalar = 0;
  return noslip_launch(p, vx, vy, pr, wp, 2, "noslip", rk);
}

int bwall_fused(spx_plan* p, void* a0, void* a1, void* a2, int ncomp, const BWallParams& bp) {
  if (p->g.oz > 16) return fail(SPX_ERR_UNSUPPORTED, "magnetic walls: matching order above 16");
  const LineFFTHost& f = p->fz;
  const long nlines = p->wlines();
  dim3 grid(unsigned((nlines + f.nl - 1) / f.nl), unsigned(ncomp));
  ProfScope ps_(p, "bwall", ncomp * double(p->wnc()) * 2.0 * 16.0);
  SPX_LAUNCH_RS(f, k_bwall, double, , grid, blk(f), f.lds_bytes(16), p->stream, C_<double>(a0),
                C_<double>(a1), C_<double>(a2), f.dev<Cd>(), nlines, p->np,
                1.0 / double(p->g.nz), bp, p->d_khom, p->d_dir, p->g.Cz, p->g.oz);
  SPX_CHECK_LAUNCH();
  return SPX_OK;
}

int sconst_fused(spx_plan* p, void* th) {
  WallParams wp;
  std::memset(&wp, 0, sizeof(wp));
  wp.scalar = 1;
  return noslip_launch(p, th, th, nullptr, wp, 1, "sconst", nullptr);
}

int proj_pressure(spx_plan* p, void* d, int k0, int k1) {
  const LineFFTHost& f = fzt(p, false);
  const long nlines = p->wlines();
  ProfScope ps_(p, "proj_pressure", 2.0 * double(p->wnc()) * 16.0);
  const dim3 grid(unsigned((nlines + f.nl - 1) / f.nl));
  const size_t lds = f.lds_bytes(16);
  SPX_LAUNCH_Z(f, k_proj_pressure, double, , grid, blk(f), lds, p->stream, C_<double>(d),
               C_<double>(p->bc), f.dev<Cd>(), nlines, p->g.ny, 1.0 / double(p->g.nz), p->g.Lz,
               p->d_khom, p->d_z, k0, k1);
  SPX_CHECK_LAUNCH();
  return SPX_OK;
}

int pressure_walls(spx_plan* p, void* d) {
  const long nlines = p->wlines();
  ProfScope ps_(p, "pressure_walls", double(nlines) * 6.0 * 16.0);
  hipLaunchKernelGGL(k_pressure_walls, dim3(unsigned((nlines + 255) / 256)), dim3(256), 0,
                     p->stream, C_<double>(d), C_<double>(p->bc), C_<double>(p->pw), nlines,
                     p->g.nz, p->np, p->g.Lz, p->d_khom, p->d_z);
  SPX_CHECK_LAUNCH();
  return SPX_OK;
}

// The two harmonic profiles of each distinct khom value, continued and
// transformed on the host in long double and rounded once to float64: rows
// 2u, 2u+1 of tab.  The table is reused by every line with that kh, every
// stage and every step, so its rounding acts as a systematic error of the
// projection; a float64 device FFT leaves several ulps in each entry and
// moves cfg1 (64^3, 100 steps) 4x further from the exact trajectory than
// per-line transforms do, a correctly rounded table does not.  Real input:
// P[nz-k] = conj(P[k]).  Threads over the profiles.
static void harmonic_table_ld(const std::vector<double>& khv, int n, int np, int C, int d,
                              const std::vector<double>& z, double Lz,
                              const std::vector<double>& dir, std::vector<double>& out) {
  const long double PI2 = 6.283185307179586476925286766559005768L;
  std::vector<long double> wr(n), wi(n);
  for (int m = 0; m < n; ++m) {
    wr[m] = cosl(PI2 * (long double)m / (long double)n);
    wi[m] = -sinl(PI2 * (long double)m / (long double)n);
  }
  const int U = int(khv.size());
  out.assign(size_t(2) * U * n * 2, 0.0);
  auto work = [&](int u0, int u1) {
    std::vector<long double> x(n);
    for (int u = u0; u < u1; ++u) {
      const long double kh = khv[u];
      for (int which = 0; which < 2; ++which) {
        for (int k = 0; k < np; ++k) {
          const long double zk = z[k];
          if (kh == 0.0L) x[k] = which == 0 ? zk : 1.0L;
          else x[k] = which == 0 ? expl(kh * (zk - (long double)Lz)) : expl(-kh * zk);
        }
        // FC-Gram continuation (fftp.fpp:755-773 order)
        const int top = np - d;
        for (int ii = 0; ii < C; ++ii) {
          const double* D0 = dir.data() + size_t(ii) * d;
          const double* D1 = dir.data() + size_t(C - 1 - ii) * d;
          long double a = (long double)D0[0] * x[top] + (long double)D1[0] * x[d - 1];
          for (int jj = 1; jj < d; ++jj) {
            a += (long double)D0[jj] * x[top + jj];
            a += (long double)D1[jj] * x[d - 1 - jj];
          }
          x[np + ii] = a;
        }
        double* o = out.data() + (size_t(2) * u + which) * n * 2;
        for (int k = 0; k <= n / 2; ++k) {
          long double sr = 0, si = 0;
          int idx = 0;
          for (int m = 0; m < n; ++m) {
            sr += x[m] * wr[idx];
            si += x[m] * wi[idx];
            idx += k;
            if (idx >= n) idx -= n;
          }
          o[2 * k] = double(sr);
          o[2 * k + 1] = double(si);
          if (k > 0 && n - k != k) {
            o[2 * (n - k)] = double(sr);
            o[2 * (n - k) + 1] = double(-si);
          }
        }
      }
    }
  };
  const int nt = std::max(1, std::min(16, int(std::thread::hardware_concurrency())));
  std::vector<std::thread> th;
  const int chunk = (U + nt - 1) / nt;
  for (int t = 0; t < nt; ++t) {
    const int u0 = t * chunk, u1 = std::min(U, u0 + chunk);
    if (u0 < u1) th.emplace_back(work, u0, u1);
  }
  for (auto& t : th) t.join();
}

int build_harmonic_table(spx_plan* p) {
  // SPX_HARMONIC_TABLE=0: no table (the harmonic kernel transforms per line;
  // for plans whose memory is tighter than 2 * (distinct kh) * nz complex values)
  if (const char* e = std::getenv("SPX_HARMONIC_TABLE"))
    if (e[0] == '0') return SPX_OK;
  const int n = p->g.nz, np = p->np, ny = p->g.ny;
  const size_t lines1 = size_t(p->nkx) * ny;
  std::unordered_map<uint64_t, int> seen;
  std::vector<double> khv;
  std::vector<int> idx(lines1 * size_t(p->nb));
  for (size_t l = 0; l < lines1; ++l) {
    uint64_t key;
    std::memcpy(&key, &p->h_khom[l], sizeof(key));
    auto it = seen.find(key);
    if (it == seen.end()) {
      it = seen.emplace(key, int(khv.size())).first;
      khv.push_back(p->h_khom[l]);
    }
    idx[l] = it->second;
  }
  for (int b = 1; b < p->nb; ++b)
    std::copy(idx.begin(), idx.begin() + long(lines1), idx.begin() + long(b * lines1));
  const int U = int(khv.size());
  void *tab = nullptr, *hid = nullptr, *kv = nullptr;
  const size_t tab_bytes = 2 * size_t(U) * n * sizeof(Cd);
  if (hipMalloc(&tab, tab_bytes) != hipSuccess) {
    (void)hipGetLastError();  // no table: proj_harmonic transforms per line
    return SPX_OK;
  }
  p->owned.push_back(tab);
  p->dev_bytes += tab_bytes;
  SPX_HIP(hipMalloc(&hid, idx.size() * sizeof(int)));
  p->owned.push_back(hid);
  p->dev_bytes += idx.size() * sizeof(int);
  SPX_HIP(hipMalloc(&kv, khv.size() * sizeof(double)));
  SPX_HIP(hipMemcpy(hid, idx.data(), idx.size() * sizeof(int), hipMemcpyHostToDevice));
  SPX_HIP(hipMemcpy(kv, khv.data(), khv.size() * sizeof(double), hipMemcpyHostToDevice));
  // long double on the host (the default; about 2 U np nz / 2 multiply-adds,
  // split over up to 16 threads), or the profiles continued and transformed
  // on the device in float64 for tables whose host cost would be too high
  // (U nz^2 > 2e11, e.g. 1024^3) or with SPX_HARMONIC_LD=0
  bool host_ld = double(U) * double(n) * double(n) <= 2e11;
  if (const char* e = std::getenv("SPX_HARMONIC_LD")) host_ld = e[0] != '0';
  if (host_ld) {
    std::vector<double> h;
    harmonic_table_ld(khv, n, np, p->g.Cz, p->g.oz, p->h_z, p->g.Lz, p->h_dir, h);
    SPX_HIP(hipMemcpy(tab, h.data(), tab_bytes, hipMemcpyHostToDevice));
  } else {
    const long tot = long(U) * np;
    hipLaunchKernelGGL(k_harm_profiles, dim3(unsigned((tot + 255) / 256)), dim3(256), 0,
                       p->stream, static_cast<Cd*>(tab), static_cast<const double*>(kv), U, n,
                       np, p->d_z, p->g.Lz);
    SPX_CHECK_LAUNCH();
    // continuation + forward z FFT of the 2U profile rows
    const LineFFTHost& f = p->fz;
    const long rows = 2L * U;
    SPX_LAUNCH_RS(f, k_z_inplace, double, , dim3(unsigned((rows + f.nl - 1) / f.nl)), blk(f),
                  f.lds_bytes(16), p->stream, static_cast<Cd*>(tab), rows, f.dev<Cd>(), np, 1,
                  1, 1.0, p->d_dir, p->g.Cz, p->g.oz);
    SPX_CHECK_LAUNCH();
  }
  SPX_HIP(hipStreamSynchronize(p->stream));
  (void)hipFree(kv);
  p->d_htab = tab;
  p->d_hidx = static_cast<int*>(hid);
  p->nhtab = U;
  return SPX_OK;
}

int project_harmonic(spx_plan* p, void* a, void* b, void* c, void* d, const RkLoad* rk,
                     bool walls_only) {
  const int nz = p->g.nz;
  if (!p->fused_project || !p->d_htab || !p->d_wtop || nz > 64 * 17) return SPX_ERR_UNSUPPORTED;
  const long nlines = p->wlines();
  RkLoad r;
  std::memset(&r, 0, sizeof(r));
  if (rk) r = *rk;
  const double extra = rk ? (rk->nl32 ? 8.0 : 16.0) + 32.0 : 0.0;
  ProfScope ps_(p, rk ? "project_harmonic_rk" : "project_harmonic",
                double(p->wnc()) * (7.0 * 16.0 + extra) + 2.0 * p->nhtab * double(nz) * 16.0);
  const dim3 grid(unsigned((nlines + 3) / 4)), block(256);
  // RK mode of the branch-free variants (lines of at least 64 (KPL - 1) points)
  const int rkm = rk ? (rk->nl32 ? 2 : 1) : 0;
#define SPX_PH(KPL)                                                                            \
  if (nz < 64 * (KPL - 1)) {                                                                   \
    SPX_PH2(KPL, -1);                                                                          \
  } else if (rkm == 0) {                                                                       \
    SPX_PH2(KPL, 0);                                                                           \
  } else if (rkm == 1) {                                                                       \
    SPX_PH2(KPL, 1);                                                                           \
  } else {                                                                                     \
    SPX_PH2(KPL, 2);                                                                           \
  }
#define SPX_PH2(KPL, RKM)                                                                      \
  hipLaunchKernelGGL((k_project_harmonic<KPL, RKM>), grid, block, 0, p->stream, C_<double>(a), \
                     C_<double>(b), C_<double>(c), C_<double>(d), C_<double>(p->bc), nlines,   \
                     nz, p->g.ny, p->d_kx, p->d_ky, p->d_kz,                                   \
                     static_cast<const Cd*>(p->d_wtop), 1.0 / double(nz), r, rk ? 1 : 0,       \
                     walls_only ? C_<double>(p->pw) : nullptr, p->np,                          \
                     static_cast<const Cd*>(p->d_htab), p->d_hidx, p->g.Lz, p->d_khom)
  if (nz <= 64 * 3) {
    SPX_PH2(3, -1);
  } else if (nz <= 64 * 5) {
    SPX_PH(5)
  } else if (nz <= 64 * 9) {
    SPX_PH(9)
  } else {
    SPX_PH(17)
  }
#undef SPX_PH
#undef SPX_PH2
  SPX_CHECK_LAUNCH();
  return SPX_OK;
}

int proj_harmonic(spx_plan* p, void* a, void* b, void* c, int k0, int k1) {
  const long nlines = p->wlines();
  if (p->d_htab) {
    ProfScope ps_(p, "proj_harmonic",
                  6.0 * double(p->wnc()) * 16.0 + 2.0 * p->nhtab * double(p->g.nz) * 16.0);
    hipLaunchKernelGGL(k_proj_harmonic_tab<double>, dim3(unsigned((nlines + 3) / 4)), dim3(256), 0,
                       p->stream, C_<double>(a), C_<double>(b), C_<double>(c),
                       C_<double>(p->bc), static_cast<const Cd*>(p->d_htab), p->d_hidx, nlines,
                       p->g.ny, p->g.nz, p->g.Lz, p->d_khom, p->d_kx, p->d_ky, k0, k1);
    SPX_CHECK_LAUNCH();
    return SPX_OK;
  }
  const LineFFTHost& f = p->fz2;
  const size_t lds = f.lds_bytes(16, 2);
  ProfScope ps_(p, "proj_harmonic", 6.0 * double(p->wnc()) * 16.0);
  SPX_LAUNCH_RS(f, k_proj_harmonic, double, , dim3(unsigned((nlines + f.nl - 1) / f.nl)), blk(f),
                lds, p->stream, C_<double>(a), C_<double>(b), C_<double>(c), C_<double>(p->bc),
                f.dev<Cd>(), nlines, p->g.ny, p->np, p->g.Lz, p->d_khom, p->d_z, p->d_kx,
                p->d_ky, p->d_dir, p->g.Cz, p->g.oz, k0, k1);
  SPX_CHECK_LAUNCH();
  return SPX_OK;
}

static int zline_launch(spx_plan* p, const LinComb& lc, const void* b, int mode, double* out) {
  const LineFFTHost& f = mode == 1 ? p->fz2 : p->fz;
  const long nlines = p->wlines();
  LinCombDev<double> d;
  for (int t = 0; t < 3; ++t) {
    d.in[t] = static_cast<const Cd*>(lc.in[t]);
    d.dir[t] = lc.dir[t];
    d.sgn[t] = lc.sgn[t];
  }
  const size_t lds = f.lds_bytes(16, mode == 1 ? 2 : 1) + size_t(f.nl) * f.tpl * sizeof(double);
  ProfScope ps_(p, "zline_reduce", double(p->wnc()) * 16.0 * (mode == 1 ? 2 : 1));
  SPX_LAUNCH_RS(f, k_zline_reduce, double, , dim3(unsigned((nlines + f.nl - 1) / f.nl)), blk(f),
                lds, p->stream, d, C_<double>(b), mode, out, f.dev<Cd>(), nlines, p->g.ny, p->np,
                p->d_kx, p->d_ky, p->d_kz);
  SPX_CHECK_LAUNCH();
  return SPX_OK;
}

int zline_sumsq(spx_plan* p, const LinComb& lc, double* line_out) {
  return zline_launch(p, lc, nullptr, 0, line_out);
}

int zline_dot(spx_plan* p, const void* a, const void* b, double* line_out) {
  LinComb lc;
  lc.in[0] = a;
  lc.in[1] = lc.in[2] = nullptr;
  lc.dir[0] = lc.dir[1] = lc.dir[2] = 0;
  lc.sgn[0] = 1;
  lc.sgn[1] = lc.sgn[2] = 0;
  return zline_launch(p, lc, b, 1, line_out);
}

int zline_dot_lc(spx_plan* p, const LinComb& lc, const void* b, double* line_out) {
  return zline_launch(p, lc, b, 1, line_out);
}

int zline_wall_values(spx_plan* p, const LinComb& lc, double* line_out) {
  return zline_launch(p, lc, nullptr, 3, line_out);
}

int zline_walls(spx_plan* p, const void* a, double* line_out) {
  LinComb lc;
  lc.in[0] = a;
  lc.in[1] = lc.in[2] = nullptr;
  lc.dir[0] = lc.dir[1] = lc.dir[2] = 0;
  lc.sgn[0] = 1;
  lc.sgn[1] = lc.sgn[2] = 0;
  return zline_launch(p, lc, nullptr, 2, line_out);
}
#endif  // SPX_PART 3 (boundary, projection, diagnostics)

}  // namespace spx
