// In-LDS batched line DFT engine for gfx950 (CDNA4, wave64).
//
// A workgroup owns nl lines of one length n; tpl threads work on each line.
// A mixed-radix Stockham autosort FFT (radices 2,3,4,5,7,8,9,11,13) runs in
// LDS with register staging: every pass reads R strided points per butterfly,
// applies twiddles from one M-entry table staged in LDS, does the radix-R DFT
// in registers, synchronises, and writes the autosorted result back.
//
// Lengths with a prime factor above 13 are reduced to a smooth FFT inside
// LDS (one HBM read and one write per line either way):
//   * Rader (n prime, n-1 smooth): cyclic convolution of length n-1;
//   * Bluestein (otherwise): chirp-z convolution with a smooth M >= 2n-1.
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

// cos(2 pi m / R) and sin(2 pi m / R) for m = 1..(R-1)/2 of the odd radices,
// correctly rounded (computed to 45 digits)
template <int R>
struct OddRoots;
template <>
struct OddRoots<5> {
  static constexpr double c[2] = {0.3090169943749474241022934, -0.8090169943749474241022934};
  static constexpr double s[2] = {0.9510565162951535721164393, 0.5877852522924731291687060};
};
template <>
struct OddRoots<7> {
  static constexpr double c[3] = {0.6234898018587335305250049, -0.2225209339563144042889026,
                                  -0.9009688679024191262361023};
  static constexpr double s[3] = {0.7818314824680298087084445, 0.9749279121818236070181317,
                                  0.4338837391175581204757683};
};
template <>
struct OddRoots<11> {
  static constexpr double c[5] = {0.8412535328311811688618116, 0.4154150130018864255292741,
                                  -0.1423148382732851404437927, -0.6548607339452850640569251,
                                  -0.9594929736144973898903681};
  static constexpr double s[5] = {0.5406408174555975821076360, 0.9096319953545183714117154,
                                  0.9898214418809327323760920, 0.7557495743542582837740358,
                                  0.2817325568414296977114179};
};
template <>
struct OddRoots<13> {
  static constexpr double c[6] = {0.8854560256532098959003755, 0.5680647467311558025118076,
                                  0.1205366802553230533490677, -0.3546048870425356259696379,
                                  -0.7485107481711010986346306, -0.9709418174260520271569823};
  static constexpr double s[6] = {0.4647231720437685456560153, 0.8229838658936563945796174,
                                  0.9927088740980539928007516, 0.9350162426854148234397846,
                                  0.6631226582407952023767855, 0.2393156642875577671487537};
};

// Radix-R DFT in registers, forward sign (odd R: r paired with R-r, the
// roots of unity as immediates).
template <typename T, int R>
struct Dft {
  static __device__ __forceinline__ void run(cpx<T>* v, const cpx<T>*, int) {
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
        const int m = (r * q) % R;  // compile-time after unrolling
        const T c = T(OddRoots<R>::c[(m <= H ? m : R - m) - 1]);
        const T sn = m <= H ? T(OddRoots<R>::s[m - 1]) : T(-OddRoots<R>::s[R - m - 1]);
        re = cadd(re, cscale(a[r], c));
        im = cadd(im, cscale(b[r], sn));
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
struct Dft<T, 3> {
  static __device__ __forceinline__ void run(cpx<T>* v, const cpx<T>*, int) {
    const T c = T(-0.5), s = T(0.86602540378443864676372317075293618347L);
    const cpx<T> t1 = cadd(v[1], v[2]);
    const cpx<T> t2 = mkc<T>(v[0].x + c * t1.x, v[0].y + c * t1.y);
    const cpx<T> d = csub(v[1], v[2]);
    const cpx<T> t3 = mkc<T>(s * d.y, -s * d.x);  // -i s d
    v[0] = cadd(v[0], t1);
    v[1] = cadd(t2, t3);
    v[2] = csub(t2, t3);
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

// radix 9 as 3 x 3 with the inner twiddles W_9^{ab}
template <typename T>
struct Dft<T, 9> {
  static __device__ __forceinline__ void run(cpx<T>* v, const cpx<T>* tw, int s) {
    cpx<T> a[3][3];
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      cpx<T> t[3] = {v[r], v[r + 3], v[r + 6]};
      Dft<T, 3>::run(t, nullptr, 0);
#pragma unroll
      for (int q = 0; q < 3; ++q) a[q][r] = t[q];
    }
    // twiddle a[q][r] by W_9^{q r}
    a[1][1] = cmul(a[1][1], tw[1 * s]);
    a[1][2] = cmul(a[1][2], tw[2 * s]);
    a[2][1] = cmul(a[2][1], tw[2 * s]);
    a[2][2] = cmul(a[2][2], tw[4 * s]);
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      cpx<T> t[3] = {a[q][0], a[q][1], a[q][2]};
      Dft<T, 3>::run(t, nullptr, 0);
      v[q] = t[0];
      v[q + 3] = t[1];
      v[q + 6] = t[2];
    }
  }
};

// Points a thread stages in registers per pass; the plan sizes threads per
// line so that no thread holds more (radix 9/11/13 passes hold one butterfly).
constexpr int kStage = 8;
template <int R>
struct MaxB {
  static constexpr int value = (kStage / R) > 0 ? (kStage / R) : 1;
};

// Synchronisation between the passes of one line.  A line held by threads of
// one wavefront (tpl <= 64) needs only the wavefront's own ordering of its LDS
// accesses (a wavefront's LDS instructions complete in order) and a compiler
// fence; a line spread over several wavefronts needs the workgroup barrier.
// line_dft ends with a workgroup barrier either way, so code between
// transforms may read other lines' buffers.
__device__ __forceinline__ void line_sync(int tpl) {
  if (tpl <= 64) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  } else {
    __syncthreads();
  }
}

// j mod Ns with the per-stage magic multiplier (j, Ns < 2^16)
__device__ __forceinline__ int fastmod(int j, int Ns, unsigned magic) {
  const int q = int(__umulhi(unsigned(j), magic));
  return j - q * Ns;
}

template <typename T, int R>
__device__ __forceinline__ void stockham_pass(cpx<T>* x, int M, int Ns, unsigned magic, int lane,
                                              int tpl, const cpx<T>* tw) {
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
        const int st = fastmod(j, Ns, magic) * tstep;
#pragma unroll
        for (int r = 1; r < R; ++r) v[b][r] = cmul(v[b][r], tw[r * st]);
      }
      Dft<T, R>::run(v[b], tw, s);
    }
  }
  line_sync(tpl);
#pragma unroll
  for (int b = 0; b < MB; ++b) {
    const int j = lane + b * tpl;
    if (j < nb) {
      const int jm = Ns > 1 ? fastmod(j, Ns, magic) : 0;
      const int base = (j - jm) * R + jm;
#pragma unroll
      for (int r = 0; r < R; ++r) x[base + r * Ns] = v[b][r];
    }
  }
  line_sync(tpl);
}

// Load / store variants of a pass: the index permutations, chirps and
// pointwise filters of the Rader and Bluestein convolutions are applied on
// the first pass's loads and the last passes' stores instead of in passes
// of their own (one LDS round trip and one barrier less each).
enum { LD_PLAIN = 0, LD_GATHER = 1, LD_CHIRP = 2 };
enum { ST_PLAIN = 0, ST_FILTER = 1, ST_SCATTER = 2, ST_CHIRP = 3 };

template <typename T>
struct PassIO {
  int n;                 // LD_CHIRP / ST_CHIRP: input / output length
  const int* gin;        // LD_GATHER: element q of the pass input is src[gin[q]]
  const int* gout;       // ST_SCATTER: output m goes to dst[gout[m]] as x0 + conj(v)
  const cpx<T>* filt;    // ST_FILTER: output m becomes conj(v * filt[m])
  const cpx<T>* chirp;   // LD_CHIRP: src[q] * chirp[q] (0 for q >= n); ST_CHIRP: conj(v) * chirp[m], m < n
  cpx<T> x0;             // Rader: the input's element 0
};

template <typename T, int R, int LD>
__device__ __forceinline__ void pass_load(cpx<T>* v, const cpx<T>* src, int j, int nb,
                                          const PassIO<T>& io) {
  if constexpr (LD == LD_GATHER) {
#pragma unroll
    for (int r = 0; r < R; ++r) v[r] = src[io.gin[j + r * nb]];
  } else if constexpr (LD == LD_CHIRP) {
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int q = j + r * nb;
      v[r] = q < io.n ? cmul(src[q], io.chirp[q]) : mkc<T>(0, 0);
    }
  } else {
#pragma unroll
    for (int r = 0; r < R; ++r) v[r] = src[j + r * nb];
  }
}

// X0 (Rader's output 0 = x0 + the first convolution FFT's element 0) is
// taken by the thread that stores element 0 under ST_FILTER.
template <typename T, int R, int ST>
__device__ __forceinline__ void pass_store(cpx<T>* dst, const cpx<T>* v, int base, int Ns,
                                           const PassIO<T>& io, cpx<T>& X0) {
  if constexpr (ST == ST_FILTER) {
    if (base == 0) X0 = cadd(io.x0, v[0]);
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int m = base + r * Ns;
      dst[m] = cconj(cmul(v[r], io.filt[m]));
    }
  } else if constexpr (ST == ST_SCATTER) {
#pragma unroll
    for (int r = 0; r < R; ++r) dst[io.gout[base + r * Ns]] = cadd(io.x0, cconj(v[r]));
  } else if constexpr (ST == ST_CHIRP) {
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int m = base + r * Ns;
      if (m < io.n) dst[m] = cmul(cconj(v[r]), io.chirp[m]);
    }
  } else {
#pragma unroll
    for (int r = 0; r < R; ++r) dst[base + r * Ns] = v[r];
  }
}

// A pass from src to a different buffer dst: no barrier is needed between
// the loads and the stores, one after them.
template <typename T, int R, int LD, int ST>
__device__ __forceinline__ void stockham_pass_pp(const cpx<T>* src, cpx<T>* dst, int M, int Ns,
                                                 unsigned magic, int lane, int tpl,
                                                 const cpx<T>* tw, const PassIO<T>& io,
                                                 cpx<T>& X0) {
  constexpr int MB = MaxB<R>::value;
  const int nb = M / R;
  const int tstep = M / (Ns * R);
  const int s = M / R;
#pragma unroll
  for (int b = 0; b < MB; ++b) {
    const int j = lane + b * tpl;
    if (j < nb) {
      cpx<T> v[R];
      pass_load<T, R, LD>(v, src, j, nb, io);
      const int jm = Ns > 1 ? fastmod(j, Ns, magic) : 0;
      if (Ns > 1) {
        const int st = jm * tstep;
#pragma unroll
        for (int r = 1; r < R; ++r) v[r] = cmul(v[r], tw[r * st]);
      }
      Dft<T, R>::run(v, tw, s);
      pass_store<T, R, ST>(dst, v, (j - jm) * R + jm, Ns, io, X0);
    }
  }
  line_sync(tpl);
}

// In-place pass with load / store variants (two barriers).
template <typename T, int R, int LD, int ST>
__device__ __forceinline__ void stockham_pass_io(cpx<T>* x, int M, int Ns, unsigned magic,
                                                 int lane, int tpl, const cpx<T>* tw,
                                                 const PassIO<T>& io, cpx<T>& X0) {
  constexpr int MB = MaxB<R>::value;
  const int nb = M / R;
  const int tstep = M / (Ns * R);
  const int s = M / R;
  cpx<T> v[MB][R];
#pragma unroll
  for (int b = 0; b < MB; ++b) {
    const int j = lane + b * tpl;
    if (j < nb) {
      pass_load<T, R, LD>(v[b], x, j, nb, io);
      if (Ns > 1) {
        const int st = fastmod(j, Ns, magic) * tstep;
#pragma unroll
        for (int r = 1; r < R; ++r) v[b][r] = cmul(v[b][r], tw[r * st]);
      }
      Dft<T, R>::run(v[b], tw, s);
    }
  }
  line_sync(tpl);
#pragma unroll
  for (int b = 0; b < MB; ++b) {
    const int j = lane + b * tpl;
    if (j < nb) {
      const int jm = Ns > 1 ? fastmod(j, Ns, magic) : 0;
      pass_store<T, R, ST>(x, v[b], (j - jm) * R + jm, Ns, io, X0);
    }
  }
  line_sync(tpl);
}

// Radix sets a kernel instance is compiled for (register pressure follows the
// largest radix pass compiled in): powers of two; {2,3,4,5,7,8,9}; all;
// {2,...,9} with the convolution tables (filter, chirp, indices) read from
// global memory instead of LDS (very long lines).  The table location is a
// compile-time property so that every table access has a known address space.
enum { RS_POW2 = 0, RS_SMOOTH = 1, RS_ALL = 2, RS_GTAB = 3 };

// Forward FFT of length f.M on the line x (LDS) with twiddles tw (LDS).
// Every thread of the workgroup must call it (it synchronises the line).
template <typename T, int RS>
__device__ __forceinline__ void fft_lds(cpx<T>* x, const LineFFT<cpx<T>>& f, const cpx<T>* tw,
                                        int lane) {
  int Ns = 1;
  for (int st = 0; st < f.nstage; ++st) {
    const int R = f.radix[st];
    const unsigned mg = f.magic[st];
    if (R == 8) stockham_pass<T, 8>(x, f.M, Ns, mg, lane, f.tpl, tw);
    else if (R == 4) stockham_pass<T, 4>(x, f.M, Ns, mg, lane, f.tpl, tw);
    else if (R == 2) stockham_pass<T, 2>(x, f.M, Ns, mg, lane, f.tpl, tw);
    else if constexpr (RS >= RS_SMOOTH) {
      if (R == 3) stockham_pass<T, 3>(x, f.M, Ns, mg, lane, f.tpl, tw);
      else if (R == 5) stockham_pass<T, 5>(x, f.M, Ns, mg, lane, f.tpl, tw);
      else if (R == 7) stockham_pass<T, 7>(x, f.M, Ns, mg, lane, f.tpl, tw);
      else if (R == 9) stockham_pass<T, 9>(x, f.M, Ns, mg, lane, f.tpl, tw);
      else if constexpr (RS == RS_ALL) {
        if (R == 11) stockham_pass<T, 11>(x, f.M, Ns, mg, lane, f.tpl, tw);
        else if (R == 13) stockham_pass<T, 13>(x, f.M, Ns, mg, lane, f.tpl, tw);
      }
    }
    Ns *= R;
  }
}

// One pass of stage st of a convolution FFT (radix dispatch); PP: from src
// to dst, else in place on src.
template <typename T, int RS, bool PP, int LD, int ST>
__device__ __forceinline__ void conv_pass(int st, cpx<T>* src, cpx<T>* dst,
                                          const LineFFT<cpx<T>>& f, const cpx<T>* tw, int lane,
                                          const PassIO<T>& io, cpx<T>& X0) {
  int Ns = 1;
  for (int u = 0; u < st; ++u) Ns *= f.radix[u];
  const int R = f.radix[st];
  const unsigned mg = f.magic[st];
#define SPX_CONV_PASS(RR)                                                                    \
  do {                                                                                       \
    if constexpr (PP)                                                                        \
      stockham_pass_pp<T, RR, LD, ST>(src, dst, f.M, Ns, mg, lane, f.tpl, tw, io, X0);       \
    else                                                                                     \
      stockham_pass_io<T, RR, LD, ST>(src, f.M, Ns, mg, lane, f.tpl, tw, io, X0);            \
  } while (0)
  if (R == 8) SPX_CONV_PASS(8);
  else if (R == 4) SPX_CONV_PASS(4);
  else if (R == 2) SPX_CONV_PASS(2);
  else if constexpr (RS >= RS_SMOOTH) {
    if (R == 3) SPX_CONV_PASS(3);
    else if (R == 5) SPX_CONV_PASS(5);
    else if (R == 7) SPX_CONV_PASS(7);
    else if (R == 9) SPX_CONV_PASS(9);
    else if constexpr (RS == RS_ALL) {
      if (R == 11) SPX_CONV_PASS(11);
      else if (R == 13) SPX_CONV_PASS(13);
    }
  }
#undef SPX_CONV_PASS
}

// The two length-M FFTs of a Rader (PP) or Bluestein (in place) convolution
// as 2P passes, P = f.nstage >= 2: pass 0 loads with LD0, pass P-1 stores
// with STM, pass 2P-1 with STE.  Ping-pong passes alternate a -> b -> a ...,
// so pass 2P-1 stores into a.
template <typename T, int RS, bool PP, int LD0, int STM, int STE>
__device__ __forceinline__ void conv_passes(cpx<T>* a, cpx<T>* b, const LineFFT<cpx<T>>& f,
                                            const cpx<T>* tw, int lane, const PassIO<T>& io,
                                            cpx<T>& X0) {
  const int P = f.nstage;
  conv_pass<T, RS, PP, LD0, ST_PLAIN>(0, a, b, f, tw, lane, io, X0);
  for (int t = 1; t < 2 * P - 1; ++t) {
    cpx<T>* src = (t & 1) ? b : a;
    cpx<T>* dst = (t & 1) ? a : b;
    if (t == P - 1) conv_pass<T, RS, PP, LD_PLAIN, STM>(t, src, dst, f, tw, lane, io, X0);
    else conv_pass<T, RS, PP, LD_PLAIN, ST_PLAIN>(t < P ? t : t - P, src, dst, f, tw, lane, io, X0);
  }
  conv_pass<T, RS, PP, LD_PLAIN, STE>(P - 1, b, a, f, tw, lane, io, X0);
}

// Copy the tables into LDS (called by all threads; the caller's next
// __syncthreads publishes them): twiddles at [0, M); Bluestein: filter at
// [M, 2M), chirp at [2M, 2M+n); Rader: filter at [M, 2M), then the 2M
// permutation indices as ints.
template <typename T>
__device__ __forceinline__ void stage_twiddles(cpx<T>* tws, const LineFFT<cpx<T>>& f) {
  for (int m = threadIdx.x; m < f.M; m += blockDim.x) tws[m] = f.tw[m];
  if (f.mode == 0 || f.tab == f.M) return;  // RS_GTAB plans read them from global memory
  for (int m = threadIdx.x; m < f.M; m += blockDim.x) tws[f.M + m] = f.bfil[m];
  if (f.mode == 1) {
    for (int m = threadIdx.x; m < f.n; m += blockDim.x) tws[2 * f.M + m] = f.chirp[m];
  } else {
    int* pr = reinterpret_cast<int*>(tws + 2 * f.M);
    for (int m = threadIdx.x; m < 2 * f.M; m += blockDim.x) pr[m] = f.perm[m];
  }
}

// Forward DFT of length f.n of the n values at x[0..n), result in x[0..n).
// The caller has synchronised after filling x; x spans f.ldl elements.
template <typename T, int RS>
__device__ __forceinline__ void line_dft(cpx<T>* x, const LineFFT<cpx<T>>& f, const cpx<T>* tw,
                                         int lane) {
  const int n = f.n, M = f.M;
  // the power-of-two family is only used for direct transforms: keep the
  // Rader / Bluestein code (and its registers) out of those instances
  if (RS == RS_POW2 || f.mode == 0) {
    fft_lds<T, RS>(x, f, tw, lane);
    __syncthreads();
    return;
  }
  // tables staged in LDS by stage_twiddles, or global (RS_GTAB)
  constexpr bool G = RS == RS_GTAB;
  PassIO<T> io;
  io.n = n;
  io.filt = G ? f.bfil : tw + M;
  cpx<T> X0 = mkc<T>(0, 0);
  if (f.mode == 1) {  // Bluestein: chirp, FFT, filter, FFT, chirp; in place
    io.gin = io.gout = nullptr;
    io.chirp = G ? f.chirp : tw + 2 * M;
    io.x0 = mkc<T>(0, 0);
    conv_passes<T, RS, false, LD_CHIRP, ST_FILTER, ST_CHIRP>(x, x, f, tw, lane, io, X0);
    __syncthreads();
    return;
  }
  // Rader: M = n - 1, second buffer x + n.  The first FFT gathers its input
  // x[g^q] on its first pass and applies the filter on its last; the second
  // scatters x0 + conj(.) to x[g^-p] on its last pass, which stores into x.
  io.gin = G ? f.perm : reinterpret_cast<const int*>(tw + 2 * M);
  io.gout = io.gin + M;
  io.chirp = nullptr;
  io.x0 = x[0];
  conv_passes<T, RS, true, LD_GATHER, ST_FILTER, ST_SCATTER>(x, x + n, f, tw, lane, io, X0);
  if (lane == 0) x[0] = X0;
  __syncthreads();
}

// RkLoad applied to component q's value v at element e = line*nz + k,
// line = i*ny + j.
// The same update from loaded values: kl2 = kx^2 + ky^2 of the line, kz, the
// filter factors fx, fy, fz and the diffusivity nu of the line.
__device__ __forceinline__ cpx<double> rk_apply(const RkLoad& r, cpx<double> v, cpx<double> nl,
                                                cpx<double> f, cpx<double> c0, double kl2,
                                                double kz, double fx, double fy, double fz,
                                                double nu) {
#pragma clang fp contract(off)
  const double kk2 = kl2 + kz * kz;
  nl = cscale(cscale(cscale(nl, fx), fy), fz);
  const cpx<double> lap = cscale(v, -kk2);
  const cpx<double> t = cadd(csub(cscale(lap, nu), nl), f);
  return cadd(c0, cscale(cscale(t, r.dt), r.rmp));
}

__device__ __forceinline__ cpx<double> rk_on_load(const RkLoad& r, int q, cpx<double> v, long e,
                                                  long i, int j, int k) {
#pragma clang fp contract(off)
  const RkLoadComp& c = r.c[q];
  const double kl2 = r.kx[i] * r.kx[i] + r.ky[j] * r.ky[j];
  cpx<double> nl;
  if (r.nl32) {
    const cpx<float> t = static_cast<const cpx<float>*>(c.nl)[e];
    nl = mkc<double>(double(t.x), double(t.y));
  } else {
    nl = static_cast<const cpx<double>*>(c.nl)[e];
  }
  return rk_apply(r, v, nl, static_cast<const cpx<double>*>(c.f)[e],
                  static_cast<const cpx<double>*>(c.c0)[e], kl2, r.kz[k], r.fkx[i], r.fky[j],
                  r.fkz[k], c.nu[i]);
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
