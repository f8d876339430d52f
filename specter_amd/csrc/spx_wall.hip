// Wall update of the velocity (and scalar) z lines as a rank-(2+C) correction
// on the matrix cores (v_mfma_f64_16x16x4f64).
//
// goto_domain_w_boundaries + noslip_z + goto_3d_fourier (boundary_mod.fpp:72-194,
// vboundary.f90:154-211; s_constant_z, sboundary.f90:122-165) turn a z line V
// into FFT(x) where x = IFFT(V)/nz on the physical points, with the two wall
// values replaced and the C continuation points recomputed from the 2d
// boundary values (fftp.fpp:755-773).  Only 2 + C points of x change, so
//
//   V' = V + FFT(delta),  delta supported on z = 0, np-1, np..nz-1,
//
// with delta from the unnormalised inverse DFT of V at the 2d + C points
// z in {0..d-1, np-d..np-1, np..nz-1}.  Both are dense products over a tile
// of 16 lines:
//
//   GEMM A  X[q][l]   = sum_kappa WinvT[q][kappa] Vr[kappa][l]      (partial IDFT)
//   GEMM C  Vr'[kappa][l] = Vr[kappa][l] + sum_rho Wf[kappa][rho] dr[rho][l]
//
// in real form (kappa = 2z + {re, im}; q and rho likewise), float64, with the
// DFT matrices formed in long double.  Vr stays in registers between the two
// products.  v_mfma_f64_16x16x4f64 takes A[i][k] from lane i + 16k, B[k][j]
// from lane j + 16k and keeps D[i][j] in register i / 4 of lane j + 16 (i % 4)
// (measured).  Each wave holds a range of 16-row kappa blocks of the 16 lines
// as GEMM C accumulators whose rows are permuted within a block (row
// 4r + g <-> kappa offset kap_off(g, r)), so that lane group g, registers
// r = 0..3, holds the complex values z = 8 block + g and z + 4 of its line (a
// load instruction reads 64 contiguous bytes of each of its 16 lines); the
// same registers are GEMM A's B operand (register s = K step s, kappa offset
// kap_off(g, s)).  A line is read once (with
// the RK update applied on the load) and written once.  The waves'
// partial X are summed in LDS in wave order (deterministic), delta is formed
// there exactly as the reference forms the continuation, and GEMM C
// accumulates onto Vr.
//
// Besides moving the transforms off the LDS FFT engine, the correction keeps
// the interior of every line exactly as the RK update left it: only delta
// (a small correction) is rounded through a transform.  The reference's
// inverse-then-forward round trip rounds the whole line twice per stage.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "spx_engine.h"

namespace spx {

typedef double v4d __attribute__((ext_vector_type(4)));

namespace {

using Cd = cpx<double>;

__device__ __forceinline__ Cd ikmul_d(double k, Cd v) { return mkc<double>(-(k * v.y), k * v.x); }

// kappa offset (2 z_off + re/im) within a 16-row block held by lane group g in
// register s: z_off = g + 4 (s / 2), re/im = s % 2
__host__ __device__ constexpr int kap_off(int g, int s) { return 2 * g + 8 * (s >> 1) + (s & 1); }

__device__ __forceinline__ v4d mfma_f64(double a, double b, v4d c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// GEMM A rows [M0, M1) (16-row tiles) of the wave's kappa blocks, then the
// waves' partial sums into the ns LDS slots in a fixed order: waves < ns
// store, waves >= ns add onto slot wave - ns (one barrier between).
// Point z of partial-IDFT index qi (0..d-1, np-d..np-1, np..n-1) and of
// correction index p (0, np-1, np..n-1).
__device__ __forceinline__ int qpoint(int qi, int d, int np, int nq) {
  if (qi >= nq) return 0;
  return qi < d ? qi : (qi < 2 * d ? np - 2 * d + qi : np - 2 * d + qi);
}
__device__ __forceinline__ int rpoint(int p, int np, int nrho) {
  if (p >= nrho) return 0;
  return p == 0 ? 0 : (p == 1 ? np - 1 : np + p - 2);
}

// x mod n for 0 <= x < 2^31 with magic = floor(2^32 / n) + 1
__device__ __forceinline__ int modn(unsigned x, int n, unsigned magic) {
  const unsigned q = __umulhi(x, magic);
  int r = int(x - q * unsigned(n));
  return r < 0 ? r + n : r;
}

// TAB 0: A operands read from per-plan fragment tables in global memory;
// TAB 1: formed from the n roots of unity staged in LDS (index z q mod n).
template <int MT, int M0, int M1, int KBW, int TAB>
__device__ __forceinline__ void gemm_a_rows(const double* __restrict__ winv, const Cd* rl,
                                            const v4d (&vr)[KBW], int b0, int nbk, int lane,
                                            int wave, int ns, double* red, int n, int np,
                                            int d, int nq) {
  v4d acc[M1 - M0];
#pragma unroll
  for (int m = 0; m < M1 - M0; ++m) acc[m] = v4d{0, 0, 0, 0};
  if constexpr (TAB == 0) {
#pragma unroll
    for (int b = 0; b < KBW; ++b) {
      const int bb = b0 + b;
      if (bb < nbk) {  // wave-uniform
        const double* wa = winv + size_t(bb) * 4 * MT * 64 + lane;
#pragma unroll
        for (int s = 0; s < 4; ++s)
#pragma unroll
          for (int m = M0; m < M1; ++m)
            acc[m - M0] = mfma_f64(wa[(s * MT + m) * 64], vr[b][s], acc[m - M0]);
      }
    }
  } else {
    // lane = i + 16k: row 16m + i = (point qi, re/im qc); K step s of block
    // bb, lane group k = (z = 8 bb + k + 4 (s/2), re/im s&1); e^{+2 pi i z q / n}
    const int i = lane & 15, k = lane >> 4;
    const int qc = i & 1;
    int q4[M1 - M0], inc[M1 - M0], idx[M1 - M0];
#pragma unroll
    for (int m = M0; m < M1; ++m) {
      const int q = qpoint((16 * m + i) >> 1, d, np, nq);
      q4[m - M0] = (4 * q) % n;
      inc[m - M0] = (8 * q) % n;
      idx[m - M0] = int((long long)(8 * b0 + k) * q % n);
    }
#pragma unroll
    for (int b = 0; b < KBW; ++b) {
      const int bb = b0 + b;
      if (bb < nbk) {
#pragma unroll
        for (int m = M0; m < M1; ++m) {
          int i1 = idx[m - M0] + q4[m - M0];
          i1 = i1 >= n ? i1 - n : i1;
          const Cd wa = rl[idx[m - M0]], wb = rl[i1];
          const double a0 = qc ? wa.y : wa.x, a1 = qc ? wa.x : -wa.y;
          const double a2 = qc ? wb.y : wb.x, a3 = qc ? wb.x : -wb.y;
          acc[m - M0] = mfma_f64(a0, vr[b][0], acc[m - M0]);
          acc[m - M0] = mfma_f64(a1, vr[b][1], acc[m - M0]);
          acc[m - M0] = mfma_f64(a2, vr[b][2], acc[m - M0]);
          acc[m - M0] = mfma_f64(a3, vr[b][3], acc[m - M0]);
          const int nx = idx[m - M0] + inc[m - M0];
          idx[m - M0] = nx >= n ? nx - n : nx;
        }
      }
    }
  }
  if (wave < ns) {
#pragma unroll
    for (int m = M0; m < M1; ++m)
#pragma unroll
      for (int r = 0; r < 4; ++r) red[((wave * MT + m) * 4 + r) * 64 + lane] = acc[m - M0][r];
  }
  __syncthreads();
  if (wave >= ns) {
    double* slot = red + (wave - ns) * MT * 256;
#pragma unroll
    for (int m = M0; m < M1; ++m)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int e = (m * 4 + r) * 64 + lane;
        slot[e] = slot[e] + acc[m - M0][r];
      }
  }
}

// Loads of the wave's kappa blocks: lane group g of block bb holds the
// complex values z = 8 bb + g and z + 4 of its line (row = first element of the
// line).  Straight-line code: the indices of points past the line end (and of
// lanes past the last line, which read line 0) are clamped to a valid element
// and their values replaced by zero, so that every load of the wave can be in
// flight at once.
template <int KBW>
__device__ __forceinline__ void wall_load(v4d (&vr)[KBW], const Cd* __restrict__ src, long row,
                                          bool okl, int b0, int g, int n) {
#pragma unroll
  for (int b = 0; b < KBW; ++b) {
    const int z0 = 8 * (b0 + b) + g;
    const bool ok0 = okl && z0 < n, ok1 = okl && z0 + 4 < n;
    const Cd v0 = src[row + (ok0 ? z0 : 0)], v1 = src[row + (ok1 ? z0 + 4 : 0)];
    vr[b] = v4d{ok0 ? v0.x : 0.0, ok0 ? v0.y : 0.0, ok1 ? v1.x : 0.0, ok1 ? v1.y : 0.0};
  }
}

// The same with the RK update of the stage applied on the load (rk_on_load with
// the line's factors read once).
template <int KBW, bool NL32>
__device__ __forceinline__ void wall_load_rk(v4d (&vr)[KBW], const Cd* __restrict__ src,
                                             const RkLoad& rk, int comp, long row, bool okl,
                                             long ki, int kj, int b0, int g, int n) {
#pragma clang fp contract(off)
  const RkLoadComp& c = rk.c[comp];
  const double kxi = rk.kx[ki], kyj = rk.ky[kj];
  const double kl2 = kxi * kxi + kyj * kyj;
  const double fx = rk.fkx[ki], fy = rk.fky[kj], nui = c.nu[ki];
  const Cd* __restrict__ c0p = static_cast<const Cd*>(c.c0);
  const Cd* __restrict__ fp = static_cast<const Cd*>(c.f);
#pragma unroll
  for (int b = 0; b < KBW; ++b) {
    const int z0 = 8 * (b0 + b) + g;
    Cd out[2];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const bool ok = okl && z0 + 4 * h < n;
      const int z = ok ? z0 + 4 * h : 0;
      const long e = row + z;
      Cd nl;
      if constexpr (NL32) {
        const cpx<float> t = static_cast<const cpx<float>*>(c.nl)[e];
        nl = mkc<double>(double(t.x), double(t.y));
      } else {
        nl = static_cast<const Cd*>(c.nl)[e];
      }
      const Cd u = rk_apply(rk, src[e], nl, fp[e], c0p[e], kl2, rk.kz[z], fx, fy, rk.fkz[z], nui);
      out[h] = ok ? u : mkc<double>(0, 0);
    }
    vr[b] = v4d{out[0].x, out[0].y, out[1].x, out[1].y};
  }
}

// GEMM A over all MT row tiles, G tiles at a time (G accumulators live).
template <int MT, int M0, int G, int KBW, int TAB>
__device__ __forceinline__ void gemm_a_groups(const double* __restrict__ winv, const Cd* rl,
                                              const v4d (&vr)[KBW], int b0, int nbk, int lane,
                                              int wave, int ns, double* red, int n, int np,
                                              int d, int nq) {
  constexpr int M1 = M0 + G < MT ? M0 + G : MT;
  gemm_a_rows<MT, M0, M1, KBW, TAB>(winv, rl, vr, b0, nbk, lane, wave, ns, red, n, np, d, nq);
  if constexpr (M1 < MT)
    gemm_a_groups<MT, M1, G, KBW, TAB>(winv, rl, vr, b0, nbk, lane, wave, ns, red, n, np, d, nq);
}

// MT: 16-row tiles of the 2(2d+C) real partial-IDFT outputs; KT: 4-column
// steps of the 2(2+C) real correction inputs; KBW: 16-row kappa blocks per
// wave; WPE: waves per SIMD the register budget is sized for.
template <int MT, int KT, int KBW, int WPE, int TAB>
__global__ void __launch_bounds__(512, WPE) k_wall_mfma(
    double* __restrict__ vx, double* __restrict__ vy, const Cd* __restrict__ pr, long nlines,
    int ny, int n, int np, int nbk, double invn, WallParams wp, const double* __restrict__ kx,
    const double* __restrict__ ky, const double* __restrict__ dir, int C, int d,
    const double* __restrict__ winv, const double* __restrict__ wfwd,
    const Cd* __restrict__ roots, unsigned magic, RkLoad rk, int rk_on) {
#pragma clang fp contract(off)
  extern __shared__ double sm[];
  const int nw = blockDim.x >> 6;
  const int ns = nw < 4 ? nw : 4;               // partial-sum slots
  Cd* const rl = reinterpret_cast<Cd*>(sm);     // [n] roots of unity (TAB 1)
  double* const red = sm + (TAB ? 2 * n : 0);   // [ns][MT][4][64] partial X
  double* const xs = red;                       // slot 0 ends up holding X
  double* const ds = red + ns * MT * 256;       // [4 KT][16] real delta
  double* const wv = ds + KT * 64;              // [16][4] wall values (re, im bottom, top)
  const int comp = blockIdx.y;
  double* const V = comp ? vy : vx;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int li = lane & 15, g = lane >> 4;
  const long lbase = long(blockIdx.x) * 16;
  const long ln = lbase + li;
  const bool okl = ln < nlines;
  const long lc = okl ? ln : 0;
  const long ki = lc / ny;
  const int kj = int(lc - ki * ny);
  const int b0 = wave * KBW;

  // ---- load the wave's kappa blocks (RK update applied on the load)
  v4d vr[KBW];
  const Cd* const src = reinterpret_cast<const Cd*>(V);
  if (!rk_on)
    wall_load<KBW>(vr, src, lc * n, okl, b0, g, n);
  else if (rk.nl32)
    wall_load_rk<KBW, true>(vr, src, rk, comp, lc * n, okl, ki, kj, b0, g, n);
  else
    wall_load_rk<KBW, false>(vr, src, rk, comp, lc * n, okl, ki, kj, b0, g, n);

  // ---- wall values of the tile's lines (vboundary.f90:195-209; zero for the
  // scalar), staged in LDS here so that the delta phase reads no global memory
  if (threadIdx.x < 16) {
    const long lt = lbase + threadIdx.x;
    Cd wb = mkc<double>(0, 0), wt = mkc<double>(0, 0);
    if (!wp.scalar && lt < nlines) {
      const long ti = lt / ny;
      const int tj = int(lt - ti * ny);
      const double kxi = kx[ti], kyj = ky[tj];
      if (kxi == 0.0 && kyj == 0.0) {
        wb = mkc<double>(wp.nxny * wp.vb[comp][0], 0.0);
        wt = mkc<double>(wp.nxny * wp.vb[comp][1], 0.0);
      } else {
        wb = cscale(ikmul_d(comp ? kyj : kxi, pr[lt * n]), wp.tau);
        wt = cscale(ikmul_d(comp ? kyj : kxi, pr[lt * n + np - 1]), wp.tau);
      }
    }
    double* w = wv + 4 * threadIdx.x;
    w[0] = wb.x;
    w[1] = wb.y;
    w[2] = wt.x;
    w[3] = wt.y;
  }

  // ---- GEMM A: partial inverse DFT at the 2d + C points, one row tile at a
  // time, summed over the waves in a fixed order; then slot 0 += slots
  // 1..ns-1
  const int nq = 2 * d + C;
  if constexpr (TAB == 1) {
    for (int e = threadIdx.x; e < n; e += blockDim.x) rl[e] = roots[e];
    __syncthreads();
  }
  // one 16-row tile at a time: 8 accumulator registers beside the wave's
  // line blocks (three at a time spilled 26 registers per lane at 4 waves per
  // SIMD; HD 512^3: 3.43 -> 3.19 ms per launch with the batched loads)
  constexpr int G = 1;
  gemm_a_groups<MT, 0, G, KBW, TAB>(winv, rl, vr, b0, nbk, lane, wave, ns, red, n, np, d, nq);
  __syncthreads();
  for (int e = threadIdx.x; e < MT * 256; e += blockDim.x) {
    double s = red[e];
    for (int w = 1; w < ns; ++w) s = s + red[w * MT * 256 + e];
    xs[e] = s;
  }
  for (int e = threadIdx.x; e < KT * 64; e += blockDim.x) ds[e] = 0.0;
  __syncthreads();

  // ---- delta at the walls and the continuation points (16 lines x (2 + C))
  // X of point index qi (0..2d+C-1) and line l: real row 2qi, imaginary 2qi+1
  auto xget = [&](int qi, int l) -> Cd {
    auto at = [&](int qr) {
      const int m = qr >> 4, rem = qr & 15;
      return xs[(m * 4 + (rem >> 2)) * 64 + (rem & 3) * 16 + l];
    };
    return mkc<double>(at(2 * qi), at(2 * qi + 1));
  };
  const int nt = 16 * (2 + C);
  const int top = np - d;
  for (int e = threadIdx.x; e < nt; e += blockDim.x) {
    const int l = e / (2 + C), t = e - l * (2 + C);
    const long lt = lbase + l;
    if (lt >= nlines) continue;
    const Cd wb = mkc<double>(wv[4 * l], wv[4 * l + 1]);
    const Cd wt = mkc<double>(wv[4 * l + 2], wv[4 * l + 3]);
    // physical-domain value of boundary point z (0..d-1 or np-d..np-1)
    auto xn = [&](int z) -> Cd {
      if (z == 0) return wb;
      if (z == np - 1) return wt;
      const int qi = z < d ? z : d + (z - top);
      return cscale(xget(qi, l), invn);
    };
    Cd delta;
    if (t == 0) {
      delta = csub(wb, cscale(xget(0, l), invn));
    } else if (t == 1) {
      delta = csub(wt, cscale(xget(2 * d - 1, l), invn));
    } else {
      const int ii = t - 2;
      const double* D0 = dir + ii * d;
      const double* D1 = dir + (C - 1 - ii) * d;
      const Cd xt = xn(top), xb = xn(d - 1);
      double ar = D0[0] * xt.x + D1[0] * xb.x;
      double ai = D0[0] * xt.y + D1[0] * xb.y;
      for (int jj = 1; jj < d; ++jj) {
        const Cd ut = xn(top + jj), ub = xn(d - 1 - jj);
        ar = ar + D0[jj] * ut.x;
        ar = ar + D1[jj] * ub.x;
        ai = ai + D0[jj] * ut.y;
        ai = ai + D1[jj] * ub.y;
      }
      delta = csub(mkc<double>(ar, ai), cscale(xget(2 * d + ii, l), invn));
    }
    ds[(2 * t) * 16 + l] = delta.x;
    ds[(2 * t + 1) * 16 + l] = delta.y;
  }
  __syncthreads();

  // ---- GEMM C: Vr += Wf * delta, then store
  if constexpr (TAB == 0) {
#pragma unroll
    for (int b = 0; b < KBW; ++b) {
      const int bb = b0 + b;
      if (bb < nbk) {
        const double* wa = wfwd + size_t(bb) * KT * 64 + lane;
#pragma unroll
        for (int t = 0; t < KT; ++t)
          vr[b] = mfma_f64(wa[t * 64], ds[(4 * t + g) * 16 + li], vr[b]);
      }
    }
  } else {
    // lane = i + 16k: row i = kappa offset kap_off(i&3, i>>2) of block bb
    // (z = 8 bb + (i&3) + 4(i>>3), re/im (i>>2)&1); column 4t + k = (point
    // p = 2t + k/2, re/im k&1); e^{-2 pi i z r_p / n}
    const int zo = (li & 3) + 4 * (li >> 3), zc = (li >> 2) & 1, pc = g & 1;
    const int nrho = 2 + C;
    // K step outer, kappa block inner: consecutive MFMAs accumulate into
    // different blocks, and one B operand (delta) is live at a time
#pragma unroll
    for (int t = 0; t < KT; ++t) {
      const double bt = ds[(4 * t + g) * 16 + li];
      const unsigned r = unsigned(rpoint(2 * t + (g >> 1), np, nrho));
#pragma unroll
      for (int b = 0; b < KBW; ++b) {
        const int bb = b0 + b;
        if (bb < nbk) {
          const unsigned z = unsigned(8 * bb + zo);
          const Cd w = rl[modn(z * r, n, magic)];
          const double a = zc == pc ? w.x : (zc ? -w.y : w.y);
          vr[b] = mfma_f64(a, bt, vr[b]);
        }
      }
    }
  }
  if (okl) {
    Cd* dst = reinterpret_cast<Cd*>(V) + ln * n;
#pragma unroll
    for (int b = 0; b < KBW; ++b) {
      const int bb = b0 + b;
      const int z0 = 8 * bb + g;
      if (bb < nbk) {
        if (z0 < n) dst[z0] = mkc<double>(vr[b][0], vr[b][1]);
        if (z0 + 4 < n) dst[z0 + 4] = mkc<double>(vr[b][2], vr[b][3]);
      }
    }
  }
}

constexpr int kMaxWaves = 8;

}  // namespace

// Tables of the plan's wall correction (float64 plans' and mixed plans' state
// are float64 alike): fragments of WinvT and Wf in the lane order of the MFMA
// operands, formed in long double.
int build_wall_mfma(spx_plan* p) {
  WallMfmaHost& w = p->wall;
  w = WallMfmaHost();
  const char* sw = std::getenv("SPX_WALL_MFMA");
  if (p->g.Cz <= 0 || (sw && std::strcmp(sw, "0") == 0)) return SPX_OK;
  const int n = p->g.nz, np = p->np, C = p->g.Cz, d = p->g.oz;
  const int nq = 2 * d + C, nrho = 2 + C;
  const int mt = (2 * nq + 15) / 16, kt = (2 * nrho + 3) / 4;
  const int nbk = (2 * n + 15) / 16;
  if (!((mt == 5 || mt == 6) && kt <= 16 && np >= 2 * d)) return SPX_OK;
  // waves per workgroup and kappa blocks per wave: the fewest registers that
  // keep the workgroup within 8 waves
  int kbw = 9;
  if ((nbk + kbw - 1) / kbw > kMaxWaves) kbw = 18;
  if ((nbk + kbw - 1) / kbw > kMaxWaves) return SPX_OK;
  WallTune tune;
  const bool tuned = get_wall_tune(n, &tune);
  if (tuned && (tune.kbw == 9 || tune.kbw == 18) && (nbk + tune.kbw - 1) / tune.kbw <= kMaxWaves)
    kbw = tune.kbw;
  if (const char* e = std::getenv("SPX_WALL_KBW")) {
    const int v = std::atoi(e);
    if ((v == 9 || v == 18) && (nbk + v - 1) / v <= kMaxWaves) kbw = v;
  }
  // register budget: 3 or 4 waves per SIMD for 9 blocks per wave (4: two
  // 8-wave workgroups per CU overlap their load and matrix phases; HD 512^3:
  // 3.64 ms per launch against 5.31 ms with 3), 2 for 18
  int wpe = 4;
  if (tuned && (tune.wpe == 3 || tune.wpe == 4)) wpe = tune.wpe;
  if (const char* e = std::getenv("SPX_WALL_WPE")) {
    const int v = std::atoi(e);
    if (v == 3 || v == 4) wpe = v;
  }
  w.wpe = kbw == 18 ? 2 : wpe;
  w.nbk = nbk;
  w.kbw = kbw;
  w.nw = (nbk + kbw - 1) / kbw;
  w.mt = mt;
  w.kt = kt;
  // points of the partial inverse DFT and of the correction
  std::vector<int> qz(nq), rz(nrho);
  for (int i = 0; i < d; ++i) qz[i] = i;
  for (int i = 0; i < d; ++i) qz[d + i] = np - d + i;
  for (int i = 0; i < C; ++i) qz[2 * d + i] = np + i;
  rz[0] = 0;
  rz[1] = np - 1;
  for (int i = 0; i < C; ++i) rz[2 + i] = np + i;
  const long double PI2 = 6.283185307179586476925286766559005768L;
  std::vector<double> cr(n), ci(n);
  for (int m = 0; m < n; ++m) {
    cr[m] = double(cosl(PI2 * (long double)m / (long double)n));
    ci[m] = double(sinl(PI2 * (long double)m / (long double)n));
  }
  // WinvT[qr][kappa]: x_q = sum_z V_z e^{+2 pi i z q / n}
  auto winvT = [&](int qr, int kap) -> double {
    const int qi = qr >> 1, qc = qr & 1, z = kap >> 1, zc = kap & 1;
    if (qi >= nq || z >= n) return 0.0;
    const int m = int((long long)z * qz[qi] % n);
    const double wr = cr[m], wi = ci[m];
    if (qc == 0) return zc == 0 ? wr : -wi;
    return zc == 0 ? wi : wr;
  };
  // Wf[kappa][rho]: V'_z += sum_p delta_p e^{-2 pi i z r_p / n}
  auto wf = [&](int kap, int rho) -> double {
    const int z = kap >> 1, zc = kap & 1, pi = rho >> 1, pc = rho & 1;
    if (pi >= nrho || z >= n) return 0.0;
    const int m = int((long long)z * rz[pi] % n);
    const double er = cr[m], ei = -ci[m];
    if (zc == 0) return pc == 0 ? er : -ei;
    return pc == 0 ? ei : er;
  };
  std::vector<double> ta(size_t(nbk) * 4 * mt * 64), tc(size_t(nbk) * kt * 64);
  for (int bb = 0; bb < nbk; ++bb)
    for (int s = 0; s < 4; ++s)
      for (int m = 0; m < mt; ++m)
        for (int l = 0; l < 64; ++l)
          ta[((size_t(bb) * 4 + s) * mt + m) * 64 + l] =
              winvT(16 * m + (l & 15), 16 * bb + kap_off(l >> 4, s));
  for (int bb = 0; bb < nbk; ++bb)
    for (int t = 0; t < kt; ++t)
      for (int l = 0; l < 64; ++l)
        tc[(size_t(bb) * kt + t) * 64 + l] =
            wf(16 * bb + kap_off((l & 15) & 3, (l & 15) >> 2), 4 * t + (l >> 4));
  std::vector<double> rt(2 * size_t(n));
  for (int m = 0; m < n; ++m) {
    rt[2 * m] = cr[m];
    rt[2 * m + 1] = ci[m];
  }
  w.magic = unsigned((uint64_t(1) << 32) / uint64_t(n) + 1);
  w.tab = 1;
  if (const char* e = std::getenv("SPX_WALL_TAB")) w.tab = std::atoi(e) == 0 ? 0 : 1;
  double** dst[3] = {&w.winv, &w.wfwd, &w.roots};
  const std::vector<double>* src[3] = {&ta, &tc, &rt};
  for (int i = 0; i < 3; ++i) {
    const size_t bytes = src[i]->size() * sizeof(double);
    void* dptr = nullptr;
    if (hipMalloc(&dptr, bytes) != hipSuccess) return fail(SPX_ERR_ALLOC, "wall correction tables");
    p->owned.push_back(dptr);
    p->dev_bytes += bytes;
    SPX_HIP(hipMemcpy(dptr, src[i]->data(), bytes, hipMemcpyHostToDevice));
    *dst[i] = static_cast<double*>(dptr);
  }
  w.on = true;
  return SPX_OK;
}

#define SPX_WALL_LAUNCH(MT, KT, KBW, WPE)                                                     \
  if (w.tab) SPX_WALL_LAUNCH_T(MT, KT, KBW, WPE, 1) else SPX_WALL_LAUNCH_T(MT, KT, KBW, WPE, 0)
#define SPX_WALL_LAUNCH_T(MT, KT, KBW, WPE, TAB)                                              \
  hipLaunchKernelGGL((k_wall_mfma<MT, KT, KBW, WPE, TAB>), grid, dim3(64 * w.nw), lds, p->stream, \
                     static_cast<double*>(vx), static_cast<double*>(vy),                      \
                     static_cast<const Cd*>(pr), nlines, p->g.ny, p->g.nz, p->np, w.nbk,      \
                     1.0 / double(p->g.nz), wp, p->d_kx, p->d_ky, p->d_dir, p->g.Cz, p->g.oz, \
                     w.winv, w.wfwd, reinterpret_cast<const Cd*>(w.roots), w.magic, r,     \
                     rk ? 1 : 0);

int wall_mfma(spx_plan* p, void* vx, void* vy, const void* pr, const WallParams& wp, int ncomp,
              const RkLoad* rk) {
  const WallMfmaHost& w = p->wall;
  if (!w.on) return SPX_ERR_UNSUPPORTED;
  const long nlines = p->wlines();
  const dim3 grid(unsigned((nlines + 15) / 16), unsigned(ncomp));
  const size_t lds = (size_t(std::min(w.nw, 4)) * w.mt * 256 + size_t(w.kt) * 64 + 64 +
                      (w.tab ? 2 * size_t(p->g.nz) : 0)) *
                     sizeof(double);
  RkLoad r;
  std::memset(&r, 0, sizeof(r));
  if (rk) r = *rk;
  const double extra = rk ? (rk->nl32 ? 8.0 : 16.0) + 32.0 : 0.0;  // nl, c0, f reads
  ProfScope ps_(p, rk ? "noslip_rk" : (wp.scalar ? "sconst" : "noslip"),
                ncomp * double(p->wnc()) * (2.0 * 16.0 + extra));
  // the partial-IDFT tile count follows (d, C): 5 for d = 5, C = 25 (70 rows),
  // 6 for d = 8, C = 27 (86 rows); the correction steps 14 / 15
  const int key = w.mt * 10000 + w.kt * 100 + w.kbw * 10 / 9 + w.wpe;  // kbw 9 -> 10, 18 -> 20
  switch (key) {
    case 51413: SPX_WALL_LAUNCH(5, 14, 9, 3); break;
    case 51414: SPX_WALL_LAUNCH(5, 14, 9, 4); break;
    case 51422: SPX_WALL_LAUNCH(5, 14, 18, 2); break;
    case 61513: SPX_WALL_LAUNCH(6, 15, 9, 3); break;
    case 61514: SPX_WALL_LAUNCH(6, 15, 9, 4); break;
    case 61522: SPX_WALL_LAUNCH(6, 15, 18, 2); break;
    default: return SPX_ERR_UNSUPPORTED;
  }
  SPX_CHECK_LAUNCH();
  return SPX_OK;
}

}  // namespace spx
