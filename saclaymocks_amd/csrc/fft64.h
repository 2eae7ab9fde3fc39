// FP64 FFTs for the fit kernels on gfx950: a register-resident Stockham FFT
// of M = 2^m points (P threads, E = M / P points each, LDS exchange between
// radix stages) and, on top of it, Bluestein's chirp-z algorithm for a DFT
// of any length N with 2N - 1 <= M:
//   X_k = sum_n y_n exp(-2 pi i k n / N)
//       = w_k sum_n (y_n w_n) conj(w_{k-n}),   w_n = exp(-i pi n^2 / N),
// i.e. a circular convolution of length M done with three M-point FFTs
// (L. Bluestein, IEEE Trans. Audio Electroacoust. 18, 451, 1970).
// The chirp's phase n^2 mod 2N is formed exactly in integers, so w_n is
// accurate to FP64 rounding for any n.
//
// This file is device-only; it is included by fit.hip.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace smk {
namespace f64 {

__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(fma(a.x, b.x, -a.y * b.y), fma(a.x, b.y, a.y * b.x));
}
__device__ __forceinline__ double2 cconj(double2 a) { return make_double2(a.x, -a.y); }
// times -i (forward) or +i (inverse)
template <bool INV> __device__ __forceinline__ double2 mul_mi(double2 a) {
  return INV ? make_double2(-a.y, a.x) : make_double2(a.y, -a.x);
}

template <int R, bool INV> struct DFT;
template <bool INV> struct DFT<2, INV> {
  static __device__ __forceinline__ void run(double2* a) {
    const double2 t = a[0];
    a[0] = cadd(t, a[1]);
    a[1] = csub(t, a[1]);
  }
};
template <bool INV> struct DFT<4, INV> {
  static __device__ __forceinline__ void run(double2* a) {
    const double2 t0 = cadd(a[0], a[2]), t1 = csub(a[0], a[2]);
    const double2 t2 = cadd(a[1], a[3]), t3 = mul_mi<INV>(csub(a[1], a[3]));
    a[0] = cadd(t0, t2);
    a[2] = csub(t0, t2);
    a[1] = cadd(t1, t3);
    a[3] = csub(t1, t3);
  }
};
template <bool INV> struct DFT<8, INV> {
  static __device__ __forceinline__ void run(double2* a) {
    const double r = 0.70710678118654752440084436210484904;
    double2 e[4] = {a[0], a[2], a[4], a[6]};
    double2 o[4] = {a[1], a[3], a[5], a[7]};
    DFT<4, INV>::run(e);
    DFT<4, INV>::run(o);
    double2 o1, o3;
    if (INV) {
      o1 = make_double2(r * (o[1].x - o[1].y), r * (o[1].x + o[1].y));
      o3 = make_double2(-r * (o[3].x + o[3].y), r * (o[3].x - o[3].y));
    } else {
      o1 = make_double2(r * (o[1].x + o[1].y), r * (o[1].y - o[1].x));
      o3 = make_double2(r * (o[3].y - o[3].x), -r * (o[3].x + o[3].y));
    }
    const double2 o2 = mul_mi<INV>(o[2]);
    a[0] = cadd(e[0], o[0]);
    a[4] = csub(e[0], o[0]);
    a[1] = cadd(e[1], o1);
    a[5] = csub(e[1], o1);
    a[2] = cadd(e[2], o2);
    a[6] = csub(e[2], o2);
    a[3] = cadd(e[3], o3);
    a[7] = csub(e[3], o3);
  }
};
// 16 = 4 x 4: four 4-point DFTs on stride-4 inputs, twiddles W16^(q r), four
// 4-point DFTs across (six-step decomposition inside the registers)
template <bool INV> struct DFT<16, INV> {
  static __device__ __forceinline__ void run(double2* a) {
    const double c1 = 0.92387953251128675612818318939678829;  // cos(pi/8)
    const double s1 = 0.38268343236508977172845998403039887;  // sin(pi/8)
    const double h = 0.70710678118654752440084436210484904;
    double2 t[4][4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      double2 u[4] = {a[q], a[q + 4], a[q + 8], a[q + 12]};
      DFT<4, INV>::run(u);
#pragma unroll
      for (int r = 0; r < 4; ++r) t[q][r] = u[r];
    }
    // W16^(q r) for q, r in 1..3 (sign of the imaginary part by direction)
    const double sg = INV ? 1.0 : -1.0;
    const double2 w[4][4] = {
        {{1, 0}, {1, 0}, {1, 0}, {1, 0}},
        {{1, 0}, {c1, sg * s1}, {h, sg * h}, {s1, sg * c1}},
        {{1, 0}, {h, sg * h}, {0, sg}, {-h, sg * h}},
        {{1, 0}, {s1, sg * c1}, {-h, sg * h}, {-c1, -sg * s1}}};
#pragma unroll
    for (int q = 1; q < 4; ++q)
#pragma unroll
      for (int r = 1; r < 4; ++r) t[q][r] = cmul(t[q][r], w[q][r]);
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      double2 u[4] = {t[0][r], t[1][r], t[2][r], t[3][r]};
      DFT<4, INV>::run(u);
#pragma unroll
      for (int q = 0; q < 4; ++q) a[r + 4 * q] = u[q];
    }
  }
};

// radix of stage S of an M-point FFT with E points per thread: the largest
// of 16, 8, 4, 2 dividing both E and what remains of M
__host__ __device__ constexpr int stage_radix(int M, int E, int S) {
  int rem = M;
  int r = 0;
  for (int s = 0; s <= S; ++s) {
    r = 0;
    const int cand[4] = {16, 8, 4, 2};
    for (int i = 0; i < 4; ++i)
      if (E % cand[i] == 0 && rem % cand[i] == 0) { r = cand[i]; break; }
    if (r == 0) return 0;
    rem /= r;
  }
  return r;
}
__host__ __device__ constexpr int n_stages(int M, int E) {
  int rem = M, s = 0;
  while (rem > 1) {
    const int r = stage_radix(M, E, s);
    if (r == 0) return -1;
    rem /= r;
    ++s;
  }
  return s;
}
// LDS slot of point `pos`: one pad slot per 32 points (double2 = 4 banks)
__device__ __forceinline__ int slot(int pos) { return pos + (pos >> 5); }
__host__ __device__ constexpr int line_len(int M) { return M + M / 32 + 1; }

// Twiddle tables of the Stockham stages: a stage of radix R after Ns
// points per sub-transform needs W_(Ns R)^(r jj), r < R, jj < Ns; these are
// stored per (R, Ns) as [r][jj] (so the lanes of a wave, consecutive jj,
// read consecutive entries), after the TW-point table W_TW^k.  Block of
// radix R at stage_tab_base(R): the tables of Ns = 1, 2, 4, ... TW / R,
// Ns R entries each.  Every value is W_TW^(r jj TW / (Ns R)), the same
// double as the TW-point table holds.
__host__ __device__ constexpr int stage_tab_base(int R, int TW) {
  int base = 0;
  for (int q = 2; q < R; q *= 2) base += 2 * TW - q;
  return base;
}
__host__ __device__ constexpr int stage_tab_len(int TW) { return stage_tab_base(32, TW); }
__host__ __device__ constexpr int stage_tab_off(int R, int Ns, int TW) { return stage_tab_base(R, TW) + R * (Ns - 1); }

// Twiddles of a Stockham stage for this thread's butterflies j = p + P b:
// w[b][r] = W_M^(r (j % Ns) M/(Ns R)), r >= 1 (conjugated for the inverse).
// twM: W_TW^k, k < TW (TW = 4096 >= M), then the stage tables above.
template <int M, int E, int R, int Ns, bool INV>
__device__ __forceinline__ void stage_twiddles(double2 (&w)[E / R][R], int p, const double2* __restrict__ twM,
                                               int TW) {
  constexpr int P = M / E;
  constexpr int B = E / R;
  static_assert(Ns * R <= 4096, "stage tables hold sub-transforms up to 4096 points");
  const double2* __restrict__ tab = twM + TW + stage_tab_off(R, Ns, 4096);
#pragma unroll
  for (int b = 0; b < B; ++b) {
    const int jj = (p + P * b) & (Ns - 1);
    w[b][0] = make_double2(1.0, 0.0);
#pragma unroll
    for (int r = 1; r < R; ++r) {
      double2 t = tab[r * Ns + jj];
      if (INV) t.y = -t.y;
      w[b][r] = t;
    }
  }
}

// Stockham stage: butterfly j = p + P b takes points j + r M/R (register
// slots b + r B), multiplies point r by w[b][r] (stage_twiddles; none for
// Ns = 1), outputs to points (j / Ns) Ns R + j % Ns + r Ns (the last stage
// keeps them in registers: slot b + r B holds point j + r M/R).
template <int M, int E, int R, int Ns, bool INV, bool LAST>
__device__ __forceinline__ void stage(double2 (&v)[E], double2* __restrict__ line, int p,
                                      const double2 (&w)[E / R][R]) {
  constexpr int P = M / E;
  constexpr int B = E / R;
#pragma unroll
  for (int b = 0; b < B; ++b) {
    const int j = p + P * b;
    double2 a[R];
#pragma unroll
    for (int r = 0; r < R; ++r) a[r] = v[b + r * B];
    if (Ns > 1) {
#pragma unroll
      for (int r = 1; r < R; ++r) a[r] = cmul(a[r], w[b][r]);
    }
    DFT<R, INV>::run(a);
    if (LAST) {
#pragma unroll
      for (int r = 0; r < R; ++r) v[b + r * B] = a[r];
    } else {
      const int d0 = (j / Ns) * Ns * R + (j & (Ns - 1));
#pragma unroll
      for (int r = 0; r < R; ++r) line[slot(d0 + r * Ns)] = a[r];
    }
  }
}

// The stages from S on; w holds stage S's twiddles.  The next stage's
// twiddles are loaded before the exchange's barriers, so that their latency
// overlaps the exchange.
template <int M, int E, bool INV, int S, int Ns>
__device__ __forceinline__ void stages(double2 (&v)[E], double2* __restrict__ line, int p,
                                       const double2* __restrict__ twM, int TW,
                                       const double2 (&w)[E / stage_radix(M, E, S)][stage_radix(M, E, S)]) {
  constexpr int NS = n_stages(M, E);
  static_assert(NS > 0, "unsupported FP64 FFT size");
  constexpr int P = M / E;
  if constexpr (S < NS) {
    constexpr int R = stage_radix(M, E, S);
    constexpr bool LAST = S + 1 == NS;
    stage<M, E, R, Ns, INV, LAST>(v, line, p, w);
    if constexpr (!LAST) {
      constexpr int R1 = stage_radix(M, E, S + 1);
      double2 w1[E / R1][R1];
      stage_twiddles<M, E, R1, Ns * R, INV>(w1, p, twM, TW);
      __syncthreads();
#pragma unroll
      for (int e = 0; e < E; ++e) v[e] = line[slot(p + P * e)];
      __syncthreads();
      stages<M, E, INV, S + 1, Ns * R>(v, line, p, twM, TW, w1);
    }
  }
}

// In-register FFT of one M-point line over the M / E threads p of the block
// (every thread of the block must call it: it holds block barriers).  On
// entry and exit thread p holds points p + P e.  Unnormalised.
template <int M, int E, bool INV>
__device__ __forceinline__ void fft(double2 (&v)[E], double2* __restrict__ line, int p,
                                    const double2* __restrict__ twM, int TW) {
  constexpr int R0 = stage_radix(M, E, 0);
  double2 w0[E / R0][R0];   // stage 0 has Ns = 1: no twiddles
  stages<M, E, INV, 0, 1>(v, line, p, twM, TW, w0);
}

// w_n = exp(-i pi n^2 / N), n^2 mod 2N exact (n < 2^31 / N)
__device__ __forceinline__ double2 chirp(int64_t n, int N) {
  const int64_t q = (n * n) % (2 * (int64_t)N);
  double s, c;
  sincospi(-(double)q / (double)N, &s, &c);
  return make_double2(c, s);
}

// Bluestein DFT of length N (2N - 1 <= M) over a block of M / E threads:
// in(n) gives y_n for n < N (called once per n by one thread); emit(k, X_k)
// is called for k < nout (nout <= N) by the thread that holds it.  `line`
// is M-point LDS scratch (line_len(M) double2); every thread of the block
// must call it.  Thread p computes the chirp w_n of its own points n = p + P e
// once and keeps it in registers; the wrapped half of b (m > M - N needs
// w_{M-m}, a point of another thread) goes through the line.
template <int M, int E, class In, class Emit>
__device__ __forceinline__ void bluestein(int N, int nout, In in, Emit emit, double2* __restrict__ line,
                                          const double2* __restrict__ twM, int TW) {
  constexpr int P = M / E;
  const int p = threadIdx.x;
  double2 ch[E], vb[E], va[E];
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int n = p + P * e;
    ch[e] = n < N ? chirp(n, N) : make_double2(0.0, 0.0);
    line[slot(n)] = ch[e];
  }
  __syncthreads();
  // B = FFT(b), b_m = conj(w_m) for |m| < N (circularly), 0 elsewhere
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int m = p + P * e;
    vb[e] = m < N ? cconj(ch[e]) : (m > M - N ? cconj(line[slot(M - m)]) : make_double2(0.0, 0.0));
  }
  __syncthreads();
  fft<M, E, false>(vb, line, p, twM, TW);
  __syncthreads();
  // A = FFT(a), a_n = y_n w_n for n < N
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int n = p + P * e;
    va[e] = n < N ? cmul(in(n), ch[e]) : make_double2(0.0, 0.0);
  }
  fft<M, E, false>(va, line, p, twM, TW);
  __syncthreads();
#pragma unroll
  for (int e = 0; e < E; ++e) va[e] = cmul(va[e], vb[e]);
  fft<M, E, true>(va, line, p, twM, TW);
  const double inv = 1.0 / (double)M;
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int k = p + P * e;
    if (k < nout) {
      const double2 c = cmul(va[e], ch[e]);
      emit(k, make_double2(c.x * inv, c.y * inv));
    }
  }
  __syncthreads();
}

// The per-length table of bluestein_tab, written at T: w_n for n < N, then
// FFT_M(b)_k / M for k < M with b as above (1/M is a power of two, so the
// products below round exactly as bluestein's).  Every thread of the block
// must call it.
template <int M, int E>
__device__ __forceinline__ void bluestein_table(int N, double2* __restrict__ T, double2* __restrict__ line,
                                                const double2* __restrict__ twM, int TW) {
  constexpr int P = M / E;
  const int p = threadIdx.x;
  double2 vb[E];
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int m = p + P * e;
    if (m < N) {
      const double2 w = chirp(m, N);
      T[m] = w;
      vb[e] = cconj(w);
    } else {
      vb[e] = m > M - N ? cconj(chirp(M - m, N)) : make_double2(0.0, 0.0);
    }
  }
  fft<M, E, false>(vb, line, p, twM, TW);
  const double inv = 1.0 / (double)M;
#pragma unroll
  for (int e = 0; e < E; ++e) T[N + p + P * e] = make_double2(vb[e].x * inv, vb[e].y * inv);
}

// Bluestein with the chirp w (N values) and FFT_M(b) / M (M values) read
// from a table made once per length (bluestein_table): one register tile,
// two transforms.  in(n) may read LDS that `line` aliases (all of it is read
// before the first exchange); emit may write there too (the line is free
// once the last transform's final stage runs in registers).
// KEEP bit 0: the chirp's transform is loaded before the first FFT (its
// latency overlaps that FFT); bit 1: the chirp is kept from the input for
// the output.  2 E more registers each, for kernels whose occupancy does not
// drop by them.
constexpr int KEEP_B = 1, KEEP_W = 2;
template <int M, int E, int KEEP = 0, class In, class Emit>
__device__ __forceinline__ void bluestein_tab(int N, int nout, In in, Emit emit, double2* __restrict__ line,
                                              const double2* __restrict__ twM, int TW,
                                              const double2* __restrict__ wtab, const double2* __restrict__ btab) {
  constexpr int P = M / E;
  const int p = threadIdx.x;
  double2 va[E];
  [[maybe_unused]] double2 wk[E], bk[E];
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int n = p + P * e;
    if constexpr ((KEEP & KEEP_W) != 0) {
      wk[e] = n < N ? wtab[n] : make_double2(0.0, 0.0);
      va[e] = n < N ? cmul(in(n), wk[e]) : make_double2(0.0, 0.0);
    } else {
      va[e] = n < N ? cmul(in(n), wtab[n]) : make_double2(0.0, 0.0);
    }
  }
  if constexpr ((KEEP & KEEP_B) != 0) {
#pragma unroll
    for (int e = 0; e < E; ++e) bk[e] = btab[p + P * e];
  }
  __syncthreads();
  fft<M, E, false>(va, line, p, twM, TW);
#pragma unroll
  for (int e = 0; e < E; ++e) va[e] = cmul(va[e], (KEEP & KEEP_B) ? bk[e] : btab[p + P * e]);
  fft<M, E, true>(va, line, p, twM, TW);
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int k = p + P * e;
    if (k < nout) emit(k, cmul(va[e], (KEEP & KEEP_W) ? wk[e] : wtab[k]));
  }
}

// Two independent M-point FFTs (va on line la, vb on line lb) in lockstep:
// each stage's butterflies of both, then one pair of barriers around both
// exchanges, so the two transforms share every barrier and every twiddle.
template <int M, int E, bool INV, int S, int Ns>
__device__ __forceinline__ void stages2(double2 (&va)[E], double2 (&vb)[E], double2* __restrict__ la,
                                        double2* __restrict__ lb, int p, const double2* __restrict__ twM, int TW,
                                        const double2 (&w)[E / stage_radix(M, E, S)][stage_radix(M, E, S)]) {
  constexpr int NS = n_stages(M, E);
  static_assert(NS > 0, "unsupported FP64 FFT size");
  constexpr int P = M / E;
  if constexpr (S < NS) {
    constexpr int R = stage_radix(M, E, S);
    constexpr bool LAST = S + 1 == NS;
    stage<M, E, R, Ns, INV, LAST>(va, la, p, w);
    stage<M, E, R, Ns, INV, LAST>(vb, lb, p, w);
    if constexpr (!LAST) {
      constexpr int R1 = stage_radix(M, E, S + 1);
      double2 w1[E / R1][R1];
      stage_twiddles<M, E, R1, Ns * R, INV>(w1, p, twM, TW);
      __syncthreads();
#pragma unroll
      for (int e = 0; e < E; ++e) {
        va[e] = la[slot(p + P * e)];
        vb[e] = lb[slot(p + P * e)];
      }
      __syncthreads();
      stages2<M, E, INV, S + 1, Ns * R>(va, vb, la, lb, p, twM, TW, w1);
    }
  }
}

template <int M, int E, bool INV>
__device__ __forceinline__ void fft2(double2 (&va)[E], double2 (&vb)[E], double2* __restrict__ la,
                                     double2* __restrict__ lb, int p, const double2* __restrict__ twM, int TW) {
  constexpr int R0 = stage_radix(M, E, 0);
  double2 w0[E / R0][R0];   // stage 0 has Ns = 1: no twiddles
  stages2<M, E, INV, 0, 1>(va, vb, la, lb, p, twM, TW, w0);
}

// bluestein_tab for two inputs in lockstep on the lines la and lb, of
// lengths Na and Nb with the tables (wa, ba) and (wb, bb) (SAME: one length,
// one table, (wa, ba), read once): in(n), n < max(Na, Nb), gives point n of
// both inputs as (x.re, x.im, y.re, y.im), the first pair used for n < Na,
// the second for n < Nb; emit_a(k, X_k) for k < Na and emit_b(k, Y_k) for
// k < Nb.  The emits may write the first Na (Nb) points of la (lb).
// KEEP_B (SAME only): the chirp's transform loaded before the first FFTs.
template <int M, int E, bool SAME, int KEEP = 0, class In, class EmitA, class EmitB>
__device__ __forceinline__ void bluestein_tab2(int Na, int Nb, In in, EmitA emit_a, EmitB emit_b,
                                               double2* __restrict__ la, double2* __restrict__ lb,
                                               const double2* __restrict__ twM, int TW,
                                               const double2* __restrict__ wa, const double2* __restrict__ ba,
                                               const double2* __restrict__ wb, const double2* __restrict__ bb) {
  constexpr int P = M / E;
  const int p = threadIdx.x;
  const int nmax = Na > Nb ? Na : Nb;
  double2 va[E], vb[E];
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int n = p + P * e;
    va[e] = vb[e] = make_double2(0.0, 0.0);
    if (n < nmax) {
      const double4 xy = in(n);
      if (SAME) {
        const double2 w = wa[n];
        va[e] = cmul(make_double2(xy.x, xy.y), w);
        vb[e] = cmul(make_double2(xy.z, xy.w), w);
      } else {
        if (n < Na) va[e] = cmul(make_double2(xy.x, xy.y), wa[n]);
        if (n < Nb) vb[e] = cmul(make_double2(xy.z, xy.w), wb[n]);
      }
    }
  }
  static_assert(KEEP == 0 || SAME, "KEEP_B reads one table");
  [[maybe_unused]] double2 bk[E];
  if constexpr ((KEEP & KEEP_B) != 0) {
#pragma unroll
    for (int e = 0; e < E; ++e) bk[e] = ba[p + P * e];
  }
  __syncthreads();
  fft2<M, E, false>(va, vb, la, lb, p, twM, TW);
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const double2 x = (KEEP & KEEP_B) ? bk[e] : ba[p + P * e];
    va[e] = cmul(va[e], x);
    vb[e] = cmul(vb[e], SAME ? x : bb[p + P * e]);
  }
  fft2<M, E, true>(va, vb, la, lb, p, twM, TW);
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int k = p + P * e;
    if (SAME) {
      if (k < Na) {
        const double2 w = wa[k];
        emit_a(k, cmul(va[e], w));
        emit_b(k, cmul(vb[e], w));
      }
    } else {
      if (k < Na) emit_a(k, cmul(va[e], wa[k]));
      if (k < Nb) emit_b(k, cmul(vb[e], wb[k]));
    }
  }
}

}  // namespace f64
}  // namespace smk
