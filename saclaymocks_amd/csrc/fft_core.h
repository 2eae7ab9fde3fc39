// Register-resident Stockham FFT engine for gfx950 (64-wide wavefronts).
//
// One line of N complex points is owned by P threads; each thread keeps
// E = N / P points in VGPRs.  Every radix stage is done in registers and the
// permutation between stages goes through LDS (one write, one read per point).
// Thread p always holds, before the first stage and after the last one, the
// points n = p + P*e (e = 0..E-1): the global load of the first stage and the
// global store of the last stage are therefore unit-stride across p, which is
// what makes the HBM side coalesced whatever the line stride is.
//
// Twiddles come from a per-length table W_N^m = exp(-2 pi i m / N), m < N,
// computed in FP64 on the host and rounded once to FP32.
//
// This file is device-only; it is included by the box and skewer kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace smk {

// ------------------------------------------------------- unfused arithmetic
// One IEEE operation each, which the compiler may not fuse with a neighbour
// into an FMA.  HIP's __fadd_rn / __dmul_rn and friends are plain operators,
// and hipcc contracts a product feeding a sum into one FMA even across those
// calls; these are for results that must match numpy's (or numba's) unfused
// evaluation bit for bit: k of a mode, pixel positions, cell indices.
#define SMK_UNFUSED(T, NAME, OP)                                  \
  __device__ __forceinline__ T NAME(T a, T b) {                   \
    _Pragma("clang fp contract(off)") return a OP b;              \
  }
SMK_UNFUSED(float, add_rn, +)
SMK_UNFUSED(float, sub_rn, -)
SMK_UNFUSED(float, mul_rn, *)
SMK_UNFUSED(double, add_rn, +)
SMK_UNFUSED(double, sub_rn, -)
SMK_UNFUSED(double, mul_rn, *)
SMK_UNFUSED(double, div_rn, /)
#undef SMK_UNFUSED

// ---------------------------------------------------------------- complex ops
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
__device__ __forceinline__ float2 cconj(float2 a) { return make_float2(a.x, -a.y); }
// multiply by -i (forward transform) or +i (inverse transform)
template <bool INV> __device__ __forceinline__ float2 mul_mi(float2 a) {
  return INV ? make_float2(-a.y, a.x) : make_float2(a.y, -a.x);
}

// ------------------------------------------------------------ radix kernels
template <int R, bool INV> struct DFT;

template <bool INV> struct DFT<2, INV> {
  static __device__ __forceinline__ void run(float2* a) {
    float2 t = a[0];
    a[0] = cadd(t, a[1]);
    a[1] = csub(t, a[1]);
  }
};

template <bool INV> struct DFT<3, INV> {
  static __device__ __forceinline__ void run(float2* a) {
    const float s = 0.866025403784438646763723f;  // sin(2 pi / 3)
    float2 b = cadd(a[1], a[2]);
    float2 d = csub(a[1], a[2]);
    float2 y0 = cadd(a[0], b);
    float2 t = make_float2(a[0].x - 0.5f * b.x, a[0].y - 0.5f * b.y);
    float2 u = mul_mi<INV>(cscale(d, s));
    a[0] = y0;
    a[1] = cadd(t, u);
    a[2] = csub(t, u);
  }
};

template <bool INV> struct DFT<4, INV> {
  static __device__ __forceinline__ void run(float2* a) {
    float2 t0 = cadd(a[0], a[2]);
    float2 t1 = csub(a[0], a[2]);
    float2 t2 = cadd(a[1], a[3]);
    float2 t3 = mul_mi<INV>(csub(a[1], a[3]));
    a[0] = cadd(t0, t2);
    a[2] = csub(t0, t2);
    a[1] = cadd(t1, t3);
    a[3] = csub(t1, t3);
  }
};

template <bool INV> struct DFT<5, INV> {
  static __device__ __forceinline__ void run(float2* a) {
    const float c1 = 0.309016994374947424102293f;   // cos(2 pi / 5)
    const float c2 = -0.809016994374947424102293f;  // cos(4 pi / 5)
    const float s1 = 0.951056516295153572116439f;   // sin(2 pi / 5)
    const float s2 = 0.587785252292473129168706f;   // sin(4 pi / 5)
    float2 b1 = cadd(a[1], a[4]);
    float2 b2 = cadd(a[2], a[3]);
    float2 d1 = csub(a[1], a[4]);
    float2 d2 = csub(a[2], a[3]);
    float2 y0 = cadd(a[0], cadd(b1, b2));
    float2 t1 = make_float2(a[0].x + c1 * b1.x + c2 * b2.x, a[0].y + c1 * b1.y + c2 * b2.y);
    float2 t2 = make_float2(a[0].x + c2 * b1.x + c1 * b2.x, a[0].y + c2 * b1.y + c1 * b2.y);
    float2 u1 = mul_mi<INV>(make_float2(s1 * d1.x + s2 * d2.x, s1 * d1.y + s2 * d2.y));
    float2 u2 = mul_mi<INV>(make_float2(s2 * d1.x - s1 * d2.x, s2 * d1.y - s1 * d2.y));
    a[0] = y0;
    a[1] = cadd(t1, u1);
    a[4] = csub(t1, u1);
    a[2] = cadd(t2, u2);
    a[3] = csub(t2, u2);
  }
};

template <bool INV> struct DFT<8, INV> {
  static __device__ __forceinline__ void run(float2* a) {
    const float r = 0.707106781186547524400844f;
    float2 e[4] = {a[0], a[2], a[4], a[6]};
    float2 o[4] = {a[1], a[3], a[5], a[7]};
    DFT<4, INV>::run(e);
    DFT<4, INV>::run(o);
    // o[k] *= W8^{+-k}
    float2 o1, o3;
    if (INV) {
      o1 = make_float2(r * (o[1].x - o[1].y), r * (o[1].x + o[1].y));
      o3 = make_float2(-r * (o[3].x + o[3].y), r * (o[3].x - o[3].y));
    } else {
      o1 = make_float2(r * (o[1].x + o[1].y), r * (o[1].y - o[1].x));
      o3 = make_float2(r * (o[3].y - o[3].x), -r * (o[3].x + o[3].y));
    }
    float2 o2 = mul_mi<INV>(o[2]);
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

// ------------------------------------------------ compile-time unit roots
// exp(-2 pi i num / den) in double at compile time (argument reduced to
// [-pi/4, pi/4], Taylor series to 1e-17), rounded once to float: the
// rotation constants of fft_stage's class-2 butterflies.
namespace detail {
constexpr double kPi = 3.14159265358979323846264338327950288;
constexpr double taylor_sin(double x) {
  double term = x, sum = x;
  for (int i = 1; i < 14; ++i) {
    term *= -x * x / ((2.0 * i) * (2.0 * i + 1.0));
    sum += term;
  }
  return sum;
}
constexpr double taylor_cos(double x) {
  double term = 1.0, sum = 1.0;
  for (int i = 1; i < 14; ++i) {
    term *= -x * x / ((2.0 * i - 1.0) * (2.0 * i));
    sum += term;
  }
  return sum;
}
// cos and sin of 2 pi m / d for integers m, d > 0, via the octant of m / d
constexpr void sincos_frac(long m, long d, double& c, double& s) {
  m %= d;
  if (m < 0) m += d;
  // q = the octant: angle = 2 pi m / d = q pi/4 + r with |r| <= pi/8
  const long q = (8 * m + d / 2) / d;
  const double r = 2.0 * kPi * ((double)m / (double)d) - (double)q * (kPi / 4.0);
  const double cr = taylor_cos(r), sr = taylor_sin(r);
  const double h = 0.70710678118654752440084436210484904;
  double cq = 1.0, sq = 0.0;
  switch (q % 8) {
    case 0: cq = 1.0; sq = 0.0; break;
    case 1: cq = h; sq = h; break;
    case 2: cq = 0.0; sq = 1.0; break;
    case 3: cq = -h; sq = h; break;
    case 4: cq = -1.0; sq = 0.0; break;
    case 5: cq = -h; sq = -h; break;
    case 6: cq = 0.0; sq = -1.0; break;
    default: cq = h; sq = -h; break;
  }
  c = cq * cr - sq * sr;
  s = sq * cr + cq * sr;
}
}  // namespace detail

struct CF {
  float x, y;
};
// W_den^num = exp(-2 pi i num / den)
constexpr CF unit_root(long num, long den) {
  double c = 0.0, s = 0.0;
  detail::sincos_frac(num, den, c, s);
  return CF{(float)c, (float)-s};
}

// --------------------------------------------------------- compile-time plan
// Every radix must divide E so that each thread owns a whole number of
// butterflies in every stage.  Power-of-two radices come first (largest
// first), the odd radix (3 or 5) last: every stage that writes LDS then has a
// power-of-two Ns, which makes its butterfly indices a per-thread base plus
// compile-time offsets (see fft_stage).
struct RadixPlan {
  int n, e, p, nstage;
  int r[16];
};

// nstage = -1 when n does not factor into radices dividing e within 16 stages
__host__ __device__ constexpr RadixPlan make_plan(int n, int e) {
  RadixPlan pl{n, e, n / e, 0, {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0}};
  int rem = n;
  const int cand[5] = {8, 4, 2, 5, 3};
  while (rem > 1 && pl.nstage < 16) {
    int pick = 0;
    for (int i = 0; i < 5; ++i) {
      if (e % cand[i] == 0 && rem % cand[i] == 0) { pick = cand[i]; break; }
    }
    if (pick == 0) { pl.nstage = -1; return pl; }
    pl.r[pl.nstage++] = pick;
    rem /= pick;
  }
  if (rem > 1) pl.nstage = -1;
  return pl;
}

// Register tile (elements per thread) of the lengths whose box passes were
// measured and tuned (2 to 4 LDS exchanges, at most 40 points per thread).
__host__ __device__ constexpr int tuned_elems(int n) {
  switch (n) {
    case 8: return 8;
    case 12: return 12;
    case 16: return 4;
    case 20: return 20;
    case 24: return 12;
    case 32: return 8;
    case 40: return 20;
    case 48: return 12;
    case 64: return 8;
    case 80: return 20;
    case 96: return 24;
    case 128: return 8;
    case 160: return 20;
    case 192: return 24;
    case 256: return 8;
    case 320: return 40;
    case 384: return 24;
    case 512: return 8;
    case 640: return 40;
    case 768: return 24;
    case 1024: return 8;
    case 1280: return 40;
    case 1536: return 24;
    case 2048: return 8;
    case 2560: return 40;
    case 4096: return 16;
    default: return 0;
  }
}

__host__ __device__ constexpr bool smooth235(int n) {
  if (n < 1) return false;
  while (n % 2 == 0) n /= 2;
  while (n % 3 == 0) n /= 3;
  while (n % 5 == 0) n /= 5;
  return n == 1;
}

// Any other 2/3/5-smooth length n <= 4096: E = (3 if 3 | n) (5 if 5 | n)
// 2^k, so that every radix of the plan divides E; among those with a plan,
// prefer 32 <= n/E <= 256 threads per line, then E <= 40, then the fewest
// stages, then the largest E.
__host__ __device__ constexpr int general_elems(int n) {
  const int odd = (n % 3 == 0 ? 3 : 1) * (n % 5 == 0 ? 5 : 1);
  int best = 0, bs0 = 9, bs1 = 9, bs2 = 99;
  for (int p2 = 1; p2 <= 64; p2 *= 2) {
    const int e = odd * p2;
    if (n % e != 0 || e > 64) continue;
    const RadixPlan pl = make_plan(n, e);
    if (pl.nstage <= 0) continue;
    const int P = n / e;
    const int s0 = (P >= 32 && P <= 256) ? 0 : 1, s1 = e <= 40 ? 0 : 1, s2 = pl.nstage;
    const bool better = s0 < bs0 || (s0 == bs0 && (s1 < bs1 || (s1 == bs1 && (s2 < bs2 || (s2 == bs2 && e > best)))));
    if (best == 0 || better) {
      best = e;
      bs0 = s0;
      bs1 = s1;
      bs2 = s2;
    }
  }
  return best;
}

// Elements per thread for every supported line length (0: unsupported):
// the tuned tile, or the general rule for the other 2/3/5-smooth lengths
// from 8 to 4096.
__host__ __device__ constexpr int elems_per_thread(int n) {
  const int t = tuned_elems(n);
  if (t) return t;
  return (n >= 8 && n <= 4096 && smooth235(n)) ? general_elems(n) : 0;
}

// LDS line stride (in float2): one pad slot every 32 points breaks the
// power-of-two strides of the early Stockham stages across banks.  The
// 2560-point line is left unpadded so that four lines fill exactly 80 KiB and
// two blocks share one CU's 160 KiB.
__host__ __device__ constexpr bool lds_padded(int n) { return n != 2560; }
template <int N> __host__ __device__ constexpr int lds_pos(int pos) {
  return lds_padded(N) ? pos + (pos >> 5) : pos;
}
__host__ __device__ constexpr int lds_line_len(int n) { return lds_padded(n) ? n + (n >> 5) + 1 : n; }

// A zero the compiler cannot see through.  Adding it to a table pointer at
// the start of each stage keeps that stage's loads inside the stage: without
// it the compiler merges the identical twiddle (and wavenumber) loads of the
// several transforms of one kernel and keeps them live across all of them,
// which spills the register tile.
__device__ __forceinline__ int opaque_zero() {
  int z = 0;
  asm volatile("" : "+s"(z));
  return z;
}

// Slot of point `pos` in an LDS line.  Padded lines (every length but 2560)
// insert one slot every 32 points.  The unpadded 2560-point line instead
// XORs bits [SWS, SWS + SWB) of the position with the lane's swizzle `sw`
// (its line index within the block, shifted by SWS): the lines of a block
// start a multiple of 64 banks apart and the lanes of a wave interleave them.
// fold(slot(x), C) = slot(x + C) for a compile-time C that is a multiple of G,
// which turns the LDS addresses of a stage into one base plus immediates.
template <int N, int SWS, int SWB> struct LdsMap {
  static constexpr bool padded = lds_padded(N);
  static constexpr int G = padded ? 32 : (1 << (SWS + SWB));
  static __device__ __forceinline__ uint32_t slot(uint32_t pos, uint32_t sw) {
    if constexpr (padded) return pos + (pos >> 5);
    else if constexpr (SWB == 0) return pos;
    else return pos ^ sw;
  }
  // c must be a multiple of G (checked through folds() at compile time by
  // the callers); c is a constant after unrolling
  static __device__ __forceinline__ uint32_t fold(uint32_t s, int c) {
    return s + (uint32_t)(padded ? c + c / 32 : c);
  }
  static constexpr bool folds(int c) { return c % G == 0; }
};

// ------------------------------------------------------------- stage driver
// One radix-R stage of a Stockham autosort FFT.  Butterfly j = p + P*b
// (b < B = E/R) takes the points j + r*N/R, i.e. register slots b + r*B;
// its twiddles are W_N^(r * (j % Ns) * N/(Ns R)) and its outputs go to
// idxD + r*Ns with idxD = (j/Ns)*Ns*R + j%Ns.  For the last stage that is
// j + r*N/R again, so the result stays in registers (slot b + r*B holds
// output r of butterfly b).
// Index classes, fixed at compile time:
//   Ns == 1       no twiddles, idxD = R*p + b*P*R
//   P % Ns == 0   j % Ns = p % Ns for every b: R-1 twiddle loads per stage,
//                 idxD = base(p) + b*P*R
//   Ns % P == 0   j % Ns = p + (P b) % Ns, j / Ns = (P b) / Ns
//   otherwise     per-butterfly division
template <int N, int P, int R, int Ns> struct StageIdx {
  static constexpr int KIND = Ns == 1 ? 0 : (P % Ns == 0 ? 1 : (Ns % P == 0 ? 2 : 3));
  static constexpr int step = N / (Ns * R);
  // compile-time parts for butterfly b (classes 0-2)
  static constexpr int offD(int b) {
    return KIND == 2 ? ((P * b) / Ns) * Ns * R + (P * b) % Ns : b * P * R;
  }
  static constexpr int offJ(int b) { return KIND == 2 ? (P * b) % Ns : 0; }
};

// Rotation constants of a class-2 stage: c[b][r] = W_N^(r offJ(b) step).
template <int N, int P, int R, int Ns, int E> struct RotTab {
  struct Tab {
    CF c[E / R][R];
  };
  static constexpr Tab make() {
    Tab t{};
    using I = StageIdx<N, P, R, Ns>;
    for (int b = 0; b < E / R; ++b)
      for (int r = 0; r < R; ++r) t.c[b][r] = unit_root((long)r * I::offJ(b) * I::step, N);
    return t;
  }
  static constexpr Tab tab = make();
};

// LDS slots of a stage's outputs: output r of butterfly b goes to point
// idxD(b) + r*Ns.  Classes 0-2 keep one per-thread base (and R per-r slots
// when an r*Ns offset does not fold); the butterfly offsets are immediates.
template <int N, int P, int R, int Ns, int E, int SWS, int SWB> struct OutSlots {
  using I = StageIdx<N, P, R, Ns>;
  using M = LdsMap<N, SWS, SWB>;
  static constexpr int B = E / R;
  static constexpr bool fold_b = []() {
    for (int b = 0; b < B; ++b)
      if (!M::folds(I::offD(b))) return false;
    return true;
  }();
  uint32_t p, sw, base0;
  uint32_t sr[R];
  __device__ __forceinline__ OutSlots(uint32_t p_, uint32_t sw_) : p(p_), sw(sw_), base0(0) {
    if constexpr (I::KIND == 0) base0 = p * R;
    else if constexpr (I::KIND == 1) base0 = (p / Ns) * (Ns * R) + p % Ns;
    else if constexpr (I::KIND == 2) base0 = p;
    if constexpr (I::KIND != 3 && fold_b) {
      if constexpr (M::folds(Ns)) {
        const uint32_t s0 = M::slot(base0, sw);
#pragma unroll
        for (int r = 0; r < R; ++r) sr[r] = s0 + (uint32_t)(M::padded ? r * Ns + (r * Ns) / 32 : r * Ns);
      } else {
#pragma unroll
        for (int r = 0; r < R; ++r) sr[r] = M::slot(base0 + (uint32_t)(r * Ns), sw);
      }
    }
  }
  __device__ __forceinline__ uint32_t at(int b, int r) const {
    if constexpr (I::KIND == 3) {
      const uint32_t j = p + (uint32_t)(P * b);
      return M::slot((j / Ns) * (Ns * R) + j % Ns + (uint32_t)(r * Ns), sw);
    } else if constexpr (fold_b) {
      return M::fold(sr[r], I::offD(b));
    } else {
      return M::slot(base0 + (uint32_t)(I::offD(b) + r * Ns), sw);
    }
  }
};

// Twiddles of butterfly b = 0 of a stage: W^(r * (p % Ns) * step), r < R
// (entry 0 unused).  For class 1 they serve every butterfly of the stage.
// fft_stages loads them for stage S+1 before the barriers of stage S, so the
// loads are in flight while the block synchronises.
template <int N, int R, int Ns, bool INV>
__device__ __forceinline__ void stage_twiddles0(float2 (&w0)[8], const float2* tw_in, uint32_t p_in) {
  if constexpr (Ns > 1) {
    constexpr int step = N / (Ns * R);
    const int z = opaque_zero();
    const float2* tw = tw_in + z;
    const uint32_t jj = (p_in + (uint32_t)z) % Ns;
#pragma unroll
    for (int r = 1; r < R; ++r) {
      w0[r] = tw[r * jj * step];
      if (INV) w0[r].y = -w0[r].y;
    }
  }
}

template <int N, int P, int R, int Ns, bool INV, bool LAST, int E, int SWS, int SWB>
__device__ __forceinline__ void fft_stage(float2 (&v)[E], float2* __restrict__ line, uint32_t p_in,
                                          const float2* tw_in, uint32_t sw, const float2 (&w0)[8]) {
  using I = StageIdx<N, P, R, Ns>;
  constexpr int B = E / R;
  const int z = opaque_zero();
  const float2* tw = tw_in + z;
  const uint32_t p = p_in + (uint32_t)z;

  // twiddle exponent base of butterfly b: p % Ns (class 1), p + offJ(b)
  // (class 2); butterfly 0 takes the preloaded w0
  [[maybe_unused]] OutSlots<N, P, R, Ns, E, SWS, SWB> os(p, sw);

#pragma unroll
  for (int b = 0; b < B; ++b) {
    float2 a[R];
#pragma unroll
    for (int r = 0; r < R; ++r) a[r] = v[b + r * B];
    if constexpr (I::KIND == 1) {
#pragma unroll
      for (int r = 1; r < R; ++r) a[r] = cmul(a[r], w0[r]);
    } else if constexpr (I::KIND == 2 || I::KIND == 3) {
      if (b == 0) {
#pragma unroll
        for (int r = 1; r < R; ++r) a[r] = cmul(a[r], w0[r]);
      } else if constexpr (I::KIND == 2) {
        // W^(r (p + offJ(b)) step) = w0[r] * W_N^(r offJ(b) step): butterfly
        // 0's twiddles rotated by a compile-time constant (no per-butterfly
        // loads, whose results the compiler would keep live in registers)
#pragma unroll
        for (int r = 1; r < R; ++r) {
          const CF c = RotTab<N, P, R, Ns, E>::tab.c[b][r];
          const float2 cw = make_float2(c.x, INV ? -c.y : c.y);
          a[r] = cmul(cmul(a[r], cw), w0[r]);
        }
      } else {
        const uint32_t jb = (p + (uint32_t)(P * b)) % Ns;
#pragma unroll
        for (int r = 1; r < R; ++r) {
          float2 w = tw[r * jb * I::step];
          if (INV) w.y = -w.y;
          a[r] = cmul(a[r], w);
        }
      }
    }
    DFT<R, INV>::run(a);
    if constexpr (LAST) {
#pragma unroll
      for (int r = 0; r < R; ++r) v[b + r * B] = a[r];
    } else {
#pragma unroll
      for (int r = 0; r < R; ++r) line[os.at(b, r)] = a[r];
    }
  }
}

// thread p loads points p + P*e of the line
template <int N, int P, int E, int SWS = 0, int SWB = 0>
__device__ __forceinline__ void lds_gather(float2 (&v)[E], const float2* __restrict__ line, uint32_t p_in,
                                           uint32_t sw = 0) {
  using M = LdsMap<N, SWS, SWB>;
  const uint32_t p = p_in + (uint32_t)opaque_zero();
  if constexpr (M::folds(P)) {
    const uint32_t s0 = M::slot(p, sw);
#pragma unroll
    for (int e = 0; e < E; ++e) v[e] = line[s0 + (uint32_t)(M::padded ? P * e + (P * e) / 32 : P * e)];
  } else {
#pragma unroll
    for (int e = 0; e < E; ++e) v[e] = line[M::slot(p + (uint32_t)(P * e), sw)];
  }
}

template <int N, int P, int E, int SWS = 0, int SWB = 0>
__device__ __forceinline__ void lds_scatter(const float2 (&v)[E], float2* __restrict__ line, uint32_t p,
                                            uint32_t sw = 0) {
  using M = LdsMap<N, SWS, SWB>;
  if constexpr (M::folds(P)) {
    const uint32_t s0 = M::slot(p, sw);
#pragma unroll
    for (int e = 0; e < E; ++e) line[s0 + (uint32_t)(M::padded ? P * e + (P * e) / 32 : P * e)] = v[e];
  } else {
#pragma unroll
    for (int e = 0; e < E; ++e) line[M::slot(p + (uint32_t)(P * e), sw)] = v[e];
  }
}

// Split exchange after a stage computed in the LAST form: output r of
// butterfly b (slot b + r*B) moves to point idxD + r*Ns.  The real parts go
// through a float LDS line, then the imaginary parts: half the LDS of the
// float2 exchange for two more barriers.
template <int N, int P, int R, int Ns, int E, int SWS, int SWB>
__device__ __forceinline__ void exchange_split(float2 (&v)[E], float* __restrict__ lf, uint32_t p_in, uint32_t sw) {
  using M = LdsMap<N, SWS, SWB>;
  constexpr int B = E / R;
  const uint32_t p = p_in + (uint32_t)opaque_zero();
  const OutSlots<N, P, R, Ns, E, SWS, SWB> os(p, sw);
  const uint32_t g0 = M::slot(p, sw);
#pragma unroll
  for (int half = 0; half < 2; ++half) {
#pragma unroll
    for (int b = 0; b < B; ++b) {
#pragma unroll
      for (int r = 0; r < R; ++r) lf[os.at(b, r)] = half ? v[b + r * B].y : v[b + r * B].x;
    }
    __syncthreads();
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const uint32_t sl = M::folds(P) ? M::fold(g0, P * e) : M::slot(p + (uint32_t)(P * e), sw);
      const float x = lf[sl];
      if (half) v[e].y = x; else v[e].x = x;
    }
    __syncthreads();
  }
}

template <int N, int E, bool INV, int S, int Ns, int SWS, int SWB>
__device__ __forceinline__ void fft_stages(float2 (&v)[E], float2* __restrict__ line, uint32_t p, const float2* tw,
                                           uint32_t sw, const float2 (&w0)[8]) {
  constexpr RadixPlan pl = make_plan(N, E);
  static_assert(pl.nstage > 0, "unsupported FFT length");
  constexpr int P = N / E;
  if constexpr (S < pl.nstage) {
    constexpr int R = pl.r[S];
    constexpr bool LAST = (S + 1 == pl.nstage);
    fft_stage<N, P, R, Ns, INV, LAST, E, SWS, SWB>(v, line, p, tw, sw, w0);
    if constexpr (!LAST) {
      float2 w1[8];
      stage_twiddles0<N, pl.r[S + 1], Ns * R, INV>(w1, tw, p);
      __syncthreads();
      lds_gather<N, P, E, SWS, SWB>(v, line, p, sw);
      __syncthreads();
      fft_stages<N, E, INV, S + 1, Ns * R, SWS, SWB>(v, line, p, tw, sw, w1);
    }
  }
}

template <int N, int E, bool INV, int S, int Ns, int SWS, int SWB>
__device__ __forceinline__ void fft_stages_split(float2 (&v)[E], float* __restrict__ lf, uint32_t p,
                                                 const float2* tw, uint32_t sw, const float2 (&w0)[8]) {
  constexpr RadixPlan pl = make_plan(N, E);
  static_assert(pl.nstage > 0, "unsupported FFT length");
  constexpr int P = N / E;
  if constexpr (S < pl.nstage) {
    constexpr int R = pl.r[S];
    fft_stage<N, P, R, Ns, INV, true, E, SWS, SWB>(v, nullptr, p, tw, sw, w0);
    if constexpr (S + 1 < pl.nstage) {
      float2 w1[8];
      stage_twiddles0<N, pl.r[S + 1], Ns * R, INV>(w1, tw, p);
      exchange_split<N, P, R, Ns, E, SWS, SWB>(v, lf, p, sw);
      fft_stages_split<N, E, INV, S + 1, Ns * R, SWS, SWB>(v, lf, p, tw, sw, w1);
    }
  }
}

// ------------------------------------------------- in-place DIF FFT in LDS
// Gentleman-Sande decimation in frequency on one N-point line held in LDS,
// with the radices of make_plan(N, E) in order: stage s (sub-transform
// length L = N / (R_1 ... R_{s-1}), stride m = L / R_s) takes the points
// base + j + r m of butterfly u (base = (u / m) L, j = u % m), computes
// their R-point DFT and multiplies output q by W_L^(j q) (conjugated for the
// inverse), in place.  The first stage reads the thread's register tile
// (natural order: thread p holds points p + P e) and every later one reads
// and writes LDS, so the data costs no registers beyond one butterfly.  The
// result is in digit-reversed order: X_k sits at dif_pos(k) = sum_s d_s(k)
// N / (R_1 ... R_s) with k = d_1 + R_1 (d_2 + R_2 (...)).  Every thread of
// the block must call it (it holds block barriers).
template <int N, int E> struct DifRev {
  static constexpr RadixPlan pl = make_plan(N, E);
  static constexpr int P = N / E;
  // J: the number of leading radices whose product is P (-1 if none)
  static constexpr int J = []() {
    int prod = 1;
    for (int s = 0; s < pl.nstage; ++s) {
      if (prod == P) return s;
      prod *= pl.r[s];
    }
    return prod == P ? pl.nstage : -1;
  }();
  static constexpr int stride(int s) {
    int prod = 1;
    for (int i = 0; i <= s; ++i) prod *= pl.r[i];
    return N / prod;
  }
  // the digits of lane p (radices 1..J)
  static __device__ __forceinline__ uint32_t lane(uint32_t p) {
    uint32_t pos = 0;
#pragma unroll
    for (int s = 0; s < J; ++s) {
      pos += (p % (uint32_t)pl.r[s]) * (uint32_t)stride(s);
      p /= (uint32_t)pl.r[s];
    }
    return pos;
  }
  // the digits of slot e (radices J+1..S)
  static constexpr uint32_t slot(int e) {
    uint32_t pos = 0;
    for (int s = J; s < pl.nstage; ++s) {
      pos += (uint32_t)(e % pl.r[s]) * (uint32_t)stride(s);
      e /= pl.r[s];
    }
    return pos;
  }
};

template <int N, int P, int R, int L, bool INV, bool FIRST, int E, int SWS, int SWB>
__device__ __forceinline__ void dif_stage(float2 (&v)[E], float2* __restrict__ line, uint32_t p_in,
                                          const float2* tw_in, uint32_t sw) {
  using M = LdsMap<N, SWS, SWB>;
  constexpr int m = L / R;
  constexpr int B = E / R;
  constexpr int TS = N / L;  // W_L^x = W_N^(x TS)
  const int z = opaque_zero();
  const float2* tw = tw_in + z;
  const uint32_t p = p_in + (uint32_t)z;
#pragma unroll
  for (int b = 0; b < B; ++b) {
    const uint32_t u = p + (uint32_t)(P * b);
    const uint32_t j = FIRST ? u : u % (uint32_t)m;
    const uint32_t base = FIRST ? u : (u / (uint32_t)m) * (uint32_t)L + j;
    float2 a[R];
#pragma unroll
    for (int r = 0; r < R; ++r) a[r] = FIRST ? v[b + r * B] : line[M::slot(base + (uint32_t)(r * m), sw)];
    DFT<R, INV>::run(a);
#pragma unroll
    for (int q = 1; q < R; ++q) {
      float2 w = tw[j * (uint32_t)(q * TS)];
      if (INV) w.y = -w.y;
      a[q] = cmul(a[q], w);
    }
#pragma unroll
    for (int q = 0; q < R; ++q) line[M::slot(base + (uint32_t)(q * m), sw)] = a[q];
  }
}

template <int N, int E, bool INV, int S, int L, int SWS, int SWB>
__device__ __forceinline__ void dif_stages(float2 (&v)[E], float2* __restrict__ line, uint32_t p, const float2* tw,
                                           uint32_t sw) {
  constexpr RadixPlan pl = make_plan(N, E);
  constexpr int P = N / E;
  if constexpr (S < pl.nstage) {
    constexpr int R = pl.r[S];
    dif_stage<N, P, R, L, INV, S == 0, E, SWS, SWB>(v, line, p, tw, sw);
    __syncthreads();
    dif_stages<N, E, INV, S + 1, L / R, SWS, SWB>(v, line, p, tw, sw);
  }
}

// The in-place DIF transform of the register tile v into `line`; on return
// (after a block barrier) X_k is at slot dif_pos(k) of the line.
template <int N, int E, bool INV, int SWS = 0, int SWB = 0>
__device__ __forceinline__ void fft_line_dif(float2 (&v)[E], float2* __restrict__ line, uint32_t p, const float2* tw,
                                             uint32_t sw = 0) {
  static_assert(make_plan(N, E).nstage > 0, "unsupported FFT length");
  dif_stages<N, E, INV, 0, N, SWS, SWB>(v, line, p, tw, sw);
}

// Full in-register FFT of one line.  All threads of the block must call it
// (it contains block barriers).  On entry and exit thread p holds points
// p + P*e.  Forward: X_k = sum_n x_n exp(-2 pi i n k / N); inverse uses
// exp(+...) and is unnormalised.
template <int N, bool INV>
__device__ __forceinline__ void fft_line(float2 (&v)[elems_per_thread(N)], float2* __restrict__ line, uint32_t p,
                                         const float2* tw) {
  const float2 w0[8] = {};
  fft_stages<N, elems_per_thread(N), INV, 0, 1, 0, 0>(v, line, p, tw, 0, w0);
}

// Same with an explicit register tile of E points per thread (P = N / E) and
// an optional line swizzle (see LdsMap; sw must be < 2^(SWS+SWB) with its
// bits in [SWS, SWS+SWB)).
template <int N, int E, bool INV, int SWS = 0, int SWB = 0>
__device__ __forceinline__ void fft_line_e(float2 (&v)[E], float2* __restrict__ line, uint32_t p,
                                           const float2* tw, uint32_t sw = 0) {
  const float2 w0[8] = {};
  fft_stages<N, E, INV, 0, 1, SWS, SWB>(v, line, p, tw, sw, w0);
}

// Same through a float LDS line of lds_line_len(N) floats (split exchange).
template <int N, int E, bool INV, int SWS = 0, int SWB = 0>
__device__ __forceinline__ void fft_line_split(float2 (&v)[E], float* __restrict__ lf, uint32_t p,
                                               const float2* tw, uint32_t sw = 0) {
  const float2 w0[8] = {};
  fft_stages_split<N, E, INV, 0, 1, SWS, SWB>(v, lf, p, tw, sw, w0);
}

// ------------------------------------------------------ counter-based noise
// Philox4x32-10 (Salmon et al., SC'11), keyed by the 64-bit seed, one counter
// per group of four consecutive cells; Box-Muller turns the four uniforms
// into four N(0,1) values.  The value at a cell depends only on (seed, cell
// index), never on the launch geometry.
struct U4 { uint32_t x, y, z, w; };

__host__ __device__ __forceinline__ U4 philox4x32_10_c4(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                                        uint64_t seed) {
  uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
  const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u, W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#pragma unroll
  for (int i = 0; i < 10; ++i) {
    uint64_t p0 = (uint64_t)M0 * c0;
    uint64_t p1 = (uint64_t)M1 * c2;
    uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    uint32_t n0 = hi1 ^ c1 ^ k0;
    uint32_t n2 = hi0 ^ c3 ^ k1;
    c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
    k0 += W0; k1 += W1;
  }
  return U4{c0, c1, c2, c3};
}

__host__ __device__ __forceinline__ U4 philox4x32_10(uint64_t ctr, uint64_t seed) {
  return philox4x32_10_c4((uint32_t)ctr, (uint32_t)(ctr >> 32), 0u, 0u, seed);
}

// uniform in (0, 1): 24 high bits plus half an ulp, exactly representable
__device__ __forceinline__ float u01(uint32_t x) {
  return (float)(x >> 8) * 5.9604644775390625e-8f + 2.98023223876953125e-8f;
}

__device__ __forceinline__ void box_muller(uint32_t a, uint32_t b, float& n0, float& n1) {
  float r = sqrtf(-2.0f * logf(u01(a)));
  float s, c;
  sincospif(2.0f * u01(b), &s, &c);
  n0 = r * c;
  n1 = r * s;
}

// The same transform on the hardware transcendentals: v_log_f32 is log2
// (ln u = ln2 * log2 u), and v_sin_f32 / v_cos_f32 take their argument in
// revolutions, so sin(2 pi u) is one instruction for u in (0, 1).  The
// error against the FP64 transform is a few 1e-7 absolute.
__device__ __forceinline__ void box_muller_hw(uint32_t a, uint32_t b, float& n0, float& n1) {
  const float r = __builtin_amdgcn_sqrtf(-1.3862943611198906f * __builtin_amdgcn_logf(u01(a)));
  const float ub = u01(b);
  n0 = r * __builtin_amdgcn_cosf(ub);
  n1 = r * __builtin_amdgcn_sinf(ub);
}

// XCD-aware tile order: consecutive hardware block ids land on different
// XCDs (round robin over 8); remap so that each XCD walks a contiguous range
// of tiles and neighbouring tiles share that XCD's L2.
__device__ __forceinline__ uint32_t xcd_tile(uint32_t b, uint32_t nblk) {
  const uint32_t nx = 8;
  uint32_t q = nblk / nx, r = nblk % nx;
  uint32_t x = b % nx, i = b / nx;
  return x * q + (x < r ? x : r) + i;
}

}  // namespace smk
