// Kernels and per-length launchers of the box passes (included by box.hip
// and by the box_len_<g>.hip units that instantiate them, one group of line
// lengths each, see fft_lengths.h).  Device and host code; namespace smk_box.
#pragma once
#include <stdint.h>

#include "fft_core.h"
#include "fft_lengths.h"
#include "smk_internal.h"

namespace smk_box {

using namespace smk;

// --------------------------------------------------------------- launch shapes
// Threads per line P = N / E.  Lines per block T: strided passes use runs of
// T consecutive lines (T*8 bytes per HBM access run, see strided_variant);
// pencil passes just batch T pencils of E points per thread.
constexpr int pencil_T(int m, int e) {
  int T = 16;
  while (T > 1 && (T * (m / e) > 512 || T * lds_line_len(m) * 8 > 65536)) T >>= 1;
  return T;
}

enum { MODE_FWD = 0, MODE_INV = 1, MODE_SPECTRAL = 2 };

struct StridedArgs {
  // line (o, i) point n is read at  in[o*in_so + i*in_si + n*in_sn]  and
  // written at out[o*out_so + i*out_si + n*out_sk]; in == out is allowed
  // (each block reads its whole tile before it writes it)
  const float2* in;
  float2* out;
  float2* aux;           // second output (MODE_SPECTRAL with aux_kind != NONE)
  const float2* tw;      // W_N^m, m < N
  uint64_t in_so, in_si, in_sn;
  uint64_t out_so, out_si, out_sk;
  uint32_t n_inner;      // extent of the inner (run) index
  uint32_t n_iblk;       // ceil(n_inner / T)
  // ---- MODE_SPECTRAL only: the line axis is x; kz_outer = 1 when the outer
  // index is kz and the inner one ky, 0 for the opposite
  int kz_outer;
  const float* kxv;
  const float* kyv;
  const float* kzv;
  float pkA, pk_k0inv, pk_n, pk_R;
  // the same P(k) in log2 form (k_xspec): sqrt(P / Vcell) * norm =
  // pk_cm * exp2(L/2 - log2(1 + x^n)/2 - pk_c2 kk), L = log2(kk)/2 - pk_l2k0
  float pk_l2k0, pk_c2, pk_cm;
  float vcell;           // float32(cell^3), as make_boxes.py:181
  float norm;            // 1 / (nx*ny*nz) folded into the spectrum
  float aux_coef;
  int aux_kind;
  int write_delta;
  uint32_t nz_nyq;       // kz index of the Nyquist plane (nz/2)
  uint32_t kz0;          // kz index of this launch's first kz (slab launches)
  double* partial;       // per-block FP64 sums (spectral: w(kz)|delta_k|^2)
};

// P(k) = A x / (1 + x^n) exp(-(k R)^2), x = k / k0, with the hardware
// transcendentals (about 1 ulp each); x = 0 gives P = 0.
__device__ __forceinline__ float pk_analytic(float k, const StridedArgs& a) {
  const float x = k * a.pk_k0inv;
  const float xn = __builtin_amdgcn_exp2f(a.pk_n * __builtin_amdgcn_logf(x));
  const float kr = k * a.pk_R;
  const float g = __builtin_amdgcn_exp2f(-1.4426950408889634f * kr * kr);
  return a.pkA * x * __builtin_amdgcn_rcpf(1.0f + xn) * g;
}

constexpr int pow2_ceil(int x) { return x <= 1 ? 1 : 2 * pow2_ceil((x + 1) / 2); }

// Wave reduction then block reduction in a fixed order (deterministic).
// Every thread must call it; `red` may alias LDS used by other phases (it
// needs NT doubles when NT is not a power of two): the function is bracketed
// by barriers.  The result is valid on thread 0.
template <int NT>
__device__ __forceinline__ double block_sum(double v, double* red) {
  if constexpr ((NT & (NT - 1)) == 0) {
    // NT is a power of two; blocks smaller than a wave shuffle only among
    // their own (active) lanes
    constexpr int W = NT < 64 ? NT : 64;
    for (int off = W / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, W);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) red[wv] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0) {
      for (int i = 0; i < (NT + 63) / 64; ++i) s += red[i];
    }
    __syncthreads();
    return s;
  } else {
    // any other block size (the general line lengths): a tree over LDS
    __syncthreads();
    red[threadIdx.x] = v;
    __syncthreads();
    for (int h = pow2_ceil(NT) / 2; h > 0; h >>= 1) {
      if ((int)threadIdx.x < h && (int)threadIdx.x + h < NT) red[threadIdx.x] += red[threadIdx.x + h];
      __syncthreads();
    }
    const double s = red[0];
    __syncthreads();
    return s;
  }
}

// One line FFT through the block's LDS in the exchange form XCH
// (0: float2 line, 1: float line, split exchange), line swizzle bits
// [SWS, SWS+SWB) (see LdsMap).
template <int N, int E, bool INV, int XCH, int SWS, int SWB>
__device__ __forceinline__ void line_fft(float2 (&v)[E], float2* lds, uint32_t t, uint32_t p, const float2* tw,
                                         uint32_t sw) {
  constexpr int LL = lds_line_len(N);
  if constexpr (XCH == 1) {
    fft_line_split<N, E, INV, SWS, SWB>(v, reinterpret_cast<float*>(lds) + t * LL, p, tw, sw);
  } else {
    fft_line_e<N, E, INV, SWS, SWB>(v, lds + t * LL, p, tw, sw);
  }
}

constexpr int ilog2(int x) { return x <= 1 ? 0 : 1 + ilog2(x / 2); }

// Aux factor of one mode: eta_ij = k_i k_j / kk, v_i = -i k_i / kk * coef.
// Operand selectors sa, sb: 0 = kx, 1 = ky, 2 = kz, 3 = 1.
struct AuxSel {
  int sa, sb;
  bool vel;
  float coef;
};
__device__ __forceinline__ AuxSel aux_sel(int ak, float aux_coef) {
  AuxSel s;
  s.sa = (ak == SMK_AUX_ETA_XX || ak == SMK_AUX_ETA_XY || ak == SMK_AUX_ETA_XZ || ak == SMK_AUX_VX) ? 0
         : (ak == SMK_AUX_ETA_YY || ak == SMK_AUX_ETA_YZ || ak == SMK_AUX_VY) ? 1 : 2;
  s.sb = (ak == SMK_AUX_ETA_XX) ? 0 : (ak == SMK_AUX_ETA_YY || ak == SMK_AUX_ETA_XY) ? 1 : (ak >= SMK_AUX_VX) ? 3 : 2;
  s.vel = ak >= SMK_AUX_VX;
  s.coef = s.vel ? aux_coef : 1.0f;
  return s;
}
__device__ __forceinline__ float2 aux_apply(float2 d, const AuxSel& s, float kx, float ky, float kz, float kk,
                                            bool origin, float scale) {
  const float kkp = origin ? 1.0f : kk;
  const float ka = s.sa == 0 ? kx : (s.sa == 1 ? ky : kz);
  const float kb = s.sb == 0 ? kx : (s.sb == 1 ? ky : (s.sb == 2 ? kz : 1.0f));
  const float2 x = cscale(d, ka * kb * __builtin_amdgcn_rcpf(kkp) * (s.coef * scale));
  return s.vel ? mul_mi<false>(x) : x;
}

// T lines of N points per block, P = N / E threads per line, lanes ordered
// line-fastest.  XCH selects the LDS exchange (float2 line, or float line
// with the split exchange: half the LDS per block).  TILES (spectral mode
// only) says where the aux spectrum waits while delta is inverse transformed:
// 1 recomputes it (FFT(IFFT(d)) = N d: 4 line FFTs), 2 keeps it in a second
// register tile (3 FFTs; used with the smaller tile E1 only: 2 x 40 points
// per thread do not fit the register file at E = 40), 3 stashes it in the aux
// output rows the same thread overwrites later (3 FFTs; the rows are still
// in L2 / MALL when they are read back).
// SWS places the line swizzle of unpadded lines (see LdsMap).  OCC > 0 asks
// the compiler for at least OCC waves per SIMD (it then caps the registers).
// Global addresses are a block-uniform base, a per-lane offset computed once
// and a uniform per-point offset; lanes past the last line read the block's
// first line (always in range) and store nothing.
template <int N, int E, int T, int MODE, int XCH, int TILES, int SWS, int OCC>
__global__ __launch_bounds__(T*(N / E)) __attribute__((amdgpu_waves_per_eu(OCC > 0 ? OCC : 1))) void k_strided(
    StridedArgs a) {
  constexpr int P = N / E;
  constexpr int LL = lds_line_len(N);
  constexpr int NT = T * P;
  constexpr int SWB = lds_padded(N) ? 0 : ilog2(T);
  __shared__ float2 lds[XCH == 1 ? (T * LL + 1) / 2 : T * LL];

  const uint32_t tid = threadIdx.x;
  const uint32_t t = tid % T, p = tid / T;
  const uint32_t sw = SWB ? (t << SWS) : 0u;
  const uint32_t tile = xcd_tile(blockIdx.x, gridDim.x);
  const uint32_t o = tile / a.n_iblk;
  const uint32_t inner0 = (tile % a.n_iblk) * T;
  const uint32_t inner = inner0 + t;
  const bool valid = inner < a.n_inner;
  const uint32_t tl = valid ? t : 0u;
  const float2* src = a.in + (uint64_t)o * a.in_so + (uint64_t)inner0 * a.in_si + (uint64_t)tl * a.in_si +
                      (uint64_t)p * a.in_sn;
  const uint64_t in_step = (uint64_t)P * a.in_sn;
  float2* dst = a.out + (uint64_t)o * a.out_so + (uint64_t)inner * a.out_si + (uint64_t)p * a.out_sk;
  const uint64_t out_step = (uint64_t)P * a.out_sk;

  float2 v[E];
#pragma unroll
  for (int e = 0; e < E; ++e) v[e] = src[(uint64_t)e * in_step];

  if constexpr (MODE == MODE_INV) {
    // optional FP64 Parseval sum of the input spectrum (natural x pass only:
    // the inner index is kz)
    if (a.partial != nullptr) {
      const double wz = (inner == 0 || inner == a.nz_nyq) ? 1.0 : 2.0;
      double acc = 0.0;
#pragma unroll
      for (int e = 0; e < E; ++e) acc += (double)v[e].x * (double)v[e].x + (double)v[e].y * (double)v[e].y;
      acc = valid ? wz * acc : 0.0;
      const double bs = block_sum<NT>(acc, reinterpret_cast<double*>(lds));
      if (threadIdx.x == 0) a.partial[tile] = bs;
    }
  }
  line_fft<N, E, MODE == MODE_INV, XCH, SWS, SWB>(v, lds, t, p, a.tw, sw);

  if constexpr (MODE != MODE_SPECTRAL) {
    if (valid) {
#pragma unroll
      for (int e = 0; e < E; ++e) dst[(uint64_t)e * out_step] = v[e];
    }
  } else {
    // v[e] holds X(kx = p + P e, ky, kz)
    const uint32_t iky = a.kz_outer ? inner : o;
    const uint32_t ikz = (a.kz_outer ? o : inner) + a.kz0;
    const float ky = valid ? a.kyv[iky] : 0.0f;
    const float kz = valid ? a.kzv[ikz] : 0.0f;
    const double wz = (ikz == 0 || ikz == a.nz_nyq) ? 1.0 : 2.0;
    const float inv_vcell = 1.0f / a.vcell;
    const AuxSel sel = aux_sel(a.aux_kind, a.aux_coef);
    const bool has_aux = a.aux_kind != SMK_AUX_NONE;
    float2* adst = has_aux ? a.aux + (uint64_t)o * a.out_so + (uint64_t)inner * a.out_si + (uint64_t)p * a.out_sk
                           : nullptr;
    double acc = 0.0;
    float2 w[TILES == 2 ? E : 1];
    {
      const int z1 = opaque_zero();
      const float* kxv = a.kxv + z1;
      const int p1 = p + z1;
#pragma unroll
      for (int e = 0; e < E; ++e) {
        const int i = p1 + P * e;
        const float kx = kxv[i];
        // kk = kx*kx + ky*ky + kz*kz in float32, numpy evaluation order
        const float kk = add_rn(add_rn(mul_rn(kx, kx), mul_rn(ky, ky)), mul_rn(kz, kz));
        const float k = __builtin_amdgcn_sqrtf(kk);
        const float sp = __builtin_amdgcn_sqrtf(pk_analytic(k, a) * inv_vcell);
        const float2 d = cscale(v[e], sp * a.norm);
        acc += (double)d.x * (double)d.x + (double)d.y * (double)d.y;
        v[e] = d;
        if constexpr (TILES == 3) {
          if (valid && has_aux) adst[(uint64_t)e * out_step] = aux_apply(d, sel, kx, ky, kz, kk,
                                                                         i == 0 && iky == 0 && ikz == 0, 1.0f);
        }
        if constexpr (TILES == 2) w[e] = aux_apply(d, sel, kx, ky, kz, kk, i == 0 && iky == 0 && ikz == 0, 1.0f);
      }
    }
    acc = valid ? wz * acc : 0.0;
    const double bs = block_sum<NT>(acc, reinterpret_cast<double*>(lds));
    if (threadIdx.x == 0) a.partial[tile] = bs;

    line_fft<N, E, true, XCH, SWS, SWB>(v, lds, t, p, a.tw, sw);
    if (valid && a.write_delta) {
      const uint64_t st = out_step + (uint64_t)opaque_zero();
#pragma unroll
      for (int e = 0; e < E; ++e) dst[(uint64_t)e * st] = v[e];
    }
    if (has_aux) {
      if constexpr (TILES == 1) {
        // FFT(IFFT(d)) = N d (exact up to rounding): recompute the spectrum,
        // scale by f/N, inverse transform
        __syncthreads();
        line_fft<N, E, false, XCH, SWS, SWB>(v, lds, t, p, a.tw, sw);
        const int z3 = opaque_zero();
        const float* kxv2 = a.kxv + z3;
        const int p3 = p + z3;
#pragma unroll
        for (int e = 0; e < E; ++e) {
          const int i = p3 + P * e;
          const float kx = kxv2[i];
          const float kk = add_rn(add_rn(mul_rn(kx, kx), mul_rn(ky, ky)), mul_rn(kz, kz));
          v[e] = aux_apply(v[e], sel, kx, ky, kz, kk, i == 0 && iky == 0 && ikz == 0, 1.0f / (float)N);
        }
        line_fft<N, E, true, XCH, SWS, SWB>(v, lds, t, p, a.tw, sw);
      } else if constexpr (TILES == 2) {
#pragma unroll
        for (int e = 0; e < E; ++e) v[e] = w[e];
        line_fft<N, E, true, XCH, SWS, SWB>(v, lds, t, p, a.tw, sw);
      } else {
        // read back the stash this thread wrote (same addresses)
        const uint64_t st = out_step + (uint64_t)opaque_zero();
#pragma unroll
        for (int e = 0; e < E; ++e) v[e] = valid ? adst[(uint64_t)e * st] : make_float2(0.f, 0.f);
        line_fft<N, E, true, XCH, SWS, SWB>(v, lds, t, p, a.tw, sw);
      }
      if (valid) {
        const uint64_t st = out_step + (uint64_t)opaque_zero();
#pragma unroll
        for (int e = 0; e < E; ++e) adst[(uint64_t)e * st] = v[e];
      }
    }
  }
}

// Spectral x pass with three line FFTs (variant 14): x FFT, then per mode
// delta_k = X sqrt(P/Vcell)/N and aux_k = delta_k * factor, both kept in
// registers (two tiles of E points), then the inverse FFT of each.  Global
// addresses are a block-uniform base (scalar registers) plus a 32-bit lane
// offset, so a load or store costs no vector address arithmetic.  The
// spectrum is evaluated as one exp2 of a sum of logarithms:
//   sqrt(P/Vcell)/N = cm x^(1/2) (1 + x^n)^(-1/2) exp(-(kR)^2 / 2),
// x = k/k0, with log2 x = log2(kk)/2 - log2 k0 (no square root of kk);
// kk = 0 gives exp2(-inf) = 0 = P(0).
// Loads and stores at a block-uniform base plus a 32-bit byte offset: the
// form global_load/store take with the base in scalar registers.
template <class V>
__device__ __forceinline__ V ld_at(const V* base, uint32_t boff) {
  return *reinterpret_cast<const V*>(reinterpret_cast<const char*>(base) + boff);
}
template <class V>
__device__ __forceinline__ void st_at(V* base, uint32_t boff, V v) {
  *reinterpret_cast<V*>(reinterpret_cast<char*>(base) + boff) = v;
}

template <int N, int E, int T, int SWS, int OCC>
__global__ __launch_bounds__(T*(N / E)) __attribute__((amdgpu_waves_per_eu(OCC))) void k_xspec(StridedArgs a) {
  constexpr int P = N / E;
  constexpr int LL = lds_line_len(N);
  constexpr int NT = T * P;
  constexpr int SWB = lds_padded(N) ? 0 : ilog2(T);
  __shared__ float2 lds[T * LL];

  const uint32_t tid = threadIdx.x;
  const uint32_t t = tid % T, p = tid / T;
  const uint32_t sw = SWB ? (t << SWS) : 0u;
  const uint32_t tile = xcd_tile(blockIdx.x, gridDim.x);
  const uint32_t o = tile / a.n_iblk;
  const uint32_t inner0 = (tile % a.n_iblk) * T;
  const uint32_t inner = inner0 + t;
  const bool valid = inner < a.n_inner;
  const uint32_t tl = valid ? t : 0u;
  const float2* src = a.in + (uint64_t)o * a.in_so + (uint64_t)inner0 * a.in_si;
  const uint32_t src_b = (tl * (uint32_t)a.in_si + p * (uint32_t)a.in_sn) * 8u;
  const uint64_t in_step = (uint64_t)P * a.in_sn;

  float2 v[E];
#pragma unroll
  for (int e = 0; e < E; ++e) v[e] = ld_at(src + (uint64_t)e * in_step, src_b);

  line_fft<N, E, false, 0, SWS, SWB>(v, lds, t, p, a.tw, sw);

  const uint32_t iky = a.kz_outer ? inner : o;
  const uint32_t ikz = (a.kz_outer ? o : inner) + a.kz0;
  const float ky = valid ? a.kyv[iky] : 0.0f;
  const float kz = valid ? a.kzv[ikz] : 0.0f;
  const float ky2 = mul_rn(ky, ky), kz2 = mul_rn(kz, kz);
  const double wz = (ikz == 0 || ikz == a.nz_nyq) ? 1.0 : 2.0;
  const AuxSel sel = aux_sel(a.aux_kind, a.aux_coef);
  const bool has_aux = a.aux_kind != SMK_AUX_NONE;
  const bool origin_line = iky == 0 && ikz == 0;
  double acc = 0.0;
  float2 w[E];
  {
    const uint32_t kb = p * 4u;
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const uint32_t i = p + (uint32_t)(P * e);
      const float kx = ld_at(a.kxv + P * e, kb);
      // kk = kx*kx + ky*ky + kz*kz in float32, numpy evaluation order
      const float kk = add_rn(add_rn(mul_rn(kx, kx), ky2), kz2);
      const float l2x = 0.5f * __builtin_amdgcn_logf(kk) - a.pk_l2k0;
      const float xn = __builtin_amdgcn_exp2f(a.pk_n * l2x);
      const float ex = 0.5f * l2x - 0.5f * __builtin_amdgcn_logf(1.0f + xn) - a.pk_c2 * kk;
      const float sp = a.pk_cm * __builtin_amdgcn_exp2f(ex);
      const float2 d = cscale(v[e], sp);
      acc += (double)d.x * (double)d.x + (double)d.y * (double)d.y;
      v[e] = d;
      w[e] = aux_apply(d, sel, kx, ky, kz, kk, origin_line && i == 0, 1.0f);
    }
  }
  acc = valid ? wz * acc : 0.0;
  const double bs = block_sum<NT>(acc, reinterpret_cast<double*>(lds));
  if (threadIdx.x == 0) a.partial[tile] = bs;

  const uint64_t out_base = (uint64_t)o * a.out_so + (uint64_t)inner0 * a.out_si;
  const uint32_t dst_b = (t * (uint32_t)a.out_si + p * (uint32_t)a.out_sk) * 8u;
  const uint64_t out_step = (uint64_t)P * a.out_sk;

  line_fft<N, E, true, 0, SWS, SWB>(v, lds, t, p, a.tw, sw);
  if (valid && a.write_delta) {
    float2* dst = a.out + out_base;
#pragma unroll
    for (int e = 0; e < E; ++e) st_at(dst + (uint64_t)e * out_step, dst_b, v[e]);
  }
  if (has_aux) {
    line_fft<N, E, true, 0, SWS, SWB>(w, lds, t, p, a.tw, sw);
    if (valid) {
      float2* adst = a.aux + out_base;
#pragma unroll
      for (int e = 0; e < E; ++e) st_at(adst + (uint64_t)e * out_step, dst_b, w[e]);
    }
  }
}

// Spectral x pass, variant 16: as k_xspec (three line FFTs, aux spectrum in
// registers), but delta's inverse FFT runs in place in LDS (fft_line_dif,
// decimation in frequency), so that while it runs the registers hold only
// the aux tile and one butterfly; delta is then stored from the
// digit-reversed line.  Needs P = N / E to be a product of the plan's
// leading radices (DifRev::J >= 0).
template <int N, int E, int T, int SWS>
__global__ __launch_bounds__(T*(N / E)) __attribute__((amdgpu_waves_per_eu(2))) void k_xspec_dif(StridedArgs a) {
  constexpr int P = N / E;
  constexpr int LL = lds_line_len(N);
  constexpr int NT = T * P;
  constexpr int SWB = lds_padded(N) ? 0 : ilog2(T);
  using Rev = DifRev<N, E>;
  static_assert(Rev::J >= 0, "threads per line must be a product of leading radices");
  using LM = LdsMap<N, SWS, SWB>;
  __shared__ float2 lds[T * LL];

  const uint32_t tid = threadIdx.x;
  const uint32_t t = tid % T, p = tid / T;
  const uint32_t sw = SWB ? (t << SWS) : 0u;
  const uint32_t tile = xcd_tile(blockIdx.x, gridDim.x);
  const uint32_t o = tile / a.n_iblk;
  const uint32_t inner0 = (tile % a.n_iblk) * T;
  const uint32_t inner = inner0 + t;
  const bool valid = inner < a.n_inner;
  const uint32_t tl = valid ? t : 0u;
  const float2* src = a.in + (uint64_t)o * a.in_so + (uint64_t)inner0 * a.in_si;
  const uint32_t src_b = (tl * (uint32_t)a.in_si + p * (uint32_t)a.in_sn) * 8u;
  const uint64_t in_step = (uint64_t)P * a.in_sn;

  float2 v[E];
#pragma unroll
  for (int e = 0; e < E; ++e) v[e] = ld_at(src + (uint64_t)e * in_step, src_b);

  line_fft<N, E, false, 0, SWS, SWB>(v, lds, t, p, a.tw, sw);

  const uint32_t iky = a.kz_outer ? inner : o;
  const uint32_t ikz = (a.kz_outer ? o : inner) + a.kz0;
  const float ky = valid ? a.kyv[iky] : 0.0f;
  const float kz = valid ? a.kzv[ikz] : 0.0f;
  const float ky2 = mul_rn(ky, ky), kz2 = mul_rn(kz, kz);
  const double wz = (ikz == 0 || ikz == a.nz_nyq) ? 1.0 : 2.0;
  const AuxSel sel = aux_sel(a.aux_kind, a.aux_coef);
  const bool has_aux = a.aux_kind != SMK_AUX_NONE;
  const bool origin_line = iky == 0 && ikz == 0;
  double acc = 0.0;
  float2 w[E];
  {
    const uint32_t kb = p * 4u;
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const uint32_t i = p + (uint32_t)(P * e);
      const float kx = ld_at(a.kxv + P * e, kb);
      // kk = kx*kx + ky*ky + kz*kz in float32, numpy evaluation order
      const float kk = add_rn(add_rn(mul_rn(kx, kx), ky2), kz2);
      const float l2x = 0.5f * __builtin_amdgcn_logf(kk) - a.pk_l2k0;
      const float xn = __builtin_amdgcn_exp2f(a.pk_n * l2x);
      const float ex = 0.5f * l2x - 0.5f * __builtin_amdgcn_logf(1.0f + xn) - a.pk_c2 * kk;
      const float sp = a.pk_cm * __builtin_amdgcn_exp2f(ex);
      const float2 d = cscale(v[e], sp);
      acc += (double)d.x * (double)d.x + (double)d.y * (double)d.y;
      v[e] = d;
      w[e] = aux_apply(d, sel, kx, ky, kz, kk, origin_line && i == 0, 1.0f);
    }
  }
  acc = valid ? wz * acc : 0.0;
  const double bs = block_sum<NT>(acc, reinterpret_cast<double*>(lds));
  if (threadIdx.x == 0) a.partial[tile] = bs;

  const uint64_t out_base = (uint64_t)o * a.out_so + (uint64_t)inner0 * a.out_si;
  const uint32_t dst_b = (t * (uint32_t)a.out_si + p * (uint32_t)a.out_sk) * 8u;
  const uint64_t out_step = (uint64_t)P * a.out_sk;

  // delta: in place in this line of the block's LDS, then stored point by
  // point from its digit-reversed slot
  float2* line = lds + t * LL;
  fft_line_dif<N, E, true, SWS, SWB>(v, line, p, a.tw, sw);
  if (valid && a.write_delta) {
    float2* dst = a.out + out_base;
    const uint32_t lp = Rev::lane(p);
#pragma unroll
    for (int e = 0; e < E; ++e) st_at(dst + (uint64_t)e * out_step, dst_b, line[LM::slot(lp + Rev::slot(e), sw)]);
  }
  __syncthreads();
  if (has_aux) {
    line_fft<N, E, true, 0, SWS, SWB>(w, lds, t, p, a.tw, sw);
    if (valid) {
      float2* adst = a.aux + out_base;
#pragma unroll
      for (int e = 0; e < E; ++e) st_at(adst + (uint64_t)e * out_step, dst_b, w[e]);
    }
  }
}

// z pencils, R2C.  M = nz/2 complex points z_n = x_{2n} + i x_{2n+1}; after
// the M-point FFT, X_k = Ze_k + W_{2M}^k Zo_k with
// Ze_k = (Z_k + conj Z_{M-k})/2, Zo_k = -i (Z_k - conj Z_{M-k})/2.
// Output: trans == 0 -> numpy layout K[x][y][kz]; trans == 1 -> K[x][kz][y]
// (the store phase then runs with lanes ordered pencil-fastest, i.e. along y).
// Pencils pencil0 .. n_pencils-1 (a launch may cover a slab of the box).
template <int M, int E, int T>
__global__ __launch_bounds__(T*(M / E)) void k_z_r2c(const float* __restrict__ noise, uint64_t seed,
                                                     float2* __restrict__ K, const float2* __restrict__ tw,
                                                     const float2* __restrict__ twr, uint64_t pencil0,
                                                     uint64_t n_pencils, int trans, uint32_t ny) {
  constexpr int P = M / E;
  constexpr int LL = lds_line_len(M);
  constexpr int NT = T * P;
  __shared__ float2 lds[T * LL];
  const int tid = threadIdx.x;
  const int p = tid % P, t = tid / P;
  const uint64_t L0 = pencil0 + (uint64_t)blockIdx.x * T;
  const uint64_t L = L0 + t;
  const bool valid = L < n_pencils;
  float2* line = lds + t * LL;
  float2 v[E];

  if (noise == nullptr) {
    // four cells per Philox call: cells 4q..4q+3 of the whole box
    constexpr int Q = M / 2;
    for (int q = tid; q < T * Q; q += NT) {
      const int tt = q / Q, qq = q % Q;
      const uint64_t LL_ = L0 + tt;
      if (LL_ < n_pencils) {
        U4 u = philox4x32_10(LL_ * (uint64_t)Q + qq, seed);
        float n0, n1, n2, n3;
        box_muller_hw(u.x, u.y, n0, n1);
        box_muller_hw(u.z, u.w, n2, n3);
        float2* ln = lds + tt * LL;
        ln[lds_pos<M>(2 * qq)] = make_float2(n0, n1);
        ln[lds_pos<M>(2 * qq + 1)] = make_float2(n2, n3);
      }
    }
    __syncthreads();
    lds_gather<M, P, E>(v, line, p);
    __syncthreads();
  } else {
    const float2* src = reinterpret_cast<const float2*>(noise) + L * M;
#pragma unroll
    for (int e = 0; e < E; ++e) v[e] = valid ? src[p + P * e] : make_float2(0.f, 0.f);
  }

  fft_line_e<M, E, false>(v, line, p, tw);

  lds_scatter<M, P, E>(v, line, p);
  __syncthreads();
  // store phase: (ts, ps) is the lane's (pencil, point) role
  const int ts = trans ? tid % T : t;
  const int ps = trans ? tid / T : p;
  const uint64_t Ls = L0 + ts;
  const bool vs = Ls < n_pencils;
  const float2* sl = lds + ts * LL;
  float2* dst;
  uint64_t sk;
  if (trans) {
    const uint64_t x = Ls / ny, y = Ls % ny;
    dst = K + x * (uint64_t)(M + 1) * ny + y;
    sk = ny;
  } else {
    dst = K + Ls * (uint64_t)(M + 1);
    sk = 1;
  }
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int k = ps + P * e;
    const float2 zk = sl[lds_pos<M>(k)];
    const float2 zc = cconj(sl[lds_pos<M>((M - k) % M)]);
    const float2 ze = cscale(cadd(zk, zc), 0.5f);
    const float2 zo = mul_mi<false>(cscale(csub(zk, zc), 0.5f));
    const float2 x = cadd(ze, cmul(twr[k], zo));
    if (vs) dst[k * sk] = x;
  }
  if (ps == 0 && vs) {
    const float2 z0 = sl[lds_pos<M>(0)];
    dst[(uint64_t)M * sk] = make_float2(z0.x - z0.y, 0.0f);
  }
}

// Z'_k = (X_k + conj X_{M-k}) + i W_{2M}^{-k} (X_k - conj X_{M-k}) of the
// C2R below for point k, with xc = X_{M-k} (X_M at k = 0); the imaginary
// parts of X_0 and X_M are ignored, as a C2R transform does.
__device__ __forceinline__ float2 c2r_fold(float2 xk, float2 xc, int k, const float2* __restrict__ twr1) {
  if (k == 0) { xk.y = 0.0f; xc.y = 0.0f; }
  const float2 sm = cadd(xk, cconj(xc));
  const float2 d = csub(xk, cconj(xc));
  const float2 wd = cmul(cconj(twr1[k]), d);
  return make_float2(sm.x - wd.y, sm.y + wd.x);  // sm + i*wd
}

// z pencils, C2R (unnormalised, times `scale`): the inverse M-point FFT of
// Z' (c2r_fold) is x_{2n} + i x_{2n+1}.  XCH 0 exchanges float2 values
// through LDS; XCH 1 the real, then the imaginary parts through a float line
// (half the LDS per pencil, so up to twice the resident waves, for twice the
// barriers).
template <int M, int E, int T, int XCH>
__global__ __launch_bounds__(T*(M / E)) void k_z_c2r(const float2* __restrict__ K, float* __restrict__ out,
                                                     const float2* __restrict__ tw, const float2* __restrict__ twr,
                                                     uint64_t n_pencils, float scale, int* __restrict__ status) {
  constexpr int P = M / E;
  constexpr int LL = lds_line_len(M);
  __shared__ float2 xnyq[T];
  const int tid = threadIdx.x;
  const int p = tid % P, t = tid / P;
  const uint64_t L = (uint64_t)blockIdx.x * T + t;
  const bool valid = L < n_pencils;
  const float2* src = K + L * (uint64_t)(M + 1);
  float2 v[E];
  // one coalesced load per point; the mirror point X_{M-k} comes from LDS
#pragma unroll
  for (int e = 0; e < E; ++e) v[e] = valid ? src[p + P * e] : make_float2(0.f, 0.f);
  if (p == 0) xnyq[t] = valid ? src[M] : make_float2(0.f, 0.f);
  const int z1 = opaque_zero();
  const int p1 = p + z1;
  const float2* twr1 = twr + z1;
  if constexpr (XCH == 0) {
    __shared__ float2 lds[T * LL];
    float2* line = lds + t * LL;
    lds_scatter<M, P, E>(v, line, p);
    __syncthreads();
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int k = p1 + P * e;
      v[e] = c2r_fold(v[e], (k == 0) ? xnyq[t] : line[lds_pos<M>(M - k)], k, twr1);
    }
    __syncthreads();
    fft_line_e<M, E, true>(v, line, p, tw);
  } else {
    __shared__ float ldf[T * LL];
    float* lf = ldf + t * LL;
    float xr[E];
#pragma unroll
    for (int e = 0; e < E; ++e) lf[lds_pos<M>(p + P * e)] = v[e].x;
    __syncthreads();
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int k = p1 + P * e;
      xr[e] = (k == 0) ? xnyq[t].x : lf[lds_pos<M>(M - k)];
    }
    __syncthreads();
#pragma unroll
    for (int e = 0; e < E; ++e) lf[lds_pos<M>(p + P * e)] = v[e].y;
    __syncthreads();
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int k = p1 + P * e;
      const float xi = (k == 0) ? xnyq[t].y : lf[lds_pos<M>(M - k)];
      v[e] = c2r_fold(v[e], make_float2(xr[e], xi), k, twr1);
    }
    __syncthreads();
    fft_line_split<M, E, true>(v, lf, p, tw);
  }
  int flags = 0;
  if (valid) {
    float2* dst = reinterpret_cast<float2*>(out + L * (uint64_t)(2 * M));
    const int p2 = p + opaque_zero();
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const float2 y = cscale(v[e], scale);
      dst[p2 + P * e] = y;
      if (isnan(y.x) || isnan(y.y)) flags |= 1;
      if (y.x != 0.0f || y.y != 0.0f) flags |= 2;
    }
  }
  const int any_nan = __syncthreads_or(flags & 1);
  const int any_nz = __syncthreads_or(flags & 2);
  if (tid == 0 && status != nullptr) {
    const int f = (any_nan ? 1 : 0) | (any_nz ? 2 : 0);
    if (f) atomicOr(status, f);
  }
}

// --------------------------------------------------------- pass variants
// A strided-pass variant: register tile E, lines per block T, LDS exchange
// form XCH (0 float2 line, 1 float line with the split exchange), in
// spectral mode where the aux spectrum waits (TILES, see k_strided), and the
// swizzle shift SWS of unpadded lines.  Variant 0 exists for every length;
// the others for n >= 512, where the passes are long enough to tune.
struct SVar {
  int e, t, xch, tiles, sws, occ;
};
constexpr int SMK_NVAR = SMK_BOX_NVAR;

constexpr bool plan_ok(int n, int e) { return e > 0 && n % e == 0 && make_plan(n, e).nstage > 0; }

// the in-place DIF variant needs n / e threads per line to be a product of
// the plan's leading radices
constexpr bool dif_ok(int n, int e) {
  if (!plan_ok(n, e)) return false;
  const RadixPlan pl = make_plan(n, e);
  int prod = 1;
  for (int s = 0; s <= pl.nstage; ++s) {
    if (prod == n / e) return true;
    if (s < pl.nstage) prod *= pl.r[s];
  }
  return false;
}

constexpr int lines_per_block(int n, int e, int xch, int max_threads) {
  const int bytes = lds_line_len(n) * (xch ? 4 : 8);
  int T = 8;
  while (T > 1 && (T * (n / e) > max_threads || T * bytes > 81920)) T >>= 1;
  return T;
}

// the smaller register tile of the alternative variants (0: none)
constexpr int half_tile(int n) {
  const int e0 = elems_per_thread(n);
  return (e0 % 2 == 0 && plan_ok(n, e0 / 2) && n / (e0 / 2) <= 512) ? e0 / 2 : 0;
}

// Variants (n >= 512; "swz" ones only for the unpadded 2560-point line,
// whose LDS exchange needs a swizzle; the stash ones only in spectral mode):
//   0  E0, float2 exchange            5  E0, split exchange, SWS 3 (swz)
//   1  E0, float2, SWS 2 (swz)        6  E1, float2, SWS 2 (swz) / SWS 0
//   2  E0, float2, SWS 3 (swz)        7  E1, float2, SWS 3 (swz)
//   3  E0, split exchange             8  E0, aux stash in the output, SWS 2 / 0
//   4  E0, split, SWS 2 (swz)         9  E1, aux stash, SWS 2 / 0
//  10  E1, aux in registers, SWS 2 / 0  11  E1, aux in registers, SWS 3 (swz)
//  12  as 10 with half the lines per block and >= 3 waves per SIMD
//  13  as 12 with >= 4 waves per SIMD
//  14  E0, k_xspec: aux in a second register tile (3 FFTs), scalar-base
//      addressing, log2-form spectrum, SWS 2 / 0 (spectral mode only)
//  15  as 14, registers capped for 2 waves per SIMD (spills what is left)
//  16  E0, k_xspec_dif: delta's inverse FFT in place in LDS, the aux
//      spectrum in registers (3 FFTs); lengths whose threads per line are a
//      product of leading radices
// E0 = elems_per_thread(n), E1 = half_tile(n).
constexpr SVar strided_variant(int n, int mode, int vid) {
  const int e0 = elems_per_thread(n);
  const int e1 = half_tile(n);
  const SVar none{0, 0, 0, 0, 0, 0};
  if (e0 == 0) return none;
  const int t0 = lines_per_block(n, e0, 0, 512);
  if (vid == 0) return SVar{e0, t0, 0, 1, 0, 0};
  if (n < 512 || tuned_elems(n) == 0) return none;
  const bool spec = mode == MODE_SPECTRAL;
  const bool swz = !lds_padded(n);
  const int ts = lines_per_block(n, e0, 1, 1024);
  const int t1 = e1 ? lines_per_block(n, e1, 0, 1024) : 0;
  switch (vid) {
    case 1: return swz ? SVar{e0, t0, 0, 1, 2, 0} : none;
    case 2: return swz ? SVar{e0, t0, 0, 1, 3, 0} : none;
    case 3: return SVar{e0, ts, 1, 1, 0, 0};
    case 4: return swz ? SVar{e0, ts, 1, 1, 2, 0} : none;
    case 5: return swz ? SVar{e0, ts, 1, 1, 3, 0} : none;
    case 6: return e1 ? SVar{e1, t1, 0, 1, swz ? 2 : 0, 0} : none;
    case 7: return (e1 && swz) ? SVar{e1, t1, 0, 1, 3, 0} : none;
    case 8: return spec ? SVar{e0, t0, 0, 3, swz ? 2 : 0, 0} : none;
    case 9: return (spec && e1) ? SVar{e1, t1, 0, 3, swz ? 2 : 0, 0} : none;
    case 10: return (spec && e1) ? SVar{e1, t1, 0, 2, swz ? 2 : 0, 0} : none;
    case 11: return (spec && e1 && swz) ? SVar{e1, t1, 0, 2, 3, 0} : none;
    case 12: return (spec && e1 && t1 >= 2) ? SVar{e1, t1 / 2, 0, 2, swz ? 2 : 0, 3} : none;
    case 13: return (spec && e1 && t1 >= 2) ? SVar{e1, t1 / 2, 0, 2, swz ? 2 : 0, 4} : none;
    case 14: return spec ? SVar{e0, t0, 0, 4, swz ? 2 : 0, 1} : none;
    case 15: return spec ? SVar{e0, t0, 0, 4, swz ? 2 : 0, 2} : none;
    case 16: return (spec && dif_ok(n, e0)) ? SVar{e0, t0, 0, 5, swz ? 2 : 0, 0} : none;
    default: return none;
  }
}

template <int N, int MODE, int VID>
hipError_t launch_strided_nv(const StridedArgs& a, uint64_t n_outer, hipStream_t s) {
  constexpr SVar sv = strided_variant(N, MODE, VID);
  if constexpr (sv.e == 0) {
    return hipErrorInvalidValue;
  } else {
    StridedArgs b = a;
    b.n_iblk = (a.n_inner + sv.t - 1) / sv.t;
    const uint64_t nblk = n_outer * b.n_iblk;
    if (nblk > 0x7fffffffull) return hipErrorInvalidConfiguration;
    if constexpr (sv.tiles == 5) {
      hipLaunchKernelGGL((k_xspec_dif<N, sv.e, sv.t, sv.sws>), dim3((uint32_t)nblk), dim3(sv.t * (N / sv.e)), 0, s,
                         b);
    } else if constexpr (sv.tiles == 4) {
      hipLaunchKernelGGL((k_xspec<N, sv.e, sv.t, sv.sws, sv.occ>), dim3((uint32_t)nblk), dim3(sv.t * (N / sv.e)), 0,
                         s, b);
    } else {
      hipLaunchKernelGGL((k_strided<N, sv.e, sv.t, MODE, sv.xch, sv.tiles, sv.sws, sv.occ>), dim3((uint32_t)nblk),
                         dim3(sv.t * (N / sv.e)), 0, s, b);
    }
    return hipGetLastError();
  }
}

template <int N, int MODE>
hipError_t launch_strided_n(int vid, const StridedArgs& a, uint64_t n_outer, hipStream_t s) {
  switch (vid) {
    case 0: return launch_strided_nv<N, MODE, 0>(a, n_outer, s);
    case 1: return launch_strided_nv<N, MODE, 1>(a, n_outer, s);
    case 2: return launch_strided_nv<N, MODE, 2>(a, n_outer, s);
    case 3: return launch_strided_nv<N, MODE, 3>(a, n_outer, s);
    case 4: return launch_strided_nv<N, MODE, 4>(a, n_outer, s);
    case 5: return launch_strided_nv<N, MODE, 5>(a, n_outer, s);
    case 6: return launch_strided_nv<N, MODE, 6>(a, n_outer, s);
    case 7: return launch_strided_nv<N, MODE, 7>(a, n_outer, s);
    case 8: return launch_strided_nv<N, MODE, 8>(a, n_outer, s);
    case 9: return launch_strided_nv<N, MODE, 9>(a, n_outer, s);
    case 10: return launch_strided_nv<N, MODE, 10>(a, n_outer, s);
    case 11: return launch_strided_nv<N, MODE, 11>(a, n_outer, s);
    case 12: return launch_strided_nv<N, MODE, 12>(a, n_outer, s);
    case 13: return launch_strided_nv<N, MODE, 13>(a, n_outer, s);
    case 14: return launch_strided_nv<N, MODE, 14>(a, n_outer, s);
    case 15: return launch_strided_nv<N, MODE, 15>(a, n_outer, s);
    case 16: return launch_strided_nv<N, MODE, 16>(a, n_outer, s);
    default: return hipErrorInvalidValue;
  }
}


inline uint64_t strided_blocks(int n, int mode, int vid, uint64_t n_outer, uint32_t n_inner) {
  const SVar sv = strided_variant(n, mode, vid);
  if (sv.e == 0) return 0;
  return n_outer * ((n_inner + sv.t - 1) / sv.t);
}

// most blocks any variant of this pass launches (sizes the partial sums)
inline uint64_t strided_blocks_max(int n, int mode, uint64_t n_outer, uint32_t n_inner) {
  uint64_t m = 0;
  for (int v = 0; v < SMK_NVAR; ++v) {
    const uint64_t b = strided_blocks(n, mode, v, n_outer, n_inner);
    m = b > m ? b : m;
  }
  return m;
}

// Pencil-pass variants: 0 = the default register tile, 1 = half of it
// (twice the threads per pencil; m >= 256 only), 2 = the default tile with
// the split LDS exchange (C2R only, m >= 256; the half tile with it measured
// slower at m = 768: 18.2 ms against 17.7 for variant 0 and 17.0 for 2).
constexpr int SMK_NPVAR = 3;
constexpr int pencil_E(int m, int vid) {
  return vid == 0 ? elems_per_thread(m)
         : (m < 256 || tuned_elems(m) == 0) ? 0
         : vid == 1 ? half_tile(m) : vid == 2 ? elems_per_thread(m) : 0;
}
constexpr int pencil_XCH(int vid) { return vid == 2 ? 1 : 0; }
inline bool pencil_variant_ok(int m, int vid, bool c2r) {
  return vid >= 0 && vid < SMK_NPVAR && pencil_E(m, vid) != 0 && (c2r || pencil_XCH(vid) == 0);
}

template <int M, int VID>
hipError_t launch_z_r2c_v(const float* noise, uint64_t seed, float2* K, const float2* tw, const float2* twr,
                          uint64_t p0, uint64_t np, int trans, uint32_t ny, hipStream_t s) {
  constexpr int E = pencil_E(M, VID);
  if constexpr (E == 0) {
    return hipErrorInvalidValue;
  } else {
    constexpr int T = pencil_T(M, E);
    if (np <= p0) return hipSuccess;
    const uint64_t nblk = (np - p0 + T - 1) / T;
    if (nblk > 0x7fffffffull) return hipErrorInvalidConfiguration;
    hipLaunchKernelGGL((k_z_r2c<M, E, T>), dim3((uint32_t)nblk), dim3(T * (M / E)), 0, s, noise, seed, K, tw, twr,
                       p0, np, trans, ny);
    return hipGetLastError();
  }
}

template <int M, int VID>
hipError_t launch_z_c2r_v(const float2* K, float* out, const float2* tw, const float2* twr, uint64_t np,
                          float scale, int* status, hipStream_t s) {
  constexpr int E = pencil_E(M, VID);
  if constexpr (E == 0) {
    return hipErrorInvalidValue;
  } else {
    constexpr int T = pencil_T(M, E);
    if ((np + T - 1) / T > 0x7fffffffull) return hipErrorInvalidConfiguration;
    hipLaunchKernelGGL((k_z_c2r<M, E, T, pencil_XCH(VID)>), dim3((uint32_t)((np + T - 1) / T)), dim3(T * (M / E)), 0, s, K, out,
                       tw, twr, np, scale, status);
    return hipGetLastError();
  }
}

// ---------------------------------------------------- per-length launchers
// One entry per line length for each pass family; box.hip dispatches on the
// length and declares these extern, the box_len_<g>.hip units instantiate
// them.
template <int N>
hipError_t launch_len_strided(int mode, int vid, const StridedArgs& a, uint64_t n_outer, hipStream_t s) {
  switch (mode) {
    case MODE_FWD: return launch_strided_n<N, MODE_FWD>(vid, a, n_outer, s);
    case MODE_INV: return launch_strided_n<N, MODE_INV>(vid, a, n_outer, s);
    case MODE_SPECTRAL: return launch_strided_n<N, MODE_SPECTRAL>(vid, a, n_outer, s);
    default: return hipErrorInvalidValue;
  }
}

template <int M>
hipError_t launch_len_r2c(int vid, const float* noise, uint64_t seed, float2* K, const float2* tw, const float2* twr,
                          uint64_t p0, uint64_t np, int trans, uint32_t ny, hipStream_t s) {
  return vid == 0   ? launch_z_r2c_v<M, 0>(noise, seed, K, tw, twr, p0, np, trans, ny, s)
         : vid == 1 ? launch_z_r2c_v<M, 1>(noise, seed, K, tw, twr, p0, np, trans, ny, s)
                    : hipErrorInvalidValue;
}

template <int M>
hipError_t launch_len_c2r(int vid, const float2* K, float* out, const float2* tw, const float2* twr, uint64_t np,
                          float scale, int* status, hipStream_t s) {
  return vid == 0   ? launch_z_c2r_v<M, 0>(K, out, tw, twr, np, scale, status, s)
         : vid == 1 ? launch_z_c2r_v<M, 1>(K, out, tw, twr, np, scale, status, s)
         : vid == 2 ? launch_z_c2r_v<M, 2>(K, out, tw, twr, np, scale, status, s)
                    : hipErrorInvalidValue;
}

#define SMK_LEN_SIGNATURES(SPEC, N)                                                                                \
  SPEC hipError_t launch_len_strided<N>(int, int, const StridedArgs&, uint64_t, hipStream_t);                    \
  SPEC hipError_t launch_len_r2c<N>(int, const float*, uint64_t, float2*, const float2*, const float2*, uint64_t, \
                                    uint64_t, int, uint32_t, hipStream_t);                                        \
  SPEC hipError_t launch_len_c2r<N>(int, const float2*, float*, const float2*, const float2*, uint64_t, float, int*, \
                                    hipStream_t);
#define SMK_LEN_EXTERN(N) SMK_LEN_SIGNATURES(extern template, N)
#define SMK_LEN_INSTANTIATE(N) SMK_LEN_SIGNATURES(template, N)

SMK_FFT_LENGTHS(SMK_LEN_EXTERN)

}  // namespace smk_box
