// Forest stage on gfx950: skewer gather, missing small-scale power and FGPA.
//
// Reference: bin/make_spectra.py:39-139,311-321,429-472 (gather along the
// line of sight) and bin/merge_spectra.py:303-339 (missing power + FGPA).
//
//   k_gather          one thread per pixel: the block's skewers by one
//                     64-wide search over the packed offsets, position and
//                     containing cell in the reference's float64 evaluation
//                     order (bit-exact cell indices), trilinear or Gaussian
//                     smoothing over (2 dmax + 1)^3 cells (343 for the
//                     default dmax = 3) of delta and eta, then the FGPA
//   k_gather_plane_tri  the same for plane-parallel skewers and trilinear
//                     interpolation, with the eight corner loads of both
//                     fields issued together
//   k_missing_power   one FFT line per skewer, T skewers per block, one
//                     launch per length class nz = 256 << c; the classes are
//                     bucketed on the device
#include <math.h>
#include <stdlib.h>

#include <vector>

#include "fft_core.h"
#include "smk_internal.h"

using namespace smk;

namespace {

constexpr int N_CLASS = 6;  // nz = 256 .. 8192

struct GatherArgs {
  smk_skewers sk;
  smk_box_view box;
  int interp;
  const float* abcg;
  int64_t abcg_rows;
  const float* delta_s;
  float* delta_l;
  float* eta_par;
  float* v_par;
  float* flux;
  int32_t* cells;
  int32_t* status;
  int64_t npix_total;
  double hx, hy, hz;  // LX/2, LY/2, LZ/2
};

// last skewer s in [lo, hi] with off[s] <= pix (skips empty skewers)
__device__ __forceinline__ int find_skewer(const int64_t* __restrict__ off, int lo, int hi, int64_t pix) {
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (off[mid] <= pix) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// find_skewer for a whole wave: a 64-ary search, one load per lane and step
// (log64 n steps).  Every lane of the wave must call it; off[lo] <= pix.
__device__ __forceinline__ int wave_find_skewer(const int64_t* __restrict__ off, int lo, int hi, int64_t pix,
                                                int lane) {
  while (lo < hi) {
    const int64_t step = ((int64_t)hi - lo + 64) / 64;
    const int64_t c = (int64_t)lo + (int64_t)lane * step;
    const uint64_t m = __ballot(c <= hi && off[c] <= pix);
    // lane 0 holds (off[lo] <= pix) for packed offsets starting at 0; the
    // clamp keeps the search in range for any offsets
    const int last = __popcll(m) > 0 ? __popcll(m) - 1 : 0;
    lo += (int)(last * step);
    const int64_t h = (int64_t)lo + step - 1;
    hi = h < hi ? (int)h : hi;
  }
  return lo;
}

// i mod n in [0, n); indices within one box length of the box need no division
__device__ __forceinline__ int wrap(int64_t i, int n) {
  if (i >= 0 && i < n) return (int)i;
  if (i >= -(int64_t)n && i < 2 * (int64_t)n) return (int)(i < 0 ? i + n : i - n);
  const int64_t r = i % n;
  return (int)(r < 0 ? r + n : r);
}

// The block's skewer range [s_rng[0], s_rng[1]] for pixels [pix0, plast]
// (pix0 < npix_total), the block covering `block_pixels` pixels: from the
// batch's per-block table when it has one and the block is exactly one table
// stride, else searched by the first wave; every thread must call it.
__device__ __forceinline__ void block_skewers(const smk_skewers& sk, int64_t pix0, int64_t plast, int* s_rng,
                                              int64_t block_pixels) {
  if (sk.blk_s0 != nullptr && block_pixels == SMK_GATHER_BLOCK_PIXELS &&
      pix0 == (int64_t)blockIdx.x * SMK_GATHER_BLOCK_PIXELS) {
    if (threadIdx.x < 2) {
      const int v = sk.blk_s0[blockIdx.x + threadIdx.x];
      s_rng[threadIdx.x] = v < 0 ? 0 : (v >= sk.n ? sk.n - 1 : v);
    }
    __syncthreads();
    return;
  }
  if (threadIdx.x < 64) {
    const int lane = threadIdx.x;
    const int s0 = wave_find_skewer(sk.off, 0, sk.n - 1, pix0, lane);
    const int64_t c = (int64_t)s0 + lane;
    const int cnt = __popcll(__ballot(c < sk.n && sk.off[c] <= plast));
    const int s1 = cnt < 64 ? s0 + (cnt > 0 ? cnt - 1 : 0)
                            : wave_find_skewer(sk.off, s0 + 63, sk.n - 1, plast, lane);
    if (lane == 0) {
      s_rng[0] = s0;
      s_rng[1] = s1;
    }
  }
  __syncthreads();
}

// NF fields: 1 (delta), 2 (delta, eta_zz), 7 (delta and the six eta_ij) or
// 10 (those and the three velocities); DM: the Gaussian stencil's radius in
// cells (make_spectra -dmax), or -1 for trilinear interpolation (one
// instance per kind, so that neither carries the other's registers)
template <int NF, int DM>
__global__ __launch_bounds__(256) void k_gather(GatherArgs a) {
  const int64_t pix0 = (int64_t)blockIdx.x * blockDim.x;
  const int64_t pix = pix0 + threadIdx.x;
  const smk_skewers& sk = a.sk;
  // the skewers of this block's pixel range, searched once by the first wave
  __shared__ int s_rng[2];
  block_skewers(sk, pix0, pix0 + blockDim.x - 1 < a.npix_total ? pix0 + blockDim.x - 1 : a.npix_total - 1, s_rng,
                (int64_t)blockDim.x);
  if (pix >= a.npix_total) return;
  const int s = find_skewer(sk.off, s_rng[0], s_rng[1], pix);
  const int i = (int)(pix - sk.off[s]);
  const int64_t g = (int64_t)sk.g0[s] + i;
  // R_g = rmin + g*dr (make_spectra.py:312), no contraction
  const double R = add_rn(sk.rmin, mul_rn((double)g, sk.dr));
  double X, Y, Z;
  if (sk.geometry == SMK_GEOM_RADIAL) {
    // Xvec = Rvec * X_QSO / R_QSO (make_spectra.py:450-452)
    const double qr = sk.qr[s];
    X = div_rn(mul_rn(R, sk.qx[s]), qr);
    Y = div_rn(mul_rn(R, sk.qy[s]), qr);
    Z = div_rn(mul_rn(R, sk.qz[s]), qr);
  } else {
    X = sk.qx[s];
    Y = sk.qy[s];
    Z = R;
  }
  const smk_box_view& b = a.box;
  const double D = b.cell;
  // make_spectra.py:47-49: int((X + LX/2)/DX), int((Z + LZ/2 - R0)/DZ)
  const int64_t ix = (int64_t)div_rn(add_rn(X, a.hx), D);
  const int64_t iy = (int64_t)div_rn(add_rn(Y, a.hy), D);
  const int64_t iz = (int64_t)div_rn(sub_rn(add_rn(Z, a.hz), b.r0), D);
  if (a.cells) {
    a.cells[3 * pix + 0] = (int32_t)ix;
    a.cells[3 * pix + 1] = (int32_t)iy;
    a.cells[3 * pix + 2] = (int32_t)iz;
  }
  const bool in_forest = i >= sk.imin[s] && i < sk.imax[s];
  const bool radial = sk.geometry == SMK_GEOM_RADIAL;
  // FGPA inputs, loaded before the field gather so their latency overlaps it
  const int64_t row = sk.pset ? (int64_t)sk.pset[s] : g;
  const bool row_ok = row >= 0 && row < a.abcg_rows;
  float4 prow = make_float4(0.f, 0.f, 0.f, 0.f);
  float ds = 0.0f;
  if (a.flux) {
    if (row_ok) prow = reinterpret_cast<const float4*>(a.abcg)[row];
    if (a.delta_s) ds = a.delta_s[pix];
  }
  float dl = -1000000.0f, eta = 0.0f, vpar = 0.0f;
  bool outside = false, off_slab = false;
  if (in_forest) {
    const int nx = b.nx, ny = b.ny, nz = b.nz;
    // resident x planes [x0, x0 + nxs); the field pointers point at plane x0
    const int64_t x0 = b.nxs > 0 ? b.x0 : 0, nxs = b.nxs > 0 ? b.nxs : nx;
    // fields: 0 = delta, 1..6 = eta_xx..eta_yz (radial) or 1 = eta_zz (plane)
    constexpr int nf = NF;
    double acc[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    const float* fld[10] = {b.delta, radial ? b.eta[0] : b.eta[2], b.eta[1], b.eta[2], b.eta[3], b.eta[4],
                            b.eta[5], b.vel[0], b.vel[1], b.vel[2]};
    if constexpr (DM >= 0) {
      // cell centres (i+0.5)D - L/2 (+R0 along Z), weights exp(-dr^2/(2 D^2)),
      // taken as the product of the three axes' factors (3 W exponentials,
      // not W^3)
      constexpr int W = 2 * (DM >= 0 ? DM : 0) + 1;
      const double cx = sub_rn(mul_rn((double)ix + 0.5, D), a.hx);
      const double cy = sub_rn(mul_rn((double)iy + 0.5, D), a.hy);
      const double cz = add_rn(sub_rn(mul_rn((double)iz + 0.5, D), a.hz), b.r0);
      const double sig2 = 2.0 * D * D;
      double wy[W], wz[W];
      int64_t jyv[W], jzv[W];
      for (int d = 0; d < W; ++d) {
        const double ey = sub_rn(add_rn((d - DM) * D, cy), Y);
        const double ez = sub_rn(add_rn((d - DM) * D, cz), Z);
        int64_t jy = iy + d - DM, jz = iz + d - DM;
        if (b.periodic) {
          jy = wrap(jy, ny);
          jz = wrap(jz, nz);
        } else if (jy < 0 || jy >= ny || jz < 0 || jz >= nz) {
          outside = true;
        }
        wy[d] = exp(-mul_rn(ey, ey) / sig2);
        wz[d] = exp(-mul_rn(ez, ez) / sig2);
        jyv[d] = jy;
        jzv[d] = jz;
      }
      double wsum = 0.0;
      for (int da = -DM; da <= DM; ++da) {
        const double ex = sub_rn(add_rn(da * D, cx), X);
        int64_t jx = ix + da;
        if (b.periodic) jx = wrap(jx, nx); else if (jx < 0 || jx >= nx) outside = true;
        // a cell inside the box but not among the resident planes
        if (jx >= 0 && jx < nx && (jx < x0 || jx >= x0 + nxs)) off_slab = true;
        if (outside || off_slab) continue;
        const double wx = exp(-mul_rn(ex, ex) / sig2);
        for (int db = 0; db < W; ++db) {
          const double wxy = wx * wy[db];
          const int64_t row = ((jx - x0) * ny + jyv[db]) * nz;
#pragma unroll
          for (int dc = 0; dc < W; ++dc) {
            const double w = wxy * wz[dc];
            wsum += w;
            const int64_t flat = row + jzv[dc];
#pragma unroll
            for (int f = 0; f < nf; ++f) acc[f] += w * (double)fld[f][flat];
          }
        }
      }
      for (int f = 0; f < nf; ++f) acc[f] /= wsum;
    } else {
      // trilinear on cell-centred values; indices wrap in a periodic box
      const double ux = div_rn(add_rn(X, a.hx), D) - 0.5;
      const double uy = div_rn(add_rn(Y, a.hy), D) - 0.5;
      const double uz = div_rn(sub_rn(add_rn(Z, a.hz), b.r0), D) - 0.5;
      const double fx0 = floor(ux), fy0 = floor(uy), fz0 = floor(uz);
      const double fx = ux - fx0, fy = uy - fy0, fz = uz - fz0;
      const int64_t i0x = (int64_t)fx0, i0y = (int64_t)fy0, i0z = (int64_t)fz0;
      for (int c = 0; c < 8; ++c) {
        const int ox = c >> 2, oy = (c >> 1) & 1, oz = c & 1;
        const double w = (ox ? fx : 1.0 - fx) * (oy ? fy : 1.0 - fy) * (oz ? fz : 1.0 - fz);
        int64_t jx = i0x + ox, jy = i0y + oy, jz = i0z + oz;
        if (b.periodic) {
          jx = wrap(jx, nx);
          jy = wrap(jy, ny);
          jz = wrap(jz, nz);
        } else if (jx < 0 || jx >= nx || jy < 0 || jy >= ny || jz < 0 || jz >= nz) {
          outside = true;
        }
        if (jx >= 0 && jx < nx && (jx < x0 || jx >= x0 + nxs)) off_slab = true;
        if (outside || off_slab) continue;
        const int64_t flat = ((jx - x0) * ny + jy) * nz + jz;
        for (int f = 0; f < nf; ++f) acc[f] += w * (double)fld[f][flat];
      }
    }
    if (outside || off_slab) {
      dl = __int_as_float(0x7fc00000);
      eta = dl;
      vpar = dl;
      if (a.status) atomicOr(&a.status[s], (outside ? 1 : 0) | (off_slab ? 4 : 0));
    } else {
      dl = (float)acc[0];
      if (nf > 1) {
        if (radial) {
          // make_spectra.py:127-129 with Xtrue = X (no slab shift)
          const double RR = X * X + Y * Y + Z * Z;
          const double e = (X * acc[1] * X + Y * acc[2] * Y + Z * acc[3] * Z + 2 * X * acc[4] * Y +
                            2 * X * acc[5] * Z + 2 * Y * acc[6] * Z) / RR;
          eta = (float)e;
          // make_spectra.py:134: v_par = (vx X + vy Y + vz Z) / sqrt(RR)
          if (nf == 10) vpar = (float)((acc[7] * X + acc[8] * Y + acc[9] * Z) / sqrt(RR));
        } else {
          eta = (float)acc[1];
        }
      }
    }
  }
  if (a.delta_l) a.delta_l[pix] = dl;
  if (a.eta_par) a.eta_par[pix] = eta;
  if (a.v_par) a.v_par[pix] = vpar;
  if (a.flux) {
    if (!row_ok) {
      a.flux[pix] = __int_as_float(0x7fc00000);
      if (a.status) atomicOr(&a.status[s], 2);
    } else {
      // F = exp(-a exp(b G (delta + c eta)))
      const float gfield = dl + ds + prow.z * eta;
      a.flux[pix] = expf(-prow.x * expf(prow.y * prow.w * gfield));
    }
  }
}

// k_gather for plane-parallel skewers with trilinear interpolation, PPT
// pixels per thread (pix0 + threadIdx.x + 256 k): positions, cells and
// corner indices of all of a thread's pixels first, then all their field
// loads at once (no early exits between them), then the sums, so that the
// loads' latency overlaps.  Same arithmetic, in the same order, as k_gather.
template <int PPT>
__global__ __launch_bounds__(256) void k_gather_plane_tri(GatherArgs a) {
  const int64_t pix0 = (int64_t)blockIdx.x * (256 * PPT);
  const smk_skewers& sk = a.sk;
  __shared__ int s_rng[2];
  {
    const int64_t pl = pix0 + 256 * PPT - 1;
    block_skewers(sk, pix0, pl < a.npix_total ? pl : a.npix_total - 1, s_rng, (int64_t)256 * PPT);
  }
  const smk_box_view& b = a.box;
  const int nx = b.nx, ny = b.ny, nz = b.nz;
  const int64_t x0 = b.nxs > 0 ? b.x0 : 0, nxs = b.nxs > 0 ? b.nxs : nx;
  const bool has_eta = b.eta[2] != nullptr;
  const double D = b.cell;

  int s_[PPT];
  bool act[PPT], inf[PPT], bad[PPT], outs[PPT], offs[PPT];
  double fx_[PPT], fy_[PPT], fz_[PPT];
  int64_t rx_[PPT][2];
  int jy_[PPT][2], jz_[PPT][2];
  float4 prow[PPT];
  float dsv[PPT];
  bool row_ok[PPT];
#pragma unroll
  for (int k = 0; k < PPT; ++k) {
    const int64_t pix = pix0 + threadIdx.x + 256 * k;
    act[k] = pix < a.npix_total;
    const int s = act[k] ? find_skewer(sk.off, s_rng[0], s_rng[1], pix) : s_rng[0];
    s_[k] = s;
    const int i = (int)(pix - sk.off[s]);
    const int64_t g = (int64_t)sk.g0[s] + i;
    const double X = sk.qx[s], Y = sk.qy[s];
    // R_g = rmin + g*dr (make_spectra.py:312), no contraction
    const double Z = add_rn(sk.rmin, mul_rn((double)g, sk.dr));
    const double qx = add_rn(X, a.hx), qy = add_rn(Y, a.hy), qz = sub_rn(add_rn(Z, a.hz), b.r0);
    // make_spectra.py:47-49
    const int64_t ix = (int64_t)div_rn(qx, D), iy = (int64_t)div_rn(qy, D), iz = (int64_t)div_rn(qz, D);
    if (a.cells && act[k]) {
      a.cells[3 * pix + 0] = (int32_t)ix;
      a.cells[3 * pix + 1] = (int32_t)iy;
      a.cells[3 * pix + 2] = (int32_t)iz;
    }
    inf[k] = act[k] && i >= sk.imin[s] && i < sk.imax[s];
    const int64_t row = sk.pset ? (int64_t)sk.pset[s] : g;
    row_ok[k] = row >= 0 && row < a.abcg_rows;
    prow[k] = make_float4(0.f, 0.f, 0.f, 0.f);
    dsv[k] = 0.0f;
    if (a.flux && act[k]) {
      if (row_ok[k]) prow[k] = reinterpret_cast<const float4*>(a.abcg)[row];
      if (a.delta_s) dsv[k] = a.delta_s[pix];
    }
    // trilinear on cell-centred values; indices wrap in a periodic box
    const double ux = div_rn(qx, D) - 0.5, uy = div_rn(qy, D) - 0.5, uz = div_rn(qz, D) - 0.5;
    const double fx0 = floor(ux), fy0 = floor(uy), fz0 = floor(uz);
    fx_[k] = ux - fx0;
    fy_[k] = uy - fy0;
    fz_[k] = uz - fz0;
    bool outside = false, off_slab = false;
    int64_t jx[2], jy[2], jz[2];
#pragma unroll
    for (int o = 0; o < 2; ++o) {
      jx[o] = (int64_t)fx0 + o;
      jy[o] = (int64_t)fy0 + o;
      jz[o] = (int64_t)fz0 + o;
      if (b.periodic) {
        jx[o] = wrap(jx[o], nx);
        jy[o] = wrap(jy[o], ny);
        jz[o] = wrap(jz[o], nz);
      } else if (jx[o] < 0 || jx[o] >= nx || jy[o] < 0 || jy[o] >= ny || jz[o] < 0 || jz[o] >= nz) {
        outside = true;
      }
      if (jx[o] >= 0 && jx[o] < nx && (jx[o] < x0 || jx[o] >= x0 + nxs)) off_slab = true;
    }
    outs[k] = outside;
    offs[k] = off_slab;
    bad[k] = outside || off_slab || !inf[k];
#pragma unroll
    for (int o = 0; o < 2; ++o) {
      // a pixel whose sum is not used loads cell 0 (always resident)
      rx_[k][o] = bad[k] ? 0 : (jx[o] - x0) * ny;
      jy_[k][o] = bad[k] ? 0 : (int)jy[o];
      jz_[k][o] = bad[k] ? 0 : (int)jz[o];
    }
  }
  float vd[PPT][8], ve[PPT][8];
#pragma unroll
  for (int k = 0; k < PPT; ++k) {
#pragma unroll
    for (int c = 0; c < 8; ++c) {
      const int64_t flat = (rx_[k][c >> 2] + jy_[k][(c >> 1) & 1]) * nz + jz_[k][c & 1];
      vd[k][c] = b.delta[flat];
      ve[k][c] = has_eta ? b.eta[2][flat] : 0.0f;
    }
  }
#pragma unroll
  for (int k = 0; k < PPT; ++k) {
    if (!act[k]) continue;
    const int64_t pix = pix0 + threadIdx.x + 256 * k;
    const int s = s_[k];
    float dl = -1000000.0f, eta = 0.0f;
    if (inf[k]) {
      if (outs[k] || offs[k]) {
        dl = __int_as_float(0x7fc00000);
        eta = dl;
        if (a.status) atomicOr(&a.status[s], (outs[k] ? 1 : 0) | (offs[k] ? 4 : 0));
      } else {
        const double fx = fx_[k], fy = fy_[k], fz = fz_[k];
        double acc0 = 0.0, acc1 = 0.0;
#pragma unroll
        for (int c = 0; c < 8; ++c) {
          const int ox = c >> 2, oy = (c >> 1) & 1, oz = c & 1;
          const double w = (ox ? fx : 1.0 - fx) * (oy ? fy : 1.0 - fy) * (oz ? fz : 1.0 - fz);
          acc0 += w * (double)vd[k][c];
          acc1 += w * (double)ve[k][c];
        }
        dl = (float)acc0;
        if (has_eta) eta = (float)acc1;
      }
    }
    if (a.delta_l) a.delta_l[pix] = dl;
    if (a.eta_par) a.eta_par[pix] = eta;
    if (a.flux) {
      if (!row_ok[k]) {
        a.flux[pix] = __int_as_float(0x7fc00000);
        if (a.status) atomicOr(&a.status[s], 2);
      } else {
        // F = exp(-a exp(b G (delta + c eta)))
        const float gfield = dl + dsv[k] + prow[k].z * eta;
        a.flux[pix] = expf(-prow[k].x * expf(prow[k].y * prow[k].w * gfield));
      }
    }
  }
}

// ------------------------------------------------------------ missing power
// Fourier modes of real N(0,1) white noise of length 2M: interior modes have
// Re, Im ~ N(0, M); k = 0 and k = M are real with variance 2M.  M/2 Philox
// calls per skewer (counter c), each two Box-Muller pairs: call 0 gives
// modes 0 and M (the first pair, one normal each) and mode 1 (the second);
// call c >= 1 gives modes 2c (first pair) and 2c + 1 (second pair).
constexpr uint32_t MP_TAG = 0x6d697373u;

__device__ __forceinline__ U4 mp_call(uint64_t seed, uint64_t id, int c) {
  return philox4x32_10_c4((uint32_t)c, (uint32_t)id, (uint32_t)(id >> 32), MP_TAG, seed);
}

__device__ __forceinline__ float2 mp_mode(uint64_t seed, uint64_t id, int k, int M, const float* __restrict__ filt) {
  const int c = (k == M || k == 1) ? 0 : k >> 1;
  const U4 u = mp_call(seed, id, c);
  float n0, n1;
  if (k == 0 || k == M || (k & 1) == 0) box_muller_hw(u.x, u.y, n0, n1);
  else box_muller_hw(u.z, u.w, n0, n1);
  if (k == 0) return make_float2(n0 * sqrtf(2.0f * M) * filt[k], 0.0f);
  if (k == M) return make_float2(n1 * sqrtf(2.0f * M) * filt[k], 0.0f);
  const float sm = sqrtf((float)M) * filt[k];
  return make_float2(n0 * sm, n1 * sm);
}

template <int M, int T>
__global__ __launch_bounds__(T* (M / elems_per_thread(M))) void k_missing_power(
    smk_skewers sk, const int32_t* __restrict__ list, const int32_t* __restrict__ count,
    const float* __restrict__ filt_all, const int32_t* __restrict__ fset, int64_t set_stride,
    const float* __restrict__ scale, uint64_t seed_arg, const uint64_t* __restrict__ seed_dev,
    const float* __restrict__ noise, const int64_t* __restrict__ noise_off, const float2* __restrict__ tw,
    const float2* __restrict__ twr, float* __restrict__ delta_s) {
  constexpr int E = elems_per_thread(M);
  constexpr int P = M / E;
  constexpr int LL = lds_line_len(M);
  __shared__ float2 lds[T * LL];
  const int ncls = *count;
  const int base = blockIdx.x * T;
  if (base >= ncls) return;  // uniform across the block
  // the seed of a graph replay is read from device memory
  const uint64_t seed = seed_dev ? *seed_dev : seed_arg;
  const int tid = threadIdx.x;
  const int p = tid % P, t = tid / P;
  const bool valid = base + t < ncls;
  const int s = valid ? list[base + t] : 0;
  // the skewer's filter set (one P1D_miss table per set)
  const float* __restrict__ filt = filt_all + (fset ? (int64_t)fset[s] * set_stride : 0);
  float2* line = lds + t * LL;
  float2 v[E];
  if (noise == nullptr) {
    const uint64_t id = sk.noise_id ? sk.noise_id[s] : (uint64_t)s;
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int k = p + P * e;
      const float2 yk = mp_mode(seed, id, k, M, filt);
      const float2 yc = cconj(mp_mode(seed, id, M - k, M, filt));
      const float2 sm = cadd(yk, yc);
      const float2 d = csub(yk, yc);
      const float2 wd = cmul(cconj(twr[k]), d);
      v[e] = make_float2(sm.x - wd.y, sm.y + wd.x);
    }
  } else {
    // forward R2C of the fed noise, times the filter, then the C2R pre-step
    const float2* src = reinterpret_cast<const float2*>(noise + (valid ? noise_off[s] : 0));
#pragma unroll
    for (int e = 0; e < E; ++e) v[e] = valid ? src[p + P * e] : make_float2(0.f, 0.f);
    fft_line<M, false>(v, line, p, tw);
    lds_scatter<M, P, E>(v, line, p);
    __syncthreads();
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int k = p + P * e;
      const float2 zk = line[lds_pos<M>(k)];
      const float2 zc = line[lds_pos<M>((M - k) % M)];
      // X_k and X_{M-k} of the real transform
      const float2 zcc = cconj(zc);
      const float2 xk = cadd(cscale(cadd(zk, zcc), 0.5f), cmul(twr[k], mul_mi<false>(cscale(csub(zk, zcc), 0.5f))));
      float2 xc;
      if (k == 0) {
        xc = make_float2(zk.x - zk.y, 0.0f);   // X_M
      } else {
        const float2 zkc = cconj(zk);
        const float2 wm = cconj(twr[k]);       // W_{2M}^{M-k} = -conj(W_{2M}^k)
        xc = csub(cscale(cadd(zc, zkc), 0.5f), cmul(wm, mul_mi<false>(cscale(csub(zc, zkc), 0.5f))));
      }
      float2 yk = cscale(xk, filt[k]);
      float2 yc = cscale(xc, filt[M - k]);
      if (k == 0) { yk.y = 0.0f; yc.y = 0.0f; }
      const float2 ycc = cconj(yc);
      const float2 sm = cadd(yk, ycc);
      const float2 d = csub(yk, ycc);
      const float2 wd = cmul(cconj(twr[k]), d);
      v[e] = make_float2(sm.x - wd.y, sm.y + wd.x);
    }
    __syncthreads();
  }
  fft_line<M, true>(v, line, p, tw);
  if (valid) {
    const int n = sk.npix[s];
    float* dst = delta_s + sk.off[s];
    const float inv = 1.0f / (2.0f * M);
    if (scale == nullptr) {
#pragma unroll
      for (int e = 0; e < E; ++e) {
        const int j = 2 * (p + P * e);
        if (j < n) dst[j] = v[e].x * inv;
        if (j + 1 < n) dst[j + 1] = v[e].y * inv;
      }
    } else {
      const float* sc = scale + sk.off[s];
#pragma unroll
      for (int e = 0; e < E; ++e) {
        const int j = 2 * (p + P * e);
        if (j < n) dst[j] = v[e].x * inv * sc[j];
        if (j + 1 < n) dst[j + 1] = v[e].y * inv * sc[j + 1];
      }
    }
  }
}

__global__ void k_classify(smk_skewers sk, const int32_t* __restrict__ fset, int32_t n_sets,
                           int32_t* __restrict__ lists, int32_t* __restrict__ counts, int32_t* __restrict__ bad) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  int c = -1;  // length class, -1: none
  if (s < sk.n) {
    if (fset && (fset[s] < 0 || fset[s] >= n_sets)) {
      atomicOr(bad, 2);
    } else {
      const int n = sk.npix[s];
      int nz = 256;
      c = 0;
      while (nz < n + 50 && c < N_CLASS) { nz <<= 1; ++c; }
      if (c >= N_CLASS) { atomicOr(bad, 1); c = -1; }
    }
  }
  // one atomic per class and wave (the order within a class list is free:
  // each skewer's missing power depends only on its own index)
  const int lane = threadIdx.x & 63;
  const uint64_t below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
  for (int k = 0; k < N_CLASS; ++k) {
    const uint64_t m = __ballot(c == k);
    if (m == 0) continue;
    const int leader = __ffsll((unsigned long long)m) - 1;
    int base = 0;
    if (lane == leader) base = atomicAdd(&counts[k], __popcll(m));
    base = __shfl(base, leader, 64);
    if (c == k) lists[(int64_t)k * sk.n + base + __popcll(m & below)] = s;
  }
}

// ------------------------------------------------------ fused forest step
// One persistent launch for a batch of plane-parallel skewers: each block
// takes skewers from a device work queue and, per skewer, (1) draws the
// missing-power Fourier modes into LDS (one Philox call per two modes),
// (2) folds them for the C2R transform and inverse-FFTs them in LDS, so that
// delta_s stays in LDS, (3) interpolates delta and eta_zz trilinearly along
// the line of sight and applies the FGPA.  The x and y terms of the
// interpolation (cells, weights, the four z columns) are computed once per
// skewer; along z the position advances by dr/D per pixel from one division
// per skewer.  Float32 weights and sums (the result is stored as float32).

struct StepArgs {
  smk_skewers sk;
  smk_box_view box;
  const float* filt;
  int64_t foff[N_CLASS];
  const int32_t* fset;
  int64_t set_stride;
  int32_t n_sets;
  const float* scale;
  const float* abcg;
  int64_t abcg_rows;
  uint64_t seed;
  const uint64_t* seed_dev;
  const float2* tw[N_CLASS];
  const float2* twr[N_CLASS];
  float* flux;
  float* delta_l;
  float* eta_par;
  float* delta_s;
  int32_t* status;
  int32_t* queue;
  int32_t max_class;
  double hx, hy, hz;
  int32_t ctab_cap;      // z cells the per-skewer column table holds (0: no table)
  int32_t ctab_off;      // its offset in the dynamic LDS, in float2
};

// Block of the fused step: one wave per skewer (NT = 64) for batches whose
// FFTs are at most 2048 points (nz <= 2048, n <= 1998), four waves (NT =
// 256) for longer ones.  Register tile of the missing-power FFT of M points:
// all NT threads on one line when M >= 2 NT, else NT / (M / E) thread groups,
// each transforming the same line in an LDS line of its own.
constexpr int step_E(int M, int NT) { return M >= 2 * NT ? M / NT : 2; }
constexpr int step_groups(int M, int NT) { return NT / (M / step_E(M, NT)); }
constexpr int step_lds_float2(int M, int NT) {
  const int lines = step_groups(M, NT) * lds_line_len(M);
  return lines > M + 1 ? lines : M + 1;
}
constexpr int step_lds_max(int c, int NT) {
  int m = 0;
  for (int i = 0; i <= c; ++i) {
    const int v = step_lds_float2(128 << i, NT);
    m = v > m ? v : m;
  }
  return m;
}

// delta_s of skewer s (n pixels) into dsf[0, n) (LDS floats); length class
// C, M = 128 << C
template <int C, int NT>
__device__ __forceinline__ void step_missing_power(const StepArgs& a, int s, int n, uint64_t seed,
                                                   float2* __restrict__ sh) {
  constexpr int M = 128 << C;
  constexpr int E = step_E(M, NT);
  constexpr int P = M / E;
  constexpr int LL = lds_line_len(M);
  static_assert(P * step_groups(M, NT) == NT, "thread groups must tile the block");
  const int tid = threadIdx.x;
  const float* __restrict__ filt = a.filt + a.foff[C] + (a.fset ? (int64_t)a.fset[s] * a.set_stride : 0);
  const uint64_t id = a.sk.noise_id ? a.sk.noise_id[s] : (uint64_t)s;
  // (1) modes k = 0..M: M/2 Philox calls (see mp_mode)
  const float sr = sqrtf((float)M), s2 = sqrtf(2.0f * M);
  // a compile-time trip count, so that the filter loads of every pass are
  // issued together ahead of the Philox rounds
  constexpr int NPASS = (M / 2 + NT - 1) / NT;
#pragma unroll
  for (int it = 0; it < NPASS; ++it) {
    const int c = tid + it * NT;
    if (NPASS * NT > M / 2 && c >= M / 2) break;
    const U4 u = mp_call(seed, id, c);
    float n0, n1, n2, n3;
    box_muller_hw(u.x, u.y, n0, n1);
    box_muller_hw(u.z, u.w, n2, n3);
    if (c == 0) {
      sh[0] = make_float2(n0 * s2 * filt[0], 0.0f);
      sh[M] = make_float2(n1 * s2 * filt[M], 0.0f);
      sh[1] = make_float2(n2 * sr * filt[1], n3 * sr * filt[1]);
    } else {
      const float f0 = filt[2 * c] * sr, f1 = filt[2 * c + 1] * sr;
      sh[2 * c] = make_float2(n0 * f0, n1 * f0);
      sh[2 * c + 1] = make_float2(n2 * f1, n3 * f1);
    }
  }
  __syncthreads();
  // (2) C2R fold: Z'_k = (Y_k + conj Y_{M-k}) + i W_2M^-k (Y_k - conj Y_{M-k})
  const int g = tid / P, p = tid % P;
  const float2* __restrict__ twr = a.twr[C];
  float2 v[E];
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int k = p + P * e;
    const float2 yk = sh[k];
    const float2 yc = cconj(sh[M - k]);
    const float2 sm = cadd(yk, yc);
    const float2 d = csub(yk, yc);
    const float2 wd = cmul(cconj(twr[k]), d);
    v[e] = make_float2(sm.x - wd.y, sm.y + wd.x);
  }
  __syncthreads();
  // (3) inverse FFT: x_2k + i x_2k+1, k = p + P e
  fft_line_e<M, E, true>(v, sh + g * LL, p, a.tw[C]);
  __syncthreads();
  if (g == 0) {
    float* dsf = reinterpret_cast<float*>(sh);
    const float inv = 1.0f / (2.0f * M);
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int j = 2 * (p + P * e);
      if (j < n) dsf[j] = v[e].x * inv;
      if (j + 1 < n) dsf[j + 1] = v[e].y * inv;
    }
  }
  __syncthreads();
}

// MAXC: the longest length class of the batch (classes above it are
// flagged); NT threads per block
template <int MAXC, int NT, int OCC>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(OCC))) void k_forest_step(StepArgs a) {
  extern __shared__ float2 sh[];
  __shared__ int s_next;
  const int tid = threadIdx.x;
  const smk_skewers& sk = a.sk;
  const smk_box_view& b = a.box;
  const uint64_t seed = a.seed_dev ? *a.seed_dev : a.seed;
  const int nx = b.nx, ny = b.ny, nz = b.nz;
  const int64_t x0 = b.nxs > 0 ? b.x0 : 0, nxs = b.nxs > 0 ? b.nxs : nx;
  const float* __restrict__ fd = b.delta;
  const float* __restrict__ fe = b.eta[2];
  const double D = b.cell;
  // the queue slot of the next skewer is taken while this one is processed
  // (its atomic's latency overlaps the work); every slot goes to one block
  int q_next = 0;
  if (tid == 0) q_next = atomicAdd(a.queue, 1);
  for (;;) {
    if (tid == 0) s_next = q_next;
    __syncthreads();
    // the same value in every lane: scalar registers for all per-skewer terms
    const int s = __builtin_amdgcn_readfirstlane(s_next);
    __syncthreads();
    if (s >= sk.n) break;
    if (tid == 0) q_next = atomicAdd(a.queue, 1);
    const int n = sk.npix[s];
    int c = 0;
    for (int nzc = 256; nzc < n + 50 && c < N_CLASS; nzc <<= 1) ++c;
    const int64_t off = sk.off[s];
    if (c > a.max_class || (a.fset && (a.fset[s] < 0 || a.fset[s] >= a.n_sets))) {
      // longer than the caller's max_npix bound, or a filter set outside the
      // tables (possible only in graph mode, which skips the host check):
      // flagged, flux NaN
      if (tid == 0 && a.status) atomicOr(&a.status[s], 8);
      for (int i = tid; i < n; i += NT) a.flux[off + i] = __int_as_float(0x7fc00000);
      continue;
    }
    if (c == 0) step_missing_power<0, NT>(a, s, n, seed, sh);
    else if (c == 1) step_missing_power<1, NT>(a, s, n, seed, sh);
    else if (c == 2) step_missing_power<2, NT>(a, s, n, seed, sh);
    else if (c == 3 || MAXC <= 3) step_missing_power<(MAXC < 3 ? MAXC : 3), NT>(a, s, n, seed, sh);
    else if (c == 4 || MAXC <= 4) step_missing_power<(MAXC < 4 ? MAXC : 4), NT>(a, s, n, seed, sh);
    else step_missing_power<MAXC, NT>(a, s, n, seed, sh);
    const float* dsf = reinterpret_cast<const float*>(sh);
    // per-skewer x and y terms (make_spectra.py:47-49 positions; trilinear on
    // cell-centred values, periodic indices in a periodic box)
    const double ux = div_rn(add_rn(sk.qx[s], a.hx), D) - 0.5;
    const double uy = div_rn(add_rn(sk.qy[s], a.hy), D) - 0.5;
    const double fx0 = floor(ux), fy0 = floor(uy);
    const double fx = ux - fx0, fy = uy - fy0;
    bool outside = false, off_slab = false;
    int64_t jx[2], jy[2];
#pragma unroll
    for (int o = 0; o < 2; ++o) {
      jx[o] = (int64_t)fx0 + o;
      jy[o] = (int64_t)fy0 + o;
      if (b.periodic) {
        jx[o] = wrap(jx[o], nx);
        jy[o] = wrap(jy[o], ny);
      } else if (jx[o] < 0 || jx[o] >= nx || jy[o] < 0 || jy[o] >= ny) {
        outside = true;
      }
      if (jx[o] >= 0 && jx[o] < nx && (jx[o] < x0 || jx[o] >= x0 + nxs)) off_slab = true;
    }
    const bool xy_bad = outside || off_slab;
    // the four z columns (ox, oy) and their weights
    int64_t col[4];
    float wxy[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int ox = q >> 1, oy = q & 1;
      col[q] = xy_bad ? 0 : ((jx[ox] - x0) * ny + jy[oy]) * nz;
      wxy[q] = (float)((ox ? fx : 1.0 - fx) * (oy ? fy : 1.0 - fy));
    }
    // z of pixel i: uz = ((rmin + g dr) + LZ/2 - R0)/D - 0.5 with g = g0 + i,
    // advanced by dr/D per pixel from pixel 0
    const int64_t g0 = sk.g0[s];
    const double z0 = add_rn(sk.rmin, mul_rn((double)g0, sk.dr));
    const double uz0 = div_rn(sub_rn(add_rn(z0, a.hz), b.r0), D) - 0.5;
    const double duz = sk.dr / D;
    const int imin = sk.imin[s], imax = sk.imax[s];
    const bool per_pixel_row = sk.pset == nullptr;
    const int64_t srow = per_pixel_row ? 0 : (int64_t)sk.pset[s];
    float4 prow = make_float4(0.f, 0.f, 0.f, 0.f);
    bool row_ok = per_pixel_row || (srow >= 0 && srow < a.abcg_rows);
    if (!per_pixel_row && row_ok) prow = reinterpret_cast<const float4*>(a.abcg)[srow];
    // the xy-interpolated delta and eta of every z cell the forest spans,
    // once per skewer into LDS (the same fused multiply-adds, in the same
    // order, as the per-pixel sums below): a pixel then reads 4 LDS values
    // instead of 16 box values.  Cells k = kmin .. kmax, unwrapped.
    float* ctab = reinterpret_cast<float*>(sh + a.ctab_off);
    int64_t kmin = 0;
    bool use_tab = false;
    if (a.ctab_cap > 0 && !xy_bad && imin < imax) {
      kmin = (int64_t)floor(fma((double)imin, duz, uz0));
      const int64_t kmax = (int64_t)floor(fma((double)(imax - 1), duz, uz0)) + 1;
      const int64_t ncell = kmax - kmin + 1;
      use_tab = ncell <= a.ctab_cap;
      if (use_tab) {
        // two cells per pass: the 16 loads of both are issued together
        for (int k0 = tid; k0 < ncell; k0 += 2 * NT) {
          float dv[2][4], evv[2][4];
          bool ok[2];
#pragma unroll
          for (int h = 0; h < 2; ++h) {
            const int k = k0 + h * NT;
            int64_t jz = kmin + k;
            ok[h] = k < ncell;
            if (b.periodic) jz = wrap(jz, nz);
            else ok[h] = ok[h] && jz >= 0 && jz < nz;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
              dv[h][q] = ok[h] ? fd[col[q] + jz] : 0.f;
              evv[h][q] = (ok[h] && fe) ? fe[col[q] + jz] : 0.f;
            }
          }
#pragma unroll
          for (int h = 0; h < 2; ++h) {
            const int k = k0 + h * NT;
            float av = 0.f, ev = 0.f;
            if (ok[h]) {
#pragma unroll
              for (int q = 0; q < 4; ++q) {
                av = fmaf(wxy[q], dv[h][q], av);
                if (fe) ev = fmaf(wxy[q], evv[h][q], ev);
              }
            }
            if (k < ncell) {
              ctab[k] = av;
              ctab[a.ctab_cap + k] = ev;
            }
          }
        }
        __syncthreads();
      }
    }
    int flags = 0;
    if (use_tab && !per_pixel_row && a.scale == nullptr) {
      // the common case (one parameter row per skewer, the column table, no
      // per-pixel scale): the loop below with its uniform branches hoisted
      // and the cell index in 32 bits (a z cell index is below nz); the same
      // floating-point operations, so the same bits
      const int kmin32 = (int)kmin;
      const bool periodic = b.periodic;
      const float* __restrict__ cd = ctab;
      const float* __restrict__ ce = ctab + a.ctab_cap;
      for (int i = tid; i < n; i += NT) {
        const float ds = dsf[i];
        float dl = -1000000.0f, eta = 0.0f;
        if (i >= imin && i < imax) {
          const double uz = fma((double)i, duz, uz0);
          const double fz0 = floor(uz);
          const float fz = (float)(uz - fz0);
          const int iz = (int)fz0;
          if (!periodic && (iz < 0 || iz + 1 >= nz)) {
            dl = __int_as_float(0x7fc00000);
            eta = dl;
            flags |= 1;
          } else {
            const int k0 = iz - kmin32;
            const float a0 = cd[k0], a1 = cd[k0 + 1];
            const float e0 = ce[k0], e1 = ce[k0 + 1];
            dl = fmaf(fz, a1 - a0, a0);
            eta = fmaf(fz, e1 - e0, e0);
          }
        }
        if (a.delta_l) a.delta_l[off + i] = dl;
        if (a.eta_par) a.eta_par[off + i] = eta;
        if (a.delta_s) a.delta_s[off + i] = ds;
        if (!row_ok) {
          a.flux[off + i] = __int_as_float(0x7fc00000);
          flags |= 2;
        } else {
          const float gfield = dl + ds + prow.z * eta;
          a.flux[off + i] = expf(-prow.x * expf(prow.y * prow.w * gfield));
        }
      }
      if (flags && a.status) atomicOr(&a.status[s], flags);
      continue;
    }
    for (int i = tid; i < n; i += NT) {
      float ds = dsf[i];
      if (a.scale) ds *= a.scale[off + i];
      float dl = -1000000.0f, eta = 0.0f;
      if (i >= imin && i < imax) {
        const double uz = fma((double)i, duz, uz0);
        const double fz0 = floor(uz);
        const float fz = (float)(uz - fz0);
        int64_t jz0 = (int64_t)fz0, jz1 = jz0 + 1;
        bool zout = false;
        if (b.periodic) {
          jz0 = wrap(jz0, nz);
          jz1 = wrap(jz1, nz);
        } else if (jz0 < 0 || jz1 >= nz) {
          zout = true;
        }
        if (xy_bad || zout) {
          dl = __int_as_float(0x7fc00000);
          eta = dl;
          flags |= (outside || zout ? 1 : 0) | (off_slab ? 4 : 0);
        } else if (use_tab) {
          const int k0 = (int)((int64_t)fz0 - kmin);
          const float a0 = ctab[k0], a1 = ctab[k0 + 1];
          const float e0 = ctab[a.ctab_cap + k0], e1 = ctab[a.ctab_cap + k0 + 1];
          dl = fmaf(fz, a1 - a0, a0);
          eta = fmaf(fz, e1 - e0, e0);
        } else {
          float a0 = 0.f, a1 = 0.f, e0 = 0.f, e1 = 0.f;
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            a0 = fmaf(wxy[q], fd[col[q] + jz0], a0);
            a1 = fmaf(wxy[q], fd[col[q] + jz1], a1);
            if (fe) {
              e0 = fmaf(wxy[q], fe[col[q] + jz0], e0);
              e1 = fmaf(wxy[q], fe[col[q] + jz1], e1);
            }
          }
          dl = fmaf(fz, a1 - a0, a0);
          eta = fmaf(fz, e1 - e0, e0);
        }
      }
      if (a.delta_l) a.delta_l[off + i] = dl;
      if (a.eta_par) a.eta_par[off + i] = eta;
      if (a.delta_s) a.delta_s[off + i] = ds;
      float4 pr = prow;
      bool rok = row_ok;
      if (per_pixel_row) {
        const int64_t row = g0 + i;
        rok = row >= 0 && row < a.abcg_rows;
        if (rok) pr = reinterpret_cast<const float4*>(a.abcg)[row];
      }
      if (!rok) {
        a.flux[off + i] = __int_as_float(0x7fc00000);
        flags |= 2;
      } else {
        // F = exp(-a exp(b G (delta + c eta)))
        const float gfield = dl + ds + pr.z * eta;
        a.flux[off + i] = expf(-pr.x * expf(pr.y * pr.w * gfield));
      }
    }
    if (flags && a.status) atomicOr(&a.status[s], flags);
  }
}

__global__ void k_zero_ints(int32_t* p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

__global__ void k_check_sets(const int32_t* __restrict__ fset, int n, int32_t n_sets, int32_t* bad) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && (fset[i] < 0 || fset[i] >= n_sets)) atomicOr(bad, 1);
}

constexpr int mp_T(int m) {
  const int P = m / elems_per_thread(m);
  return P >= 256 ? 1 : 256 / P;
}

template <int M>
hipError_t launch_mp(const smk_skewers& sk, const int32_t* list, const int32_t* count, const float* filt,
                     const int32_t* fset, int64_t set_stride, const float* scale, uint64_t seed,
                     const uint64_t* seed_dev, const float* noise, const int64_t* noise_off, const float2* tw,
                     const float2* twr, float* delta_s, hipStream_t s) {
  constexpr int T = mp_T(M);
  const int nblk = (sk.n + T - 1) / T;
  if (nblk == 0) return hipSuccess;
  hipLaunchKernelGGL((k_missing_power<M, T>), dim3(nblk), dim3(T * (M / elems_per_thread(M))), 0, s, sk, list,
                     count, filt, fset, set_stride, scale, seed, seed_dev, noise, noise_off, tw, twr, delta_s);
  return hipGetLastError();
}

std::vector<float2> twiddles_f(int n, double scale) {
  std::vector<float2> t(n);
  for (int m = 0; m < n; ++m) {
    const double ang = -2.0 * M_PI * (double)m / ((double)n * scale);
    t[m] = make_float2((float)cos(ang), (float)sin(ang));
  }
  return t;
}

}  // namespace

struct smk_forest_plan {
  float2* tw[N_CLASS];   // W_M^m, M = 128 << c
  float2* twr[N_CLASS];  // W_{2M}^k
  const uint64_t* seed_dev;  // graph mode: the seed in device memory, no host read-back
  int step_variant;          // smk_forest_step instance for batches up to class 3
};

extern "C" int smk_forest_plan_create(smk_forest_plan** plan) {
  SMK_REQUIRE(plan != nullptr, SMK_ERR_ARG, "plan pointer is NULL");
  smk_forest_plan* p = (smk_forest_plan*)calloc(1, sizeof(smk_forest_plan));
  SMK_REQUIRE(p != nullptr, SMK_ERR_OOM, "host allocation failed");
  for (int c = 0; c < N_CLASS; ++c) {
    const int M = 128 << c;
    auto a = twiddles_f(M, 1.0), b = twiddles_f(M, 2.0);
    hipError_t e = hipMalloc((void**)&p->tw[c], M * sizeof(float2));
    if (e == hipSuccess) e = hipMemcpy(p->tw[c], a.data(), M * sizeof(float2), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMalloc((void**)&p->twr[c], M * sizeof(float2));
    if (e == hipSuccess) e = hipMemcpy(p->twr[c], b.data(), M * sizeof(float2), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
      smk_forest_plan_destroy(p);
      return set_error(SMK_ERR_HIP, "forest plan tables: %s", hipGetErrorString(e));
    }
  }
  *plan = p;
  return SMK_OK;
}

extern "C" int smk_forest_plan_set_graph_mode(smk_forest_plan* p, const uint64_t* seed_dev) {
  SMK_REQUIRE(p != nullptr, SMK_ERR_ARG, "plan is NULL");
  p->seed_dev = seed_dev;
  return SMK_OK;
}

extern "C" int smk_forest_plan_set_step_variant(smk_forest_plan* p, int variant) {
  SMK_REQUIRE(p != nullptr, SMK_ERR_ARG, "plan is NULL");
  SMK_REQUIRE(variant >= 0 && variant < SMK_FOREST_STEP_NVAR, SMK_ERR_ARG, "step variant %d not in [0, %d)",
              variant, SMK_FOREST_STEP_NVAR);
  p->step_variant = variant;
  return SMK_OK;
}

extern "C" int smk_forest_plan_destroy(smk_forest_plan* p) {
  if (p == nullptr) return SMK_OK;
  for (int c = 0; c < N_CLASS; ++c) {
    if (p->tw[c]) (void)hipFree(p->tw[c]);
    if (p->twr[c]) (void)hipFree(p->twr[c]);
  }
  free(p);
  return SMK_OK;
}

// stencil radius of a Gaussian interp code, -1 if it is not one
static int gauss_radius(int interp) {
  if (interp == SMK_INTERP_GAUSS343) return 3;
  const int r = interp - SMK_INTERP_GAUSS(0);
  return r >= 0 && r <= SMK_GAUSS_DMAX ? r : -1;
}

template <int NF>
static void launch_gather(int r, dim3 grid, hipStream_t st, const GatherArgs& a) {
  switch (r) {
    case -1: hipLaunchKernelGGL((k_gather<NF, -1>), grid, dim3(256), 0, st, a); break;
    case 0: hipLaunchKernelGGL((k_gather<NF, 0>), grid, dim3(256), 0, st, a); break;
    case 1: hipLaunchKernelGGL((k_gather<NF, 1>), grid, dim3(256), 0, st, a); break;
    case 2: hipLaunchKernelGGL((k_gather<NF, 2>), grid, dim3(256), 0, st, a); break;
    case 4: hipLaunchKernelGGL((k_gather<NF, 4>), grid, dim3(256), 0, st, a); break;
    case 5: hipLaunchKernelGGL((k_gather<NF, 5>), grid, dim3(256), 0, st, a); break;
    case 6: hipLaunchKernelGGL((k_gather<NF, 6>), grid, dim3(256), 0, st, a); break;
    default: hipLaunchKernelGGL((k_gather<NF, 3>), grid, dim3(256), 0, st, a); break;
  }
}

extern "C" size_t smk_forest_workspace_bytes(int n_skewers) {
  return (size_t)(N_CLASS * (size_t)(n_skewers > 0 ? n_skewers : 0) + 16) * sizeof(int32_t);
}

extern "C" int smk_forest_gather(const smk_skewers* sk, const smk_box_view* box, int interp, const float* abcg,
                                 int64_t abcg_rows, const float* delta_s, float* delta_l, float* eta_par,
                                 float* v_par, float* flux, int32_t* cells, int32_t* status, void* stream) {
  SMK_REQUIRE(sk != nullptr && box != nullptr, SMK_ERR_ARG, "NULL skewers/box");
  SMK_REQUIRE(box->delta != nullptr, SMK_ERR_ARG, "box delta is NULL");
  SMK_REQUIRE(box->nx > 0 && box->ny > 0 && box->nz > 0 && box->cell > 0, SMK_ERR_ARG, "bad box geometry");
  SMK_REQUIRE(interp == SMK_INTERP_TRILINEAR || gauss_radius(interp) >= 0, SMK_ERR_ARG, "bad interp %d", interp);
  SMK_REQUIRE(box->nxs >= 0 && box->x0 >= 0 && (box->nxs == 0 || box->x0 + box->nxs <= box->nx), SMK_ERR_ARG,
              "bad resident x slab [%d, %d + %d) of a %d-plane box", box->x0, box->x0, box->nxs, box->nx);
  SMK_REQUIRE(sk->geometry == SMK_GEOM_PLANE || sk->geometry == SMK_GEOM_RADIAL, SMK_ERR_ARG, "bad geometry");
  SMK_REQUIRE(flux == nullptr || abcg != nullptr, SMK_ERR_ARG, "flux needs FGPA parameters");
  SMK_REQUIRE(sk->geometry != SMK_GEOM_RADIAL ||
                  (box->eta[0] == nullptr) == (box->eta[2] == nullptr),
              SMK_ERR_ARG, "radial eta_par needs all six eta boxes");
  SMK_REQUIRE(v_par == nullptr || (sk->geometry == SMK_GEOM_RADIAL && box->eta[0] != nullptr &&
                                   box->vel[0] != nullptr && box->vel[1] != nullptr && box->vel[2] != nullptr),
              SMK_ERR_ARG, "v_par needs radial skewers, the six eta boxes and the three velocity boxes");
  if (sk->n <= 0) return SMK_OK;
  GatherArgs a{};
  a.sk = *sk;
  a.box = *box;
  a.interp = interp;
  a.abcg = abcg;
  a.abcg_rows = abcg_rows;
  a.delta_s = delta_s;
  a.delta_l = delta_l;
  a.eta_par = eta_par;
  a.v_par = v_par;
  a.flux = flux;
  a.cells = cells;
  a.status = status;
  a.npix_total = sk->npix_total;
  a.hx = box->cell * box->nx / 2;
  a.hy = box->cell * box->ny / 2;
  a.hz = box->cell * box->nz / 2;
  if (a.npix_total <= 0) return SMK_OK;
  if (interp == SMK_INTERP_TRILINEAR && sk->geometry == SMK_GEOM_PLANE) {
    constexpr int PPT = 1;  // measured: 1 pixel per thread 2.9 ms, 2 4.2 ms, 4 6.5 ms (occupancy)
    const uint64_t nblk = (uint64_t)(a.npix_total + 256 * PPT - 1) / (256 * PPT);
    hipLaunchKernelGGL(k_gather_plane_tri<PPT>, dim3((uint32_t)nblk), dim3(256), 0, (hipStream_t)stream, a);
  } else {
    const uint64_t nblk = (uint64_t)(a.npix_total + 255) / 256;
    const bool radial = sk->geometry == SMK_GEOM_RADIAL;
    const int nf = box->eta[2] == nullptr ? 1 : (radial ? (v_par ? 10 : 7) : 2);
    const int r = interp == SMK_INTERP_TRILINEAR ? -1 : gauss_radius(interp);
    const dim3 grid((uint32_t)nblk);
    if (nf == 1) launch_gather<1>(r, grid, (hipStream_t)stream, a);
    else if (nf == 2) launch_gather<2>(r, grid, (hipStream_t)stream, a);
    else if (nf == 10) launch_gather<10>(r, grid, (hipStream_t)stream, a);
    else launch_gather<7>(r, grid, (hipStream_t)stream, a);
  }
  SMK_LAUNCH_CHECK();
  return SMK_OK;
}

extern "C" int smk_forest_missing_power_sets(smk_forest_plan* plan, const smk_skewers* sk, const float* filt,
                                             const int64_t* filt_off, int64_t set_stride, int32_t n_sets,
                                             const int32_t* filt_set, const float* scale, uint64_t seed,
                                             const float* noise, const int64_t* noise_off, void* workspace,
                                             float* delta_s, void* stream) {
  SMK_REQUIRE(plan != nullptr && sk != nullptr && filt != nullptr && filt_off != nullptr && delta_s != nullptr,
              SMK_ERR_ARG, "NULL argument");
  SMK_REQUIRE(workspace != nullptr, SMK_ERR_ARG, "workspace is NULL");
  SMK_REQUIRE(noise == nullptr || noise_off != nullptr, SMK_ERR_ARG, "fed noise needs noise_off");
  SMK_REQUIRE(filt_set == nullptr || (n_sets >= 1 && set_stride >= 0), SMK_ERR_ARG,
              "filter sets need n_sets >= 1 and set_stride >= 0");
  if (sk->n <= 0) return SMK_OK;
  hipStream_t s = (hipStream_t)stream;
  int32_t* counts = (int32_t*)workspace;           // [N_CLASS], [N_CLASS] = bad flags
  int32_t* lists = counts + 16;
  hipLaunchKernelGGL(k_zero_ints, dim3(1), dim3(64), 0, s, counts, 16);
  SMK_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_classify, dim3((sk->n + 255) / 256), dim3(256), 0, s, *sk, filt_set, n_sets, lists, counts,
                     counts + N_CLASS);
  SMK_LAUNCH_CHECK();
  const int64_t* foff = filt_off;
  if (plan->seed_dev == nullptr) {
    // (in graph mode the batch was checked by an eager call; a skewer that
    // fails the checks is left out of every class list and gets no delta_s)
    int32_t bad = 0;
    SMK_HIP_CHECK(hipMemcpyAsync(&bad, counts + N_CLASS, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    SMK_HIP_CHECK(hipStreamSynchronize(s));
    SMK_REQUIRE((bad & 2) == 0, SMK_ERR_RANGE, "a filter set index is outside [0, %d)", (int)n_sets);
    SMK_REQUIRE(bad == 0, SMK_ERR_RANGE, "a skewer needs an FFT longer than %d", 256 << (N_CLASS - 1));
  }
  const int64_t stride = filt_set ? set_stride : 0;
#define SMK_MP(C, M)                                                                                            \
  SMK_HIP_CHECK(launch_mp<M>(*sk, lists + (int64_t)C * sk->n, counts + C, filt + foff[C], filt_set, stride,     \
                             scale, seed, plan->seed_dev, noise, noise_off, plan->tw[C], plan->twr[C], delta_s, s))
  SMK_MP(0, 128);
  SMK_MP(1, 256);
  SMK_MP(2, 512);
  SMK_MP(3, 1024);
  SMK_MP(4, 2048);
  SMK_MP(5, 4096);
#undef SMK_MP
  return SMK_OK;
}

extern "C" int smk_forest_missing_power(smk_forest_plan* plan, const smk_skewers* sk, const float* filt,
                                        const int64_t* filt_off, uint64_t seed, const float* noise,
                                        const int64_t* noise_off, void* workspace, float* delta_s, void* stream) {
  return smk_forest_missing_power_sets(plan, sk, filt, filt_off, 0, 1, nullptr, nullptr, seed, noise, noise_off,
                                       workspace, delta_s, stream);
}

extern "C" size_t smk_forest_step_workspace_bytes(void) { return 256; }

extern "C" int smk_forest_step(smk_forest_plan* plan, const smk_skewers* sk, const smk_box_view* box,
                               const float* filt, const int64_t* filt_off, int64_t set_stride, int32_t n_sets,
                               const int32_t* filt_set, const float* scale, uint64_t seed, const float* abcg,
                               int64_t abcg_rows, int32_t max_npix, float* flux, float* delta_l, float* eta_par,
                               float* delta_s, int32_t* status, void* workspace, void* stream) {
  SMK_REQUIRE(plan != nullptr && sk != nullptr && box != nullptr, SMK_ERR_ARG, "NULL plan/skewers/box");
  SMK_REQUIRE(filt != nullptr && filt_off != nullptr, SMK_ERR_ARG, "NULL filter tables");
  SMK_REQUIRE(abcg != nullptr && flux != nullptr, SMK_ERR_ARG, "the step needs FGPA parameters and a flux output");
  SMK_REQUIRE(workspace != nullptr && ((uintptr_t)workspace & 3) == 0, SMK_ERR_ARG, "bad workspace");
  SMK_REQUIRE(box->delta != nullptr, SMK_ERR_ARG, "box delta is NULL");
  SMK_REQUIRE(box->nx > 0 && box->ny > 0 && box->nz > 0 && box->cell > 0, SMK_ERR_ARG, "bad box geometry");
  SMK_REQUIRE(box->nxs >= 0 && box->x0 >= 0 && (box->nxs == 0 || box->x0 + box->nxs <= box->nx), SMK_ERR_ARG,
              "bad resident x slab [%d, %d + %d) of a %d-plane box", box->x0, box->x0, box->nxs, box->nx);
  SMK_REQUIRE(sk->geometry == SMK_GEOM_PLANE, SMK_ERR_ARG, "the fused step takes plane-parallel skewers");
  SMK_REQUIRE(filt_set == nullptr || (n_sets >= 1 && set_stride >= 0), SMK_ERR_ARG,
              "filter sets need n_sets >= 1 and set_stride >= 0");
  SMK_REQUIRE(max_npix >= 0 && max_npix + 50 <= (256 << (N_CLASS - 1)), SMK_ERR_RANGE,
              "max_npix %d needs an FFT longer than %d", (int)max_npix, 256 << (N_CLASS - 1));
  if (sk->n <= 0) return SMK_OK;
  hipStream_t s = (hipStream_t)stream;
  StepArgs a{};
  a.sk = *sk;
  a.box = *box;
  a.filt = filt;
  for (int c = 0; c < N_CLASS; ++c) {
    a.foff[c] = filt_off[c];
    a.tw[c] = plan->tw[c];
    a.twr[c] = plan->twr[c];
  }
  a.fset = filt_set;
  a.set_stride = filt_set ? set_stride : 0;
  a.n_sets = n_sets;
  a.scale = scale;
  a.abcg = abcg;
  a.abcg_rows = abcg_rows;
  a.seed = seed;
  a.seed_dev = plan->seed_dev;
  a.flux = flux;
  a.delta_l = delta_l;
  a.eta_par = eta_par;
  a.delta_s = delta_s;
  a.status = status;
  a.queue = (int32_t*)workspace;
  int c = 0;
  for (int nzc = 256; nzc < max_npix + 50; nzc <<= 1) ++c;
  a.max_class = c;
  a.hx = box->cell * box->nx / 2;
  a.hy = box->cell * box->ny / 2;
  a.hz = box->cell * box->nz / 2;
  // a filter-set index out of range reads outside the tables: checked here
  // in normal mode (one host sync), as smk_forest_missing_power_sets does
  if (filt_set != nullptr && plan->seed_dev == nullptr) {
    int32_t* bad = a.queue + 1;
    hipLaunchKernelGGL(k_zero_ints, dim3(1), dim3(64), 0, s, a.queue, 2);
    SMK_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_check_sets, dim3((sk->n + 255) / 256), dim3(256), 0, s, filt_set, sk->n, n_sets, bad);
    SMK_LAUNCH_CHECK();
    int32_t h = 0;
    SMK_HIP_CHECK(hipMemcpyAsync(&h, bad, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    SMK_HIP_CHECK(hipStreamSynchronize(s));
    SMK_REQUIRE(h == 0, SMK_ERR_RANGE, "a filter set index is outside [0, %d)", (int)n_sets);
  }
  // kernel instance for classes up to 3 (nz <= 2048, n <= 1998): one wave
  // per skewer with at least 4 waves per SIMD (variant 0: some registers
  // spill) or 3 (variant 1: none spill), two waves per skewer with at least
  // 4 (variant 2) or 5 (variant 3: some spill); four waves per skewer with
  // all six classes.  Skewers of a class above max_npix's are flagged
  // (a.max_class)
  const bool small = c <= 3;
  const int var = plan->step_variant;
  constexpr int NT_L = 256;
  const int nt = small ? (var >= 2 ? 128 : 64) : NT_L;
  const int mp_float2 = small ? (nt == 128 ? step_lds_max(3, 128) : step_lds_max(3, 64))
                              : step_lds_max(N_CLASS - 1, NT_L);
  // column table: the z cells of the longest forest (n pixels span at most
  // n dr / D + 2 cells; 2 more for the rounding of the positions), up to
  // 4096 cells; longer spans read the box per pixel
  const double span = (double)max_npix * sk->dr / box->cell + 4.0;
  a.ctab_cap = (sk->dr > 0 && span <= 4096.0) ? (int32_t)span : 0;
  a.ctab_off = mp_float2;
  const size_t lds = ((size_t)mp_float2 + (size_t)a.ctab_cap) * sizeof(float2);
  int dev = 0, ncu = 0, per_cu = 0;
  SMK_HIP_CHECK(hipGetDevice(&dev));
  SMK_HIP_CHECK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
  void (*const small_kern[SMK_FOREST_STEP_NVAR])(StepArgs) = {k_forest_step<3, 64, 4>, k_forest_step<3, 64, 3>,
                                                                k_forest_step<3, 128, 4>, k_forest_step<3, 128, 5>};
  void (*kern)(StepArgs) = small ? small_kern[var] : k_forest_step<N_CLASS - 1, NT_L, 1>;
  SMK_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, nt, lds));
  const int64_t want = (int64_t)ncu * (per_cu > 0 ? per_cu : 1);
  const int nblk = (int)(want < sk->n ? want : sk->n);
  hipLaunchKernelGGL(k_zero_ints, dim3(1), dim3(64), 0, s, a.queue, 1);
  SMK_LAUNCH_CHECK();
  hipLaunchKernelGGL(kern, dim3(nblk), dim3(nt), lds, s, a);
  SMK_LAUNCH_CHECK();
  return SMK_OK;
}
