// FGPA fit kernels on gfx950: flux, mean flux, binned P1D and their adjoint.
//
// Reference: py/SaclayMocks/fit_az.py:194-264 (comp_spectra, compute_p1d,
// chi2), powerspectrum.py:204-238 (P1D_1spectrum, ComputeP1D) and
// util.py:29-37,46-75 (regroup, MakeProfileHisto).
//
// One block per skewer (grid-stride).  The P1D of a forest of arbitrary
// length N is a direct DFT evaluated in FP64 by Horner's rule (N
// multiply-adds per mode, one mode per thread), so any N is exact to FP64
// rounding, as numpy's rfft is.  All reductions are FP64 and deterministic
// (fixed thread-to-data assignment, fixed tree order).
#include <math.h>

#include "smk_internal.h"

namespace {

constexpr int FIT_THREADS = 256;
constexpr int FIT_NMAX = 8192;  // longest (regrouped) forest handled by the P1D kernels

// dynamic LDS of the P1D kernels for forests of at most nmax (even) points
size_t p1d_lds_bytes(int nmax) { return (size_t)nmax * 8 + (size_t)(nmax / 2 + 1) * 12; }
size_t adj_lds_bytes(int nmax) { return (size_t)nmax * 8 + (size_t)(nmax / 2 + 1) * 16; }

__device__ __forceinline__ double block_reduce_sum(double v, double* red) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red[wv] = v;
  __syncthreads();
  double s = 0.0;
  for (int i = 0; i < FIT_THREADS / 64; ++i) s += red[i];
  __syncthreads();
  return s;  // valid on every thread
}

__device__ __forceinline__ float4 prow(const float* abcg, int64_t rows, int64_t r) {
  if (r < 0 || r >= rows) return make_float4(__int_as_float(0x7fc00000), 0.f, 0.f, 0.f);
  return reinterpret_cast<const float4*>(abcg)[r];
}

// F for every pixel; per-skewer FP64 sum of F over the forest [imin, imax)
__global__ __launch_bounds__(FIT_THREADS) void k_fgpa_flux(smk_skewers sk, const float* __restrict__ delta,
                                                           const float* __restrict__ eta,
                                                           const float* __restrict__ abcg, int64_t rows,
                                                           float* __restrict__ flux, double* __restrict__ fsum) {
  __shared__ double red[FIT_THREADS / 64];
  for (int s = blockIdx.x; s < sk.n; s += gridDim.x) {
    const int64_t o = sk.off[s];
    const int n = sk.npix[s], i0 = sk.imin[s], i1 = sk.imax[s];
    const float4 p = prow(abcg, rows, sk.pset[s]);
    double acc = 0.0;
    for (int i = threadIdx.x; i < n; i += FIT_THREADS) {
      const float g = delta[o + i] + p.z * (eta ? eta[o + i] : 0.0f);
      const float f = expf(-p.x * expf(p.y * p.w * g));
      flux[o + i] = f;
      if (i >= i0 && i < i1) acc += (double)f;
    }
    const double tot = block_reduce_sum(acc, red);
    if (threadIdx.x == 0) fsum[s] = tot;
  }
}

// sums over the skewers of each parameter set: out[set][c] = sum_{s in set} in[s][c]
__global__ __launch_bounds__(FIT_THREADS) void k_reduce_by_set(const int32_t* __restrict__ pset, int n,
                                                               const double* __restrict__ in, int width,
                                                               double* __restrict__ out) {
  __shared__ double red[FIT_THREADS / 64];
  const int set = blockIdx.y, col = blockIdx.x;
  double acc = 0.0;
  for (int s = threadIdx.x; s < n; s += FIT_THREADS)
    if (pset[s] == set) acc += in[(int64_t)s * width + col];
  const double tot = block_reduce_sum(acc, red);
  if (threadIdx.x == 0) out[(int64_t)set * width + col] = tot;
}

// numpy.fft.rfftfreq(n) * 2 * k_ny: (j * (1/n)) * 2 * (pi / pixel)
__device__ __forceinline__ double kfreq(int j, int n, double k_ny) {
  return __dmul_rn(__dmul_rn(__dmul_rn((double)j, 1.0 / (double)n), 2.0), k_ny);
}

// scipy binned_statistic bin: edges[b] <= k < edges[b+1], last bin closed
__device__ __forceinline__ int find_bin(const double* __restrict__ edges, int nb, double k) {
  if (!(k >= edges[0]) || k > edges[nb]) return -1;
  if (k == edges[nb]) return nb - 1;
  int lo = 0, hi = nb - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (edges[mid] <= k) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// delta_F of the (regrouped) forest into LDS; returns the regrouped length
__device__ __forceinline__ int load_delta_f(const float* __restrict__ flux, int64_t o, int i0, int i1, int m,
                                           double inv_fmean, double* __restrict__ x) {
  const int nf = i1 - i0;
  const int nq = nf / m;
  for (int q = threadIdx.x; q < nq; q += FIT_THREADS) {
    double acc = 0.0;
    for (int r = 0; r < m; ++r) acc += (double)flux[o + i0 + q * m + r];
    x[q] = (acc / m) * inv_fmean - 1.0;
  }
  return nq;
}

// X_j = sum_n x_n exp(-2 pi i j n / N) by Horner in FP64
__device__ __forceinline__ double2 dft_mode(const double* __restrict__ x, int N, int j) {
  double s, c;
  sincospi(-2.0 * (double)j / (double)N, &s, &c);
  double re = 0.0, im = 0.0;
  for (int n = N - 1; n >= 0; --n) {
    const double r2 = re * c - im * s + x[n];
    im = re * s + im * c;
    re = r2;
  }
  return make_double2(re, im);
}

struct P1dArgs {
  smk_skewers sk;
  const float* flux;
  const double* fmean;   // per set
  double pixel;          // spectrum pixel (before regrouping)
  int regroup;           // m >= 1
  const double* edges;   // nb + 1
  int nb;
  double* partial;       // [n][nb][4]: count, sum k, sum P, sum P^2
  int32_t* status;       // per skewer: 1 if the forest is longer than nmax
  int nmax;
};

__global__ __launch_bounds__(FIT_THREADS) void k_p1d(P1dArgs a) {
  extern __shared__ double dyn[];
  double* x = dyn;
  double* pk = dyn + a.nmax;
  int* bins = reinterpret_cast<int*>(pk + a.nmax / 2 + 1);
  const smk_skewers& sk = a.sk;
  for (int s = blockIdx.x; s < sk.n; s += gridDim.x) {
    double* part = a.partial + (int64_t)s * a.nb * 4;
    const int i0 = sk.imin[s], i1 = sk.imax[s];
    const int nq = (i1 - i0) / a.regroup;
    __syncthreads();
    if (nq > a.nmax) {
      for (int b = threadIdx.x; b < 4 * a.nb; b += FIT_THREADS) part[b] = 0.0;
      if (threadIdx.x == 0 && a.status) a.status[s] = 1;
      continue;
    }
    const double inv = 1.0 / a.fmean[sk.pset[s]];
    load_delta_f(a.flux, sk.off[s], i0, i1, a.regroup, inv, x);
    __syncthreads();
    const double px = a.pixel * a.regroup;
    const double k_ny = M_PI / px;
    const int nj = nq > 0 ? nq / 2 + 1 : 0;
    for (int j = threadIdx.x; j < nj; j += FIT_THREADS) {
      const double2 X = dft_mode(x, nq, j);
      pk[j] = (X.x * X.x + X.y * X.y) * px / nq;
      bins[j] = find_bin(a.edges, a.nb, kfreq(j, nq, k_ny));
    }
    __syncthreads();
    for (int b = threadIdx.x; b < a.nb; b += FIT_THREADS) {
      double c = 0.0, sk_ = 0.0, sp = 0.0, sp2 = 0.0;
      for (int j = 0; j < nj; ++j) {
        if (bins[j] == b) {
          c += 1.0;
          sk_ += kfreq(j, nq, k_ny);
          sp += pk[j];
          sp2 += pk[j] * pk[j];
        }
      }
      part[4 * b + 0] = c;
      part[4 * b + 1] = sk_;
      part[4 * b + 2] = sp;
      part[4 * b + 3] = sp2;
    }
  }
}

struct AdjArgs {
  smk_skewers sk;
  const float* delta;
  const float* eta;
  const float* abcg;
  int64_t rows;
  const float* flux;
  const double* fmean;
  double pixel;
  int regroup;
  const double* edges;
  int nb;
  const double* gp;      // [sets][nb]: dL/dP for one P_j sample of bin b
  double* out;           // [n][7]: sum A F', sum A dF/d(a,b,c) / m, sum dF/d(a,b,c)
  int nmax;
};

// Adjoint of the binned P1D through delta_F = F/<F> - 1 (at fixed <F>) and
// of the FGPA: A_q = dL/d delta_F,q = sum_j gP_j (2 px / N) Re(conj(X_j) W^{jq}).
__global__ __launch_bounds__(FIT_THREADS) void k_p1d_adjoint(AdjArgs a) {
  extern __shared__ double dyn[];
  double* x = dyn;
  double2* cj = reinterpret_cast<double2*>(dyn + a.nmax);
  __shared__ double red[FIT_THREADS / 64];
  const smk_skewers& sk = a.sk;
  for (int s = blockIdx.x; s < sk.n; s += gridDim.x) {
    const int i0 = sk.imin[s], i1 = sk.imax[s];
    const int m = a.regroup;
    const int nq = (i1 - i0) / m;
    const int64_t o = sk.off[s];
    double* res = a.out + (int64_t)s * 7;
    __syncthreads();
    if (nq > a.nmax) {
      if (threadIdx.x < 7) res[threadIdx.x] = 0.0;
      continue;
    }
    const int set = sk.pset[s];
    const double fm = a.fmean[set];
    load_delta_f(a.flux, o, i0, i1, m, 1.0 / fm, x);
    __syncthreads();
    const double px = a.pixel * m;
    const double k_ny = M_PI / px;
    const int nj = nq > 0 ? nq / 2 + 1 : 0;
    for (int j = threadIdx.x; j < nj; j += FIT_THREADS) {
      const int b = find_bin(a.edges, a.nb, kfreq(j, nq, k_ny));
      double2 c = make_double2(0.0, 0.0);
      if (b >= 0) {
        const double g = a.gp[(int64_t)set * a.nb + b] * 2.0 * px / nq;
        if (g != 0.0) {
          const double2 X = dft_mode(x, nq, j);
          c = make_double2(g * X.x, -g * X.y);   // g conj(X_j)
        }
      }
      cj[j] = c;
    }
    __syncthreads();
    const float4 p = prow(a.abcg, a.rows, set);
    const double pa = p.x, pb = p.y, pc = p.z, pg = p.w;
    double saf = 0.0, t1a = 0.0, t1b = 0.0, t1c = 0.0, t2a = 0.0, t2b = 0.0, t2c = 0.0;
    // pass over the forest pixels: A of the pixel's regroup cell, dF/dtheta
    for (int i = threadIdx.x; i < i1 - i0; i += FIT_THREADS) {
      const int q = i / m;
      const double d = a.delta[o + i0 + i];
      const double e = a.eta ? (double)a.eta[o + i0 + i] : 0.0;
      const double gf = d + pc * e;
      const double T = exp(pb * pg * gf);
      const double F = exp(-pa * T);
      const double dfa = -T * F;
      const double dfb = -pa * T * F * pg * gf;
      const double dfc = -pa * T * F * pb * pg * e;
      t2a += dfa; t2b += dfb; t2c += dfc;
      if (q < nq) {
        // A_q = Re sum_j cj[j] W^{jq}, Horner in w = W^q
        double sn, cs;
        sincospi(-2.0 * (double)q / (double)nq, &sn, &cs);
        double re = 0.0, im = 0.0;
        for (int j = nj - 1; j >= 0; --j) {
          const double r2 = re * cs - im * sn + cj[j].x;
          im = re * sn + im * cs + cj[j].y;
          re = r2;
        }
        const double A = re;
        if (i % m == 0) saf += A * (x[q] + 1.0) * fm;   // A_q F'_q, once per regroup cell
        t1a += A * dfa / m;
        t1b += A * dfb / m;
        t1c += A * dfc / m;
      }
    }
    const double v[7] = {saf, t1a, t1b, t1c, t2a, t2b, t2c};
    for (int k = 0; k < 7; ++k) {
      const double tot = block_reduce_sum(v[k], red);
      if (threadIdx.x == 0) res[k] = tot;
    }
  }
}

int grid_for(int n) { return n < 65536 ? (n > 0 ? n : 1) : 65536; }

}  // namespace

extern "C" int smk_fgpa_flux(const smk_skewers* sk, const float* delta, const float* eta, const float* abcg,
                             int64_t abcg_rows, float* flux, double* fsum, void* stream) {
  SMK_REQUIRE(sk != nullptr && delta != nullptr && abcg != nullptr && flux != nullptr && fsum != nullptr,
              SMK_ERR_ARG, "NULL argument");
  SMK_REQUIRE(sk->pset != nullptr, SMK_ERR_ARG, "the fit kernels need skewers.pset");
  if (sk->n <= 0) return SMK_OK;
  hipLaunchKernelGGL(k_fgpa_flux, dim3(grid_for(sk->n)), dim3(FIT_THREADS), 0, (hipStream_t)stream, *sk, delta,
                     eta, abcg, abcg_rows, flux, fsum);
  SMK_LAUNCH_CHECK();
  return SMK_OK;
}

extern "C" int smk_reduce_by_set(const smk_skewers* sk, const double* per_skewer, int width, int n_sets,
                                 double* out, void* stream) {
  SMK_REQUIRE(sk != nullptr && per_skewer != nullptr && out != nullptr && sk->pset != nullptr, SMK_ERR_ARG,
              "NULL argument");
  SMK_REQUIRE(width > 0 && n_sets > 0 && width <= 65535 && n_sets <= 65535, SMK_ERR_ARG, "bad width/n_sets");
  hipLaunchKernelGGL(k_reduce_by_set, dim3(width, n_sets), dim3(FIT_THREADS), 0, (hipStream_t)stream, sk->pset,
                     sk->n, per_skewer, width, out);
  SMK_LAUNCH_CHECK();
  return SMK_OK;
}

extern "C" int smk_p1d_binned(const smk_skewers* sk, const float* flux, const double* fmean, double pixel,
                              int regroup, const double* edges, int nbins, int max_forest, double* partial,
                              int32_t* status, void* stream) {
  SMK_REQUIRE(sk != nullptr && flux != nullptr && fmean != nullptr && edges != nullptr && partial != nullptr,
              SMK_ERR_ARG, "NULL argument");
  SMK_REQUIRE(sk->pset != nullptr, SMK_ERR_ARG, "the fit kernels need skewers.pset");
  SMK_REQUIRE(regroup >= 1 && nbins >= 1 && pixel > 0, SMK_ERR_ARG, "bad regroup/nbins/pixel");
  SMK_REQUIRE(max_forest >= 0 && max_forest / regroup <= FIT_NMAX, SMK_ERR_RANGE,
              "forest of %d points exceeds %d", max_forest / regroup, FIT_NMAX);
  if (sk->n <= 0) return SMK_OK;
  int nmax = max_forest / regroup;
  nmax += nmax & 1;
  if (nmax < 2) nmax = 2;
  P1dArgs a{*sk, flux, fmean, pixel, regroup, edges, nbins, partial, status, nmax};
  const size_t lds = p1d_lds_bytes(nmax);
  SMK_HIP_CHECK(hipFuncSetAttribute((const void*)k_p1d, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  hipLaunchKernelGGL(k_p1d, dim3(grid_for(sk->n)), dim3(FIT_THREADS), lds, (hipStream_t)stream, a);
  SMK_LAUNCH_CHECK();
  return SMK_OK;
}

extern "C" int smk_p1d_adjoint(const smk_skewers* sk, const float* delta, const float* eta, const float* abcg,
                               int64_t abcg_rows, const float* flux, const double* fmean, double pixel,
                               int regroup, const double* edges, int nbins, int max_forest, const double* gp,
                               double* out, void* stream) {
  SMK_REQUIRE(sk != nullptr && delta != nullptr && abcg != nullptr && flux != nullptr && fmean != nullptr &&
                  edges != nullptr && gp != nullptr && out != nullptr,
              SMK_ERR_ARG, "NULL argument");
  SMK_REQUIRE(sk->pset != nullptr, SMK_ERR_ARG, "the fit kernels need skewers.pset");
  SMK_REQUIRE(regroup >= 1 && nbins >= 1 && pixel > 0, SMK_ERR_ARG, "bad regroup/nbins/pixel");
  SMK_REQUIRE(max_forest >= 0 && max_forest / regroup <= FIT_NMAX, SMK_ERR_RANGE,
              "forest of %d points exceeds %d", max_forest / regroup, FIT_NMAX);
  if (sk->n <= 0) return SMK_OK;
  int nmax = max_forest / regroup;
  nmax += nmax & 1;
  if (nmax < 2) nmax = 2;
  AdjArgs a{*sk, delta, eta, abcg, abcg_rows, flux, fmean, pixel, regroup, edges, nbins, gp, out, nmax};
  const size_t lds = adj_lds_bytes(nmax);
  SMK_HIP_CHECK(
      hipFuncSetAttribute((const void*)k_p1d_adjoint, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  hipLaunchKernelGGL(k_p1d_adjoint, dim3(grid_for(sk->n)), dim3(FIT_THREADS), lds, (hipStream_t)stream, a);
  SMK_LAUNCH_CHECK();
  return SMK_OK;
}
