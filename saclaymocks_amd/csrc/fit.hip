// FGPA fit kernels on gfx950: flux, mean flux, binned P1D and their adjoint.
//
// Reference: py/SaclayMocks/fit_az.py:194-264 (comp_spectra, compute_p1d,
// chi2), powerspectrum.py:204-238 (P1D_1spectrum, ComputeP1D) and
// util.py:29-37,46-75 (regroup, MakeProfileHisto).
//
// One block per skewer (grid-stride).  The P1D of a forest of N <= 2048
// points is the DFT at that exact length by Bluestein's algorithm on FP64
// power-of-two FFTs (fft64.h, O(N log N)); longer forests take a direct DFT
// evaluated in FP64 by Horner's rule (N multiply-adds per mode, one mode per
// thread).  Either is exact to FP64 rounding at any N, as numpy's rfft is.
// All reductions are FP64 and deterministic (fixed thread-to-data
// assignment, fixed tree order).
#include <math.h>

#include <mutex>
#include <vector>

#include "fft64.h"
#include "fft_core.h"
#include "smk_internal.h"

using namespace smk;

namespace {

constexpr int FIT_THREADS = 512;
constexpr int FIT_NMAX = 8192;  // longest (regrouped) forest handled by the P1D kernels
constexpr int BLUE_NMAX = 2048; // longest forest of the Bluestein path (M <= 4096)
constexpr int BLUE_MMAX = 4096;

// dynamic LDS of the P1D kernels for forests of at most nmax (even) points:
// x, then the binning arrays (forward) or the adjoint's coefficients, with
// the Bluestein FFT line aliased over them
int blue_m(int nmax) {
  const int n = nmax < BLUE_NMAX ? nmax : BLUE_NMAX;
  int M = 1024;
  while (M < 2 * n - 1) M *= 2;
  return M;
}
size_t blue_line_bytes(int nmax) { return (size_t)f64::line_len(blue_m(nmax)) * 16; }
size_t align16(size_t b) { return (b + 15) & ~(size_t)15; }
size_t p1d_x_bytes(int nmax) {
  const size_t a = (size_t)nmax * 8, b = blue_line_bytes(nmax);
  return align16(a > b ? a : b);
}
// forward: [x (Horner) or the Bluestein line][pk][bins][bin runs]
size_t p1d_lds_bytes(int nmax, int nb) {
  return p1d_x_bytes(nmax) + (size_t)(nmax / 2 + 1) * 12 + (size_t)nb * 8;
}
// adjoint: [x][cj][Bluestein line]
size_t adj_line_off(int nmax) { return align16((size_t)nmax * 8 + (size_t)(nmax / 2 + 1) * 16); }
size_t adj_lds_bytes(int nmax) { return adj_line_off(nmax) + blue_line_bytes(nmax); }

// one transform size per function, so that each gets its own register
// allocation instead of the union of all of them
template <int M, class In, class Emit>
__device__ __noinline__ void bluestein_at(int N, int nout, In in, Emit emit, double2* line, const double2* tw) {
  f64::bluestein<M, M / FIT_THREADS>(N, nout, in, emit, line, tw, BLUE_MMAX);
}

// Bluestein DFT of length N <= BLUE_NMAX with M the smallest power of two
// >= max(2N - 1, 1024): X_k for k < nout via emit(k, X_k)
template <class In, class Emit>
__device__ __forceinline__ void dft_bluestein(int N, int nout, In in, Emit emit, double2* line, const double2* tw) {
  if (2 * N - 1 <= 1024) bluestein_at<1024>(N, nout, in, emit, line, tw);
  else if (2 * N - 1 <= 2048) bluestein_at<2048>(N, nout, in, emit, line, tw);
  else bluestein_at<4096>(N, nout, in, emit, line, tw);
}

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

__device__ __forceinline__ double4 prow(const double* abcg, int64_t rows, int64_t r) {
  if (r < 0 || r >= rows) return make_double4(__longlong_as_double(0x7ff8000000000000ll), 0.0, 0.0, 0.0);
  const double* q = abcg + 4 * r;
  return make_double4(q[0], q[1], q[2], q[3]);
}

// F for every pixel, in FP64 so that chi2(theta) is smooth to FP64 rounding
// (the optimisers' line searches need that); per-skewer sum of F over the
// forest [imin, imax).  Parameter row: pset[s], or with pset NULL the pixel's
// global grid index g0[s] + i (parameters at the pixel's redshift).
__global__ __launch_bounds__(FIT_THREADS) void k_fgpa_flux(smk_skewers sk, const float* __restrict__ delta,
                                                           const float* __restrict__ eta,
                                                           const double* __restrict__ abcg, int64_t rows,
                                                           double* __restrict__ flux, double* __restrict__ fsum) {
  __shared__ double red[FIT_THREADS / 64];
  for (int s = blockIdx.x; s < sk.n; s += gridDim.x) {
    const int64_t o = sk.off[s];
    const int n = sk.npix[s], i0 = sk.imin[s], i1 = sk.imax[s];
    const bool per_pixel = sk.pset == nullptr;
    const double4 ps = per_pixel ? make_double4(0.0, 0.0, 0.0, 0.0) : prow(abcg, rows, sk.pset[s]);
    double acc = 0.0;
    for (int i = threadIdx.x; i < n; i += FIT_THREADS) {
      const double4 p = per_pixel ? prow(abcg, rows, (int64_t)sk.g0[s] + i) : ps;
      const double g = (double)delta[o + i] + p.z * (eta ? (double)eta[o + i] : 0.0);
      const double f = exp(-p.x * exp(p.y * p.w * g));
      flux[o + i] = f;
      if (i >= i0 && i < i1) acc += f;
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
  return mul_rn(mul_rn(mul_rn((double)j, 1.0 / (double)n), 2.0), k_ny);
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
__device__ __forceinline__ int load_delta_f(const double* __restrict__ flux, int64_t o, int i0, int i1, int m,
                                           double inv_fmean, double* __restrict__ x) {
  const int nf = i1 - i0;
  const int nq = nf / m;
  for (int q = threadIdx.x; q < nq; q += FIT_THREADS) {
    double acc = 0.0;
    for (int r = 0; r < m; ++r) acc += flux[o + i0 + q * m + r];
    x[q] = (acc / m) * inv_fmean - 1.0;
  }
  return nq;
}

// S_r = sum_{n<N} y_n w_r^n, w_r = exp(i th_r), for R frequencies at once:
// R independent FP64 Horner chains share each LDS read of y_n, so the
// latency of one chain is covered by the others.
template <int R, bool CPLX>
__device__ __forceinline__ void horner_multi(const double* __restrict__ yr, const double2* __restrict__ yc, int N,
                                             const double* c, const double* s, double* re, double* im) {
#pragma unroll
  for (int r = 0; r < R; ++r) { re[r] = 0.0; im[r] = 0.0; }
  for (int n = N - 1; n >= 0; --n) {
    double xr, xi = 0.0;
    if (CPLX) { xr = yc[n].x; xi = yc[n].y; } else { xr = yr[n]; }
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const double r2 = fma(re[r], c[r], fma(-im[r], s[r], xr));
      im[r] = fma(re[r], s[r], fma(im[r], c[r], xi));
      re[r] = r2;
    }
  }
}

// For frequencies f = f0 + r*FIT_THREADS (r < R), f < nf:
// out_r = sum_{n<N} y_n exp(-2 pi i f n / P); calls `emit(f, re, im)`.
template <int R, bool CPLX, class Emit>
__device__ __forceinline__ void dft_group(const double* yr, const double2* yc, int N, int P, int f0, int nf,
                                          Emit emit) {
  double c[R], s[R], re[R], im[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int f = f0 + r * FIT_THREADS;
    if (f < nf) {
      sincospi(-2.0 * (double)f / (double)P, &s[r], &c[r]);
    } else {
      s[r] = 0.0;
      c[r] = 1.0;
    }
  }
  horner_multi<R, CPLX>(yr, yc, N, c, s, re, im);
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int f = f0 + r * FIT_THREADS;
    if (f < nf) emit(f, re[r], im[r]);
  }
}

// every frequency f < nf, FIT_THREADS * (up to 4) per sweep
template <bool CPLX, class Emit>
__device__ __forceinline__ void dft_all(const double* yr, const double2* yc, int N, int P, int nf, Emit emit) {
  for (int base = 0; base < nf; base += 4 * FIT_THREADS) {
    const int left = nf - base;
    const int R = left > 3 * FIT_THREADS ? 4 : left > 2 * FIT_THREADS ? 3 : left > FIT_THREADS ? 2 : 1;
    const int f0 = base + threadIdx.x;
    switch (R) {
      case 4: dft_group<4, CPLX>(yr, yc, N, P, f0, nf, emit); break;
      case 3: dft_group<3, CPLX>(yr, yc, N, P, f0, nf, emit); break;
      case 2: dft_group<2, CPLX>(yr, yc, N, P, f0, nf, emit); break;
      default: dft_group<1, CPLX>(yr, yc, N, P, f0, nf, emit); break;
    }
  }
}

struct P1dArgs {
  smk_skewers sk;
  const double* flux;
  const double* fmean;   // per set
  double pixel;          // spectrum pixel (before regrouping)
  int regroup;           // m >= 1
  const double* edges;   // nb + 1
  int nb;
  double* partial;       // [n][nb][4]: count, sum k, sum P, sum P^2
  int32_t* status;       // per skewer: 1 if the forest is longer than nmax
  int nmax;
  double2* xj;           // NULL or the modes X_j of every skewer, at xj + xoff[s]
  const int64_t* xoff;
  const double2* tw;     // W_4096^k, k < 4096 (FP64), for the Bluestein FFTs
  size_t x_bytes;        // p1d_x_bytes(nmax)
};

__global__ __launch_bounds__(FIT_THREADS) void k_p1d(P1dArgs a) {
  extern __shared__ double dyn[];
  double* x = dyn;                                    // Horner input
  double2* bline = reinterpret_cast<double2*>(dyn);   // Bluestein FFT line (aliases x)
  double* pk = reinterpret_cast<double*>(reinterpret_cast<char*>(dyn) + a.x_bytes);
  int* bins = reinterpret_cast<int*>(pk + a.nmax / 2 + 1);
  int* bfirst = bins + a.nmax / 2 + 1;
  int* blast = bfirst + a.nb;
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
    const double px = a.pixel * a.regroup;
    const double k_ny = M_PI / px;
    const int nj = nq > 0 ? nq / 2 + 1 : 0;
    double2* xj = a.xj ? a.xj + a.xoff[s] : nullptr;
    if (nq <= BLUE_NMAX && nq > 0) {
      // delta_F read straight from the flux (regrouped), the DFT by Bluestein
      const int64_t o = sk.off[s] + i0;
      const int m = a.regroup;
      dft_bluestein(nq, nj,
                    [&](int q) {
                      double acc = 0.0;
                      for (int r = 0; r < m; ++r) acc += a.flux[o + (int64_t)q * m + r];
                      return make_double2((acc / m) * inv - 1.0, 0.0);
                    },
                    [&](int j, double2 X) {
                      pk[j] = (X.x * X.x + X.y * X.y) * px / nq;
                      bins[j] = find_bin(a.edges, a.nb, kfreq(j, nq, k_ny));
                      if (xj) xj[j] = X;
                    },
                    bline, a.tw);
      __syncthreads();
    } else {
      load_delta_f(a.flux, sk.off[s], i0, i1, a.regroup, inv, x);
      __syncthreads();
      dft_all<false>(x, nullptr, nq, nq, nj, [&](int j, double re, double im) {
        pk[j] = (re * re + im * im) * px / nq;
        bins[j] = find_bin(a.edges, a.nb, kfreq(j, nq, k_ny));
        if (xj) xj[j] = make_double2(re, im);
      });
      __syncthreads();
    }
    // k_j increases with j, so each bin holds a run of consecutive modes
    // (bins[] is non-decreasing, then -1 past the last edge): find each
    // bin's run [bfirst, blast] from the transitions, then sum it in
    // ascending j (the order of a full scan, so the sums are the same)
    for (int b = threadIdx.x; b < a.nb; b += FIT_THREADS) {
      bfirst[b] = 0;
      blast[b] = -1;
    }
    __syncthreads();
    for (int j = threadIdx.x; j < nj; j += FIT_THREADS) {
      const int b = bins[j];
      if (b < 0) continue;
      if (j == 0 || bins[j - 1] != b) bfirst[b] = j;
      if (j == nj - 1 || bins[j + 1] != b) blast[b] = j;
    }
    __syncthreads();
    for (int b = threadIdx.x; b < a.nb; b += FIT_THREADS) {
      double c = 0.0, sk_ = 0.0, sp = 0.0, sp2 = 0.0;
      for (int j = bfirst[b]; j <= blast[b]; ++j) {
        c += 1.0;
        sk_ += kfreq(j, nq, k_ny);
        sp += pk[j];
        sp2 += pk[j] * pk[j];
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
  const double* abcg;
  int64_t rows;
  const double* flux;
  const double* fmean;
  double pixel;
  int regroup;
  const double* edges;
  int nb;
  const double* gp;      // [sets][nb]: dL/dP for one P_j sample of bin b
  double* out;           // [n][7]: sum A F', sum A dF/d(a,b,c) / m, sum dF/d(a,b,c)
  int nmax;
  const double2* xj;     // NULL: recompute the modes; else those saved by k_p1d
  const int64_t* xoff;
  const double2* tw;     // W_4096^k (FP64), for the Bluestein FFTs
  size_t line_off;       // adj_line_off(nmax): byte offset of the Bluestein line
};

// Adjoint of the binned P1D through delta_F = F/<F> - 1 (at fixed <F>) and
// of the FGPA: A_q = dL/d delta_F,q = sum_j gP_j (2 px / N) Re(conj(X_j) W^{jq}).
__global__ __launch_bounds__(FIT_THREADS) void k_p1d_adjoint(AdjArgs a) {
  extern __shared__ double dyn[];
  double* x = dyn;
  double2* cj = reinterpret_cast<double2*>(dyn + a.nmax);
  double2* bline = reinterpret_cast<double2*>(reinterpret_cast<char*>(dyn) + a.line_off);
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
    const double* gps = a.gp + (int64_t)set * a.nb;
    auto coef = [&](int j) -> double {
      const int b = find_bin(a.edges, a.nb, kfreq(j, nq, k_ny));
      return b >= 0 ? gps[b] * 2.0 * px / nq : 0.0;
    };
    const bool blue = nq > 0 && nq <= BLUE_NMAX;
    if (a.xj) {
      const double2* xs = a.xj + a.xoff[s];
      for (int j = threadIdx.x; j < nj; j += FIT_THREADS) {
        const double g = coef(j);
        const double2 X = xs[j];
        cj[j] = make_double2(g * X.x, -g * X.y);   // g conj(X_j)
      }
    } else if (blue) {
      dft_bluestein(nq, nj, [&](int n) { return make_double2(x[n], 0.0); },
                    [&](int j, double2 X) {
                      const double g = coef(j);
                      cj[j] = make_double2(g * X.x, -g * X.y);
                    },
                    bline, a.tw);
    } else {
      dft_all<false>(x, nullptr, nq, nq, nj, [&](int j, double re, double im) {
        const double g = coef(j);
        cj[j] = make_double2(g * re, -g * im);
      });
    }
    __syncthreads();
    // A_q = Re sum_j cj[j] W^{jq}, q < nq (a DFT of length nq of cj padded
    // with zeros).  The thread that owns q accumulates A_q F'_q and then
    // stores A_q over x[q] (x is not read by any other thread from here on).
    double saf = 0.0;
    if (blue) {
      dft_bluestein(nq, nq, [&](int n) { return n < nj ? cj[n] : make_double2(0.0, 0.0); },
                    [&](int q, double2 X) {
                      saf += X.x * (x[q] + 1.0) * fm;
                      x[q] = X.x;
                    },
                    bline, a.tw);
    } else {
      dft_all<true>(nullptr, cj, nj, nq, nq, [&](int q, double re, double im) {
        saf += re * (x[q] + 1.0) * fm;
        x[q] = re;
      });
    }
    __syncthreads();
    const double* A = x;
    const double4 p = prow(a.abcg, a.rows, set);
    const double pa = p.x, pb = p.y, pc = p.z, pg = p.w;
    double t1a = 0.0, t1b = 0.0, t1c = 0.0, t2a = 0.0, t2b = 0.0, t2c = 0.0;
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
        const double Aq = A[q] / m;
        t1a += Aq * dfa;
        t1b += Aq * dfb;
        t1c += Aq * dfc;
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

// W_4096^k, k < 4096, FP64, one table per device (built on first use)
const double2* fp64_twiddles(hipError_t* err) {
  static std::mutex mu;
  static const double2* tabs[64] = {};
  int dev = 0;
  *err = hipGetDevice(&dev);
  if (*err != hipSuccess || dev < 0 || dev >= 64) return nullptr;
  std::lock_guard<std::mutex> lk(mu);
  if (tabs[dev] == nullptr) {
    std::vector<double2> t(BLUE_MMAX);
    for (int k = 0; k < BLUE_MMAX; ++k) {
      const double ang = -2.0 * M_PI * (double)k / (double)BLUE_MMAX;
      t[k] = make_double2(cos(ang), sin(ang));
    }
    void* d = nullptr;
    *err = hipMalloc(&d, BLUE_MMAX * sizeof(double2));
    if (*err != hipSuccess) return nullptr;
    *err = hipMemcpy(d, t.data(), BLUE_MMAX * sizeof(double2), hipMemcpyHostToDevice);
    if (*err != hipSuccess) {
      (void)hipFree(d);
      return nullptr;
    }
    tabs[dev] = (const double2*)d;
  }
  return tabs[dev];
}

}  // namespace

extern "C" int smk_fgpa_flux(const smk_skewers* sk, const float* delta, const float* eta, const double* abcg,
                             int64_t abcg_rows, double* flux, double* fsum, void* stream) {
  SMK_REQUIRE(sk != nullptr && delta != nullptr && abcg != nullptr && flux != nullptr && fsum != nullptr,
              SMK_ERR_ARG, "NULL argument");
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

extern "C" int smk_p1d_binned(const smk_skewers* sk, const double* flux, const double* fmean, double pixel,
                              int regroup, const double* edges, int nbins, int max_forest, double* partial,
                              int32_t* status, double* modes, const int64_t* modes_off, void* stream) {
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
  SMK_REQUIRE(modes == nullptr || modes_off != nullptr, SMK_ERR_ARG, "modes need modes_off");
  hipError_t te = hipSuccess;
  const double2* tw = fp64_twiddles(&te);
  SMK_REQUIRE(tw != nullptr, SMK_ERR_HIP, "FP64 twiddle table: %s", hipGetErrorString(te));
  P1dArgs a{*sk, flux, fmean, pixel, regroup, edges, nbins, partial, status, nmax, (double2*)modes, modes_off,
            tw, p1d_x_bytes(nmax)};
  const size_t lds = p1d_lds_bytes(nmax, nbins);
  SMK_HIP_CHECK(hipFuncSetAttribute((const void*)k_p1d, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  hipLaunchKernelGGL(k_p1d, dim3(grid_for(sk->n)), dim3(FIT_THREADS), lds, (hipStream_t)stream, a);
  SMK_LAUNCH_CHECK();
  return SMK_OK;
}

extern "C" int smk_p1d_adjoint(const smk_skewers* sk, const float* delta, const float* eta, const double* abcg,
                               int64_t abcg_rows, const double* flux, const double* fmean, double pixel,
                               int regroup, const double* edges, int nbins, int max_forest, const double* modes,
                               const int64_t* modes_off, const double* gp, double* out, void* stream) {
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
  SMK_REQUIRE(modes == nullptr || modes_off != nullptr, SMK_ERR_ARG, "modes need modes_off");
  hipError_t te = hipSuccess;
  const double2* tw = fp64_twiddles(&te);
  SMK_REQUIRE(tw != nullptr, SMK_ERR_HIP, "FP64 twiddle table: %s", hipGetErrorString(te));
  AdjArgs a{*sk,  delta, eta,   abcg, abcg_rows, flux, fmean, pixel, regroup, edges,
            nbins, gp,   out, nmax, (const double2*)modes, modes_off, tw, adj_line_off(nmax)};
  const size_t lds = adj_lds_bytes(nmax);
  SMK_HIP_CHECK(
      hipFuncSetAttribute((const void*)k_p1d_adjoint, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  hipLaunchKernelGGL(k_p1d_adjoint, dim3(grid_for(sk->n)), dim3(FIT_THREADS), lds, (hipStream_t)stream, a);
  SMK_LAUNCH_CHECK();
  return SMK_OK;
}
