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
// With per-length tables (smk_p1d_tables: the chirp and the chirp's
// transform of every forest length of a batch, made once, since a fit's
// forests do not change between evaluations) a Bluestein DFT is two FFTs on
// one register tile instead of three on three tiles; the kernels are then
// separate instances (TAB = true) with their own, smaller register count.
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
constexpr int FIT_NLDS = 8192;  // longest (regrouped) forest of the LDS layouts; longer ones take the
                                // direct DFT with the input staged through LDS in pieces
constexpr int FIT_XC = 2048;    // input points per staged piece of that path (complex doubles)
constexpr int BLUE_NMAX = 2048; // longest forest of the Bluestein path (M <= 4096)
constexpr int BLUE_MMAX = 4096;

// dynamic LDS of the P1D kernels for forests of at most nmax (even) points:
// x, then the binning arrays (forward) or the adjoint's coefficients, with
// the Bluestein FFT line aliased over them
// transform size of the Bluestein DFT of N points: the smallest power of
// two >= max(2N - 1, 1024)
__host__ __device__ inline int blue_m_of(int N) {
  int M = 1024;
  while (M < 2 * N - 1) M *= 2;
  return M;
}
int blue_m(int nmax) { return blue_m_of(nmax < BLUE_NMAX ? nmax : BLUE_NMAX); }
size_t blue_line_bytes(int nmax) { return (size_t)f64::line_len(blue_m(nmax)) * 16; }
size_t align16(size_t b) { return (b + 15) & ~(size_t)15; }
size_t max_sz(size_t a, size_t b) { return a > b ? a : b; }
// forward, region A: the Bluestein line with pk and bins over its start
// (written once the last transform has left the line), or for Horner the
// input x, then pk and bins; after region A the edges (nb + 1 doubles) and
// the bin runs (2 nb ints)
size_t p1d_a_bytes(int nmax) {
  const size_t nj = (size_t)(nmax / 2 + 1);
  return align16(max_sz(blue_line_bytes(nmax), (size_t)nmax * 8 + nj * 12));
}
// + the edges, the bin runs (2 nb ints) and, for forests longer than the
// layout, the running bin sums (4 nb doubles)
size_t p1d_lds_bytes(int nmax, int nb) {
  return p1d_a_bytes(nmax) + (size_t)(nb + 1) * 8 + (size_t)nb * 8 + (size_t)nb * 32;
}
// adjoint, generic kernel: x (nmax doubles) and cj (nmax / 2 + 1 complex)
// for the Horner forests; a Bluestein forest (TAB = false, N <= BLUE_NMAX)
// keeps x and cj in the first BLUE_NMAX-point layout and the Bluestein line
// after them; then the edges
size_t adj_x_bytes(int nmax) { return align16((size_t)nmax * 8); }
size_t adj_cj_bytes(int nmax) { return align16((size_t)(nmax / 2 + 1) * 16); }
size_t adj_line_off(int nmax) {
  const int nb = nmax < BLUE_NMAX ? nmax : BLUE_NMAX;
  return adj_x_bytes(nb) + adj_cj_bytes(nb);
}
size_t adj_e_off(int nmax, bool tab) {
  const size_t horner = adj_x_bytes(nmax) + adj_cj_bytes(nmax);
  return tab ? horner : max_sz(horner, adj_line_off(nmax) + blue_line_bytes(nmax));
}

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

// sum over a block of NT threads (red: NT / 64 doubles), fixed order
template <int NT = FIT_THREADS>
__device__ __forceinline__ double block_reduce_sum(double v, double* red) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red[wv] = v;
  __syncthreads();
  double s = 0.0;
  for (int i = 0; i < NT / 64; ++i) s += red[i];
  __syncthreads();
  return s;  // valid on every thread
}

__device__ __forceinline__ double4 prow(const double* abcg, int64_t rows, int64_t r) {
  if (r < 0 || r >= rows) return make_double4(__longlong_as_double(0x7ff8000000000000ll), 0.0, 0.0, 0.0);
  const double* q = abcg + 4 * r;
  return make_double4(q[0], q[1], q[2], q[3]);
}

// sum over a wave, fixed order (every lane gets the same bits: each xor
// step adds the same two operands on both lanes of a pair)
__device__ __forceinline__ double wave_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// the per-skewer FGPA kernels run one wave per skewer (a forest of ~10^3
// pixels is ~16 loads per lane; a block per skewer spent its time on block
// barriers and the skewer's descriptor loads)
constexpr int FLUX_WAVES = FIT_THREADS / 64;
__device__ __forceinline__ int flux_wave() { return __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6); }
int grid_for_waves(int n) {
  const int b = (n + FLUX_WAVES - 1) / FLUX_WAVES;
  return b < 1 ? 1 : (b < 65536 ? b : 65536);
}

// F for every pixel, in FP64 so that chi2(theta) is smooth to FP64 rounding
// (the optimisers' line searches need that); per-skewer sum of F over the
// forest [imin, imax).  Parameter row: pset[s], or with pset NULL the pixel's
// global grid index g0[s] + i (parameters at the pixel's redshift).
// pixels per lane whose loads are issued together in the FGPA kernels
constexpr int FLUX_U = 8;

// one skewer's flux and forest sum of F on one wave (lane's share, pixels
// lane + 64 k in ascending k); PP: parameters per pixel; ETA: eta given
template <bool PP, bool ETA>
__device__ __forceinline__ double flux_skewer(const smk_skewers& sk, int s, const float* __restrict__ delta,
                                              const float* __restrict__ eta, const double* __restrict__ abcg,
                                              int64_t rows, double* __restrict__ flux, int lane) {
  const int64_t o = sk.off[s];
  const int n = sk.npix[s], i0 = sk.imin[s], i1 = sk.imax[s];
  const double4 ps = PP ? make_double4(0.0, 0.0, 0.0, 0.0) : prow(abcg, rows, sk.pset[s]);
  double acc = 0.0;
  for (int i = lane; i < n; i += 64 * FLUX_U) {
    float dv[FLUX_U], ev[FLUX_U];
#pragma unroll
    for (int u = 0; u < FLUX_U; ++u) {
      // past the end: a clamped (in-bounds, unused) load, so no load waits
      // on a branch
      const int j = i + 64 * u < n ? i + 64 * u : n - 1;
      dv[u] = delta[o + j];
      ev[u] = ETA ? eta[o + j] : 0.0f;
    }
#pragma unroll
    for (int u = 0; u < FLUX_U; ++u) {
      const int j = i + 64 * u;
      if (j < n) {
        const double4 p = PP ? prow(abcg, rows, (int64_t)sk.g0[s] + j) : ps;
        const double g = (double)dv[u] + p.z * (ETA ? (double)ev[u] : 0.0);
        const double f = exp(-p.x * exp(p.y * p.w * g));
        flux[o + j] = f;
        if (j >= i0 && j < i1) acc += f;
      }
    }
  }
  return acc;
}

__global__ __launch_bounds__(FIT_THREADS) void k_fgpa_flux(smk_skewers sk, const float* __restrict__ delta,
                                                           const float* __restrict__ eta,
                                                           const double* __restrict__ abcg, int64_t rows,
                                                           double* __restrict__ flux, double* __restrict__ fsum) {
  const int lane = threadIdx.x & 63;
  const bool per_pixel = sk.pset == nullptr;
  for (int s = blockIdx.x * FLUX_WAVES + flux_wave(); s < sk.n; s += gridDim.x * FLUX_WAVES) {
    double acc;
    if (per_pixel)
      acc = eta ? flux_skewer<true, true>(sk, s, delta, eta, abcg, rows, flux, lane)
                : flux_skewer<true, false>(sk, s, delta, eta, abcg, rows, flux, lane);
    else
      acc = eta ? flux_skewer<false, true>(sk, s, delta, eta, abcg, rows, flux, lane)
                : flux_skewer<false, false>(sk, s, delta, eta, abcg, rows, flux, lane);
    const double tot = wave_sum(acc);
    if (lane == 0) fsum[s] = tot;
  }
}

// sums over the skewers of each parameter set, out[set][c] = sum_{s in set}
// in[s][c], in two launches of fixed order.  Stage 1, block (column tile of
// 64, chunk of RED_CHUNK skewers, set): each wave takes 64 skewers of the
// chunk, finds those of the set with one ballot and adds their rows in
// ascending order, four loads in flight; the four waves' sums are added in
// wave order into work[set][chunk][c].  Stage 2, block (column tile, set):
// wave w adds chunks w, w + 4, ... in order, then the four waves in order,
// and divides by divisor[set] when given.
constexpr int RED_CHUNK = 256;
int red_chunks(int n) { return (n + RED_CHUNK - 1) / RED_CHUNK; }

__global__ __launch_bounds__(RED_CHUNK) void k_reduce_chunks(const int32_t* __restrict__ pset, int n,
                                                             const double* __restrict__ in, int width,
                                                             double* __restrict__ work) {
  __shared__ double red[RED_CHUNK / 64][64];
  const int lane = threadIdx.x & 63, w = flux_wave();
  const int col = blockIdx.x * 64 + lane;
  const int k = blockIdx.y, set = blockIdx.z;
  const int s0 = k * RED_CHUNK + w * 64;
  uint64_t mask = __ballot(s0 + lane < n && pset[s0 + lane] == set);
  double acc = 0.0;
  while (mask) {
    int ss[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      ss[u] = mask ? s0 + __builtin_ctzll(mask) : -1;
      mask &= mask - 1;
    }
    double v[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) v[u] = (ss[u] >= 0 && col < width) ? in[(int64_t)ss[u] * width + col] : 0.0;
#pragma unroll
    for (int u = 0; u < 4; ++u)
      if (ss[u] >= 0) acc += v[u];
  }
  red[w][lane] = acc;
  __syncthreads();
  if (w == 0 && col < width) {
    double t = red[0][lane];
    for (int i = 1; i < RED_CHUNK / 64; ++i) t += red[i][lane];
    work[((int64_t)set * gridDim.y + k) * width + col] = t;
  }
}

__global__ __launch_bounds__(256) void k_reduce_chunk_sums(const double* __restrict__ work, int nchunk, int width,
                                                           const double* __restrict__ divisor,
                                                           double* __restrict__ out) {
  __shared__ double red[4][64];
  const int lane = threadIdx.x & 63, w = flux_wave();
  const int col = blockIdx.x * 64 + lane, set = blockIdx.y;
  double acc = 0.0;
  if (col < width) {
    const double* q = work + (int64_t)set * nchunk * width + col;
#pragma unroll 8
    for (int k = w; k < nchunk; k += 4) acc += q[(int64_t)k * width];
  }
  red[w][lane] = acc;
  __syncthreads();
  if (w == 0 && col < width) {
    double t = ((red[0][lane] + red[1][lane]) + red[2][lane]) + red[3][lane];
    if (divisor) t = t / divisor[set];
    out[(int64_t)set * width + col] = t;
  }
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
template <int NT = FIT_THREADS>
__device__ __forceinline__ int load_delta_f(const double* __restrict__ flux, int64_t o, int i0, int i1, int m,
                                           double inv_fmean, double* __restrict__ x) {
  const int nf = i1 - i0;
  const int nq = nf / m;
  if (m == 1) {
    // four points per thread in flight: loads at clamped indices, then the
    // stores of the points that exist ((F / 1) * inv - 1, as below)
    for (int q = threadIdx.x; q < nq; q += 4 * NT) {
      double v[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) v[u] = flux[o + i0 + (q + u * NT < nq ? q + u * NT : nq - 1)];
#pragma unroll
      for (int u = 0; u < 4; ++u)
        if (q + u * NT < nq) x[q + u * NT] = (v[u] / 1) * inv_fmean - 1.0;
    }
    return nq;
  }
  for (int q = threadIdx.x; q < nq; q += NT) {
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

// first mode j in [0, nj) with k_j >= e (gt: k_j > e), nj if none (k_j
// increases with j)
__device__ __forceinline__ int first_mode(int nj, int n, double k_ny, double e, bool gt) {
  int lo = 0, hi = nj;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    const double k = kfreq(mid, n, k_ny);
    if (gt ? k > e : k >= e) hi = mid;
    else lo = mid + 1;
  }
  return lo;
}

// the modes [lo, hi) of bin b: e_b <= k_j < e_b+1, the last bin closed (the
// comparisons of find_bin, so the same membership)
__device__ __forceinline__ void bin_span(const double* edges, int nb, int b, int nj, int n, double k_ny, int& lo,
                                         int& hi) {
  lo = first_mode(nj, n, k_ny, edges[b], false);
  hi = b + 1 < nb ? first_mode(nj, n, k_ny, edges[b + 1], false) : first_mode(nj, n, k_ny, edges[nb], true);
}

// Horner's rule for the frequencies f = fb + threadIdx.x + r FIT_THREADS
// (r < 4, f < fe): out_f = sum_{n<N} y_n exp(-2 pi i f n / P), y staged
// through LDS (buf, FIT_XC points) in pieces, n descending; stage(n0, n1)
// fills buf[0, n1 - n0) with y_n0 .. y_n1-1 (every thread calls it)
template <class Stage>
__device__ __forceinline__ void horner_staged(Stage stage, double2* buf, int N, int P, int fb, int fe,
                                              double (&re)[4], double (&im)[4]) {
  double c[4], sn[4];
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int f = fb + (int)threadIdx.x + r * FIT_THREADS;
    if (f < fe) {
      sincospi(-2.0 * (double)f / (double)P, &sn[r], &c[r]);
    } else {
      sn[r] = 0.0;
      c[r] = 1.0;
    }
    re[r] = 0.0;
    im[r] = 0.0;
  }
  for (int nh = N; nh > 0; nh -= FIT_XC) {
    const int nl = nh > FIT_XC ? nh - FIT_XC : 0;
    __syncthreads();
    stage(nl, nh);
    __syncthreads();
    for (int n = nh - 1; n >= nl; --n) {
      const double2 y = buf[n - nl];
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const double r2 = fma(re[r], c[r], fma(-im[r], sn[r], y.x));
        im[r] = fma(re[r], sn[r], fma(im[r], c[r], y.y));
        re[r] = r2;
      }
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
  const double2* tables; // NULL, or the per-length Bluestein tables (bluestein_table layout)
  const int64_t* tab_off;// with tables: per skewer, its table at tables + tab_off[s]; < 0: none
  size_t a_bytes;        // byte offset of the edges in the dynamic LDS
  int nlds;              // longest forest of the LDS layout (min(nmax, FIT_NLDS))
  int chunk;             // modes per pass of the long-forest path
  const int32_t* pairs;  // NULL, or [n_pairs][2] skewers of one length sharing a DFT (second -1: alone)
  int pair_lo, pair_hi;  // the pairs of the launch's transform size
};

// skewer s has a table and a forest in the Bluestein range: the per-size
// kernels take it, the generic one skips it
__device__ __forceinline__ bool uses_table(const double2* tables, const int64_t* tab_off, int s, int nq,
                                           int nmax) {
  return tables != nullptr && tab_off[s] >= 0 && nq >= 1 && nq <= BLUE_NMAX && nq <= nmax;
}

// Z_j and Z_{N-j} of Z = DFT(u + i w), u and w real: U_j and W_j
__device__ __forceinline__ void split_pair(double2 zj, double2 zn, double2& u, double2& w) {
  u = make_double2(0.5 * (zj.x + zn.x), 0.5 * (zj.y - zn.y));
  w = make_double2(0.5 * (zj.y + zn.y), 0.5 * (zn.x - zj.x));
}

// bin runs of one skewer: k_j increases with j, so each bin holds a run of
// consecutive modes (bins[] is non-decreasing, then -1 past the last edge):
// find each bin's run [bfirst, blast] from the transitions, then sum it in
// ascending j (the order of a full scan, so the sums are the same)
template <int NT>
__device__ __forceinline__ void bin_runs(const double* pk, const int* bins, int nj, int nq, double k_ny, int nb,
                                         int* bfirst, int* blast, double* part) {
  for (int b = threadIdx.x; b < nb; b += NT) {
    bfirst[b] = 0;
    blast[b] = -1;
  }
  __syncthreads();
  for (int j = threadIdx.x; j < nj; j += NT) {
    const int b = bins[j];
    if (b < 0) continue;
    if (j == 0 || bins[j - 1] != b) bfirst[b] = j;
    if (j == nj - 1 || bins[j + 1] != b) blast[b] = j;
  }
  __syncthreads();
  for (int b = threadIdx.x; b < nb; b += NT) {
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

// Generic P1D kernel.  TAB = false: every skewer, Bluestein with the chirp
// formed in place (N <= 2048) or Horner.  TAB = true: the skewers without a
// usable table only (Horner; those with one go to k_p1d_blue).  Dynamic
// LDS: region A (the Bluestein line with pk and bins over its start, or the
// Horner input x, then pk and bins), the edges, the bin runs.
// the candidates c0 + t * gridDim.x (t < NT) that take(s, nq) accepts, as
// a bit per candidate in need[wave] (ascending), checked by all threads at
// once; every thread must call it
template <int NT, class Take>
__device__ __forceinline__ void cand_masks(const smk_skewers& sk, int m, int c0, uint64_t* need, Take take) {
  __syncthreads();   // the previous masks are no longer read
  const int64_t s = c0 + (int64_t)threadIdx.x * gridDim.x;
  bool t = false;
  if (s < sk.n) t = take((int)s, (sk.imax[s] - sk.imin[s]) / m);
  const uint64_t b = __ballot(t);
  if ((threadIdx.x & 63) == 0) need[threadIdx.x >> 6] = b;
  static_assert(NT % 64 == 0, "whole waves");
  __syncthreads();
}

// the generic kernels with tables take only the skewers without a usable
// one: block b checks its candidates c0 + t * gridDim.x (t < FIT_THREADS)
// with one load round, as a bit per candidate in need[wave] (candidates in
// ascending order); every thread must call it
__device__ __forceinline__ void rest_masks(const smk_skewers& sk, int m, const double2* tables,
                                           const int64_t* tab_off, int nmax, int c0, uint64_t* need) {
  __syncthreads();   // the previous masks are no longer read
  const int64_t s = c0 + (int64_t)threadIdx.x * gridDim.x;
  bool take = false;
  if (s < sk.n) {
    const int nq = (sk.imax[s] - sk.imin[s]) / m;
    take = !uses_table(tables, tab_off, (int)s, nq, nmax);
  }
  const uint64_t b = __ballot(take);
  if ((threadIdx.x & 63) == 0) need[threadIdx.x >> 6] = b;
  __syncthreads();
}

template <bool TAB>
__global__ __launch_bounds__(FIT_THREADS) void k_p1d(P1dArgs a) {
  extern __shared__ double dyn[];
  const int njmax = a.nlds / 2 + 1;
  double* ledges = reinterpret_cast<double*>(reinterpret_cast<char*>(dyn) + a.a_bytes);
  int* bfirst = reinterpret_cast<int*>(ledges + a.nb + 1);
  int* blast = bfirst + a.nb;
  double* bsum = reinterpret_cast<double*>(blast + a.nb);
  for (int i = threadIdx.x; i <= a.nb; i += FIT_THREADS) ledges[i] = a.edges[i];
  const smk_skewers& sk = a.sk;
  // one skewer (returns early where the loop over skewers moved on)
  auto one = [&](const int s) {
    double* part = a.partial + (int64_t)s * a.nb * 4;
    const int i0 = sk.imin[s], i1 = sk.imax[s];
    const int nq = (i1 - i0) / a.regroup;
    __syncthreads();
    if (nq > a.nmax) {
      for (int b = threadIdx.x; b < 4 * a.nb; b += FIT_THREADS) part[b] = 0.0;
      if (threadIdx.x == 0 && a.status) a.status[s] = 1;
      return;
    }
    const double inv = 1.0 / a.fmean[sk.pset[s]];
    const double px = a.pixel * a.regroup;
    const double k_ny = M_PI / px;
    const int nj = nq > 0 ? nq / 2 + 1 : 0;
    double2* xj = a.xj ? a.xj + a.xoff[s] : nullptr;
    if (nq > a.nlds) {
      // longer than the LDS layout: the direct DFT with delta_F staged
      // through LDS in pieces, the modes in passes of a.chunk, each bin's
      // running sums in LDS, added in ascending j as a full scan adds them
      double2* xbuf = reinterpret_cast<double2*>(dyn);
      double* pkc = dyn + 2 * FIT_XC;
      for (int b = threadIdx.x; b < a.nb; b += FIT_THREADS) {
        bin_span(ledges, a.nb, b, nj, nq, k_ny, bfirst[b], blast[b]);
        bsum[4 * b] = bsum[4 * b + 1] = bsum[4 * b + 2] = bsum[4 * b + 3] = 0.0;
      }
      const int64_t o = sk.off[s] + i0;
      const int m = a.regroup;
      auto stage = [&](int n0, int n1) {
        for (int n = n0 + (int)threadIdx.x; n < n1; n += FIT_THREADS) {
          double acc = 0.0;
          for (int r = 0; r < m; ++r) acc += a.flux[o + (int64_t)n * m + r];
          xbuf[n - n0] = make_double2((acc / m) * inv - 1.0, 0.0);
        }
      };
      for (int j0 = 0; j0 < nj; j0 += a.chunk) {
        const int j1 = j0 + a.chunk < nj ? j0 + a.chunk : nj;
        for (int fb = j0; fb < j1; fb += 4 * FIT_THREADS) {
          double re[4], im[4];
          horner_staged(stage, xbuf, nq, nq, fb, j1, re, im);
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int j = fb + (int)threadIdx.x + r * FIT_THREADS;
            if (j < j1) {
              pkc[j - j0] = (re[r] * re[r] + im[r] * im[r]) * px / nq;
              if (xj) xj[j] = make_double2(re[r], im[r]);
            }
          }
        }
        __syncthreads();
        for (int b = threadIdx.x; b < a.nb; b += FIT_THREADS) {
          const int lo = bfirst[b] > j0 ? bfirst[b] : j0, hi = blast[b] < j1 ? blast[b] : j1;
          double c = bsum[4 * b], sk_ = bsum[4 * b + 1], sp = bsum[4 * b + 2], sp2 = bsum[4 * b + 3];
          for (int j = lo; j < hi; ++j) {
            c += 1.0;
            sk_ += kfreq(j, nq, k_ny);
            sp += pkc[j - j0];
            sp2 += pkc[j - j0] * pkc[j - j0];
          }
          bsum[4 * b] = c; bsum[4 * b + 1] = sk_; bsum[4 * b + 2] = sp; bsum[4 * b + 3] = sp2;
        }
        __syncthreads();
      }
      for (int b = threadIdx.x; b < 4 * a.nb; b += FIT_THREADS) part[b] = bsum[b];
      return;
    }
    const bool blue = !TAB && nq > 0 && nq <= BLUE_NMAX;
    double* pk = blue ? dyn : dyn + a.nlds;
    int* bins = reinterpret_cast<int*>(pk + njmax);
    if (blue) {
      if constexpr (!TAB) {
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
                        bins[j] = find_bin(ledges, a.nb, kfreq(j, nq, k_ny));
                        if (xj) xj[j] = X;
                      },
                      reinterpret_cast<double2*>(dyn), a.tw);
      }
      __syncthreads();
    } else {
      double* x = dyn;
      load_delta_f(a.flux, sk.off[s], i0, i1, a.regroup, inv, x);
      __syncthreads();
      dft_all<false>(x, nullptr, nq, nq, nj, [&](int j, double re, double im) {
        pk[j] = (re * re + im * im) * px / nq;
        bins[j] = find_bin(ledges, a.nb, kfreq(j, nq, k_ny));
        if (xj) xj[j] = make_double2(re, im);
      });
      __syncthreads();
    }
    bin_runs<FIT_THREADS>(pk, bins, nj, nq, k_ny, a.nb, bfirst, blast, part);
  };
  if constexpr (TAB) {
    __shared__ uint64_t need[FIT_THREADS / 64];
    for (int c0 = blockIdx.x; c0 < sk.n; c0 += gridDim.x * FIT_THREADS) {
      rest_masks(sk, a.regroup, a.tables, a.tab_off, a.nmax, c0, need);
      for (int w = 0; w < FIT_THREADS / 64; ++w)
        for (uint64_t mk = need[w]; mk; mk &= mk - 1) one(c0 + (w * 64 + __builtin_ctzll(mk)) * (int)gridDim.x);
    }
  } else {
    for (int s = blockIdx.x; s < sk.n; s += gridDim.x) one(s);
  }
}

// LDS of k_p1d_blue<M>: the line (pk and bins over its start), the edges,
// the bin runs
__host__ __device__ constexpr size_t blue_line_bytes_m(int M) { return (((size_t)f64::line_len(M) * 16) + 15) & ~(size_t)15; }

// The P1D of the skewers with a table and a forest of M's size class
// (blue_m_of(N) == M), M / 8 threads of 8 points each: the DFT by
// bluestein_tab (two FFTs on one register tile), then the bins.
template <int M>
__global__ __launch_bounds__(M / 8) __attribute__((amdgpu_waves_per_eu(2))) void k_p1d_blue(P1dArgs a) {
  constexpr int NT = M / 8;
  extern __shared__ double dyn[];
  double2* line = reinterpret_cast<double2*>(dyn);
  double* pk = dyn;
  int* bins = reinterpret_cast<int*>(pk + (M / 4 + 1));
  double* ledges = reinterpret_cast<double*>(reinterpret_cast<char*>(dyn) + blue_line_bytes_m(M));
  int* bfirst = reinterpret_cast<int*>(ledges + a.nb + 1);
  int* blast = bfirst + a.nb;
  for (int i = threadIdx.x; i <= a.nb; i += NT) ledges[i] = a.edges[i];
  const smk_skewers& sk = a.sk;
  for (int s = blockIdx.x; s < sk.n; s += gridDim.x) {
    const int i0 = sk.imin[s], i1 = sk.imax[s];
    const int nq = (i1 - i0) / a.regroup;
    if (!uses_table(a.tables, a.tab_off, s, nq, a.nmax) || blue_m_of(nq) != M) continue;
    __syncthreads();
    double* part = a.partial + (int64_t)s * a.nb * 4;
    const double inv = 1.0 / a.fmean[sk.pset[s]];
    const double px = a.pixel * a.regroup;
    const double k_ny = M_PI / px;
    const int nj = nq / 2 + 1;
    double2* xj = a.xj ? a.xj + a.xoff[s] : nullptr;
    const double2* T = a.tables + a.tab_off[s];
    const int64_t o = sk.off[s] + i0;
    const int m = a.regroup;
    f64::bluestein_tab<M, 8>(
        nq, nj,
        [&](int q) {
          double acc = 0.0;
          for (int r = 0; r < m; ++r) acc += a.flux[o + (int64_t)q * m + r];
          return make_double2((acc / m) * inv - 1.0, 0.0);
        },
        [&](int j, double2 X) {
          pk[j] = (X.x * X.x + X.y * X.y) * px / nq;
          bins[j] = find_bin(ledges, a.nb, kfreq(j, nq, k_ny));
          if (xj) xj[j] = X;
        },
        line, a.tw, BLUE_MMAX, T, T + nq);
    __syncthreads();
    bin_runs<NT>(pk, bins, nj, nq, k_ny, a.nb, bfirst, blast, part);
  }
}

// LDS of k_p1d_blue_pair<M>: the line, the two forests' pk (NJ each,
// NJ = M/4 + 1 >= N/2 + 1), the bins, then the edges and the bin runs
__host__ __device__ constexpr size_t pair_fixed_bytes(int M) {
  return blue_line_bytes_m(M) + (size_t)(M / 4 + 1) * 16 + (size_t)(((M / 4 + 1) + 1) & ~1) * 4;
}

// The members (s1, s2) of pair w that take M's size class: each must still
// have a table of this size class, the second one of the same length and
// table; a member that does not is left to the generic kernel and its
// partner goes alone (s2 = -1); s1 = -1 when neither does
template <int M>
__device__ __forceinline__ void pair_members(const P1dArgs& a, int w, int& s1, int& s2) {
  const smk_skewers& sk = a.sk;
  const int m = a.regroup;
  s1 = a.pairs[2 * w];
  s2 = a.pairs[2 * w + 1];
  const int nq1 = (sk.imax[s1] - sk.imin[s1]) / m;
  const bool v1 = uses_table(a.tables, a.tab_off, s1, nq1, a.nmax) && blue_m_of(nq1) == M;
  bool v2 = s2 >= 0;
  if (v2) {
    const int nq2 = (sk.imax[s2] - sk.imin[s2]) / m;
    v2 = uses_table(a.tables, a.tab_off, s2, nq2, a.nmax) && blue_m_of(nq2) == M &&
         (!v1 || (nq2 == nq1 && a.tab_off[s2] == a.tab_off[s1]));
  }
  if (!v1) {
    s1 = v2 ? s2 : -1;
    s2 = -1;
  } else if (!v2) {
    s2 = -1;
  }
}

// What a pair kernel reads about the members of one pair before it can
// start, loaded a pair ahead so that the chain pair list -> the members'
// bounds and tables -> their flux runs while the previous pair is
// transformed
struct PairDesc {
  int s1, s2;            // as in the pair list (s2 < 0: alone; s1 < 0: no pair)
  int n1, n2;            // forest lengths imax - imin
  int64_t o1, o2;        // sk.off + imin
  int64_t t1, t2;        // tab_off (-1 without tables)
  int set1, set2;        // pset
};

template <class Args>
__device__ __forceinline__ PairDesc pair_desc(const Args& a, int s1, int s2) {
  const smk_skewers& sk = a.sk;
  PairDesc d{s1, s2, 0, 0, 0, 0, -1, -1, 0, 0};
  if (s1 >= 0) {
    d.n1 = sk.imax[s1] - sk.imin[s1];
    d.o1 = sk.off[s1] + sk.imin[s1];
    d.t1 = a.tab_off ? a.tab_off[s1] : -1;
    d.set1 = sk.pset[s1];
  }
  if (s2 >= 0) {
    d.n2 = sk.imax[s2] - sk.imin[s2];
    d.o2 = sk.off[s2] + sk.imin[s2];
    d.t2 = a.tab_off ? a.tab_off[s2] : -1;
    d.set2 = sk.pset[s2];
  }
  return d;
}

// pair_members on a descriptor: the members that take M's size class (s1 <
// 0: neither); swap: the second member of the list goes alone
template <int M, class Args>
__device__ __forceinline__ void desc_members(const Args& a, const PairDesc& d, int& s1, int& s2, bool& swap) {
  const int m = a.regroup;
  const int nq1 = d.n1 / m;
  const bool v1 = d.s1 >= 0 && a.tables != nullptr && d.t1 >= 0 && nq1 >= 1 && nq1 <= BLUE_NMAX &&
                  nq1 <= a.nmax && blue_m_of(nq1) == M;
  bool v2 = d.s2 >= 0;
  if (v2) {
    const int nq2 = d.n2 / m;
    v2 = a.tables != nullptr && d.t2 >= 0 && nq2 >= 1 && nq2 <= BLUE_NMAX && nq2 <= a.nmax &&
         blue_m_of(nq2) == M && (!v1 || (nq2 == nq1 && d.t2 == d.t1));
  }
  swap = !v1 && v2;
  if (!v1) {
    s1 = v2 ? d.s2 : -1;
    s2 = -1;
  } else {
    s1 = d.s1;
    s2 = v2 ? d.s2 : -1;
  }
}

// The P1D of two forests of one length with one Bluestein DFT: the DFT of
// delta_F1 + i delta_F2 (all N outputs into the line), split by the
// symmetry of real inputs into X1 and X2, then each forest's bins.  The
// pairs [pair_lo, pair_hi) are those of M's size class.  Persistent: block
// b takes pairs pair_lo + b + k gridDim.x, with the members' descriptors
// loaded a pair ahead and the pair list entry two pairs ahead.
template <int M>
__global__ __launch_bounds__(M / 8) __attribute__((amdgpu_waves_per_eu(2))) void k_p1d_blue_pair(P1dArgs a) {
  constexpr int NT = M / 8;
  constexpr int NJ = M / 4 + 1;
  extern __shared__ double dyn[];
  double2* line = reinterpret_cast<double2*>(dyn);
  double* pk1 = reinterpret_cast<double*>(reinterpret_cast<char*>(dyn) + blue_line_bytes_m(M));
  double* pk2 = pk1 + NJ;
  int* bins = reinterpret_cast<int*>(pk2 + NJ);
  double* ledges = reinterpret_cast<double*>(bins + ((NJ + 1) & ~1));
  int* bfirst = reinterpret_cast<int*>(ledges + a.nb + 1);
  int* blast = bfirst + a.nb;
  for (int i = threadIdx.x; i <= a.nb; i += NT) ledges[i] = a.edges[i];
  const int m = a.regroup;
  const double px = a.pixel * m;
  const double k_ny = M_PI / px;
  const int G = gridDim.x;
  int w = a.pair_lo + (int)blockIdx.x;
  PairDesc cur = w < a.pair_hi ? pair_desc(a, a.pairs[2 * w], a.pairs[2 * w + 1]) : pair_desc(a, -1, -1);
  int q1 = -1, q2 = -1;   // the pair list entry of pair w + G
  if (w + G < a.pair_hi) {
    q1 = a.pairs[2 * (w + G)];
    q2 = a.pairs[2 * (w + G) + 1];
  }
  for (; w < a.pair_hi; w += G) {
    // in flight while this pair is transformed: the next pair's
    // descriptors and the pair list entry after it
    const PairDesc nxt = pair_desc(a, q1, q2);
    q1 = q2 = -1;
    if (w + 2 * G < a.pair_hi) {
      q1 = a.pairs[2 * (w + 2 * G)];
      q2 = a.pairs[2 * (w + 2 * G) + 1];
    }
    int s1, s2;
    bool swap;
    desc_members<M>(a, cur, s1, s2, swap);
    const int64_t o1 = swap ? cur.o2 : cur.o1;
    const int set1 = swap ? cur.set2 : cur.set1;
    const int64_t tb = swap ? cur.t2 : cur.t1;
    const int nq = (swap ? cur.n2 : cur.n1) / m;
    // a lone forest reads its own pixels as the (unused) partner
    const int64_t o2 = s2 >= 0 ? cur.o2 : o1;
    const int set2 = s2 >= 0 ? cur.set2 : set1;
    cur = nxt;
    if (s1 < 0) continue;
    __syncthreads();
    const double inv1 = 1.0 / a.fmean[set1];
    const double inv2 = s2 >= 0 ? 1.0 / a.fmean[set2] : 0.0;
    const int nj = nq / 2 + 1;
    const double2* T = a.tables + tb;
    f64::bluestein_tab<M, 8, (M >= 4096 ? f64::KEEP_B | f64::KEEP_W : 0)>(
        nq, nq,
        [&](int q) {
          double acc1 = 0.0, acc2 = 0.0;
          for (int r = 0; r < m; ++r) {
            acc1 += a.flux[o1 + (int64_t)q * m + r];
            acc2 += a.flux[o2 + (int64_t)q * m + r];
          }
          return make_double2((acc1 / m) * inv1 - 1.0, s2 >= 0 ? (acc2 / m) * inv2 - 1.0 : 0.0);
        },
        [&](int j, double2 Z) { line[j] = Z; }, line, a.tw, BLUE_MMAX, T, T + nq);
    __syncthreads();
    double2* xj1 = a.xj ? a.xj + a.xoff[s1] : nullptr;
    double2* xj2 = (a.xj && s2 >= 0) ? a.xj + a.xoff[s2] : nullptr;
    for (int j = threadIdx.x; j < nj; j += NT) {
      double2 X1, X2;
      split_pair(line[j], line[j == 0 ? 0 : nq - j], X1, X2);
      pk1[j] = (X1.x * X1.x + X1.y * X1.y) * px / nq;
      pk2[j] = (X2.x * X2.x + X2.y * X2.y) * px / nq;
      bins[j] = find_bin(ledges, a.nb, kfreq(j, nq, k_ny));
      if (xj1) xj1[j] = X1;
      if (xj2) xj2[j] = X2;
    }
    __syncthreads();
    bin_runs<NT>(pk1, bins, nj, nq, k_ny, a.nb, bfirst, blast, a.partial + (int64_t)s1 * a.nb * 4);
    if (s2 >= 0) {
      __syncthreads();
      bin_runs<NT>(pk2, bins, nj, nq, k_ny, a.nb, bfirst, blast, a.partial + (int64_t)s2 * a.nb * 4);
    }
  }
}

// LDS of k_p1d_blue_quad<M>: two lines, each holding past its first M/2
// points (a pair's outputs need at most N <= M/2) its two forests' pk (NJ
// each, NJ = M/4 + 1) and its bins; then the edges and the bin runs
__host__ __device__ constexpr size_t quad_fixed_bytes(int M) { return 2 * blue_line_bytes_m(M); }
static_assert((size_t)(f64::line_len(1024) - 512) * 16 >= (size_t)(1024 / 4 + 1) * 16 + (size_t)(1024 / 4 + 2) * 4,
              "a pair's pk and bins must fit past the outputs in its line");

// Two pairs of one size class (pairs w and w + 1, their lengths may differ)
// in lockstep: bluestein_tab2 runs both pairs' DFTs with one set of barriers
// and twiddles, each emitting all N outputs into its line; then each pair
// is split and binned as in k_p1d_blue_pair.
template <int M>
__global__ __launch_bounds__(M / 8) __attribute__((amdgpu_waves_per_eu(2))) void k_p1d_blue_quad(P1dArgs a) {
  constexpr int NT = M / 8;
  constexpr int NJ = M / 4 + 1;
  extern __shared__ double dyn[];
  double2* la = reinterpret_cast<double2*>(dyn);
  double2* lb = reinterpret_cast<double2*>(reinterpret_cast<char*>(dyn) + blue_line_bytes_m(M));
  double* ledges = reinterpret_cast<double*>(reinterpret_cast<char*>(dyn) + quad_fixed_bytes(M));
  int* bfirst = reinterpret_cast<int*>(ledges + a.nb + 1);
  int* blast = bfirst + a.nb;
  for (int i = threadIdx.x; i <= a.nb; i += NT) ledges[i] = a.edges[i];
  const smk_skewers& sk = a.sk;
  const int m = a.regroup;
  const double px = a.pixel * m;
  const double k_ny = M_PI / px;
  for (int w = a.pair_lo + 2 * (int)blockIdx.x; w < a.pair_hi; w += 2 * (int)gridDim.x) {
    int s1, s2, s3 = -1, s4 = -1;
    pair_members<M>(a, w, s1, s2);
    if (w + 1 < a.pair_hi) pair_members<M>(a, w + 1, s3, s4);
    if (s1 < 0) {
      s1 = s3;
      s2 = s4;
      s3 = s4 = -1;
    }
    if (s1 < 0) continue;
    const int nqa = (sk.imax[s1] - sk.imin[s1]) / m;
    const int nqb = s3 >= 0 ? (sk.imax[s3] - sk.imin[s3]) / m : 0;
    __syncthreads();
    const double inv1 = 1.0 / a.fmean[sk.pset[s1]];
    const double inv2 = s2 >= 0 ? 1.0 / a.fmean[sk.pset[s2]] : 0.0;
    const double inv3 = s3 >= 0 ? 1.0 / a.fmean[sk.pset[s3]] : 0.0;
    const double inv4 = s4 >= 0 ? 1.0 / a.fmean[sk.pset[s4]] : 0.0;
    // an absent member reads its partner's (or the first) pixels, unused
    const int64_t o1 = sk.off[s1] + sk.imin[s1];
    const int64_t o2 = s2 >= 0 ? sk.off[s2] + sk.imin[s2] : o1;
    const int64_t o3 = s3 >= 0 ? sk.off[s3] + sk.imin[s3] : o1;
    const int64_t o4 = s4 >= 0 ? sk.off[s4] + sk.imin[s4] : o3;
    const double2* Ta = a.tables + a.tab_off[s1];
    // without a second pair its line transforms zeros, with the first's table
    const double2* Tb = s3 >= 0 ? a.tables + a.tab_off[s3] : Ta;
    auto dflux = [&](int64_t o, int q, double inv) {
      double acc = 0.0;
      for (int r = 0; r < m; ++r) acc += a.flux[o + (int64_t)q * m + r];
      return (acc / m) * inv - 1.0;
    };
    f64::bluestein_tab2<M, 8, false>(
        nqa, nqb,
        [&](int q) {
          // q is below nqa or nqb; each pair reads at its own clamped index
          const int qa = q < nqa ? q : nqa - 1, qb = q < nqb ? q : (nqb > 0 ? nqb - 1 : 0);
          return make_double4(dflux(o1, qa, inv1), s2 >= 0 ? dflux(o2, qa, inv2) : 0.0,
                              s3 >= 0 ? dflux(o3, qb, inv3) : 0.0, s4 >= 0 ? dflux(o4, qb, inv4) : 0.0);
        },
        [&](int j, double2 Z) { la[j] = Z; }, [&](int j, double2 Z) { lb[j] = Z; }, la, lb, a.tw, BLUE_MMAX, Ta,
        Ta + nqa, Tb, Tb + (s3 >= 0 ? nqb : nqa));
    __syncthreads();
    for (int h = 0; h < 2; ++h) {
      const int nq = h ? nqb : nqa;
      const int sx = h ? s3 : s1, sy = h ? s4 : s2;
      if (sx < 0) continue;
      const double2* line = h ? lb : la;
      double* pk1 = reinterpret_cast<double*>(const_cast<double2*>(line) + M / 2);
      double* pk2 = pk1 + NJ;
      int* bins = reinterpret_cast<int*>(pk2 + NJ);
      double2* xj1 = a.xj ? a.xj + a.xoff[sx] : nullptr;
      double2* xj2 = (a.xj && sy >= 0) ? a.xj + a.xoff[sy] : nullptr;
      const int nj = nq / 2 + 1;
      for (int j = threadIdx.x; j < nj; j += NT) {
        double2 X1, X2;
        split_pair(line[j], line[j == 0 ? 0 : nq - j], X1, X2);
        pk1[j] = (X1.x * X1.x + X1.y * X1.y) * px / nq;
        pk2[j] = (X2.x * X2.x + X2.y * X2.y) * px / nq;
        bins[j] = find_bin(ledges, a.nb, kfreq(j, nq, k_ny));
        if (xj1) xj1[j] = X1;
        if (xj2) xj2[j] = X2;
      }
    }
    for (int f = 0; f < 4; ++f) {
      const int h = f >> 1;
      const int sf = f == 0 ? s1 : (f == 1 ? s2 : (f == 2 ? s3 : s4));
      if (sf < 0) continue;
      const int nq = h ? nqb : nqa;
      double* pk = reinterpret_cast<double*>((h ? lb : la) + M / 2) + (f & 1) * NJ;
      int* bins = reinterpret_cast<int*>(reinterpret_cast<double*>((h ? lb : la) + M / 2) + 2 * NJ);
      __syncthreads();
      bin_runs<NT>(pk, bins, nq / 2 + 1, nq, k_ny, a.nb, bfirst, blast, a.partial + (int64_t)sf * a.nb * 4);
    }
  }
}

// the Bluestein table of every length lens[d] at tables + offs[d]
// (bluestein_table layout); lengths outside [1, BLUE_NMAX] are skipped
__global__ __launch_bounds__(FIT_THREADS) void k_p1d_tables(const int32_t* __restrict__ lens,
                                                            const int64_t* __restrict__ offs, int n,
                                                            double2* __restrict__ tables,
                                                            const double2* __restrict__ tw) {
  extern __shared__ double dyn[];
  double2* line = reinterpret_cast<double2*>(dyn);
  for (int d = blockIdx.x; d < n; d += gridDim.x) {
    const int N = lens[d];
    __syncthreads();
    if (N < 1 || N > BLUE_NMAX) continue;
    double2* T = tables + offs[d];
    if (2 * N - 1 <= 1024) f64::bluestein_table<1024, 1024 / FIT_THREADS>(N, T, line, tw, BLUE_MMAX);
    else if (2 * N - 1 <= 2048) f64::bluestein_table<2048, 2048 / FIT_THREADS>(N, T, line, tw, BLUE_MMAX);
    else f64::bluestein_table<4096, 4096 / FIT_THREADS>(N, T, line, tw, BLUE_MMAX);
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
  const double2* tables; // NULL, or the per-length Bluestein tables
  const int64_t* tab_off;
  size_t line_off;       // generic kernel: byte offset of the Bluestein line
  size_t cjb_off;        // generic kernel: byte offset of cj for the Bluestein forests
  size_t e_off;          // generic kernel: byte offset of the edges
  int nlds;              // longest forest of the LDS layout (min(nmax, FIT_NLDS))
  const int32_t* pairs;  // NULL, or the pairs of smk_p1d_binned
  int pair_lo, pair_hi;
};

// res[k] = the block sum of v[k], k < K, in one pass: a shuffle tree per
// wave, then the waves' partials in a fixed order (red: K NT / 64 doubles);
// the order of block_reduce_sum, so the same bits for each k
template <int NT, int K>
__device__ __forceinline__ void reduce_k(double (&v)[K], double* red, double* res) {
  constexpr int NW = NT / 64;
#pragma unroll
  for (int k = 0; k < K; ++k)
    for (int off = 32; off > 0; off >>= 1) v[k] += __shfl_xor(v[k], off, 64);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < K; ++k) red[k * NW + wv] = v[k];
  }
  __syncthreads();
  if (threadIdx.x < K) {
    double t = 0.0;
    for (int i = 0; i < NW; ++i) t += red[threadIdx.x * NW + i];
    res[threadIdx.x] = t;
  }
  __syncthreads();
}
template <int NT>
__device__ __forceinline__ void reduce7(double (&v)[7], double* red, double* res) {
  reduce_k<NT, 7>(v, red, res);
}

// dF/d(a, b, c) of pixel i (forest offset) of skewer s at parameters p
struct DF { double a, b, c; };
__device__ __forceinline__ DF fgpa_df(const AdjArgs& a, int64_t at, double4 p) {
  const double d = a.delta[at];
  const double e = a.eta ? (double)a.eta[at] : 0.0;
  const double gf = d + p.z * e;
  const double T = exp(p.y * p.w * gf);
  const double F = a.flux[at];
  return DF{-T * F, -p.x * T * F * p.w * gf, -p.x * T * F * p.y * p.w * e};
}

// The FGPA part of the adjoint of one skewer, with A_q (q < nq) in A:
// (sum A F', sum A_q(n) dF_n/d(a, b, c) / m, sum dF_n/d(a, b, c)) into res.
// F is the forward's flux at the same parameters (the same expression, so
// the same bits as recomputing it).  The seven block sums share one pass:
// a shuffle tree per wave, then the waves' partials in a fixed order (red:
// 7 NT / 64 doubles).
template <int NT>
__device__ __forceinline__ void adj_fgpa(const AdjArgs& a, int set, int64_t o0, int nf, int nq, double saf,
                                         const double* A, double* red, double* res) {
  const int m = a.regroup;
  const int64_t o = o0;
  const double4 p = prow(a.abcg, a.rows, set);
  const double pa = p.x, pb = p.y, pc = p.z, pg = p.w;
  double v[7] = {saf, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  // two pixels per thread and pass with their loads issued together (at a
  // clamped index past the end, unused; without eta the load reads delta
  // and is not used), in the order of one pixel per pass
  const float* ep = a.eta ? a.eta : a.delta;
  for (int i = threadIdx.x; i < nf; i += 2 * NT) {
    float dv[2], ev[2];
    double fv[2];
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      const int j = i + u * NT < nf ? i + u * NT : nf - 1;
      dv[u] = a.delta[o + j];
      ev[u] = ep[o + j];
      fv[u] = a.flux[o + j];
    }
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      const int j = i + u * NT;
      if (j >= nf) break;
      const int q = j / m;
      const double d = dv[u];
      const double e = a.eta ? (double)ev[u] : 0.0;
      const double gf = d + pc * e;
      const double T = exp(pb * pg * gf);
      const double F = fv[u];
      const double dfa = -T * F;
      const double dfb = -pa * T * F * pg * gf;
      const double dfc = -pa * T * F * pb * pg * e;
      v[4] += dfa; v[5] += dfb; v[6] += dfc;
      if (q < nq) {
        const double Aq = A[q] / m;
        v[1] += Aq * dfa;
        v[2] += Aq * dfb;
        v[3] += Aq * dfc;
      }
    }
  }
  reduce7<NT>(v, red, res);
}

// Adjoint of the binned P1D through delta_F = F/<F> - 1 (at fixed <F>) and
// of the FGPA: A_q = dL/d delta_F,q = sum_j gP_j (2 px / N) Re(conj(X_j) W^{jq}).
// Generic kernel, TAB as k_p1d.  Dynamic LDS: x, cj, the Bluestein line
// (TAB = false), the edges.
template <bool TAB>
__global__ __launch_bounds__(FIT_THREADS) void k_p1d_adjoint(AdjArgs a) {
  extern __shared__ double dyn[];
  double* x = dyn;
  double2* const cjh = reinterpret_cast<double2*>(dyn + ((a.nlds + 1) & ~1));
  double2* bline = reinterpret_cast<double2*>(reinterpret_cast<char*>(dyn) + a.line_off);
  double* ledges = reinterpret_cast<double*>(reinterpret_cast<char*>(dyn) + a.e_off);
  int* blo = reinterpret_cast<int*>(ledges + a.nb + 1);
  int* bhi = blo + a.nb;
  double* gco = reinterpret_cast<double*>(bhi + a.nb);
  __shared__ double red[7 * (FIT_THREADS / 64)];
  for (int i = threadIdx.x; i <= a.nb; i += FIT_THREADS) ledges[i] = a.edges[i];
  const smk_skewers& sk = a.sk;
  // one skewer (returns early where the loop over skewers moved on)
  auto one = [&](const int s) {
    const int i0 = sk.imin[s], i1 = sk.imax[s];
    const int m = a.regroup;
    const int nq = (i1 - i0) / m;
    const int64_t o = sk.off[s];
    double* res = a.out + (int64_t)s * 7;
    __syncthreads();
    if (nq > a.nmax) {
      if (threadIdx.x < 7) res[threadIdx.x] = 0.0;
      return;
    }
    const int set = sk.pset[s];
    const double fm = a.fmean[set];
    const double px = a.pixel * m;
    const double k_ny = M_PI / px;
    const int nj = nq > 0 ? nq / 2 + 1 : 0;
    const double* gps = a.gp + (int64_t)set * a.nb;
    if (nq > a.nlds) {
      // longer than the LDS layout (the modes saved by the forward are
      // required): A_q by Horner's rule over j with c_j = g_b(j) conj(X_j)
      // staged through LDS in pieces; each thread folds its own A_q into the
      // FGPA sums of the pixels of q, then all pixels give the dF sums
      const double2* xs = a.xj + a.xoff[s];
      double2* cbuf = cjh;
      for (int b = threadIdx.x; b < a.nb; b += FIT_THREADS) {
        bin_span(ledges, a.nb, b, nj, nq, k_ny, blo[b], bhi[b]);
        gco[b] = gps[b] * 2.0 * px / nq;
      }
      auto stage = [&](int n0, int n1) {
        for (int j = n0 + (int)threadIdx.x; j < n1; j += FIT_THREADS) {
          // the bin of mode j: the last b with blo[b] <= j, if j < bhi[b]
          int lo = 0, hi = a.nb - 1;
          while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (blo[mid] <= j) lo = mid; else hi = mid - 1;
          }
          double2 c = make_double2(0.0, 0.0);
          if (blo[lo] <= j && j < bhi[lo]) {
            const double g = gco[lo];
            const double2 X = xs[j];
            c = make_double2(g * X.x, -g * X.y);   // g conj(X_j)
          }
          cbuf[j - n0] = c;
        }
      };
      const double4 p = prow(a.abcg, a.rows, set);
      const double inv = 1.0 / fm;
      double v[7] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
      for (int fb = 0; fb < nq; fb += 4 * FIT_THREADS) {
        double re[4], im[4];
        horner_staged(stage, cbuf, nj, nq, fb, nq, re, im);
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int q = fb + (int)threadIdx.x + r * FIT_THREADS;
          if (q >= nq) continue;
          const double A = re[r];
          double acc = 0.0;
          for (int k = 0; k < m; ++k) acc += a.flux[o + i0 + (int64_t)q * m + k];
          const double xq = (acc / m) * inv - 1.0;
          v[0] += A * (xq + 1.0) * fm;
          for (int k = 0; k < m; ++k) {
            const DF df = fgpa_df(a, o + i0 + (int64_t)q * m + k, p);
            v[1] += A / m * df.a;
            v[2] += A / m * df.b;
            v[3] += A / m * df.c;
          }
        }
      }
      for (int i = threadIdx.x; i < i1 - i0; i += FIT_THREADS) {
        const DF df = fgpa_df(a, o + i0 + i, p);
        v[4] += df.a; v[5] += df.b; v[6] += df.c;
      }
      reduce7<FIT_THREADS>(v, red, res);
      return;
    }
    load_delta_f(a.flux, o, i0, i1, m, 1.0 / fm, x);
    __syncthreads();
    auto coef = [&](int j) -> double {
      const int b = find_bin(ledges, a.nb, kfreq(j, nq, k_ny));
      return b >= 0 ? gps[b] * 2.0 * px / nq : 0.0;
    };
    const bool blue = !TAB && nq > 0 && nq <= BLUE_NMAX;
    double2* const cj = blue ? reinterpret_cast<double2*>(reinterpret_cast<char*>(dyn) + a.cjb_off) : cjh;
    if (a.xj) {
      const double2* xs = a.xj + a.xoff[s];
      for (int j = threadIdx.x; j < nj; j += FIT_THREADS) {
        const double g = coef(j);
        const double2 X = xs[j];
        cj[j] = make_double2(g * X.x, -g * X.y);   // g conj(X_j)
      }
    } else if (blue) {
      if constexpr (!TAB)
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
      if constexpr (!TAB)
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
    adj_fgpa<FIT_THREADS>(a, set, a.sk.off[s] + i0, i1 - i0, nq, saf, x, red, res);
  };
  if constexpr (TAB) {
    __shared__ uint64_t need[FIT_THREADS / 64];
    for (int c0 = blockIdx.x; c0 < sk.n; c0 += gridDim.x * FIT_THREADS) {
      rest_masks(sk, a.regroup, a.tables, a.tab_off, a.nmax, c0, need);
      for (int w = 0; w < FIT_THREADS / 64; ++w)
        for (uint64_t mk = need[w]; mk; mk &= mk - 1) one(c0 + (w * 64 + __builtin_ctzll(mk)) * (int)gridDim.x);
    }
  } else {
    for (int s = blockIdx.x; s < sk.n; s += gridDim.x) one(s);
  }
}

// LDS of k_p1d_adjoint_blue<M>: x (M / 2 doubles), the line with cj over its
// start (cj is read into registers before the line's first exchange), the
// edges
__host__ __device__ constexpr size_t adj_blue_x_bytes(int M) { return (size_t)(M / 2) * 8; }

// The adjoint of the skewers with a table and a forest of M's size class,
// M / 8 threads of 8 points each.
template <int M>
__global__ __launch_bounds__(M / 8) __attribute__((amdgpu_waves_per_eu(2))) void k_p1d_adjoint_blue(
    AdjArgs a) {
  constexpr int NT = M / 8;
  extern __shared__ double dyn[];
  double* x = dyn;
  double2* line = reinterpret_cast<double2*>(reinterpret_cast<char*>(dyn) + adj_blue_x_bytes(M));
  double2* cj = line;
  double* ledges =
      reinterpret_cast<double*>(reinterpret_cast<char*>(dyn) + adj_blue_x_bytes(M) + blue_line_bytes_m(M));
  __shared__ double red[7 * (NT / 64)];
  for (int i = threadIdx.x; i <= a.nb; i += NT) ledges[i] = a.edges[i];
  const smk_skewers& sk = a.sk;
  for (int s = blockIdx.x; s < sk.n; s += gridDim.x) {
    const int i0 = sk.imin[s], i1 = sk.imax[s];
    const int m = a.regroup;
    const int nq = (i1 - i0) / m;
    if (!uses_table(a.tables, a.tab_off, s, nq, a.nmax) || blue_m_of(nq) != M) continue;
    const int64_t o = sk.off[s];
    double* res = a.out + (int64_t)s * 7;
    __syncthreads();
    const int set = sk.pset[s];
    const double fm = a.fmean[set];
    load_delta_f<NT>(a.flux, o, i0, i1, m, 1.0 / fm, x);
    __syncthreads();
    const double px = a.pixel * m;
    const double k_ny = M_PI / px;
    const int nj = nq / 2 + 1;
    const double* gps = a.gp + (int64_t)set * a.nb;
    const double2* T = a.tables + a.tab_off[s];
    auto coef = [&](int j) -> double {
      const int b = find_bin(ledges, a.nb, kfreq(j, nq, k_ny));
      return b >= 0 ? gps[b] * 2.0 * px / nq : 0.0;
    };
    if (a.xj) {
      const double2* xs = a.xj + a.xoff[s];
      for (int j = threadIdx.x; j < nj; j += NT) {
        const double g = coef(j);
        const double2 X = xs[j];
        cj[j] = make_double2(g * X.x, -g * X.y);   // g conj(X_j)
      }
    } else {
      f64::bluestein_tab<M, 8>(nq, nj, [&](int n) { return make_double2(x[n], 0.0); },
                               [&](int j, double2 X) {
                                 const double g = coef(j);
                                 cj[j] = make_double2(g * X.x, -g * X.y);
                               },
                               line, a.tw, BLUE_MMAX, T, T + nq);
    }
    __syncthreads();
    double saf = 0.0;
    f64::bluestein_tab<M, 8>(nq, nq, [&](int n) { return n < nj ? cj[n] : make_double2(0.0, 0.0); },
                             [&](int q, double2 X) {
                               saf += X.x * (x[q] + 1.0) * fm;
                               x[q] = X.x;
                             },
                             line, a.tw, BLUE_MMAX, T, T + nq);
    __syncthreads();
    adj_fgpa<NT>(a, set, a.sk.off[s] + i0, i1 - i0, nq, saf, x, red, res);
  }
}

// The adjoint of two forests of one length with one Bluestein DFT per
// transform.  A_q = Re sum_{j < N/2+1} c_j W^{jq} is the DFT of the
// Hermitian spectrum H (H_0 = Re c_0, H_j = c_j / 2 and H_{N-j} = conj(c_j)
// / 2 for 0 < j < N/2, H_{N/2} = Re c_{N/2} for even N), a real sequence, so
// the DFT of H1 + i H2 is A1 + i A2.  Without saved modes, X1 and X2 come from
// one DFT of delta_F1 + i delta_F2 as in k_p1d_blue_pair.  LDS: x1, x2 (M/2
// doubles each), the line (c1 and c2 over its start), the edges.
template <int M>
__global__ __launch_bounds__(M / 8) __attribute__((amdgpu_waves_per_eu(2))) void k_p1d_adjoint_blue_pair(
    AdjArgs a) {
  constexpr int NT = M / 8;
  extern __shared__ double dyn[];
  double* x1 = dyn;
  double* x2 = dyn + M / 2;
  double2* line = reinterpret_cast<double2*>(reinterpret_cast<char*>(dyn) + 2 * adj_blue_x_bytes(M));
  double* ledges = reinterpret_cast<double*>(reinterpret_cast<char*>(dyn) + 2 * adj_blue_x_bytes(M) +
                                             blue_line_bytes_m(M));
  __shared__ double red[7 * (NT / 64)];
  for (int i = threadIdx.x; i <= a.nb; i += NT) ledges[i] = a.edges[i];
  const smk_skewers& sk = a.sk;
  const int m = a.regroup;
  const double px = a.pixel * m;
  const double k_ny = M_PI / px;
  // persistent, as k_p1d_blue_pair: the members' descriptors a pair ahead,
  // the pair list entry two pairs ahead; members that no longer take this
  // kernel are left out
  const int G = gridDim.x;
  int w = a.pair_lo + (int)blockIdx.x;
  PairDesc cur = w < a.pair_hi ? pair_desc(a, a.pairs[2 * w], a.pairs[2 * w + 1]) : pair_desc(a, -1, -1);
  int q1 = -1, q2 = -1;
  if (w + G < a.pair_hi) {
    q1 = a.pairs[2 * (w + G)];
    q2 = a.pairs[2 * (w + G) + 1];
  }
  for (; w < a.pair_hi; w += G) {
    const PairDesc nxt = pair_desc(a, q1, q2);
    q1 = q2 = -1;
    if (w + 2 * G < a.pair_hi) {
      q1 = a.pairs[2 * (w + 2 * G)];
      q2 = a.pairs[2 * (w + 2 * G) + 1];
    }
    int s1, s2;
    bool swap;
    desc_members<M>(a, cur, s1, s2, swap);
    const int64_t o1 = swap ? cur.o2 : cur.o1, o2 = cur.o2;
    const int nf1 = swap ? cur.n2 : cur.n1, nf2 = cur.n2;
    const int set1 = swap ? cur.set2 : cur.set1;
    const int set2 = s2 >= 0 ? cur.set2 : set1;
    const int64_t tb = swap ? cur.t2 : cur.t1;
    cur = nxt;
    if (s1 < 0) continue;
    const bool two = s2 >= 0;
    const int nq = nf1 / m;
    const int nj = nq / 2 + 1;
    const double fm1 = a.fmean[set1], fm2 = a.fmean[set2];
    __syncthreads();
    load_delta_f<NT>(a.flux, o1, 0, nf1, m, 1.0 / fm1, x1);
    if (two) load_delta_f<NT>(a.flux, o2, 0, nf2, m, 1.0 / fm2, x2);
    __syncthreads();
    const double* gp1 = a.gp + (int64_t)set1 * a.nb;
    const double* gp2 = a.gp + (int64_t)set2 * a.nb;
    const double2* T = a.tables + tb;
    double2* c1 = line;        // c_j = g_b(j) 2 px / N conj(X_j), j < nj
    double2* c2 = line + nj;
    auto coef = [&](const double* gps, int b) { return b >= 0 ? gps[b] * 2.0 * px / nq : 0.0; };
    if (a.xj) {
      const double2* xs1 = a.xj + a.xoff[s1];
      const double2* xs2 = two ? a.xj + a.xoff[s2] : nullptr;
      for (int j = threadIdx.x; j < nj; j += NT) {
        const int b = find_bin(ledges, a.nb, kfreq(j, nq, k_ny));
        const double g1 = coef(gp1, b);
        const double2 X1 = xs1[j];
        c1[j] = make_double2(g1 * X1.x, -g1 * X1.y);
        double2 cc = make_double2(0.0, 0.0);
        if (two) {
          const double g2 = coef(gp2, b);
          const double2 X2 = xs2[j];
          cc = make_double2(g2 * X2.x, -g2 * X2.y);
        }
        c2[j] = cc;
      }
    } else {
      f64::bluestein_tab<M, 8>(nq, nq, [&](int n) { return make_double2(x1[n], two ? x2[n] : 0.0); },
                               [&](int j, double2 Z) { line[j] = Z; }, line, a.tw, BLUE_MMAX, T, T + nq);
      __syncthreads();
      // the split reads line[j] and line[N - j]: into registers first, then
      // the block writes c1, c2 over the line's start
      double2 cr1[(M / 4 + 1 + NT - 1) / NT], cr2[(M / 4 + 1 + NT - 1) / NT];
#pragma unroll
      for (int t = 0; t < (M / 4 + 1 + NT - 1) / NT; ++t) {
        const int j = (int)threadIdx.x + t * NT;
        cr1[t] = cr2[t] = make_double2(0.0, 0.0);
        if (j < nj) {
          double2 X1, X2;
          split_pair(line[j], line[j == 0 ? 0 : nq - j], X1, X2);
          const int b = find_bin(ledges, a.nb, kfreq(j, nq, k_ny));
          const double g1 = coef(gp1, b), g2 = two ? coef(gp2, b) : 0.0;
          cr1[t] = make_double2(g1 * X1.x, -g1 * X1.y);
          cr2[t] = make_double2(g2 * X2.x, -g2 * X2.y);
        }
      }
      __syncthreads();
#pragma unroll
      for (int t = 0; t < (M / 4 + 1 + NT - 1) / NT; ++t) {
        const int j = (int)threadIdx.x + t * NT;
        if (j < nj) {
          c1[j] = cr1[t];
          c2[j] = cr2[t];
        }
      }
    }
    __syncthreads();
    // u_n = H1_n + i H2_n, n < N (the Hermitian completion above)
    auto herm = [&](const double2* c, int n) -> double2 {
      if (n == 0) return make_double2(c[0].x, 0.0);
      if (2 * n == nq) return make_double2(c[n].x, 0.0);
      if (2 * n < nq) return make_double2(0.5 * c[n].x, 0.5 * c[n].y);
      const double2 v = c[nq - n];
      return make_double2(0.5 * v.x, -0.5 * v.y);
    };
    double saf1 = 0.0, saf2 = 0.0;
    f64::bluestein_tab<M, 8, f64::KEEP_B | f64::KEEP_W>(
        nq, nq,
        [&](int n) {
          const double2 h1 = herm(c1, n), h2 = herm(c2, n);
          return make_double2(h1.x - h2.y, h1.y + h2.x);   // H1 + i H2
        },
        [&](int q, double2 Z) {
          saf1 += Z.x * (x1[q] + 1.0) * fm1;
          x1[q] = Z.x;
          if (two) {
            saf2 += Z.y * (x2[q] + 1.0) * fm2;
            x2[q] = Z.y;
          }
        },
        line, a.tw, BLUE_MMAX, T, T + nq);
    __syncthreads();
    adj_fgpa<NT>(a, set1, o1, nf1, nq, saf1, x1, red, a.out + (int64_t)s1 * 7);
    if (two) adj_fgpa<NT>(a, set2, o2, nf2, nq, saf2, x2, red, a.out + (int64_t)s2 * 7);
  }
}

// ------------------------------------------------------------- Jacobian
// Forward-mode derivatives of the fit's observables, for Gauss-Newton: with
// theta = (a, b, c) of the skewer's set, F' the regrouped flux and
// delta_F = F'/<F> - 1, the tangents t_k = d delta_F / d theta_k =
// dF'_k / <F> - F' d<F>_k / <F>^2 and dP_j / d theta_k =
// 2 px / N Re(conj(X_j) T_kj) with T_k = DFT(t_k); per bin, the sums of
// dP_j / d theta_k beside the forward's (count, sum k, sum P, sum P^2).

struct JacArgs {
  smk_skewers sk;
  const float* delta;
  const float* eta;
  const double* abcg;
  int64_t rows;
  const double* flux;
  const double* fmean;   // per set
  const double* dfmean;  // [sets][3]: d<F>/d(a, b, c)
  double pixel;
  int regroup;
  const double* edges;
  int nb;
  double* partial;       // [n][nb][7]: count, sum k, sum P, sum P^2, sum dP/d(a, b, c)
  int32_t* status;       // per skewer: 1 if the forest is longer than nmax
  int nmax;
  const double2* tw;
  const double2* tables;
  const int64_t* tab_off;
  const int32_t* pairs;  // NULL, or the pair list of smk_p1d_binned (the forests of each size class)
  int pair_lo, pair_hi;  // the pairs of the launch's transform size
};

// F and dF/d(a, b, c) summed over the m pixels of regrouped point q (the
// expressions of k_fgpa_flux and fgpa_df, so F is the same bits)
struct JacPoint { double f, da, db, dc; };
__device__ __forceinline__ JacPoint jac_point(const JacArgs& a, int64_t o, int q, int m, double4 p) {
  JacPoint r{0.0, 0.0, 0.0, 0.0};
  // without eta the load reads delta and is not used, so no load waits on
  // a branch
  const float* ep = a.eta ? a.eta : a.delta;
  for (int k = 0; k < m; ++k) {
    const int64_t at = o + (int64_t)q * m + k;
    const float ev = ep[at];
    const double e = a.eta ? (double)ev : 0.0;
    const double gf = (double)a.delta[at] + p.z * e;
    const double T = exp(p.y * p.w * gf);
    const double F = a.flux[at];
    r.f += F;
    r.da += -T * F;
    r.db += -p.x * T * F * p.w * gf;
    r.dc += -p.x * T * F * p.y * p.w * e;
  }
  return r;
}

// one skewer's flux (k_fgpa_flux's expression) and its forest sums of F
// and of dF/d(a, b, c) at the parameters of its set, on one wave
template <bool ETA>
__device__ __forceinline__ void flux_jac_skewer(const smk_skewers& sk, int s, const float* __restrict__ delta,
                                                const float* __restrict__ eta, const double* __restrict__ abcg,
                                                int64_t rows, double* __restrict__ flux, int lane, double (&v)[4]) {
  const int64_t o = sk.off[s];
  const int n = sk.npix[s], i0 = sk.imin[s], i1 = sk.imax[s];
  const double4 p = prow(abcg, rows, sk.pset[s]);
  for (int i = lane; i < n; i += 64 * FLUX_U) {
    float dv[FLUX_U], ev[FLUX_U];
#pragma unroll
    for (int u = 0; u < FLUX_U; ++u) {
      // past the end: a clamped (in-bounds, unused) load, so no load waits
      // on a branch
      const int j = i + 64 * u < n ? i + 64 * u : n - 1;
      dv[u] = delta[o + j];
      ev[u] = ETA ? eta[o + j] : 0.0f;
    }
#pragma unroll
    for (int u = 0; u < FLUX_U; ++u) {
      const int j = i + 64 * u;
      if (j < n) {
        const double e = ETA ? (double)ev[u] : 0.0;
        const double g = (double)dv[u] + p.z * e;
        const double T = exp(p.y * p.w * g);
        const double f = exp(-p.x * T);
        flux[o + j] = f;
        if (j >= i0 && j < i1) {
          v[0] += f;
          v[1] += -T * f;
          v[2] += -p.x * T * f * p.w * g;
          v[3] += -p.x * T * f * p.y * p.w * e;
        }
      }
    }
  }
}

__global__ __launch_bounds__(FIT_THREADS) void k_fgpa_flux_jac(smk_skewers sk, const float* __restrict__ delta,
                                                               const float* __restrict__ eta,
                                                               const double* __restrict__ abcg, int64_t rows,
                                                               double* __restrict__ flux,
                                                               double* __restrict__ sums) {
  const int lane = threadIdx.x & 63;
  for (int s = blockIdx.x * FLUX_WAVES + flux_wave(); s < sk.n; s += gridDim.x * FLUX_WAVES) {
    double v[4] = {0.0, 0.0, 0.0, 0.0};
    if (eta)
      flux_jac_skewer<true>(sk, s, delta, eta, abcg, rows, flux, lane, v);
    else
      flux_jac_skewer<false>(sk, s, delta, eta, abcg, rows, flux, lane, v);
#pragma unroll
    for (int k = 0; k < 4; ++k) v[k] = wave_sum(v[k]);
    if (lane == 0) {
#pragma unroll
      for (int k = 0; k < 4; ++k) sums[(int64_t)s * 4 + k] = v[k];
    }
  }
}

// bin runs as bin_runs, seven sums per bin: count, sum k, sum P, sum P^2 and
// the sums of the three dP rows (dp + r * ds)
template <int NT>
__device__ __forceinline__ void bin_runs_jac(const double* pk, const double* dp, int ds, const int* bins, int nj,
                                             int nq, double k_ny, int nb, int* bfirst, int* blast, double* part) {
  for (int b = threadIdx.x; b < nb; b += NT) {
    bfirst[b] = 0;
    blast[b] = -1;
  }
  __syncthreads();
  for (int j = threadIdx.x; j < nj; j += NT) {
    const int b = bins[j];
    if (b < 0) continue;
    if (j == 0 || bins[j - 1] != b) bfirst[b] = j;
    if (j == nj - 1 || bins[j + 1] != b) blast[b] = j;
  }
  __syncthreads();
  for (int b = threadIdx.x; b < nb; b += NT) {
    double c = 0.0, sk_ = 0.0, sp = 0.0, sp2 = 0.0, d0 = 0.0, d1 = 0.0, d2 = 0.0;
    for (int j = bfirst[b]; j <= blast[b]; ++j) {
      c += 1.0;
      sk_ += kfreq(j, nq, k_ny);
      sp += pk[j];
      sp2 += pk[j] * pk[j];
      d0 += dp[j];
      d1 += dp[ds + j];
      d2 += dp[2 * ds + j];
    }
    double* o = part + 7 * b;
    o[0] = c; o[1] = sk_; o[2] = sp; o[3] = sp2; o[4] = d0; o[5] = d1; o[6] = d2;
  }
}

// LDS of k_p1d_jac_blue<M>: two lines; over the second line past its first
// M/2 points (the emitted outputs need at most N <= M/2) pk, the three dP
// rows (NJ each, NJ = M/4 + 1 >= N/2 + 1) and the bins; then the edges and
// the bin runs
__host__ __device__ constexpr int jac_nj(int M) { return M / 4 + 1; }
__host__ __device__ constexpr size_t jac_ext_off(int M) { return blue_line_bytes_m(M) + (size_t)(M / 2) * 16; }
__host__ __device__ constexpr size_t jac_blue_fixed_bytes(int M) {
  return jac_ext_off(M) +
         ((size_t)jac_nj(M) * 32 + (size_t)((jac_nj(M) + 1) & ~1) * 4 > blue_line_bytes_m(M) - (size_t)(M / 2) * 16
              ? (size_t)jac_nj(M) * 32 + (size_t)((jac_nj(M) + 1) & ~1) * 4
              : blue_line_bytes_m(M) - (size_t)(M / 2) * 16);
}

// The Jacobian of the skewers with a table and a forest of M's size class:
// the Bluestein DFTs of delta_F + i t_c (X and T_c) and of t_a + i t_b (T_a
// and T_b) in lockstep (bluestein_tab2: one set of barriers and twiddles for
// both), each emitting all N outputs into its line, split by the symmetry
// of real inputs
template <int M>
__global__ __launch_bounds__(M / 8) __attribute__((amdgpu_waves_per_eu(2))) void k_p1d_jac_blue(JacArgs a) {
  constexpr int NT = M / 8;
  constexpr int NJ = jac_nj(M);
  extern __shared__ double dyn[];
  double2* la = reinterpret_cast<double2*>(dyn);
  double2* lb = reinterpret_cast<double2*>(reinterpret_cast<char*>(dyn) + blue_line_bytes_m(M));
  double* pk = reinterpret_cast<double*>(reinterpret_cast<char*>(dyn) + jac_ext_off(M));
  double* dp = pk + NJ;
  int* bins = reinterpret_cast<int*>(dp + 3 * NJ);
  double* ledges = reinterpret_cast<double*>(reinterpret_cast<char*>(dyn) + jac_blue_fixed_bytes(M));
  int* bfirst = reinterpret_cast<int*>(ledges + a.nb + 1);
  int* blast = bfirst + a.nb;
  for (int i = threadIdx.x; i <= a.nb; i += NT) ledges[i] = a.edges[i];
  const smk_skewers& sk = a.sk;
  // M = 4096: persistent, each block checks its candidate skewers for this
  // size class with one ballot per wave, then transforms the ones it takes
  auto one = [&](const int s) {
    const int i0 = sk.imin[s], i1 = sk.imax[s];
    const int m = a.regroup;
    const int nq = (i1 - i0) / m;
    __syncthreads();
    const int set = sk.pset[s];
    const double inv = 1.0 / a.fmean[set];
    const double g0 = a.dfmean[3 * set], g1 = a.dfmean[3 * set + 1], g2 = a.dfmean[3 * set + 2];
    const double4 p = prow(a.abcg, a.rows, set);
    const double px = a.pixel * m;
    const double k_ny = M_PI / px;
    const int nj = nq / 2 + 1;
    const double c2 = 2.0 * px / nq;
    const double2* T = a.tables + a.tab_off[s];
    const int64_t o = sk.off[s] + i0;
    auto tangent = [&](double dsum, double fsum, double g) { return (dsum / m) * inv - (fsum / m) * g * inv * inv; };
    f64::bluestein_tab2<M, 8, true, (M <= 2048 ? f64::KEEP_B : 0)>(
        nq, nq,
        [&](int q) {
          const JacPoint P = jac_point(a, o, q, m, p);
          return make_double4((P.f / m) * inv - 1.0, tangent(P.dc, P.f, g2), tangent(P.da, P.f, g0),
                              tangent(P.db, P.f, g1));
        },
        [&](int j, double2 Z) { la[j] = Z; }, [&](int j, double2 Z) { lb[j] = Z; }, la, lb, a.tw, BLUE_MMAX, T,
        T + nq, T, T + nq);
    __syncthreads();
    for (int j = threadIdx.x; j < nj; j += NT) {
      const int jn = j == 0 ? 0 : nq - j;
      double2 X, Tc, Ta, Tb;
      split_pair(la[j], la[jn], X, Tc);
      split_pair(lb[j], lb[jn], Ta, Tb);
      pk[j] = (X.x * X.x + X.y * X.y) * px / nq;
      dp[j] = c2 * (X.x * Ta.x + X.y * Ta.y);
      dp[NJ + j] = c2 * (X.x * Tb.x + X.y * Tb.y);
      dp[2 * NJ + j] = c2 * (X.x * Tc.x + X.y * Tc.y);
      bins[j] = find_bin(ledges, a.nb, kfreq(j, nq, k_ny));
    }
    __syncthreads();
    bin_runs_jac<NT>(pk, dp, NJ, bins, nj, nq, k_ny, a.nb, bfirst, blast, a.partial + (int64_t)s * a.nb * 7);
  };
  auto takes = [&](int s, int nq) { return uses_table(a.tables, a.tab_off, s, nq, a.nmax) && blue_m_of(nq) == M; };
  if (a.pairs) {
    // the forests of this size class from the pair list (two entries a
    // pair, -1: none), each checked as above
    for (int e = 2 * a.pair_lo + (int)blockIdx.x; e < 2 * a.pair_hi; e += gridDim.x) {
      const int s = a.pairs[e];
      if (s >= 0 && takes(s, (sk.imax[s] - sk.imin[s]) / a.regroup)) one(s);
    }
  } else {
    if constexpr (M >= 4096) {
      __shared__ uint64_t need[NT / 64];
      for (int c0 = blockIdx.x; c0 < sk.n; c0 += gridDim.x * NT) {
        cand_masks<NT>(sk, a.regroup, c0, need, takes);
        for (int w = 0; w < NT / 64; ++w)
          for (uint64_t mk = need[w]; mk; mk &= mk - 1) one(c0 + (w * 64 + __builtin_ctzll(mk)) * (int)gridDim.x);
      }
    } else {
      for (int s = blockIdx.x; s < sk.n; s += gridDim.x)
        if (takes(s, (sk.imax[s] - sk.imin[s]) / a.regroup)) one(s);
    }
  }
}

// Generic Jacobian kernel (any N; TAB as k_p1d: with tables, the skewers
// without a usable one): the four real series delta_F, t_a, t_b, t_c staged
// through LDS in pieces of JAC_XC points, their DFTs by Horner's rule for
// 4 FIT_THREADS modes per pass (four accumulators per mode sharing the
// twiddle), each bin's seven running sums in LDS, added in ascending j.
constexpr int JAC_XC = 1024;
constexpr int JAC_CH = 4 * FIT_THREADS;
size_t jac_generic_lds(int nb) {
  return (size_t)JAC_XC * 32 + (size_t)JAC_CH * 32 + (size_t)(nb + 1) * 8 + (size_t)nb * 8 + (size_t)nb * 56;
}

template <bool TAB>
__global__ __launch_bounds__(FIT_THREADS) void k_p1d_jac(JacArgs a) {
  extern __shared__ double dyn[];
  double2* ya = reinterpret_cast<double2*>(dyn);  // (delta_F, t_c)
  double2* yb = ya + JAC_XC;                      // (t_a, t_b)
  double* vals = reinterpret_cast<double*>(yb + JAC_XC);  // pk, dP_a, dP_b, dP_c rows of JAC_CH
  double* ledges = vals + 4 * JAC_CH;
  int* blo = reinterpret_cast<int*>(ledges + a.nb + 1);
  int* bhi = blo + a.nb;
  double* bsum = reinterpret_cast<double*>(bhi + a.nb);
  for (int i = threadIdx.x; i <= a.nb; i += FIT_THREADS) ledges[i] = a.edges[i];
  const smk_skewers& sk = a.sk;
  // one skewer (returns early where the loop over skewers moved on)
  auto one = [&](const int s) {
    const int i0 = sk.imin[s], i1 = sk.imax[s];
    const int m = a.regroup;
    const int nq = (i1 - i0) / m;
    double* part = a.partial + (int64_t)s * a.nb * 7;
    __syncthreads();
    if (nq > a.nmax) {
      for (int b = threadIdx.x; b < 7 * a.nb; b += FIT_THREADS) part[b] = 0.0;
      if (threadIdx.x == 0 && a.status) a.status[s] = 1;
      return;
    }
    const int set = sk.pset[s];
    const double inv = 1.0 / a.fmean[set];
    const double g0 = a.dfmean[3 * set], g1 = a.dfmean[3 * set + 1], g2 = a.dfmean[3 * set + 2];
    const double4 p = prow(a.abcg, a.rows, set);
    const double px = a.pixel * m;
    const double k_ny = M_PI / px;
    const int nj = nq > 0 ? nq / 2 + 1 : 0;
    const double c2 = nq > 0 ? 2.0 * px / nq : 0.0;
    const int64_t o = sk.off[s] + i0;
    for (int b = threadIdx.x; b < a.nb; b += FIT_THREADS) {
      bin_span(ledges, a.nb, b, nj, nq, k_ny, blo[b], bhi[b]);
#pragma unroll
      for (int r = 0; r < 7; ++r) bsum[7 * b + r] = 0.0;
    }
    auto tangent = [&](double dsum, double fsum, double g) { return (dsum / m) * inv - (fsum / m) * g * inv * inv; };
    for (int j0 = 0; j0 < nj; j0 += JAC_CH) {
      const int j1 = j0 + JAC_CH < nj ? j0 + JAC_CH : nj;
      double c[4], sn[4], re[4][4], im[4][4];
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int f = j0 + (int)threadIdx.x + r * FIT_THREADS;
        if (f < j1) {
          sincospi(-2.0 * (double)f / (double)nq, &sn[r], &c[r]);
        } else {
          sn[r] = 0.0;
          c[r] = 1.0;
        }
#pragma unroll
        for (int k = 0; k < 4; ++k) re[r][k] = im[r][k] = 0.0;
      }
      // Horner over n descending, the input staged in pieces
      for (int nh = nq; nh > 0; nh -= JAC_XC) {
        const int nl = nh > JAC_XC ? nh - JAC_XC : 0;
        __syncthreads();
        for (int q = nl + (int)threadIdx.x; q < nh; q += FIT_THREADS) {
          const JacPoint P = jac_point(a, o, q, m, p);
          ya[q - nl] = make_double2((P.f / m) * inv - 1.0, tangent(P.dc, P.f, g2));
          yb[q - nl] = make_double2(tangent(P.da, P.f, g0), tangent(P.db, P.f, g1));
        }
        __syncthreads();
        for (int n = nh - 1; n >= nl; --n) {
          const double2 u = ya[n - nl], w = yb[n - nl];
          const double y[4] = {u.x, w.x, w.y, u.y};   // delta_F, t_a, t_b, t_c
#pragma unroll
          for (int r = 0; r < 4; ++r) {
#pragma unroll
            for (int k = 0; k < 4; ++k) {
              const double r2 = fma(re[r][k], c[r], fma(-im[r][k], sn[r], y[k]));
              im[r][k] = fma(re[r][k], sn[r], im[r][k] * c[r]);
              re[r][k] = r2;
            }
          }
        }
      }
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int j = j0 + (int)threadIdx.x + r * FIT_THREADS;
        if (j < j1) {
          const double xr = re[r][0], xi = im[r][0];
          vals[j - j0] = (xr * xr + xi * xi) * px / nq;
#pragma unroll
          for (int k = 1; k < 4; ++k) vals[k * JAC_CH + j - j0] = c2 * (xr * re[r][k] + xi * im[r][k]);
        }
      }
      __syncthreads();
      for (int b = threadIdx.x; b < a.nb; b += FIT_THREADS) {
        const int lo = blo[b] > j0 ? blo[b] : j0, hi = bhi[b] < j1 ? bhi[b] : j1;
        double* t = bsum + 7 * b;
        double cnt = t[0], sk_ = t[1], sp = t[2], sp2 = t[3], d0 = t[4], d1 = t[5], d2 = t[6];
        for (int j = lo; j < hi; ++j) {
          const double pj = vals[j - j0];
          cnt += 1.0;
          sk_ += kfreq(j, nq, k_ny);
          sp += pj;
          sp2 += pj * pj;
          d0 += vals[JAC_CH + j - j0];
          d1 += vals[2 * JAC_CH + j - j0];
          d2 += vals[3 * JAC_CH + j - j0];
        }
        t[0] = cnt; t[1] = sk_; t[2] = sp; t[3] = sp2; t[4] = d0; t[5] = d1; t[6] = d2;
      }
    }
    __syncthreads();
    for (int b = threadIdx.x; b < 7 * a.nb; b += FIT_THREADS) part[b] = bsum[b];
  };
  if constexpr (TAB) {
    __shared__ uint64_t need[FIT_THREADS / 64];
    for (int c0 = blockIdx.x; c0 < sk.n; c0 += gridDim.x * FIT_THREADS) {
      rest_masks(sk, a.regroup, a.tables, a.tab_off, a.nmax, c0, need);
      for (int w = 0; w < FIT_THREADS / 64; ++w)
        for (uint64_t mk = need[w]; mk; mk &= mk - 1) one(c0 + (w * 64 + __builtin_ctzll(mk)) * (int)gridDim.x);
    }
  } else {
    for (int s = blockIdx.x; s < sk.n; s += gridDim.x) one(s);
  }
}

int grid_for(int n) { return n < 65536 ? (n > 0 ? n : 1) : 65536; }
// grid of a generic kernel that runs beside the per-size table kernels: its
// blocks walk the batch and take only the forests without a table, usually
// few, so a block per forest would spend the launch on blocks that exit at
// once; 2048 blocks still spread many such forests over every CU
int grid_for_rest(int n) { const int g = grid_for(n); return g < 2048 ? g : 2048; }

// propagate a nonzero SMK_ERR_* return
#define SMK_CALL(expr)          \
  do {                          \
    const int smk_rc_ = (expr); \
    if (smk_rc_) return smk_rc_; \
  } while (0)

// blocks of a persistent kernel: as many as fit on the device at once (its
// occupancy at this block size and LDS), at most n
int resident_grid(const void* kern, int threads, size_t lds, int n) {
  int dev = 0, cus = 0, per_cu = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
      hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, threads, lds) != hipSuccess || cus < 1 ||
      per_cu < 1)
    return grid_for(n);
  const int64_t g = (int64_t)cus * per_cu;
  return n < g ? (n > 0 ? n : 1) : (int)g;
}

template <int M>
int launch_p1d_blue_pair(P1dArgs a, int lo, int hi, size_t tail, hipStream_t st) {
  if (hi <= lo) return SMK_OK;
  a.pair_lo = lo;
  a.pair_hi = hi;
  const size_t lds = pair_fixed_bytes(M) + tail;
  SMK_HIP_CHECK(hipFuncSetAttribute((const void*)k_p1d_blue_pair<M>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  const int grid = resident_grid((const void*)k_p1d_blue_pair<M>, M / 8, lds, hi - lo);
  hipLaunchKernelGGL(k_p1d_blue_pair<M>, dim3(grid), dim3(M / 8), lds, st, a);
  SMK_LAUNCH_CHECK();
  return SMK_OK;
}

template <int M>
int launch_p1d_blue_quad(P1dArgs a, int lo, int hi, size_t tail, hipStream_t st) {
  if (hi <= lo) return SMK_OK;
  a.pair_lo = lo;
  a.pair_hi = hi;
  const size_t lds = quad_fixed_bytes(M) + tail;
  SMK_HIP_CHECK(hipFuncSetAttribute((const void*)k_p1d_blue_quad<M>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  hipLaunchKernelGGL(k_p1d_blue_quad<M>, dim3(grid_for((hi - lo + 1) / 2)), dim3(M / 8), lds, st, a);
  SMK_LAUNCH_CHECK();
  return SMK_OK;
}

template <int M>
int launch_p1d_blue(const P1dArgs& a, dim3 grid, size_t tail, hipStream_t st) {
  const size_t lds = blue_line_bytes_m(M) + tail;
  SMK_HIP_CHECK(hipFuncSetAttribute((const void*)k_p1d_blue<M>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                    (int)lds));
  hipLaunchKernelGGL(k_p1d_blue<M>, grid, dim3(M / 8), lds, st, a);
  SMK_LAUNCH_CHECK();
  return SMK_OK;
}

template <int M>
int launch_adj_blue(const AdjArgs& a, dim3 grid, size_t tail, hipStream_t st) {
  const size_t lds = adj_blue_x_bytes(M) + blue_line_bytes_m(M) + tail;
  SMK_HIP_CHECK(hipFuncSetAttribute((const void*)k_p1d_adjoint_blue<M>,
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  hipLaunchKernelGGL(k_p1d_adjoint_blue<M>, grid, dim3(M / 8), lds, st, a);
  SMK_LAUNCH_CHECK();
  return SMK_OK;
}

template <int M>
int launch_adj_blue_pair(AdjArgs a, int lo, int hi, size_t tail, hipStream_t st) {
  if (hi <= lo) return SMK_OK;
  a.pair_lo = lo;
  a.pair_hi = hi;
  const size_t lds = 2 * adj_blue_x_bytes(M) + blue_line_bytes_m(M) + tail;
  SMK_HIP_CHECK(hipFuncSetAttribute((const void*)k_p1d_adjoint_blue_pair<M>,
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  const int grid = resident_grid((const void*)k_p1d_adjoint_blue_pair<M>, M / 8, lds, hi - lo);
  hipLaunchKernelGGL(k_p1d_adjoint_blue_pair<M>, dim3(grid), dim3(M / 8), lds, st, a);
  SMK_LAUNCH_CHECK();
  return SMK_OK;
}

template <int M>
int launch_p1d_jac_blue(JacArgs a, dim3 grid, int lo, int hi, size_t tail, hipStream_t st) {
  const size_t lds = jac_blue_fixed_bytes(M) + tail;
  if (a.pairs) {
    // a block per pair list entry of the class (20 000 forests: 83, 366 and
    // 692 us at M = 1024, 2048, 4096, against 102, 423 and 812 us scanning
    // the batch)
    if (hi <= lo) return SMK_OK;
    a.pair_lo = lo;
    a.pair_hi = hi;
    grid = dim3(grid_for(2 * (hi - lo)));
  }
  SMK_HIP_CHECK(hipFuncSetAttribute((const void*)k_p1d_jac_blue<M>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  // M = 4096 runs one block per CU: resident blocks, each checking its
  // candidates at once, save a block dispatch per skewer (976 -> 814 us for
  // 20 000 forests).  At M <= 2048 several blocks share a CU and a block per
  // skewer balances better (resident: 125 against 103 us at 1024, 464
  // against 424 us at 2048)
  const int g = M >= 4096 ? resident_grid((const void*)k_p1d_jac_blue<M>, M / 8, lds, (int)grid.x) : (int)grid.x;
  hipLaunchKernelGGL(k_p1d_jac_blue<M>, dim3(g), dim3(M / 8), lds, st, a);
  SMK_LAUNCH_CHECK();
  return SMK_OK;
}

// W_4096^k, k < 4096, FP64, then the FFT stages' tables (f64::stage_tab_off:
// W_(Ns R)^(r jj) as W_4096^(r jj 4096 / (Ns R)), the same values); one
// table per device (built on first use)
const double2* fp64_twiddles(hipError_t* err) {
  static std::mutex mu;
  static const double2* tabs[64] = {};
  int dev = 0;
  *err = hipGetDevice(&dev);
  if (*err != hipSuccess || dev < 0 || dev >= 64) return nullptr;
  std::lock_guard<std::mutex> lk(mu);
  if (tabs[dev] == nullptr) {
    constexpr int NTAB = BLUE_MMAX + f64::stage_tab_len(BLUE_MMAX);
    std::vector<double2> t(NTAB);
    for (int k = 0; k < BLUE_MMAX; ++k) {
      const double ang = -2.0 * M_PI * (double)k / (double)BLUE_MMAX;
      t[k] = make_double2(cos(ang), sin(ang));
    }
    for (int R = 2; R <= 16; R *= 2)
      for (int Ns = 1; Ns * R <= BLUE_MMAX; Ns *= 2)
        for (int r = 0; r < R; ++r)
          for (int jj = 0; jj < Ns; ++jj)
            t[BLUE_MMAX + f64::stage_tab_off(R, Ns, BLUE_MMAX) + r * Ns + jj] =
                t[(int64_t)r * jj * (BLUE_MMAX / (Ns * R))];
    void* d = nullptr;
    *err = hipMalloc(&d, NTAB * sizeof(double2));
    if (*err != hipSuccess) return nullptr;
    *err = hipMemcpy(d, t.data(), NTAB * sizeof(double2), hipMemcpyHostToDevice);
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
  hipLaunchKernelGGL(k_fgpa_flux, dim3(grid_for_waves(sk->n)), dim3(FIT_THREADS), 0, (hipStream_t)stream, *sk, delta,
                     eta, abcg, abcg_rows, flux, fsum);
  SMK_LAUNCH_CHECK();
  return SMK_OK;
}

extern "C" int64_t smk_reduce_work_len(int n, int width, int n_sets) {
  if (n < 0 || width <= 0 || n_sets <= 0) return -1;
  return (int64_t)(n > 0 ? red_chunks(n) : 1) * width * n_sets;
}

extern "C" int smk_reduce_by_set(const smk_skewers* sk, const double* per_skewer, int width, int n_sets,
                                 const double* divisor, double* out, double* work, int64_t work_len,
                                 void* stream) {
  SMK_REQUIRE(sk != nullptr && out != nullptr && sk->pset != nullptr, SMK_ERR_ARG, "NULL argument");
  SMK_REQUIRE(width > 0 && n_sets > 0 && n_sets <= 65535 && (width + 63) / 64 <= 65535, SMK_ERR_ARG,
              "bad width/n_sets");
  SMK_REQUIRE(sk->n >= 0 && red_chunks(sk->n) <= 65535, SMK_ERR_RANGE, "too many skewers for one reduction");
  const int nchunk = sk->n > 0 ? red_chunks(sk->n) : 1;
  SMK_REQUIRE(work != nullptr && work_len >= (int64_t)nchunk * width * n_sets, SMK_ERR_ARG,
              "work needs smk_reduce_work_len doubles");
  SMK_REQUIRE(sk->n == 0 || per_skewer != nullptr, SMK_ERR_ARG, "NULL per_skewer");
  hipStream_t st = (hipStream_t)stream;
  const int ncol = (width + 63) / 64;
  if (sk->n > 0) {
    hipLaunchKernelGGL(k_reduce_chunks, dim3(ncol, nchunk, n_sets), dim3(RED_CHUNK), 0, st, sk->pset, sk->n,
                       per_skewer, width, work);
    SMK_LAUNCH_CHECK();
  }
  // no skewers: no chunks, every sum 0
  hipLaunchKernelGGL(k_reduce_chunk_sums, dim3(ncol, n_sets), dim3(256), 0, st, work, sk->n > 0 ? nchunk : 0, width,
                     divisor, out);
  SMK_LAUNCH_CHECK();
  return SMK_OK;
}

extern "C" int64_t smk_p1d_table_len(int n) {
  return (n >= 1 && n <= BLUE_NMAX) ? (int64_t)n + blue_m_of(n) : 0;
}

extern "C" int smk_p1d_tables(const int32_t* lens, const int64_t* offs, int n_lens, double* tables, void* stream) {
  SMK_REQUIRE(n_lens >= 0, SMK_ERR_ARG, "n_lens < 0");
  if (n_lens == 0) return SMK_OK;
  SMK_REQUIRE(lens != nullptr && offs != nullptr && tables != nullptr, SMK_ERR_ARG, "NULL argument");
  hipError_t te = hipSuccess;
  const double2* tw = fp64_twiddles(&te);
  SMK_REQUIRE(tw != nullptr, SMK_ERR_HIP, "FP64 twiddle table: %s", hipGetErrorString(te));
  const size_t lds = (size_t)f64::line_len(BLUE_MMAX) * 16;
  SMK_HIP_CHECK(hipFuncSetAttribute((const void*)k_p1d_tables, hipFuncAttributeMaxDynamicSharedMemorySize,
                                    (int)lds));
  hipLaunchKernelGGL(k_p1d_tables, dim3(grid_for(n_lens)), dim3(FIT_THREADS), lds, (hipStream_t)stream, lens,
                     offs, n_lens, (double2*)tables, tw);
  SMK_LAUNCH_CHECK();
  return SMK_OK;
}

extern "C" int smk_p1d_binned(const smk_skewers* sk, const double* flux, const double* fmean, double pixel,
                              int regroup, const double* edges, int nbins, int max_forest, double* partial,
                              int32_t* status, double* modes, const int64_t* modes_off, const double* tables,
                              const int64_t* tab_off, const int32_t* pairs, const int64_t* pair_off,
                              void* stream) {
  SMK_REQUIRE(sk != nullptr && flux != nullptr && fmean != nullptr && edges != nullptr && partial != nullptr,
              SMK_ERR_ARG, "NULL argument");
  SMK_REQUIRE(sk->pset != nullptr, SMK_ERR_ARG, "the fit kernels need skewers.pset");
  SMK_REQUIRE(regroup >= 1 && nbins >= 1 && pixel > 0, SMK_ERR_ARG, "bad regroup/nbins/pixel");
  SMK_REQUIRE(max_forest >= 0, SMK_ERR_ARG, "max_forest < 0");
  SMK_REQUIRE((tables == nullptr) == (tab_off == nullptr), SMK_ERR_ARG, "tables and tab_off go together");
  if (sk->n <= 0) return SMK_OK;
  int nmax = max_forest / regroup;
  nmax += nmax & 1;
  if (nmax < 2) nmax = 2;
  const int nlds = nmax < FIT_NLDS ? nmax : FIT_NLDS;
  SMK_REQUIRE(modes == nullptr || modes_off != nullptr, SMK_ERR_ARG, "modes need modes_off");
  hipError_t te = hipSuccess;
  const double2* tw = fp64_twiddles(&te);
  SMK_REQUIRE(tw != nullptr, SMK_ERR_HIP, "FP64 twiddle table: %s", hipGetErrorString(te));
  const size_t a_bytes = p1d_a_bytes(nlds);
  const int64_t chunk = ((int64_t)a_bytes - 16 * FIT_XC) / 8;   // modes per pass after the staged input
  SMK_REQUIRE(pairs == nullptr || (tables != nullptr && pair_off != nullptr), SMK_ERR_ARG,
              "pairs need tables and pair_off");
  P1dArgs a{*sk,      flux,      fmean, pixel,   regroup, edges, nbins, partial, status, nmax, (double2*)modes,
            modes_off, tw,       (const double2*)tables, tab_off, a_bytes, nlds, (int)(chunk > 1 ? chunk : 1),
            pairs,     0,        0};
  const dim3 grid(grid_for(sk->n));
  hipStream_t st = (hipStream_t)stream;
  const size_t lds = p1d_lds_bytes(nlds, nbins);
  if (tables) {
    // the forests without a table (Horner), then one launch per Bluestein
    // size class; each skips the skewers of the others
    SMK_HIP_CHECK(hipFuncSetAttribute((const void*)k_p1d<true>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                      (int)lds));
    hipLaunchKernelGGL(k_p1d<true>, dim3(grid_for_rest(sk->n)), dim3(FIT_THREADS), lds, st, a);
    SMK_LAUNCH_CHECK();
    const int nblue = nmax < BLUE_NMAX ? nmax : BLUE_NMAX;
    const size_t tail = (size_t)(nbins + 1) * 8 + (size_t)nbins * 8;
    if (pairs) {
      // pair_off (host): the pairs of M = 1024, 2048, 4096 at [pair_off[c], pair_off[c + 1])
      for (int c = 0; c < 3; ++c)
        SMK_REQUIRE(pair_off[c] >= 0 && pair_off[c] <= pair_off[c + 1] && pair_off[c + 1] <= INT32_MAX,
                    SMK_ERR_ARG, "bad pair_off");
      // two pairs per block in lockstep where that keeps the occupancy
      // (M = 1024: 0.041 against 0.050 ms for 20 000 forests); at M = 2048
      // and 4096 the second register tile costs a wave per SIMD or spills
      // (0.182 against 0.175 ms, 0.395 against 0.364 ms), so one pair
      SMK_CALL(launch_p1d_blue_quad<1024>(a, (int)pair_off[0], (int)pair_off[1], tail, st));
      SMK_CALL(launch_p1d_blue_pair<2048>(a, (int)pair_off[1], (int)pair_off[2], tail, st));
      SMK_CALL(launch_p1d_blue_pair<4096>(a, (int)pair_off[2], (int)pair_off[3], tail, st));
    } else {
      SMK_CALL(launch_p1d_blue<1024>(a, grid, tail, st));
      if (blue_m_of(nblue) >= 2048) SMK_CALL(launch_p1d_blue<2048>(a, grid, tail, st));
      if (blue_m_of(nblue) >= 4096) SMK_CALL(launch_p1d_blue<4096>(a, grid, tail, st));
    }
  } else {
    SMK_HIP_CHECK(hipFuncSetAttribute((const void*)k_p1d<false>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                      (int)lds));
    hipLaunchKernelGGL(k_p1d<false>, grid, dim3(FIT_THREADS), lds, st, a);
    SMK_LAUNCH_CHECK();
  }
  return SMK_OK;
}

extern "C" int smk_p1d_adjoint(const smk_skewers* sk, const float* delta, const float* eta, const double* abcg,
                               int64_t abcg_rows, const double* flux, const double* fmean, double pixel,
                               int regroup, const double* edges, int nbins, int max_forest, const double* modes,
                               const int64_t* modes_off, const double* gp, double* out, const double* tables,
                               const int64_t* tab_off, const int32_t* pairs, const int64_t* pair_off,
                               void* stream) {
  SMK_REQUIRE(sk != nullptr && delta != nullptr && abcg != nullptr && flux != nullptr && fmean != nullptr &&
                  edges != nullptr && gp != nullptr && out != nullptr,
              SMK_ERR_ARG, "NULL argument");
  SMK_REQUIRE(sk->pset != nullptr, SMK_ERR_ARG, "the fit kernels need skewers.pset");
  SMK_REQUIRE(regroup >= 1 && nbins >= 1 && pixel > 0, SMK_ERR_ARG, "bad regroup/nbins/pixel");
  SMK_REQUIRE(max_forest >= 0, SMK_ERR_ARG, "max_forest < 0");
  SMK_REQUIRE((tables == nullptr) == (tab_off == nullptr), SMK_ERR_ARG, "tables and tab_off go together");
  if (sk->n <= 0) return SMK_OK;
  int nmax = max_forest / regroup;
  nmax += nmax & 1;
  if (nmax < 2) nmax = 2;
  const int nlds = nmax < FIT_NLDS ? nmax : FIT_NLDS;
  SMK_REQUIRE(modes == nullptr || modes_off != nullptr, SMK_ERR_ARG, "modes need modes_off");
  SMK_REQUIRE(modes != nullptr || nmax <= FIT_NLDS, SMK_ERR_ARG,
              "forests longer than %d points need the modes saved by smk_p1d_binned", FIT_NLDS);
  hipError_t te = hipSuccess;
  const double2* tw = fp64_twiddles(&te);
  SMK_REQUIRE(tw != nullptr, SMK_ERR_HIP, "FP64 twiddle table: %s", hipGetErrorString(te));
  const bool tab = tables != nullptr;
  const size_t e_off = adj_e_off(nlds, tab);
  const size_t cjb_off = adj_x_bytes(nlds < BLUE_NMAX ? nlds : BLUE_NMAX);
  AdjArgs a{*sk,      delta, eta,          abcg,  abcg_rows, flux, fmean, pixel, regroup,
            edges,    nbins, gp,           out,   nmax,      (const double2*)modes,
            modes_off, tw,   (const double2*)tables, tab_off, adj_line_off(nlds), cjb_off, e_off, nlds,
            pairs,     0,    0};
  SMK_REQUIRE(pairs == nullptr || (tables != nullptr && pair_off != nullptr), SMK_ERR_ARG,
              "pairs need tables and pair_off");
  const dim3 grid(grid_for(sk->n));
  hipStream_t st = (hipStream_t)stream;
  // + the edges, and for forests longer than the layout the bin spans and
  // coefficients (2 nb ints, nb doubles)
  const size_t lds = e_off + (size_t)(nbins + 1) * 8 + (size_t)nbins * 16;
  if (tab) {
    SMK_HIP_CHECK(hipFuncSetAttribute((const void*)k_p1d_adjoint<true>,
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL(k_p1d_adjoint<true>, dim3(grid_for_rest(sk->n)), dim3(FIT_THREADS), lds, st, a);
    SMK_LAUNCH_CHECK();
    const int nblue = nmax < BLUE_NMAX ? nmax : BLUE_NMAX;
    const size_t tail = (size_t)(nbins + 1) * 8;
    if (pairs) {
      for (int c = 0; c < 3; ++c)
        SMK_REQUIRE(pair_off[c] >= 0 && pair_off[c] <= pair_off[c + 1] && pair_off[c + 1] <= INT32_MAX,
                    SMK_ERR_ARG, "bad pair_off");
      SMK_CALL(launch_adj_blue_pair<1024>(a, (int)pair_off[0], (int)pair_off[1], tail, st));
      SMK_CALL(launch_adj_blue_pair<2048>(a, (int)pair_off[1], (int)pair_off[2], tail, st));
      SMK_CALL(launch_adj_blue_pair<4096>(a, (int)pair_off[2], (int)pair_off[3], tail, st));
    } else {
      SMK_CALL(launch_adj_blue<1024>(a, grid, tail, st));
      if (blue_m_of(nblue) >= 2048) SMK_CALL(launch_adj_blue<2048>(a, grid, tail, st));
      if (blue_m_of(nblue) >= 4096) SMK_CALL(launch_adj_blue<4096>(a, grid, tail, st));
    }
  } else {
    SMK_HIP_CHECK(hipFuncSetAttribute((const void*)k_p1d_adjoint<false>,
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL(k_p1d_adjoint<false>, grid, dim3(FIT_THREADS), lds, st, a);
    SMK_LAUNCH_CHECK();
  }
  return SMK_OK;
}

extern "C" int smk_fgpa_flux_jac(const smk_skewers* sk, const float* delta, const float* eta, const double* abcg,
                                 int64_t abcg_rows, double* flux, double* sums, void* stream) {
  SMK_REQUIRE(sk != nullptr && delta != nullptr && abcg != nullptr && flux != nullptr && sums != nullptr,
              SMK_ERR_ARG, "NULL argument");
  SMK_REQUIRE(sk->pset != nullptr, SMK_ERR_ARG, "the fit kernels need skewers.pset");
  if (sk->n <= 0) return SMK_OK;
  hipLaunchKernelGGL(k_fgpa_flux_jac, dim3(grid_for_waves(sk->n)), dim3(FIT_THREADS), 0, (hipStream_t)stream, *sk,
                     delta, eta, abcg, abcg_rows, flux, sums);
  SMK_LAUNCH_CHECK();
  return SMK_OK;
}

extern "C" int smk_p1d_jacobian(const smk_skewers* sk, const float* delta, const float* eta, const double* abcg,
                                int64_t abcg_rows, const double* flux, const double* fmean, const double* dfmean,
                                double pixel, int regroup, const double* edges, int nbins, int max_forest,
                                double* partial, int32_t* status, const double* tables, const int64_t* tab_off,
                                const int32_t* pairs, const int64_t* pair_off, void* stream) {
  SMK_REQUIRE(sk != nullptr && delta != nullptr && abcg != nullptr && flux != nullptr && fmean != nullptr &&
                  dfmean != nullptr && edges != nullptr && partial != nullptr,
              SMK_ERR_ARG, "NULL argument");
  SMK_REQUIRE(pairs == nullptr || (tables != nullptr && pair_off != nullptr), SMK_ERR_ARG,
              "pairs need tables and pair_off");
  SMK_REQUIRE(sk->pset != nullptr, SMK_ERR_ARG, "the fit kernels need skewers.pset");
  SMK_REQUIRE(regroup >= 1 && nbins >= 1 && pixel > 0, SMK_ERR_ARG, "bad regroup/nbins/pixel");
  SMK_REQUIRE(max_forest >= 0, SMK_ERR_ARG, "max_forest < 0");
  SMK_REQUIRE((tables == nullptr) == (tab_off == nullptr), SMK_ERR_ARG, "tables and tab_off go together");
  if (sk->n <= 0) return SMK_OK;
  int nmax = max_forest / regroup;
  nmax += nmax & 1;
  if (nmax < 2) nmax = 2;
  hipError_t te = hipSuccess;
  const double2* tw = fp64_twiddles(&te);
  SMK_REQUIRE(tw != nullptr, SMK_ERR_HIP, "FP64 twiddle table: %s", hipGetErrorString(te));
  JacArgs a{*sk,    delta, eta,    abcg,    abcg_rows, flux, fmean, dfmean, pixel, regroup, edges, nbins,
            partial, status, nmax, tw,      (const double2*)tables, tab_off, pairs, 0, 0};
  int64_t po[4] = {0, 0, 0, 0};
  if (pairs) {
    for (int c = 0; c < 3; ++c)
      SMK_REQUIRE(pair_off[c] >= 0 && pair_off[c] <= pair_off[c + 1] && pair_off[c + 1] <= INT32_MAX / 2,
                  SMK_ERR_ARG, "bad pair_off");
    for (int c = 0; c < 4; ++c) po[c] = pair_off[c];
  }
  const dim3 grid(grid_for(sk->n));
  hipStream_t st = (hipStream_t)stream;
  const size_t lds = jac_generic_lds(nbins);
  if (tables) {
    SMK_HIP_CHECK(hipFuncSetAttribute((const void*)k_p1d_jac<true>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                      (int)lds));
    hipLaunchKernelGGL(k_p1d_jac<true>, dim3(grid_for_rest(sk->n)), dim3(FIT_THREADS), lds, st, a);
    SMK_LAUNCH_CHECK();
    const int nblue = nmax < BLUE_NMAX ? nmax : BLUE_NMAX;
    const size_t tail = (size_t)(nbins + 1) * 8 + (size_t)nbins * 8;
    SMK_CALL(launch_p1d_jac_blue<1024>(a, grid, (int)po[0], (int)po[1], tail, st));
    if (blue_m_of(nblue) >= 2048) SMK_CALL(launch_p1d_jac_blue<2048>(a, grid, (int)po[1], (int)po[2], tail, st));
    if (blue_m_of(nblue) >= 4096) SMK_CALL(launch_p1d_jac_blue<4096>(a, grid, (int)po[2], (int)po[3], tail, st));
  } else {
    SMK_HIP_CHECK(hipFuncSetAttribute((const void*)k_p1d_jac<false>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                      (int)lds));
    hipLaunchKernelGGL(k_p1d_jac<false>, grid, dim3(FIT_THREADS), lds, st, a);
    SMK_LAUNCH_CHECK();
  }
  return SMK_OK;
}
