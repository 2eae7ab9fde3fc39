// Per-mode P(k) tables on gfx950 (interpolate_pk.py).
//
// One thread per mode of the k-space box: k from the float32 axes with the
// reference's float32 operations (correctly rounded, as numpy's), the spline
// segment from a uniform guess table plus a short walk, a Horner cubic in
// FP64, then float32 max(P, 0) / Vcell and its square root.  The output is
// written once, so the kernel is bound by the 4 B per mode it stores.
#include <math.h>

#include "fft_core.h"
#include "smk_internal.h"

using namespace smk;

namespace {

// P(k) of a piecewise-cubic spline: segment from the guess table, then a short walk
__device__ __forceinline__ double spline_eval(const smk_pk_spline& sp, double k) {
  const double g = floor((k - sp.g_lo) / sp.g_step);
  const int s = g < 0.0 ? 0 : (g >= (double)(sp.n_guess - 1) ? sp.n_guess - 1 : (int)g);
  int iv = sp.guess[s];
  while (iv + 1 < sp.n_int && sp.x[iv + 1] <= k) ++iv;
  while (iv > 0 && sp.x[iv] > k) --iv;
  const double t = k - sp.x[iv];
  const double* cc = sp.c + 4 * (int64_t)iv;
  return ((cc[0] * t + cc[1]) * t + cc[2]) * t + cc[3];
}

__global__ __launch_bounds__(256) void k_pk_table(smk_pk_spline sp, const float* __restrict__ kx,
                                                  const float* __restrict__ ky, const float* __restrict__ kz,
                                                  int ny, int nzh, float vcell, float* __restrict__ out,
                                                  uint64_t nmodes) {
  const uint64_t m = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  if (m >= nmodes) return;
  const int l = (int)(m % (uint64_t)nzh);
  const uint64_t rest = m / (uint64_t)nzh;
  const int j = (int)(rest % (uint64_t)ny);
  const int i = (int)(rest / (uint64_t)ny);
  const float a = kx[i], b = ky[j], c = kz[l];
  // numpy float32: (kx*kx + ky*ky) + kz*kz, then sqrt, each correctly rounded
  // (the square root and the quotient below through FP64, which rounds them
  // to the correctly rounded float32 results)
  const float kk = add_rn(add_rn(mul_rn(a, a), mul_rn(b, b)), mul_rn(c, c));
  const float kf = (float)sqrt((double)kk);
  const double p = spline_eval(sp, (double)kf);
  const float pf = (float)fmax(p, 0.0);
  const float q = (float)((double)pf / (double)vcell);
  out[m] = (float)sqrt((double)q);
}

struct RsdArgs {
  smk_pk_spline sp;
  const double* k_par;
  const double* k_perp;
  int n_perp;
  double beta;
  int model;
  double q1, q2, kp, kv, av, bv, G;
  double dx, kny;
  double* out;
};

// k_perp * P3(k_par, k_perp) of p1d_missing.py:12-37, 49-52; lkpar = log(k_par)
template <int MODEL>
__device__ __forceinline__ double rsd_integrand(const RsdArgs& a, double kpar, double lkpar, double kperp) {
  const double k = sqrt(kpar * kpar + kperp * kperp);
  const double mu = kpar / fmax(k, 1e-15);
  const double p = fmax(spline_eval(a.sp, k), 0.0);
  double f = 1.0;
  if (MODEL == 1) {
    const double d2 = k * k * k * p * a.G * a.G / 2.0 / (M_PI * M_PI);
    // (k / k_v)^a_v mu^b_v as one exponential of logs, mu = k_par / k; the
    // reference's powers where a base is 0
    double pw;
    if (kpar > 0.0) {
      const double lk = log(k);
      pw = exp(a.av * (lk - log(a.kv)) + a.bv * (lkpar - lk));
    } else {
      pw = pow(k / a.kv, a.av) * pow(mu, a.bv);
    }
    f = exp((a.q1 * d2 + a.q2 * d2 * d2) * (1.0 - pw) - (k / a.kp) * (k / a.kp));
  } else if (MODEL == 2) {
    const double w = exp(-a.dx * a.dx * k * k / 2.0);
    f = k < a.kny ? w * w : 0.0;
  }
  const double r = 1.0 + a.beta * mu * mu;
  return kperp * r * r * f * p;
}

// One block per k_par row.  The trapezoid sum over k_perp is taken as
// sum_j f_j (x_{j+1} - x_{j-1}) / 2 (half intervals at the ends), so each
// point is evaluated once.
template <int MODEL>
__global__ __launch_bounds__(256) void k_p1d_rsd_rows(RsdArgs a) {
  __shared__ double red[4];
  const int i = blockIdx.x;
  const double kpar = a.k_par[i];
  const double lkpar = kpar > 0.0 ? log(kpar) : 0.0;
  const int n = a.n_perp;
  double acc = 0.0;
  for (int j = threadIdx.x; j < n; j += 256) {
    const double x = a.k_perp[j];
    const double hi = j + 1 < n ? a.k_perp[j + 1] : x;
    const double lo = j > 0 ? a.k_perp[j - 1] : x;
    acc += (hi - lo) * rsd_integrand<MODEL>(a, kpar, lkpar, x) / 2.0;
  }
  // wave then block reduction in a fixed order
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) a.out[i] = (((red[0] + red[1]) + red[2]) + red[3]) / 2.0 / M_PI;
}

}  // namespace

extern "C" int smk_p1d_rsd_rows(const smk_pk_spline* spline, const double* k_par, int n_par, const double* k_perp,
                                int n_perp, double beta, int model, const double* prats, double dx, double kny,
                                double* out, void* stream) {
  SMK_REQUIRE(spline != nullptr && k_par != nullptr && k_perp != nullptr && out != nullptr, SMK_ERR_ARG,
              "NULL argument");
  SMK_REQUIRE(n_par >= 0 && n_perp >= 2 && model >= 0 && model <= 2, SMK_ERR_ARG, "bad grid or model");
  SMK_REQUIRE(model != 1 || prats != nullptr, SMK_ERR_ARG, "model 1 needs the Prats parameters");
  if (n_par == 0) return SMK_OK;
  RsdArgs a;
  a.sp = *spline;
  a.k_par = k_par;
  a.k_perp = k_perp;
  a.n_perp = n_perp;
  a.beta = beta;
  a.model = model;
  a.q1 = a.q2 = a.kp = a.kv = a.av = a.bv = a.G = 0.0;
  if (model == 1) {
    a.q1 = prats[0]; a.q2 = prats[1]; a.kp = prats[2]; a.kv = prats[3]; a.av = prats[4]; a.bv = prats[5];
    a.G = prats[6];
  }
  a.dx = dx;
  a.kny = kny;
  a.out = out;
  const hipStream_t st = (hipStream_t)stream;
  if (model == 0) hipLaunchKernelGGL(k_p1d_rsd_rows<0>, dim3(n_par), dim3(256), 0, st, a);
  else if (model == 1) hipLaunchKernelGGL(k_p1d_rsd_rows<1>, dim3(n_par), dim3(256), 0, st, a);
  else hipLaunchKernelGGL(k_p1d_rsd_rows<2>, dim3(n_par), dim3(256), 0, st, a);
  SMK_LAUNCH_CHECK();
  return SMK_OK;
}

extern "C" int smk_pk_table(const smk_pk_spline* spline, const float* kx, const float* ky, const float* kz, int nxs,
                            int ny, int nzh, float vcell, float* table, void* stream) {
  SMK_REQUIRE(spline != nullptr && kx != nullptr && ky != nullptr && kz != nullptr && table != nullptr, SMK_ERR_ARG,
              "NULL argument");
  SMK_REQUIRE(spline->x != nullptr && spline->c != nullptr && spline->guess != nullptr && spline->n_int >= 1 &&
                  spline->n_guess >= 1 && spline->g_step > 0.0,
              SMK_ERR_ARG, "bad spline");
  SMK_REQUIRE(nxs >= 0 && ny > 0 && nzh > 0 && vcell > 0.0f, SMK_ERR_ARG, "bad table shape");
  const uint64_t nmodes = (uint64_t)nxs * (uint64_t)ny * (uint64_t)nzh;
  if (nmodes == 0) return SMK_OK;
  const uint64_t nblk = (nmodes + 255) / 256;
  SMK_REQUIRE(nblk <= 0x7fffffffull, SMK_ERR_ARG, "table too large for one launch");
  hipLaunchKernelGGL(k_pk_table, dim3((uint32_t)nblk), dim3(256), 0, (hipStream_t)stream, *spline, kx, ky, kz, ny,
                     nzh, vcell, table, nmodes);
  SMK_LAUNCH_CHECK();
  return SMK_OK;
}
