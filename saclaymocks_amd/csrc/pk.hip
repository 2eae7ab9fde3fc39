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
  const double k = (double)kf;
  double g = floor((k - sp.g_lo) / sp.g_step);
  int s = g < 0.0 ? 0 : (g >= (double)(sp.n_guess - 1) ? sp.n_guess - 1 : (int)g);
  int iv = sp.guess[s];
  while (iv + 1 < sp.n_int && sp.x[iv + 1] <= k) ++iv;
  while (iv > 0 && sp.x[iv] > k) --iv;
  const double t = k - sp.x[iv];
  const double* cc = sp.c + 4 * (int64_t)iv;
  const double p = ((cc[0] * t + cc[1]) * t + cc[2]) * t + cc[3];
  const float pf = (float)fmax(p, 0.0);
  const float q = (float)((double)pf / (double)vcell);
  out[m] = (float)sqrt((double)q);
}

}  // namespace

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
