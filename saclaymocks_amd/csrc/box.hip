// GRF box stage on gfx950: white noise -> R2C -> sqrt(P(k)/Vcell) -> C2R.
//
// Reference: bin/make_boxes.py:40-125,289-352 and bin/interpolate_pk.py:63-77.
//
// Two pass chains share one set of line-FFT kernels.  Every kernel FFTs T
// lines per block; lanes are ordered line-fastest so that an access to the
// "run" axis of a tile is T consecutive complex values.
//
// Fused job (smk_grf_box), transposing chain, 3 complex work boxes A1..A3:
//   Z    z R2C of Philox noise       -> A1 [x][kz][y]
//   Y    y FFT                A1     -> A2 [kz][ky][x]
//   X    x FFT, * sqrt(P/Vcell)/N, aux factor, x IFFT twice
//                             A2     -> A1 (delta_k) and A3 (aux_k), [kz][x][ky]
//   Y^-1 y IFFT  A1 -> A2, A3 -> A1  [x][y][kz]
//   Z^-1 z C2R   A2 -> delta, A1 -> aux, float32 [x][y][z]
// Each pass reads whole contiguous lines and writes runs of T values whose
// stride stays within one (y,kz) or (x,ky) plane, so no tile walks the full
// 40 GB box (the x axis of the numpy layout has a 15.7 MB stride at
// 2560x2560x1536).  HBM traffic: 4 + 8 + 12 + 16 + 16 = 56 bytes per cell.
//
// Drop-in steps (smk_draw_grf_boxk, smk_fft_and_store) keep boxk in the
// numpy layout complex64 [nx][ny][nz/2+1] and work in place on it.
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include <mutex>
#include <vector>

#include "box_kernels.h"
#include "fft_core.h"
#include "smk_internal.h"

using namespace smk;
using namespace smk_box;


namespace {

__global__ void k_reduce_partials(const double* __restrict__ partial, uint64_t n, double* __restrict__ out) {
  __shared__ double red[16];
  double s = 0.0;
  for (uint64_t i = threadIdx.x; i < n; i += blockDim.x) s += partial[i];
  const double tot = block_sum<1024>(s, red);
  if (threadIdx.x == 0) *out = tot;
}

__global__ void k_clear_ints(int* p, int n) {
  if ((int)threadIdx.x < n) p[threadIdx.x] = 0;
}

// --------------------------------------------------------------- dispatch
bool length_supported(int n) {
  switch (n) {
#define SMK_CASE(N) case N:
    SMK_FFT_LENGTHS(SMK_CASE)
#undef SMK_CASE
    return true;
    default: return false;
  }
}

template <int MODE>
hipError_t launch_strided(int n, int vid, const StridedArgs& a, uint64_t n_outer, hipStream_t s) {
  switch (n) {
#define SMK_CASE(N) \
  case N: return launch_len_strided<N>(MODE, vid, a, n_outer, s);
    SMK_FFT_LENGTHS(SMK_CASE)
#undef SMK_CASE
    default: return hipErrorInvalidValue;
  }
}

bool strided_variant_ok(int n, int mode, int vid) {
  return vid >= 0 && vid < SMK_NVAR && length_supported(n) && strided_variant(n, mode, vid).e != 0;
}

// pencils [p0, np)
hipError_t launch_z_r2c(int m, int vid, const float* noise, uint64_t seed, float2* K, const float2* tw,
                        const float2* twr, uint64_t p0, uint64_t np, int trans, uint32_t ny, hipStream_t s) {
  switch (m) {
#define SMK_CASE(N) \
  case N: return launch_len_r2c<N>(vid, noise, seed, K, tw, twr, p0, np, trans, ny, s);
    SMK_FFT_LENGTHS(SMK_CASE)
#undef SMK_CASE
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_z_c2r(int m, int vid, const float2* K, float* out, const float2* tw, const float2* twr,
                        uint64_t np, float scale, int* status, hipStream_t s) {
  switch (m) {
#define SMK_CASE(N) \
  case N: return launch_len_c2r<N>(vid, K, out, tw, twr, np, scale, status, s);
    SMK_FFT_LENGTHS(SMK_CASE)
#undef SMK_CASE
    default: return hipErrorInvalidValue;
  }
}

#define SMK_HIP_RET(expr)                                                                        \
  do {                                                                                           \
    hipError_t smk_e_ = (expr);                                                                  \
    if (smk_e_ != hipSuccess)                                                                    \
      return smk::set_error(SMK_ERR_HIP, "%s: %s", #expr, hipGetErrorString(smk_e_));            \
  } while (0)

std::vector<float2> twiddles(int n, double sign_angle_scale) {
  // W^m = exp(-2 pi i m / (n * sign_angle_scale)), computed in double
  std::vector<float2> t(n);
  for (int m = 0; m < n; ++m) {
    const double ang = -2.0 * M_PI * (double)m / ((double)n * sign_angle_scale);
    t[m] = make_float2((float)cos(ang), (float)sin(ang));
  }
  return t;
}

// numpy.fft.fftfreq(n) * 2 * k_ny, then float32 (make_boxes.py:299-304)
std::vector<float> fftfreq_k(int n, double cell, bool half) {
  const double k_ny = M_PI / cell;
  const double val = 1.0 / (double)n;
  const int len = half ? n / 2 + 1 : n;
  std::vector<float> k(len);
  for (int i = 0; i < len; ++i) {
    int ii = i;
    if (!half && i >= (n + 1) / 2) ii = i - n;
    k[i] = (float)(((double)ii * val) * 2.0 * k_ny);
  }
  return k;
}

}  // namespace

#define SMK_BOX_NPASS 7
#define SMK_BOX_MAXSLAB 16
// timed launches per call: 7 in the sequential schedule, at most
// 5 * slabs + 2 in the slab schedule
#define SMK_BOX_MAXMARK (5 * SMK_BOX_MAXSLAB + 2)
#define SMK_BOX_NSEV (3 * SMK_BOX_MAXSLAB + 8)

struct smk_box_plan {
  int nx, ny, nz, m, nzh;
  double cell;
  // optional per-pass timing of smk_grf_box: one (start, end) event pair
  // per launch, recorded on the launch's stream and tagged with its pass
  // slot; `pending` marks a call whose events have not been read yet
  int timing;
  int pending;
  int nmark;
  hipEvent_t tev[2 * SMK_BOX_MAXMARK];
  int tslot[SMK_BOX_MAXMARK];
  double acc_ms[SMK_BOX_NPASS];
  int ncalls;
  // slab schedule (nslab >= 2): plan-owned worker streams (default priority)
  // and dependency events; 0 = sequential schedule on the caller's stream
  int nslab;
  hipStream_t wstream[2];
  hipEvent_t sev[SMK_BOX_NSEV];
  int have_sched;
  float2* tw_x;
  float2* tw_y;
  float2* tw_z;
  float2* tw_r;
  float* kx;
  float* ky;
  float* kz;
  uint64_t n_partial;
  // pass variants, indexed by SMK_PASS_*
  int vid[SMK_PASS_COUNT];
};

extern "C" int smk_fft_length_supported(int n) { return length_supported(n) ? 1 : 0; }

extern "C" int smk_box_plan_create(smk_box_plan** plan, int nx, int ny, int nz, double cell) {
  SMK_REQUIRE(plan != nullptr, SMK_ERR_ARG, "plan pointer is NULL");
  *plan = nullptr;
  SMK_REQUIRE(nx > 0 && ny > 0 && nz > 0, SMK_ERR_ARG, "box sizes must be positive (%d, %d, %d)", nx, ny, nz);
  SMK_REQUIRE(nz % 2 == 0, SMK_ERR_ARG, "nz must be even (got %d)", nz);
  SMK_REQUIRE(cell > 0.0, SMK_ERR_ARG, "cell size must be positive");
  SMK_REQUIRE(length_supported(nx), SMK_ERR_UNSUPPORTED, "no FFT kernel for nx = %d", nx);
  SMK_REQUIRE(length_supported(ny), SMK_ERR_UNSUPPORTED, "no FFT kernel for ny = %d", ny);
  SMK_REQUIRE(length_supported(nz / 2), SMK_ERR_UNSUPPORTED, "no FFT kernel for nz/2 = %d", nz / 2);
  smk_box_plan* p = (smk_box_plan*)calloc(1, sizeof(smk_box_plan));
  SMK_REQUIRE(p != nullptr, SMK_ERR_OOM, "host allocation failed");
  p->nx = nx; p->ny = ny; p->nz = nz; p->m = nz / 2; p->nzh = nz / 2 + 1; p->cell = cell;
  p->n_partial = strided_blocks_max(nx, MODE_SPECTRAL, (uint64_t)p->nzh, (uint32_t)ny);

  auto upload = [](void** dst, const void* src, size_t bytes) -> hipError_t {
    hipError_t e = hipMalloc(dst, bytes);
    if (e != hipSuccess) return e;
    return hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice);
  };
  auto twx = twiddles(nx, 1.0), twy = twiddles(ny, 1.0), twz = twiddles(p->m, 1.0), twr = twiddles(p->m, 2.0);
  auto kx = fftfreq_k(nx, cell, false), ky = fftfreq_k(ny, cell, false), kz = fftfreq_k(nz, cell, true);
  hipError_t e = hipSuccess;
  if (e == hipSuccess) e = upload((void**)&p->tw_x, twx.data(), twx.size() * sizeof(float2));
  if (e == hipSuccess) e = upload((void**)&p->tw_y, twy.data(), twy.size() * sizeof(float2));
  if (e == hipSuccess) e = upload((void**)&p->tw_z, twz.data(), twz.size() * sizeof(float2));
  if (e == hipSuccess) e = upload((void**)&p->tw_r, twr.data(), twr.size() * sizeof(float2));
  if (e == hipSuccess) e = upload((void**)&p->kx, kx.data(), kx.size() * sizeof(float));
  if (e == hipSuccess) e = upload((void**)&p->ky, ky.data(), ky.size() * sizeof(float));
  if (e == hipSuccess) e = upload((void**)&p->kz, kz.data(), kz.size() * sizeof(float));
  if (e != hipSuccess) {
    smk_box_plan_destroy(p);
    return set_error(e == hipErrorOutOfMemory ? SMK_ERR_OOM : SMK_ERR_HIP, "plan tables: %s", hipGetErrorString(e));
  }
  *plan = p;
  return SMK_OK;
}

extern "C" int smk_box_plan_destroy(smk_box_plan* p) {
  if (p == nullptr) return SMK_OK;
  if (p->timing) {
    for (int i = 0; i < 2 * SMK_BOX_MAXMARK; ++i) (void)hipEventDestroy(p->tev[i]);
  }
  if (p->have_sched) {
    for (int i = 0; i < 2; ++i) (void)hipStreamDestroy(p->wstream[i]);
    for (int i = 0; i < SMK_BOX_NSEV; ++i) (void)hipEventDestroy(p->sev[i]);
  }
  void* bufs[] = {p->tw_x, p->tw_y, p->tw_z, p->tw_r, p->kx, p->ky, p->kz};
  for (void* b : bufs)
    if (b) (void)hipFree(b);
  free(p);
  return SMK_OK;
}

static uint64_t kbox_elems(const smk_box_plan* p) { return (uint64_t)p->nx * p->ny * p->nzh; }
static uint64_t align256(uint64_t b) { return (b + 255) & ~(uint64_t)255; }

// work boxes of smk_grf_box: 3 in the sequential schedule, 4 in the slab
// schedule (the aux branch's y inverse needs its own output box while the
// delta branch still reads A1)
static int work_boxes(const smk_box_plan* p) { return p->nslab >= 2 ? 4 : 3; }

extern "C" size_t smk_box_workspace_bytes(const smk_box_plan* p) {
  if (p == nullptr) return 0;
  const uint64_t kb = align256(kbox_elems(p) * sizeof(float2));
  return (size_t)(work_boxes(p) * kb + align256(p->n_partial * sizeof(double)));
}

namespace {

// ---- natural chain (numpy layout [x][y][kz], in place) -------------------
StridedArgs nat_y_args(const smk_box_plan* p, float2* K) {
  StridedArgs a{};
  a.in = K; a.out = K;
  a.tw = p->tw_y;
  a.in_so = a.out_so = (uint64_t)p->ny * p->nzh;  // outer = x
  a.in_si = a.out_si = 1;                         // inner = kz
  a.in_sn = a.out_sk = (uint64_t)p->nzh;          // along y
  a.n_inner = (uint32_t)p->nzh;
  return a;
}

StridedArgs nat_x_args(const smk_box_plan* p, float2* K) {
  StridedArgs a{};
  a.in = K; a.out = K;
  a.tw = p->tw_x;
  a.in_so = a.out_so = (uint64_t)p->nzh;          // outer = ky
  a.in_si = a.out_si = 1;                         // inner = kz
  a.in_sn = a.out_sk = (uint64_t)p->ny * p->nzh;  // along x
  a.n_inner = (uint32_t)p->nzh;
  return a;
}

int nat_forward(smk_box_plan* p, uint64_t seed, const float* noise, float2* K, hipStream_t s) {
  const uint64_t np = (uint64_t)p->nx * p->ny;
  SMK_HIP_RET(launch_z_r2c(p->m, p->vid[SMK_PASS_Z_R2C], noise, seed, K, p->tw_z, p->tw_r, 0, np, 0, (uint32_t)p->ny,
                           s));
  SMK_HIP_RET(launch_strided<MODE_FWD>(p->ny, p->vid[SMK_PASS_Y_FWD], nat_y_args(p, K), (uint64_t)p->nx, s));
  SMK_HIP_RET(launch_strided<MODE_FWD>(p->nx, p->vid[SMK_PASS_X_FWD], nat_x_args(p, K), (uint64_t)p->ny, s));
  return SMK_OK;
}

// ---- transposing chain (fused job) ---------------------------------------
// A1 [x][kz][y] -> Y -> A2 [kz][ky][x]
StridedArgs tr_y_fwd_args(const smk_box_plan* p, const float2* in, float2* out) {
  StridedArgs a{};
  a.in = in; a.out = out;
  a.tw = p->tw_y;
  a.in_so = (uint64_t)p->ny;                         // outer = kz
  a.in_si = (uint64_t)p->nzh * p->ny;                // inner = x
  a.in_sn = 1;                                       // along y
  a.out_so = (uint64_t)p->ny * p->nx;
  a.out_si = 1;
  a.out_sk = (uint64_t)p->nx;
  a.n_inner = (uint32_t)p->nx;
  return a;
}

// A2 [kz][ky][x] -> X (spectral) -> [kz][x][ky]
StridedArgs tr_x_args(const smk_box_plan* p, const float2* in, float2* out) {
  StridedArgs a{};
  a.in = in; a.out = out;
  a.tw = p->tw_x;
  a.in_so = (uint64_t)p->ny * p->nx;                 // outer = kz
  a.in_si = (uint64_t)p->nx;                         // inner = ky
  a.in_sn = 1;                                       // along x
  a.out_so = (uint64_t)p->nx * p->ny;
  a.out_si = 1;
  a.out_sk = (uint64_t)p->ny;
  a.n_inner = (uint32_t)p->ny;
  a.kz_outer = 1;
  return a;
}

// [kz][x][ky] -> Y^-1 -> [x][y][kz]
StridedArgs tr_y_inv_args(const smk_box_plan* p, const float2* in, float2* out) {
  StridedArgs a{};
  a.in = in; a.out = out;
  a.tw = p->tw_y;
  a.in_so = (uint64_t)p->ny;                         // outer = x
  a.in_si = (uint64_t)p->nx * p->ny;                 // inner = kz
  a.in_sn = 1;                                       // along ky
  a.out_so = (uint64_t)p->ny * p->nzh;
  a.out_si = 1;
  a.out_sk = (uint64_t)p->nzh;
  a.n_inner = (uint32_t)p->nzh;
  return a;
}

}  // namespace

namespace {

// fold the events of the previous timed call into the accumulators
int harvest_timing(smk_box_plan* p) {
  if (!p->timing || !p->pending) return SMK_OK;
  for (int k = 0; k < p->nmark; ++k) {
    SMK_HIP_CHECK(hipEventSynchronize(p->tev[2 * k + 1]));
    float ms = 0.f;
    SMK_HIP_CHECK(hipEventElapsedTime(&ms, p->tev[2 * k], p->tev[2 * k + 1]));
    p->acc_ms[p->tslot[k]] += ms;
  }
  p->ncalls += 1;
  p->pending = 0;
  p->nmark = 0;
  return SMK_OK;
}

// launch `f` on stream s, bracketed by timing events tagged `slot` when
// timing is enabled
template <class F>
hipError_t timed(smk_box_plan* p, int slot, hipStream_t s, F&& f) {
  if (!p->timing || p->nmark >= SMK_BOX_MAXMARK) return f();
  const int k = p->nmark++;
  p->tslot[k] = slot;
  hipError_t e = hipEventRecord(p->tev[2 * k], s);
  if (e != hipSuccess) return e;
  e = f();
  if (e != hipSuccess) return e;
  return hipEventRecord(p->tev[2 * k + 1], s);
}

}  // namespace

extern "C" int smk_box_enable_pass_timing(smk_box_plan* p, int enable) {
  SMK_REQUIRE(p != nullptr, SMK_ERR_ARG, "plan is NULL");
  if (enable && !p->timing) {
    for (int i = 0; i < 2 * SMK_BOX_MAXMARK; ++i) SMK_HIP_CHECK(hipEventCreate(&p->tev[i]));
    p->timing = 1;
  } else if (!enable && p->timing) {
    int rc = harvest_timing(p);
    if (rc) return rc;
    for (int i = 0; i < 2 * SMK_BOX_MAXMARK; ++i) (void)hipEventDestroy(p->tev[i]);
    p->timing = 0;
  }
  for (int i = 0; i < SMK_BOX_NPASS; ++i) p->acc_ms[i] = 0.0;
  p->ncalls = 0;
  p->pending = 0;
  p->nmark = 0;
  return SMK_OK;
}

extern "C" int smk_box_pass_times(smk_box_plan* p, double* ms_sum, int nmax, int* ncalls, int reset) {
  SMK_REQUIRE(p != nullptr && ms_sum != nullptr && ncalls != nullptr, SMK_ERR_ARG, "NULL argument");
  SMK_REQUIRE(p->timing, SMK_ERR_ARG, "pass timing is not enabled");
  int rc = harvest_timing(p);
  if (rc) return rc;
  for (int i = 0; i < nmax && i < SMK_BOX_NPASS; ++i) ms_sum[i] = p->acc_ms[i];
  *ncalls = p->ncalls;
  if (reset) {
    for (int i = 0; i < SMK_BOX_NPASS; ++i) p->acc_ms[i] = 0.0;
    p->ncalls = 0;
  }
  return SMK_OK;
}

extern "C" int smk_box_set_schedule(smk_box_plan* p, int nslab) {
  SMK_REQUIRE(p != nullptr, SMK_ERR_ARG, "plan is NULL");
  SMK_REQUIRE(nslab >= 0 && nslab <= SMK_BOX_MAXSLAB, SMK_ERR_ARG, "nslab must be in [0, %d] (got %d)",
              SMK_BOX_MAXSLAB, nslab);
  if (nslab >= 2 && !p->have_sched) {
    for (int i = 0; i < 2; ++i) SMK_HIP_CHECK(hipStreamCreateWithFlags(&p->wstream[i], hipStreamNonBlocking));
    for (int i = 0; i < SMK_BOX_NSEV; ++i) SMK_HIP_CHECK(hipEventCreateWithFlags(&p->sev[i], hipEventDisableTiming));
    p->have_sched = 1;
  }
  p->nslab = nslab >= 2 ? nslab : 0;
  return SMK_OK;
}

extern "C" int smk_box_get_schedule(const smk_box_plan* p) { return p ? p->nslab : -1; }

namespace {

// Arguments of the spectral x pass over kz planes [k0, k1).
StridedArgs spectral_args(const smk_box_plan* p, const float2* in, float2* out, float2* aux, const double* pk,
                          int aux_kind, double aux_coef, double* partial, uint32_t k0) {
  StridedArgs a = tr_x_args(p, in, out);
  a.in = in + (uint64_t)k0 * a.in_so;
  a.out = out + (uint64_t)k0 * a.out_so;
  a.aux = aux ? aux + (uint64_t)k0 * a.out_so : nullptr;
  a.kxv = p->kx; a.kyv = p->ky; a.kzv = p->kz;
  a.pkA = (float)pk[0];
  a.pk_k0inv = (float)(1.0 / pk[1]);
  a.pk_n = (float)pk[2];
  a.pk_R = (float)pk[3];
  a.pk_l2k0 = (float)log2(pk[1]);
  a.pk_c2 = (float)(pk[3] * pk[3] / (2.0 * M_LN2));
  a.vcell = (float)(p->cell * p->cell * p->cell);
  a.pk_cm = (float)(sqrt(pk[0] / (double)a.vcell) / ((double)p->nx * p->ny * p->nz));
  a.norm = (float)(1.0 / ((double)p->nx * p->ny * p->nz));
  a.aux_coef = (float)aux_coef;
  a.aux_kind = aux_kind;
  a.write_delta = 1;
  a.nz_nyq = (uint32_t)p->m;
  a.kz0 = k0;
  a.partial = partial + strided_blocks(p->nx, MODE_SPECTRAL, p->vid[SMK_PASS_X_SPECTRAL], k0, (uint32_t)p->ny);
  return a;
}

// y FFT of the lines x in [x0, x1) (all kz): A1 [x][kz][y] -> A2 [kz][ky][x]
StridedArgs y_fwd_slab(const smk_box_plan* p, const float2* A1, float2* A2, uint32_t x0, uint32_t x1) {
  StridedArgs a = tr_y_fwd_args(p, A1, A2);
  a.in += (uint64_t)x0 * a.in_si;
  a.out += (uint64_t)x0 * a.out_si;
  a.n_inner = x1 - x0;
  return a;
}

// y IFFT of the kz planes [k0, k1) (all x): [kz][x][ky] -> [x][y][kz]
StridedArgs y_inv_slab(const smk_box_plan* p, const float2* in, float2* out, uint32_t k0, uint32_t k1) {
  StridedArgs a = tr_y_inv_args(p, in, out);
  a.in += (uint64_t)k0 * a.in_si;
  a.out += (uint64_t)k0 * a.out_si;
  a.n_inner = k1 - k0;
  return a;
}

// slab boundaries: [b(i), b(i+1)) for i < n, multiples of `align` except the last
uint32_t slab_bound(uint32_t total, int n, int i, uint32_t align) {
  if (i >= n) return total;
  uint64_t b = (uint64_t)total * (uint64_t)i / (uint64_t)n;
  b = (b + align / 2) / align * align;
  return (uint32_t)(b < total ? b : total);
}

}  // namespace

extern "C" int smk_grf_box(smk_box_plan* p, uint64_t seed, const float* noise, const double* pk_params, int aux_kind,
                           double aux_coef, float* delta_out, float* aux_out, void* workspace, double* var_out,
                           int* status_out, void* stream) {
  SMK_REQUIRE(p != nullptr, SMK_ERR_ARG, "plan is NULL");
  SMK_REQUIRE(pk_params != nullptr, SMK_ERR_ARG, "pk_params is NULL");
  SMK_REQUIRE(delta_out != nullptr && workspace != nullptr, SMK_ERR_ARG, "delta_out/workspace is NULL");
  SMK_REQUIRE(aux_kind >= SMK_AUX_NONE && aux_kind <= SMK_AUX_VZ, SMK_ERR_ARG, "bad aux_kind %d", aux_kind);
  SMK_REQUIRE(aux_kind == SMK_AUX_NONE || aux_out != nullptr, SMK_ERR_ARG, "aux_out is NULL");
  SMK_REQUIRE(((uintptr_t)workspace & 255) == 0, SMK_ERR_ARG, "workspace must be 256-byte aligned");
  SMK_REQUIRE(pk_params[1] > 0.0, SMK_ERR_ARG, "pk k0 must be positive");
  hipStream_t s = (hipStream_t)stream;
  const uint64_t kb = align256(kbox_elems(p) * sizeof(float2));
  const int nbox = work_boxes(p);
  float2* A1 = (float2*)workspace;
  float2* A2 = (float2*)((char*)workspace + kb);
  float2* A3 = (float2*)((char*)workspace + 2 * kb);
  float2* A4 = nbox == 4 ? (float2*)((char*)workspace + 3 * kb) : nullptr;
  double* partial = (double*)((char*)workspace + nbox * kb);
  const uint64_t np = (uint64_t)p->nx * p->ny;
  const bool has_aux = aux_kind != SMK_AUX_NONE;
  const int vzr = p->vid[SMK_PASS_Z_R2C], vyf = p->vid[SMK_PASS_Y_FWD], vx = p->vid[SMK_PASS_X_SPECTRAL];
  const int vyi = p->vid[SMK_PASS_Y_INV], vz = p->vid[SMK_PASS_Z_C2R];

  int rc = harvest_timing(p);
  if (rc) return rc;

  if (status_out) {
    hipLaunchKernelGGL(k_clear_ints, dim3(1), dim3(64), 0, s, status_out, 2);
    SMK_LAUNCH_CHECK();
  }
  const uint64_t nb_spec = strided_blocks(p->nx, MODE_SPECTRAL, vx, (uint64_t)p->nzh, (uint32_t)p->ny);

  if (p->nslab < 2) {
    // ---- sequential schedule on the caller's stream
    SMK_HIP_RET(timed(p, 0, s, [&] {
      return launch_z_r2c(p->m, vzr, noise, seed, A1, p->tw_z, p->tw_r, 0, np, 1, (uint32_t)p->ny, s);
    }));
    SMK_HIP_RET(timed(p, 1, s, [&] {
      return launch_strided<MODE_FWD>(p->ny, vyf, tr_y_fwd_args(p, A1, A2), (uint64_t)p->nzh, s);
    }));
    const StridedArgs a = spectral_args(p, A2, A1, has_aux ? A3 : nullptr, pk_params, aux_kind, aux_coef, partial, 0);
    SMK_HIP_RET(timed(p, 2, s, [&] { return launch_strided<MODE_SPECTRAL>(p->nx, vx, a, (uint64_t)p->nzh, s); }));
    if (var_out) {
      hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(1024), 0, s, partial, nb_spec, var_out);
      SMK_LAUNCH_CHECK();
    }
    SMK_HIP_RET(timed(p, 3, s, [&] {
      return launch_strided<MODE_INV>(p->ny, vyi, tr_y_inv_args(p, A1, A2), (uint64_t)p->nx, s);
    }));
    SMK_HIP_RET(timed(p, 4, s, [&] {
      return launch_z_c2r(p->m, vz, A2, delta_out, p->tw_z, p->tw_r, np, 1.0f, status_out, s);
    }));
    if (has_aux) {
      SMK_HIP_RET(timed(p, 5, s, [&] {
        return launch_strided<MODE_INV>(p->ny, vyi, tr_y_inv_args(p, A3, A1), (uint64_t)p->nx, s);
      }));
      SMK_HIP_RET(timed(p, 6, s, [&] {
        return launch_z_c2r(p->m, vz, A1, aux_out, p->tw_z, p->tw_r, np, 1.0f, status_out ? status_out + 1 : nullptr,
                            s);
      }));
    }
  } else {
    // ---- slab schedule.  Stage A (x slabs): z R2C on stream 1 feeding the
    // y FFT on stream 2.  Stage B (kz slabs): spectral x pass on stream 1
    // feeding the delta y IFFT (A1 -> A4) on stream 2.  Stage C: once the
    // last spectral slab has read A2, the aux y IFFT (A3 -> A2) and its z C2R
    // run on the caller's stream beside the delta z C2R on stream 2.  The
    // compute-bound passes (R2C with its noise, spectral) overlap the
    // memory-bound ones.
    const int S = p->nslab;
    hipStream_t s1 = p->wstream[0], s2 = p->wstream[1];
    hipEvent_t* ev = p->sev;
    int ne = 0;
    auto rec = [&](hipStream_t st) -> hipEvent_t {
      hipEvent_t e = ev[ne++];
      (void)hipEventRecord(e, st);
      return e;
    };
    hipEvent_t start = rec(s);
    SMK_HIP_RET(hipStreamWaitEvent(s1, start, 0));
    SMK_HIP_RET(hipStreamWaitEvent(s2, start, 0));
    for (int i = 0; i < S; ++i) {
      const uint32_t x0 = slab_bound((uint32_t)p->nx, S, i, 8), x1 = slab_bound((uint32_t)p->nx, S, i + 1, 8);
      if (x1 <= x0) continue;
      SMK_HIP_RET(timed(p, 0, s1, [&] {
        return launch_z_r2c(p->m, vzr, noise, seed, A1, p->tw_z, p->tw_r, (uint64_t)x0 * p->ny, (uint64_t)x1 * p->ny,
                            1, (uint32_t)p->ny, s1);
      }));
      SMK_HIP_RET(hipStreamWaitEvent(s2, rec(s1), 0));
      SMK_HIP_RET(timed(p, 1, s2, [&] {
        return launch_strided<MODE_FWD>(p->ny, vyf, y_fwd_slab(p, A1, A2, x0, x1), (uint64_t)p->nzh, s2);
      }));
    }
    SMK_HIP_RET(hipStreamWaitEvent(s1, rec(s2), 0));
    for (int i = 0; i < S; ++i) {
      const uint32_t k0 = slab_bound((uint32_t)p->nzh, S, i, 8), k1 = slab_bound((uint32_t)p->nzh, S, i + 1, 8);
      if (k1 <= k0) continue;
      const StridedArgs a = spectral_args(p, A2, A1, has_aux ? A3 : nullptr, pk_params, aux_kind, aux_coef, partial,
                                          k0);
      SMK_HIP_RET(timed(p, 2, s1, [&] {
        return launch_strided<MODE_SPECTRAL>(p->nx, vx, a, (uint64_t)(k1 - k0), s1);
      }));
      SMK_HIP_RET(hipStreamWaitEvent(s2, rec(s1), 0));
      SMK_HIP_RET(timed(p, 3, s2, [&] {
        return launch_strided<MODE_INV>(p->ny, vyi, y_inv_slab(p, A1, A4, k0, k1), (uint64_t)p->nx, s2);
      }));
    }
    if (var_out) {
      hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(1024), 0, s1, partial, nb_spec, var_out);
      SMK_LAUNCH_CHECK();
    }
    hipEvent_t b_done = rec(s1);
    SMK_HIP_RET(timed(p, 4, s2, [&] {
      return launch_z_c2r(p->m, vz, A4, delta_out, p->tw_z, p->tw_r, np, 1.0f, status_out, s2);
    }));
    SMK_HIP_RET(hipStreamWaitEvent(s, b_done, 0));
    if (has_aux) {
      SMK_HIP_RET(timed(p, 5, s, [&] {
        return launch_strided<MODE_INV>(p->ny, vyi, tr_y_inv_args(p, A3, A2), (uint64_t)p->nx, s);
      }));
      SMK_HIP_RET(timed(p, 6, s, [&] {
        return launch_z_c2r(p->m, vz, A2, aux_out, p->tw_z, p->tw_r, np, 1.0f, status_out ? status_out + 1 : nullptr,
                            s);
      }));
    }
    // join: the caller's stream waits for stream 2 (stream 1 is already joined
    // through b_done)
    SMK_HIP_RET(hipStreamWaitEvent(s, rec(s2), 0));
  }
  if (p->timing) p->pending = 1;
  return SMK_OK;
}

namespace {

// (type, length, mode) of a pass kind: type 0 strided, 1 z pencil
struct PassKind {
  int type, n, mode;
};
PassKind pass_kind(const smk_box_plan* p, int kind) {
  switch (kind) {
    case SMK_PASS_Z_R2C: return {1, p->m, MODE_FWD};
    case SMK_PASS_Y_FWD: return {0, p->ny, MODE_FWD};
    case SMK_PASS_X_SPECTRAL: return {0, p->nx, MODE_SPECTRAL};
    case SMK_PASS_Y_INV: return {0, p->ny, MODE_INV};
    case SMK_PASS_Z_C2R: return {1, p->m, MODE_INV};
    case SMK_PASS_X_FWD: return {0, p->nx, MODE_FWD};
    case SMK_PASS_X_INV: return {0, p->nx, MODE_INV};
    default: return {-1, 0, 0};
  }
}

bool variant_valid(const smk_box_plan* p, int kind, int vid) {
  const PassKind k = pass_kind(p, kind);
  if (k.type == 0) return strided_variant_ok(k.n, k.mode, vid);
  if (k.type == 1) return pencil_variant_ok(k.n, vid, k.mode == MODE_INV);
  return false;
}

}  // namespace

extern "C" int smk_box_variant_count(void) { return SMK_BOX_NVAR; }

extern "C" int smk_box_variant_valid(const smk_box_plan* p, int kind, int vid) {
  return (p != nullptr && variant_valid(p, kind, vid)) ? 1 : 0;
}

extern "C" int smk_box_set_variant(smk_box_plan* p, int kind, int vid) {
  SMK_REQUIRE(p != nullptr, SMK_ERR_ARG, "plan is NULL");
  SMK_REQUIRE(kind >= 0 && kind < SMK_PASS_COUNT, SMK_ERR_ARG, "bad pass kind %d", kind);
  SMK_REQUIRE(variant_valid(p, kind, vid), SMK_ERR_UNSUPPORTED, "variant %d does not exist for pass kind %d", vid,
              kind);
  p->vid[kind] = vid;
  return SMK_OK;
}

extern "C" int smk_box_get_variant(const smk_box_plan* p, int kind) {
  if (p == nullptr || kind < 0 || kind >= SMK_PASS_COUNT) return -1;
  return p->vid[kind];
}

extern "C" int smk_box_autotune(smk_box_plan* p, uint64_t seed, const float* noise, const double* pk_params,
                                int aux_kind, double aux_coef, float* delta_out, float* aux_out, void* workspace,
                                double* var_out, int* status_out, void* stream, int reps, double* ms_out) {
  SMK_REQUIRE(p != nullptr, SMK_ERR_ARG, "plan is NULL");
  SMK_REQUIRE(reps >= 1 && reps <= 100, SMK_ERR_ARG, "reps must be in [1, 100] (got %d)", reps);
  constexpr int NK = 5;  // kinds of the fused job
  const int nv = SMK_NVAR > SMK_NPVAR ? SMK_NVAR : SMK_NPVAR;
  int saved[SMK_PASS_COUNT];
  memcpy(saved, p->vid, sizeof(saved));
  // variants are timed in the sequential schedule (one pass at a time); the
  // workspace of either schedule is large enough for it
  const int saved_nslab = p->nslab;
  p->nslab = 0;
  int rc = smk_box_enable_pass_timing(p, 1);
  if (rc) {
    p->nslab = saved_nslab;
    return rc;
  }
  double best_ms[NK];
  int best[NK];
  for (int k = 0; k < NK; ++k) {
    best_ms[k] = -1.0;
    best[k] = saved[k];
  }
  if (ms_out)
    for (int i = 0; i < NK * SMK_NVAR; ++i) ms_out[i] = -1.0;
  for (int v = 0; v < nv && rc == SMK_OK; ++v) {
    bool any = false;
    for (int k = 0; k < NK; ++k) {
      const bool ok = variant_valid(p, k, v);
      p->vid[k] = ok ? v : 0;
      any = any || ok;
    }
    if (!any) continue;
    for (int r = 0; r <= reps && rc == SMK_OK; ++r) {
      rc = smk_grf_box(p, seed, noise, pk_params, aux_kind, aux_coef, delta_out, aux_out, workspace, var_out,
                       status_out, stream);
      if (rc == SMK_OK && r == 0) {
        // the first call is untimed (module load, cold caches)
        double tmp[SMK_BOX_NPASS];
        int nc = 0;
        rc = smk_box_pass_times(p, tmp, SMK_BOX_NPASS, &nc, 1);
      }
    }
    if (rc) break;
    double ms[SMK_BOX_NPASS] = {0};
    int nc = 0;
    rc = smk_box_pass_times(p, ms, SMK_BOX_NPASS, &nc, 1);
    if (rc) break;
    // pass slots: 0 z R2C, 1 y FFT, 2 x spectral, 3 y IFFT, 4 z C2R, 5 y IFFT aux, 6 z C2R aux
    const double per[NK] = {ms[0], ms[1], ms[2], ms[3] + ms[5], ms[4] + ms[6]};
    for (int k = 0; k < NK; ++k) {
      if (!variant_valid(p, k, v)) continue;
      const double t = per[k] / (nc > 0 ? nc : 1);
      if (ms_out && v < SMK_NVAR) ms_out[k * SMK_NVAR + v] = t;
      if (best_ms[k] < 0.0 || t < best_ms[k]) {
        best_ms[k] = t;
        best[k] = v;
      }
    }
  }
  const int rc2 = smk_box_enable_pass_timing(p, 0);
  p->nslab = saved_nslab;
  if (rc) {
    memcpy(p->vid, saved, sizeof(saved));
    return rc;
  }
  for (int k = 0; k < NK; ++k) p->vid[k] = best[k];
  return rc2;
}

extern "C" int smk_draw_grf_boxk(smk_box_plan* p, uint64_t seed, const float* noise, float* boxk_out, void* stream) {
  SMK_REQUIRE(p != nullptr && boxk_out != nullptr, SMK_ERR_ARG, "plan/boxk_out is NULL");
  return nat_forward(p, seed, noise, (float2*)boxk_out, (hipStream_t)stream);
}

extern "C" int smk_fft_and_store(smk_box_plan* p, float* boxk_inout, float* box_out, void* workspace,
                                 double* sumsq_out, int* status_out, void* stream) {
  SMK_REQUIRE(p != nullptr && boxk_inout != nullptr && box_out != nullptr, SMK_ERR_ARG, "NULL argument");
  SMK_REQUIRE(sumsq_out == nullptr || workspace != nullptr, SMK_ERR_ARG, "sumsq_out needs a workspace");
  hipStream_t s = (hipStream_t)stream;
  float2* K = (float2*)boxk_inout;
  const uint64_t np = (uint64_t)p->nx * p->ny;
  if (status_out) {
    hipLaunchKernelGGL(k_clear_ints, dim3(1), dim3(64), 0, s, status_out, 1);
    SMK_LAUNCH_CHECK();
  }
  StridedArgs a = nat_x_args(p, K);
  a.nz_nyq = (uint32_t)p->m;
  a.partial = sumsq_out ? (double*)workspace : nullptr;
  const int vx = p->vid[SMK_PASS_X_INV];
  SMK_HIP_RET(launch_strided<MODE_INV>(p->nx, vx, a, (uint64_t)p->ny, s));
  if (sumsq_out) {
    const uint64_t nb = strided_blocks(p->nx, MODE_INV, vx, (uint64_t)p->ny, (uint32_t)p->nzh);
    hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(1024), 0, s, a.partial, nb, sumsq_out);
    SMK_LAUNCH_CHECK();
  }
  const float scale = (float)(1.0 / ((double)p->nx * p->ny * p->nz));
  SMK_HIP_RET(launch_strided<MODE_INV>(p->ny, p->vid[SMK_PASS_Y_INV], nat_y_args(p, K), (uint64_t)p->nx, s));
  SMK_HIP_RET(launch_z_c2r(p->m, p->vid[SMK_PASS_Z_C2R], K, box_out, p->tw_z, p->tw_r, np, scale, status_out, s));
  return SMK_OK;
}

extern "C" size_t smk_fft_and_store_workspace_bytes(const smk_box_plan* p) {
  if (p == nullptr) return 0;
  return (size_t)align256(strided_blocks_max(p->nx, MODE_INV, (uint64_t)p->ny, (uint32_t)p->nzh) * sizeof(double));
}
