// Quasar draw on gfx950: one thread per cell of the lognormal slab.
//
// Reference: bin/draw_qso.py:186-462.  The reference builds ptot for the
// whole slab with numpy, then loops over the z planes drawing two uniforms
// per cell and a position in the cell.  Here every cell is independent:
// the uniforms come from Philox at counters fixed by the cell, so a thread
// reads its three field values (coalesced along z), forms ptot, and tests
// the two density conditions; only the rare cells that pass go on to the
// position, the angles, the RSD redshift and the footprint and redshift cut,
// and append a record through one atomic counter.  The host orders the
// records (they carry the reference's output order as a key).
#include <math.h>

#include "fft_core.h"
#include "smk_internal.h"

using namespace smk;

namespace {

constexpr uint32_t TAG_CELL = 0x71736f63u;
constexpr uint32_t TAG_X = 0x71736f78u;
constexpr uint32_t TAG_Y = 0x71736f79u;
constexpr uint32_t TAG_Z = 0x71736f7au;
constexpr double ZB1 = 1.9, ZB2 = 2.75, ZB3 = 3.6;   // constant.py:30-32
constexpr double H0_KMS = 100.0;                     // constant.py:36
constexpr double RAND_QSO_NB = 0.006;                // constant.py:10
// (1 + z_k)^-0.7 of the three bias redshifts
constexpr float INV1 = 0.47459308f;   // 2.9^-0.7
constexpr float INV2 = 0.39644059f;   // 3.75^-0.7
constexpr float INV3 = 0.34361295f;   // 4.6^-0.7

// uniform in [0, 1) with 53 bits from two Philox words
__device__ __forceinline__ double u53(uint32_t a, uint32_t b) {
  return ((double)(a >> 5) * 67108864.0 + (double)(b >> 6)) * 1.1102230246251565e-16;
}

// numpy.interp on ascending nodes, starting the search at node j; sp: NULL,
// or the segments' slopes (fp[j+1] - fp[j]) / (xp[j+1] - xp[j]) made on the
// host (the same correctly rounded operations, so the same bits, without a
// division per call)
__device__ __forceinline__ double interp_from(double x, const double* __restrict__ xp, const double* __restrict__ fp,
                                              const double* __restrict__ sp, int n, int j) {
  if (x <= xp[0]) return fp[0];
  if (x >= xp[n - 1]) return fp[n - 1];
  j = j < 0 ? 0 : (j > n - 2 ? n - 2 : j);
  while (j > 0 && xp[j] > x) --j;
  while (j < n - 2 && xp[j + 1] <= x) ++j;
  const double slope = sp ? sp[j] : (fp[j + 1] - fp[j]) / (xp[j + 1] - xp[j]);
  return slope * (x - xp[j]) + fp[j];
}

// start node for a uniform grid of nodes x0 + i * dx (a guess: interp_from
// walks to the exact node, so a reciprocal is good enough)
__device__ __forceinline__ int uniform_guess(double x, double x0, double dx) {
  const double g = floor((x - x0) * __builtin_amdgcn_rcp(dx));
  return g < 0.0 ? 0 : (g > 2.0e9 ? 2000000000 : (int)g);
}

// z(R): the comoving-distance table through its uniform guess table
__device__ __forceinline__ double z_of_r(const smk_qso_params& p, double r) {
  int b = uniform_guess(r, p.r2z_lo, p.r2z_step);
  b = b >= p.r2z_ng ? p.r2z_ng - 1 : b;
  return interp_from(r, p.r2z_chi, p.r2z_z, p.r2z_slope, p.r2z_n, p.r2z_guess[b]);
}

__device__ __forceinline__ double r_of_z(const smk_qso_params& p, double z) {
  // inverse lookup on the same nodes (z ascending, uniform in the reference's table)
  const int n = p.r2z_n;
  const double z0 = p.r2z_z[0], dz = (p.r2z_z[n - 1] - z0) / (double)(n - 1);
  return interp_from(z, p.r2z_z, p.r2z_chi, nullptr, n, uniform_guess(z, z0, dz));
}

__device__ __forceinline__ double coef_of_z(const smk_qso_params& p, double z) {
  // nodes: 0, then a regular grid from coef_z[1], then 10
  const double z1 = p.coef_z[1], dz = p.coef_z[2] - p.coef_z[1];
  return interp_from(z, p.coef_z, p.coef_c, p.coef_slope, p.coef_n, 1 + uniform_guess(z, z1, dz));
}

// a_k(z) = b(z)(1+z_k) / (b(z_k)(1+z)) with b(z) = 3.7((1+z)/3.33)^1.7
// (util.py:508-517), i.e. ((1+z)/(1+z_k))^0.7
__device__ __forceinline__ double a_of_z(double z, double zk) { return pow((1.0 + z) / (1.0 + zk), 0.7); }

// the two linear interpolations between the bias redshifts and their mix
// (draw_qso.py:247-250), with the constant divisions as products
__device__ __forceinline__ double mix(double z, double c, double e1, double e2, double e3) {
  constexpr double R12 = 1.0 / (ZB2 - ZB1), R23 = 1.0 / (ZB3 - ZB2);
  const double p12 = (e1 * (ZB2 - z) + e2 * (z - ZB1)) * R12;
  const double p23 = (e2 * (ZB3 - z) + e3 * (z - ZB2)) * R23;
  return c * (p12 - p23) + p23;
}

// mean of ptot at redshift z: exp((a_k sigma_k)^2 / 2) mixed as ptot
__device__ __forceinline__ double mean_ptot(const smk_qso_params& p, double z) {
  const double m1 = exp(0.5 * (a_of_z(z, ZB1) * p.sigma[0]) * (a_of_z(z, ZB1) * p.sigma[0]));
  const double m2 = exp(0.5 * (a_of_z(z, ZB2) * p.sigma[1]) * (a_of_z(z, ZB2) * p.sigma[1]));
  const double m3 = exp(0.5 * (a_of_z(z, ZB3) * p.sigma[2]) * (a_of_z(z, ZB3) * p.sigma[2]));
  return mix(z, coef_of_z(p, z), m1, m2, m3);
}

__device__ __forceinline__ double diffmod(double a, double b, double c) {
  double d = fmod(a - b, c);
  if (d < 0.0) d += c;   // Python's % for c > 0
  return fmin(d, c - d);
}

__global__ __launch_bounds__(256) void k_qso_draw(smk_qso_params p, const float* __restrict__ p1,
                                                  const float* __restrict__ p2, const float* __restrict__ p3,
                                                  const float* __restrict__ vx, const float* __restrict__ vy,
                                                  const float* __restrict__ vz, uint64_t seed,
                                                  smk_qso_record* __restrict__ rec, int64_t cap,
                                                  unsigned long long* __restrict__ count) {
  // block (column ix * ny + iy, 256-cell chunk of z): the column's indices
  // are block-uniform (scalar), the cell index c = column * nz + mz
  const uint32_t col = blockIdx.x;
  const int mz = (int)(blockIdx.y * 256 + threadIdx.x);
  if (mz >= p.nz) return;
  const int ix = (int)(col / (uint32_t)p.ny), iy = (int)(col % (uint32_t)p.ny);
  const uint64_t c = (uint64_t)col * (uint64_t)p.nz + (uint64_t)mz;
  const double xc = p.x_axis[ix], yc = p.y_axis[iy], zcell = p.z_axis[mz];
  const double rc = sqrt(xc * xc + yc * yc + zcell * zcell);
  const double zc = z_of_r(p, rc / p.h);
  const U4 w = philox4x32_10_c4((uint32_t)c, (uint32_t)(c >> 32), TAG_CELL, 0u, seed);
  const double rnd1 = u53(w.x, w.y), rnd2 = u53(w.z, w.w);
  // ---- cond1 (draw_qso.py:421-429)
  if (p.randoms) {
    if (!(rnd1 > 1.0 - RAND_QSO_NB)) return;
  } else {
    // exp(p)**a_k(z), float32 as the reference's in-place float32 arrays
    // (:219-223, :236-238), evaluated as exp(a_k p) with a_k = ((1+z)/(1+z_k))^0.7
    // from one power per cell (the reference: a float32 exp, then a float32
    // power), on the hardware exp2 / log2 (a few 1e-7 relative: ptot only
    // decides the draw against a uniform)
    const float t = __builtin_amdgcn_exp2f(0.7f * __builtin_amdgcn_logf((float)(1.0 + zc)));
    const float e1 = __expf(t * INV1 * p1[c]);
    const float e2 = __expf(t * INV2 * p2[c]);
    const float e3 = __expf(t * INV3 * p3[c]);
    const double ptot = mix(zc, coef_of_z(p, zc), (double)e1, (double)e2, (double)e3);
    if (!(rnd1 < p.norm * ptot)) return;
  }
  // ---- cond2: n(z) at the cell centre, corrected for the fields' mean (:400-431)
  const double zcen = p.zfix_on ? p.zfix : zc;
  const double fi = rint((zcen - p.dn_z0) / p.dn_dz);
  double density = (fi >= 0.0 && fi < (double)p.dn_n) ? p.dn_cell[(int)fi] : 0.0;
  if (!p.randoms) density /= mean_ptot(p, zcen);
  if (!(p.density_max * rnd2 < density)) return;
  // ---- position in the cell (:433-437), angles and redshift (:438-444)
  const U4 wx = philox4x32_10_c4((uint32_t)ix, (uint32_t)mz, TAG_X, 0u, seed);
  const U4 wy = philox4x32_10_c4((uint32_t)iy, (uint32_t)mz, TAG_Y, 0u, seed);
  const U4 wz = philox4x32_10_c4((uint32_t)c, (uint32_t)(c >> 32), TAG_Z, 0u, seed);
  const double x = xc + (-0.5 * p.dx + p.dx * u53(wx.x, wx.y));
  const double y = yc + (-0.5 * p.dy + p.dy * u53(wy.x, wy.y));
  const double z = zcell + (-0.5 * p.dz + p.dz * u53(wz.x, wz.y));
  const double cr = cos(p.ra0), sr = sin(p.ra0), cd = cos(p.dec0), sd = sin(p.dec0);
  const double numra = cr * x - sd * sr * y + cd * sr * z;
  const double denra = -sr * x - sd * cr * y + cd * cr * z;
  const double numdec = cd * y + sd * z;
  const double r = sqrt(x * x + y * y + z * z);
  double ra = atan2(numra, denra);
  if (ra < 0.0) ra += 2.0 * M_PI;
  ra *= 180.0 / M_PI;
  const double dec = asin(numdec / r) * (180.0 / M_PI);
  const double red = p.zfix_on ? p.zfix : z_of_r(p, r / p.h);
  double red_rsd = red;
  if (p.rsd && !p.randoms) {
    const double vpar = (x * (double)vx[c] + y * (double)vy[c] + z * (double)vz[c]) / r;
    double rr = p.zfix_on ? r_of_z(p, red) * p.h : r;
    if (red < p.zmax + 1.0) {
      const double dg = interp_from(red, p.dg_z, p.dg_v, nullptr, p.dg_n,
                                    uniform_guess(red, p.dg_z[0], (p.dg_z[p.dg_n - 1] - p.dg_z[0]) / (p.dg_n - 1)));
      rr += vpar * (1.0 + red) * dg / (p.dgrowth0 * H0_KMS);
    }
    red_rsd = z_of_r(p, rr / p.h);
  }
  // ---- cond3 (:455-460)
  const double zsel = (p.rsd && !p.randoms) ? red_rsd : red;
  if (!(diffmod(ra, p.ra0 * (180.0 / M_PI), 360.0) < p.dra)) return;
  if (!(diffmod(dec, p.dec0 * (180.0 / M_PI), 180.0) < p.ddec)) return;
  if (!(zsel > p.zmin && zsel < p.zmax)) return;
  const unsigned long long pos = atomicAdd(count, 1ull);
  if ((int64_t)pos < cap) {
    smk_qso_record q;
    q.key = ((int64_t)mz * p.nx + ix) * p.ny + iy;
    q.x = x; q.y = y; q.z = z; q.r = r;
    q.z_qso = red; q.z_qso_rsd = red_rsd; q.ra = ra; q.dec = dec;
    rec[pos] = q;
  }
}

__global__ void k_clear_count(unsigned long long* c) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *c = 0ull;
}

}  // namespace

extern "C" int smk_qso_draw(const smk_qso_params* p, const float* p1, const float* p2, const float* p3,
                            const float* vx, const float* vy, const float* vz, uint64_t seed,
                            smk_qso_record* records, int64_t capacity, int64_t* count, void* stream) {
  SMK_REQUIRE(p != nullptr && count != nullptr, SMK_ERR_ARG, "NULL argument");
  SMK_REQUIRE(p->nx > 0 && p->ny > 0 && p->nz > 0, SMK_ERR_ARG, "empty slab");
  SMK_REQUIRE(p->randoms || (p1 != nullptr && p2 != nullptr && p3 != nullptr), SMK_ERR_ARG,
              "the three lognormal fields are needed unless randoms");
  SMK_REQUIRE(!p->rsd || p->randoms || (vx != nullptr && vy != nullptr && vz != nullptr), SMK_ERR_ARG,
              "rsd needs the three velocity boxes");
  SMK_REQUIRE(capacity >= 0 && (capacity == 0 || records != nullptr), SMK_ERR_ARG, "bad record buffer");
  SMK_REQUIRE(p->x_axis && p->y_axis && p->z_axis && p->r2z_chi && p->r2z_z && p->r2z_guess && p->coef_z &&
                  p->coef_c && p->dn_cell && (!p->rsd || (p->dg_z && p->dg_v)),
              SMK_ERR_ARG, "NULL table");
  SMK_REQUIRE(p->r2z_n >= 2 && p->r2z_ng >= 1 && p->coef_n >= 3 && p->dn_n >= 1 && (!p->rsd || p->dg_n >= 2) &&
                  p->r2z_step > 0.0 && p->dn_dz > 0.0,
              SMK_ERR_ARG, "bad table sizes");
  hipStream_t s = (hipStream_t)stream;
  unsigned long long* cnt = reinterpret_cast<unsigned long long*>(count);
  hipLaunchKernelGGL(k_clear_count, dim3(1), dim3(64), 0, s, cnt);
  SMK_LAUNCH_CHECK();
  const uint64_t ncol = (uint64_t)p->nx * (uint64_t)p->ny;
  const uint32_t nzb = (uint32_t)((p->nz + 255) / 256);
  SMK_REQUIRE(ncol <= 0x7fffffffull && nzb <= 65535u, SMK_ERR_ARG, "slab too large for one launch");
  hipLaunchKernelGGL(k_qso_draw, dim3((uint32_t)ncol, nzb), dim3(256), 0, s, *p, p1, p2, p3, vx, vy, vz, seed,
                     records, capacity, cnt);
  SMK_LAUNCH_CHECK();
  return SMK_OK;
}
