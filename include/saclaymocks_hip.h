/*
 * saclaymocks_hip.h - C ABI of libsaclay_hip.so, the MI355X (gfx950) kernels
 * behind the SaclayMocks box stage (make_boxes.py, interpolate_pk.py), the
 * quasar draw (draw_qso.py), the forest stage (make_spectra.py +
 * merge_spectra.py) and the FGPA fit (fit_az.py).
 *
 * Conventions
 *   - every entry point returns 0 (SMK_OK) or a negative SMK_ERR_* code and
 *     leaves a message readable with smk_last_error() (per host thread);
 *   - every array argument is a DEVICE pointer (e.g. torch tensor data_ptr()),
 *     except where a comment says "host";
 *   - `stream` is a hipStream_t passed as void* (0 = legacy default stream);
 *     calls are stream ordered and asynchronous;
 *   - compute entry points never allocate: scratch memory is a caller-given
 *     workspace whose size the library reports;
 *   - a plan must not be shared by host threads running calls concurrently.
 *
 * Layouts follow the reference's numpy arrays (C order):
 *   real box       float32 [nx][ny][nz]        (make_boxes.py:86)
 *   k-space box    complex64 [nx][ny][nz/2+1]  (make_boxes.py:52), stored as
 *                  interleaved float pairs.
 */
#ifndef SACLAYMOCKS_HIP_H
#define SACLAYMOCKS_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SMK_ABI_VERSION 9

#define SMK_OK 0
#define SMK_ERR_ARG (-1)         /* bad argument (null pointer, size, enum) */
#define SMK_ERR_OOM (-2)         /* device allocation failed (plan tables only) */
#define SMK_ERR_NULLBOX (-3)     /* output box all-zero or holds NaN */
#define SMK_ERR_HIP (-4)         /* HIP runtime error */
#define SMK_ERR_UNSUPPORTED (-5) /* FFT length without a compiled kernel */
#define SMK_ERR_RANGE (-6)       /* value outside a table or the box */

/* auxiliary field produced beside delta by the fused k-space pass */
#define SMK_AUX_NONE 0
#define SMK_AUX_ETA_XX 1
#define SMK_AUX_ETA_YY 2
#define SMK_AUX_ETA_ZZ 3 /* eta_par for a line of sight along z */
#define SMK_AUX_ETA_XY 4
#define SMK_AUX_ETA_XZ 5
#define SMK_AUX_ETA_YZ 6
#define SMK_AUX_VX 7
#define SMK_AUX_VY 8
#define SMK_AUX_VZ 9

int smk_abi_version(void);
const char* smk_last_error(void);

/* 1 if an FFT line of n complex points has a compiled kernel, else 0.
 * Box sizes need nx, ny and nz/2 to be supported (nz even). */
int smk_fft_length_supported(int n);

/* ------------------------------------------------------------------ box */
typedef struct smk_box_plan smk_box_plan;

/* Plan for an nx*ny*nz box of cubic cells of side `cell` (Mpc/h).  Owns the
 * FFT twiddle tables (FP64-computed, stored FP32) and the wavenumber axes
 * kx = fftfreq(nx)*2*k_ny, ky likewise, kz = rfftfreq(nz)*2*k_ny, cast to
 * float32 as in interpolate_pk.py:63-77 / make_boxes.py:299-305.
 * Replaces the FFTW plans built at make_boxes.py:53 and :87. */
int smk_box_plan_create(smk_box_plan** plan, int nx, int ny, int nz, double cell);
int smk_box_plan_destroy(smk_box_plan* plan);

/* Bytes of device workspace smk_grf_box needs under the plan's current
 * schedule: three k-space boxes (four in the slab schedule) and per-block
 * FP64 partial sums, i.e. 3 (or 4) * nx*ny*(nz/2+1)*8 bytes + small. */
size_t smk_box_workspace_bytes(const smk_box_plan* plan);

/* Fused GRF box: noise -> R2C -> x sqrt(P(k)/Vcell) -> C2R giving delta, and
 * the auxiliary field (eta_zz for SMK_AUX_ETA_ZZ) out of the same k-space
 * pass.  Replaces make_boxes.py:40-72 (DrawGRF_boxk), :289-292 (x P0 and
 * FFTandStore of the density) and :341-352 (eta_zz), with the power spectrum
 * evaluated per mode instead of read from the P0 table of
 * interpolate_pk.py:146-150.
 *   seed        Philox key for on-device N(0,1) noise (used if noise == NULL)
 *   noise       NULL, or float32 [nx][ny][nz] white noise (parity mode)
 *   pk_params   host double[4] {A, k0, n, R}:
 *               P(k) = A (k/k0) / (1 + (k/k0)^n) exp(-(k R)^2)
 *   aux_kind    SMK_AUX_*; aux_coef multiplies velocity kinds
 *               (H0 * dD/dz(z=0), make_boxes.py:403)
 *   delta_out   float32 [nx][ny][nz]
 *   aux_out     float32 [nx][ny][nz] or NULL when aux_kind == SMK_AUX_NONE
 *   workspace   smk_box_workspace_bytes(plan) bytes, 256-byte aligned
 *   var_out     device double[1]: variance of delta from k-space, FP64
 *   status_out  device int[2]: [0] bit0 = NaN seen, bit1 = non-zero seen in
 *               delta; [1] same for aux.  The null/NaN guard of
 *               make_boxes.py:100-105 is the host's check of these bits. */
int smk_grf_box(smk_box_plan* plan, uint64_t seed, const float* noise, const double* pk_params,
                int aux_kind, double aux_coef, float* delta_out, float* aux_out, void* workspace,
                double* var_out, int* status_out, void* stream);

/* Per-pass timing of smk_grf_box with HIP events recorded around each
 * launch on the stream it runs on (the reference prints wall times per
 * step: make_boxes.py:44-56,90).  Passes: 0 z R2C, 1 y FFT, 2 x spectral,
 * 3 y IFFT delta, 4 z C2R delta, 5 y IFFT aux, 6 z C2R aux; in the slab
 * schedule a pass sums its slab launches, which overlap other passes.
 * smk_box_pass_times waits for the last timed call, writes the per-pass sums
 * (ms) over the calls since the last reset into ms_sum[0..nmax) and the
 * number of calls into *ncalls. */
int smk_box_enable_pass_timing(smk_box_plan* plan, int enable);
int smk_box_pass_times(smk_box_plan* plan, double* ms_sum, int nmax, int* ncalls, int reset);

/* Schedule of smk_grf_box.  nslab = 0: the passes run one after the other on
 * the caller's stream (3 work boxes).  nslab in [2, 16]: slab schedule on
 * two plan-owned streams beside the caller's, joined back into it before
 * the call's work ends (stream order is kept for the caller): z R2C and
 * y FFT are pipelined over nslab x slabs, the spectral pass and the delta
 * y IFFT over nslab kz slabs, so that compute-bound and memory-bound passes
 * overlap.  Needs 4 work boxes: query smk_box_workspace_bytes after setting
 * the schedule.  Results are identical to the sequential schedule's.  The
 * reference's counterpart is its multiprocessing pool over x planes
 * (make_boxes.py:100-117, ncpu workers). */
int smk_box_set_schedule(smk_box_plan* plan, int nslab);
int smk_box_get_schedule(const smk_box_plan* plan);

/* Kernel variants per pass kind.  The reference leaves this choice to FFTW's
 * planner and its wisdom file (make_boxes.py:53-54,87, FFTW_MEASURE): here a
 * strided pass picks its register tile, lines per block and LDS exchange
 * form and the swizzle of its LDS lines (variants 0..SMK_BOX_NVAR-1), a
 * z-pencil pass its register tile (0..1).  Variant 0 is the default of every
 * pass; results of all variants agree to FP32 rounding.  Kinds 0-4 are the
 * passes of smk_grf_box; the drop-in steps use kinds 0, 1, 3, 4 (y and z
 * axes) and 5, 6 (x axis). */
#define SMK_PASS_Z_R2C 0
#define SMK_PASS_Y_FWD 1
#define SMK_PASS_X_SPECTRAL 2
#define SMK_PASS_Y_INV 3
#define SMK_PASS_Z_C2R 4
#define SMK_PASS_X_FWD 5
#define SMK_PASS_X_INV 6
#define SMK_PASS_COUNT 7
#define SMK_BOX_NVAR 17

/* SMK_BOX_NVAR of the library (for bindings that cannot read the header). */
int smk_box_variant_count(void);
/* 1 if variant `vid` exists for pass `kind` of this plan's sizes, else 0. */
int smk_box_variant_valid(const smk_box_plan* plan, int kind, int vid);
/* Select a variant (SMK_ERR_UNSUPPORTED if it does not exist) / read it. */
int smk_box_set_variant(smk_box_plan* plan, int kind, int vid);
int smk_box_get_variant(const smk_box_plan* plan, int kind);

/* Autotuner of the fused job (the counterpart of FFTW_MEASURE planning):
 * runs smk_grf_box `reps` + 1 times per variant index (the first untimed)
 * with the same arguments, times each pass with HIP events, and keeps for
 * each of kinds 0-4 the fastest valid variant.  The outputs then hold a
 * valid result of the call.  ms_out (host, may be NULL) receives
 * [5][SMK_BOX_NVAR] mean milliseconds per call (-1 for a variant that does
 * not exist); the passes of kinds 3 and 4 count their delta and aux runs.
 * Variants are timed in the sequential schedule whatever the plan's
 * schedule (which is kept).  Pass timing is left disabled and its sums
 * cleared. */
int smk_box_autotune(smk_box_plan* plan, uint64_t seed, const float* noise, const double* pk_params, int aux_kind,
                     double aux_coef, float* delta_out, float* aux_out, void* workspace, double* var_out,
                     int* status_out, void* stream, int reps, double* ms_out);

/* DrawGRF_boxk (make_boxes.py:40-72): unnormalised forward R2C FFT of the
 * white noise into boxk, complex64 [nx][ny][nz/2+1].  noise == NULL draws
 * the Philox noise of `seed` on the device. */
int smk_draw_grf_boxk(smk_box_plan* plan, uint64_t seed, const float* noise, float* boxk_out,
                      void* stream);

/* FFTandStore's transform (make_boxes.py:86-93): box = irfftn(boxk) / N.
 * boxk is destroyed (FFTW_DESTROY_INPUT at make_boxes.py:87).
 *   workspace   smk_fft_and_store_workspace_bytes(plan) bytes, or NULL when
 *               sumsq_out is NULL
 *   sumsq_out   NULL or device double[1]: sum over the FULL spectrum of
 *               |boxk|^2, FP64, taken before the transform (Parseval:
 *               sum(box^2) = sumsq / N); gives np.std(box) of make_boxes.py:92
 *               once the mean (Re boxk[0,0,0] / N) is removed
 *   status_out  NULL or device int[1], same bits as smk_grf_box. */
int smk_fft_and_store(smk_box_plan* plan, float* boxk_inout, float* box_out, void* workspace,
                      double* sumsq_out, int* status_out, void* stream);
size_t smk_fft_and_store_workspace_bytes(const smk_box_plan* plan);

/* ---------------------------------------------------------------- forest */
#define SMK_GEOM_PLANE 0  /* line of sight along z at fixed (X, Y) */
#define SMK_GEOM_RADIAL 1 /* line of sight from the observer, make_spectra.py:443-452 */
#define SMK_INTERP_TRILINEAR 0
#define SMK_INTERP_GAUSS343 1 /* 343-cell Gaussian kernel, make_spectra.py:39-68 */
/* Gaussian kernel over (2d + 1)^3 cells, make_spectra -dmax d
 * (make_spectra.py:147,385-390), 0 <= d <= SMK_GAUSS_DMAX;
 * SMK_INTERP_GAUSS(3) is the same kernel as SMK_INTERP_GAUSS343. */
#define SMK_GAUSS_DMAX 6
#define SMK_INTERP_GAUSS(d) (16 + (d))

/* Resident density / tidal boxes, float32 [nx][ny][nz].  Box frame of
 * make_spectra.py:196-206,300-305: X in [-LX/2, LX/2), Y likewise, Z in
 * [R0 - LZ/2, R0 + LZ/2) with L = cell * n. */
typedef struct smk_box_view {
  const float* delta;
  const float* eta[6];  /* eta_xx, eta_yy, eta_zz, eta_xy, eta_xz, eta_yz; PLANE reads eta_zz */
  int32_t nx, ny, nz;
  int32_t periodic;     /* 1: wrap cell indices; 0: a skewer leaving the box is flagged */
  double cell;          /* DX = DY = DZ, Mpc/h */
  double r0;            /* R0 = h chi(z0), Mpc/h */
  /* x slab resident on the device (make_spectra.py:220-226 reads planes
   * [iXmin, iXmax) of the box files): nxs > 0 means the field pointers point
   * at plane x0 and hold planes [x0, x0 + nxs) of the nx-plane box; a cell
   * outside them is never read (status bit 2).  nxs = 0: the whole box. */
  int32_t x0, nxs;
  const float* vel[3];  /* vx, vy, vz: the velocity boxes of v_par (RADIAL); NULL when unused */
} smk_box_view;

/* A batch of skewers on the global pixel grid R_g = rmin + g * dr (Mpc/h),
 * make_spectra.py:311-321.  Pixels are packed: skewer s owns
 * [off[s], off[s] + npix[s]) of every per-pixel array, with
 * off[s+1] = off[s] + npix[s].  All pointers are device arrays of length n. */
typedef struct smk_skewers {
  int32_t n;
  int32_t geometry;     /* SMK_GEOM_* */
  int64_t npix_total;   /* off[n-1] + npix[n-1] */
  double rmin, dr;
  const double* qx;     /* RADIAL: X_QSO; PLANE: X of the line of sight */
  const double* qy;     /* RADIAL: Y_QSO; PLANE: Y of the line of sight */
  const double* qz;     /* RADIAL: Z_QSO; PLANE: unused */
  const double* qr;     /* RADIAL: R_QSO; PLANE: unused */
  const int32_t* g0;    /* global grid index of pixel 0 */
  const int32_t* npix;
  const int32_t* imin;  /* forest range [imin, imax) in local pixels; outside it */
  const int32_t* imax;  /* delta_l = -1e6 and eta_par = 0 (make_spectra.py:99-105) */
  const int64_t* off;
  const int32_t* pset;  /* FGPA parameter row per skewer; NULL: row = global index g */
  const uint64_t* noise_id; /* missing-power noise stream per skewer (e.g. THING_ID) */
  /* NULL, or ceil(npix_total / SMK_GATHER_BLOCK_PIXELS) + 1 entries: entry b
   * is the skewer of pixel min(b * SMK_GATHER_BLOCK_PIXELS, npix_total - 1),
   * the last s with off[s] <= that pixel.  smk_forest_gather then reads each
   * block's skewer range from it instead of searching off[] on the device
   * (entries are clamped to [0, n - 1]; a wrong table gives wrong pixels,
   * never an access outside the arrays). */
  const int32_t* blk_s0;
} smk_skewers;

/* Pixels per block of smk_forest_gather (the blk_s0 table's stride). */
#define SMK_GATHER_BLOCK_PIXELS 256

/* Gather along each line of sight (make_spectra.py:90-139) and, when `flux`
 * is non-NULL, the FGPA (merge_spectra.py:335-339, util.py:421-433):
 *   F = exp(-a exp(b G (delta_l + delta_s + c eta_par)))
 *   abcg        float32 [abcg_rows][4] = (a, b, c, G), row = skewers.pset[s]
 *               or, with pset NULL, the global pixel index g
 *   delta_s     NULL or float32 per pixel (smk_forest_missing_power)
 * Outputs (NULL to skip, float32 per pixel): delta_l, eta_par, flux;
 *   v_par       the smoothed velocity along the line of sight, (vx X + vy Y
 *               + vz Z) / R (make_spectra.py:130-134, the -dla skewers); RADIAL
 *               skewers with the six eta and three velocity boxes; 0 outside
 *               the forest;
 *   cells       int32 [pixels][3], containing cell (ix, iy, iz), make_spectra.py:47-49
 *   status      int32 per skewer (zeroed by the caller): bit 0 set if a kernel
 *               cell left a non-periodic box (the reference skips such a QSO,
 *               make_spectra.py:490-494); bit 1 if an FGPA row was out of
 *               range (flux NaN); bit 2 if a kernel cell lay outside the
 *               resident x slab (smk_box_view.x0, nxs). */
int smk_forest_gather(const smk_skewers* sk, const smk_box_view* box, int interp, const float* abcg,
                      int64_t abcg_rows, const float* delta_s, float* delta_l, float* eta_par, float* v_par,
                      float* flux, int32_t* cells, int32_t* status, void* stream);

/* Missing small-scale power, merge_spectra.py:303-322: for a skewer of n
 * pixels, nz = 256 * 2^m >= n + 50; delta_s = irfft(rfft(w) * filt)[:n] with
 * filt[j] = sqrt(P1D_miss(k_j) / pixel), k_j = j 2 pi / (nz pixel).
 *   filt        float32 tables, one per length class c (nz = 256 << c, c < 6),
 *               table c at filt + filt_off[c], nz/2 + 1 entries (device)
 *   noise       NULL: the Fourier modes of w are drawn on the device (Philox,
 *               key `seed`, stream skewers.noise_id[s]; the reference's own
 *               comment at merge_spectra.py:310 plans this); else float32
 *               N(0,1) white noise, nz values per skewer at noise + noise_off[s]
 *   delta_s     float32 per pixel (output). */
typedef struct smk_forest_plan smk_forest_plan;
int smk_forest_plan_create(smk_forest_plan** plan); /* FFT tables for every class */
int smk_forest_plan_destroy(smk_forest_plan* plan);
/* Graph mode, for capturing a forest step in a HIP graph (no counterpart in
 * the reference, whose steps run eagerly): with seed_dev non-NULL, the
 * missing-power calls of this plan read the seed from *seed_dev (a device
 * uint64) when their kernels run, so a replay takes the seed the caller has
 * just written, and they do not read the length-class checks back to the host
 * (a capture may not synchronise): run the batch once eagerly first, which
 * checks it.  NULL restores the normal mode. */
int smk_forest_plan_set_graph_mode(smk_forest_plan* plan, const uint64_t* seed_dev);
/* Kernel instance of smk_forest_step for batches whose FFTs are at most 2048
 * points (the planner's choice, no counterpart in the reference): one wave
 * per skewer with at least 4 (0, the default) or 3 (1) waves per SIMD, two
 * waves per skewer with at least 4 (2) or 5 (3).  0 and 1 give the same
 * bits, as do 2 and 3; the two pairs transform the missing power with
 * different register tiles, so their delta_s and F differ by float32
 * rounding (delta_l and eta_par are the same bits). */
#define SMK_FOREST_STEP_NVAR 4
int smk_forest_plan_set_step_variant(smk_forest_plan* plan, int variant);
size_t smk_forest_workspace_bytes(int n_skewers);   /* class bucketing scratch */
/*   filt_off    HOST int64[6]
 *   workspace   smk_forest_workspace_bytes(sk->n) bytes
 * Returns SMK_ERR_RANGE if a skewer needs nz > 8192 (one host sync). */
int smk_forest_missing_power(smk_forest_plan* plan, const smk_skewers* sk, const float* filt,
                             const int64_t* filt_off, uint64_t seed, const float* noise,
                             const int64_t* noise_off, void* workspace, float* delta_s, void* stream);

/* The forest step of one batch in one persistent launch: for plane-parallel
 * skewers, the device-drawn missing power (smk_forest_missing_power_sets,
 * merge_spectra.py:303-322), trilinear delta and eta_zz along the line of
 * sight (smk_forest_gather) and the FGPA (merge_spectra.py:335-339), with
 * delta_s kept on chip.  Each block takes skewers from a work queue in the
 * workspace (smk_forest_step_workspace_bytes(), cleared by the call).
 *   filt, filt_off, set_stride, n_sets, filt_set, scale, seed: as
 *               smk_forest_missing_power_sets (device-drawn modes only; the
 *               plan's graph mode applies)
 *   abcg        float32 [abcg_rows][4] FGPA rows, as smk_forest_gather
 *   max_npix    an upper bound of skewers.npix (sizes the on-chip buffers);
 *               a longer skewer is not processed (status bit 3, flux NaN)
 *   flux        float32 per pixel (output); delta_l, eta_par, delta_s: NULL
 *               or float32 per pixel (outputs, for checks)
 *   status      int32 per skewer, zeroed by the caller: the bits of
 *               smk_forest_gather, and bit 3 for a skewer not processed
 * Box: box.delta and, if non-NULL, box.eta[2] (eta_zz); periodic or not,
 * whole or an x slab.  Returns SMK_ERR_RANGE if max_npix needs nz > 8192 or,
 * outside graph mode, if a filter set index is out of range (one host sync).
 * Replaces, for the plane-parallel bench workload, the per-skewer loop of
 * merge_spectra.py:282-339 over the spectra of make_spectra.py:429-522. */
size_t smk_forest_step_workspace_bytes(void);
int smk_forest_step(smk_forest_plan* plan, const smk_skewers* sk, const smk_box_view* box, const float* filt,
                    const int64_t* filt_off, int64_t set_stride, int32_t n_sets, const int32_t* filt_set,
                    const float* scale, uint64_t seed, const float* abcg, int64_t abcg_rows, int32_t max_npix,
                    float* flux, float* delta_l, float* eta_par, float* delta_s, int32_t* status, void* workspace,
                    void* stream);

/* smk_forest_step without a workspace argument, for callers that cannot
 * allocate device memory (code under HIP graph capture, serving loops); it
 * has no counterpart in the reference.  The step's work queue lives in a
 * static device array of smk_forest_step_noalloc_scratch_bytes() bytes per
 * GPU (held until the process ends); each call clears what it uses.  All
 * calls on one GPU share that array, so calls that overlap on one GPU race.
 * Returns SMK_ERR_RANGE if the step ever needs more scratch than the array
 * holds; otherwise as smk_forest_step.  Calls no hipMalloc/hipFree. */
size_t smk_forest_step_noalloc_scratch_bytes(void);
int smk_forest_step_noalloc(smk_forest_plan* plan, const smk_skewers* sk, const smk_box_view* box,
                            const float* filt, const int64_t* filt_off, int64_t set_stride, int32_t n_sets,
                            const int32_t* filt_set, const float* scale, uint64_t seed, const float* abcg,
                            int64_t abcg_rows, int32_t max_npix, float* flux, float* delta_l, float* eta_par,
                            float* delta_s, int32_t* status, void* stream);

/* Missing power with a P1D_miss table chosen per skewer, merge_spectra.py:311-322
 * without --fit-p1d: P1D_miss(z, k) is the table at the redshift nearest the
 * QSO's mean forest redshift z_eff (util.py:448-474, InterpP1Dmissing), and
 * delta_s is then scaled per pixel by sigma_s(z) / sigma_s(z_eff).
 *   filt        n_sets filter sets, set i at filt + i * set_stride, each laid
 *               out as the filt of smk_forest_missing_power (same filt_off)
 *   filt_set    device int32 per skewer in [0, n_sets) (SMK_ERR_RANGE else),
 *               or NULL: set 0 for every skewer
 *   scale       NULL or device float32 per pixel multiplying delta_s
 * Other arguments as smk_forest_missing_power. */
int smk_forest_missing_power_sets(smk_forest_plan* plan, const smk_skewers* sk, const float* filt,
                                  const int64_t* filt_off, int64_t set_stride, int32_t n_sets,
                                  const int32_t* filt_set, const float* scale, uint64_t seed,
                                  const float* noise, const int64_t* noise_off, void* workspace,
                                  float* delta_s, void* stream);

/* ------------------------------------------------------------- FGPA fit */
/* The fit kernels need skewers.pset (one parameter row per set, e.g. a
 * redshift bin) and use the forest range [imin, imax) of every skewer, the
 * unmasked pixels of fit_az.py:88-91. */

/* comp_spectra (fit_az.py:245-256): F = exp(-a exp(b G (delta + c eta)))
 * for every pixel (delta = delta_l + delta_s, eta may be NULL), FP64 as the
 * reference (fit_az.py:101-110 promotes the fields to float64); also the
 * FGPA of merged spectra (merge_spectra.py:335-339);
 *   abcg   device double [rows][4] = (a, b, c, G), row = skewers.pset[s], or
 *          with pset NULL the pixel's global grid index g0[s] + i (a, b, c
 *          and G at each pixel's redshift, merge_spectra.py:282-296)
 *   flux   device double per pixel
 *   fsum   device double per skewer: sum of F over the forest. */
int smk_fgpa_flux(const smk_skewers* sk, const float* delta, const float* eta, const double* abcg,
                  int64_t abcg_rows, double* flux, double* fsum, void* stream);

/* out[set][c] = sum over the skewers of that set of per_skewer[s][c] (FP64,
 * deterministic: a fixed order of additions for given n, width, n_sets),
 * divided by divisor[set] when divisor (device double [n_sets]) is not
 * NULL: mean flux, P1D bins and adjoint terms per set (fit_az.py:206-217
 * takes the means over the forests of one redshift).
 *   work  device double scratch of at least smk_reduce_work_len(skewers.n,
 *         width, n_sets) values; calls that share it must not overlap. */
int64_t smk_reduce_work_len(int n, int width, int n_sets);
int smk_reduce_by_set(const smk_skewers* sk, const double* per_skewer, int width, int n_sets,
                      const double* divisor, double* out, double* work, int64_t work_len, void* stream);

/* compute_p1d (fit_az.py:194-217): per skewer, delta_F = regroup(F, m) /
 * <F>_set - 1 over the forest (util.py:29-37), P_j = |rfft(delta_F)_j|^2 px / N
 * with px = m * pixel (powerspectrum.py:204-209), k_j = rfftfreq(N) 2 pi / px,
 * then per bin b of `edges` (scipy binned_statistic: [e_b, e_b+1), last bin
 * closed): partial[s][b] = (count, sum k, sum P, sum P^2), FP64.
 *   fmean       device double per set
 *   max_forest  longest forest (imax - imin) in the batch; regrouped
 *               forests of more than 8192 points take a direct DFT with the
 *               input staged through LDS (O(N^2), no length limit)
 *   status      NULL or int32 per skewer, 1 if its forest exceeds max_forest
 *   modes       NULL or device complex double: the rfft modes X_j of every
 *               skewer, (N/2 + 1) of them at modes + 2 * modes_off[s], kept
 *               for smk_p1d_adjoint
 *   tables      NULL, or the Bluestein tables of smk_p1d_tables for the
 *               batch's forest lengths, with tab_off[s] the offset (in
 *               complex doubles) of skewer s's length table, < 0 for none;
 *               the transform is then two FFTs on one register tile instead
 *               of three, in a kernel per transform size (same results to
 *               FP64 rounding)
 *   pairs       NULL, or (with tables) device int32 [n_pairs][2]: skewers of
 *               one forest length that share one DFT (of delta_F1 + i
 *               delta_F2, split by the symmetry of real inputs; second -1 for
 *               a skewer alone), sorted by transform size; every skewer with a
 *               table in exactly one pair
 *   pair_off    with pairs, HOST int64[4]: the pairs of M = 1024, 2048, 4096
 *               at [pair_off[c], pair_off[c + 1]) */
int smk_p1d_binned(const smk_skewers* sk, const double* flux, const double* fmean, double pixel, int regroup,
                   const double* edges, int nbins, int max_forest, double* partial, int32_t* status,
                   double* modes, const int64_t* modes_off, const double* tables, const int64_t* tab_off,
                   const int32_t* pairs, const int64_t* pair_off, void* stream);

/* Length of the Bluestein table of a forest of n (regrouped) points, in
 * complex doubles: n + M (M the smallest power of two >= max(2n - 1, 1024))
 * for 1 <= n <= 2048, else 0 (such forests take the direct DFT).  A
 * building block of powerspectrum.ComputeP1D (powerspectrum.py:204-238) on
 * the device; the reference's numpy rfft needs no table. */
int64_t smk_p1d_table_len(int n);

/* The Bluestein tables of n_lens forest lengths: lens (device int32) and
 * offs (device int64, in complex doubles) give length d's table at
 * tables + 2 * offs[d], smk_p1d_table_len(lens[d]) complex doubles: the
 * chirp w_m = exp(-i pi m^2 / N), m < N, then FFT_M(b) / M of the chirp
 * filter b (b_m = conj(w_|m|), circular).  Made once per batch (a fit's
 * forests do not change between evaluations).  Lengths outside [1, 2048]
 * are skipped. */
int smk_p1d_tables(const int32_t* lens, const int64_t* offs, int n_lens, double* tables, void* stream);

/* Adjoint of the binned P1D and of the FGPA, for a loss L(<F>_set, P1D bins).
 *   gp    device double [sets][nbins]: dL/dP for one P_j sample of bin b
 *         (dL/d mean_b divided by the bin count)
 *   modes NULL (recompute) or the modes saved by smk_p1d_binned at the same
 *         parameters; required (SMK_ERR_ARG) when max_forest / m > 8192
 *   tables, tab_off, pairs, pair_off  as smk_p1d_binned (a pair's two
 *         inverse transforms as one DFT of H1 + i H2, H the Hermitian spectra
 *         whose DFTs are the real A1, A2)
 * out[s][7] = (sum_q A_q F'_q, sum_n A_q(n) dF_n/d(a, b, c) / m,
 *              sum_n dF_n/d(a, b, c)) with A_q = dL/d delta_F,q at fixed <F>;
 * the host adds the <F> chain rule per set.  No counterpart in the
 * reference, whose fit is derivative-free (fit_az.py:266-279). */
int smk_p1d_adjoint(const smk_skewers* sk, const float* delta, const float* eta, const double* abcg,
                    int64_t abcg_rows, const double* flux, const double* fmean, double pixel, int regroup,
                    const double* edges, int nbins, int max_forest, const double* modes,
                    const int64_t* modes_off, const double* gp, double* out, const double* tables,
                    const int64_t* tab_off, const int32_t* pairs, const int64_t* pair_off, void* stream);

/* Forward-mode derivatives for Gauss-Newton fits (no counterpart in the
 * reference, whose fit is derivative-free, fit_az.py:266-279).
 * smk_fgpa_flux_jac: F per pixel as smk_fgpa_flux (same bits), and per
 * skewer sums[s][4] = forest sums of (F, dF/da, dF/db, dF/dc) at the
 * parameters of its set (skewers.pset required).
 * smk_p1d_jacobian: as smk_p1d_binned at the same flux and <F>, plus the
 * derivatives of the per-bin P sums: partial[s][b][7] = (count, sum k,
 * sum P, sum P^2, sum dP/da, sum dP/db, sum dP/dc), where, with F' the
 * regrouped flux, delta_F = F'/<F> - 1 and t = d delta_F / d theta =
 * dF'/<F> - F' d<F>/<F>^2, dP_j = 2 px / N Re(conj(X_j) T_j), T = DFT(t).
 *   dfmean  device double [sets][3]: d<F>/d(a, b, c) (from the sums above)
 * With tables the forests of a table take two Bluestein DFTs (of delta_F +
 * i t_c and of t_a + i t_b, split by the symmetry of real inputs); the others
 * a direct DFT of the four series by Horner's rule, any length.
 *   pairs, pair_off  NULL, or the pair list of smk_p1d_binned: the per-size
 *          kernels then take their forests from it (each forest of a table
 *          in exactly one pair) instead of scanning the batch. */
int smk_fgpa_flux_jac(const smk_skewers* sk, const float* delta, const float* eta, const double* abcg,
                      int64_t abcg_rows, double* flux, double* sums, void* stream);
int smk_p1d_jacobian(const smk_skewers* sk, const float* delta, const float* eta, const double* abcg,
                     int64_t abcg_rows, const double* flux, const double* fmean, const double* dfmean,
                     double pixel, int regroup, const double* edges, int nbins, int max_forest, double* partial,
                     int32_t* status, const double* tables, const int64_t* tab_off, const int32_t* pairs,
                     const int64_t* pair_off, void* stream);

/* ---------------------------------------------------------- QSO catalogue */
/* Quasar draw of one x slice of the lognormal boxes, draw_qso.py:186-462:
 * per cell, ptot = c(z)(p12 - p23) + p23 of the three fields exp(p_k)^a_k(z)
 * (float32 power as the reference, z at the cell centre), then
 *   cond1  rnd1 < norm * ptot                 (randoms: rnd1 > 1 - 0.006)
 *   cond2  density_max * rnd2 < dn_cell(z) / <ptot>(z)   (randoms: dn_cell(z))
 *   cond3  |ra - ra0| < dra, |dec - dec0| < ddec, zmin < z (RSD) < zmax
 * at a position drawn uniformly in the cell; z from R by the comoving-distance
 * table, the RSD shift R += v_par (1 + z) D'(z) / (D'(0) H0) (:443-452).
 * Uniforms: Philox4x32-10 keyed by `seed`, cell c = (ix*ny + iy)*nz + mz:
 * counter (c_lo, c_hi, 0x71736f63, 0) gives rnd1, rnd2 (53 bits each from
 * word pairs); (c_lo, c_hi, 0x71736f7a, 0) the z offset; (ix, mz, 0x71736f78, 0)
 * and (iy, mz, 0x71736f79, 0) the x and y offsets, shared along a plane's
 * rows and columns as in the reference (:433-436).  The reference draws the
 * same quantities from numpy's legacy RNG plane after plane. */
typedef struct smk_qso_params {
  int32_t nx, ny, nz;            /* slab planes, y and z cells */
  int32_t randoms;               /* 1: random catalogue (no fields) */
  int32_t rsd;                   /* 1: velocity boxes given, cut on the RSD redshift */
  int32_t zfix_on;               /* 1: every redshift is zfix */
  double zfix;
  double dx, dy, dz;             /* cell sides (Mpc/h) */
  const double* x_axis;          /* device [nx], [ny], [nz]: cell-centre coordinates */
  const double* y_axis;
  const double* z_axis;
  double h;                      /* z(R) is looked up at R / h */
  const double* r2z_chi;         /* device [r2z_n]: comoving distance nodes (Mpc), ascending */
  const double* r2z_z;           /* device [r2z_n]: their redshifts */
  int32_t r2z_n;
  const int32_t* r2z_guess;      /* device [r2z_ng]: last node <= r2z_lo + b * r2z_step */
  int32_t r2z_ng;
  double r2z_lo, r2z_step;
  const double* coef_z;          /* device [coef_n]: c(z) nodes (padded to z = 0 and 10) */
  const double* coef_c;
  int32_t coef_n;
  const double* dn_cell;         /* device [dn_n]: QSOs per cell on z = dn_z0 + i * dn_dz */
  int32_t dn_n;
  double dn_z0, dn_dz;
  double sigma[3];               /* std of the three fields over the slab */
  double norm, density_max;
  double zmin, zmax;
  double ra0, dec0;              /* radians */
  double dra, ddec;              /* degrees */
  const double* dg_z;            /* device [dg_n]: dD/dz table, ascending z */
  const double* dg_v;
  int32_t dg_n;
  double dgrowth0;               /* dD/dz at z = 0 */
  const double* r2z_slope;       /* NULL, or device [r2z_n - 1]: (z[j+1] - z[j]) / (chi[j+1] - chi[j]) */
  const double* coef_slope;      /* NULL, or device [coef_n - 1]: the same for c(z) */
} smk_qso_params;

/* One drawn QSO: key = (mz * nx + ix) * ny + iy (the reference's output
 * order), position (Mpc/h), R, redshift without and with RSD, ra, dec (deg). */
typedef struct smk_qso_record {
  int64_t key;
  double x, y, z, r, z_qso, z_qso_rsd, ra, dec;
} smk_qso_record;

/*   p1, p2, p3   float32 [nx][ny][nz] lognormal Gaussian fields (NULL for randoms)
 *   vx, vy, vz   float32 [nx][ny][nz] (used when rsd)
 *   records      device [capacity]; filled in no particular order
 *   count        device int64[1]: number of QSOs drawn (may exceed capacity:
 *                then only `capacity` are stored; call again with a larger
 *                buffer, the draw is deterministic) */
int smk_qso_draw(const smk_qso_params* p, const float* p1, const float* p2, const float* p3, const float* vx,
                 const float* vy, const float* vz, uint64_t seed, smk_qso_record* records, int64_t capacity,
                 int64_t* count, void* stream);

/* ------------------------------------------------------------ P(k) tables */
/* A cubic spline as piecewise cubics: on [x[i], x[i+1]) (positive length),
 * P(k) = ((c[i][0] t + c[i][1]) t + c[i][2]) t + c[i][3], t = k - x[i];
 * guess[b] is an interval with x[guess[b]] <= g_lo + b * g_step (a start for
 * the search).  All arrays are device arrays of doubles (guess: int32). */
typedef struct smk_pk_spline {
  const double* x;     /* [n_int + 1] breakpoints, ascending */
  const double* c;     /* [n_int][4] */
  int32_t n_int;
  const int32_t* guess;
  int32_t n_guess;
  double g_lo, g_step;
} smk_pk_spline;

/* interpolate_pk.py:17-26, 63-149: table[i][j][l] = float32
 * sqrt(float32(max(P(k), 0)) / vcell), k = sqrt(kx[i]^2 + ky[j]^2 + kz[l]^2)
 * in float32 (correctly rounded, as numpy), P the spline at k in FP64.
 *   kx, ky, kz   device float32 axes [nxs], [ny], [nzh]
 *   table        device float32 [nxs][ny][nzh] */
int smk_pk_table(const smk_pk_spline* spline, const float* kx, const float* ky, const float* kz, int nxs, int ny,
                 int nzh, float vcell, float* table, void* stream);

/* p1d_missing.py:12-37,120-140 (P_RSD, D_Prats, PW2cut) and
 * powerspectrum.py:37-49 (P_1D_RSD): for every k_par[i],
 *   out[i] = trapz_j( k_perp[j] * P3(k_par[i], k_perp[j]) ) / (2 pi)
 * over the ascending grid k_perp (trapezoid rule, numpy.trapz), with
 * k = sqrt(k_par^2 + k_perp^2), mu = k_par / max(k, 1e-15), P(k) = max(spline(k), 0)
 * and P3 = (1 + beta mu^2)^2 F(k, mu) P(k), F by `model`:
 *   0  1                                            (linear, P_RSD)
 *   1  exp((q1 D2 + q2 D2^2)(1 - (k/kv)^av mu^bv) - (k/kp)^2),
 *      D2 = k^3 P(k) G^2 / (2 pi^2)                 (P_RSD_Prats)
 *   2  exp(-dx^2 k^2) for k < kny, else 0           (PW2cut)
 * prats = {q1, q2, kp, kv, av, bv, G}; dx, kny used by model 2.
 * One block per k_par, FP64, deterministic.  out: device double [n_par]. */
int smk_p1d_rsd_rows(const smk_pk_spline* spline, const double* k_par, int n_par, const double* k_perp, int n_perp,
                     double beta, int model, const double* prats, double dx, double kny, double* out,
                     void* stream);

#ifdef __cplusplus
}
#endif
#endif /* SACLAYMOCKS_HIP_H */
