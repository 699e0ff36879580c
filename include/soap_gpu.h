/*
 * soap_gpu.h -- C ABI of libsoapgpu, the MI355X (gfx950) SOAP-GPU engine.
 *
 * The reference (YinanZhao21/SOAP_GPU) has no library interface: its stages are
 * four executables called with positional argv and exchanging raw .bin files
 * (SOAP_code/run_all_SOAP.py:226-250).  Each entry point below replaces one
 * stage (or one step inside a stage); the comment on each names the reference
 * code it stands in for.  The CLI shims built from soap_gpu_amd/driver keep
 * the reference's argv and files on top of this ABI.
 *
 * Conventions
 *  - All arithmetic is FP64.  Host arrays are caller-owned, contiguous,
 *    little-endian double (int32 for the region map), sizes in elements.
 *  - Device state lives in an opaque soap_ctx owned by the library; one ctx per
 *    host thread.  Each ctx owns one HIP stream; entry points enqueue work on it
 *    and return; soap_sync waits.  Calls that return data to the host
 *    synchronise the stream themselves.
 *  - Every call returns SOAP_OK (0) or a negative soap_status; the message of
 *    the last failure is soap_last_error(ctx).
 *  - Seed cubes are [n_mu][n] row-major (row k = mu bin k), the layout of the
 *    reference's seed-cube .bin files (SOAP_integration.cu:504-580).
 *  - Spectra produced per phase are stored [phase][type][n] with type
 *    SOAP_FLUX, SOAP_BCONV, SOAP_TOT (file names of SOAP_integration.cu:802-816).
 */
#ifndef SOAP_GPU_H
#define SOAP_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct soap_ctx soap_ctx;

enum soap_status {
  SOAP_OK = 0,
  SOAP_ERR_ARG = -1,     /* bad argument (null pointer, size, range) */
  SOAP_ERR_STATE = -2,   /* a required earlier step was not done */
  SOAP_ERR_HIP = -3,     /* HIP runtime error */
  SOAP_ERR_NOMEM = -4,   /* device allocation failed */
  SOAP_ERR_GRID = -5,    /* wavelength grid not of the declared kind */
  SOAP_ERR_LIMIT = -6,   /* size beyond a compiled limit */
  SOAP_ERR_IO = -7       /* file missing or unreadable */
};

enum soap_grid_kind { SOAP_GRID_LINEAR = 0, SOAP_GRID_LOG = 1 };

/* seed cubes, by the reference's file roles (run_all_SOAP.py:94-100) */
enum soap_cube {
  SOAP_CUBE_QUIET = 0,      /* new_noconti_convec_T_eff_{T}.bin       (normalised quiet)  */
  SOAP_CUBE_QUIET_SED = 1,  /* new_conti_convec_T_eff_{T}.bin         (SED quiet)         */
  SOAP_CUBE_SPOT = 2,       /* new_noconti_convec_T_eff_{T-dTs}.bin   (normalised spot)   */
  SOAP_CUBE_SPOT_SED = 3,   /* new_conti_T_eff_{T-dTs}.bin            (SED spot, row 0)   */
  SOAP_CUBE_FAC = 4,        /* new_noconti_convec_T_eff_{T+dTf}.bin   (normalised facula) */
  SOAP_CUBE_FAC_SED = 5     /* new_conti_T_eff_{T+dTf}.bin            (SED facula, row 0) */
};

enum soap_spec_type { SOAP_FLUX = 0, SOAP_BCONV = 1, SOAP_TOT = 2 };

/* star parameters: SOAP_initialize argv 2-8 and 20-23 (SOAP_initialize.cu:513-537) */
typedef struct {
  double prot_days;  /* argv[2] period */
  double w_ratio;    /* argv[3] differential rotation ratio omega1/omega0 */
  double incl_deg;   /* argv[4] inclination */
  double radius_km;  /* argv[5] stellar radius */
  double u1, u2;     /* argv[6,7] quadratic limb darkening */
  double lb[4];      /* argv[20..23] limb-brightening polynomial, constant term first */
  int32_t grid;      /* argv[8] disk grid size */
} soap_star;

/* activity parameters: SOAP_integration argv 2,4,5,7 (SOAP_integration.cu:469-497) */
typedef struct {
  int32_t nrho;        /* boundary points per region */
  int32_t t_star;      /* K */
  int32_t t_diff_spot; /* K */
  int32_t t_diff_fac;  /* K */
  int32_t solar_mode;  /* 1: facula intensity by Planck ratio (SOAP_integration_solar.cu:433) */
} soap_activity;

/* one active region: one line "lat lon size type" of maps{id}_for_all.txt
 * (SOAP_integration.cu:669-682); type 0 = spot, 1 = facula */
typedef struct {
  double lat_deg, lon_deg, size, type;
} soap_region;

/* ------------------------------------------------------------ context */
int soap_ctx_create(int device, soap_ctx **out);
void soap_ctx_destroy(soap_ctx *ctx);
const char *soap_last_error(const soap_ctx *ctx);
int soap_sync(soap_ctx *ctx);
/* library version string */
const char *soap_version(void);

/* ------------------------------------------------------------ inputs */
/* wavelength grid (the solar_wave .bin, SOAP_initialize.cu:554-556).  kind LOG
 * requires lambda_j = lambda_0 exp(j h) to 1e-12 relative (else SOAP_ERR_GRID);
 * step is the reference's step_size argv (round(l[3]-l[2],5),
 * run_all_SOAP.py:124) and is used for LINEAR grids only. */
int soap_set_wave(soap_ctx *ctx, int kind, const double *wave, int64_t n, double step);
/* mu boundaries (mu_boundary_array.bin, SOAP_initialize.cu:567-574); n_mu = n_bounds + 1 */
int soap_set_mu_bounds(soap_ctx *ctx, const double *bounds, int32_t n_bounds);
/* seed cube [n_mu][n] host -> device (the fread + cudaMemcpy of SOAP_integration.cu:504-580);
 * SOAP_ERR_ARG if a value is not finite */
int soap_set_cube(soap_ctx *ctx, int which, const double *cube, int32_t n_mu);
/* seed cube already resident on the device (not copied; must outlive its use) */
int soap_set_cube_device(soap_ctx *ctx, int which, const double *dev_cube, int32_t n_mu);

/* ------------------------------------------------------------ a1: disk grid */
/* itot's per-cell loop (SOAP_initialize.cu:297-395) as one kernel: mu, projected
 * velocity, limb darkening, limb brightening and mu bin for every cell */
int soap_disk_init(soap_ctx *ctx, const soap_star *star);
/* maps to host, any pointer may be NULL (the five *_map.bin, SOAP_initialize.cu:399-422) */
int soap_get_maps(soap_ctx *ctx, double *mu, double *delta, double *limb, int32_t *region, double *lb);
/* maps from host (the map .bin files read by SOAP_integration.cu:603-635) */
int soap_set_maps(soap_ctx *ctx, const soap_star *star, const double *delta, const double *limb,
                  const int32_t *region, const double *lb);

/* ------------------------------------------------------------ a2: quiet star */
/* quiet-star sums of the normalised and SED cubes (fast_shift + fast_interp_new +
 * fast_sum over every disk cell, SOAP_initialize.cu:117-176,379-381; run twice by
 * run_all_SOAP.py:226-234) */
int soap_quiet_star(soap_ctx *ctx);
/* which: 0 = normalised (_quiet_sun_spec.bin), 1 = SED (_SED_quiet_sun_spec.bin) */
int soap_get_quiet(soap_ctx *ctx, int which, double *out);
int soap_set_quiet(soap_ctx *ctx, int which, const double *q);

/* ------------------------------------------------------------ a5/a6: activity */
/* phases (rotations) and per-phase region lists (phase file + final_info.txt +
 * maps{id}_for_all.txt, SOAP_integration.cu:637-682) */
int soap_set_activity(soap_ctx *ctx, const soap_activity *act, int32_t n_phase, const double *phase,
                      const int32_t *region_count, const soap_region *regions);
/* per-phase active-region deltas flux/bconv/tot for every phase (the phase loop of
 * SOAP_integration.cu:661-845: spot_init_gpu/spot_phase_gpu/spot_area/spot_scan +
 * spot_integrate) */
int soap_integrate(soap_ctx *ctx);
/* raw spectra of one phase to host (_Spec_{flux,bconv,tot}_{id}_.bin); NULL skips */
int soap_get_raw(soap_ctx *ctx, int32_t phase, double *flux, double *bconv, double *tot);
/* declare n_phase spectrum slots without region geometry, for callers that
 * bring their own raw spectra (the phase loop of SOAP_resolution.cu:379-428) */
int soap_alloc_phases(soap_ctx *ctx, int32_t n_phase);
/* raw spectra of one phase from host (SOAP_resolution.cu:398-412) */
int soap_set_raw(soap_ctx *ctx, int32_t phase, const double *flux, const double *bconv, const double *tot);
/* diagnostic temperature / binary maps of one phase (_temp_map_, _spot_binary_map_,
 * SOAP_integration.cu:832-840) */
int soap_phase_maps(soap_ctx *ctx, int32_t phase, double *temp_map, double *bi_map);

/* ------------------------------------------------------------ a7: resolution */
/* Gaussian instrumental convolution + spec_norm for every phase and type
 * (lower_resolution_gpu, SOAP_resolution.cu:235-310) and for the two quiet
 * frames (SOAP_initialize.cu:446-505).  sigma_ms = c/R/(2 sqrt(2 ln 2));
 * quiet_weight = the wavelength_step_median argv of SOAP_initialize (LINEAR only). */
int soap_resolution(soap_ctx *ctx, double sigma_ms, double quiet_weight);
/* lowered spectrum: phase in [0, n_phase) and type, or phase = -1 for the quiet
 * frames (type SOAP_BCONV -> _quiet_sun_lower_spec, else _SED_quiet_sun_lower_spec) */
int soap_get_lower(soap_ctx *ctx, int32_t phase, int type, double *out);
/* [min, max] of spec_norm for one spectrum (diagnostic) */
int soap_get_norm(soap_ctx *ctx, int32_t phase, int type, double *minmax2);

/* ------------------------------------------------------------ a8: CCF */
/* mask lines after the driver's selection (run_all_SOAP.py:285-290), velocity grid
 * (run_all_SOAP.py:262) and the pixel trim (cutoff 200 then 10, :259-281) */
int soap_set_mask(soap_ctx *ctx, const double *line_l, const double *line_w, int32_t n_lines,
                  const double *vgrid, int32_t n_v, int32_t trim);
/* CCF of every frame and type (calculate_CCF_1/2 of CCF_func.py:5-91), normalised by
 * its max and appodised (run_all_SOAP.py:306-314,368-370).  Frames 0..n_phase-1
 * are the phases, frame n_phase the quiet star (run_all_SOAP.py:351-362). */
int soap_ccf(soap_ctx *ctx);
/* CCFs to host: out[(frame*3+type)*n_v + iv] */
int soap_get_ccf(soap_ctx *ctx, double *out);

/* ------------------------------------------------------------ a9: fit */
/* Gaussian fit + bisector of every CCF (BIS_FIT2.cpp:97-320, one per frame and type) */
int soap_fit(soap_ctx *ctx);
/* results, row = frame*3+type: par[row*5 + {c,|k|,span,v0,fwhm}], model[row*n_v],
 * depth[row*91], bis[row*91]; NULL skips */
int soap_get_fit(soap_ctx *ctx, double *par, double *model, double *depth, double *bis);
/* standalone batched fit of host CCFs (the BIS_FIT2 executable, one call per CCF
 * in run_all_SOAP.py:316-394): v[n_v], y[n_ccf*n_v] */
int soap_fit_ccf_host(soap_ctx *ctx, const double *v, int32_t n_v, const double *y, int32_t n_ccf,
                      double *par, double *model, double *depth, double *bis);
/* the same with per-point errors (CCF_error.txt, BIS_FIT2.cpp:114-119): residuals
 * (fit - y)/err and Jacobian rows divided by err (BIS_FIT2.cpp:49,74-77);
 * err[n_ccf*n_v], NULL = unit errors.  Errors must be finite and non-zero
 * (SOAP_ERR_ARG otherwise; the reference would divide by zero).  status[n_ccf]
 * (may be NULL) receives the solver's outcome per CCF, in GSL's codes: 0 = the
 * xtol or gtol test ended the fit, 27 (GSL_ENOPROG) = no step reduced the
 * residual (a flat or degenerate CCF; the outputs hold the starting point),
 * 11 (GSL_EMAXITER) = 10000 iterations.  The reference ignores this outcome
 * (BIS_FIT2.cpp:191) and so does the call: it returns SOAP_OK either way. */
int soap_fit_ccf_host_err(soap_ctx *ctx, const double *v, int32_t n_v, const double *y, const double *err,
                          int32_t n_ccf, double *par, double *model, double *depth, double *bis, int32_t *status);
/* soap_fit_ccf_host_err without any device allocation: no counterpart in the
 * reference; for callers that cannot allocate device memory (tight serving or
 * fitting loops, harnesses that forbid hipMalloc).  Inputs and outputs are the
 * caller's host arrays as above; the device staging (inputs, fit parameters,
 * model, depth, bisector) lives in a fixed 16 MiB array per GPU held by the
 * library until the process ends, so a call needs
 *   n_v + n_ccf (n_v [y] + n_v [err, if given] + 5 + n_v + 182) doubles
 * (each part rounded up to 32; + n_ccf / 2 with status) <= 2^21, else
 * SOAP_ERR_LIMIT and nothing runs.
 * Every call on one GPU uses that GPU's one array (each process holds its own)
 * and it keeps the last call's contents: calls that overlapped on one GPU would
 * race on it, so the library serialises them with a lock (a call waits for the
 * one before it) and each returns after its results are on the host.  Results
 * equal soap_fit_ccf_host_err's bit for bit. */
int soap_fit_ccf_host_noalloc(soap_ctx *ctx, const double *v, int32_t n_v, const double *y, const double *err,
                              int32_t n_ccf, double *par, double *model, double *depth, double *bis,
                              int32_t *status);

/* ------------------------------------------------------------ seed preprocessing: bisector injection */
/* The convective-blueshift injection of CB_injection_faculae.py:191-244 on one
 * spectrum (host arrays wave[n], flux[n] -> out[n]): each pixel moves by the
 * velocity v(flux) (Delta_wavelength, :119-123), the moved samples are sorted by
 * wavelength and resampled linearly (scipy interp1d) onto wave, zero outside the
 * open range of the moved wavelengths.
 *   mode 0, BIS_inject_spectral_quiet (:227-244): v = poly(up)(flux) * scale;
 *   mode 1, BIS_inject_spectral_active (:191-224): v = (poly(up)(flux) + velo_shift) * scale
 *     if 0.5 < flux < dep_cut, (poly(low)(flux) + velo_shift) * scale if flux <= 0.5,
 *     (vel_cut + velo_shift) * scale if flux > dep_cut, unmoved otherwise.
 * poly(c) = numpy.poly1d(c): coefficients highest degree first, 1..16 of them. */
typedef struct soap_bis_inject_params {
  int32_t mode, n_up, n_low;
  double up[16], low[16];
  double dep_cut, vel_cut, velo_shift, scale;
} soap_bis_inject_params;
int soap_bis_inject(soap_ctx *ctx, const soap_bis_inject_params *p, const double *wave, const double *flux, int64_t n,
                    double *out);

/* ------------------------------------------------------------ seed preprocessing: local PHOENIX spectra */
/* phoenix_extraction.py, local files only (it never downloads).  Host-only helpers
 * (no GPU needed):
 * soap_phoenix_bracket: the two library temperatures around temp (:136-146, the
 *   reference's temp_array including its entry 48 = 6800 K); SOAP_ERR_LIMIT past
 *   either end of the grid (where the reference fails or wraps to the last entry). */
int soap_phoenix_bracket(double temp, double *t_low, double *t_high);
/* the library grid values nearest to logg, [Fe/H], [alpha/Fe] (:84-104,126-131) */
int soap_phoenix_grid(double logg, double feh, double alpha, double *logg_val, double *feh_val, double *alpha_val);
/* PHOENIX_name (:36-54): the library subdirectory ("Z-0.0") and file name
 * ("/lte05700-4.50-0.0.PHOENIX-ACES-AGSS-COND-2011-HiRes.fits") */
int soap_phoenix_name(int32_t temp, double logg, double feh, double alpha, char *subdir, int32_t subdir_cap,
                      char *name, int32_t name_cap);
/* the primary image of a FITS file (astropy's fits.getdata for the PHOENIX files):
 * BITPIX -64 / -32 / 32 / 16, all NAXISn values multiplied, BSCALE / BZERO applied.
 * With out NULL or cap < n, only n and bitpix are returned.  SOAP_ERR_IO if the
 * file cannot be read, SOAP_ERR_ARG if it is not such a FITS file. */
int soap_fits_read(const char *path, double *out, int64_t cap, int64_t *n, int32_t *bitpix);
/* On the device (:170-185): for the vacuum wavelengths wave_vac[n] and the spectra
 * of the bracketing temperatures, the pixels with wave_vac in [wmin - 500, wmax +
 * 500] and air wavelength (vac2air, Birch and Downs 1994, :13-22) in [wmin, wmax],
 * in order: wave_air and flux = low + (temp - t_low) / (t_high - t_low) (high -
 * low), the difference taken in float32 when f32 is set (the library's float32
 * spectra).  Writes *n_out <= n values to wave_air and flux. */
typedef struct soap_phoenix_params {
  double temp, t_low, t_high, wmin, wmax;
  int32_t f32;
} soap_phoenix_params;
int soap_phoenix_interp(soap_ctx *ctx, const soap_phoenix_params *p, const double *wave_vac, const double *spec_low,
                        const double *spec_high, int64_t n, double *wave_air, double *flux, int64_t *n_out);

/* ------------------------------------------------------------ whole series */
/* disk init + quiet star + integrate + resolution + CCF + fit, all on device */
int soap_run_series(soap_ctx *ctx, const soap_star *star, double sigma_ms, double quiet_weight);

/* ------------------------------------------------------------ instrumentation */
/* per-stage timing with HIP events on the ctx stream: enable, then read the
 * accumulated milliseconds of each stage (names by index via soap_stage_name) */
int soap_set_timing(soap_ctx *ctx, int enable);
int soap_get_stage_ms(soap_ctx *ctx, double *ms, int32_t n_stage);
const char *soap_stage_name(int32_t stage);
int32_t soap_num_stages(void);
/* HBM bytes of the last soap_integrate: bytes_min = seed rows touched by any
 * phase (read once) + continua + 3 outputs per phase; bytes_phase_sum = the same
 * counted per phase, sum_psi 8 n (R_psi + continua + 3); rows_sum = sum_psi R_psi */
int soap_get_integrate_stats(soap_ctx *ctx, double *bytes_min, double *bytes_phase_sum, int64_t *rows_sum);
/* Launch shape of the active-phase FIR of soap_integrate (log grids, non-solar):
 * (4, -7) = a 64-pixel tile and every phase per block, all seed rows of the tile
 * staged once, taps broadcast across lanes (the automatic choice); (4, -4) = the
 * same with the taps in scalar registers; (4, 0) = four phases per block, one
 * pixel per thread, rows staged per mu bin; (G, -2) = G in {1, 2, 4} phases per
 * block reading the rows from global memory; (G, B) with B > 0 = G phases per
 * block, B pixels per thread (register-blocked), one of (1, 3), (1, 5), (2, 3),
 * (2, 1), (4, 1); (-1, -1) = automatic.  Results agree to rounding across shapes
 * (bit for bit among the one-pixel forms).  No counterpart in the reference
 * (tuning support). */
int soap_set_fir_shape(soap_ctx *ctx, int32_t phases_per_block, int32_t pixels_per_thread);
/* ------------------------------------------------------------ a11: differentiable path
 * No counterpart in the reference (a differentiable forward model with its FP64
 * adjoint, for fitting).  Model on the context's log-uniform grid, mu bounds
 * and mask (soap_set_wave, soap_set_mu_bounds, soap_set_mask): disk grid G,
 * rigid rotation delta = vsini y, quadratic limb darkening L, per-phase spot and
 * facula filling-factor maps phi_p, psi_p (M x M over [-1, 1]^2, bilinear)
 * blending the quiet seed rows Q, the spot rows S scaled by the spot contrast
 * and the facula rows Fa scaled by the facula contrast and the reference's
 * facula intensity factor 1 + LB(mu) (SOAP_integration.cu:317-323,
 * LB = lb0 + lb1 mu + lb2 mu^2 + lb3 mu^3 as in SOAP_initialize.cu:367):
 *   I_p = sum_c L_c [(1 - phi - psi) Shift(Q_k) + contrast phi Shift(S_k)
 *                    + contrast_fac (1 + LB(mu_c)) psi Shift(Fa_k)],
 *   F_p = instrumental convolution of I_p, CCF_p = raw mask CCF of F_p,
 *   (RV_p, FWHM_p) = Gaussian fit of CCF_p polished to the least-squares optimum,
 *   BIS_p = bisector span of CCF_p at that Gaussian (BIS_FIT2.cpp:214-268),
 *   loss = w_spec sum (F_p - T_p)^2 over the trimmed pixels
 *        + w_rv sum (RV_p - rv_target_p)^2 + w_fwhm sum (FWHM_p - fwhm_target_p)^2
 *        + w_bis sum (BIS_p - bis_target_p)^2.
 * The facula population is optional: without facula seeds and maps psi = 0.
 * Maps are filling factors in [0, 1].  The per-cell weights enter fixed-point tap
 * sums: G^2 max|L| max(3, |contrast|, |contrast_fac| max|1 + LB|) must stay below
 * 2^30 (SOAP_ERR_LIMIT from soap_diff_forward, a NaN loss in graph replays).
 * Array arguments named d_* are device (HIP) pointers: seeds [n_mu][n], maps
 * [P][M][M], target [P][n]; they must stay valid until soap_diff_backward. */
typedef struct soap_diff_params {
  double vsini_kms, u1, u2, contrast, contrast_fac;
} soap_diff_params;
/* lb: the facula intensity polynomial (4 values, constant term first), NULL = zeros */
int soap_diff_setup(soap_ctx *ctx, int32_t grid, int32_t n_phase, int32_t map_res, double sigma_ms,
                    const double *lb);
/* d_seed_f and d_fac_maps: both NULL (no faculae) or both set.  rv_target,
 * fwhm_target, bis_target [P], loss (scalar), rv, fwhm, bis [P]: host arrays
 * (rv, fwhm, bis may be NULL) */
int soap_diff_forward(soap_ctx *ctx, const soap_diff_params *prm, const double *d_seed_q, const double *d_seed_s,
                      const double *d_seed_f, const double *d_maps, const double *d_fac_maps, const double *d_target,
                      const double *rv_target, const double *fwhm_target, const double *bis_target, double w_spec,
                      double w_rv, double w_fwhm, double w_bis, double *loss, double *rv, double *fwhm, double *bis);
/* gradients of the last forward's loss: grad5 (host) = d/d(vsini, u1, u2, contrast,
 * contrast_fac) (the last is 0 without faculae); d_grad_q, d_grad_s, d_grad_f
 * [n_mu][n] and d_grad_maps, d_grad_fac_maps [P][M][M] on the device (overwritten;
 * d_grad_f and d_grad_fac_maps are required when the forward had faculae and
 * ignored otherwise) */
int soap_diff_backward(soap_ctx *ctx, double *grad5, double *d_grad_q, double *d_grad_s, double *d_grad_f,
                       double *d_grad_maps, double *d_grad_fac_maps);
/* host copies of the last forward: what 0 = F [P][n], 1 = raw CCF [P][n_v], 2 = (c, k, RV, FWHM) [P][4],
 * 3 = BIS span [P] */
int soap_diff_get(soap_ctx *ctx, int32_t what, double *out);

/* Algorithmic work of one run of a stage (index as soap_stage_name) on the
 * current inputs: HBM bytes that must move (each input read once, each output
 * written once) and FP64 flops (2 per multiply-add the algorithm needs, padding
 * and skipped zero taps excluded).  Stages without a model return 0 / 0.
 * No counterpart in the reference (measurement support). */
int soap_get_stage_work(soap_ctx *ctx, int32_t stage, double *bytes, double *flops);


/* ------------------------------------------------------------ a11: batched instances
 * No counterpart in the reference (which runs one star per process, stage by
 * stage).  A batch runs the per-phase step (integrate + resolution + CCF + fit)
 * of many uneven instances -- each a prepared soap_ctx on the same device with
 * its own grid size, wavelength grid, phases and regions -- with one persistent
 * launch per stage covering every instance, optionally replayed as a HIP graph.
 * Instances: log-uniform grid, non-solar activity with >= 1 phase, mask set,
 * disk_init and quiet_star done, n_mu <= 64.  soap_batch_create runs each
 * instance once through the plain path (allocating its buffers); results of a
 * step are read through each instance's soap_get_* calls after soap_batch_sync.
 * The contexts must outlive the batch and must not be changed while it exists.
 * On failure soap_batch_create still returns the batch in *out (if allocated)
 * so that soap_batch_last_error can be read; destroy it either way. */
typedef struct soap_batch soap_batch;
int soap_batch_create(soap_ctx *const *ctxs, int32_t n_inst, const double *sigma_ms, const double *quiet_weight,
                      soap_batch **out);
void soap_batch_destroy(soap_batch *b);
const char *soap_batch_last_error(const soap_batch *b);
/* enqueue one step on the batch's stream; use_graph != 0 replays a captured HIP
 * graph of the step (captured on first use after any shape change) */
int soap_batch_step(soap_batch *b, int32_t use_graph);
int soap_batch_sync(soap_batch *b);
/* one eager step with HIP events around each stage launch (synchronous) */
int soap_batch_stage_ms(soap_batch *b, double *ms, int32_t n_stage);
const char *soap_batch_stage_name(int32_t stage);
int32_t soap_batch_num_stages(void);
/* launch shapes: disk cells per taps block; FIR phases per block (fir_g) and
 * pixels per thread (fir_b, 0 = one-pixel form; (g, b) one of (2,0) (4,0) (6,0)
 * (1,3) (1,5) (2,3), or (4,-7) = the tile-stationary form of soap_set_fir_shape,
 * the default, which adds the work of stage "fir_prog"); fir_pre = 1: the
 * one-pixel form prefetches the next bin's seed rows into registers; per stage
 * the persistent block count (0 = one block per work item, else a multiple of
 * 8) and the work items */
int soap_batch_get_config(soap_batch *b, int32_t *taps_cells, int32_t *fir_g, int32_t *fir_b, int32_t *fir_pre,
                          int64_t *grid, int64_t *work, int32_t n_stage);
int soap_batch_set_config(soap_batch *b, int32_t taps_cells, int32_t fir_g, int32_t fir_b, int32_t fir_pre,
                          const int64_t *grid, int32_t n_stage);
/* autotuner: coordinate search over the shapes above, scored by stage time
 * (best of reps eager steps per candidate) */
int soap_batch_tune(soap_batch *b, int32_t reps);

/* ------------------------------------------------------------ a11: differentiable instances in one graph
 * No counterpart in the reference.  Forward + adjoint of many instances (each a
 * soap_ctx after soap_diff_setup) with every input, parameter and result in
 * device memory, so that one evaluation is one HIP graph launch:
 *   d_params [5] = (vsini, u1, u2, contrast, contrast_fac), d_q / d_s / d_f [n_mu][n],
 *   d_maps / d_fac_maps [P][M][M], d_target [P][n], d_rv_t / d_fw_t / d_bis_t [P]
 *   ->  d_loss [1], d_grad5 [5], d_grad_q / d_grad_s / d_grad_f [n_mu][n],
 *   d_grad_maps / d_grad_fac_maps [P][M][M].
 * The facula pointers d_f, d_fac_maps, d_grad_f, d_grad_fac_maps are all NULL (an
 * instance without faculae) or all set.
 * The tap range is fixed at bind time by vsini_max; an evaluation with vsini
 * outside (0, vsini_max] returns loss = NaN.  The graph is captured on the first
 * step after a bind and replayed afterwards (values at the bound addresses may
 * change between steps, the addresses may not).  A per-call soap_diff_forward
 * on a bound context invalidates its binding (the step refuses it). */
typedef struct soap_diffbatch soap_diffbatch;
int soap_diff_batch_create(soap_ctx *const *ctxs, int32_t n_inst, soap_diffbatch **out);
void soap_diff_batch_destroy(soap_diffbatch *b);
const char *soap_diff_batch_last_error(const soap_diffbatch *b);
int soap_diff_batch_bind(soap_diffbatch *b, int32_t inst, double vsini_max, const double *d_params,
                         const double *d_seed_q, const double *d_seed_s, const double *d_seed_f, const double *d_maps,
                         const double *d_fac_maps, const double *d_target, const double *d_rv_target,
                         const double *d_fwhm_target, const double *d_bis_target, double w_spec, double w_rv,
                         double w_fwhm, double w_bis, double *d_loss, double *d_grad5, double *d_grad_q,
                         double *d_grad_s, double *d_grad_f, double *d_grad_maps, double *d_grad_fac_maps);
/* enqueue forward + loss of every instance, and the adjoint when with_adjoint != 0
 * (without it d_grad* are left as they are); use_graph != 0 replays a captured
 * graph (one per with_adjoint value) */
int soap_diff_batch_step(soap_diffbatch *b, int32_t use_graph, int32_t with_adjoint);
int soap_diff_batch_sync(soap_diffbatch *b);

#ifdef __cplusplus
}
#endif
#endif
