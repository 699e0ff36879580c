// Host side of libsoapgpu: context, validation, per-phase region geometry and
// stage orchestration on one HIP stream.  See include/soap_gpu.h for the ABI.
#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>

#include <algorithm>
#include <cmath>
#include <map>
#include <mutex>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/soap_gpu.h"
#include "soap_internal.h"

using namespace soap;

namespace {

enum Stage {
  ST_DISK = 0,
  ST_TAPS_QUIET,
  ST_FIR_QUIET,
  ST_TAPS_ACTIVE,
  ST_FIR_ACTIVE,
  ST_CONV_QUIET,
  ST_CONV,
  ST_MINMAX,
  ST_NORM,
  ST_MASK,
  ST_CCF,
  ST_CCF_REDUCE,
  ST_FIT,
  ST_COUNT
};
const char *kStageNames[ST_COUNT] = {"disk_init", "taps_quiet", "fir_quiet",  "taps_active", "fir_active",
                                     "conv_quiet", "conv",       "minmax",     "norm",        "mask",
                                     "ccf",        "ccf_reduce", "fit"};

struct DevBuf {
  void *p = nullptr;
  size_t bytes = 0;
};

}  // namespace

struct soap_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  // grid
  int64_t n = 0;
  int kind = -1;
  double step = 0, h = 0, expm1_h = 0;
  std::vector<double> wave;
  double *d_wave = nullptr;
  // mu
  int n_mu = 0;
  std::vector<double> mu_bounds;
  double *d_mub = nullptr;
  // cubes
  double *d_cube[6] = {};
  bool cube_owned[6] = {};
  int cube_nmu[6] = {};
  // star + maps
  soap_star star{};
  bool have_maps = false;
  int G = 0;
  double *d_mu = nullptr, *d_delta = nullptr, *d_limb = nullptr, *d_lbm = nullptr, *d_t = nullptr;
  int32_t *d_region = nullptr, *d_m = nullptr;
  int dmin = 0, nd = 0;
  // quiet
  double *d_quiet = nullptr;  // [2][n]
  bool have_quiet = false;
  // activity
  soap_activity act{};
  int P = 0;
  std::vector<double> phases;
  std::vector<int32_t> reg_count;
  std::vector<soap_region> regions;
  std::vector<RegionDev> geo;
  std::vector<int32_t> geo_off, box;
  RegionDev *d_geo = nullptr;
  int32_t *d_geo_off = nullptr, *d_box = nullptr;
  int max_col_blocks = 0;
  bool ccf_fused = false;     // the last CCF ran with spec_norm in its launch (k_norm_ccf)
  int fir_g = 0, fir_b = -1;  // FIR launch shape (phases per block, pixels per thread); -1 = automatic
  double tap_wmax = 0;        // bound on |limb| and |limb (1 + LB)| over the disk (fixed-point tap capacity)
  double solar_leb = 1;       // max over the facula temperatures of sum_q |l_q(T)| (solar Lagrange weights)
  // differentiable path
  struct {
    bool ready = false, have_fwd = false;
    int G = 0, P = 0, M = 0, dmin = 0, nd = 0, res_blocks = 0;
    double sigma = 0, w_spec = 0;
    soap_diff_params prm{};
    double lb[4] = {0, 0, 0, 0};  // facula intensity polynomial, constant term first
    int ns = 2;                   // tap sets of the last forward: 2, or 3 with the facula population
    const double *q = nullptr, *s = nullptr, *f = nullptr, *maps = nullptr, *fmaps = nullptr, *target = nullptr;
    DevBuf ck, cm, ct, cL, cmu, cdt, acc, taps, krange, I, F, res_part, ccf, par, model, depth, bis, iters, th,
        dth, gy, GF, GI, C, Cpart, cell_part;
    DevBuf bisd;  // [P][n_v + kBisHead]: span, loss coefficient, dspan/dtheta, dspan/dy
    DevBuf tg;    // per-call targets [3][P] (RV, FWHM, BIS span), the loss [1], the scalar gradients [5]
    DevBuf mt_ptr, mt_w, mt_col, mt_val;  // the mask in sliced pixel-major form (adjoint of the CCF)
    bool mt_ready = false;
  } diff;
  // linear grids: staged source range per tile, bin-sorted disk cells, active cell
  // lists, and their shift groups with per-segment group counts
  int lin_off_lo = 0, lin_span = 0;
  bool lin_solid = false;  // delta constant along each disk column (disk_init maps, omega1 = 0)
  DevBuf qcells, qbin_off, acells, abin_off, alist_off;
  DevBuf qgroups, qg_n, agroups, ag_n, lcnt, qpart, qcolw, qcolf, qcols;
  // solar mode: Lagrange nodes over the facula temperature range
  int solar_k = 0, t_norm_fac = 0;
  double solar_t[kMaxSolarNodes] = {}, solar_w[kMaxSolarNodes] = {};
  double *d_raw = nullptr;  // [P][3][n]
  bool have_raw = false;
  // tap scratch
  DevBuf acc, taps, krange, kflags, bins, nbins;
  DevBuf taps_g;  // phase-interleaved active taps [n_mu][nd][taps_g_pad(P)][4]
  DevBuf prog_t, prog_x, prog_n, prog_u;  // entry streams of the tile-stationary FIR
  DevBuf inj;      // bisector injection: wave, flux, moved, sorted x, sorted y, out [6][n], partials, sort scratch
  DevBuf phx;      // PHOENIX interpolation: wave, low, high, air, flux, out air, out flux [7][n], flags, count, scratch
  // quiet-star tap scratch of its own, so that the quiet star can run beside the
  // active phases (soap_run_series puts it on a second stream)
  DevBuf qacc, qtaps, qkrange, qkflags, qbins, qnbins;
  hipStream_t aux = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;

  // resolution
  double *d_lower = nullptr;  // [(P+1)][3][n], after spec_norm
  double *d_lraw = nullptr;   // [(P+1)][3][n], the convolution output before spec_norm
  DevBuf part, mm, ctap, qsel, variant;
  int res_dlo = 0, res_ntp = 0, res_wid = -1;
  bool have_lower = false;
  // mask / ccf / fit
  int n_lines = 0, n_v = 0, trim = 0;
  double *d_line_l = nullptr, *d_line_w = nullptr, *d_vgrid = nullptr, *d_appod = nullptr, *d_w_sorted = nullptr;
  DevBuf m_start, m_len, m_val, m_alive, sorted, rank, ccf_part;
  DevBuf wsum;  // [n_v] live mask weight per velocity (the CCF of a spectrum of ones)
  DevBuf fp;    // [n_lines] int2 line footprints (first trimmed pixel, length)
  int ccf_cap = 0;  // LDS segment length of the CCF (ccf_lds_cap of the longest footprint)
  int mask_scan = 0;
  bool have_mask = false, mask_built = false;
  double *d_ccf = nullptr;  // [(P+1)*3][n_v]
  bool have_ccf = false;
  DevBuf f_par, f_model, f_depth, f_bis, f_iter;
  bool have_fit = false;
  // timing
  bool timing = false;
  std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> pending;
  double stage_ms[ST_COUNT] = {};
  // allocation sizes of the buffers made by dalloc
  std::map<void *, size_t> caps;
  // resolution taps last uploaded
  double res_sigma = -1;
  // stats
  double alg_bytes = 0;
  int64_t rows_touched = 0;
};

// ------------------------------------------------------------------ helpers
static int fail(soap_ctx *c, int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  if (c) c->err = buf;
  return code;
}

#define HIPCHK(ctx, expr)                                                                         \
  do {                                                                                           \
    hipError_t _e = (expr);                                                                      \
    if (_e != hipSuccess) return fail((ctx), SOAP_ERR_HIP, "%s: %s", #expr, hipGetErrorString(_e)); \
  } while (0)

// (re)allocate *p to hold at least `bytes`; an existing allocation that is
// large enough is kept, so repeated calls in a loop do not hit hipMalloc
static int dalloc(soap_ctx *c, void **p, size_t bytes) {
  if (bytes == 0) bytes = 8;
  auto it = c->caps.find(*p);
  if (*p && it != c->caps.end() && it->second >= bytes) return SOAP_OK;
  if (*p) {
    if (it != c->caps.end()) c->caps.erase(it);
    hipFree(*p);
    *p = nullptr;
  }
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) {
    *p = nullptr;
    return fail(c, SOAP_ERR_NOMEM, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
  }
  c->caps[*p] = bytes;
  return SOAP_OK;
}

template <class T>
static int dalloc(soap_ctx *c, T **p, size_t count) {
  return dalloc(c, (void **)p, count * sizeof(T));
}

static int ensure(soap_ctx *c, DevBuf &b, size_t bytes) {
  if (b.bytes >= bytes && b.p) return SOAP_OK;
  int rc = dalloc(c, &b.p, bytes);
  b.bytes = rc == SOAP_OK ? std::max(bytes, c->caps[b.p]) : 0;
  return rc;
}

static void dfree(void *p) {
  if (p) hipFree(p);
}

static void dfree_c(soap_ctx *c, void *p) {
  if (!p) return;
  c->caps.erase(p);
  hipFree(p);
}

struct StageTimer {
  soap_ctx *c;
  int st;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  StageTimer(soap_ctx *c_, int st_) : c(c_), st(st_) {
    if (c->timing) {
      hipEventCreate(&e0);
      hipEventCreate(&e1);
      hipEventRecord(e0, c->stream);
    }
  }
  ~StageTimer() {
    if (c->timing) {
      hipEventRecord(e1, c->stream);
      c->pending.push_back({st, {e0, e1}});
    }
  }
};

static void collect_timing(soap_ctx *c) {
  for (auto &p : c->pending) {
    float ms = 0;
    if (hipEventSynchronize(p.second.second) == hipSuccess &&
        hipEventElapsedTime(&ms, p.second.first, p.second.second) == hipSuccess)
      c->stage_ms[p.first] += ms;
    hipEventDestroy(p.second.first);
    hipEventDestroy(p.second.second);
  }
  c->pending.clear();
}

// ------------------------------------------------------------ region geometry
// spot_init_gpu + spot_phase_gpu + spot_area, evaluated serially (the
// reference's cross-block race in spot_area resolved to the serial bounding
// box), SOAP_integration.cu:21-161.
static int region_bbox(double lat, double lon, double s, double incl_deg, int nrho, int grid, double phase,
                       int bnd[5]) {
  const double pi = kPi;
  double longitude = lon * pi / 180.;
  double latitude = lat * pi / 180.;
  double inclination = incl_deg * pi / 180.;
  double rho_step = 2. * pi / (nrho - 1);
  double b = -latitude, g = longitude, b2 = pi / 2. - inclination;
  double R[3][3] = {{cos(b2) * cos(g) * cos(b) - sin(b) * sin(b2), -sin(g) * cos(b2),
                     cos(b2) * cos(g) * sin(b) + sin(b2) * cos(b)},
                    {sin(g) * cos(b), cos(g), sin(g) * sin(b)},
                    {-sin(b2) * cos(g) * cos(b) - cos(b2) * sin(b), sin(b2) * sin(g),
                     -sin(b2) * cos(g) * sin(b) + cos(b2) * cos(b)}};
  double psi = -phase * (2 * pi);
  double inc = incl_deg * pi / 180.;
  double ax[3] = {cos(inc), 0, sin(inc)};
  double cp = cos(psi), sp = sin(psi);
  double Pm[3][3] = {{(1 - cp) * ax[0] * ax[0] + cp, (1 - cp) * ax[0] * ax[1] + sp * ax[2],
                      (1 - cp) * ax[0] * ax[2] - sp * ax[1]},
                     {(1 - cp) * ax[1] * ax[0] - sp * ax[2], (1 - cp) * ax[1] * ax[1] + cp,
                      (1 - cp) * ax[1] * ax[2] + sp * ax[0]},
                     {(1 - cp) * ax[2] * ax[0] + sp * ax[1], (1 - cp) * ax[2] * ax[1] - sp * ax[0],
                      (1 - cp) * ax[2] * ax[2] + cp}};
  int on = 0, off = 0;
  double miny = 1, minz = 1, maxy = -1, maxz = -1;
  for (int k = 0; k < nrho; k++) {
    double rho = -pi + k * rho_step;
    double x2[3] = {sqrt(1 - s * s), s * cos(rho), s * sin(rho)};
    double x1[3], x3[3];
    for (int r = 0; r < 3; r++) x1[r] = R[r][0] * x2[0] + R[r][1] * x2[1] + R[r][2] * x2[2];
    for (int r = 0; r < 3; r++) x3[r] = Pm[r][0] * x1[0] + Pm[r][1] * x1[1] + Pm[r][2] * x1[2];
    if (x3[0] >= 0) {
      on++;
      miny = std::min(miny, x3[1]);
      minz = std::min(minz, x3[2]);
      maxy = std::max(maxy, x3[1]);
      maxz = std::max(maxz, x3[2]);
    } else {
      off++;
    }
  }
  if (on > 0 && off > 0) {
    if (miny * maxy < 0) {
      if (minz < 0)
        minz = -1;
      else
        maxz = 1;
    }
    if (minz * maxz < 0) {
      if (miny < 0)
        miny = -1;
      else
        maxy = 1;
    }
  }
  double gs = 2. / grid;
  bnd[0] = (int)floor((1. + miny) / gs);
  bnd[1] = (int)floor((1. + minz) / gs);
  bnd[2] = (int)ceil((1. + maxy) / gs);
  bnd[3] = (int)ceil((1. + maxz) / gs);
  bnd[4] = on > 0 ? 1 : 0;
  return bnd[4];
}

// Row 0 of spot_inverse_rotation's R3 = R2 R (SOAP_integration.cu:164-202),
// phase pre-decrement included; only x_back is used by the membership test.
static void inverse_rotation_row0(double lon, double lat, double incl_deg, double phase, double r0[3]) {
  const double pi = kPi;
  double g2 = (phase - 1) * (2 * pi);
  double i = incl_deg * pi / 180.;
  double b = lat * pi / 180.;
  double g = -lon * pi / 180.;
  double b2 = -(pi / 2. - i);
  double R[3][3] = {{(1 - cos(g2)) * cos(i) * cos(i) + cos(g2), sin(g2) * sin(i), (1 - cos(g2)) * cos(i) * sin(i)},
                    {-sin(g2) * sin(i), cos(g2), sin(g2) * cos(i)},
                    {(1 - cos(g2)) * sin(i) * cos(i), -sin(g2) * cos(i), (1 - cos(g2)) * sin(i) * sin(i) + cos(g2)}};
  double R2r0[3] = {cos(b) * cos(g) * cos(b2) - sin(b2) * sin(b), -sin(g) * cos(b),
                    cos(b) * cos(g) * sin(b2) + sin(b) * cos(b2)};
  for (int c = 0; c < 3; c++) r0[c] = R2r0[0] * R[0][c] + R2r0[1] * R[1][c] + R2r0[2] * R[2][c];
}

// ------------------------------------------------------------------ allocation-free staging
// Staging for soap_fit_ccf_host_noalloc: a file-scope device array, one copy per
// GPU, held until the process ends.  Host code never uses its name as an address
// (that is a host-side stand-in); each device's address comes from
// hipGetSymbolAddress, taken once per device at context creation with that
// device current.
constexpr size_t kScratchDoubles = size_t(2) << 20;  // 16 MiB per GPU
static __device__ __align__(256) double scratch[kScratchDoubles];
constexpr int kMaxDevices = 64;
static double *g_scratch_ptr[kMaxDevices];
static std::mutex g_scratch_lock[kMaxDevices];  // one call at a time per device in this process
static std::mutex g_scratch_init;

// called with `device` current
static hipError_t scratch_setup(int device) {
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> g(g_scratch_init);
  if (g_scratch_ptr[device]) return hipSuccess;
  void *ptr = nullptr;
  hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !ptr) return e != hipSuccess ? e : hipErrorInvalidSymbol;
  g_scratch_ptr[device] = (double *)ptr;
  return hipSuccess;
}

// ------------------------------------------------------------------ API
extern "C" {

const char *soap_version(void) { return "soap_gpu_amd 0.1 (gfx950)"; }

int soap_ctx_create(int device, soap_ctx **out) {
  if (!out) return SOAP_ERR_ARG;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return SOAP_ERR_HIP;
  if (device < 0 || device >= ndev) return SOAP_ERR_ARG;
  if (hipSetDevice(device) != hipSuccess) return SOAP_ERR_HIP;
  if (scratch_setup(device) != hipSuccess) return SOAP_ERR_HIP;
  soap_ctx *c = new soap_ctx();
  c->device = device;
  if (hipStreamCreate(&c->stream) != hipSuccess) {
    delete c;
    return SOAP_ERR_HIP;
  }
  *out = c;
  return SOAP_OK;
}

void soap_ctx_destroy(soap_ctx *c) {
  if (!c) return;
  hipSetDevice(c->device);
  hipStreamSynchronize(c->stream);
  if (c->aux) hipStreamSynchronize(c->aux);
  collect_timing(c);
  // every device buffer of the context came from dalloc, which records it in caps
  // (caller-owned cube pointers are not recorded and are not freed)
  for (auto &kv : c->caps) hipFree(kv.first);
  c->caps.clear();
  if (c->aux) hipStreamDestroy(c->aux);
  if (c->ev_fork) hipEventDestroy(c->ev_fork);
  if (c->ev_join) hipEventDestroy(c->ev_join);
  hipStreamDestroy(c->stream);
  delete c;
}

const char *soap_last_error(const soap_ctx *c) { return c ? c->err.c_str() : "null context"; }

int soap_sync(soap_ctx *c) {
  if (!c) return SOAP_ERR_ARG;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  collect_timing(c);
  return SOAP_OK;
}

int soap_set_timing(soap_ctx *c, int enable) {
  if (!c) return SOAP_ERR_ARG;
  soap_sync(c);
  c->timing = enable != 0;
  for (double &v : c->stage_ms) v = 0;
  return SOAP_OK;
}

int soap_get_stage_ms(soap_ctx *c, double *ms, int32_t n_stage) {
  if (!c || !ms) return SOAP_ERR_ARG;
  int rc = soap_sync(c);
  if (rc) return rc;
  for (int i = 0; i < n_stage && i < ST_COUNT; i++) ms[i] = c->stage_ms[i];
  return SOAP_OK;
}

const char *soap_stage_name(int32_t st) { return (st >= 0 && st < ST_COUNT) ? kStageNames[st] : ""; }
int32_t soap_num_stages(void) { return ST_COUNT; }

int soap_set_wave(soap_ctx *c, int kind, const double *wave, int64_t n, double step) {
  if (!c || !wave || n < 8) return fail(c, SOAP_ERR_ARG, "set_wave: need a wavelength array of >= 8 pixels");
  if (kind != SOAP_GRID_LOG && kind != SOAP_GRID_LINEAR) return fail(c, SOAP_ERR_ARG, "set_wave: bad grid kind");
  if (kind == SOAP_GRID_LINEAR && !(step > 0))
    return fail(c, SOAP_ERR_ARG, "set_wave: a linear grid needs the step of the reference's index formula");
  HIPCHK(c, hipSetDevice(c->device));
  std::vector<double> w(wave, wave + n);
  double h = log(w[n - 1] / w[0]) / (double)(n - 1);
  for (int64_t j = 0; j < n; j++) {
    if (kind == SOAP_GRID_LOG) {
      double e = w[0] * exp((double)j * h);
      if (!(fabs(w[j] - e) <= 1e-12 * fabs(w[j])))
        return fail(c, SOAP_ERR_GRID, "set_wave: pixel %lld deviates from a log-uniform grid (%.17g vs %.17g)",
                    (long long)j, w[j], e);
    } else if (j > 0 && !(w[j] > w[j - 1])) {
      return fail(c, SOAP_ERR_GRID, "set_wave: wavelengths must increase (pixel %lld)", (long long)j);
    }
  }
  if (c->n != n) {
    // every per-pixel buffer depends on n
    for (int i = 0; i < 6; i++) {
      if (c->cube_owned[i]) {
        c->caps.erase(c->d_cube[i]);
        dfree(c->d_cube[i]);
      }
      c->d_cube[i] = nullptr;
      c->cube_owned[i] = false;
    }
    c->have_quiet = c->have_raw = c->have_lower = c->have_ccf = c->have_fit = false;
    c->mask_built = false;
  }
  c->wave = w;
  c->n = n;
  c->kind = kind;
  c->step = step;
  c->h = h;
  c->expm1_h = expm1(h);
  c->res_sigma = -1;
  int rc = dalloc(c, &c->d_wave, (size_t)n);
  if (rc) return rc;
  HIPCHK(c, hipMemcpyAsync(c->d_wave, w.data(), n * sizeof(double), hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return SOAP_OK;
}

int soap_set_mu_bounds(soap_ctx *c, const double *bounds, int32_t n_bounds) {
  if (!c || (!bounds && n_bounds > 0) || n_bounds < 0) return fail(c, SOAP_ERR_ARG, "set_mu_bounds: bad arguments");
  if (n_bounds > 63) return fail(c, SOAP_ERR_LIMIT, "set_mu_bounds: at most 64 mu bins");
  HIPCHK(c, hipSetDevice(c->device));
  c->mu_bounds.assign(bounds, bounds + n_bounds);
  c->n_mu = n_bounds + 1;
  int rc = dalloc(c, &c->d_mub, (size_t)std::max(1, n_bounds));
  if (rc) return rc;
  if (n_bounds)
    HIPCHK(c, hipMemcpyAsync(c->d_mub, bounds, n_bounds * sizeof(double), hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return SOAP_OK;
}

int soap_set_cube(soap_ctx *c, int which, const double *cube, int32_t n_mu) {
  if (!c || !cube || which < 0 || which > 5) return fail(c, SOAP_ERR_ARG, "set_cube: bad arguments");
  if (c->n == 0) return fail(c, SOAP_ERR_STATE, "set_cube: set the wavelength grid first");
  if (n_mu < 1) return fail(c, SOAP_ERR_ARG, "set_cube: n_mu < 1");
  // the reference reads seed cubes without checks and carries a NaN into every
  // spectrum; here a non-finite value is refused with its position
  for (size_t i = 0, cnt = (size_t)n_mu * c->n; i < cnt; i++)
    if (!std::isfinite(cube[i]))
      return fail(c, SOAP_ERR_ARG, "set_cube: cube %d holds %g at mu row %zu, pixel %zu", which, cube[i],
                  i / (size_t)c->n, i % (size_t)c->n);
  HIPCHK(c, hipSetDevice(c->device));
  if (!c->cube_owned[which]) c->d_cube[which] = nullptr;
  int rc = dalloc(c, &c->d_cube[which], (size_t)n_mu * c->n);
  if (rc) return rc;
  c->cube_owned[which] = true;
  c->cube_nmu[which] = n_mu;
  HIPCHK(c, hipMemcpyAsync(c->d_cube[which], cube, (size_t)n_mu * c->n * sizeof(double), hipMemcpyHostToDevice,
                           c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return SOAP_OK;
}

int soap_set_cube_device(soap_ctx *c, int which, const double *dev_cube, int32_t n_mu) {
  if (!c || !dev_cube || which < 0 || which > 5 || n_mu < 1) return fail(c, SOAP_ERR_ARG, "set_cube_device: bad arguments");
  if (c->cube_owned[which]) {
    c->caps.erase(c->d_cube[which]);
    dfree(c->d_cube[which]);
  }
  c->d_cube[which] = const_cast<double *>(dev_cube);
  c->cube_owned[which] = false;
  c->cube_nmu[which] = n_mu;
  return SOAP_OK;
}

static int alloc_maps(soap_ctx *c, int G) {
  size_t g2 = (size_t)G * G;
  int rc;
  if ((rc = dalloc(c, &c->d_mu, g2)) || (rc = dalloc(c, &c->d_delta, g2)) || (rc = dalloc(c, &c->d_limb, g2)) ||
      (rc = dalloc(c, &c->d_lbm, g2)) || (rc = dalloc(c, &c->d_region, g2)) || (rc = dalloc(c, &c->d_m, g2)) ||
      (rc = dalloc(c, &c->d_t, g2)))
    return rc;
  c->G = G;
  return SOAP_OK;
}

// Tap offset range [dmin, dmin+nd) covering every cell's m-1 and m, from a bound
// on |delta| (no device round trip): |delta| <= vmax, s(delta) monotone.
static int set_tap_range(soap_ctx *c, double vmax_kms) {
  auto s_of = [&](double d) {
    double beta = d * 1000. / kC;
    double fbeta = -1 * (1 - sqrt((1 + beta) / (1 - beta)));
    return log1p(fbeta) / c->h;
  };
  vmax_kms = fabs(vmax_kms) * (1 + 1e-9) + 1e-12;
  if (!(vmax_kms < 0.1 * kC / 1000.)) return fail(c, SOAP_ERR_ARG, "projected velocity %g km/s out of range", vmax_kms);
  int lo = (int)floor(s_of(-vmax_kms)) - 2, hi = (int)ceil(s_of(vmax_kms)) + 1;
  c->dmin = lo;
  c->nd = hi - lo + 1;
  if (c->nd > 4096) return fail(c, SOAP_ERR_LIMIT, "Doppler shift range of %d pixels exceeds 4096", c->nd);
  return SOAP_OK;
}

// Linear grids, solid-body rotation: the quiet star from the disk maps by columns
// (no quiet cell lists or groups), where a bin set's rows fit the block's LDS.
static bool quiet_by_columns(const soap_ctx *c) { return c->lin_solid && quiet_columns_fit(c->lin_span); }

// Linear grids: the staged source-index range of a pixel tile for shifts up to
// vmax (i(j) - j = (-fbeta lambda_0/step - fbeta j)/(1 + fbeta) at the extreme
// fbeta and j, plus a margin), and the bin-sorted lists of the disk cells.
static int linear_prepare(soap_ctx *c, double vmax) {
  double lo = 0, hi = 0;
  for (double v : {-vmax, vmax}) {
    double beta = v * 1000. / kC;
    double fb = -1 * (1 - sqrt((1 + beta) / (1 - beta)));
    double fl = fb * c->wave[0] / c->step;
    for (double j : {0.0, (double)(c->n - 1)}) {
      double off = (-fl - fb * j) / (1 + fb);
      lo = std::min(lo, off);
      hi = std::max(hi, off);
    }
  }
  if (c->n > 0x7fffffffL - 4096)
    return fail(c, SOAP_ERR_LIMIT, "linear grid: %lld pixels (the gathers index pixels in 32 bits)", (long long)c->n);
  c->lin_off_lo = (int)floor(lo) - 3;
  c->lin_span = 256 + ((int)ceil(hi) + 4 - c->lin_off_lo);
  if (c->lin_span > 1600)
    return fail(c, SOAP_ERR_LIMIT, "linear grid: shifts of up to %.1f pixels exceed the tile window", hi - lo);
  int rc;
  if (quiet_by_columns(c)) return SOAP_OK;
  size_t g2 = (size_t)c->G * c->G;
  CellSrc src{};
  src.grid = c->G;
  src.n_mu = c->n_mu;
  src.n_phase = 1;
  src.nch = cell_list_chunks((long)g2);
  src.region = c->d_region;
  src.limb = c->d_limb;
  src.delta = c->d_delta;
  if ((rc = ensure(c, c->qbin_off, (size_t)(c->n_mu + 1) * 4)) || (rc = ensure(c, c->qcells, g2 * 4)) ||
      (rc = ensure(c, c->lcnt, (size_t)c->n_mu * src.nch * 4)))
    return rc;
  HIPCHK(c, launch_cell_lists(src, nullptr, (int32_t *)c->lcnt.p, (int32_t *)c->qcells.p, (int32_t *)c->qbin_off.p,
                              c->stream));
  return SOAP_OK;
}

static LinArgs lin_args(soap_ctx *c) {
  LinArgs a{};
  a.n = c->n;
  a.n_mu = c->n_mu;
  a.grid = c->G;
  a.lambda0 = c->wave[0];
  a.step = c->step;
  a.wave = c->d_wave;
  a.off_lo = c->lin_off_lo;
  a.span = c->lin_span;
  a.delta = c->d_delta;
  a.limb = c->d_limb;
  a.lbm = c->d_lbm;
  return a;
}

// max over mu in [0, 1] of |L| and |L (1 + LB)|, L = 1 - u1 (1-mu) - u2 (1-mu)^2,
// LB = lb0 + lb1 mu + lb2 mu^2 + lb3 mu^3 (sampled finely, with a 1 % margin)
static double limb_weight_bound(double u1, double u2, const double *lb) {
  double w = 0;
  for (int i = 0; i <= 4096; i++) {
    double mu = i / 4096.0, L = 1 - u1 * (1 - mu) - u2 * (1 - mu) * (1 - mu);
    double f = 1 + (lb[0] + lb[1] * mu + lb[2] * mu * mu + lb[3] * mu * mu * mu);
    w = std::max(w, std::max(std::fabs(L), std::fabs(L * f)));
  }
  return 1.01 * w;
}

// every contribution to a tap is a cell weight times a fraction in [0, 1]; with
// at most G^2 cells per tap the fixed-point sums need G^2 w_max < 2^30
static int check_tap_capacity(soap_ctx *c, double w_max, const char *what) {
  double ncell = (double)c->G * c->G;
  if (!fx_capacity_ok(ncell, w_max))
    return fail(c, SOAP_ERR_LIMIT,
                "%s: %d^2 cells with weights up to %g exceed the fixed-point tap capacity (cells x weight < 2^30)",
                what, c->G, w_max);
  return SOAP_OK;
}

int soap_disk_init(soap_ctx *c, const soap_star *st) {
  if (!c || !st) return fail(c, SOAP_ERR_ARG, "disk_init: null argument");
  if (c->n == 0 || c->n_mu == 0) return fail(c, SOAP_ERR_STATE, "disk_init: set wave and mu bounds first");
  if (st->grid < 1 || st->grid > 8192) return fail(c, SOAP_ERR_ARG, "disk_init: grid must be in [1, 8192]");
  if (!(st->prot_days != 0.0)) return fail(c, SOAP_ERR_ARG, "disk_init: zero rotation period");
  HIPCHK(c, hipSetDevice(c->device));
  int rc = alloc_maps(c, st->grid);
  if (rc) return rc;
  c->star = *st;
  c->tap_wmax = limb_weight_bound(st->u1, st->u2, st->lb);
  DiskArgs a{};
  a.grid = st->grid;
  a.n_mu = c->n_mu;
  a.delta_grid = 2. / st->grid;
  a.theta_grid = 180.0 / st->grid;
  a.omega0 = 360.0 / st->prot_days;
  a.omega1 = a.omega0 * st->w_ratio;
  c->lin_solid = a.omega1 == 0.0;
  a.radius = st->radius_km;
  a.sin_incl = sin(st->incl_deg * kPi / 180.);
  a.u1 = st->u1;
  a.u2 = st->u2;
  a.lb0 = st->lb[0];
  a.lb1 = st->lb[1];
  a.lb2 = st->lb[2];
  a.lb3 = st->lb[3];
  a.mu_bounds = c->d_mub;
  a.log_grid = c->kind == SOAP_GRID_LOG;
  a.h = c->h;
  a.expm1_h = c->expm1_h;
  a.mu = c->d_mu;
  a.delta = c->d_delta;
  a.limb = c->d_limb;
  a.lbm = c->d_lbm;
  a.region = c->d_region;
  a.m = c->d_m;
  a.t = c->d_t;
  // |delta| <= (|omega0| + |omega1|) pi/(180*86400) R |sin i|, SOAP_initialize.cu:348-351
  double vmax = (fabs(a.omega0) + fabs(a.omega1)) * kPi / (180 * 24 * 3600) * fabs(a.radius) * fabs(a.sin_incl);
  if (c->kind == SOAP_GRID_LOG && (rc = set_tap_range(c, vmax))) return rc;
  {
    StageTimer tm(c, ST_DISK);
    HIPCHK(c, launch_disk_init(a, c->stream));
  }
  if (c->kind == SOAP_GRID_LINEAR && (rc = linear_prepare(c, vmax * (1 + 1e-12)))) return rc;
  c->have_maps = true;
  c->have_quiet = c->have_raw = c->have_lower = c->have_ccf = c->have_fit = false;
  return SOAP_OK;
}

int soap_get_maps(soap_ctx *c, double *mu, double *delta, double *limb, int32_t *region, double *lb) {
  if (!c) return SOAP_ERR_ARG;
  if (!c->have_maps) return fail(c, SOAP_ERR_STATE, "get_maps: no maps");
  HIPCHK(c, hipSetDevice(c->device));
  size_t g2 = (size_t)c->G * c->G;
  if (mu) HIPCHK(c, hipMemcpyAsync(mu, c->d_mu, g2 * 8, hipMemcpyDeviceToHost, c->stream));
  if (delta) HIPCHK(c, hipMemcpyAsync(delta, c->d_delta, g2 * 8, hipMemcpyDeviceToHost, c->stream));
  if (limb) HIPCHK(c, hipMemcpyAsync(limb, c->d_limb, g2 * 8, hipMemcpyDeviceToHost, c->stream));
  if (region) HIPCHK(c, hipMemcpyAsync(region, c->d_region, g2 * 4, hipMemcpyDeviceToHost, c->stream));
  if (lb) HIPCHK(c, hipMemcpyAsync(lb, c->d_lbm, g2 * 8, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return SOAP_OK;
}

int soap_set_maps(soap_ctx *c, const soap_star *st, const double *delta, const double *limb, const int32_t *region,
                  const double *lb) {
  if (!c || !st || !delta || !limb || !region || !lb) return fail(c, SOAP_ERR_ARG, "set_maps: null argument");
  if (c->n == 0 || c->n_mu == 0) return fail(c, SOAP_ERR_STATE, "set_maps: set wave and mu bounds first");
  HIPCHK(c, hipSetDevice(c->device));
  int G = st->grid;
  size_t g2 = (size_t)G * G;
  c->lin_solid = false;  // external maps: no column structure assumed
  int rc = alloc_maps(c, G);
  if (rc) return rc;
  c->star = *st;
  // per-cell (m, t) from the given delta map, same expressions as k_disk_init
  std::vector<int32_t> m(g2, 0);
  std::vector<double> t(g2, 0.0), mu(g2, -999.0);
  double vmax = 0;
  for (size_t i = 0; i < g2; i++)
    if (!(limb[i] == -999.0 && delta[i] == -999.0)) vmax = std::max(vmax, fabs(delta[i]));
  if (c->kind == SOAP_GRID_LOG && (rc = set_tap_range(c, vmax))) return rc;
  for (size_t i = 0; i < g2; i++) {
    if (region[i] < 0 || region[i] >= c->n_mu) return fail(c, SOAP_ERR_ARG, "set_maps: region %d out of range", region[i]);
    if (limb[i] == -999.0 && delta[i] == -999.0) continue;
    if (c->kind != SOAP_GRID_LOG) continue;
    double beta = delta[i] * 1000. / kC;
    double fbeta = -1 * (1 - sqrt((1 + beta) / (1 - beta)));
    double sh = log1p(fbeta) / c->h;
    double mc = ceil(sh);
    m[i] = (int32_t)mc;
    t[i] = expm1((mc - sh) * c->h) / c->expm1_h;
  }
  double wmax = 0;
  for (size_t i = 0; i < g2; i++)
    if (!(limb[i] == -999.0 && delta[i] == -999.0))
      wmax = std::max(wmax, std::max(std::fabs(limb[i]), std::fabs(limb[i] * (1 + lb[i]))));
  c->tap_wmax = wmax;
  HIPCHK(c, hipMemcpyAsync(c->d_delta, delta, g2 * 8, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->d_limb, limb, g2 * 8, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->d_region, region, g2 * 4, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->d_lbm, lb, g2 * 8, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->d_mu, mu.data(), g2 * 8, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->d_m, m.data(), g2 * 4, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->d_t, t.data(), g2 * 8, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (c->kind == SOAP_GRID_LINEAR && (rc = linear_prepare(c, vmax))) return rc;
  c->have_maps = true;
  c->have_quiet = c->have_raw = c->have_lower = c->have_ccf = c->have_fit = false;
  return SOAP_OK;
}

static int check_cube(soap_ctx *c, int which, const char *what) {
  if (!c->d_cube[which]) return fail(c, SOAP_ERR_STATE, "%s: seed cube %d not set", what, which);
  if (c->cube_nmu[which] != c->n_mu)
    return fail(c, SOAP_ERR_ARG, "%s: cube %d has %d mu rows, the mu bounds give %d", what, which, c->cube_nmu[which],
                c->n_mu);
  return SOAP_OK;
}

// taps scratch for n_job jobs of nsets_in/out sets
static int ensure_taps(soap_ctx *c, int n_job, int nsets_in, int nsets_out) {
  int rc;
  size_t per = (size_t)c->n_mu * c->nd;
  if ((rc = ensure(c, c->acc, (size_t)n_job * nsets_in * per * 2 * 8)) ||
      (rc = ensure(c, c->taps, (size_t)n_job * nsets_out * per * 8)) ||
      (rc = ensure(c, c->krange, (size_t)n_job * c->n_mu * 2 * 4)) ||
      (rc = ensure(c, c->kflags, (size_t)n_job * c->n_mu * 4)) || (rc = ensure(c, c->bins, (size_t)n_job * c->n_mu * 4)) ||
      (rc = ensure(c, c->nbins, (size_t)n_job * 4)))
    return rc;
  // the phase-interleaved copy is read by the grouped FIR (four sets per entry)
  if (nsets_out == 4 && (rc = ensure(c, c->taps_g, (size_t)taps_g_pad(n_job) * nsets_out * per * 8))) return rc;
  return SOAP_OK;
}

// LDS window of tap offsets per (set, bin), kept within 64 KiB of LDS
static int tap_window(soap_ctx *c, int nsets) {
  int wd = std::min(16, c->nd);
  while (wd > 1 && (size_t)nsets * c->n_mu * wd * 16 > 65536) wd--;
  return wd;
}

int soap_quiet_star(soap_ctx *c) {
  if (!c) return SOAP_ERR_ARG;
  if (!c->have_maps) return fail(c, SOAP_ERR_STATE, "quiet_star: run disk_init first");
  int rc;
  if ((rc = check_cube(c, SOAP_CUBE_QUIET, "quiet_star")) || (rc = check_cube(c, SOAP_CUBE_QUIET_SED, "quiet_star")))
    return rc;
  HIPCHK(c, hipSetDevice(c->device));
  if ((rc = dalloc(c, &c->d_quiet, (size_t)2 * c->n))) return rc;
  if (c->kind == SOAP_GRID_LINEAR) {
    size_t g2 = (size_t)c->G * c->G;
    int ncs = quiet_column_sets(c->n_mu);
    const bool by_cols = quiet_by_columns(c);
    size_t nsets = (size_t)std::max(ncs, quiet_linear_sets(c->n_mu));
    if ((rc = ensure(c, c->qpart, nsets * 2 * c->n * 8))) return rc;
    if (!by_cols && ((rc = ensure(c, c->qgroups, g2 * sizeof(LinGroup))) || (rc = ensure(c, c->qg_n, (size_t)c->n_mu * 4))))
      return rc;
    if (by_cols && ((rc = ensure(c, c->qcolw, (size_t)ncs * c->G * 4 * 8)) ||
                         (rc = ensure(c, c->qcolf, (size_t)ncs * c->G * 4)) ||
                         (rc = ensure(c, c->qcols, (size_t)c->G * sizeof(LinGroup)))))
      return rc;
    if (!by_cols) {
      StageTimer tm(c, ST_TAPS_QUIET);
      HIPCHK(c, launch_lin_groups((const int32_t *)c->qcells.p, nullptr, (const int32_t *)c->qbin_off.p, 1, c->n_mu,
                                  c->d_delta, c->d_limb, c->d_lbm, c->wave[0], c->step, 0,
                                  (LinGroup *)c->qgroups.p, (int32_t *)c->qg_n.p, c->stream));
    }
    {
      StageTimer tm(c, ST_FIR_QUIET);
      if (by_cols)
        HIPCHK(c, launch_gather_quiet_cols(lin_args(c), c->d_cube[SOAP_CUBE_QUIET], c->d_cube[SOAP_CUBE_QUIET_SED],
                                           c->d_region, (double *)c->qcolw.p, (int32_t *)c->qcolf.p,
                                           (LinGroup *)c->qcols.p, (double *)c->qpart.p, c->d_quiet, c->stream));
      else
        HIPCHK(c, launch_gather_quiet_linear(lin_args(c), c->d_cube[SOAP_CUBE_QUIET], c->d_cube[SOAP_CUBE_QUIET_SED],
                                             (const LinGroup *)c->qgroups.p, (const int32_t *)c->qbin_off.p,
                                             (const int32_t *)c->qg_n.p, (double *)c->qpart.p, c->d_quiet,
                                             c->stream));
    }
    c->have_quiet = true;
    c->have_lower = c->have_ccf = c->have_fit = false;
    return SOAP_OK;
  }
  if ((rc = check_tap_capacity(c, c->tap_wmax, "quiet_star"))) return rc;
  {
    size_t per = (size_t)c->n_mu * c->nd;
    if ((rc = ensure(c, c->qacc, per * 2 * 8)) || (rc = ensure(c, c->qtaps, per * 8)) ||
        (rc = ensure(c, c->qkrange, (size_t)c->n_mu * 2 * 4)) || (rc = ensure(c, c->qkflags, (size_t)c->n_mu * 4)) ||
        (rc = ensure(c, c->qbins, (size_t)c->n_mu * 4)) || (rc = ensure(c, c->qnbins, 4)))
      return rc;
  }
  TapAccArgs a{};
  a.grid = c->G;
  a.n_mu = c->n_mu;
  a.dmin = c->dmin;
  a.nd = c->nd;
  a.nsets = 1;
  a.wd = tap_window(c, 1);
  a.cols_per_block = std::max(1, 4096 / std::max(1, c->G));
  a.region = c->d_region;
  a.limb = c->d_limb;
  a.delta = c->d_delta;
  a.lbm = c->d_lbm;
  a.m = c->d_m;
  a.t = c->d_t;
  a.acc = (unsigned long long *)c->qacc.p;
  {
    StageTimer tm(c, ST_TAPS_QUIET);
    HIPCHK(c, launch_fill_u64(a.acc, (size_t)c->n_mu * c->nd * 2, c->stream));
    HIPCHK(c, launch_taps_quiet(a, c->stream));
    HIPCHK(c, launch_taps_finalize(a.acc, 1, 1, 1, c->n_mu, c->nd, (double *)c->qtaps.p, (int32_t *)c->qkrange.p,
                                   (int32_t *)c->qkflags.p, (int32_t *)c->qbins.p, (int32_t *)c->qnbins.p, nullptr,
                                   0, c->stream));
  }
  FirArgs f{};
  f.q = c->d_cube[SOAP_CUBE_QUIET];
  f.qs = c->d_cube[SOAP_CUBE_QUIET_SED];
  f.n = c->n;
  f.n_mu = c->n_mu;
  f.dmin = c->dmin;
  f.nd = c->nd;
  f.taps = (const double *)c->qtaps.p;
  f.krange = (const int32_t *)c->qkrange.p;
  f.kflags = (const int32_t *)c->qkflags.p;
  f.bins = (const int32_t *)c->qbins.p;
  f.nbins = (const int32_t *)c->qnbins.p;
  f.out = c->d_quiet;
  {
    StageTimer tm(c, ST_FIR_QUIET);
    HIPCHK(c, launch_fir_quiet_lds(f, c->stream));
  }
  c->have_quiet = true;
  c->have_lower = c->have_ccf = c->have_fit = false;
  return SOAP_OK;
}

int soap_get_quiet(soap_ctx *c, int which, double *out) {
  if (!c || !out || which < 0 || which > 1) return fail(c, SOAP_ERR_ARG, "get_quiet: bad arguments");
  if (!c->have_quiet) return fail(c, SOAP_ERR_STATE, "get_quiet: no quiet star");
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipMemcpyAsync(out, c->d_quiet + (size_t)which * c->n, c->n * 8, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return SOAP_OK;
}

int soap_set_quiet(soap_ctx *c, int which, const double *q) {
  if (!c || !q || which < 0 || which > 1) return fail(c, SOAP_ERR_ARG, "set_quiet: bad arguments");
  if (c->n == 0) return fail(c, SOAP_ERR_STATE, "set_quiet: set the wavelength grid first");
  HIPCHK(c, hipSetDevice(c->device));
  if (!c->d_quiet) {
    int rc = dalloc(c, &c->d_quiet, (size_t)2 * c->n);
    if (rc) return rc;
    HIPCHK(c, launch_fill_f64(c->d_quiet, (size_t)2 * c->n, 0.0, c->stream));
  }
  HIPCHK(c, hipMemcpyAsync(c->d_quiet + (size_t)which * c->n, q, c->n * 8, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->have_quiet = true;
  return SOAP_OK;
}

// Solar mode (SOAP_integration_solar.cu:424-433): facula intensity
// B(lambda, T_cell)/B(lambda, T_norm) C_fac(lambda), T_norm = int(t_sum + 0.9 - 216.8),
// with T_cell = int(t_sum + 0.9 - 407.7 mu + 190.9 mu^2) in [t_sum - 216, t_sum].
// The kernels use the Lagrange interpolant of the ratio in T on K Chebyshev
// nodes; K is the smallest of 6, 8, 10, 12 whose interpolant matches the ratio
// to 2e-14 relative at every integer T of the range, at 9 wavelengths across the
// grid (the error is largest at the shortest wavelength).
static double planck_ratio(double lam, double t, int t_norm) {
  double hc = kPlanckH * kC;
  return std::expm1(hc / (lam * kBoltzK * (double)t_norm * 1e-10)) / std::expm1(hc / (lam * kBoltzK * t * 1e-10));
}

static int solar_nodes(soap_ctx *c, int t_sum) {
  int t_norm = (int)(t_sum + 0.9 - 216.8);
  double lo = t_sum - 216, hi = t_sum;
  if (lo < 500) return fail(c, SOAP_ERR_ARG, "set_activity: facula temperature range [%g, %g] K is not physical", lo, hi);
  for (int K = 6; K <= kMaxSolarNodes; K += 2) {
    double x[kMaxSolarNodes], w[kMaxSolarNodes];
    for (int q = 0; q < K; q++) x[q] = 0.5 * (lo + hi) + 0.5 * (hi - lo) * std::cos((2 * q + 1) * kPi / (2 * K));
    for (int q = 0; q < K; q++) {
      w[q] = 1.0;
      for (int r = 0; r < K; r++)
        if (r != q) w[q] /= x[q] - x[r];
    }
    double err = 0;
    for (int sidx = 0; sidx < 9; sidx++) {
      double lam = c->wave[(size_t)((c->n - 1) * sidx / 8)];
      double R[kMaxSolarNodes];
      for (int q = 0; q < K; q++) R[q] = planck_ratio(lam, x[q], t_norm);
      for (int T = t_sum - 216; T <= t_sum; T++) {
        double v = 0;
        for (int q = 0; q < K; q++) {
          double l = w[q];
          for (int r = 0; r < K; r++)
            if (r != q) l *= T - x[r];
          v += R[q] * l;
        }
        double ex = planck_ratio(lam, T, t_norm);
        err = std::max(err, std::fabs(v - ex) / std::fabs(ex));
      }
    }
    if (err <= 2e-14) {
      double leb = 0;  // Lebesgue function max: bounds the Lagrange weights of a cell
      for (int T = t_sum - 216; T <= t_sum; T++) {
        double sum = 0;
        for (int q = 0; q < K; q++) {
          double l = w[q];
          for (int r = 0; r < K; r++)
            if (r != q) l *= T - x[r];
          sum += std::fabs(l);
        }
        leb = std::max(leb, sum);
      }
      c->solar_leb = leb;
      c->solar_k = K;
      c->t_norm_fac = t_norm;
      std::copy(x, x + K, c->solar_t);
      std::copy(w, w + K, c->solar_w);
      return SOAP_OK;
    }
  }
  return fail(c, SOAP_ERR_LIMIT, "set_activity: Planck-ratio interpolation over [%g, %g] K misses 2e-14 at %d nodes", lo,
              hi, kMaxSolarNodes);
}

int soap_set_activity(soap_ctx *c, const soap_activity *act, int32_t n_phase, const double *phase,
                      const int32_t *region_count, const soap_region *regions) {
  if (!c || !act || n_phase < 0 || (n_phase > 0 && (!phase || !region_count)))
    return fail(c, SOAP_ERR_ARG, "set_activity: bad arguments");
  if (!c->have_maps) return fail(c, SOAP_ERR_STATE, "set_activity: run disk_init first");
  if (act->nrho < 2) return fail(c, SOAP_ERR_ARG, "set_activity: nrho must be >= 2");
  int rc;
  if (act->solar_mode && (rc = solar_nodes(c, act->t_star + act->t_diff_fac))) return rc;
  if (!act->solar_mode) c->solar_k = 0;
  HIPCHK(c, hipSetDevice(c->device));
  int64_t total = 0;
  for (int p = 0; p < n_phase; p++) {
    if (region_count[p] < 0) return fail(c, SOAP_ERR_ARG, "set_activity: negative region count");
    total += region_count[p];
  }
  if (total > 0 && !regions) return fail(c, SOAP_ERR_ARG, "set_activity: null regions");
  for (int64_t i = 0; i < total; i++)
    if (!(regions[i].size >= 0.0 && regions[i].size < 1.0))
      return fail(c, SOAP_ERR_ARG, "set_activity: region size %g outside [0, 1)", regions[i].size);
  c->act = *act;
  c->P = n_phase;
  c->phases.assign(phase, phase + n_phase);
  c->reg_count.assign(region_count, region_count + n_phase);
  c->regions.assign(regions, regions + total);
  // geometry per (phase, visible region)
  int G = c->G;
  c->geo.clear();
  c->geo_off.assign(1, 0);
  c->box.clear();
  int cpb = std::max(1, 4096 / std::max(1, G));
  c->max_col_blocks = 0;
  int64_t ri = 0;
  for (int p = 0; p < n_phase; p++) {
    int by0 = G, by1 = 0, bz0 = G, bz1 = 0;
    for (int r = 0; r < region_count[p]; r++, ri++) {
      const soap_region &rg = c->regions[ri];
      int bnd[5];
      region_bbox(rg.lat_deg, rg.lon_deg, rg.size, c->star.incl_deg, act->nrho, G, phase[p], bnd);
      if (bnd[4] != 1) continue;
      RegionDev g{};
      double r0[3];
      inverse_rotation_row0(rg.lon_deg, rg.lat_deg, c->star.incl_deg, phase[p], r0);
      g.r0 = r0[0];
      g.r1 = r0[1];
      g.r2 = r0[2];
      g.oms2 = 1. - rg.size * rg.size;
      g.y0 = std::max(0, bnd[0]);
      g.y1 = std::min(G, bnd[2]);
      g.z0 = std::max(0, bnd[1]);
      g.z1 = std::min(G, bnd[3]);
      g.type = rg.type == 0.0 ? 0 : 1;
      if (g.y0 >= g.y1 || g.z0 >= g.z1) continue;
      c->geo.push_back(g);
      by0 = std::min(by0, g.y0);
      by1 = std::max(by1, g.y1);
      bz0 = std::min(bz0, g.z0);
      bz1 = std::max(bz1, g.z1);
    }
    c->geo_off.push_back((int32_t)c->geo.size());
    if (by0 >= by1) by0 = by1 = bz0 = bz1 = 0;
    c->box.insert(c->box.end(), {by0, by1, bz0, bz1});
    c->max_col_blocks = std::max(c->max_col_blocks, (by1 - by0 + cpb - 1) / cpb);
  }
  if ((rc = dalloc(c, &c->d_geo, std::max<size_t>(1, c->geo.size()))) ||
      (rc = dalloc(c, &c->d_geo_off, c->geo_off.size())) || (rc = dalloc(c, &c->d_box, std::max<size_t>(4, c->box.size()))))
    return rc;
  if (!c->geo.empty())
    HIPCHK(c, hipMemcpyAsync(c->d_geo, c->geo.data(), c->geo.size() * sizeof(RegionDev), hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->d_geo_off, c->geo_off.data(), c->geo_off.size() * 4, hipMemcpyHostToDevice, c->stream));
  if (!c->box.empty())
    HIPCHK(c, hipMemcpyAsync(c->d_box, c->box.data(), c->box.size() * 4, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if ((rc = dalloc(c, &c->d_raw, (size_t)std::max(1, n_phase) * 3 * c->n))) return rc;
  c->have_raw = c->have_lower = c->have_ccf = c->have_fit = false;
  return SOAP_OK;
}

// Linear grids: per-phase active cell lists (by bin), then the gather kernel.
static int integrate_linear(soap_ctx *c) {
  int rc;
  std::vector<int64_t> loff(c->P + 1, 0);
  long max_box = 0;
  for (int p = 0; p < c->P; p++) {
    const int32_t *b = &c->box[(size_t)p * 4];
    long nb = (long)(b[1] - b[0]) * (b[3] - b[2]);
    loff[p + 1] = loff[p] + nb;
    max_box = std::max(max_box, nb);
  }
  CellSrc src{};
  src.grid = c->G;
  src.n_mu = c->n_mu;
  src.n_phase = c->P;
  src.nch = std::max(1, cell_list_chunks(max_box));
  src.region = c->d_region;
  src.limb = c->d_limb;
  src.delta = c->d_delta;
  src.geo = c->d_geo;
  src.geo_off = c->d_geo_off;
  src.box = c->d_box;
  size_t cap = (size_t)std::max<int64_t>(1, loff[c->P]);
  if ((rc = ensure(c, c->acells, cap * 4)) || (rc = ensure(c, c->abin_off, (size_t)c->P * (c->n_mu + 1) * 4)) ||
      (rc = ensure(c, c->alist_off, (size_t)(c->P + 1) * 8)) || (rc = ensure(c, c->agroups, cap * sizeof(LinGroup))) ||
      (rc = ensure(c, c->ag_n, (size_t)c->P * c->n_mu * 4)) ||
      (rc = ensure(c, c->lcnt, (size_t)c->P * c->n_mu * src.nch * 4)))
    return rc;
  HIPCHK(c, hipMemcpyAsync(c->alist_off.p, loff.data(), (c->P + 1) * 8, hipMemcpyHostToDevice, c->stream));
  {
    StageTimer tm(c, ST_TAPS_ACTIVE);
    HIPCHK(c, launch_cell_lists(src, (const int64_t *)c->alist_off.p, (int32_t *)c->lcnt.p, (int32_t *)c->acells.p,
                                (int32_t *)c->abin_off.p, c->stream));
    HIPCHK(c, launch_lin_groups((const int32_t *)c->acells.p, (const int64_t *)c->alist_off.p,
                                (const int32_t *)c->abin_off.p, c->P, c->n_mu, c->d_delta, c->d_limb, c->d_lbm,
                                c->wave[0], c->step, c->act.solar_mode != 0, (LinGroup *)c->agroups.p,
                                (int32_t *)c->ag_n.p, c->stream));
  }
  FirArgs f{};
  f.q = c->d_cube[SOAP_CUBE_QUIET];
  f.qs = c->d_cube[SOAP_CUBE_QUIET_SED];
  f.sp = c->d_cube[SOAP_CUBE_SPOT];
  f.sps = c->d_cube[SOAP_CUBE_SPOT_SED];
  f.fa = c->d_cube[SOAP_CUBE_FAC];
  f.fas = c->d_cube[SOAP_CUBE_FAC_SED];
  f.n = c->n;
  f.n_mu = c->n_mu;
  f.out = c->d_raw;
  LinSolar sol{};
  sol.on = c->act.solar_mode != 0;
  sol.t_sum = c->act.t_star + c->act.t_diff_fac;
  sol.t_norm = (int)(sol.t_sum + 0.9 - 216.8);
  {
    StageTimer tm(c, ST_FIR_ACTIVE);
    HIPCHK(c, launch_gather_active_linear(lin_args(c), f, (const LinGroup *)c->agroups.p,
                                          (const int64_t *)c->alist_off.p, (const int32_t *)c->abin_off.p,
                                          (const int32_t *)c->ag_n.p, c->P, sol, c->stream));
  }
  c->have_raw = true;
  c->have_lower = c->have_ccf = c->have_fit = false;
  c->alg_bytes = -1;
  return SOAP_OK;
}

int soap_integrate(soap_ctx *c) {
  if (!c) return SOAP_ERR_ARG;
  if (!c->have_maps || c->phases.size() != (size_t)c->P) return fail(c, SOAP_ERR_STATE, "integrate: set activity first");
  int rc;
  if ((rc = check_cube(c, SOAP_CUBE_QUIET, "integrate")) || (rc = check_cube(c, SOAP_CUBE_QUIET_SED, "integrate")))
    return rc;
  bool any_spot = false, any_fac = false;
  for (auto &g : c->geo) (g.type == 0 ? any_spot : any_fac) = true;
  if (any_spot && ((rc = check_cube(c, SOAP_CUBE_SPOT, "integrate")) || (rc = check_cube(c, SOAP_CUBE_SPOT_SED, "integrate"))))
    return rc;
  if (any_fac && ((rc = check_cube(c, SOAP_CUBE_FAC, "integrate")) || (rc = check_cube(c, SOAP_CUBE_FAC_SED, "integrate"))))
    return rc;
  if (c->P == 0) {
    c->have_raw = true;
    return SOAP_OK;
  }
  HIPCHK(c, hipSetDevice(c->device));
  if (c->kind == SOAP_GRID_LINEAR) return integrate_linear(c);
  // sets: H1s, H1f, H2f ((1+LB) limb) -> + H1a; solar: H1s, H1f, L_0..L_{K-1} -> + H1a
  const bool solar = c->act.solar_mode != 0;
  const int nsets = solar ? 2 + c->solar_k : 3;
  if (solar && c->solar_k == 0) return fail(c, SOAP_ERR_STATE, "integrate: solar nodes missing");
  if ((rc = check_tap_capacity(c, c->tap_wmax * (solar ? std::max(1.0, c->solar_leb) : 1.0), "integrate")))
    return rc;
  if ((rc = ensure_taps(c, c->P, nsets, nsets + 1))) return rc;
  TapAccArgs a{};
  a.grid = c->G;
  a.n_mu = c->n_mu;
  a.dmin = c->dmin;
  a.nd = c->nd;
  a.nsets = nsets;
  a.wd = tap_window(c, nsets);
  if (solar) {
    a.solar_nodes = c->solar_k;
    a.t_sum = c->act.t_star + c->act.t_diff_fac;
    std::copy(c->solar_t, c->solar_t + c->solar_k, a.node_t);
    std::copy(c->solar_w, c->solar_w + c->solar_k, a.node_w);
  }
  a.cols_per_block = std::max(1, 4096 / std::max(1, c->G));
  a.region = c->d_region;
  a.limb = c->d_limb;
  a.delta = c->d_delta;
  a.lbm = c->d_lbm;
  a.m = c->d_m;
  a.t = c->d_t;
  a.acc = (unsigned long long *)c->acc.p;
  a.geo = c->d_geo;
  a.geo_off = c->d_geo_off;
  a.box = c->d_box;
  {
    StageTimer tm(c, ST_TAPS_ACTIVE);
    HIPCHK(c, launch_fill_u64(a.acc, (size_t)c->P * nsets * c->n_mu * c->nd * 2, c->stream));
    HIPCHK(c, launch_taps_active(a, c->P, c->max_col_blocks, c->stream));
    HIPCHK(c, launch_taps_finalize(a.acc, c->P, nsets, nsets + 1, c->n_mu, c->nd, (double *)c->taps.p, (int32_t *)c->krange.p,
                                   (int32_t *)c->kflags.p, (int32_t *)c->bins.p, (int32_t *)c->nbins.p,
                                   solar ? nullptr : (double *)c->taps_g.p, taps_g_pad(c->P), c->stream));
  }
  FirArgs f{};
  f.q = c->d_cube[SOAP_CUBE_QUIET];
  f.qs = c->d_cube[SOAP_CUBE_QUIET_SED];
  f.sp = c->d_cube[SOAP_CUBE_SPOT];
  f.sps = c->d_cube[SOAP_CUBE_SPOT_SED];
  f.fa = c->d_cube[SOAP_CUBE_FAC];
  f.fas = c->d_cube[SOAP_CUBE_FAC_SED];
  f.n = c->n;
  f.n_mu = c->n_mu;
  f.dmin = c->dmin;
  f.nd = c->nd;
  f.taps = (const double *)c->taps.p;
  f.taps_g = (const double *)c->taps_g.p;
  f.p_pad = taps_g_pad(c->P);
  f.krange = (const int32_t *)c->krange.p;
  f.kflags = (const int32_t *)c->kflags.p;
  f.bins = (const int32_t *)c->bins.p;
  f.nbins = (const int32_t *)c->nbins.p;
  f.out = c->d_raw;
  {
    StageTimer tm(c, ST_FIR_ACTIVE);
    if (solar) {
      SolarFir sf{};
      sf.wave = c->d_wave;
      sf.k = c->solar_k;
      sf.t_norm = c->t_norm_fac;
      std::copy(c->solar_t, c->solar_t + c->solar_k, sf.node_t);
      hipError_t e = launch_fir_solar(f, sf, c->P, c->stream);
      if (e == hipErrorInvalidValue)
        return fail(c, SOAP_ERR_LIMIT, "integrate: solar mode needs n_mu <= 64 and a shift range of < 7000 pixels");
      HIPCHK(c, e);
    } else {
      int fg = c->fir_g, fb = c->fir_b;
      const bool automatic = fb == -1;
      if (automatic) fg = 0, fb = -7;  // the DPP tile form, else the grouped form
      const bool tile = fb == -4 || fb == -7 || fb == -8;
      FirProg pg{};
      if (tile) {
        const int ng = fir_tile_groups(c->P), cap = c->n_mu * c->nd + 6;
        int rc;
        if ((rc = ensure(c, c->prog_t, (size_t)ng * cap * 16 * 8)) || (rc = ensure(c, c->prog_x, (size_t)ng * cap * 4)) ||
            (rc = ensure(c, c->prog_n, (size_t)ng * 4)) || (rc = ensure(c, c->prog_u, (size_t)c->n_mu * 4)))
          return rc;
        pg = FirProg{(double *)c->prog_t.p, (int32_t *)c->prog_x.p, (int32_t *)c->prog_n.p, (int32_t *)c->prog_u.p, cap};
      }
      hipError_t e = fb == -2  ? launch_fir_direct(f, c->P, fg, c->stream)
                     : tile ? launch_fir_tile(f, c->P, pg, fb != -4, fb == -4 ? 8 : fg, fb == -8 ? 6 : 4, c->stream)
                     : fb == 0 ? launch_fir_active_grp(f, c->P, c->stream)
                               : launch_fir2(f, c->P, fg, fb, c->stream);
      // shift ranges too wide for a form's LDS: the grouped form, then the plain one
      if (e == hipErrorInvalidValue && automatic) e = launch_fir_active_grp(f, c->P, c->stream);
      if (e == hipErrorInvalidValue) e = launch_fir_active(f, c->P, c->stream);
      HIPCHK(c, e);
    }
  }
  c->have_raw = true;
  c->have_lower = c->have_ccf = c->have_fit = false;
  c->alg_bytes = -1;  // computed lazily
  return SOAP_OK;
}

int soap_set_fir_shape(soap_ctx *c, int32_t phases_per_block, int32_t pixels_per_thread) {
  if (!c) return SOAP_ERR_ARG;
  const int ok[][2] = {{-1, -1}, {4, 0}, {1, 3}, {1, 5}, {2, 3}, {2, 1}, {4, 1}, {1, -2}, {2, -2}, {4, -2}, {4, -4}, {0, -7}, {4, -7}, {5, -7}, {6, -7}, {7, -7}, {8, -7}, {0, -8}, {5, -8}, {8, -8}};
  for (auto &s : ok)
    if (s[0] == phases_per_block && s[1] == pixels_per_thread) {
      c->fir_g = phases_per_block;
      c->fir_b = pixels_per_thread;
      return SOAP_OK;
    }
  return fail(c, SOAP_ERR_ARG, "set_fir_shape: unsupported shape (%d, %d)", phases_per_block, pixels_per_thread);
}

int soap_get_integrate_stats(soap_ctx *c, double *bytes_min, double *bytes_phase_sum, int64_t *rows_sum) {
  if (!c) return SOAP_ERR_ARG;
  if (!c->have_raw || c->P == 0) return fail(c, SOAP_ERR_STATE, "integrate_stats: run integrate first");
  if (c->kind != SOAP_GRID_LOG) return fail(c, SOAP_ERR_STATE, "integrate_stats: the traffic model covers log grids");
  HIPCHK(c, hipSetDevice(c->device));
  std::vector<int32_t> fl((size_t)c->P * c->n_mu);
  HIPCHK(c, hipMemcpyAsync(fl.data(), c->kflags.p, fl.size() * 4, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  int64_t rows = 0;
  double per_phase = 0;
  std::vector<int> uni(c->n_mu, 0);
  bool us = false, uf = false;
  for (int p = 0; p < c->P; p++) {
    int64_t r = 0;
    bool s = false, f = false;
    for (int k = 0; k < c->n_mu; k++) {
      int v = fl[(size_t)p * c->n_mu + k];
      uni[k] |= v ? (4 | v) : 0;
      if (v) r += 2;  // quiet normalised + SED rows
      if (v & 1) r += 1, s = true;
      if (v & 2) r += 1, f = true;
    }
    us |= s;
    uf |= f;
    rows += r;
    per_phase += 8.0 * c->n * (double)(r + (s ? 1 : 0) + (f ? 1 : 0) + 3);
  }
  int64_t ru = 0;
  for (int k = 0; k < c->n_mu; k++) ru += (uni[k] & 4 ? 2 : 0) + (uni[k] & 1 ? 1 : 0) + (uni[k] & 2 ? 1 : 0);
  if (bytes_min) *bytes_min = 8.0 * c->n * (double)(ru + (us ? 1 : 0) + (uf ? 1 : 0) + 3 * (int64_t)c->P);
  if (bytes_phase_sum) *bytes_phase_sum = per_phase;
  if (rows_sum) *rows_sum = rows;
  return SOAP_OK;
}

int soap_get_stage_work(soap_ctx *c, int32_t stage, double *bytes, double *flops) {
  if (!c || !bytes || !flops || stage < 0 || stage >= ST_COUNT) return fail(c, SOAP_ERR_ARG, "stage_work: bad arguments");
  HIPCHK(c, hipSetDevice(c->device));
  double n = (double)c->n;
  *bytes = 0;
  *flops = 0;
  switch (stage) {
    case ST_FIR_ACTIVE: {
      if (!c->have_raw || c->P == 0) return fail(c, SOAP_ERR_STATE, "stage_work: run integrate first");
      int rc = soap_get_integrate_stats(c, bytes, nullptr, nullptr);
      if (rc) return rc;
      size_t nk = (size_t)c->P * c->n_mu;
      std::vector<int32_t> fl(nk), kr(nk * 2);
      HIPCHK(c, hipMemcpyAsync(fl.data(), c->kflags.p, nk * 4, hipMemcpyDeviceToHost, c->stream));
      HIPCHK(c, hipMemcpyAsync(kr.data(), c->krange.p, nk * 8, hipMemcpyDeviceToHost, c->stream));
      HIPCHK(c, hipStreamSynchronize(c->stream));
      // products per tap: H1a*SEDq always; spot: 2; facula: 4 ((1+LB) form) or 2 + 2K (solar)
      int fac_prod = c->act.solar_mode ? 2 + 2 * c->solar_k : 4;
      double fma = 0;
      for (size_t e = 0; e < nk; e++) {
        if (!fl[e] || kr[2 * e] > kr[2 * e + 1]) continue;
        int prod = 1 + ((fl[e] & 1) ? 2 : 0) + ((fl[e] & 2) ? fac_prod : 0);
        fma += (double)(kr[2 * e + 1] - kr[2 * e] + 1) * prod;
      }
      *flops = 2.0 * fma * n;
      return SOAP_OK;
    }
    case ST_CONV: {
      if (!c->have_lower) return fail(c, SOAP_ERR_STATE, "stage_work: run resolution first");
      if (c->kind != SOAP_GRID_LOG) return SOAP_OK;  // per-pixel tap widths: not modelled
      double taps = 0;  // nonzero taps of the window
      std::vector<double> ct(c->res_ntp);
      HIPCHK(c, hipMemcpyAsync(ct.data(), c->ctap.p, ct.size() * 8, hipMemcpyDeviceToHost, c->stream));
      HIPCHK(c, hipStreamSynchronize(c->stream));
      for (double v : ct) taps += v != 0.0;
      double spectra = 3.0 * c->P + 2;  // active spectra + the two quiet spectra
      // reads: raw 3P + quiet 2; writes: 3P + 3 frame slots
      *bytes = 8 * n * (3.0 * c->P + 2) + 8 * n * (3.0 * c->P + 3);
      *flops = 2 * taps * n * spectra;
      return SOAP_OK;
    }
    case ST_NORM:
      if (!c->have_lower) return fail(c, SOAP_ERR_STATE, "stage_work: run resolution first");
      *bytes = 2 * 8 * n * (3.0 * (c->P + 1));
      return SOAP_OK;
    case ST_CCF: {
      if (!c->have_ccf) return fail(c, SOAP_ERR_STATE, "stage_work: run ccf first");
      size_t ne = (size_t)c->n_v * c->n_lines;
      std::vector<int32_t> len(ne);
      HIPCHK(c, hipMemcpyAsync(len.data(), c->m_len.p, ne * 4, hipMemcpyDeviceToHost, c->stream));
      HIPCHK(c, hipStreamSynchronize(c->stream));
      double entries = 0;
      for (int32_t v : len) entries += v > 0 ? v : 0;
      double n_spec = 3.0 * (c->P + 1);
      // every trimmed spectrum read once; with spec_norm in the same launch, every
      // spectrum read and its normalised form written instead
      *bytes = c->ccf_fused ? 2 * 8 * n * n_spec : 8 * (n - 2.0 * c->trim) * n_spec;
      *flops = 2 * entries * n_spec;
      return SOAP_OK;
    }
    default:
      return SOAP_OK;
  }
}

int soap_get_raw(soap_ctx *c, int32_t phase, double *flux, double *bconv, double *tot) {
  if (!c || phase < 0 || phase >= c->P) return fail(c, SOAP_ERR_ARG, "get_raw: bad phase");
  if (!c->have_raw) return fail(c, SOAP_ERR_STATE, "get_raw: run integrate first");
  HIPCHK(c, hipSetDevice(c->device));
  double *outs[3] = {flux, bconv, tot};
  for (int t = 0; t < 3; t++)
    if (outs[t])
      HIPCHK(c, hipMemcpyAsync(outs[t], c->d_raw + ((size_t)phase * 3 + t) * c->n, c->n * 8, hipMemcpyDeviceToHost,
                               c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return SOAP_OK;
}

int soap_alloc_phases(soap_ctx *c, int32_t n_phase) {
  if (!c || n_phase < 0) return fail(c, SOAP_ERR_ARG, "alloc_phases: bad arguments");
  if (c->n == 0) return fail(c, SOAP_ERR_STATE, "alloc_phases: set the wavelength grid first");
  HIPCHK(c, hipSetDevice(c->device));
  int rc = dalloc(c, &c->d_raw, (size_t)std::max(1, n_phase) * 3 * c->n);
  if (rc) return rc;
  HIPCHK(c, launch_fill_f64(c->d_raw, (size_t)std::max(1, n_phase) * 3 * c->n, 0.0, c->stream));
  c->P = n_phase;
  c->phases.assign(n_phase, 0.0);
  c->reg_count.assign(n_phase, 0);
  c->regions.clear();
  c->geo.clear();
  c->geo_off.assign(n_phase + 1, 0);
  c->have_raw = true;
  c->have_lower = c->have_ccf = c->have_fit = false;
  return SOAP_OK;
}

int soap_set_raw(soap_ctx *c, int32_t phase, const double *flux, const double *bconv, const double *tot) {
  if (!c || phase < 0 || phase >= c->P || !flux || !bconv || !tot) return fail(c, SOAP_ERR_ARG, "set_raw: bad arguments");
  HIPCHK(c, hipSetDevice(c->device));
  const double *ins[3] = {flux, bconv, tot};
  for (int t = 0; t < 3; t++)
    HIPCHK(c, hipMemcpyAsync(c->d_raw + ((size_t)phase * 3 + t) * c->n, ins[t], c->n * 8, hipMemcpyHostToDevice,
                             c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->have_raw = true;
  return SOAP_OK;
}

int soap_phase_maps(soap_ctx *c, int32_t phase, double *temp_map, double *bi_map) {
  if (!c || phase < 0 || phase >= c->P || !temp_map || !bi_map) return fail(c, SOAP_ERR_ARG, "phase_maps: bad arguments");
  HIPCHK(c, hipSetDevice(c->device));
  size_t g2 = (size_t)c->G * c->G;
  double *d_tmp = nullptr;
  int rc = dalloc(c, &d_tmp, 2 * g2);
  if (rc) return rc;
  int g0 = c->geo_off[phase], g1 = c->geo_off[phase + 1];
  hipError_t e = launch_phase_maps(c->G, c->act.t_star, c->act.t_diff_spot, c->act.t_diff_fac, c->d_geo + g0, g1 - g0,
                                   d_tmp, d_tmp + g2, c->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(temp_map, d_tmp, g2 * 8, hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(bi_map, d_tmp + g2, g2 * 8, hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  dfree_c(c, d_tmp);
  if (e != hipSuccess) return fail(c, SOAP_ERR_HIP, "phase_maps: %s", hipGetErrorString(e));
  return SOAP_OK;
}

// ------------------------------------------------------------ resolution
// nb_pixel with Delta_lambda_cpu's single-precision sqrtf, SOAP_resolution.cu:18-49,241-244
static double rndup(double n, int nb) {
  double t = n * pow(10, nb) - floor(n * pow(10, nb));
  if (t >= 0.5) {
    n *= pow(10, nb);
    n = ceil(n);
    n /= pow(10, nb);
  } else {
    n *= pow(10, nb);
    n = floor(n);
    n /= pow(10, nb);
  }
  return n;
}

static int conv_window(const std::vector<double> &w, double sigma_ms) {
  double beta = sigma_ms / kC;
  double dl = -1 * w[0] * (1 - sqrtf((float)((1 + beta) / (1 - beta))));
  return (int)rndup((50. * dl) / (w[1] - w[0]), 0);
}

// Log-grid resolution taps c_d = exp(d h) expm1(h)/(fs sqrt(2 pi)) exp(-(expm1(d h)/fs)^2/2)
// (the generalisation of conv_inst's 0.005 A pixel), kept where c_d >= 2^-60 c_0,
// d in [-wid, wid-2] (SOAP_resolution.cu:76-95), uploaded once per (sigma, wid).
static int log_conv_taps(soap_ctx *c, double sigma_ms, int wid) {
  int rc;
  if (c->res_sigma != sigma_ms || c->res_wid != wid) {
    double fs = -1 * 1.0 * (1 - sqrt((1 + sigma_ms / kC) / (1 - sigma_ms / kC)));
    auto tap = [&](int d) {
      double u = expm1(d * c->h) / fs;
      return exp(d * c->h) * c->expm1_h / (fs * sqrt(2. * kPi)) * exp(-u * u / 2);
    };
    double c0 = tap(0);
    int dt = 0;
    for (int d = 1; d <= wid; d++)
      if (tap(d) >= c0 * 8.673617379884035e-19 || tap(-d) >= c0 * 8.673617379884035e-19) dt = d;
    int dlo = -std::min(wid, dt), dhi = std::min(wid - 2, dt);
    int ch = conv_chunk();
    int ntp = ((dhi - dlo + 1) + ch - 1) / ch * ch;
    std::vector<double> ct(ntp, 0.0);
    for (int d = dlo; d <= dhi; d++) ct[d - dlo] = tap(d);
    if ((rc = ensure(c, c->ctap, ct.size() * 8))) return rc;
    HIPCHK(c, hipMemcpyAsync(c->ctap.p, ct.data(), ct.size() * 8, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->res_sigma = sigma_ms;
    c->res_wid = wid;
    c->res_dlo = dlo;
    c->res_ntp = ntp;
  }
  return SOAP_OK;
}

// The convolution writes d_lraw and the tile min / max; spec_norm (d_lraw ->
// d_lower) runs here when norm is set, else in the same launch as the CCF (the
// series, ccf_impl).
// log grids: the convolution's taps (host, cached per sigma), its arguments and the
// output buffers of every spectrum
static int log_conv_setup(soap_ctx *c, double sigma_ms, int wid, ConvArgs &ca) {
  int rc;
  if ((rc = log_conv_taps(c, sigma_ms, wid))) return rc;
  ca = ConvArgs{};
  ca.n = c->n;
  ca.wid = wid;
  ca.dlo = c->res_dlo;
  ca.ntp = c->res_ntp;
  ca.ctap = (const double *)c->ctap.p;
  ca.n_tiles = (int)((c->n + conv_tile() - 1) / conv_tile());
  int n_spec = 3 * (c->P + 1);
  if ((rc = dalloc(c, &c->d_lower, (size_t)n_spec * c->n)) || (rc = dalloc(c, &c->d_lraw, (size_t)n_spec * c->n)) ||
      (rc = ensure(c, c->part, (size_t)n_spec * ca.n_tiles * 2 * 8)) || (rc = ensure(c, c->mm, (size_t)n_spec * 2 * 8)))
    return rc;
  return SOAP_OK;
}

// the quiet star's convolution alone, on the ctx stream (the series runs it on the
// quiet star's stream, right after the quiet star); a log grid with the quiet star set
static int resolution_quiet(soap_ctx *c, double sigma_ms) {
  if (!(sigma_ms > 0)) return fail(c, SOAP_ERR_ARG, "resolution: sigma must be > 0");
  int wid = conv_window(c->wave, sigma_ms);
  if (wid < 1 || 2 * (int64_t)wid + 2 >= c->n)
    return fail(c, SOAP_ERR_ARG, "resolution: window %d vs %lld pixels", wid, (long long)c->n);
  ConvArgs ca{};
  int rc;
  if ((rc = log_conv_setup(c, sigma_ms, wid, ca))) return rc;
  HIPCHK(c, launch_conv_quiet(ca, c->d_quiet, c->d_lraw + (size_t)3 * c->P * c->n,
                              (double *)c->part.p + (size_t)3 * c->P * ca.n_tiles * 2, c->stream));
  return SOAP_OK;
}

static int resolution_impl(soap_ctx *c, double sigma_ms, double quiet_weight, bool norm, bool quiet_conv = true) {
  if (!c) return SOAP_ERR_ARG;
  if (!c->have_quiet) return fail(c, SOAP_ERR_STATE, "resolution: quiet star missing");
  if (c->P > 0 && !c->have_raw) return fail(c, SOAP_ERR_STATE, "resolution: raw spectra missing");
  if (!(sigma_ms > 0)) return fail(c, SOAP_ERR_ARG, "resolution: sigma must be > 0");
  HIPCHK(c, hipSetDevice(c->device));
  int wid = conv_window(c->wave, sigma_ms);
  if (wid < 1 || 2 * (int64_t)wid + 2 >= c->n)
    return fail(c, SOAP_ERR_ARG, "resolution: window %d vs %lld pixels", wid, (long long)c->n);
  int rc;
  if (c->kind == SOAP_GRID_LINEAR) {
    // conv_inst with per-pixel widths, SOAP_resolution.cu:76-95 (quiet form: SOAP_initialize.cu:78-96)
    ConvLinArgs cl{};
    cl.n = c->n;
    cl.wid = wid;
    cl.n_tiles = (int)((c->n + conv_linear_tile() - 1) / conv_linear_tile());
    cl.n_active = 3 * c->P;
    cl.quiet_sel[0] = 1;  // frame slots [flux, bconv, tot] = [SED, normalised, SED]
    cl.quiet_sel[1] = 0;
    cl.quiet_sel[2] = 1;
    cl.c1 = 1 - sqrt((1 + sigma_ms / kC) / (1 - sigma_ms / kC));
    cl.quiet_weight = quiet_weight;
    cl.wave = c->d_wave;
    int n_spec = 3 * (c->P + 1);
    if ((rc = dalloc(c, &c->d_lower, (size_t)n_spec * c->n)) || (rc = dalloc(c, &c->d_lraw, (size_t)n_spec * c->n)) ||
        (rc = ensure(c, c->part, (size_t)n_spec * cl.n_tiles * 2 * 8)) || (rc = ensure(c, c->mm, (size_t)n_spec * 2 * 8)))
      return rc;
    double *part = (double *)c->part.p;
    {
      StageTimer tm(c, ST_CONV);
      hipError_t e = launch_conv_linear(cl, n_spec, c->d_quiet, c->d_raw, c->d_lraw, part, c->stream);
      if (e == hipErrorInvalidValue) return fail(c, SOAP_ERR_LIMIT, "resolution: window %d too wide", wid);
      HIPCHK(c, e);
    }
    {
      StageTimer tm(c, ST_MINMAX);
      HIPCHK(c, launch_minmax(n_spec, cl.n_tiles, part, (double *)c->mm.p, c->stream));
    }
    if (norm) {
      StageTimer tm(c, ST_NORM);
      HIPCHK(c, launch_norm(c->n, n_spec, (double *)c->mm.p, c->d_lraw, c->d_lower, c->stream));
    }
    c->have_lower = norm;
    c->have_ccf = c->have_fit = false;
    return SOAP_OK;
  }
  ConvArgs ca{};
  if ((rc = log_conv_setup(c, sigma_ms, wid, ca))) return rc;
  int n_spec = 3 * (c->P + 1);
  double *part = (double *)c->part.p;
  if (quiet_conv) {
    StageTimer tm(c, ST_CONV_QUIET);
    HIPCHK(c, launch_conv_quiet(ca, c->d_quiet, c->d_lraw + (size_t)3 * c->P * c->n,
                                part + (size_t)3 * c->P * ca.n_tiles * 2, c->stream));
  }
  {
    StageTimer tm(c, ST_CONV);
    HIPCHK(c, launch_conv_active(ca, c->d_quiet, c->d_raw, c->P, c->d_lraw, part, c->stream));
  }
  {
    StageTimer tm(c, ST_MINMAX);
    HIPCHK(c, launch_minmax(n_spec, ca.n_tiles, part, (double *)c->mm.p, c->stream));
  }
  if (norm) {
    StageTimer tm(c, ST_NORM);
    HIPCHK(c, launch_norm(c->n, n_spec, (double *)c->mm.p, c->d_lraw, c->d_lower, c->stream));
  }
  c->have_lower = norm;
  c->have_ccf = c->have_fit = false;
  return SOAP_OK;
}

int soap_resolution(soap_ctx *c, double sigma_ms, double quiet_weight) {
  return resolution_impl(c, sigma_ms, quiet_weight, true);
}

int soap_get_lower(soap_ctx *c, int32_t phase, int type, double *out) {
  if (!c || !out || type < 0 || type > 2 || phase < -1 || phase >= c->P) return fail(c, SOAP_ERR_ARG, "get_lower: bad arguments");
  if (!c->have_lower) return fail(c, SOAP_ERR_STATE, "get_lower: run resolution first");
  HIPCHK(c, hipSetDevice(c->device));
  int frame = phase < 0 ? c->P : phase;
  HIPCHK(c, hipMemcpyAsync(out, c->d_lower + ((size_t)frame * 3 + type) * c->n, c->n * 8, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return SOAP_OK;
}

int soap_get_norm(soap_ctx *c, int32_t phase, int type, double *minmax2) {
  if (!c || !minmax2 || type < 0 || type > 2 || phase < -1 || phase >= c->P) return fail(c, SOAP_ERR_ARG, "get_norm: bad arguments");
  if (!c->have_lower) return fail(c, SOAP_ERR_STATE, "get_norm: run resolution first");
  HIPCHK(c, hipSetDevice(c->device));
  int sp = (phase >= 0 ? phase : c->P) * 3 + type;
  HIPCHK(c, hipMemcpyAsync(minmax2, (double *)c->mm.p + sp * 2, 16, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return SOAP_OK;
}

// ------------------------------------------------------------------ CCF
int soap_set_mask(soap_ctx *c, const double *line_l, const double *line_w, int32_t n_lines, const double *vgrid,
                  int32_t n_v, int32_t trim) {
  if (!c || n_lines < 1 || !line_l || !line_w || !vgrid || n_v < 3 || trim < 0) return fail(c, SOAP_ERR_ARG, "set_mask: bad arguments");
  if (c->n == 0) return fail(c, SOAP_ERR_STATE, "set_mask: set the wavelength grid first");
  if (2 * (int64_t)trim + 8 > c->n) return fail(c, SOAP_ERR_ARG, "set_mask: trim %d too large", trim);
  if (n_v > 1024) return fail(c, SOAP_ERR_LIMIT, "set_mask: at most 1024 velocities");
  HIPCHK(c, hipSetDevice(c->device));
  c->n_lines = n_lines;
  c->n_v = n_v;
  c->trim = trim;
  int rc;
  if ((rc = dalloc(c, &c->d_line_l, n_lines)) || (rc = dalloc(c, &c->d_line_w, n_lines)) ||
      (rc = dalloc(c, &c->d_vgrid, n_v)) || (rc = dalloc(c, &c->d_appod, n_v)) ||
      (rc = dalloc(c, &c->d_w_sorted, n_lines)))
    return rc;
  // wavelength order of the lines and the neighbour scan depth for overwrites
  std::vector<int32_t> srt(n_lines), rnk(n_lines);
  std::iota(srt.begin(), srt.end(), 0);
  std::stable_sort(srt.begin(), srt.end(), [&](int a, int b) { return line_l[a] < line_l[b]; });
  for (int i = 0; i < n_lines; i++) rnk[srt[i]] = i;
  double lmax = 0;
  for (int i = 0; i < n_lines; i++) lmax = std::max(lmax, fabs(line_l[i]));
  double beta = 820. / kC;
  // width of a mask pixel: the log grid's, or the 0.005 A of the linear index formula
  double pix = c->kind == SOAP_GRID_LOG ? lmax * c->expm1_h : 0.005;
  double lim = 2 * lmax * (sqrt((1 + beta) / (1 - beta)) - 1) + 4 * pix + 1e-9;
  int scan = 1;
  for (int i = 0, j = 0; i < n_lines; i++) {
    while (line_l[srt[i]] - line_l[srt[j]] > lim) j++;
    scan = std::max(scan, i - j + 1);
  }
  c->mask_scan = scan;
  // a mask hole must span at most 22 pixels (24-bit alive masks per line)
  double hole_px = c->kind == SOAP_GRID_LOG ? (sqrt((1 + beta) / (1 - beta)) - 1) * 1.001 / c->h + 3
                                             : lmax * 1.001 * (sqrt((1 + beta) / (1 - beta)) - 1) / 0.005 + 3;
  if (hole_px > 22) return fail(c, SOAP_ERR_LIMIT, "set_mask: mask holes span %.1f pixels (> 22)", hole_px);
  // the reference slices the extended-grid mask at a fixed 20000 pixels (CCF_func.py:8-9)
  if (c->kind == SOAP_GRID_LINEAR && (long)(100.0 / c->step) != 20000)
    return fail(c, SOAP_ERR_LIMIT, "set_mask: the linear-grid CCF needs int(100/step) == 20000 (step 0.005)");
  // appodisation window, run_all_SOAP.py:306-311
  std::vector<double> ap;
  {
    int nb0 = 5, period = (int)((n_v - 1) / 2.), len_app = (int)(period / 2.);
    for (int i = 0; i < nb0; i++) ap.push_back(0.0);
    for (int a = 0; a < len_app; a++) ap.push_back(0.5 * cos(2 * kPi / period * a - kPi) + 0.5);
    for (int i = 0; i < n_v - period - 2 * nb0; i++) ap.push_back(1.0);
    for (int a = len_app - 1; a >= 0; a--) ap.push_back(0.5 * cos(2 * kPi / period * a - kPi) + 0.5);
    for (int i = 0; i < nb0; i++) ap.push_back(0.0);
    if ((int)ap.size() != n_v) return fail(c, SOAP_ERR_ARG, "set_mask: %d velocities too few for the appodisation", n_v);
  }
  size_t ne = (size_t)n_v * n_lines;
  if ((rc = ensure(c, c->m_start, ne * 4)) || (rc = ensure(c, c->m_len, ne * 4)) || (rc = ensure(c, c->m_val, ne * 16)) ||
      (rc = ensure(c, c->m_alive, ne * 4)) || (rc = ensure(c, c->sorted, (size_t)n_lines * 4)) ||
      (rc = ensure(c, c->rank, (size_t)n_lines * 4)))
    return rc;
  HIPCHK(c, hipMemcpyAsync(c->d_line_l, line_l, n_lines * 8, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->d_line_w, line_w, n_lines * 8, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->d_vgrid, vgrid, n_v * 8, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->d_appod, ap.data(), n_v * 8, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->sorted.p, srt.data(), n_lines * 4, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->rank.p, rnk.data(), n_lines * 4, hipMemcpyHostToDevice, c->stream));
  std::vector<double> ws(n_lines);
  for (int q = 0; q < n_lines; q++) ws[q] = line_w[srt[q]];
  HIPCHK(c, hipMemcpyAsync(c->d_w_sorted, ws.data(), n_lines * 8, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->have_mask = true;
  c->mask_built = false;
  c->have_ccf = c->have_fit = false;
  return SOAP_OK;
}

static int build_mask(soap_ctx *c) {
  StageTimer tm(c, ST_MASK);
  int64_t n_t = c->n - 2 * (int64_t)c->trim;
  if (c->kind == SOAP_GRID_LOG)
    HIPCHK(c, launch_mask_build(c->d_line_l, c->d_line_w, (const int32_t *)c->sorted.p, c->n_lines, c->d_vgrid, c->n_v,
                                c->wave[c->trim], c->h, (int32_t *)c->m_start.p, (int32_t *)c->m_len.p,
                                (double2 *)c->m_val.p, c->stream));
  else
    HIPCHK(c, launch_mask_build_linear(c->d_line_l, c->d_line_w, (const int32_t *)c->sorted.p, c->n_lines, c->d_vgrid,
                                       c->n_v, c->d_wave + c->trim, n_t, c->step, (int32_t *)c->m_start.p,
                                       (int32_t *)c->m_len.p, (double2 *)c->m_val.p, c->stream));
  HIPCHK(c, launch_mask_overwrite((const int32_t *)c->sorted.p, c->n_lines, c->n_v, c->mask_scan,
                                  (const int32_t *)c->m_start.p, (const int32_t *)c->m_len.p, (uint32_t *)c->m_alive.p,
                                  c->stream));
  int rc;
  if ((rc = ensure(c, c->wsum, (size_t)c->n_v * 8))) return rc;
  HIPCHK(c, launch_mask_wsum(c->n_lines, c->n_v, n_t, (const int32_t *)c->m_start.p, (const uint32_t *)c->m_alive.p,
                             (const double2 *)c->m_val.p, c->d_w_sorted, (double *)c->wsum.p, c->stream));
  // line footprints: the CCF stages them in LDS
  if ((rc = ensure(c, c->fp, (size_t)c->n_lines * 8))) return rc;
  HIPCHK(c, launch_mask_footprint(c->n_lines, c->n_v, n_t, (const int32_t *)c->m_start.p, (const uint32_t *)c->m_alive.p,
                                  (int2 *)c->fp.p, c->stream));
  std::vector<int32_t> fp((size_t)c->n_lines * 2);
  HIPCHK(c, hipMemcpyAsync(fp.data(), c->fp.p, fp.size() * 4, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  int fmax = 0;
  for (int q = 0; q < c->n_lines; q++) fmax = std::max(fmax, fp[2 * q + 1]);
  c->ccf_cap = ccf_lds_cap(fmax);
  c->mask_built = true;
  c->diff.mt_ready = false;
  return SOAP_OK;
}

// Pixel-major form of the built mask for the CCF adjoint: for every trimmed
// pixel t, the (velocity, weight) pairs of the live mask pixels that the CCF
// multiplies with spectrum pixel t, pairs with the same velocity merged in line
// order; stored in slices of 64 rows (one wave), column-major inside a slice.
// Built on the host once per mask.
static int diff_mask_transpose(soap_ctx *c) {
  auto &d = c->diff;
  if (d.mt_ready) return SOAP_OK;
  const int nl = c->n_lines, nv = c->n_v;
  const int64_t n_t = c->n - 2 * (int64_t)c->trim;
  const size_t ne = (size_t)nl * nv;
  std::vector<int32_t> st(ne);
  std::vector<uint32_t> pk(ne);
  std::vector<double> wv(2 * ne), ws(nl);
  HIPCHK(c, hipMemcpyAsync(st.data(), c->m_start.p, ne * 4, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemcpyAsync(pk.data(), c->m_alive.p, ne * 4, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemcpyAsync(wv.data(), c->m_val.p, ne * 16, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemcpyAsync(ws.data(), c->d_w_sorted, (size_t)nl * 8, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  // the live (pixel, velocity, weight) triples, in line order, as the CCF visits them
  auto visit = [&](auto &&f) {
    for (int q = 0; q < nl; q++)
      for (int iv = 0; iv < nv; iv++) {
        size_t e = (size_t)q * nv + iv;
        int ln = (int)(pk[e] & 0xffu);
        uint32_t al = pk[e] >> 8;
        for (int b = 0; b < ln; b++) {
          int64_t pp = (int64_t)st[e] + b;
          if (!((al >> b) & 1u) || pp < 0 || pp >= n_t) continue;
          f(pp, iv, b == ln - 1 ? wv[2 * e + 1] : (b == 0 ? wv[2 * e] : ws[q]));
        }
      }
  };
  std::vector<int64_t> cnt(n_t + 1, 0);
  visit([&](int64_t pp, int, double) { cnt[pp + 1]++; });
  for (int64_t t = 0; t < n_t; t++) cnt[t + 1] += cnt[t];
  if (cnt[n_t] > 0x7fffffffL) return fail(c, SOAP_ERR_LIMIT, "diff: mask too large for the adjoint table");
  std::vector<int32_t> col((size_t)cnt[n_t]);
  std::vector<double> val((size_t)cnt[n_t]);
  std::vector<int64_t> fillp(cnt.begin(), cnt.end() - 1);
  visit([&](int64_t pp, int iv, double w) {
    col[fillp[pp]] = iv;
    val[fillp[pp]++] = w;
  });
  // merge equal velocities within each row (stable: line order kept per velocity)
  std::vector<std::vector<std::pair<int32_t, double>>> rows(n_t);
  std::vector<std::pair<int32_t, int64_t>> idx;
  for (int64_t t = 0; t < n_t; t++) {
    idx.clear();
    for (int64_t e = cnt[t]; e < cnt[t + 1]; e++) idx.push_back({col[e], e});
    std::stable_sort(idx.begin(), idx.end(), [](const auto &x, const auto &y) { return x.first < y.first; });
    auto &r = rows[t];
    for (size_t k = 0; k < idx.size(); k++) {
      if (k > 0 && idx[k].first == idx[k - 1].first)
        r.back().second += val[idx[k].second];
      else
        r.push_back({idx[k].first, val[idx[k].second]});
    }
  }
  // slices of 64 rows, column-major within a slice, zero-padded to the slice's widest row
  const int64_t ns = (n_t + 63) / 64;
  std::vector<int64_t> sp(ns + 1, 0);
  std::vector<int32_t> sw(ns, 0);
  for (int64_t sl = 0; sl < ns; sl++) {
    int w = 0;
    for (int64_t t = sl * 64; t < std::min(n_t, sl * 64 + 64); t++) w = std::max(w, (int)rows[t].size());
    sw[sl] = w;
    sp[sl + 1] = sp[sl] + 64L * w;
  }
  std::vector<int32_t> mcol((size_t)std::max<int64_t>(1, sp[ns]), 0);
  std::vector<double> mval((size_t)std::max<int64_t>(1, sp[ns]), 0.0);
  for (int64_t t = 0; t < n_t; t++) {
    int64_t sl = t / 64;
    for (size_t k = 0; k < rows[t].size(); k++) {
      int64_t e = sp[sl] + 64L * (int64_t)k + (t & 63);
      mcol[e] = rows[t][k].first;
      mval[e] = rows[t][k].second;
    }
  }
  int rc;
  if ((rc = ensure(c, d.mt_ptr, sp.size() * 8)) || (rc = ensure(c, d.mt_w, std::max<size_t>(1, sw.size()) * 4)) ||
      (rc = ensure(c, d.mt_col, mcol.size() * 4)) || (rc = ensure(c, d.mt_val, mval.size() * 8)))
    return rc;
  HIPCHK(c, hipMemcpyAsync(d.mt_ptr.p, sp.data(), sp.size() * 8, hipMemcpyHostToDevice, c->stream));
  if (!sw.empty()) HIPCHK(c, hipMemcpyAsync(d.mt_w.p, sw.data(), sw.size() * 4, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(d.mt_col.p, mcol.data(), mcol.size() * 4, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(d.mt_val.p, mval.data(), mval.size() * 8, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  d.mt_ready = true;
  return SOAP_OK;
}


// CCF chunk partials of n_spec spectra (stride n) with the built mask, line
// footprints staged in LDS
static hipError_t ctx_ccf(soap_ctx *c, const double *spec, int n_spec, double *part, hipStream_t st) {
  return launch_ccf_lds(spec, c->n, c->trim, n_spec, c->n_lines, c->n_v, c->n - 2 * (int64_t)c->trim,
                        (const int32_t *)c->m_start.p, (const uint32_t *)c->m_alive.p, (const double2 *)c->m_val.p,
                        c->d_w_sorted, (const int2 *)c->fp.p, c->ccf_cap, part, st);
}

// The CCF runs on the unnormalised spectra d_lraw (the reduction applies the
// spec_norm map, see k_ccf_reduce).  With norm set, spec_norm itself (d_lraw ->
// d_lower) runs in the same launch (the series: resolution_impl without norm
// just before).
static int ccf_impl(soap_ctx *c, bool norm) {
  if (!c) return SOAP_ERR_ARG;
  if (!c->have_mask) return fail(c, SOAP_ERR_STATE, "ccf: set the mask first");
  if (!norm && !c->have_lower) return fail(c, SOAP_ERR_STATE, "ccf: run resolution first");
  HIPCHK(c, hipSetDevice(c->device));
  int rc;
  if (!c->mask_built && (rc = build_mask(c))) return rc;
  int n_spec = (c->P + 1) * 3;
  if ((rc = dalloc(c, &c->d_ccf, (size_t)n_spec * c->n_v)) ||
      (rc = ensure(c, c->ccf_part, (size_t)ccf_chunks(c->n_lines) * n_spec * c->n_v * 8)))
    return rc;
  if (norm) {
    NormCcfArgs a{};
    a.n = c->n;
    a.trim = c->trim;
    a.n_t = c->n - 2 * (int64_t)c->trim;
    a.n_spec = n_spec;
    a.n_lines = c->n_lines;
    a.n_v = c->n_v;
    a.mm = (const double *)c->mm.p;
    a.raw = c->d_lraw;
    a.out = c->d_lower;
    a.m_start = (const int32_t *)c->m_start.p;
    a.m_pk = (const uint32_t *)c->m_alive.p;
    a.m_wv = (const double2 *)c->m_val.p;
    a.w_sorted = c->d_w_sorted;
    a.fp = (const int2 *)c->fp.p;
    a.cap = c->ccf_cap;
    a.part = (double *)c->ccf_part.p;
    StageTimer tm(c, ST_CCF);
    HIPCHK(c, launch_norm_ccf(a, c->stream));
    c->have_lower = true;
    c->ccf_fused = true;
  } else {
    c->ccf_fused = false;
    StageTimer tm(c, ST_CCF);
    HIPCHK(c, ctx_ccf(c, c->d_lraw, n_spec, (double *)c->ccf_part.p, c->stream));
  }
  {
    StageTimer tm(c, ST_CCF_REDUCE);
    HIPCHK(c, launch_ccf_reduce((const double *)c->ccf_part.p, ccf_chunks(c->n_lines), n_spec, c->n_v, c->d_appod,
                                (const double *)c->mm.p, (const double *)c->wsum.p, c->d_ccf, c->stream));
  }
  c->have_ccf = true;
  c->have_fit = false;
  return SOAP_OK;
}

int soap_ccf(soap_ctx *c) { return ccf_impl(c, false); }

int soap_get_ccf(soap_ctx *c, double *out) {
  if (!c || !out) return SOAP_ERR_ARG;
  if (!c->have_ccf) return fail(c, SOAP_ERR_STATE, "get_ccf: run ccf first");
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipMemcpyAsync(out, c->d_ccf, (size_t)(c->P + 1) * 3 * c->n_v * 8, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return SOAP_OK;
}

static int ensure_fit(soap_ctx *c, int n_ccf, int n_v) {
  int rc;
  if ((rc = ensure(c, c->f_par, (size_t)n_ccf * 5 * 8)) || (rc = ensure(c, c->f_model, (size_t)n_ccf * n_v * 8)) ||
      (rc = ensure(c, c->f_depth, (size_t)n_ccf * 91 * 8)) || (rc = ensure(c, c->f_bis, (size_t)n_ccf * 91 * 8)) ||
      (rc = ensure(c, c->f_iter, (size_t)n_ccf * 4)))
    return rc;
  return SOAP_OK;
}

int soap_fit(soap_ctx *c) {
  if (!c) return SOAP_ERR_ARG;
  if (!c->have_ccf) return fail(c, SOAP_ERR_STATE, "fit: run ccf first");
  HIPCHK(c, hipSetDevice(c->device));
  int n_ccf = (c->P + 1) * 3;
  int rc = ensure_fit(c, n_ccf, c->n_v);
  if (rc) return rc;
  {
    StageTimer tm(c, ST_FIT);
    HIPCHK(c, launch_fit(c->d_vgrid, c->n_v, c->d_ccf, nullptr, n_ccf, (double *)c->f_par.p, (double *)c->f_model.p,
                         (double *)c->f_depth.p, (double *)c->f_bis.p, (int32_t *)c->f_iter.p, c->stream));
  }
  c->have_fit = true;
  return SOAP_OK;
}

int soap_get_fit(soap_ctx *c, double *par, double *model, double *depth, double *bis) {
  if (!c) return SOAP_ERR_ARG;
  if (!c->have_fit) return fail(c, SOAP_ERR_STATE, "get_fit: run fit first");
  HIPCHK(c, hipSetDevice(c->device));
  size_t n_ccf = (size_t)(c->P + 1) * 3;
  if (par) HIPCHK(c, hipMemcpyAsync(par, c->f_par.p, n_ccf * 5 * 8, hipMemcpyDeviceToHost, c->stream));
  if (model) HIPCHK(c, hipMemcpyAsync(model, c->f_model.p, n_ccf * c->n_v * 8, hipMemcpyDeviceToHost, c->stream));
  if (depth) HIPCHK(c, hipMemcpyAsync(depth, c->f_depth.p, n_ccf * 91 * 8, hipMemcpyDeviceToHost, c->stream));
  if (bis) HIPCHK(c, hipMemcpyAsync(bis, c->f_bis.p, n_ccf * 91 * 8, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return SOAP_OK;
}

int soap_bis_inject(soap_ctx *c, const soap_bis_inject_params *p, const double *wave, const double *flux, int64_t n,
                    double *out) {
  if (!c || !p || !wave || !flux || !out) return fail(c, SOAP_ERR_ARG, "bis_inject: null argument");
  if (n < 2 || n > (int64_t)1 << 30) return fail(c, SOAP_ERR_ARG, "bis_inject: n = %lld", (long long)n);
  if ((p->mode != 0 && p->mode != 1) || p->n_up < 1 || p->n_up > kInjMaxCoef ||
      (p->mode == 1 && (p->n_low < 1 || p->n_low > kInjMaxCoef)))
    return fail(c, SOAP_ERR_ARG, "bis_inject: mode %d with %d / %d coefficients", p->mode, p->n_up, p->n_low);
  HIPCHK(c, hipSetDevice(c->device));
  InjectArgs a{};
  a.mode = p->mode;
  a.n_up = p->n_up;
  a.n_low = p->mode == 1 ? p->n_low : 0;
  for (int i = 0; i < kInjMaxCoef; i++) {
    a.up[i] = p->up[i];
    a.low[i] = p->low[i];
  }
  a.dep_cut = p->dep_cut;
  a.vel_cut = p->vel_cut;
  a.velo_shift = p->velo_shift;
  a.scale = p->scale;
  size_t sort_bytes = 0;
  HIPCHK(c, inject_sort_bytes(n, &sort_bytes));
  const size_t nb = (size_t)inject_blocks(n);
  const size_t arrays = (size_t)6 * n * 8, parts = (nb * 3 + 4) * 8;
  int rc;
  if ((rc = ensure(c, c->inj, arrays + parts + sort_bytes + 256))) return rc;
  double *d_w = (double *)c->inj.p, *d_f = d_w + n, *d_m = d_f + n, *d_xs = d_m + n, *d_ys = d_xs + n,
         *d_out = d_ys + n, *d_part = d_out + n, *d_sum = d_part + nb * 3;
  void *d_tmp = (void *)(((uintptr_t)(d_sum + 4) + 255) & ~(uintptr_t)255);
  HIPCHK(c, hipMemcpyAsync(d_w, wave, n * 8, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(d_f, flux, n * 8, hipMemcpyHostToDevice, c->stream));
  hipError_t e = launch_inject_move(a, d_w, d_f, n, d_m, d_part, d_sum, c->stream);
  if (e == hipErrorInvalidValue) return fail(c, SOAP_ERR_ARG, "bis_inject: bad sizes");
  HIPCHK(c, e);
  double summary[3];
  HIPCHK(c, hipMemcpyAsync(summary, d_sum, sizeof summary, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  const double *xs = d_m, *ys = d_f;
  if (summary[2] != 0.0) {  // some moved pixel passed its neighbour: sort (stable) by moved wavelength
    HIPCHK(c, launch_inject_sort(d_tmp, sort_bytes, d_m, d_xs, d_f, d_ys, n, c->stream));
    xs = d_xs;
    ys = d_ys;
  }
  HIPCHK(c, launch_inject_interp(d_w, xs, ys, n, d_sum, d_out, c->stream));
  HIPCHK(c, hipMemcpyAsync(out, d_out, n * 8, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return SOAP_OK;
}

int soap_phoenix_interp(soap_ctx *c, const soap_phoenix_params *p, const double *wave_vac, const double *spec_low,
                        const double *spec_high, int64_t n, double *wave_air, double *flux, int64_t *n_out) {
  if (!c || !p || !wave_vac || !spec_low || !spec_high || !wave_air || !flux || !n_out)
    return fail(c, SOAP_ERR_ARG, "phoenix_interp: null argument");
  if (n < 1 || n > (int64_t)1 << 30) return fail(c, SOAP_ERR_ARG, "phoenix_interp: n = %lld", (long long)n);
  if (!(p->t_high != p->t_low) || !std::isfinite(p->temp) || !std::isfinite(p->t_low) || !std::isfinite(p->t_high))
    return fail(c, SOAP_ERR_ARG, "phoenix_interp: temperatures %g in [%g, %g]", p->temp, p->t_low, p->t_high);
  HIPCHK(c, hipSetDevice(c->device));
  PhoenixArgs a{};
  a.frac = (p->temp - p->t_low) / (p->t_high - p->t_low);
  a.vlo = p->wmin - 500;
  a.vhi = p->wmax + 500;
  a.wmin = p->wmin;
  a.wmax = p->wmax;
  a.f32 = p->f32 != 0;
  size_t sel_bytes = 0;
  HIPCHK(c, phoenix_select_bytes(n, &sel_bytes));
  const size_t arrays = (size_t)7 * n * 8, flags = ((size_t)n + 255) / 256 * 256;
  int rc;
  if ((rc = ensure(c, c->phx, arrays + flags + 256 + sel_bytes + 256))) return rc;
  double *d_w = (double *)c->phx.p, *d_lo = d_w + n, *d_hi = d_lo + n, *d_air = d_hi + n, *d_fl = d_air + n,
         *d_oair = d_fl + n, *d_ofl = d_oair + n;
  char *d_flag = (char *)(d_ofl + n);
  int64_t *d_cnt = (int64_t *)(d_flag + flags);
  void *d_tmp = (void *)(((uintptr_t)(d_cnt + 2) + 255) & ~(uintptr_t)255);
  HIPCHK(c, hipMemcpyAsync(d_w, wave_vac, n * 8, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(d_lo, spec_low, n * 8, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(d_hi, spec_high, n * 8, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, launch_phoenix(a, d_w, d_lo, d_hi, n, d_air, d_fl, d_flag, c->stream));
  HIPCHK(c, launch_phoenix_select(d_tmp, sel_bytes, d_air, d_flag, d_oair, d_cnt, n, c->stream));
  HIPCHK(c, launch_phoenix_select(d_tmp, sel_bytes, d_fl, d_flag, d_ofl, d_cnt + 1, n, c->stream));
  int64_t cnt[2] = {0, 0};
  HIPCHK(c, hipMemcpyAsync(cnt, d_cnt, sizeof cnt, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (cnt[0] != cnt[1] || cnt[0] < 0 || cnt[0] > n)
    return fail(c, SOAP_ERR_HIP, "phoenix_interp: selection counts %lld / %lld", (long long)cnt[0], (long long)cnt[1]);
  if (cnt[0] > 0) {
    HIPCHK(c, hipMemcpyAsync(wave_air, d_oair, cnt[0] * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(flux, d_ofl, cnt[0] * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
  }
  *n_out = cnt[0];
  return SOAP_OK;
}

int soap_fit_ccf_host_err(soap_ctx *c, const double *v, int32_t n_v, const double *y, const double *err,
                          int32_t n_ccf, double *par, double *model, double *depth, double *bis, int32_t *status) {
  if (!c || !v || !y || n_v < 3 || n_v > 1024 || n_ccf < 1 || !par) return fail(c, SOAP_ERR_ARG, "fit_ccf_host: bad arguments");
  if (err)
    for (size_t i = 0; i < (size_t)n_ccf * n_v; i++)
      if (!std::isfinite(err[i]) || err[i] == 0.0)
        return fail(c, SOAP_ERR_ARG, "fit_ccf_host: error %zu is %g (must be finite and non-zero)", i, err[i]);
  HIPCHK(c, hipSetDevice(c->device));
  double *dv = nullptr, *dy = nullptr, *de = nullptr, *dpar = nullptr, *dmod = nullptr, *ddep = nullptr, *dbis = nullptr;
  int32_t *dst = nullptr;
  int rc = SOAP_OK;
  if ((rc = dalloc(c, &dv, n_v)) || (rc = dalloc(c, &dy, (size_t)n_ccf * n_v)) ||
      (err && (rc = dalloc(c, &de, (size_t)n_ccf * n_v))) || (rc = dalloc(c, &dpar, (size_t)n_ccf * 5)) ||
      (rc = dalloc(c, &dmod, (size_t)n_ccf * n_v)) || (rc = dalloc(c, &ddep, (size_t)n_ccf * 91)) ||
      (rc = dalloc(c, &dbis, (size_t)n_ccf * 91)) || (status && (rc = dalloc(c, &dst, (size_t)n_ccf)))) {
    for (double *p : {dv, dy, de, dpar, dmod, ddep, dbis}) dfree_c(c, p);
    dfree_c(c, dst);
    return rc;
  }
  hipError_t e = hipMemcpyAsync(dv, v, n_v * 8, hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(dy, y, (size_t)n_ccf * n_v * 8, hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess && err) e = hipMemcpyAsync(de, err, (size_t)n_ccf * n_v * 8, hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess) e = launch_fit(dv, n_v, dy, de, n_ccf, dpar, dmod, ddep, dbis, nullptr, c->stream, dst);
  if (e == hipSuccess) e = hipMemcpyAsync(par, dpar, (size_t)n_ccf * 5 * 8, hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess && status) e = hipMemcpyAsync(status, dst, (size_t)n_ccf * 4, hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess && model) e = hipMemcpyAsync(model, dmod, (size_t)n_ccf * n_v * 8, hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess && depth) e = hipMemcpyAsync(depth, ddep, (size_t)n_ccf * 91 * 8, hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess && bis) e = hipMemcpyAsync(bis, dbis, (size_t)n_ccf * 91 * 8, hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  for (double *p : {dv, dy, de, dpar, dmod, ddep, dbis}) dfree_c(c, p);
  dfree_c(c, dst);
  if (e != hipSuccess) return fail(c, SOAP_ERR_HIP, "fit_ccf_host: %s", hipGetErrorString(e));
  return SOAP_OK;
}

int soap_fit_ccf_host(soap_ctx *c, const double *v, int32_t n_v, const double *y, int32_t n_ccf, double *par,
                      double *model, double *depth, double *bis) {
  return soap_fit_ccf_host_err(c, v, n_v, y, nullptr, n_ccf, par, model, depth, bis, nullptr);
}

int soap_fit_ccf_host_noalloc(soap_ctx *c, const double *v, int32_t n_v, const double *y, const double *err,
                              int32_t n_ccf, double *par, double *model, double *depth, double *bis,
                              int32_t *status) {
  if (!c || !v || !y || n_v < 3 || n_v > 1024 || n_ccf < 1 || !par)
    return fail(c, SOAP_ERR_ARG, "fit_ccf_host_noalloc: bad arguments");
  if (err)
    for (size_t i = 0; i < (size_t)n_ccf * n_v; i++)
      if (!std::isfinite(err[i]) || err[i] == 0.0)
        return fail(c, SOAP_ERR_ARG, "fit_ccf_host_noalloc: error %zu is %g (must be finite and non-zero)", i,
                    err[i]);
  // staging layout, each part rounded up to 32 doubles (256 bytes)
  auto up = [](size_t d) { return (d + 31) / 32 * 32; };
  const size_t nv = (size_t)n_v, nc = (size_t)n_ccf;
  const size_t o_v = 0, o_y = o_v + up(nv), o_e = o_y + up(nc * nv), o_par = o_e + (err ? up(nc * nv) : 0),
               o_mod = o_par + up(nc * 5), o_dep = o_mod + up(nc * nv), o_bis = o_dep + up(nc * 91),
               o_st = o_bis + up(nc * 91), total = o_st + (status ? up((nc + 1) / 2) : 0);
  if (total > kScratchDoubles)
    return fail(c, SOAP_ERR_LIMIT,
                "fit_ccf_host_noalloc: %d CCFs of %d points need %zu staging doubles, the limit is %zu "
                "(16 MiB per GPU); use soap_fit_ccf_host_err or fewer CCFs per call",
                n_ccf, n_v, total, kScratchDoubles);
  double *base = g_scratch_ptr[c->device];
  if (!base) return fail(c, SOAP_ERR_STATE, "fit_ccf_host_noalloc: no staging address for device %d", c->device);
  std::lock_guard<std::mutex> g(g_scratch_lock[c->device]);
  HIPCHK(c, hipSetDevice(c->device));
  hipError_t e = hipMemcpyAsync(base + o_v, v, nv * 8, hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(base + o_y, y, nc * nv * 8, hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess && err) e = hipMemcpyAsync(base + o_e, err, nc * nv * 8, hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess)
    e = launch_fit(base + o_v, n_v, base + o_y, err ? base + o_e : nullptr, n_ccf, base + o_par, base + o_mod,
                   base + o_dep, base + o_bis, nullptr, c->stream, status ? (int32_t *)(base + o_st) : nullptr);
  if (e == hipSuccess) e = hipMemcpyAsync(par, base + o_par, nc * 5 * 8, hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess && model) e = hipMemcpyAsync(model, base + o_mod, nc * nv * 8, hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess && depth) e = hipMemcpyAsync(depth, base + o_dep, nc * 91 * 8, hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess && bis) e = hipMemcpyAsync(bis, base + o_bis, nc * 91 * 8, hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess && status) e = hipMemcpyAsync(status, base + o_st, nc * 4, hipMemcpyDeviceToHost, c->stream);
  // the staging is reused by the next call on this device: wait before releasing it
  hipError_t e2 = hipStreamSynchronize(c->stream);
  if (e == hipSuccess) e = e2;
  if (e != hipSuccess) return fail(c, SOAP_ERR_HIP, "fit_ccf_host_noalloc: %s", hipGetErrorString(e));
  return SOAP_OK;
}

// ------------------------------------------------------------------ a11: differentiable path
int soap_diff_setup(soap_ctx *c, int32_t grid, int32_t n_phase, int32_t map_res, double sigma_ms, const double *lb) {
  if (!c || grid < 2 || grid > 8192 || n_phase < 1 || map_res < 2 || map_res > 8192 || !(sigma_ms > 0))
    return fail(c, SOAP_ERR_ARG, "diff_setup: bad arguments");
  if (c->kind != SOAP_GRID_LOG) return fail(c, SOAP_ERR_STATE, "diff_setup: the differentiable path needs a log grid");
  if (c->n == 0 || c->n_mu == 0) return fail(c, SOAP_ERR_STATE, "diff_setup: set wave and mu bounds first");
  if (!c->have_mask) return fail(c, SOAP_ERR_STATE, "diff_setup: set the mask first");
  if (c->n_v < 3 || c->n_v > kMaxVDiff) return fail(c, SOAP_ERR_LIMIT, "diff_setup: %d CCF velocities", c->n_v);
  if (c->n_mu > 1024) return fail(c, SOAP_ERR_LIMIT, "diff_setup: %d mu bins (at most 1024)", c->n_mu);
  HIPCHK(c, hipSetDevice(c->device));
  auto &d = c->diff;
  size_t g2 = (size_t)grid * grid;
  int rc;
  if ((rc = ensure(c, d.ck, g2 * 4)) || (rc = ensure(c, d.cm, g2 * 4)) || (rc = ensure(c, d.ct, g2 * 8)) ||
      (rc = ensure(c, d.cL, g2 * 8)) || (rc = ensure(c, d.cmu, g2 * 8)) || (rc = ensure(c, d.cdt, g2 * 8)) ||
      (rc = ensure(c, d.I, (size_t)n_phase * c->n * 8)) || (rc = ensure(c, d.F, (size_t)n_phase * c->n * 8)) ||
      (rc = ensure(c, d.GF, (size_t)n_phase * c->n * 8)) || (rc = ensure(c, d.GI, (size_t)n_phase * c->n * 8)) ||
      (rc = ensure(c, d.res_part, (size_t)n_phase * diff_conv_blocks(c->n) * 8)) ||
      (rc = ensure(c, d.ccf, (size_t)n_phase * c->n_v * 8)) || (rc = ensure(c, d.par, (size_t)n_phase * 5 * 8)) ||
      (rc = ensure(c, d.model, (size_t)n_phase * c->n_v * 8)) || (rc = ensure(c, d.depth, (size_t)n_phase * 91 * 8)) ||
      (rc = ensure(c, d.bis, (size_t)n_phase * 91 * 8)) || (rc = ensure(c, d.iters, (size_t)n_phase * 4)) ||
      (rc = ensure(c, d.th, (size_t)n_phase * 4 * 8)) || (rc = ensure(c, d.dth, (size_t)n_phase * 4 * 8)) ||
      (rc = ensure(c, d.gy, (size_t)n_phase * c->n_v * 8)) ||
      (rc = ensure(c, d.bisd, (size_t)n_phase * (c->n_v + kBisHead) * 8)) ||
      (rc = ensure(c, d.tg, ((size_t)n_phase * 3 + 1 + kDiffScalars) * 8)) ||
      (rc = ensure(c, d.cell_part, (size_t)diff_cell_blocks((int64_t)g2) * kDiffScalars * 8)) ||
      (rc = ensure(c, c->ccf_part, (size_t)ccf_chunks(c->n_lines) * n_phase * c->n_v * 8)))
    return rc;
  int wid = conv_window(c->wave, sigma_ms);
  if (wid < 1 || 2 * (int64_t)wid + 2 >= c->n) return fail(c, SOAP_ERR_ARG, "diff_setup: window %d vs %lld pixels", wid, (long long)c->n);
  if ((rc = log_conv_taps(c, sigma_ms, wid))) return rc;
  d.G = grid;
  d.P = n_phase;
  d.M = map_res;
  d.sigma = sigma_ms;
  for (int q = 0; q < 4; q++) d.lb[q] = lb ? lb[q] : 0.0;
  d.ready = true;
  d.have_fwd = false;
  return SOAP_OK;
}

// max over mu in [0, 1] of |1 + LB(mu)| (sampled, 1 % margin)
static double lb_factor_bound(const double *lb) {
  double w = 0;
  for (int i = 0; i <= 4096; i++) {
    double mu = i / 4096.0;
    w = std::max(w, std::fabs(1 + (lb[0] + lb[1] * mu + lb[2] * mu * mu + lb[3] * mu * mu * mu)));
  }
  return 1.01 * w;
}

static DiffArgs diff_args(soap_ctx *c) {
  auto &d = c->diff;
  DiffArgs a{};
  a.n = c->n;
  a.grid = d.G;
  a.n_mu = c->n_mu;
  a.n_phase = d.P;
  a.map_res = d.M;
  a.dmin = d.dmin;
  a.nd = d.nd;
  a.h = c->h;
  a.expm1_h = c->expm1_h;
  a.vsini = d.prm.vsini_kms;
  a.u1 = d.prm.u1;
  a.u2 = d.prm.u2;
  a.kappa = d.prm.contrast;
  a.kappa_f = d.prm.contrast_fac;
  a.ns = d.ns;
  for (int q = 0; q < 4; q++) a.lb[q] = d.lb[q];
  a.mu_bounds = c->d_mub;
  a.ck = (int32_t *)d.ck.p;
  a.cm = (int32_t *)d.cm.p;
  a.ct = (double *)d.ct.p;
  a.cL = (double *)d.cL.p;
  a.cmu = (double *)d.cmu.p;
  a.cdt = (double *)d.cdt.p;
  a.maps = d.maps;
  a.fmaps = d.fmaps;
  a.acc = (unsigned long long *)d.acc.p;
  a.taps = (double *)d.taps.p;
  a.krange = (int32_t *)d.krange.p;
  a.q = d.q;
  a.s = d.s;
  a.f = d.f;
  a.I = (double *)d.I.p;
  a.F = (double *)d.F.p;
  return a;
}

static ConvArgs diff_conv(soap_ctx *c) {
  ConvArgs ca{};
  ca.n = c->n;
  ca.wid = c->res_wid;
  ca.dlo = c->res_dlo;
  ca.ntp = c->res_ntp;
  ca.ctap = (const double *)c->ctap.p;
  return ca;
}

int soap_diff_forward(soap_ctx *c, const soap_diff_params *prm, const double *d_q, const double *d_s,
                      const double *d_f, const double *d_maps, const double *d_fmaps, const double *d_target,
                      const double *rv_t, const double *fw_t, const double *bis_t, double w_spec, double w_rv,
                      double w_fw, double w_bis, double *loss, double *rv, double *fw, double *bis) {
  if (!c || !prm || !d_q || !d_s || !d_maps || !d_target || !rv_t || !fw_t || !bis_t || !loss)
    return fail(c, SOAP_ERR_ARG, "diff_forward: null argument");
  if (!d_f != !d_fmaps) return fail(c, SOAP_ERR_ARG, "diff_forward: facula seeds and facula maps go together");
  auto &d = c->diff;
  if (!d.ready) return fail(c, SOAP_ERR_STATE, "diff_forward: run diff_setup first");
  if (!(prm->vsini_kms > 0 && prm->vsini_kms < 2000))
    return fail(c, SOAP_ERR_ARG, "diff_forward: vsini %g km/s out of range", prm->vsini_kms);
  {
    // cell weights L (1 - phi - psi), contrast L phi, contrast_fac L (1 + LB) psi, maps in [0, 1]
    const double zero4[4] = {0, 0, 0, 0};
    double lmax = limb_weight_bound(prm->u1, prm->u2, zero4);
    double w = lmax * std::max(3.0, std::max(std::fabs(prm->contrast),
                                             d_f ? std::fabs(prm->contrast_fac) * lb_factor_bound(d.lb) : 0.0));
    if (!fx_capacity_ok((double)d.G * d.G, w))
      return fail(c, SOAP_ERR_LIMIT, "diff_forward: %d^2 cells with weights up to %g exceed the fixed-point tap "
                  "capacity (cells x weight < 2^30)", d.G, w);
  }
  HIPCHK(c, hipSetDevice(c->device));
  // taps span the shifts of delta in [-vsini, vsini] (m monotone in delta), plus margins
  auto mshift = [&](double v) {
    double beta = v * 1000. / kC;
    double fb = -1 * (1 - sqrt((1 + beta) / (1 - beta)));
    return (int)ceil(log1p(fb) / c->h);
  };
  int mlo = mshift(-prm->vsini_kms), mhi = mshift(prm->vsini_kms);
  d.dmin = mlo - 2;
  d.nd = mhi - d.dmin + 3;
  if (d.nd > 2048) return fail(c, SOAP_ERR_LIMIT, "diff_forward: shift range of %d pixels", d.nd);
  const int ns = d_f ? 3 : 2;
  int rc;
  if ((rc = ensure(c, d.acc, (size_t)d.P * c->n_mu * d.nd * ns * 2 * 8)) ||
      (rc = ensure(c, d.taps, (size_t)d.P * c->n_mu * d.nd * ns * 8)) ||
      (rc = ensure(c, d.krange, (size_t)d.P * c->n_mu * 2 * 4)) ||
      (rc = ensure(c, d.C, (size_t)d.P * c->n_mu * d.nd * ns * 8)) ||
      (rc = ensure(c, d.Cpart, (size_t)d.P * c->n_mu * diff_corr_chunks(c->n) * d.nd * ns * 8)))
    return rc;
  if (!c->mask_built && (rc = build_mask(c))) return rc;
  d.prm = *prm;
  d.ns = ns;
  d.q = d_q;
  d.s = d_s;
  d.f = d_f;
  d.maps = d_maps;
  d.fmaps = d_fmaps;
  d.target = d_target;
  d.w_spec = w_spec;
  double *tg = (double *)d.tg.p;  // [rv_t | fw_t | bis_t | loss]
  HIPCHK(c, hipMemcpyAsync(tg, rv_t, (size_t)d.P * 8, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(tg + d.P, fw_t, (size_t)d.P * 8, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(tg + 2 * d.P, bis_t, (size_t)d.P * 8, hipMemcpyHostToDevice, c->stream));
  DiffArgs a = diff_args(c);
  ConvArgs ca = diff_conv(c);
  {
    StageTimer tm(c, ST_FIR_ACTIVE);
    hipError_t e = launch_diff_forward_spectra(a, ca, d_target, c->trim, (double *)d.res_part.p, &d.res_blocks,
                                               c->stream);
    if (e == hipErrorInvalidValue) return fail(c, SOAP_ERR_LIMIT, "diff_forward: window or shift range too wide");
    HIPCHK(c, e);
  }
  {
    StageTimer tm(c, ST_CCF);
    HIPCHK(c, ctx_ccf(c, (const double *)d.F.p, d.P, (double *)c->ccf_part.p, c->stream));
    HIPCHK(c, launch_ccf_sum((const double *)c->ccf_part.p, ccf_chunks(c->n_lines), d.P, c->n_v, (double *)d.ccf.p,
                             c->stream));
  }
  {
    StageTimer tm(c, ST_FIT);
    HIPCHK(c, launch_fit(c->d_vgrid, c->n_v, (const double *)d.ccf.p, nullptr, d.P, (double *)d.par.p, (double *)d.model.p,
                         (double *)d.depth.p, (double *)d.bis.p, (int32_t *)d.iters.p, c->stream));
    HIPCHK(c, launch_fit_polish(c->d_vgrid, c->n_v, (const double *)d.ccf.p, (const double *)d.par.p, d.P,
                                (double *)d.th.p, c->stream));
    HIPCHK(c, launch_diff_bis(c->d_vgrid, c->n_v, (const double *)d.ccf.p, (const double *)d.th.p, d.P,
                              (double *)d.bisd.p, c->stream));
    // the loss and the fit-parameter derivatives for the backward stay on the device
    HIPCHK(c, launch_diff_loss(d.P, d.res_blocks, (const double *)d.res_part.p, (const double *)d.th.p, tg, tg + d.P,
                               tg + 2 * d.P, w_spec, w_rv, w_fw, w_bis, (double *)d.bisd.p, c->n_v, nullptr, 0.0,
                               0.0, 0.0, tg + 3 * d.P, (double *)d.dth.p, c->stream));
  }
  std::vector<double> th((size_t)d.P * 4), bd((size_t)d.P * (c->n_v + kBisHead));
  HIPCHK(c, hipMemcpyAsync(loss, tg + 3 * d.P, 8, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemcpyAsync(th.data(), d.th.p, th.size() * 8, hipMemcpyDeviceToHost, c->stream));
  if (bis) HIPCHK(c, hipMemcpyAsync(bd.data(), d.bisd.p, bd.size() * 8, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  for (int p = 0; p < d.P; p++) {
    if (rv) rv[p] = th[p * 4 + 2];
    if (fw) fw[p] = th[p * 4 + 3];
    if (bis) bis[p] = bd[(size_t)p * (c->n_v + kBisHead)];
  }
  d.have_fwd = true;
  return SOAP_OK;
}

int soap_diff_backward(soap_ctx *c, double *grad5, double *d_gq, double *d_gs, double *d_gf, double *d_gmaps,
                       double *d_gfmaps) {
  if (!c || !grad5 || !d_gq || !d_gs || !d_gmaps) return fail(c, SOAP_ERR_ARG, "diff_backward: null argument");
  auto &d = c->diff;
  if (!d.have_fwd) return fail(c, SOAP_ERR_STATE, "diff_backward: run diff_forward first");
  if (d.ns == 3 && (!d_gf || !d_gfmaps))
    return fail(c, SOAP_ERR_ARG, "diff_backward: the forward had faculae: facula gradients need buffers");
  HIPCHK(c, hipSetDevice(c->device));
  int rc0 = diff_mask_transpose(c);
  if (rc0) return rc0;
  DiffArgs a = diff_args(c);
  ConvArgs ca = diff_conv(c);
  {
    StageTimer tm(c, ST_CONV);
    hipError_t e = launch_diff_backward(
        a, ca, d.target, c->trim, d.w_spec, c->d_vgrid, c->n_v, (const double *)d.ccf.p, (const double *)d.th.p,
        (const double *)d.dth.p, (double *)d.gy.p, c->n - 2 * (int64_t)c->trim, (const int64_t *)d.mt_ptr.p,
        (const int32_t *)d.mt_w.p, (const int32_t *)d.mt_col.p, (const double *)d.mt_val.p, (double *)d.GF.p,
        (double *)d.GI.p, (double *)d.Cpart.p, (double *)d.C.p, d_gmaps, d.ns == 3 ? d_gfmaps : nullptr,
        (double *)d.cell_part.p, d_gq, d_gs, d.ns == 3 ? d_gf : nullptr, (const double *)d.bisd.p, c->stream);
    if (e == hipErrorInvalidValue) return fail(c, SOAP_ERR_LIMIT, "diff_backward: window or shift range too wide");
    HIPCHK(c, e);
  }
  // the scalar gradients: the same device reduction as the graph path; the sums
  // land in the tail of the loss slot buffer
  double *g5 = (double *)d.tg.p + 3 * d.P + 1;
  HIPCHK(c, launch_diff_grad(diff_cell_blocks((int64_t)d.G * d.G), (const double *)d.cell_part.p, g5, c->stream));
  HIPCHK(c, hipMemcpyAsync(grad5, g5, kDiffScalars * 8, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return SOAP_OK;
}

int soap_diff_get(soap_ctx *c, int32_t what, double *out) {
  if (!c || !out || what < 0 || what > 3) return fail(c, SOAP_ERR_ARG, "diff_get: bad arguments");
  auto &d = c->diff;
  if (!d.have_fwd) return fail(c, SOAP_ERR_STATE, "diff_get: run diff_forward first");
  HIPCHK(c, hipSetDevice(c->device));
  if (what == 3) {  // the span of each phase: the head of its bisd record
    HIPCHK(c, hipMemcpy2DAsync(out, 8, d.bisd.p, (size_t)(c->n_v + kBisHead) * 8, 8, d.P, hipMemcpyDeviceToHost,
                               c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return SOAP_OK;
  }
  const void *src = what == 0 ? d.F.p : what == 1 ? d.ccf.p : d.th.p;
  size_t bytes = (size_t)d.P * (what == 0 ? c->n : what == 1 ? c->n_v : 4) * 8;
  HIPCHK(c, hipMemcpyAsync(out, src, bytes, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return SOAP_OK;
}

int soap_run_series(soap_ctx *c, const soap_star *star, double sigma_ms, double quiet_weight) {
  int rc;
  if ((rc = soap_disk_init(c, star))) return rc;
  // the quiet star (its own tap scratch) runs on a second stream beside the active
  // phases; the resolution waits for it.  With stage timing on, stages run in order.
  const bool fork = !c->timing && c->kind == SOAP_GRID_LOG;
  if (fork) {
    HIPCHK(c, hipSetDevice(c->device));
    if (!c->aux) {
      HIPCHK(c, hipStreamCreateWithFlags(&c->aux, hipStreamNonBlocking));
      HIPCHK(c, hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming));
      HIPCHK(c, hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming));
    }
    HIPCHK(c, hipEventRecord(c->ev_fork, c->stream));
    HIPCHK(c, hipStreamWaitEvent(c->aux, c->ev_fork, 0));
    hipStream_t main_stream = c->stream;
    c->stream = c->aux;
    rc = soap_quiet_star(c);
    if (!rc) rc = resolution_quiet(c, sigma_ms);  // its convolution too, off the active path
    c->stream = main_stream;
    hipError_t e = hipEventRecord(c->ev_join, c->aux);  // recorded even after a failure, so the wait below ends
    if (rc) {
      hipStreamWaitEvent(c->stream, c->ev_join, 0);
      return rc;
    }
    HIPCHK(c, e);
    rc = soap_integrate(c);
    hipError_t ew = hipStreamWaitEvent(c->stream, c->ev_join, 0);
    if (rc) return rc;
    HIPCHK(c, ew);
  } else {
    if ((rc = soap_quiet_star(c))) return rc;
    if ((rc = soap_integrate(c))) return rc;
  }
  // spec_norm shares its launch with the CCF (stage timing then reports the pair
  // under the CCF stage, as it runs)
  const bool fuse = c->have_mask;
  if ((rc = resolution_impl(c, sigma_ms, quiet_weight, !fuse, !fork))) return rc;
  if (c->have_mask) {
    if ((rc = ccf_impl(c, fuse))) return rc;
    if ((rc = soap_fit(c))) return rc;
  }
  return SOAP_OK;
}

// ------------------------------------------------------------------ a11: batched instances
}  // extern "C"

namespace {
const char *kBatchStageNames[BS_COUNT] = {"fill", "taps", "finalize", "fir_prog", "fir", "conv_quiet", "conv",
                                          "minmax", "norm", "ccf",      "ccf_reduce", "fit"};

struct BatchStageBuf {
  void *jobs = nullptr;       // device job array [n_inst]
  int64_t *prefix = nullptr;  // device [n_inst + 1]
  size_t job_bytes = 0;
  int64_t total = 0;
  int threads = 64;
  size_t lds = 0;
  int64_t grid = 0;  // persistent block count; 0 = one block per work item
};
}  // namespace

struct soap_batch {
  int device = 0;
  int n_cu = 256;
  hipStream_t stream = nullptr;
  std::string err;
  std::vector<soap_ctx *> inst;
  BatchStageBuf st[BS_COUNT];
  int taps_cells = 4096;  // disk cells per taps block (cells-per-block shape)
  int fir_g = 4;          // phases per FIR block
  int fir_b = -7;         // pixels per FIR thread (0: one-pixel form; -7: the tile-stationary DPP form)
  int fir_pre = 0;        // one-pixel form: prefetch the next bin's seed rows into registers
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipEvent_t ev[BS_COUNT + 1] = {};
  double stage_ms[BS_COUNT] = {};
  std::vector<std::vector<uintptr_t>> sig;  // per instance: buffers and sizes the job tables point at
};

namespace {
int bfail(soap_batch *b, int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  if (b) b->err = buf;
  return code;
}
#define BHIP(b, x)                                                                                     \
  do {                                                                                                 \
    hipError_t e_ = (x);                                                                               \
    if (e_ != hipSuccess) return bfail(b, SOAP_ERR_HIP, "%s: %s", #x, hipGetErrorString(e_));          \
  } while (0)

void batch_drop_graph(soap_batch *b) {
  if (b->exec) hipGraphExecDestroy(b->exec);
  if (b->graph) hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
}

template <class Job>
int batch_upload(soap_batch *b, int stage, const std::vector<Job> &jobs, const std::vector<int64_t> &work, int threads,
                 size_t lds) {
  BatchStageBuf &sb = b->st[stage];
  size_t nb = jobs.size() * sizeof(Job);
  if (!sb.jobs || sb.job_bytes < nb) {
    if (sb.jobs) hipFree(sb.jobs);
    sb.jobs = nullptr;
    BHIP(b, hipMalloc(&sb.jobs, nb));
    sb.job_bytes = nb;
  }
  std::vector<int64_t> pre(jobs.size() + 1, 0);
  for (size_t i = 0; i < jobs.size(); i++) pre[i + 1] = pre[i] + work[i];
  if (!sb.prefix) BHIP(b, hipMalloc(&sb.prefix, pre.size() * sizeof(int64_t)));
  BHIP(b, hipMemcpy(sb.jobs, jobs.data(), nb, hipMemcpyHostToDevice));
  BHIP(b, hipMemcpy(sb.prefix, pre.data(), pre.size() * sizeof(int64_t), hipMemcpyHostToDevice));
  sb.total = pre.back();
  sb.threads = threads;
  sb.lds = lds;
  return SOAP_OK;
}

// Everything of an instance that the job tables capture: a context changed
// after the batch was built (reallocated buffers, other sizes or taps) must not
// be stepped through stale pointers.
std::vector<uintptr_t> batch_signature(const soap_ctx *c) {
  auto u = [](const void *p) { return (uintptr_t)p; };
  return {u(c->d_raw),    u(c->d_lower), u(c->d_lraw),   u(c->wsum.p),   u(c->fp.p),     (uintptr_t)c->ccf_cap,
          u(c->d_quiet),   u(c->acc.p),    u(c->taps.p),   u(c->taps_g.p), u(c->krange.p),
          u(c->kflags.p), u(c->bins.p),  u(c->nbins.p),   u(c->part.p),   u(c->mm.p),     u(c->ctap.p),
          u(c->d_ccf),    u(c->ccf_part.p), u(c->f_par.p), u(c->f_model.p), u(c->f_depth.p), u(c->f_bis.p),
          u(c->f_iter.p), u(c->m_start.p), u(c->m_alive.p), u(c->m_val.p), u(c->d_geo),   u(c->d_box),
          u(c->d_region), u(c->d_m),     u(c->d_t),       u(c->prog_t.p), u(c->prog_x.p), u(c->prog_n.p),
          u(c->prog_u.p), (uintptr_t)c->P, (uintptr_t)c->n, (uintptr_t)c->G,
          (uintptr_t)c->nd, (uintptr_t)c->res_ntp, (uintptr_t)c->res_wid, (uintptr_t)c->n_lines,
          (uintptr_t)c->n_v, (uintptr_t)c->geo.size(), (uintptr_t)c->act.solar_mode};
}

int batch_check(soap_batch *b) {
  for (size_t i = 0; i < b->inst.size(); i++)
    if (batch_signature(b->inst[i]) != b->sig[i])
      return bfail(b, SOAP_ERR_STATE, "batch: instance %zu was changed after the batch was created", i);
  return SOAP_OK;
}

// Job tables of every stage from the prepared contexts and the current shapes.
int batch_build(soap_batch *b) {
  size_t ni = b->inst.size();
  std::vector<FillJob> fill(ni);
  std::vector<TapsJob> taps(ni);
  std::vector<FinalizeJob> fin(ni);
  std::vector<FirJob> fir(ni);
  std::vector<FirProgJob> fprog(ni);
  std::vector<FirTileJob> ftile(ni);
  const bool tile = b->fir_b == -7;
  std::vector<ConvQJob> cq(ni);
  std::vector<ConvAJob> ca(ni);
  std::vector<MinmaxJob> mm(ni);
  std::vector<NormJob> nm(ni);
  std::vector<CcfJob> cc(ni);
  std::vector<CcfRedJob> cr(ni);
  std::vector<FitJob> ft(ni);
  std::vector<int64_t> w[BS_COUNT];
  for (auto &v : w) v.assign(ni, 0);
  size_t lds_taps = 0, lds_fir = 0, lds_conv = 0, lds_ccf = 0;
  int ccf_threads = 64, red_threads = 64;
  for (size_t i = 0; i < ni; i++) {
    soap_ctx *c = b->inst[i];
    const int P = c->P, nsets = 3, n_spec = 3 * (P + 1);
    // fill + taps + finalize (soap_integrate, non-solar log grid)
    TapAccArgs a{};
    a.grid = c->G;
    a.n_mu = c->n_mu;
    a.dmin = c->dmin;
    a.nd = c->nd;
    a.nsets = nsets;
    a.wd = tap_window(c, nsets);
    a.cols_per_block = std::max(1, b->taps_cells / std::max(1, c->G));
    a.region = c->d_region;
    a.limb = c->d_limb;
    a.delta = c->d_delta;
    a.lbm = c->d_lbm;
    a.m = c->d_m;
    a.t = c->d_t;
    a.acc = (unsigned long long *)c->acc.p;
    a.geo = c->d_geo;
    a.geo_off = c->d_geo_off;
    a.box = c->d_box;
    int mcb = 0;
    for (int p = 0; p < P; p++) mcb = std::max(mcb, (c->box[p * 4 + 1] - c->box[p * 4] + a.cols_per_block - 1) / a.cols_per_block);
    fill[i] = FillJob{a.acc, (int64_t)P * nsets * c->n_mu * c->nd * 2};
    w[BS_FILL][i] = (fill[i].count + batch_fill_chunk() - 1) / batch_fill_chunk();
    taps[i] = TapsJob{a, std::max(1, mcb)};
    w[BS_TAPS][i] = (int64_t)mcb * P;
    lds_taps = std::max(lds_taps, taps_lds_bytes(a));
    fin[i] = FinalizeJob{a.acc, nsets, nsets + 1, c->n_mu, c->nd, (double *)c->taps.p, (int32_t *)c->krange.p,
                         (int32_t *)c->kflags.p, (int32_t *)c->bins.p, (int32_t *)c->nbins.p,
                         (double *)c->taps_g.p, P, taps_g_pad(P)};
    w[BS_FINALIZE][i] = taps_g_pad(P);
    FirArgs f{};
    f.q = c->d_cube[SOAP_CUBE_QUIET];
    f.qs = c->d_cube[SOAP_CUBE_QUIET_SED];
    f.sp = c->d_cube[SOAP_CUBE_SPOT];
    f.sps = c->d_cube[SOAP_CUBE_SPOT_SED];
    f.fa = c->d_cube[SOAP_CUBE_FAC];
    f.fas = c->d_cube[SOAP_CUBE_FAC_SED];
    f.n = c->n;
    f.n_mu = c->n_mu;
    f.dmin = c->dmin;
    f.nd = c->nd;
    f.taps = (const double *)c->taps.p;
    f.taps_g = (const double *)c->taps_g.p;
    f.p_pad = taps_g_pad(P);
    f.krange = (const int32_t *)c->krange.p;
    f.kflags = (const int32_t *)c->kflags.p;
    f.bins = (const int32_t *)c->bins.p;
    f.nbins = (const int32_t *)c->nbins.p;
    f.out = c->d_raw;
    if (tile) {  // the tile-stationary DPP form: entry streams per phase group, then 64-pixel tiles
      const int ng = fir_tile_groups(P), cap = c->n_mu * c->nd + 6;
      int rc;
      if ((rc = ensure(c, c->prog_t, (size_t)ng * cap * 16 * 8)) || (rc = ensure(c, c->prog_x, (size_t)ng * cap * 4)) ||
          (rc = ensure(c, c->prog_n, (size_t)ng * 4)) || (rc = ensure(c, c->prog_u, (size_t)c->n_mu * 4)))
        return bfail(b, rc, "batch: instance %zu: %s", i, c->err.c_str());
      const FirProg pg{(double *)c->prog_t.p, (int32_t *)c->prog_x.p, (int32_t *)c->prog_n.p, (int32_t *)c->prog_u.p,
                       cap};
      // wider tiles for few phase groups only while two blocks still fit per CU:
      // the launch's LDS is the largest instance's for every block
      const int lg = fir_tile_t64_log2(P, c->n_mu, c->nd, 81920), tp = 64 << lg;
      const int n_tiles = (int)((c->n + tp - 1) / tp), tpx = (n_tiles + 7) / 8;
      fprog[i] = FirProgJob{f, pg, P, ng, tp + c->nd - 1};
      w[BS_FIR_PROG][i] = ng + 1;
      ftile[i] = FirTileJob{f, pg, P, ng, n_tiles, tpx, lg};
      w[BS_FIR][i] = 8L * tpx;
      lds_fir = std::max(lds_fir, fir_tile_lds(c->n_mu, c->nd, 1 << lg));
    } else {
      int fir_tiles = b->fir_b > 0 ? fir2_tiles(c->n, b->fir_b) : (int)((c->n + fir_tile() - 1) / fir_tile());
      int groups = (P + b->fir_g - 1) / b->fir_g;
      fir[i] = FirJob{f, P, fir_tiles, groups};
      w[BS_FIR][i] = (int64_t)(fir_tiles + 7) / 8 * 8 * groups;
      lds_fir = std::max(lds_fir, b->fir_b > 0 ? fir2_lds(c->nd, b->fir_g, b->fir_b) : fir_lds(c->nd));
    }
    // resolution (log grid)
    ConvArgs cv{};
    cv.n = c->n;
    cv.wid = c->res_wid;
    cv.dlo = c->res_dlo;
    cv.ntp = c->res_ntp;
    cv.ctap = (const double *)c->ctap.p;
    cv.n_tiles = (int)((c->n + conv_tile() - 1) / conv_tile());
    double *part = (double *)c->part.p;
    cq[i] = ConvQJob{cv, c->d_quiet, c->d_lraw + (size_t)3 * P * c->n, part + (size_t)3 * P * cv.n_tiles * 2};
    w[BS_CONV_QUIET][i] = 2L * cv.n_tiles;
    // one block per (tile, phase, type); XCD groups of 8 tiles only for spectra of
    // at least 8 tiles (below that, the padding blocks would outnumber the tiles)
    const int xcd = cv.n_tiles >= 8;
    ca[i] = ConvAJob{cv, c->d_quiet, c->d_raw, c->d_lraw, part, P, xcd};
    w[BS_CONV][i] = xcd ? (int64_t)(cv.n_tiles + 7) / 8 * 8 * P * 3 : ((int64_t)cv.n_tiles * P * 3 + 7) / 8 * 8;
    lds_conv = std::max(lds_conv, conv_lds_bytes(cv.ntp));
    mm[i] = MinmaxJob{part, cv.n_tiles, (double *)c->mm.p};
    w[BS_MINMAX][i] = n_spec;
    int64_t nbx = (c->n + norm_px_per_block() - 1) / norm_px_per_block();
    nm[i] = NormJob{c->n, nbx, (const double *)c->mm.p, c->d_lraw, c->d_lower};
    w[BS_NORM][i] = nbx * n_spec;
    // CCF + fit
    int nc = ccf_chunks(c->n_lines);
    cc[i] = CcfJob{c->d_lraw, c->n, c->trim, c->n - 2 * (int64_t)c->trim, n_spec, c->n_lines, c->n_v, nc,
                   (const int32_t *)c->m_start.p, (const uint32_t *)c->m_alive.p, (const double2 *)c->m_val.p,
                   c->d_w_sorted, (const int2 *)c->fp.p, c->ccf_cap, (double *)c->ccf_part.p};
    lds_ccf = std::max(lds_ccf, ccf_lds_bytes(c->ccf_cap));
    w[BS_CCF][i] = (int64_t)nc * ((n_spec + ccf_spec_chunk() - 1) / ccf_spec_chunk());
    ccf_threads = std::max(ccf_threads, (c->n_v + 63) / 64 * 64);
    int rt = 64;
    while (rt < c->n_v) rt <<= 1;
    red_threads = std::max(red_threads, rt);
    cr[i] = CcfRedJob{(const double *)c->ccf_part.p, nc,        n_spec, c->n_v, c->d_appod, (const double *)c->mm.p,
                      (const double *)c->wsum.p,    c->d_ccf};
    w[BS_CCF_REDUCE][i] = n_spec;
    ft[i] = FitJob{c->d_vgrid, c->n_v, c->d_ccf, (double *)c->f_par.p, (double *)c->f_model.p,
                   (double *)c->f_depth.p, (double *)c->f_bis.p, (int32_t *)c->f_iter.p};
    w[BS_FIT][i] = n_spec;
  }
  for (auto &v : w)
    for (int64_t x : v)
      if (x > 0x7fffffffL) return bfail(b, SOAP_ERR_LIMIT, "batch: an instance has more than 2^31 blocks in a stage");
  if (lds_fir > (tile ? 163840u : 65536u) || lds_conv > 65536 || lds_taps > 65536 || ccf_threads > 1024 ||
      red_threads > 1024)
    return bfail(b, SOAP_ERR_LIMIT, "batch: an instance exceeds a per-block limit (lds fir %zu conv %zu taps %zu)",
                 lds_fir, lds_conv, lds_taps);
  int rc;
  if ((rc = batch_upload(b, BS_FILL, fill, w[BS_FILL], 256, 0)) ||
      (rc = batch_upload(b, BS_TAPS, taps, w[BS_TAPS], 256, lds_taps)) ||
      (rc = batch_upload(b, BS_FINALIZE, fin, w[BS_FINALIZE], 256, 0)) ||
      (rc = batch_upload(b, BS_FIR_PROG, fprog, w[BS_FIR_PROG], 1024, 0)) ||
      (rc = tile ? batch_upload(b, BS_FIR, ftile, w[BS_FIR], 512, lds_fir)
                 : batch_upload(b, BS_FIR, fir, w[BS_FIR], fir_tile(), lds_fir)) ||
      (rc = batch_upload(b, BS_CONV_QUIET, cq, w[BS_CONV_QUIET], 256, lds_conv)) ||
      (rc = batch_upload(b, BS_CONV, ca, w[BS_CONV], 256, lds_conv)) ||
      (rc = batch_upload(b, BS_MINMAX, mm, w[BS_MINMAX], 256, 0)) ||
      (rc = batch_upload(b, BS_NORM, nm, w[BS_NORM], 256, 0)) ||
      (rc = batch_upload(b, BS_CCF, cc, w[BS_CCF], ccf_threads, lds_ccf)) ||
      (rc = batch_upload(b, BS_CCF_REDUCE, cr, w[BS_CCF_REDUCE], red_threads, 0)) ||
      (rc = batch_upload(b, BS_FIT, ft, w[BS_FIT], 64, 0)))
    return rc;
  b->sig.resize(ni);
  for (size_t i = 0; i < ni; i++) b->sig[i] = batch_signature(b->inst[i]);
  batch_drop_graph(b);
  return SOAP_OK;
}

// The step's launches in stage order; with events, each stage is bracketed.
int batch_enqueue(soap_batch *b, bool events) {
  int rc = batch_check(b);
  if (rc) return rc;
  for (int s = 0; s < BS_COUNT; s++) {
    BatchStageBuf &sb = b->st[s];
    if (events) BHIP(b, hipEventRecord(b->ev[s], b->stream));
    hipError_t e = launch_batch(s, sb.jobs, sb.prefix, (int)b->inst.size(), sb.total, sb.grid, sb.threads, sb.lds,
                                b->fir_g, b->fir_b, b->fir_pre, b->stream);
    if (e != hipSuccess) return bfail(b, SOAP_ERR_HIP, "batch stage %s: %s", kBatchStageNames[s], hipGetErrorString(e));
  }
  if (events) BHIP(b, hipEventRecord(b->ev[BS_COUNT], b->stream));
  return SOAP_OK;
}

int batch_timed(soap_batch *b, double *ms) {
  int rc = batch_enqueue(b, true);
  if (rc) return rc;
  BHIP(b, hipEventSynchronize(b->ev[BS_COUNT]));
  for (int s = 0; s < BS_COUNT; s++) {
    float t = 0;
    BHIP(b, hipEventElapsedTime(&t, b->ev[s], b->ev[s + 1]));
    ms[s] = t;
  }
  return SOAP_OK;
}
}  // namespace

extern "C" {

int soap_batch_create(soap_ctx *const *ctxs, int32_t n_inst, const double *sigma_ms, const double *quiet_weight,
                      soap_batch **out) {
  if (!out) return SOAP_ERR_ARG;
  *out = nullptr;
  if (!ctxs || n_inst < 1 || !sigma_ms || !quiet_weight) return SOAP_ERR_ARG;
  auto *b = new soap_batch();
  b->device = ctxs[0] ? ctxs[0]->device : 0;
  int rc = SOAP_OK;
  auto bad = [&](int code, const char *msg, int i) {
    rc = bfail(b, code, "batch_create: instance %d: %s", i, msg);
  };
  for (int i = 0; i < n_inst && !rc; i++) {
    soap_ctx *c = ctxs[i];
    if (!c) bad(SOAP_ERR_ARG, "null context", i);
    else if (c->device != b->device) bad(SOAP_ERR_ARG, "contexts must share one device", i);
    else if (c->kind != SOAP_GRID_LOG) bad(SOAP_ERR_GRID, "batched instances need log-uniform grids", i);
    else if (c->act.solar_mode) bad(SOAP_ERR_LIMIT, "solar mode is not batched", i);
    else if (!c->have_quiet || !c->have_maps) bad(SOAP_ERR_STATE, "run disk_init and quiet_star first", i);
    else if (!c->have_mask) bad(SOAP_ERR_STATE, "set the mask first", i);
    else if (c->P < 1) bad(SOAP_ERR_STATE, "set an activity with at least one phase", i);
    else if (c->n_mu > 64) bad(SOAP_ERR_LIMIT, "n_mu > 64", i);
    if (rc) break;
    for (int j = 0; j < i; j++)
      if (ctxs[j] == c) bad(SOAP_ERR_ARG, "context listed twice", i);
    if (rc) break;
    // one plain run allocates the instance's buffers, builds its mask tables and
    // resolution taps, and leaves results the batched step must reproduce
    int r;
    if ((r = soap_integrate(c)) || (r = soap_resolution(c, sigma_ms[i], quiet_weight[i])) || (r = soap_ccf(c)) ||
        (r = soap_fit(c)) || (r = soap_sync(c)))
      rc = bfail(b, r, "batch_create: instance %d: %s", i, c->err.c_str());
    b->inst.push_back(c);
  }
  if (!rc) {
    hipError_t e = hipSetDevice(b->device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
    for (int s = 0; s <= BS_COUNT && e == hipSuccess; s++) e = hipEventCreate(&b->ev[s]);
    int ncu = 0;
    if (e == hipSuccess) e = hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, b->device);
    if (e != hipSuccess) rc = bfail(b, SOAP_ERR_HIP, "batch_create: %s", hipGetErrorString(e));
    b->n_cu = ncu > 0 ? ncu : 256;
  }
  if (!rc) rc = batch_build(b);
  if (rc) {
    // the error string lives in the batch; keep it for soap_batch_last_error
    *out = b;
    return rc;
  }
  *out = b;
  return SOAP_OK;
}

void soap_batch_destroy(soap_batch *b) {
  if (!b) return;
  hipSetDevice(b->device);
  if (b->stream) hipStreamSynchronize(b->stream);
  batch_drop_graph(b);
  for (auto &sb : b->st) {
    if (sb.jobs) hipFree(sb.jobs);
    if (sb.prefix) hipFree(sb.prefix);
  }
  for (auto &e : b->ev)
    if (e) hipEventDestroy(e);
  if (b->stream) hipStreamDestroy(b->stream);
  delete b;
}

const char *soap_batch_last_error(const soap_batch *b) { return b ? b->err.c_str() : "null batch"; }

int soap_batch_step(soap_batch *b, int32_t use_graph) {
  if (!b || b->inst.empty()) return SOAP_ERR_ARG;
  BHIP(b, hipSetDevice(b->device));
  if (!use_graph) return batch_enqueue(b, false);
  int rc0 = batch_check(b);
  if (rc0) return rc0;
  if (!b->exec) {
    BHIP(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    int rc = batch_enqueue(b, false);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(b->stream, &g);
    if (rc) {
      if (g) hipGraphDestroy(g);
      return rc;
    }
    BHIP(b, e);
    b->graph = g;
    BHIP(b, hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
  }
  BHIP(b, hipGraphLaunch(b->exec, b->stream));
  return SOAP_OK;
}

int soap_batch_sync(soap_batch *b) {
  if (!b) return SOAP_ERR_ARG;
  BHIP(b, hipSetDevice(b->device));
  BHIP(b, hipStreamSynchronize(b->stream));
  return SOAP_OK;
}

int soap_batch_stage_ms(soap_batch *b, double *ms, int32_t n_stage) {
  if (!b || !ms || n_stage < BS_COUNT) return SOAP_ERR_ARG;
  BHIP(b, hipSetDevice(b->device));
  return batch_timed(b, ms);
}

const char *soap_batch_stage_name(int32_t s) { return (s >= 0 && s < BS_COUNT) ? kBatchStageNames[s] : ""; }
int32_t soap_batch_num_stages(void) { return BS_COUNT; }

int soap_batch_get_config(soap_batch *b, int32_t *taps_cells, int32_t *fir_g, int32_t *fir_b, int32_t *fir_pre,
                          int64_t *grid, int64_t *work, int32_t n_stage) {
  if (!b || n_stage < BS_COUNT) return SOAP_ERR_ARG;
  if (taps_cells) *taps_cells = b->taps_cells;
  if (fir_g) *fir_g = b->fir_g;
  if (fir_b) *fir_b = b->fir_b;
  if (fir_pre) *fir_pre = b->fir_pre;
  for (int s = 0; s < BS_COUNT; s++) {
    if (grid) grid[s] = b->st[s].grid;
    if (work) work[s] = b->st[s].total;
  }
  return SOAP_OK;
}

static bool batch_fir_shape_ok(int g, int bb) {
  const int ok[][2] = {{2, 0}, {4, 0}, {6, 0}, {1, 3}, {1, 5}, {2, 3}, {4, -7}};
  for (auto &s : ok)
    if (s[0] == g && s[1] == bb) return true;
  return false;
}

int soap_batch_set_config(soap_batch *b, int32_t taps_cells, int32_t fir_g, int32_t fir_b, int32_t fir_pre,
                          const int64_t *grid, int32_t n_stage) {
  if (!b || n_stage < BS_COUNT || !grid) return SOAP_ERR_ARG;
  if (taps_cells < 64 || taps_cells > (1 << 20)) return bfail(b, SOAP_ERR_ARG, "set_config: taps_cells out of range");
  if (!batch_fir_shape_ok(fir_g, fir_b))
    return bfail(b, SOAP_ERR_ARG,
                 "set_config: FIR shape (%d, %d) not one of (2,0) (4,0) (6,0) (1,3) (1,5) (2,3) (4,-7)", fir_g, fir_b);
  if (fir_pre != 0 && (fir_pre != 1 || fir_b != 0))
    return bfail(b, SOAP_ERR_ARG, "set_config: fir_pre is 0 or 1, and 1 only with the one-pixel form (fir_b 0)");
  for (int s = 0; s < BS_COUNT; s++)
    if (grid[s] < 0 || grid[s] % 8) return bfail(b, SOAP_ERR_ARG, "set_config: grid sizes are multiples of 8 (0 = all)");
  BHIP(b, hipSetDevice(b->device));
  BHIP(b, hipStreamSynchronize(b->stream));
  b->taps_cells = taps_cells;
  b->fir_g = fir_g;
  b->fir_b = fir_b;
  b->fir_pre = fir_pre;
  for (int s = 0; s < BS_COUNT; s++) b->st[s].grid = grid[s];
  return batch_build(b);
}

// Coordinate search over the launch shapes: the taps block size in disk cells,
// the FIR phase group, then each stage's persistent grid (0 = one block per
// work item, else k * CUs).  Each candidate runs whole steps (so every stage
// sees its real inputs) and is scored by its stage's fastest of `reps` runs.
int soap_batch_tune(soap_batch *b, int32_t reps) {
  if (!b || reps < 1) return SOAP_ERR_ARG;
  BHIP(b, hipSetDevice(b->device));
  BHIP(b, hipStreamSynchronize(b->stream));
  double ms[BS_COUNT];
  auto score = [&](int s0, int s1, double *best) -> int {
    *best = HUGE_VAL;
    for (int r = 0; r < reps; r++) {
      int rc = batch_timed(b, ms);
      if (rc) return rc;
      double t = 0;
      for (int s = s0; s <= s1; s++) t += ms[s];
      *best = std::min(*best, t);
    }
    return SOAP_OK;
  };
  int rc;
  double t, best;
  // taps block shape (taps stage only; fill and finalize do not depend on it)
  const int cells_c[] = {1024, 2048, 4096, 8192, 16384};
  int best_cells = b->taps_cells;
  best = HUGE_VAL;
  for (int cells : cells_c) {
    b->taps_cells = cells;
    if ((rc = batch_build(b)) || (rc = score(BS_TAPS, BS_TAPS, &t))) return rc;
    if (t < best) best = t, best_cells = cells;
  }
  b->taps_cells = best_cells;
  // FIR shape: phases per block, pixels per thread (the lambda tile) and, for the
  // one-pixel form, the seed-row prefetch
  const int g_c[][3] = {{2, 0, 0}, {4, 0, 0}, {6, 0, 0}, {2, 0, 1}, {4, 0, 1},
                        {6, 0, 1}, {1, 3, 0}, {1, 5, 0}, {2, 3, 0}, {4, -7, 0}};
  int best_g = b->fir_g, best_b = b->fir_b, best_pre = b->fir_pre;
  best = HUGE_VAL;
  for (auto &gb : g_c) {
    b->fir_g = gb[0];
    b->fir_b = gb[1];
    b->fir_pre = gb[2];
    if ((rc = batch_build(b)) || (rc = score(BS_FIR_PROG, BS_FIR, &t))) return rc;
    if (t < best) best = t, best_g = gb[0], best_b = gb[1], best_pre = gb[2];
  }
  b->fir_g = best_g;
  b->fir_b = best_b;
  b->fir_pre = best_pre;
  if ((rc = batch_build(b))) return rc;
  const int64_t mult[] = {0, 1, 2, 4, 8, 16};
  for (int s = 0; s < BS_COUNT; s++) {
    int64_t best_grid = 0;
    best = HUGE_VAL;
    for (int64_t m : mult) {
      int64_t g = m * b->n_cu / 8 * 8;
      if (m > 0 && g >= b->st[s].total) continue;  // no smaller than one block per item
      b->st[s].grid = g;
      if ((rc = score(s, s, &t))) return rc;
      if (t < best * 0.98 || (m == 0 && t < best)) best = t, best_grid = g;
    }
    b->st[s].grid = best_grid;
  }
  batch_drop_graph(b);
  return SOAP_OK;
}

}  // extern "C"

// ------------------------------------------------------------------ a11: differentiable instances in one graph
// Forward + adjoint of many differentiable instances (each a soap_ctx after
// soap_diff_setup) enqueued on one stream with every input, parameter and result
// in device memory, so that the whole evaluation can be captured once as a HIP
// graph and replayed with new parameter values: the tap range of each instance
// is fixed by a bound on vsini, the loss, the fit-parameter derivatives and the
// scalar gradients are reduced on the device.
constexpr int kDiffLanes = 4;

struct soap_diffbatch {
  struct Item {
    soap_ctx *c = nullptr;
    bool bound = false;
    double vsini_max = 0;
    const double *prm = nullptr, *q = nullptr, *s = nullptr, *f = nullptr, *maps = nullptr, *fmaps = nullptr,
                 *target = nullptr, *rv_t = nullptr, *fw_t = nullptr, *bis_t = nullptr;
    double w_spec = 0, w_rv = 0, w_fw = 0, w_bis = 0;
    double *loss = nullptr, *g5 = nullptr, *gq = nullptr, *gs = nullptr, *gf = nullptr, *gmaps = nullptr,
           *gfmaps = nullptr;
    int dmin = 0, nd = 0;
    std::vector<uintptr_t> sig;  // buffers and sizes the enqueued launches use
  };
  int device = 0;
  hipStream_t stream = nullptr;
  // instances run on kDiffLanes streams forked from `stream` and joined back, so
  // that small instances overlap; the captured graph has one branch per lane
  hipStream_t lane[kDiffLanes] = {};
  hipEvent_t fork = nullptr, join[kDiffLanes] = {};
  std::vector<int> lane_of;  // lane of each instance, balanced by estimated work
  std::vector<Item> items;
  hipGraph_t graph[2] = {};  // [with_adjoint]
  hipGraphExec_t exec[2] = {};
  std::string err;
};

namespace {
std::vector<uintptr_t> dbatch_signature(const soap_ctx *c) {
  auto &d = c->diff;
  auto u = [](const void *p) { return (uintptr_t)p; };
  return {u(d.acc.p),  u(d.taps.p),  u(d.krange.p), u(d.C.p),     u(d.Cpart.p),  u(d.res_part.p), u(d.ccf.p),
          u(d.par.p),  u(d.model.p), u(d.depth.p),  u(d.bis.p),   u(d.iters.p),  u(d.th.p),       u(d.dth.p),
          u(d.gy.p),   u(d.GF.p),    u(d.GI.p),     u(d.cell_part.p), u(d.I.p),  u(d.F.p),        u(d.ck.p),
          u(d.cm.p),   u(d.ct.p),    u(d.cL.p),     u(d.cmu.p),   u(d.cdt.p),    u(d.mt_ptr.p),   u(d.mt_w.p),
          u(d.mt_col.p), u(d.mt_val.p), u(c->ccf_part.p), u(c->m_start.p), u(c->m_alive.p), u(c->m_val.p),
          u(c->d_w_sorted), u(c->ctap.p), u(d.bisd.p), u(d.maps), u(d.q), u(d.s), u(d.target), u(d.f), u(d.fmaps),
          (uintptr_t)d.ns, (uintptr_t)d.nd,
          (uintptr_t)d.dmin, (uintptr_t)d.P, (uintptr_t)d.G, (uintptr_t)d.M, (uintptr_t)c->n, (uintptr_t)c->res_ntp,
          (uintptr_t)c->n_lines, (uintptr_t)c->n_v};
}

void dbatch_drop_graph(soap_diffbatch *b) {
  for (int k = 0; k < 2; k++) {
    if (b->exec[k]) hipGraphExecDestroy(b->exec[k]);
    if (b->graph[k]) hipGraphDestroy(b->graph[k]);
    b->exec[k] = nullptr;
    b->graph[k] = nullptr;
  }
}

int dfail(soap_diffbatch *b, int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  if (b) b->err = buf;
  return code;
}

// Lanes by estimated work (spectra pixels and disk cells per phase), largest
// instance first onto the least-loaded lane.
void dbatch_assign_lanes(soap_diffbatch *b) {
  const size_t n = b->items.size();
  std::vector<double> work(n);
  std::vector<size_t> order(n);
  for (size_t i = 0; i < n; i++) {
    const soap_ctx *c = b->items[i].c;
    work[i] = (double)c->diff.P * ((double)c->n + (double)c->diff.G * c->diff.G);
    order[i] = i;
  }
  std::stable_sort(order.begin(), order.end(), [&](size_t x, size_t y) { return work[x] > work[y]; });
  double load[kDiffLanes] = {};
  b->lane_of.assign(n, 0);
  for (size_t i : order) {
    int best = 0;
    for (int k = 1; k < kDiffLanes; k++)
      if (load[k] < load[best]) best = k;
    b->lane_of[i] = best;
    load[best] += work[i];
  }
}

hipError_t dbatch_enqueue_one(soap_diffbatch::Item &it, hipStream_t st, bool adjoint);

int dbatch_enqueue(soap_diffbatch *b, bool adjoint) {
  for (size_t i = 0; i < b->items.size(); i++)
    if (!b->items[i].bound) return dfail(b, SOAP_ERR_STATE, "diff batch: instance %zu not bound", i);
  hipError_t e = hipEventRecord(b->fork, b->stream);
  for (int k = 0; k < kDiffLanes && e == hipSuccess; k++) e = hipStreamWaitEvent(b->lane[k], b->fork, 0);
  if (e != hipSuccess) return dfail(b, SOAP_ERR_HIP, "diff batch: fork: %s", hipGetErrorString(e));
  int status = SOAP_OK;
  for (size_t i = 0; i < b->items.size() && status == SOAP_OK; i++) {
    e = dbatch_enqueue_one(b->items[i], b->lane[b->lane_of[i]], adjoint);
    if (e != hipSuccess)
      status = dfail(b, e == hipErrorInvalidValue ? SOAP_ERR_LIMIT : SOAP_ERR_HIP, "diff batch: instance %zu: %s", i,
                     hipGetErrorString(e));
  }
  // join every lane back, also after a failed launch, so that a capture can end
  for (int k = 0; k < kDiffLanes; k++) {
    hipError_t ej = hipEventRecord(b->join[k], b->lane[k]);
    if (ej == hipSuccess) ej = hipStreamWaitEvent(b->stream, b->join[k], 0);
    if (ej != hipSuccess && status == SOAP_OK)
      status = dfail(b, SOAP_ERR_HIP, "diff batch: join: %s", hipGetErrorString(ej));
  }
  return status;
}

// Forward, loss and (if adjoint) the adjoint of one bound instance on stream st.
hipError_t dbatch_enqueue_one(soap_diffbatch::Item &it, hipStream_t st, bool adjoint) {
  {
    soap_ctx *c = it.c;
    auto &d = c->diff;
    DiffArgs a = diff_args(c);
    a.prm = it.prm;
    ConvArgs ca = diff_conv(c);
    int rb = 0;
    hipError_t e = launch_diff_forward_spectra(a, ca, it.target, c->trim, (double *)d.res_part.p, &rb, st);
    if (e == hipSuccess) e = ctx_ccf(c, (const double *)d.F.p, d.P, (double *)c->ccf_part.p, st);
    if (e == hipSuccess)
      e = launch_ccf_sum((const double *)c->ccf_part.p, ccf_chunks(c->n_lines), d.P, c->n_v, (double *)d.ccf.p,
                         st);
    if (e == hipSuccess)
      e = launch_fit(c->d_vgrid, c->n_v, (const double *)d.ccf.p, nullptr, d.P, (double *)d.par.p, (double *)d.model.p,
                     (double *)d.depth.p, (double *)d.bis.p, (int32_t *)d.iters.p, st);
    if (e == hipSuccess)
      e = launch_fit_polish(c->d_vgrid, c->n_v, (const double *)d.ccf.p, (const double *)d.par.p, d.P,
                            (double *)d.th.p, st);
    if (e == hipSuccess)
      e = it.w_bis != 0.0 ? launch_diff_bis(c->d_vgrid, c->n_v, (const double *)d.ccf.p, (const double *)d.th.p, d.P,
                                            (double *)d.bisd.p, st)
                          : hipSuccess;
    double *bisd = it.w_bis != 0.0 ? (double *)d.bisd.p : nullptr;  // no span term: no span kernel
    if (e == hipSuccess)
      e = launch_diff_loss(d.P, rb, (const double *)d.res_part.p, (const double *)d.th.p, it.rv_t, it.fw_t, it.bis_t,
                           it.w_spec, it.w_rv, it.w_fw, it.w_bis, bisd, c->n_v, it.prm, it.vsini_max,
                           (double)d.G * d.G, d.ns == 3 ? lb_factor_bound(d.lb) : 0.0, it.loss, (double *)d.dth.p,
                           st);
    if (!adjoint) return e;
    if (e == hipSuccess)
      e = launch_diff_backward(a, ca, it.target, c->trim, it.w_spec, c->d_vgrid, c->n_v, (const double *)d.ccf.p,
                               (const double *)d.th.p, (const double *)d.dth.p, (double *)d.gy.p,
                               c->n - 2 * (int64_t)c->trim, (const int64_t *)d.mt_ptr.p, (const int32_t *)d.mt_w.p, (const int32_t *)d.mt_col.p,
                               (const double *)d.mt_val.p, (double *)d.GF.p, (double *)d.GI.p, (double *)d.Cpart.p,
                               (double *)d.C.p, it.gmaps, it.gfmaps, (double *)d.cell_part.p, it.gq, it.gs, it.gf,
                               bisd, st);
    if (e == hipSuccess) e = launch_diff_grad(diff_cell_blocks((int64_t)d.G * d.G), (const double *)d.cell_part.p,
                                              it.g5, st);
    return e;
  }
}
}  // namespace

extern "C" {

int soap_diff_batch_create(soap_ctx *const *ctxs, int32_t n_inst, soap_diffbatch **out) {
  if (!out) return SOAP_ERR_ARG;
  *out = nullptr;
  if (!ctxs || n_inst < 1) return SOAP_ERR_ARG;
  auto *b = new soap_diffbatch();
  *out = b;
  b->device = ctxs[0] ? ctxs[0]->device : 0;
  for (int i = 0; i < n_inst; i++) {
    soap_ctx *c = ctxs[i];
    if (!c) return dfail(b, SOAP_ERR_ARG, "diff_batch_create: instance %d: null context", i);
    if (c->device != b->device) return dfail(b, SOAP_ERR_ARG, "diff_batch_create: contexts must share one device");
    if (!c->diff.ready) return dfail(b, SOAP_ERR_STATE, "diff_batch_create: instance %d: run diff_setup first", i);
    for (int j = 0; j < i; j++)
      if (ctxs[j] == c) return dfail(b, SOAP_ERR_ARG, "diff_batch_create: instance %d listed twice", i);
    soap_diffbatch::Item it;
    it.c = c;
    b->items.push_back(it);
  }
  hipError_t e = hipSetDevice(b->device);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&b->fork, hipEventDisableTiming);
  for (int k = 0; k < kDiffLanes && e == hipSuccess; k++) {
    e = hipStreamCreateWithFlags(&b->lane[k], hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&b->join[k], hipEventDisableTiming);
  }
  if (e != hipSuccess) return dfail(b, SOAP_ERR_HIP, "diff_batch_create: %s", hipGetErrorString(e));
  dbatch_assign_lanes(b);
  return SOAP_OK;
}

void soap_diff_batch_destroy(soap_diffbatch *b) {
  if (!b) return;
  hipSetDevice(b->device);
  if (b->stream) hipStreamSynchronize(b->stream);
  for (int k = 0; k < kDiffLanes; k++)
    if (b->lane[k]) hipStreamSynchronize(b->lane[k]);
  dbatch_drop_graph(b);
  for (int k = 0; k < kDiffLanes; k++) {
    if (b->join[k]) hipEventDestroy(b->join[k]);
    if (b->lane[k]) hipStreamDestroy(b->lane[k]);
  }
  if (b->fork) hipEventDestroy(b->fork);
  if (b->stream) hipStreamDestroy(b->stream);
  delete b;
}

const char *soap_diff_batch_last_error(const soap_diffbatch *b) { return b ? b->err.c_str() : "null batch"; }

int soap_diff_batch_bind(soap_diffbatch *b, int32_t i, double vsini_max, const double *d_params, const double *d_q,
                         const double *d_s, const double *d_f, const double *d_maps, const double *d_fmaps,
                         const double *d_target, const double *d_rv_t, const double *d_fw_t, const double *d_bis_t,
                         double w_spec, double w_rv, double w_fw, double w_bis, double *d_loss, double *d_grad5,
                         double *d_gq, double *d_gs, double *d_gf, double *d_gmaps, double *d_gfmaps) {
  if (!b || i < 0 || i >= (int32_t)b->items.size()) return SOAP_ERR_ARG;
  if (!d_params || !d_q || !d_s || !d_maps || !d_target || !d_rv_t || !d_fw_t || !d_bis_t || !d_loss || !d_grad5 ||
      !d_gq || !d_gs || !d_gmaps)
    return dfail(b, SOAP_ERR_ARG, "diff_batch_bind: null pointer");
  const bool fac = d_f != nullptr;
  if (fac != (d_fmaps != nullptr) || fac != (d_gf != nullptr) || fac != (d_gfmaps != nullptr))
    return dfail(b, SOAP_ERR_ARG, "diff_batch_bind: facula seeds, maps and their gradients go together");
  const int ns = fac ? 3 : 2;
  if (!(vsini_max > 0 && vsini_max < 2000)) return dfail(b, SOAP_ERR_ARG, "diff_batch_bind: vsini_max out of range");
  auto &it = b->items[i];
  soap_ctx *c = it.c;
  auto &d = c->diff;
  if (hipSetDevice(c->device) != hipSuccess) return dfail(b, SOAP_ERR_HIP, "diff_batch_bind: hipSetDevice");
  // fixed tap range: the shifts of delta in [-vsini_max, vsini_max] plus margins
  auto mshift = [&](double v) {
    double beta = v * 1000. / kC;
    double fb = -1 * (1 - sqrt((1 + beta) / (1 - beta)));
    return (int)ceil(log1p(fb) / c->h);
  };
  int dmin = mshift(-vsini_max) - 2;
  int nd = mshift(vsini_max) - dmin + 3;
  if (nd > 2048) return dfail(b, SOAP_ERR_LIMIT, "diff_batch_bind: shift range of %d pixels", nd);
  int rc;
  if ((rc = ensure(c, d.acc, (size_t)d.P * c->n_mu * nd * ns * 2 * 8)) ||
      (rc = ensure(c, d.taps, (size_t)d.P * c->n_mu * nd * ns * 8)) ||
      (rc = ensure(c, d.krange, (size_t)d.P * c->n_mu * 2 * 4)) ||
      (rc = ensure(c, d.C, (size_t)d.P * c->n_mu * nd * ns * 8)) ||
      (rc = ensure(c, d.Cpart, (size_t)d.P * c->n_mu * diff_corr_chunks(c->n) * nd * ns * 8)) ||
      (!c->mask_built && (rc = build_mask(c))) || (rc = diff_mask_transpose(c)))
    return dfail(b, rc, "diff_batch_bind: instance %d: %s", i, c->err.c_str());
  d.dmin = dmin;
  d.nd = nd;
  d.ns = ns;
  d.q = d_q;
  d.s = d_s;
  d.f = d_f;
  d.maps = d_maps;
  d.fmaps = d_fmaps;
  d.target = d_target;
  d.have_fwd = false;  // the per-call forward state no longer matches the buffers
  it.vsini_max = vsini_max;
  it.dmin = dmin;
  it.nd = nd;
  it.prm = d_params;
  it.q = d_q;
  it.s = d_s;
  it.f = d_f;
  it.maps = d_maps;
  it.fmaps = d_fmaps;
  it.target = d_target;
  it.rv_t = d_rv_t;
  it.fw_t = d_fw_t;
  it.bis_t = d_bis_t;
  it.w_bis = w_bis;
  it.w_spec = w_spec;
  it.w_rv = w_rv;
  it.w_fw = w_fw;
  it.loss = d_loss;
  it.g5 = d_grad5;
  it.gq = d_gq;
  it.gs = d_gs;
  it.gf = d_gf;
  it.gmaps = d_gmaps;
  it.gfmaps = d_gfmaps;
  it.bound = true;
  it.sig = dbatch_signature(c);
  dbatch_drop_graph(b);
  return SOAP_OK;
}

int soap_diff_batch_step(soap_diffbatch *b, int32_t use_graph, int32_t with_adjoint) {
  if (!b) return SOAP_ERR_ARG;
  if (hipSetDevice(b->device) != hipSuccess) return dfail(b, SOAP_ERR_HIP, "diff_batch_step: hipSetDevice");
  for (auto &it : b->items) {
    // the buffers were sized for the bound tap range; a per-call forward in between changes it
    if (it.bound && dbatch_signature(it.c) != it.sig)
      return dfail(b, SOAP_ERR_STATE, "diff_batch_step: an instance was changed after binding");
  }
  const bool adj = with_adjoint != 0;
  if (!use_graph) return dbatch_enqueue(b, adj);
  if (!b->exec[adj]) {
    hipError_t e = hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return dfail(b, SOAP_ERR_HIP, "diff_batch_step: begin capture: %s", hipGetErrorString(e));
    int rc = dbatch_enqueue(b, adj);
    hipGraph_t g = nullptr;
    e = hipStreamEndCapture(b->stream, &g);
    if (rc) {
      if (g) hipGraphDestroy(g);
      return rc;
    }
    if (e != hipSuccess) return dfail(b, SOAP_ERR_HIP, "diff_batch_step: end capture: %s", hipGetErrorString(e));
    b->graph[adj] = g;
    e = hipGraphInstantiate(&b->exec[adj], b->graph[adj], nullptr, nullptr, 0);
    if (e != hipSuccess) return dfail(b, SOAP_ERR_HIP, "diff_batch_step: instantiate: %s", hipGetErrorString(e));
  }
  hipError_t e = hipGraphLaunch(b->exec[adj], b->stream);
  if (e != hipSuccess) return dfail(b, SOAP_ERR_HIP, "diff_batch_step: launch: %s", hipGetErrorString(e));
  return SOAP_OK;
}

int soap_diff_batch_sync(soap_diffbatch *b) {
  if (!b) return SOAP_ERR_ARG;
  if (hipSetDevice(b->device) != hipSuccess || hipStreamSynchronize(b->stream) != hipSuccess)
    return dfail(b, SOAP_ERR_HIP, "diff_batch_sync failed");
  return SOAP_OK;
}

}  // extern "C"
