/*
 * soap_oracle.c -- FP64 CPU restatement of the SOAP-GPU reference hot path.
 *
 * TEST INFRASTRUCTURE ONLY: the parity checker for the HIP engine.  It is never
 * linked into, loaded by or called from the engine (soap_gpu_amd/).
 *
 * Structure follows the reference: per-cell loops in cell order, per-pixel
 * formulas as written there.  Parallelism is OpenMP over wavelength chunks, so
 * every pixel's accumulation order equals the reference's cell order and the
 * result does not depend on the thread count.
 *
 * Build: see oracle/Makefile (gcc -O3 -fopenmp -ffp-contract=off).
 *
 * Resolved undefined behaviour of the reference (each choice is deterministic):
 *  - scatter interpolation gaps/collisions -> gather: pixel j takes the last
 *    source interval i with ceil(p_i) <= j (SOAP_integration.cu:256-279);
 *  - source pixels outside [0, n) read as 0 (zero padding);
 *  - mu beyond the last boundary -> last bin (SOAP_initialize.cu:370-377);
 *  - spot_area cross-block race -> serial bounding box (SOAP_integration.cu:51-120);
 *  - uninitialised accumulators -> zero (SOAP_resolution.cu:252).
 */
#include "soap_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static const double ORC_PI = 3.141592653589793238463;
static const double ORC_C = 299792458.;

/* ------------------------------------------------------------------ helpers */

/* SOAP_resolution.cu:18-35 */
double orc_rndup(double n, int nb_decimal) {
  double t = n * pow(10, nb_decimal) - floor(n * pow(10, nb_decimal));
  if (t >= 0.5) {
    n *= pow(10, nb_decimal);
    n = ceil(n);
    n /= pow(10, nb_decimal);
  } else {
    n *= pow(10, nb_decimal);
    n = floor(n);
    n /= pow(10, nb_decimal);
  }
  return n;
}

/* device Delta_lambda, SOAP_integration.cu:216-226 */
double orc_delta_lambda(double v_ms, double lambda0) {
  double beta = v_ms / ORC_C;
  return -1 * lambda0 * (1 - sqrt((1 + beta) / (1 - beta)));
}

/* host Delta_lambda_cpu with its single-precision sqrtf, SOAP_resolution.cu:38-49 */
double orc_delta_lambda_cpu(double v_ms, double lambda0) {
  double beta = v_ms / ORC_C;
  return -1 * lambda0 * (1 - sqrtf((float)((1 + beta) / (1 - beta))));
}

/* log step of a log-uniform grid (first and last pixel define it) */
double orc_grid_h(const double *wave, long n) { return log(wave[n - 1] / wave[0]) / (double)(n - 1); }

/* fbeta of fast_shift / fast_interp_new, SOAP_integration.cu:236-237 */
static double orc_fbeta(double delta_kms) {
  double beta = delta_kms * 1000. / ORC_C;
  return -1 * (1 - sqrt((1 + beta) / (1 - beta)));
}

/* lambda + Delta_lambda(delta*1000, lambda), SOAP_integration.cu:240 */
static double orc_lshift(double lambda, double delta_kms) {
  return lambda + orc_delta_lambda(delta_kms * 1000., lambda);
}

/* Least squares polynomial fit (degree deg) by Householder QR on the
 * Vandermonde matrix, highest power first like numpy.polyfit. */
static void orc_polyfit(const double *x, const double *y, int n, int deg, double *coef) {
  int p = deg + 1;
  double A[16][8], b[16];
  for (int i = 0; i < n; i++) {
    for (int j = 0; j < p; j++) A[i][j] = pow(x[i], deg - j);
    b[i] = y[i];
  }
  for (int j = 0; j < p; j++) {
    double nrm = 0;
    for (int i = j; i < n; i++) nrm += A[i][j] * A[i][j];
    nrm = sqrt(nrm);
    double alpha = A[j][j] > 0 ? -nrm : nrm;
    double v[16];
    for (int i = 0; i < n; i++) v[i] = 0;
    v[j] = A[j][j] - alpha;
    for (int i = j + 1; i < n; i++) v[i] = A[i][j];
    double vv = 0;
    for (int i = j; i < n; i++) vv += v[i] * v[i];
    if (vv == 0) continue;
    for (int k = j; k < p; k++) {
      double s = 0;
      for (int i = j; i < n; i++) s += v[i] * A[i][k];
      s = 2 * s / vv;
      for (int i = j; i < n; i++) A[i][k] -= s * v[i];
    }
    double s = 0;
    for (int i = j; i < n; i++) s += v[i] * b[i];
    s = 2 * s / vv;
    for (int i = j; i < n; i++) b[i] -= s * v[i];
  }
  for (int j = p - 1; j >= 0; j--) {
    double s = b[j];
    for (int k = j + 1; k < p; k++) s -= A[j][k] * coef[k];
    coef[j] = s / A[j][j];
  }
}

/* LB_boundary + LB_params, run_all_SOAP.py:8-34,140-145.
 * out4 = [LB0, LB1, LB2, LB3] (constant term first, as passed on argv). */
int orc_lb_params(int t_star, const double *lb_g2, const double *lb_k0, const double *lb_m0,
                  const double *temps3, double *out4) {
  int b0, b1;
  if (t_star > 5778 || (t_star <= 5778 && t_star >= 5240)) {
    b0 = 0;
    b1 = 1;
  } else {
    b0 = 1;
    b1 = 2;
  }
  double mu[10], lb[10];
  for (int k = 0; k < 10; k++) mu[k] = (k == 9) ? 1.0 : 0.1 + k * (0.9 / 9);
  for (int id = 0; id < 10; id++) {
    double yv[3] = {lb_g2[id], lb_k0[id], lb_m0[id]};
    double xs[2] = {temps3[b0], temps3[b1]}, ys[2] = {yv[b0], yv[b1]}, c1[2];
    orc_polyfit(xs, ys, 2, 1, c1);
    lb[id] = c1[0] * t_star + c1[1];
  }
  double c3[4];
  orc_polyfit(mu, lb, 10, 3, c3);
  for (int k = 0; k < 4; k++) out4[k] = c3[3 - k];
  return 0;
}

/* --------------------------------------------------------------- a1: disk */

/* itot cell loop, SOAP_initialize.cu:297-395.  pow(x,.5) is written sqrt(x)
 * and pow(s,2) is written s*s (both exact restatements of the mathematics). */
void orc_disk_init(int grid, double prot, double w_ratio, double incl_deg, double radius_km, double u1,
                   double u2, const double *lb4, const double *mu_bounds, int n_mu, double *mu_map,
                   double *delta_map, double *limb_map, int *region_map, double *lb_map) {
  const double pi = ORC_PI;
  double incl = incl_deg * pi / 180.;
  double delta_grid = 2. / grid;
  double theta_grid = 180.0 / grid;
  double omega0 = 360.0 / prot;
  double omega1 = omega0 * w_ratio;
  double sin_incl = sin(incl);
  for (int iy = 0; iy < grid; iy++) {
    double y = -1. + iy * delta_grid;
    for (int iz = 0; iz < grid; iz++) {
      double z = -1. + iz * delta_grid;
      double theta_z = -90.0 + iz * theta_grid;
      double s = sin(theta_z * pi / 180);
      double omega_velo = (omega0 + omega1 * (s * s)) * pi / (180 * 24 * 3600) * radius_km;
      double delta = y * omega_velo * sin_incl;
      long idx = (long)iy * grid + iz;
      mu_map[idx] = -999.0;
      delta_map[idx] = -999.0;
      limb_map[idx] = -999.0;
      lb_map[idx] = -999.0;
      region_map[idx] = 0;
      if ((y * y + z * z) <= 1.) {
        double r_cos = sqrt(1. - (y * y + z * z));
        double limb = 1. - u1 * (1 - r_cos) - u2 * (1 - r_cos) * (1 - r_cos);
        double lb = lb4[0] + lb4[1] * r_cos + lb4[2] * r_cos * r_cos + lb4[3] * r_cos * r_cos * r_cos;
        int k = n_mu - 1;
        for (int b = 0; b < n_mu - 1; b++)
          if (r_cos > mu_bounds[b]) {
            k = b;
            break;
          }
        mu_map[idx] = r_cos;
        delta_map[idx] = delta;
        limb_map[idx] = limb;
        lb_map[idx] = lb;
        region_map[idx] = k;
      }
    }
  }
}

/* ------------------------------------------------------- a4: shift/interp */

static inline double orc_row_at(const double *row, long n, long i) {
  return (i >= 0 && i < n) ? row[i] : 0.0;
}

/* wavelength of pixel i, with linear continuation outside [0, n) */
static inline double orc_wave_at(const double *wave, long n, long i) {
  if (i < 0) return wave[0] + (double)i * (wave[1] - wave[0]);
  if (i >= n) return wave[n - 1] + (double)(i - (n - 1)) * (wave[n - 1] - wave[n - 2]);
  return wave[i];
}

/* ceil(p_i) of the reference scatter index, SOAP_integration.cu:258 */
static inline long orc_ceil_p(long i, double fbeta, double lambda0, double step) {
  return (long)ceil(i * (1 + fbeta) + fbeta * lambda0 / step);
}

/* gather source index on a linear grid: max{i : ceil(p_i) <= j} (may be -1 or >= n-1) */
static long orc_gather_index_linear(long j, double fbeta, double lambda0, double step) {
  double est = (j - fbeta * lambda0 / step) / (1 + fbeta);
  long i = (long)floor(est);
  while (orc_ceil_p(i, fbeta, lambda0, step) > j) i--;
  while (orc_ceil_p(i + 1, fbeta, lambda0, step) <= j) i++;
  return i;
}

/* Value of Shift(row, delta) at pixel j, times limb.
 *  linear: fast_interp_new's expression at the gather index, SOAP_integration.cu:246-281;
 *  log:    the same linear interpolation with the exact fractional weight
 *          t = (lambda_j - lambda_s,i)/(lambda_s,i+1 - lambda_s,i) = expm1((m-s)h)/expm1(h),
 *          s = log1p(fbeta)/h, m = ceil(s), i = j - m. */
typedef struct {
  int kind;
  double delta, fbeta, lambda0, step;
  long m; /* log */
  double t; /* log */
} orc_shift;

static orc_shift orc_shift_setup(int kind, const double *wave, long n, double step, double delta) {
  orc_shift s;
  s.kind = kind;
  s.delta = delta;
  s.fbeta = orc_fbeta(delta);
  s.lambda0 = wave[0];
  s.step = step;
  s.m = 0;
  s.t = 0;
  if (kind == ORC_GRID_LOG) {
    double h = orc_grid_h(wave, n);
    double sh = log1p(s.fbeta) / h;
    double m = ceil(sh);
    s.m = (long)m;
    s.t = expm1((m - sh) * h) / expm1(h);
  }
  return s;
}

static inline double orc_shift_value(const orc_shift *s, const double *wave, long n, const double *row,
                                     long j, double limb) {
  if (s->delta == 0.0) return orc_row_at(row, n, j) * limb; /* SOAP_integration.cu:262-265 */
  if (s->kind == ORC_GRID_LOG) {
    long i = j - s->m;
    double si = orc_row_at(row, n, i);
    double g = orc_row_at(row, n, i + 1) - si;
    return (si + g * s->t) * limb;
  }
  long i = orc_gather_index_linear(j, s->fbeta, s->lambda0, s->step);
  double si = orc_row_at(row, n, i);
  double g = orc_row_at(row, n, i + 1) - si;
  double ls0 = orc_lshift(orc_wave_at(wave, n, i), s->delta);
  double ls1 = orc_lshift(orc_wave_at(wave, n, i + 1), s->delta);
  return (si + g / (ls1 - ls0) * (wave[j] - ls0)) * limb;
}

void orc_shift_row(int kind, const double *wave, long n, double step, const double *row,
                   double delta_kms, double limb, double *out) {
  orc_shift s = orc_shift_setup(kind, wave, n, step, delta_kms);
  for (long j = 0; j < n; j++) out[j] = orc_shift_value(&s, wave, n, row, j, limb);
}

/* literal serial scatter (fast_interp_new executed in ascending index order),
 * SOAP_integration.cu:246-281; buf keeps values of earlier cells in the gaps */
void orc_scatter_row_ref(const double *wave, long n, double step, const double *row, double delta_kms,
                         double limb, double *buf) {
  double fbeta = orc_fbeta(delta_kms);
  double lambda0 = wave[0];
  for (long i = 0; i < n; i++) {
    if (delta_kms == 0.0) {
      buf[i] = row[i] * limb;
      continue;
    }
    long j = orc_ceil_p(i, fbeta, lambda0, step);
    if (j < 0 || j >= n) continue;
    double g = orc_row_at(row, n, i + 1) - row[i];
    double ls0 = orc_lshift(wave[i], delta_kms);
    double ls1 = orc_lshift(orc_wave_at(wave, n, i + 1), delta_kms);
    buf[j] = (row[i] + g / (ls1 - ls0) * (wave[j] - ls0)) * limb;
  }
}

/* --------------------------------------------------------- a2: quiet star */

static void orc_chunk(long n, int nchunk, int c, long *j0, long *j1) {
  *j0 = n * c / nchunk;
  *j1 = n * (c + 1) / nchunk;
}

/* itot + fast_sum accumulation, SOAP_initialize.cu:339-395,171-176 */
void orc_quiet_star(int kind, const double *wave, long n, double step, int grid,
                    const double *delta_map, const double *limb_map, const int *region_map,
                    const double *cube, int n_mu, double *out, int nthreads) {
  (void)n_mu;
  long ncell = (long)grid * grid;
  for (long j = 0; j < n; j++) out[j] = 0.0;
  int nchunk = nthreads > 0 ? nthreads * 4 : 32;
#pragma omp parallel for schedule(dynamic) num_threads(nthreads > 0 ? nthreads : 1)
  for (int c = 0; c < nchunk; c++) {
    long j0, j1;
    orc_chunk(n, nchunk, c, &j0, &j1);
    for (long idx = 0; idx < ncell; idx++) {
      if (limb_map[idx] == -999.0 && delta_map[idx] == -999.0) continue; /* off disk */
      orc_shift s = orc_shift_setup(kind, wave, n, step, delta_map[idx]);
      const double *row = cube + (long)region_map[idx] * n;
      double limb = limb_map[idx];
      for (long j = j0; j < j1; j++) out[j] = out[j] + orc_shift_value(&s, wave, n, row, j, limb);
    }
  }
}

/* 'ref' mode: literal scatter with the carried output buffer, linear grids */
void orc_quiet_star_ref(const double *wave, long n, double step, int grid, const double *delta_map,
                        const double *limb_map, const int *region_map, const double *cube, int n_mu,
                        double *out) {
  (void)n_mu;
  double *buf = (double *)calloc((size_t)n, sizeof(double));
  for (long j = 0; j < n; j++) out[j] = 0.0;
  long ncell = (long)grid * grid;
  for (long idx = 0; idx < ncell; idx++) {
    if (limb_map[idx] == -999.0 && delta_map[idx] == -999.0) continue;
    orc_scatter_row_ref(wave, n, step, cube + (long)region_map[idx] * n, delta_map[idx], limb_map[idx],
                        buf);
    for (long j = 0; j < n; j++) out[j] = out[j] + buf[j];
  }
  free(buf);
}

/* --------------------------------------------------- a5: region geometry */

/* spot_init_gpu + spot_phase_gpu + spot_area (serial), SOAP_integration.cu:21-161 */
int orc_region_bbox(double lat, double lon, double size, double incl_deg, int nrho, int grid,
                    double phase, int *bnd5) {
  const double pi = ORC_PI;
  double s = size;
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
  double axe[3] = {cos(inc), 0, sin(inc)};
  double P[3][3] = {{(1 - cos(psi)) * axe[0] * axe[0] + cos(psi), (1 - cos(psi)) * axe[0] * axe[1] + sin(psi) * axe[2],
                     (1 - cos(psi)) * axe[0] * axe[2] - sin(psi) * axe[1]},
                    {(1 - cos(psi)) * axe[1] * axe[0] - sin(psi) * axe[2], (1 - cos(psi)) * axe[1] * axe[1] + cos(psi),
                     (1 - cos(psi)) * axe[1] * axe[2] + sin(psi) * axe[0]},
                    {(1 - cos(psi)) * axe[2] * axe[0] + sin(psi) * axe[1],
                     (1 - cos(psi)) * axe[2] * axe[1] - sin(psi) * axe[0],
                     (1 - cos(psi)) * axe[2] * axe[2] + cos(psi)}};
  int sum_on = 0, sum_off = 0;
  double miny = 1, minz = 1, maxy = -1, maxz = -1;
  for (int k = 0; k < nrho; k++) {
    double rho = -pi + k * rho_step;
    double x2[3] = {sqrt(1 - s * s), s * cos(rho), s * sin(rho)};
    double x1[3], x3[3];
    for (int r = 0; r < 3; r++) x1[r] = R[r][0] * x2[0] + R[r][1] * x2[1] + R[r][2] * x2[2];
    for (int r = 0; r < 3; r++) x3[r] = P[r][0] * x1[0] + P[r][1] * x1[1] + P[r][2] * x1[2];
    if (x3[0] >= 0) {
      sum_on++;
      if (x3[1] < miny) miny = x3[1];
      if (x3[2] < minz) minz = x3[2];
      if (x3[1] > maxy) maxy = x3[1];
      if (x3[2] > maxz) maxz = x3[2];
    } else {
      sum_off++;
    }
  }
  if ((sum_on > 0) && (sum_off > 0)) {
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
  double grid_step = 2. / grid;
  bnd5[0] = (int)floor((1. + miny) / grid_step);
  bnd5[1] = (int)floor((1. + minz) / grid_step);
  bnd5[2] = (int)ceil((1. + maxy) / grid_step);
  bnd5[3] = (int)ceil((1. + maxz) / grid_step);
  bnd5[4] = sum_on > 0 ? 1 : 0;
  return bnd5[4];
}

/* spot_inverse_rotation matrix R3 = R2 R, SOAP_integration.cu:164-202 (phase pre-decrement kept) */
static void orc_inverse_rotation(double lon, double lat, double incl_deg, double phase, double R3[3][3]) {
  const double pi = ORC_PI;
  double g2 = (phase - 1) * (2 * pi);
  double i = incl_deg * pi / 180.;
  double b = lat * pi / 180.;
  double g = -lon * pi / 180.;
  double b2 = -(pi / 2. - i);
  double R[3][3] = {{(1 - cos(g2)) * cos(i) * cos(i) + cos(g2), sin(g2) * sin(i), (1 - cos(g2)) * cos(i) * sin(i)},
                    {-sin(g2) * sin(i), cos(g2), sin(g2) * cos(i)},
                    {(1 - cos(g2)) * sin(i) * cos(i), -sin(g2) * cos(i), (1 - cos(g2)) * sin(i) * sin(i) + cos(g2)}};
  double R2[3][3] = {{cos(b) * cos(g) * cos(b2) - sin(b2) * sin(b), -sin(g) * cos(b), cos(b) * cos(g) * sin(b2) + sin(b) * cos(b2)},
                     {sin(g) * cos(b2), cos(g), sin(g) * sin(b2)},
                     {-sin(b) * cos(g) * cos(b2) - cos(b) * sin(b2), sin(b) * sin(g), -sin(b) * cos(g) * sin(b2) + cos(b) * cos(b2)}};
  for (int r = 0; r < 3; r++)
    for (int c = 0; c < 3; c++) R3[r][c] = R2[r][0] * R[0][c] + R2[r][1] * R[1][c] + R2[r][2] * R[2][c];
}

/* per-phase temperature and binary maps: loop over regions in order, spot_scan
 * for each visible one (later regions overwrite), SOAP_integration.cu:686-785,326-390 */
void orc_phase_maps(int grid, double incl_deg, int nrho, int t_star, int t_diff_spot,
                    int t_diff_fac, double phase, const orc_region *regs, int n_reg,
                    double *temp_map, double *bi_map) {
  long ncell = (long)grid * grid;
  for (long k = 0; k < ncell; k++) {
    temp_map[k] = -9999.0;
    bi_map[k] = -9999.0;
  }
  double delta_grid = 2. / grid;
  for (int r = 0; r < n_reg; r++) {
    int bnd[5];
    orc_region_bbox(regs[r].lat, regs[r].lon, regs[r].size, incl_deg, nrho, grid, phase, bnd);
    if (bnd[4] != 1) continue;
    double R3[3][3];
    orc_inverse_rotation(regs[r].lon, regs[r].lat, incl_deg, phase, R3);
    double s = regs[r].size;
    int y0 = bnd[0] < 0 ? 0 : bnd[0], y1 = bnd[2] > grid ? grid : bnd[2];
    int z0 = bnd[1] < 0 ? 0 : bnd[1], z1 = bnd[3] > grid ? grid : bnd[3];
    for (int iy = y0; iy < y1; iy++) {
      double y = -1. + iy * delta_grid;
      for (int iz = z0; iz < z1; iz++) {
        double z = -1. + iz * delta_grid;
        if (z * z + y * y < 1.) {
          double xa[3] = {sqrt(1. - (y * y + z * z)), y, z};
          double xi0 = R3[0][0] * xa[0] + R3[0][1] * xa[1] + R3[0][2] * xa[2];
          if (xi0 * xi0 >= 1. - s * s) {
            long idx = (long)iy * grid + iz;
            double r_cos = sqrt(1. - (y * y + z * z));
            if (regs[r].type == 0.0) {
              int T_spot = t_star - t_diff_spot;
              temp_map[idx] = T_spot;
              bi_map[idx] = 0.0;
            } else {
              int T_plage = (int)(t_star + t_diff_fac + 0.9 - 407.7 * r_cos + 190.9 * (r_cos * r_cos));
              temp_map[idx] = T_plage;
              bi_map[idx] = 1.0;
            }
          }
        }
      }
    }
  }
}

/* ------------------------------------------------ a6: active integration */

/* loi_Planck, SOAP_integration.cu:205-212 */
static double orc_planck(double lambda0, double temp) {
  double c = 299792458.;
  double h = 6.62606896e-34;
  double k_b = 1.380e-23;
  double scale = 1e-10;
  return 2 * h * pow(c, 2) * 1. / pow(lambda0 * scale, 5) * 1. / (exp((h * c) / (lambda0 * k_b * temp * scale)) - 1);
}

/* spot_integrate, SOAP_integration.cu:392-447 (solar variant: SOAP_integration_solar.cu:433) */
void orc_integrate_phase(int kind, const double *wave, long n, double step, int grid,
                         const double *delta_map, const double *limb_map, const int *region_map,
                         const double *lb_map, const double *temp_map, const double *bi_map,
                         const double *cube_q, const double *cube_qsed, const double *cube_sp,
                         const double *cube_spsed, const double *cube_fac, const double *cube_facsed,
                         int n_mu, int t_star, int t_diff_fac, int solar_mode, double *flux,
                         double *bconv, double *tot, int nthreads) {
  (void)n_mu;
  long ncell = (long)grid * grid;
  for (long j = 0; j < n; j++) flux[j] = bconv[j] = tot[j] = 0.0;
  /* active cell list in map order */
  long nact = 0;
  for (long idx = 0; idx < ncell; idx++)
    if (bi_map[idx] == 0.0 || bi_map[idx] == 1.0) nact++;
  long *act = (long *)malloc(sizeof(long) * (nact > 0 ? nact : 1));
  nact = 0;
  for (long idx = 0; idx < ncell; idx++)
    if (bi_map[idx] == 0.0 || bi_map[idx] == 1.0) act[nact++] = idx;
  int t_norm_fac = (int)(t_star + t_diff_fac + 0.9 - 216.8);
  int nchunk = nthreads > 0 ? nthreads * 4 : 32;
#pragma omp parallel for schedule(dynamic) num_threads(nthreads > 0 ? nthreads : 1)
  for (int c = 0; c < nchunk; c++) {
    long j0, j1;
    orc_chunk(n, nchunk, c, &j0, &j1);
    for (long a = 0; a < nact; a++) {
      long idx = act[a];
      int is_spot = bi_map[idx] == 0.0;
      orc_shift s = orc_shift_setup(kind, wave, n, step, delta_map[idx]);
      long roff = (long)region_map[idx] * n;
      double limb = limb_map[idx];
      const double *act_cube = is_spot ? cube_sp : cube_fac;
      for (long j = j0; j < j1; j++) {
        double spec_out = orc_shift_value(&s, wave, n, cube_q + roff, j, limb);
        double sed_out = orc_shift_value(&s, wave, n, cube_qsed + roff, j, limb);
        double intensity;
        if (is_spot) {
          int t_norm = (int)temp_map[idx];
          intensity = orc_planck(wave[j], temp_map[idx]) / orc_planck(wave[j], t_norm) * cube_spsed[j];
        } else if (solar_mode) {
          intensity = orc_planck(wave[j], temp_map[idx]) / orc_planck(wave[j], t_norm_fac) * cube_facsed[j];
        } else {
          intensity = (1 + lb_map[idx]) * cube_facsed[j];
        }
        double spot_out = orc_shift_value(&s, wave, n, act_cube + roff, j, limb);
        flux[j] = flux[j] + sed_out - intensity * spec_out;
        bconv[j] = bconv[j] + spec_out - spot_out;
        tot[j] = tot[j] + sed_out - spot_out * intensity;
      }
    }
  }
  free(act);
}

/* -------------------------------------------------------- a7: resolution */

/* nb_pixel of lower_resolution_gpu, SOAP_resolution.cu:241-244 */
int orc_conv_window(const double *wave, double sigma_ms) {
  double window_width = 50.;
  double wavelength_step = wave[1] - wave[0];
  return (int)orc_rndup((window_width * orc_delta_lambda_cpu(sigma_ms, wave[0])) / wavelength_step, 0);
}

/* conv_inst (SOAP_initialize.cu:78-96, SOAP_resolution.cu:76-95): the
 * unnormalised output o = sum_{m=i-w}^{i+w-2} weight (1 - A[m]) G_i(m - i),
 * zero outside w < i < n-1-w (A = spec for variant 0, quiet - spec for
 * variant 1; weight = the argument for variant 0, 0.005 for variant 1).
 * Linear grids use conv_inst literally, sigma_i = Delta_lambda(sigma, lambda_i).
 * Log grids replace the hard-coded 0.005 A pixel by the exact geometry of the
 * grid: offset lambda_{i+d} - lambda_i = lambda_i expm1(d h), sigma_i = lambda_i fs,
 * quadrature weight lambda_{i+d} expm1(h); every factor lambda_i cancels and the
 * tap is c_d = exp(d h) expm1(h)/(fs sqrt(2 pi)) exp(-(expm1(d h)/fs)^2/2). */
void orc_conv_raw(int kind, const double *wave, long n, double sigma_ms, const double *A, int variant,
                  double weight, double *o, int nthreads) {
  const double pi = 3.141592653589793238463;
  int wid = orc_conv_window(wave, sigma_ms);
  for (long i = 0; i < n; i++) o[i] = 0.0;
  double w = variant == 0 ? weight : 0.005;
  double *ctap = NULL;
  if (kind == ORC_GRID_LOG) {
    double h = orc_grid_h(wave, n);
    double fs = orc_delta_lambda(sigma_ms, 1.0);
    ctap = (double *)malloc(sizeof(double) * (2 * (size_t)wid + 1));
    for (int d = -wid; d <= wid; d++) {
      double u = expm1(d * h) / fs;
      ctap[d + wid] = exp(d * h) * expm1(h) / (fs * sqrt(2. * pi)) * exp(-u * u / 2);
    }
  }
#pragma omp parallel for schedule(static) num_threads(nthreads > 0 ? nthreads : 1)
  for (long i = 0; i < n; i++) {
    if (!((i > wid) && (i + wid < n - 1))) continue;
    double acc = 0.0;
    if (kind == ORC_GRID_LOG) {
      for (long m = i - wid; m < i + wid - 1; m++) acc += ctap[m - i + wid] * (1.0 - A[m]);
    } else {
      double sig = orc_delta_lambda(sigma_ms, wave[i]);
      for (long m = i - wid; m < i + wid - 1; m++) {
        int d = (int)(m - i);
        acc += w * (1.0 - A[m]) * 1. / (sig * sqrt(2. * pi)) * exp(-d * d * 0.005 * 0.005 / (2 * sig * sig));
      }
    }
    o[i] = acc;
  }
  free(ctap);
}

/* lower_resolution_gpu: conv_inst, then spec_norm with the min of A and the max
 * of the non-zero outputs.
 *  variant 0: quiet-star form, SOAP_initialize.cu:446-505 (A = spec, weight = argument)
 *  variant 1: active form, SOAP_resolution.cu:235-310 (A = quiet - spec, weight = 0.005)
 * minmax2 (optional) receives [min, max] as used by spec_norm. */
void orc_lower_resolution(int kind, const double *wave, long n, double sigma_ms,
                          const double *quiet, const double *spec, int variant, double weight,
                          double *out, double *minmax2, int nthreads) {
  if (n <= 0) return;
  double *A = (double *)malloc(sizeof(double) * n);
  double *o = (double *)calloc((size_t)n, sizeof(double));
  for (long m = 0; m < n; m++) A[m] = variant == 0 ? spec[m] : quiet[m] - spec[m];
  orc_conv_raw(kind, wave, n, sigma_ms, A, variant, weight, o, nthreads);
  /* reduceMinIdxOptimizedShared over A, reduceMax over o with the -60000 seed
   * and zero filter: SOAP_resolution.cu:123-214 */
  double mn = A[0];
  for (long m = 0; m < n; m++)
    if (mn > A[m]) mn = A[m];
  double mx = -60000.0;
  for (long i = 0; i < n; i++)
    if (o[i] != 0.0 && mx < o[i]) mx = o[i];
  for (long i = 0; i < n; i++) out[i] = 1.0 - o[i] * (1 - mn) / mx; /* spec_norm, :225-232 */
  if (minmax2) {
    minmax2[0] = mn;
    minmax2[1] = mx;
  }
  free(A);
  free(o);
}

/* ---------------------------------------------------------------- a8: CCF */

/* line selection, run_all_SOAP.py:285-290 */
long orc_select_lines(const double *wave_t, long n_t, const double *line_l, const double *line_w,
                      long n_lines, double *sel_l, double *sel_w) {
  double lo = wave_t[0], hi = wave_t[0];
  for (long i = 0; i < n_t; i++) {
    if (wave_t[i] < lo) lo = wave_t[i];
    if (wave_t[i] > hi) hi = wave_t[i];
  }
  long k = 0;
  for (long i = 0; i < n_lines; i++)
    if (line_l[i] > lo && line_l[i] < hi && line_w[i] > 0.1) {
      sel_l[k] = line_l[i];
      sel_w[k] = line_w[i];
      k++;
    }
  return k;
}

/* numpy.linspace(start, stop, num)[k] */
static double orc_linspace(double start, double stop, long num, long k) {
  if (k == num - 1) return stop;
  double st = (stop - start) / (double)(num - 1);
  return (double)k * st + start;
}

/* calculate_CCF_1 / calc_mask for one velocity + calculate_CCF_2 dot product,
 * CCF_func.py:5-36,45-91, extended grid of run_all_SOAP.py:292-295.
 * linear: literal (requires int(100/step) == 20000, the slicing offset of CCF_func.py:8-9);
 * log:    the /0.005 index arithmetic is replaced by the exact inverse grid map
 *         x(l) = ln(l/l_t0)/h on the trimmed grid (step carries h), fractions
 *         by their exact log-grid values, the grid continued analytically. */
void orc_ccf(int kind, const double *wave_t, long n_t, double step, const double *vgrid, int n_v,
             const double *line_l, const double *line_w, long n_lines, const double *spec_t,
             double *ccf, int nthreads) {
  const double c = 299792458.;
  /* the extended grid exists for the linear form only (the log form's step is h) */
  long no_points = kind == ORC_GRID_LOG ? 0 : (long)(100.0 / step);
  long n_ext = n_t + 2 * no_points;
  double begin_wave = wave_t[0] - step * no_points;
#pragma omp parallel for schedule(dynamic) num_threads(nthreads > 0 ? nthreads : 1)
  for (int iv = 0; iv < n_v; iv++) {
    double *mask = (double *)calloc((size_t)(kind == ORC_GRID_LOG ? n_t : n_ext), sizeof(double));
    for (long li = 0; li < n_lines; li++) {
      double ls = line_l[li] * (1 + vgrid[iv] / c);
      double beta = 820. / c;
      double hw = ls * (sqrt((1 + beta) / (1 - beta)) - 1);
      double bgn = ls - hw / 2., end = ls + hw / 2.;
      double w = line_w[li];
      if (kind == ORC_GRID_LOG) {
        double h = step;
        double xb = log(bgn / wave_t[0]) / h, xe = log(end / wave_t[0]) / h;
        long aa = (long)ceil(xb);
        long bb = (long)(ceil(xe) - 1);
        double fb = -expm1(-(aa - xb) * h) / -expm1(-h);
        double fa = fabs(expm1((xe - bb) * h)) / expm1(h);
        for (long p = aa; p < bb; p++)
          if (p >= 0 && p < n_t) mask[p] = w;
        if (aa - 1 >= 0 && aa - 1 < n_t) mask[aa - 1] = w * fb;
        if (bb >= 0 && bb < n_t) mask[bb] = w * fa;
      } else {
        long aa = (long)ceil((bgn - begin_wave) / 0.005);
        long bb = (long)(ceil((end - begin_wave) / 0.005) - 1);
#define EXT(k)                                                                                    \
  ((k) < no_points ? orc_linspace(begin_wave, wave_t[0], no_points, (k))                          \
                   : ((k) < no_points + n_t ? wave_t[(k)-no_points]                               \
                                            : orc_linspace(wave_t[n_t - 1], wave_t[n_t - 1] + step * no_points, \
                                                           no_points, (k)-no_points - n_t)))
        if (aa < 1 || bb + 1 >= n_ext) continue;
        double fb = fabs(EXT(aa) - bgn) / (EXT(aa) - EXT(aa - 1));
        double fa = fabs(EXT(bb) - end) / (EXT(bb + 1) - EXT(bb));
#undef EXT
        for (long p = aa; p < bb; p++) mask[p] = w;
        mask[aa - 1] = w * fb;
        mask[bb] = w * fa;
      }
    }
    const double *mk = kind == ORC_GRID_LOG ? mask : mask + 20000;
    double acc = 0.0;
    for (long p = 0; p < n_t; p++) acc += spec_t[p] * mk[p];
    ccf[iv] = acc;
    free(mask);
  }
}

/* appodisation window, run_all_SOAP.py:306-311 */
void orc_appodisation(int n_v, double *appod) {
  int nb_zeros = 5;
  int period = (int)((n_v - 1) / 2.);
  int len_app = (int)(period / 2.);
  const double pi = 3.141592653589793238463;
  int k = 0;
  for (int i = 0; i < nb_zeros; i++) appod[k++] = 0.0;
  for (int a = 0; a < len_app; a++) appod[k++] = 0.5 * cos(2 * pi / period * a - pi) + 0.5;
  for (int i = 0; i < n_v - period - 2 * nb_zeros; i++) appod[k++] = 1.0;
  for (int a = len_app - 1; a >= 0; a--) appod[k++] = 0.5 * cos(2 * pi / period * a - pi) + 0.5;
  for (int i = 0; i < nb_zeros; i++) appod[k++] = 0.0;
}

/* CCF /= max(CCF); CCF = 1-((-CCF+1)*appod), run_all_SOAP.py:368-370 */
void orc_ccf_finish(int n_v, const double *appod, double *ccf) {
  double mx = ccf[0];
  for (int i = 1; i < n_v; i++)
    if (ccf[i] > mx) mx = ccf[i];
  for (int i = 0; i < n_v; i++) {
    double cc = ccf[i] / mx;
    ccf[i] = 1 - ((-cc + 1) * appod[i]);
  }
}

/* --------------------------------------------------- a9: Gaussian + BIS */

/* gauss_f / gauss_df, BIS_FIT2.cpp:31-81: residual (fit - y)/err, Jacobian rows
 * divided by err; err NULL = unit errors (the driver writes ones,
 * run_all_SOAP.py:317) */
static void orc_gauss(const double *px, const double *y, const double *err, int n, const double *v, double *f,
                      double *fit, double (*J)[4]) {
  double c = v[0], k = v[1], x0 = v[2], fwhm = v[3];
  double sigma = fwhm / 2 / sqrt(2 * log(2));
  for (int i = 0; i < n; i++) {
    double ei = err ? err[i] : 1.0;
    double e = exp(-(px[i] - x0) * (px[i] - x0) / 2 / sigma / sigma);
    double fi = c + k * e;
    if (fit) fit[i] = fi;
    if (f) f[i] = (fi - y[i]) / ei;
    if (J) {
      J[i][0] = 1 / ei;
      J[i][1] = e / ei;
      J[i][2] = k / ei * e * (px[i] - x0) / sigma / sigma;
      J[i][3] = k / ei * e * (px[i] - x0) * (px[i] - x0) / sigma / sigma / sigma / 2 / sqrt(2 * log(2));
    }
  }
}

/* solve (A + mu diag(D^2)) dx = -g by Cholesky (4x4); returns 0 on success */
static int orc_solve4(const double A[4][4], const double *D, double mu, const double *g, double *dx) {
  double L[4][4] = {{0}};
  for (int i = 0; i < 4; i++) {
    for (int j = 0; j <= i; j++) {
      double s = A[i][j] + (i == j ? mu * D[i] * D[i] : 0.0);
      for (int k = 0; k < j; k++) s -= L[i][k] * L[j][k];
      if (i == j) {
        if (!(s > 0)) return -1;
        L[i][i] = sqrt(s);
      } else {
        L[i][j] = s / L[j][j];
      }
    }
  }
  double z[4];
  for (int i = 0; i < 4; i++) {
    double s = -g[i];
    for (int k = 0; k < i; k++) s -= L[i][k] * z[k];
    z[i] = s / L[i][i];
  }
  for (int i = 3; i >= 0; i--) {
    double s = z[i];
    for (int k = i + 1; k < 4; k++) s -= L[k][i] * dx[k];
    dx[i] = s / L[i][i];
  }
  return 0;
}

/* Bisector and span of a CCF y over velocities v for the Gaussian (c, k, v0, fwhm)
 * with signed amplitude k and absolute centre v0, BIS_FIT2.cpp:214-268: per-interval
 * quadratic inverse p0 + p1 d + p2 d^2 of the normalised CCF norm = -c / k (1 - y / c),
 * p2 from the Gaussian's inverse curvature at the interval centre; for each depth d = (i + 5) / 100, i = 0..90, the intervals
 * reached by walking out from argmax(norm) while norm > d; bis = the mean of the
 * two sides; span = mean(bis, d in [0.1, 0.4]) - mean(bis, d in [0.6, 0.9]).
 * depth91 = 1 - |k| d and bis91 optional.  Returns the span. */
double orc_bisector(const double *v, const double *y, int n, double c, double k, double v0, double fwhm,
                    double *depth91, double *bis91) {
  int nstep = 100, margin = 5, len_depth = nstep - 2 * margin + 1;
  double sigma = fwhm / 2. / sqrt(2. * log(2.));
  double *norm_ccf = (double *)malloc(sizeof(double) * n);
  double *p0 = (double *)calloc((size_t)n, sizeof(double));
  double *p1 = (double *)calloc((size_t)n, sizeof(double));
  double *p2 = (double *)calloc((size_t)n, sizeof(double));
  double depth[91], bis[91];
  for (int i = 0; i < n; i++) norm_ccf[i] = -c / k * (1. - y[i] / c);
  for (int i = 0; i < len_depth; i++) depth[i] = (double)(i + margin) / nstep;
  for (int i = 0; i < n - 1; i++) {
    double vr = (v[i] + v[i + 1]) / 2.;
    double s2 = sigma * sigma;
    double dd = (vr - v0) * (vr - v0);
    double dccf = -(vr - v0) / s2 * exp(-dd / 2. / s2);
    double d2ccf = (dd / s2 - 1) / s2 * exp(-dd / 2. / s2);
    double d2rv = -d2ccf / (dccf * dccf * dccf);
    p2[i] = d2rv / 2.;
    p1[i] = (v[i + 1] - v[i] - p2[i] * (norm_ccf[i + 1] * norm_ccf[i + 1] - norm_ccf[i] * norm_ccf[i])) /
            (norm_ccf[i + 1] - norm_ccf[i]);
    p0[i] = v[i] - p1[i] * norm_ccf[i] - p2[i] * (norm_ccf[i] * norm_ccf[i]);
  }
  int ind_max = 0;
  for (int i = 0; i < n; i++)
    if (norm_ccf[i] > norm_ccf[ind_max]) ind_max = i;
  for (int i = 0; i < len_depth; i++) {
    int i_b = ind_max, i_r = ind_max;
    while ((norm_ccf[i_b] > depth[i]) && (i_b > 1)) i_b--;
    while ((norm_ccf[i_r + 1] > depth[i]) && (i_r < (n - 2))) i_r++;
    bis[i] = (p0[i_b] + p0[i_r]) + (p1[i_b] + p1[i_r]) * depth[i] + (p2[i_b] + p2[i_r]) * (depth[i] * depth[i]);
    bis[i] /= 2.;
  }
  int n1 = 0, n2 = 0;
  double rv_top = 0., rv_bottom = 0.;
  for (int i = 0; i < len_depth; i++) {
    if ((depth[i] >= 0.1) && (depth[i] <= 0.4)) {
      n1++;
      rv_top += bis[i];
    }
    if ((depth[i] >= 0.6) && (depth[i] <= 0.9)) {
      n2++;
      rv_bottom += bis[i];
    }
  }
  rv_top /= n1;
  rv_bottom /= n2;
  if (depth91)
    for (int i = 0; i < len_depth; i++) depth91[i] = 1.0 - fabs(k) * depth[i];
  if (bis91)
    for (int i = 0; i < len_depth; i++) bis91[i] = bis[i];
  free(norm_ccf);
  free(p0);
  free(p1);
  free(p2);
  return rv_top - rv_bottom;
}

/* BIS_FIT2 main, BIS_FIT2.cpp:97-320.  The GSL trust-region solver
 * (gsl_multifit_nlinear_trust, default parameters) is restated as a
 * Levenberg-Marquardt iteration with More scaling, Nielsen damping updates and
 * GSL's xtol/gtol tests (xtol = gtol = 1e-8, ftol = 0, max 10000 iterations).
 * par5 = [c, |k|, span, v0, fwhm]; model = fit at the final parameters;
 * depth91 = 1 - |k| depth; bis91 = bisector. */
int orc_fit_ccf_err(const double *v, const double *y, const double *err, int n, double *par5, double *model,
                    double *depth91, double *bis91, int *n_iter) {
  const double xtol = 1.0e-8, gtol = 1.0e-8;
  const int max_iter = 10000;
  double *px = (double *)malloc(sizeof(double) * n);
  double *f = (double *)malloc(sizeof(double) * n);
  double *ft = (double *)malloc(sizeof(double) * n);
  double (*J)[4] = (double(*)[4])malloc(sizeof(double) * 4 * n);
  for (int i = 0; i < n; i++) px[i] = v[i] - v[0];
  double c = (y[0] + y[n - 1]) / 2;
  double fwhm = (px[n - 1] - px[0]) / 5.;
  double k = 0., x0 = px[n / 2];
  for (int i = 0; i < n; i++)
    if (fabs(y[i] - c) > fabs(k)) {
      k = y[i] - c;
      x0 = px[i];
    }
  double x[4] = {c, k, x0, fwhm};
  orc_gauss(px, y, err, n, x, f, NULL, J);
  double D[4];
  for (int j = 0; j < 4; j++) {
    double s = 0;
    for (int i = 0; i < n; i++) s += J[i][j] * J[i][j];
    D[j] = sqrt(s) > 0 ? sqrt(s) : 1.0;
  }
  double mu = 1.0e-3, nu = 2.0;
  int iter = 0, status = 0, done = 0;
  for (iter = 0; iter < max_iter; iter++) {
    double A[4][4], g[4];
    for (int a = 0; a < 4; a++) {
      g[a] = 0;
      for (int i = 0; i < n; i++) g[a] += J[i][a] * f[i];
      for (int b = 0; b < 4; b++) {
        double s = 0;
        for (int i = 0; i < n; i++) s += J[i][a] * J[i][b];
        A[a][b] = s;
      }
    }
    double fnorm2 = 0;
    for (int i = 0; i < n; i++) fnorm2 += f[i] * f[i];
    double dx[4], xt[4];
    int bad = 0, accepted = 0;
    while (!accepted) {
      if (orc_solve4(A, D, mu, g, dx) != 0) {
        mu *= nu;
        nu *= 2;
        if (++bad > 15) break;
        continue;
      }
      for (int a = 0; a < 4; a++) xt[a] = x[a] + dx[a];
      orc_gauss(px, y, err, n, xt, ft, NULL, NULL);
      double ft2 = 0;
      for (int i = 0; i < n; i++) ft2 += ft[i] * ft[i];
      double pred = 0;
      for (int a = 0; a < 4; a++) {
        double Adx = 0;
        for (int b = 0; b < 4; b++) Adx += A[a][b] * dx[b];
        pred += dx[a] * Adx + 2 * mu * D[a] * D[a] * dx[a] * dx[a];
      }
      double rho = pred > 0 ? (fnorm2 - ft2) / pred : -1.0;
      if (rho > 0) {
        accepted = 1;
        double tt = 2 * rho - 1;
        double fac = 1 - tt * tt * tt;
        mu *= fac > 1. / 3. ? fac : 1. / 3.;
        nu = 2.0;
      } else {
        mu *= nu;
        nu *= 2;
        if (++bad > 15) break;
      }
    }
    if (!accepted) {
      status = 27; /* GSL_ENOPROG */
      break;
    }
    for (int a = 0; a < 4; a++) x[a] = xt[a];
    orc_gauss(px, y, err, n, x, f, NULL, J);
    for (int j = 0; j < 4; j++) {
      double s = 0;
      for (int i = 0; i < n; i++) s += J[i][j] * J[i][j];
      if (sqrt(s) > D[j]) D[j] = sqrt(s);
    }
    int small = 1;
    for (int a = 0; a < 4; a++)
      if (!(fabs(dx[a]) < xtol * xtol + xtol * fabs(x[a]))) small = 0;
    if (small) {
      iter++;
      done = 1;
      break;
    }
    double gn = 0, f2 = 0;
    for (int i = 0; i < n; i++) f2 += f[i] * f[i];
    for (int a = 0; a < 4; a++) {
      double ga = 0;
      for (int i = 0; i < n; i++) ga += J[i][a] * f[i];
      double sc = fabs(x[a]) > 1.0 ? fabs(x[a]) : 1.0;
      if (fabs(ga * sc) > gn) gn = fabs(ga * sc);
    }
    double phi = 0.5 * f2;
    if (gn <= gtol * (phi > 1.0 ? phi : 1.0)) {
      iter++;
      done = 1;
      break;
    }
  }
  if (status == 0 && !done) status = 11; /* GSL_EMAXITER */
  if (n_iter) *n_iter = iter;
  c = x[0];
  k = x[1];
  x0 = x[2];
  fwhm = x[3];
  if (model) orc_gauss(px, y, err, n, x, NULL, model, NULL);

  double span = orc_bisector(v, y, n, c, k, x0 + v[0], fwhm, depth91, bis91);
  par5[0] = c;
  par5[1] = fabs(k);
  par5[2] = span;
  par5[3] = x0 + v[0];
  par5[4] = fwhm;
  free(px);
  free(f);
  free(ft);
  free(J);
  return status;
}

int orc_fit_ccf(const double *v, const double *y, int n, double *par5, double *model, double *depth91,
                double *bis91, int *n_iter) {
  return orc_fit_ccf_err(v, y, NULL, n, par5, model, depth91, bis91, n_iter);
}

/* ----------------------------------- a11: differentiable model (no reference counterpart) */

/* Exact gradient and Hessian of 0.5 sum r^2 for the Gaussian model of BIS_FIT2.cpp:31-81,
 * f = c + k exp(-(x-x0)^2 / (2 s^2)), s = fwhm / (2 sqrt(2 ln 2)), theta = (c, k, x0, fwhm),
 * over absolute velocities x.  J (n x 4) optional. */
void orc_gauss_hess(const double *x, const double *y, int n, const double *th, double *g4, double H[4][4],
                    double *J, double *res_sum2) {
  double c = th[0], k = th[1], x0 = th[2], fw = th[3];
  double s = fw / (2 * sqrt(2 * log(2.0)));
  double r2 = 0;
  for (int a = 0; a < 4; a++) {
    g4[a] = 0;
    for (int b = 0; b < 4; b++) H[a][b] = 0;
  }
  for (int i = 0; i < n; i++) {
    double u = (x[i] - x0) / s, e = exp(-u * u / 2);
    double f = c + k * e, r = f - y[i];
    double j4[4] = {1.0, e, k * e * u / s, k * e * u * u / fw};
    double h[4][4] = {{0}};
    h[1][2] = h[2][1] = e * u / s;
    h[1][3] = h[3][1] = e * u * u / fw;
    h[2][2] = k * e * (u * u - 1) / (s * s);
    h[2][3] = h[3][2] = k * e * u * (u * u - 2) / (s * fw);
    h[3][3] = k * e * u * u * (u * u - 3) / (fw * fw);
    for (int a = 0; a < 4; a++) {
      g4[a] += j4[a] * r;
      for (int b = 0; b < 4; b++) H[a][b] += j4[a] * j4[b] + r * h[a][b];
      if (J) J[i * 4 + a] = j4[a];
    }
    r2 += r * r;
  }
  if (res_sum2) *res_sum2 = r2;
}

/* 4x4 solve by Gaussian elimination with partial pivoting; returns 0 on success */
int orc_solve4_dense(const double A_in[4][4], const double *b, double *x) {
  double A[4][5];
  for (int i = 0; i < 4; i++) {
    for (int j = 0; j < 4; j++) A[i][j] = A_in[i][j];
    A[i][4] = b[i];
  }
  for (int col = 0; col < 4; col++) {
    int piv = col;
    for (int r = col + 1; r < 4; r++)
      if (fabs(A[r][col]) > fabs(A[piv][col])) piv = r;
    if (A[piv][col] == 0.0) return -1;
    if (piv != col)
      for (int j = 0; j < 5; j++) {
        double t = A[col][j];
        A[col][j] = A[piv][j];
        A[piv][j] = t;
      }
    for (int r = col + 1; r < 4; r++) {
      double f = A[r][col] / A[col][col];
      for (int j = col; j < 5; j++) A[r][j] -= f * A[col][j];
    }
  }
  for (int i = 3; i >= 0; i--) {
    double s = A[i][4];
    for (int j = i + 1; j < 4; j++) s -= A[i][j] * x[j];
    x[i] = s / A[i][i];
  }
  return 0;
}

/* Newton steps on the exact Hessian from theta (damped when a step does not
 * reduce the residual), then two undamped Newton steps on the gradient, which
 * fix the optimum to rounding (a residual test alone leaves it loose by
 * ~sqrt(eps) in flat directions). */
static void orc_gauss_newton2(const double *x, const double *y, int n, double *th) {
  for (int it = 0; it < 2; it++) {
    double g[4], H[4][4], r2, mg[4], dx[4];
    orc_gauss_hess(x, y, n, th, g, H, NULL, &r2);
    for (int a = 0; a < 4; a++) mg[a] = -g[a];
    if (orc_solve4_dense(H, mg, dx) != 0) return;
    for (int a = 0; a < 4; a++) th[a] += dx[a];
  }
}

int orc_gauss_polish(const double *x, const double *y, int n, double *th) {
  double g[4], H[4][4], r2;
  orc_gauss_hess(x, y, n, th, g, H, NULL, &r2);
  for (int it = 0; it < 60; it++) {
    double mu = 0.0, dx[4], tn[4], r2n = r2;
    int ok = 0;
    for (int tries = 0; tries < 30; tries++) {
      double Hd[4][4], mg[4];
      for (int a = 0; a < 4; a++) {
        mg[a] = -g[a];
        for (int b = 0; b < 4; b++) Hd[a][b] = H[a][b] + (a == b ? mu * (fabs(H[a][a]) + 1e-300) : 0.0);
      }
      if (orc_solve4_dense(Hd, mg, dx) == 0) {
        for (int a = 0; a < 4; a++) tn[a] = th[a] + dx[a];
        double gn[4], Hn[4][4];
        orc_gauss_hess(x, y, n, tn, gn, Hn, NULL, &r2n);
        if (r2n <= r2 * (1 + 1e-15)) {
          ok = 1;
          break;
        }
      }
      mu = mu == 0.0 ? 1e-6 : mu * 10;
    }
    if (!ok) {
      orc_gauss_newton2(x, y, n, th);
      return it;
    }
    int small = 1;
    for (int a = 0; a < 4; a++)
      if (fabs(dx[a]) > 1e-12 * fabs(th[a]) + 1e-300) small = 0;
    for (int a = 0; a < 4; a++) th[a] = tn[a];
    orc_gauss_hess(x, y, n, th, g, H, NULL, &r2);
    if (small) {
      orc_gauss_newton2(x, y, n, th);
      return it + 1;
    }
  }
  orc_gauss_newton2(x, y, n, th);
  return 60;
}

/* Forward model of the differentiable path (the adjoint's reference):
 *   cells of the G x G disk grid on the disk (y^2 + z^2 <= 1), y = -1 + 2 iy/G,
 *   z = -1 + 2 iz/G, mu = sqrt(1 - y^2 - z^2), mu bin as orc_disk_init,
 *   limb L = 1 - u1 (1-mu) - u2 (1-mu)^2, rigid rotation delta = vsini y [km/s];
 *   filling factors phi_pc (spots) and psi_pc (faculae) = bilinear samples of map p
 *   and facula map p (M x M over [-1, 1]^2; psi = 0 without facula maps);
 *   I_p = sum_c [(1 - phi - psi) Shift(L Q_k) + kappa phi Shift(L S_k)
 *                + kappa_f (1 + LB(mu)) psi Shift(L Fa_k)]
 *   with LB(mu) = lb0 + lb1 mu + lb2 mu^2 + lb3 mu^3 and (1 + LB) the facula intensity
 *   factor of SOAP_initialize.cu:367 / SOAP_integration.cu:317-323, and the log-grid
 *   shift of orc_shift_value (SOAP_integration.cu:232-281);
 *   F_p = conv(I_p) with the log-grid taps of orc_lower_resolution (no spec_norm),
 *   zero outside wid < i < n-1-wid;
 *   CCF_p = raw mask CCF of F_p[trim, n-trim) (orc_ccf, CCF_func.py:5-91);
 *   theta_p = (c, k, x0, fwhm) of the Gaussian fit (orc_fit_ccf) polished to the
 *   optimum (orc_gauss_polish); RV_p = x0, FWHM_p = fwhm; BIS_p = the bisector span of
 *   the CCF at that optimum (orc_bisector);
 *   loss = w_spec sum_p sum_{trim <= j < n-trim} (F_p - T_p)^2
 *        + w_rv sum_p (RV_p - rv_t_p)^2 + w_fw sum_p (FWHM_p - fw_t_p)^2
 *        + w_bis sum_p (BIS_p - bis_t_p)^2   (the last term only when w_bis != 0).
 * Outputs F [P][n], ccf [P][n_v], theta [P][4], span [P] (each optional) and loss. */
int orc_diff_forward(const double *wave, long n, int grid, int n_mu, const double *mu_bounds, double vsini,
                     double u1, double u2, double kappa, double kappa_f, const double *lb, const double *cube_q,
                     const double *cube_s, const double *cube_f, int n_phase, int map_res, const double *maps,
                     const double *fmaps, double sigma_ms, int trim, const double *vgrid, int n_v,
                     const double *line_l, const double *line_w, long n_lines, const double *target,
                     const double *rv_t, const double *fw_t, const double *bis_t, double w_spec, double w_rv,
                     double w_fw, double w_bis, double *F_out, double *ccf_out, double *theta_out, double *span_out,
                     double *loss_out, int nthreads) {
  const double pi = ORC_PI;
  long ncell = (long)grid * grid;
  int M = map_res;
  /* cells */
  long non = 0;
  for (long idx = 0; idx < ncell; idx++) {
    int iy = (int)(idx / grid), iz = (int)(idx % grid);
    double y = -1. + iy * (2. / grid), z = -1. + iz * (2. / grid);
    if (y * y + z * z <= 1.) non++;
  }
  long *cidx = (long *)malloc(sizeof(long) * (non > 0 ? non : 1));
  int *ck = (int *)malloc(sizeof(int) * (non > 0 ? non : 1));
  double *cL = (double *)malloc(sizeof(double) * (non > 0 ? non : 1));
  double *cy = (double *)malloc(sizeof(double) * (non > 0 ? non : 1));
  double *cz = (double *)malloc(sizeof(double) * (non > 0 ? non : 1));
  double *clb = (double *)malloc(sizeof(double) * (non > 0 ? non : 1));
  orc_shift *csh = (orc_shift *)malloc(sizeof(orc_shift) * (non > 0 ? non : 1));
  non = 0;
  for (long idx = 0; idx < ncell; idx++) {
    int iy = (int)(idx / grid), iz = (int)(idx % grid);
    double y = -1. + iy * (2. / grid), z = -1. + iz * (2. / grid);
    if (!(y * y + z * z <= 1.)) continue;
    double mu = sqrt(1. - (y * y + z * z));
    int k = n_mu - 1;
    for (int b = 0; b < n_mu - 1; b++)
      if (mu > mu_bounds[b]) {
        k = b;
        break;
      }
    cidx[non] = idx;
    ck[non] = k;
    cL[non] = 1. - u1 * (1 - mu) - u2 * (1 - mu) * (1 - mu);
    cy[non] = y;
    cz[non] = z;
    clb[non] = lb ? 1. + (lb[0] + lb[1] * mu + lb[2] * mu * mu + lb[3] * mu * mu * mu) : 1.;
    csh[non] = orc_shift_setup(ORC_GRID_LOG, wave, n, 0.0, vsini * y);
    non++;
  }
  int wid = orc_conv_window(wave, sigma_ms);
  double h = orc_grid_h(wave, n);
  double fs = orc_delta_lambda(sigma_ms, 1.0);
  double *ctap = (double *)malloc(sizeof(double) * (2 * (size_t)wid + 1));
  for (int d = -wid; d <= wid; d++) {
    double u = expm1(d * h) / fs;
    ctap[d + wid] = exp(d * h) * expm1(h) / (fs * sqrt(2. * pi)) * exp(-u * u / 2);
  }
  long n_t = n - 2 * (long)trim;
  double loss = 0.0;
  double *I = (double *)malloc(sizeof(double) * n);
  double *F = (double *)malloc(sizeof(double) * n);
  double *cc = (double *)malloc(sizeof(double) * n_v);
  double *phi = (double *)malloc(sizeof(double) * (non > 0 ? non : 1));
  double *psi = (double *)malloc(sizeof(double) * (non > 0 ? non : 1));
  const int fac = cube_f && fmaps;
  for (int p = 0; p < n_phase; p++) {
    const double *mp = maps + (size_t)p * M * M;
    const double *fp = fac ? fmaps + (size_t)p * M * M : NULL;
    for (long c = 0; c < non; c++) {
      double gy = (cy[c] + 1) * 0.5 * (M - 1), gz = (cz[c] + 1) * 0.5 * (M - 1);
      int i0 = (int)floor(gy), j0 = (int)floor(gz);
      if (i0 > M - 2) i0 = M - 2;
      if (j0 > M - 2) j0 = M - 2;
      if (i0 < 0) i0 = 0;
      if (j0 < 0) j0 = 0;
      double fy = gy - i0, fz = gz - j0;
      phi[c] = (1 - fy) * (1 - fz) * mp[i0 * M + j0] + fy * (1 - fz) * mp[(i0 + 1) * M + j0] +
               (1 - fy) * fz * mp[i0 * M + j0 + 1] + fy * fz * mp[(i0 + 1) * M + j0 + 1];
      psi[c] = fac ? (1 - fy) * (1 - fz) * fp[i0 * M + j0] + fy * (1 - fz) * fp[(i0 + 1) * M + j0] +
                         (1 - fy) * fz * fp[i0 * M + j0 + 1] + fy * fz * fp[(i0 + 1) * M + j0 + 1]
                   : 0.0;
    }
    int nchunk = nthreads > 0 ? nthreads * 4 : 32;
#pragma omp parallel for schedule(dynamic) num_threads(nthreads > 0 ? nthreads : 1)
    for (int ch = 0; ch < nchunk; ch++) {
      long j0, j1;
      orc_chunk(n, nchunk, ch, &j0, &j1);
      for (long j = j0; j < j1; j++) I[j] = 0.0;
      for (long c = 0; c < non; c++) {
        const double *rq = cube_q + (long)ck[c] * n, *rs = cube_s + (long)ck[c] * n;
        double a = 1 - phi[c] - psi[c], b = kappa * phi[c];
        if (fac) {
          const double *rf = cube_f + (long)ck[c] * n;
          double e = kappa_f * clb[c] * psi[c];
          for (long j = j0; j < j1; j++)
            I[j] += a * orc_shift_value(&csh[c], wave, n, rq, j, cL[c]) +
                    b * orc_shift_value(&csh[c], wave, n, rs, j, cL[c]) +
                    e * orc_shift_value(&csh[c], wave, n, rf, j, cL[c]);
        } else {
          for (long j = j0; j < j1; j++)
            I[j] += a * orc_shift_value(&csh[c], wave, n, rq, j, cL[c]) +
                    b * orc_shift_value(&csh[c], wave, n, rs, j, cL[c]);
        }
      }
    }
#pragma omp parallel for schedule(static) num_threads(nthreads > 0 ? nthreads : 1)
    for (long i = 0; i < n; i++) {
      double acc = 0.0;
      if ((i > wid) && (i + wid < n - 1))
        for (long m = i - wid; m < i + wid - 1; m++) acc += ctap[m - i + wid] * I[m];
      F[i] = acc;
    }
    if (F_out) memcpy(F_out + (size_t)p * n, F, sizeof(double) * n);
    for (long j = trim; j < n - trim; j++) {
      double r = F[j] - target[(size_t)p * n + j];
      loss += w_spec * r * r;
    }
    orc_ccf(ORC_GRID_LOG, wave + trim, n_t, h, vgrid, n_v, line_l, line_w, n_lines, F + trim, cc, nthreads);
    if (ccf_out) memcpy(ccf_out + (size_t)p * n_v, cc, sizeof(double) * n_v);
    double par[5];
    orc_fit_ccf(vgrid, cc, n_v, par, NULL, NULL, NULL, NULL);
    int inear = 0;
    for (int i = 1; i < n_v; i++)
      if (fabs(vgrid[i] - par[3]) < fabs(vgrid[inear] - par[3])) inear = i;
    double th[4] = {par[0], cc[inear] < par[0] ? -par[1] : par[1], par[3], par[4]};
    orc_gauss_polish(vgrid, cc, n_v, th);
    if (theta_out) memcpy(theta_out + (size_t)p * 4, th, sizeof(th));
    loss += w_rv * (th[2] - rv_t[p]) * (th[2] - rv_t[p]) + w_fw * (th[3] - fw_t[p]) * (th[3] - fw_t[p]);
    double span = orc_bisector(vgrid, cc, n_v, th[0], th[1], th[2], th[3], NULL, NULL);
    if (span_out) span_out[p] = span;
    if (w_bis != 0.0) loss += w_bis * (span - bis_t[p]) * (span - bis_t[p]);
  }
  *loss_out = loss;
  free(cidx);
  free(ck);
  free(cL);
  free(cy);
  free(cz);
  free(clb);
  free(csh);
  free(ctap);
  free(I);
  free(F);
  free(cc);
  free(phi);
  free(psi);
  return 0;
}
