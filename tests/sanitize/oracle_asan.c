/* Driver for the oracle built under AddressSanitizer and UndefinedBehaviorSanitizer
 * (tests/test_sanitize.py): every restated stage on small synthetic inputs, log and
 * linear grids, spot / facula / solar modes, the mask CCF, the fit with errors and
 * the differentiable model.  Exit status 0 and no sanitizer report is the pass. */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include "soap_oracle.h"

static double *dalloc(long n) {
  double *p = (double *)calloc((size_t)(n > 0 ? n : 1), sizeof(double));
  if (!p) exit(3);
  return p;
}

/* n_mu rows of absorption lines (depth varying with the row) */
static double *cube(const double *wave, long n, int n_mu, double scale, double shift_a) {
  double *c = dalloc((long)n_mu * n);
  for (int k = 0; k < n_mu; k++)
    for (long i = 0; i < n; i++) {
      double v = 1.0;
      for (int l = 0; l < 6; l++) {
        const double c0 = wave[n / 8] + l * (wave[n - n / 8] - wave[n / 8]) / 6.0 + shift_a;
        v -= (0.3 + 0.05 * k + 0.02 * l) * exp(-0.5 * pow((wave[i] - c0) / 0.03, 2));
      }
      c[(long)k * n + i] = scale * v;
    }
  return c;
}

static double checksum(const double *x, long n) {
  double s = 0;
  for (long i = 0; i < n; i++) s += fabs(x[i]);
  return s;
}

int main(void) {
  const int grid = 20, n_mu = 4, nrho = 20, trim = 60, n_v = 41;
  const long n = 900, g2 = (long)grid * grid;
  const double mu_b[3] = {0.75, 0.5, 0.25};
  const double lb4[4] = {0.23161761, -0.11674743, -0.36866954, 0.25537337};
  double total = 0;
  double *wl = dalloc(n), *wlin = dalloc(n);
  for (long i = 0; i < n; i++) {
    wl[i] = 5000.0 * exp(i * 1.0e-6);
    wlin[i] = 5000.0 + 0.005 * i;
  }
  double *mu = dalloc(g2), *delta = dalloc(g2), *limb = dalloc(g2), *lbm = dalloc(g2);
  int *region = (int *)calloc((size_t)g2, sizeof(int));
  orc_disk_init(grid, 25.05, 0.1, 80.0, 696000.0, 0.29, 0.34, lb4, mu_b, n_mu, mu, delta, limb, region, lbm);
  orc_region regs[3] = {{20.0, 10.0, 0.25, 1.0}, {-15.0, 40.0, 0.15, 0.0}, {5.0, 200.0, 0.2, 1.0}};
  double *temp = dalloc(g2), *bi = dalloc(g2);
  orc_phase_maps(grid, 80.0, nrho, 5778, 663, 250, 0.05, regs, 3, temp, bi);
  double vgrid[41], appod[41], err[41], par[5], depth[91], bis[91], model[41];
  for (int v = 0; v < n_v; v++) {
    vgrid[v] = -10000.0 + 500.0 * v;
    err[v] = 1e-3 * (1.0 + 0.1 * (v % 3));
  }
  orc_appodisation(n_v, appod);
  double line_l[8], line_w[8], sel_l[8], sel_w[8];
  for (int kind = 0; kind < 2; kind++) {
    const double *wave = kind == ORC_GRID_LOG ? wl : wlin;
    const double step = kind == ORC_GRID_LOG ? 0.0 : 0.005;
    double *q = cube(wave, n, n_mu, 1.0, 0.0), *qs = cube(wave, n, n_mu, 2.0, 0.0);
    double *sp = cube(wave, n, n_mu, 0.4, 0.01), *sps = cube(wave, n, n_mu, 0.8, 0.01);
    double *fa = cube(wave, n, n_mu, 1.1, -0.01), *fas = cube(wave, n, n_mu, 2.2, -0.01);
    double *quiet = dalloc(n), *flux = dalloc(n), *bconv = dalloc(n), *tot = dalloc(n), *low = dalloc(n);
    double mm[2], ccf[41];
    orc_quiet_star(kind, wave, n, step, grid, delta, limb, region, q, n_mu, quiet, 1);
    for (int solar = 0; solar <= kind; solar++) {
      orc_integrate_phase(kind, wave, n, step, grid, delta, limb, region, lbm, temp, bi, q, qs, sp, sps, fa, fas, n_mu,
                          5778, 250, solar, flux, bconv, tot, 1);
      total += checksum(flux, n) + checksum(bconv, n) + checksum(tot, n);
    }
    orc_lower_resolution(kind, wave, n, 1107.0, quiet, flux, 1, 0.005, low, mm, 1);
    orc_lower_resolution(kind, wave, n, 1107.0, quiet, quiet, 0, 0.005, low, mm, 1);
    for (int l = 0; l < 8; l++) {
      line_l[l] = wave[trim + 40] + l * (wave[n - trim - 40] - wave[trim + 40]) / 7.0;
      line_w[l] = l == 3 ? 0.05 : 0.5 + 0.1 * l; /* one line below the weight cut */
    }
    const long k = orc_select_lines(wave + trim, n - 2 * trim, line_l, line_w, 8, sel_l, sel_w);
    /* the log form's CCF takes the grid's h as its step */
    orc_ccf(kind, wave + trim, n - 2 * trim, kind == ORC_GRID_LOG ? 1.0e-6 : step, vgrid, n_v, sel_l, sel_w, k,
            low + trim, ccf, 1);
    orc_ccf_finish(n_v, appod, ccf);
    int it = 0;
    orc_fit_ccf_err(vgrid, ccf, err, n_v, par, model, depth, bis, &it);
    orc_fit_ccf(vgrid, ccf, n_v, par, model, depth, bis, &it);
    total += checksum(quiet, n) + checksum(low, n) + checksum(ccf, n_v) + checksum(par, 5);
    free(q), free(qs), free(sp), free(sps), free(fa), free(fas);
    free(quiet), free(flux), free(bconv), free(tot), free(low);
  }
  { /* differentiable model with both populations */
    const int P = 2, M = 6;
    double *q = cube(wl, n, n_mu, 1.0, 0.0), *s = cube(wl, n, n_mu, 0.5, 0.01), *f = cube(wl, n, n_mu, 1.2, -0.01);
    double *maps = dalloc((long)P * M * M), *fmaps = dalloc((long)P * M * M), *target = dalloc((long)P * n);
    for (long i = 0; i < (long)P * M * M; i++) {
      maps[i] = (i % 7) / 7.0;
      fmaps[i] = (i % 5) / 10.0;
    }
    double rv_t[2] = {10.0, -5.0}, fw_t[2] = {7000.0, 7100.0}, bis_t[2] = {0.0, 1.0};
    double *F = dalloc((long)P * n), *ccf = dalloc((long)P * n_v), theta[8], span[2], loss = 0;
    for (int l = 0; l < 8; l++) {
      line_l[l] = wl[trim + 40] + l * (wl[n - trim - 40] - wl[trim + 40]) / 7.0;
      line_w[l] = 0.6;
    }
    orc_diff_forward(wl, n, 16, n_mu, mu_b, 3.0, 0.29, 0.34, 0.3, 1.15, lb4, q, s, f, P, M, maps, fmaps, 1107.0, trim,
                     vgrid, n_v, line_l, line_w, 8, target, rv_t, fw_t, bis_t, 1e-3, 1e-3, 1e-5, 1e-3, F, ccf, theta,
                     span, &loss, 1);
    total += checksum(F, (long)P * n) + fabs(loss);
    free(q), free(s), free(f), free(maps), free(fmaps), free(target), free(F), free(ccf);
  }
  free(wl), free(wlin), free(mu), free(delta), free(limb), free(lbm), free(region), free(temp), free(bi);
  printf("oracle under sanitizers: ok (checksum %.6g)\n", total);
  return isfinite(total) ? 0 : 2;
}
