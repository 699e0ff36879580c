/*
 * soap_oracle.h -- FP64 CPU restatement of the SOAP-GPU reference hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker for the HIP
 * engine in soap_gpu_amd/.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it; the engine never calls it.
 *
 * Every function cites the reference file:line it restates (paths relative to
 * the reference repository root).  Grid kinds:
 *   ORC_GRID_LINEAR (0): the reference's own format, constant step; formulas are
 *                        the reference's literal ones (hard-coded 0.005 A where
 *                        the reference hard-codes it).
 *   ORC_GRID_LOG    (1): log-uniform grid lambda_j = lambda_0 exp(j h); the
 *                        reference's index arithmetic generalised through the
 *                        exact inverse grid map (see each function).
 * Out-of-range source pixels read as 0 (zero padding) in every kind.
 */
#ifndef SOAP_ORACLE_H
#define SOAP_ORACLE_H
#ifdef __cplusplus
extern "C" {
#endif

#define ORC_GRID_LINEAR 0
#define ORC_GRID_LOG 1

typedef struct {
  double lat, lon, size, type; /* one line of maps{id}_for_all.txt */
} orc_region;

/* scalar helpers */
double orc_rndup(double n, int nb_decimal);
double orc_delta_lambda(double v_ms, double lambda0);
double orc_delta_lambda_cpu(double v_ms, double lambda0);
double orc_grid_h(const double *wave, long n);
int orc_lb_params(int t_star, const double *lb_g2, const double *lb_k0, const double *lb_m0,
                  const double *temps3, double *out4);

/* a1: disk grid maps */
void orc_disk_init(int grid, double prot, double w_ratio, double incl_deg, double radius_km, double u1,
                   double u2, const double *lb4, const double *mu_bounds, int n_mu, double *mu_map,
                   double *delta_map, double *limb_map, int *region_map, double *lb_map);

/* a4: one Doppler-shifted, limb-weighted seed row (gather definition) */
void orc_shift_row(int kind, const double *wave, long n, double step, const double *row,
                   double delta_kms, double limb, double *out);
/* a4 'ref' mode: literal serial scatter of one cell into a carried buffer (linear grids) */
void orc_scatter_row_ref(const double *wave, long n, double step, const double *row,
                         double delta_kms, double limb, double *buf);

/* a2: quiet star */
void orc_quiet_star(int kind, const double *wave, long n, double step, int grid,
                    const double *delta_map, const double *limb_map, const int *region_map,
                    const double *cube, int n_mu, double *out, int nthreads);
void orc_quiet_star_ref(const double *wave, long n, double step, int grid, const double *delta_map,
                        const double *limb_map, const int *region_map, const double *cube, int n_mu,
                        double *out);

/* a5: region geometry and scan */
int orc_region_bbox(double lat, double lon, double size, double incl_deg, int nrho, int grid,
                    double phase, int *bnd5);
void orc_phase_maps(int grid, double incl_deg, int nrho, int t_star, int t_diff_spot,
                    int t_diff_fac, double phase, const orc_region *regs, int n_reg,
                    double *temp_map, double *bi_map);

/* a6: active-region deltas for one phase */
void orc_integrate_phase(int kind, const double *wave, long n, double step, int grid,
                         const double *delta_map, const double *limb_map, const int *region_map,
                         const double *lb_map, const double *temp_map, const double *bi_map,
                         const double *cube_q, const double *cube_qsed, const double *cube_sp,
                         const double *cube_spsed, const double *cube_fac, const double *cube_facsed,
                         int n_mu, int t_star, int t_diff_fac, int solar_mode, double *flux,
                         double *bconv, double *tot, int nthreads);

/* a7: instrumental convolution + normalisation */
int orc_conv_window(const double *wave, double sigma_ms);
void orc_conv_raw(int kind, const double *wave, long n, double sigma_ms, const double *A, int variant,
                  double weight, double *o, int nthreads);
void orc_lower_resolution(int kind, const double *wave, long n, double sigma_ms,
                          const double *quiet, const double *spec, int variant, double weight,
                          double *out, double *minmax2, int nthreads);

/* a8: mask CCF on a trimmed spectrum */
long orc_select_lines(const double *wave_t, long n_t, const double *line_l, const double *line_w,
                      long n_lines, double *sel_l, double *sel_w);
void orc_ccf(int kind, const double *wave_t, long n_t, double step, const double *vgrid, int n_v,
             const double *line_l, const double *line_w, long n_lines, const double *spec_t,
             double *ccf, int nthreads);
void orc_appodisation(int n_v, double *appod);
void orc_ccf_finish(int n_v, const double *appod, double *ccf);

/* a9: Gaussian fit + bisector (BIS_FIT2) */
int orc_fit_ccf(const double *v, const double *y, int n, double *par5, double *model, double *depth91,
                double *bis91, int *n_iter);
/* the same with per-point errors err[n] (CCF_error.txt), NULL = unit errors */
int orc_fit_ccf_err(const double *v, const double *y, const double *err, int n, double *par5, double *model,
                    double *depth91, double *bis91, int *n_iter);

/* a11: differentiable model (no reference counterpart): Gaussian-fit Hessian,
 * Newton polish, forward model and loss (definitions in soap_oracle.c) */
void orc_gauss_hess(const double *x, const double *y, int n, const double *th, double *g4, double H[4][4],
                    double *J, double *res_sum2);
int orc_solve4_dense(const double A_in[4][4], const double *b, double *x);
int orc_gauss_polish(const double *x, const double *y, int n, double *th);
double orc_bisector(const double *v, const double *y, int n, double c, double k, double v0, double fwhm,
                    double *depth91, double *bis91);
int orc_diff_forward(const double *wave, long n, int grid, int n_mu, const double *mu_bounds, double vsini,
                     double u1, double u2, double kappa, double kappa_f, const double *lb, const double *cube_q,
                     const double *cube_s, const double *cube_f, int n_phase, int map_res, const double *maps,
                     const double *fmaps, double sigma_ms, int trim, const double *vgrid, int n_v,
                     const double *line_l, const double *line_w, long n_lines, const double *target,
                     const double *rv_t, const double *fw_t, const double *bis_t, double w_spec, double w_rv,
                     double w_fw, double w_bis, double *F_out, double *ccf_out, double *theta_out, double *span_out,
                     double *loss_out, int nthreads);

#ifdef __cplusplus
}
#endif
#endif
