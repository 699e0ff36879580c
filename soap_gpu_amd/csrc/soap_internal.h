// Internal declarations shared by the host engine (soap_engine.hip) and the
// gfx950 kernels (soap_kernels.hip).  Not part of the public ABI.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace soap {

constexpr double kC = 299792458.;
constexpr double kPi = 3.141592653589793238463;
constexpr int kMaxSolarNodes = 12;
// loi_Planck constants (SOAP_integration.cu:205-212)
constexpr double kPlanckH = 6.62606896e-34;
constexpr double kBoltzK = 1.380e-23;

// read-only for the whole kernel (the constant address space): uniform loads from it
// go through the scalar unit even where the pointer came from memory and the kernel
// stores elsewhere through pointers that could alias it
#define CONSTANT __attribute__((address_space(4)))
template <class T>
__device__ __forceinline__ const CONSTANT T *gc(const T *p) {
  return (const CONSTANT T *)(uintptr_t)p;
}

// Region geometry of one (phase, region), evaluated on the host once per phase
// (spot_init_gpu/spot_phase_gpu/spot_area + the first row of spot_inverse_rotation).
struct RegionDev {
  double r0, r1, r2;  // row 0 of the inverse rotation R3 = R2 R
  double oms2;        // 1. - s*s
  int32_t y0, y1, z0, z1;  // scan box [y0,y1) x [z0,z1), clamped to the grid
  int32_t type;       // 0 spot, 1 facula
  int32_t pad;
};

struct DiskArgs {
  int32_t grid, n_mu;
  double delta_grid, theta_grid, omega0, omega1, radius, sin_incl, u1, u2;
  double lb0, lb1, lb2, lb3;
  const double *mu_bounds;  // n_mu - 1
  int32_t log_grid;
  double h, expm1_h;
  double *mu, *delta, *limb, *lbm;
  int32_t *region;
  int32_t *m;  // per-cell integer pixel shift (log grids)
  double *t;   // per-cell interpolation fraction (log grids)
};

// Fixed-point tap accumulation (exact, order independent): value =
// hi * 2^-32 + lo * 2^-72, hi signed, lo unsigned.
struct TapAccArgs {
  int32_t grid, n_mu, dmin, nd, nsets, wd, cols_per_block;
  const int32_t *region;
  const double *limb, *delta, *lbm;
  const int32_t *m;
  const double *t;
  unsigned long long *acc;  // [job][nsets][n_mu][nd][2]
  // active mode
  const RegionDev *geo;
  const int32_t *geo_off;   // [n_job + 1]
  const int32_t *box;       // [n_job][4]: cy0, cy1, cz0, cz1
  // solar mode (facula intensity by Planck ratio): facula cells also add
  // limb * l_q(T_cell) to set 2+q, l_q the Lagrange basis on node_t
  int32_t solar_nodes;      // K, 0 = off
  int32_t t_sum;            // t_star + t_diff_fac
  double node_t[kMaxSolarNodes], node_w[kMaxSolarNodes];  // nodes, barycentric weights
};

struct FirArgs {
  const double *q, *qs, *sp, *sps, *fa, *fas;  // cube bases [n_mu][n]
  int64_t n;
  int32_t n_mu, dmin, nd;
  const double *taps;    // quiet: [n_mu][nd]; active: [job][n_mu][nd][4] (sets contiguous)
  const double *taps_g;  // active, phase-interleaved: [n_mu][nd][p_pad][4], jobs >= n_job zero
  int32_t p_pad;         // job stride of taps_g (>= n_job + 8)
  const int32_t *krange; // [job][n_mu][2] inclusive d-index range, lo > hi: empty
  const int32_t *kflags; // [job][n_mu]: bit0 spot sets nonzero, bit1 facula sets nonzero
  const int32_t *bins;   // [job][n_mu]
  const int32_t *nbins;  // [job]
  double *out;           // quiet: [2][n]; active: [job][3][n]
};

// Entry streams of the tile-stationary FIR (built on the device per launch).
struct FirProg {
  double *taps;    // [n_groups][cap][16]: the group's 4 phases x (H1s, H1f, H2f, H1a)
  int32_t *xoff;   // [n_groups][cap]: LDS byte offset of lane 0's source pixel
  int32_t *count;  // [n_groups]: entries in use, padded with zero entries to the consumer's step
  int32_t *ufl;    // [n_mu]: bin flags over every phase (bit 0 spot, 1 facula, 2 used)
  int32_t cap;     // entries per group, >= n_mu * nd + 6
};

// Solar-mode epilogue: the Planck ratio B(lambda, T)/B(lambda, t_norm) at the
// K temperature nodes multiplies the Lagrange-set FIR sums.
struct SolarFir {
  const double *wave;  // [n]
  int32_t k;           // number of nodes
  int32_t t_norm;
  double node_t[kMaxSolarNodes];
};

struct ConvArgs {
  int64_t n;
  int32_t wid;           // reference window half width (nb_pixel)
  int32_t dlo;           // first tap offset
  int32_t ntp;           // taps used (zero padded to a multiple of conv_chunk())
  const double *ctap;    // [ntp]: c_{dlo + k}
  int32_t n_tiles;       // ceil(n / conv_tile())
};

// ------------------------------------------------------------ device helpers
// ------------------------------------------------------------ fixed point taps
// A contribution x is split into hi = floor(x 2^32) and lo = (x 2^32 - hi) 2^32
// (exact: power-of-two scalings and the subtraction of an integer part do not
// round); sums of 64-bit integers are associative, so atomics give the exact,
// order-independent sum of all contributions truncated at 2^-64 each.
// Capacity: lo < 2^32 per contribution, so the lo word holds 2^32 contributions
// to one tap before it could wrap (a disk of 8192^2 cells has 5.3e7); the hi word
// holds |sum| < 2^63 / 2^32 = 2^31, and the host keeps the summed weights of one
// tap below 2^30 (fx_capacity_ok).
__device__ __forceinline__ void fx_split(double x, unsigned long long &hi, unsigned long long &lo) {
  double y = x * 4294967296.0;
  double fl = floor(y);
  hi = (unsigned long long)(long long)fl;
  lo = (unsigned long long)((y - fl) * 4294967296.0);
}

// P: a pointer to unsigned long long in any address space (LDS, global, generic)
template <class P>
__device__ __forceinline__ void fx_add(P p, double x) {
  unsigned long long hi, lo;
  fx_split(x, hi, lo);
  __hip_atomic_fetch_add(p, hi, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __hip_atomic_fetch_add(p + 1, lo, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ double fx_value(unsigned long long hi, unsigned long long lo) {
  long long h = (long long)hi + (long long)(lo >> 32);
  unsigned long long r = lo & ((1ull << 32) - 1);
  return (double)h * 2.3283064365386963e-10 + (double)r * 5.421010862427522e-20;  // 2^-32, 2^-64
}

// Host bound for the fixed-point tap sums: every tap receives at most n_contrib
// contributions of magnitude <= w_max; the hi word's total stays below 2^62.
inline bool fx_capacity_ok(double n_contrib, double w_max) { return n_contrib * w_max < 1073741824.0; }

// Cell classification of one phase: loop over the phase's visible regions in
// order, later ones overwriting earlier ones (SOAP_integration.cu:724-785,
// spot_scan :346-387).  Returns -1 (quiet), 0 (spot) or 1 (facula).
template <class GeoPtr>
__device__ __forceinline__ int classify_cell(int iy, int iz, double delta_grid, GeoPtr geo, int g0, int g1,
                                             double *r_cos_out) {
#pragma clang fp contract(off)
  double y = -1. + iy * delta_grid;
  double z = -1. + iz * delta_grid;
  if (!(z * z + y * y < 1.)) return -1;
  double xa0 = sqrt(1. - (y * y + z * z));
  int res = -1;
  for (int r = g0; r < g1; r++) {
    const auto &g = geo[r];
    if (iy < g.y0 || iy >= g.y1 || iz < g.z0 || iz >= g.z1) continue;
    double xi0 = g.r0 * xa0 + g.r1 * y + g.r2 * z;
    if (xi0 * xi0 >= g.oms2) res = g.type;
  }
  *r_cos_out = xa0;
  return res;
}

// Facula temperature of a cell at r_cos (spot_scan, SOAP_integration.cu:376).
__device__ __forceinline__ int plage_temp(int t_sum, double rc) {
#pragma clang fp contract(off)
  return (int)(t_sum + 0.9 - 407.7 * rc + 190.9 * (rc * rc));
}


// Linear-grid gather path (soap_linear.hip)
struct LinArgs {
  int64_t n;
  int32_t n_mu, grid;
  double lambda0, step;       // step of the reference's index formula (argv step_size)
  const double *wave;         // [n]
  int32_t off_lo, span;       // source indices staged per tile: [j0 + off_lo, j0 + off_lo + span)
  const double *delta, *limb, *lbm;
};

// A run of cells of one (phase, mu bin) cell list with bitwise-equal shifts and
// one population, merged: the shift constants of the index formula, the summed
// limb weights w = sum L and w2 = sum (1 + LB) L, and the first cell's list entry
// (index | type << 30).  64 bytes, read with one scalar load per group.
struct alignas(16) LinGroup {
  double opf, fl, c1, rinv;  // 1 + fbeta, fbeta lambda_0 / step, 1 - sqrt((1+beta)/(1-beta)), 1 / opf
  double w, w2;
  int32_t e, ident;          // ident: delta == 0 (the row itself)
  int32_t pad[2];
};

// Source of the linear path's cell lists: the quiet disk (box NULL: the whole map,
// n_phase 1) or the active regions of each phase (scan boxes + region geometry);
// nch = chunks of 4096 box cells of the largest box (cell_list_chunks).
struct CellSrc {
  int32_t grid, n_mu, n_phase, nch;
  const int32_t *region;
  const double *limb, *delta;
  const RegionDev *geo;
  const int32_t *geo_off, *box;
};

struct LinSolar {
  int32_t on, t_sum, t_norm;  // Planck-ratio facula intensity (SOAP_integration_solar.cu:433)
};

// Differentiable path (soap_diff.hip)
struct DiffArgs {
  int64_t n;
  int32_t grid, n_mu, n_phase, map_res, dmin, nd;
  int32_t ns;                    // tap sets: 2 (quiet, spot) or 3 (quiet, spot, facula)
  double h, expm1_h, vsini, u1, u2, kappa, kappa_f;
  double lb[4];                  // facula intensity 1 + LB(mu), LB = lb0 + lb1 mu + lb2 mu^2 + lb3 mu^3
  const double *prm;             // device (vsini, u1, u2, kappa, kappa_f) if not null (graph replays), else the fields above
  const double *mu_bounds;       // n_mu - 1
  int32_t *ck, *cm;              // per cell [G*G]: bin (-1 off disk), shift m
  double *ct, *cL, *cmu, *cdt;   // weight t, limb, mu, dt/d(vsini)
  const double *maps, *fmaps;    // spot / facula filling factors [P][M][M] (fmaps null when ns == 2)
  unsigned long long *acc;       // [P][n_mu][nd][ns][hi, lo]
  double *taps;                  // [P][n_mu][nd][ns]
  int32_t *krange;               // [P][n_mu][2]
  const double *q, *s, *f;       // quiet / spot / facula seed rows [n_mu][n] (f null when ns == 2)
  double *I, *F;                 // [P][n]
};
constexpr int kDiffScalars = 5;  // gradients of (vsini, u1, u2, kappa, kappa_f)

struct ConvLinArgs {
  int64_t n;
  int32_t wid, n_tiles;
  int32_t n_active;           // 3 P active spectra first, then the quiet frame slots
  int32_t quiet_sel[3];       // quiet spectrum of each frame slot
  double c1;                  // 1 - sqrt((1+beta)/(1-beta)), beta = sigma/c
  double quiet_weight;        // w of the quiet form (active form: 0.005)
  const double *wave;
};

// Batched instances: one persistent launch per stage covers every instance.
// Work item w of the flattened index space belongs to the instance i with
// prefix[i] <= w < prefix[i + 1]; local = w - prefix[i] is the block index the
// single-instance kernel would have had.  Counts of the XCD-ordered stages
// (FIR, active convolution) are multiples of 8, so local % 8 == w % 8.
enum BatchStage {
  BS_FILL = 0, BS_TAPS, BS_FINALIZE, BS_FIR_PROG, BS_FIR, BS_CONV_QUIET, BS_CONV, BS_MINMAX, BS_NORM, BS_CCF,
  BS_CCF_REDUCE, BS_FIT, BS_COUNT
};
struct FillJob { unsigned long long *p; int64_t count; };           // work: ceil(count / batch_fill_chunk())
struct TapsJob { TapAccArgs a; int32_t mcb; };                        // work: mcb * P, local = p * mcb + col block
struct FinalizeJob {                                                  // work: p_pad
  const unsigned long long *acc;
  int32_t nsets_in, nsets_out, n_mu, nd;
  double *taps;
  int32_t *krange, *kflags, *bins, *nbins;
  double *taps_g;
  int32_t n_job, p_pad;
};
struct FirJob { FirArgs a; int32_t n_job, n_tiles, n_groups; };      // work: tiles8 * n_groups
struct FirProgJob { FirArgs a; FirProg pg; int32_t n_job, n_groups, lmax; };  // work: n_groups + 1
struct FirTileJob {  // work: 8 tiles_per_xcd
  FirArgs a;
  FirProg pg;
  int32_t n_job, n_groups, n_tiles, tiles_per_xcd, t64_log2;
};
struct ConvQJob { ConvArgs c; const double *quiet; double *frame, *part; };  // work: 2 n_tiles
// work: 3P tiles8 in XCD groups (xcd), else 3P n_tiles
struct ConvAJob { ConvArgs c; const double *quiet, *raw; double *out, *part; int32_t n_phase, xcd; };
struct MinmaxJob { const double *part; int32_t n_tiles; double *mm; };    // work: n_spec
struct NormJob { int64_t n, bx; const double *mm, *in; double *o; };     // work: bx * n_spec
struct CcfJob {                                                          // work: nc * ceil(n_spec / ccf_spec_chunk())
  const double *spec;
  int64_t stride, trim, n_t;
  int32_t n_spec, n_lines, n_v, nc;
  const int32_t *m_start;
  const uint32_t *m_pk;
  const double2 *m_wv;
  const double *w_sorted;
  const int2 *fp;  // line footprints, staged in LDS
  int32_t cap;     // LDS segment length (ccf_lds_cap)
  double *part;
};
struct CcfRedJob {  // work: n_spec
  const double *part;
  int32_t nc, n_spec, n_v;
  const double *appod, *mm, *wsum;
  double *ccf;
};
struct FitJob {                                                          // work: n_ccf
  const double *v;
  int32_t n_v;
  const double *y;
  double *par, *model, *depth, *bis;
  int32_t *iters;
};
}  // namespace soap

// launchers (soap_kernels.hip)
namespace soap {
hipError_t launch_fill_u64(unsigned long long *p, size_t count, hipStream_t s);
hipError_t launch_fill_f64(double *p, size_t count, double v, hipStream_t s);
hipError_t launch_disk_init(const DiskArgs &a, hipStream_t s);
hipError_t launch_taps_quiet(const TapAccArgs &a, hipStream_t s);
hipError_t launch_taps_active(const TapAccArgs &a, int n_job, int max_col_blocks, hipStream_t s);
// taps_g (optional): the phase-interleaved copy [n_mu][nd][p_pad][nsets_out] of the
// FIR, zero for jobs n_job..p_pad-1
hipError_t launch_taps_finalize(const unsigned long long *acc, int n_job, int nsets_in, int nsets_out, int n_mu,
                                int nd, double *taps, int32_t *krange, int32_t *kflags, int32_t *bins,
                                int32_t *nbins, double *taps_g, int p_pad, hipStream_t s);
// job stride of the phase-interleaved tap table for n_job jobs
inline int taps_g_pad(int n_job) { return (n_job + 7) / 8 * 8 + 8; }
hipError_t launch_fir_quiet(const FirArgs &a, hipStream_t s);
hipError_t launch_fir_active(const FirArgs &a, int n_job, hipStream_t s);
hipError_t launch_fir_quiet_lds(const FirArgs &a, hipStream_t s);
// phase-group form (bin outer, phases inner, LDS-staged rows and taps)
hipError_t launch_fir_active_grp(const FirArgs &a, int n_job, hipStream_t s);
// register-blocked form: G phases per block, B pixels per thread (G, B in
// {(1,3), (1,5), (2,3), (2,1), (4,1)})
hipError_t launch_fir2(const FirArgs &a, int n_job, int G, int B, hipStream_t s);
// direct form (no LDS staging): G phases per block, G in {1, 2, 4}
hipError_t launch_fir_direct(const FirArgs &a, int n_job, int G, hipStream_t s);
// tile-stationary form: 64 pixels and every phase per block; the taps in scalar
// registers (8 waves), or (dpp) in vector registers broadcast by DPP with `waves`
// waves per block (4..8; 0: chosen from the phase-group count) and `ahead` (4 or
// 6) stream entries in flight per wave
hipError_t launch_fir_tile(const FirArgs &a, int n_job, const FirProg &pg, bool dpp, int waves, int ahead,
                           hipStream_t s);
size_t fir_tile_lds(int n_mu, int nd, int t64);
int fir_tile_ahead();  // entries in flight per wave of the batched DPP form (stream padding)
int fir_tile_groups(int n_job);
// sub-tiles of 64 pixels per tile (log2) of the DPP form for this shape and LDS limit
int fir_tile_t64_log2(int n_job, int n_mu, int nd, size_t lds_max);
size_t fir2_lds(int nd, int G, int B);
int fir2_tiles(int64_t n, int B);
// solar mode: taps [job][n_mu][nd][K+3] = H1s, H1f, L_0..L_{K-1}, H1a
hipError_t launch_fir_solar(const FirArgs &a, const SolarFir &sf, int n_job, hipStream_t s);
hipError_t launch_phase_maps(int grid, int t_star, int t_diff_spot, int t_diff_fac, const RegionDev *geo,
                             int n_geo, double *temp, double *bi, hipStream_t s);
// resolution
int conv_tile();
int conv_chunk();
// the quiet-star frame (3 slots of the lower layout) from the two quiet spectra
hipError_t launch_conv_quiet(const ConvArgs &c, const double *quiet, double *frame, double *part, hipStream_t s);
// 3 spectra per phase from A = quiet[sel] - raw
hipError_t launch_conv_active(const ConvArgs &c, const double *quiet, const double *raw, int n_phase, double *o_out,
                              double *part, hipStream_t s);
// spectra sp = 0..n_spec-1 stored contiguously at o + sp*n, partials part[sp][n_tiles][2]
hipError_t launch_minmax(int n_spec, int n_tiles, const double *part, double *mm, hipStream_t s);
hipError_t launch_norm(int64_t n, int n_spec, const double *mm, const double *in, double *o, hipStream_t s);
// spec_norm (raw -> out) and the CCF partials of raw in one launch; the launcher
// fills nc, n_ccf, nbx, n_norm and ratio
struct NormCcfArgs {
  int64_t n, trim, n_t;
  int32_t n_spec, n_lines, n_v;
  const double *mm, *raw;
  double *out;
  const int32_t *m_start;
  const uint32_t *m_pk;
  const double2 *m_wv;
  const double *w_sorted;
  const int2 *fp;  // line footprints, staged in LDS by the CCF
  int32_t cap;     // LDS segment length (ccf_lds_cap)
  double *part;
  int32_t nc, n_ccf, nbx, n_norm, ratio;
};
hipError_t launch_norm_ccf(NormCcfArgs a, hipStream_t s);
// the CCF with line footprints staged in LDS (fp from launch_mask_footprint) in
// segments of cap = ccf_lds_cap(longest footprint) pixels, ccf_lds_bytes(cap) of
// LDS per block
hipError_t launch_mask_footprint(int n_lines, int n_v, int64_t n_t, const int32_t *m_start, const uint32_t *m_pk,
                                 int2 *fp, hipStream_t s);
int ccf_lds_cap(int fmax);
size_t ccf_lds_bytes(int cap);
hipError_t launch_ccf_lds(const double *spec, int64_t spec_stride, int64_t trim, int n_spec, int n_lines, int n_v,
                          int64_t n_t, const int32_t *m_start, const uint32_t *m_pk, const double2 *m_wv,
                          const double *w_sorted, const int2 *fp, int cap, double *part, hipStream_t s);
hipError_t launch_mask_wsum(int n_lines, int n_v, int64_t n_t, const int32_t *m_start, const uint32_t *m_pk,
                            const double2 *m_wv, const double *w_sorted, double *wsum, hipStream_t s);
// CCF
// mask tables [q][v], q = position in wavelength order (sorted[q] = line index)
hipError_t launch_mask_build(const double *line_l, const double *line_w, const int32_t *sorted, int n_lines,
                             const double *vgrid, int n_v, double lam_t0, double h, int32_t *m_start, int32_t *m_len,
                             double2 *m_wv, hipStream_t s);
hipError_t launch_mask_overwrite(const int32_t *sorted, int n_lines, int n_v, int scan, const int32_t *m_start,
                                 const int32_t *m_len, uint32_t *m_pk, hipStream_t s);
int ccf_chunks(int n_lines);
// CCF of n_spec spectra (spec + sp*spec_stride + trim), normalised and appodised;
// part: ccf_chunks(n_lines) * n_spec * n_v doubles; w_sorted[q] = weight of line sorted[q]
hipError_t launch_ccf_reduce(const double *part, int nc, int n_spec, int n_v, const double *appod, const double *mm,
                             const double *wsum, double *ccf, hipStream_t s);
// err: per-point errors [n_ccf][n_v] or NULL (unit errors); status (optional): per
// CCF 0 (converged), 27 (no progress), 11 (iteration limit)
hipError_t launch_fit(const double *v, int n_v, const double *y, const double *err, int n_ccf, double *par,
                      double *model, double *depth, double *bis, int32_t *iters, hipStream_t s,
                      int32_t *status = nullptr);
// linear grids
// cell lists of every (phase, bin) in map order: phase p's list from cells +
// list_off[p] (list_off NULL: offset 0), bin k of it at bin_off[p][k];
// cnt: n_phase * n_mu * nch ints of scratch
hipError_t launch_cell_lists(const CellSrc &src, const int64_t *list_off, int32_t *cnt, int32_t *cells,
                             int32_t *bin_off, hipStream_t s);
int cell_list_chunks(long box_cells);
// Groups of every (phase, bin) segment of a cell list (list_off NULL: one phase at
// offset 0), written at the segment's own start; g_n[p * n_mu + k] = group count.
// split_fac: facula cells stay single (solar mode's per-cell intensity).
hipError_t launch_lin_groups(const int32_t *cells, const int64_t *list_off, const int32_t *bin_off, int n_phase,
                             int n_mu, const double *delta, const double *limb, const double *lbm, double lambda0,
                             double step, int split_fac, LinGroup *groups, int32_t *g_n, hipStream_t s);
// part: quiet_linear_sets(n_mu) * 2 * n doubles of scratch
hipError_t launch_gather_quiet_linear(const LinArgs &a, const double *q, const double *qs, const LinGroup *groups,
                                      const int32_t *bin_off, const int32_t *g_n, double *part, double *out,
                                      hipStream_t s);
int quiet_linear_sets(int n_mu);
// Solid-body rotation only (delta constant along each disk column): the quiet star
// from the disk maps' limb weights summed per (bin, column); colw:
// quiet_column_sets(n_mu) * grid * 4 doubles, colf: quiet_column_sets(n_mu) * grid
// ints, cols: grid groups, part: quiet_column_sets(n_mu) * 2 * n doubles
hipError_t launch_gather_quiet_cols(const LinArgs &a, const double *q, const double *qs, const int32_t *region,
                                    double *colw, int32_t *colf, LinGroup *cols, double *part, double *out,
                                    hipStream_t s);
int quiet_column_sets(int n_mu);
bool quiet_columns_fit(int span);  // the staged rows of a bin set fit the block's LDS
hipError_t launch_gather_active_linear(const LinArgs &a, const FirArgs &f, const LinGroup *groups,
                                       const int64_t *list_off, const int32_t *bin_off, const int32_t *g_n,
                                       int n_phase, const LinSolar &sol, hipStream_t s);
hipError_t launch_conv_linear(const ConvLinArgs &c, int n_spec, const double *quiet, const double *raw, double *o_out,
                              double *part, hipStream_t s);
int conv_linear_tile();
hipError_t launch_mask_build_linear(const double *line_l, const double *line_w, const int32_t *sorted, int n_lines,
                                    const double *vgrid, int n_v, const double *wave_t, int64_t n_t, double step,
                                    int32_t *m_start, int32_t *m_len, double2 *m_wv, hipStream_t s);
// differentiable path
hipError_t launch_diff_forward_spectra(const DiffArgs &a, const ConvArgs &c, const double *target, int trim,
                                       double *res_part, int *res_blocks, hipStream_t s);
int diff_conv_blocks(int64_t n);
int diff_cell_blocks(int64_t ncell);
hipError_t launch_fit_polish(const double *v, int n_v, const double *ccf, const double *par, int n_ccf, double *theta,
                             hipStream_t s);
// mt_*: the mask in pixel-major form, sliced by 64 rows (row t = trimmed pixel
// t; columns = velocities; values = the CCF pixel weights, merged per (pixel,
// velocity)): entry k of row t at mt_ptr[t / 64] + 64 k + t % 64, mt_w[t / 64]
// entries per row of the slice (zero-padded)
hipError_t launch_diff_backward(const DiffArgs &a, const ConvArgs &c, const double *target, int trim, double w_spec,
                                const double *v, int n_v, const double *ccf, const double *theta, const double *dth,
                                double *gy, int64_t n_t, const int64_t *mt_ptr, const int32_t *mt_w,
                                const int32_t *mt_col, const double *mt_val, double *GF, double *GI, double *Cpart,
                                double *C, double *gmaps, double *gfmaps, double *cell_part, double *gq, double *gs,
                                double *gf, const double *bisd, hipStream_t s);
// bisector span of each CCF at the polished Gaussian and its partials (k_diff_bis):
// bisd [n_ccf][n_v + kBisHead] = (span, loss coefficient, dspan/dtheta[4], dspan/dy[n_v])
constexpr int kBisHead = 6;
constexpr int kMaxVDiff = 1024;
hipError_t launch_diff_bis(const double *v, int n_v, const double *ccf, const double *theta, int n_ccf, double *bisd,
                           hipStream_t s);
// chunks of the lag-correlation pass (Cpart holds [P][n_mu][chunks][nd][ns])
int diff_corr_chunks(int64_t n);
hipError_t launch_ccf_sum(const double *part, int nc, int n_spec, int n_v, double *ccf, hipStream_t s);
// device-side loss of the differentiable model: per phase the residual partials
// (res_blocks per phase, in order), the RV / FWHM terms and, when w_bis != 0, the
// bisector-span term from bisd; writes loss[0], the fit-parameter derivatives
// dth [P][4] and the span coefficients of bisd; loss = NaN when prm (optional) has
// vsini outside (0, vsini_max] (the fixed tap range of a graph) or weights beyond
// the fixed-point tap capacity of n_cells cells (lbmax = max |1 + LB(mu)|)
hipError_t launch_diff_loss(int n_phase, int res_blocks, const double *res_part, const double *theta,
                            const double *rv_t, const double *fw_t, const double *bis_t, double w_spec, double w_rv,
                            double w_fw, double w_bis, double *bisd, int n_v, const double *prm, double vsini_max,
                            double n_cells, double lbmax, double *loss, double *dth, hipStream_t s);
// g[q] = sum over blocks (in order) of cell_part[b][q], q < kDiffScalars
hipError_t launch_diff_grad(int n_blocks, const double *cell_part, double *g, hipStream_t s);
// batched instances: stage st over jobs[n_inst] (device), prefix[n_inst + 1]
// (device, prefix[n_inst] = total); grid = persistent block count (0: one block
// per work item), threads / lds as the stage needs; FIR shape: fir_g phases per
// block, fir_b pixels per thread (0: the one-pixel-per-thread form), fir_pre
// (one-pixel form only): prefetch the next bin's seed rows into registers
hipError_t launch_batch(int st, const void *jobs, const int64_t *prefix, int n_inst, int64_t total, int64_t grid,
                        int threads, size_t lds, int fir_g, int fir_b, int fir_pre, hipStream_t s);
int batch_fill_chunk();
int norm_px_per_block();
int ccf_spec_chunk();
int fir_tile();
size_t fir_lds(int nd);
size_t conv_lds_bytes(int ntp);
size_t taps_lds_bytes(const TapAccArgs &a);
// ------------------------------------------------------------ bisector injection (seed preprocessing)
constexpr int kInjMaxCoef = 16;
struct InjectArgs {
  int32_t mode;  // 0: BIS_inject_spectral_quiet, 1: BIS_inject_spectral_active
  int32_t n_up, n_low;
  double up[kInjMaxCoef], low[kInjMaxCoef];
  double dep_cut, vel_cut, velo_shift, scale;
};
int inject_blocks(int64_t n);
// moved wavelengths, and summary[3] = (min, max, 1 if they need sorting); part: 3 per block
hipError_t launch_inject_move(const InjectArgs &a, const double *wave, const double *flux, int64_t n, double *moved,
                              double *part, double *summary, hipStream_t s);
hipError_t inject_sort_bytes(int64_t n, size_t *bytes);
hipError_t launch_inject_sort(void *tmp, size_t tmp_bytes, const double *moved, double *xs, const double *flux,
                              double *ys, int64_t n, hipStream_t s);
hipError_t launch_inject_interp(const double *wave, const double *xs, const double *ys, int64_t n,
                                const double *summary, double *out, hipStream_t s);
// ------------------------------------------------------------ local PHOENIX spectra (seed preprocessing)
struct PhoenixArgs {
  double frac;        // (temp - t_low) / (t_high - t_low)
  double vlo, vhi;    // vacuum window wmin - 500, wmax + 500
  double wmin, wmax;  // air window
  int32_t f32;        // high - low in float32
};
hipError_t launch_phoenix(const PhoenixArgs &a, const double *wave, const double *lo, const double *hi, int64_t n,
                          double *air, double *flux, char *flag, hipStream_t s);
hipError_t phoenix_select_bytes(int64_t n, size_t *bytes);
hipError_t launch_phoenix_select(void *tmp, size_t bytes, const double *in, const char *flag, double *out,
                                 int64_t *count, int64_t n, hipStream_t s);

}  // namespace soap
