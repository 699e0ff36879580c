// gfx950 kernels of the linear-wavelength-grid path (the reference's native
// 0.005 A seed format).  On a linear grid a Doppler shift is not a constant
// pixel offset, so the per-cell shift + linear interpolation
// (fast_shift / fast_interp_new, SOAP_integration.cu:232-281) is evaluated in
// gather form for every (cell, pixel): pixel j takes the source index
// i(j) = max{i : ceil(i (1 + fbeta) + fbeta lambda_0 / step) <= j} and the
// reference's interpolation expression at i.  Cells are processed per mu bin
// from bin-sorted cell lists, with the bin's seed-row segments and the
// wavelengths of the tile staged in LDS.  The gathered value depends on a cell
// only through its shift, so each run of cells with bitwise-equal shifts in a
// bin's list (a disk column under solid-body rotation) is gathered once with the
// run's summed weights.  The convolution (conv_inst,
// SOAP_resolution.cu:76-95) has per-pixel Gaussian widths; the mask CCF uses
// the reference's literal index arithmetic (CCF_func.py:45-91).
#include <hip/hip_runtime.h>
#include <math.h>

#include "soap_internal.h"

namespace soap {

constexpr int kLTPB = 256;
constexpr int kQuietSets = 16;  // bin sets of the quiet-star gather (blocks per tile)

// ------------------------------------------------------------ cell lists
// Stable compaction of the block's predicate over 256 candidates: returns this
// thread's slot among the block's selected candidates, *total the count.
__device__ __forceinline__ int block_compact(bool pred, int *s_wave, int *total) {
  unsigned long long m = __ballot(pred);
  int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  int before = __popcll(m & ((1ull << lane) - 1ull));
  if (lane == 0) s_wave[wv] = __popcll(m);
  __syncthreads();
  int off = 0, tot = 0;
  for (int w = 0; w < (int)(blockDim.x >> 6); w++) {
    if (w < wv) off += s_wave[w];
    tot += s_wave[w];
  }
  __syncthreads();
  *total = tot;
  return off + before;
}

// Exclusive block-wide prefix sum of v (in thread order); *total = the block's sum.
__device__ __forceinline__ int block_scan_excl(int v, int *s_wave, int *total) {
  int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  int x = v;
  for (int o = 1; o < 64; o <<= 1) {
    int y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) s_wave[wv] = x;
  __syncthreads();
  int off = 0, tot = 0;
  for (int w = 0; w < (int)(blockDim.x >> 6); w++) {
    if (w < wv) off += s_wave[w];
    tot += s_wave[w];
  }
  __syncthreads();
  *total = tot;
  return off + x - v;
}

// Cell lists by (phase, mu bin), each in map order: the quiet disk (one phase whose
// box is the whole map; disk cells of bin k, entry = index) and the active regions
// (phase p's scan box; cells of bin k that classify as spot or facula, entry =
// index | type << 30).  Three passes over blocks (p, k, ch), ch = chunk of kChunk
// box cells in map order: count, per-phase offsets, write.
constexpr int kChunk = 4096;

// entry of box cell c of phase p for bin k, or -1
__device__ __forceinline__ int cell_entry(const CellSrc &s, int p, int k, long c) {
  int y0 = 0, z0 = 0, nz = s.grid;
  if (s.box) {
    y0 = s.box[p * 4];
    z0 = s.box[p * 4 + 2];
    nz = s.box[p * 4 + 3] - z0;
  }
  int iy = y0 + (int)(c / nz), iz = z0 + (int)(c % nz);
  int idx = iy * s.grid + iz;
  if (s.region[idx] != k) return -1;
  if (!s.box) return (s.limb[idx] == -999.0 && s.delta[idx] == -999.0) ? -1 : idx;
  double rc;
  int ty = classify_cell(iy, iz, 2. / s.grid, s.geo, s.geo_off[p], s.geo_off[p + 1], &rc);
  return ty >= 0 ? (idx | (ty << 30)) : -1;
}

__device__ __forceinline__ long cell_box_size(const CellSrc &s, int p) {
  if (!s.box) return (long)s.grid * s.grid;
  const int32_t *bx = s.box + p * 4;
  return (long)(bx[1] - bx[0]) * (bx[3] - bx[2]);
}

__global__ __launch_bounds__(kLTPB) void k_cells_count(CellSrc s, int32_t *cnt) {
  __shared__ int s_n;
  long b = blockIdx.x;
  int ch = (int)(b % s.nch);
  long pk = b / s.nch;
  int k = (int)(pk % s.n_mu), p = (int)(pk / s.n_mu);
  long lo = (long)ch * kChunk, nbox = cell_box_size(s, p);
  long hi = lo + kChunk < nbox ? lo + kChunk : nbox;
  if (threadIdx.x == 0) s_n = 0;
  __syncthreads();
  int n = 0;
  for (long c = lo + threadIdx.x; c < hi; c += kLTPB) n += cell_entry(s, p, k, c) >= 0;
  if (n) atomicAdd(&s_n, n);
  __syncthreads();
  if (threadIdx.x == 0) cnt[b] = s_n;
}

// Phase p = blockIdx.x: exclusive offsets over (k, ch) within the phase, in place;
// bin_off[p][k] = offset of (k, 0), bin_off[p][n_mu] = the phase's total.
__global__ __launch_bounds__(kLTPB) void k_cells_offsets(int n_mu, int nch, int32_t *cnt, int32_t *bin_off) {
  __shared__ int s_wave[kLTPB / 64];
  int p = blockIdx.x;
  long ne = (long)n_mu * nch;
  int32_t *c = cnt + (long)p * ne;
  long per = (ne + kLTPB - 1) / kLTPB;
  long e0 = threadIdx.x * per, e1 = e0 + per < ne ? e0 + per : ne;
  int sum = 0;
  for (long e = e0; e < e1; e++) sum += c[e];
  int tot;
  int run = block_scan_excl(sum, s_wave, &tot);
  for (long e = e0; e < e1; e++) {
    int v = c[e];
    c[e] = run;
    if (e % nch == 0) bin_off[(long)p * (n_mu + 1) + e / nch] = run;
    run += v;
  }
  if (threadIdx.x == 0) bin_off[(long)p * (n_mu + 1) + n_mu] = tot;
}

__global__ __launch_bounds__(kLTPB) void k_cells_write(CellSrc s, const int32_t *off, const int64_t *list_off,
                                                       int32_t *cells) {
  __shared__ int s_wave[kLTPB / 64];
  long b = blockIdx.x;
  int ch = (int)(b % s.nch);
  long pk = b / s.nch;
  int k = (int)(pk % s.n_mu), p = (int)(pk / s.n_mu);
  long lo = (long)ch * kChunk, nbox = cell_box_size(s, p);
  if (lo >= nbox) return;  // uniform over the block
  long hi = lo + kChunk < nbox ? lo + kChunk : nbox;
  int32_t *out = cells + (list_off ? list_off[p] : 0) + off[b];
  int run = 0;
  for (long c0 = lo; c0 < hi; c0 += kLTPB) {
    long c = c0 + threadIdx.x;
    int e = c < hi ? cell_entry(s, p, k, c) : -1;
    int tot;
    int slot = block_compact(e >= 0, s_wave, &tot);
    if (e >= 0) out[run + slot] = e;
    run += tot;
  }
}

hipError_t launch_cell_lists(const CellSrc &src, const int64_t *list_off, int32_t *cnt, int32_t *cells,
                             int32_t *bin_off, hipStream_t st) {
  long blocks = (long)src.n_phase * src.n_mu * src.nch;
  if (blocks == 0) return hipSuccess;
  if (blocks > 0x7fffffffL) return hipErrorInvalidValue;
  k_cells_count<<<(unsigned)blocks, kLTPB, 0, st>>>(src, cnt);
  k_cells_offsets<<<src.n_phase, kLTPB, 0, st>>>(src.n_mu, src.nch, cnt, bin_off);
  k_cells_write<<<(unsigned)blocks, kLTPB, 0, st>>>(src, cnt, list_off, cells);
  return hipGetLastError();
}

int cell_list_chunks(long box_cells) { return (int)((box_cells + kChunk - 1) / kChunk); }

// mu of a cell from its grid position (as classify_cell computes it)
__device__ __forceinline__ double cell_rcos(int idx, int grid) {
#pragma clang fp contract(off)
  int iy = idx / grid, iz = idx % grid;
  double dg = 2. / grid;
  double y = -1. + iy * dg, z = -1. + iz * dg;
  return sqrt(1. - (y * y + z * z));
}

// ------------------------------------------------------------ shift on a linear grid
struct LinShift {
  bool ident;            // delta == 0: the row itself (SOAP_integration.cu:262-265)
  double opf, fl, c1;    // 1 + fbeta, fbeta lambda_0 / step, 1 - sqrt((1+beta)/(1-beta))
  double rinv;           // 1 / opf (first guess of the gather index only)
};

__device__ __forceinline__ LinShift lin_shift(double delta_kms, double lambda0, double step) {
#pragma clang fp contract(off)
  LinShift s;
  s.ident = delta_kms == 0.0;
  double beta = delta_kms * 1000. / kC;
  double sq = sqrt((1 + beta) / (1 - beta));
  double fbeta = -1 * (1 - sq);
  s.c1 = 1 - sq;
  s.opf = 1 + fbeta;
  s.fl = fbeta * lambda0 / step;
  s.rinv = 1 / s.opf;
  return s;
}

__device__ __forceinline__ LinShift group_shift(const CONSTANT LinGroup &g) {
  LinShift s;
  s.ident = g.ident != 0;
  s.opf = g.opf;
  s.fl = g.fl;
  s.c1 = g.c1;
  s.rinv = g.rinv;
  return s;
}

// lambda + Delta_lambda(delta 1000, lambda) (SOAP_integration.cu:240)
__device__ __forceinline__ double lin_lshift(double lam, const LinShift &s) {
#pragma clang fp contract(off)
  return lam + -1 * lam * s.c1;
}

// max{i : ceil(i (1+fbeta) + fbeta lambda_0/step) <= j}: a first guess from the
// reciprocal, then the reference's own expression decides
// (32-bit indices: the launchers refuse grids of 2^31 - 2^12 pixels or more)
__device__ __forceinline__ int lin_gather_index(int j, const LinShift &s) {
#pragma clang fp contract(off)
  int i = (int)floor((j - s.fl) * s.rinv);
  for (int it = 0; it < 8 && (int)ceil(i * s.opf + s.fl) > j; it++) i--;
  for (int it = 0; it < 8 && (int)ceil((i + 1) * s.opf + s.fl) <= j; it++) i++;
  return i;
}

// Staged tile: source indices [base, base + span) of the rows and the wavelengths
// (linear continuation outside the grid, rows 0 there).  The host sizes the window
// from the extreme shifts on the disk with 3-pixel margins on both sides, so every
// i, i + 1 a gather of the tile reads lies inside it; local() clamps all the same,
// so that no index can leave the LDS arrays.
__device__ __forceinline__ int tile_local(int i, int base, int span) {
  int li = i - base;
  return li < 0 ? 0 : (li > span - 2 ? span - 2 : li);
}
constexpr long kLinMaxN = 0x7fffffffL - 4096;

__device__ void lin_stage(double *dst, const double *grow, long base, int span, long n) {
  for (int e = threadIdx.x; e < span; e += blockDim.x) {
    long i = base + e;
    dst[e] = (i >= 0 && i < n) ? grow[i] : 0.0;
  }
}

__device__ void lin_stage_wave(double *dst, const double *wave, long base, int span, long n) {
  for (int e = threadIdx.x; e < span; e += blockDim.x) {
    long i = base + e;
    double v;
    if (i < 0)
      v = wave[0] + (double)i * (wave[1] - wave[0]);
    else if (i >= n)
      v = wave[n - 1] + (double)(i - (n - 1)) * (wave[n - 1] - wave[n - 2]);
    else
      v = wave[i];
    dst[e] = v;
  }
}

// ------------------------------------------------------------ shift groups
// Cell c of a segment starts a new group unless the previous cell has the same
// population and a bitwise-equal shift (facula cells in solar mode always start
// one: their intensity depends on the cell's own mu).
__device__ __forceinline__ bool starts_group(const int32_t *cells, long c, const double *delta, int split_fac) {
  int e = cells[c], ep = cells[c - 1];
  int ty = e >> 30;
  if (ty != (ep >> 30) || (split_fac && ty == 1)) return true;
  return delta[e & 0x3fffffff] != delta[ep & 0x3fffffff];
}

// Segment (p, k) = blockIdx.x of a bin-sorted cell list: its runs of equal shift
// become groups with w = sum L and w2 = sum (1 + LB) L over the run, stored from
// the segment's start.  sum_c L_c v(delta) = v(delta) sum_c L_c, so the gathers
// over groups equal the per-cell sums of SOAP_integration.cu:285-307,392-447
// (and itot's, SOAP_initialize.cu:339-395) up to rounding.  The run sums come from
// a segmented scan over each 256-cell chunk (restarting at run heads, in a fixed
// order) plus the open run carried from the chunk before; TPB cells per chunk.
template <int TPB>
__global__ __launch_bounds__(TPB) void k_lin_groups(const int32_t *cells, const int64_t *list_off,
                                                    const int32_t *bin_off, int n_mu, const double *delta,
                                                    const double *limb, const double *lbm, double lambda0,
                                                    double step, int split_fac, LinGroup *groups, int32_t *g_n) {
#pragma clang fp contract(off)
  __shared__ int s_wave[TPB / 64];
  __shared__ double s_w[TPB], s_w2[TPB];
  __shared__ int s_f[TPB];
  __shared__ double s_carry[2];
  const int tid = threadIdx.x;
  int sg = blockIdx.x, p = sg / n_mu, k = sg % n_mu;
  long base = list_off ? list_off[p] : 0;
  const int32_t *bo = bin_off + (long)p * (n_mu + 1);
  long c0 = base + bo[k], c1 = base + bo[k + 1];
  LinGroup *gs = groups + c0;
  int run = 0;                  // groups begun before this chunk
  double cw = 0.0, cw2 = 0.0;   // sums of the run open at the chunk's start
  for (long cb = c0; cb < c1; cb += TPB) {
    long c = cb + tid;
    bool in = c < c1;
    bool head = in && (c == c0 || starts_group(cells, c, delta, split_fac));
    bool last = in && (c + 1 == c1 || starts_group(cells, c + 1, delta, split_fac));
    int e = in ? cells[c] : 0;
    double v = 0.0, v2 = 0.0;
    if (in) {
      int id = e & 0x3fffffff;
      v = limb[id];
      v2 = (1 + lbm[id]) * v;
    }
    int tot;
    int before = block_compact(head, s_wave, &tot);  // heads before this cell in the chunk
    int f = head;
    for (int o = 1; o < TPB; o <<= 1) {
      s_w[tid] = v;
      s_w2[tid] = v2;
      s_f[tid] = f;
      __syncthreads();
      if (tid >= o && !f) {
        v = s_w[tid - o] + v;
        v2 = s_w2[tid - o] + v2;
        f = s_f[tid - o];
      }
      __syncthreads();
    }
    if (!f) {  // no head at or before this cell in the chunk: the carried run
      v = cw + v;
      v2 = cw2 + v2;
    }
    if (head) {
      LinShift sh = lin_shift(delta[e & 0x3fffffff], lambda0, step);
      LinGroup &g = gs[run + before];
      g.opf = sh.opf;
      g.fl = sh.fl;
      g.c1 = sh.c1;
      g.rinv = sh.rinv;
      g.e = e;
      g.ident = sh.ident ? 1 : 0;
      g.pad[0] = g.pad[1] = 0;
    }
    if (last) {  // the run's group: the last head at or before this cell
      LinGroup &g = gs[run + before + (head ? 1 : 0) - 1];
      g.w = v;
      g.w2 = v2;
    }
    long lastc = (cb + TPB < c1 ? cb + TPB : c1) - 1;
    if (c == lastc) {
      s_carry[0] = last ? 0.0 : v;
      s_carry[1] = last ? 0.0 : v2;
    }
    __syncthreads();
    cw = s_carry[0];
    cw2 = s_carry[1];
    run += tot;
  }
  if (tid == 0) g_n[sg] = run;
}

hipError_t launch_lin_groups(const int32_t *cells, const int64_t *list_off, const int32_t *bin_off, int n_phase,
                             int n_mu, const double *delta, const double *limb, const double *lbm, double lambda0,
                             double step, int split_fac, LinGroup *groups, int32_t *g_n, hipStream_t s) {
  long nseg = (long)n_phase * n_mu;
  if (nseg == 0) return hipSuccess;
  if (nseg > 0x7fffffffL) return hipErrorInvalidValue;
  // 1024-cell chunks: the time goes to the longest segments (the active cells of
  // a phase crowd into the few bins its regions cover)
  k_lin_groups<1024><<<(unsigned)nseg, 1024, 0, s>>>(cells, list_off, bin_off, n_mu, delta, limb, lbm, lambda0, step,
                                                     split_fac, groups, g_n);
  return hipGetLastError();
}

// Quiet star on a linear grid: out[0][j] = sum_cells limb Shift(Sq_k)(j), out[1] for
// the SED cube (itot's cell loop, SOAP_initialize.cu:339-395), per tile of kLTPB
// pixels, bins outer, shift groups of the bin inner.  Block (tile, s) takes bins
// s, s + ns, ... and writes its partial sums to part[s][2][n]; k_quiet_sum adds
// the ns partials in order.
__global__ __launch_bounds__(kLTPB) void k_gather_quiet_linear(LinArgs a, const double *q, const double *qs,
                                                               const LinGroup *groups, const int32_t *bin_off,
                                                               const int32_t *g_n, double *part) {
  extern __shared__ double lds[];  // rows [2][span], wave [span]
  const int ns = gridDim.y, set = blockIdx.y;
  const int j0 = blockIdx.x * kLTPB;
  const int j = j0 + threadIdx.x;
  const int base = j0 + a.off_lo;
  const int span = a.span;
  double *rw = lds, *wl = lds + 2 * span;
  lin_stage_wave(wl, a.wave, base, span, a.n);
  const int lj = j - base;  // inside the window: -off_lo .. kLTPB - 1 - off_lo
  double lam_j = a.wave[j < a.n ? j : a.n - 1];
  double acc_q = 0.0, acc_s = 0.0;
  const CONSTANT LinGroup *gp = gc(groups);
  for (int k = set; k < a.n_mu; k += ns) {
    int c0 = bin_off[k], ng = g_n[k];
    if (ng == 0) continue;  // uniform
    const double *gq = q + (long)k * a.n, *gs = qs + (long)k * a.n;
    __syncthreads();
    lin_stage(rw, gq, base, span, a.n);
    lin_stage(rw + span, gs, base, span, a.n);
    __syncthreads();
    for (int c = c0; c < c0 + ng; c++) {
      const CONSTANT LinGroup &G = gp[c];
      double L = G.w;
      LinShift s = group_shift(G);
      if (s.ident) {
        acc_q += rw[lj] * L;
        acc_s += rw[span + lj] * L;
        continue;
      }
      const int li = tile_local(lin_gather_index(j, s), base, span);
      double ls0 = lin_lshift(wl[li], s), ls1 = lin_lshift(wl[li + 1], s);
      double t = (lam_j - ls0) / (ls1 - ls0);
      double a0 = rw[li], a1 = rw[li + 1];
      double b0 = rw[span + li], b1 = rw[span + li + 1];
      acc_q += (a0 + (a1 - a0) * t) * L;
      acc_s += (b0 + (b1 - b0) * t) * L;
    }
  }
  if (j < a.n) {
    double *o = part + (long)set * 2 * a.n;
    o[j] = acc_q;
    o[a.n + j] = acc_s;
  }
}

__global__ __launch_bounds__(kLTPB) void k_quiet_sum(const double *part, int ns, long n2, double *out) {
  long j = (long)blockIdx.x * kLTPB + threadIdx.x;
  if (j >= n2) return;
  double acc = part[j];
  for (int t = 1; t < ns; t++) acc += part[(long)t * n2 + j];
  out[j] = acc;
}

int quiet_linear_sets(int n_mu) { return n_mu < kQuietSets ? n_mu : kQuietSets; }

hipError_t launch_gather_quiet_linear(const LinArgs &a, const double *q, const double *qs, const LinGroup *groups,
                                      const int32_t *bin_off, const int32_t *g_n, double *part, double *out,
                                      hipStream_t s) {
  size_t lds = (size_t)3 * a.span * sizeof(double);
  if (lds > 65536 || a.span < kLTPB || a.n_mu < 1 || a.n > kLinMaxN) return hipErrorInvalidValue;
  unsigned nb = (unsigned)((a.n + kLTPB - 1) / kLTPB);
  int ns = quiet_linear_sets(a.n_mu);
  k_gather_quiet_linear<<<dim3(nb, ns), kLTPB, lds, s>>>(a, q, qs, groups, bin_off, g_n, part);
  long n2 = 2 * a.n;
  k_quiet_sum<<<(unsigned)((n2 + kLTPB - 1) / kLTPB), kLTPB, 0, s>>>(part, ns, n2, out);
  return hipGetLastError();
}

// ------------------------------------------------------------ quiet star by columns
// Under solid-body rotation delta = y omega sin i is the same along a disk column
// in every mu bin, so the quiet star's shift groups are (bin, column) pairs that
// share their column's index and weight.  k_quiet_columns_maps lays them out as a
// table: per (bin set s, column iy) the weights of the set's kQColBins bins (0
// where the column does not meet the bin), a flag for columns with any weight, and
// the column's shift; k_gather_quiet_cols then computes the index and weight once
// per (pixel, column) for all bins of its set.
constexpr int kQColBins = 4;

// The same table straight from the disk maps: block = column iy (one wave), the
// lanes over the column's cells; per bin, each lane sums its cells (z = lane,
// lane + 64, ...) and a fixed butterfly adds the 64 partials.  The column's shift
// comes from its first disk cell.
__global__ __launch_bounds__(64) void k_quiet_columns_maps(const double *limb, const double *delta,
                                                           const int32_t *region, int n_mu, int grid, double lambda0,
                                                           double step, double *colw, int32_t *colf,
                                                           LinGroup *cols) {
#pragma clang fp contract(off)
  const int iy = blockIdx.x, lane = threadIdx.x;
  const long row = (long)iy * grid;
  const int ns = (n_mu + kQColBins - 1) / kQColBins;
  int first = -1;
  for (int z0 = 0; z0 < grid && first < 0; z0 += 64) {
    const int iz = z0 + lane;
    const bool disk = iz < grid && !(limb[row + iz] == -999.0 && delta[row + iz] == -999.0);
    const unsigned long long m = __ballot(disk);
    if (m) first = z0 + __ffsll((long long)m) - 1;
  }
  for (int s = 0; s < ns; s++) {
    int any = 0;
    for (int kk = 0; kk < kQColBins; kk++) {
      const int k = s * kQColBins + kk;
      double w = 0.0;
      if (k < n_mu) {
        for (int z0 = 0; z0 < grid; z0 += 64) {
          const long idx = row + z0 + lane;
          if (z0 + lane < grid && region[idx] == k && !(limb[idx] == -999.0 && delta[idx] == -999.0)) w += limb[idx];
        }
        for (int o = 32; o > 0; o >>= 1) w += __shfl_xor(w, o, 64);
      }
      any |= w != 0.0;
      if (lane == 0) colw[((long)s * grid + iy) * kQColBins + kk] = w;
    }
    if (lane == 0) colf[(long)s * grid + iy] = any;
  }
  if (lane == 0) {
    LinGroup g{};
    if (first >= 0) {
      LinShift sh = lin_shift(delta[row + first], lambda0, step);
      g.opf = sh.opf;
      g.fl = sh.fl;
      g.c1 = sh.c1;
      g.rinv = sh.rinv;
      g.e = (int)(row + first);
      g.ident = sh.ident ? 1 : 0;
    } else {  // no disk cell in this column: an identity shift, never used
      g.opf = 1.0;
      g.rinv = 1.0;
      g.ident = 1;
    }
    cols[iy] = g;
  }
}

// Block (tile, s): the rows of bins s * kQColBins .. + kQColBins - 1 staged
// together (rows of bins past n_mu staged as zeros, with weight 0); per column
// with weight, one gather index and weight for all of them.
__global__ __launch_bounds__(kLTPB) void k_gather_quiet_cols(LinArgs a, const double *q, const double *qs,
                                                             const double *colw, const int32_t *colf,
                                                             const LinGroup *cols, double *part) {
  extern __shared__ double lds[];  // rows [kQColBins][2][span], wave [span]
  const int set = blockIdx.y;
  const int j0 = blockIdx.x * kLTPB;
  const int j = j0 + threadIdx.x;
  const int base = j0 + a.off_lo;
  const int span = a.span;
  double *rw = lds, *wl = lds + 2 * kQColBins * span;
  lin_stage_wave(wl, a.wave, base, span, a.n);
  for (int kk = 0; kk < kQColBins; kk++) {
    int k = set * kQColBins + kk;
    double *dq = rw + (2 * kk) * span, *ds = dq + span;
    if (k < a.n_mu) {
      lin_stage(dq, q + (long)k * a.n, base, span, a.n);
      lin_stage(ds, qs + (long)k * a.n, base, span, a.n);
    } else {
      for (int e = threadIdx.x; e < 2 * span; e += blockDim.x) dq[e] = 0.0;
    }
  }
  __syncthreads();
  const int lj = j - base;  // inside the window: -off_lo .. kLTPB - 1 - off_lo
  double lam_j = a.wave[j < a.n ? j : a.n - 1];
  double acc_q = 0.0, acc_s = 0.0;
  const CONSTANT LinGroup *cp = gc(cols);
  const CONSTANT double *wp = gc(colw) + (long)set * a.grid * kQColBins;
  const CONSTANT int32_t *fp = gc(colf) + (long)set * a.grid;
  for (int iy = 0; iy < a.grid; iy++) {
    if (!fp[iy]) continue;  // uniform
    LinShift s = group_shift(cp[iy]);
    const CONSTANT double *w = wp + (long)iy * kQColBins;
    if (s.ident) {
#pragma unroll
      for (int kk = 0; kk < kQColBins; kk++) {
        acc_q += rw[(2 * kk) * span + lj] * w[kk];
        acc_s += rw[(2 * kk + 1) * span + lj] * w[kk];
      }
      continue;
    }
    const int li = tile_local(lin_gather_index(j, s), base, span);
    double ls0 = lin_lshift(wl[li], s), ls1 = lin_lshift(wl[li + 1], s);
    double t = (lam_j - ls0) / (ls1 - ls0);
#pragma unroll
    for (int kk = 0; kk < kQColBins; kk++) {
      const double *dq = rw + (2 * kk) * span, *ds = dq + span;
      double a0 = dq[li], a1 = dq[li + 1];
      double b0 = ds[li], b1 = ds[li + 1];
      acc_q += (a0 + (a1 - a0) * t) * w[kk];
      acc_s += (b0 + (b1 - b0) * t) * w[kk];
    }
  }
  if (j < a.n) {
    double *o = part + (long)set * 2 * a.n;
    o[j] = acc_q;
    o[a.n + j] = acc_s;
  }
}

int quiet_column_sets(int n_mu) { return (n_mu + kQColBins - 1) / kQColBins; }

bool quiet_columns_fit(int span) { return (size_t)(2 * kQColBins + 1) * span * sizeof(double) <= 65536; }

hipError_t launch_gather_quiet_cols(const LinArgs &a, const double *q, const double *qs, const int32_t *region,
                                    double *colw, int32_t *colf, LinGroup *cols, double *part, double *out,
                                    hipStream_t s) {
  size_t lds = (size_t)(2 * kQColBins + 1) * a.span * sizeof(double);
  if (lds > 65536 || a.span < kLTPB || a.n_mu < 1 || a.n > kLinMaxN || a.grid < 1) return hipErrorInvalidValue;
  k_quiet_columns_maps<<<(unsigned)a.grid, 64, 0, s>>>(a.limb, a.delta, region, a.n_mu, a.grid, a.lambda0, a.step,
                                                       colw, colf, cols);
  unsigned nb = (unsigned)((a.n + kLTPB - 1) / kLTPB);
  int ns = quiet_column_sets(a.n_mu);
  k_gather_quiet_cols<<<dim3(nb, ns), kLTPB, lds, s>>>(a, q, qs, colw, colf, cols, part);
  long n2 = 2 * a.n;
  k_quiet_sum<<<(unsigned)((n2 + kLTPB - 1) / kLTPB), kLTPB, 0, s>>>(part, ns, n2, out);
  return hipGetLastError();
}

// Active deltas of phase p on a linear grid (spot_integrate + fast_flux / bconv /
// tot, SOAP_integration.cu:285-307,392-447; solar variant :433 of
// SOAP_integration_solar.cu): per active cell, Q = Shift(Sq_k), QS = Shift(SEDq_k),
// P = Shift(Sspot_k or Sfac_k), all x limb; intensity I = C_spot (spot),
// (1+LB) C_fac or B(T_cell)/B(T_norm) C_fac (facula); flux += QS - I Q,
// bconv += Q - P, tot += QS - I P.  Over a shift group the limb weights sum
// (w = sum L) and the intensity weights too (spot: C_spot w, facula: C_fac
// sum (1+LB) L; solar-mode facula groups are single cells).  Blocks are numbered
// so the phases of a tile run back to back on one XCD.
__global__ __launch_bounds__(kLTPB) void k_gather_active_linear(LinArgs a, FirArgs f, const LinGroup *groups,
                                                                const int64_t *list_off, const int32_t *bin_off,
                                                                const int32_t *g_n, int n_phase, int n_tiles,
                                                                LinSolar sol) {
  extern __shared__ double lds[];  // rows [4][span]: q, qs, spot, facula; wave [span]
  long b = blockIdx.x;
  long slot = b / 8;
  long tile = (slot / n_phase) * 8 + b % 8;
  int p = (int)(slot % n_phase);
  if (tile >= n_tiles) return;  // uniform over the block
  const int j0 = (int)tile * kLTPB;
  const int j = j0 + threadIdx.x;
  const int jc = j < a.n ? j : (int)a.n - 1;
  const int base = j0 + a.off_lo;
  const int span = a.span;
  double *rw = lds, *wl = lds + 4 * span;
  lin_stage_wave(wl, a.wave, base, span, a.n);
  const int lj = j - base;  // inside the window: -off_lo .. kLTPB - 1 - off_lo
  double lam_j = a.wave[jc];
  double cs = f.sps ? f.sps[jc] : 0.0, cf = f.fas ? f.fas[jc] : 0.0;
  double hc = kPlanckH * kC;
  double en = sol.on ? expm1(hc / (lam_j * kBoltzK * (double)sol.t_norm * 1e-10)) : 0.0;
  double flux = 0.0, bconv = 0.0, tot = 0.0;
  const CONSTANT LinGroup *gp = gc(groups + list_off[p]);
  const int32_t *bo = bin_off + (long)p * (a.n_mu + 1);
  const int32_t *gn = g_n + (long)p * a.n_mu;
  for (int k = 0; k < a.n_mu; k++) {
    int c0 = bo[k], ng = gn[k];
    if (ng == 0) continue;  // uniform
    const long ko = (long)k * a.n;
    const double *g[4] = {f.q + ko, f.qs + ko, f.sp ? f.sp + ko : f.q + ko, f.fa ? f.fa + ko : f.q + ko};
    __syncthreads();
    for (int r = 0; r < 4; r++) lin_stage(rw + (long)r * span, g[r], base, span, a.n);
    __syncthreads();
    for (int c = c0; c < c0 + ng; c++) {
      const CONSTANT LinGroup &G = gp[c];
      int e = G.e;
      int idx = e & 0x3fffffff, ty = e >> 30;
      double L = G.w;
      int r_act = ty == 0 ? 2 : 3;
      LinShift s = group_shift(G);
      const double *ra = rw + r_act * span;
      double iq, is, ia;  // shifted rows at pixel j
      if (s.ident) {
        iq = rw[lj];
        is = rw[span + lj];
        ia = ra[lj];
      } else {
        const int li = tile_local(lin_gather_index(j, s), base, span);
        double ls0 = lin_lshift(wl[li], s), ls1 = lin_lshift(wl[li + 1], s);
        double t = (lam_j - ls0) / (ls1 - ls0);
        double q0 = rw[li], q1 = rw[li + 1];
        double s0 = rw[span + li], s1 = rw[span + li + 1];
        double a0 = ra[li], a1 = ra[li + 1];
        iq = q0 + (q1 - q0) * t;
        is = s0 + (s1 - s0) * t;
        ia = a0 + (a1 - a0) * t;
      }
      // IW = sum over the group of I L
      double IW;
      if (ty == 0) {
        IW = cs * L;  // B(T_spot)/B(int T_spot) = 1
      } else if (sol.on) {
        double T_cell = (double)plage_temp(sol.t_sum, cell_rcos(idx, a.grid));
        IW = en / expm1(hc / (lam_j * kBoltzK * T_cell * 1e-10)) * cf * L;
      } else {
        IW = cf * G.w2;
      }
      double vq = iq * L, vs = is * L, va = ia * L;
      flux = flux + vs - IW * iq;
      bconv = bconv + vq - va;
      tot = tot + vs - IW * ia;
    }
  }
  if (j < a.n) {
    double *o = f.out + (long)p * 3 * a.n;
    o[j] = flux;
    o[a.n + j] = bconv;
    o[2 * a.n + j] = tot;
  }
}

hipError_t launch_gather_active_linear(const LinArgs &a, const FirArgs &f, const LinGroup *groups,
                                       const int64_t *list_off, const int32_t *bin_off, const int32_t *g_n,
                                       int n_phase, const LinSolar &sol, hipStream_t s) {
  if (n_phase == 0) return hipSuccess;
  size_t lds = (size_t)5 * a.span * sizeof(double);
  if (lds > 65536 || a.span < kLTPB || a.n > kLinMaxN) return hipErrorInvalidValue;
  int n_tiles = (int)((a.n + kLTPB - 1) / kLTPB);
  long tiles8 = ((long)n_tiles + 7) / 8 * 8;
  long blocks = tiles8 * n_phase;
  if (blocks > 0x7fffffffL) return hipErrorInvalidValue;
  k_gather_active_linear<<<(unsigned)blocks, kLTPB, lds, s>>>(a, f, groups, list_off, bin_off, g_n, n_phase,
                                                              n_tiles, sol);
  return hipGetLastError();
}

// ------------------------------------------------------------ resolution
// conv_inst on a linear grid (SOAP_initialize.cu:78-96, SOAP_resolution.cu:76-95):
// o_i = sum_{d=-wid}^{wid-2} w (1 - A_{i+d}) / (sig_i sqrt(2 pi)) exp(-d^2 0.005^2 / (2 sig_i^2)),
// sig_i = Delta_lambda(sigma, lambda_i), for wid < i < n-1-wid (0 elsewhere).
// The Gaussian factors follow g_{d+1} = g_d u_d, u_{d+1} = u_d v from g_0 = 1,
// u_0 = exp(-a), v = exp(-2a), a = 0.005^2 / (2 sig_i^2) (relative error of
// g_d below |d| ulp).  Block = kLTPB pixels of one spectrum; spectrum sp =
// blockIdx.y: A = quiet[sel] - raw (active) or the quiet spectrum itself.
__global__ __launch_bounds__(kLTPB) void k_conv_linear(ConvLinArgs c, const double *quiet, const double *raw,
                                                       double *o_out, double *part) {
  extern __shared__ double xs[];  // 1 - A over [i0 - wid, i0 + kLTPB + wid)
  long sp = blockIdx.y;
  long i0 = (long)blockIdx.x * kLTPB;
  int span = kLTPB + 2 * c.wid;
  const double *Q, *D = nullptr;
  double w;
  if (sp < c.n_active) {
    int t = (int)(sp % 3);
    Q = quiet + (long)(t == 1 ? 0 : 1) * c.n;  // bconv uses the normalised quiet star
    D = raw + sp * c.n;
    w = 0.005;
  } else {
    Q = quiet + (long)c.quiet_sel[sp - c.n_active] * c.n;
    w = c.quiet_weight;
  }
  double amin = HUGE_VAL, omax = -60000.0;
  for (int e = threadIdx.x; e < span; e += blockDim.x) {
    long m = i0 - c.wid + e;
    double v = 1.0;
    if (m >= 0 && m < c.n) {
      double A = D ? Q[m] - D[m] : Q[m];
      v = 1.0 - A;
      if (m >= i0 && m < i0 + kLTPB) amin = A < amin ? A : amin;
    }
    xs[e] = v;
  }
  __syncthreads();
  long i = i0 + threadIdx.x;
  double o = 0.0;
  if (i < c.n && i > c.wid && i + c.wid < c.n - 1) {
    double sig = -1 * c.wave[i] * c.c1;  // Delta_lambda(sigma, lambda_i)
    double aa = 0.005 * 0.005 / (2 * sig * sig);
    double v = exp(-2 * aa);
    const double *x = xs + threadIdx.x + c.wid;  // x[d] = 1 - A_{i+d}
    double acc = x[0];
    double g = 1.0, u = exp(-aa);
    for (int d = 1; d <= c.wid; d++) {
      g *= u;
      u *= v;
      double pair = x[-d];
      if (d <= c.wid - 2) pair += x[d];
      acc = fma(g, pair, acc);
    }
    o = w * acc / (sig * sqrt(2. * kPi));
  }
  if (i < c.n) {
    o_out[sp * c.n + i] = o;
    if (o != 0.0 && omax < o) omax = o;
  }
  // block min / max partials
  __shared__ double s_min[kLTPB / 64], s_max[kLTPB / 64];
  for (int off = 32; off > 0; off >>= 1) {
    double x1 = __shfl_xor(amin, off), x2 = __shfl_xor(omax, off);
    amin = x1 < amin ? x1 : amin;
    omax = x2 > omax ? x2 : omax;
  }
  int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane == 0) {
    s_min[wv] = amin;
    s_max[wv] = omax;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double a0 = s_min[0], b0 = s_max[0];
    for (int k = 1; k < kLTPB / 64; k++) {
      a0 = s_min[k] < a0 ? s_min[k] : a0;
      b0 = s_max[k] > b0 ? s_max[k] : b0;
    }
    part[(sp * c.n_tiles + blockIdx.x) * 2] = a0;
    part[(sp * c.n_tiles + blockIdx.x) * 2 + 1] = b0;
  }
}

hipError_t launch_conv_linear(const ConvLinArgs &c, int n_spec, const double *quiet, const double *raw, double *o_out,
                              double *part, hipStream_t s) {
  if (n_spec == 0) return hipSuccess;
  size_t lds = (size_t)(kLTPB + 2 * c.wid) * sizeof(double);
  if (lds > 65536) return hipErrorInvalidValue;
  dim3 grid((unsigned)c.n_tiles, (unsigned)n_spec);
  k_conv_linear<<<grid, kLTPB, lds, s>>>(c, quiet, raw, o_out, part);
  return hipGetLastError();
}

int conv_linear_tile() { return kLTPB; }

// ------------------------------------------------------------ CCF mask
// calc_mask on a linear grid with the reference's literal index arithmetic
// (CCF_func.py:45-91): pixel positions on the extended grid of run_all_SOAP.py:
// 292-295 (no_points = 20000 on each side, step 0.005 in the index formula),
// converted to trimmed-spectrum positions (the slice mask[20000:]); lines with
// aa < 1 or bb + 1 >= n_ext are skipped (len 0).
__device__ __forceinline__ double ext_wave(const double *wt, long n_t, double step, long np, long k) {
  double begin = wt[0] - step * np;
  if (k < np) {
    if (k == np - 1) return wt[0];
    double st = (wt[0] - begin) / (double)(np - 1);
    return (double)k * st + begin;
  }
  if (k < np + n_t) return wt[k - np];
  long kk = k - np - n_t;
  double stop = wt[n_t - 1] + step * np;
  if (kk == np - 1) return stop;
  double st = (stop - wt[n_t - 1]) / (double)(np - 1);
  return (double)kk * st + wt[n_t - 1];
}

__global__ void k_mask_build_linear(const double *line_l, const double *line_w, const int32_t *sorted, int n_lines,
                                    const double *vgrid, int n_v, const double *wave_t, long n_t, double step,
                                    int32_t *m_start, int32_t *m_len, double2 *m_wv) {
  long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long)n_v * n_lines) return;
  int q = (int)(e / n_v), iv = (int)(e % n_v);
  int li = sorted[q];
  const long np = (long)(100.0 / step);
  const long n_ext = n_t + 2 * np;
  double begin_wave = wave_t[0] - step * np;
  double ls = line_l[li] * (1 + vgrid[iv] / kC);
  double beta = 820. / kC;
  double hw = ls * (sqrt((1 + beta) / (1 - beta)) - 1);
  double bgn = ls - hw / 2., end = ls + hw / 2.;
  long aa = (long)ceil((bgn - begin_wave) / 0.005);
  long bb = (long)(ceil((end - begin_wave) / 0.005) - 1);
  double w = line_w[li];
  if (aa < 1 || bb + 1 >= n_ext || bb - aa + 2 > 24 || bb - aa + 2 < 1) {
    m_start[e] = 0;
    m_len[e] = 0;
    m_wv[e] = make_double2(0.0, 0.0);
    return;
  }
  double fb = fabs(ext_wave(wave_t, n_t, step, np, aa) - bgn) /
              (ext_wave(wave_t, n_t, step, np, aa) - ext_wave(wave_t, n_t, step, np, aa - 1));
  double fa = fabs(ext_wave(wave_t, n_t, step, np, bb) - end) /
              (ext_wave(wave_t, n_t, step, np, bb + 1) - ext_wave(wave_t, n_t, step, np, bb));
  m_start[e] = (int32_t)(aa - 1 - 20000);
  m_len[e] = (int32_t)(bb - aa + 2);
  m_wv[e] = make_double2(w * fb, w * fa);
}

hipError_t launch_mask_build_linear(const double *line_l, const double *line_w, const int32_t *sorted, int n_lines,
                                    const double *vgrid, int n_v, const double *wave_t, int64_t n_t, double step,
                                    int32_t *m_start, int32_t *m_len, double2 *m_wv, hipStream_t s) {
  long tot = (long)n_v * n_lines;
  if (tot == 0) return hipSuccess;
  k_mask_build_linear<<<(unsigned)((tot + kLTPB - 1) / kLTPB), kLTPB, 0, s>>>(
      line_l, line_w, sorted, n_lines, vgrid, n_v, wave_t, n_t, step, m_start, m_len, m_wv);
  return hipGetLastError();
}

}  // namespace soap
