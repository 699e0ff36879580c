// gfx950 kernels of the SOAP-GPU engine.  All arithmetic FP64.
//
// Hot-path design (log-uniform grids): on lambda_j = lambda_0 exp(j h) a Doppler
// shift is a constant fractional pixel offset s = log1p(fbeta)/h, so the
// reference's per-cell "shift + linear interpolation" (SOAP_integration.cu:232-281)
// equals (1-t) S[j-m] + t S[j-m+1] with m = ceil(s), t = expm1((m-s)h)/expm1(h)
// for every pixel j.  Summing cells therefore reduces to a short FIR filter per
// mu bin whose taps are the limb-weighted (1-t, t) of all cells with that shift:
//   1. k_disk_init   : per-cell maps + (m, t)                       (a1)
//   2. k_taps_*      : exact fixed-point accumulation of the taps   (a2/a5/a6)
//   3. k_fir_*       : lambda-streaming FIR over the seed rows       (a2/a6)
//   4. k_conv/k_norm : Gaussian instrumental convolution + spec_norm (a7)
//   5. k_mask_*/k_ccf_lds: sparse mask CCF                           (a8)
//   6. k_fit         : Gaussian LM fit + bisector, one wave per CCF  (a9)
#include <hip/hip_runtime.h>
#include <math.h>

#include "soap_internal.h"

namespace soap {

constexpr int kTPB = 256;

// Global-memory views of pointers.  A pointer loaded from memory (a field of a
// batch job) is generic to the compiler, whose accesses then become flat
// instructions and uniform loads cannot use the scalar unit; the kernel bodies
// below access device buffers through these views so that the single-instance
// kernels (pointers in kernel arguments) and the batched ones compile alike.
#define GLOBAL __attribute__((address_space(1)))
template <class T>
__device__ __forceinline__ const GLOBAL T *gl(const T *p) {
  return (const GLOBAL T *)p;
}
template <class T>
__device__ __forceinline__ GLOBAL T *gw(T *p) {
  return (GLOBAL T *)p;
}

// ------------------------------------------------------------------ utilities
__global__ void k_fill_u64(unsigned long long *p, size_t count) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < count; i += stride) p[i] = 0ull;
}

__global__ void k_fill_f64(double *p, size_t count, double v) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < count; i += stride) p[i] = v;
}

static unsigned grid_for(size_t count) {
  size_t b = (count + kTPB - 1) / kTPB;
  if (b > 65535 * 8) b = 65535 * 8;
  if (b == 0) b = 1;
  return (unsigned)b;
}

hipError_t launch_fill_u64(unsigned long long *p, size_t count, hipStream_t s) {
  if (count == 0) return hipSuccess;
  k_fill_u64<<<grid_for(count), kTPB, 0, s>>>(p, count);
  return hipGetLastError();
}

hipError_t launch_fill_f64(double *p, size_t count, double v, hipStream_t s) {
  if (count == 0) return hipSuccess;
  k_fill_f64<<<grid_for(count), kTPB, 0, s>>>(p, count, v);
  return hipGetLastError();
}

// ------------------------------------------------------------------ a1: disk
// itot's cell loop, SOAP_initialize.cu:339-395.  Contraction is off so every
// map value rounds exactly as the reference's separate multiply/add sequence.
__global__ void k_disk_init(DiskArgs a) {
#pragma clang fp contract(off)
  long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  long ncell = (long)a.grid * a.grid;
  if (idx >= ncell) return;
  int iy = (int)(idx / a.grid), iz = (int)(idx % a.grid);
  double y = -1. + iy * a.delta_grid;
  double z = -1. + iz * a.delta_grid;
  double theta_z = -90.0 + iz * a.theta_grid;
  double s = sin(theta_z * kPi / 180);
  double omega_velo = (a.omega0 + a.omega1 * (s * s)) * kPi / (180 * 24 * 3600) * a.radius;
  double delta = y * omega_velo * a.sin_incl;
  double mu = -999.0, dl = -999.0, limb = -999.0, lb = -999.0;
  int32_t k = 0, m = 0;
  double t = 0.0;
  if ((y * y + z * z) <= 1.) {
    double r_cos = sqrt(1. - (y * y + z * z));
    limb = 1. - a.u1 * (1 - r_cos) - a.u2 * (1 - r_cos) * (1 - r_cos);
    lb = a.lb0 + a.lb1 * r_cos + a.lb2 * r_cos * r_cos + a.lb3 * r_cos * r_cos * r_cos;
    k = a.n_mu - 1;
    for (int b = 0; b < a.n_mu - 1; b++)
      if (r_cos > a.mu_bounds[b]) {
        k = b;
        break;
      }
    mu = r_cos;
    dl = delta;
    if (a.log_grid) {
      double beta = delta * 1000. / kC;
      double fbeta = -1 * (1 - sqrt((1 + beta) / (1 - beta)));
      double sh = log1p(fbeta) / a.h;
      double mc = ceil(sh);
      m = (int32_t)mc;
      t = expm1((mc - sh) * a.h) / a.expm1_h;
    }
  }
  a.mu[idx] = mu;
  a.delta[idx] = dl;
  a.limb[idx] = limb;
  a.lbm[idx] = lb;
  a.region[idx] = k;
  if (a.m) {
    a.m[idx] = m;
    a.t[idx] = t;
  }
}

hipError_t launch_disk_init(const DiskArgs &a, hipStream_t s) {
  long ncell = (long)a.grid * a.grid;
  k_disk_init<<<(unsigned)((ncell + kTPB - 1) / kTPB), kTPB, 0, s>>>(a);
  return hipGetLastError();
}

// Shared part of the tap kernels: accumulate into an LDS window of WD tap offsets
// [base, base+WD) per (set, bin); offsets outside it go straight to global.
#define LDSMEM __attribute__((address_space(3)))
struct TapWindow {
  LDSMEM unsigned long long *lds;  // [nsets][n_mu][wd][2]
  GLOBAL unsigned long long *gacc; // [nsets][n_mu][nd][2] of this job
  int base, wd, nd, dmin, n_mu;
  __device__ void add(int set, int k, int d, double x) const {
    if (x == 0.0) return;
    int w = d - base;
    if (w >= 0 && w < wd)
      fx_add(lds + (((long)set * n_mu + k) * wd + w) * 2, x);
    else if (d - dmin >= 0 && d - dmin < nd)  // always true: the host bounds every shift
      fx_add(gacc + (((long)set * n_mu + k) * nd + (d - dmin)) * 2, x);
  }
};

__device__ int block_min_int(int v, int *scratch) {
  for (int off = 32; off > 0; off >>= 1) {
    int o = __shfl_xor(v, off);
    v = o < v ? o : v;
  }
  int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane == 0) scratch[wv] = v;
  __syncthreads();
  int nw = (blockDim.x + 63) >> 6;
  int r = scratch[0];
  for (int i = 1; i < nw; i++) r = scratch[i] < r ? scratch[i] : r;
  __syncthreads();
  return r;
}

__device__ void flush_window(const TapWindow &w, int nsets) {
  __syncthreads();
  long total = (long)nsets * w.n_mu * w.wd;
  for (long e = threadIdx.x; e < total; e += blockDim.x) {
    unsigned long long hi = w.lds[e * 2], lo = w.lds[e * 2 + 1];
    if (hi == 0ull && lo == 0ull) continue;
    int wi = (int)(e % w.wd);
    long sk = e / w.wd;  // set * n_mu + k
    long g = (sk * w.nd + (w.base + wi - w.dmin)) * 2;
    __hip_atomic_fetch_add(w.gacc + g, hi, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_fetch_add(w.gacc + g + 1, lo, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

// Quiet-star taps: every disk cell contributes limb*(1-t) at offset m and limb*t
// at m-1 of its mu bin (itot's cell loop, SOAP_initialize.cu:339-395).
__global__ void k_taps_quiet(TapAccArgs a) {
  extern __shared__ unsigned long long lds_u64[];
  __shared__ int scratch[16];
  int G = a.grid;
  int iy0 = blockIdx.x * a.cols_per_block;
  int iy1 = min(G, iy0 + a.cols_per_block);
  long ncell = (long)(iy1 - iy0) * G;
  int mn = 1 << 30;
  for (long c = threadIdx.x; c < ncell; c += blockDim.x) {
    long idx = (long)iy0 * G + c;
    if (!(a.limb[idx] == -999.0 && a.delta[idx] == -999.0)) mn = min(mn, a.m[idx] - 1);
  }
  int base = block_min_int(mn, scratch);
  TapWindow w{(LDSMEM unsigned long long *)lds_u64, gw(a.acc), base, a.wd, a.nd, a.dmin, a.n_mu};
  long lds_n = (long)a.nsets * a.n_mu * a.wd * 2;
  for (long e = threadIdx.x; e < lds_n; e += blockDim.x) lds_u64[e] = 0ull;
  __syncthreads();
  for (long c = threadIdx.x; c < ncell; c += blockDim.x) {
    long idx = (long)iy0 * G + c;
    if (a.limb[idx] == -999.0 && a.delta[idx] == -999.0) continue;
    double L = a.limb[idx], t = a.t[idx];
    int k = a.region[idx], m = a.m[idx];
    w.add(0, k, m, L * (1.0 - t));
    w.add(0, k, m - 1, L * t);
  }
  flush_window(w, a.nsets);
}

hipError_t launch_taps_quiet(const TapAccArgs &a, hipStream_t s) {
  int nb = (a.grid + a.cols_per_block - 1) / a.cols_per_block;
  size_t lds = (size_t)a.nsets * a.n_mu * a.wd * 2 * sizeof(unsigned long long);
  k_taps_quiet<<<nb, kTPB, lds, s>>>(a);
  return hipGetLastError();
}

// Active taps of every phase (job = phase): spot cells add limb to set 0,
// facula cells add limb to set 1 and (1+LB) limb to set 2 (the intensity factor
// of fast_intensity_fac, SOAP_integration.cu:317-323,424-440).
// Solar mode (SOAP_integration_solar.cu:424-433): the facula intensity is
// B(lambda, T_cell)/B(lambda, T_norm) C_fac(lambda), a smooth function of T_cell
// over [t_sum - 216, t_sum].  It is replaced by its Lagrange interpolant on K
// Chebyshev nodes T_q (the host checks the interpolation error), so facula cells
// add limb to set 1 and limb * l_q(T_cell) to set 2+q; the FIR epilogue then
// weights set 2+q by B(lambda, T_q)/B(lambda, T_norm).
__device__ __forceinline__ void taps_active_body(const TapAccArgs &a, int bx_, int by_) {
  extern __shared__ unsigned long long lds_u64[];
  __shared__ int scratch[16];
  int job = by_;
  int G = a.grid;
  const GLOBAL int32_t *region = gl(a.region), *mcell = gl(a.m), *geo_off = gl(a.geo_off);
  const GLOBAL double *limb = gl(a.limb), *delta = gl(a.delta), *tcell = gl(a.t), *lbm = gl(a.lbm);
  const GLOBAL RegionDev *geo = gl(a.geo);
  const GLOBAL int32_t *bx = gl(a.box) + job * 4;
  int cy0 = bx[0], cy1 = bx[1], cz0 = bx[2], cz1 = bx[3];
  int iy0 = cy0 + bx_ * a.cols_per_block;
  if (iy0 >= cy1) return;  // uniform over the block
  int iy1 = min(cy1, iy0 + a.cols_per_block);
  int nz = cz1 - cz0;
  long ncell = (long)(iy1 - iy0) * nz;
  int g0 = geo_off[job], g1 = geo_off[job + 1];
  double dg = 2. / G;
  int mn = 1 << 30;
  for (long c = threadIdx.x; c < ncell; c += blockDim.x) {
    long idx = (long)(iy0 + c / nz) * G + (cz0 + c % nz);
    if (!(limb[idx] == -999.0 && delta[idx] == -999.0)) mn = min(mn, mcell[idx] - 1);
  }
  int base = block_min_int(mn, scratch);
  TapWindow w{(LDSMEM unsigned long long *)lds_u64, gw(a.acc) + (long)job * a.nsets * a.n_mu * a.nd * 2, base, a.wd,
              a.nd, a.dmin, a.n_mu};
  long lds_n = (long)a.nsets * a.n_mu * a.wd * 2;
  for (long e = threadIdx.x; e < lds_n; e += blockDim.x) lds_u64[e] = 0ull;
  __syncthreads();
  for (long c = threadIdx.x; c < ncell; c += blockDim.x) {
    int iy = iy0 + (int)(c / nz), iz = cz0 + (int)(c % nz);
    long idx = (long)iy * G + iz;
    double rc;
    int ty = classify_cell(iy, iz, dg, geo, g0, g1, &rc);
    if (ty < 0) continue;
    double L = limb[idx], t = tcell[idx];
    int k = region[idx], m = mcell[idx];
    if (ty == 0) {
      w.add(0, k, m, L * (1.0 - t));
      w.add(0, k, m - 1, L * t);
    } else if (a.solar_nodes == 0) {
      double L2 = (1 + lbm[idx]) * L;
      w.add(1, k, m, L * (1.0 - t));
      w.add(1, k, m - 1, L * t);
      w.add(2, k, m, L2 * (1.0 - t));
      w.add(2, k, m - 1, L2 * t);
    } else {
      w.add(1, k, m, L * (1.0 - t));
      w.add(1, k, m - 1, L * t);
      double T = (double)plage_temp(a.t_sum, rc);
      for (int q = 0; q < a.solar_nodes; q++) {
        double l = a.node_w[q];
        for (int r = 0; r < a.solar_nodes; r++)
          if (r != q) l *= T - a.node_t[r];
        double Lq = L * l;
        w.add(2 + q, k, m, Lq * (1.0 - t));
        w.add(2 + q, k, m - 1, Lq * t);
      }
    }
  }
  flush_window(w, a.nsets);
}

__global__ void k_taps_active(TapAccArgs a) {
  taps_active_body(a, blockIdx.x, blockIdx.y);
}

hipError_t launch_taps_active(const TapAccArgs &a, int n_job, int max_col_blocks, hipStream_t s) {
  if (n_job == 0 || max_col_blocks == 0) return hipSuccess;
  size_t lds = (size_t)a.nsets * a.n_mu * a.wd * 2 * sizeof(unsigned long long);
  dim3 grid(max_col_blocks, n_job);
  k_taps_active<<<grid, kTPB, lds, s>>>(a);
  return hipGetLastError();
}

// Fixed point -> double taps, per-(job, bin) nonzero d range, flags, bin list.
// Active taps get one more output set than they accumulate: the last one is
// set 0 + set 1 (all active cells, weight limb), used on the SED quiet rows.
// Flags: bit 0 the spot set is nonzero, bit 1 a facula set is nonzero.
__device__ __forceinline__ void taps_finalize_body(const unsigned long long *acc_, int nsets_in, int nsets_out, int n_mu,
                                                   int nd, double *taps_, int32_t *krange_, int32_t *kflags_,
                                                   int32_t *bins_, int32_t *nbins_, double *taps_g_, int n_job,
                                                   int p_pad, int bx_, int by_) {
  const GLOBAL unsigned long long *acc = gl(acc_);
  GLOBAL double *taps = gw(taps_);
  GLOBAL int32_t *krange = gw(krange_), *kflags = gw(kflags_), *bins = gw(bins_), *nbins = gw(nbins_);
  int job = bx_;
  if (job >= n_job) {  // padding jobs of the interleaved table: zero taps
    if (taps_g_)
      for (long e = threadIdx.x; e < (long)n_mu * nd; e += blockDim.x)
        for (int st = 0; st < nsets_out; st++) gw(taps_g_)[(e * p_pad + job) * nsets_out + st] = 0.0;
    return;  // uniform over the block
  }
  // one thread per (bin, offset): the nonzero range and the populations of a bin
  // gathered with LDS atomics (min, max and or do not depend on the order)
  __shared__ int s_lo[256], s_hi[256], s_fl[256];
  for (int k = threadIdx.x; k < n_mu; k += blockDim.x) {
    s_lo[k] = nd;
    s_hi[k] = -1;
    s_fl[k] = 0;
  }
  __syncthreads();
  for (int e = threadIdx.x; e < n_mu * nd; e += blockDim.x) {
    const int k = e / nd, d = e - k * nd;
    bool nz = false;
    int fl = 0;
    double v01 = 0.0;
    GLOBAL double *out = taps + (((long)job * n_mu + k) * nd + d) * nsets_out;
    GLOBAL double *outg = taps_g_ ? gw(taps_g_) + (((long)k * nd + d) * p_pad + job) * nsets_out : nullptr;
    for (int st = 0; st < nsets_in; st++) {
      long ea = ((((long)job * nsets_in + st) * n_mu + k) * nd + d) * 2;
      double v = fx_value(acc[ea], acc[ea + 1]);
      out[st] = v;
      if (outg) outg[st] = v;
      if (st < 2) v01 += v;
      if (v != 0.0) {
        nz = true;
        fl |= (st == 0 || nsets_in == 1) ? 1 : 2;
      }
    }
    if (nsets_out > nsets_in) {
      out[nsets_in] = v01;
      if (outg) outg[nsets_in] = v01;
    }
    if (nz) {
      atomicMin(&s_lo[k], d);
      atomicMax(&s_hi[k], d);
      atomicOr(&s_fl[k], fl);
    }
  }
  __syncthreads();
  for (int k = threadIdx.x; k < n_mu; k += blockDim.x) {
    krange[((long)job * n_mu + k) * 2] = s_lo[k];
    krange[((long)job * n_mu + k) * 2 + 1] = s_hi[k];
    kflags[(long)job * n_mu + k] = s_fl[k];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    int c = 0;
    for (int k = 0; k < n_mu; k++)
      if (s_lo[k] <= s_hi[k]) bins[(long)job * n_mu + c++] = k;
    nbins[job] = c;
  }
}

__global__ void k_taps_finalize(const unsigned long long *acc, int nsets_in, int nsets_out, int n_mu, int nd,
                                double *taps, int32_t *krange, int32_t *kflags, int32_t *bins, int32_t *nbins,
                                double *taps_g, int n_job, int p_pad) {
  taps_finalize_body(acc, nsets_in, nsets_out, n_mu, nd, taps, krange, kflags, bins, nbins, taps_g, n_job, p_pad,
                     blockIdx.x, blockIdx.y);
}

hipError_t launch_taps_finalize(const unsigned long long *acc, int n_job, int nsets_in, int nsets_out, int n_mu,
                                int nd, double *taps, int32_t *krange, int32_t *kflags, int32_t *bins,
                                int32_t *nbins, double *taps_g, int p_pad, hipStream_t s) {
  if (n_mu > 256 || (taps_g && p_pad < n_job + 8)) return hipErrorInvalidValue;
  k_taps_finalize<<<taps_g ? p_pad : n_job, 256, 0, s>>>(acc, nsets_in, nsets_out, n_mu, nd, taps, krange, kflags,
                                                        bins, nbins, taps_g, n_job, p_pad);
  return hipGetLastError();
}

// ------------------------------------------------------------------ a2/a6: FIR
__device__ __forceinline__ double row_at(const double *row, long i, long n) {
  return (i >= 0 && i < n) ? row[i] : 0.0;
}

// Quiet star: Q_j = sum_k sum_d H_k[d] Sq_k[j-d], the SED sum likewise.
__global__ void k_fir_quiet(FirArgs a) {
  long j = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= a.n) return;
  double accq = 0.0, accs = 0.0;
  int nb = a.nbins[0];
  for (int kk = 0; kk < nb; kk++) {
    int k = a.bins[kk];
    int lo = a.krange[k * 2], hi = a.krange[k * 2 + 1];
    const double *T = a.taps + (long)k * a.nd;
    const double *rq = a.q + (long)k * a.n;
    const double *rs = a.qs ? a.qs + (long)k * a.n : nullptr;
    for (int di = lo; di <= hi; di++) {
      double hv = T[di];
      long i = j - (a.dmin + di);
      accq += hv * row_at(rq, i, a.n);
      if (rs) accs += hv * row_at(rs, i, a.n);
    }
  }
  a.out[j] = accq;
  a.out[a.n + j] = accs;
}

hipError_t launch_fir_quiet(const FirArgs &a, hipStream_t s) {
  k_fir_quiet<<<(unsigned)((a.n + kTPB - 1) / kTPB), kTPB, 0, s>>>(a);
  return hipGetLastError();
}

// Active deltas of one phase per blockIdx.y (spot_integrate's fast_flux /
// fast_bconv / fast_tot summed over cells, SOAP_integration.cu:285-307,392-447):
//   flux  = H1a*SEDq - Cs (H1s*Sq) - Cf (H2f*Sq)
//   bconv = H1s*Sq + H1f*Sq - H1s*Ssp - H1f*Sfac
//   tot   = H1a*SEDq - Cs (H1s*Ssp) - Cf (H2f*Sfac)
// with Cs, Cf the unshifted row 0 of the SED spot / facula cubes.
__global__ void k_fir_active(FirArgs a) {
  int job = blockIdx.y;
  long j = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= a.n) return;
  double F1 = 0, F2 = 0, F3 = 0, F4 = 0, F5 = 0, F6 = 0, F7 = 0;
  int nb = a.nbins[job];
  const int32_t *bl = a.bins + (long)job * a.n_mu;
  for (int kk = 0; kk < nb; kk++) {
    int k = bl[kk];
    long jk = (long)job * a.n_mu + k;
    int lo = a.krange[jk * 2], hi = a.krange[jk * 2 + 1];
    int fl = a.kflags[jk];
    const double *T = a.taps + jk * 4 * a.nd;
    const double *rq = a.q + (long)k * a.n, *rs = a.qs + (long)k * a.n;
    for (int di = lo; di <= hi; di++) {
      long i = j - (a.dmin + di);
      if (i < 0 || i >= a.n) continue;
      double h0 = T[di * 4], h1 = T[di * 4 + 1], h2 = T[di * 4 + 2], h3 = T[di * 4 + 3];
      double vq = rq[i];
      F1 += h3 * rs[i];
      F2 += h0 * vq;
      F4 += h1 * vq;
      F3 += h2 * vq;
      if (fl & 1) F5 += h0 * a.sp[(long)k * a.n + i];
      if (fl & 2) {
        double vf = a.fa[(long)k * a.n + i];
        F6 += h1 * vf;
        F7 += h2 * vf;
      }
    }
  }
  double cs = a.sps ? a.sps[j] : 0.0;
  double cf = a.fas ? a.fas[j] : 0.0;
  double *o = a.out + (long)job * 3 * a.n;
  o[j] = F1 - cs * F2 - cf * F3;
  o[a.n + j] = (F2 + F4) - F5 - F6;
  o[2 * a.n + j] = F1 - cs * F5 - cf * F7;
}

hipError_t launch_fir_active(const FirArgs &a, int n_job, hipStream_t s) {
  if (n_job == 0) return hipSuccess;
  dim3 grid((unsigned)((a.n + kTPB - 1) / kTPB), n_job);
  k_fir_active<<<grid, kTPB, 0, s>>>(a);
  return hipGetLastError();
}

// Phase-group FIR (bin outer, phases inner).  Block = 256-pixel tile x group of
// kFirG phases; every pixel's 7 partial sums of the kFirG phases live in
// registers.  For each mu bin touched by any phase of the group, the block
// stages that bin's seed-row segments (quiet, quiet SED, and spot / facula rows
// when present) in LDS; each staged row value then feeds the FMAs of all phases
// of the group, with the taps read by uniform (scalar) loads from the
// phase-interleaved copy of the tap table (the group's taps at one offset are
// contiguous: two wide scalar loads from one pointer).  The inner loop is
// branch-free: all seven products of every phase slot, over the union of the
// group's tap ranges (taps are zero where a phase has no cells), which costs
// some zero FMAs but no per-phase flag branches.  Blocks are numbered so the
// groups of a tile run back to back on one XCD (b % 8), keeping the tile's rows
// in its L2.
#ifndef SOAP_FIR_G
#define SOAP_FIR_G 4
#endif

constexpr int kFirG = SOAP_FIR_G;  // phases per block
constexpr int kFirTile = kTPB;

__device__ __forceinline__ int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }

template <int G, bool PRE>
__device__ __forceinline__ void fir_active_grp_body(const FirArgs &a, int n_job, int n_tiles, int n_groups, int bx_) {
  extern __shared__ double rows[];  // [4][row_len]: quiet, quiet SED, spot, facula rows of one bin
  __shared__ int4 s_meta[64];  // bins in use, in order: (bin, union flags, union tap range)
  __shared__ int s_nb;
  // 32-bit block arithmetic (block counts are < 2^31): keeps the division short
  const int xcd = bx_ & 7, slot = bx_ >> 3;
  const int tile = (slot / n_groups) * 8 + xcd;
  const int grp = slot % n_groups;
  if (tile >= n_tiles) return;  // uniform over the block
  int p0 = grp * G;
  int ng = min(G, n_job - p0);
  int nd = a.nd, n_mu = a.n_mu;
  int row_len = kFirTile + nd - 1;
  const GLOBAL int32_t *kflags = gl(a.kflags), *krange = gl(a.krange);
  long j0 = (long)tile * kFirTile;
  long src0 = j0 - a.dmin - nd + 1;
  // the spot and facula continua of this pixel, for the epilogue (loaded now, so
  // that they arrive during the bins)
  const long j = j0 + threadIdx.x;
  const double cs = (j < a.n && a.sps) ? gl(a.sps)[j] : 0.0;
  const double cf = (j < a.n && a.fas) ? gl(a.fas)[j] : 0.0;
  // group metadata, in the first wave: lane k (mu bin k) loads the flags and tap
  // ranges of the group's phases (all loads issued together), forms the union flags
  // and range, and the bins in use are listed in order through a ballot
  if (threadIdx.x < 64) {
    const int k = threadIdx.x;
    int fl[G], lo[G], hi[G];
#pragma unroll
    for (int g = 0; g < G; g++) {
      fl[g] = 0;
      lo[g] = 0;
      hi[g] = -1;
      if (g < ng && k < n_mu) {
        const long jk = (long)(p0 + g) * n_mu + k;
        fl[g] = kflags[jk];
        lo[g] = krange[jk * 2];
        hi[g] = krange[jk * 2 + 1];
      }
    }
    int ufl = 0, ulo = nd, uhi = -1;
#pragma unroll
    for (int g = 0; g < G; g++)
      if (fl[g] && lo[g] <= hi[g]) {
        ufl |= fl[g] | 4;
        ulo = min(ulo, lo[g]);
        uhi = max(uhi, hi[g]);
      }
    const unsigned long long used = __ballot(ufl != 0);
    if (ufl) s_meta[__popcll(used & ((1ull << k) - 1ull))] = make_int4(k, ufl, ulo, uhi);
    if (k == 0) s_nb = __popcll(used);
  }
  __syncthreads();
  double F[G][7];
#pragma unroll
  for (int g = 0; g < G; g++)
#pragma unroll
    for (int f = 0; f < 7; f++) F[g][f] = 0.0;
  const int nb = uni(s_nb);
  // PRE: the next bin's rows are loaded into registers (the first kPre x kTPB values
  // of each row) while the current bin computes, hiding their latency when the tap
  // ranges are short; costs registers (occupancy) when they are long
  constexpr int kPre = 2;
  double pre[4][kPre];
  auto row_src = [&](int r, int k, int has_sp) {
    return gl(r == 0 ? a.q : r == 1 ? a.qs : (r == 2 && has_sp) ? a.sp : a.fa) + (long)k * a.n;
  };
  auto prefetch = [&](int kk) {
    const int4 m = s_meta[kk];
    const int k = uni(m.x);
    const int ufl = uni(m.y);
    const int has_sp = ufl & 1, nrow = 2 + has_sp + ((ufl >> 1) & 1);
#pragma unroll
    for (int r = 0; r < 4; r++) {
      if (r >= nrow) break;
      const GLOBAL double *src = row_src(r, k, has_sp);
#pragma unroll
      for (int m = 0; m < kPre; m++) {
        int x = threadIdx.x + m * kTPB;
        long i = src0 + x;
        pre[r][m] = (x < row_len && i >= 0 && i < a.n) ? src[i] : 0.0;
      }
    }
  };
  if (PRE && nb > 0) prefetch(0);
  for (int kk = 0; kk < nb; kk++) {
    const int4 m = s_meta[kk];
    const int k = uni(m.x);
    const int ufl = uni(m.y), ulo = uni(m.z), uhi = uni(m.w);
    if (kk > 0) __syncthreads();
    // rows present: quiet, quiet SED, then spot (if any spot cell) and facula (if any)
    const int has_sp = ufl & 1, has_fa = (ufl >> 1) & 1;
    const int nrow = 2 + has_sp + has_fa;
    if constexpr (PRE) {
#pragma unroll
      for (int r = 0; r < 4; r++) {
        if (r >= nrow) break;
#pragma unroll
        for (int m = 0; m < kPre; m++) {
          int x = threadIdx.x + m * kTPB;
          if (x < row_len) rows[r * row_len + x] = pre[r][m];
        }
        const GLOBAL double *src = row_src(r, k, has_sp);
        for (int x = threadIdx.x + kPre * kTPB; x < row_len; x += blockDim.x) {
          long i = src0 + x;
          rows[r * row_len + x] = (i >= 0 && i < a.n) ? src[i] : 0.0;
        }
      }
      __syncthreads();
      if (kk + 1 < nb) prefetch(kk + 1);
    } else {
      for (int r = 0; r < nrow; r++) {
        const GLOBAL double *src = row_src(r, k, has_sp);
        for (int x = threadIdx.x; x < row_len; x += blockDim.x) {
          long i = src0 + x;
          rows[r * row_len + x] = (i >= 0 && i < a.n) ? src[i] : 0.0;
        }
      }
      __syncthreads();
    }
    // taps of the group's phases at offset di: G x 4 consecutive doubles of the
    // phase-interleaved table (zero for phases past the last), so one uniform base
    // pointer per offset serves every phase slot
    const GLOBAL double *Tg = gl(a.taps_g) + (((long)k * nd + ulo) * a.p_pad + p0) * 4;
    const long dstride = (long)a.p_pad * 4;
    const double *rq = rows, *rs = rows + row_len, *rp = rows + 2 * row_len, *rf = rows + (2 + has_sp) * row_len;
    const int e0 = threadIdx.x + nd - 1 - ulo, n_off = uhi - ulo + 1;
    for (int i = 0; i < n_off; i++, Tg += dstride) {
      const int e = e0 - i;
      double vq = rq[e], vs = rs[e];
      double vp = (ufl & 1) ? rp[e] : 0.0;
      double vf = (ufl & 2) ? rf[e] : 0.0;
      // every product of every phase slot (taps are zero where a phase has no cells)
#pragma unroll
      for (int g = 0; g < G; g++) {
        const GLOBAL double *T = Tg + g * 4;
        const double t0 = T[0], t1 = T[1], t2 = T[2], t3 = T[3];
        F[g][0] = fma(t3, vs, F[g][0]);
        F[g][1] = fma(t0, vq, F[g][1]);
        F[g][4] = fma(t0, vp, F[g][4]);
        F[g][2] = fma(t2, vq, F[g][2]);
        F[g][3] = fma(t1, vq, F[g][3]);
        F[g][5] = fma(t1, vf, F[g][5]);
        F[g][6] = fma(t2, vf, F[g][6]);
      }
    }
  }
  if (j < a.n) {
#pragma unroll
    for (int g = 0; g < G; g++) {
      if (g >= ng) break;
      GLOBAL double *o = gw(a.out) + (long)(p0 + g) * 3 * a.n;
      // F: 0 = H1a*SEDq, 1 = H1s*Sq, 2 = H2f*Sq, 3 = H1f*Sq, 4 = H1s*Ssp, 5 = H1f*Sfac, 6 = H2f*Sfac
      o[j] = F[g][0] - cs * F[g][1] - cf * F[g][2];
      o[a.n + j] = (F[g][1] + F[g][3]) - F[g][4] - F[g][5];
      o[2 * a.n + j] = F[g][0] - cs * F[g][4] - cf * F[g][6];
    }
  }
}

template <bool PRE>
__global__ __launch_bounds__(kTPB) void k_fir_active_grp(FirArgs a, int n_job, int n_tiles, int n_groups) {
  fir_active_grp_body<kFirG, PRE>(a, n_job, n_tiles, n_groups, blockIdx.x);
}

// shift ranges up to this many offsets take the row-prefetching form
constexpr int kFirPrefetchMaxNd = 64;

hipError_t launch_fir_active_grp(const FirArgs &a, int n_job, hipStream_t s) {
  if (n_job == 0) return hipSuccess;
  if (a.n_mu > 64) return hipErrorInvalidValue;
  int n_tiles = (int)((a.n + kFirTile - 1) / kFirTile);
  int n_groups = (n_job + kFirG - 1) / kFirG;
  long tiles8 = ((long)n_tiles + 7) / 8 * 8;
  long blocks = tiles8 * n_groups;
  size_t lds = (size_t)4 * (kFirTile + a.nd - 1) * sizeof(double);
  if (lds > 60000) return hipErrorInvalidValue;
  if (a.nd <= kFirPrefetchMaxNd)
    k_fir_active_grp<true><<<(unsigned)blocks, kTPB, lds, s>>>(a, n_job, n_tiles, n_groups);
  else
    k_fir_active_grp<false><<<(unsigned)blocks, kTPB, lds, s>>>(a, n_job, n_tiles, n_groups);
  return hipGetLastError();
}

// Tile-stationary FIR.  Block = a tile of 64 x T64 pixels and every phase.  The
// seed rows of all mu bins that any phase uses are staged in LDS once per block,
// interleaved per source pixel as (quiet, quiet SED, spot, facula) so that one
// pixel's four values are two 16-byte reads; after one barrier each wave walks its
// own phase groups (kTileG phases each) without further synchronisation.  A
// group's work is a stream of (bin, offset) entries in bin-then-offset order (the
// order of every other FIR form, so the sums are the same bit for bit), built once
// per launch by k_fir_prog: the group's 4 x kTileG taps at that (bin, offset) and
// the LDS byte offset of lane 0's source pixel.  Scalar (SMEM) and LDS loads share
// one counter and scalar loads return out of order, so each entry is waited for in
// full at the top of its step, the next entry's loads are issued, and the FMAs of
// the current one then run while those loads are in flight (registers A/B
// alternate).  Blocks are numbered so that each XCD takes a contiguous run of
// tiles, whose halos then come from its L2.
constexpr int kTileG = 4;      // phases per group

constexpr int kProgThreads = 1024;

// xscale: bytes per staged source pixel and bin row (32: the four rows interleaved;
// 16: two planes of row pairs); mult: the stream length is padded with zero entries
// (reading the last entry's staged pixels) to a multiple of mult, the entries a
// consuming loop takes per step
__device__ __forceinline__ void fir_prog_body(const FirArgs &a, const FirProg &pg, int n_job, int n_groups, int lmax,
                                              int xscale, int mult, int g) {
  const int n_mu = a.n_mu, nd = a.nd;
  const GLOBAL int32_t *kflags = gl(a.kflags), *krange = gl(a.krange);
  if (g == n_groups) {  // union over every phase: which rows the tiles stage
    __shared__ int s_u[64];
    const int k = threadIdx.x & 63, slice = threadIdx.x >> 6;  // 16 phase slices per bin
    if (threadIdx.x < 64) s_u[k] = 0;
    __syncthreads();
    int u = 0;
    if (k < n_mu)
      for (int p = slice; p < n_job; p += kProgThreads / 64) {
        const long jk = (long)p * n_mu + k;
        const int fl = kflags[jk];
        if (fl && krange[jk * 2] <= krange[jk * 2 + 1]) u |= fl | 4;
      }
    if (u) atomicOr(&s_u[k], u);
    __syncthreads();
    if (threadIdx.x < n_mu) gw(pg.ufl)[threadIdx.x] = s_u[threadIdx.x];
    return;
  }
  // the group's bins: wave 0, lane = mu bin, forms the union ranges and their
  // prefix (entry positions, bins in order) in LDS; then every thread of the block
  // builds whole entries (their 16 taps loaded together, then stored)
  __shared__ int s_pos[65], s_lo[64];
  const int p0 = g * kTileG;
  if (threadIdx.x < 64) {
    const int k = threadIdx.x;
    int ufl = 0, ulo = nd, uhi = -1;
    if (k < n_mu)
      for (int q = 0; q < kTileG; q++) {
        const int p = p0 + q;
        if (p >= n_job) break;
        const long jk = (long)p * n_mu + k;
        const int fl = kflags[jk], lo = krange[jk * 2], hi = krange[jk * 2 + 1];
        if (fl && lo <= hi) {
          ufl |= fl;
          ulo = min(ulo, lo);
          uhi = max(uhi, hi);
        }
      }
    const int cnt = ufl ? uhi - ulo + 1 : 0;
    int pos = cnt;  // inclusive prefix over the lanes (bins in order)
    for (int o = 1; o < 64; o <<= 1) {
      const int v = __shfl_up(pos, o);
      if (k >= o) pos += v;
    }
    const int total = __shfl(pos, 63);
    s_pos[k] = pos - cnt;
    s_lo[k] = ulo;
    if (k == 63) s_pos[64] = total;
    // zero entries up to a multiple of mult, reading the last entry's pixels
    const unsigned long long used = __ballot(cnt > 0);
    const int klast = used ? 63 - __clzll(used) : 0;
    const int hlast = __shfl(uhi, klast);
    const int padded = total ? (total + mult - 1) / mult * mult : 0;
    GLOBAL double *T = gw(pg.taps) + (long)g * pg.cap * (4 * kTileG);
    GLOBAL int32_t *X = gw(pg.xoff) + (long)g * pg.cap;
    if (k == 0) gw(pg.count)[g] = padded;
    if (k < padded - total) {
      X[total + k] = (klast * lmax + nd - 1 - hlast) * xscale;
      for (int t = 0; t < 4 * kTileG; t++) T[(long)(total + k) * (4 * kTileG) + t] = 0.0;
    }
  }
  __syncthreads();
  const int total = s_pos[64];
  const GLOBAL double *taps = gl(a.taps);
  GLOBAL double *T = gw(pg.taps) + (long)g * pg.cap * (4 * kTileG);
  GLOBAL int32_t *X = gw(pg.xoff) + (long)g * pg.cap;
  for (int e = threadIdx.x; e < total; e += blockDim.x) {
    int k = 0;  // the bin of entry e: the last bin whose first entry is <= e
    for (int step = 32; step > 0; step >>= 1)
      if (k + step < 64 && s_pos[k + step] <= e) k += step;
    const int di = s_lo[k] + (e - s_pos[k]);
    double v[4 * kTileG];
#pragma unroll
    for (int q = 0; q < kTileG; q++) {
      const int p = p0 + q;
      const GLOBAL double *src = taps + (((long)(p < n_job ? p : 0) * n_mu + k) * nd + di) * 4;
#pragma unroll
      for (int t = 0; t < 4; t++) v[q * 4 + t] = p < n_job ? src[t] : 0.0;
    }
    int kind = 0;  // bit 0: a spot tap (H1s) of the group is nonzero, bit 1: a facula tap (H1f, H2f)
#pragma unroll
    for (int q = 0; q < kTileG; q++) {
      kind |= v[q * 4] != 0.0 ? 1 : 0;
      kind |= (v[q * 4 + 1] != 0.0 || v[q * 4 + 2] != 0.0) ? 2 : 0;
    }
#pragma unroll
    for (int t = 0; t < 4 * kTileG; t++) T[(long)e * (4 * kTileG) + t] = v[t];
    // the DPP form carries the kind in bits 28-29 of the offset word (the
    // offsets are below 2^28 bytes; the scalar-tap form reads plain offsets)
    X[e] = (k * lmax + nd - 1 - di) * xscale | (xscale == 16 ? kind << 28 : 0);
  }
}

__global__ __launch_bounds__(kProgThreads) void k_fir_prog(FirArgs a, FirProg pg, int n_job, int n_groups, int lmax,
                                                           int xscale, int mult) {
  fir_prog_body(a, pg, n_job, n_groups, lmax, xscale, mult, blockIdx.x);
}

__device__ __forceinline__ void lgkm_wait_all() { __builtin_amdgcn_s_waitcnt(0xc07f); }  // lgkmcnt(0)

template <int T64, int NW>
__device__ __forceinline__ void fir_tile_body(const FirArgs &a, const double *__restrict__ ptaps,
                                              const int32_t *__restrict__ pxoff, const int32_t *__restrict__ pcount,
                                              const int32_t *__restrict__ pufl, double *__restrict__ out, int cap,
                                              int n_job, int n_groups, int n_tiles, int tiles_per_xcd) {
  extern __shared__ double2 srow[];  // [n_mu][lmax][2]: (q, qs), (sp, fa)
  constexpr int TP = 64 * T64;
  const int b = blockIdx.x;
  const int tile = (b & 7) * tiles_per_xcd + (b >> 3);
  if (tile >= n_tiles) return;  // uniform over the block
  const int nd = a.nd, n_mu = a.n_mu, lmax = TP + nd - 1;
  const long n = a.n;
  const long j0 = (long)tile * TP;
  const long s0 = j0 - a.dmin - (nd - 1);
  // staging: every (bin, source pixel) of the bins in use; rows a bin does not use
  // are zero (their taps are zero too)
  {
    const GLOBAL double *q = gl(a.q), *qs = gl(a.qs), *sp = gl(a.sp), *fa = gl(a.fa);
    const int tot = n_mu * lmax;
    for (int e = threadIdx.x; e < tot; e += 64 * NW) {
      const int k = e / lmax, x = e - k * lmax;
      const int u = pufl[k];
      if (!u) continue;
      const long i = s0 + x;
      const bool in = i >= 0 && i < n;
      const long o = (long)k * n + i;
      const double vq = in ? q[o] : 0.0, vs = in ? qs[o] : 0.0;
      const double vp = (in && (u & 1)) ? sp[o] : 0.0, vf = (in && (u & 2)) ? fa[o] : 0.0;
      srow[2 * e] = make_double2(vq, vs);
      srow[2 * e + 1] = make_double2(vp, vf);
    }
  }
  __syncthreads();
  const int w = uni(threadIdx.x >> 6), lane = threadIdx.x & 63;  // w: the same in every lane of a wave
  const int n_items = T64 * n_groups;
  for (int it = w; it < n_items; it += NW) {
    const int sub = T64 == 1 ? 0 : it % T64, g = T64 == 1 ? it : it / T64;
    const int px = sub * 64 + lane;
    const long j = j0 + px;
    const double cs = (j < n && a.sps) ? gl(a.sps)[j] : 0.0;
    const double cf = (j < n && a.fas) ? gl(a.fas)[j] : 0.0;
    const char *lb = (const char *)srow + px * 32;
    const double *T = ptaps + (long)g * cap * (4 * kTileG);
    const int32_t *X = pxoff + (long)g * cap;
    const int ne = pcount[g];
    double F[kTileG][7];
#pragma unroll
    for (int q = 0; q < kTileG; q++)
#pragma unroll
      for (int f = 0; f < 7; f++) F[q][f] = 0.0;
    auto loadt = [&](double(&t)[4 * kTileG], int m) {
      const double *p = T + (long)m * (4 * kTileG);
#pragma unroll
      for (int i = 0; i < 4 * kTileG; i++) t[i] = p[i];
    };
    auto loadr = [&](double2 &r0, double2 &r1, int x) {
      const double2 *p = (const double2 *)(lb + x);
      r0 = p[0];
      r1 = p[1];
    };
    auto fmas = [&](const double(&t)[4 * kTileG], const double2 r0, const double2 r1) {
      const double vq = r0.x, vs = r0.y, vp = r1.x, vf = r1.y;
#pragma unroll
      for (int q = 0; q < kTileG; q++) {
        const double t0 = t[q * 4], t1 = t[q * 4 + 1], t2 = t[q * 4 + 2], t3 = t[q * 4 + 3];
        F[q][0] = fma(t3, vs, F[q][0]);
        F[q][1] = fma(t0, vq, F[q][1]);
        F[q][4] = fma(t0, vp, F[q][4]);
        F[q][2] = fma(t2, vq, F[q][2]);
        F[q][3] = fma(t1, vq, F[q][3]);
        F[q][5] = fma(t1, vf, F[q][5]);
        F[q][6] = fma(t2, vf, F[q][6]);
      }
    };
    if (ne > 0) {  // ne is even
      double ta[4 * kTileG], tb[4 * kTileG];
      double2 a0, a1, b0, b1;
      const int last = ne - 1;
      loadt(ta, 0);
      loadr(a0, a1, X[0]);
      int xn = X[1];
      for (int m = 0; m < ne; m += 2) {
        lgkm_wait_all();
        loadt(tb, m + 1);
        loadr(b0, b1, xn);
        const int xn2 = X[min(m + 2, last)];
        __builtin_amdgcn_sched_barrier(0);
        fmas(ta, a0, a1);
        lgkm_wait_all();
        loadt(ta, min(m + 2, last));
        loadr(a0, a1, xn2);
        xn = X[min(m + 3, last)];
        __builtin_amdgcn_sched_barrier(0);
        fmas(tb, b0, b1);
      }
    }
    if (j < n) {
      const int p0 = g * kTileG;
#pragma unroll
      for (int q = 0; q < kTileG; q++) {
        if (p0 + q >= n_job) break;
        double *o = out + (long)(p0 + q) * 3 * n;
        // F: 0 = H1a*SEDq, 1 = H1s*Sq, 2 = H2f*Sq, 3 = H1f*Sq, 4 = H1s*Ssp, 5 = H1f*Sfac, 6 = H2f*Sfac
        o[j] = F[q][0] - cs * F[q][1] - cf * F[q][2];
        o[n + j] = (F[q][1] + F[q][3]) - F[q][4] - F[q][5];
        o[2 * n + j] = F[q][0] - cs * F[q][4] - cf * F[q][6];
      }
    }
  }
}

template <int T64, int NW>
__global__ __launch_bounds__(64 * NW) void k_fir_tile(FirArgs a, const double *__restrict__ ptaps,
                                                      const int32_t *__restrict__ pxoff,
                                                      const int32_t *__restrict__ pcount,
                                                      const int32_t *__restrict__ pufl, double *__restrict__ out,
                                                      int cap, int n_job, int n_groups, int n_tiles,
                                                      int tiles_per_xcd) {
  fir_tile_body<T64, NW>(a, ptaps, pxoff, pcount, pufl, out, cap, n_job, n_groups, n_tiles, tiles_per_xcd);
}

// Tile-stationary FIR with the taps in vector registers.  Scalar loads of the
// taps from L2 cost about a thousand cycles each and the scalar registers hold one
// entry ahead at most, so here every lane loads one tap of the entry (lane l: tap
// l mod 16, the four 16-lane rows of a wave alike) with vector loads issued four
// entries ahead, and each FMA broadcasts its tap from lane n of the lane's row
// (v_fmac_f64 with DPP row_newbcast:n, a single rounding like fma()).  The LDS
// holds two planes per bin: (quiet, quiet SED) and (spot, facula) pairs, so the
// two 16-byte reads of a step are each 64 consecutive pairs.
// The FMAs of one entry with the entry's kind (bit 0: a spot tap of the group is
// nonzero, bit 1: a facula tap) in a scalar register, the four phases together:
// H1a x quiet SED (F0) always, the spot products (F1, F4) behind one scalar branch
// and the facula products (F2, F3, F5, F6) behind another, so an entry of one
// population skips the other's FMAs (their taps are zero: the skipped products are
// exact zeros and every sum is the same).  Tap lanes: phase q's (H1s, H1f, H2f, H1a)
// in lanes 4q .. 4q + 3 of each 16-lane row.  Each accumulator takes one FMA per
// entry, so the sums are the same as with the FMAs in any other order within it.
__device__ __forceinline__ void fir_dpp_entry(double (&F)[kTileG][7], double t, const double2 r0, const double2 r1,
                                              int kind) {
  static_assert(kTileG == 4, "four phases per entry");
  // two wait states between a VALU write of the tap register and its DPP read
  asm volatile("s_nop 1" ::: "memory");
  asm("v_fmac_f64_dpp %0, %12, %14 row_newbcast:3 row_mask:0xf bank_mask:0xf\n\t"
      "v_fmac_f64_dpp %3, %12, %14 row_newbcast:7 row_mask:0xf bank_mask:0xf\n\t"
      "v_fmac_f64_dpp %6, %12, %14 row_newbcast:11 row_mask:0xf bank_mask:0xf\n\t"
      "v_fmac_f64_dpp %9, %12, %14 row_newbcast:15 row_mask:0xf bank_mask:0xf\n\t"
      "s_bitcmp1_b32 %16, 0\n\t"
      "s_cbranch_scc0 1f\n\t"
      "v_fmac_f64_dpp %1, %12, %13 row_newbcast:0 row_mask:0xf bank_mask:0xf\n\t"
      "v_fmac_f64_dpp %2, %12, %15 row_newbcast:0 row_mask:0xf bank_mask:0xf\n\t"
      "v_fmac_f64_dpp %4, %12, %13 row_newbcast:4 row_mask:0xf bank_mask:0xf\n\t"
      "v_fmac_f64_dpp %5, %12, %15 row_newbcast:4 row_mask:0xf bank_mask:0xf\n\t"
      "v_fmac_f64_dpp %7, %12, %13 row_newbcast:8 row_mask:0xf bank_mask:0xf\n\t"
      "v_fmac_f64_dpp %8, %12, %15 row_newbcast:8 row_mask:0xf bank_mask:0xf\n\t"
      "v_fmac_f64_dpp %10, %12, %13 row_newbcast:12 row_mask:0xf bank_mask:0xf\n\t"
      "v_fmac_f64_dpp %11, %12, %15 row_newbcast:12 row_mask:0xf bank_mask:0xf\n"
      "1:"
      : "+v"(F[0][0]), "+v"(F[0][1]), "+v"(F[0][4]), "+v"(F[1][0]), "+v"(F[1][1]), "+v"(F[1][4]), "+v"(F[2][0]),
        "+v"(F[2][1]), "+v"(F[2][4]), "+v"(F[3][0]), "+v"(F[3][1]), "+v"(F[3][4])
      : "v"(t), "v"(r0.x), "v"(r0.y), "v"(r1.x), "s"(kind)
      : "scc");
  asm("s_nop 1\n\t"  // the tap register may have been copied by a VALU move since the first block
      "s_bitcmp1_b32 %19, 1\n\t"
      "s_cbranch_scc0 2f\n\t"
      "v_fmac_f64_dpp %0, %16, %17 row_newbcast:2 row_mask:0xf bank_mask:0xf\n\t"
      "v_fmac_f64_dpp %1, %16, %17 row_newbcast:1 row_mask:0xf bank_mask:0xf\n\t"
      "v_fmac_f64_dpp %2, %16, %18 row_newbcast:1 row_mask:0xf bank_mask:0xf\n\t"
      "v_fmac_f64_dpp %3, %16, %18 row_newbcast:2 row_mask:0xf bank_mask:0xf\n\t"
      "v_fmac_f64_dpp %4, %16, %17 row_newbcast:6 row_mask:0xf bank_mask:0xf\n\t"
      "v_fmac_f64_dpp %5, %16, %17 row_newbcast:5 row_mask:0xf bank_mask:0xf\n\t"
      "v_fmac_f64_dpp %6, %16, %18 row_newbcast:5 row_mask:0xf bank_mask:0xf\n\t"
      "v_fmac_f64_dpp %7, %16, %18 row_newbcast:6 row_mask:0xf bank_mask:0xf\n\t"
      "v_fmac_f64_dpp %8, %16, %17 row_newbcast:10 row_mask:0xf bank_mask:0xf\n\t"
      "v_fmac_f64_dpp %9, %16, %17 row_newbcast:9 row_mask:0xf bank_mask:0xf\n\t"
      "v_fmac_f64_dpp %10, %16, %18 row_newbcast:9 row_mask:0xf bank_mask:0xf\n\t"
      "v_fmac_f64_dpp %11, %16, %18 row_newbcast:10 row_mask:0xf bank_mask:0xf\n\t"
      "v_fmac_f64_dpp %12, %16, %17 row_newbcast:14 row_mask:0xf bank_mask:0xf\n\t"
      "v_fmac_f64_dpp %13, %16, %17 row_newbcast:13 row_mask:0xf bank_mask:0xf\n\t"
      "v_fmac_f64_dpp %14, %16, %18 row_newbcast:13 row_mask:0xf bank_mask:0xf\n\t"
      "v_fmac_f64_dpp %15, %16, %18 row_newbcast:14 row_mask:0xf bank_mask:0xf\n"
      "2:"
      : "+v"(F[0][2]), "+v"(F[0][3]), "+v"(F[0][5]), "+v"(F[0][6]), "+v"(F[1][2]), "+v"(F[1][3]), "+v"(F[1][5]),
        "+v"(F[1][6]), "+v"(F[2][2]), "+v"(F[2][3]), "+v"(F[2][5]), "+v"(F[2][6]), "+v"(F[3][2]), "+v"(F[3][3]),
        "+v"(F[3][5]), "+v"(F[3][6])
      : "v"(t), "v"(r0.x), "v"(r1.y), "s"(kind)
      : "scc");
}

constexpr int kDppAheadMax = 6;  // entries in flight per wave, at most (the stream is padded to a multiple)

// NW waves per block, each taking phase groups w, w + NW, ...; D entries in flight
// one phase group's stream over the staged tile: the FMAs of each entry, then the
// three outputs of each phase for this lane's pixel
template <int D>
__device__ __forceinline__ void fir_dpp_group(const FirArgs &a, const double *T, const int32_t *X, int cap, int ne,
                                              const char *la, const char *lb, int p0, int n_job, long j, double cs,
                                              double cf, GLOBAL double *out) {
  constexpr int kDppAhead = D;
  const long n = a.n;
  double F[kTileG][7];
#pragma unroll
  for (int q = 0; q < kTileG; q++)
#pragma unroll
    for (int f = 0; f < 7; f++) F[q][f] = 0.0;
  if (ne > 0) {
    const int last = ne - 1;
    double t[kDppAhead];
    int x[kDppAhead];
    // slot s holds the taps of entry m + s and the offset of entry m + s + 1 (the
    // rows are read one entry ahead); rows in two buffers by entry parity.  A slot
    // is refilled after its last use, so each load writes the register that its
    // use reads and nothing waits for it before then
    // buffer loads: the stream bases in scalar registers (resources), the entry's
    // offset a scalar operand and the lane's tap a constant vector offset, so a step
    // spends no vector instructions on addresses; the offsets are uniform but come
    // through the vector unit too (a scalar load would share the LDS reads' wait
    // counter and, returning out of order, force a full wait each step)
    const __amdgpu_buffer_rsrc_t rt = __builtin_amdgcn_make_buffer_rsrc((void *)T, 0, cap * 8 * (4 * kTileG), 0x00020000);
    const __amdgpu_buffer_rsrc_t rx = __builtin_amdgcn_make_buffer_rsrc((void *)X, 0, cap * 4, 0x00020000);
    const int tl = (threadIdx.x & 15) * 8;
    auto ldt = [&](int e) {
      return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(rt, tl, e * (8 * 4 * kTileG), 0));
    };
    auto ldx = [&](int e) { return __builtin_amdgcn_raw_buffer_load_b32(rx, 0, e * 4, 0); };
    const int x0 = ldx(0);
    // issued in the loop's own order (taps then offset, slot by slot), so that the
    // wait counts at the loop head are the same on entry as around the loop
#pragma unroll
    for (int s = 0; s < kDppAhead; s++) {
      __builtin_amdgcn_sched_barrier(0);
      t[s] = ldt(min(s, last));
      x[s] = ldx(min(s + 1, last));
    }
    __builtin_amdgcn_sched_barrier(0);
    // the offset word: LDS byte offset in bits 0-27, the entry's kind in bits 28-29
    constexpr int kOff = (1 << 28) - 1;
    double2 r0[2], r1[2];
    r0[0] = *(const double2 *)(la + (x0 & kOff));
    r1[0] = *(const double2 *)(lb + (x0 & kOff));
    int kind = uni(x0) >> 28;  // the kind of entry m + s (uniform: every lane loaded the same word)
    for (int m = 0; m < ne; m += kDppAhead) {
#pragma unroll
      for (int s = 0; s < kDppAhead; s++) {
        // each step's instructions stay in their step: hoisting the next step's LDS
        // address (a just-loaded offset) above these FMAs would wait for that load
        __builtin_amdgcn_sched_barrier(0);
        r0[(s + 1) & 1] = *(const double2 *)(la + (x[s] & kOff));  // rows of entry m + s + 1
        r1[(s + 1) & 1] = *(const double2 *)(lb + (x[s] & kOff));
        const int kind_next = uni(x[s]) >> 28;
        fir_dpp_entry(F, t[s], r0[s & 1], r1[s & 1], kind);
        kind = kind_next;
        t[s] = ldt(min(m + s + kDppAhead, last));
        x[s] = ldx(min(m + s + kDppAhead + 1, last));
      }
    }
  }
  if (j < n) {
#pragma unroll
    for (int q = 0; q < kTileG; q++) {
      if (p0 + q >= n_job) break;
      GLOBAL double *o = out + (long)(p0 + q) * 3 * n;
      // F: 0 = H1a*SEDq, 1 = H1s*Sq, 2 = H2f*Sq, 3 = H1f*Sq, 4 = H1s*Ssp, 5 = H1f*Sfac, 6 = H2f*Sfac
      // The three planes are written once and read back only by the convolution, which runs after
      // the whole grid: streaming stores keep them from evicting the tap and offset streams.
      __builtin_nontemporal_store(F[q][0] - cs * F[q][1] - cf * F[q][2], &o[j]);
      __builtin_nontemporal_store((F[q][1] + F[q][3]) - F[q][4] - F[q][5], &o[n + j]);
      __builtin_nontemporal_store(F[q][0] - cs * F[q][4] - cf * F[q][6], &o[2 * n + j]);
    }
  }
}

template <int NW, int D>
__device__ __forceinline__ void fir_tile_dpp_body(const FirArgs &a, const double *__restrict__ ptaps,
                                                  const int32_t *__restrict__ pxoff,
                                                  const int32_t *__restrict__ pcount,
                                                  const int32_t *__restrict__ pufl, double *__restrict__ out, int cap,
                                                  int n_job, int n_groups, int n_tiles, int tiles_per_xcd, int t64_log2,
                                                  int b) {
  extern __shared__ double2 srow[];  // plane A [n_mu][lmax] (q, qs), then plane B (sp, fa)
  const int TP = 64 << t64_log2;    // tile pixels: 64, 128 or 256 (few groups: waves take (sub-tile, group))
  const int tile = (b & 7) * tiles_per_xcd + (b >> 3);
  if (tile >= n_tiles) return;  // uniform over the block
  const int nd = a.nd, n_mu = a.n_mu, lmax = TP + nd - 1;
  const long n = a.n;
  const long j0 = (long)tile * TP;
  const long s0 = j0 - a.dmin - (nd - 1);
  const int tot = n_mu * lmax;
  {
    // kStage elements per thread with all their loads in flight before the first
    // use: clamped addresses (a row the bin lacks reads the quiet row instead), so
    // every load is unconditional, then selects
    constexpr int kStage = 4;
    const GLOBAL double *q = gl(a.q), *qs = gl(a.qs), *sp = gl(a.sp), *fa = gl(a.fa);
    for (int e0 = threadIdx.x; e0 < tot; e0 += kStage * 64 * NW) {
      double vq[kStage], vs[kStage], vp[kStage], vf[kStage];
      bool in[kStage];
      int fl[kStage];
#pragma unroll
      for (int u = 0; u < kStage; u++) {
        const int e = min(e0 + u * 64 * NW, tot - 1);
        const int k = e / lmax, x = e - k * lmax;
        fl[u] = gl(pufl)[k];
        const long i = s0 + x;
        in[u] = fl[u] && i >= 0 && i < n;
        const long o = (long)k * n + (in[u] ? i : 0);
        vq[u] = q[o];
        vs[u] = qs[o];
        vp[u] = ((fl[u] & 1) ? sp : q)[o];
        vf[u] = ((fl[u] & 2) ? fa : q)[o];
      }
      __builtin_amdgcn_sched_barrier(0);  // every load above the first use
#pragma unroll
      for (int u = 0; u < kStage; u++) {
        const int e = e0 + u * 64 * NW;
        if (e < tot) {
          srow[e] = make_double2(in[u] ? vq[u] : 0.0, in[u] ? vs[u] : 0.0);
          srow[tot + e] = make_double2((in[u] && (fl[u] & 1)) ? vp[u] : 0.0, (in[u] && (fl[u] & 2)) ? vf[u] : 0.0);
        }
      }
    }
  }
  __syncthreads();
  const int w = uni(threadIdx.x >> 6), lane = threadIdx.x & 63;
  // items (64-pixel sub-tile, group) in turn over the waves; the spot and facula
  // continua of the lane's pixel are loaded when the sub-tile changes
  int cur = -1;
  double cs = 0.0, cf = 0.0;
  for (int it = w; it < (n_groups << t64_log2); it += NW) {
    const int sub = it & ((1 << t64_log2) - 1), g = it >> t64_log2;
    const int px = sub * 64 + lane;
    const long j = j0 + px;
    if (sub != cur) {
      cs = (j < n && a.sps) ? gl(a.sps)[j] : 0.0;
      cf = (j < n && a.fas) ? gl(a.fas)[j] : 0.0;
      cur = sub;
    }
    const char *la = (const char *)srow + px * 16, *lb = la + (size_t)tot * 16;
    const int ne = gl(pcount)[g];  // a multiple of D
    fir_dpp_group<D>(a, ptaps + (long)g * cap * (4 * kTileG), pxoff + (long)g * cap, cap, ne, la, lb, g * kTileG,
                     n_job, j, cs, cf, gw(out));
  }
}

template <int NW, int D>
__global__ __launch_bounds__(64 * NW) void k_fir_tile_dpp(FirArgs a, const double *__restrict__ ptaps,
                                                          const int32_t *__restrict__ pxoff,
                                                          const int32_t *__restrict__ pcount,
                                                          const int32_t *__restrict__ pufl, double *__restrict__ out,
                                                          int cap, int n_job, int n_groups, int n_tiles,
                                                          int tiles_per_xcd, int t64_log2) {
  fir_tile_dpp_body<NW, D>(a, ptaps, pxoff, pcount, pufl, out, cap, n_job, n_groups, n_tiles, tiles_per_xcd,
                           t64_log2, blockIdx.x);
}

size_t fir_tile_lds(int n_mu, int nd, int t64) { return (size_t)n_mu * (64 * t64 + nd - 1) * 32; }
int fir_tile_ahead() { return 4; }

int fir_tile_groups(int n_job) { return (n_job + kTileG - 1) / kTileG; }


template <int D>
static void launch_tile_dpp(int nw, unsigned blocks, size_t lds, hipStream_t s, const FirArgs &a, const FirProg &pg,
                            int n_job, int n_groups, int nt, int tx, int t64_log2) {
#define SOAP_TILE_DPP(W)                                                                                        \
  k_fir_tile_dpp<W, D><<<blocks, 64 * W, lds, s>>>(a, pg.taps, pg.xoff, pg.count, pg.ufl, a.out, pg.cap, n_job, \
                                                   n_groups, nt, tx, t64_log2)
  switch (nw) {
    case 4: SOAP_TILE_DPP(4); break;
    case 5: SOAP_TILE_DPP(5); break;
    case 6: SOAP_TILE_DPP(6); break;
    case 7: SOAP_TILE_DPP(7); break;
    default: SOAP_TILE_DPP(8); break;
  }
#undef SOAP_TILE_DPP
}

// the device's LDS limit per block, read once per device
static size_t lds_limit() {
  static int lds_max[64];
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 0;
  if (!lds_max[dev]) {
    int v = 0;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess) return 0;
    lds_max[dev] = v;
  }
  return (size_t)lds_max[dev];
}

// sub-tiles of 64 pixels per tile (log2) for the DPP form: enough (sub-tile, group)
// items for eight waves when the phase groups are few, within the LDS limit
int fir_tile_t64_log2(int n_job, int n_mu, int nd, size_t lds_max) {
  const int ng = fir_tile_groups(n_job);
  int lg = 0;
  while (lg < 2 && (ng << lg) < 8 && fir_tile_lds(n_mu, nd, 2 << lg) <= lds_max) lg++;
  return lg;
}

hipError_t launch_fir_tile(const FirArgs &a, int n_job, const FirProg &pg, bool dpp, int waves, int ahead,
                           hipStream_t s) {
  if (n_job == 0) return hipSuccess;
  if (a.n_mu > 64 || pg.cap < a.n_mu * a.nd + kDppAheadMax || (waves != 0 && (waves < 4 || waves > 8)) ||
      (ahead != 4 && ahead != 6))
    return hipErrorInvalidValue;
  const size_t lds_max = lds_limit();
  const int lg = dpp ? fir_tile_t64_log2(n_job, a.n_mu, a.nd, lds_max) : 0;
  const int t64 = 1 << lg;
  const size_t lds = fir_tile_lds(a.n_mu, a.nd, t64);
  if (lds > lds_max) return hipErrorInvalidValue;
  const int TP = 64 * t64, lmax = TP + a.nd - 1;
  const int n_groups = fir_tile_groups(n_job);
  const int nw = waves ? waves : 8;
  const long n_tiles = (a.n + TP - 1) / TP;
  const long tiles_per_xcd = (n_tiles + 7) / 8;
  if (8 * tiles_per_xcd >= (1L << 31)) return hipErrorInvalidValue;
  k_fir_prog<<<n_groups + 1, kProgThreads, 0, s>>>(a, pg, n_job, n_groups, lmax, dpp ? 16 : 32, dpp ? ahead : 2);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  const unsigned blocks = (unsigned)(8 * tiles_per_xcd);
  const int nt = (int)n_tiles, tx = (int)tiles_per_xcd;
  if (dpp && ahead == 6)
    launch_tile_dpp<6>(nw, blocks, lds, s, a, pg, n_job, n_groups, nt, tx, lg);
  else if (dpp)
    launch_tile_dpp<4>(nw, blocks, lds, s, a, pg, n_job, n_groups, nt, tx, lg);
  else
    k_fir_tile<1, 8><<<blocks, 512, lds, s>>>(a, pg.taps, pg.xoff, pg.count, pg.ufl, a.out, pg.cap, n_job, n_groups,
                                             nt, tx);
  return hipGetLastError();
}

// Direct form of the phase-group FIR: no LDS staging and no barrier after the
// group metadata.  Each thread reads the seed-row values of its pixel straight
// from global memory (a wave's loads at one offset are 64 consecutive doubles;
// the rows of a tile are shared by all its phase groups, whose blocks run back
// to back on one XCD, so they come from L2 / the vector L1), and the inner loop
// is specialised by the row kinds a bin uses for the group (spot rows, facula
// rows, both), so absent rows cost neither loads nor FMAs.  Blocks whose source
// range is inside [0, n) skip the per-load bounds selects.
template <int G, bool SP, bool FA, bool CHECK>
__device__ __forceinline__ void fir_direct_bin(const GLOBAL double *rq, const GLOBAL double *rs,
                                               const GLOBAL double *rp, const GLOBAL double *rf,
                                               const GLOBAL double *Tg, long dstride, long i0, long n, int n_off,
                                               double (&F)[G][7]) {
  // unrolled so that the loads of several offsets are in flight together; the
  // unroll is bounded by the scalar registers holding the group's taps
  constexpr int U = G >= 2 ? 2 : 4;
#pragma unroll U
  for (int i = 0; i < n_off; i++) {
    const long x = i0 - i;
    bool in = true;
    long xc = x;
    if (CHECK) {
      in = x >= 0 && x < n;
      xc = in ? x : 0;
    }
    double vq = rq[xc], vs = rs[xc];
    double vp = SP ? rp[xc] : 0.0, vf = FA ? rf[xc] : 0.0;
    if (CHECK) {
      vq = in ? vq : 0.0;
      vs = in ? vs : 0.0;
      if (SP) vp = in ? vp : 0.0;
      if (FA) vf = in ? vf : 0.0;
    }
    const GLOBAL double *T = Tg + (long)i * dstride;
#pragma unroll
    for (int g = 0; g < G; g++) {
      // taps per phase: 0 = H1s, 1 = H1f, 2 = H2f, 3 = H1a (zero where a phase has no cells)
      const double t0 = T[g * 4 + 0], t1 = T[g * 4 + 1], t2 = T[g * 4 + 2], t3 = T[g * 4 + 3];
      F[g][0] = fma(t3, vs, F[g][0]);
      if (SP) {
        F[g][1] = fma(t0, vq, F[g][1]);
        F[g][4] = fma(t0, vp, F[g][4]);
      }
      if (FA) {
        F[g][2] = fma(t2, vq, F[g][2]);
        F[g][3] = fma(t1, vq, F[g][3]);
        F[g][5] = fma(t1, vf, F[g][5]);
        F[g][6] = fma(t2, vf, F[g][6]);
      }
    }
  }
}

template <int G>
__global__ __launch_bounds__(kTPB) void k_fir_direct(FirArgs a, int n_job, int n_tiles, int n_groups) {
  __shared__ int4 s_meta[64];
  __shared__ int s_nb;
  const int bx = blockIdx.x;
  const int xcd = bx & 7, slot = bx >> 3;
  const int tile = (slot / n_groups) * 8 + xcd;
  const int grp = slot % n_groups;
  if (tile >= n_tiles) return;  // uniform over the block
  const int p0 = grp * G;
  const int ng = min(G, n_job - p0);
  const int nd = a.nd, n_mu = a.n_mu;
  const GLOBAL int32_t *kflags = gl(a.kflags), *krange = gl(a.krange);
  const long j0 = (long)tile * kFirTile;
  const long j = j0 + threadIdx.x;
  const double cs = (j < a.n && a.sps) ? gl(a.sps)[j] : 0.0;
  const double cf = (j < a.n && a.fas) ? gl(a.fas)[j] : 0.0;
  if (threadIdx.x < 64) {  // group metadata: union flags and tap range per bin, bins in use in order
    const int k = threadIdx.x;
    int ufl = 0, ulo = nd, uhi = -1;
#pragma unroll
    for (int g = 0; g < G; g++)
      if (g < ng && k < n_mu) {
        const long jk = (long)(p0 + g) * n_mu + k;
        const int fl = kflags[jk], lo = krange[jk * 2], hi = krange[jk * 2 + 1];
        if (fl && lo <= hi) {
          ufl |= fl | 4;
          ulo = min(ulo, lo);
          uhi = max(uhi, hi);
        }
      }
    const unsigned long long used = __ballot(ufl != 0);
    if (ufl) s_meta[__popcll(used & ((1ull << k) - 1ull))] = make_int4(k, ufl, ulo, uhi);
    if (k == 0) s_nb = __popcll(used);
  }
  __syncthreads();
  double F[G][7];
#pragma unroll
  for (int g = 0; g < G; g++)
#pragma unroll
    for (int f = 0; f < 7; f++) F[g][f] = 0.0;
  const int nb = uni(s_nb);
  const long dstride = (long)a.p_pad * 4;
  for (int kk = 0; kk < nb; kk++) {
    const int4 m = s_meta[kk];
    const int k = uni(m.x), ufl = uni(m.y), ulo = uni(m.z), uhi = uni(m.w);
    const long roff = (long)k * a.n;
    const GLOBAL double *rq = gl(a.q) + roff, *rs = gl(a.qs) + roff;
    const GLOBAL double *rp = (ufl & 1) ? gl(a.sp) + roff : rq, *rf = (ufl & 2) ? gl(a.fa) + roff : rq;
    const GLOBAL double *Tg = gl(a.taps_g) + (((long)k * nd + ulo) * a.p_pad + p0) * 4;
    const int n_off = uhi - ulo + 1;
    // source index of offset ulo + i: x = j - dmin - ulo - i
    const long i0 = j - a.dmin - ulo;
    const bool inside = (j0 - a.dmin - uhi >= 0) && (j0 + kFirTile - 1 - a.dmin - ulo < a.n);  // uniform
    const int kind = ufl & 3;
    if (inside) {
      if (kind == 3) fir_direct_bin<G, true, true, false>(rq, rs, rp, rf, Tg, dstride, i0, a.n, n_off, F);
      else if (kind == 1) fir_direct_bin<G, true, false, false>(rq, rs, rp, rf, Tg, dstride, i0, a.n, n_off, F);
      else fir_direct_bin<G, false, true, false>(rq, rs, rp, rf, Tg, dstride, i0, a.n, n_off, F);
    } else {
      if (kind == 3) fir_direct_bin<G, true, true, true>(rq, rs, rp, rf, Tg, dstride, i0, a.n, n_off, F);
      else if (kind == 1) fir_direct_bin<G, true, false, true>(rq, rs, rp, rf, Tg, dstride, i0, a.n, n_off, F);
      else fir_direct_bin<G, false, true, true>(rq, rs, rp, rf, Tg, dstride, i0, a.n, n_off, F);
    }
  }
  if (j < a.n) {
#pragma unroll
    for (int g = 0; g < G; g++) {
      if (g >= ng) break;
      GLOBAL double *o = gw(a.out) + (long)(p0 + g) * 3 * a.n;
      o[j] = F[g][0] - cs * F[g][1] - cf * F[g][2];
      o[a.n + j] = (F[g][1] + F[g][3]) - F[g][4] - F[g][5];
      o[2 * a.n + j] = F[g][0] - cs * F[g][4] - cf * F[g][6];
    }
  }
}

hipError_t launch_fir_direct(const FirArgs &a, int n_job, int G, hipStream_t s) {
  if (n_job == 0) return hipSuccess;
  if (a.n_mu > 64) return hipErrorInvalidValue;
  int n_tiles = (int)((a.n + kFirTile - 1) / kFirTile);
  int n_groups = (n_job + G - 1) / G;
  long tiles8 = ((long)n_tiles + 7) / 8 * 8;
  long blocks = tiles8 * n_groups;
  if (blocks >= (1L << 31)) return hipErrorInvalidValue;
  switch (G) {
    case 1: k_fir_direct<1><<<(unsigned)blocks, kTPB, 0, s>>>(a, n_job, n_tiles, n_groups); break;
    case 2: k_fir_direct<2><<<(unsigned)blocks, kTPB, 0, s>>>(a, n_job, n_tiles, n_groups); break;
    case 4: k_fir_direct<4><<<(unsigned)blocks, kTPB, 0, s>>>(a, n_job, n_tiles, n_groups); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

// Register-blocked form of the active FIR.  Per block: a lambda tile of
// kTPB * B pixels and a group of G phases; each thread owns B consecutive
// output pixels of every phase of the group.  For each mu bin used by the
// group, the group's taps (reversed, per set, zero outside each phase's range)
// and then each needed seed row segment are staged in LDS; the row is read
// into registers in windows of kF2CH + B - 1 values and every tap (an LDS
// broadcast) feeds B FMAs per phase.  Six accumulators per (phase, pixel):
//   0: F0 = H1a * Sq_SED      1: F1 = H1s * Sq      2: F2 = H2f * Sq
//   3: F4 = H1s * S_spot      4: F6 = H2f * S_fac   5: Bc = H1f * (Sq - S_fac)
// flux = F0 - cs F1 - cf F2, bconv = F1 + Bc - F4, tot = F0 - cs F4 - cf F6
// (cs, cf the spot / facula SED continua; the products of the single-phase
// form in the same combination).  B is odd: a thread stride of B doubles is
// conflict-free for 64-bit LDS reads.  Blocks are XCD-ordered as in the other
// forms (all phase groups of a tile on XCD b % 8).
constexpr int kF2CH = 8;

template <int G, int B, int R>
__device__ __forceinline__ void fir2_row(const double *xs, const double *tp, int Lc, int lcs, double (&F)[G][6][B]) {
  const double *base = xs + threadIdx.x * B;
  for (int k0 = 0; k0 < Lc; k0 += kF2CH) {
    double win[kF2CH + B - 1];
#pragma unroll
    for (int e = 0; e < kF2CH + B - 1; e++) win[e] = base[k0 + e];
#pragma unroll
    for (int k = 0; k < kF2CH; k++) {
#pragma unroll
      for (int g = 0; g < G; g++) {
        const double *t = tp + (g * 4) * lcs + k0 + k;
        if (R == 0) {  // quiet SED row: H1a
          const double c = t[3 * lcs];
#pragma unroll
          for (int b = 0; b < B; b++) F[g][0][b] = fma(c, win[b + k], F[g][0][b]);
        } else if (R == 1) {  // quiet row: H1s, H1f, H2f
          const double c0 = t[0], c1 = t[lcs], c2 = t[2 * lcs];
#pragma unroll
          for (int b = 0; b < B; b++) {
            F[g][1][b] = fma(c0, win[b + k], F[g][1][b]);
            F[g][5][b] = fma(c1, win[b + k], F[g][5][b]);
            F[g][2][b] = fma(c2, win[b + k], F[g][2][b]);
          }
        } else if (R == 2) {  // spot row: H1s
          const double c0 = t[0];
#pragma unroll
          for (int b = 0; b < B; b++) F[g][3][b] = fma(c0, win[b + k], F[g][3][b]);
        } else {  // facula row: H1f (subtracted), H2f
          const double c1 = t[lcs], c2 = t[2 * lcs];
#pragma unroll
          for (int b = 0; b < B; b++) {
            F[g][5][b] = fma(-c1, win[b + k], F[g][5][b]);
            F[g][4][b] = fma(c2, win[b + k], F[g][4][b]);
          }
        }
      }
    }
  }
}

template <int G, int B>
__device__ __forceinline__ void fir2_body(const FirArgs &a, int n_job, int n_tiles, int n_groups, int bx_) {
  constexpr int T = kTPB * B;
  extern __shared__ double f2s[];
  __shared__ int s_fl[G][64];
  __shared__ int s_bins[64], s_ufl[64], s_ulo[64], s_uhi[64];
  __shared__ int s_nb;
  const int xcd = bx_ & 7, slot = bx_ >> 3;
  const int tile = (slot / n_groups) * 8 + xcd;
  const int grp = slot % n_groups;
  if (tile >= n_tiles) return;  // uniform over the block
  const int p0 = grp * G;
  const int ng = min(G, n_job - p0);
  const int nd = a.nd, n_mu = a.n_mu;
  const int lcs = (nd + kF2CH - 1) / kF2CH * kF2CH + kF2CH;  // LDS stride of one tap set
  double *tp = f2s;                    // [G][4][lcs]
  double *xs = f2s + (long)G * 4 * lcs;  // [T + lcs]
  const GLOBAL int32_t *kflags = gl(a.kflags), *krange = gl(a.krange);
  const GLOBAL double *taps = gl(a.taps);
  const long j0 = (long)tile * T;
  const long src0 = j0 - a.dmin - (nd - 1);
  // group metadata: per-phase flags, union flags and union tap range per bin
  for (int e = threadIdx.x; e < G * 64; e += blockDim.x) {
    int g = e / 64, k = e % 64;
    int fl = 0;
    if (g < ng && k < n_mu) {
      long jk = (long)(p0 + g) * n_mu + k;
      fl = kflags[jk];
      if (krange[jk * 2] > krange[jk * 2 + 1]) fl = 0;
    }
    s_fl[g][k] = fl;
  }
  __syncthreads();
  if (threadIdx.x < 64) {
    int k = threadIdx.x;
    int ufl = 0, ulo = nd, uhi = -1;
    for (int g = 0; g < ng; g++)
      if (s_fl[g][k]) {
        long jk = (long)(p0 + g) * n_mu + k;
        ufl |= s_fl[g][k] | 4;
        ulo = min(ulo, nd - 1 - krange[jk * 2 + 1]);
        uhi = max(uhi, nd - 1 - krange[jk * 2]);
      }
    s_ufl[k] = ufl;
    s_ulo[k] = ulo;
    s_uhi[k] = uhi;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    int c = 0;
    for (int k = 0; k < n_mu; k++)
      if (s_ufl[k]) s_bins[c++] = k;
    s_nb = c;
  }
  __syncthreads();
  double F[G][6][B];
#pragma unroll
  for (int g = 0; g < G; g++)
#pragma unroll
    for (int f = 0; f < 6; f++)
#pragma unroll
      for (int b = 0; b < B; b++) F[g][f][b] = 0.0;
  const int nb = uni(s_nb);
  for (int kk = 0; kk < nb; kk++) {
    const int k = uni(s_bins[kk]);
    const int ufl = uni(s_ufl[k]), ulo = uni(s_ulo[k]), uhi = uni(s_uhi[k]);
    const int Lc = (uhi - ulo + kF2CH) / kF2CH * kF2CH;  // taps used, padded with zeros
    __syncthreads();  // the previous bin's LDS reads are done
    // taps of the group, reversed (u = nd - 1 - d), zero outside each phase's range
    for (int e = threadIdx.x; e < G * 4 * Lc; e += blockDim.x) {
      int g = e / (4 * Lc), r = e % (4 * Lc), st = r / Lc, u = ulo + r % Lc;
      double v = 0.0;
      if (g < ng && s_fl[g][k]) {
        long jk = (long)(p0 + g) * n_mu + k;
        int di = nd - 1 - u;
        if (di >= krange[jk * 2] && di <= krange[jk * 2 + 1]) v = taps[(jk * nd + di) * 4 + st];
      }
      tp[(g * 4 + st) * lcs + (r % Lc)] = v;
    }
    const int span = T + Lc - 1;
    const long rs0 = src0 + ulo;
#pragma unroll 1
    for (int r = 0; r < 4; r++) {
      if ((r == 2 && !(ufl & 1)) || (r == 3 && !(ufl & 2))) continue;  // no spot / facula cells in this bin
      const GLOBAL double *src = gl(r == 0 ? a.qs : r == 1 ? a.q : r == 2 ? a.sp : a.fa) + (long)k * a.n;
      if (r > 0) __syncthreads();  // the previous row's reads are done
      for (int x = threadIdx.x; x < span; x += blockDim.x) {
        long i = rs0 + x;
        xs[x] = (i >= 0 && i < a.n) ? src[i] : 0.0;
      }
      __syncthreads();
      if (r == 0) fir2_row<G, B, 0>(xs, tp, Lc, lcs, F);
      else if (r == 1) fir2_row<G, B, 1>(xs, tp, Lc, lcs, F);
      else if (r == 2) fir2_row<G, B, 2>(xs, tp, Lc, lcs, F);
      else fir2_row<G, B, 3>(xs, tp, Lc, lcs, F);
    }
  }
  // epilogue: the three spectra of each phase through LDS for coalesced stores
  double cs[B], cf[B];
#pragma unroll
  for (int b = 0; b < B; b++) {
    long j = j0 + threadIdx.x * B + b;
    cs[b] = (a.sps && j < a.n) ? gl(a.sps)[j] : 0.0;
    cf[b] = (a.fas && j < a.n) ? gl(a.fas)[j] : 0.0;
  }
#pragma unroll
  for (int g = 0; g < G; g++) {
    if (g >= ng) break;
    GLOBAL double *o = gw(a.out) + (long)(p0 + g) * 3 * a.n;
#pragma unroll
    for (int t = 0; t < 3; t++) {
      __syncthreads();
#pragma unroll
      for (int b = 0; b < B; b++) {
        double v = t == 0 ? F[g][0][b] - cs[b] * F[g][1][b] - cf[b] * F[g][2][b]
                 : t == 1 ? F[g][1][b] + F[g][5][b] - F[g][3][b]
                          : F[g][0][b] - cs[b] * F[g][3][b] - cf[b] * F[g][4][b];
        xs[threadIdx.x * B + b] = v;
      }
      __syncthreads();
      for (int e = threadIdx.x; e < T; e += blockDim.x) {
        long j = j0 + e;
        if (j < a.n) o[(long)t * a.n + j] = xs[e];
      }
    }
  }
}

template <int G, int B>
__global__ __launch_bounds__(kTPB) void k_fir2(FirArgs a, int n_job, int n_tiles, int n_groups) {
  fir2_body<G, B>(a, n_job, n_tiles, n_groups, blockIdx.x);
}

size_t fir2_lds(int nd, int G, int B) {
  int lcs = (nd + kF2CH - 1) / kF2CH * kF2CH + kF2CH;
  return ((size_t)G * 4 * lcs + (size_t)kTPB * B + lcs) * sizeof(double);
}

int fir2_tiles(int64_t n, int B) { return (int)((n + (int64_t)kTPB * B - 1) / ((int64_t)kTPB * B)); }

hipError_t launch_fir2(const FirArgs &a, int n_job, int G, int B, hipStream_t s) {
  if (n_job == 0) return hipSuccess;
  if (a.n_mu > 64) return hipErrorInvalidValue;
  int n_tiles = fir2_tiles(a.n, B);
  int n_groups = (n_job + G - 1) / G;
  long blocks = ((long)n_tiles + 7) / 8 * 8 * n_groups;
  size_t lds = fir2_lds(a.nd, G, B);
  if (lds > 65536 || blocks > 0x7fffffffL) return hipErrorInvalidValue;
#define SOAP_FIR2(GG, BB)                                                                                  \
  if (G == GG && B == BB) {                                                                                \
    k_fir2<GG, BB><<<(unsigned)blocks, kTPB, lds, s>>>(a, n_job, n_tiles, n_groups);                       \
    return hipGetLastError();                                                                              \
  }
  SOAP_FIR2(1, 3)
  SOAP_FIR2(1, 5)
  SOAP_FIR2(2, 3)
  SOAP_FIR2(2, 1)
  SOAP_FIR2(4, 1)
#undef SOAP_FIR2
  return hipErrorInvalidValue;
}

// Solar-mode active deltas, one phase per block (taps [job][n_mu][nd][K+3]:
// H1s, H1f, L_0..L_{K-1}, H1a).  With R_q = B(lambda, T_q)/B(lambda, T_norm):
//   flux  = H1a*SEDq - Cs (H1s*Sq) - Cf sum_q R_q (L_q*Sq)
//   bconv = H1s*Sq + H1f*Sq - H1s*Ssp - H1f*Sfac
//   tot   = H1a*SEDq - Cs (H1s*Ssp) - Cf sum_q R_q (L_q*Sfac)
// Rows are staged in LDS per mu bin as in k_fir_active_grp; blocks of one tile
// run on one XCD.
template <int K>
__global__ __launch_bounds__(kTPB) void k_fir_solar(FirArgs a, SolarFir sf, int n_job, int n_tiles) {
  constexpr int NS = K + 3;
  extern __shared__ double rows[];
  __shared__ int s_bins[64];
  __shared__ int s_nb;
  long b = blockIdx.x;
  long xcd = b % 8, slot = b / 8;
  long tile = (slot / n_job) * 8 + xcd;
  int job = (int)(slot % n_job);
  if (tile >= n_tiles) return;  // uniform over the block
  int nd = a.nd, n_mu = a.n_mu;
  int row_len = kFirTile + nd - 1;
  long j0 = tile * kFirTile;
  long src0 = j0 - a.dmin - nd + 1;
  const long jb = (long)job * n_mu;
  if (threadIdx.x == 0) {
    int c = 0;
    for (int k = 0; k < n_mu; k++)
      if (a.kflags[jb + k] && a.krange[(jb + k) * 2] <= a.krange[(jb + k) * 2 + 1]) s_bins[c++] = k;
    s_nb = c;
  }
  __syncthreads();
  double F0 = 0, F1 = 0, F4 = 0, Fq = 0, Ff = 0, Gq[K], Gf[K];
#pragma unroll
  for (int q = 0; q < K; q++) Gq[q] = Gf[q] = 0.0;
  const int nb = uni(s_nb);
  for (int kk = 0; kk < nb; kk++) {
    const int k = uni(s_bins[kk]);
    const int fl = uni(a.kflags[jb + k]);
    const int lo = uni(a.krange[(jb + k) * 2]), hi = uni(a.krange[(jb + k) * 2 + 1]);
    if (kk > 0) __syncthreads();
    const int has_sp = fl & 1, has_fa = (fl >> 1) & 1;
    const int nrow = 2 + has_sp + has_fa;
    const double *src[4] = {a.q + (long)k * a.n, a.qs + (long)k * a.n,
                            has_sp ? a.sp + (long)k * a.n : a.fa + (long)k * a.n, a.fa + (long)k * a.n};
    for (int e = threadIdx.x; e < nrow * row_len; e += blockDim.x) {
      int r = e / row_len, x = e - r * row_len;
      long i = src0 + x;
      rows[e] = (i >= 0 && i < a.n) ? src[r][i] : 0.0;
    }
    __syncthreads();
    const double *Tk = a.taps + (jb + k) * nd * NS;
    const double *rq = rows, *rs = rows + row_len, *rp = rows + 2 * row_len, *rf = rows + (2 + has_sp) * row_len;
    for (int di = lo; di <= hi; di++) {
      int e = threadIdx.x + nd - 1 - di;
      const double *T = Tk + di * NS;
      double vq = rq[e];
      F0 = fma(T[NS - 1], rs[e], F0);
      if (has_sp) {
        F1 = fma(T[0], vq, F1);
        F4 = fma(T[0], rp[e], F4);
      }
      if (has_fa) {
        double vf = rf[e];
        Fq = fma(T[1], vq, Fq);
        Ff = fma(T[1], vf, Ff);
#pragma unroll
        for (int q = 0; q < K; q++) {
          Gq[q] = fma(T[2 + q], vq, Gq[q]);
          Gf[q] = fma(T[2 + q], vf, Gf[q]);
        }
      }
    }
  }
  long j = j0 + threadIdx.x;
  if (j < a.n) {
    double cs = a.sps ? a.sps[j] : 0.0;
    double cf = a.fas ? a.fas[j] : 0.0;
    double sq = 0.0, sfa = 0.0;
    if (cf != 0.0) {
      // B(T)/B(Tn) = (exp(x_n) - 1)/(exp(x) - 1), x = h c/(lambda k T scale)
      double lam = sf.wave[j];
      double hc = kPlanckH * kC;
      double en = expm1(hc / (lam * kBoltzK * (double)sf.t_norm * 1e-10));
#pragma unroll
      for (int q = 0; q < K; q++) {
        double R = en / expm1(hc / (lam * kBoltzK * sf.node_t[q] * 1e-10));
        sq = fma(R, Gq[q], sq);
        sfa = fma(R, Gf[q], sfa);
      }
    }
    double *o = a.out + (long)job * 3 * a.n;
    o[j] = F0 - cs * F1 - cf * sq;
    o[a.n + j] = (F1 + Fq) - F4 - Ff;
    o[2 * a.n + j] = F0 - cs * F4 - cf * sfa;
  }
}

hipError_t launch_fir_solar(const FirArgs &a, const SolarFir &sf, int n_job, hipStream_t s) {
  if (n_job == 0) return hipSuccess;
  if (a.n_mu > 64) return hipErrorInvalidValue;
  int n_tiles = (int)((a.n + kFirTile - 1) / kFirTile);
  long tiles8 = ((long)n_tiles + 7) / 8 * 8;
  long blocks = tiles8 * n_job;
  size_t lds = (size_t)4 * (kFirTile + a.nd - 1) * sizeof(double);
  if (lds > 60000 || blocks > 0x7fffffffL) return hipErrorInvalidValue;
  switch (sf.k) {
    case 6: k_fir_solar<6><<<(unsigned)blocks, kTPB, lds, s>>>(a, sf, n_job, n_tiles); break;
    case 8: k_fir_solar<8><<<(unsigned)blocks, kTPB, lds, s>>>(a, sf, n_job, n_tiles); break;
    case 10: k_fir_solar<10><<<(unsigned)blocks, kTPB, lds, s>>>(a, sf, n_job, n_tiles); break;
    case 12: k_fir_solar<12><<<(unsigned)blocks, kTPB, lds, s>>>(a, sf, n_job, n_tiles); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

// Quiet star with LDS staging per mu bin: 256 pixels per block, the two rows of
// a bin staged together, one pixel per thread.
__global__ __launch_bounds__(kTPB) void k_fir_quiet_lds(FirArgs a) {
  extern __shared__ double rows[];
  long j0 = (long)blockIdx.x * kTPB;
  int row_len = kTPB + a.nd - 1;
  long src0 = j0 - a.dmin - a.nd + 1;
  long j = j0 + threadIdx.x;
  double accq = 0.0, accs = 0.0;
  int nb = a.nbins[0];
  for (int kk = 0; kk < nb; kk++) {
    int k = a.bins[kk];
    int lo = a.krange[k * 2], hi = a.krange[k * 2 + 1];
    const double *T = a.taps + (long)k * a.nd;
    __syncthreads();
    for (int e = threadIdx.x; e < row_len; e += blockDim.x) {
      long i = src0 + e;
      bool in = i >= 0 && i < a.n;
      rows[e] = in ? a.q[(long)k * a.n + i] : 0.0;
      rows[row_len + e] = in ? a.qs[(long)k * a.n + i] : 0.0;
    }
    __syncthreads();
    for (int di = lo; di <= hi; di++) {
      int e = threadIdx.x + a.nd - 1 - di;
      double hv = T[di];
      accq = fma(hv, rows[e], accq);
      accs = fma(hv, rows[row_len + e], accs);
    }
  }
  if (j < a.n) {
    a.out[j] = accq;
    a.out[a.n + j] = accs;
  }
}

// The quiet star's FIR with the seed rows of the next bin in flight: a block of
// kQuietPx * 256 pixels keeps two LDS buffers of the rows of one bin; while it
// sums bin kk from one, the rows of bin kk + 1 go from registers into the other
// and the loads of bin kk + 2 are issued, so one bin's load latency hides behind
// another's FMAs (k_fir_quiet_lds waited on each bin's loads).  The sums per pixel
// are the same fmas in the same order (bins in list order, offsets ascending).
constexpr int kQuietPx = 2;
constexpr int kQuietTile = kTPB * kQuietPx;
constexpr int kQuietSlots = 3;  // row values per thread: rows of up to 3 * 256 pixels (nd <= 257)
__global__ __launch_bounds__(kTPB) void k_fir_quiet_pipe(FirArgs a) {
  extern __shared__ double rows[];  // [2 buffers][q row, qs row] of row_len each
  const long n = a.n;
  const long j0 = (long)blockIdx.x * kQuietTile;
  const int row_len = kQuietTile + a.nd - 1;
  const long src0 = j0 - a.dmin - a.nd + 1;
  const CONSTANT int32_t *bins = gc(a.bins), *krange = gc(a.krange);
  const CONSTANT double *taps = gc(a.taps);
  const GLOBAL double *q = gl(a.q), *qs = gl(a.qs);
  const int nb = gc(a.nbins)[0];
  double rq[kQuietSlots], rs[kQuietSlots];
  auto load = [&](int kk) {  // bin kk's rows into registers, 0 outside the spectrum
    const long kb = (long)bins[kk] * n;
#pragma unroll
    for (int u = 0; u < kQuietSlots; u++) {
      const int e = threadIdx.x + u * kTPB;
      const long i = src0 + e;
      const bool in = e < row_len && i >= 0 && i < n;
      const long ic = in ? i : 0;
      const double vq = q[kb + ic], vs = qs[kb + ic];
      rq[u] = in ? vq : 0.0;
      rs[u] = in ? vs : 0.0;
    }
  };
  auto put = [&](int buf) {
    double *R = rows + (long)buf * 2 * row_len;
#pragma unroll
    for (int u = 0; u < kQuietSlots; u++) {
      const int e = threadIdx.x + u * kTPB;
      if (e < row_len) {
        R[e] = rq[u];
        R[row_len + e] = rs[u];
      }
    }
  };
  double accq[kQuietPx], accs[kQuietPx];
#pragma unroll
  for (int p = 0; p < kQuietPx; p++) accq[p] = accs[p] = 0.0;
  if (nb > 0) {
    load(0);
    put(0);
  }
  __syncthreads();
  if (nb > 1) load(1);
  for (int kk = 0; kk < nb; kk++) {
    const int k = bins[kk];
    const int lo = krange[k * 2], hi = krange[k * 2 + 1];
    const CONSTANT double *T = taps + (long)k * a.nd;
    const double *R = rows + (long)(kk & 1) * 2 * row_len;
#pragma unroll 4
    for (int di = lo; di <= hi; di++) {
      const double hv = T[di];
#pragma unroll
      for (int p = 0; p < kQuietPx; p++) {
        const int e = threadIdx.x + p * kTPB + a.nd - 1 - di;
        accq[p] = fma(hv, R[e], accq[p]);
        accs[p] = fma(hv, R[row_len + e], accs[p]);
      }
    }
    if (kk + 1 < nb) put((kk + 1) & 1);  // that buffer was last read in iteration kk - 1
    if (kk + 2 < nb) load(kk + 2);
    __syncthreads();
  }
  GLOBAL double *out = gw(a.out);
#pragma unroll
  for (int p = 0; p < kQuietPx; p++) {
    const long j = j0 + threadIdx.x + p * kTPB;
    if (j < n) {
      out[j] = accq[p];
      out[n + j] = accs[p];
    }
  }
}

hipError_t launch_fir_quiet_lds(const FirArgs &a, hipStream_t s) {
  const long row_pipe = kQuietTile + a.nd - 1;
  if (row_pipe <= (long)kQuietSlots * kTPB) {
    k_fir_quiet_pipe<<<(unsigned)((a.n + kQuietTile - 1) / kQuietTile), kTPB, (size_t)4 * row_pipe * sizeof(double),
                       s>>>(a);
    return hipGetLastError();
  }
  size_t lds = (size_t)2 * (kTPB + a.nd - 1) * sizeof(double);
  if (lds > 65536) return launch_fir_quiet(a, s);
  k_fir_quiet_lds<<<(unsigned)((a.n + kTPB - 1) / kTPB), kTPB, lds, s>>>(a);
  return hipGetLastError();
}

// Temperature / binary maps of one phase (spot_scan, SOAP_integration.cu:326-390;
// initial -9999 fill :686-700).
__global__ void k_phase_maps(int grid, int t_star, int t_diff_spot, int t_diff_fac, const RegionDev *geo, int n_geo,
                             double *temp, double *bi) {
#pragma clang fp contract(off)
  long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long)grid * grid) return;
  int iy = (int)(idx / grid), iz = (int)(idx % grid);
  double rc = 0;
  int ty = classify_cell(iy, iz, 2. / grid, geo, 0, n_geo, &rc);
  if (ty < 0) {
    temp[idx] = -9999.0;
    bi[idx] = -9999.0;
  } else if (ty == 0) {
    int T_spot = t_star - t_diff_spot;
    temp[idx] = T_spot;
    bi[idx] = 0.0;
  } else {
    temp[idx] = plage_temp(t_star + t_diff_fac, rc);
    bi[idx] = 1.0;
  }
}

hipError_t launch_phase_maps(int grid, int t_star, int t_diff_spot, int t_diff_fac, const RegionDev *geo,
                             int n_geo, double *temp, double *bi, hipStream_t s) {
  long ncell = (long)grid * grid;
  k_phase_maps<<<(unsigned)((ncell + kTPB - 1) / kTPB), kTPB, 0, s>>>(grid, t_star, t_diff_spot, t_diff_fac,
                                                                     geo, n_geo, temp, bi);
  return hipGetLastError();
}

// ------------------------------------------------------------ a7: resolution
// conv_inst (SOAP_resolution.cu:76-95): o_i = sum_d c_d (1 - A_{i+d}) with A = Q
// for the quiet star (SOAP_initialize.cu:78-96) and A = Q - D for an active
// spectrum, as the reference forms it; outputs exist for wid < i < n-1-wid only
// (zero elsewhere).  Tiles: kConvB outputs per thread (odd, so the per-lane LDS
// stride is bank-conflict free for 64-bit reads), the window in register slots
// and the taps in scalar registers, kConvCH at a time (the host pads the tap
// vector with zeros to a multiple of kConvCH).
#ifndef SOAP_CONV_B
#define SOAP_CONV_B 11
#endif
constexpr int kConvB = SOAP_CONV_B;  // outputs per thread (odd: conflict-free LDS stride)
#ifndef SOAP_CONV_CH
#define SOAP_CONV_CH 12
#endif
constexpr int kConvCH = SOAP_CONV_CH;
constexpr int kConvTile = kTPB * kConvB;
constexpr int kStageU = 8;  // global loads per thread issued together when staging
constexpr int kStageA = 6;  // slots per thread and pass of the active convolution's buffer-load staging
constexpr long kConvBufN = 1L << 27;  // spectra shorter than this stage through buffer loads (8n < 2^31)
constexpr int kConvMaxTaps = 4095;

// stage x[i0 + dlo + e], e in [0, kConvTile + ntp - 1), into LDS; g(x) = 1 - x for
// the quiet term, x otherwise; zero outside [0, n)
template <bool ONE_MINUS>
__device__ __forceinline__ void conv_stage(double *xs, const GLOBAL double *X, long i0, int dlo, int span, long n) {
  for (int e0 = 0; e0 < span; e0 += kStageU * kTPB) {
    double xv[kStageU];
#pragma unroll
    for (int u = 0; u < kStageU; u++) {  // all loads in flight before the first use
      long m = i0 + dlo + e0 + threadIdx.x + u * kTPB;
      xv[u] = X[m < 0 ? 0 : (m >= n ? n - 1 : m)];
    }
    __builtin_amdgcn_sched_barrier(0);  // keep every load above the first use
#pragma unroll
    for (int u = 0; u < kStageU; u++) {  // selects, not branches: the loads stay in flight together
      int e = e0 + threadIdx.x + u * kTPB;
      long m = i0 + dlo + e;
      const double x = xv[u];
      double v = (m >= 0 && m < n) ? x : 0.0;
      if (e < span) xs[e] = ONE_MINUS ? 1.0 - v : v;
    }
  }
}

// acc[b] = sum_k c_k xs[tid*B + b + k], k in [0, ntp), in increasing k.  The
// window x[k .. k+B-1] of tap k lives in kConvCH > B register slots (x[m] in
// slot m mod kConvCH): after tap k, x[k] is dead and its slot is refilled with
// x[k + kConvCH], first needed kConvCH - B + 1 taps later, so each tap costs one
// LDS read for B FMAs.  The taps go through the scalar unit, kConvCH at a time
// (uniform loads; ntp is a multiple of kConvCH).  The refills run up to kConvCH
// values past the staged window: the LDS allocation holds that slack.
__device__ __forceinline__ void conv_accumulate(const double *xs, const CONSTANT double *ctg, int ntp,
                                                double acc[kConvB]) {
  static_assert(kConvCH > kConvB, "the window needs more slots than outputs");
#pragma unroll
  for (int b = 0; b < kConvB; b++) acc[b] = 0.0;
  const double *base = xs + threadIdx.x * kConvB;
  double w[kConvCH];
#pragma unroll
  for (int e = 0; e < kConvCH; e++) w[e] = base[e];
  for (int k0 = 0; k0 < ntp; k0 += kConvCH) {
    const CONSTANT double *cp = ctg + uni(k0);  // uniform: the taps go through the scalar unit
    double cv[kConvCH];
#pragma unroll
    for (int k = 0; k < kConvCH; k++) cv[k] = cp[k];
    const double *nx = base + k0 + kConvCH;
#pragma unroll
    for (int k = 0; k < kConvCH; k++) {
#pragma unroll
      for (int b = 0; b < kConvB; b++) acc[b] = fma(cv[k], w[(k + b) % kConvCH], acc[b]);
      w[k] = nx[k];
    }
  }
}

__device__ __forceinline__ void block_minmax_store(double amin, double omax, GLOBAL double *part_slot) {
  __shared__ double s_min[kTPB / 64], s_max[kTPB / 64];
  for (int off = 32; off > 0; off >>= 1) {
    double a = __shfl_xor(amin, off), b = __shfl_xor(omax, off);
    amin = a < amin ? a : amin;
    omax = b > omax ? b : omax;
  }
  int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane == 0) {
    s_min[wv] = amin;
    s_max[wv] = omax;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double a = s_min[0], b = s_max[0];
    for (int k = 1; k < kTPB / 64; k++) {
      a = s_min[k] < a ? s_min[k] : a;
      b = s_max[k] > b ? s_max[k] : b;
    }
    part_slot[0] = a;
    part_slot[1] = b;
  }
  __syncthreads();
}

// Quiet-star frame (the quiet form of lower_resolution_gpu, SOAP_initialize.cu:
// 446-505, A = Q): o = sum_d c_d (1 - Q_{i+d}) for q = 0 (normalised) and 1 (SED),
// written to the frame slots of the lower-spectrum layout [flux, bconv, tot] =
// [SED, normalised, SED] (run_all_SOAP.py:359-362), with the partial min(Q) and
// max(o != 0) (seed -60000) of each tile per slot.
__device__ __forceinline__ void conv_quiet_body(const ConvArgs &c, const double *ctap, const double *quiet_, double *frame_,
                                                double *part_, int bx_, int by_) {
  const GLOBAL double *quiet = gl(quiet_);
  GLOBAL double *frame = gw(frame_), *part = gw(part_);
  extern __shared__ double xs[];
  int q = by_;
  long i0 = (long)bx_ * kConvTile;
  int span = kConvTile + c.ntp - 1;
  const GLOBAL double *Q = quiet + (long)q * c.n;
  conv_stage<true>(xs, Q, i0, c.dlo, span, c.n);
  __syncthreads();
  double acc[kConvB];
  conv_accumulate(xs, gc(ctap), c.ntp, acc);
  __syncthreads();
  // results through LDS so the global stores and loads below are coalesced
#pragma unroll
  for (int b = 0; b < kConvB; b++) xs[threadIdx.x * kConvB + b] = acc[b];
  __syncthreads();
  double amin = HUGE_VAL, omax = -60000.0;
  int s0 = q == 0 ? 1 : 0;  // slot of this quiet spectrum; SED also fills slot 2
  for (int e = threadIdx.x; e < kConvTile; e += blockDim.x) {
    long i = i0 + e;
    if (i >= c.n) break;
    double o = ((i > c.wid) && (i + c.wid < c.n - 1)) ? xs[e] : 0.0;
    frame[(long)s0 * c.n + i] = o;
    if (q == 1) frame[2L * c.n + i] = o;
    double Ai = Q[i];
    amin = Ai < amin ? Ai : amin;
    if (o != 0.0 && omax < o) omax = o;
  }
  block_minmax_store(amin, omax, part + ((long)s0 * c.n_tiles + bx_) * 2);
  if (q == 1 && threadIdx.x == 0) {
    part[(2L * c.n_tiles + bx_) * 2] = part[((long)s0 * c.n_tiles + bx_) * 2];
    part[(2L * c.n_tiles + bx_) * 2 + 1] = part[((long)s0 * c.n_tiles + bx_) * 2 + 1];
  }
}

// the tap vector is a separate restrict argument: no store of the kernel can reach
// it, so its uniform loads go through the scalar unit
__global__ __launch_bounds__(kTPB) void k_conv_quiet(ConvArgs c, const double *__restrict__ ctap, const double *quiet,
                                                     double *frame, double *part) {
  conv_quiet_body(c, ctap, quiet, frame, part, blockIdx.x, blockIdx.y);
}

// Active spectra of phase p = blockIdx.y, types 0..2 (flux, bconv, tot):
// o = Qc[sel] + sum_d c_d D_{i+d}; partial min(Q[sel] - D) over every pixel and
// max(o != 0).  sel = 0 for bconv, 1 otherwise (SOAP_resolution.cu:414-416).
// Active spectra of phase p, types 0..2 (flux, bconv, tot), as in
// lower_resolution_gpu (SOAP_resolution.cu:235-310): A = Q[sel] - D is staged as
// 1 - A, o = sum_d c_d (1 - A_{i+d}) for wid < i < n-1-wid (0 elsewhere), with
// the partial min(A) over the tile's pixels and max(o != 0) (seed -60000).
// sel = 0 for bconv, 1 otherwise (SOAP_resolution.cu:414-416).  One block per
// (lambda tile, phase, type); blocks are numbered so that all spectra of one
// lambda tile run back to back on one XCD (block b -> XCD b % 8): the tile's
// quiet rows stay in that XCD's L2.  Staging issues kStageU loads of Q and D per
// thread before their first use.
// xcd: blocks in groups of 8 tiles (tiles padded to a multiple of 8, block b on XCD
// b % 8 keeps a tile's spectra on one XCD); otherwise block = tile + n_tiles (3 p + t)
// (a spectrum of fewer than 8 tiles), the count padded to a multiple of 8 only
__device__ __forceinline__ void conv_active_body(const ConvArgs &c, const double *ctap, const double *quiet_,
                                                 const double *raw_, double *o_out_, double *part_, int n_phase,
                                                 int bx_, int by_, bool xcd = true) {
  const GLOBAL double *quiet = gl(quiet_), *raw = gl(raw_);
  GLOBAL double *o_out = gw(o_out_), *part = gw(part_);
  extern __shared__ double xs[];
  int tile, p, t;
  if (xcd) {
    const int slot = bx_ >> 3;
    tile = (slot / (3 * n_phase)) * 8 + (bx_ & 7);
    p = (slot % (3 * n_phase)) / 3;
    t = slot % 3;
  } else {
    tile = bx_ % c.n_tiles;
    const int slot = bx_ / c.n_tiles;
    p = slot / 3;
    t = slot % 3;
    if (p >= n_phase) return;  // the count is padded to a multiple of 8 (uniform over the block)
  }
  if (tile >= c.n_tiles) return;  // uniform over the block
  long i0 = (long)tile * kConvTile;
  int span = kConvTile + c.ntp - 1;
  {
    int sel = t == 1 ? 0 : 1;
    long sp = (long)p * 3 + t;
    const GLOBAL double *D = raw + sp * c.n;
    const GLOBAL double *Q = quiet + (long)sel * c.n;
    double amin = HUGE_VAL, omax = -60000.0;
    // kStageU loads of Q and of D in flight per thread before the first use
    // (clamped addresses, so every load is unconditional)
    const long s0 = i0 + c.dlo;  // source index of span position 0
    if (c.n < kConvBufN) {
      // buffer loads with 32-bit byte offsets from the spectrum's start: an offset
      // outside [0, 8n) (a negative index wraps above it) reads 0, so A = 0 and the
      // staged 1 - A is exactly the 1 the reference pads with; every tile takes this
      // one path, kStageA slots of kTPB per pass (the LDS holds the rounded span)
      const __amdgpu_buffer_rsrc_t rq = __builtin_amdgcn_make_buffer_rsrc((void *)Q, 0, (int)(c.n * 8), 0x00020000);
      const __amdgpu_buffer_rsrc_t rd = __builtin_amdgcn_make_buffer_rsrc((void *)D, 0, (int)(c.n * 8), 0x00020000);
      const int m0 = (int)s0 + (int)threadIdx.x;     // source index of this thread's slot 0
      const int lo = (int)i0 - m0;                    // m in the tile <=> (unsigned)(m - i0) < tl
      const unsigned tl = (unsigned)min((long)kConvTile, c.n - i0);
      const int nu = (span + kTPB - 1) / kTPB;
      for (int u0 = 0; u0 < nu; u0 += kStageA) {
        double qv[kStageA], dv[kStageA];
#pragma unroll
        for (int u = 0; u < kStageA; u++) {
          const int off = (m0 + (u0 + u) * kTPB) * 8;
          qv[u] = __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(rq, off, 0, 0));
          dv[u] = __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(rd, off, 0, 0));
        }
        __builtin_amdgcn_sched_barrier(0);  // keep every load above the first use
#pragma unroll
        for (int u = 0; u < kStageA; u++) {
          const double A = qv[u] - dv[u];
          const bool in = (unsigned)((u0 + u) * kTPB - lo) < tl;
          amin = (in && A < amin) ? A : amin;
          xs[(u0 + u) * kTPB + threadIdx.x] = 1.0 - A;
        }
      }
    } else if (s0 >= 0 && s0 + span <= c.n) {
      // the whole span inside the spectrum (every tile but the first and last):
      // 32-bit offsets from the span's start, no range selects
      const GLOBAL double *Qs = Q + s0, *Ds = D + s0;
      const int t0 = -c.dlo;  // span position of the tile's first pixel
      for (int e0 = 0; e0 < span; e0 += kStageU * kTPB) {
        double qv[kStageU], dv[kStageU];
#pragma unroll
        for (int u = 0; u < kStageU; u++) {
          const int e = min(e0 + (int)threadIdx.x + u * kTPB, span - 1);
          qv[u] = Qs[e];
          dv[u] = Ds[e];
        }
        __builtin_amdgcn_sched_barrier(0);  // keep every load above the first use
#pragma unroll
        for (int u = 0; u < kStageU; u++) {
          const int e = e0 + threadIdx.x + u * kTPB;
          const double A = qv[u] - dv[u];
          amin = (e >= t0 && e < t0 + kConvTile && A < amin) ? A : amin;
          if (e < span) xs[e] = 1.0 - A;
        }
      }
    } else {
      for (int e0 = 0; e0 < span; e0 += kStageU * kTPB) {
        double qv[kStageU], dv[kStageU];
#pragma unroll
        for (int u = 0; u < kStageU; u++) {
          long m = s0 + e0 + threadIdx.x + u * kTPB;
          long mc = m < 0 ? 0 : (m >= c.n ? c.n - 1 : m);
          qv[u] = Q[mc];
          dv[u] = D[mc];
        }
        __builtin_amdgcn_sched_barrier(0);  // keep every load above the first use
#pragma unroll
        for (int u = 0; u < kStageU; u++) {  // selects, not branches: the loads stay in flight together
          int e = e0 + threadIdx.x + u * kTPB;
          long m = s0 + e;
          const bool in = m >= 0 && m < c.n;
          const double A = qv[u] - dv[u];
          const double v = in ? 1.0 - A : 1.0;
          amin = (in && m >= i0 && m < i0 + kConvTile && A < amin) ? A : amin;
          if (e < span) xs[e] = v;
        }
      }
    }
    __syncthreads();
    double acc[kConvB];
    conv_accumulate(xs, gc(ctap), c.ntp, acc);
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kConvB; k++) xs[threadIdx.x * kConvB + k] = acc[k];
    __syncthreads();
    GLOBAL double *ob = o_out + sp * c.n + i0;
    if (i0 > c.wid && i0 + kConvTile + c.wid < c.n - 1) {  // every output of the tile exists
#pragma unroll
      for (int k = 0; k < kConvB; k++) {
        const int e = k * kTPB + threadIdx.x;
        const double o = xs[e];
        __builtin_nontemporal_store(o, &ob[e]);  // read back only by the normalisation pass
        if (o != 0.0 && omax < o) omax = o;
      }
    } else {
      for (int e = threadIdx.x; e < kConvTile; e += blockDim.x) {
        long i = i0 + e;
        if (i >= c.n) break;
        double o = ((i > c.wid) && (i + c.wid < c.n - 1)) ? xs[e] : 0.0;
        ob[e] = o;
        if (o != 0.0 && omax < o) omax = o;
      }
    }
    block_minmax_store(amin, omax, part + (sp * c.n_tiles + tile) * 2);  // ends with a barrier
  }
}

__global__ __launch_bounds__(kTPB) void k_conv_active(ConvArgs c, const double *__restrict__ ctap, const double *quiet,
                                                      const double *raw, double *o_out, double *part, int n_phase) {
  conv_active_body(c, ctap, quiet, raw, o_out, part, n_phase, blockIdx.x, blockIdx.y);
}

// the window slack of conv_accumulate (ntp + 1 past the span), or the span rounded up to
// whole passes of the buffer-load staging, whichever is larger
static size_t conv_lds_doubles(int ntp) {
  const long span = kConvTile + ntp - 1;
  const long pass = (long)kStageA * kTPB;
  return (size_t)max((long)kConvTile + 2 * ntp + 1, (span + pass - 1) / pass * pass);
}
static size_t conv_lds(const ConvArgs &c) { return conv_lds_doubles(c.ntp) * sizeof(double); }

hipError_t launch_conv_quiet(const ConvArgs &c, const double *quiet, double *frame, double *part, hipStream_t s) {
  if (c.ntp % kConvCH || c.ntp > kConvMaxTaps) return hipErrorInvalidValue;
  size_t lds = conv_lds(c);
  if (lds > 65536) return hipErrorInvalidValue;
  dim3 grid(c.n_tiles, 2);
  k_conv_quiet<<<grid, kTPB, lds, s>>>(c, c.ctap, quiet, frame, part);
  return hipGetLastError();
}

hipError_t launch_conv_active(const ConvArgs &c, const double *quiet, const double *raw, int n_phase, double *o_out,
                              double *part, hipStream_t s) {
  if (n_phase == 0) return hipSuccess;
  if (c.ntp % kConvCH || c.ntp > kConvMaxTaps) return hipErrorInvalidValue;
  size_t lds = conv_lds(c);
  long tiles8 = ((long)c.n_tiles + 7) / 8 * 8;
  long blocks = tiles8 * n_phase * 3;
  if (blocks > 0x7fffffffL || lds > 65536) return hipErrorInvalidValue;
  k_conv_active<<<(unsigned)blocks, kTPB, lds, s>>>(c, c.ctap, quiet, raw, o_out, part, n_phase);
  return hipGetLastError();
}

int conv_tile() { return kConvTile; }
int conv_chunk() { return kConvCH; }

// min / max of each spectrum over its tile partials (exact: order independent)
__device__ __forceinline__ void minmax_body(const double *part_, int n_tiles, double *mm_, int bx_, int by_) {
  const GLOBAL double *part = gl(part_);
  GLOBAL double *mm = gw(mm_);
  __shared__ double s_min[kTPB], s_max[kTPB];
  int sp = bx_;
  double a = HUGE_VAL, b = -HUGE_VAL;
  for (int t = threadIdx.x; t < n_tiles; t += blockDim.x) {
    double pa = part[((long)sp * n_tiles + t) * 2], pb = part[((long)sp * n_tiles + t) * 2 + 1];
    a = pa < a ? pa : a;
    b = pb > b ? pb : b;
  }
  s_min[threadIdx.x] = a;
  s_max[threadIdx.x] = b;
  __syncthreads();
  for (int off = kTPB / 2; off > 0; off >>= 1) {
    if (threadIdx.x < off) {
      s_min[threadIdx.x] = s_min[threadIdx.x + off] < s_min[threadIdx.x] ? s_min[threadIdx.x + off] : s_min[threadIdx.x];
      s_max[threadIdx.x] = s_max[threadIdx.x + off] > s_max[threadIdx.x] ? s_max[threadIdx.x + off] : s_max[threadIdx.x];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    mm[sp * 2] = s_min[0];
    mm[sp * 2 + 1] = s_max[0];
  }
}

__global__ __launch_bounds__(kTPB) void k_minmax_reduce(const double *part, int n_tiles, double *mm) {
  minmax_body(part, n_tiles, mm, blockIdx.x, blockIdx.y);
}

// spec_norm, SOAP_resolution.cu:225-232: out = 1 - in (1 - min) / max, from the
// unnormalised spectra into the output buffer.  Spectrum sp = by_ at in + sp*n and
// out + sp*n; a block of nt threads covers kNormPer*nt pixels as 16-byte
// (two-pixel) loads, all issued before the stores.  Spectra start at multiples of
// n pixels, so an odd n leaves odd-aligned spectra: those take the one-pixel path
// for their first pixel (in and out have the same 16-byte alignment).
constexpr int kNormPer = 8;
typedef double dvec2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ void norm_body(long n, const double *mm_, const double *in_, double *o_, int bx_, int by_,
                                          int nt) {
  const GLOBAL double *mm = gl(mm_);
  const long sp = by_;
  const double mn = mm[sp * 2], mx = mm[sp * 2 + 1];
  const GLOBAL double *x = gl(in_) + sp * n;
  GLOBAL double *y = gw(o_) + sp * n;
  const long blk = (long)bx_ * (nt * kNormPer);
  const long end = min(n, blk + (long)nt * kNormPer);
  // pairs start at the first 16-byte boundary of the block; a block that starts
  // off the boundary (odd n, odd spectrum) does its first pixel alone
  const int s0 = ((uintptr_t)(x + blk) & 15) ? 1 : 0;
  constexpr int V = kNormPer / 2;
  dvec2 v[V];
#pragma unroll
  for (int k = 0; k < V; k++) {
    long i = blk + s0 + 2L * (threadIdx.x + (long)k * nt);
    if (i + 1 < end) {
      v[k] = *(const GLOBAL dvec2 *)(x + i);
    } else {
      v[k].x = i < end ? x[i] : 0.0;
      v[k].y = 0.0;
    }
  }
  double head = (s0 && threadIdx.x == 0 && blk < n) ? x[blk] : 0.0;
#pragma unroll
  for (int k = 0; k < V; k++) {
    long i = blk + s0 + 2L * (threadIdx.x + (long)k * nt);
    dvec2 r;
    r.x = 1.0 - v[k].x * (1 - mn) / mx;
    r.y = 1.0 - v[k].y * (1 - mn) / mx;
    if (i + 1 < end)
      *(GLOBAL dvec2 *)(y + i) = r;
    else if (i < end)
      y[i] = r.x;
  }
  if (s0 && threadIdx.x == 0 && blk < n) y[blk] = 1.0 - head * (1 - mn) / mx;
}

__global__ __launch_bounds__(kTPB) void k_norm(long n, const double *mm, const double *in, double *o) {
  norm_body(n, mm, in, o, blockIdx.x, blockIdx.y, kTPB);
}

hipError_t launch_minmax(int n_spec, int n_tiles, const double *part, double *mm, hipStream_t s) {
  if (n_spec == 0) return hipSuccess;
  k_minmax_reduce<<<n_spec, kTPB, 0, s>>>(part, n_tiles, mm);
  return hipGetLastError();
}

hipError_t launch_norm(int64_t n, int n_spec, const double *mm, const double *in, double *o, hipStream_t s) {
  if (n_spec == 0) return hipSuccess;
  if (((uintptr_t)in ^ (uintptr_t)o) & 15) return hipErrorInvalidValue;
  long bx = (n + (long)kTPB * kNormPer - 1) / ((long)kTPB * kNormPer);
  dim3 grid((unsigned)bx, n_spec);
  k_norm<<<grid, kTPB, 0, s>>>(n, mm, in, o);
  return hipGetLastError();
}

// ------------------------------------------------------------------ a8: CCF
// Mask of calc_mask (CCF_func.py:45-91) for every (velocity, line), log grid:
// index arithmetic through x(l) = ln(l / l_t0)/h on the trimmed grid.  Tables
// are [q][v] with q the line's position in wavelength order (line = sorted[q]):
// start = aa-1, len = bb-aa+2, wv = {w fb, w fa}.
__global__ void k_mask_build(const double *line_l, const double *line_w, const int32_t *sorted, int n_lines,
                             const double *vgrid, int n_v, double lam_t0, double h, int32_t *m_start, int32_t *m_len,
                             double2 *m_wv) {
  long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long)n_v * n_lines) return;
  int q = (int)(e / n_v), iv = (int)(e % n_v);
  int li = sorted[q];
  double ls = line_l[li] * (1 + vgrid[iv] / kC);
  double beta = 820. / kC;
  double hw = ls * (sqrt((1 + beta) / (1 - beta)) - 1);
  double bgn = ls - hw / 2., end = ls + hw / 2.;
  double xb = log(bgn / lam_t0) / h, xe = log(end / lam_t0) / h;
  double aa = ceil(xb);
  double bb = ceil(xe) - 1;
  double fb = -expm1(-(aa - xb) * h) / -expm1(-h);
  double fa = fabs(expm1((xe - bb) * h)) / expm1(h);
  double w = line_w[li];
  m_start[e] = (int32_t)aa - 1;
  m_len[e] = (int32_t)(bb - aa) + 2;
  m_wv[e] = make_double2(w * fb, w * fa);
}

hipError_t launch_mask_build(const double *line_l, const double *line_w, const int32_t *sorted, int n_lines,
                             const double *vgrid, int n_v, double lam_t0, double h, int32_t *m_start, int32_t *m_len,
                             double2 *m_wv, hipStream_t s) {
  long tot = (long)n_v * n_lines;
  if (tot == 0) return hipSuccess;
  k_mask_build<<<(unsigned)((tot + kTPB - 1) / kTPB), kTPB, 0, s>>>(line_l, line_w, sorted, n_lines, vgrid, n_v,
                                                                    lam_t0, h, m_start, m_len, m_wv);
  return hipGetLastError();
}

// Later lines overwrite earlier ones (CCF_func.py:85-89): pixel p of a line is
// alive unless a line with a larger input index writes p.  Candidates are the
// wavelength-order neighbours whose pixel spans overlap.  Output
// m_pk = (alive << 8) | len (len <= 24).
__global__ void k_mask_overwrite(const int32_t *sorted, int n_lines, int n_v, int scan, const int32_t *m_start,
                                 const int32_t *m_len, uint32_t *m_pk) {
  long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long)n_v * n_lines) return;
  int q = (int)(e / n_v), iv = (int)(e % n_v);
  const int32_t *st = m_start + iv;  // element q' at st[q' * n_v]
  const int32_t *ln = m_len + iv;
  int s0 = st[(long)q * n_v], l0 = ln[(long)q * n_v];
  int li = sorted[q];
  uint32_t alive = (1u << l0) - 1u;
  for (int dir = -1; dir <= 1; dir += 2) {
    for (int step = 1; step <= scan; step++) {
      int qq = q + dir * step;
      if (qq < 0 || qq >= n_lines) break;
      int s1 = st[(long)qq * n_v], l1 = ln[(long)qq * n_v];
      if (s1 + l1 - 1 < s0 || s1 > s0 + l0 - 1) continue;
      if (sorted[qq] <= li) continue;
      for (int b = 0; b < l0; b++) {
        int p = s0 + b;
        if (p >= s1 && p <= s1 + l1 - 1) alive &= ~(1u << b);
      }
    }
  }
  m_pk[e] = (alive << 8) | (uint32_t)l0;
}

hipError_t launch_mask_overwrite(const int32_t *sorted, int n_lines, int n_v, int scan, const int32_t *m_start,
                                 const int32_t *m_len, uint32_t *m_pk, hipStream_t s) {
  long tot = (long)n_v * n_lines;
  if (tot == 0) return hipSuccess;
  k_mask_overwrite<<<(unsigned)((tot + kTPB - 1) / kTPB), kTPB, 0, s>>>(sorted, n_lines, n_v, scan, m_start, m_len,
                                                                        m_pk);
  return hipGetLastError();
}

// calculate_CCF_2 (CCF_func.py:27-36) on the sparse mask.  Block = (chunk of
// kCcfLines lines in wavelength order, chunk of kCcfSpec spectra); thread = one
// velocity, summing its lines in order for all spectra of the chunk (each mask
// entry is read once per block).  Each line's pixel footprint is staged in LDS
// first: consecutive velocities' windows overlap, so a pixel is loaded once per
// line instead of once per window.  Partials per line chunk are summed in order
// by k_ccf_reduce, which also normalises by the maximum and appodises
// (run_all_SOAP.py:368-370).
//
// The series takes the CCF of the spec_norm output y = 1 - x (1 - min) / max.
// That map is affine per spectrum, so the CCF of y is W(v) - C_x(v) (1 - min) /
// max with C_x the CCF of the unnormalised spectrum x and W(v) the sum of the
// live mask weights at velocity v (the CCF of a spectrum of ones, k_mask_wsum).
// The CCF therefore runs on x, and spec_norm does not have to finish before it:
// the two share one launch (k_norm_ccf).
constexpr int kCcfLines = 32;
constexpr int kCcfSpec = 8;
constexpr int kCcfLdsCap = 512;  // LDS segment of the footprint form: 8 spectra x 512 pixels = 32 KiB

// Line footprints for the LDS form below: fp[q] = (lo, len), the trimmed pixels
// [lo, lo + len) that the window of line q touches at any velocity, clipped to
// [0, n_t) (len 0: none).
__global__ __launch_bounds__(kTPB) void k_mask_footprint(int n_lines, int n_v, int64_t n_t, const int32_t *m_start,
                                                         const uint32_t *m_pk, int2 *fp) {
  const int q = blockIdx.x * kTPB + threadIdx.x;
  if (q >= n_lines) return;
  long lo = n_t, hi = -1;
  for (int iv = 0; iv < n_v; iv++) {
    const long e = (long)q * n_v + iv;
    const int ln = (int)(m_pk[e] & 0xffu);
    if (ln == 0) continue;
    const long st = m_start[e];
    const long a = st > 0 ? st : 0, b = st + ln - 1 < n_t - 1 ? st + ln - 1 : n_t - 1;
    if (a > b) continue;
    lo = a < lo ? a : lo;
    hi = b > hi ? b : hi;
  }
  fp[q] = hi >= lo ? make_int2((int)lo, (int)(hi - lo + 1)) : make_int2(0, 0);
}

hipError_t launch_mask_footprint(int n_lines, int n_v, int64_t n_t, const int32_t *m_start, const uint32_t *m_pk,
                                 int2 *fp, hipStream_t s) {
  if (n_lines == 0) return hipSuccess;
  k_mask_footprint<<<(n_lines + kTPB - 1) / kTPB, kTPB, 0, s>>>(n_lines, n_v, n_t, m_start, m_pk, fp);
  return hipGetLastError();
}

// Per line, the block copies the footprint pixels of its spectra with coalesced
// loads (8 in flight per thread), then every thread (one velocity; blockDim.x >=
// n_v) adds its window's pixels from LDS, in pixel order.  A footprint longer
// than cap pixels goes through in segments of cap (a window that straddles two
// takes its low pixels from the first, so the order of the terms is kept).
// lds: kCcfSpec * cap doubles.
__device__ __forceinline__ void ccf_lds_body(const double *spec_, int64_t spec_stride, int64_t trim, int n_spec,
                                             int n_lines, int n_v, int64_t n_t, const int32_t *m_start_,
                                             const uint32_t *m_pk_, const double2 *m_wv_, const double *w_sorted_,
                                             const int2 *fp_, int cap, double *part_, int bx_, int by_,
                                             double *lds) {
  const GLOBAL double *spec = gl(spec_), *w_sorted = gl(w_sorted_);
  const GLOBAL int32_t *m_start = gl(m_start_);
  const GLOBAL uint32_t *m_pk = gl(m_pk_);
  const GLOBAL double *m_wv = gl((const double *)m_wv_);
  const GLOBAL int32_t *fp = gl((const int32_t *)fp_);  // {lo, len} pairs
  GLOBAL double *part = gw(part_);
  const int chunk = bx_;
  const int s0 = by_ * kCcfSpec;
  const int ns = min(kCcfSpec, n_spec - s0);
  const int q0 = chunk * kCcfLines, q1 = min(n_lines, q0 + kCcfLines);
  const GLOBAL double *S[kCcfSpec];
#pragma unroll
  for (int k = 0; k < kCcfSpec; k++) S[k] = spec + (long)(s0 + (k < ns ? k : 0)) * spec_stride + trim;
  const int iv = threadIdx.x;
  const bool act = iv < n_v;
  double acc[kCcfSpec];
#pragma unroll
  for (int k = 0; k < kCcfSpec; k++) acc[k] = 0.0;
  for (int q = q0; q < q1; q++) {
    // this line's mask entry, loaded before the staging so that the two overlap
    int st = 0;
    uint32_t pk = 0;
    double wx = 0.0, wy = 0.0;
    if (act) {
      const long e = (long)q * n_v + iv;
      st = m_start[e];
      pk = m_pk[e];
      wx = m_wv[2 * e];
      wy = m_wv[2 * e + 1];
    }
    const double w = w_sorted[q];
    const int f_lo = fp[2 * q], f_len = fp[2 * q + 1];
    for (int g0 = 0; g0 < f_len; g0 += cap) {
      const int seg = min(cap, f_len - g0);
      const long base = (long)f_lo + g0;
      __syncthreads();  // the previous segment's readers are done with the buffer
      for (int j0 = 0; j0 < seg; j0 += blockDim.x) {
        const int j = j0 + threadIdx.x;
        double v[kCcfSpec];
#pragma unroll
        for (int k = 0; k < kCcfSpec; k++) v[k] = (j < seg && k < ns) ? S[k][base + j] : 0.0;
        if (j < seg) {
#pragma unroll
          for (int k = 0; k < kCcfSpec; k++) lds[k * cap + j] = v[k];
        }
      }
      __syncthreads();
      if (act) {
        const int ln = (int)(pk & 0xffu);
        const uint32_t al = pk >> 8;
        for (int b = 0; b < ln; b++) {
          const long pp = st + b;
          if (!((al >> b) & 1u) || pp < 0 || pp >= n_t || pp < base || pp >= base + seg) continue;
          const double val = (b == ln - 1) ? wy : ((b == 0) ? wx : w);
          const double *L = lds + (pp - base);
#pragma unroll
          for (int k = 0; k < kCcfSpec; k++) acc[k] = fma(L[k * cap], val, acc[k]);
        }
      }
    }
  }
  if (act) {
#pragma unroll
    for (int k = 0; k < kCcfSpec; k++)  // fixed trip count: the accumulators stay in registers
      if (k < ns) part[((long)chunk * n_spec + s0 + k) * n_v + iv] = acc[k];
  }
}

__global__ void k_ccf_lds(const double *spec, int64_t spec_stride, int64_t trim, int n_spec, int n_lines, int n_v,
                          int64_t n_t, const int32_t *m_start, const uint32_t *m_pk, const double2 *m_wv,
                          const double *w_sorted, const int2 *fp, int cap, double *part) {
  extern __shared__ double ccf_lds[];
  ccf_lds_body(spec, spec_stride, trim, n_spec, n_lines, n_v, n_t, m_start, m_pk, m_wv, w_sorted, fp, cap, part,
               blockIdx.x, blockIdx.y, ccf_lds);
}

// LDS segment length for a longest footprint of fmax pixels
int ccf_lds_cap(int fmax) { return fmax < 1 ? 1 : (fmax < kCcfLdsCap ? fmax : kCcfLdsCap); }
size_t ccf_lds_bytes(int cap) { return (size_t)kCcfSpec * (cap > 0 ? cap : 1) * 8; }

hipError_t launch_ccf_lds(const double *spec, int64_t spec_stride, int64_t trim, int n_spec, int n_lines, int n_v,
                          int64_t n_t, const int32_t *m_start, const uint32_t *m_pk, const double2 *m_wv,
                          const double *w_sorted, const int2 *fp, int cap, double *part, hipStream_t s) {
  if (n_spec == 0) return hipSuccess;
  if (n_v > 1024 || cap < 1 || cap > kCcfLdsCap) return hipErrorInvalidValue;
  int nc = ccf_chunks(n_lines);
  int threads = ((n_v + 63) / 64) * 64;
  dim3 grid(nc, (n_spec + kCcfSpec - 1) / kCcfSpec);
  k_ccf_lds<<<grid, threads, ccf_lds_bytes(cap), s>>>(spec, spec_stride, trim, n_spec, n_lines, n_v, n_t, m_start,
                                                       m_pk, m_wv, w_sorted, fp, cap, part);
  return hipGetLastError();
}

// ccf = W - (sum of the chunk partials of C_x) (1 - min) / max, then the
// normalisation by the maximum and the appodisation
__device__ __forceinline__ void ccf_reduce_body(const double *part_, int n_chunks, int n_spec, int n_v,
                                                const double *appod_, const double *mm_, const double *wsum_,
                                                double *ccf_, int bx_, int by_) {
  const GLOBAL double *part = gl(part_), *appod = gl(appod_), *mm = gl(mm_), *wsum = gl(wsum_);
  GLOBAL double *ccf = gw(ccf_);
  __shared__ double s_red[1024];
  int sp = bx_;
  const double s_min = mm[sp * 2], s_max = mm[sp * 2 + 1];
  double v = -HUGE_VAL;
  for (int iv = threadIdx.x; iv < n_v; iv += blockDim.x) {
    // chunks added in order; kRedU loads in flight ahead of their adds
    constexpr int kRedU = 16;
    const GLOBAL double *pv = part + (long)sp * n_v + iv;
    const long cs = (long)n_spec * n_v;
    double a = 0.0;
    int c0 = 0;
    for (; c0 + kRedU <= n_chunks; c0 += kRedU) {
      double t[kRedU];
#pragma unroll
      for (int u = 0; u < kRedU; u++) t[u] = pv[(c0 + u) * cs];
#pragma unroll
      for (int u = 0; u < kRedU; u++) a += t[u];
    }
    for (; c0 < n_chunks; c0++) a += pv[c0 * cs];
    ccf[(long)sp * n_v + iv] = wsum[iv] - a * (1 - s_min) / s_max;
  }
  __syncthreads();
  // max over velocities (np.max), then normalise and appodise
  for (int iv = threadIdx.x; iv < n_v; iv += blockDim.x) v = fmax(v, ccf[(long)sp * n_v + iv]);
  s_red[threadIdx.x] = v;
  __syncthreads();
  for (int off = blockDim.x / 2; off > 0; off >>= 1) {
    if (threadIdx.x < off) s_red[threadIdx.x] = fmax(s_red[threadIdx.x], s_red[threadIdx.x + off]);
    __syncthreads();
  }
  double mx = s_red[0];
  for (int iv = threadIdx.x; iv < n_v; iv += blockDim.x) {
    double cc = ccf[(long)sp * n_v + iv] / mx;
    ccf[(long)sp * n_v + iv] = 1 - ((-cc + 1) * appod[iv]);
  }
}

__global__ void k_ccf_reduce(const double *part, int n_chunks, int n_spec, int n_v, const double *appod,
                             const double *mm, const double *wsum, double *ccf) {
  ccf_reduce_body(part, n_chunks, n_spec, n_v, appod, mm, wsum, ccf, blockIdx.x, blockIdx.y);
}

// W(v): per velocity, the live mask weights the CCF multiplies in-range pixels
// with, summed over the lines (one block per velocity, lines strided over the
// threads, then a fixed-order tree)
__global__ __launch_bounds__(kTPB) void k_mask_wsum(int n_lines, int n_v, int64_t n_t, const int32_t *m_start,
                                                    const uint32_t *m_pk, const double2 *m_wv, const double *w_sorted,
                                                    double *wsum) {
  __shared__ double s_red[kTPB];
  const int iv = blockIdx.x;
  const double *wv = (const double *)m_wv;
  double acc = 0.0;
  for (int q = threadIdx.x; q < n_lines; q += kTPB) {
    long e = (long)q * n_v + iv;
    int st = m_start[e];
    uint32_t pk = m_pk[e];
    int ln = (int)(pk & 0xffu);
    uint32_t al = pk >> 8;
    for (int b = 0; b < ln; b++) {
      long pp = st + b;
      if (!((al >> b) & 1u) || pp < 0 || pp >= n_t) continue;
      acc += (b == ln - 1) ? wv[2 * e + 1] : ((b == 0) ? wv[2 * e] : w_sorted[q]);
    }
  }
  s_red[threadIdx.x] = acc;
  __syncthreads();
  for (int off = kTPB / 2; off > 0; off >>= 1) {
    if (threadIdx.x < off) s_red[threadIdx.x] += s_red[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) wsum[iv] = s_red[0];
}

hipError_t launch_mask_wsum(int n_lines, int n_v, int64_t n_t, const int32_t *m_start, const uint32_t *m_pk,
                            const double2 *m_wv, const double *w_sorted, double *wsum, hipStream_t s) {
  if (n_v == 0) return hipSuccess;
  k_mask_wsum<<<n_v, kTPB, 0, s>>>(n_lines, n_v, n_t, m_start, m_pk, m_wv, w_sorted, wsum);
  return hipGetLastError();
}

// spec_norm and the CCF of the unnormalised spectra in one launch: the blocks come
// in groups of one CCF block followed by `ratio` norm blocks (ratio = n_norm /
// n_ccf), the remaining norm blocks after the groups, so that every CU holds both
// kinds and the norm's streaming runs beside the CCF's gathers.  blockDim.x is the
// CCF's thread count (velocities rounded up to a wave); a norm block covers
// kNormPer * blockDim.x pixels of one spectrum.
__device__ __forceinline__ void norm_ccf_body(const NormCcfArgs &a, long b, double *ccf_lds) {
  const long grp = (long)a.ratio + 1;
  const long mixed = (long)a.n_ccf * grp;
  long ci = -1, ni;
  if (b < mixed) {
    long g = b / grp, r = b - g * grp;
    if (r == 0)
      ci = g;
    else
      ni = g * a.ratio + r - 1;
  } else {
    ni = (long)a.n_ccf * a.ratio + (b - mixed);
  }
  if (ci >= 0)
    ccf_lds_body(a.raw, a.n, a.trim, a.n_spec, a.n_lines, a.n_v, a.n_t, a.m_start, a.m_pk, a.m_wv, a.w_sorted, a.fp,
                 a.cap, a.part, (int)(ci % a.nc), (int)(ci / a.nc), ccf_lds);
  else if (ni < a.n_norm)
    norm_body(a.n, a.mm, a.raw, a.out, (int)(ni % a.nbx), (int)(ni / a.nbx), blockDim.x);
}

__global__ void k_norm_ccf(NormCcfArgs a) {
  extern __shared__ double ccf_lds[];
  norm_ccf_body(a, blockIdx.x, ccf_lds);
}

hipError_t launch_norm_ccf(NormCcfArgs a, hipStream_t s) {
  if (a.n_spec == 0) return hipSuccess;
  if (a.n_v > 1024 || (((uintptr_t)a.raw ^ (uintptr_t)a.out) & 15)) return hipErrorInvalidValue;
  if (a.cap < 1 || a.cap > kCcfLdsCap) return hipErrorInvalidValue;
  const int threads = ((a.n_v + 63) / 64) * 64;
  a.nc = ccf_chunks(a.n_lines);
  a.n_ccf = a.nc * ((a.n_spec + kCcfSpec - 1) / kCcfSpec);
  a.nbx = (int)((a.n + (int64_t)threads * kNormPer - 1) / ((int64_t)threads * kNormPer));
  a.n_norm = a.nbx * a.n_spec;
  a.ratio = a.n_ccf > 0 ? a.n_norm / a.n_ccf : 0;
  const long blocks = (long)a.n_ccf + a.n_norm;
  if (blocks > 0x7fffffffL) return hipErrorInvalidValue;
  k_norm_ccf<<<(unsigned)blocks, threads, ccf_lds_bytes(a.cap), s>>>(a);
  return hipGetLastError();
}

int ccf_chunks(int n_lines) { return (n_lines + kCcfLines - 1) / kCcfLines; }

hipError_t launch_ccf_reduce(const double *part, int nc, int n_spec, int n_v, const double *appod, const double *mm,
                             const double *wsum, double *ccf, hipStream_t s) {
  if (n_spec == 0) return hipSuccess;
  if (n_v > 1024) return hipErrorInvalidValue;
  int rt = 1;
  while (rt < n_v && rt < 1024) rt <<= 1;
  rt = rt < 64 ? 64 : rt;
  k_ccf_reduce<<<n_spec, rt, 0, s>>>(part, nc, n_spec, n_v, appod, mm, wsum, ccf);
  return hipGetLastError();
}

// raw CCF (no normalisation): ccf[sp][v] = sum over line chunks, in chunk order
__global__ void k_ccf_sum(const double *part, int n_chunks, int n_spec, int n_v, double *ccf) {
  long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long)n_spec * n_v) return;
  long sp = e / n_v, iv = e % n_v;
  double a = 0.0;
  for (int c = 0; c < n_chunks; c++) a += part[((long)c * n_spec + sp) * n_v + iv];
  ccf[e] = a;
}

hipError_t launch_ccf_sum(const double *part, int nc, int n_spec, int n_v, double *ccf, hipStream_t s) {
  long tot = (long)n_spec * n_v;
  if (tot == 0) return hipSuccess;
  k_ccf_sum<<<(unsigned)((tot + kTPB - 1) / kTPB), kTPB, 0, s>>>(part, nc, n_spec, n_v, ccf);
  return hipGetLastError();
}

// ------------------------------------------------------------------ a9: fit
constexpr int kMaxV = 1024;

__device__ __forceinline__ double wave_sum(double x) {
  for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off);
  return x;
}

struct GaussEval {
  double sum_f2;
  double A[10];  // upper triangle of J^T J: 00 01 02 03 11 12 13 22 23 33
  double g[4];
  double colsq[4];
};

// gauss_f / gauss_df of BIS_FIT2.cpp:31-81, lane-strided over points: residual
// (fit - y) / err and Jacobian rows divided by err (BIS_FIT2.cpp:49,74-77); err
// NULL = unit errors (the driver's CCF_error.txt of ones, run_all_SOAP.py:317).
// with_jac = false evaluates only sum f^2.
__device__ void gauss_eval(const double *px, const double *y, const double *err, int n, const double *x,
                           bool with_jac, GaussEval &r) {
  double c = x[0], k = x[1], x0 = x[2], fwhm = x[3];
  double sigma = fwhm / 2 / sqrt(2 * log(2.0));
  double f2 = 0, A[10] = {0}, g[4] = {0};
  for (int i = threadIdx.x; i < n; i += 64) {
    double e = exp(-(px[i] - x0) * (px[i] - x0) / 2 / sigma / sigma);
    double fi = err ? ((c + k * e) - y[i]) / err[i] : (c + k * e) - y[i];
    f2 += fi * fi;
    if (with_jac) {
      double J[4];
      if (err) {
        double ei = err[i];
        J[0] = 1 / ei;
        J[1] = e / ei;
        J[2] = k / ei * e * (px[i] - x0) / sigma / sigma;
        J[3] = k / ei * e * (px[i] - x0) * (px[i] - x0) / sigma / sigma / sigma / 2 / sqrt(2 * log(2.0));
      } else {
        J[0] = 1.0;
        J[1] = e;
        J[2] = k * e * (px[i] - x0) / sigma / sigma;
        J[3] = k * e * (px[i] - x0) * (px[i] - x0) / sigma / sigma / sigma / 2 / sqrt(2 * log(2.0));
      }
      int q = 0;
      for (int a = 0; a < 4; a++) {
        g[a] += J[a] * fi;
        for (int b = a; b < 4; b++) A[q++] += J[a] * J[b];
      }
    }
  }
  r.sum_f2 = wave_sum(f2);
  if (with_jac) {
    for (int q = 0; q < 10; q++) r.A[q] = wave_sum(A[q]);
    for (int a = 0; a < 4; a++) r.g[a] = wave_sum(g[a]);
    r.colsq[0] = r.A[0];
    r.colsq[1] = r.A[4];
    r.colsq[2] = r.A[7];
    r.colsq[3] = r.A[9];
  }
}

__device__ __forceinline__ double symA(const double *A, int a, int b) {
  if (a > b) {
    int t = a;
    a = b;
    b = t;
  }
  const int idx[4][4] = {{0, 1, 2, 3}, {1, 4, 5, 6}, {2, 5, 7, 8}, {3, 6, 8, 9}};
  return A[idx[a][b]];
}

__device__ int solve4(const double *A, const double *D, double mu, const double *g, double *dx) {
  double L[4][4] = {{0}};
  for (int i = 0; i < 4; i++) {
    for (int j = 0; j <= i; j++) {
      double s = symA(A, i, j) + (i == j ? mu * D[i] * D[i] : 0.0);
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

// BIS_FIT2 (BIS_FIT2.cpp:97-320) for one CCF per 64-lane block: the same
// Levenberg-Marquardt restatement of GSL's trust-region solver as the oracle
// (More scaling, Nielsen damping, xtol = gtol = 1e-8), then the bisector.
// E_: per-point errors [n_ccf][n] (CCF_error.txt), NULL = unit errors; held in
// s_p0 during the fit (s_p0 is bisector scratch afterwards).
// status_ (optional): 0 when the xtol or gtol test ended the fit, 27 when no step
// reduced the residual (GSL_ENOPROG: a flat or degenerate CCF), 11 after 10000
// iterations (GSL_EMAXITER).
__device__ __forceinline__ void fit_body(const double *v_, int n, const double *Y_, const double *E_, double *par_,
                                         double *model_, double *depth_out_, double *bis_out_, int32_t *iters_,
                                         int bx_, int32_t *status_ = nullptr) {
  const GLOBAL double *v = gl(v_), *Y = gl(Y_);
  GLOBAL double *par = gw(par_), *model = gw(model_), *depth_out = gw(depth_out_), *bis_out = gw(bis_out_);
  GLOBAL int32_t *iters = gw(iters_);
  __shared__ double s_px[kMaxV], s_y[kMaxV], s_nc[kMaxV], s_p0[kMaxV], s_p1[kMaxV], s_p2[kMaxV];
  __shared__ double s_bis[91];
  __shared__ double s_init[4];
  int cidx = bx_;
  const GLOBAL double *y = Y + (long)cidx * n;
  for (int i = threadIdx.x; i < n; i += 64) {
    s_px[i] = v[i] - v[0];
    s_y[i] = y[i];
  }
  const double *s_err = nullptr;
  if (E_) {
    const GLOBAL double *er = gl(E_) + (long)cidx * n;
    for (int i = threadIdx.x; i < n; i += 64) s_p0[i] = er[i];
    s_err = s_p0;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double c = (s_y[0] + s_y[n - 1]) / 2;
    double fwhm = (s_px[n - 1] - s_px[0]) / 5.;
    double k = 0., x0 = s_px[n / 2];
    for (int i = 0; i < n; i++)
      if (fabs(s_y[i] - c) > fabs(k)) {
        k = s_y[i] - c;
        x0 = s_px[i];
      }
    s_init[0] = c;
    s_init[1] = k;
    s_init[2] = x0;
    s_init[3] = fwhm;
  }
  __syncthreads();
  double x[4] = {s_init[0], s_init[1], s_init[2], s_init[3]};
  const double xtol = 1.0e-8, gtol = 1.0e-8;
  GaussEval ev;
  gauss_eval(s_px, s_y, s_err, n, x, true, ev);
  double D[4];
  for (int a = 0; a < 4; a++) D[a] = sqrt(ev.colsq[a]) > 0 ? sqrt(ev.colsq[a]) : 1.0;
  double mu = 1.0e-3, nu = 2.0;
  int iter = 0, fit_status = 11;
  GaussEval et;
  for (iter = 0; iter < 10000; iter++) {
    double dx[4], xt[4];
    int bad = 0;
    bool accepted = false;
    while (!accepted) {
      if (solve4(ev.A, D, mu, ev.g, dx) != 0) {
        mu *= nu;
        nu *= 2;
        if (++bad > 15) break;
        continue;
      }
      for (int a = 0; a < 4; a++) xt[a] = x[a] + dx[a];
      // the trial point with its Jacobian sums: an accepted trial's evaluation is
      // the next iteration's (the same sums at the same point)
      gauss_eval(s_px, s_y, s_err, n, xt, true, et);
      double pred = 0;
      for (int a = 0; a < 4; a++) {
        double Adx = 0;
        for (int b = 0; b < 4; b++) Adx += symA(ev.A, a, b) * dx[b];
        pred += dx[a] * Adx + 2 * mu * D[a] * D[a] * dx[a] * dx[a];
      }
      double rho = pred > 0 ? (ev.sum_f2 - et.sum_f2) / pred : -1.0;
      if (rho > 0) {
        accepted = true;
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
      fit_status = 27;
      break;
    }
    for (int a = 0; a < 4; a++) x[a] = xt[a];
    ev = et;
    for (int a = 0; a < 4; a++)
      if (sqrt(ev.colsq[a]) > D[a]) D[a] = sqrt(ev.colsq[a]);
    bool small = true;
    for (int a = 0; a < 4; a++)
      if (!(fabs(dx[a]) < xtol * xtol + xtol * fabs(x[a]))) small = false;
    if (small) {
      iter++;
      fit_status = 0;
      break;
    }
    double gn = 0;
    for (int a = 0; a < 4; a++) {
      double sc = fabs(x[a]) > 1.0 ? fabs(x[a]) : 1.0;
      if (fabs(ev.g[a] * sc) > gn) gn = fabs(ev.g[a] * sc);
    }
    double phi = 0.5 * ev.sum_f2;
    if (gn <= gtol * (phi > 1.0 ? phi : 1.0)) {
      iter++;
      fit_status = 0;
      break;
    }
  }
  double c = x[0], k = x[1], x0 = x[2], fwhm = x[3];
  double sigma_m = fwhm / 2 / sqrt(2 * log(2.0));
  if (model)
    for (int i = threadIdx.x; i < n; i += 64)
      model[(long)cidx * n + i] = c + k * exp(-(s_px[i] - x0) * (s_px[i] - x0) / 2 / sigma_m / sigma_m);
  // bisector, BIS_FIT2.cpp:214-268
  double sigma = fwhm / 2. / sqrt(2. * log(2.));
  double v0 = x0 + v[0];
  for (int i = threadIdx.x; i < n; i += 64) s_nc[i] = -c / k * (1. - s_y[i] / c);
  __syncthreads();  // also: every lane is done with s_err (s_p0) before s_p0 is rewritten
  for (int i = threadIdx.x; i < n - 1; i += 64) {
    double vr = (v[i] + v[i + 1]) / 2.;
    double s2 = sigma * sigma;
    double dd = (vr - v0) * (vr - v0);
    double dccf = -(vr - v0) / s2 * exp(-dd / 2. / s2);
    double d2ccf = (dd / s2 - 1) / s2 * exp(-dd / 2. / s2);
    double d2rv = -d2ccf / (dccf * dccf * dccf);
    double p2 = d2rv / 2.;
    double p1 = (v[i + 1] - v[i] - p2 * (s_nc[i + 1] * s_nc[i + 1] - s_nc[i] * s_nc[i])) / (s_nc[i + 1] - s_nc[i]);
    s_p2[i] = p2;
    s_p1[i] = p1;
    s_p0[i] = v[i] - p1 * s_nc[i] - p2 * (s_nc[i] * s_nc[i]);
  }
  if (threadIdx.x == 0) {
    s_p0[n - 1] = 0;
    s_p1[n - 1] = 0;
    s_p2[n - 1] = 0;
  }
  __syncthreads();
  __shared__ int s_imax;
  if (threadIdx.x == 0) {
    int im = 0;
    for (int i = 0; i < n; i++)
      if (s_nc[i] > s_nc[im]) im = i;
    s_imax = im;
  }
  __syncthreads();
  int ind_max = s_imax;
  for (int i = threadIdx.x; i < 91; i += 64) {
    double dep = (double)(i + 5) / 100;
    int i_b = ind_max, i_r = ind_max;
    while ((s_nc[i_b] > dep) && (i_b > 1)) i_b--;
    while ((s_nc[i_r + 1] > dep) && (i_r < (n - 2))) i_r++;
    double b = (s_p0[i_b] + s_p0[i_r]) + (s_p1[i_b] + s_p1[i_r]) * dep + (s_p2[i_b] + s_p2[i_r]) * (dep * dep);
    b /= 2.;
    s_bis[i] = b;
    if (bis_out) bis_out[(long)cidx * 91 + i] = b;
    if (depth_out) depth_out[(long)cidx * 91 + i] = 1.0 - fabs(k) * dep;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    int n1 = 0, n2 = 0;
    double top = 0., bot = 0.;
    for (int i = 0; i < 91; i++) {
      double dep = (double)(i + 5) / 100;
      if ((dep >= 0.1) && (dep <= 0.4)) {
        n1++;
        top += s_bis[i];
      }
      if ((dep >= 0.6) && (dep <= 0.9)) {
        n2++;
        bot += s_bis[i];
      }
    }
    top /= n1;
    bot /= n2;
    GLOBAL double *pp = par + (long)cidx * 5;
    pp[0] = c;
    pp[1] = fabs(k);
    pp[2] = top - bot;
    pp[3] = v0;
    pp[4] = fwhm;
    if (iters) iters[cidx] = iter;
    if (status_) gw(status_)[cidx] = fit_status;
  }
}

__global__ __launch_bounds__(64) void k_fit(const double *v, int n, const double *Y, const double *E, double *par,
                                            double *model, double *depth_out, double *bis_out, int32_t *iters,
                                            int32_t *status) {
  fit_body(v, n, Y, E, par, model, depth_out, bis_out, iters, blockIdx.x, status);
}

hipError_t launch_fit(const double *v, int n_v, const double *y, const double *err, int n_ccf, double *par,
                      double *model, double *depth, double *bis, int32_t *iters, hipStream_t s, int32_t *status) {
  if (n_ccf == 0) return hipSuccess;
  if (n_v > kMaxV || n_v < 3) return hipErrorInvalidValue;
  k_fit<<<n_ccf, 64, 0, s>>>(v, n_v, y, err, par, model, depth, bis, iters, status);
  return hipGetLastError();
}


// ------------------------------------------------------------------ a11: batched instances
// One launch per stage for all instances of a batch.  Each block walks the
// flattened work space with a grid stride; the owning instance of item w is
// found by binary search over the prefix of per-instance counts (uniform over
// the block, so its loads are scalar).  The body is the single-instance
// kernel's, given the item's local block index.  A persistent grid whose size
// is a multiple of 8 keeps item w on XCD w % 8, as in the single-instance
// launch of the XCD-ordered stages.
constexpr int kBatchFillChunk = kTPB * 16;

__device__ __forceinline__ int batch_find(const int64_t *__restrict__ prefix, int n_inst, int64_t w) {
  int lo = 0, hi = n_inst;  // prefix[lo] <= w < prefix[hi]
  while (hi - lo > 1) {
    int mid = (lo + hi) >> 1;
    if (prefix[mid] <= w) lo = mid;
    else hi = mid;
  }
  return lo;
}

template <bool PERSIST, class Job, class F>
__device__ __forceinline__ void batch_loop(const Job *__restrict__ jobs, const int64_t *__restrict__ prefix, int n_inst,
                                           int64_t total, F body) {
  if (!PERSIST) {  // one work item per block: no loop state live across items
    int64_t w = blockIdx.x;
    int i = batch_find(prefix, n_inst, w);
    body(jobs[i], (int)(w - prefix[i]));  // per-instance counts are < 2^31
    return;
  }
  for (int64_t w = blockIdx.x; w < total; w += gridDim.x) {
    int i = batch_find(prefix, n_inst, w);
    body(jobs[i], (int)(w - prefix[i]));
    __syncthreads();  // the next item reuses the block's LDS
  }
}

#define SOAP_BATCH_ARGS(JOB) const JOB *__restrict__ jobs, const int64_t *__restrict__ prefix, int n_inst, int64_t total

template <bool P>
__global__ __launch_bounds__(kTPB) void kb_fill(SOAP_BATCH_ARGS(FillJob)) {
  batch_loop<P>(jobs, prefix, n_inst, total, [&](const FillJob &j, int local) {
    int64_t e0 = (int64_t)local * kBatchFillChunk;
    int64_t e1 = min(j.count, e0 + kBatchFillChunk);
    GLOBAL unsigned long long *p = gw(j.p);
    for (int64_t e = e0 + threadIdx.x; e < e1; e += blockDim.x) p[e] = 0ull;
  });
}

template <bool P>
__global__ __launch_bounds__(kTPB) void kb_taps(SOAP_BATCH_ARGS(TapsJob)) {
  batch_loop<P>(jobs, prefix, n_inst, total,
                [&](const TapsJob &j, int local) { taps_active_body(j.a, local % j.mcb, local / j.mcb); });
}

template <bool P>
__global__ __launch_bounds__(256) void kb_finalize(SOAP_BATCH_ARGS(FinalizeJob)) {
  batch_loop<P>(jobs, prefix, n_inst, total, [&](const FinalizeJob &j, int local) {
    taps_finalize_body(j.acc, j.nsets_in, j.nsets_out, j.n_mu, j.nd, j.taps, j.krange, j.kflags, j.bins, j.nbins,
                       j.taps_g, j.n_job, j.p_pad, local, 0);
  });
}

// occupancy floor of the FIR blocks: the single-instance kernel's register budget
// (more waves per SIMD would spill the accumulators of the inner loop)
// (the row-prefetching form holds 16 more registers per thread)
template <bool P, int G, bool PRE>
struct FirWaves {
  static constexpr int v = PRE ? (G == 2 ? (P ? 4 : 6) : G == 4 ? 4 : 3) : G == 2 ? (P ? 7 : 8) : G == 4 ? (P ? 4 : 6) : 4;
};

template <bool P, int G, bool PRE>
__global__ __launch_bounds__(kTPB) __attribute__((amdgpu_waves_per_eu(FirWaves<P, G, PRE>::v, 8))) void kb_fir(
    SOAP_BATCH_ARGS(FirJob)) {
  batch_loop<P>(jobs, prefix, n_inst, total, [&](const FirJob &j, int local) {
    fir_active_grp_body<G, PRE>(j.a, j.n_job, j.n_tiles, j.n_groups, local);
  });
}

// the tile-stationary DPP form: entry streams, then the tiles (8 waves, 4 entries
// ahead; every instance's block count is a multiple of 8, so a block's local index
// keeps the XCD order of the single-instance launch)
template <bool P>
__global__ __launch_bounds__(kProgThreads) void kb_fir_prog(SOAP_BATCH_ARGS(FirProgJob)) {
  batch_loop<P>(jobs, prefix, n_inst, total, [&](const FirProgJob &j, int local) {
    fir_prog_body(j.a, j.pg, j.n_job, j.n_groups, j.lmax, 16, 4, local);
  });
}

template <bool P>
__global__ __launch_bounds__(512) void kb_fir_tile(SOAP_BATCH_ARGS(FirTileJob)) {
  batch_loop<P>(jobs, prefix, n_inst, total, [&](const FirTileJob &j, int local) {
    fir_tile_dpp_body<8, 4>(j.a, j.pg.taps, j.pg.xoff, j.pg.count, j.pg.ufl, j.a.out, j.pg.cap, j.n_job, j.n_groups,
                            j.n_tiles, j.tiles_per_xcd, j.t64_log2, local);
  });
}

template <bool P, int G, int B>
__global__ __launch_bounds__(kTPB) void kb_fir2(SOAP_BATCH_ARGS(FirJob)) {
  batch_loop<P>(jobs, prefix, n_inst, total, [&](const FirJob &j, int local) {
    fir2_body<G, B>(j.a, j.n_job, j.n_tiles, j.n_groups, local);
  });
}

template <bool P>
__global__ __launch_bounds__(kTPB) __attribute__((amdgpu_waves_per_eu(6, 8))) void kb_conv_quiet(
    SOAP_BATCH_ARGS(ConvQJob)) {
  batch_loop<P>(jobs, prefix, n_inst, total, [&](const ConvQJob &j, int local) {
    conv_quiet_body(j.c, j.c.ctap, j.quiet, j.frame, j.part, local % j.c.n_tiles, local / j.c.n_tiles);
  });
}

template <bool P>
__global__ __launch_bounds__(kTPB) __attribute__((amdgpu_waves_per_eu(6, 8))) void kb_conv(SOAP_BATCH_ARGS(ConvAJob)) {
  batch_loop<P>(jobs, prefix, n_inst, total, [&](const ConvAJob &j, int local) {
    conv_active_body(j.c, j.c.ctap, j.quiet, j.raw, j.out, j.part, j.n_phase, local, 0, j.xcd != 0);
  });
}

template <bool P>
__global__ __launch_bounds__(kTPB) void kb_minmax(SOAP_BATCH_ARGS(MinmaxJob)) {
  batch_loop<P>(jobs, prefix, n_inst, total,
                [&](const MinmaxJob &j, int local) { minmax_body(j.part, j.n_tiles, j.mm, local, 0); });
}

template <bool P>
__global__ __launch_bounds__(kTPB) void kb_norm(SOAP_BATCH_ARGS(NormJob)) {
  batch_loop<P>(jobs, prefix, n_inst, total, [&](const NormJob &j, int local) {
    norm_body(j.n, j.mm, j.in, j.o, local % (int)j.bx, local / (int)j.bx, kTPB);
  });
}

template <bool P>
__global__ __launch_bounds__(1024) void kb_ccf(SOAP_BATCH_ARGS(CcfJob)) {
  extern __shared__ double ccf_lds[];
  batch_loop<P>(jobs, prefix, n_inst, total, [&](const CcfJob &j, int local) {
    ccf_lds_body(j.spec, j.stride, j.trim, j.n_spec, j.n_lines, j.n_v, j.n_t, j.m_start, j.m_pk, j.m_wv, j.w_sorted,
                 j.fp, j.cap, j.part, local % j.nc, local / j.nc, ccf_lds);
  });
}

template <bool P>
__global__ __launch_bounds__(1024) void kb_ccf_reduce(SOAP_BATCH_ARGS(CcfRedJob)) {
  batch_loop<P>(jobs, prefix, n_inst, total, [&](const CcfRedJob &j, int local) {
    ccf_reduce_body(j.part, j.nc, j.n_spec, j.n_v, j.appod, j.mm, j.wsum, j.ccf, local, 0);
  });
}

template <bool P>
__global__ __launch_bounds__(64) void kb_fit(SOAP_BATCH_ARGS(FitJob)) {
  batch_loop<P>(jobs, prefix, n_inst, total, [&](const FitJob &j, int local) {
    fit_body(j.v, j.n_v, j.y, nullptr, j.par, j.model, j.depth, j.bis, j.iters, local);
  });
}

int batch_fill_chunk() { return kBatchFillChunk; }
int norm_px_per_block() { return kTPB * kNormPer; }
int ccf_spec_chunk() { return kCcfSpec; }
int fir_tile() { return kFirTile; }
size_t fir_lds(int nd) { return (size_t)4 * (kFirTile + nd - 1) * sizeof(double); }
size_t conv_lds_bytes(int ntp) { return conv_lds_doubles(ntp) * sizeof(double); }
size_t taps_lds_bytes(const TapAccArgs &a) { return (size_t)a.nsets * a.n_mu * a.wd * 2 * sizeof(unsigned long long); }

hipError_t launch_batch(int st, const void *jobs, const int64_t *prefix, int n_inst, int64_t total, int64_t grid,
                        int threads, size_t lds, int fir_g, int fir_b, int fir_pre, hipStream_t s) {
  if (total == 0) return hipSuccess;
  const bool tile = st == BS_FIR && fir_b == -7;
  if (n_inst < 1 || total < 0 || grid < 0 || lds > (tile ? 163840u : 65536u)) return hipErrorInvalidValue;
  int64_t g = grid == 0 ? total : std::min<int64_t>(grid, total);
  if (g > 0x7fffffffL) return hipErrorInvalidValue;
  unsigned nb = (unsigned)g;
  const bool pers = grid > 0 && grid < total;
  switch (st) {
#define SOAP_BATCH_CASE(ID, K, JOB, TPB)                                                                 \
  case ID:                                                                                               \
    if (threads > TPB || threads < 64 || threads % 64) return hipErrorInvalidValue;                     \
    if (pers) K<true><<<nb, threads, lds, s>>>((const JOB *)jobs, prefix, n_inst, total);                \
    else K<false><<<nb, threads, lds, s>>>((const JOB *)jobs, prefix, n_inst, total);                    \
    break;
    SOAP_BATCH_CASE(BS_FILL, kb_fill, FillJob, kTPB)
    SOAP_BATCH_CASE(BS_TAPS, kb_taps, TapsJob, kTPB)
    SOAP_BATCH_CASE(BS_FINALIZE, kb_finalize, FinalizeJob, kTPB)
    SOAP_BATCH_CASE(BS_FIR_PROG, kb_fir_prog, FirProgJob, kProgThreads)
    SOAP_BATCH_CASE(BS_CONV_QUIET, kb_conv_quiet, ConvQJob, kTPB)
    SOAP_BATCH_CASE(BS_CONV, kb_conv, ConvAJob, kTPB)
    SOAP_BATCH_CASE(BS_MINMAX, kb_minmax, MinmaxJob, kTPB)
    SOAP_BATCH_CASE(BS_NORM, kb_norm, NormJob, kTPB)
    SOAP_BATCH_CASE(BS_CCF, kb_ccf, CcfJob, 1024)
    SOAP_BATCH_CASE(BS_CCF_REDUCE, kb_ccf_reduce, CcfRedJob, 1024)
    SOAP_BATCH_CASE(BS_FIT, kb_fit, FitJob, 64)
#undef SOAP_BATCH_CASE
    case BS_FIR:
      if (tile) {
        if (threads != 512) return hipErrorInvalidValue;
        if (pers) kb_fir_tile<true><<<nb, threads, lds, s>>>((const FirTileJob *)jobs, prefix, n_inst, total);
        else kb_fir_tile<false><<<nb, threads, lds, s>>>((const FirTileJob *)jobs, prefix, n_inst, total);
        break;
      }
      if (threads != kTPB) return hipErrorInvalidValue;
#define SOAP_FIR_CASE(G)                                                                                   \
  case G:                                                                                                  \
    if (fir_pre) {                                                                                         \
      if (pers) kb_fir<true, G, true><<<nb, threads, lds, s>>>((const FirJob *)jobs, prefix, n_inst, total); \
      else kb_fir<false, G, true><<<nb, threads, lds, s>>>((const FirJob *)jobs, prefix, n_inst, total);     \
    } else {                                                                                               \
      if (pers) kb_fir<true, G, false><<<nb, threads, lds, s>>>((const FirJob *)jobs, prefix, n_inst, total); \
      else kb_fir<false, G, false><<<nb, threads, lds, s>>>((const FirJob *)jobs, prefix, n_inst, total);     \
    }                                                                                                      \
    break;
#define SOAP_FIR2_CASE(GG, BB)                                                                            \
  if (fir_g == GG && fir_b == BB) {                                                                      \
    if (pers) kb_fir2<true, GG, BB><<<nb, threads, lds, s>>>((const FirJob *)jobs, prefix, n_inst, total); \
    else kb_fir2<false, GG, BB><<<nb, threads, lds, s>>>((const FirJob *)jobs, prefix, n_inst, total);     \
    break;                                                                                               \
  }
      if (fir_b > 0) {
        SOAP_FIR2_CASE(1, 3)
        SOAP_FIR2_CASE(1, 5)
        SOAP_FIR2_CASE(2, 3)
        return hipErrorInvalidValue;
      }
      switch (fir_g) {
        SOAP_FIR_CASE(2)
        SOAP_FIR_CASE(4)
        SOAP_FIR_CASE(6)
        default: return hipErrorInvalidValue;
      }
#undef SOAP_FIR_CASE
#undef SOAP_FIR2_CASE
      break;
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace soap
