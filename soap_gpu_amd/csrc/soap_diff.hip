// gfx950 kernels of the differentiable path (no counterpart in the reference).
//
// Model (log-uniform grid): disk cell c with bin k, limb L = 1 - u1(1-mu) -
// u2(1-mu)^2 and rigid-rotation velocity delta = vsini y has the integer shift m
// and weight t of the forward engine; phi_pc is the bilinear sample of phase p's
// M x M filling-factor map over [-1, 1]^2 (psi_pc likewise from the facula map,
// when there is one); the phase spectrum is
//   I_p = sum_k sum_d HQ_pk[d] Q_k[j - d] + HS_pk[d] S_k[j - d] + HF_pk[d] Fa_k[j - d],
//   HQ_pk = sum_c L (1 - phi - psi) {(1-t) at d = m, t at d = m-1},  HS_pk likewise with
//   kappa phi, HF_pk with kappa_f (1 + LB(mu)) psi (the reference's facula intensity
//   factor, SOAP_integration.cu:317-323),
// F_p = conv(I_p) with the resolution taps, CCF_p = raw mask CCF of F_p, and the
// Gaussian fit of CCF_p polished to the least-squares optimum gives RV_p, FWHM_p.
// The adjoint runs the same structures backwards: the implicit derivative of
// the fit (H^-1 J^T), the transposed mask and convolution, lag correlations of
// the adjoint spectrum with the seed rows (from which every per-cell derivative
// follows), and the transposed FIR for the seed gradients.  Shifted spectra are
// never stored: the backward recomputes from taps and correlations.
#include <hip/hip_runtime.h>
#include <math.h>

#include "soap_internal.h"

namespace soap {

constexpr int kDT = 256;

// ------------------------------------------------------------ per cell
__global__ void k_diff_cells(DiffArgs a) {
#pragma clang fp contract(off)
  long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  long ncell = (long)a.grid * a.grid;
  if (idx >= ncell) return;
  int iy = (int)(idx / a.grid), iz = (int)(idx % a.grid);
  double y = -1. + iy * (2. / a.grid), z = -1. + iz * (2. / a.grid);
  if (!(y * y + z * z <= 1.)) {
    a.ck[idx] = -1;
    return;
  }
  double mu = sqrt(1. - (y * y + z * z));
  int k = a.n_mu - 1;
  for (int b = 0; b < a.n_mu - 1; b++)
    if (mu > a.mu_bounds[b]) {
      k = b;
      break;
    }
  const double u1 = a.prm ? a.prm[1] : a.u1, u2 = a.prm ? a.prm[2] : a.u2, vsini = a.prm ? a.prm[0] : a.vsini;
  double L = 1. - u1 * (1 - mu) - u2 * (1 - mu) * (1 - mu);
  double delta = vsini * y;
  double beta = delta * 1000. / kC;
  double g = sqrt((1 + beta) / (1 - beta));
  double fbeta = -1 * (1 - g);
  double sh = log1p(fbeta) / a.h;
  double mc = ceil(sh);
  double t = expm1((mc - sh) * a.h) / a.expm1_h;
  a.ck[idx] = k;
  a.cm[idx] = (int32_t)mc;
  a.ct[idx] = t;
  a.cL[idx] = L;
  a.cmu[idx] = mu;
  // dt/d(delta) = -exp((m-s)h)/expm1(h) / (g^2 (1-beta)^2) * 1000/c (m piecewise constant)
  a.cdt[idx] = -exp((mc - sh) * a.h) / a.expm1_h / (g * g * (1 - beta) * (1 - beta)) * (1000. / kC) * y;
}

// bilinear sample of map p at the cell; i0, j0, fy, fz of the sample
__device__ __forceinline__ double map_sample(const double *mp, int M, double y, double z, int &i0, int &j0,
                                             double &fy, double &fz) {
#pragma clang fp contract(off)
  double gy = (y + 1) * 0.5 * (M - 1), gz = (z + 1) * 0.5 * (M - 1);
  i0 = (int)floor(gy);
  j0 = (int)floor(gz);
  i0 = i0 > M - 2 ? M - 2 : (i0 < 0 ? 0 : i0);
  j0 = j0 > M - 2 ? M - 2 : (j0 < 0 ? 0 : j0);
  fy = gy - i0;
  fz = gz - j0;
  return (1 - fy) * (1 - fz) * mp[i0 * M + j0] + fy * (1 - fz) * mp[(i0 + 1) * M + j0] +
         (1 - fy) * fz * mp[i0 * M + j0 + 1] + fy * fz * mp[(i0 + 1) * M + j0 + 1];
}

__device__ __forceinline__ void cell_yz(long idx, int grid, double &y, double &z) {
#pragma clang fp contract(off)
  int iy = (int)(idx / grid), iz = (int)(idx % grid);
  y = -1. + iy * (2. / grid);
  z = -1. + iz * (2. / grid);
}

// facula intensity factor 1 + LB(mu) (SOAP_initialize.cu:367, SOAP_integration.cu:317-323)
__device__ __forceinline__ double lb_factor(const DiffArgs &a, double mu) {
#pragma clang fp contract(off)
  return 1. + (a.lb[0] + a.lb[1] * mu + a.lb[2] * mu * mu + a.lb[3] * mu * mu * mu);
}

// taps of every phase (blockIdx.y): fixed-point sums in LDS when they fit, then global.
// All cells of one disk row share the shift (m, t), and neighbouring cells of a row
// share the mu bin, so each wave first sums its lanes' weights over runs of equal
// (row, bin) (a segmented shuffle scan, lanes = consecutive cells); the last lane of
// each run adds the run's four tap terms.  This removes the many-way atomic
// contention of one atomic per cell on the same few tap addresses.
template <int NS>
__global__ __launch_bounds__(kDT) void k_diff_taps(DiffArgs a, int cells_per_block, int use_lds) {
  const double kappa = a.prm ? a.prm[3] : a.kappa;
  const double kappa_f = a.prm ? a.prm[4] : a.kappa_f;
  extern __shared__ unsigned long long lacc[];  // [n_mu][nd][NS sets][2]
  int p = blockIdx.y;
  long ncell = (long)a.grid * a.grid;
  long c0 = (long)blockIdx.x * cells_per_block, c1 = min(ncell, c0 + cells_per_block);
  long nacc = (long)a.n_mu * a.nd * NS * 2;
  unsigned long long *gacc = a.acc + (long)p * nacc;
  if (use_lds) {
    for (long e = threadIdx.x; e < nacc; e += blockDim.x) lacc[e] = 0ull;
    __syncthreads();
  }
  unsigned long long *dst = use_lds ? lacc : gacc;
  const double *mp = a.maps + (long)p * a.map_res * a.map_res;
  const double *fp = NS == 3 ? a.fmaps + (long)p * a.map_res * a.map_res : nullptr;
  const int lane = threadIdx.x & 63;
  // every lane of a wave runs every iteration (the shuffles need the whole wave)
  for (long base = c0; base < c1; base += blockDim.x) {
    long idx = base + threadIdx.x;
    int key = -1, k = -1, d0 = 0;
    double w[NS], t = 0.0;
#pragma unroll
    for (int q = 0; q < NS; q++) w[q] = 0.0;
    if (idx < c1) k = a.ck[idx];
    if (k >= 0) {
      d0 = a.cm[idx] - a.dmin;  // tap index of offset m; m-1 at d0-1
      if (d0 >= 1 && d0 < a.nd) {  // the host bounds every shift
        double y, z, fy, fz;
        int i0, j0;
        cell_yz(idx, a.grid, y, z);
        double phi = map_sample(mp, a.map_res, y, z, i0, j0, fy, fz);
        double psi = NS == 3 ? map_sample(fp, a.map_res, y, z, i0, j0, fy, fz) : 0.0;
        double L = a.cL[idx];
        t = a.ct[idx];
        w[0] = L * (1 - phi - psi);
        w[1] = kappa * L * phi;
        if (NS == 3) w[NS - 1] = kappa_f * lb_factor(a, a.cmu[idx]) * L * psi;
        key = (int)(idx / a.grid) * 1024 + k;  // (row, bin): rows share (m, t); k < 1024
      }
    }
    const int kprev = __shfl_up(key, 1), knext = __shfl_down(key, 1);
    bool f = lane == 0 || kprev != key;  // run head
    const bool tail = lane == 63 || knext != key;
    for (int off = 1; off < 64; off <<= 1) {  // inclusive segmented scan of the weights
      double x[NS];
#pragma unroll
      for (int q = 0; q < NS; q++) x[q] = __shfl_up(w[q], off);
      int fo = __shfl_up((int)f, off);
      if (lane >= off && !f) {
#pragma unroll
        for (int q = 0; q < NS; q++) w[q] += x[q];
        f = fo != 0;
      }
    }
    if (tail && key >= 0) {
      unsigned long long *row = dst + (long)k * a.nd * NS * 2;
#pragma unroll
      for (int q = 0; q < NS; q++) {
        fx_add(row + ((long)d0 * NS + q) * 2, w[q] * (1.0 - t));
        fx_add(row + ((long)(d0 - 1) * NS + q) * 2, w[q] * t);
      }
    }
  }
  if (use_lds) {
    __syncthreads();
    for (long e = threadIdx.x; e < nacc; e += blockDim.x) {
      unsigned long long v = lacc[e];
      if (v) atomicAdd(gacc + e, v);
    }
  }
}

// fixed point -> taps [p][k][d][ns] and the nonzero range per (p, k)
__global__ void k_diff_taps_finalize(DiffArgs a) {
  int p = blockIdx.x;
  const int ns = a.ns;
  for (int k = threadIdx.x; k < a.n_mu; k += blockDim.x) {
    int lo = a.nd, hi = -1;
    for (int d = 0; d < a.nd; d++) {
      const unsigned long long *e = a.acc + (((long)p * a.n_mu + k) * a.nd + d) * ns * 2;
      double *o = a.taps + (((long)p * a.n_mu + k) * a.nd + d) * ns;
      bool nz = false;
      for (int q = 0; q < ns; q++) {
        double v = fx_value(e[q * 2], e[q * 2 + 1]);
        o[q] = v;
        nz |= v != 0.0;
      }
      if (nz) {
        lo = d < lo ? d : lo;
        hi = d;
      }
    }
    a.krange[((long)p * a.n_mu + k) * 2] = lo;
    a.krange[((long)p * a.n_mu + k) * 2 + 1] = hi;
  }
}

// I_p[j] = sum_k sum_d HQ[d] Q_k[j - dmin - d] + HS[d] S_k[j - dmin - d] (+ HF[d] Fa_k[j - dmin - d]);
// block = (tile, p)
template <int NS>
__global__ __launch_bounds__(kDT) void k_diff_fir(DiffArgs a) {
  extern __shared__ double rows[];  // [NS][row_len]
  int p = blockIdx.y;
  long j0 = (long)blockIdx.x * kDT;
  int nd = a.nd, row_len = kDT + nd - 1;
  long src0 = j0 - a.dmin - nd + 1;
  double acc = 0.0;
  for (int k = 0; k < a.n_mu; k++) {
    int lo = a.krange[((long)p * a.n_mu + k) * 2], hi = a.krange[((long)p * a.n_mu + k) * 2 + 1];
    if (lo > hi) continue;  // uniform
    __syncthreads();
    for (int e = threadIdx.x; e < row_len; e += blockDim.x) {
      long i = src0 + e;
      bool in = i >= 0 && i < a.n;
      rows[e] = in ? a.q[(long)k * a.n + i] : 0.0;
      rows[row_len + e] = in ? a.s[(long)k * a.n + i] : 0.0;
      if (NS == 3) rows[2 * row_len + e] = in ? a.f[(long)k * a.n + i] : 0.0;
    }
    __syncthreads();
    const double *T = a.taps + ((long)p * a.n_mu + k) * nd * NS;
    for (int di = lo; di <= hi; di++) {
      int e = threadIdx.x + nd - 1 - di;
#pragma unroll
      for (int q = 0; q < NS; q++) acc = fma(T[di * NS + q], rows[q * row_len + e], acc);
    }
  }
  long j = j0 + threadIdx.x;
  if (j < a.n) a.I[(long)p * a.n + j] = acc;
}

// F_p = conv(I_p) (zero outside wid < i < n-1-wid); residual partials of the
// trimmed range per block: part[p][block]
__global__ __launch_bounds__(kDT) void k_diff_conv(DiffArgs a, ConvArgs c, const double *target, int trim,
                                                   double *part) {
  extern __shared__ double xs[];  // I over [i0 + dlo, i0 + 4*kDT + ntp - 1), taps after
  int p = blockIdx.y;
  long i0 = (long)blockIdx.x * (4 * kDT);
  int span = 4 * kDT + c.ntp - 1;
  double *ct = xs + span;
  const double *X = a.I + (long)p * a.n;
  for (int e = threadIdx.x; e < c.ntp; e += blockDim.x) ct[e] = c.ctap[e];
  for (int e = threadIdx.x; e < span; e += blockDim.x) {
    long m = i0 + c.dlo + e;
    xs[e] = (m >= 0 && m < a.n) ? X[m] : 0.0;
  }
  __syncthreads();
  double acc[4] = {0, 0, 0, 0};
  for (int k = 0; k < c.ntp; k++) {
    double cv = ct[k];
#pragma unroll
    for (int b = 0; b < 4; b++) acc[b] = fma(cv, xs[threadIdx.x + b * kDT + k], acc[b]);
  }
  double r2 = 0.0;
#pragma unroll
  for (int b = 0; b < 4; b++) {
    long i = i0 + threadIdx.x + b * kDT;
    if (i >= a.n) continue;
    double o = ((i > c.wid) && (i + c.wid < a.n - 1)) ? acc[b] : 0.0;
    a.F[(long)p * a.n + i] = o;
    if (i >= trim && i < a.n - trim) {
      double r = o - target[(long)p * a.n + i];
      r2 += r * r;
    }
  }
  __shared__ double s_r[kDT / 64];
  for (int off = 32; off > 0; off >>= 1) r2 += __shfl_xor(r2, off);
  if ((threadIdx.x & 63) == 0) s_r[threadIdx.x >> 6] = r2;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0;
    for (int w = 0; w < kDT / 64; w++) s += s_r[w];
    part[(long)p * gridDim.x + blockIdx.x] = s;
  }
}

// ------------------------------------------------------------ Gaussian fit: polish and adjoint
// gradient (4) and Hessian (10, upper) of 0.5 sum r^2 (BIS_FIT2 model, exact second
// derivatives), summed over the wave's lanes; every lane receives the totals
__device__ void gauss_gh(const double *x, const double *y, int n, const double th[4], double g[4], double H[4][4],
                         double *r2out) {
  double s = th[3] / (2 * sqrt(2 * log(2.0)));
  double acc[15];
#pragma unroll
  for (int i = 0; i < 15; i++) acc[i] = 0.0;
  for (int i = threadIdx.x & 63; i < n; i += 64) {
    double u = (x[i] - th[2]) / s, e = exp(-u * u / 2);
    double r = th[0] + th[1] * e - y[i];
    double j4[4] = {1.0, e, th[1] * e * u / s, th[1] * e * u * u / th[3]};
    double h12 = e * u / s, h13 = e * u * u / th[3];
    double h22 = th[1] * e * (u * u - 1) / (s * s), h23 = th[1] * e * u * (u * u - 2) / (s * th[3]);
    double h33 = th[1] * e * u * u * (u * u - 3) / (th[3] * th[3]);
    int q = 0;
    for (int a = 0; a < 4; a++) acc[q++] += j4[a] * r;
    double hh[4][4] = {{0, 0, 0, 0}, {0, 0, h12, h13}, {0, h12, h22, h23}, {0, h13, h23, h33}};
    for (int a = 0; a < 4; a++)
      for (int b = a; b < 4; b++) acc[q++] += j4[a] * j4[b] + r * hh[a][b];
    acc[14] += r * r;
  }
#pragma unroll
  for (int i = 0; i < 15; i++)
    for (int off = 32; off > 0; off >>= 1) acc[i] += __shfl_xor(acc[i], off);
  int q = 0;
  for (int a = 0; a < 4; a++) g[a] = acc[q++];
  for (int a = 0; a < 4; a++)
    for (int b = a; b < 4; b++) H[a][b] = H[b][a] = acc[q++];
  *r2out = acc[14];
}

__device__ int solve4(const double A_in[4][4], const double *b, double *x) {
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

// one wave per CCF: theta from the fit parameters (signed amplitude), then
// damped Newton steps on the exact Hessian to the least-squares optimum
__global__ void k_fit_polish(const double *v, int n_v, const double *ccf, const double *par, double *theta) {
  int sp = blockIdx.x;
  const double *y = ccf + (long)sp * n_v;
  const double *pp = par + (long)sp * 5;
  int inear = 0;
  for (int i = 1; i < n_v; i++)
    if (fabs(v[i] - pp[3]) < fabs(v[inear] - pp[3])) inear = i;
  double th[4] = {pp[0], y[inear] < pp[0] ? -pp[1] : pp[1], pp[3], pp[4]};
  double g[4], H[4][4], r2;
  gauss_gh(v, y, n_v, th, g, H, &r2);
  for (int it = 0; it < 60; it++) {
    double mu = 0.0, dx[4] = {0, 0, 0, 0}, tn[4] = {th[0], th[1], th[2], th[3]}, r2n = r2;
    bool ok = false;
    for (int tries = 0; tries < 30; tries++) {
      double Hd[4][4], mg[4];
      for (int a = 0; a < 4; a++) {
        mg[a] = -g[a];
        for (int b = 0; b < 4; b++) Hd[a][b] = H[a][b] + (a == b ? mu * (fabs(H[a][a]) + 1e-300) : 0.0);
      }
      if (solve4(Hd, mg, dx) == 0) {
        for (int a = 0; a < 4; a++) tn[a] = th[a] + dx[a];
        double gn[4], Hn[4][4];
        gauss_gh(v, y, n_v, tn, gn, Hn, &r2n);
        if (r2n <= r2 * (1 + 1e-15)) {
          ok = true;
          break;
        }
      }
      mu = mu == 0.0 ? 1e-6 : mu * 10;
    }
    if (!ok) break;
    bool small = true;
    for (int a = 0; a < 4; a++)
      if (fabs(dx[a]) > 1e-12 * fabs(th[a]) + 1e-300) small = false;
    for (int a = 0; a < 4; a++) th[a] = tn[a];
    gauss_gh(v, y, n_v, th, g, H, &r2);
    if (small) break;
  }
  // two undamped Newton steps on the gradient fix the optimum to rounding
  for (int it = 0; it < 2; it++) {
    double mg[4], dx[4];
    gauss_gh(v, y, n_v, th, g, H, &r2);
    for (int a = 0; a < 4; a++) mg[a] = -g[a];
    if (solve4(H, mg, dx) != 0) break;
    for (int a = 0; a < 4; a++) th[a] += dx[a];
  }
  if (threadIdx.x == 0)
    for (int a = 0; a < 4; a++) theta[(long)sp * 4 + a] = th[a];
}

// dl/dy_v = J_v . z with H z = dl/dtheta (implicit derivative of the optimum), plus
// the bisector span's direct term coef_p dspan_p/dy_v when bisd holds one
__global__ void k_fit_adjoint(const double *v, int n_v, const double *ccf, const double *theta, const double *dth,
                              const double *bisd, double *gy) {
  int sp = blockIdx.x;
  const double *y = ccf + (long)sp * n_v;
  double th[4], g[4], H[4][4], r2, z[4] = {0, 0, 0, 0};
  for (int a = 0; a < 4; a++) th[a] = theta[(long)sp * 4 + a];
  gauss_gh(v, y, n_v, th, g, H, &r2);
  double b[4];
  for (int a = 0; a < 4; a++) b[a] = dth[(long)sp * 4 + a];
  solve4(H, b, z);
  double s = th[3] / (2 * sqrt(2 * log(2.0)));
  const double *bp = bisd ? bisd + (long)sp * (n_v + kBisHead) : nullptr;
  const double coef = bp ? bp[1] : 0.0;
  for (int i = threadIdx.x; i < n_v; i += blockDim.x) {
    double u = (v[i] - th[2]) / s, e = exp(-u * u / 2);
    double g = z[0] + z[1] * e + z[2] * th[1] * e * u / s + z[3] * th[1] * e * u * u / th[3];
    if (coef != 0.0) g += coef * bp[kBisHead + i];
    gy[(long)sp * n_v + i] = g;
  }
}

// ------------------------------------------------------------ bisector span
// p2 of the interval centred at vr (BIS_FIT2.cpp:233-236: half the inverse
// curvature of the Gaussian at vr) and its derivatives in the centre v0 and in
// s2 = sigma^2.  With u = vr - v0, E = exp(-u^2 / 2 s2): dccf = -u E / s2,
// d2ccf = (u^2 / s2 - 1) E / s2, p2 = -d2ccf / (2 dccf^3); d dccf / du = d2ccf,
// d d2ccf / du = E u (3 - u^2 / s2) / s2^2, d dccf / ds2 = u E (1 - u^2 / 2 s2) / s2^2,
// d d2ccf / ds2 = E [(s2 - 2 u^2) / s2^3 + (u^2 - s2) u^2 / (2 s2^4)].
__device__ void bis_p2(double vr, double v0, double s2, double *p2, double *dp2_dv0, double *dp2_ds2) {
  double dd = (vr - v0) * (vr - v0);
  double dccf = -(vr - v0) / s2 * exp(-dd / 2. / s2);
  double d2ccf = (dd / s2 - 1) / s2 * exp(-dd / 2. / s2);
  double d2rv = -d2ccf / (dccf * dccf * dccf);
  *p2 = d2rv / 2.;
  double u = vr - v0, E = exp(-dd / 2. / s2);
  double d2ccf_u = E * u * (3 - dd / s2) / (s2 * s2);
  double dccf_s = u * E / (s2 * s2) * (1 - dd / (2 * s2));
  double d2ccf_s = E * ((s2 - 2 * dd) / (s2 * s2 * s2) + (dd - s2) * dd / (2 * s2 * s2 * s2 * s2));
  double i3 = 1.0 / (dccf * dccf * dccf);
  double dp2_u = -d2ccf_u * i3 / 2 + 1.5 * d2ccf * d2ccf * i3 / dccf;
  *dp2_dv0 = -dp2_u;
  *dp2_ds2 = -d2ccf_s * i3 / 2 + 1.5 * d2ccf * dccf_s * i3 / dccf;
}

// One 64-lane block per CCF: the span of BIS_FIT2.cpp:214-268 at the polished
// Gaussian theta = (c, k, v0, fwhm) (the forward of the plain fit's bisector,
// with the same per-interval quadratic inverse and walk), and its partial
// derivatives at fixed theta: dspan/dy_v and dspan/dtheta.  Per interval s with
// a = norm_s, b = norm_{s+1}, the quadratic inverse at depth d is
// v_s + dv (d - a) / (b - a) + p2 (d - a) (d - b), so
//   dterm/da = (d - b) (dv / (b - a)^2 - p2), dterm/db = -(d - a) (dv / (b - a)^2 + p2),
//   dterm/dp2 = (d - a) (d - b);  norm_j = (y_j - c) / k.
// Output per CCF (stride n_v + kBisHead): [0] span, [1] loss coefficient (written
// by k_diff_loss), [2..5] dspan/dtheta, [kBisHead + v] dspan/dy_v.
constexpr int kBisLevels = 91;
__global__ __launch_bounds__(64) void k_diff_bis(const double *v, int n_v, const double *ccf, const double *theta,
                                                 double *bisd) {
  __shared__ double s_nc[kMaxVDiff], s_p0[kMaxVDiff], s_p1[kMaxVDiff], s_p2[kMaxVDiff];
  __shared__ double s_bis[kBisLevels], s_dA[kBisLevels][2], s_dB[kBisLevels][2], s_th[kBisLevels][4];
  __shared__ int s_side[kBisLevels][2];
  __shared__ int s_imax;
  const int sp = blockIdx.x;
  const double *y = ccf + (long)sp * n_v;
  double *out = bisd + (long)sp * (n_v + kBisHead);
  const double c = theta[sp * 4 + 0], k = theta[sp * 4 + 1], v0 = theta[sp * 4 + 2], fwhm = theta[sp * 4 + 3];
  const double sigma = fwhm / 2. / sqrt(2. * log(2.));
  const double s2 = sigma * sigma;
  for (int i = threadIdx.x; i < n_v; i += 64) s_nc[i] = -c / k * (1. - y[i] / c);
  __syncthreads();
  for (int i = threadIdx.x; i < n_v - 1; i += 64) {
    double p2, d0, d1;
    bis_p2((v[i] + v[i + 1]) / 2., v0, s2, &p2, &d0, &d1);
    double p1 = (v[i + 1] - v[i] - p2 * (s_nc[i + 1] * s_nc[i + 1] - s_nc[i] * s_nc[i])) / (s_nc[i + 1] - s_nc[i]);
    s_p2[i] = p2;
    s_p1[i] = p1;
    s_p0[i] = v[i] - p1 * s_nc[i] - p2 * (s_nc[i] * s_nc[i]);
  }
  if (threadIdx.x == 0) {
    int im = 0;
    for (int i = 0; i < n_v; i++)
      if (s_nc[i] > s_nc[im]) im = i;
    s_imax = im;
  }
  __syncthreads();
  const int ind_max = s_imax;
  for (int i = threadIdx.x; i < kBisLevels; i += 64) {
    double dep = (double)(i + 5) / 100;
    int i_b = ind_max, i_r = ind_max;
    while ((s_nc[i_b] > dep) && (i_b > 1)) i_b--;
    while ((s_nc[i_r + 1] > dep) && (i_r < (n_v - 2))) i_r++;
    double b = (s_p0[i_b] + s_p0[i_r]) + (s_p1[i_b] + s_p1[i_r]) * dep + (s_p2[i_b] + s_p2[i_r]) * (dep * dep);
    s_bis[i] = b / 2.;
    // weight of this level in the span, halved for the mean of the two sides
    double w = ((dep >= 0.1) && (dep <= 0.4)) ? 0.5 / 31 : ((dep >= 0.6) && (dep <= 0.9)) ? -0.5 / 31 : 0.0;
    double th4[4] = {0, 0, 0, 0};
    for (int side = 0; side < 2; side++) {
      int sx = side ? i_r : i_b;
      s_side[i][side] = sx;
      if (w == 0.0 || sx + 1 >= n_v) {
        s_dA[i][side] = 0.0;
        s_dB[i][side] = 0.0;
        continue;
      }
      double a = s_nc[sx], bb = s_nc[sx + 1], dv = v[sx + 1] - v[sx];
      double p2, dp2_v0, dp2_s2;
      bis_p2((v[sx] + v[sx + 1]) / 2., v0, s2, &p2, &dp2_v0, &dp2_s2);
      double q = dv / ((bb - a) * (bb - a));
      double dA = w * (dep - bb) * (q - p2), dB = -w * (dep - a) * (q + p2), dP = w * (dep - a) * (dep - bb);
      s_dA[i][side] = dA / k;
      s_dB[i][side] = dB / k;
      th4[0] += -(dA + dB) / k;
      th4[1] += -(dA * a + dB * bb) / k;
      th4[2] += dP * dp2_v0;
      th4[3] += dP * dp2_s2 * 2 * s2 / fwhm;
    }
    for (int q4 = 0; q4 < 4; q4++) s_th[i][q4] = th4[q4];
  }
  __syncthreads();
  for (int j = threadIdx.x; j < n_v; j += 64) {
    double g = 0.0;
    for (int i = 0; i < kBisLevels; i++)
      for (int side = 0; side < 2; side++) {
        int sx = s_side[i][side];
        if (sx == j) g += s_dA[i][side];
        if (sx + 1 == j) g += s_dB[i][side];
      }
    out[kBisHead + j] = g;
  }
  if (threadIdx.x == 0) {
    int n1 = 0, n2 = 0;
    double top = 0., bot = 0.;
    double th4[4] = {0, 0, 0, 0};
    for (int i = 0; i < kBisLevels; i++) {
      double dep = (double)(i + 5) / 100;
      if ((dep >= 0.1) && (dep <= 0.4)) {
        n1++;
        top += s_bis[i];
      }
      if ((dep >= 0.6) && (dep <= 0.9)) {
        n2++;
        bot += s_bis[i];
      }
      for (int q4 = 0; q4 < 4; q4++) th4[q4] += s_th[i][q4];
    }
    out[0] = top / n1 - bot / n2;
    out[1] = 0.0;
    for (int q4 = 0; q4 < 4; q4++) out[2 + q4] = th4[q4];
  }
}

hipError_t launch_diff_bis(const double *v, int n_v, const double *ccf, const double *theta, int n_ccf, double *bisd,
                           hipStream_t s) {
  if (n_ccf == 0) return hipSuccess;
  if (n_v < 3 || n_v > kMaxVDiff) return hipErrorInvalidValue;
  k_diff_bis<<<n_ccf, 64, 0, s>>>(v, n_v, ccf, theta, bisd);
  return hipGetLastError();
}

// ------------------------------------------------------------ backward spectra
// G_F = 2 w_spec (F - T) on [trim, n - trim), 0 elsewhere
__global__ void k_diff_gradF(DiffArgs a, const double *target, int trim, double w_spec, double *GF) {
  long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long)a.n_phase * a.n) return;
  long i = e % a.n;
  GF[e] = (i >= trim && i < a.n - trim) ? 2 * w_spec * (a.F[e] - target[e]) : 0.0;
}

// transposed mask as a gather: G_F[p][trim + t] += sum_e gy[p][col_e] val_e over the
// entries e of row t of the pixel-major table (the (velocity, weight) pairs of
// every live mask pixel of the CCF, merged per pixel).  The table is sliced by 64
// rows (one wave): entry k of row t sits at slice_ptr[t / 64] + 64 k + t % 64,
// rows padded with zero weights to their slice's width, so every load is
// coalesced.  Block = (kDT pixels, chunk of kMtP phases), the chunk's gy rows in
// LDS, one owner per output: no atomics.
constexpr int kMtP = 8;
__global__ __launch_bounds__(kDT) void k_mask_transpose(int64_t n_t, int trim, int64_t n, int n_phase, int n_v,
                                                        const int64_t *slice_ptr, const int32_t *slice_w,
                                                        const int32_t *col, const double *val, const double *gy,
                                                        double *GF) {
  extern __shared__ double sg[];  // [kMtP][n_v]
  const int p0 = blockIdx.y * kMtP;
  const int np = min(kMtP, n_phase - p0);
  for (int e = threadIdx.x; e < kMtP * n_v; e += blockDim.x) {
    int c = e / n_v;
    sg[e] = c < np ? gy[(long)(p0 + c) * n_v + e % n_v] : 0.0;
  }
  __syncthreads();
  const long t = (long)blockIdx.x * kDT + threadIdx.x;
  const long sl = t >> 6;
  if (sl * 64 >= n_t) return;  // whole wave past the end (n_t rounds up to slices)
  const int W = slice_w[sl];
  const int64_t base = slice_ptr[sl] + (t & 63);
  double acc[kMtP];
#pragma unroll
  for (int c = 0; c < kMtP; c++) acc[c] = 0.0;
  for (int k = 0; k < W; k++) {
    const int v = col[base + 64L * k];
    const double w = val[base + 64L * k];
#pragma unroll
    for (int c = 0; c < kMtP; c++) acc[c] = fma(sg[c * n_v + v], w, acc[c]);
  }
  if (t >= n_t) return;
#pragma unroll
  for (int c = 0; c < kMtP; c++)
    if (c < np) GF[(long)(p0 + c) * n + trim + t] += acc[c];
}

// G_I[m] = sum_k c_k G_F[m - dlo - k] over the valid outputs (G_F is zero elsewhere)
__global__ __launch_bounds__(kDT) void k_diff_convT(DiffArgs a, ConvArgs c, const double *GF, double *GI) {
  extern __shared__ double xs[];  // G_F over [m0 - dlo - ntp + 1, m0 + 4 kDT - dlo)
  int p = blockIdx.y;
  long m0 = (long)blockIdx.x * (4 * kDT);
  int span = 4 * kDT + c.ntp - 1;
  double *ct = xs + span;
  const double *X = GF + (long)p * a.n;
  long base = m0 - c.dlo - c.ntp + 1;
  for (int e = threadIdx.x; e < c.ntp; e += blockDim.x) ct[e] = c.ctap[e];
  for (int e = threadIdx.x; e < span; e += blockDim.x) {
    long i = base + e;
    xs[e] = (i >= 0 && i < a.n && i > c.wid && i + c.wid < a.n - 1) ? X[i] : 0.0;
  }
  __syncthreads();
  double acc[4] = {0, 0, 0, 0};
  for (int k = 0; k < c.ntp; k++) {
    double cv = ct[k];
    // output m = m0 + tid + b kDT reads G_F[m - dlo - k] at xs[tid + b kDT + ntp - 1 - k]
#pragma unroll
    for (int b = 0; b < 4; b++) acc[b] = fma(cv, xs[threadIdx.x + b * kDT + c.ntp - 1 - k], acc[b]);
  }
#pragma unroll
  for (int b = 0; b < 4; b++) {
    long m = m0 + threadIdx.x + b * kDT;
    if (m < a.n) GI[(long)p * a.n + m] = acc[b];
  }
}

// lag correlations C[p][k][d][set] = sum_j G_I[p][j] X_k[j - dmin - d], in two
// passes.  Pass 1: block = (chunk of kCorrT pixels, bin k, group of kCorrPG
// phases); the bin's two seed-row segments are staged once for the group.  Each
// thread owns kCorrPer consecutive pixels (G_I of the phase in registers) and,
// per group of kCorrL lags, reads one register window of the staged row and
// forms kCorrL partial sums; the block adds them up through LDS ([lag][thread]
// then 32-lane shuffles) into the chunk's partial.  Pass 2 sums the chunk
// partials in chunk order (deterministic, no atomics).
constexpr int kCorrPer = 9;  // odd: a thread stride of kCorrPer doubles is free of LDS bank conflicts
constexpr int kCorrT = kDT * kCorrPer;
constexpr int kCorrL = 8;
constexpr int kCorrPG = 10;
// Lags the cells can use: a cell's shift index d0 (k_diff_cells: m = ceil(log1p(fb)/h),
// fb monotone in delta = vsini y, |y| <= 1) lies in [m(-vsini), m(vsini)] - dmin, and the
// per-cell derivatives read lags d0 and d0 - 1; two lags of margin on each side.
// Lags outside need no correlation (the tap range fixed by a bound on vsini can be
// much wider than the current vsini needs).
__device__ __forceinline__ void diff_lag_range(const DiffArgs &a, int &lo, int &hi) {
  const double vs = a.prm ? a.prm[0] : a.vsini;
  auto m = [&](double v) {
    double beta = v * 1000. / kC;
    double fb = -1 * (1 - sqrt((1 + beta) / (1 - beta)));
    return (int)ceil(log1p(fb) / a.h);
  };
  lo = max(0, m(-vs) - a.dmin - 3);
  hi = min(a.nd - 1, m(vs) - a.dmin + 2);
}

// One pass covers nsp <= 2 seed cubes (x0, x1), written to sets soff, soff + 1 of
// the ns-set tables; the facula set, when present, is a second pass (three staged
// rows would not fit the LDS budget).
__global__ __launch_bounds__(kDT) void k_diff_corr(DiffArgs a, const double *GI, double *Cpart, const double *x0,
                                                   const double *x1, int nsp, int soff) {
  extern __shared__ double lds[];  // row(s) [kCorrT + nd - 1 + kCorrL] each, then [kCorrL][kDT] partials
  const int k = blockIdx.y, chunk = blockIdx.x, nchunk = gridDim.x;
  const long j0 = (long)chunk * kCorrT;
  const int nd = a.nd, rl = kCorrT + nd - 1 + kCorrL, ns = a.ns;
  double *rq = lds, *rs = rq + rl, *red = rq + nsp * rl;
  const long src0 = j0 - a.dmin - nd + 1;
  for (int e = threadIdx.x; e < rl; e += blockDim.x) {
    long i = src0 + e;
    bool in = i >= 0 && i < a.n && e < kCorrT + nd - 1;
    rq[e] = in ? x0[(long)k * a.n + i] : 0.0;
    if (nsp == 2) rs[e] = in ? x1[(long)k * a.n + i] : 0.0;
  }
  __syncthreads();
  const int p1 = min(a.n_phase, (int)(blockIdx.z + 1) * kCorrPG);
  for (int p = blockIdx.z * kCorrPG; p < p1; p++) {
    double g[kCorrPer];
#pragma unroll
    for (int u = 0; u < kCorrPer; u++) {
      long j = j0 + threadIdx.x * kCorrPer + u;
      g[u] = j < a.n ? GI[(long)p * a.n + j] : 0.0;
    }
    double *out = Cpart + (((long)p * a.n_mu + k) * nchunk + chunk) * nd * ns;
    int dlo, dhi;
    diff_lag_range(a, dlo, dhi);  // uniform
    for (int set = 0; set < nsp; set++) {
      const double *row = (set ? rs : rq) + threadIdx.x * kCorrPer;
      for (int d0 = dlo; d0 <= dhi; d0 += kCorrL) {
        // lag d = d0 + l reads X[j - dmin - d] = row[u + nd - 1 - d]; window base nd - d0 - kCorrL
        const int wb = nd - d0 - kCorrL;
        double w[kCorrPer + kCorrL - 1];
#pragma unroll
        for (int e = 0; e < kCorrPer + kCorrL - 1; e++) {
          int x = wb + e;
          w[e] = x >= 0 ? row[x] : 0.0;  // x < 0 only for lags d >= nd (discarded)
        }
        double acc[kCorrL];
#pragma unroll
        for (int l = 0; l < kCorrL; l++) {
          acc[l] = 0.0;
#pragma unroll
          for (int u = 0; u < kCorrPer; u++) acc[l] = fma(g[u], w[u + kCorrL - 1 - l], acc[l]);
        }
        __syncthreads();  // red is free
#pragma unroll
        for (int l = 0; l < kCorrL; l++) red[l * kDT + threadIdx.x] = acc[l];
        __syncthreads();
        // kDT threads = kCorrL lags x 32 lanes: each lane sums kDT / 32 partials
        const int l = threadIdx.x >> 5, part = threadIdx.x & 31;
        double x = 0.0;
#pragma unroll
        for (int r = 0; r < kDT / 32; r++) x += red[l * kDT + r * 32 + part];
        for (int off = 16; off > 0; off >>= 1) x += __shfl_xor(x, off);
        if (part == 0 && d0 + l < nd) out[(d0 + l) * ns + soff + set] = x;
      }
    }
  }
}

__global__ void k_diff_corr_reduce(DiffArgs a, int n_pk, int nd, int nchunk, const double *Cpart, double *C) {
  const int ns = a.ns;
  long e = (long)blockIdx.x * blockDim.x + threadIdx.x;  // (p, k) x (d, set)
  if (e >= (long)n_pk * nd * ns) return;
  long pk = e / (nd * ns), r = e % (nd * ns);
  int dlo, dhi;
  diff_lag_range(a, dlo, dhi);
  const int d = (int)(r / ns);
  double x = 0.0;
  if (d >= dlo && d <= dhi)  // k_diff_corr computes only these lags
    for (int c = 0; c < nchunk; c++) x += Cpart[(pk * nchunk + c) * nd * ns + r];
  C[e] = x;
}

int diff_corr_chunks(int64_t n) { return (int)((n + kCorrT - 1) / kCorrT); }

// per-cell derivatives from the correlations, phases summed per cell; scalar
// partials per block: [gvsini, gu1, gu2, gkappa, gkappa_f]
__global__ __launch_bounds__(kDT) void k_diff_cellgrad(DiffArgs a, const double *C, double *part) {
  const double kappa = a.prm ? a.prm[3] : a.kappa;
  const double kappa_f = a.prm ? a.prm[4] : a.kappa_f;
  const int ns = a.ns;
  long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  long ncell = (long)a.grid * a.grid;
  double sv = 0, su1 = 0, su2 = 0, sk = 0, sf = 0;
  if (idx < ncell && a.ck[idx] >= 0) {
    int k = a.ck[idx];
    int d0 = a.cm[idx] - a.dmin;
    if (d0 >= 1 && d0 < a.nd) {
      double y, z;
      cell_yz(idx, a.grid, y, z);
      double L = a.cL[idx], t = a.ct[idx], mu = a.cmu[idx];
      const double lbf = ns == 3 ? lb_factor(a, mu) : 0.0;
      double gL = 0, gt = 0;
      int M = a.map_res;
      for (int p = 0; p < a.n_phase; p++) {
        const double *Cp = C + (((long)p * a.n_mu + k) * a.nd) * ns;
        double cq0 = Cp[d0 * ns], cq1 = Cp[(d0 - 1) * ns], cs0 = Cp[d0 * ns + 1], cs1 = Cp[(d0 - 1) * ns + 1];
        double aq = (1 - t) * cq0 + t * cq1, as = (1 - t) * cs0 + t * cs1;
        double bq = cq1 - cq0, bs = cs1 - cs0;
        int i0, j0;
        double fy, fz;
        double phi = map_sample(a.maps + (long)p * M * M, M, y, z, i0, j0, fy, fz);
        if (ns == 3) {
          double cf0 = Cp[d0 * ns + 2], cf1 = Cp[(d0 - 1) * ns + 2];
          double af = (1 - t) * cf0 + t * cf1, bf = cf1 - cf0;
          double psi = map_sample(a.fmaps + (long)p * M * M, M, y, z, i0, j0, fy, fz);
          double ef = kappa_f * lbf * psi;
          gL += (1 - phi - psi) * aq + kappa * phi * as + ef * af;
          gt += L * ((1 - phi - psi) * bq + kappa * phi * bs + ef * bf);
          sf += L * lbf * psi * af;
        } else {
          gL += (1 - phi) * aq + kappa * phi * as;
          gt += L * ((1 - phi) * bq + kappa * phi * bs);
        }
        sk += L * phi * as;
      }
      su1 = -gL * (1 - mu);
      su2 = -gL * (1 - mu) * (1 - mu);
      sv = gt * a.cdt[idx];  // dt/d(delta) * d(delta)/d(vsini)
    }
  }
  __shared__ double s_p[kDiffScalars][kDT / 64];
  double v5[kDiffScalars] = {sv, su1, su2, sk, sf};
  for (int q = 0; q < kDiffScalars; q++) {
    double x = v5[q];
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off);
    if ((threadIdx.x & 63) == 0) s_p[q][threadIdx.x >> 6] = x;
  }
  __syncthreads();
  if (threadIdx.x < kDiffScalars) {
    double s = 0;
    for (int w = 0; w < kDT / 64; w++) s += s_p[threadIdx.x][w];
    part[(long)blockIdx.x * kDiffScalars + threadIdx.x] = s;
  }
}

// map-pixel gradients as a gather (no atomics): node (i, j) of phase p sums, over
// the cells whose bilinear stencil in map_sample holds it, weight x L (kappa a_S -
// a_Q) for the spot map and weight x L (kappa_f (1 + LB) a_F - a_Q) for the facula
// map, with a_Q, a_S, a_F the cell's interpolated correlations.  A cell at map
// coordinate gy = iy (M - 1) / G touches nodes floor(gy) and floor(gy) + 1, so
// node i looks at the rows with floor(gy) in {i - 1, i} (a superset is scanned
// and filtered with the same arithmetic as map_sample).
__global__ __launch_bounds__(kDT) void k_diff_mapgrad(DiffArgs a, const double *C, double *gmaps, double *gfmaps) {
#pragma clang fp contract(off)
  const double kappa = a.prm ? a.prm[3] : a.kappa;
  const double kappa_f = a.prm ? a.prm[4] : a.kappa_f;
  const int ns = a.ns;
  const int M = a.map_res, G = a.grid;
  const int j = blockIdx.x * blockDim.x + threadIdx.x;  // map column
  const int i = blockIdx.y;                              // map row
  const int p = blockIdx.z;
  if (j >= M) return;
  const double sc = (double)G / (M - 1), dg = 2. / G;
  const int y_lo = max(0, (int)floor((i - 1) * sc) - 1), y_hi = min(G - 1, (int)ceil((i + 1) * sc) + 1);
  const int z_lo = max(0, (int)floor((j - 1) * sc) - 1), z_hi = min(G - 1, (int)ceil((j + 1) * sc) + 1);
  double acc = 0.0, accf = 0.0;
  for (int iy = y_lo; iy <= y_hi; iy++) {
    // the row half of map_sample: gy, its clamped floor and fraction
    double y = -1. + iy * dg;
    double gy = (y + 1) * 0.5 * (M - 1);
    int i0 = (int)floor(gy);
    i0 = i0 > M - 2 ? M - 2 : (i0 < 0 ? 0 : i0);
    if (i0 != i && i0 + 1 != i) continue;
    double fy = gy - i0;
    double wy = i0 == i ? 1 - fy : fy;
    for (int iz = z_lo; iz <= z_hi; iz++) {
      double z = -1. + iz * dg;
      double gz = (z + 1) * 0.5 * (M - 1);
      int j0 = (int)floor(gz);
      j0 = j0 > M - 2 ? M - 2 : (j0 < 0 ? 0 : j0);
      if (j0 != j && j0 + 1 != j) continue;
      long idx = (long)iy * G + iz;
      int k = a.ck[idx];
      if (k < 0) continue;
      int d0 = a.cm[idx] - a.dmin;
      if (d0 < 1 || d0 >= a.nd) continue;
      double fz = gz - j0;
      double wz = j0 == j ? 1 - fz : fz;
      const double *Cp = C + (((long)p * a.n_mu + k) * a.nd) * ns;
      double t = a.ct[idx];
      double aq = (1 - t) * Cp[d0 * ns] + t * Cp[(d0 - 1) * ns];
      double as = (1 - t) * Cp[d0 * ns + 1] + t * Cp[(d0 - 1) * ns + 1];
      acc += wy * wz * (a.cL[idx] * (kappa * as - aq));
      if (ns == 3) {
        double af = (1 - t) * Cp[d0 * ns + 2] + t * Cp[(d0 - 1) * ns + 2];
        accf += wy * wz * (a.cL[idx] * (kappa_f * lb_factor(a, a.cmu[idx]) * af - aq));
      }
    }
  }
  gmaps[((long)p * M + i) * M + j] = acc;
  if (ns == 3) gfmaps[((long)p * M + i) * M + j] = accf;
}

// seed gradients gX_k[i] = sum_p sum_d HX_pk[d] G_I,p[i + dmin + d]; block = (tile, bin)
template <int NS>
__global__ __launch_bounds__(kDT) void k_diff_seedgrad(DiffArgs a, const double *GI, double *gq, double *gs,
                                                       double *gf) {
  extern __shared__ double gl[];  // G_I,p over [i0 + dmin, i0 + kDT + dmin + nd - 1)
  int k = blockIdx.y;
  long i0 = (long)blockIdx.x * kDT;
  int nd = a.nd, len = kDT + nd - 1;
  double ax[NS];
#pragma unroll
  for (int q = 0; q < NS; q++) ax[q] = 0.0;
  for (int p = 0; p < a.n_phase; p++) {
    int lo = a.krange[((long)p * a.n_mu + k) * 2], hi = a.krange[((long)p * a.n_mu + k) * 2 + 1];
    if (lo > hi) continue;  // uniform
    __syncthreads();
    for (int e = threadIdx.x; e < len; e += blockDim.x) {
      long j = i0 + a.dmin + e;
      gl[e] = (j >= 0 && j < a.n) ? GI[(long)p * a.n + j] : 0.0;
    }
    __syncthreads();
    const double *T = a.taps + ((long)p * a.n_mu + k) * nd * NS;
    for (int d = lo; d <= hi; d++) {
      double g = gl[threadIdx.x + d];
#pragma unroll
      for (int q = 0; q < NS; q++) ax[q] = fma(T[d * NS + q], g, ax[q]);
    }
  }
  long i = i0 + threadIdx.x;
  if (i < a.n) {
    gq[(long)k * a.n + i] = ax[0];
    gs[(long)k * a.n + i] = ax[1];
    if (NS == 3) gf[(long)k * a.n + i] = ax[NS - 1];
  }
}

// ------------------------------------------------------------ launchers
hipError_t launch_diff_forward_spectra(const DiffArgs &a, const ConvArgs &c, const double *target, int trim,
                                       double *res_part, int *res_blocks, hipStream_t s) {
  long ncell = (long)a.grid * a.grid;
  k_diff_cells<<<(unsigned)((ncell + kDT - 1) / kDT), kDT, 0, s>>>(a);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  if (a.ns != 2 && a.ns != 3) return hipErrorInvalidValue;
  if (a.ns == 3 && (!a.f || !a.fmaps)) return hipErrorInvalidValue;
  long nacc = (long)a.n_mu * a.nd * a.ns * 2;
  e = launch_fill_u64(a.acc, (size_t)a.n_phase * nacc, s);
  if (e != hipSuccess) return e;
  int use_lds = nacc * 8 <= 49152;
  int cpb = 4096;
  dim3 gt((unsigned)((ncell + cpb - 1) / cpb), a.n_phase);
  if (a.ns == 3)
    k_diff_taps<3><<<gt, kDT, use_lds ? nacc * 8 : 0, s>>>(a, cpb, use_lds);
  else
    k_diff_taps<2><<<gt, kDT, use_lds ? nacc * 8 : 0, s>>>(a, cpb, use_lds);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  k_diff_taps_finalize<<<a.n_phase, 64, 0, s>>>(a);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  size_t lds_fir = (size_t)a.ns * (kDT + a.nd - 1) * sizeof(double);
  if (lds_fir > 65536) return hipErrorInvalidValue;
  dim3 gf((unsigned)((a.n + kDT - 1) / kDT), a.n_phase);
  if (a.ns == 3)
    k_diff_fir<3><<<gf, kDT, lds_fir, s>>>(a);
  else
    k_diff_fir<2><<<gf, kDT, lds_fir, s>>>(a);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  size_t lds_conv = (size_t)(4 * kDT + 2 * c.ntp) * sizeof(double);
  if (lds_conv > 65536) return hipErrorInvalidValue;
  int nb = (int)((a.n + 4 * kDT - 1) / (4 * kDT));
  *res_blocks = nb;
  dim3 gc((unsigned)nb, a.n_phase);
  k_diff_conv<<<gc, kDT, lds_conv, s>>>(a, c, target, trim, res_part);
  return hipGetLastError();
}

int diff_conv_blocks(int64_t n) { return (int)((n + 4 * kDT - 1) / (4 * kDT)); }

// loss = sum_p [w_spec r2_p + w_rv (RV_p - rv_t)^2 + w_fw (FWHM_p - fw_t)^2], r2_p the
// residual partials of phase p summed in order; the same order as the host sum
// With device parameters (graph replays) the loss is NaN when vsini leaves the
// bound tap range or when the cell weights could exceed the fixed-point tap
// capacity: n_cells max|L| max(3, |kappa|, |kappa_f| lbmax) >= 2^30, lbmax =
// max |1 + LB(mu)| (maps are filling factors in [0, 1]).
__global__ __launch_bounds__(kDT) void k_diff_loss(int n_phase, int res_blocks, const double *res_part,
                                                   const double *theta, const double *rv_t, const double *fw_t,
                                                   const double *bis_t, double w_spec, double w_rv, double w_fw,
                                                   double w_bis, double *bisd, int n_v, const double *prm,
                                                   double vsini_max, double n_cells, double lbmax, double *loss,
                                                   double *dth) {
  extern __shared__ double term[];  // [n_phase]
  for (int p = threadIdx.x; p < n_phase; p += blockDim.x) {
    double r2 = 0;
    for (int b = 0; b < res_blocks; b++) r2 += res_part[(long)p * res_blocks + b];
    double drv = theta[p * 4 + 2] - rv_t[p], dfw = theta[p * 4 + 3] - fw_t[p];
    double t = w_spec * r2 + w_rv * drv * drv + w_fw * dfw * dfw;
    double d4[4] = {0.0, 0.0, 2 * w_rv * drv, 2 * w_fw * dfw};
    if (w_bis != 0.0) {  // span term; its coefficient feeds k_fit_adjoint's direct term
      double *bp = bisd + (long)p * (n_v + kBisHead);
      double db = bp[0] - bis_t[p];
      t += w_bis * db * db;
      double coef = 2 * w_bis * db;
      for (int q = 0; q < 4; q++) d4[q] += coef * bp[2 + q];
      bp[1] = coef;
    }
    term[p] = t;
    for (int q = 0; q < 4; q++) dth[p * 4 + q] = d4[q];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double l = 0;
    for (int p = 0; p < n_phase; p++) l += term[p];
    double vs = prm ? prm[0] : 0.0;
    if (prm && !(vs > 0 && vs <= vsini_max)) l = __longlong_as_double(0x7ff8000000000000LL);  // outside the tap range
    if (prm) {
      // max |L| over mu in [0, 1]: the ends and the vertex of the quadratic in (1 - mu)
      const double u1 = prm[1], u2 = prm[2];
      double lmax = fmax(1.0, fabs(1 - u1 - u2));
      if (u2 != 0.0) {
        double x = -u1 / (2 * u2);
        if (x > 0 && x < 1) lmax = fmax(lmax, fabs(1 - u1 * x - u2 * x * x));
      }
      double w = lmax * fmax(3.0, fmax(fabs(prm[3]), fabs(prm[4]) * lbmax));
      if (!(n_cells * w < 1073741824.0)) l = __longlong_as_double(0x7ff8000000000000LL);
    }
    loss[0] = l;
  }
}

// g[q] = sum over blocks of cell_part[b][q]: thread t sums blocks t, t + kDT, ...
// (independent loads), then a fixed-order tree over the threads
__global__ __launch_bounds__(kDT) void k_diff_grad(int n_blocks, const double *cell_part, double *g) {
  __shared__ double red[kDiffScalars][kDT];
  double sq[kDiffScalars];
#pragma unroll
  for (int q = 0; q < kDiffScalars; q++) sq[q] = 0.0;
  for (int b = threadIdx.x; b < n_blocks; b += kDT)
#pragma unroll
    for (int q = 0; q < kDiffScalars; q++) sq[q] += cell_part[(long)b * kDiffScalars + q];
#pragma unroll
  for (int q = 0; q < kDiffScalars; q++) red[q][threadIdx.x] = sq[q];
  __syncthreads();
  for (int off = kDT / 2; off > 0; off >>= 1) {
    if (threadIdx.x < off)
#pragma unroll
      for (int q = 0; q < kDiffScalars; q++) red[q][threadIdx.x] += red[q][threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x < kDiffScalars) g[threadIdx.x] = red[threadIdx.x][0];
}

hipError_t launch_diff_loss(int n_phase, int res_blocks, const double *res_part, const double *theta,
                            const double *rv_t, const double *fw_t, const double *bis_t, double w_spec, double w_rv,
                            double w_fw, double w_bis, double *bisd, int n_v, const double *prm, double vsini_max,
                            double n_cells, double lbmax, double *loss, double *dth, hipStream_t s) {
  if (n_phase < 1 || (size_t)n_phase * 8 > 65536) return hipErrorInvalidValue;
  if (w_bis != 0.0 && (!bisd || !bis_t)) return hipErrorInvalidValue;
  k_diff_loss<<<1, kDT, (size_t)n_phase * 8, s>>>(n_phase, res_blocks, res_part, theta, rv_t, fw_t, bis_t, w_spec,
                                                  w_rv, w_fw, w_bis, bisd, n_v, prm, vsini_max, n_cells, lbmax, loss,
                                                  dth);
  return hipGetLastError();
}

hipError_t launch_diff_grad(int n_blocks, const double *cell_part, double *g, hipStream_t s) {
  k_diff_grad<<<1, kDT, 0, s>>>(n_blocks, cell_part, g);
  return hipGetLastError();
}
int diff_cell_blocks(int64_t ncell) { return (int)((ncell + kDT - 1) / kDT); }

hipError_t launch_fit_polish(const double *v, int n_v, const double *ccf, const double *par, int n_ccf, double *theta,
                             hipStream_t s) {
  if (n_ccf == 0) return hipSuccess;
  k_fit_polish<<<n_ccf, 64, 0, s>>>(v, n_v, ccf, par, theta);
  return hipGetLastError();
}

hipError_t launch_diff_backward(const DiffArgs &a, const ConvArgs &c, const double *target, int trim, double w_spec,
                                const double *v, int n_v, const double *ccf, const double *theta, const double *dth,
                                double *gy, int64_t n_t, const int64_t *mt_ptr, const int32_t *mt_w,
                                const int32_t *mt_col, const double *mt_val, double *GF, double *GI, double *Cpart,
                                double *C, double *gmaps, double *gfmaps, double *cell_part, double *gq, double *gs,
                                double *gf, const double *bisd, hipStream_t s) {
  hipError_t e;
  if (a.ns == 3 && (!gf || !gfmaps)) return hipErrorInvalidValue;
  k_fit_adjoint<<<a.n_phase, 64, 0, s>>>(v, n_v, ccf, theta, dth, bisd, gy);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  long tot = (long)a.n_phase * a.n;
  k_diff_gradF<<<(unsigned)((tot + kDT - 1) / kDT), kDT, 0, s>>>(a, target, trim, w_spec, GF);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  if (n_t > 0) {
    size_t lds_mt = (size_t)kMtP * n_v * sizeof(double);
    if (lds_mt > 65536) return hipErrorInvalidValue;
    dim3 gm((unsigned)((n_t + kDT - 1) / kDT), (a.n_phase + kMtP - 1) / kMtP);
    k_mask_transpose<<<gm, kDT, lds_mt, s>>>(n_t, trim, a.n, a.n_phase, n_v, mt_ptr, mt_w, mt_col, mt_val, gy, GF);
    if ((e = hipGetLastError()) != hipSuccess) return e;
  }
  size_t lds_conv = (size_t)(4 * kDT + 2 * c.ntp) * sizeof(double);
  if (lds_conv > 65536) return hipErrorInvalidValue;
  dim3 gc((unsigned)((a.n + 4 * kDT - 1) / (4 * kDT)), a.n_phase);
  k_diff_convT<<<gc, kDT, lds_conv, s>>>(a, c, GF, GI);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  size_t lds_corr = (size_t)(2 * (kCorrT + a.nd - 1 + kCorrL) + kCorrL * kDT) * sizeof(double);
  if (lds_corr > 65536) return hipErrorInvalidValue;
  const int nchunk = diff_corr_chunks(a.n);
  dim3 gr((unsigned)nchunk, a.n_mu, (a.n_phase + kCorrPG - 1) / kCorrPG);
  k_diff_corr<<<gr, kDT, lds_corr, s>>>(a, GI, Cpart, a.q, a.s, 2, 0);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  if (a.ns == 3) {
    k_diff_corr<<<gr, kDT, lds_corr, s>>>(a, GI, Cpart, a.f, nullptr, 1, 2);
    if ((e = hipGetLastError()) != hipSuccess) return e;
  }
  long nc = (long)a.n_phase * a.n_mu * a.nd * a.ns;
  k_diff_corr_reduce<<<(unsigned)((nc + kDT - 1) / kDT), kDT, 0, s>>>(a, a.n_phase * a.n_mu, a.nd, nchunk, Cpart, C);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  long ncell = (long)a.grid * a.grid;
  k_diff_cellgrad<<<(unsigned)((ncell + kDT - 1) / kDT), kDT, 0, s>>>(a, C, cell_part);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  dim3 gmg((unsigned)((a.map_res + kDT - 1) / kDT), a.map_res, a.n_phase);
  k_diff_mapgrad<<<gmg, kDT, 0, s>>>(a, C, gmaps, gfmaps);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  size_t lds_seed = (size_t)(kDT + a.nd - 1) * sizeof(double);
  dim3 gsd((unsigned)((a.n + kDT - 1) / kDT), a.n_mu);
  if (a.ns == 3)
    k_diff_seedgrad<3><<<gsd, kDT, lds_seed, s>>>(a, GI, gq, gs, gf);
  else
    k_diff_seedgrad<2><<<gsd, kDT, lds_seed, s>>>(a, GI, gq, gs, gf);
  return hipGetLastError();
}

}  // namespace soap
