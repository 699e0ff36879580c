// Bisector injection into seed spectra (the preprocessing of
// SOAP_code/CB_injection_faculae.py:191-244, BIS_inject_spectral_quiet / _active):
// every pixel is moved by a velocity that depends on its own flux (a polynomial of
// the flux, numpy.poly1d coefficient order), then the moved samples are resampled
// linearly back onto the original wavelengths (scipy interp1d, linear, on the
// samples sorted by moved wavelength), zero outside the open range of the moved
// wavelengths.  Arithmetic follows the Python expressions term by term without
// contraction, so the results are those of the numpy restatement to rounding.
#include <hip/hip_runtime.h>
#include <math.h>

#include <hipcub/hipcub.hpp>

#include "soap_internal.h"

namespace soap {

constexpr int kInjT = 256;

// numpy.poly1d evaluation (polyval): y = y * x + c over the coefficients, highest
// degree first
__device__ __forceinline__ double inj_poly(const double *c, int nc, double x) {
#pragma clang fp contract(off)
  double y = 0.0;
  for (int i = 0; i < nc; i++) y = y * x + c[i];
  return y;
}

// moved wavelength of pixel k (CB_injection_faculae.py:119-123 Delta_wavelength)
__device__ __forceinline__ double inj_moved(const InjectArgs &a, double w, double f) {
#pragma clang fp contract(off)
  double rv;
  bool shift = true;
  if (a.mode == 0) {
    rv = inj_poly(a.up, a.n_up, f) * a.scale;
  } else if ((f < a.dep_cut) && (f > 0.5)) {
    rv = (inj_poly(a.up, a.n_up, f) + a.velo_shift) * a.scale;
  } else if (f <= 0.5) {
    rv = (inj_poly(a.low, a.n_low, f) + a.velo_shift) * a.scale;
  } else if (f > a.dep_cut) {
    rv = (a.vel_cut + a.velo_shift) * a.scale;
  } else {  // f == dep_cut (or NaN): unmoved
    rv = 0.0;
    shift = false;
  }
  if (!shift) return w;
  const double beta = rv / kC;
  const double dw = w * (sqrt((1 + beta) / (1 - beta)) - 1);
  return w + dw;
}

// moved wavelengths, with per-block min, max and a flag for a descending neighbour
// pair (the moved samples then need sorting)
__global__ __launch_bounds__(kInjT) void k_inject_move(InjectArgs a, const double *wave, const double *flux, int64_t n,
                                                       double *moved, double *part) {
  __shared__ double s_min[kInjT], s_max[kInjT];
  __shared__ int s_desc;
  if (threadIdx.x == 0) s_desc = 0;
  __syncthreads();
  int64_t k = (int64_t)blockIdx.x * kInjT + threadIdx.x;
  double mn = HUGE_VAL, mx = -HUGE_VAL;
  if (k < n) {
    double x = inj_moved(a, wave[k], flux[k]);
    moved[k] = x;
    mn = x;
    mx = x;
    if (k + 1 < n) {
      double x1 = inj_moved(a, wave[k + 1], flux[k + 1]);
      if (x1 < x) s_desc = 1;  // benign race: every writer stores 1
    }
  }
  s_min[threadIdx.x] = mn;
  s_max[threadIdx.x] = mx;
  __syncthreads();
  for (int off = kInjT / 2; off > 0; off >>= 1) {
    if (threadIdx.x < off) {
      s_min[threadIdx.x] = fmin(s_min[threadIdx.x], s_min[threadIdx.x + off]);
      s_max[threadIdx.x] = fmax(s_max[threadIdx.x], s_max[threadIdx.x + off]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[blockIdx.x * 3] = s_min[0];
    part[blockIdx.x * 3 + 1] = s_max[0];
    part[blockIdx.x * 3 + 2] = s_desc ? 1.0 : 0.0;
  }
}

// summary[3] = (min, max, 1 if any descending pair) over the block partials
__global__ __launch_bounds__(kInjT) void k_inject_reduce(const double *part, int nb, double *summary) {
  __shared__ double s_min[kInjT], s_max[kInjT], s_d[kInjT];
  double mn = HUGE_VAL, mx = -HUGE_VAL, d = 0.0;
  for (int b = threadIdx.x; b < nb; b += kInjT) {
    mn = fmin(mn, part[b * 3]);
    mx = fmax(mx, part[b * 3 + 1]);
    d = fmax(d, part[b * 3 + 2]);
  }
  s_min[threadIdx.x] = mn;
  s_max[threadIdx.x] = mx;
  s_d[threadIdx.x] = d;
  __syncthreads();
  for (int off = kInjT / 2; off > 0; off >>= 1) {
    if (threadIdx.x < off) {
      s_min[threadIdx.x] = fmin(s_min[threadIdx.x], s_min[threadIdx.x + off]);
      s_max[threadIdx.x] = fmax(s_max[threadIdx.x], s_max[threadIdx.x + off]);
      s_d[threadIdx.x] = fmax(s_d[threadIdx.x], s_d[threadIdx.x + off]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    summary[0] = s_min[0];
    summary[1] = s_max[0];
    summary[2] = s_d[0];
  }
}

// Linear resampling of (xs, ys) (xs non-decreasing) at the original wavelengths in
// the open range (xmin, xmax), zero elsewhere: the knot interval j with
// xs[j] <= x < xs[j + 1], y = slope (x - xs[j]) + ys[j] (numpy.interp, which scipy's
// linear interp1d calls for float64 data), with its fall-backs for a NaN result.
__global__ __launch_bounds__(kInjT) void k_inject_interp(const double *wave, const double *xs, const double *ys,
                                                         int64_t n, const double *summary, double *out) {
#pragma clang fp contract(off)
  int64_t k = (int64_t)blockIdx.x * kInjT + threadIdx.x;
  if (k >= n) return;
  const double x = wave[k], xmin = summary[0], xmax = summary[1];
  if (!((x < xmax) && (x > xmin))) {
    out[k] = 0.0;
    return;
  }
  // largest j with xs[j] <= x (exists: x > xmin = xs[0]); x < xmax keeps j <= n - 2
  int64_t lo = 0, hi = n - 1;  // invariant: xs[lo] <= x < xs[hi]
  while (hi - lo > 1) {
    int64_t mid = lo + (hi - lo) / 2;
    if (xs[mid] <= x)
      lo = mid;
    else
      hi = mid;
  }
  const int64_t j = lo;
  const double slope = (ys[j + 1] - ys[j]) / (xs[j + 1] - xs[j]);
  double y = slope * (x - xs[j]) + ys[j];
  if (isnan(y)) {
    y = slope * (x - xs[j + 1]) + ys[j + 1];
    if (isnan(y) && ys[j] == ys[j + 1]) y = ys[j];
  }
  out[k] = y;
}

hipError_t launch_inject_move(const InjectArgs &a, const double *wave, const double *flux, int64_t n, double *moved,
                              double *part, double *summary, hipStream_t s) {
  if (n < 2 || a.n_up < 1 || a.n_up > kInjMaxCoef || a.n_low < 0 || a.n_low > kInjMaxCoef) return hipErrorInvalidValue;
  const int nb = inject_blocks(n);
  k_inject_move<<<nb, kInjT, 0, s>>>(a, wave, flux, n, moved, part);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  k_inject_reduce<<<1, kInjT, 0, s>>>(part, nb, summary);
  return hipGetLastError();
}

hipError_t inject_sort_bytes(int64_t n, size_t *bytes) {
  *bytes = 0;
  return hipcub::DeviceRadixSort::SortPairs(nullptr, *bytes, (const double *)nullptr, (double *)nullptr,
                                            (const double *)nullptr, (double *)nullptr, (int)n);
}

// stable sort of (moved, flux) by moved wavelength (the argsort of interp1d)
hipError_t launch_inject_sort(void *tmp, size_t tmp_bytes, const double *moved, double *xs, const double *flux,
                              double *ys, int64_t n, hipStream_t s) {
  return hipcub::DeviceRadixSort::SortPairs(tmp, tmp_bytes, moved, xs, flux, ys, (int)n, 0, 64, s);
}

hipError_t launch_inject_interp(const double *wave, const double *xs, const double *ys, int64_t n,
                                const double *summary, double *out, hipStream_t s) {
  k_inject_interp<<<inject_blocks(n), kInjT, 0, s>>>(wave, xs, ys, n, summary, out);
  return hipGetLastError();
}

int inject_blocks(int64_t n) { return (int)((n + kInjT - 1) / kInjT); }

}  // namespace soap
