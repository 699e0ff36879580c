// Local PHOENIX spectra for the seed preprocessing (SOAP_code/phoenix_extraction.py):
// the effective-temperature grid of the PHOENIX-ACES library and its bracketing
// rule, the library's file names, a reader for the primary image of a FITS file,
// and on the device the vacuum-to-air conversion (Birch and Downs 1994), the
// wavelength windows and the linear interpolation in temperature between the two
// bracketing spectra.  Nothing is downloaded: the files must be local.
#include <hip/hip_runtime.h>
#include <math.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include <hipcub/hipcub.hpp>

#include "soap_gpu.h"
#include "soap_internal.h"

namespace soap {

constexpr int kPhxT = 256;

// vac2air, phoenix_extraction.py:13-22: lambda / n_air(s), s = 1e4 / lambda, the
// terms added left to right as the Python expression does
__device__ __forceinline__ double phx_vac2air(double w) {
#pragma clang fp contract(off)
  const double s = 1.e4 / w;
  const double s2 = s * s;
  const double n_air = 1 + 0.0000834254 + 0.02406147 / (130 - s2) + 0.00015998 / (38.9 - s2);
  return w / n_air;
}

// per pixel: air wavelength, the interpolated flux and the selection flag of
// phoenix_extraction.py:170-185 (vacuum window [wmin - 500, wmax + 500], then air
// window [wmin, wmax]); flux = low + frac (high - low), the difference in float32
// when the spectra are float32 (numpy's float32 - float32)
__global__ __launch_bounds__(kPhxT) void k_phoenix(PhoenixArgs a, const double *wave, const double *lo,
                                                   const double *hi, int64_t n, double *air, double *flux,
                                                   char *flag) {
#pragma clang fp contract(off)
  const int64_t i = (int64_t)blockIdx.x * kPhxT + threadIdx.x;
  if (i >= n) return;
  const double w = wave[i];
  const double x = phx_vac2air(w);
  const double l = lo[i], h = hi[i];
  const double d = a.f32 ? (double)((float)h - (float)l) : h - l;
  air[i] = x;
  flux[i] = l + a.frac * d;
  flag[i] = (w >= a.vlo) && (w <= a.vhi) && (x >= a.wmin) && (x <= a.wmax);
}

hipError_t launch_phoenix(const PhoenixArgs &a, const double *wave, const double *lo, const double *hi, int64_t n,
                          double *air, double *flux, char *flag, hipStream_t s) {
  if (n < 1) return hipErrorInvalidValue;
  k_phoenix<<<(unsigned)((n + kPhxT - 1) / kPhxT), kPhxT, 0, s>>>(a, wave, lo, hi, n, air, flux, flag);
  return hipGetLastError();
}

// compaction of the selected pixels, in order (the boolean-mask indexing of numpy)
hipError_t phoenix_select_bytes(int64_t n, size_t *bytes) {
  *bytes = 0;
  return hipcub::DeviceSelect::Flagged(nullptr, *bytes, (const double *)nullptr, (const char *)nullptr,
                                       (double *)nullptr, (int64_t *)nullptr, (int)n);
}

hipError_t launch_phoenix_select(void *tmp, size_t bytes, const double *in, const char *flag, double *out,
                                 int64_t *count, int64_t n, hipStream_t s) {
  return hipcub::DeviceSelect::Flagged(tmp, bytes, in, flag, out, count, (int)n, s);
}

}  // namespace soap

namespace {

// temp_array of phoenix_extraction.py:92-96: 2300 + 100 j for j < 48, then
// 7000 + 200 (j - 49), so entry 48 is 6800 (the grid is not monotonic there)
double phx_temp(int j) { return j < 48 ? 2300.0 + 100.0 * j : 7000.0 + 200.0 * (j - 49); }
constexpr int kPhxNT = 73;

// numpy.argmin of |grid - x|: the first index of the smallest distance
int argmin_abs(const double *g, int n, double x) {
  int best = 0;
  for (int i = 1; i < n; i++)
    if (std::fabs(g[i] - x) < std::fabs(g[best] - x)) best = i;
  return best;
}

}  // namespace

extern "C" {

// phoenix_extraction.py:136-146
int soap_phoenix_bracket(double temp, double *t_low, double *t_high) {
  if (!t_low || !t_high || !std::isfinite(temp)) return SOAP_ERR_ARG;
  double g[kPhxNT];
  for (int j = 0; j < kPhxNT; j++) g[j] = phx_temp(j);
  const int k = argmin_abs(g, kPhxNT, temp);
  int lo, hi;
  if (temp > g[k]) {
    lo = k;
    hi = k + 1;
  } else {
    lo = k - 1;
    hi = k;
  }
  // beyond the grid the reference indexes past the end (an error) or wraps to the
  // last entry (index -1); both are refused here
  if (lo < 0 || hi >= kPhxNT) return SOAP_ERR_LIMIT;
  *t_low = g[lo];
  *t_high = g[hi];
  return SOAP_OK;
}

// phoenix_extraction.py:84-89,100-104,126-131: the grid values nearest to logg,
// [Fe/H] and [alpha/Fe]; alpha is 0 for [Fe/H] > 0
int soap_phoenix_grid(double logg, double feh, double alpha, double *logg_val, double *feh_val, double *alpha_val) {
  if (!logg_val || !feh_val || !alpha_val) return SOAP_ERR_ARG;
  double lg[13], fe[9], al[8];
  for (int i = 0; i < 13; i++) lg[i] = 6.0 * i / 12;  // linspace(0, 6, 13)
  const double fev[9] = {-4.0, -3.0, -2.0, -1.5, -1.0, -0.5, 0.0, 0.5, 1.0};
  for (int i = 0; i < 9; i++) fe[i] = fev[i];
  const double alv[8] = {-0.2, 0.0, 0.2, 0.4, 0.6, 0.8, 1.0, 1.2};
  for (int i = 0; i < 8; i++) al[i] = alv[i];
  *logg_val = lg[argmin_abs(lg, 13, logg)];
  *feh_val = fe[argmin_abs(fe, 9, feh)];
  *alpha_val = *feh_val > 0.0 ? 0.0 : al[argmin_abs(al, 8, alpha)];
  return SOAP_OK;
}

// PHOENIX_name, phoenix_extraction.py:36-54
int soap_phoenix_name(int32_t temp, double logg, double feh, double alpha, char *subdir, int32_t subdir_cap,
                      char *name, int32_t name_cap) {
  if (!subdir || !name || subdir_cap < 1 || name_cap < 1) return SOAP_ERR_ARG;
  char sd[64], nm[160];
  const char *tail = ".PHOENIX-ACES-AGSS-COND-2011-HiRes.fits";
  if (feh <= 0.0) {
    std::snprintf(sd, sizeof sd, "Z-%.1f", std::fabs(feh));
    if (alpha == 0.0)
      std::snprintf(nm, sizeof nm, "/lte0%d-%.2f-%.1f%s", temp, logg, std::fabs(feh), tail);
    else if (alpha > 0.0)
      std::snprintf(nm, sizeof nm, "/lte0%d-%.2f-%.1f.Alpha=+%.2f%s", temp, logg, std::fabs(feh), alpha, tail);
    else
      std::snprintf(nm, sizeof nm, "/lte0%d-%.2f-%.1f.Alpha=%.2f%s", temp, logg, std::fabs(feh), alpha, tail);
  } else {
    std::snprintf(sd, sizeof sd, "Z+%.1f", std::fabs(feh));
    std::snprintf(nm, sizeof nm, "/lte0%d-%.2f+%.1f%s", temp, logg, std::fabs(feh), tail);
  }
  if ((int)std::strlen(sd) >= subdir_cap || (int)std::strlen(nm) >= name_cap) return SOAP_ERR_LIMIT;
  std::strcpy(subdir, sd);
  std::strcpy(name, nm);
  return SOAP_OK;
}

// The primary image of a FITS file (what astropy's fits.getdata returns for the
// PHOENIX files): 2880-byte header blocks of 80-character cards up to END, then
// big-endian data of BITPIX -64, -32, 32 or 16, product of NAXISn values,
// BSCALE / BZERO applied.  With out null (or cap too small) only n and bitpix are
// reported.
int soap_fits_read(const char *path, double *out, int64_t cap, int64_t *n, int32_t *bitpix) {
  if (!path || !n) return SOAP_ERR_ARG;
  FILE *f = std::fopen(path, "rb");
  if (!f) return SOAP_ERR_IO;
  int bp = 0, naxis = -1;
  int64_t dims[9] = {0};
  double bscale = 1.0, bzero = 0.0;
  bool simple = false, end = false;
  int64_t header_bytes = 0;
  char card[81];
  card[80] = 0;
  for (int c = 0; !end; c++) {
    if (std::fread(card, 1, 80, f) != 80) {
      std::fclose(f);
      return SOAP_ERR_IO;
    }
    header_bytes += 80;
    std::string key(card, 8);
    key.erase(key.find_last_not_of(' ') + 1);
    const char *val = card + 10;
    if (c == 0) simple = key == "SIMPLE" && std::strchr(val, 'T') != nullptr;
    if (key == "END") end = true;
    else if (key == "BITPIX") bp = std::atoi(val);
    else if (key == "NAXIS") naxis = std::atoi(val);
    else if (key.size() > 5 && key.compare(0, 5, "NAXIS") == 0) {
      int k = std::atoi(key.c_str() + 5);
      if (k >= 1 && k <= 8) dims[k] = std::atoll(val);
    } else if (key == "BSCALE") bscale = std::atof(val);
    else if (key == "BZERO") bzero = std::atof(val);
    if (c > 100000) break;
  }
  if (!simple || !end || naxis < 1 || naxis > 8 || (bp != -64 && bp != -32 && bp != 32 && bp != 16)) {
    std::fclose(f);
    return SOAP_ERR_ARG;
  }
  int64_t count = 1;
  for (int k = 1; k <= naxis; k++) {
    if (dims[k] < 1) {
      std::fclose(f);
      return SOAP_ERR_ARG;
    }
    count *= dims[k];
  }
  *n = count;
  if (bitpix) *bitpix = bp;
  if (!out || cap < count) {
    std::fclose(f);
    return SOAP_OK;
  }
  const int64_t data_at = (header_bytes + 2879) / 2880 * 2880;
  const int w = std::abs(bp) / 8;
  std::vector<unsigned char> buf((size_t)count * w);
  if (std::fseek(f, (long)data_at, SEEK_SET) != 0 || std::fread(buf.data(), 1, buf.size(), f) != buf.size()) {
    std::fclose(f);
    return SOAP_ERR_IO;
  }
  std::fclose(f);
  for (int64_t i = 0; i < count; i++) {
    unsigned char b[8];
    for (int k = 0; k < w; k++) b[k] = buf[(size_t)i * w + (w - 1 - k)];  // big-endian -> host order
    double v;
    if (bp == -64) {
      double x;
      std::memcpy(&x, b, 8);
      v = x;
    } else if (bp == -32) {
      float x;
      std::memcpy(&x, b, 4);
      v = x;
    } else if (bp == 32) {
      int32_t x;
      std::memcpy(&x, b, 4);
      v = x;
    } else {
      int16_t x;
      std::memcpy(&x, b, 2);
      v = x;
    }
    out[i] = (bscale == 1.0 && bzero == 0.0) ? v : bzero + bscale * v;
  }
  return SOAP_OK;
}

}  // extern "C"
