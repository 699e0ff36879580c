// phoenix_extraction drop-in (SOAP_code/phoenix_extraction.py) for local files:
// argv[1] = config (default config_conti_phoenix.cfg).  For the star, spot and
// facula temperatures (Tstar, Tstar - Tdiff_spot, Tstar + Tdiff_faculae) it finds
// the two bracketing PHOENIX-ACES spectra and the library's wavelength file in
// input_prefix + input_dir_source, converts the vacuum wavelengths to air, keeps
// [minwave, maxwave] and interpolates linearly in temperature (on the GPU), and
// writes New_Phoenix_<T>.npz (arrays wave, flux) there.  Nothing is downloaded: a
// missing file is an error.  The reference's pandas pickle and its RASSINE call
// are not reproduced.
#include <cstdio>
#include <iostream>

#include "soap_host.hpp"

using namespace soaphost;

namespace {

std::vector<double> read_fits(const std::string &path, int32_t *bitpix) {
  int64_t n = 0;
  int rc = soap_fits_read(path.c_str(), nullptr, 0, &n, bitpix);
  if (rc == SOAP_ERR_IO) die("local PHOENIX file missing or unreadable: " + path + " (nothing is downloaded)");
  if (rc != SOAP_OK) die("not a single-image FITS file: " + path);
  std::vector<double> v((size_t)n);
  if (soap_fits_read(path.c_str(), v.data(), n, &n, bitpix) != SOAP_OK) die("cannot read " + path);
  return v;
}

}  // namespace

int main(int argc, char **argv) {
  std::string cfg_name = argc > 1 ? argv[1] : "config_conti_phoenix.cfg";
  Ini cfg(cfg_name);
  const long tstar = cfg.integer("star", "tstar");
  const long tds = cfg.integer("star", "tdiff_spot"), tdf = cfg.integer("star", "tdiff_faculae");
  const double grav = cfg.num("star", "logg"), wvmin = cfg.num("star", "minwave"), wvmax = cfg.num("star", "maxwave");
  const std::string out_path = cfg.get("data_io", "input_prefix") + cfg.get("data_io", "input_dir_source");
  double logg_val, feh_val, alpha_val;  // [Fe/H] = [alpha/Fe] = 0, as the reference fixes them
  if (soap_phoenix_grid(grav, 0.0, 0.0, &logg_val, &feh_val, &alpha_val) != SOAP_OK) die("phoenix_grid");
  int32_t bp_wave = 0;
  const std::vector<double> wave = read_fits(out_path + "WAVE_PHOENIX-ACES-AGSS-COND-2011.fits", &bp_wave);
  soap_ctx *ctx = nullptr;
  if (soap_ctx_create(0, &ctx) != SOAP_OK) die("no HIP device");
  for (long temp : {tstar, tstar - tds, tstar + tdf}) {
    double t_low, t_high;
    if (soap_phoenix_bracket((double)temp, &t_low, &t_high) != SOAP_OK)
      die("temperature " + std::to_string(temp) + " K is outside the PHOENIX grid");
    std::string files[2];
    for (int k = 0; k < 2; k++) {
      char sub[64], name[256];
      if (soap_phoenix_name((int32_t)(k == 0 ? t_low : t_high), logg_val, feh_val, alpha_val, sub, 64, name, 256) !=
          SOAP_OK)
        die("phoenix_name");
      files[k] = out_path + name;
    }
    int32_t bp_lo = 0, bp_hi = 0;
    const std::vector<double> lo = read_fits(files[0], &bp_lo), hi = read_fits(files[1], &bp_hi);
    if (lo.size() != wave.size() || hi.size() != wave.size())
      die("spectra and wavelength file differ in length (" + files[0] + ", " + files[1] + ")");
    soap_phoenix_params p{};
    p.temp = (double)temp;
    p.t_low = t_low;
    p.t_high = t_high;
    p.wmin = wvmin;
    p.wmax = wvmax;
    p.f32 = bp_lo == -32 && bp_hi == -32;
    std::vector<double> air(wave.size()), flux(wave.size());
    int64_t m = 0;
    check(ctx, soap_phoenix_interp(ctx, &p, wave.data(), lo.data(), hi.data(), (int64_t)wave.size(), air.data(),
                                   flux.data(), &m),
          "phoenix_interp");
    air.resize((size_t)m);
    flux.resize((size_t)m);
    const std::string out = out_path + "New_Phoenix_" + std::to_string(temp) + ".npz";
    write_npz(out, {{"wave", npy_f64(air)}, {"flux", npy_f64(flux)}});
    std::printf("%ld K from %.0f K and %.0f K: %lld pixels -> %s\n", temp, t_low, t_high, (long long)m, out.c_str());
  }
  soap_ctx_destroy(ctx);
  return 0;
}
