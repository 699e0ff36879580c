// SOAP_resolution drop-in: the reference's 9 positional arguments and files
// (SOAP_code/SOAP_resolution.cu:318-440), computed by libsoapgpu.
//   argv: n_phase N sigma[m/s] raw_dir spec_dir no_block no_thread prefix quiet_prefix
// reads {quiet_prefix}_quiet_sun_spec.bin, _SED_quiet_sun_spec.bin,
// _quiet_sun_wave.bin and {raw_dir}{prefix}_Spec_{type}_{i}_.bin; writes
// {spec_dir}{prefix}_lower_spec_{type}_{i}_.bin (no_block*no_thread values).
#include <chrono>
#include <iostream>

#include "soap_host.hpp"

using namespace soaphost;

int main(int argc, char **argv) {
  if (argc != 10) die("usage: SOAP_resolution <9 positional arguments of the reference>");
  auto t0 = std::chrono::system_clock::now();
  int npsi = std::atoi(argv[1]);
  long n = std::atol(argv[2]);
  double sigma = std::atof(argv[3]);
  std::string raw_dir = argv[4], spec_dir = argv[5];
  long no_block = std::atol(argv[6]), no_thread = std::atol(argv[7]);
  std::string prefix = argv[8], qprefix = argv[9];
  long nout = no_block * no_thread;
  if (nout > n || nout < 8) die("no_block*no_thread must be in [8, N]");
  // the stage files hold no_block*no_thread values; the tail of the grid is
  // continued from them and the tail of each spectrum is zero
  std::vector<double> wave = read_bin<double>(qprefix + "_quiet_sun_wave.bin", nout);
  int kind = grid_kind(wave);
  if (kind < 0) die("wavelength grid is neither linear nor log-uniform");
  double h = std::log(wave[nout - 1] / wave[0]) / (double)(nout - 1), st = wave[1] - wave[0];
  for (long j = nout; j < n; j++) wave.push_back(kind == SOAP_GRID_LOG ? wave[0] * std::exp((double)j * h) : wave[0] + j * st);
  auto padded = [&](const std::string &p) {
    std::vector<double> v = read_bin<double>(p, nout);
    v.resize(n, 0.0);
    return v;
  };
  soap_ctx *ctx = nullptr;
  if (soap_ctx_create(0, &ctx) != SOAP_OK) die("no HIP device");
  check(ctx, soap_set_wave(ctx, kind, wave.data(), n, st), "set_wave");
  std::vector<double> q = padded(qprefix + "_quiet_sun_spec.bin"), qs = padded(qprefix + "_SED_quiet_sun_spec.bin");
  check(ctx, soap_set_quiet(ctx, 0, q.data()), "set_quiet");
  check(ctx, soap_set_quiet(ctx, 1, qs.data()), "set_quiet");
  check(ctx, soap_alloc_phases(ctx, npsi), "alloc_phases");
  for (int i = 0; i < npsi; i++) {
    std::string id = std::to_string(i);
    std::vector<double> f = padded(raw_dir + prefix + "_Spec_flux_" + id + "_.bin");
    std::vector<double> b = padded(raw_dir + prefix + "_Spec_bconv_" + id + "_.bin");
    std::vector<double> t = padded(raw_dir + prefix + "_Spec_tot_" + id + "_.bin");
    check(ctx, soap_set_raw(ctx, i, f.data(), b.data(), t.data()), "set_raw");
  }
  check(ctx, soap_resolution(ctx, sigma, 0.005), "resolution");
  const char *tn[3] = {"flux", "bconv", "tot"};
  std::vector<double> out(n);
  for (int i = 0; i < npsi; i++) {
    std::printf("PHASE = %d over %d\n", i, npsi);
    for (int t = 0; t < 3; t++) {
      check(ctx, soap_get_lower(ctx, i, t, out.data()), "get_lower");
      write_bin(spec_dir + prefix + "_lower_spec_" + tn[t] + "_" + std::to_string(i) + "_.bin", out.data(), nout);
    }
  }
  soap_ctx_destroy(ctx);
  std::chrono::duration<double> dt = std::chrono::system_clock::now() - t0;
  std::cout << "finished in " << dt.count() << " seconds [wall clock]" << std::endl;
  return 0;
}
