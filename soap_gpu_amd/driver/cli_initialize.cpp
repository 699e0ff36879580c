// SOAP_initialize drop-in: the reference's 23 positional arguments and output
// files (SOAP_code/SOAP_initialize.cu:508-610), computed by libsoapgpu.
//   argv: N P W_ratio incl R[km] u1 u2 grid sigma[m/s] step step_med window
//         no_block no_thread seed_cube out_prefix n_mu mu_bounds wave LB0..LB3
// writes {prefix}{mu,delta,limb,region,LB}_map.bin and
// {prefix}_quiet_sun_{spec,wave,lower_spec}.bin (no_block*no_thread values).
#include <chrono>
#include <iostream>

#include "soap_host.hpp"

using namespace soaphost;

int main(int argc, char **argv) {
  if (argc != 24) die("usage: SOAP_initialize <23 positional arguments of the reference>");
  auto t0 = std::chrono::system_clock::now();
  long n = std::atol(argv[1]);
  soap_star st{};
  st.prot_days = std::atof(argv[2]);
  st.w_ratio = std::atof(argv[3]);
  st.incl_deg = std::atof(argv[4]);
  st.radius_km = std::atof(argv[5]);
  st.u1 = std::atof(argv[6]);
  st.u2 = std::atof(argv[7]);
  st.grid = std::atoi(argv[8]);
  double sigma = std::atof(argv[9]);
  double step = std::atof(argv[10]);
  double step_med = std::atof(argv[11]);
  double window = std::atof(argv[12]);
  long no_block = std::atol(argv[13]), no_thread = std::atol(argv[14]);
  std::string cube_name = argv[15], prefix = argv[16];
  int n_mu = std::atoi(argv[17]);
  std::string mub_name = argv[18], wave_name = argv[19];
  for (int k = 0; k < 4; k++) st.lb[k] = std::atof(argv[20 + k]);
  if (window != 50.0) die("window width other than 50 is not supported");
  long nout = no_block * no_thread;
  if (nout > n || nout < 0) die("no_block*no_thread exceeds the vector size");

  std::vector<double> wave = read_bin<double>(wave_name, n);
  std::vector<double> cube = read_bin<double>(cube_name, (size_t)n * n_mu);
  std::vector<double> mub = read_bin<double>(mub_name, n_mu - 1);
  int kind = grid_kind(wave);
  if (kind < 0) die("wavelength grid is neither linear nor log-uniform");

  soap_ctx *ctx = nullptr;
  if (soap_ctx_create(0, &ctx) != SOAP_OK) die("no HIP device");
  check(ctx, soap_set_wave(ctx, kind, wave.data(), n, step), "set_wave");
  check(ctx, soap_set_mu_bounds(ctx, mub.data(), n_mu - 1), "set_mu_bounds");
  check(ctx, soap_set_cube(ctx, SOAP_CUBE_QUIET, cube.data(), n_mu), "set_cube");
  check(ctx, soap_set_cube(ctx, SOAP_CUBE_QUIET_SED, cube.data(), n_mu), "set_cube");
  check(ctx, soap_disk_init(ctx, &st), "disk_init");
  size_t g2 = (size_t)st.grid * st.grid;
  std::vector<double> mu(g2), delta(g2), limb(g2), lb(g2);
  std::vector<int32_t> region(g2);
  check(ctx, soap_get_maps(ctx, mu.data(), delta.data(), limb.data(), region.data(), lb.data()), "get_maps");
  write_bin(prefix + "mu_map.bin", mu.data(), g2);
  write_bin(prefix + "delta_map.bin", delta.data(), g2);
  write_bin(prefix + "limb_map.bin", limb.data(), g2);
  write_bin(prefix + "region_map.bin", region.data(), g2);
  write_bin(prefix + "LB_map.bin", lb.data(), g2);
  check(ctx, soap_quiet_star(ctx), "quiet_star");
  std::vector<double> q(n), low(n);
  check(ctx, soap_get_quiet(ctx, 0, q.data()), "get_quiet");
  check(ctx, soap_resolution(ctx, sigma, step_med), "resolution");
  check(ctx, soap_get_lower(ctx, -1, SOAP_BCONV, low.data()), "get_lower");
  write_bin(prefix + "_quiet_sun_lower_spec.bin", low.data(), nout);
  write_bin(prefix + "_quiet_sun_spec.bin", q.data(), nout);
  write_bin(prefix + "_quiet_sun_wave.bin", wave.data(), nout);
  soap_ctx_destroy(ctx);
  std::chrono::duration<double> dt = std::chrono::system_clock::now() - t0;
  std::cout << "finished in " << dt.count() << "seconds [wall clock]" << std::endl;
  return 0;
}
