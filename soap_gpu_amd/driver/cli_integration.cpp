// SOAP_integration / SOAP_integration_solar drop-in: the reference's 24
// positional arguments, input files and per-phase outputs
// (SOAP_code/SOAP_integration.cu:464-860), computed by libsoapgpu.
//   argv: grid nrho n_phase T* dTspot incl dTfac phase_file fac_cube N raw_dir
//         maps_dir no_block no_thread quiet_cube spot_cube prefix step n_mu
//         wave quietSED spotSED facSED map_dir
// writes {raw_dir}{prefix}_Spec_{flux,bconv,tot}_{id}_.bin (no_block*no_thread
// values) and _temp_map_{id}_.bin, _spot_binary_map_{id}_.bin (grid^2 values).
// Built twice: SOAP_integration (solar_mode 0) and SOAP_integration_solar (1).
#include <chrono>
#include <iostream>

#include "soap_host.hpp"

#ifndef SOAP_SOLAR_MODE
#define SOAP_SOLAR_MODE 0
#endif

using namespace soaphost;

int main(int argc, char **argv) {
  if (argc != 25) die("usage: SOAP_integration <24 positional arguments of the reference>");
  auto t0 = std::chrono::system_clock::now();
  soap_star st{};
  st.grid = std::atoi(argv[1]);
  soap_activity act{};
  act.nrho = std::atoi(argv[2]);
  int npsi = std::atoi(argv[3]);
  act.t_star = std::atoi(argv[4]);
  act.t_diff_spot = std::atoi(argv[5]);
  st.incl_deg = std::atof(argv[6]);
  act.t_diff_fac = std::atoi(argv[7]);
  act.solar_mode = SOAP_SOLAR_MODE;
  std::string phase_file = argv[8], fac_name = argv[9];
  long n = std::atol(argv[10]);
  std::string out_dir = argv[11], map_dir = argv[12];
  long no_block = std::atol(argv[13]), no_thread = std::atol(argv[14]);
  std::string q_name = argv[15], sp_name = argv[16], prefix = argv[17];
  double step = std::atof(argv[18]);
  int n_mu = std::atoi(argv[19]);
  std::string wave_name = argv[20], qsed_name = argv[21], spsed_name = argv[22], facsed_name = argv[23];
  std::string map_soap_dir = argv[24];
  long nout = no_block * no_thread;
  if (nout > n || nout < 0) die("no_block*no_thread exceeds the vector size");
  st.prot_days = 1.0;  // unused here: the maps come from files
  st.radius_km = 1.0;

  std::vector<double> wave = read_bin<double>(wave_name, n);
  int kind = grid_kind(wave);
  if (kind < 0) die("wavelength grid is neither linear nor log-uniform");
  soap_ctx *ctx = nullptr;
  if (soap_ctx_create(0, &ctx) != SOAP_OK) die("no HIP device");
  check(ctx, soap_set_wave(ctx, kind, wave.data(), n, step), "set_wave");
  std::vector<double> bounds(n_mu - 1, 0.0);  // mu bins come from region_map.bin
  check(ctx, soap_set_mu_bounds(ctx, bounds.data(), n_mu - 1), "set_mu_bounds");
  const char *names[6] = {q_name.c_str(), qsed_name.c_str(), sp_name.c_str(), spsed_name.c_str(), fac_name.c_str(),
                          facsed_name.c_str()};
  for (int c = 0; c < 6; c++) {
    std::vector<double> cube = read_bin<double>(names[c], (size_t)n * n_mu);
    check(ctx, soap_set_cube(ctx, c, cube.data(), n_mu), "set_cube");
  }
  size_t g2 = (size_t)st.grid * st.grid;
  std::string mp = map_soap_dir + prefix;
  std::vector<int32_t> region = read_bin<int32_t>(mp + "region_map.bin", g2);
  std::vector<double> delta = read_bin<double>(mp + "delta_map.bin", g2);
  std::vector<double> limb = read_bin<double>(mp + "limb_map.bin", g2);
  std::vector<double> lb = read_bin<double>(mp + "LB_map.bin", g2);
  check(ctx, soap_set_maps(ctx, &st, delta.data(), limb.data(), region.data(), lb.data()), "set_maps");

  std::vector<double> psi = read_numbers(phase_file);
  if ((int)psi.size() < npsi) die(phase_file + ": fewer phases than n_phase");
  psi.resize(npsi);
  std::vector<double> info = read_numbers(map_dir + "final_info.txt");
  if ((int)info.size() < 2 * npsi) die(map_dir + "final_info.txt: fewer rows than n_phase");
  std::vector<int32_t> counts(npsi), ids(npsi);
  std::vector<soap_region> regs;
  for (int i = 0; i < npsi; i++) {
    ids[i] = (int32_t)info[2 * i];
    counts[i] = (int32_t)info[2 * i + 1];
    std::vector<double> m = read_numbers(map_dir + "maps" + std::to_string(ids[i]) + "_for_all.txt");
    if ((int)m.size() < 4 * counts[i]) die("region file of phase " + std::to_string(ids[i]) + " is short");
    for (int r = 0; r < counts[i]; r++) regs.push_back({m[4 * r], m[4 * r + 1], m[4 * r + 2], m[4 * r + 3]});
  }
  check(ctx, soap_set_activity(ctx, &act, npsi, psi.data(), counts.data(), regs.data()), "set_activity");
  check(ctx, soap_integrate(ctx), "integrate");
  std::vector<double> f(n), b(n), t(n), tm(g2), bm(g2);
  for (int i = 0; i < npsi; i++) {
    std::string id = std::to_string(ids[i]);
    check(ctx, soap_get_raw(ctx, i, f.data(), b.data(), t.data()), "get_raw");
    write_bin(out_dir + prefix + "_Spec_flux_" + id + "_.bin", f.data(), nout);
    write_bin(out_dir + prefix + "_Spec_bconv_" + id + "_.bin", b.data(), nout);
    write_bin(out_dir + prefix + "_Spec_tot_" + id + "_.bin", t.data(), nout);
    std::cout << out_dir + prefix + "_Spec_tot_" + id + "_.bin" << std::endl;
    check(ctx, soap_phase_maps(ctx, i, tm.data(), bm.data()), "phase_maps");
    write_bin(out_dir + prefix + "_temp_map_" + id + "_.bin", tm.data(), g2);
    write_bin(out_dir + prefix + "_spot_binary_map_" + id + "_.bin", bm.data(), g2);
  }
  soap_ctx_destroy(ctx);
  std::chrono::duration<double> dt = std::chrono::system_clock::now() - t0;
  std::cout << "finished in " << dt.count() << " seconds [wall clock]" << std::endl;
  return 0;
}
