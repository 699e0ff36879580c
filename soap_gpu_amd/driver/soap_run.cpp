// soap_run: the config_hpc.cfg-driven host driver, a drop-in for
// SOAP_code/run_all_SOAP.py (:49-455).  All stages run in one process on one
// HIP device through libsoapgpu; the stage files and the .npz products keep
// the reference's names and keys.
//
//   soap_run [config_hpc.cfg] [--no-maps] [--no-stage-files] [--size-evolution[=RATE]]
//
// --size-evolution: the spot-size evolution of SOAP_code/run_all_SOAP_evo.py
// (:214-220, spot_evo): with t = (phase - first phase) * prot days, each generated
// region's size at a phase is sqrt(size^2 - RATE t) (RATE defaults to -0.002, the
// value hard-coded there: the area grows by 0.002 per day).  The reference applies
// it to a single region; here it applies to every region of the configuration.
// Like the reference, it acts on generated maps only (not keyword 'real').
//
// Differences from the Python driver, by design: the output directory is not
// deleted first (files are overwritten), no PNG plot is drawn, and the stage
// executables are not spawned (the same computation runs in-process).
#include <sys/stat.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <chrono>
#include <iostream>

#include "soap_host.hpp"

using namespace soaphost;

static void mkdir_p(const std::string &d) {
  std::string cur;
  for (size_t i = 0; i < d.size(); i++) {
    cur += d[i];
    if (d[i] == '/' || i + 1 == d.size()) mkdir(cur.c_str(), 0755);
  }
}

// np.savetxt default format
static void savetxt(const std::string &path, const std::vector<std::vector<double>> &rows) {
  FILE *f = std::fopen(path.c_str(), "w");
  if (!f) die("cannot write " + path);
  for (auto &r : rows) {
    for (size_t i = 0; i < r.size(); i++) std::fprintf(f, i ? " %.18e" : "%.18e", r[i]);
    std::fprintf(f, "\n");
  }
  std::fclose(f);
}

// mask_templates, run_all_SOAP.py:38-46
static std::string mask_template(int teff) {
  const char *types[] = {"F9", "G2", "G8", "G9", "K2", "K5", "K6", "M0", "M2", "M3", "M4", "M5"};
  const int temps[] = {6050, 5770, 5480, 5380, 5100, 4400, 4300, 3800, 3400, 3250, 3100, 2800};
  int best = 0;
  for (int i = 1; i < 12; i++)
    if (std::abs(temps[i] - teff) < std::abs(temps[best] - teff)) best = i;
  return std::string(types[best]) + "_mask.txt";
}

int main(int argc, char **argv) {
  std::string cfg_name = "config_hpc.cfg";
  bool write_maps = true, stage_files = true, evolve = false;
  double evo_rate = -0.002;
  for (int i = 1; i < argc; i++) {
    std::string a = argv[i];
    if (a == "--no-maps")
      write_maps = false;
    else if (a == "--no-stage-files")
      stage_files = false;
    else if (a == "--size-evolution")
      evolve = true;
    else if (a.rfind("--size-evolution=", 0) == 0) {
      evolve = true;
      char *end = nullptr;
      evo_rate = std::strtod(a.c_str() + 17, &end);
      if (!end || *end) die("--size-evolution=RATE: RATE is not a number");
    } else
      cfg_name = a;
  }
  auto t_start = std::chrono::system_clock::now();
  Ini cfg(cfg_name);
  // run_all_SOAP.py:56-121
  int grid = (int)cfg.integer("main", "grid");
  int nrho = (int)cfg.integer("main", "nrho");
  long inst_reso = cfg.integer("main", "instrument_reso");
  double rot_num = cfg.num("main", "rot_num"), rot_step = cfg.num("main", "rot_step");
  if (inst_reso <= 0) die("instrument_reso must be > 0");
  double sig = sigma_inst((double)inst_reso);
  double rad_sun = (double)cfg.integer("star", "radius_sun");
  double rad = cfg.num("star", "radius") * rad_sun;
  double prot = cfg.num("star", "prot"), w_ratio = cfg.num("star", "w_ratio");
  double incl = cfg.num("star", "i"), u1 = cfg.num("star", "limb1"), u2 = cfg.num("star", "limb2");
  int temp = (int)cfg.integer("star", "tstar");
  int tds = (int)cfg.integer("star", "tdiff_spot"), tdf = (int)cfg.integer("star", "tdiff_faculae");
  int n_act = (int)cfg.integer("star", "n_act");
  std::string input_prefix = cfg.get("data_io", "input_prefix");
  std::string input_dir = input_prefix + cfg.get("data_io", "input_dir");
  std::string output_prefix = cfg.get("data_io", "output_prefix") + "_phoenix" + std::to_string(temp);
  std::string maps_dir = output_prefix + cfg.get("data_io", "maps_dir");
  std::string raw_dir = output_prefix + cfg.get("data_io", "raw_dir");
  std::string spec_dir = output_prefix + cfg.get("data_io", "spec_dir");
  std::string fprefix = cfg.get("data_io", "file_prefix");
  std::string T = std::to_string(temp), Ts = std::to_string(temp - tds), Tf = std::to_string(temp + tdf);
  std::string cube_names[6] = {"new_noconti_convec_T_eff_" + T + ".bin", "new_conti_convec_T_eff_" + T + ".bin",
                               "new_noconti_convec_T_eff_" + Ts + ".bin", "new_conti_T_eff_" + Ts + ".bin",
                               "new_noconti_convec_T_eff_" + Tf + ".bin", "new_conti_T_eff_" + Tf + ".bin"};
  double ccf_window = cfg.num("data_io", "ccf_window"), ccf_size = cfg.num("data_io", "ccf_size");
  std::string keyword = cfg.get("data_io", "keyword");
  long no_thread = cfg.integer("data_io", "no_thread");

  // inputs, run_all_SOAP.py:123-145
  std::vector<double> wave = read_bin_all<double>(input_dir + cfg.get("data_io", "wave_array_name"));
  long n = (long)wave.size();
  if (n < 8) die("wavelength file too short");
  double step_size = std::round((wave[3] - wave[2]) * 1e5) / 1e5;
  long no_block = n / no_thread, nout = no_block * no_thread;
  std::vector<double> mub = read_bin_all<double>(input_dir + cfg.get("data_io", "mu_array_name"));
  int n_mu = (int)mub.size() + 1;
  auto lbf = read_npz_f64(input_dir + "Limb_brightening_info.npz");
  std::vector<double> lb = lb_params(temp, lbf["LB_G2"], lbf["LB_K0"], lbf["LB_M0"], lbf["Temp_array"]);
  int kind = grid_kind(wave);
  if (kind < 0) die("wavelength grid is neither linear nor log-uniform");

  // phases and regions, run_all_SOAP.py:148-220
  std::string ploc = output_prefix + "/" + fprefix + "_" + "phase_locator.txt";
  std::vector<double> phases;
  std::vector<int32_t> counts, ids;
  std::vector<soap_region> regs;
  mkdir_p(output_prefix);
  mkdir_p(raw_dir);
  mkdir_p(spec_dir);
  if (keyword == "real") {
    phases = read_numbers(ploc);
    std::vector<double> info = read_numbers(maps_dir + "final_info.txt");
    for (size_t i = 0; i < phases.size(); i++) {
      if (info.size() < 2 * i + 2) die("final_info.txt shorter than the phase list");
      ids.push_back((int32_t)info[2 * i]);
      counts.push_back((int32_t)info[2 * i + 1]);
      std::vector<double> m = read_numbers(maps_dir + "maps" + std::to_string(ids.back()) + "_for_all.txt");
      if ((long)m.size() < 4L * counts.back()) die("region file too short");
      for (int r = 0; r < counts.back(); r++) regs.push_back({m[4 * r], m[4 * r + 1], m[4 * r + 2], m[4 * r + 3]});
    }
  } else {
    mkdir_p(maps_dir);
    std::vector<soap_region> list;
    for (int a = 0; a < n_act; a++) {
      std::string key = "active_regions_" + std::to_string(a + 1);
      list.push_back({cfg.num(key, "latitude"), cfg.num(key, "longitude"), cfg.num(key, "size"), cfg.num(key, "type")});
    }
    phases = arange(0.0, rot_num, rot_step);
    std::vector<std::vector<double>> prow, info;
    for (double p : phases) prow.push_back({p});
    savetxt(ploc, prow);
    for (size_t i = 0; i < phases.size(); i++) {
      std::vector<soap_region> at = list;
      if (evolve) {  // run_all_SOAP_evo.py:214-220: size = sqrt(spot_evo(size^2, rate, t))
        double t = (phases[i] - phases[0]) * prot;
        for (auto &r : at) {
          double area = r.size * r.size - evo_rate * t;
          if (!(area >= 0)) die("size evolution: negative area at phase " + std::to_string(phases[i]));
          r.size = std::sqrt(area);
        }
      }
      std::vector<std::vector<double>> rows;
      for (auto &r : at) rows.push_back({r.lat_deg, r.lon_deg, r.size, r.type});
      savetxt(maps_dir + "maps" + std::to_string(i) + "_for_all.txt", rows);
      info.push_back({(double)i, (double)at.size()});
      ids.push_back((int32_t)i);
      counts.push_back((int32_t)at.size());
      regs.insert(regs.end(), at.begin(), at.end());
    }
    savetxt(maps_dir + "final_info.txt", info);
  }
  int P = (int)phases.size();

  soap_ctx *ctx = nullptr;
  if (soap_ctx_create(0, &ctx) != SOAP_OK) die("no HIP device");
  check(ctx, soap_set_wave(ctx, kind, wave.data(), n, step_size), "set_wave");
  check(ctx, soap_set_mu_bounds(ctx, mub.data(), n_mu - 1), "set_mu_bounds");
  for (int k = 0; k < 6; k++) {
    std::vector<double> cube = read_bin<double>(input_dir + cube_names[k], (size_t)n * n_mu);
    check(ctx, soap_set_cube(ctx, k, cube.data(), n_mu), "set_cube");
  }
  soap_star st{};
  st.prot_days = prot;
  st.w_ratio = w_ratio;
  st.incl_deg = incl;
  st.radius_km = rad;
  st.u1 = u1;
  st.u2 = u2;
  for (int k = 0; k < 4; k++) st.lb[k] = lb[k];
  st.grid = grid;
  soap_activity act{nrho, temp, tds, tdf, 0};
  check(ctx, soap_disk_init(ctx, &st), "disk_init");
  check(ctx, soap_quiet_star(ctx), "quiet_star");
  check(ctx, soap_set_activity(ctx, &act, P, phases.data(), counts.data(), regs.data()), "set_activity");
  check(ctx, soap_integrate(ctx), "integrate");
  check(ctx, soap_resolution(ctx, sig, step_size), "resolution");

  // mask lines, run_all_SOAP.py:262-296
  std::vector<double> vgrid = arange(-1 * ccf_window, ccf_window + 1., ccf_size);
  int n_v = (int)vgrid.size();
  const int trim = 210;
  double wmin = wave[trim], wmax = wave[trim];
  for (long j = trim; j < nout - trim; j++) wmin = std::min(wmin, wave[j]), wmax = std::max(wmax, wave[j]);
  std::vector<double> tmpl = read_numbers(input_prefix + cfg.get("data_io", "mask_prefix") + mask_template(temp));
  std::vector<double> ll, lw;
  for (size_t i = 0; i + 1 < tmpl.size(); i += 2)
    if (tmpl[i] > wmin && tmpl[i] < wmax && tmpl[i + 1] > 0.1) ll.push_back(tmpl[i]), lw.push_back(tmpl[i + 1]);
  check(ctx, soap_set_mask(ctx, ll.data(), lw.data(), (int)ll.size(), vgrid.data(), n_v, trim), "set_mask");
  check(ctx, soap_ccf(ctx), "ccf");
  check(ctx, soap_fit(ctx), "fit");
  auto t_gpu = std::chrono::system_clock::now();

  // stage files of the reference's executables
  std::string op = output_prefix + "/" + fprefix;
  std::vector<double> buf(n);
  if (stage_files) {
    size_t g2 = (size_t)grid * grid;
    std::vector<double> mu(g2), dl(g2), li(g2), lbm(g2);
    std::vector<int32_t> rg(g2);
    check(ctx, soap_get_maps(ctx, mu.data(), dl.data(), li.data(), rg.data(), lbm.data()), "get_maps");
    for (std::string pre : {op, op + "_SED"}) {
      write_bin(pre + "mu_map.bin", mu.data(), g2);
      write_bin(pre + "delta_map.bin", dl.data(), g2);
      write_bin(pre + "limb_map.bin", li.data(), g2);
      write_bin(pre + "region_map.bin", rg.data(), g2);
      write_bin(pre + "LB_map.bin", lbm.data(), g2);
      write_bin(pre + "_quiet_sun_wave.bin", wave.data(), nout);
    }
    check(ctx, soap_get_quiet(ctx, 0, buf.data()), "get_quiet");
    write_bin(op + "_quiet_sun_spec.bin", buf.data(), nout);
    check(ctx, soap_get_quiet(ctx, 1, buf.data()), "get_quiet");
    write_bin(op + "_SED_quiet_sun_spec.bin", buf.data(), nout);
    check(ctx, soap_get_lower(ctx, -1, SOAP_BCONV, buf.data()), "get_lower");
    write_bin(op + "_quiet_sun_lower_spec.bin", buf.data(), nout);
    check(ctx, soap_get_lower(ctx, -1, SOAP_TOT, buf.data()), "get_lower");
    write_bin(op + "_SED_quiet_sun_lower_spec.bin", buf.data(), nout);
    const char *tn[3] = {"flux", "bconv", "tot"};
    std::vector<double> f(n), b(n), t(n), tm(g2), bm(g2);
    for (int i = 0; i < P; i++) {
      std::string id = std::to_string(ids[i]);
      check(ctx, soap_get_raw(ctx, i, f.data(), b.data(), t.data()), "get_raw");
      write_bin(raw_dir + fprefix + "_Spec_flux_" + id + "_.bin", f.data(), nout);
      write_bin(raw_dir + fprefix + "_Spec_bconv_" + id + "_.bin", b.data(), nout);
      write_bin(raw_dir + fprefix + "_Spec_tot_" + id + "_.bin", t.data(), nout);
      if (write_maps) {
        check(ctx, soap_phase_maps(ctx, i, tm.data(), bm.data()), "phase_maps");
        write_bin(raw_dir + fprefix + "_temp_map_" + id + "_.bin", tm.data(), g2);
        write_bin(raw_dir + fprefix + "_spot_binary_map_" + id + "_.bin", bm.data(), g2);
      }
      for (int ty = 0; ty < 3; ty++) {
        check(ctx, soap_get_lower(ctx, i, ty, buf.data()), "get_lower");
        write_bin(spec_dir + fprefix + "_lower_spec_" + tn[ty] + "_" + std::to_string(i) + "_.bin", buf.data(), nout);
      }
    }
  }

  // RV products, run_all_SOAP.py:327-449
  size_t nf = (size_t)P + 1;
  std::vector<double> ccf(nf * 3 * n_v), par(nf * 3 * 5), model(nf * 3 * n_v), depth(nf * 3 * 91), bis(nf * 3 * 91);
  check(ctx, soap_get_ccf(ctx, ccf.data()), "get_ccf");
  check(ctx, soap_get_fit(ctx, par.data(), model.data(), depth.data(), bis.data()), "get_fit");
  soap_ctx_destroy(ctx);
  const char *types[3] = {"tot", "flux", "bconv"};
  const int tidx[3] = {SOAP_TOT, SOAP_FLUX, SOAP_BCONV};
  std::vector<double> rvs[3];
  for (int ti = 0; ti < 3; ti++) {
    int t = tidx[ti];
    std::vector<double> cc, md, cont, contr, span, rv, fwhm, dp, bs;
    for (size_t f = 0; f < nf; f++) {
      size_t row = f * 3 + t;
      cc.insert(cc.end(), &ccf[row * n_v], &ccf[row * n_v] + n_v);
      md.insert(md.end(), &model[row * n_v], &model[row * n_v] + n_v);
      dp.insert(dp.end(), &depth[row * 91], &depth[row * 91] + 91);
      bs.insert(bs.end(), &bis[row * 91], &bis[row * 91] + 91);
      cont.push_back(par[row * 5 + 0]);
      contr.push_back(par[row * 5 + 1]);
      span.push_back(par[row * 5 + 2]);
      rv.push_back(par[row * 5 + 3]);
      fwhm.push_back(par[row * 5 + 4]);
    }
    double ref = rv.back();
    for (double &x : rv) x -= ref;
    std::vector<double> rv_out(rv.begin(), rv.end() - 1);
    rvs[ti] = rv_out;
    write_npz(op + "_RV_" + types[ti] + "_GPU.npz",
              {{"time", npy_f64(phases)},
               {"RV", npy_f64(rv_out)},
               {"ccf", npy_f64(cc, {nf, (size_t)n_v})},
               {"model", npy_f64(md, {nf, (size_t)n_v})},
               {"continuum", npy_f64(cont)},
               {"contrast", npy_f64(contr)},
               {"span", npy_f64(span)},
               {"fwhm", npy_f64(fwhm)},
               {"depth", npy_f64(dp, {nf, (size_t)91})},
               {"bis", npy_f64(bs, {nf, (size_t)91})}});
  }
  write_npz(op + "_GPU_SOAP_RV_data.npz",
            {{"time", npy_f64(phases)}, {"tot", npy_f64(rvs[0])}, {"flux", npy_f64(rvs[1])}, {"bconv", npy_f64(rvs[2])}});
  // keep a copy of the configuration next to the products (run_all_SOAP.py:452-453)
  {
    std::ifstream in(cfg_name, std::ios::binary);
    std::ofstream out(op + "_config_file.cfg", std::ios::binary);
    out << in.rdbuf();
  }
  std::chrono::duration<double> g = t_gpu - t_start, tot = std::chrono::system_clock::now() - t_start;
  std::cout << "GPU Calculation time (s): " << g.count() << std::endl;
  std::cout << "Computing time (s): " << tot.count() << std::endl;
  return 0;
}
