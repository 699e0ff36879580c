// BIS_FIT2 drop-in (SOAP_code/BIS_FIT2.cpp:97-320): argv[1] = number of CCF
// points; reads CCF_data.txt ("v ccf") and CCF_error.txt from the working
// directory, writes ccf_model.bin, ccf_parameter.bin (c, |k|, span, v0, fwhm),
// ccf_depth.bin and ccf_bis.bin.  The fit runs on the GPU (soap_fit_ccf_host).
#include <iostream>

#include "soap_host.hpp"

using namespace soaphost;

int main(int argc, char **argv) {
  if (argc != 2) die("usage: BIS_FIT2 <n>");
  int n = std::atoi(argv[1]);
  std::vector<double> d = read_numbers("CCF_data.txt"), e = read_numbers("CCF_error.txt");
  if ((int)d.size() < 2 * n || (int)e.size() < n) die("CCF_data.txt / CCF_error.txt shorter than n");
  std::vector<double> v(n), y(n), err(n);
  for (int i = 0; i < n; i++) {
    v[i] = d[2 * i];
    y[i] = d[2 * i + 1];
    err[i] = e[i];  // residuals and Jacobian rows are divided by these (BIS_FIT2.cpp:49,74-77)
  }
  soap_ctx *ctx = nullptr;
  if (soap_ctx_create(0, &ctx) != SOAP_OK) die("no HIP device");
  std::vector<double> par(5), model(n), depth(91), bis(91);
  int32_t status = 0;
  check(ctx, soap_fit_ccf_host_err(ctx, v.data(), n, y.data(), err.data(), 1, par.data(), model.data(), depth.data(),
                                   bis.data(), &status),
        "fit_ccf_host_err");
  // the reference writes its outputs whatever the solver's outcome (BIS_FIT2.cpp:191); so does this tool
  if (status == 27) std::cerr << "BIS_FIT2: the fit made no progress (flat or degenerate CCF)\n";
  if (status == 11) std::cerr << "BIS_FIT2: the fit stopped at 10000 iterations\n";
  soap_ctx_destroy(ctx);
  write_bin("ccf_model.bin", model.data(), n);
  write_bin("ccf_parameter.bin", par.data(), 5);
  write_bin("ccf_depth.bin", depth.data(), 91);
  write_bin("ccf_bis.bin", bis.data(), 91);
  return 0;
}
