// Driver for the host helpers of the drop-in executables (soap_gpu_amd/driver/
// soap_host.hpp) built under AddressSanitizer and UndefinedBehaviorSanitizer
// (tests/test_sanitize.py): .bin and text reading, the grid-kind check, the .npz
// writer and reader, polyfit and the limb-brightening fit.  Exit status 0 and no
// sanitizer report is the pass.
#include <cmath>
#include <cstdio>
#include <string>
#include <vector>

#include "soap_host.hpp"

using namespace soaphost;

int main(int argc, char **argv) {
  const std::string dir = argc > 1 ? argv[1] : ".";
  std::vector<double> lin(5000), lg(5000);
  for (size_t i = 0; i < lin.size(); i++) {
    lin[i] = 5000.0 + 0.005 * i;
    lg[i] = 5000.0 * std::exp(1e-6 * i);
  }
  write_bin(dir + "/w.bin", lin.data(), lin.size());
  auto back = read_bin<double>(dir + "/w.bin", lin.size());
  auto all = read_bin_all<double>(dir + "/w.bin");
  if (back != lin || all != lin) return 2;
  std::FILE *f = std::fopen((dir + "/n.txt").c_str(), "w");
  std::fprintf(f, "1.5 2.5\n3.5\n# comment\n4.5\n");
  std::fclose(f);
  auto nums = read_numbers(dir + "/n.txt");
  if (nums.empty()) return 3;
  if (grid_kind(lin) == grid_kind(lg)) return 4;
  std::vector<double> x{0, 1, 2, 3, 4, 5}, y;
  for (double v : x) y.push_back(1.0 - 2.0 * v + 0.5 * v * v);
  auto c = polyfit(x, y, 2);
  if (c.size() != 3) return 5;
  NpyArray a = npy_f64(lin, {lin.size()});
  NpyArray b = npy_f64(std::vector<double>(y), {2, 3});
  write_npz(dir + "/t.npz", {{"wave", a}, {"flux", b}});
  auto m = read_npz_f64(dir + "/t.npz");
  if (m["wave"] != lin || m["flux"].size() != 6) return 6;
  auto r = arange(0.0, 1.0, 0.125);
  std::printf("host helpers under sanitizers: ok (%zu numbers, %zu arange, %.3g)\n", nums.size(), r.size(), c[0]);
  return 0;
}
