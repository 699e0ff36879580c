// Host helpers shared by the SOAP-GPU command-line tools: raw .bin I/O, the text
// files of the reference's stage interface, an INI reader with configparser's
// key syntax, a stored (uncompressed) .npz writer/reader and the derived
// parameters of the driver (limb-brightening fit, instrumental sigma).
#pragma once

#include <cctype>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "soap_gpu.h"

namespace soaphost {

[[noreturn]] inline void die(const std::string &msg) {
  std::fprintf(stderr, "error: %s\n", msg.c_str());
  std::exit(1);
}

inline void check(soap_ctx *ctx, int rc, const char *what) {
  if (rc != SOAP_OK) die(std::string(what) + " failed (" + std::to_string(rc) + "): " + soap_last_error(ctx));
}

// ------------------------------------------------------------------ .bin files
template <class T>
std::vector<T> read_bin(const std::string &path, size_t count) {
  FILE *f = std::fopen(path.c_str(), "rb");
  if (!f) die("cannot open " + path);
  std::vector<T> v(count);
  size_t got = std::fread(v.data(), sizeof(T), count, f);
  std::fclose(f);
  if (got != count) die(path + ": expected " + std::to_string(count) + " values, read " + std::to_string(got));
  return v;
}

template <class T>
std::vector<T> read_bin_all(const std::string &path) {
  FILE *f = std::fopen(path.c_str(), "rb");
  if (!f) die("cannot open " + path);
  std::fseek(f, 0, SEEK_END);
  long bytes = std::ftell(f);
  std::fseek(f, 0, SEEK_SET);
  std::vector<T> v(bytes / sizeof(T));
  size_t got = std::fread(v.data(), sizeof(T), v.size(), f);
  std::fclose(f);
  if (got != v.size()) die("short read of " + path);
  return v;
}

template <class T>
void write_bin(const std::string &path, const T *data, size_t count) {
  FILE *f = std::fopen(path.c_str(), "wb");
  if (!f) die("cannot write " + path);
  if (std::fwrite(data, sizeof(T), count, f) != count) die("short write of " + path);
  std::fclose(f);
}

// whitespace-separated numbers of a text file
inline std::vector<double> read_numbers(const std::string &path) {
  std::ifstream in(path);
  if (!in) die("cannot open " + path);
  std::vector<double> v;
  double x;
  while (in >> x) v.push_back(x);
  return v;
}

// ------------------------------------------------------------------ grid kind
// 1 (log-uniform) if every pixel is lambda_0 exp(j h) to 1e-12 relative,
// 0 (linear) if every step equals the first to 1e-9 relative, -1 otherwise
inline int grid_kind(const std::vector<double> &w) {
  size_t n = w.size();
  if (n < 8) return -1;
  double h = std::log(w[n - 1] / w[0]) / (double)(n - 1);
  bool is_log = true;
  for (size_t j = 0; j < n && is_log; j++)
    if (!(std::fabs(w[j] - w[0] * std::exp((double)j * h)) <= 1e-12 * std::fabs(w[j]))) is_log = false;
  if (is_log) return SOAP_GRID_LOG;
  double st = (w[n - 1] - w[0]) / (double)(n - 1);
  for (size_t j = 0; j < n; j++)
    if (!(std::fabs(w[j] - (w[0] + (double)j * st)) <= 1e-9 * std::fabs(w[j]))) return -1;
  return SOAP_GRID_LINEAR;
}

// ------------------------------------------------------------------ INI
// configparser-style: [section], "key : value" or "key = value", '#'/';' comments,
// keys lower-cased, values stripped
struct Ini {
  std::map<std::string, std::map<std::string, std::string>> sec;
  static std::string strip(const std::string &s) {
    size_t a = s.find_first_not_of(" \t\r\n"), b = s.find_last_not_of(" \t\r\n");
    return a == std::string::npos ? std::string() : s.substr(a, b - a + 1);
  }
  explicit Ini(const std::string &path) {
    std::ifstream in(path);
    if (!in) die("cannot open config " + path);
    std::string line, cur;
    while (std::getline(in, line)) {
      std::string t = strip(line);
      if (t.empty() || t[0] == '#' || t[0] == ';') continue;
      if (t[0] == '[') {
        cur = strip(t.substr(1, t.find(']') - 1));
        sec[cur];
        continue;
      }
      size_t p = t.find_first_of(":=");
      if (p == std::string::npos) continue;
      std::string k = strip(t.substr(0, p));
      for (auto &ch : k) ch = (char)std::tolower((unsigned char)ch);
      sec[cur][k] = strip(t.substr(p + 1));
    }
  }
  bool has(const std::string &s, const std::string &k) const {
    auto it = sec.find(s);
    return it != sec.end() && it->second.count(k);
  }
  std::string get(const std::string &s, const std::string &k) const {
    auto it = sec.find(s);
    if (it == sec.end() || !it->second.count(k)) die("config: missing [" + s + "] " + k);
    return it->second.at(k);
  }
  // float(config.get(...)) of Python: the whole value must be a number
  double num(const std::string &s, const std::string &k) const {
    std::string v = get(s, k);
    char *end = nullptr;
    double r = std::strtod(v.c_str(), &end);
    if (v.empty() || !end || *end != '\0') die("config: [" + s + "] " + k + " = '" + v + "' is not a number");
    return r;
  }
  // int(config.get(...)) of Python: integer literal only
  long integer(const std::string &s, const std::string &k) const {
    std::string v = get(s, k);
    char *end = nullptr;
    long r = std::strtol(v.c_str(), &end, 10);
    if (!end || *end != '\0') die("config: [" + s + "] " + k + " = '" + v + "' is not an integer");
    return r;
  }
};

// ------------------------------------------------------------------ npz
inline uint32_t crc32(const uint8_t *d, size_t n) {
  static uint32_t table[256];
  static bool init = false;
  if (!init) {
    for (uint32_t i = 0; i < 256; i++) {
      uint32_t c = i;
      for (int k = 0; k < 8; k++) c = (c & 1) ? 0xEDB88320u ^ (c >> 1) : c >> 1;
      table[i] = c;
    }
    init = true;
  }
  uint32_t c = 0xFFFFFFFFu;
  for (size_t i = 0; i < n; i++) c = table[(c ^ d[i]) & 0xFF] ^ (c >> 8);
  return c ^ 0xFFFFFFFFu;
}

struct NpyArray {
  std::string descr;          // "<f8" or "<i8"
  std::vector<size_t> shape;  // C order
  std::vector<uint8_t> bytes;
};

inline NpyArray npy_f64(const std::vector<double> &v, std::vector<size_t> shape = {}) {
  NpyArray a;
  a.descr = "<f8";
  a.shape = shape.empty() ? std::vector<size_t>{v.size()} : shape;
  a.bytes.resize(v.size() * 8);
  std::memcpy(a.bytes.data(), v.data(), a.bytes.size());
  return a;
}

inline std::vector<uint8_t> npy_encode(const NpyArray &a) {
  std::string shp = "(";
  for (size_t i = 0; i < a.shape.size(); i++) shp += std::to_string(a.shape[i]) + (a.shape.size() == 1 ? "," : (i + 1 < a.shape.size() ? ", " : ""));
  shp += ")";
  std::string hdr = "{'descr': '" + a.descr + "', 'fortran_order': False, 'shape': " + shp + ", }";
  size_t total = 10 + hdr.size() + 1;
  size_t pad = (64 - total % 64) % 64;
  hdr += std::string(pad, ' ') + "\n";
  std::vector<uint8_t> out = {0x93, 'N', 'U', 'M', 'P', 'Y', 1, 0};
  uint16_t hl = (uint16_t)hdr.size();
  out.push_back(hl & 0xFF);
  out.push_back(hl >> 8);
  out.insert(out.end(), hdr.begin(), hdr.end());
  out.insert(out.end(), a.bytes.begin(), a.bytes.end());
  return out;
}

inline void put16(std::vector<uint8_t> &b, uint32_t v) {
  b.push_back(v & 0xFF);
  b.push_back((v >> 8) & 0xFF);
}
inline void put32(std::vector<uint8_t> &b, uint32_t v) {
  put16(b, v & 0xFFFF);
  put16(b, v >> 16);
}

// np.savez-compatible archive: one stored .npy member per key
inline void write_npz(const std::string &path, const std::vector<std::pair<std::string, NpyArray>> &arrays) {
  std::vector<uint8_t> file, central;
  for (auto &kv : arrays) {
    std::string name = kv.first + ".npy";
    std::vector<uint8_t> data = npy_encode(kv.second);
    uint32_t crc = crc32(data.data(), data.size());
    uint32_t off = (uint32_t)file.size();
    put32(file, 0x04034b50);
    put16(file, 20);
    put16(file, 0);
    put16(file, 0);
    put16(file, 0);
    put16(file, 0x21);  // 1980-01-01
    put32(file, crc);
    put32(file, (uint32_t)data.size());
    put32(file, (uint32_t)data.size());
    put16(file, (uint32_t)name.size());
    put16(file, 0);
    file.insert(file.end(), name.begin(), name.end());
    file.insert(file.end(), data.begin(), data.end());
    put32(central, 0x02014b50);
    put16(central, 20);
    put16(central, 20);
    put16(central, 0);
    put16(central, 0);
    put16(central, 0);
    put16(central, 0x21);
    put32(central, crc);
    put32(central, (uint32_t)data.size());
    put32(central, (uint32_t)data.size());
    put16(central, (uint32_t)name.size());
    put16(central, 0);
    put16(central, 0);
    put16(central, 0);
    put16(central, 0);
    put32(central, 0);
    put32(central, off);
    central.insert(central.end(), name.begin(), name.end());
  }
  uint32_t cd_off = (uint32_t)file.size();
  file.insert(file.end(), central.begin(), central.end());
  put32(file, 0x06054b50);
  put16(file, 0);
  put16(file, 0);
  put16(file, (uint32_t)arrays.size());
  put16(file, (uint32_t)arrays.size());
  put32(file, (uint32_t)central.size());
  put32(file, cd_off);
  put16(file, 0);
  write_bin(path, file.data(), file.size());
}

// read the stored (uncompressed) members of an .npz as float64 arrays
inline std::map<std::string, std::vector<double>> read_npz_f64(const std::string &path) {
  std::vector<uint8_t> b = read_bin_all<uint8_t>(path);
  std::map<std::string, std::vector<double>> out;
  size_t p = 0;
  auto u16 = [&](size_t q) { return (uint32_t)b[q] | ((uint32_t)b[q + 1] << 8); };
  auto u32 = [&](size_t q) { return u16(q) | (u16(q + 2) << 16); };
  while (p + 30 <= b.size() && u32(p) == 0x04034b50) {
    uint32_t method = u16(p + 8), csize = u32(p + 18);
    uint32_t nlen = u16(p + 26), xlen = u16(p + 28);
    std::string name((const char *)&b[p + 30], nlen);
    size_t data = p + 30 + nlen + xlen;
    if (method != 0) die(path + ": compressed npz members are not supported");
    if (b[data] != 0x93 || std::memcmp(&b[data + 1], "NUMPY", 5) != 0) die(path + ": bad npy member " + name);
    uint32_t hl = b[data + 6] == 1 ? u16(data + 8) : u32(data + 8);
    size_t hstart = data + (b[data + 6] == 1 ? 10 : 12);
    std::string hdr((const char *)&b[hstart], hl);
    size_t body = hstart + hl, nbytes = csize - (body - data);
    std::vector<double> v;
    if (hdr.find("'<f8'") != std::string::npos) {
      v.resize(nbytes / 8);
      std::memcpy(v.data(), &b[body], v.size() * 8);
    } else if (hdr.find("'<i8'") != std::string::npos) {
      std::vector<int64_t> t(nbytes / 8);
      std::memcpy(t.data(), &b[body], t.size() * 8);
      v.assign(t.begin(), t.end());
    } else {
      die(path + ": unsupported dtype in " + name);
    }
    if (name.size() > 4 && name.substr(name.size() - 4) == ".npy") name = name.substr(0, name.size() - 4);
    out[name] = v;
    p = body + nbytes;
  }
  return out;
}

// ------------------------------------------------------------------ derived parameters
// least squares polynomial fit, highest power first (numpy.polyfit), by
// Householder QR of the Vandermonde matrix
inline std::vector<double> polyfit(const std::vector<double> &x, const std::vector<double> &y, int deg) {
  int n = (int)x.size(), p = deg + 1;
  std::vector<std::vector<double>> A(n, std::vector<double>(p));
  std::vector<double> b = y, coef(p);
  for (int i = 0; i < n; i++)
    for (int j = 0; j < p; j++) A[i][j] = std::pow(x[i], deg - j);
  for (int j = 0; j < p; j++) {
    double nrm = 0;
    for (int i = j; i < n; i++) nrm += A[i][j] * A[i][j];
    nrm = std::sqrt(nrm);
    double alpha = A[j][j] > 0 ? -nrm : nrm;
    std::vector<double> v(n, 0.0);
    v[j] = A[j][j] - alpha;
    for (int i = j + 1; i < n; i++) v[i] = A[i][j];
    double vv = 0;
    for (int i = j; i < n; i++) vv += v[i] * v[i];
    if (vv == 0) continue;
    for (int k = j; k < p; k++) {
      double s = 0;
      for (int i = j; i < n; i++) s += v[i] * A[i][k];
      s = 2 * s / vv;
      for (int i = j; i < n; i++) A[i][k] -= s * v[i];
    }
    double s = 0;
    for (int i = j; i < n; i++) s += v[i] * b[i];
    s = 2 * s / vv;
    for (int i = j; i < n; i++) b[i] -= s * v[i];
  }
  for (int j = p - 1; j >= 0; j--) {
    double s = b[j];
    for (int k = j + 1; k < p; k++) s -= A[j][k] * coef[k];
    coef[j] = s / A[j][j];
  }
  return coef;
}

// LB_boundary + LB_params, run_all_SOAP.py:8-34,140-145: [LB0..LB3], constant first
inline std::vector<double> lb_params(int t_star, const std::vector<double> &g2, const std::vector<double> &k0,
                                     const std::vector<double> &m0, const std::vector<double> &temps) {
  int b0 = (t_star >= 5240) ? 0 : 1, b1 = b0 + 1;
  std::vector<double> mu(10), lb(10);
  for (int k = 0; k < 10; k++) mu[k] = (k == 9) ? 1.0 : 0.1 + k * (0.9 / 9);
  for (int id = 0; id < 10; id++) {
    double yv[3] = {g2[id], k0[id], m0[id]};
    std::vector<double> c = polyfit({temps[b0], temps[b1]}, {yv[b0], yv[b1]}, 1);
    lb[id] = c[0] * t_star + c[1];
  }
  std::vector<double> c3 = polyfit(mu, lb, 3);
  return {c3[3], c3[2], c3[1], c3[0]};
}

// c / R / (2 sqrt(2 ln 2)), run_all_SOAP.py:63-65
inline double sigma_inst(double resolution) { return 299792458. / resolution / (2 * std::sqrt(2 * std::log(2.0))); }

// numpy.arange(start, stop, step) for floats
inline std::vector<double> arange(double start, double stop, double step) {
  std::vector<double> v;
  long n = (long)std::ceil((stop - start) / step);
  for (long i = 0; i < n; i++) v.push_back(start + (double)i * step);
  return v;
}

// python str(float) of the reference's argv/file naming is not needed here: ids are ints
}  // namespace soaphost
