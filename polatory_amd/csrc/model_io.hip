// Reader and writer of polatory's binary model and interpolant files (host code).
#include "model_io.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <fstream>
#include <istream>
#include <ostream>
#include <type_traits>
#include <vector>

namespace pam {

namespace {

// Bounds that a well-formed file never exceeds; a corrupt count fails here
// instead of allocating without limit.
constexpr uint64_t kMaxName = 64, kMaxParams = 64, kMaxRbfs = 1024;

template <class T>
void put(std::ostream& os, const T& v) {
  static_assert(std::is_trivially_copyable_v<T>);
  os.write(reinterpret_cast<const char*>(&v), sizeof(T));
}

template <class T>
T get(std::istream& is) {
  static_assert(std::is_trivially_copyable_v<T>);
  T v{};
  is.read(reinterpret_cast<char*>(&v), sizeof(T));
  if (!is) throw_runtime("unexpected end of file");
  return v;
}

void get_bytes(std::istream& is, void* p, uint64_t n) {
  is.read(static_cast<char*>(p), static_cast<std::streamsize>(n));
  if (!is) throw_runtime("unexpected end of file");
}

// Bytes left in the stream (for seekable streams), to check counts against.
uint64_t remaining(std::istream& is) {
  const auto pos = is.tellg();
  if (pos < 0) return UINT64_MAX;
  is.seekg(0, std::ios::end);
  const auto end = is.tellg();
  is.seekg(pos);
  return end > pos ? static_cast<uint64_t>(end - pos) : 0;
}

// Eigen::Matrix<double, Dynamic, 3, RowMajor>: rows (Eigen::Index), then the
// coefficients row by row.
void put_points(std::ostream& os, const std::vector<double>& xyz) {
  put<int64_t>(os, static_cast<int64_t>(xyz.size() / 3));
  if (!xyz.empty()) os.write(reinterpret_cast<const char*>(xyz.data()), sizeof(double) * xyz.size());
}

std::vector<double> get_points(std::istream& is) {
  const int64_t rows = get<int64_t>(is);
  if (rows < 0 || static_cast<uint64_t>(rows) > remaining(is) / 24) throw_runtime("corrupt point count");
  std::vector<double> v(static_cast<size_t>(rows) * 3);
  if (rows > 0) get_bytes(is, v.data(), sizeof(double) * v.size());
  return v;
}

const char* short_name(int kind) {
  switch (kind) {
    case kBh3: return "bh3";
    case kTh3: return "th3";
    case kExp: return "exp";
  }
  throw_invalid("unknown rbf kind");
}

// rbf/make_rbf.hpp:23-49, restricted to the kernels on this library's path.
int kind_of(const std::string& name) {
  if (name == "bh3") return kBh3;
  if (name == "th3") return kTh3;
  if (name == "exp") return kExp;
  static const char* const kOther[] = {"bh2", "th2", "cub", "gau", "gc3", "gc5", "gc7",
                                       "gc9", "sph", "sp3", "sp5", "sp7", "sp9"};
  for (const char* o : kOther)
    if (name == o)
      throw Error(kNotImplemented, "RBF '" + name + "' is not implemented by this library (bh3, th3 and exp are)");
  throw_runtime("unknown RBF name: '" + name + "'");
}

void write_rbf(std::ostream& os, const RbfDesc& r) {
  const std::string name = short_name(r.kind);
  put<uint64_t>(os, name.size());
  os.write(name.data(), static_cast<std::streamsize>(name.size()));
  put<uint64_t>(os, 2);
  put<double>(os, r.p0);
  put<double>(os, r.p1);
  for (int i = 0; i < 9; i++) put<double>(os, r.aniso[i]);
}

RbfDesc read_rbf(std::istream& is) {
  const uint64_t len = get<uint64_t>(is);
  if (len > kMaxName) throw_runtime("corrupt RBF name");
  std::string name(len, '\0');
  if (len) get_bytes(is, name.data(), len);
  const uint64_t np = get<uint64_t>(is);
  if (np > kMaxParams) throw_runtime("corrupt RBF parameter count");
  std::vector<double> params(np);
  if (np) get_bytes(is, params.data(), sizeof(double) * np);
  RbfDesc r{};
  get_bytes(is, r.aniso, sizeof(r.aniso));
  r.kind = kind_of(name);
  // RbfBase::set_parameters (rbf/rbf_base.hpp:81-87)
  if (np != 2) throw_invalid("params.size() must be 2");
  r.p0 = params[0];
  r.p1 = params[1];
  return r;
}

template <class F>
void with_file(const std::string& filename, std::ios::openmode mode, F&& f) {
  std::fstream fs(filename, mode | std::ios::binary);
  if (!fs) throw_runtime("cannot open file '" + filename + "'");
  f(fs);
  if (mode & std::ios::out) {
    fs.flush();
    if (!fs) throw_runtime("cannot write file '" + filename + "'");
  }
}

}  // namespace

void write_model(std::ostream& os, const Model& m) {
  put<uint64_t>(os, m.rbfs.size());
  for (const RbfDesc& r : m.rbfs) write_rbf(os, r);
  put<int32_t>(os, m.poly_degree);
  put<double>(os, m.nugget);
}

Model read_model(std::istream& is) {
  Model m;
  const uint64_t n = get<uint64_t>(is);
  if (n > kMaxRbfs) throw_runtime("corrupt RBF count");
  for (uint64_t i = 0; i < n; i++) m.rbfs.push_back(read_rbf(is));
  m.poly_degree = get<int32_t>(is);
  m.nugget = get<double>(is);
  m.validate();
  return m;
}

void save_model(const std::string& filename, const Model& m) {
  m.validate();
  with_file(filename, std::ios::out | std::ios::trunc, [&](std::ostream& os) { write_model(os, m); });
}

Model load_model(const std::string& filename) {
  Model m;
  with_file(filename, std::ios::in, [&](std::istream& is) { m = read_model(is); });
  return m;
}

void save_interpolant(const std::string& filename, const Interpolant& ip) {
  with_file(filename, std::ios::out | std::ios::trunc, [&](std::ostream& os) {
    write_model(os, ip.model());
    const bool fitted = ip.fitted();
    put<bool>(os, fitted);
    const std::vector<double>& c = ip.centers();
    const std::vector<double>& g = ip.grad_centers();
    put_points(os, fitted ? c : std::vector<double>{});
    put_points(os, fitted ? g : std::vector<double>{});
    // bbox_ = the box of the centres and the gradient centres
    // (interpolant.hpp:112, 190); the default Bbox (+inf, -inf) when not
    // fitted (geometry/bbox3d.hpp:18-20).
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    if (fitted) {
      for (const std::vector<double>* v : {&c, &g})
        for (size_t i = 0; i < v->size() / 3; i++)
          for (int k = 0; k < 3; k++) {
            lo[k] = std::min(lo[k], (*v)[3 * i + k]);
            hi[k] = std::max(hi[k], (*v)[3 * i + k]);
          }
    }
    for (double v : lo) put<double>(os, v);
    for (double v : hi) put<double>(os, v);
    const std::vector<double>& w = ip.weights();
    const size_t nw = fitted ? w.size() : 0;
    put<int64_t>(os, static_cast<int64_t>(nw));
    if (nw) os.write(reinterpret_cast<const char*>(w.data()), sizeof(double) * nw);
  });
}

std::unique_ptr<Interpolant> load_interpolant(const std::string& filename) {
  std::unique_ptr<Interpolant> ip;
  with_file(filename, std::ios::in, [&](std::istream& is) {
    Model m = read_model(is);
    const bool fitted = get<uint8_t>(is) != 0;
    std::vector<double> centers = get_points(is);
    std::vector<double> grad = get_points(is);
    double box[6];
    get_bytes(is, box, sizeof(box));
    const int64_t nw = get<int64_t>(is);
    if (nw < 0 || static_cast<uint64_t>(nw) > remaining(is) / 8) throw_runtime("corrupt weight count");
    std::vector<double> w(static_cast<size_t>(nw));
    if (nw > 0) get_bytes(is, w.data(), sizeof(double) * w.size());
    ip = std::make_unique<Interpolant>(m);
    if (fitted) ip->set_fitted(centers, w, grad);
  });
  return ip;
}

}  // namespace pam
