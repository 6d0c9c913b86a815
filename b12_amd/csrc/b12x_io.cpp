// Files of the reference's formats, written and read on the host through the C ABI:
//  - the tree's JSON file (ImplicitBoxTree.hpp:435-524 save, :142-339 the loading constructor)
//    and the legacy binary file of B12a/B12b it also reads (:174-257);
//  - GAIO's binary box list (ImplicitBoxTree.hpp:526-569 saveForGaio);
//  - the matrix's JSON file (CooMatrix.hpp:473-534 save, :235-284 the loading constructor).
// The writers reproduce the reference's bytes: a "{", then per entry a newline, two spaces, the
// quoted name, ": " and the value (arrays: "[", each element on its own line indented by four
// spaces, separated by commas, "]" on its own line indented by two), entries separated by
// commas, and "\n}" at the end. Doubles are printed as iostreams print them at
// max_digits10 = 17 significant digits (%.17g), integers in decimal, and the tree's uint8 arrays
// (sdScheme, depthVector, flagsVector: Depth, Dimension and Flags are uint8_t, TypeDefinitions.h:
// 19,37, Flags.h:9) as one raw byte per element, which is what `os << value` inserts for a
// uint8_t (helpFunctions.hpp:128). The readers follow the reference's tolerant scan: the value
// after `"name"` and a colon, an array as the comma-separated tokens between the next '[' and
// ']' (so the matrix file's trailing comma is fine), each token trimmed of whitespace and quotes;
// a uint8 element is the first character of its token, as `iss >> value` extracts it
// (helpFunctions.hpp:197-205).
// That reader loses elements whose byte is whitespace (9-13, 32) or a quote (34), splits at a
// comma byte (44) and ends the array at a ']' byte (93), so B12 cannot read back such trees
// (depths 9-13, 32, 34 or 44; flags such as SD|HIT = 9); the reference then indexes its shorter
// arrays past their end. Here the loader reads the same way and reports the short arrays as an
// error instead.
#include <cerrno>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "b12x_host.hpp"

namespace {

using b12x::HipFail;

constexpr uint32_t kVersionMajor = 3;  // 'c' - 'a' + 1 (TypeDefinitions.h:11-12)
constexpr uint32_t kVersionMinor = 1;
constexpr uint64_t kHostBufferBytes = uint64_t(4) << 30;  // ImplicitBoxTree.hpp:38 default
constexpr int kMaxDepthN = 64;                            // getMaxDepth<64>()

void check(int rc) {
  if (rc != B12X_OK) throw HipFail(rc, b12x::lastError());
}

std::string fmtDouble(double v) {
  char buf[64];
  std::snprintf(buf, sizeof(buf), "%.17g", v);
  return buf;
}

void putName(std::ostream& os, const char* name) { os << "\n  \"" << name << "\": "; }

template <class T>
void putNumber(std::ostream& os, const char* name, T v) {
  putName(os, name);
  os << v;
}

template <class T, class F>
void putVector(std::ostream& os, const char* name, const std::vector<T>& v, F fmt) {
  putName(os, name);
  os << "[";
  for (size_t i = 0; i < v.size(); ++i) {
    if (i) os << ",";
    os << "\n    " << fmt(v[i]);
  }
  os << "\n  ]";
}

bool endsWithJson(const std::string& s) { return s.size() >= 5 && s.compare(s.size() - 5, 5, ".json") == 0; }

// position just after the colon that follows "name", or npos
size_t valueStart(const std::string& str, const std::string& name) {
  const std::string key = "\"" + name + "\"";
  size_t pos = 0;
  for (;;) {
    pos = str.find(key, pos);
    if (pos == std::string::npos) return pos;
    pos += key.size();
    const size_t q = str.find_first_not_of(" \t\n\v\f\r", pos);
    if (q != std::string::npos && str[q] == ':') return q + 1;
  }
}

std::string token(const std::string& s) {
  const char* ws = " \t\n\v\f\r\"";
  const size_t a = s.find_first_not_of(ws);
  if (a == std::string::npos) return "";
  const size_t b = s.find_first_of(std::string(ws) + ",", a);
  return s.substr(a, b == std::string::npos ? std::string::npos : b - a);
}

uint64_t parseU64(const std::string& t, const std::string& name) {
  if (t.empty() || t[0] == '-') throw HipFail(B12X_E_IO, "reading \"" + name + "\": bad value '" + t + "'");
  char* end = nullptr;
  errno = 0;
  const unsigned long long v = std::strtoull(t.c_str(), &end, 10);
  if (errno || end == t.c_str()) throw HipFail(B12X_E_IO, "reading \"" + name + "\": bad value '" + t + "'");
  return v;
}

double parseDouble(const std::string& t, const std::string& name) {
  char* end = nullptr;
  const double v = std::strtod(t.c_str(), &end);
  if (end == t.c_str()) throw HipFail(B12X_E_IO, "reading \"" + name + "\": bad value '" + t + "'");
  return v;
}

std::string scalarToken(const std::string& str, const std::string& name) {
  const size_t p = valueStart(str, name);
  if (p == std::string::npos) throw HipFail(B12X_E_IO, "\"" + name + "\" could not be found");
  return token(str.substr(p, 64));
}

std::vector<std::string> arrayTokens(const std::string& str, const std::string& name) {
  const size_t p = valueStart(str, name);
  const size_t a = p == std::string::npos ? p : str.find('[', p);
  const size_t b = a == std::string::npos ? a : str.find(']', a);
  if (a == std::string::npos || b == std::string::npos)
    throw HipFail(B12X_E_IO, "\"" + name + "\" could not be found as an array");
  std::vector<std::string> out;
  std::istringstream iss(str.substr(a + 1, b - a - 1));
  std::string part;
  while (std::getline(iss, part, ',')) {
    std::string t = token(part);
    if (!t.empty()) out.push_back(t);
  }
  return out;
}

std::vector<uint64_t> u64Array(const std::string& str, const std::string& name) {
  std::vector<uint64_t> v;
  for (const auto& t : arrayTokens(str, name)) v.push_back(parseU64(t, name));
  return v;
}

// a uint8 array: the first character of every non-empty token
std::vector<uint8_t> byteArray(const std::string& str, const std::string& name) {
  std::vector<uint8_t> v;
  for (const auto& t : arrayTokens(str, name)) v.push_back(uint8_t(t[0]));
  return v;
}

std::vector<double> doubleArray(const std::string& str, const std::string& name) {
  std::vector<double> v;
  for (const auto& t : arrayTokens(str, name)) v.push_back(parseDouble(t, name));
  return v;
}

struct Leaves {
  std::vector<uint64_t> bits;
  std::vector<uint8_t> depth, flags;
};

// ImplicitBoxTree.hpp:297-330: depths above 64 truncated to 64, then equal consecutive boxes
// merged, the first keeping the AND of their flags
void mergeEqual(Leaves& L) {
  Leaves out;
  for (size_t i = 0; i < L.bits.size(); ++i) {
    const bool same = !out.bits.empty() && out.bits.back() == L.bits[i] && out.depth.back() == L.depth[i];
    if (same) {
      out.flags.back() &= L.flags[i];
    } else {
      out.bits.push_back(L.bits[i]);
      out.depth.push_back(L.depth[i]);
      out.flags.push_back(L.flags[i]);
    }
  }
  L = std::move(out);
}

template <class T>
bool readRaw(std::istream& is, T* p, size_t n) {
  is.read(reinterpret_cast<char*>(p), std::streamsize(n * sizeof(T)));
  return bool(is);
}

}  // namespace

extern "C" {

int b12x_tree_save_json(b12x_tree t, const char* fileName) {
  return b12x::guarded([&]() -> int {
    if (!t || !fileName) return b12x::setErr(B12X_E_ARG, "null argument");
    int dim = 0, nsd = 0;
    double c[8], r[8];
    uint8_t sd[kMaxDepthN + 1];
    check(b12x_tree_geometry(t, &dim, c, r, sd, &nsd));
    uint32_t n = 0;
    check(b12x_size(t, &n));
    std::vector<uint64_t> bits(n);
    std::vector<uint8_t> depth(n), flags(n);
    check(b12x_tree_download(t, bits.data(), depth.data(), flags.data()));
    std::string name(fileName);
    if (!endsWithJson(name)) name += ".json";
    std::ofstream ofs(name, std::ofstream::out | std::ofstream::trunc);
    if (!ofs.good()) return b12x::setErr(B12X_E_IO, "file could not be opened: " + name);
    auto asInt = [](auto v) { return std::to_string(uint64_t(v)); };
    auto asByte = [](uint8_t v) { return std::string(1, char(v)); };
    auto asDouble = [](double v) { return fmtDouble(v); };
    ofs << "{";
    putNumber(ofs, "B12_VERSION_MAJOR", kVersionMajor);
    ofs << ",";
    putNumber(ofs, "B12_VERSION_MINOR", kVersionMinor);
    ofs << ",";
    putVector(ofs, "center", std::vector<double>(c, c + dim), asDouble);
    ofs << ",";
    putVector(ofs, "radius", std::vector<double>(r, r + dim), asDouble);
    ofs << ",";
    putVector(ofs, "sdScheme", std::vector<uint8_t>(sd, sd + nsd), asByte);
    ofs << ",";
    putNumber(ofs, "nLeaves", uint64_t(n));
    ofs << ",";
    putVector(ofs, "bitPatternVector", bits, asInt);
    ofs << ",";
    putVector(ofs, "depthVector", depth, asByte);
    ofs << ",";
    putVector(ofs, "flagsVector", flags, asByte);
    ofs << ",";
    uint64_t hostBytes = kHostBufferBytes;
    check(b12x_get_host_buffer_bytes(t, &hostBytes));
    putNumber(ofs, "nHostBufferBytes", hostBytes);
    ofs << "\n}";
    if (!ofs.good()) return b12x::setErr(B12X_E_IO, "an error occurred during writing: " + name);
    return B12X_OK;
  });
}

int b12x_tree_save_gaio(b12x_tree t, const char* fileName) {
  return b12x::guarded([&]() -> int {
    if (!t || !fileName) return b12x::setErr(B12X_E_ARG, "null argument");
    int dim = 0, nsd = 0;
    double c[8], r[8];
    uint8_t sd[kMaxDepthN + 1];
    check(b12x_tree_geometry(t, &dim, c, r, sd, &nsd));
    uint32_t n = 0;
    check(b12x_size(t, &n));
    std::vector<uint64_t> bits(n);
    std::vector<uint8_t> depth(n), flags(n);
    check(b12x_tree_download(t, bits.data(), depth.data(), flags.data()));
    uint64_t nb = 0;
    check(b12x_boxes(t, B12X_LEAF_DEPTH, nullptr, &nb));
    std::vector<double> rows(nb * uint64_t(2 * dim + 1));
    if (nb) check(b12x_boxes(t, B12X_LEAF_DEPTH, rows.data(), &nb));
    std::ofstream f(fileName, std::ofstream::binary | std::ofstream::trunc);
    if (!f.good()) return b12x::setErr(B12X_E_IO, std::string("file could not be opened: ") + fileName);
    const int sizeOfReal = int(sizeof(double));
    f.write(reinterpret_cast<const char*>(&dim), sizeof(int));
    f.write(reinterpret_cast<const char*>(&sizeOfReal), sizeof(int));
    f.write(reinterpret_cast<const char*>(c), std::streamsize(dim * sizeof(double)));
    f.write(reinterpret_cast<const char*>(r), std::streamsize(dim * sizeof(double)));
    for (int i = 0; i <= 500 * dim; ++i) {
      const int s = sd[i % nsd];
      f.write(reinterpret_cast<const char*>(&s), sizeof(int));
    }
    const double nBoxes = double(n);
    f.write(reinterpret_cast<const char*>(&nBoxes), sizeof(double));
    for (uint64_t i = 0; i < nb; ++i) {
      f.write(reinterpret_cast<const char*>(&rows[i * (2 * dim + 1)]), std::streamsize(dim * sizeof(double)));
      const double d = double(depth[i]), fl = double(flags[i]);
      f.write(reinterpret_cast<const char*>(&d), sizeof(double));
      f.write(reinterpret_cast<const char*>(&fl), sizeof(double));
    }
    if (!f.good()) return b12x::setErr(B12X_E_IO, std::string("an error occurred during writing: ") + fileName);
    return B12X_OK;
  });
}

int b12x_tree_load(const char* fileName, int dim, b12x_tree* out) {
  return b12x::guarded([&]() -> int {
    if (!fileName || !out) return b12x::setErr(B12X_E_ARG, "null argument");
    if (dim < 1 || dim > 4) return b12x::setErr(B12X_E_ARG, "dim must be 1..4");
    const std::string fn(fileName);
    std::ifstream ifs(fn, std::ifstream::in);
    bool legacy = false;
    if (ifs.good()) {
      legacy = !endsWithJson(fn);  // B12a / B12b files carry no ".json"
      if (legacy) {
        ifs.close();
        ifs.open(fn, std::ifstream::in | std::ifstream::binary);
      }
    } else {
      ifs.close();
      ifs.open(fn + ".json", std::ifstream::in);
    }
    if (!ifs.good()) return b12x::setErr(B12X_E_IO, "file could not be opened: " + fn);
    std::vector<double> center, radius;
    std::vector<uint8_t> sd;
    Leaves L;
    uint64_t hostBytes = kHostBufferBytes;
    if (legacy) {  // ImplicitBoxTree.hpp:174-257
      uint32_t vmaj = 0, vmin = 0;
      uint64_t fdim = 0, n = 0;
      bool ok = readRaw(ifs, &vmaj, 1) && readRaw(ifs, &vmin, 1) && readRaw(ifs, &fdim, 1);
      if (ok && fdim != uint64_t(dim)) return b12x::setErr(B12X_E_IO, "file holds a tree of another dimension");
      center.resize(dim);
      radius.resize(dim);
      sd.resize(kMaxDepthN + 1);
      ok = ok && readRaw(ifs, center.data(), dim) && readRaw(ifs, radius.data(), dim) &&
           readRaw(ifs, sd.data(), sd.size()) && readRaw(ifs, &n, 1);
      if (ok) {
        std::vector<uint64_t> wide(2 * n);  // BitPattern<128>: high word, then low word
        L.depth.resize(n);
        L.flags.resize(n);
        ok = readRaw(ifs, wide.data(), wide.size()) && readRaw(ifs, L.depth.data(), n) &&
             readRaw(ifs, L.flags.data(), n) && readRaw(ifs, &hostBytes, 1);
        L.bits.resize(n);
        for (uint64_t i = 0; ok && i < n; ++i) {
          const uint64_t hi = wide[2 * i], lo = wide[2 * i + 1];
          const int d = L.depth[i];
          if (d > kMaxDepthN) {  // DecrementDepthToFunctor(64): keep the top 64 bits
            const int sh = d - kMaxDepthN;
            L.bits[i] = sh >= 64 ? (hi >> (sh - 64)) : ((lo >> sh) | (sh ? hi << (64 - sh) : 0));
            L.depth[i] = uint8_t(kMaxDepthN);
          } else {
            L.bits[i] = lo;
          }
        }
      }
      if (!ok) return b12x::setErr(B12X_E_IO, "an error occurred during reading: " + fn);
    } else {
      std::stringstream ss;
      ss << ifs.rdbuf();
      const std::string str = ss.str();
      center = doubleArray(str, "center");
      radius = doubleArray(str, "radius");
      if (center.size() < size_t(dim) || radius.size() < size_t(dim))
        return b12x::setErr(B12X_E_IO, "centre or radius shorter than the dimension");
      center.resize(dim);
      radius.resize(dim);
      sd = byteArray(str, "sdScheme");
      const uint64_t n = parseU64(scalarToken(str, "nLeaves"), "nLeaves");
      L.bits = u64Array(str, "bitPatternVector");
      L.depth = byteArray(str, "depthVector");
      L.flags = byteArray(str, "flagsVector");
      if (L.bits.size() != n)
        return b12x::setErr(B12X_E_IO, "nLeaves does not match the leaf arrays");
      if (L.depth.size() != n || L.flags.size() != n)
        return b12x::setErr(B12X_E_IO,
                            "depthVector or flagsVector holds fewer entries than nLeaves: the reference's "
                            "reader drops bytes 9-13, 32 and 34, splits at 44 and stops at 93");
      if (valueStart(str, "nHostBufferBytes") != std::string::npos)
        hostBytes = parseU64(scalarToken(str, "nHostBufferBytes"), "nHostBufferBytes");
      for (uint64_t i = 0; i < n; ++i)
        if (L.depth[i] > kMaxDepthN) {
          L.bits[i] >>= (L.depth[i] - kMaxDepthN);
          L.depth[i] = uint8_t(kMaxDepthN);
        }
    }
    mergeEqual(L);
    b12x_tree t = nullptr;
    check(b12x_tree_create(dim, center.data(), radius.data(), sd.empty() ? nullptr : sd.data(), int(sd.size()), &t));
    int rc = b12x_tree_upload(t, uint32_t(L.bits.size()), L.bits.data(), L.depth.data(), L.flags.data());
    if (rc == B12X_OK) rc = b12x_set_host_buffer_bytes(t, hostBytes);
    if (rc != B12X_OK) {
      const std::string msg = b12x::lastError();
      b12x_tree_destroy(t);
      return b12x::setErr(rc, msg);
    }
    *out = t;
    return B12X_OK;
  });
}

int b12x_coo_save_json(b12x_coo m, const char* fileName) {
  return b12x::guarded([&]() -> int {
    if (!m || !fileName) return b12x::setErr(B12X_E_ARG, "null argument");
    uint64_t nnz = 0;
    uint32_t nr = 0, nc = 0;
    int rm = 0;
    check(b12x_coo_info(m, &nnz, &nr, &nc, &rm));
    std::vector<uint32_t> rows(nnz), cols(nnz);
    std::vector<double> vals(nnz);
    if (nnz) check(b12x_coo_download(m, rows.data(), cols.data(), vals.data()));
    std::string name(fileName);
    if (!endsWithJson(name)) name += ".json";
    std::ofstream ofs(name, std::ofstream::out | std::ofstream::trunc);
    if (!ofs.good()) return b12x::setErr(B12X_E_IO, "file could not be opened: " + name);
    auto asInt = [](auto v) { return std::to_string(uint64_t(v)); };
    ofs << "{";
    putNumber(ofs, "nRows", nr);
    ofs << ",";
    putNumber(ofs, "nColumns", nc);
    ofs << ",";
    putNumber(ofs, "nnz", nnz);
    ofs << ",";
    putVector(ofs, "rowIndices", rows, asInt);
    ofs << ",";
    putVector(ofs, "columnIndices", cols, asInt);
    ofs << ",";
    putVector(ofs, "values", vals, [](double v) { return fmtDouble(v); });
    ofs << ",";  // the reference writes this comma too
    ofs << "\n}";
    if (!ofs.good()) return b12x::setErr(B12X_E_IO, "an error occurred during writing: " + name);
    return B12X_OK;
  });
}

int b12x_coo_load_json(const char* fileName, b12x_coo* out) {
  return b12x::guarded([&]() -> int {
    if (!fileName || !out) return b12x::setErr(B12X_E_ARG, "null argument");
    const std::string fn(fileName);
    std::ifstream ifs(fn, std::ifstream::in);
    if (!ifs.good()) {
      ifs.close();
      ifs.open(fn + ".json", std::ifstream::in);
    }
    if (!ifs.good()) return b12x::setErr(B12X_E_IO, "file could not be opened: " + fn);
    std::stringstream ss;
    ss << ifs.rdbuf();
    const std::string str = ss.str();
    const uint64_t nr = parseU64(scalarToken(str, "nRows"), "nRows");
    const uint64_t nc = parseU64(scalarToken(str, "nColumns"), "nColumns");
    const uint64_t nnz = parseU64(scalarToken(str, "nnz"), "nnz");
    std::vector<uint32_t> rows, cols;
    for (uint64_t v : u64Array(str, "rowIndices")) rows.push_back(uint32_t(v));
    for (uint64_t v : u64Array(str, "columnIndices")) cols.push_back(uint32_t(v));
    const std::vector<double> vals = doubleArray(str, "values");
    if (rows.size() != nnz || cols.size() != nnz || vals.size() != nnz)
      return b12x::setErr(B12X_E_IO, "nnz does not match the index and value arrays");
    return b12x_coo_upload(rows.data(), cols.data(), vals.data(), nnz, uint32_t(nr), uint32_t(nc), out);
  });
}

}  // extern "C"
