// Writes and reads the tree's JSON file with C++ iostreams, to check b12x's files against what
// B12's stream helpers produce (helpFunctions.hpp:101-216, ImplicitBoxTree.hpp:436-524 save and
// :260-309 load). Those helpers put every element with `os << value` and read it back with
// `iss >> value`. For uint8_t elements (sdScheme, depthVector, flagsVector: TypeDefinitions.h:19,37,
// Flags.h:9) these are the character overloads of operator<< and operator>>: one raw byte per
// element is written, and one character per comma-separated token is read. Test program only;
// standalone (standard library), no b12x code.
//
//   json_stream write  < description  > file.json
//     description: dim, centre (dim), radius (dim), n_sd and the scheme, n, bits (n), depths (n),
//     flags (n), nHostBufferBytes, all as whitespace-separated decimal numbers
//   json_stream read NAME < file.json
//     prints the decimal values of the uint8 array NAME as the stream extraction reads them
#include <cstdint>
#include <iomanip>
#include <iostream>
#include <limits>
#include <sstream>
#include <string>
#include <vector>

namespace {

template <class T>
void putVector(std::ostream& os, const char* name, const std::vector<T>& v) {
  if (std::is_floating_point<T>::value) os << std::setprecision(std::numeric_limits<T>::max_digits10);
  os << "\n  \"" << name << "\": [";
  for (size_t i = 0; i < v.size(); ++i) os << (i ? "," : "") << "\n    " << v[i];
  os << "\n  ]";
}

template <class T>
void putNumber(std::ostream& os, const char* name, T v) {
  os << "\n  \"" << name << "\": " << v;
}

template <class T>
std::vector<T> readN(std::istream& is, size_t n) {
  std::vector<T> v(n);
  for (auto& x : v) {
    uint64_t w = 0;
    double d = 0.0;
    if (std::is_floating_point<T>::value) {
      is >> d;
      x = T(d);
    } else {
      is >> w;
      x = T(w);
    }
  }
  return v;
}

int write() {
  size_t dim = 0, nsd = 0, n = 0;
  std::cin >> dim;
  auto c = readN<double>(std::cin, dim), r = readN<double>(std::cin, dim);
  std::cin >> nsd;
  auto sd = readN<uint8_t>(std::cin, nsd);
  std::cin >> n;
  auto bits = readN<uint64_t>(std::cin, n);
  auto depth = readN<uint8_t>(std::cin, n), flags = readN<uint8_t>(std::cin, n);
  uint64_t host = 0;
  std::cin >> host;
  if (!std::cin) return 2;
  std::ostringstream os;
  os << "{";
  putNumber(os, "B12_VERSION_MAJOR", uint32_t(3));
  os << ",";
  putNumber(os, "B12_VERSION_MINOR", uint32_t(1));
  os << ",";
  putVector(os, "center", c);
  os << ",";
  putVector(os, "radius", r);
  os << ",";
  putVector(os, "sdScheme", sd);
  os << ",";
  putNumber(os, "nLeaves", n);
  os << ",";
  putVector(os, "bitPatternVector", bits);
  os << ",";
  putVector(os, "depthVector", depth);
  os << ",";
  putVector(os, "flagsVector", flags);
  os << ",";
  putNumber(os, "nHostBufferBytes", host);
  os << "\n}";
  std::cout << os.str();
  return 0;
}

// the array after "NAME": between the next '[' and ']', split at commas, each part trimmed of
// whitespace and quotes, one uint8_t extracted from every non-empty part
int read(const std::string& name) {
  std::stringstream ss;
  ss << std::cin.rdbuf();
  const std::string s = ss.str(), ws = " \t\n\v\f\r\"";
  const size_t at = s.find("\"" + name + "\"");
  if (at == std::string::npos) return 2;
  const size_t a = s.find('[', at), b = s.find(']', at);
  std::istringstream iss(s.substr(a + 1, b - a - 1));
  std::string part;
  while (std::getline(iss, part, ',')) {
    const size_t p = part.find_first_not_of(ws);
    if (p == std::string::npos) continue;
    const size_t q = part.find_first_of(ws, p);
    std::istringstream one(part.substr(p, q == std::string::npos ? std::string::npos : q - p));
    uint8_t v = 0;
    one >> v;
    std::cout << unsigned(v) << "\n";
  }
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  const std::string mode = argc > 1 ? argv[1] : "";
  if (mode == "write") return write();
  if (mode == "read" && argc > 2) return read(argv[2]);
  std::cerr << "usage: json_stream write | read NAME\n";
  return 2;
}
