"""Writes tests/golden/uniform_grid.json: draws of std::mt19937(seed + i) followed by
uniform_real_distribution<double>(-1, 1) from this container's libstdc++ (GCC 11.4), the library
routines B12's host UniformGrid calls (Grids.hpp:152-158). Run: python tests/golden/make_golden.py
"""
import json
import os
import subprocess
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = r"""
#include <cstdio>
#include <random>
int main() {
  const unsigned long long seeds[] = {0ull, 1ull, 12345ull, 4294967295ull, 4294967296ull, 9007199254740993ull};
  for (unsigned long long s : seeds) {
    for (unsigned long long i : {0ull, 1ull, 2ull, 63ull, 64ull, 1000003ull}) {
      std::mt19937 gen(s + i);
      std::uniform_real_distribution<double> dis(-1.0, 1.0);
      std::printf("%llu %llu", s, i);
      for (int j = 0; j < 4; ++j) std::printf(" %a", dis(gen));
      std::printf("\n");
    }
  }
  std::mt19937 d;  // default seed 5489
  unsigned v = 0;
  for (int i = 0; i < 10000; ++i) v = d();
  std::printf("KAT %u\n", v);
}
"""


def main():
    with tempfile.TemporaryDirectory() as tmp:
        src = os.path.join(tmp, "g.cpp")
        exe = os.path.join(tmp, "g")
        open(src, "w").write(SRC)
        subprocess.run(["g++", "-O2", "-std=c++17", "-o", exe, src], check=True)
        out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout
    rows, kat = [], None
    for line in out.splitlines():
        p = line.split()
        if p[0] == "KAT":
            kat = int(p[1])
        else:
            rows.append({"seed": int(p[0]), "i": int(p[1]), "draws_hex": p[2:]})
    json.dump({"generator": "g++ 11.4 libstdc++ std::mt19937 + uniform_real_distribution<double>(-1,1)",
               "rows": rows, "mt19937_default_10000th": kat}, open(os.path.join(HERE, "uniform_grid.json"), "w"),
              indent=1)


if __name__ == "__main__":
    main()
