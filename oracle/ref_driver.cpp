// Test-infrastructure driver around the reference's host-only helper sources.
//
// It is linked (by oracle/build_ref.sh) against four files compiled straight
// from the reference tree, never copied:
//   cpp-src/Utilities/BSpline.cpp    B-spline weights (host twin of Splines.cu:11-74)
//   cpp-src/Utilities/BSpmod.cpp     modulus tables lambda^2/|b(k)|^2 (BSpmod.cpp:45-129)
//   cpp-src/Utilities/Scattering.cpp Cromer-Mann coefficient table (Scattering.cpp:9-189)
//   cpp-src/System/Cell.cpp          box -> inverse matrix oc (Cell.cpp:111-126)
//   cpp-src/System/AtomCounter.cpp   explicit-solvent atoms per grid point
//                                    (AtomCounter.cpp:56-103)
// The output is plain text read by tests/test_oracle_vs_ref.py to pin the
// FP64 oracle and the engine's host tables.  Nothing in the product links it.
//
// usage:
//   ref_driver bspline ORDER W...        -> per W: "x0 .. x{p-1} | dx0 .. dx{p-1}"
//   ref_driver bspmod ORDER NX NY NZ     -> three lines: modX, modY, modZ
//   ref_driver scatter SYM...            -> per SYM: 9 coefficients
//   ref_driver cell B00 B01 .. B22       -> 9 entries of oc, row major
//   ref_driver atomcount LX LY LZ MODEL GX GY GZ [ION COUNT]...
//                                        -> "SYM value" per line, or "error"
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

#include "AtomCounter.h"
#include "BSpline.h"
#include "BSpmod.h"
#include "Cell.h"
#include "Scattering.h"

static void print_vec(const std::vector<float>& v) {
  for (size_t i = 0; i < v.size(); ++i) std::printf(i ? " %.9e" : "%.9e", (double)v[i]);
}

int main(int argc, char** argv) {
  if (argc < 2) { std::fprintf(stderr, "missing mode\n"); return 2; }
  std::string mode = argv[1];
  if (mode == "bspline" && argc >= 4) {
    int order = std::atoi(argv[2]);
    BSpline::BSpline::SetOrder(order);
    for (int i = 3; i < argc; ++i) {
      BSpline::BSpline b;
      float w = std::strtof(argv[i], nullptr);
      BSpline::spline& s = b(w);
      print_vec(s.x);
      std::printf(" |");
      for (float d : s.dx) std::printf(" %.9e", (double)d);
      std::printf("\n");
    }
    return 0;
  }
  if (mode == "bspmod" && argc == 6) {
    BSpline::BSpline::SetOrder(std::atoi(argv[2]));
    BSpline::BSpmod m(std::atoi(argv[3]), std::atoi(argv[4]), std::atoi(argv[5]));
    print_vec(m.ModX()); std::printf("\n");
    print_vec(m.ModY()); std::printf("\n");
    print_vec(m.ModZ()); std::printf("\n");
    return 0;
  }
  if (mode == "scatter") {
    for (int i = 2; i < argc; ++i) {
      print_vec(Scattering::getScattering(argv[i]));
      std::printf("\n");
    }
    return 0;
  }
  if (mode == "cell" && argc == 11) {
    std::vector<std::vector<float>> box(3, std::vector<float>(3));
    for (int i = 0; i < 9; ++i) box[i / 3][i % 3] = std::strtof(argv[2 + i], nullptr);
    Cell::calculateMatrices(box);
    const auto& oc = Cell::getOC();
    for (int i = 0; i < 9; ++i) std::printf(i ? " %.9e" : "%.9e", (double)oc[i / 3][i % 3]);
    std::printf("\n");
    return 0;
  }
  if (mode == "atomcount" && argc >= 9 && (argc - 9) % 2 == 0) {
    std::map<std::string, int> ions;
    for (int i = 9; i + 1 < argc; i += 2) ions[argv[i]] = std::atoi(argv[i + 1]);
    AtomCounter ac(std::strtof(argv[2], nullptr), std::strtof(argv[3], nullptr),
                   std::strtof(argv[4], nullptr), ions, argv[5], std::atoi(argv[6]),
                   std::atoi(argv[7]), std::atoi(argv[8]));
    try {
      for (const auto& kv : ac.calculateAtomCounts())
        std::printf("%s %.9e\n", kv.first.c_str(), (double)kv.second);
    } catch (const std::exception&) {
      std::printf("error\n");
    }
    return 0;
  }
  std::fprintf(stderr, "bad arguments\n");
  return 2;
}
