// Corrupted-input driver for the trajectory reader (tests/test_trajectory.py
// builds it with AddressSanitizer and UndefinedBehaviorSanitizer, host code
// only): a valid file's bytes with random bit flips, random bytes or a
// truncation are opened, indexed and decoded; errors are expected, memory
// errors are not.  Usage: xdr_fuzz <seed file> <cases> <scratch file>.
#include "saxs_xdr.h"
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>
#include <string>
#include <fstream>
#include <iterator>
int main(int argc, char** argv) {
  std::ifstream in(argv[1], std::ios::binary);
  std::vector<unsigned char> seed((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
  const int iters = atoi(argv[2]);
  std::mt19937 g(42);
  int opened = 0, read_ok = 0;
  for (int it = 0; it < iters; ++it) {
    std::vector<unsigned char> d = seed;
    const int mode = it % 3;
    if (mode == 0) { for (int k = 0; k < 1 + (int)(g() % 8); ++k) d[g() % d.size()] ^= (unsigned char)(1u << (g() % 8)); }
    else if (mode == 1) { d.resize(g() % d.size() + 1); }
    else { d[g() % d.size()] = (unsigned char)g(); d[g() % d.size()] = (unsigned char)g(); }
    const std::string path = argv[3];
    FILE* f = fopen(path.c_str(), "wb"); fwrite(d.data(), 1, d.size(), f); fclose(f);
    saxs_xdr* h = nullptr;
    if (saxs_xdr_open(path.c_str(), &h) != 0) continue;
    ++opened;
    int kind, na; int64_t nf;
    saxs_xdr_info(h, &kind, &na, &nf);
    if (na > 0 && nf > 0 && (long long)na * nf <= 2000000) {
      std::vector<int64_t> fr(nf); for (int64_t k = 0; k < nf; ++k) fr[k] = k;
      std::vector<float> xyz((size_t)nf * na * 3), box(nf * 9); std::vector<double> t(nf); std::vector<int64_t> st(nf);
      if (saxs_xdr_read(h, fr.data(), (int)nf, xyz.data(), box.data(), t.data(), st.data(), 1 + it % 4) == 0) ++read_ok;
    }
    saxs_xdr_close(h);
  }
  printf("%d cases, %d opened, %d decoded without error\n", iters, opened, read_ok);
  return 0;
}
