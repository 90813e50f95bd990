// expf_table_gen.cpp -- build step: list every float x in [-104, 0] where glibc expf(x) differs
// from tamc::expf_cr(x), and write the list as a C++ include for the runtime library.
//
// The reference accepts a Metropolis move iff u < f32::exp(-beta*dE), and f32::exp is glibc expf on
// Linux.  glibc expf is not correctly rounded everywhere; this table records exactly where it
// differs from the library's own double-precision evaluation, so the device can reproduce glibc's
// value bit for bit.  Every listed difference must be one ulp: the accept test relies on it.
//
// usage: expf_table_gen <output.inc>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <thread>
#include <vector>

#include "tamc_math.hpp"

int main(int argc, char **argv) {
  if (argc != 2) {
    fprintf(stderr, "usage: %s <output.inc>\n", argv[0]);
    return 2;
  }
  const uint32_t first = 0x80000000u;             // -0.0
  const uint32_t last = tamc::f2u(-104.0f);        // most negative input that can be nonzero
  unsigned nt = std::thread::hardware_concurrency();
  if (nt == 0) nt = 4;
  if (nt > 16) nt = 16;
  std::vector<std::vector<std::pair<uint32_t, uint32_t>>> found(nt);
  std::vector<int> bad(nt, 0);
  std::vector<std::thread> th;
  const uint64_t total = (uint64_t)last - first + 1;
  for (unsigned t = 0; t < nt; ++t) {
    th.emplace_back([&, t]() {
      const uint64_t b = first + total * t / nt, e = first + total * (t + 1) / nt;
      for (uint64_t u = b; u < e; ++u) {
        const float x = tamc::u2f((uint32_t)u);
        const float g = expf(x);
        const float m = tamc::expf_cr(x);
        if (tamc::f2u(g) != tamc::f2u(m)) {
          const uint32_t a = tamc::f2u(g), c = tamc::f2u(m);
          if ((a > c ? a - c : c - a) != 1u) bad[t]++;
          found[t].emplace_back((uint32_t)u, a);
        }
      }
    });
  }
  for (auto &x : th) x.join();
  std::vector<std::pair<uint32_t, uint32_t>> all;
  int nbad = 0;
  for (unsigned t = 0; t < nt; ++t) {
    all.insert(all.end(), found[t].begin(), found[t].end());
    nbad += bad[t];
  }
  std::sort(all.begin(), all.end());
  if (nbad) {
    fprintf(stderr, "expf_table_gen: %d differences exceed one ulp; the accept test needs a wider band\n", nbad);
    return 1;
  }
  FILE *f = fopen(argv[1], "w");
  if (!f) {
    perror(argv[1]);
    return 1;
  }
  fprintf(f, "// generated by expf_table_gen.cpp from this host's glibc expf; do not edit\n");
  fprintf(f, "#ifndef TAMC_EXC_DECL\n#define TAMC_EXC_DECL static const\n#endif\n");
  fprintf(f, "static constexpr uint32_t kExpfExcCount = %zuu;\n", all.size());
  fprintf(f, "TAMC_EXC_DECL uint32_t kExpfExcX[%zu] = {", all.size() ? all.size() : 1);
  for (size_t k = 0; k < all.size(); ++k) fprintf(f, "%s0x%08xu", k == 0 ? "\n" : ((k % 8) ? "," : ",\n"), all[k].first);
  if (all.empty()) fprintf(f, "0");
  fprintf(f, "};\nTAMC_EXC_DECL uint32_t kExpfExcY[%zu] = {", all.size() ? all.size() : 1);
  for (size_t k = 0; k < all.size(); ++k) fprintf(f, "%s0x%08xu", k == 0 ? "\n" : ((k % 8) ? "," : ",\n"), all[k].second);
  if (all.empty()) fprintf(f, "0");
  fprintf(f, "};\n");
  fclose(f);
  printf("expf_table_gen: %zu exceptions over %llu inputs\n", all.size(), (unsigned long long)total);
  return 0;
}
