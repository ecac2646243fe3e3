// Robustness check of the netCDF classic reader (rrtmg_sw_amd/csrc/rrtmgsw_nc.cpp) on damaged
// files, for a build with AddressSanitizer and UndefinedBehaviorSanitizer: every mutation of a
// valid file (bytes flipped, the file cut short, bytes inserted or deleted) must either be
// refused by open() or be read in full, variable by variable, without an access outside the
// file image. Usage: nc_fuzz_check <valid.nc> <work file> <mutations>
#include <cstdio>
#include <random>
#include <string>
#include <vector>

#include "rrtmgsw_nc.h"

int main(int argc, char **argv) {
  if (argc < 4) return 2;
  std::vector<unsigned char> orig;
  {
    FILE *f = std::fopen(argv[1], "rb");
    if (!f) return 2;
    int ch;
    while ((ch = std::fgetc(f)) != EOF) orig.push_back((unsigned char)ch);
    std::fclose(f);
  }
  const int n = std::atoi(argv[3]);
  std::mt19937_64 rng(12345);
  int refused = 0, read = 0;
  std::vector<double> out;
  for (int it = 0; it < n; it++) {
    std::vector<unsigned char> b = orig;
    const int kind = it % 4;
    // the header holds the structure: aim most mutations at the first kilobytes
    const size_t span = std::min<size_t>(b.size(), it % 3 ? 4096 : b.size());
    if (kind == 0) {
      for (int k = 0; k < 1 + (int)(rng() % 8); k++) b[rng() % span] ^= (unsigned char)(1u << (rng() % 8));
    } else if (kind == 1) {
      b.resize(rng() % b.size());
    } else if (kind == 2) {
      b.insert(b.begin() + rng() % span, 1 + rng() % 16, (unsigned char)(rng() & 0xff));
    } else {
      size_t at = rng() % span, len = std::min<size_t>(1 + rng() % 16, b.size() - at);
      b.erase(b.begin() + at, b.begin() + at + len);
    }
    FILE *f = std::fopen(argv[2], "wb");
    if (!f) return 2;
    std::fwrite(b.data(), 1, b.size(), f);
    std::fclose(f);
    rrtmgsw::NcFile nc;
    std::string err;
    if (!nc.open(argv[2], err)) {
      refused++;
      continue;
    }
    read++;
    for (const auto &v : nc.vars()) {
      std::vector<uint64_t> start(v.shape.size(), 0), count(v.shape);
      uint64_t total = 1;
      for (uint64_t c : count) total *= c;
      if (total > (1u << 26)) continue;   // a damaged header may describe a huge (empty) slab
      out.assign(total ? total : 1, 0.0);
      nc.get(v, start.data(), count.data(), out.data(), err);
    }
  }
  std::printf("mutations=%d refused=%d read=%d\n", n, refused, read);
  return 0;
}
