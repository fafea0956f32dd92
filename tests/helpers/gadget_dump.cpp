// Test helper: reads a snapshot with the gen-pk C++ reader and prints, per
// type, the particle count and the positions and masses (as read in chunks
// of `chunk` particles) as text, for comparison with the Python reader.
#include "../../genpk_amd/cli/gadget.h"

#include <cstdio>
#include <cstdlib>
#include <vector>

int main(int argc, char **argv) {
  if (argc < 3) return 2;
  try {
    gadget::Snapshot s(argv[1]);
    const int64_t chunk = std::atoll(argv[2]);
    const gadget::Header &h = s.header();
    std::printf("files %d box %.17g z %.17g om %.17g posbytes %lld\n", s.num_files(), h.box, h.redshift, h.omega0,
                (long long)s.bytes_per_particle("POS "));
    for (int t = 0; t < gadget::N_TYPE; t++) {
      const int64_t n = s.npart(t);
      std::printf("type %d n %lld hmass %.17g\n", t, (long long)n, h.mass[t]);
      std::vector<float> p(3 * (size_t)chunk), m((size_t)chunk);
      for (int64_t at = 0; at < n; at += chunk) {
        const int64_t k = std::min(chunk, n - at);
        if (s.read("POS ", t, at, k, p.data(), 12) != k) return 3;
        const bool mb = h.mass[t] == 0;
        if (mb && s.read("MASS", t, at, k, m.data(), 4) != k) return 4;
        for (int64_t i = 0; i < k; i++)
          std::printf("%a %a %a %a\n", p[3 * i], p[3 * i + 1], p[3 * i + 2], mb ? m[i] : (float)h.mass[t]);
      }
    }
  } catch (const std::exception &e) {
    std::fprintf(stderr, "%s\n", e.what());
    return 1;
  }
  return 0;
}
