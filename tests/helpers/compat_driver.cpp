// Test helper: the reference driver's forward for one species through the
// drop-in shims of include/genpk_compat.h (gen-pk.cpp:208-237 with
// read_fieldize.cpp:43-94): zero the grid, fieldize once per chunk of at
// most CHUNK particles (constant mass 1, +1 per read), in-place FFT, then
// powerspectrum.  Linked against libgenpk_amd.so only through that header.
//
//   compat_driver POSFILE N DIMS BOX CHUNK OUTFILE [resident]
//
// POSFILE holds float32 [N][3].  OUTFILE receives power[DIMS] (double),
// count[DIMS] (int), keffs[DIMS] (double) and the total mass (double).  One
// line of timings is printed.  "resident" calls genpk_compat_resident(1).
#include "genpk_compat.h"

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

extern "C" const char *genpk_last_error(void);

static double now() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

int main(int argc, char **argv) {
  if (argc < 7) {
    std::fprintf(stderr, "usage: %s POSFILE N DIMS BOX CHUNK OUTFILE [resident]\n", argv[0]);
    return 2;
  }
  const int64_t n = std::atoll(argv[2]);
  const int dims = std::atoi(argv[3]);
  const double box = std::atof(argv[4]);
  const int64_t chunk = std::atoll(argv[5]);
  const bool resident = argc > 7 && std::string(argv[7]) == "resident";
  std::vector<float> pos((size_t)n * 3);
  FILE *f = std::fopen(argv[1], "rb");
  if (!f || std::fread(pos.data(), 12, (size_t)n, f) != (size_t)n) {
    std::fprintf(stderr, "cannot read %s\n", argv[1]);
    return 1;
  }
  std::fclose(f);
  if (resident && genpk_compat_resident(1)) return 1;
  const size_t field_size = (size_t)2 * dims * dims * (dims / 2 + 1);
  std::vector<double> field(field_size, 0.0);  // memset(field, 0, ...) of gen-pk.cpp:208
  double total_mass = 0;
  const double t0 = now();
  for (int64_t read = 0; read < n;) {
    const int64_t parts = n - read < chunk ? n - read : chunk;
    if (fieldize(box, dims, field.data(), parts, pos.data() + 3 * read, nullptr, 1.0, 1)) {
      std::fprintf(stderr, "fieldize: %s\n", genpk_last_error());
      return 1;
    }
    total_mass += 1.0 * parts;
    read += parts;
  }
  total_mass += 1;  // read_fieldize.cpp:94
  const double t1 = now();
  if (genpk_fft_r2c_3d_inplace(dims, field.data())) {
    std::fprintf(stderr, "fft: %s\n", genpk_last_error());
    return 1;
  }
  const double t2 = now();
  std::vector<double> power(dims), keffs(dims);
  std::vector<int> count(dims);
  auto *spec = reinterpret_cast<double(*)[2]>(field.data());
  if (powerspectrum(dims, spec, spec, dims, power.data(), count.data(), keffs.data(), total_mass, total_mass)) {
    std::fprintf(stderr, "powerspectrum: %s\n", genpk_last_error());
    return 1;
  }
  const double t3 = now();
  FILE *o = std::fopen(argv[6], "wb");
  if (!o) return 1;
  std::fwrite(power.data(), 8, dims, o);
  std::fwrite(count.data(), 4, dims, o);
  std::fwrite(keffs.data(), 8, dims, o);
  std::fwrite(&total_mass, 8, 1, o);
  std::fclose(o);
  std::printf("{\"fieldize_s\": %.6f, \"fft_s\": %.6f, \"powerspectrum_s\": %.6f, \"total_s\": %.6f}\n", t1 - t0,
              t2 - t1, t3 - t2, t3 - t0);
  return 0;
}
