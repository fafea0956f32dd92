// gen-pk: matter power spectra of Gadget snapshots on the MI355X, with the
// reference driver's options, console output and output files
// (gen-pk.cpp:74-360, read_fieldize.cpp:18-97, utils.cpp:7-72).
//
//   gen-pk -i snap -o outdir            auto spectrum of every particle type: PK-<type>-<snap>
//   gen-pk -i snap -j other -o outdir   cross spectrum of each type between two snapshots: PX-<type>-<snap>
//   gen-pk -i snap -c T -o outdir       cross spectrum of dark matter (type 1) with type T: PK-DMx<T>-<snap>
//   -s 1                                 add the stars (type 4) to the baryons (type 0)
//   -p 32|64                             grid/FFT precision (default 64, the reference's double)
//
// Particles are read on the host in chunks of at most 2^26 (as the reference
// does), copied to the device and deposited into a plan; the fused
// transform + binning runs on the device.  The total mass of a species is
// the sum of its particle masses (header mass x count, or the MASS block)
// plus 1 per read of a species, the reference's convention
// (read_fieldize.cpp:94).  HDF5 and bigfile snapshots are not supported.
#include "../../include/genpk_amd.h"
#include "gadget.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <unistd.h>
#include <vector>

namespace {

constexpr int64_t FIELD_DIMS = 3072;
constexpr int64_t CHUNK = (int64_t)1 << 26;
constexpr int BARYON_TYPE = 0, STARS_TYPE = 4;

void help() {
  std::fprintf(stderr,
               "Usage: ./gen-pk -i filenames -j other_filenames -o outdir -c (optional) cross-corr type\n"
               "Outputs one file per particle type, with the name PK-$TYPE-$INPUT\n"
               "Each output file has three columns, for each bin, k_eff, P(k) and N_modes\n"
               "if -j is specified it cross-correlates files specified under -i with files specifed under\n"
               "-j (one output per particle type).\n"
               "If -c is specified the code computes the cross-correlation of that particle type with \n"
               "the CDM (type 1) (within the file specified by -i)\n"
               "-s (1,0) Determines whether stars are included in the baryon type.\n"
               "-p (32,64) grid and FFT precision (default 64).\n");
}

std::string type_str(int type) {
  char b[8];
  genpk_type_str(type, b, sizeof b);
  return b;
}

[[noreturn]] void die(const std::string &msg) {
  std::fprintf(stderr, "%s\n", msg.c_str());
  std::exit(1);
}

void ok(int status, const char *what) {
  if (status) die(std::string(what) + ": " + genpk_last_error());
}

void hip_ok(hipError_t e, const char *what) {
  if (e != hipSuccess) die(std::string(what) + ": " + hipGetErrorString(e));
}

// Device staging buffers for one chunk of particles.
struct Staging {
  float *pos = nullptr, *mass = nullptr;
  int64_t cap = 0;
  std::vector<float> hpos, hmass;
  void ensure(int64_t n) {
    if (n <= cap) return;
    (void)hipFree(pos);
    (void)hipFree(mass);
    hip_ok(hipMalloc(&pos, n * 12), "hipMalloc positions");
    hip_ok(hipMalloc(&mass, n * 4), "hipMalloc masses");
    cap = n;
  }
  ~Staging() {
    (void)hipFree(pos);
    (void)hipFree(mass);
  }
};

// read_fieldize (read_fieldize.cpp:18-97): deposit every particle of `type`
// into the plan, accumulating; adds the species' mass plus 1 to *total_mass.
// Returns 1 (nothing read, no +1) when the snapshot holds none of the type.
int read_fieldize(genpk_plan *plan, const gadget::Snapshot &snap, int type, double box, double *total_mass,
                  Staging &st, hipStream_t s) {
  const int64_t npart = snap.npart(type);
  if (npart == 0) return 1;
  const double mass = snap.header().mass[type];
  const int64_t partlen = snap.bytes_per_particle("POS ");
  if (partlen != 12) {
    std::fprintf(stderr,
                 "The pos array uses %ld bytes per particle, instead of %lu.\n You probably want to recompile a "
                 "different DOUBLE_PRECISION_SNAP define.\n",
                 (long)partlen, (unsigned long)12);
    std::exit(1);
  }
  int64_t parts = std::min(npart, CHUNK);
  st.ensure(parts);
  st.hpos.resize((size_t)parts * 3);
  if (mass == 0) st.hmass.resize((size_t)parts);
  int64_t toread = npart, read = 0;
  while (toread > 0) {
    if (toread < parts) parts = toread;
    if (snap.read("POS ", type, read, parts, st.hpos.data(), 12) != parts) die("Error reading particle data");
    hip_ok(hipMemcpyAsync(st.pos, st.hpos.data(), (size_t)parts * 12, hipMemcpyHostToDevice, s), "copy positions");
    const float *dm = nullptr;
    if (mass == 0) {
      if (snap.read("MASS", type, read, parts, st.hmass.data(), 4) != parts) die("Error reading mass data");
      double this_file = 0;
      for (int64_t i = 0; i < parts; i++) this_file += st.hmass[(size_t)i];
      *total_mass += this_file;
      hip_ok(hipMemcpyAsync(st.mass, st.hmass.data(), (size_t)parts * 4, hipMemcpyHostToDevice, s), "copy masses");
      dm = st.mass;
    }
    *total_mass += mass * parts;
    ok(genpk_deposit(plan, box, parts, st.pos, dm, mass), "deposit");
    // the host buffers are refilled next: wait for the copies
    hip_ok(hipStreamSynchronize(s), "synchronize");
    toread -= parts;
    read += parts;
  }
  *total_mass += 1;
  return 0;
}

struct Spectrum {
  int nrbins;
  double *dpow_aa, *dpow_ab, *dkeff;
  int *dcount;
  std::vector<double> power, keffs;
  std::vector<int> count;
  explicit Spectrum(int n) : nrbins(n), power(n), keffs(n), count(n) {
    hip_ok(hipMalloc(&dpow_aa, n * 8), "hipMalloc");
    hip_ok(hipMalloc(&dpow_ab, n * 8), "hipMalloc");
    hip_ok(hipMalloc(&dkeff, n * 8), "hipMalloc");
    hip_ok(hipMalloc(&dcount, n * 4), "hipMalloc");
  }
  ~Spectrum() {
    (void)hipFree(dpow_aa);
    (void)hipFree(dpow_ab);
    (void)hipFree(dkeff);
    (void)hipFree(dcount);
  }
  // powerspectrum(dims, Fa, Fb, ...) of powerspectrum.c:35-110: auto when b is null, else Re(Fa Fb*)/(ma mb)
  void run(genpk_plan *a, genpk_plan *b, double ma, double mb, hipStream_t s) {
    if (b)
      ok(genpk_power(a, b, nrbins, nullptr, nullptr, dpow_ab, dcount, dkeff, ma, mb), "power");
    else
      ok(genpk_power(a, nullptr, nrbins, dpow_aa, nullptr, nullptr, dcount, dkeff, ma, ma), "power");
    hip_ok(hipMemcpyAsync(power.data(), b ? dpow_ab : dpow_aa, nrbins * 8, hipMemcpyDeviceToHost, s), "copy");
    hip_ok(hipMemcpyAsync(keffs.data(), dkeff, nrbins * 8, hipMemcpyDeviceToHost, s), "copy");
    hip_ok(hipMemcpyAsync(count.data(), dcount, nrbins * 4, hipMemcpyDeviceToHost, s), "copy");
    hip_ok(hipStreamSynchronize(s), "synchronize");
  }
  void write(const std::string &file) {
    ok(genpk_print_pk(file.c_str(), nrbins, keffs.data(), power.data(), count.data()), "print_pk");
  }
};

}  // namespace

int main(int argc, char *argv[]) {
  std::string infiles, jinfiles, outdir;
  int crosstype = -1;
  bool stars_are_baryons = false;
  int precision = GENPK_FLOAT64;
  int c;
  while ((c = getopt(argc, argv, "i:j:o:c:s:p:h")) != -1) {
    switch (c) {
      case 'o': outdir = optarg; break;
      case 'i': infiles = optarg; break;
      case 'j': jinfiles = optarg; break;
      case 'c': crosstype = std::atoi(optarg); break;
      case 's': stars_are_baryons = std::atoi(optarg) != 0; break;
      case 'p': precision = std::atoi(optarg) == 32 ? GENPK_FLOAT32 : GENPK_FLOAT64; break;
      case 'h':
      default: help(); return 0;
    }
  }
  if (outdir.empty() || infiles.empty()) {
    help();
    return 0;
  }
  try {
    gadget::Snapshot snap(infiles);
    const gadget::Header &h = snap.header();
    int64_t npart_total[gadget::N_TYPE];
    for (int t = 0; t < gadget::N_TYPE; t++) npart_total[t] = snap.npart(t);
    const double box = h.box;
    std::printf("Boxsize=%g, ", box);
    std::printf("NPart=(%g,%g,%g,%g,%g,%g)**3\n", std::cbrt((double)npart_total[0]), std::cbrt((double)npart_total[1]),
                std::cbrt((double)npart_total[2]), std::cbrt((double)npart_total[3]),
                std::cbrt((double)npart_total[4]), std::cbrt((double)npart_total[5]));
    std::printf("Masses=[%g %g %g ]\n", h.mass[0], h.mass[1], h.mass[2]);
    std::printf("redshift=%g, \xce\xa9_M=%g\n", h.redshift, h.omega0);
    // grid: max over types of min(2*nexttwo((int)cbrt(N)), FIELD_DIMS) (gen-pk.cpp:169-173)
    int64_t field_dims = 0;
    for (int t = 0; t < gadget::N_TYPE; t++) {
      const int64_t tmp = 2 * (int64_t)genpk_nexttwo((int)std::cbrt((double)npart_total[t]));
      field_dims = std::max(field_dims, std::min(tmp, FIELD_DIMS));
    }
    const int nrbins = (int)field_dims;
    std::printf("FFT grid dimension: %lu\n", (unsigned long)field_dims);
    std::fflush(stdout);
    if (field_dims < 2) die("no particles in the snapshot");

    hipStream_t s;
    hip_ok(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "stream");
    genpk_plan *a = nullptr, *b = nullptr;
    ok(genpk_plan_create(&a, field_dims, precision, s), "plan");
    Staging st;
    Spectrum sp(nrbins);
    const size_t last = infiles.find_last_of("/\\");
    const std::string base = infiles.substr(last == std::string::npos ? 0 : last + 1);

    if (crosstype < 0 && jinfiles.empty()) {
      for (int type = 0; type < gadget::N_TYPE; type++) {
        if (npart_total[type] == 0) continue;
        ok(genpk_clear(a), "clear");
        double total_mass = 0;
        read_fieldize(a, snap, type, box, &total_mass, st, s);
        if (type == BARYON_TYPE && stars_are_baryons) read_fieldize(a, snap, STARS_TYPE, box, &total_mass, st, s);
        std::printf("total_mass in type %d = %g\n", type, total_mass);
        std::fflush(stdout);
        sp.run(a, nullptr, total_mass, total_mass, s);
        sp.write(outdir + "/PK-" + type_str(type) + "-" + base);
      }
    } else if (!jinfiles.empty()) {
      gadget::Snapshot snap2(jinfiles);
      ok(genpk_plan_create(&b, field_dims, precision, s), "plan");
      for (int type = 0; type < gadget::N_TYPE; type++) {
        if (npart_total[type] == 0) continue;
        ok(genpk_clear(a), "clear");
        ok(genpk_clear(b), "clear");
        std::fprintf(stderr, "Reading...\n");
        double total_mass = 0, total_mass2 = 0;
        if (read_fieldize(a, snap, type, box, &total_mass, st, s)) continue;
        if (read_fieldize(b, snap2, type, box, &total_mass2, st, s)) continue;
        sp.run(a, b, total_mass, total_mass2, s);
        sp.write(outdir + "/PX-" + type_str(type) + "-" + base);
      }
    } else {
      if (crosstype >= gadget::N_TYPE || npart_total[1] == 0 || npart_total[crosstype] == 0) {
        std::fprintf(stderr, "Can't cross-correlate types not present in snapshot\n");
        return 1;
      }
      ok(genpk_plan_create(&b, field_dims, precision, s), "plan");
      ok(genpk_clear(a), "clear");
      ok(genpk_clear(b), "clear");
      double total_mass = 0, total_mass2 = 0;
      read_fieldize(a, snap, 1, box, &total_mass, st, s);
      read_fieldize(b, snap, crosstype, box, &total_mass2, st, s);
      sp.run(a, b, total_mass, total_mass2, s);
      sp.write(outdir + "/PK-DMx" + type_str(crosstype) + "-" + base);
    }
    if (b) genpk_plan_destroy(b);
    genpk_plan_destroy(a);
    (void)hipStreamDestroy(s);
  } catch (const std::exception &e) {
    std::fprintf(stderr, "%s\n", e.what());
    return 1;
  }
  return 0;
}
