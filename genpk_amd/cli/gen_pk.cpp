// gen-pk: matter power spectra of Gadget snapshots on the MI355X, with the
// reference driver's options, console output and output files
// (gen-pk.cpp:74-360, read_fieldize.cpp:18-340, utils.cpp:7-72).
//
//   gen-pk -i snap -o outdir            auto spectrum of every particle type: PK-<type>-<snap>
//   gen-pk -i snap -j other -o outdir   cross spectrum of each type between two snapshots: PX-<type>-<snap>
//   gen-pk -i snap -c T -o outdir       cross spectrum of dark matter (type 1) with type T: PK-DMx<T>-<snap>
//   -s 1                                 add the stars (type 4) to the baryons (type 0)
//   -p 32|64                             grid/FFT precision (default 64, the reference's double)
//
// Snapshots are Gadget format I/II file sets, MP-Gadget bigfile snapshots, or
// Gadget-HDF5 file sets when built with the HDF5 library (GENPK_HDF5).  Particles are read on the host
// (Gadget: chunks of at most 2^26, as the reference; HDF5: a file at a time),
// copied to the device in chunks of at most 2^26 and deposited into a plan;
// the fused transform + binning runs on the device.  The total mass of a
// species is the sum of its particle masses (header mass x count, or the
// MASS block / Masses dataset), plus 1 per read of a species from a Gadget
// set (read_fieldize.cpp:94); HDF5 and bigfile reads add nothing
// (read_fieldize.cpp:320-331, read_fieldize_bigfile.cpp:94-114).
#include "../../include/genpk_amd.h"
#include "bigfile.h"
#include "gadget.h"
#ifdef GENPK_HDF5
#include "hdf5snap.h"
#endif

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <memory>
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

// Deposit n particles held on the host (masses null: constant mass).
void deposit_host(genpk_plan *plan, double box, int64_t n, const float *pos, const float *masses, double mass,
                  Staging &st, hipStream_t s) {
  for (int64_t at = 0; at < n; at += CHUNK) {
    const int64_t parts = std::min(CHUNK, n - at);
    st.ensure(parts);
    hip_ok(hipMemcpyAsync(st.pos, pos + 3 * at, (size_t)parts * 12, hipMemcpyHostToDevice, s), "copy positions");
    if (masses)
      hip_ok(hipMemcpyAsync(st.mass, masses + at, (size_t)parts * 4, hipMemcpyHostToDevice, s), "copy masses");
    ok(genpk_deposit(plan, box, parts, st.pos, masses ? st.mass : nullptr, mass), "deposit");
    hip_ok(hipStreamSynchronize(s), "synchronize");
  }
}

#ifdef GENPK_HDF5
// read_fieldize_hdf5 (read_fieldize.cpp:258-339) over the files of a set:
// every file's particles of `type`; per file the masses summed in double in
// file order, then + MassTable x count.  A file that cannot be read is
// skipped, as the reference's caller does (gen-pk.cpp:213-216).
void read_fieldize_hdf5(genpk_plan *plan, const std::vector<std::string> &files, int type, double box,
                        double *total_mass, Staging &st, hipStream_t s) {
  hdf5snap::FileTypeData d;
  for (size_t fileno = 0; fileno < files.size(); fileno++) {
    if (hdf5snap::read_type(files[fileno], type, (int)fileno, d) < 0 || d.n == 0) continue;
    const bool per_particle = d.mass_table == 0;
    if (per_particle) {
      double this_file = 0;
      for (int64_t i = 0; i < d.n; i++) this_file += d.masses[(size_t)i];
      *total_mass += this_file;
    }
    *total_mass += d.mass_table * d.n;
    deposit_host(plan, box, d.n, d.pos.data(), per_particle ? d.masses.data() : nullptr, d.mass_table, st, s);
  }
}
#endif

// read_fieldize_bigfile (read_fieldize_bigfile.cpp:63-125): every particle
// of `type`; the masses summed in double in block order, then + MassTable x
// count.  A block that cannot be read ends the program, as the reference
// cannot go on without its coordinates.
void read_fieldize_bigfile(genpk_plan *plan, const std::string &name, const bigfile::Header &h, int type,
                           double box, double *total_mass, Staging &st, hipStream_t s) {
  std::vector<float> pos, masses;
  std::string err;
  if (!bigfile::read_type(name, type, h, pos, masses, err)) die("Failed to read from block: " + err);
  const int64_t n = h.npart_total[type];
  if (n == 0) return;
  const bool per_particle = h.mass[type] == 0;
  if (per_particle) {
    double this_file = 0;
    for (int64_t i = 0; i < n; i++) this_file += masses[(size_t)i];
    *total_mass += this_file;
  }
  *total_mass += h.mass[type] * n;
  deposit_host(plan, box, n, pos.data(), per_particle ? masses.data() : nullptr, h.mass[type], st, s);
}

// A snapshot: a Gadget set, an HDF5 set (found first, as the reference does)
// or a bigfile snapshot.
struct Source {
  std::unique_ptr<gadget::Snapshot> gadget;
  std::vector<std::string> hdf5;
  std::string big;        // bigfile snapshot directory
  bigfile::Header bighdr;
  explicit Source(const std::string &name) {
#ifdef GENPK_HDF5
    hdf5 = hdf5snap::find_set(name);
    if (!hdf5.empty()) return;
#endif
    if (bigfile::is_bigfile(name)) {
      std::string err;
      if (!bigfile::load_header(name, bighdr, err)) die(err + "\nCould not load header");
      big = name;
      return;
    }
    gadget = std::make_unique<gadget::Snapshot>(name);
  }
  bool is_hdf5() const { return !hdf5.empty(); }
  bool is_bigfile() const { return !big.empty(); }
  // Gadget: read_fieldize, returning 1 when the snapshot holds none of the
  // type; HDF5 and bigfile: their readers, returning 0.
  int read(genpk_plan *plan, int type, double box, double *total_mass, Staging &st, hipStream_t s) const {
#ifdef GENPK_HDF5
    if (is_hdf5()) {
      read_fieldize_hdf5(plan, hdf5, type, box, total_mass, st, s);
      return 0;
    }
#endif
    if (is_bigfile()) {
      read_fieldize_bigfile(plan, big, bighdr, type, box, total_mass, st, s);
      return 0;
    }
    return read_fieldize(plan, *gadget, type, box, total_mass, st, s);
  }
};

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
    const Source snap(infiles);
    int64_t npart_total[gadget::N_TYPE];
    double box = 0;
    if (snap.is_hdf5()) {
#ifdef GENPK_HDF5
      // load_hdf5_header (read_fieldize.cpp:209-254) and its console lines
      std::fprintf(stderr, "Found %lu hdf5 files in snapshot\n", (unsigned long)snap.hdf5.size());
      hdf5snap::Header h;
      if (!hdf5snap::load_header(snap.hdf5[0], h)) {
        std::fprintf(stderr, "Could not load header\n");
        return 1;
      }
      for (int t = 0; t < gadget::N_TYPE; t++) npart_total[t] = h.npart_total[t];
      box = h.box;
      std::printf("NumPart=[%ld,%ld,%ld,%ld,%ld,%ld], ", (long)npart_total[0], (long)npart_total[1],
                  (long)npart_total[2], (long)npart_total[3], (long)npart_total[4], (long)npart_total[5]);
      std::printf("Masses=[%g %g %g %g %g %g], ", h.mass[0], h.mass[1], h.mass[2], h.mass[3], h.mass[4], h.mass[5]);
      std::printf("Redshift=%g, \xce\xa9_M=%g \xce\xa9_L=%g\n", h.redshift, h.omega0, h.omegal);
      std::printf("Expansion factor = %f\n", h.time);
      std::printf("Hubble = %g Box=%g \n", h.h100, h.box);
#endif
    } else if (snap.is_bigfile()) {
      // load_bigfile_header's console lines (read_fieldize_bigfile.cpp:53-58)
      const bigfile::Header &h = snap.bighdr;
      for (int t = 0; t < gadget::N_TYPE; t++) npart_total[t] = h.npart_total[t];
      box = h.box;
      std::printf("NumPart=[%ld,%ld,%ld,%ld,%ld,%ld], ", (long)npart_total[0], (long)npart_total[1],
                  (long)npart_total[2], (long)npart_total[3], (long)npart_total[4], (long)npart_total[5]);
      std::printf("Masses=[%g %g %g %g %g %g], ", h.mass[0], h.mass[1], h.mass[2], h.mass[3], h.mass[4], h.mass[5]);
      std::printf("Redshift=%g, \xce\xa9_M=%g\n", h.redshift, h.omega0);
      std::printf("Expansion factor = %f\n", h.time);
      std::printf("Hubble = %g Box=%g \n", h.h100, h.box);
    } else {
      const gadget::Header &h = snap.gadget->header();
      for (int t = 0; t < gadget::N_TYPE; t++) npart_total[t] = snap.gadget->npart(t);
      box = h.box;
      std::printf("Boxsize=%g, ", box);
      std::printf("NPart=(%g,%g,%g,%g,%g,%g)**3\n", std::cbrt((double)npart_total[0]),
                  std::cbrt((double)npart_total[1]), std::cbrt((double)npart_total[2]),
                  std::cbrt((double)npart_total[3]), std::cbrt((double)npart_total[4]),
                  std::cbrt((double)npart_total[5]));
      std::printf("Masses=[%g %g %g ]\n", h.mass[0], h.mass[1], h.mass[2]);
      std::printf("redshift=%g, \xce\xa9_M=%g\n", h.redshift, h.omega0);
    }
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
        snap.read(a, type, box, &total_mass, st, s);
        if (type == BARYON_TYPE && stars_are_baryons) snap.read(a, STARS_TYPE, box, &total_mass, st, s);
        std::printf("total_mass in type %d = %g\n", type, total_mass);
        std::fflush(stdout);
        sp.run(a, nullptr, total_mass, total_mass, s);
        sp.write(outdir + "/PK-" + type_str(type) + "-" + base);
      }
    } else if (!jinfiles.empty()) {
      const Source snap2(jinfiles);
      // With an HDF5 second set the reference crosses type 1 of the two sets
      // under every type's name (gen-pk.cpp:271-288); so does this driver
      // (the reference reads nothing for the first set when that one is a
      // Gadget set).  Otherwise each type is read from both sets, whatever
      // their formats (with an HDF5 first set and a Gadget second set the
      // reference dereferences a missing snapshot; with a bigfile first set
      // it reads the first set twice into the first field).
      const bool dm_only = snap2.is_hdf5();
      ok(genpk_plan_create(&b, field_dims, precision, s), "plan");
      for (int type = 0; type < gadget::N_TYPE; type++) {
        if (npart_total[type] == 0) continue;
        ok(genpk_clear(a), "clear");
        ok(genpk_clear(b), "clear");
        std::fprintf(stderr, "Reading...\n");
        double total_mass = 0, total_mass2 = 0;
        const int rt = dm_only ? 1 : type;
        if (snap.read(a, rt, box, &total_mass, st, s)) continue;
        if (snap2.read(b, rt, box, &total_mass2, st, s)) continue;
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
      snap.read(a, 1, box, &total_mass, st, s);
      snap.read(b, crosstype, box, &total_mass2, st, s);
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
