// Test helper for Gadget-HDF5 snapshot sets (built with the HDF5 library).
//
//   hdf5_snap write OUT NFILES BOX REDSHIFT IN
//     Writes OUT.<i>.hdf5 for i < NFILES (NFILES = 0: the single file
//     OUT.hdf5) from IN.table (6 float64: the MassTable) and, per type t,
//     IN.pos<t> (float32 [n][3]) and, when its MassTable entry is 0,
//     IN.mass<t> (float32 [n]); absent IN.pos<t>: no particles.  Each type's
//     particles are split into contiguous ranges over the files.
//
//   hdf5_snap dump NAME OUTPREFIX
//     Reads the set named NAME with the gen-pk reader (hdf5snap.h) and writes
//     OUTPREFIX.info (files, header values, per type the particles read and
//     the total mass as gen-pk accumulates it) and per type OUTPREFIX.pos<t>,
//     OUTPREFIX.mass<t> (the particles of all files in file order; masses
//     only for per-particle masses).
#include "hdf5snap.h"

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

namespace {

std::vector<char> slurp(const std::string &f, bool &found) {
  std::vector<char> d;
  FILE *fp = std::fopen(f.c_str(), "rb");
  found = fp != nullptr;
  if (!fp) return d;
  std::fseek(fp, 0, SEEK_END);
  d.resize((size_t)std::ftell(fp));
  std::fseek(fp, 0, SEEK_SET);
  if (!d.empty() && std::fread(d.data(), 1, d.size(), fp) != d.size()) std::exit(2);
  std::fclose(fp);
  return d;
}

void spit(const std::string &f, const void *p, size_t n) {
  FILE *fp = std::fopen(f.c_str(), "wb");
  if (!fp || (n && std::fwrite(p, 1, n, fp) != n)) std::exit(2);
  std::fclose(fp);
}

int write_set(const std::string &out, int nfiles, double box, double redshift, const std::string &in) {
  bool found;
  const std::vector<char> table = slurp(in + ".table", found);
  if (!found || table.size() != 48) return 2;
  double mass[6];
  std::memcpy(mass, table.data(), 48);
  std::vector<float> pos[6], ms[6];
  unsigned int ntot[6] = {0, 0, 0, 0, 0, 0}, hw[6] = {0, 0, 0, 0, 0, 0};
  for (int t = 0; t < 6; t++) {
    std::vector<char> p = slurp(in + ".pos" + std::to_string(t), found);
    if (!found) continue;
    pos[t].resize(p.size() / 4);
    std::memcpy(pos[t].data(), p.data(), p.size());
    ntot[t] = (unsigned int)(pos[t].size() / 3);
    if (mass[t] == 0) {
      std::vector<char> m = slurp(in + ".mass" + std::to_string(t), found);
      if (!found || m.size() != (size_t)ntot[t] * 4) return 2;
      ms[t].resize(ntot[t]);
      std::memcpy(ms[t].data(), m.data(), m.size());
    }
  }
  const int nf = nfiles > 0 ? nfiles : 1;
  for (int i = 0; i < nf; i++) {
    const std::string name = nfiles > 0 ? out + "." + std::to_string(i) + ".hdf5" : out + ".hdf5";
    const hid_t f = H5Fcreate(name.c_str(), H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
    if (f < 0) return 2;
    H5Gclose(H5Gcreate2(f, "/Header", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT));
    int nthis[6];
    size_t lo[6];
    for (int t = 0; t < 6; t++) {
      lo[t] = (size_t)ntot[t] * i / nf;
      nthis[t] = (int)((size_t)ntot[t] * (i + 1) / nf - lo[t]);
    }
    const double time = 1.0 / (1.0 + redshift), h100 = 0.7, om = 0.3, ol = 0.7;
    const int cooling = 0, nfs = nf;
    H5LTset_attribute_int(f, "/Header", "NumPart_ThisFile", nthis, 6);
    H5LTset_attribute_uint(f, "/Header", "NumPart_Total", ntot, 6);
    H5LTset_attribute_uint(f, "/Header", "NumPart_Total_HighWord", hw, 6);
    H5LTset_attribute_double(f, "/Header", "MassTable", mass, 6);
    H5LTset_attribute_double(f, "/Header", "Time", &time, 1);
    H5LTset_attribute_double(f, "/Header", "Redshift", &redshift, 1);
    H5LTset_attribute_double(f, "/Header", "BoxSize", &box, 1);
    H5LTset_attribute_double(f, "/Header", "HubbleParam", &h100, 1);
    H5LTset_attribute_double(f, "/Header", "Omega0", &om, 1);
    H5LTset_attribute_double(f, "/Header", "OmegaLambda", &ol, 1);
    H5LTset_attribute_int(f, "/Header", "Flag_Cooling", &cooling, 1);
    H5LTset_attribute_int(f, "/Header", "NumFilesPerSnapshot", &nfs, 1);
    for (int t = 0; t < 6; t++) {
      if (nthis[t] == 0) continue;
      const std::string g = "/PartType" + std::to_string(t);
      H5Gclose(H5Gcreate2(f, g.c_str(), H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT));
      const hsize_t d2[2] = {(hsize_t)nthis[t], 3}, d1[1] = {(hsize_t)nthis[t]};
      H5LTmake_dataset_float(f, (g + "/Coordinates").c_str(), 2, d2, pos[t].data() + 3 * lo[t]);
      if (mass[t] == 0) H5LTmake_dataset_float(f, (g + "/Masses").c_str(), 1, d1, ms[t].data() + lo[t]);
    }
    H5Fclose(f);
  }
  return 0;
}

int dump_set(const std::string &name, const std::string &out) {
  const std::vector<std::string> files = hdf5snap::find_set(name);
  FILE *info = std::fopen((out + ".info").c_str(), "w");
  if (!info) return 2;
  std::fprintf(info, "files %zu\n", files.size());
  for (const std::string &f : files) std::fprintf(info, "file %s\n", f.c_str());
  if (!files.empty()) {
    hdf5snap::Header h;
    if (!hdf5snap::load_header(files[0], h)) {
      std::fprintf(info, "header failed\n");
    } else {
      std::fprintf(info, "box %.17g\nredshift %.17g\ntime %.17g\n", h.box, h.redshift, h.time);
      for (int t = 0; t < 6; t++)
        std::fprintf(info, "total %d %ld %.17g\n", t, (long)h.npart_total[t], h.mass[t]);
    }
  }
  for (int t = 0; t < 6; t++) {
    std::vector<float> pos, ms;
    double total_mass = 0;
    int64_t n = 0;
    hdf5snap::FileTypeData d;
    for (size_t i = 0; i < files.size(); i++) {
      if (hdf5snap::read_type(files[i], t, (int)i, d) < 0 || d.n == 0) continue;
      if (d.mass_table == 0) {
        double this_file = 0;
        for (int64_t j = 0; j < d.n; j++) this_file += d.masses[(size_t)j];
        total_mass += this_file;
        ms.insert(ms.end(), d.masses.begin(), d.masses.end());
      }
      total_mass += d.mass_table * d.n;
      pos.insert(pos.end(), d.pos.begin(), d.pos.end());
      n += d.n;
    }
    std::fprintf(info, "type %d %ld %.17g\n", t, (long)n, total_mass);
    spit(out + ".pos" + std::to_string(t), pos.data(), pos.size() * 4);
    spit(out + ".mass" + std::to_string(t), ms.data(), ms.size() * 4);
  }
  std::fclose(info);
  return 0;
}

}  // namespace

int main(int argc, char **argv) {
  const std::string mode = argc > 1 ? argv[1] : "";
  if (mode == "write" && argc == 7) return write_set(argv[2], std::atoi(argv[3]), std::atof(argv[4]), std::atof(argv[5]), argv[6]);
  if (mode == "dump" && argc == 4) return dump_set(argv[2], argv[3]);
  std::fprintf(stderr, "usage: hdf5_snap write OUT NFILES BOX REDSHIFT IN | hdf5_snap dump NAME OUTPREFIX\n");
  return 2;
}
