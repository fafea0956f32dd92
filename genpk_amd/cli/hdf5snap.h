// Gadget-HDF5 snapshot sets for gen-pk: discovery of the file set, the
// header of the first file and the coordinates (and masses) of one particle
// type in one file, following the reference's HDF5 path
// (read_fieldize.cpp:99-340).
//
//  - A set is found from a name as given, else name.0.hdf5, else name.hdf5
//    (read_fieldize.cpp:101-127); the files of a set are name.N.hdf5 for
//    N = 1, 2, ... while they exist and are HDF5 (:142-167).
//  - The header (first file, /Header attributes) must hold Time, Redshift,
//    BoxSize, HubbleParam, Omega0, OmegaLambda and Flag_Cooling; totals are
//    NumPart_Total + 2^32 NumPart_Total_HighWord; MassTable (:209-254).
//  - Per file: NumPart_ThisFile, Omega0 and MassTable must be present;
//    /PartType<t>/Coordinates is a rank-2 [n][3] 4-byte float dataset
//    with n = NumPart_ThisFile[t]; for a type whose MassTable entry is 0,
//    /PartType<t>/Masses is a rank-1 4-byte float dataset of the same
//    length (:258-339).
#pragma once

#include <hdf5.h>
#include <hdf5_hl.h>

#include <cstdint>
#include <cstdio>
#include <sstream>
#include <string>
#include <vector>

namespace hdf5snap {

constexpr int N_TYPE = 6;

// H5Fis_hdf5 with the library's error printing switched off.
inline bool is_hdf5(const std::string &f) {
  H5E_auto2_t old_func;
  void *old_data;
  H5Eget_auto2(H5E_DEFAULT, &old_func, &old_data);
  H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
  const bool yes = H5Fis_hdf5(f.c_str()) > 0;
  H5Eset_auto2(H5E_DEFAULT, old_func, old_data);
  return yes;
}

inline bool readable(const std::string &f) {
  if (FILE *fp = std::fopen(f.c_str(), "r")) {
    std::fclose(fp);
    return true;
  }
  return false;
}

// First file of a set: the name itself, name.0.hdf5 or name.hdf5; "" if none.
inline std::string first_file(const std::string &name) {
  for (const std::string &f : {name, name + ".0.hdf5", name + ".hdf5"})
    if (is_hdf5(f)) return f;
  return std::string();
}

// The files of a set.  The later files replace the number in ".0.hdf5";
// a first file without ".0.hdf5" in its name is a set of one (the
// reference's search there only looks for a name with its first character
// replaced, read_fieldize.cpp:152-165).
inline std::vector<std::string> find_set(const std::string &name) {
  std::vector<std::string> files;
  const std::string first = first_file(name);
  if (first.empty()) return files;
  files.push_back(first);
  const size_t at = first.find(".0.hdf5");
  if (at == std::string::npos) return files;
  for (int n = 1; n < 1000; n++) {
    std::ostringstream f;
    f << first.substr(0, at + 1) << n << first.substr(at + 2);
    if (!(readable(f.str()) && is_hdf5(f.str()))) break;
    files.push_back(f.str());
  }
  return files;
}

struct Header {
  double time = 0, redshift = 0, box = 0, h100 = 0, omega0 = 0, omegal = 0;
  int flag_cooling = 0;
  int64_t npart_total[N_TYPE] = {0, 0, 0, 0, 0, 0};
  double mass[N_TYPE] = {0, 0, 0, 0, 0, 0};
};

// Header of the first file; false (after the reference's message) when a
// required attribute is missing.
inline bool load_header(const std::string &file, Header &h) {
  const hid_t f = H5Fopen(file.c_str(), H5F_ACC_RDONLY, H5P_DEFAULT);
  if (f < 0) return false;
  const hid_t g = H5Gopen2(f, "/Header", H5P_DEFAULT);
  if (g < 0) {
    H5Fclose(f);
    return false;
  }
  if (H5LTget_attribute_double(g, ".", "Time", &h.time) || H5LTget_attribute_double(g, ".", "Redshift", &h.redshift) ||
      H5LTget_attribute_double(g, ".", "BoxSize", &h.box) || H5LTget_attribute_double(g, ".", "HubbleParam", &h.h100) ||
      H5LTget_attribute_double(g, ".", "Omega0", &h.omega0) ||
      H5LTget_attribute_double(g, ".", "OmegaLambda", &h.omegal) ||
      H5LTget_attribute_int(g, ".", "Flag_Cooling", &h.flag_cooling)) {
    std::fprintf(stderr, "Failed to read some header value\n");
    H5Gclose(g);
    H5Fclose(f);
    return false;
  }
  unsigned int np[N_TYPE] = {0, 0, 0, 0, 0, 0}, hw[N_TYPE] = {0, 0, 0, 0, 0, 0};
  H5LTget_attribute(g, ".", "NumPart_Total", H5T_NATIVE_UINT, np);
  H5LTget_attribute(g, ".", "NumPart_Total_HighWord", H5T_NATIVE_UINT, hw);
  H5LTget_attribute(g, ".", "MassTable", H5T_NATIVE_DOUBLE, h.mass);
  for (int t = 0; t < N_TYPE; t++) h.npart_total[t] = (int64_t)np[t] + ((int64_t)hw[t] << 32);
  H5Gclose(g);
  H5Fclose(f);
  return true;
}

// Particles of one type in one file.
struct FileTypeData {
  int64_t n = 0;              // particles read
  double mass_table = 0;      // this file's MassTable entry for the type
  std::vector<float> pos;     // [n][3]
  std::vector<float> masses;  // [n] when mass_table == 0
};

// Dataset of the given rank and 4-byte float class: its first dimension
// (and the second, rank 2) through dims; false when absent or different.
inline bool float_dataset(hid_t g, const char *name, int want_rank, hsize_t *dims) {
  int rank = 0;
  if (H5LTget_dataset_ndims(g, name, &rank) < 0 || rank != want_rank) return false;
  H5T_class_t cls;
  size_t size = 0;
  if (H5LTget_dataset_info(g, name, dims, &cls, &size) < 0) return false;
  return cls == H5T_FLOAT && size == 4;
}

// Reads type `type` of one file (file number `fileno` for messages).
// Returns 0 on success (n = 0 when the file holds none of the type), -1
// when the file or its data cannot be read (the caller skips the file).
inline int read_type(const std::string &file, int type, int fileno, FileTypeData &out) {
  out = FileTypeData();
  const hid_t f = H5Fopen(file.c_str(), H5F_ACC_RDONLY, H5P_DEFAULT);
  if (f < 0) return -1;
  hid_t g = H5Gopen2(f, "/Header", H5P_DEFAULT);
  if (g < 0) {
    H5Fclose(f);
    return -1;
  }
  int np[N_TYPE] = {0, 0, 0, 0, 0, 0};
  double omega0 = 0, mass[N_TYPE] = {0, 0, 0, 0, 0, 0};
  if (H5LTget_attribute(g, ".", "NumPart_ThisFile", H5T_NATIVE_INT, np) ||
      H5LTget_attribute_double(g, ".", "Omega0", &omega0) ||
      H5LTget_attribute(g, ".", "MassTable", H5T_NATIVE_DOUBLE, mass)) {
    std::fprintf(stderr, "Failed to read some header value\n");
    H5Gclose(g);
    H5Fclose(f);
    return -1;
  }
  H5Gclose(g);
  out.mass_table = mass[type];
  if (np[type] <= 0) {
    H5Fclose(f);
    return 0;
  }
  char name[16];
  std::snprintf(name, sizeof name, "/PartType%d", type);
  g = H5Gopen2(f, name, H5P_DEFAULT);
  if (g < 0) {
    H5Fclose(f);
    return -1;
  }
  int ret = 0;
  hsize_t dims[2] = {0, 0};
  // (the reference deposits NumPart_ThisFile particles whatever the length;
  // a different length is taken as a failed read here)
  if (!float_dataset(g, "Coordinates", 2, dims) || dims[1] != 3 || dims[0] != (hsize_t)np[type]) {
    std::fprintf(stderr, "File %d: Failed reading Coordinates (%lu)\n", fileno, (unsigned long)dims[0]);
    ret = -1;
  } else {
    out.n = (int64_t)dims[0];
    out.pos.resize((size_t)out.n * 3);
    if (H5LTread_dataset_float(g, "Coordinates", out.pos.data()) < 0) ret = -1;
    if (ret == 0 && mass[type] == 0) {
      hsize_t md[1] = {0};
      out.masses.resize((size_t)out.n);
      if (!float_dataset(g, "Masses", 1, md) || md[0] != (hsize_t)out.n ||
          H5LTread_dataset_float(g, "Masses", out.masses.data()) < 0) {
        std::fprintf(stderr, "File %d: Failed reading Masses (%lu)\n", fileno, (unsigned long)md[0]);
        ret = -1;
      }
    }
  }
  if (ret) out = FileTypeData();
  H5Gclose(g);
  H5Fclose(f);
  return ret;
}

}  // namespace hdf5snap
