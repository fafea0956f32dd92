// MP-Gadget bigfile snapshots for gen-pk, following the reference's bigfile
// path (read_fieldize_bigfile.cpp:7-125).  The reference reads them through
// the bigfile library (a submodule it does not ship); this reader restates
// the library's on-disk layout:
//
//   <snapshot>/<block>/header    text: "DTYPE: <f8", "NMEMB: 3", "NFILE: n",
//                                then per data file "%06X: rows : cksum : sysv"
//   <snapshot>/<block>/%06X      the rows of each data file, raw, in DTYPE
//   <snapshot>/<block>/attr-v2   text, one attribute per line:
//                                "name dtype nmemb HEX [#HUMANE ...]", HEX the
//                                attribute's bytes as two hex digits each
//
// dtype strings are a byte order ('<', '>', '=', '|' or none), a kind
// ('f', 'i', 'u') and an item size in bytes.
//
// The snapshot is a bigfile when its "Header" block opens (is_bigfile).  The
// header holds TotNumPart (6 integers), MassTable (6 floats), Time,
// HubbleParam, Omega0 and BoxSize; the redshift is 1/Time - 1
// (load_bigfile_header).  Per type t, block "t/Position" (NMEMB 3) is read
// over its first TotNumPart[t] rows as float64, then rounded to float32; when
// MassTable[t] is 0, block "t/Mass" is read as float32
// (read_fieldize_bigfile).
#pragma once

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

namespace bigfile {

constexpr int N_TYPE = 6;

struct Dtype {
  char kind = 0;     // 'f', 'i', 'u'
  int size = 0;      // bytes per item
  bool swap = false;  // stored big-endian
};

// Parses "<f8", ">i4", "f4", "=u8", ...; size 0 when not understood.
inline Dtype parse_dtype(const std::string &s) {
  Dtype d;
  size_t at = 0;
  if (!s.empty() && (s[0] == '<' || s[0] == '>' || s[0] == '=' || s[0] == '|')) {
    d.swap = s[0] == '>';
    at = 1;
  }
  if (at >= s.size()) return Dtype();
  d.kind = s[at];
  if (d.kind != 'f' && d.kind != 'i' && d.kind != 'u') return Dtype();
  d.size = std::atoi(s.c_str() + at + 1);
  if (!((d.kind == 'f' && (d.size == 4 || d.size == 8)) ||
        (d.kind != 'f' && (d.size == 1 || d.size == 2 || d.size == 4 || d.size == 8))))
    return Dtype();
  return d;
}

// One item of dtype d at p as a double (integers up to 2^53 exactly).
inline double item_value(const unsigned char *p, const Dtype &d) {
  unsigned char b[8];
  for (int i = 0; i < d.size; i++) b[i] = d.swap ? p[d.size - 1 - i] : p[i];
  if (d.kind == 'f') {
    if (d.size == 4) {
      float f;
      std::memcpy(&f, b, 4);
      return f;
    }
    double v;
    std::memcpy(&v, b, 8);
    return v;
  }
  uint64_t u = 0;
  for (int i = d.size - 1; i >= 0; i--) u = (u << 8) | b[i];
  if (d.kind == 'i' && d.size < 8 && (u >> (8 * d.size - 1)) & 1) u |= ~uint64_t(0) << (8 * d.size);
  return d.kind == 'i' ? (double)(int64_t)u : (double)u;
}

struct Block {
  std::string dir;
  Dtype dtype;
  int nmemb = 0;
  std::vector<int64_t> rows;  // rows of each data file
  int64_t size() const {
    int64_t s = 0;
    for (int64_t r : rows) s += r;
    return s;
  }
};

// Opens <snapshot>/<name>/header; false when absent or malformed.
inline bool open_block(const std::string &snapshot, const std::string &name, Block &b, std::string &err) {
  b = Block();
  b.dir = snapshot + "/" + name;
  std::ifstream f(b.dir + "/header");
  if (!f) {
    err = "cannot open block " + name;
    return false;
  }
  std::string key, dtype;
  int nfile = -1;
  if (!(f >> key >> dtype) || key != "DTYPE:" || !(f >> key >> b.nmemb) || key != "NMEMB:" || !(f >> key >> nfile) ||
      key != "NFILE:" || nfile < 0) {
    err = "malformed header of block " + name;
    return false;
  }
  b.dtype = parse_dtype(dtype);
  b.rows.assign((size_t)nfile, 0);
  for (int i = 0; i < nfile; i++) {
    std::string id, colon1, colon2;
    long long rows = 0;
    unsigned long ck = 0, sysv = 0;
    // "%06X: rows : cksum : sysv"
    if (!(f >> id >> rows >> colon1 >> ck >> colon2 >> sysv) || id.size() < 2 || id.back() != ':') {
      err = "malformed file list of block " + name;
      return false;
    }
    const long fid = std::strtol(id.substr(0, id.size() - 1).c_str(), nullptr, 16);
    if (fid < 0 || fid >= nfile || rows < 0) {
      err = "malformed file list of block " + name;
      return false;
    }
    b.rows[(size_t)fid] = rows;
  }
  return true;
}

// Reads rows [0, n) of a block as doubles (n * nmemb values), over its data
// files in order.
inline bool read_rows(const Block &b, int64_t n, std::vector<double> &out, std::string &err) {
  if (b.dtype.size == 0) {
    err = "unsupported dtype in " + b.dir;
    return false;
  }
  if (n > b.size()) {
    err = "block " + b.dir + " holds fewer rows than requested";
    return false;
  }
  out.resize((size_t)(n * b.nmemb));
  const int64_t rowbytes = (int64_t)b.nmemb * b.dtype.size;
  std::vector<unsigned char> buf;
  int64_t done = 0;
  for (size_t fid = 0; fid < b.rows.size() && done < n; fid++) {
    const int64_t take = std::min<int64_t>(b.rows[fid], n - done);
    if (take == 0) continue;
    char name[16];
    std::snprintf(name, sizeof name, "%06X", (unsigned)fid);
    FILE *fp = std::fopen((b.dir + "/" + name).c_str(), "rb");
    if (!fp) {
      err = "cannot open " + b.dir + "/" + name;
      return false;
    }
    buf.resize((size_t)(take * rowbytes));
    const size_t got = std::fread(buf.data(), 1, buf.size(), fp);
    std::fclose(fp);
    if (got != buf.size()) {
      err = "short data file " + b.dir + "/" + name;
      return false;
    }
    for (int64_t i = 0; i < take * b.nmemb; i++)
      out[(size_t)(done * b.nmemb + i)] = item_value(buf.data() + i * b.dtype.size, b.dtype);
    done += take;
  }
  return true;
}

// Attribute name -> values (as doubles) of a block's attr-v2 file.
inline bool read_attrs(const Block &b, std::map<std::string, std::vector<double>> &attrs, std::string &err) {
  std::ifstream f(b.dir + "/attr-v2");
  if (!f) {
    err = "cannot open the attributes of " + b.dir;
    return false;
  }
  std::string line;
  while (std::getline(f, line)) {
    std::istringstream ls(line);
    std::string name, dtype, hex;
    int nmemb = 0;
    if (!(ls >> name >> dtype >> nmemb >> hex)) continue;
    const Dtype d = parse_dtype(dtype);
    if (d.size == 0 || nmemb < 0 || hex.size() != (size_t)(2 * d.size * nmemb)) continue;  // e.g. strings
    std::vector<unsigned char> raw(hex.size() / 2);
    for (size_t i = 0; i < raw.size(); i++) raw[i] = (unsigned char)std::strtoul(hex.substr(2 * i, 2).c_str(), nullptr, 16);
    std::vector<double> v((size_t)nmemb);
    for (int i = 0; i < nmemb; i++) v[(size_t)i] = item_value(raw.data() + (size_t)i * d.size, d);
    attrs[name] = v;
  }
  return true;
}

// is_bigfile (read_fieldize_bigfile.cpp:7-20): the "Header" block opens.
inline bool is_bigfile(const std::string &snapshot) {
  Block b;
  std::string err;
  return open_block(snapshot, "Header", b, err);
}

struct Header {
  int64_t npart_total[N_TYPE] = {0, 0, 0, 0, 0, 0};
  double mass[N_TYPE] = {0, 0, 0, 0, 0, 0};
  double time = 0, redshift = 0, h100 = 0, omega0 = 0, box = 0;
};

// load_bigfile_header (read_fieldize_bigfile.cpp:23-60).
inline bool load_header(const std::string &snapshot, Header &h, std::string &err) {
  Block b;
  if (!open_block(snapshot, "Header", b, err)) return false;
  std::map<std::string, std::vector<double>> a;
  if (!read_attrs(b, a, err)) return false;
  auto get = [&](const char *name, int n, double *out) {
    auto it = a.find(name);
    if (it == a.end() || (int)it->second.size() < n) {
      err = std::string("Failed to read attr: ") + name;
      return false;
    }
    for (int i = 0; i < n; i++) out[i] = it->second[(size_t)i];
    return true;
  };
  double np[N_TYPE];
  if (!get("TotNumPart", N_TYPE, np) || !get("MassTable", N_TYPE, h.mass) || !get("Time", 1, &h.time) ||
      !get("HubbleParam", 1, &h.h100) || !get("Omega0", 1, &h.omega0) || !get("BoxSize", 1, &h.box))
    return false;
  for (int t = 0; t < N_TYPE; t++) h.npart_total[t] = (int64_t)np[t];
  h.redshift = 1 / h.time - 1;
  return true;
}

// Particles of one type: positions float32 [n][3] (read as float64 and
// rounded) and, for MassTable[type] == 0, masses float32 [n].
inline bool read_type(const std::string &snapshot, int type, const Header &h, std::vector<float> &pos,
                      std::vector<float> &masses, std::string &err) {
  const int64_t n = h.npart_total[type];
  pos.clear();
  masses.clear();
  if (n == 0) return true;
  Block b;
  std::vector<double> v;
  if (!open_block(snapshot, std::to_string(type) + "/Position", b, err)) return false;
  if (b.nmemb != 3) {
    err = "Position block of type " + std::to_string(type) + " does not hold 3 values per particle";
    return false;
  }
  if (!read_rows(b, n, v, err)) return false;
  pos.resize(v.size());
  for (size_t i = 0; i < v.size(); i++) pos[i] = (float)v[i];
  if (h.mass[type] == 0) {
    if (!open_block(snapshot, std::to_string(type) + "/Mass", b, err)) return false;
    if (b.nmemb != 1 || !read_rows(b, n, v, err)) {
      if (err.empty()) err = "Mass block of type " + std::to_string(type) + " is not one value per particle";
      return false;
    }
    masses.resize(v.size());
    for (size_t i = 0; i < v.size(); i++) masses[i] = (float)v[i];
  }
  return true;
}

}  // namespace bigfile
