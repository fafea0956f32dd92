// Reader of Gadget-2 snapshots (format II with 4-character block labels, and
// format I with the standard block order), single or multi-file, for the
// gen-pk driver.  Replaces what the reference takes from the GadgetReader
// library (GSnap: GetNpart, GetHeader, GetBlock with a type mask,
// GetBlockSize/GetBlockParts; used by read_fieldize.cpp:18-97 and
// gen-pk.cpp:150-166).
//
// Layout: every record is framed by 4-byte length markers.  Format II puts a
// 8-byte record (label + next block's framed size) before each block.  The
// 256-byte header holds npart[6] (int32), mass[6] (double), time, redshift,
// flag_sfr, flag_feedback, npartTotal[6] (uint32), flag_cooling, num_files,
// BoxSize, Omega0, OmegaLambda, HubbleParam, ...  Particle blocks list the
// types in order 0..5; the MASS block holds only the types whose header mass
// is 0.  Little-endian files only.
#pragma once
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

namespace gadget {

constexpr int N_TYPE = 6;

struct Header {
  int32_t npart[N_TYPE] = {0};
  double mass[N_TYPE] = {0};
  double time = 0, redshift = 0;
  int32_t flag_sfr = 0, flag_feedback = 0;
  uint32_t npart_total[N_TYPE] = {0};
  int32_t flag_cooling = 0, num_files = 1;
  double box = 0, omega0 = 0, omega_lambda = 0, hubble = 0;
};

struct Block {
  int64_t offset;  // file offset of the block's data (after its leading marker)
  int64_t size;    // bytes of data
};

struct File {
  std::string path;
  Header h;
  bool format2 = false;
  std::map<std::string, Block> blocks;
};

inline bool exists(const std::string &p) {
  FILE *f = std::fopen(p.c_str(), "rb");
  if (!f) return false;
  std::fclose(f);
  return true;
}

class Snapshot {
 public:
  // name: a file, or the stem of a set name.0, name.1, ... (also accepted:
  // name.0 itself).  Throws std::runtime_error on unreadable input.
  explicit Snapshot(const std::string &name) {
    std::vector<std::string> paths;
    std::string first;
    if (exists(name))
      first = name;
    else if (exists(name + ".0"))
      first = name + ".0";
    else
      throw std::runtime_error("cannot open snapshot " + name + " (nor " + name + ".0)");
    File f0 = open(first);
    const int nf = f0.h.num_files > 1 ? f0.h.num_files : 1;
    if (nf > 1 && first.size() > 2 && first.compare(first.size() - 2, 2, ".0") == 0) {
      const std::string stem = first.substr(0, first.size() - 2);
      files_.push_back(std::move(f0));
      for (int i = 1; i < nf; i++) files_.push_back(open(stem + "." + std::to_string(i)));
    } else {
      files_.push_back(std::move(f0));
    }
    for (int t = 0; t < N_TYPE; t++) {
      npart_total_[t] = 0;
      for (const File &f : files_) npart_total_[t] += f.h.npart[t];
    }
  }

  int num_files() const { return (int)files_.size(); }
  const Header &header() const { return files_[0].h; }
  int64_t npart(int type) const { return npart_total_[type]; }

  // Bytes per particle of a block (POS: 12 for float32 positions).
  int64_t bytes_per_particle(const std::string &label) const {
    int64_t bytes = 0, parts = 0;
    for (const File &f : files_) {
      auto it = f.blocks.find(label);
      if (it == f.blocks.end()) continue;
      bytes += it->second.size;
      parts += parts_in_block(f, label);
    }
    return parts ? bytes / parts : 0;
  }

  // Reads `count` records of `type` starting at the type's particle `start`
  // (counted across the files in order) from block `label` into out
  // (rec_bytes bytes each).  Returns the number read.
  int64_t read(const std::string &label, int type, int64_t start, int64_t count, void *out, int64_t rec_bytes) const {
    unsigned char *dst = static_cast<unsigned char *>(out);
    int64_t done = 0, base = 0;
    for (const File &f : files_) {
      const int64_t nt = f.h.npart[type];
      if (start + done < base + nt && done < count) {
        auto it = f.blocks.find(label);
        if (it == f.blocks.end()) throw std::runtime_error("block " + label + " missing in " + f.path);
        int64_t before = 0;  // records of earlier types in this block
        for (int t = 0; t < type; t++)
          if (in_block(f, label, t)) before += f.h.npart[t];
        const int64_t first = start + done - base;
        const int64_t take = std::min<int64_t>(nt - first, count - done);
        const int64_t off = it->second.offset + (before + first) * rec_bytes;
        if (off + take * rec_bytes > it->second.offset + it->second.size)
          throw std::runtime_error("block " + label + " shorter than its header says in " + f.path);
        FILE *fp = std::fopen(f.path.c_str(), "rb");
        if (!fp) throw std::runtime_error("cannot open " + f.path);
        const bool ok = fseeko(fp, off, SEEK_SET) == 0 &&
                        std::fread(dst + done * rec_bytes, (size_t)rec_bytes, (size_t)take, fp) == (size_t)take;
        std::fclose(fp);
        if (!ok) throw std::runtime_error("short read of " + label + " in " + f.path);
        done += take;
      }
      base += nt;
    }
    return done;
  }

 private:
  std::vector<File> files_;
  int64_t npart_total_[N_TYPE];

  static bool in_block(const File &f, const std::string &label, int t) {
    if (label == "MASS") return f.h.npart[t] > 0 && f.h.mass[t] == 0;
    return true;
  }
  static int64_t parts_in_block(const File &f, const std::string &label) {
    int64_t n = 0;
    for (int t = 0; t < N_TYPE; t++)
      if (in_block(f, label, t)) n += f.h.npart[t];
    return n;
  }

  static int32_t marker(FILE *fp, bool &eof) {
    int32_t m;
    eof = std::fread(&m, 4, 1, fp) != 1;
    return m;
  }

  static File open(const std::string &path) {
    File f;
    f.path = path;
    FILE *fp = std::fopen(path.c_str(), "rb");
    if (!fp) throw std::runtime_error("cannot open " + path);
    bool eof = false;
    int32_t m = marker(fp, eof);
    if (eof) {
      std::fclose(fp);
      throw std::runtime_error("empty file " + path);
    }
    f.format2 = (m == 8);
    if (!f.format2 && m != 256) {
      std::fclose(fp);
      throw std::runtime_error(path + " is not a little-endian Gadget-2 snapshot");
    }
    std::fseek(fp, 0, SEEK_SET);
    static const char *order1[] = {"HEAD", "POS ", "VEL ", "ID  ", "MASS"};
    int nblock = 0;
    for (;;) {
      std::string label;
      if (f.format2) {
        int32_t a = marker(fp, eof);
        if (eof) break;
        char lab[5] = {0};
        int32_t nxt = 0, b = 0;
        if (a != 8 || std::fread(lab, 1, 4, fp) != 4 || std::fread(&nxt, 4, 1, fp) != 1 ||
            std::fread(&b, 4, 1, fp) != 1 || b != 8) {
          std::fclose(fp);
          throw std::runtime_error("bad block label record in " + path);
        }
        label = lab;
      } else {
        if (nblock >= 5) break;
        label = order1[nblock];
      }
      int32_t s1 = marker(fp, eof);
      if (eof) break;
      const int64_t off = ftello(fp);
      if (fseeko(fp, s1, SEEK_CUR) != 0) break;
      int32_t s2 = marker(fp, eof);
      if (eof || s1 != s2) {
        std::fclose(fp);
        throw std::runtime_error("inconsistent record markers of block " + label + " in " + path);
      }
      f.blocks[label] = Block{off, s1};
      if (label == "HEAD") {
        if (s1 < 256) {
          std::fclose(fp);
          throw std::runtime_error("short header in " + path);
        }
        unsigned char h[256];
        const int64_t here = ftello(fp);
        fseeko(fp, off, SEEK_SET);
        if (std::fread(h, 1, 256, fp) != 256) {
          std::fclose(fp);
          throw std::runtime_error("short header in " + path);
        }
        fseeko(fp, here, SEEK_SET);
        std::memcpy(f.h.npart, h + 0, 24);
        std::memcpy(f.h.mass, h + 24, 48);
        std::memcpy(&f.h.time, h + 72, 8);
        std::memcpy(&f.h.redshift, h + 80, 8);
        std::memcpy(&f.h.flag_sfr, h + 88, 4);
        std::memcpy(&f.h.flag_feedback, h + 92, 4);
        std::memcpy(f.h.npart_total, h + 96, 24);
        std::memcpy(&f.h.flag_cooling, h + 120, 4);
        std::memcpy(&f.h.num_files, h + 124, 4);
        std::memcpy(&f.h.box, h + 128, 8);
        std::memcpy(&f.h.omega0, h + 136, 8);
        std::memcpy(&f.h.omega_lambda, h + 144, 8);
        std::memcpy(&f.h.hubble, h + 152, 8);
      }
      nblock++;
      // format I: MASS exists only when some type has no header mass
      if (!f.format2 && nblock == 4) {
        bool any = false;
        for (int t = 0; t < N_TYPE; t++) any |= f.h.npart[t] > 0 && f.h.mass[t] == 0;
        if (!any) break;
      }
    }
    std::fclose(fp);
    if (!f.blocks.count("HEAD")) throw std::runtime_error("no header in " + path);
    return f;
  }
};

}  // namespace gadget
