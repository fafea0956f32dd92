// Plan (context) of the power-spectrum library: geometry, device buffers and
// cached tables for one grid size and precision.
#pragma once
#include "common.h"
#include "fft_core.h"
#include <map>
#include <string>
#include <memory>
#include <vector>

namespace genpk {

// Grow-only device buffer owned by a plan.
struct DevBuf {
  void *ptr = nullptr;
  size_t bytes = 0;
  bool owned = true;
  DevBuf() = default;
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  ~DevBuf() { release(); }
  void release() {
    if (ptr && owned) (void)hipFree(ptr);
    ptr = nullptr;
    bytes = 0;
    owned = true;
  }
  void attach(void *p, size_t b) {
    release();
    ptr = p;
    bytes = b;
    owned = false;
  }
  void *ensure(size_t want) {
    if (want == 0) want = 16;
    if (want > bytes) {
      release();
      GENPK_HIP(hipMalloc(&ptr, want));
      bytes = want;
    }
    return ptr;
  }
  template <typename T> T *as() const { return static_cast<T *>(ptr); }
};

// Deposit tiling.  The grid is cut into tiles of TX x TY x TZ cells
// (clamped to d); each tile is stored with a one-cell halo on its high side
// in x, y and z ("halo'd tile", HX x HY x HZ = (TX+1)(TY+1)(TZ+1) cells,
// z fastest), so a tile receives every CIC corner of the particles whose
// floor cell lies inside it.  Halo cells are folded into the neighbouring
// tile's low face when the real-space grid is assembled.
struct TileGeom {
  int TX = 16, TY = 32, TZ = 32;   // powers of two
  int sx = 4, sy = 5, sz = 5;      // log2 of TX, TY, TZ
  int ntx = 1, nty = 1, ntz = 1;
  int HX = 17, HY = 33, HZ = 33;
  int64_t K = 1;        // number of tiles
  int64_t cells = 0;    // HX*HY*HZ: accumulator cells of one tile
  int64_t cm = 0;       // HX*HY*TZ: row-store values of one tile
  int64_t cz = 0;       // HX*HY: z-halo values of one tile
  int kbits = 0;        // bits of the tile key
  int passes = 1;       // radix passes (1 or 2)
  int lo_bits = 0;      // bits of the low digit (pass 2), 0 if one pass
  int hi_bits = 0;      // bits of the high digit (pass 1)
};

// Per-plan FFT twiddle tables.
struct FftTables {
  DevBuf tw_d_f64, tw_d_f32;      // exp(-2 pi i m / d), m < d
  DevBuf tw_h_f64, tw_h_f32;      // exp(-2 pi i m / (d/2)), m < d/2
};

// Bin tables for one nrbins: q -> bin and the data-independent mode counts
// and effective k.
struct BinTables {
  int nrbins = 0;
  DevBuf q2bin;     // int32 [qmax+1], -1 for q = 0
  DevBuf thr;       // int64 [nrbins+1]: smallest q >= 1 whose bin is >= b (qmax+1 past the end)
  DevBuf count;     // int32 [nrbins]
  DevBuf keffsum;   // double [nrbins], unnormalised sum of h*|k|
  float lq = 0;     // bin(q) ~ lq * log2(q)
  bool counts_ready = false;
};

enum class FieldState { None, Real, Spectrum, YDone };

// Particles partitioned by tile (plan scratch; see sort_by_tile): records
// of rw floats (x, y, z[, mass]), tile starts, work items, and optionally
// the permutations input -> sorted position.
struct TileOrder {
  const float *rec = nullptr;
  int rw = 0;
  bool has_mass = false, has_perm = false;
  const uint32_t *perm1 = nullptr, *perm2 = nullptr;
  const uint32_t *tile_start = nullptr, *item_start = nullptr, *part_start = nullptr;
  const uint32_t *item_tile = nullptr;  // tile of each work item (max_items entries)
  const uint32_t *nheavy = nullptr, *heavy = nullptr;
  // with masses: [0] the bits of the largest |mass|, [1] those of the
  // smallest nonzero |mass| (0xffffffff when every mass is 0)
  const uint32_t *mrange = nullptr;
  int64_t ks = 0, n = 0, max_items = 0;
  uint32_t item_max = 0;  // records per work item at most (deposit_item_max)
};

// Tile order kept from the last deposit made with retain = true, for the
// adjoint gather of the same particle arrays and for depositing them again
// (redeposit).  every_tile: the order's work items cover every tile (it was
// made by a deposit into an empty plan).
struct RetainedOrder {
  bool valid = false;
  const float *pos = nullptr, *masses = nullptr;
  int64_t n = 0;
  double boxsize = 0;
  double mass = 1;
  bool every_tile = false;
  TileOrder to;
};

// Stage timing: hipEvents recorded on the plan's stream around each stage
// while enabled; read back (and cleared) by genpk_timing_read().
struct StageTimer {
  bool enabled = false;
  std::vector<hipEvent_t> pool;
  struct Mark { std::string name; hipEvent_t a, b; };
  std::vector<Mark> marks;
  hipEvent_t get() {
    if (!pool.empty()) {
      hipEvent_t e = pool.back();
      pool.pop_back();
      return e;
    }
    hipEvent_t e;
    GENPK_HIP(hipEventCreate(&e));
    return e;
  }
  ~StageTimer() {
    for (auto &m : marks) {
      (void)hipEventDestroy(m.a);
      (void)hipEventDestroy(m.b);
    }
    for (auto e : pool) (void)hipEventDestroy(e);
  }
};

struct Plan {
  int64_t d = 0;
  int prec = 8;             // bytes per real: 4 or 8
  int device = 0;
  hipStream_t stream = nullptr;
  int64_t hz = 0;           // d/2 + 1
  int64_t rowlen = 0;       // 2*hz
  size_t field_elems = 0;   // d*d*rowlen reals
  TileGeom tg;
  int fft_variant = 0;      // radix ordering of the transforms (make_radix_plan)
  RadixPlan rp_d, rp_h;     // radix plans of length d (y, x passes) and d/2 (z passes)

  DevBuf field;             // padded grid, real or half-complex
  FieldState fstate = FieldState::None;
  DevBuf tiles;             // K*(cm + cz) reals: row store, then z-halo planes
  bool tiles_live = false;  // tiles hold deposits not yet folded

  // sort / deposit scratch
  DevBuf recA, recB, table1, table2, seginfo, tile_start, item_start, part_start, heavy, partials, item_tile;
  DevBuf permA, permB;      // sort permutations kept for the adjoint gather
  DevBuf dig2;              // low tile-key digit of each pass-1 output record (one byte)
  DevBuf pages;             // paged first sort pass: block digits, fills and lists
  DevBuf mrange;            // largest and smallest nonzero |mass| of the last sort (float bits)
  DevBuf nonfinite;         // flag of the last deposit: a particle position is not finite
  DevBuf scan_tmp;
  RetainedOrder order;

  // window factors: float(onedinvwindow(KVAL(k), d)), k < d
  DevBuf wtab;
  FftTables fft;
  std::map<int, std::unique_ptr<BinTables>> bins;
  DevBuf bin_part;          // per-workgroup partial bins
  DevBuf bin_tmp;           // power sums before normalisation
  StageTimer timer;

  Plan(int64_t dims, int precision, hipStream_t s);
};

// Deposit tile edges of a plan (powers of two; each clamped to the largest
// power of two <= d; TZ >= 2; the halo'd tile (TX+1)(TY+1)(TZ+1) of u64
// must fit in LDS).  Drops the plan's deposits and kept tile order.
void plan_set_tiles(Plan &p, int tx, int ty, int tz);
// Radix ordering of the plan's transforms (make_radix_plan's variants).
void plan_set_fft_variant(Plan &p, int variant);

// RAII stage scope: records start/stop events when timing is enabled.
struct Stage {
  Plan &p;
  hipEvent_t a = nullptr, b = nullptr;
  const char *name;
  Stage(Plan &pl, const char *nm) : p(pl), name(nm) {
    if (p.timer.enabled) {
      a = p.timer.get();
      b = p.timer.get();
      GENPK_HIP(hipEventRecord(a, p.stream));
    }
  }
  ~Stage() {
    if (a) {
      (void)hipEventRecord(b, p.stream);
      p.timer.marks.push_back({name, a, b});
    }
  }
};

// The field holds no density (None), a density (Real), or the forward state
// left by the fused power (YDone), which counts as no density.
inline bool density_state(const Plan &p) {
  return p.fstate == FieldState::None || p.fstate == FieldState::Real || p.fstate == FieldState::YDone;
}

struct FoldGeom;
FoldGeom fold_geom(const Plan &p);
// capi.hip: dims the FFT supports (even, dims/2 = 2^a 3^b 5^c)
bool plan_fft_ok(const Plan &p);

// deposit.hip
void deposit(Plan &p, double boxsize, int64_t n, const float *pos, const float *masses, double mass,
             bool retain = false);
// Deposit again the particles of the kept tile order (p.order) from its
// sorted records, skipping the partition.
void redeposit(Plan &p);
// Compute units of the current device.
int num_cus();
void fold_tiles(Plan &p, void *out, int64_t out_rowlen, bool accumulate);
void fill_zero(void *ptr, size_t bytes, hipStream_t s);
void tiles_to_field(Plan &p);
// scan.hip
void exclusive_scan_u32(const uint32_t *in, uint32_t *out, int64_t n, DevBuf &tmp, hipStream_t s);
// fft.hip
void fft_tables_init(Plan &p);
void r2c_zpass(Plan &p);
void c2c_ypass(Plan &p, int sign);
void c2c_xpass(Plan &p, int sign);
void c2r_zpass(Plan &p);
void r2c(Plan &p);
void c2r(Plan &p);
// bin.hip
BinTables &bin_tables(Plan &p, int nrbins);
// The bin tables of nrbins if an earlier call built them, else an error
// (builds nothing, allocates nothing).
const BinTables &built_bin_tables(const Plan &p, int nrbins);
// Bin the spectra (fused_xpass: the fields after the z and y passes, the x
// pass run in the binning).  scratch: caller memory for the per-workgroup
// partial bins (scratch_doubles long) instead of the plan's buffer; with
// it the bin tables must already be built and nothing is allocated.
void bin_spectra(Plan &a, Plan *b, int nrbins, bool fused_xpass, double *pow_aa, double *pow_bb,
                 double *pow_ab, int *count, double *keffs, double ma, double mb, double *scratch = nullptr,
                 size_t scratch_doubles = 0);
// Doubles of per-workgroup partial bins the fused binning of a (and b) with
// nrbins bins writes.
int64_t fused_partial_doubles(const Plan &a, const Plan *b, int nrbins);
// adjoint.hip
void adjoint_density(Plan &p, int nrbins, const double *lambda, double m);
void adjoint_gather(Plan &p, double boxsize, int64_t n, const float *pos, const float *masses, double mass,
                    float *grad_pos, float *grad_mass);

}  // namespace genpk

// The C ABI's opaque plan handle (include/genpk_amd.h).
struct genpk_plan {
  std::unique_ptr<genpk::Plan> p;
};
