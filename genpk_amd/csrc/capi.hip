// C ABI of the library and the drop-in replacements of the reference's
// host-pointer call surface (gen-pk.h:93,104,119 and the FFTW step of
// gen-pk.cpp:193,233).
#include "fft_core.h"
#include "plan.h"
#include "../../include/genpk_amd.h"
#include "../../include/genpk_compat.h"

#include <cmath>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>

namespace genpk {

static thread_local std::string g_last_error;

void set_last_error(const std::string &msg) { g_last_error = msg; }

static int bits_for(int64_t k) {
  int b = 0;
  while (((int64_t)1 << b) < k) b++;
  return b;
}

Plan::Plan(int64_t dims, int precision, hipStream_t s) {
  GENPK_REQUIRE(dims >= 1 && dims <= 4096, "dims must be in [1, 4096]");
  GENPK_REQUIRE(precision == 4 || precision == 8, "precision must be GENPK_FLOAT32 (4) or GENPK_FLOAT64 (8)");
  d = dims;
  prec = precision;
  stream = s;
  GENPK_HIP(hipGetDevice(&device));
  hz = d / 2 + 1;
  rowlen = 2 * hz;
  field_elems = (size_t)d * d * rowlen;
  plan_set_tiles(*this, 16, 32, 32);
  plan_set_fft_variant(*this, default_fft_variant(prec));
  field.ensure(field_elems * prec);
  const bool fftable = (d % 2 == 0) && fft_length_ok(d) && fft_length_ok(d / 2) && d >= 2;
  if (fftable) fft_tables_init(*this);
}

void plan_set_tiles(Plan &p, int tx, int ty, int tz) {
  auto pow2 = [](int v) { return v >= 1 && (v & (v - 1)) == 0; };
  GENPK_REQUIRE(pow2(tx) && pow2(ty) && pow2(tz) && tz >= 2, "tile edges must be powers of two (z edge >= 2)");
  GENPK_REQUIRE((int64_t)(tx + 1) * (ty + 1) * (tz + 1) * 8 <= 160 * 1024,
                "a halo'd tile of 64-bit accumulators must fit in LDS (160 KB)");
  const int64_t d = p.d;
  TileGeom t;
  // each edge: the requested power of two, at most the largest power of two <= d
  auto clamp = [d](int cap, int &lg) {
    int64_t e = 1;
    lg = 0;
    while (e * 2 <= std::min<int64_t>(cap, d)) {
      e *= 2;
      lg++;
    }
    return (int)e;
  };
  t.TX = clamp(tx, t.sx);
  t.TY = clamp(ty, t.sy);
  t.TZ = clamp(tz, t.sz);
  t.ntx = (int)ceil_div(d, t.TX);
  t.nty = (int)ceil_div(d, t.TY);
  t.ntz = (int)ceil_div(d, t.TZ);
  t.HX = t.TX + 1;
  t.HY = t.TY + 1;
  t.HZ = t.TZ + 1;
  t.K = (int64_t)t.ntx * t.nty * t.ntz;
  t.cells = (int64_t)t.HX * t.HY * t.HZ;
  t.cm = (int64_t)t.HX * t.HY * t.TZ;
  t.cz = (int64_t)t.HX * t.HY;
  t.kbits = bits_for(t.K);
  if (t.kbits <= 9) {
    t.passes = 1;
    t.hi_bits = t.kbits;
    t.lo_bits = 0;
  } else {
    t.passes = 2;
    t.lo_bits = (t.kbits + 1) / 2;
    t.hi_bits = t.kbits - t.lo_bits;
  }
  GENPK_REQUIRE(t.hi_bits <= 11 && t.lo_bits <= 11, "tile key too wide: use larger tiles");
  if (p.tiles.ptr) GENPK_HIP(hipStreamSynchronize(p.stream));  // queued kernels may still use the store
  p.tg = t;
  p.tiles.release();  // sized for the old tiling
  p.tiles_live = false;
  p.order.valid = false;
}

// Variant 3 takes one radix 4 first in the column passes (n = d) only; the
// z passes (n = d/2) keep ordering 0.
void plan_set_fft_variant(Plan &p, int variant) {
  GENPK_REQUIRE(variant >= 0 && variant < FFT_VARIANTS, "unknown FFT radix variant");
  p.fft_variant = variant;
  p.rp_d = make_radix_plan((int)p.d, variant);
  p.rp_h = make_radix_plan((int)std::max<int64_t>(p.d / 2, 1), variant == 3 ? 0 : variant);
}

bool plan_fft_ok(const Plan &p) {
  return (p.d % 2 == 0) && p.d >= 2 && fft_length_ok(p.d) && fft_length_ok(p.d / 2);
}


// ---- cached host-surface plans for the compatibility shims ----
// Resident mode (off unless GENPK_COMPAT_RESIDENT=1 or genpk_compat_resident
// turns it on): the grid of the shims stays in the compat plan between
// calls on the same host array.  State:
//   Density: `host`'s contents were loaded into the plan's grid and later
//            fieldize calls added deposits there (the host array lacks them);
//   Spectrum: the plan's grid holds the spectrum the last in-place FFT of
//            `host` wrote back to it (both copies equal).
enum class Resident { None, Density, Spectrum };

struct CompatCache {
  std::mutex mu;
  std::map<int64_t, std::unique_ptr<Plan>> a, b;
  DevBuf pos, masses, grid;
  DevBuf pw, ke, ct;
  bool resident_mode = false;
  Resident res = Resident::None;
  double *res_host = nullptr;
  int64_t res_dims = 0;
};
static CompatCache &compat() {
  static CompatCache c;
  static bool init = [] {
    const char *e = std::getenv("GENPK_COMPAT_RESIDENT");
    c.resident_mode = e && e[0] == '1';
    return true;
  }();
  (void)init;
  return c;
}

static Plan &compat_plan(std::map<int64_t, std::unique_ptr<Plan>> &m, int64_t dims);

// Bring the host array of a resident density up to date (deposits merged
// into the grid, then copied back) and leave resident mode's state empty.
static void flush_resident(CompatCache &cc) {
  if (cc.res == Resident::Density) {
    Plan &pa = compat_plan(cc.a, cc.res_dims);
    tiles_to_field(pa);
    GENPK_HIP(hipStreamSynchronize(pa.stream));
    GENPK_HIP(hipMemcpy(cc.res_host, pa.field.ptr, pa.field_elems * 8, hipMemcpyDeviceToHost));
  }
  cc.res = Resident::None;
  cc.res_host = nullptr;
  cc.res_dims = 0;
}
static Plan &compat_plan(std::map<int64_t, std::unique_ptr<Plan>> &m, int64_t dims) {
  auto it = m.find(dims);
  if (it != m.end()) return *it->second;
  auto p = std::make_unique<Plan>(dims, 8, (hipStream_t) nullptr);
  Plan &r = *p;
  m[dims] = std::move(p);
  return r;
}

}  // namespace genpk

using namespace genpk;

extern "C" {

const char *genpk_last_error(void) { return g_last_error.c_str(); }

int genpk_plan_create(genpk_plan **plan, int64_t dims, int precision, void *stream) {
  return guarded([&] {
    GENPK_REQUIRE(plan != nullptr, "null plan pointer");
    auto h = std::make_unique<genpk_plan>();
    h->p = std::make_unique<Plan>(dims, precision, (hipStream_t)stream);
    *plan = h.release();
  });
}

int genpk_plan_destroy(genpk_plan *plan) {
  return guarded([&] {
    if (!plan) return;
    if (plan->p) GENPK_HIP(hipStreamSynchronize(plan->p->stream));
    delete plan;
  });
}

int genpk_plan_set_stream(genpk_plan *plan, void *stream) {
  return guarded([&] {
    GENPK_REQUIRE(plan && plan->p, "null plan");
    plan->p->stream = (hipStream_t)stream;
  });
}

int genpk_plan_field_bytes(const genpk_plan *plan, size_t *bytes) {
  return guarded([&] {
    GENPK_REQUIRE(plan && plan->p && bytes, "null argument");
    *bytes = plan->p->field_elems * plan->p->prec;
  });
}

int genpk_plan_field_ptr(genpk_plan *plan, void **ptr) {
  return guarded([&] {
    GENPK_REQUIRE(plan && plan->p && ptr, "null argument");
    *ptr = plan->p->field.ptr;
  });
}

int genpk_plan_attach_field(genpk_plan *plan, void *ptr) {
  return guarded([&] {
    GENPK_REQUIRE(plan && plan->p && ptr, "null argument");
    Plan &p = *plan->p;
    GENPK_HIP(hipStreamSynchronize(p.stream));
    p.field.attach(ptr, p.field_elems * p.prec);
    p.fstate = FieldState::None;
  });
}

int genpk_plan_set_options(genpk_plan *plan, int tile_x, int tile_y, int tile_z, int fft_variant) {
  return guarded([&] {
    GENPK_REQUIRE(plan && plan->p, "null plan");
    Plan &p = *plan->p;
    GENPK_REQUIRE(fft_variant >= 0 && fft_variant < FFT_VARIANTS, "unknown FFT radix variant");
    plan_set_tiles(p, tile_x, tile_y, tile_z);
    plan_set_fft_variant(p, fft_variant);
  });
}

int genpk_radix_plan(int64_t n, int fft_variant, int *radices, int max, int *count) {
  return guarded([&] {
    GENPK_REQUIRE(radices && count && max >= 0, "null argument");
    GENPK_REQUIRE(fft_variant >= -1 && fft_variant < FFT_VARIANTS, "unknown FFT radix variant");
    GENPK_REQUIRE(fft_length_ok(n), "transform length must be 2^a 3^b 5^c, at most 2^20");
    const RadixPlan rp = fft_variant < 0 ? radix_plan_8_to_2((int)n) : make_radix_plan((int)n, fft_variant);
    GENPK_REQUIRE(rp.nst >= 0, "no ordering with a first radix 8, a last radix 2 and an even number of stages "
                               "for this length");
    GENPK_REQUIRE(rp.nst <= max, "radices buffer too short");
    for (int s = 0; s < rp.nst; s++) radices[s] = rp.st[s].R;
    *count = rp.nst;
  });
}

int genpk_plan_get_options(const genpk_plan *plan, int *tile_x, int *tile_y, int *tile_z, int *fft_variant) {
  return guarded([&] {
    GENPK_REQUIRE(plan && plan->p && tile_x && tile_y && tile_z && fft_variant, "null argument");
    const Plan &p = *plan->p;
    *tile_x = p.tg.TX;
    *tile_y = p.tg.TY;
    *tile_z = p.tg.TZ;
    *fft_variant = p.fft_variant;
  });
}

int genpk_clear(genpk_plan *plan) {
  return guarded([&] {
    GENPK_REQUIRE(plan && plan->p, "null plan");
    plan->p->tiles_live = false;
    plan->p->fstate = FieldState::None;
  });
}

int genpk_deposit(genpk_plan *plan, double boxsize, int64_t n, const float *pos, const float *masses, double mass) {
  return guarded([&] {
    GENPK_REQUIRE(plan && plan->p, "null plan");
    Plan &p = *plan->p;
    GENPK_REQUIRE(density_state(p), "deposit after a transform: clear the plan first");
    deposit(p, boxsize, n, pos, masses, mass);
  });
}

int genpk_deposit_retain(genpk_plan *plan, double boxsize, int64_t n, const float *pos, const float *masses,
                         double mass) {
  return guarded([&] {
    GENPK_REQUIRE(plan && plan->p, "null plan");
    Plan &p = *plan->p;
    GENPK_REQUIRE(density_state(p), "deposit after a transform: clear the plan first");
    deposit(p, boxsize, n, pos, masses, mass, true);
  });
}

int genpk_redeposit(genpk_plan *plan) {
  return guarded([&] {
    GENPK_REQUIRE(plan && plan->p, "null plan");
    Plan &p = *plan->p;
    GENPK_REQUIRE(density_state(p), "deposit after a transform: clear the plan first");
    redeposit(p);
  });
}

int genpk_drop_order(genpk_plan *plan) {
  return guarded([&] {
    GENPK_REQUIRE(plan && plan->p, "null plan");
    plan->p->order.valid = false;
  });
}

int genpk_export_density(genpk_plan *plan, void *out, int64_t out_rowlen, int accumulate) {
  return guarded([&] {
    GENPK_REQUIRE(plan && plan->p && out, "null argument");
    Plan &p = *plan->p;
    GENPK_REQUIRE(density_state(p), "the plan holds no density");
    GENPK_REQUIRE(out_rowlen >= p.d, "out_rowlen shorter than dims");
    if (p.fstate == FieldState::Real) {
      // imported density plus deposits: assemble in the plan's grid first
      tiles_to_field(p);
      const size_t row_bytes = (size_t)p.d * p.prec;
      if (accumulate) {
        GENPK_REQUIRE(false, "accumulating export of an imported density is not supported");
      }
      GENPK_HIP(hipMemcpy2DAsync(out, out_rowlen * p.prec, p.field.ptr, p.rowlen * p.prec, row_bytes,
                                 (size_t)(p.d * p.d), hipMemcpyDeviceToDevice, p.stream));
    } else {
      fold_tiles(p, out, out_rowlen, accumulate != 0);
    }
  });
}

int genpk_import_density(genpk_plan *plan, const void *grid) {
  return guarded([&] {
    GENPK_REQUIRE(plan && plan->p && grid, "null argument");
    Plan &p = *plan->p;
    GENPK_HIP(hipMemcpyAsync(p.field.ptr, grid, p.field_elems * p.prec, hipMemcpyDeviceToDevice, p.stream));
    p.fstate = FieldState::Real;
    p.tiles_live = false;
  });
}

int genpk_import_spectrum(genpk_plan *plan, const void *spectrum) {
  return guarded([&] {
    GENPK_REQUIRE(plan && plan->p && spectrum, "null argument");
    Plan &p = *plan->p;
    if (spectrum != p.field.ptr)
      GENPK_HIP(hipMemcpyAsync(p.field.ptr, spectrum, p.field_elems * p.prec, hipMemcpyDeviceToDevice, p.stream));
    p.fstate = FieldState::Spectrum;
    p.tiles_live = false;
  });
}

int genpk_r2c(genpk_plan *plan) {
  return guarded([&] {
    GENPK_REQUIRE(plan && plan->p, "null plan");
    Plan &p = *plan->p;
    GENPK_REQUIRE(plan_fft_ok(p), "the FFT needs an even dims with dims/2 = 2^a 3^b 5^c");
    GENPK_REQUIRE(density_state(p), "the plan holds no density");
    r2c(p);
  });
}

int genpk_c2r(genpk_plan *plan) {
  return guarded([&] {
    GENPK_REQUIRE(plan && plan->p, "null plan");
    GENPK_REQUIRE(plan_fft_ok(*plan->p), "the FFT needs an even dims with dims/2 = 2^a 3^b 5^c");
    c2r(*plan->p);
  });
}

int genpk_bin(genpk_plan *a, genpk_plan *b, int nrbins, double *power_aa, double *power_bb, double *power_ab,
              int *count, double *keffs, double ma, double mb) {
  return guarded([&] {
    GENPK_REQUIRE(a && a->p, "null plan");
    bin_spectra(*a->p, b ? b->p.get() : nullptr, nrbins, false, power_aa, power_bb, power_ab, count, keffs, ma,
                b ? mb : ma);
  });
}

int genpk_power(genpk_plan *a, genpk_plan *b, int nrbins, double *power_aa, double *power_bb, double *power_ab,
                int *count, double *keffs, double ma, double mb) {
  return guarded([&] {
    GENPK_REQUIRE(a && a->p, "null plan");
    Plan &pa = *a->p;
    Plan *pb = b ? b->p.get() : nullptr;
    GENPK_REQUIRE(plan_fft_ok(pa), "the FFT needs an even dims with dims/2 = 2^a 3^b 5^c");
    GENPK_REQUIRE(density_state(pa), "plan a holds no density");
    if (pb) GENPK_REQUIRE(density_state(*pb), "plan b holds no density");
    (void)bin_tables(pa, nrbins);
    for (Plan *q : {&pa, pb}) {
      if (!q) continue;
      r2c_zpass(*q);
      c2c_ypass(*q, -1);
      q->fstate = FieldState::YDone;
    }
    bin_spectra(pa, pb, nrbins, true, power_aa, power_bb, power_ab, count, keffs, ma, pb ? mb : ma);
  });
}

int genpk_adjoint_density(genpk_plan *plan, int nrbins, const double *lambda, double m) {
  return guarded([&] {
    GENPK_REQUIRE(plan && plan->p && lambda, "null argument");
    GENPK_REQUIRE(plan_fft_ok(*plan->p), "the FFT needs an even dims with dims/2 = 2^a 3^b 5^c");
    adjoint_density(*plan->p, nrbins, lambda, m);
  });
}

int genpk_adjoint_gather(genpk_plan *plan, double boxsize, int64_t n, const float *pos, const float *masses,
                         double mass, float *grad_pos, float *grad_mass) {
  return guarded([&] {
    GENPK_REQUIRE(plan && plan->p, "null plan");
    adjoint_gather(*plan->p, boxsize, n, pos, masses, mass, grad_pos, grad_mass);
  });
}

int genpk_timing_enable(genpk_plan *plan, int enable) {
  return guarded([&] {
    GENPK_REQUIRE(plan && plan->p, "null plan");
    plan->p->timer.enabled = enable != 0;
  });
}

int genpk_timing_read(genpk_plan *plan, char *names, size_t names_len, double *ms, int *launches, int max,
                      int *count) {
  return guarded([&] {
    GENPK_REQUIRE(plan && plan->p && names && ms && launches && count, "null argument");
    Plan &p = *plan->p;
    GENPK_HIP(hipStreamSynchronize(p.stream));
    std::vector<std::string> order;
    std::map<std::string, std::pair<double, int>> acc;
    for (auto &m : p.timer.marks) {
      float t = 0;
      GENPK_HIP(hipEventElapsedTime(&t, m.a, m.b));
      if (!acc.count(m.name)) order.push_back(m.name);
      acc[m.name].first += t;
      acc[m.name].second += 1;
      p.timer.pool.push_back(m.a);
      p.timer.pool.push_back(m.b);
    }
    p.timer.marks.clear();
    size_t off = 0;
    int k = 0;
    for (auto &nm : order) {
      if (k >= max || off + nm.size() + 1 > names_len) break;
      std::memcpy(names + off, nm.c_str(), nm.size() + 1);
      off += nm.size() + 1;
      ms[k] = acc[nm].first;
      launches[k] = acc[nm].second;
      k++;
    }
    *count = k;
  });
}

// utils.cpp:35-42
int genpk_nexttwo(int n) {
  unsigned int i;
  n--;
  for (i = 1; i < sizeof(int) * 8; i <<= 1) n |= n >> i;
  return ++n;
}

// utils.cpp:57-72 (type numbers of the Gadget format: 0 gas, 1 halo/DM,
// 2 disk/neutrino, 4 stars)
int genpk_type_str(int type, char *out, size_t len) {
  const char *s = "xx";
  switch (type) {
    case 0: s = "by"; break;
    case 1: s = "DM"; break;
    case 2: s = "nu"; break;
    case 4: s = "st"; break;
    default: break;
  }
  if (!out || len < 3) {
    g_last_error = "type_str output buffer too small";
    return 1;
  }
  std::memcpy(out, s, 3);
  return 0;
}

// utils.cpp:7-21
int genpk_print_pk(const char *filename, int nrbins, const double *keffs, const double *power, const int *count) {
  FILE *fd = std::fopen(filename, "w");
  if (!fd) {
    g_last_error = std::string("Error opening file: ") + filename;
    std::fprintf(stderr, "Error opening file: %s\n", filename);
    return 1;
  }
  for (int i = 0; i < nrbins; i++)
    if (count[i]) std::fprintf(fd, "%e\t%e\t%d\n", keffs[i], power[i], count[i]);
  std::fclose(fd);
  return 0;
}

// ---------------- drop-in replacements (host pointers) ----------------

// fieldize.cpp:117-133, evaluated on the host exactly as the reference does
double invwindow(int64_t kx, int64_t ky, int64_t kz, int64_t n) {
  if (n == 0) return 0;
  auto one = [n](int64_t k) -> double { return k ? M_PI * k / (n * sin(M_PI * k / (float)n)) : 1.0; };
  const float iwx = (float)one(kx), iwy = (float)one(ky), iwz = (float)one(kz);
  const float prod = (iwx * iwy) * iwz;
  return (double)prod * (double)prod;
}

int powerspectrum(int64_t dims, double (*outfield)[2], double (*outfield2)[2], int nrbins, double *power, int *count,
                  double *keffs, double total_mass, double total_mass2) {
  CompatCache &cc = compat();
  std::lock_guard<std::mutex> lk(cc.mu);
  return guarded([&] {
    GENPK_REQUIRE(outfield && outfield2 && power && count && keffs, "null argument");
    Plan &pa = compat_plan(cc.a, dims);
    const size_t bytes = pa.field_elems * 8;
    const bool kept = cc.res == Resident::Spectrum && cc.res_host == (double *)outfield && cc.res_dims == dims &&
                      outfield2 == outfield;
    if (!kept) {
      flush_resident(cc);
      GENPK_HIP(hipMemcpy(pa.field.ptr, outfield, bytes, hipMemcpyHostToDevice));
    }
    pa.fstate = FieldState::Spectrum;
    pa.tiles_live = false;
    double *dp = static_cast<double *>(cc.pw.ensure((size_t)nrbins * 8 * 3));
    double *dk = static_cast<double *>(cc.ke.ensure((size_t)nrbins * 8));
    int *dc = static_cast<int *>(cc.ct.ensure((size_t)nrbins * 4));
    if (outfield2 == outfield) {
      // auto spectrum, normalised by total_mass*total_mass2 as the reference does
      bin_spectra(pa, nullptr, nrbins, false, dp, nullptr, nullptr, dc, dk, total_mass, total_mass2);
      GENPK_HIP(hipStreamSynchronize(pa.stream));
    } else {
      Plan &pb = compat_plan(cc.b, dims);
      GENPK_HIP(hipMemcpy(pb.field.ptr, outfield2, bytes, hipMemcpyHostToDevice));
      pb.fstate = FieldState::Spectrum;
      pb.tiles_live = false;
      bin_spectra(pa, &pb, nrbins, false, nullptr, nullptr, dp + 2 * nrbins, dc, dk, total_mass, total_mass2);
      GENPK_HIP(hipStreamSynchronize(pa.stream));
      dp += 2 * nrbins;
    }
    GENPK_HIP(hipMemcpy(power, dp, (size_t)nrbins * 8, hipMemcpyDeviceToHost));
    GENPK_HIP(hipMemcpy(count, dc, (size_t)nrbins * 4, hipMemcpyDeviceToHost));
    GENPK_HIP(hipMemcpy(keffs, dk, (size_t)nrbins * 8, hipMemcpyDeviceToHost));
  });
}

int genpk_fft_r2c_3d_inplace(int64_t dims, double *field) {
  CompatCache &cc = compat();
  std::lock_guard<std::mutex> lk(cc.mu);
  return guarded([&] {
    GENPK_REQUIRE(field, "null field");
    Plan &pa = compat_plan(cc.a, dims);
    GENPK_REQUIRE(plan_fft_ok(pa), "the FFT needs an even dims with dims/2 = 2^a 3^b 5^c");
    const size_t bytes = pa.field_elems * 8;
    if (!(cc.res == Resident::Density && cc.res_host == field && cc.res_dims == dims)) {
      flush_resident(cc);
      GENPK_HIP(hipMemcpy(pa.field.ptr, field, bytes, hipMemcpyHostToDevice));
      pa.fstate = FieldState::Real;
      pa.tiles_live = false;
    }
    // resident density: the loaded grid plus the deposits kept in the tiles
    r2c(pa);
    GENPK_HIP(hipStreamSynchronize(pa.stream));
    GENPK_HIP(hipMemcpy(field, pa.field.ptr, bytes, hipMemcpyDeviceToHost));
    cc.res = cc.resident_mode ? Resident::Spectrum : Resident::None;
    cc.res_host = cc.resident_mode ? field : nullptr;
    cc.res_dims = cc.resident_mode ? dims : 0;
  });
}

int genpk_compat_resident(int on) {
  CompatCache &cc = compat();
  std::lock_guard<std::mutex> lk(cc.mu);
  return guarded([&] {
    if (!on) flush_resident(cc);
    cc.resident_mode = on != 0;
  });
}

int genpk_compat_flush(void) {
  CompatCache &cc = compat();
  std::lock_guard<std::mutex> lk(cc.mu);
  return guarded([&] { flush_resident(cc); });
}

}  // extern "C"

// fieldize.cpp:46 (C++ linkage, as declared at gen-pk.h:93)
int fieldize(double boxsize, int dims, double *out, int64_t segment_particles, float *positions, float *masses,
             double mass, int extra) {
  CompatCache &cc = compat();
  std::lock_guard<std::mutex> lk(cc.mu);
  return guarded([&] {
    GENPK_REQUIRE(out != nullptr, "null grid");
    GENPK_REQUIRE(extra == 0 || extra == 1, "extra must be 0 or 1");
    const int64_t rowlen = 2 * (dims / 2 + extra);
    GENPK_REQUIRE(rowlen >= dims, "extra = 0 with odd dims gives overlapping rows; use extra = 1");
    Plan &p = compat_plan(cc.a, dims);
    const int64_t n = segment_particles;
    if (n <= 0) return;
    GENPK_REQUIRE(positions != nullptr, "null positions");
    // the previous call's deposit may still read the staging buffers
    GENPK_HIP(hipStreamSynchronize(p.stream));
    float *dpos = static_cast<float *>(cc.pos.ensure((size_t)n * 12));
    GENPK_HIP(hipMemcpy(dpos, positions, (size_t)n * 12, hipMemcpyHostToDevice));
    float *dm = nullptr;
    if (masses) {
      dm = static_cast<float *>(cc.masses.ensure((size_t)n * 4));
      GENPK_HIP(hipMemcpy(dm, masses, (size_t)n * 4, hipMemcpyHostToDevice));
    }
    if (cc.resident_mode && extra == 1) {
      // the grid stays in the plan: loaded once, then deposits accumulate in
      // its tiles until the in-place FFT or a flush
      if (!(cc.res == Resident::Density && cc.res_host == out && cc.res_dims == dims)) {
        flush_resident(cc);
        GENPK_HIP(hipMemcpy(p.field.ptr, out, p.field_elems * 8, hipMemcpyHostToDevice));
        p.fstate = FieldState::Real;
        p.tiles_live = false;
        cc.res = Resident::Density;
        cc.res_host = out;
        cc.res_dims = dims;
      }
      deposit(p, boxsize, n, dpos, dm, mass);
      return;
    }
    flush_resident(cc);
    const size_t gbytes = (size_t)dims * dims * rowlen * 8;
    void *dg = cc.grid.ensure(gbytes);
    GENPK_HIP(hipMemcpy(dg, out, gbytes, hipMemcpyHostToDevice));
    p.tiles_live = false;
    p.fstate = FieldState::None;
    deposit(p, boxsize, n, dpos, dm, mass);
    fold_tiles(p, dg, rowlen, true);
    p.tiles_live = false;
    GENPK_HIP(hipStreamSynchronize(p.stream));
    GENPK_HIP(hipMemcpy(out, dg, gbytes, hipMemcpyDeviceToHost));
  });
}
