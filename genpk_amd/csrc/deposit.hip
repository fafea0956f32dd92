// CIC deposit (the reference's fieldize(), fieldize.cpp:46-114) on gfx950.
//
// 1. Tile sort.  Particles are partitioned by the tile of their floor cell
//    (sort.hip).
// 2. Tile accumulation.  One workgroup per (tile, particle range) item
//    accumulates the 8 CIC weights of each particle into the halo'd tile in
//    LDS as 64-bit fixed point (ds_add_u64), then flushes the whole tile with
//    coalesced stores.  Tiles with more particles than the item cap
//    (deposit_item_max: DEPOSIT_ITEM_MAX, less when tiles are few) are
//    split; their items write their raw accumulators as partial tiles, which a
//    reduction sums exactly (fixed point) whatever records formed each item.
// 3. Fold.  Grid cell values are assembled from the tile store by adding the
//    halo cells of the previous tiles onto each tile's low faces: y and x
//    halos in place after the deposit (k_fold_axis), z halos by the readers
//    (the transform's z pass while loading, export with fold_value).
//
// Per-particle index and weight arithmetic follows fieldize.cpp:63-92:
// u = (double)pos * (d/box); i0 = floor(u) wrapped into [0,d), i1 = i0+1
// wrapped; f = u - floor(u); corner c has bit0 = x, bit1 = y, bit2 = z.
#include "fft_core.h"
#include "tiles.h"
#include "tilesort.h"

#include <algorithm>
#include <cmath>
#include <type_traits>

namespace genpk {

namespace {

constexpr int DEP_THREADS = 1024;
constexpr int FOLD_THREADS = 256;

// 2^36 * (a * b) rounded to nearest once, for a in [0, 1] and b = (a weight
// in [0, 1]) * 2^36: the integer lands in the low mantissa bits of
// a * b + 2^52, which lies in [2^52, 2^53), so masking off the exponent
// leaves it (one AND on the high word).  bias = 2^52 comes from the kernel
// arguments, so it stays in a scalar register and the product is one FMA.
__device__ __forceinline__ unsigned long long to_fixed(double a, double b, double bias) {
  const double t = fma(a, b, bias);
  return (unsigned long long)__double_as_longlong(t) & 0x000fffffffffffffull;
}

// round(a * b) as a two's complement u64 for |a * b| < 2^51: a * b + 1.5 * 2^52
// lies in [2^52, 2^53), where the double's ulp is 1 (one rounding, by the FMA)
__device__ __forceinline__ unsigned long long to_fixed_signed(double a, double b) {
  const double t = fma(a, b, 6755399441055744.0);
  return (unsigned long long)__double_as_longlong(t) - 0x4338000000000000ull;
}

struct DepParams {
  KeyGeom kg;
  int HX, HY, HZ;
  int64_t cells;       // accumulator cells HX*HY*HZ
  int64_t cm, cz, zoff;  // tile store layout
  int64_t ks;          // key space size (tile_start has ks+1 entries)
  int64_t max_items;   // entries of the item-to-tile table
  double mass;         // constant mass (u64 path)
  double bias;         // 2^52 (to_fixed)
  const uint32_t *mrange;  // per-particle masses: the sort's |mass| range (float bits)
  uint32_t *nonfinite;     // set to 1 when a particle's position is not finite
  int store;           // 1: write tiles, 0: add to tiles
  FastDiv fntz, fnty;  // tile index -> (tx, ty, tz) by multiply-high (tiles < 2^31)
  FastDiv fnblk;       // k_reduce_partials: unit -> (heavy tile, 256-cell block)
};

// Records are RW floats: x, y, z, then the mass when NM.  Without masses the
// weights accumulate in u64 fixed point with 2^36 per unit weight (times the
// constant mass at the flush).  With masses they accumulate as signed fixed
// point with 2^s per unit mass, s = 45 - e where every |mass| < 2^e: an item's
// <= 2^17 terms per cell stay below 2^62 and each term below 2^45.  The
// absolute error per term, 2^(e-46), stays below the constant-mass path's
// 2^-37 of a particle's mass when the masses span at most 2^10 (exponents of
// the largest and smallest nonzero |mass|); wider ranges and non-finite
// masses accumulate in FP64 LDS atomics instead.  STD: tiles of 16 x 32 x 32
// cells (every grid of d >= 32), whose strides are constants, so the 8 corner
// offsets fold into the LDS instructions' immediate offsets.
template <typename Real, int NM, int RW, bool STD>
__global__ void __launch_bounds__(DEP_THREADS) k_deposit(const float *__restrict__ rec,
                                                         const uint32_t *__restrict__ tile_start,
                                                         const uint32_t *__restrict__ item_start,
                                                         const uint32_t *__restrict__ item_tile,
                                                         const uint32_t *__restrict__ part_start, DepParams P,
                                                         Real *__restrict__ tiles, double *__restrict__ partials) {
  using Acc = unsigned long long;
  constexpr int UNR = 4;
  constexpr int BATCH = DEP_THREADS * UNR;
  extern __shared__ unsigned long long lds_raw[];
  Acc *acc = reinterpret_cast<Acc *>(lds_raw);
  const int TX = STD ? 16 : P.kg.TX, TY = STD ? 32 : P.kg.TY, TZ = STD ? 32 : P.kg.TZ;
  const int sy = STD ? 33 : P.HZ, sx = STD ? 33 * 33 : P.HY * P.HZ;
  const double scale = (NM == 1) ? 1.0 : P.mass * FIXED_INV;
  bool fixed = true;            // masses: fixed point (else FP64 atomics)
  double m2s = 1.0, m2inv = 1.0;  // masses: 2^s, 2^-s
  if constexpr (NM == 1) {
    const uint32_t mxb = P.mrange[0], mnb = P.mrange[1];
    fixed = mxb < 0x7f800000u && (mxb == 0u || (int)(mxb >> 23) - (int)(mnb >> 23) <= 10);
    const int sc = 45 - ((int)(mxb >> 23) - 126);  // every |mass| < 2^((mxb >> 23) - 126)
    m2s = ldexp(1.0, sc);
    m2inv = ldexp(1.0, -sc);
  }
  // a cell's value at the flush, which clears the cell for the next item
  // (every cell is read once, by one thread)
  auto cell = [&](int c) -> double {
    const Acc a = acc[c];
    acc[c] = 0;
    if constexpr (NM == 1)
      return fixed ? (double)(long long)a * m2inv : __longlong_as_double((long long)a);
    else
      return (double)a * scale;
  };
  float r[UNR][RW], nx[UNR][RW];
  auto load_batch = [&](uint32_t i0, uint32_t pe) {
#pragma unroll
    for (int u = 0; u < UNR; u++) {
      const uint32_t i = i0 + u * DEP_THREADS + threadIdx.x;
      if (i < pe) {
#pragma unroll
        for (int q = 0; q < RW; q++) nx[u][q] = rec[(int64_t)i * RW + q];
      }
    }
  };
  auto take_batch = [&]() {
#pragma unroll
    for (int u = 0; u < UNR; u++)
#pragma unroll
      for (int q = 0; q < RW; q++) r[u][q] = nx[u][q];
  };
  // Persistent over the items (item, item + gridDim.x, ...).  The next
  // item's first records load during this item's last batch and are in
  // registers before the flush: the flush's stores then drain while the next
  // item accumulates (a wait for loads issued after them would wait for the
  // stores too, the counter being in order).
  uint32_t item = blockIdx.x;
  ItemRange ir;
  bool have = item_range(item_tile, item_start, tile_start, P.ks, P.max_items, item, ir);
  if (have) load_batch(ir.pb, ir.pe);
  // the tile starts cleared; each flush clears the cells it reads
  for (int64_t c = threadIdx.x; c < P.cells; c += DEP_THREADS) acc[c] = 0;
  take_batch();
  __syncthreads();
  while (have) {
    const int64_t k = ir.k;
    const uint32_t m = ir.m, j = ir.j, pb = ir.pb, pe = ir.pe;
    const int ntz = P.kg.ntz, nty = P.kg.nty;
    // 32-bit multiply-high, not 64-bit divisions (~130 scalar instructions each)
    const uint32_t kxy = P.fntz.div((uint32_t)k), kx = P.fnty.div(kxy);
    const int tz = (int)((uint32_t)k - kxy * (uint32_t)ntz);
    const int ty = (int)(kxy - kx * (uint32_t)nty);
    const int tx = (int)kx;
    const int ox = tx * TX, oy = ty * TY, oz = tz * TZ;
    const int wx = min(TX, P.kg.d - ox), wy = min(TY, P.kg.d - oy), wz = min(TZ, P.kg.d - oz);  // the tile's own cells
    // the next item (uniform over the workgroup)
    const uint32_t nitem = item + gridDim.x;
    ItemRange nir;
    const bool next = item_range(item_tile, item_start, tile_start, P.ks, P.max_items, nitem, nir);
    bool pre = false;  // the next item's first records are loading
    for (uint32_t i0 = pb; i0 < pe; i0 += BATCH) {
      // loads of the next batch (this item's, or the next item's first) are
      // issued before this batch (in r) is accumulated; one load site (two
      // would merge their registers with copies that wait for the loads)
      const bool more = i0 + BATCH < pe;
      if (more || next) {
        load_batch(more ? i0 + BATCH : nir.pb, more ? pe : nir.pe);
        pre = !more;
      }
#pragma unroll
      for (int u = 0; u < UNR; u++) {
        const uint32_t i = i0 + u * DEP_THREADS + threadIdx.x;
        if (i >= pe) continue;
        // cell_of without its wrap: a floor cell inside the tile's own cells
        // [o, o + w) lies in [0, d), where cell_of does not wrap
        double fx, fy, fz;
        int lx, ly, lz;
        {
          const double ux = (double)r[u][0] * P.kg.units, uy = (double)r[u][1] * P.kg.units;
          const double uz = (double)r[u][2] * P.kg.units;
          const double flx = floor(ux), fly = floor(uy), flz = floor(uz);
          fx = ux - flx;
          fy = uy - fly;
          fz = uz - flz;
          lx = (int)flx - ox;
          ly = (int)fly - oy;
          lz = (int)flz - oz;
        }
        if ((unsigned)lx >= (unsigned)wx || (unsigned)ly >= (unsigned)wy || (unsigned)lz >= (unsigned)wz) {
          // a position outside [0, box) (its cells wrap), or a non-finite one
          lx = cell_of(r[u][0], P.kg.units, P.kg.d, fx) - ox;
          ly = cell_of(r[u][1], P.kg.units, P.kg.d, fy) - oy;
          lz = cell_of(r[u][2], P.kg.units, P.kg.d, fz) - oz;
          if ((unsigned)lx >= (unsigned)TX || (unsigned)ly >= (unsigned)TY || (unsigned)lz >= (unsigned)TZ)
            continue;  // cannot happen for records of this tile; guards the LDS bounds
        }
        if (!(fx + fy + fz <= 3.0)) {
          // a non-finite position: k_poison sets its eight cells to NaN
          *P.nonfinite = 1u;
          continue;
        }
        const double gx = 1.0 - fx, gy = 1.0 - fy, gz = 1.0 - fz;
        const double wxy[4] = {gx * gy, fx * gy, gx * fy, fx * fy};
        const double gzs = gz * FIXED_ONE, fzs = fz * FIXED_ONE;  // exact: a power-of-two scale
        // 24-bit products (one full-rate instruction each; every factor is below 2^24)
        const int base = (int)(__umul24((unsigned)lx, (unsigned)sx) + __umul24((unsigned)ly, (unsigned)sy)) + lz;
        const int off[4] = {0, sx, sy, sx + sy};
#pragma unroll
        for (int c = 0; c < 4; c++) {
          const int at = base + off[c];
          if constexpr (NM == 1) {
            const double mw = (double)r[u][3] * wxy[c];
            if (fixed) {
              atomicAdd(&acc[at], to_fixed_signed(mw, gz * m2s));
              atomicAdd(&acc[at + 1], to_fixed_signed(mw, fz * m2s));
            } else {
              atomicAdd(reinterpret_cast<double *>(&acc[at]), mw * gz);
              atomicAdd(reinterpret_cast<double *>(&acc[at + 1]), mw * fz);
            }
          } else {
            atomicAdd(&acc[at], to_fixed(wxy[c], gzs, P.bias));
            atomicAdd(&acc[at + 1], to_fixed(wxy[c], fzs, P.bias));
          }
        }
      }
      take_batch();
    }
    // an item without records: the next item's first records load here
    // (taken before the flush too, so that no load is left in flight across it)
    if (next && !pre) {
      load_batch(nir.pb, nir.pe);
      take_batch();
    }
    __syncthreads();
    if (m == 1) {
      // rows of TZ values (aligned), then the z-halo plane
      // (32-bit cell indices: a tile's accumulator holds fewer than 2^24 cells)
      Real *rows = tiles + k * P.cm;
      Real *zh = tiles + P.zoff + k * P.cz;
      const int sz = STD ? 5 : P.kg.sz, hz = STD ? 33 : P.HZ;
      const int cm = (int)P.cm, cz = (int)P.cz;
      if (P.store && STD && sizeof(Real) == 8) {
        // FP64: two cells of a row per thread (rows of 32 values, 16-byte
        // aligned pairs), half the store and LDS instructions (5.85 -> 5.76
        // ms at 2^30; FP32 measured slower with pairs: 5.52 -> 5.57 ms)
        using R2 = typename cplx_of<Real>::type;
        R2 *rows2 = reinterpret_cast<R2 *>(rows);
#pragma unroll 4
        for (int c2 = threadIdx.x; c2 < cm / 2; c2 += DEP_THREADS) {
          const int at = (int)__umul24((unsigned)(c2 >> 4), 33u) + 2 * (c2 & 15);
          R2 v;
          v.x = (Real)cell(at);
          v.y = (Real)cell(at + 1);
          rows2[c2] = v;
        }
        for (int c = threadIdx.x; c < cz; c += DEP_THREADS) zh[c] = (Real)cell((int)__umul24((unsigned)c, 33u) + 32);
      } else if (P.store) {
#pragma unroll 4
        for (int c = threadIdx.x; c < cm; c += DEP_THREADS)
          rows[c] = (Real)cell((int)__umul24((unsigned)(c >> sz), (unsigned)hz) + (c & (TZ - 1)));
        for (int c = threadIdx.x; c < cz; c += DEP_THREADS) zh[c] = (Real)cell((int)__umul24((unsigned)c, (unsigned)hz) + TZ);
      } else {
        for (int c = threadIdx.x; c < cm; c += DEP_THREADS)
          rows[c] = (Real)((double)rows[c] + cell((int)__umul24((unsigned)(c >> sz), (unsigned)hz) + (c & (TZ - 1))));
        for (int c = threadIdx.x; c < cz; c += DEP_THREADS)
          zh[c] = (Real)((double)zh[c] + cell((int)__umul24((unsigned)c, (unsigned)hz) + TZ));
      }
    } else {
      // a split tile: this item's raw accumulators (fixed-point integers, or
      // the FP64 sums), which k_reduce_partials adds up
      Acc *out = reinterpret_cast<Acc *>(partials) + (int64_t)(part_start[k] + j) * P.cells;
      for (int64_t c = threadIdx.x; c < P.cells; c += DEP_THREADS) {
        out[c] = acc[c];
        acc[c] = 0;
      }
    }
    __syncthreads();  // the flush has read and cleared the tile before the next item adds to it
    item = nitem;
    ir = nir;
    have = next;
  }
}

// Sum the partial tiles of split tiles into the tile store.  Fixed-point
// partials add exactly in a 128-bit integer and convert once, so a cell's
// value does not depend on which records of the tile formed which item (the
// order of the records within a tile is not fixed); FP64 partials (masses
// beyond the fixed-point range) add in item order.
template <typename Real, int NM>
__global__ void __launch_bounds__(256) k_reduce_partials(const uint32_t *__restrict__ nheavy,
                                                         const uint32_t *__restrict__ heavy,
                                                         const uint32_t *__restrict__ item_start,
                                                         const uint32_t *__restrict__ part_start,
                                                         const double *__restrict__ partials, DepParams P,
                                                         Real *__restrict__ tiles) {
  bool fixed = true;
  double scale = P.mass * FIXED_INV;  // constant mass: unsigned sums of 2^36 per unit weight
  if constexpr (NM == 1) {
    const uint32_t mxb = P.mrange[0], mnb = P.mrange[1];
    fixed = mxb < 0x7f800000u && (mxb == 0u || (int)(mxb >> 23) - (int)(mnb >> 23) <= 10);
    scale = ldexp(1.0, -(45 - ((int)(mxb >> 23) - 126)));  // signed sums of 2^s per unit mass
  }
  const uint64_t *raw = reinterpret_cast<const uint64_t *>(partials);
  const int64_t cells = P.cells;
  const int64_t nblk = ceil_div(cells, 256);
  const int64_t units = (int64_t)(*nheavy) * nblk;
  for (int64_t u = blockIdx.x; u < units; u += gridDim.x) {
    // units < 2^31 (heavy tiles < 2^22, blocks per tile < 2^8): 32-bit multiply-high
    const uint32_t h = P.fnblk.div((uint32_t)u);
    const uint32_t k = heavy[h];
    const int64_t c = (int64_t)((uint32_t)u - h * (uint32_t)nblk) * 256 + threadIdx.x;
    if (c >= cells) continue;
    const uint32_t m = item_start[k + 1] - item_start[k];
    const int64_t at = (int64_t)part_start[k] * cells + c;
    double s;
    if (fixed) {
      // 128-bit two's-complement sum (hi:lo); eight loads in flight per thread
      uint64_t lo = 0, hi = 0;
      auto add = [&](uint64_t v) {
        lo += v;
        hi += (lo < v ? 1u : 0u) + (NM == 1 && (int64_t)v < 0 ? ~0ull : 0ull);
      };
      uint32_t j = 0;
      for (; j + 8 <= m; j += 8) {
        uint64_t t[8];
#pragma unroll
        for (int q = 0; q < 8; q++) t[q] = raw[at + (int64_t)(j + q) * cells];
#pragma unroll
        for (int q = 0; q < 8; q++) add(t[q]);
      }
      for (; j < m; j++) add(raw[at + (int64_t)j * cells]);
      const bool neg = (int64_t)hi < 0;
      if (neg) {
        lo = ~lo + 1;
        hi = ~hi + (lo == 0 ? 1u : 0u);
      }
      s = ((double)hi * 18446744073709551616.0 + (double)lo) * scale;
      if (neg) s = -s;
    } else {
      // eight loads in flight per thread; the additions keep the order j = 0, 1, ...
      const double *pp = partials + at;
      s = 0;
      uint32_t j = 0;
      for (; j + 8 <= m; j += 8) {
        double t[8];
#pragma unroll
        for (int q = 0; q < 8; q++) t[q] = pp[(int64_t)(j + q) * cells];
#pragma unroll
        for (int q = 0; q < 8; q++) s += t[q];
      }
      for (; j < m; j++) s += pp[(int64_t)j * cells];
    }
    const int64_t row = c / P.HZ;
    const int lz = (int)(c - row * P.HZ);
    Real *out = lz < P.kg.TZ ? tiles + (int64_t)k * P.cm + row * P.kg.TZ + lz : tiles + P.zoff + (int64_t)k * P.cz + row;
    *out = P.store ? (Real)s : (Real)((double)*out + s);
  }
}

// Particles with a non-finite position (found by k_deposit, which skips
// them) make the reference's eight CIC weights non-finite (fieldize.cpp:
// 67-84: the fractional part is NaN or infinite), so their eight cells end
// up NaN or infinite there.  Here those cells are set to NaN in the tile
// store, at the cells cell_of_x86 gives (those of the reference on x86-64).
// The records are scanned only when k_deposit raised the flag.  A cell's
// own store location is written (never a halo), so the order against the
// halo folds does not matter, and later deposits keep it NaN.
template <typename Real>
__global__ void __launch_bounds__(256) k_poison(const float *__restrict__ rec, int rw, int64_t n, KeyGeom kg,
                                                FoldGeom fg, const uint32_t *__restrict__ nonfinite,
                                                Real *__restrict__ tiles) {
  if (*nonfinite == 0u) return;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const float *r = rec + i * rw;
    double fx, fy, fz;
    const int cx = cell_of_x86(r[0], kg.units, kg.d, fx);
    const int cy = cell_of_x86(r[1], kg.units, kg.d, fy);
    const int cz = cell_of_x86(r[2], kg.units, kg.d, fz);
    if (fx + fy + fz <= 3.0) continue;
    const int nx = cx + 1 == kg.d ? 0 : cx + 1, ny = cy + 1 == kg.d ? 0 : cy + 1, nz = cz + 1 == kg.d ? 0 : cz + 1;
#pragma unroll
    for (int c = 0; c < 8; c++)
      tiles[cell_at(fg, (c & 1) ? nx : cx, (c & 2) ? ny : cy, (c & 4) ? nz : cz)] = (Real)NAN;
  }
}

// In-place halo folds along y, then x.  fold_y: one workgroup per tile t adds
// the y-halo rows (local y = width of the previous tile along y) of that
// tile into t's rows at local y = 0, for every local x including t's x-halo
// row, and for the z-halo plane entries of those rows, zeroing the sources.
// fold_x then adds the previous tile along x's x-halo rows (every local y)
// into t's rows at local x = 0.  Corner contributions travel through both
// steps; z halos stay in place for the readers.  Each source row has one
// destination and sources and destinations never coincide (a width is >= 1),
// so workgroups do not race; zeroed sources let later deposits accumulate.
template <typename Real, int AXIS>
__global__ void __launch_bounds__(256) k_fold_axis(Real *__restrict__ tiles, FoldGeom P) {
  const int64_t t = blockIdx.x;
  const int tz = (int)(t % P.ntz);
  const int ty = (int)((t / P.ntz) % P.nty);
  const int tx = (int)(t / ((int64_t)P.ntz * P.nty));
  const int HX = P.TX + 1;
  int64_t src;
  int h;
  if (AXIS == 1) {
    const int typ = ty == 0 ? P.nty - 1 : ty - 1;
    h = min(P.TY, P.d - typ * P.TY);
    src = ((int64_t)tx * P.nty + typ) * P.ntz + tz;
  } else {
    const int txp = tx == 0 ? P.ntx - 1 : tx - 1;
    h = min(P.TX, P.d - txp * P.TX);
    src = ((int64_t)txp * P.nty + ty) * P.ntz + tz;
  }
  // rows: AXIS 1 -> (lx, h) onto (lx, 0), lx < HX; AXIS 0 -> (h, ly) onto (0, ly), ly < HY
  const int nrows = AXIS == 1 ? HX : P.HY;
  const int per_row = P.TZ + 1;  // TZ row values, then the z-halo plane entry
  for (int i = threadIdx.x; i < nrows * per_row; i += 256) {
    const int r = i / per_row, z = i - r * per_row;
    const int sl = AXIS == 1 ? r * P.HY + h : h * P.HY + r;   // source row (lx*HY + ly)
    const int dl = AXIS == 1 ? r * P.HY : r;                    // destination row
    const int64_t sa = z < P.TZ ? src * P.cm + (int64_t)sl * P.TZ + z : P.zoff + src * P.cz + sl;
    const int64_t da = z < P.TZ ? t * P.cm + (int64_t)dl * P.TZ + z : P.zoff + t * P.cz + dl;
    const Real v = tiles[sa];
    tiles[sa] = (Real)0;
    tiles[da] = (Real)((double)tiles[da] + (double)v);
  }
}

// Fold the halo'd tiles into a padded real grid [d][d][out_rowlen].
template <typename Real, typename Out>
__global__ void __launch_bounds__(FOLD_THREADS) k_fold(const Real *__restrict__ tiles, FoldGeom P, int64_t out_rowlen,
                                                       int accumulate, Out *__restrict__ out) {
  const int64_t line = blockIdx.y + (int64_t)blockIdx.z * gridDim.y;  // x*d + y
  const int z = blockIdx.x * FOLD_THREADS + threadIdx.x;
  if (z >= P.d || line >= (int64_t)P.d * P.d) return;
  const int x = (int)(line / P.d), y = (int)(line % P.d);
  const double v = fold_value(tiles, P, x, y, z);
  Out *o = out + line * out_rowlen + z;
  *o = accumulate ? (Out)((double)*o + v) : (Out)v;
}

// Persistent grid of the deposit: the workgroups that stay resident (one
// per CU with a 147 KB tile), at most one per item.
template <typename Real, int NM, int RW, bool STD>
int64_t deposit_grid_of(size_t lds, int64_t max_items) {
  int per_cu = 0;
  GENPK_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(
      &per_cu, reinterpret_cast<const void *>(&k_deposit<Real, NM, RW, STD>), DEP_THREADS, lds));
  return std::max<int64_t>(1, std::min<int64_t>((int64_t)std::max(per_cu, 1) * num_cus(), max_items));
}

template <int NM, int RW>
int64_t deposit_grid(const Plan &p, bool std_tiles, size_t lds, int64_t max_items) {
  if (p.prec == 4)
    return std_tiles ? deposit_grid_of<float, NM, RW, true>(lds, max_items)
                     : deposit_grid_of<float, NM, RW, false>(lds, max_items);
  return std_tiles ? deposit_grid_of<double, NM, RW, true>(lds, max_items)
                   : deposit_grid_of<double, NM, RW, false>(lds, max_items);
}

template <int NM, int RW>
void deposit_impl(Plan &p, const KeyGeom &kg, const TileOrder &o, double mass) {
  hipStream_t s = p.stream;
  const TileGeom &tg = p.tg;
  const int store = p.tiles_live ? 0 : 1;
  const float *sorted = o.rec;
  const uint32_t *tile_start = o.tile_start, *item_start = o.item_start, *part_start = o.part_start;
  const uint32_t *nheavy = o.nheavy, *heavy = o.heavy;
  const int64_t ks = o.ks, n = o.n, max_items = o.max_items;
  const int64_t max_parts = 2 * ceil_div(n, (int64_t)o.item_max);
  double *partials = static_cast<double *>(p.partials.ensure((size_t)max_parts * tg.cells * 8 + 64));

  DepParams P;
  P.kg = kg;
  P.HX = tg.HX; P.HY = tg.HY; P.HZ = tg.HZ;
  P.cells = tg.cells;
  P.cm = tg.cm;
  P.cz = tg.cz;
  P.zoff = tg.K * tg.cm;
  P.ks = ks;
  P.max_items = max_items;
  P.mass = mass;
  P.bias = 4503599627370496.0;
  P.mrange = o.mrange;
  P.nonfinite = static_cast<uint32_t *>(p.nonfinite.ensure(64));
  P.store = store;
  P.fntz = FastDiv((uint32_t)tg.ntz);
  P.fnty = FastDiv((uint32_t)tg.nty);
  P.fnblk = FastDiv((uint32_t)ceil_div(tg.cells, 256));
  fill_zero(P.nonfinite, 4, s);
  const size_t lds = (size_t)tg.cells * 8;
  {
    Stage st(p, "deposit_tiles");
    const bool std_tiles = tg.TX == 16 && tg.TY == 32 && tg.TZ == 32;
    const dim3 grid((unsigned)deposit_grid<NM, RW>(p, std_tiles, lds, max_items)), block(DEP_THREADS);
    if (p.prec == 4 && std_tiles)
      hipLaunchKernelGGL((k_deposit<float, NM, RW, true>), grid, block, lds, s, sorted, tile_start, item_start,
                         o.item_tile, part_start, P, p.tiles.as<float>(), partials);
    else if (p.prec == 4)
      hipLaunchKernelGGL((k_deposit<float, NM, RW, false>), grid, block, lds, s, sorted, tile_start, item_start,
                         o.item_tile, part_start, P, p.tiles.as<float>(), partials);
    else if (std_tiles)
      hipLaunchKernelGGL((k_deposit<double, NM, RW, true>), grid, block, lds, s, sorted, tile_start, item_start,
                         o.item_tile, part_start, P, p.tiles.as<double>(), partials);
    else
      hipLaunchKernelGGL((k_deposit<double, NM, RW, false>), grid, block, lds, s, sorted, tile_start, item_start,
                         o.item_tile, part_start, P, p.tiles.as<double>(), partials);
    GENPK_LAUNCH_CHECK();
  }
  {
    Stage st(p, "reduce_partials");
    if (p.prec == 4)
      hipLaunchKernelGGL((k_reduce_partials<float, NM>), dim3(1024), dim3(256), 0, s, nheavy, heavy, item_start,
                         part_start, partials, P, p.tiles.as<float>());
    else
      hipLaunchKernelGGL((k_reduce_partials<double, NM>), dim3(1024), dim3(256), 0, s, nheavy, heavy, item_start,
                         part_start, partials, P, p.tiles.as<double>());
    GENPK_LAUNCH_CHECK();
  }
  {
    // returns at once unless a position was not finite
    const FoldGeom fg = fold_geom(p);
    const unsigned grid = (unsigned)std::min<int64_t>(ceil_div(n, 256), 2048);
    if (p.prec == 4)
      hipLaunchKernelGGL(k_poison<float>, dim3(grid), dim3(256), 0, s, sorted, o.rw, n, kg, fg, P.nonfinite,
                         p.tiles.as<float>());
    else
      hipLaunchKernelGGL(k_poison<double>, dim3(grid), dim3(256), 0, s, sorted, o.rw, n, kg, fg, P.nonfinite,
                         p.tiles.as<double>());
    GENPK_LAUNCH_CHECK();
  }
  {
    Stage st(p, "fold_y");
    const FoldGeom fg = fold_geom(p);
    if (p.prec == 4)
      hipLaunchKernelGGL((k_fold_axis<float, 1>), dim3((unsigned)tg.K), dim3(256), 0, s, p.tiles.as<float>(), fg);
    else
      hipLaunchKernelGGL((k_fold_axis<double, 1>), dim3((unsigned)tg.K), dim3(256), 0, s, p.tiles.as<double>(), fg);
    GENPK_LAUNCH_CHECK();
  }
  {
    Stage st(p, "fold_x");
    const FoldGeom fg = fold_geom(p);
    if (p.prec == 4)
      hipLaunchKernelGGL((k_fold_axis<float, 0>), dim3((unsigned)tg.K), dim3(256), 0, s, p.tiles.as<float>(), fg);
    else
      hipLaunchKernelGGL((k_fold_axis<double, 0>), dim3((unsigned)tg.K), dim3(256), 0, s, p.tiles.as<double>(), fg);
    GENPK_LAUNCH_CHECK();
  }
}

__global__ void k_fill_zero(uint32_t *__restrict__ a, int64_t words) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < words; i += (int64_t)gridDim.x * blockDim.x)
    a[i] = 0u;
}

}  // namespace

// Zero a device buffer with a kernel (a memset node replays wrongly inside
// captured graphs on this platform; a kernel is safe there).
void fill_zero(void *ptr, size_t bytes, hipStream_t s) {
  GENPK_REQUIRE(bytes % 4 == 0, "fill_zero: size must be a multiple of 4");
  const int64_t words = (int64_t)(bytes / 4);
  if (words == 0) return;
  const unsigned grid = (unsigned)std::min<int64_t>(ceil_div(words, 256), 65536);
  hipLaunchKernelGGL(k_fill_zero, dim3(grid), dim3(256), 0, s, static_cast<uint32_t *>(ptr), words);
  GENPK_LAUNCH_CHECK();
}

FoldGeom fold_geom(const Plan &p) {
  const TileGeom &tg = p.tg;
  FoldGeom F;
  F.d = (int)p.d;
  F.TX = tg.TX; F.TY = tg.TY; F.TZ = tg.TZ;
  F.sx = tg.sx; F.sy = tg.sy; F.sz = tg.sz;
  F.ntx = tg.ntx; F.nty = tg.nty; F.ntz = tg.ntz;
  F.HY = tg.HY; F.HZ = tg.HZ;
  F.cm = tg.cm;
  F.cz = tg.cz;
  F.zoff = tg.K * tg.cm;
  return F;
}

void deposit(Plan &p, double boxsize, int64_t n, const float *pos, const float *masses, double mass, bool retain) {
  GENPK_REQUIRE(boxsize > 0 && std::isfinite(boxsize), "boxsize must be positive and finite");
  GENPK_REQUIRE(n >= 0, "negative particle count");
  GENPK_REQUIRE(n == 0 || pos != nullptr, "null positions");
  p.order.valid = false;  // the sort scratch is reused below
  if (n == 0) return;     // adds nothing; an untouched density stays zero
  p.tiles.ensure((size_t)p.tg.K * (p.tg.cm + p.tg.cz) * p.prec);
  const KeyGeom kg = key_geom(p, boxsize);
  // one call handles at most 2^31 particles (32-bit record offsets); the
  // tile order is retained only when one call covers all of them
  const int64_t cap = (int64_t)1 << 31;
  const bool keep = retain && n <= cap;
  int64_t done = 0;
  do {
    const int64_t m = std::min(n - done, cap);
    const float *mp = masses ? masses + done : nullptr;
    const bool live_before = p.tiles_live;
    const TileOrder o = sort_by_tile(p, kg, m, pos + 3 * done, mp, keep, !live_before);
    if (masses)
      deposit_impl<1, 4>(p, kg, o, mass);
    else
      deposit_impl<0, 3>(p, kg, o, mass);
    p.tiles_live = true;
    if (keep) {
      p.order.valid = true;
      p.order.pos = pos;
      p.order.masses = masses;
      p.order.n = n;
      p.order.boxsize = boxsize;
      p.order.mass = mass;
      p.order.every_tile = !live_before;
      p.order.to = o;
    }
    done += m;
  } while (done < n);
}

void redeposit(Plan &p) {
  GENPK_REQUIRE(p.order.valid, "no kept tile order: deposit with retain first");
  const RetainedOrder &r = p.order;
  p.tiles.ensure((size_t)p.tg.K * (p.tg.cm + p.tg.cz) * p.prec);
  if (!p.tiles_live && !r.every_tile) {
    // the kept items skip empty tiles: start from a zeroed store and accumulate
    fill_zero(p.tiles.ptr, (size_t)p.tg.K * (p.tg.cm + p.tg.cz) * p.prec, p.stream);
    p.tiles_live = true;
  }
  const KeyGeom kg = key_geom(p, r.boxsize);
  if (r.to.has_mass)
    deposit_impl<1, 4>(p, kg, r.to, r.mass);
  else
    deposit_impl<0, 3>(p, kg, r.to, r.mass);
  p.tiles_live = true;
}

void fold_tiles(Plan &p, void *out, int64_t out_rowlen, bool accumulate) {
  GENPK_REQUIRE(out_rowlen >= p.d, "output row length shorter than dims");
  if (!p.tiles_live) {
    // nothing deposited: the density is zero
    if (!accumulate) fill_zero(out, (size_t)p.d * p.d * out_rowlen * p.prec, p.stream);
    return;
  }
  const FoldGeom F = fold_geom(p);
  const int64_t lines = p.d * p.d;
  const unsigned gy = (unsigned)std::min<int64_t>(lines, 65535);
  const unsigned gz = (unsigned)ceil_div(lines, gy);
  dim3 grid((unsigned)ceil_div(p.d, FOLD_THREADS), gy, gz);
  Stage st(p, "fold");
  if (p.prec == 4)
    hipLaunchKernelGGL((k_fold<float, float>), grid, dim3(FOLD_THREADS), 0, p.stream, p.tiles.as<float>(), F,
                       out_rowlen, accumulate ? 1 : 0, static_cast<float *>(out));
  else
    hipLaunchKernelGGL((k_fold<double, double>), grid, dim3(FOLD_THREADS), 0, p.stream, p.tiles.as<double>(), F,
                       out_rowlen, accumulate ? 1 : 0, static_cast<double *>(out));
  GENPK_LAUNCH_CHECK();
}

void tiles_to_field(Plan &p) {
  if (!p.tiles_live) {
    if (p.fstate != FieldState::Real) {
      fill_zero(p.field.ptr, p.field_elems * p.prec, p.stream);
      p.fstate = FieldState::Real;
    }
    return;
  }
  fold_tiles(p, p.field.ptr, p.rowlen, p.fstate == FieldState::Real);
  p.fstate = FieldState::Real;
  p.tiles_live = false;
}

}  // namespace genpk
