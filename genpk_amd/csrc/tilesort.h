// Partition of particles by deposit tile (sort.hip) and the per-particle
// cell arithmetic shared by the deposit and the adjoint gather.
#pragma once
#include "plan.h"

namespace genpk {

struct KeyGeom {
  double units;
  int d;
  int TX, TY, TZ;      // tile edge, powers of two
  int sx, sy, sz;      // log2 of the tile edges
  int ntx, nty, ntz;
};

KeyGeom key_geom(const Plan &p, double boxsize);

// Floor cell of one coordinate, wrapped into [0, d), and the fractional part
// (fieldize.cpp:66-75).  The floor is converted to a 32-bit integer as the
// reference does; positions more than one box outside [0, box) take the
// general modulo.  A floor outside the int range converts with the GPU's
// saturation (NaN to 0), so every record still gets a valid tile; the
// deposit skips a record whose fractional part is not finite, and k_poison
// finds its cells with cell_of_x86.
__device__ __forceinline__ int cell_of(float p, double units, int d, double &frac) {
  const double u = (double)p * units;
  const double f = floor(u);
  frac = u - f;
  int c = (int)f;
  if ((unsigned)c >= (unsigned)d) {
    c %= d;
    if (c < 0) c += d;
  }
  return c;
}

// cell_of with the conversion x86-64 makes (cvttsd2si): a floor outside the
// int range, NaN or an infinity converts to INT_MIN (the "integer
// indefinite"), so a non-finite position maps to the cells the reference
// gives it.
__device__ __forceinline__ int cell_of_x86(float p, double units, int d, double &frac) {
  const double u = (double)p * units;
  const double f = floor(u);
  frac = u - f;
  int c = (f >= -2147483648.0 && f < 2147483648.0) ? (int)f : (int)0x80000000u;
  if ((unsigned)c >= (unsigned)d) {
    c %= d;
    if (c < 0) c += d;
  }
  return c;
}

// Tile key of a position: cell_of's cells along the three axes, with one
// branch for the rare wrap of any of them into [0, d).
__device__ __forceinline__ uint32_t tile_key(const float *r, const KeyGeom &g) {
  int c[3];
#pragma unroll
  for (int a = 0; a < 3; a++) c[a] = (int)floor((double)r[a] * g.units);
  const unsigned ud = (unsigned)g.d;
  if (((unsigned)c[0] >= ud) | ((unsigned)c[1] >= ud) | ((unsigned)c[2] >= ud)) {
#pragma unroll
    for (int a = 0; a < 3; a++)
      if ((unsigned)c[a] >= ud) {
        c[a] %= g.d;
        if (c[a] < 0) c[a] += g.d;
      }
  }
  return (uint32_t)((((c[0] >> g.sx) * g.nty + (c[1] >> g.sy)) * g.ntz) + (c[2] >> g.sz));
}

// Work item -> tile k and its particle range [pb, pe): tile k has m items
// (item_start[k+1] - item_start[k]) that split its records evenly; j is the
// item's number within the tile.  item_tile (ntab entries, at least the
// item count) holds each item's tile; it is read at min(item, ntab - 1)
// before the item count is known, and used only below the count.  Returns
// false past the last item.
struct ItemRange {
  int64_t k;
  uint32_t m, j, pb, pe;
};
__device__ __forceinline__ bool item_range(const uint32_t *__restrict__ item_tile,
                                           const uint32_t *__restrict__ item_start,
                                           const uint32_t *__restrict__ tile_start, int64_t ks, int64_t ntab,
                                           uint32_t item, ItemRange &r) {
  const uint32_t total = item_start[ks];
  const uint32_t k = item_tile[min((int64_t)item, ntab - 1)];
  if (item >= total) return false;
  r.k = k;
  const uint32_t i0 = item_start[k], i1 = item_start[k + 1];
  const uint32_t t0 = tile_start[k], t1 = tile_start[k + 1];
  r.m = i1 - i0;
  r.j = item - i0;
  const uint32_t per = r.m == 1 ? t1 - t0 : (t1 - t0 + r.m - 1) / r.m;  // most tiles are one item
  r.pb = min(t1, t0 + r.j * per);
  r.pe = min(t1, r.pb + per);
  return true;
}

// Records per work item at most for n particles over `tiles` tiles:
// DEPOSIT_ITEM_MAX, halved (down to DEPOSIT_ITEM_MIN) while the items would
// number fewer than four per CU or one full item would hold more than half
// of a CU's share of the particles, so that a few heavy tiles still spread
// over the CUs (their items' partial tiles are summed afterwards).
uint32_t deposit_item_max(int64_t n, int64_t tiles);

// Sort the n particles of one call (n < 2^31) by tile into plan scratch and
// build the work items (at most deposit_item_max records each).  Records are
// x, y, z, then the mass when masses != nullptr.  with_perm: also record
// where every particle went (perm1: input -> pass-1 position; perm2, two
// passes only: pass-1 -> final position), for taking results back to the
// input order without scattered writes.  every_tile: every tile gets at
// least one item (the deposit then writes untouched tiles as zeros);
// otherwise empty tiles get none.
TileOrder sort_by_tile(Plan &p, const KeyGeom &kg, int64_t n, const float *pos, const float *masses, bool with_perm,
                       bool every_tile);

}  // namespace genpk
