// Halo'd tile store of the deposit and its fold into the real-space grid.
//
// Tile t = (tx*nty + ty)*ntz + tz covers cells [tx*TX, tx*TX+wx) x ... with
// wx = min(TX, d - tx*TX); its accumulator has (TX+1)(TY+1)(TZ+1) cells
// (local index 0..TX along x, and so on).  Local index wx (the halo layer)
// holds the CIC weight that belongs to the first cell of the next tile along
// that axis (periodically), so the value of grid cell (x, y, z) is the sum
// over the one or two (tile, local) sources per axis.
//
// Store layout: rows of TZ values for local z < TZ, HX*HY rows per tile,
// tile after tile (each row is 32 floats = 128 bytes, aligned), followed by
// the z-halo plane (local z = TZ) of every tile, HX*HY values per tile.  The
// y and x halos are folded in place after each deposit (fold_y, fold_x); the
// z halos by the readers (the z pass, fold_value).
#pragma once
#include "common.h"

namespace genpk {

struct FoldGeom {
  int d;
  int TX, TY, TZ, sx, sy, sz;
  int ntx, nty, ntz;
  int HY, HZ;
  int64_t cm;     // row-store values per tile: HX*HY*TZ
  int64_t cz;     // z-halo values per tile: HX*HY
  int64_t zoff;   // start of the z-halo planes: K*cm
};

// Store offset of (tile t, local lx, ly, lz), lz <= TZ.
__device__ __forceinline__ int64_t tile_at(const FoldGeom &P, int64_t t, int lx, int ly, int lz) {
  const int64_t r = (int64_t)lx * P.HY + ly;
  return lz < P.TZ ? t * P.cm + r * P.TZ + lz : P.zoff + t * P.cz + r;
}

// Store offset of grid cell (x, y, z) in the tile that contains it (its own
// accumulator cell; halo contributions of the previous tiles are separate).
__device__ __forceinline__ int64_t cell_at(const FoldGeom &P, int x, int y, int z) {
  const int64_t t = (((int64_t)(x >> P.sx) * P.nty + (y >> P.sy)) * P.ntz) + (z >> P.sz);
  return tile_at(P, t, x & (P.TX - 1), y & (P.TY - 1), z & (P.TZ - 1));
}

// After the in-place x/y halo folds, a grid cell's value is its own cell plus,
// for the first cell of a tile along z (local z = 0), the z-halo cell of the
// previous tile along z (at local z = that tile's width: the z-halo plane for
// full tiles, the row store for a partial last tile).  zhalo_at gives the
// store offset of that z-halo cell.
__device__ __forceinline__ int64_t zhalo_at(const FoldGeom &P, int x, int y, int z) {
  const int tx = x >> P.sx, ty = y >> P.sy, tz = z >> P.sz;
  const int tzp = tz == 0 ? P.ntz - 1 : tz - 1;
  const int hz = min(P.TZ, P.d - tzp * P.TZ);
  return tile_at(P, ((int64_t)tx * P.nty + ty) * P.ntz + tzp, x & (P.TX - 1), y & (P.TY - 1), hz);
}

// Value of grid cell (x, y, z) from an x/y-folded tile store, summed as the
// transform's z pass does.
template <typename Real>
__device__ __forceinline__ double fold_value(const Real *__restrict__ tiles, const FoldGeom &P, int x, int y, int z) {
  double v = (double)tiles[cell_at(P, x, y, z)];
  if ((z & (P.TZ - 1)) == 0) v += (double)tiles[zhalo_at(P, x, y, z)];
  return v;
}

}  // namespace genpk
