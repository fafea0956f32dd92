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
// the z-halo plane (local z = TZ) of every tile, HX*HY values per tile.
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

// Store offset of grid cell (x, y, z) in the tile that contains it.  After
// the halo fold this is the whole value of the cell.
__device__ __forceinline__ int64_t cell_at(const FoldGeom &P, int x, int y, int z) {
  const int64_t t = (((int64_t)(x >> P.sx) * P.nty + (y >> P.sy)) * P.ntz) + (z >> P.sz);
  return tile_at(P, t, x & (P.TX - 1), y & (P.TY - 1), z & (P.TZ - 1));
}

// Value of grid cell (x, y, z) assembled from a tile store whose halo cells
// have not been folded.
template <typename Real>
__device__ __forceinline__ double fold_value(const Real *__restrict__ tiles, const FoldGeom &P, int x, int y, int z) {
  int stx[2], slx[2], sty[2], sly[2], stz[2], slz[2];
  int nxs = 1, nys = 1, nzs = 1;
  stx[0] = x >> P.sx; slx[0] = x - (stx[0] << P.sx);
  sty[0] = y >> P.sy; sly[0] = y - (sty[0] << P.sy);
  stz[0] = z >> P.sz; slz[0] = z - (stz[0] << P.sz);
  if (slx[0] == 0) {
    const int t = stx[0] == 0 ? P.ntx - 1 : stx[0] - 1;
    stx[1] = t; slx[1] = min(P.TX, P.d - t * P.TX); nxs = 2;
  }
  if (sly[0] == 0) {
    const int t = sty[0] == 0 ? P.nty - 1 : sty[0] - 1;
    sty[1] = t; sly[1] = min(P.TY, P.d - t * P.TY); nys = 2;
  }
  if (slz[0] == 0) {
    const int t = stz[0] == 0 ? P.ntz - 1 : stz[0] - 1;
    stz[1] = t; slz[1] = min(P.TZ, P.d - t * P.TZ); nzs = 2;
  }
  double v = 0;
#pragma unroll
  for (int a = 0; a < 2; a++) {
    if (a >= nxs) break;
#pragma unroll
    for (int b = 0; b < 2; b++) {
      if (b >= nys) break;
      const int64_t t0 = ((int64_t)stx[a] * P.nty + sty[b]) * P.ntz;
#pragma unroll
      for (int c = 0; c < 2; c++) {
        if (c >= nzs) break;
        v += (double)tiles[tile_at(P, t0 + stz[c], slx[a], sly[b], slz[c])];
      }
    }
  }
  return v;
}

}  // namespace genpk
