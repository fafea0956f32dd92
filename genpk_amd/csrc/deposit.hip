// CIC deposit (the reference's fieldize(), fieldize.cpp:46-114) on gfx950.
//
// 1. Tile sort.  Particles are partitioned by the tile of their floor cell
//    with an MSD radix partition of one or two passes (upsweep histogram in
//    LDS -> offset scan -> downsweep that ranks records with wavefront
//    ballots, stages them in LDS by digit and writes contiguous runs).
//    Global atomics are avoided: on this chip they run at ~27 Gop/s.
// 2. Tile accumulation.  One workgroup per (tile, particle range) item
//    accumulates the 8 CIC weights of each particle into the halo'd tile in
//    LDS as 64-bit fixed point (ds_add_u64), then flushes the whole tile with
//    coalesced stores.  Tiles with more than DEPOSIT_ITEM_MAX particles are
//    split; their items write FP64 partial tiles that a reduction sums.
// 3. Fold.  The real-space padded grid [d][d][2(d/2+1)] is assembled by
//    adding each tile's halo layer onto the neighbouring tile's low face.
//
// Per-particle index and weight arithmetic follows fieldize.cpp:63-92:
// u = (double)pos * (d/box); i0 = floor(u) wrapped into [0,d), i1 = i0+1
// wrapped; f = u - floor(u); corner c has bit0 = x, bit1 = y, bit2 = z.
#include "plan.h"
#include "tiles.h"

#include <algorithm>
#include <cmath>
#include <type_traits>

namespace genpk {

namespace {

constexpr int SORT_THREADS = 256;
constexpr int SORT_WAVES = SORT_THREADS / 64;
constexpr int SORT_PER_LANE = 16;                                   // iterations per wave per sub-chunk
constexpr int SORT_SUB = SORT_THREADS * SORT_PER_LANE;              // 4096 records per sub-chunk
constexpr int SORT_CHUNK = 8 * SORT_SUB;                            // 32768 records per workgroup
constexpr int DEP_THREADS = 1024;
constexpr int FOLD_THREADS = 256;
constexpr int MAX_DIGIT_BITS = 11;

struct KeyGeom {
  double units;
  int d;
  int TX, TY, TZ;      // tile edge, powers of two
  int sx, sy, sz;      // log2 of the tile edges
  int ntx, nty, ntz;
};

// Floor cell of one coordinate, wrapped into [0, d), and the fractional part
// (fieldize.cpp:66-75).  The floor is converted to a 32-bit integer as the
// reference does; positions more than one box outside [0, box) take the
// general modulo.
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

__device__ __forceinline__ uint32_t tile_key(const float *r, const KeyGeom &g) {
  double fr;
  const int cx = cell_of(r[0], g.units, g.d, fr);
  const int cy = cell_of(r[1], g.units, g.d, fr);
  const int cz = cell_of(r[2], g.units, g.d, fr);
  return (uint32_t)((((cx >> g.sx) * g.nty + (cy >> g.sy)) * g.ntz) + (cz >> g.sz));
}

// 2^36 * w rounded to nearest, for w in [0, 1]: the integer lands in the low
// mantissa bits of w * 2^36 + 2^52.
__device__ __forceinline__ unsigned long long to_fixed(double w) {
  const double t = fma(w, FIXED_ONE, 4503599627370496.0);
  return (unsigned long long)__double_as_longlong(t) - 0x4330000000000000ull;
}

// Source of records: either the caller's arrays (positions [n][3] and an
// optional masses [n]) or a packed record buffer of RW floats per record.
template <int RW>
struct Src {
  const float *pos;     // raw positions (stride 3) or packed records (stride RW)
  const float *masses;  // raw masses when reading the caller's arrays
  bool raw;
  __device__ __forceinline__ void load(int64_t i, float *r) const {
    if (raw) {
      r[0] = pos[3 * i + 0];
      r[1] = pos[3 * i + 1];
      r[2] = pos[3 * i + 2];
      if (RW == 4) r[3] = masses[i];
    } else {
#pragma unroll
      for (int k = 0; k < RW; k++) r[k] = pos[RW * i + k];
    }
  }
};

// Chunk range of workgroup g.  Pass 1 chunks tile [0, n); pass 2 chunks
// tile each high-digit segment separately.
struct ChunkMap {
  int64_t n;
  int nseg;                    // 0: plain chunks of [0, n)
  const uint32_t *seg_start;   // [nseg+1] record offsets of the segments
  const uint32_t *chunk_start; // [nseg+1] first chunk of each segment
  __device__ __forceinline__ bool range(int64_t g, int64_t &b, int64_t &e) const {
    if (nseg == 0) {
      b = g * SORT_CHUNK;
      e = min(n, b + SORT_CHUNK);
      return b < e;
    }
    if (g >= (int64_t)chunk_start[nseg]) return false;
    int lo = 0, hi = nseg;  // find s with chunk_start[s] <= g < chunk_start[s+1]
    while (hi - lo > 1) {
      const int mid = (lo + hi) >> 1;
      if ((int64_t)chunk_start[mid] <= g) lo = mid; else hi = mid;
    }
    const int s = lo;
    b = (int64_t)seg_start[s] + (g - (int64_t)chunk_start[s]) * SORT_CHUNK;
    e = min((int64_t)seg_start[s + 1], b + SORT_CHUNK);
    return b < e;
  }
};

struct DigitSel {
  int shift;
  uint32_t mask;
  __device__ __forceinline__ uint32_t of(uint32_t key) const { return (key >> shift) & mask; }
};

__device__ __forceinline__ uint64_t lanemask_lt() {
  const int lane = threadIdx.x & 63;
  return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

// counts[dg * sd + g * sg] = number of records of chunk g with digit dg
template <int RW>
__global__ void __launch_bounds__(SORT_THREADS) k_upsweep(Src<RW> src, ChunkMap cm, KeyGeom kg, DigitSel ds,
                                                          int nb, uint32_t *__restrict__ counts, int64_t sd,
                                                          int64_t sg) {
  constexpr int UNR = 8;
  extern __shared__ uint32_t hist[];
  int64_t b, e;
  if (!cm.range(blockIdx.x, b, e)) return;
  for (int i = threadIdx.x; i < nb; i += SORT_THREADS) hist[i] = 0;
  __syncthreads();
  for (int64_t base = b; base < e; base += SORT_THREADS * UNR) {
    float r[UNR][4];
#pragma unroll
    for (int u = 0; u < UNR; u++) {
      const int64_t i = base + u * SORT_THREADS + threadIdx.x;
      if (i < e) src.load(i, r[u]);
    }
#pragma unroll
    for (int u = 0; u < UNR; u++) {
      const int64_t i = base + u * SORT_THREADS + threadIdx.x;
      const bool valid = i < e;
      const uint32_t dg = valid ? ds.of(tile_key(r[u], kg)) : 0u;
      // a wavefront whose records share one digit adds them with one atomic
      const uint64_t vmask = __ballot(valid);
      const uint32_t d0 = __shfl(dg, (int)__builtin_ctzll(vmask | (1ull << 63)), 64);
      if (__ballot(valid && dg == d0) == vmask) {
        if (valid && (vmask & lanemask_lt()) == 0) atomicAdd(&hist[d0], (uint32_t)__popcll(vmask));
      } else if (valid) {
        atomicAdd(&hist[dg], 1u);
      }
    }  }
  __syncthreads();
  for (int i = threadIdx.x; i < nb; i += SORT_THREADS)
    counts[(int64_t)i * sd + (int64_t)blockIdx.x * sg] = hist[i];
}

// Stable partition of each chunk by digit into dst at offsets[dg*sd + g*sg].
// Per sub-chunk of SORT_SUB records (each wave owns 1024 consecutive ones):
// wavefront ballots give every record its rank among equal digits, LDS
// prefix sums turn ranks into positions of a digit-sorted staging copy, and
// the staging copy is written out as contiguous runs.  The next sub-chunk's
// records are loaded while the current one is written out.
template <int RW>
__global__ void __launch_bounds__(SORT_THREADS) k_downsweep(Src<RW> src, float *__restrict__ dst, ChunkMap cm,
                                                            KeyGeom kg, DigitSel ds, int nb, int dbits,
                                                            const uint32_t *__restrict__ offsets, int64_t sd,
                                                            int64_t sg) {
  extern __shared__ uint32_t sm[];
  uint32_t *whist = sm;                         // [SORT_WAVES][nb]
  uint32_t *lstart = whist + SORT_WAVES * nb;   // [nb]
  uint32_t *gbase = lstart + nb;                // [nb]
  uint32_t *subcnt = gbase + nb;                // [nb]
  uint16_t *sdig = reinterpret_cast<uint16_t *>(subcnt + nb);          // [SORT_SUB]
  float *stage = reinterpret_cast<float *>(sdig + SORT_SUB);           // [SORT_SUB][RW]
  __shared__ uint32_t wtot[SORT_WAVES];

  int64_t cb, ce;
  if (!cm.range(blockIdx.x, cb, ce)) return;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  for (int i = threadIdx.x; i < nb; i += SORT_THREADS)
    gbase[i] = offsets[(int64_t)i * sd + (int64_t)blockIdx.x * sg];
  for (int i = threadIdx.x; i < SORT_WAVES * nb; i += SORT_THREADS) whist[i] = 0;

  float rec[SORT_PER_LANE][RW];
  const int wbase = wave * (SORT_PER_LANE * 64) + lane;
#pragma unroll
  for (int it = 0; it < SORT_PER_LANE; it++) {
    const int64_t i = cb + wbase + it * 64;
    if (i < ce) src.load(i, rec[it]);
  }
  __syncthreads();

  for (int64_t s0 = cb; s0 < ce; s0 += SORT_SUB) {
    const int nsub = (int)min((int64_t)SORT_SUB, ce - s0);
    uint32_t dig[SORT_PER_LANE];
    uint32_t rnk[SORT_PER_LANE];
#pragma unroll
    for (int it = 0; it < SORT_PER_LANE; it++) {
      const int loc = wbase + it * 64;
      const bool valid = loc < nsub;
      const uint32_t dg = valid ? ds.of(tile_key(rec[it], kg)) : 0u;
      uint64_t peers = __ballot(valid);
      // wavefronts whose records share one digit skip the per-bit match
      const uint32_t d0 = __shfl(dg, (int)__builtin_ctzll(peers | (1ull << 63)), 64);
      if (__ballot(valid && dg == d0) != peers) {
        for (int bit = 0; bit < dbits; bit++) {
          const bool set = (dg >> bit) & 1u;
          const uint64_t bb = __ballot(set);
          peers &= set ? bb : ~bb;
        }
      }
      dig[it] = dg;
      rnk[it] = 0;
      if (valid) {
        const uint32_t before = whist[wave * nb + dg];
        rnk[it] = before + (uint32_t)__popcll(peers & lanemask_lt());
        // every lane has read the running count before the group leader updates it
        __builtin_amdgcn_wave_barrier();
        if ((peers & lanemask_lt()) == 0) whist[wave * nb + dg] = before + (uint32_t)__popcll(peers);
      }
      __builtin_amdgcn_wave_barrier();
    }
    __syncthreads();
    // per-digit exclusive prefix over waves and sub-chunk totals
    for (int dgi = threadIdx.x; dgi < nb; dgi += SORT_THREADS) {
      uint32_t tot = 0;
#pragma unroll
      for (int w = 0; w < SORT_WAVES; w++) {
        const uint32_t t = whist[w * nb + dgi];
        whist[w * nb + dgi] = tot;
        tot += t;
      }
      subcnt[dgi] = tot;
    }
    __syncthreads();
    // exclusive scan of subcnt -> lstart (nb <= 2048)
    {
      const int per = (nb + SORT_THREADS - 1) / SORT_THREADS;
      const int first = threadIdx.x * per;
      uint32_t tsum = 0;
      for (int i = 0; i < per; i++)
        if (first + i < nb) tsum += subcnt[first + i];
      uint32_t incl = tsum;
#pragma unroll
      for (int off = 1; off < 64; off <<= 1) {
        const uint32_t t = __shfl_up(incl, off, 64);
        if (lane >= off) incl += t;
      }
      if (lane == 63) wtot[wave] = incl;
      __syncthreads();
      uint32_t run = incl - tsum;
      for (int w = 0; w < wave; w++) run += wtot[w];
      for (int i = 0; i < per; i++)
        if (first + i < nb) {
          lstart[first + i] = run;
          run += subcnt[first + i];
        }
    }
    __syncthreads();
#pragma unroll
    for (int it = 0; it < SORT_PER_LANE; it++) {
      const int loc = wbase + it * 64;
      if (loc < nsub) {
        const uint32_t dg = dig[it];
        const uint32_t p = lstart[dg] + whist[wave * nb + dg] + rnk[it];
#pragma unroll
        for (int k = 0; k < RW; k++) stage[p * RW + k] = rec[it][k];
        sdig[p] = (uint16_t)dg;
      }
    }
    __syncthreads();
    // prefetch the next sub-chunk while this one is written out
    {
      const int64_t n0 = s0 + SORT_SUB;
#pragma unroll
      for (int it = 0; it < SORT_PER_LANE; it++) {
        const int64_t i = n0 + wbase + it * 64;
        if (i < ce) src.load(i, rec[it]);
      }
    }
    for (int p = threadIdx.x; p < nsub; p += SORT_THREADS) {
      const uint32_t dg = sdig[p];
      const int64_t dest = (int64_t)gbase[dg] + (p - (int64_t)lstart[dg]);
#pragma unroll
      for (int k = 0; k < RW; k++) dst[dest * RW + k] = stage[p * RW + k];
    }
    __syncthreads();
    for (int i = threadIdx.x; i < nb; i += SORT_THREADS) gbase[i] += subcnt[i];
    for (int i = threadIdx.x; i < SORT_WAVES * nb; i += SORT_THREADS) whist[i] = 0;
    __syncthreads();
  }
}

// Pass-2 segment bookkeeping from the scanned pass-1 table.
__global__ void k_seginfo(const uint32_t *__restrict__ t1scan, int64_t g1, int nseg, int64_t n,
                          uint32_t *__restrict__ seg_start, uint32_t *__restrict__ nchunks) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s > nseg) return;
  if (s == nseg) {
    seg_start[s] = (uint32_t)n;
    nchunks[s] = 0;
    return;
  }
  const uint32_t st = t1scan[(int64_t)s * g1];
  const uint32_t en = (s + 1 < nseg) ? t1scan[(int64_t)(s + 1) * g1] : (uint32_t)n;
  seg_start[s] = st;
  nchunks[s] = (en - st + SORT_CHUNK - 1) / SORT_CHUNK;
}

// tot[hi*nlo+lo] = sum over the chunks of segment hi of cnt2[g*nlo+lo]
__global__ void k_seg_total(const uint32_t *__restrict__ cnt2, const uint32_t *__restrict__ chunk_start, int nhi,
                            int nlo, uint32_t *__restrict__ tot) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)nhi * nlo) return;
  const int hi = (int)(t / nlo), lo = (int)(t % nlo);
  uint32_t s = 0;
  for (uint32_t g = chunk_start[hi]; g < chunk_start[hi + 1]; g++) s += cnt2[(int64_t)g * nlo + lo];
  tot[t] = s;
}

// cnt2 -> per-chunk output offsets, starting at the tile start
__global__ void k_seg_offsets(uint32_t *__restrict__ cnt2, const uint32_t *__restrict__ chunk_start, int nhi, int nlo,
                              const uint32_t *__restrict__ tile_start) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)nhi * nlo) return;
  const int hi = (int)(t / nlo), lo = (int)(t % nlo);
  uint32_t run = tile_start[t];
  for (uint32_t g = chunk_start[hi]; g < chunk_start[hi + 1]; g++) {
    const uint32_t c = cnt2[(int64_t)g * nlo + lo];
    cnt2[(int64_t)g * nlo + lo] = run;
    run += c;
  }
}

// Single-pass tile starts: tile_start[k] = t1scan[k*g1] (chunk 0), tile_start[ks] = n
__global__ void k_tile_start_1pass(const uint32_t *__restrict__ t1scan, int64_t g1, int ks, int64_t n,
                                   uint32_t *__restrict__ tile_start) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k > ks) return;
  tile_start[k] = (k == ks) ? (uint32_t)n : t1scan[(int64_t)k * g1];
}

__global__ void k_set_last(uint32_t *__restrict__ a, int64_t at, uint32_t v) {
  if (threadIdx.x == 0 && blockIdx.x == 0) a[at] = v;
}

// Work items per tile.  store mode: every real tile gets at least one item
// (so untouched tiles are written as zeros); add mode: empty tiles are skipped.
__global__ void k_items(const uint32_t *__restrict__ tile_start, int64_t ks, int64_t kreal, int store,
                        uint32_t *__restrict__ mcnt, uint32_t *__restrict__ pcnt, uint32_t *__restrict__ nheavy) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k > ks) return;
  if (k == ks) {
    mcnt[k] = 0;
    pcnt[k] = 0;
    return;
  }
  const uint32_t c = tile_start[k + 1] - tile_start[k];
  uint32_t m = 0;
  if (k < kreal) m = c == 0 ? (store ? 1u : 0u) : (c + DEPOSIT_ITEM_MAX - 1) / DEPOSIT_ITEM_MAX;
  mcnt[k] = m;
  pcnt[k] = m > 1 ? m : 0;
  if (k == 0) *nheavy = 0;
}

__global__ void k_heavy_list(const uint32_t *__restrict__ pcnt, int64_t ks, uint32_t *__restrict__ nheavy,
                             uint32_t *__restrict__ heavy) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= ks) return;
  if (pcnt[k] > 1) heavy[atomicAdd(nheavy, 1u)] = (uint32_t)k;
}

struct DepParams {
  KeyGeom kg;
  int HX, HY, HZ;
  int64_t cells;       // accumulator cells HX*HY*HZ
  int64_t cm, cz, zoff;  // tile store layout
  int64_t ks;          // key space size (tile_start has ks+1 entries)
  double mass;         // constant mass (u64 path)
  int store;           // 1: write tiles, 0: add to tiles
};

template <typename Real, int RW>
__global__ void __launch_bounds__(DEP_THREADS) k_deposit(const float *__restrict__ rec,
                                                         const uint32_t *__restrict__ tile_start,
                                                         const uint32_t *__restrict__ item_start,
                                                         const uint32_t *__restrict__ part_start, DepParams P,
                                                         Real *__restrict__ tiles, double *__restrict__ partials) {
  using Acc = typename std::conditional<RW == 4, double, unsigned long long>::type;
  constexpr int UNR = 4;
  extern __shared__ unsigned long long lds_raw[];
  Acc *acc = reinterpret_cast<Acc *>(lds_raw);
  const uint32_t item = blockIdx.x;
  if (item >= item_start[P.ks]) return;
  // tile k with item_start[k] <= item < item_start[k+1]
  int64_t lo = 0, hi = P.ks;
  while (hi - lo > 1) {
    const int64_t mid = (lo + hi) >> 1;
    if (item_start[mid] <= item) lo = mid; else hi = mid;
  }
  const int64_t k = lo;
  const uint32_t m = item_start[k + 1] - item_start[k];
  const uint32_t j = item - item_start[k];
  const uint32_t t0 = tile_start[k], t1 = tile_start[k + 1];
  const uint32_t per = (t1 - t0 + m - 1) / m;
  const uint32_t pb = t0 + j * per;
  const uint32_t pe = min(t1, pb + per);

  for (int64_t c = threadIdx.x; c < P.cells; c += DEP_THREADS) acc[c] = 0;
  __syncthreads();

  const int ntz = P.kg.ntz, nty = P.kg.nty;
  const int tz = (int)(k % ntz);
  const int ty = (int)((k / ntz) % nty);
  const int tx = (int)(k / ((int64_t)ntz * nty));
  const int ox = tx << P.kg.sx, oy = ty << P.kg.sy, oz = tz << P.kg.sz;
  const int sy = P.HZ, sx = P.HY * P.HZ;
  float r[UNR][RW], nx[UNR][RW];
#pragma unroll
  for (int u = 0; u < UNR; u++) {
    const uint32_t i = pb + u * DEP_THREADS + threadIdx.x;
    if (i < pe) {
#pragma unroll
      for (int q = 0; q < RW; q++) nx[u][q] = rec[(int64_t)i * RW + q];
    }
  }
  for (uint32_t i0 = pb; i0 < pe; i0 += DEP_THREADS * UNR) {
#pragma unroll
    for (int u = 0; u < UNR; u++)
#pragma unroll
      for (int q = 0; q < RW; q++) r[u][q] = nx[u][q];
    // loads of the next batch are issued before this batch is accumulated
#pragma unroll
    for (int u = 0; u < UNR; u++) {
      const uint32_t i = i0 + (UNR + u) * DEP_THREADS + threadIdx.x;
      if (i < pe) {
#pragma unroll
        for (int q = 0; q < RW; q++) nx[u][q] = rec[(int64_t)i * RW + q];
      }
    }
#pragma unroll
    for (int u = 0; u < UNR; u++) {
      const uint32_t i = i0 + u * DEP_THREADS + threadIdx.x;
      if (i >= pe) continue;
      double fx, fy, fz;
      const int lx = cell_of(r[u][0], P.kg.units, P.kg.d, fx) - ox;
      const int ly = cell_of(r[u][1], P.kg.units, P.kg.d, fy) - oy;
      const int lz = cell_of(r[u][2], P.kg.units, P.kg.d, fz) - oz;
      if ((unsigned)lx >= (unsigned)P.kg.TX || (unsigned)ly >= (unsigned)P.kg.TY ||
          (unsigned)lz >= (unsigned)P.kg.TZ)
        continue;  // cannot happen for records of this tile; guards the LDS bounds
      const double gx = 1.0 - fx, gy = 1.0 - fy, gz = 1.0 - fz;
      const double wxy[4] = {gx * gy, fx * gy, gx * fy, fx * fy};
      const int base = lx * sx + ly * sy + lz;
      const int off[4] = {0, sx, sy, sx + sy};
#pragma unroll
      for (int c = 0; c < 4; c++) {
        const int at = base + off[c];
        if constexpr (RW == 4) {
          const double mw = (double)r[u][3] * wxy[c];
          atomicAdd(&acc[at], mw * gz);
          atomicAdd(&acc[at + 1], mw * fz);
        } else {
          atomicAdd(&acc[at], to_fixed(wxy[c] * gz));
          atomicAdd(&acc[at + 1], to_fixed(wxy[c] * fz));
        }
      }
    }
  }
  __syncthreads();
  const double scale = (RW == 4) ? 1.0 : P.mass * FIXED_INV;
  if (m == 1) {
    // rows of TZ values (aligned), then the z-halo plane
    Real *rows = tiles + k * P.cm;
    Real *zh = tiles + P.zoff + k * P.cz;
    const int sz = P.kg.sz, TZ = P.kg.TZ;
    if (P.store) {
      for (int64_t c = threadIdx.x; c < P.cm; c += DEP_THREADS) {
        const int64_t a = (c >> sz) * P.HZ + (c & (TZ - 1));
        rows[c] = (Real)((double)acc[a] * scale);
      }
      for (int64_t c = threadIdx.x; c < P.cz; c += DEP_THREADS) zh[c] = (Real)((double)acc[c * P.HZ + TZ] * scale);
    } else {
      for (int64_t c = threadIdx.x; c < P.cm; c += DEP_THREADS) {
        const int64_t a = (c >> sz) * P.HZ + (c & (TZ - 1));
        rows[c] = (Real)((double)rows[c] + (double)acc[a] * scale);
      }
      for (int64_t c = threadIdx.x; c < P.cz; c += DEP_THREADS)
        zh[c] = (Real)((double)zh[c] + (double)acc[c * P.HZ + TZ] * scale);
    }
  } else {
    double *out = partials + (int64_t)(part_start[k] + j) * P.cells;
    for (int64_t c = threadIdx.x; c < P.cells; c += DEP_THREADS) out[c] = (double)acc[c] * scale;
  }
}

// Sum the partial tiles of split tiles into the tile store.
template <typename Real>
__global__ void __launch_bounds__(256) k_reduce_partials(const uint32_t *__restrict__ nheavy,
                                                         const uint32_t *__restrict__ heavy,
                                                         const uint32_t *__restrict__ item_start,
                                                         const uint32_t *__restrict__ part_start,
                                                         const double *__restrict__ partials, DepParams P,
                                                         Real *__restrict__ tiles) {
  const int64_t cells = P.cells;
  const int64_t nblk = ceil_div(cells, 256);
  const int64_t units = (int64_t)(*nheavy) * nblk;
  for (int64_t u = blockIdx.x; u < units; u += gridDim.x) {
    const uint32_t k = heavy[u / nblk];
    const int64_t c = (u % nblk) * 256 + threadIdx.x;
    if (c >= cells) continue;
    const uint32_t m = item_start[k + 1] - item_start[k];
    const double *pp = partials + (int64_t)part_start[k] * cells + c;
    double s = 0;
    for (uint32_t j = 0; j < m; j++) s += pp[(int64_t)j * cells];
    const int64_t row = c / P.HZ;
    const int lz = (int)(c - row * P.HZ);
    Real *out = lz < P.kg.TZ ? tiles + (int64_t)k * P.cm + row * P.kg.TZ + lz : tiles + P.zoff + (int64_t)k * P.cz + row;
    *out = P.store ? (Real)s : (Real)((double)*out + s);
  }
}

// Halo fold: every cell on a tile's low face (local x, y or z = 0) adds the
// halo cells of the previous tiles that belong to it and zeroes them, so the
// store holds each grid cell's full value at its own (tile, local) place.
// Each halo cell maps to exactly one face cell, so the kernel is race free
// and later accumulating deposits start from zero halos.
template <typename Real>
__global__ void __launch_bounds__(256) k_halo_fold(Real *__restrict__ tiles, FoldGeom P) {
  const int64_t t = blockIdx.x;
  const int tz = (int)(t % P.ntz);
  const int ty = (int)((t / P.ntz) % P.nty);
  const int tx = (int)(t / ((int64_t)P.ntz * P.nty));
  const int wx = min(P.TX, P.d - tx * P.TX), wy = min(P.TY, P.d - ty * P.TY), wz = min(P.TZ, P.d - tz * P.TZ);
  const int txp = tx == 0 ? P.ntx - 1 : tx - 1, typ = ty == 0 ? P.nty - 1 : ty - 1, tzp = tz == 0 ? P.ntz - 1 : tz - 1;
  const int hx = min(P.TX, P.d - txp * P.TX), hy = min(P.TY, P.d - typ * P.TY), hz = min(P.TZ, P.d - tzp * P.TZ);
  // face cells: plane lx = 0 (wy*wz), plane ly = 0 with lx >= 1 ((wx-1)*wz), line-free rest lz = 0 ((wx-1)(wy-1))
  const int nA = wy * wz, nB = (wx - 1) * wz, nC = (wx - 1) * (wy - 1);
  for (int f = threadIdx.x; f < nA + nB + nC; f += 256) {
    int lx, ly, lz;
    if (f < nA) { lx = 0; ly = f / wz; lz = f - ly * wz; }
    else if (f < nA + nB) { const int g = f - nA; lx = 1 + g / wz; ly = 0; lz = g - (lx - 1) * wz; }
    else { const int g = f - nA - nB; lx = 1 + g / (wy - 1); ly = 1 + g - (lx - 1) * (wy - 1); lz = 0; }
    const int nxs = lx == 0 ? 2 : 1, nys = ly == 0 ? 2 : 1, nzs = lz == 0 ? 2 : 1;
    double v = 0;
    for (int a = 0; a < nxs; a++)
      for (int b = 0; b < nys; b++)
        for (int c = 0; c < nzs; c++) {
          if (a == 0 && b == 0 && c == 0) continue;
          const int64_t st = ((int64_t)(a ? txp : tx) * P.nty + (b ? typ : ty)) * P.ntz + (c ? tzp : tz);
          const int64_t at = tile_at(P, st, a ? hx : lx, b ? hy : ly, c ? hz : lz);
          v += (double)tiles[at];
          tiles[at] = (Real)0;
        }
    const int64_t me = tile_at(P, t, lx, ly, lz);
    tiles[me] = (Real)((double)tiles[me] + v);
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
  const double v = (double)tiles[cell_at(P, x, y, z)];
  Out *o = out + line * out_rowlen + z;
  *o = accumulate ? (Out)((double)*o + v) : (Out)v;
}

KeyGeom key_geom(const Plan &p, double boxsize) {
  KeyGeom g;
  g.units = (double)p.d / boxsize;
  g.d = (int)p.d;
  g.TX = p.tg.TX; g.TY = p.tg.TY; g.TZ = p.tg.TZ;
  g.sx = p.tg.sx; g.sy = p.tg.sy; g.sz = p.tg.sz;
  g.ntx = p.tg.ntx; g.nty = p.tg.nty; g.ntz = p.tg.ntz;
  return g;
}

size_t downsweep_lds(int nb, int rw) {
  return (size_t)(SORT_WAVES * nb + 3 * nb) * 4 + (size_t)SORT_SUB * 2 + (size_t)SORT_SUB * rw * 4;
}

template <int RW>
void sort_and_deposit(Plan &p, double boxsize, int64_t n, const float *pos, const float *masses, double mass) {
  hipStream_t s = p.stream;
  const TileGeom &tg = p.tg;
  const KeyGeom kg = key_geom(p, boxsize);

  float *A = p.recA.as<float>();
  float *B = p.recB.as<float>();
  const int64_t g1 = ceil_div(n, SORT_CHUNK);
  Src<RW> raw{pos, masses, true};
  const float *sorted = nullptr;
  uint32_t *tile_start = nullptr;
  int64_t ks = 0;

  if (tg.passes == 1) {
    const int nb = 1 << tg.hi_bits;
    ks = nb;
    uint32_t *t1 = static_cast<uint32_t *>(p.table1.ensure((size_t)nb * g1 * 4));
    DigitSel dsel{0, (uint32_t)(nb - 1)};
    ChunkMap cm{n, 0, nullptr, nullptr};
    {
      Stage st(p, "sort_upsweep1");
      hipLaunchKernelGGL(k_upsweep<RW>, dim3((unsigned)g1), dim3(SORT_THREADS), nb * 4, s, raw, cm, kg, dsel, nb, t1,
                         g1, (int64_t)1);
      GENPK_LAUNCH_CHECK();
    }
    exclusive_scan_u32(t1, t1, (int64_t)nb * g1, p.scan_tmp, s);
    {
      Stage st(p, "sort_downsweep1");
      hipLaunchKernelGGL(k_downsweep<RW>, dim3((unsigned)g1), dim3(SORT_THREADS), downsweep_lds(nb, RW), s, raw, A,
                         cm, kg, dsel, nb, tg.hi_bits, t1, g1, (int64_t)1);
      GENPK_LAUNCH_CHECK();
    }
    tile_start = static_cast<uint32_t *>(p.tile_start.ensure((size_t)(ks + 1) * 4));
    hipLaunchKernelGGL(k_tile_start_1pass, dim3((unsigned)ceil_div(ks + 1, 256)), dim3(256), 0, s, t1, g1, (int)ks,
                       n, tile_start);
    GENPK_LAUNCH_CHECK();
    sorted = A;
  } else {
    const int nhi = 1 << tg.hi_bits, nlo = 1 << tg.lo_bits;
    ks = (int64_t)nhi * nlo;
    // pass 1: high digit over [0, n)
    uint32_t *t1 = static_cast<uint32_t *>(p.table1.ensure((size_t)nhi * g1 * 4));
    DigitSel d1{tg.lo_bits, (uint32_t)(nhi - 1)};
    ChunkMap cm1{n, 0, nullptr, nullptr};
    {
      Stage st(p, "sort_upsweep1");
      hipLaunchKernelGGL(k_upsweep<RW>, dim3((unsigned)g1), dim3(SORT_THREADS), nhi * 4, s, raw, cm1, kg, d1, nhi,
                         t1, g1, (int64_t)1);
      GENPK_LAUNCH_CHECK();
    }
    exclusive_scan_u32(t1, t1, (int64_t)nhi * g1, p.scan_tmp, s);
    {
      Stage st(p, "sort_downsweep1");
      hipLaunchKernelGGL(k_downsweep<RW>, dim3((unsigned)g1), dim3(SORT_THREADS), downsweep_lds(nhi, RW), s, raw, A,
                         cm1, kg, d1, nhi, tg.hi_bits, t1, g1, (int64_t)1);
      GENPK_LAUNCH_CHECK();
    }
    // segments of pass 2
    uint32_t *seg = static_cast<uint32_t *>(p.seginfo.ensure((size_t)(nhi + 1) * 3 * 4));
    uint32_t *seg_start = seg, *nchunks = seg + (nhi + 1), *chunk_start = seg + 2 * (nhi + 1);
    hipLaunchKernelGGL(k_seginfo, dim3((unsigned)ceil_div(nhi + 1, 256)), dim3(256), 0, s, t1, g1, nhi, n, seg_start,
                       nchunks);
    GENPK_LAUNCH_CHECK();
    exclusive_scan_u32(nchunks, chunk_start, nhi + 1, p.scan_tmp, s);
    const int64_t g2 = g1 + nhi;  // upper bound of pass-2 chunks
    uint32_t *t2 = static_cast<uint32_t *>(p.table2.ensure((size_t)g2 * nlo * 4 + (size_t)(ks + 1) * 4 * 2));
    uint32_t *tot = t2 + g2 * nlo;
    DigitSel d2{0, (uint32_t)(nlo - 1)};
    ChunkMap cm2{n, nhi, seg_start, chunk_start};
    Src<RW> pa{A, nullptr, false};
    {
      Stage st(p, "sort_upsweep2");
      hipLaunchKernelGGL(k_upsweep<RW>, dim3((unsigned)g2), dim3(SORT_THREADS), nlo * 4, s, pa, cm2, kg, d2, nlo, t2,
                         (int64_t)1, (int64_t)nlo);
      GENPK_LAUNCH_CHECK();
    }
    hipLaunchKernelGGL(k_seg_total, dim3((unsigned)ceil_div(ks, 256)), dim3(256), 0, s, t2, chunk_start, nhi, nlo,
                       tot);
    GENPK_LAUNCH_CHECK();
    tile_start = static_cast<uint32_t *>(p.tile_start.ensure((size_t)(ks + 1) * 4));
    hipLaunchKernelGGL(k_set_last, dim3(1), dim3(1), 0, s, tot, ks, 0u);
    GENPK_LAUNCH_CHECK();
    exclusive_scan_u32(tot, tile_start, ks + 1, p.scan_tmp, s);
    hipLaunchKernelGGL(k_seg_offsets, dim3((unsigned)ceil_div(ks, 256)), dim3(256), 0, s, t2, chunk_start, nhi, nlo,
                       tile_start);
    GENPK_LAUNCH_CHECK();
    {
      Stage st(p, "sort_downsweep2");
      hipLaunchKernelGGL(k_downsweep<RW>, dim3((unsigned)g2), dim3(SORT_THREADS), downsweep_lds(nlo, RW), s, pa, B,
                         cm2, kg, d2, nlo, tg.lo_bits, t2, (int64_t)1, (int64_t)nlo);
      GENPK_LAUNCH_CHECK();
    }
    sorted = B;
  }

  // work items
  const int store = p.tiles_live ? 0 : 1;
  uint32_t *ib = static_cast<uint32_t *>(p.item_start.ensure((size_t)(ks + 1) * 4 * 4 + 64));
  uint32_t *mcnt = ib, *item_start = ib + (ks + 1), *pcnt = ib + 2 * (ks + 1), *part_start = ib + 3 * (ks + 1);
  uint32_t *hv = static_cast<uint32_t *>(p.heavy.ensure((size_t)(ks + 2) * 4));
  uint32_t *nheavy = hv, *heavy = hv + 1;
  hipLaunchKernelGGL(k_items, dim3((unsigned)ceil_div(ks + 1, 256)), dim3(256), 0, s, tile_start, ks, tg.K, store,
                     mcnt, pcnt, nheavy);
  GENPK_LAUNCH_CHECK();
  exclusive_scan_u32(mcnt, item_start, ks + 1, p.scan_tmp, s);
  exclusive_scan_u32(pcnt, part_start, ks + 1, p.scan_tmp, s);
  hipLaunchKernelGGL(k_heavy_list, dim3((unsigned)ceil_div(ks, 256)), dim3(256), 0, s, pcnt, ks, nheavy, heavy);
  GENPK_LAUNCH_CHECK();

  const int64_t max_items = tg.K + ceil_div(n, DEPOSIT_ITEM_MAX);
  const int64_t max_parts = 2 * ceil_div(n, DEPOSIT_ITEM_MAX);
  double *partials = static_cast<double *>(p.partials.ensure((size_t)max_parts * tg.cells * 8 + 64));

  DepParams P;
  P.kg = kg;
  P.HX = tg.HX; P.HY = tg.HY; P.HZ = tg.HZ;
  P.cells = tg.cells;
  P.cm = tg.cm;
  P.cz = tg.cz;
  P.zoff = tg.K * tg.cm;
  P.ks = ks;
  P.mass = mass;
  P.store = store;
  const size_t lds = (size_t)tg.cells * 8;
  Stage st(p, "deposit_tiles");
  if (p.prec == 4) {
    hipLaunchKernelGGL((k_deposit<float, RW>), dim3((unsigned)max_items), dim3(DEP_THREADS), lds, s, sorted,
                       tile_start, item_start, part_start, P, p.tiles.as<float>(), partials);
    GENPK_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_reduce_partials<float>, dim3(1024), dim3(256), 0, s, nheavy, heavy, item_start, part_start,
                       partials, P, p.tiles.as<float>());
  } else {
    hipLaunchKernelGGL((k_deposit<double, RW>), dim3((unsigned)max_items), dim3(DEP_THREADS), lds, s, sorted,
                       tile_start, item_start, part_start, P, p.tiles.as<double>(), partials);
    GENPK_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_reduce_partials<double>, dim3(1024), dim3(256), 0, s, nheavy, heavy, item_start, part_start,
                       partials, P, p.tiles.as<double>());
  }
  GENPK_LAUNCH_CHECK();
  {
    Stage hst(p, "halo_fold");
    const FoldGeom fg = fold_geom(p);
    if (p.prec == 4)
      hipLaunchKernelGGL(k_halo_fold<float>, dim3((unsigned)tg.K), dim3(256), 0, s, p.tiles.as<float>(), fg);
    else
      hipLaunchKernelGGL(k_halo_fold<double>, dim3((unsigned)tg.K), dim3(256), 0, s, p.tiles.as<double>(), fg);
    GENPK_LAUNCH_CHECK();
  }
}

}  // namespace

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

void deposit(Plan &p, double boxsize, int64_t n, const float *pos, const float *masses, double mass) {
  GENPK_REQUIRE(boxsize > 0 && std::isfinite(boxsize), "boxsize must be positive and finite");
  GENPK_REQUIRE(n >= 0, "negative particle count");
  GENPK_REQUIRE(n == 0 || pos != nullptr, "null positions");
  if (n == 0) return;  // adds nothing; an untouched density stays zero
  p.tiles.ensure((size_t)p.tg.K * (p.tg.cm + p.tg.cz) * p.prec);
  // one call handles at most 2^31 particles (32-bit record offsets)
  const int64_t cap = (int64_t)1 << 31;
  int64_t done = 0;
  do {
    const int64_t m = std::min(n - done, cap);
    const int rw = masses ? 4 : 3;
    const size_t rec_bytes = (size_t)std::max<int64_t>(m, 1) * rw * 4;
    p.recA.ensure(rec_bytes);
    if (p.tg.passes == 2) p.recB.ensure(rec_bytes);
    if (masses)
      sort_and_deposit<4>(p, boxsize, m, pos + 3 * done, masses + done, mass);
    else
      sort_and_deposit<3>(p, boxsize, m, pos + 3 * done, nullptr, mass);
    p.tiles_live = true;
    done += m;
  } while (done < n);
}

void fold_tiles(Plan &p, void *out, int64_t out_rowlen, bool accumulate) {
  GENPK_REQUIRE(out_rowlen >= p.d, "output row length shorter than dims");
  if (!p.tiles_live) {
    // nothing deposited: the density is zero
    if (!accumulate) GENPK_HIP(hipMemsetAsync(out, 0, (size_t)p.d * p.d * out_rowlen * p.prec, p.stream));
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
      GENPK_HIP(hipMemsetAsync(p.field.ptr, 0, p.field_elems * p.prec, p.stream));
      p.fstate = FieldState::Real;
    }
    return;
  }
  fold_tiles(p, p.field.ptr, p.rowlen, p.fstate == FieldState::Real);
  p.fstate = FieldState::Real;
  p.tiles_live = false;
}

}  // namespace genpk
