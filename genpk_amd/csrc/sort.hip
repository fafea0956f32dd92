// Partition of particles by deposit tile on gfx950: an MSD radix partition
// of one or two passes over the tile key (upsweep histogram in LDS ->
// offset scan -> downsweep that ranks records with wavefront ballots,
// stages them in LDS by digit and writes contiguous runs), then the work
// items of the tile kernels.  Global atomics are avoided: on this chip they
// run at ~27 Gop/s.  The partition is stable, so the records of one tile
// keep their input order.
#include "tilesort.h"

namespace genpk {

namespace {

// 8 waves per workgroup.  The downsweep's records per lane per sub-chunk (PL)
// is the largest of 8, 4, 2 whose LDS staging leaves room for two workgroups
// (16 waves) per CU: at 1024^3 this measured faster than 4 waves x 16 records
// (8.3 -> 7.2 ms per downsweep) or 16 waves x 4 records, and one workgroup
// per CU (16-byte records staged 4096 at a time) doubles the time.
constexpr int SORT_THREADS = 512;
constexpr int SORT_WAVES = SORT_THREADS / 64;
constexpr int SORT_CHUNK = 65536;                                   // records per workgroup

// Source of records: either the caller's arrays (positions [n][3], optional
// masses [n]) or a packed record buffer of RW floats per record.
template <int NM>
struct Src {
  static constexpr int RW = 3 + NM;
  const float *pos;     // raw positions (stride 3) or packed records (stride RW)
  const float *masses;  // raw masses when reading the caller's arrays
  bool raw;
  __device__ __forceinline__ void load_pos(int64_t i, float *r) const {
    const int64_t st = raw ? 3 : RW;
    r[0] = pos[st * i + 0];
    r[1] = pos[st * i + 1];
    r[2] = pos[st * i + 2];
  }
  __device__ __forceinline__ void load(int64_t i, float *r) const {
    if (raw) {
      r[0] = pos[3 * i + 0];
      r[1] = pos[3 * i + 1];
      r[2] = pos[3 * i + 2];
      if (NM) r[3] = masses[i];
    } else {
#pragma unroll
      for (int k = 0; k < RW; k++) r[k] = pos[RW * i + k];
    }
  }
};

// Chunk range of workgroup g and the place of its digit counts in the
// count table: counts of digit dg at tb + dg * ts.  Pass 1 chunks tile
// [0, n) and its table is digit-major over all chunks (tb = g, ts = number
// of chunks).  Pass 2 chunks tile each high-digit segment separately; its
// table is segment-major, then digit, then chunk (tb = first chunk of the
// segment * nlo + chunk within the segment, ts = chunks of the segment), so
// one exclusive scan of the table gives every chunk's output offsets and
// every tile's start.
struct ChunkMap {
  int64_t n;
  int nseg;                    // 0: plain chunks of [0, n)
  const uint32_t *seg_start;   // [nseg+1] record offsets of the segments
  const uint32_t *chunk_start; // [nseg+1] first chunk of each segment
  int64_t nchunks;             // pass 1: number of chunks
  int nlo;                     // pass 2: digits
  __device__ __forceinline__ bool range(int64_t g, int64_t &b, int64_t &e, int64_t &tb, int64_t &ts) const {
    if (nseg == 0) {
      b = g * SORT_CHUNK;
      e = min(n, b + SORT_CHUNK);
      tb = g;
      ts = nchunks;
      return b < e;
    }
    if (g >= (int64_t)chunk_start[nseg]) return false;
    int lo = 0, hi = nseg;  // find s with chunk_start[s] <= g < chunk_start[s+1]
    while (hi - lo > 1) {
      const int mid = (lo + hi) >> 1;
      if ((int64_t)chunk_start[mid] <= g) lo = mid; else hi = mid;
    }
    const int s = lo;
    const int64_t c0 = chunk_start[s];
    b = (int64_t)seg_start[s] + (g - c0) * SORT_CHUNK;
    e = min((int64_t)seg_start[s + 1], b + SORT_CHUNK);
    tb = c0 * nlo + (g - c0);
    ts = (int64_t)chunk_start[s + 1] - c0;
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

// counts[tb + dg * ts] = number of records of chunk g with digit dg
template <int NM>
__global__ void __launch_bounds__(SORT_THREADS) k_upsweep(Src<NM> src, ChunkMap cm, KeyGeom kg, DigitSel ds,
                                                          int nb, uint32_t *__restrict__ counts) {
  constexpr int UNR = 8;
  extern __shared__ uint32_t hist[];
  int64_t b, e, tb, ts;
  if (!cm.range(blockIdx.x, b, e, tb, ts)) return;
  for (int i = threadIdx.x; i < nb; i += SORT_THREADS) hist[i] = 0;
  __syncthreads();
  for (int64_t base = b; base < e; base += SORT_THREADS * UNR) {
    float r[UNR][3];
#pragma unroll
    for (int u = 0; u < UNR; u++) {
      const int64_t i = base + u * SORT_THREADS + threadIdx.x;
      if (i < e) src.load_pos(i, r[u]);
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
    counts[tb + (int64_t)i * ts] = hist[i];
}

// Stable partition of each chunk by digit into dst at offsets[tb + dg * ts].
// perm (optional): perm[i] = output position of input record i.
// Per sub-chunk of SORT_THREADS * PL records (each wave owns 64 * PL consecutive ones):
// wavefront ballots give every record its rank among equal digits, LDS
// prefix sums turn ranks into positions of a digit-sorted staging copy, and
// the staging copy is written out as contiguous runs.  The next sub-chunk's
// records are loaded while the current one is written out.
template <int NM, int PL>
__global__ void __launch_bounds__(SORT_THREADS) k_downsweep(Src<NM> src, float *__restrict__ dst, ChunkMap cm,
                                                            KeyGeom kg, DigitSel ds, int nb, int dbits,
                                                            const uint32_t *__restrict__ offsets,
                                                            uint32_t *__restrict__ perm) {
  constexpr int RW = Src<NM>::RW;
  constexpr int SUBN = SORT_THREADS * PL;  // records per sub-chunk
  extern __shared__ uint32_t sm[];
  uint32_t *whist = sm;                         // [SORT_WAVES][nb]
  uint32_t *lstart = whist + SORT_WAVES * nb;   // [nb]
  uint32_t *gbase = lstart + nb;                // [nb]
  uint32_t *subcnt = gbase + nb;                // [nb]
  uint16_t *sdig = reinterpret_cast<uint16_t *>(subcnt + nb);          // [SUBN]
  float *stage = reinterpret_cast<float *>(sdig + SUBN);           // [SUBN][RW]
  __shared__ uint32_t wtot[SORT_WAVES];

  int64_t cb, ce, tb, ts;
  if (!cm.range(blockIdx.x, cb, ce, tb, ts)) return;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  for (int i = threadIdx.x; i < nb; i += SORT_THREADS)
    gbase[i] = offsets[tb + (int64_t)i * ts];
  for (int i = threadIdx.x; i < SORT_WAVES * nb; i += SORT_THREADS) whist[i] = 0;

  float rec[PL][RW];
  const int wbase = wave * (PL * 64) + lane;
#pragma unroll
  for (int it = 0; it < PL; it++) {
    const int64_t i = cb + wbase + it * 64;
    if (i < ce) src.load(i, rec[it]);
  }
  __syncthreads();

  for (int64_t s0 = cb; s0 < ce; s0 += SUBN) {
    const int nsub = (int)min((int64_t)SUBN, ce - s0);
    uint32_t dig[PL];
    uint32_t rnk[PL];
#pragma unroll
    for (int it = 0; it < PL; it++) {
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
    for (int it = 0; it < PL; it++) {
      const int loc = wbase + it * 64;
      if (loc < nsub) {
        const uint32_t dg = dig[it];
        const uint32_t within = whist[wave * nb + dg] + rnk[it];  // place in the digit's run of this sub-chunk
        const uint32_t p = lstart[dg] + within;
#pragma unroll
        for (int k = 0; k < RW; k++) stage[p * RW + k] = rec[it][k];
        sdig[p] = (uint16_t)dg;
        if (perm) perm[s0 + loc] = gbase[dg] + within;  // where this input record goes (coalesced)
      }
    }
    __syncthreads();
    // prefetch the next sub-chunk while this one is written out
    {
      const int64_t n0 = s0 + SUBN;
#pragma unroll
      for (int it = 0; it < PL; it++) {
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

// Two-pass tile starts from the scanned pass-2 table: the offset of a tile's
// first chunk, or the segment start when the segment is empty; tile_start[ks] = n.
__global__ void k_tile_start_2pass(const uint32_t *__restrict__ t2scan, const uint32_t *__restrict__ chunk_start,
                                   const uint32_t *__restrict__ seg_start, int nhi, int nlo, int64_t n,
                                   uint32_t *__restrict__ tile_start) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t ks = (int64_t)nhi * nlo;
  if (t > ks) return;
  if (t == ks) {
    tile_start[t] = (uint32_t)n;
    return;
  }
  const int hi = (int)(t / nlo), lo = (int)(t % nlo);
  const int64_t c0 = chunk_start[hi], nc = (int64_t)chunk_start[hi + 1] - c0;
  tile_start[t] = nc ? t2scan[c0 * nlo + lo * nc] : seg_start[hi];
}

// Single-pass tile starts: tile_start[k] = t1scan[k*g1] (chunk 0), tile_start[ks] = n
__global__ void k_tile_start_1pass(const uint32_t *__restrict__ t1scan, int64_t g1, int ks, int64_t n,
                                   uint32_t *__restrict__ tile_start) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k > ks) return;
  tile_start[k] = (k == ks) ? (uint32_t)n : t1scan[(int64_t)k * g1];
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

size_t downsweep_lds(int nb, int rw, int pl) {
  const size_t sub = (size_t)SORT_THREADS * pl;
  return (size_t)(SORT_WAVES * nb + 3 * nb) * 4 + sub * 2 + sub * rw * 4;
}

int downsweep_pl(int nb, int rw) {
  for (int pl : {8, 4})
    if (downsweep_lds(nb, rw, pl) <= 80 * 1024) return pl;
  return 2;
}

template <int NM>
void launch_downsweep(int64_t grid, hipStream_t s, Src<NM> src, float *dst, const ChunkMap &cm, const KeyGeom &kg,
                      const DigitSel &ds, int nb, int dbits, const uint32_t *offsets, uint32_t *perm) {
  constexpr int RW = Src<NM>::RW;
  const int pl = downsweep_pl(nb, RW);
  const size_t lds = downsweep_lds(nb, RW, pl);
  if (pl == 8)
    hipLaunchKernelGGL((k_downsweep<NM, 8>), dim3((unsigned)grid), dim3(SORT_THREADS), lds, s, src, dst, cm, kg, ds,
                       nb, dbits, offsets, perm);
  else if (pl == 4)
    hipLaunchKernelGGL((k_downsweep<NM, 4>), dim3((unsigned)grid), dim3(SORT_THREADS), lds, s, src, dst, cm, kg, ds,
                       nb, dbits, offsets, perm);
  else
    hipLaunchKernelGGL((k_downsweep<NM, 2>), dim3((unsigned)grid), dim3(SORT_THREADS), lds, s, src, dst, cm, kg, ds,
                       nb, dbits, offsets, perm);
  GENPK_LAUNCH_CHECK();
}

template <int NM>
TileOrder sort_impl(Plan &p, const KeyGeom &kg, int64_t n, const float *pos, const float *masses, bool with_perm,
                    bool every_tile) {
  hipStream_t s = p.stream;
  const TileGeom &tg = p.tg;

  float *A = p.recA.as<float>();
  float *B = p.recB.as<float>();
  const int64_t g1 = ceil_div(n, SORT_CHUNK);
  Src<NM> raw{pos, masses, true};
  uint32_t *perm1 = with_perm ? p.permA.as<uint32_t>() : nullptr;
  uint32_t *perm2 = (with_perm && tg.passes == 2) ? p.permB.as<uint32_t>() : nullptr;
  const float *sorted = nullptr;
  uint32_t *tile_start = nullptr;
  int64_t ks = 0;

  if (tg.passes == 1) {
    const int nb = 1 << tg.hi_bits;
    ks = nb;
    uint32_t *t1 = static_cast<uint32_t *>(p.table1.ensure((size_t)nb * g1 * 4));
    DigitSel dsel{0, (uint32_t)(nb - 1)};
    ChunkMap cm{n, 0, nullptr, nullptr, g1, 0};
    {
      Stage st(p, "sort_upsweep1");
      hipLaunchKernelGGL((k_upsweep<NM>), dim3((unsigned)g1), dim3(SORT_THREADS), nb * 4, s, raw, cm, kg, dsel, nb, t1);
      GENPK_LAUNCH_CHECK();
    }
    exclusive_scan_u32(t1, t1, (int64_t)nb * g1, p.scan_tmp, s);
    {
      Stage st(p, "sort_downsweep1");
      launch_downsweep<NM>(g1, s, raw, A, cm, kg, dsel, nb, tg.hi_bits, t1, perm1);
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
    ChunkMap cm1{n, 0, nullptr, nullptr, g1, 0};
    {
      Stage st(p, "sort_upsweep1");
      hipLaunchKernelGGL((k_upsweep<NM>), dim3((unsigned)g1), dim3(SORT_THREADS), nhi * 4, s, raw, cm1, kg, d1, nhi,
                         t1);
      GENPK_LAUNCH_CHECK();
    }
    exclusive_scan_u32(t1, t1, (int64_t)nhi * g1, p.scan_tmp, s);
    {
      Stage st(p, "sort_downsweep1");
      launch_downsweep<NM>(g1, s, raw, A, cm1, kg, d1, nhi, tg.hi_bits, t1, perm1);
    }
    // segments of pass 2
    uint32_t *seg = static_cast<uint32_t *>(p.seginfo.ensure((size_t)(nhi + 1) * 3 * 4));
    uint32_t *seg_start = seg, *nchunks = seg + (nhi + 1), *chunk_start = seg + 2 * (nhi + 1);
    hipLaunchKernelGGL(k_seginfo, dim3((unsigned)ceil_div(nhi + 1, 256)), dim3(256), 0, s, t1, g1, nhi, n, seg_start,
                       nchunks);
    GENPK_LAUNCH_CHECK();
    exclusive_scan_u32(nchunks, chunk_start, nhi + 1, p.scan_tmp, s);
    const int64_t g2 = g1 + nhi;  // upper bound of pass-2 chunks
    uint32_t *t2 = static_cast<uint32_t *>(p.table2.ensure((size_t)g2 * nlo * 4));
    DigitSel d2{0, (uint32_t)(nlo - 1)};
    ChunkMap cm2{n, nhi, seg_start, chunk_start, 0, nlo};
    Src<NM> pa{A, nullptr, false};
    {
      Stage st(p, "sort_upsweep2");
      hipLaunchKernelGGL((k_upsweep<NM>), dim3((unsigned)g2), dim3(SORT_THREADS), nlo * 4, s, pa, cm2, kg, d2, nlo,
                         t2);
      GENPK_LAUNCH_CHECK();
    }
    // the written part of the table is its first (pass-2 chunks) * nlo entries;
    // the rest, scanned along, is never read
    exclusive_scan_u32(t2, t2, g2 * nlo, p.scan_tmp, s);
    tile_start = static_cast<uint32_t *>(p.tile_start.ensure((size_t)(ks + 1) * 4));
    hipLaunchKernelGGL(k_tile_start_2pass, dim3((unsigned)ceil_div(ks + 1, 256)), dim3(256), 0, s, t2, chunk_start,
                       seg_start, nhi, nlo, n, tile_start);
    GENPK_LAUNCH_CHECK();
    {
      Stage st(p, "sort_downsweep2");
      launch_downsweep<NM>(g2, s, pa, B, cm2, kg, d2, nlo, tg.lo_bits, t2, perm2);
    }
    sorted = B;
  }

  // work items
  const int store = every_tile ? 1 : 0;
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
  TileOrder o;
  o.rec = sorted;
  o.rw = Src<NM>::RW;
  o.has_mass = NM != 0;
  o.has_perm = with_perm;
  o.perm1 = perm1;
  o.perm2 = perm2;
  o.tile_start = tile_start;
  o.item_start = item_start;
  o.part_start = part_start;
  o.nheavy = nheavy;
  o.heavy = heavy;
  o.ks = ks;
  o.n = n;
  o.max_items = max_items;
  return o;
}
}  // namespace

KeyGeom key_geom(const Plan &p, double boxsize) {
  KeyGeom g;
  g.units = (double)p.d / boxsize;
  g.d = (int)p.d;
  g.TX = p.tg.TX; g.TY = p.tg.TY; g.TZ = p.tg.TZ;
  g.sx = p.tg.sx; g.sy = p.tg.sy; g.sz = p.tg.sz;
  g.ntx = p.tg.ntx; g.nty = p.tg.nty; g.ntz = p.tg.ntz;
  return g;
}

TileOrder sort_by_tile(Plan &p, const KeyGeom &kg, int64_t n, const float *pos, const float *masses, bool with_perm,
                       bool every_tile) {
  GENPK_REQUIRE(n > 0 && n <= ((int64_t)1 << 31), "sort_by_tile: particle count out of range");
  const int rw = 3 + (masses ? 1 : 0);
  // with the permutations, both record buffers can also take the adjoint
  // gather's output records (up to 4 floats)
  const size_t rec_bytes = (size_t)n * (with_perm ? 4 : rw) * 4;
  p.recA.ensure(rec_bytes);
  if (p.tg.passes == 2 || with_perm) p.recB.ensure(rec_bytes);
  if (with_perm) {
    p.permA.ensure((size_t)n * 4);
    if (p.tg.passes == 2) p.permB.ensure((size_t)n * 4);
  }
  if (masses) return sort_impl<1>(p, kg, n, pos, masses, with_perm, every_tile);
  return sort_impl<0>(p, kg, n, pos, nullptr, with_perm, every_tile);
}

}  // namespace genpk
