// Partition of particles by deposit tile on gfx950: an MSD radix partition
// of one or two passes over the tile key (upsweep histogram in LDS ->
// offset scan -> downsweep that ranks records with LDS atomics, stages them
// in LDS by digit and writes contiguous runs), then the work items of the
// tile kernels.  Large two-pass sorts run their first pass paged: each
// workgroup appends each digit's records to blocks of its own, so the
// positions are read once (no histogram pass), and the second pass reads
// every digit's blocks through a block list.  Global atomics are avoided:
// on this chip they run at ~27 Gop/s.  The order of the records within one
// tile is not specified (the tile kernels do not depend on it).
#include "tilesort.h"

#include <algorithm>

namespace genpk {

namespace {

// 16 waves per workgroup, one workgroup per CU in the downsweep: a
// sub-chunk of 1024 * PL records is staged in LDS at a time and written out
// as one run per digit, and longer runs write faster (at 2^30 records and
// 256 digits, 8192-record sub-chunks take 6.9 ms per downsweep against 7.2
// ms for two 512-thread workgroups per CU staging 4096 records each).
constexpr int SORT_THREADS = 1024;
constexpr int SORT_WAVES = SORT_THREADS / 64;
// Records per workgroup: from SORT_CHUNK_MIN (one sub-chunk of the
// downsweep) up to SORT_CHUNK_MAX, the smallest that keeps at most four
// chunks per CU (small sets then still spread over every CU).  32 Ki rather
// than 64 Ki: the paged second pass's count kernel runs more, shorter
// workgroups (2^27 particles at 512^3: 0.23 -> 0.17 ms) and the downsweeps
// are unchanged or slightly faster; 16 Ki is no faster.
constexpr int64_t SORT_CHUNK_MIN = 8192, SORT_CHUNK_MAX = 32768;
// The paged second pass's count kernel: four waves per workgroup, so that
// eight workgroups (chunks) per CU overlap their memory round trips.
constexpr int UP_THREADS = 256;
constexpr int UP_WAVES = UP_THREADS / 64;

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
  int64_t chunk;               // records per chunk (a multiple of the downsweep's sub-chunk)
  __device__ __forceinline__ bool range(int64_t g, int64_t &b, int64_t &e, int64_t &tb, int64_t &ts) const {
    if (nseg == 0) {
      b = g * chunk;
      e = min(n, b + chunk);
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
    b = (int64_t)seg_start[s] + (g - c0) * chunk;
    e = min((int64_t)seg_start[s + 1], b + chunk);
    tb = c0 * nlo + (g - c0);
    ts = (int64_t)chunk_start[s + 1] - c0;
    return b < e;
  }
};

// Write target of the downsweep's staging positions past a sub-chunk's last
// record (the write-out issues the same stores on every path): many
// workgroups may write it at once, and nothing reads it.
__device__ float sort_sink[SORT_THREADS * 8 * 4];
__device__ uint8_t sort_sink_digits[SORT_THREADS * 8];

// Paged pass 1 (MODE 1 and 2 of k_downsweep, two-pass sorts of many
// records): the pass-1 output pool is cut into blocks of PG_BLOCK record
// slots, Q per pass-1 workgroup.
constexpr int PG_LOG = 10;
constexpr uint32_t PG_BLOCK = 1u << PG_LOG;

// A block-list entry: the block and its fill (1 to PG_BLOCK records) in one
// word, so that one load gives both.
__host__ __device__ constexpr uint32_t pg_entry(uint32_t blk, uint32_t fill) { return blk | (fill - 1) << 22; }
__device__ __forceinline__ uint32_t pg_blk(uint32_t e) { return e & ((1u << 22) - 1); }
__device__ __forceinline__ uint32_t pg_fill(uint32_t e) { return (e >> 22) + 1; }

struct PgArgs {
  uint32_t *hdr = nullptr;          // [G * Q] of each taken block: digit | (its number among the digit's blocks) << 8
  uint32_t *fill = nullptr;         // [G * Q] records in each taken block
  uint32_t *used = nullptr;         // [G] blocks taken by each pass-1 workgroup
  int64_t Q = 0;                    // blocks per pass-1 workgroup
  const uint32_t *blist = nullptr;  // pass 2: the taken blocks, segment (pass-1 digit) by segment (pg_entry)
};

// Tile key of a record with the cells along the axes not in `axes` (bit a:
// axis a of x, y, z) taken as 0 and not computed.
__device__ __forceinline__ uint32_t key_on_axes(const float *r, const KeyGeom &g, int axes) {
  double fr;
  const int cx = (axes & 1) ? cell_of(r[0], g.units, g.d, fr) : 0;
  const int cy = (axes & 2) ? cell_of(r[1], g.units, g.d, fr) : 0;
  const int cz = (axes & 4) ? cell_of(r[2], g.units, g.d, fr) : 0;
  return (uint32_t)((((cx >> g.sx) * g.nty + (cy >> g.sy)) * g.ntz) + (cz >> g.sz));
}

struct DigitSel {
  int shift;
  uint32_t mask;
  int axes;  // bit a set: the digit depends on the tile index along axis a (x, y, z)
  __device__ __forceinline__ uint32_t of_key(uint32_t key) const { return (key >> shift) & mask; }
  // digit of a record; cells along axes the digit does not depend on are not computed
  __device__ __forceinline__ uint32_t of(const float *r, const KeyGeom &g) const {
    return of_key(key_on_axes(r, g, axes));
  }
};

int log2_floor(int v) {
  int l = 0;
  while ((1 << (l + 1)) <= v) l++;
  return l;
}

// Digit of key bits [shift, shift + width).  The key is (tx * nty + ty) * ntz
// + tz; when nty and ntz are powers of two its bits are tz, then ty, then tx,
// and a digit needs only the axes whose bits it covers.
DigitSel digit_sel(const KeyGeom &g, int shift, int width) {
  DigitSel s{shift, (uint32_t)((1u << width) - 1), 7};
  if ((g.nty & (g.nty - 1)) != 0 || (g.ntz & (g.ntz - 1)) != 0) return s;
  const int bz = log2_floor(g.ntz), by = log2_floor(g.nty);
  const int lo[3] = {bz + by, bz, 0}, hi[3] = {32, bz + by, bz};
  s.axes = 0;
  for (int a = 0; a < 3; a++)
    if (lo[a] < hi[a] && lo[a] < shift + width && shift < hi[a]) s.axes |= 1 << a;
  return s;
}

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
      const uint32_t dg = valid ? ds.of(r[u], kg) : 0u;
      // a wavefront whose records share one digit adds them with one atomic
      const uint64_t vmask = __ballot(valid);
      const uint32_t d0 = __builtin_amdgcn_readlane(dg, (int)__builtin_ctzll(vmask | (1ull << 63)));
      if (__ballot(valid && dg == d0) == vmask) {
        if (valid && (vmask & lanemask_lt()) == 0) atomicAdd(&hist[d0], (uint32_t)__popcll(vmask));
      } else if (valid) {
        atomicAdd(&hist[dg], 1u);
      }
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < nb; i += SORT_THREADS)
    counts[tb + (int64_t)i * ts] = hist[i];
}

// k_upsweep's counts for a digit of at most 8 bits from the digit bytes the
// previous downsweep wrote beside its output records (digs[i]: digit of
// record i), so the records themselves are not read.  Reads the 4-byte words
// covering [b, e): digs holds at least 4 bytes past the last record.
__global__ void __launch_bounds__(SORT_THREADS) k_upsweep_bytes(const uint8_t *__restrict__ digs, ChunkMap cm,
                                                                int nb, uint32_t *__restrict__ counts) {
  constexpr int UNR = 4;
  extern __shared__ uint32_t hist[];
  int64_t b, e, tb, ts;
  if (!cm.range(blockIdx.x, b, e, tb, ts)) return;
  for (int i = threadIdx.x; i < nb; i += SORT_THREADS) hist[i] = 0;
  __syncthreads();
  const uint32_t *words = reinterpret_cast<const uint32_t *>(digs);
  const int64_t w0 = b >> 2, w1 = (e + 3) >> 2;
  for (int64_t base = w0; base < w1; base += SORT_THREADS * UNR) {
    uint32_t v[UNR];
#pragma unroll
    for (int u = 0; u < UNR; u++) {
      const int64_t w = base + u * SORT_THREADS + threadIdx.x;
      v[u] = w < w1 ? words[w] : 0u;
    }
#pragma unroll
    for (int u = 0; u < UNR; u++) {
      const int64_t w = base + u * SORT_THREADS + threadIdx.x;
      // a word of four equal in-range digits adds them with one atomic, and
      // a wavefront of such words of one digit with one atomic for all
      const uint32_t d0 = v[u] & 0xffu;
      const bool same4 = 4 * w >= b && 4 * w + 4 <= e && v[u] == d0 * 0x01010101u;
      const uint32_t dw = __builtin_amdgcn_readfirstlane(d0);
      if (__ballot(same4 && d0 == dw) == ~0ull) {
        if ((threadIdx.x & 63) == 0) atomicAdd(&hist[dw], 256u);
      } else if (same4) {
        atomicAdd(&hist[d0], 4u);
      } else {
#pragma unroll
        for (int j = 0; j < 4; j++) {
          const int64_t i = 4 * w + j;
          if (i >= b && i < e) atomicAdd(&hist[(v[u] >> (8 * j)) & 0xffu], 1u);
        }
      }
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < nb; i += SORT_THREADS)
    counts[tb + (int64_t)i * ts] = hist[i];
}

// Rank of each of a lane's PL record slots among the wave's slots of its
// digit in this sub-chunk, from LDS atomic adds that return the old count on
// the wave's count row (a wave whose 64 slots share one digit adds them with
// one atomic, from lane 0).  Slots it >= nv hold no record: they take the
// digit `trash` (one past the last), so that they rank and stage like
// records, after every real one.  Equal digits are ordered by the LDS unit's
// processing order of one instruction's lanes and, where waves share a count
// row, by the order in which they run: the order of the records within a
// tile is not specified (the tile kernels do not depend on it).  pk
// receives each slot's rank (bits 0-12, below the sub-chunk's 8192 slots),
// its digit from bit 13 and, with `two` (digits of at most 8 bits, so the
// digit takes bits 13-21), its digit of ds2 (from the same key) from bit 22:
// one register per slot.  Called by whole waves.
template <int NM, int PL>
__device__ __forceinline__ void rank_records(const float (&rec)[PL][Src<NM>::RW], uint32_t nv, const KeyGeom &kg,
                                             const DigitSel &ds, const DigitSel &ds2, bool two, uint32_t trash,
                                             uint32_t *wrow, uint32_t (&pk)[PL]) {
  const uint32_t lane = threadIdx.x & 63;
  uint32_t rnk[PL];
  uint32_t uni = 0;  // bit it: the wave's slots `it` share one digit
#pragma unroll
  for (int it = 0; it < PL; it++) {
    const uint32_t key = tile_key(rec[it], kg);
    const uint32_t dg = (uint32_t)it < nv ? ds.of_key(key) : trash;
    const uint32_t d0 = __builtin_amdgcn_readfirstlane(dg);
    uint32_t r = 0;
    if (__ballot(dg == d0) == ~0ull) {
      if (lane == 0) r = atomicAdd(&wrow[d0], 64u);
      uni |= 1u << it;
    } else {
      r = atomicAdd(&wrow[dg], 1u);
    }
    pk[it] = dg << 13 | (two ? ds2.of_key(key) << 22 : 0u);
    rnk[it] = r;
  }
  // a one-digit wave's slots follow lane 0's old count
#pragma unroll
  for (int it = 0; it < PL; it++)
    pk[it] |= (uni >> it) & 1u ? __builtin_amdgcn_readlane(rnk[it], 0) + lane : rnk[it];
}

// Per-row counts of digit dg -> per-row exclusive offsets in place; returns the total.
__device__ __forceinline__ uint32_t wave_offsets(uint32_t *whist, int nb, int hw, int dg) {
  uint32_t tot = 0;
  for (int w = 0; w < hw; w++) {
    const uint32_t t = whist[w * nb + dg];
    whist[w * nb + dg] = tot;
    tot += t;
  }
  return tot;
}

// One round of a block-wide exclusive scan (one value per thread), added to
// and advancing the running total 'carry'.  Contains a barrier; wtot rows
// alternate between rounds.
__device__ __forceinline__ uint32_t block_scan_round(uint32_t v, uint32_t (*wtot)[SORT_WAVES], int rd,
                                                     uint32_t &carry) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  uint32_t incl = v;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const uint32_t t = __shfl_up(incl, off, 64);
    if (lane >= off) incl += t;
  }
  if (lane == 63) wtot[rd & 1][wave] = incl;
  __syncthreads();
  uint32_t run = carry + incl - v, all = 0;
#pragma unroll
  for (int w = 0; w < SORT_WAVES; w++) {
    const uint32_t t = wtot[rd & 1][w];
    if (w < wave) run += t;
    all += t;
  }
  carry += all;
  return run;
}

// Partition of each chunk by digit.  perm (optional): perm[i] = output
// position of input record i.  Per sub-chunk of SORT_THREADS * PL record
// slots (each wave owns 64 * PL consecutive ones): ranks (rank_records; the
// slots past the chunk's end take the digit nb, one past the last), then one
// thread per digit turns the per-wave counts into offsets and a block scan
// gives the digit's start in a digit-sorted staging copy and dbase = (global
// output offset - staging start) mod 2^32, records go to the staging copy,
// and the staging copy's records (every slot before digit nb's start) are
// written out as contiguous runs (record p goes to dbase[its digit] + p)
// while the next sub-chunk's records load.  With DB the per-wave counts are
// double-buffered: this sub-chunk's buffer is cleared during the write-out
// and the next sub-chunk ranks into the other one, so no barrier separates
// them.  hw count rows (a divisor of the wave count) serve the waves, wave w
// on row w % hw.  A digit's staging start, global base, dbase and second
// dbase sit together in one 16-byte LDS word (dinfo).
//   MODE 0: the chunk's output offsets per digit are offsets[tb + dg * ts].
//   MODE 1 (paged pass 1): workgroup g appends each digit's records to
//     blocks of PG_BLOCK slots that it takes in order from its own range
//     [g Q, g Q + Q) of the pool: the digit's current block until it is
//     full, then new ones (a run then has two parts, the second from
//     staging position dinfo[dg].y on, at dbase dinfo[dg].w).  Each block's
//     digit and number among the digit's blocks go to pg.hdr, its fill to
//     pg.fill, the number of blocks taken to pg.used[g].  No histogram pass
//     and no global atomics.
//   MODE 2 (paged pass 2): the chunk covers list slots of its segment's
//     blocks (pg.blist); slot v is record (block of blist[v / PG_BLOCK], v %
//     PG_BLOCK) when below that block's fill, and is skipped otherwise.
//     perm is indexed by pool slot.
template <int NM, int PL, bool DB, int MODE>
__global__ void __launch_bounds__(SORT_THREADS, 4) k_downsweep(Src<NM> src, float *__restrict__ dst, ChunkMap cm,
                                                               KeyGeom kg, DigitSel ds, int nb, int hw,
                                                               const uint32_t *__restrict__ offsets,
                                                               uint32_t *__restrict__ perm,
                                                               uint8_t *__restrict__ dig2, DigitSel ds2,
                                                               uint32_t *__restrict__ mrange, PgArgs pg) {
  constexpr int RW = Src<NM>::RW;
  constexpr int SUBN = SORT_THREADS * PL;
  constexpr int NBUF = DB ? 2 : 1;
  constexpr uint32_t NONE = 0xffffffffu;
  static_assert(MODE != 2 || PL * 64 <= (int)PG_BLOCK, "a wave's records of a sub-chunk lie in one block");
  static_assert(SUBN <= 8192, "ranks take 13 bits (rank_records)");
  extern __shared__ uint4 sm4[];
  const int nbt = nb + 1;                                    // digits and the empty slots' digit nb
  uint4 *dinfo = sm4;                                        // [nbt] staging start, gbase, dbase, dbase2
  uint32_t *whist_all = reinterpret_cast<uint32_t *>(dinfo + nbt);  // [NBUF][hw][nbt]
  uint32_t *cur_blk = whist_all + NBUF * hw * nbt;           // MODE 1 [nb]: the digit's current block
  uint32_t *cur_fill = cur_blk + (MODE == 1 ? nb : 0);       // MODE 1 [nb]: its fill (PG_BLOCK: none open)
  uint32_t *dblocks = cur_fill + (MODE == 1 ? nb : 0);       // MODE 1 [nb]: blocks the digit took
  uint16_t *sdig = reinterpret_cast<uint16_t *>(dblocks + (MODE == 1 ? nb : 0));  // [SUBN]
  float *stage = reinterpret_cast<float *>(sdig + SUBN);      // [SUBN][RW]
  __shared__ uint32_t wtot[2][SORT_WAVES];
  __shared__ uint32_t smr[2];  // masses: largest, smallest nonzero |mass| bits of this chunk

  int64_t cb, ce, tb, ts;
  if (!cm.range(blockIdx.x, cb, ce, tb, ts)) return;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  // with dig2 (digits of at most 8 bits), sdig holds the digit and the ds2
  // digit (paged pass 1 always writes dig2, paged pass 2 never)
  const bool two = MODE == 1 ? true : MODE == 2 ? false : dig2 != nullptr;
  if (MODE != 1)
    for (int i = threadIdx.x; i < nb; i += SORT_THREADS) dinfo[i].y = offsets[tb + (int64_t)i * ts];
  if (MODE == 1)
    for (int i = threadIdx.x; i < nb; i += SORT_THREADS) {
      cur_blk[i] = NONE;
      cur_fill[i] = PG_BLOCK;
      dblocks[i] = 0;
    }
  const uint32_t blk0 = MODE == 1 ? (uint32_t)(blockIdx.x * pg.Q) : 0u;  // this workgroup's block range
  uint32_t taken = 0;                                                    // MODE 1: blocks taken so far
  for (int i = threadIdx.x; i < NBUF * hw * nbt; i += SORT_THREADS) whist_all[i] = 0;
  const int row = (wave % hw) * nbt;
  if (threadIdx.x == 0) {
    smr[0] = 0u;
    smr[1] = 0xffffffffu;
  }
  uint32_t mx = 0u, mn = 0xffffffffu;

  float rec[PL][RW] = {};
  uint32_t nv = 0;  // this lane's slots of the sub-chunk that hold records (a prefix: it < nv)
  const int wbase = wave * (PL * 64) + lane;
  // MODE 2: a wave's 64 * PL consecutive list slots (PL * 64 <= PG_BLOCK,
  // sub-chunks start at multiples of PG_BLOCK) lie in one block: its pool
  // offset and fill are looked up once per wave and sub-chunk
  uint32_t wblk = 0, wfill = 0;
  // records of the sub-chunk at s0 (MODE 2: through the block list)
  auto load_sub = [&](int64_t s0) {
    const int64_t i0 = s0 + wbase;
    int64_t lim = ce - i0;  // slot i0 + 64 it holds a record iff 64 it < lim
    if constexpr (MODE == 2) {
      // the wave's first slot as a uniform value: the entry is a scalar load
      const int64_t w0 = s0 + __builtin_amdgcn_readfirstlane(wave) * (PL * 64);
      wfill = 0;
      if (w0 < ce) {
        const uint32_t ent = pg.blist[w0 >> PG_LOG];
        wblk = pg_blk(ent);
        wfill = pg_fill(ent);
      }
      lim = min(lim, (int64_t)wfill - (int64_t)((uint32_t)i0 & (PG_BLOCK - 1)));
    }
    nv = lim <= 0 ? 0u : (uint32_t)min((int64_t)PL, (lim + 63) >> 6);
#pragma unroll
    for (int it = 0; it < PL; it++) {
      const int64_t i = i0 + it * 64;
      if ((uint32_t)it < nv) {
        if constexpr (MODE == 2) src.load(wblk * PG_BLOCK + ((uint32_t)i & (PG_BLOCK - 1)), rec[it]);
        else src.load(i, rec[it]);
      }
    }
  };
  load_sub(cb);
  // nothing in flight on entry to the loop, so that the wait for a
  // sub-chunk's records there is not a wait for the write-out's stores
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();

  int sub = 0;
  for (int64_t s0 = cb; s0 < ce; s0 += SUBN, sub++) {
    uint32_t *whist = whist_all + (DB ? (sub & 1) * hw * nbt : 0);
    uint32_t pk[PL];
    rank_records<NM, PL>(rec, nv, kg, ds, ds2, two, (uint32_t)nb, whist + row, pk);
    __syncthreads();
    {
      uint32_t total = 0;  // slots of this sub-chunk (MODE 1: and blocks taken << 16)
      for (int r0 = 0, rd = 0; r0 < nbt; r0 += SORT_THREADS, rd++) {
        const int dg = r0 + threadIdx.x;
        const uint32_t tot = dg < nbt ? wave_offsets(whist, nbt, hw, dg) : 0u;
        if constexpr (MODE == 1) {
          // the run fills the current block (a records), then `need` new
          // blocks; counts and block numbers scan together (nb <= 256, one
          // round: at most SUBN slots and SUBN / PG_BLOCK + nb blocks)
          const bool real = dg < nb;
          const uint32_t cf = real ? cur_fill[dg] : PG_BLOCK;
          const uint32_t a = min(tot, PG_BLOCK - cf);
          const uint32_t need = real ? (tot - a + PG_BLOCK - 1) >> PG_LOG : 0u;
          const uint32_t run = block_scan_round(tot | need << 16, wtot, rd, total);
          if (real) {
            const uint32_t lr = run & 0xffffu, first = blk0 + taken + (run >> 16);
            const uint32_t cbk = cur_blk[dg];
            // dbase used only when a > 0, i.e. cbk is open
            dinfo[dg] = make_uint4(lr, lr + a, cbk * PG_BLOCK + cf - lr, first * PG_BLOCK - (lr + a));
            if (need) {
              if (cbk != NONE) pg.fill[cbk] = PG_BLOCK;  // filled by this run
              const uint32_t nd = dblocks[dg];
              for (uint32_t k = 0; k < need; k++) {
                pg.hdr[first + k] = (uint32_t)dg | (nd + k) << 8;
                if (k + 1 < need) pg.fill[first + k] = PG_BLOCK;
              }
              dblocks[dg] = nd + need;
              cur_blk[dg] = first + need - 1;
              cur_fill[dg] = tot - a - (need - 1) * PG_BLOCK;
            } else {
              cur_fill[dg] = cf + tot;
            }
          } else if (dg == nb) {
            dinfo[nb].x = run & 0xffffu;
          }
        } else {
          const uint32_t run = block_scan_round(tot, wtot, rd, total);
          if (dg < nb) {
            const uint32_t g = dinfo[dg].y;
            dinfo[dg].x = run;
            dinfo[dg].y = g + tot;
            dinfo[dg].z = g - run;
          } else if (dg == nb) {
            dinfo[nb].x = run;
          }
        }
      }
      if constexpr (MODE == 1) taken += total >> 16;
    }
    __syncthreads();
    // output position of the record at staging position p of digit d
    auto dest_of = [&](uint32_t p, uint32_t d) -> uint32_t {
      if constexpr (MODE == 1) {
        const uint4 di = dinfo[d];
        return p < di.y ? di.z + p : di.w + p;
      } else {
        return dinfo[d].z + p;
      }
    };
#pragma unroll
    for (int it = 0; it < PL; it++) {
      const int loc = wbase + it * 64;
      const uint32_t dg = two ? (pk[it] >> 13) & 0x1ffu : pk[it] >> 13;
      const uint32_t p = dinfo[dg].x + whist[row + dg] + (pk[it] & 0x1fffu);
#pragma unroll
      for (int k = 0; k < RW; k++) stage[p * RW + k] = rec[it][k];
      sdig[p] = (uint16_t)(two ? dg | (pk[it] >> 22) << 8 : dg);
      if (dg < (uint32_t)nb) {
        if constexpr (NM == 1) {
          // |mass| as bits orders like the value (a NaN above every finite one)
          const uint32_t mb = __float_as_uint(rec[it][3]) & 0x7fffffffu;
          mx = max(mx, mb);
          if (mb) mn = min(mn, mb);
        }
        // where this record goes (coalesced writes, by input position or pool slot)
        if (perm) perm[MODE == 2 ? (int64_t)(wblk * PG_BLOCK + ((uint32_t)(s0 + loc) & (PG_BLOCK - 1))) : s0 + loc] = dest_of(p, dg);
      }
    }
    // the next sub-chunk loads during the barrier and the write-out
    load_sub(s0 + SUBN);
    __syncthreads();
    const uint32_t nrec = dinfo[nb].x;  // records of this sub-chunk: staging positions [0, nrec)
    // record_p: staging position p holds a record (p < nrec); the others
    // are written to the sink
    auto put = [&](int p, bool record_p) {
      const uint32_t sd = sdig[p];
      const uint32_t dest = dest_of((uint32_t)p, record_p ? (two ? sd & 0xffu : sd) : 0u);
      float v[RW];
#pragma unroll
      for (int k = 0; k < RW; k++) v[k] = stage[p * RW + k];
      // the sink position from p computed in the loop (not hoisted out of it)
      int ps = p;
      asm volatile("" : "+v"(ps));
      float *to = record_p ? dst + (int64_t)dest * RW : sort_sink + ps * RW;
#pragma unroll
      for (int k = 0; k < RW; k++) to[k] = v[k];
      if (two) *(record_p ? dig2 + dest : sort_sink_digits + ps) = (uint8_t)(sd >> 8);
    };
    // With the sink (paged passes; pass 1 only with masses, where it measured
    // faster), every lane stores for each of its staging positions, so that
    // every path issues the same stores and the next sub-chunk's loads
    // (issued before them) are waited for without waiting for these writes;
    // elsewhere the extra registers cost more than that wait.  Pass 1's
    // two-part runs write two positions at a time to keep the registers in
    // bounds.
    constexpr bool SINK = MODE == 2 || (MODE == 1 && NM == 1);
    if constexpr (SINK && MODE == 1) {
#pragma unroll 2
      for (int it = 0; it < PL; it++) {
        const int p = threadIdx.x + it * SORT_THREADS;
        put(p, p < (int)nrec);
      }
    } else if constexpr (SINK) {
#pragma unroll
      for (int it = 0; it < PL; it++) {
        const int p = threadIdx.x + it * SORT_THREADS;
        put(p, p < (int)nrec);
      }
    } else if (nrec == SUBN) {
      if constexpr (MODE == 1) {
#pragma unroll 2
        for (int it = 0; it < PL; it++) put(threadIdx.x + it * SORT_THREADS, true);
      } else {
#pragma unroll
        for (int it = 0; it < PL; it++) put(threadIdx.x + it * SORT_THREADS, true);
      }
    } else {
      for (int p = threadIdx.x; p < (int)nrec; p += SORT_THREADS) put(p, true);
    }
    for (int i = threadIdx.x; i < hw * nbt; i += SORT_THREADS) whist[i] = 0;
    if (!DB) __syncthreads();
  }
  if constexpr (MODE == 1) {
    // the open blocks' fills (each digit's state is its own thread's)
    for (int dg = threadIdx.x; dg < nb; dg += SORT_THREADS)
      if (cur_blk[dg] != NONE) pg.fill[cur_blk[dg]] = cur_fill[dg];
    if (threadIdx.x == 0) pg.used[blockIdx.x] = taken;
  }
  if constexpr (NM == 1) {
    if (mrange) {
      atomicMax(&smr[0], mx);
      atomicMin(&smr[1], mn);
      __syncthreads();
      if (threadIdx.x == 0) {
        atomicMax(&mrange[0], smr[0]);
        atomicMin(&mrange[1], smr[1]);
      }
    }
  }
}

__global__ void k_mrange_init(uint32_t *mrange) {
  mrange[0] = 0u;
  mrange[1] = 0xffffffffu;
}

// Pass-2 segment bookkeeping from the scanned pass-1 table.
__global__ void k_seginfo(const uint32_t *__restrict__ t1scan, int64_t g1, int nseg, int64_t n, int64_t chunk,
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
  nchunks[s] = (uint32_t)(((int64_t)en - st + chunk - 1) / chunk);
}

// Two-pass tile starts from the scanned pass-2 table: the offset of a tile's
// first chunk, or the segment's first record (rec_start) when the segment is
// empty; tile_start[ks] = n.
__global__ void k_tile_start_2pass(const uint32_t *__restrict__ t2scan, const uint32_t *__restrict__ chunk_start,
                                   const uint32_t *__restrict__ rec_start, int nhi, int nlo, int64_t n,
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
  tile_start[t] = nc ? t2scan[c0 * nlo + lo * nc] : rec_start[hi];
}

// Paged pass 1's block directory.  Per pass-1 workgroup w (one workgroup
// each): the blocks and records of each digit among the blocks it took, at
// bcnt / rcnt[d * G + w]; entry nhi * G is zero, so that exclusive scans of
// the nhi * G + 1 entries end in the totals.
__global__ void __launch_bounds__(256) k_pg_count(PgArgs pg, int nhi, int64_t G, uint32_t *__restrict__ bcnt,
                                                  uint32_t *__restrict__ rcnt) {
  __shared__ uint32_t cb[256], cr[256];
  const int64_t w = blockIdx.x;
  for (int i = threadIdx.x; i < nhi; i += 256) cb[i] = cr[i] = 0;
  __syncthreads();
  const uint32_t used = (uint32_t)min((int64_t)pg.used[w], pg.Q);
  const int64_t b0 = w * pg.Q;
  for (uint32_t k = threadIdx.x; k < used; k += 256) {
    const uint32_t d = pg.hdr[b0 + k] & 0xffu;
    if (d < (uint32_t)nhi) {
      atomicAdd(&cb[d], 1u);
      atomicAdd(&cr[d], pg.fill[b0 + k]);
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < nhi; i += 256) {
    bcnt[(int64_t)i * G + w] = cb[i];
    rcnt[(int64_t)i * G + w] = cr[i];
  }
  if (w == 0 && threadIdx.x == 0) {
    bcnt[(int64_t)nhi * G] = 0;
    rcnt[(int64_t)nhi * G] = 0;
  }
}

// The block list: digit d's blocks from pass-1 workgroup w at boff[d * G +
// w] onwards, in the order the workgroup took them (each block's number
// among its digit's blocks is in its header), so a segment lists its records
// in input order, as the plain pass 1 writes them.
__global__ void __launch_bounds__(256) k_pg_list(PgArgs pg, int nhi, int64_t G, const uint32_t *__restrict__ boff,
                                                 uint32_t *__restrict__ blist) {
  __shared__ uint32_t at[256], cnt[256];
  const int64_t w = blockIdx.x;
  for (int i = threadIdx.x; i < nhi; i += 256) {
    at[i] = boff[(int64_t)i * G + w];
    cnt[i] = boff[(int64_t)i * G + w + 1] - at[i];
  }
  __syncthreads();
  const uint32_t used = (uint32_t)min((int64_t)pg.used[w], pg.Q);
  const int64_t b0 = w * pg.Q;
  for (uint32_t k = threadIdx.x; k < used; k += 256) {
    const uint32_t h = pg.hdr[b0 + k], d = h & 0xffu;
    if (d < (uint32_t)nhi && (h >> 8) < cnt[d]) blist[at[d] + (h >> 8)] = pg_entry((uint32_t)(b0 + k), pg.fill[b0 + k]);
  }
}

// Pass-2 segments over the block list: segment s (pass-1 digit s) covers
// list slots [seg_start[s], seg_start[s+1]) (PG_BLOCK per listed block), its
// records start at rec_start[s] of the output, and it has nchunks[s] chunks
// of `chunk` list slots.
__global__ void k_pg_seginfo(const uint32_t *__restrict__ boff, const uint32_t *__restrict__ roff, int nhi,
                             int64_t G, int64_t chunk, uint32_t *__restrict__ seg_start,
                             uint32_t *__restrict__ rec_start, uint32_t *__restrict__ nchunks) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s > nhi) return;
  const uint32_t b0 = boff[(int64_t)s * G];
  seg_start[s] = b0 * PG_BLOCK;
  rec_start[s] = roff[(int64_t)s * G];
  nchunks[s] = s < nhi ? (uint32_t)(((int64_t)(boff[(int64_t)(s + 1) * G] - b0) * PG_BLOCK + chunk - 1) / chunk) : 0u;
}

// k_upsweep_bytes over the list slots of a paged pass 2: the digit bytes of
// the listed blocks' filled slots.  Chunks start at multiples of PG_BLOCK
// list slots; each wave takes whole blocks (one lookup of the block and its
// fill), each lane 16 consecutive slots of it.
__global__ void __launch_bounds__(UP_THREADS) k_upsweep_pages(const uint8_t *__restrict__ digs, ChunkMap cm,
                                                                PgArgs pg, int nb, uint32_t *__restrict__ counts) {
  static_assert(PG_BLOCK == 64 * 16, "one block per wave, 16 slots per lane");
  extern __shared__ uint32_t hist[];
  int64_t b, e, tb, ts;
  if (!cm.range(blockIdx.x, b, e, tb, ts)) return;
  for (int i = threadIdx.x; i < nb; i += UP_THREADS) hist[i] = 0;
  __syncthreads();
  // the wave as a uniform value: the block-list entries are scalar loads
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  const int64_t k0 = b >> PG_LOG, k1 = (e + PG_BLOCK - 1) >> PG_LOG;  // list entries of the chunk
  constexpr int UNR = 4;  // blocks per wave in flight
  for (int64_t kb = k0 + wave; kb < k1; kb += UNR * UP_WAVES) {
    uint32_t blk[UNR];
    int nv[UNR];
    uint4 v[UNR];
#pragma unroll
    for (int u = 0; u < UNR; u++) {
      const int64_t k = kb + u * UP_WAVES;
      const uint32_t ent = k < k1 ? pg.blist[k] : 0u;
      blk[u] = pg_blk(ent);
      // filled slots of this block inside the chunk, from this lane's 16
      const int64_t lim = k < k1 ? min((int64_t)pg_fill(ent), e - (k << PG_LOG)) : 0;
      nv[u] = (int)max((int64_t)0, min((int64_t)16, lim - 16 * lane));
    }
#pragma unroll
    for (int u = 0; u < UNR; u++) {
      v[u] = make_uint4(0u, 0u, 0u, 0u);
      if (nv[u] > 0) v[u] = *reinterpret_cast<const uint4 *>(digs + (int64_t)blk[u] * PG_BLOCK + 16 * lane);
    }
#pragma unroll
    for (int u = 0; u < UNR; u++) {
      // 16 equal digits add with one atomic, and a wavefront of such lanes
      // of one digit with one atomic for all
      const uint32_t d0 = v[u].x & 0xffu;
      const uint32_t rep = d0 * 0x01010101u;
      const bool same16 = nv[u] == 16 && v[u].x == rep && v[u].y == rep && v[u].z == rep && v[u].w == rep;
      const uint32_t dw = __builtin_amdgcn_readfirstlane(d0);
      if (__ballot(same16 && d0 == dw) == ~0ull) {
        if (lane == 0) atomicAdd(&hist[dw], 1024u);
      } else if (same16) {
        atomicAdd(&hist[d0], 16u);
      } else {
#pragma unroll
        for (int q = 0; q < 4; q++) {
          const uint32_t wq = q == 0 ? v[u].x : q == 1 ? v[u].y : q == 2 ? v[u].z : v[u].w;
#pragma unroll
          for (int j = 0; j < 4; j++)
            if (4 * q + j < nv[u]) atomicAdd(&hist[(wq >> (8 * j)) & 0xffu], 1u);
        }
      }
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < nb; i += UP_THREADS)
    counts[tb + (int64_t)i * ts] = hist[i];
}

// Single-pass tile starts: tile_start[k] = t1scan[k*g1] (chunk 0), tile_start[ks] = n
__global__ void k_tile_start_1pass(const uint32_t *__restrict__ t1scan, int64_t g1, int ks, int64_t n,
                                   uint32_t *__restrict__ tile_start) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k > ks) return;
  tile_start[k] = (k == ks) ? (uint32_t)n : t1scan[(int64_t)k * g1];
}

// Work items per tile, at most imax records each.  store mode: every real
// tile gets at least one item (so untouched tiles are written as zeros); add
// mode: empty tiles are skipped.
__global__ void k_items(const uint32_t *__restrict__ tile_start, int64_t ks, int64_t kreal, int store, uint32_t imax,
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
  if (k < kreal) m = c == 0 ? (store ? 1u : 0u) : (c + imax - 1) / imax;
  mcnt[k] = m;
  pcnt[k] = m > 1 ? m : 0;
  if (k == 0) *nheavy = 0;
}

// item_tile[i] = tile of work item i
__global__ void k_item_tiles(const uint32_t *__restrict__ item_start, int64_t ks, uint32_t *__restrict__ item_tile) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= ks) return;
  for (uint32_t i = item_start[k]; i < item_start[k + 1]; i++) item_tile[i] = (uint32_t)k;
}

__global__ void k_heavy_list(const uint32_t *__restrict__ pcnt, int64_t ks, uint32_t *__restrict__ nheavy,
                             uint32_t *__restrict__ heavy) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= ks) return;
  if (pcnt[k] > 1) heavy[atomicAdd(nheavy, 1u)] = (uint32_t)k;
}

struct DsShape {
  int pl;
  bool db;
  int hw;
};

// LDS of a downsweep shape (nb digits and the empty slots' digit); MODE 1
// keeps three more arrays per digit.
size_t downsweep_lds(int nb, int rw, const DsShape &h, int mode = 0) {
  const size_t sub = (size_t)SORT_THREADS * h.pl, nbt = (size_t)nb + 1;
  return ((4 + (h.db ? 2 : 1) * h.hw) * nbt + (mode == 1 ? 3 * (size_t)nb : 0)) * 4 + sub * 2 + sub * rw * 4;
}

// The first shape whose LDS fits one workgroup per CU: four count rows (each
// shared by four waves: 16 rows measured 0.1 ms slower per downsweep at 2^30
// records, the per-digit offsets running over more rows), then the most
// records per lane, then double-buffered counts.
DsShape downsweep_shape(int nb, int rw, int mode = 0) {
  const size_t cap = 160 * 1024 - 256;
  for (int hw : {SORT_WAVES / 4, 1})
    for (int pl : {8, 4, 2})
      for (bool db : {true, false}) {
        const DsShape h{pl, db, hw};
        if (downsweep_lds(nb, rw, h, mode) <= cap) return h;
      }
  return DsShape{1, false, 1};
}

// MODE 1 and 2 run with 8 records per lane only (sub-chunks of 8192 records)
template <int NM, int MODE = 0>
void launch_downsweep(int64_t grid, hipStream_t s, Src<NM> src, float *dst, const ChunkMap &cm, const KeyGeom &kg,
                      const DigitSel &ds, int nb, const uint32_t *offsets, uint32_t *perm,
                      uint8_t *dig2 = nullptr, const DigitSel &ds2 = DigitSel{0, 0, 0},
                      uint32_t *mrange = nullptr, const PgArgs &pg = PgArgs{}) {
  GENPK_REQUIRE(nb <= 2 * SORT_THREADS, "sort digit too wide");
  GENPK_REQUIRE(!dig2 || (nb <= 256 && ds2.mask <= 255u), "sort digit bytes need digits of at most 8 bits");
  const DsShape h = downsweep_shape(nb, Src<NM>::RW, MODE);
  const size_t lds = downsweep_lds(nb, Src<NM>::RW, h, MODE);
  const dim3 g((unsigned)grid), b(SORT_THREADS);
  if constexpr (MODE != 0) {
    GENPK_REQUIRE(h.pl == 8 && nb <= 256 && (MODE != 1 || dig2), "paged sort pass: unsupported shape");
    if (h.db)
      hipLaunchKernelGGL((k_downsweep<NM, 8, true, MODE>), g, b, lds, s, src, dst, cm, kg, ds, nb, h.hw, offsets,
                         perm, dig2, ds2, mrange, pg);
    else
      hipLaunchKernelGGL((k_downsweep<NM, 8, false, MODE>), g, b, lds, s, src, dst, cm, kg, ds, nb, h.hw, offsets,
                         perm, dig2, ds2, mrange, pg);
  } else if (h.pl == 8 && h.db) {
    hipLaunchKernelGGL((k_downsweep<NM, 8, true, 0>), g, b, lds, s, src, dst, cm, kg, ds, nb, h.hw, offsets, perm,
                       dig2, ds2, mrange, pg);
  } else if (h.pl == 8) {
    hipLaunchKernelGGL((k_downsweep<NM, 8, false, 0>), g, b, lds, s, src, dst, cm, kg, ds, nb, h.hw, offsets, perm,
                       dig2, ds2, mrange, pg);
  } else if (h.pl == 4 && h.db) {
    hipLaunchKernelGGL((k_downsweep<NM, 4, true, 0>), g, b, lds, s, src, dst, cm, kg, ds, nb, h.hw, offsets, perm,
                       dig2, ds2, mrange, pg);
  } else if (h.pl == 4) {
    hipLaunchKernelGGL((k_downsweep<NM, 4, false, 0>), g, b, lds, s, src, dst, cm, kg, ds, nb, h.hw, offsets, perm,
                       dig2, ds2, mrange, pg);
  } else if (h.pl == 2) {
    hipLaunchKernelGGL((k_downsweep<NM, 2, false, 0>), g, b, lds, s, src, dst, cm, kg, ds, nb, h.hw, offsets, perm,
                       dig2, ds2, mrange, pg);
  } else {
    hipLaunchKernelGGL((k_downsweep<NM, 1, false, 0>), g, b, lds, s, src, dst, cm, kg, ds, nb, h.hw, offsets, perm,
                       dig2, ds2, mrange, pg);
  }
  GENPK_LAUNCH_CHECK();
}

// Whether and how a sort runs its first pass paged: two passes with digits
// of at most 8 bits, and so many records that the partly filled blocks (at
// most one per pass-1 workgroup and digit) leave fewer empty slots than 1/4
// of the records (half that on average).  G = twice the CU count workgroups
// (two rounds of one per CU even out their finishing times: 7.8 -> 7.5 ms
// at 2^30 records) take `chunk` records each (whole sub-chunks) and Q =
// chunk / PG_BLOCK + nhi blocks: every digit's blocks but its last are
// full.  With masses (16-byte records) the paged downsweeps keep a few
// registers in scratch and are still faster than the histogram pass they
// replace (37.5 -> 36.3 ms per retaining forward at 2^30).
struct PgLayout {
  bool on = false;
  int64_t G = 0, chunk = 0, Q = 0;
  int64_t slots = 0;  // G * Q * PG_BLOCK record slots of the pass-1 output
};

PgLayout pg_layout(const Plan &p, int64_t n, int rw) {
  PgLayout L;
  const TileGeom &tg = p.tg;
  if (tg.passes != 2 || tg.hi_bits > 8 || tg.lo_bits > 8) return L;
  const int64_t nhi = (int64_t)1 << tg.hi_bits, cus = num_cus();
  if (2 * cus * nhi * (int64_t)PG_BLOCK > n / 4) return L;
  for (int mode : {1, 2})
    if (downsweep_shape((int)nhi, rw, mode).pl != 8) return L;
  constexpr int64_t SUBN = (int64_t)SORT_THREADS * 8;
  L.chunk = ceil_div(ceil_div(n, 2 * cus), SUBN) * SUBN;
  L.G = ceil_div(n, L.chunk);
  L.Q = L.chunk / PG_BLOCK + nhi;
  L.slots = L.G * L.Q * PG_BLOCK;
  L.on = L.slots <= ((int64_t)1 << 32);  // block numbers below 2^22 (pg_entry)
  return L;
}

template <int NM>
TileOrder sort_impl(Plan &p, const KeyGeom &kg, int64_t n, const float *pos, const float *masses, bool with_perm,
                    bool every_tile, const PgLayout &L) {
  hipStream_t s = p.stream;
  const TileGeom &tg = p.tg;

  float *A = p.recA.as<float>();
  float *B = p.recB.as<float>();
  int64_t chunk = SORT_CHUNK_MIN;
  while (chunk < SORT_CHUNK_MAX && chunk * 4 * num_cus() < n) chunk *= 2;
  const int64_t g1 = ceil_div(n, chunk);
  Src<NM> raw{pos, masses, true};
  uint32_t *perm1 = with_perm ? p.permA.as<uint32_t>() : nullptr;
  uint32_t *perm2 = (with_perm && tg.passes == 2) ? p.permB.as<uint32_t>() : nullptr;
  const float *sorted = nullptr;
  uint32_t *tile_start = nullptr;
  int64_t ks = 0;
  // with masses, pass 1 (which reads every mass) records their range for the deposit
  uint32_t *mrange = nullptr;
  if (NM == 1) {
    mrange = static_cast<uint32_t *>(p.mrange.ensure(64));
    hipLaunchKernelGGL(k_mrange_init, dim3(1), dim3(1), 0, s, mrange);
    GENPK_LAUNCH_CHECK();
  }

  if (tg.passes == 1) {
    const int nb = 1 << tg.hi_bits;
    ks = nb;
    uint32_t *t1 = static_cast<uint32_t *>(p.table1.ensure((size_t)nb * g1 * 4));
    const DigitSel dsel = digit_sel(kg, 0, tg.hi_bits);
    ChunkMap cm{n, 0, nullptr, nullptr, g1, 0, chunk};
    {
      Stage st(p, "sort_upsweep1");
      hipLaunchKernelGGL((k_upsweep<NM>), dim3((unsigned)g1), dim3(SORT_THREADS), nb * 4, s, raw, cm, kg, dsel, nb, t1);
      GENPK_LAUNCH_CHECK();
    }
    exclusive_scan_u32(t1, t1, (int64_t)nb * g1, p.scan_tmp, s);
    {
      Stage st(p, "sort_downsweep1");
      launch_downsweep<NM>(g1, s, raw, A, cm, kg, dsel, nb, t1, perm1, nullptr, DigitSel{0, 0, 0}, mrange);
    }
    tile_start = static_cast<uint32_t *>(p.tile_start.ensure((size_t)(ks + 1) * 4));
    hipLaunchKernelGGL(k_tile_start_1pass, dim3((unsigned)ceil_div(ks + 1, 256)), dim3(256), 0, s, t1, g1, (int)ks,
                       n, tile_start);
    GENPK_LAUNCH_CHECK();
    sorted = A;
  } else {
    const int nhi = 1 << tg.hi_bits, nlo = 1 << tg.lo_bits;
    ks = (int64_t)nhi * nlo;
    const DigitSel d1 = digit_sel(kg, tg.lo_bits, tg.hi_bits);
    const DigitSel d2 = digit_sel(kg, 0, tg.lo_bits);
    // with digits of at most 8 bits, pass 1 writes each record's low digit as
    // a byte beside it (in its output slot), and pass 2 counts those bytes
    // instead of reading the records
    const int64_t slots = L.on ? L.slots : n;  // pass-1 output slots
    uint8_t *dig2 = tg.lo_bits <= 8 && tg.hi_bits <= 8
                        ? static_cast<uint8_t *>(p.dig2.ensure((size_t)slots + 64))
                        : nullptr;
    // segments of pass 2 (one per pass-1 digit): start (records, or list
    // slots when paged), chunk count and first chunk, first record
    uint32_t *seg = static_cast<uint32_t *>(p.seginfo.ensure((size_t)(nhi + 1) * 4 * 4));
    uint32_t *seg_start = seg, *nchunks = seg + (nhi + 1), *chunk_start = seg + 2 * (nhi + 1);
    uint32_t *rec_start = seg_start;
    int64_t g2 = 0;  // upper bound of pass-2 chunks
    PgArgs pg;
    if (L.on) {
      // pass 1, paged: the positions are read once
      GENPK_REQUIRE(chunk % PG_BLOCK == 0, "paged sort: pass-2 chunks must be whole blocks");
      const int64_t nbk = L.G * L.Q, nt = (int64_t)nhi * L.G + 1;
      uint32_t *pb = static_cast<uint32_t *>(p.pages.ensure((size_t)(4 * nbk + L.G + 4 * nt) * 4));
      pg.hdr = pb;
      pg.fill = pb + nbk;
      pg.used = pb + 2 * nbk;
      pg.Q = L.Q;
      uint32_t *blist = pg.used + L.G;
      uint32_t *bcnt = blist + nbk, *rcnt = bcnt + nt, *boff = rcnt + nt, *roff = boff + nt;
      rec_start = seg + 3 * (nhi + 1);
      ChunkMap cm1{n, 0, nullptr, nullptr, L.G, 0, L.chunk};
      {
        Stage st(p, "sort_downsweep1");
        launch_downsweep<NM, 1>(L.G, s, raw, A, cm1, kg, d1, nhi, nullptr, perm1, dig2, d2, mrange, pg);
      }
      {
        Stage st(p, "sort_pages");
        hipLaunchKernelGGL(k_pg_count, dim3((unsigned)L.G), dim3(256), 0, s, pg, nhi, L.G, bcnt, rcnt);
        GENPK_LAUNCH_CHECK();
        exclusive_scan_u32(bcnt, boff, nt, p.scan_tmp, s);
        exclusive_scan_u32(rcnt, roff, nt, p.scan_tmp, s);
        hipLaunchKernelGGL(k_pg_list, dim3((unsigned)L.G), dim3(256), 0, s, pg, nhi, L.G, boff, blist);
        GENPK_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_pg_seginfo, dim3((unsigned)ceil_div(nhi + 1, 256)), dim3(256), 0, s, boff, roff, nhi,
                           L.G, chunk, seg_start, rec_start, nchunks);
        GENPK_LAUNCH_CHECK();
      }
      pg.blist = blist;
      g2 = ceil_div(nbk * PG_BLOCK, chunk) + nhi;
    } else {
      // pass 1: high digit over [0, n)
      uint32_t *t1 = static_cast<uint32_t *>(p.table1.ensure((size_t)nhi * g1 * 4));
      ChunkMap cm1{n, 0, nullptr, nullptr, g1, 0, chunk};
      {
        Stage st(p, "sort_upsweep1");
        hipLaunchKernelGGL((k_upsweep<NM>), dim3((unsigned)g1), dim3(SORT_THREADS), nhi * 4, s, raw, cm1, kg, d1,
                           nhi, t1);
        GENPK_LAUNCH_CHECK();
      }
      exclusive_scan_u32(t1, t1, (int64_t)nhi * g1, p.scan_tmp, s);
      {
        Stage st(p, "sort_downsweep1");
        launch_downsweep<NM>(g1, s, raw, A, cm1, kg, d1, nhi, t1, perm1, dig2, d2, mrange);
      }
      hipLaunchKernelGGL(k_seginfo, dim3((unsigned)ceil_div(nhi + 1, 256)), dim3(256), 0, s, t1, g1, nhi, n, chunk,
                         seg_start, nchunks);
      GENPK_LAUNCH_CHECK();
      g2 = g1 + nhi;
    }
    exclusive_scan_u32(nchunks, chunk_start, nhi + 1, p.scan_tmp, s);
    uint32_t *t2 = static_cast<uint32_t *>(p.table2.ensure((size_t)g2 * nlo * 4));
    ChunkMap cm2{n, nhi, seg_start, chunk_start, 0, nlo, chunk};
    Src<NM> pa{A, nullptr, false};
    {
      Stage st(p, "sort_upsweep2");
      if (L.on)
        hipLaunchKernelGGL(k_upsweep_pages, dim3((unsigned)g2), dim3(UP_THREADS), nlo * 4, s, dig2, cm2, pg, nlo,
                           t2);
      else if (dig2)
        hipLaunchKernelGGL(k_upsweep_bytes, dim3((unsigned)g2), dim3(SORT_THREADS), nlo * 4, s, dig2, cm2, nlo, t2);
      else
        hipLaunchKernelGGL((k_upsweep<NM>), dim3((unsigned)g2), dim3(SORT_THREADS), nlo * 4, s, pa, cm2, kg, d2, nlo,
                           t2);
      GENPK_LAUNCH_CHECK();
    }
    // the written part of the table is its first (pass-2 chunks) * nlo entries;
    // the rest, scanned along, is never read
    exclusive_scan_u32(t2, t2, g2 * nlo, p.scan_tmp, s);
    tile_start = static_cast<uint32_t *>(p.tile_start.ensure((size_t)(ks + 1) * 4));
    hipLaunchKernelGGL(k_tile_start_2pass, dim3((unsigned)ceil_div(ks + 1, 256)), dim3(256), 0, s, t2, chunk_start,
                       rec_start, nhi, nlo, n, tile_start);
    GENPK_LAUNCH_CHECK();
    {
      Stage st(p, "sort_downsweep2");
      if (L.on)
        launch_downsweep<NM, 2>(g2, s, pa, B, cm2, kg, d2, nlo, t2, perm2, nullptr, DigitSel{0, 0, 0}, nullptr, pg);
      else
        launch_downsweep<NM>(g2, s, pa, B, cm2, kg, d2, nlo, t2, perm2);
    }
    sorted = B;
  }

  // work items
  const int store = every_tile ? 1 : 0;
  uint32_t *ib = static_cast<uint32_t *>(p.item_start.ensure((size_t)(ks + 1) * 4 * 4 + 64));
  uint32_t *mcnt = ib, *item_start = ib + (ks + 1), *pcnt = ib + 2 * (ks + 1), *part_start = ib + 3 * (ks + 1);
  uint32_t *hv = static_cast<uint32_t *>(p.heavy.ensure((size_t)(ks + 2) * 4));
  uint32_t *nheavy = hv, *heavy = hv + 1;
  const uint32_t imax = deposit_item_max(n, tg.K);
  hipLaunchKernelGGL(k_items, dim3((unsigned)ceil_div(ks + 1, 256)), dim3(256), 0, s, tile_start, ks, tg.K, store,
                     imax, mcnt, pcnt, nheavy);
  GENPK_LAUNCH_CHECK();
  exclusive_scan_u32(mcnt, item_start, ks + 1, p.scan_tmp, s);
  exclusive_scan_u32(pcnt, part_start, ks + 1, p.scan_tmp, s);
  hipLaunchKernelGGL(k_heavy_list, dim3((unsigned)ceil_div(ks, 256)), dim3(256), 0, s, pcnt, ks, nheavy, heavy);
  GENPK_LAUNCH_CHECK();
  const int64_t max_items = tg.K + ceil_div(n, (int64_t)imax);
  uint32_t *item_tile = static_cast<uint32_t *>(p.item_tile.ensure((size_t)max_items * 4));
  hipLaunchKernelGGL(k_item_tiles, dim3((unsigned)ceil_div(ks, 256)), dim3(256), 0, s, item_start, ks, item_tile);
  GENPK_LAUNCH_CHECK();

  TileOrder o;
  o.rec = sorted;
  o.rw = Src<NM>::RW;
  o.has_mass = NM != 0;
  o.has_perm = with_perm;
  o.perm1 = perm1;
  o.perm2 = perm2;
  o.tile_start = tile_start;
  o.item_start = item_start;
  o.item_tile = item_tile;
  o.part_start = part_start;
  o.nheavy = nheavy;
  o.heavy = heavy;
  o.mrange = mrange;
  o.ks = ks;
  o.n = n;
  o.max_items = max_items;
  o.item_max = imax;
  return o;
}
}  // namespace

uint32_t deposit_item_max(int64_t n, int64_t tiles) {
  const int64_t cus = num_cus();
  uint32_t m = DEPOSIT_ITEM_MAX;
  while (m > DEPOSIT_ITEM_MIN && (tiles + n / m < 4 * cus || (int64_t)m * 2 * cus > n)) m /= 2;
  return m;
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

TileOrder sort_by_tile(Plan &p, const KeyGeom &kg, int64_t n, const float *pos, const float *masses, bool with_perm,
                       bool every_tile) {
  GENPK_REQUIRE(n > 0 && n <= ((int64_t)1 << 31), "sort_by_tile: particle count out of range");
  const int rw = 3 + (masses ? 1 : 0);
  const PgLayout L = pg_layout(p, n, rw);
  // with the permutations, both record buffers can also take the adjoint
  // gather's output records (up to 4 floats); recA holds the pass-1 output
  // (paged: L.slots record slots), perm2 is indexed by its slots
  const size_t rec_bytes = (size_t)n * (with_perm ? 4 : rw) * 4;
  p.recA.ensure(L.on ? std::max(rec_bytes, (size_t)L.slots * rw * 4) : rec_bytes);
  if (p.tg.passes == 2 || with_perm) p.recB.ensure(rec_bytes);
  if (with_perm) {
    p.permA.ensure((size_t)n * 4);
    if (p.tg.passes == 2) p.permB.ensure((size_t)(L.on ? L.slots : n) * 4);
  }
  if (masses) return sort_impl<1>(p, kg, n, pos, masses, with_perm, every_tile, L);
  return sort_impl<0>(p, kg, n, pos, nullptr, with_perm, every_tile, L);
}

}  // namespace genpk
