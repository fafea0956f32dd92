// 3-D real-to-complex FFT on the padded in-place layout the reference hands
// to FFTW (gen-pk.cpp:176-193: real [d][d][2(d/2+1)] in, complex
// [d][d][d/2+1] out, unnormalised, forward sign -1), and its unnormalised
// inverse.  Three passes over HBM:
//   z: r2c of each contiguous row; d reals are packed as d/2 complex, one
//      d/2-point Stockham FFT in LDS, then the even/odd split to d/2+1 bins;
//   y, x: length-d c2c FFTs over column blocks of B consecutive kz, loaded
//      as B-wide rows (coalesced) into LDS, transformed in place.
#include "fft_core.h"
#include "plan.h"
#include "tiles.h"

#include <cmath>
#include <vector>

namespace genpk {

namespace {

// Workgroup shapes (measured at 1024^3): the z passes run 512 threads and
// the column passes 1024 threads over the same LDS footprint (ELEMS complex
// values per workgroup, two workgroups per CU), i.e. MAXE staged values per
// thread; more waves per CU hide the global-memory latency of the staged
// loads (z pass 3.7 -> 3.0 ms, y pass 2.9 -> 2.0 ms against 256 threads).
constexpr int FFT_THREADS = 512;   // z passes
constexpr int COL_THREADS = 1024;  // y and x passes
template <typename T> struct FftCfg;
template <> struct FftCfg<float> { static constexpr int ELEMS = 8192; static constexpr int MAXE = 16; static constexpr int COLE = 8; };
template <> struct FftCfg<double> { static constexpr int ELEMS = 4096; static constexpr int MAXE = 8; static constexpr int COLE = 4; };

// z pass, forward: rows of d reals -> d/2+1 complex, L rows per workgroup.
// FROM_TILES: the rows are assembled from the deposit's tile store while
// loading, halo cells folded in (the fold), so the real-space grid is never
// written.
// The LDS footprint (up to 64 KB) allows two workgroups per CU, so the
// register budget of two waves per SIMD is available.  The rows sit in LDS
// with one element of padding per 32 (lds_ix), which spreads the strided
// writes of the FFT stages over the banks.
// BFLY (from the tile store, a first stage of radix 8, L * M / 8 <= 2 *
// FFT_THREADS): each thread loads the eight inputs of up to two butterflies
// of the first stage straight from the tile store and runs that stage from
// registers.  CT = 512: the plan is M = 512 = 8 * 8 * 8 (d = 1024), run with
// compile-time stages (fft_stage_ct); 0: any plan.  STD (with CT = 512):
// the default 16 x 32 x 32 tiles, whose store offsets (below 2^31 at d =
// 1024) are computed in 32 bits from compile-time strides, in the butterfly
// loader (FP32) or the element-wise one (FP64).
template <typename T, bool FROM_TILES, bool BFLY = false, int CT = 0, bool STD = false>
__global__ void __launch_bounds__(FFT_THREADS) __attribute__((amdgpu_waves_per_eu(1, 2))) k_r2c_z(T *__restrict__ field, const T *__restrict__ tiles,
                                                      FoldGeom fg, FastDiv fdd, int64_t rowlen, int64_t nrows, int L,
                                                      FastDiv fdM, FastDiv fdM1, RadixPlan rp,
                                                      const typename Cx<T>::C *__restrict__ twM,
                                                      const typename Cx<T>::C *__restrict__ twd) {
  using C = typename Cx<T>::C;
  constexpr int MAXE = FftCfg<T>::MAXE;
  extern __shared__ unsigned char lds_raw[];
  C *buf = reinterpret_cast<C *>(lds_raw);
  const int M = CT ? CT : rp.n;
  const int64_t row0 = (int64_t)xcd_block(blockIdx.x, gridDim.x) * L;
  const int nl = (int)min((int64_t)L, nrows - row0);
  const int tot = L * M;
  // twiddles of the M-point FFT and of the even/odd split, staged in LDS
  C *stw = buf + lds_len<true>((int64_t)L * M);  // [M]
  C *sxw = stw + M;               // [M+1]
  if constexpr (BFLY) {
    // butterfly u = threadIdx.x + FFT_THREADS * i: row u / NQ, inputs
    // u % NQ + NQ * r (NQ = M / 8)
    const int NQ = M / 8;
#pragma unroll
    for (int i = 0; i < 2; i++) {
      const int u = threadIdx.x + FFT_THREADS * i;
      const int b = CT ? u / (CT / 8) : (int)fdM.div((uint32_t)(8 * u)), j = u - b * NQ;
      if (b < L) {
        C w[8];
        if (STD && b < nl) {
          static_assert(!STD || CT == 512, "standard tiles with the d = 1024 loaders");
          // d = 1024, tiles 16 x 32 x 32 (17 x 33 halo'd rows of 32 per tile):
          // z = 2 (j + 64 r) lies in tile column tz = 4 r + j / 16 at lz = 2 (j % 16)
          constexpr uint32_t CM = 17 * 33 * 32, CZ = 17 * 33;
          const uint32_t row = (uint32_t)(row0 + b);
          const uint32_t x = row >> 10, y = row & 1023u;
          const uint32_t txy = (x >> 4) * 32u + (y >> 5), rr = (x & 15u) * 33u + (y & 31u);
          const uint32_t jz = (uint32_t)j >> 4, lz = 2u * ((uint32_t)j & 15u);
          const uint32_t base = (txy * 32u + jz) * CM + rr * 32u + lz;
          const T *tb = tiles + base;
#pragma unroll
          for (int r = 0; r < 8; r++) w[r] = C{tb[r * 4 * CM], tb[r * 4 * CM + 1]};
          if (lz == 0) {
            // the previous tile along z's z-halo cell: tile column 4 r + jz - 1
            // (for r = jz = 0 the last column, 31, periodically)
#pragma unroll
            for (int r = 0; r < 8; r++) {
              const uint32_t tzp = (4u * r + jz + 31u) & 31u;
              w[r].x = w[r].x + tiles[(uint32_t)fg.zoff + (txy * 32u + tzp) * CZ + rr];  // one rounding, as below
            }
          }
        } else if (b < nl) {
          const int64_t row = row0 + b;
          const int x = (int)fdd.div((uint32_t)row), y = (int)(row - (int64_t)x * fg.d);
          const int64_t txy = (int64_t)(x >> fg.sx) * fg.nty + (y >> fg.sy);
          const int64_t rr = (int64_t)(x & (fg.TX - 1)) * fg.HY + (y & (fg.TY - 1));
          T zh8[8];
#pragma unroll
          for (int r = 0; r < 8; r++) {
            const int z = 2 * (j + NQ * r);
            const int tz = z >> fg.sz, lz = z & (fg.TZ - 1);
            const int64_t at = txy * fg.ntz * fg.cm + (int64_t)tz * fg.cm + rr * fg.TZ + lz;
            w[r] = C{tiles[at], tiles[at + 1]};
            zh8[r] = (T)0;
            if (lz == 0) {
              // the previous tile along z's z-halo cell (in its row store when partial)
              const int tzp = tz == 0 ? fg.ntz - 1 : tz - 1;
              const int hw = min(fg.TZ, fg.d - tzp * fg.TZ);
              const int64_t h = hw < fg.TZ ? (int64_t)tzp * fg.cm + hw + txy * fg.ntz * fg.cm + rr * fg.TZ
                                           : fg.zoff + (int64_t)tzp * fg.cz + txy * fg.ntz * fg.cz + rr;
              zh8[r] = tiles[h];
            }
          }
#pragma unroll
          for (int r = 0; r < 8; r++) w[r].x = w[r].x + zh8[r];  // as the element-wise load: one rounding
        } else {
#pragma unroll
          for (int r = 0; r < 8; r++) w[r] = C{0, 0};
        }
        Dft<8, -1, C>::run(w);
#pragma unroll
        for (int r = 0; r < 8; r++) buf[lds_ix<true>(b * M + 8 * j + r)] = w[r];
      }
    }
    for (int i = threadIdx.x; i < M; i += FFT_THREADS) stw[i] = twM[i];
    for (int i = threadIdx.x; i <= M; i += FFT_THREADS) sxw[i] = twd[i];
    __syncthreads();
    if constexpr (CT == 512)
      lds_fft_ct<512, 8, -1, false, MAXE, false, C, true, 8, 8>(buf, buf, L, 0, stw);
    else
      for (int st = 1; st < rp.nst; st++)
        fft_stage_r<-1, false, MAXE, false, C, true>(rp.st[st].R, buf, buf, L, 0, 1, M, rp.st[st], stw);
  } else {
  C v[MAXE];
  T zh[FROM_TILES ? MAXE : 1];
  if (FROM_TILES && M == FFT_THREADS && STD) {
    // d = 1024 with the default tiles (the FP64 plan): as below, with 32-bit
    // store offsets from compile-time strides (the store holds fewer than
    // 2^31 values; every tile is full, so a z halo sits in the z-halo plane)
    constexpr uint32_t CM = 17 * 33 * 32, CZ = 17 * 33;
    const uint32_t z = 2u * threadIdx.x, tz = z >> 5, lz = z & 31u, tzp = (tz + 31u) & 31u;
#pragma unroll
    for (int q = 0; q < MAXE; q++) {
      v[q] = C{0, 0};
      zh[FROM_TILES ? q : 0] = (T)0;
      if (q < L && q < nl) {
        const uint32_t row = (uint32_t)(row0 + q);
        const uint32_t x = row >> 10, y = row & 1023u;
        const uint32_t txy = (x >> 4) * 32u + (y >> 5), rr = (x & 15u) * 33u + (y & 31u);
        const uint32_t at = (txy * 32u + tz) * CM + rr * 32u + lz;
        v[q] = C{tiles[at], tiles[at + 1]};
        if (lz == 0) zh[FROM_TILES ? q : 0] = tiles[(uint32_t)fg.zoff + (txy * 32u + tzp) * CZ + rr];
      }
    }
  } else if (FROM_TILES && M == FFT_THREADS) {
    // element q of this thread is row row0 + q, reals 2m and 2m + 1 with m =
    // threadIdx.x: the row's part of the store offset is the same for the
    // whole workgroup (scalar), the z part is the lane's own
    const int z = 2 * (int)threadIdx.x;
    const int tz = z >> fg.sz, lz = z & (fg.TZ - 1);
    const int64_t lane_off = (int64_t)tz * fg.cm + lz;
    // z-halo cell added to the first cell of a tile along z: local z = width
    // of the previous tile, in its row store when that tile is partial
    const int tzp = tz == 0 ? fg.ntz - 1 : tz - 1;
    const int hw = min(fg.TZ, fg.d - tzp * fg.TZ);
    const bool hrow = hw < fg.TZ;
    const int64_t h_off = hrow ? (int64_t)tzp * fg.cm + hw : fg.zoff + (int64_t)tzp * fg.cz;
    const int64_t h_tmul = hrow ? (int64_t)fg.ntz * fg.cm : (int64_t)fg.ntz * fg.cz;
    const int64_t h_rmul = hrow ? fg.TZ : 1;
#pragma unroll
    for (int q = 0; q < MAXE; q++) {
      v[q] = C{0, 0};
      zh[FROM_TILES ? q : 0] = (T)0;
      if (q < L && q < nl) {
        const int64_t row = row0 + q;
        const int x = (int)fdd.div((uint32_t)row), y = (int)(row - (int64_t)x * fg.d);
        const int64_t txy = (int64_t)(x >> fg.sx) * fg.nty + (y >> fg.sy);
        const int64_t r = (int64_t)(x & (fg.TX - 1)) * fg.HY + (y & (fg.TY - 1));
        const int64_t at = txy * fg.ntz * fg.cm + r * fg.TZ + lane_off;
        v[q] = C{tiles[at], tiles[at + 1]};
        if (lz == 0) zh[FROM_TILES ? q : 0] = tiles[h_off + txy * h_tmul + r * h_rmul];
      }
    }
  } else {
#pragma unroll
  for (int q = 0; q < MAXE; q++) {
    const int i = threadIdx.x + q * FFT_THREADS;
    v[q] = C{0, 0};
    if (FROM_TILES) zh[FROM_TILES ? q : 0] = (T)0;
    if (i < tot) {
      const int l = (int)fdM.div((uint32_t)i), m = i - l * M;
      if (l < nl) {
        const int64_t row = row0 + l;
        if (FROM_TILES) {
          // both reals sit in one aligned tile row (TZ even); the first cell
          // of a tile along z adds the previous tile's z-halo cell
          const int x = (int)fdd.div((uint32_t)row), y = (int)(row - (int64_t)x * fg.d);
          const int64_t at = cell_at(fg, x, y, 2 * m);
          v[q] = C{tiles[at], tiles[at + 1]};
          if (((2 * m) & (fg.TZ - 1)) == 0) zh[FROM_TILES ? q : 0] = tiles[zhalo_at(fg, x, y, 2 * m)];
        } else {
          const T *r = field + row * rowlen;
          v[q] = C{r[2 * m], r[2 * m + 1]};
        }
      }
    }
  }
  }
#pragma unroll
  for (int q = 0; q < MAXE; q++) {
    const int i = threadIdx.x + q * FFT_THREADS;
    // z-halo added after all loads are in flight (a + b rounds once, as fold_value)
    if (FROM_TILES) v[q].x = v[q].x + zh[FROM_TILES ? q : 0];
    if (i < tot) buf[lds_ix<true>(i)] = v[q];
  }
  for (int i = threadIdx.x; i < M; i += FFT_THREADS) stw[i] = twM[i];
  for (int i = threadIdx.x; i <= M; i += FFT_THREADS) sxw[i] = twd[i];
  __syncthreads();
  if constexpr (CT == 512)
    lds_fft_ct<512, 1, -1, false, MAXE, false, C, true, 8, 8, 8>(buf, buf, L, 0, stw);
  else
    lds_fft<-1, false, MAXE, C, true>(buf, rp, L, 0, 1, M, stw);
  }
  // X_k = (Z_k + conj Z_{M-k})/2 - i/2 * w^k (Z_k - conj Z_{M-k}),  w = exp(-2 pi i / d)
  auto split = [&](C zk, C zc, C w) {
    const C e = C{(zk.x + zc.x) * (T)0.5, (zk.y + zc.y) * (T)0.5};
    const C o = C{(zk.y - zc.y) * (T)0.5, -(zk.x - zc.x) * (T)0.5};  // (zk - zc) / (2i)
    return cadd(e, cmul(o, w));
  };
  if (M == FFT_THREADS) {
    // output k = threadIdx.x of every row (L <= MAXE rows), and k = M by thread 0
    const int k = threadIdx.x;
    const C w = sxw[k], wm = sxw[M];
#pragma unroll
    for (int l = 0; l < MAXE; l++) {
      if (l < nl) {
        const C zk = buf[lds_ix<true>(l * M + k)];
        const C zc = cconj(buf[lds_ix<true>(l * M + (k == 0 ? 0 : M - k))]);
        C *row = reinterpret_cast<C *>(field + (row0 + l) * rowlen);
        row[k] = split(zk, zc, w);
        if (k == 0) row[M] = split(zk, zc, wm);
      }
    }
    return;
  }
  if (FFT_THREADS % M == 0) {
    // M a smaller power of two: output k = threadIdx.x % M of rows
    // threadIdx.x / M + FFT_THREADS / M * i, and k = M by the threads of k = 0
    const int k = threadIdx.x & (M - 1), rpi = FFT_THREADS / M;
    const C w = sxw[k], wm = sxw[M];
#pragma unroll 4
    for (int l = threadIdx.x / M; l < nl; l += rpi) {
      const C zk = buf[lds_ix<true>(l * M + k)];
      const C zc = cconj(buf[lds_ix<true>(l * M + (k == 0 ? 0 : M - k))]);
      C *row = reinterpret_cast<C *>(field + (row0 + l) * rowlen);
      row[k] = split(zk, zc, w);
      if (k == 0) row[M] = split(zk, zc, wm);
    }
    return;
  }
  const int nout = M + 1;
  for (int i = threadIdx.x; i < nl * nout; i += FFT_THREADS) {
    const int l = (int)fdM1.div((uint32_t)i), k = i - l * nout;
    const C zk = buf[lds_ix<true>(l * M + (k == M ? 0 : k))];
    const C zc = cconj(buf[lds_ix<true>(l * M + (k == 0 ? 0 : M - k))]);
    reinterpret_cast<C *>(field + (row0 + l) * rowlen)[k] = split(zk, zc, sxw[k]);
  }
}

// Elements of the inverse z pass's row buffer: L rows of M + 1 while the
// half spectrum is unfolded, then L padded rows of M for the transform.
__host__ __device__ inline int64_t c2r_span(int64_t L, int64_t M) {
  return L * (M + 1) > lds_len<true>(L * M) ? L * (M + 1) : lds_len<true>(L * M);
}

// z pass, inverse: d/2+1 complex -> d reals (unnormalised c2r); the
// imaginary parts of kz = 0 and kz = d/2 do not contribute.
template <typename T>
__global__ void __launch_bounds__(FFT_THREADS) k_c2r_z(T *__restrict__ field, int64_t rowlen, int64_t nrows, int L,
                                                      FastDiv fdM, FastDiv fdM1, RadixPlan rp,
                                                      const typename Cx<T>::C *__restrict__ twM,
                                                      const typename Cx<T>::C *__restrict__ twd) {
  using C = typename Cx<T>::C;
  constexpr int MAXE = FftCfg<T>::MAXE;
  extern __shared__ unsigned char lds_raw[];
  C *buf = reinterpret_cast<C *>(lds_raw);  // [L][M+1]
  const int M = rp.n;
  const int W = M + 1;
  const int64_t row0 = (int64_t)xcd_block(blockIdx.x, gridDim.x) * L;
  const int nl = (int)min((int64_t)L, nrows - row0);
  // one row element is one thread's column (M == FFT_THREADS, L <= MAXE
  // rows): element k = threadIdx.x of every row as one 8- or 16-byte load,
  // k = M by thread 0; no index division
  const bool colwise = M == FFT_THREADS && L <= MAXE;
  if (colwise) {
    const int k = threadIdx.x;
    C v[MAXE];
#pragma unroll
    for (int l = 0; l < MAXE; l++) {
      v[l] = C{0, 0};
      if (l < nl) v[l] = reinterpret_cast<const C *>(field + (row0 + l) * rowlen)[k];
    }
#pragma unroll
    for (int l = 0; l < MAXE; l++) {
      if (l < L) {
        if (k == 0) v[l].y = (T)0;
        buf[l * W + k] = v[l];
      }
    }
    if (k == 0)
      for (int l = 0; l < L; l++)
        buf[l * W + M] = l < nl ? C{field[(row0 + l) * rowlen + 2 * M], (T)0} : C{0, 0};
  } else {
    for (int i = threadIdx.x; i < L * W; i += FFT_THREADS) {
      const int l = (int)fdM1.div((uint32_t)i), k = i - l * W;
      C v = C{0, 0};
      if (l < nl) {
        const T *row = field + (row0 + l) * rowlen;
        v = C{row[2 * k], (k == 0 || k == M) ? (T)0 : row[2 * k + 1]};
      }
      buf[i] = v;
    }
  }
  __syncthreads();
  // Z_k = (X_k + conj X_{M-k}) + i w^-k (X_k - conj X_{M-k}), k < M; staged in registers
  C z[MAXE];
#pragma unroll
  for (int q = 0; q < MAXE; q++) {
    const int i = threadIdx.x + q * FFT_THREADS;
    if (i < L * M) {
      const int l = colwise ? q : (int)fdM.div((uint32_t)i), k = colwise ? (int)threadIdx.x : i - l * M;
      const C xk = buf[l * W + k];
      const C xc = cconj(buf[l * W + (M - k)]);
      const C s = cadd(xk, xc);
      C dd = csub(xk, xc);
      const C t = twd[k];
      dd = cmul(dd, C{t.x, -t.y});                 // times exp(+2 pi i k / d)
      z[q] = C{s.x - dd.y, s.y + dd.x};            // s + i*dd
    }
  }
  __syncthreads();
  // the transform runs on padded rows (lds_ix), as the forward z pass
#pragma unroll
  for (int q = 0; q < MAXE; q++) {
    const int i = threadIdx.x + q * FFT_THREADS;
    if (i < L * M) buf[lds_ix<true>(i)] = z[q];
  }
  C *stw = buf + c2r_span(L, M);
  for (int i = threadIdx.x; i < M; i += FFT_THREADS) stw[i] = twM[i];
  __syncthreads();
  lds_fft<+1, false, MAXE, C, true>(buf, rp, L, 0, 1, M, stw);
  if (colwise) {
    const int m = threadIdx.x;
#pragma unroll
    for (int l = 0; l < MAXE; l++) {
      if (l < nl) {
        C *row = reinterpret_cast<C *>(field + (row0 + l) * rowlen);
        row[m] = buf[lds_ix<true>(l * M + m)];
        if (m == 0) row[M] = C{0, 0};
      }
    }
    return;
  }
  for (int i = threadIdx.x; i < nl * M; i += FFT_THREADS) {
    const int l = (int)fdM.div((uint32_t)i), m = i - l * M;
    const C v = buf[lds_ix<true>(i)];
    T *row = field + (row0 + l) * rowlen;
    row[2 * m] = v.x;
    row[2 * m + 1] = v.y;
    if (m == 0) {
      row[2 * M] = 0;
      row[2 * M + 1] = 0;
    }
  }
}

// c2c along y (axis = 1) or x (axis = 0) over column blocks of B = 1 << lgb
// consecutive kz values.  CT = 1024: the plan is n = 1024 = 8 * 8 * 8 * 2,
// run with compile-time stages; 0: any plan.
template <typename T, int SIGN, int CT = 0>
__global__ void __launch_bounds__(COL_THREADS) k_cols(typename Cx<T>::C *__restrict__ f, int64_t d, int64_t hz,
                                                     int lgb, int nkb, int axis, RadixPlan rp,
                                                     const typename Cx<T>::C *__restrict__ tw) {
  using C = typename Cx<T>::C;
  constexpr int MAXE = FftCfg<T>::COLE;
  extern __shared__ unsigned char lds_raw[];
  C *buf = reinterpret_cast<C *>(lds_raw);  // [n][B]
  const int n = CT ? CT : rp.n;
  const int B = 1 << lgb;
  const uint32_t blk = xcd_block(blockIdx.x, gridDim.x);
  const int64_t outer = blk / nkb;
  const int kb = (int)(blk - outer * nkb);
  const int64_t kz0 = (int64_t)kb * B;
  const int nb = (int)min((int64_t)B, hz - kz0);
  const int64_t estride = axis == 1 ? hz : d * hz;
  C *base = f + (axis == 1 ? outer * d * hz : outer * hz) + kz0;
  const int tot = n * B;
  C v[MAXE];
#pragma unroll
  for (int q = 0; q < MAXE; q++) {
    const int i = threadIdx.x + q * COL_THREADS;
    const int e = i >> lgb, b = i & (B - 1);
    v[q] = (i < tot && b < nb) ? base[e * estride + b] : C{0, 0};
  }
#pragma unroll
  for (int q = 0; q < MAXE; q++) {
    const int i = threadIdx.x + q * COL_THREADS;
    if (i < tot) buf[i] = v[q];
  }
  C *stw = buf + tot;  // twiddles staged in LDS
  for (int i = threadIdx.x; i < n; i += COL_THREADS) stw[i] = tw[i];
  __syncthreads();
  if constexpr (CT == 1024)
    lds_fft_ct<1024, 1, SIGN, true, MAXE, false, C, false, 8, 8, 8, 2>(buf, buf, B, lgb, stw);
  else
    lds_fft<SIGN, true, MAXE>(buf, rp, B, lgb, B, 1, stw);
#pragma unroll
  for (int q = 0; q < MAXE; q++) {
    const int i = threadIdx.x + q * COL_THREADS;
    const int e = i >> lgb, b = i & (B - 1);
    if (i < tot && b < nb) base[e * estride + b] = buf[i];
  }
}

// k_cols for n * B = COL_THREADS * COLE and a first radix of COLE: every
// thread loads exactly the inputs of one first-stage butterfly (elements
// j + q * n/R of column b, j = tid >> lgb, b = tid & (B-1): the same
// addresses as k_cols' element-wise load), so the first stage runs from
// registers and the load's LDS round trip and two barriers are saved.
// L2 (last radix 2): the last stage's butterflies (rows k and k + n/2 of a
// column, twiddle w^k) write their outputs straight to global memory.
// CT = 1024: compile-time stages for n = 1024 as 8, 8, 8, 2 (FP32, with L2)
// or 4, 8, 8, 4 (FP64 without L2).
template <typename T, int SIGN, bool L2, int CT = 0>
__global__ void __launch_bounds__(COL_THREADS) k_cols_r1(typename Cx<T>::C *__restrict__ f, int64_t d, int64_t hz,
                                                        int lgb, int nkb, int axis, RadixPlan rp,
                                                        const typename Cx<T>::C *__restrict__ tw) {
  using C = typename Cx<T>::C;
  constexpr int R = FftCfg<T>::COLE;
  extern __shared__ unsigned char lds_raw[];
  C *buf = reinterpret_cast<C *>(lds_raw);  // [n][B]
  const int n = CT ? CT : rp.n;
  const int B = 1 << lgb;
  const uint32_t blk = xcd_block(blockIdx.x, gridDim.x);
  const int64_t outer = blk / nkb;
  const int64_t kz0 = (int64_t)(blk - outer * nkb) * B;
  const int nb = (int)min((int64_t)B, hz - kz0);
  const int64_t estride = axis == 1 ? hz : d * hz;
  C *base = f + (axis == 1 ? outer * d * hz : outer * hz) + kz0;
  const int b = threadIdx.x & (B - 1), j = threadIdx.x >> lgb, nq = n / R;
  C v[R];
#pragma unroll
  for (int q = 0; q < R; q++) v[q] = b < nb ? base[(int64_t)(j + q * nq) * estride + b] : C{0, 0};
  C *stw = buf + (int64_t)n * B;  // twiddles staged in LDS
  for (int i = threadIdx.x; i < n; i += COL_THREADS) stw[i] = tw[i];
  fft_first_regs<R, SIGN>(v, buf, B, lgb, B, 1);
  __syncthreads();
  if constexpr (CT == 1024 && L2 && R == 8)
    lds_fft_ct<1024, 8, SIGN, true, R, false, C, false, 8, 8>(buf, buf, B, lgb, stw);
  else if constexpr (CT == 1024 && !L2 && R == 4)
    lds_fft_ct<1024, 4, SIGN, true, R, false, C, false, 8, 8, 4>(buf, buf, B, lgb, stw);
  else
    for (int s = 1; s < rp.nst - (L2 ? 1 : 0); s++)
      fft_stage_r<SIGN, true, R, false, C, false>(rp.st[s].R, buf, buf, B, lgb, B, 1, rp.st[s], stw);
  if constexpr (L2) {
    const int h = n / 2;
#pragma unroll
    for (int i = 0; i < R / 2; i++) {
      const int k = j + i * (COL_THREADS >> lgb);  // butterfly b + k * B of B * n/2
      C t = stw[k];
      if (SIGN > 0) t.y = -t.y;
      C w[2] = {buf[k * B + b], cmul(buf[(k + h) * B + b], t)};
      Dft<2, SIGN, C>::run(w);
      if (b < nb) {
        base[(int64_t)k * estride + b] = w[0];
        base[(int64_t)(k + h) * estride + b] = w[1];
      }
    }
  } else {
#pragma unroll
    for (int q = 0; q < R; q++) {
      const int e = (threadIdx.x + q * COL_THREADS) >> lgb;
      if (b < nb) base[(int64_t)e * estride + b] = buf[threadIdx.x + q * COL_THREADS];
    }
  }
}

template <typename T>
int rows_per_block(int64_t M) {
  return (int)std::max<int64_t>(1, FftCfg<T>::ELEMS / std::max<int64_t>(M + 1, 1));
}

}  // namespace

int col_block(int prec, int64_t d, int64_t hz) {
  const int elems = prec == 4 ? FftCfg<float>::ELEMS : FftCfg<double>::ELEMS;
  int B = 1;
  while (B * 2 * d <= elems && B * 2 <= 64) B *= 2;
  if (B > hz) {
    B = 1;
    while (B < hz) B *= 2;
  }
  return B;
}

void fft_tables_init(Plan &p) {
  const int64_t d = p.d, M = d / 2;
  std::vector<double> td(2 * d), tm(2 * std::max<int64_t>(M, 1));
  std::vector<float> fd(2 * d), fm(2 * std::max<int64_t>(M, 1));
  const long double two_pi = 6.283185307179586476925286766559005768L;
  for (int64_t m = 0; m < d; m++) {
    const long double a = two_pi * (long double)m / (long double)d;
    td[2 * m] = (double)cosl(a);
    td[2 * m + 1] = (double)(-sinl(a));
    fd[2 * m] = (float)cosl(a);
    fd[2 * m + 1] = (float)(-sinl(a));
  }
  for (int64_t m = 0; m < M; m++) {
    const long double a = two_pi * (long double)m / (long double)M;
    tm[2 * m] = (double)cosl(a);
    tm[2 * m + 1] = (double)(-sinl(a));
    fm[2 * m] = (float)cosl(a);
    fm[2 * m + 1] = (float)(-sinl(a));
  }
  auto up = [&](DevBuf &b, const void *src, size_t bytes) {
    b.ensure(bytes);
    GENPK_HIP(hipMemcpy(b.ptr, src, bytes, hipMemcpyHostToDevice));
  };
  up(p.fft.tw_d_f64, td.data(), td.size() * 8);
  up(p.fft.tw_d_f32, fd.data(), fd.size() * 4);
  up(p.fft.tw_h_f64, tm.data(), tm.size() * 8);
  up(p.fft.tw_h_f32, fm.data(), fm.size() * 4);
}

template <typename T>
static void r2c_z_impl(Plan &p, bool from_tiles) {
  using C = typename Cx<T>::C;
  const int64_t M = p.d / 2;
  const int L = rows_per_block<T>(M);
  const int64_t nrows = p.d * p.d;
  const RadixPlan &rp = p.rp_h;
  const C *twM = (p.prec == 4) ? (const C *)p.fft.tw_h_f32.ptr : (const C *)p.fft.tw_h_f64.ptr;
  const C *twd = (p.prec == 4) ? (const C *)p.fft.tw_d_f32.ptr : (const C *)p.fft.tw_d_f64.ptr;
  const FoldGeom fg = fold_geom(p);
  const dim3 grid((unsigned)ceil_div(nrows, L));
  const size_t lds = ((size_t)lds_len<true>((int64_t)L * M) + 2 * M + 1) * sizeof(C);
  // the d = 1024 plan of radices 8, 8, 8 runs compile-time stages
  const bool ct512 = M == 512 && rp.nst == 3 && rp.st[0].R == 8 && rp.st[1].R == 8 && rp.st[2].R == 8;
  // FP64 at d = 1024 (seven rows a workgroup) reads the rows element-wise:
  // the butterfly loader measured 3.50 -> 3.73 ms there (0.52 -> 0.47 ms at 512)
  const bool bfly = rp.st[0].R == 8 && rp.nst >= 2 && (int64_t)L * (M / 8) <= 2 * FFT_THREADS &&
                    (sizeof(T) == 4 || !ct512);
  const bool std_tiles = p.d == 1024 && fg.TX == 16 && fg.TY == 32 && fg.TZ == 32;
  if (from_tiles && bfly && ct512 && std_tiles) {
    Stage st(p, "fold_fft_z");
    hipLaunchKernelGGL((k_r2c_z<T, true, true, 512, true>), grid, dim3(FFT_THREADS), lds, p.stream, p.field.as<T>(),
                       p.tiles.as<const T>(), fg, FastDiv((uint32_t)p.d), p.rowlen, nrows, L, FastDiv((uint32_t)M),
                       FastDiv((uint32_t)(M + 1)), rp, twM, twd);
  } else if (from_tiles && bfly && ct512) {
    Stage st(p, "fold_fft_z");
    hipLaunchKernelGGL((k_r2c_z<T, true, true, 512>), grid, dim3(FFT_THREADS), lds, p.stream, p.field.as<T>(),
                       p.tiles.as<const T>(), fg, FastDiv((uint32_t)p.d), p.rowlen, nrows, L, FastDiv((uint32_t)M),
                       FastDiv((uint32_t)(M + 1)), rp, twM, twd);
  } else if (from_tiles && bfly) {
    Stage st(p, "fold_fft_z");
    hipLaunchKernelGGL((k_r2c_z<T, true, true>), grid, dim3(FFT_THREADS), lds, p.stream, p.field.as<T>(),
                       p.tiles.as<const T>(), fg, FastDiv((uint32_t)p.d), p.rowlen, nrows, L, FastDiv((uint32_t)M),
                       FastDiv((uint32_t)(M + 1)), rp, twM, twd);
  } else if (from_tiles && ct512 && std_tiles) {
    Stage st(p, "fold_fft_z");
    hipLaunchKernelGGL((k_r2c_z<T, true, false, 512, true>), grid, dim3(FFT_THREADS), lds, p.stream, p.field.as<T>(),
                       p.tiles.as<const T>(), fg, FastDiv((uint32_t)p.d), p.rowlen, nrows, L, FastDiv((uint32_t)M),
                       FastDiv((uint32_t)(M + 1)), rp, twM, twd);
  } else if (from_tiles && ct512) {
    Stage st(p, "fold_fft_z");
    hipLaunchKernelGGL((k_r2c_z<T, true, false, 512>), grid, dim3(FFT_THREADS), lds, p.stream, p.field.as<T>(),
                       p.tiles.as<const T>(), fg, FastDiv((uint32_t)p.d), p.rowlen, nrows, L, FastDiv((uint32_t)M),
                       FastDiv((uint32_t)(M + 1)), rp, twM, twd);
  } else if (from_tiles) {
    Stage st(p, "fold_fft_z");
    hipLaunchKernelGGL((k_r2c_z<T, true>), grid, dim3(FFT_THREADS), lds, p.stream, p.field.as<T>(),
                       p.tiles.as<const T>(), fg, FastDiv((uint32_t)p.d), p.rowlen, nrows, L, FastDiv((uint32_t)M),
                       FastDiv((uint32_t)(M + 1)), rp, twM, twd);
  } else {
    Stage st(p, "fft_z");
    hipLaunchKernelGGL((k_r2c_z<T, false>), grid, dim3(FFT_THREADS), lds, p.stream, p.field.as<T>(),
                       (const T *)nullptr, fg, FastDiv((uint32_t)p.d), p.rowlen, nrows, L, FastDiv((uint32_t)M),
                       FastDiv((uint32_t)(M + 1)), rp, twM, twd);
  }
  GENPK_LAUNCH_CHECK();
}

template <typename T>
static void c2r_z_impl(Plan &p) {
  using C = typename Cx<T>::C;
  const int64_t M = p.d / 2;
  const int L = rows_per_block<T>(M);
  const int64_t nrows = p.d * p.d;
  const RadixPlan &rp = p.rp_h;
  const C *twM = (p.prec == 4) ? (const C *)p.fft.tw_h_f32.ptr : (const C *)p.fft.tw_h_f64.ptr;
  const C *twd = (p.prec == 4) ? (const C *)p.fft.tw_d_f32.ptr : (const C *)p.fft.tw_d_f64.ptr;
  Stage st(p, "ifft_z");
  hipLaunchKernelGGL(k_c2r_z<T>, dim3((unsigned)ceil_div(nrows, L)), dim3(FFT_THREADS),
                     ((size_t)c2r_span(L, M) + M) * sizeof(C), p.stream, p.field.as<T>(), p.rowlen, nrows, L,
                     FastDiv((uint32_t)M), FastDiv((uint32_t)(M + 1)), rp, twM, twd);
  GENPK_LAUNCH_CHECK();
}

template <typename T, int SIGN>
static void cols_impl(Plan &p, int axis) {
  using C = typename Cx<T>::C;
  const int B = col_block(p.prec, p.d, p.hz);
  const int nkb = (int)ceil_div(p.hz, B);
  const RadixPlan &rp = p.rp_d;
  const C *tw = (p.prec == 4) ? (const C *)p.fft.tw_d_f32.ptr : (const C *)p.fft.tw_d_f64.ptr;
  const size_t lds = (size_t)p.d * (B + 1) * sizeof(C);
  const int64_t nblk = p.d * nkb;
  Stage st(p, axis == 1 ? (SIGN < 0 ? "fft_y" : "ifft_y") : (SIGN < 0 ? "fft_x" : "ifft_x"));
  const bool ct1024 = is_plan_8882(rp);
  if (rp.st[0].R == FftCfg<T>::COLE && p.d * B == (int64_t)COL_THREADS * FftCfg<T>::COLE) {
    if (rp.st[rp.nst - 1].R == 2 && ct1024)
      hipLaunchKernelGGL((k_cols_r1<T, SIGN, true, 1024>), dim3((unsigned)nblk), dim3(COL_THREADS), lds, p.stream,
                         p.field.as<C>(), p.d, p.hz, ilog2(B), nkb, axis, rp, tw);
    else if (rp.st[rp.nst - 1].R == 2)
      hipLaunchKernelGGL((k_cols_r1<T, SIGN, true>), dim3((unsigned)nblk), dim3(COL_THREADS), lds, p.stream,
                         p.field.as<C>(), p.d, p.hz, ilog2(B), nkb, axis, rp, tw);
    else if (FftCfg<T>::COLE == 4 && is_plan_4884(rp))
      hipLaunchKernelGGL((k_cols_r1<T, SIGN, false, FftCfg<T>::COLE == 4 ? 1024 : 0>), dim3((unsigned)nblk),
                         dim3(COL_THREADS), lds, p.stream, p.field.as<C>(), p.d, p.hz, ilog2(B), nkb, axis, rp, tw);
    else
      hipLaunchKernelGGL((k_cols_r1<T, SIGN, false>), dim3((unsigned)nblk), dim3(COL_THREADS), lds, p.stream,
                         p.field.as<C>(), p.d, p.hz, ilog2(B), nkb, axis, rp, tw);
  } else if (ct1024) {
    hipLaunchKernelGGL((k_cols<T, SIGN, 1024>), dim3((unsigned)nblk), dim3(COL_THREADS), lds, p.stream,
                       p.field.as<C>(), p.d, p.hz, ilog2(B), nkb, axis, rp, tw);
  } else {
    hipLaunchKernelGGL((k_cols<T, SIGN>), dim3((unsigned)nblk), dim3(COL_THREADS), lds, p.stream, p.field.as<C>(), p.d,
                       p.hz, ilog2(B), nkb, axis, rp, tw);
  }
  GENPK_LAUNCH_CHECK();
}

void r2c_zpass(Plan &p) {
  // a fresh deposit (no imported density) is folded while the z pass loads it
  const bool from_tiles = p.tiles_live && p.fstate != FieldState::Real;
  if (!from_tiles) tiles_to_field(p);
  GENPK_REQUIRE(from_tiles || p.fstate == FieldState::Real, "r2c needs a real-space density");
  if (p.prec == 4) r2c_z_impl<float>(p, from_tiles); else r2c_z_impl<double>(p, from_tiles);
  p.tiles_live = false;
  p.fstate = FieldState::Real;
}
void c2r_zpass(Plan &p) {
  if (p.prec == 4) c2r_z_impl<float>(p); else c2r_z_impl<double>(p);
}
void c2c_ypass(Plan &p, int sign) {
  if (p.prec == 4) {
    if (sign < 0) cols_impl<float, -1>(p, 1); else cols_impl<float, 1>(p, 1);
  } else {
    if (sign < 0) cols_impl<double, -1>(p, 1); else cols_impl<double, 1>(p, 1);
  }
}
void c2c_xpass(Plan &p, int sign) {
  if (p.prec == 4) {
    if (sign < 0) cols_impl<float, -1>(p, 0); else cols_impl<float, 1>(p, 0);
  } else {
    if (sign < 0) cols_impl<double, -1>(p, 0); else cols_impl<double, 1>(p, 0);
  }
}

void r2c(Plan &p) {
  r2c_zpass(p);
  c2c_ypass(p, -1);
  c2c_xpass(p, -1);
  p.fstate = FieldState::Spectrum;
}

void c2r(Plan &p) {
  GENPK_REQUIRE(p.fstate == FieldState::Spectrum, "c2r needs a spectrum");
  c2c_xpass(p, +1);
  c2c_ypass(p, +1);
  c2r_zpass(p);
  p.fstate = FieldState::Real;
}

}  // namespace genpk
