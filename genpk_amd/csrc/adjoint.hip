// Hand-written adjoint of the auto power spectrum (no counterpart in the
// reference, which has no gradients).
//
// Forward: P_b = sum_{k in H, bin(k) = b} h_k g_k |F_k|^2 / (M^2 C_b), with
// F = r2c(rho), H the half spectrum, h_k the Hermitian weight, g_k the CIC
// deconvolution factor and C_b the mode count.  For lambda_b = dL/dP_b:
//   A_k      = 2 lambda_{b(k)} g_k F_k / (M^2 C_{b(k)}),  A_0 = 0
//   dL/drho  = c2r(A)   (unnormalised inverse; interior kz implicitly doubled)
//   dL/dm_j  = sum_c w_jc G(c)
//   dL/dx_ja = m_j (d/box) sum_c (dw_jc/du_a) G(c)
// where w_jc are the CIC weights of fieldize.cpp:77-84 and G = dL/drho.
//
// dL/drho is formed from whichever forward state the plan holds: a kept
// spectrum (weight, then the three inverse passes), the field after the
// forward z and y passes (left there by the fused power), or a deposit or
// density (z and y passes first).  In the last two cases the x pass runs
// forward FFT -> weight -> inverse FFT on each column block in LDS, so the
// spectrum is never written.  The gather visits the particles in tile order
// (the deposit's partition), reads dL/drho from a halo'd tile staged in LDS
// and writes the gradients in tile order; the sort's permutations then take
// them back to the particles' order.
#include "binlookup.h"
#include "fft_core.h"
#include "tilesort.h"

#include <algorithm>
#include <cmath>

namespace genpk {

int col_block(int prec, int64_t d, int64_t hz);

namespace {

constexpr int ADJ_THREADS = 1024;  // as the transform's column passes
constexpr int GATHER_THREADS = 512;
constexpr int GATHER_RUN = 32;
template <typename T> struct AdjCfg;
template <> struct AdjCfg<float> { static constexpr int MAXE = 8; };
template <> struct AdjCfg<double> { static constexpr int MAXE = 4; };

__device__ __forceinline__ int64_t kv(int64_t n, int64_t d) { return n <= d / 2 ? n : n - d; }

struct AdjWeight {
  const int *q2bin;
  const int *count;
  const double *lambda;
  const float *wtab;
  double inv_m2;
  const int64_t *thr;  // bin thresholds [nrbins + 1] (smallest q whose bin is >= b)
  int nrbins;
  float lq;            // bin(q) ~ lq * log2(q)
  // 2 lambda_b g_k / (M^2 C_b) of the mode at array indices (x, y, kz); 0 at k = 0
  __device__ __forceinline__ double at(int64_t x, int64_t y, int64_t kz, int64_t d) const {
    const int64_t kx = kv(x, d), ky = kv(y, d);
    const int64_t q = kx * kx + ky * ky + kz * kz;
    if (q == 0) return 0.0;
    const int b = q2bin[q];
    const int c = count[b];
    const float gw = (wtab[x] * wtab[y]) * wtab[kz];
    double g = (double)gw * (double)gw;
    g *= g;
    return c ? 2.0 * lambda[b] * g * inv_m2 / (double)c : 0.0;
  }
};

template <typename T>
__global__ void k_weight(typename Cx<T>::C *__restrict__ f, int64_t d, int64_t hz, AdjWeight w) {
  using C = typename Cx<T>::C;
  const int64_t total = d * d * hz;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t kz = i % hz;
    const int64_t row = i / hz;
    const double s = w.at(row / d, row % d, kz, d);
    const C v = f[i];
    f[i] = C{(T)((double)v.x * s), (T)((double)v.y * s)};
  }
}

// x pass of the adjoint on the field after the forward z and y passes:
// per column block (fixed y, B consecutive kz), forward length-d FFT along
// x, multiply by the adjoint weight, inverse FFT, store back.  One block
// per workgroup, the stages in place in one LDS buffer (k_adjoint_x_pp
// covers the usual shapes).
template <typename T>
__global__ void __launch_bounds__(ADJ_THREADS) k_adjoint_x(typename Cx<T>::C *__restrict__ f, int64_t d, int64_t hz,
                                                          int lgb, int nkb, RadixPlan rp,
                                                          const typename Cx<T>::C *__restrict__ tw, AdjWeight w) {
  using C = typename Cx<T>::C;
  constexpr int MAXE = AdjCfg<T>::MAXE;
  extern __shared__ unsigned char lds_raw[];
  C *buf = reinterpret_cast<C *>(lds_raw);  // [n][B]
  const int n = rp.n;
  const int B = 1 << lgb;
  const uint32_t blk = xcd_block(blockIdx.x, gridDim.x);
  const int64_t y = blk / nkb;
  const int kb = (int)(blk - y * nkb);
  const int64_t kz0 = (int64_t)kb * B;
  const int nb = (int)min((int64_t)B, hz - kz0);
  const int64_t estride = d * hz;
  C *base = f + y * hz + kz0;
  const int tot = n * B;
#pragma unroll
  for (int q = 0; q < MAXE; q++) {
    const int i = threadIdx.x + q * ADJ_THREADS;
    const int e = i >> lgb, b = i & (B - 1);
    if (i < tot) buf[i] = b < nb ? base[e * estride + b] : C{0, 0};
  }
  C *stw = buf + tot;
  for (int i = threadIdx.x; i < n; i += ADJ_THREADS) stw[i] = tw[i];
  __syncthreads();
  lds_fft<-1, true, MAXE>(buf, rp, B, lgb, B, 1, stw);
  __syncthreads();
  for (int i = threadIdx.x; i < tot; i += ADJ_THREADS) {
    const int e = i >> lgb, b = i & (B - 1);
    if (b < nb) {
      const double s = w.at(e, y, kz0 + b, d);
      const C a = buf[i];
      buf[i] = C{(T)((double)a.x * s), (T)((double)a.y * s)};
    }
  }
  __syncthreads();
  lds_fft<+1, true, MAXE>(buf, rp, B, lgb, B, 1, stw);
#pragma unroll
  for (int q = 0; q < MAXE; q++) {
    const int i = threadIdx.x + q * ADJ_THREADS;
    const int e = i >> lgb, b = i & (B - 1);
    if (i < tot && b < nb) base[e * estride + b] = buf[i];
  }
}

// k_adjoint_x for n * B = ADJ_THREADS * MAXE and a first radix of MAXE,
// persistent over the column blocks (v, v + gridDim.x, ...): the first
// stage runs from the loaded registers, the stages alternate between two
// LDS buffers (one barrier each), and the next block's loads are in flight
// during this block's inverse transform and stores.
// LT (a last radix 2): the last forward stage runs in registers with the
// weight; the per-bin factors 2 lambda_b / (M^2 C_b), the window factors
// and the bin thresholds are staged in LDS, and each thread's modes find
// their bin by tracking the thresholds along kx, so a mode's weight costs
// no global read and no division.  (Compile-time stages for n = 1024 spill
// registers here: 5.2 -> 7.3 ms at 1024^3, against 4.3 ms without.)
template <typename T, bool LT>
__global__ void __launch_bounds__(ADJ_THREADS) k_adjoint_x_pp(typename Cx<T>::C *__restrict__ f, int64_t d,
                                                             int64_t hz, int lgb, int nkb, FastDiv fnkb, RadixPlan rp,
                                                             const typename Cx<T>::C *__restrict__ tw, AdjWeight w) {
  using C = typename Cx<T>::C;
  constexpr int MAXE = AdjCfg<T>::MAXE;
  extern __shared__ unsigned char lds_raw[];
  const int n = rp.n;
  const int B = 1 << lgb;
  const int tot = n * B;
  C *buf = reinterpret_cast<C *>(lds_raw);  // [n][B]
  C *buf2 = buf + tot;
  C *stw = buf2 + tot;
  for (int i = threadIdx.x; i < n; i += ADJ_THREADS) stw[i] = tw[i];
  double *sfb = reinterpret_cast<double *>(stw + n);  // [nrbins]
  float *swt = reinterpret_cast<float *>(sfb + w.nrbins);  // [d]
  int *sthr = reinterpret_cast<int *>(swt + d);  // [nrbins + 1]
  if constexpr (LT) {
    for (int i = threadIdx.x; i < w.nrbins; i += ADJ_THREADS) {
      const int c = w.count[i];
      sfb[i] = c ? 2.0 * w.lambda[i] * w.inv_m2 / (double)c : 0.0;
    }
    for (int i = threadIdx.x; i < d; i += ADJ_THREADS) swt[i] = w.wtab[i];
    for (int i = threadIdx.x; i <= w.nrbins; i += ADJ_THREADS) sthr[i] = (int)w.thr[i];
  }
  const int64_t estride = d * hz;
  const int64_t nblk = d * nkb;
  const int b = threadIdx.x & (B - 1);
  C v[MAXE];
  // column block vb's loads: the inputs of this thread's first-stage butterfly
  // (a grid that is a multiple of the XCD count keeps vb on this workgroup's XCD)
  auto load = [&](int64_t vb) {
    const uint32_t blk = xcd_block((uint32_t)vb, (uint32_t)nblk);
    const uint32_t yb = fnkb.div(blk);  // a multiply-high, not a division loop
    const int64_t y = yb;
    const int64_t kz0 = (int64_t)(blk - yb * (uint32_t)nkb) * B;
    const int nb = (int)min((int64_t)B, hz - kz0);
    const C *src = f + y * hz + kz0;
#pragma unroll
    for (int q = 0; q < MAXE; q++) {
      const int i = threadIdx.x + q * ADJ_THREADS;
      const int e = i >> lgb;
      v[q] = b < nb ? src[e * estride + b] : C{0, 0};
    }
  };
  if ((int64_t)blockIdx.x < nblk) load(blockIdx.x);
  for (int64_t vb = blockIdx.x; vb < nblk; vb += gridDim.x) {
    const uint32_t blk = xcd_block((uint32_t)vb, (uint32_t)nblk);
    const uint32_t yb = fnkb.div(blk);
    const int64_t y = yb;
    const int64_t kz0 = (int64_t)(blk - yb * (uint32_t)nkb) * B;
    const int nb = (int)min((int64_t)B, hz - kz0);
    C *base = f + y * hz + kz0;
    // the previous block's stores have read their buffer before this barrier
    fft_first_regs<MAXE, -1>(v, buf2, B, lgb, B, 1);
    __syncthreads();
    C *res;
    if (LT) {
      // the last forward stage (radix 2) and the adjoint weight in registers:
      // butterfly jj pairs stage inputs jj and jj + n/2 (the latter times
      // w^jj) into modes jj and jj + n/2, written weighted to the other buffer
      C *x = lds_fft_pp<-1, true, MAXE>(buf2, buf, rp, B, lgb, B, 1, stw, 1, rp.nst - 1);
      res = x == buf ? buf2 : buf;
      constexpr int HB = MAXE / 2;
      const int half = n / 2;
      const int jb = (threadIdx.x >> lgb) * HB;
      // modes jj (kx = jj) and jj + n/2 (kx = jj - n/2, kx = n/2 at jj = 0)
      // each keep a bin tracker along their run
      const int kz = (int)kz0 + b, ky = (int)kv(y, d);
      const int a = ky * ky + kz * kz;
      const float wy = swt[y], wz = swt[min(kz, (int)d - 1)];
      BinTrackLds t0, t1;
      auto lds_weight = [&](int xi, BinTrackLds &bt) -> double {
        const int kx = xi <= half ? xi : xi - n;
        const int qk = a + kx * kx;
        if (qk == 0) return 0.0;
        const float gw = (swt[xi] * wy) * wz;  // the forward's ((wx*wy)*wz) in float
        double g = (double)gw * (double)gw;
        g *= g;
        return g * sfb[bt.of(qk, sthr, w.nrbins, w.lq)];
      };
#pragma unroll
      for (int q = 0; q < HB; q++) {
        const int jj = jb + q;
        const C a0 = x[jj * B + b];
        const C v1 = cmul(x[(jj + half) * B + b], stw[jj]);
        C o0 = cadd(a0, v1), o1 = csub(a0, v1);
        if (b < nb) {
          const double s0 = lds_weight(jj, t0);
          const double s1 = lds_weight(jj + half, t1);
          o0 = C{(T)((double)o0.x * s0), (T)((double)o0.y * s0)};
          o1 = C{(T)((double)o1.x * s1), (T)((double)o1.y * s1)};
        }
        res[jj * B + b] = o0;
        res[(jj + half) * B + b] = o1;
      }
    } else {
      res = lds_fft_pp<-1, true, MAXE>(buf2, buf, rp, B, lgb, B, 1, stw, 1);
      for (int i = threadIdx.x; i < tot; i += ADJ_THREADS) {
        const int e = i >> lgb;
        if (b < nb) {
          const double s = w.at(e, y, kz0 + b, d);
          const C a = res[i];
          res[i] = C{(T)((double)a.x * s), (T)((double)a.y * s)};
        }
      }
    }
    __syncthreads();
    // the next block's loads (v is free since the first stage)
    if (vb + gridDim.x < nblk) load(vb + gridDim.x);
    res = lds_fft_pp<+1, true, MAXE>(res, res == buf ? buf2 : buf, rp, B, lgb, B, 1, stw);
#pragma unroll
    for (int q = 0; q < MAXE; q++) {
      const int i = threadIdx.x + q * ADJ_THREADS;
      const int e = i >> lgb;
      if (b < nb) base[e * estride + b] = res[i];
    }
  }
}

struct GatherParams {
  KeyGeom kg;
  int HY, HZ;
  int64_t cells;
  int64_t ks;
  int64_t max_items;  // entries of the item-to-tile table
  int64_t rowlen;
  double mass;
  FastDiv fhz, fhy;
  FastDiv fntz, fnty;  // tile index -> (tx, ty, tz)
};

// One workgroup per work item of the tile order: stage the halo'd tile of
// G = dL/drho in LDS, then the gradients of the item's records, written at
// the records' own (tile-order) positions as (dL/dx, dL/dy, dL/dz[, dL/dm]):
// coalesced.
template <typename T, int NM, int GM>
__global__ void __launch_bounds__(GATHER_THREADS) k_gather_tiles(const T *__restrict__ G, const float *__restrict__ rec,
                                                                 const uint32_t *__restrict__ tile_start,
                                                                 const uint32_t *__restrict__ item_start,
                                                                 const uint32_t *__restrict__ item_tile,
                                                                 GatherParams P, float *__restrict__ out) {
  constexpr int RW = 3 + NM;
  constexpr int RWO = 3 + GM;
  extern __shared__ unsigned char lds_raw[];
  T *g = reinterpret_cast<T *>(lds_raw);
  // items in runs of GATHER_RUN per XCD (workgroup id mod 8), the XCDs' runs
  // side by side: consecutive tiles along z share the lines at their z edges
  // in G, which one L2 then serves to both (the grid is a multiple of
  // 8 * GATHER_RUN; ids past the last item return)
  const uint32_t xcd = blockIdx.x % NUM_XCD, slot = blockIdx.x / NUM_XCD;
  const uint32_t item = ((slot / GATHER_RUN) * NUM_XCD + xcd) * GATHER_RUN + slot % GATHER_RUN;
  ItemRange ir;
  if (!item_range(item_tile, item_start, tile_start, P.ks, P.max_items, item, ir)) return;
  if (ir.pb >= ir.pe) return;
  const int64_t k = ir.k;
  const int ntz = P.kg.ntz, nty = P.kg.nty, d = P.kg.d;
  const uint32_t kxy = P.fntz.div((uint32_t)k), kx = P.fnty.div(kxy);
  const int tz = (int)((uint32_t)k - kxy * (uint32_t)ntz);
  const int ty = (int)(kxy - kx * (uint32_t)nty);
  const int tx = (int)kx;
  const int ox = tx << P.kg.sx, oy = ty << P.kg.sy, oz = tz << P.kg.sz;
  // four loads in flight per thread while staging, and four records while gathering
  constexpr int UNR = 4;
  for (int c0 = threadIdx.x; c0 < (int)P.cells; c0 += GATHER_THREADS * UNR) {
    T t[UNR];
#pragma unroll
    for (int u = 0; u < UNR; u++) {
      const int c = c0 + u * GATHER_THREADS;
      if (c < (int)P.cells) {
        const int row = (int)P.fhz.div((uint32_t)c);
        const int lz = c - row * P.HZ;
        const int lx = (int)P.fhy.div((uint32_t)row);
        const int ly = row - lx * P.HY;
        int x = ox + lx, y = oy + ly, z = oz + lz;
        if (x >= d) x -= d;
        if (y >= d) y -= d;
        if (z >= d) z -= d;
        t[u] = G[((int64_t)x * d + y) * P.rowlen + z];
      }
    }
#pragma unroll
    for (int u = 0; u < UNR; u++) {
      const int c = c0 + u * GATHER_THREADS;
      if (c < (int)P.cells) g[c] = t[u];
    }
  }
  __syncthreads();
  const int sy = P.HZ, sx = P.HY * P.HZ;
  const double units = P.kg.units;
  for (uint32_t i0 = ir.pb + threadIdx.x; i0 < ir.pe; i0 += GATHER_THREADS * UNR) {
    float rr[UNR][RW];
#pragma unroll
    for (int u = 0; u < UNR; u++) {
      const uint32_t i = i0 + u * GATHER_THREADS;
      if (i < ir.pe) {
#pragma unroll
        for (int q = 0; q < RW; q++) rr[u][q] = rec[(int64_t)i * RW + q];
      }
    }
#pragma unroll
    for (int u = 0; u < UNR; u++) {
      const uint32_t i = i0 + u * GATHER_THREADS;
      if (i >= ir.pe) break;
      const float *r = rr[u];
      double fx, fy, fz;
      const int lx = cell_of(r[0], units, d, fx) - ox;
      const int ly = cell_of(r[1], units, d, fy) - oy;
      const int lz = cell_of(r[2], units, d, fz) - oz;
      if ((unsigned)lx >= (unsigned)P.kg.TX || (unsigned)ly >= (unsigned)P.kg.TY || (unsigned)lz >= (unsigned)P.kg.TZ)
        continue;  // cannot happen for records of this tile; guards the LDS bounds
      const int at = lx * sx + ly * sy + lz;
      const double g000 = (double)g[at], g001 = (double)g[at + 1];
      const double g100 = (double)g[at + sx], g101 = (double)g[at + sx + 1];
      const double g010 = (double)g[at + sy], g011 = (double)g[at + sy + 1];
      const double g110 = (double)g[at + sx + sy], g111 = (double)g[at + sx + sy + 1];
      // corners (bx, by, bz); weights (1-f or f) per axis
      const double ex = 1.0 - fx, ey = 1.0 - fy, ez = 1.0 - fz;
      // interpolate along z first
      const double a00 = ez * g000 + fz * g001, a10 = ez * g100 + fz * g101;
      const double a01 = ez * g010 + fz * g011, a11 = ez * g110 + fz * g111;
      const double dgx = ey * (a10 - a00) + fy * (a11 - a01);
      const double dgy = ex * (a01 - a00) + fx * (a11 - a10);
      const double b00 = ex * ey, b10 = fx * ey, b01 = ex * fy, b11 = fx * fy;
      const double dgz = b00 * (g001 - g000) + b10 * (g101 - g100) + b01 * (g011 - g010) + b11 * (g111 - g110);
      const double gm = b00 * a00 + b10 * a10 + b01 * a01 + b11 * a11;
      const double m = NM ? (double)r[3] : P.mass;
      const double s = m * units;
      float *o = out + (int64_t)i * RWO;
      o[0] = (float)(dgx * s);
      o[1] = (float)(dgy * s);
      o[2] = (float)(dgz * s);
      if (GM) o[3] = (float)gm;
    }
  }
}

// Gradients back to the particles' input order: particle j sits at sorted
// position perm2[perm1[j]] (perm1 alone after a one-pass sort).  Reads of
// perm1 and writes are coalesced; the two lookups are random reads, which
// cost far less than scattered partial-line writes.  Each XCD (workgroup id
// mod 8) takes runs of UNP_RUN consecutive blocks, the eight XCDs' runs side
// by side: both partition passes are stable, so a run's particles of one
// bucket sit together in perm2 and its lines are fetched by one L2, while
// all XCDs stay within one window of the particles.  The grid is a multiple
// of 8 * UNP_RUN.
constexpr int UNP_RUN = 256;
template <int RWO, int TWO>
__global__ void k_unpermute(const float *__restrict__ rec, const uint32_t *__restrict__ perm1,
                            const uint32_t *__restrict__ perm2, int64_t n, float *__restrict__ gpos,
                            float *__restrict__ gmass) {
  const uint32_t x = blockIdx.x % NUM_XCD, slot = blockIdx.x / NUM_XCD;
  const int64_t lb = ((int64_t)(slot / UNP_RUN) * NUM_XCD + x) * UNP_RUN + slot % UNP_RUN;
  for (int64_t j = lb * blockDim.x + threadIdx.x; j < n; j += (int64_t)gridDim.x * blockDim.x) {
    int64_t at = perm1[j];
    if (TWO) at = perm2[at];
    const float *r = rec + at * RWO;
    gpos[3 * j + 0] = r[0];
    gpos[3 * j + 1] = r[1];
    gpos[3 * j + 2] = r[2];
    if (RWO == 4) gmass[j] = r[3];
  }
}

AdjWeight adj_weight(Plan &p, int nrbins, const double *lambda, double m) {
  const BinTables &bt = bin_tables(p, nrbins);
  AdjWeight w;
  w.q2bin = bt.q2bin.as<int>();
  w.count = bt.count.as<int>();
  w.lambda = lambda;
  w.wtab = p.wtab.as<float>();
  w.inv_m2 = 1.0 / (m * m);
  w.thr = bt.thr.as<int64_t>();
  w.nrbins = bt.nrbins;
  w.lq = bt.lq;
  return w;
}

template <typename T>
void adjoint_x_impl(Plan &p, const AdjWeight &w) {
  using C = typename Cx<T>::C;
  const int B = col_block(p.prec, p.d, p.hz);
  const int nkb = (int)ceil_div(p.hz, B);
  const RadixPlan &rp = p.rp_d;
  const C *tw = (p.prec == 4) ? (const C *)p.fft.tw_d_f32.ptr : (const C *)p.fft.tw_d_f64.ptr;
  GENPK_REQUIRE((int64_t)p.d * B <= (int64_t)AdjCfg<T>::MAXE * ADJ_THREADS, "adjoint x pass: column block too large");
  Stage st(p, "adjoint_x");
  const int64_t tot = p.d * B;
  const bool pp = rp.st[0].R == AdjCfg<T>::MAXE && tot == (int64_t)ADJ_THREADS * AdjCfg<T>::MAXE &&
                  (size_t)(2 * tot + p.d) * sizeof(C) <= 160 * 1024;
  const dim3 grid((unsigned)(p.d * nkb));
  // the LDS tables of the fused weight: per-bin factors, window factors, thresholds
  const size_t lds_lt = (size_t)(2 * tot + p.d) * sizeof(C) + (size_t)w.nrbins * 8 + (size_t)p.d * 4 +
                        (size_t)(w.nrbins + 1) * 4;
  const bool lt = pp && rp.st[rp.nst - 1].R == 2 && w.nrbins >= 1 && lds_lt <= 160 * 1024;
  if (pp) {
    // persistent: as many workgroups as stay resident, at most one per column block
    // (4.3 ms at 1024^3 against 5.2 ms for one workgroup per block)
    const void *fn = lt ? reinterpret_cast<const void *>(&k_adjoint_x_pp<T, true>)
                        : reinterpret_cast<const void *>(&k_adjoint_x_pp<T, false>);
    const size_t lds = lt ? lds_lt : (size_t)(2 * tot + p.d) * sizeof(C);
    int per_cu = 0;
    GENPK_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, ADJ_THREADS, lds));
    const dim3 g((unsigned)std::max<int64_t>(1, std::min<int64_t>((int64_t)std::max(per_cu, 1) * num_cus(), p.d * nkb)));
    if (lt)
      hipLaunchKernelGGL((k_adjoint_x_pp<T, true>), g, dim3(ADJ_THREADS), lds, p.stream, p.field.as<C>(), p.d, p.hz,
                         ilog2(B), nkb, FastDiv((uint32_t)nkb), rp, tw, w);
    else
      hipLaunchKernelGGL((k_adjoint_x_pp<T, false>), g, dim3(ADJ_THREADS), lds, p.stream, p.field.as<C>(), p.d, p.hz,
                         ilog2(B), nkb, FastDiv((uint32_t)nkb), rp, tw, w);
  } else {
    hipLaunchKernelGGL((k_adjoint_x<T>), grid, dim3(ADJ_THREADS), (size_t)(tot + p.d) * sizeof(C), p.stream,
                       p.field.as<C>(), p.d, p.hz, ilog2(B), nkb, rp, tw, w);
  }
  GENPK_LAUNCH_CHECK();
}

template <typename T, int NM, int GM>
void gather_impl(Plan &p, const KeyGeom &kg, const TileOrder &o, double mass, float *out) {
  const TileGeom &tg = p.tg;
  GatherParams P;
  P.kg = kg;
  P.HY = tg.HY;
  P.HZ = tg.HZ;
  P.cells = tg.cells;
  P.ks = o.ks;
  P.max_items = o.max_items;
  P.rowlen = p.rowlen;
  P.mass = mass;
  P.fhz = FastDiv((uint32_t)tg.HZ);
  P.fhy = FastDiv((uint32_t)tg.HY);
  P.fntz = FastDiv((uint32_t)tg.ntz);
  P.fnty = FastDiv((uint32_t)tg.nty);
  const int64_t grid = ceil_div(o.max_items, 8 * GATHER_RUN) * 8 * GATHER_RUN;
  hipLaunchKernelGGL((k_gather_tiles<T, NM, GM>), dim3((unsigned)grid), dim3(GATHER_THREADS),
                     (size_t)tg.cells * sizeof(T), p.stream, p.field.as<const T>(), o.rec, o.tile_start, o.item_start,
                     o.item_tile, P, out);
  GENPK_LAUNCH_CHECK();
}

}  // namespace

void adjoint_density(Plan &p, int nrbins, const double *lambda, double m) {
  GENPK_REQUIRE(m != 0, "total mass must be nonzero");
  GENPK_REQUIRE(lambda != nullptr, "null lambda");
  const AdjWeight w = adj_weight(p, nrbins, lambda, m);
  if (!p.tiles_live && p.fstate == FieldState::Spectrum) {
    // kept spectrum: weight every mode, then the inverse passes
    const int64_t total = p.d * p.d * p.hz;
    const unsigned grid = (unsigned)std::min<int64_t>(ceil_div(total, 256), 65536);
    {
      Stage st(p, "adjoint_weight");
      if (p.prec == 4)
        hipLaunchKernelGGL(k_weight<float>, dim3(grid), dim3(256), 0, p.stream, p.field.as<float2>(), p.d, p.hz, w);
      else
        hipLaunchKernelGGL(k_weight<double>, dim3(grid), dim3(256), 0, p.stream, p.field.as<double2>(), p.d, p.hz, w);
      GENPK_LAUNCH_CHECK();
    }
    c2r(p);
    return;
  }
  if (p.tiles_live || p.fstate == FieldState::Real) {
    // forward density (deposit or imported): forward z and y passes first
    r2c_zpass(p);
    c2c_ypass(p, -1);
    p.fstate = FieldState::YDone;
  }
  GENPK_REQUIRE(p.fstate == FieldState::YDone,
                "adjoint needs the forward spectrum, the forward density or the state left by the fused power");
  if (p.prec == 4) adjoint_x_impl<float>(p, w); else adjoint_x_impl<double>(p, w);
  c2c_ypass(p, +1);
  c2r_zpass(p);
  p.fstate = FieldState::Real;
}

void adjoint_gather(Plan &p, double boxsize, int64_t n, const float *pos, const float *masses, double mass,
                    float *grad_pos, float *grad_mass) {
  GENPK_REQUIRE(p.fstate == FieldState::Real && !p.tiles_live,
                "gather needs dL/drho in the plan (run the adjoint first)");
  GENPK_REQUIRE(boxsize > 0 && std::isfinite(boxsize), "boxsize must be positive and finite");
  GENPK_REQUIRE(n >= 0, "negative particle count");
  if (n == 0) return;
  GENPK_REQUIRE(pos && grad_pos, "null positions or gradient output");
  const int64_t cap = (int64_t)1 << 31;
  const RetainedOrder &r = p.order;
  const bool reuse = r.valid && r.pos == pos && r.masses == masses && r.n == n && r.boxsize == boxsize &&
                     r.to.has_perm;
  const KeyGeom kg = key_geom(p, boxsize);
  int64_t done = 0;
  do {
    const int64_t m = std::min(n - done, cap);
    const float *mp = masses ? masses + done : nullptr;
    TileOrder o;
    if (reuse) {
      o = r.to;
    } else {
      o = sort_by_tile(p, kg, m, pos + 3 * done, mp, true, false);
      p.order.valid = n <= cap;
      p.order.pos = pos;
      p.order.masses = masses;
      p.order.n = n;
      p.order.boxsize = boxsize;
      p.order.mass = mass;
      p.order.every_tile = false;
      p.order.to = o;
    }
    float *gp = grad_pos + 3 * done;
    float *gmp = grad_mass ? grad_mass + done : nullptr;
    // output records go to the record buffer that does not hold the sorted records
    float *out = (o.rec == p.recA.as<float>()) ? p.recB.as<float>() : p.recA.as<float>();
    const int gm = grad_mass ? 1 : 0;
    {
      Stage st(p, "adjoint_gather");
      if (p.prec == 4) {
        if (masses)
          gm ? gather_impl<float, 1, 1>(p, kg, o, mass, out) : gather_impl<float, 1, 0>(p, kg, o, mass, out);
        else
          gm ? gather_impl<float, 0, 1>(p, kg, o, mass, out) : gather_impl<float, 0, 0>(p, kg, o, mass, out);
      } else {
        if (masses)
          gm ? gather_impl<double, 1, 1>(p, kg, o, mass, out) : gather_impl<double, 1, 0>(p, kg, o, mass, out);
        else
          gm ? gather_impl<double, 0, 1>(p, kg, o, mass, out) : gather_impl<double, 0, 0>(p, kg, o, mass, out);
      }
    }
    {
      Stage st(p, "adjoint_unpermute");
      const unsigned grid = (unsigned)(ceil_div(std::min<int64_t>(ceil_div(m, 256), 1 << 22), 8 * UNP_RUN) * 8 * UNP_RUN);
      const bool two = o.perm2 != nullptr;
      if (gm && two)
        hipLaunchKernelGGL((k_unpermute<4, 1>), dim3(grid), dim3(256), 0, p.stream, out, o.perm1, o.perm2, m, gp, gmp);
      else if (gm)
        hipLaunchKernelGGL((k_unpermute<4, 0>), dim3(grid), dim3(256), 0, p.stream, out, o.perm1, o.perm2, m, gp, gmp);
      else if (two)
        hipLaunchKernelGGL((k_unpermute<3, 1>), dim3(grid), dim3(256), 0, p.stream, out, o.perm1, o.perm2, m, gp, gmp);
      else
        hipLaunchKernelGGL((k_unpermute<3, 0>), dim3(grid), dim3(256), 0, p.stream, out, o.perm1, o.perm2, m, gp, gmp);
      GENPK_LAUNCH_CHECK();
    }
    done += m;
  } while (done < n);
}

}  // namespace genpk
