// CIC-deconvolved, log-binned power (the reference's powerspectrum(),
// powerspectrum.c:35-110) on gfx950.
//
// * Bin index: the reference computes floor(binsperunit * log(|k|)) per mode
//   (powerspectrum.c:38,66,75,84).  It depends only on the integer
//   q = |k|^2, so the host evaluates that exact FP64 formula once per q and
//   the kernels look the bin up (no device log, identical rounding).
// * Deconvolution: invwindow() squared (fieldize.cpp:125-133,
//   powerspectrum.c:68,77,86) = ((float)(wx*wy)*wz)^4 with per-axis float
//   factors; the factor tables are built on the host with the reference's
//   double formula and rounded to float as it does.
// * Weights: 1 on the planes kz = 0 (DC skipped) and kz = d/2, 2 elsewhere
//   (powerspectrum.c:61-89).
// * Counts and effective k depend only on (d, nrbins): computed once per plan.
// * Power: FP64 accumulation.  Each thread sums a run of modes that share a
//   bin in registers and adds to per-workgroup LDS bins only at bin changes;
//   per-workgroup partials are summed in a fixed order and normalised as in
//   powerspectrum.c:102-108.
// The fused kernel performs the last (x) FFT pass and bins the modes while
// they are still in LDS, so the spectrum is never written back.
#include "binlookup.h"
#include "fft_core.h"
#include "plan.h"

#include <cmath>
#include <vector>

namespace genpk {

int col_block(int prec, int64_t d, int64_t hz);

namespace {

// 16 waves per workgroup with 8 staged values per thread (MAXE): few enough
// registers (~80) for 16 resident waves per CU, twice the waves of 256 threads
// x 32 values (x pass + bin 4.0 -> 3.4 ms at 1024^3).
constexpr int BIN_THREADS = 1024;
constexpr int LDS_LIMIT = 160 * 1024;
template <typename T> struct BinCfg;
template <> struct BinCfg<float> { static constexpr int ELEMS = 8192; static constexpr int MAXE = 8; };
template <> struct BinCfg<double> { static constexpr int ELEMS = 4096; static constexpr int MAXE = 4; };

__device__ __forceinline__ int64_t kval(int64_t n, int64_t d) { return n <= d / 2 ? n : n - d; }

__device__ __forceinline__ double deconv_of(const float *__restrict__ wtab, int64_t kx, int64_t ky, int64_t kz) {
  const float g = (wtab[kx] * wtab[ky]) * wtab[kz];
  double w = (double)g * (double)g;  // invwindow()
  return w * w;                      // its square, as applied per mode
}

// Mode count per |k|^2 = q over the half spectrum, weighted by the Hermitian
// weight h (1 on the planes kz = 0 and kz = d/2, else 2): with r2[a] the
// number of (x, y) array indices whose kx^2 + ky^2 = a,
//   qcnt[q] = sum_kz h(kz) r2[q - kz^2].
// Integer sums: the result is independent of scheduling.
__global__ void k_qcount(int64_t d, int64_t qmax, const uint32_t *__restrict__ r2, int64_t amax,
                         uint32_t *__restrict__ qcnt) {
  const int64_t hz = d / 2 + 1;
  for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q <= qmax; q += (int64_t)gridDim.x * blockDim.x) {
    uint32_t s = 0;
    for (int64_t kz = 0; kz < hz && kz * kz <= q; kz++) {
      const int64_t a = q - kz * kz;
      if (a > amax) continue;
      s += ((kz == 0 || kz == d / 2) ? 1u : 2u) * r2[a];
    }
    qcnt[q] = s;
  }
}

// Accumulates one spectrum contribution run-by-run in registers.
struct RunAcc {
  int cur = -1;
  double v[3] = {0, 0, 0};
};

template <int NS>
__device__ __forceinline__ void run_flush(RunAcc &r, double *bins, int nrbins) {
  if (r.cur >= 0) {
#pragma unroll
    for (int s = 0; s < NS; s++) atomicAdd(&bins[s * nrbins + r.cur], r.v[s]);
  }
}

template <int NS>
__device__ __forceinline__ void run_add(RunAcc &r, int b, const double *c, double *bins, int nrbins) {
  if (b != r.cur) {
    run_flush<NS>(r, bins, nrbins);
    r.cur = b;
#pragma unroll
    for (int s = 0; s < NS; s++) r.v[s] = 0;
  }
#pragma unroll
  for (int s = 0; s < NS; s++) r.v[s] += c[s];
}

template <typename C, int NF>
__device__ __forceinline__ void mode_terms(const C &fa, const C &fb, double h, double dc, double *c) {
  const double ax = (double)fa.x, ay = (double)fa.y;
  if (NF == 1) {
    c[0] = h * (ax * ax + ay * ay) * dc;
  } else {
    const double bx = (double)fb.x, by = (double)fb.y;
    c[0] = h * (ax * ax + ay * ay) * dc;
    c[1] = h * (bx * bx + by * by) * dc;
    c[2] = h * (ax * bx + ay * by) * dc;
  }
}

struct BinGeom {
  int64_t d, hz;
  int nrbins;
  const int *q2bin;
  const int64_t *thr;
  const float *wtab;
  float lq;  // bin(q) ~ lq * log2(q): binsperunit * ln(2) / 2
};

// Bin of q, tracked along a run of modes: the table is read only when q
// leaves the current bin's [thr[b], thr[b+1]) range.
struct BinTrack {
  int cur = -1;
  int64_t lo = 0, hi = 0;
  __device__ __forceinline__ int of(int64_t q, const BinGeom &g) {
    if (cur < 0 || q < lo || q >= hi) {
      cur = g.q2bin[q];
      lo = g.thr[cur];
      hi = g.thr[cur + 1];
    }
    return cur;
  }
};

// Fused final FFT pass along x + deconvolved binning.  NF fields (1: auto,
// 2: auto a, auto b and cross a.b*).  Persistent over column blocks.
//
// REG1 (one field, ping-pong buffers, n * B <= BIN_THREADS * MAXE, a first
// stage of radix MAXE and an even number of stages): the MAXE loaded
// elements of each thread below n * B / MAXE are the inputs of its butterfly
// in the first stage, which then runs from registers into buf2 (the threads
// above it load nothing and bin nothing).  The previous block no longer reads
// buf2 (an even number of stages leaves its result in buf), so no barrier
// comes before that stage.
//
// R2L (with REG1, a last stage of radix 2): the last stage runs in the
// binning, each thread pairing stage inputs j and j + n/2 into modes j and
// j + n/2 in registers; the two buffers then swap roles between blocks so
// that the next first stage never writes the buffer being binned.
template <typename T, int NF, bool REG1, bool R2L = false, int CT = 0>
__global__ void __launch_bounds__(BIN_THREADS) k_xpass_bin(typename Cx<T>::C *__restrict__ fa,
                                                            typename Cx<T>::C *__restrict__ fb, BinGeom g, int B,
                                                            int nkb, FastDiv fnkb, RadixPlan rp,
                                                            const typename Cx<T>::C *__restrict__ tw, int keep,
                                                            int pp, double *__restrict__ part) {
  using C = typename Cx<T>::C;
  constexpr int NS = NF == 1 ? 1 : 3;
  static_assert(!R2L || REG1, "the radix-2 last stage in the binning needs the register first stage");
  extern __shared__ unsigned char lds_raw[];
  double *bins = reinterpret_cast<double *>(lds_raw);
  C *buf = reinterpret_cast<C *>(bins + ((NS * g.nrbins + 1) & ~1));  // [NF][n][B], 16-byte aligned
  const int n = CT ? CT : rp.n;
  const int lgb = ilog2(B);
  // pp (one field): a second buffer lets the FFT stages alternate (one barrier each)
  C *buf2 = buf + (size_t)NF * n * B;
  C *stw = buf2 + (pp ? (size_t)n * B : 0);  // twiddles staged in LDS
  float *wt = reinterpret_cast<float *>(stw + n);  // deconvolution weights [d]
  int *thr32 = reinterpret_cast<int *>(wt + g.d);  // bin thresholds [nrbins]
  for (int i = threadIdx.x; i < n; i += BIN_THREADS) stw[i] = tw[i];
  for (int i = threadIdx.x; i < NS * g.nrbins; i += BIN_THREADS) bins[i] = 0;
  for (int i = threadIdx.x; i < g.d; i += BIN_THREADS) wt[i] = g.wtab[i];
  for (int i = threadIdx.x; i < g.nrbins; i += BIN_THREADS) thr32[i] = (int)g.thr[i];
  const int64_t d = g.d, hz = g.hz;
  const int64_t estride = d * hz;
  const int64_t nblk = d * nkb;
  // bin-phase mapping: thread -> column b, contiguous run of e
  const int tpc = BIN_THREADS / B;               // threads per column (B divides BIN_THREADS)
  const int my_b = threadIdx.x % B;
  const int per = (n + tpc - 1) / tpc;
  const int e0 = (threadIdx.x / B) * per;
  // register-staged loads of column block v: all its global reads in flight
  // together; with REG1 at a stride of the first stage's butterfly count
  constexpr int MAXE = BinCfg<T>::MAXE;
  const int tot = n * B;
  const int lst = REG1 ? tot / MAXE : BIN_THREADS;
  const bool loader = !REG1 || (int)threadIdx.x < lst;
  C va[MAXE], vb[NF == 2 ? MAXE : 1];
  auto load = [&](int64_t v) {
    // a grid that is a multiple of the XCD count keeps v on this workgroup's XCD
    const uint32_t blk = xcd_block((uint32_t)v, (uint32_t)nblk);
    const uint32_t yb = fnkb.div(blk);  // 32-bit: a 64-bit division costs ~130 scalar instructions
    const int64_t y = yb;
    const int64_t kz0 = (int64_t)(blk - yb * (uint32_t)nkb) * B;
    const int nb = (int)min((int64_t)B, hz - kz0);
    const int64_t off = y * hz + kz0;
#pragma unroll
    for (int q = 0; q < MAXE; q++) {
      const int i = threadIdx.x + q * lst;
      const int e = i >> lgb, b = i & (B - 1);
      const bool ok = loader && i < tot && b < nb;
      va[q] = ok ? fa[off + e * estride + b] : C{0, 0};
      if (NF == 2) vb[NF == 2 ? q : 0] = ok ? fb[off + e * estride + b] : C{0, 0};
    }
  };
  // block v + gridDim.x loads into the registers while block v is binned
  // (their live range misses the FFT stages, where the registers peak)
  if ((int64_t)blockIdx.x < nblk) load(blockIdx.x);
  for (int64_t v = blockIdx.x; v < nblk; v += gridDim.x) {
    const uint32_t blk = xcd_block((uint32_t)v, (uint32_t)nblk);
    const uint32_t yb = fnkb.div(blk);
    const int64_t y = yb;
    const int kb = (int)(blk - yb * (uint32_t)nkb);
    const int64_t kz0 = (int64_t)kb * B;
    const int nb = (int)min((int64_t)B, hz - kz0);
    const int64_t off = y * hz + kz0;
    C *res = buf;
    if constexpr (REG1) {
      if (loader) fft_first_regs<MAXE, -1>(va, buf2, B, lgb, B, 1);
      __syncthreads();
      if constexpr (CT == 1024 && R2L && MAXE == 8)
        res = lds_fft_ct<1024, 8, -1, true, MAXE, true, C, false, 8, 8>(buf2, buf, B, lgb, stw);
      else
        res = lds_fft_pp<-1, true, MAXE>(buf2, buf, rp, B, lgb, B, 1, stw, 1, R2L ? rp.nst - 1 : -1);
    } else {
      __syncthreads();  // the previous block's binning is done with buf
#pragma unroll
      for (int q = 0; q < MAXE; q++) {
        const int i = threadIdx.x + q * BIN_THREADS;
        if (i < tot) {
          buf[i] = va[q];
          if (NF == 2) buf[tot + i] = vb[NF == 2 ? q : 0];
        }
      }
      __syncthreads();
      if (NF == 1 && pp) {
        res = lds_fft_pp<-1, true, MAXE>(buf, buf2, rp, B, lgb, B, 1, stw);
      } else {
        lds_fft<-1, true, MAXE>(buf, rp, B, lgb, B, 1, stw);
        if (NF == 2) lds_fft<-1, true, MAXE>(buf + n * B, rp, B, lgb, B, 1, stw);
      }
    }
    if (keep) {
      for (int i = threadIdx.x; i < n * B; i += BIN_THREADS) {
        const int e = i >> lgb, b = i & (B - 1);
        if (b < nb) {
          fa[off + e * estride + b] = res[i];
          if (NF == 2) fb[off + e * estride + b] = res[n * B + i];
        }
      }
    }
    if (v + gridDim.x < nblk) load(v + gridDim.x);
    if (my_b < nb) {
      const int kz = (int)(kz0 + my_b);
      const int ky = (int)kval(y, d);
      const double h = (kz == 0 || kz == d / 2) ? 1.0 : 2.0;
      const int a = ky * ky + kz * kz;
      const float wy = wt[y], wz = wt[kz];
      RunAcc r;
      BinTrackLds bt;
      // this thread's run of at most MAXE modes along x (per <= MAXE as
      // n * B <= BIN_THREADS * MAXE); with one field all its LDS reads are
      // issued first (two fields would not fit the registers)
      if constexpr (R2L) {
        // the last stage: butterfly jj pairs stage inputs jj and jj + n/2
        // (the latter times w^jj) into modes jj and jj + n/2
        constexpr int HB = MAXE / 2;
        const int half = n / 2;
        const int jb = (threadIdx.x / B) * HB;  // below half exactly for the loaders
        C mo[MAXE];
        if (loader) {
#pragma unroll
        for (int q = 0; q < HB; q++) {
          const int jj = jb + q;
          const C a0 = res[jj * B + my_b];
          const C v1 = cmul(res[(jj + half) * B + my_b], stw[jj]);
          mo[q] = cadd(a0, v1);
          mo[q + HB] = csub(a0, v1);
        }
#pragma unroll
        for (int q = 0; q < MAXE; q++) {
          const int e = q < HB ? jb + q : jb + (q - HB) + half;
          const int kx = e <= d / 2 ? e : e - (int)d;
          const int qk = a + kx * kx;
          if (qk != 0) {
            const int bn = bt.of(qk, thr32, g.nrbins, g.lq);
            const float gw = (wt[e] * wy) * wz;  // the reference's ((wx*wy)*wz) in float
            double dc = (double)gw * (double)gw;
            dc *= dc;
            double c[3];
            mode_terms<C, NF>(mo[q], mo[q], h, dc, c);
            run_add<NS>(r, bn, c, bins, g.nrbins);
          }
        }
        }
        run_flush<NS>(r, bins, g.nrbins);
      } else {
      constexpr bool BATCH = NF == 1;
      C ma[BATCH ? MAXE : 1];
      float wx[BATCH ? MAXE : 1];
      if (BATCH) {
#pragma unroll
        for (int j = 0; j < MAXE; j++) {
          const int e = e0 + j;
          if (j < per && e < n) {
            ma[BATCH ? j : 0] = res[e * B + my_b];
            if (!REG1) wx[BATCH ? j : 0] = wt[e];
          }
        }
      }
#pragma unroll
      for (int j = 0; j < MAXE; j++) {
        const int e = e0 + j;
        if (j < per && e < n) {
          const int kx = e <= d / 2 ? e : e - (int)d;
          const int q = a + kx * kx;
          if (q != 0) {
            const int b = bt.of(q, thr32, g.nrbins, g.lq);
            const float gw = ((BATCH && !REG1 ? wx[BATCH ? j : 0] : wt[e]) * wy) * wz;  // the reference's ((wx*wy)*wz) in float
            double dc = (double)gw * (double)gw;
            dc *= dc;
            const C fa_m = BATCH ? ma[BATCH ? j : 0] : res[e * B + my_b];
            const C fb_m = NF == 2 ? res[n * B + e * B + my_b] : fa_m;
            double c[3];
            mode_terms<C, NF>(fa_m, fb_m, h, dc, c);
            run_add<NS>(r, b, c, bins, g.nrbins);
          }
        }
      }
      run_flush<NS>(r, bins, g.nrbins);
      }
    }
    if constexpr (R2L) {
      // the next first stage writes the buffer the binning did not read
      if (res == buf2) {
        C *t = buf;
        buf = buf2;
        buf2 = t;
      }
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < NS * g.nrbins; i += BIN_THREADS) part[(int64_t)blockIdx.x * NS * g.nrbins + i] = bins[i];
}

// Binning of stored spectra [d][d][hz] (no FFT): one wave per (kx, ky) row,
// each lane a contiguous run of kz.
template <typename T, int NF>
__global__ void __launch_bounds__(BIN_THREADS) k_bin_stored(const typename Cx<T>::C *__restrict__ fa,
                                                             const typename Cx<T>::C *__restrict__ fb, BinGeom g,
                                                             double *__restrict__ part) {
  using C = typename Cx<T>::C;
  constexpr int NS = NF == 1 ? 1 : 3;
  extern __shared__ unsigned char lds_raw[];
  double *bins = reinterpret_cast<double *>(lds_raw);
  for (int i = threadIdx.x; i < NS * g.nrbins; i += BIN_THREADS) bins[i] = 0;
  __syncthreads();
  const int64_t d = g.d, hz = g.hz;
  const int lane = threadIdx.x & 63;
  const int64_t wave = (int64_t)blockIdx.x * (BIN_THREADS / 64) + (threadIdx.x >> 6);
  const int64_t nwaves = (int64_t)gridDim.x * (BIN_THREADS / 64);
  const int64_t per = (hz + 63) / 64;
  for (int64_t row = wave; row < d * d; row += nwaves) {
    const int64_t x = row / d, y = row % d;
    const int64_t kx = kval(x, d), ky = kval(y, d);
    const int64_t a = kx * kx + ky * ky;
    RunAcc r;
    BinTrack bt;
    const int64_t z0 = lane * per, z1 = min(hz, z0 + per);
    for (int64_t kz = z0; kz < z1; kz++) {
      const int64_t q = a + kz * kz;
      if (q == 0) continue;
      const int b = bt.of(q, g);
      const double h = (kz == 0 || kz == d / 2) ? 1.0 : 2.0;
      const double dc = deconv_of(g.wtab, x, y, kz);
      const C va = fa[row * hz + kz];
      const C vb = NF == 2 ? fb[row * hz + kz] : va;
      double c[3];
      mode_terms<C, NF>(va, vb, h, dc, c);
      run_add<NS>(r, b, c, bins, g.nrbins);
    }
    run_flush<NS>(r, bins, g.nrbins);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < NS * g.nrbins; i += BIN_THREADS) part[(int64_t)blockIdx.x * NS * g.nrbins + i] = bins[i];
}

// Sum per-workgroup partials (fixed order) and normalise (powerspectrum.c:102-108).
__global__ void k_bin_finish(int nrbins, int ns, int G, const double *__restrict__ part,
                             const int *__restrict__ count, const double *__restrict__ keffsum, double m_aa,
                             double m_bb, double m_ab, double *__restrict__ p_aa, double *__restrict__ p_bb,
                             double *__restrict__ p_ab, int *__restrict__ count_out, double *__restrict__ keffs_out) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nrbins) return;
  // eight partials in flight per thread; each sum adds in the order gi = 0, 1, ...
  double s[3] = {0, 0, 0};
  const int64_t gs = (int64_t)ns * nrbins;
#pragma unroll
  for (int k = 0; k < 3; k++) {
    if (k >= ns) break;
    const double *pk = part + (int64_t)k * nrbins + b;
    double acc = 0;
    int gi = 0;
    for (; gi + 8 <= G; gi += 8) {
      double t[8];
#pragma unroll
      for (int q = 0; q < 8; q++) t[q] = pk[(int64_t)(gi + q) * gs];
#pragma unroll
      for (int q = 0; q < 8; q++) acc += t[q];
    }
    for (; gi < G; gi++) acc += pk[(int64_t)gi * gs];
    s[k] = acc;
  }
  const int c = count[b];
  double *outs[3] = {p_aa, p_bb, p_ab};
  const double ms[3] = {m_aa, m_bb, m_ab};
  for (int k = 0; k < ns; k++) {
    if (!outs[k]) continue;
    double v = 0;
    if (c) {
      v = s[k] / ms[k];
      v /= c;
    }
    outs[k][b] = v;
  }
  if (count_out) count_out[b] = c;
  if (keffs_out) keffs_out[b] = c ? keffsum[b] / c : 0.0;
}

}  // namespace

int num_cus() {
  static int cus = 0;
  if (!cus) {
    int dev = 0;
    GENPK_HIP(hipGetDevice(&dev));
    GENPK_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  }
  return cus;
}

// float(onedinvwindow(KVAL(k), d)) for k < d, the factor of fieldize.cpp:117-121
static double onedinvwindow_host(int64_t kx, int64_t n) {
  return kx ? M_PI * kx / (n * sin(M_PI * kx / (float)n)) : 1.0;
}

BinTables &bin_tables(Plan &p, int nrbins) {
  GENPK_REQUIRE(nrbins >= 1 && nrbins <= 8192, "nrbins must be in [1, 8192]");
  auto it = p.bins.find(nrbins);
  if (it != p.bins.end() && it->second->counts_ready) return *it->second;
  auto bt = std::make_unique<BinTables>();
  bt->nrbins = nrbins;
  const int64_t d = p.d, h = d / 2;
  const int64_t qmax = 3 * h * h;
  std::vector<int> tab((size_t)qmax + 1);
  // binsperunit and the per-mode index exactly as powerspectrum.c:38,66
  const double binsperunit = (nrbins - 1) / log(sqrt(3) * d / 2.0);
  bt->lq = (float)(binsperunit * log(2.0) / 2);
  tab[0] = -1;
  for (int64_t q = 1; q <= qmax; q++) {
    const double kk = sqrt((double)q);
    const int64_t b = (int64_t)floor(binsperunit * log(kk));
    GENPK_REQUIRE(b >= 0 && b < nrbins, "a mode falls outside [0, nrbins): nrbins too small for dims");
    tab[q] = (int)b;
  }
  bt->q2bin.ensure(tab.size() * 4);
  GENPK_HIP(hipMemcpy(bt->q2bin.ptr, tab.data(), tab.size() * 4, hipMemcpyHostToDevice));
  // thr[b] = smallest q >= 1 with bin(q) >= b (bins are non-decreasing in q)
  std::vector<int64_t> thr((size_t)nrbins + 1, qmax + 1);
  {
    int64_t q = 1;
    for (int b = 0; b <= nrbins; b++) {
      while (q <= qmax && tab[q] < b) q++;
      thr[b] = q;
    }
    for (int64_t q2 = 2; q2 <= qmax; q2++)
      GENPK_REQUIRE(tab[q2] >= tab[q2 - 1], "bin index not monotonic in |k|^2");
  }
  bt->thr.ensure(thr.size() * 8);
  GENPK_HIP(hipMemcpy(bt->thr.ptr, thr.data(), thr.size() * 8, hipMemcpyHostToDevice));
  if (!p.wtab.ptr) {
    std::vector<float> w(d);
    for (int64_t k = 0; k < d; k++) {
      const int64_t kv = k <= d / 2 ? k : k - d;
      w[k] = (float)onedinvwindow_host(kv, d);
    }
    p.wtab.ensure(d * 4);
    GENPK_HIP(hipMemcpy(p.wtab.ptr, w.data(), d * 4, hipMemcpyHostToDevice));
  }
  // counts and keffs, once: per-q mode counts on the device, then per-bin
  // sums on the host in ascending q (deterministic)
  {
    const int64_t amax = 2 * h * h;
    std::vector<uint32_t> r2((size_t)amax + 1, 0);
    for (int64_t x = 0; x < d; x++) {
      const int64_t kx = x <= h ? x : x - d;
      for (int64_t y = 0; y < d; y++) {
        const int64_t ky = y <= h ? y : y - d;
        r2[(size_t)(kx * kx + ky * ky)]++;
      }
    }
    DevBuf dr2, dq;
    dr2.ensure(r2.size() * 4);
    dq.ensure(((size_t)qmax + 1) * 4);
    GENPK_HIP(hipMemcpy(dr2.ptr, r2.data(), r2.size() * 4, hipMemcpyHostToDevice));
    const unsigned grid = (unsigned)std::min<int64_t>(ceil_div(qmax + 1, 256), 65536);
    hipLaunchKernelGGL(k_qcount, dim3(grid), dim3(256), 0, p.stream, d, qmax, dr2.as<uint32_t>(), amax,
                       dq.as<uint32_t>());
    GENPK_LAUNCH_CHECK();
    std::vector<uint32_t> qcnt((size_t)qmax + 1);
    GENPK_HIP(hipMemcpyAsync(qcnt.data(), dq.ptr, qcnt.size() * 4, hipMemcpyDeviceToHost, p.stream));
    GENPK_HIP(hipStreamSynchronize(p.stream));
    std::vector<int64_t> cnt(nrbins, 0);
    std::vector<double> ke(nrbins, 0.0);
    for (int64_t q = 1; q <= qmax; q++) {
      if (!qcnt[q]) continue;
      cnt[tab[q]] += qcnt[q];
      ke[tab[q]] += (double)qcnt[q] * sqrt((double)q);
    }
    std::vector<int> cnt32(nrbins);
    for (int b = 0; b < nrbins; b++) {
      GENPK_REQUIRE(cnt[b] <= 2147483647, "a bin holds more than 2^31-1 modes");
      cnt32[b] = (int)cnt[b];
    }
    bt->count.ensure((size_t)nrbins * 4);
    bt->keffsum.ensure((size_t)nrbins * 8);
    GENPK_HIP(hipMemcpy(bt->count.ptr, cnt32.data(), (size_t)nrbins * 4, hipMemcpyHostToDevice));
    GENPK_HIP(hipMemcpy(bt->keffsum.ptr, ke.data(), (size_t)nrbins * 8, hipMemcpyHostToDevice));
  }
  bt->counts_ready = true;
  BinTables &ref = *bt;
  p.bins[nrbins] = std::move(bt);
  return ref;
}

const BinTables &built_bin_tables(const Plan &p, int nrbins) {
  GENPK_REQUIRE(nrbins >= 1 && nrbins <= 8192, "nrbins must be in [1, 8192]");
  auto it = p.bins.find(nrbins);
  GENPK_REQUIRE(it != p.bins.end() && it->second->counts_ready && p.wtab.ptr,
                "the bin tables of this nrbins are not built on the plan: call genpk_power or genpk_bin with it once "
                "first");
  return *it->second;
}

// Largest column block for the fused x-pass + bin kernel, 0 if none fits.
static int fused_block(int prec, int64_t d, int64_t hz, int nf, int ns, int nrbins) {
  const int csize = 2 * prec;
  const int elems = prec == 4 ? BinCfg<float>::ELEMS : BinCfg<double>::ELEMS;
  int B = col_block(prec, d, hz);
  while (B >= 1) {
    const int64_t lds = ((int64_t)nf * B + 1) * d * csize + (int64_t)((ns * nrbins + 1) & ~1) * 8 + d * 4 +
                        nrbins * 4;
    if (lds <= LDS_LIMIT && d * B <= elems && BIN_THREADS % B == 0) return B;
    B /= 2;
  }
  return 0;
}

template <typename T, int NF>
static size_t fused_lds(const Plan &a, int B, int nrbins, bool pp) {
  using C = typename Cx<T>::C;
  constexpr int NS = NF == 1 ? 1 : 3;
  return ((size_t)NF * B + (pp ? B : 0) + 1) * a.d * sizeof(C) + (size_t)((NS * nrbins + 1) & ~1) * 8 +
         (size_t)a.d * 4 + (size_t)nrbins * 4;
}

// Whether the REG1 instance applies to plan rp (one field with ping-pong
// buffers): every thread holds one first-stage butterfly, or (some_idle)
// at most one.
template <typename T>
static bool fused_reg1(const Plan &a, int B, bool pp, const RadixPlan &rp, bool some_idle = false) {
  const int64_t tot = a.d * B, cap = (int64_t)BIN_THREADS * BinCfg<T>::MAXE;
  return pp && rp.st[0].R == BinCfg<T>::MAXE && (some_idle ? tot <= cap : tot == cap) && rp.nst % 2 == 0;
}

// Persistent grid of the fused kernel: as many workgroups as stay resident
// (occupancy of this instance x CUs), at most one per column block.
template <typename T, int NF, bool REG1, bool R2L = false, int CT = 0>
static int fused_grid(const Plan &a, int B, int nrbins, bool pp) {
  int per_cu = 0;
  GENPK_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu,
                                                          reinterpret_cast<const void *>(&k_xpass_bin<T, NF, REG1, R2L, CT>),
                                                          BIN_THREADS, fused_lds<T, NF>(a, B, nrbins, pp)));
  const int64_t nblk = a.d * ceil_div(a.hz, B);
  return (int)std::max<int64_t>(1, std::min<int64_t>((int64_t)std::max(per_cu, 1) * num_cus(), nblk));
}

template <typename T, int NF, bool REG1, bool R2L = false, int CT = 0>
static void launch_fused(Plan &a, Plan *b, const BinTables &bt, const RadixPlan &rp, int B, bool keep, bool pp,
                         double *part, int G) {
  using C = typename Cx<T>::C;
  constexpr int NS = NF == 1 ? 1 : 3;
  // the R2L instance finishes the transform only in registers for the binning
  GENPK_REQUIRE(!(R2L && keep), "the fused pass cannot keep the spectrum with the last stage in the binning");
  const int nkb = (int)ceil_div(a.hz, B);
  const C *tw = (a.prec == 4) ? (const C *)a.fft.tw_d_f32.ptr : (const C *)a.fft.tw_d_f64.ptr;
  BinGeom g{a.d, a.hz, bt.nrbins, bt.q2bin.as<int>(), bt.thr.as<int64_t>(), a.wtab.as<float>(), bt.lq};
  const size_t lds = fused_lds<T, NF>(a, B, bt.nrbins, pp);
  Stage st(a, "xpass_bin");
  hipLaunchKernelGGL((k_xpass_bin<T, NF, REG1, R2L, CT>), dim3(G), dim3(BIN_THREADS), lds, a.stream, a.field.as<C>(),
                     b ? b->field.as<C>() : (C *)nullptr, g, B, nkb, FastDiv((uint32_t)nkb), rp, tw, keep ? 1 : 0, pp ? 1 : 0, part);
  GENPK_LAUNCH_CHECK();
}

template <typename T, int NF>
static void launch_stored(Plan &a, Plan *b, const BinTables &bt, double *part, int G) {
  using C = typename Cx<T>::C;
  constexpr int NS = NF == 1 ? 1 : 3;
  BinGeom g{a.d, a.hz, bt.nrbins, bt.q2bin.as<int>(), bt.thr.as<int64_t>(), a.wtab.as<float>(), bt.lq};
  Stage st(a, "bin_stored");
  hipLaunchKernelGGL((k_bin_stored<T, NF>), dim3(G), dim3(BIN_THREADS), (size_t)NS * bt.nrbins * 8, a.stream,
                     a.field.as<C>(), b ? b->field.as<C>() : (const C *)nullptr, g, part);
  GENPK_LAUNCH_CHECK();
}

// Launch shape of the fused x pass + bin: the radix plan of its transforms
// (rp), column block B, ping-pong buffers (pp), the register first stage
// (reg1), the radix-2 last stage in the binning (r2l), compile-time stages
// for the n = 1024 plan 8, 8, 8, 2 (ct, with r2l), and the persistent grid G.
struct FusedShape {
  RadixPlan rp;
  int B = 0;
  bool pp = false, reg1 = false, r2l = false, ct = false;
  int G = 0;
};

template <typename T>
static int fused_grid_of(const Plan &a, bool cross, const FusedShape &s, int nrbins) {
  if (cross) return fused_grid<T, 2, false>(a, s.B, nrbins, false);
  if (s.r2l && s.ct) return fused_grid<T, 1, true, true, 1024>(a, s.B, nrbins, s.pp);
  if (s.r2l) return fused_grid<T, 1, true, true>(a, s.B, nrbins, s.pp);
  if (s.reg1) return fused_grid<T, 1, true>(a, s.B, nrbins, s.pp);
  return fused_grid<T, 1, false>(a, s.B, nrbins, s.pp);
}

template <typename T>
static void launch_fused_of(Plan &a, Plan *b, const BinTables &bt, const FusedShape &s, double *part) {
  if (b) launch_fused<T, 2, false>(a, b, bt, s.rp, s.B, false, false, part, s.G);
  else if (s.r2l && s.ct) launch_fused<T, 1, true, true, 1024>(a, b, bt, s.rp, s.B, false, s.pp, part, s.G);
  else if (s.r2l) launch_fused<T, 1, true, true>(a, b, bt, s.rp, s.B, false, s.pp, part, s.G);
  else if (s.reg1) launch_fused<T, 1, true>(a, b, bt, s.rp, s.B, false, s.pp, part, s.G);
  else launch_fused<T, 1, false>(a, b, bt, s.rp, s.B, false, s.pp, part, s.G);
}

static FusedShape fused_shape(const Plan &a, const Plan *b, int nrbins) {
  const int nf = b ? 2 : 1, ns = b ? 3 : 1;
  FusedShape s;
  s.B = fused_block(a.prec, a.d, a.hz, nf, ns, nrbins);
  GENPK_REQUIRE(s.B > 0, "fused x-pass does not fit in LDS for this size");
  const int B = s.B;
  // one field: ping-pong FFT buffers where the LDS has room for them
  s.pp = !b && (a.prec == 4 ? fused_lds<float, 1>(a, B, nrbins, true)
                            : fused_lds<double, 1>(a, B, nrbins, true)) <= (size_t)LDS_LIMIT;
  s.rp = a.rp_d;
  s.reg1 = !b && (a.prec == 4 ? fused_reg1<float>(a, B, s.pp, s.rp) : fused_reg1<double>(a, B, s.pp, s.rp));
  // FP32 with a last stage of radix 2: that stage runs in the binning
  s.r2l = s.reg1 && a.prec == 4 && s.rp.st[s.rp.nst - 1].R == 2;
  if (!b && a.prec == 4 && !s.r2l) {
    // the plan's own ordering (the column passes') does not start with a
    // radix 8 and end with a radix 2 in an even number of stages, or its
    // columns leave threads idle: the x pass takes a plan of its length that
    // does, where the register first stage applies (512 = 8 8 8 -> 8 8 4 2:
    // 0.40 -> 0.30 ms at 512^3)
    const RadixPlan xp = radix_plan_8_to_2((int)a.d);
    if (xp.nst > 0 && fused_reg1<float>(a, B, s.pp, xp, true)) {
      s.rp = xp;
      s.reg1 = s.r2l = true;
    }
  }
  // compile-time stages where they pay: the FP32 register-first instance at
  // n = 1024 (the others would spill registers)
  s.ct = s.r2l && is_plan_8882(s.rp);
  const bool cross = b != nullptr;
  s.G = a.prec == 4 ? fused_grid_of<float>(a, cross, s, nrbins) : fused_grid_of<double>(a, cross, s, nrbins);
  return s;
}

int64_t fused_partial_doubles(const Plan &a, const Plan *b, int nrbins) {
  return (int64_t)fused_shape(a, b, nrbins).G * (b ? 3 : 1) * nrbins;
}

void bin_spectra(Plan &a, Plan *b, int nrbins, bool fused_xpass, double *pow_aa, double *pow_bb, double *pow_ab,
                 int *count, double *keffs, double ma, double mb, double *scratch, size_t scratch_doubles) {
  if (b) {
    GENPK_REQUIRE(b->d == a.d && b->prec == a.prec, "cross spectrum needs plans of equal dims and precision");
    GENPK_REQUIRE(b->stream == a.stream, "cross spectrum needs plans on the same stream");
  }
  const BinTables &bt = scratch ? built_bin_tables(a, nrbins) : bin_tables(a, nrbins);
  const int nf = b ? 2 : 1, ns = b ? 3 : 1;
  int G = std::max(1, std::min<int>(2 * num_cus(), (int)std::min<int64_t>(a.d * a.d, 1 << 20)));
  double *part = nullptr;
  // per-workgroup partial bins: the caller's scratch, or the plan's buffer
  auto partials = [&](int64_t doubles) -> double * {
    if (!scratch) return static_cast<double *>(a.bin_part.ensure((size_t)doubles * 8));
    GENPK_REQUIRE((size_t)doubles <= scratch_doubles,
                  "the partial bins (" + std::to_string(doubles) + " doubles: workgroups x spectra x nrbins) exceed "
                  "the scratch of " + std::to_string(scratch_doubles) + " doubles: use fewer bins");
    return scratch;
  };
  if (fused_xpass) {
    GENPK_REQUIRE(a.fstate == FieldState::YDone && (!b || b->fstate == FieldState::YDone),
                  "fused binning needs fields after the z and y passes");
    const FusedShape fs = fused_shape(a, b, nrbins);
    G = fs.G;
    part = partials((int64_t)G * ns * nrbins);
    if (a.prec == 4)
      launch_fused_of<float>(a, b, bt, fs, part);
    else
      launch_fused_of<double>(a, b, bt, fs, part);
    // the field keeps the forward state after the z and y passes (an adjoint can start from it)
    a.fstate = FieldState::YDone;
    if (b) b->fstate = FieldState::YDone;
  } else {
    GENPK_REQUIRE(a.fstate == FieldState::Spectrum && (!b || b->fstate == FieldState::Spectrum),
                  "binning needs transformed fields");
    part = partials((int64_t)G * ns * nrbins);
    if (a.prec == 4) {
      if (b) launch_stored<float, 2>(a, b, bt, part, G); else launch_stored<float, 1>(a, b, bt, part, G);
    } else {
      if (b) launch_stored<double, 2>(a, b, bt, part, G); else launch_stored<double, 1>(a, b, bt, part, G);
    }
  }
  hipLaunchKernelGGL(k_bin_finish, dim3((unsigned)ceil_div(nrbins, 256)), dim3(256), 0, a.stream, nrbins, ns, G, part,
                     bt.count.as<int>(), bt.keffsum.as<double>(), b ? ma * ma : ma * mb, mb * mb, ma * mb, pow_aa,
                     b ? pow_bb : (double *)nullptr, b ? pow_ab : (double *)nullptr, count, keffs);
  GENPK_LAUNCH_CHECK();
}

}  // namespace genpk
