// Exclusive prefix sum of uint32 arrays: the offset tables of the tile sort.
#include "plan.h"

namespace genpk {

namespace {

constexpr int SCAN_THREADS = 256;
constexpr int SCAN_PER_THREAD = 8;
constexpr int SCAN_BLOCK = SCAN_THREADS * SCAN_PER_THREAD;  // elements per workgroup

__device__ inline uint32_t wave_incl_scan(uint32_t v) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    uint32_t t = __shfl_up(v, off, 64);
    if (lane >= off) v += t;
  }
  return v;
}

// Scans SCAN_BLOCK elements per workgroup in place into out (exclusive) and
// writes the block total to sums[blockIdx.x] (if sums != nullptr).
__global__ void __launch_bounds__(SCAN_THREADS) k_scan_block(const uint32_t *__restrict__ in,
                                                             uint32_t *__restrict__ out, int64_t n,
                                                             uint32_t *__restrict__ sums) {
  __shared__ uint32_t wsum[SCAN_THREADS / 64];
  const int64_t base = (int64_t)blockIdx.x * SCAN_BLOCK + (int64_t)threadIdx.x * SCAN_PER_THREAD;
  uint32_t v[SCAN_PER_THREAD];
  uint32_t tot = 0;
#pragma unroll
  for (int i = 0; i < SCAN_PER_THREAD; i++) {
    const int64_t at = base + i;
    v[i] = at < n ? in[at] : 0u;
    tot += v[i];
  }
  uint32_t incl = wave_incl_scan(tot);
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (lane == 63) wsum[wave] = incl;
  __syncthreads();
  uint32_t wpre = 0;
  for (int w = 0; w < wave; w++) wpre += wsum[w];
  uint32_t run = wpre + incl - tot;
#pragma unroll
  for (int i = 0; i < SCAN_PER_THREAD; i++) {
    const int64_t at = base + i;
    const uint32_t x = v[i];
    if (at < n) out[at] = run;
    run += x;
  }
  if (sums && threadIdx.x == SCAN_THREADS - 1) sums[blockIdx.x] = run;
}

__global__ void __launch_bounds__(SCAN_THREADS) k_scan_add(uint32_t *__restrict__ out, int64_t n,
                                                           const uint32_t *__restrict__ sums) {
  const uint32_t add = sums[blockIdx.x];
  const int64_t base = (int64_t)blockIdx.x * SCAN_BLOCK;
  for (int i = threadIdx.x; i < SCAN_BLOCK; i += SCAN_THREADS) {
    const int64_t at = base + i;
    if (at < n) out[at] += add;
  }
}

void scan_rec(const uint32_t *in, uint32_t *out, int64_t n, uint32_t *tmp, hipStream_t s) {
  const int64_t nb = ceil_div(n, SCAN_BLOCK);
  if (nb <= 1) {
    hipLaunchKernelGGL(k_scan_block, dim3(1), dim3(SCAN_THREADS), 0, s, in, out, n, (uint32_t *)nullptr);
    GENPK_LAUNCH_CHECK();
    return;
  }
  uint32_t *sums = tmp;
  uint32_t *sums_scanned = tmp + nb;
  hipLaunchKernelGGL(k_scan_block, dim3((unsigned)nb), dim3(SCAN_THREADS), 0, s, in, out, n, sums);
  GENPK_LAUNCH_CHECK();
  scan_rec(sums, sums_scanned, nb, tmp + 2 * nb, s);
  hipLaunchKernelGGL(k_scan_add, dim3((unsigned)nb), dim3(SCAN_THREADS), 0, s, out, n, sums_scanned);
  GENPK_LAUNCH_CHECK();
}

}  // namespace

void exclusive_scan_u32(const uint32_t *in, uint32_t *out, int64_t n, DevBuf &tmp, hipStream_t s) {
  if (n <= 0) return;
  // temporary space: 2 * (sum over levels of block counts)
  int64_t need = 0, m = n;
  while (m > SCAN_BLOCK) {
    m = ceil_div(m, SCAN_BLOCK);
    need += 2 * m;
  }
  uint32_t *t = static_cast<uint32_t *>(tmp.ensure((size_t)(need + 16) * sizeof(uint32_t)));
  scan_rec(in, out, n, t, s);
}

}  // namespace genpk
