// Bin lookup of an integer |k|^2 = q from thresholds staged in LDS, shared
// by the fused binning and the adjoint's fused weight.
#pragma once
#include <hip/hip_runtime.h>

namespace genpk {

// Bin of q from LDS copies of the thresholds (thr32[b] = smallest q whose
// bin is >= b, b < nrbins; the last bin holds every larger q): a candidate
// from log2 (off by at most one near a bin edge) moved to the bin whose
// [thr32[b], thr32[b+1]) holds q.
struct BinTrackLds {
  int cur = -1, lo = 0, hi = 0;
  __device__ __forceinline__ int of(int q, const int *thr32, int nrbins, float lq) {
    if (cur < 0 || q < lo || q >= hi) {
      int b = (int)(lq * __log2f((float)q));
      if (nrbins >= 4) {
        // the estimate is the bin or a neighbour: the four thresholds around
        // it load together (one LDS round trip) and settle those cases (the
        // last bin, past thr32[nrbins - 1], is left to the search below)
        b = min(max(b, 1), nrbins - 3);
        const int t0 = thr32[b - 1], t1 = thr32[b], t2 = thr32[b + 1], t3 = thr32[b + 2];
        if (q >= t1 && q < t2) {
          cur = b, lo = t1, hi = t2;
          return cur;
        }
        if (q >= t0 && q < t1) {
          cur = b - 1, lo = t0, hi = t1;
          return cur;
        }
        if (q >= t2 && q < t3) {
          cur = b + 1, lo = t2, hi = t3;
          return cur;
        }
      }
      b = min(max(b, 0), nrbins - 1);
      while (b > 0 && q < thr32[b]) b--;
      while (b < nrbins - 1 && q >= thr32[b + 1]) b++;
      cur = b;
      lo = thr32[b];
      hi = b + 1 < nrbins ? thr32[b + 1] : 0x7fffffff;
    }
    return cur;
  }
};

}  // namespace genpk
