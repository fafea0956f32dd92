// LDS-resident mixed-radix Stockham FFT building blocks for gfx950.
//
// A workgroup holds a batch of transforms of length n in LDS, element e of
// transform b at buf[e*se + b*sb].  Each stage of radix R reads R values
// per butterfly (twiddled), all threads pass a barrier, then write the
// DFT_R outputs to their autosorted positions; the buffer is reused in
// place.  Twiddles come from a table tw[m] = exp(-2 pi i m / n), m < n,
// conjugated for the inverse direction (SIGN = +1).
#pragma once
#include "common.h"

namespace genpk {

template <typename T> struct Cx;
template <> struct Cx<float> { using C = float2; };
template <> struct Cx<double> { using C = double2; };

template <typename C> __device__ __forceinline__ C cadd(C a, C b) { return C{a.x + b.x, a.y + b.y}; }
template <typename C> __device__ __forceinline__ C csub(C a, C b) { return C{a.x - b.x, a.y - b.y}; }
template <typename C> __device__ __forceinline__ C cmul(C a, C b) {
  return C{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <typename C> __device__ __forceinline__ C cconj(C a) { return C{a.x, -a.y}; }
// multiply by SIGN*i
template <int SIGN, typename C> __device__ __forceinline__ C cmul_si(C a) {
  return SIGN < 0 ? C{a.y, -a.x} : C{-a.y, a.x};
}
template <typename C, typename S> __device__ __forceinline__ C cscale(C a, S s) { return C{a.x * s, a.y * s}; }

// In-place DFT of length R on v[0..R), exponent sign SIGN.
template <int R, int SIGN, typename C> struct Dft;

template <int SIGN, typename C> struct Dft<2, SIGN, C> {
  __device__ __forceinline__ static void run(C *v) {
    const C a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
  }
};

template <int SIGN, typename C> struct Dft<4, SIGN, C> {
  __device__ __forceinline__ static void run(C *v) {
    const C s02 = cadd(v[0], v[2]), d02 = csub(v[0], v[2]);
    const C s13 = cadd(v[1], v[3]), d13 = cmul_si<SIGN>(csub(v[1], v[3]));
    v[0] = cadd(s02, s13);
    v[2] = csub(s02, s13);
    v[1] = cadd(d02, d13);
    v[3] = csub(d02, d13);
  }
};

template <int SIGN, typename C> struct Dft<8, SIGN, C> {
  __device__ __forceinline__ static void run(C *v) {
    using S = decltype(v[0].x);
    const S h = (S)0.70710678118654752440084436210484903928;  // sqrt(1/2)
    // radix-2 over stride 4, then two radix-4 with the 8th-root twiddles
    C a[4], b[4];
#pragma unroll
    for (int i = 0; i < 4; i++) {
      a[i] = cadd(v[i], v[i + 4]);
      b[i] = csub(v[i], v[i + 4]);
    }
    // b[k] *= w8^k, w8 = exp(SIGN * 2 pi i / 8)
    b[1] = C{(b[1].x - (S)SIGN * b[1].y) * h, ((S)SIGN * b[1].x + b[1].y) * h};
    b[2] = cmul_si<SIGN>(b[2]);
    b[3] = C{(-b[3].x - (S)SIGN * b[3].y) * h, ((S)SIGN * b[3].x - b[3].y) * h};
    Dft<4, SIGN, C>::run(a);
    Dft<4, SIGN, C>::run(b);
    v[0] = a[0]; v[2] = a[1]; v[4] = a[2]; v[6] = a[3];
    v[1] = b[0]; v[3] = b[1]; v[5] = b[2]; v[7] = b[3];
  }
};

template <int SIGN, typename C> struct Dft<3, SIGN, C> {
  __device__ __forceinline__ static void run(C *v) {
    using S = decltype(v[0].x);
    const S c = (S)-0.5;
    const S s = (S)SIGN * (S)0.86602540378443864676372317075293618347;  // sin(2 pi / 3)
    const C t = cadd(v[1], v[2]);
    const C u = csub(v[1], v[2]);
    const C m = C{v[0].x + c * t.x, v[0].y + c * t.y};
    const C r = C{-s * u.y, s * u.x};  // i*s*u
    v[0] = cadd(v[0], t);
    v[1] = cadd(m, r);
    v[2] = csub(m, r);
  }
};

template <int SIGN, typename C> struct Dft<5, SIGN, C> {
  __device__ __forceinline__ static void run(C *v) {
    using S = decltype(v[0].x);
    const S c1 = (S)0.30901699437494742410229341718281905886;   // cos(2 pi/5)
    const S c2 = (S)-0.80901699437494742410229341718281905886;  // cos(4 pi/5)
    const S s1 = (S)SIGN * (S)0.95105651629515357211643933337938214340;  // sin(2 pi/5)
    const S s2 = (S)SIGN * (S)0.58778525229247312916870595463907276860;  // sin(4 pi/5)
    const C t1 = cadd(v[1], v[4]), t2 = cadd(v[2], v[3]);
    const C u1 = csub(v[1], v[4]), u2 = csub(v[2], v[3]);
    const C a1 = C{v[0].x + c1 * t1.x + c2 * t2.x, v[0].y + c1 * t1.y + c2 * t2.y};
    const C a2 = C{v[0].x + c2 * t1.x + c1 * t2.x, v[0].y + c2 * t1.y + c1 * t2.y};
    // i*(s1 u1 + s2 u2), i*(s2 u1 - s1 u2)
    const C b1 = C{-(s1 * u1.y + s2 * u2.y), s1 * u1.x + s2 * u2.x};
    const C b2 = C{-(s2 * u1.y - s1 * u2.y), s2 * u1.x - s1 * u2.x};
    v[0] = cadd(v[0], cadd(t1, t2));
    v[1] = cadd(a1, b1);
    v[4] = csub(a1, b1);
    v[2] = cadd(a2, b2);
    v[3] = csub(a2, b2);
  }
};

// Division by a runtime constant d (< 2^31) with a multiply-high, an add
// and a shift (Granlund-Montgomery round-up method), exact for n < 2^31.
struct FastDiv {
  uint32_t d = 1, mul = 0, shift = 0;
  FastDiv() = default;
  explicit FastDiv(uint32_t dv) : d(dv) {
    if (dv <= 1) return;
    uint32_t l = 0;
    while ((1ull << l) < dv) l++;
    const unsigned __int128 one = 1;
    const unsigned __int128 m = ((one << (32 + l)) + dv - 1) / dv;
    mul = (uint32_t)(m - (one << 32));
    shift = l - 1;
  }
  __host__ __device__ __forceinline__ uint32_t div(uint32_t n) const {
    if (d <= 1) return n;
#ifdef __HIP_DEVICE_COMPILE__
    const uint32_t t = __umulhi(n, mul);
#else
    const uint32_t t = (uint32_t)(((uint64_t)n * mul) >> 32);
#endif
    return (t + ((n - t) >> 1)) >> shift;
  }
};

// One Stockham stage: radix R, ns = product of the previous radices,
// nq = n / R butterflies per transform, twstep = n / (ns R).
struct StagePlan {
  int R = 1, ns = 1, nq = 1, twstep = 1;
  FastDiv fns, fnq;
};

// Radix sequence of a length with its per-stage index dividers.
struct RadixPlan {
  int n = 1;
  int nst = 0;
  StagePlan st[14];
};

// Radix orderings (variant): 0 = radix 8 first, then 4, 2, 3, 5 (the
// fewest stages; the leftover small radices last); 1 = the same radices
// with the leftover powers of two, 3s and 5s first and the 8s last; 2 =
// radix 4 first (no 8), then 2, 3, 5 (more stages, smaller butterflies);
// 3 = one radix 4 first, then as 0 (1024 = 4 8 8 4: a first stage that the
// FP64 column kernels, four values per thread, run from registers; plans
// use it for the column passes only).  Plans start with 3 in FP64 (at 1024^3
// the y and fused x passes take 3.92 + 4.67 -> 3.83 + 4.15 ms) and 0 in FP32.
constexpr int FFT_VARIANTS = 4;
inline int default_fft_variant(int prec) { return prec == 8 ? 3 : 0; }

// Stockham stage plans of the radix sequence r[0..nr) (product n).
inline RadixPlan radix_plan_of(int n, const int *r, int nr) {
  RadixPlan p;
  p.n = n;
  int ns = 1;
  for (int s = 0; s < nr; s++) {
    StagePlan &sp = p.st[s];
    sp.R = r[s];
    sp.ns = ns;
    sp.nq = n / r[s];
    sp.twstep = n / (ns * r[s]);
    sp.fns = FastDiv((uint32_t)ns);
    sp.fnq = FastDiv((uint32_t)sp.nq);
    ns *= r[s];
  }
  p.nst = nr;
  return p;
}

inline RadixPlan make_radix_plan(int n, int variant = 0) {
  RadixPlan p;
  p.n = n;
  int r[32], nr = 0, m = n;
  if (variant == 3 && m % 4 == 0) { r[nr++] = 4; m /= 4; }
  if (variant != 2)
    while (m % 8 == 0) { r[nr++] = 8; m /= 8; }
  while (m % 4 == 0) { r[nr++] = 4; m /= 4; }
  while (m % 2 == 0) { r[nr++] = 2; m /= 2; }
  while (m % 3 == 0) { r[nr++] = 3; m /= 3; }
  while (m % 5 == 0) { r[nr++] = 5; m /= 5; }
  if (m != 1 || nr > 14) {
    p.nst = -1;  // unsupported prime factor
    return p;
  }
  if (variant == 1) {
    // the 8s (leading) move behind the other radices
    int e = 0;
    while (e < nr && r[e] == 8) e++;
    int t[32], k = 0;
    for (int s = e; s < nr; s++) t[k++] = r[s];
    for (int s = 0; s < e; s++) t[k++] = 8;
    for (int s = 0; s < nr; s++) r[s] = t[s];
  }
  return radix_plan_of(n, r, nr);
}

// Plan of a length n = 16 m with a first radix 8, an even number of stages
// and a last radix 2.  Powers of two take the fewest such stages (the middle
// radices as even as possible, larger first): 16 = 8 2, 64 = 8 2 2 2, 128 =
// 8 4 2 2, 256 = 8 4 4 2, 512 = 8 8 4 2, 1024 = 8 8 8 2.  Other lengths take
// ordering 0 of m in the middle, its first 8 (else 4) split in two where
// that leaves an odd count: 384 = 8 8 3 2, 96 = 8 2 3 2.  nst = -1 where no
// such plan exists (32, 48, lengths not divisible by 16).
inline RadixPlan radix_plan_8_to_2(int n) {
  RadixPlan p;
  p.n = n;
  p.nst = -1;
  if (n < 16 || n % 16 != 0) return p;
  int r[14], nr = 0;
  r[nr++] = 8;
  if ((n & (n - 1)) == 0) {
    int t = -4;  // log2 of the middle stages' product
    for (int m = n; m > 1; m >>= 1) t++;
    int k = (t + 2) / 3;  // fewest middle stages of radix <= 8 ...
    if (k % 2) k++;       // ... in an even number
    if (k > t) return p;
    for (int i = 0; i < k; i++) r[nr++] = 1 << (t / k + (i < t % k ? 1 : 0));
  } else {
    const RadixPlan mp = make_radix_plan(n / 16, 0);
    if (mp.nst < 0 || mp.nst > 10) return p;
    int split = -1;
    if (mp.nst % 2) {
      for (int s = 0; s < mp.nst && split < 0; s++)
        if (mp.st[s].R == 8) split = s;
      for (int s = 0; s < mp.nst && split < 0; s++)
        if (mp.st[s].R == 4) split = s;
      if (split < 0) return p;
    }
    for (int s = 0; s < mp.nst; s++) {
      if (s == split) {
        r[nr++] = mp.st[s].R / 2;
        r[nr++] = 2;
      } else {
        r[nr++] = mp.st[s].R;
      }
    }
  }
  r[nr++] = 2;
  return radix_plan_of(n, r, nr);
}

// The column-pass plan that runs compile-time stages (fft_stage_ct): n =
// 1024 as radices 8, 8, 8, 2 (d = 1024, radix ordering 0).
inline bool is_plan_8882(const RadixPlan &p) {
  return p.n == 1024 && p.nst == 4 && p.st[0].R == 8 && p.st[1].R == 8 && p.st[2].R == 8 && p.st[3].R == 2;
}

// The FP64 column-pass plan that runs compile-time stages: n = 1024 as
// radices 4, 8, 8, 4 (d = 1024, radix ordering 3, the FP64 default).
inline bool is_plan_4884(const RadixPlan &p) {
  return p.n == 1024 && p.nst == 4 && p.st[0].R == 4 && p.st[1].R == 8 && p.st[2].R == 8 && p.st[3].R == 4;
}

inline bool fft_length_ok(int64_t n) {
  if (n < 1 || n > (1 << 20)) return false;
  return make_radix_plan((int)n).nst >= 0;
}

// LDS position of element index i: PAD inserts one element after every 32,
// so that accesses at a stride of R elements (the writes of a Stockham
// stage) spread over the LDS banks.
template <bool PAD> __device__ __forceinline__ int lds_ix(int i) { return PAD ? i + (i >> 5) : i; }
template <bool PAD> __host__ __device__ inline int64_t lds_len(int64_t n) { return PAD ? n + (n >> 5) + 1 : n; }

// One Stockham stage of radix R on nbatch transforms held in LDS, element
// e of transform b at in[e*se + b*sb], results to out (same layout; with
// PAD, index i at lds_ix(i)).  MAXE bounds the elements a thread owns per
// stage.  BFAST: the batch index varies fastest across threads
// (interleaved layout, nbatch = 1 << lgb); otherwise the element index
// does.  In place (in == out) all reads pass a barrier before the writes;
// out of place (PP) one barrier ends the stage.
template <int R, int SIGN, bool BFAST, int MAXE, bool PP, typename C, bool PAD = false>
__device__ __forceinline__ void fft_stage(const C *in, C *out, int nbatch, int lgb, int se, int sb,
                                          const StagePlan &sp, const C *__restrict__ tw) {
  constexpr int MAXB = (MAXE + R - 1) / R;
  const int nq = sp.nq, ns = sp.ns;
  const int nbf = nbatch * nq;
  C v[MAXB][R];
  int dsto[MAXB];
#pragma unroll
  for (int i = 0; i < MAXB; i++) {
    const int u = threadIdx.x + i * blockDim.x;
    if (u < nbf) {
      int b, j;
      if (BFAST) {
        b = u & (nbatch - 1);
        j = u >> lgb;
      } else {
        b = (int)sp.fnq.div((uint32_t)u);
        j = u - b * nq;
      }
      const int jq = (int)sp.fns.div((uint32_t)j);
      const int k = j - jq * ns;
      const int i0 = b * sb + j * se;
      auto src = [&](int r) -> C {
        if constexpr (PAD) return in[lds_ix<true>(i0 + r * nq * se)];
        else return in[i0 + r * nq * se];
      };
      v[i][0] = src(0);
      if (ns == 1) {
        // first stage: every twiddle is 1
#pragma unroll
        for (int r = 1; r < R; r++) v[i][r] = src(r);
      } else {
#pragma unroll
        for (int r = 1; r < R; r++) {
          C t = tw[r * k * sp.twstep];
          if (SIGN > 0) t.y = -t.y;
          v[i][r] = cmul(src(r), t);
        }
      }
      dsto[i] = b * sb + (jq * ns * R + k) * se;
    }
  }
  if (!PP) __syncthreads();
#pragma unroll
  for (int i = 0; i < MAXB; i++) {
    const int u = threadIdx.x + i * blockDim.x;
    if (u < nbf) {
      Dft<R, SIGN, C>::run(v[i]);
#pragma unroll
      for (int r = 0; r < R; r++) {
        if constexpr (PAD) out[lds_ix<true>(dsto[i] + r * ns * se)] = v[i][r];
        else out[dsto[i] + r * ns * se] = v[i][r];
      }
    }
  }
  __syncthreads();
}

template <int SIGN, bool BFAST, int MAXE, bool PP, typename C, bool PAD = false>
__device__ __forceinline__ void fft_stage_r(int R, const C *in, C *out, int nbatch, int lgb, int se, int sb,
                                            const StagePlan &sp, const C *__restrict__ tw) {
  switch (R) {
    case 8: fft_stage<8, SIGN, BFAST, MAXE, PP, C, PAD>(in, out, nbatch, lgb, se, sb, sp, tw); break;
    case 4: fft_stage<4, SIGN, BFAST, MAXE, PP, C, PAD>(in, out, nbatch, lgb, se, sb, sp, tw); break;
    case 2: fft_stage<2, SIGN, BFAST, MAXE, PP, C, PAD>(in, out, nbatch, lgb, se, sb, sp, tw); break;
    case 3: fft_stage<3, SIGN, BFAST, MAXE, PP, C, PAD>(in, out, nbatch, lgb, se, sb, sp, tw); break;
    case 5: fft_stage<5, SIGN, BFAST, MAXE, PP, C, PAD>(in, out, nbatch, lgb, se, sb, sp, tw); break;
    default: break;
  }
}

// Full transform in place; the caller has filled buf and synchronised.
template <int SIGN, bool BFAST, int MAXE, typename C, bool PAD = false>
__device__ __forceinline__ void lds_fft(C *buf, const RadixPlan &rp, int nbatch, int lgb, int se, int sb,
                                        const C *__restrict__ tw) {
  for (int s = 0; s < rp.nst; s++)
    fft_stage_r<SIGN, BFAST, MAXE, false, C, PAD>(rp.st[s].R, buf, buf, nbatch, lgb, se, sb, rp.st[s], tw);
}

// Stages s0 .. s1 - 1 (s1 < 0: to the last) of a transform alternating
// between buffers a and b (one barrier per stage); returns the buffer that
// holds the result.  The caller has filled a and synchronised.
template <int SIGN, bool BFAST, int MAXE, typename C>
__device__ __forceinline__ C *lds_fft_pp(C *a, C *b, const RadixPlan &rp, int nbatch, int lgb, int se, int sb,
                                         const C *__restrict__ tw, int s0 = 0, int s1 = -1) {
  const int se1 = s1 < 0 ? rp.nst : s1;
  for (int s = s0; s < se1; s++) {
    fft_stage_r<SIGN, BFAST, MAXE, true>(rp.st[s].R, a, b, nbatch, lgb, se, sb, rp.st[s], tw);
    C *t = a;
    a = b;
    b = t;
  }
  return a;
}

// fft_stage with the radix, the length N and the span NS (product of the
// previous radices) known at compile time, all powers of two: the butterfly
// and twiddle indices are shifts and masks, and in the row layout (element
// e of transform b at b * N + e) the R reads of a butterfly sit at
// compile-time offsets from one address.  Same layouts and barriers as
// fft_stage (se = nbatch, sb = 1 when BFAST; se = 1, sb = N otherwise).
template <int R, int N, int NS, int SIGN, bool BFAST, int MAXE, bool PP, typename C, bool PAD>
__device__ __forceinline__ void fft_stage_ct(const C *in, C *out, int nbatch, int lgb, const C *__restrict__ tw) {
  static_assert((N & (N - 1)) == 0 && (NS & (NS - 1)) == 0 && N % (NS * R) == 0, "powers of two");
  constexpr uint32_t NQ = N / R, TWS = N / (NS * R);
  constexpr int MAXB = (MAXE + R - 1) / R;
  const uint32_t nbf = (uint32_t)nbatch * NQ;
  const uint32_t se = BFAST ? (uint32_t)nbatch : 1u;
  C v[MAXB][R];
  uint32_t dsto[MAXB];
#pragma unroll
  for (int i = 0; i < MAXB; i++) {
    const uint32_t u = threadIdx.x + i * blockDim.x;
    if (u < nbf) {
      uint32_t b, j;
      if (BFAST) {
        b = u & (uint32_t)(nbatch - 1);
        j = u >> lgb;
      } else {
        b = u / NQ;
        j = u % NQ;
      }
      const uint32_t jq = j / NS, k = j % NS;
      const uint32_t i0 = BFAST ? b + j * se : b * N + j;
      auto src = [&](int r) -> C {
        if constexpr (PAD) return in[lds_ix<true>((int)(i0 + r * NQ * se))];
        else return in[i0 + r * NQ * se];
      };
      v[i][0] = src(0);
#pragma unroll
      for (int r = 1; r < R; r++) {
        if constexpr (NS == 1) {
          v[i][r] = src(r);
        } else {
          C t = tw[r * k * TWS];
          if (SIGN > 0) t.y = -t.y;
          v[i][r] = cmul(src(r), t);
        }
      }
      dsto[i] = (BFAST ? b : b * N) + (jq * NS * R + k) * se;
    }
  }
  if (!PP) __syncthreads();
#pragma unroll
  for (int i = 0; i < MAXB; i++) {
    const uint32_t u = threadIdx.x + i * blockDim.x;
    if (u < nbf) {
      Dft<R, SIGN, C>::run(v[i]);
#pragma unroll
      for (int r = 0; r < R; r++) {
        if constexpr (PAD) out[lds_ix<true>((int)(dsto[i] + r * NS * se))] = v[i][r];
        else out[dsto[i] + r * NS * se] = v[i][r];
      }
    }
  }
  __syncthreads();
}

// Stages R, RS... of a length-N transform starting at span NS (fft_stage_ct
// each), alternating between buffers a and b when PP, in place otherwise;
// returns the buffer that holds the result.
template <int N, int NS, int SIGN, bool BFAST, int MAXE, bool PP, typename C, bool PAD, int R, int... RS>
__device__ __forceinline__ C *lds_fft_ct(C *a, C *b, int nbatch, int lgb, const C *__restrict__ tw) {
  C *dst = PP ? b : a;
  fft_stage_ct<R, N, NS, SIGN, BFAST, MAXE, PP, C, PAD>(a, dst, nbatch, lgb, tw);
  if constexpr (sizeof...(RS) > 0)
    return lds_fft_ct<N, NS * R, SIGN, BFAST, MAXE, PP, C, PAD, RS...>(dst, PP ? a : b, nbatch, lgb, tw);
  else
    return dst;
}

// First stage (every twiddle 1) of a radix-R transform whose R inputs this
// thread holds in v: elements j + r * nq (nq = n / R) of transform b, with
// b = threadIdx.x & (nbatch - 1) and j = threadIdx.x >> lgb (one butterfly
// per thread, batch index fastest, as fft_stage<R, SIGN, true>).  Writes the
// outputs to out, with no barrier.
template <int R, int SIGN, typename C>
__device__ __forceinline__ void fft_first_regs(C (&v)[R], C *out, int nbatch, int lgb, int se, int sb) {
  Dft<R, SIGN, C>::run(v);
  const int b = threadIdx.x & (nbatch - 1), j = threadIdx.x >> lgb;
  C *dst = out + b * sb + j * R * se;
#pragma unroll
  for (int r = 0; r < R; r++) dst[r * se] = v[r];
}

__host__ __device__ inline int ilog2(int64_t v) {
  int l = 0;
  while (((int64_t)1 << l) < v) l++;
  return l;
}

}  // namespace genpk
