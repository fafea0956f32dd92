// Shared definitions for the MI355X (gfx950) power-spectrum library.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstddef>
#include <cstdio>
#include <string>
#include <new>
#include <stdexcept>

namespace genpk {

// Error reporting.  Every C entry point catches genpk::Error and stores the
// message in a thread-local string returned by genpk_last_error().
struct Error : public std::runtime_error {
  explicit Error(const std::string &m) : std::runtime_error(m) {}
};

void set_last_error(const std::string &msg);

#define GENPK_HIP(call)                                                                 \
  do {                                                                                  \
    hipError_t _e = (call);                                                             \
    if (_e != hipSuccess)                                                               \
      throw ::genpk::Error(std::string(#call) + " failed: " + hipGetErrorString(_e) +   \
                           " (" __FILE__ ":" + std::to_string(__LINE__) + ")");         \
  } while (0)

#define GENPK_REQUIRE(cond, msg)                    \
  do {                                              \
    if (!(cond)) throw ::genpk::Error(msg);         \
  } while (0)

#define GENPK_LAUNCH_CHECK() GENPK_HIP(hipGetLastError())

// Body of every C entry point: runs f, maps exceptions to a nonzero status
// and the thread's last-error message.
template <typename F>
int guarded(F &&f) {
  try {
    f();
    set_last_error("");
    return 0;
  } catch (const Error &e) {
    set_last_error(e.what());
    return 1;
  } catch (const std::bad_alloc &) {
    set_last_error("host allocation failed");
    return 2;
  } catch (const std::exception &e) {
    set_last_error(e.what());
    return 3;
  } catch (...) {
    set_last_error("unknown error");
    return 4;
  }
}

// Fixed-point scale of the LDS deposit accumulators: CIC weights in [0, 1]
// are accumulated as unsigned 64-bit integers with 36 fractional bits
// (resolution 1.5e-11 of a particle's mass).  A work item holds at most
// DEPOSIT_ITEM_MAX = 2^17 particles, so a cell sum stays at or below 2^53
// and converts to double exactly.
constexpr int FIXED_FRAC_BITS = 36;
constexpr double FIXED_ONE = 68719476736.0;            // 2^36
constexpr double FIXED_INV = 1.0 / 68719476736.0;      // 2^-36
constexpr int DEPOSIT_ITEM_MAX = 131072;
// Smallest item cap that deposit_item_max (tilesort.h) takes for few tiles.
constexpr int DEPOSIT_ITEM_MIN = 16384;

template <typename T> struct cplx_of;
template <> struct cplx_of<float> { using type = float2; };
template <> struct cplx_of<double> { using type = double2; };

__host__ __device__ inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

// Workgroups are dispatched to the 8 XCDs (each with its own L2) round
// robin by workgroup id.  xcd_block maps the dispatched id p of a grid of g
// workgroups to a logical block so that every XCD runs one contiguous range
// of logical blocks, in order: neighbouring blocks, which share cache lines
// at their edges, then meet in the same L2.
constexpr int NUM_XCD = 8;
__device__ __forceinline__ uint32_t xcd_block(uint32_t p, uint32_t g) {
  const uint32_t q = g / NUM_XCD, r = g % NUM_XCD;
  const uint32_t x = p % NUM_XCD, slot = p / NUM_XCD;
  return x * q + min(x, r) + slot;
}

}  // namespace genpk
