// genpk_power_noalloc: genpk_power for callers that cannot allocate device
// memory (code under HIP graph capture, tight serving loops, harnesses that
// forbid hipMalloc).  No counterpart in the reference.
//
// genpk_power's only device allocation is the per-workgroup partial bins of
// the fused x pass + binning (a plan buffer, grown on first use; the bin
// tables are built once per nrbins).  Here the partial bins live in a
// file-scope device array of fixed size, and the bin tables must have been
// built by an earlier genpk_power or genpk_bin of that nrbins on the plan.
#include "plan.h"
#include "../../include/genpk_amd.h"

#include <map>
#include <mutex>

namespace genpk {
namespace {

// The partial bins: 2 Mi doubles = 16 MiB (resident workgroups x spectra x
// nrbins: e.g. 512 x 1 x 4096 or 256 x 3 x 2730), each GPU the library runs
// on holding its own copy until the process ends.  In host code the name
// `scratch` is a stand-in, not a device address: the device address comes
// from hipGetSymbolAddress, once per device with that device current, and
// only that address is given to kernels.  Every call on a GPU shares that
// GPU's array; the binning writes every partial it later reads, so nothing
// needs clearing, and calls that overlap on one GPU race.
constexpr size_t NOALLOC_SCRATCH_DOUBLES = (size_t)2 << 20;
static __device__ __align__(256) double scratch[NOALLOC_SCRATCH_DOUBLES];

double *scratch_address(int device) {
  static std::mutex mu;
  static std::map<int, double *> addr;
  int cur = -1;
  GENPK_HIP(hipGetDevice(&cur));
  GENPK_REQUIRE(cur == device, "call with the plan's device current");
  std::lock_guard<std::mutex> lk(mu);
  auto it = addr.find(device);
  if (it != addr.end()) return it->second;
  void *p = nullptr;
  GENPK_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  GENPK_REQUIRE(p != nullptr, "no device address for the scratch of genpk_power_noalloc");
  addr[device] = static_cast<double *>(p);
  return static_cast<double *>(p);
}

}  // namespace
}  // namespace genpk

using namespace genpk;

extern "C" int genpk_power_noalloc(genpk_plan *a, genpk_plan *b, int nrbins, double *power_aa, double *power_bb,
                                   double *power_ab, int *count, double *keffs, double ma, double mb) {
  return guarded([&] {
    GENPK_REQUIRE(a && a->p, "null plan");
    Plan &pa = *a->p;
    Plan *pb = b ? b->p.get() : nullptr;
    GENPK_REQUIRE(plan_fft_ok(pa), "the FFT needs an even dims with dims/2 = 2^a 3^b 5^c");
    GENPK_REQUIRE(density_state(pa), "plan a holds no density");
    if (pb) GENPK_REQUIRE(density_state(*pb), "plan b holds no density");
    if (pb) {
      GENPK_REQUIRE(pb->d == pa.d && pb->prec == pa.prec, "cross spectrum needs plans of equal dims and precision");
      GENPK_REQUIRE(pb->stream == pa.stream, "cross spectrum needs plans on the same stream");
    }
    (void)built_bin_tables(pa, nrbins);
    const int64_t need = fused_partial_doubles(pa, pb, nrbins);
    GENPK_REQUIRE(need <= (int64_t)NOALLOC_SCRATCH_DOUBLES,
                  "the partial bins (" + std::to_string(need) + " doubles: workgroups x spectra x nrbins) exceed the "
                  "scratch of " + std::to_string(NOALLOC_SCRATCH_DOUBLES) + " doubles: use fewer bins");
    double *part = scratch_address(pa.device);
    for (Plan *q : {&pa, pb}) {
      if (!q) continue;
      r2c_zpass(*q);
      c2c_ypass(*q, -1);
      q->fstate = FieldState::YDone;
    }
    bin_spectra(pa, pb, nrbins, true, power_aa, power_bb, power_ab, count, keffs, ma, pb ? mb : ma, part,
                NOALLOC_SCRATCH_DOUBLES);
  });
}
