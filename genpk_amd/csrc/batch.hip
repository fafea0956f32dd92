// Batched forward: many independent realisations (uneven particle counts and
// grid sizes) captured into one hipGraph and launched with one call.
//
// Jobs are spread over `branches` parallel graph branches, each with its own
// stream and its own plans (one per grid size and precision used on the
// branch), so small jobs that cannot fill the 256 CUs run side by side.
// Build: assign jobs to branches (longest-processing-time first unless the
// caller gives the assignment), run every job once eagerly (this sizes every
// plan buffer and builds the bin tables, so the capture allocates nothing),
// then capture fork -> per-branch job sequences -> join on the batch stream
// and instantiate.  Every kernel of a job rewrites what it reads later (the
// deposit writes every tile, the binning every partial), so replays are
// independent of earlier launches.  No counterpart in the reference.
#include "fft_core.h"
#include "plan.h"
#include "../../include/genpk_amd.h"

#include <algorithm>
#include <array>
#include <map>
#include <memory>
#include <tuple>
#include <vector>

namespace genpk {

struct BatchJob {
  int64_t d;
  int prec;
  double box;
  int64_t n;
  const float *pos;
  const float *masses;
  double mass;
  int nrbins;
  double total_mass;
  double *power;
  int *count;
  double *keffs;
  int branch = 0;
};

struct Batch {
  hipStream_t stream = nullptr;   // the caller's stream: launches go there
  hipStream_t cap = nullptr;      // capture stream (the legacy default stream cannot be captured)
  int branches = 1;
  std::vector<BatchJob> jobs;
  std::vector<hipStream_t> bstreams;
  std::vector<hipEvent_t> events;
  std::map<std::tuple<int, int64_t, int>, std::unique_ptr<Plan>> plans;
  // per (dims, precision): deposit tile edges x, y, z and FFT radix variant
  std::map<std::pair<int64_t, int>, std::array<int, 4>> options;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;

  ~Batch() { reset(); }
  void reset() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
    plans.clear();
    for (auto e : events) (void)hipEventDestroy(e);
    events.clear();
    for (auto s : bstreams) (void)hipStreamDestroy(s);
    bstreams.clear();
    if (cap) (void)hipStreamDestroy(cap);
    cap = nullptr;
  }
  Plan &plan_for(const BatchJob &j) {
    auto key = std::make_tuple(j.branch, j.d, j.prec);
    auto it = plans.find(key);
    if (it != plans.end()) return *it->second;
    auto p = std::make_unique<Plan>(j.d, j.prec, bstreams[j.branch]);
    auto o = options.find(std::make_pair(j.d, j.prec));
    if (o != options.end()) {
      plan_set_tiles(*p, o->second[0], o->second[1], o->second[2]);
      plan_set_fft_variant(*p, o->second[3]);
    }
    Plan &r = *p;
    plans[key] = std::move(p);
    return r;
  }
};

// Relative cost used to balance branches: particles moved by the sort and
// deposit plus the grid passes.
static double job_cost(const BatchJob &j) {
  return 60.0 * (double)j.n + 6.0 * (double)j.prec * (double)j.d * (double)j.d * (double)(j.d + 2) + 2e6;
}

// One forward: clear, deposit, fused r2c + bin (genpk_clear, genpk_deposit, genpk_power).
static void forward_job(Plan &p, const BatchJob &j) {
  p.tiles_live = false;
  p.fstate = FieldState::None;
  deposit(p, j.box, j.n, j.pos, j.masses, j.mass);
  (void)bin_tables(p, j.nrbins);
  r2c_zpass(p);
  c2c_ypass(p, -1);
  p.fstate = FieldState::YDone;
  bin_spectra(p, nullptr, j.nrbins, true, j.power, nullptr, nullptr, j.count, j.keffs, j.total_mass, j.total_mass);
}

static void batch_build(Batch &b, const int *assign) {
  GENPK_REQUIRE(!b.jobs.empty(), "batch has no jobs");
  b.reset();
  GENPK_HIP(hipStreamSynchronize(b.stream));  // the jobs' inputs are ready
  const int B = b.branches;
  if (assign) {
    for (size_t i = 0; i < b.jobs.size(); i++) {
      GENPK_REQUIRE(assign[i] >= 0 && assign[i] < B, "branch assignment out of range");
      b.jobs[i].branch = assign[i];
    }
  } else {
    std::vector<size_t> order(b.jobs.size());
    for (size_t i = 0; i < order.size(); i++) order[i] = i;
    std::stable_sort(order.begin(), order.end(),
                     [&](size_t x, size_t y) { return job_cost(b.jobs[x]) > job_cost(b.jobs[y]); });
    std::vector<double> load(B, 0.0);
    for (size_t i : order) {
      const int k = (int)(std::min_element(load.begin(), load.end()) - load.begin());
      b.jobs[i].branch = k;
      load[k] += job_cost(b.jobs[i]);
    }
  }
  GENPK_HIP(hipStreamCreateWithFlags(&b.cap, hipStreamNonBlocking));
  for (int k = 0; k < B; k++) {
    hipStream_t s;
    GENPK_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    b.bstreams.push_back(s);
  }
  for (int k = 0; k < 2 * B + 1; k++) {
    hipEvent_t e;
    GENPK_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    b.events.push_back(e);
  }
  // eager pass: sizes every buffer, builds bin tables and counts
  for (const BatchJob &j : b.jobs) forward_job(b.plan_for(j), j);
  for (hipStream_t s : b.bstreams) GENPK_HIP(hipStreamSynchronize(s));
  GENPK_HIP(hipStreamSynchronize(b.stream));

  GENPK_HIP(hipStreamBeginCapture(b.cap, hipStreamCaptureModeThreadLocal));
  bool capturing = true;
  try {
    hipEvent_t fork = b.events[0];
    GENPK_HIP(hipEventRecord(fork, b.cap));
    for (int k = 0; k < B; k++) GENPK_HIP(hipStreamWaitEvent(b.bstreams[k], fork, 0));
    for (const BatchJob &j : b.jobs) forward_job(b.plan_for(j), j);
    for (int k = 0; k < B; k++) {
      hipEvent_t join = b.events[1 + k];
      GENPK_HIP(hipEventRecord(join, b.bstreams[k]));
      GENPK_HIP(hipStreamWaitEvent(b.cap, join, 0));
    }
    capturing = false;
    GENPK_HIP(hipStreamEndCapture(b.cap, &b.graph));
  } catch (...) {
    if (capturing) {
      hipGraph_t g = nullptr;
      (void)hipStreamEndCapture(b.cap, &g);
      if (g) (void)hipGraphDestroy(g);
    }
    throw;
  }
  GENPK_HIP(hipGraphInstantiate(&b.exec, b.graph, nullptr, nullptr, 0));
}

}  // namespace genpk

using namespace genpk;

struct genpk_batch {
  Batch b;
};

extern "C" {

int genpk_batch_create(genpk_batch **batch, int branches, void *stream) {
  return guarded([&] {
    GENPK_REQUIRE(batch != nullptr, "null batch pointer");
    GENPK_REQUIRE(branches >= 1 && branches <= 64, "branches must be in [1, 64]");
    auto h = std::make_unique<genpk_batch>();
    h->b.stream = (hipStream_t)stream;
    h->b.branches = branches;
    *batch = h.release();
  });
}

int genpk_batch_destroy(genpk_batch *batch) {
  return guarded([&] {
    if (!batch) return;
    GENPK_HIP(hipStreamSynchronize(batch->b.stream));
    delete batch;
  });
}

int genpk_batch_add(genpk_batch *batch, int64_t dims, int precision, double boxsize, int64_t n, const float *pos,
                    const float *masses, double mass, int nrbins, double total_mass, double *power, int *count,
                    double *keffs) {
  return guarded([&] {
    GENPK_REQUIRE(batch != nullptr, "null batch");
    GENPK_REQUIRE(batch->b.exec == nullptr, "batch already built");
    GENPK_REQUIRE(precision == GENPK_FLOAT32 || precision == GENPK_FLOAT64, "bad precision");
    GENPK_REQUIRE(dims >= 2 && dims <= 4096 && dims % 2 == 0 && fft_length_ok(dims) && fft_length_ok(dims / 2),
                  "dims must be even, in [2, 4096], with dims/2 = 2^a 3^b 5^c");
    GENPK_REQUIRE(nrbins >= 1, "nrbins must be positive");
    GENPK_REQUIRE(boxsize > 0, "boxsize must be positive");
    GENPK_REQUIRE(n >= 0 && (n == 0 || pos), "bad particles");
    GENPK_REQUIRE(total_mass != 0, "total mass must be nonzero");
    GENPK_REQUIRE(power != nullptr, "null power output");
    BatchJob j{dims, precision, boxsize, n, pos, masses, mass, nrbins, total_mass, power, count, keffs};
    batch->b.jobs.push_back(j);
  });
}

int genpk_batch_plan_options(genpk_batch *batch, int64_t dims, int precision, int tile_x, int tile_y, int tile_z,
                             int fft_variant) {
  return guarded([&] {
    GENPK_REQUIRE(batch != nullptr, "null batch");
    GENPK_REQUIRE(batch->b.exec == nullptr, "batch already built");
    GENPK_REQUIRE(precision == GENPK_FLOAT32 || precision == GENPK_FLOAT64, "bad precision");
    // the rules of genpk_plan_set_options
    auto pow2 = [](int v) { return v >= 1 && (v & (v - 1)) == 0; };
    GENPK_REQUIRE(pow2(tile_x) && pow2(tile_y) && pow2(tile_z) && tile_z >= 2,
                  "tile edges must be powers of two (z edge >= 2)");
    GENPK_REQUIRE((int64_t)(tile_x + 1) * (tile_y + 1) * (tile_z + 1) * 8 <= 160 * 1024,
                  "a halo'd tile of 64-bit accumulators must fit in LDS (160 KB)");
    GENPK_REQUIRE(fft_variant >= 0 && fft_variant < FFT_VARIANTS, "unknown FFT radix variant");
    batch->b.options[std::make_pair(dims, precision)] = {tile_x, tile_y, tile_z, fft_variant};
  });
}

int genpk_batch_build(genpk_batch *batch, const int *branch_of_job) {
  return guarded([&] {
    GENPK_REQUIRE(batch != nullptr, "null batch");
    batch_build(batch->b, branch_of_job);
  });
}

int genpk_batch_launch(genpk_batch *batch) {
  return guarded([&] {
    GENPK_REQUIRE(batch != nullptr && batch->b.exec != nullptr, "batch not built");
    GENPK_HIP(hipGraphLaunch(batch->b.exec, batch->b.stream));
  });
}

int genpk_batch_branch_of(const genpk_batch *batch, int64_t job, int *branch) {
  return guarded([&] {
    GENPK_REQUIRE(batch != nullptr && branch != nullptr, "null argument");
    GENPK_REQUIRE(job >= 0 && job < (int64_t)batch->b.jobs.size(), "job index out of range");
    *branch = batch->b.jobs[job].branch;
  });
}

}  // extern "C"
