// Host-side internals shared by the translation units of libfhhip.so.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <memory>
#include <string>
#include <utility>
#include <vector>

#include "../../include/fh_hair.h"

namespace fh {
// Launches a kernel for many instances at once (fh_launch.hpp): instance i
// owns blocks [dStarts[i], dStarts[i + 1]) and reads the i-th argument struct.
using BatchLauncher = int (*)(hipStream_t s, const void* dArgs, const uint32_t* dStarts, uint32_t nInst,
                              uint32_t totalBlocks, uint32_t shmem);
// One clear or copy collected by an open clear batch (fh_clear_batch_begin):
// n u32 words of dst set to `value`, or copied from src when src is set.
struct FillJob {
  void* dst;
  const void* src;
  uint64_t n;
  uint32_t value;
};
// Generation of a context's scratch: bumped whenever the scratch (or the
// branch scratch) is reallocated, and set to kGenGone when the context is
// destroyed. Recorded programs and captured graphs keep the generation of
// every context whose scratch their launches address, and refuse to run
// once it has changed. The cell outlives the context (shared ownership).
struct ScratchGen {
  uint64_t gen = 0;
};
constexpr uint64_t kGenGone = ~0ull;
using ScratchDep = std::pair<std::shared_ptr<ScratchGen>, uint64_t>;
// Appends `deps` to `out`, once per cell.
void add_deps(std::vector<ScratchDep>& out, const std::vector<ScratchDep>& deps);
// FH_OK, or FH_ERR_CAPACITY naming `what` when a dependency's scratch changed.
int check_deps(const std::vector<ScratchDep>& deps, const char* what);
// Looks up, on `device` (current), the device address of the fixed scratch of
// fh_grid_post_sim_noalloc; called once per device by fh_create.
int setup_noalloc_scratch(int device);
// One launch kept by a recording context instead of being issued.
struct RecLaunch {
  const char* name;
  BatchLauncher fn;  // also names the kernel: equal for launches of one kernel
  uint32_t blocks, shmem;
  std::vector<unsigned char> args;
};
}  // namespace fh

struct fh_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  bool owns_stream = false;
  int num_cus = 0;
  // Workgroups of a grid-stride ("persistent") launch; 0: 8 per CU.
  uint32_t persist_blocks = 0;
  // Bands of large triangles the software rasterizer queues at most (0: the
  // default, 2^20); the triangles whose bands do not fit draw themselves.
  uint32_t band_cap = 0;
  size_t max_lds = 64 * 1024;  // LDS bytes one workgroup may allocate
  // Grow-on-demand scratch owned by the context (fine-pass gather lists,
  // per-bin offsets, prefix-scan temporaries).
  void* scratch = nullptr;
  size_t scratch_bytes = 0;
  std::shared_ptr<fh::ScratchGen> sgen = std::make_shared<fh::ScratchGen>();
  // Branch (fh_branch_begin .. fh_branch_end): calls that may overlap the
  // main stream's later work run on a second stream, ordered after the main
  // stream's work at fh_branch_begin and joined back by fh_branch_join. A
  // branch has its own scratch, swapped into `scratch` while it is open, so
  // the two streams never share temporaries. While recording or timing
  // kernels the calls stay on the main stream (the scratch is still swapped).
  // Two branch lanes (fh_branch_begin: lane 0, fh_branch_begin_lane), each
  // with its own stream, events and scratch; at most one is open at a time,
  // and both may be pending (running) beside the main stream.
  static constexpr int kLanes = 2;
  hipStream_t branch_stream[kLanes] = {nullptr, nullptr};
  hipStream_t main_stream = nullptr;  // the context's stream while a branch runs live
  hipEvent_t branch_fork[kLanes] = {nullptr, nullptr}, branch_done[kLanes] = {nullptr, nullptr};
  void* branch_scratch[kLanes] = {nullptr, nullptr};
  size_t branch_scratch_bytes[kLanes] = {0, 0};
  bool branch_open = false;                     // between fh_branch_begin* and fh_branch_end
  int branch_lane = 0;                          // the open branch's lane
  bool branch_live = false;                     // the open branch runs on its lane's stream
  bool branch_pending[kLanes] = {false, false};  // lane work not yet joined into the main stream
  bool any_pending() const { return branch_pending[0] || branch_pending[1]; }
  // Between fh_clear_batch_begin and _end the clears and copies of the entry
  // points are collected here and issued as one launch by _end.
  bool fill_batch = false;
  std::vector<fh::FillJob> fill_jobs;
  // The 256 thresholds of the rgba8unorm-srgb store (fh_present), uploaded
  // at creation so that no call allocates or copies them later.
  float* srgb_thresholds = nullptr;
  // Device words the passes count into (cleared by the pass that counts):
  // [0] bins whose slice memory the last fine pass used up.
  uint32_t* stats = nullptr;
  // Kernel timing (fh_profiler_enable): an event pair per timed launch.
  std::vector<hipEvent_t> prof_events;
  std::vector<const char*> prof_names;  // kernel of each timed launch
  uint32_t prof_cap = 0;                // 0: timing off
  bool capturing = false;               // between fh_capture_begin and _end
  // Kernels launched while capturing with timing on, in launch order: the
  // graph gets event-record nodes around them, rebound on every replay.
  std::vector<const char*> cap_names;
  // Scratch generations that the calls captured so far depend on (this
  // context's, and those of the merged programs launched in the capture).
  std::vector<fh::ScratchDep> cap_deps;
  // Between fh_record_begin and _end the launches are kept here, not issued.
  bool recording = false;
  bool rec_grew = false;  // the scratch was reallocated while recording
  std::vector<fh::RecLaunch> rec;
};

// A recorded launch sequence of one context, or several merged into batched
// launches whose arguments and block ranges live in device memory.
struct fh_program {
  std::vector<fh::RecLaunch> rec;
  struct Step {
    const char* name;
    fh::BatchLauncher fn;
    uint32_t nInst, totalBlocks, shmem;
    size_t argsOff, startsOff;  // byte offsets into dmem
  };
  std::vector<Step> steps;
  void* dmem = nullptr;
  uint32_t instances = 0;
  bool merged = false;
  std::vector<fh::ScratchDep> deps;  // scratch generations of the recording contexts
};

struct fh_graph {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  std::vector<hipGraphNode_t> timedBegin, timedEnd;  // event-record nodes per timed kernel
  std::vector<const char*> timedNames;
  std::vector<hipEvent_t> capEvents;                  // owned: the nodes' own events (never read)
  std::vector<fh::ScratchDep> deps;                   // scratch generations the replay addresses
};

namespace fh {

void set_error(const std::string& msg);
int fail(int code, const std::string& msg);
int check_launch(const char* what);
// Ensure ctx->scratch holds at least `bytes`; frees and reallocates if not.
int ensure_scratch(fh_ctx* ctx, size_t bytes);
// Kernel timing: prof_begin records the start event of the next launch on
// the context's stream and returns its index (-1 when timing is off or the
// event pool is used up); prof_end records the matching end event.
int prof_begin(fh_ctx* ctx, const char* kernel);
// The host copy of the sRGB store thresholds (fh_post.hip).
const float* srgb_thresholds();
// Zeroes up to kMaxFillJobs buffers (n_words u32 each) in one launch.
constexpr int kMaxFillJobs = 4;
int zero_many(fh_ctx* ctx, int count, void* const* dst, const uint64_t* n_words,
              const char* kernel_name = "fill_many");
// While a clear batch is open: keeps the job and returns true.
bool batch_fill(fh_ctx* ctx, void* dst, const void* src, uint64_t n_words, uint32_t value);
void prof_end(fh_ctx* ctx, int k);

inline uint32_t div_ceil(uint32_t a, uint32_t b) { return (a + b - 1u) / b; }
inline uint32_t persistent_blocks(const fh_ctx* ctx) {
  return ctx->persist_blocks ? ctx->persist_blocks : 8u * (uint32_t)(ctx->num_cus > 0 ? ctx->num_cus : 1);
}
inline size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

}  // namespace fh

#define FH_CHECK_CTX(ctx)                                              \
  do {                                                                 \
    if (!(ctx)) return fh::fail(FH_ERR_INVALID_ARG, "null fh_ctx");    \
  } while (0)

#define FH_REQUIRE(cond, msg)                                          \
  do {                                                                 \
    if (!(cond)) return fh::fail(FH_ERR_INVALID_ARG, (msg));           \
  } while (0)

#define FH_HIP(call)                                                   \
  do {                                                                 \
    hipError_t e_ = (call);                                            \
    if (e_ != hipSuccess)                                              \
      return fh::fail(FH_ERR_HIP, std::string(#call) + ": " +          \
                                      hipGetErrorString(e_));          \
  } while (0)
