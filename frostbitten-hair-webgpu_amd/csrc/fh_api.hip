// Context, memory and error handling of the C-ABI (include/fh_hair.h).
#include <algorithm>
#include <cstring>
#include <memory>

#include "fh_launch.hpp"

namespace {
thread_local std::string g_last_error;

// GPUCommandEncoder.clearBuffer as a kernel: graph-capture safe, unlike a
// memset node.
struct FillU32 {
  struct Args {
    uint32_t* dst;
    uint64_t n;
    uint32_t value;
  };
  static constexpr uint32_t kThreads = 256;
  static constexpr const char* kName = "fill_u32";
  static __device__ void run(const Args& a, uint32_t block, uint32_t blocks) {
    uint64_t i = (uint64_t)block * kThreads + threadIdx.x;
    const uint64_t stride = (uint64_t)blocks * kThreads;
    for (; i < a.n; i += stride) a.dst[i] = a.value;
  }
};

// Several clears in one launch: `per` consecutive blocks per buffer.
struct FillMany {
  struct Args {
    uint32_t* dst[fh::kMaxFillJobs];
    uint64_t n[fh::kMaxFillJobs];
    uint32_t per;
  };
  static constexpr uint32_t kThreads = 256;
  static constexpr const char* kName = "fill_many";
  static __device__ void run(const Args& a, uint32_t block, uint32_t blocks) {
    (void)blocks;
    const uint32_t j = block / a.per;
    uint32_t* __restrict__ dst = a.dst[j];
    const uint64_t n = a.n[j];
    uint64_t i = (uint64_t)(block - j * a.per) * kThreads + threadIdx.x;
    const uint64_t stride = (uint64_t)a.per * kThreads;
    for (; i < n; i += stride) dst[i] = 0u;
  }
};
// The jobs of a clear batch in one launch: `per` consecutive blocks per job;
// a job whose buffers are 16-byte aligned and whose length is a multiple of
// four words moves 16 bytes per access.
struct FillJobs {
  static constexpr int kMax = 6;
  struct Args {
    uint32_t* dst[kMax];
    const uint32_t* src[kMax];  // null: fill with value
    uint64_t n[kMax];           // u32 words
    uint32_t value[kMax];
    uint32_t vec;               // bit j: job j moves uint4
    uint32_t per;
  };
  static constexpr uint32_t kThreads = 256;
  static constexpr const char* kName = "clear_batch";
  static __device__ void run(const Args& a, uint32_t block, uint32_t blocks) {
    (void)blocks;
    const uint32_t j = block / a.per;
    const uint64_t stride = (uint64_t)a.per * kThreads;
    uint64_t i = (uint64_t)(block - j * a.per) * kThreads + threadIdx.x;
    // the job's fields by constant index: a run-time index into the argument
    // struct would copy it to per-lane scratch in the batched launch
    uint32_t* dst = nullptr;
    const uint32_t* src = nullptr;
    uint64_t n = 0;
    uint32_t v = 0;
#pragma unroll
    for (int q = 0; q < kMax; q++) {
      if ((uint32_t)q == j) {
        dst = a.dst[q];
        src = a.src[q];
        n = a.n[q];
        v = a.value[q];
      }
    }
    if ((a.vec >> j) & 1u) {
      uint4* __restrict__ d = reinterpret_cast<uint4*>(dst);
      const uint4* __restrict__ s = reinterpret_cast<const uint4*>(src);
      const uint64_t n4 = n / 4u;
      if (s) for (; i < n4; i += stride) d[i] = s[i];
      else for (; i < n4; i += stride) d[i] = make_uint4(v, v, v, v);
    } else {
      uint32_t* __restrict__ d = dst;
      const uint32_t* __restrict__ s = src;
      if (s) for (; i < n; i += stride) d[i] = s[i];
      else for (; i < n; i += stride) d[i] = v;
    }
  }
};
}  // namespace

namespace fh {

bool batch_fill(fh_ctx* ctx, void* dst, const void* src, uint64_t n_words, uint32_t value) {
  if (!ctx->fill_batch) return false;
  if (n_words) ctx->fill_jobs.push_back(FillJob{dst, src, n_words, value});
  return true;
}

void set_error(const std::string& msg) { g_last_error = msg; }

int fail(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

int check_launch(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(FH_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
  return FH_OK;
}

void add_deps(std::vector<ScratchDep>& out, const std::vector<ScratchDep>& deps) {
  for (const ScratchDep& d : deps) {
    bool seen = false;
    for (const ScratchDep& o : out) seen = seen || o.first == d.first;
    if (!seen) out.push_back(d);
  }
}

int check_deps(const std::vector<ScratchDep>& deps, const char* what) {
  for (const ScratchDep& d : deps)
    if (d.first->gen != d.second)
      return fail(FH_ERR_CAPACITY, std::string(what) +
                                       (d.first->gen == kGenGone
                                            ? ": a context whose scratch it addresses was destroyed"
                                            : ": a context's scratch was reallocated after the recording; record "
                                              "or capture again"));
  return FH_OK;
}

int ensure_scratch(fh_ctx* ctx, size_t bytes) {
  if (ctx->scratch_bytes >= bytes) return FH_OK;
  if (ctx->capturing)
    return fail(FH_ERR_UNSUPPORTED, "scratch would grow while capturing: run the calls once before capture");
  // launches recorded so far keep the old scratch address: fh_record_end
  // discards such a recording
  if (ctx->recording) ctx->rec_grew = true;
  if (ctx->scratch) {
    // every stream that may still read it: inside a branch the context's own
    // stream too (calls made while the context records or times kernels ran
    // a branch's scratch on it)
    FH_HIP(hipStreamSynchronize(ctx->stream));
    if (ctx->main_stream) FH_HIP(hipStreamSynchronize(ctx->main_stream));
    for (hipStream_t b : ctx->branch_stream)
      if (b) FH_HIP(hipStreamSynchronize(b));
    FH_HIP(hipFree(ctx->scratch));
    ctx->scratch = nullptr;
    ctx->scratch_bytes = 0;
  }
  size_t sz = align_up(bytes + bytes / 4, 1 << 20);
  ctx->sgen->gen++;  // programs and graphs recorded so far address the old scratch
  FH_HIP(hipMalloc(&ctx->scratch, sz));
  ctx->scratch_bytes = sz;
  return FH_OK;
}

int prof_begin(fh_ctx* ctx, const char* kernel) {
  if (ctx->prof_cap == 0) return -1;
  if (ctx->capturing) {
    // stream capture keeps no event records: fh_capture_end puts event-record
    // nodes around the captured kernel nodes, named in launch order
    ctx->cap_names.push_back(kernel);
    return -1;
  }
  if (ctx->prof_names.size() >= ctx->prof_cap) return -1;
  const int k = (int)ctx->prof_names.size();
  if (hipEventRecord(ctx->prof_events[2 * k], ctx->stream) != hipSuccess) return -1;
  ctx->prof_names.push_back(kernel);
  return k;
}

void prof_end(fh_ctx* ctx, int k) {
  if (k >= 0) (void)hipEventRecord(ctx->prof_events[2 * k + 1], ctx->stream);
}

// Puts an event-record node before and after each kernel node of a captured
// single-stream recording (a chain), so that replays can be timed. Leaves the
// graph untouched when its kernels do not match the recorded names.
int add_timing_nodes(fh_graph* g, const std::vector<const char*>& names) {
  size_t nn = 0;
  FH_HIP(hipGraphGetNodes(g->graph, nullptr, &nn));
  std::vector<hipGraphNode_t> nodes(nn);
  FH_HIP(hipGraphGetNodes(g->graph, nodes.data(), &nn));
  hipGraphNode_t cur = nullptr;
  int roots = 0;
  for (hipGraphNode_t x : nodes) {
    size_t nd = 0;
    FH_HIP(hipGraphNodeGetDependencies(x, nullptr, &nd));
    if (nd == 0) { cur = x; roots++; }
  }
  if (roots != 1) return FH_OK;
  std::vector<hipGraphNode_t> kernels;
  for (size_t steps = 0; cur && steps <= nn; steps++) {
    hipGraphNodeType t;
    FH_HIP(hipGraphNodeGetType(cur, &t));
    if (t == hipGraphNodeTypeKernel) kernels.push_back(cur);
    size_t no = 0;
    FH_HIP(hipGraphNodeGetDependentNodes(cur, nullptr, &no));
    if (no == 0) break;
    if (no != 1) return FH_OK;  // not a chain
    FH_HIP(hipGraphNodeGetDependentNodes(cur, &cur, &no));
  }
  if (kernels.size() != names.size()) return FH_OK;
  for (size_t i = 0; i < kernels.size(); i++) {
    hipGraphNode_t K = kernels[i];
    hipEvent_t e0, e1;
    FH_HIP(hipEventCreate(&e0));
    g->capEvents.push_back(e0);
    FH_HIP(hipEventCreate(&e1));
    g->capEvents.push_back(e1);
    size_t nd = 0, no = 0;
    FH_HIP(hipGraphNodeGetDependencies(K, nullptr, &nd));
    std::vector<hipGraphNode_t> deps(nd);
    if (nd) FH_HIP(hipGraphNodeGetDependencies(K, deps.data(), &nd));
    FH_HIP(hipGraphNodeGetDependentNodes(K, nullptr, &no));
    std::vector<hipGraphNode_t> outs(no);
    if (no) FH_HIP(hipGraphNodeGetDependentNodes(K, outs.data(), &no));
    for (hipGraphNode_t d : deps) FH_HIP(hipGraphRemoveDependencies(g->graph, &d, &K, 1));
    for (hipGraphNode_t o : outs) FH_HIP(hipGraphRemoveDependencies(g->graph, &K, &o, 1));
    hipGraphNode_t B, E;
    FH_HIP(hipGraphAddEventRecordNode(&B, g->graph, deps.data(), nd, e0));
    FH_HIP(hipGraphAddDependencies(g->graph, &B, &K, 1));
    FH_HIP(hipGraphAddEventRecordNode(&E, g->graph, &K, 1, e1));
    for (hipGraphNode_t o : outs) FH_HIP(hipGraphAddDependencies(g->graph, &E, &o, 1));
    g->timedBegin.push_back(B);
    g->timedEnd.push_back(E);
    g->timedNames.push_back(names[i]);
  }
  return FH_OK;
}

int zero_many(fh_ctx* ctx, int count, void* const* dst, const uint64_t* n_words, const char* kernel_name) {
  if (ctx->fill_batch) {
    for (int i = 0; i < count; i++) batch_fill(ctx, dst[i], nullptr, n_words[i], 0u);
    return FH_OK;
  }
  FillMany::Args jobs{};
  uint64_t most = 0;
  int k = 0;
  for (int i = 0; i < count && k < kMaxFillJobs; i++) {
    if (n_words[i] == 0) continue;
    jobs.dst[k] = (uint32_t*)dst[i];
    jobs.n[k] = n_words[i];
    most = std::max(most, n_words[i]);
    k++;
  }
  if (k == 0) return FH_OK;
  jobs.per = (uint32_t)std::min<uint64_t>((most + 255) / 256, 2048);
  return launch<FillMany>(ctx, jobs.per * k, 0, jobs, kernel_name);
}

int fill_u32(fh_ctx* ctx, void* dst, uint64_t n_words, uint32_t value) {
  if (n_words == 0 || batch_fill(ctx, dst, nullptr, n_words, value)) return FH_OK;
  const uint32_t blocks = (uint32_t)std::min<uint64_t>((n_words + 255) / 256, 4096);
  return launch<FillU32>(ctx, blocks, 0, FillU32::Args{(uint32_t*)dst, n_words, value});
}

int record_launch(fh_ctx* ctx, const char* name, BatchLauncher fn, uint32_t blocks, uint32_t shmem,
                  const void* args, size_t bytes) {
  RecLaunch r;
  r.name = name;
  r.fn = fn;
  r.blocks = blocks;
  r.shmem = shmem;
  r.args.assign(static_cast<const unsigned char*>(args), static_cast<const unsigned char*>(args) + bytes);
  ctx->rec.push_back(std::move(r));
  return FH_OK;
}

}  // namespace fh

extern "C" {

const char* fh_last_error(void) { return g_last_error.c_str(); }

int fh_abi_version(void) { return FH_ABI_VERSION; }

int fh_create(int device, fh_ctx** out) {
  FH_REQUIRE(out != nullptr, "fh_create: out is null");
  *out = nullptr;
  int count = 0;
  FH_HIP(hipGetDeviceCount(&count));
  FH_REQUIRE(device >= 0 && device < count, "fh_create: no such device");
  FH_HIP(hipSetDevice(device));
  fh_ctx* c = new fh_ctx();
  c->device = device;
  hipDeviceProp_t prop;
  hipError_t e = hipGetDeviceProperties(&prop, device);
  if (e != hipSuccess) {
    delete c;
    return fh::fail(FH_ERR_HIP, std::string("hipGetDeviceProperties: ") + hipGetErrorString(e));
  }
  c->num_cus = prop.multiProcessorCount;
  c->max_lds = prop.sharedMemPerBlock;
  if (int rc = fh::setup_noalloc_scratch(device)) {
    delete c;
    return rc;
  }
  e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete c;
    return fh::fail(FH_ERR_HIP, std::string("hipStreamCreate: ") + hipGetErrorString(e));
  }
  c->owns_stream = true;
  // the sRGB table, then the pass statistics words (fh_fine_exhausted_bins)
  e = hipMalloc(reinterpret_cast<void**>(&c->srgb_thresholds), 256 * sizeof(float) + 64 * sizeof(uint32_t));
  if (e == hipSuccess)
    e = hipMemcpy(c->srgb_thresholds, fh::srgb_thresholds(), 256 * sizeof(float), hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    c->stats = reinterpret_cast<uint32_t*>(c->srgb_thresholds + 256);
    e = hipMemset(c->stats, 0, 64 * sizeof(uint32_t));
  }
  if (e != hipSuccess) {
    if (c->srgb_thresholds) hipFree(c->srgb_thresholds);
    hipStreamDestroy(c->stream);
    delete c;
    return fh::fail(FH_ERR_HIP, std::string("fh_create: sRGB table: ") + hipGetErrorString(e));
  }
  *out = c;
  return FH_OK;
}

int fh_destroy(fh_ctx* ctx) {
  if (!ctx) return FH_OK;
  hipSetDevice(ctx->device);
  if (ctx->branch_open) fh_branch_end(ctx);
  if (ctx->stream) hipStreamSynchronize(ctx->stream);
  for (int l = 0; l < fh_ctx::kLanes; l++) {
    if (ctx->branch_stream[l]) {
      hipStreamSynchronize(ctx->branch_stream[l]);
      hipStreamDestroy(ctx->branch_stream[l]);
    }
    if (ctx->branch_fork[l]) hipEventDestroy(ctx->branch_fork[l]);
    if (ctx->branch_done[l]) hipEventDestroy(ctx->branch_done[l]);
  }
  for (hipEvent_t e : ctx->prof_events) hipEventDestroy(e);
  ctx->sgen->gen = fh::kGenGone;
  if (ctx->scratch) hipFree(ctx->scratch);
  for (void* b : ctx->branch_scratch)
    if (b) hipFree(b);
  if (ctx->srgb_thresholds) hipFree(ctx->srgb_thresholds);
  if (ctx->owns_stream && ctx->stream) hipStreamDestroy(ctx->stream);
  delete ctx;
  return FH_OK;
}

int fh_set_stream(fh_ctx* ctx, void* stream) {
  FH_CHECK_CTX(ctx);
  FH_REQUIRE(!ctx->branch_open && !ctx->any_pending(), "fh_set_stream: a branch is open or not joined");
  if (ctx->owns_stream && ctx->stream) {
    FH_HIP(hipStreamSynchronize(ctx->stream));
    FH_HIP(hipStreamDestroy(ctx->stream));
  }
  ctx->stream = (hipStream_t)stream;
  ctx->owns_stream = false;
  return FH_OK;
}

void* fh_get_stream(fh_ctx* ctx) { return ctx ? (void*)ctx->stream : nullptr; }

int fh_profiler_enable(fh_ctx* ctx, uint32_t max_launches) {
  FH_CHECK_CTX(ctx);
  FH_REQUIRE(max_launches <= (1u << 20), "fh_profiler_enable: max_launches must be <= 2^20");
  FH_REQUIRE(!ctx->branch_open && !ctx->any_pending(), "fh_profiler_enable: a branch is open or not joined");
  FH_HIP(hipStreamSynchronize(ctx->stream));
  ctx->prof_names.clear();
  while (ctx->prof_events.size() < 2ull * max_launches) {
    hipEvent_t e;
    FH_HIP(hipEventCreate(&e));
    ctx->prof_events.push_back(e);
  }
  ctx->prof_names.reserve(max_launches);
  ctx->prof_cap = max_launches;
  return FH_OK;
}

int fh_profiler_read(fh_ctx* ctx, fh_kernel_time* out, uint32_t max_out, uint32_t* n_out) {
  FH_CHECK_CTX(ctx);
  FH_REQUIRE(n_out != nullptr && (out != nullptr || max_out == 0), "fh_profiler_read: null output");
  FH_HIP(hipStreamSynchronize(ctx->stream));
  uint32_t n = 0;
  for (size_t k = 0; k < ctx->prof_names.size(); k++) {
    float ms = 0.0f;
    FH_HIP(hipEventElapsedTime(&ms, ctx->prof_events[2 * k], ctx->prof_events[2 * k + 1]));
    uint32_t i = 0;
    while (i < n && std::strncmp(out[i].name, ctx->prof_names[k], sizeof(out[i].name)) != 0) i++;
    if (i == n) {
      if (n == max_out) continue;
      std::memset(&out[n], 0, sizeof(out[n]));
      std::strncpy(out[n].name, ctx->prof_names[k], sizeof(out[n].name) - 1);
      n++;
    }
    out[i].launches++;
    out[i].total_ms += (double)ms;
  }
  ctx->prof_names.clear();
  *n_out = n;
  return FH_OK;
}

int fh_scratch_bytes(fh_ctx* ctx, uint64_t* bytes) {
  FH_CHECK_CTX(ctx);
  FH_REQUIRE(bytes != nullptr, "fh_scratch_bytes: null argument");
  *bytes = ctx->scratch_bytes + ctx->branch_scratch_bytes[0] + ctx->branch_scratch_bytes[1];
  return FH_OK;
}

int fh_clear_batch_begin(fh_ctx* ctx) {
  FH_CHECK_CTX(ctx);
  FH_REQUIRE(!ctx->fill_batch, "fh_clear_batch_begin: a clear batch is already open");
  ctx->fill_batch = true;
  ctx->fill_jobs.clear();
  return FH_OK;
}

int fh_clear_batch_end(fh_ctx* ctx) {
  FH_CHECK_CTX(ctx);
  FH_REQUIRE(ctx->fill_batch, "fh_clear_batch_end: no clear batch is open");
  ctx->fill_batch = false;
  std::vector<fh::FillJob> jobs;
  jobs.swap(ctx->fill_jobs);
  for (size_t k = 0; k < jobs.size(); k += FillJobs::kMax) {
    FillJobs::Args a{};
    const int m = (int)std::min<size_t>(FillJobs::kMax, jobs.size() - k);
    uint64_t most = 0;
    for (int j = 0; j < m; j++) {
      const fh::FillJob& f = jobs[k + j];
      a.dst[j] = static_cast<uint32_t*>(f.dst);
      a.src[j] = static_cast<const uint32_t*>(f.src);
      a.n[j] = f.n;
      a.value[j] = f.value;
      const bool vec = f.n % 4 == 0 && reinterpret_cast<uintptr_t>(f.dst) % 16 == 0 &&
                       reinterpret_cast<uintptr_t>(f.src) % 16 == 0;
      if (vec) a.vec |= 1u << j;
      most = std::max(most, vec ? f.n / 4 : f.n);
    }
    a.per = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>((most + 255) / 256, 2048));
    int rc = fh::launch<FillJobs>(ctx, a.per * (uint32_t)m, 0, a);
    if (rc) return rc;
  }
  return FH_OK;
}

int fh_branch_begin_lane(fh_ctx* ctx, uint32_t lane) {
  FH_CHECK_CTX(ctx);
  FH_REQUIRE(lane < (uint32_t)fh_ctx::kLanes, "fh_branch_begin_lane: lane must be 0 or 1");
  FH_REQUIRE(!ctx->branch_open, "fh_branch_begin: a branch is already open");
  FH_REQUIRE(!ctx->fill_batch, "fh_branch_begin: a clear batch is open");
  const bool live = !ctx->recording && ctx->prof_cap == 0;
  if (live) {
    if (!ctx->branch_stream[lane]) {
      FH_HIP(hipSetDevice(ctx->device));
      FH_HIP(hipStreamCreateWithFlags(&ctx->branch_stream[lane], hipStreamNonBlocking));
      FH_HIP(hipEventCreateWithFlags(&ctx->branch_fork[lane], hipEventDisableTiming));
      FH_HIP(hipEventCreateWithFlags(&ctx->branch_done[lane], hipEventDisableTiming));
    }
    FH_HIP(hipEventRecord(ctx->branch_fork[lane], ctx->stream));
    FH_HIP(hipStreamWaitEvent(ctx->branch_stream[lane], ctx->branch_fork[lane], 0));
    ctx->main_stream = ctx->stream;
    ctx->stream = ctx->branch_stream[lane];
  }
  std::swap(ctx->scratch, ctx->branch_scratch[lane]);
  std::swap(ctx->scratch_bytes, ctx->branch_scratch_bytes[lane]);
  ctx->branch_open = true;
  ctx->branch_lane = (int)lane;
  ctx->branch_live = live;
  return FH_OK;
}

int fh_branch_begin(fh_ctx* ctx) { return fh_branch_begin_lane(ctx, 0); }

int fh_branch_end(fh_ctx* ctx) {
  FH_CHECK_CTX(ctx);
  FH_REQUIRE(ctx->branch_open, "fh_branch_end: no branch is open");
  FH_REQUIRE(!ctx->fill_batch, "fh_branch_end: a clear batch is open");
  const int lane = ctx->branch_lane;
  std::swap(ctx->scratch, ctx->branch_scratch[lane]);
  std::swap(ctx->scratch_bytes, ctx->branch_scratch_bytes[lane]);
  ctx->branch_open = false;
  if (!ctx->branch_live) return FH_OK;
  ctx->branch_live = false;
  ctx->stream = ctx->main_stream;
  ctx->main_stream = nullptr;
  ctx->branch_pending[lane] = true;
  FH_HIP(hipEventRecord(ctx->branch_done[lane], ctx->branch_stream[lane]));
  return FH_OK;
}

int fh_branch_join(fh_ctx* ctx) {
  FH_CHECK_CTX(ctx);
  FH_REQUIRE(!ctx->branch_open, "fh_branch_join: the branch is still open");
  FH_REQUIRE(!ctx->fill_batch, "fh_branch_join: a clear batch is open");
  for (int l = 0; l < fh_ctx::kLanes; l++) {
    if (!ctx->branch_pending[l]) continue;
    ctx->branch_pending[l] = false;
    FH_HIP(hipStreamWaitEvent(ctx->stream, ctx->branch_done[l], 0));
  }
  return FH_OK;
}

int fh_scratch_reserve(fh_ctx* ctx, uint64_t bytes) {
  FH_CHECK_CTX(ctx);
  FH_REQUIRE(!ctx->recording && !ctx->capturing, "fh_scratch_reserve: the context is recording or capturing");
  return fh::ensure_scratch(ctx, (size_t)bytes);
}

int fh_set_raster_band_cap(fh_ctx* ctx, uint32_t bands) {
  FH_CHECK_CTX(ctx);
  FH_REQUIRE(bands <= (1u << 20), "fh_set_raster_band_cap: at most 2^20 bands");
  ctx->band_cap = bands;
  return FH_OK;
}

int fh_set_persistent_blocks(fh_ctx* ctx, uint32_t blocks) {
  FH_CHECK_CTX(ctx);
  ctx->persist_blocks = blocks;
  return FH_OK;
}

int fh_record_begin(fh_ctx* ctx) {
  FH_CHECK_CTX(ctx);
  FH_REQUIRE(!ctx->recording && !ctx->capturing, "fh_record_begin: already recording or capturing");
  FH_REQUIRE(!ctx->branch_open, "fh_record_begin: a branch is open");
  ctx->recording = true;
  ctx->rec_grew = false;
  ctx->rec.clear();
  return FH_OK;
}

int fh_record_end(fh_ctx* ctx, fh_program** out) {
  FH_CHECK_CTX(ctx);
  FH_REQUIRE(ctx->recording, "fh_record_end: not recording");
  FH_REQUIRE(!ctx->fill_batch, "fh_record_end: a clear batch is open");
  FH_REQUIRE(out != nullptr, "fh_record_end: out is null");
  ctx->recording = false;
  *out = nullptr;
  if (ctx->rec_grew) {
    ctx->rec.clear();
    return fh::fail(FH_ERR_CAPACITY, "fh_record_end: the context's scratch grew while recording; record again");
  }
  fh_program* p = new fh_program();
  p->rec.swap(ctx->rec);
  p->instances = 1;
  p->deps.emplace_back(ctx->sgen, ctx->sgen->gen);
  *out = p;
  return FH_OK;
}

int fh_program_merge(fh_ctx* ctx, const fh_program* const* programs, uint32_t count, fh_program** out) {
  FH_CHECK_CTX(ctx);
  FH_REQUIRE(programs && out && count > 0, "fh_program_merge: null argument");
  for (uint32_t i = 0; i < count; i++)
    FH_REQUIRE(programs[i] && !programs[i]->merged, "fh_program_merge: inputs must be recorded programs");
  // Greedy common supersequence of the launch sequences: among the kernels
  // at the heads, take the one whose instances have the most launches left
  // (ties: the kernel more heads share), launch it for all those instances.
  std::vector<size_t> at(count, 0);
  std::vector<unsigned char> blob;
  auto align = [&](size_t a) { blob.resize(fh::align_up(blob.size(), a)); };
  std::unique_ptr<fh_program> m(new fh_program());
  m->merged = true;
  m->instances = count;
  while (true) {
    fh::BatchLauncher best = nullptr;
    size_t bestLeft = 0, bestShare = 0;
    for (uint32_t i = 0; i < count; i++) {
      if (at[i] >= programs[i]->rec.size()) continue;
      const fh::BatchLauncher fn = programs[i]->rec[at[i]].fn;
      size_t left = 0, share = 0;
      for (uint32_t j = 0; j < count; j++) {
        if (at[j] >= programs[j]->rec.size() || programs[j]->rec[at[j]].fn != fn) continue;
        left = std::max(left, programs[j]->rec.size() - at[j]);
        share++;
      }
      if (!best || left > bestLeft || (left == bestLeft && share > bestShare)) {
        best = fn;
        bestLeft = left;
        bestShare = share;
      }
    }
    if (!best) break;
    fh_program::Step st{};
    std::vector<uint32_t> starts(1, 0u);
    std::vector<const fh::RecLaunch*> members;
    uint64_t total = 0;
    for (uint32_t j = 0; j < count; j++) {
      if (at[j] >= programs[j]->rec.size() || programs[j]->rec[at[j]].fn != best) continue;
      const fh::RecLaunch& r = programs[j]->rec[at[j]++];
      if (!members.empty() && r.args.size() != members[0]->args.size())
        return fh::fail(FH_ERR_INVALID_ARG, std::string("fh_program_merge: argument sizes differ for ") + r.name);
      members.push_back(&r);
      fh::add_deps(m->deps, programs[j]->deps);
      total += r.blocks;
      st.shmem = std::max(st.shmem, r.shmem);
      starts.push_back((uint32_t)total);
    }
    if (total >= (1ull << 31)) return fh::fail(FH_ERR_INVALID_ARG, "fh_program_merge: too many blocks in one launch");
    st.name = members[0]->name;
    st.fn = best;
    st.nInst = (uint32_t)members.size();
    st.totalBlocks = (uint32_t)total;
    align(256);
    st.argsOff = blob.size();
    for (const fh::RecLaunch* r : members) blob.insert(blob.end(), r->args.begin(), r->args.end());
    align(256);
    st.startsOff = blob.size();
    const unsigned char* sb = reinterpret_cast<const unsigned char*>(starts.data());
    blob.insert(blob.end(), sb, sb + starts.size() * sizeof(uint32_t));
    m->steps.push_back(st);
  }
  if (!blob.empty()) {
    FH_HIP(hipMalloc(&m->dmem, blob.size()));
    FH_HIP(hipMemcpy(m->dmem, blob.data(), blob.size(), hipMemcpyHostToDevice));
  }
  *out = m.release();
  return FH_OK;
}

int fh_program_launches(const fh_program* program, uint32_t* n_out) {
  FH_REQUIRE(program && n_out, "fh_program_launches: null argument");
  *n_out = (uint32_t)(program->merged ? program->steps.size() : program->rec.size());
  return FH_OK;
}

int fh_program_launch(fh_ctx* ctx, const fh_program* program) {
  FH_CHECK_CTX(ctx);
  FH_REQUIRE(program && program->merged, "fh_program_launch: not a merged program");
  FH_REQUIRE(!ctx->recording, "fh_program_launch: the context is recording");
  if (int rc = fh::check_deps(program->deps, "fh_program_launch")) return rc;
  if (ctx->capturing) fh::add_deps(ctx->cap_deps, program->deps);
  for (const fh_program::Step& st : program->steps) {
    const char* base = static_cast<const char*>(program->dmem);
    const int pk = fh::prof_begin(ctx, st.name);
    const int rc = st.fn(ctx->stream, base + st.argsOff, reinterpret_cast<const uint32_t*>(base + st.startsOff),
                         st.nInst, st.totalBlocks, st.shmem);
    fh::prof_end(ctx, pk);
    if (rc) return rc;
  }
  return FH_OK;
}

int fh_program_destroy(fh_program* program) {
  if (!program) return FH_OK;
  if (program->dmem) FH_HIP(hipFree(program->dmem));
  delete program;
  return FH_OK;
}

int fh_capture_begin(fh_ctx* ctx) {
  FH_CHECK_CTX(ctx);
  FH_REQUIRE(!ctx->capturing, "fh_capture_begin: already capturing");
  FH_REQUIRE(!ctx->branch_open && !ctx->any_pending(), "fh_capture_begin: a branch is open or not joined");
  FH_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeRelaxed));
  ctx->capturing = true;
  ctx->cap_names.clear();
  ctx->cap_deps.clear();
  ctx->cap_deps.emplace_back(ctx->sgen, ctx->sgen->gen);
  return FH_OK;
}

int fh_capture_end(fh_ctx* ctx, fh_graph** out) {
  FH_CHECK_CTX(ctx);
  FH_REQUIRE(ctx->capturing, "fh_capture_end: not capturing");
  FH_REQUIRE(out != nullptr, "fh_capture_end: out is null");
  *out = nullptr;
  // a capture must end on its origin stream with every branch joined into it
  if (ctx->fill_batch) fh_clear_batch_end(ctx);
  const bool was_open = ctx->branch_open;
  if (was_open) fh_branch_end(ctx);
  fh_branch_join(ctx);
  ctx->capturing = false;
  fh_graph* g = new fh_graph();
  hipError_t e = hipStreamEndCapture(ctx->stream, &g->graph);
  int rc = was_open ? fh::fail(FH_ERR_INVALID_ARG, "a branch was still open") : FH_OK;
  if (rc == FH_OK && e == hipSuccess && !ctx->cap_names.empty()) rc = fh::add_timing_nodes(g, ctx->cap_names);
  ctx->cap_names.clear();
  g->deps.swap(ctx->cap_deps);
  if (e == hipSuccess && rc == FH_OK) e = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0);
  if (e != hipSuccess || rc != FH_OK) {
    const std::string msg = rc != FH_OK ? fh_last_error() : hipGetErrorString(e);
    fh_graph_destroy(g);
    return fh::fail(FH_ERR_HIP, std::string("fh_capture_end: ") + msg);
  }
  *out = g;
  return FH_OK;
}

int fh_graph_launch(fh_ctx* ctx, fh_graph* graph) {
  FH_CHECK_CTX(ctx);
  FH_REQUIRE(graph && graph->exec, "fh_graph_launch: null graph");
  if (int rc = fh::check_deps(graph->deps, "fh_graph_launch")) return rc;
  // timed kernels: bind this replay's event pairs from the profiler's pool, or
  // the capture's own events once the pool is used up (never read back)
  for (size_t k = 0; k < graph->timedNames.size(); k++) {
    if (!graph->timedBegin[k] || !graph->timedEnd[k]) continue;
    hipEvent_t e0 = graph->capEvents[2 * k], e1 = graph->capEvents[2 * k + 1];
    const bool timed = ctx->prof_names.size() < ctx->prof_cap;
    if (timed) {
      const size_t j = ctx->prof_names.size();
      e0 = ctx->prof_events[2 * j];
      e1 = ctx->prof_events[2 * j + 1];
    }
    FH_HIP(hipGraphExecEventRecordNodeSetEvent(graph->exec, graph->timedBegin[k], e0));
    FH_HIP(hipGraphExecEventRecordNodeSetEvent(graph->exec, graph->timedEnd[k], e1));
    if (timed) ctx->prof_names.push_back(graph->timedNames[k]);
  }
  FH_HIP(hipGraphLaunch(graph->exec, ctx->stream));
  return FH_OK;
}

int fh_graph_destroy(fh_graph* graph) {
  if (!graph) return FH_OK;
  if (graph->exec) hipGraphExecDestroy(graph->exec);
  if (graph->graph) hipGraphDestroy(graph->graph);
  for (hipEvent_t e : graph->capEvents) hipEventDestroy(e);
  delete graph;
  return FH_OK;
}

int fh_sync(fh_ctx* ctx) {
  FH_CHECK_CTX(ctx);
  FH_REQUIRE(!ctx->fill_batch, "fh_sync: a clear batch is open");
  FH_HIP(hipStreamSynchronize(ctx->stream));
  for (hipStream_t b : ctx->branch_stream)
    if (b) FH_HIP(hipStreamSynchronize(b));
  return FH_OK;
}

int fh_alloc(fh_ctx* ctx, uint64_t bytes, void** out) {
  FH_CHECK_CTX(ctx);
  FH_REQUIRE(out != nullptr, "fh_alloc: out is null");
  FH_HIP(hipSetDevice(ctx->device));
  FH_HIP(hipMalloc(out, bytes ? bytes : 4));
  return FH_OK;
}

int fh_free(fh_ctx* ctx, void* ptr) {
  FH_CHECK_CTX(ctx);
  if (ptr) FH_HIP(hipFree(ptr));
  return FH_OK;
}

int fh_upload(fh_ctx* ctx, void* dst, const void* src, uint64_t bytes) {
  FH_CHECK_CTX(ctx);
  FH_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
  FH_HIP(hipStreamSynchronize(ctx->stream));
  return FH_OK;
}

int fh_download(fh_ctx* ctx, void* dst, const void* src, uint64_t bytes) {
  FH_CHECK_CTX(ctx);
  FH_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
  FH_HIP(hipStreamSynchronize(ctx->stream));
  return FH_OK;
}

int fh_clear(fh_ctx* ctx, void* dst, uint64_t bytes) {
  FH_CHECK_CTX(ctx);
  FH_REQUIRE(bytes % 4 == 0, "fh_clear: bytes must be a multiple of 4");
  FH_REQUIRE(dst != nullptr || bytes == 0, "fh_clear: null buffer");
  return fh::fill_u32(ctx, dst, bytes / 4, 0u);
}

uint32_t fh_tile_count(const fh_render_cfg* cfg, uint32_t vp_w, uint32_t vp_h) {
  if (!cfg || cfg->tileSize == 0) return 0;
  return fh::div_ceil(vp_w, cfg->tileSize) * fh::div_ceil(vp_h, cfg->tileSize);
}

uint32_t fh_segments_buffer_entries(const fh_render_cfg* cfg, uint32_t vp_w, uint32_t vp_h,
                                    float avg_segments_per_tile) {
  double tiles = (double)fh_tile_count(cfg, vp_w, vp_h);
  double cnt = tiles * (double)avg_segments_per_tile;
  uint64_t c = (uint64_t)cnt;
  if ((double)c < cnt) c++;
  return (uint32_t)(1 + c);
}

uint32_t fh_slice_data_per_processor(const fh_render_cfg* cfg) {
  if (!cfg) return 0;
  return cfg->avgFragmentsPerSlice * cfg->slicesPerPixel * cfg->tileSize * cfg->tileSize;
}

}  // extern "C"
