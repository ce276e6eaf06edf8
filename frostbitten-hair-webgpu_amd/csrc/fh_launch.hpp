// Kernel launches of libfhhip.so, issued directly or recorded for batching.
//
// Every kernel is a struct K with
//   using Args = ...;                  // one POD argument struct
//   static constexpr uint32_t kThreads;
//   static constexpr const char* kName;
//   static __device__ void run(const Args& a, uint32_t block, uint32_t blocks);
// where `block` / `blocks` stand for blockIdx.x / gridDim.x. A plain launch
// runs K over one grid. A batched launch runs K for many hair instances at
// once: instance i owns the blocks [starts[i], starts[i + 1]) of one large
// grid, reads its own Args and sees its own block index and grid size, so the
// kernel body cannot tell the two apart.
//
// While a context records (fh_record_begin), launch<K> stores the kernel, its
// grid, its LDS bytes and a copy of its Args instead of launching; the
// recorded sequences of several contexts are then merged into one sequence of
// batched launches (fh_program_merge).
#pragma once

#include <type_traits>

#include "fh_internal.hpp"

namespace fh {

template <class K>
__global__ __launch_bounds__(K::kThreads) void k_plain(typename K::Args a) {
  K::run(a, blockIdx.x, gridDim.x);
}

// The argument structs and block ranges of a batched launch are read through
// the constant address space: the loads are scalar, like those of kernel
// arguments, and the pointers read from them are known to be global.
template <class T>
using const_as = const __attribute__((address_space(4))) T;

template <class K>
__global__ __launch_bounds__(K::kThreads) void k_batch(uint64_t argsAddr, uint64_t startsAddr, uint32_t n) {
#if defined(__HIP_DEVICE_COMPILE__)  // the host pass only needs the kernel's signature
  const_as<typename K::Args>* as = reinterpret_cast<const_as<typename K::Args>*>(argsAddr);
  const_as<uint32_t>* starts = reinterpret_cast<const_as<uint32_t>*>(startsAddr);
  const uint32_t b = blockIdx.x;
  uint32_t lo = 0, hi = n;  // the last instance whose first block is <= b
  while (hi - lo > 1u) {
    const uint32_t mid = (lo + hi) >> 1;
    if (starts[mid] <= b) lo = mid;
    else hi = mid;
  }
  const typename K::Args a = as[lo];
  K::run(a, b - starts[lo], starts[lo + 1] - starts[lo]);
#endif
}

// The device hands out workgroups round-robin over its 8 XCDs (each with its
// own L2): workgroup b runs on XCD b % 8. xcd_order(block, blocks) renumbers
// the blocks of a (plain or batched) grid so that the blocks on one XCD take
// consecutive numbers: a kernel whose neighbouring blocks touch neighbouring
// data (strands next to each other, hence screen tiles, grid cells and
// shadow texels next to each other) then keeps that data in one L2. It is a
// bijection on [0, blocks).
constexpr uint32_t kXcds = 8;

__device__ __forceinline__ uint32_t xcd_order(uint32_t block, uint32_t blocks) {
  const uint32_t o = (blockIdx.x - block) % kXcds;  // XCD of the grid's (instance's) first block
  const uint32_t v = (block + o) % kXcds;           // XCD of this block
  uint32_t before = 0;
  for (uint32_t x = 0; x < v; x++) {
    const uint32_t f = (x + kXcds - o) % kXcds;      // first block on XCD x
    if (f < blocks) before += (blocks - f + kXcds - 1u) / kXcds;
  }
  const uint32_t f = (v + kXcds - o) % kXcds;
  return before + (block - f) / kXcds;
}

// Dynamic LDS above the default limit must be allowed per kernel.
template <class Fn>
inline void allow_lds(Fn* fn, uint32_t shmem, uint32_t& allowed) {
  if (shmem > 64u * 1024u && shmem > allowed) {
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(fn), hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)shmem);
    allowed = shmem;
  }
}

template <class K>
int batch_launch(hipStream_t s, const void* dArgs, const uint32_t* dStarts, uint32_t nInst, uint32_t totalBlocks,
                 uint32_t shmem) {
  static uint32_t allowed = 0;
  allow_lds(&k_batch<K>, shmem, allowed);
  k_batch<K><<<totalBlocks, K::kThreads, shmem, s>>>(reinterpret_cast<uint64_t>(dArgs),
                                                     reinterpret_cast<uint64_t>(dStarts), nInst);
  return check_launch(K::kName);
}

int record_launch(fh_ctx* ctx, const char* name, BatchLauncher fn, uint32_t blocks, uint32_t shmem,
                  const void* args, size_t bytes);

// `name` labels the launch for kernel timing (a kernel shared by several
// stages is timed under each stage's name).
template <class K>
int launch(fh_ctx* ctx, uint32_t blocks, uint32_t shmem, const typename K::Args& a, const char* name = K::kName) {
  static_assert(std::is_trivially_copyable<typename K::Args>::value, "kernel arguments are copied as bytes");
  // a batched launch copies each instance's Args into registers: large
  // arrays belong in device memory, with a pointer here
  static_assert(sizeof(typename K::Args) <= 256, "kernel argument struct too large for a batched launch");
  // a clear batch holds clears and copies only: a kernel of another entry
  // point would run before the fills it depends on
  if (ctx->fill_batch) return fail(FH_ERR_INVALID_ARG, std::string(name) + ": only clears and copies inside a clear batch");
  if (blocks == 0) return FH_OK;
  if (ctx->recording) return record_launch(ctx, name, &batch_launch<K>, blocks, shmem, &a, sizeof(a));
  static uint32_t allowed = 0;
  allow_lds(&k_plain<K>, shmem, allowed);
  const int pk = prof_begin(ctx, name);
  k_plain<K><<<blocks, K::kThreads, shmem, ctx->stream>>>(a);
  prof_end(ctx, pk);
  return check_launch(K::kName);
}

}  // namespace fh
