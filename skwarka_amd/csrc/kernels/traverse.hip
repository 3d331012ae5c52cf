// Single-scene closest-hit and occlusion kernels on gfx950 (the walk itself is in
// trace_device.hpp).
//
// Scheduling: one ray per lane, 64-lane workgroups, a persistent grid sized to the chip. A wave
// iteration has three phases: refill (idle lanes load a new ray and test the root), leaf (the
// triangles of the leaves the lanes hold, spread over all 64 lanes) and advance (every other busy
// lane pops or descends one internal node). Refill and leaf phases are
// the expensive, rarely needed ones, so a wave defers them until refill_min / leaf_min lanes wait
// for them (or no lane could move otherwise): each then runs for many lanes at once instead of
// for one or two lanes per iteration. Leaves are no longer deferred once the rays run out.
// Idle lanes take the next unstarted rays of the wave's pool (ballot + rank); when the pool runs
// short, one lane claims the next `claim` rays (at least 64) with a single atomicAdd on the launch
// counter. One device-wide counter serves about one atomic per 11 ns, so claims are kept rare.
// The first gridDim*blockDim rays are handed out statically; the counter starts at that value.
// Rays are 64-byte records read as four 16-byte loads.
#include <hip/hip_runtime.h>

#include "kernels.hpp"
#include "trace_device.hpp"

namespace skw {
namespace {

using namespace trace;

// STATS: count node boxes and triangles tested per ray (the oracle counts the same events).
template <bool ANY, bool INST, bool STATS, class Stack>
__device__ void run_lane(const TraceArgs &args, const Stack &stk, uint8_t *slots) {
  const uint32_t lane = threadIdx.x & (kWave - 1);
  const uint64_t n = args.n_dev ? *args.n_dev : args.n;
  uint64_t idx = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  bool busy = false;
  bool first = true;
  unsigned long long sum_pops = 0, sum_tests = 0;
  Walker<ANY, INST, STATS, Stack> w;
  w.set_root(args.nodes, args.root);
  w.tris = args.tris;
  w.insts = args.insts;
  w.inst_range = args.inst_range;
  w.stk = stk;
  // The wave's pool of claimed, unstarted rays [pool, pool_end): the same value in every lane.
  uint64_t pool = 0, pool_end = 0;
  bool drained = false;  // some lane found no ray left (the same value in every lane)
  // Every lane stays in the loop until the wave is done (the pooled leaf tests read other lanes'
  // registers); a lane without a ray left stays idle.
  while (true) {
    // Refill: idle lanes take new rays once at least refill_min of them wait (or nothing else can
    // move), so the ray load and root test run for many lanes at a time.
    const uint64_t idle = __ballot(!busy);
    if (idle && !drained) {
      const uint64_t others = __ballot(busy);
      if (first || (uint64_t)__popcll(idle) >= args.refill_min || others == 0) {
        if (!first) {
          const uint64_t n_idle = (uint64_t)__popcll(idle);
          const uint64_t rank = (uint64_t)__popcll(idle & ((1ull << lane) - 1ull));
          const uint64_t avail = pool_end - pool;
          if (avail >= n_idle) {
            if (!busy) idx = pool + rank;
            pool += n_idle;
          } else {
            // Idle lanes take the rest of the pool, then the start of a newly claimed chunk.
            const int leader = __ffsll((unsigned long long)idle) - 1;
            unsigned long long base = 0;
            if ((int)lane == leader)
              base = atomicAdd(reinterpret_cast<unsigned long long *>(args.counter), (unsigned long long)args.claim);
            base = __shfl(base, leader);
            if (!busy) idx = rank < avail ? pool + rank : base + (rank - avail);
            pool = base + (n_idle - avail);
            pool_end = base + args.claim;
          }
        }
        // Once the launch's rays run out, deferring leaves only delays the last walks.
        drained = __ballot(!busy && idx >= n) != 0;
        if (!busy && idx < n) {
          RayState r;
          double tmax;
          load_ray(args.rays, idx, r, tmax);
          w.start(r, tmax);
          busy = true;
        }
      }
    }
    if (__ballot(busy) == 0) break;  // nothing to walk and (the refill ran) no ray left
    first = false;
    // Leaves: lanes holding an accepted leaf test it together once at least leaf_min of them
    // hold one (or no other lane can move).
    const uint64_t held = __ballot(busy && w.at_leaf);
    if (held) {
      const uint64_t moving = __ballot(busy && !w.at_leaf);
      if ((uint64_t)__popcll(held) >= args.leaf_min || moving == 0 || drained) w.leaves_pooled(busy && w.at_leaf, slots);
    }
    // Between leaves: one move per lane; the NaN-free slab test (and pops without the re-test of
    // an internal node) when the tree's boxes nest and every busy lane's ray allows it.
    const bool fast = args.nested && __ballot(busy && !w.nan_free) == 0;
    if (busy && !w.at_leaf) {
      const bool done = fast ? w.template advance<true>() : w.template advance<false>();
      if (done) {
        if (ANY)
          args.occluded[idx] = w.found ? 1 : 0;
        else
          store_hit(args.hits + idx, args.tris, w.found, w.best, w.tmax, w.r);
        if (STATS) {
          if (args.pops) args.pops[idx] = w.pops;
          if (args.tests) args.tests[idx] = w.tests;
          sum_pops += w.pops;
          sum_tests += w.tests;
        }
        busy = false;
      }
    }
  }
  if (STATS && args.sums) {
    // all lanes have left the loop: one atomic per wave
    for (int off = 32; off > 0; off >>= 1) {
      sum_pops += __shfl_xor(sum_pops, off);
      sum_tests += __shfl_xor(sum_tests, off);
    }
    if (lane == 0) {
      atomicAdd(reinterpret_cast<unsigned long long *>(args.sums), (unsigned long long)sum_pops);
      atomicAdd(reinterpret_cast<unsigned long long *>(args.sums + 1), (unsigned long long)sum_tests);
    }
  }
}

template <bool ANY, bool INST, bool STATS, int S>
__global__ __launch_bounds__(64) void k_trace_lds(TraceArgs args) {
  __shared__ uint32_t lds[S * kWave];
  __shared__ uint8_t slots[kWave * Walker<ANY, INST, STATS, LdsStack<S>>::kPoolMax];
  LdsStack<S> stk{lds + (threadIdx.x & (kWave - 1))};
  run_lane<ANY, INST, STATS>(args, stk, slots);
}

// The same walk compiled for five resident waves per SIMD (96 VGPRs, 16 of them spilled to scratch,
// where k_trace_lds holds 116 and four waves): for launches with work for many rounds of waves,
// where the extra resident waves hide more of the walk's latency than the spills cost.
// S = 24 (6,400 bytes of LDS per wave) lets all 20 such waves of a CU be resident; S = 32 (8,448
// bytes) holds 19.
template <bool ANY, bool INST, int S>
__global__ __launch_bounds__(64, 5) void k_trace_lds_wide(TraceArgs args) {
  __shared__ uint32_t lds[S * kWave];
  __shared__ uint8_t slots[kWave * Walker<ANY, INST, false, LdsStack<S>>::kPoolMax];
  LdsStack<S> stk{lds + (threadIdx.x & (kWave - 1))};
  run_lane<ANY, INST, false>(args, stk, slots);
}

template <bool ANY, bool INST, bool STATS>
__global__ __launch_bounds__(64) void k_trace_global(TraceArgs args) {
  uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  GlobalStack stk{args.gstack + g, (uint64_t)gridDim.x * blockDim.x};
  __shared__ uint8_t slots[kWave * Walker<ANY, INST, STATS, GlobalStack>::kPoolMax];
  run_lane<ANY, INST, STATS>(args, stk, slots);
}

__global__ void k_set_counter(uint64_t *counter, uint64_t value) { *counter = value; }

}  // namespace

template <bool ANY, bool INST, bool STATS>
static hipError_t launch_trace(const TraceArgs &args, uint32_t stack_need, uint32_t grid, bool wide,
                               hipStream_t stream) {
  dim3 blk(kWave);
  if (stack_need <= 32) {
    if (!STATS && wide && stack_need <= 24)
      hipLaunchKernelGGL((k_trace_lds_wide<ANY, INST, 24>), dim3(grid), blk, 0, stream, args);
    else if (!STATS && wide)
      hipLaunchKernelGGL((k_trace_lds_wide<ANY, INST, 32>), dim3(grid), blk, 0, stream, args);
    else
      hipLaunchKernelGGL((k_trace_lds<ANY, INST, STATS, 32>), dim3(grid), blk, 0, stream, args);
  } else if (stack_need <= 64) {
    hipLaunchKernelGGL((k_trace_lds<ANY, INST, STATS, 64>), dim3(grid), blk, 0, stream, args);
  } else if (stack_need <= 128) {
    hipLaunchKernelGGL((k_trace_lds<ANY, INST, STATS, 128>), dim3(grid), blk, 0, stream, args);
  } else {
    hipLaunchKernelGGL((k_trace_global<ANY, INST, STATS>), dim3(grid), blk, 0, stream, args);
  }
  return hipGetLastError();
}

hipError_t launch_closest_hit(const TraceArgs &args, bool inst, bool stats, uint32_t stack_need, uint32_t grid,
                              bool wide, hipStream_t stream) {
  if (stats)
    return inst ? launch_trace<false, true, true>(args, stack_need, grid, false, stream)
                : launch_trace<false, false, true>(args, stack_need, grid, false, stream);
  return inst ? launch_trace<false, true, false>(args, stack_need, grid, wide, stream)
              : launch_trace<false, false, false>(args, stack_need, grid, wide, stream);
}

hipError_t launch_any_hit(const TraceArgs &args, bool inst, bool stats, uint32_t stack_need, uint32_t grid, bool wide,
                          hipStream_t stream) {
  if (stats)
    return inst ? launch_trace<true, true, true>(args, stack_need, grid, false, stream)
                : launch_trace<true, false, true>(args, stack_need, grid, false, stream);
  return inst ? launch_trace<true, true, false>(args, stack_need, grid, wide, stream)
              : launch_trace<true, false, false>(args, stack_need, grid, wide, stream);
}

hipError_t launch_set_counter(uint64_t *counter, uint64_t value, hipStream_t stream) {
  hipLaunchKernelGGL(k_set_counter, dim3(1), dim3(1), 0, stream, counter, value);
  return hipGetLastError();
}

}  // namespace skw
