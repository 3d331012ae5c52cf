// Many scenes in one persistent launch (gfx950).
//
// Work is a queue of chunks (scene, ray range); a wave takes one chunk at a time with a single
// atomicAdd on the queue head, and its lanes pull rays from the chunk as they finish (ballot +
// prefix rank, as in the single-scene kernel), so a wave only waits for stragglers at chunk ends.
// All lanes of a wave work on the same scene, so scene pointers stay wave-uniform. Camera rays are
// generated in the kernel (no ray buffer): PerspectiveCamera::GenerateRay
// (Raytracer/Cameras/PerspectiveCamera.cpp:36-64) at pixel + splitmix64 jitter, the same stream as
// the direct-lighting camera. The walk is trace_device.hpp's (TriangleAccelerator::_NodeIntersect
// semantics). Photon-gather jobs (k nearest + disc estimate, gather_device.hpp) can share the
// queue: their chunks are query ranges, which a wave takes like a ray chunk and serves 64 queries
// at a time before it takes its next chunk. When an item's last chunk completes, the wave records
// the wall clock for it.
#include <hip/hip_runtime.h>

#include "gather_device.hpp"
#include "geom_device.hpp"
#include "kernels.hpp"
#include "trace_device.hpp"

namespace skw {
namespace {

using namespace trace;

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
__device__ __forceinline__ double rnd(uint64_t seed, uint64_t sample, uint32_t dim) {
  uint64_t h = mix64(seed ^ mix64(sample * 64ull + dim));
  return (double)(h >> 11) * 0x1.0p-53;
}

__device__ __forceinline__ void camera_ray(const DevBatchScene &s, uint64_t idx, RayState &r) {
  const uint64_t pix = idx / s.spp;
  const uint32_t px = (uint32_t)(pix % s.width), py = (uint32_t)(pix / s.width);
  double ix = (double)px + rnd(s.seed, idx, 0), iy = (double)py + rnd(s.seed, idx, 1);
  double cx = 2.0 * s.x_tan * (ix - s.width * 0.5) / s.width;
  double cy = 2.0 * s.y_tan * (iy - s.height * 0.5) / s.height;
  dev::D3 d{s.cam_r[0] * cx + s.cam_r[1] * cy + s.cam_r[2], s.cam_r[3] * cx + s.cam_r[4] * cy + s.cam_r[5],
            s.cam_r[6] * cx + s.cam_r[7] * cy + s.cam_r[8]};
  dev::normalize3(d);
  make_state(r, s.cam_o[0], s.cam_o[1], s.cam_o[2], d.x, d.y, d.z, 0.0);
}

__global__ void k_batch_reset(BatchArgs a) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0) {
    *a.queue_head = 0;
    *a.frame_start = wall_clock64();
    *a.stalled = 0;
  }
  if (i < a.n_scenes + a.n_jobs) {
    a.item_done[i] = 0;
    a.item_end[i] = 0;
  }
}

// Lanes a wave gathers before it starts new rays / tests leaves (as the single-scene kernel's
// defaults).
constexpr uint32_t kRefillMin = 8;
constexpr uint32_t kLeafMin = 16;

// Marks len units of item done (lead: the one lane of the wave that does it); the wave that
// completes the item records the wall clock.
__device__ __forceinline__ void account(const BatchArgs &a, uint32_t item, uint64_t len, uint64_t total, bool lead) {
  if (lead) {
    unsigned long long old =
        atomicAdd(reinterpret_cast<unsigned long long *>(a.item_done + item), (unsigned long long)len);
    if (old + len == total) a.item_end[item] = wall_clock64();
  }
}

// GATHER: the queue may hold photon-gather chunks (an instance of its own, so the gather's
// registers do not lower the occupancy of ray-only batches).
template <int S, int WAVES, bool GATHER>
__global__ __launch_bounds__(64 * WAVES) void k_batch_trace(BatchArgs a) {
  __shared__ uint32_t lds[S * 64 * WAVES];
  const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  LdsStack<S> stk{lds + wave * (S * 64) + lane};
  Walker<false, false, false, LdsStack<S>> w;
  w.insts = nullptr;
  w.inst_range = nullptr;
  w.stk = stk;
  bool busy = false;
  uint64_t idx = 0, cursor = 0, end = 0, chunk_len = 0;
  uint32_t sid = 0xFFFFFFFFu;
  const DevBatchScene *sc = nullptr;
  // Every iteration moves the wave forward: with no lane busy it takes a chunk or starts rays;
  // otherwise a busy lane between leaves advances, or, when every busy lane holds a leaf
  // (moving == 0), the leaves are tested. The conditions below are ballots, the same in every
  // lane, so an iteration that moves nothing can only mean that lanes and ballots disagree: the
  // wave then records the stall and leaves instead of spinning.
  // Nothing here assumes that all 64 lanes run the iteration together: the compiler may let a
  // subset of the lanes go round the loop on its own (a loop back edge threaded past a branch
  // that only lane 0 takes did exactly that to an earlier form of this kernel, which then tested
  // `idle == ~0ull` among 63 lanes forever). "Every lane idle" means every lane of `active`, the
  // lanes running this iteration, and the claim's leader and broadcast come from that set, so
  // such a subset takes and finishes chunks of its own.
  while (true) {
    bool moved = false;
    const uint64_t active = __ballot(1);
    const int leader = __ffsll((unsigned long long)active) - 1;
    const bool lead = (int)lane == leader;
    const uint64_t idle = __ballot(!busy);
    if (idle == active && cursor >= end) {
      moved = true;
      // The wave's chunk is done: account for it, then take the next one.
      if (sid != 0xFFFFFFFFu) account(a, sid, chunk_len, sc->n_rays, lead);
      if (GATHER) sid = 0xFFFFFFFFu;  // a gather chunk taken next must not account this scene again
      unsigned long long q = 0;
      if (lead) q = atomicAdd(reinterpret_cast<unsigned long long *>(a.queue_head), 1ull);
      q = __shfl(q, leader);
      if (q >= a.n_chunks) break;
      const DevChunk c = a.chunks[q];
      if (GATHER) {
        const uint32_t item = __builtin_amdgcn_readfirstlane(c.item);
        if (item >= a.n_scenes) {
          // A photon-gather chunk: its queries 64 at a time, the lookup stack in the LDS rows the
          // ray walk uses (no lane of the wave is walking a ray now).
          const GatherArgs &ga = a.jobs[item - a.n_scenes];
          const gath::LdsStack<S> gstk{stk.base};
          for (uint64_t qi = c.begin + lane; qi < c.end; qi += 64) gath::gather_one<0, false>(ga, qi, gstk);
          account(a, item, c.end - c.begin, ga.nq, lead);
          cursor = 0;
          end = 0;
          continue;
        }
      }
      sid = c.item;
      cursor = c.begin;
      end = c.end;
      chunk_len = c.end - c.begin;
      sc = a.scenes + sid;
      w.set_root(sc->nodes, sc->root);
      w.tris = sc->tris;
    }
    // Refill, leaf and advance phases as in the single-scene kernel (traverse.hip): refills and
    // leaf tests wait until kRefillMin / kLeafMin lanes need them, unless nothing else can move
    // or the chunk has no rays left to start.
    if (idle != 0 && cursor < end && (idle == active || (uint32_t)__popcll(idle) >= kRefillMin)) {
      moved = true;
      const uint64_t avail = end - cursor;
      const uint64_t rank = (uint64_t)__popcll(idle & ((1ull << lane) - 1ull));
      if (!busy && rank < avail) {
        idx = cursor + rank;
        RayState r;
        camera_ray(*sc, idx, r);
        w.start(r, __builtin_huge_val());
        busy = true;
      }
      const uint64_t took = (uint64_t)__popcll(idle);
      cursor += took < avail ? took : avail;
    }
    const uint64_t held = __ballot(busy && w.at_leaf);
    if (held) {
      const uint64_t moving = __ballot(busy && !w.at_leaf);
      if ((uint32_t)__popcll(held) >= kLeafMin || moving == 0 || cursor >= end) {
        moved = true;
        if (busy && w.at_leaf) w.leaf();
      }
    }
    const bool fast = sc->nested && __ballot(busy && !w.nan_free) == 0;
    moved = moved || __ballot(busy && !w.at_leaf) != 0;
    if (!moved) {
      if (lead) atomicOr(reinterpret_cast<unsigned long long *>(a.stalled), 1ull);
      break;
    }
    if (busy && !w.at_leaf) {
      const bool done = fast ? w.template advance<true>() : w.template advance<false>();
      if (done) {
        store_hit(a.hits + sc->hit_offset + idx, w.tris, w.found, w.best, w.tmax, w.r);
        busy = false;
      }
    }
  }
}

__global__ __launch_bounds__(256) void k_batch_rays(const DevBatchScene *s, skw_ray *rays, uint64_t n) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  RayState r;
  camera_ray(*s, i, r);
  skw_ray &o = rays[i];
  o.o[0] = r.o0;
  o.o[1] = r.o1;
  o.o[2] = r.o2;
  o.d[0] = r.d0;
  o.d[1] = r.d1;
  o.d[2] = r.d2;
  o.tmin = 0.0;
  o.tmax = __builtin_huge_val();
}

}  // namespace

hipError_t launch_batch_reset(const BatchArgs &a, hipStream_t stream) {
  uint32_t grid = (a.n_scenes + a.n_jobs + 255) / 256;
  hipLaunchKernelGGL(k_batch_reset, dim3(grid > 0 ? grid : 1), dim3(256), 0, stream, a);
  return hipGetLastError();
}

hipError_t launch_batch_trace(const BatchArgs &a, uint32_t waves, uint32_t stack, uint32_t grid, hipStream_t stream) {
  dim3 g(grid);
#define SKW_BATCH_LAUNCH(S, W)                                                           \
  do {                                                                                   \
    if (a.n_jobs > 0)                                                                    \
      hipLaunchKernelGGL((k_batch_trace<S, W, true>), g, dim3(64 * (W)), 0, stream, a);  \
    else                                                                                 \
      hipLaunchKernelGGL((k_batch_trace<S, W, false>), g, dim3(64 * (W)), 0, stream, a); \
  } while (0)
  if (stack <= 32) {
    if (waves == 1) SKW_BATCH_LAUNCH(32, 1);
    else if (waves == 2) SKW_BATCH_LAUNCH(32, 2);
    else SKW_BATCH_LAUNCH(32, 4);
  } else {
    if (waves == 1) SKW_BATCH_LAUNCH(64, 1);
    else if (waves == 2) SKW_BATCH_LAUNCH(64, 2);
    else SKW_BATCH_LAUNCH(64, 4);
  }
#undef SKW_BATCH_LAUNCH
  return hipGetLastError();
}

hipError_t launch_batch_rays(const DevBatchScene *scene, skw_ray *rays, uint64_t n, hipStream_t stream) {
  if (n == 0) return hipSuccess;
  uint32_t grid = (uint32_t)((n + 255) / 256);
  hipLaunchKernelGGL(k_batch_rays, dim3(grid), dim3(256), 0, stream, scene, rays, n);
  return hipGetLastError();
}

}  // namespace skw
