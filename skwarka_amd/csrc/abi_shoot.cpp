// C ABI of photon shooting (skw_photon_shoot, skw_photon_set_*): the host loop over rounds and
// bounces around the kernels of shoot.hip.
//
// PhotonLTEIntegrator::ShootPhotons (Raytracer/LTEIntegrators/PhotonLTEIntegrator/
// PhotonLTEIntegrator.cpp:247-288) with the input filter's stop rules and the merging filter's
// per-map bookkeeping (PhotonShootingPipeline.cpp:48-92, 254-310) over rounds of paths (up to three
// in flight; a round's done flags are checked against exact counts when it is sorted, in round
// order); the lights' power CDF of _GetLightsPowerCDF (PhotonLTEIntegrator.cpp:80-124).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <memory>
#include <mutex>
#include <vector>

#include "internal.hpp"

using namespace skw;

struct skw_dl_scene;

namespace skw {
void dl_tables(const skw_dl_scene *dl, const DevMat **mats, const double **rad);
void dl_lights(const skw_dl_scene *dl, const std::vector<DevLight> **host, const DevLight **d_lights,
               const DevLightTri **d_ltris);
}  // namespace skw

namespace {

constexpr double kPi = 3.14159265358979323846;
constexpr double kDblEps = 2.220446049250313080847e-16;
const double kLumW[3] = {0x1.b38293da589a6p-5, 0x1.2b993cc221566p-1, 0x1.725d3400723f6p-2};

// A map's photon records (SoA) owned by a photon set, in the library's device memory pool (a
// set's arrays are reused by the next set instead of being mapped again).
struct PhotonArrays {
  float *pos = nullptr, *w = nullptr;
  uint32_t *codes = nullptr;
  uint64_t n = 0;
  void release() {
    for (void *p : {(void *)pos, (void *)w, (void *)codes})
      if (p) pool_free(p, nullptr);
    pos = w = nullptr;
    codes = nullptr;
    n = 0;
  }
  hipError_t alloc(uint64_t count, hipStream_t st) {
    release();
    hipError_t e = pool_alloc((void **)&pos, std::max<uint64_t>(count, 1) * 3 * sizeof(float), st);
    if (e == hipSuccess) e = pool_alloc((void **)&w, std::max<uint64_t>(count, 1) * 3 * sizeof(float), st);
    if (e == hipSuccess) e = pool_alloc((void **)&codes, std::max<uint64_t>(count, 1) * sizeof(uint32_t), st);
    return e;
  }
};

}  // namespace

struct skw_photon_set {
  PhotonArrays m[3];
  uint64_t paths[3] = {0, 0, 0};
  uint64_t total_paths = 0;
  int device = 0;
  ~skw_photon_set() {
    // as hipFree would: no kernel on the device may still read the arrays
    int cur = 0;
    const bool ok = hipGetDevice(&cur) == hipSuccess && hipSetDevice(device) == hipSuccess;
    if (ok) (void)hipDeviceSynchronize();
    for (auto &a : m) a.release();
    if (ok) (void)hipSetDevice(cur);
  }
};

namespace {

// A device buffer the library keeps between calls: it only grows (keeping its first `keep`
// items), so repeated calls of one size allocate and free nothing.
template <class T>
struct Keep {
  T *p = nullptr;
  uint64_t cap = 0;
  hipError_t grow(uint64_t need, uint64_t keep, hipStream_t s) {
    if (p && need <= cap) return hipSuccess;
    const uint64_t c = std::max<uint64_t>(std::max<uint64_t>(need, cap * 2), 1);
    T *np = nullptr;
    hipError_t e = hipMalloc((void **)&np, c * sizeof(T));
    if (e != hipSuccess) return e;
    keep = std::min(keep, cap);
    if (keep > 0) e = hipMemcpyAsync(np, p, keep * sizeof(T), hipMemcpyDeviceToDevice, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
      (void)hipFree(np);
      return e;
    }
    if (p) (void)hipFree(p);
    p = np;
    cap = c;
    return hipSuccess;
  }
};

// Photon records (SoA) of one map, with the path keys while they are in deposit order.
struct Records {
  Keep<float> pos, w;  // 3 per record
  Keep<uint32_t> codes, key;
  hipError_t reserve(uint64_t need, uint64_t keep, hipStream_t s, bool with_key) {
    hipError_t e = pos.grow(3 * need, 3 * keep, s);
    if (e == hipSuccess) e = w.grow(3 * need, 3 * keep, s);
    if (e == hipSuccess) e = codes.grow(need, keep, s);
    if (e == hipSuccess && with_key) e = key.grow(need, keep, s);
    return e;
  }
};

// One round of paths in flight: its stream, path lists, counters and photon records.
struct Lane {
  hipStream_t st = nullptr;
  Keep<skw_ray> rays[2];
  Keep<ShootPath> state[2];
  Keep<skw_hit> hits;
  Keep<uint64_t> counts;
  Records round[3];       // the round's photons of each map, in deposit order
  uint64_t *h = nullptr;  // two pinned copies of the five counters
  hipEvent_t ev[2] = {nullptr, nullptr};
  hipEvent_t fin = nullptr;  // the lane's last round has been sorted into the staging arrays
};

// The working memory of photon shooting on one device, kept for the life of the process; calls
// on one device take turns on it. Several lanes let a round start while earlier rounds' last few
// paths still bounce.
constexpr int kLanes = 3;

struct ShootWorkspace {
  std::mutex mu;
  bool ready = false;
  hipStream_t aux[kLanes - 1] = {};  // the other lanes' streams (default priority)
  hipEvent_t start = nullptr;
  Keep<double> cdf;
  Lane lane[kLanes];
  Keep<uint32_t> iota, order, keys_sorted;
  Keep<uint8_t> sort_tmp;
  Records out[3];  // every round's photons of each map, in (path, intersection) order
  hipError_t init() {
    if (ready) return hipSuccess;
    hipError_t e = hipSuccess;
    for (int i = 0; i < kLanes - 1 && e == hipSuccess; ++i) e = hipStreamCreateWithFlags(&aux[i], hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&start, hipEventDisableTiming);
    for (Lane &L : lane) {
      if (e == hipSuccess) e = hipHostMalloc((void **)&L.h, 10 * sizeof(uint64_t), hipHostMallocDefault);
      for (int i = 0; i < 2 && e == hipSuccess; ++i) e = hipEventCreateWithFlags(&L.ev[i], hipEventDisableTiming);
      if (e == hipSuccess) e = hipEventCreateWithFlags(&L.fin, hipEventDisableTiming);
    }
    ready = e == hipSuccess;
    return e;
  }
};

constexpr int kWorkspaceDevices = 64;

hipError_t shoot_workspace(ShootWorkspace **out) {
  static ShootWorkspace ws[kWorkspaceDevices];
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= kWorkspaceDevices) return hipErrorInvalidDevice;
  *out = &ws[dev];
  return hipSuccess;
}

// Host state of the round a lane runs.
struct RoundState {
  bool busy = false;      // paths still bounce
  bool finished = false;  // every path ended; waiting to be sorted in round order
  uint64_t index = 0, n = 0;
  bool done[3] = {false, false, false};
  ShootArgs a;
  uint64_t A = 0, P[3] = {0, 0, 0};  // exact counts at step s_known (see step_lane)
  int64_t s_known = -1, m = 1;
};


// The lane's next bounce. The host reads the counters of step k (k = 0 the start, k >= 1 bounce k)
// from a pinned copy two steps late: active paths never increase and a path deposits at most one
// photon per bounce, so from exact counts at step s (A paths active for bounce s + 1, P[j]
// records) bounce m runs a grid for A paths and holds at most P[j] + (m - s) A records; the kernels
// read the exact path count on the device. Without `block` it returns false when the counters it
// needs have not arrived.
bool step_lane(const skw_scene *scene, Lane &L, RoundState &R, bool block, hipError_t &e, skw_status &rs) {
  if (R.m >= 2) {
    hipEvent_t ev = L.ev[R.m & 1];
    if (!block) {
      const hipError_t q = hipEventQuery(ev);
      if (q == hipErrorNotReady) return false;
      if (q != hipSuccess) {
        e = q;
        return true;
      }
    } else if ((e = hipEventSynchronize(ev)) != hipSuccess) {
      return true;
    }
    const uint64_t *c = L.h + 5 * (R.m & 1);  // step m - 2
    R.A = c[(R.m - 2) & 1];
    for (int j = 0; j < 3; ++j) R.P[j] = c[2 + j];
    R.s_known = R.m - 2;
  }
  if (R.A == 0) {  // no path reaches bounce s_known + 1: every later step is empty
    e = hipStreamSynchronize(L.st);  // the copies still in flight write the pinned counters
    R.busy = false;
    R.finished = true;
    return true;
  }
  const uint64_t steps = (uint64_t)(R.m - R.s_known);
  for (int j = 0; j < 3 && e == hipSuccess; ++j)
    if (!R.done[j] && j != 1) e = L.round[j].reserve(R.P[j] + steps * R.A, R.P[j] + (steps - 1) * R.A, L.st, true);
  if (e != hipSuccess) return true;
  ShootArgs &a = R.a;
  for (int j = 0; j < 3; ++j) {
    a.ph_pos[j] = L.round[j].pos.p;
    a.ph_w[j] = L.round[j].w.p;
    a.ph_codes[j] = L.round[j].codes.p;
    a.ph_key[j] = L.round[j].key.p;
  }
  rs = enqueue_trace(scene, L.rays[a.cur].p, R.A, L.counts.p + a.cur, L.hits.p, nullptr, L.st, false, nullptr,
                     nullptr, false);
  if (rs != SKW_OK) return true;
  if ((e = hipMemsetAsync(L.counts.p + (a.cur ^ 1u), 0, sizeof(uint64_t), L.st)) != hipSuccess) return true;
  if ((e = launch_shoot_bounce(a, R.A, L.st)) != hipSuccess) return true;
  a.cur ^= 1u;
  if ((e = hipMemcpyAsync(L.h + 5 * (R.m & 1), L.counts.p, 5 * sizeof(uint64_t), hipMemcpyDeviceToHost, L.st)) !=
      hipSuccess)
    return true;
  e = hipEventRecord(L.ev[R.m & 1], L.st);
  ++R.m;
  return true;
}

// The shooting loop over rounds, up to three in flight.
hipError_t shoot_rounds(const skw_scene *scene, const skw_dl_scene *dl, const skw_shoot_params *prm,
                        const std::vector<DevLight> &lights, const DevLight *d_lights, const DevLightTri *d_ltris,
                        const DevMat *mats, skw_photon_set *set, hipStream_t st, skw_status *rs) {
  const uint32_t nl = (uint32_t)lights.size();
  // _GetLightsPowerCDF: luminance of (pi * area) * radiance, accumulated, divided by the total.
  std::vector<double> cdf(nl, 0.0);
  for (uint32_t l = 0; l < nl; ++l) {
    double y = 0.0;
    for (int c = 0; c < 3; ++c) y += kLumW[c] * ((kPi * lights[l].area) * lights[l].radiance[c]);
    cdf[l] = y;
  }
  for (uint32_t i = 1; i < nl; ++i) cdf[i] += cdf[i - 1];
  const double total = cdf[nl - 1];
  if (total > kDblEps)
    for (uint32_t i = 0; i < nl; ++i) cdf[i] /= total;
  else
    cdf[nl - 1] = 1.0;

  const uint64_t R = std::min<uint64_t>(prm->paths_per_round, prm->max_paths);
  ShootWorkspace *ws = nullptr;
  hipError_t e = shoot_workspace(&ws);
  if (e != hipSuccess) return e;
  std::lock_guard<std::mutex> lock(ws->mu);
  if ((e = ws->init()) != hipSuccess) return e;
  ws->lane[0].st = st;
  for (int i = 1; i < kLanes; ++i) ws->lane[i].st = ws->aux[i - 1];
  e = ws->cdf.grow(nl, 0, st);
  for (Lane &L : ws->lane) {
    for (int i = 0; i < 2 && e == hipSuccess; ++i) {
      e = L.rays[i].grow(R, 0, L.st);
      if (e == hipSuccess) e = L.state[i].grow(R, 0, L.st);
    }
    if (e == hipSuccess) e = L.hits.grow(R, 0, L.st);
    if (e == hipSuccess) e = L.counts.grow(5, 0, L.st);
  }
  if (e == hipSuccess) e = hipMemcpyAsync(ws->cdf.p, cdf.data(), nl * sizeof(double), hipMemcpyHostToDevice, st);
  // the other lanes start after everything the caller's stream holds
  if (e == hipSuccess) e = hipEventRecord(ws->start, st);
  for (int i = 0; i < kLanes - 1 && e == hipSuccess; ++i) e = hipStreamWaitEvent(ws->aux[i], ws->start, 0);
  if (e != hipSuccess) return e;
  int key_bits = 1;
  while ((1ull << key_bits) < R) ++key_bits;

  ShootArgs base;
  std::memset(&base, 0, sizeof(base));
  base.geom = scene->geom_args();
  base.mats = mats;
  base.lights = d_lights;
  base.ltris = d_ltris;
  base.light_cdf = ws->cdf.p;
  base.n_lights = nl;
  base.seed = prm->seed;

  RoundState rst[kLanes];
  uint64_t have[3] = {0, 0, 0};  // photons of the rounds sorted so far
  uint64_t started = 0, sorted = 0;
  bool stop = false;
  hipEvent_t last_fin = nullptr;

  // Whether a round may start now, and with which maps done. The rule (a map is done for a round
  // once the rounds before it hold max_photons) needs the exact counts of every earlier round; a
  // round starts before the rounds in flight are counted, with a map marked done only when the
  // sorted rounds already fill it (counts only grow, so such a mark is final). The prediction is
  // checked when the round is sorted, in round order, with exact counts: a map that turned out full
  // drops the round's photons and does not count its paths. That equals shooting the round with
  // the map marked done, because a done map only suppresses its own deposits (and, for the
  // indirect map, the bounces after the first intersection, which deposit into that map alone);
  // the other maps' photons are the same. A round that finds every map full is dropped whole.
  // Returns 1 start, 0 wait, -1 no more rounds.
  auto decide = [&](bool done[3]) -> int {
    bool other = false;
    for (const RoundState &r : rst) other = other || r.busy || r.finished;
    bool all_done = true;
    for (int j = 0; j < 3; ++j) {
      done[j] = have[j] >= prm->max_photons[j];
      all_done = all_done && done[j];
    }
    if (set->total_paths >= prm->max_paths) return -1;
    if (all_done) return -1;
    if (set->total_paths > 10000000ull) {
      if (other) return 0;  // the stop rule below needs exact counts
      bool stuck = true;
      for (int j = 0; j < 3; ++j) stuck = stuck && (done[j] || have[j] == 0);
      if (stuck) return -1;
    }
    return 1;
  };

  while (e == hipSuccess && *rs == SKW_OK) {
    bool progress = false;
    // sort the finished rounds into the staging arrays, in round order
    for (int l = 0; l < kLanes && e == hipSuccess; ++l) {
      RoundState &Rs = rst[l];
      if (!Rs.finished || Rs.index != sorted) continue;
      Lane &L = ws->lane[l];
      if (last_fin && (e = hipStreamWaitEvent(L.st, last_fin, 0)) != hipSuccess) break;
      // the done flags the round should have had: exact counts of every earlier round
      bool done_now[3], all_now = true;
      for (int j = 0; j < 3; ++j) {
        done_now[j] = have[j] >= prm->max_photons[j];
        all_now = all_now && done_now[j];
      }
      if (all_now) set->total_paths -= Rs.n;  // no round would have started: its paths do not count
      for (int j = 0; j < 3 && e == hipSuccess && !all_now; ++j) {
        if (done_now[j]) continue;
        set->paths[j] += Rs.n;
        const uint64_t k = Rs.P[j];
        if (k == 0) continue;
        size_t sort_bytes = 0;
        if ((e = sort_pairs_u32(nullptr, sort_bytes, nullptr, nullptr, nullptr, nullptr, k, 0, key_bits, L.st)) !=
            hipSuccess)
          break;
        if ((e = ws->sort_tmp.grow(sort_bytes, 0, L.st)) != hipSuccess) break;
        if ((e = ws->iota.grow(k, 0, L.st)) != hipSuccess) break;
        if ((e = ws->order.grow(k, 0, L.st)) != hipSuccess) break;
        if ((e = ws->keys_sorted.grow(k, 0, L.st)) != hipSuccess) break;
        if ((e = launch_iota_u32(ws->iota.p, k, L.st)) != hipSuccess) break;
        if ((e = sort_pairs_u32(ws->sort_tmp.p, sort_bytes, L.round[j].key.p, ws->keys_sorted.p, ws->iota.p,
                                ws->order.p, k, 0, key_bits, L.st)) != hipSuccess)
          break;
        Records &o = ws->out[j];
        if ((e = o.reserve(have[j] + k, have[j], L.st, false)) != hipSuccess) break;
        e = launch_photon_gather_records(L.round[j].pos.p, L.round[j].w.p, L.round[j].codes.p, ws->order.p, k,
                                         o.pos.p + 3 * have[j], o.w.p + 3 * have[j], o.codes.p + have[j], L.st);
        have[j] += k;
      }
      if (e == hipSuccess) e = hipEventRecord(L.fin, L.st);
      last_fin = L.fin;
      Rs.finished = false;
      ++sorted;
      progress = true;
    }
    if (e != hipSuccess) break;
    // start rounds on free lanes
    for (int l = 0; l < kLanes && !stop && e == hipSuccess; ++l) {
      RoundState &Rs = rst[l];
      if (Rs.busy || Rs.finished) continue;
      bool done[3];
      const int d = decide(done);
      if (d < 0) stop = true;
      if (d <= 0) break;
      Lane &L = ws->lane[l];
      Rs = RoundState();
      Rs.busy = true;
      Rs.index = started++;
      Rs.n = std::min<uint64_t>(R, prm->max_paths - set->total_paths);
      ShootArgs &a = Rs.a;
      a = base;
      a.first = set->total_paths;
      a.n = Rs.n;
      set->total_paths += Rs.n;
      for (int j = 0; j < 3; ++j) {
        Rs.done[j] = done[j];
        a.done[j] = done[j] ? 1u : 0u;
      }
      for (int i = 0; i < 2; ++i) {
        a.rays[i] = L.rays[i].p;
        a.state[i] = L.state[i].p;
      }
      a.hits = L.hits.p;
      a.counts = L.counts.p;
      a.cur = 0;
      Rs.A = Rs.n;
      // direct photons: at most one per path
      if (!done[1] && (e = L.round[1].reserve(Rs.n, 0, L.st, true)) != hipSuccess) break;
      for (int j = 0; j < 3; ++j) {
        a.ph_pos[j] = L.round[j].pos.p;
        a.ph_w[j] = L.round[j].w.p;
        a.ph_codes[j] = L.round[j].codes.p;
        a.ph_key[j] = L.round[j].key.p;
      }
      if ((e = hipMemsetAsync(L.counts.p, 0, 5 * sizeof(uint64_t), L.st)) != hipSuccess) break;
      if ((e = launch_shoot_start(a, L.st)) != hipSuccess) break;
      if ((e = hipMemcpyAsync(L.h, L.counts.p, 5 * sizeof(uint64_t), hipMemcpyDeviceToHost, L.st)) != hipSuccess)
        break;
      if ((e = hipEventRecord(L.ev[0], L.st)) != hipSuccess) break;
      progress = true;
    }
    if (e != hipSuccess) break;
    // the lanes' next bounces
    int oldest = -1;
    for (int l = 0; l < kLanes && e == hipSuccess && *rs == SKW_OK; ++l) {
      if (!rst[l].busy) continue;
      if (oldest < 0 || rst[l].index < rst[oldest].index) oldest = l;
      progress = step_lane(scene, ws->lane[l], rst[l], false, e, *rs) || progress;
    }
    if (e != hipSuccess || *rs != SKW_OK) break;
    bool idle = true;
    for (const RoundState &r : rst) idle = idle && !r.busy && !r.finished;
    if (stop && idle) break;
    if (!progress && oldest >= 0) step_lane(scene, ws->lane[oldest], rst[oldest], true, e, *rs);
  }
  if (*rs != SKW_OK || e != hipSuccess) {
    // no copy may still target the pinned counters
    for (hipStream_t x : ws->aux) (void)hipStreamSynchronize(x);
    (void)hipStreamSynchronize(st);
    return e;
  }
  // each map's rounds, concatenated into arrays of the exact size, on the caller's stream
  if (last_fin && (e = hipStreamWaitEvent(st, last_fin, 0)) != hipSuccess) return e;
  for (int j = 0; j < 3 && e == hipSuccess; ++j) {
    if (have[j] == 0) continue;
    if ((e = set->m[j].alloc(have[j], st)) != hipSuccess) break;
    const Records &o = ws->out[j];
    e = hipMemcpyAsync(set->m[j].pos, o.pos.p, 3 * have[j] * sizeof(float), hipMemcpyDeviceToDevice, st);
    if (e == hipSuccess)
      e = hipMemcpyAsync(set->m[j].w, o.w.p, 3 * have[j] * sizeof(float), hipMemcpyDeviceToDevice, st);
    if (e == hipSuccess)
      e = hipMemcpyAsync(set->m[j].codes, o.codes.p, have[j] * sizeof(uint32_t), hipMemcpyDeviceToDevice, st);
    set->m[j].n = have[j];
  }
  for (int i = 0; i < kLanes - 1 && e == hipSuccess; ++i) e = hipStreamSynchronize(ws->aux[i]);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  return e;
}

}  // namespace

extern "C" {

skw_status skw_photon_shoot(const skw_scene *scene, const skw_dl_scene *dl, const skw_shoot_params *prm,
                            skw_photon_set **out, void *stream) {
  if (!scene || !dl || !prm || !out) return fail(SKW_INVALID_ARG, "arguments must not be NULL");
  *out = nullptr;
  if (!scene->on_device) return fail(SKW_UNSUPPORTED, "scene was created host-only");
  if (prm->paths_per_round == 0) return fail(SKW_INVALID_ARG, "paths_per_round must be > 0");
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  std::unique_ptr<skw_photon_set> set(new skw_photon_set());
  SKW_HIP(hipGetDevice(&set->device), "current device");
  const std::vector<DevLight> *lights = nullptr;
  const DevLight *d_lights = nullptr;
  const DevLightTri *d_ltris = nullptr;
  dl_lights(dl, &lights, &d_lights, &d_ltris);
  const DevMat *mats = nullptr;
  const double *rad = nullptr;
  dl_tables(dl, &mats, &rad);
  if (lights->empty() || prm->max_paths == 0) {
    *out = set.release();
    return SKW_OK;
  }
  skw_status rs = SKW_OK;
  const hipError_t e = shoot_rounds(scene, dl, prm, *lights, d_lights, d_ltris, mats, set.get(), st, &rs);
  if (rs != SKW_OK) return rs;
  if (e != hipSuccess) return hip_fail(e, "photon shooting");
  *out = set.release();
  return SKW_OK;
}

skw_status skw_photon_set_info(const skw_photon_set *set, uint64_t counts[3], uint64_t paths[3],
                               uint64_t *total_paths) {
  if (!set) return fail(SKW_INVALID_ARG, "set must not be NULL");
  for (int j = 0; j < 3; ++j) {
    if (counts) counts[j] = set->m[j].n;
    if (paths) paths[j] = set->paths[j];
  }
  if (total_paths) *total_paths = set->total_paths;
  return SKW_OK;
}

skw_status skw_photon_set_export(const skw_photon_set *set, uint32_t which, float *pos, float *weight, uint16_t *dir,
                                 uint16_t *normal) {
  if (!set) return fail(SKW_INVALID_ARG, "set must not be NULL");
  if (which > 2) return fail(SKW_INVALID_ARG, "which must be 0, 1 or 2");
  const PhotonArrays &m = set->m[which];
  if (m.n == 0) return SKW_OK;
  if (pos) SKW_HIP(hipMemcpy(pos, m.pos, m.n * 3 * sizeof(float), hipMemcpyDeviceToHost), "photon export");
  if (weight) SKW_HIP(hipMemcpy(weight, m.w, m.n * 3 * sizeof(float), hipMemcpyDeviceToHost), "photon export");
  if (dir || normal) {
    std::vector<uint32_t> codes(m.n);
    SKW_HIP(hipMemcpy(codes.data(), m.codes, m.n * sizeof(uint32_t), hipMemcpyDeviceToHost), "photon export");
    for (uint64_t i = 0; i < m.n; ++i) {
      if (dir) dir[i] = (uint16_t)(codes[i] & 0xFFFFu);
      if (normal) normal[i] = (uint16_t)(codes[i] >> 16);
    }
  }
  return SKW_OK;
}

skw_status skw_photon_set_map(const skw_photon_set *set, uint32_t which, skw_photon_map **out, void *stream) {
  if (!set || !out) return fail(SKW_INVALID_ARG, "set and out must not be NULL");
  if (which > 2) return fail(SKW_INVALID_ARG, "which must be 0, 1 or 2");
  const PhotonArrays &m = set->m[which];
  return photon_map_build_device(m.n, m.pos, m.w, m.codes, nullptr, nullptr, out, reinterpret_cast<hipStream_t>(stream));
}

skw_status skw_photon_set_destroy(skw_photon_set *set) {
  delete set;
  return SKW_OK;
}

}  // extern "C"
