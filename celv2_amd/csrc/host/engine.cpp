// The C ABI of libcelv2hip.so (include/celv2hip.h): device memory, reference-key caches, the
// per-launch queue of kernels on the engine's main HIP stream, prepare() on a side stream, and row
// completion events.
//
// Queue order of a launch (VulkanEngine::launch, src/vk/vk_engine.cpp:173-278, re-planned for HIP):
//   source: the prepare()d pyramid and intra scores under src_key (built on the side stream, which
//   the main stream waits for), else import + pyramid + intra now;
//   references: a slot holding the key, else a free slot: import + 15 phase planes + pyramid;
//   motion search: levels 4..0, each one launch over every reference (kernel boundaries order them);
//   rows in row-stream order (or all rows at once): inter scores, finalize, copy of the rows' records
//   to pinned host memory, an event per row that wait_row() waits on.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

#include "celv2hip.h"
#include "../kernels/launch.h"
#include "pruner.h"
#include "status.h"

using namespace celv2hip;
using namespace celv2host;

namespace {

size_t round_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

// Row groups of a launch with one row stream (the wavefront form) unless set otherwise
// (celv2hip_set_row_groups, chosen by the autotuner): each group's records are downloaded while the
// next groups compute.
constexpr int kDefaultRowGroups = 4;

// A picture's pyramid in one allocation: level 0 (a reference's followed by its 15 phase planes,
// phase_bytes apart) then levels 1..4, every plane bordered.
struct Pyramid {
  uint8_t* mem = nullptr;
  size_t bytes = 0;
  DPlane lv[kLevels] = {};
  long long phase_bytes = 0;
};

size_t plane_bytes(int w, int h, int* stride) {
  *stride = static_cast<int>(round_up(static_cast<size_t>(w) + 2 * kBorder, 64));
  return static_cast<size_t>(h + 2 * kBorder) * *stride + 2 * kSlack;
}

void layout(Pyramid& p, int pw, int ph, bool phases) {
  size_t off = 0;
  int w = pw, h = ph;
  for (int l = 0; l < kLevels; l++) {
    int stride;
    const size_t b = plane_bytes(w, h, &stride);
    p.lv[l] = DPlane{p.mem + off + kSlack + static_cast<size_t>(kBorder) * stride + kBorder, stride, w, h};
    if (l == 0 && phases) {
      // The kernels address phase planes as (phase x phase_bytes / 64) << 6 with a 24-bit multiply.
      if (b % 64 != 0 || b / 64 >= (size_t{1} << 24)) throw Failure(CELV2_ERR_ARG, "frame too large for the phase-plane layout");
      p.phase_bytes = static_cast<long long>(b);
    }
    off += b * ((l == 0 && phases) ? kPhases : 1);
    w /= 2, h /= 2;
  }
}

size_t pyramid_bytes(int pw, int ph, bool phases) {
  size_t total = 0;
  int w = pw, h = ph, stride;
  for (int l = 0; l < kLevels; l++) {
    total += plane_bytes(w, h, &stride) * ((l == 0 && phases) ? kPhases : 1);
    w /= 2, h /= 2;
  }
  return total;
}

struct Upload {  // pinned staging + device copy of one frame's luma
  void* host = nullptr;
  void* dev = nullptr;
  size_t bytes = 0;
  hipEvent_t done = nullptr;  // the last copy out of `host` finished
};

struct RefSlot {
  long long key = -1, used = -1;
  Pyramid p;
  Upload up;
};

struct SrcSlot {
  long long key = -1;
  Pyramid p;
  Upload up;
  uint32_t* intra = nullptr;
  uint8_t* imask_used = nullptr;  // the intra mode groups `intra` was made with, one byte per superblock
  // prepare() fills a slot on the side stream while launches run on the main stream: each stream
  // waits for the other's last work on the slot before touching it.
  hipEvent_t side_done = nullptr, main_done = nullptr;
  bool side_pending = false, main_pending = false;
  // prepare()'s pyramid done (before its intra scores): a launch of the wavefront form reads only the
  // pyramid on the launch stream (its finalize, on the side stream, follows the intra scores there).
  hipEvent_t pyr_done = nullptr;
  bool pyr_pending = false;
  // prepare() left the intra scores to the launch (the pruner chooses their mode groups there).
  bool intra_deferred = false;
};

}  // namespace

struct celv2hip_engine {
  int width, height, sb_size, device, pw, ph, sb_cols, sb_rows;
  hipStream_t stream = nullptr;  // launches
  hipStream_t side = nullptr;    // prepare(): the next source's planes and intra scores
  // A new reference's 15 phase planes, built beside the coarse pyramid levels and the motion search
  // of levels 4..1 (only level 0's refinement, inter and compound scores read them): sub_start is
  // recorded on the launch stream after the import, sub_done on this stream after the phase planes.
  hipStream_t aux = nullptr;
  // Record downloads (rec_dev -> rec_host) run on `aux` too, late in the launch, long after its phase
  // planes: each group of rows is copied once its finalize is done, while the launch stream goes on
  // with the next rows (a stream of their own would put every engine's launch stream on the same of
  // the process's 4 hardware queues). group_done[g] is recorded on aux after group g's copy, and row
  // r of the last launch waits for group_done[row_group[r]]; copy_done
  // after a launch's last copy (the next launch waits for it before finalize overwrites rec_dev).
  // rows_ev: one per group, recorded on the launch stream after the group's finalize.
  std::vector<hipEvent_t> rows_ev;
  std::vector<hipEvent_t> inter_ev;  // wavefront form: group g's inter scores done (launch stream)
  hipEvent_t copy_done = nullptr;
  bool copy_pending = false;
  hipEvent_t sub_start = nullptr, sub_done = nullptr;
  hipEvent_t order_ev = nullptr;  // orders this engine's stream around batched launches
  // The last launch ran in a batch on another engine's stream (batch_on), batch_done recorded there
  // behind it: later work of this engine that is not ordered by that stream waits for it then. (A
  // wait put on this engine's own stream at once would block the hardware queue that stream shares
  // with other streams, and so their work, until the whole batch is done.)
  hipEvent_t batch_done = nullptr;
  hipStream_t batch_on = nullptr;
  // Recorded once the last launch's planes are built: prepare() starts the next source's intra scores
  // there, beside the motion search. (Recorded before level 0 instead, the launch stream idled a few
  // microseconds at that event between level 1 and level 0: about 1 % slower, measured.)
  hipEvent_t deep_ev = nullptr;
  bool deep_pending = false;
  std::string name;
  int row_streams = 1;
  int row_groups = kDefaultRowGroups;  // the wavefront form's row groups
  bool resident = false;
  // Launches of resident engines captured as a HIP graph and replayed (celv2hip_set_graph): the
  // executable graph of the last launch run on this engine's stream, updated in place when the next
  // launch has the same shape, instantiated again when it has not.
  bool graph_on = false;
  hipGraphExec_t gexec = nullptr;
  long long graph_launches = 0, graph_builds = 0;
  bool counting = false;  // count non-repeated inter candidates (costs the inter kernel LDS atomics)
  SrcSlot srcs[2];
  int cur_set = 0;
  std::vector<std::unique_ptr<RefSlot>> slots;
  long long launches = 0;
  uint32_t* fields[kMaxRefs][kLevels] = {};
  int fcols[kLevels], frows[kLevels];
  uint2* best = nullptr;
  uint8_t* kmap = nullptr;       // per-superblock reference limits on the device (when kmap_on)
  bool kmap_on = false;
  uint8_t* smap = nullptr;       // per-superblock motion-search modes on the device (when smap_on)
  bool smap_on = false;
  uint8_t* imask = nullptr;      // per-superblock intra mode groups on the device (when imask_on)
  bool imask_on = false;
  uint8_t* nmask = nullptr;      // the pruner's intra mode groups of the last launch
  int64_t* sb_cost = nullptr;    // per-superblock decision cost of the last launch
  // The learned pruner (celv2hip_set_pruner) and the per-superblock features it reads.
  const celv2hip_pruner* pruner = nullptr;
  float prune_threshold = 0.5f, prune_band = 0.f;
  // > 0: the pruner also chooses each superblock's intra mode groups (celv2hip_set_pruner_groups);
  // the source's intra scores then run after the network, on the side stream (intra_ev behind them).
  float prune_gthreshold = 0.f;
  // > 0: the pruner also chooses each superblock's motion-search modes (celv2hip_set_pruner_search).
  float prune_sthreshold = 0.f;
  bool modes_last = false;  // the last launch's levels 1 and 0 obeyed per-superblock search modes (smap)
  hipEvent_t infer_ev = nullptr, intra_ev = nullptr;
  bool intra_ev_pending = false;
  bool prune_groups() const { return pruner && prune_gthreshold > 0.f; }
  bool want_features = false, features_valid = false;
  bool limits_last = false;  // the last launch obeyed per-superblock reference limits (kmap)
  uint32_t* cost2[kMaxRefs] = {};  // level-2 best search cost per block and reference
  CnnWork infer;                   // features (infer.x) and the network's activations, one frame
  void* rec_dev = nullptr;   // celv2_sb per superblock
  void* rec_host = nullptr;  // pinned
  std::vector<int> order;
  std::vector<hipEvent_t> group_done;  // one per download group (at most one per row)
  std::vector<int> row_group;          // the group that downloads row r in the last launch
  // Per-launch stage events in a ring (kRing launches): start, references built, motion search,
  // inter scores, finalize, end, and (6) the motion search's level 0 starting; and per prepare():
  // intra scores start / end.
  static constexpr int kRing = 512;
  // Stage events are recorded only with timing on (celv2hip_set_timing): each record between two
  // kernels costs the stream a few microseconds. Launches since the last reset without them.
  bool timing = false;
  long long window_untimed = 0;
  std::vector<std::array<hipEvent_t, 7>> ring;
  std::vector<std::array<hipEvent_t, 2>> pring;
  hipEvent_t* ev = nullptr;  // the current launch's 6 events
  long long prepares = 0, window_launch = 0, window_prepare = 0;
  // Device event counters, each 64 slots 16 words apart (kernels/common.h): counter k at
  // counters + k * kCounterWords: 0 unique MVs scored, 1 compound blocks, 2 inter candidates,
  // 3 pruner fallbacks, 4 ME centres searched.
  unsigned long long* counters = nullptr;
  unsigned long long* counters_host = nullptr;
  int nrefs_last = 0;
  bool launched = false;
  std::vector<RefSlot*> last_refs;  // the last launch's reference slots, nearest first

  ~celv2hip_engine() {
    if (device >= 0) (void)hipSetDevice(device);
    if (batch_on) (void)hipEventSynchronize(batch_done);
    if (stream) (void)hipStreamSynchronize(stream);
    if (side) (void)hipStreamSynchronize(side);
    if (aux) (void)hipStreamSynchronize(aux);
    if (gexec) (void)hipGraphExecDestroy(gexec);
    auto free_up = [](Upload& u) {
      if (u.host) (void)hipHostFree(u.host);
      if (u.dev) (void)hipFree(u.dev);
      if (u.done) (void)hipEventDestroy(u.done);
    };
    for (SrcSlot& s : srcs) {
      if (s.p.mem) (void)hipFree(s.p.mem);
      if (s.intra) (void)hipFree(s.intra);
      if (s.imask_used) (void)hipFree(s.imask_used);
      if (s.side_done) (void)hipEventDestroy(s.side_done);
      if (s.main_done) (void)hipEventDestroy(s.main_done);
      if (s.pyr_done) (void)hipEventDestroy(s.pyr_done);
      free_up(s.up);
    }
    for (auto& s : slots) {
      if (s->p.mem) (void)hipFree(s->p.mem);
      free_up(s->up);
    }
    for (auto& r : fields)
      for (uint32_t* f : r)
        if (f) (void)hipFree(f);
    if (best) (void)hipFree(best);
    if (kmap) (void)hipFree(kmap);
    if (smap) (void)hipFree(smap);
    if (imask) (void)hipFree(imask);
    if (nmask) (void)hipFree(nmask);
    for (uint32_t* c : cost2)
      if (c) (void)hipFree(c);
    cnn_free(infer);
    if (sb_cost) (void)hipFree(sb_cost);
    if (rec_dev) (void)hipFree(rec_dev);
    if (rec_host) (void)hipHostFree(rec_host);
    if (counters) (void)hipFree(counters);
    if (counters_host) (void)hipHostFree(counters_host);
    for (hipEvent_t e : group_done) (void)hipEventDestroy(e);
    for (hipEvent_t e : rows_ev) (void)hipEventDestroy(e);
    for (hipEvent_t e : inter_ev) (void)hipEventDestroy(e);
    if (copy_done) (void)hipEventDestroy(copy_done);
    for (auto& a : ring)
      for (hipEvent_t e : a) (void)hipEventDestroy(e);
    for (auto& a : pring)
      for (hipEvent_t e : a) (void)hipEventDestroy(e);
    if (stream) (void)hipStreamDestroy(stream);
    if (side) (void)hipStreamDestroy(side);
    if (aux) (void)hipStreamDestroy(aux);
    if (sub_start) (void)hipEventDestroy(sub_start);
    if (sub_done) (void)hipEventDestroy(sub_done);
    if (order_ev) (void)hipEventDestroy(order_ev);
    if (batch_done) (void)hipEventDestroy(batch_done);
    if (infer_ev) (void)hipEventDestroy(infer_ev);
    if (intra_ev) (void)hipEventDestroy(intra_ev);
    if (deep_ev) (void)hipEventDestroy(deep_ev);
  }

  void sync_all() {
    if (batch_on) check(hipEventSynchronize(batch_done), "hipEventSynchronize batch");
    check(hipStreamSynchronize(stream), "hipStreamSynchronize");
    check(hipStreamSynchronize(side), "hipStreamSynchronize side");
    check(hipStreamSynchronize(aux), "hipStreamSynchronize aux");
  }

  size_t nsb() const { return static_cast<size_t>(sb_cols) * sb_rows; }

  // A pyramid is set up whole or not at all: a layout that fails frees its memory and leaves p empty.
  void ensure_pyramid(Pyramid& p, bool phases) {
    if (p.mem) return;
    Pyramid q;
    q.bytes = pyramid_bytes(pw, ph, phases);
    check(hipMalloc(&q.mem, q.bytes), "hipMalloc pyramid");
    try {
      layout(q, pw, ph, phases);
    } catch (...) {
      (void)hipFree(q.mem);
      throw;
    }
    p = q;
  }

  // Puts a frame's luma where the import kernel reads it: in place if on the device, else through
  // the slot's pinned staging buffer (copied now, so the caller may reuse its memory at once).
  const void* stage(const celv2_luma& f, Upload& up, hipStream_t st) {
    const bool wide = f.y16 != nullptr;
    if (!wide && !f.y8) throw Failure(CELV2_ERR_ARG, "frame has no luma");
    if (wide && (f.bit_depth < 9 || f.bit_depth > 16)) throw Failure(CELV2_ERR_ARG, "bit depth must be 9..16 with 16-bit samples");
    if (f.on_device) return wide ? static_cast<const void*>(f.y16) : static_cast<const void*>(f.y8);
    const size_t n = static_cast<size_t>(pw) * ph * (wide ? 2 : 1);
    if (up.bytes < n) {
      if (up.host) check(hipHostFree(up.host), "hipHostFree");
      if (up.dev) check(hipFree(up.dev), "hipFree");
      up.host = up.dev = nullptr;
      check(hipHostMalloc(&up.host, n, hipHostMallocDefault), "hipHostMalloc staging");
      check(hipMalloc(&up.dev, n), "hipMalloc staging");
      up.bytes = n;
      if (!up.done) check(hipEventCreateWithFlags(&up.done, hipEventDisableTiming), "hipEventCreate");
    } else if (up.done) {
      check(hipEventSynchronize(up.done), "hipEventSynchronize staging");
    }
    std::memcpy(up.host, wide ? static_cast<const void*>(f.y16) : static_cast<const void*>(f.y8), n);
    check(hipMemcpyAsync(up.dev, up.host, n, hipMemcpyHostToDevice, st), "hipMemcpyAsync upload");
    check(hipEventRecord(up.done, st), "hipEventRecord");
    return up.dev;
  }

  // A plane build for the batched plane launches: import (rounding >8-bit luma), the 15 phase planes
  // of a reference, the 4 coarser levels.
  struct BuildJob {
    PyramidItem pyr;
    Pyramid* p;
    bool phases;
  };
  BuildJob build_job(const celv2_luma& f, Pyramid& p, Upload& up, bool phases, hipStream_t st) {
    ensure_pyramid(p, phases);
    const void* src = stage(f, up, st);
    PyramidItem it{src, pw, f.y16 ? f.bit_depth : 8, f.y16 ? 1 : 0, {}};
    for (int l = 0; l < kLevels; l++) it.lv[l] = p.lv[l];
    return BuildJob{it, &p, phases};
  }

  IntraItem intra_item(SrcSlot& s) {
    if (!s.intra) check(hipMalloc(&s.intra, nsb() * kBlocks * kModes * sizeof(uint32_t)), "hipMalloc intra");
    if (!s.imask_used) check(hipMalloc(&s.imask_used, nsb()), "hipMalloc intra masks");
    return IntraItem{s.p.lv[0], sb_cols, 0, sb_rows, s.intra, imask_on ? imask : nullptr, s.imask_used};
  }

  // Orders stream st's coming work on slot s after the other stream's last work on it (the side
  // stream is prepare()'s, any other stream is a launch's).
  void claim(SrcSlot& s, hipStream_t st) {
    if (st == side && s.main_pending) check(hipStreamWaitEvent(side, s.main_done, 0), "hipStreamWaitEvent");
    if (st != side && s.side_pending) check(hipStreamWaitEvent(st, s.side_done, 0), "hipStreamWaitEvent");
  }
  void release(SrcSlot& s, hipStream_t st) {
    if (st == side) check(hipEventRecord(s.side_done, side), "hipEventRecord"), s.side_pending = true;
    else check(hipEventRecord(s.main_done, st), "hipEventRecord"), s.main_pending = true;
  }
};

namespace {

// The kernels' constant tables on `device` (the current device), uploaded by the first engine created
// there; a failed upload leaves the device's flag unset, so the next creation tries again.
constexpr int kMaxDevices = 64;
void upload_tables_once(int device) {
  static std::once_flag flags[kMaxDevices];
  if (device >= kMaxDevices) throw Failure(CELV2_ERR_DEVICE, "HipEngine: device index above 63");
  std::call_once(flags[device], [] {
    check(upload_intra_tiles(), "intra tables");
    check(upload_inter_tables(), "inter tables");
    check(upload_final_tiles(), "compound tables");
  });
}

// The checks of an entry point that only waits on events or reads host memory: no current device needed.
int validate_host(celv2hip_engine* e) {
  if (!e) {
    g_error = "null engine";
    return CELV2_ERR_ARG;
  }
  return CELV2_OK;
}

int validate(celv2hip_engine* e) {
  if (!e) {
    g_error = "null engine";
    return CELV2_ERR_ARG;
  }
  if (hipSetDevice(e->device) != hipSuccess) {
    g_error = "hipSetDevice failed";
    return CELV2_ERR_DEVICE;
  }
  return CELV2_OK;
}

}  // namespace

extern "C" {

const char* celv2hip_last_error(void) { return g_error.c_str(); }

int celv2hip_device_count(int* n) {
  return guarded([&] {
    int c = 0;
    if (hipGetDeviceCount(&c) != hipSuccess) c = 0;
    *n = c;
  });
}

int celv2hip_create(int width, int height, int sb_size, int device, celv2hip_engine** out) {
  if (!out) return CELV2_ERR_ARG;
  *out = nullptr;
  return guarded([&] {
    if (width <= 0 || height <= 0) throw Failure(CELV2_ERR_SIZE, "HipEngine: frame size must be positive");
    if (sb_size != 64) throw Failure(CELV2_ERR_SIZE, "HipEngine: only 64x64 superblocks are supported");
    // The kernels address a reference's 16 quarter-pel planes with 32-bit signed byte offsets from
    // its level-0 corner (buffer resources of 2^31 - 1 bytes): all 16 must span less than that.
    // Checked before any device call.
    {
      int stride;
      const unsigned long long span = static_cast<unsigned long long>(plane_bytes((width + 7) & ~7, (height + 7) & ~7, &stride)) * kPhases;
      if (span > 0x7fffffffull)
        throw Failure(CELV2_ERR_SIZE, "HipEngine: frame too large (a reference's 16 quarter-pel planes must span under 2 GiB; about 11,500 x 11,500 at most)");
    }
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) throw Failure(CELV2_ERR_DEVICE, "HipEngine: no HIP device");
    if (device < 0 || device >= count) throw Failure(CELV2_ERR_DEVICE, "HipEngine: no such HIP device");
    check(hipSetDevice(device), "hipSetDevice");
    auto e = std::make_unique<celv2hip_engine>();
    e->width = width, e->height = height, e->sb_size = sb_size, e->device = device;
    e->pw = (width + 7) & ~7, e->ph = (height + 7) & ~7;
    e->sb_cols = (width + 63) / 64, e->sb_rows = (height + 63) / 64;
    hipDeviceProp_t prop;
    check(hipGetDeviceProperties(&prop, device), "hipGetDeviceProperties");
    e->name = prop.name[0] ? prop.name : prop.gcnArchName;
    if (e->name.empty()) e->name = "HIP device " + std::to_string(device);
    check(hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking), "hipStreamCreate");
    check(hipStreamCreateWithFlags(&e->side, hipStreamNonBlocking), "hipStreamCreate");
    check(hipStreamCreateWithFlags(&e->aux, hipStreamNonBlocking), "hipStreamCreate");
    check(hipEventCreateWithFlags(&e->copy_done, hipEventDisableTiming), "hipEventCreate");
    check(hipEventCreateWithFlags(&e->sub_start, hipEventDisableTiming), "hipEventCreate");
    check(hipEventCreateWithFlags(&e->sub_done, hipEventDisableTiming), "hipEventCreate");
    check(hipEventCreateWithFlags(&e->order_ev, hipEventDisableTiming), "hipEventCreate");
    check(hipEventCreateWithFlags(&e->batch_done, hipEventDisableTiming), "hipEventCreate");
    check(hipEventCreateWithFlags(&e->infer_ev, hipEventDisableTiming), "hipEventCreate");
    check(hipEventCreateWithFlags(&e->intra_ev, hipEventDisableTiming), "hipEventCreate");
    check(hipEventCreateWithFlags(&e->deep_ev, hipEventDisableTiming), "hipEventCreate");
    for (SrcSlot& s : e->srcs) {
      check(hipEventCreateWithFlags(&s.side_done, hipEventDisableTiming), "hipEventCreate");
      check(hipEventCreateWithFlags(&s.main_done, hipEventDisableTiming), "hipEventCreate");
      check(hipEventCreateWithFlags(&s.pyr_done, hipEventDisableTiming), "hipEventCreate");
    }
    int w = e->pw, h = e->ph;
    for (int l = 0; l < kLevels; l++) {
      e->fcols[l] = (w + 7) / 8, e->frows[l] = (h + 7) / 8;
      w /= 2, h /= 2;
    }
    check(hipMalloc(&e->best, e->nsb() * kMaxRefs * kBlocks * sizeof(uint2)), "hipMalloc best");
    check(hipMalloc(&e->kmap, e->nsb()), "hipMalloc kmap");
    check(hipMalloc(&e->smap, e->nsb()), "hipMalloc search modes");
    check(hipMalloc(&e->sb_cost, e->nsb() * sizeof(int64_t)), "hipMalloc sb_cost");
    check(hipMalloc(&e->rec_dev, e->nsb() * sizeof(celv2_sb)), "hipMalloc records");
    check(hipHostMalloc(&e->rec_host, e->nsb() * sizeof(celv2_sb), hipHostMallocDefault), "hipHostMalloc records");
    check(hipMalloc(&e->counters, kCounterBytes), "hipMalloc counters");
    check(hipHostMalloc(reinterpret_cast<void**>(&e->counters_host), kCounterBytes, hipHostMallocDefault), "hipHostMalloc counters");
    e->group_done.resize(e->sb_rows);
    e->row_group.assign(e->sb_rows, 0);
    e->rows_ev.resize(e->sb_rows);
    for (hipEvent_t& ev : e->rows_ev) check(hipEventCreateWithFlags(&ev, hipEventDisableTiming), "hipEventCreate");
    e->inter_ev.resize(e->sb_rows);
    for (hipEvent_t& ev : e->inter_ev) check(hipEventCreateWithFlags(&ev, hipEventDisableTiming), "hipEventCreate");
    for (hipEvent_t& ev : e->group_done) check(hipEventCreateWithFlags(&ev, hipEventDisableTiming), "hipEventCreate");
    e->ring.resize(celv2hip_engine::kRing);
    e->pring.resize(celv2hip_engine::kRing);
    for (auto& a : e->ring)
      for (hipEvent_t& ev : a) check(hipEventCreate(&ev), "hipEventCreate");
    for (auto& a : e->pring)
      for (hipEvent_t& ev : a) check(hipEventCreate(&ev), "hipEventCreate");
    e->ev = e->ring[0].data();
    check(hipMemset(e->counters, 0, kCounterBytes), "hipMemset counters");
    upload_tables_once(device);
    *out = e.release();
  });
}

int celv2hip_destroy(celv2hip_engine* e) {
  delete e;
  return CELV2_OK;
}

int celv2hip_device_name(const celv2hip_engine* e, char* buf, int cap) {
  if (!e || !buf || cap <= 0) return CELV2_ERR_ARG;
  std::snprintf(buf, static_cast<size_t>(cap), "%s", e->name.c_str());
  return CELV2_OK;
}

int celv2hip_set_row_streams(celv2hip_engine* e, int n) {
  if (!e) return CELV2_ERR_ARG;
  e->row_streams = n;
  return CELV2_OK;
}

int celv2hip_set_row_groups(celv2hip_engine* e, int groups) {
  if (!e) return CELV2_ERR_ARG;
  if (groups < 0 || groups > 64) {
    g_error = "row groups are 1..64 (0: the default)";
    return CELV2_ERR_ARG;
  }
  e->row_groups = groups == 0 ? kDefaultRowGroups : groups;
  return CELV2_OK;
}

int celv2hip_row_groups(const celv2hip_engine* e) { return e ? e->row_groups : -CELV2_ERR_ARG; }

int celv2hip_set_resident(celv2hip_engine* e, int on) {
  if (!e) return CELV2_ERR_ARG;
  e->resident = on != 0;
  return CELV2_OK;
}

int celv2hip_set_graph(celv2hip_engine* e, int on) {
  if (!e) return CELV2_ERR_ARG;
  e->graph_on = on != 0;
  return CELV2_OK;
}

int celv2hip_graph_stats(const celv2hip_engine* e, long long* launches, long long* builds) {
  if (!e) return CELV2_ERR_ARG;
  if (launches) *launches = e->graph_launches;
  if (builds) *builds = e->graph_builds;
  return CELV2_OK;
}

int celv2hip_set_ref_limits(celv2hip_engine* e, const uint8_t* k, int count) {
  if (int rc = validate(e)) return rc;
  return guarded([&] {
    if (!k || count == 0) {
      e->kmap_on = false;
      return;
    }
    if (count != static_cast<int>(e->nsb())) throw Failure(CELV2_ERR_SIZE, "one limit per superblock");
    for (int i = 0; i < count; i++)
      if (k[i] < 1 || k[i] > kMaxRefs) throw Failure(CELV2_ERR_ARG, "reference limits are 1..7");
    e->sync_all();
    check(hipMemcpy(e->kmap, k, static_cast<size_t>(count), hipMemcpyHostToDevice), "hipMemcpy kmap");
    e->kmap_on = true;
  });
}

int celv2hip_set_search_modes(celv2hip_engine* e, const uint8_t* modes, int count) {
  if (int rc = validate(e)) return rc;
  return guarded([&] {
    if (!modes || count == 0) {
      e->smap_on = false;
      return;
    }
    if (count != static_cast<int>(e->nsb())) throw Failure(CELV2_ERR_SIZE, "one search mode per superblock");
    for (int i = 0; i < count; i++)
      if (modes[i] > 3) throw Failure(CELV2_ERR_ARG, "search modes are 2-bit level sets");
    e->sync_all();
    check(hipMemcpy(e->smap, modes, static_cast<size_t>(count), hipMemcpyHostToDevice), "hipMemcpy search modes");
    e->smap_on = true;
  });
}

int celv2hip_set_intra_masks(celv2hip_engine* e, const uint8_t* masks, int count) {
  if (int rc = validate(e)) return rc;
  return guarded([&] {
    if (!masks || count == 0) {
      e->imask_on = false;
      return;
    }
    if (count != static_cast<int>(e->nsb())) throw Failure(CELV2_ERR_SIZE, "one mask per superblock");
    for (int i = 0; i < count; i++)
      if (masks[i] > kAllGroups) throw Failure(CELV2_ERR_ARG, "intra masks are 4-bit group sets");
    e->sync_all();
    if (!e->imask) check(hipMalloc(&e->imask, e->nsb()), "hipMalloc intra masks");
    check(hipMemcpy(e->imask, masks, static_cast<size_t>(count), hipMemcpyHostToDevice), "hipMemcpy intra masks");
    e->imask_on = true;
  });
}

long long celv2hip_fetch_sb_costs(celv2hip_engine* e, int64_t* out, long long cap) {
  if (int rc = validate(e)) return rc;
  long long n = 0;
  const int rc = guarded([&] {
    if (!e->launched) throw Failure(CELV2_ERR_STATE, "no launch");
    n = static_cast<long long>(e->nsb());
    if (!out || cap < n) throw Failure(CELV2_ERR_SIZE, "cost buffer too small");
    e->sync_all();
    check(hipMemcpy(out, e->sb_cost, static_cast<size_t>(n) * sizeof(int64_t), hipMemcpyDeviceToHost), "hipMemcpy sb_cost");
  });
  return rc ? rc : n;
}

int celv2hip_pruner_param_count(void) { return kParams; }

int celv2hip_pruner_create(int device, int max_positions, unsigned seed, celv2hip_pruner** out) {
  if (!out || max_positions <= 0) return CELV2_ERR_ARG;
  *out = nullptr;
  return guarded([&] {
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) throw Failure(CELV2_ERR_DEVICE, "pruner: no HIP device");
    if (device < 0 || device >= count) throw Failure(CELV2_ERR_DEVICE, "pruner: no such HIP device");
    check(hipSetDevice(device), "hipSetDevice");
    auto p = std::make_unique<celv2hip_pruner>();
    p->device = device;
    check(cnn_scratch_init(device), "hipGetSymbolAddress pruner scratch");
    check(hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking), "hipStreamCreate");
    check(hipMalloc(&p->params, kParams * sizeof(float)), "hipMalloc params");
    check(hipMalloc(&p->pb, kParams * sizeof(uint16_t)), "hipMalloc params");
    check(hipMalloc(&p->m1, kParams * sizeof(float)), "hipMalloc adam");
    check(hipMalloc(&p->m2, kParams * sizeof(float)), "hipMalloc adam");
    check(hipMalloc(&p->grads, kParams * sizeof(float)), "hipMalloc grads");
    cnn_alloc(p->work, static_cast<size_t>(max_positions));
    // He-normal weights (fan-in = the im2col row), zero biases.
    std::vector<float> w(kParams, 0.f);
    std::mt19937 rng(seed);
    auto fill = [&](int off, int fan_in, int count) {
      std::normal_distribution<float> d(0.f, std::sqrt(2.f / static_cast<float>(fan_in)));
      for (int k = 0; k < count; k++) w[off + k] = d(rng);
    };
    fill(kW1, 9 * kFeatures, 9 * kFeatures * kHidden);
    fill(kW2, kK2, kK2 * kHidden);
    fill(kW3, kHidden, kHidden * kOut);
    check(hipMemcpy(p->params, w.data(), kParams * sizeof(float), hipMemcpyHostToDevice), "hipMemcpy params");
    check(hipMemset(p->m1, 0, kParams * sizeof(float)), "hipMemset");
    check(hipMemset(p->m2, 0, kParams * sizeof(float)), "hipMemset");
    cnn_params_to_bf16(p->params, p->pb, p->stream);
    check(hipStreamSynchronize(p->stream), "pruner init");
    *out = p.release();
  });
}

int celv2hip_pruner_destroy(celv2hip_pruner* p) {
  delete p;
  return CELV2_OK;
}

int celv2hip_pruner_get_params(celv2hip_pruner* p, float* out, int cap) {
  if (!p || !out || cap < kParams) return CELV2_ERR_ARG;
  return guarded([&] {
    check(hipSetDevice(p->device), "hipSetDevice");
    check(hipStreamSynchronize(p->stream), "sync");
    check(hipMemcpy(out, p->params, kParams * sizeof(float), hipMemcpyDeviceToHost), "hipMemcpy params");
  });
}

int celv2hip_pruner_set_params(celv2hip_pruner* p, const float* in, int n) {
  if (!p || !in || n != kParams) return CELV2_ERR_ARG;
  return guarded([&] {
    check(hipSetDevice(p->device), "hipSetDevice");
    check(hipStreamSynchronize(p->stream), "sync");
    check(hipMemcpy(p->params, in, kParams * sizeof(float), hipMemcpyHostToDevice), "hipMemcpy params");
    check(hipMemset(p->m1, 0, kParams * sizeof(float)), "hipMemset");
    check(hipMemset(p->m2, 0, kParams * sizeof(float)), "hipMemset");
    p->step = 0;
    cnn_params_to_bf16(p->params, p->pb, p->stream);
    check(hipStreamSynchronize(p->stream), "sync");
  });
}

namespace {

// Uploads b frames of h x w superblock features (and labels) into the pruner's workspace.
void pruner_upload(celv2hip_pruner* p, const float* feats, const uint8_t* kstar, int b, int h, int w) {
  if (!p || !feats || b <= 0 || h <= 0 || w <= 0) throw Failure(CELV2_ERR_ARG, "pruner: bad arguments");
  const size_t m = static_cast<size_t>(b) * h * w;
  if (m > p->work.cap) throw Failure(CELV2_ERR_SIZE, "pruner: more positions than the workspace holds");
  check(hipSetDevice(p->device), "hipSetDevice");
  check(hipMemcpyAsync(p->work.x, feats, m * kFeatures * sizeof(float), hipMemcpyHostToDevice, p->stream), "hipMemcpy feats");
  if (kstar) check(hipMemcpyAsync(p->work.kstar, kstar, m, hipMemcpyHostToDevice, p->stream), "hipMemcpy labels");
}

float pruner_loss(celv2hip_pruner* p, size_t m) {
  const size_t parts = (m * kOut + 255) / 256;
  std::vector<float> part(parts);
  check(hipMemcpyAsync(part.data(), p->work.loss_part, parts * sizeof(float), hipMemcpyDeviceToHost, p->stream), "hipMemcpy loss");
  check(hipStreamSynchronize(p->stream), "sync");
  double l = 0;
  for (float v : part) l += v;
  return static_cast<float>(l / static_cast<double>(m * kOut));
}

}  // namespace

int celv2hip_pruner_forward(celv2hip_pruner* p, const float* feats, int b, int h, int w, float* logits) {
  if (!logits) return CELV2_ERR_ARG;
  return guarded([&] {
    pruner_upload(p, feats, nullptr, b, h, w);
    cnn_infer(p->work.x, b, h, w, p->params, p->pb, p->work.z3, kMaxRefs, 0.5f, 0.f, nullptr, nullptr, p->stream);
    const size_t m = static_cast<size_t>(b) * h * w;
    check(hipMemcpyAsync(logits, p->work.z3, m * kOut * sizeof(float), hipMemcpyDeviceToHost, p->stream), "hipMemcpy logits");
    check(hipStreamSynchronize(p->stream), "sync");
  });
}

int celv2hip_pruner_forward_static(const celv2hip_pruner* p, const float* feats, int b, int h, int w, float* logits) {
  if (!p || !feats || !logits || b <= 0 || h <= 0 || w <= 0) {
    g_error = "pruner_forward_static: bad arguments";
    return CELV2_ERR_ARG;
  }
  return guarded([&] {
    const size_t m = static_cast<size_t>(b) * h * w;
    if (m > static_cast<size_t>(kStaticPositions))
      throw Failure(CELV2_ERR_SIZE, "pruner_forward_static: at most 16384 superblock positions (b x h x w) per call");
    check(hipSetDevice(p->device), "hipSetDevice");
    float* const s = cnn_scratch(p->device);
    if (!s) throw Failure(CELV2_ERR_STATE, "pruner_forward_static: no scratch on this device");
    float* const x = s;
    float* const z = s + static_cast<size_t>(kStaticPositions) * kFeatures;
    check(hipMemcpyAsync(x, feats, m * kFeatures * sizeof(float), hipMemcpyHostToDevice, p->stream), "hipMemcpy feats");
    cnn_infer(x, b, h, w, p->params, p->pb, z, kMaxRefs, 0.5f, 0.f, nullptr, nullptr, p->stream);
    check(hipGetLastError(), "pruner kernels");
    check(hipMemcpyAsync(logits, z, m * kOut * sizeof(float), hipMemcpyDeviceToHost, p->stream), "hipMemcpy logits");
    check(hipStreamSynchronize(p->stream), "sync");
  });
}

int celv2hip_pruner_gradients(celv2hip_pruner* p, const float* feats, const uint8_t* kstar, int b, int h, int w,
                              float* grads, float* loss) {
  if (!kstar || !grads) return CELV2_ERR_ARG;
  return guarded([&] {
    pruner_upload(p, feats, kstar, b, h, w);
    cnn_forward(p->work, p->params, p->pb, b, h, w, p->stream);
    cnn_backward(p->work, p->pb, p->grads, b, h, w, p->stream);
    check(hipMemcpyAsync(grads, p->grads, kParams * sizeof(float), hipMemcpyDeviceToHost, p->stream), "hipMemcpy grads");
    const float l = pruner_loss(p, static_cast<size_t>(b) * h * w);
    if (loss) *loss = l;
  });
}

int celv2hip_pruner_train_step(celv2hip_pruner* p, const float* feats, const uint8_t* kstar, int b, int h, int w,
                               float lr, float weight_decay, float* loss) {
  if (!kstar) return CELV2_ERR_ARG;
  return guarded([&] {
    pruner_upload(p, feats, kstar, b, h, w);
    cnn_forward(p->work, p->params, p->pb, b, h, w, p->stream);
    cnn_backward(p->work, p->pb, p->grads, b, h, w, p->stream);
    cnn_adamw(p->params, p->grads, p->m1, p->m2, p->pb, ++p->step, lr, weight_decay, p->stream);
    check(hipGetLastError(), "pruner kernels");
    const float l = pruner_loss(p, static_cast<size_t>(b) * h * w);
    if (loss) *loss = l;
  });
}

int celv2hip_set_pruner(celv2hip_engine* e, const celv2hip_pruner* p, float threshold, float band) {
  if (int rc = validate(e)) return rc;
  if (p && p->device != e->device) {
    g_error = "pruner and engine on different devices";
    return CELV2_ERR_ARG;
  }
  e->pruner = p, e->prune_threshold = threshold, e->prune_band = band;
  return CELV2_OK;
}

int celv2hip_set_pruner_groups(celv2hip_engine* e, float gthreshold) {
  if (!e) return CELV2_ERR_ARG;
  if (!(gthreshold >= 0.f && gthreshold <= 1.f)) {
    g_error = "group threshold is 0..1 (0: every group)";
    return CELV2_ERR_ARG;
  }
  e->prune_gthreshold = gthreshold;
  return CELV2_OK;
}

int celv2hip_set_pruner_search(celv2hip_engine* e, float sthreshold) {
  if (!e) return CELV2_ERR_ARG;
  if (!(sthreshold >= 0.f && sthreshold <= 1.f)) {
    g_error = "search threshold is 0..1 (0: every search full)";
    return CELV2_ERR_ARG;
  }
  e->prune_sthreshold = sthreshold;
  return CELV2_OK;
}

long long celv2hip_fetch_search_modes(celv2hip_engine* e, uint8_t* out, long long cap) {
  if (int rc = validate(e)) return rc;
  long long n = 0;
  const int rc = guarded([&] {
    n = static_cast<long long>(e->nsb());
    if (!out || cap < n) throw Failure(CELV2_ERR_SIZE, "mode buffer too small");
    if (!e->launched) throw Failure(CELV2_ERR_STATE, "no launch");
    e->sync_all();
    if (e->modes_last) check(hipMemcpy(out, e->smap, static_cast<size_t>(n), hipMemcpyDeviceToHost), "hipMemcpy search modes");
    else std::fill(out, out + n, static_cast<uint8_t>(0));  // every search full
  });
  return rc ? rc : n;
}

int celv2hip_set_features(celv2hip_engine* e, int on) {
  if (!e) return CELV2_ERR_ARG;
  e->want_features = on != 0;
  return CELV2_OK;
}

long long celv2hip_fetch_features(celv2hip_engine* e, float* out, long long cap) {
  if (int rc = validate(e)) return rc;
  long long n = 0;
  const int rc = guarded([&] {
    if (!e->features_valid) throw Failure(CELV2_ERR_STATE, "no features: the last launch had none (set_features or a pruner, and references)");
    n = static_cast<long long>(e->nsb()) * kFeatures;
    if (!out || cap < n) throw Failure(CELV2_ERR_SIZE, "feature buffer too small");
    e->sync_all();
    check(hipMemcpy(out, e->infer.x, static_cast<size_t>(n) * sizeof(float), hipMemcpyDeviceToHost), "hipMemcpy features");
  });
  return rc ? rc : n;
}

long long celv2hip_fetch_ref_limits(celv2hip_engine* e, uint8_t* out, long long cap) {
  if (int rc = validate(e)) return rc;
  long long n = 0;
  const int rc = guarded([&] {
    n = static_cast<long long>(e->nsb());
    if (!out || cap < n) throw Failure(CELV2_ERR_SIZE, "limit buffer too small");
    if (!e->launched) throw Failure(CELV2_ERR_STATE, "no launch");
    e->sync_all();
    if (e->limits_last) check(hipMemcpy(out, e->kmap, static_cast<size_t>(n), hipMemcpyDeviceToHost), "hipMemcpy kmap");
    else std::fill(out, out + n, static_cast<uint8_t>(e->nrefs_last));  // every superblock kept them all
  });
  return rc ? rc : n;
}

long long celv2hip_fetch_intra_masks(celv2hip_engine* e, uint8_t* out, long long cap) {
  if (int rc = validate(e)) return rc;
  long long n = 0;
  const int rc = guarded([&] {
    n = static_cast<long long>(e->nsb());
    if (!out || cap < n) throw Failure(CELV2_ERR_SIZE, "mask buffer too small");
    if (!e->launched) throw Failure(CELV2_ERR_STATE, "no launch");
    const SrcSlot& s = e->srcs[e->cur_set];
    if (!s.imask_used) throw Failure(CELV2_ERR_STATE, "no intra scores");
    e->sync_all();
    check(hipMemcpy(out, s.imask_used, static_cast<size_t>(n), hipMemcpyDeviceToHost), "hipMemcpy intra masks");
  });
  return rc ? rc : n;
}

int celv2hip_set_counting(celv2hip_engine* e, int on) {
  if (!e) return CELV2_ERR_ARG;
  e->counting = on != 0;
  return CELV2_OK;
}

namespace {

// Plane builds of any engines in batched launches: every import with its pyramid (one launch), then
// every reference's phase planes. With `e` (the launching engine), the phase planes go to its aux
// stream after the pyramids and the caller waits for e->sub_done before they are read; returns whether
// any were queued there.
bool run_builds(const std::vector<celv2hip_engine::BuildJob>& jobs, hipStream_t st, celv2hip_engine* e = nullptr) {
  bool on_aux = false;
  for (size_t j0 = 0; j0 < jobs.size(); j0 += kMaxItems) {
    const size_t j1 = std::min(jobs.size(), j0 + kMaxItems);
    Batch<PyramidItem> bp;
    Batch<SubpelItem> bs;
    for (size_t j = j0; j < j1; j++) {
      add_pyramid(bp, jobs[j].pyr);
      if (jobs[j].phases) add_subpel(bs, SubpelItem{jobs[j].p->lv[0], jobs[j].p->phase_bytes});
    }
    launch_pyramid(bp, st);
    if (e && bs.total() > 0) {
      check(hipEventRecord(e->sub_start, st), "hipEventRecord");
      check(hipStreamWaitEvent(e->aux, e->sub_start, 0), "hipStreamWaitEvent");
      launch_subpel(bs, e->aux);
      check(hipEventRecord(e->sub_done, e->aux), "hipEventRecord");
      on_aux = true;
    } else {
      launch_subpel(bs, st);
    }
  }
  check(hipGetLastError(), "plane kernels");
  return on_aux;
}

// Intra scores of the given engines' source slots, with each engine's prepare-ring events around.
void run_intra(const std::vector<std::pair<celv2hip_engine*, SrcSlot*>>& items, hipStream_t st) {
  for (size_t j0 = 0; j0 < items.size(); j0 += kMaxItems) {
    const size_t j1 = std::min(items.size(), j0 + kMaxItems);
    Batch<IntraItem> b;
    std::vector<std::array<hipEvent_t, 2>*> rings;
    for (size_t j = j0; j < j1; j++) {
      celv2hip_engine* e = items[j].first;
      add_intra(b, e->intra_item(*items[j].second));
      rings.push_back(&e->pring[e->prepares++ % celv2hip_engine::kRing]);
    }
    const bool timed = items[j0].first->timing;
    if (timed)
      for (auto* r : rings) check(hipEventRecord((*r)[0], st), "hipEventRecord");
    launch_intra(b, st);
    check(hipGetLastError(), "intra kernel");
    if (timed)
      for (auto* r : rings) check(hipEventRecord((*r)[1], st), "hipEventRecord");
  }
}

}  // namespace

int celv2hip_prepare(celv2hip_engine* e, const celv2_luma* next, long long key) {
  if (int rc = validate(e)) return rc;
  return guarded([&] {
    if (key < 0 || !next) return;
    // The slot the last launch does not use, filled on the side stream once the main stream's
    // earlier launches are done with it; the launch that takes it waits for the side stream.
    SrcSlot& s = e->srcs[1 - e->cur_set];
    s.key = -1;
    e->claim(s, e->side);
    run_builds({e->build_job(*next, s.p, s.up, false, e->side)}, e->side);
    check(hipEventRecord(s.pyr_done, e->side), "hipEventRecord"), s.pyr_pending = true;
    s.intra_deferred = e->prune_groups();
    if (!s.intra_deferred) {
      if (e->deep_pending) check(hipStreamWaitEvent(e->side, e->deep_ev, 0), "hipStreamWaitEvent");
      run_intra({{e, &s}}, e->side);
    }
    e->release(s, e->side);
    s.key = key;
  });
}

namespace {

struct LaunchIn {
  celv2hip_engine* e;
  const celv2_luma* src;
  long long src_key;
  const celv2_luma* refs;
  const long long* ref_keys;
  int nrefs;
  int64_t lam;
};

// Ends a stream capture that a throw left open (the captured graph is dropped).
struct CaptureGuard {
  hipStream_t s = nullptr;
  ~CaptureGuard() {
    if (!s) return;
    hipGraph_t g = nullptr;
    (void)hipStreamEndCapture(s, &g);
    if (g) (void)hipGraphDestroy(g);
  }
};

// One frame of each engine: every stage is one launch (or one per kMaxItems engines) on the first
// engine's stream; each engine's own stream orders its earlier work before and its later work after.
void launch_batch(const LaunchIn* in, int n) {
  celv2hip_engine* const e0 = in[0].e;
  const hipStream_t S = e0->stream;
  for (int i = 0; i < n; i++) {
    const LaunchIn& x = in[i];
    if (!x.src) throw Failure(CELV2_ERR_ARG, "HipEngine: no source");
    if (x.nrefs < 0 || x.nrefs > kMaxRefs) throw Failure(CELV2_ERR_SIZE, "HipEngine: too many references");
    if (x.nrefs > 0 && !x.refs) throw Failure(CELV2_ERR_ARG, "HipEngine: no references");
    if (x.e->device != e0->device) throw Failure(CELV2_ERR_ARG, "batched engines must share a device");
    for (int k = 0; k < i; k++)
      if (in[k].e == x.e) throw Failure(CELV2_ERR_ARG, "an engine appears twice in one batch");
  }
  for (int i = 0; i < n; i++) {
    celv2hip_engine* e = in[i].e;
    if (i > 0) {  // the engine's own earlier launches
      check(hipEventRecord(e->order_ev, e->stream), "hipEventRecord");
      check(hipStreamWaitEvent(S, e->order_ev, 0), "hipStreamWaitEvent");
    }
    if (e->batch_on && e->batch_on != S) check(hipStreamWaitEvent(S, e->batch_done, 0), "hipStreamWaitEvent");
  }
  std::vector<celv2hip_engine::BuildJob> jobs;
  std::vector<std::pair<celv2hip_engine*, SrcSlot*>> intra_now;
  // Engines whose source's intra scores wait for the network's mode groups (after level 2).
  std::vector<char> intra_after(n);
  std::vector<SrcSlot*> cur(n);
  std::vector<std::vector<RefSlot*>> used(n);
  std::vector<char> use_kmap(n), use_smap(n);
  for (int i = 0; i < n; i++) use_kmap[i] = in[i].e->kmap_on, use_smap[i] = in[i].e->smap_on;
  // Prepared sources whose intra scores the launch stream waits for only before finalize.
  std::vector<char> late_claim(n);
  // One engine with one row stream whose records go to the host (the wavefront form): level 0 of the
  // motion search runs in row groups interleaved with the inter scores of the groups above them; each
  // group's finalize runs on the side stream, its record download on the aux stream.
  const bool wavefront = n == 1 && !e0->resident && std::max(1, std::min(e0->row_streams, e0->sb_rows)) == 1;
  // The graph form: every engine resident with its graph on and no stage events. The launch's kernels
  // are captured on S (the aux stream's phase planes joined back in), and the graph replayed; every
  // wait on work outside the launch is placed on S before the capture, every event that work outside
  // waits for is recorded after the replay.
  bool use_graph = true;
  for (int i = 0; i < n; i++) use_graph = use_graph && in[i].e->graph_on && in[i].e->resident && !in[i].e->timing;
  for (int i = 0; i < n; i++) {
    const LaunchIn& x = in[i];
    celv2hip_engine* e = x.e;
    e->ev = e->ring[e->launches % celv2hip_engine::kRing].data();
    if (!e->timing) e->window_untimed++;
    if (e->timing) check(hipEventRecord(e->ev[0], S), "hipEventRecord");
    int use = -1;
    for (int k = 0; k < 2; k++)
      if (x.src_key >= 0 && e->srcs[k].key == x.src_key) use = k;
    SrcSlot& claimed = e->srcs[use < 0 ? 1 - e->cur_set : use];
    if (use >= 0 && (wavefront || n > 1) && !use_graph) {
      // A prepared source: the motion search reads its pyramid; finalize waits for its intra scores
      // (on the side stream in the wavefront form, which computes them there first).
      if (claimed.pyr_pending) check(hipStreamWaitEvent(S, claimed.pyr_done, 0), "hipStreamWaitEvent");
      late_claim[i] = !wavefront;
    } else {
      e->claim(claimed, S);
    }
    const bool groups_here = e->prune_groups() && x.nrefs >= 1;
    if (use < 0) {
      use = 1 - e->cur_set;
      SrcSlot& s = e->srcs[use];
      s.key = -1;
      jobs.push_back(e->build_job(*x.src, s.p, s.up, false, S));
      if (groups_here) intra_after[i] = true;
      else intra_now.push_back({e, &s});
    } else if (claimed.intra_deferred) {
      // prepare() left the intra scores to this launch: after the network, or now with every group.
      if (groups_here) intra_after[i] = true;
      else intra_now.push_back({e, &claimed});
    }
    e->srcs[use].intra_deferred = false;
    e->srcs[use].key = -1;  // taken: a later prepare() may overwrite it
    e->cur_set = use;
    cur[i] = &e->srcs[use];
    const long long now = ++e->launches;
    used[i].assign(x.nrefs, nullptr);
    for (int ri = 0; ri < x.nrefs; ri++) {
      const long long key = x.ref_keys ? x.ref_keys[ri] : -1;
      RefSlot* slot = nullptr;
      for (auto& sl : e->slots)
        if (key >= 0 && sl->key == key) slot = sl.get();
      if (!slot) {
        for (auto& sl : e->slots) {
          const bool wanted = x.ref_keys && std::find(x.ref_keys, x.ref_keys + x.nrefs, sl->key) != x.ref_keys + x.nrefs && sl->key >= 0;
          if (sl->used != now && !wanted) {
            slot = sl.get();
            break;
          }
        }
        if (!slot) e->slots.push_back(std::make_unique<RefSlot>()), slot = e->slots.back().get();
        slot->key = -1;
        jobs.push_back(e->build_job(x.refs[ri], slot->p, slot->up, true, S));
        slot->key = key;
      }
      slot->used = now;
      used[i][ri] = slot;
    }
  }
  CaptureGuard capture;
  if (use_graph) {
    // Buffers a launch allocates on first use, allocated now (no allocation inside a capture), and
    // the waits on earlier downloads.
    for (int i = 0; i < n; i++) {
      celv2hip_engine* e = in[i].e;
      const bool feats = e->pruner || e->want_features;
      for (int r = 0; r < in[i].nrefs; r++) {
        for (int l = 0; l < kLevels; l++)
          if (!e->fields[r][l]) check(hipMalloc(&e->fields[r][l], static_cast<size_t>(e->fcols[l]) * e->frows[l] * sizeof(uint32_t)), "hipMalloc field");
        if (feats && !e->cost2[r]) check(hipMalloc(&e->cost2[r], static_cast<size_t>(e->fcols[2]) * e->frows[2] * sizeof(uint32_t)), "hipMalloc cost2");
      }
      if (feats && in[i].nrefs >= 1 && !e->infer.cap) cnn_alloc(e->infer, e->nsb());
      if (e->copy_pending) check(hipStreamWaitEvent(S, e->copy_done, 0), "hipStreamWaitEvent"), e->copy_pending = false;
    }
    for (auto& it : intra_now) (void)it.first->intra_item(*it.second);
    for (int i = 0; i < n; i++)
      if (intra_after[i]) {
        if (!in[i].e->nmask) check(hipMalloc(&in[i].e->nmask, in[i].e->nsb()), "hipMalloc intra masks");
        (void)in[i].e->intra_item(*cur[i]);
      }
    check(hipStreamBeginCapture(S, hipStreamCaptureModeRelaxed), "hipStreamBeginCapture");
    capture.s = S;
  }
  // Only one batch of jobs per aux stream use: more than kMaxItems jobs build on S itself, and so does
  // the graph form (one chain of kernels: a branch on another stream makes the graph's replay take
  // more of the process's hardware queues, which the engines' own streams then share).
  const bool phases_pending = jobs.size() <= static_cast<size_t>(kMaxItems) && !use_graph ? run_builds(jobs, S, e0) : run_builds(jobs, S);
  run_intra(intra_now, S);
  // The next source's intra scores (prepare(), side stream) may start once this launch's planes are
  // built: the motion search's coarse levels leave most of the GPU idle (about 1 % faster than waiting
  // for level 0, measured).
  for (int i = 0; i < n && !use_graph; i++) {
    check(hipEventRecord(in[i].e->deep_ev, S), "hipEventRecord");
    in[i].e->deep_pending = true;
  }
  for (int i = 0; i < n; i++) if (in[i].e->timing) check(hipEventRecord(in[i].e->ev[1], S), "hipEventRecord");

  Batch<MeLevelArgs> level0;  // the level-0 search of the wavefront form, launched by row groups
  // Motion search, top level first, every reference of every engine per launch.
  for (int l = kLevels - 1; l >= 0; l--) {
    Batch<MeLevelArgs> b;
    for (int i = 0; i < n; i++) {
      celv2hip_engine* e = in[i].e;
      const int nr = in[i].nrefs;
      if (!nr) continue;
      MeLevelArgs a{};
      a.cur = cur[i]->p.lv[l];
      a.nrefs = nr, a.cols = e->fcols[l], a.rows = e->frows[l];
      a.pcols = l + 1 < kLevels ? e->fcols[l + 1] : 0, a.prows = l + 1 < kLevels ? e->frows[l + 1] : 0;
      a.level0 = l == 0, a.phase_bytes = used[i][0]->p.phase_bytes;
      a.searches = e->counters + (l == 0 ? 5 : 4) * kCounterWords;  // level 0's centres counted apart
      a.kmap = use_kmap[i] ? e->kmap : nullptr, a.sb_cols = e->sb_cols, a.sb_rows = e->sb_rows, a.level = l;
      a.smap = use_smap[i] ? e->smap : nullptr;
      const bool feats = l == 2 && (e->pruner || e->want_features);
      for (int r = 0; r < nr; r++) {
        if (feats && !e->cost2[r]) check(hipMalloc(&e->cost2[r], static_cast<size_t>(e->fcols[2]) * e->frows[2] * sizeof(uint32_t)), "hipMalloc cost2");
        a.cost_out[r] = feats ? e->cost2[r] : nullptr;
        if (!e->fields[r][l]) check(hipMalloc(&e->fields[r][l], static_cast<size_t>(e->fcols[l]) * e->frows[l] * sizeof(uint32_t)), "hipMalloc field");
        a.ref[r] = used[i][r]->p.lv[l];
        a.coarser[r] = l + 1 < kLevels ? e->fields[r][l + 1] : nullptr;
        a.out[r] = e->fields[r][l];
      }
      add_me_level(b, a);
      if (b.n == kMaxItems) launch_me_level(b, S), b = Batch<MeLevelArgs>();
    }
    if (l == 0) {
      if (phases_pending) check(hipStreamWaitEvent(S, e0->sub_done, 0), "hipStreamWaitEvent");  // level 0 refines on them
      for (int i = 0; i < n; i++) if (in[i].e->timing) check(hipEventRecord(in[i].e->ev[6], S), "hipEventRecord");
      if (wavefront) {
        level0 = b;
        continue;
      }
    }
    launch_me_level(b, S);
    check(hipGetLastError(), "me kernel");
    if (l == 2) {
      // The pruner: features from the coarse levels (searched in full), the network, then the
      // per-superblock reference limits that levels 1 and 0, inter scores and finalize obey.
      for (int i = 0; i < n; i++) {
        celv2hip_engine* e = in[i].e;
        e->features_valid = false;
        if (!(e->pruner || e->want_features) || in[i].nrefs < 1) continue;
        if (!e->infer.cap) cnn_alloc(e->infer, e->nsb());
        FeatureArgs fa{};
        fa.src2 = cur[i]->p.lv[2];
        for (int r = 0; r < in[i].nrefs; r++) fa.cost2[r] = e->cost2[r];
        fa.nrefs = in[i].nrefs, fa.cols2 = e->fcols[2], fa.rows2 = e->frows[2], fa.sb_cols = e->sb_cols, fa.sb_rows = e->sb_rows;
        fa.lambda_q8 = in[i].lam;
        fa.feats = e->infer.x;
        launch_features(fa, S, /*context=*/!e->pruner);  // with a pruner the network's launch computes it
        e->features_valid = true;
        if (e->pruner) {
          const bool groups = intra_after[i] != 0;
          if (groups && !e->nmask) check(hipMalloc(&e->nmask, e->nsb()), "hipMalloc intra masks");
          const bool modes = e->prune_sthreshold > 0.f;
          cnn_infer(e->infer.x, 1, e->sb_rows, e->sb_cols, e->pruner->params, e->pruner->pb, nullptr, in[i].nrefs,
                    e->prune_threshold, e->prune_band, e->kmap, e->counters + 3 * kCounterWords, S,
                    groups ? e->nmask : nullptr, e->prune_gthreshold, modes ? e->smap : nullptr, e->prune_sthreshold,
                    /*context=*/true);
          use_kmap[i] = true;
          if (modes) use_smap[i] = true;
          if (groups) {
            // The source's intra scores with the network's groups, beside levels 1 and 0 (on the launch
            // stream in the graph form); finalize waits for them.
            const hipStream_t I = use_graph ? S : e->side;
            if (I != S) {
              check(hipEventRecord(e->infer_ev, S), "hipEventRecord");
              check(hipStreamWaitEvent(I, e->infer_ev, 0), "hipStreamWaitEvent");
            }
            IntraItem it = e->intra_item(*cur[i]);
            it.mask = e->nmask;
            Batch<IntraItem> b;
            add_intra(b, it);
            launch_intra(b, I);
            if (I != S) check(hipEventRecord(e->intra_ev, I), "hipEventRecord"), e->intra_ev_pending = true;
          }
        }
        check(hipGetLastError(), "pruner kernels");
      }
    }
  }
  for (int i = 0; i < n; i++) if (in[i].e->timing) check(hipEventRecord(in[i].e->ev[2], S), "hipEventRecord");

  auto inter_args = [&](int i) {
    celv2hip_engine* e = in[i].e;
    InterArgs ia{};
    ia.src = cur[i]->p.lv[0];
    ia.nrefs = in[i].nrefs, ia.cols = e->fcols[0], ia.rows = e->frows[0], ia.sb_cols = e->sb_cols;
    ia.phase_bytes = in[i].nrefs ? used[i][0]->p.phase_bytes : 0;
    ia.best = e->best;
    ia.counter = e->counters, ia.candidates = e->counting ? e->counters + 2 * kCounterWords : nullptr;
    ia.kmap = use_kmap[i] ? e->kmap : nullptr;
    for (int r = 0; r < in[i].nrefs; r++) ia.ref[r] = used[i][r]->p.lv[0], ia.field[r] = e->fields[r][0];
    return ia;
  };
  auto finalize_args = [&](int i) {
    celv2hip_engine* e = in[i].e;
    FinalizeArgs fa{};
    fa.src = cur[i]->p.lv[0];
    fa.nrefs = in[i].nrefs, fa.sb_cols = e->sb_cols, fa.lam = in[i].lam;
    fa.phase_bytes = in[i].nrefs ? used[i][0]->p.phase_bytes : 0;
    fa.best = e->best, fa.intra = cur[i]->intra, fa.out = e->rec_dev, fa.counter = e->counters + kCounterWords;
    fa.kmap = use_kmap[i] ? e->kmap : nullptr, fa.sb_cost = e->sb_cost;
    fa.imask = cur[i]->imask_used;
    for (int r = 0; r < in[i].nrefs; r++) fa.ref[r] = used[i][r]->p.lv[0];
    return fa;
  };
  // Rows: row streams as VulkanEngine orders them (src/vk/vk_engine.cpp:254-260).
  for (int i = 0; i < n; i++) {
    celv2hip_engine* e = in[i].e;
    const int streams = std::max(1, std::min(e->row_streams, e->sb_rows));
    const int run = (e->sb_rows + streams - 1) / streams;
    e->order.clear();
    for (int k = 0; k < run; k++)
      for (int t = 0; t < streams; t++)
        if (t * run + k < e->sb_rows) e->order.push_back(t * run + k);
  }
  // Records of rows r0 .. r0 + nrows - 1 of engine e to the host once the launch stream has finished
  // them (group g's event), on the engine's aux stream, group_done[g] behind the copy.
  auto queue_copy = [&](celv2hip_engine* e, int g, int r0, int nrows, hipStream_t done_on) {
    const size_t row_bytes = static_cast<size_t>(e->sb_cols) * sizeof(celv2_sb);
    check(hipEventRecord(e->rows_ev[g], done_on), "hipEventRecord");
    check(hipStreamWaitEvent(e->aux, e->rows_ev[g], 0), "hipStreamWaitEvent");
    const size_t off = static_cast<size_t>(r0) * row_bytes;
    // The copy engine, not a copy kernel: a kernel writing to host memory beside the inter scores
    // slowed them by 60-100 % (measured), the copy engine's transfers barely.
    check(hipMemcpyAsync(static_cast<char*>(e->rec_host) + off, static_cast<char*>(e->rec_dev) + off, row_bytes * nrows,
                         hipMemcpyDeviceToHost, e->aux), "hipMemcpyAsync records");
    check(hipEventRecord(e->group_done[g], e->aux), "hipEventRecord");
    for (int r = r0; r < r0 + nrows; r++) e->row_group[r] = g;
  };
  // The previous launch's downloads read rec_dev: this launch's finalize writes it after them (the
  // graph form waited before its capture).
  for (int i = 0; i < n; i++)
    if (in[i].e->copy_pending && !use_graph) check(hipStreamWaitEvent(S, in[i].e->copy_done, 0), "hipStreamWaitEvent");
  const bool grouped = n == 1 && !e0->resident;
  if (!grouped) {
    // All rows in one launch per kernel (batches, and resident mode).
    Batch<InterArgs> bi;
    for (int i = 0; i < n; i++) {
      if (!in[i].nrefs) continue;
      InterArgs ia = inter_args(i);
      ia.row0 = 0, ia.nrows = in[i].e->sb_rows;
      add_inter(bi, ia);
      if (bi.n == kMaxItems) launch_inter(bi, S), bi = Batch<InterArgs>();
    }
    launch_inter(bi, S);
    for (int i = 0; i < n; i++) if (in[i].e->timing) check(hipEventRecord(in[i].e->ev[3], S), "hipEventRecord");
    for (int i = 0; i < n; i++) if (late_claim[i]) in[i].e->claim(*cur[i], S);
    for (int i = 0; i < n; i++)
      if (in[i].e->intra_ev_pending) check(hipStreamWaitEvent(S, in[i].e->intra_ev, 0), "hipStreamWaitEvent"), in[i].e->intra_ev_pending = false;
    Batch<FinalizeArgs> bf;
    for (int i = 0; i < n; i++) {
      FinalizeArgs fa = finalize_args(i);
      fa.row0 = 0, fa.nrows = in[i].e->sb_rows;
      add_finalize(bf, fa);
      if (bf.n == kMaxItems) launch_finalize(bf, S), bf = Batch<FinalizeArgs>();
    }
    launch_finalize(bf, S);
    check(hipGetLastError(), "row kernels");
    for (int i = 0; i < n; i++) {
      celv2hip_engine* e = in[i].e;
      if (e->timing) check(hipEventRecord(e->ev[4], S), "hipEventRecord");
      if (!e->resident) queue_copy(e, 0, 0, e->sb_rows, S);
    }
  } else {
    // One engine: its rows in groups, each group's inter scores and finalize in turn, each group's
    // download on the aux stream while the next group computes. One row stream (the wavefront form):
    // groups of about a quarter of the rows, level 0 of the search by the same groups one ahead, so
    // each download overlaps the next group's search; several row streams: one row at a time in
    // row-stream order after the whole search.
    celv2hip_engine* e = e0;
    std::vector<std::pair<int, int>> groups;  // (first row, rows)
    if (!wavefront) {
      for (int r : e->order) groups.push_back({r, 1});
    } else {
      const int rows = std::max(1, (e->sb_rows + e->row_groups - 1) / e->row_groups);
      for (int r0 = 0; r0 < e->sb_rows; r0 += rows) groups.push_back({r0, std::min(rows, e->sb_rows - r0)});
    }
    // The level-0 search of group g's block rows. A superblock row's inter scores read the level-0
    // field one cell beyond its own 8 block rows, so group g's rows are scored once group g + 1's
    // search is done: search(0), search(1), rows(0), search(2), rows(1), ..., rows(last).
    auto search_rows = [&](size_t g) {
      if (!wavefront || g >= groups.size() || level0.n == 0) return;
      Batch<MeLevelArgs> b0;
      MeLevelArgs a = level0.item[0];
      a.row0 = 8 * groups[g].first;
      a.nrun = std::min(8 * groups[g].second, e->frows[0] - a.row0);
      if (a.nrun <= 0) return;
      add_me_level(b0, a);
      launch_me_level(b0, S);
      check(hipGetLastError(), "me kernel");
    };
    search_rows(0);
    for (size_t g = 0; g < groups.size(); g++) {
      search_rows(g + 1);
      if (wavefront && g + 1 == groups.size() && e->timing) check(hipEventRecord(e->ev[2], S), "hipEventRecord");
      const auto [r0, nr] = groups[g];
      if (in[0].nrefs) {
        Batch<InterArgs> bi;
        InterArgs ia = inter_args(0);
        ia.row0 = r0, ia.nrows = nr;
        add_inter(bi, ia);
        launch_inter(bi, S);
      }
      // Wavefront form: the group's finalize on the side stream once its inter scores are done.
      const hipStream_t F = wavefront ? e->side : S;
      if (!wavefront && e->intra_ev_pending) check(hipStreamWaitEvent(S, e->intra_ev, 0), "hipStreamWaitEvent"), e->intra_ev_pending = false;
      if (wavefront) {
        check(hipEventRecord(e->inter_ev[g], S), "hipEventRecord");
        check(hipStreamWaitEvent(F, e->inter_ev[g], 0), "hipStreamWaitEvent");
      }
      Batch<FinalizeArgs> bf;
      FinalizeArgs fa = finalize_args(0);
      fa.row0 = r0, fa.nrows = nr;
      add_finalize(bf, fa);
      launch_finalize(bf, F);
      check(hipGetLastError(), "row kernels");
      queue_copy(e, static_cast<int>(g), r0, nr, F);
    }
    if (e->timing) check(hipEventRecord(e->ev[3], S), "hipEventRecord");
    if (e->timing) check(hipEventRecord(e->ev[4], S), "hipEventRecord");
  }
  if (use_graph) {
    hipGraph_t g = nullptr;
    capture.s = nullptr;
    check(hipStreamEndCapture(S, &g), "hipStreamEndCapture");
    hipError_t rc = hipErrorUnknown;
    if (e0->gexec) {
      hipGraphNode_t bad = nullptr;
      hipGraphExecUpdateResult why;
      rc = hipGraphExecUpdate(e0->gexec, g, &bad, &why);
      if (rc != hipSuccess || why != hipGraphExecUpdateSuccess) {
        (void)hipGetLastError();
        (void)hipGraphExecDestroy(e0->gexec);
        e0->gexec = nullptr;
        rc = hipErrorUnknown;
      }
    }
    if (rc != hipSuccess) {
      const hipError_t irc = hipGraphInstantiate(&e0->gexec, g, nullptr, nullptr, 0);
      if (irc != hipSuccess) (void)hipGraphDestroy(g), e0->gexec = nullptr;
      check(irc, "hipGraphInstantiate");
      e0->graph_builds++;
    }
    (void)hipGraphDestroy(g);
    check(hipGraphLaunch(e0->gexec, S), "hipGraphLaunch");
    e0->graph_launches++;
    for (int i = 0; i < n; i++) {
      check(hipEventRecord(in[i].e->deep_ev, S), "hipEventRecord");
      in[i].e->deep_pending = true;
    }
  }
  for (int i = 0; i < n; i++) {
    celv2hip_engine* e = in[i].e;
    if (e->resident) continue;
    check(hipEventRecord(e->copy_done, e->aux), "hipEventRecord");
    e->copy_pending = true;
  }
  for (int i = 0; i < n; i++) {
    celv2hip_engine* e = in[i].e;
    if (e->timing) check(hipEventRecord(e->ev[5], e->resident ? S : e->aux), "hipEventRecord");
    e->release(*cur[i], S);
    e->nrefs_last = in[i].nrefs;
    e->limits_last = use_kmap[i] != 0;
    e->modes_last = use_smap[i] != 0;
    e->last_refs = used[i];
    e->launched = true;
  }
  // The other engines' later work waits for the batch when it comes (batch_done).
  e0->batch_on = nullptr;
  for (int i = 1; i < n; i++) {
    check(hipEventRecord(in[i].e->batch_done, S), "hipEventRecord");
    in[i].e->batch_on = S;
  }
}

}  // namespace

int celv2hip_launch(celv2hip_engine* e, const celv2_luma* src, long long src_key, const celv2_luma* refs,
                    const long long* ref_keys, int nrefs, int64_t lambda_q8) {
  if (int rc = validate(e)) return rc;
  return guarded([&] {
    const LaunchIn x{e, src, src_key, refs, ref_keys, nrefs, lambda_q8};
    launch_batch(&x, 1);
  });
}

int celv2hip_launch_batch(celv2hip_engine* const* engines, int n, const celv2_luma* srcs, const long long* src_keys,
                          const celv2_luma* const* refs, const long long* const* ref_keys, const int* nrefs,
                          const int64_t* lambdas_q8) {
  if (!engines || n <= 0 || n > 64 || !srcs || !nrefs || !lambdas_q8) {
    g_error = "launch_batch: bad arguments";
    return CELV2_ERR_ARG;
  }
  for (int i = 0; i < n; i++)
    if (int rc = validate(engines[i])) return rc;
  return guarded([&] {
    std::vector<LaunchIn> in(n);
    for (int i = 0; i < n; i++)
      in[i] = LaunchIn{engines[i], &srcs[i], src_keys ? src_keys[i] : -1, refs ? refs[i] : nullptr,
                       ref_keys ? ref_keys[i] : nullptr, nrefs[i], lambdas_q8[i]};
    launch_batch(in.data(), n);
  });
}

int celv2hip_row_order(const celv2hip_engine* e, int* rows, int cap) {
  if (!e) return -CELV2_ERR_ARG;
  const int n = static_cast<int>(e->order.size());
  for (int i = 0; i < n && i < cap && rows; i++) rows[i] = e->order[i];
  return n;
}

int celv2hip_wait_row(celv2hip_engine* e, int row, celv2_sb* out) {
  if (int rc = validate_host(e)) return rc;
  return guarded([&] {
    if (!e->launched) throw Failure(CELV2_ERR_STATE, "HipEngine: no launch to collect");
    if (e->resident) throw Failure(CELV2_ERR_STATE, "HipEngine: resident mode keeps results on the device");
    if (row < 0 || row >= e->sb_rows) throw Failure(CELV2_ERR_ARG, "HipEngine: no such row");
    check(hipEventSynchronize(e->group_done[e->row_group[row]]), "row wait");
    if (out) std::memcpy(out, static_cast<char*>(e->rec_host) + static_cast<size_t>(row) * e->sb_cols * sizeof(celv2_sb), e->sb_cols * sizeof(celv2_sb));
  });
}

int celv2hip_row_records(celv2hip_engine* e, int row, const celv2_sb** out) {
  if (int rc = validate_host(e)) return rc;
  return guarded([&] {
    if (!out) throw Failure(CELV2_ERR_ARG, "HipEngine: no output pointer");
    if (!e->launched) throw Failure(CELV2_ERR_STATE, "HipEngine: no launch to collect");
    if (e->resident) throw Failure(CELV2_ERR_STATE, "HipEngine: resident mode keeps results on the device");
    if (row < 0 || row >= e->sb_rows) throw Failure(CELV2_ERR_ARG, "HipEngine: no such row");
    check(hipEventSynchronize(e->group_done[e->row_group[row]]), "row wait");
    *out = reinterpret_cast<const celv2_sb*>(static_cast<const char*>(e->rec_host) + static_cast<size_t>(row) * e->sb_cols * sizeof(celv2_sb));
  });
}

int celv2hip_fetch_row(celv2hip_engine* e, int row, celv2_sb* out) {
  if (int rc = validate(e)) return rc;
  return guarded([&] {
    if (row < 0 || row >= e->sb_rows) throw Failure(CELV2_ERR_ARG, "HipEngine: no such row");
    e->sync_all();
    check(hipMemcpy(out, static_cast<char*>(e->rec_dev) + static_cast<size_t>(row) * e->sb_cols * sizeof(celv2_sb),
                    e->sb_cols * sizeof(celv2_sb), hipMemcpyDeviceToHost), "hipMemcpy row");
  });
}

static int fetch_plane(celv2hip_engine* e, int ref, int level, int phase, bool bordered, uint8_t* out, int* w, int* h) {
  if (int rc = validate(e)) return rc;
  return guarded([&] {
    if (!e->launched) throw Failure(CELV2_ERR_STATE, "no launch");
    if (level < 0 || level >= kLevels || phase < 0 || phase >= kPhases || (phase && level)) throw Failure(CELV2_ERR_ARG, "no such plane");
    const Pyramid* p = nullptr;
    if (ref < 0) p = &e->srcs[e->cur_set].p;
    else if (ref < static_cast<int>(e->last_refs.size())) p = &e->last_refs[ref]->p;
    else throw Failure(CELV2_ERR_ARG, "no such reference");
    if (phase && !p->phase_bytes) throw Failure(CELV2_ERR_ARG, "no phases");
    const DPlane& d = p->lv[level];
    const int b = bordered ? kBorder : 0;
    *w = d.w + 2 * b, *h = d.h + 2 * b;
    e->sync_all();
    const uint8_t* first = d.p + static_cast<ptrdiff_t>(phase) * p->phase_bytes - static_cast<ptrdiff_t>(b) * d.stride - b;
    check(hipMemcpy2D(out, *w, first, d.stride, *w, *h, hipMemcpyDeviceToHost), "hipMemcpy2D");
  });
}
int celv2hip_fetch_plane(celv2hip_engine* e, int ref, int level, int phase, uint8_t* out, int* w, int* h) {
  return fetch_plane(e, ref, level, phase, false, out, w, h);
}
int celv2hip_fetch_plane_bordered(celv2hip_engine* e, int ref, int level, int phase, uint8_t* out, int* w, int* h) {
  return fetch_plane(e, ref, level, phase, true, out, w, h);
}

int celv2hip_fetch_field(celv2hip_engine* e, int ref, int level, int32_t* mvx, int32_t* mvy, int* cols, int* rows) {
  if (int rc = validate(e)) return rc;
  return guarded([&] {
    if (!e->launched || ref < 0 || ref >= e->nrefs_last || level < 0 || level >= kLevels) throw Failure(CELV2_ERR_ARG, "no such field");
    const int n = e->fcols[level] * e->frows[level];
    std::vector<uint32_t> packed(n);
    e->sync_all();
    check(hipMemcpy(packed.data(), e->fields[ref][level], n * sizeof(uint32_t), hipMemcpyDeviceToHost), "hipMemcpy field");
    for (int i = 0; i < n; i++) mvx[i] = static_cast<int16_t>(packed[i] & 0xffffu), mvy[i] = static_cast<int16_t>(packed[i] >> 16);
    *cols = e->fcols[level], *rows = e->frows[level];
  });
}

long long celv2hip_fetch_intra(celv2hip_engine* e, uint32_t* out, long long cap) {
  if (int rc = validate(e)) return rc;
  long long n = 0;
  const int rc = guarded([&] {
    const SrcSlot& s = e->srcs[e->cur_set];
    if (!e->launched || !s.intra) throw Failure(CELV2_ERR_STATE, "no launch");
    n = static_cast<long long>(e->nsb()) * kBlocks * kModes;
    if (!out || cap < n) throw Failure(CELV2_ERR_SIZE, "intra buffer too small");
    e->sync_all();
    check(hipMemcpy(out, s.intra, static_cast<size_t>(n) * sizeof(uint32_t), hipMemcpyDeviceToHost), "hipMemcpy intra");
  });
  return rc ? rc : n;
}

int celv2hip_sync(celv2hip_engine* e) {
  if (int rc = validate(e)) return rc;
  return guarded([&] { e->sync_all(); });
}

int celv2hip_stats_reset(celv2hip_engine* e) {
  if (int rc = validate(e)) return rc;
  return guarded([&] {
    e->sync_all();
    check(hipMemset(e->counters, 0, kCounterBytes), "hipMemset counters");
    e->window_launch = e->launches, e->window_prepare = e->prepares, e->window_untimed = 0;
  });
}

int celv2hip_stats(celv2hip_engine* e, double* ms, int cap_ms, long long* counts, int cap_counts) {
  if (int rc = validate(e)) return -rc;
  long long n = 0;
  const int rc = guarded([&] {
    e->sync_all();
    n = e->launches - e->window_launch;
    const long long np = e->prepares - e->window_prepare;
    if (n > celv2hip_engine::kRing || np > celv2hip_engine::kRing) throw Failure(CELV2_ERR_STATE, "stats window longer than the event ring");
    double acc[8] = {};
    for (long long i = e->window_launch; i < e->launches && !e->window_untimed; i++) {
      auto& a = e->ring[i % celv2hip_engine::kRing];
      const int pairs[7][2] = {{0, 1}, {1, 2}, {2, 3}, {3, 4}, {4, 5}, {0, 5}, {6, 2}};
      for (int k = 0; k < 7; k++) {
        float t = 0;
        check(hipEventElapsedTime(&t, a[pairs[k][0]], a[pairs[k][1]]), "hipEventElapsedTime");
        acc[k < 6 ? k : 7] += t;
      }
    }
    for (long long i = e->window_prepare; i < e->prepares && !e->window_untimed; i++) {
      auto& a = e->pring[i % celv2hip_engine::kRing];
      float t = 0;
      check(hipEventElapsedTime(&t, a[0], a[1]), "hipEventElapsedTime");
      acc[6] += t;
    }
    for (int k = 0; k < 8 && k < cap_ms; k++) ms[k] = acc[k];
    check(hipMemcpy(e->counters_host, e->counters, kCounterBytes, hipMemcpyDeviceToHost), "hipMemcpy counters");
    auto total = [&](int k) {
      long long t = 0;
      for (int i = 0; i < kCounterSlots; i++) t += static_cast<long long>(e->counters_host[k * kCounterWords + kCounterStride * i]);
      return t;
    };
    const long long v[7] = {n, total(0), total(4) + total(5), total(1), total(2), total(3), total(5)};
    for (int k = 0; k < 7 && k < cap_counts; k++) counts[k] = v[k];
  });
  return rc ? -rc : static_cast<int>(n);
}

int celv2hip_set_timing(celv2hip_engine* e, int on) {
  if (int rc = validate(e)) return rc;
  return guarded([&] { e->timing = on != 0; });
}

int celv2hip_device_alloc(int device, long long bytes, void** out) {
  return guarded([&] {
    check(hipSetDevice(device), "hipSetDevice");
    check(hipMalloc(out, static_cast<size_t>(bytes)), "hipMalloc");
  });
}
int celv2hip_device_free(int device, void* p) {
  return guarded([&] {
    check(hipSetDevice(device), "hipSetDevice");
    check(hipFree(p), "hipFree");
  });
}
int celv2hip_copy_to_device(int device, void* dst, const void* src, long long bytes) {
  return guarded([&] {
    check(hipSetDevice(device), "hipSetDevice");
    check(hipMemcpy(dst, src, static_cast<size_t>(bytes), hipMemcpyHostToDevice), "hipMemcpy");
  });
}

int celv2hip_copy_from_device(int device, void* dst, const void* src, long long bytes) {
  return guarded([&] {
    check(hipSetDevice(device), "hipSetDevice");
    check(hipMemcpy(dst, src, static_cast<size_t>(bytes), hipMemcpyDeviceToHost), "hipMemcpy D2H");
  });
}

int celv2hip_copy_on_device(int device, void* dst, const void* src, long long bytes) {
  return guarded([&] {
    check(hipSetDevice(device), "hipSetDevice");
    check(hipMemcpy(dst, src, static_cast<size_t>(bytes), hipMemcpyDeviceToDevice), "hipMemcpy D2D");
    check(hipDeviceSynchronize(), "hipDeviceSynchronize");
  });
}

}  // extern "C"
