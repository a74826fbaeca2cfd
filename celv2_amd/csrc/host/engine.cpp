// The C ABI of libcelv2hip.so (include/celv2hip.h): device memory, reference-key caches, the
// per-launch queue of kernels on the engine's main HIP stream, prepare() on a side stream, and row
// completion events.
//
// Queue order of a launch (VulkanEngine::launch, src/vk/vk_engine.cpp:207-312, re-planned for HIP):
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
#include <stdexcept>
#include <string>
#include <vector>

#include "../../../include/celv2hip.h"
#include "../kernels/launch.h"

namespace celv2hip {
void upload_intra_tiles(hipStream_t s);
void upload_inter_tables(hipStream_t s);
void upload_final_tiles(hipStream_t s);
}  // namespace celv2hip

using namespace celv2hip;

namespace {

thread_local std::string g_error;

struct Failure : std::runtime_error {
  int code;
  Failure(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

void check(hipError_t e, const char* what) {
  if (e != hipSuccess) throw Failure(e == hipErrorOutOfMemory ? CELV2_ERR_MEMORY : CELV2_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

template <typename F>
int guarded(F&& f) {
  try {
    f();
    return CELV2_OK;
  } catch (const Failure& e) {
    g_error = e.what();
    return e.code;
  } catch (const std::exception& e) {
    g_error = e.what();
    return CELV2_ERR_STATE;
  }
}

size_t round_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

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
    if (l == 0 && phases) p.phase_bytes = static_cast<long long>(b);
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
  // prepare() fills a slot on the side stream while launches run on the main stream: each stream
  // waits for the other's last work on the slot before touching it.
  hipEvent_t side_done = nullptr, main_done = nullptr;
  bool side_pending = false, main_pending = false;
};

}  // namespace

struct celv2hip_engine {
  int width, height, sb_size, device, pw, ph, sb_cols, sb_rows;
  hipStream_t stream = nullptr;  // launches
  hipStream_t side = nullptr;    // prepare(): the next source's planes and intra scores
  std::string name;
  int row_streams = 1;
  bool resident = false;
  bool counting = false;  // count non-repeated inter candidates (costs the inter kernel LDS atomics)
  SrcSlot srcs[2];
  int cur_set = 0;
  std::vector<std::unique_ptr<RefSlot>> slots;
  long long launches = 0;
  uint32_t* fields[kMaxRefs][kLevels] = {};
  int fcols[kLevels], frows[kLevels];
  uint2* best = nullptr;
  void* rec_dev = nullptr;   // celv2_sb per superblock
  void* rec_host = nullptr;  // pinned
  std::vector<int> order;
  std::vector<hipEvent_t> row_done;
  // Per-launch stage events in a ring (kRing launches): start, references built, motion search,
  // inter scores, finalize, end; and per prepare(): intra scores start / end.
  static constexpr int kRing = 512;
  std::vector<std::array<hipEvent_t, 6>> ring;
  std::vector<std::array<hipEvent_t, 2>> pring;
  hipEvent_t* ev = nullptr;  // the current launch's 6 events
  long long prepares = 0, window_launch = 0, window_prepare = 0;
  unsigned long long* counters = nullptr;  // device: [0] unique MVs scored, [1] compound blocks, [2] inter candidates, [16..]: ME searches
  unsigned long long* counters_host = nullptr;
  int nrefs_last = 0;
  bool launched = false;
  std::vector<RefSlot*> last_refs;  // the last launch's reference slots, nearest first

  ~celv2hip_engine() {
    if (device >= 0) (void)hipSetDevice(device);
    if (stream) (void)hipStreamSynchronize(stream);
    if (side) (void)hipStreamSynchronize(side);
    auto free_up = [](Upload& u) {
      if (u.host) (void)hipHostFree(u.host);
      if (u.dev) (void)hipFree(u.dev);
      if (u.done) (void)hipEventDestroy(u.done);
    };
    for (SrcSlot& s : srcs) {
      if (s.p.mem) (void)hipFree(s.p.mem);
      if (s.intra) (void)hipFree(s.intra);
      if (s.side_done) (void)hipEventDestroy(s.side_done);
      if (s.main_done) (void)hipEventDestroy(s.main_done);
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
    if (rec_dev) (void)hipFree(rec_dev);
    if (rec_host) (void)hipHostFree(rec_host);
    if (counters) (void)hipFree(counters);
    if (counters_host) (void)hipHostFree(counters_host);
    for (hipEvent_t e : row_done) (void)hipEventDestroy(e);
    for (auto& a : ring)
      for (hipEvent_t e : a) (void)hipEventDestroy(e);
    for (auto& a : pring)
      for (hipEvent_t e : a) (void)hipEventDestroy(e);
    if (stream) (void)hipStreamDestroy(stream);
    if (side) (void)hipStreamDestroy(side);
  }

  void sync_all() {
    check(hipStreamSynchronize(stream), "hipStreamSynchronize");
    check(hipStreamSynchronize(side), "hipStreamSynchronize side");
  }

  size_t nsb() const { return static_cast<size_t>(sb_cols) * sb_rows; }

  void ensure_pyramid(Pyramid& p, bool phases) {
    if (p.mem) return;
    p.bytes = pyramid_bytes(pw, ph, phases);
    check(hipMalloc(&p.mem, p.bytes), "hipMalloc pyramid");
    layout(p, pw, ph, phases);
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

  void build(const celv2_luma& f, Pyramid& p, Upload& up, bool phases, hipStream_t st) {
    ensure_pyramid(p, phases);
    const void* src = stage(f, up, st);
    if (f.y16) launch_import16(static_cast<const uint16_t*>(src), pw, f.bit_depth, p.lv[0], st);
    else launch_import8(static_cast<const uint8_t*>(src), pw, p.lv[0], st);
    if (phases) launch_subpel(p.lv[0], p.phase_bytes, st);
    for (int l = 1; l < kLevels; l++) launch_downsample(p.lv[l - 1], p.lv[l], st);
    check(hipGetLastError(), "plane kernels");
  }

  void intra_scores(SrcSlot& s, hipStream_t st) {
    if (!s.intra) check(hipMalloc(&s.intra, nsb() * kBlocks * kModes * sizeof(uint32_t)), "hipMalloc intra");
    std::array<hipEvent_t, 2>& pe = pring[prepares++ % kRing];
    check(hipEventRecord(pe[0], st), "hipEventRecord");
    launch_intra(s.p.lv[0], sb_cols, 0, sb_rows, s.intra, st);
    check(hipGetLastError(), "intra kernel");
    check(hipEventRecord(pe[1], st), "hipEventRecord");
  }

  // Orders stream `st`'s coming work on slot s after the other stream's last work on it.
  void claim(SrcSlot& s, bool on_side) {
    if (on_side && s.main_pending) check(hipStreamWaitEvent(side, s.main_done, 0), "hipStreamWaitEvent");
    if (!on_side && s.side_pending) check(hipStreamWaitEvent(stream, s.side_done, 0), "hipStreamWaitEvent");
  }
  void release(SrcSlot& s, bool on_side) {
    if (on_side) check(hipEventRecord(s.side_done, side), "hipEventRecord"), s.side_pending = true;
    else check(hipEventRecord(s.main_done, stream), "hipEventRecord"), s.main_pending = true;
  }
};

namespace {

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
    for (SrcSlot& s : e->srcs) {
      check(hipEventCreateWithFlags(&s.side_done, hipEventDisableTiming), "hipEventCreate");
      check(hipEventCreateWithFlags(&s.main_done, hipEventDisableTiming), "hipEventCreate");
    }
    int w = e->pw, h = e->ph;
    for (int l = 0; l < kLevels; l++) {
      e->fcols[l] = (w + 7) / 8, e->frows[l] = (h + 7) / 8;
      w /= 2, h /= 2;
    }
    check(hipMalloc(&e->best, e->nsb() * kMaxRefs * kBlocks * sizeof(uint2)), "hipMalloc best");
    check(hipMalloc(&e->rec_dev, e->nsb() * sizeof(celv2_sb)), "hipMalloc records");
    check(hipHostMalloc(&e->rec_host, e->nsb() * sizeof(celv2_sb), hipHostMallocDefault), "hipHostMalloc records");
    check(hipMalloc(&e->counters, 80 * 16 * sizeof(unsigned long long)), "hipMalloc counters");
    check(hipHostMalloc(reinterpret_cast<void**>(&e->counters_host), 80 * 16 * sizeof(unsigned long long), hipHostMallocDefault), "hipHostMalloc counters");
    e->row_done.resize(e->sb_rows);
    for (hipEvent_t& ev : e->row_done) check(hipEventCreateWithFlags(&ev, hipEventDisableTiming), "hipEventCreate");
    e->ring.resize(celv2hip_engine::kRing);
    e->pring.resize(celv2hip_engine::kRing);
    for (auto& a : e->ring)
      for (hipEvent_t& ev : a) check(hipEventCreate(&ev), "hipEventCreate");
    for (auto& a : e->pring)
      for (hipEvent_t& ev : a) check(hipEventCreate(&ev), "hipEventCreate");
    e->ev = e->ring[0].data();
    check(hipMemset(e->counters, 0, 80 * 16 * sizeof(unsigned long long)), "hipMemset counters");
    upload_intra_tiles(e->stream);
    upload_inter_tables(e->stream);
    upload_final_tiles(e->stream);
    check(hipStreamSynchronize(e->stream), "table upload");
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

int celv2hip_set_resident(celv2hip_engine* e, int on) {
  if (!e) return CELV2_ERR_ARG;
  e->resident = on != 0;
  return CELV2_OK;
}

int celv2hip_set_counting(celv2hip_engine* e, int on) {
  if (!e) return CELV2_ERR_ARG;
  e->counting = on != 0;
  return CELV2_OK;
}

int celv2hip_prepare(celv2hip_engine* e, const celv2_luma* next, long long key) {
  if (int rc = validate(e)) return rc;
  return guarded([&] {
    if (key < 0 || !next) return;
    // The slot the last launch does not use, filled on the side stream once the main stream's
    // earlier launches are done with it; the launch that takes it waits for the side stream.
    SrcSlot& s = e->srcs[1 - e->cur_set];
    s.key = -1;
    e->claim(s, true);
    e->build(*next, s.p, s.up, false, e->side);
    e->intra_scores(s, e->side);
    e->release(s, true);
    s.key = key;
  });
}

int celv2hip_launch(celv2hip_engine* e, const celv2_luma* src, long long src_key, const celv2_luma* refs,
                    const long long* ref_keys, int nrefs, int64_t lambda_q8) {
  if (int rc = validate(e)) return rc;
  return guarded([&] {
    if (!src) throw Failure(CELV2_ERR_ARG, "HipEngine: no source");
    if (nrefs < 0 || nrefs > kMaxRefs) throw Failure(CELV2_ERR_SIZE, "HipEngine: too many references");
    if (nrefs > 0 && !refs) throw Failure(CELV2_ERR_ARG, "HipEngine: no references");
    e->ev = e->ring[e->launches % celv2hip_engine::kRing].data();
    check(hipEventRecord(e->ev[0], e->stream), "hipEventRecord");
    int use = -1;
    for (int i = 0; i < 2; i++)
      if (src_key >= 0 && e->srcs[i].key == src_key) use = i;
    e->claim(e->srcs[use < 0 ? 1 - e->cur_set : use], false);
    if (use < 0) {
      use = 1 - e->cur_set;
      SrcSlot& s = e->srcs[use];
      s.key = -1;
      e->build(*src, s.p, s.up, false, e->stream);
      e->intra_scores(s, e->stream);
    }
    e->srcs[use].key = -1;  // taken: a later prepare() may overwrite it
    e->cur_set = use;
    SrcSlot& cs = e->srcs[use];
    const long long now = ++e->launches;
    std::vector<RefSlot*> used(nrefs, nullptr);
    for (int ri = 0; ri < nrefs; ri++) {
      const long long key = ref_keys ? ref_keys[ri] : -1;
      RefSlot* slot = nullptr;
      for (auto& s : e->slots)
        if (key >= 0 && s->key == key) slot = s.get();
      if (!slot) {
        for (auto& s : e->slots) {
          const bool wanted = ref_keys && std::find(ref_keys, ref_keys + nrefs, s->key) != ref_keys + nrefs && s->key >= 0;
          if (s->used != now && !wanted) {
            slot = s.get();
            break;
          }
        }
        if (!slot) e->slots.push_back(std::make_unique<RefSlot>()), slot = e->slots.back().get();
        slot->key = -1;
        e->build(refs[ri], slot->p, slot->up, true, e->stream);
        slot->key = key;
      }
      slot->used = now;
      used[ri] = slot;
    }
    check(hipEventRecord(e->ev[1], e->stream), "hipEventRecord");
    // Motion search, top level first, every reference per launch.
    for (int l = kLevels - 1; l >= 0 && nrefs > 0; l--) {
      MeLevelArgs a{};
      a.cur = cs.p.lv[l];
      a.nrefs = nrefs, a.cols = e->fcols[l], a.rows = e->frows[l];
      a.pcols = l + 1 < kLevels ? e->fcols[l + 1] : 0, a.prows = l + 1 < kLevels ? e->frows[l + 1] : 0;
      a.level0 = l == 0, a.phase_bytes = used[0]->p.phase_bytes;
      a.searches = e->counters + 16;
      for (int r = 0; r < nrefs; r++) {
        if (!e->fields[r][l]) check(hipMalloc(&e->fields[r][l], static_cast<size_t>(e->fcols[l]) * e->frows[l] * sizeof(uint32_t)), "hipMalloc field");
        a.ref[r] = used[r]->p.lv[l];
        a.coarser[r] = l + 1 < kLevels ? e->fields[r][l + 1] : nullptr;
        a.out[r] = e->fields[r][l];
      }
      launch_me_level(a, e->stream);
      check(hipGetLastError(), "me kernel");
    }
    check(hipEventRecord(e->ev[2], e->stream), "hipEventRecord");
    // Rows: row streams as VulkanEngine orders them (src/vk/vk_engine.cpp:288-294).
    const int streams = std::max(1, std::min(e->row_streams, e->sb_rows));
    const int run = (e->sb_rows + streams - 1) / streams;
    e->order.clear();
    for (int i = 0; i < run; i++)
      for (int s = 0; s < streams; s++)
        if (s * run + i < e->sb_rows) e->order.push_back(s * run + i);
    InterArgs ia{};
    ia.src = cs.p.lv[0];
    ia.nrefs = nrefs, ia.cols = e->fcols[0], ia.rows = e->frows[0], ia.sb_cols = e->sb_cols;
    ia.phase_bytes = nrefs ? used[0]->p.phase_bytes : 0;
    ia.best = e->best;
    ia.counter = e->counters, ia.candidates = e->counting ? e->counters + 2 : nullptr;
    FinalizeArgs fa{};
    fa.src = cs.p.lv[0];
    fa.nrefs = nrefs, fa.sb_cols = e->sb_cols, fa.phase_bytes = ia.phase_bytes, fa.lam = lambda_q8;
    fa.best = e->best, fa.intra = cs.intra, fa.out = e->rec_dev, fa.counter = e->counters + 1;
    for (int r = 0; r < nrefs; r++) ia.ref[r] = fa.ref[r] = used[r]->p.lv[0], ia.field[r] = e->fields[r][0];
    const size_t row_bytes = static_cast<size_t>(e->sb_cols) * sizeof(celv2_sb);
    if (e->resident || streams == 1) {
      // All rows in one launch per kernel (resident: results stay on the device).
      ia.row0 = 0, ia.nrows = e->sb_rows;
      if (nrefs) launch_inter(ia, e->stream);
      check(hipEventRecord(e->ev[3], e->stream), "hipEventRecord");
      fa.row0 = 0, fa.nrows = e->sb_rows;
      launch_finalize(fa, e->stream);
      check(hipGetLastError(), "row kernels");
      check(hipEventRecord(e->ev[4], e->stream), "hipEventRecord");
      if (!e->resident) {
        check(hipMemcpyAsync(e->rec_host, e->rec_dev, row_bytes * e->sb_rows, hipMemcpyDeviceToHost, e->stream), "hipMemcpyAsync records");
        for (int r = 0; r < e->sb_rows; r++) check(hipEventRecord(e->row_done[r], e->stream), "hipEventRecord");
      }
    } else {
      for (int r : e->order) {
        ia.row0 = r, ia.nrows = 1;
        if (nrefs) launch_inter(ia, e->stream);
        fa.row0 = r, fa.nrows = 1;
        launch_finalize(fa, e->stream);
        check(hipGetLastError(), "row kernels");
        const size_t off = static_cast<size_t>(r) * row_bytes;
        check(hipMemcpyAsync(static_cast<char*>(e->rec_host) + off, static_cast<char*>(e->rec_dev) + off, row_bytes,
                             hipMemcpyDeviceToHost, e->stream), "hipMemcpyAsync row");
        check(hipEventRecord(e->row_done[r], e->stream), "hipEventRecord");
      }
      check(hipEventRecord(e->ev[3], e->stream), "hipEventRecord");
      check(hipEventRecord(e->ev[4], e->stream), "hipEventRecord");
    }
    check(hipEventRecord(e->ev[5], e->stream), "hipEventRecord");
    e->release(cs, false);
    e->nrefs_last = nrefs;
    e->last_refs = used;
    e->launched = true;
  });
}

int celv2hip_row_order(const celv2hip_engine* e, int* rows, int cap) {
  if (!e) return -CELV2_ERR_ARG;
  const int n = static_cast<int>(e->order.size());
  for (int i = 0; i < n && i < cap && rows; i++) rows[i] = e->order[i];
  return n;
}

int celv2hip_wait_row(celv2hip_engine* e, int row, celv2_sb* out) {
  if (int rc = validate(e)) return rc;
  return guarded([&] {
    if (!e->launched) throw Failure(CELV2_ERR_STATE, "HipEngine: no launch to collect");
    if (e->resident) throw Failure(CELV2_ERR_STATE, "HipEngine: resident mode keeps results on the device");
    if (row < 0 || row >= e->sb_rows) throw Failure(CELV2_ERR_ARG, "HipEngine: no such row");
    check(hipEventSynchronize(e->row_done[row]), "row wait");
    if (out) std::memcpy(out, static_cast<char*>(e->rec_host) + static_cast<size_t>(row) * e->sb_cols * sizeof(celv2_sb), e->sb_cols * sizeof(celv2_sb));
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

int celv2hip_fetch_plane(celv2hip_engine* e, int ref, int level, int phase, uint8_t* out, int* w, int* h) {
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
    *w = d.w, *h = d.h;
    e->sync_all();
    check(hipMemcpy2D(out, d.w, d.p + static_cast<ptrdiff_t>(phase) * p->phase_bytes, d.stride, d.w, d.h, hipMemcpyDeviceToHost), "hipMemcpy2D");
  });
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
    check(hipMemset(e->counters, 0, 80 * 16 * sizeof(unsigned long long)), "hipMemset counters");
    e->window_launch = e->launches, e->window_prepare = e->prepares;
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
    double acc[7] = {};
    for (long long i = e->window_launch; i < e->launches; i++) {
      auto& a = e->ring[i % celv2hip_engine::kRing];
      const int pairs[6][2] = {{0, 1}, {1, 2}, {2, 3}, {3, 4}, {4, 5}, {0, 5}};
      for (int k = 0; k < 6; k++) {
        float t = 0;
        check(hipEventElapsedTime(&t, a[pairs[k][0]], a[pairs[k][1]]), "hipEventElapsedTime");
        acc[k] += t;
      }
    }
    for (long long i = e->window_prepare; i < e->prepares; i++) {
      auto& a = e->pring[i % celv2hip_engine::kRing];
      float t = 0;
      check(hipEventElapsedTime(&t, a[0], a[1]), "hipEventElapsedTime");
      acc[6] += t;
    }
    for (int k = 0; k < 7 && k < cap_ms; k++) ms[k] = acc[k];
    check(hipMemcpy(e->counters_host, e->counters, 80 * 16 * sizeof(unsigned long long), hipMemcpyDeviceToHost), "hipMemcpy counters");
    unsigned long long searches = 0;
    for (int i = 0; i < 64; i++) searches += e->counters_host[16 + 16 * i];
    const long long v[5] = {n, static_cast<long long>(e->counters_host[0]), static_cast<long long>(searches),
                            static_cast<long long>(e->counters_host[1]), static_cast<long long>(e->counters_host[2])};
    for (int k = 0; k < 5 && k < cap_counts; k++) counts[k] = v[k];
  });
  return rc ? -rc : static_cast<int>(n);
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

int celv2hip_copy_on_device(int device, void* dst, const void* src, long long bytes) {
  return guarded([&] {
    check(hipSetDevice(device), "hipSetDevice");
    check(hipMemcpy(dst, src, static_cast<size_t>(bytes), hipMemcpyDeviceToDevice), "hipMemcpy D2D");
    check(hipDeviceSynchronize(), "hipDeviceSynchronize");
  });
}

}  // extern "C"
