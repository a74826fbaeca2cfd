// C entry points over celv2::HipEngine (the C++ DecisionEngine), so tests and the benchmark drive
// the same object the encoder would: launch / prepare / collect with a RowGate, and the resulting
// DecisionMap exported as flat records (include/celv2hip.h).
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "celv2hip.h"
#include "celv2/hip_engine.h"

namespace {

thread_local std::string g_err;

struct Handle {
  std::unique_ptr<celv2::HipEngine> e;
  DecisionMap dm;
  int width, height;
  // Frames by key, as a pipeline holds its sources and reconstructions: a keyed frame is copied in
  // once; a launch keeps its source and references, prepare() adds the next source.
  std::map<long long, Frame420> frames;
};

Frame420 frame_of(int w, int h, const uint8_t* y8, const uint16_t* y16, int bit_depth) {
  Frame420 f;
  f.width = w, f.height = h, f.pw = (w + 7) & ~7, f.ph = (h + 7) & ~7;
  const size_t n = static_cast<size_t>(f.pw) * f.ph;
  f.bit_depth = y16 ? bit_depth : 8;
  if (y8) f.y.assign(y8, y8 + n);
  if (y16) f.hy.assign(y16, y16 + n);
  return f;
}

}  // namespace

extern "C" {

const char* celv2_engine_last_error(void) { return g_err.c_str(); }

void* celv2_engine_create(int width, int height, int device, int row_streams, int full_depth) {
  try {
    auto h = std::make_unique<Handle>();
    h->e = std::make_unique<celv2::HipEngine>(width, height, 64, device);
    h->e->set_row_streams(row_streams);
    h->e->set_upload_full_depth(full_depth != 0);
    h->width = width, h->height = height;
    return h.release();
  } catch (const std::exception& ex) {
    g_err = ex.what();
    return nullptr;
  }
}

void celv2_engine_destroy(void* p) { delete static_cast<Handle*>(p); }

// Frames: y8[i] (padded 8-bit luma) and, when y16 is given, y16[i] at bit_depth. lam >= 0 overrides
// lambda_q8(qindex).
int celv2_engine_launch(void* p, const uint8_t* src8, const uint16_t* src16, int bit_depth, const uint8_t* const* refs8,
                        const uint16_t* const* refs16, const long long* keys, int nrefs, int qindex, long long lam,
                        long long src_key) {
  Handle* h = static_cast<Handle*>(p);
  try {
    std::map<long long, Frame420> keep;
    std::vector<Frame420> unkeyed;
    unkeyed.reserve(static_cast<size_t>(nrefs) + 1);
    // The frame under key (copied in unless held already), or an unkeyed copy.
    auto frame = [&](long long key, const uint8_t* y8, const uint16_t* y16) -> const Frame420* {
      if (key < 0) return &unkeyed.emplace_back(frame_of(h->width, h->height, y8, y16, bit_depth));
      auto it = keep.find(key);
      if (it != keep.end()) return &it->second;
      auto held = h->frames.find(key);
      if (held != h->frames.end()) return &keep.emplace(key, std::move(held->second)).first->second;
      return &keep.emplace(key, frame_of(h->width, h->height, y8, y16, bit_depth)).first->second;
    };
    const Frame420* src = frame(src_key, src8, src16);
    std::vector<const Frame420*> rp;
    for (int i = 0; i < nrefs; i++) rp.push_back(frame(keys ? keys[i] : -1, refs8 ? refs8[i] : nullptr, refs16 ? refs16[i] : nullptr));
    std::vector<long long> k;
    if (keys) k.assign(keys, keys + nrefs);
    h->e->set_lambda(lam);
    h->e->launch(*src, rp, qindex, h->dm, k, src_key);
    h->frames = std::move(keep);
    return 0;
  } catch (const std::invalid_argument& ex) {
    g_err = ex.what();
    return 1;
  } catch (const std::exception& ex) {
    g_err = ex.what();
    return 2;
  }
}

// Frames in device memory (celv2_luma with on_device set), through HipEngine::launch_device.
int celv2_engine_launch_device(void* p, const celv2_luma* src, const celv2_luma* refs, const long long* keys, int nrefs,
                               int qindex, long long lam, long long src_key) {
  Handle* h = static_cast<Handle*>(p);
  try {
    if (!src || nrefs < 0 || (nrefs > 0 && !refs)) throw std::invalid_argument("launch_device: bad arguments");
    std::vector<celv2_luma> r(refs, refs + nrefs);
    std::vector<long long> k;
    if (keys) k.assign(keys, keys + nrefs);
    h->e->set_lambda(lam);
    h->e->launch_device(*src, r, qindex, h->dm, k, src_key);
    return 0;
  } catch (const std::invalid_argument& ex) {
    g_err = ex.what();
    return 1;
  } catch (const std::exception& ex) {
    g_err = ex.what();
    return 2;
  }
}

int celv2_engine_prepare_device(void* p, const celv2_luma* next, long long key) {
  Handle* h = static_cast<Handle*>(p);
  try {
    if (!next) throw std::invalid_argument("prepare_device: no frame");
    h->e->prepare_device(*next, key);
    return 0;
  } catch (const std::exception& ex) {
    g_err = ex.what();
    return 2;
  }
}

int celv2_engine_prepare(void* p, const uint8_t* y8, const uint16_t* y16, int bit_depth, long long key) {
  Handle* h = static_cast<Handle*>(p);
  try {
    if (key >= 0) {
      auto it = h->frames.find(key);
      if (it == h->frames.end()) it = h->frames.emplace(key, frame_of(h->width, h->height, y8, y16, bit_depth)).first;
      h->e->prepare(it->second, key);
    } else {
      h->e->prepare(frame_of(h->width, h->height, y8, y16, bit_depth), key);
    }
    return 0;
  } catch (const std::exception& ex) {
    g_err = ex.what();
    return 2;
  }
}

// Collects into the DecisionMap and exports it (out NULL: not exported); rows_ready: the gate's ready
// rows afterwards.
int celv2_engine_collect(void* p, celv2_sb* out, int* rows_ready) {
  Handle* h = static_cast<Handle*>(p);
  RowGate gate;
  try {
    h->e->collect(h->dm, gate);
    if (rows_ready) *rows_ready = gate.ready_rows();
    if (out)
      for (size_t i = 0; i < h->dm.sbs.size(); i++) celv2::decision_to_sb_record(h->dm.sbs[i], out[i]);
    return 0;
  } catch (const std::exception& ex) {
    if (rows_ready) *rows_ready = gate.ready_rows();
    g_err = ex.what();
    return 2;
  }
}

int celv2_engine_set_row_groups(void* p, int groups) {
  try {
    static_cast<Handle*>(p)->e->set_row_groups(groups);
    return 0;
  } catch (const std::invalid_argument& ex) {
    g_err = ex.what();
    return 1;
  }
}

int celv2_engine_row_groups(void* p) { return static_cast<Handle*>(p)->e->row_groups(); }

// HipEngine::autotune_device; writes the chosen count, then (row groups, median ms) per candidate timed
// (none when the count came from the process's earlier tuning) up to cap, and their number.
int celv2_engine_autotune_device(void* p, const celv2_luma* src, const celv2_luma* refs, const long long* keys, int nrefs,
                                 int qindex, long long lam, long long src_key, int frames, int force, int* groups,
                                 int* cand_groups, double* cand_ms, int cap, int* ncand) {
  Handle* h = static_cast<Handle*>(p);
  try {
    if (!src || nrefs < 0 || (nrefs > 0 && !refs)) throw std::invalid_argument("autotune_device: bad arguments");
    std::vector<celv2_luma> r(refs, refs + nrefs);
    std::vector<long long> k;
    if (keys) k.assign(keys, keys + nrefs);
    h->e->set_lambda(lam);
    const celv2::HipEngine::Tuning t = h->e->autotune_device(*src, r, qindex, k, src_key, frames, force != 0);
    if (groups) *groups = t.groups;
    int n = 0;
    for (const auto& [g, ms] : t.ms) {
      if (n >= cap) break;
      if (cand_groups) cand_groups[n] = g;
      if (cand_ms) cand_ms[n] = ms;
      n++;
    }
    if (ncand) *ncand = n;
    return 0;
  } catch (const std::invalid_argument& ex) {
    g_err = ex.what();
    return 1;
  } catch (const std::exception& ex) {
    g_err = ex.what();
    return 2;
  }
}

int celv2_engine_sync(void* p) {
  try {
    static_cast<Handle*>(p)->e->synchronize();
    return 0;
  } catch (const std::exception& ex) {
    g_err = ex.what();
    return 2;
  }
}

int celv2_engine_description(void* p, char* buf, int cap) {
  const std::string d = static_cast<Handle*>(p)->e->description();
  std::snprintf(buf, static_cast<size_t>(cap), "%s", d.c_str());
  return static_cast<int>(d.size());
}

int celv2_sizeof_sb(void) { return static_cast<int>(sizeof(celv2_sb)); }

#ifndef CELV2_SOURCE_HASH
#define CELV2_SOURCE_HASH "unknown"
#endif
const char* celv2hip_source_hash(void) { return CELV2_SOURCE_HASH; }

}  // extern "C"
