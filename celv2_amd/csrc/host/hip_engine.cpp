// celv2::HipEngine (include/celv2/hip_engine.h) over the C ABI of include/celv2hip.h.
#include "celv2/hip_engine.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstddef>
#include <condition_variable>
#include <cstring>
#include <functional>
#include <iterator>
#include <map>
#include <stdexcept>
#include <string>
#include <mutex>
#include <thread>
#include <tuple>
#include <type_traits>
#include <vector>

#include "celv2/lambda.h"
#include "celv2hip.h"
#ifdef CELV2_IN_TREE
#include "gpu/kernels_ref.h"  // CelV2's lambda_q8 (defined in src/core/quant.cpp)
#endif

namespace celv2 {

namespace {

[[noreturn]] void raise(int rc, const char* what) {
  const std::string msg = std::string(what) + ": " + celv2hip_last_error();
  if (rc == CELV2_ERR_ARG) throw std::invalid_argument(msg);
  throw std::runtime_error(msg);
}

celv2_luma luma_of(const Frame420& f, bool full_depth) {
  celv2_luma l{};
  if (full_depth && f.bit_depth > 8 && !f.hy.empty()) {
    l.y16 = f.hy.data(), l.bit_depth = f.bit_depth;
  } else {
    l.y8 = f.y.data(), l.bit_depth = 8;
  }
  return l;
}

BlockDecision block_of(const celv2_block& b) {
  BlockDecision d;
  d.w4 = b.w4, d.h4 = b.h4, d.inter = b.inter != 0, d.intra = static_cast<IntraMode>(b.intra);
  d.mv_row = b.mv_row, d.mv_col = b.mv_col, d.ref = b.ref, d.ref2 = b.ref2;
  d.mv2_row = b.mv2_row, d.mv2_col = b.mv2_col;
  return d;
}

celv2_block record_of(const BlockDecision& d) {
  celv2_block b{};
  b.w4 = d.w4, b.h4 = d.h4, b.inter = d.inter ? 1 : 0, b.intra = static_cast<uint8_t>(d.intra);
  b.mv_row = d.mv_row, b.mv_col = d.mv_col, b.ref = d.ref, b.ref2 = d.ref2;
  b.mv2_row = d.mv2_row, b.mv2_col = d.mv2_col;
  return b;
}

}  // namespace

// The records' structs have the DecisionMap types' layout (same sizes, field offsets and 1-byte
// enums), so a record converts by copying bytes; where a compiler lays them out otherwise, field by
// field.
constexpr bool kSameLayout =
    std::is_trivially_copyable_v<BlockDecision> && std::is_trivially_copyable_v<BlockCandidates> &&
    sizeof(Part) == 1 && sizeof(IntraMode) == 1 && sizeof(bool) == 1 && sizeof(BlockDecision) == sizeof(celv2_block) &&
    offsetof(BlockDecision, w4) == offsetof(celv2_block, w4) && offsetof(BlockDecision, h4) == offsetof(celv2_block, h4) &&
    offsetof(BlockDecision, inter) == offsetof(celv2_block, inter) && offsetof(BlockDecision, intra) == offsetof(celv2_block, intra) &&
    offsetof(BlockDecision, mv_row) == offsetof(celv2_block, mv_row) && offsetof(BlockDecision, mv_col) == offsetof(celv2_block, mv_col) &&
    offsetof(BlockDecision, ref) == offsetof(celv2_block, ref) && offsetof(BlockDecision, ref2) == offsetof(celv2_block, ref2) &&
    offsetof(BlockDecision, mv2_row) == offsetof(celv2_block, mv2_row) && offsetof(BlockDecision, mv2_col) == offsetof(celv2_block, mv2_col) &&
    sizeof(BlockCandidates) == sizeof(celv2_cands) && offsetof(BlockCandidates, inter_cost) == offsetof(celv2_cands, inter_cost) &&
    offsetof(BlockCandidates, intra_cost) == offsetof(celv2_cands, intra_cost) && offsetof(BlockCandidates, inter) == offsetof(celv2_cands, inter) &&
    offsetof(BlockCandidates, intra) == offsetof(celv2_cands, intra) && offsetof(BlockCandidates, ref_cost) == offsetof(celv2_cands, ref_cost) &&
    offsetof(BlockCandidates, ref_mv_col) == offsetof(celv2_cands, ref_mv_col) && offsetof(BlockCandidates, ref_mv_row) == offsetof(celv2_cands, ref_mv_row) &&
    BlockCandidates::kRefs == CELV2_MAX_REFS;

namespace {

// A forward iterator over records that yields each one as the DecisionMap type with the same layout
// (byte copy), so a vector is built from records in one pass: assign() constructs each element from
// its record, where resize() + memcpy would write every element twice.
template <class To, class From>
struct AsIter {
  using iterator_category = std::forward_iterator_tag;
  using value_type = To;
  using difference_type = std::ptrdiff_t;
  using pointer = const To*;
  using reference = To;
  const From* p;
  To operator*() const {
    To t;
    std::memcpy(static_cast<void*>(&t), p, sizeof(To));
    return t;
  }
  AsIter& operator++() {
    ++p;
    return *this;
  }
  AsIter operator++(int) {
    AsIter o = *this;
    ++p;
    return o;
  }
  bool operator==(const AsIter& o) const { return p == o.p; }
  bool operator!=(const AsIter& o) const { return p != o.p; }
};

template <class To, class From>
void assign_records(std::vector<To>& v, const From* first, int n) {
  static_assert(sizeof(To) == sizeof(From) && std::is_trivially_copyable_v<To>, "record layouts differ");
  v.assign(AsIter<To, From>{first}, AsIter<To, From>{first + n});
}

}  // namespace

// Worker threads that run one job at a time beside the calling thread: run(job) starts job on every
// worker, runs it on the caller, and returns once every worker has finished it (so no worker touches
// a job, or what it refers to, after run returns). The job divides its work itself.
class CollectPool {
 public:
  explicit CollectPool(int workers) {
    for (int k = 0; k < workers; k++) threads_.emplace_back([this] { loop(); });
  }
  ~CollectPool() {
    {
      std::lock_guard<std::mutex> lock(mu_);
      stop_ = true;
    }
    start_.notify_all();
    for (auto& t : threads_) t.join();
  }
  void run(const std::function<void()>& job) {
    {
      std::lock_guard<std::mutex> lock(mu_);
      job_ = &job;
      finished_ = 0;
      gen_++;
    }
    start_.notify_all();
    job();
    std::unique_lock<std::mutex> lock(mu_);
    done_.wait(lock, [&] { return finished_ == static_cast<int>(threads_.size()); });
    job_ = nullptr;
  }

 private:
  void loop() {
    long long seen = 0;
    for (;;) {
      const std::function<void()>* job;
      {
        std::unique_lock<std::mutex> lock(mu_);
        start_.wait(lock, [&] { return stop_ || gen_ != seen; });
        if (stop_) return;
        seen = gen_;
        job = job_;
      }
      (*job)();
      {
        std::lock_guard<std::mutex> lock(mu_);
        finished_++;
      }
      done_.notify_all();
    }
  }
  std::vector<std::thread> threads_;
  std::mutex mu_;
  std::condition_variable start_, done_;
  const std::function<void()>* job_ = nullptr;
  long long gen_ = 0;
  int finished_ = 0;
  bool stop_ = false;
};

void sb_record_to_decision(const celv2_sb& rec, SbDecision& d) {
  if constexpr (kSameLayout) {
    // inter / intra bytes of the records are 0 or 1 and a mode index: valid values of bool and IntraMode.
    assign_records(d.tree, reinterpret_cast<const Part*>(rec.tree), rec.ntree);
    assign_records(d.blocks, rec.blocks, rec.nblocks);
    assign_records(d.cands, rec.cands, rec.ncands);
    return;
  }
  d.tree.resize(rec.ntree);
  for (int k = 0; k < rec.ntree; k++) d.tree[k] = static_cast<Part>(rec.tree[k]);
  d.blocks.resize(rec.nblocks);
  for (int k = 0; k < rec.nblocks; k++) d.blocks[k] = block_of(rec.blocks[k]);
  d.cands.resize(rec.ncands);
  for (int k = 0; k < rec.ncands; k++) {
    const celv2_cands& c = rec.cands[k];
    BlockCandidates& o = d.cands[k];
    o.inter_cost = c.inter_cost, o.intra_cost = c.intra_cost;
    o.inter = block_of(c.inter);
    o.intra = static_cast<IntraMode>(c.intra);
    for (int r = 0; r < BlockCandidates::kRefs; r++)
      o.ref_cost[r] = c.ref_cost[r], o.ref_mv_col[r] = c.ref_mv_col[r], o.ref_mv_row[r] = c.ref_mv_row[r];
  }
}

void decision_to_sb_record(const SbDecision& d, celv2_sb& rec) {
  std::memset(&rec, 0, sizeof(rec));
  if (d.tree.size() > CELV2_SB_CAP || d.blocks.size() > CELV2_SB_CAP || d.cands.size() > CELV2_SB_BLOCKS)
    throw std::runtime_error("superblock decision exceeds the record");
  rec.ntree = static_cast<int32_t>(d.tree.size());
  rec.nblocks = static_cast<int32_t>(d.blocks.size());
  rec.ncands = static_cast<int32_t>(d.cands.size());
  for (size_t k = 0; k < d.tree.size(); k++) rec.tree[k] = static_cast<uint8_t>(d.tree[k]);
  for (size_t k = 0; k < d.blocks.size(); k++) rec.blocks[k] = record_of(d.blocks[k]);
  for (size_t k = 0; k < d.cands.size(); k++) {
    const BlockCandidates& c = d.cands[k];
    celv2_cands& o = rec.cands[k];
    o.inter_cost = c.inter_cost, o.intra_cost = c.intra_cost;
    o.inter = record_of(c.inter);
    o.intra = static_cast<uint8_t>(c.intra);
    for (int r = 0; r < BlockCandidates::kRefs; r++)
      o.ref_cost[r] = c.ref_cost[r], o.ref_mv_col[r] = c.ref_mv_col[r], o.ref_mv_row[r] = c.ref_mv_row[r];
  }
}

HipEngine::HipEngine(int width, int height, int sb_size, int device)
    : width_(width), height_(height), sb_cols_((width + 63) / 64), sb_rows_((height + 63) / 64), device_(device) {
  if (width <= 0 || height <= 0) throw std::runtime_error("HipEngine: frame size must be positive");
  if (sb_size != 64) throw std::runtime_error("HipEngine: only 64x64 superblocks are supported");
  const int rc = celv2hip_create(width, height, sb_size, device, &h_);
  if (rc != CELV2_OK) throw std::runtime_error(std::string("HipEngine: ") + celv2hip_last_error());
  char name[256] = {};
  celv2hip_device_name(h_, name, sizeof(name));
  name_ = name;
}

HipEngine::~HipEngine() {
  pool_.reset();
  celv2hip_destroy(h_);
}

void HipEngine::set_row_streams(int n) {
  celv2hip_set_row_streams(h_, n);
  row_streams_ = n;
}

void HipEngine::set_row_groups(int groups) {
  const int rc = celv2hip_set_row_groups(h_, groups);
  if (rc != CELV2_OK) throw std::invalid_argument(std::string("HipEngine::set_row_groups: ") + celv2hip_last_error());
}

int HipEngine::row_groups() const { return celv2hip_row_groups(h_); }

HipEngine::Tuning HipEngine::autotune(const Frame420& src, const std::vector<const Frame420*>& refs, int qindex,
                                      const std::vector<long long>& ref_keys, long long src_key, int frames, bool force) {
  return tune(
      static_cast<int>(refs.size()), frames, force,
      [&](DecisionMap& dm) { launch(src, refs, qindex, dm, ref_keys, src_key); },
      [&] { prepare(src, src_key); });
}

HipEngine::Tuning HipEngine::autotune_device(const celv2_luma& src, const std::vector<celv2_luma>& refs, int qindex,
                                             const std::vector<long long>& ref_keys, long long src_key, int frames,
                                             bool force) {
  return tune(
      static_cast<int>(refs.size()), frames, force,
      [&](DecisionMap& dm) { launch_device(src, refs, qindex, dm, ref_keys, src_key); },
      [&] { prepare_device(src, src_key); });
}

HipEngine::Tuning HipEngine::tune(int nrefs, int frames, bool force, const std::function<void(DecisionMap&)>& launch_one,
                                  const std::function<void()>& prepare_one) {
  // Best counts by (width, height, references, device), shared by the process's engines.
  static std::mutex mu;
  static std::map<std::tuple<int, int, int, int>, int> best;
  const auto key = std::make_tuple(width_, height_, nrefs, device_);
  Tuning t;
  if (std::max(1, std::min(row_streams_, sb_rows_)) != 1) {
    t.groups = row_groups();
    return t;
  }
  if (!force) {
    std::lock_guard<std::mutex> lock(mu);
    auto it = best.find(key);
    if (it != best.end()) {
      set_row_groups(it->second);
      t.groups = it->second, t.cached = true;
      return t;
    }
  }
  // Candidates: distinct rows per group (the launch splits the rows into runs of ceil(rows / groups)).
  std::vector<int> cands;
  std::vector<int> seen_rows;
  for (int g : {1, 2, 3, 4, 5, 6, 8, 12, 16}) {
    if (g > sb_rows_) break;
    const int rows = (sb_rows_ + g - 1) / g;
    if (std::find(seen_rows.begin(), seen_rows.end(), rows) != seen_rows.end()) continue;
    seen_rows.push_back(rows);
    cands.push_back((sb_rows_ + rows - 1) / rows);
  }
  const int was = row_groups();
  DecisionMap dm;
  RowGate gate;
  auto frame = [&] {
    launch_one(dm);
    prepare_one();
    collect(dm, gate);
  };
  std::vector<std::vector<double>> ms(cands.size());
  try {
    frames = std::max(1, frames);
    prepare_one();
    for (size_t c = 0; c < cands.size(); c++) {  // one untimed frame at each count
      set_row_groups(cands[c]);
      frame();
    }
    for (int f = 0; f < frames; f++)
      for (size_t c = 0; c < cands.size(); c++) {
        set_row_groups(cands[c]);
        const auto t0 = std::chrono::steady_clock::now();
        frame();
        ms[c].push_back(std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count());
      }
    synchronize();
  } catch (...) {
    set_row_groups(was);
    throw;
  }
  size_t pick = 0;
  for (size_t c = 0; c < cands.size(); c++) {
    std::sort(ms[c].begin(), ms[c].end());
    t.ms.emplace_back(cands[c], ms[c][ms[c].size() / 2]);
    if (t.ms[c].second < t.ms[pick].second) pick = c;
  }
  t.groups = cands[pick];
  set_row_groups(t.groups);
  std::lock_guard<std::mutex> lock(mu);
  best[key] = t.groups;
  return t;
}

void HipEngine::synchronize() {
  const int rc = celv2hip_sync(h_);
  if (rc != CELV2_OK) raise(rc, "HipEngine::synchronize");
}

void HipEngine::prepare(const Frame420& next, long long key) {
  if (key < 0 || next.width != width_ || next.height != height_) return;
  const celv2_luma l = luma_of(next, full_depth_);
  const int rc = celv2hip_prepare(h_, &l, key);
  if (rc != CELV2_OK) raise(rc, "HipEngine::prepare");
}

void HipEngine::launch(const Frame420& src, const std::vector<const Frame420*>& refs, int qindex, DecisionMap& dm,
                       const std::vector<long long>& ref_keys, long long src_key) {
  if (src.width != width_ || src.height != height_) throw std::runtime_error("HipEngine: source size does not match the engine");
  for (const Frame420* r : refs)
    if (r->width != width_ || r->height != height_) throw std::runtime_error("HipEngine: reference size does not match the engine");
  if (refs.size() > 7) throw std::runtime_error("HipEngine: too many references");
  if (!ref_keys.empty() && ref_keys.size() != refs.size()) throw std::invalid_argument("HipEngine: one key per reference");
  const celv2_luma s = luma_of(src, full_depth_);
  std::vector<celv2_luma> r;
  for (const Frame420* f : refs) r.push_back(luma_of(*f, full_depth_));
  launch_luma(s, r.data(), static_cast<int>(r.size()), qindex, dm, ref_keys, src_key);
}

void HipEngine::launch_device(const celv2_luma& src, const std::vector<celv2_luma>& refs, int qindex, DecisionMap& dm,
                              const std::vector<long long>& ref_keys, long long src_key) {
  if (refs.size() > 7) throw std::runtime_error("HipEngine: too many references");
  if (!ref_keys.empty() && ref_keys.size() != refs.size()) throw std::invalid_argument("HipEngine: one key per reference");
  if (!src.on_device) throw std::invalid_argument("HipEngine::launch_device: the source is not in device memory");
  for (const celv2_luma& l : refs)
    if (!l.on_device) throw std::invalid_argument("HipEngine::launch_device: a reference is not in device memory");
  launch_luma(src, refs.data(), static_cast<int>(refs.size()), qindex, dm, ref_keys, src_key);
}

void HipEngine::prepare_device(const celv2_luma& next, long long key) {
  if (key < 0) return;
  if (!next.on_device) throw std::invalid_argument("HipEngine::prepare_device: the frame is not in device memory");
  const int rc = celv2hip_prepare(h_, &next, key);
  if (rc != CELV2_OK) raise(rc, "HipEngine::prepare");
}

void HipEngine::launch_luma(const celv2_luma& src, const celv2_luma* refs, int nrefs, int qindex, DecisionMap& dm,
                            const std::vector<long long>& ref_keys, long long src_key) {
  const int64_t lam = lambda_ >= 0 ? lambda_ : lambda_q8(qindex);
  const int rc = celv2hip_launch(h_, &src, src_key, refs, ref_keys.empty() ? nullptr : ref_keys.data(), nrefs, lam);
  if (rc != CELV2_OK) raise(rc, "HipEngine::launch");
  dm.sb_size = 64, dm.sb_cols = sb_cols_, dm.sb_rows = sb_rows_;
  dm.sbs.assign(static_cast<size_t>(sb_cols_) * sb_rows_, SbDecision{});
}

void HipEngine::collect(DecisionMap& dm, RowGate& gate) {
  std::vector<int> order(sb_rows_);
  const int n = celv2hip_row_order(h_, order.data(), sb_rows_);
  if (n != sb_rows_) {
    const auto e = std::make_exception_ptr(std::runtime_error("HipEngine::collect: no launch to collect"));
    gate.fail(e);
    celv2hip_sync(h_);
    std::rethrow_exception(e);
  }
  // Rows are converted in chunks of up to 8 superblocks, taken in queue order by up to 8 threads (this
  // one included): each waits for its row's records, converts them straight from the engine's pinned
  // buffer, and the thread that converts a row's last chunk publishes the row. The first failure stops
  // the others taking chunks, fails the gate and is rethrown once every thread has finished.
  constexpr int kChunk = 8, kThreads = 8;
  struct Chunk {
    int row, sb0, sb1;
  };
  std::vector<Chunk> chunks;
  std::vector<std::atomic<int>> left(static_cast<size_t>(sb_rows_));
  for (int i = 0; i < n; i++) {
    int k = 0;
    for (int c = 0; c < sb_cols_; c += kChunk, k++) chunks.push_back({order[i], c, std::min(sb_cols_, c + kChunk)});
    left[order[i]] = k;
  }
  std::atomic<int> next{0};
  std::atomic<bool> failed{false};
  std::exception_ptr err;
  std::mutex err_mu;
  const int total = static_cast<int>(chunks.size());
  const std::function<void()> work = [&] {
    for (int i; !failed.load() && (i = next.fetch_add(1)) < total;) {
      try {
        const Chunk& ch = chunks[i];
        const celv2_sb* row = nullptr;
        const int rc = celv2hip_row_records(h_, ch.row, &row);
        if (rc != CELV2_OK) raise(rc, "HipEngine::collect");
        for (int c = ch.sb0; c < ch.sb1; c++) sb_record_to_decision(row[c], dm.sbs[static_cast<size_t>(ch.row) * sb_cols_ + c]);
        if (left[ch.row].fetch_sub(1) == 1) gate.publish_one(ch.row);
      } catch (...) {
        std::lock_guard<std::mutex> lock(err_mu);
        if (!err) err = std::current_exception();
        failed = true;
      }
    }
  };
  if (!pool_) {
    const unsigned hw = std::thread::hardware_concurrency();
    pool_ = std::make_unique<CollectPool>(std::max(0, std::min(kThreads, static_cast<int>(hw ? hw : 1)) - 1));
  }
  pool_->run(work);
  if (err) {
    gate.fail(err);
    celv2hip_sync(h_);  // returns with no work pending
    std::rethrow_exception(err);
  }
}

}  // namespace celv2

#ifndef CELV2_IN_TREE
// The mirror build has no AVM: lambda from the stated step model (include/celv2/lambda.h).
int64_t lambda_q8(int qindex) { return celv2_lambda_q8(qindex); }
#endif

extern "C" int64_t celv2_lambda_q8(int qindex) {
  auto step = [](int q) { return (4.0 + q + static_cast<double>(q) * q * q / 9000.0) / 8.0; };
  const double qstep = step(qindex), pivot = step(120);
  return std::llround(0.002 * qstep * qstep * std::pow(pivot / qstep, 0.8) * 256.0);
}
