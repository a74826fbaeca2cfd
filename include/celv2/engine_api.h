// CelV2's decision-engine interface, declared for building the HIP engine outside the CelV2 tree.
//
// The types keep the reference's names, fields and meanings so that celv2::HipEngine compiles
// unchanged against CelV2's own headers (build with -DCELV2_IN_TREE and CelV2's src/ on the
// include path):
//   Frame420 (the fields the engines read)   src/io/y4m.h:11-23
//   Part, IntraMode, BlockDecision,
//   BlockCandidates, SbDecision, DecisionMap src/gpu/decision_map.h:17-78
//   RowGate                                  src/gpu/decision_map.h:83-108, decision_map.cpp:19-61
//   DecisionEngine                           src/gpu/decision_engine.h:9-51
#pragma once
#include <chrono>
#include <climits>
#include <condition_variable>
#include <cstdint>
#include <exception>
#include <mutex>
#include <vector>

// A frame as the engines see it: the padded 8-bit luma (pw x ph, stride pw) the decisions read,
// and above 8 bits the same plane at full depth in hy.
struct Frame420 {
  int width = 0, height = 0;
  int pw = 0, ph = 0;
  std::vector<uint8_t> y, u, v;
  int bit_depth = 8;
  std::vector<uint16_t> hy, hu, hv;
  int stride_y() const { return pw; }
  int stride_uv() const { return pw / 2; }
};

enum class Part : uint8_t { None, Horz, Vert, Split };
enum class IntraMode : uint8_t { DC, V, H, Smooth, SmoothV, SmoothH, Paeth, D45, D67, D113, D135, D157, D203 };
constexpr int kIntraModes = 13;

struct BlockDecision {
  uint8_t w4 = 0, h4 = 0;  // block size in 4-pixel units
  bool inter = false;
  IntraMode intra = IntraMode::DC;
  int16_t mv_row = 0, mv_col = 0;  // 1/8 pel
  uint8_t ref = 0;
  static constexpr uint8_t kSingleRef = 255;
  uint8_t ref2 = kSingleRef;
  int16_t mv2_row = 0, mv2_col = 0;
};

struct BlockCandidates {
  int64_t inter_cost = INT64_MAX, intra_cost = INT64_MAX;
  BlockDecision inter;
  IntraMode intra = IntraMode::DC;
  static constexpr int kRefs = 7;
  uint32_t ref_cost[kRefs] = {UINT32_MAX, UINT32_MAX, UINT32_MAX, UINT32_MAX, UINT32_MAX, UINT32_MAX, UINT32_MAX};
  int16_t ref_mv_col[kRefs] = {}, ref_mv_row[kRefs] = {};
};

struct SbDecision {
  std::vector<Part> tree;             // preorder; SPLIT has 4 children, HORZ/VERT 2
  std::vector<BlockDecision> blocks;  // leaves in coding order
  std::vector<BlockCandidates> cands; // the 425 candidate blocks (64x64 superblocks)
};

struct DecisionMap {
  int sb_size = 0;
  int sb_cols = 0, sb_rows = 0;
  std::vector<SbDecision> sbs;  // raster order
};

// Rows of a DecisionMap handed from the collector to the tile threads; a failure is recorded once
// and rethrown to every waiter. This standalone build keeps the ready rows as a bitmap and the
// length of its leading run of set bits (ready_rows()); in-tree builds use CelV2's own RowGate.
class RowGate {
 public:
  RowGate() = default;
  RowGate(const RowGate&) = delete;
  RowGate& operator=(const RowGate&) = delete;
  void publish(int row);            // rows 0..row are ready
  void publish_one(int row);        // row is ready (any order)
  void wait_row(int row) const;     // blocks until row is ready; rethrows a failure
  void fail(std::exception_ptr e);  // records e, wakes every waiter
  int ready_rows() const;
  std::chrono::steady_clock::time_point first_ready() const;

 private:
  void mark_locked(int first, int last);  // sets rows first..last, then extends the leading run
  mutable std::mutex mu_;
  mutable std::condition_variable changed_;
  std::vector<uint64_t> bits_;  // bit r of word r / 64: row r is ready
  int run_ = 0;                 // rows 0 .. run_ - 1 are all ready
  std::exception_ptr failure_;
  std::chrono::steady_clock::time_point first_ready_{};
};

class DecisionEngine {
 public:
  virtual ~DecisionEngine() = default;
  virtual void launch(const Frame420& src, const std::vector<const Frame420*>& refs, int qindex, DecisionMap& dm,
                      const std::vector<long long>& ref_keys = {}, long long src_key = -1) = 0;
  virtual void prepare(const Frame420& next, long long key) { (void)next, (void)key; }
  void launch(const Frame420& src, const Frame420* ref, int qindex, DecisionMap& dm) {
    launch(src, ref ? std::vector<const Frame420*>{ref} : std::vector<const Frame420*>{}, qindex, dm);
  }
  virtual void collect(DecisionMap& dm, RowGate& gate) = 0;
  virtual void set_row_streams(int n) { (void)n; }
  DecisionMap decide_inter(const Frame420& src, const std::vector<const Frame420*>& refs, int qindex) {
    DecisionMap dm;
    RowGate gate;
    launch(src, refs, qindex, dm);
    collect(dm, gate);
    return dm;
  }
  DecisionMap decide_inter(const Frame420& src, const Frame420& prev, int qindex) {
    return decide_inter(src, std::vector<const Frame420*>{&prev}, qindex);
  }
  DecisionMap decide_intra(const Frame420& src, int qindex) { return decide_inter(src, std::vector<const Frame420*>{}, qindex); }
};

// lambda * 256 for a qindex on the 8-bit scale (the reference's core/quant.cpp:31-39). The HIP
// engine takes it from here; see celv2/lambda.h for the table this tree provides.
int64_t lambda_q8(int qindex);
