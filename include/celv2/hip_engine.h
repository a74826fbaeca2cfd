// celv2::HipEngine — CelV2's DecisionEngine on one AMD Instinct MI355X (HIP, gfx950).
//
// A third implementation of the engine interface beside CpuEngine (src/gpu/cpu_engine.h:21) and
// VulkanEngine (src/vk/vk_engine.h:15-34), chosen by make_engine(Device::Hip, ...) / --device hip.
// It produces the same DecisionMap as CpuEngine, bit for bit: every superblock's tree, leaves and all
// 425 candidate blocks' choices and costs. Host code is C++; the device work goes through the C ABI of
// include/celv2hip.h.
#pragma once
#include <cstdint>
#include <functional>
#include <memory>
#include <string>
#include <utility>
#include <vector>

#ifdef CELV2_IN_TREE
#include "gpu/decision_engine.h"
#include "gpu/decision_map.h"
#include "io/y4m.h"
#else
#include "celv2/engine_api.h"
#endif

struct celv2hip_engine;
struct celv2_sb;
struct celv2_luma;

namespace celv2 {

class CollectPool;  // collect()'s worker threads (hip_engine.cpp)

class HipEngine : public DecisionEngine {
 public:
  // Throws std::runtime_error when there is no HIP device, the size is not positive or sb_size is not
  // 64 (VulkanEngine::VulkanEngine, src/vk/vk_engine.cpp:138-146).
  HipEngine(int width, int height, int sb_size, int device = 0);
  ~HipEngine() override;
  HipEngine(const HipEngine&) = delete;
  HipEngine& operator=(const HipEngine&) = delete;

  // Copies what it needs from src and refs before returning; std::runtime_error for a size mismatch or
  // more than 7 references, std::invalid_argument for a key count that is not 0 or nrefs.
  using DecisionEngine::launch;
  void launch(const Frame420& src, const std::vector<const Frame420*>& refs, int qindex, DecisionMap& dm,
              const std::vector<long long>& ref_keys = {}, long long src_key = -1) override;
  void prepare(const Frame420& next, long long key) override;
  // launch / prepare for frames already in this engine's device memory (celv2_luma with on_device
  // set: the padded luma, 8-bit or deeper, include/celv2hip.h): no copy from the host, the same
  // decisions and the same collect. No counterpart in CelV2's interface: for callers that keep their
  // frames in HBM.
  void launch_device(const celv2_luma& src, const std::vector<celv2_luma>& refs, int qindex, DecisionMap& dm,
                     const std::vector<long long>& ref_keys = {}, long long src_key = -1);
  void prepare_device(const celv2_luma& next, long long key);
  // Publishes rows in the order they were queued (publish_one); on failure gate.fail + rethrow.
  void collect(DecisionMap& dm, RowGate& gate) override;
  void set_row_streams(int n) override;

  // Waits for all work queued on the engine's streams (prepare()'s included); for timing.
  void synchronize();

  // Row groups of the one-row-stream launch (celv2hip_set_row_groups; 0: the default). The decisions
  // do not depend on them.
  void set_row_groups(int groups);
  int row_groups() const;
  // Picks the row-group count by timing: `frames` steady-state frames (launch with the source
  // prepared -> prepare of the same source under the same key -> collect, after one untimed frame) at
  // each candidate count, round-robin; keeps the count with the fastest median frame and remembers it
  // for every engine of this frame size, reference count and device in the process (a later call
  // finds it there and times nothing; force re-times). With more than one row stream there is
  // nothing to pick. No counterpart in CelV2: its launch shapes are hand-set (src/vk/vk_kernels.cpp:193-209).
  struct Tuning {
    int groups = 0;
    bool cached = false;
    std::vector<std::pair<int, double>> ms;  // (row groups, median ms per frame) per candidate
  };
  Tuning autotune(const Frame420& src, const std::vector<const Frame420*>& refs, int qindex,
                  const std::vector<long long>& ref_keys, long long src_key, int frames = 5, bool force = false);
  Tuning autotune_device(const celv2_luma& src, const std::vector<celv2_luma>& refs, int qindex,
                         const std::vector<long long>& ref_keys, long long src_key, int frames = 5, bool force = false);
  const std::string& device_name() const { return name_; }
  std::string description() const { return "decisions on HIP (" + name_ + ")"; }
  // Frames above 8 bits: send the full-depth luma (hy) and round it on the device instead of
  // sending the 8-bit plane (the same decisions; twice the upload).
  void set_upload_full_depth(bool on) { full_depth_ = on; }
  // Overrides lambda_q8(qindex) (tests that pin the cost scale).
  void set_lambda(int64_t lam) { lambda_ = lam; }

 private:
  void launch_luma(const celv2_luma& src, const celv2_luma* refs, int nrefs, int qindex, DecisionMap& dm,
                   const std::vector<long long>& ref_keys, long long src_key);
  Tuning tune(int nrefs, int frames, bool force, const std::function<void(DecisionMap&)>& launch,
              const std::function<void()>& prepare);
  int width_, height_, sb_cols_, sb_rows_, device_;
  int row_streams_ = 1;
  celv2hip_engine* h_ = nullptr;
  std::unique_ptr<CollectPool> pool_;  // created by the first collect
  std::string name_;
  bool full_depth_ = false;
  int64_t lambda_ = -1;
};

// One exported superblock record as the encoder takes it (the role of sb_out_to_decision,
// src/gpu/decision_map.cpp:65-107): tree, leaves and the 425 candidates; MVs are already 1/8 pel.
void sb_record_to_decision(const celv2_sb& rec, SbDecision& d);
// The inverse, for tests that compare maps as flat records.
void decision_to_sb_record(const SbDecision& d, celv2_sb& rec);

}  // namespace celv2
