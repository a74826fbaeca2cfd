// RowGate for the build outside the CelV2 tree (in-tree the reference's src/gpu/decision_map.cpp
// provides it): rows become readable one at a time or as a prefix; a failure is recorded once and
// rethrown to every waiter. Semantics of src/gpu/decision_map.h:139-160.
#include "celv2/engine_api.h"

#include <algorithm>

void RowGate::publish(int row) {
  {
    std::lock_guard<std::mutex> lk(m_);
    if (row + 1 > ready_) {
      if (ready_ == 0) first_ready_ = std::chrono::steady_clock::now();
      ready_ = row + 1;
    }
  }
  cv_.notify_all();
}

void RowGate::publish_one(int row) {
  {
    std::lock_guard<std::mutex> lk(m_);
    if (one_.size() <= static_cast<size_t>(row)) one_.resize(static_cast<size_t>(row) + 1, 0);
    one_[row] = 1;
    const int was = ready_;
    while (static_cast<size_t>(ready_) < one_.size() && one_[ready_]) ready_++;
    if (was == 0 && ready_ > 0) first_ready_ = std::chrono::steady_clock::now();
  }
  cv_.notify_all();
}

void RowGate::wait_row(int row) const {
  std::unique_lock<std::mutex> lk(m_);
  cv_.wait(lk, [&] { return error_ || ready_ > row || (static_cast<size_t>(row) < one_.size() && one_[row]); });
  if (error_) std::rethrow_exception(error_);
}

void RowGate::fail(std::exception_ptr e) {
  {
    std::lock_guard<std::mutex> lk(m_);
    if (!error_) error_ = e;
  }
  cv_.notify_all();
}

int RowGate::ready_rows() const {
  std::lock_guard<std::mutex> lk(m_);
  return ready_;
}

std::chrono::steady_clock::time_point RowGate::first_ready() const {
  std::lock_guard<std::mutex> lk(m_);
  return first_ready_;
}
