// RowGate for builds outside the CelV2 tree (in-tree builds link CelV2's own, whose interface and
// meaning this follows: src/gpu/decision_map.h:83-108). Ready rows are bits of a bitmap; the length
// of its leading run of set bits is ready_rows(). publish(row) sets bits 0..row, publish_one(row) one
// bit; a waiter sleeps until its row's bit is set or a failure is recorded.
#include "celv2/engine_api.h"

namespace {

bool test_bit(const std::vector<uint64_t>& bits, int row) {
  const size_t w = static_cast<size_t>(row) >> 6;
  return w < bits.size() && ((bits[w] >> (row & 63)) & 1u);
}

}  // namespace

void RowGate::mark_locked(int first, int last) {
  if (first < 0) first = 0;
  if (last < first) return;
  const size_t words = (static_cast<size_t>(last) >> 6) + 1;
  if (bits_.size() < words) bits_.resize(words, 0);
  for (int r = first; r <= last;) {
    const int lo = r & 63, hi = (last >> 6) == (r >> 6) ? (last & 63) : 63;
    const uint64_t mask = (hi == 63 ? ~uint64_t{0} : ((uint64_t{1} << (hi + 1)) - 1)) & ~((uint64_t{1} << lo) - 1);
    bits_[static_cast<size_t>(r) >> 6] |= mask;
    r += hi - lo + 1;
  }
  const int before = run_;
  // Extend the leading run a word at a time: count the set bits above run_ in its word.
  for (size_t w = static_cast<size_t>(run_) >> 6; w < bits_.size(); w++) {
    const uint64_t rest = ~(bits_[w] >> (run_ & 63));
    const int ones = rest == 0 ? 64 - (run_ & 63) : __builtin_ctzll(rest);
    const int room = 64 - (run_ & 63);
    run_ += ones < room ? ones : room;
    if (ones < room) break;
  }
  if (before == 0 && run_ > 0) first_ready_ = std::chrono::steady_clock::now();
}

void RowGate::publish(int row) {
  {
    std::lock_guard<std::mutex> lk(mu_);
    mark_locked(run_, row);
  }
  changed_.notify_all();
}

void RowGate::publish_one(int row) {
  {
    std::lock_guard<std::mutex> lk(mu_);
    mark_locked(row, row);
  }
  changed_.notify_all();
}

void RowGate::wait_row(int row) const {
  std::unique_lock<std::mutex> lk(mu_);
  while (!failure_ && !test_bit(bits_, row)) changed_.wait(lk);
  if (failure_) std::rethrow_exception(failure_);
}

void RowGate::fail(std::exception_ptr e) {
  {
    std::lock_guard<std::mutex> lk(mu_);
    if (!failure_) failure_ = std::move(e);
  }
  changed_.notify_all();
}

int RowGate::ready_rows() const {
  std::lock_guard<std::mutex> lk(mu_);
  return run_;
}

std::chrono::steady_clock::time_point RowGate::first_ready() const {
  std::lock_guard<std::mutex> lk(mu_);
  return first_ready_;
}
