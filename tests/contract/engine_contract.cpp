// The engine contract of CelV2's DecisionEngine (src/gpu/decision_engine.h:9-51) exercised on
// celv2::HipEngine, the way CelV2's own GPU tests exercise its Vulkan engine; each case compares the
// decisions with the oracle's CPU engine (oracle/, test infrastructure):
//   concurrent_creation          engines created on 8 threads at once decide as the oracle does
//   launch_collect_matches_decide tests/test_gpu_intra.cpp:95-114: launch copies its frames (the
//                                caller overwrites them before collect), key and inter frames
//   prepare_copies_its_frame     decision_engine.h:19-21: next may change once prepare returns
//   rows_publish_before_collect_ends tests/test_gpu_intra.cpp:194-231 (default and 3 row streams)
//   rejects_bad_sizes            tests/test_gpu_intra.cpp:233-242, plus sizes the engine cannot hold
//   exception_mapping            vk_engine.cpp:176-183: size / too many references -> runtime_error,
//                                key count -> invalid_argument; collect without a launch fails the
//                                gate and rethrows; the engine still decides afterwards
//   three_engines_match_cpu      tests/test_gpu_me.cpp:195-257: 3 engines on 3 threads, 30
//                                iterations each, 7 references, sources alternating
// Prints "CASE <name> ok" or "CASE <name> FAIL <detail>" per case; exit status 0 when all pass.
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <exception>
#include <functional>
#include <mutex>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "celv2/hip_engine.h"
#include "celv2/lambda.h"
#include "celv2hip.h"
#include "celv2synth.h"

extern "C" {
const char* orc_last_error(void);
int orc_decide_frame(int w, int h, const uint8_t* src, const uint8_t* const* refs, int nrefs, int64_t lam, int sb_size,
                     celv2_sb* out);
void* orc_engine_create(int w, int h, int sb_size, int threads, int fail_row);
void orc_engine_destroy(void* p);
int orc_engine_launch(void* p, int w, int h, const uint8_t* src, const uint8_t* const* refs, int nrefs, int64_t lam);
int orc_engine_collect(void* p, celv2_sb* out, int* rows_ready);
}

namespace {

using Records = std::vector<celv2_sb>;

struct Fail : std::runtime_error {
  using std::runtime_error::runtime_error;
};

void expect(bool ok, const std::string& what) {
  if (!ok) throw Fail(what);
}

// Frame t of a synthetic clip (include/celv2synth.h), the 8-bit luma the engines read.
Frame420 synth(int workload, uint64_t seed, int w, int h, int t) {
  celv2_synth_params p;
  if (celv2_synth_config(workload, seed, w, h, &p) != 0) throw std::runtime_error("synth config");
  Frame420 f;
  f.width = w, f.height = h, f.pw = (w + 7) & ~7, f.ph = (h + 7) & ~7;
  f.y.resize(static_cast<size_t>(f.pw) * f.ph);
  if (celv2_synth_frame(&p, t, nullptr, nullptr, nullptr, f.y.data(), 4) != 0) throw std::runtime_error("synth frame");
  return f;
}

Records records(const DecisionMap& dm) {
  Records r(dm.sbs.size());
  for (size_t i = 0; i < dm.sbs.size(); i++) celv2::decision_to_sb_record(dm.sbs[i], r[i]);
  return r;
}

// The oracle's single-threaded decision of the frame.
Records oracle(const Frame420& src, const std::vector<const Frame420*>& refs, int64_t lam) {
  std::vector<const uint8_t*> rp;
  for (const Frame420* f : refs) rp.push_back(f->y.data());
  Records out(static_cast<size_t>((src.width + 63) / 64) * ((src.height + 63) / 64));
  if (orc_decide_frame(src.width, src.height, src.y.data(), rp.data(), static_cast<int>(rp.size()), lam, 64, out.data()) != 0)
    throw std::runtime_error(std::string("oracle: ") + orc_last_error());
  return out;
}

// The oracle's threaded CPU engine (CpuEngine), for the larger frames.
Records oracle_threaded(const Frame420& src, const std::vector<const Frame420*>& refs, int64_t lam, int threads) {
  void* e = orc_engine_create(src.width, src.height, 64, threads, -1);
  if (!e) throw std::runtime_error(std::string("oracle engine: ") + orc_last_error());
  std::vector<const uint8_t*> rp;
  for (const Frame420* f : refs) rp.push_back(f->y.data());
  Records out(static_cast<size_t>((src.width + 63) / 64) * ((src.height + 63) / 64));
  int ready = 0;
  const bool ok = orc_engine_launch(e, src.width, src.height, src.y.data(), rp.data(), static_cast<int>(rp.size()), lam) == 0 &&
                  orc_engine_collect(e, out.data(), &ready) == 0;
  orc_engine_destroy(e);
  if (!ok) throw std::runtime_error(std::string("oracle engine: ") + orc_last_error());
  return out;
}

// Superblocks whose records differ (0: the maps are identical).
int differing(const Records& a, const Records& b) {
  if (a.size() != b.size()) return -1;
  int n = 0;
  for (size_t i = 0; i < a.size(); i++) n += std::memcmp(&a[i], &b[i], sizeof(celv2_sb)) != 0;
  return n;
}

void same(const Records& got, const Records& want, const std::string& what) {
  const int d = differing(got, want);
  expect(d == 0, what + ": " + std::to_string(d) + " superblocks differ from the oracle");
}

DecisionMap decide(celv2::HipEngine& e, const Frame420& src, const std::vector<const Frame420*>& refs, int qindex,
                   const std::vector<long long>& keys = {}) {
  DecisionMap dm;
  RowGate gate;
  e.launch(src, refs, qindex, dm, keys);
  e.collect(dm, gate);
  expect(gate.ready_rows() == dm.sb_rows, "collect left rows unpublished");
  return dm;
}

// ---- cases ---------------------------------------------------------------------------------------

void concurrent_creation() {
  const int w = 352, h = 288, n = 8;
  const Frame420 a = synth(0, 31, w, h, 1), b = synth(0, 31, w, h, 0);
  const Records want = oracle(a, {&b}, lambda_q8(120));
  std::mutex mu;
  std::condition_variable cv;
  int arrived = 0;
  std::vector<std::string> errors(n);
  std::vector<std::thread> ts;
  for (int t = 0; t < n; t++)
    ts.emplace_back([&, t] {
      try {
        {
          std::unique_lock<std::mutex> lk(mu);
          if (++arrived == n) cv.notify_all();
          cv.wait(lk, [&] { return arrived == n; });
        }
        celv2::HipEngine e(w, h, 64);  // every thread creates its engine at once
        const int d = differing(records(decide(e, a, {&b}, 120)), want);
        if (d) errors[t] = std::to_string(d) + " superblocks differ";
      } catch (const std::exception& ex) {
        errors[t] = ex.what();
      }
    });
  for (auto& t : ts) t.join();
  for (int t = 0; t < n; t++) expect(errors[t].empty(), "thread " + std::to_string(t) + ": " + errors[t]);
}

void launch_collect_matches_decide() {
  for (auto [w, h] : {std::pair<int, int>{202, 122}, std::pair<int, int>{640, 360}}) {
    const Frame420 src = synth(0, 7, w, h, 1), prev = synth(0, 7, w, h, 0);
    celv2::HipEngine a(w, h, 64), b(w, h, 64);
    for (bool key : {false, true}) {
      std::vector<const Frame420*> refs;
      if (!key) refs.push_back(&prev);
      const Records want = records(decide(a, src, refs, 120));
      same(want, oracle(src, refs, lambda_q8(120)), "engine a");
      // launch must not read the caller's frames after it returns.
      Frame420 s = src, p = prev;
      DecisionMap dm;
      b.launch(s, key ? nullptr : &p, 120, dm);
      std::fill(s.y.begin(), s.y.end(), uint8_t{0});
      std::fill(p.y.begin(), p.y.end(), uint8_t{255});
      RowGate gate;
      b.collect(dm, gate);
      same(records(dm), want, std::string(key ? "key" : "inter") + " frame overwritten after launch");
      expect(gate.ready_rows() == dm.sb_rows, "rows left unpublished");
    }
  }
}

void prepare_copies_its_frame() {
  const int w = 352, h = 288;
  const Frame420 f0 = synth(1, 9, w, h, 0), f1 = synth(1, 9, w, h, 1), f2 = synth(1, 9, w, h, 2);
  celv2::HipEngine e(w, h, 64);
  DecisionMap dm;
  RowGate g1;
  e.launch(f1, {&f0}, 110, dm, {0}, 1);
  Frame420 next = f2;
  e.prepare(next, 2);
  std::fill(next.y.begin(), next.y.end(), uint8_t{77});
  e.collect(dm, g1);
  same(records(dm), oracle(f1, {&f0}, lambda_q8(110)), "frame 1");
  RowGate g2;
  e.launch(next, {&f1, &f0}, 110, dm, {1, 0}, 2);  // the prepared copy under key 2 is used
  e.collect(dm, g2);
  same(records(dm), oracle(f2, {&f1, &f0}, lambda_q8(110)), "prepared frame 2 (caller's copy overwritten)");
}

void rows_publish_before_collect_ends() {
  const int w = 640, h = 360;
  const Frame420 src = synth(0, 3, w, h, 1), prev = synth(0, 3, w, h, 0);
  for (int streams : {1, 3}) {
    celv2::HipEngine gpu(w, h, 64);
    gpu.set_row_streams(streams);
    DecisionMap dm;
    gpu.launch(src, &prev, 120, dm);
    RowGate gate;
    std::atomic<bool> finished{false};
    // A waiter that never wakes is released by failing the gate (no abort, no hang of this thread).
    std::thread watchdog([&] {
      const auto end = std::chrono::steady_clock::now() + std::chrono::seconds(20);
      while (!finished && std::chrono::steady_clock::now() < end) std::this_thread::sleep_for(std::chrono::milliseconds(1));
      if (!finished) gate.fail(std::make_exception_ptr(std::runtime_error("no row within 20 s")));
    });
    std::exception_ptr err;
    std::thread collector([&] {
      try {
        gpu.collect(dm, gate);
      } catch (...) {
        err = std::current_exception();
      }
    });
    int first = -1;
    std::string wait_error;
    try {
      gate.wait_row(0);
      first = gate.ready_rows();
    } catch (const std::exception& e) {
      wait_error = e.what();
    }
    collector.join();
    finished = true;
    watchdog.join();
    if (err) std::rethrow_exception(err);
    expect(wait_error.empty(), "wait_row: " + wait_error);
    expect(first >= 1, "row 0 published but ready_rows() " + std::to_string(first));
    expect(gate.ready_rows() == dm.sb_rows, "not every row published");
    same(records(dm), oracle(src, {&prev}, lambda_q8(120)), "rows with " + std::to_string(streams) + " row streams");
  }
}

template <typename E, typename F>
std::string throws(F&& f) {
  try {
    f();
  } catch (const E& e) {
    return e.what()[0] ? "" : "empty message";
  } catch (const std::exception& e) {
    return std::string("wrong exception type: ") + e.what();
  }
  return "no exception";
}

void rejects_bad_sizes() {
  std::string r;
  r = throws<std::runtime_error>([] { celv2::HipEngine g(0, 0, 64); });
  expect(r.empty(), "0x0: " + r);
  r = throws<std::runtime_error>([] { celv2::HipEngine g(-5, 64, 64); });
  expect(r.empty(), "-5x64: " + r);
  r = throws<std::runtime_error>([] { celv2::HipEngine g(352, 288, 128); });
  expect(r.empty(), "128 superblocks: " + r);
  // A reference's 16 quarter-pel planes must span under 2 GiB: no memory is allocated for these.
  r = throws<std::runtime_error>([] { celv2::HipEngine g(12000, 12000, 64); });
  expect(r.empty(), "12000x12000: " + r);
  r = throws<std::runtime_error>([] { celv2::HipEngine g(16384, 8192, 64); });
  expect(r.empty(), "16384x8192: " + r);
}

void exception_mapping() {
  const int w = 202, h = 122;
  const Frame420 src = synth(0, 5, w, h, 1), ref = synth(0, 5, w, h, 0), other = synth(0, 5, 160, 96, 0);
  celv2::HipEngine e(w, h, 64);
  DecisionMap dm;
  std::string r;
  r = throws<std::runtime_error>([&] { e.launch(other, std::vector<const Frame420*>{&ref}, 120, dm); });
  expect(r.empty(), "source size: " + r);
  r = throws<std::runtime_error>([&] { e.launch(src, std::vector<const Frame420*>{&other}, 120, dm); });
  expect(r.empty(), "reference size: " + r);
  std::vector<const Frame420*> eight(8, &ref);
  r = throws<std::runtime_error>([&] { e.launch(src, eight, 120, dm); });
  expect(r.empty(), "8 references: " + r);
  r = throws<std::invalid_argument>([&] { e.launch(src, {&ref}, 120, dm, {1, 2}); });
  expect(r.empty(), "key count: " + r);
  {
    // collect with nothing launched: fails the gate, rethrows, leaves no work pending.
    celv2::HipEngine idle(w, h, 64);
    DecisionMap none;
    RowGate gate;
    r = throws<std::runtime_error>([&] { idle.collect(none, gate); });
    expect(r.empty(), "collect without launch: " + r);
    r = throws<std::runtime_error>([&] { gate.wait_row(0); });
    expect(r.empty(), "gate after a failed collect: " + r);
  }
  // A frame of the wrong size given to prepare is ignored (decision_engine.h: nothing to prepare).
  e.prepare(other, 9);
  // The engine still decides after the rejected calls.
  same(records(decide(e, src, {&ref}, 120)), oracle(src, {&ref}, lambda_q8(120)), "after rejected calls");
}

void three_engines_match_cpu() {
  const int w = 640, h = 360, threads = 3, iters = 30;
  std::vector<Frame420> pics;
  for (int i = 0; i < 9; i++) pics.push_back(synth(1, 17, w, h, i));
  std::vector<const Frame420*> refs;
  std::vector<long long> keys;
  for (int i = 1; i < 8; i++) refs.push_back(&pics[i]), keys.push_back(i);
  const int64_t lam = lambda_q8(90);
  const Records want[2] = {oracle_threaded(pics[0], refs, lam, 8), oracle_threaded(pics[8], refs, lam, 8)};
  expect(differing(want[0], want[1]) > 0, "the two sources decide alike");
  std::atomic<int> bad{0}, errors{0};
  std::mutex log;
  std::vector<std::thread> ts;
  for (int t = 0; t < threads; t++)
    ts.emplace_back([&, t] {
      try {
        celv2::HipEngine e(w, h, 64);
        for (int i = 0; i < iters; i++) {
          const int s = (i + t) % 2;
          const int d = differing(records(decide(e, pics[s ? 8 : 0], refs, 90, keys)), want[s]);
          if (d) {
            std::lock_guard<std::mutex> lk(log);
            if (++bad <= 3) std::fprintf(stderr, "thread %d iteration %d: %d superblocks differ\n", t, i, d);
          }
        }
      } catch (const std::exception& ex) {
        std::lock_guard<std::mutex> lk(log);
        std::fprintf(stderr, "thread %d: %s\n", t, ex.what());
        errors++;
      }
    });
  for (auto& t : ts) t.join();
  expect(bad == 0 && errors == 0, std::to_string(bad.load()) + " bad decisions of " + std::to_string(threads * iters) +
                                      ", " + std::to_string(errors.load()) + " threads threw");
}

}  // namespace

int main(int argc, char** argv) {
  const std::vector<std::pair<const char*, std::function<void()>>> cases = {
      {"concurrent_creation", concurrent_creation},  // first: the process's first engines
      {"launch_collect_matches_decide", launch_collect_matches_decide},
      {"prepare_copies_its_frame", prepare_copies_its_frame},
      {"rows_publish_before_collect_ends", rows_publish_before_collect_ends},
      {"rejects_bad_sizes", rejects_bad_sizes},
      {"exception_mapping", exception_mapping},
      {"three_engines_match_cpu", three_engines_match_cpu},
  };
  int failed = 0;
  for (const auto& [name, fn] : cases) {
    if (argc > 1 && std::strcmp(argv[1], name) != 0) continue;
    std::printf("CASE %s start\n", name);
    std::fflush(stdout);
    try {
      fn();
      std::printf("CASE %s ok\n", name);
    } catch (const std::exception& e) {
      failed++;
      std::printf("CASE %s FAIL %s\n", name, e.what());
    }
    std::fflush(stdout);
  }
  return failed ? 1 : 0;
}
