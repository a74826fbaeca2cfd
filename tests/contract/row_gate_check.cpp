// RowGate of the standalone build (celv2_amd/csrc/host/row_gate.cpp) against the semantics CelV2's
// RowGate documents (src/gpu/decision_map.h:83-108): publish(row) readies rows 0..row (smaller values
// change nothing), publish_one(row) one row in any order, ready_rows() the leading run of ready rows,
// wait_row blocks until its row is ready and rethrows a recorded failure, fail keeps the first one.
// Exit status 0 when every check holds.
#include <atomic>
#include <chrono>
#include <cstdio>
#include <stdexcept>
#include <thread>
#include <vector>

#include "celv2/engine_api.h"

static int failures = 0;
#define CHECK(c)                                                   \
  do {                                                             \
    if (!(c)) std::printf("FAIL line %d: %s\n", __LINE__, #c), failures++; \
  } while (0)

int main() {
  {
    RowGate g;
    CHECK(g.ready_rows() == 0);
    CHECK(g.first_ready() == std::chrono::steady_clock::time_point{});
    g.publish_one(2);
    CHECK(g.ready_rows() == 0);
    g.publish_one(0);
    CHECK(g.ready_rows() == 1);
    CHECK(g.first_ready() != std::chrono::steady_clock::time_point{});
    g.publish_one(1);
    CHECK(g.ready_rows() == 3);
    g.publish(1);  // smaller than what is ready: nothing changes
    CHECK(g.ready_rows() == 3);
    g.publish(9);
    CHECK(g.ready_rows() == 10);
    g.wait_row(7);  // ready: returns at once
  }
  {
    // Long runs across 64-row words, published out of order.
    RowGate g;
    std::vector<int> order;
    for (int r = 199; r >= 0; r -= 2) order.push_back(r);
    for (int r = 0; r < 200; r += 2) order.push_back(r);
    int expect_run = 0;
    std::vector<char> ready(200, 0);
    for (int r : order) {
      g.publish_one(r);
      ready[r] = 1;
      while (expect_run < 200 && ready[expect_run]) expect_run++;
      CHECK(g.ready_rows() == expect_run);
    }
    CHECK(g.ready_rows() == 200);
    RowGate h;
    h.publish(63);
    CHECK(h.ready_rows() == 64);
    h.publish_one(65);
    CHECK(h.ready_rows() == 64);
    h.publish(64);
    CHECK(h.ready_rows() == 66);
    h.publish(127);
    h.publish_one(129);
    h.publish(128);
    CHECK(h.ready_rows() == 130);
  }
  {
    // A waiter wakes when its row is published by another thread, in any order.
    RowGate g;
    std::atomic<int> woke{0};
    std::thread t([&] {
      g.wait_row(5);
      woke = 1;
    });
    std::this_thread::sleep_for(std::chrono::milliseconds(20));
    CHECK(woke == 0);
    g.publish_one(4);
    std::this_thread::sleep_for(std::chrono::milliseconds(20));
    CHECK(woke == 0);
    g.publish_one(5);
    t.join();
    CHECK(woke == 1);
    CHECK(g.ready_rows() == 0);  // rows 0..3 never published
  }
  {
    // A failure wakes every waiter with the first error, also for rows already ready.
    RowGate g;
    std::atomic<int> errors{0};
    std::vector<std::thread> ts;
    for (int i = 0; i < 4; i++)
      ts.emplace_back([&, i] {
        try {
          g.wait_row(10 + i);
        } catch (const std::runtime_error& e) {
          if (std::string(e.what()) == "first") errors++;
        }
      });
    std::this_thread::sleep_for(std::chrono::milliseconds(20));
    g.fail(std::make_exception_ptr(std::runtime_error("first")));
    g.fail(std::make_exception_ptr(std::runtime_error("second")));
    for (auto& t : ts) t.join();
    CHECK(errors == 4);
    g.publish_one(0);
    bool threw = false;
    try {
      g.wait_row(0);
    } catch (const std::runtime_error& e) {
      threw = std::string(e.what()) == "first";
    }
    CHECK(threw);
  }
  std::printf(failures ? "row gate: %d failures\n" : "row gate: ok\n", failures);
  return failures ? 1 : 0;
}
