// celv2hip_fwd_tx: the C entry point of the forward transforms (include/celv2hip.h). Validates the
// jobs, groups them by size and kind pair into units of one wave each (kernels/transform.hip),
// uploads the job and unit lists through a pinned staging buffer and waits for the kernel. Each
// device has its own staging, list buffer and stream, reused by later calls (calls on one device
// take turns).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>
#include <vector>

#include "../kernels/launch.h"
#include "celv2hip.h"
#include "status.h"

using namespace celv2hip;
using namespace celv2host;

namespace {

constexpr int kMaxDevices = 64;

struct TxDevice {
  std::once_flag tables;
  std::mutex mu;
  hipStream_t stream = nullptr;
  hipEvent_t t0 = nullptr, t1 = nullptr;
  void* host = nullptr;  // pinned staging of the lists
  void* dev = nullptr;
  size_t cap = 0;
};

TxDevice& tx_device(int device) {
  static TxDevice devices[kMaxDevices];
  return devices[device];
}

bool kind_fits(int kind, int lg) {
  if (lg < 2 || lg > 6) return false;
  switch (kind) {
    case CELV2_TX_DCT: return true;
    case CELV2_TX_DST:
    case CELV2_TX_FLIPDST: return lg <= 4;
    case CELV2_TX_IDT: return lg <= 5;
    default: return false;
  }
}

size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

}  // namespace

extern "C" int celv2hip_fwd_tx(int device, const celv2_tx_job* jobs, int njobs, float* kernel_ms) {
  return guarded([&] {
    int count = 0;
    check(hipGetDeviceCount(&count), "hipGetDeviceCount");
    if (device < 0 || device >= count || device >= kMaxDevices) throw Failure(CELV2_ERR_DEVICE, "fwd_tx: no such device");
    if (njobs < 0 || (njobs > 0 && !jobs)) throw Failure(CELV2_ERR_ARG, "fwd_tx: bad job list");
    for (int i = 0; i < njobs; i++) {
      const celv2_tx_job& j = jobs[i];
      if (!j.res || !j.coeff || !kind_fits(j.hkind, j.w_log2) || !kind_fits(j.vkind, j.h_log2) || j.stride < (1 << j.w_log2) ||
          !(j.scale > 0) || !std::isfinite(j.scale))
        throw Failure(CELV2_ERR_ARG, "fwd_tx: job " + std::to_string(i) + " has a size, kind, stride, scale or pointer out of range");
    }
    if (kernel_ms) *kernel_ms = 0.0f;
    if (njobs == 0) return;
    // Jobs grouped by (w, h, vertical kind, horizontal kind), in their order within a group.
    constexpr int kKeys = 8 * 8 * 4 * 4;
    auto key = [](const celv2_tx_job& j) { return ((j.w_log2 * 8 + j.h_log2) * 4 + j.vkind) * 4 + j.hkind; };
    std::vector<int> start(kKeys + 1, 0);
    for (int i = 0; i < njobs; i++) start[key(jobs[i]) + 1]++;
    for (int k = 0; k < kKeys; k++) start[k + 1] += start[k];
    std::vector<int> fill(start.begin(), start.end() - 1);
    std::vector<TxJob> tj(static_cast<size_t>(njobs));
    for (int i = 0; i < njobs; i++) {
      const celv2_tx_job& j = jobs[i];
      const int h = 1 << j.h_log2;
      TxJob& t = tj[static_cast<size_t>(fill[key(j)]++)];
      t.res = j.res, t.coeff = j.coeff, t.dropped = reinterpret_cast<long long*>(j.dropped), t.scale = j.scale;
      t.stride = j.stride;
      const bool dct = j.vkind == CELV2_TX_DCT;
      t.kv0 = dct ? static_cast<float>(j.scale * std::sqrt(1.0 / h)) : static_cast<float>(j.scale);
      t.kv = dct ? static_cast<float>(j.scale * std::sqrt(2.0 / h)) : static_cast<float>(j.scale);
    }
    std::vector<TxUnit> units;
    for (int k = 0; k < kKeys; k++) {
      const int hk = k % 4, vk = (k / 4) % 4, hl = (k / 16) % 8, wl = k / 128;
      const int per = tx_jobs_per_unit(hl);
      for (int f = start[k]; f < start[k + 1]; f += per) units.push_back({f, std::min(per, start[k + 1] - f), wl, hl, vk, hk});
    }
    TxDevice& d = tx_device(device);
    check(hipSetDevice(device), "hipSetDevice");
    std::call_once(d.tables, [] { check(upload_tx_tables(), "transform tables"); });
    std::lock_guard<std::mutex> lock(d.mu);
    if (!d.stream) {
      check(hipStreamCreateWithFlags(&d.stream, hipStreamNonBlocking), "hipStreamCreate");
      check(hipEventCreate(&d.t0), "hipEventCreate");
      check(hipEventCreate(&d.t1), "hipEventCreate");
    }
    const size_t jb = align_up(tj.size() * sizeof(TxJob), 256), ub = units.size() * sizeof(TxUnit);
    if (jb + ub > d.cap) {
      if (d.host) check(hipHostFree(d.host), "hipHostFree"), d.host = nullptr;
      if (d.dev) check(hipFree(d.dev), "hipFree"), d.dev = nullptr;
      d.cap = 0;
      const size_t cap = align_up((jb + ub) * 5 / 4, 1 << 20);
      check(hipHostMalloc(&d.host, cap, hipHostMallocDefault), "hipHostMalloc transform lists");
      check(hipMalloc(&d.dev, cap), "hipMalloc transform lists");
      d.cap = cap;
    }
    std::memcpy(d.host, tj.data(), tj.size() * sizeof(TxJob));
    std::memcpy(static_cast<char*>(d.host) + jb, units.data(), ub);
    check(hipMemcpyAsync(d.dev, d.host, jb + ub, hipMemcpyHostToDevice, d.stream), "hipMemcpyAsync transform lists");
    if (kernel_ms) check(hipEventRecord(d.t0, d.stream), "hipEventRecord");
    launch_fwd_tx(static_cast<const TxJob*>(d.dev), reinterpret_cast<const TxUnit*>(static_cast<char*>(d.dev) + jb),
                  static_cast<int>(units.size()), d.stream);
    check(hipGetLastError(), "fwd_tx kernel");
    if (kernel_ms) check(hipEventRecord(d.t1, d.stream), "hipEventRecord");
    check(hipStreamSynchronize(d.stream), "fwd_tx");
    if (kernel_ms) check(hipEventElapsedTime(kernel_ms, d.t0, d.t1), "hipEventElapsedTime");
  });
}
