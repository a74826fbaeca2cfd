// Issue-rate microbenchmark of v_sad_u8 on the current GPU: the ceiling of the motion search's
// pixel-SAD work. Eight independent accumulation chains per lane, a fixed number of iterations, a
// grid of 8 waves per SIMD of a 256-CU device; prints wave-instructions/s and pixel-SADs/s.
//
//   celv2_amd/lib/sad_rate
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>

__global__ void k_sad(const uint32_t* in, uint32_t* out, int iters) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t x0 = in[i & 1023], x1 = in[(i + 1) & 1023], x2 = in[(i + 2) & 1023], x3 = in[(i + 3) & 1023];
  uint32_t s0 = 0, s1 = 0, s2 = 0, s3 = 0, s4 = 0, s5 = 0, s6 = 0, s7 = 0;
  for (int it = 0; it < iters; it++) {
    s0 = __builtin_amdgcn_sad_u8(x0, x2, s0), s1 = __builtin_amdgcn_sad_u8(x0, x3, s1);
    s2 = __builtin_amdgcn_sad_u8(x1, x2, s2), s3 = __builtin_amdgcn_sad_u8(x1, x3, s3);
    s4 = __builtin_amdgcn_sad_u8(x0 ^ 5u, x2, s4), s5 = __builtin_amdgcn_sad_u8(x0 ^ 5u, x3, s5);
    s6 = __builtin_amdgcn_sad_u8(x1 ^ 5u, x2, s6), s7 = __builtin_amdgcn_sad_u8(x1 ^ 5u, x3, s7);
  }
  out[i] = s0 ^ s1 ^ s2 ^ s3 ^ s4 ^ s5 ^ s6 ^ s7;
}

int main() {
  const int threads = 256 * 4 * 64 * 8, iters = 2000;
  static uint32_t host[1024];
  for (int i = 0; i < 1024; i++) host[i] = 0x9e3779b9u * static_cast<uint32_t>(i + 1);
  uint32_t *din = nullptr, *dout = nullptr;
  if (hipMalloc(&din, sizeof(host)) != hipSuccess || hipMalloc(&dout, sizeof(uint32_t) * threads) != hipSuccess) return 1;
  if (hipMemcpy(din, host, sizeof(host), hipMemcpyHostToDevice) != hipSuccess) return 1;
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) return 1;
  double best = 0;
  for (int rep = 0; rep < 3; rep++) {
    if (hipEventRecord(e0) != hipSuccess) return 2;
    k_sad<<<threads / 256, 256>>>(din, dout, iters);
    if (hipEventRecord(e1) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) return 2;
    float ms = 0;
    if (hipEventElapsedTime(&ms, e0, e1) != hipSuccess) return 2;
    const double inst = static_cast<double>(threads) / 64 * iters * 8;
    const double px = inst * 64 * 4;
    std::printf("v_sad_u8: %.3f ms, %.1f G wave-instructions/s, %.1f T pixel-SADs/s\n", ms, inst / ms / 1e6, px / ms / 1e9);
    if (px / ms / 1e9 > best) best = px / ms / 1e9;
  }
  std::printf("best: %.1f T pixel-SADs/s\n", best);
  return 0;
}
