// The learned pruner's per-superblock inputs. From the source's level 2 (a superblock is 16x16 there)
// and each reference's level-2 search: c[q][r] = the best SAD + distance penalty of the superblock's
// block q (2x2 blocks of 8x8) against reference r. E_k = sum over q of min over r < k of c[q][r] is
// the coarse cost with the nearest k references (E_k = E_nrefs for k > nrefs).
//   f[0]     log2(1 + E_nrefs) / 16
//   f[k]     min(4, 8 (log2(1 + E_k) - log2(1 + E_k+1))), k = 1..6: what reference k + 1 gains
//   f[7]     the 16x16 level-2 source mean / 255
//   f[8]     log2(1 + the source variance) / 16
//   f[9]     log2(1 + 16 x the mean absolute horizontal + vertical difference) / 16
//   f[10]    log2(1 + lambda_q8) / 32 (the rate multiplier: how much a reference's extra bits weigh)
//   f[11..13] the frame's means of f[0], f[8] and f[1] + ... + f[6] (the content's kind: flat-shaded
//             or natural, and how much further references gain across the frame)
// The gains are differences taken before rounding, so the network sees them at full relative
// precision in BF16. Positions beyond the frame read the clamped plane, as every kernel does.
#include "common.h"
#include "launch.h"

namespace celv2hip {

namespace {

// One wave per superblock. Lanes 0..3 hold block q's costs (q = lane) and take their running minimum
// over the references; lane = 4 y + x / 4 takes 4 source samples of row y. Every sum is of integers
// below 2^24, exact in any order.
__global__ void __launch_bounds__(256) k_features(FeatureArgs a) {
  const int lane = threadIdx.x & 63;
  const int sb = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (sb >= a.sb_cols * a.sb_rows) return;  // whole wave
  const int sx = sb % a.sb_cols, sy = sb / a.sb_cols;
  uint32_t cost[kMaxRefs];
  {
    const int q = lane & 3;
    const int bx = min(2 * sx + (q & 1), a.cols2 - 1), by = min(2 * sy + (q >> 1), a.rows2 - 1);
#pragma unroll
    for (int r = 0; r < kMaxRefs; r++) cost[r] = lane < 4 && r < a.nrefs ? a.cost2[r][by * a.cols2 + bx] : 0u;
  }
  const DPlane& p = a.src2;
  const int py = min(16 * sy + (lane >> 2), p.h - 1), pd = min(py + 1, p.h - 1);
  float sum = 0.f, sq = 0.f, grad = 0.f;
#pragma unroll
  for (int i = 0; i < 4; i++) {
    const int px = min(16 * sx + 4 * (lane & 3) + i, p.w - 1);
    const float v = p.p[static_cast<ptrdiff_t>(py) * p.stride + px];
    const float right = p.p[static_cast<ptrdiff_t>(py) * p.stride + min(px + 1, p.w - 1)];
    const float down = p.p[static_cast<ptrdiff_t>(pd) * p.stride + px];
    sum += v, sq += v * v;
    grad += fabsf(right - v) + fabsf(down - v);
  }
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) {
    sum += __shfl_xor(sum, m, 64);
    sq += __shfl_xor(sq, m, 64);
    grad += __shfl_xor(grad, m, 64);
  }
  float f[kFeatures] = {};
  uint32_t envelope = 0;
  float prev = 0.f;
#pragma unroll
  for (int r = 0; r < kMaxRefs; r++) {
    if (r >= a.nrefs) break;
    envelope = r == 0 ? cost[r] : min(envelope, cost[r]);
    uint32_t e = envelope;  // lanes 4.. hold 0
    e += static_cast<uint32_t>(__shfl_xor(static_cast<int>(e), 1, 64));
    e += static_cast<uint32_t>(__shfl_xor(static_cast<int>(e), 2, 64));
    const float l = __log2f(1.f + static_cast<float>(e));
    if (r > 0) f[r] = fminf(4.f, 8.f * (prev - l));
    prev = l;
  }
  f[0] = a.nrefs > 0 ? prev / 16.f : 0.f;
  const float mean = sum / 256.f, var = fmaxf(0.f, sq / 256.f - mean * mean);
  f[7] = mean / 255.f;
  f[8] = __log2f(1.f + var) / 16.f;
  f[9] = __log2f(1.f + 16.f * grad / 512.f) / 16.f;
  f[10] = __log2f(1.f + static_cast<float>(a.lambda_q8)) / 32.f;
  if (lane == 0) {
    float* out = a.feats + static_cast<size_t>(sb) * kFeatures;
#pragma unroll
    for (int k = 0; k < kFeatures; k++) out[k] = f[k];
  }
}

// The frame context f[11..13]: one workgroup sums the three per-superblock quantities in a fixed
// order (a strided pass, then a tree), so the means are the same on every run.
__global__ void __launch_bounds__(256) k_frame_context(float* feats, int n) {
  __shared__ float part[3][256];
  const int tid = static_cast<int>(threadIdx.x);
  float s0 = 0.f, s1 = 0.f, s2 = 0.f;
  for (int sb = tid; sb < n; sb += 256) {
    const float* f = feats + static_cast<size_t>(sb) * kFeatures;
    s0 += f[0], s1 += f[8], s2 += f[1] + f[2] + f[3] + f[4] + f[5] + f[6];
  }
  part[0][tid] = s0, part[1][tid] = s1, part[2][tid] = s2;
  __syncthreads();
  for (int w = 128; w >= 1; w >>= 1) {
    if (tid < w)
      for (int c = 0; c < 3; c++) part[c][tid] += part[c][tid + w];
    __syncthreads();
  }
  const float inv = 1.f / static_cast<float>(n);
  for (int sb = tid; sb < n; sb += 256) {
    float* f = feats + static_cast<size_t>(sb) * kFeatures;
    f[11] = part[0][0] * inv, f[12] = part[1][0] * inv, f[13] = part[2][0] * inv;
  }
}

}  // namespace

void launch_features(const FeatureArgs& a, hipStream_t s, bool context) {
  const int n = a.sb_cols * a.sb_rows;
  if (n > 0) {
    k_features<<<(n + 3) / 4, 256, 0, s>>>(a);
    if (context) k_frame_context<<<1, 256, 0, s>>>(a.feats, n);
  }
}

}  // namespace celv2hip
