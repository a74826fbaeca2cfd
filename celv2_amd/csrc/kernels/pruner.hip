// The learned pruner's per-superblock inputs. From the source's level 2 (a superblock is 16x16 there)
// and each reference's level-2 search (the superblock's 2x2 blocks' best SAD + distance penalty):
//   f[r]     log2(1 + the 4 blocks' summed cost) / 16 for reference r (0 when r >= nrefs)
//   f[7]     the 16x16 level-2 source mean / 255
//   f[8]     log2(1 + the source variance) / 16
//   f[9]     log2(1 + 16 x the mean absolute horizontal + vertical difference) / 16
// Positions beyond the frame read the clamped plane, as every kernel does.
#include "common.h"
#include "launch.h"

namespace celv2hip {

namespace {

__global__ void __launch_bounds__(256) k_features(FeatureArgs a) {
  const int sb = blockIdx.x * 256 + threadIdx.x;
  if (sb >= a.sb_cols * a.sb_rows) return;
  const int sx = sb % a.sb_cols, sy = sb / a.sb_cols;
  float* f = a.feats + static_cast<size_t>(sb) * kFeatures;
  for (int r = 0; r < kMaxRefs; r++) {
    float v = 0.f;
    if (r < a.nrefs) {
      uint32_t c = 0;
      for (int q = 0; q < 4; q++) {
        const int bx = min(2 * sx + (q & 1), a.cols2 - 1), by = min(2 * sy + (q >> 1), a.rows2 - 1);
        c += a.cost2[r][by * a.cols2 + bx];
      }
      v = __log2f(1.f + static_cast<float>(c)) / 16.f;
    }
    f[r] = v;
  }
  const DPlane& p = a.src2;
  float sum = 0.f, sq = 0.f, grad = 0.f;
  for (int y = 0; y < 16; y++)
    for (int x = 0; x < 16; x++) {
      const int py = min(16 * sy + y, p.h - 1), px = min(16 * sx + x, p.w - 1);
      const float v = p.p[static_cast<ptrdiff_t>(py) * p.stride + px];
      sum += v, sq += v * v;
      const float right = p.p[static_cast<ptrdiff_t>(py) * p.stride + min(px + 1, p.w - 1)];
      const float down = p.p[static_cast<ptrdiff_t>(min(py + 1, p.h - 1)) * p.stride + px];
      grad += fabsf(right - v) + fabsf(down - v);
    }
  const float mean = sum / 256.f, var = fmaxf(0.f, sq / 256.f - mean * mean);
  f[7] = mean / 255.f;
  f[8] = __log2f(1.f + var) / 16.f;
  f[9] = __log2f(1.f + 16.f * grad / 512.f) / 16.f;
}

}  // namespace

void launch_features(const FeatureArgs& a, hipStream_t s) {
  const int n = a.sb_cols * a.sb_rows;
  if (n > 0) k_features<<<(n + 255) / 256, 256, 0, s>>>(a);
}

}  // namespace celv2hip
