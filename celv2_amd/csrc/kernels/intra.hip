// Intra scores of a superblock: SATD of the 13 intra modes (block_mode_satd,
// src/gpu/intra_common.h:149-163) for each of the 420 candidate blocks of 32x32 and below.
//
// One workgroup per superblock. The source around it (rows and columns -1 .. 95 from its origin, all
// the neighbours any of its blocks reads) is staged in LDS through clamped reads; the mode loop is
// uniform over the workgroup and lanes take 4x4 tiles from a static (block, tile) list, so a wave
// works on one block of one mode; per-(block, mode) sums are LDS integer atomics (order-free).
#include <stdexcept>

#include "common.h"
#include "launch.h"

namespace celv2hip {

namespace {

constexpr int kRegionSide = 97;  // -1 .. 95
constexpr int kRegionStride = 100;
constexpr int kIntraTiles = 2304;  // 4x4 tiles of the blocks of levels 1-3 (768 per level)

__device__ uint32_t g_intra_tiles[kIntraTiles];  // b | tx << 16 | ty << 24 (tile origin in 4-px units)

struct Region {
  const uint8_t* r;  // LDS, sample (y, x) relative to the superblock at r[(y + 1) * stride + x + 1]
  __device__ __forceinline__ int at(int y, int x) const { return r[(y + 1) * kRegionStride + x + 1]; }
};

// intra_common.h:47-56 with the superblock origin (ox, oy): the frame's top/left edges read 128.
__device__ __forceinline__ int edge_above(const Region& s, int ox, int oy, int x, int y, int k) {
  if (oy + y == 0) return 128;
  if (k < 0 && ox + x == 0) return s.at(y - 1, x);
  return s.at(y - 1, x + k);
}
__device__ __forceinline__ int edge_left(const Region& s, int ox, int oy, int x, int y, int k) {
  if (ox + x == 0) return 128;
  if (k < 0 && oy + y == 0) return s.at(y, x - 1);
  return s.at(y + k, x - 1);
}

__device__ __forceinline__ int floor_div64(int v) { return v >= 0 ? v >> 6 : -((-v + 63) >> 6); }

// dir_pred (intra_common.h:61-85) for the angles 45, 67 (z1), 113, 135, 157 (z2), 203 (z3).
__device__ __forceinline__ int directional(const Region& s, int ox, int oy, int x, int y, int w, int h, int angle, int i,
                                           int j) {
  const int last = w + h - 1;
  auto deriv = [](int a) { return a == 23 ? 151 : a == 45 ? 64 : 27; };
  if (angle < 90) {
    const int pos = (i + 1) * deriv(angle), base = (pos >> 6) + j, f = (pos & 63) >> 1;
    const int a0 = edge_above(s, ox, oy, x, y, min(base, last)), a1 = edge_above(s, ox, oy, x, y, min(base + 1, last));
    return (a0 * (32 - f) + a1 * f + 16) >> 5;
  }
  if (angle > 180) {
    const int pos = (j + 1) * deriv(270 - angle), base = (pos >> 6) + i, f = (pos & 63) >> 1;
    const int l0 = edge_left(s, ox, oy, x, y, min(base, last)), l1 = edge_left(s, ox, oy, x, y, min(base + 1, last));
    return (l0 * (32 - f) + l1 * f + 16) >> 5;
  }
  const int px = 64 * j - (i + 1) * deriv(180 - angle), bx = floor_div64(px);
  if (bx >= -1) {
    const int f = (px - 64 * bx) >> 1;
    return (edge_above(s, ox, oy, x, y, bx) * (32 - f) + edge_above(s, ox, oy, x, y, bx + 1) * f + 16) >> 5;
  }
  const int py = 64 * i - (j + 1) * deriv(angle - 90), by = floor_div64(py);
  const int f = (py - 64 * by) >> 1;
  return (edge_left(s, ox, oy, x, y, max(by, -1)) * (32 - f) + edge_left(s, ox, oy, x, y, max(by + 1, -1)) * f + 16) >> 5;
}

// intra_pred (intra_common.h:89-118), modes 1..12; mode 0 (DC) is the block's dc value.
__device__ __forceinline__ int predict(const Region& s, int ox, int oy, int x, int y, int w, int h, int mode, int i,
                                       int j, int dc) {
  const bool up = oy + y > 0, lf = ox + x > 0;
  switch (mode) {
    case 0: return dc;
    case 1: return up ? s.at(y - 1, x + j) : 128;
    case 2: return lf ? s.at(y + i, x - 1) : 128;
    case 3: {
      const int a = up ? s.at(y - 1, x + j) : 128, l = lf ? s.at(y + i, x - 1) : 128;
      const int tr = up ? s.at(y - 1, x + w - 1) : 128, bl = lf ? s.at(y + h - 1, x - 1) : 128;
      const int den = w + h + 2;
      return ((h - i) * a + (i + 1) * bl + (w - j) * l + (j + 1) * tr + den / 2) / den;
    }
    case 4:
      return ((h - 1 - i) * edge_above(s, ox, oy, x, y, j) + (i + 1) * edge_left(s, ox, oy, x, y, h - 1) + h / 2) / h;
    case 5:
      return ((w - 1 - j) * edge_left(s, ox, oy, x, y, i) + (j + 1) * edge_above(s, ox, oy, x, y, w - 1) + w / 2) / w;
    case 6: {
      const int a = edge_above(s, ox, oy, x, y, j), l = edge_left(s, ox, oy, x, y, i), tl = edge_above(s, ox, oy, x, y, -1);
      const int base = a + l - tl, pa = abs(base - a), pl = abs(base - l), ptl = abs(base - tl);
      return (pl <= pa && pl <= ptl) ? l : (pa <= ptl ? a : tl);
    }
    case 7: return directional(s, ox, oy, x, y, w, h, 45, i, j);
    case 8: return directional(s, ox, oy, x, y, w, h, 67, i, j);
    case 9: return directional(s, ox, oy, x, y, w, h, 113, i, j);
    case 10: return directional(s, ox, oy, x, y, w, h, 135, i, j);
    case 11: return directional(s, ox, oy, x, y, w, h, 157, i, j);
    default: return directional(s, ox, oy, x, y, w, h, 203, i, j);
  }
}

__global__ void __launch_bounds__(256) k_intra(DPlane src, int sb_cols, int row0, uint32_t* __restrict__ out) {
  __shared__ uint8_t reg[kRegionSide * kRegionStride];
  __shared__ int dcv[kBlocks];
  __shared__ uint32_t acc[kBlocks * kModes];
  const int tid = threadIdx.x;
  const int sbx = blockIdx.x, sby = row0 + blockIdx.y;
  const int ox = 64 * sbx, oy = 64 * sby;
  for (int k = tid; k < kRegionSide * kRegionSide; k += 256) {
    const int ry = k / kRegionSide - 1, rx = k % kRegionSide - 1;
    const int yy = clampi(oy + ry, 0, src.h - 1), xx = clampi(ox + rx, 0, src.w - 1);
    reg[(ry + 1) * kRegionStride + rx + 1] = src.p[static_cast<ptrdiff_t>(yy) * src.stride + xx];
  }
  for (int k = tid; k < kBlocks * kModes; k += 256) acc[k] = 0;
  __syncthreads();
  const Region s{reg};
  // dc_value (intra_common.h:120-128) of every block of levels 1-3.
  for (int b = 5 + tid; b < kBlocks; b += 256) {
    int x, y, w, h;
    block_geom(b, x, y, w, h);
    int sum = 0, n = 0;
    if (oy + y > 0)
      for (int j = 0; j < w; j++) sum += s.at(y - 1, x + j), n++;
    if (ox + x > 0)
      for (int i = 0; i < h; i++) sum += s.at(y + i, x - 1), n++;
    dcv[b] = n ? (sum + n / 2) / n : 128;
  }
  __syncthreads();
  for (int m = 0; m < kModes; m++) {
    for (int slot = tid; slot < kIntraTiles; slot += 256) {
      const uint32_t e = g_intra_tiles[slot];
      const int b = static_cast<int>(e & 0xffff), tx = static_cast<int>((e >> 16) & 255), ty = static_cast<int>(e >> 24);
      int x, y, w, h;
      block_geom(b, x, y, w, h);
      int d[16];
#pragma unroll
      for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) {
          const int pi = 4 * ty + i, pj = 4 * tx + j;
          d[4 * i + j] = s.at(y + pi, x + pj) - predict(s, ox, oy, x, y, w, h, m, pi, pj, dcv[b]);
        }
      atomicAdd(&acc[b * kModes + m], static_cast<uint32_t>(wht4_abs_sum(d)));
    }
  }
  __syncthreads();
  uint32_t* o = out + (static_cast<size_t>(sby) * sb_cols + sbx) * kBlocks * kModes;
  for (int k = 5 * kModes + tid; k < kBlocks * kModes; k += 256) o[k] = acc[k];
}

}  // namespace

void launch_intra(const DPlane& src, int sb_cols, int row0, int nrows, uint32_t* out, hipStream_t s) {
  if (nrows <= 0) return;
  k_intra<<<dim3(sb_cols, nrows), 256, 0, s>>>(src, sb_cols, row0, out);
}

// The static (block, tile) list of levels 1-3 for k_intra.
void upload_intra_tiles(hipStream_t s) {
  static uint32_t table[kIntraTiles];
  int n = 0;
  for (int b = 5; b < kBlocks; b++) {
    const int sq = b / 5, kind = b % 5;
    const int level = sq >= 21 ? 3 : sq >= 5 ? 2 : 1;
    const int size = 64 >> level;
    const int w = (kind == 3 || kind == 4) ? size / 2 : size, h = (kind == 1 || kind == 2) ? size / 2 : size;
    for (int ty = 0; ty < h / 4; ty++)
      for (int tx = 0; tx < w / 4; tx++) table[n++] = static_cast<uint32_t>(b) | (static_cast<uint32_t>(tx) << 16) | (static_cast<uint32_t>(ty) << 24);
  }
  if (n != kIntraTiles) throw std::logic_error("intra tile table size");
  (void)hipMemcpyToSymbolAsync(HIP_SYMBOL(g_intra_tiles), table, sizeof(table), 0, hipMemcpyHostToDevice, s);
}

}  // namespace celv2hip
