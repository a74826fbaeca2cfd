// Plane building: import of a frame's luma into a bordered plane with the 2x2 pyramid levels above
// it (one launch), and the 15 quarter-pel phase planes of a reference.
//
// Every kernel writes the whole bordered plane, computing a border sample from the clamped position
// (so the border equals what the reference's clamped reads return). Threads own 4 horizontally
// adjacent samples of one row and store them as one dword (bordered rows start 64-byte aligned).
#include "batch.h"
#include "common.h"
#include "launch.h"

namespace celv2hip {

namespace {

// 8-bit value of a frame sample: itself, or rounded from bit_depth bits exactly as round_to_8bit
// (src/io/y4m.cpp:28-34): min(255, (v + 2^(d-9)) >> (d-8)).
__device__ __forceinline__ uint32_t to8(uint32_t v, int bit_depth) {
  if (bit_depth <= 8) return v & 255;
  const int s = bit_depth - 8;
  const uint32_t r = (v + (1u << (s - 1))) >> s;
  return r > 255 ? 255 : r;
}

// Workgroups of one plane: 64 x 4-sample groups of 4 bordered rows each.
__device__ __forceinline__ int plane_cols(const DPlane& p) { return ((p.w + 2 * kBorder + 3) / 4 + 63) / 64; }

// This workgroup's item of a batch over planes and its plane-local block (bx, by).
#define CELV2_PLANE_BLOCK(b, plane_of, t, bx, by)                                   \
  const int item_ = batch_item(b, static_cast<int>(blockIdx.x));                 \
  const auto& t = b.item[item_];                                                  \
  const int lw_ = static_cast<int>(blockIdx.x) - b.first[item_], cols_ = plane_cols(plane_of); \
  const int by = lw_ / cols_, bx = lw_ - by * cols_;

// The import and the pyramid in one launch (round_to_8bit, src/io/y4m.cpp:28-34; downsample_px,
// src/gpu/inter_common.h:91-94). A workgroup takes a 64x64 tile of level 0 and the tiles above it
// (32x32 at level 1 ... 4x4 at level 4, aligned the same way): the frame's samples rounded to 8 bits,
// then each level's rounded 2x2 means from the level below, all in LDS. A level-l sample inside the
// level (x < w_l, y < h_l) reads only inside samples of level l - 1 (2x + 1 < w_{l-1}), so a tile
// needs no neighbour. Each level is then written with its border: the tile writes its inside samples,
// and a tile at an edge of the level also the border beyond that edge (and the corner), every border
// sample taking the value at its clamped position, which lies in the tile. Rows are written as
// dwords from 4-aligned columns; a group that runs past the right border lands in the row's padding
// (a row's stride rounds w + 128 up to 64 bytes, which leaves at least the 4 - w mod 4 bytes needed).
constexpr int kTile = 64;

__device__ __forceinline__ void write_level(const DPlane& out, const uint8_t* lds, int level, int x0, int y0, int tid) {
  const int T = kTile >> level, xl = x0 >> level, yl = y0 >> level, W = out.w, H = out.h;
  if (xl >= W || yl >= H) return;  // the tile holds none of this level
  const int xa = xl == 0 ? -kBorder : xl, xb = xl + T >= W ? W + kBorder : xl + T;
  const int ya = yl == 0 ? -kBorder : yl, yb = yl + T >= H ? H + kBorder : yl + T;
  const int groups = (xb - xa + 3) >> 2, n = groups * (yb - ya);
  for (int i = tid; i < n; i += 256) {
    const int r = i / groups, y = ya + r, x = xa + 4 * (i - r * groups);
    const uint8_t* row = lds + (clampi(y, 0, H - 1) - yl) * T;
    uint32_t word = 0;
#pragma unroll
    for (int k = 0; k < 4; k++) word |= static_cast<uint32_t>(row[clampi(x + k, 0, W - 1) - xl]) << (8 * k);
    *reinterpret_cast<uint32_t*>(out.p + static_cast<ptrdiff_t>(y) * out.stride + x) = word;
  }
}

// Level l's inside samples of the tile (T = 64 >> l) from level l - 1's tile in LDS.
__device__ __forceinline__ void down_level(const uint8_t* in, uint8_t* out, int level, int x0, int y0, int W, int H, int tid) {
  const int T = kTile >> level, xl = x0 >> level, yl = y0 >> level;
  for (int i = tid; i < T * T; i += 256) {
    const int r = i / T, c = i - r * T;
    if (yl + r < H && xl + c < W) {
      const uint8_t* a = in + (2 * r) * (2 * T) + 2 * c;
      out[i] = static_cast<uint8_t>((a[0] + a[1] + a[2 * T] + a[2 * T + 1] + 2) >> 2);
    }
  }
}

__global__ void __launch_bounds__(256) k_pyramid(Batch<PyramidItem> b) {
  __shared__ uint8_t l0[kTile * kTile], l1[(kTile / 2) * (kTile / 2)], l2[(kTile / 4) * (kTile / 4)],
      l3[(kTile / 8) * (kTile / 8)], l4[(kTile / 16) * (kTile / 16)];
  const int item = batch_item(b, static_cast<int>(blockIdx.x));
  const PyramidItem& t = b.item[item];
  const int tiles_x = (t.lv[0].w + kTile - 1) / kTile;
  const int tl = static_cast<int>(blockIdx.x) - b.first[item], ty = tl / tiles_x, tx = tl - ty * tiles_x;
  const int x0 = kTile * tx, y0 = kTile * ty, tid = static_cast<int>(threadIdx.x);
  const DPlane& p0 = t.lv[0];
  // Level 0: 16 groups of 4 samples per tile row (level 0's width is a multiple of 8).
  for (int i = tid; i < kTile * (kTile / 4); i += 256) {
    const int r = i >> 4, x = x0 + 4 * (i & 15), y = y0 + r;
    if (y < p0.h && x < p0.w) {
      uint32_t word = 0;
      if (t.wide) {
        const uint16_t* row = static_cast<const uint16_t*>(t.src) + static_cast<size_t>(y) * t.sstride + x;
#pragma unroll
        for (int k = 0; k < 4; k++) word |= to8(row[k], t.bit_depth) << (8 * k);
      } else {
        const uint8_t* row = static_cast<const uint8_t*>(t.src) + static_cast<size_t>(y) * t.sstride + x;
#pragma unroll
        for (int k = 0; k < 4; k++) word |= static_cast<uint32_t>(row[k]) << (8 * k);
      }
      *reinterpret_cast<uint32_t*>(l0 + r * kTile + 4 * (i & 15)) = word;
    }
  }
  __syncthreads();
  write_level(p0, l0, 0, x0, y0, tid);
  down_level(l0, l1, 1, x0, y0, t.lv[1].w, t.lv[1].h, tid);
  __syncthreads();
  write_level(t.lv[1], l1, 1, x0, y0, tid);
  down_level(l1, l2, 2, x0, y0, t.lv[2].w, t.lv[2].h, tid);
  __syncthreads();
  write_level(t.lv[2], l2, 2, x0, y0, tid);
  down_level(l2, l3, 3, x0, y0, t.lv[3].w, t.lv[3].h, tid);
  __syncthreads();
  write_level(t.lv[3], l3, 3, x0, y0, tid);
  down_level(l3, l4, 4, x0, y0, t.lv[4].w, t.lv[4].h, tid);
  __syncthreads();
  write_level(t.lv[4], l4, 4, x0, y0, tid);
}

// AV2 EIGHTTAP_REGULAR at quarter-pel phase q (subpel_tap, src/gpu/inter_common.h:12-19).
constexpr int kTaps[4][8] = {{0, 0, 0, 128, 0, 0, 0, 0},
                                {0, 2, -14, 110, 38, -10, 2, 0},
                                {0, 2, -14, 76, 76, -14, 2, 0},
                                {0, 2, -10, 38, 110, -14, 2, 0}};

// Phases 1..15 (phase k = (qy, qx) = (k / 4, k % 4)) of level 0: horizontal then vertical 8-tap with
// one rounding by 2^14, clamped to 0..255 (subpel_px, src/gpu/inter_common.h:20-33), at each
// sample's clamped position. Phase k's plane starts phase_bytes * k after level 0's.
//
// A group of 4 samples lies wholly left of, inside or right of the frame (x0 is 4-aligned, w a
// multiple of 8). Inside, the 11 source columns x0-3 .. x0+7 of each row come from three aligned dword
// loads (columns x0-4 .. x0+7); in the border all 4 samples are the edge column's value.
// Keeps a value in a register as computed: the byte packing below must stay plain shifts and ORs.
// Left to itself the compiler packs two saturated shifts with v_ashr_pk_u8_i32, and on the MI355X
// boxes measured the packed word then carried a wrong third byte (every 4th sample of a phase plane
// differed from the CPU rule).
__device__ __forceinline__ void opaque(int& v) { asm volatile("" : "+v"(v)); }

template <bool kInside>
__device__ __forceinline__ void subpel_group(const DPlane& in, int cy, int c0, uint32_t out[16]) {
  int acc[16][4];
#pragma unroll
  for (int ph = 0; ph < 16; ph++)
#pragma unroll
    for (int k = 0; k < 4; k++) acc[ph][k] = 0;
#pragma unroll
  for (int i = 0; i < 8; i++) {
    const uint8_t* row = in.p + static_cast<ptrdiff_t>(cy + i - 3) * in.stride + c0 - 4;
    int s[12];
    if (kInside) {
      const uint32_t* q = reinterpret_cast<const uint32_t*>(row);
      const uint32_t d[3] = {q[0], q[1], q[2]};
#pragma unroll
      for (int o = 0; o < 12; o++) s[o] = static_cast<int>((d[o >> 2] >> (8 * (o & 3))) & 255);
    } else {
#pragma unroll
      for (int o = 0; o < 12; o++) s[o] = row[o];
    }
#pragma unroll
    for (int qx = 0; qx < 4; qx++) {
      int h[4];
#pragma unroll
      for (int k = 0; k < 4; k++) {
        int a = 0;
#pragma unroll
        for (int j = 0; j < 8; j++)
          if (kTaps[qx][j]) a += kTaps[qx][j] * s[(kInside ? k : 0) + j + 1];
        h[k] = a;
      }
#pragma unroll
      for (int qy = 0; qy < 4; qy++)
#pragma unroll
        for (int k = 0; k < 4; k++)
          if (kTaps[qy][i]) acc[4 * qy + qx][k] += kTaps[qy][i] * h[k];
    }
  }
#pragma unroll
  for (int ph = 1; ph < 16; ph++) {
    uint32_t word = 0;
#pragma unroll
    for (int k = 0; k < 4; k++) {
      int v = (acc[ph][k] + 8192) >> 14;
      v = v < 0 ? 0 : (v > 255 ? 255 : v);
      opaque(v);
      word |= static_cast<uint32_t>(v) << (8 * k);
    }
    out[ph] = word;
  }
}

__global__ void __launch_bounds__(256) k_subpel(Batch<SubpelItem> b) {
  CELV2_PLANE_BLOCK(b, t.in, t, bx, by)
  const DPlane& in = t.in;
  const long long phase_bytes = t.phase_bytes;
  const int gx = bx * 64 + (threadIdx.x & 63);
  const int y = by * 4 + (threadIdx.x >> 6) - kBorder;
  const int x0 = 4 * gx - kBorder;
  if (y >= in.h + kBorder || x0 >= in.w + kBorder) return;
  const int cy = clampi(y, 0, in.h - 1);
  uint32_t word[16];
  if (x0 >= 0 && x0 < in.w) subpel_group<true>(in, cy, x0, word);
  else subpel_group<false>(in, cy, x0 < 0 ? 0 : in.w - 1, word);
  uint8_t* const out = in.p + static_cast<ptrdiff_t>(y) * in.stride + x0;
#pragma unroll
  for (int ph = 1; ph < 16; ph++) *reinterpret_cast<uint32_t*>(out + phase_bytes * ph) = word[ph];
}

int plane_groups(const DPlane& p) {
  const int groups = (p.w + 2 * kBorder + 3) / 4;
  return ((groups + 63) / 64) * ((p.h + 2 * kBorder + 3) / 4);
}

}  // namespace

void launch_pyramid(const Batch<PyramidItem>& b, hipStream_t s) {
  if (b.total()) k_pyramid<<<b.total(), 256, 0, s>>>(b);
}
void launch_subpel(const Batch<SubpelItem>& b, hipStream_t s) {
  if (b.total()) k_subpel<<<b.total(), 256, 0, s>>>(b);
}
void add_pyramid(Batch<PyramidItem>& b, const PyramidItem& t) {
  b.add(t, ((t.lv[0].w + kTile - 1) / kTile) * ((t.lv[0].h + kTile - 1) / kTile));
}
void add_subpel(Batch<SubpelItem>& b, const SubpelItem& t) { b.add(t, plane_groups(t.in)); }

}  // namespace celv2hip
