// Plane building: import of a frame's luma into a bordered plane, the 2x2 pyramid levels and the
// 15 quarter-pel phase planes of a reference.
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

__global__ void __launch_bounds__(256) k_import(Batch<ImportItem> b) {
  CELV2_PLANE_BLOCK(b, t.out, t, bx, by)
  const DPlane& out = t.out;
  const int gx = bx * 64 + (threadIdx.x & 63);  // group of 4 samples
  const int y = by * 4 + (threadIdx.x >> 6) - kBorder;
  const int x0 = 4 * gx - kBorder;
  if (y >= out.h + kBorder || x0 >= out.w + kBorder) return;
  const int sy = clampi(y, 0, out.h - 1);
  uint32_t word = 0;
  if (t.wide) {
    const uint16_t* row = static_cast<const uint16_t*>(t.src) + static_cast<size_t>(sy) * t.sstride;
#pragma unroll
    for (int k = 0; k < 4; k++) word |= to8(row[clampi(x0 + k, 0, out.w - 1)], t.bit_depth) << (8 * k);
  } else {
    const uint8_t* row = static_cast<const uint8_t*>(t.src) + static_cast<size_t>(sy) * t.sstride;
#pragma unroll
    for (int k = 0; k < 4; k++) word |= static_cast<uint32_t>(row[clampi(x0 + k, 0, out.w - 1)]) << (8 * k);
  }
  *reinterpret_cast<uint32_t*>(out.p + static_cast<ptrdiff_t>(y) * out.stride + x0) = word;
}

// Level l+1 from level l: the rounded 2x2 mean (downsample_px, src/gpu/inter_common.h:91-94).
__global__ void __launch_bounds__(256) k_downsample(Batch<DownItem> b) {
  CELV2_PLANE_BLOCK(b, t.out, t, bx, by)
  const DPlane &in = t.in, &out = t.out;
  const int gx = bx * 64 + (threadIdx.x & 63);
  const int y = by * 4 + (threadIdx.x >> 6) - kBorder;
  const int x0 = 4 * gx - kBorder;
  if (y >= out.h + kBorder || x0 >= out.w + kBorder) return;
  const int cy = clampi(y, 0, out.h - 1);
  const uint8_t* r0 = in.p + static_cast<ptrdiff_t>(2 * cy) * in.stride;
  const uint8_t* r1 = r0 + in.stride;
  uint32_t word = 0;
#pragma unroll
  for (int k = 0; k < 4; k++) {
    const int cx = clampi(x0 + k, 0, out.w - 1);
    const uint32_t v = (r0[2 * cx] + r0[2 * cx + 1] + r1[2 * cx] + r1[2 * cx + 1] + 2) >> 2;
    word |= v << (8 * k);
  }
  *reinterpret_cast<uint32_t*>(out.p + static_cast<ptrdiff_t>(y) * out.stride + x0) = word;
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

void launch_import(const Batch<ImportItem>& b, hipStream_t s) {
  if (b.total()) k_import<<<b.total(), 256, 0, s>>>(b);
}
void launch_downsample(const Batch<DownItem>& b, hipStream_t s) {
  if (b.total()) k_downsample<<<b.total(), 256, 0, s>>>(b);
}
void launch_subpel(const Batch<SubpelItem>& b, hipStream_t s) {
  if (b.total()) k_subpel<<<b.total(), 256, 0, s>>>(b);
}
void add_import(Batch<ImportItem>& b, const ImportItem& t) { b.add(t, plane_groups(t.out)); }
void add_downsample(Batch<DownItem>& b, const DownItem& t) { b.add(t, plane_groups(t.out)); }
void add_subpel(Batch<SubpelItem>& b, const SubpelItem& t) { b.add(t, plane_groups(t.in)); }

}  // namespace celv2hip
