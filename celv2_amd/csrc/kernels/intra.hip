// Intra scores of a superblock: SATD of the 13 intra modes (block_mode_satd,
// src/gpu/intra_common.h:149-163) for each of the 420 candidate blocks of 32x32 and below.
//
// The 4x4 Walsh-Hadamard transform is linear: for 16 tiles at once, (H (x) H) vec(cur - pred) is one
// 16x16x16 matrix product. A sample v in 0..255 becomes the exact half 1024 + v by placing the byte
// 0x64 above it, so v_mfma_f32_16x16x16_f16 computes every coefficient exactly in f32 (integers
// below 2^24). Per 16 tiles: the source tiles' transform C = HH * (1024 + cur) once, then per mode
// D = C - HH * (1024 + pred); |D| costs nothing on the f32 add.
//
// Work layout. The 2304 (block, 4x4 tile) pairs of a superblock are sorted by block shape (nine
// shapes of 256 tiles each), a block's tiles contiguous, so every 16-tile "step" has one shape and
// a block's tiles fill an aligned group of 2..16 columns of the product. Lane (j, g) = (lane & 15,
// lane >> 4) holds row g of tile j: it predicts that row's four samples in each mode. A superblock is
// split over 4 workgroups of 36 steps, at block boundaries; each wave takes every 4th step. Per step
// the 13 modes' sums are reduced two at a time (row swaps, then DPP within rows), and one lane per
// block and mode adds the total to the block's LDS accumulator.
//
// Edges. The source around the superblock (rows and columns -1 .. 95) is staged in LDS twice. The
// copies differ only in row -1 and column -1, so that every edge sample the reference reads through
// edge_above / edge_left (intra_common.h:47-56) is a plain LDS byte: the above copy has row -1 = 128
// at the frame top and column -1 = the clamped sample; the left copy has column -1 = 128 at the frame
// left and row -1 = the clamped sample. The directional modes read two samples per predicted sample
// at offsets from a static per-(shape, mode, row, column) table built from dir_pred's geometry.
#include <algorithm>
#include <stdexcept>
#include <type_traits>
#include <vector>

#include "common.h"
#include "hadamard.h"
#include "launch.h"

namespace celv2hip {

namespace {

constexpr int kRS = 104;               // region row stride (bytes): columns -4 .. 99
constexpr int kRRows = 97;             // rows -1 .. 95
constexpr int kRBytes = kRRows * kRS;  // one copy of the region
constexpr int kShapes = 9;
constexpr int kTilesPerShape = 256;
constexpr int kIntraTiles = kShapes * kTilesPerShape;  // 2304
constexpr int kSteps = kIntraTiles / 16;               // 144
constexpr int kGroups = 4;                             // workgroups per superblock
constexpr int kGroupSteps = kSteps / kGroups;          // 36
constexpr int kMaxGroupBlocks = 272;
constexpr int kDirModes = 6;                           // modes 7..12
constexpr int kDirEntries = kDirModes * 2688;          // 2688 = samples of the nine shapes
constexpr int kMultiBlocks = 20;                       // 32x32, 32x16, 16x32: the first 20 positions

struct ShapeInfo {
  int w, h, lw, lh;     // size and log2 size
  int den;              // smooth: w + h + 2
  uint32_t magic;       // ceil(2^32 / den): floor(n / den) = mulhi(n, magic) for n < 2^16
  int seg;              // tiles per block, at most 16
  int dir;              // first directional table entry
};

__constant__ ShapeInfo g_shape[kShapes];
__constant__ int g_group_pos[kGroups + 1];  // block positions (shape-sorted order) of each group
__device__ uint32_t g_itile[kIntraTiles];   // x/4 | y/4 << 4 | tx << 8 | ty << 11 | pos << 14 | block << 23
__device__ uint16_t g_pos_block[420];       // block index of each position
__device__ __align__(16) uint32_t g_dir[kDirEntries];  // uint16 offset | f << 16 | step << 21

__global__ void __launch_bounds__(256) k_intra(Batch<IntraItem> batch) {
  __shared__ __align__(16) uint8_t reg[2 * kRBytes];  // above copy, left copy
  __shared__ float macc[kMultiBlocks * kModes];  // 32x32, 32x16, 16x32 blocks (their tiles span steps)
  __shared__ uint32_t dcv[kMaxGroupBlocks];
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int bi = batch_item(batch, static_cast<int>(blockIdx.x));
  const IntraItem& it = batch.item[bi];
  const DPlane& src = it.src;
  const int lw = static_cast<int>(blockIdx.x) - batch.first[bi], sb_cols = it.sb_cols;
  const int q = lw & (kGroups - 1), sbx = (lw >> 2) % sb_cols, sby = it.row0 + (lw >> 2) / sb_cols;
  uint32_t* __restrict__ const out = it.out;
  const int ox = 64 * sbx, oy = 64 * sby;
  const int p0 = g_group_pos[q], np = g_group_pos[q + 1] - p0;
  // The superblock's scored mode groups (uniform over the workgroup); group 0 is always scored.
  const uint32_t gm = (it.mask ? it.mask[sby * sb_cols + sbx] : kAllGroups) | 1u;
  const bool g1 = gm & 2u, g2 = gm & 4u, g3 = gm & 8u;
  if (it.used_mask && q == 0 && tid == 0) it.used_mask[sby * sb_cols + sbx] = static_cast<uint8_t>(gm);

  // Stage: region row r (frame row oy + r - 1, clamped), dword d = columns ox - 4 + 4d .. + 3 (a
  // dword entirely right of the frame reads the replicated border at w + 60).
  uint32_t* ra = reinterpret_cast<uint32_t*>(reg);
  uint32_t* rl = reinterpret_cast<uint32_t*>(reg + kRBytes);
  for (int k = tid; k < kRRows * (kRS / 4); k += 256) {
    const int r = k / (kRS / 4), d = k - r * (kRS / 4);
    const int yy = clampi(oy + r - 1, 0, src.h - 1), xx = min(ox - 4 + 4 * d, src.w + 60);
    const uint32_t v = *reinterpret_cast<const uint32_t*>(src.p + static_cast<ptrdiff_t>(yy) * src.stride + xx);
    ra[k] = v, rl[k] = v;
  }
  for (int k = tid; k < kMultiBlocks * kModes; k += 256) macc[k] = 0.f;
  __syncthreads();
  if (oy == 0 && tid < kRS / 4) ra[tid] = 0x80808080u;    // above copy: row -1 is 128 at the frame top
  if (ox == 0 && tid < kRRows) reg[kRBytes + tid * kRS + 3] = 128;  // left copy: column -1 at the frame left
  __syncthreads();

  // dc_value (intra_common.h:120-128) of the group's blocks.
  for (int i = tid; i < np; i += 256) {
    int x, y, w, h;
    block_geom(g_pos_block[p0 + i], x, y, w, h);
    const int blk = (y + 1) * kRS + x + 4;
    uint32_t sum = 0, n = 0;
    if (oy + y > 0) {
      for (int j = 0; j < w; j += 4) sum = sad4(*reinterpret_cast<const uint32_t*>(reg + blk - kRS + j), 0u, sum);
      n += w;
    }
    if (ox + x > 0) {
      for (int r = 0; r < h; r++) sum += reg[kRBytes + blk + r * kRS - 1];
      n += h;
    }
    dcv[i] = n ? (sum + n / 2) / n : 128u;
  }
  __syncthreads();

  const half4 hh = hadamard_operand(lane, false), nhh = hadamard_operand(lane, true);
  const f32x4 zero = {0.f, 0.f, 0.f, 0.f};
  const int j16 = lane & 15, g = lane >> 4;
  uint32_t* __restrict__ const o = out + (static_cast<size_t>(sby) * sb_cols + sbx) * kBlocks * kModes;
  for (int step = kGroupSteps * q + wave; step < kGroupSteps * (q + 1); step += 4) {
    const ShapeInfo sh = g_shape[step >> 4];
    const uint32_t e = g_itile[16 * step + j16];
    const int x = 4 * static_cast<int>(e & 15), y = 4 * static_cast<int>((e >> 4) & 15);
    const int pj0 = 4 * static_cast<int>((e >> 8) & 7), pi = 4 * static_cast<int>((e >> 11) & 7) + g;
    const int pos = static_cast<int>((e >> 14) & 511), block = static_cast<int>(e >> 23);
    const int w = sh.w, h = sh.h;
    const int blk = (y + 1) * kRS + x + 4;  // offset of sample (y, x) in the above copy
    const uint32_t cur = *reinterpret_cast<const uint32_t*>(reg + blk + pi * kRS + pj0);
    const uint32_t a4 = *reinterpret_cast<const uint32_t*>(reg + blk - kRS + pj0);  // above[pj0 .. pj0 + 3]
    const int lft = reg[kRBytes + blk + pi * kRS - 1];                              // left[pi]
    const int tr = reg[blk - kRS + w - 1], bl = reg[kRBytes + blk + (h - 1) * kRS - 1];
    const int tl = reg[blk - kRS - 1];
    const int dc = static_cast<int>(dcv[pos - p0]);
    const f32x4 ccur = mfma16(hh, row_halves(cur), zero);
    int a[4];
#pragma unroll
    for (int jj = 0; jj < 4; jj++) a[jj] = static_cast<int>((a4 >> (8 * jj)) & 255);

    // The step's work for blocks of SEG tiles (the shape's; uniform over the step).
    auto body = [&](auto seg_c) {
      constexpr int SEG = decltype(seg_c)::value;
      // A mode's SATD per lane is a quarter tile (|coefficients| of row g of tile j). Two modes at a
      // time: a tile's 16 coefficients sit in lanes j, j + 16, j + 32, j + 48, so the row sums leave
      // mode m's tile totals in row 0 and mode m + 1's in row 1; the block's SEG columns (an aligned
      // group) are then summed within the row, and one lane per block and mode keeps the total: a
      // store for a block whose tiles are all in this step, an add to the LDS accumulator for the
      // 32x32, 32x16 and 16x32 blocks, whose tiles span several steps.
      auto emit = [&](int m, float pa, float pb) {
        float t = sum_rows16(pa, pb);
        t = sum_halves32(t, t);
        if constexpr (SEG >= 2) t += as_f32(dpp<kDppXor1>(as_u32f(t)));
        if constexpr (SEG >= 4) t += as_f32(dpp<kDppXor2>(as_u32f(t)));
        if constexpr (SEG >= 8) t += as_f32(dpp<kDppHalfMirror>(as_u32f(t)));
        if constexpr (SEG >= 16) t += as_f32(dpp<kDppMirror>(as_u32f(t)));
        const int mode = m + g;
        if (g < 2 && (j16 & (SEG - 1)) == 0 && mode < kModes) {
          if (SEG == 16 && pos < kMultiBlocks)
            atomicAdd(&macc[pos * kModes + mode], t);
          else
            o[block * kModes + mode] = static_cast<uint32_t>(t);
        }
      };
      auto sat = [&](uint32_t p01, uint32_t p23) { return abs_sum4(mfma16(nhh, as_half4(p01, p23), ccur)); };

      // 0 DC, 1 V, 2 H (intra_common.h:89-96).
      const uint32_t pd = (static_cast<uint32_t>(dc) | 0x6400u) * 0x10001u;
      emit(0, sat(pd, pd), sat(halves01(a4), halves23(a4)));
      const uint32_t ph = (static_cast<uint32_t>(lft) | 0x6400u) * 0x10001u;
      const float s2 = sat(ph, ph);
      // Modes of a group that is not scored are not computed (0 is left in their place; finalize
      // skips them), and a pair of them is not summed.
      // 3 Smooth: (h-i) a[j] + (i+1) bl + (w-j) l[i] + (j+1) tr, rounded, / (w + h + 2).
      float s3 = 0.f;
      if (g1) {
        const int base = (pi + 1) * bl + (w - pj0) * lft + (pj0 + 1) * tr + sh.den / 2, dl = tr - lft, al = h - pi;
        uint32_t v[4];
#pragma unroll
        for (int jj = 0; jj < 4; jj++) v[jj] = __umulhi(static_cast<uint32_t>(al * a[jj] + base + jj * dl), sh.magic);
        s3 = sat(pack_lo16(v[0], v[1]) | 0x64006400u, pack_lo16(v[2], v[3]) | 0x64006400u);
      }
      emit(2, s2, s3);
      // 4 Smooth V: ((h-1-i) a[j] + (i+1) left[h-1] + h/2) / h.
      float s4 = 0.f;
      if (g1) {
        const int c = (pi + 1) * bl + h / 2 + (0x6400 << sh.lh), al = h - 1 - pi;
        uint32_t v[4];
#pragma unroll
        for (int jj = 0; jj < 4; jj++) v[jj] = static_cast<uint32_t>((al * a[jj] + c) >> sh.lh);
        s4 = sat(pack_lo16(v[0], v[1]), pack_lo16(v[2], v[3]));
      }
      // 5 Smooth H: ((w-1-j) l[i] + (j+1) above[w-1] + w/2) / w.
      float s5 = 0.f;
      if (g1) {
        const int c = (w - 1 - pj0) * lft + (pj0 + 1) * tr + w / 2 + (0x6400 << sh.lw), dl = tr - lft;
        uint32_t v[4];
#pragma unroll
        for (int jj = 0; jj < 4; jj++) v[jj] = static_cast<uint32_t>((c + jj * dl) >> sh.lw);
        s5 = sat(pack_lo16(v[0], v[1]), pack_lo16(v[2], v[3]));
      }
      if (g1) emit(4, s4, s5);
      // 6 Paeth: the neighbour nearest to above + left - top-left, ties to left, then above.
      float s6 = 0.f;
      if (g1) {
        const int pa = abs(lft - tl);
        uint32_t v[4];
#pragma unroll
        for (int jj = 0; jj < 4; jj++) {
          const int pl = abs(a[jj] - tl), ptl = abs(a[jj] + lft - 2 * tl);
          const int p = (pl <= pa && pl <= ptl) ? lft : (pa <= ptl ? a[jj] : tl);
          v[jj] = static_cast<uint32_t>(p) | 0x6400u;
        }
        s6 = sat(pack_lo16(v[0], v[1]), pack_lo16(v[2], v[3]));
      }
      // 7..12 directional (intra_common.h:61-85): two samples and a 5-bit weight per predicted sample.
      // The table through a buffer resource: the lane's entry offset once, each mode's part in the
      // scalar offset (no per-load address arithmetic).
      const __amdgpu_buffer_rsrc_t dtab = __builtin_amdgcn_make_buffer_rsrc(g_dir, static_cast<short>(0), static_cast<int>(sizeof(g_dir)), 0x00020000);
      const int dlane = 4 * (static_cast<int>(umul24_sv(static_cast<uint32_t>(w), static_cast<uint32_t>(pi))) + pj0);
      auto dir = [&](int mm) {
        const auto t = __builtin_amdgcn_raw_buffer_load_b128(dtab, dlane, 4 * (sh.dir + mm * h * w), 0);
        const uint32_t te[4] = {t[0], t[1], t[2], t[3]};
        uint32_t v[4];
#pragma unroll
        for (int jj = 0; jj < 4; jj++) {
          const int off = blk + static_cast<int>(static_cast<int16_t>(te[jj] & 0xffffu));
          const int f = static_cast<int>((te[jj] >> 16) & 31), st = static_cast<int>(te[jj] >> 21);
          const int e0 = reg[off], e1 = reg[off + st];
          v[jj] = static_cast<uint32_t>(e0 + (((e1 - e0) * f + 16) >> 5) + 0x6400);
        }
        return sat(pack_lo16(v[0], v[1]), pack_lo16(v[2], v[3]));
      };
      if (g1 || g2) emit(6, s6, g2 ? dir(0) : 0.f);
      if (g2) emit(8, dir(1), dir(2));
      if (g3) {
        emit(10, dir(3), dir(4));
        emit(12, dir(5), 0.f);
      }
    };
    switch (sh.seg) {
      case 16: body(std::integral_constant<int, 16>{}); break;
      case 8: body(std::integral_constant<int, 8>{}); break;
      case 4: body(std::integral_constant<int, 4>{}); break;
      default: body(std::integral_constant<int, 2>{}); break;
    }
  }
  __syncthreads();
  // The multi-step blocks (positions 0 .. kMultiBlocks - 1, in the first groups).
  for (int k = tid; k < np * kModes; k += 256) {
    const int i = k / kModes, m = k - i * kModes;
    if (p0 + i < kMultiBlocks) o[g_pos_block[p0 + i] * kModes + m] = static_cast<uint32_t>(macc[(p0 + i) * kModes + m]);
  }
}

// dir_pred's sample choice (intra_common.h:61-85) for sample (i, j) of a w x h block: the edge
// (0 above, 1 left), the two indices along it, and the 5-bit weight of the second.
struct DirSample {
  int edge, k0, k1, f;
};
DirSample dir_sample(int w, int h, int angle, int i, int j) {
  const int last = w + h - 1;
  auto deriv = [](int a) { return a == 23 ? 151 : a == 45 ? 64 : 27; };
  auto fdiv64 = [](int v) { return v >= 0 ? v / 64 : -((-v + 63) / 64); };
  if (angle < 90) {
    const int pos = (i + 1) * deriv(angle), base = (pos >> 6) + j;
    return {0, std::min(base, last), std::min(base + 1, last), (pos & 63) >> 1};
  }
  if (angle > 180) {
    const int pos = (j + 1) * deriv(270 - angle), base = (pos >> 6) + i;
    return {1, std::min(base, last), std::min(base + 1, last), (pos & 63) >> 1};
  }
  const int px = 64 * j - (i + 1) * deriv(180 - angle), bx = fdiv64(px);
  if (bx >= -1) return {0, bx, bx + 1, (px - 64 * bx) >> 1};
  const int py = 64 * i - (j + 1) * deriv(angle - 90), by = fdiv64(py);
  return {1, std::max(by, -1), std::max(by + 1, -1), (py - 64 * by) >> 1};
}

}  // namespace

void add_intra(Batch<IntraItem>& b, const IntraItem& t) {
  if (t.nrows > 0) b.add(t, kGroups * t.sb_cols * t.nrows);
}
void launch_intra(const Batch<IntraItem>& b, hipStream_t s) {
  if (b.total() > 0) k_intra<<<b.total(), 256, 0, s>>>(b);
}

// The shape-sorted tile list, the block order, the group ranges and the directional sample table,
// built in local memory and copied to this device's symbols (synchronous).
hipError_t upload_intra_tiles() {
  static const int kShapeWH[kShapes][2] = {{32, 32}, {32, 16}, {16, 32}, {16, 16}, {16, 8}, {8, 16}, {8, 8}, {8, 4}, {4, 8}};
  static const int kAngles[kDirModes] = {45, 67, 113, 135, 157, 203};
  std::vector<uint32_t> tiles(kIntraTiles);
  std::vector<uint16_t> pos_block(420);
  std::vector<uint32_t> dir(kDirEntries);
  ShapeInfo shapes[kShapes] = {};
  int group_pos[kGroups + 1] = {};
  std::vector<int> tile_pos;
  int nt = 0, np = 0, nd = 0;
  for (int si = 0; si < kShapes; si++) {
    const int w = kShapeWH[si][0], h = kShapeWH[si][1];
    ShapeInfo& sh = shapes[si];
    sh.w = w, sh.h = h, sh.lw = __builtin_ctz(w), sh.lh = __builtin_ctz(h), sh.den = w + h + 2;
    sh.magic = static_cast<uint32_t>(((uint64_t(1) << 32) + sh.den - 1) / sh.den);
    sh.seg = std::min(16, (w / 4) * (h / 4)), sh.dir = nd;
    for (int b = 5; b < kBlocks; b++) {
      const int sq = b / 5, kind = b % 5, level = sq >= 21 ? 3 : sq >= 5 ? 2 : 1;
      const int size = 64 >> level, first = ((1 << (2 * level)) - 1) / 3, i = sq - first, n = 1 << level;
      const int bw = (kind == 3 || kind == 4) ? size / 2 : size, bh = (kind == 1 || kind == 2) ? size / 2 : size;
      if (bw != w || bh != h) continue;
      const int bx = (i & (n - 1)) * size + (kind == 4 ? size / 2 : 0), by = (i >> level) * size + (kind == 2 ? size / 2 : 0);
      for (int ty = 0; ty < h / 4; ty++)
        for (int tx = 0; tx < w / 4; tx++) {
          tiles[nt++] = static_cast<uint32_t>(bx / 4) | static_cast<uint32_t>(by / 4) << 4 | static_cast<uint32_t>(tx) << 8 |
                        static_cast<uint32_t>(ty) << 11 | static_cast<uint32_t>(np) << 14 | static_cast<uint32_t>(b) << 23;
          tile_pos.push_back(np);
        }
      pos_block[np++] = static_cast<uint16_t>(b);
    }
    if (nt != kTilesPerShape * (si + 1)) throw std::logic_error("intra shape tile count");
    // The shapes whose blocks span several 16-tile steps come first: positions 0 .. kMultiBlocks - 1.
    if (si == 2 && np != kMultiBlocks) throw std::logic_error("intra multi-step block count");
    for (int mm = 0; mm < kDirModes; mm++)
      for (int i = 0; i < h; i++)
        for (int j = 0; j < w; j++) {
          const DirSample d = dir_sample(w, h, kAngles[mm], i, j);
          // Offsets from sample (0, 0) of the block in the above copy; the left copy follows it.
          const int off = d.edge == 0 ? -kRS + d.k0 : kRBytes + d.k0 * kRS - 1;
          const int step = d.edge == 0 ? d.k1 - d.k0 : (d.k1 - d.k0) * kRS;
          if (off < -32768 || off > 32767 || step < 0 || step > 127 || d.f < 0 || d.f > 31) throw std::logic_error("intra direction entry");
          dir[nd++] = (static_cast<uint32_t>(off) & 0xffffu) | static_cast<uint32_t>(d.f) << 16 | static_cast<uint32_t>(step) << 21;
        }
  }
  if (nt != kIntraTiles || np != 420 || nd != kDirEntries) throw std::logic_error("intra table size");
  for (int q = 0; q <= kGroups; q++) group_pos[q] = q < kGroups ? tile_pos[16 * kGroupSteps * q] : np;
  for (int q = 0; q < kGroups; q++) {
    // A block's tiles must not straddle two groups, and a group must fit the LDS accumulators.
    const int last = tile_pos[16 * kGroupSteps * (q + 1) - 1];
    if (last + 1 != group_pos[q + 1] || group_pos[q + 1] - group_pos[q] > kMaxGroupBlocks) throw std::logic_error("intra group split");
  }
  hipError_t rc = hipMemcpyToSymbol(HIP_SYMBOL(g_itile), tiles.data(), tiles.size() * sizeof(uint32_t));
  if (rc == hipSuccess) rc = hipMemcpyToSymbol(HIP_SYMBOL(g_pos_block), pos_block.data(), pos_block.size() * sizeof(uint16_t));
  if (rc == hipSuccess) rc = hipMemcpyToSymbol(HIP_SYMBOL(g_dir), dir.data(), dir.size() * sizeof(uint32_t));
  if (rc == hipSuccess) rc = hipMemcpyToSymbol(HIP_SYMBOL(g_shape), shapes, sizeof(shapes));
  if (rc == hipSuccess) rc = hipMemcpyToSymbol(HIP_SYMBOL(g_group_pos), group_pos, sizeof(group_pos));
  return rc;
}

}  // namespace celv2hip
