// Hierarchical motion search, one pyramid level for every reference in one launch.
//
// One wave (64 lanes) searches one 8x8 block against one reference (me_cell + me_refine,
// src/gpu/inter_common.h:306-358):
//   1. the six candidate centres' SADs, 8 lanes per centre (one block row each), summed by
//      cross-lane adds; the centre ranking is then the reference's sequential rule, wave-uniform;
//   2. each searched centre's 17 x 17 positions: lane = (dx, third of the dy range), 51 lanes; a lane
//      keeps its window rows (13 rows x 8 bytes) in registers and scores its 5-6 positions with
//      v_sad_u8 against the block held in scalar registers; the first least cost in dy-major order is
//      the minimum of the packed key (cost << 9 | position index) over the wave;
//   3. level 0: the half-pel then quarter-pel refinement, 36 lanes = 9 positions x 4 tiles, each lane
//      one 4x4 Hadamard, keyed (cost << 4 | order) so the centre wins ties as in me_refine.
#include "common.h"
#include "launch.h"

namespace celv2hip {

namespace {

struct Item {
  int r, bx, by;
};

// Row i (0..7), dword half (0/1) of the 8x8 block at (x, y) of plane p.
__device__ __forceinline__ uint32_t block_word(const DPlane& p, int x, int y, int i, int half) {
  return *reinterpret_cast<const uint32_t*>(p.p + static_cast<ptrdiff_t>(y + i) * p.stride + x + 4 * half);
}

// SAD of the block (8 rows, words cur[2i], cur[2i+1]) against the 8x8 window at (x, y) of ref (origin
// already clamped), one row per lane group of 8: returns the group's sum in every lane of the group.
__device__ __forceinline__ uint32_t centre_sad(const DPlane& ref, int x, int y, int row, uint32_t c0, uint32_t c1) {
  uint32_t a, b;
  load8u(ref.p + static_cast<ptrdiff_t>(y + row) * ref.stride + x, a, b);
  return oct_sum(sad4(c1, b, sad4(c0, a, 0)));
}

// Full search of +-kRange around (px, py) (me_search, src/gpu/inter_common.h:273-284): the wave's
// least key cost << 9 | ((dy + 8) * 17 + dx + 8).
__device__ __forceinline__ uint32_t full_search(const DPlane& ref, int bx, int by, int px, int py, const uint32_t cur[16],
                                                int lane) {
  const bool act = lane < 51;
  const int li = act ? lane : 0;
  const int dxi = li % 17, g = li / 17, dy0 = 6 * g, ndy = g == 2 ? 5 : 6;
  const int X = clamp_origin(8 * bx + px - kRange, ref.w, 2 * kRange + 8);
  const int Y = clamp_origin(8 * by + py - kRange, ref.h, 2 * kRange + 8);
  const uint8_t* base = ref.p + static_cast<ptrdiff_t>(Y) * ref.stride + X + dxi;
  uint32_t w[13][2];
#pragma unroll
  for (int rr = 0; rr < 13; rr++) {
    const int row = min(dy0 + rr, 2 * kRange + 7);
    load8u(base + static_cast<ptrdiff_t>(row) * ref.stride, w[rr][0], w[rr][1]);
  }
  uint32_t best = 0xffffffffu;
#pragma unroll
  for (int k = 0; k < 6; k++) {
    if (k < ndy) {
      uint32_t s = 0;
#pragma unroll
      for (int i = 0; i < 8; i++) s = sad4(cur[2 * i + 1], w[k + i][1], sad4(cur[2 * i], w[k + i][0], s));
      const int dyi = dy0 + k;
      const uint32_t cost = s + kPelPenalty * (abs(dxi - kRange) + abs(dyi - kRange));
      const uint32_t key = (cost << 9) | static_cast<uint32_t>(dyi * 17 + dxi);
      best = key < best ? key : best;
    }
  }
  return wave_min_u32(act ? best : 0xffffffffu);
}

__constant__ int8_t kRefineDx[9] = {0, -1, 0, 1, -1, 1, -1, 0, 1};
__constant__ int8_t kRefineDy[9] = {0, -1, -1, -1, 0, 0, 1, 1, 1};

__global__ void __launch_bounds__(256) k_me_level(MeLevelArgs a) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int per_ref = a.rows * a.cols, nitems = a.nrefs * per_ref, ngroups = (nitems + 3) >> 2;
  const int lg = xcd_logical(blockIdx.x, ngroups);
  if (lg >= ngroups) return;
  const int item = 4 * lg + wave;
  if (item >= nitems) return;  // whole wave
  const int r = item / per_ref, rem = item - r * per_ref, by = rem / a.cols, bx = rem - by * a.cols;
  const DPlane cur = a.cur, ref = a.ref[r];

  // The block, every lane holding every word (scalar registers after readlane).
  const uint32_t mine = lane < 16 ? block_word(cur, 8 * bx, 8 * by, lane >> 1, lane & 1) : 0u;
  uint32_t c[16];
#pragma unroll
  for (int i = 0; i < 16; i++) c[i] = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(mine), i));

  int p1x = 0, p1y = 0, p2x = 0, p2y = 0;
  bool two = false;
  if (a.coarser[r]) {
    // Centres k = 0..4: the coarser MV under the block and its left, right, above and below
    // neighbours (clamped into the coarser field), doubled; k = 5: zero.
    uint32_t cm = 0;
    if (lane < 5) {
      const int ox = lane == 1 ? -1 : lane == 2 ? 1 : 0, oy = lane == 3 ? -1 : lane == 4 ? 1 : 0;
      const int qx = clampi(bx / 2 + ox, 0, a.pcols - 1), qy = clampi(by / 2 + oy, 0, a.prows - 1);
      cm = a.coarser[r][qy * a.pcols + qx];
    }
    int cx[6], cy[6];
#pragma unroll
    for (int k = 0; k < 5; k++) {
      const uint32_t m = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(cm), k));
      cx[k] = 2 * mv_x(m), cy[k] = 2 * mv_y(m);
    }
    cx[5] = 0, cy[5] = 0;
    // Each centre's SAD (lanes 8k .. 8k+7; lanes 48.. repeat centre 5).
    const int kk = min(lane >> 3, 5), row = lane & 7;
    int mx = cx[0], my = cy[0];
#pragma unroll
    for (int k = 1; k < 6; k++)
      if (kk == k) mx = cx[k], my = cy[k];
    const uint32_t rowc0 = block_word(cur, 8 * bx, 8 * by, row, 0), rowc1 = block_word(cur, 8 * bx, 8 * by, row, 1);
    const uint32_t s = centre_sad(ref, clamp_origin(8 * bx + mx, ref.w, 8), clamp_origin(8 * by + my, ref.h, 8), row,
                                  rowc0, rowc1);
    // me_cell's ranking (src/gpu/inter_common.h:331-350): each distinct centre once, best and runner-up.
    int s1 = 0x7fffffff, s2 = 0x7fffffff;
#pragma unroll
    for (int k = 0; k < 6; k++) {
      const bool dup = (k > 0 && cx[k] == p1x && cy[k] == p1y) || (two && cx[k] == p2x && cy[k] == p2y);
      if (dup) continue;
      const int sk = __builtin_amdgcn_readlane(static_cast<int>(s), 8 * k);
      if (sk < s1) {
        if (k > 0) s2 = s1, p2x = p1x, p2y = p1y, two = true;
        s1 = sk, p1x = cx[k], p1y = cy[k];
      } else if (sk < s2) {
        s2 = sk, p2x = cx[k], p2y = cy[k], two = true;
      }
    }
  }

  const uint32_t k1 = full_search(ref, bx, by, p1x, p1y, c, lane);
  int i1 = static_cast<int>(k1 & 511);
  int mvx = p1x + i1 % 17 - kRange, mvy = p1y + i1 / 17 - kRange;
  if (two) {
    const uint32_t k2 = full_search(ref, bx, by, p2x, p2y, c, lane);
    if ((k2 >> 9) < (k1 >> 9)) {
      const int i2 = static_cast<int>(k2 & 511);
      mvx = p2x + i2 % 17 - kRange, mvy = p2y + i2 / 17 - kRange;
    }
  }

  if (a.level0) {
    // me_refine (src/gpu/inter_common.h:306-319) from the integer MV, in quarter pels.
    int qcx = 4 * mvx, qcy = 4 * mvy;
    const bool act = lane < 36;
    const int o = act ? (lane >> 2) : 8, t = lane & 3, tx = t & 1, ty = t >> 1;
    uint32_t cw[4];
#pragma unroll
    for (int i = 0; i < 4; i++)
      cw[i] = *reinterpret_cast<const uint32_t*>(cur.p + static_cast<ptrdiff_t>(8 * by + 4 * ty + i) * cur.stride + 8 * bx + 4 * tx);
#pragma unroll
    for (int step = 2; step >= 1; step--) {
      const int qx = qcx + kRefineDx[o] * step, qy = qcy + kRefineDy[o] * step;
      const uint8_t* plane = ref.p + a.phase_bytes * ((qy & 3) * 4 + (qx & 3));
      const int X = clamp_origin(8 * bx + 4 * tx + (qx >> 2), ref.w, 4);
      const int Y = clamp_origin(8 * by + 4 * ty + (qy >> 2), ref.h, 4);
      uint32_t pw[4];
#pragma unroll
      for (int i = 0; i < 4; i++) pw[i] = load4u(plane + static_cast<ptrdiff_t>(Y + i) * ref.stride + X);
      const uint32_t s = quad_sum(static_cast<uint32_t>(satd4x4_rows(cw, pw)));
      const uint32_t cost = s + (o ? kQpelPenalty : 0);
      const uint32_t key = wave_min_u32(act ? ((cost << 4) | static_cast<uint32_t>(o)) : 0xffffffffu);
      const int ob = static_cast<int>(key & 15);
      qcx += kRefineDx[ob] * step, qcy += kRefineDy[ob] * step;
    }
    mvx = qcx, mvy = qcy;
  }
  if (lane == 0) {
    a.out[r][by * a.cols + bx] = pack_mv(mvx, mvy);
    if (a.searches) atomicAdd(a.searches + 16 * (item & 63), two ? 2ull : 1ull);
  }
}

}  // namespace

void launch_me_level(const MeLevelArgs& a, hipStream_t s) {
  const int nitems = a.nrefs * a.rows * a.cols;
  if (nitems <= 0) return;
  const int ngroups = (nitems + 3) / 4;
  const int grid = ((ngroups + 7) / 8) * 8;
  k_me_level<<<grid, 256, 0, s>>>(a);
}

}  // namespace celv2hip
