// Hierarchical motion search, one pyramid level for every reference in one launch.
//
// One wave (64 lanes) searches one 8x8 block against one reference (me_cell + me_refine,
// src/gpu/inter_common.h:306-358):
//   1. the six candidate centres' SADs, 8 lanes per centre (one block row each), summed by DPP;
//      the centre ranking is then the reference's sequential rule, wave-uniform (scalar code);
//   2. each searched centre's 17 x 17 positions: lane = (dx, third of the dy range), 51 lanes; a lane
//      loads its 13 window rows with one aligned buffer_load_dwordx3 each (the row's offset in an
//      SGPR) and scores its 5-6 positions with v_sad_u8 against the block held in SGPRs, the
//      distance penalty folded into the first v_sad_u8's accumulator; the first least cost in
//      dy-major order is the minimum of the key cost << 9 | position index (DPP + 4 readlanes).
//      With two centres, the second window's loads are issued before the first is scored;
//   3. level 0: the half-pel then quarter-pel refinement, 18 lanes = 9 positions x 2 tile rows,
//      each lane two 4x4 tiles in packed 16-bit arithmetic, keyed (cost << 4 | order) so the
//      centre wins ties as in me_refine.
#include "common.h"
#include "launch.h"

namespace celv2hip {

namespace {

constexpr int kWin = 2 * kRange + 8;  // 24: a search window's side

// One lane's share of a search window: rows dy0 .. dy0 + 12 of the window at (X, Y), 8 bytes from
// column dxi, as 13 x 2 words.
struct Window {
  uint32_t w[13][2];
};

__device__ __forceinline__ void load_window(const BPlane& rb, int X, int Y, int dxi, int dy0, Window& win) {
  const int off = corner_offset(rb, Y + dy0, X + dxi);
  const int base = off & ~3, sh = off & 3;
#pragma unroll
  for (int rr = 0; rr < 13; rr++) {
    uint32_t d0, d1, d2;
    bload12(rb, base, rr * rb.stride, d0, d1, d2);
    win.w[rr][0] = __builtin_amdgcn_alignbyte(d1, d0, sh);
    win.w[rr][1] = __builtin_amdgcn_alignbyte(d2, d1, sh);
  }
}

// The lane's best key over its positions: pen[k] / idx[k] hold each position's penalty and index
// (kill5: the lane has only 5 positions).
__device__ __forceinline__ uint32_t score_window(const Window& win, const uint32_t c[16], const uint32_t pen[6],
                                                 const uint32_t idx[6], uint32_t kill5) {
  uint32_t best = 0xffffffffu;
#pragma unroll
  for (int k = 0; k < 6; k++) {
    uint32_t s = pen[k];
#pragma unroll
    for (int i = 0; i < 8; i++) s = sad4(c[2 * i + 1], win.w[k + i][1], sad4(c[2 * i], win.w[k + i][0], s));
    uint32_t key = (s << 9) | idx[k];
    if (k == 5) key |= kill5;
    best = min(best, key);
  }
  return best;
}

__constant__ int8_t kRefineDx[9] = {0, -1, 0, 1, -1, 1, -1, 0, 1};
__constant__ int8_t kRefineDy[9] = {0, -1, -1, -1, 0, 0, 1, 1, 1};

__global__ void __launch_bounds__(256) k_me_level(MeLevelArgs a) {
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // wave-uniform: SGPR addressing
  const int per_ref = a.rows * a.cols, nitems = a.nrefs * per_ref, ngroups = (nitems + 3) >> 2;
  const int lg = xcd_logical(blockIdx.x, ngroups);
  if (lg >= ngroups) return;
  const int item = 4 * lg + wave;
  if (item >= nitems) return;  // whole wave
  const int r = item / per_ref, rem = item - r * per_ref, by = rem / a.cols, bx = rem - by * a.cols;
  const DPlane cur = a.cur, ref = a.ref[r];
  const BPlane rb = as_buffer(ref), cb = as_buffer(cur);

  // The block, every lane holding every word (scalar registers after readlane).
  const uint32_t mine = lane < 16 ? bload4(cb, corner_offset(cb, 8 * by + (lane >> 1), 8 * bx + 4 * (lane & 1)), 0) : 0u;
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
    // Each centre's SAD (lanes 8k .. 8k+7, one block row each; lanes 48.. repeat centre 5).
    const int kk = min(lane >> 3, 5), row = lane & 7;
    int mx = cx[0], my = cy[0];
#pragma unroll
    for (int k = 1; k < 6; k++)
      if (kk == k) mx = cx[k], my = cy[k];
    const int off = corner_offset(rb, clamp_origin(8 * by + my, ref.h, 8) + row, clamp_origin(8 * bx + mx, ref.w, 8));
    uint32_t d0, d1, d2;
    bload12(rb, off & ~3, 0, d0, d1, d2);
    const uint32_t w0 = __builtin_amdgcn_alignbyte(d1, d0, off & 3), w1 = __builtin_amdgcn_alignbyte(d2, d1, off & 3);
    uint32_t c0, c1;
    bload8(cb, corner_offset(cb, 8 * by + row, 8 * bx), 0, c0, c1);
    const uint32_t s = oct_sum(sad4(c1, w1, sad4(c0, w0, 0)));
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

  // Full searches (me_search, src/gpu/inter_common.h:273-284).
  const bool act = lane < 51;
  const int li = act ? lane : 0;
  const int dxi = li % 17, g = li / 17, dy0 = 6 * g;
  uint32_t pen[6], idx[6];
#pragma unroll
  for (int k = 0; k < 6; k++) {
    pen[k] = kPelPenalty * (abs(dxi - kRange) + abs(dy0 + k - kRange));
    idx[k] = static_cast<uint32_t>((dy0 + k) * 17 + dxi);
  }
  const uint32_t kill5 = g == 2 ? 0xffffffffu : 0u;
  Window w1, w2;
  load_window(rb, clamp_origin(8 * bx + p1x - kRange, ref.w, kWin), clamp_origin(8 * by + p1y - kRange, ref.h, kWin), dxi, dy0, w1);
  if (two) load_window(rb, clamp_origin(8 * bx + p2x - kRange, ref.w, kWin), clamp_origin(8 * by + p2y - kRange, ref.h, kWin), dxi, dy0, w2);
  const uint32_t k1 = wave_min_u32(act ? score_window(w1, c, pen, idx, kill5) : 0xffffffffu);
  const int i1 = static_cast<int>(k1 & 511);
  int mvx = p1x + i1 % 17 - kRange, mvy = p1y + i1 / 17 - kRange;
  if (two) {
    const uint32_t k2 = wave_min_u32(act ? score_window(w2, c, pen, idx, kill5) : 0xffffffffu);
    if ((k2 >> 9) < (k1 >> 9)) {
      const int i2 = static_cast<int>(k2 & 511);
      mvx = p2x + i2 % 17 - kRange, mvy = p2y + i2 / 17 - kRange;
    }
  }

  if (a.level0) {
    // me_refine (src/gpu/inter_common.h:306-319) from the integer MV, in quarter pels: lanes 2o, 2o+1
    // score position o's top and bottom tile rows (two 4x4 tiles each, packed), 18 lanes in all.
    int qcx = 4 * mvx, qcy = 4 * mvy;
    const bool on = lane < 18;
    const int o = on ? (lane >> 1) : 8, ty = lane & 1;
    uint32_t ca[4], cbw[4];
#pragma unroll
    for (int i = 0; i < 4; i++) bload8(cb, corner_offset(cb, 8 * by + 4 * ty + i, 8 * bx), 0, ca[i], cbw[i]);
    const TilePair cp = tile_pair(ca, cbw);
#pragma unroll
    for (int step = 2; step >= 1; step--) {
      const int qx = qcx + kRefineDx[o] * step, qy = qcy + kRefineDy[o] * step;
      const int X = clamp_origin(8 * bx + (qx >> 2), ref.w, 8);
      const int Y = clamp_origin(8 * by + 4 * ty + (qy >> 2), ref.h, 4);
      const int off = static_cast<int>(a.phase_bytes) * ((qy & 3) * 4 + (qx & 3)) + corner_offset(rb, Y, X);
      const int base = off & ~3, sh = off & 3;
      uint32_t pa[4], pb[4];
#pragma unroll
      for (int i = 0; i < 4; i++) {
        uint32_t d0, d1, d2;
        bload12(rb, base, i * rb.stride, d0, d1, d2);
        pa[i] = __builtin_amdgcn_alignbyte(d1, d0, sh), pb[i] = __builtin_amdgcn_alignbyte(d2, d1, sh);
      }
      const uint32_t two_t = satd4x4_pair(cp, pa, pb);
      uint32_t s = (two_t & 0xffffu) + (two_t >> 16);
      s += dpp<kDppXor1>(s);  // the other tile row of the same position
      const uint32_t cost = s + (o ? kQpelPenalty : 0);
      const uint32_t key = wave_min_u32(on ? ((cost << 4) | static_cast<uint32_t>(o)) : 0xffffffffu);
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
