// Hierarchical motion search, one pyramid level for every reference in one launch.
//
// One wave (64 lanes) searches one 8x8 block against one reference (me_cell + me_refine,
// src/gpu/inter_common.h:135-193):
//   1. the six candidate centres' SADs, 8 lanes per centre (one block row each), summed by DPP and
//      read into scalar registers; the centre ranking is then scalar code;
//   2. each searched centre's 24 x 24 window is read from the plane once (144 word pairs, 6 per row)
//      into the wave's LDS area as 4 byte-shifted copies; both windows' positions but their centres
//      (2 x 288, whose centre SADs step 1 has) are then scored in one pass by all 64 lanes, 9 each:
//      lane (window, half of the dy range, dx < 16) takes 15 rows as aligned words of the copy its
//      offset selects for 8 positions, and 8 more rows for one position of column 16 or row 8, with
//      v_sad_u8 against the block held in SGPRs (scalar loads). A position's key is
//      (SAD << 10) + (penalty << 10 | window << 9 | dy * 17 + dx), its lane constant computed per
//      block, so the wave minimum (DPP + 4 readlanes, then the centres' keys on the scalar unit) is
//      the first least cost in dy-major order, window 1 before window 2 at equal cost;
//   3. level 0: the half-pel then quarter-pel refinement: the half-pel step's 9 positions x 4 tiles
//      are 36 columns of three 16x16x16 Hadamard products on the matrix cores, accumulated onto the
//      block's own transform (hadamard.h), the quarter-pel step's 8 new positions two products (its
//      centre is the half-pel winner, already scored); the key (cost << 4 | order) makes the centre
//      win ties as in me_refine.
#include <stdexcept>

#include "common.h"
#include "hadamard.h"
#include "launch.h"

namespace celv2hip {

namespace {

constexpr int kWin = 2 * kRange + 8;  // 24: a search window's side
// A window is staged as 4 copies of its 24 rows, copy s holding the aligned row's bytes from s on, so
// a scoring lane reads its 8 samples at any byte offset o (0..19) as 2 aligned words of copy o & 3.
// The copies are interleaved by word: word q of copy s of row r at r * kRowWords + 4 q + s, so the
// lane's two words sit at o and o + 4 from its row's start, and 16 lanes with consecutive offsets
// read 16 consecutive words (16 LDS banks). Rows are an odd number of words apart, so lanes reading
// one column down 17 rows meet 17 different banks; the second window starts 16 banks after the
// first, so the two windows' 16-lane groups of one half-wave read disjoint banks (up to 3 shared when
// the first window's byte shift is the larger).
constexpr int kCopyWords = 6;               // words per copied row: words 0..5 cover offsets up to 19 + 8
constexpr int kRowWords = 4 * kCopyWords + 1;  // 25
constexpr int kWinWords = kWin * kRowWords;    // 600
constexpr int kWinB = 624;                     // the second window's start: 624 = 16 (mod 32)
// A wave's area: both windows (24 rows each; the second 24 words after the first's end).
constexpr int kWaveLds = kWinB + kWinWords;    // 1224

// Stages a window: lane (row group q = lane >> 3, word d = lane & 7) copies word d of rows q, q + 8
// and q + 16 into every copy, from the aligned row's words d and d + 1 (rows from the window's first
// byte rounded down); lanes with d > 5 idle. off = the lane's row and word offset in the plane, dst =
// its LDS word in the window (both for row q), base = the window's aligned first byte in the plane
// (wave-uniform: the buffer's scalar offset).
__device__ __forceinline__ void stage_window(const BPlane& rb, int base, int off, bool active, uint32_t* dst) {
  if (active) {
#pragma unroll
    for (int i = 0; i < 3; i++) {
      uint32_t lo, hi;
      bload8(rb, off, base + 8 * i * rb.stride, lo, hi);
      uint32_t* o = dst + 8 * i * kRowWords;
      o[0] = lo;
      o[1] = __builtin_amdgcn_alignbyte(hi, lo, 1);
      o[2] = __builtin_amdgcn_alignbyte(hi, lo, 2);
      o[3] = __builtin_amdgcn_alignbyte(hi, lo, 3);
    }
  }
}

// Lane constants of the search, computed per block (a table of them cost more in loads than the
// arithmetic). Lane (dy group g = lane >> 5, window w = (lane >> 4) & 1, column dx = lane & 15)
// scores the 8 main positions (dx, 9 g + k), k < 8, from the 15 rows of its strip, and one extra
// position (ex, ey): (dx, 8) for g = 0, except (16, 16) in place of the centre (8, 8) for dx = 8, and
// (16, dx) for g = 1. Over the 64 lanes that is every position of both windows but the two centres,
// whose SADs the centre ranking has already: 2 x 288 = 9 x 64.
//   t[k], k < 8: the main keys (kPelPenalty (|dx - 8| + |dy - 8|)) << 10 | w << 9 | dy * 17 + dx;
//     along k the penalty falls by kPelPenalty while dy < 8 (group 0) and rises after (group 1):
//     each key is the last plus a step;
//   t[8]: the extra position's key;
//   t[9]: w; t[10]: the strip's first word, dx + 9 g kRowWords; t[11]: the extra position's first
//     word, ex + ey kRowWords (both before the window's byte shift: a sample at byte offset o of row y
//     is word o + y kRowWords of the copies, words o and o + 4 holding its 8 bytes).
__device__ __forceinline__ void search_lanes(int lane, uint32_t (&t)[12]) {
  const int g = lane >> 5, w = (lane >> 4) & 1, dx = lane & 15;
  const uint32_t pen0 = static_cast<uint32_t>(kPelPenalty * (abs(dx - kRange) + (g ? 1 : kRange)));
  t[0] = pen0 << 10 | static_cast<uint32_t>(w) << 9 | static_cast<uint32_t>(153 * g + dx);
  const uint32_t step = g ? (static_cast<uint32_t>(kPelPenalty) << 10) + 17u : 17u - (static_cast<uint32_t>(kPelPenalty) << 10);
#pragma unroll
  for (int k = 1; k < 8; k++) t[k] = t[k - 1] + step;
  const int ex = g || dx == kRange ? 2 * kRange : dx, ey = g ? dx : dx == kRange ? 2 * kRange : kRange;
  const uint32_t xpen = static_cast<uint32_t>(kPelPenalty * (abs(ex - kRange) + abs(ey - kRange)));
  t[8] = xpen << 10 | static_cast<uint32_t>(w) << 9 | static_cast<uint32_t>(17 * ey + ex);
  t[9] = static_cast<uint32_t>(w), t[10] = static_cast<uint32_t>(dx + 9 * kRowWords * g), t[11] = static_cast<uint32_t>(ex + kRowWords * ey);
}

// The lane's least key over its positions of the two staged windows A and B (LDS word addresses,
// byte shifts shA / shB of their first samples; B may be A).
__device__ __forceinline__ uint32_t score_pair(const uint32_t (&t)[12], const uint32_t* winA, const uint32_t* winB, int shA, int shB,
                                               const uint32_t c[16]) {
  const uint32_t* win = t[9] ? winB : winA;
  const int sh = t[9] ? shB : shA;
  const uint32_t* row = win + sh + t[10];
  uint32_t w[15][2];
#pragma unroll
  for (int rr = 0; rr < 15; rr++) w[rr][0] = row[rr * kRowWords], w[rr][1] = row[rr * kRowWords + 4];
  uint32_t best = 0xffffffffu;
#pragma unroll
  for (int k = 0; k < 8; k++) {
    uint32_t acc = 0;
#pragma unroll
    for (int i = 0; i < 8; i++) acc = sad4(c[2 * i + 1], w[k + i][1], sad4(c[2 * i], w[k + i][0], acc));
    best = min(best, (acc << 10) + t[k]);
  }
  const uint32_t* xrow = win + sh + t[11];
  uint32_t acc = 0;
#pragma unroll
  for (int i = 0; i < 8; i++) acc = sad4(c[2 * i + 1], xrow[i * kRowWords + 4], sad4(c[2 * i], xrow[i * kRowWords], acc));
  return min(best, (acc << 10) + t[8]);
}

// The narrow form of the two full searches (the pruned form's search mode): me_search with range
// kNarrowRange instead of kRange around each centre, read straight from the plane (the two windows are
// small enough to stay in the cache). Lane (window w, column dx, row group g) scores kNarrowRows
// vertically adjacent positions from the kNarrowRows + 7 rows it loads; the last group starts early
// enough to stay inside the window (positions scored twice give the same key). Keys as in the full
// form: (SAD << 10) + (penalty << 10 | w << 9 | (dy + kRange - R) * 17 + dx + kRange - R), so the wave
// minimum is the first least cost in dy-major order, window 1 before window 2 at equal cost, and
// decodes as the full form's keys do.
constexpr uint32_t kNoPosition = 0x7fffffffu;  // above any key (SAD << 10 < 2^24)
constexpr int kNarrowSide = 2 * kNarrowRange + 1;                         // positions per window side
constexpr int kNarrowRows = kNarrowRange <= 2 ? 1 : kNarrowRange == 3 ? 2 : 3;  // positions per lane
constexpr int kNarrowGroups = (kNarrowSide + kNarrowRows - 1) / kNarrowRows;
constexpr int kNarrowLanes = kNarrowSide * kNarrowGroups;                // lanes per window
constexpr int kNarrowWin = 2 * kNarrowRange + 8;                          // a narrow window's side
static_assert(2 * kNarrowLanes <= 64 && kNarrowRange < kRange, "narrow search lanes");

__device__ __forceinline__ uint32_t search_narrow(const BPlane& rb, int X1, int Y1, int X2, int Y2, bool two,
                                                  const uint32_t c[16], int lane) {
  const int w = lane >= kNarrowLanes ? 1 : 0, q = lane - kNarrowLanes * w;
  const int g = q / kNarrowSide, dx = q - kNarrowSide * g;
  const int dy0 = min(g * kNarrowRows, kNarrowSide - kNarrowRows);
  const bool active = lane < 2 * kNarrowLanes && (w == 0 || two);
  // Inactive lanes read window 1's first rows (in range) and are given no key.
  const int X = w && two ? X2 : X1, Y = w && two ? Y2 : Y1;
  const int off = corner_offset_lane(rb, Y + (active ? dy0 : 0), X + (active ? dx : 0));
  const int base = off & ~3, sh = off & 3;
  uint32_t rw[kNarrowRows + 7][2];
#pragma unroll
  for (int i = 0; i < kNarrowRows + 7; i++) {
    uint32_t d0, d1, d2;
    bload12(rb, base, i * rb.stride, d0, d1, d2);
    rw[i][0] = __builtin_amdgcn_alignbyte(d1, d0, sh);
    rw[i][1] = __builtin_amdgcn_alignbyte(d2, d1, sh);
  }
  uint32_t best = kNoPosition;
#pragma unroll
  for (int k = 0; k < kNarrowRows; k++) {
    uint32_t acc = 0;
#pragma unroll
    for (int i = 0; i < 8; i++) acc = sad4(c[2 * i + 1], rw[k + i][1], sad4(c[2 * i], rw[k + i][0], acc));
    const int dy = dy0 + k;
    const uint32_t pen = static_cast<uint32_t>(kPelPenalty * (abs(dx - kNarrowRange) + abs(dy - kNarrowRange)));
    const uint32_t idx = static_cast<uint32_t>((dy + kRange - kNarrowRange) * 17 + dx + kRange - kNarrowRange);
    best = min(best, ((acc + pen) << 10) | static_cast<uint32_t>(w) << 9 | idx);
  }
  return active ? best : kNoPosition;
}

// me_refine's 9 positions (src/gpu/inter_common.h:135-154), 2 bits each as offset + 1; 12..15 pad.
constexpr uint32_t kRefDx = (1u << 0) | (0u << 2) | (1u << 4) | (2u << 6) | (0u << 8) | (2u << 10) | (0u << 12) | (1u << 14) | (2u << 16) | (1u << 18) | (1u << 20) | (1u << 22);
constexpr uint32_t kRefDy = (1u << 0) | (0u << 2) | (0u << 4) | (0u << 6) | (1u << 8) | (1u << 10) | (2u << 12) | (2u << 14) | (2u << 16) | (1u << 18) | (1u << 20) | (1u << 22);

// The loads that open a block's search, issued together: lanes 0..8 the reference limits of the 3x3
// superblocks around it (0 outside the frame or without limits), lanes 0..15 the block's 16 words
// (row i's two words in lanes 2i, 2i + 1; every later use of the block's samples takes them from
// here by lane shuffles, not from memory again: the search is bound by the vector-memory data path),
// the coarser MVs under and beside it are read by scalar loads where they are used.
struct Opening {
  uint32_t km = 0, sm = 0, mine = 0;
};

// v in lanes 8k .. 8k+7 (row k >> 1 of the wave, banks 0-1 for even k, 2-3 for odd), old elsewhere:
// one DPP move with an identity lane pattern and a row and bank mask.
template <int k>
__device__ __forceinline__ uint32_t put_group(uint32_t old, uint32_t v) {
  return static_cast<uint32_t>(__builtin_amdgcn_update_dpp(static_cast<int>(old), static_cast<int>(v), 0xE4, 1 << (k >> 1), (k & 1) ? 0xC : 0x3, false));
}

struct BlockPos {
  int r, bx, by;
};
// n / d for 0 <= n < 2^24 (wave-uniform: scalar multiply-high and shift).
__device__ __forceinline__ int div_exact(int n, const Divisor& d) {
  return static_cast<int>(__umulhi(static_cast<uint32_t>(n) << 8, d.m) >> d.l);
}
__device__ __forceinline__ bool limited(const MeLevelArgs& a, int r) { return a.kmap && a.level <= 1 && r > 0; }
__device__ __forceinline__ bool moded(const MeLevelArgs& a) { return a.smap && a.level <= 1; }

__device__ __forceinline__ Opening open_block(const MeLevelArgs& a, const BlockPos& q, int lane) {
  Opening o;
  if (limited(a, q.r) || moded(a)) {
    // A superblock's decisions read the level-0 field on it and a ring of 1 block (its inter
    // candidates), and those blocks' searches read the level-1 field on them and 1 more block each
    // way (their coarser centres): a ring of 2 at level 1. Lane 3 (dy + 1) + dx + 1 asks the
    // neighbour (dx, dy) whether this block lies in its ring.
    const int shift = 3 - a.level, n = 1 << shift;  // blocks per superblock side: 8 at level 0, 4 at level 1
    const int reach = a.level == 0 ? 1 : 2;
    const int dx = lane % 3 - 1, dy = lane / 3 - 1;
    const int sx = (q.bx >> shift) + dx, sy = (q.by >> shift) + dy, lx = q.bx & (n - 1), ly = q.by & (n - 1);
    const bool in_x = dx == 0 || (dx < 0 ? lx < reach : lx >= n - reach);
    const bool in_y = dy == 0 || (dy < 0 ? ly < reach : ly >= n - reach);
    if (lane < 9 && in_x && in_y && sx >= 0 && sy >= 0 && sx < a.sb_cols && sy < a.sb_rows) {
      const int s = sy * a.sb_cols + sx;
      o.km = a.kmap ? a.kmap[s] : static_cast<uint32_t>(kMaxRefs);
      o.sm = a.smap ? a.smap[s] : 0u;
    }
  }
  const BPlane cb = as_buffer(a.cur);
  if (lane < 16) o.mine = bload4(cb, corner_offset(cb, 8 * q.by + (lane >> 1), 8 * q.bx + 4 * (lane & 1)), 0);
  return o;
}

// The refinement's per-lane constants: lane (j, g) supplies row g of tile t = j & 3 (t = 2 ty + tx)
// of product column j; off = that row's offset within the block; nhh = -HH as the A operand; ccur =
// the block's own tiles' transforms, accumulated onto.
struct RefineLane {
  int j, g, tx, ty, off;
  half4 nhh;
  f32x4 ccur;
};

// One me_refine step of kStep quarter pels around (qcx, qcy), which it moves to the winner; returns
// the winner's key (cost << 4 | order). Column j' = 16 c + j of product c is position
// o = first + (j' >> 2), tile j' & 3. kStep 2 scores positions 0..8 in three products (9..11 padding,
// reading position 8's rows); kStep 1 scores positions 1..8 in two, the centre's key given
// (centre_key: its cost << 4).
template <int kStep>
__device__ __forceinline__ uint32_t refine_step(const MeLevelArgs& a, const BPlane& rb, const DPlane& ref, int bx, int by, const RefineLane& L,
                                                int& qcx, int& qcy, uint32_t centre_key) {
  constexpr int kProducts = kStep == 2 ? 3 : 2, kFirst = kStep == 2 ? 0 : 1;
  int odx[kProducts], ody[kProducts];
#pragma unroll
  for (int cc = 0; cc < kProducts; cc++) {
    const int o = min(kFirst + 4 * cc + (L.j >> 2), 8);
    odx[cc] = static_cast<int>((kRefDx >> (2 * o)) & 3u) - 1, ody[cc] = static_cast<int>((kRefDy >> (2 * o)) & 3u) - 1;
  }
  // The products' rows first, then the products back to back, then their sums (every partial sum an
  // integer below 2^24: any order is exact). Row offsets: when no position of the step moves a window
  // origin into its clamp (checked once for the wave), the offset is the centre's, plus the lane's
  // row and tile, plus the position's phase and whole-sample carry; otherwise each origin is clamped
  // as at_q() reads.
  int off[kProducts];
  const bool inner = 8 * bx + ((qcx - kStep) >> 2) >= -kBorder && 8 * bx + ((qcx + kStep) >> 2) <= ref.w + kBorder - 8 &&
                     8 * by + ((qcy - kStep) >> 2) >= -kBorder && 8 * by + 4 + ((qcy + kStep) >> 2) <= ref.h + kBorder - 4;
  if (inner) {
    const int centre = corner_offset(rb, 8 * by + (qcy >> 2), 8 * bx + (qcx >> 2)) + L.off;
#pragma unroll
    for (int cc = 0; cc < kProducts; cc++) {
      const int ex = (qcx & 3) + odx[cc] * kStep, ey = (qcy & 3) + ody[cc] * kStep;  // -2 .. 5
      off[cc] = centre + phase_offset((ey & 3) * 4 + (ex & 3), a.phase_bytes) + __mul24(ey >> 2, rb.stride) + (ex >> 2);
    }
  } else {
#pragma unroll
    for (int cc = 0; cc < kProducts; cc++) {
      const int qx = qcx + odx[cc] * kStep, qy = qcy + ody[cc] * kStep;
      const int X = clamp_origin(8 * bx + (qx >> 2), ref.w, 8) + 4 * L.tx;
      const int Y = clamp_origin(8 * by + 4 * L.ty + (qy >> 2), ref.h, 4) + L.g;
      off[cc] = phase_offset((qy & 3) * 4 + (qx & 3), a.phase_bytes) + corner_offset_lane(rb, Y, X);
    }
  }
  uint32_t prow[kProducts];
#pragma unroll
  for (int cc = 0; cc < kProducts; cc++) {
    uint32_t d0, d1;
    bload8(rb, off[cc] & ~3, 0, d0, d1);
    prow[cc] = __builtin_amdgcn_alignbyte(d1, d0, off[cc] & 3);
  }
  f32x4 m[kProducts];
#pragma unroll
  for (int cc = 0; cc < kProducts; cc++) m[cc] = mfma16(L.nhh, row_halves(prow[cc]), L.ccur);
  float sv[kProducts];
#pragma unroll
  for (int cc = 0; cc < kProducts; cc++) {
    float v = abs_sum4(m[cc]);
    v += as_f32(dpp<kDppXor1>(as_u32f(v)));  // the position's 4 tiles (lanes 4o' .. 4o'+3)
    sv[cc] = v + as_f32(dpp<kDppXor2>(as_u32f(v)));
  }
  // A position's 4 rows of coefficients: products 0 and 1 in one pass (rows 0 and 2 of the result
  // hold product 0's totals, rows 1 and 3 product 1's), product 2 in another.
  const float s01 = sum_halves32(sum_rows16(sv[0], sv[1]), sum_rows16(sv[0], sv[1]));
  const int o01 = kFirst + 4 * (L.g & 1) + (L.j >> 2);
  const uint32_t c01 = static_cast<uint32_t>(s01) + (o01 ? kQpelPenalty : 0);
  uint32_t best = (c01 << 4) | static_cast<uint32_t>(o01);
  if constexpr (kProducts == 3) {
    const float s2 = sum_halves32(sum_rows16(sv[2], sv[2]), sum_rows16(sv[2], sv[2]));
    const int o2 = 8 + (L.j >> 2);
    const uint32_t c2 = static_cast<uint32_t>(s2) + kQpelPenalty;
    if (o2 < 9) best = min(best, (c2 << 4) | static_cast<uint32_t>(o2));
  }
  const uint32_t key = scalar_min(wave_min_quads_u32(best), centre_key);  // best is equal over each quad
  const int ob = static_cast<int>(key & 15);
  qcx += (static_cast<int>((kRefDx >> (2 * ob)) & 3u) - 1) * kStep;
  qcy += (static_cast<int>((kRefDy >> (2 * ob)) & 3u) - 1) * kStep;
  return key;
}

// Block (q.bx, q.by) against reference q.r of one launch item, by one wave; win1 / win2 are the
// wave's two LDS windows.
__device__ __forceinline__ void search_block(const MeLevelArgs& a, const BlockPos& q, int lane, uint32_t* win1, uint32_t* win2) {
  const int r = q.r, bx = q.bx, by = q.by;
  const Opening o = open_block(a, q, lane);
  // Searched only near a superblock that keeps reference r.
  if (limited(a, r) && __ballot(o.km > static_cast<uint32_t>(r)) == 0) return;  // whole wave
  // Narrow when no superblock that keeps r and reads the block asks for the full search.
  const bool narrow = moded(a) && __ballot(o.km > static_cast<uint32_t>(r) && ((o.sm >> a.level) & 1u) == 0) == 0;
  const DPlane ref = a.ref[r];
  const BPlane rb = as_buffer(ref);

  // The block, every lane holding every word: scalar loads of its 8 rows (its address is uniform and
  // the kernel writes no source plane), no VALU.
  uint32_t c[16];
  {
    const uint8_t* base = a.cur.p + static_cast<ptrdiff_t>(8 * by) * a.cur.stride + 8 * bx;
#pragma unroll
    for (int i = 0; i < 8; i++) {
      const uint32_t* row = reinterpret_cast<const uint32_t*>(base + static_cast<uint32_t>(i * a.cur.stride));
      c[2 * i] = row[0], c[2 * i + 1] = row[1];
    }
  }

  int p1x = 0, p1y = 0, p2x = 0, p2y = 0;
  uint32_t s1, s2 = 0;  // the searched centres' SADs
  bool two = false;
  {
    // Centres k = 0..4: the coarser MV under the block and its left, right, above and below
    // neighbours (clamped into the coarser field), doubled; k = 5: zero (pack_mv(0, 0) = 0). Uniform
    // addresses, and the coarser field is the previous launch's output: scalar loads. The top level
    // has no coarser field: every centre is zero, one window.
    uint32_t m[6] = {0u, 0u, 0u, 0u, 0u, 0u};
    if (const uint32_t* cf = a.coarser[r]) {
      // Each centre's cell clamped into the coarser field; one 64-bit base, 32-bit byte offsets
      // (scalar loads with an offset register).
      const auto at = [&](int ox, int oy) {
        const int qx = clamp_uniform(bx / 2 + ox, 0, a.pcols - 1), qy = clamp_uniform(by / 2 + oy, 0, a.prows - 1);
        return *reinterpret_cast<const uint32_t*>(reinterpret_cast<const uint8_t*>(cf) + static_cast<uint32_t>(4 * (qy * a.pcols + qx)));
      };
      m[0] = at(0, 0), m[1] = at(-1, 0), m[2] = at(1, 0), m[3] = at(0, -1), m[4] = at(0, 1);
    }
    // Each centre's SAD: lanes 8k .. 8k+7, one block row each (lanes 48.. repeat centre 0, unread).
    // Group k's lanes are row k >> 1 of the wave, banks 0-1 (k even) or 2-3 (k odd): each takes its
    // centre's MV by one masked DPP move.
    uint32_t mv = m[0];
    mv = put_group<1>(mv, m[1]), mv = put_group<2>(mv, m[2]), mv = put_group<3>(mv, m[3]);
    mv = put_group<4>(mv, m[4]), mv = put_group<5>(mv, m[5]);
    const int row = lane & 7;
    const int X = clamp_origin(8 * bx + 2 * mv_x(mv), ref.w, 8), Y = clamp_origin(8 * by + 2 * mv_y(mv), ref.h, 8);
    const int off = corner_offset_lane(rb, Y + row, X);
    uint32_t d0, d1, d2;
    bload12(rb, off & ~3, 0, d0, d1, d2);
    const uint32_t w0 = __builtin_amdgcn_alignbyte(d1, d0, off & 3), w1 = __builtin_amdgcn_alignbyte(d2, d1, off & 3);
    const int baddr = row << 3;  // lane 2 row's byte address
    const uint32_t c0 = static_cast<uint32_t>(__builtin_amdgcn_ds_bpermute(baddr, static_cast<int>(o.mine)));
    const uint32_t c1 = static_cast<uint32_t>(__builtin_amdgcn_ds_bpermute(baddr + 4, static_cast<int>(o.mine)));
    const uint32_t s = oct_sum(sad4(c1, w1, sad4(c0, w0, 0)));
    // me_cell's ranking (src/gpu/inter_common.h:164-186) walks the centres in order, skipping one equal
    // to the current best or runner-up, and keeps the best and the runner-up, each the first strict
    // minimum. With keys SAD << 3 | k that is: the best is the least key (a centre equal to an
    // earlier one has its SAD and a larger k), the runner-up the least key among the centres that
    // differ from the best (one equal to an earlier centre other than the best has a larger key than
    // that one, which is eligible too). Scalar code on the six group sums.
    uint32_t key[6];
#pragma unroll
    for (int k = 0; k < 6; k++) key[k] = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(s), 8 * k)) << 3 | static_cast<uint32_t>(k);
    uint32_t k1 = key[0];
#pragma unroll
    for (int k = 1; k < 6; k++) k1 = scalar_min(k1, key[k]);
    const int b1 = static_cast<int>(k1 & 7);
    uint32_t m1 = m[0];
#pragma unroll
    for (int k = 1; k < 6; k++) m1 = b1 == k ? m[k] : m1;
    uint32_t k2 = 0xffffffffu;
#pragma unroll
    for (int k = 0; k < 6; k++) k2 = m[k] != m1 ? min(k2, key[k]) : k2;
    p1x = 2 * mv_x(m1), p1y = 2 * mv_y(m1), s1 = k1 >> 3;
    two = k2 != 0xffffffffu;
    if (two) {
      const int b2 = static_cast<int>(k2 & 7);
      s2 = k2 >> 3;
      uint32_t m2 = m[0];
#pragma unroll
      for (int k = 1; k < 6; k++) m2 = b2 == k ? m[k] : m2;
      p2x = 2 * mv_x(m2), p2y = 2 * mv_y(m2);
    }
  }

  uint32_t key;
  if (narrow) {
    const int R = kNarrowRange;
    const int X1 = clamp_origin_uniform(8 * bx + p1x - R, ref.w, kNarrowWin), Y1 = clamp_origin_uniform(8 * by + p1y - R, ref.h, kNarrowWin);
    const int X2 = clamp_origin_uniform(8 * bx + p2x - R, ref.w, kNarrowWin), Y2 = clamp_origin_uniform(8 * by + p2y - R, ref.h, kNarrowWin);
    key = wave_min_u32(search_narrow(rb, X1, Y1, X2, Y2, two, c, lane));
  } else {
  // Full searches (me_search, src/gpu/inter_common.h:106-119) from windows staged in LDS.
  const int q = lane >> 3, d = lane & 7;
  const int soff = static_cast<int>(umul24_sv(static_cast<uint32_t>(rb.stride), static_cast<uint32_t>(q))) + 4 * d;
  const int sdst = q * kRowWords + 4 * d;
  const int X1 = clamp_origin_uniform(8 * bx + p1x - kRange, ref.w, kWin), Y1 = clamp_origin_uniform(8 * by + p1y - kRange, ref.h, kWin);
  const int X2 = clamp_origin_uniform(8 * bx + p2x - kRange, ref.w, kWin), Y2 = clamp_origin_uniform(8 * by + p2y - kRange, ref.h, kWin);
  stage_window(rb, corner_offset(rb, Y1, X1) & ~3, soff, d < kCopyWords, win1 + sdst);
  if (two) stage_window(rb, corner_offset(rb, Y2, X2) & ~3, soff, d < kCopyWords, win2 + sdst);
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");

  // A window's first sample sits at byte X & 3 of its aligned row (rows are a multiple of 64 bytes
  // apart, the border 64 samples wide). With one centre, the second window's lanes score the first
  // window again (its keys lose every tie to the first window's). The centres (8, 8) come in with
  // the SADs the ranking found (the same samples: both origins clamp to reads the border makes equal
  // to the reference's per-sample clamping) and no penalty.
  uint32_t tk[12];
  search_lanes(lane, tk);
  key = wave_min_u32(score_pair(tk, win1, two ? win2 : win1, X1 & 3, two ? X2 & 3 : X1 & 3, c));
  key = scalar_min(key, s1 << 10 | (kRange * 17 + kRange));
  if (two) key = scalar_min(key, s2 << 10 | 1u << 9 | (kRange * 17 + kRange));
  }
  const uint32_t best_cost = key >> 10;
  const int ip = static_cast<int>(key & 511);
  const bool second = ((key >> 9) & 1) != 0;
  int mvx = (second ? p2x : p1x) + ip % 17 - kRange, mvy = (second ? p2y : p1y) + ip / 17 - kRange;

  if (a.level0) {
    // me_refine (src/gpu/inter_common.h:135-154) from the integer MV, in quarter pels: the half-pel
    // step scores its 9 positions, the quarter-pel step the 8 around the half-pel winner, whose
    // cost without the step penalty is its SATD from the half-pel step (the same position read
    // the same way).
    const int j = lane & 15, g = lane >> 4, t = j & 3, tx = t & 1, ty = t >> 1;
    half4 hh, nhh;
    hadamard_operands(lane, hh, nhh);
    const f32x4 zero = {0.f, 0.f, 0.f, 0.f};
    // Row g of the block's tile (tx, ty): word 2 (4 ty + g) + tx of the block.
    const uint32_t crow = static_cast<uint32_t>(__builtin_amdgcn_ds_bpermute(4 * (2 * (4 * ty + g) + tx), static_cast<int>(o.mine)));
    const RefineLane L{j, g, tx, ty, static_cast<int>(umul24_sv(static_cast<uint32_t>(rb.stride), static_cast<uint32_t>(4 * ty + g))) + 4 * tx,
                       nhh, mfma16(hh, row_halves(crow), zero)};
    int qcx = 4 * mvx, qcy = 4 * mvy;
    const uint32_t k2 = refine_step<2>(a, rb, ref, bx, by, L, qcx, qcy, 0xffffffffu);
    const uint32_t satd = (k2 >> 4) - ((k2 & 15) ? static_cast<uint32_t>(kQpelPenalty) : 0u);
    refine_step<1>(a, rb, ref, bx, by, L, qcx, qcy, satd << 4);
    mvx = qcx, mvy = qcy;
  }
  if (lane == 0) {
    if (a.cost_out[r]) a.cost_out[r][by * a.cols + bx] = best_cost;
    a.out[r][by * a.cols + bx] = pack_mv(mvx, mvy);
    if (a.searches) count(a.searches, bx + 7 * by + r, two ? 2ull : 1ull);
  }
}

// One workgroup per unit: 4 blocks of one reference, one per wave. kOne: the batch has one item (its
// fields then sit at constant kernel-argument offsets: no per-field address arithmetic).
template <bool kOne>
__global__ void __launch_bounds__(256) k_me_level(Batch<MeLevelArgs> batch) {
  __shared__ uint32_t wins[4 * kWaveLds];
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // wave-uniform: SGPR addressing
  const int total = batch.total();
  const int g = xcd_logical(blockIdx.x, static_cast<int>(gridDim.x));
  if (g >= total) return;  // whole workgroup
  // Units are (group of 4 consecutive blocks of a reference, reference), the reference cycling
  // fastest: the XCD-major numbering then gives every XCD all references of its blocks, however the
  // reference limits thin out the far ones, and a workgroup's 4 waves search neighbouring blocks.
  const int bi = kOne ? 0 : batch_item(batch, g);
  const MeLevelArgs& a = batch.item[bi];
  const int u = g - batch.first[bi], grp = div_exact(u, a.by_nrefs), r = u - grp * a.nrefs;
  const int block = a.row0 * a.cols + 4 * grp + wave;
  if (block >= (a.row0 + a.nrun) * a.cols) return;  // whole wave
  const int by = div_exact(block, a.by_cols);
  search_block(a, BlockPos{r, block - by * a.cols, by}, lane, wins + wave * kWaveLds, wins + wave * kWaveLds + kWinB);
}

}  // namespace

Divisor make_divisor(int d) {
  if (d < 1 || d >= (1 << 16)) throw std::logic_error("divisor outside 1 .. 2^16 - 1");
  int l = 0;
  while ((1 << l) < d) l++;
  const unsigned long long p = 1ull << (24 + l);
  return Divisor{static_cast<uint32_t>((p + static_cast<unsigned long long>(d) - 1) / static_cast<unsigned long long>(d)), l};
}

void add_me_level(Batch<MeLevelArgs>& b, const MeLevelArgs& a) {
  MeLevelArgs x = a;
  if (x.nrun <= 0) x.row0 = 0, x.nrun = a.rows;
  if (x.row0 < 0 || x.row0 + x.nrun > a.rows) throw std::logic_error("motion search rows outside the level");
  const int units = a.nrefs * ((x.nrun * a.cols + 3) / 4);
  if (units <= 0) return;
  // The kernel divides unit numbers (< units) and block numbers (< rows cols + 4) exactly below 2^24.
  if (units >= (1 << 24) || static_cast<long long>(a.rows) * a.cols + 4 >= (1 << 24) || a.cols >= (1 << 16))
    throw std::logic_error("motion search: level too large for the launch's unit numbering");
  x.by_nrefs = make_divisor(a.nrefs), x.by_cols = make_divisor(a.cols);
  b.add(x, units);
}
void launch_me_level(const Batch<MeLevelArgs>& b, hipStream_t s) {
  if (b.total() <= 0) return;
  if (b.n == 1)
    k_me_level<true><<<((b.total() + 7) / 8) * 8, 256, 0, s>>>(b);
  else
    k_me_level<false><<<((b.total() + 7) / 8) * 8, 256, 0, s>>>(b);
}

}  // namespace celv2hip
