// ORACLE — test infrastructure, not product code (see oracle.h). Restates CelV2's CPU decision
// rules; each function cites the reference lines it follows.
#include "oracle.h"

#include <emmintrin.h>

#include <algorithm>
#include <cstring>
#include <stdexcept>

namespace orc {

// ---- primitives ----------------------------------------------------------------------------------

// intra_common.h:32 — DC 1 bit, V/H/Smooth 3, the rest 9.
int mode_bits(int mode) {
  if (mode == 0) return 1;
  return mode < 4 ? 3 : 9;
}

// intra_common.h:131-146 — rows first, then columns; sum of |coefficients|, unnormalised.
int wht4_abs_sum(const int d[16]) {
  int r[16];
  for (int row = 0; row < 4; row++) {
    const int* v = d + 4 * row;
    const int s01 = v[0] + v[1], d01 = v[0] - v[1], s23 = v[2] + v[3], d23 = v[2] - v[3];
    r[4 * row + 0] = s01 + s23;
    r[4 * row + 1] = d01 + d23;
    r[4 * row + 2] = s01 - s23;
    r[4 * row + 3] = d01 - d23;
  }
  int total = 0;
  for (int col = 0; col < 4; col++) {
    const int s01 = r[col] + r[4 + col], d01 = r[col] - r[4 + col];
    const int s23 = r[8 + col] + r[12 + col], d23 = r[8 + col] - r[12 + col];
    const int out[4] = {s01 + s23, d01 + d23, s01 - s23, d01 - d23};
    for (int v : out) total += v < 0 ? -v : v;
  }
  return total;
}

// inter_common.h:12-19 — AV2 EIGHTTAP_REGULAR at quarter-pel phases 0..3.
int qpel_tap(int q, int k) {
  static const int taps[4][8] = {{0, 0, 0, 128, 0, 0, 0, 0},
                                 {0, 2, -14, 110, 38, -10, 2, 0},
                                 {0, 2, -14, 76, 76, -14, 2, 0},
                                 {0, 2, -10, 38, 110, -14, 2, 0}};
  return taps[q][k];
}

// inter_common.h:20-33 — horizontal then vertical 8-tap, one rounding by 2^14, clamped to 0..255;
// rows with a zero vertical tap are skipped (their product is zero).
int qpel_sample(const Plane8& s, int y, int x, int qy, int qx) {
  int acc = 0;
  if (y >= 3 && x >= 3 && y + 5 <= s.h && x + 5 <= s.w) {  // every tap inside: plain reads
    const uint8_t* r = s.px + static_cast<size_t>(y - 3) * s.stride + (x - 3);
    for (int i = 0; i < 8; i++, r += s.stride) {
      const int tv = qpel_tap(qy, i);
      if (!tv) continue;
      int row = 0;
      for (int j = 0; j < 8; j++) row += qpel_tap(qx, j) * r[j];
      acc += tv * row;
    }
    return std::clamp((acc + 8192) >> 14, 0, 255);
  }
  for (int i = 0; i < 8; i++) {
    const int tv = qpel_tap(qy, i);
    if (!tv) continue;
    int row = 0;
    for (int j = 0; j < 8; j++) row += qpel_tap(qx, j) * s.at(y + i - 3, x + j - 3);
    acc += tv * row;
  }
  acc = (acc + 8192) >> 14;
  return std::clamp(acc, 0, 255);
}

// inter_common.h:91-94 — rounded 2x2 mean.
int half_sample(const Plane8& s, int y, int x) {
  return (s.at(2 * y, 2 * x) + s.at(2 * y, 2 * x + 1) + s.at(2 * y + 1, 2 * x) + s.at(2 * y + 1, 2 * x + 1) + 2) >> 2;
}

// Direct addressing of the h x w area at (y, x) where it lies inside the plane, else null: clamped
// reads there are plain reads. Only a speed path; the clamped form below each use covers the rest.
static const uint8_t* direct(const Plane8& p, int y, int x, int h, int w) {
  return y >= 0 && x >= 0 && y + h <= p.h && x + w <= p.w ? p.px + static_cast<size_t>(y) * p.stride + x : nullptr;
}
// The same for the area read by atq(y + i, x + j, qy, qx): phase plane (qy & 3) * 4 + (qx & 3).
static const uint8_t* direct_q(const Plane8& p, int y, int x, int h, int w, int qy, int qx) {
  Plane8 q = p;
  q.px = p.px + static_cast<size_t>((qy & 3) * 4 + (qx & 3)) * p.phase;
  return direct(q, y + (qy >> 2), x + (qx >> 2), h, w);
}

// inter_common.h:96-104.
int sad_8x8(const Plane8& cur, int x, int y, const Plane8& ref, int mvx, int mvy) {
  int sad = 0;
  const uint8_t* a = direct(cur, y, x, 8, 8);
  const uint8_t* b = direct(ref, y + mvy, x + mvx, 8, 8);
  if (a && b) {  // two 8-byte rows per psadbw (SSE2): the sum of |a - b| over the 64 samples
    __m128i acc = _mm_setzero_si128();
    for (int i = 0; i < 8; i += 2, a += 2 * cur.stride, b += 2 * ref.stride) {
      const __m128i va = _mm_unpacklo_epi64(_mm_loadl_epi64(reinterpret_cast<const __m128i*>(a)),
                                            _mm_loadl_epi64(reinterpret_cast<const __m128i*>(a + cur.stride)));
      const __m128i vb = _mm_unpacklo_epi64(_mm_loadl_epi64(reinterpret_cast<const __m128i*>(b)),
                                            _mm_loadl_epi64(reinterpret_cast<const __m128i*>(b + ref.stride)));
      acc = _mm_add_epi64(acc, _mm_sad_epu8(va, vb));
    }
    return _mm_cvtsi128_si32(acc) + _mm_cvtsi128_si32(_mm_unpackhi_epi64(acc, acc));
  }
  for (int i = 0; i < 8; i++)
    for (int j = 0; j < 8; j++) sad += std::abs(cur.at(y + i, x + j) - ref.at(y + i + mvy, x + j + mvx));
  return sad;
}

// inter_common.h:106-119 — +-range around (px, py), dy-major, first strict minimum of
// SAD + 4 * (|dx| + |dy|).
int full_search(const Plane8& cur, const Plane8& ref, int bx, int by, int px, int py, int range, int* mvx, int* mvy) {
  int best = 0x7fffffff;
  *mvx = px, *mvy = py;
  for (int dy = -range; dy <= range; dy++)
    for (int dx = -range; dx <= range; dx++) {
      const int c = sad_8x8(cur, 8 * bx, 8 * by, ref, px + dx, py + dy) + kPelPenalty * (std::abs(dx) + std::abs(dy));
      if (c < best) best = c, *mvx = px + dx, *mvy = py + dy;
    }
  return best;
}

// inter_common.h:121-133 — SATD of an 8x8 block against ref at a quarter-pel MV.
int satd_8x8_q(const Plane8& cur, int x, int y, const Plane8& ref, int qmx, int qmy) {
  return static_cast<int>(inter_satd(cur, ref, x, y, 8, 8, qmx, qmy));
}

// inter_common.h:135-154 — half-pel step then quarter-pel step; the centre first, each of the 8
// neighbours (dy-major) charged 16 and taken only if strictly cheaper.
void refine_qpel(const Plane8& cur, const Plane8& ref, int bx, int by, int* mvx, int* mvy) {
  int cx = 4 * *mvx, cy = 4 * *mvy;
  for (int step = 2; step >= 1; step--) {
    int best = satd_8x8_q(cur, 8 * bx, 8 * by, ref, cx, cy), nx = cx, ny = cy;
    for (int dy = -1; dy <= 1; dy++)
      for (int dx = -1; dx <= 1; dx++) {
        if (!dx && !dy) continue;
        const int c = satd_8x8_q(cur, 8 * bx, 8 * by, ref, cx + step * dx, cy + step * dy) + kQpelPenalty;
        if (c < best) best = c, nx = cx + step * dx, ny = cy + step * dy;
      }
    cx = nx, cy = ny;
  }
  *mvx = cx, *mvy = cy;
}

// inter_common.h:156-193 — six candidate centres (coarser MV x2 under the block, its left, right,
// above and below neighbours, then zero), each distinct one ranked by SAD; the best and the
// runner-up are both searched and the runner-up's result kept only if strictly cheaper.
int search_cell(const Plane8& cur, const Plane8& ref, int bx, int by, const int* pmx, const int* pmy, int pcols,
                int prows, int* mvx, int* mvy, bool /*level0*/, int range) {
  int c1x = 0, c1y = 0, c2x = 0, c2y = 0;
  bool have2 = false;
  if (pmx) {
    static const int offx[5] = {0, -1, 1, 0, 0}, offy[5] = {0, 0, 0, -1, 1};
    int s1 = 0x7fffffff, s2 = 0x7fffffff;
    for (int k = 0; k < 6; k++) {
      int cx = 0, cy = 0;
      if (k < 5) {
        const int qx = std::clamp(bx / 2 + offx[k], 0, pcols - 1), qy = std::clamp(by / 2 + offy[k], 0, prows - 1);
        cx = 2 * pmx[qy * pcols + qx], cy = 2 * pmy[qy * pcols + qx];
      }
      const bool dup1 = k > 0 && cx == c1x && cy == c1y, dup2 = have2 && cx == c2x && cy == c2y;
      if (dup1 || dup2) continue;
      const int s = sad_8x8(cur, 8 * bx, 8 * by, ref, cx, cy);
      if (s < s1) {
        if (k > 0) s2 = s1, c2x = c1x, c2y = c1y, have2 = true;
        s1 = s, c1x = cx, c1y = cy;
      } else if (s < s2) {
        s2 = s, c2x = cx, c2y = cy, have2 = true;
      }
    }
  }
  const int cost1 = full_search(cur, ref, bx, by, c1x, c1y, range, mvx, mvy);
  if (have2) {
    int mx, my;
    const int cost2 = full_search(cur, ref, bx, by, c2x, c2y, range, &mx, &my);
    if (cost2 < cost1) {
      *mvx = mx, *mvy = my;
      return cost2;
    }
  }
  return cost1;
}

// intra_common.h:47-56 — source neighbours of the block at (x, y); 128 outside the frame's top/left.
int edge_above(const Plane8& s, int x, int y, int k) {
  if (y == 0) return 128;
  if (k < 0 && x == 0) return s.at(y - 1, x);
  return s.at(y - 1, x + k);
}
int edge_left(const Plane8& s, int x, int y, int k) {
  if (x == 0) return 128;
  if (k < 0 && y == 0) return s.at(y, x - 1);
  return s.at(y + k, x - 1);
}

static int derivative(int a) { return a == 23 ? 151 : a == 45 ? 64 : 27; }  // intra_common.h:60

// intra_common.h:61-85 — z1 (above), z3 (left) and z2 (both) directional prediction, 1/64-pel
// positions with 5-bit interpolation, no edge filter or upsampling.
int directional(const Plane8& s, int x, int y, int w, int h, int angle, int i, int j) {
  const int last = w + h - 1;
  if (angle < 90) {
    const int pos = (i + 1) * derivative(angle), base = (pos >> 6) + j, f = (pos & 63) >> 1;
    const int a0 = edge_above(s, x, y, std::min(base, last)), a1 = edge_above(s, x, y, std::min(base + 1, last));
    return (a0 * (32 - f) + a1 * f + 16) >> 5;
  }
  if (angle > 180) {
    const int pos = (j + 1) * derivative(270 - angle), base = (pos >> 6) + i, f = (pos & 63) >> 1;
    const int l0 = edge_left(s, x, y, std::min(base, last)), l1 = edge_left(s, x, y, std::min(base + 1, last));
    return (l0 * (32 - f) + l1 * f + 16) >> 5;
  }
  const int px = 64 * j - (i + 1) * derivative(180 - angle);
  const int bx = px >= 0 ? px >> 6 : -((-px + 63) >> 6);  // floor division by 64
  if (bx >= -1) {
    const int f = (px - 64 * bx) >> 1;
    return (edge_above(s, x, y, bx) * (32 - f) + edge_above(s, x, y, bx + 1) * f + 16) >> 5;
  }
  const int py = 64 * i - (j + 1) * derivative(angle - 90);
  const int by = py >= 0 ? py >> 6 : -((-py + 63) >> 6);
  const int f = (py - 64 * by) >> 1;
  const int b0 = std::max(by, -1), b1 = std::max(by + 1, -1);
  return (edge_left(s, x, y, b0) * (32 - f) + edge_left(s, x, y, b1) * f + 16) >> 5;
}

// intra_common.h:89-118 — modes in IntraMode order; DC (mode 0) is the caller's (dc_mean).
int intra_predict(const Plane8& s, int x, int y, int w, int h, int mode, int i, int j) {
  const bool up = y > 0, lf = x > 0;
  switch (mode) {
    case 1: return up ? s.at(y - 1, x + j) : 128;
    case 2: return lf ? s.at(y + i, x - 1) : 128;
    case 3: {
      const int a = up ? s.at(y - 1, x + j) : 128, l = lf ? s.at(y + i, x - 1) : 128;
      const int tr = up ? s.at(y - 1, x + w - 1) : 128, bl = lf ? s.at(y + h - 1, x - 1) : 128;
      const int den = w + h + 2;
      return ((h - i) * a + (i + 1) * bl + (w - j) * l + (j + 1) * tr + den / 2) / den;
    }
    case 4: return ((h - 1 - i) * edge_above(s, x, y, j) + (i + 1) * edge_left(s, x, y, h - 1) + h / 2) / h;
    case 5: return ((w - 1 - j) * edge_left(s, x, y, i) + (j + 1) * edge_above(s, x, y, w - 1) + w / 2) / w;
    case 6: {
      const int a = edge_above(s, x, y, j), l = edge_left(s, x, y, i), tl = edge_above(s, x, y, -1);
      const int base = a + l - tl;
      const int pa = std::abs(base - a), pl = std::abs(base - l), ptl = std::abs(base - tl);
      if (pl <= pa && pl <= ptl) return l;
      return pa <= ptl ? a : tl;
    }
    default: break;
  }
  if (mode >= 7) {
    static const int angles[6] = {45, 67, 113, 135, 157, 203};
    return directional(s, x, y, w, h, angles[mode - 7], i, j);
  }
  return -1;
}

// intra_common.h:120-128.
int dc_mean(const Plane8& s, int x, int y, int w, int h) {
  int sum = 0, n = 0;
  if (y > 0)
    for (int j = 0; j < w; j++) sum += s.at(y - 1, x + j), n++;
  if (x > 0)
    for (int i = 0; i < h; i++) sum += s.at(y + i, x - 1), n++;
  return n ? (sum + n / 2) / n : 128;
}

// intra_common.h:149-163.
int64_t intra_satd(const Plane8& s, int x, int y, int w, int h, int mode) {
  const int dc = mode == 0 ? dc_mean(s, x, y, w, h) : 0;
  int64_t total = 0;
  for (int ty = 0; ty < h; ty += 4)
    for (int tx = 0; tx < w; tx += 4) {
      int d[16];
      for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) {
          const int p = mode == 0 ? dc : intra_predict(s, x, y, w, h, mode, ty + i, tx + j);
          d[4 * i + j] = s.at(y + ty + i, x + tx + j) - p;
        }
      total += wht4_abs_sum(d);
    }
  return total;
}

// inter_common.h:222-235.
int64_t inter_satd(const Plane8& cur, const Plane8& ref, int x, int y, int w, int h, int mvx, int mvy) {
  int64_t total = 0;
  const uint8_t* a = direct(cur, y, x, h, w);
  const uint8_t* b = direct_q(ref, y, x, h, w, mvy, mvx);
  if (a && b) {
    for (int ty = 0; ty < h; ty += 4)
      for (int tx = 0; tx < w; tx += 4) {
        int d[16];
        for (int i = 0; i < 4; i++)
          for (int j = 0; j < 4; j++)
            d[4 * i + j] = a[static_cast<size_t>(ty + i) * cur.stride + tx + j] - b[static_cast<size_t>(ty + i) * ref.stride + tx + j];
        total += wht4_abs_sum(d);
      }
    return total;
  }
  for (int ty = 0; ty < h; ty += 4)
    for (int tx = 0; tx < w; tx += 4) {
      int d[16];
      for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++)
          d[4 * i + j] = cur.at(y + ty + i, x + tx + j) - ref.atq(y + ty + i, x + tx + j, mvy, mvx);
      total += wht4_abs_sum(d);
    }
  return total;
}

// inter_common.h:205-220 — the two predictions averaged with +1 rounding.
int64_t compound_satd(const Plane8& cur, const Plane8& r0, int mx0, int my0, const Plane8& r1, int mx1, int my1,
                      int x, int y, int w, int h) {
  int64_t total = 0;
  const uint8_t* a = direct(cur, y, x, h, w);
  const uint8_t* p0 = direct_q(r0, y, x, h, w, my0, mx0);
  const uint8_t* p1 = direct_q(r1, y, x, h, w, my1, mx1);
  if (a && p0 && p1) {
    for (int ty = 0; ty < h; ty += 4)
      for (int tx = 0; tx < w; tx += 4) {
        int d[16];
        for (int i = 0; i < 4; i++)
          for (int j = 0; j < 4; j++) {
            const size_t o = static_cast<size_t>(ty + i) * r0.stride + tx + j;
            d[4 * i + j] = a[static_cast<size_t>(ty + i) * cur.stride + tx + j] - ((p0[o] + p1[static_cast<size_t>(ty + i) * r1.stride + tx + j] + 1) >> 1);
          }
        total += wht4_abs_sum(d);
      }
    return total;
  }
  for (int ty = 0; ty < h; ty += 4)
    for (int tx = 0; tx < w; tx += 4) {
      int d[16];
      for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) {
          const int yy = y + ty + i, xx = x + tx + j;
          d[4 * i + j] = cur.at(yy, xx) - ((r0.atq(yy, xx, my0, mx0) + r1.atq(yy, xx, my1, mx1) + 1) >> 1);
        }
      total += wht4_abs_sum(d);
    }
  return total;
}

// inter_common.h:195-204.
bool pair_allowed(int a, int b, int nrefs) { return a != b && std::min(a, b) <= std::min(nrefs - 2, 2); }
int pair_bits(int a, int b) { return 20 + a + b; }

// inter_common.h:57-78 — candidates: the block's 8x8 cells in raster order, then left, above,
// right, below, above-left, above-right, below-left, below-right; field coordinates clamped.
int cells_covered(int x, int y, int w, int h) { return ((x + w - 1) / 8 - x / 8 + 1) * ((y + h - 1) / 8 - y / 8 + 1); }
int cand_count(int x, int y, int w, int h) { return cells_covered(x, y, w, h) + 8; }
int cand_cell(int x, int y, int w, int h, int k, int cols, int rows) {
  const int n = cells_covered(x, y, w, h), across = (x + w - 1) / 8 - x / 8 + 1;
  int cx, cy;
  if (k < n) {
    cx = x / 8 + k % across, cy = y / 8 + k / across;
  } else {
    const int m = k - n;
    const int left = x / 8 - 1, right = (x + w) / 8, top = y / 8 - 1, bottom = (y + h) / 8;
    static const int8_t hx[8] = {-1, 0, 1, 0, -1, 1, -1, 1};  // -1 left, 0 same, 1 right
    static const int8_t vy[8] = {0, -1, 0, 1, -1, -1, 1, 1};  // -1 above, 0 same, 1 below
    cx = hx[m] < 0 ? left : hx[m] > 0 ? right : x / 8;
    cy = vy[m] < 0 ? top : vy[m] > 0 ? bottom : y / 8;
  }
  cx = std::clamp(cx, 0, cols - 1);
  cy = std::clamp(cy, 0, rows - 1);
  return cy * cols + cx;
}
// inter_common.h:80-89.
bool cand_repeats(const int* fx, const int* fy, int cols, int rows, int x, int y, int w, int h, int k) {
  const int i = cand_cell(x, y, w, h, k, cols, rows);
  for (int j = 0; j < k; j++) {
    const int c = cand_cell(x, y, w, h, j, cols, rows);
    if (fx[c] == fx[i] && fy[c] == fy[i]) return true;
  }
  return false;
}

// sb_blocks.h:1-37 — 85 squares (64, 32, 16, 8; raster within a level), each whole then its
// top, bottom, left and right halves: block b = square * 5 + kind.
int level_first(int level) { return ((1 << (2 * level)) - 1) / 3; }
int square_id(int level, int qx, int qy) { return level_first(level) + (qy << level) + qx; }
void block_geom(int b, int* x, int* y, int* w, int* h) {
  const int sq = b / 5, kind = b % 5;
  int level = 0;
  while (level < 3 && sq >= level_first(level + 1)) level++;
  const int i = sq - level_first(level), n = 1 << level, size = 64 >> level;
  const int sx = (i % n) * size, sy = (i / n) * size;
  *x = sx, *y = sy, *w = size, *h = size;
  if (kind == 1 || kind == 2) *h = size / 2, *y = sy + (kind == 2 ? size / 2 : 0);
  if (kind == 3 || kind == 4) *w = size / 2, *x = sx + (kind == 4 ? size / 2 : 0);
}
int block_index(int x, int y, int w, int h) {
  const int size = std::max(w, h);
  int level = 0;
  while ((64 >> level) > size) level++;
  const int sx = x / size * size, sy = y / size * size;
  int kind = 0;
  if (w > h) kind = y == sy ? 1 : 2;
  if (w < h) kind = x == sx ? 3 : 4;
  return square_id(level, sx / size, sy / size) * 5 + kind;
}

// ---- frames, pyramids, motion fields ----------------------------------------------------------------

Luma make_luma(int width, int height, const uint8_t* y) {
  Luma f;
  f.width = width, f.height = height, f.pw = (width + 7) & ~7, f.ph = (height + 7) & ~7;
  f.y.assign(y, y + static_cast<size_t>(f.pw) * f.ph);
  return f;
}

void round_plane(const uint16_t* in, uint8_t* out, size_t n, int bit_depth) {
  const int shift = bit_depth - 8;
  for (size_t i = 0; i < n; i++)
    out[i] = shift <= 0 ? static_cast<uint8_t>(in[i]) : static_cast<uint8_t>(std::min(255, (in[i] + (1 << (shift - 1))) >> shift));
}

// kernels_ref.cpp:213-238 — level 0 the frame (a reference's followed by phases 1..15 at the
// clamped positions, one plane apart), levels 1..4 halved (floor) by half_sample.
Pyramid build_pyramid(const Luma& f, bool phases) {
  Pyramid p;
  p.px.resize(kLevels);
  const size_t n = f.y.size();
  p.px[0] = f.y;
  if (phases) {
    // qpel_sample at every position, computed separably: the horizontal 8-tap sum of each clamped
    // row at each horizontal phase once (the same integers qpel_sample forms for that row), then the
    // vertical taps over the clamped rows' sums and the one rounding.
    p.px[0].resize(n * kPhases);
    const Plane8 src{f.y.data(), f.pw, f.pw, f.ph, 0};
    std::vector<int> rows(n);
    for (int qx = 0; qx < 4; qx++) {
      for (int y = 0; y < f.ph; y++)
        for (int x = 0; x < f.pw; x++) {
          int r = 0;
          for (int j = 0; j < 8; j++) r += qpel_tap(qx, j) * src.at(y, x + j - 3);
          rows[static_cast<size_t>(y) * f.pw + x] = r;
        }
      for (int qy = 0; qy < 4; qy++) {
        const int k = qy * 4 + qx;
        if (k == 0) continue;
        uint8_t* out = p.px[0].data() + k * n;
        for (int y = 0; y < f.ph; y++)
          for (int x = 0; x < f.pw; x++) {
            int acc = 0;
            for (int i = 0; i < 8; i++) {
              const int tv = qpel_tap(qy, i);
              if (tv) acc += tv * rows[static_cast<size_t>(std::clamp(y + i - 3, 0, f.ph - 1)) * f.pw + x];
            }
            out[static_cast<size_t>(y) * f.pw + x] = static_cast<uint8_t>(std::clamp((acc + 8192) >> 14, 0, 255));
          }
      }
    }
  }
  p.plane[0] = Plane8{p.px[0].data(), f.pw, f.pw, f.ph, phases ? n : 0};
  for (int l = 1; l < kLevels; l++) {
    const Plane8& up = p.plane[l - 1];
    const int w = up.w / 2, h = up.h / 2;
    p.px[l].resize(static_cast<size_t>(w) * h);
    for (int y = 0; y < h; y++)
      for (int x = 0; x < w; x++) p.px[l][static_cast<size_t>(y) * w + x] = static_cast<uint8_t>(half_sample(up, y, x));
    p.plane[l] = Plane8{p.px[l].data(), w, w, h, 0};
  }
  return p;
}

// kernels_ref.cpp:240-251.
// The pruned form's restrictions for the decision in progress on this thread (null: none). Reference
// limits: superblock s is decided with its nearest g_ref_limits[s] references. Search modes: bit L of
// g_search_modes[s] (L = 0, 1) asks for the narrow full search (+-kNarrowRange) at level L. A level-L
// block is read by the superblock it lies in and by each neighbour whose ring holds it (a superblock's
// inter candidates read the level-0 field one block beyond it, and those blocks' centres the level-1
// field one block beyond them: a ring of 1 block at level 0, 2 at level 1); it is searched narrow for
// reference r when every such superblock that keeps r asks for it.
thread_local const uint8_t* g_ref_limits = nullptr;
thread_local const uint8_t* g_search_modes = nullptr;
thread_local int g_sb_cols = 0, g_sb_rows = 0;

bool narrow_search(int level, int bx, int by, int r) {
  if (!g_search_modes || level > 1) return false;
  const int shift = 3 - level, n = 1 << shift, reach = level == 0 ? 1 : 2;
  const int lx = bx & (n - 1), ly = by & (n - 1);
  for (int dy = -1; dy <= 1; dy++)
    for (int dx = -1; dx <= 1; dx++) {
      const int sx = (bx >> shift) + dx, sy = (by >> shift) + dy;
      if (sx < 0 || sy < 0 || sx >= g_sb_cols || sy >= g_sb_rows) continue;
      const bool in_x = dx == 0 || (dx < 0 ? lx < reach : lx >= n - reach);
      const bool in_y = dy == 0 || (dy < 0 ? ly < reach : ly >= n - reach);
      if (!in_x || !in_y) continue;
      const int s = sy * g_sb_cols + sx;
      const int k = g_ref_limits ? g_ref_limits[s] : kMaxRefs;
      if (k > r && !((g_search_modes[s] >> level) & 1)) return false;
    }
  return true;
}

void search_rows(const Pyramid& cur, const Pyramid& ref, int level, const Field* coarser, int row0, int row1,
                 Field& out, int ref_index) {
  for (int by = row0; by < row1; by++)
    for (int bx = 0; bx < out.cols; bx++) {
      const size_t i = static_cast<size_t>(by) * out.cols + bx;
      search_cell(cur.plane[level], ref.plane[level], bx, by, coarser ? coarser->mvx.data() : nullptr,
                  coarser ? coarser->mvy.data() : nullptr, coarser ? coarser->cols : 0, coarser ? coarser->rows : 0,
                  &out.mvx[i], &out.mvy[i], level == 0, narrow_search(level, bx, by, ref_index) ? kNarrowRange : kRange);
      if (level == 0) refine_qpel(cur.plane[0], ref.plane[0], bx, by, &out.mvx[i], &out.mvy[i]);
    }
}

void search_all(const Pyramid& cur, const Pyramid& ref, std::vector<Field>& lv, int ref_index) {
  lv.assign(kLevels, Field{});
  for (int l = kLevels - 1; l >= 0; l--) {
    Field& o = lv[l];
    o.cols = (cur.plane[l].w + 7) / 8, o.rows = (cur.plane[l].h + 7) / 8;
    o.mvx.assign(static_cast<size_t>(o.cols) * o.rows, 0), o.mvy = o.mvx;
    search_rows(cur, ref, l, l == kLevels - 1 ? nullptr : &lv[l + 1], 0, o.rows, o, ref_index);
  }
}

// ---- per-block choice and partition DP (kernels_ref.cpp:26-198) --------------------------------------

// ORC_PLANT (test builds only, tests/test_oracle_choice.py): a deliberate mistake in the per-block
// choice below, 1..8 as listed there; 0, the default, is the rule itself.
#ifndef ORC_PLANT
#define ORC_PLANT 0
#endif

namespace {

struct Pick {
  int64_t cost;
  int mode;  // intra 0..12, kModes = inter
  int mvx, mvy, ref;
  int ref2 = 255, mvx2 = 0, mvy2 = 0;
};

struct Scores {
  Pick intra{INT64_MAX, 0, 0, 0, 0}, inter{INT64_MAX, kModes, 0, 0, 0};
  Pick per_ref[kMaxRefs];
  int nrefs = 0;
};

}  // namespace

// Intra mode groups per superblock for the decision in progress on this thread (the pruned form's
// restriction; null: every mode): group 0 = modes 0..2 (always), 1 = 3..6, 2 = 7..9, 3 = 10..12.
thread_local const uint8_t* g_intra_masks = nullptr;
thread_local int g_intra_mask_cols = 0;

namespace {

bool mode_allowed(int x, int y, int m) {
  if (!g_intra_masks) return true;
  const int g = m < 3 ? 0 : m < 7 ? 1 : m < 10 ? 2 : 3;
  return ((g_intra_masks[(y / 64) * g_intra_mask_cols + x / 64] | 1) >> g) & 1;
}

// kernels_ref.cpp:41-79 — intra modes in order; each reference's candidates in order (repeats
// skipped); the best reference, the runner-up and their compound (first strict minimum throughout).
Scores score_block(const Plane8& s, const RefSet& rs, int x, int y, int w, int h, int64_t lam, bool with_intra) {
  Scores sc;
  const int n = rs.size();
  if (with_intra)
    for (int m = 0; m < kModes; m++) {
      if (!mode_allowed(x, y, m)) continue;
      const int64_t c = intra_satd(s, x, y, w, h, m) * 256 + lam * (mode_bits(m) + (n > 0 ? kIntraOnInter - (ORC_PLANT == 1) : 0));
      if (c < sc.intra.cost) sc.intra = Pick{c, m, 0, 0, 0};
    }
  int best = -1, runner = -1;
  for (int r = 0; r < n; r++) {
    const Field& f = *rs.field[r];
    sc.per_ref[r] = Pick{INT64_MAX, kModes, 0, 0, r};
    const int count = cand_count(x, y, w, h);
    for (int k = 0; k < count; k++) {
      if (cand_repeats(f.mvx.data(), f.mvy.data(), f.cols, f.rows, x, y, w, h, k)) continue;
      const int c = cand_cell(x, y, w, h, k, f.cols, f.rows);
      const int64_t cost = inter_satd(s, *rs.plane[r], x, y, w, h, f.mvx[c], f.mvy[c]) * 256 +
                           lam * (kNewmvBits + (ORC_PLANT == 2 ? 0 : r) + (ORC_PLANT == 3));
      if (cost < sc.per_ref[r].cost || (ORC_PLANT == 4 && cost == sc.per_ref[r].cost))
        sc.per_ref[r] = Pick{cost, kModes, f.mvx[c], f.mvy[c], r};
    }
    if (best < 0 || sc.per_ref[r].cost < sc.per_ref[best].cost) best = r;
  }
  for (int r = 0; r < n; r++)
    if (r != best && (runner < 0 || sc.per_ref[r].cost < sc.per_ref[runner].cost || (ORC_PLANT == 5 && sc.per_ref[r].cost == sc.per_ref[runner].cost)))
      runner = r;
  if (best >= 0) sc.inter = sc.per_ref[best];
  sc.nrefs = n;
  if (runner >= 0 && pair_allowed(best, runner, n)) {
    const Pick &p = sc.per_ref[best], &q = sc.per_ref[runner];
    const int64_t c = compound_satd(s, *rs.plane[best], p.mvx, p.mvy, *rs.plane[runner], q.mvx, q.mvy, x, y, w, h) * 256 +
                      lam * (pair_bits(best, runner) + (ORC_PLANT == 6 ? runner : 0));
    if (c < sc.inter.cost) sc.inter = Pick{c, kModes, p.mvx, p.mvy, best, runner, q.mvx, q.mvy};
  }
  return sc;
}

// kernels_ref.cpp:87-104 — the block's choice (inter only when strictly cheaper) and, for 64x64
// superblocks, its candidate record.
Pick choose(const Plane8& s, const RefSet& rs, int x, int y, int w, int h, int64_t lam, std::vector<Cands>* rec,
            int ox, int oy, bool with_intra) {
  const Scores sc = score_block(s, rs, x, y, w, h, lam, with_intra);
  if (rec) {
    Cands& c = (*rec)[block_index(x - ox, y - oy, w, h)];
    c.inter_cost = sc.inter.cost, c.intra_cost = sc.intra.cost, c.intra = static_cast<uint8_t>(sc.intra.mode);
    c.inter.mv_col = static_cast<int16_t>(2 * sc.inter.mvx), c.inter.mv_row = static_cast<int16_t>(2 * sc.inter.mvy);
    c.inter.ref = static_cast<uint8_t>(sc.inter.ref), c.inter.ref2 = static_cast<uint8_t>(sc.inter.ref2);
    c.inter.mv2_col = static_cast<int16_t>(2 * sc.inter.mvx2), c.inter.mv2_row = static_cast<int16_t>(2 * sc.inter.mvy2);
    for (int r = 0; r < sc.nrefs; r++)
      if (sc.per_ref[r].cost != INT64_MAX) {
        c.ref_cost[r] = static_cast<uint32_t>(std::min<int64_t>(sc.per_ref[r].cost >> (ORC_PLANT == 7 ? 9 : 8), UINT32_MAX - 1));
        c.ref_mv_col[r] = static_cast<int16_t>((ORC_PLANT == 8 ? 1 : 2) * sc.per_ref[r].mvx);
        c.ref_mv_row[r] = static_cast<int16_t>(2 * sc.per_ref[r].mvy);
      }
  }
  return sc.inter.cost < sc.intra.cost ? sc.inter : sc.intra;
}

// kernels_ref.cpp:106-122 — a leaf's BlockDecision (quarter pels to 1/8).
void add_leaf(Sb& sb, int w, int h, const Pick& p) {
  sb.tree.push_back(kNone);
  Block b;
  b.w4 = static_cast<uint8_t>(w / 4), b.h4 = static_cast<uint8_t>(h / 4);
  b.inter = p.mode == kModes;
  b.intra = b.inter ? 0 : static_cast<uint8_t>(p.mode);
  if (b.inter) {
    b.mv_col = static_cast<int16_t>(2 * p.mvx), b.mv_row = static_cast<int16_t>(2 * p.mvy);
    b.ref = static_cast<uint8_t>(p.ref);
    if (p.ref2 != 255) {
      b.ref2 = static_cast<uint8_t>(p.ref2);
      b.mv2_col = static_cast<int16_t>(2 * p.mvx2), b.mv2_row = static_cast<int16_t>(2 * p.mvy2);
    }
  }
  sb.blocks.push_back(b);
}

void append(Sb& out, const Sb& sub) {
  out.tree.insert(out.tree.end(), sub.tree.begin(), sub.tree.end());
  out.blocks.insert(out.blocks.end(), sub.blocks.begin(), sub.blocks.end());
}

}  // namespace

// kernels_ref.cpp:124-198 — bottom-up; NONE, HORZ, VERT, SPLIT with the first strict minimum.
// Above 32x32 only inter blocks (inter frames, up to 64x64) compete with SPLIT; 8x8 never splits.
int64_t decide_square(const Plane8& s, const RefSet& rs, int x, int y, int size, int64_t lam, Sb& out,
                      std::vector<Cands>* rec, int ox, int oy) {
  const int half = size / 2;
  const int64_t part = lam * kSplitBits;
  const bool big = size > kIntraMaxSize;
  Pick none{}, h0{}, h1{}, v0{}, v1{};
  int64_t best = INT64_MAX;
  int choice = 3;
  const bool leaves = !big || (rs.size() > 0 && size <= kInterMaxSize);
  if (leaves) {
    none = choose(s, rs, x, y, size, size, lam, rec, ox, oy, !big);
    h0 = choose(s, rs, x, y, size, half, lam, rec, ox, oy, !big);
    h1 = choose(s, rs, x, y + half, size, half, lam, rec, ox, oy, !big);
    v0 = choose(s, rs, x, y, half, size, lam, rec, ox, oy, !big);
    v1 = choose(s, rs, x + half, y, half, size, lam, rec, ox, oy, !big);
    const int64_t c_h = part + h0.cost + h1.cost, c_v = part + v0.cost + v1.cost;
    best = none.cost, choice = 0;
    if (c_h < best) best = c_h, choice = 1;
    if (c_v < best) best = c_v, choice = 2;
  }
  Sb split;
  if (size > 8) {
    int64_t c_s = part;
    for (int q = 0; q < 4; q++) c_s += decide_square(s, rs, x + (q & 1) * half, y + (q >> 1) * half, half, lam, split, rec, ox, oy);
    if (c_s < best) best = c_s, choice = 3;
  }
  switch (choice) {
    case 0: add_leaf(out, size, size, none); break;
    case 1: out.tree.push_back(kHorz), add_leaf(out, size, half, h0), add_leaf(out, size, half, h1); break;
    case 2: out.tree.push_back(kVert), add_leaf(out, half, size, v0), add_leaf(out, half, size, v1); break;
    default: out.tree.push_back(kSplit), append(out, split); break;
  }
  return best;
}

// kernels_ref.cpp:253-270 — candidates are recorded only for 64x64 superblocks, pre-filled with each
// block's size and the inter flag.
Sb decide_sb(const Luma& src, const RefSet& rs, int sb_col, int sb_row, int sb_size, int64_t lam) {
  const Plane8 s{src.y.data(), src.pw, src.pw, src.ph, 0};
  Sb sb;
  if (sb_size == 64) {
    sb.cands.resize(kBlocks);
    for (int b = 0; b < kBlocks; b++) {
      int x, y, w, h;
      block_geom(b, &x, &y, &w, &h);
      sb.cands[b].inter.w4 = static_cast<uint8_t>(w / 4), sb.cands[b].inter.h4 = static_cast<uint8_t>(h / 4);
      sb.cands[b].inter.inter = true;
    }
  }
  const int ox = sb_col * sb_size, oy = sb_row * sb_size;
  RefSet near = rs;  // the superblock's nearest references under a limit
  if (g_ref_limits) {
    const size_t k = std::min<size_t>(near.plane.size(), g_ref_limits[sb_row * g_sb_cols + sb_col]);
    near.plane.resize(k), near.field.resize(k);
  }
  sb.cost = decide_square(s, near, ox, oy, sb_size, lam, sb, sb.cands.empty() ? nullptr : &sb.cands, ox, oy);
  return sb;
}

Map decide_frame(const Luma& src, const std::vector<const Luma*>& refs, int sb_size, int64_t lam) {
  const Pyramid cur = build_pyramid(src, false);
  std::vector<Pyramid> pyr;
  std::vector<std::vector<Field>> lv(refs.size());
  RefSet rs;
  for (const Luma* f : refs) pyr.push_back(build_pyramid(*f, true));
  for (size_t r = 0; r < refs.size(); r++) {
    search_all(cur, pyr[r], lv[r], static_cast<int>(r));
    rs.plane.push_back(&pyr[r].plane[0]);
    rs.field.push_back(&lv[r][0]);
  }
  Map m;
  m.sb_size = sb_size;
  m.sb_cols = (src.width + sb_size - 1) / sb_size;
  m.sb_rows = (src.height + sb_size - 1) / sb_size;
  for (int r = 0; r < m.sb_rows; r++)
    for (int c = 0; c < m.sb_cols; c++) m.sbs.push_back(decide_sb(src, rs, c, r, sb_size, lam));
  return m;
}

// ---- RowGate (decision_map.cpp:19-61) ---------------------------------------------------------------

void RowGate::publish(int row) {
  {
    std::lock_guard<std::mutex> lk(m_);
    ready_ = std::max(ready_, row + 1);
  }
  cv_.notify_all();
}
void RowGate::publish_one(int row) {
  {
    std::lock_guard<std::mutex> lk(m_);
    if (one_.size() <= static_cast<size_t>(row)) one_.resize(row + 1, 0);
    one_[row] = 1;
    while (static_cast<size_t>(ready_) < one_.size() && one_[ready_]) ready_++;
  }
  cv_.notify_all();
}
void RowGate::wait_row(int row) const {
  std::unique_lock<std::mutex> lk(m_);
  cv_.wait(lk, [&] { return error_ || ready_ > row || (static_cast<size_t>(row) < one_.size() && one_[row]); });
  if (error_) std::rethrow_exception(error_);
}
void RowGate::fail(std::exception_ptr e) {
  {
    std::lock_guard<std::mutex> lk(m_);
    if (!error_) error_ = e;
  }
  cv_.notify_all();
}
int RowGate::ready_rows() const {
  std::lock_guard<std::mutex> lk(m_);
  return ready_;
}

// ---- CpuEngine (cpu_engine.cpp) -------------------------------------------------------------------

int decision_threads(int cores, int tile_threads) { return std::max(1, (cores > 0 ? cores : 1) - tile_threads); }

CpuEngine::CpuEngine(int width, int height, int sb_size, int threads, std::function<void(int)> row_hook)
    : width_(width), height_(height), sb_size_(sb_size), threads_(std::max(1, threads)),
      sb_cols_((width + sb_size - 1) / sb_size), sb_rows_((height + sb_size - 1) / sb_size),
      row_hook_(std::move(row_hook)) {
  for (auto& n : next_me_) n = 0;
}

CpuEngine::~CpuEngine() { stop_and_join(); }

void CpuEngine::launch(const Luma& src, const std::vector<const Luma*>& refs, int64_t lam, Map& dm) {
  if (src.width != width_ || src.height != height_) throw std::runtime_error("CpuEngine: source size does not match the engine");
  for (const Luma* r : refs)
    if (r->width != width_ || r->height != height_) throw std::runtime_error("CpuEngine: reference size does not match the engine");
  if (refs.size() > static_cast<size_t>(kMaxRefs)) throw std::runtime_error("CpuEngine: too many references");
  stop_and_join();
  src_ = src;
  lam_ = lam;
  nrefs_ = static_cast<int>(refs.size());
  if (nrefs_ > 0) {
    cur_ = build_pyramid(src_, false);
    ref_.clear();
    for (const Luma* r : refs) ref_.push_back(build_pyramid(*r, true));
    lv_.assign(nrefs_, std::vector<Field>(kLevels));
    for (int l = 0; l < kLevels; l++) {
      for (auto& per : lv_) {
        Field& o = per[l];
        o.cols = (cur_.plane[l].w + 7) / 8, o.rows = (cur_.plane[l].h + 7) / 8;
        o.mvx.assign(static_cast<size_t>(o.cols) * o.rows, 0), o.mvy = o.mvx;
      }
      next_me_[l] = 0;
      arrived_[l] = 0;
    }
  }
  const size_t nsb = static_cast<size_t>(sb_cols_) * sb_rows_;
  out_.assign(nsb, Sb{});
  row_done_.assign(sb_rows_, 0);
  next_row_ = 0;
  error_ = nullptr;
  abort_ = false;
  dm.sb_size = sb_size_, dm.sb_cols = sb_cols_, dm.sb_rows = sb_rows_;
  dm.sbs.assign(nsb, Sb{});
  try {
    for (int i = 0; i < threads_; i++) workers_.emplace_back([this] { work(); });
  } catch (...) {
    stop_and_join();
    throw;
  }
}

void CpuEngine::work() {
  try {
    if (nrefs_ > 0)
      for (int l = kLevels - 1; l >= 0; l--) {
        const int rows = lv_[0][l].rows;
        for (int t; !abort_ && (t = next_me_[l]++) < rows * nrefs_;) {
          const int r = t / rows, row = t % rows;
          search_rows(cur_, ref_[r], l, l == kLevels - 1 ? nullptr : &lv_[r][l + 1], row, row + 1, lv_[r][l]);
        }
        std::unique_lock<std::mutex> lk(m_);
        if (++arrived_[l] == threads_) cv_.notify_all();
        cv_.wait(lk, [&] { return arrived_[l] == threads_ || abort_; });
        if (abort_) return;
      }
    RefSet rs;
    for (int i = 0; i < nrefs_; i++) rs.plane.push_back(&ref_[i].plane[0]), rs.field.push_back(&lv_[i][0]);
    for (int r; !abort_ && (r = next_row_++) < sb_rows_;) {
      if (row_hook_) row_hook_(r);
      for (int c = 0; c < sb_cols_; c++) out_[static_cast<size_t>(r) * sb_cols_ + c] = decide_sb(src_, rs, c, r, sb_size_, lam_);
      std::lock_guard<std::mutex> lk(m_);
      row_done_[r] = 1;
      cv_.notify_all();
    }
  } catch (...) {
    std::lock_guard<std::mutex> lk(m_);
    if (!error_) error_ = std::current_exception();
    abort_ = true;
    cv_.notify_all();
  }
}

void CpuEngine::collect(Map& dm, RowGate& gate) {
  std::exception_ptr err;
  try {
    for (int next = 0; next < sb_rows_;) {
      int end = next;
      {
        std::unique_lock<std::mutex> lk(m_);
        cv_.wait(lk, [&] { return row_done_[next] || error_; });
        if (error_) {
          err = error_;
          break;
        }
        while (end < sb_rows_ && row_done_[end]) end++;
      }
      for (size_t i = static_cast<size_t>(next) * sb_cols_; i < static_cast<size_t>(end) * sb_cols_; i++)
        dm.sbs[i] = std::move(out_[i]);
      gate.publish(end - 1);
      next = end;
    }
  } catch (...) {
    err = std::current_exception();
  }
  if (err) gate.fail(err);
  stop_and_join();
  if (err) std::rethrow_exception(err);
}

void CpuEngine::stop_and_join() {
  {
    std::lock_guard<std::mutex> lk(m_);
    abort_ = true;
  }
  cv_.notify_all();
  for (std::thread& t : workers_) t.join();
  workers_.clear();
}

}  // namespace orc
