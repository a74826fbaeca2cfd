// The frame loop's coding schedule (include/celv2/gop.h): which frame is coded when, hidden or shown,
// at which qindex delta, and from which references.
//
// The schedule is the part of ChunkCoder::code (src/chunk_coder.cpp:155-616) that does not look at the
// pictures; the reference slots follow CodecState::begin_frame (src/state/codec_state.cpp:476-498) and
// the ranking Pipeline::references hands the engine (src/pipeline.cpp:298-309).
#include "celv2/gop.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <stdexcept>
#include <string>
#include <utility>

namespace celv2 {
namespace {

struct Planner {
  const celv2_gop_config& c;
  std::vector<celv2_gop_frame> out;
  int slot[CELV2_GOP_SLOTS];  // display index each slot holds (-1: empty)
  int turns_since_key = 0;

  explicit Planner(const celv2_gop_config& cfg) : c(cfg) { std::fill(slot, slot + CELV2_GOP_SLOTS, -1); }

  // The distinct frames the slots hold, nearest to `display` first (the earlier of two equally
  // distant ones first), at most refs of them.
  int ranked(int display, int32_t* refs) const {
    int held[CELV2_GOP_SLOTS], n = 0;
    for (int s : slot)
      if (s >= 0 && std::find(held, held + n, s) == held + n) held[n++] = s;
    std::sort(held, held + n, [display](int a, int b) {
      const int da = std::abs(display - a), db = std::abs(display - b);
      return da != db ? da < db : a < b;
    });
    n = std::min(n, c.refs);
    std::copy(held, held + n, refs);
    return n;
  }

  // Codes frame `display`: takes its references from the slots, then refreshes them (a frame never
  // predicts from itself).
  void code(int display, int type, bool key, bool shown, bool reference, bool keep_shown, int level, int delta) {
    celv2_gop_frame f{};
    f.display = display, f.type = type, f.coded = 1, f.shown = shown, f.reference = reference || key;
    f.keep_shown = keep_shown, f.level = level, f.qindex_delta = delta;
    f.qindex = std::max(1, std::min(255, c.qindex + delta));
    f.nrefs = key ? 0 : ranked(display, f.refs);
    const int all = (1 << CELV2_GOP_SLOTS) - 1, turns = (1 << c.refs) - 1;
    int refresh;
    if (key) {
      turns_since_key = 0;
      refresh = all;
    } else if (reference) {
      // The turn holding the earliest frame, from the turn count since the keyframe on.
      int s = ++turns_since_key % c.refs;
      for (int i = 0; i < c.refs; i++)
        if (slot[i] < slot[s]) s = i;
      refresh = (1 << s) | (shown ? all & ~turns : 0);
    } else {
      refresh = shown && keep_shown ? all & ~turns : 0;
    }
    f.refresh = refresh;
    for (int i = 0; i < CELV2_GOP_SLOTS; i++)
      if (refresh >> i & 1) slot[i] = display;
    out.push_back(f);
  }

  void show_existing(int display, int type, int level, int delta) {
    celv2_gop_frame f{};
    f.display = display, f.type = type, f.shown = 1, f.reference = 0, f.level = level, f.qindex_delta = delta;
    f.qindex = std::max(1, std::min(255, c.qindex + delta));
    out.push_back(f);
  }

  // The pyramid of a mini-GOP of m frames (positions 0..m-1, the anchor at m-1): the middle of every
  // gap of 3 or more frames between references, depth first, (position, level) from level 1; none
  // with fewer than 3 references, and no more than leave the anchor and the last shown frame a slot.
  std::vector<std::pair<int, int>> pyramid(int m) const {
    std::vector<std::pair<int, int>> mids;
    if (!c.pyramid) return mids;
    struct Gap {
      int lo, hi, level;
    };
    std::vector<Gap> todo{{-1, m - 1, 1}};  // a stack: the left half is split before the right
    while (!todo.empty()) {
      const Gap g = todo.back();
      todo.pop_back();
      if (g.hi - g.lo - 1 < 3 || c.refs < 3 || static_cast<int>(mids.size()) + 2 >= c.refs) continue;
      const int mid = (g.lo + g.hi) / 2;
      mids.emplace_back(mid, g.level);
      todo.push_back({mid, g.hi, g.level + 1});
      todo.push_back({g.lo, mid, g.level + 1});
    }
    return mids;
  }

  void minigop(int start, int m) {
    const int step = m > 8 && c.level_step16 >= 0 ? c.level_step16 : c.level_step;
    const auto mids = pyramid(m);
    code(start + m - 1, CELV2_GOP_ANCHOR, false, false, true, true, 0, c.aq);
    std::vector<int> level(m, 0);
    for (const auto& [pos, lv] : mids) {
      level[pos] = lv;
      code(start + pos, CELV2_GOP_MID, false, false, true, true, lv, c.mq + (lv - 1) * step);
    }
    int last_leaf = -1;
    for (int i = 0; i + 1 < m; i++)
      if (!level[i]) last_leaf = i;
    for (int i = 0; i + 1 < m; i++) {
      if (level[i])
        show_existing(start + i, CELV2_GOP_MID, level[i], c.mq + (level[i] - 1) * step);
      else
        code(start + i, CELV2_GOP_LEAF, false, true, false, !c.parallel_b || i == last_leaf, 0, c.bq);
    }
    show_existing(start + m - 1, CELV2_GOP_ANCHOR, 0, c.aq);
  }

  void run(int frames, const std::vector<bool>& key) {
    int i = 0;
    while (i < frames) {
      if (key[i]) {
        // The frames it leads, as far as the loop looks ahead (bframes + 1, at most 8): a shot
        // shorter than 9 frames takes (ahead + 1) / 9 of a negative keyframe delta.
        int ahead = 0;
        while (ahead < 8 && ahead < c.bframes + 1 && i + ahead + 1 < frames && !key[i + ahead + 1]) ahead++;
        int kq = c.kq;
        if (ahead < 8 && kq < 0) kq = static_cast<int>(std::lround(kq * (ahead + 1) / 9.0));
        code(i, CELV2_GOP_KEY, true, true, true, true, 0, kq);
        i++;
        continue;
      }
      int m = 0;
      while (m < c.bframes + 1 && i + m < frames && !key[i + m]) m++;
      if (m <= 1)
        code(i, CELV2_GOP_LONE, false, true, true, true, 0, 0);
      else
        minigop(i, m);
      i += m;
    }
  }
};

void check(const celv2_gop_config& c, int frames, const std::vector<bool>& key) {
  if (c.refs < 1 || c.refs > CELV2_GOP_MAX_REFS) throw std::invalid_argument("gop: refs must be 1..7");
  if (c.bframes < 0 || c.bframes > 15) throw std::invalid_argument("gop: bframes must be 0..15");
  if (c.qindex < 1 || c.qindex > 255) throw std::invalid_argument("gop: qindex must be 1..255");
  if (c.level_step16 < -1) throw std::invalid_argument("gop: level_step16 must be -1 or more");
  if (frames < 0 || static_cast<int>(key.size()) != frames) throw std::invalid_argument("gop: one key flag per frame");
  if (frames > 0 && !key[0]) throw std::invalid_argument("gop: an inter frame needs a preceding keyframe");
}

}  // namespace

std::vector<celv2_gop_frame> gop_plan(const celv2_gop_config& c, int frames, const std::vector<bool>& key) {
  check(c, frames, key);
  Planner p(c);
  p.run(frames, key);
  return std::move(p.out);
}

}  // namespace celv2

namespace {
thread_local std::string g_gop_err;
}

extern "C" {

void celv2_gop_defaults(celv2_gop_config* c, int refs, int qindex, int width, int height, int long_gop) {
  if (!c) return;
  c->refs = refs;
  c->qindex = qindex;
  c->bframes = refs < 2 ? 0 : refs >= 5 && qindex >= 100 ? (long_gop ? 15 : 7) : 3;
  c->parallel_b = static_cast<long long>(width) * height <= 2LL * 3840 * 2160;
  c->pyramid = 1;
  c->bq = 16, c->aq = -16, c->mq = 0, c->kq = 0;
  c->level_step = 8, c->level_step16 = -1;
}

int celv2_gop_plan(const celv2_gop_config* c, int frames, const uint8_t* key, celv2_gop_frame* out, int cap,
                   int* count) {
  try {
    if (!c || !count || cap < 0 || (cap > 0 && !out) || frames < 0)
      throw std::invalid_argument("gop: null or negative argument");
    std::vector<bool> k(frames, false);
    for (int i = 0; i < frames; i++) k[i] = key ? key[i] != 0 : i == 0;
    const auto plan = celv2::gop_plan(*c, frames, k);
    *count = static_cast<int>(plan.size());
    std::copy_n(plan.begin(), std::min<size_t>(plan.size(), static_cast<size_t>(cap)), out);
    if (static_cast<int>(plan.size()) > cap) {
      g_gop_err = "gop: " + std::to_string(plan.size()) + " entries do not fit in " + std::to_string(cap);
      return 2;
    }
    return 0;
  } catch (const std::exception& e) {
    g_gop_err = e.what();
    return 1;
  }
}

const char* celv2_gop_last_error(void) { return g_gop_err.c_str(); }

}  // extern "C"
