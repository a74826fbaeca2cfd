// ORACLE — test infrastructure, not product code (see oracle.h). C entry points for the tests and
// for bench.py's CPU baseline; results are exported in the flat records of include/celv2hip.h so they
// compare field by field with the HIP engine's.
#include <cstring>
#include <exception>
#include <memory>
#include <stdexcept>
#include <string>

#include "../include/celv2hip.h"
#include "oracle.h"

using namespace orc;

namespace {

thread_local std::string g_err;

int fail(const std::exception& e) {
  g_err = e.what();
  return 1;
}

void export_block(const Block& b, celv2_block* o) {
  o->w4 = b.w4, o->h4 = b.h4, o->inter = b.inter ? 1 : 0, o->intra = b.intra;
  o->mv_row = b.mv_row, o->mv_col = b.mv_col, o->ref = b.ref, o->ref2 = b.ref2;
  o->mv2_row = b.mv2_row, o->mv2_col = b.mv2_col;
}

void export_sb(const Sb& s, celv2_sb* o) {
  std::memset(o, 0, sizeof(*o));
  o->ntree = static_cast<int32_t>(s.tree.size());
  o->nblocks = static_cast<int32_t>(s.blocks.size());
  o->ncands = static_cast<int32_t>(s.cands.size());
  if (s.tree.size() > CELV2_SB_CAP || s.blocks.size() > CELV2_SB_CAP) throw std::runtime_error("superblock exceeds the record");
  for (size_t i = 0; i < s.tree.size(); i++) o->tree[i] = s.tree[i];
  for (size_t i = 0; i < s.blocks.size(); i++) export_block(s.blocks[i], &o->blocks[i]);
  for (size_t k = 0; k < s.cands.size() && k < CELV2_SB_BLOCKS; k++) {
    const Cands& c = s.cands[k];
    celv2_cands& d = o->cands[k];
    d.inter_cost = c.inter_cost, d.intra_cost = c.intra_cost, d.intra = c.intra;
    export_block(c.inter, &d.inter);
    for (int r = 0; r < kMaxRefs; r++) d.ref_cost[r] = c.ref_cost[r], d.ref_mv_col[r] = c.ref_mv_col[r], d.ref_mv_row[r] = c.ref_mv_row[r];
  }
}

void export_map(const Map& m, celv2_sb* out) {
  for (size_t i = 0; i < m.sbs.size(); i++) export_sb(m.sbs[i], &out[i]);
}

std::vector<Luma> lumas(int w, int h, const uint8_t* const* ys, int n) {
  std::vector<Luma> v;
  for (int i = 0; i < n; i++) v.push_back(make_luma(w, h, ys[i]));
  return v;
}

struct Engine {
  std::unique_ptr<CpuEngine> e;
  Map dm;
  Luma src;
  std::vector<Luma> refs;
};

}  // namespace

extern "C" {

const char* orc_last_error(void) { return g_err.c_str(); }

// Whole-frame decision with the single-threaded driver (decide_inter_ref / decide_intra_ref).
int orc_decide_frame(int w, int h, const uint8_t* src, const uint8_t* const* refs, int nrefs, int64_t lam, int sb_size,
                     celv2_sb* out) {
  try {
    const Luma s = make_luma(w, h, src);
    const std::vector<Luma> r = lumas(w, h, refs, nrefs);
    std::vector<const Luma*> rp;
    for (const Luma& l : r) rp.push_back(&l);
    export_map(decide_frame(s, rp, sb_size, lam), out);
    return 0;
  } catch (const std::exception& e) {
    return fail(e);
  }
}

// The same restricted to per-superblock intra mode groups (bit g of masks[s]: group g scored; group
// 0, modes 0..2, always): the pruned form's intra dimension.
int orc_decide_frame_masked(int w, int h, const uint8_t* src, const uint8_t* const* refs, int nrefs, int64_t lam,
                            int sb_size, const uint8_t* masks, celv2_sb* out) {
  g_intra_masks = masks, g_intra_mask_cols = (w + sb_size - 1) / sb_size;
  const int rc = orc_decide_frame(w, h, src, refs, nrefs, lam, sb_size, out);
  g_intra_masks = nullptr;
  return rc;
}

// The pruned form's restrictions together (each may be null): per-superblock reference limits
// (1..7), motion-search modes (bit L: narrow full searches at level L = 0, 1) and intra mode groups;
// costs (may be null): each superblock's decision cost.
int orc_decide_frame_pruned(int w, int h, const uint8_t* src, const uint8_t* const* refs, int nrefs, int64_t lam,
                            int sb_size, const uint8_t* limits, const uint8_t* modes, const uint8_t* masks,
                            celv2_sb* out, int64_t* costs) {
  g_sb_cols = (w + sb_size - 1) / sb_size, g_sb_rows = (h + sb_size - 1) / sb_size;
  g_ref_limits = limits, g_search_modes = modes;
  g_intra_masks = masks, g_intra_mask_cols = g_sb_cols;
  int rc = 0;
  try {
    const Luma s = make_luma(w, h, src);
    const std::vector<Luma> r = lumas(w, h, refs, nrefs);
    std::vector<const Luma*> rp;
    for (const Luma& l : r) rp.push_back(&l);
    const Map m = decide_frame(s, rp, sb_size, lam);
    export_map(m, out);
    if (costs)
      for (size_t i = 0; i < m.sbs.size(); i++) costs[i] = m.sbs[i].cost;
  } catch (const std::exception& e) {
    rc = fail(e);
  }
  g_ref_limits = g_search_modes = g_intra_masks = nullptr;
  return rc;
}

// The threaded CpuEngine. fail_row >= 0: the row hook throws std::runtime_error("row hook") there.
void* orc_engine_create(int w, int h, int sb_size, int threads, int fail_row) {
  try {
    auto* x = new Engine;
    std::function<void(int)> hook;
    if (fail_row >= 0)
      hook = [fail_row](int r) {
        if (r == fail_row) throw std::runtime_error("row hook");
      };
    x->e = std::make_unique<CpuEngine>(w, h, sb_size, threads, hook);
    return x;
  } catch (const std::exception& e) {
    fail(e);
    return nullptr;
  }
}
void orc_engine_destroy(void* p) { delete static_cast<Engine*>(p); }
int orc_engine_launch(void* p, int w, int h, const uint8_t* src, const uint8_t* const* refs, int nrefs, int64_t lam) {
  try {
    Engine* x = static_cast<Engine*>(p);
    x->src = make_luma(w, h, src);
    x->refs = lumas(w, h, refs, nrefs);
    std::vector<const Luma*> rp;
    for (const Luma& l : x->refs) rp.push_back(&l);
    x->e->launch(x->src, rp, lam, x->dm);
    return 0;
  } catch (const std::exception& e) {
    return fail(e);
  }
}
// rows_ready: the gate's ready rows when collect returned (or failed).
int orc_engine_collect(void* p, celv2_sb* out, int* rows_ready) {
  Engine* x = static_cast<Engine*>(p);
  RowGate gate;
  try {
    x->e->collect(x->dm, gate);
    if (rows_ready) *rows_ready = gate.ready_rows();
    if (out) export_map(x->dm, out);
    return 0;
  } catch (const std::exception& e) {
    if (rows_ready) *rows_ready = gate.ready_rows();
    return fail(e);
  }
}
int orc_decision_threads(int cores, int tile_threads) { return decision_threads(cores, tile_threads); }

// ---- primitives, for the pinning tests -----------------------------------------------------------
int orc_wht4(const int* d) { return wht4_abs_sum(d); }
int orc_mode_bits(int m) { return mode_bits(m); }
void orc_round_plane(const uint16_t* in, uint8_t* out, long long n, int bit_depth) { round_plane(in, out, static_cast<size_t>(n), bit_depth); }
int orc_block_geom(int b, int* xywh) { block_geom(b, xywh, xywh + 1, xywh + 2, xywh + 3); return 0; }
int orc_block_index(int x, int y, int w, int h) { return block_index(x, y, w, h); }
int orc_cand_count(int x, int y, int w, int h) { return cand_count(x, y, w, h); }
int orc_cand_cell(int x, int y, int w, int h, int k, int cols, int rows) { return cand_cell(x, y, w, h, k, cols, rows); }
int orc_cand_repeats(const int* fx, const int* fy, int cols, int rows, int x, int y, int w, int h, int k) {
  return cand_repeats(fx, fy, cols, rows, x, y, w, h, k) ? 1 : 0;
}
int orc_pair_allowed(int a, int b, int n) { return pair_allowed(a, b, n) ? 1 : 0; }

// Pyramid of a pw x ph plane: level sizes and pixels (levels concatenated; a reference's level 0
// followed by its 15 phase planes). Returns the bytes written (0 when cap is too small).
long long orc_pyramid(int w, int h, const uint8_t* y, int phases, uint8_t* out, long long cap, int* dims) {
  const Pyramid p = build_pyramid(make_luma(w, h, y), phases != 0);
  long long total = 0;
  for (int l = 0; l < kLevels; l++) total += static_cast<long long>(p.px[l].size());
  if (total > cap) return 0;
  long long off = 0;
  for (int l = 0; l < kLevels; l++) {
    std::memcpy(out + off, p.px[l].data(), p.px[l].size());
    off += static_cast<long long>(p.px[l].size());
    dims[2 * l] = p.plane[l].w, dims[2 * l + 1] = p.plane[l].h;
  }
  return total;
}

// Motion fields of cur against ref, every level (level l at offsets given by the field sizes):
// mvx/mvy sized to the sum over levels of cols * rows; dims[3*l..] = cols, rows, offset.
int orc_fields(int w, int h, const uint8_t* cur, const uint8_t* ref, int* mvx, int* mvy, int* dims) {
  try {
    const Pyramid pc = build_pyramid(make_luma(w, h, cur), false), pr = build_pyramid(make_luma(w, h, ref), true);
    std::vector<Field> lv;
    search_all(pc, pr, lv);
    int off = 0;
    for (int l = 0; l < kLevels; l++) {
      dims[3 * l] = lv[l].cols, dims[3 * l + 1] = lv[l].rows, dims[3 * l + 2] = off;
      std::memcpy(mvx + off, lv[l].mvx.data(), lv[l].mvx.size() * sizeof(int));
      std::memcpy(mvy + off, lv[l].mvy.data(), lv[l].mvy.size() * sizeof(int));
      off += static_cast<int>(lv[l].mvx.size());
    }
    return off;
  } catch (const std::exception& e) {
    return -fail(e);
  }
}

// One search_cell (+ refine at level 0) on raw planes: cur/ref are w x h (stride w); ref_phases
// non-null: 16 planes of w*h for the refinement. pmx/pmy: coarser field (NULL at the top).
// Returns the integer search's winning cost (SAD + distance penalty).
int orc_search_cell(const uint8_t* cur, const uint8_t* ref, const uint8_t* ref_phases, int w, int h, int bx, int by,
                    const int* pmx, const int* pmy, int pcols, int prows, int level0, int* mv) {
  const Plane8 pc{cur, w, w, h, 0};
  const Plane8 pr{ref_phases ? ref_phases : ref, w, w, h, ref_phases ? static_cast<size_t>(w) * h : 0};
  const int cost = search_cell(pc, pr, bx, by, pmx, pmy, pcols, prows, &mv[0], &mv[1], level0 != 0);
  if (level0 && ref_phases) refine_qpel(pc, pr, bx, by, &mv[0], &mv[1]);
  return cost;
}

// The same search with the full searches' range given (the pruned form's narrow search: range
// kNarrowRange), and me_search alone: +-range around (px, py), the winning cost returned.
int orc_search_cell_range(const uint8_t* cur, const uint8_t* ref, const uint8_t* ref_phases, int w, int h, int bx, int by,
                          const int* pmx, const int* pmy, int pcols, int prows, int level0, int range, int* mv) {
  const Plane8 pc{cur, w, w, h, 0};
  const Plane8 pr{ref_phases ? ref_phases : ref, w, w, h, ref_phases ? static_cast<size_t>(w) * h : 0};
  const int cost = search_cell(pc, pr, bx, by, pmx, pmy, pcols, prows, &mv[0], &mv[1], level0 != 0, range);
  if (level0 && ref_phases) refine_qpel(pc, pr, bx, by, &mv[0], &mv[1]);
  return cost;
}
int orc_full_search(const uint8_t* cur, const uint8_t* ref, int w, int h, int bx, int by, int px, int py, int range, int* mv) {
  return full_search(Plane8{cur, w, w, h, 0}, Plane8{ref, w, w, h, 0}, bx, by, px, py, range, &mv[0], &mv[1]);
}
// Which blocks of motion-search level `level` (cols x rows) are searched narrow for reference r under
// per-superblock limits (may be null) and search modes on a sb_cols x sb_rows grid: out[i] = 1 / 0.
int orc_narrow_map(int level, int cols, int rows, int sb_cols, int sb_rows, const uint8_t* limits, const uint8_t* modes,
                   int r, uint8_t* out) {
  g_sb_cols = sb_cols, g_sb_rows = sb_rows, g_ref_limits = limits, g_search_modes = modes;
  for (int by = 0; by < rows; by++)
    for (int bx = 0; bx < cols; bx++) out[by * cols + bx] = narrow_search(level, bx, by, r) ? 1 : 0;
  g_ref_limits = g_search_modes = nullptr;
  return 0;
}

// SATDs on raw planes (w x h, stride w; ref with phases if ref_phases).
long long orc_intra_satd(const uint8_t* s, int w, int h, int x, int y, int bw, int bh, int mode) {
  return intra_satd(Plane8{s, w, w, h, 0}, x, y, bw, bh, mode);
}
int orc_intra_pred(const uint8_t* s, int w, int h, int x, int y, int bw, int bh, int mode, int i, int j) {
  const Plane8 p{s, w, w, h, 0};
  return mode == 0 ? dc_mean(p, x, y, bw, bh) : intra_predict(p, x, y, bw, bh, mode, i, j);
}
long long orc_inter_satd(const uint8_t* cur, const uint8_t* ref_phases, int w, int h, int x, int y, int bw, int bh,
                         int mvx, int mvy) {
  return inter_satd(Plane8{cur, w, w, h, 0}, Plane8{ref_phases, w, w, h, static_cast<size_t>(w) * h}, x, y, bw, bh, mvx, mvy);
}
long long orc_compound_satd(const uint8_t* cur, const uint8_t* r0, const uint8_t* r1, int w, int h, int x, int y, int bw,
                            int bh, int mx0, int my0, int mx1, int my1) {
  const size_t n = static_cast<size_t>(w) * h;
  return compound_satd(Plane8{cur, w, w, h, 0}, Plane8{r0, w, w, h, n}, mx0, my0, Plane8{r1, w, w, h, n}, mx1, my1, x, y, bw, bh);
}
int orc_qpel_sample(const uint8_t* s, int w, int h, int y, int x, int qy, int qx) {
  return qpel_sample(Plane8{s, w, w, h, 0}, y, x, qy, qx);
}

// decide_square on one square of a frame with nrefs references and their level-0 fields (cols x rows
// each): returns the cost; the subtree goes to tree / blocks (counts in n[0], n[1]); cands (425,
// may be NULL) receives the candidate records relative to (ox, oy).
long long orc_decide_square(int w, int h, const uint8_t* src, const uint8_t* const* ref_phases, const int* const* fmx,
                            const int* const* fmy, int cols, int rows, int nrefs, int64_t lam, int x, int y, int size,
                            int ox, int oy, uint8_t* tree, celv2_block* blocks, int* n, celv2_sb* cands) {
  const Plane8 s{src, w, w, h, 0};
  std::vector<Plane8> planes(nrefs);
  std::vector<Field> fields(nrefs);
  RefSet rs;
  for (int r = 0; r < nrefs; r++) {
    planes[r] = Plane8{ref_phases[r], w, w, h, static_cast<size_t>(w) * h};
    fields[r].cols = cols, fields[r].rows = rows;
    fields[r].mvx.assign(fmx[r], fmx[r] + cols * rows), fields[r].mvy.assign(fmy[r], fmy[r] + cols * rows);
  }
  for (int r = 0; r < nrefs; r++) rs.plane.push_back(&planes[r]), rs.field.push_back(&fields[r]);
  Sb out;
  std::vector<Cands> rec(kBlocks);
  for (int b = 0; b < kBlocks; b++) {
    int bx, by, bw, bh;
    block_geom(b, &bx, &by, &bw, &bh);
    rec[b].inter.w4 = static_cast<uint8_t>(bw / 4), rec[b].inter.h4 = static_cast<uint8_t>(bh / 4), rec[b].inter.inter = true;
  }
  const int64_t c = decide_square(s, rs, x, y, size, lam, out, cands ? &rec : nullptr, ox, oy);
  n[0] = static_cast<int>(out.tree.size()), n[1] = static_cast<int>(out.blocks.size());
  for (size_t i = 0; i < out.tree.size(); i++) tree[i] = out.tree[i];
  for (size_t i = 0; i < out.blocks.size(); i++) export_block(out.blocks[i], &blocks[i]);
  if (cands) {
    Sb tmp;
    tmp.cands = rec;
    export_sb(tmp, cands);
  }
  return c;
}

}  // extern "C"

extern "C" int orc_sizeof_sb(void) { return static_cast<int>(sizeof(celv2_sb)); }
