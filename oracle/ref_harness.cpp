// ORACLE PINNING HARNESS — test infrastructure. Compiled only by `make -C oracle ref` in a tree where
// the CelV2 reference is mounted (REF=/root/reference); it includes the reference's own header-only
// decision primitives from $(REF)/src/gpu and exports them under ref_* names with the same
// signatures as the oracle's orc_* entry points, so tests can compare the two on the same inputs.
// The reference's kernels_ref.cpp (per-block choice, partition DP, pyramid) is not built: it
// includes AVM's av2/common/quant_common.h, which this tree does not have.
#include <cstdint>
#include <cstring>
#include <vector>

#include "gpu/inter_common.h"
#include "gpu/intra_common.h"
#include "gpu/sb_blocks.h"

using celv2::Plane;

extern "C" {

int ref_wht4(const int* d) { return celv2::hadamard4x4(d); }
int ref_mode_bits(int m) { return celv2::mode_bits(m); }
int ref_block_geom(int b, int* xywh) {
  celv2::sb_block_geom(b, xywh, xywh + 1, xywh + 2, xywh + 3);
  return 0;
}
int ref_block_index(int x, int y, int w, int h) { return celv2::sb_block_index(x, y, w, h); }
int ref_cand_count(int x, int y, int w, int h) { return celv2::mv_candidates(x, y, w, h); }
int ref_cand_cell(int x, int y, int w, int h, int k, int cols, int rows) {
  return celv2::candidate_mv_index(x, y, w, h, k, cols, rows);
}
int ref_cand_repeats(const int* fx, const int* fy, int cols, int rows, int x, int y, int w, int h, int k) {
  return celv2::repeated_candidate(fx, fy, cols, rows, x, y, w, h, k) ? 1 : 0;
}
int ref_pair_allowed(int a, int b, int n) { return celv2::compound_pair_ok(a, b, n) ? 1 : 0; }

int ref_qpel_sample(const uint8_t* s, int w, int h, int y, int x, int qy, int qx) {
  return celv2::subpel_px(Plane{s, w, w, h}, y, x, qy, qx);
}
int ref_half_sample(const uint8_t* s, int w, int h, int y, int x) { return celv2::downsample_px(Plane{s, w, w, h}, y, x); }

int ref_search_cell(const uint8_t* cur, const uint8_t* ref, const uint8_t* ref_phases, int w, int h, int bx, int by,
                    const int* pmx, const int* pmy, int pcols, int prows, int level0, int* mv) {
  const Plane pc{cur, w, w, h};
  Plane pr{ref_phases ? ref_phases : ref, w, w, h};
  pr.phase = ref_phases ? static_cast<size_t>(w) * h : 0;
  celv2::me_cell(pc, pr, bx, by, pmx, pmy, pcols, prows, &mv[0], &mv[1], level0 != 0);
  if (level0 && ref_phases) celv2::me_refine(pc, pr, bx, by, &mv[0], &mv[1]);
  return 0;
}

// me_search alone (the reference's range parameter): +-range around (px, py); returns its cost.
int ref_me_search(const uint8_t* cur, const uint8_t* ref, int w, int h, int bx, int by, int px, int py, int range, int* mv) {
  return celv2::me_search(Plane{cur, w, w, h}, Plane{ref, w, w, h}, bx, by, px, py, range, &mv[0], &mv[1]);
}

long long ref_intra_satd(const uint8_t* s, int w, int h, int x, int y, int bw, int bh, int mode) {
  return celv2::block_mode_satd(Plane{s, w, w, h}, x, y, bw, bh, mode);
}
int ref_intra_pred(const uint8_t* s, int w, int h, int x, int y, int bw, int bh, int mode, int i, int j) {
  const Plane p{s, w, w, h};
  return mode == 0 ? celv2::dc_value(p, x, y, bw, bh) : celv2::intra_pred(p, x, y, bw, bh, mode, i, j);
}
long long ref_inter_satd(const uint8_t* cur, const uint8_t* ref_phases, int w, int h, int x, int y, int bw, int bh, int mvx,
                         int mvy) {
  Plane pr{ref_phases, w, w, h};
  pr.phase = static_cast<size_t>(w) * h;
  return celv2::block_inter_satd(Plane{cur, w, w, h}, pr, x, y, bw, bh, mvx, mvy);
}
long long ref_compound_satd(const uint8_t* cur, const uint8_t* r0, const uint8_t* r1, int w, int h, int x, int y, int bw,
                            int bh, int mx0, int my0, int mx1, int my1) {
  Plane p0{r0, w, w, h}, p1{r1, w, w, h};
  p0.phase = p1.phase = static_cast<size_t>(w) * h;
  return celv2::block_compound_satd(Plane{cur, w, w, h}, p0, mx0, my0, p1, mx1, my1, x, y, bw, bh);
}

// The reference primitives composed into a whole pyramid search (levels 4..0) by this harness's own
// loop, as kernels_ref.cpp's build_pyramid_ref / me_level_rows_ref order them: the output layout is
// orc_fields'. Pyramid: level 0 the plane (ref: + 15 phases by subpel_px), levels 1..4 by downsample_px.
int ref_fields(int w, int h, const uint8_t* cur, const uint8_t* ref, int* mvx, int* mvy, int* dims) {
  const int pw = (w + 7) & ~7, ph = (h + 7) & ~7;
  const size_t n = static_cast<size_t>(pw) * ph;
  std::vector<std::vector<uint8_t>> pc(5), pr(5);
  pc[0].assign(cur, cur + n);
  pr[0].assign(ref, ref + n);
  pr[0].resize(n * 16);
  const Plane src{ref, pw, pw, ph};
  for (int k = 1; k < 16; k++)
    for (int y = 0; y < ph; y++)
      for (int x = 0; x < pw; x++) pr[0][k * n + static_cast<size_t>(y) * pw + x] = static_cast<uint8_t>(celv2::subpel_px(src, y, x, k / 4, k % 4));
  int lw[5], lh[5];
  lw[0] = pw, lh[0] = ph;
  for (int l = 1; l < 5; l++) {
    lw[l] = lw[l - 1] / 2, lh[l] = lh[l - 1] / 2;
    pc[l].resize(static_cast<size_t>(lw[l]) * lh[l]);
    pr[l].resize(static_cast<size_t>(lw[l]) * lh[l]);
    const Plane uc{pc[l - 1].data(), lw[l - 1], lw[l - 1], lh[l - 1]}, ur{pr[l - 1].data(), lw[l - 1], lw[l - 1], lh[l - 1]};
    for (int y = 0; y < lh[l]; y++)
      for (int x = 0; x < lw[l]; x++) {
        pc[l][static_cast<size_t>(y) * lw[l] + x] = static_cast<uint8_t>(celv2::downsample_px(uc, y, x));
        pr[l][static_cast<size_t>(y) * lw[l] + x] = static_cast<uint8_t>(celv2::downsample_px(ur, y, x));
      }
  }
  std::vector<std::vector<int>> fx(5), fy(5);
  int cols[5], rows[5];
  for (int l = 4; l >= 0; l--) {
    cols[l] = (lw[l] + 7) / 8, rows[l] = (lh[l] + 7) / 8;
    fx[l].assign(static_cast<size_t>(cols[l]) * rows[l], 0), fy[l] = fx[l];
    const Plane c{pc[l].data(), lw[l], lw[l], lh[l]};
    Plane r{pr[l].data(), lw[l], lw[l], lh[l]};
    if (l == 0) r.phase = n;
    for (int by = 0; by < rows[l]; by++)
      for (int bx = 0; bx < cols[l]; bx++) {
        int* mx = &fx[l][static_cast<size_t>(by) * cols[l] + bx];
        int* my = &fy[l][static_cast<size_t>(by) * cols[l] + bx];
        celv2::me_cell(c, r, bx, by, l == 4 ? nullptr : fx[l + 1].data(), l == 4 ? nullptr : fy[l + 1].data(),
                       l == 4 ? 0 : cols[l + 1], l == 4 ? 0 : rows[l + 1], mx, my, l == 0);
        if (l == 0) celv2::me_refine(c, r, bx, by, mx, my);
      }
  }
  int off = 0;
  for (int l = 0; l < 5; l++) {
    dims[3 * l] = cols[l], dims[3 * l + 1] = rows[l], dims[3 * l + 2] = off;
    std::memcpy(mvx + off, fx[l].data(), fx[l].size() * sizeof(int));
    std::memcpy(mvy + off, fy[l].data(), fy[l].size() * sizeof(int));
    off += static_cast<int>(fx[l].size());
  }
  return off;
}

}  // extern "C"
