// Per superblock: each candidate block's reference choice and compound (best_both,
// src/gpu/kernels_ref.cpp:43-104), its intra choice, its record (BlockCandidates), the bottom-up
// partition decision (decide_square, src/gpu/kernels_ref.cpp:124-198) and the exported superblock
// record (celv2_sb, include/celv2hip.h).
//
// One workgroup per superblock. Compound SATDs run on the matrix cores (hadamard.h): the 3072 tiles of
// all 425 blocks, sorted by block size so that every 16-tile product holds whole blocks or a part of
// one block in an aligned group of columns; lane (j, g) loads row g of tile j from the source and from
// both predictions and averages the predictions bytewise. Per-block sums are LDS float atomics of
// exact integers (order-free). The partition DP runs level by level with one thread per square, and
// one thread walks the chosen tree.
#include <stdexcept>
#include <vector>

#include "celv2hip.h"
#include "common.h"
#include "hadamard.h"
#include "launch.h"

namespace celv2hip {

namespace {

constexpr int kAllTiles = 3072;  // 4x4 tiles of all 425 blocks: per level 256 (whole) + 4 x 128 (halves)
constexpr int64_t kMax = INT64_MAX;
constexpr int kThreads = 1024;  // 16 waves: the compound rows' gathers spread over more waves (2 superblocks per CU)

__device__ uint32_t g_all_tiles[kAllTiles];  // b | tx << 16 | ty << 24
__device__ uint8_t g_tile_seg[kAllTiles / 16];  // tiles per block (at most 16) of each 16-tile step

__device__ __forceinline__ bool pair_allowed(int a, int b, int n) {  // compound_pair_ok, inter_common.h:199-202
  const int lo = a < b ? a : b, lim = n - 2 < 2 ? n - 2 : 2;
  return a != b && lo <= lim;
}

// (p + q + 1) >> 1 of each byte pair (the compound average, block_compound_satd,
// src/gpu/inter_common.h:205-220): (p | q) - ((p ^ q) >> 1).
__device__ __forceinline__ uint32_t avg_bytes(uint32_t p, uint32_t q) { return (p | q) - (((p ^ q) >> 1) & 0x7f7f7f7fu); }

// Row g of the 4x4 tile at (x, y) of a bordered plane, the tile's origin clamped as a 4x4 window.
__device__ __forceinline__ uint32_t tile_row(const uint8_t* plane, int stride, int w, int h, int x, int y, int g) {
  const int X = clamp_origin(x, w, 4), Y = clamp_origin(y, h, 4);
  return load4u(plane + static_cast<ptrdiff_t>(Y + g) * stride + X);
}

__global__ void __launch_bounds__(kThreads) k_finalize(Batch<FinalizeArgs> batch) {
  __shared__ int64_t rc[kMaxRefs][kBlocks];    // per-reference inter cost
  __shared__ uint32_t rmv[kMaxRefs][kBlocks];  // per-reference MV (quarter pel, packed)
  __shared__ int8_t ra[kBlocks], rs[kBlocks];  // best and second reference (-1: none)
  __shared__ float comp[kBlocks];              // compound SATD (blocks with an allowed pair)
  __shared__ int64_t bcost[kBlocks];           // the block's choice's cost
  __shared__ uint8_t bmode[kBlocks];           // its intra mode, kModes = inter
  __shared__ uint8_t bref[kBlocks], bref2[kBlocks];
  __shared__ uint32_t bmv[kBlocks], bmv2[kBlocks];  // the inter choice's MVs (quarter pel, packed)
  __shared__ int64_t sq_cost[kSquares];
  __shared__ uint8_t sq_choice[kSquares];
  __shared__ int16_t sq_nt[kSquares], sq_nl[kSquares];      // the chosen subtree's tree entries and leaves
  __shared__ int16_t sq_toff[kSquares], sq_loff[kSquares];  // its first tree entry and leaf (-1: not in the tree)
  __shared__ uint8_t tree[CELV2_SB_CAP];
  __shared__ uint16_t leaf[CELV2_SB_CAP];
  __shared__ int counts[2];
  __shared__ const uint8_t* refp[kMaxRefs];  // reference planes (level 0, phase 0), indexable per lane

  const int tid = threadIdx.x;
  const int bi = batch_item(batch, static_cast<int>(blockIdx.x));
  const FinalizeArgs& a = batch.item[bi];
  const int lw = static_cast<int>(blockIdx.x) - batch.first[bi];
  const int sbx = lw % a.sb_cols, sby = a.row0 + lw / a.sb_cols;
  const size_t sb = static_cast<size_t>(sby) * a.sb_cols + sbx;
  const int ox = 64 * sbx, oy = 64 * sby;
  // With a reference limit, the superblock is decided exactly as with its nearest kmap[sb] references.
  const int n = a.kmap ? min(a.nrefs, static_cast<int>(a.kmap[sb])) : a.nrefs;
  const int64_t lam = a.lam;
  celv2_sb* const out = static_cast<celv2_sb*>(a.out) + sb;
  if (tid < kMaxRefs) refp[tid] = tid < n ? a.ref[tid].p : nullptr;

  // Per block: each reference's best (cost and MV), the best and the runner-up reference.
  for (int b = tid; b < kBlocks; b += kThreads) {
    int best = -1, second = -1;
    for (int r = 0; r < n; r++) {
      const uint2 e = a.best[(sb * kMaxRefs + r) * kBlocks + b];
      const int64_t c = static_cast<int64_t>(e.x) * 256 + lam * (kNewmvBits + r);
      rc[r][b] = c, rmv[r][b] = e.y;
      if (best < 0 || c < rc[best][b]) best = r;
    }
    for (int r = 0; r < n; r++)
      if (r != best && (second < 0 || rc[r][b] < rc[second][b])) second = r;
    ra[b] = static_cast<int8_t>(best), rs[b] = static_cast<int8_t>(second);
    comp[b] = 0.f;
  }
  __syncthreads();

  // Compound SATDs of the blocks whose two best references form an allowed pair (the products of
  // other blocks' tiles are computed and dropped).
  if (n >= 2) {
    const int lane = tid & 63, j = lane & 15, g = lane >> 4;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const half4 hh = hadamard_operand(lane, false);
    const f32x4 zero = {0.f, 0.f, 0.f, 0.f};
    const DPlane& r0 = a.ref[0];
    // Four steps at a time: every load of the four is in flight before the first product.
    constexpr int kWaves = kThreads / 64, kBatch = 4;
    static_assert((kAllTiles / 16) % (kWaves * kBatch) == 0, "compound steps per batch");
    constexpr int kRounds = kAllTiles / 16 / (kWaves * kBatch);
#pragma unroll
    for (int it = 0; it < kRounds; it++) {
      const int step0 = wave + it * kWaves * kBatch;
      uint32_t cur[kBatch], avg[kBatch];
      int blk[kBatch];
      bool on[kBatch];
#pragma unroll
      for (int u = 0; u < kBatch; u++) {
        const uint32_t e = g_all_tiles[16 * (step0 + kWaves * u) + j];
        const int b = static_cast<int>(e & 0xffff), tx = static_cast<int>((e >> 16) & 255), ty = static_cast<int>(e >> 24);
        int x, y, w, h;
        block_geom(b, x, y, w, h);
        const int px = ox + x + 4 * tx, py = oy + y + 4 * ty;
        const int p = ra[b], q = rs[b];
        blk[u] = b, on[u] = q >= 0 && pair_allowed(p, q, n);
        cur[u] = *reinterpret_cast<const uint32_t*>(a.src.p + static_cast<ptrdiff_t>(py + g) * a.src.stride + px);
        avg[u] = cur[u];
        if (on[u]) {
          const uint32_t mp = rmv[p][b], mq = rmv[q][b];
          const uint8_t* pp = refp[p] + a.phase_bytes * ((mv_y(mp) & 3) * 4 + (mv_x(mp) & 3));
          const uint8_t* qp = refp[q] + a.phase_bytes * ((mv_y(mq) & 3) * 4 + (mv_x(mq) & 3));
          avg[u] = avg_bytes(tile_row(pp, r0.stride, r0.w, r0.h, px + (mv_x(mp) >> 2), py + (mv_y(mp) >> 2), g),
                             tile_row(qp, r0.stride, r0.w, r0.h, px + (mv_x(mq) >> 2), py + (mv_y(mq) >> 2), g));
        }
      }
#pragma unroll
      for (int u = 0; u < kBatch; u++) {
        const int seg = g_tile_seg[step0 + kWaves * u];
        float s = abs_sum4(mfma16(hh, row_halves(cur[u]) - row_halves(avg[u]), zero));
        // Sum over the block's tiles in this product (an aligned group of seg columns of this row).
        if (seg >= 2) s += as_f32(dpp<kDppXor1>(as_u32f(s)));
        if (seg >= 4) s += as_f32(dpp<kDppXor2>(as_u32f(s)));
        if (seg >= 8) s += as_f32(dpp<kDppHalfMirror>(as_u32f(s)));
        if (seg >= 16) s += as_f32(dpp<kDppMirror>(as_u32f(s)));
        if (on[u] && (j & (seg - 1)) == 0) atomicAdd(&comp[blk[u]], s);
      }
    }
  }
  __syncthreads();

  // Each block's choices, its candidate record, and its cost for the DP.
  for (int b = tid; b < kBlocks; b += kThreads) {
    int x, y, w, h;
    block_geom(b, x, y, w, h);
    const int p = ra[b], q = rs[b];
    int64_t ci = kMax;
    uint32_t m1 = 0, m2 = 0;
    int r1 = 0, r2 = CELV2_SINGLE_REF;
    if (p >= 0) ci = rc[p][b], m1 = rmv[p][b], r1 = p;
    if (q >= 0 && pair_allowed(p, q, n)) {
      const int64_t cc = static_cast<int64_t>(comp[b]) * 256 + lam * (20 + p + q);  // comp: an exact integer
      if (cc < ci) ci = cc, r2 = q, m2 = rmv[q][b];
    }
    int64_t c = kMax;
    int m = 0;
    if (b >= 5) {  // blocks above 32x32 are inter only (intra_common.h:39)
      const uint32_t* is = a.intra + (sb * kBlocks + b) * kModes;
      const uint32_t gm = a.imask ? a.imask[sb] : kAllGroups;
      for (int k = 0; k < kModes; k++) {
        if (!mode_scored(gm, k)) continue;
        const int64_t ck = static_cast<int64_t>(is[k]) * 256 + lam * (mode_bits(k) + (n > 0 ? kIntraOnInter : 0));
        if (ck < c) c = ck, m = k;
      }
    }
    celv2_cands rec;
    rec.inter_cost = ci, rec.intra_cost = c;
    rec.inter.w4 = static_cast<uint8_t>(w / 4), rec.inter.h4 = static_cast<uint8_t>(h / 4), rec.inter.inter = 1, rec.inter.intra = 0;
    rec.inter.mv_col = static_cast<int16_t>(2 * mv_x(m1)), rec.inter.mv_row = static_cast<int16_t>(2 * mv_y(m1));
    rec.inter.ref = static_cast<uint8_t>(r1), rec.inter.ref2 = static_cast<uint8_t>(r2);
    rec.inter.mv2_col = static_cast<int16_t>(2 * mv_x(m2)), rec.inter.mv2_row = static_cast<int16_t>(2 * mv_y(m2));
    rec.intra = static_cast<uint8_t>(m), rec.pad_ = 0;
    for (int r = 0; r < CELV2_MAX_REFS; r++) {
      const bool has = r < n;
      const int64_t v = has ? (rc[r][b] >> 8) : 0;
      rec.ref_cost[r] = has ? static_cast<uint32_t>(v < 0xfffffffeLL ? v : 0xfffffffeLL) : 0xffffffffu;
      rec.ref_mv_col[r] = static_cast<int16_t>(has ? 2 * mv_x(rmv[r][b]) : 0);
      rec.ref_mv_row[r] = static_cast<int16_t>(has ? 2 * mv_y(rmv[r][b]) : 0);
    }
    out->cands[b] = rec;
    const bool inter = ci < c;  // inter only when strictly cheaper (kernels_ref.cpp:92, 103)
    bcost[b] = inter ? ci : c;
    bmode[b] = static_cast<uint8_t>(inter ? kModes : m);
    bref[b] = static_cast<uint8_t>(r1), bref2[b] = static_cast<uint8_t>(r2);
    bmv[b] = m1, bmv2[b] = m2;
  }
  __syncthreads();

  // Partition DP, 8x8 squares up to the 64x64: NONE, HORZ, VERT, then SPLIT, first strict minimum.
  const int64_t part = lam * kSplitBits;
  for (int level = 3; level >= 0; level--) {
    const int cnt = 1 << (2 * level);
    if (tid < cnt) {
      const int sq = level_first(level) + tid, b = 5 * sq;
      int64_t best = kMax;
      int choice = 3;
      if ((64 >> level) <= 32 || n > 0) {
        const int64_t ch = part + bcost[b + 1] + bcost[b + 2], cv = part + bcost[b + 3] + bcost[b + 4];
        best = bcost[b], choice = 0;
        if (ch < best) best = ch, choice = 1;
        if (cv < best) best = cv, choice = 2;
      }
      if (level < 3) {
        const int nn = 1 << level, qx = tid % nn, qy = tid / nn;
        const int child = level_first(level + 1) + (2 * qy) * (2 * nn) + 2 * qx;
        const int64_t cs = part + sq_cost[child] + sq_cost[child + 1] + sq_cost[child + 2 * nn] + sq_cost[child + 2 * nn + 1];
        if (cs < best) best = cs, choice = 3;
        if (choice == 3) {
          sq_nt[sq] = static_cast<int16_t>(1 + sq_nt[child] + sq_nt[child + 1] + sq_nt[child + 2 * nn] + sq_nt[child + 2 * nn + 1]);
          sq_nl[sq] = static_cast<int16_t>(sq_nl[child] + sq_nl[child + 1] + sq_nl[child + 2 * nn] + sq_nl[child + 2 * nn + 1]);
        }
      }
      if (choice != 3) sq_nt[sq] = static_cast<int16_t>(choice == 0 ? 1 : 3), sq_nl[sq] = static_cast<int16_t>(choice == 0 ? 1 : 2);
      sq_cost[sq] = best, sq_choice[sq] = static_cast<uint8_t>(choice);
    }
    __syncthreads();
  }

  // The chosen tree in preorder, built in parallel: each square's first tree entry and leaf from its
  // parent's (a split's children follow it in quadrant order 0..3, each after the previous child's
  // whole subtree), then every square in the tree writes its own entries.
  for (int i = tid; i < kSquares; i += kThreads) sq_toff[i] = static_cast<int16_t>(i == 0 ? 0 : -1), sq_loff[i] = 0;
  __syncthreads();
  for (int level = 0; level < 3; level++) {
    const int cnt = 1 << (2 * level);
    if (tid < cnt) {
      const int sq = level_first(level) + tid;
      if (sq_toff[sq] >= 0 && sq_choice[sq] == 3) {
        const int nn = 1 << level, qx = tid % nn, qy = tid / nn;
        const int child = level_first(level + 1) + (2 * qy) * (2 * nn) + 2 * qx;
        const int kids[4] = {child, child + 1, child + 2 * nn, child + 2 * nn + 1};
        int t = sq_toff[sq] + 1, l = sq_loff[sq];
#pragma unroll
        for (int q = 0; q < 4; q++) {
          sq_toff[kids[q]] = static_cast<int16_t>(t), sq_loff[kids[q]] = static_cast<int16_t>(l);
          t += sq_nt[kids[q]], l += sq_nl[kids[q]];
        }
      }
    }
    __syncthreads();
  }
  for (int sq = tid; sq < kSquares; sq += kThreads) {
    const int t = sq_toff[sq], l = sq_loff[sq], ch = sq_choice[sq], b = 5 * sq;
    if (t < 0) continue;
    tree[t] = static_cast<uint8_t>(ch);
    if (ch == 0) leaf[l] = static_cast<uint16_t>(b);
    if (ch == 1 || ch == 2) {
      tree[t + 1] = 0, tree[t + 2] = 0;
      leaf[l] = static_cast<uint16_t>(b + 2 * ch - 1), leaf[l + 1] = static_cast<uint16_t>(b + 2 * ch);
    }
  }
  if (tid == 0) {
    if (a.sb_cost) a.sb_cost[sb] = sq_cost[0];
    counts[0] = sq_nt[0], counts[1] = sq_nl[0];
  }
  __syncthreads();
  const int nt = counts[0], nl = counts[1];
  if (tid == 0) out->ntree = nt, out->nblocks = nl, out->ncands = kBlocks, out->pad_ = 0;
  for (int i = tid; i < CELV2_SB_CAP; i += kThreads) out->tree[i] = i < nt ? tree[i] : 0;
  for (int i = tid; i < CELV2_SB_CAP; i += kThreads) {
    celv2_block d = {};
    if (i < nl) {
      const int b = leaf[i];
      int x, y, w, h;
      block_geom(b, x, y, w, h);
      const bool inter = bmode[b] == kModes;
      d.w4 = static_cast<uint8_t>(w / 4), d.h4 = static_cast<uint8_t>(h / 4);
      d.inter = inter ? 1 : 0, d.intra = inter ? 0 : bmode[b];
      d.ref2 = CELV2_SINGLE_REF;
      if (inter) {
        d.mv_col = static_cast<int16_t>(2 * mv_x(bmv[b])), d.mv_row = static_cast<int16_t>(2 * mv_y(bmv[b]));
        d.ref = bref[b];
        if (bref2[b] != CELV2_SINGLE_REF)
          d.ref2 = bref2[b], d.mv2_col = static_cast<int16_t>(2 * mv_x(bmv2[b])), d.mv2_row = static_cast<int16_t>(2 * mv_y(bmv2[b]));
      }
    }
    out->blocks[i] = d;
  }
  if (a.counter) {
    // Compound blocks scored: each thread's blocks, summed over its wave.
    int k = 0;
    for (int b = tid; b < kBlocks; b += kThreads) k += (rs[b] >= 0 && pair_allowed(ra[b], rs[b], n)) ? 1 : 0;
    for (int m = 32; m >= 1; m >>= 1) k += __shfl_xor(k, m, 64);
    if ((tid & 63) == 0 && k) count(a.counter, static_cast<int>(sb) * (kThreads / 64) + (tid >> 6), static_cast<unsigned long long>(k));
  }
}

}  // namespace

void add_finalize(Batch<FinalizeArgs>& b, const FinalizeArgs& a) {
  if (a.nrows > 0) b.add(a, a.sb_cols * a.nrows);
}
void launch_finalize(const Batch<FinalizeArgs>& b, hipStream_t s) {
  if (b.total() > 0) k_finalize<<<b.total(), kThreads, 0, s>>>(b);
}

// All blocks' tiles, blocks in decreasing tile count (256 .. 2, every class a multiple of 16 tiles),
// each block's tiles contiguous; seg[step] = min(16, tiles per block) of the step's blocks.
// Built in local memory and copied to this device's symbols (synchronous).
hipError_t upload_final_tiles() {
  std::vector<uint32_t> table(kAllTiles);
  std::vector<uint8_t> seg(kAllTiles / 16);
  int n = 0;
  for (int count = 256; count >= 2; count /= 2)
    for (int b = 0; b < kBlocks; b++) {
      const int sq = b / 5, kind = b % 5;
      const int level = sq >= 21 ? 3 : sq >= 5 ? 2 : sq >= 1 ? 1 : 0;
      const int size = 64 >> level;
      const int w = (kind == 3 || kind == 4) ? size / 2 : size, h = (kind == 1 || kind == 2) ? size / 2 : size;
      if ((w / 4) * (h / 4) != count) continue;
      if (n % 16 != 0 && count >= 16) throw std::logic_error("compound tile alignment");
      for (int ty = 0; ty < h / 4; ty++)
        for (int tx = 0; tx < w / 4; tx++) {
          if (n % 16 == 0) seg[n / 16] = static_cast<uint8_t>(count < 16 ? count : 16);
          table[n++] = static_cast<uint32_t>(b) | (static_cast<uint32_t>(tx) << 16) | (static_cast<uint32_t>(ty) << 24);
        }
    }
  if (n != kAllTiles) throw std::logic_error("compound tile table size");
  hipError_t rc = hipMemcpyToSymbol(HIP_SYMBOL(g_all_tiles), table.data(), table.size() * sizeof(uint32_t));
  if (rc == hipSuccess) rc = hipMemcpyToSymbol(HIP_SYMBOL(g_tile_seg), seg.data(), seg.size());
  return rc;
}

}  // namespace celv2hip
