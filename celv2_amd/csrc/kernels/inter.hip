// Single-reference inter scores of a superblock: for each of the 425 candidate blocks, the first
// least SATD over its MV candidates (best_both's per-reference loop, src/gpu/kernels_ref.cpp:104-119)
// and that MV.
//
// The candidates of all 425 blocks are MVs of the 10 x 10 cells of the level-0 field around the
// superblock (its 64 cells and the ring, src/gpu/inter_common.h:225-243), and every candidate block's
// 4x4 tiles lie on the superblock's 4x4 grid. So per distinct MV the kernel scores the whole 64x64
// area once - 256 tile SATDs - and sums tiles up the block quadtree; each block's cost for any
// candidate is then a lookup. A workgroup takes one (superblock, reference); each of its 4 waves
// takes every fourth distinct MV on its own, with no workgroup barrier between MVs, the ref rows
// through buffer loads and the cell -> (block, candidate) tables from LDS. Within a block, candidate
// k's cost is (SATD << 7 | k) and the minimum of
// those keys (LDS atomicMin, order-free) is the first strict minimum in candidate order, exactly the
// reference's choice (a repeated MV costs the same as its first occurrence and never wins; the
// lambda * (3 + r) term is the same for every candidate of a reference and cannot change the order).
#include <stdexcept>

#include "common.h"
#include "launch.h"

namespace celv2hip {

namespace {

__device__ uint16_t g_cell_first[kRegion * kRegion + 1];  // tasks of region cell c: [first[c], first[c + 1])
__device__ uint16_t g_cell_tasks[kTasks];                 // b << 7 | k
__device__ uint8_t g_cand_cell[kBlocks * kMaxCands];     // region cell of candidate k of block b

// v + v of lane ^ 16 (m = 16) or lane ^ 32 (m = 32) by gfx950's row swaps (VALU, no LDS).
__device__ __forceinline__ uint32_t xor_add(uint32_t v, int m) {
  const auto r = m == 16 ? __builtin_amdgcn_permlane16_swap(v, v, false, false) : __builtin_amdgcn_permlane32_swap(v, v, false, false);
  return r[0] + r[1];
}
// Lane grid of the superblock's 8x8 squares: lane = 8 * sqy + sqx. Neighbour sums within a 16-lane
// row by DPP: lane ^ 1 (next square across), lane ^ 8 (next square down), lane ^ 2, and lane ^ 4 for
// values already equal over each quad.

__global__ void __launch_bounds__(256) k_inter(Batch<InterArgs> batch) {
  __shared__ uint32_t mv[kRegion * kRegion];
  __shared__ int cuid[kRegion * kRegion];
  __shared__ int urep[kRegion * kRegion];
  __shared__ uint32_t best[kBlocks];
  __shared__ unsigned long long seen[kBlocks][2];  // MV classes among each block's candidates (counting only)
  __shared__ uint32_t sc[4][kBlocks];
  __shared__ uint16_t cell_first[kRegion * kRegion + 1];
  __shared__ uint16_t cell_tasks[kTasks];
  __shared__ int nuniq;

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int total = batch.total();
  const int glg = xcd_logical(blockIdx.x, total);
  if (glg >= total) return;  // whole workgroup
  const int bi = batch_item(batch, glg);
  const InterArgs& a = batch.item[bi];
  const int lg = glg - batch.first[bi];
  // Superblock-major (the references of a superblock adjacent): the XCD-major numbering then gives
  // every XCD whole superblocks, however the reference limits thin out the far references.
  const int r = lg % a.nrefs, rem = lg / a.nrefs;
  const int sby = a.row0 + rem / a.sb_cols, sbx = rem % a.sb_cols;
  const int ox = 64 * sbx, oy = 64 * sby;
  if (a.kmap && r >= a.kmap[sby * a.sb_cols + sbx]) return;  // reference pruned here: whole workgroup
  const uint32_t* field = a.field[r];

  if (tid < kRegion * kRegion) {
    const int cx = clampi(8 * sbx - 1 + tid % kRegion, 0, a.cols - 1), cy = clampi(8 * sby - 1 + tid / kRegion, 0, a.rows - 1);
    mv[tid] = field[cy * a.cols + cx];
  }
  for (int b = tid; b < kBlocks; b += 256) best[b] = 0xffffffffu, seen[b][0] = seen[b][1] = 0;
  // The cell -> (block, candidate) tables, read once per MV below: from LDS.
  for (int i = tid; i < kRegion * kRegion + 1; i += 256) cell_first[i] = g_cell_first[i];
  for (int i = tid; i < kTasks; i += 256) cell_tasks[i] = g_cell_tasks[i];
  __syncthreads();
  // Distinct MVs, numbered in cell order (wave 0); each cell's class is its first cell with that MV.
  const int c0 = lane, c1 = lane + 64;
  int rep0 = c0, rep1 = c1;
  if (wave == 0) {
    // Each cell's first cell with the same MV: cell c's MV broadcast from the lane that holds it.
    const uint32_t m0 = mv[c0], m1 = c1 < 100 ? mv[c1] : 0u;
    for (int c = 0; c < 99; c++) {
      const uint32_t mc = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(c < 64 ? m0 : m1), c & 63));
      if (c < c0 && rep0 == c0 && m0 == mc) rep0 = c;
      if (c < c1 && rep1 == c1 && m1 == mc) rep1 = c;
    }
    const uint64_t b0 = __ballot(rep0 == c0), b1 = __ballot(c1 < 100 && rep1 == c1);
    const uint64_t below = (1ull << lane) - 1ull;
    if (rep0 == c0) {
      const int u = __popcll(b0 & below);
      cuid[c0] = u, urep[u] = c0;
    }
    if (c1 < 100 && rep1 == c1) {
      const int u = __popcll(b0) + __popcll(b1 & below);
      cuid[c1] = u, urep[u] = c1;
    }
    if (lane == 0) nuniq = __popcll(b0) + __popcll(b1);
  }
  __syncthreads();
  if (wave == 0) {
    if (rep0 != c0) cuid[c0] = cuid[rep0];
    if (c1 < 100 && rep1 != c1) cuid[c1] = cuid[rep1];
  }
  __syncthreads();

  // This lane's 8x8 square of the superblock: rows as two words.
  const int sqx = lane & 7, sqy = lane >> 3;
  uint32_t cw[8][2];
#pragma unroll
  for (int i = 0; i < 8; i++) {
    const uint32_t* q = reinterpret_cast<const uint32_t*>(a.src.p + static_cast<ptrdiff_t>(oy + 8 * sqy + i) * a.src.stride + ox + 8 * sqx);
    cw[i][0] = q[0], cw[i][1] = q[1];
  }
  TilePair cpair[2];  // the square's tile rows as packed pairs (tiles (ty, 0) low, (ty, 1) high)
#pragma unroll
  for (int ty = 0; ty < 2; ty++) {
    const uint32_t ca[4] = {cw[4 * ty][0], cw[4 * ty + 1][0], cw[4 * ty + 2][0], cw[4 * ty + 3][0]};
    const uint32_t cb[4] = {cw[4 * ty][1], cw[4 * ty + 1][1], cw[4 * ty + 2][1], cw[4 * ty + 3][1]};
    cpair[ty] = tile_pair(ca, cb);
  }
  const DPlane ref = a.ref[r];
  const BPlane rb = as_buffer(ref);
  const int U = nuniq;
  uint32_t* const mysc = sc[wave];
  // The 8 rows of the lane's square under MV m: 12 aligned bytes each (row i at soffset i * stride,
  // a multiple of 4), and the byte shift of the square's first sample.
  auto load_rows = [&](uint32_t m, uint32_t (&d)[8][3], int& sh) {
    const int mx = mv_x(m), my = mv_y(m);
    const int X = clamp_origin(ox + (mx >> 2), ref.w, 64), Y = clamp_origin(oy + (my >> 2), ref.h, 64);
    const int off = static_cast<int>(a.phase_bytes) * ((my & 3) * 4 + (mx & 3)) + corner_offset(rb, Y + 8 * sqy, X + 8 * sqx);
    sh = off & 3;
#pragma unroll
    for (int i = 0; i < 8; i++) bload12(rb, off & ~3, i * rb.stride, d[i][0], d[i][1], d[i][2]);
  };
  // Each wave takes the distinct MVs u = wave, wave + 4, ... on its own (its own LDS scores; the
  // block minima are order-free atomics).
  for (int u = wave; u < U; u += 4) {
    uint32_t dn[8][3];
    int shn = 0;
    load_rows(mv[urep[u]], dn, shn);
    uint32_t pw[8][2];
#pragma unroll
    for (int i = 0; i < 8; i++) pw[i][0] = __builtin_amdgcn_alignbyte(dn[i][1], dn[i][0], shn), pw[i][1] = __builtin_amdgcn_alignbyte(dn[i][2], dn[i][1], shn);
    // Tiles (ty, 0) and (ty, 1) of the square in one packed pass per tile row.
    uint32_t t[2][2];
#pragma unroll
    for (int ty = 0; ty < 2; ty++) {
      const uint32_t pa[4] = {pw[4 * ty][0], pw[4 * ty + 1][0], pw[4 * ty + 2][0], pw[4 * ty + 3][0]};
      const uint32_t pb[4] = {pw[4 * ty][1], pw[4 * ty + 1][1], pw[4 * ty + 2][1], pw[4 * ty + 3][1]};
      const uint32_t two = satd4x4_pair(cpair[ty], pa, pb);
      t[ty][0] = two & 0xffffu, t[ty][1] = two >> 16;
    }
    // Level 3 (8x8 and its halves): this lane's square.
    const int b3 = (21 + 8 * sqy + sqx) * 5;
    const uint32_t W = t[0][0] + t[0][1] + t[1][0] + t[1][1];
    mysc[b3] = W, mysc[b3 + 1] = t[0][0] + t[0][1], mysc[b3 + 2] = t[1][0] + t[1][1];
    mysc[b3 + 3] = t[0][0] + t[1][0], mysc[b3 + 4] = t[0][1] + t[1][1];
    // Level 2 (16x16): lanes (2a..2a+1, 2b..2b+1).
    const uint32_t hp = W + dpp<kDppXor1>(W), vp = W + dpp<kDppRor8>(W), quad = hp + dpp<kDppRor8>(hp);
    const int b2 = (5 + 4 * (sqy >> 1) + (sqx >> 1)) * 5;
    if (!(sqx & 1) && !(sqy & 1)) mysc[b2] = quad, mysc[b2 + 1] = hp, mysc[b2 + 3] = vp;
    if (!(sqx & 1) && (sqy & 1)) mysc[b2 + 2] = hp;
    if ((sqx & 1) && !(sqy & 1)) mysc[b2 + 4] = vp;
    // Level 1 (32x32).
    const uint32_t hp2 = quad + dpp<kDppXor2>(quad), vp2 = xor_add(quad, 16), quad2 = xor_add(hp2, 16);
    const int b1 = (1 + 2 * (sqy >> 2) + (sqx >> 2)) * 5;
    if (!(sqx & 3) && !(sqy & 3)) mysc[b1] = quad2, mysc[b1 + 1] = hp2, mysc[b1 + 3] = vp2;
    if (!(sqx & 3) && (sqy & 3) == 2) mysc[b1 + 2] = hp2;
    if ((sqx & 3) == 2 && !(sqy & 3)) mysc[b1 + 4] = vp2;
    // Level 0 (64x64).
    const uint32_t hp1 = quad2 + dpp<kDppHalfMirror>(quad2), vp1 = xor_add(quad2, 32), quad1 = xor_add(hp1, 32);
    if (lane == 0) mysc[0] = quad1, mysc[1] = hp1, mysc[3] = vp1;
    if (lane == 32) mysc[2] = hp1;
    if (lane == 4) mysc[4] = vp1;
    // The wave's scores are in LDS before any lane reads another lane's.
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    // Every (block, candidate) whose cell has this MV.
    const uint64_t in0 = __ballot(cuid[lane] == u), in1 = __ballot(lane + 64 < 100 && cuid[min(lane + 64, 99)] == u);
    for (int half = 0; half < 2; half++) {
      uint64_t bits = half ? in1 : in0;
      while (bits) {
        const int c = __ffsll(static_cast<long long>(bits)) - 1 + 64 * half;
        bits &= bits - 1;
        const int f0 = cell_first[c], f1 = cell_first[c + 1];
        for (int i = f0 + lane; i < f1; i += 64) {
          const uint32_t e = cell_tasks[i];
          const int b = static_cast<int>(e >> 7);
          atomicMin(&best[b], (mysc[b] << 7) | (e & 127u));
          if (a.candidates) atomicOr(&seen[b][u >> 6], 1ull << (u & 63));
        }
      }
    }
    // Every lane's reads of the scores come before the next MV's writes.
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
  __syncthreads();
  uint2* out = a.best + ((static_cast<size_t>(sby) * a.sb_cols + sbx) * kMaxRefs + r) * kBlocks;
  for (int b = tid; b < kBlocks; b += 256) {
    const uint32_t key = best[b];
    const int cell = g_cand_cell[b * kMaxCands + (key & 127u)];
    out[b] = make_uint2(key >> 7, mv[cell]);
  }
  if (tid == 0 && a.counter) count(a.counter, sbx + 7 * sby + r, static_cast<unsigned long long>(U));
  if (a.candidates) {
    // Non-repeated candidates: per block, the distinct MVs among its candidates (a repeat is skipped
    // by the reference, repeated_candidate, src/gpu/inter_common.h:65-89), marked in the task loop.
    unsigned long long n = 0;
    for (int b = tid; b < kBlocks; b += 256) n += static_cast<unsigned long long>(__popcll(seen[b][0]) + __popcll(seen[b][1]));
    for (int m = 32; m >= 1; m >>= 1) n += static_cast<unsigned long long>(__shfl_xor(static_cast<long long>(n), m, 64));
    if (lane == 0) count(a.candidates, 4 * (sbx + 7 * sby + r) + wave, n);
  }
}

}  // namespace

void add_inter(Batch<InterArgs>& b, const InterArgs& a) {
  const int nwg = a.nrefs * a.nrows * a.sb_cols;
  if (nwg > 0) b.add(a, nwg);
}
void launch_inter(const Batch<InterArgs>& b, hipStream_t s) {
  if (b.total() <= 0) return;
  k_inter<<<((b.total() + 7) / 8) * 8, 256, 0, s>>>(b);
}

// Candidate k of block b as a cell of the 10 x 10 region around the superblock (region column / row
// = superblock cell + 1): the block's cells in raster order, then left, above, right, below,
// above-left, above-right, below-left, below-right (candidate_mv_index, src/gpu/inter_common.h:230-243,
// before its clamping, which the region's clamped load reproduces).
static int region_cell(int b, int k) {
  const int sq = b / 5, kind = b % 5;
  const int level = sq >= 21 ? 3 : sq >= 5 ? 2 : sq >= 1 ? 1 : 0;
  const int first = ((1 << (2 * level)) - 1) / 3, i = sq - first, n = 1 << level, size = 64 >> level;
  int x = (i % n) * size, y = (i / n) * size, w = size, h = size;
  if (kind == 1 || kind == 2) h = size / 2, y += kind == 2 ? size / 2 : 0;
  if (kind == 3 || kind == 4) w = size / 2, x += kind == 4 ? size / 2 : 0;
  const int across = (x + w - 1) / 8 - x / 8 + 1, down = (y + h - 1) / 8 - y / 8 + 1, cells = across * down;
  int cx, cy;
  if (k < cells) {
    cx = x / 8 + k % across, cy = y / 8 + k / across;
  } else {
    const int m = k - cells, left = x / 8 - 1, right = (x + w) / 8, top = y / 8 - 1, bottom = (y + h) / 8;
    static const int hx[8] = {-1, 0, 1, 0, -1, 1, -1, 1}, vy[8] = {0, -1, 0, 1, -1, -1, 1, 1};
    cx = hx[m] < 0 ? left : hx[m] > 0 ? right : x / 8;
    cy = vy[m] < 0 ? top : vy[m] > 0 ? bottom : y / 8;
  }
  return (cy + 1) * kRegion + (cx + 1);
}

static int cand_total(int b) {
  const int sq = b / 5, kind = b % 5;
  const int level = sq >= 21 ? 3 : sq >= 5 ? 2 : sq >= 1 ? 1 : 0;
  const int first = ((1 << (2 * level)) - 1) / 3, i = sq - first, n = 1 << level, size = 64 >> level;
  int x = (i % n) * size, y = (i / n) * size, w = size, h = size;
  if (kind == 1 || kind == 2) h = size / 2, y += kind == 2 ? size / 2 : 0;
  if (kind == 3 || kind == 4) w = size / 2, x += kind == 4 ? size / 2 : 0;
  return ((x + w - 1) / 8 - x / 8 + 1) * ((y + h - 1) / 8 - y / 8 + 1) + 8;
}

void upload_inter_tables(hipStream_t s) {
  static uint16_t first[kRegion * kRegion + 1], tasks[kTasks];
  static uint8_t cand[kBlocks * kMaxCands];
  int ccount[kRegion * kRegion] = {};
  int total = 0;
  for (int b = 0; b < kBlocks; b++)
    for (int k = 0; k < cand_total(b); k++) {
      const int c = region_cell(b, k);
      cand[b * kMaxCands + k] = static_cast<uint8_t>(c);
      ccount[c]++, total++;
    }
  int off = 0;
  for (int c = 0; c < kRegion * kRegion; c++) first[c] = static_cast<uint16_t>(off), off += ccount[c];
  first[kRegion * kRegion] = static_cast<uint16_t>(off);
  int fill[kRegion * kRegion] = {};
  for (int b = 0; b < kBlocks; b++)
    for (int k = 0; k < cand_total(b); k++) {
      const int c = region_cell(b, k);
      tasks[first[c] + fill[c]++] = static_cast<uint16_t>((b << 7) | k);
    }
  if (total != kTasks || off != kTasks) throw std::logic_error("candidate task table size");
  (void)hipMemcpyToSymbolAsync(HIP_SYMBOL(g_cell_first), first, sizeof(first), 0, hipMemcpyHostToDevice, s);
  (void)hipMemcpyToSymbolAsync(HIP_SYMBOL(g_cell_tasks), tasks, sizeof(tasks), 0, hipMemcpyHostToDevice, s);
  (void)hipMemcpyToSymbolAsync(HIP_SYMBOL(g_cand_cell), cand, sizeof(cand), 0, hipMemcpyHostToDevice, s);
}

}  // namespace celv2hip
