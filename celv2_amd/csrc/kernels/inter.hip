// Single-reference inter scores of a superblock: for each of the 425 candidate blocks, the first
// least SATD over its MV candidates (best_both's per-reference loop, src/gpu/kernels_ref.cpp:51-66)
// and that MV.
//
// The candidates of all 425 blocks are MVs of the 10 x 10 cells of the level-0 field around the
// superblock (its 64 cells and the ring, src/gpu/inter_common.h:57-78), and every candidate block's
// 4x4 tiles lie on the superblock's 4x4 grid. So per distinct MV the kernel scores the whole 64x64
// area once - 256 tile SATDs - and sums tiles up the block quadtree; each block's cost for any
// candidate is then a lookup. The tile transforms run on the matrix cores (hadamard.h): per MV, 16
// products of HH with the source tiles' halves, each continued by the product of -HH with the
// prediction's, so the coefficients of cur - pred come out exact in f32 and the VALU only forms the
// prediction operands and sums |coefficients| (by v_permlane16/32_swap across lane rows and DPP).
// A workgroup takes one (superblock, reference); each of its 4 waves takes every fourth distinct MV
// on its own, with no workgroup barrier between MVs, the ref rows through buffer loads and the
// cell -> (block, candidate) tables from LDS. Within a block, candidate k's cost is (SATD << 7 | k)
// and the minimum of those keys (LDS atomicMin, order-free) is the first strict minimum in candidate
// order, exactly the
// reference's choice (a repeated MV costs the same as its first occurrence and never wins; the
// lambda * (3 + r) term is the same for every candidate of a reference and cannot change the order).
#include <stdexcept>
#include <vector>

#include "common.h"
#include "hadamard.h"
#include "launch.h"

namespace celv2hip {

namespace {

__device__ uint16_t g_cell_first[kRegion * kRegion + 1];  // tasks of region cell c: [first[c], first[c + 1])
__device__ uint16_t g_cell_tasks[kTasks];                 // b << 7 | k
__device__ uint8_t g_cand_cell[kBlocks * kMaxCands];     // region cell of candidate k of block b

template <int kCtrl>
__device__ __forceinline__ float fdpp(float v) { return as_f32(dpp<kCtrl>(as_u32f(v))); }
// v + v of lane ^ 16 (m = 16) or lane ^ 32 (m = 32), for the integer-valued float scores.
__device__ __forceinline__ float fswap_add(float v, int m) {
  const uint32_t x = as_u32f(v);
  const auto r = m == 16 ? __builtin_amdgcn_permlane16_swap(x, x, false, false) : __builtin_amdgcn_permlane32_swap(x, x, false, false);
  return as_f32(r[0]) + as_f32(r[1]);
}

__global__ void __launch_bounds__(256, 6) k_inter(Batch<InterArgs> batch) {
  __shared__ uint32_t mv[kRegion * kRegion];
  __shared__ int cuid[kRegion * kRegion];
  __shared__ int urep[kRegion * kRegion];
  __shared__ uint32_t best[kBlocks];
  __shared__ unsigned long long seen[kBlocks][2];  // MV classes among each block's candidates (counting only)
  __shared__ uint32_t sc[4][kBlocks];
  __shared__ uint16_t cell_first[kRegion * kRegion + 1];
  __shared__ uint16_t cell_tasks[kTasks];
  __shared__ uint32_t ukey[kRegion * kRegion];
  __shared__ uint8_t uord[kRegion * kRegion];
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
    // The order the distinct MVs are scored in: by quarter-pel phase, then the prediction's row and
    // column, so the MVs scored together (the 4 waves take consecutive ones) read the same phase
    // plane at nearby rows and share cache lines. A distinct MV has a distinct key (|MV| < 256
    // samples at level 0: 10 bits each for the row and column offsets). The per-block minima are
    // order-free, so the order changes no decision.
    const int U = nuniq;
    for (int u = lane; u < U; u += 64) {
      const uint32_t m = mv[urep[u]];
      const int mx = mv_x(m), my = mv_y(m);
      ukey[u] = static_cast<uint32_t>((my & 3) * 4 + (mx & 3)) << 20 | static_cast<uint32_t>((my >> 2) + 512) << 10 |
                static_cast<uint32_t>((mx >> 2) + 512);
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const uint32_t k0 = lane < U ? ukey[lane] : 0xffffffffu, k1 = lane + 64 < U ? ukey[lane + 64] : 0xffffffffu;
    int r0 = 0, r1 = 0;
    for (int v = 0; v < U; v++) {
      const uint32_t kv = ukey[v];
      r0 += kv < k0, r1 += kv < k1;
    }
    if (lane < U) uord[r0] = static_cast<uint8_t>(lane);
    if (lane + 64 < U) uord[r1] = static_cast<uint8_t>(lane + 64);
  }
  __syncthreads();

  // MV-independent operands of the products. Lane l = j + 16 g (j = l & 15, g = l >> 4) feeds
  // column j of products 2p and 2p + 1 (p = 0..7) with row g of a tile: the left (e = 0) and right
  // (e = 1) tile of tile row j >> 3 of 8x8 square (j & 7, p), superblock row 8p + 4 (j >> 3) + g,
  // columns 8 (j & 7) + 4e .. + 3. One 12-byte load per lane and p covers both products, and each
  // load instruction spans 8 rows of 64 (+4) bytes.
  const int j = lane & 15, g = lane >> 4;
  const bool upper = j < 8;  // the square's upper tile row; lanes j and j + 8 share a square
  const int prow = 4 * (j >> 3) + g, pcol = 8 * (j & 7);
  const half4 hh = hadamard_operand(lane, false), nhh = hadamard_operand(lane, true);
  half4 cb[16];  // the source tiles' rows as halves 1024 + v
#pragma unroll
  for (int p = 0; p < 8; p++) {
    const uint32_t* q = reinterpret_cast<const uint32_t*>(a.src.p + static_cast<ptrdiff_t>(oy + 8 * p + prow) * a.src.stride + ox + pcol);
    cb[2 * p] = row_halves(q[0]), cb[2 * p + 1] = row_halves(q[1]);
  }
  // After the reductions below, lane (j, g) of register set s holds square (j & 7, 4s + g); for the
  // quadtree above 8x8 the two sets merge into one register, lane (j, g) = square (sx, sy):
  const int sx = j & 7, sy = g + 4 * (j >> 3);
  const DPlane ref = a.ref[r];
  const BPlane rb = as_buffer(ref);
  const int U = nuniq;
  uint32_t* const mysc = sc[wave];
  // Each wave takes the distinct MVs si = wave, wave + 4, ... of the scoring order on its own (its own
  // LDS scores; the block minima are order-free atomics).
  for (int si = wave; si < U; si += 4) {
    const int u = uord[si];
    const uint32_t m = mv[urep[u]];
    const int mx = mv_x(m), my = mv_y(m);
    const int X = clamp_origin(ox + (mx >> 2), ref.w, 64), Y = clamp_origin(oy + (my >> 2), ref.h, 64);
    const int off = phase_offset((my & 3) * 4 + (mx & 3), a.phase_bytes) + corner_offset_lane(rb, Y + prow, X + pcol);
    const int sh = off & 3;  // the same for every lane (X + pcol = X mod 4)
    // Per square row p: the lane's quarter (4 coefficients) of the left tile's SATD, of the right
    // tile's, and their sum; each tile's transform is HH (1024 + cur) - HH (1024 + pred), exact in
    // f32. Then the sums over the row groups g, 4 square rows per register set.
    float rl[2], rr[2], rh[2];
#pragma unroll
    for (int s = 0; s < 2; s++) {
      uint32_t d[4][3];
#pragma unroll
      for (int i = 0; i < 4; i++) bload12(rb, off & ~3, 8 * (4 * s + i) * rb.stride, d[i][0], d[i][1], d[i][2]);
      float ql[4], qr[4], qh[4];
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const int p = 4 * s + i;
        const uint32_t w0 = __builtin_amdgcn_alignbyte(d[i][1], d[i][0], sh), w1 = __builtin_amdgcn_alignbyte(d[i][2], d[i][1], sh);
        const f32x4 z = {0.f, 0.f, 0.f, 0.f};
        const f32x4 c0 = mfma16(nhh, row_halves(w0), mfma16(hh, cb[2 * p], z));
        const f32x4 c1 = mfma16(nhh, row_halves(w1), mfma16(hh, cb[2 * p + 1], z));
        ql[i] = abs_sum4(c0), qr[i] = abs_sum4(c1), qh[i] = ql[i] + qr[i];
      }
      rl[s] = sum_halves32(sum_rows16(ql[0], ql[1]), sum_rows16(ql[2], ql[3]));
      rr[s] = sum_halves32(sum_rows16(qr[0], qr[1]), sum_rows16(qr[2], qr[3]));
      rh[s] = sum_halves32(sum_rows16(qh[0], qh[1]), sum_rows16(qh[2], qh[3]));
    }
    // Level 3 (8x8 and its halves). Lane j of a set holds the upper tile row's sums (rh: top half;
    // rl, rr: the left and right tiles), lane j + 8 the lower row's.
    float w[2];
#pragma unroll
    for (int s = 0; s < 2; s++) {
      w[s] = rh[s] + fdpp<kDppRor8>(rh[s]);
      // Left half in the upper lane, right half in the lower: each lane sends its partner the tile
      // the partner needs.
      const float lr = (upper ? rl[s] : rr[s]) + fdpp<kDppRor8>(upper ? rr[s] : rl[s]);
      const int b3 = (21 + 8 * (4 * s + g) + sx) * 5;
      if (upper) mysc[b3] = as_u32f(w[s]);
      mysc[b3 + (upper ? 1 : 2)] = as_u32f(rh[s]);
      mysc[b3 + (upper ? 3 : 4)] = as_u32f(lr);
    }
    const float W = upper ? w[0] : w[1];
    // Level 2 (16x16): squares sx ^ 1 at lane ^ 1, sy ^ 1 at lane ^ 16.
    const float hp = W + fdpp<kDppXor1>(W), vp = fswap_add(W, 16), quad = fswap_add(hp, 16);
    const int b2 = (5 + 4 * (sy >> 1) + (sx >> 1)) * 5;
    if (!(sx & 1) && !(sy & 1)) mysc[b2] = as_u32f(quad), mysc[b2 + 1] = as_u32f(hp), mysc[b2 + 3] = as_u32f(vp);
    if (!(sx & 1) && (sy & 1)) mysc[b2 + 2] = as_u32f(hp);
    if ((sx & 1) && !(sy & 1)) mysc[b2 + 4] = as_u32f(vp);
    // Level 1 (32x32): sx ^ 2 at lane ^ 2, sy ^ 2 at lane ^ 32.
    const float hp2 = quad + fdpp<kDppXor2>(quad), vp2 = fswap_add(quad, 32), quad2 = fswap_add(hp2, 32);
    const int b1 = (1 + 2 * (sy >> 2) + (sx >> 2)) * 5;
    if (!(sx & 3) && !(sy & 3)) mysc[b1] = as_u32f(quad2), mysc[b1 + 1] = as_u32f(hp2), mysc[b1 + 3] = as_u32f(vp2);
    if (!(sx & 3) && (sy & 3) == 2) mysc[b1 + 2] = as_u32f(hp2);
    if ((sx & 3) == 2 && !(sy & 3)) mysc[b1 + 4] = as_u32f(vp2);
    // Level 0 (64x64): the other 4 squares across within the lane's 8 (values are uniform within
    // quads), the other 4 square rows at lane ^ 8.
    const float hp1 = quad2 + fdpp<kDppHalfMirror>(quad2), vp1 = quad2 + fdpp<kDppRor8>(quad2), quad1 = hp1 + fdpp<kDppRor8>(hp1);
    if (lane == 0) mysc[0] = as_u32f(quad1), mysc[1] = as_u32f(hp1), mysc[3] = as_u32f(vp1);
    if (lane == 8) mysc[2] = as_u32f(hp1);
    if (lane == 4) mysc[4] = as_u32f(vp1);
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
          atomicMin(&best[b], (static_cast<uint32_t>(as_f32(mysc[b])) << 7) | (e & 127u));
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
// above-left, above-right, below-left, below-right (candidate_mv_index, src/gpu/inter_common.h:65-78,
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

// Builds the tables in local memory and copies them to this device's symbols (synchronous).
hipError_t upload_inter_tables() {
  std::vector<uint16_t> first(kRegion * kRegion + 1), tasks(kTasks);
  std::vector<uint8_t> cand(kBlocks * kMaxCands);
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
  hipError_t rc = hipMemcpyToSymbol(HIP_SYMBOL(g_cell_first), first.data(), first.size() * sizeof(uint16_t));
  if (rc == hipSuccess) rc = hipMemcpyToSymbol(HIP_SYMBOL(g_cell_tasks), tasks.data(), tasks.size() * sizeof(uint16_t));
  if (rc == hipSuccess) rc = hipMemcpyToSymbol(HIP_SYMBOL(g_cand_cell), cand.data(), cand.size());
  return rc;
}

}  // namespace celv2hip
