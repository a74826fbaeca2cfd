// Device-side definitions shared by the decision kernels (gfx950, wave64).
//
// Planes live in HBM with a replicated border of kBorder samples on every side, so a window of up
// to kBorder x kBorder samples at any origin reads exactly what the reference's clamped accessor
// (Plane::at, src/gpu/intra_common.h:17-21) reads once its origin is clamped into the bordered range
// (clamp_origin). Quarter-pel phase planes (src/gpu/intra_common.h:22-25) are bordered the same way:
// their border holds the phase value at the clamped position, as at_q() reads it.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace celv2hip {

constexpr int kBorder = 64;         // >= the widest window read at one origin (a 64x64 block)
constexpr int kSlack = 256;         // bytes before / after each plane allocation for aligned over-reads
constexpr int kLevels = 5;          // src/gpu/inter_common.h:209
constexpr int kPhases = 16;         // src/gpu/inter_common.h:176
constexpr int kModes = 13;          // src/gpu/intra_common.h:42
constexpr int kMaxRefs = 7;         // src/gpu/inter_common.h:211
constexpr int kBlocks = 425;        // src/gpu/sb_blocks.h:10
constexpr int kSquares = 85;        // src/gpu/sb_blocks.h:9
constexpr int kRange = 8;           // src/gpu/inter_common.h:200, 205
constexpr int kPelPenalty = 4;      // src/gpu/inter_common.h:220
constexpr int kQpelPenalty = 16;    // src/gpu/inter_common.h:305
constexpr int kSplitBits = 2;       // src/gpu/intra_common.h:33
constexpr int kIntraOnInter = 32;   // src/gpu/inter_common.h:215
constexpr int kNewmvBits = 3;       // src/gpu/inter_common.h:210
constexpr int kMaxCands = 72;       // candidates of the 64x64 block: 64 cells + 8 neighbours
constexpr int kRegion = 10;         // 10x10 8x8-cells: a superblock's cells and the ring around it
constexpr int kTasks = 4296;        // (block, candidate) pairs of a superblock

// A bordered plane: p points at sample (0, 0); rows are `stride` bytes apart; samples at
// y in [-kBorder, h + kBorder), x in [-kBorder, w + kBorder) are readable.
struct DPlane {
  uint8_t* p;
  int stride, w, h;
};

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// Origin of a size-wide window moved into the bordered range; reads through it equal clamped reads
// (every sample beyond the border repeats the edge it would clamp to).
__device__ __forceinline__ int clamp_origin(int o, int extent, int size) {
  return clampi(o, -kBorder, extent + kBorder - size);
}

// 4 bytes starting at an arbitrary address, from two aligned dword loads (never past the bytes the
// two dwords cover, which the plane's slack keeps inside the allocation).
__device__ __forceinline__ uint32_t load4u(const uint8_t* a) {
  const uintptr_t u = reinterpret_cast<uintptr_t>(a);
  const uint32_t* q = reinterpret_cast<const uint32_t*>(u & ~uintptr_t(3));
  const uint32_t lo = q[0], hi = q[1];
  return __builtin_amdgcn_alignbyte(hi, lo, static_cast<uint32_t>(u & 3));
}
// 8 bytes starting at an arbitrary address (three aligned dwords).
__device__ __forceinline__ void load8u(const uint8_t* a, uint32_t& x0, uint32_t& x1) {
  const uintptr_t u = reinterpret_cast<uintptr_t>(a);
  const uint32_t* q = reinterpret_cast<const uint32_t*>(u & ~uintptr_t(3));
  const uint32_t d0 = q[0], d1 = q[1], d2 = q[2];
  const uint32_t s = static_cast<uint32_t>(u & 3);
  x0 = __builtin_amdgcn_alignbyte(d1, d0, s);
  x1 = __builtin_amdgcn_alignbyte(d2, d1, s);
}

// A bordered plane as a buffer resource: byte offsets count from its (-kBorder, -kBorder) corner, so
// every offset inside the bordered range is non-negative; loads take the row part in an SGPR.
struct BPlane {
  __amdgpu_buffer_rsrc_t r;
  int stride;
};
__device__ __forceinline__ BPlane as_buffer(const DPlane& p) {
  uint8_t* corner = p.p - static_cast<ptrdiff_t>(kBorder) * p.stride - kBorder;
  return BPlane{__builtin_amdgcn_make_buffer_rsrc(corner, static_cast<short>(0), 0x7fffffff, 0x00020000), p.stride};
}
// Offset of sample (y, x) of a bordered plane from its corner.
__device__ __forceinline__ int corner_offset(const BPlane& b, int y, int x) { return (y + kBorder) * b.stride + x + kBorder; }
// 12 aligned bytes at voffset + soffset (both multiples of 4).
__device__ __forceinline__ void bload12(const BPlane& b, int voff, int soff, uint32_t& d0, uint32_t& d1, uint32_t& d2) {
  const auto v = __builtin_amdgcn_raw_buffer_load_b96(b.r, voff, soff, 0);
  d0 = v[0], d1 = v[1], d2 = v[2];
}
__device__ __forceinline__ void bload8(const BPlane& b, int voff, int soff, uint32_t& d0, uint32_t& d1) {
  const auto v = __builtin_amdgcn_raw_buffer_load_b64(b.r, voff, soff, 0);
  d0 = v[0], d1 = v[1];
}
__device__ __forceinline__ uint32_t bload4(const BPlane& b, int voff, int soff) {
  return __builtin_amdgcn_raw_buffer_load_b32(b.r, voff, soff, 0);
}

__device__ __forceinline__ uint32_t sad4(uint32_t a, uint32_t b, uint32_t acc) { return __builtin_amdgcn_sad_u8(a, b, acc); }

// Sum of |4x4 Walsh-Hadamard coefficients| of d (row-major), unnormalised
// (hadamard4x4, src/gpu/intra_common.h:131-146).
__device__ __forceinline__ int wht4_abs_sum(const int d[16]) {
  int r[16];
#pragma unroll
  for (int i = 0; i < 4; i++) {
    const int s01 = d[4 * i] + d[4 * i + 1], d01 = d[4 * i] - d[4 * i + 1];
    const int s23 = d[4 * i + 2] + d[4 * i + 3], d23 = d[4 * i + 2] - d[4 * i + 3];
    r[4 * i] = s01 + s23, r[4 * i + 1] = d01 + d23, r[4 * i + 2] = s01 - s23, r[4 * i + 3] = d01 - d23;
  }
  int t = 0;
#pragma unroll
  for (int c = 0; c < 4; c++) {
    const int s01 = r[c] + r[4 + c], d01 = r[c] - r[4 + c], s23 = r[8 + c] + r[12 + c], d23 = r[8 + c] - r[12 + c];
    t += abs(s01 + s23) + abs(d01 + d23) + abs(s01 - s23) + abs(d01 - d23);
  }
  return t;
}

// SATD of one 4x4 tile from four rows of packed bytes: cur c[i] against prediction p[i].
__device__ __forceinline__ int satd4x4_rows(const uint32_t c[4], const uint32_t p[4]) {
  int d[16];
#pragma unroll
  for (int i = 0; i < 4; i++)
#pragma unroll
    for (int j = 0; j < 4; j++) d[4 * i + j] = static_cast<int>((c[i] >> (8 * j)) & 255) - static_cast<int>((p[i] >> (8 * j)) & 255);
  return wht4_abs_sum(d);
}

// ---- Two 4x4 SATDs at once in packed 16-bit arithmetic ----------------------------------------------
// Tile A's samples go to the low halves, tile B's to the high halves; every butterfly is one
// v_pk_add/sub_u16 (int16 wrap-around equals the exact value: |coefficient| <= 4080). The last
// stage uses |a + b| + |a - b| = 2 max(|a|, |b|), so each column needs 2 maxima of absolute values.
typedef short s16x2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ s16x2 as_s16x2(uint32_t v) { return __builtin_bit_cast(s16x2, v); }
__device__ __forceinline__ uint32_t as_u32(s16x2 v) { return __builtin_bit_cast(uint32_t, v); }
// Byte j of a into the low half, byte j of b into the high half (zero-extended).
template <int j>
__device__ __forceinline__ s16x2 pair_bytes(uint32_t a, uint32_t b) {
  constexpr uint32_t sel = static_cast<uint32_t>(j) | (0x0Cu << 8) | ((4u + j) << 16) | (0x0Cu << 24);
  return as_s16x2(__builtin_amdgcn_perm(b, a, sel));
}
__device__ __forceinline__ s16x2 abs2(s16x2 v) {
  const s16x2 n = s16x2{0, 0} - v;
  return __builtin_elementwise_max(v, n);
}
// The current block's two tiles as packed pairs (hoisted out of loops over predictions).
struct TilePair {
  s16x2 v[4][4];
};
__device__ __forceinline__ TilePair tile_pair(const uint32_t ca[4], const uint32_t cb[4]) {
  TilePair t;
#pragma unroll
  for (int i = 0; i < 4; i++) {
    t.v[i][0] = pair_bytes<0>(ca[i], cb[i]), t.v[i][1] = pair_bytes<1>(ca[i], cb[i]);
    t.v[i][2] = pair_bytes<2>(ca[i], cb[i]), t.v[i][3] = pair_bytes<3>(ca[i], cb[i]);
  }
  return t;
}
// Packed sums of |coefficients| / 2 of the differences cur - pred (tile A low half, B high half).
__device__ __forceinline__ s16x2 wht_pair_half(const s16x2 d[4][4]) {
  s16x2 r[4][4];
#pragma unroll
  for (int i = 0; i < 4; i++) {
    const s16x2 s01 = d[i][0] + d[i][1], d01 = d[i][0] - d[i][1], s23 = d[i][2] + d[i][3], d23 = d[i][2] - d[i][3];
    r[i][0] = s01 + s23, r[i][1] = d01 + d23, r[i][2] = s01 - s23, r[i][3] = d01 - d23;
  }
  s16x2 acc = {0, 0};
#pragma unroll
  for (int c = 0; c < 4; c++) {
    const s16x2 s01 = r[0][c] + r[1][c], d01 = r[0][c] - r[1][c], s23 = r[2][c] + r[3][c], d23 = r[2][c] - r[3][c];
    acc += __builtin_elementwise_max(abs2(s01), abs2(s23)) + __builtin_elementwise_max(abs2(d01), abs2(d23));
  }
  return acc;
}
__device__ __forceinline__ uint32_t unpack_satd_pair(s16x2 acc) {
  const uint32_t a = as_u32(acc);
  return ((a & 0xffffu) << 1) | ((a >> 16) << 17);
}
// SATD(A) | SATD(B) << 16 of the current tiles c against predicted rows pa (tile A) and pb (tile B).
__device__ __forceinline__ uint32_t satd4x4_pair(const TilePair& c, const uint32_t pa[4], const uint32_t pb[4]) {
  s16x2 d[4][4];
#pragma unroll
  for (int i = 0; i < 4; i++) {
    d[i][0] = c.v[i][0] - pair_bytes<0>(pa[i], pb[i]), d[i][1] = c.v[i][1] - pair_bytes<1>(pa[i], pb[i]);
    d[i][2] = c.v[i][2] - pair_bytes<2>(pa[i], pb[i]), d[i][3] = c.v[i][3] - pair_bytes<3>(pa[i], pb[i]);
  }
  return unpack_satd_pair(wht_pair_half(d));
}
__device__ __forceinline__ uint32_t satd4x4_pair(const uint32_t ca[4], const uint32_t pa[4], const uint32_t cb[4],
                                                 const uint32_t pb[4]) {
  return satd4x4_pair(tile_pair(ca, cb), pa, pb);
}

// Quarter-pel MVs are packed as (int16 mvx) | (int16 mvy) << 16.
__device__ __forceinline__ uint32_t pack_mv(int x, int y) { return (static_cast<uint32_t>(x) & 0xffffu) | (static_cast<uint32_t>(y) << 16); }
__device__ __forceinline__ int mv_x(uint32_t m) { return static_cast<int16_t>(m & 0xffffu); }
__device__ __forceinline__ int mv_y(uint32_t m) { return static_cast<int16_t>(m >> 16); }

// Candidate-block geometry (src/gpu/sb_blocks.h:172-187).
// Event counters are spread over 64 slots 16 words (128 bytes) apart, so that workgroups finishing
// together do not queue on one address; the host sums the slots.
constexpr int kCounterSlots = 64, kCounterStride = 16;
constexpr int kCounterWords = kCounterSlots * kCounterStride;  // one counter
constexpr int kCounterKinds = 5;                                // the engine's counters
constexpr size_t kCounterBytes = static_cast<size_t>(kCounterKinds) * kCounterWords * sizeof(unsigned long long);
__device__ __forceinline__ void count(unsigned long long* base, int slot, unsigned long long v) {
  atomicAdd(base + kCounterStride * (slot & (kCounterSlots - 1)), v);
}
__device__ __forceinline__ int level_first(int level) { return ((1 << (2 * level)) - 1) / 3; }
__device__ __forceinline__ void block_geom(int b, int& x, int& y, int& w, int& h) {
  const int sq = b / 5, kind = b - 5 * (b / 5);
  const int level = sq >= 21 ? 3 : sq >= 5 ? 2 : sq >= 1 ? 1 : 0;
  const int i = sq - level_first(level), n = 1 << level, size = 64 >> level;
  const int sx = (i & (n - 1)) * size, sy = (i >> level) * size;
  x = sx, y = sy, w = size, h = size;
  if (kind == 1 || kind == 2) h = size / 2, y = sy + (kind == 2 ? size / 2 : 0);
  if (kind == 3 || kind == 4) w = size / 2, x = sx + (kind == 4 ? size / 2 : 0);
}

__device__ __forceinline__ int mode_bits(int m) { return m == 0 ? 1 : (m < 4 ? 3 : 9); }  // intra_common.h:32

// Cross-lane moves within 16-lane rows (DPP, no LDS traffic): quad_perm [1,0,3,2] = lane ^ 1,
// quad_perm [2,3,0,1] = lane ^ 2, row_ror:8 = lane ^ 8 (within a row), row_half_mirror maps lane i of
// each 8 to 7 - i (the other quad of the 8, for values uniform within quads), row_ror:4 = lane + 4
// mod 16, row_mirror maps lane i of each 16 to 15 - i (the other 8 of the row).
constexpr int kDppXor1 = 0xB1, kDppXor2 = 0x4E, kDppRor4 = 0x124, kDppRor8 = 0x128, kDppMirror = 0x140,
              kDppHalfMirror = 0x141;
template <int kCtrl>
__device__ __forceinline__ uint32_t dpp(uint32_t v) {
  return static_cast<uint32_t>(__builtin_amdgcn_mov_dpp(static_cast<int>(v), kCtrl, 0xf, 0xf, false));
}

// Wave-wide minimum of a 32-bit key (all 64 lanes take part; the result is wave-uniform).
__device__ __forceinline__ uint32_t wave_min_u32(uint32_t v) {
  v = min(v, dpp<kDppXor1>(v));
  v = min(v, dpp<kDppXor2>(v));
  v = min(v, dpp<kDppRor4>(v));
  v = min(v, dpp<kDppRor8>(v));  // every lane: its row's minimum
  const uint32_t r0 = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(v), 0));
  const uint32_t r1 = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(v), 16));
  const uint32_t r2 = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(v), 32));
  const uint32_t r3 = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(v), 48));
  return min(min(r0, r1), min(r2, r3));
}

// Sums over aligned lane groups of 4 and of 8 (every lane of the group gets the sum).
__device__ __forceinline__ uint32_t quad_sum(uint32_t v) {
  v += dpp<kDppXor1>(v);
  return v + dpp<kDppXor2>(v);
}
__device__ __forceinline__ uint32_t oct_sum(uint32_t v) {
  v = quad_sum(v);
  return v + dpp<kDppHalfMirror>(v);
}

// XCD-aware workgroup order: consecutive logical workgroups (neighbouring blocks, which read
// overlapping reference windows) land on the same XCD and share its L2. The hardware deals
// workgroups round-robin over the 8 XCDs; the grid is padded to a multiple of 8 and a logical index
// >= n means "no work".
__device__ __forceinline__ int xcd_logical(int bid, int n) {
  const int per = (n + 7) >> 3;
  return (bid & 7) * per + (bid >> 3);
}

}  // namespace celv2hip
