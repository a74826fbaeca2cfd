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
constexpr int kLevels = 5;          // src/gpu/inter_common.h:44
constexpr int kPhases = 16;         // src/gpu/inter_common.h:11
constexpr int kModes = 13;          // src/gpu/intra_common.h:42
constexpr int kMaxRefs = 7;         // src/gpu/inter_common.h:46
constexpr int kBlocks = 425;        // src/gpu/sb_blocks.h:10
constexpr int kSquares = 85;        // src/gpu/sb_blocks.h:9
constexpr int kRange = 8;           // src/gpu/inter_common.h:35, 40
constexpr int kPelPenalty = 4;      // src/gpu/inter_common.h:55
constexpr int kNarrowRange = 2;     // the pruned form's narrow full search (+-2 instead of +-kRange)
constexpr int kQpelPenalty = 16;    // src/gpu/inter_common.h:140
constexpr int kSplitBits = 2;       // src/gpu/intra_common.h:33
constexpr int kIntraOnInter = 32;   // src/gpu/inter_common.h:50
constexpr int kNewmvBits = 3;       // src/gpu/inter_common.h:45
constexpr int kMaxCands = 72;       // candidates of the 64x64 block: 64 cells + 8 neighbours
constexpr int kRegion = 10;         // 10x10 8x8-cells: a superblock's cells and the ring around it
constexpr int kTasks = 4296;        // (block, candidate) pairs of a superblock

// A bordered plane: p points at sample (0, 0); rows are `stride` bytes apart; samples at
// y in [-kBorder, h + kBorder), x in [-kBorder, w + kBorder) are readable.
struct DPlane {
  uint8_t* p;
  int stride, w, h;
};

// v clamped into [lo, hi] (lo <= hi at every use), written as the median of (v, lo, hi) so that a
// per-lane clamp is one v_med3_i32.
__device__ __forceinline__ int clampi(int v, int lo, int hi) { return max(min(v, lo), min(max(v, lo), hi)); }
// The same for wave-uniform values, in the form the compiler keeps in scalar instructions.
__device__ __forceinline__ int clamp_uniform(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// Origin of a size-wide window moved into the bordered range; reads through it equal clamped reads
// (every sample beyond the border repeats the edge it would clamp to).
__device__ __forceinline__ int clamp_origin(int o, int extent, int size) {
  return clampi(o, -kBorder, extent + kBorder - size);
}
__device__ __forceinline__ int clamp_origin_uniform(int o, int extent, int size) {
  return clamp_uniform(o, -kBorder, extent + kBorder - size);
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
// s * v for a wave-uniform s and a per-lane v, both below 2^24, as the full-rate v_mul_u32_u24.
// (Written as the instruction itself: from __umul24 the compiler emits the quarter-rate 32-bit
// multiply whenever it can bound one operand.)
__device__ __forceinline__ uint32_t umul24_sv(uint32_t s, uint32_t v) {
  uint32_t r;
  asm("v_mul_u32_u24 %0, %1, %2" : "=v"(r) : "s"(s), "v"(v));
  return r;
}
// The same for per-lane coordinates inside the bordered range (0 <= y + kBorder < 2^24).
__device__ __forceinline__ int corner_offset_lane(const BPlane& b, int y, int x) {
  return static_cast<int>(umul24_sv(static_cast<uint32_t>(b.stride), static_cast<uint32_t>(y + kBorder))) + x + kBorder;
}
// Byte offset of quarter-pel phase `phase` (0..15) from phase 0, for planes phase_bytes apart;
// phase_bytes is a multiple of 64 (a plane's rows are a multiple of 64 bytes, its slack 512), and
// phase_bytes / 64 < 2^24.
__device__ __forceinline__ int phase_offset(int phase, long long phase_bytes) {
  return static_cast<int>(umul24_sv(static_cast<uint32_t>(phase_bytes >> 6), static_cast<uint32_t>(phase)) << 6);
}
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

// Quarter-pel MVs are packed as (int16 mvx) | (int16 mvy) << 16.
__device__ __forceinline__ uint32_t pack_mv(int x, int y) { return (static_cast<uint32_t>(x) & 0xffffu) | (static_cast<uint32_t>(y) << 16); }
__device__ __forceinline__ int mv_x(uint32_t m) { return static_cast<int16_t>(m & 0xffffu); }
__device__ __forceinline__ int mv_y(uint32_t m) { return static_cast<int16_t>(m >> 16); }

// Candidate-block geometry (src/gpu/sb_blocks.h:17-27).
// Event counters are spread over 64 slots 16 words (128 bytes) apart, so that workgroups finishing
// together do not queue on one address; the host sums the slots.
constexpr int kCounterSlots = 64, kCounterStride = 16;
constexpr int kCounterWords = kCounterSlots * kCounterStride;  // one counter
constexpr int kCounterKinds = 6;                                // the engine's counters
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

// Intra mode groups of the pruned form: 0 = DC, V, H (always scored); 1 = smooth, smooth V, smooth H,
// Paeth; 2 = the directional modes at 45, 67 and 113 degrees; 3 = at 135, 157 and 203 degrees. A
// superblock's mask has bit g set when group g is scored; a block's intra choice is then the least
// cost over the scored modes.
constexpr int kIntraGroups = 4;
constexpr uint8_t kAllGroups = 0xF;
__host__ __device__ __forceinline__ int mode_group(int m) { return m < 3 ? 0 : m < 7 ? 1 : m < 10 ? 2 : 3; }
__host__ __device__ __forceinline__ bool mode_scored(uint32_t mask, int m) { return ((mask | 1u) >> mode_group(m)) & 1u; }

// Cross-lane moves within 16-lane rows (DPP, no LDS traffic): quad_perm [1,0,3,2] = lane ^ 1,
// quad_perm [2,3,0,1] = lane ^ 2, row_ror:8 = lane ^ 8 (within a row), row_half_mirror maps lane i of
// each 8 to 7 - i (the other quad of the 8, for values uniform within quads), row_ror:4 = lane + 4
// mod 16, row_mirror maps lane i of each 16 to 15 - i (the other 8 of the row).
constexpr int kDppXor1 = 0xB1, kDppXor2 = 0x4E, kDppRor4 = 0x124, kDppRor8 = 0x128, kDppMirror = 0x140,
              kDppHalfMirror = 0x141;
// Every control here reads a lane of the same row, so no lane is out of range and the zero fill
// (bound_ctrl) never applies; with it, the compiler folds the move into the consuming v_add / v_min
// as a DPP operand instead of issuing a separate v_mov_b32_dpp.
template <int kCtrl>
__device__ __forceinline__ uint32_t dpp(uint32_t v) {
  return static_cast<uint32_t>(__builtin_amdgcn_update_dpp(0, static_cast<int>(v), kCtrl, 0xf, 0xf, true));
}

// Wave-wide minimum of a 32-bit key (all 64 lanes take part; the result is wave-uniform).
// min(a, b) of wave-uniform values on the scalar unit (left to itself the compiler takes the vector
// unit's three-operand minimum, which costs the busier unit two copies and the minimum).
__device__ __forceinline__ uint32_t scalar_min(uint32_t a, uint32_t b) {
  uint32_t r;
  asm("s_min_u32 %0, %1, %2" : "=s"(r) : "s"(a), "s"(b));
  return r;
}
__device__ __forceinline__ uint32_t wave_min_u32(uint32_t v) {
  v = min(v, dpp<kDppXor1>(v));
  v = min(v, dpp<kDppXor2>(v));
  v = min(v, dpp<kDppRor4>(v));
  v = min(v, dpp<kDppRor8>(v));  // every lane: its row's minimum
  const uint32_t r0 = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(v), 0));
  const uint32_t r1 = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(v), 16));
  const uint32_t r2 = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(v), 32));
  const uint32_t r3 = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(v), 48));
  return scalar_min(scalar_min(r0, r1), scalar_min(r2, r3));
}

// The same when v is already equal over each aligned quad of lanes (the xor-1 and xor-2 steps would
// change nothing).
__device__ __forceinline__ uint32_t wave_min_quads_u32(uint32_t v) {
  v = min(v, dpp<kDppRor4>(v));
  v = min(v, dpp<kDppRor8>(v));  // every lane: its row's minimum
  const uint32_t r0 = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(v), 0));
  const uint32_t r1 = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(v), 16));
  const uint32_t r2 = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(v), 32));
  const uint32_t r3 = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(v), 48));
  return scalar_min(scalar_min(r0, r1), scalar_min(r2, r3));
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
