// 4x4 Walsh-Hadamard SATDs on the matrix cores (gfx950 v_mfma_f32_16x16x16_f16).
//
// The reference's SATD (hadamard4x4, src/gpu/intra_common.h:131-146) is the sum of |H d H^T| over a
// 4x4 difference tile d, H the natural-order 4x4 Hadamard matrix; row-major vec(H d H^T) =
// (H (x) H) vec(d). For 16 tiles at once that is one 16x16x16 product with HH = H (x) H as the A
// operand and the tiles as the 16 columns of B. A sample v in 0..255 becomes the exact half
// 1024 + v by writing the byte 0x64 above it; the bias cancels in cur - pred, every partial sum is an
// integer below 2^24, so the f32 coefficients are exact. |x| is a free source modifier on v_add_f32.
// Where both operands are at hand, (1024 + c) - (1024 + p) = c - p is itself an exact half
// (|c - p| <= 255), so one v_pk_add_f16 per pair gives the difference tile directly.
//
// Operand layout of v_mfma_f32_16x16x16_f16 for lane l = j + 16 g (j = l & 15, g = l >> 4):
//   A: HH[j][4g .. 4g + 3]; B: column j's elements 4g .. 4g + 3 (for a tile: its row g);
//   C / D: element (4g + r, j) in component r - the coefficients 4g .. 4g + 3 of tile j.
#pragma once
#include "common.h"

namespace celv2hip {

typedef _Float16 half4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ half4 as_half4(uint32_t lo, uint32_t hi) {
  return __builtin_bit_cast(half4, static_cast<uint64_t>(hi) << 32 | lo);
}
__device__ __forceinline__ float as_f32(uint32_t v) { return __builtin_bit_cast(float, v); }
__device__ __forceinline__ uint32_t as_u32f(float v) { return __builtin_bit_cast(uint32_t, v); }

// Bytes 0, 1 (halves01) or 2, 3 (halves23) of x as the halves 1024 + byte.
constexpr uint32_t kHalfBias = 0x64646464u;
__device__ __forceinline__ uint32_t halves01(uint32_t x) { return __builtin_amdgcn_perm(kHalfBias, x, 0x04010400u); }
__device__ __forceinline__ uint32_t halves23(uint32_t x) { return __builtin_amdgcn_perm(kHalfBias, x, 0x04030402u); }
// Four packed bytes (one tile row) as the B operand.
__device__ __forceinline__ half4 row_halves(uint32_t x) { return as_half4(halves01(x), halves23(x)); }
// Low 16 bits of a and b as one dword (a low).
__device__ __forceinline__ uint32_t pack_lo16(uint32_t a, uint32_t b) { return __builtin_amdgcn_perm(b, a, 0x05040100u); }

__device__ __forceinline__ f32x4 mfma16(half4 a, half4 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x16f16(a, b, c, 0, 0, 0);
}

// This lane's A operand: HH[j][4g + k], k = 0..3, as halves +-1 (negated: -HH).
__device__ __forceinline__ half4 hadamard_operand(int lane, bool negate) {
  const int i = lane & 15, g = lane >> 4;
  uint32_t w[2];
#pragma unroll
  for (int p = 0; p < 2; p++) {
    uint32_t v = 0;
#pragma unroll
    for (int h = 0; h < 2; h++) {
      const int k = 2 * p + h;
      const int par = (__builtin_popcount((i >> 2) & g) + __builtin_popcount((i & 3) & k) + (negate ? 1 : 0)) & 1;
      v |= (par ? 0xBC00u : 0x3C00u) << (16 * h);
    }
    w[p] = v;
  }
  return as_half4(w[0], w[1]);
}

// The A operands of all 64 lanes as a constant table: one 8-byte load per lane instead of the
// parity arithmetic, for kernels that build them once per small unit of work. -HH is HH with every
// half's sign bit flipped.
struct HadamardLanes {
  uint32_t w[64][2];
  constexpr HadamardLanes() : w{} {
    for (int lane = 0; lane < 64; lane++)
      for (int p = 0; p < 2; p++) {
        uint32_t v = 0;
        for (int h = 0; h < 2; h++) {
          const int i = lane & 15, g = lane >> 4, k = 2 * p + h;
          const int par = (__builtin_popcount((i >> 2) & g) + __builtin_popcount((i & 3) & k)) & 1;
          v |= (par ? 0xBC00u : 0x3C00u) << (16 * h);
        }
        w[lane][p] = v;
      }
  }
};
__device__ constexpr HadamardLanes kHadamardLanes{};
__device__ __forceinline__ void hadamard_operands(int lane, half4& hh, half4& nhh) {
  const uint32_t w0 = kHadamardLanes.w[lane][0], w1 = kHadamardLanes.w[lane][1];
  hh = as_half4(w0, w1), nhh = as_half4(w0 ^ 0x80008000u, w1 ^ 0x80008000u);
}

// Sum of |coefficients| this lane holds (a quarter of one tile's SATD).
__device__ __forceinline__ float abs_sum4(f32x4 d) {
  return __builtin_fabsf(d[0]) + __builtin_fabsf(d[1]) + __builtin_fabsf(d[2]) + __builtin_fabsf(d[3]);
}

// Cross-row sums without LDS (gfx950 v_permlane16_swap / v_permlane32_swap exchange rows of 16
// lanes between two registers). sum_rows16(a, b): row 0 gets a0 + a1, row 1 b0 + b1, row 2 a2 + a3,
// row 3 b2 + b3 (an = row n of a). sum_halves32(a, b): lanes 0..31 get a(l) + a(l + 32), lanes
// 32..63 get b(l - 32) + b(l).
__device__ __forceinline__ float sum_rows16(float a, float b) {
  const auto r = __builtin_amdgcn_permlane16_swap(as_u32f(a), as_u32f(b), false, false);
  return as_f32(r[0]) + as_f32(r[1]);
}
__device__ __forceinline__ float sum_halves32(float a, float b) {
  const auto r = __builtin_amdgcn_permlane32_swap(as_u32f(a), as_u32f(b), false, false);
  return as_f32(r[0]) + as_f32(r[1]);
}

}  // namespace celv2hip
