// Forward transforms of residual blocks, the transform step of the RD pass's trial scoring
// (tx_cost_uncached, src/core/rd_pass.cpp:277-300: fwd_dct, then dropped_energy for 64-point sizes).
//
// Every size from 4x4 to 64x64 (sides 4..64, any ratio), per direction one of four 1D kinds:
//   - DCT-II: Lee's recursion in float, the steps of fwd_dct (src/core/transform.cpp:43-88 and
//     170-214): the rows' full transforms first, then the columns of the kept (low 32) horizontal
//     frequencies, the orthonormal factors and the caller's scale multiplied in at the end, rounded
//     half away from zero. Every product and sum is rounded on its own, in the reference's order
//     (contraction is off in this file), so each coefficient is the reference's value;
//   - DST-VII (up to 16 points), the same of the reversed samples, and the identity (up to 32): the
//     orthonormal float forms, as AVM's integer kernels (fwd_transform, src/core/transform.cpp:
//     219-262) would be without their tables' rounding;
// with the coefficients in AVM's layout (row = vertical frequency, min(w, 32) x min(h, 32) kept) and,
// for a 64-point side, the energy the transform leaves out (dropped_energy, rd_pass.cpp:268-276).
//
// One wave per unit: up to 64 / h blocks of one size and kind pair. The unit's residual rows are
// read once, coalesced, into the wave's LDS area; lane (block, row) transforms a row in registers
// (fully unrolled for its size) into the same area as floats; then lane (block, kept column) takes a
// column, transforms it and writes its coefficients (lanes of one row of frequencies write
// neighbouring words). The kernel streams: about 2 bytes read and 2-4 written per sample.
#include <stdexcept>

#include "launch.h"

#pragma clang fp contract(off)

namespace celv2hip {

namespace {

__constant__ float g_lee[7][32];      // Lee's factors 1 / (2 cos(pi (2i + 1) / (2n))), by log2 n
__constant__ float g_dst[3][16 * 16];  // orthonormal DST-VII, n = 4, 8, 16: row k = frequency
__constant__ float g_ku[7][2];         // sqrt(1 / w), sqrt(2 / w) as floats, by log2 w

constexpr int kDct = 0, kDst = 1, kFlip = 2;  // kinds (CELV2_TX_*); 3, the identity, leaves a pass as it is
constexpr int kUnitLds = 64 * 33 * 4;  // bytes of a wave's area: 64 rows of 66 int16 or 33 floats

template <int N>
struct Log2 {
  static constexpr int value = 1 + Log2<N / 2>::value;
};
template <>
struct Log2<1> {
  static constexpr int value = 0;
};

// The unnormalised DCT-II of x[0..N) in place (Lee's recursion, the reference's order).
template <int N>
__device__ __forceinline__ void lee(float* x) {
  constexpr int lg = Log2<N>::value;
  if constexpr (N == 2) {
    const float a = x[0], b = x[1];
    x[0] = a + b;
    x[1] = (a - b) * g_lee[1][0];
  } else {
    constexpr int H = N / 2;
    float a[H], c[H];
#pragma unroll
    for (int i = 0; i < H; i++) {
      a[i] = x[i] + x[N - 1 - i];
      c[i] = (x[i] - x[N - 1 - i]) * g_lee[lg][i];
    }
    lee<H>(a);
    lee<H>(c);
#pragma unroll
    for (int m = 0; m < H; m++) {
      x[2 * m] = a[m];
      x[2 * m + 1] = m + 1 < H ? c[m] + c[m + 1] : c[m];
    }
  }
}

// DST-VII of x (flip: of the reversed samples) in place, sums in sample order.
template <int N>
__device__ __forceinline__ void dst7(float* x, bool flip) {
  if constexpr (N <= 16) {
    constexpr int lg = Log2<N>::value;
    float in[N];
#pragma unroll
    for (int i = 0; i < N; i++) in[i] = flip ? x[N - 1 - i] : x[i];
#pragma unroll
    for (int k = 0; k < N; k++) {
      float acc = 0.0f;
#pragma unroll
      for (int i = 0; i < N; i++) acc += g_dst[lg - 2][k * N + i] * in[i];
      x[k] = acc;
    }
  }
}

template <int N>
__device__ __forceinline__ void pass1d(float* x, int kind) {
  if (kind == kDct) lee<N>(x);
  else if (kind == kDst || kind == kFlip) dst7<N>(x, kind == kFlip);
}

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <int W, int H>
__device__ void run_unit(const TxJob* jobs, const TxUnit& u, int lane, char* area, unsigned long long* energy) {
  constexpr int CW = W > 32 ? 32 : W, CH = H > 32 ? 32 : H;
  constexpr int RS = W + 2, FS = CW + 1;  // LDS row strides: int16 residual rows, float rows
  constexpr bool kLong = W > 32 || H > 32;
  const int nj = u.count;
  int16_t* res = reinterpret_cast<int16_t*>(area);
  float* rows = reinterpret_cast<float*>(area);

  // The unit's residual rows into LDS, 64 samples at a time in row order.
  const int total = nj * H * W;
  for (int e = lane; e < total; e += 64) {
    const int j = e / (H * W), rem = e - j * (H * W), y = rem / W, x = rem - y * W;
    const TxJob& b = jobs[u.first + j];
    res[(j * H + y) * RS + x] = b.res[static_cast<ptrdiff_t>(y) * b.stride + x];
  }
  if (kLong && lane < nj) energy[lane] = 0, energy[16 + lane] = 0;
  wave_sync();

  // Rows: lane (j, y).
  const bool row_lane = lane < nj * H;
  float x[W];
  long long e_row = 0;
  if (row_lane) {
#pragma unroll
    for (int i = 0; i < W; i++) {
      const int v = res[lane * RS + i];
      x[i] = static_cast<float>(v);
      if (kLong) e_row += static_cast<long long>(v) * v;
    }
  }
  wave_sync();  // every row read before the floats overwrite the area
  if (row_lane) {
    pass1d<W>(x, u.hk);
#pragma unroll
    for (int i = 0; i < CW; i++) rows[lane * FS + i] = x[i];
    if (kLong) atomicAdd(&energy[lane / H], static_cast<unsigned long long>(e_row));
  }
  wave_sync();

  // Columns: lane (j, c) over the unit's kept columns.
  const float ku0 = u.hk == kDct ? g_ku[Log2<W>::value][0] : 1.0f, ku1 = u.hk == kDct ? g_ku[Log2<W>::value][1] : 1.0f;
  for (int t = lane; t < nj * CW; t += 64) {
    const int j = t / CW, c = t - j * CW;
    const TxJob& b = jobs[u.first + j];
    float col[H];
#pragma unroll
    for (int y = 0; y < H; y++) col[y] = rows[(j * H + y) * FS + c];
    pass1d<H>(col, u.vk);
    const float ku = c ? ku1 : ku0;
    long long kept = 0;
#pragma unroll
    for (int v = 0; v < CH; v++) {
      const float k = v ? b.kv : b.kv0;
      const float s = col[v] * ku * k;
      const int q = static_cast<int>(s + (s >= 0.0f ? 0.5f : -0.5f));
      b.coeff[v * CW + c] = q;
      if (kLong) kept += static_cast<long long>(q) * q;
    }
    if (kLong) atomicAdd(&energy[16 + j], static_cast<unsigned long long>(kept));
  }
  if (kLong) {
    wave_sync();
    if (lane < nj) {
      const TxJob& b = jobs[u.first + lane];
      if (b.dropped) {
        const double all = static_cast<double>(static_cast<long long>(energy[lane]));
        const double d = all * b.scale * b.scale - static_cast<double>(static_cast<long long>(energy[16 + lane]));
        const long long r = static_cast<long long>(round(d));
        *b.dropped = r > 0 ? r : 0;
      }
    }
  } else if (lane < nj) {
    const TxJob& b = jobs[u.first + lane];
    if (b.dropped) *b.dropped = 0;
  }
}

// One unit per wave, 4 waves per workgroup.
__global__ void __launch_bounds__(256) k_fwd_tx(const TxJob* __restrict__ jobs, const TxUnit* __restrict__ units, int nunits) {
  __shared__ __align__(16) char areas[4][kUnitLds];
  __shared__ unsigned long long energy[4][32];
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int ui = blockIdx.x * 4 + wave;
  if (ui >= nunits) return;  // whole wave
  const TxUnit u = units[ui];
  char* area = areas[wave];
  unsigned long long* en = energy[wave];
#define CELV2_TX_CASE(wl, hl) \
  case (wl) * 8 + (hl): run_unit<1 << (wl), 1 << (hl)>(jobs, u, lane, area, en); break;
#define CELV2_TX_ROW(wl) CELV2_TX_CASE(wl, 2) CELV2_TX_CASE(wl, 3) CELV2_TX_CASE(wl, 4) CELV2_TX_CASE(wl, 5) CELV2_TX_CASE(wl, 6)
  switch (u.wl * 8 + u.hl) {
    CELV2_TX_ROW(2)
    CELV2_TX_ROW(3)
    CELV2_TX_ROW(4)
    CELV2_TX_ROW(5)
    CELV2_TX_ROW(6)
    default: break;
  }
#undef CELV2_TX_ROW
#undef CELV2_TX_CASE
}

}  // namespace

hipError_t upload_tx_tables() {
  const double pi = 3.14159265358979323846;
  float lee[7][32] = {}, dst[3][16 * 16] = {}, ku[7][2] = {};
  for (int lg = 1; lg <= 6; lg++) {
    const int n = 1 << lg;
    for (int i = 0; i < n / 2; i++) lee[lg][i] = static_cast<float>(0.5 / std::cos(pi * (2 * i + 1) / (2.0 * n)));
    ku[lg][0] = static_cast<float>(std::sqrt(1.0 / n)), ku[lg][1] = static_cast<float>(std::sqrt(2.0 / n));
  }
  for (int lg = 2; lg <= 4; lg++) {
    const int n = 1 << lg;
    for (int k = 0; k < n; k++)
      for (int i = 0; i < n; i++)
        dst[lg - 2][k * n + i] = static_cast<float>(2.0 / std::sqrt(2.0 * n + 1) * std::sin(pi * (2 * k + 1) * (i + 1) / (2.0 * n + 1)));
  }
  hipError_t rc = hipMemcpyToSymbol(HIP_SYMBOL(g_lee), lee, sizeof(lee));
  if (rc == hipSuccess) rc = hipMemcpyToSymbol(HIP_SYMBOL(g_dst), dst, sizeof(dst));
  if (rc == hipSuccess) rc = hipMemcpyToSymbol(HIP_SYMBOL(g_ku), ku, sizeof(ku));
  return rc;
}

int tx_jobs_per_unit(int hl) { return hl >= 6 ? 1 : 64 >> hl; }

void launch_fwd_tx(const TxJob* jobs, const TxUnit* units, int nunits, hipStream_t s) {
  if (nunits <= 0) return;
  k_fwd_tx<<<(nunits + 3) / 4, 256, 0, s>>>(jobs, units, nunits);
}

}  // namespace celv2hip
