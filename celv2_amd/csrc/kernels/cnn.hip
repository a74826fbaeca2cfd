// The learned pruner's network on the matrix cores: BF16 operands, FP32 accumulation, FP32 master
// weights with AdamW.
//
// Input: a frame's superblock grid (h x w, batch b) with kFeatures features per superblock. Layers:
// 3x3 conv (kFeatures -> 32) + ReLU, 3x3 conv (32 -> 32) + ReLU, 1x1 conv (32 -> 6); output j is the
// logit of "this superblock needs more than its nearest j + 1 references". Every layer is a GEMM
// C[m][n] = sum_k A(m, k) B(k, n) on v_mfma_f32_16x16x16bf16_1k (one wave per 16x16 tile of C), over
// an im2col of its input (tap-major: k = ((dy + 1) * 3 + dx + 1) * C + c, zero outside the grid).
// The backward pass reuses the GEMM with strides: dW = col^T dZ, dcol = dZ W^T (gathered back to
// the grid), and reductions for the biases.
#include <cmath>
#include <mutex>

#include "cnn.h"
#include "common.h"

namespace celv2hip {

namespace {

typedef short v4i16 __attribute__((ext_vector_type(4)));
typedef float f32x4v __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint16_t bf16_bits(float f) {  // round to nearest even
  uint32_t u = __builtin_bit_cast(uint32_t, f);
  u += 0x7fffu + ((u >> 16) & 1u);
  return static_cast<uint16_t>(u >> 16);
}
__device__ __forceinline__ float bf16_value(uint16_t h) { return __builtin_bit_cast(float, static_cast<uint32_t>(h) << 16); }

struct Gemm {
  const uint16_t* a;  // A(m, k) = a[m * sam + k * sak]
  const uint16_t* b;  // B(k, n) = b[k * sbk + n * sbn]
  long long sam, sak, sbk, sbn;
  int m, n, k;
  const float* bias;  // per column (may be null)
  int relu;
  float* c;           // C[m * ldc + n] (may be null)
  uint16_t* cb;       // bf16 copy (may be null)
  int ldc;
};

__global__ void __launch_bounds__(256) k_gemm(Gemm g) {
  const int wave = static_cast<int>(threadIdx.x >> 6), lane = static_cast<int>(threadIdx.x & 63);
  const int tiles_n = (g.n + 15) / 16, tiles_m = (g.m + 15) / 16;
  const int t = static_cast<int>(blockIdx.x) * 4 + wave;
  if (t >= tiles_m * tiles_n) return;  // whole wave
  const int tm = t / tiles_n, tn = t - tm * tiles_n, i = lane & 15, q = lane >> 4;
  const int row = 16 * tm + i, col = 16 * tn + i;
  f32x4v acc = {0.f, 0.f, 0.f, 0.f};
  for (int k0 = 0; k0 < g.k; k0 += 16) {
    v4i16 av, bv;
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const int kk = k0 + 4 * q + j;
      av[j] = static_cast<short>(row < g.m && kk < g.k ? g.a[row * g.sam + kk * g.sak] : 0);
      bv[j] = static_cast<short>(col < g.n && kk < g.k ? g.b[kk * g.sbk + col * g.sbn] : 0);
    }
    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, bv, acc, 0, 0, 0);
  }
#pragma unroll
  for (int r = 0; r < 4; r++) {
    const int m = 16 * tm + 4 * q + r, n = 16 * tn + i;
    if (m >= g.m || n >= g.n) continue;
    float v = acc[r] + (g.bias ? g.bias[n] : 0.f);
    if (g.relu) v = fmaxf(v, 0.f);
    if (g.c) g.c[static_cast<size_t>(m) * g.ldc + n] = v;
    if (g.cb) g.cb[static_cast<size_t>(m) * g.ldc + n] = bf16_bits(v);
  }
}

// Inference in one launch: a workgroup takes a 4 x 16 tile of one frame's grid. Layer 1 runs on the
// tile plus a ring of 1 (6 x 18 positions) from the features staged as BF16 with a ring of 2, layer 2
// and 3 on the tile, all through LDS. Each output element is the same chain of
// v_mfma_f32_16x16x16bf16_1k over the same BF16 operands, K-chunk by K-chunk, as k_gemm over the
// im2col buffers, so the logits are those of the layer-by-layer pass bit for bit. Then each
// position's reference limit (as k_limits).
constexpr int kTileH = 4, kTileW = 16;
constexpr int kZ1H = kTileH + 2, kZ1W = kTileW + 2, kZ1 = kZ1H * kZ1W;  // 108 positions
constexpr int kXH = kTileH + 4, kXW = kTileW + 4;

__device__ __forceinline__ int limit_of(const float* z, int nrefs, float threshold, float band, bool& unsure) {
  int k = 1;
  bool open = true;
  unsure = false;
  for (int j = 0; j < kRefOut; j++) {
    const float s = 1.f / (1.f + expf(-z[j]));
    if (fabsf(s - threshold) < band) unsure = true;
    if (open && s >= threshold) k = j + 2;
    else open = false;
  }
  if (unsure) k = kMaxRefs;
  return k < nrefs ? k : (nrefs < 1 ? 1 : nrefs);
}

struct Infer {
  const float* x;  // [b][h][w][kFeatures]
  int h, w;
  const uint16_t* pb;
  const float* params;
  float* logits;  // [b][h][w][kOut] (may be null)
  uint8_t* kmap;  // [b][h][w] (may be null)
  unsigned long long* fallbacks;
  int nrefs;
  float threshold, band;
  uint8_t* imask;  // [b][h][w] intra mode groups (may be null)
  float gthreshold;
  uint8_t* smap;   // [b][h][w] motion-search modes (may be null)
  float sthreshold;
  // The features' frame context (f[11..13], the frame's means) is not in x yet: each workgroup
  // computes it from every position of its frame and writes it for its own tile's positions.
  bool context;
};

// A superblock's intra mode groups from the group outputs (group 0 always; every group on a fallback).
__device__ __forceinline__ uint8_t groups_of(const float* z, float gthreshold, bool unsure) {
  if (unsure) return kAllGroups;
  uint8_t m = 1;
  for (int g = 0; g < kGroupOut; g++)
    if (1.f / (1.f + expf(-z[kRefOut + g])) >= gthreshold) m |= static_cast<uint8_t>(2u << g);
  return m;
}

static_assert(kOut <= 16, "the output layer is one 16-column tile");
constexpr int kInferThreads = 512;  // 8 waves: layer 2's 8 tiles one per wave
__global__ void __launch_bounds__(kInferThreads) k_infer(Infer a) {
  __shared__ uint16_t xs[kXH * kXW * kFeatures];
  __shared__ uint16_t z1s[kZ1 * kHidden];
  __shared__ uint16_t z2s[kTileH * kTileW * kHidden];
  __shared__ float z3s[kTileH * kTileW * 16];
  // The BF16 weights in LDS (the products read them at every step).
  __shared__ uint16_t w1s[9 * kFeatures * kHidden], w2s[kK2 * kHidden], w3s[kHidden * kOut];
  const int tid = static_cast<int>(threadIdx.x), wave = tid >> 6, lane = tid & 63, i = lane & 15, q = lane >> 4;
  constexpr int kWaves = kInferThreads / 64;
  for (int e = tid; e < 9 * kFeatures * kHidden; e += kInferThreads) w1s[e] = a.pb[kW1 + e];
  for (int e = tid; e < kK2 * kHidden; e += kInferThreads) w2s[e] = a.pb[kW2 + e];
  for (int e = tid; e < kHidden * kOut; e += kInferThreads) w3s[e] = a.pb[kW3 + e];
  const int tx0 = static_cast<int>(blockIdx.x) * kTileW, ty0 = static_cast<int>(blockIdx.y) * kTileH;
  const size_t frame = static_cast<size_t>(blockIdx.z) * a.h * a.w;
  __shared__ float part[3][256];
  float ctx[3] = {0.f, 0.f, 0.f};
  if (a.context) {
    // The frame's means of f[0], f[8] and f[1] + ... + f[6]: a strided pass by 256 threads, then a
    // tree, in a fixed order (the same sums k_frame_context forms).
    const int n = a.h * a.w;
    if (tid < 256) {
      float s0 = 0.f, s1 = 0.f, s2 = 0.f;
      for (int sb = tid; sb < n; sb += 256) {
        const float* f = a.x + (frame + static_cast<size_t>(sb)) * kFeatures;
        s0 += f[0], s1 += f[8], s2 += f[1] + f[2] + f[3] + f[4] + f[5] + f[6];
      }
      part[0][tid] = s0, part[1][tid] = s1, part[2][tid] = s2;
    }
    __syncthreads();
    for (int w = 128; w >= 1; w >>= 1) {
      if (tid < w)
        for (int c = 0; c < 3; c++) part[c][tid] += part[c][tid + w];
      __syncthreads();
    }
    const float inv = 1.f / static_cast<float>(n);
    ctx[0] = part[0][0] * inv, ctx[1] = part[1][0] * inv, ctx[2] = part[2][0] * inv;
    if (tid < kTileH * kTileW) {
      const int yy = ty0 + tid / kTileW, xx = tx0 + tid % kTileW;
      if (yy < a.h && xx < a.w) {
        float* f = const_cast<float*>(a.x) + (frame + static_cast<size_t>(yy) * a.w + xx) * kFeatures;
        f[11] = ctx[0], f[12] = ctx[1], f[13] = ctx[2];
      }
    }
  }
  for (int e = tid; e < kXH * kXW * kFeatures; e += kInferThreads) {
    const int c = e % kFeatures, pos = e / kFeatures, yy = ty0 - 2 + pos / kXW, xx = tx0 - 2 + pos % kXW;
    const bool in = yy >= 0 && yy < a.h && xx >= 0 && xx < a.w;
    const float v = !in ? 0.f : a.context && c >= 11 ? ctx[c - 11] : a.x[(frame + static_cast<size_t>(yy) * a.w + xx) * kFeatures + c];
    xs[e] = bf16_bits(v);
  }
  __syncthreads();
  // Layer 1 on the 6 x 18 region (7 row tiles x 2 column tiles); 0 outside the grid (layer 2's padding).
  for (int t = wave; t < 14; t += kWaves) {
    const int tm = t >> 1, tn = t & 1, m = 16 * tm + i, col = 16 * tn + i;
    const int ry = m / kZ1W, rx = m - ry * kZ1W;
    f32x4v acc = {0.f, 0.f, 0.f, 0.f};
    for (int k0 = 0; k0 < 9 * kFeatures; k0 += 16) {
      v4i16 av, bv;
#pragma unroll
      for (int j = 0; j < 4; j++) {
        const int kk = k0 + 4 * q + j, tap = kk / kFeatures, c = kk - tap * kFeatures, dy = tap / 3 - 1, dx = tap % 3 - 1;
        const bool ok = kk < 9 * kFeatures;
        av[j] = static_cast<short>(ok && m < kZ1 ? xs[((ry + 1 + dy) * kXW + rx + 1 + dx) * kFeatures + c] : 0);
        bv[j] = static_cast<short>(ok ? w1s[kk * kHidden + col] : 0);
      }
      acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, bv, acc, 0, 0, 0);
    }
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int mm = 16 * tm + 4 * q + r;
      if (mm >= kZ1) continue;
      const int yy = ty0 - 1 + mm / kZ1W, xx = tx0 - 1 + mm % kZ1W;
      const bool in = yy >= 0 && yy < a.h && xx >= 0 && xx < a.w;
      z1s[mm * kHidden + col] = in ? bf16_bits(fmaxf(acc[r] + a.params[kB1 + col], 0.f)) : 0;
    }
  }
  __syncthreads();
  // Layer 2: row tile tm = tile row, its 16 rows = the tile's columns.
  for (int t = wave; t < 2 * kTileH; t += kWaves) {
    const int tm = t >> 1, tn = t & 1, col = 16 * tn + i;
    f32x4v acc = {0.f, 0.f, 0.f, 0.f};
    for (int k0 = 0; k0 < 9 * kHidden; k0 += 16) {
      v4i16 av, bv;
#pragma unroll
      for (int j = 0; j < 4; j++) {
        const int kk = k0 + 4 * q + j, tap = kk / kHidden, c = kk - tap * kHidden, dy = tap / 3 - 1, dx = tap % 3 - 1;
        av[j] = static_cast<short>(z1s[((tm + 1 + dy) * kZ1W + i + 1 + dx) * kHidden + c]);
        bv[j] = static_cast<short>(w2s[kk * kHidden + col]);
      }
      acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, bv, acc, 0, 0, 0);
    }
#pragma unroll
    for (int r = 0; r < 4; r++) z2s[(tm * kTileW + 4 * q + r) * kHidden + col] = bf16_bits(fmaxf(acc[r] + a.params[kB2 + col], 0.f));
  }
  __syncthreads();
  // Layer 3 (1x1, 6 outputs): one row tile per wave.
  if (wave < kTileH) {
    const int tm = wave;
    f32x4v acc = {0.f, 0.f, 0.f, 0.f};
    for (int k0 = 0; k0 < kHidden; k0 += 16) {
      v4i16 av, bv;
#pragma unroll
      for (int j = 0; j < 4; j++) {
        const int kk = k0 + 4 * q + j;
        av[j] = static_cast<short>(z2s[(tm * kTileW + i) * kHidden + kk]);
        bv[j] = static_cast<short>(i < kOut ? w3s[kk * kOut + i] : 0);
      }
      acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, bv, acc, 0, 0, 0);
    }
    if (i < kOut)
#pragma unroll
      for (int r = 0; r < 4; r++) z3s[(tm * kTileW + 4 * q + r) * 16 + i] = acc[r] + a.params[kB3 + i];
  }
  __syncthreads();
  if (tid < kTileH * kTileW) {
    const int yy = ty0 + tid / kTileW, xx = tx0 + tid % kTileW;
    if (yy < a.h && xx < a.w) {
      const size_t p = frame + static_cast<size_t>(yy) * a.w + xx;
      const float* z = z3s + tid * 16;
      if (a.logits)
        for (int j = 0; j < kOut; j++) a.logits[p * kOut + j] = z[j];
      if (a.kmap) {
        bool unsure;
        a.kmap[p] = static_cast<uint8_t>(limit_of(z, a.nrefs, a.threshold, a.band, unsure));
        if (unsure && a.fallbacks) count(a.fallbacks, static_cast<int>(p), 1ull);
        if (a.imask) a.imask[p] = groups_of(z, a.gthreshold, unsure);
        if (a.smap) a.smap[p] = unsure || 1.f / (1.f + expf(-z[kRefOut + kGroupOut])) >= a.sthreshold ? 0 : 3;
      }
    }
  }
}

// col[p][k] (k < kp) from x[p][c] (f32, c channels), 3x3 taps around p on each frame's grid.
__global__ void k_im2col(const float* x, int b, int h, int w, int c, int kp, uint16_t* col) {
  const long long idx = static_cast<long long>(blockIdx.x) * 256 + threadIdx.x;
  const long long total = static_cast<long long>(b) * h * w * kp;
  if (idx >= total) return;
  const int k = static_cast<int>(idx % kp);
  const long long p = idx / kp;
  const int px = static_cast<int>(p % w), py = static_cast<int>((p / w) % h);
  float v = 0.f;
  if (k < 9 * c) {
    const int tap = k / c, ch = k - tap * c, dy = tap / 3 - 1, dx = tap % 3 - 1;
    const int yy = py + dy, xx = px + dx;
    if (yy >= 0 && yy < h && xx >= 0 && xx < w) v = x[(p + static_cast<long long>(dy) * w + dx) * c + ch];
  }
  col[idx] = bf16_bits(v);
}

// dx[p][c] = sum over taps of dcol[p - tap offset][tap * c + ch] (the transpose of k_im2col), times
// the ReLU mask of the layer's output y (y > 0).
__global__ void k_col2im_relu(const float* dcol, int b, int h, int w, int c, int kp, const float* y, float* dx,
                              uint16_t* dxb) {
  const long long idx = static_cast<long long>(blockIdx.x) * 256 + threadIdx.x;
  const long long total = static_cast<long long>(b) * h * w * c;
  if (idx >= total) return;
  const int ch = static_cast<int>(idx % c);
  const long long p = idx / c;
  const int px = static_cast<int>(p % w), py = static_cast<int>((p / w) % h);
  float s = 0.f;
  for (int tap = 0; tap < 9; tap++) {
    const int dy = tap / 3 - 1, dx2 = tap % 3 - 1, sy = py - dy, sx = px - dx2;
    if (sy >= 0 && sy < h && sx >= 0 && sx < w) s += dcol[(p - static_cast<long long>(dy) * w - dx2) * kp + tap * c + ch];
  }
  if (y[idx] <= 0.f) s = 0.f;
  dx[idx] = s;
  dxb[idx] = bf16_bits(s);
}

// g[p][c] *= (y[p][c] > 0), with its bf16 copy.
__global__ void k_relu_mask(float* g, const float* y, long long n, uint16_t* gb) {
  const long long i = static_cast<long long>(blockIdx.x) * 256 + threadIdx.x;
  if (i >= n) return;
  const float v = y[i] > 0.f ? g[i] : 0.f;
  g[i] = v;
  gb[i] = bf16_bits(v);
}

// db[n] = sum_p g[p][n].
__global__ void k_bias_grad(const float* g, long long m, int n, float* db) {
  __shared__ float part[256];
  const int col = static_cast<int>(blockIdx.x);
  float s = 0.f;
  for (long long p = threadIdx.x; p < m; p += 256) s += g[p * n + col];
  part[threadIdx.x] = s;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (static_cast<int>(threadIdx.x) < st) part[threadIdx.x] += part[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) db[col] = part[0];
}

// Binary cross-entropy with logits, labels y_j = (kstar > j + 1) for the reference outputs (kstar =
// label bits 0..2), y = label bit 3 + g for group output 6 + g and y = label bit 6 for the search
// output 9 (the same rule: bit 3 + j - 6); mean over m x kOut. Writes the
// gradient (sigma(z) - y) / (m kOut), its bf16 copy, and per-block partial losses.
__global__ void k_bce(const float* z, const uint8_t* kstar, long long m, float* dz, uint16_t* dzb, float* loss_part) {
  __shared__ float part[256];
  const long long i = static_cast<long long>(blockIdx.x) * 256 + threadIdx.x;
  float l = 0.f;
  if (i < m * kOut) {
    const long long p = i / kOut;
    const int j = static_cast<int>(i - p * kOut);
    const uint32_t lab = kstar[p];
    const float x = z[i], y = (j < kRefOut ? (lab & 7u) > static_cast<uint32_t>(j + 1) : ((lab >> (3 + j - kRefOut)) & 1u)) ? 1.f : 0.f;
    l = fmaxf(x, 0.f) - x * y + log1pf(expf(-fabsf(x)));
    const float s = 1.f / (1.f + expf(-x));
    const float g = (s - y) / static_cast<float>(m * kOut);
    dz[i] = g;
    dzb[i] = bf16_bits(g);
  }
  part[threadIdx.x] = l;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (static_cast<int>(threadIdx.x) < st) part[threadIdx.x] += part[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) loss_part[blockIdx.x] = part[0];
}

// AdamW on the flat parameter vector; refreshes the bf16 copy.
__global__ void k_adamw(float* w, const float* g, float* m1, float* m2, int n, float lr, float wd, float b1, float b2,
                        float eps, float c1, float c2, uint16_t* wb) {
  const int i = static_cast<int>(blockIdx.x) * 256 + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float a = b1 * m1[i] + (1.f - b1) * gi, v = b2 * m2[i] + (1.f - b2) * gi * gi;
  m1[i] = a, m2[i] = v;
  const float upd = (a / c1) / (sqrtf(v / c2) + eps) + wd * w[i];
  w[i] -= lr * upd;
  wb[i] = bf16_bits(w[i]);
}

__global__ void k_to_bf16(const float* x, long long n, uint16_t* y) {
  const long long i = static_cast<long long>(blockIdx.x) * 256 + threadIdx.x;
  if (i < n) y[i] = bf16_bits(x[i]);
}

// Per superblock: the reference limit from the network's outputs. k = 1 + the number of leading
// outputs j with sigma(z_j) >= threshold (ordinal), at most nrefs; any output within `band` of the
// threshold makes the superblock fall back to all references.
__global__ void k_limits(const float* z, int m, int nrefs, float threshold, float band, uint8_t* kmap,
                         unsigned long long* fallbacks) {
  const int p = static_cast<int>(blockIdx.x) * 256 + threadIdx.x;
  if (p >= m) return;
  bool unsure;
  kmap[p] = static_cast<uint8_t>(limit_of(z + static_cast<size_t>(p) * kOut, nrefs, threshold, band, unsure));
  if (unsure && fallbacks) count(fallbacks, p, 1ull);
}

unsigned grid(long long n) { return static_cast<unsigned>((n + 255) / 256); }

void gemm(hipStream_t s, const uint16_t* a, long long sam, long long sak, const uint16_t* b, long long sbk, long long sbn,
          int m, int n, int k, const float* bias, bool relu, float* c, uint16_t* cb, int ldc) {
  Gemm g{a, b, sam, sak, sbk, sbn, m, n, k, bias, relu ? 1 : 0, c, cb, ldc};
  const int tiles = ((m + 15) / 16) * ((n + 15) / 16);
  k_gemm<<<(tiles + 3) / 4, 256, 0, s>>>(g);
}

}  // namespace

size_t CnnWork::positions_capacity() const { return cap; }

void cnn_alloc(CnnWork& w, size_t positions) {
  w.cap = positions;
  auto f = [&](float** p, size_t n) { if (hipMalloc(p, n * sizeof(float)) != hipSuccess) throw std::runtime_error("hipMalloc cnn"); };
  auto h = [&](uint16_t** p, size_t n) { if (hipMalloc(p, n * sizeof(uint16_t)) != hipSuccess) throw std::runtime_error("hipMalloc cnn"); };
  f(&w.x, positions * kFeatures);
  h(&w.col1, positions * kK1);
  f(&w.z1, positions * kHidden);
  h(&w.col2, positions * kK2);
  f(&w.z2, positions * kHidden);
  h(&w.z2b, positions * kHidden);
  f(&w.z3, positions * kOut);
  f(&w.dz3, positions * kOut);
  h(&w.dz3b, positions * kOut);
  f(&w.dz2, positions * kHidden);
  h(&w.dz2b, positions * kHidden);
  f(&w.dcol2, positions * kK2);
  f(&w.dz1, positions * kHidden);
  h(&w.dz1b, positions * kHidden);
  f(&w.loss_part, (positions * kOut + 255) / 256 + 1);
  if (hipMalloc(&w.kstar, positions) != hipSuccess) throw std::runtime_error("hipMalloc cnn");
}

void cnn_free(CnnWork& w) {
  for (void* p : {static_cast<void*>(w.x), static_cast<void*>(w.col1), static_cast<void*>(w.z1), static_cast<void*>(w.col2),
                  static_cast<void*>(w.z2), static_cast<void*>(w.z2b), static_cast<void*>(w.z3), static_cast<void*>(w.dz3),
                  static_cast<void*>(w.dz3b), static_cast<void*>(w.dz2), static_cast<void*>(w.dz2b), static_cast<void*>(w.dcol2),
                  static_cast<void*>(w.dz1), static_cast<void*>(w.dz1b), static_cast<void*>(w.loss_part), static_cast<void*>(w.kstar)})
    if (p) (void)hipFree(p);
  w = CnnWork{};
}

void cnn_params_to_bf16(const float* params, uint16_t* pb, hipStream_t s) {
  k_to_bf16<<<grid(kParams), 256, 0, s>>>(params, kParams, pb);
}

// Forward over w.x (b frames of h x w superblocks): logits in w.z3.
void cnn_forward(const CnnWork& w, const float* params, const uint16_t* pb, int b, int h, int wd, hipStream_t s) {
  const long long m = static_cast<long long>(b) * h * wd;
  k_im2col<<<grid(m * kK1), 256, 0, s>>>(w.x, b, h, wd, kFeatures, kK1, w.col1);
  gemm(s, w.col1, kK1, 1, pb + kW1, kHidden, 1, static_cast<int>(m), kHidden, 9 * kFeatures, params + kB1, true, w.z1, nullptr, kHidden);
  k_im2col<<<grid(m * kK2), 256, 0, s>>>(w.z1, b, h, wd, kHidden, kK2, w.col2);
  gemm(s, w.col2, kK2, 1, pb + kW2, kHidden, 1, static_cast<int>(m), kHidden, 9 * kHidden, params + kB2, true, w.z2, w.z2b, kHidden);
  gemm(s, w.z2b, kHidden, 1, pb + kW3, kOut, 1, static_cast<int>(m), kOut, kHidden, params + kB3, false, w.z3, nullptr, kOut);
}

// Loss and gradients (into grads, kParams floats) for labels w.kstar; the loss's per-block partial
// sums are in w.loss_part (blocks: grid(m * kOut)).
void cnn_backward(const CnnWork& w, const uint16_t* pb, float* grads, int b, int h, int wd, hipStream_t s) {
  const long long m = static_cast<long long>(b) * h * wd;
  k_bce<<<grid(m * kOut), 256, 0, s>>>(w.z3, w.kstar, m, w.dz3, w.dz3b, w.loss_part);
  // Layer 3 (1x1): dW3 = z2^T dz3, db3, dz2 = dz3 W3^T masked by ReLU(z2).
  gemm(s, w.z2b, 1, kHidden, w.dz3b, kOut, 1, kHidden, kOut, static_cast<int>(m), nullptr, false, grads + kW3, nullptr, kOut);
  k_bias_grad<<<kOut, 256, 0, s>>>(w.dz3, m, kOut, grads + kB3);
  gemm(s, w.dz3b, kOut, 1, pb + kW3, 1, kOut, static_cast<int>(m), kHidden, kOut, nullptr, false, w.dz2, nullptr, kHidden);
  k_relu_mask<<<grid(m * kHidden), 256, 0, s>>>(w.dz2, w.z2, m * kHidden, w.dz2b);
  // Layer 2 (3x3 over z1): dW2 = col2^T dz2, db2, dcol2 = dz2 W2^T, dz1 = col2im(dcol2) masked.
  gemm(s, w.col2, 1, kK2, w.dz2b, kHidden, 1, 9 * kHidden, kHidden, static_cast<int>(m), nullptr, false, grads + kW2, nullptr, kHidden);
  k_bias_grad<<<kHidden, 256, 0, s>>>(w.dz2, m, kHidden, grads + kB2);
  gemm(s, w.dz2b, kHidden, 1, pb + kW2, 1, kHidden, static_cast<int>(m), 9 * kHidden, kHidden, nullptr, false, w.dcol2, nullptr, kK2);
  k_col2im_relu<<<grid(m * kHidden), 256, 0, s>>>(w.dcol2, b, h, wd, kHidden, kK2, w.z1, w.dz1, w.dz1b);
  // Layer 1 (3x3 over the features): dW1 = col1^T dz1, db1.
  gemm(s, w.col1, 1, kK1, w.dz1b, kHidden, 1, 9 * kFeatures, kHidden, static_cast<int>(m), nullptr, false, grads + kW1, nullptr, kHidden);
  k_bias_grad<<<kHidden, 256, 0, s>>>(w.dz1, m, kHidden, grads + kB1);
}

void cnn_adamw(float* params, const float* grads, float* m1, float* m2, uint16_t* pb, long long step, float lr, float wd,
               hipStream_t s) {
  const float b1 = 0.9f, b2 = 0.999f;
  const float c1 = 1.f - std::pow(b1, static_cast<float>(step)), c2 = 1.f - std::pow(b2, static_cast<float>(step));
  k_adamw<<<grid(kParams), 256, 0, s>>>(params, grads, m1, m2, kParams, lr, wd, b1, b2, 1e-8f, c1, c2, pb);
}

void cnn_infer(const float* x, int b, int h, int wd, const float* params, const uint16_t* pb, float* logits, int nrefs,
               float threshold, float band, uint8_t* kmap, unsigned long long* fallbacks, hipStream_t s, uint8_t* imask,
               float gthreshold, uint8_t* smap, float sthreshold, bool context) {
  if (b <= 0 || h <= 0 || wd <= 0) return;
  const Infer a{x, h, wd, pb, params, logits, kmap, fallbacks, nrefs, threshold, band, imask, gthreshold, smap, sthreshold, context};
  k_infer<<<dim3((wd + kTileW - 1) / kTileW, (h + kTileH - 1) / kTileH, b), kInferThreads, 0, s>>>(a);
}

// Scratch of celv2hip_pruner_forward_static: the features and logits of up to kStaticPositions
// positions in a device array of this file, so that entry point allocates nothing. Each device holds
// its own copy until the process ends; its address is taken once per device (cnn_scratch_init, from
// celv2hip_pruner_create with that device current) and the kernels receive that address, never the
// array's name.
static __device__ __align__(256) float scratch[static_cast<size_t>(kStaticPositions) * (kFeatures + kOut)];

namespace {
std::mutex g_scratch_mu;
float* g_scratch[kScratchDevices] = {};
}  // namespace

hipError_t cnn_scratch_init(int device) {
  if (device < 0 || device >= kScratchDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch[device]) return hipSuccess;
  void* p = nullptr;
  const hipError_t rc = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (rc == hipSuccess) g_scratch[device] = static_cast<float*>(p);
  return rc;
}

float* cnn_scratch(int device) {
  if (device < 0 || device >= kScratchDevices) return nullptr;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  return g_scratch[device];
}

void cnn_limits(const float* logits, int m, int nrefs, float threshold, float band, uint8_t* kmap,
                unsigned long long* fallbacks, hipStream_t s) {
  if (m > 0) k_limits<<<grid(m), 256, 0, s>>>(logits, m, nrefs, threshold, band, kmap, fallbacks);
}

}  // namespace celv2hip
