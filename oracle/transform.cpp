// ORACLE — test infrastructure, not product code. A CPU restatement of the forward transforms of
// CelV2's RD pass, used only by tests/ and bench.py's CPU leg to check and time the HIP transform
// kernel (celv2_amd/csrc/kernels/transform.hip). The product never includes, links or calls it.
//
// What it restates (paths relative to the CelV2 tree):
//   - fwd_dct (src/core/transform.h:12; the AVX2 path src/core/transform.cpp:170-214, which the
//     AVX-512 one equals): a separable DCT-II in float, rows first then the kept columns, each 1D
//     pass Lee's recursion (src/core/transform.cpp:23-31 factors, 43-88 the recursion), the
//     orthonormal factors and the caller's scale multiplied in at the end, rounded half away from
//     zero; AVM's layout (row = vertical frequency), only the low min(w, 32) x min(h, 32) kept.
//     The float operations are the reference's, one for one and in its order (the reference builds
//     with MSVC /fp:precise and no contraction, CMakeLists.txt:7-8; this file builds with
//     -ffp-contract=off), so the coefficients are the reference's values.
//   - dropped_energy (src/core/rd_pass.cpp:268-276): what a 64-point transform leaves out, as the
//     residual's energy times scale^2 less the kept coefficients' squares.
//   - The other 1D kinds of fwd_transform (src/core/transform.cpp:219-262: DST-VII, its flip and
//     the identity, by vertical and horizontal kind) are AVM's integer kernels there; AVM's tables
//     are absent, so here they are the orthonormal float forms (DST-VII T[k][i] =
//     2 / sqrt(2n + 1) sin(pi (2k + 1)(i + 1) / (2n + 1)), k the frequency, i the sample; the
//     identity x[i]) summed in index
//     order, on the same rows-then-columns plan and scale as fwd_dct. Parity with AVM's integer
//     transforms is unpinned.
//
// Pinning (tests/test_oracle_transform.py): fwd_dct against the double-precision DCT-II at every
// size on the reference's own residual kinds (constant 255, random +-255, random in [-255, 255]),
// each coefficient within +-1 (tests/test_transform.cpp:81-100's bar), DC closed forms,
// Parseval at sizes up to 32, the dropped energy against double sums, the matrix kinds against
// float64 products and orthonormality; deliberate mistakes (ORC_TX_PLANT) each fail one of them.
#include <cmath>
#include <cstdint>
#include <cstdlib>

namespace {

constexpr int kDct = 0, kDst = 1, kFlipDst = 2, kIdt = 3;

// Lee's odd-half factors 1 / (2 cos(pi (2i + 1) / (2n))) for n = 2..64 (by log2 n), computed in
// double and stored as float, as the reference's table is.
struct LeeTable {
  float f[7][32] = {};
  LeeTable() {
    const double pi = 3.14159265358979323846;
    for (int lg = 1; lg <= 6; lg++) {
      const int n = 1 << lg;
      for (int i = 0; i < n / 2; i++) f[lg][i] = static_cast<float>(0.5 / std::cos(pi * (2 * i + 1) / (2.0 * n)));
    }
#if ORC_TX_PLANT == 1
    f[3][1] = f[3][2];  // a factor of the 8-point step taken from the wrong index
#endif
  }
};
const LeeTable& lee() {
  static const LeeTable t;
  return t;
}

int log2i(int n) {
  int l = 0;
  while ((1 << l) < n) l++;
  return l;
}

// The unnormalised DCT-II X[k] = sum_i x[i] cos(pi (2i + 1) k / (2n)) of one signal of n samples,
// in place, by Lee's recursion: even outputs are the half-length DCT of x[i] + x[n-1-i], odd outputs
// X[2m+1] = C[m] + C[m+1] with C the half-length DCT of (x[i] - x[n-1-i]) f[i]. tmp: 2n floats.
void dct_lee(float* x, int n, float* tmp) {
  const float* f = lee().f[log2i(n)];
  if (n == 2) {
    const float a = x[0], b = x[1];
    x[0] = a + b;
    x[1] = (a - b) * f[0];
    return;
  }
  const int h = n / 2;
  float* a = tmp;
  float* c = tmp + h;
  for (int i = 0; i < h; i++) {
    a[i] = x[i] + x[n - 1 - i];
    c[i] = (x[i] - x[n - 1 - i]) * f[i];
  }
  dct_lee(a, h, tmp + n);
  dct_lee(c, h, tmp + n);
  for (int m = 0; m < h; m++) {
    x[2 * m] = a[m];
#if ORC_TX_PLANT == 2
    x[2 * m + 1] = m + 1 < h ? c[m] - c[m + 1] : c[m];  // the odd outputs' sum with the wrong sign
#else
    x[2 * m + 1] = m + 1 < h ? c[m] + c[m + 1] : c[m];
#endif
  }
}

// The orthonormal DST-VII matrix of n points (n <= 16; row k = frequency), float entries computed in
// double.
struct DstTable {
  float t[5][16 * 16] = {};  // by log2 n
  DstTable() {
    const double pi = 3.14159265358979323846;
    for (int lg = 2; lg <= 4; lg++) {
      const int n = 1 << lg;
      for (int k = 0; k < n; k++)
        for (int i = 0; i < n; i++)
          t[lg][k * n + i] = static_cast<float>(2.0 / std::sqrt(2.0 * n + 1) * std::sin(pi * (2 * k + 1) * (i + 1) / (2.0 * n + 1)));
    }
  }
};
const DstTable& dst() {
  static const DstTable t;
  return t;
}

// One 1D pass of kind `kind` over one signal in place (x: n floats; tmp: 2n floats). DCT2: Lee's
// unnormalised sums (the orthonormal factors come at the end); the others orthonormal already.
void pass1d(float* x, int n, int kind, float* tmp) {
  if (kind == kDct) return dct_lee(x, n, tmp);
  if (kind == kIdt) return;
  const float* t = dst().t[log2i(n)];
  for (int i = 0; i < n; i++) tmp[i] = kind == kFlipDst ? x[n - 1 - i] : x[i];
  for (int k = 0; k < n; k++) {
    float acc = 0.0f;
    for (int i = 0; i < n; i++) acc += t[k * n + i] * tmp[i];
    x[k] = acc;
  }
}

bool kind_ok(int kind, int n) {
  if (n != 4 && n != 8 && n != 16 && n != 32 && n != 64) return false;
  if (kind == kDct) return true;
  if (kind == kIdt) return n <= 32;
  return (kind == kDst || kind == kFlipDst) && n <= 16;
}

}  // namespace

extern "C" {

// res: h rows of w residual samples, `stride` apart. out: min(h, 32) rows of min(w, 32)
// coefficients (row v = vertical frequency). vkind / hkind: 0 DCT-II, 1 DST-VII, 2 flipped
// DST-VII, 3 identity. *dropped (when given): dropped_energy's value. Returns 0, or -1 for a size
// or kind the transforms do not have.
int orc_fwd_tx(const int16_t* res, int stride, int w, int h, int vkind, int hkind, double scale, int32_t* out,
               int64_t* dropped) {
  if (!kind_ok(hkind, w) || !kind_ok(vkind, h)) return -1;
  const int cw = w > 32 ? 32 : w, ch = h > 32 ? 32 : h;
  float rows[64 * 64];  // rows[y * w + u]: horizontal frequency u of row y
  float col[64], tmp[256];
  for (int y = 0; y < h; y++) {
    float* r = rows + y * w;
    for (int x = 0; x < w; x++) r[x] = static_cast<float>(res[y * stride + x]);
    pass1d(r, w, hkind, tmp);
  }
  float cols[32 * 32];  // cols[v * cw + u]
  for (int u = 0; u < cw; u++) {
    for (int y = 0; y < h; y++) col[y] = rows[y * w + u];
    pass1d(col, h, vkind, tmp);
    for (int v = 0; v < ch; v++) cols[v * cw + u] = col[v];
  }
  // The orthonormal factors: per column (horizontal DCT) and, with the scale, per row.
  float ku[32];
  for (int u = 0; u < cw; u++)
    ku[u] = hkind != kDct ? 1.0f : static_cast<float>(std::sqrt((u ? 2.0 : 1.0) / w));
#if ORC_TX_PLANT == 3
  if (hkind == kDct) ku[0] = static_cast<float>(std::sqrt(2.0 / w));  // DC normalised as an AC column
#endif
  const float kv0 = vkind != kDct ? static_cast<float>(scale) : static_cast<float>(scale * std::sqrt(1.0 / h));
  const float kv = vkind != kDct ? static_cast<float>(scale) : static_cast<float>(scale * std::sqrt(2.0 / h));
  for (int v = 0; v < ch; v++) {
    const float k = v ? kv : kv0;
    for (int u = 0; u < cw; u++) {
      const float x = cols[v * cw + u] * ku[u] * k;
      out[v * cw + u] = static_cast<int32_t>(x + (x >= 0 ? 0.5f : -0.5f));
    }
  }
  if (dropped) {
    *dropped = 0;
    if (w > 32 || h > 32) {
      int64_t all = 0, kept = 0;
      for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++) all += static_cast<int64_t>(res[y * stride + x]) * res[y * stride + x];
      for (int i = 0; i < cw * ch; i++) kept += static_cast<int64_t>(out[i]) * out[i];
      const int64_t d = std::llround(static_cast<double>(all) * scale * scale - static_cast<double>(kept));
      *dropped = d > 0 ? d : 0;
    }
  }
  return 0;
}

}  // extern "C"

extern "C" {

// The candidate blocks of superblocks in raster order, 425 each (every square of 64, 32, 16 and 8
// samples, then its two horizontal and two vertical halves), jobs first .. first + n - 1 of a
// residual plane (sb_rows * 64 rows, `stride` samples a row): each block's DCT-II at its own size,
// coefficients packed one job after another into out. The CPU side of the transform benchmark.
long long orc_fwd_tx_frame(const int16_t* plane, int stride, int sb_cols, int sb_rows, long long first, long long n,
                           double scale, int32_t* out) {
  int geo[425][4];
  int k = 0;
  for (int s = 64; s >= 8; s /= 2)
    for (int y = 0; y < 64; y += s)
      for (int x = 0; x < 64; x += s) {
        const int b[5][4] = {{x, y, s, s}, {x, y, s, s / 2}, {x, y + s / 2, s, s / 2}, {x, y, s / 2, s}, {x + s / 2, y, s / 2, s}};
        for (const auto& g : b) geo[k][0] = g[0], geo[k][1] = g[1], geo[k][2] = g[2], geo[k][3] = g[3], k++;
      }
  long long written = 0;
  for (long long j = first; j < first + n; j++) {
    const long long sb = (j / 425) % (static_cast<long long>(sb_cols) * sb_rows);
    const int* g = geo[j % 425];
    const long long sx = (sb % sb_cols) * 64 + g[0], sy = (sb / sb_cols) * 64 + g[1];
    orc_fwd_tx(plane + sy * stride + sx, stride, g[2], g[3], 0, 0, scale, out + written, nullptr);
    written += static_cast<long long>(g[2] > 32 ? 32 : g[2]) * (g[3] > 32 ? 32 : g[3]);
  }
  return written;
}

}  // extern "C"
