// Deterministic synthetic clips for the decision-engine workloads (include/celv2synth.h).
// Integer arithmetic throughout: hash-based value noise with a fixed-point smoothstep, fixed-point
// object motion and small-angle rotation, Irwin-Hall (sum of 4 uniforms) grain.
#include "celv2synth.h"

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

namespace {

inline uint64_t mix64(uint64_t z) {
  z += 0x9e3779b97f4a7c15ull;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}
inline uint32_t hash4(uint64_t seed, int64_t a, int64_t b, int64_t c) {
  uint64_t h = mix64(seed ^ 0x51ed270b27e5d3a1ull);
  h = mix64(h ^ static_cast<uint64_t>(a));
  h = mix64(h ^ static_cast<uint64_t>(b) * 0x9e3779b97f4a7c15ull);
  h = mix64(h ^ static_cast<uint64_t>(c) * 0xc2b2ae3d27d4eb4full);
  return static_cast<uint32_t>(h >> 32);
}

// A small deterministic generator for per-clip parameters.
struct Rng {
  uint64_t s;
  uint32_t next() { return static_cast<uint32_t>((s = mix64(s)) >> 32); }
  int uniform(int lo, int hi) { return lo + static_cast<int>(next() % static_cast<uint32_t>(hi - lo + 1)); }
};

inline int64_t floor_div(int64_t a, int64_t b) { return a >= 0 ? a / b : -((-a + b - 1) / b); }

// Smoothstep of f in Q16 (0..65536) -> Q16.
inline int64_t ease_q16(int64_t f) { return (f * f >> 16) * (3 * 65536 - 2 * f) >> 16; }

// Value noise in [0, 1023] at (x, y) given in 1/256 px, lattice period `period` px.
int value_noise(uint64_t seed, int octave, int64_t xq, int64_t yq, int period) {
  const int64_t pq = static_cast<int64_t>(period) * 256;
  const int64_t gx = floor_div(xq, pq), gy = floor_div(yq, pq);
  const int64_t fx = ((xq - gx * pq) << 16) / pq, fy = ((yq - gy * pq) << 16) / pq;
  const int64_t tx = ease_q16(fx), ty = ease_q16(fy);
  const int64_t v00 = hash4(seed, octave, gx, gy) & 1023, v10 = hash4(seed, octave, gx + 1, gy) & 1023;
  const int64_t v01 = hash4(seed, octave, gx, gy + 1) & 1023, v11 = hash4(seed, octave, gx + 1, gy + 1) & 1023;
  const int64_t top = v00 * (65536 - tx) + v10 * tx, bot = v01 * (65536 - tx) + v11 * tx;
  return static_cast<int>((top * (65536 - ty) + bot * ty + (1ll << 31)) >> 32);
}

// Four octaves (periods P, P/2, P/4, P/8; weights 8, 4, 2, 1), in [0, 1023].
int fractal_noise(uint64_t seed, int64_t xq, int64_t yq, int base) {
  int acc = 0;
  for (int o = 0; o < 4; o++) acc += (8 >> o) * value_noise(seed, o, xq, yq, std::max(1, base >> o));
  return (acc + 7) / 15;
}

// Gaussian-like grain: (sum of 4 uniform bytes - 510) has sigma 147.8; scaled to sigma_x16 / 16.
inline int grain(uint64_t seed, int t, int x, int y, int sigma_x16) {
  if (!sigma_x16) return 0;
  const uint32_t h = hash4(seed ^ 0x6a09e667f3bcc909ull, t, x, y);
  const int s = static_cast<int>((h & 255) + ((h >> 8) & 255) + ((h >> 16) & 255) + (h >> 24)) - 510;
  const int64_t v = static_cast<int64_t>(s) * sigma_x16 * 443;  // 443 / 2^16 ~ 1 / 147.8
  return static_cast<int>((v + (1 << 19)) >> 20);
}

struct Object {
  bool ellipse;
  int w, h;               // px
  int64_t cx0, cy0;       // centre at frame 0, 1/256 px
  int vx8, vy8;           // 1/8 px per frame
  int64_t theta0, omega;  // Q16 radians, per frame
  int level, contrast, period;
  uint64_t tex_seed;
};

struct Clip {
  celv2_synth_params p;
  std::vector<Object> obj;
};

Clip make_clip(const celv2_synth_params& p) {
  Clip c{p, {}};
  Rng r{p.seed * 0x2545f4914f6cdd1dull + 17};
  for (int i = 0; i < p.nobjects; i++) {
    Object o{};
    o.ellipse = p.ellipses && (r.next() & 1);
    o.w = r.uniform(p.obj_min, p.obj_max);
    o.h = r.uniform(p.obj_min, std::max(p.obj_min, p.obj_max * 2 / 3));
    o.cx0 = static_cast<int64_t>(r.uniform(0, p.width - 1)) * 256;
    o.cy0 = static_cast<int64_t>(r.uniform(0, p.height - 1)) * 256;
    o.vx8 = r.uniform(-p.vmax8, p.vmax8);
    o.vy8 = r.uniform(-p.vmax8, p.vmax8);
    o.theta0 = p.rotate ? r.uniform(-3000, 3000) : 0;  // up to ~2.6 degrees
    o.omega = p.rotate ? r.uniform(-400, 400) : 0;     // up to ~0.35 degrees per frame
    o.level = r.uniform(200, 820);
    o.contrast = r.uniform(60, 200);
    o.period = r.uniform(6, 16);
    o.tex_seed = r.next() * 0x100000001ull + i;
    c.obj.push_back(o);
  }
  return c;
}

// sin / cos of a small angle (Q16 radians) by their Taylor polynomials, Q16.
inline void sincos_q16(int64_t a, int64_t* s, int64_t* c) {
  const int64_t a2 = a * a >> 16, a3 = a2 * a >> 16, a4 = a2 * a2 >> 16;
  *s = a - a3 / 6;
  *c = 65536 - a2 / 2 + a4 / 24;
}

// 10-bit luma at visible pixel (x, y) of frame t, natural content.
int natural_luma(const Clip& c, int t, int x, int y) {
  const celv2_synth_params& p = c.p;
  const int64_t bx = static_cast<int64_t>(x) * 256 + static_cast<int64_t>(p.pan_x8) * 32 * t;
  const int64_t by = static_cast<int64_t>(y) * 256 + static_cast<int64_t>(p.pan_y8) * 32 * t;
  int v = fractal_noise(p.seed, bx, by, p.base_period);
  v = 512 + (v - 512) * 600 / 1024;  // 0..1023 mapped to 600 peak to peak around 512
  for (const Object& o : c.obj) {
    const int64_t cx = o.cx0 + static_cast<int64_t>(o.vx8) * 32 * t, cy = o.cy0 + static_cast<int64_t>(o.vy8) * 32 * t;
    const int64_t dx = static_cast<int64_t>(x) * 256 + 128 - cx, dy = static_cast<int64_t>(y) * 256 + 128 - cy;
    int64_t sn, cs;
    sincos_q16(o.theta0 + o.omega * t, &sn, &cs);
    const int64_t lx = (dx * cs + dy * sn) >> 16, ly = (-dx * sn + dy * cs) >> 16;  // 1/256 px
    const int64_t hw = static_cast<int64_t>(o.w) * 128, hh = static_cast<int64_t>(o.h) * 128;
    bool in;
    if (o.ellipse) {
      const int64_t nx = (lx << 12) / hw, ny = (ly << 12) / hh;  // Q12
      in = nx * nx + ny * ny < (1ll << 24);
    } else {
      in = lx > -hw && lx < hw && ly > -hh && ly < hh;
    }
    if (!in) continue;
    const int tex = value_noise(o.tex_seed, 0, lx + hw, ly + hh, o.period);
    v = o.level + (tex - 512) * o.contrast / 512;
  }
  if (p.fade_first >= 0 && t >= p.fade_first) {
    const int k = std::min(4, t - p.fade_first + 1);
    v = 64 + (v - 64) * (4 - k) / 4;
  }
  v += grain(p.seed, t, x, y, p.sigma_x16);
  return std::clamp(v, 0, 1023);
}

int natural_chroma(const Clip& c, int t, int x, int y, int plane) {
  const celv2_synth_params& p = c.p;
  const int64_t bx = static_cast<int64_t>(x) * 512 + static_cast<int64_t>(p.pan_x8) * 16 * t;
  const int64_t by = static_cast<int64_t>(y) * 512 + static_cast<int64_t>(p.pan_y8) * 16 * t;
  int v = 512 + (fractal_noise(p.seed ^ (0x1234567ull + plane), bx, by, p.base_period) - 512) / 3;
  v += grain(p.seed + 1 + plane, t, x, y, p.sigma_x16 / 2);
  return std::clamp(v, 0, 1023);
}

// ---- flat-shaded cel content ----------------------------------------------------------------------

inline int64_t isqrt64(int64_t v) {
  if (v <= 0) return 0;
  int64_t r = static_cast<int64_t>(__builtin_sqrt(static_cast<double>(v)));
  while (r * r > v) r--;
  while ((r + 1) * (r + 1) <= v) r++;
  return r;
}

int cel_luma(const celv2_synth_params& p, int t, int x, int y) {
  static const int palette[16] = {96, 160, 224, 288, 352, 416, 480, 544, 608, 672, 736, 800, 864, 896, 928, 960};
  const int scene = p.cut_every > 0 ? t / p.cut_every : 0, ts = p.cut_every > 0 ? t % p.cut_every : t;
  Rng r{p.seed * 0x9e3779b97f4a7c15ull + static_cast<uint64_t>(scene) * 7919 + 3};
  const int n = r.uniform(2, 6);
  int px[6], py[6], col[6];
  for (int i = 0; i < n; i++) px[i] = r.uniform(0, p.width - 1), py[i] = r.uniform(0, p.height - 1), col[i] = palette[r.next() & 15];
  int sky = 0;
  for (int i = 1; i < n; i++)
    if (py[i] < py[sky]) sky = i;
  int64_t d1 = INT64_MAX, d2 = INT64_MAX;
  int a = 0, b = 0;
  for (int i = 0; i < n; i++) {
    const int64_t d = static_cast<int64_t>(x - px[i]) * (x - px[i]) + static_cast<int64_t>(y - py[i]) * (y - py[i]);
    if (d < d1) d2 = d1, b = a, d1 = d, a = i;
    else if (d < d2) d2 = d, b = i;
  }
  int v = col[a];
  if (a == sky) v += static_cast<int>(static_cast<int64_t>(y) * 96 / std::max(1, p.height));  // slow gradient
  if (n > 1) {  // distance to the bisector of the two nearest points: an outline within 1 px
    const int64_t sep2 = static_cast<int64_t>(px[a] - px[b]) * (px[a] - px[b]) + static_cast<int64_t>(py[a] - py[b]) * (py[a] - py[b]);
    const int64_t sep = std::max<int64_t>(1, isqrt64(sep2));
    if ((d2 - d1) < 2 * sep) v = 48;
  }
  const int steps = (ts + 1) / 2;  // moves on even-to-odd frame steps, holds on the others
  const int nobj = r.uniform(3, 5);
  for (int i = 0; i < nobj; i++) {
    const int w = r.uniform(24, 160), h = r.uniform(24, 120);
    const int ox = r.uniform(0, p.width - 1) + r.uniform(-8, 8) * steps;
    const int oy = r.uniform(0, p.height - 1) + r.uniform(-8, 8) * steps;
    const int c = palette[r.next() & 15];
    const bool ell = r.next() & 1;
    const int dx = x - ox, dy = y - oy;
    bool in, edge;
    if (ell) {
      const int64_t q = static_cast<int64_t>(dx) * dx * h * h + static_cast<int64_t>(dy) * dy * w * w, lim = static_cast<int64_t>(w) * w * h * h;
      const int64_t wi = w - 4, hi = h - 4;
      const int64_t qi = static_cast<int64_t>(dx) * dx * hi * hi + static_cast<int64_t>(dy) * dy * wi * wi, limi = wi * wi * hi * hi;
      in = 4 * q < lim;
      edge = in && !(4 * qi < limi);
    } else {
      in = 2 * std::abs(dx) < w && 2 * std::abs(dy) < h;
      edge = in && !(2 * std::abs(dx) < w - 4 && 2 * std::abs(dy) < h - 4);
    }
    if (in) v = edge ? 40 : c;
  }
  v += grain(p.seed, t, x, y, p.sigma_x16);
  return std::clamp(v, 0, 1023);
}

template <typename T>
void pad_plane(T* q, int stride, int w, int h, int pw, int ph) {
  for (int r = 0; r < h; r++)
    for (int c = w; c < pw; c++) q[static_cast<size_t>(r) * stride + c] = q[static_cast<size_t>(r) * stride + w - 1];
  for (int r = h; r < ph; r++) std::memcpy(q + static_cast<size_t>(r) * stride, q + static_cast<size_t>(h - 1) * stride, sizeof(T) * pw);
}

}  // namespace

extern "C" int celv2_synth_config(int workload, uint64_t seed, int width, int height, celv2_synth_params* p) {
  if (!p) return -1;
  std::memset(p, 0, sizeof(*p));
  p->seed = seed;
  p->bit_depth = 10;
  p->fade_first = -1;
  switch (workload) {
    case 0:
      p->kind = CELV2_SYNTH_NATURAL, p->width = 352, p->height = 288, p->base_period = 64, p->nobjects = 3;
      p->obj_min = 32, p->obj_max = 96, p->vmax8 = 48, p->sigma_x16 = 64;
      break;
    case 1:
      p->kind = CELV2_SYNTH_NATURAL, p->width = 1920, p->height = 1080, p->base_period = 128, p->nobjects = 12;
      p->obj_min = 24, p->obj_max = 256, p->vmax8 = 128, p->pan_x8 = 18, p->pan_y8 = -6, p->rotate = 1, p->ellipses = 1;
      p->sigma_x16 = 96, p->fade_first = 24;
      break;
    case 4:
      p->kind = CELV2_SYNTH_CEL, p->width = 3840, p->height = 2160, p->sigma_x16 = 16, p->cut_every = 16, p->hold_odd = 1;
      break;
    default: return -1;
  }
  if (width > 0) p->width = width;
  if (height > 0) p->height = height;
  return 0;
}

extern "C" int celv2_synth_frame(const celv2_synth_params* p, int t, uint16_t* y16, uint16_t* u16, uint16_t* v16,
                                 uint8_t* y8, int threads) {
  if (!p || p->width <= 0 || p->height <= 0 || p->bit_depth < 8 || p->bit_depth > 16) return -1;
  const Clip clip = make_clip(*p);
  const int w = p->width, h = p->height, pw = (w + 7) & ~7, ph = (h + 7) & ~7;
  const int cw = (w + 1) / 2, ch = (h + 1) / 2, cpw = pw / 2, cph = ph / 2;
  std::vector<uint16_t> Y(static_cast<size_t>(pw) * ph), U(static_cast<size_t>(cpw) * cph), V(U.size());
  const int down = 10 - p->bit_depth;  // samples are made at 10 bits
  auto scale = [&](int v) { return down > 0 ? v >> down : v << -down; };
  auto rows = [&](int r0, int r1) {
    for (int y = r0; y < r1; y++) {
      for (int x = 0; x < w; x++)
        Y[static_cast<size_t>(y) * pw + x] = static_cast<uint16_t>(scale(p->kind == CELV2_SYNTH_CEL ? cel_luma(*p, t, x, y) : natural_luma(clip, t, x, y)));
      if (y < ch)
        for (int x = 0; x < cw; x++) {
          U[static_cast<size_t>(y) * cpw + x] = static_cast<uint16_t>(scale(natural_chroma(clip, t, x, y, 0)));
          V[static_cast<size_t>(y) * cpw + x] = static_cast<uint16_t>(scale(natural_chroma(clip, t, x, y, 1)));
        }
    }
  };
  int n = threads > 0 ? threads : static_cast<int>(std::thread::hardware_concurrency());
  n = std::max(1, std::min(n, 64));
  std::vector<std::thread> pool;
  for (int i = 0; i < n; i++) pool.emplace_back(rows, h * i / n, h * (i + 1) / n);
  for (auto& th : pool) th.join();
  pad_plane(Y.data(), pw, w, h, pw, ph);
  pad_plane(U.data(), cpw, cw, ch, cpw, cph);
  pad_plane(V.data(), cpw, cw, ch, cpw, cph);
  if (y16) std::memcpy(y16, Y.data(), Y.size() * 2);
  if (u16) std::memcpy(u16, U.data(), U.size() * 2);
  if (v16) std::memcpy(v16, V.data(), V.size() * 2);
  if (y8) {
    const int shift = p->bit_depth - 8;
    for (size_t i = 0; i < Y.size(); i++)
      y8[i] = static_cast<uint8_t>(shift ? std::min(255, (Y[i] + (1 << (shift - 1))) >> shift) : Y[i]);
  }
  return 0;
}
