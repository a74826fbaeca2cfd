// y4m input and output (include/celv2/y4m.h): the behaviour of CelV2's src/io/y4m.cpp for builds
// outside the CelV2 tree.
#include "celv2/y4m.h"

#include <algorithm>
#include <cstring>
#include <stdexcept>

namespace {

// Right columns from the last visible one, then bottom rows from the last visible row.
template <typename T>
void replicate_edges(T* p, int stride, int w, int h, int pw, int ph) {
  for (int r = 0; r < h; r++) std::fill(p + static_cast<size_t>(r) * stride + w, p + static_cast<size_t>(r) * stride + pw, p[static_cast<size_t>(r) * stride + w - 1]);
  for (int r = h; r < ph; r++) std::copy_n(p + static_cast<size_t>(h - 1) * stride, pw, p + static_cast<size_t>(r) * stride);
}

// The integer after a one-letter tag ("W1920"); std::runtime_error when there is none.
int tag_int(const std::string& t, size_t from) {
  size_t used = 0;
  const int v = std::stoi(t.substr(from), &used);
  (void)used;
  return v;
}

}  // namespace

Frame420 make_frame(int width, int height, int bit_depth) {
  if (width <= 0 || height <= 0) throw std::runtime_error("frame size must be positive");
  if (bit_depth != 8 && bit_depth != 10) throw std::runtime_error("bit depth must be 8 or 10");
  Frame420 f;
  f.width = width, f.height = height, f.bit_depth = bit_depth;
  f.pw = (width + 7) & ~7, f.ph = (height + 7) & ~7;
  const size_t luma = static_cast<size_t>(f.pw) * f.ph, chroma = static_cast<size_t>(f.pw / 2) * (f.ph / 2);
  f.y.assign(luma, 0), f.u.assign(chroma, 0), f.v.assign(chroma, 0);
  if (bit_depth > 8) f.hy.assign(luma, 0), f.hu.assign(chroma, 0), f.hv.assign(chroma, 0);
  return f;
}

void round_to_8bit(Frame420& f) {
  if (f.bit_depth <= 8) return;
  const int s = f.bit_depth - 8, half = 1 << (s - 1);
  const std::vector<uint16_t>* in[3] = {&f.hy, &f.hu, &f.hv};
  std::vector<uint8_t>* out[3] = {&f.y, &f.u, &f.v};
  for (int k = 0; k < 3; k++) {
    const uint16_t* a = in[k]->data();
    uint8_t* b = out[k]->data();
    for (size_t i = 0, n = in[k]->size(); i < n; i++) b[i] = static_cast<uint8_t>(std::min(255, (a[i] + half) >> s));
  }
}

void pad_frame(Frame420& f) {
  const int cw = (f.width + 1) / 2, ch = (f.height + 1) / 2, cs = f.stride_uv();
  replicate_edges(f.y.data(), f.stride_y(), f.width, f.height, f.pw, f.ph);
  replicate_edges(f.u.data(), cs, cw, ch, f.pw / 2, f.ph / 2);
  replicate_edges(f.v.data(), cs, cw, ch, f.pw / 2, f.ph / 2);
  if (f.bit_depth > 8) {
    replicate_edges(f.hy.data(), f.stride_y(), f.width, f.height, f.pw, f.ph);
    replicate_edges(f.hu.data(), cs, cw, ch, f.pw / 2, f.ph / 2);
    replicate_edges(f.hv.data(), cs, cw, ch, f.pw / 2, f.ph / 2);
  }
}

Y4mReader::Y4mReader(const std::string& path) {
  if (path == "-") {
    f_ = stdin;
  } else {
    f_ = std::fopen(path.c_str(), "rb");
    if (!f_) throw std::runtime_error("cannot open " + path);
    owned_ = true;
  }
  try {
    std::string line;
    int c;
    while ((c = std::fgetc(f_)) != EOF && c != '\n') line.push_back(static_cast<char>(c));
    if (line.empty() && c == EOF) throw std::runtime_error("empty y4m input: " + path);
    // Space-separated tokens: the signature, then one-letter tags (W, H, F num:den, C colour space;
    // the others are ignored).
    std::vector<std::string> tok;
    for (size_t i = 0; i < line.size();) {
      const size_t j = std::min(line.find(' ', i), line.size());
      if (j > i) tok.push_back(line.substr(i, j - i));
      i = j + 1;
    }
    if (tok.empty() || tok[0] != "YUV4MPEG2") throw std::runtime_error("not a y4m file: " + path);
    std::string colour = "420";
    for (size_t k = 1; k < tok.size(); k++) {
      const std::string& t = tok[k];
      if (t[0] == 'W') {
        w_ = tag_int(t, 1);
      } else if (t[0] == 'H') {
        h_ = tag_int(t, 1);
      } else if (t[0] == 'F') {
        const size_t colon = t.find(':');
        fps_num_ = tag_int(t.substr(0, colon), 1);
        fps_den_ = tag_int(t, colon + 1);
      } else if (t[0] == 'C') {
        colour = t.substr(1);
      }
    }
    if (colour == "420p10") {
      depth_ = 10;
    } else if (colour != "420" && colour != "420jpeg" && colour != "420mpeg2" && colour != "420paldv") {
      throw std::runtime_error("unsupported y4m colorspace C" + colour + " (only 4:2:0 at 8 or 10 bits)");
    }
    if (w_ <= 0 || h_ <= 0) throw std::runtime_error("y4m header has no size");
    out_depth_ = depth_;
  } catch (...) {
    if (owned_) std::fclose(f_);
    throw;
  }
}

Y4mReader::~Y4mReader() {
  if (owned_) std::fclose(f_);
}

void Y4mReader::set_bit_depth(int d) {
  if (d != 8 && d != 10) throw std::invalid_argument("bit depth must be 8 or 10");
  out_depth_ = d;
}

bool Y4mReader::read(Frame420& out) {
  int c = std::fgetc(f_);
  if (c == EOF) return false;
  std::string marker;
  for (; c != EOF && c != '\n'; c = std::fgetc(f_)) marker.push_back(static_cast<char>(c));
  if (marker.compare(0, 5, "FRAME") != 0) throw std::runtime_error("bad y4m frame marker");
  if (out.width != w_ || out.height != h_ || out.bit_depth != depth_) out = make_frame(w_, h_, depth_);
  const int cw = (w_ + 1) / 2, ch = (h_ + 1) / 2, bytes = depth_ > 8 ? 2 : 1;
  const size_t payload = (static_cast<size_t>(w_) * h_ + 2 * static_cast<size_t>(cw) * ch) * bytes;
  raw_.resize(payload);
  if (std::fread(raw_.data(), 1, payload, f_) != payload) throw std::runtime_error("truncated y4m frame");
  // Rows of the payload into the padded planes.
  const uint8_t* src = raw_.data();
  auto place = [&](auto* dst, int stride, int w, int h) {
    const size_t row = static_cast<size_t>(w) * bytes;
    for (int r = 0; r < h; r++, src += row) std::memcpy(dst + static_cast<size_t>(r) * stride, src, row);
  };
  if (depth_ > 8) {
    place(out.hy.data(), out.stride_y(), w_, h_);
    place(out.hu.data(), out.stride_uv(), cw, ch);
    place(out.hv.data(), out.stride_uv(), cw, ch);
    const int limit = 1 << depth_;
    for (const std::vector<uint16_t>* p : {&out.hy, &out.hu, &out.hv})
      for (uint16_t s : *p)
        if (s >= limit) throw std::runtime_error("y4m sample above the bit depth");
  } else {
    place(out.y.data(), out.stride_y(), w_, h_);
    place(out.u.data(), out.stride_uv(), cw, ch);
    place(out.v.data(), out.stride_uv(), cw, ch);
  }
  pad_frame(out);
  round_to_8bit(out);
  if (out_depth_ == depth_) return true;
  if (out_depth_ > 8) {  // 8 -> 10: the samples shifted left, the 8-bit planes as they are
    const int s = out_depth_ - 8;
    out.bit_depth = out_depth_;
    out.hy.resize(out.y.size()), out.hu.resize(out.u.size()), out.hv.resize(out.v.size());
    for (size_t i = 0; i < out.y.size(); i++) out.hy[i] = static_cast<uint16_t>(out.y[i] << s);
    for (size_t i = 0; i < out.u.size(); i++) out.hu[i] = static_cast<uint16_t>(out.u[i] << s);
    for (size_t i = 0; i < out.v.size(); i++) out.hv[i] = static_cast<uint16_t>(out.v[i] << s);
  } else {  // 10 -> 8: the rounded planes
    out.bit_depth = 8;
    out.hy.clear(), out.hu.clear(), out.hv.clear();
  }
  return true;
}

std::vector<uint8_t> write_y4m_frame_bytes(const Frame420& f) {
  const int cw = (f.width + 1) / 2, ch = (f.height + 1) / 2;
  const bool wide = f.bit_depth > 8;
  std::vector<uint8_t> b;
  b.reserve((static_cast<size_t>(f.width) * f.height + 2 * static_cast<size_t>(cw) * ch) * (wide ? 2 : 1));
  auto put = [&](const auto* p, int stride, int w, int h) {
    for (int r = 0; r < h; r++) {
      const uint8_t* row = reinterpret_cast<const uint8_t*>(p + static_cast<size_t>(r) * stride);
      b.insert(b.end(), row, row + static_cast<size_t>(w) * sizeof(*p));
    }
  };
  if (wide) {
    put(f.hy.data(), f.stride_y(), f.width, f.height);
    put(f.hu.data(), f.stride_uv(), cw, ch);
    put(f.hv.data(), f.stride_uv(), cw, ch);
  } else {
    put(f.y.data(), f.stride_y(), f.width, f.height);
    put(f.u.data(), f.stride_uv(), cw, ch);
    put(f.v.data(), f.stride_uv(), cw, ch);
  }
  return b;
}

struct celv2_y4m {
  Y4mReader r;
  Frame420 f;
  explicit celv2_y4m(const char* path) : r(path) {}
};

namespace {
thread_local std::string g_y4m_err;

template <typename F>
int y4m_guard(F&& fn) {
  try {
    fn();
    return 0;
  } catch (const std::invalid_argument& e) {
    g_y4m_err = e.what();
    return 1;
  } catch (const std::exception& e) {
    g_y4m_err = e.what();
    return 2;
  }
}
}  // namespace

extern "C" {

int celv2_y4m_open(const char* path, celv2_y4m** out, int32_t info[5]) {
  return y4m_guard([&] {
    if (!path || !out) throw std::invalid_argument("y4m: null argument");
    *out = nullptr;
    auto* r = new celv2_y4m(path);
    if (info) {
      info[0] = r->r.width(), info[1] = r->r.height(), info[2] = r->r.bit_depth();
      info[3] = r->r.fps_num(), info[4] = r->r.fps_den();
    }
    *out = r;
  });
}

int celv2_y4m_set_bit_depth(celv2_y4m* r, int bit_depth) {
  return y4m_guard([&] {
    if (!r) throw std::invalid_argument("y4m: null reader");
    r->r.set_bit_depth(bit_depth);
  });
}

int celv2_y4m_read(celv2_y4m* r, uint8_t* planes8, uint16_t* planes16, int32_t* got) {
  return y4m_guard([&] {
    if (!r || !planes8 || !got) throw std::invalid_argument("y4m: null argument");
    *got = 0;
    if (!r->r.read(r->f)) return;
    const Frame420& f = r->f;
    std::memcpy(planes8, f.y.data(), f.y.size());
    std::memcpy(planes8 + f.y.size(), f.u.data(), f.u.size());
    std::memcpy(planes8 + f.y.size() + f.u.size(), f.v.data(), f.v.size());
    if (planes16 && f.bit_depth > 8) {
      std::copy(f.hy.begin(), f.hy.end(), planes16);
      std::copy(f.hu.begin(), f.hu.end(), planes16 + f.hy.size());
      std::copy(f.hv.begin(), f.hv.end(), planes16 + f.hy.size() + f.hu.size());
    }
    *got = 1;
  });
}

void celv2_y4m_close(celv2_y4m* r) { delete r; }

int celv2_y4m_frame_bytes(const uint8_t* planes8, const uint16_t* planes16, int width, int height, int bit_depth,
                          uint8_t* out, size_t cap, size_t* n) {
  return y4m_guard([&] {
    if (!n || (bit_depth > 8 ? !planes16 : !planes8)) throw std::invalid_argument("y4m: null argument");
    if (bit_depth != 8 && bit_depth != 10) throw std::invalid_argument("bit depth must be 8 or 10");
    Frame420 f = make_frame(width, height, bit_depth);
    const size_t ys = f.y.size(), cs = f.u.size();
    if (bit_depth > 8) {
      std::copy_n(planes16, ys, f.hy.begin());
      std::copy_n(planes16 + ys, cs, f.hu.begin());
      std::copy_n(planes16 + ys + cs, cs, f.hv.begin());
    } else {
      std::copy_n(planes8, ys, f.y.begin());
      std::copy_n(planes8 + ys, cs, f.u.begin());
      std::copy_n(planes8 + ys + cs, cs, f.v.begin());
    }
    const std::vector<uint8_t> b = write_y4m_frame_bytes(f);
    *n = b.size();
    if (b.size() > cap || !out) throw std::invalid_argument("y4m: output buffer too small");
    std::memcpy(out, b.data(), b.size());
  });
}

const char* celv2_y4m_last_error(void) { return g_y4m_err.c_str(); }

}  // extern "C"
