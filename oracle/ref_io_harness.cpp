// ORACLE PINNING HARNESS — test infrastructure. Compiled only by `make -C oracle ref` where the CelV2
// reference is mounted (REF=/root/reference), together with the reference's own src/io/y4m.cpp (it
// builds from its own two files); it exports the reference's y4m reader under ref_y4m_* names with the
// same argument meaning as the product's celv2_y4m_* C ABI, so tests can compare the two on the same
// files.
#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>

#include "io/y4m.h"

namespace {
thread_local std::string g_err;
}

extern "C" {

// Reads up to max_frames frames of path at out_depth (0: the file's) into planes8 / planes16 (each
// frame's y, u, v padded planes back to back; planes16 may be NULL). info: width, height, depth,
// fps_num, fps_den. Returns the frames read, or -1 (ref_y4m_error() describes it).
int ref_y4m_read(const char* path, int out_depth, int max_frames, uint8_t* planes8, uint16_t* planes16, int32_t* info) {
  try {
    Y4mReader r(path);
    if (out_depth) r.set_bit_depth(out_depth);
    info[0] = r.width(), info[1] = r.height(), info[2] = r.bit_depth(), info[3] = r.fps_num(), info[4] = r.fps_den();
    Frame420 f;
    int n = 0;
    for (; n < max_frames && r.read(f); n++) {
      const size_t per = f.y.size() + f.u.size() + f.v.size();
      uint8_t* p8 = planes8 + per * n;
      std::memcpy(p8, f.y.data(), f.y.size());
      std::memcpy(p8 + f.y.size(), f.u.data(), f.u.size());
      std::memcpy(p8 + f.y.size() + f.u.size(), f.v.data(), f.v.size());
      if (planes16 && f.bit_depth > 8) {
        uint16_t* p16 = planes16 + per * n;
        std::memcpy(p16, f.hy.data(), f.hy.size() * 2);
        std::memcpy(p16 + f.hy.size(), f.hu.data(), f.hu.size() * 2);
        std::memcpy(p16 + f.hy.size() + f.hu.size(), f.hv.data(), f.hv.size() * 2);
      }
    }
    return n;
  } catch (const std::exception& e) {
    g_err = e.what();
    return -1;
  }
}

const char* ref_y4m_error(void) { return g_err.c_str(); }

}  // extern "C"
