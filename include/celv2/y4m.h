// y4m input and output for the engine outside the CelV2 tree: 4:2:0 at 8 or 10 bits, frames padded to
// a multiple of 8 by edge replication, plus the 8-bit planes the decisions read.
//
// The same functions and behaviour as CelV2's src/io/y4m.{h,cpp} (make_frame, round_to_8bit,
// pad_frame: y4m.cpp:12-55; Y4mReader: y4m.h:32-56, y4m.cpp:56-158; write_y4m_frame_bytes:
// y4m.cpp:160-179), so a standalone caller feeds HipEngine the frames CelV2's pipeline would. In-tree
// builds use CelV2's own (this file is not compiled with -DCELV2_IN_TREE).
#pragma once
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
#include <cstdio>
#include <string>
#include <vector>

#include "celv2/engine_api.h"

// Planes for a visible size padded to a multiple of 8, zeroed; above 8 bits also hy/hu/hv.
// std::runtime_error for a size that is not positive or a depth other than 8 or 10.
Frame420 make_frame(int width, int height, int bit_depth = 8);
// y/u/v from hy/hu/hv, each sample rounded to 8 bits (nothing at 8 bits).
void round_to_8bit(Frame420& f);
// Edge replication from the visible area into the padding, every plane.
void pad_frame(Frame420& f);

class Y4mReader {
 public:
  // path "-": standard input. std::runtime_error unless a y4m stream of 4:2:0 at 8 bits (C420,
  // C420jpeg, C420mpeg2, C420paldv, or no C tag) or 10 bits (C420p10, 16-bit little-endian samples).
  explicit Y4mReader(const std::string& path);
  ~Y4mReader();
  Y4mReader(const Y4mReader&) = delete;
  Y4mReader& operator=(const Y4mReader&) = delete;
  int width() const { return w_; }
  int height() const { return h_; }
  int fps_num() const { return fps_num_; }
  int fps_den() const { return fps_den_; }
  // The depth read() returns: the file's unless set (8 -> 10 shifts left by 2, 10 -> 8 rounds).
  int bit_depth() const { return out_depth_; }
  void set_bit_depth(int d);  // std::invalid_argument unless 8 or 10
  // The next frame, padded, with its 8-bit planes; false at the end of the stream. std::runtime_error
  // for a bad frame marker, a truncated frame or a sample above the file's depth.
  bool read(Frame420& out);

 private:
  std::FILE* f_ = nullptr;
  bool owned_ = false;
  int w_ = 0, h_ = 0, fps_num_ = 30, fps_den_ = 1, depth_ = 8, out_depth_ = 8;
  std::vector<uint8_t> raw_;  // one frame's payload
};

// The visible area as one y4m frame payload (no "FRAME" line); 16-bit little-endian above 8 bits.
std::vector<uint8_t> write_y4m_frame_bytes(const Frame420& f);

extern "C" {
#endif

/* C ABI over the reader (0 on success, 1 for invalid arguments, 2 for a stream error;
 * celv2_y4m_last_error() describes the failure). */
typedef struct celv2_y4m celv2_y4m;
/* info: width, height, bit depth, fps numerator, fps denominator. */
int celv2_y4m_open(const char* path, celv2_y4m** out, int32_t info[5]);
int celv2_y4m_set_bit_depth(celv2_y4m* r, int bit_depth);
/* The next frame into planes8 (y, then u, then v, padded: pw*ph + 2*(pw/2)*(ph/2) bytes) and, at a
 * depth above 8, planes16 (the same layout in 16-bit samples; may be NULL). *got = 0 at the end. */
int celv2_y4m_read(celv2_y4m* r, uint8_t* planes8, uint16_t* planes16, int32_t* got);
void celv2_y4m_close(celv2_y4m* r);
/* write_y4m_frame_bytes of a frame given as padded planes (the layout celv2_y4m_read fills; planes16
 * is read above 8 bits, planes8 at 8): the payload into out (cap bytes), its size in *n. */
int celv2_y4m_frame_bytes(const uint8_t* planes8, const uint16_t* planes16, int width, int height, int bit_depth,
                          uint8_t* out, size_t cap, size_t* n);
const char* celv2_y4m_last_error(void);

#ifdef __cplusplus
}
#endif
