/* The frame loop's coding schedule: keyframes, lone frames and mini-GOPs with a hidden anchor and a
 * pyramid of hidden middle frames, each coded frame's qindex delta and its ranked references.
 *
 * It restates what CelV2's frame loop decides before any frame is coded (ChunkCoder::code,
 * src/chunk_coder.cpp:155-616, its LoopConfig src/chunk_coder.h:85-100, and the defaults main derives,
 * src/main.cpp:739-740) and what CodecState does with the reference slots (begin_frame,
 * src/state/codec_state.cpp:476-498):
 *   - the chunk's frames are taken in display order; a keyframe is coded alone, a run of 2 or more
 *     non-key frames (at most bframes + 1) is one mini-GOP, a run of 1 a lone frame;
 *   - a mini-GOP codes its last frame first, hidden (the anchor), then the middle frames of each gap
 *     of 3 or more frames, hidden, depth first (the pyramid), then the others in display order, shown
 *     and not kept as references; a hidden frame is shown later by a show-existing entry;
 *   - there are 8 slots; slots 0..refs-1 take turns: a reference frame refreshes the turn holding the
 *     earliest frame in display order (starting from the turn count since the keyframe), and a shown
 *     one also every slot from refs up, which a shown non-reference frame refreshes only when kept;
 *     a keyframe refreshes all;
 *   - a frame predicts from the distinct frames the slots hold, nearest in display order first, at
 *     most refs of them.
 * The ranking of equally distant frames (this code: the earlier frame first) follows AV2's reference
 * ranking (av2_get_ref_frames), which is not part of this tree.
 *
 * Left to the caller: content analysis (the still-share qindex boosts of keyframes and anchors,
 * --gop-still, --gop-change, TPL's qindex and superblock offsets, the GOP trial); scene cuts come in
 * as keyframe flags.
 */
#ifndef CELV2_GOP_H_
#define CELV2_GOP_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CELV2_GOP_SLOTS 8
#define CELV2_GOP_MAX_REFS 7

/* LoopConfig's schedule fields (src/chunk_coder.h:86-92) and SeqConfig::refs. */
typedef struct celv2_gop_config {
  int32_t refs;         /* reference frames a frame predicts from, and turn slots (1..7) */
  int32_t bframes;      /* frames between reference frames (0..15): mini-GOPs of up to bframes + 1 */
  int32_t pyramid;      /* nonzero: hidden middle frames */
  int32_t parallel_b;   /* nonzero: only the last shown B frame of a mini-GOP is kept */
  int32_t bq, aq, mq, kq; /* qindex deltas: B frames, anchors, the first pyramid level, keyframes */
  int32_t level_step;   /* qindex step between pyramid levels */
  int32_t level_step16; /* the same in mini-GOPs over 8 frames (-1: level_step) */
  int32_t qindex;       /* the sequence qindex (1..255) */
} celv2_gop_config;

/* The defaults celv2enc derives for a sequence (src/main.cpp:739-740, LoopConfig's initialisers):
 * bframes 0 with one reference, else 7 with 5 or more references at qindex 100 or more (15 with
 * long_gop, as --tune and --preset max set it), else 3; parallel B frames up to twice 3840x2160;
 * bq 16, aq -16, mq 0, kq 0, level_step 8, level_step16 -1, pyramid on. */
void celv2_gop_defaults(celv2_gop_config* c, int refs, int qindex, int width, int height, int long_gop);

enum celv2_gop_type {
  CELV2_GOP_KEY = 'K',    /* keyframe: intra, refreshes every slot */
  CELV2_GOP_LONE = 'P',   /* lone inter frame: shown, a reference */
  CELV2_GOP_ANCHOR = 'A', /* a mini-GOP's last frame, coded first, hidden, a reference */
  CELV2_GOP_MID = 'M',    /* a pyramid frame, hidden, a reference */
  CELV2_GOP_LEAF = 'B'    /* a mini-GOP's other frames: shown, not references */
};

/* One entry of the stream, in stream order: a coded frame (coded = 1, the engine decides it) or the
 * show-existing entry of a hidden one (coded = 0). */
typedef struct celv2_gop_frame {
  int32_t display;      /* display index in the chunk */
  int32_t type;         /* celv2_gop_type (of the frame shown, for a show-existing entry) */
  int32_t coded, shown, reference, keep_shown;
  int32_t level;        /* pyramid level of a mid (1: the coarsest), else 0 */
  int32_t qindex_delta; /* added to the sequence qindex */
  int32_t qindex;       /* clamp(qindex + qindex_delta, 1, 255): what the engine decides at */
  int32_t refresh;      /* slots refreshed (bit i: slot i) */
  int32_t nrefs;        /* references, ranked (0 for a keyframe or a show-existing entry) */
  int32_t refs[CELV2_GOP_MAX_REFS]; /* their display indices */
} celv2_gop_frame;

/* The schedule of a chunk of `frames` frames; key[i] nonzero marks frame i a keyframe (NULL: frame
 * 0 only). Frame 0 must be a keyframe. Writes up to cap entries to out and the number of entries to
 * *count (also when cap is too small). Returns 0, 1 for invalid arguments, 2 when cap is too small;
 * celv2_gop_last_error() describes the failure. */
int celv2_gop_plan(const celv2_gop_config* c, int frames, const uint8_t* key, celv2_gop_frame* out, int cap,
                   int* count);
const char* celv2_gop_last_error(void);

#ifdef __cplusplus
}

#include <vector>

namespace celv2 {

// The C++ form: throws std::invalid_argument for a configuration or key pattern the loop rejects.
std::vector<celv2_gop_frame> gop_plan(const celv2_gop_config& c, int frames, const std::vector<bool>& key);

}  // namespace celv2
#endif

#endif /* CELV2_GOP_H_ */
