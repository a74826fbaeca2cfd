/* C ABI of libcelv2synth.so: deterministic synthetic 4:2:0 clips for the decision-engine workloads.
 *
 * Integer arithmetic only (hash-based value noise, fixed-point motion, Irwin-Hall grain), so the
 * same parameters give the same samples on every machine. Frames are written padded to a multiple
 * of 8 by edge replication (make_frame / pad_frame, src/io/y4m.cpp:12-54), at the requested bit
 * depth, plus the 8-bit luma the decisions read (round_to_8bit, src/io/y4m.cpp:28-34).
 */
#ifndef CELV2SYNTH_H_
#define CELV2SYNTH_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum celv2_synth_kind {
  CELV2_SYNTH_NATURAL = 0, /* value-noise background, textured moving objects, grain */
  CELV2_SYNTH_CEL = 1      /* flat Voronoi cells, outlines, sky gradient, cuts (animation) */
};

typedef struct celv2_synth_params {
  int32_t kind;
  uint64_t seed;
  int32_t width, height, bit_depth;
  int32_t base_period;    /* background noise base period, px (4 octaves) */
  int32_t nobjects;       /* moving foreground objects */
  int32_t obj_min, obj_max; /* object side range, px */
  int32_t vmax8;          /* max |velocity| per axis, 1/8 px per frame */
  int32_t pan_x8, pan_y8; /* global pan, 1/8 px per frame */
  int32_t rotate;         /* nonzero: objects turn by a small seeded angle per frame */
  int32_t ellipses;       /* nonzero: half the objects are ellipses */
  int32_t sigma_x16;      /* grain standard deviation in 1/16 of a 10-bit code value */
  int32_t fade_first;     /* first frame of a 4-frame fade to black (-1: none) */
  int32_t cut_every;      /* CEL: frames per scene (hard cuts), 0: one scene */
  int32_t hold_odd;       /* CEL: objects hold still on odd frames */
} celv2_synth_params;

/* Fills p with the workload's generator settings:
 *  0: 352x288 10-bit, noise period 64, 3 textured rectangles 32x32..96x64, |v| <= 6 px, sigma 4;
 *  1: 1920x1080 10-bit (or the given size), period 128, pan (2.25, -0.75), 12 ellipses/rectangles
 *     24..256 px, |v| <= 16 px, slight rotation, sigma 6, fade to black at frames 24-27;
 *  4: 3840x2160 (or the given size) flat cells, 16-colour palette, 2-px outlines, cuts every 16
 *     frames, integer motion |v| <= 8 px held every other frame, sigma 1.
 * width/height 0: the workload's own. Returns 0, or -1 for an unknown workload. */
int celv2_synth_config(int workload, uint64_t seed, int width, int height, celv2_synth_params* p);

/* Frame t of the clip. Luma planes are pw x ph (pw = (width+7)&~7), chroma (pw/2) x (ph/2).
 * y16/u16/v16 (bit_depth > 8) and y8 (rounded luma) may each be NULL. threads <= 0: one per core. */
int celv2_synth_frame(const celv2_synth_params* p, int t, uint16_t* y16, uint16_t* u16, uint16_t* v16, uint8_t* y8,
                      int threads);

#ifdef __cplusplus
}
#endif

#endif /* CELV2SYNTH_H_ */
