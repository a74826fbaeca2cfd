// Host-side launchers of the decision kernels (one per .hip file), called by the engine.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "common.h"

namespace celv2hip {

// planes.hip
void launch_import8(const uint8_t* src, int sstride, const DPlane& out, hipStream_t s);
void launch_import16(const uint16_t* src, int sstride, int bit_depth, const DPlane& out, hipStream_t s);
void launch_downsample(const DPlane& in, const DPlane& out, hipStream_t s);
void launch_subpel(const DPlane& level0, long long phase_bytes, hipStream_t s);

// me.hip: one pyramid level's motion fields for n references at once. cur: the source's level;
// ref[r]: reference r's level (at level 0 with its phases, phase_bytes apart); coarser[r]: the level
// above's field of reference r (nullptr at the top), pcols x prows; out[r]: cols x rows packed MVs.
struct MeLevelArgs {
  DPlane cur;
  DPlane ref[kMaxRefs];
  const uint32_t* coarser[kMaxRefs];
  uint32_t* out[kMaxRefs];
  int nrefs, cols, rows, pcols, prows, level0;
  long long phase_bytes;
  unsigned long long* searches;  // centres searched, spread over 64 slots 16 words apart (may be null)
};
void launch_me_level(const MeLevelArgs& a, hipStream_t s);

// intra.hip: SATD of the 13 intra modes of the 420 candidate blocks of levels 1-3 of superblock
// rows [row0, row0 + nrows): out[(sb * 425 + b) * 13 + m] (blocks 0..4 untouched).
void launch_intra(const DPlane& src, int sb_cols, int row0, int nrows, uint32_t* out, hipStream_t s);

// inter.hip: per superblock (rows [row0, row0 + nrows)) and reference, every candidate block's best
// single-reference SATD and quarter-pel MV: best[((sb * kMaxRefs) + r) * 425 + b] = {satd, packed mv}.
struct InterArgs {
  DPlane src;
  DPlane ref[kMaxRefs];  // level 0 with phases
  const uint32_t* field[kMaxRefs];  // level-0 fields, cols x rows
  int nrefs, cols, rows, sb_cols, row0, nrows;
  long long phase_bytes;
  uint2* best;
  unsigned long long* counter;     // unique MVs scored (may be null)
  unsigned long long* candidates;  // non-repeated (block, MV) candidates scored (may be null)
};
void launch_inter(const InterArgs& a, hipStream_t s);

// finalize.hip: per superblock, the reference choice, compound, intra choice, partition DP and the
// exported record (include/celv2hip.h celv2_sb) at out[sb].
struct FinalizeArgs {
  DPlane src;
  DPlane ref[kMaxRefs];
  int nrefs, sb_cols, row0, nrows;
  long long phase_bytes;
  int64_t lam;
  const uint2* best;       // from launch_inter (unused when nrefs == 0)
  const uint32_t* intra;   // from launch_intra
  void* out;               // celv2_sb records, one per superblock of the frame
  unsigned long long* counter;  // compound blocks scored (may be null)
};
void launch_finalize(const FinalizeArgs& a, hipStream_t s);

// Static tables the kernels read (built once per process on the host, copied to the device).
void upload_tables(hipStream_t s);

}  // namespace celv2hip
