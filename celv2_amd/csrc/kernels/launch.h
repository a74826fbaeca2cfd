// Host-side launchers of the decision kernels (one per .hip file), called by the engine.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "batch.h"
#include "cnn.h"
#include "common.h"

namespace celv2hip {

// planes.hip: every launch takes a batch of planes (add_* appends one with its workgroup count).
struct PyramidItem {  // a frame's luma (8-bit, or 9..16-bit in uint16) into bordered levels 0..4
  const void* src;
  int sstride, bit_depth, wide;
  DPlane lv[kLevels];  // level l + 1 is level l's size halved (rounded down)
};
struct SubpelItem {  // the 15 quarter-pel phase planes of a level-0 plane, phase_bytes apart
  DPlane in;
  long long phase_bytes;
};
void add_pyramid(Batch<PyramidItem>& b, const PyramidItem& t);
void add_subpel(Batch<SubpelItem>& b, const SubpelItem& t);
void launch_pyramid(const Batch<PyramidItem>& b, hipStream_t s);
void launch_subpel(const Batch<SubpelItem>& b, hipStream_t s);

// me.hip: one pyramid level's motion fields for n references at once. cur: the source's level;
// ref[r]: reference r's level (at level 0 with its phases, phase_bytes apart); coarser[r]: the level
// above's field of reference r (nullptr at the top), pcols x prows; out[r]: cols x rows packed MVs.
// Exact unsigned division n / d for n < 2^24 and 1 <= d < 2^16 on the scalar unit:
// q = mulhi(n << 8, m) >> l with l = ceil(log2 d), m = ceil(2^(24 + l) / d). Since d <= 2^l,
// m d - 2^(24 + l) < d <= 2^l, the condition under which floor(n m / 2^(24 + l)) = floor(n / d) for
// every 24-bit n (Granlund and Montgomery, "Division by invariant integers using multiplication",
// 1994, theorem 4.2); m < 2^25 + 1 fits 32 bits.
struct Divisor {
  uint32_t m;
  int l;
};
Divisor make_divisor(int d);
struct MeLevelArgs {
  DPlane cur;
  DPlane ref[kMaxRefs];
  const uint32_t* coarser[kMaxRefs];
  uint32_t* out[kMaxRefs];
  int nrefs, cols, rows, pcols, prows, level0;
  long long phase_bytes;
  unsigned long long* searches;  // centres searched, a counter of 64 slots (count(), common.h) (may be null)
  // Per-superblock reference limits (may be null): superblock s uses references 0 .. kmap[s] - 1.
  // At levels 0 and 1 a block is searched for reference r only when a superblock that keeps r reads
  // it: its own, or a neighbour whose ring it lies in (1 block at level 0, 2 at level 1).
  const uint8_t* kmap;
  // Per-superblock search modes (may be null): bit L set narrows level L's full searches (L = 0, 1)
  // to +-kNarrowRange around each centre. A block is searched narrow at level L only when every
  // superblock that keeps the reference and reads the block (the same ring as above) has bit L set.
  const uint8_t* smap;
  int sb_cols, sb_rows, level;
  uint32_t* cost_out[kMaxRefs];  // the searched best cost (SAD + distance penalty) per block (may be null)
  // The block rows this launch searches: [row0, row0 + nrun) (nrun 0: every row). rows stays the
  // level's row count (neighbour clamps, the coarser level's rows).
  int row0, nrun;
  Divisor by_nrefs, by_cols;     // exact division by nrefs and by cols (set by add_me_level)
};
void add_me_level(Batch<MeLevelArgs>& b, const MeLevelArgs& a);
void launch_me_level(const Batch<MeLevelArgs>& b, hipStream_t s);

// intra.hip: SATD of the 13 intra modes of the 420 candidate blocks of levels 1-3 of superblock
// rows [row0, row0 + nrows): out[(sb * 425 + b) * 13 + m] (blocks 0..4 untouched).
struct IntraItem {
  DPlane src;
  int sb_cols, row0, nrows;
  uint32_t* out;
  // Intra mode groups per superblock (may be null: every group); the masks used are written to
  // used_mask (one byte per superblock), which finalize then obeys.
  const uint8_t* mask;
  uint8_t* used_mask;
};
void add_intra(Batch<IntraItem>& b, const IntraItem& t);
void launch_intra(const Batch<IntraItem>& b, hipStream_t s);

// inter.hip: per superblock (rows [row0, row0 + nrows)) and reference, every candidate block's best
// single-reference SATD and quarter-pel MV: best[((sb * kMaxRefs) + r) * 425 + b] = {satd, packed mv}.
struct InterArgs {
  DPlane src;
  DPlane ref[kMaxRefs];  // level 0 with phases
  const uint32_t* field[kMaxRefs];  // level-0 fields, cols x rows
  int nrefs, cols, rows, sb_cols, row0, nrows;
  long long phase_bytes;
  uint2* best;
  unsigned long long* counter;     // unique MVs scored, a counter of 64 slots (may be null)
  unsigned long long* candidates;  // non-repeated (block, MV) candidates scored, 64 slots (may be null)
  const uint8_t* kmap;             // per-superblock reference limits (may be null)
};
void add_inter(Batch<InterArgs>& b, const InterArgs& a);
void launch_inter(const Batch<InterArgs>& b, hipStream_t s);

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
  unsigned long long* counter;  // compound blocks scored, a counter of 64 slots (may be null)
  const uint8_t* kmap;     // per-superblock reference limits (may be null): n = min(nrefs, kmap[sb])
  const uint8_t* imask;    // the intra mode groups the source's intra scores were made with (k_intra's used_mask)
  int64_t* sb_cost;        // the superblock's decision cost (the 64x64 square's DP cost) (may be null)
};
void add_finalize(Batch<FinalizeArgs>& b, const FinalizeArgs& a);
void launch_finalize(const Batch<FinalizeArgs>& b, hipStream_t s);

// pruner.hip: the learned pruner's per-superblock inputs, from the source's level 2 and each
// reference's level-2 search costs: feats[sb * kFeatures + f].
struct FeatureArgs {
  DPlane src2;                      // the source's level 2
  const uint32_t* cost2[kMaxRefs];  // level-2 best cost per block, cols2 x rows2
  int nrefs, cols2, rows2, sb_cols, sb_rows;
  long long lambda_q8;  // the frame's rate multiplier
  float* feats;
};
// context: also the frame context f[11..13] (else the network's launch computes it, cnn_infer).
void launch_features(const FeatureArgs& a, hipStream_t s, bool context = true);

// Constant tables the kernels read: each builds its tables on the host and copies them to the
// current device's symbols before returning (once per device, from the engine's creation).
hipError_t upload_intra_tiles();
hipError_t upload_inter_tables();
hipError_t upload_final_tiles();
hipError_t upload_tx_tables();

// transform.hip: forward transforms of residual blocks. A job is one block (res: h rows of w int16
// samples, stride apart; coeff: min(h, 32) rows of min(w, 32) coefficients; dropped may be null);
// kv0 / kv: the vertical factor of frequency 0 / the others with the scale in it (as floats; the
// scale itself for a vertical kind other than DCT-II). A unit is `count` consecutive jobs of one size
// and kind pair, count <= tx_jobs_per_unit(hl); one wave per unit.
struct TxJob {
  const int16_t* res;
  int32_t* coeff;
  long long* dropped;
  double scale;
  int stride;
  float kv0, kv;
};
struct TxUnit {
  int first, count;
  int wl, hl, vk, hk;  // log2 w, log2 h (2..6); vertical and horizontal kinds
};
int tx_jobs_per_unit(int hl);
void launch_fwd_tx(const TxJob* jobs, const TxUnit* units, int nunits, hipStream_t s);

}  // namespace celv2hip
