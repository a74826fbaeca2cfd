/* C ABI of libcelv2hip.so: the CelV2 decision engine on one AMD Instinct MI355X (gfx950).
 *
 * The engine replaces the Vulkan engine behind CelV2's engine interface `DecisionEngine`
 * (src/gpu/decision_engine.h:9-51). Its C++ face is `celv2::HipEngine` (include/celv2/hip_engine.h);
 * this header is the thin layer underneath it: plain pointers and sizes, an opaque handle,
 * integer status codes and a thread-local error string. Nothing here names a C++ or torch type.
 *
 * Units and conventions follow the reference engine:
 *   - frames are the padded 8-bit luma plane the decisions use (Frame420::y, src/io/y4m.h:11-23):
 *     pw x ph samples, pw = (width + 7) & ~7, ph = (height + 7) & ~7, stride pw; or the same plane
 *     at more than 8 bits (Frame420::hy), rounded on the device by min(255, (v + 2^(d-9)) >> (d-8))
 *     exactly as round_to_8bit (src/io/y4m.cpp:28-34);
 *   - motion vectors in the exported records are 1/8 pel (BlockDecision, src/gpu/decision_map.h:22-36);
 *   - costs are SATD * 256 + lambda_q8 * bits (src/gpu/kernels_ref.cpp:46-77); lambda_q8 is an input
 *     (the reference computes it on the host, src/vk/vk_engine.cpp:253).
 */
#ifndef CELV2HIP_H_
#define CELV2HIP_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CELV2_SB_CAP 256    /* tree entries / leaves per 64x64 superblock (src/gpu/sb_out.h:11) */
#define CELV2_SB_BLOCKS 425 /* candidate blocks per 64x64 superblock (src/gpu/sb_blocks.h:9-10) */
#define CELV2_MAX_REFS 7    /* AV2's INTER_REFS_PER_FRAME (src/gpu/inter_common.h:46) */
#define CELV2_NUM_MODES 13  /* intra modes decided (src/gpu/decision_map.h:19-20) */
#define CELV2_SINGLE_REF 255

/* Status codes returned by every entry point; celv2hip_last_error() describes the last failure on
 * the calling thread. The C++ layer maps them to the reference's exceptions:
 * SIZE -> std::runtime_error (size mismatch, too many references, not enough device memory),
 * ARG -> std::invalid_argument (key count mismatch), src/vk/vk_engine.cpp:176-183. */
enum celv2_status {
  CELV2_OK = 0,
  CELV2_ERR_ARG = 1,
  CELV2_ERR_SIZE = 2,
  CELV2_ERR_DEVICE = 3,
  CELV2_ERR_MEMORY = 4,
  CELV2_ERR_STATE = 5
};

/* ---- The decision map as flat records (the engine's whole output) ---------------------------- */

/* BlockDecision (src/gpu/decision_map.h:22-36). inter, intra: 0/1 and IntraMode. */
typedef struct celv2_block {
  uint8_t w4, h4, inter, intra;
  int16_t mv_row, mv_col;
  uint8_t ref, ref2;
  int16_t mv2_row, mv2_col;
} celv2_block;

/* BlockCandidates (src/gpu/decision_map.h:38-50). */
typedef struct celv2_cands {
  int64_t inter_cost, intra_cost;
  celv2_block inter;
  uint8_t intra;
  uint8_t pad_;
  uint32_t ref_cost[CELV2_MAX_REFS];
  int16_t ref_mv_col[CELV2_MAX_REFS], ref_mv_row[CELV2_MAX_REFS];
} celv2_cands;

/* SbDecision (src/gpu/decision_map.h:52-72) of one 64x64 superblock: preorder tree (Part values
 * None 0, Horz 1, Vert 2, Split 3), leaves in coding order, and every candidate block's choices in
 * src/gpu/sb_blocks.h order (ncands 0 where the superblock is not 64x64). */
typedef struct celv2_sb {
  int32_t ntree, nblocks, ncands, pad_;
  uint8_t tree[CELV2_SB_CAP];
  celv2_block blocks[CELV2_SB_CAP];
  celv2_cands cands[CELV2_SB_BLOCKS];
} celv2_sb;

/* ---- Frames -------------------------------------------------------------------------------- */

/* One frame's padded luma. Exactly one of y8 / y16 is set. on_device != 0: the pointer is a device
 * (HBM) address on the engine's GPU, read in place (frames already resident); else host memory,
 * copied before the call returns (the caller may change it afterwards, src/gpu/decision_engine.h:12-13). */
typedef struct celv2_luma {
  const uint8_t* y8;
  const uint16_t* y16;
  int32_t bit_depth; /* 8 with y8; 9..16 with y16 */
  int32_t on_device;
} celv2_luma;

/* ---- The engine ---------------------------------------------------------------------------- */

typedef struct celv2hip_engine celv2hip_engine;

/* Last failure on this thread ("" if none). */
const char* celv2hip_last_error(void);

/* Number of HIP devices visible (0 without a GPU). */
int celv2hip_device_count(int* n);

/* VulkanEngine::VulkanEngine (src/vk/vk_engine.cpp:138-146): an engine for width x height frames,
 * 64x64 superblocks only, on HIP device `device`. Fails with SIZE for a bad size. */
int celv2hip_create(int width, int height, int sb_size, int device, celv2hip_engine** out);
int celv2hip_destroy(celv2hip_engine* e);

/* VulkanEngine::device_name (src/vk/vk_engine.h:29). */
int celv2hip_device_name(const celv2hip_engine* e, char* buf, int cap);

/* DecisionEngine::set_row_streams (src/gpu/decision_engine.h:30-33). */
int celv2hip_set_row_streams(celv2hip_engine* e, int n);

/* Launch shape of a single engine with one row stream whose records go to the host: the superblock
 * rows fall into `groups` equal runs; level 0 of the motion search runs run by run, and each run's
 * inter scores, finalize and record download overlap the next run's search (0: the default, 4). The
 * decisions do not depend on it; celv2::HipEngine::autotune picks it by timing. No counterpart in the
 * reference (its hand-set launch rules, src/vk/vk_kernels.cpp:193-209, are the closest). */
int celv2hip_set_row_groups(celv2hip_engine* e, int groups);
int celv2hip_row_groups(const celv2hip_engine* e);

/* DecisionEngine::prepare (src/gpu/decision_engine.h:19-22; src/vk/vk_engine.cpp:154-171): queues the
 * next source's import, pyramid and intra scores behind the current work. key < 0: nothing. */
int celv2hip_prepare(celv2hip_engine* e, const celv2_luma* next, long long key);

/* DecisionEngine::launch (src/gpu/decision_engine.h:17-18; src/vk/vk_engine.cpp:173-278): queues the
 * decision of every superblock row. refs nearest first, nrefs 0..7 (0: a key frame); ref_keys NULL
 * or one per reference (equal keys: the same picture, its pyramid and phases are reused). */
int celv2hip_launch(celv2hip_engine* e, const celv2_luma* src, long long src_key, const celv2_luma* refs,
                    const long long* ref_keys, int nrefs, int64_t lambda_q8);

/* One frame of each of n engines (1..64, one device) in shared launches: each stage runs once over
 * every engine's work (at most 8 engines per kernel launch), on the first engine's stream; each
 * engine's decisions are exactly those of celv2hip_launch with the same inputs. Per engine i:
 * srcs[i], src_keys[i] (src_keys may be NULL: no keys), nrefs[i] references refs[i][0..nrefs[i]) with
 * keys ref_keys[i] (ref_keys or ref_keys[i] may be NULL), lambdas_q8[i]. No counterpart in the
 * reference's DecisionEngine: the batched form for several streams sharing one GPU. */
int celv2hip_launch_batch(celv2hip_engine* const* engines, int n, const celv2_luma* srcs, const long long* src_keys,
                          const celv2_luma* const* refs, const long long* const* ref_keys, const int* nrefs,
                          const int64_t* lambdas_q8);

/* The superblock rows in the order the last launch queued them (row streams); returns the count. */
int celv2hip_row_order(const celv2hip_engine* e, int* rows, int cap);

/* Part of DecisionEngine::collect (src/vk/vk_engine.cpp:280-294): waits for superblock row `row` of
 * the last launch and writes its sb_cols records to out. */
int celv2hip_wait_row(celv2hip_engine* e, int row, celv2_sb* out);

/* The same wait without the copy: *out points at the row's sb_cols records in the engine's pinned
 * host buffer, valid until the engine's next launch. No counterpart in the reference (the C++ engine's
 * collect converts records to a DecisionMap straight from there). */
int celv2hip_row_records(celv2hip_engine* e, int row, const celv2_sb** out);

/* Waits for all queued work (used before timing and in tests). */
int celv2hip_sync(celv2hip_engine* e);

/* Statistics over the launches since the last reset (at most 512), measured on the engine's stream
 * with HIP events, without host waits between launches. ms[0..6] sums, in milliseconds:
 * [0] import + pyramids + phase planes queued by launch (a source not prepare()d includes its intra
 * scores), [1] motion search (5 levels), [2] inter scores, [3] compound + finalize, [4] record copy,
 * [5] launch start to end, [6] intra scores queued by prepare(), [7] the motion search's level-0 launch
 * (in the wavefront form with the inter scores interleaved). counts: [0] launches, [1] distinct
 * (MV, reference, superblock) tile evaluations, [2] motion-search centres searched, [3] compound
 * blocks scored, [4] non-repeated (block, MV, reference) inter candidates scored (while counting is on,
 * celv2hip_set_counting), [5] superblocks the pruner sent to its fallback, [6] the centres of [2]
 * searched at level 0. Returns the number of launches (negative status on failure). */
int celv2hip_stats_reset(celv2hip_engine* e);
/* on != 0: later launches and prepares record the stage events the ms figures come from (off by
 * default: each record between two kernels costs the stream a few microseconds). ms are all zero
 * for a window with any launch made with timing off; counts are kept either way. */
int celv2hip_set_timing(celv2hip_engine* e, int on);
int celv2hip_stats(celv2hip_engine* e, double* ms, int cap_ms, long long* counts, int cap_counts);

/* ---- Timing mode: device-resident throughput ------------------------------------------------ */

/* Keeps results on the device: wait_row is not available and no device-to-host copy is queued.
 * on != 0: device-resident mode; 0: normal. */
int celv2hip_set_resident(celv2hip_engine* e, int on);
/* on != 0: a launch whose engines are all resident (and not timing stages) is captured as a HIP graph
 * on the launching engine's stream and replayed; the next launch of the same shape updates that graph
 * in place instead of instantiating it again. The same kernels and decisions; the host queues one
 * graph instead of each kernel. No counterpart in the reference (its command batches,
 * src/vk/vk_engine.cpp:261-268, are recorded anew for every frame). graph_stats: launches replayed as
 * graphs, and of them the ones that instantiated a graph. */
int celv2hip_set_graph(celv2hip_engine* e, int on);
int celv2hip_graph_stats(const celv2hip_engine* e, long long* launches, long long* builds);

/* Per-superblock reference limits for later launches (the pruned form): superblock s (row-major) is
 * decided exactly as by a launch with only its nearest k[s] references (1..7; references beyond
 * nrefs are absent anyway); the work for the pruned (superblock, reference) pairs is skipped. count =
 * superblocks of the frame, or k NULL / count 0 to decide every superblock with all references. No
 * counterpart in the reference's DecisionEngine: the hook of the learned pruner. */
int celv2hip_set_ref_limits(celv2hip_engine* e, const uint8_t* k, int count);
/* Per-superblock motion-search modes for later launches (the pruned form's search dimension): bit L
 * of modes[s] (L = 0, 1) narrows the full searches of motion-search level L from +-8 to +-2 samples
 * around each of the two centres (me_search, src/gpu/inter_common.h:106-119, with a smaller range;
 * the centres, their ranking and the refinement are unchanged). A block at level L is searched narrow
 * only when every superblock that keeps the reference and reads the block (its own, or a neighbour
 * whose ring it lies in: 1 block at level 0, 2 at level 1) has bit L set; so each superblock is decided
 * exactly as the oracle decides it under the same modes and limits. modes NULL / count 0: every
 * search full. No counterpart in the reference: the motion-search dimension of the pruned form. */
int celv2hip_set_search_modes(celv2hip_engine* e, const uint8_t* modes, int count);
/* The search modes the last launch's levels 1 and 0 obeyed (set, or chosen by the pruner; 0 where
 * every search was full) into out[0 .. superblocks); returns the count or a negative status. */
long long celv2hip_fetch_search_modes(celv2hip_engine* e, uint8_t* out, long long cap);
/* Per-superblock intra mode groups for the sources scored from now on (the next prepare() or
 * launch): bit g of masks[s] set = group g scored (0: DC, V, H, always scored; 1: smooth, smooth V,
 * smooth H, Paeth; 2: the directional modes at 45, 67, 113 degrees; 3: at 135, 157, 203 degrees).
 * Every block of superblock s takes the least intra cost over its scored modes, as the reference's
 * choice (src/gpu/kernels_ref.cpp:46-50) restricted to them; the other modes' scores are not
 * computed. A source keeps the masks it was scored with. masks NULL / count 0: every group. No
 * counterpart in the reference: the intra dimension of the pruned form. */
int celv2hip_set_intra_masks(celv2hip_engine* e, const uint8_t* masks, int count);
/* The intra mode groups the last launch's source was scored with (one byte per superblock; the
 * network's when it chose them); returns the count or a negative status. */
long long celv2hip_fetch_intra_masks(celv2hip_engine* e, uint8_t* out, long long cap);

/* The last launch's per-superblock decision cost (the 64x64 square's partition DP cost, SATD*256 +
 * lambda*bits units: J of the superblock) into out[0 .. superblocks); returns the count or a negative
 * status. No counterpart in the reference's DecisionEngine (cost regret of the pruned form). */
long long celv2hip_fetch_sb_costs(celv2hip_engine* e, int64_t* out, long long cap);

/* ---- The learned pruner (the pruned form; no counterpart in the reference's DecisionEngine) ------
 * A small network over a frame's superblock grid: inputs are CELV2HIP_PRUNER_FEATURES (14) features
 * per superblock from the source's level 2, each reference's level-2 motion search, the frame's
 * lambda and frame-wide means; output j (0..5) is the logit of "the
 * superblock needs more than its nearest j + 1 references". BF16 operands on the matrix cores, FP32
 * accumulation and master weights, AdamW. Attached to an engine, every launch computes the features
 * after motion-search level 2, runs the network and sets each superblock's reference limit (as
 * celv2hip_set_ref_limits) before levels 1 and 0, inter scores and finalize. A pruner may be
 * attached to several engines; it must not be trained while their launches run. */
typedef struct celv2hip_pruner celv2hip_pruner;
#define CELV2HIP_PRUNER_FEATURES 14
/* Outputs per superblock: 6 reference outputs (j: needs more than j + 1 references), then 3 intra
 * mode-group outputs (6 + g - 1: needs group g, g = 1..3), then the search output (9: needs the full
 * motion search at levels 1 and 0). */
#define CELV2HIP_PRUNER_OUTPUTS 10

/* Number of FP32 parameters of the network. */
int celv2hip_pruner_param_count(void);
/* A pruner on `device` with a workspace for max_positions superblocks, He-initialised from seed. */
int celv2hip_pruner_create(int device, int max_positions, unsigned seed, celv2hip_pruner** out);
int celv2hip_pruner_destroy(celv2hip_pruner* p);
int celv2hip_pruner_get_params(celv2hip_pruner* p, float* out, int cap);
/* Loads parameters (n = param count) and resets the AdamW state. */
int celv2hip_pruner_set_params(celv2hip_pruner* p, const float* in, int n);
/* b frames of h x w superblocks, feats[((f * h + y) * w + x) * 14 + i]: the 10 logits per superblock. */
int celv2hip_pruner_forward(celv2hip_pruner* p, const float* feats, int b, int h, int w, float* logits);
/* The same forward pass without any device allocation: features and logits are staged in a static
 * device array of the library (one per device, taken at celv2hip_pruner_create), so a caller that
 * cannot allocate (a serving loop, code around graph capture) may call it. At most 16384 positions
 * (b * h * w) per call, else CELV2_ERR_SIZE and nothing runs. Calls that overlap on one device share
 * that array and race; the array keeps the last call's contents. No counterpart in the reference: an
 * entry point for callers that cannot allocate device memory. */
int celv2hip_pruner_forward_static(const celv2hip_pruner* p, const float* feats, int b, int h, int w, float* logits);
/* Mean binary cross-entropy against labels kstar (per superblock: the references it needs, 1..7, in
 * bits 0..2, bit 2 + g set when it needs intra mode group g = 1..3, and bit 6 set when it needs the
 * full motion search) and its gradient with respect
 * to every parameter, without updating them. */
int celv2hip_pruner_gradients(celv2hip_pruner* p, const float* feats, const uint8_t* kstar, int b, int h, int w,
                              float* grads, float* loss);
/* One AdamW step (beta 0.9 / 0.999, eps 1e-8) on that loss; *loss is the loss before the step. */
int celv2hip_pruner_train_step(celv2hip_pruner* p, const float* feats, const uint8_t* kstar, int b, int h, int w,
                               float lr, float weight_decay, float* loss);
/* Attaches p (NULL detaches). A superblock keeps 1 + the number of leading outputs whose probability
 * is >= threshold references (at most nrefs); if any output lies within `band` of the threshold the
 * superblock falls back to all references (counted in celv2hip_stats counts[5]). */
int celv2hip_set_pruner(celv2hip_engine* e, const celv2hip_pruner* p, float threshold, float band);
/* gthreshold > 0: the attached pruner also chooses each superblock's intra mode groups (as
 * celv2hip_set_intra_masks): group g when the probability of output 5 + g is >= gthreshold, every
 * group where the superblock fell back to all references; the source's intra scores then run after
 * the network (prepare() leaves them to the launch). 0: every group (the default). No counterpart in
 * the reference. */
int celv2hip_set_pruner_groups(celv2hip_engine* e, float gthreshold);
/* sthreshold > 0: the attached pruner also chooses each superblock's motion-search modes (as
 * celv2hip_set_search_modes): the full search when the probability of output 9 is >= sthreshold or
 * the superblock fell back to all references, else narrow at levels 1 and 0. 0: every search full
 * (the default). No counterpart in the reference: the network's fallback to the full search. */
int celv2hip_set_pruner_search(celv2hip_engine* e, float sthreshold);
/* on != 0: launches compute the pruner's features even without a pruner (training data). */
int celv2hip_set_features(celv2hip_engine* e, int on);
/* The last launch's features (superblocks x 14 floats); returns the count or a negative status. */
long long celv2hip_fetch_features(celv2hip_engine* e, float* out, long long cap);
/* The reference limits the last launch used (one byte per superblock; its reference count for every
 * superblock when it had no limits). */
long long celv2hip_fetch_ref_limits(celv2hip_engine* e, uint8_t* out, long long cap);

/* on != 0: later launches count the non-repeated inter candidates they score (stats counts[4]);
 * off by default, as the count costs the inter kernel extra LDS atomics. */
int celv2hip_set_counting(celv2hip_engine* e, int on);

/* Copies the records of superblock row `row` of the last launch in resident mode (tests). */
int celv2hip_fetch_row(celv2hip_engine* e, int row, celv2_sb* out);


/* ---- Stage outputs of the last launch, for per-kernel tests (tests/test_vk_kernels.cpp's
 * downsample / subpel / me checks of CelV2 compare each kernel with its CPU rule the same way) ---- */

/* Interior samples (w x h of that level) of the source pyramid (ref < 0) or of reference `ref`'s
 * (its index in the last launch); phase 1..15: that quarter-pel plane of level 0. */
int celv2hip_fetch_plane(celv2hip_engine* e, int ref, int level, int phase, uint8_t* out, int* w, int* h);
/* The same plane with its border: (w + 128) x (h + 128) samples, the 64-sample border that holds each
 * clamped position's value. */
int celv2hip_fetch_plane_bordered(celv2hip_engine* e, int ref, int level, int phase, uint8_t* out, int* w, int* h);

/* Motion field of reference `ref` at `level` (cols x rows, level 0 in quarter pels). */
int celv2hip_fetch_field(celv2hip_engine* e, int ref, int level, int32_t* mvx, int32_t* mvy, int* cols, int* rows);

/* Intra scores of the last launch's source: per superblock (row-major), per candidate block 5..424,
 * the 13 modes' SATD (block_mode_satd, src/gpu/intra_common.h:149-163); entries of blocks 0..4 (the
 * 64x64 level, inter only) are undefined. `cap` counts uint32 entries; returns the entries written
 * (superblocks x 425 x 13) or a negative status. */
long long celv2hip_fetch_intra(celv2hip_engine* e, uint32_t* out, long long cap);

/* ---- Forward transforms of the RD pass's trial scoring ------------------------------------------
 * The transform step of tx_cost (src/core/rd_pass.cpp:231-300): per residual block, fwd_dct's
 * coefficients (src/core/transform.h:12, transform.cpp:170-214: a separable DCT-II in float, rows then
 * the kept columns, orthonormal times `scale`, rounded half away from zero; AVM's layout, row v =
 * vertical frequency, only min(w, 32) x min(h, 32) kept) and, for a 64-point side, dropped_energy
 * (rd_pass.cpp:268-276); 0 below 64 points. Kinds per direction as fwd_transform
 * (transform.cpp:250-262) takes them from a TX_TYPE (g_ver_tx_type, g_hor_tx_type): DCT-II at every
 * size, DST-VII and its flip up to 16 points and the identity up to 32, these three as orthonormal
 * float forms (AVM's integer kernels and their tables are not part of this library). */
enum {
  CELV2_TX_DCT = 0,     /* DCT-II (fwd_dct's values) */
  CELV2_TX_DST = 1,     /* DST-VII */
  CELV2_TX_FLIPDST = 2, /* DST-VII of the reversed samples */
  CELV2_TX_IDT = 3      /* identity */
};
typedef struct celv2_tx_job {
  const int16_t* res; /* device memory: the block's first residual sample */
  int32_t* coeff;     /* device memory: min(h, 32) rows of min(w, 32) coefficients */
  int64_t* dropped;   /* device memory or NULL */
  double scale;       /* > 0 */
  int32_t stride;     /* samples from one residual row to the next, >= w */
  uint8_t w_log2, h_log2; /* 2..6 each (4..64 samples) */
  uint8_t vkind, hkind;   /* vertical and horizontal kind (CELV2_TX_*) */
} celv2_tx_job;
/* Transforms every job on `device` and returns when they are done; kernel_ms (may be NULL) gets the
 * kernel's own duration (HIP events around it, the job list's upload excluded). Jobs may come in any
 * order and mix sizes and kinds; one job's result does not depend on the others. CELV2_ERR_ARG
 * (nothing runs) for a job with a size, kind, stride, scale or pointer outside the above. */
int celv2hip_fwd_tx(int device, const celv2_tx_job* jobs, int njobs, float* kernel_ms);

/* ---- Device memory helpers for callers without another allocator --------------------------- */
int celv2hip_device_alloc(int device, long long bytes, void** out);
int celv2hip_device_free(int device, void* p);
int celv2hip_copy_to_device(int device, void* dst, const void* src, long long bytes);
int celv2hip_copy_from_device(int device, void* dst, const void* src, long long bytes);
/* Device-to-device copy of `bytes` (a plain copy kernel of known size, for calibrating traffic counters). */
int celv2hip_copy_on_device(int device, void* dst, const void* src, long long bytes);


/* ---- The C++ engine (celv2::HipEngine, include/celv2/hip_engine.h) behind a handle ------------
 * For callers that cannot link C++ (tests, the benchmark): DecisionEngine::launch / prepare /
 * collect with a RowGate, the DecisionMap exported as flat records. Frames are padded luma planes
 * (y8 and/or y16 at bit_depth); lam >= 0 overrides lambda_q8(qindex). Returns 0, 1 for
 * std::invalid_argument, 2 for other failures (celv2_engine_last_error describes them). */
const char* celv2_engine_last_error(void);
void* celv2_engine_create(int width, int height, int device, int row_streams, int full_depth);
void celv2_engine_destroy(void* engine);
int celv2_engine_launch(void* engine, const uint8_t* src8, const uint16_t* src16, int bit_depth,
                        const uint8_t* const* refs8, const uint16_t* const* refs16, const long long* keys, int nrefs,
                        int qindex, long long lam, long long src_key);
int celv2_engine_prepare(void* engine, const uint8_t* y8, const uint16_t* y16, int bit_depth, long long key);
/* The same launch / prepare for frames already in device memory (celv2_luma with on_device set):
 * HipEngine::launch_device / prepare_device. */
int celv2_engine_launch_device(void* engine, const celv2_luma* src, const celv2_luma* refs, const long long* keys,
                               int nrefs, int qindex, long long lam, long long src_key);
int celv2_engine_prepare_device(void* engine, const celv2_luma* next, long long key);
int celv2_engine_collect(void* engine, celv2_sb* out, int* rows_ready);
int celv2_engine_description(void* engine, char* buf, int cap);
/* Waits for all work queued on the engine's streams (timing). */
int celv2_engine_sync(void* engine);
/* HipEngine::set_row_groups / row_groups / autotune_device: the one-row-stream launch's row groups, and
 * the tuner that picks them by timing steady-state frames. The tuner writes the chosen count and
 * (row groups, median ms per frame) for each candidate it timed, up to cap (none when the process
 * had tuned this size, reference count and device already). No counterpart in the reference. */
int celv2_engine_set_row_groups(void* engine, int groups);
int celv2_engine_row_groups(void* engine);
int celv2_engine_autotune_device(void* engine, const celv2_luma* src, const celv2_luma* refs, const long long* keys,
                                 int nrefs, int qindex, long long lam, long long src_key, int frames, int force,
                                 int* groups, int* cand_groups, double* cand_ms, int cap, int* ncand);
int celv2_sizeof_sb(void);

/* Digest of the native sources the library was built from (the build's own check against a stale
 * library; no counterpart in the reference). */
const char* celv2hip_source_hash(void);

#ifdef __cplusplus
}
#endif

#endif /* CELV2HIP_H_ */
