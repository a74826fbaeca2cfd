// ORACLE — test infrastructure, not product code. A CPU restatement of CelV2's decision-engine
// rules, used only by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg to check and
// time the HIP engine. The product (celv2_amd/) never includes, links or calls anything here.
//
// Every rule cites the reference function it restates (paths relative to the CelV2 tree):
//   src/gpu/intra_common.h, src/gpu/inter_common.h, src/gpu/sb_blocks.h, src/gpu/kernels_ref.cpp,
//   src/gpu/cpu_engine.cpp, src/gpu/decision_map.cpp, src/io/y4m.cpp.
// Pinning (see tests/test_oracle_*.py): the header-only primitives are compared with the
// reference's own headers compiled from /root/reference (oracle/_ref, recipe oracle/Makefile) and
// with fixtures under tests/golden/. The per-block choice (kernels_ref.cpp:41-104), which lives in a
// file that cannot be built here (it includes AVM's quant_common.h), is pinned by rebuilding every
// candidate record of several superblocks from those pinned primitives, by known answers (identical
// references, a source averaging two references, flat pictures, tied runners-up) and by deliberate
// mistakes (ORC_PLANT in oracle.cpp) that each make one of those checks fail; the partition DP by
// brute force over every partition tree of small squares, forced ties and invariants. lambda_q8 is
// an input here: its parity with CelV2 is unpinned (AVM's quantiser table is absent).
// Speed paths (direct addressing where a block lies inside its plane, two-row SSE2 SADs, separable
// quarter-pel planes) compute the same integers as the clamped per-sample forms they sit beside.
#pragma once
#include <atomic>
#include <condition_variable>
#include <cstdint>
#include <exception>
#include <functional>
#include <mutex>
#include <thread>
#include <vector>

namespace orc {

constexpr int kLevels = 5;         // inter_common.h:44
constexpr int kPhases = 16;        // inter_common.h:11
constexpr int kModes = 13;         // intra_common.h:42
constexpr int kMaxRefs = 7;        // inter_common.h:46
constexpr int kSquares = 85;       // sb_blocks.h:9
constexpr int kBlocks = 425;       // sb_blocks.h:10
constexpr int kRange = 8;          // inter_common.h:35, 40 (both levels)
constexpr int kPelPenalty = 4;     // inter_common.h:55
constexpr int kNarrowRange = 2;    // the pruned form's narrow full search (no counterpart in the reference)
constexpr int kQpelPenalty = 16;   // inter_common.h:140
constexpr int kIntraMaxSize = 32;  // intra_common.h:39
constexpr int kInterMaxSize = 64;  // intra_common.h:40
constexpr int kSplitBits = 2;      // intra_common.h:33
constexpr int kIntraOnInter = 32;  // inter_common.h:50
constexpr int kNewmvBits = 3;      // inter_common.h:45
constexpr int kCap = 256;          // sb_out.h:11

// A read-only 8-bit plane read through clamped coordinates (intra_common.h:11-26). A reference's
// level 0 also carries its 16 quarter-pel phase planes `phase` bytes apart (0: none).
struct Plane8 {
  const uint8_t* px = nullptr;
  int stride = 0, w = 0, h = 0;
  size_t phase = 0;
  int at(int y, int x) const {
    if (y < 0) y = 0;
    if (y > h - 1) y = h - 1;
    if (x < 0) x = 0;
    if (x > w - 1) x = w - 1;
    return px[static_cast<size_t>(y) * stride + x];
  }
  // Sample at (y + qy/4, x + qx/4), qy / qx in quarter pels, from phase plane (qy&3)*4 + (qx&3).
  int atq(int y, int x, int qy, int qx) const {
    Plane8 p = *this;
    p.px = px + static_cast<size_t>((qy & 3) * 4 + (qx & 3)) * phase;
    return p.at(y + (qy >> 2), x + (qx >> 2));
  }
};

// ---- primitives (header-only rules of the reference) -------------------------------------------
int mode_bits(int mode);                                         // intra_common.h:32
int wht4_abs_sum(const int d[16]);                               // intra_common.h:131-146
int qpel_tap(int q, int k);                                      // inter_common.h:12-19
int qpel_sample(const Plane8& s, int y, int x, int qy, int qx);  // inter_common.h:20-33
int half_sample(const Plane8& s, int y, int x);                  // inter_common.h:91-94
int sad_8x8(const Plane8& cur, int x, int y, const Plane8& ref, int mvx, int mvy);   // inter_common.h:96-104
int full_search(const Plane8& cur, const Plane8& ref, int bx, int by, int px, int py, int range, int* mvx,
                int* mvy);                                                           // inter_common.h:106-119
int satd_8x8_q(const Plane8& cur, int x, int y, const Plane8& ref, int qmx, int qmy); // inter_common.h:121-133
void refine_qpel(const Plane8& cur, const Plane8& ref, int bx, int by, int* mvx, int* mvy);  // inter_common.h:135-154
// Returns the integer search's winning cost (SAD + distance penalty from its centre).
int search_cell(const Plane8& cur, const Plane8& ref, int bx, int by, const int* pmx, const int* pmy, int pcols,
                int prows, int* mvx, int* mvy, bool level0, int range = kRange);                          // inter_common.h:156-193
int edge_above(const Plane8& s, int x, int y, int k);                                 // intra_common.h:47-51
int edge_left(const Plane8& s, int x, int y, int k);                                  // intra_common.h:52-56
int directional(const Plane8& s, int x, int y, int w, int h, int angle, int i, int j); // intra_common.h:61-85
int intra_predict(const Plane8& s, int x, int y, int w, int h, int mode, int i, int j); // intra_common.h:89-118
int dc_mean(const Plane8& s, int x, int y, int w, int h);                             // intra_common.h:120-128
int64_t intra_satd(const Plane8& s, int x, int y, int w, int h, int mode);             // intra_common.h:149-163
int64_t inter_satd(const Plane8& cur, const Plane8& ref, int x, int y, int w, int h, int mvx, int mvy);  // inter_common.h:222-235
int64_t compound_satd(const Plane8& cur, const Plane8& r0, int mx0, int my0, const Plane8& r1, int mx1, int my1,
                      int x, int y, int w, int h);                                    // inter_common.h:205-220
bool pair_allowed(int a, int b, int nrefs);                                           // inter_common.h:199-202
int pair_bits(int a, int b);                                                          // inter_common.h:204
int cells_covered(int x, int y, int w, int h);                                        // inter_common.h:60-62
int cand_count(int x, int y, int w, int h);                                           // inter_common.h:63
int cand_cell(int x, int y, int w, int h, int k, int cols, int rows);                 // inter_common.h:65-78
bool cand_repeats(const int* fx, const int* fy, int cols, int rows, int x, int y, int w, int h, int k);  // inter_common.h:82-89
int level_first(int level);                                                           // sb_blocks.h:12
int square_id(int level, int qx, int qy);                                             // sb_blocks.h:14
void block_geom(int b, int* x, int* y, int* w, int* h);                               // sb_blocks.h:17-27
int block_index(int x, int y, int w, int h);                                          // sb_blocks.h:30-37

// ---- frames, pyramids, motion fields -----------------------------------------------------------
// The padded 8-bit luma of a frame (Frame420::y, src/io/y4m.h:11-23): pw x ph, stride pw.
struct Luma {
  int width = 0, height = 0, pw = 0, ph = 0;
  std::vector<uint8_t> y;
};
Luma make_luma(int width, int height, const uint8_t* y);  // copies a pw x ph plane
// round_to_8bit for one plane (src/io/y4m.cpp:28-34).
void round_plane(const uint16_t* in, uint8_t* out, size_t n, int bit_depth);

struct Pyramid {  // build_pyramid_ref (kernels_ref.cpp:213-238)
  std::vector<std::vector<uint8_t>> px;
  Plane8 plane[kLevels];
  Pyramid() = default;
  Pyramid(const Pyramid&) = delete;
  Pyramid& operator=(const Pyramid&) = delete;
  Pyramid(Pyramid&&) = default;
  Pyramid& operator=(Pyramid&&) = default;
};
Pyramid build_pyramid(const Luma& f, bool phases);

struct Field {  // RefMvLevel (kernels_ref.h:33-36)
  std::vector<int> mvx, mvy;
  int cols = 0, rows = 0;
};
void search_rows(const Pyramid& cur, const Pyramid& ref, int level, const Field* coarser, int row0, int row1,
                 Field& out, int ref_index = 0);  // me_level_rows_ref (kernels_ref.cpp:240-251)
void search_all(const Pyramid& cur, const Pyramid& ref, std::vector<Field>& levels, int ref_index = 0);  // levels 4..0

// ---- the decision map --------------------------------------------------------------------------
enum Part : uint8_t { kNone = 0, kHorz = 1, kVert = 2, kSplit = 3 };
struct Block {  // BlockDecision (decision_map.h:22-36)
  uint8_t w4 = 0, h4 = 0;
  bool inter = false;
  uint8_t intra = 0;
  int16_t mv_row = 0, mv_col = 0;
  uint8_t ref = 0, ref2 = 255;
  int16_t mv2_row = 0, mv2_col = 0;
};
struct Cands {  // BlockCandidates (decision_map.h:38-50)
  int64_t inter_cost = INT64_MAX, intra_cost = INT64_MAX;
  Block inter;
  uint8_t intra = 0;
  uint32_t ref_cost[kMaxRefs] = {UINT32_MAX, UINT32_MAX, UINT32_MAX, UINT32_MAX, UINT32_MAX, UINT32_MAX, UINT32_MAX};
  int16_t ref_mv_col[kMaxRefs] = {}, ref_mv_row[kMaxRefs] = {};
};
struct Sb {  // SbDecision (decision_map.h:52-72), the fields the engines fill
  std::vector<uint8_t> tree;
  std::vector<Block> blocks;
  std::vector<Cands> cands;
  int64_t cost = 0;  // the superblock's decision cost (the root square's DP cost)
};
struct Map {
  int sb_size = 0, sb_cols = 0, sb_rows = 0;
  std::vector<Sb> sbs;
};

// References nearest first: level-0 planes (with phases) and level-0 fields.
struct RefSet {
  std::vector<const Plane8*> plane;
  std::vector<const Field*> field;
  int size() const { return static_cast<int>(plane.size()); }
};

Sb decide_sb(const Luma& src, const RefSet& rs, int sb_col, int sb_row, int sb_size, int64_t lam);  // kernels_ref.cpp:253-270
Map decide_frame(const Luma& src, const std::vector<const Luma*>& refs, int sb_size, int64_t lam);
// The pruned form's intra restriction for decisions made on this thread (test of the engine's intra
// mode groups; null: every mode): one 4-bit group set per superblock of `cols` columns.
// The pruned form's reference limits and motion-search modes (oracle.cpp), with the superblock grid
// they index; null: none.
extern thread_local const uint8_t* g_ref_limits;
extern thread_local const uint8_t* g_search_modes;
extern thread_local int g_sb_cols, g_sb_rows;
// Whether block (bx, by) of motion-search level `level` is searched narrow for reference r under them.
bool narrow_search(int level, int bx, int by, int r);
extern thread_local const uint8_t* g_intra_masks;
extern thread_local int g_intra_mask_cols;  // kernels_ref.cpp:200-209, 282-301

// One square's bottom-up decision (kernels_ref.cpp:124-198), exposed for the brute-force tests: the
// cost it returns and the subtree it appends.
int64_t decide_square(const Plane8& s, const RefSet& rs, int x, int y, int size, int64_t lam, Sb& out,
                      std::vector<Cands>* rec, int ox, int oy);

// ---- row gate and the threaded engine ----------------------------------------------------------
class RowGate {  // decision_map.h:87-108, decision_map.cpp:19-61
 public:
  void publish(int row);
  void publish_one(int row);
  void wait_row(int row) const;
  void fail(std::exception_ptr e);
  int ready_rows() const;

 private:
  mutable std::mutex m_;
  mutable std::condition_variable cv_;
  int ready_ = 0;
  std::vector<char> one_;
  std::exception_ptr error_;
};

int decision_threads(int cores, int tile_threads);  // cpu_engine.cpp:7-9

class CpuEngine {  // cpu_engine.h:21-59, cpu_engine.cpp
 public:
  CpuEngine(int width, int height, int sb_size, int threads, std::function<void(int)> row_hook = {});
  ~CpuEngine();
  CpuEngine(const CpuEngine&) = delete;
  CpuEngine& operator=(const CpuEngine&) = delete;
  void launch(const Luma& src, const std::vector<const Luma*>& refs, int64_t lam, Map& dm);
  void collect(Map& dm, RowGate& gate);

 private:
  void work();
  void stop_and_join();
  int width_, height_, sb_size_, threads_, sb_cols_, sb_rows_;
  std::function<void(int)> row_hook_;
  int64_t lam_ = 0;
  int nrefs_ = 0;
  Luma src_;
  Pyramid cur_;
  std::vector<Pyramid> ref_;
  std::vector<std::vector<Field>> lv_;
  std::vector<Sb> out_;
  std::vector<std::thread> workers_;
  std::atomic<int> next_me_[kLevels];
  std::atomic<int> next_row_{0};
  std::atomic<bool> abort_{false};
  std::mutex m_;
  std::condition_variable cv_;
  int arrived_[kLevels] = {};
  std::vector<char> row_done_;
  std::exception_ptr error_;
};

}  // namespace orc
