// The learned pruner's network (cnn.hip): sizes, the flat FP32 parameter layout and the device
// workspace of a forward / backward pass.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include <stdexcept>

namespace celv2hip {

constexpr int kFeatures = 14;       // inputs per superblock (pruner.hip)
constexpr int kHidden = 32;         // channels of the two 3x3 layers
constexpr int kRefOut = 6;          // output j < 6: needs more than j + 1 references
constexpr int kGroupOut = 3;        // output 6 + g: needs intra mode group g + 1 (groups 1..3; group 0 always scored)
constexpr int kSearchOut = 1;       // output 9: needs the full motion search at levels 1 and 0 (else narrow)
constexpr int kOut = kRefOut + kGroupOut + kSearchOut;
constexpr int kK1 = 128;            // im2col row of layer 1: 9 * kFeatures = 126, padded
constexpr int kK2 = 9 * kHidden;    // im2col row of layer 2
// Parameters: W1[9 kFeatures][kHidden], b1, W2[kK2][kHidden], b2, W3[kHidden][kOut], b3 (row k of a
// weight matrix is im2col column k: k = tap * channels + channel, tap = (dy + 1) * 3 + dx + 1).
constexpr int kW1 = 0, kB1 = kW1 + 9 * kFeatures * kHidden, kW2 = kB1 + kHidden, kB2 = kW2 + kK2 * kHidden,
              kW3 = kB2 + kHidden, kB3 = kW3 + kHidden * kOut, kParams = kB3 + kOut;

struct CnnWork {
  float* x = nullptr;  // [positions][kFeatures]
  uint16_t* col1 = nullptr;
  float* z1 = nullptr;
  uint16_t* col2 = nullptr;
  float* z2 = nullptr;
  uint16_t* z2b = nullptr;
  float* z3 = nullptr;  // logits [positions][kOut]
  float* dz3 = nullptr;
  uint16_t* dz3b = nullptr;
  float* dz2 = nullptr;
  uint16_t* dz2b = nullptr;
  float* dcol2 = nullptr;
  float* dz1 = nullptr;
  uint16_t* dz1b = nullptr;
  float* loss_part = nullptr;
  // labels: the references a superblock needs (1..7, bits 0..2), the intra groups it needs (bit 3 + g - 1
  // for group g) and whether it needs the full motion search (bit 6)
  uint8_t* kstar = nullptr;
  size_t cap = 0;
  size_t positions_capacity() const;
};

void cnn_alloc(CnnWork& w, size_t positions);
void cnn_free(CnnWork& w);
void cnn_params_to_bf16(const float* params, uint16_t* pb, hipStream_t s);
void cnn_forward(const CnnWork& w, const float* params, const uint16_t* pb, int b, int h, int wd, hipStream_t s);
void cnn_backward(const CnnWork& w, const uint16_t* pb, float* grads, int b, int h, int wd, hipStream_t s);
void cnn_adamw(float* params, const float* grads, float* m1, float* m2, uint16_t* pb, long long step, float lr, float wd,
               hipStream_t s);
void cnn_limits(const float* logits, int m, int nrefs, float threshold, float band, uint8_t* kmap,
                unsigned long long* fallbacks, hipStream_t s);
// Inference in one launch over x (b frames of h x wd superblocks): logits (if non-null, as cnn_forward
// leaves them in w.z3, bit for bit), reference limits into kmap (if non-null, as cnn_limits) and, if
// imask is non-null, each superblock's intra mode groups: group g (1..3) when sigma(logit 5 + g) >=
// gthreshold, every group when a reference output was within band of its threshold (a fallback);
// if smap is non-null, each superblock's motion-search modes: the full search (0) when sigma(logit 9)
// >= sthreshold or on a fallback, else narrow at levels 1 and 0 (3). context: x lacks the frame
// context f[11..13] (launch_features without it); the launch computes it and writes it into x.
void cnn_infer(const float* x, int b, int h, int wd, const float* params, const uint16_t* pb, float* logits, int nrefs,
               float threshold, float band, uint8_t* kmap, unsigned long long* fallbacks, hipStream_t s,
               uint8_t* imask = nullptr, float gthreshold = 0.f, uint8_t* smap = nullptr, float sthreshold = 0.f,
               bool context = false);

// The static scratch of the no-allocation forward pass: features [kStaticPositions][kFeatures]
// followed by logits [kStaticPositions][kOut] (1.3 MB), one copy per device. cnn_scratch_init takes
// the current device's address once (call with `device` current); cnn_scratch returns it (null before).
constexpr int kStaticPositions = 16384;
constexpr int kScratchDevices = 64;
hipError_t cnn_scratch_init(int device);
float* cnn_scratch(int device);

}  // namespace celv2hip
