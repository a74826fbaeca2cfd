// The learned pruner object behind celv2hip_pruner_* (include/celv2hip.h): the network's FP32
// parameters, their BF16 copy, AdamW state and a training workspace on one device.
#pragma once
#include <hip/hip_runtime.h>

#include "../kernels/cnn.h"

struct celv2hip_pruner {
  int device = -1;
  hipStream_t stream = nullptr;
  float* params = nullptr;
  uint16_t* pb = nullptr;
  float* m1 = nullptr;
  float* m2 = nullptr;
  float* grads = nullptr;
  celv2hip::CnnWork work;
  long long step = 0;

  ~celv2hip_pruner() {
    if (device >= 0) (void)hipSetDevice(device);
    if (stream) (void)hipStreamSynchronize(stream);
    for (void* p : {static_cast<void*>(params), static_cast<void*>(pb), static_cast<void*>(m1), static_cast<void*>(m2),
                    static_cast<void*>(grads)})
      if (p) (void)hipFree(p);
    celv2hip::cnn_free(work);
    if (stream) (void)hipStreamDestroy(stream);
  }
};
