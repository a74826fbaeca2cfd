// Launch batches: one kernel launch serves up to kMaxItems independent work items (frames of
// different engines, possibly of different sizes). The items' arguments travel by value in the
// kernel's argument segment; `first[i]` is item i's first workgroup, so a workgroup finds its item
// with a short scalar scan and computes exactly what a launch of that item alone would.
#pragma once
#include <hip/hip_runtime.h>

namespace celv2hip {

constexpr int kMaxItems = 8;

template <class T>
struct Batch {
  T item[kMaxItems];
  int first[kMaxItems + 1];  // first[n] = total workgroups
  int n = 0;

  // Host: appends an item that needs `groups` workgroups.
  void add(const T& t, int groups) {
    if (n == 0) first[0] = 0;
    item[n] = t;
    first[n + 1] = first[n] + groups;
    n++;
  }
  __host__ __device__ int total() const { return n ? first[n] : 0; }
};

// The item that owns workgroup g (g < total).
template <class T>
__device__ __forceinline__ int batch_item(const Batch<T>& b, int g) {
  int i = 0;
#pragma unroll
  for (int k = 1; k < kMaxItems; k++)
    if (k < b.n && g >= b.first[k]) i = k;
  return i;
}

}  // namespace celv2hip
