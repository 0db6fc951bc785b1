// Grid sizes of the grid-stride node kernels: resident-block counts per
// kernel and the launch policies the BP autotuner chooses between.
#pragma once

#include <hip/hip_runtime.h>

namespace bqa {

// Blocks of kernel K (THREADS per block) resident on the device at once: CUs x
// resident blocks per CU, computed once per kernel.  The node kernels loop over
// their nodes with a grid stride, so one resident round of blocks keeps every
// CU busy to the end, where a larger grid leaves a partly filled last round.
template <auto K, int THREADS, int LDS = 0>
inline int resident_raw() {
  static const int blocks = [] {
    int dev = 0, cus = 0, per = 0;
    if (LDS > 65536) (void)hipFuncSetAttribute(reinterpret_cast<const void*>(K),
                                               hipFuncAttributeMaxDynamicSharedMemorySize, LDS);
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
      return -1;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, K, THREADS, LDS) != hipSuccess || per <= 0) return -1;
    return cus * per;
  }();
  return blocks;
}

// The same with the waves per CU rounded down to a multiple of the CU's 4 SIMDs:
// with a static split of the nodes, a SIMD holding one wave more than the others
// sets the kernel's time (chi = 8 BP sweep at 100k nodes: 10 one-wave blocks per
// CU 1.14 ms, 9 blocks 0.87 ms, 8 blocks 0.77 ms).
template <auto K, int THREADS>
inline int resident_blocks() {
  static const int blocks = [] {
    int dev = 0, cus = 0, per = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
      return 4096;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, K, THREADS, 0) != hipSuccess || per <= 0) return 4096;
    const int wpb = (THREADS + 63) / 64;
    const int waves = per * wpb;
    if (waves >= 4 && wpb <= 4 && 4 % wpb == 0) per = (waves - waves % 4) / wpb;
    return cus * per;
  }();
  return blocks;
}

template <auto K, int THREADS>
inline int node_grid(int n) {
  const int r = resident_blocks<K, THREADS>();
  return n < r ? n : r;
}

// Grid of a grid-stride node kernel under launch policy `policy`:
//   0  one resident round, waves per CU rounded to a multiple of 4 (node_grid);
//   1  one resident round, unrounded;
//   2  two resident rounds.
// Every policy computes each node in one block with the same code, so the
// results are identical; only the time differs.
constexpr int kGridPolicies = 3;
inline const char* grid_policy_name(int policy) {
  return policy == 0 ? "resident, SIMD-rounded" : policy == 1 ? "resident" : "2x resident";
}

template <auto K, int THREADS>
inline int policy_grid(int n, int policy) {
  int g = node_grid<K, THREADS>(n);
  if (policy != 0) {
    const int r = resident_raw<K, THREADS>();
    if (r > 0) g = policy == 1 ? r : 2 * r;
  }
  g = n < g ? n : g;
  return g > 0 ? g : 1;
}

// Kernels with more than 64 KB of dynamic LDS: the size is allowed once per kernel, and the
// resident-block count takes it into account.
template <auto K, int THREADS, int LDS>
inline int node_grid_lds(int n, int policy = 1) {
  int r = resident_raw<K, THREADS, LDS>();
  if (r <= 0) r = 1024;
  if (policy == 2) r *= 2;
  return n < r ? n : r;
}

}  // namespace bqa
