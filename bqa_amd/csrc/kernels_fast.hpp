// Register/LDS-tiled kernels for the hot (degree, chi) shapes.  Each
// launcher returns false when it has no kernel for the shape; the caller then
// runs the generic kernel of kernels_node.hpp.
#pragma once

#include "common.hpp"
#include "kernels_node.hpp"
#include "kernels_mfma.hpp"
#include "kernels_mfma16.hpp"
#include "kernels_small.hpp"

namespace bqa {

// Blocks of kernel K (THREADS per block) resident on the device at once: CUs x
// resident blocks per CU, computed once per kernel.  The node kernels loop over
// their nodes with a grid stride, so one resident round of blocks keeps every
// CU busy to the end, where a larger grid leaves a partly filled last round.
// The waves per CU are rounded down to a multiple of the CU's 4 SIMDs: with a
// static split of the nodes, a SIMD holding one wave more than the others sets
// the kernel's time (chi = 8 BP sweep at 100k nodes: 10 one-wave blocks per CU
// 1.14 ms, 9 blocks 0.87 ms, 8 blocks 0.77 ms).
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

// Kernels with more than 64 KB of dynamic LDS: the size is allowed once per kernel, and the
// resident-block count takes it into account.
template <auto K, int THREADS, int LDS>
inline int node_grid_lds(int n) {
  static const int blocks = [] {
    int dev = 0, cus = 0, per = 0;
    if (LDS > 65536) (void)hipFuncSetAttribute(reinterpret_cast<const void*>(K),
                                               hipFuncAttributeMaxDynamicSharedMemorySize, LDS);
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
      return 1024;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, K, THREADS, LDS) != hipSuccess || per <= 0) per = 1;
    return cus * per;
  }();
  return n < blocks ? n : blocks;
}

// Compile-time-shaped one-wave kernels (kernels_small.hpp) for every (d, chi)
// they serve; the chi = 8 specialisations below (FP32 and FP64) use MFMA instead.
template <typename R>
bool fast_bp_launch(ClassView cv, int chi, const cplx<R>* T, cplx<R>* mA, cplx<R>* mB, BpStatus* st,
                    hipStream_t s) {
  return launch_small<kModeBP, R, R>(cv, chi, T, mA, mB, st, nullptr, nullptr, nullptr, s);
}

// The fused absorb/gates/gauge step at a fixed bond dimension: one-wave
// kernels (kernels_small.hpp); the chi = 8, degree-3 case uses MFMA (both precisions).
template <typename R>
bool fast_local_launch(ClassView cv, int chi_in, int chi_out, const cplx<R>* Tin, cplx<R>* Tout,
                       const cplx<double>* canon, int rows, int cols, double, int flags, double z_time, double x_time,
                       const double* lmbd, int lstride, cplx<R>* msgs_init, hipStream_t s) {
  if (chi_in != chi_out || flags != (kAbsorb | kGates | kGauge)) return false;
  return launch_small_local<R>(cv, chi_in, Tin, Tout, canon, rows, cols, z_time, x_time, lmbd, lstride, msgs_init, s);
}

template <typename R>
bool fast_density_launch(ClassView cv, int chi, const cplx<R>* T, const cplx<R>* msgs, double* rho, double* bloch,
                         hipStream_t s) {
  cplx<R>* m = const_cast<cplx<R>*>(msgs);
  return launch_small<kModeDensity, R, R>(cv, chi, T, m, m, nullptr, nullptr, rho, bloch, s);
}

// Extended messages in FP64 from R storage: MFMA for chi = 8 (degree 2, 3),
// k_small for the other served shapes; false leaves the generic kernel.  The
// ZZ-split gates of the step come from cv.zz.
template <typename R>
bool fast_ext_launch(ClassView cv, int chi, const cplx<R>* T, const cplx<R>* msgs, cplx<double>* ext,
                     hipStream_t s) {
  if (cv.n == 0 || cv.zz == nullptr) return false;
  if (chi == 8 && (cv.d == 2 || cv.d == 3)) {
    if constexpr (sizeof(R) == sizeof(float)) {
      // FP32 job: FP32 MFMA mode products, FP64 MFMA Gram
      if (cv.d == 3)
        hipLaunchKernelGGL((k_mfma8_ext32<3, 4>), dim3(node_grid<k_mfma8_ext32<3, 4>, 256>(cv.n)), dim3(256), 0, s,
                           cv, T, msgs, ext);
      else
        hipLaunchKernelGGL((k_mfma8_ext32<2, 1>), dim3(node_grid<k_mfma8_ext32<2, 1>, 64>(cv.n)), dim3(64), 0, s, cv,
                           T, msgs, ext);
    } else {
      if (cv.d == 3)
        hipLaunchKernelGGL((k_mfma8_ext64<3, 4, R>), dim3(node_grid<k_mfma8_ext64<3, 4, R>, 256>(cv.n)), dim3(256),
                           0, s, cv, T, msgs, ext);
      else
        hipLaunchKernelGGL((k_mfma8_ext64<2, 1, R>), dim3(node_grid<k_mfma8_ext64<2, 1, R>, 64>(cv.n)), dim3(64), 0,
                           s, cv, T, msgs, ext);
    }
    return true;
  }
  cplx<R>* m = const_cast<cplx<R>*>(msgs);
  return launch_small<kModeExt, R, double>(cv, chi, T, m, m, nullptr, ext, nullptr, nullptr, s);
}

// FP32, chi = 8, degree 2 or 3: MFMA kernels (kernels_mfma.hpp), one wave per node (two and
// four waves per node for degree 3 measured no faster).
template <int MODE>
inline void launch_mfma8(ClassView cv, const cplx<float>* T, cplx<float>* mA, cplx<float>* mB, BpStatus* st,
                         double* rho, double* bloch, hipStream_t s) {
  if (cv.d == 2)
    hipLaunchKernelGGL((k_mfma8<2, MODE, 1>), dim3(node_grid<k_mfma8<2, MODE, 1>, 64>(cv.n)), dim3(64), 0, s, cv, T,
                       mA, mB, st, rho, bloch);
  else
    hipLaunchKernelGGL((k_mfma8<3, MODE, 1>), dim3(node_grid<k_mfma8<3, MODE, 1>, 64>(cv.n)), dim3(64), 0, s, cv, T,
                       mA, mB, st, rho, bloch);
}

// FP32, chi = 16, degree 2 or 3: k_mfma16_bp, eight waves per degree-3 node (128 KB of LDS,
// one node per CU), one wave per degree-2 node.
template <int D, int W>
inline void launch_mfma16_bp(ClassView cv, const cplx<float>* T, cplx<float>* mA, cplx<float>* mB, BpStatus* st,
                             hipStream_t s) {
  constexpr int LDS = Mf16Lds<D, W>::BYTES;
  const int grid = node_grid_lds<k_mfma16_bp<D, W>, 64 * W, LDS>(cv.n);
  hipLaunchKernelGGL((k_mfma16_bp<D, W>), dim3(grid), dim3(64 * W), LDS, s, cv, T, mA, mB, st);
}

template <>
inline bool fast_bp_launch<float>(ClassView cv, int chi, const cplx<float>* T, cplx<float>* mA, cplx<float>* mB,
                                  BpStatus* st, hipStream_t s) {
  if (chi == 16 && cv.n > 0 && (cv.d == 2 || cv.d == 3)) {
    if (cv.d == 3)
      launch_mfma16_bp<3, 8>(cv, T, mA, mB, st, s);
    else
      launch_mfma16_bp<2, 1>(cv, T, mA, mB, st, s);
    return true;
  }
  if (chi != 8 || cv.n == 0 || (cv.d != 2 && cv.d != 3))
    return launch_small<kModeBP, float, float>(cv, chi, T, mA, mB, st, nullptr, nullptr, nullptr, s);
  launch_mfma8<kModeBP>(cv, T, mA, mB, st, nullptr, nullptr, s);
  return true;
}

// FP64, chi = 8: FP64 MFMA; degree 3 in six contractions per node over four
// waves (k_mfma8_bp64_3p), degree 2 one wave per node (k_mfma8_bp64).
template <>
inline bool fast_bp_launch<double>(ClassView cv, int chi, const cplx<double>* T, cplx<double>* mA, cplx<double>* mB,
                                   BpStatus* st, hipStream_t s) {
  if (chi != 8 || cv.n == 0 || (cv.d != 2 && cv.d != 3))
    return launch_small<kModeBP, double, double>(cv, chi, T, mA, mB, st, nullptr, nullptr, nullptr, s);
  if (cv.d == 3)
    hipLaunchKernelGGL((k_mfma8_bp64_3p<4>), dim3(node_grid<k_mfma8_bp64_3p<4>, 256>(cv.n)), dim3(256), 0, s, cv, T,
                       mA, mB, st);
  else
    hipLaunchKernelGGL((k_mfma8_bp64<2, 1>), dim3(node_grid<k_mfma8_bp64<2, 1>, 64>(cv.n)), dim3(64), 0, s, cv, T,
                       mA, mB, st);
  return true;
}

template <>
inline bool fast_density_launch<float>(ClassView cv, int chi, const cplx<float>* T, const cplx<float>* msgs,
                                       double* rho, double* bloch, hipStream_t s) {
  cplx<float>* m = const_cast<cplx<float>*>(msgs);
  if (chi != 8 || cv.n == 0 || (cv.d != 2 && cv.d != 3))
    return launch_small<kModeDensity, float, float>(cv, chi, T, m, m, nullptr, nullptr, rho, bloch, s);
  launch_mfma8<kModeDensity>(cv, T, m, m, nullptr, rho, bloch, s);
  return true;
}

// chi = 8, degree 3 (FP32 and FP64 storage): k_mfma8_local64; other shapes as above.
template <typename R>
inline bool fast_local_mfma(ClassView cv, int chi_in, int chi_out, const cplx<R>* Tin, cplx<R>* Tout,
                            const cplx<double>* canon, int rows, int cols, double zz_time, int flags, double z_time,
                            double x_time, const double* lmbd, int lstride, cplx<R>* msgs_init, hipStream_t s) {
  if (chi_in != chi_out || flags != (kAbsorb | kGates | kGauge) || cv.zz == nullptr) return false;
  if (chi_in != 8 || cv.d != 3 || cv.n == 0)
    return launch_small_local<R>(cv, chi_in, Tin, Tout, canon, rows, cols, z_time, x_time, lmbd, lstride, msgs_init,
                                 s);
  hipLaunchKernelGGL((k_mfma8_local64<3, 4, R>), dim3(node_grid<k_mfma8_local64<3, 4, R>, 256>(cv.n)), dim3(256), 0,
                     s, cv, Tin, Tout, canon, rows, cols, zz_time, z_time, x_time, lmbd, lstride, msgs_init);
  return true;
}

template <>
inline bool fast_local_launch<float>(ClassView cv, int chi_in, int chi_out, const cplx<float>* Tin,
                                     cplx<float>* Tout, const cplx<double>* canon, int rows, int cols,
                                     double zz_time, int flags, double z_time, double x_time, const double* lmbd,
                                     int lstride, cplx<float>* msgs_init, hipStream_t s) {
  return fast_local_mfma<float>(cv, chi_in, chi_out, Tin, Tout, canon, rows, cols, zz_time, flags, z_time, x_time,
                                lmbd, lstride, msgs_init, s);
}

template <>
inline bool fast_local_launch<double>(ClassView cv, int chi_in, int chi_out, const cplx<double>* Tin,
                                      cplx<double>* Tout, const cplx<double>* canon, int rows, int cols,
                                      double zz_time, int flags, double z_time, double x_time, const double* lmbd,
                                      int lstride, cplx<double>* msgs_init, hipStream_t s) {
  return fast_local_mfma<double>(cv, chi_in, chi_out, Tin, Tout, canon, rows, cols, zz_time, flags, z_time, x_time,
                                 lmbd, lstride, msgs_init, s);
}

}  // namespace bqa
