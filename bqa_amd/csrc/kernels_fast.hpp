// Register/LDS-tiled kernels for the hot (degree, chi) shapes.  Each
// launcher returns false when it has no kernel for the shape; the caller then
// runs the generic kernel of kernels_node.hpp.
#pragma once

#include "common.hpp"
#include "launch_grid.hpp"
#include "kernels_node.hpp"
#include "kernels_mfma.hpp"
#include "kernels_mfma16.hpp"
#include "kernels_local.hpp"
#include "kernels_small.hpp"
#include "kernels_d4.hpp"

namespace bqa {

// Compile-time-shaped one-wave kernels (kernels_small.hpp) for every (d, chi)
// they serve; the chi = 8 specialisations below (FP32 and FP64) use MFMA instead.
//
// The BP sweep has launch variants per (precision, degree, chi): kernels that split a node over
// a different number of waves, and the grid policies of launch_grid.hpp.  bp_variant_count and
// bp_variant_name describe them; fast_bp_launch(..., v) launches variant v (0 is the default),
// and the autotuner in bqa_api.hip times them on the job's first sweep at each shape.
template <typename R>
bool fast_bp_launch(ClassView cv, int chi, const cplx<R>* T, cplx<R>* mA, cplx<R>* mB, BpStatus* st,
                    hipStream_t s, int v = 0);

template <typename R>
int bp_variant_count(int d, int chi);

// Shapes whose BP sweep has a specialised kernel (fast_bp_launch returns true for them).
template <typename R>
inline bool bp_fast_shape(int d, int chi) {
  if ((chi == 8 || chi == 16) && (d == 2 || d == 3)) return true;
  if (chi == 8 && d == 4) return true;
  return small_bp_shape<R>(d, chi);
}

template <typename R>
const char* bp_variant_name(int d, int chi, int v);

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

// FP32, chi = 16, degree 2 or 3: k_mfma16_ext, W waves per node (128 KB of LDS at degree 3).
template <int D, int W>
inline void launch_mfma16_ext(ClassView cv, const cplx<float>* T, const cplx<float>* msgs, cplx<double>* ext,
                              hipStream_t s) {
  constexpr int LDS = Mf16ExtLds<D, W>::BYTES;
  const int grid = node_grid_lds<k_mfma16_ext<D, W>, 64 * W, LDS>(cv.n);
  hipLaunchKernelGGL((k_mfma16_ext<D, W>), dim3(grid), dim3(64 * W), LDS, s, cv, T, msgs, ext);
}

// Extended messages in FP64 from R storage: MFMA for chi = 8 and chi = 16 (degree 2, 3), the
// degree-4 chi = 8 MFMA kernel, k_small for the other served shapes; false
// leaves the generic kernel.  The ZZ-split gates of the step come from cv.zz.
// FP64, chi = 16, degree 2 or 3: k_mfma16_ext64, one physical half at a time.
template <int D, int W>
inline void launch_mfma16_ext64(ClassView cv, const cplx<double>* T, const cplx<double>* msgs, cplx<double>* ext,
                                hipStream_t s) {
  constexpr int LDS = Mf16DExt<D, W>::BYTES;
  const int grid = node_grid_lds<k_mfma16_ext64<D, W>, 64 * W, LDS>(cv.n);
  hipLaunchKernelGGL((k_mfma16_ext64<D, W>), dim3(grid), dim3(64 * W), LDS, s, cv, T, msgs, ext);
}

template <typename R>
bool fast_ext_launch(ClassView cv, int chi, const cplx<R>* T, const cplx<R>* msgs, cplx<double>* ext,
                     hipStream_t s) {
  if (cv.n == 0 || cv.zz == nullptr) return false;
  if (chi == 16 && (cv.d == 2 || cv.d == 3)) {
    if constexpr (sizeof(R) == sizeof(float)) {
      if (cv.d == 3)
        launch_mfma16_ext<3, 8>(cv, T, msgs, ext, s);
      else
        launch_mfma16_ext<2, 1>(cv, T, msgs, ext, s);
    } else {
      if (cv.d == 3)
        launch_mfma16_ext64<3, 8>(cv, T, msgs, ext, s);
      else
        launch_mfma16_ext64<2, 1>(cv, T, msgs, ext, s);
    }
    return true;
  }
  if (chi == 8 && cv.d == 4) {
    launch_ext_d4c8m<R, 4>(cv, T, msgs, ext, s);
    return true;
  }
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

// FP32, chi = 16, degree 2 or 3: k_mfma16_bp, eight waves per degree-3 node (128 KB of LDS,
// one node per CU), one wave per degree-2 node.
template <int D, int W>
inline void launch_mfma16_bp(ClassView cv, const cplx<float>* T, cplx<float>* mA, cplx<float>* mB, BpStatus* st,
                             hipStream_t s, int policy) {
  constexpr int LDS = Mf16Lds<D, W>::BYTES;
  const int grid = node_grid_lds<k_mfma16_bp<D, W>, 64 * W, LDS>(cv.n, policy);
  hipLaunchKernelGGL((k_mfma16_bp<D, W>), dim3(grid), dim3(64 * W), LDS, s, cv, T, mA, mB, st);
}

// FP32, chi = 16, degree 3: k_mfma16_bp3h (three mode products per node, one physical half at a time).
template <int W>
inline void launch_mfma16_bp3h(ClassView cv, const cplx<float>* T, cplx<float>* mA, cplx<float>* mB, BpStatus* st,
                               hipStream_t s) {
  constexpr int LDS = Mf16HLds<W>::BYTES;
  const int grid = node_grid_lds<k_mfma16_bp3h<W>, 64 * W, LDS>(cv.n, 1);
  hipLaunchKernelGGL((k_mfma16_bp3h<W>), dim3(grid), dim3(64 * W), LDS, s, cv, T, mA, mB, st);
}

// FP32 BP sweep variants.
//   chi = 16, d = 3: k_mfma16_bp<3, W>, W = 8 or 4 waves per node (one node per CU, 128 KB LDS),
//   k_mfma16_bp3h<8> (three products per node, 96 KB LDS);
//   chi = 16, d = 2: k_mfma16_bp<2, 1>, one or two resident rounds;
//   chi = 8, d = 2, 3: k_mfma8 (one wave per node) under the three grid policies;
//   chi = 8, d = 4: k_bp_d4c8 with 256 or 512 threads per node, k_bp_d4c8m (MFMA);
//   other served shapes: k_small under the three grid policies, and for degree 2, 3 with chi <= 4
//   also k_small_npw with 2 and 4 nodes per wave.
template <>
inline int bp_variant_count<float>(int d, int chi) {
  if (chi == 16 && d == 3) return 3;
  if (chi == 16 && d == 2) return 2;
  if (chi == 8 && d == 4) return 3;
  if (chi == 8 && (d == 2 || d == 3)) return kGridPolicies;
  if (npw_shape(d, chi)) return kGridPolicies + 2;
  return small_bp_shape<float>(d, chi) ? kGridPolicies : 1;
}

template <>
inline const char* bp_variant_name<float>(int d, int chi, int v) {
  if (chi == 16 && d == 3)
    return v == 0 ? "k_mfma16_bp<3,8>" : v == 1 ? "k_mfma16_bp<3,4>" : "k_mfma16_bp3h<8> (three products)";
  if (chi == 16 && d == 2) return v == 0 ? "k_mfma16_bp<2,1>, resident" : "k_mfma16_bp<2,1>, 2x resident";
  if (chi == 8 && d == 4)
    return v == 0 ? "k_bp_d4c8<256 threads>" : v == 1 ? "k_bp_d4c8<512 threads>" : "k_bp_d4c8m (FP32 MFMA)";
  static const char* m8[kGridPolicies] = {"k_mfma8, resident SIMD-rounded", "k_mfma8, resident",
                                          "k_mfma8, 2x resident"};
  static const char* sm[kGridPolicies + 2] = {"k_small, 8192 blocks", "k_small, resident", "k_small, 2x resident",
                                              "k_small_npw<2> (2 nodes per wave), resident",
                                              "k_small_npw<4> (4 nodes per wave), resident"};
  if (chi == 8 && (d == 2 || d == 3)) return m8[v % kGridPolicies];
  return small_bp_shape<float>(d, chi) ? sm[v % (kGridPolicies + 2)] : "generic";
}

// FP32, chi = 8, degree 2 or 3: k_mfma8, one wave per node (two and four waves per node for
// degree 3 measured no faster).
template <int MODE, int D>
inline void launch_mfma8_d(ClassView cv, const cplx<float>* T, cplx<float>* mA, cplx<float>* mB, BpStatus* st,
                           double* rho, double* bloch, hipStream_t s, int policy) {
  hipLaunchKernelGGL((k_mfma8<D, MODE, 1>), dim3(policy_grid<k_mfma8<D, MODE, 1>, 64>(cv.n, policy)), dim3(64), 0, s,
                     cv, T, mA, mB, st, rho, bloch);
}

template <>
inline bool fast_bp_launch<float>(ClassView cv, int chi, const cplx<float>* T, cplx<float>* mA, cplx<float>* mB,
                                  BpStatus* st, hipStream_t s, int v) {
  if (chi == 16 && cv.n > 0 && (cv.d == 2 || cv.d == 3)) {
    if (cv.d == 3) {
      if (v == 2)
        launch_mfma16_bp3h<8>(cv, T, mA, mB, st, s);
      else if (v == 1)
        launch_mfma16_bp<3, 4>(cv, T, mA, mB, st, s, 1);
      else
        launch_mfma16_bp<3, 8>(cv, T, mA, mB, st, s, 1);
    } else {
      launch_mfma16_bp<2, 1>(cv, T, mA, mB, st, s, v == 1 ? 2 : 1);
    }
    return true;
  }
  if (v >= kGridPolicies && npw_shape(cv.d, chi))
    return v == kGridPolicies ? launch_small_npw<2, float>(cv, chi, T, mA, mB, st, s)
                              : launch_small_npw<4, float>(cv, chi, T, mA, mB, st, s);
  if (chi == 8 && cv.d == 4 && cv.n > 0) {
    if (v == 2)
      launch_bp_d4c8m(cv, T, mA, mB, st, s);
    else if (v == 1)
      launch_bp_d4c8<float, 512>(cv, T, mA, mB, st, s);
    else
      launch_bp_d4c8<float, 256>(cv, T, mA, mB, st, s);
    return true;
  }
  if (chi != 8 || cv.n == 0 || (cv.d != 2 && cv.d != 3))
    return launch_small<kModeBP, float, float>(cv, chi, T, mA, mB, st, nullptr, nullptr, nullptr, s, v);
  if (cv.d == 2)
    launch_mfma8_d<kModeBP, 2>(cv, T, mA, mB, st, nullptr, nullptr, s, v);
  else
    launch_mfma8_d<kModeBP, 3>(cv, T, mA, mB, st, nullptr, nullptr, s, v);
  return true;
}

// FP64 BP sweep variants.
//   chi = 8, d = 3: k_mfma8_bp64_3p<4> (six contractions per node over four waves) under the three
//     grid policies, and k_mfma8_bp64_3p<2> (two waves per node);
//   chi = 8, d = 2: k_mfma8_bp64<2, 1> (one wave per node) under the three grid policies;
//   chi = 8, d = 4: k_bp_d4c8 with 256 or 512 threads per node, and k_bp_d4c8m64 (FP64 MFMA) with
//     four or eight waves per node;
//   other served shapes: k_small under the three grid policies, and for degree 2, 3 with chi <= 4
//   also k_small_npw with 2 and 4 nodes per wave.
template <>
inline int bp_variant_count<double>(int d, int chi) {
  if (chi == 8 && d == 4) return 4;
  if (chi == 8 && d == 3) return kGridPolicies + 1;
  if (chi == 8 && d == 2) return kGridPolicies;
  if (npw_shape(d, chi)) return kGridPolicies + 2;
  return small_shape(d, chi) ? kGridPolicies : 1;
}

template <>
inline const char* bp_variant_name<double>(int d, int chi, int v) {
  static const char* b3[kGridPolicies + 1] = {"k_mfma8_bp64_3p<4>, resident SIMD-rounded",
                                              "k_mfma8_bp64_3p<4>, resident", "k_mfma8_bp64_3p<4>, 2x resident",
                                              "k_mfma8_bp64_3p<2>, resident SIMD-rounded"};
  static const char* b2[kGridPolicies] = {"k_mfma8_bp64<2,1>, resident SIMD-rounded", "k_mfma8_bp64<2,1>, resident",
                                          "k_mfma8_bp64<2,1>, 2x resident"};
  static const char* sm[kGridPolicies + 2] = {"k_small, 8192 blocks", "k_small, resident", "k_small, 2x resident",
                                              "k_small_npw<2> (2 nodes per wave), resident",
                                              "k_small_npw<4> (4 nodes per wave), resident"};
  static const char* b4[4] = {"k_bp_d4c8<256 threads>", "k_bp_d4c8<512 threads>", "k_bp_d4c8m64<4 waves>",
                              "k_bp_d4c8m64<8 waves>"};
  if (chi == 8 && d == 4) return b4[v % 4];
  if (chi == 16 && (d == 2 || d == 3)) return d == 3 ? "k_mfma16_bp64<3,8>" : "k_mfma16_bp64<2,1>";
  if (chi == 8 && d == 3) return b3[v % (kGridPolicies + 1)];
  if (chi == 8 && d == 2) return b2[v % kGridPolicies];
  return small_shape(d, chi) ? sm[v % (kGridPolicies + 2)] : "generic";
}

// FP64, chi = 16, degree 2 or 3: k_mfma16_bp64 (eight waves per degree-3 node, one node per CU;
// one wave per degree-2 node).
template <int D, int W>
inline void launch_mfma16_bp64(ClassView cv, const cplx<double>* T, cplx<double>* mA, cplx<double>* mB,
                               BpStatus* st, hipStream_t s) {
  constexpr int LDS = Mf16D<D, W>::BYTES;
  const int grid = node_grid_lds<k_mfma16_bp64<D, W>, 64 * W, LDS>(cv.n, 1);
  hipLaunchKernelGGL((k_mfma16_bp64<D, W>), dim3(grid), dim3(64 * W), LDS, s, cv, T, mA, mB, st);
}

template <>
inline bool fast_bp_launch<double>(ClassView cv, int chi, const cplx<double>* T, cplx<double>* mA, cplx<double>* mB,
                                   BpStatus* st, hipStream_t s, int v) {
  if (chi == 16 && cv.n > 0 && (cv.d == 2 || cv.d == 3)) {
    if (cv.d == 3)
      launch_mfma16_bp64<3, 8>(cv, T, mA, mB, st, s);
    else
      launch_mfma16_bp64<2, 1>(cv, T, mA, mB, st, s);
    return true;
  }
  if (v >= kGridPolicies && npw_shape(cv.d, chi))
    return v == kGridPolicies ? launch_small_npw<2, double>(cv, chi, T, mA, mB, st, s)
                              : launch_small_npw<4, double>(cv, chi, T, mA, mB, st, s);
  if (chi == 8 && cv.d == 4 && cv.n > 0) {
    if (v == 1)
      launch_bp_d4c8<double, 512>(cv, T, mA, mB, st, s);
    else if (v == 2)
      launch_bp_d4c8m64<4>(cv, T, mA, mB, st, s);
    else if (v == 3)
      launch_bp_d4c8m64<8>(cv, T, mA, mB, st, s);
    else
      launch_bp_d4c8<double, 256>(cv, T, mA, mB, st, s);
    return true;
  }
  if (chi != 8 || cv.n == 0 || (cv.d != 2 && cv.d != 3))
    return launch_small<kModeBP, double, double>(cv, chi, T, mA, mB, st, nullptr, nullptr, nullptr, s, v);
  if (cv.d == 3) {
    if (v == kGridPolicies)
      hipLaunchKernelGGL((k_mfma8_bp64_3p<2>), dim3(node_grid<k_mfma8_bp64_3p<2>, 128>(cv.n)), dim3(128), 0, s, cv,
                         T, mA, mB, st);
    else
      hipLaunchKernelGGL((k_mfma8_bp64_3p<4>), dim3(policy_grid<k_mfma8_bp64_3p<4>, 256>(cv.n, v)), dim3(256), 0, s,
                         cv, T, mA, mB, st);
  } else {
    hipLaunchKernelGGL((k_mfma8_bp64<2, 1>), dim3(policy_grid<k_mfma8_bp64<2, 1>, 64>(cv.n, v)), dim3(64), 0, s, cv,
                       T, mA, mB, st);
  }
  return true;
}

// Multi-job BP launches (the serving batch): which kernel serves a class of a job of
// precision R in a launch shared with other jobs -- 1 k_small_multi, 2 k_mfma8_multi, or 0
// when the class keeps its own launches.  Only shapes whose single-job launch runs the same
// node code are shared, so a job's results are the same in a batch as alone.
// Shared-launch families: 1 k_small_multi (variant 0's arithmetic), 2 k_mfma8_multi, kFamNpw
// k_small_npw_multi with kNpwShared nodes per wave (for jobs whose BP variants are autotuned: the
// Grams' fibre split, and so the rounding, differs from variant 0), 0 none.
constexpr int kFamNpw = 3;
constexpr int kNpwShared = 4;

template <typename R>
inline int multi_family(int d, int chi) {
  if (sizeof(R) == sizeof(float) && chi == 8 && (d == 2 || d == 3)) return 2;
  if (sizeof(R) == sizeof(double) && chi == 8 && (d == 2 || d == 3)) return 0;  // FP64 MFMA kernels
  return small_bp_shape<R>(d, chi) ? 1 : 0;
}

template <typename R, int D, int CHI>
inline void launch_small_multi_dc(const BpEntry* tab, int entries, int total, int sweep, hipStream_t s) {
  hipLaunchKernelGGL((k_small_multi<D, CHI, R>), dim3(policy_grid<k_small_multi<D, CHI, R>, 64>(total, 0)), dim3(64),
                     0, s, tab, entries, sweep);
}

template <typename R, int D>
inline bool launch_small_multi_d(int chi, const BpEntry* tab, int entries, int total, int sweep, hipStream_t s) {
  switch (chi) {
    case 2: launch_small_multi_dc<R, D, 2>(tab, entries, total, sweep, s); return true;
    case 3: launch_small_multi_dc<R, D, 3>(tab, entries, total, sweep, s); return true;
    case 4: launch_small_multi_dc<R, D, 4>(tab, entries, total, sweep, s); return true;
    default: break;
  }
  if constexpr (D <= 3) {
    switch (chi) {
      case 5: launch_small_multi_dc<R, D, 5>(tab, entries, total, sweep, s); return true;
      case 6: launch_small_multi_dc<R, D, 6>(tab, entries, total, sweep, s); return true;
      case 7: launch_small_multi_dc<R, D, 7>(tab, entries, total, sweep, s); return true;
      case 8: launch_small_multi_dc<R, D, 8>(tab, entries, total, sweep, s); return true;
      default: break;
    }
  }
  return false;
}

// One shared launch of sweep `sweep` over `entries` classes (same family, degree and chi)
// holding `total` nodes (node groups for kFamNpw); false if the family has no such kernel.
template <typename R>
inline bool launch_bp_multi(int fam, int d, int chi, const BpEntry* tab, int entries, int total, int sweep,
                            hipStream_t s) {
  if (fam == 2) {
    if constexpr (sizeof(R) == sizeof(float)) {
      if (d == 2)
        hipLaunchKernelGGL((k_mfma8_multi<2>), dim3(policy_grid<k_mfma8_multi<2>, 64>(total, 0)), dim3(64), 0, s, tab,
                           entries, sweep);
      else
        hipLaunchKernelGGL((k_mfma8_multi<3>), dim3(policy_grid<k_mfma8_multi<3>, 64>(total, 0)), dim3(64), 0, s, tab,
                           entries, sweep);
      return true;
    }
    return false;
  }
  if (fam == kFamNpw) {
    // total counts node groups of kNpwShared nodes
#define BQA_NPW_MULTI(DD, CC)                                                                                   \
  if (d == DD && chi == CC) {                                                                                   \
    hipLaunchKernelGGL((k_small_npw_multi<DD, CC, kNpwShared, R>),                                              \
                       dim3(policy_grid<k_small_npw_multi<DD, CC, kNpwShared, R>, 64>(total, 0)), dim3(64), 0, s, \
                       tab, entries, sweep);                                                                    \
    return true;                                                                                                \
  }
    BQA_NPW_MULTI(2, 2) BQA_NPW_MULTI(2, 3) BQA_NPW_MULTI(2, 4) BQA_NPW_MULTI(3, 2) BQA_NPW_MULTI(3, 3)
    BQA_NPW_MULTI(3, 4)
#undef BQA_NPW_MULTI
    return false;
  }
  if (fam != 1) return false;
  switch (d) {
    case 2: return launch_small_multi_d<R, 2>(chi, tab, entries, total, sweep, s);
    case 3: return launch_small_multi_d<R, 3>(chi, tab, entries, total, sweep, s);
    case 4: return launch_small_multi_d<R, 4>(chi, tab, entries, total, sweep, s);
    case 5:
      if constexpr (sizeof(R) == sizeof(float)) return launch_small_multi_d<R, 5>(chi, tab, entries, total, sweep, s);
      return false;
    default: return false;
  }
}

template <>
inline bool fast_density_launch<float>(ClassView cv, int chi, const cplx<float>* T, const cplx<float>* msgs,
                                       double* rho, double* bloch, hipStream_t s) {
  cplx<float>* m = const_cast<cplx<float>*>(msgs);
  if (chi != 8 || cv.n == 0 || (cv.d != 2 && cv.d != 3))
    return launch_small<kModeDensity, float, float>(cv, chi, T, m, m, nullptr, nullptr, rho, bloch, s);
  if (cv.d == 2)
    launch_mfma8_d<kModeDensity, 2>(cv, T, m, m, nullptr, rho, bloch, s, 0);
  else
    launch_mfma8_d<kModeDensity, 3>(cv, T, m, m, nullptr, rho, bloch, s, 0);
  return true;
}

// chi = 8, degree 3 (FP32 and FP64 storage): k_mfma8_local64; chi = 16 at degrees 2-3 and
// chi = 8 at degree 4: k_local_lds; other shapes as above.
template <typename R>
inline bool fast_local_mfma(ClassView cv, int chi_in, int chi_out, const cplx<R>* Tin, cplx<R>* Tout,
                            const cplx<double>* canon, int rows, int cols, double zz_time, int flags, double z_time,
                            double x_time, const double* lmbd, int lstride, cplx<R>* msgs_init, hipStream_t s) {
  if (chi_in != chi_out || flags != (kAbsorb | kGates | kGauge) || cv.zz == nullptr) return false;
  if (launch_local_big<R>(cv, chi_in, Tin, Tout, canon, rows, cols, z_time, x_time, lmbd, lstride, msgs_init, s))
    return true;
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
