// Register/LDS-tiled kernels for the hot (degree, chi) shapes.  Each
// launcher returns false when it has no kernel for the shape; the caller then
// runs the generic kernel of kernels_node.hpp.
#pragma once

#include "common.hpp"
#include "kernels_node.hpp"
#include "kernels_mfma.hpp"

namespace bqa {

template <typename R>
bool fast_bp_launch(ClassView, int, const cplx<R>*, cplx<R>*, cplx<R>*, BpStatus*, hipStream_t) {
  return false;
}

template <typename R>
bool fast_local_launch(ClassView, int, int, const cplx<R>*, cplx<R>*, const cplx<double>*, int, int, double, int,
                       double, double, const double*, int, cplx<R>*, hipStream_t) {
  return false;
}

template <typename R>
bool fast_density_launch(ClassView, int, const cplx<R>*, const cplx<R>*, double*, double*, hipStream_t) {
  return false;
}

// FP32, chi = 8, degree 2 or 3: MFMA kernels (kernels_mfma.hpp).
inline int mfma8_grid(int n) { return n < 4096 ? n : 4096; }

template <>
inline bool fast_bp_launch<float>(ClassView cv, int chi, const cplx<float>* T, cplx<float>* mA, cplx<float>* mB,
                                  BpStatus* st, hipStream_t s) {
  if (chi != 8 || cv.n == 0 || (cv.d != 2 && cv.d != 3)) return false;
  if (cv.d == 3)
    hipLaunchKernelGGL((k_mfma8<3, kModeBP>), dim3(mfma8_grid(cv.n)), dim3(64), 0, s, cv, T, mA, mB, st,
                       (double*)nullptr, (double*)nullptr);
  else
    hipLaunchKernelGGL((k_mfma8<2, kModeBP>), dim3(mfma8_grid(cv.n)), dim3(64), 0, s, cv, T, mA, mB, st,
                       (double*)nullptr, (double*)nullptr);
  return true;
}

template <>
inline bool fast_density_launch<float>(ClassView cv, int chi, const cplx<float>* T, const cplx<float>* msgs,
                                       double* rho, double* bloch, hipStream_t s) {
  if (chi != 8 || cv.n == 0 || (cv.d != 2 && cv.d != 3)) return false;
  cplx<float>* m = const_cast<cplx<float>*>(msgs);
  if (cv.d == 3)
    hipLaunchKernelGGL((k_mfma8<3, kModeDensity>), dim3(mfma8_grid(cv.n)), dim3(64), 0, s, cv, T, m, m,
                       (BpStatus*)nullptr, rho, bloch);
  else
    hipLaunchKernelGGL((k_mfma8<2, kModeDensity>), dim3(mfma8_grid(cv.n)), dim3(64), 0, s, cv, T, m, m,
                       (BpStatus*)nullptr, rho, bloch);
  return true;
}

template <>
inline bool fast_local_launch<float>(ClassView cv, int chi_in, int chi_out, const cplx<float>* Tin,
                                     cplx<float>* Tout, const cplx<double>* canon, int rows, int cols,
                                     double zz_time, int flags, double z_time, double x_time, const double* lmbd,
                                     int lstride, cplx<float>* msgs_init, hipStream_t s) {
  if (chi_in != 8 || chi_out != 8 || cv.d != 3 || cv.n == 0 || flags != (kAbsorb | kGates | kGauge)) return false;
  hipLaunchKernelGGL((k_mfma8_local64<3, float>), dim3(mfma8_grid(cv.n)), dim3(64), 0, s, cv, Tin, Tout, canon, rows,
                     cols, zz_time, z_time, x_time, lmbd, lstride, msgs_init);
  return true;
}

}  // namespace bqa
