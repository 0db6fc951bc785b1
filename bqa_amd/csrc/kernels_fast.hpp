// Register/LDS-tiled kernels for the hot (degree, chi) shapes.  Each
// launcher returns false when it has no kernel for the shape; the caller then
// runs the generic kernel of kernels_node.hpp.
#pragma once

#include "common.hpp"
#include "kernels_node.hpp"

namespace bqa {

template <typename R>
bool fast_bp_launch(ClassView, int, const cplx<R>*, cplx<R>*, cplx<R>*, BpStatus*, hipStream_t) {
  return false;
}

template <typename R>
bool fast_ext_launch(ClassView, int, const cplx<R>*, const cplx<R>*, cplx<R>*, double, hipStream_t) {
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

}  // namespace bqa
