// Shared device helpers: interleaved complex arithmetic, reductions, ZZ split.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace bqa {

template <typename R>
struct alignas(2 * sizeof(R)) cplx {
  R x, y;
};

template <typename R>
__device__ __host__ __forceinline__ cplx<R> mk(R a, R b) {
  cplx<R> r;
  r.x = a;
  r.y = b;
  return r;
}

template <typename R>
__device__ __host__ __forceinline__ cplx<R> operator+(cplx<R> a, cplx<R> b) {
  return mk(a.x + b.x, a.y + b.y);
}
template <typename R>
__device__ __host__ __forceinline__ cplx<R> operator-(cplx<R> a, cplx<R> b) {
  return mk(a.x - b.x, a.y - b.y);
}
template <typename R>
__device__ __host__ __forceinline__ cplx<R> operator*(cplx<R> a, cplx<R> b) {
  return mk(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
template <typename R>
__device__ __host__ __forceinline__ cplx<R> scale(cplx<R> a, R s) {
  return mk(a.x * s, a.y * s);
}
template <typename R>
__device__ __host__ __forceinline__ cplx<R> conj(cplx<R> a) {
  return mk(a.x, -a.y);
}
// acc += a * b
template <typename R>
__device__ __forceinline__ void cfma(cplx<R>& acc, cplx<R> a, cplx<R> b) {
  acc.x = fma(a.x, b.x, acc.x);
  acc.x = fma(-a.y, b.y, acc.x);
  acc.y = fma(a.x, b.y, acc.y);
  acc.y = fma(a.y, b.x, acc.y);
}
// acc += conj(a) * b
template <typename R>
__device__ __forceinline__ void cfma_cj(cplx<R>& acc, cplx<R> a, cplx<R> b) {
  acc.x = fma(a.x, b.x, acc.x);
  acc.x = fma(a.y, b.y, acc.x);
  acc.y = fma(a.x, b.y, acc.y);
  acc.y = fma(-a.y, b.x, acc.y);
}
template <typename R>
__device__ __host__ __forceinline__ R abs2(cplx<R> a) {
  return a.x * a.x + a.y * a.y;
}
template <typename R>
__device__ __forceinline__ R cabs(cplx<R> a) {
  return hypot(a.x, a.y);
}
// 1 / z
template <typename R>
__device__ __forceinline__ cplx<R> cinv(cplx<R> z) {
  R d = z.x * z.x + z.y * z.y;
  return mk(z.x / d, -z.y / d);
}

template <typename To, typename From>
__device__ __host__ __forceinline__ cplx<To> cast(cplx<From> a) {
  return mk((To)a.x, (To)a.y);
}

// Principal square root of a real number as a complex number: sqrt(x) for
// x >= 0, i*sqrt(-x) for x < 0.
__device__ __host__ inline cplx<double> csqrt_real(double x) {
  return x >= 0.0 ? mk(sqrt(x), 0.0) : mk(0.0, sqrt(-x));
}

// Ket-side split of exp(-i theta ZZ) over one bond:
// g0 = sqrt(cos theta), g1 = e^{-i pi/4} sqrt(sin theta), so that
// g0^2 I(x)I + g1^2 Z(x)Z = exp(-i theta ZZ).  Any fixed branch works as long
// as both bond ends and every use within a step take the same one.
__device__ __host__ inline void zz_split(double theta, cplx<double>& g0, cplx<double>& g1) {
  g0 = csqrt_real(cos(theta));
  cplx<double> s = csqrt_real(sin(theta));
  const double h = 0.70710678118654752440;
  g1 = s * mk(h, -h);
}

// Non-negative double compares like its bit pattern as uint64.
__device__ __forceinline__ void atomic_max_nonneg(unsigned long long* addr, double v) {
  if (!(v >= 0.0)) v = __longlong_as_double(0x7ff8000000000000ULL);  // NaN marker: largest pattern
  atomicMax(addr, (unsigned long long)__double_as_longlong(v));
}

// Block-wide max of a non-negative double (blockDim.x <= 1024, multiple of 64).
__device__ __forceinline__ double block_max(double v, double* sh) {
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) sh[wid] = v;
  __syncthreads();
  if (wid == 0) {
    const int nw = (blockDim.x + 63) >> 6;
    v = lane < nw ? sh[lane] : 0.0;
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  }
  return v;  // valid in wave 0
}

__device__ __forceinline__ double block_sum(double v, double* sh) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) sh[wid] = v;
  __syncthreads();
  const int nw = (blockDim.x + 63) >> 6;
  double s = 0.0;
  for (int w = 0; w < nw; ++w) s += sh[w];
  __syncthreads();
  return s;  // valid in every thread
}

__host__ __device__ inline int ipow(int b, int e) {
  int r = 1;
  for (int i = 0; i < e; ++i) r *= b;
  return r;
}

constexpr int kSweepSlots = 128;

// BP control block living in device memory.  Two protocols:
//  - status (sweep index < 0; damping runs): the sweep kernels read `cur` and
//    `done`, and the finish kernel (one thread) updates them after every sweep;
//  - sweep slots (sweep index >= 0): each sweep writes its distance maxima to
//    its own slot, each launch decides from the previous sweep's slot whether
//    the run is over, and the host applies the rules to the slots after every
//    chunk of sweeps, so no kernel runs between sweeps.
struct BpStatus {
  unsigned long long num_bits;  // max |new - old| as double bits
  unsigned long long den_bits;  // max |new + old|
  int cur;                      // index of the current message buffer (slots: of the run's first sweep)
  int done;                     // 1 once converged or out of iterations
  int converged;
  int iters;                    // sweeps performed
  int have_best;                // some sweep gave a finite distance
  int damp_now;                 // the damping kernel must run after this sweep
  int max_iters;
  int marker;                   // slots: a launch found the run finished before its sweep
  int base;                     // slots: added to every launch's sweep index (graph replays advance it)
  double bp_eps;
  double last_dist;
  double best_dist;
  unsigned long long slot[kSweepSlots][2];  // slots: the two maxima of each sweep
};

// Buffer index a BP launch of sweep `sweep` reads, or -1 when the run is over.
// With slots (sweep >= 0), the launch's sweep is st->base + sweep; it runs
// unless it is beyond the run's max_iters or an earlier sweep converged (the
// previous sweep's distance below bp_eps, or a launch before it said so), and
// the buffers alternate from the run's first one.
__device__ inline int bp_begin(BpStatus* st, int sweep) {
  if (sweep < 0) return st->done ? -1 : st->cur;
  sweep += st->base;
  if (sweep >= st->max_iters) return -1;
  if (sweep > 0) {
    // blocks of one launch may set the marker while others read it; either
    // outcome gives the same decision, since both follow from the same slot
    if (__hip_atomic_load(&st->marker, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return -1;
    const double num = __longlong_as_double((long long)st->slot[sweep - 1][0]);
    const double den = __longlong_as_double((long long)st->slot[sweep - 1][1]);
    if (num / den < st->bp_eps) {
      if (threadIdx.x == 0) __hip_atomic_store(&st->marker, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      return -1;
    }
  }
  return st->cur ^ (sweep & 1);
}

// This block's distance maxima of sweep `sweep` (one thread per block).
__device__ inline void bp_end(BpStatus* st, int sweep, double dnum, double dden) {
  if (sweep >= 0) sweep += st->base;
  atomic_max_nonneg(sweep < 0 ? &st->num_bits : &st->slot[sweep][0], dnum);
  atomic_max_nonneg(sweep < 0 ? &st->den_bits : &st->slot[sweep][1], dden);
}

}  // namespace bqa
