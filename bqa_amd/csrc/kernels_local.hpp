// Fused absorb + Z/X gates + symmetric gauge + message re-initialisation for
// shapes with large node tensors: chi = 16 at degrees 2 and 3 (the heavy-hex
// chi = 16 anneal), chi = 8 at degree 4 (the square grid's interior) and chi = 4
// at degree 5 (serving jobs).  Same operation as k_node_local with all three flags and chi_in =
// chi_out (backends.py:416-432, state.py:142-156, 219-227):
//
//   absorb, bond k:  T[p, .., a_k, ..] <- sum_c T[p, .., c, ..] W_p[c][a_k],
//                    W_p[c][a] = g_0 C_k[c][a] + (-1)^p g_1 C_k[chi + c][a]
//                    (C_k the bond's canonicaliser, (g_0, g_1) its ZZ-split gates);
//   gates:           Z phase exp(-+ i h_i z_time) on p, then cos(x) T - i sin(x) X T;
//   gauge:           T *= prod_k sqrt(lambda_k[a_k]), normalised to unit L2 norm;
//   messages:        the node's incoming messages become diag(lambda_k) / tr.
//
// The tensor lives in LDS in FP64 (both storage precisions: FP32 rounding inside
// the absorb is amplified by the canonicalisers' S^{-1/2}); a mode product runs in
// place, each thread owning whole fibres along the bond (it reads a fibre, then
// writes it back), with the bond's 2 x chi x chi operand W built in LDS just before.
// VALU FP64: the absorb is 2 chi^(d+1) d complex multiply-adds per node, far from
// the FP64 peak at these node counts.
#pragma once

#include "common.hpp"
#include "kernels_node.hpp"
#include "launch_grid.hpp"

namespace bqa {

constexpr int cpow_i(int b, int e) { return e <= 0 ? 1 : b * cpow_i(b, e - 1); }

template <int D, int CHI>
struct LocalShape {
  static constexpr int A = cpow_i(CHI, D);   // elements of one physical half
  static constexpr int S = 2 * A;
  static constexpr int FIB = S / CHI;        // fibres along one bond
  static constexpr int L = CHI == 16 ? 4 : CHI == 8 ? 3 : 2;   // log2 CHI
  static_assert(CHI == 16 || CHI == 8 || CHI == 4, "power-of-two bond dimension");
  // LDS: the tensor, one bond's W_0 and W_1, sqrt(lambda) per bond, the norm partials
  static constexpr int BYTES = S * 16 + 2 * CHI * CHI * 16 + D * CHI * 8 + 16 * 8;
};

// LDS slot of element idx: the lowest bond's index XORed with the next one's, a
// permutation inside each aligned row of CHI elements, so a fibre along the last bond
// (read whole by one thread) falls on different banks across the lanes.
template <int CHI>
__device__ __forceinline__ int lsw(int idx) {
  constexpr int L = CHI == 16 ? 4 : CHI == 8 ? 3 : 2;
  return idx ^ ((idx >> L) & (CHI - 1));
}

// W_p[c][a] of bond k at node i (p = 0, 1) into Wt[p * CHI * CHI + c * CHI + a].
template <int D, int CHI, int NT>
__device__ __forceinline__ void local_operand(cplx<double>* Wt, const ClassView& cv, int i, int k,
                                              const zd* __restrict__ canon, int canon_rows, int canon_cols) {
  const int n = cv.n;
  const zd* C = canon + (size_t)cv.in_pos[k * n + i] * canon_rows * canon_cols;
  const cplx<double>* zg = cv.zz + 2 * (size_t)cv.lmbd_pos[k * n + i];
  const cplx<double> g0 = zg[0], g1 = zg[1];
  for (int q = threadIdx.x; q < CHI * CHI; q += NT) {
    const int c = q / CHI, a = q % CHI;
    const cplx<double> u = g0 * C[c * canon_cols + a], v = g1 * C[(CHI + c) * canon_cols + a];
    Wt[q] = u + v;
    Wt[CHI * CHI + q] = u - v;
  }
}

// In-place absorb along bond J.  The threads of each physical half (the first NT / 2 hold
// p = 0) take its fibres in turn, FPT fibres at a time, so each W entry read from LDS
// serves FPT fibres (the W broadcasts, not the arithmetic, bound this loop).  SPLIT
// threads share a fibre (each computing CHI / SPLIT outputs); they are lanes of one wave,
// which reads the fibre whole before writing it.
template <int D, int CHI, int J, int NT, int SPLIT, int FPT>
__device__ __forceinline__ void local_mp(cplx<double>* X, const cplx<double>* Wt) {
  using Sh = LocalShape<D, CHI>;
  constexpr int st = cpow_i(CHI, D - 1 - J);
  constexpr int HF = Sh::FIB / 2;            // fibres with p = 0 (p is the most significant index)
  constexpr int TH = NT / 2;                 // threads per physical half
  constexpr int PER = HF * SPLIT / TH;       // (fibre, share) items per thread
  constexpr int OUT = CHI / SPLIT;
  static_assert(NT % 128 == 0 || NT == 64, "whole waves, or one wave, per physical half");
  static_assert((HF * SPLIT) % TH == 0 && PER % FPT == 0 && TH % SPLIT == 0, "whole rounds of fibres");
  static_assert(SPLIT == 1 || NT == 64, "a shared fibre stays inside one wave");
  const int g = threadIdx.x / TH, r = threadIdx.x % TH;
  const int h = r % SPLIT;
  const cplx<double>* W = Wt + g * CHI * CHI + h * OUT;
#pragma unroll 1
  for (int q0 = 0; q0 < PER; q0 += FPT) {
    int base[FPT];
    cplx<double> x[FPT][CHI];
#pragma unroll
    for (int u = 0; u < FPT; ++u) {
      const int f = g * HF + (r + TH * (q0 + u)) / SPLIT;
      base[u] = (f / st) * CHI * st + f % st;
#pragma unroll
      for (int c = 0; c < CHI; ++c) x[u][c] = X[lsw<CHI>(base[u] + c * st)];
    }
    cplx<double> y[FPT][OUT];
#pragma unroll
    for (int u = 0; u < FPT; ++u)
#pragma unroll
      for (int a = 0; a < OUT; ++a) y[u][a] = mk(0.0, 0.0);
#pragma unroll
    for (int c = 0; c < CHI; ++c)
#pragma unroll
      for (int a = 0; a < OUT; ++a) {
        const cplx<double> w = W[c * CHI + a];
#pragma unroll
        for (int u = 0; u < FPT; ++u) cfma(y[u][a], x[u][c], w);
      }
#pragma unroll
    for (int u = 0; u < FPT; ++u)
#pragma unroll
      for (int a = 0; a < OUT; ++a) X[lsw<CHI>(base[u] + (h * OUT + a) * st)] = y[u][a];
  }
}

template <int D, int CHI, int NT, int SPLIT, int FPT, typename R>
__global__ __launch_bounds__(NT) void k_local_lds(ClassView cv, const cplx<R>* __restrict__ Tin, cplx<R>* Tout,
                                                  const zd* __restrict__ canon, int canon_rows, int canon_cols,
                                                  double z_time, double x_time, const double* __restrict__ lmbd,
                                                  int lstride, cplx<R>* msgs_init) {
  using Sh = LocalShape<D, CHI>;
  constexpr int S = Sh::S, A = Sh::A, L = Sh::L;
  static_assert(S % NT == 0 && A % NT == 0, "tensor passes in whole rounds");
  extern __shared__ __align__(16) unsigned char smem_local[];
  cplx<double>* X = reinterpret_cast<cplx<double>*>(smem_local);
  cplx<double>* Wt = X + S;
  double* sl = reinterpret_cast<double*>(Wt + 2 * CHI * CHI);   // sqrt(lambda_k[a]) at [k * CHI + a]
  double* red = sl + D * CHI;
  const int n = cv.n, t = threadIdx.x;
  const double cx = cos(x_time), sx = sin(x_time);
  for (int i = blockIdx.x; i < n; i += gridDim.x) {
    const cplx<R>* Tg = Tin + (size_t)i * S;
#pragma unroll 4
    for (int k = 0; k < S / NT; ++k) X[lsw<CHI>(t + NT * k)] = cast<double>(Tg[t + NT * k]);
    for (int q = t; q < D * CHI; q += NT) {
      const int k = q / CHI, a = q % CHI;
      sl[q] = sqrt(lmbd[(size_t)cv.lmbd_pos[k * n + i] * lstride + a]);
    }
    local_operand<D, CHI, NT>(Wt, cv, i, 0, canon, canon_rows, canon_cols);
    __syncthreads();
    local_mp<D, CHI, 0, NT, SPLIT, FPT>(X, Wt);
    __syncthreads();
    local_operand<D, CHI, NT>(Wt, cv, i, 1, canon, canon_rows, canon_cols);
    __syncthreads();
    local_mp<D, CHI, 1, NT, SPLIT, FPT>(X, Wt);
    if constexpr (D >= 3) {
      __syncthreads();
      local_operand<D, CHI, NT>(Wt, cv, i, 2, canon, canon_rows, canon_cols);
      __syncthreads();
      local_mp<D, CHI, 2, NT, SPLIT, FPT>(X, Wt);
    }
    if constexpr (D >= 4) {
      __syncthreads();
      local_operand<D, CHI, NT>(Wt, cv, i, 3, canon, canon_rows, canon_cols);
      __syncthreads();
      local_mp<D, CHI, 3, NT, SPLIT, FPT>(X, Wt);
    }
    if constexpr (D >= 5) {
      __syncthreads();
      local_operand<D, CHI, NT>(Wt, cv, i, 4, canon, canon_rows, canon_cols);
      __syncthreads();
      local_mp<D, CHI, 4, NT, SPLIT, FPT>(X, Wt);
    }
    static_assert(D >= 2 && D <= 5, "degrees 2 to 5");
    __syncthreads();
    // gates and gauge on the element pairs (a, A + a); element a's bond indices are its
    // base-CHI digits, the last bond lowest
    const double phi = cv.h[i] * z_time;
    const double cz = cos(phi), sz = sin(phi);
    const cplx<double> f0 = mk(cz, -sz), f1 = mk(cz, sz);
    double ss = 0.0;
#pragma unroll 4
    for (int k = 0; k < A / NT; ++k) {
      const int a = t + NT * k;
      const int s0 = lsw<CHI>(a), s1 = lsw<CHI>(A + a);
      const cplx<double> t0 = X[s0] * f0, t1 = X[s1] * f1;
      double f = 1.0;
#pragma unroll
      for (int b = 0; b < D; ++b) f *= sl[b * CHI + ((a >> (L * (D - 1 - b))) & (CHI - 1))];
      const cplx<double> y0 = mk(f * (cx * t0.x + sx * t1.y), f * (cx * t0.y - sx * t1.x));
      const cplx<double> y1 = mk(f * (cx * t1.x + sx * t0.y), f * (cx * t1.y - sx * t0.x));
      X[s0] = y0;
      X[s1] = y1;
      ss += abs2(y0) + abs2(y1);
    }
    for (int o = 32; o > 0; o >>= 1) ss += __shfl_xor(ss, o, 64);
    if ((t & 63) == 0) red[t >> 6] = ss;
    __syncthreads();
    double tot = 0.0;
#pragma unroll
    for (int w = 0; w < NT / 64; ++w) tot += red[w];
    const double inv = 1.0 / sqrt(tot);
    cplx<R>* To = Tout + (size_t)i * S;
#pragma unroll 4
    for (int k = 0; k < S / NT; ++k) To[t + NT * k] = cast<R>(scale(X[lsw<CHI>(t + NT * k)], inv));
    for (int q = t; q < D * CHI * CHI; q += NT) {
      const int k = q / (CHI * CHI), a = (q / CHI) % CHI, b = q % CHI;
      const double* lam = lmbd + (size_t)cv.lmbd_pos[k * n + i] * lstride;
      double tr = 0.0;
#pragma unroll
      for (int e = 0; e < CHI; ++e) tr += lam[e];
      msgs_init[(size_t)cv.in_pos[k * n + i] * CHI * CHI + (q % (CHI * CHI))] =
          mk<R>(a == b ? (R)(lam[a] / tr) : (R)0, (R)0);
    }
    __syncthreads();
  }
}

template <int D, int CHI, int NT, int SPLIT, int FPT, typename R>
inline void launch_local_lds(ClassView cv, const cplx<R>* Tin, cplx<R>* Tout, const zd* canon, int rows, int cols,
                             double z_time, double x_time, const double* lmbd, int lstride, cplx<R>* msgs_init,
                             hipStream_t s) {
  constexpr int LDS = LocalShape<D, CHI>::BYTES;
  const int grid = node_grid_lds<k_local_lds<D, CHI, NT, SPLIT, FPT, R>, NT, LDS>(cv.n);
  hipLaunchKernelGGL((k_local_lds<D, CHI, NT, SPLIT, FPT, R>), dim3(grid), dim3(NT), LDS, s, cv, Tin, Tout, canon, rows,
                     cols, z_time, x_time, lmbd, lstride, msgs_init);
}

// The shapes above (chi_in = chi_out); false leaves the caller's other kernels.
template <typename R>
inline bool launch_local_big(ClassView cv, int chi, const cplx<R>* Tin, cplx<R>* Tout, const zd* canon, int rows,
                             int cols, double z_time, double x_time, const double* lmbd, int lstride,
                             cplx<R>* msgs_init, hipStream_t s) {
  if (cv.n == 0 || cv.zz == nullptr || cols < chi || rows != 2 * chi) return false;
  if (chi == 16 && cv.d == 3)
    launch_local_lds<3, 16, 512, 1, 1, R>(cv, Tin, Tout, canon, rows, cols, z_time, x_time, lmbd, lstride, msgs_init, s);
  else if (chi == 16 && cv.d == 2)
    launch_local_lds<2, 16, 64, 2, 1, R>(cv, Tin, Tout, canon, rows, cols, z_time, x_time, lmbd, lstride, msgs_init, s);
  else if (chi == 8 && cv.d == 4)
    launch_local_lds<4, 8, 512, 1, 2, R>(cv, Tin, Tout, canon, rows, cols, z_time, x_time, lmbd, lstride, msgs_init, s);
  else if (chi == 4 && cv.d == 5)
    launch_local_lds<5, 4, 256, 1, 2, R>(cv, Tin, Tout, canon, rows, cols, z_time, x_time, lmbd, lstride, msgs_init, s);
  else
    return false;
  return true;
}

}  // namespace bqa
