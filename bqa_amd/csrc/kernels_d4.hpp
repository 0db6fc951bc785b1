// BP sweep and extended messages for degree-4 nodes at chi = 8 (the 100x100
// grid's interior, the d = 4 serving jobs): k_bp_d4c8 on VALU, k_bp_d4c8m (FP32)
// and k_bp_d4c8m64 (FP64) on MFMA, k_ext_d4c8m on MFMA.  A node tensor is
// 2 x 8^4 complex elements (128 KB in FP64), so the kernels work on one physical
// half T_p (4096 elements) at a time, in two LDS buffers A and B, with mode
// products done in place (a thread or a wave owns whole fibres along the
// product's bond, reads them, then writes them back).
//
// The reference computes out_k = Gram over every bond but k of conj(T) against
// T contracted with the three other incoming messages (backends.py:381-404).
// Every message is Hermitian, so a message can act on the bra side instead:
// sum_a conj(T[..a..]) m[a][b] = conj((T x m)[..b..]).  With that, per half:
//   B = T x2 m2 x3 m3;  out_0 += conj(T x1 m1) . B;  out_1 += conj(T x0 m0) . B;
//   B = T x0 m0 x1 m1;  out_2 += conj(T x3 m3) . B;  out_3 += conj(T x2 m2) . B;
// eight mode products and four Grams per half (the reference's schedule: twelve
// and four).  T_p is read again from global memory (L2) for each product chain.
#pragma once

#include "common.hpp"
#include "kernels_mfma.hpp"
#include "kernels_node.hpp"
#include "launch_grid.hpp"

namespace bqa {

// LDS position of half-tensor element idx: the low three bits (the last bond's index) are XORed with the
// next three, a permutation inside each aligned row of eight, so the fibres along the last bond, which one
// thread reads whole, fall on different banks across a wave; every other access pattern here reads whole
// aligned rows and stays conflict-free.
__device__ __forceinline__ int d4_sw(int idx) { return idx ^ ((idx >> 3) & 7); }

template <typename C>
struct D4Lds {
  static constexpr int H = 4096;                  // elements of one physical half
  static constexpr int BYTES = 2 * H * (int)sizeof(cplx<C>) + 4 * 64 * (int)sizeof(cplx<C>);
};

// In-place mode product along bond J (stride 8^(3-J) within a half): X[.., a, ..] <- sum_b m[a][b] X[.., b, ..].
// A thread owns 512 / NT fibres and updates them together, so each message entry read from LDS (a
// broadcast) serves all of them.
template <int J, int NT, typename C>
__device__ __forceinline__ void d4_mp(cplx<C>* X, const cplx<C>* __restrict__ m) {
  constexpr int st = J == 0 ? 512 : J == 1 ? 64 : J == 2 ? 8 : 1;
  constexpr int FPT = 512 / NT;
  int base[FPT];
  cplx<C> x[FPT][8];
#pragma unroll
  for (int u = 0; u < FPT; ++u) {
    const int f = threadIdx.x + NT * u;   // fibre
    base[u] = (f / st) * 8 * st + f % st;
#pragma unroll
    for (int b = 0; b < 8; ++b) x[u][b] = X[d4_sw(base[u] + b * st)];
  }
#pragma unroll
  for (int a = 0; a < 8; ++a) {
    cplx<C> acc[FPT];
#pragma unroll
    for (int u = 0; u < FPT; ++u) acc[u] = mk<C>(0, 0);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      const cplx<C> w = m[a * 8 + b];
#pragma unroll
      for (int u = 0; u < FPT; ++u) cfma(acc[u], w, x[u][b]);
    }
#pragma unroll
    for (int u = 0; u < FPT; ++u) X[d4_sw(base[u] + a * st)] = acc[u];
  }
}

// Gram along bond K: g += sum over this thread's fibre group r of conj(A[r, c]) B[r, e], entry (c, e) =
// (threadIdx % 64) / 8, % 8; the NT / 64 groups of 64 threads take every (NT / 64)-th fibre.
template <int K, int NT, typename C>
__device__ __forceinline__ void d4_gram(const cplx<C>* A, const cplx<C>* B, cplx<double>& g) {
  constexpr int st = K == 0 ? 512 : K == 1 ? 64 : K == 2 ? 8 : 1;
  const int o = threadIdx.x & 63, grp = threadIdx.x >> 6, c = o >> 3, e = o & 7;
  cplx<C> acc = mk<C>(0, 0);
#pragma unroll 8
  for (int r = grp; r < 512; r += NT / 64) {
    const int base = (r / st) * 8 * st + r % st;
    cfma_cj(acc, A[d4_sw(base + c * st)], B[d4_sw(base + e * st)]);
  }
  g = g + cast<double>(acc);
}

__device__ __forceinline__ void d4_sync() { __syncthreads(); }

// One physical half T_p of a node: the schedule above, adding the half's four Grams into g[k].
// M holds the four incoming messages; on return A and B are still being read by other threads'
// last Gram (the caller synchronises before reusing them).
template <int NT, typename R, typename C>
__device__ __forceinline__ void d4_half(cplx<C>* A, cplx<C>* B, const cplx<C>* M, const cplx<R>* __restrict__ Tp,
                                        cplx<double> (&g)[4]) {
  constexpr int H = D4Lds<C>::H;
  const int t = threadIdx.x;
  auto load = [&](cplx<C>* X) {
#pragma unroll
    for (int q = 0; q < H / NT; ++q) X[d4_sw(t + NT * q)] = cast<C>(Tp[t + NT * q]);
  };
  load(B);
  d4_sync();
  d4_mp<2, NT, C>(B, M + 2 * 64);
  d4_sync();
  d4_mp<3, NT, C>(B, M + 3 * 64);   // B = T x2 m2 x3 m3
  load(A);
  d4_sync();
  d4_mp<1, NT, C>(A, M + 1 * 64);   // A = T x1 m1
  d4_sync();
  d4_gram<0, NT, C>(A, B, g[0]);
  d4_sync();
  load(A);
  d4_sync();
  d4_mp<0, NT, C>(A, M + 0 * 64);   // A = T x0 m0
  d4_sync();
  d4_gram<1, NT, C>(A, B, g[1]);
  d4_sync();
  load(B);
  d4_sync();
  d4_mp<0, NT, C>(B, M + 0 * 64);
  d4_sync();
  d4_mp<1, NT, C>(B, M + 1 * 64);   // B = T x0 m0 x1 m1
  load(A);
  d4_sync();
  d4_mp<3, NT, C>(A, M + 3 * 64);   // A = T x3 m3
  d4_sync();
  d4_gram<2, NT, C>(A, B, g[2]);
  d4_sync();
  load(A);
  d4_sync();
  d4_mp<2, NT, C>(A, M + 2 * 64);   // A = T x2 m2
  d4_sync();
  d4_gram<3, NT, C>(A, B, g[3]);
}

// NT threads per node (256 or 512: one or two waves per SIMD, the node filling the CU's LDS in FP64).
template <typename R, int NT>
__global__ __launch_bounds__(NT) void k_bp_d4c8(ClassView cv, const cplx<R>* __restrict__ T,
                                                cplx<R>* msgA, cplx<R>* msgB, BpStatus* st) {
  static_assert(NT == 256 || NT == 512, "four or eight waves per node");
  using C = R;
  constexpr int H = D4Lds<C>::H;
  extern __shared__ __align__(16) unsigned char d4_lds[];
  cplx<C>* A = reinterpret_cast<cplx<C>*>(d4_lds);
  cplx<C>* B = A + H;
  cplx<C>* M = B + H;  // the four incoming messages, 64 entries each
  const int cur = bp_begin(st, cv.sweep);
  if (cur < 0) return;
  const cplx<R>* min = cur ? msgB : msgA;
  cplx<R>* mout = cur ? msgA : msgB;
  const int n = cv.n;
  const int t = threadIdx.x, o = t & 63;
  double dnum = 0.0, dden = 0.0;
  for (int i = blockIdx.x; i < n; i += gridDim.x) {
    d4_sync();  // the previous node's last readers of A, B and M are done
    if (t < 256) {
      const int j = t >> 6;
      M[j * 64 + o] = cast<C>(min[(size_t)cv.in_pos[j * n + i] * 64 + o]);
    }
    cplx<double> g[4] = {mk(0.0, 0.0), mk(0.0, 0.0), mk(0.0, 0.0), mk(0.0, 0.0)};
#pragma unroll 1
    for (int p = 0; p < 2; ++p) {
      if (p) d4_sync();  // the last Gram of the first half has read B
      d4_half<NT, R, C>(A, B, M, T + (size_t)i * 2 * H + (size_t)p * H, g);
    }
    // partial Grams of the four fibre groups into the dead A buffer (FP64), then one wave finishes
    d4_sync();
    constexpr int NG = NT / 64;  // fibre groups
    cplx<double>* red = reinterpret_cast<cplx<double>*>(A);  // [k][grp][64]: at most 32 KB <= the A buffer
#pragma unroll
    for (int k = 0; k < 4; ++k) red[(k * NG + (t >> 6)) * 64 + o] = g[k];
    d4_sync();
    if (t < 64) {
      const int c = o >> 3, e = o & 7;
#pragma unroll 1
      for (int k = 0; k < 4; ++k) {
        auto entry = [&](int x) {
          cplx<double> s = mk(0.0, 0.0);
#pragma unroll
          for (int q = 0; q < NG; ++q) s = s + red[(k * NG + q) * 64 + x];
          return s;
        };
        double tr = 0.0;
#pragma unroll
        for (int d = 0; d < 8; ++d) tr += entry(d * 9).x;
        const double inv = 1.0 / tr;
        const cplx<double> a = entry(o), b = entry(e * 8 + c);
        const cplx<R> v = cast<R>(mk(0.5 * (a.x + b.x) * inv, 0.5 * (a.y - b.y) * inv));
        const size_t pos = (size_t)cv.out_pos[k * n + i];
        const cplx<double> vd = cast<double>(v), od = cast<double>(min[pos * 64 + o]);
        dnum = fmax(dnum, abs2(vd - od));
        dden = fmax(dden, abs2(vd + od));
        mout[pos * 64 + o] = v;
      }
    }
  }
  if (t < 64) {
    for (int s = 32; s > 0; s >>= 1) {
      dnum = fmax(dnum, __shfl_xor(dnum, s, 64));
      dden = fmax(dden, __shfl_xor(dden, s, 64));
    }
    if (t == 0) bp_end(st, cv.sweep, sqrt(dnum), sqrt(dden));
  }
}

template <typename R, int NT>
inline void launch_bp_d4c8(ClassView cv, const cplx<R>* T, cplx<R>* mA, cplx<R>* mB, BpStatus* st, hipStream_t s) {
  constexpr int LDS = D4Lds<R>::BYTES;
  const int grid = node_grid_lds<k_bp_d4c8<R, NT>, NT, LDS>(cv.n, 1);
  hipLaunchKernelGGL((k_bp_d4c8<R, NT>), dim3(grid), dim3(NT), LDS, s, cv, T, mA, mB, st);
}


// ---------------------------------------------------------------------------
// The same sweep for FP32 jobs on v_mfma_f32_16x16x4f32 (k_bp_d4c8m; four waves per node, same
// schedule, LDS layout and swizzle).  A mode product along bond J is, per tile of 16 fibres, the
// real GEMM [Yr | Yi] = [Xr | Xi] [[m_r^T, m_i^T], [-m_i^T, m_r^T]] (K = 16: the bond index with
// its component; N = 16: the output index with its component), four k-steps; a Gram along bond K
// is one 16 x 16 tile D[(comp, c)][(comp', e)] = sum over fibres of A[f][c].comp B[f][e].comp'
// (K = the 512 fibres of the half, four per k-step), from which
//   Gram[c][e] = D[(0,c)][(0,e)] + D[(1,c)][(1,e)] + i (D[(0,c)][(1,e)] - D[(1,c)][(0,e)]).
// Operand lane layouts: A[i][k] at lane i + 16 k, B[k][j] at lane j + 16 k, D[4 (l / 16) + v][l % 16]
// in result register v of lane l.  Mode products read a tile's operands, then write its results in
// place (a wave's tiles are disjoint fibres); each wave's Gram partials stay in FP32 MFMA chains for
// one half and are added in FP64.

// In-place mode product along bond J with message m (64 complex entries, m[a][b] at a * 8 + b).
template <int J>
__device__ __forceinline__ void d4m_mp(cplx<float>* X, const cplx<float>* m) {
  constexpr int st = J == 0 ? 512 : J == 1 ? 64 : J == 2 ? 8 : 1;
  constexpr int TPW = 8;  // 32 tiles of 16 fibres per half, four waves
  // the lane's indices through opaque(): its addresses are formed here, not hoisted out of the node loop
  const int tid = opaque((int)threadIdx.x);
  const int l = tid & 63, w = tid >> 6, i16 = l & 15, kq = l >> 4;
  const int a = i16 & 7, co = i16 >> 3;
  float bm[4];
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    const int k = 4 * s + kq, ci = k >> 3, b = k & 7;
    const cplx<float> mv = m[a * 8 + b];
    bm[s] = ci == co ? mv.x : (ci ? -mv.y : mv.y);
  }
  float xa[TPW][4];
#pragma unroll
  for (int t = 0; t < TPW; ++t) {
    const int f = (w * TPW + t) * 16 + i16;
    const int base = (f / st) * 8 * st + f % st;
    const cplx<float> x0 = X[d4_sw(base + kq * st)], x1 = X[d4_sw(base + (kq + 4) * st)];
    xa[t][0] = x0.x;
    xa[t][1] = x1.x;
    xa[t][2] = x0.y;
    xa[t][3] = x1.y;
  }
  f4v acc[TPW];
#pragma unroll
  for (int t = 0; t < TPW; ++t) acc[t] = f4v{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int s = 0; s < 4; ++s)
#pragma unroll
    for (int t = 0; t < TPW; ++t) acc[t] = mfma4(xa[t][s], bm[s], acc[t]);
  float* Xf = reinterpret_cast<float*>(X);
#pragma unroll
  for (int t = 0; t < TPW; ++t)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int f = (w * TPW + t) * 16 + 4 * kq + v;
      const int base = (f / st) * 8 * st + f % st;
      Xf[2 * d4_sw(base + a * st) + co] = acc[t][v];
    }
}

// This wave's k-steps of the Gram tile along bond K, added (FP64) into g[v] = D[4 (l / 16) + v][l % 16].
template <int K>
__device__ __forceinline__ void d4m_gram(const cplx<float>* A, const cplx<float>* B, double (&g)[4]) {
  constexpr int st = K == 0 ? 512 : K == 1 ? 64 : K == 2 ? 8 : 1;
  constexpr int KPW = 32;  // 128 k-steps of four fibres per half, four waves
  const int tid = opaque((int)threadIdx.x);
  const int l = tid & 63, w = tid >> 6, x = l & 7, comp = (l >> 3) & 1, kq = l >> 4;
  f4v acc[4] = {f4v{0.f, 0.f, 0.f, 0.f}, f4v{0.f, 0.f, 0.f, 0.f}, f4v{0.f, 0.f, 0.f, 0.f},
                f4v{0.f, 0.f, 0.f, 0.f}};
#pragma unroll 4
  for (int s = 0; s < KPW; ++s) {
    const int f = 4 * (w * KPW + s) + kq;
    const int idx = d4_sw((f / st) * 8 * st + f % st + x * st);
    const cplx<float> av = A[idx], bv = B[idx];
    acc[s & 3] = mfma4(comp ? av.y : av.x, comp ? bv.y : bv.x, acc[s & 3]);
  }
#pragma unroll
  for (int v = 0; v < 4; ++v)
    g[v] += ((double)acc[0][v] + (double)acc[1][v]) + ((double)acc[2][v] + (double)acc[3][v]);
}

// One physical half: d4_half's schedule on the MFMA products.
__device__ __forceinline__ void d4m_half(cplx<float>* A, cplx<float>* B, const cplx<float>* M,
                                         const cplx<float>* __restrict__ Tp, double (&g)[4][4]) {
  constexpr int H = D4Lds<float>::H, NT = 256;
  auto load = [&](cplx<float>* X) {
    const int t = opaque((int)threadIdx.x);
#pragma unroll
    for (int q = 0; q < H / NT; ++q) X[d4_sw(t + NT * q)] = Tp[t + NT * q];
  };
  load(B);
  d4_sync();
  d4m_mp<2>(B, M + 2 * 64);
  d4_sync();
  d4m_mp<3>(B, M + 3 * 64);  // B = T x2 m2 x3 m3
  load(A);
  d4_sync();
  d4m_mp<1>(A, M + 1 * 64);  // A = T x1 m1
  d4_sync();
  d4m_gram<0>(A, B, g[0]);
  d4_sync();
  load(A);
  d4_sync();
  d4m_mp<0>(A, M + 0 * 64);  // A = T x0 m0
  d4_sync();
  d4m_gram<1>(A, B, g[1]);
  d4_sync();
  load(B);
  d4_sync();
  d4m_mp<0>(B, M + 0 * 64);
  d4_sync();
  d4m_mp<1>(B, M + 1 * 64);  // B = T x0 m0 x1 m1
  load(A);
  d4_sync();
  d4m_mp<3>(A, M + 3 * 64);  // A = T x3 m3
  d4_sync();
  d4m_gram<2>(A, B, g[2]);
  d4_sync();
  load(A);
  d4_sync();
  d4m_mp<2>(A, M + 2 * 64);  // A = T x2 m2
  d4_sync();
  d4m_gram<3>(A, B, g[3]);
}

// The Gram of d4m_gram for the extended messages of FP32 jobs: the FP32 operands widened on load
// and multiplied and summed on v_mfma_f64_16x16x4f64 (exact products, FP64 sums), so
// g[v] = D[l / 16 + 4 v][l % 16] (the FP64 result layout).
template <int K>
__device__ __forceinline__ void d4m_gram_w(const cplx<float>* A, const cplx<float>* B, double (&g)[4]) {
  constexpr int st = K == 0 ? 512 : K == 1 ? 64 : K == 2 ? 8 : 1;
  constexpr int KPW = 32;  // 128 k-steps of four fibres per half, four waves
  const int tid = opaque((int)threadIdx.x);
  const int l = tid & 63, w = tid >> 6, x = l & 7, comp = (l >> 3) & 1, kq = l >> 4;
  const float* Af = reinterpret_cast<const float*>(A);
  const float* Bf = reinterpret_cast<const float*>(B);
  d4v acc[4] = {d4v{0.0, 0.0, 0.0, 0.0}, d4v{0.0, 0.0, 0.0, 0.0}, d4v{0.0, 0.0, 0.0, 0.0},
                d4v{0.0, 0.0, 0.0, 0.0}};
#pragma unroll 4
  for (int s = 0; s < KPW; ++s) {
    const int f = 4 * (w * KPW + s) + kq;
    const int idx = 2 * d4_sw((f / st) * 8 * st + f % st + x * st) + comp;
    acc[s & 3] = mfma4d((double)Af[idx], (double)Bf[idx], acc[s & 3]);
  }
#pragma unroll
  for (int v = 0; v < 4; ++v) g[v] += (acc[0][v] + acc[1][v]) + (acc[2][v] + acc[3][v]);
}

// One physical half for the extended messages of FP32 jobs: d4m_half's FP32 MFMA mode products with
// the FP64 MFMA Grams of d4m_gram_w.
__device__ __forceinline__ void d4m_ext_half(cplx<float>* A, cplx<float>* B, const cplx<float>* M,
                                             const cplx<float>* __restrict__ Tp, double (&g)[4][4]) {
  constexpr int H = D4Lds<float>::H, NT = 256;
  auto load = [&](cplx<float>* X) {
    const int t = opaque((int)threadIdx.x);
#pragma unroll
    for (int q = 0; q < H / NT; ++q) X[d4_sw(t + NT * q)] = Tp[t + NT * q];
  };
  load(B);
  d4_sync();
  d4m_mp<2>(B, M + 2 * 64);
  d4_sync();
  d4m_mp<3>(B, M + 3 * 64);  // B = T x2 m2 x3 m3
  load(A);
  d4_sync();
  d4m_mp<1>(A, M + 1 * 64);  // A = T x1 m1
  d4_sync();
  d4m_gram_w<0>(A, B, g[0]);
  d4_sync();
  load(A);
  d4_sync();
  d4m_mp<0>(A, M + 0 * 64);  // A = T x0 m0
  d4_sync();
  d4m_gram_w<1>(A, B, g[1]);
  d4_sync();
  load(B);
  d4_sync();
  d4m_mp<0>(B, M + 0 * 64);
  d4_sync();
  d4m_mp<1>(B, M + 1 * 64);  // B = T x0 m0 x1 m1
  load(A);
  d4_sync();
  d4m_mp<3>(A, M + 3 * 64);  // A = T x3 m3
  d4_sync();
  d4m_gram_w<2>(A, B, g[2]);
  d4_sync();
  load(A);
  d4_sync();
  d4m_mp<2>(A, M + 2 * 64);  // A = T x2 m2
  d4_sync();
  d4m_gram_w<3>(A, B, g[3]);
}

__global__ __launch_bounds__(256, 2) void k_bp_d4c8m(ClassView cv, const cplx<float>* __restrict__ T,
                                                  cplx<float>* msgA, cplx<float>* msgB, BpStatus* st) {
  constexpr int H = D4Lds<float>::H, NW = 4;
  extern __shared__ __align__(16) unsigned char d4_lds[];
  cplx<float>* A = reinterpret_cast<cplx<float>*>(d4_lds);
  cplx<float>* B = A + H;
  cplx<float>* M = B + H;
  const int cur = bp_begin(st, cv.sweep);
  if (cur < 0) return;
  const cplx<float>* min = cur ? msgB : msgA;
  cplx<float>* mout = cur ? msgA : msgB;
  const int n = cv.n;
  const int t = threadIdx.x, o = t & 63;
  double dnum = 0.0, dden = 0.0;
  for (int i = blockIdx.x; i < n; i += gridDim.x) {
    d4_sync();  // the previous node's last readers of A, B and M are done
    M[(t >> 6) * 64 + o] = min[(size_t)cv.in_pos[(t >> 6) * n + i] * 64 + o];
    double g[4][4];
#pragma unroll
    for (int k = 0; k < 4; ++k)
#pragma unroll
      for (int v = 0; v < 4; ++v) g[k][v] = 0.0;
#pragma unroll 1
    for (int p = 0; p < 2; ++p) {
      if (p) d4_sync();  // the last Gram of the first half has read B
      d4m_half(A, B, M, T + (size_t)i * 2 * H + (size_t)p * H, g);
    }
    // the waves' Gram tiles into the dead A buffer (FP64, [k][wave][256]: 32 KB), then one wave finishes
    d4_sync();
    double* red = reinterpret_cast<double*>(A);
#pragma unroll
    for (int k = 0; k < 4; ++k)
#pragma unroll
      for (int v = 0; v < 4; ++v) red[(k * NW + (t >> 6)) * 256 + (4 * (o >> 4) + v) * 16 + (o & 15)] = g[k][v];
    d4_sync();
    if (t < 64) {
      const int c = o >> 3, e = o & 7;
#pragma unroll 1
      for (int k = 0; k < 4; ++k) {
        auto dsum = [&](int r, int col) {
          double s = 0.0;
#pragma unroll
          for (int q = 0; q < NW; ++q) s += red[(k * NW + q) * 256 + r * 16 + col];
          return s;
        };
        auto entry = [&](int cc, int ee) {
          return mk(dsum(cc, ee) + dsum(8 + cc, 8 + ee), dsum(cc, 8 + ee) - dsum(8 + cc, ee));
        };
        double tr = 0.0;
#pragma unroll
        for (int d = 0; d < 8; ++d) tr += dsum(d, d) + dsum(8 + d, 8 + d);
        const double inv = 1.0 / tr;
        const cplx<double> a = entry(c, e), b = entry(e, c);
        const cplx<float> v = cast<float>(mk(0.5 * (a.x + b.x) * inv, 0.5 * (a.y - b.y) * inv));
        const size_t pos = (size_t)cv.out_pos[k * n + i];
        const cplx<double> vd = cast<double>(v), od = cast<double>(min[pos * 64 + o]);
        dnum = fmax(dnum, abs2(vd - od));
        dden = fmax(dden, abs2(vd + od));
        mout[pos * 64 + o] = v;
      }
    }
  }
  if (t < 64) {
    for (int s = 32; s > 0; s >>= 1) {
      dnum = fmax(dnum, __shfl_xor(dnum, s, 64));
      dden = fmax(dden, __shfl_xor(dden, s, 64));
    }
    if (t == 0) bp_end(st, cv.sweep, sqrt(dnum), sqrt(dden));
  }
}

inline void launch_bp_d4c8m(ClassView cv, const cplx<float>* T, cplx<float>* mA, cplx<float>* mB, BpStatus* st,
                            hipStream_t s) {
  constexpr int LDS = D4Lds<float>::BYTES;
  const int grid = node_grid_lds<k_bp_d4c8m, 256, LDS>(cv.n, 1);
  hipLaunchKernelGGL(k_bp_d4c8m, dim3(grid), dim3(256), LDS, s, cv, T, mA, mB, st);
}

// ---------------------------------------------------------------------------
// The same sweep for FP64 jobs on v_mfma_f64_16x16x4f64 (k_bp_d4c8m64<NW>; NW waves per node, the
// schedule, LDS layout and swizzle of k_bp_d4c8m).  The A and B operand lane layouts are those of
// the FP32 instruction; result register v of lane l holds D[l / 16 + 4 v][l % 16].  The two FP64
// halves take 128 KB of LDS, so one block runs per CU; NW = 8 gives two waves per SIMD.

// In-place mode product along bond J with message m (64 complex entries, m[a][b] at a * 8 + b).
template <int J, int NW>
__device__ __forceinline__ void d4m64_mp(cplx<double>* X, const cplx<double>* m) {
  constexpr int st = J == 0 ? 512 : J == 1 ? 64 : J == 2 ? 8 : 1;
  constexpr int TPW = 32 / NW;  // 32 tiles of 16 fibres per half
  const int tid = opaque((int)threadIdx.x);
  const int l = tid & 63, w = tid >> 6, i16 = l & 15, kq = l >> 4;
  const int a = i16 & 7, co = i16 >> 3;
  double bm[4];
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    const int k = 4 * s + kq, ci = k >> 3, b = k & 7;
    const cplx<double> mv = m[a * 8 + b];
    bm[s] = ci == co ? mv.x : (ci ? -mv.y : mv.y);
  }
  double xa[TPW][4];
#pragma unroll
  for (int t = 0; t < TPW; ++t) {
    const int f = (w * TPW + t) * 16 + i16;
    const int base = (f / st) * 8 * st + f % st;
    const cplx<double> x0 = X[d4_sw(base + kq * st)], x1 = X[d4_sw(base + (kq + 4) * st)];
    xa[t][0] = x0.x;
    xa[t][1] = x1.x;
    xa[t][2] = x0.y;
    xa[t][3] = x1.y;
  }
  d4v acc[TPW];
#pragma unroll
  for (int t = 0; t < TPW; ++t) acc[t] = d4v{0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int s = 0; s < 4; ++s)
#pragma unroll
    for (int t = 0; t < TPW; ++t) acc[t] = mfma4d(xa[t][s], bm[s], acc[t]);
  double* Xd = reinterpret_cast<double*>(X);
#pragma unroll
  for (int t = 0; t < TPW; ++t)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int f = (w * TPW + t) * 16 + kq + 4 * v;
      const int base = (f / st) * 8 * st + f % st;
      Xd[2 * d4_sw(base + a * st) + co] = acc[t][v];
    }
}

// This wave's k-steps of the Gram tile along bond K, added into g[v] = D[l / 16 + 4 v][l % 16]
// (D[(comp, c)][(comp', e)] = sum over fibres of A[f][c].comp B[f][e].comp').
template <int K, int NW>
__device__ __forceinline__ void d4m64_gram(const cplx<double>* A, const cplx<double>* B, double (&g)[4]) {
  constexpr int st = K == 0 ? 512 : K == 1 ? 64 : K == 2 ? 8 : 1;
  constexpr int KPW = 128 / NW;  // 128 k-steps of four fibres per half
  const int tid = opaque((int)threadIdx.x);
  const int l = tid & 63, w = tid >> 6, x = l & 7, comp = (l >> 3) & 1, kq = l >> 4;
  const double* Ad = reinterpret_cast<const double*>(A);
  const double* Bd = reinterpret_cast<const double*>(B);
  d4v acc[4] = {d4v{0.0, 0.0, 0.0, 0.0}, d4v{0.0, 0.0, 0.0, 0.0}, d4v{0.0, 0.0, 0.0, 0.0},
                d4v{0.0, 0.0, 0.0, 0.0}};
#pragma unroll 4
  for (int s = 0; s < KPW; ++s) {
    const int f = 4 * (w * KPW + s) + kq;
    const int idx = 2 * d4_sw((f / st) * 8 * st + f % st + x * st) + comp;
    acc[s & 3] = mfma4d(Ad[idx], Bd[idx], acc[s & 3]);
  }
#pragma unroll
  for (int v = 0; v < 4; ++v) g[v] += (acc[0][v] + acc[1][v]) + (acc[2][v] + acc[3][v]);
}

// One physical half: d4_half's schedule on the FP64 MFMA products.
template <int NW>
__device__ __forceinline__ void d4m64_half(cplx<double>* A, cplx<double>* B, const cplx<double>* M,
                                           const cplx<double>* __restrict__ Tp, double (&g)[4][4]) {
  constexpr int H = D4Lds<double>::H, NT = 64 * NW;
  auto load = [&](cplx<double>* X) {
    const int t = opaque((int)threadIdx.x);
#pragma unroll
    for (int q = 0; q < H / NT; ++q) X[d4_sw(t + NT * q)] = Tp[t + NT * q];
  };
  load(B);
  d4_sync();
  d4m64_mp<2, NW>(B, M + 2 * 64);
  d4_sync();
  d4m64_mp<3, NW>(B, M + 3 * 64);  // B = T x2 m2 x3 m3
  load(A);
  d4_sync();
  d4m64_mp<1, NW>(A, M + 1 * 64);  // A = T x1 m1
  d4_sync();
  d4m64_gram<0, NW>(A, B, g[0]);
  d4_sync();
  load(A);
  d4_sync();
  d4m64_mp<0, NW>(A, M + 0 * 64);  // A = T x0 m0
  d4_sync();
  d4m64_gram<1, NW>(A, B, g[1]);
  d4_sync();
  load(B);
  d4_sync();
  d4m64_mp<0, NW>(B, M + 0 * 64);
  d4_sync();
  d4m64_mp<1, NW>(B, M + 1 * 64);  // B = T x0 m0 x1 m1
  load(A);
  d4_sync();
  d4m64_mp<3, NW>(A, M + 3 * 64);  // A = T x3 m3
  d4_sync();
  d4m64_gram<2, NW>(A, B, g[2]);
  d4_sync();
  load(A);
  d4_sync();
  d4m64_mp<2, NW>(A, M + 2 * 64);  // A = T x2 m2
  d4_sync();
  d4m64_gram<3, NW>(A, B, g[3]);
}

template <int NW>
__global__ __launch_bounds__(64 * NW) void k_bp_d4c8m64(ClassView cv, const cplx<double>* __restrict__ T,
                                                        cplx<double>* msgA, cplx<double>* msgB, BpStatus* st) {
  static_assert(NW == 4 || NW == 8, "four or eight waves per node");
  constexpr int H = D4Lds<double>::H;
  extern __shared__ __align__(16) unsigned char d4_lds[];
  cplx<double>* A = reinterpret_cast<cplx<double>*>(d4_lds);
  cplx<double>* B = A + H;
  cplx<double>* M = B + H;
  const int cur = bp_begin(st, cv.sweep);
  if (cur < 0) return;
  const cplx<double>* min = cur ? msgB : msgA;
  cplx<double>* mout = cur ? msgA : msgB;
  const int n = cv.n;
  const int t = threadIdx.x, o = t & 63;
  double dnum = 0.0, dden = 0.0;
  for (int i = blockIdx.x; i < n; i += gridDim.x) {
    d4_sync();  // the previous node's last readers of A, B and M are done
    if (t < 256) M[(t >> 6) * 64 + o] = min[(size_t)cv.in_pos[(t >> 6) * n + i] * 64 + o];
    double g[4][4];
#pragma unroll
    for (int k = 0; k < 4; ++k)
#pragma unroll
      for (int v = 0; v < 4; ++v) g[k][v] = 0.0;
#pragma unroll 1
    for (int p = 0; p < 2; ++p) {
      if (p) d4_sync();  // the last Gram of the first half has read B
      d4m64_half<NW>(A, B, M, T + (size_t)i * 2 * H + (size_t)p * H, g);
    }
    // the waves' Gram tiles into the dead A buffer ([k][wave][256] doubles: 32 or 64 KB), then one
    // wave finishes
    d4_sync();
    double* red = reinterpret_cast<double*>(A);
#pragma unroll
    for (int k = 0; k < 4; ++k)
#pragma unroll
      for (int v = 0; v < 4; ++v) red[(k * NW + (t >> 6)) * 256 + ((o >> 4) + 4 * v) * 16 + (o & 15)] = g[k][v];
    d4_sync();
    if (t < 64) {
      const int c = o >> 3, e = o & 7;
#pragma unroll 1
      for (int k = 0; k < 4; ++k) {
        auto dsum = [&](int r, int col) {
          double s = 0.0;
#pragma unroll
          for (int q = 0; q < NW; ++q) s += red[(k * NW + q) * 256 + r * 16 + col];
          return s;
        };
        auto entry = [&](int cc, int ee) {
          return mk(dsum(cc, ee) + dsum(8 + cc, 8 + ee), dsum(cc, 8 + ee) - dsum(8 + cc, ee));
        };
        double tr = 0.0;
#pragma unroll
        for (int d = 0; d < 8; ++d) tr += dsum(d, d) + dsum(8 + d, 8 + d);
        const double inv = 1.0 / tr;
        const cplx<double> a = entry(c, e), b = entry(e, c);
        const cplx<double> v = mk(0.5 * (a.x + b.x) * inv, 0.5 * (a.y - b.y) * inv);
        const size_t pos = (size_t)cv.out_pos[k * n + i];
        const cplx<double> od = min[pos * 64 + o];
        dnum = fmax(dnum, abs2(v - od));
        dden = fmax(dden, abs2(v + od));
        mout[pos * 64 + o] = v;
      }
    }
  }
  if (t < 64) {
    for (int s = 32; s > 0; s >>= 1) {
      dnum = fmax(dnum, __shfl_xor(dnum, s, 64));
      dden = fmax(dden, __shfl_xor(dden, s, 64));
    }
    if (t == 0) bp_end(st, cv.sweep, sqrt(dnum), sqrt(dden));
  }
}

template <int NW>
inline void launch_bp_d4c8m64(ClassView cv, const cplx<double>* T, cplx<double>* mA, cplx<double>* mB,
                              BpStatus* st, hipStream_t s) {
  constexpr int LDS = D4Lds<double>::BYTES;
  const int grid = node_grid_lds<k_bp_d4c8m64<NW>, 64 * NW, LDS>(cv.n, 1);
  hipLaunchKernelGGL((k_bp_d4c8m64<NW>), dim3(grid), dim3(64 * NW), LDS, s, cv, T, mA, mB, st);
}

// Extended messages of degree-4 nodes at chi = 8 (state.py:127-139, backends.py:393-404,519-526) on
// MFMA (k_ext_d4c8m<R, NW>): the BP kernels' products per physical half p (FP64 jobs: d4m64_half;
// FP32 jobs, four waves: d4m_ext_half, FP32 mode products and FP64 Grams), with the Grams kept per
// half.  The waves' 16 x 16 FP64 D tiles are reduced over the waves into complex 8 x 8 Grams G_p[k]
// at the start of A, and each 16 x 16 extended message is assembled with the bond's ZZ-split gates
// g0, g1: ext[(s,c),(s',e)] = conj(g_s) g_s' (G_0 +- G_1)[c][e] / (trace * (|g0|^2 + |g1|^2)), with
// + for s + s' even.  FP32 jobs run two blocks per CU (66 KB of LDS each; held to 256
// registers, a spill of about 280 bytes per lane: 100x100 grid step 14.5 -> 13.7 ms against one
// block per CU without the spill).
template <typename R, int NW>
__global__ __launch_bounds__(64 * NW, sizeof(R) == sizeof(float) ? 2 : 1) void k_ext_d4c8m(ClassView cv, const cplx<R>* __restrict__ T,
                                                       const cplx<R>* __restrict__ msgs, cplx<double>* ext_out) {
  static_assert(NW == 4 || (NW == 8 && sizeof(R) == sizeof(double)), "four waves (or eight for FP64 jobs)");
  constexpr int H = D4Lds<R>::H, NT = 64 * NW;
  // the D tiles [k][p][wave][256] of the reduction fill the A and B buffers at most
  static_assert(4 * 2 * NW * 256 * (int)sizeof(double) <= 2 * H * (int)sizeof(cplx<R>), "reduction fits A and B");
  extern __shared__ __align__(16) unsigned char d4_lds[];
  cplx<R>* A = reinterpret_cast<cplx<R>*>(d4_lds);
  cplx<R>* B = A + H;
  cplx<R>* M = B + H;
  const int n = cv.n;
  const int t = threadIdx.x, o = t & 63;
  for (int i = blockIdx.x; i < n; i += gridDim.x) {
    d4_sync();  // the previous node's last readers of A, B and M are done
    if (t < 256) M[(t >> 6) * 64 + o] = msgs[(size_t)cv.in_pos[(t >> 6) * n + i] * 64 + o];
    double g[2][4][4];
#pragma unroll
    for (int p = 0; p < 2; ++p)
#pragma unroll
      for (int k = 0; k < 4; ++k)
#pragma unroll
        for (int v = 0; v < 4; ++v) g[p][k][v] = 0.0;
#pragma unroll
    for (int p = 0; p < 2; ++p) {  // unrolled: g[p] stays in registers
      if (p) d4_sync();  // the last Gram of the first half has read B
      if constexpr (sizeof(R) == sizeof(double))
        d4m64_half<NW>(A, B, M, T + (size_t)i * 2 * H + (size_t)p * H, g[p]);
      else
        d4m_ext_half(A, B, M, T + (size_t)i * 2 * H + (size_t)p * H, g[p]);
    }
    d4_sync();
    // the D tiles [k][p][wave][256] (16 NW KB, in the dead A and B buffers, contiguous)
    double* red = reinterpret_cast<double*>(A);
#pragma unroll
    for (int k = 0; k < 4; ++k)
#pragma unroll
      for (int p = 0; p < 2; ++p)
#pragma unroll
        for (int v = 0; v < 4; ++v)
          red[((k * 2 + p) * NW + (t >> 6)) * 256 + ((o >> 4) + 4 * v) * 16 + (o & 15)] = g[p][k][v];
    d4_sync();
    // complex Gram entry (c, e) of bond k, half p: D[(0,c)][(0,e)] + D[(1,c)][(1,e)]
    // + i (D[(0,c)][(1,e)] - D[(1,c)][(0,e)]), summed over the waves
    constexpr int PER = 512 / NT;
    cplx<double> ge[PER];
#pragma unroll
    for (int u = 0; u < PER; ++u) {
      const int x = t + NT * u, kp = x >> 6, c = (x >> 3) & 7, e = x & 7;
      auto dsum = [&](int r, int col) {
        double sum = 0.0;
#pragma unroll
        for (int q = 0; q < NW; ++q) sum += red[(kp * NW + q) * 256 + r * 16 + col];
        return sum;
      };
      ge[u] = mk(dsum(c, e) + dsum(8 + c, 8 + e), dsum(c, 8 + e) - dsum(8 + c, e));
    }
    d4_sync();
    cplx<double>* gm = reinterpret_cast<cplx<double>*>(A);  // [k][p][64]
#pragma unroll
    for (int u = 0; u < PER; ++u) gm[t + NT * u] = ge[u];
    d4_sync();
    for (int x = t; x < 4 * 256; x += NT) {
      const int k = x >> 8, e16 = x & 255, row = e16 >> 4, col = e16 & 15;
      const int s = row >> 3, c = row & 7, s2 = col >> 3, e = col & 7;
      cplx<double> tr = mk(0.0, 0.0);
#pragma unroll
      for (int d = 0; d < 8; ++d) tr = tr + gm[(k * 2) * 64 + d * 9] + gm[(k * 2 + 1) * 64 + d * 9];
      const cplx<double>* zg = cv.zz + 2 * (size_t)cv.lmbd_pos[k * n + i];
      const cplx<double> g0 = zg[0], g1 = zg[1];
      const cplx<double> inv = cinv(tr * mk(abs2(g0) + abs2(g1), 0.0));
      const cplx<double> a0 = gm[(k * 2) * 64 + c * 8 + e], a1 = gm[(k * 2 + 1) * 64 + c * 8 + e];
      const cplx<double> gs = ((s + s2) & 1) ? a0 - a1 : a0 + a1;
      const cplx<double> ga = s ? g1 : g0, gb = s2 ? g1 : g0;
      ext_out[(size_t)cv.out_pos[k * n + i] * 256 + e16] = conj(ga) * gb * gs * inv;
    }
  }
}

template <typename R, int NW>
inline void launch_ext_d4c8m(ClassView cv, const cplx<R>* T, const cplx<R>* msgs, cplx<double>* ext, hipStream_t s) {
  constexpr int LDS = D4Lds<R>::BYTES;
  const int grid = node_grid_lds<k_ext_d4c8m<R, NW>, 64 * NW, LDS>(cv.n, 1);
  hipLaunchKernelGGL((k_ext_d4c8m<R, NW>), dim3(grid), dim3(64 * NW), LDS, s, cv, T, msgs, ext);
}

}  // namespace bqa
