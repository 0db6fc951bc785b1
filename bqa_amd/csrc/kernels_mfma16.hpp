// FP32 MFMA BP sweep for bond dimension chi = 16 and degree 2 or 3 (FP32 jobs).
//
// One block of W waves per node.  The node tensor T[p, b_1..b_d] (complex64,
// 64 KB at degree 3) and one work buffer live in LDS; the reference schedule
// (backends.py:381-404) runs with the mode products in place along their bond
// and T x_{d-1} m_{d-1} computed twice at degree 3, as in the chi = 8 kernel.
// Complex contractions are real GEMMs on v_mfma_f32_16x16x4f32 (A[i][k] at lane
// i + 16 k, B[k][j] at lane j + 16 k, D[4 (l / 16) + v][l % 16] in register v):
//
//   mode product along bond J: rows = 16 fibres, K = 32 = (component, b),
//     N = 32 = (component, a): [Yr | Yi] = [Xr | Xi] [[m_r^T, m_i^T], [-m_i^T, m_r^T]];
//   Gram with bond K free: K = fibres, rows (component, c) x columns (component, e),
//     out[c][e] = G00 + G11 + i (G01 - G10) of the four 16 x 16 tiles.
//
// LDS layout: complex element e = ((p 16 + b_1) 16 + b_2) 16 + b_3 (degree 3)
// is stored at sw16(e), which flips bit 0 of b_2 by bit 0 of b_1 and the four
// bits of b_3 by b_1 ^ b_2 (degree 2: the same with (p, b_1, b_2)).  The map is
// linear over GF(2).  With it, the 32 lanes of a half-wave touch 32 distinct
// complex slots mod 32 in every access below: mode-product operand reads along
// each bond, output writes (the 16 rows of a tile are ordered so that rows
// 4 kq + v and 4 (kq + 1) + v differ in parity) and Gram operand reads.
#pragma once

#include "common.hpp"
#include "kernels_mfma.hpp"
#include "kernels_node.hpp"

namespace bqa {

constexpr int cpow16(int e) { return e <= 0 ? 1 : 16 * cpow16(e - 1); }

template <int D>
struct Mf16 {
  static constexpr int A = cpow16(D);      // chi^d
  static constexpr int S = 2 * A;          // complex elements per node
  static constexpr int FIB = S / 16;       // fibres along one bond
  static constexpr int TILES = FIB / 16;   // 16-fibre tiles of a mode product
  static constexpr int KS = FIB / 4;       // k-steps of a Gram
};

template <int D>
__host__ __device__ constexpr int sw16(int e) {
  if (D == 3) {
    const int b1 = (e >> 8) & 15, b2 = (e >> 4) & 15;
    return e ^ ((b1 & 1) << 4) ^ ((b1 ^ b2) & 15);
  }
  const int p = (e >> 8) & 1, b1 = (e >> 4) & 15;
  return e ^ ((p & 1) << 4) ^ ((p ^ b1) & 15);
}

// First element of fibre f along bond J (a zero field inserted at J).
template <int D, int J>
__device__ __forceinline__ int fbase16(int f) {
  constexpr int st = cpow16(D - 1 - J);
  return (f / st) * 16 * st + (f % st);
}

// Row r of a mode-product tile -> fibre offset within the tile: rows 4 kq + v and
// 4 (kq + 1) + v land on fibres of different parity.
__device__ __forceinline__ constexpr int row_perm16(int r) { return ((r & 3) << 1) | ((r >> 2) & 1) | (r & 8); }

__device__ __forceinline__ f4v mfma4f(float a, float b, f4v c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

// B operands of message m (16 x 16 complex64) for a mode product: bm[s][ct] for k-step
// s (k = 4 s + kq) and column tile ct, from [[m_r^T, m_i^T], [-m_i^T, m_r^T]].
__device__ __forceinline__ void msg_operand16(const cplx<float>* __restrict__ m, float (&bm)[8][2]) {
  const int l = threadIdx.x & 63, a = l & 15, kq = l >> 4;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const cplx<float> v = m[a * 16 + 4 * q + kq];   // b = 4 q + kq
    bm[q][0] = v.x;          // k < 16 (Re X): column Re Y -> m_r, column Im Y -> m_i
    bm[q][1] = v.y;
    bm[q + 4][0] = -v.y;     // k >= 16 (Im X): column Re Y -> -m_i, column Im Y -> m_r
    bm[q + 4][1] = v.x;
  }
}

__device__ __forceinline__ float2 lds_ld2(const float2* base, int byte_off) {
  return *reinterpret_cast<const float2*>(reinterpret_cast<const char*>(base) + byte_off);
}

__device__ __forceinline__ void lds_st2(float2* base, int byte_off, float2 v) {
  *reinterpret_cast<float2*>(reinterpret_cast<char*>(base) + byte_off) = v;
}

// Every LDS index below is built from disjoint bit fields (fibre fields and the element
// along the bond), and fbase16 and sw16 are linear over GF(2), so an address is
// sw16(lane part) ^ sw16(constant part): the lane part (with the wave's offset) is formed
// once per call and hidden from loop-invariant code motion by opaque(), the constant part
// folds at compile time, and each access costs one XOR.

// Y = X x_J m over this wave's tiles (float2 views of the swizzled LDS buffers); X may equal Y.
// HALF: X and Y hold one physical half (the first half of the fibres, p being their most
// significant index), so only its tiles are done.
template <int D, int W, int J, bool HALF = false>
__device__ __forceinline__ void mp16(const float2* X, float2* Y, const cplx<float>* __restrict__ m) {
  float bm[8][2];                           // loaded per product: the node's operands do not stay live
  msg_operand16(m, bm);
  constexpr int st = cpow16(D - 1 - J);
  constexpr int TPW = Mf16<D>::TILES / (HALF ? 2 : 1) / W;
  const int l = threadIdx.x & 63, i16 = l & 15, kq = l >> 4, w = threadIdx.x >> 6;
  const int wbase = sw16<D>(fbase16<D, J>(w * TPW * 16));
  // A reads: fibre row_perm16(i16), element kq along J
  const int rd = opaque(8 * (wbase ^ sw16<D>(fbase16<D, J>(row_perm16(i16)) ^ (kq * st))));
  // D writes: element i16 along J (the row's fibre is a per-(kq, v) part below)
  const int wr = opaque(8 * (wbase ^ sw16<D>(i16 * st)));
  const int kqr = 4 * kq;
  float a[TPW][8];
#pragma unroll
  for (int t = 0; t < TPW; ++t)
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const float2 v = lds_ld2(X, rd ^ (8 * sw16<D>(fbase16<D, J>(t * 16) ^ (4 * q * st))));
      a[t][q] = v.x;
      a[t][q + 4] = v.y;
    }
  f4v acc[TPW][2];
#pragma unroll
  for (int t = 0; t < TPW; ++t) acc[t][0] = acc[t][1] = f4v{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int s = 0; s < 8; ++s)
#pragma unroll
    for (int t = 0; t < TPW; ++t) {
      acc[t][0] = mfma4f(a[t][s], bm[s][0], acc[t][0]);
      acc[t][1] = mfma4f(a[t][s], bm[s][1], acc[t][1]);
    }
  // row 4 kq + v of tile t: fibre t * 16 + row_perm16(4 kq + v); row_perm16 is linear too,
  // row_perm16(4 kq + v) = row_perm16(4 kq) ^ row_perm16(v)
  const int rk = opaque(8 * sw16<D>(fbase16<D, J>(row_perm16(kqr))));
#pragma unroll
  for (int t = 0; t < TPW; ++t)
#pragma unroll
    for (int v = 0; v < 4; ++v)
      lds_st2(Y, wr ^ rk ^ (8 * sw16<D>(fbase16<D, J>(t * 16 + row_perm16(v)))),
              make_float2(acc[t][0][v], acc[t][1][v]));
}

// This wave's k-steps of out[c][e] = sum_f conj(X[f][c]) Y[f][e] over the fibres along bond K:
// lane (c = 4 kq + v, e = l % 16) gets its four entries, FP32 chains combined in FP64.
template <int D, int W, int K, bool HALF = false>
__device__ __forceinline__ void gram16(const float2* X, const float2* Y, double (&gre)[4], double (&gim)[4]) {
  constexpr int st = cpow16(D - 1 - K);
  constexpr int KPW = Mf16<D>::KS / (HALF ? 2 : 1) / W;
  constexpr int B = KPW < 4 ? KPW : 4;
  const int l = threadIdx.x & 63, i16 = l & 15, kq = l >> 4, w = threadIdx.x >> 6;
  // fibre 4 (w KPW + s) + kq, element i16 along K
  const int gl = opaque(8 * (sw16<D>(fbase16<D, K>(4 * w * KPW)) ^ sw16<D>(fbase16<D, K>(kq) ^ (i16 * st))));
  f4v g[2][2][2];                          // [chain][row tile][column tile]
#pragma unroll
  for (int c = 0; c < 2; ++c)
    g[c][0][0] = g[c][0][1] = g[c][1][0] = g[c][1][1] = f4v{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int s0 = 0; s0 < KPW; s0 += B) {
    float2 xa[B], ya[B];
#pragma unroll
    for (int b = 0; b < B; ++b) {
      const int off = gl ^ (8 * sw16<D>(fbase16<D, K>(4 * (s0 + b))));
      xa[b] = lds_ld2(X, off);
      ya[b] = lds_ld2(Y, off);
    }
#pragma unroll
    for (int b = 0; b < B; ++b) {
      const int c = b & 1;
      g[c][0][0] = mfma4f(xa[b].x, ya[b].x, g[c][0][0]);
      g[c][1][1] = mfma4f(xa[b].y, ya[b].y, g[c][1][1]);
      g[c][0][1] = mfma4f(xa[b].x, ya[b].y, g[c][0][1]);
      g[c][1][0] = mfma4f(xa[b].y, ya[b].x, g[c][1][0]);
    }
  }
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    gre[v] = ((double)g[0][0][0][v] + (double)g[1][0][0][v]) + ((double)g[0][1][1][v] + (double)g[1][1][1][v]);
    gim[v] = ((double)g[0][0][1][v] + (double)g[1][0][1][v]) - ((double)g[0][1][0][v] + (double)g[1][1][0][v]);
  }
}

// LDS: the node tensor, the work buffer, and a finishing area (per-wave Gram partials and the
// summed Gram in FP64) that overlays the work buffer when it is large enough.
template <int D, int W>
struct Mf16Lds {
  static constexpr int TBYTES = Mf16<D>::S * 8;
  static constexpr int FIN = W * 256 * 16 + 256 * 16;   // partials (double2) + summed Gram (double2)
  static constexpr int BYTES = 2 * TBYTES + (FIN > TBYTES ? FIN - TBYTES : 0);
};

// Sum the W partials, keep the Hermitian part normalised by the real trace, store the
// message at position pos and update this thread's distance maxima (squared magnitudes).
template <int W>
__device__ __forceinline__ void finish16(double2* part, double2* S, long pos, const cplx<float>* __restrict__ min,
                                         cplx<float>* mout, double& dnum2, double& dden2) {
  for (int o = threadIdx.x; o < 256; o += 64 * W) {
    double re = 0.0, im = 0.0;
#pragma unroll
    for (int w = 0; w < W; ++w) {
      re += part[w * 256 + o].x;
      im += part[w * 256 + o].y;
    }
    S[o] = make_double2(re, im);
  }
  __syncthreads();
  double tr = 0.0;
#pragma unroll
  for (int c = 0; c < 16; ++c) tr += S[c * 17].x;
  const double inv = 1.0 / tr;
  for (int o = threadIdx.x; o < 256; o += 64 * W) {
    const int c = o >> 4, e = o & 15;
    const double2 a = S[o], b = S[e * 16 + c];
    const cplx<float> v = mk((float)(0.5 * (a.x + b.x) * inv), (float)(0.5 * (a.y - b.y) * inv));
    const cplx<double> vd = cast<double>(v), od = cast<double>(min[pos * 256 + o]);
    dnum2 = fmax(dnum2, abs2(vd - od));
    dden2 = fmax(dden2, abs2(vd + od));
    mout[pos * 256 + o] = v;
  }
}

template <int D, int W, int K>
__device__ __forceinline__ void gram_finish16(const float2* T, float2* Wb, long pos, const cplx<float>* min,
                                              cplx<float>* mout, double& dnum2, double& dden2) {
  double gre[4], gim[4];
  gram16<D, W, K>(T, Wb, gre, gim);
  __syncthreads();                                      // every wave has read Wb
  double2* part = reinterpret_cast<double2*>(Wb);
  double2* S = part + W * 256;
  const int l = threadIdx.x & 63, w = threadIdx.x >> 6, kq = l >> 4, e = l & 15;
#pragma unroll
  for (int v = 0; v < 4; ++v) part[w * 256 + (4 * kq + v) * 16 + e] = make_double2(gre[v], gim[v]);
  __syncthreads();
  finish16<W>(part, S, pos, min, mout, dnum2, dden2);
  __syncthreads();                                      // the next mode product overwrites Wb
}

template <int D, int W>
__global__ __launch_bounds__(64 * W) void k_mfma16_bp(ClassView cv, const cplx<float>* __restrict__ T,
                                                      cplx<float>* msgA, cplx<float>* msgB, BpStatus* st) {
  static_assert(D == 2 || D == 3, "degree 2 or 3");
  static_assert(Mf16<D>::TILES % W == 0 && Mf16<D>::KS % W == 0, "tiles and k-steps split evenly over the waves");
  static_assert(Mf16<D>::S % (128 * W) == 0, "tensor load in whole rounds of float4");
  extern __shared__ __align__(16) unsigned char smem16[];
  float2* Ts = reinterpret_cast<float2*>(smem16);
  float2* Wb = Ts + Mf16<D>::S;
  const int cur = bp_begin(st, cv.sweep);
  if (cur < 0) return;
  const int n = cv.n;
  const cplx<float>* min = cur ? msgB : msgA;
  cplx<float>* mout = cur ? msgA : msgB;
  double dnum = 0.0, dden = 0.0;
  constexpr int S = Mf16<D>::S;
  constexpr int NT = 64 * W;
  for (int i = blockIdx.x; i < n; i += gridDim.x) {
    // tensor: float4 q holds elements 2q and 2q + 1, whose swizzled slots are one aligned pair
    const float4* Tg = reinterpret_cast<const float4*>(T + (size_t)i * S);
    float4* T4 = reinterpret_cast<float4*>(Ts);
#pragma unroll
    for (int k = 0; k < S / (2 * NT); ++k) {
      const int q = threadIdx.x + NT * k;
      const float4 v = Tg[q];
      const int sw = sw16<D>(2 * q);
      T4[sw >> 1] = (sw & 1) ? make_float4(v.z, v.w, v.x, v.y) : v;
    }
    const cplx<float>* m[D];
#pragma unroll
    for (int j = 0; j < D; ++j) m[j] = min + (size_t)cv.in_pos[j * n + i] * 256;
    __syncthreads();
    if constexpr (D == 2) {
      mp16<D, W, 1>(Ts, Wb, m[1]);                     // W = T x1 m1
      __syncthreads();
      gram_finish16<D, W, 0>(Ts, Wb, cv.out_pos[i], min, mout, dnum, dden);
      mp16<D, W, 0>(Ts, Wb, m[0]);                     // W = T x0 m0
      __syncthreads();
      gram_finish16<D, W, 1>(Ts, Wb, cv.out_pos[n + i], min, mout, dnum, dden);
    } else {
      mp16<D, W, 2>(Ts, Wb, m[2]);                     // W = T x2 m2
      __syncthreads();
      mp16<D, W, 1>(Wb, Wb, m[1]);                     // W = W x1 m1
      __syncthreads();
      gram_finish16<D, W, 0>(Ts, Wb, cv.out_pos[i], min, mout, dnum, dden);
      mp16<D, W, 2>(Ts, Wb, m[2]);                     // W = T x2 m2
      __syncthreads();
      mp16<D, W, 0>(Wb, Wb, m[0]);                     // W = W x0 m0
      __syncthreads();
      gram_finish16<D, W, 1>(Ts, Wb, cv.out_pos[n + i], min, mout, dnum, dden);
      mp16<D, W, 0>(Ts, Wb, m[0]);                     // W = T x0 m0
      __syncthreads();
      mp16<D, W, 1>(Wb, Wb, m[1]);                     // W = W x1 m1
      __syncthreads();
      gram_finish16<D, W, 2>(Ts, Wb, cv.out_pos[2 * n + i], min, mout, dnum, dden);
    }
  }
  // distance maxima of squared magnitudes: block reduction, one square root per block
  __shared__ double red[2][W];
  const int l = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (int o = 32; o > 0; o >>= 1) {
    dnum = fmax(dnum, __shfl_xor(dnum, o, 64));
    dden = fmax(dden, __shfl_xor(dden, o, 64));
  }
  if (l == 0) {
    red[0][w] = dnum;
    red[1][w] = dden;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double a = 0.0, b = 0.0;
    for (int k = 0; k < W; ++k) {
      a = fmax(a, red[0][k]);
      b = fmax(b, red[1][k]);
    }
    bp_end(st, cv.sweep, sqrt(a), sqrt(b));
  }
}

// ---------------------------------------------------------------------------
// The degree-3 sweep with the messages moved to the bra side (k_mfma16_bp3h).  Every message is
// Hermitian, so sum_a conj(T[..a..]) m[a][b] = conj((T x m)[..b..]); with P_j = T x_j m_j,
//   out_0 = Gram_0(P_1, P_2),  out_1 = Gram_1(P_0, P_2),  out_2 = Gram_2(P_0, P_1)
// (bra first): three mode products and three Grams per node instead of six and three.  The three
// products of a node do not fit in LDS together, so the node runs one physical half at a time (the
// Grams sum over p): T_p in buffer 0 (32 KB), P_2 and P_1 into buffers 1 and 2, then P_0 in place of
// T_p; the Gram partials of both halves add up in FP64 registers.  96 KB of LDS, one node per CU.
template <int W>
struct Mf16HLds {
  static constexpr int HB = Mf16<3>::S / 2 * 8;             // bytes of one half buffer
  static constexpr int FIN = W * 256 * 16 + 256 * 16;       // partials + summed Gram, over the buffers
  static_assert(FIN <= 3 * HB, "finishing area fits in the three buffers");
  static constexpr int BYTES = 3 * HB;
};

template <int W>
__global__ __launch_bounds__(64 * W) void k_mfma16_bp3h(ClassView cv, const cplx<float>* __restrict__ T,
                                                        cplx<float>* msgA, cplx<float>* msgB, BpStatus* st) {
  constexpr int D = 3, S = Mf16<D>::S, SH = S / 2, NT = 64 * W;
  static_assert(Mf16<D>::TILES % (2 * W) == 0 && Mf16<D>::KS % (2 * W) == 0, "half tiles split over the waves");
  static_assert(SH % (2 * NT) == 0, "half loads in whole rounds of float4");
  extern __shared__ __align__(16) unsigned char smem16[];
  float2* B0 = reinterpret_cast<float2*>(smem16);   // T_p, then P_0 in place
  float2* B1 = B0 + SH;                             // P_2
  float2* B2 = B1 + SH;                             // P_1
  const int cur = bp_begin(st, cv.sweep);
  if (cur < 0) return;
  const int n = cv.n;
  const cplx<float>* min = cur ? msgB : msgA;
  cplx<float>* mout = cur ? msgA : msgB;
  double dnum = 0.0, dden = 0.0;
  const int l = threadIdx.x & 63, w = threadIdx.x >> 6, kq = l >> 4, e16 = l & 15;
  for (int i = blockIdx.x; i < n; i += gridDim.x) {
    const cplx<float>* m[D];
#pragma unroll
    for (int j = 0; j < D; ++j) m[j] = min + (size_t)cv.in_pos[j * n + i] * 256;
    double gr[D][4], gi[D][4];
#pragma unroll
    for (int k = 0; k < D; ++k)
#pragma unroll
      for (int v = 0; v < 4; ++v) gr[k][v] = gi[k][v] = 0.0;
#pragma unroll 1
    for (int p = 0; p < 2; ++p) {
      __syncthreads();  // the previous half's (or node's) readers of the buffers are done
      const float4* Tg = reinterpret_cast<const float4*>(T + (size_t)i * S + (size_t)p * SH);
      float4* T4 = reinterpret_cast<float4*>(B0);
#pragma unroll
      for (int k = 0; k < SH / (2 * NT); ++k) {
        const int q = threadIdx.x + NT * k;
        const float4 v = Tg[q];
        const int sw = sw16<D>(2 * q);
        T4[sw >> 1] = (sw & 1) ? make_float4(v.z, v.w, v.x, v.y) : v;
      }
      __syncthreads();
      mp16<D, W, 2, true>(B0, B1, m[2]);             // P_2
      mp16<D, W, 1, true>(B0, B2, m[1]);             // P_1
      __syncthreads();
      double a[4], b[4];
      gram16<D, W, 0, true>(B2, B1, a, b);           // out_0 += Gram_0(P_1, P_2)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        gr[0][v] += a[v];
        gi[0][v] += b[v];
      }
      mp16<D, W, 0, true>(B0, B0, m[0]);             // P_0 in place (the Gram above reads B1, B2 only)
      __syncthreads();
      gram16<D, W, 1, true>(B0, B1, a, b);           // out_1 += Gram_1(P_0, P_2)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        gr[1][v] += a[v];
        gi[1][v] += b[v];
      }
      gram16<D, W, 2, true>(B0, B2, a, b);           // out_2 += Gram_2(P_0, P_1)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        gr[2][v] += a[v];
        gi[2][v] += b[v];
      }
    }
    __syncthreads();  // every wave has read the three buffers
    double2* part = reinterpret_cast<double2*>(B0);
    double2* Sg = part + W * 256;
#pragma unroll 1
    for (int k = 0; k < D; ++k) {
#pragma unroll
      for (int v = 0; v < 4; ++v) part[w * 256 + (4 * kq + v) * 16 + e16] = make_double2(gr[k][v], gi[k][v]);
      __syncthreads();
      finish16<W>(part, Sg, cv.out_pos[k * n + i], min, mout, dnum, dden);
      __syncthreads();
    }
  }
  __shared__ double red[2][W];
  for (int o = 32; o > 0; o >>= 1) {
    dnum = fmax(dnum, __shfl_xor(dnum, o, 64));
    dden = fmax(dden, __shfl_xor(dden, o, 64));
  }
  if (l == 0) {
    red[0][w] = dnum;
    red[1][w] = dden;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double x = 0.0, y = 0.0;
    for (int k = 0; k < W; ++k) {
      x = fmax(x, red[0][k]);
      y = fmax(y, red[1][k]);
    }
    bp_end(st, cv.sweep, sqrt(x), sqrt(y));
  }
}

// ---------------------------------------------------------------------------
// BP sweep of an FP64 job at chi = 16, degree 2 or 3 (k_mfma16_bp64), on v_mfma_f64_16x16x4f64:
// the reference's nine-contraction schedule with the FP32 kernel's tiling, complex128 elements
// (double2) at the same swizzled positions.  A degree-3 node (128 KB) runs one physical half at
// a time through a 64 KB tensor buffer and a 64 KB work buffer, the Grams (which sum over p)
// adding up in FP64 registers; a degree-2 node (8 KB) runs whole.  FP64 MFMA result register v of
// lane l holds D[l / 16 + 4 v][l % 16].

// B operand of message m for a mode product (as msg_operand16, in FP64).
__device__ __forceinline__ void msg_operand16d(const cplx<double>* __restrict__ m, double (&bm)[8][2]) {
  const int l = threadIdx.x & 63, a = l & 15, kq = l >> 4;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const cplx<double> v = m[a * 16 + 4 * q + kq];
    bm[q][0] = v.x;
    bm[q][1] = v.y;
    bm[q + 4][0] = -v.y;
    bm[q + 4][1] = v.x;
  }
}

// Y = X x_J m over this wave's tiles of a buffer holding H = 1 or 2 halves' worth of fibres (the
// first 1/H of them); X may equal Y (a wave reads its tiles whole before writing them).
template <int D, int W, int J, int H>
__device__ __forceinline__ void mp16d(const double2* X, double2* Y, const cplx<double>* __restrict__ m) {
  double bm[8][2];
  msg_operand16d(m, bm);
  constexpr int st = cpow16(D - 1 - J);
  constexpr int TPW = Mf16<D>::TILES / H / W;
  static_assert(TPW >= 1 && Mf16<D>::TILES % (H * W) == 0, "tiles split evenly over the waves");
  const int tid = opaque((int)threadIdx.x);
  const int l = tid & 63, i16 = l & 15, kq = l >> 4, w = tid >> 6;
  double a[TPW][8];
#pragma unroll
  for (int t = 0; t < TPW; ++t) {
    const int f0 = (w * TPW + t) * 16;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const double2 v = X[sw16<D>(fbase16<D, J>(f0 + i16) + (4 * q + kq) * st)];
      a[t][q] = v.x;
      a[t][q + 4] = v.y;
    }
  }
  d4v acc[TPW][2];
#pragma unroll
  for (int t = 0; t < TPW; ++t) acc[t][0] = acc[t][1] = d4v{0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int s = 0; s < 8; ++s)
#pragma unroll
    for (int t = 0; t < TPW; ++t) {
      acc[t][0] = mfma4d(a[t][s], bm[s][0], acc[t][0]);
      acc[t][1] = mfma4d(a[t][s], bm[s][1], acc[t][1]);
    }
  // D[kq + 4 v][i16]: fibre f0 + kq + 4 v, element i16 along J
#pragma unroll
  for (int t = 0; t < TPW; ++t) {
    const int f0 = (w * TPW + t) * 16;
#pragma unroll
    for (int v = 0; v < 4; ++v)
      Y[sw16<D>(fbase16<D, J>(f0 + kq + 4 * v) + i16 * st)] = make_double2(acc[t][0][v], acc[t][1][v]);
  }
}

// This wave's k-steps of out[c][e] = sum_f conj(X[f][c]) Y[f][e] over the buffer's fibres along
// bond K, added into gre/gim[v] = entry (c = l / 16 + 4 v, e = l % 16).
template <int D, int W, int K, int H>
__device__ __forceinline__ void gram16d(const double2* X, const double2* Y, double (&gre)[4], double (&gim)[4]) {
  constexpr int st = cpow16(D - 1 - K);
  constexpr int KPW = Mf16<D>::KS / H / W;
  static_assert(KPW >= 2 && KPW % 2 == 0, "two chains per wave");
  const int tid = opaque((int)threadIdx.x);
  const int l = tid & 63, i16 = l & 15, kq = l >> 4, w = tid >> 6;
  d4v gr[2], gi[2];
  gr[0] = gr[1] = gi[0] = gi[1] = d4v{0.0, 0.0, 0.0, 0.0};
#pragma unroll 4
  for (int s = 0; s < KPW; ++s) {
    const int idx = sw16<D>(fbase16<D, K>(4 * (w * KPW + s) + kq) + i16 * st);
    const double2 x = X[idx], y = Y[idx];
    const int c = s & 1;
    gr[c] = mfma4d(x.x, y.x, gr[c]);
    gr[c] = mfma4d(x.y, y.y, gr[c]);
    gi[c] = mfma4d(x.x, y.y, gi[c]);
    gi[c] = mfma4d(-x.y, y.x, gi[c]);
  }
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    gre[v] += gr[0][v] + gr[1][v];
    gim[v] += gi[0][v] + gi[1][v];
  }
}

template <int D, int W>
struct Mf16D {
  static constexpr int H = D == 3 ? 2 : 1;                   // physical halves run one at a time
  static constexpr int SB = Mf16<D>::S / H;                  // complex elements per buffer
  static constexpr int FIN = W * 256 * 16 + 256 * 16;        // partials + summed Gram over the work buffer
  static constexpr int BYTES = 2 * SB * 16 + (FIN > SB * 16 ? FIN - SB * 16 : 0);
};

// One buffer's share of the nine-contraction schedule (the whole node, or one physical half),
// adding its three (two) Grams into gr/gi[k]; on return the last Gram may still be reading Wb.
template <int D, int W, int H>
__device__ __forceinline__ void bp16d_schedule(const double2* Ts, double2* Wb, const cplx<double>* const (&m)[D],
                                               double (&gr)[D][4], double (&gi)[D][4]) {
  if constexpr (D == 2) {
    mp16d<D, W, 1, H>(Ts, Wb, m[1]);                       // W = T x1 m1
    __syncthreads();
    gram16d<D, W, 0, H>(Ts, Wb, gr[0], gi[0]);
    __syncthreads();
    mp16d<D, W, 0, H>(Ts, Wb, m[0]);                       // W = T x0 m0
    __syncthreads();
    gram16d<D, W, 1, H>(Ts, Wb, gr[1], gi[1]);
  } else {
    mp16d<D, W, 2, H>(Ts, Wb, m[2]);                       // W = T x2 m2
    __syncthreads();
    mp16d<D, W, 1, H>(Wb, Wb, m[1]);                       // W = W x1 m1
    __syncthreads();
    gram16d<D, W, 0, H>(Ts, Wb, gr[0], gi[0]);
    __syncthreads();
    mp16d<D, W, 2, H>(Ts, Wb, m[2]);
    __syncthreads();
    mp16d<D, W, 0, H>(Wb, Wb, m[0]);                       // W = T x2 m2 x0 m0
    __syncthreads();
    gram16d<D, W, 1, H>(Ts, Wb, gr[1], gi[1]);
    __syncthreads();
    mp16d<D, W, 0, H>(Ts, Wb, m[0]);
    __syncthreads();
    mp16d<D, W, 1, H>(Wb, Wb, m[1]);                       // W = T x0 m0 x1 m1
    __syncthreads();
    gram16d<D, W, 2, H>(Ts, Wb, gr[2], gi[2]);
  }
}

template <int D, int W>
__global__ __launch_bounds__(64 * W) void k_mfma16_bp64(ClassView cv, const cplx<double>* __restrict__ T,
                                                        cplx<double>* msgA, cplx<double>* msgB, BpStatus* st) {
  static_assert(D == 2 || D == 3, "degree 2 or 3");
  using L = Mf16D<D, W>;
  constexpr int S = Mf16<D>::S, H = L::H, SB = L::SB, NT = 64 * W;
  static_assert(SB % NT == 0, "buffer loads in whole rounds");
  extern __shared__ __align__(16) unsigned char smem16[];
  double2* Ts = reinterpret_cast<double2*>(smem16);
  double2* Wb = Ts + SB;
  const int cur = bp_begin(st, cv.sweep);
  if (cur < 0) return;
  const int n = cv.n;
  const cplx<double>* min = cur ? msgB : msgA;
  cplx<double>* mout = cur ? msgA : msgB;
  double dnum = 0.0, dden = 0.0;
  const int l = threadIdx.x & 63, w = threadIdx.x >> 6, kq = l >> 4, e16 = l & 15;
  for (int i = blockIdx.x; i < n; i += gridDim.x) {
    const cplx<double>* m[D];
#pragma unroll
    for (int j = 0; j < D; ++j) m[j] = min + (size_t)cv.in_pos[j * n + i] * 256;
    double gr[D][4], gi[D][4];
#pragma unroll
    for (int k = 0; k < D; ++k)
#pragma unroll
      for (int v = 0; v < 4; ++v) gr[k][v] = gi[k][v] = 0.0;
#pragma unroll 1
    for (int p = 0; p < H; ++p) {
      __syncthreads();  // the previous half's (or node's) readers of the buffers are done
      const cplx<double>* Tg = T + (size_t)i * S + (size_t)p * SB;
#pragma unroll 4
      for (int k = 0; k < SB / NT; ++k) {
        const int x = threadIdx.x + NT * k;
        const cplx<double> v = Tg[x];
        Ts[sw16<D>(x)] = make_double2(v.x, v.y);
      }
      __syncthreads();
      bp16d_schedule<D, W, H>(Ts, Wb, m, gr, gi);
    }
    __syncthreads();  // every wave has read the work buffer
    double2* part = Wb;
    double2* Sg = part + W * 256;
#pragma unroll
    for (int k = 0; k < D; ++k) {
#pragma unroll
      for (int v = 0; v < 4; ++v) part[w * 256 + (kq + 4 * v) * 16 + e16] = make_double2(gr[k][v], gi[k][v]);
      __syncthreads();
      for (int o = threadIdx.x; o < 256; o += NT) {
        double re = 0.0, im = 0.0;
#pragma unroll
        for (int q = 0; q < W; ++q) {
          re += part[q * 256 + o].x;
          im += part[q * 256 + o].y;
        }
        Sg[o] = make_double2(re, im);
      }
      __syncthreads();
      double tr = 0.0;
#pragma unroll
      for (int c = 0; c < 16; ++c) tr += Sg[c * 17].x;
      const double inv = 1.0 / tr;
      const long pos = cv.out_pos[k * n + i];
      for (int o = threadIdx.x; o < 256; o += NT) {
        const int c = o >> 4, e = o & 15;
        const double2 a = Sg[o], b = Sg[e * 16 + c];
        const cplx<double> v = mk(0.5 * (a.x + b.x) * inv, 0.5 * (a.y - b.y) * inv);
        const cplx<double> od = min[pos * 256 + o];
        dnum = fmax(dnum, abs2(v - od));
        dden = fmax(dden, abs2(v + od));
        mout[pos * 256 + o] = v;
      }
      __syncthreads();
    }
  }
  __shared__ double red[2][W];
  for (int o = 32; o > 0; o >>= 1) {
    dnum = fmax(dnum, __shfl_xor(dnum, o, 64));
    dden = fmax(dden, __shfl_xor(dden, o, 64));
  }
  if (l == 0) {
    red[0][w] = dnum;
    red[1][w] = dden;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double x = 0.0, y = 0.0;
    for (int k = 0; k < W; ++k) {
      x = fmax(x, red[0][k]);
      y = fmax(y, red[1][k]);
    }
    bp_end(st, cv.sweep, sqrt(x), sqrt(y));
  }
}

// Extended messages of an FP64 job at chi = 16, degree 2 or 3 (k_mfma16_ext64): the schedule of
// k_mfma16_bp64 run one physical half at a time (degree 2 too), so each half's Grams are the
// p-resolved Grams G_p; they are summed over the waves into LDS after each half, and every bond's
// 32 x 32 message is assembled from G_0, G_1 and its ZZ-split gates as in k_mfma16_ext.
template <int D, int W>
struct Mf16DExt {
  static constexpr int SB = Mf16<D>::S / 2;                  // one physical half
  static constexpr int PART = W * 256 * 16;                  // per-wave partials over the work buffer
  static constexpr int BYTES = 2 * SB * 16 + (PART > SB * 16 ? PART - SB * 16 : 0) + 2 * D * 256 * 16;
};

template <int D, int W>
__global__ __launch_bounds__(64 * W) void k_mfma16_ext64(ClassView cv, const cplx<double>* __restrict__ T,
                                                         const cplx<double>* __restrict__ msgs,
                                                         cplx<double>* __restrict__ ext_out) {
  static_assert(D == 2 || D == 3, "degree 2 or 3");
  using L = Mf16DExt<D, W>;
  constexpr int S = Mf16<D>::S, H = 2, SB = L::SB, NT = 64 * W;
  static_assert(SB % NT == 0, "half loads in whole rounds");
  extern __shared__ __align__(16) unsigned char smem16[];
  double2* Ts = reinterpret_cast<double2*>(smem16);
  double2* Wb = Ts + SB;
  double2* Gs = reinterpret_cast<double2*>(smem16 + L::BYTES - 2 * D * 256 * 16);  // [p][k][256]
  const int n = cv.n;
  const int l = threadIdx.x & 63, w = threadIdx.x >> 6, kq = l >> 4, e16 = l & 15;
  for (int i = blockIdx.x; i < n; i += gridDim.x) {
    const cplx<double>* m[D];
#pragma unroll
    for (int j = 0; j < D; ++j) m[j] = msgs + (size_t)cv.in_pos[j * n + i] * 256;
#pragma unroll 1
    for (int p = 0; p < H; ++p) {
      __syncthreads();  // the previous half's (or node's) readers of the buffers are done
      const cplx<double>* Tg = T + (size_t)i * S + (size_t)p * SB;
#pragma unroll 4
      for (int k = 0; k < SB / NT; ++k) {
        const int x = threadIdx.x + NT * k;
        const cplx<double> v = Tg[x];
        Ts[sw16<D>(x)] = make_double2(v.x, v.y);
      }
      double gr[D][4], gi[D][4];
#pragma unroll
      for (int k = 0; k < D; ++k)
#pragma unroll
        for (int v = 0; v < 4; ++v) gr[k][v] = gi[k][v] = 0.0;
      __syncthreads();
      bp16d_schedule<D, W, H>(Ts, Wb, m, gr, gi);
      __syncthreads();  // every wave has read the work buffer
      double2* part = Wb;
#pragma unroll
      for (int k = 0; k < D; ++k) {
#pragma unroll
        for (int v = 0; v < 4; ++v) part[w * 256 + (kq + 4 * v) * 16 + e16] = make_double2(gr[k][v], gi[k][v]);
        __syncthreads();
        for (int o = threadIdx.x; o < 256; o += NT) {
          double re = 0.0, im = 0.0;
#pragma unroll
          for (int q = 0; q < W; ++q) {
            re += part[q * 256 + o].x;
            im += part[q * 256 + o].y;
          }
          Gs[(p * D + k) * 256 + o] = make_double2(re, im);
        }
        __syncthreads();
      }
    }
#pragma unroll
    for (int k = 0; k < D; ++k) {
      const double2* G0 = Gs + k * 256;
      const double2* G1 = Gs + (D + k) * 256;
      cplx<double> tr = mk(0.0, 0.0);
#pragma unroll
      for (int c = 0; c < 16; ++c) tr = tr + mk(G0[c * 17].x + G1[c * 17].x, G0[c * 17].y + G1[c * 17].y);
      const cplx<double>* zg = cv.zz + 2 * (size_t)cv.lmbd_pos[k * n + i];
      const cplx<double> g0 = zg[0], g1 = zg[1];
      const cplx<double> inv = cinv(tr * mk(abs2(g0) + abs2(g1), 0.0));
      cplx<double>* out = ext_out + (size_t)cv.out_pos[k * n + i] * 1024;
      for (int o = threadIdx.x; o < 1024; o += NT) {
        const int row = o >> 5, col = o & 31;
        const int s_ = row >> 4, c = row & 15, s2 = col >> 4, e = col & 15;
        const double2 a = G0[c * 16 + e], b = G1[c * 16 + e];
        const cplx<double> gsum = ((s_ + s2) & 1) ? mk(a.x - b.x, a.y - b.y) : mk(a.x + b.x, a.y + b.y);
        out[o] = conj(s_ ? g1 : g0) * (s2 ? g1 : g0) * gsum * inv;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Extended messages of an FP32 job at chi = 16, degree 2 or 3 (state.py:127-139,
// backends.py:391-408): the mode products of k_mfma16_bp in FP32 MFMA, then for
// every bond K the two physical-index-resolved Grams
//   G_p[c][e] = sum over fibres with physical index p of conj(T[f][c]) Y[f][e]
// on v_mfma_f64_16x16x4f64 from the FP32 operands (the products are exact in FP64
// and the sums keep FP64 rounding: the canonicaliser's S^{-1/2} amplifies Gram
// rounding on small Schmidt values), and the 32 x 32 FP64 message of the bond's
// ZZ-split gates (g_0, g_1):
//   ext[(s, c)][(s', e)] = conj(g_s) g_s' (G_0 + (-1)^(s + s') G_1)[c][e] / (tr (|g_0|^2 + |g_1|^2)),
// tr the complex trace of G_0 + G_1 (the ket side carries the conditional Z gate
// sqrt(cos) I, sqrt(-i sin) Z of backends.py:592-600, the bra side its conjugate).
//
// The fibres of a bond are ordered with the physical index most significant, so the
// first half of the k-steps is p = 0: with W >= 2 waves, waves [0, W/2) hold p = 0;
// one wave keeps a separate accumulator per half.

template <int D, int W>
struct Mf16ExtLds {
  static constexpr int TBYTES = Mf16<D>::S * 8;
  static constexpr int PS = W == 1 ? 2 : W;                 // Gram partial slots (double2 x 256 each)
  static constexpr int FIN = PS * 256 * 16 + 2 * 256 * 16;  // partials + G_0, G_1
  static constexpr int BYTES = 2 * TBYTES + (FIN > TBYTES ? FIN - TBYTES : 0);
};

// This wave's p-resolved Gram partials of bond K into its slot(s) of `part`
// (after a barrier: `part` overlays Y).
template <int D, int W, int K>
__device__ __forceinline__ void gram16_ext(const float2* X, const float2* Y, double2* part) {
  constexpr int st = cpow16(D - 1 - K);
  constexpr int KS = Mf16<D>::KS, KPW = KS / W;
  constexpr int B = KPW < 4 ? KPW : 4;
  constexpr int NP = W == 1 ? 2 : 1;                      // accumulator sets per wave
  static_assert(W == 1 || (W % 2 == 0 && KPW <= KS / 2), "one physical index per wave");
  const int l = threadIdx.x & 63, i16 = l & 15, kq = l >> 4, w = threadIdx.x >> 6;
  const int gl = opaque(8 * (sw16<D>(fbase16<D, K>(4 * w * KPW)) ^ sw16<D>(fbase16<D, K>(kq) ^ (i16 * st))));
  d4v gr[NP][2], gi[NP][2];                               // [p][chain]
#pragma unroll
  for (int p = 0; p < NP; ++p)
#pragma unroll
    for (int c = 0; c < 2; ++c) gr[p][c] = gi[p][c] = d4v{0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int s0 = 0; s0 < KPW; s0 += B) {
    float2 xa[B], ya[B];
#pragma unroll
    for (int b = 0; b < B; ++b) {
      const int off = gl ^ (8 * sw16<D>(fbase16<D, K>(4 * (s0 + b))));
      xa[b] = lds_ld2(X, off);
      ya[b] = lds_ld2(Y, off);
    }
#pragma unroll
    for (int b = 0; b < B; ++b) {
      const int p = NP == 2 ? (s0 + b) / (KS / 2) : 0;
      const int c = b & 1;
      const double xr = xa[b].x, xi = xa[b].y, yr = ya[b].x, yi = ya[b].y;
      gr[p][c] = mfma4d(xr, yr, gr[p][c]);
      gi[p][c] = mfma4d(xr, yi, gi[p][c]);
      gr[p][c] = mfma4d(xi, yi, gr[p][c]);
      gi[p][c] = mfma4d(-xi, yr, gi[p][c]);
    }
  }
  __syncthreads();                                        // every wave has read Y
  // result register v of lane l: entry [c = l / 16 + 4 v][e = l % 16]
#pragma unroll
  for (int p = 0; p < NP; ++p) {
    const int slot = NP == 2 ? p : w;
#pragma unroll
    for (int v = 0; v < 4; ++v)
      part[slot * 256 + (kq + 4 * v) * 16 + i16] =
          make_double2(gr[p][0][v] + gr[p][1][v], gi[p][0][v] + gi[p][1][v]);
  }
}

// Sum the partial slots into G_0, G_1 and write the 32 x 32 extended message of bond K.
template <int D, int W, int K>
__device__ __forceinline__ void ext_finish16(double2* part, const ClassView& cv, int i, cplx<double>* ext_out) {
  constexpr int PS = Mf16ExtLds<D, W>::PS, H = PS / 2, NT = 64 * W;
  double2* G = part + PS * 256;                           // G_0 at [0, 256), G_1 at [256, 512)
  __syncthreads();
  for (int o = threadIdx.x; o < 512; o += NT) {
    const int p = o >> 8, x = o & 255;
    double re = 0.0, im = 0.0;
#pragma unroll
    for (int q = 0; q < H; ++q) {
      re += part[(p * H + q) * 256 + x].x;
      im += part[(p * H + q) * 256 + x].y;
    }
    G[o] = make_double2(re, im);
  }
  __syncthreads();
  cplx<double> tr = mk(0.0, 0.0);
#pragma unroll
  for (int c = 0; c < 16; ++c) {
    const double2 a = G[c * 17], b = G[256 + c * 17];
    tr = tr + mk(a.x + b.x, a.y + b.y);
  }
  const cplx<double>* zg = cv.zz + 2 * (size_t)cv.lmbd_pos[K * cv.n + i];
  const cplx<double> g0 = zg[0], g1 = zg[1];
  const cplx<double> inv = cinv(tr * mk(abs2(g0) + abs2(g1), 0.0));
  cplx<double>* out = ext_out + (size_t)cv.out_pos[K * cv.n + i] * 1024;
  for (int o = threadIdx.x; o < 1024; o += NT) {
    const int row = o >> 5, col = o & 31;
    const int s = row >> 4, c = row & 15, s2 = col >> 4, e = col & 15;
    const double2 a = G[c * 16 + e], b = G[256 + c * 16 + e];
    const cplx<double> gsum = ((s + s2) & 1) ? mk(a.x - b.x, a.y - b.y) : mk(a.x + b.x, a.y + b.y);
    out[o] = conj(s ? g1 : g0) * (s2 ? g1 : g0) * gsum * inv;
  }
  __syncthreads();                                        // the next mode product overwrites the work buffer
}

template <int D, int W>
__global__ __launch_bounds__(64 * W) void k_mfma16_ext(ClassView cv, const cplx<float>* __restrict__ T,
                                                       const cplx<float>* __restrict__ msgs,
                                                       cplx<double>* __restrict__ ext_out) {
  static_assert(D == 2 || D == 3, "degree 2 or 3");
  static_assert(Mf16<D>::TILES % W == 0 && Mf16<D>::KS % W == 0, "tiles and k-steps split evenly over the waves");
  static_assert(Mf16<D>::S % (128 * W) == 0, "tensor load in whole rounds of float4");
  extern __shared__ __align__(16) unsigned char smem16[];
  float2* Ts = reinterpret_cast<float2*>(smem16);
  float2* Wb = Ts + Mf16<D>::S;
  double2* part = reinterpret_cast<double2*>(Wb);
  const int n = cv.n;
  constexpr int S = Mf16<D>::S;
  constexpr int NT = 64 * W;
  for (int i = blockIdx.x; i < n; i += gridDim.x) {
    const float4* Tg = reinterpret_cast<const float4*>(T + (size_t)i * S);
    float4* T4 = reinterpret_cast<float4*>(Ts);
#pragma unroll
    for (int k = 0; k < S / (2 * NT); ++k) {
      const int q = threadIdx.x + NT * k;
      const float4 v = Tg[q];
      const int sw = sw16<D>(2 * q);
      T4[sw >> 1] = (sw & 1) ? make_float4(v.z, v.w, v.x, v.y) : v;
    }
    const cplx<float>* m[D];
#pragma unroll
    for (int j = 0; j < D; ++j) m[j] = msgs + (size_t)cv.in_pos[j * n + i] * 256;
    __syncthreads();
    if constexpr (D == 2) {
      mp16<D, W, 1>(Ts, Wb, m[1]);                     // W = T x1 m1
      __syncthreads();
      gram16_ext<D, W, 0>(Ts, Wb, part);
      ext_finish16<D, W, 0>(part, cv, i, ext_out);
      mp16<D, W, 0>(Ts, Wb, m[0]);                     // W = T x0 m0
      __syncthreads();
      gram16_ext<D, W, 1>(Ts, Wb, part);
      ext_finish16<D, W, 1>(part, cv, i, ext_out);
    } else {
      mp16<D, W, 2>(Ts, Wb, m[2]);                     // W = T x2 m2
      __syncthreads();
      mp16<D, W, 1>(Wb, Wb, m[1]);                     // W = W x1 m1
      __syncthreads();
      gram16_ext<D, W, 0>(Ts, Wb, part);
      ext_finish16<D, W, 0>(part, cv, i, ext_out);
      mp16<D, W, 2>(Ts, Wb, m[2]);                     // W = T x2 m2
      __syncthreads();
      mp16<D, W, 0>(Wb, Wb, m[0]);                     // W = W x0 m0
      __syncthreads();
      gram16_ext<D, W, 1>(Ts, Wb, part);
      ext_finish16<D, W, 1>(part, cv, i, ext_out);
      mp16<D, W, 0>(Ts, Wb, m[0]);                     // W = T x0 m0
      __syncthreads();
      mp16<D, W, 1>(Wb, Wb, m[1]);                     // W = W x1 m1
      __syncthreads();
      gram16_ext<D, W, 2>(Ts, Wb, part);
      ext_finish16<D, W, 2>(part, cv, i, ext_out);
    }
  }
}

}  // namespace bqa
