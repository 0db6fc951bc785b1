// FP32 MFMA node kernels for bond dimension chi = 8 and degree 2 or 3.
//
// One wave per node.  The node tensor T[p, b_1..b_d] (complex64) is staged in
// LDS with two work buffers.  Complex contractions run as real GEMMs on
// v_mfma_f32_16x16x4f32 with complex numbers embedded as 2x2 real blocks:
//
//   mode product along bond j:  rows = the 2 chi^(d-1) fibres along j,
//     [Yr | Yi] = [Xr | Xi] [[m_r^T, m_i^T], [-m_i^T, m_r^T]]      (K = N = 16)
//   Gram with bond k free:  G = [Tr | Ti]^T [Yr | Yi]  (16 x 16, K = fibres),
//     out_r = G[c][e] + G[8+c][8+e],  out_i = G[c][8+e] - G[8+c][e].
//
// MFMA 16x16x4 lane layout (checked on gfx950): A[i][k] at lane i + 16k,
// B[k][j] at lane j + 16k, D[4 (l / 16) + v][l % 16] in result register v.
// A mode product reads a 16-fibre tile completely before writing it, so it
// may run in place.
#pragma once

#include "common.hpp"
#include "kernels_node.hpp"

namespace bqa {

typedef float f4v __attribute__((ext_vector_type(4)));

constexpr int cpow8(int e) { return e <= 0 ? 1 : 8 * cpow8(e - 1); }

template <int D>
struct Mf8 {
  static constexpr int A = cpow8(D);      // chi^d
  static constexpr int S = 2 * A;         // complex elements per node
  static constexpr int ROWS = S / 8;      // fibres along one bond
  static constexpr int TILES = ROWS / 16;
  static constexpr int KS = ROWS / 4;     // k-steps of a Gram
};

template <int D, int J>
__device__ __forceinline__ int fbase8(int r) {
  constexpr int st = cpow8(D - 1 - J);
  return (r / st) * 8 * st + (r % st);
}

__device__ __forceinline__ f4v mfma4(float a, float b, f4v c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

// LDS float-index swizzle of the node tensor and work buffers: bit j (j < 6)
// of a stored index is flipped by the parity of (index & swz_mask<D>(j)).  The
// map is linear over GF(2), so an address built from disjoint bit fields
// (lane part | unrolled-loop constant) swizzles as swz(lane) ^ swz(constant),
// and the constant folds at compile time.  With it every access pattern below
// (mode-product reads and writes along each bond, Gram reads) touches distinct
// banks for 32 banks per half-wave and for 64 banks per wave, where the natural
// layout had up to 8-way conflicts.  The masks never contain bits 0 or 1:
// complex pairs stay adjacent and each aligned float4 stays one aligned float4
// with its elements permuted by the low two bits of swz(4 x).
template <int D>
__host__ __device__ constexpr int swz_mask(int j) {
  constexpr int m3[6] = {0x144, 0x124, 0x3C0, 0x490, 0x260, 0x580};
  constexpr int m2[6] = {0xE8, 0xC0, 0xF0, 0x60, 0xC0, 0x80};
  return D == 3 ? m3[j] : m2[j];
}

template <int D>
__host__ __device__ constexpr int swz(int w) {
  int out = w;
  for (int j = 0; j < 6; ++j) out ^= (__builtin_popcount(swz_mask<D>(j) & w) & 1) << j;
  return out;
}

// Work split of the FP32 BP/density kernel over the W waves of a node's block:
// each wave takes TPW of the 16-fibre tiles of every mode product and KPW of
// the k-steps of every Gram.
template <int D, int W>
struct Mf8W {
  static_assert(Mf8<D>::TILES % W == 0, "tiles split evenly over the waves");
  static constexpr int TPW = Mf8<D>::TILES / W;  // mode-product tiles per wave
  static constexpr int KPW = Mf8<D>::KS / W;     // Gram k-steps per wave
  static_assert(KPW % 4 == 0, "four Gram chains per wave");
};

// Per-lane parts of the swizzled addresses of mp8 and gram8 along each bond,
// including this wave's fibre offset, as byte offsets (4 x the float index;
// the swizzle commutes with the scaling), so an access costs one XOR and the
// array's LDS base rides in the instruction's offset.  Computed once per kernel;
// opaque() hides them from loop-invariant code motion at each use, so the
// unrolled (lane ^ constant) addresses are formed next to their LDS access
// instead of being hoisted out of the node loop into hundreds of live registers.
template <int D>
struct Swz8Lanes {
  int rd[D], wr[D], gl[D];
};

__device__ __forceinline__ int opaque(int x) {
  asm volatile("" : "+v"(x));
  return x;
}

template <int D, int W, int J>
__device__ __forceinline__ void swz8_lanes_j(Swz8Lanes<D>& z) {
  constexpr int st = cpow8(D - 1 - J);
  const int l = threadIdx.x & 63, i16 = l & 15, kq = l >> 4, w = threadIdx.x >> 6;
  // mode products: this wave's tiles start at fibre 16 TPW w
  const int fw = swz<D>(2 * fbase8<D, J>(16 * Mf8W<D, W>::TPW * w));
  // A[i][k]: fibre t*16 + i, element (4 s + kq) & 7 along J, component (4 s + kq) >> 3
  z.rd[J] = 4 * (fw ^ swz<D>(2 * (fbase8<D, J>(i16) + kq * st)));
  // D[4 kq + v][i]: fibre t*16 + 4 kq + v, element i & 7 along J, component i >> 3
  z.wr[J] = 4 * (fw ^ swz<D>(2 * (fbase8<D, J>(4 * kq) + (i16 & 7) * st) + (i16 >> 3)));
  // Gram operands: fibre 4 s + kq, element i & 7 along J, component i >> 3; this
  // wave's k-steps start at fibre 4 KPW w
  z.gl[J] = 4 * (swz<D>(2 * fbase8<D, J>(4 * Mf8W<D, W>::KPW * w)) ^
                 swz<D>(2 * (fbase8<D, J>(kq) + (i16 & 7) * st) + (i16 >> 3)));
}

__device__ __forceinline__ float lds_ld(const float* base, int byte_off) {
  return *reinterpret_cast<const float*>(reinterpret_cast<const char*>(base) + byte_off);
}

__device__ __forceinline__ void lds_st(float* base, int byte_off, float v) {
  *reinterpret_cast<float*>(reinterpret_cast<char*>(base) + byte_off) = v;
}

template <int D, int W>
__device__ __forceinline__ Swz8Lanes<D> swz8_lanes() {
  Swz8Lanes<D> z;
  swz8_lanes_j<D, W, 0>(z);
  swz8_lanes_j<D, W, 1>(z);
  if constexpr (D == 3) swz8_lanes_j<D, W, 2>(z);
  return z;
}

// Y = X x_J m on this wave's tiles (float views of swizzled complex64 LDS
// arrays).  All A operands of the wave's tiles are read first, then the MFMAs
// run interleaved over the independent tile accumulators.  The waves' tiles
// are disjoint and every read precedes every write, so X may equal Y.
template <int D, int W, int J>
__device__ __forceinline__ void mp8(const float* X, float* Y, const float (&bm)[4], const Swz8Lanes<D>& z) {
  constexpr int st = cpow8(D - 1 - J);
  constexpr int TPW = Mf8W<D, W>::TPW;
  const int rd = opaque(z.rd[J]), wr = opaque(z.wr[J]);
  float a[TPW][4];
#pragma unroll
  for (int t = 0; t < TPW; ++t)
#pragma unroll
    for (int s = 0; s < 4; ++s)
      a[t][s] = lds_ld(X, rd ^ (4 * swz<D>(2 * (fbase8<D, J>(t * 16) + 4 * (s & 1) * st) + (s >> 1))));
  f4v acc[TPW];
#pragma unroll
  for (int t = 0; t < TPW; ++t) acc[t] = f4v{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int s = 0; s < 4; ++s)
#pragma unroll
    for (int t = 0; t < TPW; ++t) acc[t] = mfma4(a[t][s], bm[s], acc[t]);
#pragma unroll
  for (int t = 0; t < TPW; ++t)
#pragma unroll
    for (int v = 0; v < 4; ++v) lds_st(Y, wr ^ (4 * swz<D>(2 * fbase8<D, J>(t * 16 + v))), acc[t][v]);
}

// Gram operands of k-steps [s0, s0 + B) along bond J from swizzled T and Y.
template <int D, int J, int B, typename F>
__device__ __forceinline__ void lds_batch(const F* T, const F* Y, int gl, int s0, F (&ta)[B], F (&ya)[B]) {
#pragma unroll
  for (int b = 0; b < B; ++b) {
    const int off = gl ^ (4 * swz<D>(2 * fbase8<D, J>(4 * (s0 + b))));
    ta[b] = lds_ld(T, off);
    ya[b] = lds_ld(Y, off);
  }
}

// Orders the LDS reads issued so far ahead of every use of (ta, ya): the empty
// asm takes the operands as read-write and clobbers memory, so the reads of the
// next batch stay above it and this batch's MFMAs below it.  Without it the
// scheduler sinks each read next to its MFMA, and every k-step waits out a full
// LDS round trip.
template <int B, typename F>
__device__ __forceinline__ void pin_after_reads(F (&ta)[B], F (&ya)[B]) {
#pragma unroll
  for (int b = 0; b < B; ++b) asm volatile("" : "+v"(ta[b]), "+v"(ya[b]) : : "memory");
}

// This wave's part of G = [Tr | Ti]^T [Yr | Yi] over the fibres along bond J:
// four interleaved MFMA chains of short FP32 sums, operands read one batch of
// eight k-steps ahead, combined in FP64 into Gp (16 x 16, row-major).  One
// sequential FP32 chain over all fibres measurably degrades the FP32 BP fixed
// point.
template <int D, int W, int J>
__device__ __forceinline__ void gram8(const float* T, const float* Y, double* Gp, const Swz8Lanes<D>& z) {
  constexpr int KPW = Mf8W<D, W>::KPW;
  constexpr int B = KPW < 8 ? KPW : 8;
  const int l = threadIdx.x & 63, i16 = l & 15, kq = l >> 4;
  const int gl = opaque(z.gl[J]);
  f4v g[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) g[q] = f4v{0.f, 0.f, 0.f, 0.f};
  float ta[2][B], ya[2][B];
  lds_batch<D, J, B>(T, Y, gl, 0, ta[0], ya[0]);
#pragma unroll
  for (int s0 = 0; s0 < KPW; s0 += B) {
    const int c = (s0 / B) & 1;
    if (s0 + B < KPW) lds_batch<D, J, B>(T, Y, gl, s0 + B, ta[c ^ 1], ya[c ^ 1]);
    pin_after_reads<B>(ta[c], ya[c]);
#pragma unroll
    for (int b = 0; b < B; ++b) g[b & 3] = mfma4(ta[c][b], ya[c][b], g[b & 3]);
  }
#pragma unroll
  for (int v = 0; v < 4; ++v)
    Gp[(4 * kq + v) * 16 + i16] = ((double)g[0][v] + (double)g[1][v]) + ((double)g[2][v] + (double)g[3][v]);
}

// Complex element ci of a swizzled LDS array.
template <int D>
__device__ __forceinline__ cplx<float> lds_c(const float* F, int ci) {
  const int w = swz<D>(2 * ci);
  return mk(F[w], F[w ^ 1]);
}

// Stage one node tensor (S complex64, contiguous in HBM) into swizzled LDS with
// all threads of the block: float4 block x lands on block swz(4 x) >> 2 with
// element e at position e ^ (swz(4 x) & 3).
template <int D, int W>
__device__ __forceinline__ void load_tensor8(const cplx<float>* __restrict__ Tnode, float* Ts) {
  constexpr int NT = 64 * W;
  const float4* Tg = reinterpret_cast<const float4*>(Tnode);
  float4* T4 = reinterpret_cast<float4*>(Ts);
  const int t = threadIdx.x;
  const int wl = swz<D>(4 * t);
#pragma unroll
  for (int k = 0; k < Mf8<D>::S / 2 / NT; ++k) {
    const float4 v = Tg[t + NT * k];
    const int w = wl ^ swz<D>(4 * NT * k);
    const int xr = w & 3;
    const float b0 = (xr & 1) ? v.y : v.x, b1 = (xr & 1) ? v.x : v.y;
    const float b2 = (xr & 1) ? v.w : v.z, b3 = (xr & 1) ? v.z : v.w;
    T4[w >> 2] = (xr & 2) ? make_float4(b2, b3, b0, b1) : make_float4(b0, b1, b2, b3);
  }
}

// out[c][e] from a 16x16 real Gram stored row-major in Gs.
__device__ __forceinline__ cplx<double> gram_entry64(const double* Gs, int c, int e) {
  return mk(Gs[c * 16 + e] + Gs[(8 + c) * 16 + 8 + e], Gs[c * 16 + 8 + e] - Gs[(8 + c) * 16 + e]);
}

__device__ __forceinline__ cplx<double> gram_trace64(const double* Gs) {
  cplx<double> t = mk(0.0, 0.0);
#pragma unroll
  for (int c = 0; c < 8; ++c) t = t + gram_entry64(Gs, c, c);
  return t;
}

// Sum of the W per-wave partial Grams, entry (r, c).
template <int W>
__device__ __forceinline__ double gsum(const double (*Gp)[256], int r, int c) {
  double v = Gp[0][r * 16 + c];
#pragma unroll
  for (int w = 1; w < W; ++w) v += Gp[w][r * 16 + c];
  return v;
}

template <int W>
__device__ __forceinline__ cplx<double> gram_entry_w(const double (*Gp)[256], int c, int e) {
  return mk(gsum<W>(Gp, c, e) + gsum<W>(Gp, 8 + c, 8 + e), gsum<W>(Gp, c, 8 + e) - gsum<W>(Gp, 8 + c, e));
}

template <int D, int W>
struct Mf8Smem {
  float T[2 * Mf8<D>::S];
  float W1[2 * Mf8<D>::S];
  double Gp[W][256];
};

// Finish one outgoing BP message from the per-wave Grams (first wave only):
// the Hermitian part (the message is Hermitian by construction), normalised by
// its real trace in FP64, stored as R at message position pos, with the
// partial maxima of |new - old| and |new + old| (old: this lane's entry of the
// previous message).  Call between the barrier after the Gram and the next one.
template <int W, typename R = float>
__device__ __forceinline__ void finish8(const double (*Gp)[256], int pos, cplx<R> old, cplx<R>* mout,
                                        double& dnum2, double& dden2) {
  const int l = threadIdx.x & 63, c = l >> 3, e = l & 7;
  // real trace: lanes 0..7 take one diagonal entry each, then a shuffle sum
  double d = l < 8 ? gsum<W>(Gp, l, l) + gsum<W>(Gp, 8 + l, 8 + l) : 0.0;
#pragma unroll
  for (int o = 4; o > 0; o >>= 1) d += __shfl_xor(d, o, 64);
  const double inv = 1.0 / __shfl(d, 0, 64);
  // entry (e, c) from its own lane: the transposed LDS reads had 4-way bank conflicts
  const cplx<double> a = gram_entry_w<W>(Gp, c, e);
  const int tl = e * 8 + c;
  const cplx<double> b = mk(__shfl(a.x, tl, 64), __shfl(a.y, tl, 64));
  const cplx<R> v = cast<R>(mk(0.5 * (a.x + b.x) * inv, 0.5 * (a.y - b.y) * inv));
  // squared magnitudes; the square roots are taken once per block
  const cplx<double> vd = cast<double>(v), od = cast<double>(old);
  dnum2 = fmax(dnum2, abs2(vd - od));
  dden2 = fmax(dden2, abs2(vd + od));
  mout[(size_t)pos * 64 + l] = v;
}

// Lane (i16, kq) of the B operand of message m for k-step s reads m[a][b] with
// a = i16 & 7, b = (4 s + kq) & 7; the real embedding is
// [[m_r^T, m_i^T], [-m_i^T, m_r^T]].
__device__ __forceinline__ void msg_operand8(const cplx<float>* __restrict__ m, float (&bm)[4]) {
  const int l = threadIdx.x & 63, a = l & 7, cn = (l & 15) >> 3, kq = l >> 4;
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    const int ck = (4 * s + kq) >> 3;
    const cplx<float> v = m[a * 8 + ((4 * s + kq) & 7)];
    bm[s] = ck == 0 ? (cn == 0 ? v.x : v.y) : (cn == 0 ? -v.y : v.x);
  }
}

// MODE kModeBP: one BP sweep over the class; kModeDensity: rho and Bloch
// vectors.  One block of W waves per node.
// Barrier between the phases of a node.  A one-wave block needs no hardware
// barrier: its LDS operations complete in issue order, so a compiler fence
// keeps the accesses in program order while VALU and MFMA work may move across
// phases (the wave barrier a block barrier becomes would stop that); wider
// blocks synchronise their waves.
template <int W>
__device__ __forceinline__ void node_sync() {
  if constexpr (W == 1)
    asm volatile("" ::: "memory");
  else
    __syncthreads();
}

// One block's share of a class: nodes start, start + gridDim.x, ...
template <int D, int MODE, int W>
__device__ __forceinline__ void mf8_run(const ClassView& cv, const cplx<float>* __restrict__ T, cplx<float>* msgA,
                                        cplx<float>* msgB, BpStatus* st, double* rho_out, double* bloch_out,
                                        int start) {
  static_assert(D == 2 || D == 3, "degree 2 or 3");
  static_assert(MODE == kModeBP || MODE == kModeDensity, "BP or density");
  __shared__ __align__(16) Mf8Smem<D, W> sm;
  const int n = cv.n;
  const int l = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const bool lead = W == 1 || wave == 0;  // constant for one wave: no branch around the finish
  cplx<float>* min = msgA;
  cplx<float>* mout = msgB;
  if (MODE == kModeBP) {
    const int cur = bp_begin(st, cv.sweep);
    if (cur < 0) return;
    min = cur ? msgB : msgA;
    mout = cur ? msgA : msgB;
  }
  double dnum = 0.0, dden = 0.0;
  constexpr int S = Mf8<D>::S;
  const Swz8Lanes<D> zl = swz8_lanes<D, W>();
  for (int i = start; i < n; i += gridDim.x) {
    load_tensor8<D, W>(T + (size_t)i * S, sm.T);
    float bm[D][4];
#pragma unroll
    for (int j = 0; j < D; ++j) msg_operand8(min + (size_t)cv.in_pos[j * n + i] * 64, bm[j]);
    int pos[D];
    cplx<float> old[D];
    if (MODE == kModeBP && lead) {
#pragma unroll
      for (int k = 0; k < D; ++k) {
        pos[k] = cv.out_pos[k * n + i];
        old[k] = min[(size_t)pos[k] * 64 + l];
      }
    }
    node_sync<W>();
    if (MODE == kModeDensity) {
      mp8<D, W, 0>(sm.T, sm.W1, bm[0], zl);
      node_sync<W>();
      mp8<D, W, 1>(sm.W1, sm.W1, bm[1], zl);
      node_sync<W>();
      if constexpr (D == 3) {
        mp8<D, W, 2>(sm.W1, sm.W1, bm[2], zl);
        node_sync<W>();
      }
      // rho[p][q] = sum_a TM[p, a] conj(T[q, a]), the a range split over the block
      constexpr int A = Mf8<D>::A;
      cplx<float> r[4] = {mk(0.f, 0.f), mk(0.f, 0.f), mk(0.f, 0.f), mk(0.f, 0.f)};
      for (int a = threadIdx.x; a < A; a += 64 * W) {
        const cplx<float> t0 = lds_c<D>(sm.T, a), t1 = lds_c<D>(sm.T, A + a);
        const cplx<float> m0 = lds_c<D>(sm.W1, a), m1 = lds_c<D>(sm.W1, A + a);
        cfma_cj(r[0], t0, m0);
        cfma_cj(r[1], t1, m0);
        cfma_cj(r[2], t0, m1);
        cfma_cj(r[3], t1, m1);
      }
#pragma unroll
      for (int q = 0; q < 4; ++q)
        for (int o = 32; o > 0; o >>= 1) {
          r[q].x += __shfl_xor(r[q].x, o, 64);
          r[q].y += __shfl_xor(r[q].y, o, 64);
        }
      if (l == 0) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          sm.Gp[wave][2 * q] = r[q].x;
          sm.Gp[wave][2 * q + 1] = r[q].y;
        }
      }
      node_sync<W>();
      if (threadIdx.x == 0) {
        cplx<double> rr[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          rr[q] = mk(0.0, 0.0);
#pragma unroll
          for (int w = 0; w < W; ++w) rr[q] = rr[q] + mk(sm.Gp[w][2 * q], sm.Gp[w][2 * q + 1]);
        }
        const cplx<double> inv = cinv(rr[0] + rr[3]);
        const cplx<double> r00 = rr[0] * inv, r01 = rr[1] * inv, r10 = rr[2] * inv, r11 = rr[3] * inv;
        const long id = cv.node_ids[i];
        if (rho_out) {
          double* ro = rho_out + id * 8;
          ro[0] = r00.x; ro[1] = r00.y; ro[2] = r01.x; ro[3] = r01.y;
          ro[4] = r10.x; ro[5] = r10.y; ro[6] = r11.x; ro[7] = r11.y;
        }
        if (bloch_out) {
          double* bo = bloch_out + id * 3;
          bo[0] = r01.x + r10.x;
          bo[1] = r10.y - r01.y;
          bo[2] = r00.x - r11.x;
        }
      }
      node_sync<W>();
      continue;
    }
    // Each gram8 is followed by a barrier; the first wave then finishes the
    // message while the others start the next mode product, whose own barrier
    // keeps the partial Grams alive until the first wave has read them.  One
    // work buffer (mode products run in place along their bond): degree 3
    // computes T x2 m2 twice, which costs one mode product of 32 MFMAs and
    // saves 8 KB of LDS per node, so 8 nodes fit per CU instead of 6.
    if constexpr (D == 2) {
      mp8<D, W, 1>(sm.T, sm.W1, bm[1], zl);          // Y0 = T x1 m1
      node_sync<W>();
      gram8<D, W, 0>(sm.T, sm.W1, sm.Gp[wave], zl);
      node_sync<W>();
      if (lead) finish8<W>(sm.Gp, pos[0], old[0], mout, dnum, dden);
      mp8<D, W, 0>(sm.T, sm.W1, bm[0], zl);          // Y1 = T x0 m0
      node_sync<W>();
      gram8<D, W, 1>(sm.T, sm.W1, sm.Gp[wave], zl);
      node_sync<W>();
      if (lead) finish8<W>(sm.Gp, pos[1], old[1], mout, dnum, dden);
    } else {
      mp8<D, W, 2>(sm.T, sm.W1, bm[2], zl);          // W1 = T x2 m2
      node_sync<W>();
      mp8<D, W, 1>(sm.W1, sm.W1, bm[1], zl);         // Y0 = W1 x1 m1
      node_sync<W>();
      gram8<D, W, 0>(sm.T, sm.W1, sm.Gp[wave], zl);
      node_sync<W>();
      if (lead) finish8<W>(sm.Gp, pos[0], old[0], mout, dnum, dden);
      mp8<D, W, 2>(sm.T, sm.W1, bm[2], zl);          // W1 = T x2 m2
      node_sync<W>();
      mp8<D, W, 0>(sm.W1, sm.W1, bm[0], zl);         // Y1 = W1 x0 m0
      node_sync<W>();
      gram8<D, W, 1>(sm.T, sm.W1, sm.Gp[wave], zl);
      node_sync<W>();
      if (lead) finish8<W>(sm.Gp, pos[1], old[1], mout, dnum, dden);
      mp8<D, W, 0>(sm.T, sm.W1, bm[0], zl);          // W1 = T x0 m0
      node_sync<W>();
      mp8<D, W, 1>(sm.W1, sm.W1, bm[1], zl);         // Y2 = W1 x1 m1
      node_sync<W>();
      gram8<D, W, 2>(sm.T, sm.W1, sm.Gp[wave], zl);
      node_sync<W>();
      if (lead) finish8<W>(sm.Gp, pos[2], old[2], mout, dnum, dden);
    }
    // the next node's tensor load overwrites T, which the last Gram read
    node_sync<W>();
  }
  if (MODE == kModeBP && lead && start < n) {  // a block without nodes here adds nothing to the maxima
    for (int o = 32; o > 0; o >>= 1) {
      dnum = fmax(dnum, __shfl_xor(dnum, o, 64));
      dden = fmax(dden, __shfl_xor(dden, o, 64));
    }
    dnum = sqrt(dnum);  // the partial maxima were of squared magnitudes
    dden = sqrt(dden);
    if (l == 0) {
      bp_end(st, cv.sweep, dnum, dden);
    }
  }
}

template <int D, int MODE, int W>
__global__ __launch_bounds__(64 * W) void k_mfma8(ClassView cv, const cplx<float>* __restrict__ T,
                                               cplx<float>* msgA, cplx<float>* msgB, BpStatus* st,
                                               double* rho_out, double* bloch_out) {
  mf8_run<D, MODE, W>(cv, T, msgA, msgB, st, rho_out, bloch_out, blockIdx.x);
}

// One BP sweep of the chi = 8 classes of several FP32 jobs in one launch, nodes shared over
// the blocks as in k_small_multi.
template <int D>
__global__ __launch_bounds__(64) void k_mfma8_multi(const BpEntry* __restrict__ tab, int entries, int sweep) {
  for (int e = 0; e < entries; ++e) {
    ClassView cv = tab[e].cv;
    cv.sweep = sweep;
    const long g = gridDim.x;
    const int start = (int)((((long)blockIdx.x - tab[e].prefix) % g + g) % g);
    if (start >= cv.n) continue;  // none of this entry's nodes falls to this block
    mf8_run<D, kModeBP, 1>(cv, reinterpret_cast<const cplx<float>*>(tab[e].T),
                           reinterpret_cast<cplx<float>*>(tab[e].mA), reinterpret_cast<cplx<float>*>(tab[e].mB),
                           tab[e].st, nullptr, nullptr, start);
  }
}

}  // namespace bqa

namespace bqa {

// ---------------------------------------------------------------------------
// Extended messages in FP64 arithmetic (FP32 or FP64 inputs), chi = 8.
//
// FP32 rounding inside the extended-message contraction is amplified by the
// S^{-1/2} of the canonicalisers on small Schmidt values; computed in FP64
// from the stored tensors and messages, the FP32 path's error drops to the
// level set by the storage precision.  v_mfma_f64_16x16x4f64 has the same A
// and B lane layout as the FP32 instruction; its result register v of lane l
// holds D[l / 16 + 4 v][l % 16].
typedef double d4v __attribute__((ext_vector_type(4)));

__device__ __forceinline__ d4v mfma4d(double a, double b, d4v c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// FP64 16x16x4 MFMA lane layout (checked on gfx950): A[i][k] and B[k][j] as in
// FP32, D[l / 16 + 4 v][l % 16] in result register v.  RS: the row stride of m
// in complex elements (8 for a message; 9 for an LDS table, where rows 8
// elements apart would put the 8 row reads of a k-step on 2 bank groups).
template <int RS = 8, typename R>
__device__ __forceinline__ void msg_operand8d(const cplx<R>* __restrict__ m, double (&bm)[4]) {
  const int l = threadIdx.x & 63, i16 = l & 15, kq = l >> 4;
  const int a = i16 & 7, cn = i16 >> 3;
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    const int kr = 4 * s + kq;
    const int b = kr & 7, ck = kr >> 3;
    const cplx<R> v = m[a * RS + b];
    const double vr = (double)v.x, vi = (double)v.y;
    bm[s] = ck == 0 ? (cn == 0 ? vr : vi) : (cn == 0 ? -vi : vr);
  }
}

// msg_operand8d in two halves, so the loads can be issued long before their
// values are turned into MFMA operands: the raw message entries this lane needs,
// then the real-embedded operand from them.
__device__ __forceinline__ void msg_raw8d(const cplx<double>* __restrict__ m, double2 (&raw)[4]) {
  const int l = threadIdx.x & 63, a = l & 7, kq = l >> 4;
  const double2* m2 = reinterpret_cast<const double2*>(m);
#pragma unroll
  for (int s = 0; s < 4; ++s) raw[s] = m2[a * 8 + ((4 * s + kq) & 7)];
}

__device__ __forceinline__ void msg_operand_from_raw8d(const double2 (&raw)[4], double (&bm)[4]) {
  const int l = threadIdx.x & 63, cn = (l & 15) >> 3, kq = l >> 4;
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    const int ck = (4 * s + kq) >> 3;
    // the empty asm keeps both components in registers: a lane-dependent choice
    // between them would otherwise become an indexed access to a stack copy
    double re = raw[s].x, im = raw[s].y;
    asm volatile("" : "+v"(re), "+v"(im));
    bm[s] = ck == 0 ? (cn == 0 ? re : im) : (cn == 0 ? -im : re);
  }
}

// LDS double-index swizzle of the FP64 kernels' tensor buffers (natural index
// 2 x element + component): bit j (j < 5) of a stored index is flipped by the
// parity of (index & swz64_mask<D>(j)).  Linear over GF(2) like swz, so an
// index built from disjoint bit fields swizzles field by field.  Every
// mode-product read and write and every Gram read along each bond touches
// distinct banks for 64 banks per 32 lanes and for 32 banks per 16 lanes of a
// 64-bit access (the natural layout had up to 16-way conflicts).  No mask
// contains bit 0, so the two components of an element stay one aligned pair,
// swapped where bit 0 is flipped.
template <int D>
__host__ __device__ constexpr int swz64_mask(int j) {
  constexpr int m3[5] = {0x278, 0x6f4, 0x1a8, 0xc0, 0x740};
  constexpr int m2[5] = {0x48, 0xb4, 0x58, 0xc0, 0xe0};
  return D == 3 ? m3[j] : m2[j];
}

template <int D>
__host__ __device__ constexpr int swz64(int a) {
  int out = a;
  for (int j = 0; j < 5; ++j) out ^= (__builtin_popcount(swz64_mask<D>(j) & a) & 1) << j;
  return out;
}

__device__ __forceinline__ double lds_ld64(const double* base, int byte_off) {
  return *reinterpret_cast<const double*>(reinterpret_cast<const char*>(base) + byte_off);
}

__device__ __forceinline__ void lds_st64(double* base, int byte_off, double v) {
  *reinterpret_cast<double*>(reinterpret_cast<char*>(base) + byte_off) = v;
}

// Complex element with swizzled pair index sw = swz64(2 x) of a tensor buffer.
__device__ __forceinline__ cplx<double> ld_pair64(const double* buf, int sw) {
  const double2 p = *reinterpret_cast<const double2*>(buf + (sw & ~1));
  return (sw & 1) ? mk(p.y, p.x) : mk(p.x, p.y);
}

__device__ __forceinline__ void st_pair64(double* buf, int sw, cplx<double> v) {
  *reinterpret_cast<double2*>(buf + (sw & ~1)) = (sw & 1) ? make_double2(v.y, v.x) : make_double2(v.x, v.y);
}

// Per-lane (and per-wave) parts of the FP64 kernels' swizzled LDS indices
// along each bond, as byte offsets; the remaining parts of every index are
// compile-time constants on disjoint bits of the natural index, so each access
// is one XOR with a folded constant.  Computed once per kernel and passed
// through opaque() at each use (see Swz8Lanes).
template <int D>
struct Lin8Lanes {
  int rd[D], wr[D], gl[D];
};

template <int D, int W, int J>
__device__ __forceinline__ void lin8_lanes_j(Lin8Lanes<D>& z) {
  constexpr int st = cpow8(D - 1 - J);
  constexpr int TPW = Mf8<D>::TILES / W, KPW = Mf8<D>::KS / W;
  const int l = threadIdx.x & 63, i16 = l & 15, kq = l >> 4, w = threadIdx.x >> 6;
  // A[i][k]: fibre (w TPW + t) 16 + i, element (4 s + kq) & 7 along J, component (4 s + kq) >> 3
  z.rd[J] = 8 * swz64<D>(2 * (fbase8<D, J>(w * TPW * 16 + i16) + kq * st));
  // D[kq + 4 v][i] (FP64 layout): fibre (w TPW + t) 16 + kq + 4 v, element i & 7, component i >> 3
  z.wr[J] = 8 * swz64<D>(2 * (fbase8<D, J>(w * TPW * 16 + kq) + (i16 & 7) * st) + (i16 >> 3));
  // Gram operands: fibre 4 (w KPW + s) + kq, element i & 7 along J, component i >> 3
  z.gl[J] = 8 * swz64<D>(2 * (fbase8<D, J>(4 * w * KPW + kq) + (i16 & 7) * st) + (i16 >> 3));
}

// Element x = t + NT k of a tensor buffer (t < NT, NT a power of two): the
// swizzled pair index from the thread's part and the loop constant.
template <int D, int NT>
__device__ __forceinline__ int pair64(int t_sw, int k) {
  return t_sw ^ swz64<D>(2 * NT * k);
}

template <int D, int W>
__device__ __forceinline__ Lin8Lanes<D> lin8_lanes() {
  Lin8Lanes<D> z;
  lin8_lanes_j<D, W, 0>(z);
  lin8_lanes_j<D, W, 1>(z);
  if constexpr (D == 3) lin8_lanes_j<D, W, 2>(z);
  return z;
}

// Y = X x_J m over this wave's TPW tiles (double views of complex128 LDS
// arrays); with P, tiles in the p = 1 half take the operand b1.  All A operands
// of the wave's tiles are read before the interleaved MFMAs and every read
// precedes every write, so X may equal Y (the waves' tiles are disjoint).
template <int D, int W, int J, bool P>
__device__ __forceinline__ void mp8dw(const double* X, double* Y, const double (&b0)[4], const double (&b1)[4],
                                      const Lin8Lanes<D>& z) {
  constexpr int st = cpow8(D - 1 - J);
  constexpr int TPW = Mf8<D>::TILES / W;
  const int w = threadIdx.x >> 6;
  const int rd = opaque(z.rd[J]), wr = opaque(z.wr[J]);
  double a[TPW][4];
#pragma unroll
  for (int tt = 0; tt < TPW; ++tt)
#pragma unroll
    for (int s = 0; s < 4; ++s)
      a[tt][s] = lds_ld64(X, rd ^ (8 * swz64<D>(2 * (fbase8<D, J>(tt * 16) + 4 * (s & 1) * st) + (s >> 1))));
  d4v acc[TPW];
#pragma unroll
  for (int tt = 0; tt < TPW; ++tt) acc[tt] = d4v{0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int s = 0; s < 4; ++s)
#pragma unroll
    for (int tt = 0; tt < TPW; ++tt) {
      const bool p1 = P && (w * TPW + tt) >= Mf8<D>::TILES / 2;
      acc[tt] = mfma4d(a[tt][s], p1 ? b1[s] : b0[s], acc[tt]);
    }
#pragma unroll
  for (int tt = 0; tt < TPW; ++tt)
#pragma unroll
    for (int v = 0; v < 4; ++v) lds_st64(Y, wr ^ (8 * swz64<D>(2 * fbase8<D, J>(tt * 16 + 4 * v))), acc[tt][v]);
}

// This wave's k-steps of the p-resolved Gram along bond J (fibres ordered with
// p slowest, so k-step s belongs to p = s / (KS / 2)): the p = 0 and p = 1
// partials, each from two interleaved chains (with W >= 2 each wave's k-steps
// lie in one half and the p = 1 register partial stays zero).
template <int D, int W, int J>
__device__ __forceinline__ void gram8dw(const double* T, const double* Y, d4v& g0, d4v& g1,
                                        const Lin8Lanes<D>& z) {
  constexpr int KS = Mf8<D>::KS, KPW = KS / W;
  constexpr int B = KPW < 8 ? KPW : 8;
  const int gl = opaque(z.gl[J]);
  d4v c[2][2] = {{d4v{0.0, 0.0, 0.0, 0.0}, d4v{0.0, 0.0, 0.0, 0.0}},
                 {d4v{0.0, 0.0, 0.0, 0.0}, d4v{0.0, 0.0, 0.0, 0.0}}};
#pragma unroll
  for (int s0 = 0; s0 < KPW; s0 += B) {
    double ta[B], ya[B];
#pragma unroll
    for (int b = 0; b < B; ++b) {
      const int off = gl ^ (8 * swz64<D>(2 * fbase8<D, J>(4 * (s0 + b))));
      ta[b] = lds_ld64(T, off);
      ya[b] = lds_ld64(Y, off);
    }
#pragma unroll
    for (int b = 0; b < B; ++b) {
      const int p = W == 1 ? (s0 + b) / (KS / 2) : 0;
      c[p][b & 1] = mfma4d(ta[b], ya[b], c[p][b & 1]);
    }
  }
  g0 = c[0][0] + c[0][1];
  g1 = c[1][0] + c[1][1];
}

// Store this wave's Gram partial(s) (16 x 16, row-major) into Gp, one slot per
// wave, or slots 0 and 1 (p = 0, 1) for a single wave.
template <int W>
__device__ __forceinline__ void store_gd(double* Gp, const d4v& g0, const d4v& g1) {
  const int l = threadIdx.x & 63, i16 = l & 15, kq = l >> 4, w = threadIdx.x >> 6;
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    Gp[w * 256 + (kq + 4 * v) * 16 + i16] = g0[v] + g1[v];
    if (W == 1) {
      Gp[(kq + 4 * v) * 16 + i16] = g0[v];
      Gp[256 + (kq + 4 * v) * 16 + i16] = g1[v];
    }
  }
}

// Extended message (2chi x 2chi, FP64) at position pos from the partial Grams
// in Gp, all threads of the block writing.  Partial slots [0, H) hold p = 0 and
// [H, 2H) hold p = 1 (H = W / 2; one slot each for a single wave); they are
// first summed into slots 0 and H.
template <int W>
__device__ __forceinline__ void finish8_ext64(double* Gp, long pos, cplx<double>* ext_out, cplx<double> g0c,
                                              cplx<double> g1c) {
  constexpr int H = W == 1 ? 1 : W / 2;
  const int t = threadIdx.x, l = t & 63;
  if constexpr (H > 1) {
    for (int o = t; o < 256; o += 64 * W) {
      double a = Gp[o], b = Gp[H * 256 + o];
#pragma unroll
      for (int q = 1; q < H; ++q) {
        a += Gp[q * 256 + o];
        b += Gp[(H + q) * 256 + o];
      }
      Gp[o] = a;
      Gp[H * 256 + o] = b;
    }
    __syncthreads();
  }
  const double* G0 = Gp;
  const double* G1 = Gp + H * 256;
  // trace over both p blocks: lanes 0..7 of every wave, then a shuffle sum
  cplx<double> d = mk(0.0, 0.0);
  if (l < 8) d = gram_entry64(G0, l, l) + gram_entry64(G1, l, l);
#pragma unroll
  for (int o = 4; o > 0; o >>= 1) {
    d.x += __shfl_xor(d.x, o, 64);
    d.y += __shfl_xor(d.y, o, 64);
  }
  const cplx<double> tr = mk(__shfl(d.x, 0, 64), __shfl(d.y, 0, 64));
  const cplx<double> inv = cinv(tr * mk(abs2(g0c) + abs2(g1c), 0.0));
  for (int o = t; o < 256; o += 64 * W) {
    const int row = o >> 4, col = o & 15;
    const int s = row >> 3, c = row & 7, s2 = col >> 3, e = col & 7;
    const cplx<double> a0 = gram_entry64(G0, c, e), a1 = gram_entry64(G1, c, e);
    const cplx<double> gsum = ((s + s2) & 1) ? a0 - a1 : a0 + a1;
    const cplx<double> ga = s ? g1c : g0c, gb = s2 ? g1c : g0c;
    ext_out[pos * 256 + o] = conj(ga) * gb * gsum * inv;
  }
}

// Work buffers of the extended-message kernel; for degree 3 the Gram partials
// (at most 4 x 256 doubles) go into the dead Y buffer, for degree 2 (Y holds 256
// doubles) into G.
template <int D>
struct Mf8SmemD {
  double T[2 * Mf8<D>::S];
  double W1[2 * Mf8<D>::S];
  double W2[2 * Mf8<D>::S];
  double G[D == 2 ? 512 : 1];
};

// One p-resolved Gram and its extended message; Gp receives the partials once
// every wave has finished reading Y (Gp may be Y).
// (edge: the lambda position of the bond, which indexes the ZZ-split gates cv.zz)
template <int D, int W, int J>
__device__ __forceinline__ void ext_out8(const double* T, const double* Y, double* Gp, const ClassView& cv,
                                         int i, cplx<double>* ext_out, const Lin8Lanes<D>& z) {
  d4v g0, g1;
  gram8dw<D, W, J>(T, Y, g0, g1, z);
  __syncthreads();
  store_gd<W>(Gp, g0, g1);
  __syncthreads();
  const cplx<double>* zg = cv.zz + 2 * (size_t)cv.lmbd_pos[J * cv.n + i];
  finish8_ext64<W>(Gp, cv.out_pos[J * cv.n + i], ext_out, zg[0], zg[1]);
  __syncthreads();
}

// Extended messages in FP64 arithmetic (FP32 or FP64 inputs), chi = 8, one
// block of W waves per node; the ZZ-split gates come from cv.zz.
template <int D, int W, typename R>
__global__ __launch_bounds__(64 * W) void k_mfma8_ext64(ClassView cv, const cplx<R>* __restrict__ T,
                                                        const cplx<R>* __restrict__ msgs, cplx<double>* ext_out) {
  static_assert(D == 2 || D == 3, "degree 2 or 3");
  static_assert(Mf8<D>::TILES % W == 0 && (W == 1 || W % 2 == 0), "wave split");
  __shared__ __align__(16) Mf8SmemD<D> sm;
  const int n = cv.n;
  constexpr int S = Mf8<D>::S;
  const Lin8Lanes<D> zl = lin8_lanes<D, W>();
  static_assert(S % (64 * W) == 0, "tensor load in whole rounds");
  const int tsw = swz64<D>(2 * (int)threadIdx.x);
  for (int i = blockIdx.x; i < n; i += gridDim.x) {
    const cplx<R>* Tg = T + (size_t)i * S;
#pragma unroll
    for (int k = 0; k < S / (64 * W); ++k)
      st_pair64(sm.T, pair64<D, 64 * W>(tsw, k), cast<double>(Tg[threadIdx.x + 64 * W * k]));
    double bm[D][4];
#pragma unroll
    for (int j = 0; j < D; ++j) msg_operand8d(msgs + (size_t)cv.in_pos[j * n + i] * 64, bm[j]);
    __syncthreads();
    if constexpr (D == 2) {
      mp8dw<D, W, 1, false>(sm.T, sm.W1, bm[1], bm[1], zl);
      __syncthreads();
      ext_out8<D, W, 0>(sm.T, sm.W1, sm.G, cv, i, ext_out, zl);
      mp8dw<D, W, 0, false>(sm.T, sm.W2, bm[0], bm[0], zl);
      __syncthreads();
      ext_out8<D, W, 1>(sm.T, sm.W2, sm.G, cv, i, ext_out, zl);
    } else {
      mp8dw<D, W, 2, false>(sm.T, sm.W1, bm[2], bm[2], zl);
      __syncthreads();
      mp8dw<D, W, 1, false>(sm.W1, sm.W2, bm[1], bm[1], zl);
      __syncthreads();
      ext_out8<D, W, 0>(sm.T, sm.W2, sm.W2, cv, i, ext_out, zl);
      mp8dw<D, W, 0, false>(sm.W1, sm.W2, bm[0], bm[0], zl);
      __syncthreads();
      ext_out8<D, W, 1>(sm.T, sm.W2, sm.W2, cv, i, ext_out, zl);
      mp8dw<D, W, 0, false>(sm.T, sm.W1, bm[0], bm[0], zl);
      __syncthreads();
      mp8dw<D, W, 1, false>(sm.W1, sm.W1, bm[1], bm[1], zl);
      __syncthreads();
      ext_out8<D, W, 2>(sm.T, sm.W1, sm.W1, cv, i, ext_out, zl);
    }
  }
}

// ---------------------------------------------------------------------------
// BP sweep of an FP64 job, chi = 8, degree 2: per output, one mode product and
// one Gram on v_mfma_f64_16x16x4f64 over the swizzled FP64 layout of the
// extended-message kernel (degree 3 runs k_mfma8_bp64_3p below).  The first
// wave finishes the message while the others start the next mode product,
// whose barrier keeps the partials alive until it has read them.
template <int D, int W>
struct Mf8Bp64Smem {
  double T[2 * Mf8<D>::S];
  double W1[2 * Mf8<D>::S];
  double Gp[W][256];
};

// This wave's Gram partial along bond J into its slot of Gp (FP64 D layout).
template <int D, int W, int J>
__device__ __forceinline__ void gram8_bp64(const double* T, const double* Y, double (*Gp)[256],
                                           const Lin8Lanes<D>& z) {
  d4v g0, g1;
  gram8dw<D, W, J>(T, Y, g0, g1, z);
  const int l = threadIdx.x & 63, i16 = l & 15, kq = l >> 4, w = threadIdx.x >> 6;
#pragma unroll
  for (int v = 0; v < 4; ++v) Gp[w][(kq + 4 * v) * 16 + i16] = g0[v] + g1[v];
}

template <int D, int W>
__global__ __launch_bounds__(64 * W) void k_mfma8_bp64(ClassView cv, const cplx<double>* __restrict__ T,
                                                       cplx<double>* msgA, cplx<double>* msgB, BpStatus* st) {
  static_assert(D == 2, "degree 2 (degree 3: k_mfma8_bp64_3p)");
  static_assert(Mf8<D>::TILES % W == 0 && Mf8<D>::KS % W == 0, "tiles and k-steps split evenly over the waves");
  static_assert(Mf8<D>::S % (64 * W) == 0, "tensor load in whole rounds");
  __shared__ __align__(16) Mf8Bp64Smem<D, W> sm;
  const int cur = bp_begin(st, cv.sweep);
  if (cur < 0) return;
  const int n = cv.n;
  const int l = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const bool lead = W == 1 || wave == 0;
  const cplx<double>* min = cur ? msgB : msgA;
  cplx<double>* mout = cur ? msgA : msgB;
  double dnum = 0.0, dden = 0.0;
  constexpr int S = Mf8<D>::S;
  const Lin8Lanes<D> zl = lin8_lanes<D, W>();
  const int tsw = swz64<D>(2 * (int)threadIdx.x);
  for (int i = blockIdx.x; i < n; i += gridDim.x) {
    const cplx<double>* Tg = T + (size_t)i * S;
#pragma unroll
    for (int k = 0; k < S / (64 * W); ++k) st_pair64(sm.T, pair64<D, 64 * W>(tsw, k), Tg[threadIdx.x + 64 * W * k]);
    double bm[D][4];
#pragma unroll
    for (int j = 0; j < D; ++j) msg_operand8d(min + (size_t)cv.in_pos[j * n + i] * 64, bm[j]);
    int pos[D];
    cplx<double> old[D];
    if (lead) {
#pragma unroll
      for (int k = 0; k < D; ++k) {
        pos[k] = cv.out_pos[k * n + i];
        old[k] = min[(size_t)pos[k] * 64 + l];
      }
    }
    node_sync<W>();
    mp8dw<D, W, 1, false>(sm.T, sm.W1, bm[1], bm[1], zl);   // Y0 = T x1 m1
    node_sync<W>();
    gram8_bp64<D, W, 0>(sm.T, sm.W1, sm.Gp, zl);
    node_sync<W>();
    if (lead) finish8<W, double>(sm.Gp, pos[0], old[0], mout, dnum, dden);
    mp8dw<D, W, 0, false>(sm.T, sm.W1, bm[0], bm[0], zl);   // Y1 = T x0 m0
    node_sync<W>();
    gram8_bp64<D, W, 1>(sm.T, sm.W1, sm.Gp, zl);
    node_sync<W>();
    if (lead) finish8<W, double>(sm.Gp, pos[1], old[1], mout, dnum, dden);
    // the next node's tensor load overwrites T, which the last Gram read
    node_sync<W>();
  }
  if (lead) {
    for (int o = 32; o > 0; o >>= 1) {
      dnum = fmax(dnum, __shfl_xor(dnum, o, 64));
      dden = fmax(dden, __shfl_xor(dden, o, 64));
    }
    dnum = sqrt(dnum);  // the partial maxima were of squared magnitudes
    dden = sqrt(dden);
    if (l == 0) {
      bp_end(st, cv.sweep, dnum, dden);
    }
  }
}

// ---------------------------------------------------------------------------
// BP sweep of an FP64 job, chi = 8, degree 3, in six contractions per node
// instead of the reference schedule's nine (backends.py:381-404 applies d - 1
// messages to the ket side of each output).  The incoming messages are
// Hermitian (every BP kernel stores the Hermitian part, the gauge writes real
// diagonals), so a message on the bra side can be moved onto the ket side of a
// conjugated factor: sum_a conj(T[..a..]) m[a][b] = conj(Y[..b..]) with
// Y = T x m.  With Y_j = T x_j m_j for each bond j,
//   out_0 = sum conj(Y_2) Y_1 over bonds 1, 2;  out_1 = sum conj(Y_2) Y_0 over bonds 0, 2;
//   out_2 = sum conj(Y_1) Y_0 over bonds 0, 1,
// three mode products and three Grams.  LDS holds T (then Y_2 in place), Y_0
// and Y_1: 48 KB of tensors plus the Gram partials per node, so two blocks of
// four waves per CU (100k qubits: 1.21 ms per sweep against 1.35 ms for the
// nine-contraction schedule with two waves per node; 1.28 ms without the
// register prefetch below).
// Two sets of Gram partials alternate, so a Gram may start while the first wave still
// finishes the message of the previous one (64 KB per block, two blocks per CU).
template <int W>
struct Mf8Bp64S3Smem {
  double T[2 * Mf8<3>::S];
  double Y0[2 * Mf8<3>::S];
  double Y1[2 * Mf8<3>::S];
  double Gp[2][W][256];
};

template <int W>
__global__ __launch_bounds__(64 * W) void k_mfma8_bp64_3p(ClassView cv, const cplx<double>* __restrict__ T,
                                                          cplx<double>* msgA, cplx<double>* msgB, BpStatus* st) {
  constexpr int D = 3;
  static_assert(Mf8<D>::TILES % W == 0 && Mf8<D>::KS % W == 0, "tiles and k-steps split evenly over the waves");
  static_assert(Mf8<D>::S % (64 * W) == 0, "tensor load in whole rounds");
  __shared__ __align__(16) Mf8Bp64S3Smem<W> sm;
  const int cur = bp_begin(st, cv.sweep);
  if (cur < 0) return;
  const int n = cv.n;
  const int l = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const bool lead = W == 1 || wave == 0;
  const cplx<double>* min = cur ? msgB : msgA;
  cplx<double>* mout = cur ? msgA : msgB;
  double dnum = 0.0, dden = 0.0;
  constexpr int S = Mf8<D>::S;
  const Lin8Lanes<D> zl = lin8_lanes<D, W>();
  const int tsw = swz64<D>(2 * (int)threadIdx.x);
  // Everything a node reads from HBM (its tensor, the three incoming messages as
  // MFMA operands, the old messages at its outputs) is loaded into registers one
  // node ahead, while the previous node computes, and the message positions two
  // nodes ahead, so no load waits on another load issued in the same node: with
  // three or fewer nodes per CU the latency of these loads is as long as a
  // node's arithmetic.
  constexpr int TPL = S / (64 * W);       // complex tensor elements per thread
  const int G = gridDim.x;
  // registers of the next node (nt, nm, nold, npos) and the positions of the one after it (pin, pout)
  cplx<double> nt[TPL], nold[D];
  double2 nm[D][4];
  int npos[D], pin[D], pout[D];
  if (blockIdx.x < n) {
    const int i0 = blockIdx.x;
#pragma unroll
    for (int k = 0; k < TPL; ++k) nt[k] = T[(size_t)i0 * S + threadIdx.x + 64 * W * k];
#pragma unroll
    for (int k = 0; k < D; ++k) {
      msg_raw8d(min + (size_t)cv.in_pos[k * n + i0] * 64, nm[k]);
      npos[k] = cv.out_pos[k * n + i0];
      nold[k] = lead ? min[(size_t)npos[k] * 64 + l] : mk(0.0, 0.0);
    }
  }
  if (blockIdx.x + G < n) {
#pragma unroll
    for (int k = 0; k < D; ++k) {
      pin[k] = cv.in_pos[k * n + blockIdx.x + G];
      pout[k] = cv.out_pos[k * n + blockIdx.x + G];
    }
  }
  for (int i = blockIdx.x; i < n; i += G) {
#pragma unroll
    for (int k = 0; k < TPL; ++k) st_pair64(sm.T, pair64<D, 64 * W>(tsw, k), nt[k]);
    double bm[D][4];
    int pos[D];
    cplx<double> old[D];
#pragma unroll
    for (int k = 0; k < D; ++k) {
      msg_operand_from_raw8d(nm[k], bm[k]);
      pos[k] = npos[k];
      old[k] = nold[k];
    }
    const int i1 = i + G, i2 = i + 2 * G;
    if (i1 < n) {
#pragma unroll
      for (int k = 0; k < TPL; ++k) nt[k] = T[(size_t)i1 * S + threadIdx.x + 64 * W * k];
#pragma unroll
      for (int k = 0; k < D; ++k) {
        msg_raw8d(min + (size_t)pin[k] * 64, nm[k]);
        npos[k] = pout[k];
        nold[k] = lead ? min[(size_t)pout[k] * 64 + l] : mk(0.0, 0.0);
      }
    }
    if (i2 < n) {
#pragma unroll
      for (int k = 0; k < D; ++k) {
        pin[k] = cv.in_pos[k * n + i2];
        pout[k] = cv.out_pos[k * n + i2];
      }
    }
    node_sync<W>();
    mp8dw<D, W, 0, false>(sm.T, sm.Y0, bm[0], bm[0], zl);   // Y0 = T x0 m0
    mp8dw<D, W, 1, false>(sm.T, sm.Y1, bm[1], bm[1], zl);   // Y1 = T x1 m1
    node_sync<W>();
    gram8_bp64<D, W, 2>(sm.Y1, sm.Y0, sm.Gp[0], zl);        // out_2: conj(Y1) Y0
    node_sync<W>();
    if (lead) finish8<W, double>(sm.Gp[0], pos[2], old[2], mout, dnum, dden);
    mp8dw<D, W, 2, false>(sm.T, sm.T, bm[2], bm[2], zl);    // T <- Y2 = T x2 m2, in place
    node_sync<W>();
    gram8_bp64<D, W, 0>(sm.T, sm.Y1, sm.Gp[1], zl);         // out_0: conj(Y2) Y1
    node_sync<W>();
    if (lead) finish8<W, double>(sm.Gp[1], pos[0], old[0], mout, dnum, dden);
    gram8_bp64<D, W, 1>(sm.T, sm.Y0, sm.Gp[0], zl);         // out_1: conj(Y2) Y0 (the other partials)
    node_sync<W>();
    if (lead) finish8<W, double>(sm.Gp[0], pos[1], old[1], mout, dnum, dden);
    // the next node's tensor load overwrites T, which the last Gram read
    node_sync<W>();
  }
  if (lead) {
    for (int o = 32; o > 0; o >>= 1) {
      dnum = fmax(dnum, __shfl_xor(dnum, o, 64));
      dden = fmax(dden, __shfl_xor(dden, o, 64));
    }
    dnum = sqrt(dnum);  // the partial maxima were of squared magnitudes
    dden = sqrt(dden);
    if (l == 0) {
      bp_end(st, cv.sweep, dnum, dden);
    }
  }
}

}  // namespace bqa

namespace bqa {

// ---------------------------------------------------------------------------
// Extended messages of an FP32 job, chi = 8: FP32 MFMA mode products on the
// swizzled float layout of k_mfma8, and the p-resolved Gram on FP64 MFMA with
// the FP32 operands widened on load, so every product and every partial sum of
// the Gram is FP64.  (The Gram's small eigenvalues come from cancellation over
// the fibres: with FP32 partial sums the anneal's <Z> error grows by three
// orders of magnitude; FP32 rounding of the mode products does not matter.)
// One block of W waves per node; with W >= 2 each wave's Gram k-steps lie in
// one p half.  The partials go into the dead Y buffer (degree 3) or G.
template <int D, int W>
struct Mf8ExtSmem {
  float T[2 * Mf8<D>::S];
  float W1[2 * Mf8<D>::S];
  float W2[2 * Mf8<D>::S];
  double G[D == 2 ? 512 : 1];
};

// This wave's k-steps of the p-resolved Gram along bond J from swizzled float
// T and Y, accumulated on FP64 MFMA (two chains per p for issue overlap).
template <int D, int W, int J>
__device__ __forceinline__ void gram8x(const float* T, const float* Y, d4v& g0, d4v& g1, const Swz8Lanes<D>& z) {
  constexpr int KS = Mf8<D>::KS, KPW = Mf8W<D, W>::KPW;
  constexpr int B = KPW < 8 ? KPW : 8;
  const int gl = opaque(z.gl[J]);
  d4v c[2][2] = {{d4v{0.0, 0.0, 0.0, 0.0}, d4v{0.0, 0.0, 0.0, 0.0}},
                 {d4v{0.0, 0.0, 0.0, 0.0}, d4v{0.0, 0.0, 0.0, 0.0}}};
  float ta[2][B], ya[2][B];
  lds_batch<D, J, B>(T, Y, gl, 0, ta[0], ya[0]);
#pragma unroll
  for (int s0 = 0; s0 < KPW; s0 += B) {
    const int q = (s0 / B) & 1;
    if (s0 + B < KPW) lds_batch<D, J, B>(T, Y, gl, s0 + B, ta[q ^ 1], ya[q ^ 1]);
    pin_after_reads<B>(ta[q], ya[q]);
#pragma unroll
    for (int b = 0; b < B; ++b) {
      const int p = W == 1 ? (s0 + b) / (KS / 2) : 0;
      c[p][b & 1] = mfma4d((double)ta[q][b], (double)ya[q][b], c[p][b & 1]);
    }
  }
  g0 = c[0][0] + c[0][1];
  g1 = c[1][0] + c[1][1];
}

template <int D, int W, int J>
__device__ __forceinline__ void ext_out8x(const float* T, const float* Y, double* Gp, const ClassView& cv, int i,
                                          cplx<double>* ext_out, const Swz8Lanes<D>& z) {
  d4v g0, g1;
  gram8x<D, W, J>(T, Y, g0, g1, z);
  __syncthreads();
  store_gd<W>(Gp, g0, g1);
  __syncthreads();
  const cplx<double>* zg = cv.zz + 2 * (size_t)cv.lmbd_pos[J * cv.n + i];
  finish8_ext64<W>(Gp, cv.out_pos[J * cv.n + i], ext_out, zg[0], zg[1]);
  __syncthreads();
}

template <int D, int W>
__global__ __launch_bounds__(64 * W, 6) void k_mfma8_ext32(ClassView cv, const cplx<float>* __restrict__ T,
                                                        const cplx<float>* __restrict__ msgs, cplx<double>* ext_out) {
  static_assert(D == 2 || D == 3, "degree 2 or 3");
  static_assert(W == 1 || W % 2 == 0, "single-p Gram slices per wave");
  static_assert(D == 2 || W * 256 <= Mf8<D>::S, "Gram partials fit in a dead work buffer");
  __shared__ __align__(16) Mf8ExtSmem<D, W> sm;
  const int n = cv.n;
  constexpr int S = Mf8<D>::S;
  const Swz8Lanes<D> zl = swz8_lanes<D, W>();
  for (int i = blockIdx.x; i < n; i += gridDim.x) {
    load_tensor8<D, W>(T + (size_t)i * S, sm.T);
    float bm[D][4];
#pragma unroll
    for (int j = 0; j < D; ++j) msg_operand8(msgs + (size_t)cv.in_pos[j * n + i] * 64, bm[j]);
    __syncthreads();
    if constexpr (D == 2) {
      mp8<D, W, 1>(sm.T, sm.W1, bm[1], zl);          // Y0 = T x1 m1
      __syncthreads();
      ext_out8x<D, W, 0>(sm.T, sm.W1, sm.G, cv, i, ext_out, zl);
      mp8<D, W, 0>(sm.T, sm.W2, bm[0], zl);          // Y1 = T x0 m0
      __syncthreads();
      ext_out8x<D, W, 1>(sm.T, sm.W2, sm.G, cv, i, ext_out, zl);
    } else {
      mp8<D, W, 2>(sm.T, sm.W1, bm[2], zl);          // W1 = T x2 m2
      __syncthreads();
      mp8<D, W, 1>(sm.W1, sm.W2, bm[1], zl);         // Y0 = W1 x1 m1
      __syncthreads();
      ext_out8x<D, W, 0>(sm.T, sm.W2, reinterpret_cast<double*>(sm.W2), cv, i, ext_out, zl);
      mp8<D, W, 0>(sm.W1, sm.W2, bm[0], zl);         // Y1 = W1 x0 m0
      __syncthreads();
      ext_out8x<D, W, 1>(sm.T, sm.W2, reinterpret_cast<double*>(sm.W2), cv, i, ext_out, zl);
      mp8<D, W, 0>(sm.T, sm.W1, bm[0], zl);          // W1 = T x0 m0
      __syncthreads();
      mp8<D, W, 1>(sm.W1, sm.W1, bm[1], zl);         // Y2 = W1 x1 m1 (in place)
      __syncthreads();
      ext_out8x<D, W, 2>(sm.T, sm.W1, reinterpret_cast<double*>(sm.W1), cv, i, ext_out, zl);
    }
  }
}

}  // namespace bqa

namespace bqa {

// ---------------------------------------------------------------------------
// Fused absorb + Z/X gates + symmetric gauge + message re-initialisation,
// chi_in = chi_out = 8, degree 3 (src/bqa/backends.py:416-432,
// state.py:142-156, 219-227), tensors stored in R (FP32 or FP64) and all
// arithmetic in FP64: the absorb multiplies by canonicalisers whose entries
// carry S^{-1/2} factors, and FP32 rounding there dominates the FP32 error over
// a full anneal.  Absorbing the canonicaliser C_k of bond k with the ZZ split
// is a mode product with W_k^(p)[c][n] = g0 C_k[c][n] + g1 (1 - 2p) C_k[8 + c][n];
// its matrix depends on the physical index p, so each 16-fibre tile (all of
// one p for degree 3) takes the operand of its p.  One block of W waves per
// node; the tensor is dead after the first mode product, so two buffers
// suffice: T -> X -> T -> X.
template <int D>
struct Mf8LocalSmemD {
  double T[2 * Mf8<D>::S];  // the mode products run in place
  cplx<double> Wt[D][2][72];  // [bond][p][n * 9 + c]: rows padded to 9 elements
  double sl[D][8];
  double red[4];
  double zf[2];  // cos, sin of the node's Z-gate angle
};

template <int D, int W, typename R>
__global__ __launch_bounds__(64 * W, 5) void k_mfma8_local64(ClassView cv, const cplx<R>* __restrict__ Tin,
                                                          cplx<R>* Tout, const zd* __restrict__ canon,
                                                          int canon_rows, int canon_cols, double zz_time,
                                                          double z_time, double x_time,
                                                          const double* __restrict__ lmbd, int lstride,
                                                          cplx<R>* msgs_init) {
  static_assert(D == 3, "degree 3");
  (void)zz_time;  // the ZZ-split gates come from cv.zz
  __shared__ __align__(16) Mf8LocalSmemD<D> sm;
  constexpr int S = Mf8<D>::S, A = Mf8<D>::A, NT = 64 * W;
  const int n = cv.n;
  const int t = threadIdx.x, l = t & 63, wave = t >> 6;
  const double cx = cos(x_time), sx = sin(x_time);
  const Lin8Lanes<D> zl = lin8_lanes<D, W>();
  static_assert(A % NT == 0, "tensor passes in whole rounds");
  const int tsw = swz64<D>(2 * t);
  for (int i = blockIdx.x; i < n; i += gridDim.x) {
    const cplx<R>* Tg = Tin + (size_t)i * S;
#pragma unroll
    for (int k = 0; k < S / NT; ++k) st_pair64(sm.T, pair64<D, NT>(tsw, k), cast<double>(Tg[t + NT * k]));
    for (int q = t; q < D * 64; q += NT) {
      const int k = q >> 6, nn = (q >> 3) & 7, c = q & 7;
      const zd* C = canon + (size_t)cv.in_pos[k * n + i] * canon_rows * canon_cols;
      const cplx<double>* zg = cv.zz + 2 * (size_t)cv.lmbd_pos[k * n + i];
      const cplx<double> u = zg[0] * C[c * canon_cols + nn], v = zg[1] * C[(8 + c) * canon_cols + nn];
      sm.Wt[k][0][nn * 9 + c] = u + v;
      sm.Wt[k][1][nn * 9 + c] = u - v;
      if (c == 0 && nn == 0)
        for (int e = 0; e < 8; ++e) sm.sl[k][e] = sqrt(lmbd[(size_t)cv.lmbd_pos[k * n + i] * lstride + e]);
    }
    if (t == 0) {
      const double phi = cv.h[i] * z_time;
      sm.zf[0] = cos(phi);
      sm.zf[1] = sin(phi);
    }
    __syncthreads();
    // each bond's operands are built from LDS right before its mode product,
    // and the products run in place (a wave reads its tiles whole before
    // writing them; the waves' tiles are disjoint): fewer registers and one
    // tensor buffer, so more nodes are resident
    {
      double b0[4], b1[4];
      msg_operand8d<9>(sm.Wt[0][0], b0);
      msg_operand8d<9>(sm.Wt[0][1], b1);
      mp8dw<D, W, 0, true>(sm.T, sm.T, b0, b1, zl);
    }
    __syncthreads();
    {
      double b0[4], b1[4];
      msg_operand8d<9>(sm.Wt[1][0], b0);
      msg_operand8d<9>(sm.Wt[1][1], b1);
      mp8dw<D, W, 1, true>(sm.T, sm.T, b0, b1, zl);
    }
    __syncthreads();
    {
      double b0[4], b1[4];
      msg_operand8d<9>(sm.Wt[2][0], b0);
      msg_operand8d<9>(sm.Wt[2][1], b1);
      mp8dw<D, W, 2, true>(sm.T, sm.T, b0, b1, zl);
    }
    __syncthreads();
    const double cz = sm.zf[0], sz = sm.zf[1];
    const cplx<double> f0 = mk(cz, -sz), f1 = mk(cz, sz);
    double ss = 0.0;
#pragma unroll
    for (int k = 0; k < A / NT; ++k) {
      const int a = t + NT * k;
      const int s0 = pair64<D, NT>(tsw, k), s1 = s0 ^ swz64<D>(2 * A);  // elements a and A + a
      const cplx<double> t0 = ld_pair64(sm.T, s0) * f0, t1 = ld_pair64(sm.T, s1) * f1;
      const double f = sm.sl[0][a >> 6] * sm.sl[1][(a >> 3) & 7] * sm.sl[2][a & 7];
      const cplx<double> y0 = mk(f * (cx * t0.x + sx * t1.y), f * (cx * t0.y - sx * t1.x));
      const cplx<double> y1 = mk(f * (cx * t1.x + sx * t0.y), f * (cx * t1.y - sx * t0.x));
      st_pair64(sm.T, s0, y0);
      st_pair64(sm.T, s1, y1);
      ss += abs2(y0) + abs2(y1);
    }
    for (int o = 32; o > 0; o >>= 1) ss += __shfl_xor(ss, o, 64);
    if (l == 0) sm.red[wave] = ss;
    __syncthreads();
    double tot = 0.0;
#pragma unroll
    for (int w = 0; w < W; ++w) tot += sm.red[w];
    const double inv = 1.0 / sqrt(tot);
    cplx<R>* To = Tout + (size_t)i * S;
#pragma unroll
    for (int k = 0; k < S / NT; ++k) To[t + NT * k] = cast<R>(scale(ld_pair64(sm.T, pair64<D, NT>(tsw, k)), inv));
    for (int q = t; q < D * 64; q += NT) {
      const int k = q >> 6, a = (q >> 3) & 7, bcol = q & 7;
      const double* lam = lmbd + (size_t)cv.lmbd_pos[k * n + i] * lstride;
      double tr = 0.0;
#pragma unroll
      for (int e = 0; e < 8; ++e) tr += lam[e];
      msgs_init[(size_t)cv.in_pos[k * n + i] * 64 + (q & 63)] = mk<R>(a == bcol ? (R)(lam[a] / tr) : (R)0, (R)0);
    }
    __syncthreads();
  }
}

}  // namespace bqa
