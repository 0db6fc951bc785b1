// Node kernels with compile-time degree D and bond dimension CHI (CHI <= 8):
// one wave per node, the node tensor and two work buffers in LDS, unrolled
// VALU mode products and lane-split Grams combined with wave shuffles.
//
// They replace the generic k_node_contract (one 256-thread workgroup per node,
// runtime index arithmetic, a barrier around every small step) on the shapes
// that occur while the bond dimension ramps up to chi_max, on chi_max = 4, and
// on FP64 jobs below chi = 8 (and for their readout at chi = 8).  Same
// contractions and results as k_node_contract:
//   BP       (state.py:97-124, backends.py:381-408): new messages, Hermitian
//            part over its real trace, dist partial maxima;
//   Ext      (state.py:127-139, backends.py:393-404): 2chi x 2chi extended
//            messages in FP64 from the p-resolved Gram and the ZZ split;
//   Density  (state.py:77-94, utils.py:23-27): rho and Bloch vectors.
// R is the storage type; C the arithmetic type (C = R for BP and density,
// double for the extended messages).
#pragma once

#include <type_traits>

#include "common.hpp"
#include "launch_grid.hpp"
#include "kernels_node.hpp"

namespace bqa {

constexpr int cpow(int b, int e) { return e <= 0 ? 1 : b * cpow(b, e - 1); }

// Phase boundary of these one-wave blocks (every launch here is 64 threads):
// the wave's LDS operations complete in issue order, so wavefront-scope fences
// around a wave barrier order the accesses without a hardware barrier and its
// wait for every outstanding LDS operation.
__device__ __forceinline__ void small_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <int D, int CHI>
struct Sm {
  static constexpr int A = cpow(CHI, D);   // chi^d
  static constexpr int S = 2 * A;          // complex elements per node
  static constexpr int Q = CHI * CHI;      // Gram entries
  static constexpr int NG = 64 / Q;        // lane groups splitting the fibres of a Gram
  static constexpr int RT = S / CHI;       // fibres along one bond
  static constexpr bool POW2 = (Q & (Q - 1)) == 0;
};

// C: the buffers' element type; GC: the Grams' (double for the extended messages).
template <int D, int CHI, typename C, typename GC = C>
struct SmallSmem {
  cplx<C> T[Sm<D, CHI>::S];
  cplx<C> X[Sm<D, CHI>::S];
  cplx<C> Y[Sm<D, CHI>::S];
  cplx<C> m[D][CHI * CHI];
  cplx<GC> red[2][64];
};

// Y = X x_J m (m: CHI x CHI in LDS, row a gives output index a along J).
template <int D, int CHI, int J, typename C>
__device__ __forceinline__ void smp(const cplx<C>* X, cplx<C>* Y, const cplx<C>* m) {
  constexpr int S = Sm<D, CHI>::S, st = cpow(CHI, D - 1 - J);
  const int l = threadIdx.x & 63;
#pragma unroll 4
  for (int q = 0; q < (S + 63) / 64; ++q) {
    const int idx = l + 64 * q;
    if (S % 64 == 0 || idx < S) {
      const int a = (idx / st) % CHI;
      const int base = idx - a * st;
      cplx<C> acc = mk<C>(0, 0);
#pragma unroll
      for (int b = 0; b < CHI; ++b) cfma(acc, m[a * CHI + b], X[base + b * st]);
      Y[idx] = acc;
    }
  }
}

// X <- X x_{D-1} m in place along the last bond, over GL lanes (gl: this lane's index among them):
// each lane owns whole fibres (reads its fibres, then writes them), so no other lane's reads are
// affected.
template <int D, int CHI, int GL, typename C>
__device__ __forceinline__ void smp_last_inplace(cplx<C>* X, const cplx<C>* m, int gl) {
  constexpr int RT = Sm<D, CHI>::RT;
#pragma unroll
  for (int q = 0; q < (RT + GL - 1) / GL; ++q) {
    const int f = gl + GL * q;
    if (RT % GL == 0 || f < RT) {
      cplx<C> x[CHI];
#pragma unroll
      for (int b = 0; b < CHI; ++b) x[b] = X[f * CHI + b];
#pragma unroll
      for (int a = 0; a < CHI; ++a) {
        cplx<C> acc = mk<C>(0, 0);
#pragma unroll
        for (int b = 0; b < CHI; ++b) cfma(acc, m[a * CHI + b], x[b]);
        X[f * CHI + a] = acc;
      }
    }
  }
}


// Gram along bond K: G[c][e] = sum over fibres r of conj(T[r, c]) Y[r, e], with
// PSPLIT separate sums over the p = 0 and p = 1 halves of the fibres (p is the
// slowest index).  Lane l takes entry o = l % Q and the fibres r = g mod NG,
// g = l / Q; the partials are combined across g.  On return lanes l < Q hold
// the full sums g[0] (and g[1] with PSPLIT).  The products and sums are formed in GC.
template <int D, int CHI, int K, bool PSPLIT, typename C, typename GC = C>
__device__ __forceinline__ void sgram(const cplx<C>* T, const cplx<C>* Y, cplx<GC> (&red)[2][64],
                                      cplx<GC> (&g)[2]) {
  using M = Sm<D, CHI>;
  constexpr int sk = cpow(CHI, D - 1 - K);
  const int l = threadIdx.x & 63;
  const int o = l % M::Q, grp = l / M::Q;
  const int c = o / CHI, e = o % CHI;
  g[0] = mk<GC>(0, 0);
  g[1] = mk<GC>(0, 0);
  if (grp < M::NG) {
#pragma unroll 8
    for (int r0 = 0; r0 < M::RT; r0 += M::NG) {
      const int r = r0 + grp;
      if (M::RT % M::NG == 0 || r < M::RT) {
        const int base = (r / sk) * CHI * sk + r % sk;
        const cplx<GC> a = cast<GC>(T[base + c * sk]), b = cast<GC>(Y[base + e * sk]);
        if (PSPLIT && r >= M::RT / 2)
          cfma_cj(g[1], a, b);
        else
          cfma_cj(g[0], a, b);
      }
    }
  }
  if constexpr (M::POW2) {
#pragma unroll
    for (int off = M::Q; off < 64; off <<= 1) {
#pragma unroll
      for (int p = 0; p < (PSPLIT ? 2 : 1); ++p) {
        g[p].x += __shfl_xor(g[p].x, off, 64);
        g[p].y += __shfl_xor(g[p].y, off, 64);
      }
    }
  } else {
#pragma unroll
    for (int p = 0; p < (PSPLIT ? 2 : 1); ++p) red[p][l] = g[p];
    small_sync();
    if (l < M::Q) {
#pragma unroll
      for (int p = 0; p < (PSPLIT ? 2 : 1); ++p) {
        cplx<GC> s = red[p][l];
#pragma unroll
        for (int q = 1; q < M::NG; ++q) s = s + red[p][l + q * M::Q];
        g[p] = s;
      }
    }
    small_sync();
  }
}

// New BP message at position pos from the Gram held by lanes l < Q (old: this
// lane's entry of the previous message at pos).
template <int CHI, typename R, typename C>
__device__ __forceinline__ void sfinish_bp(const cplx<C> (&g)[2], cplx<C> (&red)[2][64], long pos, cplx<R> old,
                                           cplx<R>* mout, double& dnum, double& dden) {
  constexpr int Q = CHI * CHI;
  const int l = threadIdx.x & 63;
  if (l < Q) red[0][l] = g[0];
  small_sync();
  if (l < Q) {
    const int c = l / CHI, e = l % CHI;
    double tr = 0.0;
#pragma unroll
    for (int k = 0; k < CHI; ++k) tr += (double)red[0][k * CHI + k].x;
    const double inv = 1.0 / tr;
    const cplx<double> a = cast<double>(red[0][l]), b = cast<double>(red[0][e * CHI + c]);
    const cplx<R> v = cast<R>(mk(0.5 * (a.x + b.x) * inv, 0.5 * (a.y - b.y) * inv));
    // squared magnitudes; the square roots are taken once per block
    const cplx<double> vd = cast<double>(v), od = cast<double>(old);
    dnum = fmax(dnum, abs2(vd - od));
    dden = fmax(dden, abs2(vd + od));
    mout[pos * Q + l] = v;
  }
  small_sync();
}

// Extended message (2chi x 2chi, FP64) at position pos from the p-resolved Gram
// held by lanes l < Q and the bond's ZZ-split gates g0c, g1c.
template <int CHI>
__device__ __forceinline__ void sfinish_ext(const cplx<double> (&g)[2], cplx<double> (&red)[2][64], long pos,
                                            cplx<double>* ext_out, cplx<double> g0c, cplx<double> g1c) {
  constexpr int Q = CHI * CHI, N2 = 2 * CHI;
  const int l = threadIdx.x & 63;
  if (l < Q) {
    red[0][l] = g[0];
    red[1][l] = g[1];
  }
  small_sync();
  cplx<double> tr = mk(0.0, 0.0);
#pragma unroll
  for (int k = 0; k < CHI; ++k) tr = tr + red[0][k * CHI + k] + red[1][k * CHI + k];
  const cplx<double> inv = cinv(tr * mk(abs2(g0c) + abs2(g1c), 0.0));
#pragma unroll
  for (int q = 0; q < (N2 * N2 + 63) / 64; ++q) {
    const int o = l + 64 * q;
    if (o < N2 * N2) {
      const int row = o / N2, col = o % N2;
      const int s = row / CHI, c = row % CHI, s2 = col / CHI, e = col % CHI;
      const cplx<double> a0 = red[0][c * CHI + e], a1 = red[1][c * CHI + e];
      const cplx<double> gsum = ((s + s2) & 1) ? a0 - a1 : a0 + a1;
      const cplx<double> ga = s ? g1c : g0c, gb = s2 ? g1c : g0c;
      ext_out[pos * N2 * N2 + o] = conj(ga) * gb * gsum * inv;
    }
  }
  small_sync();
}

// Grams of the extended messages are FP64 whatever the buffers hold.
template <int MODE, typename C>
using SmallGC = typename std::conditional<MODE == kModeExt, double, C>::type;

// Outgoing message (BP) or extended message (Ext) along bond K from Y_K.
template <int D, int CHI, int K, int MODE, typename R, typename C>
__device__ __forceinline__ void sout(SmallSmem<D, CHI, C, SmallGC<MODE, C>>& sm, const cplx<C>* A,
                                     const cplx<C>* Y, const ClassView& cv, int i, const cplx<R> (&old)[D],
                                     cplx<R>* mout, cplx<double>* ext_out, double& dnum, double& dden) {
  using GC = SmallGC<MODE, C>;
  cplx<GC> g[2];
  sgram<D, CHI, K, MODE == kModeExt, C, GC>(A, Y, sm.red, g);
  const long pos = cv.out_pos[K * cv.n + i];
  if constexpr (MODE == kModeBP) {
    sfinish_bp<CHI, R, C>(g, sm.red, pos, old[K], mout, dnum, dden);
  } else {
    const cplx<double>* zg = cv.zz + 2 * (size_t)cv.lmbd_pos[K * cv.n + i];
    sfinish_ext<CHI>(g, sm.red, pos, ext_out, zg[0], zg[1]);
  }
}

// One block's share of a class: nodes start, start + gridDim.x, ...
template <int D, int CHI, int MODE, typename R, typename C>
__device__ __forceinline__ void small_run(const ClassView& cv, const cplx<R>* __restrict__ T, cplx<R>* msgA,
                                          cplx<R>* msgB, BpStatus* st, cplx<double>* ext_out, double* rho_out,
                                          double* bloch_out, int start) {
  static_assert(D >= 2 && D <= 5 && CHI >= 2 && CHI <= 8, "shape");
  static_assert(D <= 4 || MODE != kModeDensity, "degree 5: BP sweeps and extended messages");
  using M = Sm<D, CHI>;
  __shared__ __align__(16) SmallSmem<D, CHI, C, SmallGC<MODE, C>> sm;
  const int n = cv.n;
  const int l = threadIdx.x & 63;
  const cplx<R>* min = msgA;
  cplx<R>* mout = msgB;
  if (MODE == kModeBP) {
    const int cur = bp_begin(st, cv.sweep);
    if (cur < 0) return;
    min = cur ? msgB : msgA;
    mout = cur ? msgA : msgB;
  }
  double dnum = 0.0, dden = 0.0;
  for (int i = start; i < n; i += gridDim.x) {
    const cplx<R>* Tg = T + (size_t)i * M::S;
#pragma unroll
    for (int q = 0; q < (M::S + 63) / 64; ++q) {
      const int x = l + 64 * q;
      if (M::S % 64 == 0 || x < M::S) sm.T[x] = cast<C>(Tg[x]);
    }
#pragma unroll
    for (int j = 0; j < D; ++j) {
      const cplx<R>* mj = min + (size_t)cv.in_pos[j * n + i] * M::Q;
      if (l < M::Q) sm.m[j][l] = cast<C>(mj[l]);
    }
    // this lane's entries of the previous outgoing messages, loaded with the inputs
    cplx<R> old[D];
#pragma unroll
    for (int k = 0; k < D; ++k)
      old[k] = (MODE == kModeBP && l < M::Q) ? min[(size_t)cv.out_pos[k * n + i] * M::Q + l] : mk<R>(0, 0);
    small_sync();
    if constexpr (MODE == kModeDensity) {
      smp<D, CHI, 0, C>(sm.T, sm.X, sm.m[0]);
      small_sync();
      smp<D, CHI, 1, C>(sm.X, sm.Y, sm.m[1]);
      small_sync();
      const cplx<C>* Yf = sm.Y;
      if constexpr (D >= 3) {
        smp<D, CHI, 2, C>(sm.Y, sm.X, sm.m[2]);
        small_sync();
        Yf = sm.X;
      }
      if constexpr (D == 4) {
        smp<D, CHI, 3, C>(sm.X, sm.Y, sm.m[3]);
        small_sync();
        Yf = sm.Y;
      }
      // rho[p][q] = sum_a Y[p, a] conj(T[q, a])
      cplx<C> r[4] = {mk<C>(0, 0), mk<C>(0, 0), mk<C>(0, 0), mk<C>(0, 0)};
      for (int a = l; a < M::A; a += 64) {
        cfma_cj(r[0], sm.T[a], Yf[a]);
        cfma_cj(r[1], sm.T[M::A + a], Yf[a]);
        cfma_cj(r[2], sm.T[a], Yf[M::A + a]);
        cfma_cj(r[3], sm.T[M::A + a], Yf[M::A + a]);
      }
#pragma unroll
      for (int q = 0; q < 4; ++q)
        for (int o = 32; o > 0; o >>= 1) {
          r[q].x += __shfl_xor(r[q].x, o, 64);
          r[q].y += __shfl_xor(r[q].y, o, 64);
        }
      if (l == 0) {
        cplx<double> r00 = cast<double>(r[0]), r01 = cast<double>(r[1]);
        cplx<double> r10 = cast<double>(r[2]), r11 = cast<double>(r[3]);
        const cplx<double> inv = cinv(r00 + r11);
        r00 = r00 * inv;
        r01 = r01 * inv;
        r10 = r10 * inv;
        r11 = r11 * inv;
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
      small_sync();
      continue;
    } else if constexpr (D == 2) {
      smp<D, CHI, 1, C>(sm.T, sm.X, sm.m[1]);  // Y0 = T x1 m1
      small_sync();
      sout<D, CHI, 0, MODE, R, C>(sm, sm.T, sm.X, cv, i, old, mout, ext_out, dnum, dden);
      smp<D, CHI, 0, C>(sm.T, sm.X, sm.m[0]);  // Y1 = T x0 m0
      small_sync();
      sout<D, CHI, 1, MODE, R, C>(sm, sm.T, sm.X, cv, i, old, mout, ext_out, dnum, dden);
    } else if constexpr (D == 3) {
      // The incoming messages are Hermitian, so one moves to the bra side as a product of the
      // conjugated factor: with P_j = T x_j m_j, out_0 = sum conj(P_2) P_1, out_1 = sum conj(P_2) P_0,
      // out_2 = sum conj(P_1) P_0 (each over the other bonds, per physical half for the extended
      // messages): three mode products instead of the reference schedule's five
      // (backends.py:381-404), P_2 in place of T.
      smp<D, CHI, 0, C>(sm.T, sm.X, sm.m[0]);  // P_0 = T x0 m0
      smp<D, CHI, 1, C>(sm.T, sm.Y, sm.m[1]);  // P_1 = T x1 m1
      small_sync();
      sout<D, CHI, 2, MODE, R, C>(sm, sm.Y, sm.X, cv, i, old, mout, ext_out, dnum, dden);
      smp_last_inplace<D, CHI, 64, C>(sm.T, sm.m[2], threadIdx.x & 63);  // T <- P_2 = T x2 m2
      small_sync();
      sout<D, CHI, 0, MODE, R, C>(sm, sm.T, sm.Y, cv, i, old, mout, ext_out, dnum, dden);
      sout<D, CHI, 1, MODE, R, C>(sm, sm.T, sm.X, cv, i, old, mout, ext_out, dnum, dden);
    } else if constexpr (D == 4) {
      // degree 4: W = T x3 m3 x2 m2 shared by outputs 0 and 1, V = T x0 m0 x1 m1
      // shared by outputs 2 and 3 (8 mode products instead of 12).
      smp<D, CHI, 3, C>(sm.T, sm.X, sm.m[3]);
      small_sync();
      smp<D, CHI, 2, C>(sm.X, sm.Y, sm.m[2]);  // W in Y
      small_sync();
      smp<D, CHI, 1, C>(sm.Y, sm.X, sm.m[1]);  // Y0 = W x1 m1
      small_sync();
      sout<D, CHI, 0, MODE, R, C>(sm, sm.T, sm.X, cv, i, old, mout, ext_out, dnum, dden);
      smp<D, CHI, 0, C>(sm.Y, sm.X, sm.m[0]);  // Y1 = W x0 m0
      small_sync();
      sout<D, CHI, 1, MODE, R, C>(sm, sm.T, sm.X, cv, i, old, mout, ext_out, dnum, dden);
      smp<D, CHI, 0, C>(sm.T, sm.X, sm.m[0]);
      small_sync();
      smp<D, CHI, 1, C>(sm.X, sm.Y, sm.m[1]);  // V in Y
      small_sync();
      smp<D, CHI, 3, C>(sm.Y, sm.X, sm.m[3]);  // Y2 = V x3 m3
      small_sync();
      sout<D, CHI, 2, MODE, R, C>(sm, sm.T, sm.X, cv, i, old, mout, ext_out, dnum, dden);
      smp<D, CHI, 2, C>(sm.Y, sm.X, sm.m[2]);  // Y3 = V x2 m2
      small_sync();
      sout<D, CHI, 3, MODE, R, C>(sm, sm.T, sm.X, cv, i, old, mout, ext_out, dnum, dden);
    } else {
      // degree 5: A = T x2 m2 x3 m3 x4 m4 shared by outputs 0 and 1, B = T x0 m0 x1 m1 x2 m2
      // shared by outputs 3 and 4, output 2 on its own (14 mode products instead of 20).
      smp<D, CHI, 2, C>(sm.T, sm.X, sm.m[2]);
      small_sync();
      smp<D, CHI, 3, C>(sm.X, sm.Y, sm.m[3]);
      small_sync();
      smp<D, CHI, 4, C>(sm.Y, sm.X, sm.m[4]);  // A in X
      small_sync();
      smp<D, CHI, 1, C>(sm.X, sm.Y, sm.m[1]);  // Y0 = A x1 m1
      small_sync();
      sout<D, CHI, 0, MODE, R, C>(sm, sm.T, sm.Y, cv, i, old, mout, ext_out, dnum, dden);
      smp<D, CHI, 0, C>(sm.X, sm.Y, sm.m[0]);  // Y1 = A x0 m0
      small_sync();
      sout<D, CHI, 1, MODE, R, C>(sm, sm.T, sm.Y, cv, i, old, mout, ext_out, dnum, dden);
      smp<D, CHI, 0, C>(sm.T, sm.X, sm.m[0]);
      small_sync();
      smp<D, CHI, 1, C>(sm.X, sm.Y, sm.m[1]);
      small_sync();
      smp<D, CHI, 3, C>(sm.Y, sm.X, sm.m[3]);
      small_sync();
      smp<D, CHI, 4, C>(sm.X, sm.Y, sm.m[4]);  // Y2 = T x0 m0 x1 m1 x3 m3 x4 m4
      small_sync();
      sout<D, CHI, 2, MODE, R, C>(sm, sm.T, sm.Y, cv, i, old, mout, ext_out, dnum, dden);
      smp<D, CHI, 0, C>(sm.T, sm.X, sm.m[0]);
      small_sync();
      smp<D, CHI, 1, C>(sm.X, sm.Y, sm.m[1]);
      small_sync();
      smp<D, CHI, 2, C>(sm.Y, sm.X, sm.m[2]);  // B in X
      small_sync();
      smp<D, CHI, 4, C>(sm.X, sm.Y, sm.m[4]);  // Y3 = B x4 m4
      small_sync();
      sout<D, CHI, 3, MODE, R, C>(sm, sm.T, sm.Y, cv, i, old, mout, ext_out, dnum, dden);
      smp<D, CHI, 3, C>(sm.X, sm.Y, sm.m[3]);  // Y4 = B x3 m3
      small_sync();
      sout<D, CHI, 4, MODE, R, C>(sm, sm.T, sm.Y, cv, i, old, mout, ext_out, dnum, dden);
    }
  }
  if (MODE == kModeBP && start < n) {  // a block without nodes here adds nothing to the maxima
    for (int o = 32; o > 0; o >>= 1) {
      dnum = fmax(dnum, __shfl_xor(dnum, o, 64));
      dden = fmax(dden, __shfl_xor(dden, o, 64));
    }
    if (l == 0) {
      bp_end(st, cv.sweep, sqrt(dnum), sqrt(dden));
    }
  }
}

template <int D, int CHI, int MODE, typename R, typename C>
__global__ __launch_bounds__(64) void k_small(ClassView cv, const cplx<R>* __restrict__ T, cplx<R>* msgA,
                                              cplx<R>* msgB, BpStatus* st, cplx<double>* ext_out,
                                              double* rho_out, double* bloch_out) {
  small_run<D, CHI, MODE, R, C>(cv, T, msgA, msgB, st, ext_out, rho_out, bloch_out, blockIdx.x);
}

// One BP sweep of the classes of several jobs (same degree and chi) in one launch: entry e's
// node i is flat node prefix_e + i, taken by block (prefix_e + i) mod gridDim.x, so the
// blocks share the nodes of all entries evenly.  Each entry has its own status block, so
// every job keeps its own convergence and buffer bookkeeping.
template <int D, int CHI, typename R>
__global__ __launch_bounds__(64) void k_small_multi(const BpEntry* __restrict__ tab, int entries, int sweep) {
  for (int e = 0; e < entries; ++e) {
    ClassView cv = tab[e].cv;
    cv.sweep = sweep;
    const long g = gridDim.x;
    const int start = (int)((((long)blockIdx.x - tab[e].prefix) % g + g) % g);
    if (start >= cv.n) continue;  // none of this entry's nodes falls to this block
    small_run<D, CHI, kModeBP, R, R>(cv, reinterpret_cast<const cplx<R>*>(tab[e].T),
                                     reinterpret_cast<cplx<R>*>(tab[e].mA), reinterpret_cast<cplx<R>*>(tab[e].mB),
                                     tab[e].st, nullptr, nullptr, nullptr, start);
  }
}

// ---------------------------------------------------------------------------
// BP sweep with NPW nodes per wave (chi <= 4, degree 2 and 3): each node gets
// GL = 64 / NPW lanes and its own LDS tensor and work buffers; the wave's phases
// (and their barriers) serve NPW nodes at once, and at GL >= Q a Gram entry is
// summed by one lane over all fibres (no cross-lane reduction).  Degree 2 has
// k_small's schedule; degree 3 moves one message of each output to the bra side
// (three mode products instead of five), so its messages equal k_small's to
// rounding.
template <int D, int CHI, int NPW, typename C>
struct NpwSmem {
  cplx<C> T[NPW][Sm<D, CHI>::S];
  cplx<C> X[NPW][Sm<D, CHI>::S];
  cplx<C> Y[NPW][Sm<D, CHI>::S];
  cplx<C> m[NPW][D][CHI * CHI];
  cplx<C> red[NPW][CHI * CHI];
};

// Y = X x_J m for this lane group's node (gl: lane in the group of GL lanes).
template <int D, int CHI, int J, int GL, typename C>
__device__ __forceinline__ void smp_g(const cplx<C>* X, cplx<C>* Y, const cplx<C>* m, int gl) {
  constexpr int S = Sm<D, CHI>::S, st = cpow(CHI, D - 1 - J);
#pragma unroll
  for (int q = 0; q < (S + GL - 1) / GL; ++q) {
    const int idx = gl + GL * q;
    if (S % GL == 0 || idx < S) {
      const int a = (idx / st) % CHI;
      const int base = idx - a * st;
      cplx<C> acc = mk<C>(0, 0);
#pragma unroll
      for (int b = 0; b < CHI; ++b) cfma(acc, m[a * CHI + b], X[base + b * st]);
      Y[idx] = acc;
    }
  }
}

// Gram along bond K of this group's node; lanes gl < Q end with the full sum.
template <int D, int CHI, int K, int GL, typename C>
__device__ __forceinline__ cplx<C> sgram_g(const cplx<C>* T, const cplx<C>* Y, int gl) {
  using M = Sm<D, CHI>;
  // fibre groups per entry: GL / Q for a power-of-two Q (then a power of two), else one
  constexpr int Q = M::Q, NG = (Q & (Q - 1)) == 0 ? GL / Q : 1, sk = cpow(CHI, D - 1 - K);
  static_assert(NG >= 1 && (NG & (NG - 1)) == 0, "a power-of-two number of fibre groups per Gram entry");
  const int o = gl % Q, grp = gl / Q;
  const int c = o / CHI, e = o % CHI;
  cplx<C> g = mk<C>(0, 0);
  if (grp < NG) {
#pragma unroll 8
    for (int r0 = 0; r0 < M::RT; r0 += NG) {
      const int r = r0 + grp;
      if (M::RT % NG == 0 || r < M::RT) {
        const int base = (r / sk) * CHI * sk + r % sk;
        cfma_cj(g, T[base + c * sk], Y[base + e * sk]);
      }
    }
  }
#pragma unroll
  for (int off = Q; off < Q * NG; off <<= 1) {  // NG > 1 only for a power-of-two Q: stays inside the group
    g.x += __shfl_xor(g.x, off, 64);
    g.y += __shfl_xor(g.y, off, 64);
  }
  return g;
}

// Output message K of this group's node from the Gram sum over every bond but K of conj(A) B.
template <int D, int CHI, int K, int NPW, typename R, typename C>
__device__ __forceinline__ void sout_g(NpwSmem<D, CHI, NPW, C>& sm, const cplx<C>* A, const cplx<C>* Y,
                                       const ClassView& cv, int i, bool valid, int gi, int gl,
                                       const cplx<R> (&old)[D], cplx<R>* mout, double& dnum, double& dden) {
  constexpr int GL = 64 / NPW, Q = CHI * CHI;
  const cplx<C> g = sgram_g<D, CHI, K, GL, C>(A, Y, gl);
  if (gl < Q) sm.red[gi][gl] = g;
  small_sync();
  if (valid && gl < Q) {
    const int c = gl / CHI, e = gl % CHI;
    double tr = 0.0;
#pragma unroll
    for (int k = 0; k < CHI; ++k) tr += (double)sm.red[gi][k * CHI + k].x;
    const double inv = 1.0 / tr;
    const cplx<double> a = cast<double>(sm.red[gi][gl]), b = cast<double>(sm.red[gi][e * CHI + c]);
    const cplx<R> v = cast<R>(mk(0.5 * (a.x + b.x) * inv, 0.5 * (a.y - b.y) * inv));
    const cplx<double> vd = cast<double>(v), od = cast<double>(old[K]);
    dnum = fmax(dnum, abs2(vd - od));
    dden = fmax(dden, abs2(vd + od));
    mout[(long)cv.out_pos[K * cv.n + i] * Q + gl] = v;
  }
  small_sync();
}

// The node groups start, start + gridDim.x, ... (NPW nodes each) of one BP sweep; the
// block adds its distance maxima only if it had a group (start < groups).
template <int D, int CHI, int NPW, typename R>
__device__ __forceinline__ void npw_run(NpwSmem<D, CHI, NPW, R>& sm, const ClassView& cv,
                                        const cplx<R>* __restrict__ T, cplx<R>* msgA, cplx<R>* msgB, BpStatus* st,
                                        long start) {
  static_assert((D == 2 || D == 3) && CHI >= 2 && CHI <= 4, "shape");
  using C = R;
  using M = Sm<D, CHI>;
  constexpr int GL = 64 / NPW;
  const int cur = bp_begin(st, cv.sweep);
  if (cur < 0) return;
  const cplx<R>* min = cur ? msgB : msgA;
  cplx<R>* mout = cur ? msgA : msgB;
  const int n = cv.n;
  const int l = threadIdx.x & 63, gi = l / GL, gl = l % GL;
  double dnum = 0.0, dden = 0.0;
  for (long i0 = start * NPW; i0 < n; i0 += (long)gridDim.x * NPW) {
    const int i = (int)i0 + gi;
    const bool valid = i < n;
    const int ic = valid ? i : n - 1;  // a group past the end repeats the last node and stores nothing
    const cplx<R>* Tg = T + (size_t)ic * M::S;
#pragma unroll
    for (int q = 0; q < (M::S + GL - 1) / GL; ++q) {
      const int x = gl + GL * q;
      if (M::S % GL == 0 || x < M::S) sm.T[gi][x] = cast<C>(Tg[x]);
    }
#pragma unroll
    for (int j = 0; j < D; ++j) {
      const cplx<R>* mj = min + (size_t)cv.in_pos[j * n + ic] * M::Q;
      if (gl < M::Q) sm.m[gi][j][gl] = cast<C>(mj[gl]);
    }
    cplx<R> old[D];
#pragma unroll
    for (int k = 0; k < D; ++k) old[k] = gl < M::Q ? min[(size_t)cv.out_pos[k * n + ic] * M::Q + gl] : mk<R>(0, 0);
    small_sync();
    if constexpr (D == 2) {
      smp_g<D, CHI, 1, GL, C>(sm.T[gi], sm.X[gi], sm.m[gi][1], gl);  // Y0 = T x1 m1
      small_sync();
      sout_g<D, CHI, 0, NPW, R, C>(sm, sm.T[gi], sm.X[gi], cv, ic, valid, gi, gl, old, mout, dnum, dden);
      smp_g<D, CHI, 0, GL, C>(sm.T[gi], sm.X[gi], sm.m[gi][0], gl);  // Y1 = T x0 m0
      small_sync();
      sout_g<D, CHI, 1, NPW, R, C>(sm, sm.T[gi], sm.X[gi], cv, ic, valid, gi, gl, old, mout, dnum, dden);
    } else {
      // The incoming messages are Hermitian, so one moves to the bra side as a product of the
      // conjugated factor: with P_j = T x_j m_j, out_0 = sum conj(P_2) P_1, out_1 = sum conj(P_2) P_0,
      // out_2 = sum conj(P_1) P_0 (each over the other bonds): three mode products instead of the
      // reference schedule's five (backends.py:381-404), P_2 in place of T.
      smp_g<D, CHI, 0, GL, C>(sm.T[gi], sm.X[gi], sm.m[gi][0], gl);  // P_0 = T x0 m0
      smp_g<D, CHI, 1, GL, C>(sm.T[gi], sm.Y[gi], sm.m[gi][1], gl);  // P_1 = T x1 m1
      small_sync();
      sout_g<D, CHI, 2, NPW, R, C>(sm, sm.Y[gi], sm.X[gi], cv, ic, valid, gi, gl, old, mout, dnum, dden);
      smp_last_inplace<D, CHI, GL, C>(sm.T[gi], sm.m[gi][2], gl);  // T <- P_2 = T x2 m2
      small_sync();
      sout_g<D, CHI, 0, NPW, R, C>(sm, sm.T[gi], sm.Y[gi], cv, ic, valid, gi, gl, old, mout, dnum, dden);
      sout_g<D, CHI, 1, NPW, R, C>(sm, sm.T[gi], sm.X[gi], cv, ic, valid, gi, gl, old, mout, dnum, dden);
    }
  }
  for (int o = 32; o > 0; o >>= 1) {
    dnum = fmax(dnum, __shfl_xor(dnum, o, 64));
    dden = fmax(dden, __shfl_xor(dden, o, 64));
  }
  if (l == 0 && start * NPW < n) bp_end(st, cv.sweep, sqrt(dnum), sqrt(dden));
}

template <int D, int CHI, int NPW, typename R>
__global__ __launch_bounds__(64) void k_small_npw(ClassView cv, const cplx<R>* __restrict__ T, cplx<R>* msgA,
                                                  cplx<R>* msgB, BpStatus* st) {
  __shared__ __align__(16) NpwSmem<D, CHI, NPW, R> sm;
  npw_run<D, CHI, NPW, R>(sm, cv, T, msgA, msgB, st, blockIdx.x);
}

// One BP sweep of several jobs' classes (same degree and chi) with NPW nodes per wave: entry e's
// node group q (NPW nodes) is flat group prefix_e + q (prefix counted in groups), taken by block
// (prefix_e + q) mod gridDim.x, as in k_small_multi.
template <int D, int CHI, int NPW, typename R>
__global__ __launch_bounds__(64) void k_small_npw_multi(const BpEntry* __restrict__ tab, int entries, int sweep) {
  __shared__ __align__(16) NpwSmem<D, CHI, NPW, R> sm;
  for (int e = 0; e < entries; ++e) {
    ClassView cv = tab[e].cv;
    cv.sweep = sweep;
    const long g = gridDim.x;
    const long start = (((long)blockIdx.x - tab[e].prefix) % g + g) % g;
    if (start * NPW >= cv.n) continue;  // none of this entry's groups falls to this block
    npw_run<D, CHI, NPW, R>(sm, cv, reinterpret_cast<const cplx<R>*>(tab[e].T),
                            reinterpret_cast<cplx<R>*>(tab[e].mA), reinterpret_cast<cplx<R>*>(tab[e].mB), tab[e].st,
                            start);
    small_sync();  // the next entry's loads overwrite the LDS this one's last readers use
  }
}

// Shapes served: degree 2 and 3 with chi 2..8, degree 4 with chi 2..4.
inline bool small_shape(int d, int chi) {
  return chi >= 2 && ((d == 2 || d == 3) ? chi <= 8 : (d == 4 && chi <= 4));
}

// BP sweeps also serve degree 5 with chi 2..4 for FP32 storage (three 16 KB complex64 node
// buffers at chi = 4; FP64 would need 96 KB of LDS per one-wave block).
template <typename R>
inline bool small_bp_shape(int d, int chi) {
  return small_shape(d, chi) || (sizeof(R) == sizeof(float) && d == 5 && chi >= 2 && chi <= 4);
}

inline int small_grid(int n) { return n < 8192 ? n : 8192; }

// Grid of k_small under launch policy `policy`: 0 at most 8192 one-wave blocks,
// otherwise the grid policies of launch_grid.hpp (1 one resident round, 2 two).
template <auto K>
inline int small_policy_grid(int n, int policy) {
  return policy == 0 ? small_grid(n) : policy_grid<K, 64>(n, policy);
}

template <int D, int CHI, int MODE, typename R, typename C>
inline void launch_small_dc(ClassView cv, const cplx<R>* T, cplx<R>* mA, cplx<R>* mB, BpStatus* st,
                            cplx<double>* ext, double* rho, double* bloch, hipStream_t s, int policy) {
  hipLaunchKernelGGL((k_small<D, CHI, MODE, R, C>), dim3(small_policy_grid<k_small<D, CHI, MODE, R, C>>(cv.n, policy)),
                     dim3(64), 0, s, cv, T, mA, mB, st, ext, rho, bloch);
}

template <int D, int MODE, typename R, typename C>
inline bool launch_small_d(int chi, ClassView cv, const cplx<R>* T, cplx<R>* mA, cplx<R>* mB, BpStatus* st,
                           cplx<double>* ext, double* rho, double* bloch, hipStream_t s, int policy) {
  switch (chi) {
    case 2: launch_small_dc<D, 2, MODE, R, C>(cv, T, mA, mB, st, ext, rho, bloch, s, policy); return true;
    case 3: launch_small_dc<D, 3, MODE, R, C>(cv, T, mA, mB, st, ext, rho, bloch, s, policy); return true;
    case 4: launch_small_dc<D, 4, MODE, R, C>(cv, T, mA, mB, st, ext, rho, bloch, s, policy); return true;
    default: break;
  }
  if constexpr (D <= 3) {
    switch (chi) {
      case 5: launch_small_dc<D, 5, MODE, R, C>(cv, T, mA, mB, st, ext, rho, bloch, s, policy); return true;
      case 6: launch_small_dc<D, 6, MODE, R, C>(cv, T, mA, mB, st, ext, rho, bloch, s, policy); return true;
      case 7: launch_small_dc<D, 7, MODE, R, C>(cv, T, mA, mB, st, ext, rho, bloch, s, policy); return true;
      case 8: launch_small_dc<D, 8, MODE, R, C>(cv, T, mA, mB, st, ext, rho, bloch, s, policy); return true;
      default: break;
    }
  }
  return false;
}

// Launch k_small for (d, chi) if the shape is served; false otherwise.  The
// extended-message mode reads the ZZ-split gates from cv.zz.
template <int MODE, typename R, typename C>
inline bool launch_small(ClassView cv, int chi, const cplx<R>* T, cplx<R>* mA, cplx<R>* mB, BpStatus* st,
                         cplx<double>* ext, double* rho, double* bloch, hipStream_t s, int policy = 0) {
  if (cv.n == 0) return false;
  if (cv.d == 5) {
    // FP32 storage, chi <= 4: BP in FP32, extended messages from FP32 mode products with FP64 Grams
    // (three 16 KB buffers at chi = 4; FP64 buffers would need 96 KB per one-wave block)
    if constexpr (sizeof(R) == sizeof(float)) {
      if (small_bp_shape<R>(cv.d, chi)) {
        if constexpr (MODE == kModeBP && sizeof(C) == sizeof(float))
          return launch_small_d<5, MODE, R, C>(chi, cv, T, mA, mB, st, ext, rho, bloch, s, policy);
        if constexpr (MODE == kModeExt)
          return launch_small_d<5, MODE, R, float>(chi, cv, T, mA, mB, st, ext, rho, bloch, s, policy);
      }
    }
    return false;
  }
  if (!small_shape(cv.d, chi)) return false;
  switch (cv.d) {
    case 2: return launch_small_d<2, MODE, R, C>(chi, cv, T, mA, mB, st, ext, rho, bloch, s, policy);
    case 3: return launch_small_d<3, MODE, R, C>(chi, cv, T, mA, mB, st, ext, rho, bloch, s, policy);
    case 4: return launch_small_d<4, MODE, R, C>(chi, cv, T, mA, mB, st, ext, rho, bloch, s, policy);
    default: return false;
  }
}

// Shapes of k_small_npw: degree 2 and 3 with chi 2..4.
inline bool npw_shape(int d, int chi) { return (d == 2 || d == 3) && chi >= 2 && chi <= 4; }

template <int D, int CHI, int NPW, typename R>
inline void launch_small_npw_dc(ClassView cv, const cplx<R>* T, cplx<R>* mA, cplx<R>* mB, BpStatus* st,
                                hipStream_t s) {
  const int groups = (cv.n + NPW - 1) / NPW;
  hipLaunchKernelGGL((k_small_npw<D, CHI, NPW, R>), dim3(policy_grid<k_small_npw<D, CHI, NPW, R>, 64>(groups, 0)),
                     dim3(64), 0, s, cv, T, mA, mB, st);
}

template <int NPW, typename R>
inline bool launch_small_npw(ClassView cv, int chi, const cplx<R>* T, cplx<R>* mA, cplx<R>* mB, BpStatus* st,
                             hipStream_t s) {
  if (cv.n == 0 || !npw_shape(cv.d, chi)) return false;
#define BQA_NPW(DD, CC)                                                        \
  if (cv.d == DD && chi == CC) {                                               \
    launch_small_npw_dc<DD, CC, NPW, R>(cv, T, mA, mB, st, s);                 \
    return true;                                                               \
  }
  BQA_NPW(2, 2) BQA_NPW(2, 3) BQA_NPW(2, 4) BQA_NPW(3, 2) BQA_NPW(3, 3) BQA_NPW(3, 4)
#undef BQA_NPW
  return false;
}

// Fused absorb + Z/X gates + symmetric gauge + message re-initialisation at
// a fixed bond dimension (chi_in = chi_out = CHI), one wave per node, FP64
// arithmetic on R storage; same operation as k_node_local with all three
// flags (backends.py:416-432, state.py:142-156, 219-227).  Absorbing the
// canonicaliser C_k of bond k with the ZZ split is a mode product with the
// p-dependent matrix W_k^(p)[c][n] = g0 C_k[c][n] + (1 - 2p) g1 C_k[CHI + c][n]
// (gates from cv.zz), built once per node in LDS.
template <int D, int CHI>
struct SmallLocalSmem {
  cplx<double> X[Sm<D, CHI>::S];
  cplx<double> Y[Sm<D, CHI>::S];
  cplx<double> W[D][2][CHI * CHI];  // [bond][p][c * CHI + n]
  double sl[D][CHI];                // sqrt(lambda) per bond
  double red[2];                    // cos, sin of the Z-gate angle
};

template <int D, int CHI, int K>
__device__ __forceinline__ void smp_absorb(const cplx<double>* X, cplx<double>* Y, const cplx<double> (&W)[2][CHI * CHI]) {
  constexpr int S = Sm<D, CHI>::S, A = Sm<D, CHI>::A, st = cpow(CHI, D - 1 - K);
  const int l = threadIdx.x & 63;
#pragma unroll 4
  for (int q = 0; q < (S + 63) / 64; ++q) {
    const int y = l + 64 * q;
    if (S % 64 == 0 || y < S) {
      const int nk = (y / st) % CHI;
      const int base = y - nk * st;
      const cplx<double>* w = W[y >= A ? 1 : 0];
      cplx<double> acc = mk(0.0, 0.0);
#pragma unroll
      for (int c = 0; c < CHI; ++c) cfma(acc, X[base + c * st], w[c * CHI + nk]);
      Y[y] = acc;
    }
  }
}

template <int D, int CHI, typename R>
__global__ __launch_bounds__(64) void k_small_local(ClassView cv, const cplx<R>* __restrict__ Tin, cplx<R>* Tout,
                                                    const cplx<double>* __restrict__ canon, int canon_rows,
                                                    int canon_cols, double z_time, double x_time,
                                                    const double* __restrict__ lmbd, int lstride,
                                                    cplx<R>* msgs_init) {
  using M = Sm<D, CHI>;
  __shared__ __align__(16) SmallLocalSmem<D, CHI> sm;
  const int n = cv.n;
  const int l = threadIdx.x & 63;
  const double cx = cos(x_time), sx = sin(x_time);
  for (int i = blockIdx.x; i < n; i += gridDim.x) {
    const cplx<R>* Tg = Tin + (size_t)i * M::S;
#pragma unroll 4
    for (int q = 0; q < (M::S + 63) / 64; ++q) {
      const int x = l + 64 * q;
      if (M::S % 64 == 0 || x < M::S) sm.X[x] = cast<double>(Tg[x]);
    }
    for (int q = l; q < D * M::Q; q += 64) {
      const int k = q / M::Q, o = q % M::Q, c = o / CHI, nn = o % CHI;
      const cplx<double>* C = canon + (size_t)cv.in_pos[k * n + i] * canon_rows * canon_cols;
      const cplx<double>* zg = cv.zz + 2 * (size_t)cv.lmbd_pos[k * n + i];
      const cplx<double> u = zg[0] * C[c * canon_cols + nn], v = zg[1] * C[(CHI + c) * canon_cols + nn];
      sm.W[k][0][o] = u + v;
      sm.W[k][1][o] = u - v;
    }
    for (int q = l; q < D * CHI; q += 64) {
      const int k = q / CHI, e = q % CHI;
      sm.sl[k][e] = sqrt(lmbd[(size_t)cv.lmbd_pos[k * n + i] * lstride + e]);
    }
    if (l == 0) {
      const double phi = cv.h[i] * z_time;
      sm.red[0] = cos(phi);
      sm.red[1] = sin(phi);
    }
    small_sync();
    smp_absorb<D, CHI, 0>(sm.X, sm.Y, sm.W[0]);
    small_sync();
    smp_absorb<D, CHI, 1>(sm.Y, sm.X, sm.W[1]);
    small_sync();
    cplx<double>* F = sm.X;
    if constexpr (D >= 3) {
      smp_absorb<D, CHI, 2>(sm.X, sm.Y, sm.W[2]);
      small_sync();
      F = sm.Y;
    }
    if constexpr (D == 4) {
      smp_absorb<D, CHI, 3>(sm.Y, sm.X, sm.W[3]);
      small_sync();
      F = sm.X;
    }
    // Z then X on the physical index, then sqrt(lambda) on every bond and the L2 norm
    const double cz = sm.red[0], sz = sm.red[1];
    const cplx<double> f0 = mk(cz, -sz), f1 = mk(cz, sz);
    double ss = 0.0;
    for (int a = l; a < M::A; a += 64) {
      const cplx<double> t0 = F[a] * f0, t1 = F[M::A + a] * f1;
      double f = 1.0;
      int rem = a;
#pragma unroll
      for (int k = D - 1; k >= 0; --k) {
        f *= sm.sl[k][rem % CHI];
        rem /= CHI;
      }
      const cplx<double> y0 = mk(f * (cx * t0.x + sx * t1.y), f * (cx * t0.y - sx * t1.x));
      const cplx<double> y1 = mk(f * (cx * t1.x + sx * t0.y), f * (cx * t1.y - sx * t0.x));
      F[a] = y0;
      F[M::A + a] = y1;
      ss += abs2(y0) + abs2(y1);
    }
    for (int o = 32; o > 0; o >>= 1) ss += __shfl_xor(ss, o, 64);
    const double inv = 1.0 / sqrt(ss);
    small_sync();
    cplx<R>* To = Tout + (size_t)i * M::S;
#pragma unroll 4
    for (int q = 0; q < (M::S + 63) / 64; ++q) {
      const int x = l + 64 * q;
      if (M::S % 64 == 0 || x < M::S) To[x] = cast<R>(scale(F[x], inv));
    }
    for (int q = l; q < D * M::Q; q += 64) {
      const int k = q / M::Q, o = q % M::Q, a = o / CHI, b = o % CHI;
      const double* lam = lmbd + (size_t)cv.lmbd_pos[k * n + i] * lstride;
      double tr = 0.0;
#pragma unroll
      for (int e = 0; e < CHI; ++e) tr += lam[e];
      msgs_init[(size_t)cv.in_pos[k * n + i] * M::Q + o] = mk<R>(a == b ? (R)(lam[a] / tr) : (R)0, (R)0);
    }
    small_sync();
  }
}

template <int D, typename R>
inline bool launch_small_local_d(int chi, ClassView cv, const cplx<R>* Tin, cplx<R>* Tout, const cplx<double>* canon,
                                 int rows, int cols, double z_time, double x_time, const double* lmbd, int lstride,
                                 cplx<R>* msgs, hipStream_t s) {
#define BQA_SMALL_LOCAL(CH)                                                                                      \
  case CH:                                                                                                       \
    hipLaunchKernelGGL((k_small_local<D, CH, R>), dim3(small_grid(cv.n)), dim3(64), 0, s, cv, Tin, Tout, canon,  \
                       rows, cols, z_time, x_time, lmbd, lstride, msgs);                                         \
    return true;
  switch (chi) {
    BQA_SMALL_LOCAL(2)
    BQA_SMALL_LOCAL(3)
    BQA_SMALL_LOCAL(4)
    default: break;
  }
  if constexpr (D <= 3) {
    switch (chi) {
      BQA_SMALL_LOCAL(5)
      BQA_SMALL_LOCAL(6)
      BQA_SMALL_LOCAL(7)
      BQA_SMALL_LOCAL(8)
      default: break;
    }
  }
#undef BQA_SMALL_LOCAL
  return false;
}

// The fused simple-update tail at a fixed bond dimension; false when the shape
// is not served (the caller runs k_node_local).
template <typename R>
inline bool launch_small_local(ClassView cv, int chi, const cplx<R>* Tin, cplx<R>* Tout, const cplx<double>* canon,
                               int rows, int cols, double z_time, double x_time, const double* lmbd, int lstride,
                               cplx<R>* msgs, hipStream_t s) {
  if (cv.n == 0 || cv.zz == nullptr || !small_shape(cv.d, chi)) return false;
  switch (cv.d) {
    case 2: return launch_small_local_d<2, R>(chi, cv, Tin, Tout, canon, rows, cols, z_time, x_time, lmbd, lstride, msgs, s);
    case 3: return launch_small_local_d<3, R>(chi, cv, Tin, Tout, canon, rows, cols, z_time, x_time, lmbd, lstride, msgs, s);
    case 4: return launch_small_local_d<4, R>(chi, cv, Tin, Tout, canon, rows, cols, z_time, x_time, lmbd, lstride, msgs, s);
    default: return false;
  }
}

}  // namespace bqa
