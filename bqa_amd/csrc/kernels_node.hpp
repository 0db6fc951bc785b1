// Generic per-node kernels for any degree d and bond dimension chi.
//
// One workgroup processes one node at a time (grid-stride over the class).
// The node tensor T[p, b_1..b_d] (p physical, last bond fastest) and two
// intermediates live in LDS when they fit, otherwise in a per-workgroup slice
// of a global workspace.  These kernels are the correctness baseline for every
// (d, chi); the hot shapes have register-tiled kernels in kernels_fast.hpp.
#pragma once

#include "common.hpp"

namespace bqa {

constexpr int kNT = 256;  // threads per workgroup of the generic kernels

struct ClassView {
  int n, d;
  const int32_t* node_ids;  // (n)
  const int32_t* in_pos;    // [d][n] position of message w_k -> v
  const int32_t* out_pos;   // [d][n] position of message v -> w_k
  const int32_t* lmbd_pos;  // [d][n] lambda position of edge (v, w_k)
  const double* h;          // (n)
  const double* J;          // [d][n]
  // ZZ-split gates of the current simple update per edge (lambda position e):
  // g0 = zz[2 e], g1 = zz[2 e + 1] (zz_split of J_e ztime); nullptr outside it
  const cplx<double>* zz;
  // BP launches: the sweep's index in its run with the sweep-slot protocol,
  // -1 for the status protocol (BpStatus)
  int sweep;
};

// One degree class of one job in a launch that covers several jobs (the serving
// batch): the class, its tensors and message buffers, its job's status block, and
// the nodes of the launch's earlier entries.
struct BpEntry {
  ClassView cv;
  const void* T;
  void* mA;
  void* mB;
  BpStatus* st;
  long prefix;
};

enum NodeMode { kModeBP = 0, kModeExt = 1, kModeDensity = 2 };

// dst[.., a, ..] = sum_b m[a, b] src[.., b, ..] along the axis of `stride`.
template <typename R>
__device__ __forceinline__ void mode_product(const cplx<R>* src, cplx<R>* dst,
                                             const cplx<R>* __restrict__ m, int S, int chi,
                                             int stride) {
  for (int idx = threadIdx.x; idx < S; idx += blockDim.x) {
    const int a = (idx / stride) % chi;
    const int base = idx - a * stride;
    const cplx<R>* mr = m + a * chi;
    cplx<R> acc = mk<R>(0, 0);
    for (int b = 0; b < chi; ++b) cfma(acc, mr[b], src[base + b * stride]);
    dst[idx] = acc;
  }
}

// res[o] for o < Q:  o = (p * chi + c) * chi + e when PSPLIT, else o = c * chi + e;
// res = sum over (p, all bonds but the free one) of conj(T[..c..]) Y[..e..].
// `sk` is the stride of the free bond, `rt` = S / chi the number of summed terms.
template <typename R, bool PSPLIT>
__device__ __forceinline__ void gram(const cplx<R>* T, const cplx<R>* Y, int chi, int sk, int rt,
                                     cplx<R>* part, cplx<R>* res) {
  const int cc = chi * chi;
  const int Q = PSPLIT ? 2 * cc : cc;
  const int t = threadIdx.x;
  if (Q <= (int)blockDim.x) {
    const int G = blockDim.x / Q;
    const int o = t % Q, g = t / Q;
    cplx<R> acc = mk<R>(0, 0);
    if (g < G) {
      const int p = PSPLIT ? o / cc : 0;
      const int oo = PSPLIT ? o % cc : o;
      const int c = oo / chi, e = oo % chi;
      const int r0 = PSPLIT ? p * (rt / 2) : 0;
      const int r1 = PSPLIT ? r0 + rt / 2 : rt;
      for (int r = r0 + g; r < r1; r += G) {
        const int hi = r / sk, lo = r - hi * sk;
        const int base = hi * chi * sk + lo;
        cfma_cj(acc, T[base + c * sk], Y[base + e * sk]);
      }
    }
    part[t] = acc;
    __syncthreads();
    if (t < Q) {
      cplx<R> s = part[t];
      for (int gg = 1; gg < G; ++gg) s = s + part[t + gg * Q];
      res[t] = s;
    }
    __syncthreads();
  } else {
    for (int o = t; o < Q; o += blockDim.x) {
      const int p = PSPLIT ? o / cc : 0;
      const int oo = PSPLIT ? o % cc : o;
      const int c = oo / chi, e = oo % chi;
      const int r0 = PSPLIT ? p * (rt / 2) : 0;
      const int r1 = PSPLIT ? r0 + rt / 2 : rt;
      cplx<R> acc = mk<R>(0, 0);
      for (int r = r0; r < r1; ++r) {
        const int hi = r / sk, lo = r - hi * sk;
        const int base = hi * chi * sk + lo;
        cfma_cj(acc, T[base + c * sk], Y[base + e * sk]);
      }
      res[o] = acc;
    }
    __syncthreads();
  }
}

// Dynamic LDS: part[kNT] | res[4 chi^2] | (T, X, Y)[S each] when ws == nullptr.
template <typename R>
__host__ __device__ inline size_t node_lds_bytes(int chi, int S, bool with_buffers) {
  size_t b = sizeof(cplx<R>) * (size_t)(kNT + 4 * chi * chi);
  if (with_buffers) b += sizeof(cplx<R>) * 3 * (size_t)S;
  return b;
}

// MODE kModeBP: new messages into the other buffer, dist partial maxima.
// MODE kModeExt: 2chi x 2chi messages with the output bond extended by the ZZ split.
// MODE kModeDensity: rho (n, 2, 2) complex128 and Bloch vectors by qubit id.
template <typename R, int MODE>
__global__ __launch_bounds__(kNT) void k_node_contract(
    ClassView cv, int chi, const cplx<R>* __restrict__ T, cplx<R>* msgA, cplx<R>* msgB,
    BpStatus* st, cplx<R>* ext_out, double ztime, double* rho_out, double* bloch_out,
    cplx<R>* gws, long ws_per_block) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  __shared__ double shred[16];
  cplx<R>* part = reinterpret_cast<cplx<R>*>(smem_raw);
  cplx<R>* res = part + kNT;
  const int d = cv.d, n = cv.n;
  const int A = ipow(chi, d);
  const int S = 2 * A;
  const int cc = chi * chi;
  cplx<R>* base = gws ? gws + (size_t)blockIdx.x * ws_per_block : res + 4 * cc;
  cplx<R>* bufT = base;
  cplx<R>* bufX = base + S;
  cplx<R>* bufY = base + 2 * S;

  const cplx<R>* min = msgA;
  cplx<R>* mout = msgB;
  if (MODE == kModeBP) {
    const int cur = bp_begin(st, cv.sweep);
    if (cur < 0) return;
    min = cur ? msgB : msgA;
    mout = cur ? msgA : msgB;
  }
  double dnum = 0.0, dden = 0.0;

  for (int i = blockIdx.x; i < n; i += gridDim.x) {
    const cplx<R>* Ti = T + (size_t)i * S;
    for (int x = threadIdx.x; x < S; x += blockDim.x) bufT[x] = Ti[x];
    __syncthreads();
    if (MODE == kModeDensity) {
      const cplx<R>* src = bufT;
      for (int j = 0; j < d; ++j) {
        cplx<R>* dst = (src == bufX) ? bufY : bufX;
        mode_product(src, dst, min + (size_t)cv.in_pos[j * n + i] * cc, S, chi, ipow(chi, d - 1 - j));
        __syncthreads();
        src = dst;
      }
      // rho[p, q] = sum_a Y[p, a] conj(T[q, a])
      const int t = threadIdx.x;
      const int o = t & 3, g = t >> 2, G = blockDim.x >> 2;
      const int p = o >> 1, q = o & 1;
      cplx<R> acc = mk<R>(0, 0);
      for (int a = g; a < A; a += G) cfma_cj(acc, bufT[q * A + a], src[p * A + a]);
      part[t] = acc;
      __syncthreads();
      if (t < 4) {
        cplx<R> s = part[t];
        for (int gg = 1; gg < G; ++gg) s = s + part[t + 4 * gg];
        res[t] = s;
      }
      __syncthreads();
      if (t == 0) {
        cplx<double> r00 = cast<double>(res[0]), r01 = cast<double>(res[1]);
        cplx<double> r10 = cast<double>(res[2]), r11 = cast<double>(res[3]);
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
      __syncthreads();
      continue;
    }
    for (int k = 0; k < d; ++k) {
      const cplx<R>* src = bufT;
      for (int j = 0; j < d; ++j) {
        if (j == k) continue;
        cplx<R>* dst = (src == bufX) ? bufY : bufX;
        mode_product(src, dst, min + (size_t)cv.in_pos[j * n + i] * cc, S, chi, ipow(chi, d - 1 - j));
        __syncthreads();
        src = dst;
      }
      const int sk = ipow(chi, d - 1 - k);
      const long pos = cv.out_pos[k * n + i];
      if (MODE == kModeBP) {
        gram<R, false>(bufT, src, chi, sk, S / chi, part, res);
        // The message is Hermitian by construction: keep the Hermitian part of
        // the Gram, normalised by its (real) trace.
        double tr = 0.0;
        for (int c = 0; c < chi; ++c) tr += (double)res[c * chi + c].x;
        const double inv = 1.0 / tr;
        for (int o = threadIdx.x; o < cc; o += blockDim.x) {
          const int c = o / chi, e = o % chi;
          const cplx<double> a = cast<double>(res[o]), b = cast<double>(res[e * chi + c]);
          const cplx<R> v = cast<R>(mk(0.5 * (a.x + b.x) * inv, 0.5 * (a.y - b.y) * inv));
          const cplx<R> old = min[pos * cc + o];
          dnum = fmax(dnum, (double)cabs(v - old));
          dden = fmax(dden, (double)cabs(v + old));
          mout[pos * cc + o] = v;
        }
      } else {
        gram<R, true>(bufT, src, chi, sk, S / chi, part, res);
        cplx<double> g0, g1;
        zz_split(cv.J[k * n + i] * ztime, g0, g1);
        const cplx<double> gs[2] = {g0, g1};
        const int n2 = 2 * chi;
        cplx<double> tr = mk(0.0, 0.0);
        for (int s = 0; s < 2; ++s) {
          const double w = abs2(gs[s]);
          for (int c = 0; c < chi; ++c) {
            const cplx<double> v = cast<double>(res[c * chi + c]) +
                                   (s == 0 ? cast<double>(res[cc + c * chi + c])
                                           : cast<double>(res[cc + c * chi + c]));
            tr = tr + scale(v, w);
          }
        }
        const cplx<double> inv = cinv(tr);
        for (int o = threadIdx.x; o < n2 * n2; o += blockDim.x) {
          const int row = o / n2, col = o % n2;
          const int s = row / chi, c = row % chi, s2 = col / chi, e = col % chi;
          const cplx<double> g0v = cast<double>(res[c * chi + e]);
          const cplx<double> g1v = cast<double>(res[cc + c * chi + e]);
          const cplx<double> gsum = ((s + s2) & 1) ? g0v - g1v : g0v + g1v;
          const cplx<double> v = conj(gs[s]) * gs[s2] * gsum * inv;
          ext_out[pos * n2 * n2 + o] = cast<R>(v);
        }
      }
      __syncthreads();
    }
  }
  if (MODE == kModeBP) {
    dnum = block_max(dnum, shred);
    dden = block_max(dden, shred);
    if (threadIdx.x == 0) {
      bp_end(st, cv.sweep, dnum, dden);
    }
  }
}

enum LocalFlags { kAbsorb = 1, kGates = 2, kGauge = 4 };

// Absorb canonicalisers (with the ZZ split), then the Z and X layers, then the
// symmetric gauge with per-node L2 normalisation and message re-initialisation.
// Tin: (n, 2, chi_in^d) -> Tout: (n, 2, chi_out^d).  Dynamic LDS:
// 2 * max(S) complex buffers when ws == nullptr.
// R: storage type of tensors and messages; RC: arithmetic type (double for both
// job precisions: FP32 rounding inside the absorb contraction is amplified by
// the S^{-1/2} of the canonicalisers).
template <typename R, typename RC>
__global__ __launch_bounds__(kNT) void k_node_local(
    ClassView cv, int chi_in, int chi_out, const cplx<R>* __restrict__ Tin, cplx<R>* Tout,
    const cplx<double>* __restrict__ canon, int canon_rows, int canon_cols, double zz_time,
    int flags, double z_time, double x_time, const double* __restrict__ lmbd, int lstride,
    cplx<R>* msgs_init, cplx<RC>* gws, long ws_per_block) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  __shared__ double shred[16];
  const int d = cv.d, n = cv.n;
  const int cmax = chi_in > chi_out ? chi_in : chi_out;
  const int Smax = 2 * ipow(cmax, d);
  cplx<RC>* base = gws ? gws + (size_t)blockIdx.x * ws_per_block : reinterpret_cast<cplx<RC>*>(smem_raw);
  cplx<RC>* X = base;
  cplx<RC>* Y = base + Smax;
  const int S_in = 2 * ipow(chi_in, d);
  const int A_out = ipow(chi_out, d);
  const int S_out = 2 * A_out;

  for (int i = blockIdx.x; i < n; i += gridDim.x) {
    const cplx<R>* Ti = Tin + (size_t)i * S_in;
    for (int x = threadIdx.x; x < S_in; x += blockDim.x) X[x] = cast<RC>(Ti[x]);
    __syncthreads();
    if (flags & kAbsorb) {
      for (int k = 0; k < d; ++k) {
        const cplx<double>* C = canon + (size_t)cv.in_pos[k * n + i] * canon_rows * canon_cols;
        cplx<double> g0, g1;
        zz_split(cv.J[k * n + i] * zz_time, g0, g1);
        const int pk = ipow(chi_out, k);            // digits of the bonds already done
        const int outer = 2 * pk;
        const int inner = ipow(chi_in, d - k - 1);
        const int size_out = outer * chi_out * inner;
        for (int y = threadIdx.x; y < size_out; y += blockDim.x) {
          const int lo = y % inner;
          const int rest = y / inner;
          const int nk = rest % chi_out;
          const int oi = rest / chi_out;
          const double z = (oi / pk) ? -1.0 : 1.0;
          const cplx<double> gz = scale(g1, z);
          cplx<RC> acc = mk<RC>(0, 0);
          const cplx<RC>* xs = X + (size_t)oi * chi_in * inner + lo;
          for (int c = 0; c < chi_in; ++c) {
            const cplx<double> w = g0 * C[c * canon_cols + nk] + gz * C[(chi_in + c) * canon_cols + nk];
            cfma(acc, xs[c * inner], cast<RC>(w));
          }
          Y[y] = acc;
        }
        __syncthreads();
        cplx<RC>* tmp = X;
        X = Y;
        Y = tmp;
      }
    }
    if (flags & kGates) {
      const double phi = cv.h[i] * z_time;
      const double cz = cos(phi), sz = sin(phi);
      const double cx = cos(x_time), sx = sin(x_time);
      const cplx<RC> f0 = mk<RC>((RC)cz, (RC)-sz), f1 = mk<RC>((RC)cz, (RC)sz);
      for (int a = threadIdx.x; a < A_out; a += blockDim.x) {
        const cplx<RC> t0 = X[a] * f0, t1 = X[A_out + a] * f1;
        // (cos x) t - i (sin x) X t
        X[a] = mk<RC>((RC)(cx * t0.x + sx * t1.y), (RC)(cx * t0.y - sx * t1.x));
        X[A_out + a] = mk<RC>((RC)(cx * t1.x + sx * t0.y), (RC)(cx * t1.y - sx * t0.x));
      }
      __syncthreads();
    }
    if (flags & kGauge) {
      double ss = 0.0;
      for (int y = threadIdx.x; y < S_out; y += blockDim.x) {
        double f = 1.0;
        int rem = y;
        for (int k = d - 1; k >= 0; --k) {
          const int nk = rem % chi_out;
          rem /= chi_out;
          f *= sqrt(lmbd[(size_t)cv.lmbd_pos[k * n + i] * lstride + nk]);
        }
        const cplx<RC> v = scale(X[y], (RC)f);
        X[y] = v;
        ss += (double)abs2(v);
      }
      ss = block_sum(ss, shred);
      const RC inv = (RC)(1.0 / sqrt(ss));
      for (int y = threadIdx.x; y < S_out; y += blockDim.x) X[y] = scale(X[y], inv);
      const int cc = chi_out * chi_out;
      for (int k = 0; k < d; ++k) {
        const double* lam = lmbd + (size_t)cv.lmbd_pos[k * n + i] * lstride;
        double tr = 0.0;
        for (int a = 0; a < chi_out; ++a) tr += lam[a];
        const long pos = cv.in_pos[k * n + i];
        for (int o = threadIdx.x; o < cc; o += blockDim.x) {
          const int a = o / chi_out, b = o % chi_out;
          msgs_init[pos * cc + o] = mk<R>(a == b ? (R)(lam[a] / tr) : (R)0, (R)0);
        }
      }
      __syncthreads();
    }
    cplx<R>* To = Tout + (size_t)i * S_out;
    for (int x = threadIdx.x; x < S_out; x += blockDim.x) To[x] = cast<R>(X[x]);
    __syncthreads();
  }
}

}  // namespace bqa
