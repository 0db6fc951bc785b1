// Exact state-vector simulation of the Trotterised anneal (src/bqa/exact_sim.py:14-88),
// FP64, up to 32 qubits on one GPU (2^n complex128 amplitudes + 2^n doubles of
// diagonal energies: 24 GB at 30 qubits, 96 GB at 32).
//
// Qubit q is bit q of the basis index; Z_q = +1 on bit 0.  One Trotter layer is
//   psi <- prod_q exp(-i xtime X_q) exp(-i ztime Ez) psi,
//   Ez(x) = sum_q h_q z_q(x) + sum_(ab) J_ab z_a(x) z_b(x),
// which is the reference's order (every Z gate, every ZZ gate, then every X
// gate, exact_sim.py:51-63); the Z and ZZ gates are diagonal and the X gates on
// different qubits commute.
//
// The X layer runs in passes over groups of qubits, each pass a tile kernel
// that stages 1024 amplitudes in LDS and applies the rotations of its group
// there: the first pass takes 1024 contiguous amplitudes (qubits 0..9, and the
// diagonal phase on the way in); later passes take 16 contiguous amplitudes
// (qubits 0..3, for 256-byte rows) times the 2^t values of t <= 6 higher
// qubits.  The readout uses the same tiles: each pass accumulates, for the
// qubits of its group, |psi|^2 split by the qubit's value and the off-diagonal
// sum conj(psi[x]) psi[x | 2^q] over pairs inside the tile.
#pragma once

#include "common.hpp"

namespace bqa {

constexpr int kSvTileBits = 10;            // amplitudes per tile: 1024 (16 KB of complex128)
constexpr int kSvRowBits = 4;              // contiguous amplitudes per row in the high-qubit passes
constexpr int kSvThreads = 256;
constexpr int kSvMaxEdges = 2048;          // edges kept in LDS by the energy kernel

// Ez(x) for every basis state; fields and edges staged in LDS.
__global__ __launch_bounds__(kSvThreads) void k_sv_energy(double* __restrict__ ez, int n, const double* __restrict__ h,
                                                          int E, const int32_t* __restrict__ lhs,
                                                          const int32_t* __restrict__ rhs,
                                                          const double* __restrict__ J) {
  __shared__ double sh[32];
  __shared__ int sa[kSvMaxEdges], sb[kSvMaxEdges];
  __shared__ double sj[kSvMaxEdges];
  for (int q = threadIdx.x; q < n; q += blockDim.x) sh[q] = h[q];
  for (int e = threadIdx.x; e < E; e += blockDim.x) {
    sa[e] = lhs[e];
    sb[e] = rhs[e];
    sj[e] = J[e];
  }
  __syncthreads();
  const unsigned long long dim = 1ULL << n;
  for (unsigned long long x = blockIdx.x * (unsigned long long)blockDim.x + threadIdx.x; x < dim;
       x += (unsigned long long)gridDim.x * blockDim.x) {
    double e = 0.0;
    for (int q = 0; q < n; ++q) e += ((x >> q) & 1ULL) ? -sh[q] : sh[q];
    for (int k = 0; k < E; ++k) e += (((x >> sa[k]) ^ (x >> sb[k])) & 1ULL) ? -sj[k] : sj[k];
    ez[x] = e;
  }
}

// |->^n: psi[x] = (-1)^popcount(x) 2^(-n/2).
__global__ void k_sv_init(cplx<double>* __restrict__ psi, int n) {
  const unsigned long long dim = 1ULL << n;
  const double a = exp2(-0.5 * n);
  for (unsigned long long x = blockIdx.x * (unsigned long long)blockDim.x + threadIdx.x; x < dim;
       x += (unsigned long long)gridDim.x * blockDim.x)
    psi[x] = mk(__popcll(x) & 1 ? -a : a, 0.0);
}

// Global index of tile element k (0 <= k < 16 * 2^t) in a pass over qubits b0..b0+t-1
// (b0 >= kSvRowBits): row r = k >> 4 carries the t target bits, c = k & 15 the low bits,
// the tile number T the remaining bits (those between 4 and b0, then those above b0 + t).
__device__ __forceinline__ unsigned long long sv_high_index(unsigned long long T, int k, int b0, int t) {
  const unsigned long long c = (unsigned long long)(k & ((1 << kSvRowBits) - 1));
  const unsigned long long r = (unsigned long long)(k >> kSvRowBits);
  const int mid = b0 - kSvRowBits;
  const unsigned long long lo = T & ((1ULL << mid) - 1ULL);
  const unsigned long long hi = T >> mid;
  return c | (lo << kSvRowBits) | (r << b0) | (hi << (b0 + t));
}

// Rotations exp(-i xtime X) on the tile-local bits [0, nb) of an LDS tile: amplitude k's
// partner for local bit b is k ^ (1 << (b + shift)).
__device__ __forceinline__ void sv_rotate_tile(cplx<double>* s, int size, int nb, int shift, double c, double sn) {
  const int half = size >> 1;
  for (int b = 0; b < nb; ++b) {
    const int m = 1 << (b + shift);
    for (int p = threadIdx.x; p < half; p += blockDim.x) {
      // p-th pair: insert a zero at bit position (b + shift)
      const int k0 = ((p >> (b + shift)) << (b + shift + 1)) | (p & (m - 1));
      const int k1 = k0 | m;
      const cplx<double> a0 = s[k0], a1 = s[k1];
      // (cos t) a - i (sin t) X a
      s[k0] = mk(c * a0.x + sn * a1.y, c * a0.y - sn * a1.x);
      s[k1] = mk(c * a1.x + sn * a0.y, c * a1.y - sn * a0.x);
    }
    __syncthreads();
  }
}

// Per-qubit readout sums of the tile-local bits [0, nb) (partner bit b + shift), added to this
// wave's slots wacc[4 b + k]: r00, r11, Re r10, Im r10 with r10 = sum conj(psi0) psi1.  Each
// wave reduces its lanes with shuffles and its lane 0 alone writes its slots, so the order of
// every sum is fixed.
__device__ __forceinline__ void sv_read_tile(const cplx<double>* s, int size, int nb, int shift, double* wacc) {
  const int half = size >> 1;
  const int l = threadIdx.x & 63;
  for (int b = 0; b < nb; ++b) {
    const int m = 1 << (b + shift);
    double v[4] = {0.0, 0.0, 0.0, 0.0};
    for (int p = threadIdx.x; p < half; p += blockDim.x) {
      const int k0 = ((p >> (b + shift)) << (b + shift + 1)) | (p & (m - 1));
      const cplx<double> a0 = s[k0], a1 = s[k0 | m];
      v[0] += abs2(a0);
      v[1] += abs2(a1);
      v[2] += a0.x * a1.x + a0.y * a1.y;
      v[3] += a0.x * a1.y - a0.y * a1.x;
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      for (int o = 32; o > 0; o >>= 1) v[k] += __shfl_xor(v[k], o, 64);
      if (l == 0) wacc[4 * b + k] += v[k];
    }
  }
}

// Block partials part[block][4 nb] from the per-wave slots.
__device__ __forceinline__ void sv_store_partials(const double (*wacc)[4 * kSvTileBits], int nb, double* part) {
  __syncthreads();
  for (int k = threadIdx.x; k < 4 * nb; k += blockDim.x) {
    double v = 0.0;
    for (int w = 0; w < kSvThreads / 64; ++w) v += wacc[w][k];
    part[(size_t)blockIdx.x * 4 * kSvTileBits + k] = v;
  }
}

// First pass: tiles of 2^tb contiguous amplitudes (tb = min(n, 10)).  MODE 0: diagonal phase
// exp(-i ztime Ez) (when ez != nullptr), then the X rotations of qubits 0..tb-1.  MODE 1: readout
// sums of qubits 0..tb-1 into part.
template <int MODE>
__global__ __launch_bounds__(kSvThreads) void k_sv_low(cplx<double>* __restrict__ psi, int n,
                                                       const double* __restrict__ ez, double ztime, double c,
                                                       double sn, double* part) {
  __shared__ cplx<double> s[1 << kSvTileBits];
  const int tb = n < kSvTileBits ? n : kSvTileBits;
  const int size = 1 << tb;
  const unsigned long long tiles = 1ULL << (n - tb);
  __shared__ double wacc[kSvThreads / 64][4 * kSvTileBits];
  for (int k = threadIdx.x; k < kSvThreads / 64 * 4 * kSvTileBits; k += blockDim.x) (&wacc[0][0])[k] = 0.0;
  __syncthreads();
  for (unsigned long long T = blockIdx.x; T < tiles; T += gridDim.x) {
    const unsigned long long base = T << tb;
    for (int k = threadIdx.x; k < size; k += blockDim.x) {
      cplx<double> a = psi[base + k];
      if (MODE == 0 && ez) {
        double sp, cp;
        sincos(-ztime * ez[base + k], &sp, &cp);
        a = a * mk(cp, sp);
      }
      s[k] = a;
    }
    __syncthreads();
    if (MODE == 0) {
      sv_rotate_tile(s, size, tb, 0, c, sn);
      for (int k = threadIdx.x; k < size; k += blockDim.x) psi[base + k] = s[k];
    } else {
      sv_read_tile(s, size, tb, 0, wacc[threadIdx.x >> 6]);
    }
    __syncthreads();
  }
  if (MODE == 1) sv_store_partials(wacc, tb, part);
}

// Later passes: qubits b0..b0+t-1 (t <= 6, b0 >= 4), tiles of 16 x 2^t amplitudes.
template <int MODE>
__global__ __launch_bounds__(kSvThreads) void k_sv_high(cplx<double>* __restrict__ psi, int n, int b0, int t,
                                                        double c, double sn, double* part) {
  __shared__ cplx<double> s[1 << kSvTileBits];
  const int size = 1 << (kSvRowBits + t);
  const unsigned long long tiles = 1ULL << (n - kSvRowBits - t);
  __shared__ double wacc[kSvThreads / 64][4 * kSvTileBits];
  for (int k = threadIdx.x; k < kSvThreads / 64 * 4 * kSvTileBits; k += blockDim.x) (&wacc[0][0])[k] = 0.0;
  __syncthreads();
  for (unsigned long long T = blockIdx.x; T < tiles; T += gridDim.x) {
    for (int k = threadIdx.x; k < size; k += blockDim.x) s[k] = psi[sv_high_index(T, k, b0, t)];
    __syncthreads();
    if (MODE == 0) {
      sv_rotate_tile(s, size, t, kSvRowBits, c, sn);
      for (int k = threadIdx.x; k < size; k += blockDim.x) psi[sv_high_index(T, k, b0, t)] = s[k];
    } else {
      sv_read_tile(s, size, t, kSvRowBits, wacc[threadIdx.x >> 6]);
    }
    __syncthreads();
  }
  if (MODE == 1) sv_store_partials(wacc, t, part);
}

// Bloch vectors and densities of qubits q0..q0+nb-1 from the block partials of one readout pass.
__global__ void k_sv_finish(const double* __restrict__ part, int blocks, int q0, int nb, double* rho, double* bloch) {
  const int b = threadIdx.x;
  if (b >= nb) return;
  double s[4] = {0.0, 0.0, 0.0, 0.0};
  for (int blk = 0; blk < blocks; ++blk)
    for (int k = 0; k < 4; ++k) s[k] += part[(size_t)blk * 4 * kSvTileBits + 4 * b + k];
  const double tr = s[0] + s[1];
  const double r00 = s[0] / tr, r11 = s[1] / tr, r10r = s[2] / tr, r10i = s[3] / tr;
  const int q = q0 + b;
  // rho[p][q'] = sum psi_p conj(psi_q'): rho10 = r10, rho01 = conj(r10)
  if (rho) {
    double* ro = rho + 8 * (size_t)q;
    ro[0] = r00; ro[1] = 0.0; ro[2] = r10r; ro[3] = -r10i;
    ro[4] = r10r; ro[5] = r10i; ro[6] = r11; ro[7] = 0.0;
  }
  if (bloch) {
    double* bo = bloch + 3 * (size_t)q;
    bo[0] = 2.0 * r10r;
    bo[1] = 2.0 * r10i;
    bo[2] = r00 - r11;
  }
}

}  // namespace bqa
