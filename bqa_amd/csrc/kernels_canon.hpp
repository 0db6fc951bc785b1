// Bond canonicalisers and Vidal lambdas in FP64 (src/bqa/state.py:171-200).
//
// Per undirected edge i (forward message i, backward message E + i), with
// n = 2 chi:
//   M = U S V^H        (masked thin SVD, s > pinv_eps; backends.py:709-717)
//   ul = U S^{-1/2},  lu = S^{1/2} V^H                       (backends.py:483-490)
//   K = lu_f lu_b^T  (no conjugation),  K = Us L Vs^H (masked)   (state.py:186-189)
//   C_f = ul_f Us,  C_b = ul_b conj(Vs),  lambda = L / |L|_2    (state.py:196-200)
// C_b is stored at position i and C_f at position E + i (state.py:182-183), so
// the bond of node v towards u uses the canonicaliser built from m_{v->u}.
//
// Every SVD is a one-sided (Hestenes) Jacobi on the columns of the matrix,
// parallel round-robin ordering, one workgroup (a single wave) per edge, with
// singular values sorted in descending order like LAPACK's.
#pragma once

#include "common.hpp"

namespace bqa {

constexpr int kCanonNT = 64;

using zd = cplx<double>;

// Pairs of round `r` in the round-robin tournament over n (even) columns.
__device__ __forceinline__ void rr_pair(int k, int r, int n, int& p, int& q) {
  const int m = n - 1;
  p = k == 0 ? 0 : ((k - 1 + r) % m) + 1;
  q = ((n - 2 - k + r) % m) + 1;
}

// One-sided Jacobi: A (n x n, column-major A[col*n + row]) is rotated in
// place so its columns become orthogonal; V accumulates the rotations
// (A_in V = A_out).  Columns of A_out are U S.
__device__ void jacobi_svd(zd* A, zd* V, int n, int* flag) {
  const int P = n / 2;
  int L = 1;
  while (L * 2 * P <= (int)blockDim.x && L * 2 <= 64) L *= 2;
  const int t = threadIdx.x;
  const int pair = t / L, lane = t % L;
  const bool active = pair < P;
  for (int x = t; x < n * n; x += blockDim.x) V[x] = mk((x / n) == (x % n) ? 1.0 : 0.0, 0.0);
  const double tol = 2.2204460492503131e-16 * n;
  __syncthreads();
  for (int sweep = 0; sweep < 40; ++sweep) {
    if (t == 0) *flag = 0;
    __syncthreads();
    for (int r = 0; r < n - 1; ++r) {
      int p = 0, q = 1;
      if (active) rr_pair(pair, r, n, p, q);
      double a = 0.0, b = 0.0;
      zd g = mk(0.0, 0.0);
      if (active) {
        for (int row = lane; row < n; row += L) {
          const zd ap = A[p * n + row], aq = A[q * n + row];
          a += abs2(ap);
          b += abs2(aq);
          cfma_cj(g, ap, aq);
        }
      }
      for (int o = L >> 1; o > 0; o >>= 1) {
        a += __shfl_xor(a, o, 64);
        b += __shfl_xor(b, o, 64);
        g.x += __shfl_xor(g.x, o, 64);
        g.y += __shfl_xor(g.y, o, 64);
      }
      if (active) {
        const double ga = hypot(g.x, g.y);
        if (ga > tol * sqrt(a * b) && ga > 1e-300) {
          if (lane == 0) *flag = 1;
          const double zeta = (b - a) / (2.0 * ga);
          const double tt = (zeta >= 0.0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
          const double c = 1.0 / sqrt(1.0 + tt * tt);
          const double s = c * tt;
          const zd e = mk(g.x / ga, -g.y / ga);  // e^{-i phi}
          for (int row = lane; row < n; row += L) {
            const zd ap = A[p * n + row], aq = e * A[q * n + row];
            A[p * n + row] = scale(ap, c) - scale(aq, s);
            A[q * n + row] = scale(ap, s) + scale(aq, c);
            const zd vp = V[p * n + row], vq = e * V[q * n + row];
            V[p * n + row] = scale(vp, c) - scale(vq, s);
            V[q * n + row] = scale(vp, s) + scale(vq, c);
          }
        }
      }
      __syncthreads();
    }
    const int f = *flag;
    __syncthreads();
    if (!f) break;
  }
}

// Column norms of A -> sig, and perm[rank] = column in descending order.
__device__ void sorted_norms(const zd* A, int n, double* sig, int* perm) {
  for (int j = threadIdx.x; j < n; j += blockDim.x) {
    double s = 0.0;
    for (int r = 0; r < n; ++r) s += abs2(A[j * n + r]);
    sig[j] = sqrt(s);
  }
  __syncthreads();
  for (int j = threadIdx.x; j < n; j += blockDim.x) {
    int rank = 0;
    for (int k = 0; k < n; ++k) rank += (sig[k] > sig[j]) || (sig[k] == sig[j] && k < j);
    perm[rank] = j;
  }
  __syncthreads();
}

// Workspace doubles per edge: 6 n^2 complex + small arrays.
__host__ __device__ inline size_t canon_ws_bytes(int n) {
  return sizeof(zd) * 6 * (size_t)n * n + sizeof(double) * 2 * n + sizeof(int) * 2 * n + 64;
}

template <typename R>
__global__ __launch_bounds__(kCanonNT) void k_canonicalize(
    const cplx<R>* __restrict__ ext, int E, int chi, int ccols, double pinv_eps, double dtype_eps,
    zd* canon, double* lfull, unsigned char* gws, long ws_per_block) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  __shared__ int flag;
  const int n = 2 * chi;
  const int nn = n * n;
  unsigned char* wsb = gws ? gws + (size_t)blockIdx.x * ws_per_block : smem_raw;
  zd* A = reinterpret_cast<zd*>(wsb);
  zd* V = A + nn;
  zd* ULf = V + nn;
  zd* LUf = ULf + nn;
  zd* ULb = LUf + nn;
  zd* LUb = ULb + nn;
  double* sig = reinterpret_cast<double*>(LUb + nn);
  double* sval = sig + n;
  int* perm = reinterpret_cast<int*>(sval + n);
  const int t = threadIdx.x;

  for (int i = blockIdx.x; i < E; i += gridDim.x) {
    for (int side = 0; side < 2; ++side) {
      // side 0: forward message i; side 1: backward message E + i
      const cplx<R>* M = ext + (size_t)(side ? E + i : i) * nn;
      for (int x = t; x < nn; x += blockDim.x) {
        const int row = x / n, col = x % n;
        A[col * n + row] = cast<double>(M[x]);
      }
      __syncthreads();
      jacobi_svd(A, V, n, &flag);
      sorted_norms(A, n, sig, perm);
      zd* UL = side ? ULb : ULf;
      zd* LU = side ? LUb : LUf;
      for (int x = t; x < nn; x += blockDim.x) {
        const int jj = x / n, r = x % n;     // sorted column jj, row r
        const int j = perm[jj];
        const double s = sig[j];
        const bool keep = s > pinv_eps;
        const double sq = keep ? sqrt(s) : 0.0;
        const double pinv = sq > dtype_eps ? 1.0 / sq : 0.0;
        // ul[:, jj] = U[:, jj] pinv(sqrt(s)),  U = A / s
        UL[jj * n + r] = keep ? scale(A[j * n + r], pinv / s) : mk(0.0, 0.0);
        // lu[jj, r] = sqrt(s) conj(V[r, j])
        LU[jj * n + r] = keep ? scale(conj(V[j * n + r]), sq) : mk(0.0, 0.0);
      }
      __syncthreads();
    }
    // K[a][b] = sum_r lu_f[a, r] lu_b[b, r], stored column-major in A
    for (int x = t; x < nn; x += blockDim.x) {
      const int b = x / n, a = x % n;
      zd acc = mk(0.0, 0.0);
      for (int r = 0; r < n; ++r) cfma(acc, LUf[a * n + r], LUb[b * n + r]);
      A[b * n + a] = acc;
    }
    __syncthreads();
    jacobi_svd(A, V, n, &flag);
    sorted_norms(A, n, sig, perm);
    for (int jj = t; jj < n; jj += blockDim.x) {
      const double s = sig[perm[jj]];
      sval[jj] = s > pinv_eps ? s : 0.0;
    }
    __syncthreads();
    double nrm = 0.0;
    for (int k = 0; k < n; ++k) nrm += sval[k] * sval[k];
    nrm = sqrt(nrm);
    for (int jj = t; jj < n; jj += blockDim.x) lfull[(size_t)i * n + jj] = sval[jj] / nrm;
    // C_f[a][c] = sum_r ul_f[a, r] Us[r, c];  C_b[a][c] = sum_r ul_b[a, r] conj(Vs[r, c])
    zd* Cb = canon + (size_t)i * n * ccols;
    zd* Cf = canon + (size_t)(E + i) * n * ccols;
    for (int x = t; x < n * ccols; x += blockDim.x) {
      const int a = x / ccols, c = x % ccols;
      const int j = perm[c];
      const double s = sig[j];
      zd af = mk(0.0, 0.0), ab = mk(0.0, 0.0);
      if (s > pinv_eps) {
        for (int r = 0; r < n; ++r) {
          cfma(af, ULf[r * n + a], A[j * n + r]);
          cfma(ab, ULb[r * n + a], conj(V[j * n + r]));
        }
        af = scale(af, 1.0 / s);
      }
      Cf[x] = af;
      Cb[x] = ab;
    }
    __syncthreads();
  }
}

// lmax[k] = max_e lambda_e[k] over all edges (row stride n).
__global__ void k_lambda_max(const double* __restrict__ lfull, int E, int n,
                             unsigned long long* lmax_bits) {
  __shared__ unsigned long long sm[128];
  for (int k = threadIdx.x; k < n; k += blockDim.x) sm[k] = 0ULL;
  __syncthreads();
  const long total = (long)E * n;
  for (long x = blockIdx.x * (long)blockDim.x + threadIdx.x; x < total;
       x += (long)gridDim.x * blockDim.x) {
    double v = lfull[x];
    if (!(v >= 0.0)) v = 0.0;
    atomicMax(&sm[x % n], (unsigned long long)__double_as_longlong(v));
  }
  __syncthreads();
  for (int k = threadIdx.x; k < n; k += blockDim.x) atomicMax(&lmax_bits[k], sm[k]);
}

}  // namespace bqa
