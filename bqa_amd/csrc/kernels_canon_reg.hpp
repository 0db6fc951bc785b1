// Register-resident FP64 canonicalisers for bond dimensions chi <= 16 (n = 2 chi <= 32).
//
// Same mathematics and masks as k_canonicalize (kernels_canon.hpp), laid out
// for throughput: P = n/2 lanes per edge, each lane holding two matrix columns
// (and two columns of V) in registers for the whole one-sided Jacobi.  A round
// rotates the lane's column pair, then the columns move between neighbouring
// lanes (round-robin tournament: lane 0's first column stays, the others
// rotate), so n - 1 rounds visit every pair without touching LDS.  Several
// edges share a wave (64 / P2 of them, P2 = P rounded up to a power of two).
//
// The two extended messages are Hermitian positive semi-definite, so their
// left and right singular vectors coincide on the kept subspace; their
// Jacobi runs without accumulating V and uses U for both ul and lu.  Columns
// whose norm is below 1e-13 of the largest column are left out of rotations
// (rank-deficient matrices otherwise keep rotating rounding noise).
#pragma once

#include "common.hpp"

namespace bqa {

template <int N>
struct RegCanonShape {
  static constexpr int P = N / 2;
  static constexpr int P2 = P <= 1 ? 1 : P <= 2 ? 2 : P <= 4 ? 4 : P <= 8 ? 8 : 16;
  static constexpr int EPW = 64 / P2;                 // edges per wave
  static constexpr int WS = 4 * N * N + 4 * N;        // workspace doubles per edge
};

template <int N>
struct RCol {
  double re[N], im[N];
};

template <int N>
__device__ __forceinline__ double col_norm2(const RCol<N>& c) {
  double s = 0.0;
#pragma unroll
  for (int r = 0; r < N; ++r) s = fma(c.re[r], c.re[r], fma(c.im[r], c.im[r], s));
  return s;
}

template <int N>
__device__ __forceinline__ void col_zero(RCol<N>& c) {
#pragma unroll
  for (int r = 0; r < N; ++r) {
    c.re[r] = 0.0;
    c.im[r] = 0.0;
  }
}

// Move columns to the next round's owners: new_top[k] = top[k+1] (k >= 1,
// k < P-1), top[P-1] <- bot[P-1], top[0] fixed; new_bot[k] = bot[k-1]
// (k >= 1), bot[0] <- top[1].
template <int N>
__device__ __forceinline__ void rotate_owners(RCol<N>& top, RCol<N>& bot, int k, int base) {
  constexpr int P = N / 2;
  if (P == 1) return;
  const int sx = base + (k + 1 < P ? k + 1 : k);
  const int sy = base + (k >= 1 ? k - 1 : k);
#pragma unroll
  for (int r = 0; r < N; ++r) {
    const double xr = __shfl(top.re[r], sx, 64), xi = __shfl(top.im[r], sx, 64);
    const double yr = __shfl(bot.re[r], sy, 64), yi = __shfl(bot.im[r], sy, 64);
    const double tr = k == 0 ? top.re[r] : (k < P - 1 ? xr : bot.re[r]);
    const double ti = k == 0 ? top.im[r] : (k < P - 1 ? xi : bot.im[r]);
    const double br = k == 0 ? xr : yr;
    const double bi = k == 0 ? xi : yi;
    top.re[r] = tr;
    top.im[r] = ti;
    bot.re[r] = br;
    bot.im[r] = bi;
  }
}

// One Jacobi rotation orthogonalising columns p and q; returns whether it rotated.
template <int N, bool WITHV>
__device__ __forceinline__ bool jac_rotate(RCol<N>& ap, RCol<N>& aq, RCol<N>& vp, RCol<N>& vq, double tol,
                                           double small2) {
  double a = 0.0, b = 0.0, gr = 0.0, gi = 0.0;
#pragma unroll
  for (int r = 0; r < N; ++r) {
    a = fma(ap.re[r], ap.re[r], fma(ap.im[r], ap.im[r], a));
    b = fma(aq.re[r], aq.re[r], fma(aq.im[r], aq.im[r], b));
    gr = fma(ap.re[r], aq.re[r], fma(ap.im[r], aq.im[r], gr));
    gi = fma(ap.re[r], aq.im[r], fma(-ap.im[r], aq.re[r], gi));
  }
  const double ga = sqrt(gr * gr + gi * gi);
  if (fmin(a, b) < small2 || !(ga > tol * sqrt(a * b))) return false;
  const double zeta = (b - a) / (2.0 * ga);
  const double t = (zeta >= 0.0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
  const double c = 1.0 / sqrt(1.0 + t * t);
  const double s = c * t;
  const double er = gr / ga, ei = -gi / ga;  // e^{-i phi} = conj(g) / |g|
#pragma unroll
  for (int r = 0; r < N; ++r) {
    const double xr = er * aq.re[r] - ei * aq.im[r];
    const double xi = er * aq.im[r] + ei * aq.re[r];
    const double pr = ap.re[r], pi = ap.im[r];
    ap.re[r] = c * pr - s * xr;
    ap.im[r] = c * pi - s * xi;
    aq.re[r] = s * pr + c * xr;
    aq.im[r] = s * pi + c * xi;
  }
  if (WITHV) {
#pragma unroll
    for (int r = 0; r < N; ++r) {
      const double xr = er * vq.re[r] - ei * vq.im[r];
      const double xi = er * vq.im[r] + ei * vq.re[r];
      const double pr = vp.re[r], pi = vp.im[r];
      vp.re[r] = c * pr - s * xr;
      vp.im[r] = c * pi - s * xi;
      vq.re[r] = s * pr + c * xr;
      vq.im[r] = s * pi + c * xi;
    }
  }
  return true;
}

template <int N>
__device__ __forceinline__ double group_max(double v) {
  constexpr int P2 = RegCanonShape<N>::P2;
#pragma unroll
  for (int o = P2 >> 1; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  return v;
}

template <int N>
__device__ __forceinline__ double group_sum(double v) {
  constexpr int P2 = RegCanonShape<N>::P2;
#pragma unroll
  for (int o = P2 >> 1; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

template <int N, bool WITHV>
__device__ void jacobi_cols(RCol<N>& top, RCol<N>& bot, RCol<N>& vt, RCol<N>& vb, int k, int base, bool active) {
  const double scale2 = group_max<N>(fmax(col_norm2(top), col_norm2(bot)));
  const double small2 = 1e-26 * scale2;
  const double tol = 2.2204460492503131e-16 * N;
  for (int sweep = 0; sweep < 30; ++sweep) {
    int rotated = 0;
#pragma unroll 1
    for (int r = 0; r < N - 1; ++r) {
      if (active) rotated |= jac_rotate<N, WITHV>(top, bot, vt, vb, tol, small2) ? 1 : 0;
      rotate_owners<N>(top, bot, k, base);
      if (WITHV) rotate_owners<N>(vt, vb, k, base);
    }
    if (!__any(rotated)) break;
  }
}

template <typename R, int N>
__global__ __launch_bounds__(64) void k_canon_reg(const cplx<R>* __restrict__ ext, int E, int ccols,
                                                  double pinv_eps, double deps, zd* __restrict__ canon,
                                                  double* __restrict__ lfull, double* __restrict__ ws) {
  using S = RegCanonShape<N>;
  constexpr int P = S::P, P2 = S::P2;
  const int lane = threadIdx.x;
  const int k = lane % P2;
  const int base = lane - k;
  const int edge = blockIdx.x * S::EPW + lane / P2;
  const bool valid = edge < E;
  const bool active = valid && k < P;
  double* wsE = ws + (size_t)(valid ? edge : 0) * S::WS;
  double* Uf = wsE;
  double* Ub = wsE + 2 * N * N;
  double* sqf = wsE + 4 * N * N;
  double* ipf = sqf + N;
  double* sqb = ipf + N;
  double* ipb = sqb + N;

  RCol<N> top, bot, vt, vb;
  for (int side = 0; side < 2; ++side) {
    col_zero(top);
    col_zero(bot);
    if (active) {
      const cplx<R>* M = ext + (size_t)(side ? E + edge : edge) * N * N;
      const int ct = k, cb = N - 1 - k;
#pragma unroll
      for (int r = 0; r < N; ++r) {
        const cplx<R> x = M[r * N + ct], y = M[r * N + cb];
        top.re[r] = (double)x.x;
        top.im[r] = (double)x.y;
        bot.re[r] = (double)y.x;
        bot.im[r] = (double)y.y;
      }
    }
    jacobi_cols<N, false>(top, bot, vt, vb, k, base, active);
    // singular values = column norms; U = A / s; masks of backends.py:709-727
    double* U = side ? Ub : Uf;
    double* sq = side ? sqb : sqf;
    double* ip = side ? ipb : ipf;
    if (active) {
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const RCol<N>& c = h ? bot : top;
        const int slot = 2 * k + h;
        const double s = sqrt(col_norm2(c));
        const double inv = s > 0.0 ? 1.0 / s : 0.0;
        const bool keep = s > pinv_eps;
        const double sqv = keep ? sqrt(s) : 0.0;
        sq[slot] = sqv;
        ip[slot] = sqv > deps ? 1.0 / sqv : 0.0;
        double* uc = U + slot * 2 * N;
#pragma unroll
        for (int r = 0; r < N; ++r) {
          uc[2 * r] = c.re[r] * inv;
          uc[2 * r + 1] = c.im[r] * inv;
        }
      }
    }
  }
  __syncthreads();

  // K[i][j] = sq_f[i] sq_b[j] conj(sum_r U_f[r, i] U_b[r, j]); this lane builds columns j = 2k, 2k+1.
  RCol<N> kt, kb;
  col_zero(kt);
  col_zero(kb);
  if (active) {
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int j = 2 * k + h;
      const double* ub = Ub + j * 2 * N;
      const double sj = sqb[j];
      RCol<N>& kc = h ? kb : kt;
#pragma unroll
      for (int i = 0; i < N; ++i) {
        const double* uf = Uf + i * 2 * N;
        double ar = 0.0, ai = 0.0;
#pragma unroll
        for (int r = 0; r < N; ++r) {
          const double fr = uf[2 * r], fi = uf[2 * r + 1], br = ub[2 * r], bi = ub[2 * r + 1];
          ar = fma(fr, br, fma(-fi, bi, ar));
          ai = fma(fr, bi, fma(fi, br, ai));
        }
        const double w = sqf[i] * sj;
        kc.re[i] = w * ar;
        kc.im[i] = -w * ai;
      }
    }
  }
  col_zero(vt);
  col_zero(vb);
#pragma unroll
  for (int r = 0; r < N; ++r) {
    if (r == 2 * k) vt.re[r] = 1.0;
    if (r == 2 * k + 1) vb.re[r] = 1.0;
  }
  jacobi_cols<N, true>(kt, kb, vt, vb, k, base, active);

  // singular values of K, descending ranks, masks and the normalised lambdas
  const double lt = sqrt(col_norm2(kt)), lb = sqrt(col_norm2(kb));
  int rt = 0, rb = 0;
#pragma unroll 1
  for (int q = 0; q < P; ++q) {
    const double ot = __shfl(lt, base + q, 64), ob = __shfl(lb, base + q, 64);
    const int st = 2 * q, sb = 2 * q + 1;
    rt += (ot > lt || (ot == lt && st < 2 * k)) ? 1 : 0;
    rt += (ob > lt || (ob == lt && sb < 2 * k)) ? 1 : 0;
    rb += (ot > lb || (ot == lb && st < 2 * k + 1)) ? 1 : 0;
    rb += (ob > lb || (ob == lb && sb < 2 * k + 1)) ? 1 : 0;
  }
  const bool keept = lt > pinv_eps, keepb = lb > pinv_eps;
  const double mt = keept ? lt : 0.0, mb = keepb ? lb : 0.0;
  const double nrm = sqrt(group_sum<N>(active ? mt * mt + mb * mb : 0.0));
  if (active) {
    lfull[(size_t)edge * N + rt] = mt / nrm;
    lfull[(size_t)edge * N + rb] = mb / nrm;
    zd* Cb = canon + (size_t)edge * N * ccols;
    zd* Cf = canon + (size_t)(E + edge) * N * ccols;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int rank = h ? rb : rt;
      if (rank >= ccols) continue;
      const bool keep = h ? keepb : keept;
      const RCol<N>& us = h ? kb : kt;
      const RCol<N>& vs = h ? vb : vt;
      const double il = keep ? 1.0 / (h ? lb : lt) : 0.0;
      // C_f[:, rank] = sum_r U_f[:, r] ip_f[r] Us[r]   (Us = K column / lambda)
      // C_b[:, rank] = sum_r U_b[:, r] ip_b[r] conj(Vs[r])
      double wfr[N], wfi[N], wbr[N], wbi[N];
#pragma unroll
      for (int r = 0; r < N; ++r) {
        wfr[r] = ipf[r] * us.re[r] * il;
        wfi[r] = ipf[r] * us.im[r] * il;
        wbr[r] = keep ? ipb[r] * vs.re[r] : 0.0;
        wbi[r] = keep ? -ipb[r] * vs.im[r] : 0.0;
      }
#pragma unroll 1
      for (int a = 0; a < N; ++a) {
        double cfr = 0.0, cfi = 0.0, cbr = 0.0, cbi = 0.0;
#pragma unroll
        for (int r = 0; r < N; ++r) {
          const double fr = Uf[r * 2 * N + 2 * a], fi = Uf[r * 2 * N + 2 * a + 1];
          cfr = fma(fr, wfr[r], fma(-fi, wfi[r], cfr));
          cfi = fma(fr, wfi[r], fma(fi, wfr[r], cfi));
          const double br = Ub[r * 2 * N + 2 * a], bi = Ub[r * 2 * N + 2 * a + 1];
          cbr = fma(br, wbr[r], fma(-bi, wbi[r], cbr));
          cbi = fma(br, wbi[r], fma(bi, wbr[r], cbi));
        }
        Cf[a * ccols + rank] = mk(cfr, cfi);
        Cb[a * ccols + rank] = mk(cbr, cbi);
      }
    }
  }
}

}  // namespace bqa
