// Canonicalisers with split columns for bond dimensions chi <= 8 (n = 2 chi <= 16).
//
// Same results as k_canonicalize / k_canon_reg, organised for occupancy and
// fewer Jacobi sweeps:
//   * 2 * P2 lanes per edge (P2 = n/2 rounded up to a power of two): lane
//     (k, h) holds rows [h n/2, (h+1) n/2) of the two columns of pair k, so a
//     column dot product is a lane-local sum plus one exchange with the other
//     half (lane ^ P2); columns move between pairs with DPP row shifts;
//   * the Hermitian PSD extended messages are factored M = L L^H by pivoted
//     Cholesky first; one-sided Jacobi on the columns of L then converges in
//     about half the sweeps of Jacobi on M and yields V_M Sigma_L directly,
//     with sigma_L^2 the eigenvalues (= singular values) of M;
//   * the bond matrix K runs Jacobi without accumulating V: with K = S_f X S_b,
//     X = conj(U_f^T U_b), the right singular vectors are
//     Vs = S_b U_b^T U_f S_f Us / Lambda for the kept columns.
#pragma once

#include "common.hpp"

namespace bqa {

template <int N>
struct HCanon {
  static constexpr int P = N / 2;                                   // column pairs
  static constexpr int P2 = P <= 1 ? 1 : P <= 2 ? 2 : P <= 4 ? 4 : 8;
  static constexpr int G = 2 * P2;                                  // lanes per edge
  static constexpr int EPW = 64 / G;                                // edges per wave
  static constexpr int NH = N / 2;                                  // rows per lane
  // workspace doubles per edge: U_f, U_b, Q (N x N complex each), sorted Us and Z columns
  // (N x N complex each), sq/ip of both sides, 1/Lambda of the sorted columns
  static constexpr int WS = 10 * N * N + 5 * N;
};

template <int NH>
struct HCol {
  double re[NH], im[NH];
};

__device__ __forceinline__ double dpp_from_next(double x) {  // value of lane + 1 (row_shl:1)
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(x), 0x101, 0xf, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(x), 0x101, 0xf, 0xf, false);
  return __hiloint2double(hi, lo);
}

__device__ __forceinline__ double dpp_from_prev(double x) {  // value of lane - 1 (row_shr:1)
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(x), 0x111, 0xf, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(x), 0x111, 0xf, 0xf, false);
  return __hiloint2double(hi, lo);
}

template <int N>
__device__ __forceinline__ double half_sum(double v) {
  return v + __shfl_xor(v, HCanon<N>::P2, 64);
}

template <int N>
__device__ __forceinline__ void hrotate_owners(HCol<N / 2>& top, HCol<N / 2>& bot, int k) {
  constexpr int P = N / 2, NH = N / 2;
  if (P == 1) return;
#pragma unroll
  for (int r = 0; r < NH; ++r) {
    const double xr = dpp_from_next(top.re[r]), xi = dpp_from_next(top.im[r]);
    const double yr = dpp_from_prev(bot.re[r]), yi = dpp_from_prev(bot.im[r]);
    const double tr = k == 0 ? top.re[r] : (k < P - 1 ? xr : bot.re[r]);
    const double ti = k == 0 ? top.im[r] : (k < P - 1 ? xi : bot.im[r]);
    bot.re[r] = k == 0 ? xr : yr;
    bot.im[r] = k == 0 ? xi : yi;
    top.re[r] = tr;
    top.im[r] = ti;
  }
}

template <int NH>
__device__ __forceinline__ double hnorm2(const HCol<NH>& c) {
  double s = 0.0;
#pragma unroll
  for (int r = 0; r < NH; ++r) s = fma(c.re[r], c.re[r], fma(c.im[r], c.im[r], s));
  return s;
}

// One-sided Jacobi on the pair of split columns held by this lane pair.
template <int N>
__device__ void hjacobi(HCol<N / 2>& top, HCol<N / 2>& bot, int k, bool active) {
  constexpr int NH = N / 2, P2 = HCanon<N>::P2;
  double sc = active ? fmax(half_sum<N>(hnorm2(top)), half_sum<N>(hnorm2(bot))) : 0.0;
#pragma unroll
  for (int o = 1; o < P2; o <<= 1) sc = fmax(sc, __shfl_xor(sc, o, 64));
  const double small2 = 1e-26 * sc;
  const double tol = 2.2204460492503131e-16 * N;
  // A sweep whose rotations all had column cosine |g| / sqrt(a b) < 1e-8 and
  // sine |s| < 1e-6 leaves every cosine below ~1e-14 (each later rotation of
  // the sweep changes a settled pair by at most the product of the two), so
  // the sweep that would only confirm convergence is skipped; near-degenerate
  // pairs, where tiny cosines give large rotations, keep the full test.
  constexpr double kSettledCos2 = 1e-16, kSettledSin2 = 1e-12;
  for (int sweep = 0; sweep < 30; ++sweep) {
    int rotated = 0, unsettled = 0;
#pragma unroll 1
    for (int rnd = 0; rnd < N - 1; ++rnd) {
      double a = 0.0, b = 0.0, gr = 0.0, gi = 0.0;
#pragma unroll
      for (int r = 0; r < NH; ++r) {
        a = fma(top.re[r], top.re[r], fma(top.im[r], top.im[r], a));
        b = fma(bot.re[r], bot.re[r], fma(bot.im[r], bot.im[r], b));
        gr = fma(top.re[r], bot.re[r], fma(top.im[r], bot.im[r], gr));
        gi = fma(top.re[r], bot.im[r], fma(-top.im[r], bot.re[r], gi));
      }
      a = half_sum<N>(a);
      b = half_sum<N>(b);
      gr = half_sum<N>(gr);
      gi = half_sum<N>(gi);
      const double g2 = gr * gr + gi * gi;
      if (active && fmin(a, b) >= small2 && g2 > tol * tol * a * b) {
        rotated = 1;
        // rotation parameters from reciprocal square roots (c^2 + s^2 = 1 to a few ulp)
        const double rga = rsqrt(g2);
        const double zeta = 0.5 * (b - a) * rga;
        const double w = fabs(zeta) + sqrt(fma(zeta, zeta, 1.0));
        double rw = __builtin_amdgcn_rcp(w);
        rw = fma(fma(-w, rw, 1.0), rw, rw);
        rw = fma(fma(-w, rw, 1.0), rw, rw);
        const double t = zeta >= 0.0 ? rw : -rw;
        const double c = rsqrt(fma(t, t, 1.0));
        const double s = c * t;
        const double er = gr * rga, ei = -gi * rga;
        unsettled |= (g2 > kSettledCos2 * a * b || s * s > kSettledSin2) ? 1 : 0;
#pragma unroll
        for (int r = 0; r < NH; ++r) {
          const double xr = er * bot.re[r] - ei * bot.im[r];
          const double xi = er * bot.im[r] + ei * bot.re[r];
          const double pr = top.re[r], pi = top.im[r];
          top.re[r] = c * pr - s * xr;
          top.im[r] = c * pi - s * xi;
          bot.re[r] = s * pr + c * xr;
          bot.im[r] = s * pi + c * xi;
        }
      }
      hrotate_owners<N>(top, bot, k);
    }
    if (!__any(rotated) || !__any(unsettled)) break;
  }
}

// Element `idx` (0..N-1) of a split column: the half holding it contributes, the other adds 0.
template <int N>
__device__ __forceinline__ void helem(const HCol<N / 2>& c, int idx, int h, double& re, double& im) {
  constexpr int NH = N / 2;
  double xr = 0.0, xi = 0.0;
  const int loc = idx - h * NH;
#pragma unroll
  for (int r = 0; r < NH; ++r) {
    xr = r == loc ? c.re[r] : xr;
    xi = r == loc ? c.im[r] : xi;
  }
  re = half_sum<N>(xr);
  im = half_sum<N>(xi);
}

// Pivoted Cholesky M = L L^H in place: each column ends up holding one column of
// L (in pivot order) or zero beyond the numerical rank.
template <int N>
__device__ void hcholesky(HCol<N / 2>& top, HCol<N / 2>& bot, int k, int h, int base, bool active) {
  constexpr int NH = N / 2, P2 = HCanon<N>::P2;
  const int ct = k, cb = N - 1 - k;  // matrix columns held (fixed during the factorisation)
  bool dt = !active, db = !active;   // done flags
  double d0 = 0.0;
  for (int step = 0; step < N; ++step) {
    double tr, ti, br, bi;
    helem<N>(top, ct, h, tr, ti);
    helem<N>(bot, cb, h, br, bi);
    // best undone diagonal over the group: value and the matrix column holding it
    double bv = -1.0;
    int bc = -1;
    if (!dt && tr > bv) { bv = tr; bc = ct; }
    if (!db && br > bv) { bv = br; bc = cb; }
#pragma unroll
    for (int o = 1; o < P2; o <<= 1) {
      const double ov = __shfl_xor(bv, o, 64);
      const int oc = __shfl_xor(bc, o, 64);
      if (ov > bv || (ov == bv && oc >= 0 && (bc < 0 || oc < bc))) { bv = ov; bc = oc; }
    }
    if (step == 0) d0 = bv;
    if (!(bv > 1e-20 * d0) || bc < 0) {  // numerical rank reached: zero the rest
      if (!dt) {
#pragma unroll
        for (int r = 0; r < NH; ++r) { top.re[r] = 0.0; top.im[r] = 0.0; }
      }
      if (!db) {
#pragma unroll
        for (int r = 0; r < NH; ++r) { bot.re[r] = 0.0; bot.im[r] = 0.0; }
      }
      break;
    }
    // pivot column, scaled: l = R[:, bc] / sqrt(R[bc][bc]); owner lane pair of column bc
    const int owner_k = bc < N / 2 ? bc : N - 1 - bc;
    const bool own_top = bc == owner_k;
    const double inv = 1.0 / sqrt(bv);
    const int src = base + owner_k + h * P2;
    HCol<NH> l;
#pragma unroll
    for (int r = 0; r < NH; ++r) {
      const double atr = __shfl(top.re[r], src, 64), ati = __shfl(top.im[r], src, 64);
      const double abr = __shfl(bot.re[r], src, 64), abi = __shfl(bot.im[r], src, 64);
      l.re[r] = (own_top ? atr : abr) * inv;
      l.im[r] = (own_top ? ati : abi) * inv;
    }
    // conj(l[j]) for this lane's two columns j
    double ltr, lti, lbr, lbi;
    helem<N>(l, ct, h, ltr, lti);
    helem<N>(l, cb, h, lbr, lbi);
    const bool pt = active && bc == ct, pb = active && bc == cb;
    if (!dt && !pt) {
#pragma unroll
      for (int r = 0; r < NH; ++r) {  // R[:, ct] -= l conj(l[ct])
        top.re[r] -= l.re[r] * ltr + l.im[r] * lti;
        top.im[r] -= l.im[r] * ltr - l.re[r] * lti;
      }
    }
    if (!db && !pb) {
#pragma unroll
      for (int r = 0; r < NH; ++r) {
        bot.re[r] -= l.re[r] * lbr + l.im[r] * lbi;
        bot.im[r] -= l.im[r] * lbr - l.re[r] * lbi;
      }
    }
    if (pt) {
#pragma unroll
      for (int r = 0; r < NH; ++r) { top.re[r] = l.re[r]; top.im[r] = l.im[r]; }
      dt = true;
    }
    if (pb) {
#pragma unroll
      for (int r = 0; r < NH; ++r) { bot.re[r] = l.re[r]; bot.im[r] = l.im[r]; }
      db = true;
    }
  }
}

template <int N>
__global__ __launch_bounds__(64) void k_canon_h(const cplx<double>* __restrict__ ext, int E, int ccols,
                                                double pinv_eps, double deps, zd* __restrict__ canon,
                                                double* __restrict__ lfull, double* __restrict__ ws) {
  using S = HCanon<N>;
  constexpr int P = S::P, P2 = S::P2, G = S::G, NH = S::NH;
  const int lane = threadIdx.x;
  const int w = lane % G;
  const int base = lane - w;
  const int h = w / P2, k = w % P2;
  const int edge = blockIdx.x * S::EPW + lane / G;
  const bool valid = edge < E;
  const bool active = valid && k < P;
  double* wsE = ws + (size_t)(valid ? edge : 0) * S::WS;
  double* Uf = wsE;
  double* Ub = wsE + 2 * N * N;
  double* sqf = wsE + 4 * N * N;
  double* ipf = sqf + N;
  double* sqb = ipf + N;
  double* ipb = sqb + N;
  const int r0 = h * NH;

  HCol<NH> top, bot;
  for (int side = 0; side < 2; ++side) {
    const int ct = k, cb = N - 1 - k;
#pragma unroll
    for (int r = 0; r < NH; ++r) {
      top.re[r] = top.im[r] = bot.re[r] = bot.im[r] = 0.0;
    }
    if (active) {
      // Hermitian part (M + M^H) / 2: the extended message is Hermitian by
      // construction, and its anti-Hermitian rounding part (from FP32 messages)
      // is dropped rather than folded into the factorisation.
      const cplx<double>* M = ext + (size_t)(side ? E + edge : edge) * N * N;
#pragma unroll
      for (int r = 0; r < NH; ++r) {
        const cplx<double> x = M[(r0 + r) * N + ct], y = M[(r0 + r) * N + cb];
        const cplx<double> xt = M[ct * N + r0 + r], yt = M[cb * N + r0 + r];
        top.re[r] = 0.5 * (x.x + xt.x);
        top.im[r] = 0.5 * (x.y - xt.y);
        bot.re[r] = 0.5 * (y.x + yt.x);
        bot.im[r] = 0.5 * (y.y - yt.y);
      }
    }
    hcholesky<N>(top, bot, k, h, base, active);
    hjacobi<N>(top, bot, k, active);
    // columns are V_M sigma: eigenvalue s = sigma^2, U = column / sigma
    double* U = side ? Ub : Uf;
    double* sq = side ? sqb : sqf;
    double* ip = side ? ipb : ipf;
#pragma unroll
    for (int hh = 0; hh < 2; ++hh) {
      const HCol<NH>& c = hh ? bot : top;
      const double sig = sqrt(half_sum<N>(hnorm2(c)));
      if (active) {
        const int slot = 2 * k + hh;
        const double s = sig * sig;
        const bool keep = s > pinv_eps;
        const double sqv = keep ? sig : 0.0;
        if (h == 0) {
          sq[slot] = sqv;
          ip[slot] = sqv > deps ? 1.0 / sqv : 0.0;
        }
        const double inv = sig > 0.0 ? 1.0 / sig : 0.0;
        double* uc = U + slot * 2 * N + 2 * r0;
#pragma unroll
        for (int r = 0; r < NH; ++r) {
          uc[2 * r] = c.re[r] * inv;
          uc[2 * r + 1] = c.im[r] * inv;
        }
      }
    }
  }
  __syncthreads();

  // K[i][j] = sq_f[i] sq_b[j] conj(sum_t U_f[t, i] U_b[t, j]); this lane: columns 2k, 2k+1, rows of half h
#pragma unroll
  for (int hh = 0; hh < 2; ++hh) {
    HCol<NH>& kc = hh ? bot : top;
    const int j = 2 * k + hh;
#pragma unroll
    for (int r = 0; r < NH; ++r) {
      kc.re[r] = 0.0;
      kc.im[r] = 0.0;
    }
    if (active) {
      // t outermost with the NH row accumulators inside: each sum keeps its
      // order while the workspace loads stream through a small register window
      const double* ub = Ub + j * 2 * N;
      double ar[NH], ai[NH];
#pragma unroll
      for (int r = 0; r < NH; ++r) ar[r] = ai[r] = 0.0;
#pragma unroll 2
      for (int t = 0; t < N; ++t) {
        const double br = ub[2 * t], bi = ub[2 * t + 1];
#pragma unroll
        for (int r = 0; r < NH; ++r) {
          const double fr = Uf[(r0 + r) * 2 * N + 2 * t], fi = Uf[(r0 + r) * 2 * N + 2 * t + 1];
          ar[r] = fma(fr, br, fma(-fi, bi, ar[r]));
          ai[r] = fma(fr, bi, fma(fi, br, ai[r]));
        }
      }
      const double sj = sqb[j];
#pragma unroll
      for (int r = 0; r < NH; ++r) {
        const double wgt = sqf[r0 + r] * sj;
        kc.re[r] = wgt * ar[r];
        kc.im[r] = -wgt * ai[r];
      }
    }
  }
  hjacobi<N>(top, bot, k, active);

  // singular values, descending ranks, masks, normalised lambdas
  const double lt = sqrt(half_sum<N>(hnorm2(top))), lb = sqrt(half_sum<N>(hnorm2(bot)));
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
  double nn = (active && h == 0) ? mt * mt + mb * mb : 0.0;
#pragma unroll
  for (int o = 1; o < G; o <<= 1) nn += __shfl_xor(nn, o, 64);
  const double nrm = sqrt(nn);
  if (active && h == 0) {
    lfull[(size_t)edge * N + rt] = mt / nrm;
    lfull[(size_t)edge * N + rb] = mb / nrm;
  }
  // Stage the sorted, masked Us columns (Us[:, c] = K column / Lambda) and 1/Lambda.
  double* Q = wsE + 4 * N * N + 4 * N;         // Q[t][i] = sum_u U_b[u][t] U_f[u][i] sq_f[i]
  double* Us = Q + 2 * N * N;                  // Us[c][row], c = rank
  double* Zc = Us + 2 * N * N;                 // Z[c][j] = keep_b[j] conj(sum_i Q[j][i] Us[c][i]) / Lambda_c
  double* il = Zc + 2 * N * N;                 // 1 / Lambda_c (0 when masked)
  if (active) {
#pragma unroll
    for (int hh = 0; hh < 2; ++hh) {
      const HCol<NH>& c = hh ? bot : top;
      const int rank = hh ? rb : rt;
      const double inv = (hh ? keepb : keept) ? 1.0 / (hh ? lb : lt) : 0.0;
      if (h == 0) il[rank] = inv;
      double* uc = Us + rank * 2 * N + 2 * r0;
#pragma unroll
      for (int r = 0; r < NH; ++r) {
        uc[2 * r] = c.re[r] * inv;
        uc[2 * r + 1] = c.im[r] * inv;
      }
    }
  }
  // Q, one entry per lane per step over the group
  if (valid) {
    for (int x = w; x < N * N; x += G) {
      const int t = x / N, i = x % N;
      double ar = 0.0, ai = 0.0;
      for (int u = 0; u < N; ++u) {
        const double br = Ub[t * 2 * N + 2 * u], bi = Ub[t * 2 * N + 2 * u + 1];
        const double fr = Uf[i * 2 * N + 2 * u], fi = Uf[i * 2 * N + 2 * u + 1];
        ar = fma(br, fr, fma(-bi, fi, ar));
        ai = fma(br, fi, fma(bi, fr, ai));
      }
      Q[2 * x] = ar * sqf[i];
      Q[2 * x + 1] = ai * sqf[i];
    }
  }
  __syncthreads();
  if (valid) {
    for (int x = w; x < ccols * N; x += G) {
      const int c = x / N, jj = x % N;
      double ar = 0.0, ai = 0.0;
      if (ipb[jj] > 0.0) {
        for (int i = 0; i < N; ++i) {
          const double qr = Q[(jj * N + i) * 2], qi = Q[(jj * N + i) * 2 + 1];
          const double ur = Us[c * 2 * N + 2 * i], ui = Us[c * 2 * N + 2 * i + 1];
          ar = fma(qr, ur, fma(-qi, ui, ar));
          ai = fma(qr, ui, fma(qi, ur, ai));
        }
      }
      Zc[2 * x] = ar * il[c];
      Zc[2 * x + 1] = -ai * il[c];
    }
  }
  __syncthreads();
  // C_f[a][c] = sum_t U_f[a][t] ip_f[t] Us[c][t];  C_b[a][c] = sum_j U_b[a][j] Z[c][j]
  if (valid) {
    zd* Cb = canon + (size_t)edge * N * ccols;
    zd* Cf = canon + (size_t)(E + edge) * N * ccols;
    for (int x = w; x < N * ccols; x += G) {
      const int a = x / ccols, c = x % ccols;
      double fr_ = 0.0, fi_ = 0.0, br_ = 0.0, bi_ = 0.0;
      for (int t = 0; t < N; ++t) {
        const double ur = Uf[t * 2 * N + 2 * a], ui = Uf[t * 2 * N + 2 * a + 1];
        const double wr = ipf[t] * Us[c * 2 * N + 2 * t], wi = ipf[t] * Us[c * 2 * N + 2 * t + 1];
        fr_ = fma(ur, wr, fma(-ui, wi, fr_));
        fi_ = fma(ur, wi, fma(ui, wr, fi_));
        const double vr = Ub[t * 2 * N + 2 * a], vi = Ub[t * 2 * N + 2 * a + 1];
        const double zr = Zc[(c * N + t) * 2], zi = Zc[(c * N + t) * 2 + 1];
        br_ = fma(vr, zr, fma(-vi, zi, br_));
        bi_ = fma(vr, zi, fma(vi, zr, bi_));
      }
      Cf[x] = mk(fr_, fi_);
      Cb[x] = mk(br_, bi_);
    }
  }
}

}  // namespace bqa
