// Canonicalisers with split columns for bond dimensions chi <= 16 (n = 2 chi <= 32).
//
// Same results as k_canonicalize, organised for occupancy and
// fewer Jacobi sweeps:
//   * 2 * P2 lanes per edge (P2 = n/2 rounded up to a power of two): lane
//     (k, h) = k + P2 h holds rows [h n/2, (h+1) n/2) of the two columns of
//     pair k, so a column dot product is a lane-local sum plus one DPP
//     exchange with the other half (lane ^ P2);
//   * when n/2 is a power of two the Jacobi sweep pairs columns in a recursive
//     xor order (all top x bot pairs with only the bot columns moving by one
//     lane-xor per round, then each half split in two by one re-layout, and
//     so on: n - 1 rounds), so a round moves one column per lane, through the
//     LDS crossbar (ds_bpermute) rather than VALU moves and selects; other n
//     use a round-robin with DPP row shifts;
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
  static constexpr int P2 = P <= 1 ? 1 : P <= 2 ? 2 : P <= 4 ? 4 : P <= 8 ? 8 : 16;
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

// 32-bit DPP move within 16-lane rows (every source lane valid for the
// patterns used here).
template <int CTRL>
__device__ __forceinline__ int dpp32(int v) {
  return __builtin_amdgcn_mov_dpp(v, CTRL, 0xf, 0xf, true);
}

// Value of lane ^ M (M = 1, 2, 4, 8) within 16-lane rows: quad permutations
// for 1 and 2, half-row mirror (lane ^ 7) then lane ^ 3 for 4, rotation by 8.
template <int M>
__device__ __forceinline__ int xor_lane32(int v) {
  static_assert(M == 1 || M == 2 || M == 4 || M == 8, "lane xor within a row");
  if constexpr (M == 1) return dpp32<0xB1>(v);
  else if constexpr (M == 2) return dpp32<0x4E>(v);
  else if constexpr (M == 4) return dpp32<0x1B>(dpp32<0x141>(v));
  else return dpp32<0x128>(v);
}

template <int M>
__device__ __forceinline__ double xor_lane(double x) {
  const int lo = xor_lane32<M>(__double2loint(x)), hi = xor_lane32<M>(__double2hiint(x));
  return __hiloint2double(hi, lo);
}

template <int N>
__device__ __forceinline__ double half_sum(double v) {
  if constexpr (HCanon<N>::P2 == 16)
    return v + __shfl_xor(v, 16, 64);     // the partner half is in the other 16-lane row
  else
    return v + xor_lane<HCanon<N>::P2>(v);
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

// One Jacobi round: rotate the pair of split columns held by this lane pair
// (top, bot) so that they become orthogonal; `rotated` / `unsettled` collect
// whether the round rotated at all and whether a rotation was still large.
// With TRACK the squared column norms a, b are carried (exact at each sweep's
// start, updated by the rotation's a - t|g|, b + t|g|, moved with their
// columns), otherwise they are recomputed here.
template <int N, bool TRACK>
__device__ __forceinline__ void hjacobi_round(HCol<N / 2>& top, HCol<N / 2>& bot, double& a, double& b, bool active,
                                              double small2, double tol, int& rotated, int& unsettled) {
  constexpr int NH = N / 2;
  // A sweep whose rotations all had column cosine |g| / sqrt(a b) < 1e-8 and
  // sine |s| < 1e-6 leaves every cosine below ~1e-14 (each later rotation of
  // the sweep changes a settled pair by at most the product of the two), so
  // the sweep that would only confirm convergence is skipped; near-degenerate
  // pairs, where tiny cosines give large rotations, keep the full test.
  constexpr double kSettledCos2 = 1e-16, kSettledSin2 = 1e-12;
  double gr = 0.0, gi = 0.0;
  if constexpr (!TRACK) {
    a = hnorm2(top);
    b = hnorm2(bot);
  }
#pragma unroll
  for (int r = 0; r < NH; ++r) {
    gr = fma(top.re[r], bot.re[r], fma(top.im[r], bot.im[r], gr));
    gi = fma(top.re[r], bot.im[r], fma(-top.im[r], bot.re[r], gi));
  }
  if constexpr (!TRACK) {
    a = half_sum<N>(a);
    b = half_sum<N>(b);
  }
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
    if constexpr (TRACK) {
      const double tg = t * (g2 * rga);  // t |g|
      a -= tg;
      b += tg;
    }
#pragma unroll
    for (int r = 0; r < NH; ++r) {
      const double xr = er * bot.re[r] - ei * bot.im[r];
      const double xi = er * bot.im[r] + ei * bot.re[r];
      const double pr = top.re[r], pi = top.im[r];
      bot.re[r] = s * pr + c * xr;
      bot.im[r] = s * pi + c * xi;
      top.re[r] = c * pr - s * xr;
      top.im[r] = c * pi - s * xi;
    }
  }
}

// Value of lane `src` (a wave lane, passed as 4 * src) through the LDS
// crossbar: no VALU work beyond the address, for the column moves.
__device__ __forceinline__ double bperm(double x, int addr4) {
  const int lo = __builtin_amdgcn_ds_bpermute(addr4, __double2loint(x));
  const int hi = __builtin_amdgcn_ds_bpermute(addr4, __double2hiint(x));
  return __hiloint2double(hi, lo);
}

// One-sided Jacobi on the columns held by the lane pairs of an edge.
template <int N>
__device__ void hjacobi(HCol<N / 2>& top, HCol<N / 2>& bot, int k, bool active) {
  constexpr int NH = N / 2, P = N / 2, P2 = HCanon<N>::P2;
  double sc = active ? fmax(half_sum<N>(hnorm2(top)), half_sum<N>(hnorm2(bot))) : 0.0;
#pragma unroll
  for (int o = 1; o < P2; o <<= 1) sc = fmax(sc, __shfl_xor(sc, o, 64));
  const double small2 = 1e-26 * sc;
  const double tol = 2.2204460492503131e-16 * N;
  for (int sweep = 0; sweep < 30; ++sweep) {
    int rotated = 0, unsettled = 0;
    if constexpr (P == P2) {
      // recursive xor order: in the phase of group size L the tops and the
      // bots of each group of L lanes meet all pairs (bot at lane k holds the
      // bot that started at lane k ^ gray(t) in round t, so round t moves it by
      // the lowest set bit of t); the re-layout at bit L / 2 then leaves the
      // pairs within each half for size L / 2: lanes with the bit clear take
      // (own top, partner's top), lanes with it set (own bot, partner's bot)
      const int lane = threadIdx.x & 63;
      double a = half_sum<N>(hnorm2(top)), b = half_sum<N>(hnorm2(bot));
#pragma unroll 1
      for (int L = P; L >= 1; L >>= 1) {
#pragma unroll 1
        for (int t = 0; t < L; ++t) {
          if (t > 0) {
            const int addr = 4 * (lane ^ (t & -t));
#pragma unroll
            for (int r = 0; r < NH; ++r) {
              bot.re[r] = bperm(bot.re[r], addr);
              bot.im[r] = bperm(bot.im[r], addr);
            }
            b = bperm(b, addr);
          }
          hjacobi_round<N, true>(top, bot, a, b, active, small2, tol, rotated, unsettled);
        }
        if (L > 1) {
          const int m = L >> 1, addr = 4 * (lane ^ m);
          const bool hi = (k & m) != 0;
#pragma unroll
          for (int r = 0; r < NH; ++r) {
            const double fr = bperm(hi ? top.re[r] : bot.re[r], addr);
            const double fi = bperm(hi ? top.im[r] : bot.im[r], addr);
            top.re[r] = hi ? bot.re[r] : top.re[r];
            top.im[r] = hi ? bot.im[r] : top.im[r];
            bot.re[r] = fr;
            bot.im[r] = fi;
          }
          const double fn = bperm(hi ? a : b, addr);
          a = hi ? b : a;
          b = fn;
        }
      }
    } else {
#pragma unroll 1
      for (int rnd = 0; rnd < N - 1; ++rnd) {
        double a, b;
        hjacobi_round<N, false>(top, bot, a, b, active, small2, tol, rotated, unsettled);
        hrotate_owners<N>(top, bot, k);
      }
    }
    if (!__any(rotated) || !__any(unsettled)) break;
  }
}

// Orders this wave's LDS accesses around a cross-lane exchange (the block is
// one wave, whose LDS operations complete in issue order).
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// One butterfly step of the pivot search: the larger diagonal of this lane and lane ^ M, ties to
// the lower column (-1: none).
template <int M>
__device__ __forceinline__ void pivot_step(double& bv, int& bc) {
  const double ov = xor_lane<M>(bv);
  const int oc = xor_lane32<M>(bc);
  if (ov > bv || (ov == bv && oc >= 0 && (bc < 0 || oc < bc))) {
    bv = ov;
    bc = oc;
  }
}

// Pivoted Cholesky M = L L^H in place: each column ends up holding one column of
// L (in pivot order) or zero beyond the numerical rank.  The scaled pivot
// column goes through the edge's LDS slot `lcol` (N complex).  dgt / dgb are the
// residual diagonals of the lane's columns ct = k and cb = N - 1 - k, updated
// with the arithmetic the column update applies to its diagonal element.
template <int N>
__device__ void hcholesky(HCol<N / 2>& top, HCol<N / 2>& bot, int k, int h, bool active, double dgt, double dgb,
                          double2* lcol) {
  constexpr int NH = N / 2, P2 = HCanon<N>::P2;
  const int ct = k, cb = N - 1 - k, r0 = h * NH;
  bool dt = !active, db = !active;  // done flags
  double d0 = 0.0;
  for (int step = 0; step < N; ++step) {
    // best undone diagonal over the group: value and the matrix column holding it
    double bv = -1.0;
    int bc = -1;
    if (!dt && dgt > bv) { bv = dgt; bc = ct; }
    if (!db && dgb > bv) { bv = dgb; bc = cb; }
    // butterfly over the column pairs by DPP lane xors (a step's latency is a few cycles, where a
    // shuffle through the LDS crossbar waits for its round trip); the same order as a shuffle loop
    if constexpr (P2 > 1) pivot_step<1>(bv, bc);
    if constexpr (P2 > 2) pivot_step<2>(bv, bc);
    if constexpr (P2 > 4) pivot_step<4>(bv, bc);
    if constexpr (P2 > 8) pivot_step<8>(bv, bc);
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
    // pivot column, scaled: l = R[:, bc] / sqrt(R[bc][bc]), written by the lane
    // pair holding column bc
    const int owner_k = bc < N / 2 ? bc : N - 1 - bc;
    const bool own_top = bc == owner_k;
    const double inv = 1.0 / sqrt(bv);
    if (k == owner_k) {
#pragma unroll
      for (int r = 0; r < NH; ++r)
        lcol[r0 + r] = make_double2((own_top ? top.re[r] : bot.re[r]) * inv, (own_top ? top.im[r] : bot.im[r]) * inv);
    }
    wave_lds_sync();
    HCol<NH> l;
#pragma unroll
    for (int r = 0; r < NH; ++r) {
      const double2 v = lcol[r0 + r];
      l.re[r] = v.x;
      l.im[r] = v.y;
    }
    // conj(l[j]) for this lane's two columns j
    const double2 lt = lcol[ct], lb = lcol[cb];
    wave_lds_sync();
    // one in-place form for every column: R[:, j] <- a R[:, j] - l conj(b l[j]),
    // with (a, b) = (1, 1) to update, (1 / sqrt(R[bc][bc]), 0) for the pivot
    // column (which becomes l) and (1, 0) once done
    const bool pt = active && bc == ct, pb = active && bc == cb;
    const double at = pt ? inv : 1.0, ab = pb ? inv : 1.0;
    const double ut = (!dt && !pt) ? 1.0 : 0.0, ub = (!db && !pb) ? 1.0 : 0.0;
    const double ltr = ut * lt.x, lti = ut * lt.y, lbr = ub * lb.x, lbi = ub * lb.y;
#pragma unroll
    for (int r = 0; r < NH; ++r) {
      const double tr = fma(l.re[r], ltr, l.im[r] * lti), ti = fma(l.im[r], ltr, -(l.re[r] * lti));
      const double br = fma(l.re[r], lbr, l.im[r] * lbi), bi = fma(l.im[r], lbr, -(l.re[r] * lbi));
      top.re[r] = fma(at, top.re[r], -tr);
      top.im[r] = fma(at, top.im[r], -ti);
      bot.re[r] = fma(ab, bot.re[r], -br);
      bot.im[r] = fma(ab, bot.im[r], -bi);
    }
    // the diagonal element's update: (l[j] conj(l[j])) in the form above
    dgt = fma(1.0, dgt, -fma(ltr, ltr, lti * lti));
    dgb = fma(1.0, dgb, -fma(lbr, lbr, lbi * lbi));
    dt = dt || pt;
    db = db || pb;
  }
}

// The edges of block `blk` of one job's launch: blk * EPW .. blk * EPW + EPW - 1.
template <int N>
__device__ __forceinline__ void canon_h_run(const cplx<double>* __restrict__ ext, int E, int ccols, double pinv_eps,
                                            double deps, zd* __restrict__ canon, double* __restrict__ lfull,
                                            double* __restrict__ ws, int blk) {
  using S = HCanon<N>;
  constexpr int P = S::P, P2 = S::P2, G = S::G, NH = S::NH;
  const int lane = threadIdx.x;
  const int w = lane % G;
  const int base = lane - w;
  const int h = w / P2, k = w % P2;
  const int edge = blk * S::EPW + lane / G;
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

  __shared__ double2 lslots[S::EPW * N];
  double2* lcol = lslots + (lane / G) * N;
  HCol<NH> top, bot;
  for (int side = 0; side < 2; ++side) {
    const int ct = k, cb = N - 1 - k;
    double dgt = 0.0, dgb = 0.0;
#pragma unroll
    for (int r = 0; r < NH; ++r) {
      top.re[r] = top.im[r] = bot.re[r] = bot.im[r] = 0.0;
    }
    if (active) {
      // Hermitian part (M + M^H) / 2: the extended message is Hermitian by
      // construction, and its anti-Hermitian rounding part (from FP32 messages)
      // is dropped rather than folded into the factorisation.
      // (columns and rows from four base pointers with constant offsets, one
      // column at a time, so few loads and no per-row addresses stay live)
      const cplx<double>* M = ext + (size_t)(side ? E + edge : edge) * N * N;
      const cplx<double>* Mt = M + r0 * N + ct;  // column ct from row r0
      const cplx<double>* Rt = M + ct * N + r0;  // row ct from column r0
      const cplx<double>* Mb = M + r0 * N + cb;
      const cplx<double>* Rb = M + cb * N + r0;
#pragma unroll
      for (int r = 0; r < NH; ++r) {
        const cplx<double> x = Mt[r * N], xt = Rt[r];
        top.re[r] = 0.5 * (x.x + xt.x);
        top.im[r] = 0.5 * (x.y - xt.y);
      }
      dgt = M[ct * N + ct].x;
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int r = 0; r < NH; ++r) {
        const cplx<double> y = Mb[r * N], yt = Rb[r];
        bot.re[r] = 0.5 * (y.x + yt.x);
        bot.im[r] = 0.5 * (y.y - yt.y);
      }
      dgb = M[cb * N + cb].x;
    }
    hcholesky<N>(top, bot, k, h, active, dgt, dgb, lcol);
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
  if constexpr (N == 16) {
    if (ccols <= 8) {
      // chi = 8: the same products regrouped without the 16 x 16 Q,
      //   V = U_f diag(sq_f) Us and C_f = U_f diag(ip_f) Us in one pass,
      //   Z[c][j] = keep_b[j] conj((U_b^T V)[j][c]) / Lambda_c,  C_b = U_b Z,
      // each lane taking a block of 4 rows x 2 columns so every row and column
      // it loads serves several outputs (about a quarter of the workspace
      // reads of the entry-per-lane form below).
      __syncthreads();
      const int a0 = 4 * (w >> 2), c0 = 2 * (w & 3);
      double* V = Q;  // V[c][a], 16 x ccols
      double xr[4][2] = {}, xi[4][2] = {}, yr[4][2] = {}, yi[4][2] = {};
      if (valid) {
#pragma unroll 2
        for (int t = 0; t < N; ++t) {
          double sr[2], si[2], pr[2], pi[2];
#pragma unroll
          for (int cc = 0; cc < 2; ++cc) {
            const int c = c0 + cc;
            const double ur = c < ccols ? Us[c * 2 * N + 2 * t] : 0.0, ui = c < ccols ? Us[c * 2 * N + 2 * t + 1] : 0.0;
            sr[cc] = sqf[t] * ur;
            si[cc] = sqf[t] * ui;
            pr[cc] = ipf[t] * ur;
            pi[cc] = ipf[t] * ui;
          }
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const double fr = Uf[t * 2 * N + 2 * (a0 + r)], fi = Uf[t * 2 * N + 2 * (a0 + r) + 1];
#pragma unroll
            for (int cc = 0; cc < 2; ++cc) {
              xr[r][cc] = fma(fr, sr[cc], fma(-fi, si[cc], xr[r][cc]));
              xi[r][cc] = fma(fr, si[cc], fma(fi, sr[cc], xi[r][cc]));
              yr[r][cc] = fma(fr, pr[cc], fma(-fi, pi[cc], yr[r][cc]));
              yi[r][cc] = fma(fr, pi[cc], fma(fi, pr[cc], yi[r][cc]));
            }
          }
        }
        zd* Cf = canon + (size_t)(E + edge) * N * ccols;
#pragma unroll
        for (int cc = 0; cc < 2; ++cc) {
          const int c = c0 + cc;
          if (c < ccols) {
#pragma unroll
            for (int r = 0; r < 4; ++r) {
              V[c * 2 * N + 2 * (a0 + r)] = xr[r][cc];
              V[c * 2 * N + 2 * (a0 + r) + 1] = xi[r][cc];
              Cf[(a0 + r) * ccols + c] = mk(yr[r][cc], yi[r][cc]);
            }
          }
        }
      }
      __syncthreads();
      // Z rows j = a0..a0+3, columns c0, c0+1
      double wr[4][2] = {}, wi[4][2] = {};
      if (valid) {
#pragma unroll 2
        for (int u = 0; u < N; ++u) {
          double vr[2], vi[2];
#pragma unroll
          for (int cc = 0; cc < 2; ++cc) {
            const int c = c0 + cc;
            vr[cc] = c < ccols ? V[c * 2 * N + 2 * u] : 0.0;
            vi[cc] = c < ccols ? V[c * 2 * N + 2 * u + 1] : 0.0;
          }
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const double br = Ub[(a0 + r) * 2 * N + 2 * u], bi = Ub[(a0 + r) * 2 * N + 2 * u + 1];
#pragma unroll
            for (int cc = 0; cc < 2; ++cc) {
              wr[r][cc] = fma(br, vr[cc], fma(-bi, vi[cc], wr[r][cc]));
              wi[r][cc] = fma(br, vi[cc], fma(bi, vr[cc], wi[r][cc]));
            }
          }
        }
#pragma unroll
        for (int cc = 0; cc < 2; ++cc) {
          const int c = c0 + cc;
          if (c < ccols) {
#pragma unroll
            for (int r = 0; r < 4; ++r) {
              const int j = a0 + r;
              const double g = ipb[j] > 0.0 ? il[c] : 0.0;
              Zc[(c * N + j) * 2] = wr[r][cc] * g;
              Zc[(c * N + j) * 2 + 1] = -wi[r][cc] * g;
            }
          }
        }
      }
      __syncthreads();
      // C_b rows a0..a0+3, columns c0, c0+1
      if (valid) {
        double br_[4][2] = {}, bi_[4][2] = {};
#pragma unroll 2
        for (int t = 0; t < N; ++t) {
          double zr[2], zi[2];
#pragma unroll
          for (int cc = 0; cc < 2; ++cc) {
            const int c = c0 + cc;
            zr[cc] = c < ccols ? Zc[(c * N + t) * 2] : 0.0;
            zi[cc] = c < ccols ? Zc[(c * N + t) * 2 + 1] : 0.0;
          }
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const double vr = Ub[t * 2 * N + 2 * (a0 + r)], vi = Ub[t * 2 * N + 2 * (a0 + r) + 1];
#pragma unroll
            for (int cc = 0; cc < 2; ++cc) {
              br_[r][cc] = fma(vr, zr[cc], fma(-vi, zi[cc], br_[r][cc]));
              bi_[r][cc] = fma(vr, zi[cc], fma(vi, zr[cc], bi_[r][cc]));
            }
          }
        }
        zd* Cb = canon + (size_t)edge * N * ccols;
#pragma unroll
        for (int cc = 0; cc < 2; ++cc) {
          const int c = c0 + cc;
          if (c < ccols) {
#pragma unroll
            for (int r = 0; r < 4; ++r) Cb[(a0 + r) * ccols + c] = mk(br_[r][cc], bi_[r][cc]);
          }
        }
      }
      return;
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

template <int N>
__global__ __launch_bounds__(64, N <= 16 ? 3 : 1) void k_canon_h(const cplx<double>* __restrict__ ext, int E, int ccols,
                                                double pinv_eps, double deps, zd* __restrict__ canon,
                                                double* __restrict__ lfull, double* __restrict__ ws) {
  canon_h_run<N>(ext, E, ccols, pinv_eps, deps, canon, lfull, ws, blockIdx.x);
}

// One job's edges in a launch that covers several jobs' canonicalisers (the serving batch):
// its buffers and parameters, and its first block in the launch.
struct CanonEntry {
  const cplx<double>* ext;
  zd* canon;
  double* lfull;
  double* ws;
  int E, ccols;
  double pinv_eps, deps;
  int block0;
};

// The canonicalisers of several jobs (same n) in one launch: entry e takes blocks
// block0_e .. block0_{e+1} - 1 (found by binary search; block0 ascending, block0_0 = 0).
template <int N>
__global__ __launch_bounds__(64, N <= 16 ? 3 : 1) void k_canon_h_multi(const CanonEntry* __restrict__ tab,
                                                                      int entries) {
  int lo = 0, hi = entries - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) / 2;
    if (tab[mid].block0 <= (int)blockIdx.x)
      lo = mid;
    else
      hi = mid - 1;
  }
  const CanonEntry& e = tab[lo];
  canon_h_run<N>(e.ext, e.E, e.ccols, e.pinv_eps, e.deps, e.canon, e.lfull, e.ws, (int)blockIdx.x - e.block0);
}

}  // namespace bqa
