// Screen-window arithmetic of the local search, shared by the kernels
// (hea_mc.hip) and host-side checks: the safe step of a gamma entry and the
// exact replays behind the fast path of the FP64 bookkeeping, each entry's
// screen error bounds and the window half-width beta built from them, the
// division-free screen of one shell, and the exact new gamma entries of a
// proposal.  The file header of hea_mc.hip states the bookkeeping argument;
// the bound argument sits beside refresh_entry_core.
//
// The functions are templates over the argument and table types: `Args`
// provides N, K, S, Z[], delta[], weight[] (and nbr, stride, off[] for the
// replay path), `Tab` provides the tables G, E0, R, tgt, SC (with .x, .y),
// H, conc, idx(s, i, j), hidx(s, i, j) and sh->bmax[], sh->ubmax[] (bit
// patterns of non-negative doubles).
#ifndef HEA_SCREEN_H
#define HEA_SCREEN_H

#include <math.h>
#include <stdint.h>
#include <type_traits>

#if defined(__HIPCC__)
#define HEA_SCREEN_HD __host__ __device__ __forceinline__
#define HEA_SCREEN_MEMBER __host__ __device__ __forceinline__
#else
#define HEA_SCREEN_HD static inline
#define HEA_SCREEN_MEMBER inline
#endif

#ifndef HEA_MAX_S
#define HEA_MAX_S 3
#endif

constexpr double kHeaU = 0x1p-53;  // unit roundoff of binary64

HEA_SCREEN_HD double hea_bits_f64(unsigned long long b) { return __builtin_bit_cast(double, b); }

// Species histogram of one shell walk, packed one field per species:
// 5-bit fields in 32 bits for K <= 6, 8-bit fields in 64 bits for K <= 8.
template <int KMAX>
struct Pack {
  using T = typename std::conditional<(KMAX <= 6), uint32_t, uint64_t>::type;
  static constexpr int W = (KMAX <= 6) ? 5 : 8;
  static HEA_SCREEN_MEMBER int fld(T m, int x) {
    return (int)((m >> (W * x)) & (T)((1u << W) - 1u));
  }
  static HEA_SCREEN_MEMBER T one(int x) { return (T)1 << (W * x); }
};

// Step r of a gamma entry valid for any sequence of <= 2*Z additions of
// +-delta, or 0 when the entry must be replayed exactly (see file header).
HEA_SCREEN_HD double safe_step(double g, double delta, int z) {
  if (!(g > 0.0)) return 0.0;
  int ex;
  (void)frexp(g, &ex);
  const double lo = ldexp(1.0, ex - 1), hi = ldexp(1.0, ex), u = ldexp(1.0, ex - 53);
  const double B = (2.0 * z + 3.0) * (delta + u) * (1.0 + 0x1p-20);
  if (g - B < lo || g + B >= hi) return 0.0;
  const double t = g + delta;
  const double r = t - g;      // exact: t and g share a binade
  const double e = delta - r;  // exact (Fast2Sum, g > delta)
  if (fabs(e) == 0.5 * u) return 0.0;
  return r;
}

// Exact replay of "n1 additions of d1, then n2 additions of -d1".
HEA_SCREEN_HD double replay(double v, double d1, int n1, int n2) {
  for (int i = 0; i < n1; ++i) v = v + d1;
  for (int i = 0; i < n2; ++i) v = v - d1;
  return v;
}

// Exact replay of entry (a,b) for proposal (a1, a2) over one shell
// (accelerate.cpp:517-549 restricted to the updates that hit (a,b)).
HEA_SCREEN_HD double replay_ab(double v, double delta, const int32_t* __restrict__ r1,
                                            const int32_t* __restrict__ r2, int z,
                                            const uint8_t* __restrict__ sp, int a1, int a2, int a,
                                            int b) {
  for (int k = 0; k < z; ++k) {
    const int nb = r1[k];
    const int n = sp[nb];
    const int na = (nb == a2) ? a : (nb == a1) ? b : n;
    if (n == b) v = v - delta;   // gamma[a][n]  -= d
    if (na == a) v = v + delta;  // gamma[na][b] += d
  }
  for (int k = 0; k < z; ++k) {
    const int nb = r2[k];
    const int m = sp[nb];
    const int ma = (nb == a1) ? b : (nb == a2) ? a : m;
    if (m == a) v = v - delta;   // gamma[m][b]  -= d
    if (ma == b) v = v + delta;  // gamma[a][ma] += d
  }
  return v;
}

// Refreshes the tables of entry (s,i,j) from prev_gamma and records its
// screen error bounds.
//
// Screen error bound.  Write q = G/dd, d = alpha - t with the reference's
// roundings (d-hat below), and for a proposal P let Delta = -(G' - G)/dd be
// the exact change of alpha of an entry, |Delta| <= A (a net count of at
// most 6 Z_s steps of delta_s; at most 6 Z_s additions, each rounding by at
// most 2^-52 since |gamma| < 2).  The reference's
// d-hat' - d-hat differs from Delta by e1 <= u (m + m'), m = |q| + |alpha| +
// |d| (one rounding each in q, alpha and d); the screen's
// Delta~ = fl(net * fl(-delta/dd)) differs from Delta by e2 <= (12 Z u delta
// + (6Z + 4) 2^-52) / dd.  The true change of the entry pair's two terms,
// fl(d'^2) - fl(d^2) summed over (i,j) and (j,i), then differs from the
// screen's 2 Delta~ (D + Delta~), D = fl(d_ij + d_ji), by at most
//   P = 2 e2 (D1 + A) + 2 (A + e2)(u D1 + e2) + 2 e1 (D1 + 2A + e1)
//       + 2 u (D1 + A + e1)^2 + 2 u (A + e2)(D1 + A + e2),   D1 = |d_ij|+|d_ji|
// (diagonal entries: the same with D1 = 2|d_ii|).  A proposal changes at most
// 2K-1 such pairs per shell.  Adding the screen's own summation rounding and
// the reference's rounding of its K*K-term sums (relative (K*K + S + 2) u of
// err <= E_up = sum_s |w_s| K^2 ub_s^2, ub_s >= |d'| of every entry) bounds
// |err(P) - (E + X(P))| by beta0 for one E common to all proposals.  With
// beta = beta0 + 4 u E_up, any proposal with X > Xmin + 2 beta has
// err > err(Pmin)(1 + 8u), hence sqrt(err) rounds strictly above the
// minimum's F and cannot be the reference's choice.
struct EntryBounds {
  double P, ub;  // the entry's pair error bound and its bound on |d'|
  float hz;      // guided mode: Z_s |H_ij|
};

template <int KMAX, bool GUIDED, class Args, class Tab>
HEA_SCREEN_HD EntryBounds refresh_entry_core(const Args& A, const Tab& T, int s, int i, int j) {
  const int e = T.idx(s, i, j);
  const double g = T.G[e];
  const double dd = T.conc[i] * T.conc[j];
  const double q = g / dd;
  const double al = 1.0 - q;
  const double dij = al - T.tgt[e];
  const double dji = al - T.tgt[T.idx(s, j, i)];
  T.E0[e] = dij * dij;
  T.R[e] = safe_step(g, A.delta[s], A.Z[s]);
  T.SC[e].x = dij + dji;
  // error bounds (each computed with a few roundings; the 1.0001 factors
  // cover them)
  const double z = (double)A.Z[s], delta = A.delta[s];
  const double D1 = fabs(dij) + fabs(dji);
  // 1/dd from the screen coefficient |c| = fl(delta/dd) and 1/delta = Z N
  // (each within a few u; the 1.0001 factors cover them)
  const double idd = fabs(T.SC[e].y) * (z * (double)A.N);
  const double Ae = (6.0 * z * delta + (6.0 * z + 4.0) * 0x1p-52) * idd * 1.0001;
  const double e2 = (12.0 * z * kHeaU * delta + (6.0 * z + 4.0) * 0x1p-52) * idd * 1.0001;
  const double m = 2.0 * fabs(q) + 2.0 + D1 + 3.0 * Ae;
  const double e1 = 2.0 * kHeaU * m * 1.0001;
  const double t1 = D1 + Ae + e1;
  const double P = 2.0 * e2 * (D1 + Ae) + 2.0 * (Ae + e2) * (kHeaU * D1 + e2) +
                   2.0 * e1 * (D1 + 2.0 * Ae + e1) + 2.0 * kHeaU * t1 * t1 +
                   2.0 * kHeaU * (Ae + e2) * (D1 + Ae + e2);
  const double ub = D1 + Ae + e1 + e2;
  EntryBounds r{P * 1.0001, ub * 1.0001, 0.0f};
  if constexpr (GUIDED) {
    // H_ij = dF2/dn_ij + dF2/dn_ji of the fitness squared
    // (accelerate.cpp:556-569) with gamma = n delta: dF2/dn_ij =
    // -2 w (alpha_ij - t_ij) delta / dd_ij = 2 w d_ij c_ij, c_ij = -delta/dd_ij
    const float h = (float)(2.0 * A.weight[s] * (dij + dji) * T.SC[e].y);
    T.H[T.hidx(s, i, j)] = h;
    r.hz = fabsf(h) * (float)A.Z[s];
  }
  return r;
}

// beta of the refreshed tables (see refresh_entry).
template <class Args, class Tab>
HEA_SCREEN_HD double screen_beta(const Args& A, const Tab& T) {
  const double K = (double)A.K, S = (double)A.S;
  double bs = 0.0, eup = 0.0;
  for (int s = 0; s < A.S; ++s) {
    const double bm = hea_bits_f64(T.sh->bmax[s]);
    const double ub = hea_bits_f64(T.sh->ubmax[s]);
    const double w = fabs(A.weight[s]);
    // pair errors, then the screen's fma chain and shell/weight combination
    // (|each screen term| <= 4 ub^2)
    bs += w * ((2.0 * K) * bm + (2.0 * K + S + 2.0) * kHeaU * (2.0 * K) * 4.0 * ub * ub);
    eup += w * (K * K) * ub * ub;
  }
  eup *= 1.01;
  return (bs + (K * K + S + 2.0) * kHeaU * eup + 4.0 * kHeaU * eup) * 1.01;
}

// Screen of shell s for swapping species a (at a1) and b (at a2): the
// division-free change of the shell's sum of squares, sum over the changed
// entry pairs of 2 Delta~ (D + Delta~) (diagonal entries once).
template <int KMAX, bool EXACT, class Tab>
HEA_SCREEN_HD double screen_shell(const Tab& T, int K, int s, int a, int b,
                                               typename Pack<KMAX>::T MA,
                                               typename Pack<KMAX>::T MB, int c11, int c12,
                                               int c21, int c22) {
  using PK = Pack<KMAX>;
  const int maA = PK::fld(MA, a), maB = PK::fld(MA, b);
  const int mbA = PK::fld(MB, a), mbB = PK::fld(MB, b);
  const int nAA = 2 * (mbA - c21 + c22) - 2 * maA;
  const int nBB = 2 * (maB - c12 + c11) - 2 * mbB;
  const int nAB = (maA + c12 - c11) + (mbB + c21 - c22) - maB - mbA;
  const auto* ra = T.SC + T.idx(s, a, 0);
  const auto* rb = T.SC + T.idx(s, b, 0);
  double x = 0.0;
#pragma unroll
  for (int q = 0; q < KMAX; ++q) {
    if (EXACT || q < K) {
      const int dm = PK::fld(MB, q) - PK::fld(MA, q);
      const int na = (q == a) ? nAA : (q == b) ? nAB : dm;
      const int nb = (q == b) ? nBB : -dm;
      const auto ea = ra[q];
      const auto eb = rb[q];
      const double ga = (double)na * ea.y;
      const double gb = (double)nb * eb.y;
      // (b,a) is the mirror of (a,b): row b skips it
      const double fa = (q == a) ? ga : 2.0 * ga;
      const double fb = (q == a) ? 0.0 : (q == b) ? gb : 2.0 * gb;
      x = fma(fa, ea.x + ga, x);
      x = fma(fb, eb.x + gb, x);
    }
  }
  return x;
}

// Exact new values of gamma entries (a,x) and (b,x) of shell s for swapping
// a1 <-> a2 (species a != b), from the stored walk of the two neighbour
// lists (file header: fast path, or replay when the entry's step is 0).
template <int KMAX, class Args, class Tab>
HEA_SCREEN_HD void exact_entries(const Args& A, const Tab& T,
                                              const uint8_t* __restrict__ sp, int s, int a1,
                                              int a2, int a, int b, int x,
                                              typename Pack<KMAX>::T MA,
                                              typename Pack<KMAX>::T MB, int c11, int c12,
                                              int c21, int c22, double& nA, double& nB) {
  using PK = Pack<KMAX>;
  const double delta = A.delta[s];
  const int maA = PK::fld(MA, a), maB = PK::fld(MA, b);
  const int mbA = PK::fld(MB, a), mbB = PK::fld(MB, b);
  const int max_ = PK::fld(MA, x), mbx = PK::fld(MB, x);
  const int pAA = 2 * (mbA - c21 + c22), mAA = 2 * maA;
  const int pBB = 2 * (maB - c12 + c11), mBB = 2 * mbB;
  const int netAB = (maA + c12 - c11) + (mbB + c21 - c22) - maB - mbA;
  const int eA = T.idx(s, a, x), eB = T.idx(s, b, x);
  const double gA = T.G[eA], rA = T.R[eA];
  const double gB = T.G[eB], rB = T.R[eB];
  // (a,a) and (b,a) = (a,b) for x = a; (a,b) and (b,b) for x = b
  const int netA = (x == a) ? pAA - mAA : (x == b) ? netAB : mbx - max_;
  const int netB = (x == a) ? netAB : (x == b) ? pBB - mBB : max_ - mbx;
  nA = fma((double)netA, rA, gA);
  nB = fma((double)netB, rB, gB);
  if (rA == 0.0 || rB == 0.0) {  // rare: near a binade edge, tied step or zero count
    const int32_t* r1 = A.nbr + (size_t)a1 * A.stride + A.off[s];
    const int32_t* r2 = A.nbr + (size_t)a2 * A.stride + A.off[s];
    if (rA == 0.0)
      nA = (x == a)   ? replay(gA, -delta, mAA, pAA)
           : (x == b) ? replay_ab(gA, delta, r1, r2, A.Z[s], sp, a1, a2, a, b)
                      : replay(gA, -delta, max_, mbx);
    if (rB == 0.0)
      nB = (x == a)   ? replay_ab(gB, delta, r1, r2, A.Z[s], sp, a1, a2, a, b)
           : (x == b) ? replay(gB, delta, pBB, mBB)
                      : replay(gB, delta, max_, mbx);
  }
}

#endif  // HEA_SCREEN_H
