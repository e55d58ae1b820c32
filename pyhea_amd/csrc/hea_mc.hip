// Greedy best-of-T local search (the hot path), one workgroup per lattice.
//
// Semantics follow pyhea/cpp/accelerate.cpp:601-694 (local_parallel_monte_carlo)
// and :498-570 (calculate_fitness_incremental):
//   * gamma starts from a fresh recount, gamma = R_s(pair count)
//     (accelerate.cpp:636-640), while prev_fit starts from the stored fitness
//     (accelerate.cpp:632);
//   * each round draws `task` valid proposals (unlike species) from the
//     lattice's stream, evaluates each incrementally from prev_gamma, keeps
//     the first strict minimum (accelerate.cpp:663-674), accepts it iff it is
//     strictly below prev_fit (accelerate.cpp:678) and carries the winner's
//     drifted gamma forward (accelerate.cpp:682);
//   * the search ends after `depth` consecutive rejected rounds.
//
// Work layout.  A workgroup has S x TP lanes: lane (s, j) evaluates shell s
// of proposal j (TP = proposals per pass, a multiple of 64, so every wave
// holds one shell and walks neighbour lists of one uniform length).  Each
// lane produces its shell's partial sum e_s = sum_ij (alpha_sij - t_sij)^2 in
// the reference's (i, j) order; the shell-0 lane then forms
// err = ((0 + w0 e0) + w1 e1) + w2 e2 and F = sqrt(err) exactly as
// accelerate.cpp:557-569 does.  The lattice's species sit in LDS when they
// fit, otherwise they are read from HBM.
//
// Exact FP64 bookkeeping.  The reference updates gamma by +-delta_s with
// delta_s = 1/(Z_s*N) (accelerate.cpp:526-548); the result depends on the
// order of those additions.  Every update to entry (p,q) is paired with the
// same update to (q,p), so gamma stays bit-symmetric and only rows a = sp[a1]
// and b = sp[a2] change.  For one proposal the updates an entry receives are:
//   (a,x), x not in {a,b}: -delta  MA[x] times, then +delta MB[x] times
//   (b,x), x not in {a,b}: +delta  MA[x] times, then -delta MB[x] times
//   (a,a): -delta 2*MA[a] times, then +delta 2*(MB[a]-c21+c22) times
//   (b,b): +delta 2*(MA[b]-c12+c11) times, then -delta 2*MB[b] times
//   (a,b)=(b,a): interleaved in neighbour order (a2 adjacent to a1 gives -,+)
// where MA/MB count the species among a1's/a2's shell neighbours and cXY
// counts site Y in X's list.  Inside one binade of the entry's value g with
// no rounding tie, fl(w +- delta) = w +- r for the fixed step
// r = fl(g + delta) - g, so any such sequence of at most 2*Z_s steps ends at
// g + net*r exactly, computed by one fma.  Entries whose window of reachable
// values is not provably inside one binade, or whose step ties, are marked
// r = 0 and replayed addition by addition.  Either way the result equals the
// reference's doubles bit for bit.  No FP contraction is allowed in this file
// (-ffp-contract=off).

#include <type_traits>

#include "hea_internal.h"

namespace {

constexpr int kMaxWaves = 16;  // 1024 lanes

struct McShared {
  int wcount[kMaxWaves];
  unsigned long long wkey[kMaxWaves];
  int wj[kMaxWaves];
  uint32_t next_pos;
  int pad;
  double gscale;  // guided mode: scale of the site gains
};

// Species histogram of one shell walk, packed one field per species:
// 5-bit fields in 32 bits for K <= 6, 8-bit fields in 64 bits for K <= 8.
template <int KMAX>
struct Pack {
  using T = typename std::conditional<(KMAX <= 6), uint32_t, uint64_t>::type;
  static constexpr int W = (KMAX <= 6) ? 5 : 8;
  static __device__ __forceinline__ int fld(T m, int x) {
    return (int)((m >> (W * x)) & (T)((1u << W) - 1u));
  }
  static __device__ __forceinline__ T one(int x) { return (T)1 << (W * x); }
};

// Step r of a gamma entry valid for any sequence of <= 2*Z additions of
// +-delta, or 0 when the entry must be replayed exactly (see file header).
__device__ __forceinline__ double safe_step(double g, double delta, int z) {
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
__device__ __forceinline__ double replay(double v, double d1, int n1, int n2) {
  for (int i = 0; i < n1; ++i) v = v + d1;
  for (int i = 0; i < n2; ++i) v = v - d1;
  return v;
}

// Exact replay of entry (a,b) for proposal (a1, a2) over one shell
// (accelerate.cpp:517-549 restricted to the updates that hit (a,b)).
__device__ __forceinline__ double replay_ab(double v, double delta, const int32_t* __restrict__ r1,
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

// Species histogram of one neighbour list: M += 1<<(W*species); cs counts
// `self`, co counts `other`.  Rows are padded to a multiple of 4 entries.
template <int KMAX, bool SELF>
__device__ __forceinline__ void walk(const int32_t* __restrict__ row, int z,
                                     const uint8_t* __restrict__ sp, int self, int other,
                                     typename Pack<KMAX>::T& M, int& cs, int& co) {
  M = 0;
  cs = 0;
  co = 0;
  for (int q = 0; q < z; q += 4) {
    const int4 v = *reinterpret_cast<const int4*>(row + q);
    const int nb[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      if (q + t < z) {
        const int n = nb[t];
        M += Pack<KMAX>::one((int)sp[n]);
        if (SELF) cs += (n == self);
        co += (n == other);
      }
    }
  }
}

// LDS carve-up of one workgroup.
template <int KMAX>
struct Lds {
  static constexpr int TAB = HEA_MAX_S * KMAX * KMAX;
  double* G;     // [S][KMAX][KMAX] prev_gamma
  double* E0;    // [S][KMAX][KMAX] (1 - G/d - t)^2 of prev_gamma
  double* R;     // [S][KMAX][KMAX] safe step or 0
  double* tgt;   // [S][KMAX][KMAX] target alpha
  double* H;     // [S][KMAX][KMAX] guided mode: dF2/dn + its transpose
  double* conc;  // [KMAX] fl(count/N)
  double* esum;  // [S][TP] per-shell partial sums
  int2* prop;    // [task]
  McShared* sh;
  uint8_t* sp;   // [N] when resident
  static __host__ __device__ size_t bytes(int task, int tp, int n, bool sp_lds) {
    size_t b = sizeof(double) * (5 * TAB + KMAX + HEA_MAX_S * (size_t)tp);
    b += sizeof(int2) * (size_t)task + sizeof(McShared);
    if (sp_lds) b += ((size_t)n + 15) & ~(size_t)15;
    return b;
  }
  __device__ void carve(unsigned char* base, int task, int tp, bool sp_lds) {
    G = reinterpret_cast<double*>(base);
    E0 = G + TAB;
    R = E0 + TAB;
    tgt = R + TAB;
    H = tgt + TAB;
    conc = H + TAB;
    esum = conc + KMAX;
    prop = reinterpret_cast<int2*>(esum + HEA_MAX_S * tp);
    sh = reinterpret_cast<McShared*>(prop + task);
    sp = sp_lds ? reinterpret_cast<uint8_t*>(sh + 1) : nullptr;
  }
  __device__ __forceinline__ int idx(int s, int i, int j) const { return (s * KMAX + i) * KMAX + j; }
};

template <int KMAX>
__device__ void refresh_entry(const HeaMcArgs& A, const Lds<KMAX>& T, int s, int i, int j) {
  const int e = T.idx(s, i, j);
  const double g = T.G[e];
  const double d = T.conc[i] * T.conc[j];
  const double sro = 1.0 - g / d;
  const double df = sro - T.tgt[e];
  T.E0[e] = df * df;
  T.R[e] = safe_step(g, A.delta[s], A.Z[s]);
}

// Shell s of the incremental fitness of swapping a1 <-> a2 (species a != b)
// against prev_gamma.  Returns the shell's partial sum and fills the swapped
// gamma rows a (nA) and b (nB) of that shell.
template <int KMAX, bool EXACT, bool SELF>
__device__ __forceinline__ double eval_shell(const HeaMcArgs& A, const Lds<KMAX>& T,
                                             const uint8_t* __restrict__ sp, int s, int a1,
                                             int a2, double (&nA)[KMAX], double (&nB)[KMAX]) {
  using PK = Pack<KMAX>;
  const int K = EXACT ? KMAX : A.K;
  const int z = A.Z[s];
  const double delta = A.delta[s];
  const int a = sp[a1], b = sp[a2];
  const int32_t* r1 = A.nbr + (size_t)a1 * A.stride + A.off[s];
  const int32_t* r2 = A.nbr + (size_t)a2 * A.stride + A.off[s];
  typename PK::T MA, MB;
  int c11, c12, c22, c21;
  walk<KMAX, SELF>(r1, z, sp, a1, a2, MA, c11, c12);
  walk<KMAX, SELF>(r2, z, sp, a2, a1, MB, c22, c21);
  const int maA = PK::fld(MA, a), maB = PK::fld(MA, b);
  const int mbA = PK::fld(MB, a), mbB = PK::fld(MB, b);

  // fast path for every touched entry; collect the ones needing replay
  const int eAA = T.idx(s, a, a), eBB = T.idx(s, b, b), eAB = T.idx(s, a, b);
  const int pAA = 2 * (mbA - c21 + c22), mAA = 2 * maA;
  const int pBB = 2 * (maB - c12 + c11), mBB = 2 * mbB;
  const int netAB = (maA + c12 - c11) + (mbB + c21 - c22) - maB - mbA;
  const double gAA = T.G[eAA], rAA = T.R[eAA];
  const double gBB = T.G[eBB], rBB = T.R[eBB];
  const double gAB = T.G[eAB], rAB = T.R[eAB];
  double vAA = fma((double)(pAA - mAA), rAA, gAA);
  double vBB = fma((double)(pBB - mBB), rBB, gBB);
  double vAB = fma((double)netAB, rAB, gAB);
  unsigned need = (rAA == 0.0 ? 1u : 0u) | (rBB == 0.0 ? 2u : 0u) | (rAB == 0.0 ? 4u : 0u);
  int ma[KMAX], mb[KMAX];
#pragma unroll
  for (int x = 0; x < KMAX; ++x) {
    ma[x] = PK::fld(MA, x);
    mb[x] = PK::fld(MB, x);
    if (EXACT || x < K) {
      const int eA = T.idx(s, a, x), eB = T.idx(s, b, x);
      const double gA = T.G[eA], rA = T.R[eA];
      const double gB = T.G[eB], rB = T.R[eB];
      nA[x] = fma((double)(mb[x] - ma[x]), rA, gA);
      nB[x] = fma((double)(ma[x] - mb[x]), rB, gB);
      const bool gen = (x != a) && (x != b);
      need |= (gen && rA == 0.0) ? (8u << x) : 0u;
      need |= (gen && rB == 0.0) ? (2048u << x) : 0u;
    }
  }
  if (need) {  // rare: entries near a binade edge, tied steps or zero counts
    if (need & 1u) vAA = replay(gAA, -delta, mAA, pAA);
    if (need & 2u) vBB = replay(gBB, delta, pBB, mBB);
    if (need & 4u) vAB = replay_ab(gAB, delta, r1, r2, z, sp, a1, a2, a, b);
#pragma unroll
    for (int x = 0; x < KMAX; ++x) {
      if (need & (8u << x)) nA[x] = replay(T.G[T.idx(s, a, x)], -delta, ma[x], mb[x]);
      if (need & (2048u << x)) nB[x] = replay(T.G[T.idx(s, b, x)], delta, ma[x], mb[x]);
    }
  }

  const double conc_a = T.conc[a], conc_b = T.conc[b];
  double sA[KMAX], sB[KMAX];
#pragma unroll
  for (int x = 0; x < KMAX; ++x) {
    if (EXACT || x < K) {
      nA[x] = (x == a) ? vAA : (x == b) ? vAB : nA[x];
      nB[x] = (x == b) ? vBB : (x == a) ? vAB : nB[x];
      sA[x] = 1.0 - nA[x] / (conc_a * T.conc[x]);
      sB[x] = 1.0 - nB[x] / (conc_b * T.conc[x]);
    }
  }

  // all K*K terms in row-major order (accelerate.cpp:560-566)
  const double* tg = T.tgt + T.idx(s, 0, 0);
  const double* e0 = T.E0 + T.idx(s, 0, 0);
  double es = 0.0;
#pragma unroll
  for (int i = 0; i < KMAX; ++i) {
    if (EXACT || i < K) {
      const bool ia = (i == a), ib = (i == b);
#pragma unroll
      for (int j = 0; j < KMAX; ++j) {
        if (EXACT || j < K) {
          const bool ja = (j == a), jb = (j == b);
          const double sro = ia ? sA[j] : ib ? sB[j] : ja ? sA[i] : sB[i];
          const double df = sro - tg[i * KMAX + j];
          const double et = df * df;
          const double e = (ia || ib || ja || jb) ? et : e0[i * KMAX + j];
          es = es + e;
        }
      }
    }
  }
  return es;
}

// Guided mode: site gains u[k][c] = g[k][c] - g[k][sp[k]] of lattice l at
// its current gamma, with g[k][c] = sum_s sum_b H_s[c][b] cnt_s[k][b] and
// H_s = dF2/dn_s + (dF2/dn_s)^T, dF2/dn_sab = -2 w_s (alpha_sab - t_sab) /
// (Z_s N c_a c_b) (the closed-form derivative of the reference fitness
// squared, accelerate.cpp:556-569, at gamma = n / (Z_s N)).
template <int KMAX>
__device__ void refresh_gain(const HeaMcArgs& A, const Lds<KMAX>& T, const uint8_t* sp, int l) {
  const int K = A.K, S = A.S, N = A.N, nthr = blockDim.x, tid = threadIdx.x;
  for (int e = tid; e < S * K * K; e += nthr) {
    const int s = e / (K * K), c = (e / K) % K, b = e % K;
    const double zn = (double)A.Z[s] * N;
    const double ccb = T.conc[c] * T.conc[b];
    const double dcb = -2.0 * A.weight[s] * ((1.0 - T.G[T.idx(s, c, b)] / ccb) - T.tgt[T.idx(s, c, b)]) /
                       (zn * ccb);
    const double dbc = -2.0 * A.weight[s] * ((1.0 - T.G[T.idx(s, b, c)] / ccb) - T.tgt[T.idx(s, b, c)]) /
                       (zn * ccb);
    T.H[T.idx(s, c, b)] = dcb + dbc;
  }
  __syncthreads();
  if (tid == 0) {
    double sc = 0.0;
    for (int s = 0; s < S; ++s) {
      double m = 0.0;
      for (int c = 0; c < K; ++c)
        for (int b = 0; b < K; ++b) m = fmax(m, fabs(T.H[T.idx(s, c, b)]));
      sc += A.Z[s] * m;
    }
    T.sh->gscale = sc > 0.0 ? sc : 1.0;
  }
  float* u = A.gain + (size_t)l * N * K;
  for (int k = tid; k < N; k += nthr) {
    const int32_t* row = A.nbr + (size_t)k * A.stride;
    double g[KMAX];
    for (int c = 0; c < KMAX; ++c) g[c] = 0.0;
    for (int s = 0; s < S; ++s) {
      uint64_t m = 0;  // 8-bit species counts of this shell
      for (int q = 0; q < A.Z[s]; ++q) m += 1ull << (8 * sp[row[A.off[s] + q]]);
      for (int b = 0; b < K; ++b) {
        const double cb = (double)((m >> (8 * b)) & 255u);
        if (cb != 0.0)
          for (int c = 0; c < K; ++c) g[c] += T.H[T.idx(s, c, b)] * cb;
      }
    }
    const int a = sp[k];
    double ga = 0.0;
    for (int c = 0; c < K; ++c) ga = (c == a) ? g[c] : ga;
    for (int c = 0; c < K; ++c) u[(size_t)k * K + c] = (float)(g[c] - ga);
  }
  __syncthreads();
}

__device__ __forceinline__ bool key_less(unsigned long long f1, int j1, unsigned long long f2,
                                         int j2) {
  return (f1 < f2) || (f1 == f2 && j1 < j2);
}

// Local search of lattice l of the population described by A.
template <int KMAX, bool SP_LDS, bool EXACT, bool SELF>
__device__ __forceinline__ void mc_lattice(const HeaMcArgs& A, int l, unsigned char* smem) {
  const HeaConst& P = *A.pc;
  const int nthr = blockDim.x;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int N = A.N, K = A.K, S = A.S;
  const int task = A.task, TP = A.tp;
  const int nwaves = nthr >> 6;
  const int s_me = __builtin_amdgcn_readfirstlane(tid / TP);  // this wave's shell
  const int jl = tid - s_me * TP;                             // proposal slot in a pass
  const int nq = (task + TP - 1) / TP;                        // passes per round

  Lds<KMAX> T;
  T.carve(smem, task, TP, SP_LDS);
  uint8_t* lat_g = A.lattices + (size_t)l * N;
  uint8_t* sp;
  if constexpr (SP_LDS) {
    sp = T.sp;
    for (int i = tid; i < N; i += nthr) sp[i] = lat_g[i];
  } else {
    sp = lat_g;
  }

  // prev_gamma = R_s(pair counts) (accelerate.cpp:636-640) and constants
  const int32_t* cnt = A.counts + (size_t)l * S * K * K;
  for (int e = tid; e < S * K * K; e += nthr) {
    const int s = e / (K * K), i = (e / K) % K, j = e % K;
    T.G[T.idx(s, i, j)] = hea_r_of_count(P, s, cnt[e]);
    T.tgt[T.idx(s, i, j)] = P.target[s][i][j];
  }
  if (tid < K) T.conc[tid] = P.conc[tid];
  __syncthreads();
  for (int e = tid; e < S * K * K; e += nthr) {
    const int s = e / (K * K), i = (e / K) % K, j = e % K;
    refresh_entry(A, T, s, i, j);
  }
  __syncthreads();

  double prev_fit = A.fitness[l];
  const uint32_t iter = *A.iter_ptr;
  uint32_t pos = 0;
  int acc = 0;
  long long rounds = 0, accepts = 0;
  bool hit_thr = false;
  double nA[KMAX], nB[KMAX];  // this lane's shell rows after its last proposal

  const bool guided = A.gain != nullptr;
  if (guided) refresh_gain<KMAX>(A, T, sp, l);

  while (acc < A.depth) {
    if (guided) {
      // ---- guided: each slot takes the Gumbel-max of -gain/(tau*scale)
      // over g_cand keyed candidate pairs (same-species candidates skipped)
      if (rounds > 0 && rounds % A.g_every == 0) refresh_gain<KMAX>(A, T, sp, l);
      const float* u = A.gain + (size_t)l * N * K;
      const float inv = (float)(1.0 / (A.g_tau * T.sh->gscale));
      for (int j = tid; j < task; j += nthr) {
        float best = -3.0e38f;
        int2 bp = make_int2(-1, -1);
        for (int c = 0; c < A.g_cand; ++c) {
          const uint32_t pi = pos + (uint32_t)(j * A.g_cand + c);
          uint32_t s1, s2;
          hea_stream_pair(A.seed, iter, A.phase, (uint32_t)l, pi, (uint32_t)N, &s1, &s2);
          const int a = sp[s1], b = sp[s2];
          if (a != b) {
            const hea_u32x4 w = hea_stream_block(A.seed, iter, A.phase | 0x100u, (uint32_t)l, pi >> 2);
            const uint32_t wv = (pi & 3u) == 0 ? w.x : (pi & 3u) == 1 ? w.y : (pi & 3u) == 2 ? w.z : w.w;
            const float U = ((float)wv + 0.5f) * (1.0f / 2147483648.0f);
            const float sc = -(u[(size_t)s1 * K + b] + u[(size_t)s2 * K + a]) * inv -
                             logf(-logf(U));
            if (sc > best) {
              best = sc;
              bp = make_int2((int)s1, (int)s2);
            }
          }
        }
        T.prop[j] = bp;
      }
      pos += (uint32_t)(task * A.g_cand);
      __syncthreads();
    }
    // ---- draw `task` valid proposals in stream order (accelerate.cpp:649-655)
    int found = guided ? task : 0;
    while (found < task) {
      const uint32_t p = pos + tid;
      uint32_t s1, s2;
      hea_stream_pair(A.seed, iter, A.phase, (uint32_t)l, p, (uint32_t)N, &s1, &s2);
      const bool valid = sp[s1] != sp[s2];
      const unsigned long long bal = __ballot(valid);
      const int pre = __popcll(bal & ((1ull << lane) - 1ull));
      if (lane == 0) T.sh->wcount[wave] = __popcll(bal);
      __syncthreads();
      int off = 0, tot = 0;
      for (int w = 0; w < nwaves; ++w) {
        const int c = T.sh->wcount[w];
        off += (w < wave) ? c : 0;
        tot += c;
      }
      const int rank = found + off + pre;
      if (valid && rank < task) T.prop[rank] = make_int2((int)s1, (int)s2);
      if (valid && rank == task - 1) T.sh->next_pos = p + 1;
      found += tot;
      pos += nthr;
      __syncthreads();
      if (found >= task) pos = T.sh->next_pos;
    }

    // ---- evaluate; keep the first strict minimum (accelerate.cpp:663-674)
    unsigned long long bkey = ~0ull;
    int bj = 0x7fffffff;
    for (int q = 0; q < nq; ++q) {
      const int j = jl + q * TP;
      if (j < task && T.prop[j].x >= 0) {
        const int2 pr = T.prop[j];
        T.esum[s_me * TP + jl] =
            eval_shell<KMAX, EXACT, SELF>(A, T, sp, s_me, pr.x, pr.y, nA, nB);
      }
      __syncthreads();
      if (s_me == 0 && j < task && T.prop[j].x >= 0) {
        double err = 0.0;
        for (int s = 0; s < S; ++s) err = err + A.weight[s] * T.esum[s * TP + jl];
        const unsigned long long fk = __double_as_longlong(sqrt(err));
        if (fk < bkey) {
          bkey = fk;
          bj = j;
        }
      }
      if (q + 1 < nq) __syncthreads();
    }
    // argmin of (fitness bits, proposal index) over the shell-0 lanes;
    // fitness >= 0, so its bit pattern orders like the value
    const int nw0 = TP >> 6;
    if (s_me == 0) {
      unsigned long long rk = bkey;
      int rj = bj;
#pragma unroll
      for (int m = 32; m >= 1; m >>= 1) {
        const unsigned long long ok =
            ((unsigned long long)(uint32_t)__shfl_xor((int)(rk >> 32), m) << 32) |
            (uint32_t)__shfl_xor((int)(uint32_t)rk, m);
        const int oj = __shfl_xor(rj, m);
        if (key_less(ok, oj, rk, rj)) {
          rk = ok;
          rj = oj;
        }
      }
      if (lane == 0) {
        T.sh->wkey[wave] = rk;
        T.sh->wj[wave] = rj;
      }
    }
    __syncthreads();
    unsigned long long wk = T.sh->wkey[0];
    int wj = T.sh->wj[0];
    for (int w = 1; w < nw0; ++w)
      if (key_less(T.sh->wkey[w], T.sh->wj[w], wk, wj)) {
        wk = T.sh->wkey[w];
        wj = T.sh->wj[w];
      }
    const double best_fit = __longlong_as_double((long long)wk);
    ++rounds;

    // ---- accept iff strictly better (accelerate.cpp:678-689)
    if (best_fit < prev_fit) {
      const int2 pr = T.prop[wj];
      const int a = sp[pr.x], b = sp[pr.y];
      if (jl == wj % TP) {  // the winner's lane of every shell writes its rows
        if (nq > 1) (void)eval_shell<KMAX, EXACT, SELF>(A, T, sp, s_me, pr.x, pr.y, nA, nB);
#pragma unroll
        for (int x = 0; x < KMAX; ++x) {
          if (EXACT || x < K) {
            T.G[T.idx(s_me, a, x)] = nA[x];
            T.G[T.idx(s_me, x, a)] = nA[x];
            T.G[T.idx(s_me, b, x)] = nB[x];
            T.G[T.idx(s_me, x, b)] = nB[x];
          }
        }
      }
      prev_fit = best_fit;
      acc = 0;
      ++accepts;
      if (best_fit <= A.thr) hit_thr = true;
      __syncthreads();
      if (tid == 0) {
        sp[pr.x] = (uint8_t)b;
        sp[pr.y] = (uint8_t)a;
      }
      for (int e = tid; e < S * K * K; e += nthr) {
        const int s = e / (K * K), i = (e / K) % K, j = e % K;
        refresh_entry(A, T, s, i, j);
      }
      __syncthreads();
      if (hit_thr && tid == 0 && A.first_hit) {
        atomicMin(A.first_hit, (unsigned long long)wall_clock64());
        hit_thr = false;
      }
    } else {
      ++acc;
    }
  }

  if constexpr (SP_LDS) {
    for (int i = tid; i < N; i += nthr) lat_g[i] = sp[i];
  }
  if (tid == 0) {
    A.fitness[l] = prev_fit;
    int64_t* st = A.stats + (size_t)l * 4;
    st[0] += rounds;
    st[1] += accepts;
    st[2] += rounds * (long long)task;
    st[3] += (long long)pos;
  }
}

template <int KMAX, bool SP_LDS, bool EXACT, bool SELF>
__global__ __launch_bounds__(1024) void k_local_mc(HeaMcArgs A) {
  extern __shared__ __align__(16) unsigned char smem[];
  mc_lattice<KMAX, SP_LDS, EXACT, SELF>(A, blockIdx.x, smem);
}

// Persistent local search over many problems: every workgroup takes
// (problem, lattice) items from a device queue, longest first, until the
// queue is empty.  Per-problem fields come from B.probs; run fields from
// B.run.  Problem p draws from seed B.run.seed + p.
template <int KMAX, bool SP_LDS>
__global__ __launch_bounds__(1024) void k_local_mc_batch(HeaMcBatchArgs B) {
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ int item_sh;
  for (;;) {
    if (threadIdx.x == 0) item_sh = (int)atomicAdd(B.queue, 1u);
    __syncthreads();
    const int item = __builtin_amdgcn_readfirstlane(item_sh);
    __syncthreads();
    if (item >= B.n_items) break;
    const int2 it = B.items[item];
    HeaMcArgs A = B.probs[it.x];
    A.task = B.run.task;
    A.depth = B.run.depth;
    A.tp = B.run.tp;
    A.phase = B.run.phase;
    A.thr = B.run.thr;
    A.iter_ptr = B.run.iter_ptr;
    A.seed = B.run.seed + (uint64_t)it.x;
    A.g_every = B.run.g_every;
    A.g_cand = B.run.g_cand;
    A.g_tau = B.run.g_tau;
    mc_lattice<KMAX, SP_LDS, false, false>(A, it.y, smem);
    __syncthreads();
  }
}

template <int KMAX, bool SP_LDS, bool EXACT, bool SELF>
const void* mc_fn() {
  return reinterpret_cast<const void*>(k_local_mc<KMAX, SP_LDS, EXACT, SELF>);
}

// Kernel variant: exact-K instantiations for the common case, a guarded
// K <= KMAX one for K = 7 and mixed batches, and a self-neighbour one for
// cells so small that a site lists itself.
const void* mc_fn_for(int K, bool sp_lds, bool exact_ok, bool self) {
  if (self) return sp_lds ? mc_fn<8, true, false, true>() : mc_fn<8, false, false, true>();
  if (exact_ok) {
    switch (K) {
      case 2: return sp_lds ? mc_fn<2, true, true, false>() : mc_fn<2, false, true, false>();
      case 3: return sp_lds ? mc_fn<3, true, true, false>() : mc_fn<3, false, true, false>();
      case 4: return sp_lds ? mc_fn<4, true, true, false>() : mc_fn<4, false, true, false>();
      case 5: return sp_lds ? mc_fn<5, true, true, false>() : mc_fn<5, false, true, false>();
      case 6: return sp_lds ? mc_fn<6, true, true, false>() : mc_fn<6, false, true, false>();
      case 8: return sp_lds ? mc_fn<8, true, true, false>() : mc_fn<8, false, true, false>();
      default: break;
    }
  }
  if (K <= 6) return sp_lds ? mc_fn<6, true, false, false>() : mc_fn<6, false, false, false>();
  if (K <= 8) return sp_lds ? mc_fn<8, true, false, false>() : mc_fn<8, false, false, false>();
  return nullptr;
}

const void* mc_batch_fn_for(int kmax, bool sp_lds) {
  if (kmax <= 6)
    return sp_lds ? reinterpret_cast<const void*>(k_local_mc_batch<6, true>)
                  : reinterpret_cast<const void*>(k_local_mc_batch<6, false>);
  return sp_lds ? reinterpret_cast<const void*>(k_local_mc_batch<8, true>)
                : reinterpret_cast<const void*>(k_local_mc_batch<8, false>);
}

}  // namespace

hipError_t hea_prepare_local_mc_batch(int kmax, bool sp_lds, size_t lds) {
  return hipFuncSetAttribute(mc_batch_fn_for(kmax, sp_lds),
                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
}

hipError_t hea_launch_local_mc_batch(const HeaMcBatchArgs& b, int kmax, bool sp_lds, int grid,
                                     int threads, size_t lds, hipStream_t st) {
  void* args[] = {const_cast<HeaMcBatchArgs*>(&b)};
  return hipLaunchKernel(mc_batch_fn_for(kmax, sp_lds), dim3(grid), dim3(threads), args, lds, st);
}

// Table width of the kernel chosen for K species (see mc_fn_for).
int hea_mc_kmax(int K, bool self) {
  if (self || K == 7) return 8;
  return K <= 6 ? K : 8;
}

// Proposals evaluated per pass (lanes per shell): a multiple of 64, at most
// 256, with S * TP <= 1024 lanes.
int hea_mc_tp(int task, int S) {
  const int want = ((task + 63) / 64) * 64;
  const int cap = std::min(256, (1024 / S) / 64 * 64);
  return std::min(want, cap);
}

size_t hea_mc_lds_bytes(int kmax, int task, int tp, int n_sites, bool sp_lds) {
  switch (kmax) {
    case 2: return Lds<2>::bytes(task, tp, n_sites, sp_lds);
    case 3: return Lds<3>::bytes(task, tp, n_sites, sp_lds);
    case 4: return Lds<4>::bytes(task, tp, n_sites, sp_lds);
    case 5: return Lds<5>::bytes(task, tp, n_sites, sp_lds);
    case 6: return Lds<6>::bytes(task, tp, n_sites, sp_lds);
    default: return Lds<8>::bytes(task, tp, n_sites, sp_lds);
  }
}

// Raises the kernel's dynamic-LDS limit; call before launching (and before
// any graph capture that contains the launch).
hipError_t hea_prepare_local_mc(int K, bool sp_lds, bool self, size_t lds) {
  const void* fn = mc_fn_for(K, sp_lds, true, self);
  if (!fn) return hipErrorInvalidValue;
  return hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
}

hipError_t hea_launch_local_mc(const HeaMcArgs& a, int K, bool sp_lds, bool self, size_t lds,
                               hipStream_t st) {
  const void* fn = mc_fn_for(K, sp_lds, true, self);
  if (!fn) return hipErrorInvalidValue;
  void* args[] = {const_cast<HeaMcArgs*>(&a)};
  return hipLaunchKernel(fn, dim3(a.L), dim3(a.S * a.tp), args, lds, st);
}
