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
  unsigned gmax;  // guided mode: float bits of max_s Z_s |D| (gain scale)
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
    size_t b = sizeof(double) * (5 * TAB + KMAX + HEA_MAX_S * (size_t)tp +
                                  (HEA_MAX_S * KMAX * KMAX * KMAX + 1) / 2);
    b += sizeof(int2) * (size_t)task + sizeof(McShared);
    if (sp_lds) b += ((size_t)n + 15) & ~(size_t)15;
    return b;
  }
  __device__ void carve(unsigned char* base, int task, int tp, bool sp_lds) {
    G = reinterpret_cast<double*>(base);
    E0 = G + TAB;
    R = E0 + TAB;
    tgt = R + TAB;
    H = tgt + TAB;  // followed by D: [S][KMAX][KMAX][KMAX] floats
    conc = H + TAB + (HEA_MAX_S * KMAX * KMAX * KMAX + 1) / 2;
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

// Guided mode: H_s = dF2/dn_s + (dF2/dn_s)^T with the closed-form
// dF2/dn_sab = -2 w_s (alpha_sab - t_sab) / (Z_s N c_a c_b) of the reference
// fitness squared (accelerate.cpp:556-569) at gamma = n / (Z_s N); then
// D[s][a][b][d] = H_s[b][d] - H_s[a][d], so that swapping i (species a) with
// j (species b) changes F^2 by sum_s sum_d D[s][a][b][d] (cnt_s[i][d] -
// cnt_s[j][d]) to first order (cnt: neighbour species counts).
template <int KMAX>
__device__ void guide_tables(const HeaMcArgs& A, const Lds<KMAX>& T) {
  const int K = A.K, S = A.S, N = A.N, nthr = blockDim.x, tid = threadIdx.x;
  for (int e = tid; e < S * K * K; e += nthr) {
    const int s = e / (K * K), c = (e / K) % K, b = e % K;
    const double zn = (double)A.Z[s] * N;
    const double ccb = T.conc[c] * T.conc[b];
    const double dcb =
        -2.0 * A.weight[s] * ((1.0 - T.G[T.idx(s, c, b)] / ccb) - T.tgt[T.idx(s, c, b)]) / (zn * ccb);
    const double dbc =
        -2.0 * A.weight[s] * ((1.0 - T.G[T.idx(s, b, c)] / ccb) - T.tgt[T.idx(s, b, c)]) / (zn * ccb);
    T.H[T.idx(s, c, b)] = dcb + dbc;
  }
  if (tid == 0) T.sh->gmax = 0u;
  __syncthreads();
  float* D = reinterpret_cast<float*>(T.H + Lds<KMAX>::TAB);
  unsigned mx = 0u;
  for (int e = tid; e < S * K * K * K; e += nthr) {
    const int s = e / (K * K * K), a = (e / (K * K)) % K, b = (e / K) % K, d = e % K;
    const float v = (float)(T.H[T.idx(s, b, d)] - T.H[T.idx(s, a, d)]);
    D[((s * KMAX + a) * KMAX + b) * KMAX + d] = v;
    mx = max(mx, __float_as_uint(fabsf(v) * (float)A.Z[s]));
  }
  atomicMax(&T.sh->gmax, mx);
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

  const bool guided = A.hist != nullptr;
  // guided mode: hist[k] = packed neighbour species counts of site k, one
  // 32-bit word per shell (5-bit fields), 16 bytes per site
  uint32_t* hist = guided ? A.hist + (size_t)l * N * 4 : nullptr;
  if (guided) {
    for (int k = tid; k < N; k += nthr) {
      const int32_t* row = A.nbr + (size_t)k * A.stride;
      uint32_t m[4] = {0u, 0u, 0u, 0u};
      for (int s = 0; s < S; ++s)
        for (int q = 0; q < A.Z[s]; ++q) m[s] += 1u << (5 * sp[row[A.off[s] + q]]);
      *reinterpret_cast<uint4*>(hist + (size_t)k * 4) = make_uint4(m[0], m[1], m[2], m[3]);
    }
    guide_tables<KMAX>(A, T);
  }

  while (acc < A.depth) {
    if (guided) {
      // ---- guided: each slot takes the Gumbel-max of -gain/(tau*scale)
      // over g_cand keyed candidate pairs (same-species candidates skipped)
      const float* D = reinterpret_cast<const float*>(T.H + Lds<KMAX>::TAB);
      const float gs = __uint_as_float(T.sh->gmax);
      const float inv = (float)(1.0 / (A.g_tau * (gs > 0.0f ? (double)gs : 1.0)));
      const int c4 = (A.g_cand + 3) & ~3;  // candidates per slot, in groups of 4
      for (int j = tid; j < task; j += nthr) {
        float best = -3.0e38f;
        int2 bp = make_int2(-1, -1);
        for (int c0 = 0; c0 < A.g_cand; c0 += 4) {
          const uint32_t pi0 = pos + (uint32_t)(j * c4 + c0);  // multiple of 4
          const hea_u32x4 w01 = hea_stream_block(A.seed, iter, A.phase, (uint32_t)l, pi0 >> 1);
          const hea_u32x4 w23 = hea_stream_block(A.seed, iter, A.phase, (uint32_t)l, (pi0 >> 1) + 1);
          const hea_u32x4 nz = hea_stream_block(A.seed, iter, A.phase | 0x100u, (uint32_t)l, pi0 >> 2);
          const uint32_t ws[8] = {w01.x, w01.y, w01.z, w01.w, w23.x, w23.y, w23.z, w23.w};
          const uint32_t ns[4] = {nz.x, nz.y, nz.z, nz.w};
#pragma unroll
          for (int c = 0; c < 4; ++c) {
            if (c0 + c < A.g_cand) {
              const uint32_t s1 = ws[2 * c] % (uint32_t)N, s2 = ws[2 * c + 1] % (uint32_t)N;
              const int a = sp[s1], b = sp[s2];
              if (a != b) {
                const uint4 h1 = *reinterpret_cast<const uint4*>(hist + (size_t)s1 * 4);
                const uint4 h2 = *reinterpret_cast<const uint4*>(hist + (size_t)s2 * 4);
                const uint32_t hh1[3] = {h1.x, h1.y, h1.z}, hh2[3] = {h2.x, h2.y, h2.z};
                float d = 0.0f;
#pragma unroll
                for (int s = 0; s < HEA_MAX_S; ++s) {
                  if (s < S) {
                    const float* Dr = D + ((s * KMAX + a) * KMAX + b) * KMAX;
#pragma unroll
                    for (int q = 0; q < KMAX; ++q)
                      if (q < K)
                        d += Dr[q] * (float)((int)((hh1[s] >> (5 * q)) & 31u) -
                                             (int)((hh2[s] >> (5 * q)) & 31u));
                  }
                }
                const float U = ((float)ns[c] + 0.5f) * (1.0f / 2147483648.0f);
                const float sc = -d * inv - __logf(-__logf(U));
                if (sc > best) {
                  best = sc;
                  bp = make_int2((int)s1, (int)s2);
                }
              }
            }
          }
        }
        T.prop[j] = bp;
      }
      pos += (uint32_t)(task * ((A.g_cand + 3) & ~3));
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
      if (guided) {
        // neighbours of the swapped sites see a -> b (around pr.x) and
        // b -> a (around pr.y) in their histograms
        int zs = 0;
        for (int s = 0; s < S; ++s) zs += A.Z[s];
        for (int t = tid; t < 2 * zs; t += nthr) {
          const int site = t < zs ? pr.x : pr.y;
          int q = t < zs ? t : t - zs, s = 0;
          while (q >= A.Z[s]) q -= A.Z[s++];
          const int nb = A.nbr[(size_t)site * A.stride + A.off[s] + q];
          const uint32_t dv = t < zs ? (1u << (5 * b)) - (1u << (5 * a)) : (1u << (5 * a)) - (1u << (5 * b));
          atomicAdd(hist + (size_t)nb * 4 + s, dv);
        }
      }
      for (int e = tid; e < S * K * K; e += nthr) {
        const int s = e / (K * K), i = (e / K) % K, j = e % K;
        refresh_entry(A, T, s, i, j);
      }
      __syncthreads();
      if (guided) {
        // the histogram atomics went to L2: drop this CU's L1 copies
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        guide_tables<KMAX>(A, T);
      }
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
// queue is empty.  B.probs[p] holds problem p's complete launch arguments
// (its seed is the batch seed + p).
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
    const HeaMcArgs& A = B.probs[it.x];
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
