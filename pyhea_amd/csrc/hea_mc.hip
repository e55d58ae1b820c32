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
// reference's doubles bit for bit.
//
// The fitness sum runs over all K*K*S terms in the reference's (s, i, j)
// order (accelerate.cpp:557-569); untouched terms come from a per-round
// table.  No FP contraction is allowed in this file (-ffp-contract=off).

#include <type_traits>

#include "hea_internal.h"

namespace {

constexpr int kWaves = HEA_MC_THREADS / 64;

struct McShared {
  int wcount[kWaves];
  unsigned long long wkey[kWaves];
  int wj[kWaves];
  uint32_t next_pos;
  int pad;
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

template <int KMAX>
struct Tables {
  double* G;   // [3][KMAX][KMAX] prev_gamma
  double* E0;  // [3][KMAX][KMAX] (1 - G/d - t)^2 of prev_gamma
  double* R;   // [3][KMAX][KMAX] safe step or 0
  double* conc;  // [KMAX] fl(count/N)
  __device__ int idx(int s, int i, int j) const { return (s * KMAX + i) * KMAX + j; }
};

template <int KMAX>
__device__ void refresh_entry(const HeaConst& P, const Tables<KMAX>& T, int s, int i, int j) {
  const int e = T.idx(s, i, j);
  const double g = T.G[e];
  const double d = T.conc[i] * T.conc[j];
  const double sro = 1.0 - g / d;
  const double df = sro - P.target[s][i][j];
  T.E0[e] = df * df;
  T.R[e] = safe_step(g, P.delta[s], P.Z[s]);
}

// Counts of one neighbour walk: M += 1<<(5*species), cs counts `self`,
// co counts `other`.
template <int KMAX>
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
        cs += (n == self);
        co += (n == other);
      }
    }
  }
}

// Exact replay of entry (a,b) of shell s for proposal (a1, a2)
// (accelerate.cpp:517-549 restricted to the updates that hit (a,b)).
__device__ __noinline__ double replay_ab(double v, double delta, const int32_t* __restrict__ r1,
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

// Incremental fitness of swapping a1 <-> a2 (species a != b) against
// prev_gamma; fills the swapped gamma rows a (nA) and b (nB) per shell.
template <int KMAX, bool SP_LDS>
__device__ __forceinline__ double eval_swap(const HeaConst& P, const Tables<KMAX>& T,
                                            const int32_t* __restrict__ nbr,
                                            const uint8_t* __restrict__ sp, int a1, int a2,
                                            double (&nA)[HEA_MAX_S][KMAX],
                                            double (&nB)[HEA_MAX_S][KMAX]) {
  const int K = P.K;
  const int a = sp[a1], b = sp[a2];
  const double conc_a = T.conc[a], conc_b = T.conc[b];
  double err = 0.0;
#pragma unroll
  for (int s = 0; s < HEA_MAX_S; ++s) {
    if (s < P.S) {
      const int z = P.Z[s];
      const double delta = P.delta[s];
      const int32_t* r1 = nbr + (size_t)a1 * P.stride + P.off[s];
      const int32_t* r2 = nbr + (size_t)a2 * P.stride + P.off[s];
      using PK = Pack<KMAX>;
      typename PK::T MA, MB;
      int c11, c12, c22, c21;
      walk<KMAX>(r1, z, sp, a1, a2, MA, c11, c12);
      walk<KMAX>(r2, z, sp, a2, a1, MB, c22, c21);
      const int maA = PK::fld(MA, a), maB = PK::fld(MA, b);
      const int mbA = PK::fld(MB, a), mbB = PK::fld(MB, b);

      // the three entries with interleaved or doubled updates
      const int eAA = T.idx(s, a, a), eBB = T.idx(s, b, b), eAB = T.idx(s, a, b);
      const int pAA = 2 * (mbA - c21 + c22), mAA = 2 * maA;
      const int pBB = 2 * (maB - c12 + c11), mBB = 2 * mbB;
      const int netAB = (maA + c12 - c11) + (mbB + c21 - c22) - maB - mbA;
      double gAA = T.G[eAA], rAA = T.R[eAA];
      double gBB = T.G[eBB], rBB = T.R[eBB];
      double gAB = T.G[eAB], rAB = T.R[eAB];
      double vAA = fma((double)(pAA - mAA), rAA, gAA);
      if (rAA == 0.0) vAA = replay(gAA, -delta, mAA, pAA);
      double vBB = fma((double)(pBB - mBB), rBB, gBB);
      if (rBB == 0.0) vBB = replay(gBB, delta, pBB, mBB);
      double vAB = fma((double)netAB, rAB, gAB);
      if (rAB == 0.0) vAB = replay_ab(gAB, delta, r1, r2, z, sp, a1, a2, a, b);

      double sA[KMAX], sB[KMAX];
#pragma unroll
      for (int x = 0; x < KMAX; ++x) {
        if (x < K) {
          const int ma = PK::fld(MA, x), mb = PK::fld(MB, x);
          const int eA = T.idx(s, a, x), eB = T.idx(s, b, x);
          const double gA = T.G[eA], rA = T.R[eA];
          const double gB = T.G[eB], rB = T.R[eB];
          double vA = fma((double)(mb - ma), rA, gA);
          double vB = fma((double)(ma - mb), rB, gB);
          const bool gen = (x != a) && (x != b);
          if (gen && rA == 0.0) vA = replay(gA, -delta, ma, mb);
          if (gen && rB == 0.0) vB = replay(gB, delta, ma, mb);
          vA = (x == a) ? vAA : (x == b) ? vAB : vA;
          vB = (x == b) ? vBB : (x == a) ? vAB : vB;
          nA[s][x] = vA;
          nB[s][x] = vB;
          sA[x] = 1.0 - vA / (conc_a * T.conc[x]);
          sB[x] = 1.0 - vB / (conc_b * T.conc[x]);
        }
      }

      // sum over all K*K terms in row-major order (accelerate.cpp:560-566)
      double es = 0.0;
#pragma unroll
      for (int i = 0; i < KMAX; ++i) {
        if (i < K) {
          const bool ia = (i == a), ib = (i == b);
#pragma unroll
          for (int j = 0; j < KMAX; ++j) {
            if (j < K) {
              const bool ja = (j == a), jb = (j == b);
              const double sro = ia ? sA[j] : ib ? sB[j] : ja ? sA[i] : sB[i];
              const double df = sro - P.target[s][i][j];
              const double et = df * df;
              const double e = (ia || ib || ja || jb) ? et : T.E0[T.idx(s, i, j)];
              es = es + e;
            }
          }
        }
      }
      err = err + P.weight[s] * es;
    }
  }
  return sqrt(err);
}

__device__ __forceinline__ bool key_less(unsigned long long f1, int j1, unsigned long long f2,
                                         int j2) {
  return (f1 < f2) || (f1 == f2 && j1 < j2);
}

template <int KMAX, bool SP_LDS>
__global__ __launch_bounds__(HEA_MC_THREADS) void k_local_mc(HeaMcArgs A) {
  extern __shared__ __align__(16) unsigned char smem[];
  const HeaConst& P = *A.pc;
  const int l = blockIdx.x;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int N = P.N, K = P.K, S = P.S;
  const int task = A.task;

  constexpr int TAB = HEA_MAX_S * KMAX * KMAX;
  Tables<KMAX> T;
  T.G = reinterpret_cast<double*>(smem);
  T.E0 = T.G + TAB;
  T.R = T.E0 + TAB;
  T.conc = T.R + TAB;
  int2* prop = reinterpret_cast<int2*>(T.conc + KMAX);
  McShared* sh = reinterpret_cast<McShared*>(prop + task);
  uint8_t* lat_g = A.lattices + (size_t)l * N;
  uint8_t* sp;
  if constexpr (SP_LDS) {
    sp = reinterpret_cast<uint8_t*>(sh + 1);
    for (int i = tid; i < N; i += HEA_MC_THREADS) sp[i] = lat_g[i];
  } else {
    sp = lat_g;
  }

  // prev_gamma = R_s(pair counts) (accelerate.cpp:636-640)
  const int32_t* cnt = A.counts + (size_t)l * S * K * K;
  for (int e = tid; e < S * K * K; e += HEA_MC_THREADS) {
    const int s = e / (K * K), i = (e / K) % K, j = e % K;
    T.G[T.idx(s, i, j)] = hea_r_of_count(P, s, cnt[e]);
  }
  if (tid < K) T.conc[tid] = P.conc[tid];
  __syncthreads();
  for (int e = tid; e < S * K * K; e += HEA_MC_THREADS) {
    const int s = e / (K * K), i = (e / K) % K, j = e % K;
    refresh_entry(P, T, s, i, j);
  }
  __syncthreads();

  double prev_fit = A.fitness[l];
  const uint32_t iter = *A.iter_ptr;
  uint32_t pos = 0;
  int acc = 0;
  long long rounds = 0, accepts = 0;
  bool hit_thr = false;

  // gamma rows a and b after the lane's proposal; with more proposals than
  // lanes only the last one is kept and the owner re-evaluates on accept.
  double nA[HEA_MAX_S][KMAX], nB[HEA_MAX_S][KMAX];

  while (acc < A.depth) {
    // ---- draw `task` valid proposals in stream order (accelerate.cpp:649-655)
    int found = 0;
    while (found < task) {
      const uint32_t p = pos + tid;
      uint32_t s1, s2;
      hea_stream_pair(A.seed, iter, A.phase, (uint32_t)l, p, (uint32_t)N, &s1, &s2);
      const bool valid = sp[s1] != sp[s2];
      const unsigned long long bal = __ballot(valid);
      const int pre = __popcll(bal & ((1ull << lane) - 1ull));
      if (lane == 0) sh->wcount[wave] = __popcll(bal);
      __syncthreads();
      int off = 0, tot = 0;
#pragma unroll
      for (int w = 0; w < kWaves; ++w) {
        const int c = sh->wcount[w];
        off += (w < wave) ? c : 0;
        tot += c;
      }
      const int rank = found + off + pre;
      if (valid && rank < task) prop[rank] = make_int2((int)s1, (int)s2);
      if (valid && rank == task - 1) sh->next_pos = p + 1;
      found += tot;
      pos += HEA_MC_THREADS;
      __syncthreads();
    }
    pos = sh->next_pos;

    // ---- evaluate; keep the first strict minimum (accelerate.cpp:663-674)
    unsigned long long bkey = ~0ull;
    int bj = 0x7fffffff;
    for (int j = tid; j < task; j += HEA_MC_THREADS) {
      const int2 pr = prop[j];
      const double f = eval_swap<KMAX, SP_LDS>(P, T, A.nbr, sp, pr.x, pr.y, nA, nB);
      const unsigned long long fk = __double_as_longlong(f);
      if (fk < bkey) {
        bkey = fk;
        bj = j;
      }
    }
    // workgroup argmin of (fitness bits, proposal index); fitness >= 0 so
    // its bit pattern orders like the value
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
      sh->wkey[wave] = rk;
      sh->wj[wave] = rj;
    }
    __syncthreads();
    unsigned long long wk = sh->wkey[0];
    int wj = sh->wj[0];
#pragma unroll
    for (int w = 1; w < kWaves; ++w)
      if (key_less(sh->wkey[w], sh->wj[w], wk, wj)) {
        wk = sh->wkey[w];
        wj = sh->wj[w];
      }
    const double best_fit = __longlong_as_double((long long)wk);
    ++rounds;

    // ---- accept iff strictly better (accelerate.cpp:678-689)
    if (best_fit < prev_fit) {
      if (bj == wj) {  // the owner of the winning proposal writes it back
        const int2 pr = prop[wj];
        if (task > HEA_MC_THREADS)
          (void)eval_swap<KMAX, SP_LDS>(P, T, A.nbr, sp, pr.x, pr.y, nA, nB);
        const int a = sp[pr.x], b = sp[pr.y];
#pragma unroll
        for (int s = 0; s < HEA_MAX_S; ++s) {
          if (s < S) {
#pragma unroll
            for (int x = 0; x < KMAX; ++x) {
              if (x < K) {
                T.G[T.idx(s, a, x)] = nA[s][x];
                T.G[T.idx(s, x, a)] = nA[s][x];
                T.G[T.idx(s, b, x)] = nB[s][x];
                T.G[T.idx(s, x, b)] = nB[s][x];
              }
            }
          }
        }
        sp[pr.x] = (uint8_t)b;
        sp[pr.y] = (uint8_t)a;
      }
      prev_fit = best_fit;
      acc = 0;
      ++accepts;
      if (best_fit <= A.thr) hit_thr = true;
      __syncthreads();
      for (int e = tid; e < S * K * K; e += HEA_MC_THREADS) {
        const int s = e / (K * K), i = (e / K) % K, j = e % K;
        refresh_entry(P, T, s, i, j);
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
    for (int i = tid; i < N; i += HEA_MC_THREADS) lat_g[i] = sp[i];
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

}  // namespace

// Shared-memory bytes of one local-search workgroup.
size_t hea_mc_lds_bytes(int kmax, int task, int n_sites, bool sp_lds) {
  size_t b = sizeof(double) * (3 * HEA_MAX_S * kmax * kmax + kmax);
  b += sizeof(int2) * task + sizeof(McShared);
  if (sp_lds) b += ((size_t)n_sites + 15) & ~(size_t)15;
  return b;
}

int hea_mc_kmax(int K) {
  return K <= 2 ? 2 : K <= 3 ? 3 : K <= 4 ? 4 : K <= 5 ? 5 : K <= 6 ? 6 : K <= 8 ? 8 : 0;
}

template <int KMAX, bool SP_LDS>
static const void* mc_fn() {
  return reinterpret_cast<const void*>(k_local_mc<KMAX, SP_LDS>);
}

static const void* mc_fn_for(int K, bool sp_lds) {
  switch (hea_mc_kmax(K)) {
    case 2: return sp_lds ? mc_fn<2, true>() : mc_fn<2, false>();
    case 3: return sp_lds ? mc_fn<3, true>() : mc_fn<3, false>();
    case 4: return sp_lds ? mc_fn<4, true>() : mc_fn<4, false>();
    case 5: return sp_lds ? mc_fn<5, true>() : mc_fn<5, false>();
    case 6: return sp_lds ? mc_fn<6, true>() : mc_fn<6, false>();
    case 8: return sp_lds ? mc_fn<8, true>() : mc_fn<8, false>();
    default: return nullptr;
  }
}

// Raises the kernel's dynamic-LDS limit; call before launching (and before
// any graph capture that contains the launch).
hipError_t hea_prepare_local_mc(int K, bool sp_lds, size_t lds) {
  const void* fn = mc_fn_for(K, sp_lds);
  if (!fn) return hipErrorInvalidValue;
  return hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
}

hipError_t hea_launch_local_mc(const HeaMcArgs& a, int K, bool sp_lds, size_t lds,
                               hipStream_t st) {
  const void* fn = mc_fn_for(K, sp_lds);
  if (!fn) return hipErrorInvalidValue;
  void* args[] = {const_cast<HeaMcArgs*>(&a)};
  return hipLaunchKernel(fn, dim3(a.L), dim3(HEA_MC_THREADS), args, lds, st);
}
