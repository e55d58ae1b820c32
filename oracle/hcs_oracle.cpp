// TEST INFRASTRUCTURE ONLY: the CPU oracle of the SRO lattice search.
//
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
// load this library, and only as the checker.  Product code never links it.
//
// It restates, function by function, the reference CPU engine
// pyhea/cpp/accelerate.cpp (caimeiniu/pyhea 1.0.0) with two deliberate
// substitutions that make a run replayable:
//   * proposals come from the keyed stream (seed, iter, phase, slot) instead
//     of the global std::rand() (accelerate.cpp:649-654);
//   * permutations come from the keyed Feistel permutation instead of
//     std::shuffle(mt19937(random_device())) (accelerate.cpp:240-248).
// Everything else -- the FP64 operation order of the gamma recount, the
// +-delta incremental update, the fitness sum, the strict-< first-minimum
// choice and the accept rule -- follows the reference statement by statement,
// so that the same stream gives bit-identical doubles.  Pair counts are kept
// as integers beside gamma to check the bookkeeping invariants.
//
// The stream generator here is written independently of the product's
// (pyhea_amd/csrc/hea_stream.h) from the published Philox4x32-10 definition
// (Salmon et al., SC'11) and is pinned by the Random123 known-answer vectors.
//
// Parity status per function (how each one is pinned) is listed in the
// header of oracle/oracle.py.
//
// Build: g++ -O2 -std=c++17 -ffp-contract=off -fPIC -shared (no -ffast-math,
// no -march: the reference is x86-64 SSE2 doubles without FMA).

#include <stdint.h>
#include <math.h>
#include <string.h>
#include <algorithm>
#include <numeric>
#include <vector>

#include "oracle_stream.h"

namespace {

using namespace hea_oracle;

// --------------------------------------------------------------- problem ---
struct Problem {
  int N, K, S, sumZ;
  int Z[3], off[3];
  const int32_t* nbr;     // [N][sumZ]
  const int32_t* counts;  // [K]
  const double* weight;   // [S]
  const double* target;   // [S][K*K]
  int nb(int site, int s, int k) const { return nbr[(size_t)site * sumZ + off[s] + k]; }
};

// gamma stored like the reference: gamma[a][b][s] at ((a*K)+b)*S+s
struct Gamma {
  std::vector<double> g;
  int K, S;
  double& at(int a, int b, int s) { return g[((size_t)a * K + b) * S + s]; }
  double at(int a, int b, int s) const { return g[((size_t)a * K + b) * S + s]; }
};

// accelerate.cpp:370-399 (calculate_sro_coefficient): every pair adds
// 1/(double(Z)*N) to its entry, in site, shell, neighbour order.
void recount_gamma(const Problem& P, const uint8_t* lat, Gamma& G) {
  G.K = P.K;
  G.S = P.S;
  G.g.assign((size_t)P.K * P.K * P.S, 0.0);
  for (int i = 0; i < P.N; ++i) {
    const int a = lat[i];
    for (int s = 0; s < P.S; ++s) {
      const int z = P.Z[s];
      for (int k = 0; k < z; ++k) {
        const int b = lat[P.nb(i, s, k)];
        G.at(a, b, s) += 1 / (static_cast<double>(z) * P.N);
      }
    }
  }
}

// accelerate.cpp:556-569 (fitness tail of calculate_fitness_incremental) and
// accelerate.cpp:166-178 (calculate_fitness_of_solutions, whose
// std::pow(x, 2) the compiler folds to x*x; pinned by the golden vectors).
double error_of_gamma(const Problem& P, const Gamma& G) {
  double error = 0.0;
  for (int s = 0; s < P.S; ++s) {
    const double w = P.weight[s];
    double es = 0.0;
    for (int i = 0; i < P.K; ++i) {
      for (int j = 0; j < P.K; ++j) {
        const double sro =
            1 - G.at(i, j, s) / ((static_cast<double>(P.counts[i]) / static_cast<double>(P.N)) *
                                 (static_cast<double>(P.counts[j]) / static_cast<double>(P.N)));
        const double t = P.target[(size_t)s * P.K * P.K + i * P.K + j];
        es += (sro - t) * (sro - t);
      }
    }
    error += w * es;
  }
  return error;
}

double fitness_of_gamma(const Problem& P, const Gamma& G) { return sqrt(error_of_gamma(P, G)); }

// accelerate.cpp:498-570 (calculate_fitness_incremental).  Mutates gamma and
// leaves `lat` with a1/a2 swapped.  Note the a2 loop bound is a1's shell
// size (accelerate.cpp:516, 535); all sites of a regular lattice share it.
double incremental(const Problem& P, uint8_t* lat, Gamma& G, int a1, int a2) {
  for (int s = 0; s < P.S; ++s) {
    int atom = lat[a1];
    int atom_after = lat[a2];
    const int z = P.Z[s];
    const double d = 1.0 / (z * P.N);
    for (int k = 0; k < z; ++k) {
      const int nb = P.nb(a1, s, k);
      const int n = lat[nb];
      int na = n;
      if (nb == a2) na = lat[a1];
      else if (nb == a1) na = lat[a2];
      G.at(atom, n, s) -= d;
      G.at(n, atom, s) -= d;
      G.at(atom_after, na, s) += d;
      G.at(na, atom_after, s) += d;
    }
    atom = lat[a2];
    atom_after = lat[a1];
    for (int k = 0; k < z; ++k) {
      const int nb = P.nb(a2, s, k);
      const int n = lat[nb];
      int na = n;
      if (nb == a1) na = lat[a2];
      else if (nb == a2) na = lat[a1];
      G.at(atom, n, s) -= d;
      G.at(n, atom, s) -= d;
      G.at(atom_after, na, s) += d;
      G.at(na, atom_after, s) += d;
    }
  }
  std::swap(lat[a1], lat[a2]);
  return fitness_of_gamma(P, G);
}

// Integer pair counts n[s][a][b] (directed pairs), used for invariants.
void recount_ints(const Problem& P, const uint8_t* lat, int64_t* n) {
  memset(n, 0, sizeof(int64_t) * P.S * P.K * P.K);
  for (int i = 0; i < P.N; ++i)
    for (int s = 0; s < P.S; ++s)
      for (int k = 0; k < P.Z[s]; ++k)
        n[((size_t)s * P.K + lat[i]) * P.K + lat[P.nb(i, s, k)]] += 1;
}

struct McStats {
  int64_t evals, rounds, accepts, draws;
};

// accelerate.cpp:601-691 (local_parallel_monte_carlo) for one lattice.
// `lat` is updated in place; returns the final stored fitness.
double local_mc_one(const Problem& P, uint8_t* lat, double fit, int task,
                    int depth, Stream& rs, McStats& st, int32_t* accept_log,
                    int64_t accept_cap) {
  const int N = P.N;
  Gamma prev_gamma, gamma, best_gamma;
  recount_gamma(P, lat, prev_gamma);  // accelerate.cpp:636-640
  gamma = prev_gamma;
  best_gamma = prev_gamma;
  std::fill(best_gamma.g.begin(), best_gamma.g.end(), 0.0);
  double prev_fit = fit;
  int accumulate = 0;
  std::vector<uint8_t> scratch(lat, lat + N);
  while (accumulate < depth) {
    double best_fit = 0;
    int best_a1 = -1, best_a2 = -1;
    for (int jj = 0; jj < task; ++jj) {
      // accelerate.cpp:649-655: both sites redrawn while the species agree
      int a1 = (int)(rs.next() % (uint32_t)N);
      int a2 = (int)(rs.next() % (uint32_t)N);
      st.draws += 1;
      while (lat[a1] == lat[a2]) {
        a1 = (int)(rs.next() % (uint32_t)N);
        a2 = (int)(rs.next() % (uint32_t)N);
        st.draws += 1;
      }
      gamma = prev_gamma;  // accelerate.cpp:657
      const double f = incremental(P, lat, gamma, a1, a2);
      std::swap(lat[a1], lat[a2]);  // back to prev (the reference copies)
      st.evals += 1;
      if (jj == 0 || f < best_fit) {  // accelerate.cpp:663-674
        best_fit = f;
        best_a1 = a1;
        best_a2 = a2;
        best_gamma = gamma;
      }
    }
    st.rounds += 1;
    if (best_fit < prev_fit) {  // accelerate.cpp:678-684
      if (best_a1 >= 0) std::swap(lat[best_a1], lat[best_a2]);
      prev_fit = best_fit;
      prev_gamma = best_gamma;
      accumulate = 0;
      if (accept_log && st.accepts < accept_cap) {
        accept_log[2 * st.accepts] = best_a1;
        accept_log[2 * st.accepts + 1] = best_a2;
      }
      st.accepts += 1;
    } else {
      accumulate += 1;
    }
  }
  return prev_fit;
}

// The reference's population order (accelerate.cpp:37-64): std::sort of the
// index vector 0..n-1 by fitness.  std::sort is not stable, so on equal keys
// the order is whatever the standard library's introsort leaves; calling the
// same library routine on the same int indices with the same comparator
// reproduces it (libstdc++ of GCC 11.4 here and in the reference build).
// Reports whether two keys were equal.
int sort_order(const double* f, int n, std::vector<int>& idx) {
  idx.resize(n);
  std::iota(idx.begin(), idx.end(), 0);
  std::sort(idx.begin(), idx.end(), [&](int a, int b) { return f[a] < f[b]; });
  int ties = 0;
  for (int i = 1; i < n; ++i)
    if (f[idx[i]] == f[idx[i - 1]]) ties = 1;
  return ties;
}

void reorder(uint8_t* lats, double* fit, int L, int N, const std::vector<int>& idx) {
  std::vector<uint8_t> tl((size_t)L * N);
  std::vector<double> tf(L);
  for (int r = 0; r < L; ++r) {
    memcpy(&tl[(size_t)r * N], lats + (size_t)idx[r] * N, N);
    tf[r] = fit[idx[r]];
  }
  memcpy(lats, tl.data(), (size_t)L * N);
  memcpy(fit, tf.data(), sizeof(double) * L);
}

Problem make_problem(int N, int K, int S, const int32_t* Z, const int32_t* nbr,
                     const int32_t* counts, const double* weight,
                     const double* target) {
  Problem P;
  P.N = N;
  P.K = K;
  P.S = S;
  P.sumZ = 0;
  for (int s = 0; s < 3; ++s) {
    P.Z[s] = s < S ? Z[s] : 0;
    P.off[s] = P.sumZ;
    P.sumZ += P.Z[s];
  }
  P.nbr = nbr;
  P.counts = counts;
  P.weight = weight;
  P.target = target;
  return P;
}

}  // namespace

extern "C" {

// ------------------------------------------------------------ ABI (tests) --
void oracle_philox(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  U4 c = {{ctr[0], ctr[1], ctr[2], ctr[3]}};
  U4 o = philox(c, key[0], key[1]);
  for (int i = 0; i < 4; ++i) out[i] = o.v[i];
}

void oracle_stream_words(uint64_t seed, uint32_t iter, uint32_t phase,
                         uint32_t slot, uint64_t start, int64_t n, uint32_t* out) {
  Stream s{seed, iter, phase, slot, start};
  for (int64_t i = 0; i < n; ++i) out[i] = s.next();
}

void oracle_permute(uint64_t seed, uint32_t iter, uint32_t phase, uint32_t slot,
                    const int32_t* counts, int K, int N, uint8_t* out) {
  permute_lattice(seed, iter, phase, slot, counts, K, N, out);
}

// gamma out layout [K][K][S] (reference layout).
void oracle_gamma(int N, int K, int S, const int32_t* Z, const int32_t* nbr,
                  const uint8_t* lat, double* gamma_out) {
  Problem P = make_problem(N, K, S, Z, nbr, nullptr, nullptr, nullptr);
  Gamma G;
  recount_gamma(P, lat, G);
  memcpy(gamma_out, G.g.data(), sizeof(double) * G.g.size());
}

void oracle_pair_counts(int N, int K, int S, const int32_t* Z, const int32_t* nbr,
                        const uint8_t* lat, int64_t* out) {
  Problem P = make_problem(N, K, S, Z, nbr, nullptr, nullptr, nullptr);
  recount_ints(P, lat, out);
}

// accelerate.cpp:118-180 for L lattices.
void oracle_fitness(int N, int K, int S, const int32_t* Z, const int32_t* nbr,
                    const int32_t* counts, const double* weight, const double* target,
                    int L, const uint8_t* lats, double* fit) {
  Problem P = make_problem(N, K, S, Z, nbr, counts, weight, target);
  Gamma G;
  for (int l = 0; l < L; ++l) {
    recount_gamma(P, lats + (size_t)l * N, G);
    fit[l] = fitness_of_gamma(P, G);
  }
}

// Applies a chain of swaps (a1[i], a2[i]) with gamma carried forward, as an
// accepted proposal carries best_gamma into prev_gamma
// (accelerate.cpp:682).  Returns each step's fitness; gamma_out is the final
// drifted gamma ([K][K][S]); lat is left with every swap applied.
void oracle_incremental_chain(int N, int K, int S, const int32_t* Z, const int32_t* nbr,
                              const int32_t* counts, const double* weight,
                              const double* target, uint8_t* lat, int n_swaps,
                              const int32_t* a1, const int32_t* a2, double* fits,
                              double* gamma_out) {
  Problem P = make_problem(N, K, S, Z, nbr, counts, weight, target);
  Gamma G;
  recount_gamma(P, lat, G);
  for (int i = 0; i < n_swaps; ++i) fits[i] = incremental(P, lat, G, a1[i], a2[i]);
  memcpy(gamma_out, G.g.data(), sizeof(double) * G.g.size());
}

// accelerate.cpp:601-694 on a population of L lattices (slot = position),
// stream (seed, iter, phase, slot).  Ends with the population sort
// (accelerate.cpp:693).  stats: [evals, rounds, accepts, draws, ties];
// per_slot (optional) [L][4]: rounds, accepts, draws, evals before the sort.
int oracle_local_mc(int N, int K, int S, const int32_t* Z, const int32_t* nbr,
                    const int32_t* counts, const double* weight, const double* target,
                    int L, uint8_t* lats, double* fit, int task, int depth,
                    uint64_t seed, uint32_t iter, uint32_t phase, int do_sort,
                    int64_t* stats, int64_t* per_slot) {
  Problem P = make_problem(N, K, S, Z, nbr, counts, weight, target);
  McStats tot = {0, 0, 0, 0};
  for (int l = 0; l < L; ++l) {
    Stream rs{seed, iter, phase, (uint32_t)l, 0};
    McStats st = {0, 0, 0, 0};
    fit[l] = local_mc_one(P, lats + (size_t)l * N, fit[l], task, depth, rs, st,
                          nullptr, 0);
    if (per_slot) {
      per_slot[4 * l + 0] = st.rounds;
      per_slot[4 * l + 1] = st.accepts;
      per_slot[4 * l + 2] = st.draws;
      per_slot[4 * l + 3] = st.evals;
    }
    tot.evals += st.evals;
    tot.rounds += st.rounds;
    tot.accepts += st.accepts;
    tot.draws += st.draws;
  }
  int ties = 0;
  if (do_sort) {
    std::vector<int> idx;
    ties = sort_order(fit, L, idx);
    reorder(lats, fit, L, N, idx);
  }
  if (stats) {
    stats[0] = tot.evals;
    stats[1] = tot.rounds;
    stats[2] = tot.accepts;
    stats[3] = tot.draws;
    stats[4] = ties;
  }
  return 0;
}

// One lattice, with the ordered list of accepted swaps (for trajectory tests).
double oracle_local_mc_trace(int N, int K, int S, const int32_t* Z, const int32_t* nbr,
                             const int32_t* counts, const double* weight,
                             const double* target, uint8_t* lat, double fit, int task,
                             int depth, uint64_t seed, uint32_t iter, uint32_t phase,
                             uint32_t slot, int32_t* accept_log, int64_t accept_cap,
                             int64_t* stats) {
  Problem P = make_problem(N, K, S, Z, nbr, counts, weight, target);
  Stream rs{seed, iter, phase, slot, 0};
  McStats st = {0, 0, 0, 0};
  double f = local_mc_one(P, lat, fit, task, depth, rs, st, accept_log, accept_cap);
  if (stats) {
    stats[0] = st.evals;
    stats[1] = st.rounds;
    stats[2] = st.accepts;
    stats[3] = st.draws;
  }
  return f;
}

// accelerate.cpp:713-763 (run_local_parallel_hcs) on one rank with the keyed
// streams.  best_trace[it] = population best after iteration it.
// stats: [evals, rounds, accepts, draws, ties].
int oracle_run_hcs(int N, int K, int S, const int32_t* Z, const int32_t* nbr,
                   const int32_t* counts, const double* weight, const double* target,
                   int L, int total_iter, int task, int depth, uint64_t seed,
                   uint8_t* out_lats, double* out_fit, double* best_trace,
                   int64_t* stats) {
  Problem P = make_problem(N, K, S, Z, nbr, counts, weight, target);
  std::vector<uint8_t> lat((size_t)L * N), nl((size_t)L * N);
  std::vector<double> fit(L), nf(L);
  int64_t st5[5], tot[5] = {0, 0, 0, 0, 0};
  Gamma G;
  for (int l = 0; l < L; ++l) {  // accelerate.cpp:737
    permute_lattice(seed, 0, 0, l, counts, K, N, &lat[(size_t)l * N]);
    recount_gamma(P, &lat[(size_t)l * N], G);
    fit[l] = fitness_of_gamma(P, G);
  }
  for (int it = 0; it < total_iter; ++it) {
    for (int l = 0; l < L; ++l) {  // accelerate.cpp:743
      permute_lattice(seed, it, 1, l, counts, K, N, &nl[(size_t)l * N]);
      recount_gamma(P, &nl[(size_t)l * N], G);
      nf[l] = fitness_of_gamma(P, G);
    }
    oracle_local_mc(N, K, S, Z, nbr, counts, weight, target, L, nl.data(), nf.data(),
                    task, depth, seed, it, 2, 1, st5, nullptr);  // accelerate.cpp:745
    for (int q = 0; q < 5; ++q) tot[q] += st5[q];
    {  // accelerate.cpp:747 / 305-330: old ++ new, sorted, first L kept
      std::vector<uint8_t> cl((size_t)2 * L * N);
      std::vector<double> cf(2 * L);
      memcpy(cl.data(), lat.data(), (size_t)L * N);
      memcpy(cl.data() + (size_t)L * N, nl.data(), (size_t)L * N);
      memcpy(cf.data(), fit.data(), sizeof(double) * L);
      memcpy(cf.data() + L, nf.data(), sizeof(double) * L);
      std::vector<int> idx;
      tot[4] |= sort_order(cf.data(), 2 * L, idx);
      for (int r = 0; r < L; ++r) {
        memcpy(&lat[(size_t)r * N], &cl[(size_t)idx[r] * N], N);
        fit[r] = cf[idx[r]];
      }
    }
    oracle_local_mc(N, K, S, Z, nbr, counts, weight, target, L, lat.data(), fit.data(),
                    task, depth, seed, it, 3, 1, st5, nullptr);  // accelerate.cpp:749
    for (int q = 0; q < 5; ++q) tot[q] += st5[q];
    if (best_trace) best_trace[it] = fit[0];  // accelerate.cpp:752
  }
  memcpy(out_lats, lat.data(), (size_t)L * N);
  memcpy(out_fit, fit.data(), sizeof(double) * L);
  if (stats)
    for (int q = 0; q < 5; ++q) stats[q] = tot[q];
  return 0;
}

}  // extern "C"
