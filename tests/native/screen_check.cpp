// Test helper: the local search's screen window (pyhea_amd/csrc/hea_screen.h,
// the same templates the kernels instantiate) against the oracle's exact
// arithmetic (oracle/hcs_oracle.cpp, the restatement of the reference's
// calculate_fitness_incremental, accelerate.cpp:498-570).
//
//   screen_check <problem.bin> <lattices> <drift> <seed> <beta_scale>
//   screen_check --steps <trials> <seed>
//
// For each of <lattices> random lattices of the problem, <drift> random
// swaps are applied with gamma carried forward (as accepted proposals carry
// their drifted gamma, accelerate.cpp:682), so the tables hold drifted
// doubles.  Then, for every unlike-species pair (a1, a2) of the lattice:
//   * exact_entries' new gamma entries (a,x) and (b,x) of every shell (the
//     one-fma fast path where the safe step allows it, else the replay)
//     must equal the oracle's sequentially updated gamma bit for bit;
//   * |err(P) - err(G) - X(P)| <= 2 beta, with err the reference's fitness
//     squared before the square root, X the screen's estimate and beta the
//     window half-width (times <beta_scale>): the bound behind the
//     window-decided accepts and rejects;
//   * the first strict minimum of F over all pairs (the reference's choice
//     among them) lies in the window X <= X_min + 2 beta.
// Prints "ok <pairs> <entries> <replayed> <max ratio>" (ratio = largest
// |err(P) - err(G) - X(P)| / (2 beta)) or the first failure, exit 1.
//
// The problem file: int32 N, K, S, Z[3]; int32 nbr [N][sum Z] (each shell's
// list ascending); int32 counts [K]; double weight [S]; double target
// [S][K*K].
#include <stdio.h>
#include <stdlib.h>

#include <algorithm>
#include <random>
#include <string>
#include <vector>

#include "hcs_oracle.cpp"  // the oracle (anonymous namespace: Problem, Gamma, incremental, ...)
#include "hea_screen.h"

namespace {

constexpr int KM = 8;  // table width (every K of the problem files)

struct D2 {
  double x, y;
};

struct Shared {
  unsigned long long bmax[HEA_MAX_S], ubmax[HEA_MAX_S];
};

// The kernels' argument fields the templates read.
struct Args {
  int N, K, S, stride;
  int Z[HEA_MAX_S], off[HEA_MAX_S];
  double delta[HEA_MAX_S], weight[HEA_MAX_S];
  const int32_t* nbr;  // [N][stride], shells 4-aligned and padded with -1
};

// The kernels' LDS tables, on the host.
struct Tab {
  double *G, *E0, *R, *tgt, *conc;
  D2* SC;
  float* H;
  Shared* sh;
  int idx(int s, int i, int j) const { return (s * KM + i) * KM + j; }
  int hidx(int s, int i, int j) const { return (s * KM + i) * 8 + j; }
};

using PK = Pack<KM>;

bool read_all(FILE* f, void* p, size_t n) { return fread(p, 1, n, f) == n; }

}  // namespace

// --steps mode: the safe step of a gamma entry.  Whenever safe_step(g,
// delta, z) returns r != 0, every sequence of +-delta additions from g whose
// partial sums stay within 2z steps of g ends at fma(net, r, g), the one-fma
// fast path of exact_entries.  Inputs: random g and delta = 1/(Z N); values
// just below and above powers of two (binade edges); and rounding ties,
// delta an odd multiple of half an ulp of g (the first step rounds to
// even).  Prints "ok <cases> <fast> <naive misses>": naive misses count the
// sequences where the step without the tie and edge guards, fl(g + delta) -
// g, would have given a wrong result (so the adversarial inputs matter).
int steps_mode(long trials, uint64_t seed) {
  std::mt19937_64 rng(seed);
  std::uniform_real_distribution<double> U(0.0, 1.0);
  long fast = 0, naive_miss = 0, cases = 0;
  for (long t = 0; t < trials; ++t) {
    const int z = 1 + (int)(rng() % 31);
    double g, delta;
    const int kind = (int)(t % 3);
    if (kind == 0) {  // random entry value and step
      const int n = 8 + (int)(rng() % 600000);
      delta = 1.0 / (double)(z * n);
      g = delta * (1.0 + U(rng) * (double)(z * n) * 0.9);
    } else if (kind == 1) {  // near a binade edge
      const int p = -(int)(rng() % 12);
      const int n = 8 + (int)(rng() % 6000);
      delta = 1.0 / (double)(z * n);
      g = ldexp(1.0, p) + (double)((long)(rng() % 64) - 32) * delta * U(rng);
    } else {  // a rounding tie on the first step: delta = (2k+1) ulp(g) / 2
      const int e = -30 - (int)(rng() % 20);
      const double m = ldexp(1.0, 52) + (double)(rng() % (1ull << 52));
      g = ldexp(m, e);
      const double u = ldexp(1.0, e);
      delta = (2.0 * (double)(rng() % (1u << 16)) + 1.0) * 0.5 * u;
    }
    if (!(g > delta)) continue;
    ++cases;
    const double r = safe_step(g, delta, z);
    const double rn = (g + delta) - g;  // the step without the guards
    for (int q = 0; q < 8; ++q) {
      // a sequence with partial sums in [-2z, 2z]: n1 steps of -delta then n2
      // of +delta (replay), or an interleaving (replay_ab)
      double v = g;
      int net = 0;
      const int len = 1 + (int)(rng() % (4 * z));
      for (int i = 0; i < len; ++i) {
        int st = (q < 4) ? (i < len / 2 ? -1 : 1) : ((rng() & 1) ? 1 : -1);
        if (net + st > 2 * z || net + st < -2 * z) st = -st;
        v = st > 0 ? v + delta : v - delta;
        net += st;
      }
      const double want = v;
      if (fma((double)net, rn, g) != want) ++naive_miss;
      if (r != 0.0) {
        ++fast;
        const double got = fma((double)net, r, g);
        if (__builtin_bit_cast(uint64_t, got) != __builtin_bit_cast(uint64_t, want)) {
          printf("step: g %.17g delta %.17g z %d net %d: fma %.17g, sequential %.17g\n", g, delta,
                 z, net, got, want);
          return 1;
        }
      }
    }
  }
  printf("ok %ld %ld %ld\n", cases, fast, naive_miss);
  return 0;
}

int main(int argc, char** argv) {
  if (argc == 4 && std::string(argv[1]) == "--steps")
    return steps_mode(atol(argv[2]), strtoull(argv[3], nullptr, 10));
  if (argc != 6) {
    fprintf(stderr, "usage: screen_check <problem.bin> <lattices> <drift> <seed> <beta_scale>\n");
    return 2;
  }
  const int n_lat = atoi(argv[2]), n_drift = atoi(argv[3]);
  std::mt19937_64 rng(strtoull(argv[4], nullptr, 10));
  const double beta_scale = atof(argv[5]);
  FILE* f = fopen(argv[1], "rb");
  if (!f) return 2;
  int32_t hdr[6];
  if (!read_all(f, hdr, sizeof hdr)) return 2;
  const int N = hdr[0], K = hdr[1], S = hdr[2];
  const int32_t Z[3] = {hdr[3], hdr[4], hdr[5]};
  int sumZ = 0;
  for (int s = 0; s < S; ++s) sumZ += Z[s];
  std::vector<int32_t> nbr((size_t)N * sumZ), counts(K);
  std::vector<double> weight(S), target((size_t)S * K * K);
  if (!read_all(f, nbr.data(), 4 * nbr.size()) || !read_all(f, counts.data(), 4 * K) ||
      !read_all(f, weight.data(), 8 * S) || !read_all(f, target.data(), 8 * target.size()))
    return 2;
  fclose(f);
  if (K > KM || S > HEA_MAX_S) return 2;

  const Problem P = make_problem(N, K, S, Z, nbr.data(), counts.data(), weight.data(),
                                 target.data());
  // the kernels' padded neighbour table and launch scalars
  Args A{};
  A.N = N;
  A.K = K;
  A.S = S;
  int off = 0;
  for (int s = 0; s < S; ++s) {
    A.Z[s] = Z[s];
    A.off[s] = off;
    off += (Z[s] + 3) & ~3;
    A.delta[s] = 1.0 / (double)(Z[s] * N);
    A.weight[s] = weight[s];
  }
  A.stride = off;
  std::vector<int32_t> padded((size_t)N * A.stride, -1);
  for (int i = 0; i < N; ++i)
    for (int s = 0; s < S; ++s)
      for (int k = 0; k < Z[s]; ++k)
        padded[(size_t)i * A.stride + A.off[s] + k] = P.nb(i, s, k);
  A.nbr = padded.data();

  const int TAB = HEA_MAX_S * KM * KM;
  std::vector<double> G(TAB), E0(TAB), R(TAB), tgt(TAB), conc(KM);
  std::vector<D2> SC(TAB);
  std::vector<float> H(HEA_MAX_S * KM * 8);
  Shared sh{};
  Tab T{G.data(), E0.data(), R.data(), tgt.data(), conc.data(), SC.data(), H.data(), &sh};
  for (int a = 0; a < K; ++a) conc[a] = (double)counts[a] / (double)N;

  long long pairs = 0, entries = 0, replayed = 0;
  double max_ratio = 0.0;
  std::vector<uint8_t> lat(N), lat2(N);
  for (int li = 0; li < n_lat; ++li) {
    // random lattice of the composition, then a drift chain
    int k = 0;
    for (int a = 0; a < K; ++a)
      for (int c = 0; c < counts[a]; ++c) lat[k++] = (uint8_t)a;
    std::shuffle(lat.begin(), lat.end(), rng);
    Gamma g;
    recount_gamma(P, lat.data(), g);
    for (int d = 0; d < n_drift; ++d) {
      int i = (int)(rng() % N), j = (int)(rng() % N);
      while (lat[i] == lat[j]) j = (int)(rng() % N);
      incremental(P, lat.data(), g, i, j);
    }
    // the kernel's tables of this gamma (hea_mc.hip, start of a local search)
    for (int s = 0; s < S; ++s)
      for (int i = 0; i < K; ++i)
        for (int j = 0; j < K; ++j) {
          const int e = T.idx(s, i, j);
          G[e] = g.at(i, j, s);
          tgt[e] = target[(size_t)s * K * K + i * K + j];
          SC[e].y = -A.delta[s] / (conc[i] * conc[j]);
        }
    for (int s = 0; s < HEA_MAX_S; ++s) sh.bmax[s] = sh.ubmax[s] = 0;
    for (int s = 0; s < S; ++s)
      for (int i = 0; i < K; ++i)
        for (int j = 0; j < K; ++j) {
          const EntryBounds r = refresh_entry_core<KM, false>(A, T, s, i, j);
          sh.bmax[s] = std::max(sh.bmax[s], (unsigned long long)__builtin_bit_cast(uint64_t, r.P));
          sh.ubmax[s] = std::max(sh.ubmax[s], (unsigned long long)__builtin_bit_cast(uint64_t, r.ub));
          replayed += (R[T.idx(s, i, j)] == 0.0);
        }
    const double beta = screen_beta(A, T) * beta_scale;
    const double errG = error_of_gamma(P, g);

    double xmin = INFINITY, best_f = INFINITY, x_best = 0.0;
    for (int a1 = 0; a1 < N; ++a1)
      for (int a2 = 0; a2 < N; ++a2) {
        const int a = lat[a1], b = lat[a2];
        if (a == b) continue;
        ++pairs;
        // the walk of both lists, as the screen lanes make it
        double X = 0.0;
        PK::T MA[HEA_MAX_S], MB[HEA_MAX_S];
        int cc[HEA_MAX_S][4];
        for (int s = 0; s < S; ++s) {
          MA[s] = MB[s] = 0;
          int c11 = 0, c12 = 0, c21 = 0, c22 = 0;
          for (int q = 0; q < Z[s]; ++q) {
            const int n1 = P.nb(a1, s, q), n2 = P.nb(a2, s, q);
            MA[s] += PK::one(lat[n1]);
            MB[s] += PK::one(lat[n2]);
            c11 += n1 == a1;
            c12 += n1 == a2;
            c21 += n2 == a1;
            c22 += n2 == a2;
          }
          cc[s][0] = c11;
          cc[s][1] = c12;
          cc[s][2] = c21;
          cc[s][3] = c22;
          X = X + A.weight[s] * screen_shell<KM, false>(T, K, s, a, b, MA[s], MB[s], c11, c12,
                                                        c21, c22);
        }
        // the reference: sequential +-delta updates of a copy
        Gamma g2 = g;
        lat2 = lat;
        const double F = incremental(P, lat2.data(), g2, a1, a2);
        const double errP = error_of_gamma(P, g2);
        for (int s = 0; s < S; ++s)
          for (int x = 0; x < K; ++x) {
            double nA, nB;
            exact_entries<KM>(A, T, lat.data(), s, a1, a2, a, b, x, MA[s], MB[s], cc[s][0],
                              cc[s][1], cc[s][2], cc[s][3], nA, nB);
            entries += 2;
            if (__builtin_bit_cast(uint64_t, nA) != __builtin_bit_cast(uint64_t, g2.at(a, x, s)) ||
                __builtin_bit_cast(uint64_t, nB) != __builtin_bit_cast(uint64_t, g2.at(b, x, s))) {
              printf("entry mismatch: lattice %d pair (%d,%d) shell %d x %d: %.17g %.17g vs "
                     "%.17g %.17g\n", li, a1, a2, s, x, nA, nB, g2.at(a, x, s), g2.at(b, x, s));
              return 1;
            }
          }
        const double dev = fabs(errP - errG - X);
        max_ratio = std::max(max_ratio, dev / (2.0 * beta));
        if (!(dev <= 2.0 * beta)) {
          printf("bound: lattice %d pair (%d,%d): |err(P) - err(G) - X| = %.6g > 2 beta = %.6g\n",
                 li, a1, a2, dev, 2.0 * beta);
          return 1;
        }
        xmin = std::min(xmin, X);
        if (F < best_f) {  // first strict minimum in enumeration order
          best_f = F;
          x_best = X;
        }
      }
    if (!(x_best <= xmin + 2.0 * beta)) {
      printf("window: lattice %d: the first strict minimum has X = %.17g > X_min + 2 beta = "
             "%.17g\n", li, x_best, xmin + 2.0 * beta);
      return 1;
    }
  }
  printf("ok %lld %lld %lld %.6g\n", pairs, entries, replayed, max_ratio);
  return 0;
}
