// TEST INFRASTRUCTURE ONLY: drives the reference CPU engine itself.
//
// Built by oracle/Makefile (target `ref`) into oracle/_ref/ref_hcs from the
// reference source where it lies (pyhea/cpp/accelerate.cpp is #included by
// path at build time; neither its text nor the binary is kept in this
// repository's history: oracle/_ref/ is git-ignored).  The binary is a
// checker and the "reference" CPU baseline; product code never runs it.
//
// Two inputs of the reference are not reproducible, so the harness injects
// them, leaving every reference function untouched:
//   * std::rand() (accelerate.cpp:649-654) is interposed: the executable's
//     own `rand` wins over glibc's and returns the keyed stream of the lattice
//     currently being searched;
//   * generate_random_lattices (accelerate.cpp:218-252) seeds from
//     std::random_device, so the harness writes the keyed permutations itself
//     and then calls the reference's calculate_fitness_of_solutions.
// Local search is called one lattice at a time so that each lattice draws
// from its own stream; the population sort that ends the reference's
// local_parallel_monte_carlo (accelerate.cpp:693) is then applied by calling
// the reference's sort_lattices_by_fitness on the whole population.  The
// per-lattice state the reference shares across lattices (gamma, prev_gamma,
// best_gamma) is re-initialised for every lattice before use, so the split
// changes nothing.
//
// Usage: ref_hcs <input.bin> <output.bin>; the formats are written and read
// by oracle/oracle.py (RefHarness).

#include "accelerate.cpp"  // the reference engine, from -I<reference>/pyhea/cpp

#include <stdio.h>
#include <stdlib.h>
#include <chrono>

#include "oracle_stream.h"

static hea_oracle::Stream g_stream = {0, 0, 0, 0, 0};
static long long g_draw_words = 0;

extern "C" int rand(void) noexcept {
  ++g_draw_words;
  return (int)g_stream.next();
}

namespace {

using accelerate::cpu::calculate_best_lattices;
using accelerate::cpu::calculate_fitness_incremental;
using accelerate::cpu::calculate_fitness_of_solutions;
using accelerate::cpu::calculate_sro_coefficient;
using accelerate::cpu::local_parallel_monte_carlo;
using accelerate::cpu::sort_lattices_by_fitness;

using Lat = std::vector<std::vector<int>>;
using Nbr = std::vector<std::vector<std::vector<int>>>;

struct Header {
  int32_t magic, mode, N, K, S, Z0, Z1, Z2, L, I, T, D, n_swaps, pad;
  uint64_t seed;
  uint32_t iter, phase;
};

FILE* g_out = nullptr;

void die(const char* m) {
  fprintf(stderr, "ref_hcs: %s\n", m);
  exit(2);
}

template <typename T>
void rd(FILE* f, T* p, size_t n) {
  if (n && fread(p, sizeof(T), n, f) != n) die("short read");
}
template <typename T>
void wr(const T* p, size_t n) {
  if (n && fwrite(p, sizeof(T), n, g_out) != n) die("short write");
}

void put_population(const Lat& lat, const std::vector<double>& fit) {
  for (const auto& l : lat) {
    std::vector<uint8_t> b(l.begin(), l.end());
    wr(b.data(), b.size());
  }
  wr(fit.data(), fit.size());
}

void local_search(Lat& lat, std::vector<double>& fit, int T, int D, const Nbr& nbor,
                  const std::vector<int>& species, const std::vector<std::vector<double>>& tgt,
                  const std::vector<double>& w, uint64_t seed, uint32_t iter,
                  uint32_t phase, std::vector<int64_t>& words) {
  for (size_t l = 0; l < lat.size(); ++l) {
    g_stream = hea_oracle::Stream{seed, iter, phase, (uint32_t)l, 0};
    g_draw_words = 0;
    Lat one(1, lat[l]);
    std::vector<double> f1(1, fit[l]);
    local_parallel_monte_carlo<int, double>(one, f1, T, D, nbor, species, tgt, w);
    lat[l] = one[0];
    fit[l] = f1[0];
    words.push_back(g_draw_words);
  }
  sort_lattices_by_fitness<int, double>(fit, lat);
}

}  // namespace

int main(int argc, char** argv) {
  if (argc != 3) die("usage: ref_hcs <in> <out>");
  FILE* in = fopen(argv[1], "rb");
  if (!in) die("cannot open input");
  Header h;
  rd(in, &h, 1);
  if (h.magic != 0x52414548) die("bad magic");
  const int N = h.N, K = h.K, S = h.S;
  const int Z[3] = {h.Z0, h.Z1, h.Z2};
  int sumZ = 0;
  for (int s = 0; s < S; ++s) sumZ += Z[s];
  std::vector<int32_t> flat((size_t)N * sumZ);
  rd(in, flat.data(), flat.size());
  std::vector<int32_t> counts(K);
  rd(in, counts.data(), K);
  std::vector<double> weight(S);
  rd(in, weight.data(), S);
  std::vector<std::vector<double>> target(3, std::vector<double>((size_t)K * K, 0.0));
  for (int s = 0; s < 3; ++s) rd(in, target[s].data(), (size_t)K * K);

  Nbr nbor(N, std::vector<std::vector<int>>(S));
  for (int i = 0; i < N; ++i) {
    int o = 0;
    for (int s = 0; s < S; ++s) {
      nbor[i][s].assign(flat.begin() + (size_t)i * sumZ + o,
                        flat.begin() + (size_t)i * sumZ + o + Z[s]);
      o += Z[s];
    }
  }
  std::vector<int> species(counts.begin(), counts.end());
  std::vector<double> sro(S, 0.0);

  g_out = fopen(argv[2], "wb");
  if (!g_out) die("cannot open output");

  if (h.mode == 0) {
    // Full search: run_local_parallel_hcs (accelerate.cpp:713-763) on one
    // rank with keyed streams.  Snapshots after every stage.
    const int L = h.L;
    Lat lat(L, std::vector<int>(N)), nl(L, std::vector<int>(N));
    std::vector<double> fit(L), nf(L);
    std::vector<uint8_t> buf(N);
    std::vector<int64_t> words;
    auto t0 = std::chrono::steady_clock::now();
    for (int l = 0; l < L; ++l) {
      hea_oracle::permute_lattice(h.seed, 0, 0, l, counts.data(), K, N, buf.data());
      lat[l].assign(buf.begin(), buf.end());
    }
    calculate_fitness_of_solutions<int, double>(fit, lat, sro, weight, species, nbor, target);
    put_population(lat, fit);
    for (int it = 0; it < h.I; ++it) {
      for (int l = 0; l < L; ++l) {
        hea_oracle::permute_lattice(h.seed, it, 1, l, counts.data(), K, N, buf.data());
        nl[l].assign(buf.begin(), buf.end());
      }
      calculate_fitness_of_solutions<int, double>(nf, nl, sro, weight, species, nbor, target);
      local_search(nl, nf, h.T, h.D, nbor, species, target, weight, h.seed, it, 2, words);
      put_population(nl, nf);
      calculate_best_lattices<int, double>(lat, nl, fit, nf);
      put_population(lat, fit);
      local_search(lat, fit, h.T, h.D, nbor, species, target, weight, h.seed, it, 3, words);
      put_population(lat, fit);
    }
    double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    wr(words.data(), words.size());
    wr(&secs, 1);
  } else if (h.mode == 1) {
    // One local-search phase on a given population.
    const int L = h.L;
    Lat lat(L, std::vector<int>(N));
    std::vector<double> fit(L);
    std::vector<uint8_t> buf(N);
    for (int l = 0; l < L; ++l) {
      rd(in, buf.data(), N);
      lat[l].assign(buf.begin(), buf.end());
    }
    rd(in, fit.data(), L);
    std::vector<int64_t> words;
    auto t0 = std::chrono::steady_clock::now();
    local_search(lat, fit, h.T, h.D, nbor, species, target, weight, h.seed, h.iter, h.phase,
                 words);
    double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    put_population(lat, fit);
    wr(words.data(), words.size());
    wr(&secs, 1);
  } else if (h.mode == 2) {
    // calculate_fitness_of_solutions (accelerate.cpp:118-180).
    const int L = h.L;
    Lat lat(L, std::vector<int>(N));
    std::vector<double> fit(L);
    std::vector<uint8_t> buf(N);
    for (int l = 0; l < L; ++l) {
      rd(in, buf.data(), N);
      lat[l].assign(buf.begin(), buf.end());
    }
    calculate_fitness_of_solutions<int, double>(fit, lat, sro, weight, species, nbor, target);
    wr(fit.data(), fit.size());
  } else if (h.mode == 3) {
    // calculate_sro_coefficient + a chain of calculate_fitness_incremental
    // calls with gamma carried forward (accelerate.cpp:370-399, 498-570).
    std::vector<uint8_t> buf(N);
    rd(in, buf.data(), N);
    std::vector<int> lat(buf.begin(), buf.end());
    std::vector<int32_t> sw(2 * (size_t)h.n_swaps);
    rd(in, sw.data(), sw.size());
    std::vector<std::vector<std::vector<double>>> gamma(
        K, std::vector<std::vector<double>>(K, std::vector<double>(S, 0.0)));
    const int nat = N, nty = K, nsh = S;
    calculate_sro_coefficient<int, double>(gamma, nat, nty, nsh, lat, species, nbor);
    std::vector<double> fits(h.n_swaps);
    for (int q = 0; q < h.n_swaps; ++q) {
      const int a1 = sw[2 * q], a2 = sw[2 * q + 1];
      fits[q] = calculate_fitness_incremental<int, double>(lat, gamma, a1, a2, nat, nty, nsh,
                                                            species, nbor, target, weight);
    }
    wr(fits.data(), fits.size());
    for (int a = 0; a < K; ++a)
      for (int b = 0; b < K; ++b) wr(gamma[a][b].data(), S);
    std::vector<uint8_t> ob(lat.begin(), lat.end());
    wr(ob.data(), ob.size());
  } else {
    die("bad mode");
  }
  fclose(g_out);
  fclose(in);
  return 0;
}
