// Profiling driver: runs the search session of a dumped problem through the
// C ABI with no Python in the process, so hardware-counter profilers can
// attach to a plain executable.
//
//   hea_profile <problem.bin> <solutions> <task> <depth> <seed> <warmup> <steps>
//               [copy_mib [tile [candidates accepts]]]
//
// problem.bin (written by bench.py --dump-problem): int32 N, K, S, Z[3];
// int32 nbr[N][sum Z]; int32 counts[K]; float64 weight[S];
// float64 target[S][K*K].  After the timed steps an optional device-to-device
// copy of copy_mib MiB gives a kernel of known size for calibrating byte
// counters (0: none).  tile > 0 sets the proposals per pass; candidates > 0
// switches to guided proposals (tau 0.1).  Prints one line: evals, seconds,
// evals/s.
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <stdlib.h>

#include <chrono>
#include <vector>

#include "../../include/pyhea_hcs.h"

static void need(int rc, const char* what) {
  if (rc != HEA_OK) {
    fprintf(stderr, "%s failed (%d): %s\n", what, rc, hea_last_error());
    exit(1);
  }
}

template <typename T>
static void rd(FILE* f, T* p, size_t n) {
  if (n && fread(p, sizeof(T), n, f) != n) {
    fprintf(stderr, "short problem file\n");
    exit(1);
  }
}

int main(int argc, char** argv) {
  if (argc < 8) {
    fprintf(stderr,
            "usage: hea_profile problem.bin L T D seed warmup steps [copy_mib [tile [cand acc]]]\n");
    return 2;
  }
  FILE* f = fopen(argv[1], "rb");
  if (!f) return 2;
  int32_t hdr[6];
  rd(f, hdr, 6);
  const int N = hdr[0], K = hdr[1], S = hdr[2];
  int sumZ = 0;
  for (int s = 0; s < S; ++s) sumZ += hdr[3 + s];
  std::vector<int32_t> nbr((size_t)N * sumZ), counts(K);
  std::vector<double> weight(S), target((size_t)S * K * K);
  rd(f, nbr.data(), nbr.size());
  rd(f, counts.data(), counts.size());
  rd(f, weight.data(), weight.size());
  rd(f, target.data(), target.size());
  fclose(f);
  hea_problem_desc d;
  d.n_sites = N;
  d.n_species = K;
  d.n_shells = S;
  for (int s = 0; s < 3; ++s) d.shell_size[s] = hdr[3 + s];
  d.nbr = nbr.data();
  d.counts = counts.data();
  d.weight = weight.data();
  d.target = target.data();
  const int L = atoi(argv[2]), T = atoi(argv[3]), D = atoi(argv[4]);
  const unsigned long long seed = strtoull(argv[5], nullptr, 0);
  const int warm = atoi(argv[6]), steps = atoi(argv[7]);
  hea_problem p;
  need(hea_problem_create(&d, 0, &p), "problem");
  hea_session s;
  need(hea_session_create(p, L, T, D, 1e-3, seed, &s), "session");
  if (argc > 9 && atoi(argv[9]) > 0) need(hea_session_tile(s, atoi(argv[9])), "tile");
  if (argc > 11 && atoi(argv[10]) > 0)
    need(hea_session_guide(s, atoi(argv[10]), 0.1, atoi(argv[11])), "guide");
  need(hea_session_init(s), "init");
  need(hea_session_step(s, warm), "warmup");
  need(hea_session_sync(s), "sync");
  hea_stats st0, st1;
  need(hea_session_read(s, nullptr, nullptr, &st0), "read");
  auto t0 = std::chrono::steady_clock::now();
  need(hea_session_step(s, steps), "step");
  need(hea_session_sync(s), "sync");
  const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  need(hea_session_read(s, nullptr, nullptr, &st1), "read");
  const long long ev = st1.evals - st0.evals;
  printf("evals %lld seconds %.6f evals_per_s %.6e\n", ev, secs, ev / secs);
  if (argc > 8 && atoi(argv[8]) > 0) {
    const size_t bytes = (size_t)atoi(argv[8]) << 20;
    void *a = nullptr, *b = nullptr;
    if (hipMalloc(&a, bytes) == hipSuccess && hipMalloc(&b, bytes) == hipSuccess) {
      (void)hipMemset(a, 1, bytes);
      (void)hipMemcpy(b, a, bytes, hipMemcpyDeviceToDevice);
      (void)hipDeviceSynchronize();
      printf("copied %zu bytes device to device\n", bytes);
    }
    (void)hipFree(a);
    (void)hipFree(b);
  }
  hea_session_destroy(s);
  hea_problem_destroy(p);
  return 0;
}
