/*
 * pyhea_hcs.h -- C ABI of the MI355X (gfx950) SRO-targeted lattice search.
 *
 * This is the drop-in boundary for PyHEA's native engine.  The reference
 * binds its engine to Python with pybind11 (pyhea/cpp/bind.cpp:6-122,
 * declarations pyhea/cpp/accelerate.h:23-66): run_local_parallel_hcs,
 * run_local_parallel_hcs_cuda and cuda_available.  Those three names are kept
 * by the pybind11 module pyhea_amd/cpp/accelerate, which is a thin shim over
 * the functions below; any other FFI (ctypes, cffi) binds them directly.
 *
 * Conventions
 *   - plain C types only; all buffers are caller-owned and nothing is retained
 *     after a call returns unless a handle says so;
 *   - every function returns HEA_OK (0) or a nonzero HEA_E* code and never
 *     exits the process; hea_last_error() gives the message (thread-local);
 *   - lattices are uint8 species indices, [solutions][n_sites] row-major;
 *     fitness values are FP64 and bit-identical to the reference CPU engine
 *     (pyhea/cpp/accelerate.cpp) driven by the same keyed streams;
 *   - calls on one handle are not reentrant; distinct handles are independent.
 */
#ifndef PYHEA_HCS_H
#define PYHEA_HCS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HEA_OK 0
#define HEA_EINVAL 1       /* invalid argument                              */
#define HEA_ENODEV 2       /* no usable GPU                                 */
#define HEA_EHIP 3         /* HIP runtime error                             */
#define HEA_EUNSUPPORTED 4 /* valid input outside what the kernels support  */
#define HEA_ENOMEM 5       /* device allocation failed                      */

/* Stream phases: which draw of the search a keyed stream belongs to. */
#define HEA_PHASE_INIT 0 /* initial population (accelerate.cpp:737)       */
#define HEA_PHASE_NEW 1  /* per-iteration new population (:743)           */
#define HEA_PHASE_LS1 2  /* local search I proposals (:745)               */
#define HEA_PHASE_LS2 3  /* local search II proposals (:749)              */

/* A search problem.  Replaces the reference's nested-vector arguments
 * `neighbor_list`, `species`, `weight`, `target_sro` of
 * accelerate::cpu::run_local_parallel_hcs (pyhea/cpp/accelerate.h:23-32). */
typedef struct hea_problem_desc {
  int32_t n_sites;      /* N                                                  */
  int32_t n_species;    /* K, 2..8                                            */
  int32_t n_shells;     /* S, 1..3 (len(weight), pyhea/model/opt.py:49-52)    */
  int32_t shell_size[3];/* Z_s, 1..31; every site has Z_s members in shell s  */
  const int32_t* nbr;   /* [N][sum_s Z_s]: shell s in columns
                           [sum_{t<s} Z_t, +Z_s), each list ascending as
                           pyhea/lattice/lattice.py:236-320 builds it          */
  const int32_t* counts;/* [K] atoms per species, each >= 1, sum = N          */
  const double* weight; /* [S] shell weights, finite and >= 0                 */
  const double* target; /* [S][K*K] target Warren-Cowley alpha, row-major     */
} hea_problem_desc;

typedef struct hea_problem_s* hea_problem;
typedef struct hea_session_s* hea_session;
typedef struct hea_batch_s* hea_batch;

/* Counters of the work done; one swap eval = one incremental delta-F of a
 * valid proposal. */
typedef struct hea_stats {
  int64_t evals;     /* swap evaluations                                   */
  int64_t rounds;    /* local-search rounds                                */
  int64_t accepts;   /* accepted swaps                                     */
  int64_t draws;     /* proposal draws including same-species redraws     */
  double first_hit_s;/* seconds from session start to the first fitness
                        <= thr seen by local search, or -1 if never        */
} hea_stats;

/* Number of visible GPUs (0 if none or the runtime is unusable).
 * Backs cuda_available() (pyhea/cpp/bind.cpp:90-120, pyhea/model/opt.py:87). */
int hea_device_count(void);

/* Message of the last failed call on this thread. */
const char* hea_last_error(void);

/* Validates a problem and makes it resident on `device`: neighbour table,
 * per-problem constants and the exact R_s(k) segment tables.  The neighbour
 * relation must be symmetric with multiplicity (true for every periodic
 * lattice built by pyhea/lattice/lattice.py). */
int hea_problem_create(const hea_problem_desc* desc, int device, hea_problem* out);
int hea_problem_destroy(hea_problem p);

/* Whole search: replaces accelerate::cpu::run_local_parallel_hcs and
 * accelerate::gpu::run_local_parallel_hcs_cuda (pyhea/cpp/accelerate.h:23-32,
 * 51-60; pyhea/cpp/accelerate.cpp:713-763).  Arguments keep the reference's
 * meaning: solutions = nnet, total_iter = step, parallel_task = task
 * (1..2048; a round's per-proposal records must also fit the workgroup's
 * 160 KiB of LDS, about 80 bytes per proposal for K <= 6 and 104 for K 7-8,
 * so about 1900 and 1450 at most; larger rounds return HEA_EUNSUPPORTED),
 * converge_depth = depth, thr only sets the time-to-threshold
 * statistic (the reference ignores it, accelerate.cpp:612, 718).  The final
 * population is sorted by ascending fitness.  out_lattices [solutions][N],
 * out_fitness [solutions], stats may be NULL. */
int hea_hcs_run(hea_problem p, int32_t solutions, int32_t total_iter, int32_t parallel_task,
                int32_t converge_depth, double thr, uint64_t seed, uint8_t* out_lattices,
                double* out_fitness, hea_stats* stats);

/* One local search over a caller-provided population (host buffers, updated
 * in place), with streams (seed, iter, phase, slot = position).  Replaces
 * accelerate::cpu::local_parallel_monte_carlo (accelerate.cpp:601-694);
 * sort != 0 applies its closing sort by fitness (accelerate.cpp:693). */
int hea_local_mc(hea_problem p, int32_t solutions, uint8_t* lattices, double* fitness,
                 int32_t parallel_task, int32_t converge_depth, uint64_t seed, uint32_t iter,
                 uint32_t phase, int32_t sort, hea_stats* stats);

/* Same on device-resident buffers (d_lattices [solutions][N] uint8,
 * d_fitness [solutions] double) on `stream` (a hipStream_t, NULL = default);
 * asynchronous: no host synchronisation.
 *
 * Device memory: when a lattice's species do not fit in a workgroup's LDS
 * (beyond about 110000 sites, depending on parallel_task and K) the search
 * reads them from HBM and keeps one 64-bit neighbour-count record per site
 * and solution (8 * solutions * N bytes) for its walk; hea_local_mc and
 * hea_local_mc_device grow the problem's buffer for it on demand, sessions
 * and batches allocate theirs when created (or when a tile or guide call
 * changes the placement).  Results are the same with or without it. */
int hea_local_mc_device(hea_problem p, int32_t solutions, void* d_lattices, void* d_fitness,
                        int32_t parallel_task, int32_t converge_depth, uint64_t seed,
                        uint32_t iter, uint32_t phase, void* stream);

/* Bytes of the fixed per-device workspace of hea_local_mc_device_noalloc. */
#define HEA_NOALLOC_SCRATCH_BYTES (4u << 20)

/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  Same arguments and results as hea_local_mc_device,
 * but the search's workspace (iteration word, per-lattice counters, pair
 * counts) lives in a fixed array of HEA_NOALLOC_SCRATCH_BYTES that each
 * device holds for the life of the process, and the call allocates nothing
 * (no hipMalloc, hipFree or allocating library routine).  It needs
 * 256 + roundup(32 * solutions, 256) + 4 * solutions * S * K * K bytes;
 * larger calls return HEA_EINVAL.  The workspace keeps the last call's
 * contents and is shared by every call on the device: calls that overlap on
 * one device (other streams, other threads, other problems) race and must
 * be serialised by the caller. */
int hea_local_mc_device_noalloc(hea_problem p, int32_t solutions, void* d_lattices,
                                void* d_fitness, int32_t parallel_task, int32_t converge_depth,
                                uint64_t seed, uint32_t iter, uint32_t phase, void* stream);

/* Fitness by full recount: accelerate::cpu::calculate_fitness_of_solutions
 * (accelerate.cpp:118-180). */
int hea_fitness(hea_problem p, int32_t solutions, const uint8_t* lattices, double* out_fitness);

/* Integer pair counts n[l][s][a][b] of each lattice (directed pairs). */
int hea_pair_counts(hea_problem p, int32_t solutions, const uint8_t* lattices, int32_t* out);

/* Keyed random lattices with their recount fitness: the role of
 * accelerate::cpu::generate_random_lattices (accelerate.cpp:218-252). */
int hea_generate(hea_problem p, int32_t solutions, uint64_t seed, uint32_t iter, uint32_t phase,
                 uint8_t* out_lattices, double* out_fitness);

/* Device-resident search session: the population stays in HBM and every
 * global iteration (generate -> local search I -> merge -> local search II,
 * accelerate.cpp:741-761) runs as one replayable HIP graph. */
int hea_session_create(hea_problem p, int32_t solutions, int32_t parallel_task,
                       int32_t converge_depth, double thr, uint64_t seed, hea_session* out);
/* Initial population (accelerate.cpp:737); starts the session clock. */
int hea_session_init(hea_session s);
/* Launch shape of the session's local searches, no counterpart in the
 * reference: `tile` proposal slots per pass (64, 128, 192 or 256).  A shell
 * whose neighbour list is at least twice as long as the others' is walked by
 * 2 or 4 lanes per slot (at most 1024 lanes).  Results do not depend on it;
 * the autotuner (pyhea_amd/autotune.py) picks it. */
int hea_session_tile(hea_session s, int32_t tile);
/* Guided proposals, no counterpart in the reference.  The local search keeps
 * every site's neighbour species counts and the gradient of the fitness
 * squared with respect to the pair counts (the relaxation's gradient at the
 * current lattice) up to date; the predicted first-order change of swapping
 * i (species a) with j (species b) is u[i][b] + u[j][a].  Each of the
 * parallel_task proposals of a round is the Gumbel-max of
 * -(u[i][b] + u[j][a]) / (tau * scale) over `candidates` hashed random pairs:
 * the slot draws two blocks of 16 consecutive sites and each candidate one
 * site in each (so a slot's candidates read two 128-byte lines of records;
 * the environment variable PYHEA_AMD_GUIDE_BLOCK=b, read when a problem is
 * created, sets blocks of 2^b sites, 0 draws both sites over the lattice).
 * The round's best proposal is accepted by the exact rule of
 * accelerate.cpp:663-689; then up to accepts - 1 more come from the same
 * round: the proposals that touch none of the swapped sites or their
 * neighbours are screened again against the new lattice, and the best is
 * accepted iff its exact fitness is lower.  Needs K <= 6, no self
 * neighbours, and a site's counts in one 64-bit record: (K-1) * sum over
 * shells of the bit width of Z_s <= 64 (FCC 12/6/24: 12 bits per species).
 * Call before hea_session_init/step. */
int hea_session_guide(hea_session s, int32_t candidates, double tau, int32_t accepts);
/* Test capture with no counterpart in the reference: the first candidate
 * phase of a guided local search (as hea_session_guide describes it) on each
 * of the caller's lattices (host [solutions][N]), streams (seed, iter, phase,
 * slot = position).  out_pairs [solutions][parallel_task * candidates][2]
 * receives every scored candidate pair (i, j), out_pred
 * [solutions][parallel_task * candidates] its predicted first-order change of
 * the fitness squared, u[i][b] + u[j][a] with a, b the species of i and j and
 * u the site gains of hea_site_gain at the lattice's dn = dF^2/dn (NaN for
 * same-species pairs).  No search step is taken.  Synchronous. */
int hea_guided_scores(hea_problem p, int32_t solutions, const uint8_t* lattices,
                      int32_t parallel_task, int32_t candidates, uint64_t seed, uint32_t iter,
                      uint32_t phase, int32_t* out_pairs, float* out_pred);
/* Runs `n_iter` global iterations asynchronously on the session stream. */
int hea_session_step(hea_session s, int32_t n_iter);
/* Measurement helper with no counterpart in the reference: runs n_iter
 * global iterations eagerly (no graph) with HIP events around each local
 * search.  out_ms[2i], out_ms[2i+1]: durations (ms) of local search I and
 * II of iteration i; out_evals[2i], out_evals[2i+1]: their swap evals.
 * Synchronous. */
int hea_session_step_timed(hea_session s, int32_t n_iter, float* out_ms, int64_t* out_evals);
/* Waits for the session stream. */
int hea_session_sync(hea_session s);
/* Copies out the population (sorted) and the accumulated counters. */
int hea_session_read(hea_session s, uint8_t* out_lattices, double* out_fitness, hea_stats* stats);
/* Population best fitness after the last completed iteration. */
int hea_session_best(hea_session s, double* best);
int hea_session_destroy(hea_session s);

/* Batched search with no counterpart in the reference (which runs one
 * problem per call, accelerate.cpp:713-763): n_probs independent problems on
 * one device, each with its own population of `solutions` lattices, all
 * searched by one persistent local-search launch per phase.  Problem p
 * draws from seed `seed + p`, so its population equals
 * hea_hcs_run(probs[p], ..., seed + p, ...).  Problems must share the
 * device, the shell count and have no self neighbours; parallel_task and
 * converge_depth are common; at most 65535 problems.  The problems must
 * outlive the batch. */
int hea_batch_create(const hea_problem* probs, int32_t n_probs, int32_t solutions,
                     int32_t parallel_task, int32_t converge_depth, double thr, uint64_t seed,
                     hea_batch* out);
/* Sets the number of persistent workgroups (launch shape chosen by the
 * autotuner; results do not depend on it). */
int hea_batch_grid(hea_batch b, int32_t grid);
/* Proposals per pass of the batch's persistent launch (see hea_session_tile). */
int hea_batch_tile(hea_batch b, int32_t tile);
/* Guided proposals for every problem of the batch (see hea_session_guide). */
int hea_batch_guide(hea_batch b, int32_t candidates, double tau, int32_t accepts);
int hea_batch_init(hea_batch b);
int hea_batch_step(hea_batch b, int32_t n_iter);
/* Measurement helper with no counterpart in the reference: runs n_iter
 * global iterations eagerly (no graph) with HIP events around each
 * persistent local-search launch.  out_ms[2i], out_ms[2i+1]: durations (ms)
 * of the local search I and II launches of iteration i; out_stats[2i],
 * out_stats[2i+1]: their counters summed over the batch's problems
 * (first_hit_s = -1).  Synchronous. */
int hea_batch_step_timed(hea_batch b, int32_t n_iter, float* out_ms, hea_stats* out_stats);
int hea_batch_sync(hea_batch b);
/* Population (sorted) and counters of problem p. */
int hea_batch_read(hea_batch b, int32_t p, uint8_t* out_lattices, double* out_fitness,
                   hea_stats* stats);
int hea_batch_destroy(hea_batch b);

/* Differentiable relaxation, no counterpart in the reference (which searches
 * only discrete lattices).  With p = softmax(logits) per site, the soft pair
 * counts n[s][a][b] = sum_i sum_{j in N_s(i)} p[i][a] p[j][b] equal the integer
 * pair counts when p is one-hot.  All pointers are device memory; dtype 0 =
 * float32, 1 = float64 for logits [N][K], grad [N][K], and the occupancies
 * p (N * KP elements) and neighbour sums q (S * N * KP) that the forward
 * writes for the backward, KP = hea_relax_cols(K) (p and q 16-byte aligned):
 * for K <= 4 rows of 4 columns (zeros beyond K); for K > 4 columns 0-3 of
 * every row first ([N][4]; q: [S][N][4]), then the remaining W2 = KP - 4
 * columns ([N][W2]; q: [S][N][W2]), zeros beyond K.  n and dn are float64
 * [S][K][K].  Asynchronous on
 * `stream`.  Calls on one problem must not overlap.  The first
 * hea_relax_forward or hea_site_gain on a problem allocates its work buffers
 * (about 4 * N * row ints bytes) and waits for `stream` once, so it must not
 * be captured into a graph. */
int hea_relax_forward(hea_problem p, int32_t dtype, const void* d_logits, void* d_p, void* d_q,
                      void* d_n, void* stream);
/* KP, the elements of one site's p row in the layout above, for n_species
 * species (4 for K <= 4, 5, 6, 8 for K = 5, 6, 7-8); -1 outside 1..8.  Size
 * the buffers with it: p holds N * KP elements, q holds S * N * KP. */
int32_t hea_relax_cols(int32_t n_species);
/* Neighbour-gather adjoint: d loss / d logits from dn = d loss / d n. */
int hea_relax_backward(hea_problem p, int32_t dtype, const void* d_p, const void* d_q,
                       const void* d_dn, void* d_grad, void* stream);
/* Swap guidance at discrete lattices: u[l][k][b] (float32) = first-order
 * loss change of turning site k of lattice l into species b, from
 * dn[l] = d loss / d n (float64 [L][S][K][K]). */
int hea_site_gain(hea_problem p, int32_t L, const void* d_lattices, const void* d_dn, void* d_u,
                  void* stream);

#ifdef __cplusplus
}
#endif
#endif /* PYHEA_HCS_H */
