// pybind11 module `accelerate`: the reference's Python-facing names over the
// C ABI in include/pyhea_hcs.h.
//
// Mirrors pyhea/cpp/bind.cpp:6-122 (module `pyhea.cpp.accelerate`):
//   run_local_parallel_hcs(nnet, step, task, depth, thr, nbor, element,
//                          weight, target_sro)        bind.cpp:7-46
//   run_local_parallel_hcs_cuda(<same>)                bind.cpp:49-88
//   cuda_available()                                   bind.cpp:90-120
// with the same keyword names and return shape (list[nnet][N] of int,
// list[nnet] of float, ascending fitness).  Differences, all additive:
//   * both run names execute the MI355X engine; it reproduces the reference
//     CPU engine's doubles bit for bit on the same keyed streams, so the CPU
//     name keeps its meaning;
//   * keyword-only `seed` and `device` (GPU ordinal); without a seed every
//     call draws one from std::random_device, as the reference seeds its
//     permutations (accelerate.cpp:241-242), so repeated calls differ;
//   * `nbor` may also be a 2-D integer numpy array [N][sum Z_s] with the
//     keyword `shell_size`, which avoids building nested lists at scale;
//   * errors raise ValueError / RuntimeError instead of exiting the process
//     (accelerate.cu:981-991, 1082-1091 call exit(1));
//   * the GIL is released while the GPU runs.
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <random>
#include <stdexcept>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/pyhea_hcs.h"

namespace py = pybind11;

namespace {


void check(int rc) {
  if (rc == HEA_OK) return;
  const std::string msg = hea_last_error();
  if (rc == HEA_EINVAL || rc == HEA_EUNSUPPORTED) throw py::value_error(msg);
  throw std::runtime_error(msg);
}

struct Flat {
  std::vector<int32_t> nbr;
  int z[3] = {0, 0, 0};
  int n = 0;
};

// nbor[i][s][k] for s < n_shells -> [N][sum Z] (pyhea/lattice/lattice.py:196-323)
Flat flatten(const py::object& nbor, int n_shells, const py::object& shell_size) {
  Flat f;
  if (py::isinstance<py::array>(nbor)) {
    auto arr = py::array_t<int32_t, py::array::c_style | py::array::forcecast>::ensure(nbor);
    if (!arr || arr.ndim() != 2) throw py::value_error("nbor array must be 2-D [N][sum Z]");
    if (shell_size.is_none()) throw py::value_error("nbor array needs shell_size=[Z_0, ...]");
    auto zs = shell_size.cast<std::vector<int>>();
    if ((int)zs.size() < n_shells) throw py::value_error("shell_size shorter than weight");
    int tot = 0;
    for (size_t s = 0; s < zs.size(); ++s) tot += zs[s];
    if (tot != arr.shape(1)) throw py::value_error("shell_size does not match nbor columns");
    f.n = (int)arr.shape(0);
    int used = 0;
    for (int s = 0; s < n_shells; ++s) used += (f.z[s] = zs[s]);
    f.nbr.resize((size_t)f.n * used);
    auto r = arr.unchecked<2>();
    for (int i = 0; i < f.n; ++i)
      for (int c = 0; c < used; ++c) f.nbr[(size_t)i * used + c] = r(i, c);
    return f;
  }
  auto nested = nbor.cast<std::vector<std::vector<std::vector<int>>>>();
  f.n = (int)nested.size();
  if (f.n == 0) throw py::value_error("empty neighbour list");
  if ((int)nested[0].size() < n_shells)
    throw py::value_error("neighbour list has fewer shells than weights");
  int used = 0;
  for (int s = 0; s < n_shells; ++s) used += (f.z[s] = (int)nested[0][s].size());
  f.nbr.resize((size_t)f.n * used);
  for (int i = 0; i < f.n; ++i) {
    if ((int)nested[i].size() < n_shells) throw py::value_error("ragged neighbour list");
    int c = 0;
    for (int s = 0; s < n_shells; ++s) {
      if ((int)nested[i][s].size() != f.z[s])
        throw py::value_error("every site needs the same shell sizes");
      for (int k = 0; k < f.z[s]; ++k) f.nbr[(size_t)i * used + c++] = nested[i][s][k];
    }
  }
  return f;
}

std::tuple<std::vector<std::vector<int>>, std::vector<double>> run(
    int nnet, int step, int task, int depth, double thr, const py::object& nbor,
    const std::vector<int>& element, const std::vector<double>& weight,
    const py::object& target_sro, const py::object& seed_arg, int device,
    const py::object& shell_size) {
  uint64_t seed;
  if (seed_arg.is_none()) {
    std::random_device rd;  // a fresh key per call (accelerate.cpp:241-242)
    seed = ((uint64_t)rd() << 32) | rd();
  } else {
    seed = seed_arg.cast<uint64_t>();
  }
  const int S = (int)weight.size();
  if (S < 1 || S > 3) throw py::value_error("len(weight) must be 1..3");
  const int K = (int)element.size();
  Flat f = flatten(nbor, S, shell_size);
  auto tgt = py::array_t<double, py::array::c_style | py::array::forcecast>::ensure(target_sro);
  if (!tgt || tgt.ndim() != 2 || tgt.shape(0) < S || tgt.shape(1) != (py::ssize_t)K * K)
    throw py::value_error("target_sro must be [>= len(weight)][K*K]");
  std::vector<double> target((size_t)S * K * K);
  auto t = tgt.unchecked<2>();
  for (int s = 0; s < S; ++s)
    for (int e = 0; e < K * K; ++e) target[(size_t)s * K * K + e] = t(s, e);
  std::vector<int32_t> counts(element.begin(), element.end());

  hea_problem_desc d;
  d.n_sites = f.n;
  d.n_species = K;
  d.n_shells = S;
  for (int s = 0; s < 3; ++s) d.shell_size[s] = f.z[s];
  d.nbr = f.nbr.data();
  d.counts = counts.data();
  d.weight = weight.data();
  d.target = target.data();

  std::vector<uint8_t> lat((size_t)nnet * f.n);
  std::vector<double> fit(nnet);
  {
    py::gil_scoped_release nogil;
    hea_problem p = nullptr;
    check(hea_problem_create(&d, device, &p));
    const int rc = hea_hcs_run(p, nnet, step, task, depth, thr, seed, lat.data(), fit.data(),
                               nullptr);
    hea_problem_destroy(p);
    check(rc);
  }
  std::vector<std::vector<int>> out(nnet, std::vector<int>(f.n));
  for (int l = 0; l < nnet; ++l)
    for (int i = 0; i < f.n; ++i) out[l][i] = lat[(size_t)l * f.n + i];
  return std::make_tuple(std::move(out), std::move(fit));
}

const char* kDoc = R"pbdoc(
    Run the local parallel hybrid cuckoo search (SRO-targeted lattice search).

    Parameters
    ----------
    nnet : int
        Number of solutions (lattices) in the population.
    step : int
        Global iterations (generate, local search I, merge, local search II).
    task : int
        Proposals evaluated per local-search round (parallel_task), 1..2048.
    depth : int
        Consecutive rejected rounds that end a local search (converge_depth).
    thr : float
        Fitness threshold; only reported, as in the reference.
    nbor : list[N][S][Z_s] of int, or int array [N][sum Z_s] with shell_size
        Neighbour lists, each shell ascending.
    element : list of int
        Atoms per species.
    weight : list of float
        Shell weights; len(weight) selects the number of shells.
    target_sro : array-like [>= len(weight)][K*K]
        Target Warren-Cowley parameters per shell.
    seed : int or None, keyword-only
        Key of every random stream of the run; None (the default) draws a
        fresh key from std::random_device, so calls differ like the
        reference's.
    device : int, keyword-only
        GPU ordinal.

    Returns
    -------
    tuple
        (lattices, fitness): list[nnet][N] of species indices and list[nnet]
        of fitness values, ascending.
)pbdoc";

}  // namespace

PYBIND11_MODULE(accelerate, m) {
  m.doc() = "MI355X SRO lattice search (drop-in for pyhea.cpp.accelerate)";
  for (const char* name : {"run_local_parallel_hcs", "run_local_parallel_hcs_cuda"}) {
    m.def(name, &run, py::arg("nnet"), py::arg("step"), py::arg("task"), py::arg("depth"),
          py::arg("thr"), py::arg("nbor"), py::arg("element"), py::arg("weight"),
          py::arg("target_sro"), py::kw_only(), py::arg("seed") = py::none(),
          py::arg("device") = 0, py::arg("shell_size") = py::none(), kDoc);
  }
  m.def(
      "cuda_available", []() { return hea_device_count() > 0; },
      "True if a GPU is visible to the HIP runtime (bind.cpp:90-120).");
  m.def("device_count", []() { return hea_device_count(); });
}
