// Differentiable relaxation of the lattice search and its adjoint.
//
// No counterpart in the reference (pyhea/cpp/accelerate.cpp searches only
// discrete lattices).  Each site i carries logits theta[i][K]; its occupancy
// is p[i] = softmax(theta[i]).  The soft pair counts
//     n[s][a][b] = sum_i sum_{j in N_s(i)} p[i][a] p[j][b]
// reduce to the integer pair counts of calculate_sro_coefficient
// (accelerate.cpp:370-399, times Z_s N) when p is one-hot, so
// gamma = n / (Z_s N) and alpha = 1 - gamma / (c_a c_b) relax the
// Warren-Cowley fitness smoothly.
//
// Forward: softmax, neighbour gather Q_s[i][b] = sum_{j in N_s(i)} p[j][b],
// then n_s = P^T Q_s, reduced deterministically (fixed-size site chunks, then
// chunk partials in order) in FP64.
// Backward (neighbour-gather adjoint): the neighbour relation is symmetric,
// so dn[s][a][b]/dp[k][c] = [a=c] Q_s[k][b] + [b=c] Q_s[k][a] and
//     g[k][c] = sum_s sum_b (dn[s][c][b] + dn[s][b][c]) Q_s[k][b],
// followed by the softmax adjoint p (g - <p, g>).
//
// Guidance at a discrete lattice: with p one-hot, Q_s[k][b] is the integer
// count of species b among k's shell-s neighbours and g[k][c] is the
// first-order change of the loss per unit of occupancy c at site k.
// k_site_gain writes u[k][b] = g[k][b] - g[k][a_k], the predicted loss change
// of turning site k from its species a_k into b.

#include "hea_internal.h"

namespace {

constexpr int kChunk = 1024;  // sites per reduction chunk

template <typename T>
__global__ void k_softmax(const T* __restrict__ th, T* __restrict__ p, int N, int K) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const T* r = th + (size_t)i * K;
  double m = -1e300;
  for (int a = 0; a < K; ++a) m = fmax(m, (double)r[a]);
  double e[HEA_MAX_K], z = 0.0;
  for (int a = 0; a < K; ++a) {
    e[a] = exp((double)r[a] - m);
    z += e[a];
  }
  for (int a = 0; a < K; ++a) p[(size_t)i * K + a] = (T)(e[a] / z);
}

// Q[s][i][b] = sum over the shell-s neighbours j of i of p[j][b].
template <typename T>
__global__ void k_gather(const HeaConst* __restrict__ pc, const int32_t* __restrict__ nbr,
                         const T* __restrict__ p, T* __restrict__ Q) {
  const HeaConst& P = *pc;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P.N) return;
  const int K = P.K;
  const int32_t* row = nbr + (size_t)i * P.stride;
  for (int s = 0; s < P.S; ++s) {
    double acc[HEA_MAX_K];
    for (int b = 0; b < K; ++b) acc[b] = 0.0;
    for (int k = 0; k < P.Z[s]; ++k) {
      const T* pj = p + (size_t)row[P.off[s] + k] * K;
      for (int b = 0; b < K; ++b) acc[b] += (double)pj[b];
    }
    T* q = Q + ((size_t)s * P.N + i) * K;
    for (int b = 0; b < K; ++b) q[b] = (T)acc[b];
  }
}

// Partial n over one chunk of sites: part[chunk][s][a][b], one thread per
// (s, a, b), summing the chunk's sites in order.
template <typename T>
__global__ void k_pair_partial(const HeaConst* __restrict__ pc, const T* __restrict__ p,
                               const T* __restrict__ Q, double* __restrict__ part) {
  const HeaConst& P = *pc;
  const int K = P.K, S = P.S;
  const int e = threadIdx.x;
  if (e >= S * K * K) return;
  const int s = e / (K * K), a = (e / K) % K, b = e % K;
  const int i0 = blockIdx.x * kChunk, i1 = min(P.N, i0 + kChunk);
  double acc = 0.0;
  for (int i = i0; i < i1; ++i)
    acc += (double)p[(size_t)i * K + a] * (double)Q[((size_t)s * P.N + i) * K + b];
  part[(size_t)blockIdx.x * S * K * K + e] = acc;
}

__global__ void k_pair_reduce(const double* __restrict__ part, int n_chunks, int n_ent,
                              double* __restrict__ n) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n_ent) return;
  double acc = 0.0;
  for (int c = 0; c < n_chunks; ++c) acc += part[(size_t)c * n_ent + e];
  n[e] = acc;
}

// grad[k][c] = p[k][c] (g[k][c] - sum_d p[k][d] g[k][d]),
// g[k][c] = sum_s sum_b (dn[s][c][b] + dn[s][b][c]) Q[s][k][b].
template <typename T>
__global__ void k_adjoint(const HeaConst* __restrict__ pc, const T* __restrict__ p,
                          const T* __restrict__ Q, const double* __restrict__ dn,
                          T* __restrict__ grad) {
  const HeaConst& P = *pc;
  const int K = P.K, S = P.S;
  __shared__ double H[HEA_MAX_S * HEA_MAX_K * HEA_MAX_K];
  for (int e = threadIdx.x; e < S * K * K; e += blockDim.x) {
    const int s = e / (K * K), c = (e / K) % K, b = e % K;
    H[e] = dn[(s * K + c) * K + b] + dn[(s * K + b) * K + c];
  }
  __syncthreads();
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= P.N) return;
  double g[HEA_MAX_K];
  for (int c = 0; c < K; ++c) g[c] = 0.0;
  for (int s = 0; s < S; ++s) {
    const T* q = Q + ((size_t)s * P.N + k) * K;
    for (int b = 0; b < K; ++b) {
      const double qb = (double)q[b];
      for (int c = 0; c < K; ++c) g[c] += H[(s * K + c) * K + b] * qb;
    }
  }
  double dot = 0.0;
  for (int c = 0; c < K; ++c) dot += (double)p[(size_t)k * K + c] * g[c];
  for (int c = 0; c < K; ++c)
    grad[(size_t)k * K + c] = (T)((double)p[(size_t)k * K + c] * (g[c] - dot));
}

// u[l][k][b] = g[k][b] - g[k][a_k] at the one-hot lattice l (see file header).
__global__ void k_site_gain(const HeaConst* __restrict__ pc, const int32_t* __restrict__ nbr,
                            const uint8_t* __restrict__ lattices, const double* __restrict__ dn,
                            float* __restrict__ u) {
  const HeaConst& P = *pc;
  const int K = P.K, S = P.S;
  const int l = blockIdx.y;
  __shared__ double H[HEA_MAX_S * HEA_MAX_K * HEA_MAX_K];
  const double* d = dn + (size_t)l * S * K * K;
  for (int e = threadIdx.x; e < S * K * K; e += blockDim.x) {
    const int s = e / (K * K), c = (e / K) % K, b = e % K;
    H[e] = d[(s * K + c) * K + b] + d[(s * K + b) * K + c];
  }
  __syncthreads();
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= P.N) return;
  const uint8_t* sp = lattices + (size_t)l * P.N;
  const int32_t* row = nbr + (size_t)k * P.stride;
  double g[HEA_MAX_K];
  for (int c = 0; c < K; ++c) g[c] = 0.0;
  for (int s = 0; s < S; ++s)
    for (int q = 0; q < P.Z[s]; ++q) {
      const int b = sp[row[P.off[s] + q]];
      for (int c = 0; c < K; ++c) g[c] += H[(s * K + c) * K + b];
    }
  const int a = sp[k];
  float* out = u + ((size_t)l * P.N + k) * K;
  for (int b = 0; b < K; ++b) out[b] = (float)(g[b] - g[a]);
}

}  // namespace

int hea_relax_chunks(int n_sites) { return (n_sites + kChunk - 1) / kChunk; }

template <typename T>
static hipError_t relax_forward_t(const HeaConst* dc, const HeaConst& hc, const int32_t* nbr,
                                  const void* th, void* p, void* Q, double* part, double* n,
                                  hipStream_t st) {
  const int N = hc.N, K = hc.K, S = hc.S;
  const int g = (N + 255) / 256;
  hipLaunchKernelGGL(k_softmax<T>, dim3(g), dim3(256), 0, st, static_cast<const T*>(th),
                     static_cast<T*>(p), N, K);
  hipLaunchKernelGGL(k_gather<T>, dim3(g), dim3(256), 0, st, dc, nbr, static_cast<const T*>(p),
                     static_cast<T*>(Q));
  const int nch = hea_relax_chunks(N);
  hipLaunchKernelGGL(k_pair_partial<T>, dim3(nch), dim3(256), 0, st, dc,
                     static_cast<const T*>(p), static_cast<const T*>(Q), part);
  const int ne = S * K * K;
  hipLaunchKernelGGL(k_pair_reduce, dim3((ne + 63) / 64), dim3(64), 0, st, part, nch, ne, n);
  return hipGetLastError();
}

hipError_t hea_relax_forward_impl(const HeaConst* dc, const HeaConst& hc, const int32_t* nbr,
                                  int dtype, const void* th, void* p, void* Q, double* part,
                                  double* n, hipStream_t st) {
  return dtype == 1 ? relax_forward_t<double>(dc, hc, nbr, th, p, Q, part, n, st)
                    : relax_forward_t<float>(dc, hc, nbr, th, p, Q, part, n, st);
}

hipError_t hea_relax_backward_impl(const HeaConst* dc, const HeaConst& hc, int dtype,
                                   const void* p, const void* Q, const double* dn, void* grad,
                                   hipStream_t st) {
  const int g = (hc.N + 255) / 256;
  if (dtype == 1)
    hipLaunchKernelGGL(k_adjoint<double>, dim3(g), dim3(256), 0, st, dc,
                       static_cast<const double*>(p), static_cast<const double*>(Q), dn,
                       static_cast<double*>(grad));
  else
    hipLaunchKernelGGL(k_adjoint<float>, dim3(g), dim3(256), 0, st, dc,
                       static_cast<const float*>(p), static_cast<const float*>(Q), dn,
                       static_cast<float*>(grad));
  return hipGetLastError();
}

hipError_t hea_site_gain_impl(const HeaConst* dc, const HeaConst& hc, const int32_t* nbr,
                              const uint8_t* lattices, int L, const double* dn, float* u,
                              hipStream_t st) {
  hipLaunchKernelGGL(k_site_gain, dim3((hc.N + 255) / 256, L), dim3(256), 0, st, dc, nbr,
                     lattices, dn, u);
  return hipGetLastError();
}
