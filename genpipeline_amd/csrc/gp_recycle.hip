// Recycled starting guess for a sequence of related solves.
//
// Consecutive OC iterations solve K(E_k) u = f with slowly changing moduli.
// Instead of starting PCG from the last solution alone, the start is the
// Galerkin (energy-optimal) combination of the last n solutions
// V = [v_1 .. v_n]:
//   u0 = V c,   (V^T K V) c = V^T f,
// which minimises the K-norm of the error over span(V) (span(V) contains the
// last solution, so u0 is never worse than it in that norm). Three launches:
//   k_gram      FP64 per-block partials of v_i.(K v_j) (i <= j) and f.v_i,
//   k_gram_solve one block: fixed-order sums, then a Cholesky factorisation
//               with diagonal pivoting that stops at pivots below 1e-12 of
//               the largest diagonal (nearly dependent solutions are left
//               out), and the triangular solves for c,
//   k_combine   u0 = sum_i c_i v_i in FP64, rounded once to float.
// Every sum runs in a fixed order, so the start is bitwise reproducible.
#pragma clang fp contract(on)
#include "gp_bodies.hpp"

namespace gp {

namespace {

constexpr int kGramMax = kRecycleMax * (kRecycleMax + 1) / 2 + kRecycleMax;

constexpr int gram_vals(int n) { return n * (n + 1) / 2 + n; }

// Index of v_i.(K v_j), i <= j, in the packed list; f.v_i follow at
// n(n+1)/2 + i.
__host__ __device__ inline int gram_idx(int n, int i, int j) {
  return i * n - i * (i - 1) / 2 + (j - i);
}

template <int N>
__global__ __launch_bounds__(kBlock) void k_gram(RecycleSet rs, const double* __restrict__ f,
                                                long long nd, double* __restrict__ part) {
  constexpr int NV = gram_vals(N);
  double acc[NV];
#pragma unroll
  for (int t = 0; t < NV; ++t) acc[t] = 0.0;
  const long long stride = (long long)gridDim.x * kBlock;
  for (long long d = blockIdx.x * (long long)kBlock + threadIdx.x; d < nd; d += stride) {
    double v[N], kv[N];
#pragma unroll
    for (int i = 0; i < N; ++i) {
      v[i] = (double)rs.v[i][d];
      kv[i] = (double)rs.kv[i][d];
    }
    const double fd = f[d];
    int t = 0;
#pragma unroll
    for (int i = 0; i < N; ++i)
#pragma unroll
      for (int j = i; j < N; ++j) acc[t++] += v[i] * kv[j];
#pragma unroll
    for (int i = 0; i < N; ++i) acc[t++] += fd * v[i];
  }
  __shared__ double sh[NV][kBlock / 64];
#pragma unroll
  for (int t = 0; t < NV; ++t) {
    const double s = wave_sum_nb<kBlock>(acc[t]);
    if ((threadIdx.x & 63) == 0) sh[t][threadIdx.x >> 6] = s;
  }
  __syncthreads();
  if (threadIdx.x < NV) {
    double s = 0.0;
    for (int w = 0; w < kBlock / 64; ++w) s += sh[threadIdx.x][w];
    part[(long long)threadIdx.x * gridDim.x + blockIdx.x] = s;
  }
}

// One block: every thread sums its share of the partials of each value
// (blocks t, t + kBlock, ...), then wave sums and the waves in order give the
// Gram values; thread 0 solves the small system with register arrays.
template <int N>
__global__ __launch_bounds__(kBlock) void k_gram_solve(int nb, const double* __restrict__ part,
                                                      double* __restrict__ coef) {
  constexpr int NV = gram_vals(N);
  double acc[NV];
#pragma unroll
  for (int t = 0; t < NV; ++t) acc[t] = 0.0;
  for (int b = threadIdx.x; b < nb; b += kBlock) {
#pragma unroll
    for (int t = 0; t < NV; ++t) acc[t] += part[(long long)t * nb + b];
  }
  __shared__ double sh[NV][kBlock / 64];
#pragma unroll
  for (int t = 0; t < NV; ++t) {
    const double v = wave_sum_nb<kBlock>(acc[t]);
    if ((threadIdx.x & 63) == 0) sh[t][threadIdx.x >> 6] = v;
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  double val[NV];
#pragma unroll
  for (int t = 0; t < NV; ++t) {
    double v = 0.0;
    for (int w = 0; w < kBlock / 64; ++w) v += sh[t][w];
    val[t] = v;
  }
  double M[N][N], r[N], L[N][N], res[N], c[N], z[N];
  int sel[N];
  bool used[N];
  double dmax = 0.0;
#pragma unroll
  for (int i = 0; i < N; ++i) {
#pragma unroll
    for (int j = i; j < N; ++j) M[i][j] = M[j][i] = val[gram_idx(N, i, j)];
    r[i] = val[N * (N + 1) / 2 + i];
    res[i] = M[i][i];
    used[i] = false;
    c[i] = 0.0;
    sel[i] = 0;
    dmax = fmax(dmax, M[i][i]);
#pragma unroll
    for (int k = 0; k < N; ++k) L[i][k] = 0.0;
  }
  // Cholesky with diagonal pivoting: column m takes the unused vector with
  // the largest remaining (Schur complement) diagonal
  const double drop = 1e-12 * dmax;
  int m = 0;
#pragma unroll
  for (int col = 0; col < N; ++col) {
    int p = -1;
    double best = 0.0;
#pragma unroll
    for (int i = 0; i < N; ++i)
      if (!used[i] && res[i] > best) {
        best = res[i];
        p = i;
      }
    if (p < 0 || !(best > drop)) break;
    used[p] = true;
    sel[col] = p;
    const double lpp = sqrt(best);
#pragma unroll
    for (int q = 0; q < N; ++q) {
      if (q == p) L[q][col] = lpp;
      if (used[q]) continue;
      double s = M[q][p];
#pragma unroll
      for (int k = 0; k < col; ++k) {
        double lp = 0.0;
#pragma unroll
        for (int i = 0; i < N; ++i) lp = (i == p) ? L[i][k] : lp;
        s -= L[q][k] * lp;
      }
      L[q][col] = s / lpp;
      res[q] -= L[q][col] * L[q][col];
    }
    m = col + 1;
  }
  // (L_s L_s^T) y = r_s over the selected vectors, rows sel[0..m)
  for (int a = 0; a < m; ++a) {
    double s = r[sel[a]];
    for (int k = 0; k < a; ++k) s -= L[sel[a]][k] * z[k];
    z[a] = s / L[sel[a]][a];
  }
  for (int a = m - 1; a >= 0; --a) {
    double s = z[a];
    for (int k = a + 1; k < m; ++k) s -= L[sel[k]][a] * c[sel[k]];
    c[sel[a]] = s / L[sel[a]][a];
  }
  for (int i = 0; i < kRecycleMax; ++i) coef[i] = i < N ? c[i] : 0.0;
  coef[kRecycleMax] = (double)m;  // vectors kept
}

template <int N>
__global__ __launch_bounds__(kBlock) void k_combine(RecycleSet rs, const double* __restrict__ coef,
                                                   long long nd, float* __restrict__ u0) {
  double c[N];
#pragma unroll
  for (int i = 0; i < N; ++i) c[i] = coef[i];
  const long long stride = (long long)gridDim.x * kBlock;
  for (long long d = blockIdx.x * (long long)kBlock + threadIdx.x; d < nd; d += stride) {
    double s = 0.0;
#pragma unroll
    for (int i = 0; i < N; ++i) s += c[i] * (double)rs.v[i][d];
    u0[d] = (float)s;
  }
}

template <int N>
void launch_n(const RecycleSet& rs, const double* f, long long nd, double* part, double* coef,
              float* u0, hipStream_t st) {
  int nb = grid_blocks(nd);
  nb = nb < kVecBlocks ? nb : kVecBlocks;
  hipLaunchKernelGGL((k_gram<N>), dim3(nb), dim3(kBlock), 0, st, rs, f, nd, part);
  hipLaunchKernelGGL((k_gram_solve<N>), dim3(1), dim3(kBlock), 0, st, nb, part, coef);
  hipLaunchKernelGGL((k_combine<N>), dim3(nb), dim3(kBlock), 0, st, rs, coef, nd, u0);
}

}  // namespace

size_t recycle_part_doubles() { return (size_t)kGramMax * kVecBlocks + kRecycleMax + 1; }

bool launch_recycle_guess(const RecycleSet& rs, const double* f, long long nd, double* part,
                          double* coef, float* u0, hipStream_t st) {
  switch (rs.n) {
    case 1: launch_n<1>(rs, f, nd, part, coef, u0, st); return true;
    case 2: launch_n<2>(rs, f, nd, part, coef, u0, st); return true;
    case 3: launch_n<3>(rs, f, nd, part, coef, u0, st); return true;
    case 4: launch_n<4>(rs, f, nd, part, coef, u0, st); return true;
    case 5: launch_n<5>(rs, f, nd, part, coef, u0, st); return true;
    case 6: launch_n<6>(rs, f, nd, part, coef, u0, st); return true;
    case 7: launch_n<7>(rs, f, nd, part, coef, u0, st); return true;
    case 8: launch_n<8>(rs, f, nd, part, coef, u0, st); return true;
    default: return false;
  }
}

}  // namespace gp
