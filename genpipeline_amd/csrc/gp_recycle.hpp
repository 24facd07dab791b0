// Device bodies of the recycled start (the Galerkin combination of stored
// solutions; see gp_recycle.hip). They are written against a block index lb
// out of nb blocks and take their LDS from the caller, so the single-problem
// kernels (one block per lb) and the scheduler's phases (virtual blocks
// looped on a job's blocks, LDS lent by the sweep's corner buffer) run the
// same code on the same block layout: the start is bitwise the same in both.
#pragma once
#pragma clang fp contract(on)
#include "gp_bodies.hpp"

namespace gp {

constexpr int kGramMax = kRecycleMax * (kRecycleMax + 1) / 2 + kRecycleMax;

constexpr int gram_vals(int n) { return n * (n + 1) / 2 + n; }

// Index of v_i.(K v_j), i <= j, in the packed list; f.v_i follow at
// n(n+1)/2 + i.
__host__ __device__ inline int gram_idx(int n, int i, int j) {
  return i * n - i * (i - 1) / 2 + (j - i);
}

// Blocks of the Gram and combination passes: the layout of the partials.
__host__ __device__ inline int gram_blocks(long long nd) {
  long long b = (nd + kBlock - 1) / kBlock;
  if (b < 1) b = 1;
  return b < kVecBlocks ? (int)b : kVecBlocks;
}

// LDS of the bodies: wave sums of every value, then thread 0's small solve.
struct GramLds {
  double sh[kGramMax][kBlock / 64];
  double val[kGramMax];
  double M[kRecycleMax][kRecycleMax], L[kRecycleMax][kRecycleMax];
  double r[kRecycleMax], res[kRecycleMax], c[kRecycleMax], z[kRecycleMax];
  int sel[kRecycleMax], used[kRecycleMax];
};

// Block lb of nb: FP64 partials of v_i.(K v_j) (i <= j) and f.v_i over the
// DOFs lb*kBlock + t, + nb*kBlock, ... in order, into part[value * nb + lb].
// Ends with a barrier, so a block may run several lb in a row.
template <int N>
__device__ inline void body_gram(const RecycleSet& rs, const double* __restrict__ f,
                                 long long nd, double* __restrict__ part, int lb, int nb,
                                 GramLds& gl) {
  constexpr int NV = gram_vals(N);
  double acc[NV];
#pragma unroll
  for (int t = 0; t < NV; ++t) acc[t] = 0.0;
  const long long stride = (long long)nb * kBlock;
  for (long long d = lb * (long long)kBlock + threadIdx.x; d < nd; d += stride) {
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
#pragma unroll
  for (int t = 0; t < NV; ++t) {
    const double s = wave_sum_nb<kBlock>(acc[t]);
    if ((threadIdx.x & 63) == 0) gl.sh[t][threadIdx.x >> 6] = s;
  }
  __syncthreads();
  if (threadIdx.x < NV) {
    double s = 0.0;
    for (int w = 0; w < kBlock / 64; ++w) s += gl.sh[threadIdx.x][w];
    part[(long long)threadIdx.x * nb + lb] = s;
  }
  __syncthreads();
}

// One block: every thread sums its share of the nb partials of each value
// (partials t, t + kBlock, ...), then wave sums and the waves in order give
// the Gram values; thread 0 solves the small system. Cholesky with diagonal
// pivoting: each column takes the unused solution with the largest remaining
// (Schur complement) diagonal and the factorisation stops at pivots below
// 1e-12 of the largest diagonal, so nearly dependent solutions are left out.
// coef receives c (0 for solutions left out) and then the number kept.
template <int N>
__device__ inline void body_gram_solve(int nb, const double* __restrict__ part,
                                       double* __restrict__ coef, GramLds& gl) {
  constexpr int NV = gram_vals(N);
  double acc[NV];
#pragma unroll
  for (int t = 0; t < NV; ++t) acc[t] = 0.0;
  for (int b = threadIdx.x; b < nb; b += kBlock) {
#pragma unroll
    for (int t = 0; t < NV; ++t) acc[t] += part[(long long)t * nb + b];
  }
#pragma unroll
  for (int t = 0; t < NV; ++t) {
    const double v = wave_sum_nb<kBlock>(acc[t]);
    if ((threadIdx.x & 63) == 0) gl.sh[t][threadIdx.x >> 6] = v;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    auto& M = gl.M;
    auto& L = gl.L;
    for (int t = 0; t < NV; ++t) {
      double v = 0.0;
      for (int w = 0; w < kBlock / 64; ++w) v += gl.sh[t][w];
      gl.val[t] = v;
    }
    double dmax = 0.0;
    for (int i = 0; i < N; ++i) {
      for (int j = i; j < N; ++j) M[i][j] = M[j][i] = gl.val[gram_idx(N, i, j)];
      gl.r[i] = gl.val[N * (N + 1) / 2 + i];
      gl.res[i] = M[i][i];
      gl.used[i] = 0;
      gl.c[i] = 0.0;
      gl.sel[i] = 0;
      dmax = fmax(dmax, M[i][i]);
      for (int k = 0; k < N; ++k) L[i][k] = 0.0;
    }
    const double drop = 1e-12 * dmax;
    int m = 0;
    for (int col = 0; col < N; ++col) {
      int p = -1;
      double best = 0.0;
      for (int i = 0; i < N; ++i)
        if (!gl.used[i] && gl.res[i] > best) {
          best = gl.res[i];
          p = i;
        }
      if (p < 0 || !(best > drop)) break;
      gl.used[p] = 1;
      gl.sel[col] = p;
      const double lpp = sqrt(best);
      L[p][col] = lpp;
      for (int q = 0; q < N; ++q) {
        if (gl.used[q]) continue;
        double s = M[q][p];
        for (int k = 0; k < col; ++k) s -= L[q][k] * L[p][k];
        L[q][col] = s / lpp;
        gl.res[q] -= L[q][col] * L[q][col];
      }
      m = col + 1;
    }
    // (L_s L_s^T) y = r_s over the kept solutions, rows sel[0..m)
    for (int a = 0; a < m; ++a) {
      double s = gl.r[gl.sel[a]];
      for (int k = 0; k < a; ++k) s -= L[gl.sel[a]][k] * gl.z[k];
      gl.z[a] = s / L[gl.sel[a]][a];
    }
    for (int a = m - 1; a >= 0; --a) {
      double s = gl.z[a];
      for (int k = a + 1; k < m; ++k) s -= L[gl.sel[k]][a] * gl.c[gl.sel[k]];
      gl.c[gl.sel[a]] = s / L[gl.sel[a]][a];
    }
    for (int i = 0; i < kRecycleMax; ++i) coef[i] = i < N ? gl.c[i] : 0.0;
    coef[kRecycleMax] = (double)m;
  }
  __syncthreads();
}

// u0 = sum_i c_i v_i in FP64, rounded once to float, over DOFs lb*kBlock + t,
// + nb*kBlock, ... Into u0 if given; into u64 as the float value widened, with
// supported DOFs set to zero, if given (an FP64 start that a refined solve
// uses as it stands).
template <int N>
__device__ inline void body_combine(const RecycleSet& rs, const double* __restrict__ coef,
                                    long long nd, float* __restrict__ u0,
                                    double* __restrict__ u64,
                                    const uint8_t* __restrict__ fixmask, int lb, int nb) {
  double c[N];
#pragma unroll
  for (int i = 0; i < N; ++i) c[i] = coef[i];
  const long long stride = (long long)nb * kBlock;
  for (long long d = lb * (long long)kBlock + threadIdx.x; d < nd; d += stride) {
    double s = 0.0;
#pragma unroll
    for (int i = 0; i < N; ++i) s += c[i] * (double)rs.v[i][d];
    const float u = (float)s;
    if (u0) u0[d] = u;
    if (u64) u64[d] = (fixmask[d / 3] >> (d % 3)) & 1 ? 0.0 : (double)u;
  }
}

}  // namespace gp
