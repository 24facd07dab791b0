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
// Every sum runs in a fixed order, so the start is bitwise reproducible. The
// bodies (gp_recycle.hpp) also run as phases of the device scheduler's jobs.
#include "gp_recycle.hpp"

namespace gp {

namespace {

template <int N>
__global__ __launch_bounds__(kBlock) void k_gram(RecycleSet rs, const double* __restrict__ f,
                                                long long nd, double* __restrict__ part) {
  __shared__ GramLds gl;
  body_gram<N>(rs, f, nd, part, blockIdx.x, gridDim.x, gl);
}

template <int N>
__global__ __launch_bounds__(kBlock) void k_gram_solve(int nb, const double* __restrict__ part,
                                                      double* __restrict__ coef) {
  __shared__ GramLds gl;
  body_gram_solve<N>(nb, part, coef, gl);
}

template <int N>
__global__ __launch_bounds__(kBlock) void k_combine(RecycleSet rs, const double* __restrict__ coef,
                                                   long long nd, float* __restrict__ u0) {
  body_combine<N>(rs, coef, nd, u0, nullptr, nullptr, blockIdx.x, gridDim.x);
}

template <int N>
void launch_n(const RecycleSet& rs, const double* f, long long nd, double* part, double* coef,
              float* u0, hipStream_t st) {
  const int nb = gram_blocks(nd);
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
