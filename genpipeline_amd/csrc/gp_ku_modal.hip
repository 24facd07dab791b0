// Matrix-free hex8 K.u in the Walsh-Hadamard (modal) basis, swept along i.
//
// For the trilinear hex, KE = T Khat T^T / 64 with T = kron(W, I3) and W the
// 8x8 corner/mode sign matrix (W[c,m] = prod_d s_d^{m_d}, s_d = 2 c_d - 1).
// Khat is 45-sparse and, for an isotropic material, carries 8 distinct
// values (normal, lateral, shear, two bilinear-pair, two bilinear-triple and
// one trilinear coefficient). The element product is therefore
//   b = T^T u_e      (tensor-product sums/differences, 3 comps x 7 modes)
//   v = E_e Khat b / 64
//   y_e = T v        (the same butterfly, transposed)
// and along i the plane-wise 2-D transforms are shared by neighbouring
// elements: each step transforms one new node plane.
//
// Work split: a workgroup holds TJ x TK element columns (j,k) and sweeps NI
// node planes of i. Thread (a,b) owns element column (J0-1+a, K0-1+b); the
// (TJ-1)(TK-1) node columns (J0+a, K0+b), a<TJ-1, b<TK-1, are written by
// this workgroup and no other. Each output node sums its four element
// columns' corner values from LDS in a fixed order: no atomics, bitwise
// deterministic, independent of how many problems share the launch.
#include "gp_modal.hpp"

namespace gp {

template <typename T, int TJ, int TK, int MODE>
__global__ __launch_bounds__(TJ* TK) void k_ku_modal(Grid g, int NI, int ntj, int ntk,
                                                    ModalCoef<T> kc,
                                                    const float* __restrict__ E,
                                                    const T* __restrict__ u,
                                                    const T* __restrict__ u2,
                                                    T* __restrict__ y, T* __restrict__ pout,
                                                    const uint8_t* __restrict__ fixmask,
                                                    DevScalars* sc, Parts parts, int it,
                                                    int nV) {
  __shared__ CornerLds<T, TJ * TK> cv;
  if constexpr (MODE == 0)
    body_ku_sweep<T, TJ, TK>(g, NI, ntj, ntk, kc, E, u, y, fixmask,
                             xcd_block(blockIdx.x, gridDim.x), cv);
  else
    body_ku_modal<T, TJ, TK, MODE>(g, NI, ntj, ntk, kc, E, u, u2, y, pout, fixmask, sc, parts, it,
                                   nV, xcd_block(blockIdx.x, gridDim.x), cv);
}

template <int TJ, int TK>
static void launch_modal_t(const Grid& g, int NI, const void* kc, int prec, const float* E,
                           const void* u, const void* u2, void* y, void* pout,
                           const uint8_t* fixmask, DevScalars* sc, Parts parts, int it, int nV,
                           int mode, hipStream_t st) {
  const int ntj = (g.ny + 1 + (TJ - 2)) / (TJ - 1);
  const int ntk = (g.nz + 1 + (TK - 2)) / (TK - 1);
  const int nti = (g.nx + 1 + NI - 1) / NI;
  const dim3 grid(ntj * ntk * nti), block(TJ * TK);
  if (prec == GP_F32) {
    const ModalCoef<float>& c = *reinterpret_cast<const ModalCoef<float>*>(kc);
    if (mode == 1)
      hipLaunchKernelGGL((k_ku_modal<float, TJ, TK, 1>), grid, block, 0, st, g, NI, ntj, ntk, c,
                         E, (const float*)u, (const float*)u2, (float*)y, (float*)pout, fixmask,
                         sc, parts, it, nV);
    else
      hipLaunchKernelGGL((k_ku_modal<float, TJ, TK, 0>), grid, block, 0, st, g, NI, ntj, ntk, c,
                         E, (const float*)u, (const float*)u2, (float*)y, (float*)pout, fixmask,
                         sc, parts, it, nV);
  } else {
    const ModalCoef<double>& c = *reinterpret_cast<const ModalCoef<double>*>(kc);
    if (mode == 1)
      hipLaunchKernelGGL((k_ku_modal<double, TJ, TK, 1>), grid, block, 0, st, g, NI, ntj, ntk, c,
                         E, (const double*)u, (const double*)u2, (double*)y, (double*)pout,
                         fixmask, sc, parts, it, nV);
    else
      hipLaunchKernelGGL((k_ku_modal<double, TJ, TK, 0>), grid, block, 0, st, g, NI, ntj, ntk, c,
                         E, (const double*)u, (const double*)u2, (double*)y, (double*)pout,
                         fixmask, sc, parts, it, nV);
  }
}

int modal_blocks(const Grid& g, const KuTile& t) {
  const int ntj = (g.ny + 1 + (t.tj - 2)) / (t.tj - 1);
  const int ntk = (g.nz + 1 + (t.tk - 2)) / (t.tk - 1);
  const int nti = (g.nx + 1 + t.ni - 1) / t.ni;
  return ntj * ntk * nti;
}

bool launch_ku_modal(const Grid& g, const KuTile& t, const void* kc, int prec, const float* E,
                     const void* u, const void* u2, void* y, void* pout, const uint8_t* fixmask,
                     DevScalars* sc, Parts parts, int it, int nV, int mode, hipStream_t st) {
#define GP_TILE(TJ_, TK_)                                                                    \
  if (t.tj == TJ_ && t.tk == TK_) {                                                          \
    launch_modal_t<TJ_, TK_>(g, t.ni, kc, prec, E, u, u2, y, pout, fixmask, sc, parts, it, nV, \
                             mode, st);                                                      \
    return true;                                                                             \
  }
  GP_TILE(8, 32)
  GP_TILE(16, 16)
  GP_TILE(14, 18)
  GP_TILE(6, 34)
  GP_TILE(14, 34)
  GP_TILE(4, 64)
  GP_TILE(10, 26)
  GP_TILE(7, 34)
#undef GP_TILE
  return false;
}

}  // namespace gp
