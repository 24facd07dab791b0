// One-kernel-per-iteration Jacobi-PCG on the modal hex8 K.u sweep.
//
// The recurrence is the Chronopoulos-Gear form of preconditioned CG: both
// inner products of an iteration, gamma = (r,z) and delta = (z,Az), come from
// one pass, so a single global reduction separates iterations:
//   z_k = M r_k, w_k = A z_k, gamma_k = (r_k,z_k), delta_k = (z_k,w_k)
//   beta_k  = gamma_k / gamma_{k-1}                       (beta_0 = 0)
//   alpha_k = gamma_k / (delta_k - beta_k gamma_k / alpha_{k-1})
//   p_k = z_k + beta_k p_{k-1};  s_k = w_k + beta_k s_{k-1}   (s_k = A p_k)
//   x_{k+1} = x_k + alpha_k p_k;  r_{k+1} = r_k - alpha_k s_k
// In exact arithmetic these are the iterates of standard Jacobi-PCG; the
// denominator of alpha_k equals p_k^T A p_k and its sign is the breakdown
// test.
//
// Kernel k first reduces the partials of kernel k-1 (gamma, delta, rho =
// (r,r)), applies the convergence test on rho_{k-1}, and forms alpha_{k-1},
// beta_{k-1}. While sweeping it builds s_{k-1}, r_k and z_k on every loaded
// node (halo included, since z_k is needed at all element corners), computes
// w_k = A z_k, and at each node it owns writes r_k, s_{k-1}, w_k,
// p_{k-1} = z_{k-1} + beta p_{k-2} and x_k = x_{k-1} + alpha p_{k-1}.
// r, w and s alternate between two buffers by iteration parity so that no
// block reads a value another block has already overwritten.
//
// Storage: GP_F32 / GP_F64 keep every vector in the arithmetic type. GP_BF16
// / GP_F16 store the iterate x and the search direction p in 16 bits (float
// arithmetic; r, w, s and M^-1 in float; FP64 dot products); that variant
// runs as the inner solve of FP64 iterative refinement.
#include "gp_cg_body.hpp"

namespace gp {

// One workgroup of TJ x TK lanes rounded up to whole waves (the block sums
// use the DPP wave sum); the extra lanes idle.
template <typename T, int TJ, int TK, typename S, int XP>
__global__ __launch_bounds__(whole_waves(TJ* TK)) void k_cg_modal(Grid g, int NI, int ntj, int ntk,
                                                    ModalCoef<T> kc,
                                                    const float* __restrict__ E,
                                                    const T* __restrict__ rp,
                                                    const T* __restrict__ wp,
                                                    const T* __restrict__ sp,
                                                    const T* __restrict__ minv,
                                                    T* __restrict__ rn, T* __restrict__ wn,
                                                    T* __restrict__ sn, S* __restrict__ x,
                                                    S* __restrict__ p,
                                                    const uint8_t* __restrict__ fixmask,
                                                    DevScalars* sc, Parts parts, int k) {
  __shared__ CornerLds<T, TJ * TK> cv;
  body_cg_modal<T, TJ, TK, S, whole_waves(TJ * TK), true, XP>(
      g, NI, ntj, ntk, kc, E, rp, wp, sp, minv, rn, wn, sn, x, p, fixmask, sc, parts, k,
      xcd_block(blockIdx.x, gridDim.x), gridDim.x, cv);
}

template <int TJ, int TK, typename T, typename S, int XP = 0>
static void launch_cg_ts(const dim3& grid, int NI, int ntj, int ntk, const Grid& g, const void* kc,
                         const float* E, const void* rp, const void* wp, const void* sp,
                         const void* minv, void* rn, void* wn, void* sn, void* x, void* p,
                         const uint8_t* fixmask, DevScalars* sc, Parts parts, int k,
                         hipStream_t st) {
  hipLaunchKernelGGL((k_cg_modal<T, TJ, TK, S, XP>), grid, dim3(whole_waves(TJ * TK)), 0, st, g,
                     NI, ntj, ntk,
                     *reinterpret_cast<const ModalCoef<T>*>(kc), E, (const T*)rp, (const T*)wp,
                     (const T*)sp, (const T*)minv, (T*)rn, (T*)wn, (T*)sn, (S*)x, (S*)p, fixmask,
                     sc, parts, k);
}

// Launch k of a solve: with two_step, k = 0 updates x and p as usual, odd k
// leaves them, even k >= 2 applies both pending updates.
template <int TJ, int TK, typename T, typename S = T>
static void launch_cg_xp(bool two_step, const dim3& grid, int NI, int ntj, int ntk, const Grid& g,
                         const void* kc, const float* E, const void* rp, const void* wp,
                         const void* sp, const void* minv, void* rn, void* wn, void* sn, void* x,
                         void* p, const uint8_t* fixmask, DevScalars* sc, Parts parts, int k,
                         hipStream_t st) {
#define GP_XARGS grid, NI, ntj, ntk, g, kc, E, rp, wp, sp, minv, rn, wn, sn, x, p, fixmask, sc, parts, k, st
  if (!two_step || k == 0)
    launch_cg_ts<TJ, TK, T, S, 0>(GP_XARGS);
  else if (k & 1)
    launch_cg_ts<TJ, TK, T, S, 1>(GP_XARGS);
  else
    launch_cg_ts<TJ, TK, T, S, 2>(GP_XARGS);
#undef GP_XARGS
}

template <int TJ, int TK>
static void launch_cg_t(const Grid& g, int NI, const void* kc, int prec, const float* E,
                        const void* rp, const void* wp, const void* sp, const void* minv,
                        void* rn, void* wn, void* sn, void* x, void* p, const uint8_t* fixmask,
                        DevScalars* sc, Parts parts, int k, hipStream_t st, bool two_step) {
  const int ntj = (g.ny + 1 + (TJ - 2)) / (TJ - 1);
  const int ntk = (g.nz + 1 + (TK - 2)) / (TK - 1);
  const int nti = (g.nx + 1 + NI - 1) / NI;
  const dim3 grid(ntj * ntk * nti);
#define GP_ARGS \
  grid, NI, ntj, ntk, g, kc, E, rp, wp, sp, minv, rn, wn, sn, x, p, fixmask, sc, parts, k, st
  switch (prec) {
    case GP_F32: launch_cg_xp<TJ, TK, float>(two_step, GP_ARGS); break;
    case GP_F64: launch_cg_xp<TJ, TK, double>(two_step, GP_ARGS); break;
    case GP_F16: launch_cg_xp<TJ, TK, float, _Float16>(two_step, GP_ARGS); break;
    case GP_BF16: launch_cg_xp<TJ, TK, float, bf16>(two_step, GP_ARGS); break;
    default: break;
  }
#undef GP_ARGS
}

bool launch_cg_modal(const Grid& g, const KuTile& t, const void* kc, int prec, const float* E,
                     const void* rp, const void* wp, const void* sp, const void* minv, void* rn,
                     void* wn, void* sn, void* x, void* p, const uint8_t* fixmask,
                     DevScalars* sc, Parts parts, int k, hipStream_t st, bool two_step) {
#define GP_TILE(TJ_, TK_)                                                                   \
  if (t.tj == TJ_ && t.tk == TK_) {                                                         \
    launch_cg_t<TJ_, TK_>(g, t.ni, kc, prec, E, rp, wp, sp, minv, rn, wn, sn, x, p, fixmask, \
                          sc, parts, k, st, two_step);                                      \
    return true;                                                                            \
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

// The pending x and p updates of an odd final iteration of a two-step solve
// (see launch_cg_modal).
template <typename T, typename S>
__global__ __launch_bounds__(kBlock) void k_cg_xp_fix(int nn, const DevScalars* __restrict__ sc,
                                                     const T* __restrict__ r0,
                                                     const T* __restrict__ minv,
                                                     S* __restrict__ x, S* __restrict__ p) {
  body_cg_xp_fix<T, S>(nn, sc, r0, minv, x, p, blockIdx.x, gridDim.x);
}

template <typename T, typename S>
void launch_cg_xp_fix(const Grid& g, const DevScalars* sc, const T* r0, const T* minv, S* x,
                      S* p, hipStream_t st) {
  hipLaunchKernelGGL((k_cg_xp_fix<T, S>), dim3(vec_blocks(g)), dim3(kBlock), 0, st, g.nn, sc,
                     r0, minv, x, p);
}
#define GP_INST(T, S)                                                                        \
  template void launch_cg_xp_fix<T, S>(const Grid&, const DevScalars*, const T*, const T*, S*, \
                                       S*, hipStream_t);
GP_INST(float, float)
GP_INST(double, double)
GP_INST(float, _Float16)
GP_INST(float, bf16)
#undef GP_INST

template <typename T, typename S>
__global__ __launch_bounds__(kBlock) void k_cg_init(int nn, const double* __restrict__ f,
                                                   const T* __restrict__ Ax, int use_x0,
                                                   S* __restrict__ x, T* __restrict__ r,
                                                   T* __restrict__ s0, T* __restrict__ w1,
                                                   S* __restrict__ p, const T* __restrict__ minv,
                                                   const uint8_t* __restrict__ fixmask,
                                                   Parts parts) {
  // float sweeps read M^-1 per node (launch_diag_node)
  body_cg_init<T, S>(nn, f, Ax, use_x0, x, r, s0, w1, p, minv, fixmask, sizeof(T) == 4 ? 1 : 0,
                     parts, blockIdx.x, gridDim.x);
}

template <typename T, typename S>
void launch_cg_init(const Grid& g, const double* f, const T* Ax, bool use_x0, S* x, T* r, T* s0,
                    T* w1, S* p, const T* minv, const uint8_t* fixmask, Parts parts,
                    DevScalars* sc, double tol, int max_iter, hipStream_t st) {
  const int nV = vec_blocks(g);
  hipLaunchKernelGGL((k_cg_init<T, S>), dim3(nV), dim3(kBlock), 0, st, g.nn, f, Ax,
                     use_x0 ? 1 : 0, x, r, s0, w1, p, minv, fixmask, parts);
  launch_pcg_init_fin(parts, nV, tol, max_iter, sc, st);
}

#define GP_INST(T, S)                                                                         \
  template void launch_cg_init<T, S>(const Grid&, const double*, const T*, bool, S*, T*, T*, \
                                     T*, S*, const T*, const uint8_t*, Parts, DevScalars*,   \
                                     double, int, hipStream_t);
GP_INST(float, float)
GP_INST(double, double)
GP_INST(float, _Float16)
GP_INST(float, bf16)
#undef GP_INST

}  // namespace gp
