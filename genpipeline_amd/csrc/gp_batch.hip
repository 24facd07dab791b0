// Batched SIMP kernels: one launch covers every problem of a batch.
//
// Each kernel kind has a block-offset table off[0..n] (problem b owns blocks
// [off[b], off[b+1]) of the launch); a block finds its problem by binary
// search and runs the same device body as the single-problem kernel on its
// local block range. Problems of different grid sizes share one launch; a
// problem whose solve has finished returns at once. Per-problem reductions
// never mix problems, so every result is bitwise the same as when the
// problem runs alone with the same workgroup shape.
#include "gp_cg_body.hpp"

namespace gp {

#define GP_BATCH_BLOCK(OFF)                                   \
  const int gb_ = xcd_block(blockIdx.x, gridDim.x);           \
  const int b_ = find_prob(bl.OFF, bl.n, gb_);                \
  const BatchProb& P = bl.probs[b_];                          \
  const int lb = gb_ - bl.OFF[b_];                            \
  const int nb = bl.OFF[b_ + 1] - bl.OFF[b_];

__global__ __launch_bounds__(kBlock) void k_b_stiffness(BatchLaunch bl) {
  GP_BATCH_BLOCK(off_elem)
  body_stiffness(P.g.ne, P.xPhys, P.E, P.law, P.penal, P.E0, P.Emin, lb, nb);
}

__global__ __launch_bounds__(kBlock) void k_b_diag(BatchLaunch bl) {
  GP_BATCH_BLOCK(off_node)
  body_diag<float>(P.g, P.ke32, P.E, P.fixmask, P.minv, 2, lb, nb);  // per node
}

__global__ __launch_bounds__(kBlock) void k_b_cg_init(BatchLaunch bl, int use_x0) {
  GP_BATCH_BLOCK(off_vec)
  body_cg_init<float>(P.g.nn, P.f, P.wb[0], use_x0, P.u, P.rb[1], P.sb[0], P.wb[1], P.pc, P.minv,
                      P.fixmask, 1, P.parts, lb, nb);
}

__global__ __launch_bounds__(kBlock) void k_b_init_fin(BatchLaunch bl, double tol, int max_iter) {
  const BatchProb& P = bl.probs[blockIdx.x];
  body_pcg_init_fin(P.parts, bl.off_vec[blockIdx.x + 1] - bl.off_vec[blockIdx.x], tol, max_iter,
                    P.sc);
}

template <int TJ, int TK>
__global__ __launch_bounds__(TJ* TK) void k_b_ku(BatchLaunch bl) {
  GP_BATCH_BLOCK(off_cg)
  (void)nb;
  __shared__ CornerLds<float, TJ * TK> cv;
  body_ku_sweep<float, TJ, TK>(P.g, P.ni, P.ntj, P.ntk, P.mc, P.E, P.u, P.wb[0], P.fixmask, lb,
                               cv);
}

template <int TJ, int TK>
__global__ __launch_bounds__(whole_waves(TJ* TK)) void k_b_cg(BatchLaunch bl, int k) {
  GP_BATCH_BLOCK(off_cg)
  const int c = k & 1, o = c ^ 1;
  __shared__ CornerLds<float, TJ * TK> cv;
  body_cg_modal<float, TJ, TK>(P.g, P.ni, P.ntj, P.ntk, P.mc, P.E, P.rb[o], P.wb[o], P.sb[c],
                               P.minv, P.rb[c], P.wb[c], P.sb[o], P.u, P.pc, P.fixmask, P.sc,
                               P.parts, k, lb, nb, cv);
}

__global__ __launch_bounds__(kBlock) void k_b_sens(BatchLaunch bl) {
  GP_BATCH_BLOCK(off_elem)
  body_sens<float, true>(P.g, P.ke64, P.mc64, P.xPhys, P.u, P.law, P.penal, P.E0, P.Emin, P.dc, 0,
                         nullptr, P.sc_oc, lb, nb);
}

// mode 1: dcf = H^T (dc/Hs); mode 0: xphys_new = (H xnew)/Hs
__global__ __launch_bounds__(kBlock) void k_b_filter(BatchLaunch bl, int mode) {
  GP_BATCH_BLOCK(off_elem)
  if (mode)
    body_filter(P.g, P.st, P.Hs, P.dc, P.dcf, 1, lb, nb);
  else
    body_filter(P.g, P.st, P.Hs, P.xnew, P.xphys_new, 0, lb, nb);
}

__global__ __launch_bounds__(kBlock) void k_b_oc_max(BatchLaunch bl) {
  GP_BATCH_BLOCK(off_oc)
  if (P.oc_variant == GP_OC_GPU) body_oc_max(P.g.ne, P.dcf, P.parts.M, lb, nb);
}

__global__ __launch_bounds__(kBlock) void k_b_oc_init_fin(BatchLaunch bl) {
  const BatchProb& P = bl.probs[blockIdx.x];
  body_oc_init_fin(P.parts, bl.off_oc[blockIdx.x + 1] - bl.off_oc[blockIdx.x], P.oc_variant,
                   P.volfrac, P.move, P.sc_oc);
}

__global__ __launch_bounds__(kBlock) void k_b_oc_step(BatchLaunch bl, int s) {
  GP_BATCH_BLOCK(off_oc)
  body_oc_step(s, P.g.ne, P.x, P.dcf, P.w, P.pres, P.n_pres, P.parts, P.sc_oc, lb, nb);
}

__global__ __launch_bounds__(kBlock) void k_b_oc_final(BatchLaunch bl) {
  GP_BATCH_BLOCK(off_elem)
  body_oc_final(P.g.ne, P.x, P.dcf, P.pres, P.sc_oc, P.xnew, lb, nb);
}

__global__ __launch_bounds__(kBlock) void k_b_set_pres(BatchLaunch bl) {
  GP_BATCH_BLOCK(off_elem)
  if (P.pres != nullptr) body_set_preserved(P.g.ne, P.pres, P.xphys_new, lb, nb);
}

// x <- xnew, xPhys <- xphys_new
__global__ __launch_bounds__(kBlock) void k_b_commit(BatchLaunch bl) {
  GP_BATCH_BLOCK(off_elem)
  for (int e = lb * kBlock + threadIdx.x; e < P.g.ne; e += nb * kBlock) {
    P.x[e] = P.xnew[e];
    P.xPhys[e] = P.xphys_new[e];
  }
}

__global__ __launch_bounds__(kBlock) void k_b_dot(BatchLaunch bl) {
  GP_BATCH_BLOCK(off_vec)
  body_dot<double, float>(P.g.ndof, P.f, P.u, P.parts.S, lb, nb);
}

__global__ __launch_bounds__(kBlock) void k_b_sum_fin(BatchLaunch bl) {
  const BatchProb& P = bl.probs[blockIdx.x];
  body_sum_fin(P.parts.S, bl.off_vec[blockIdx.x + 1] - bl.off_vec[blockIdx.x], &P.sc_oc->sum);
}

// dir 0: u32 <- (float) u64 ; dir 1: u64 <- (double) u32
__global__ __launch_bounds__(kBlock) void k_b_cast(BatchLaunch bl, int dir) {
  GP_BATCH_BLOCK(off_vec)
  for (long long d = lb * (long long)kBlock + threadIdx.x; d < P.g.ndof;
       d += (long long)nb * kBlock) {
    if (dir == 0)
      P.u[d] = (float)P.u64[d];
    else
      P.u64[d] = (double)P.u[d];
  }
}

// status of every problem into one contiguous array (one copy to the host)
__global__ void k_b_status(BatchLaunch bl, BatchStatus* out) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= bl.n) return;
  const DevScalars* s = bl.probs[b].sc;
  const DevScalars* so = bl.probs[b].sc_oc;
  out[b].done = s->done;
  out[b].iter = s->iter;
  out[b].status = s->status;
  out[b].oc_steps = so->oc_steps;
  out[b].rr = s->rr;
  out[b].bb = s->bb;
  out[b].sum = so->sum;
  out[b].lmid = so->lmid;
  out[b].nan_flag = so->nan_flag;
}

__global__ void k_b_reset(BatchLaunch bl) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b < bl.n) bl.probs[b].sc_oc->nan_flag = 0;
}

// ------------------------------------------------------------- launchers
static dim3 G(int blocks) { return dim3(blocks > 0 ? blocks : 1); }

void batch_stiffness(const BatchLaunch& bl, hipStream_t st) {
  hipLaunchKernelGGL(k_b_stiffness, G(bl.tot_elem), dim3(kBlock), 0, st, bl);
}
void batch_pcg_setup(const BatchLaunch& bl, bool use_x0, double tol, int max_iter,
                     hipStream_t st) {
  hipLaunchKernelGGL(k_b_diag, G(bl.tot_node), dim3(kBlock), 0, st, bl);
  if (use_x0) {
#define GP_BT(TJ_, TK_)                                                                     \
  if (bl.tj == TJ_ && bl.tk == TK_)                                                         \
    hipLaunchKernelGGL((k_b_ku<TJ_, TK_>), G(bl.tot_cg), dim3(TJ_ * TK_), 0, st, bl);
    GP_BT(8, 32) GP_BT(16, 16) GP_BT(14, 18) GP_BT(6, 34) GP_BT(14, 34) GP_BT(4, 64) GP_BT(10, 26) GP_BT(7, 34)
#undef GP_BT
  }
  hipLaunchKernelGGL(k_b_cg_init, G(bl.tot_vec), dim3(kBlock), 0, st, bl, use_x0 ? 1 : 0);
  hipLaunchKernelGGL(k_b_init_fin, G(bl.n), dim3(kBlock), 0, st, bl, tol, max_iter);
}
void batch_cg(const BatchLaunch& bl, int k, hipStream_t st) {
#define GP_BT(TJ_, TK_)                                                                     \
  if (bl.tj == TJ_ && bl.tk == TK_)                                                         \
    hipLaunchKernelGGL((k_b_cg<TJ_, TK_>), G(bl.tot_cg), dim3(whole_waves(TJ_ * TK_)), 0, st, \
                       bl, k);
  GP_BT(8, 32) GP_BT(16, 16) GP_BT(14, 18) GP_BT(6, 34) GP_BT(14, 34) GP_BT(4, 64) GP_BT(10, 26) GP_BT(7, 34)
#undef GP_BT
}
void batch_post_solve(const BatchLaunch& bl, int oc_max_steps, hipStream_t st) {
  hipLaunchKernelGGL(k_b_reset, G((bl.n + 63) / 64), dim3(64), 0, st, bl);
  hipLaunchKernelGGL(k_b_sens, G(bl.tot_elem), dim3(kBlock), 0, st, bl);
  hipLaunchKernelGGL(k_b_filter, G(bl.tot_elem), dim3(kBlock), 0, st, bl, 1);
  hipLaunchKernelGGL(k_b_oc_max, G(bl.tot_oc), dim3(kBlock), 0, st, bl);
  hipLaunchKernelGGL(k_b_oc_init_fin, G(bl.n), dim3(kBlock), 0, st, bl);
  for (int j = 0; j < oc_launches(oc_max_steps); ++j)
    hipLaunchKernelGGL(k_b_oc_step, G(bl.tot_oc), dim3(kBlock), 0, st, bl, j);
  hipLaunchKernelGGL(k_b_oc_final, G(bl.tot_elem), dim3(kBlock), 0, st, bl);
  hipLaunchKernelGGL(k_b_filter, G(bl.tot_elem), dim3(kBlock), 0, st, bl, 0);
  hipLaunchKernelGGL(k_b_set_pres, G(bl.tot_elem), dim3(kBlock), 0, st, bl);
  hipLaunchKernelGGL(k_b_commit, G(bl.tot_elem), dim3(kBlock), 0, st, bl);
}
void batch_compliance(const BatchLaunch& bl, hipStream_t st) {
  hipLaunchKernelGGL(k_b_dot, G(bl.tot_vec), dim3(kBlock), 0, st, bl);
  hipLaunchKernelGGL(k_b_sum_fin, G(bl.n), dim3(kBlock), 0, st, bl);
}
void batch_cast(const BatchLaunch& bl, int dir, hipStream_t st) {
  hipLaunchKernelGGL(k_b_cast, G(bl.tot_vec), dim3(kBlock), 0, st, bl, dir);
}
void batch_status(const BatchLaunch& bl, BatchStatus* out, hipStream_t st) {
  hipLaunchKernelGGL(k_b_status, G((bl.n + 63) / 64), dim3(64), 0, st, bl, out);
}

}  // namespace gp
