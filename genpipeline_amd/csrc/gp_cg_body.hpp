// Device body of one single-kernel Jacobi-PCG iteration on the modal sweep
// (see gp_cg_modal.hip for the recurrence), shared by the single-problem
// and batched launches.
#pragma once
// Multiply-adds are fused only where one source expression allows it, so the
// rounding of a body does not depend on where the compiler inlines it (the
// single-problem and batched kernels then agree bit for bit).
#pragma clang fp contract(on)
#include <type_traits>

#include "gp_bodies.hpp"
#include "gp_modal.hpp"

namespace gp {

// One node's update for iteration k: s_{k-1} = w_{k-1} + beta s_{k-2},
// r_k = r_{k-1} - alpha s_{k-1}, z_k = M r_k, and z_{k-1} = M r_{k-1} (for p).
template <typename T>
struct NodeCG {
  T r[3], s[3], z[3], zp[3];
};

// Raw inputs of one node (r_{k-1}, w_{k-1}, s_{k-2}, M).
template <typename T>
struct RawNode {
  T r[3], w[3], s[3], m[3];
};

// A node outside the grid reads node 0 instead: such a node is a corner only
// of elements outside the grid, whose modulus is exactly 0, and it is never
// stored or summed, so its (finite) values do not reach any result. Leaving
// the loaded values unselected lets all of a node's loads issue back to back.
// NM: M^-1 is one value per node (float sweeps; body_diag's invert 2), else
// one per DOF.
template <typename T, bool NM>
__device__ inline void load_node(const Grid& g, const T* __restrict__ rp, const T* __restrict__ wp,
                                 const T* __restrict__ sp, const T* __restrict__ m, int ip, int jj,
                                 int kk, RawNode<T>& o) {
  const bool ok = (unsigned)ip <= (unsigned)g.nx && (unsigned)jj <= (unsigned)g.ny &&
                  (unsigned)kk <= (unsigned)g.nz;
  const long long n = ok ? (long long)ip * g.sI + (long long)jj * g.sJ + kk : 0;
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    o.r[q] = rp[3 * n + q];
    o.w[q] = wp[3 * n + q];
    o.s[q] = sp[3 * n + q];
  }
  if constexpr (NM) {
    const T mv = m[n];
#pragma unroll
    for (int q = 0; q < 3; ++q) o.m[q] = mv;
  } else {
#pragma unroll
    for (int q = 0; q < 3; ++q) o.m[q] = m[3 * n + q];
  }
}

template <typename T>
__device__ inline void finish_node(const RawNode<T>& in, T alpha, T beta, NodeCG<T>& o) {
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    const T sn = in.w[q] + beta * in.s[q];
    const T rn = in.r[q] - alpha * sn;
    o.s[q] = sn;
    o.r[q] = rn;
    o.z[q] = in.m[q] * rn;
    o.zp[q] = in.m[q] * in.r[q];
  }
}


// 2-D transform of z over the element column's four corners, read from the
// plane's LDS tile zs[q][aa][bb] (aa, bb = node column - (J0-1, K0-1)).
template <typename T, int TJ, int TK>
__device__ inline void corners_transform(const T (&zs)[3][TJ + 1][TK + 1], int a, int b,
                                         Plane<T>& out) {
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    const T c00 = zs[q][a][b], c01 = zs[q][a][b + 1];
    const T c10 = zs[q][a + 1][b], c11 = zs[q][a + 1][b + 1];
    const T a0 = c00 + c01, d0 = c01 - c00;
    const T a1 = c10 + c11, d1 = c11 - c10;
    out.P[q][0] = a0 + a1;
    out.P[q][1] = d0 + d1;
    out.P[q][2] = a1 - a0;
    out.P[q][3] = d1 - d0;
  }
}

// Block sums of three values (NB threads, NT of them carrying values; the
// others pass 0) into *p0, *p1, *p2: wave sums, then waves in order.
template <int NT, int NB = NT>
__device__ inline void block_sum3_to(double v0, double v1, double v2, double* p0, double* p1,
                                     double* p2) {
  constexpr int NW = (NT + 63) / 64;
  __shared__ double red[3][NW];
  v0 = wave_sum_nb<NB>(v0);
  v1 = wave_sum_nb<NB>(v1);
  v2 = wave_sum_nb<NB>(v2);
  if ((threadIdx.x & 63) == 0) {
    red[0][threadIdx.x >> 6] = v0;
    red[1][threadIdx.x >> 6] = v1;
    red[2][threadIdx.x >> 6] = v2;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
    for (int w = 0; w < NW; ++w) {
      s0 += red[0][w];
      s1 += red[1][w];
      s2 += red[2][w];
    }
    *p0 = s0;
    *p1 = s1;
    *p2 = s2;
  }
}

// Inputs of one node plane that do not depend on alpha, beta.
template <typename T>
struct StageIn {
  RawNode<T> ro, rh;  // owned node, halo node
  T pv[3], xv[3];     // p, x of the owned node (owned planes only)
  T rq[3];            // r two iterations back (two-step x and p update)
  uint8_t fm;         // its support bits
};

// Node loading: thread (a,b) updates its own node column (J0+a, K0+b); the
// TJ+TK+1 halo columns with aa = 0 or bb = 0 are spread over the first
// threads. Each plane's z goes to an LDS tile; element columns then read
// their four corners from it.
// Returns true when the solve is finished at this launch (converged, capped
// or broken down at iteration k-1, or already finished): no work was done.
// T: arithmetic and storage of the recurrence vectors r, w, s (and M^-1);
// S: storage of the iterate x and the search direction p. These two feed only
// the x update, so a 16-bit S limits the accuracy of x (recovered by FP64
// refinement) without disturbing the recurrence.
// NB: threads per block; NT = TJ x TK of them are active and the rest only
// join the barriers and add zero to the block sums (a padded block gives the
// same results as one of exactly NT threads).
// XP: how this launch treats the iterate x and the search direction p, which
// feed only the x update. 0: both updated every launch (p_{k-1}, x_k). 1: not
// touched (no load, no store). 2: two updates at once, p_{k-2}, x_{k-1} from
// the memory's p_{k-3}, x_{k-2} and from r_{k-2}, still in the r buffer this
// launch writes, then p_{k-1}, x_k as with XP 0; each value is formed with the
// same operations as by the launch that would have stored it, so alternating
// XP 1 and 2 gives bitwise the x of XP 0 at every stored launch, moving 18 B
// per node and iteration less (S == T). With a 16-bit S the pair of updates is
// formed in T and rounded to S once, every second iteration.
template <typename T, int TJ, int TK, typename S = T, int NB = whole_waves(TJ * TK),
          bool UNROLL2 = true, int XP = 0>
__device__ inline bool body_cg_modal(const Grid& g, int NI, int ntj, int ntk,
                                     const ModalCoef<T>& kc, const float* __restrict__ E,
                                     const T* __restrict__ rp, const T* __restrict__ wp,
                                     const T* __restrict__ sp, const T* __restrict__ minv,
                                     T* __restrict__ rn, T* __restrict__ wn, T* __restrict__ sn,
                                     S* __restrict__ x, S* __restrict__ p,
                                     const uint8_t* __restrict__ fixmask, DevScalars* sc,
                                     const Parts& parts, int k, int lb, int nb,
                                     CornerLds<T, TJ * TK>& cv) {
  constexpr int NT = TJ * TK;
  constexpr int NH = TJ + TK + 1;  // halo node columns (aa = 0 or bb = 0)
  // float sweeps take M^-1 per node (minv has n_nodes entries), FP64 per DOF
  constexpr bool NM = sizeof(T) == 4;
  static_assert(NB >= NT && (NB + 63) / 64 == (NT + 63) / 64, "padding within the last wave");
  static_assert(NB % 64 == 0, "whole waves: the block sums use the DPP wave sum");
  __shared__ T zs[2][3][TJ + 1][TK + 1];
  const int tid = threadIdx.x;
  const bool act = NB == NT || tid < NT;
  const int tl = act ? tid : 0;
  const int a = tl / TK, b = tl - (tl / TK) * TK;
  int bid = lb;
  const int tk_ = bid % ntk;
  bid /= ntk;
  const int tj_ = bid % ntj;
  const int ti_ = bid / ntj;
  const int J0 = tj_ * (TJ - 1), K0 = tk_ * (TK - 1), I0 = ti_ * NI;
  const int je = J0 - 1 + a, ke = K0 - 1 + b;
  const bool col_ok = (unsigned)je < (unsigned)g.ny && (unsigned)ke < (unsigned)g.nz;
  const bool owner =
      act && a < TJ - 1 && b < TK - 1 && (je + 1) <= g.ny && (ke + 1) <= g.nz;
  // halo assignment
  const bool has_h = tid < NH;
  const int ha = tid <= TK ? 0 : tid - TK;
  const int hb = tid <= TK ? tid : 0;

  T alpha = T(0), beta = T(0);
  T alpha2 = T(0), beta2 = T(0);  // alpha_{k-2}, beta_{k-2} (XP 2)
  auto Eat = [&](int ie) -> T {
    const bool ok = col_ok && (unsigned)ie < (unsigned)g.nx;
    const int e = ok ? ie * g.eI + je * g.eJ + ke : 0;
    const float v = E[e];
    return ok ? (T)v : T(0);
  };
  double gk = 0.0, dk = 0.0, rk = 0.0;
  // Loads of one plane that do not depend on alpha, beta: the owned and halo
  // node, and (when the plane is owned) p, x and the support bits.
  auto load_stage = [&](int ip, bool own_plane, StageIn<T>& in) {
    const bool fin = owner && own_plane;
    const long long n = (long long)ip * g.sI + (long long)(je + 1) * g.sJ + (ke + 1);
    load_node<T, NM>(g, rp, wp, sp, minv, ip, je + 1, ke + 1, in.ro);
    if (has_h) load_node<T, NM>(g, rp, wp, sp, minv, ip, J0 - 1 + ha, K0 - 1 + hb, in.rh);
#pragma unroll
    for (int q = 0; q < 3; ++q) in.pv[q] = in.xv[q] = T(0);
    in.fm = 0;
    if (fin) {
      if constexpr (XP != 1) {
#pragma unroll
        for (int q = 0; q < 3; ++q) {
          in.pv[q] = ld_s<T>(p[3 * n + q]);
          in.xv[q] = ld_s<T>(x[3 * n + q]);
        }
      }
      if constexpr (XP == 2) {
        // r_{k-2}: this launch overwrites it with r_k only after this load
#pragma unroll
        for (int q = 0; q < 3; ++q) in.rq[q] = rn[3 * n + q];
      }
      if (fixmask != nullptr) in.fm = fixmask[n];
    }
  };
  // Updates plane ip's own and halo nodes and publishes their z in zs[buf].
  // When the plane is owned, the owner finishes its node's iteration here:
  // p_{k-1} = z_{k-1} + beta p_{k-2}, x_k = x_{k-1} + alpha p_{k-1}, r_k and
  // s_{k-1} are stored, and (r,z), (r,r) accumulated; only z (for (z,Az)) and
  // the support bits are kept for the end of the step that owns the plane.
  auto finish_stage = [&](int ip, bool own_plane, const StageIn<T>& in, int buf, T (&zo)[3]) {
    const bool fin = owner && own_plane;
    const long long n = (long long)ip * g.sI + (long long)(je + 1) * g.sJ + (ke + 1);
    NodeCG<T> o;
    finish_node(in.ro, alpha, beta, o);
#pragma unroll
    for (int q = 0; q < 3; ++q) zo[q] = o.z[q];
    if (act) {
#pragma unroll
      for (int q = 0; q < 3; ++q) zs[buf][q][a + 1][b + 1] = o.z[q];
    }
    if (has_h) {
      NodeCG<T> h;
      finish_node(in.rh, alpha, beta, h);
#pragma unroll
      for (int q = 0; q < 3; ++q) zs[buf][q][ha][hb] = h.z[q];
    }
    if (fin) {
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        if constexpr (XP == 2) {
          const T zq = in.ro.m[q] * in.rq[q];      // z_{k-2}
          const T pm = zq + beta2 * in.pv[q];      // p_{k-2}
          const T xm = in.xv[q] + alpha2 * pm;     // x_{k-1}
          const T pk = o.zp[q] + beta * pm;        // p_{k-1}
          const T xk = xm + alpha * pk;            // x_k
          p[3 * n + q] = st_s<S>(pk);
          x[3 * n + q] = st_s<S>(xk);
        } else if constexpr (XP == 0) {
          const T pk = o.zp[q] + beta * in.pv[q];  // p_{k-1}
          const T xk = in.xv[q] + alpha * pk;      // x_k
          p[3 * n + q] = st_s<S>(pk);
          x[3 * n + q] = st_s<S>(xk);
        }
        rn[3 * n + q] = o.r[q];
        sn[3 * n + q] = o.s[q];
        gk += (double)o.r[q] * (double)o.z[q];
        rk += (double)o.r[q] * (double)o.r[q];
      }
    }
  };

  const int iend = min(I0 + NI, g.nx + 1);
  // The first two planes' loads do not depend on the previous iteration's
  // reduction: they are issued before it (and before the finished-flag
  // check), so their latency overlaps that of the partials.
  StageIn<T> i0, i1;
  load_stage(I0 - 1, false, i0);
  load_stage(I0, true, i1);
  if (sc->done) return true;
  if (k >= 1) {
    const int q = (k - 1) & 1;
    double s3[3];
    reduce3_all<NT, NB>(parts.G[q], parts.D[q], parts.R[q], nb, s3);  // gamma, delta, rho
    const double gam = s3[0], del = s3[1], rho = s3[2];
    const double bet = (k == 1) ? 0.0 : gam / sc->cgg[k & 1];
    const double den = (k == 1) ? del : del - bet * gam / sc->cga[k & 1];
    const bool conv = rho <= sc->tol2;
    const bool cap = k - 1 >= sc->max_iter;
    const bool brk = !(den > 0.0);
    if (conv || cap || brk) {
      if (lb == 0 && tid == 0) {
        sc->iter = k - 1;
        sc->rr = rho;
        sc->status = conv ? GP_OK : (brk ? GP_BREAKDOWN : GP_NOT_CONVERGED);
        sc->done = 1;
      }
      return true;
    }
    const double alp = gam / den;
    if constexpr (XP == 2) {
      // written by launch k-1 into the slots this launch does not write
      alpha2 = (T)sc->cga[k & 1];
      beta2 = (T)sc->cgb[k & 1];
    }
    if (lb == 0 && tid == 0) {
      sc->cgg[(k - 1) & 1] = gam;
      sc->cga[(k - 1) & 1] = alp;
      sc->cgb[(k - 1) & 1] = bet;
      sc->iter = k - 1;
      sc->rr = rho;
    }
    alpha = (T)alp;
    beta = (T)bet;
  }

  Plane<T> Pold, Pnew, topp, bot, top;
  T zc[3], zn[3];            // z of the owned node on the current / next plane
  uint8_t fmc = 0, fmn = 0;  // its support bits
  {
    T zm1[3];
    finish_stage(I0 - 1, false, i0, 0, zm1);
    finish_stage(I0, true, i1, 1, zc);
    fmc = i1.fm;
    __syncthreads();
    corners_transform<T, TJ, TK>(zs[0], a, b, Pold);
    corners_transform<T, TJ, TK>(zs[1], a, b, Pnew);
    __syncthreads();  // step 0 rewrites zs[0]
    element(kc, Eat(I0 - 1), Pold, Pnew, bot, topp);
    Pold = Pnew;
  }

  // One element plane ie (step t): the owned node's plane ie+1 is staged and
  // its corners transformed into pnew, the element between pold and pnew
  // gives bot and top, and bot plus the previous step's top (topp) complete
  // node plane ie. The caller passes the planes of consecutive steps with
  // their roles swapped (pnew becomes pold, top becomes topp, zn becomes zc),
  // so nothing is copied between steps.
  auto step = [&](int ie, int t, Plane<T>& pold, Plane<T>& pnew, Plane<T>& topp_,
                  Plane<T>& top_, T(&zc_)[3], T(&zn_)[3], uint8_t& fmc_, uint8_t& fmn_) {
    const T Ee = Eat(ie);
    const int zb = t & 1;  // buffer of plane ie+1 (planes I0-1, I0 used 0, 1)
    Plane<T> bot_;
    if (ie < g.nx) {
      // zs[zb] was last read two steps earlier (or in the prologue), before a
      // barrier every thread has since passed
      StageIn<T> cur;
      load_stage(ie + 1, ie + 1 < iend, cur);
      finish_stage(ie + 1, ie + 1 < iend, cur, zb, zn_);
      fmn_ = cur.fm;
      __syncthreads();
      corners_transform<T, TJ, TK>(zs[zb], a, b, pnew);
      element(kc, Ee, pold, pnew, bot_, top_);
    } else {
#pragma unroll
      for (int q = 0; q < 3; ++q)
#pragma unroll
        for (int m = 0; m < 4; ++m) {
          bot_.P[q][m] = 0;
          top_.P[q][m] = 0;
        }
    }
    const int par = t & 1;
    if (act) faces_to_corners<T, NT>(topp_, bot_, cv[par], tid);
    __syncthreads();
    if (owner) {
      const long long n = (long long)ie * g.sI + (long long)(je + 1) * g.sJ + (ke + 1);
      T w[3];
#pragma unroll
      for (int q = 0; q < 3; ++q) w[q] = gather_node<T, TK, NT>(cv[par], q, tid);
      if (fmc_ & 1) w[0] = 0;
      if (fmc_ & 2) w[1] = 0;
      if (fmc_ & 4) w[2] = 0;
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        wn[3 * n + q] = w[q];
        dk += (double)zc_[q] * (double)w[q];
      }
    }
  };
  if constexpr (UNROLL2) {
    int ie = I0, t = 0;
    for (; ie + 1 < iend; ie += 2, t += 2) {
      step(ie, t, Pold, Pnew, topp, top, zc, zn, fmc, fmn);
      step(ie + 1, t + 1, Pnew, Pold, top, topp, zn, zc, fmn, fmc);
    }
    if (ie < iend) step(ie, t, Pold, Pnew, topp, top, zc, zn, fmc, fmn);
  } else {
    // the same steps as one loop with the plane roles rotated by copies (the
    // form that compiles best under a 128-register budget)
    for (int ie = I0, t = 0; ie < iend; ++ie, ++t) {
      const T Ee = Eat(ie);
      const int zb = t & 1;  // buffer of plane ie+1 (planes I0-1, I0 used 0, 1)
      if (ie < g.nx) {
        // zs[zb] was last read two steps earlier (or in the prologue), before a
        // barrier every thread has since passed
        StageIn<T> cur;
        load_stage(ie + 1, ie + 1 < iend, cur);
        finish_stage(ie + 1, ie + 1 < iend, cur, zb, zn);
        fmn = cur.fm;
        __syncthreads();
        corners_transform<T, TJ, TK>(zs[zb], a, b, Pnew);
        element(kc, Ee, Pold, Pnew, bot, top);
      } else {
#pragma unroll
        for (int q = 0; q < 3; ++q)
#pragma unroll
          for (int m = 0; m < 4; ++m) {
            bot.P[q][m] = 0;
            top.P[q][m] = 0;
          }
      }
      const int par = t & 1;
      if (act) faces_to_corners<T, NT>(topp, bot, cv[par], tid);
      __syncthreads();
      if (owner) {
        const long long n = (long long)ie * g.sI + (long long)(je + 1) * g.sJ + (ke + 1);
        T w[3];
#pragma unroll
        for (int q = 0; q < 3; ++q) w[q] = gather_node<T, TK, NT>(cv[par], q, tid);
        if (fmc & 1) w[0] = 0;
        if (fmc & 2) w[1] = 0;
        if (fmc & 4) w[2] = 0;
#pragma unroll
        for (int q = 0; q < 3; ++q) {
          wn[3 * n + q] = w[q];
          dk += (double)zc[q] * (double)w[q];
        }
      }
      topp = top;
      Pold = Pnew;
#pragma unroll
      for (int q = 0; q < 3; ++q) zc[q] = zn[q];
      fmc = fmn;
    }
  }
  block_sum3_to<NT, NB>(gk, dk, rk, parts.G[k & 1] + lb, parts.D[k & 1] + lb,
                        parts.R[k & 1] + lb);
  return false;
}

// The pending x and p updates after a solve whose launches alternate XP 1
// (odd) and XP 2 (even): when its last working launch, K = sc->iter, is odd,
// p_{K-1} and x_K are formed with the operations the sweep's XP 2 uses, from
// alpha_{K-1}, beta_{K-1} (stored by launch K) and r_{K-1}, which launch K-1
// wrote to r0; nodes lb*kBlock + t, + nb*kBlock, ... Nothing is pending after
// an even K. S: storage of x and p, as in the sweep.
template <typename T, typename S = T>
__device__ inline void body_cg_xp_fix(int nn, const DevScalars* __restrict__ sc,
                                      const T* __restrict__ r0, const T* __restrict__ minv,
                                      S* __restrict__ x, S* __restrict__ p, int lb, int nb) {
  const int K = sc->iter;
  if ((K & 1) == 0) return;
  const T alpha = (T)sc->cga[(K - 1) & 1], beta = (T)sc->cgb[(K - 1) & 1];
  constexpr bool NM = sizeof(T) == 4;  // float sweeps: M^-1 per node
  for (int n = lb * kBlock + threadIdx.x; n < nn; n += nb * kBlock) {
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      const long long d = 3ll * n + q;
      const T m = NM ? minv[n] : minv[d];
      const T z = m * r0[d];
      const T pk = z + beta * ld_s<T>(p[d]);
      const T xk = ld_s<T>(x[d]) + alpha * pk;
      p[d] = st_s<S>(pk);
      x[d] = st_s<S>(xk);
    }
  }
}

}  // namespace gp
