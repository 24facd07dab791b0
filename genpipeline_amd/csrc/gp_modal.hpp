// Device building blocks of the modal (Walsh-Hadamard) hex8 K.u sweep,
// shared by the K.u kernel (gp_ku_modal.hip) and the single-kernel PCG
// iteration (gp_cg_modal.hip). See gp_ku_modal.hip for the method.
#pragma once
// Multiply-adds are fused only where one source expression allows it, so the
// rounding of a body does not depend on where the compiler inlines it (the
// single-problem and batched kernels then agree bit for bit).
#pragma clang fp contract(on)
#include "gp_internal.hpp"

namespace gp {

// LDS staging of element-corner contributions, [buffer][corner group][component][thread].
// Declared by the calling kernel so that kernels running several bodies hold one copy.
template <typename T, int NT>
using CornerLds = T[2][4][3][NT];

template <typename T>
struct Plane {   // 2-D modal coefficients of one node plane of an element column
  T P[3][4];     // [component][mj*2+mk]
};

// Raw corner values of one node plane of an element column. In PCG mode
// the plane holds z and p_old, combined later as p = z + beta p_old.
template <typename T, bool PCG>
struct Raw {
  T a[2][2][3];
  T b[PCG ? 2 : 1][2][3];
};

// A corner outside the grid reads node 0 instead: it is a corner only of
// elements outside the grid, whose modulus is exactly 0, and it is never
// stored or summed (PCG mode writes and sums the owned node only). Leaving
// the loaded values unselected lets the corner loads issue back to back.
template <typename T, bool PCG>
__device__ inline void load_raw(const Grid& g, const T* __restrict__ u, const T* __restrict__ u2,
                                int ip, int je, int ke, Raw<T, PCG>& out) {
#pragma unroll
  for (int cj = 0; cj < 2; ++cj)
#pragma unroll
    for (int ck = 0; ck < 2; ++ck) {
      const int jj = je + cj, kk = ke + ck;
      const bool ok = (unsigned)ip <= (unsigned)g.nx && (unsigned)jj <= (unsigned)g.ny &&
                      (unsigned)kk <= (unsigned)g.nz;
      const long long n = ok ? (long long)ip * g.sI + (long long)jj * g.sJ + kk : 0;
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        out.a[cj][ck][q] = u[3 * n + q];
        if (PCG) out.b[cj & (PCG ? 1 : 0)][ck][q] = u2[3 * n + q];
      }
    }
}

// 2-D transform of a plane; c11 receives the (combined) corner (1,1) value.
template <typename T, bool PCG>
__device__ inline void transform(const Raw<T, PCG>& r, T beta, Plane<T>& out, T (&c11)[3]) {
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    T c[2][2];
#pragma unroll
    for (int cj = 0; cj < 2; ++cj)
#pragma unroll
      for (int ck = 0; ck < 2; ++ck)
        c[cj][ck] = PCG ? r.a[cj][ck][q] + beta * r.b[cj & (PCG ? 1 : 0)][ck][q] : r.a[cj][ck][q];
    const T a0 = c[0][0] + c[0][1], d0 = c[0][1] - c[0][0];
    const T a1 = c[1][0] + c[1][1], d1 = c[1][1] - c[1][0];
    out.P[q][0] = a0 + a1;  // (mj,mk) = (0,0)
    out.P[q][1] = d0 + d1;  // (0,1)
    out.P[q][2] = a1 - a0;  // (1,0)
    out.P[q][3] = d1 - d0;  // (1,1)
    c11[q] = c[1][1];
  }
}

// Modal inputs b = T^T u_e (b0: mode mi = 0, b1: mi = 1, [component][m2])
// and outputs v = Ee Khat b / 64 of the element between planes `lo` (i) and
// `hi` (i+1).
template <typename T>
__device__ inline void element_modes(const ModalCoef<T>& k, T Ee, const Plane<T>& lo,
                                     const Plane<T>& hi, T (&b0)[3][4], T (&b1)[3][4],
                                     T (&v0)[3][4], T (&v1)[3][4]) {
#pragma unroll
  for (int q = 0; q < 3; ++q)
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      b0[q][m] = lo.P[q][m] + hi.P[q][m];
      b1[q][m] = hi.P[q][m] - lo.P[q][m];
    }
  // named modal inputs (x = i, y = j, z = k; m2 index = mj*2 + mk)
  const T bxx = b1[0][0], byy = b0[1][2], bzz = b0[2][1];
  const T bxy = b1[1][0], byx = b0[0][2];
  const T bxz = b1[2][0], bzx = b0[0][1];
  const T byz = b0[2][2], bzy = b0[1][1];
  const T bXYx = b1[0][2], bXYy = b1[1][2], bXYz = b1[2][2];   // mode xy
  const T bYZx = b0[0][3], bYZy = b0[1][3], bYZz = b0[2][3];   // mode yz
  const T bXZx = b1[0][1], bXZy = b1[1][1], bXZz = b1[2][1];   // mode xz
  const T bXYZx = b1[0][3], bXYZy = b1[1][3], bXYZz = b1[2][3];
  const T An = Ee * k.An, Ln = Ee * k.Ln, S = Ee * k.S, B1 = Ee * k.B1, B2 = Ee * k.B2;
  const T T1 = Ee * k.T1, T2 = Ee * k.T2, X = Ee * k.X;
  const T tr = bxx + byy + bzz;
  const T vxx = An * bxx + Ln * tr, vyy = An * byy + Ln * tr, vzz = An * bzz + Ln * tr;
  const T sxy = S * (bxy + byx), sxz = S * (bxz + bzx), syz = S * (byz + bzy);
  const T vXYx = B1 * bXYx + B2 * bYZz, vYZz = B2 * bXYx + B1 * bYZz;
  const T vXYy = B1 * bXYy + B2 * bXZz, vXZz = B2 * bXYy + B1 * bXZz;
  const T vYZy = B1 * bYZy + B2 * bXZx, vXZx = B2 * bYZy + B1 * bXZx;
  const T s3 = bXYz + bYZx + bXZy;
  const T vXYz = T1 * bXYz + T2 * s3, vYZx = T1 * bYZx + T2 * s3, vXZy = T1 * bXZy + T2 * s3;
  // v[mi][q][m2]; v[0][q][0] is the rigid translation (zero output)
  v0[0][0] = 0; v0[1][0] = 0; v0[2][0] = 0;
  v1[0][0] = vxx; v1[1][0] = sxy; v1[2][0] = sxz;      // mode x
  v0[0][2] = sxy; v0[1][2] = vyy; v0[2][2] = syz;      // mode y
  v0[0][1] = sxz; v0[1][1] = syz; v0[2][1] = vzz;      // mode z
  v1[0][2] = vXYx; v1[1][2] = vXYy; v1[2][2] = vXYz;   // mode xy
  v0[0][3] = vYZx; v0[1][3] = vYZy; v0[2][3] = vYZz;   // mode yz
  v1[0][1] = vXZx; v1[1][1] = vXZy; v1[2][1] = vXZz;   // mode xz
  v1[0][3] = X * bXYZx; v1[1][3] = X * bXYZy; v1[2][3] = X * bXYZz;
}

// Element between planes `lo` (i) and `hi` (i+1): returns the 2-D modal
// outputs of its bottom (plane i) and top (plane i+1) faces.
template <typename T>
__device__ inline void element(const ModalCoef<T>& k, T Ee, const Plane<T>& lo, const Plane<T>& hi,
                               Plane<T>& bot, Plane<T>& top) {
  T b0[3][4], b1[3][4], v0[3][4], v1[3][4];
  element_modes(k, Ee, lo, hi, b0, b1, v0, v1);
#pragma unroll
  for (int q = 0; q < 3; ++q)
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      top.P[q][m] = v0[q][m] + v1[q][m];
      bot.P[q][m] = v0[q][m] - v1[q][m];
    }
}

// Fixed-order block sums of three partial arrays, returned to every thread
// (NB threads per block, the first NT of them loading; any NB, with the DPP
// wave sum when NB is a multiple of the wave size).
template <int NT, int NB = NT>
__device__ inline void reduce3_all(const double* a, const double* b, const double* c, int n,
                                   double (&out)[3]) {
  constexpr int NW = (NT + 63) / 64;
  __shared__ double sh[3][NW];
  // thread t sums entries t, t+NT, t+2NT, ... in that order; each chunk of
  // four is loaded before any is added, so the loads overlap
  constexpr int U = 4;
  double s0 = 0.0, s1 = 0.0, s2 = 0.0;
  // threads past NT (a block padded to more lanes than its tile) add nothing
  for (int i0 = threadIdx.x; threadIdx.x < NT && i0 < n; i0 += U * NT) {
    double va[U], vb[U], vc[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int i = i0 + u * NT < n ? i0 + u * NT : i0;
      va[u] = a[i];
      vb[u] = b[i];
      vc[u] = c[i];
    }
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (i0 + u * NT < n) {
        s0 += va[u];
        s1 += vb[u];
        s2 += vc[u];
      }
  }
  s0 = wave_sum_nb<NB>(s0);
  s1 = wave_sum_nb<NB>(s1);
  s2 = wave_sum_nb<NB>(s2);
  if ((threadIdx.x & 63) == 0) {
    sh[0][threadIdx.x >> 6] = s0;
    sh[1][threadIdx.x >> 6] = s1;
    sh[2][threadIdx.x >> 6] = s2;
  }
  __syncthreads();
  out[0] = out[1] = out[2] = 0.0;
  for (int w = 0; w < NW; ++w) {
    out[0] += sh[0][w];
    out[1] += sh[1][w];
    out[2] += sh[2][w];
  }
  __syncthreads();
}


// 2-D inverse of a node plane's summed face coefficients into the four
// corner values of the element column, written to cv[corner][q][slot].
template <typename T, int NT>
__device__ inline void faces_to_corners(const Plane<T>& topp, const Plane<T>& bot,
                                        T (&cv)[4][3][NT], int slot) {
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    const T Q00 = topp.P[q][0] + bot.P[q][0], Q01 = topp.P[q][1] + bot.P[q][1];
    const T Q10 = topp.P[q][2] + bot.P[q][2], Q11 = topp.P[q][3] + bot.P[q][3];
    const T yj0k0 = Q00 - Q10, yj0k1 = Q01 - Q11;  // cj = 0 (mk = 0, 1)
    const T yj1k0 = Q00 + Q10, yj1k1 = Q01 + Q11;  // cj = 1
    cv[0][q][slot] = yj0k0 - yj0k1;                // corner (0,0)
    cv[1][q][slot] = yj0k0 + yj0k1;                // corner (0,1)
    cv[2][q][slot] = yj1k0 - yj1k1;                // corner (1,0)
    cv[3][q][slot] = yj1k0 + yj1k1;                // corner (1,1)
  }
}

// Node (J,K) owned by thread `tid` (element column (J-1,K-1)): sum of the four
// element columns' corner values, in a fixed order.
template <typename T, int TK, int NT>
__device__ inline T gather_node(const T (&cv)[4][3][NT], int q, int tid) {
  return cv[3][q][tid] + cv[2][q][tid + 1] + cv[1][q][tid + TK] + cv[0][q][tid + TK + 1];
}

// MODE 0: y = K(E) u.
// MODE 1: PCG iteration `it` front half. Convergence test on the previous
//   iteration's partials (||r||^2 <= tol^2 ||b||^2, or the iteration cap),
//   beta = rz'/rz, p = z + beta p_old formed while loading (u = z,
//   u2 = p_old), y = K p, p written to pout by the owner of each node, and
//   the block's p.Ap partial.
template <typename T, int TJ, int TK, int MODE>
__device__ inline void body_ku_modal(const Grid& g, int NI, int ntj, int ntk,
                                     const ModalCoef<T>& kc, const float* __restrict__ E,
                                     const T* __restrict__ u, const T* __restrict__ u2,
                                     T* __restrict__ y, T* __restrict__ pout,
                                     const uint8_t* __restrict__ fixmask, DevScalars* sc,
                                     const Parts& parts, int it, int nV, int lb,
                                     CornerLds<T, TJ * TK>& cv) {
  constexpr int NT = TJ * TK;
  constexpr bool PCG = MODE == 1;
  if (PCG && sc->done) return;
  const int tid = threadIdx.x;
  const int a = tid / TK, b = tid - (tid / TK) * TK;
  int bid = lb;
  const int tk_ = bid % ntk;
  bid /= ntk;
  const int tj_ = bid % ntj;
  const int ti_ = bid / ntj;
  const int J0 = tj_ * (TJ - 1), K0 = tk_ * (TK - 1), I0 = ti_ * NI;
  const int je = J0 - 1 + a, ke = K0 - 1 + b;
  const bool col_ok = (unsigned)je < (unsigned)g.ny && (unsigned)ke < (unsigned)g.nz;
  const bool owner = a < TJ - 1 && b < TK - 1 && (je + 1) <= g.ny && (ke + 1) <= g.nz;

  auto Eat = [&](int ie) -> T {
    const bool ok = col_ok && (unsigned)ie < (unsigned)g.nx;
    const int e = ok ? ie * g.eI + je * g.eJ + ke : 0;
    const float v = E[e];
    return ok ? (T)v : T(0);
  };

  Raw<T, PCG> ra, rb, rn;
  load_raw<T, PCG>(g, u, u2, I0 - 1, je, ke, ra);
  load_raw<T, PCG>(g, u, u2, I0, je, ke, rb);
  load_raw<T, PCG>(g, u, u2, I0 + 1, je, ke, rn);
  T beta = T(0);
  if (PCG && it > 0) {
    const int q = it & 1;
    double s3[3];
    reduce3_all<NT>(parts.RZ[q], parts.RR, parts.RZ[q ^ 1], nV, s3);  // rz', rr, rz
    const bool conv = s3[1] <= sc->tol2;
    const bool cap = it >= sc->max_iter;
    if (conv || cap) {
      if (lb == 0 && tid == 0) {
        sc->iter = it;
        sc->rr = s3[1];
        sc->rz = s3[0];
        sc->status = conv ? GP_OK : GP_NOT_CONVERGED;
        sc->done = 1;
      }
      return;
    }
    if (lb == 0 && tid == 0) {
      sc->iter = it;
      sc->rr = s3[1];
    }
    beta = (T)(s3[0] / s3[2]);
  }

  Plane<T> Pold, Pnew, topp, bot, top;
  T pc_old[3], pc_new[3];
  transform<T, PCG>(ra, beta, Pold, pc_old);
  transform<T, PCG>(rb, beta, Pnew, pc_new);
  element(kc, Eat(I0 - 1), Pold, Pnew, bot, topp);
  Pold = Pnew;
  pc_old[0] = pc_new[0];
  pc_old[1] = pc_new[1];
  pc_old[2] = pc_new[2];

  double dot = 0.0;
  const int iend = min(I0 + NI, g.nx + 1);
  for (int ie = I0, t = 0; ie < iend; ++ie, ++t) {
    const T Ee = Eat(ie);
    if (ie < g.nx) {
      transform<T, PCG>(rn, beta, Pnew, pc_new);
      if (ie + 1 < iend) load_raw<T, PCG>(g, u, u2, ie + 2, je, ke, rn);  // prefetch
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
    // node plane ie: this column's faces, then 2-D inverse to 4 corners
    const int par = t & 1;
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      const T Q00 = topp.P[q][0] + bot.P[q][0], Q01 = topp.P[q][1] + bot.P[q][1];
      const T Q10 = topp.P[q][2] + bot.P[q][2], Q11 = topp.P[q][3] + bot.P[q][3];
      const T yj0k0 = Q00 - Q10, yj0k1 = Q01 - Q11;  // cj = 0 (mk = 0, 1)
      const T yj1k0 = Q00 + Q10, yj1k1 = Q01 + Q11;  // cj = 1
      cv[par][0][q][tid] = yj0k0 - yj0k1;             // corner (0,0)
      cv[par][1][q][tid] = yj0k0 + yj0k1;             // corner (0,1)
      cv[par][2][q][tid] = yj1k0 - yj1k1;             // corner (1,0)
      cv[par][3][q][tid] = yj1k0 + yj1k1;             // corner (1,1)
    }
    __syncthreads();
    if (owner) {
      const long long n = (long long)ie * g.sI + (long long)(je + 1) * g.sJ + (ke + 1);
      T o[3];
#pragma unroll
      for (int q = 0; q < 3; ++q)
        o[q] = cv[par][3][q][tid] + cv[par][2][q][tid + 1] + cv[par][1][q][tid + TK] +
               cv[par][0][q][tid + TK + 1];
      if (fixmask != nullptr) {
        const uint8_t fm = fixmask[n];
        if (fm & 1) o[0] = 0;
        if (fm & 2) o[1] = 0;
        if (fm & 4) o[2] = 0;
      }
      y[3 * n] = o[0];
      y[3 * n + 1] = o[1];
      y[3 * n + 2] = o[2];
      if (PCG) {
        pout[3 * n] = pc_old[0];
        pout[3 * n + 1] = pc_old[1];
        pout[3 * n + 2] = pc_old[2];
        dot += (double)pc_old[0] * (double)o[0] + (double)pc_old[1] * (double)o[1] +
               (double)pc_old[2] * (double)o[2];
      }
    }
    topp = top;
    Pold = Pnew;
    pc_old[0] = pc_new[0];
    pc_old[1] = pc_new[1];
    pc_old[2] = pc_new[2];
  }
  if (PCG) {
    // block sum of p.Ap in a fixed order (any block size)
    __shared__ double red[(NT + 63) / 64];
    dot = wave_sum64(dot);
    if ((tid & 63) == 0) red[tid >> 6] = dot;
    __syncthreads();
    if (tid == 0) {
      double s = 0.0;
      for (int w = 0; w < (NT + 63) / 64; ++w) s += red[w];
      parts.A[lb] = s;
    }
  }
}

// y = K(E) u (fixed DOFs zeroed when fixmask is given) with one load per node:
// thread (a,b) loads its own node column (J0+a, K0+b) and the TJ+TK+1 halo
// columns with aa = 0 or bb = 0 go to the first threads; element columns read
// their four corners from an LDS tile of the plane. The arithmetic, and each
// node's summation order, are those of body_ku_modal in MODE 0, so the result
// is the same bit for bit; each node is read once per plane instead of once
// per element column that touches it.
// NB: threads per block, NT = TJ x TK of them active (the rest only join the
// barriers).
template <typename T, int TJ, int TK, int NB = TJ * TK>
__device__ inline void body_ku_sweep(const Grid& g, int NI, int ntj, int ntk,
                                     const ModalCoef<T>& kc, const float* __restrict__ E,
                                     const T* __restrict__ u, T* __restrict__ y,
                                     const uint8_t* __restrict__ fixmask, int lb,
                                     CornerLds<T, TJ * TK>& cv) {
  constexpr int NT = TJ * TK;
  constexpr int NH = TJ + TK + 1;
  static_assert(NB >= NT && (NB + 63) / 64 == (NT + 63) / 64, "padding within the last wave");
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
  const bool has_h = tid < NH;
  const int ha = tid <= TK ? 0 : tid - TK;
  const int hb = tid <= TK ? tid : 0;

  auto Eat = [&](int ie) -> T {
    const bool ok = col_ok && (unsigned)ie < (unsigned)g.nx;
    const int e = ok ? ie * g.eI + je * g.eJ + ke : 0;
    const float v = E[e];
    return ok ? (T)v : T(0);
  };
  // a node outside the grid reads node 0: it is a corner only of elements
  // outside the grid, whose modulus is exactly 0
  auto node = [&](int ip, int jj, int kk, T (&v)[3]) {
    const bool ok = (unsigned)ip <= (unsigned)g.nx && (unsigned)jj <= (unsigned)g.ny &&
                    (unsigned)kk <= (unsigned)g.nz;
    const long long n = ok ? (long long)ip * g.sI + (long long)jj * g.sJ + kk : 0;
#pragma unroll
    for (int q = 0; q < 3; ++q) v[q] = u[3 * n + q];
  };
  auto stage = [&](int ip, int buf) {
    T vo[3], vh[3];
    node(ip, je + 1, ke + 1, vo);
    if (has_h) node(ip, J0 - 1 + ha, K0 - 1 + hb, vh);
    if (act) {
#pragma unroll
      for (int q = 0; q < 3; ++q) zs[buf][q][a + 1][b + 1] = vo[q];
    }
    if (has_h) {
#pragma unroll
      for (int q = 0; q < 3; ++q) zs[buf][q][ha][hb] = vh[q];
    }
  };
  // 2-D transform of the element column's four corners from the LDS tile
  auto corners = [&](int buf, Plane<T>& out) {
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      const T c00 = zs[buf][q][a][b], c01 = zs[buf][q][a][b + 1];
      const T c10 = zs[buf][q][a + 1][b], c11 = zs[buf][q][a + 1][b + 1];
      const T a0 = c00 + c01, d0 = c01 - c00;
      const T a1 = c10 + c11, d1 = c11 - c10;
      out.P[q][0] = a0 + a1;
      out.P[q][1] = d0 + d1;
      out.P[q][2] = a1 - a0;
      out.P[q][3] = d1 - d0;
    }
  };

  Plane<T> Pold, Pnew, topp, bot, top;
  stage(I0 - 1, 0);
  stage(I0, 1);
  __syncthreads();
  corners(0, Pold);
  corners(1, Pnew);
  __syncthreads();  // step 0 rewrites zs[0]
  element(kc, Eat(I0 - 1), Pold, Pnew, bot, topp);
  Pold = Pnew;
  const int iend = min(I0 + NI, g.nx + 1);
  for (int ie = I0, t = 0; ie < iend; ++ie, ++t) {
    const T Ee = Eat(ie);
    const int zb = t & 1;
    if (ie < g.nx) {
      // zs[zb] was last read two steps earlier (or in the prologue), before a
      // barrier every thread has since passed
      stage(ie + 1, zb);
      __syncthreads();
      corners(zb, Pnew);
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
      T o[3];
#pragma unroll
      for (int q = 0; q < 3; ++q) o[q] = gather_node<T, TK, NT>(cv[par], q, tid);
      if (fixmask != nullptr) {
        const uint8_t fm = fixmask[n];
        if (fm & 1) o[0] = 0;
        if (fm & 2) o[1] = 0;
        if (fm & 4) o[2] = 0;
      }
#pragma unroll
      for (int q = 0; q < 3; ++q) y[3 * n + q] = o[q];
    }
    topp = top;
    Pold = Pnew;
  }
}

}  // namespace gp
