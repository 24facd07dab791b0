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
#include "gp_internal.hpp"

namespace gp {

template <typename T>
struct Plane {   // 2-D modal coefficients of one node plane of an element column
  T P[3][4];     // [component][mj*2+mk]
};

template <typename T>
__device__ inline void load_plane(const Grid& g, const T* __restrict__ u, int ip, int je, int ke,
                                  Plane<T>& out, T (&c11)[3]) {
  T c[2][2][3];
#pragma unroll
  for (int cj = 0; cj < 2; ++cj)
#pragma unroll
    for (int ck = 0; ck < 2; ++ck) {
      const int jj = je + cj, kk = ke + ck;
      const bool ok = (unsigned)ip <= (unsigned)g.nx && (unsigned)jj <= (unsigned)g.ny &&
                      (unsigned)kk <= (unsigned)g.nz;
      const long long n = ok ? (long long)ip * g.sI + (long long)jj * g.sJ + kk : 0;
      const T v0 = u[3 * n], v1 = u[3 * n + 1], v2 = u[3 * n + 2];
      c[cj][ck][0] = ok ? v0 : T(0);
      c[cj][ck][1] = ok ? v1 : T(0);
      c[cj][ck][2] = ok ? v2 : T(0);
    }
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    const T a0 = c[0][0][q] + c[0][1][q], d0 = c[0][1][q] - c[0][0][q];
    const T a1 = c[1][0][q] + c[1][1][q], d1 = c[1][1][q] - c[1][0][q];
    out.P[q][0] = a0 + a1;  // (mj,mk) = (0,0)
    out.P[q][1] = d0 + d1;  // (0,1)
    out.P[q][2] = a1 - a0;  // (1,0)
    out.P[q][3] = d1 - d0;  // (1,1)
    c11[q] = c[1][1][q];
  }
}

// Element between planes `lo` (i) and `hi` (i+1): returns the 2-D modal
// outputs of its bottom (plane i) and top (plane i+1) faces.
template <typename T>
__device__ inline void element(const ModalCoef<T>& k, T Ee, const Plane<T>& lo, const Plane<T>& hi,
                               Plane<T>& bot, Plane<T>& top) {
  // b[mi][q][m2]
  T b0[3][4], b1[3][4];
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
  T v0[3][4], v1[3][4];
  v0[0][0] = 0; v0[1][0] = 0; v0[2][0] = 0;
  v1[0][0] = vxx; v1[1][0] = sxy; v1[2][0] = sxz;      // mode x
  v0[0][2] = sxy; v0[1][2] = vyy; v0[2][2] = syz;      // mode y
  v0[0][1] = sxz; v0[1][1] = syz; v0[2][1] = vzz;      // mode z
  v1[0][2] = vXYx; v1[1][2] = vXYy; v1[2][2] = vXYz;   // mode xy
  v0[0][3] = vYZx; v0[1][3] = vYZy; v0[2][3] = vYZz;   // mode yz
  v1[0][1] = vXZx; v1[1][1] = vXZy; v1[2][1] = vXZz;   // mode xz
  v1[0][3] = X * bXYZx; v1[1][3] = X * bXYZy; v1[2][3] = X * bXYZz;
#pragma unroll
  for (int q = 0; q < 3; ++q)
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      top.P[q][m] = v0[q][m] + v1[q][m];
      bot.P[q][m] = v0[q][m] - v1[q][m];
    }
}

template <typename T, int TJ, int TK, bool FUSE>
__global__ __launch_bounds__(TJ* TK) void k_ku_modal(Grid g, int NI, int ntj, int ntk,
                                                    ModalCoef<T> kc,
                                                    const float* __restrict__ E,
                                                    const T* __restrict__ u,
                                                    T* __restrict__ y,
                                                    const uint8_t* __restrict__ fixmask,
                                                    const DevScalars* sc, double* partA) {
  constexpr int NT = TJ * TK;
  if (FUSE && sc->done) return;
  __shared__ T cv[2][4][3][NT];
  const int tid = threadIdx.x;
  const int a = tid / TK, b = tid - (tid / TK) * TK;
  int bid = blockIdx.x;
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

  Plane<T> Pold, Pnew, topp, bot, top;
  T pc_old[3], pc_new[3];
  load_plane(g, u, I0 - 1, je, ke, Pold, pc_old);
  load_plane(g, u, I0, je, ke, Pnew, pc_new);
  element(kc, Eat(I0 - 1), Pold, Pnew, bot, topp);
  Pold = Pnew;
  pc_old[0] = pc_new[0];
  pc_old[1] = pc_new[1];
  pc_old[2] = pc_new[2];

  double dot = 0.0;
  const int iend = min(I0 + NI, g.nx + 1);
  for (int ie = I0, t = 0; ie < iend; ++ie, ++t) {
    if (ie < g.nx) {
      load_plane(g, u, ie + 1, je, ke, Pnew, pc_new);
      element(kc, Eat(ie), Pold, Pnew, bot, top);
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
      if (FUSE)
        dot += (double)pc_old[0] * (double)o[0] + (double)pc_old[1] * (double)o[1] +
               (double)pc_old[2] * (double)o[2];
    }
    topp = top;
    Pold = Pnew;
    pc_old[0] = pc_new[0];
    pc_old[1] = pc_new[1];
    pc_old[2] = pc_new[2];
  }
  if (FUSE) {
    // block sum of p.Ap in a fixed order (any block size)
    __shared__ double red[(NT + 63) / 64];
    dot = wave_sum64(dot);
    if ((tid & 63) == 0) red[tid >> 6] = dot;
    __syncthreads();
    if (tid == 0) {
      double s = 0.0;
      for (int w = 0; w < (NT + 63) / 64; ++w) s += red[w];
      partA[blockIdx.x] = s;
    }
  }
}

template <int TJ, int TK>
static void launch_modal_t(const Grid& g, int NI, const void* kc, int prec, const float* E,
                           const void* u, void* y, const uint8_t* fixmask, const DevScalars* sc,
                           double* partA, bool fuse, hipStream_t st) {
  const int ntj = (g.ny + 1 + (TJ - 2)) / (TJ - 1);
  const int ntk = (g.nz + 1 + (TK - 2)) / (TK - 1);
  const int nti = (g.nx + 1 + NI - 1) / NI;
  const dim3 grid(ntj * ntk * nti), block(TJ * TK);
  if (prec == GP_F32) {
    const ModalCoef<float>& c = *reinterpret_cast<const ModalCoef<float>*>(kc);
    if (fuse)
      hipLaunchKernelGGL((k_ku_modal<float, TJ, TK, true>), grid, block, 0, st, g, NI, ntj, ntk, c,
                         E, (const float*)u, (float*)y, fixmask, sc, partA);
    else
      hipLaunchKernelGGL((k_ku_modal<float, TJ, TK, false>), grid, block, 0, st, g, NI, ntj, ntk,
                         c, E, (const float*)u, (float*)y, fixmask, sc, partA);
  } else {
    const ModalCoef<double>& c = *reinterpret_cast<const ModalCoef<double>*>(kc);
    if (fuse)
      hipLaunchKernelGGL((k_ku_modal<double, TJ, TK, true>), grid, block, 0, st, g, NI, ntj, ntk,
                         c, E, (const double*)u, (double*)y, fixmask, sc, partA);
    else
      hipLaunchKernelGGL((k_ku_modal<double, TJ, TK, false>), grid, block, 0, st, g, NI, ntj, ntk,
                         c, E, (const double*)u, (double*)y, fixmask, sc, partA);
  }
}

int modal_blocks(const Grid& g, const KuTile& t) {
  const int ntj = (g.ny + 1 + (t.tj - 2)) / (t.tj - 1);
  const int ntk = (g.nz + 1 + (t.tk - 2)) / (t.tk - 1);
  const int nti = (g.nx + 1 + t.ni - 1) / t.ni;
  return ntj * ntk * nti;
}

bool launch_ku_modal(const Grid& g, const KuTile& t, const void* kc, int prec, const float* E,
                     const void* u, void* y, const uint8_t* fixmask, const DevScalars* sc,
                     double* partA, bool fuse, hipStream_t st) {
#define GP_TILE(TJ_, TK_)                                                                 \
  if (t.tj == TJ_ && t.tk == TK_) {                                                       \
    launch_modal_t<TJ_, TK_>(g, t.ni, kc, prec, E, u, y, fixmask, sc, partA, fuse, st);   \
    return true;                                                                          \
  }
  GP_TILE(8, 32)
  GP_TILE(16, 16)
  GP_TILE(14, 18)
  GP_TILE(6, 34)
  GP_TILE(14, 34)
  GP_TILE(4, 64)
  GP_TILE(10, 26)
#undef GP_TILE
  return false;
}

}  // namespace gp
