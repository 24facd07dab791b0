// HIP kernels of the SIMP inner loop for gfx950 (CDNA4, wave64).
//
// Grid numbering follows genpipeline/topology/simp_solver.py:251-282 exactly;
// nothing is gathered through an edof table on the structured path -- every
// index is derived arithmetically from (i,j,k).
//
// Reductions are deterministic: each block reduces in a fixed butterfly order,
// writes one FP64 partial, and the last block to finish (device-scope counter)
// sums the partials in index order. No FP64 atomics touch a result.
#include "gp_internal.hpp"

namespace gp {

// C3D8 corner offsets (simp_solver.py:255-267) and the inverse map
// (cx,cy,cz) -> local corner index, keyed by cx*4 + cy*2 + cz.
__device__ constexpr int CX[8] = {0, 1, 1, 0, 0, 1, 1, 0};
__device__ constexpr int CY[8] = {0, 0, 1, 1, 0, 0, 1, 1};
__device__ constexpr int CZ[8] = {0, 0, 0, 0, 1, 1, 1, 1};
__device__ constexpr int LOC[8] = {0, 4, 3, 7, 1, 5, 2, 6};

// ------------------------------------------------------------ reductions
__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}
__device__ inline double wave_max(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
  return v;
}

// Sums NV values over the block; the result is valid in thread 0.
template <int NV>
__device__ inline void block_sum(double (&v)[NV]) {
  __shared__ double sh[NV][kBlock / 64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < NV; ++q) {
    double s = wave_sum(v[q]);
    if (lane == 0) sh[q][w] = s;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int q = 0; q < NV; ++q) {
      double s = 0.0;
      for (int i = 0; i < kBlock / 64; ++i) s += sh[q][i];
      v[q] = s;
    }
  }
}

// Thread 0 of every block has written its partials; returns true in every
// thread of the last block to arrive. Release/acquire through device-scope
// fences so the last block sees all partials.
__device__ inline bool last_block(unsigned int* counter) {
  __shared__ int s_last;
  if (threadIdx.x == 0) {
    __threadfence();
    unsigned int prev = atomicAdd(counter, 1u);
    s_last = (prev == gridDim.x - 1);
  }
  __syncthreads();
  const bool last = s_last != 0;
  if (last) __threadfence();
  return last;
}

// Sums `n` partials (NV interleaved streams, stride kMaxPartials) in index
// order; result valid in thread 0.
template <int NV>
__device__ inline void sum_partials(const double* partials, int n, double (&out)[NV]) {
  // Each thread sums a fixed strided subset, then the block tree combines:
  // the order depends only on n, never on timing.
#pragma unroll
  for (int q = 0; q < NV; ++q) {
    double s = 0.0;
    for (int i = threadIdx.x; i < n; i += kBlock) s += partials[q * kMaxPartials + i];
    out[q] = s;
  }
  block_sum<NV>(out);
}

// --------------------------------------------------------------- helpers
template <typename T> struct V3 { T x, y, z; };

template <typename T>
__device__ inline V3<T> ld3(const T* __restrict__ a, long long n) {
  V3<T> r;
  r.x = a[3 * n];
  r.y = a[3 * n + 1];
  r.z = a[3 * n + 2];
  return r;
}
template <typename T>
__device__ inline void st3(T* __restrict__ a, long long n, T x, T y, T z) {
  a[3 * n] = x;
  a[3 * n + 1] = y;
  a[3 * n + 2] = z;
}

// --------------------------------------------------------------- K . u
// One thread per node. The node's 27-node neighbourhood is held in
// registers; each of its (up to) 8 elements contributes E_e * KE[3L+r, :] u_e,
// with L the node's corner index in that element. KE is read with uniform
// (scalar-cache) loads because every index is a compile-time constant.
// FUSE: the PCG form -- returns early once the solve is done and reduces
// p.Ap into alpha = rz / pAp through the last block.
template <typename T, bool FUSE>
__global__ __launch_bounds__(kBlock) void k_ku(Grid g, const T* __restrict__ ke,
                                              const float* __restrict__ E,
                                              const T* __restrict__ u, T* __restrict__ y,
                                              const uint8_t* __restrict__ fixmask,
                                              DevScalars* sc, double* partials) {
  if (FUSE && sc->done) return;
  const int n = blockIdx.x * kBlock + threadIdx.x;
  const bool valid = n < g.nn;
  const int nc = valid ? n : 0;
  const int i = nc / g.sI;
  const int rem = nc - i * g.sI;
  const int j = rem / g.sJ;
  const int k = rem - j * g.sJ;

  T un[27][3];
#pragma unroll
  for (int di = 0; di < 3; ++di)
#pragma unroll
    for (int dj = 0; dj < 3; ++dj)
#pragma unroll
      for (int dk = 0; dk < 3; ++dk) {
        const int ii = i + di - 1, jj = j + dj - 1, kk = k + dk - 1;
        const bool ok = (unsigned)ii <= (unsigned)g.nx && (unsigned)jj <= (unsigned)g.ny &&
                        (unsigned)kk <= (unsigned)g.nz;
        const int m = ok ? nc + (di - 1) * g.sI + (dj - 1) * g.sJ + (dk - 1) : nc;
        const int q = di * 9 + dj * 3 + dk;
        const V3<T> v = ld3(u, m);
        un[q][0] = ok ? v.x : T(0);
        un[q][1] = ok ? v.y : T(0);
        un[q][2] = ok ? v.z : T(0);
      }

  T acc0 = 0, acc1 = 0, acc2 = 0;
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int c = 0; c < 2; ++c) {
        const int ei = i - 1 + a, ej = j - 1 + b, ek = k - 1 + c;
        const bool ok = (unsigned)ei < (unsigned)g.nx && (unsigned)ej < (unsigned)g.ny &&
                        (unsigned)ek < (unsigned)g.nz;
        const int e = ok ? ei * g.eI + ej * g.eJ + ek : 0;
        const T Ee = ok ? (T)E[e] : T(0);
        const int L = LOC[(1 - a) * 4 + (1 - b) * 2 + (1 - c)];
        T s0 = 0, s1 = 0, s2 = 0;
#pragma unroll
        for (int m = 0; m < 8; ++m) {
          const int q = (a + CX[m]) * 9 + (b + CY[m]) * 3 + (c + CZ[m]);
          const T* r0 = ke + (3 * L + 0) * 24 + 3 * m;
          const T* r1 = ke + (3 * L + 1) * 24 + 3 * m;
          const T* r2 = ke + (3 * L + 2) * 24 + 3 * m;
          s0 += r0[0] * un[q][0] + r0[1] * un[q][1] + r0[2] * un[q][2];
          s1 += r1[0] * un[q][0] + r1[1] * un[q][1] + r1[2] * un[q][2];
          s2 += r2[0] * un[q][0] + r2[1] * un[q][1] + r2[2] * un[q][2];
        }
        acc0 += Ee * s0;
        acc1 += Ee * s1;
        acc2 += Ee * s2;
      }
  if (fixmask != nullptr && valid) {
    const uint8_t fm = fixmask[n];
    if (fm & 1) acc0 = 0;
    if (fm & 2) acc1 = 0;
    if (fm & 4) acc2 = 0;
  }
  if (valid) st3(y, n, acc0, acc1, acc2);

  if (FUSE) {
    double v[1] = {valid ? (double)un[13][0] * (double)acc0 + (double)un[13][1] * (double)acc1 +
                               (double)un[13][2] * (double)acc2
                         : 0.0};
    block_sum<1>(v);
    if (threadIdx.x == 0) partials[blockIdx.x] = v[0];
    if (last_block(&sc->cnt[CNT_A])) {
      double s[1];
      sum_partials<1>(partials, gridDim.x, s);
      if (threadIdx.x == 0) {
        const double pAp = s[0];
        sc->pAp = pAp;
        if (!(pAp > 0.0)) {
          sc->status = GP_BREAKDOWN;
          sc->done = 1;
        } else {
          sc->alpha = sc->rz / pAp;
        }
        sc->cnt[CNT_A] = 0;
      }
    }
  }
}

template <typename T>
void launch_ku(const Grid& g, const T* ke, const float* E, const T* u, T* y,
               const uint8_t* fixmask, DevScalars* sc, double* partials, bool fuse_pap,
               hipStream_t st) {
  const int nb = grid_blocks(g.nn);
  if (fuse_pap)
    hipLaunchKernelGGL((k_ku<T, true>), dim3(nb), dim3(kBlock), 0, st, g, ke, E, u, y, fixmask,
                       sc, partials);
  else
    hipLaunchKernelGGL((k_ku<T, false>), dim3(nb), dim3(kBlock), 0, st, g, ke, E, u, y, fixmask,
                       sc, partials);
}

// diag(K) per DOF; with invert: Jacobi M^-1 (0 on fixed DOFs and on DOFs
// with a non-positive diagonal).
template <typename T>
__global__ __launch_bounds__(kBlock) void k_diag(Grid g, const T* __restrict__ ke,
                                                const float* __restrict__ E,
                                                const uint8_t* __restrict__ fixmask,
                                                T* __restrict__ out, int invert) {
  const int n = blockIdx.x * kBlock + threadIdx.x;
  if (n >= g.nn) return;
  const int i = n / g.sI, rem = n - i * g.sI, j = rem / g.sJ, k = rem - j * g.sJ;
  T d0 = 0, d1 = 0, d2 = 0;
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int c = 0; c < 2; ++c) {
        const int ei = i - 1 + a, ej = j - 1 + b, ek = k - 1 + c;
        const bool ok = (unsigned)ei < (unsigned)g.nx && (unsigned)ej < (unsigned)g.ny &&
                        (unsigned)ek < (unsigned)g.nz;
        const int e = ok ? ei * g.eI + ej * g.eJ + ek : 0;
        const T Ee = ok ? (T)E[e] : T(0);
        const int L = LOC[(1 - a) * 4 + (1 - b) * 2 + (1 - c)];
        d0 += Ee * ke[(3 * L + 0) * 24 + 3 * L + 0];
        d1 += Ee * ke[(3 * L + 1) * 24 + 3 * L + 1];
        d2 += Ee * ke[(3 * L + 2) * 24 + 3 * L + 2];
      }
  if (invert) {
    const uint8_t fm = fixmask ? fixmask[n] : 0;
    d0 = ((fm & 1) || !(d0 > 0)) ? T(0) : T(1) / d0;
    d1 = ((fm & 2) || !(d1 > 0)) ? T(0) : T(1) / d1;
    d2 = ((fm & 4) || !(d2 > 0)) ? T(0) : T(1) / d2;
  }
  st3(out, n, d0, d1, d2);
}

template <typename T>
void launch_diag_minv(const Grid& g, const T* ke, const float* E, const uint8_t* fixmask,
                      T* out, bool invert, hipStream_t st) {
  hipLaunchKernelGGL((k_diag<T>), dim3(grid_blocks(g.nn)), dim3(kBlock), 0, st, g, ke, E, fixmask,
                     out, invert ? 1 : 0);
}

template <typename T>
__global__ void k_mask_fixed(Grid g, T* x, const uint8_t* __restrict__ fixmask) {
  const int n = blockIdx.x * kBlock + threadIdx.x;
  if (n >= g.nn) return;
  const uint8_t fm = fixmask[n];
  if (fm & 1) x[3ll * n] = 0;
  if (fm & 2) x[3ll * n + 1] = 0;
  if (fm & 4) x[3ll * n + 2] = 0;
}
template <typename T>
void launch_mask_fixed(const Grid& g, T* x, const uint8_t* fixmask, hipStream_t st) {
  hipLaunchKernelGGL((k_mask_fixed<T>), dim3(grid_blocks(g.nn)), dim3(kBlock), 0, st, g, x,
                     fixmask);
}

// ------------------------------------------------------------ PCG vectors
// Fixed DOFs are exactly the DOFs with M^-1 == 0 (set by k_diag), so the
// vector kernels need no separate mask: r = 0 there, hence z = p = 0 there.

// r = b - A x0 (or b), z = M^-1 r, p = z; partials rz, rr, bb.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_pcg_init(int nn, const double* __restrict__ f,
                                                    const T* __restrict__ Ax, int use_x0,
                                                    T* __restrict__ x, T* __restrict__ r,
                                                    T* __restrict__ z, T* __restrict__ p,
                                                    const T* __restrict__ minv, DevScalars* sc,
                                                    double* partials, double tol,
                                                    int max_iter) {
  double acc[3] = {0.0, 0.0, 0.0};  // rz, rr, bb
  for (long long d = blockIdx.x * (long long)kBlock + threadIdx.x; d < 3ll * nn;
       d += (long long)gridDim.x * kBlock) {
    const T mi = minv[d];
    const bool freed = mi != T(0);
    const double b = freed ? f[d] : 0.0;
    T xd = (use_x0 && freed) ? x[d] : T(0);
    x[d] = xd;
    const T rd = freed ? (T)(b - (use_x0 ? (double)Ax[d] : 0.0)) : T(0);
    const T zd = mi * rd;
    r[d] = rd;
    z[d] = zd;
    p[d] = zd;
    acc[0] += (double)rd * (double)zd;
    acc[1] += (double)rd * (double)rd;
    acc[2] += b * b;
  }
  block_sum<3>(acc);
  if (threadIdx.x == 0) {
    partials[blockIdx.x] = acc[0];
    partials[kMaxPartials + blockIdx.x] = acc[1];
    partials[2 * kMaxPartials + blockIdx.x] = acc[2];
  }
  if (last_block(&sc->cnt[CNT_INIT])) {
    double s[3];
    sum_partials<3>(partials, gridDim.x, s);
    if (threadIdx.x == 0) {
      sc->rz = s[0];
      sc->rr = s[1];
      sc->rr0 = s[1];
      sc->bb = s[2];
      sc->tol2 = tol * tol * s[2];
      sc->max_iter = max_iter;
      sc->iter = 0;
      sc->status = GP_OK;
      sc->done = (s[2] == 0.0 || s[1] <= sc->tol2) ? 1 : 0;
      sc->cnt[CNT_INIT] = 0;
    }
  }
}

template <typename T>
void launch_pcg_init(const Grid& g, const double* f, const T* Ax, bool use_x0, T* x, T* r, T* z,
                     T* p, const T* minv, DevScalars* sc, double* partials, double tol,
                     int max_iter, hipStream_t st) {
  hipLaunchKernelGGL((k_pcg_init<T>), dim3(reduce_blocks(3ll * g.nn)), dim3(kBlock), 0, st, g.nn,
                     f, Ax, use_x0 ? 1 : 0, x, r, z, p, minv, sc, partials, tol, max_iter);
}

// x += a p; r -= a Ap; z = M^-1 r; then beta = rz_new / rz and the
// convergence test ||r||^2 <= tol^2 ||b||^2 in the last block.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_pcg_b(int nn, T* __restrict__ x, T* __restrict__ r,
                                                 T* __restrict__ z, const T* __restrict__ p,
                                                 const T* __restrict__ Ap,
                                                 const T* __restrict__ minv, DevScalars* sc,
                                                 double* partials) {
  if (sc->done) return;
  const T alpha = (T)sc->alpha;
  double acc[2] = {0.0, 0.0};  // rz, rr
  for (long long d = blockIdx.x * (long long)kBlock + threadIdx.x; d < 3ll * nn;
       d += (long long)gridDim.x * kBlock) {
    const T pd = p[d];
    x[d] += alpha * pd;
    const T rd = r[d] - alpha * Ap[d];
    r[d] = rd;
    const T zd = minv[d] * rd;
    z[d] = zd;
    acc[0] += (double)rd * (double)zd;
    acc[1] += (double)rd * (double)rd;
  }
  block_sum<2>(acc);
  if (threadIdx.x == 0) {
    partials[blockIdx.x] = acc[0];
    partials[kMaxPartials + blockIdx.x] = acc[1];
  }
  if (last_block(&sc->cnt[CNT_B])) {
    double s[2];
    sum_partials<2>(partials, gridDim.x, s);
    if (threadIdx.x == 0) {
      const int it = sc->iter + 1;
      sc->iter = it;
      sc->rr = s[1];
      if (s[1] <= sc->tol2) {
        sc->done = 1;
      } else if (it >= sc->max_iter) {
        sc->done = 1;
        sc->status = GP_NOT_CONVERGED;
      } else {
        sc->beta = s[0] / sc->rz;
        sc->rz = s[0];
      }
      sc->cnt[CNT_B] = 0;
    }
  }
}

template <typename T>
void launch_pcg_b(const Grid& g, T* x, T* r, T* z, const T* p, const T* Ap, const T* minv,
                  DevScalars* sc, double* partials, hipStream_t st) {
  hipLaunchKernelGGL((k_pcg_b<T>), dim3(reduce_blocks(3ll * g.nn)), dim3(kBlock), 0, st, g.nn, x,
                     r, z, p, Ap, minv, sc, partials);
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_pcg_c(int nn, T* __restrict__ p,
                                                 const T* __restrict__ z, DevScalars* sc) {
  if (sc->done) return;
  const T beta = (T)sc->beta;
  for (long long d = blockIdx.x * (long long)kBlock + threadIdx.x; d < 3ll * nn;
       d += (long long)gridDim.x * kBlock)
    p[d] = z[d] + beta * p[d];
}
template <typename T>
void launch_pcg_c(const Grid& g, T* p, const T* z, DevScalars* sc, hipStream_t st) {
  hipLaunchKernelGGL((k_pcg_c<T>), dim3(grid_blocks(3ll * g.nn)), dim3(kBlock), 0, st, g.nn, p,
                     z, sc);
}

// ------------------------------------------------------- FP64 refinement
__global__ __launch_bounds__(kBlock) void k_residual64(int nn, const double* __restrict__ f,
                                                      const double* __restrict__ Ku,
                                                      const uint8_t* __restrict__ fixmask,
                                                      double* __restrict__ r, DevScalars* sc,
                                                      double* partials) {
  double acc[1] = {0.0};
  for (long long d = blockIdx.x * (long long)kBlock + threadIdx.x; d < 3ll * nn;
       d += (long long)gridDim.x * kBlock) {
    const bool fixed = (fixmask[d / 3] >> (d % 3)) & 1;
    const double rd = fixed ? 0.0 : f[d] - Ku[d];
    r[d] = rd;
    acc[0] += rd * rd;
  }
  block_sum<1>(acc);
  if (threadIdx.x == 0) partials[blockIdx.x] = acc[0];
  if (last_block(&sc->cnt[CNT_RES])) {
    double s[1];
    sum_partials<1>(partials, gridDim.x, s);
    if (threadIdx.x == 0) {
      sc->dot = s[0];
      sc->cnt[CNT_RES] = 0;
    }
  }
}
void launch_residual64(const Grid& g, const double* f, const double* Ku, const uint8_t* fixmask,
                       double* r, DevScalars* sc, double* partials, hipStream_t st) {
  hipLaunchKernelGGL(k_residual64, dim3(reduce_blocks(3ll * g.nn)), dim3(kBlock), 0, st, g.nn, f,
                     Ku, fixmask, r, sc, partials);
}

template <typename T>
__global__ void k_add_correction(long long n, double* __restrict__ u, const T* __restrict__ d) {
  const long long i = blockIdx.x * (long long)kBlock + threadIdx.x;
  if (i < n) u[i] += (double)d[i];
}
template <typename T>
void launch_add_correction(const Grid& g, double* u, const T* d, hipStream_t st) {
  hipLaunchKernelGGL((k_add_correction<T>), dim3(grid_blocks(g.ndof)), dim3(kBlock), 0, st,
                     g.ndof, u, d);
}

__global__ void k_cast64to32(const double* a, float* b, long long n) {
  const long long i = blockIdx.x * (long long)kBlock + threadIdx.x;
  if (i < n) b[i] = (float)a[i];
}
__global__ void k_cast32to64(const float* a, double* b, long long n) {
  const long long i = blockIdx.x * (long long)kBlock + threadIdx.x;
  if (i < n) b[i] = (double)a[i];
}
void launch_cast_f64_to_f32(const double* a, float* b, long long n, hipStream_t st) {
  hipLaunchKernelGGL(k_cast64to32, dim3(grid_blocks(n)), dim3(kBlock), 0, st, a, b, n);
}
void launch_cast_f32_to_f64(const float* a, double* b, long long n, hipStream_t st) {
  hipLaunchKernelGGL(k_cast32to64, dim3(grid_blocks(n)), dim3(kBlock), 0, st, a, b, n);
}

// --------------------------------------------------------- material law
__device__ inline double law_E(double x, int law, double p, double E0, double Emin) {
  if (law == GP_LAW_CLAMP) return pow(fmax(x, 1e-3), p);
  return Emin + pow(x, p) * (E0 - Emin);
}
__device__ inline double law_dE(double x, int law, double p, double E0, double Emin) {
  if (law == GP_LAW_CLAMP) return p * pow(fmax(x, 1e-3), p - 1.0);
  return p * pow(x, p - 1.0) * (E0 - Emin);
}

__global__ void k_stiffness(int ne, const double* __restrict__ x, float* __restrict__ E, int law,
                            double p, double E0, double Emin) {
  const int e = blockIdx.x * kBlock + threadIdx.x;
  if (e < ne) E[e] = (float)law_E(x[e], law, p, E0, Emin);
}
void launch_stiffness(const Grid& g, const double* xPhys, float* E, int law, double penal,
                      double E0, double Emin, hipStream_t st) {
  hipLaunchKernelGGL(k_stiffness, dim3(grid_blocks(g.ne)), dim3(kBlock), 0, st, g.ne, xPhys, E,
                     law, penal, E0, Emin);
}

// ----------------------------------------------------------- sensitivity
// One thread per element, FP64 energy u_e^T KE u_e (simp_solver.py:215-218).
template <typename T>
__global__ __launch_bounds__(kBlock) void k_sens(Grid g, const double* __restrict__ ke,
                                                const double* __restrict__ xPhys,
                                                const T* __restrict__ u, int law, double p,
                                                double E0, double Emin, double* __restrict__ dc,
                                                double* ce_sum, DevScalars* sc,
                                                double* partials) {
  const int e = blockIdx.x * kBlock + threadIdx.x;
  const bool valid = e < g.ne;
  const int ec = valid ? e : 0;
  const int i = ec / g.eI, rem = ec - i * g.eI, j = rem / g.eJ, k = rem - j * g.eJ;
  const int n0 = i * g.sI + j * g.sJ + k;
  double ue[24];
#pragma unroll
  for (int m = 0; m < 8; ++m) {
    const int nm = n0 + CX[m] * g.sI + CY[m] * g.sJ + CZ[m];
    const V3<T> v = ld3(u, nm);
    ue[3 * m] = (double)v.x;
    ue[3 * m + 1] = (double)v.y;
    ue[3 * m + 2] = (double)v.z;
  }
  double ce = 0.0;
#pragma unroll
  for (int b = 0; b < 24; ++b) {
    double t = 0.0;
#pragma unroll
    for (int a = 0; a < 24; ++a) t += ue[a] * ke[a * 24 + b];
    ce += t * ue[b];
  }
  const double x = valid ? xPhys[e] : 0.0;
  const double dcv = -law_dE(x, law, p, E0, Emin) * ce;
  if (valid) {
    dc[e] = dcv;
    if (dcv != dcv) sc->nan_flag = 1;
  }
  if (ce_sum != nullptr) {
    double v[1] = {valid ? law_E(x, law, p, E0, Emin) * ce : 0.0};
    block_sum<1>(v);
    if (threadIdx.x == 0) partials[blockIdx.x] = v[0];
    if (last_block(&sc->cnt[CNT_SENS])) {
      double s[1];
      sum_partials<1>(partials, gridDim.x, s);
      if (threadIdx.x == 0) {
        *ce_sum = s[0];
        sc->cnt[CNT_SENS] = 0;
      }
    }
  }
}
template <typename T>
void launch_sens(const Grid& g, const double* ke64, const double* xPhys, const T* u, int law,
                 double penal, double E0, double Emin, double* dc, double* ce_sum,
                 DevScalars* sc, double* partials, hipStream_t st) {
  const int nb = grid_blocks(g.ne);  // partials buffer holds >= nb entries
  hipLaunchKernelGGL((k_sens<T>), dim3(nb), dim3(kBlock), 0, st, g, ke64, xPhys, u, law, penal,
                     E0, Emin, dc, ce_sum, sc, partials);
}

// ---------------------------------------------------------------- filter
// out_e = (sum_o w_o in_{e+o}) / Hs_e            (mode 0, simp_solver.py:240)
// out_e =  sum_o w_o (in_{e+o} / Hs_{e+o})       (mode 1, simp_solver.py:223)
// Offsets run in lexicographic (di,dj,dk) order, i.e. ascending neighbour
// index, and products are not contracted into FMAs: the same operation order
// as a CSR row product of the assembled filter matrix.
__global__ __launch_bounds__(kBlock) void k_filter(Grid g, const Stencil* __restrict__ s,
                                                  const double* __restrict__ Hs,
                                                  const double* __restrict__ in,
                                                  double* __restrict__ out, int mode) {
#pragma clang fp contract(off)
  const int e = blockIdx.x * kBlock + threadIdx.x;
  if (e >= g.ne) return;
  const int i = e / g.eI, rem = e - i * g.eI, j = rem / g.eJ, k = rem - j * g.eJ;
  const int no = s->n;
  double acc = 0.0;
  for (int o = 0; o < no; ++o) {
    const int ii = i + s->di[o], jj = j + s->dj[o], kk = k + s->dk[o];
    if ((unsigned)ii < (unsigned)g.nx && (unsigned)jj < (unsigned)g.ny &&
        (unsigned)kk < (unsigned)g.nz) {
      const int f = ii * g.eI + jj * g.eJ + kk;
      const double v = mode ? in[f] / Hs[f] : in[f];
      const double t = s->w[o] * v;
      acc = acc + t;
    }
  }
  out[e] = mode ? acc : acc / Hs[e];
}
void launch_filter(const Grid& g, const Stencil* s, const double* Hs, const double* in,
                   double* out, int mode, hipStream_t st) {
  hipLaunchKernelGGL(k_filter, dim3(grid_blocks(g.ne)), dim3(kBlock), 0, st, g, s, Hs, in, out,
                     mode);
}

__global__ void k_hs(Grid g, const Stencil* __restrict__ s, double* __restrict__ Hs) {
#pragma clang fp contract(off)
  const int e = blockIdx.x * kBlock + threadIdx.x;
  if (e >= g.ne) return;
  const int i = e / g.eI, rem = e - i * g.eI, j = rem / g.eJ, k = rem - j * g.eJ;
  double acc = 0.0;
  for (int o = 0; o < s->n; ++o) {
    const int ii = i + s->di[o], jj = j + s->dj[o], kk = k + s->dk[o];
    if ((unsigned)ii < (unsigned)g.nx && (unsigned)jj < (unsigned)g.ny &&
        (unsigned)kk < (unsigned)g.nz)
      acc = acc + s->w[o] * 1.0;
  }
  Hs[e] = acc;
}
void launch_hs(const Grid& g, const Stencil* s, double* Hs, hipStream_t st) {
  hipLaunchKernelGGL(k_hs, dim3(grid_blocks(g.ne)), dim3(kBlock), 0, st, g, s, Hs);
}

// OC volume weights: mean(H xnew / Hs) over non-preserved rows equals
// sum_f w_f xnew_f / n with w_f = sum_{e not preserved} H_ef / Hs_e.
__global__ void k_oc_weights(Grid g, const Stencil* __restrict__ s, const double* __restrict__ Hs,
                             const uint8_t* __restrict__ pres, double* __restrict__ w) {
  const int f = blockIdx.x * kBlock + threadIdx.x;
  if (f >= g.ne) return;
  const int i = f / g.eI, rem = f - i * g.eI, j = rem / g.eJ, k = rem - j * g.eJ;
  double acc = 0.0;
  for (int o = 0; o < s->n; ++o) {
    const int ii = i + s->di[o], jj = j + s->dj[o], kk = k + s->dk[o];
    if ((unsigned)ii < (unsigned)g.nx && (unsigned)jj < (unsigned)g.ny &&
        (unsigned)kk < (unsigned)g.nz) {
      const int e = ii * g.eI + jj * g.eJ + kk;
      if (pres == nullptr || !pres[e]) acc += s->w[o] / Hs[e];
    }
  }
  w[f] = acc;
}
void launch_oc_weights(const Grid& g, const Stencil* s, const double* Hs, const uint8_t* pres,
                       double* w, hipStream_t st) {
  hipLaunchKernelGGL(k_oc_weights, dim3(grid_blocks(g.ne)), dim3(kBlock), 0, st, g, s, Hs, pres,
                     w);
}

// ------------------------------------------------------------------- OC
// l1 = 0; l2 = 1e9 (cpu) or 2 max(max(-dc), 1) (gpu variant).
__global__ __launch_bounds__(kBlock) void k_oc_init(int ne, const double* __restrict__ dc,
                                                   int variant, double volfrac, double move,
                                                   DevScalars* sc, double* partials) {
  double m = -INFINITY;
  if (variant == GP_OC_GPU)
    for (int e = blockIdx.x * kBlock + threadIdx.x; e < ne; e += gridDim.x * kBlock)
      m = fmax(m, -dc[e]);
  m = wave_max(m);
  __shared__ double sh[kBlock / 64];
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int i = 0; i < kBlock / 64; ++i) m = fmax(m, sh[i]);
    partials[blockIdx.x] = m;
  }
  if (last_block(&sc->cnt[CNT_MAX])) {
    if (threadIdx.x == 0) {
      double mm = -INFINITY;
      for (int b = 0; b < (int)gridDim.x; ++b) mm = fmax(mm, partials[b]);
      sc->l1 = 0.0;
      sc->l2 = (variant == GP_OC_GPU) ? fmax(mm, 1.0) * 2.0 : 1e9;
      sc->oc_eps = (variant == GP_OC_GPU) ? 1e-15 : 1e-12;
      sc->oc_tol = (variant == GP_OC_GPU) ? 1e-5 : 1e-4;
      sc->oc_floor = (variant == GP_OC_GPU) ? 1e-15 : 1e-12;
      sc->oc_max = (variant == GP_OC_GPU) ? 30 : 50;
      sc->volfrac = volfrac;
      sc->move = move;
      sc->oc_steps = 0;
      sc->oc_done = 0;
      sc->cnt[CNT_MAX] = 0;
    }
  }
}
void launch_oc_init(const Grid& g, const double* dc, int variant, double volfrac, double move,
                    DevScalars* sc, double* partials, hipStream_t st) {
  hipLaunchKernelGGL(k_oc_init, dim3(reduce_blocks(g.ne)), dim3(kBlock), 0, st, g.ne, dc, variant,
                     volfrac, move, sc, partials);
}

__device__ inline double oc_xnew(double x, double dc, double lmid, double floor_, double move) {
  double v = x * sqrt(fmax(-dc / lmid, floor_));
  v = fmin(fmax(v, 0.0), 1.0);
  return fmin(fmax(v, x - move), x + move);
}

// One bisection step (simp_solver.py:230-248): the stop test on (l1,l2)
// precedes the volume branch, exactly as in the reference loop body.
__global__ __launch_bounds__(kBlock) void k_oc_step(int ne, const double* __restrict__ x,
                                                   const double* __restrict__ dc,
                                                   const double* __restrict__ w,
                                                   const uint8_t* __restrict__ pres,
                                                   double n_pres, DevScalars* sc,
                                                   double* partials) {
  if (sc->oc_done) return;
  const double l1 = sc->l1, l2 = sc->l2;
  const double lmid = 0.5 * (l2 + l1);
  if ((l2 - l1) / (l1 + l2 + sc->oc_eps) < sc->oc_tol) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      sc->lmid = lmid;
      sc->oc_steps += 1;
      sc->oc_done = 1;
    }
    return;
  }
  const double fl = sc->oc_floor, mv = sc->move;
  double acc[1] = {0.0};
  for (int e = blockIdx.x * kBlock + threadIdx.x; e < ne; e += gridDim.x * kBlock) {
    const double xn = (pres != nullptr && pres[e]) ? 1.0 : oc_xnew(x[e], dc[e], lmid, fl, mv);
    acc[0] += w[e] * xn;
  }
  block_sum<1>(acc);
  if (threadIdx.x == 0) partials[blockIdx.x] = acc[0];
  if (last_block(&sc->cnt[CNT_OC])) {
    double s[1];
    sum_partials<1>(partials, gridDim.x, s);
    if (threadIdx.x == 0) {
      const double mean = (s[0] + n_pres) / (double)ne;
      if (mean > sc->volfrac)
        sc->l1 = lmid;
      else
        sc->l2 = lmid;
      sc->lmid = lmid;
      const int st = sc->oc_steps + 1;
      sc->oc_steps = st;
      if (st >= sc->oc_max) sc->oc_done = 1;
      sc->cnt[CNT_OC] = 0;
    }
  }
}
void launch_oc_step(const Grid& g, const double* x, const double* dc, const double* w,
                    const uint8_t* pres, double n_pres, DevScalars* sc, double* partials,
                    hipStream_t st) {
  hipLaunchKernelGGL(k_oc_step, dim3(reduce_blocks(g.ne)), dim3(kBlock), 0, st, g.ne, x, dc, w,
                     pres, n_pres, sc, partials);
}

__global__ void k_oc_final(int ne, const double* __restrict__ x, const double* __restrict__ dc,
                           const uint8_t* __restrict__ pres, const DevScalars* sc,
                           double* __restrict__ xnew) {
  const int e = blockIdx.x * kBlock + threadIdx.x;
  if (e >= ne) return;
  xnew[e] = (pres != nullptr && pres[e]) ? 1.0 : oc_xnew(x[e], dc[e], sc->lmid, sc->oc_floor, sc->move);
}
void launch_oc_final(const Grid& g, const double* x, const double* dc, const uint8_t* pres,
                     DevScalars* sc, double* xnew, hipStream_t st) {
  hipLaunchKernelGGL(k_oc_final, dim3(grid_blocks(g.ne)), dim3(kBlock), 0, st, g.ne, x, dc, pres,
                     sc, xnew);
}

__global__ void k_set_preserved(int ne, const uint8_t* __restrict__ pres, double* xPhys) {
  const int e = blockIdx.x * kBlock + threadIdx.x;
  if (e < ne && pres[e]) xPhys[e] = 1.0;
}
void launch_set_preserved(const Grid& g, const uint8_t* pres, double* xPhys, hipStream_t st) {
  if (pres == nullptr) return;
  hipLaunchKernelGGL(k_set_preserved, dim3(grid_blocks(g.ne)), dim3(kBlock), 0, st, g.ne, pres,
                     xPhys);
}

// ------------------------------------------------------------ dot (FP64)
template <typename TA, typename T>
__global__ __launch_bounds__(kBlock) void k_dot(long long n, const TA* __restrict__ a,
                                               const T* __restrict__ b, DevScalars* sc,
                                               double* partials) {
  double acc[1] = {0.0};
  for (long long i = blockIdx.x * (long long)kBlock + threadIdx.x; i < n;
       i += (long long)gridDim.x * kBlock)
    acc[0] += (double)a[i] * (double)b[i];
  block_sum<1>(acc);
  if (threadIdx.x == 0) partials[blockIdx.x] = acc[0];
  if (last_block(&sc->cnt[CNT_DOT])) {
    double s[1];
    sum_partials<1>(partials, gridDim.x, s);
    if (threadIdx.x == 0) {
      sc->dot = s[0];
      sc->cnt[CNT_DOT] = 0;
    }
  }
}
template <typename TA, typename T>
void launch_dot(const TA* a, const T* b, long long n, DevScalars* sc, double* partials,
                hipStream_t st) {
  hipLaunchKernelGGL((k_dot<TA, T>), dim3(reduce_blocks(n)), dim3(kBlock), 0, st, n, a, b, sc,
                     partials);
}

// ------------------------------------------------- generic-table kernels
// Reference semantics (cuda_kernels/simp_sensitivity_ptx.cu:163-270, 301-373):
// arbitrary edof rows, E = max(x,1e-3)^p, FP64 element arithmetic.
__global__ __launch_bounds__(kBlock) void k_generic_sens(const float* __restrict__ xPhys,
                                                        const float* __restrict__ u,
                                                        const float* __restrict__ ke,
                                                        const int32_t* __restrict__ edof,
                                                        long long ne, double p,
                                                        float* __restrict__ dc) {
  const long long e = blockIdx.x * (long long)kBlock + threadIdx.x;
  if (e >= ne) return;
  double ue[24];
#pragma unroll
  for (int a = 0; a < 24; ++a) ue[a] = (double)u[edof[e * 24 + a]];
  double ce = 0.0;
#pragma unroll
  for (int a = 0; a < 24; ++a) {
    double t = 0.0;
#pragma unroll
    for (int b = 0; b < 24; ++b) t += (double)ke[a * 24 + b] * ue[b];
    ce += ue[a] * t;
  }
  dc[e] = (float)(-p * pow(fmax((double)xPhys[e], 1e-3), p - 1.0) * ce);
}
void launch_generic_sens(const float* xPhys, const float* u, const float* ke, const int32_t* edof,
                         long long ne, double penal, float* dc, hipStream_t st) {
  hipLaunchKernelGGL(k_generic_sens, dim3(grid_blocks(ne)), dim3(kBlock), 0, st, xPhys, u, ke,
                     edof, ne, penal, dc);
}

__global__ __launch_bounds__(kBlock) void k_generic_spmv(const float* __restrict__ xPhys,
                                                        const float* __restrict__ u,
                                                        const float* __restrict__ ke,
                                                        const int32_t* __restrict__ edof,
                                                        long long ne, double p,
                                                        float* __restrict__ y) {
  const long long e = blockIdx.x * (long long)kBlock + threadIdx.x;
  if (e >= ne) return;
  const double Ee = pow(fmax((double)xPhys[e], 1e-3), p);
  double ue[24];
#pragma unroll
  for (int a = 0; a < 24; ++a) ue[a] = (double)u[edof[e * 24 + a]];
#pragma unroll 4
  for (int a = 0; a < 24; ++a) {
    double t = 0.0;
#pragma unroll
    for (int b = 0; b < 24; ++b) t += (double)ke[a * 24 + b] * ue[b];
    atomicAdd(&y[edof[e * 24 + a]], (float)(Ee * t));
  }
}
void launch_generic_spmv(const float* xPhys, const float* u, const float* ke, const int32_t* edof,
                         long long ne, double penal, float* y, hipStream_t st) {
  hipLaunchKernelGGL(k_generic_spmv, dim3(grid_blocks(ne)), dim3(kBlock), 0, st, xPhys, u, ke,
                     edof, ne, penal, y);
}

__global__ void k_csr_spmv(const int64_t* __restrict__ rp, const int64_t* __restrict__ ci,
                           const float* __restrict__ v, const float* __restrict__ x,
                           float* __restrict__ y, long long n) {
  const long long r = blockIdx.x * (long long)kBlock + threadIdx.x;
  if (r >= n) return;
  float s = 0.0f;
  for (long long q = rp[r]; q < rp[r + 1]; ++q) s += v[q] * x[ci[q]];
  y[r] = s;
}
void launch_csr_spmv(const int64_t* rp, const int64_t* ci, const float* v, const float* x, float* y,
                     long long n, hipStream_t st) {
  hipLaunchKernelGGL(k_csr_spmv, dim3(grid_blocks(n)), dim3(kBlock), 0, st, rp, ci, v, x, y, n);
}

__global__ void k_reset_flags(DevScalars* sc) { sc->nan_flag = 0; }
void launch_reset_flags(DevScalars* sc, hipStream_t st) {
  hipLaunchKernelGGL(k_reset_flags, dim3(1), dim3(1), 0, st, sc);
}

// ---------------------------------------------------- CSR PCG helpers
__global__ void k_zero_list(float* v, const int64_t* __restrict__ idx, long long n) {
  const long long i = blockIdx.x * (long long)kBlock + threadIdx.x;
  if (i < n) v[idx[i]] = 0.0f;
}
void launch_zero_list(float* v, const int64_t* idx, long long n, hipStream_t st) {
  if (n > 0)
    hipLaunchKernelGGL(k_zero_list, dim3(grid_blocks(n)), dim3(kBlock), 0, st, v, idx, n);
}
// mode 0: y = x + a*y ; mode 1: y = y + a*x ; mode 2: y = x / d
__global__ void k_vec(float* __restrict__ y, const float* __restrict__ x,
                      const float* __restrict__ d, float a, long long n, int mode) {
  const long long i = blockIdx.x * (long long)kBlock + threadIdx.x;
  if (i >= n) return;
  if (mode == 0) y[i] = x[i] + a * y[i];
  else if (mode == 1) y[i] = y[i] + a * x[i];
  else y[i] = x[i] / d[i];
}
void launch_vec(float* y, const float* x, const float* d, float a, long long n, int mode,
                hipStream_t st) {
  hipLaunchKernelGGL(k_vec, dim3(grid_blocks(n)), dim3(kBlock), 0, st, y, x, d, a, n, mode);
}

// ------------------------------------------------ explicit instantiations
template void launch_ku<float>(const Grid&, const float*, const float*, const float*, float*,
                               const uint8_t*, DevScalars*, double*, bool, hipStream_t);
template void launch_ku<double>(const Grid&, const double*, const float*, const double*, double*,
                                const uint8_t*, DevScalars*, double*, bool, hipStream_t);
template void launch_diag_minv<float>(const Grid&, const float*, const float*, const uint8_t*,
                                      float*, bool, hipStream_t);
template void launch_diag_minv<double>(const Grid&, const double*, const float*, const uint8_t*,
                                       double*, bool, hipStream_t);
template void launch_mask_fixed<float>(const Grid&, float*, const uint8_t*, hipStream_t);
template void launch_mask_fixed<double>(const Grid&, double*, const uint8_t*, hipStream_t);
template void launch_pcg_init<float>(const Grid&, const double*, const float*, bool, float*,
                                     float*, float*, float*, const float*, DevScalars*, double*,
                                     double, int, hipStream_t);
template void launch_pcg_init<double>(const Grid&, const double*, const double*, bool, double*,
                                      double*, double*, double*, const double*, DevScalars*,
                                      double*, double, int, hipStream_t);
template void launch_pcg_b<float>(const Grid&, float*, float*, float*, const float*, const float*,
                                  const float*, DevScalars*, double*, hipStream_t);
template void launch_pcg_b<double>(const Grid&, double*, double*, double*, const double*,
                                   const double*, const double*, DevScalars*, double*,
                                   hipStream_t);
template void launch_pcg_c<float>(const Grid&, float*, const float*, DevScalars*, hipStream_t);
template void launch_pcg_c<double>(const Grid&, double*, const double*, DevScalars*, hipStream_t);
template void launch_add_correction<float>(const Grid&, double*, const float*, hipStream_t);
template void launch_add_correction<double>(const Grid&, double*, const double*, hipStream_t);
template void launch_sens<float>(const Grid&, const double*, const double*, const float*, int,
                                 double, double, double, double*, double*, DevScalars*, double*,
                                 hipStream_t);
template void launch_sens<double>(const Grid&, const double*, const double*, const double*, int,
                                  double, double, double, double*, double*, DevScalars*, double*,
                                  hipStream_t);
template void launch_dot<double, float>(const double*, const float*, long long, DevScalars*,
                                       double*, hipStream_t);
template void launch_dot<double, double>(const double*, const double*, long long, DevScalars*,
                                        double*, hipStream_t);
template void launch_dot<float, float>(const float*, const float*, long long, DevScalars*, double*,
                                      hipStream_t);

}  // namespace gp
