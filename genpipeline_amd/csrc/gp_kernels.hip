// HIP kernels of the SIMP inner loop for gfx950 (CDNA4, wave64): the dense
// gather K.u (any 24x24 table), Jacobi-PCG vector kernels, FP64 refinement,
// material law, sensitivity, density filter, OC bisection and the
// generic-table kernels of the reference kernel API.
//
// Grid numbering follows genpipeline/topology/simp_solver.py:251-282 exactly;
// nothing is gathered through an edof table on the structured path -- every
// index is derived arithmetically from (i,j,k).
//
// Reductions: a producing kernel writes one FP64 partial per block (fixed
// butterfly order inside the block); each block of the consuming kernel sums
// the producer's partials in index order. All blocks therefore agree bit for
// bit on every scalar, and no kernel uses device-scope fences or atomics.
#include "gp_bodies.hpp"

namespace gp {

// One block: *out = sum of p[0..n).
__global__ __launch_bounds__(kBlock) void k_sum_fin(const double* __restrict__ p, int n,
                                                   double* out) {
  body_sum_fin(p, n, out);
}

// ------------------------------------------------------- K . u (gather)
// Any 24x24 table: one thread per node, the 27-node neighbourhood in
// registers, each of the (up to) 8 elements adding E_e * KE[3L+r, :] u_e,
// with L the node's corner index in that element.
template <typename T, bool FUSE>
__global__ __launch_bounds__(kBlock) void k_ku(Grid g, const T* __restrict__ ke,
                                              const float* __restrict__ E,
                                              const T* __restrict__ u, T* __restrict__ y,
                                              const uint8_t* __restrict__ fixmask,
                                              const DevScalars* sc, double* partA) {
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
    const double d = valid ? (double)un[13][0] * (double)acc0 + (double)un[13][1] * (double)acc1 +
                                 (double)un[13][2] * (double)acc2
                           : 0.0;
    const double s = block_sum0(d);
    if (threadIdx.x == 0) partA[blockIdx.x] = s;
  }
}

int gather_blocks(const Grid& g) { return grid_blocks(g.nn); }

template <typename T>
void launch_ku(const Grid& g, const T* ke, const float* E, const T* u, T* y,
               const uint8_t* fixmask, const DevScalars* sc, double* partA, bool fuse_pap,
               hipStream_t st) {
  const int nb = gather_blocks(g);
  if (fuse_pap)
    hipLaunchKernelGGL((k_ku<T, true>), dim3(nb), dim3(kBlock), 0, st, g, ke, E, u, y, fixmask,
                       sc, partA);
  else
    hipLaunchKernelGGL((k_ku<T, false>), dim3(nb), dim3(kBlock), 0, st, g, ke, E, u, y, fixmask,
                       sc, partA);
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_diag(Grid g, const T* __restrict__ ke,
                                                const float* __restrict__ E,
                                                const uint8_t* __restrict__ fixmask,
                                                T* __restrict__ out, int invert) {
  body_diag<T>(g, ke, E, fixmask, out, invert, blockIdx.x, gridDim.x);
}

template <typename T>
void launch_diag_minv(const Grid& g, const T* ke, const float* E, const uint8_t* fixmask,
                      T* out, bool invert, hipStream_t st) {
  hipLaunchKernelGGL((k_diag<T>), dim3(grid_blocks(g.nn)), dim3(kBlock), 0, st, g, ke, E, fixmask,
                     out, invert ? 1 : 0);
}

void launch_diag_node(const Grid& g, const float* ke, const float* E, float* out,
                      hipStream_t st) {
  hipLaunchKernelGGL((k_diag<float>), dim3(grid_blocks(g.nn)), dim3(kBlock), 0, st, g, ke, E,
                     nullptr, out, 2);
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
// Vector kernels run a fixed grid of vec_blocks() blocks, grid-striding.
int vec_blocks(const Grid& g) {
  const int b = grid_blocks(3ll * g.nn);
  return b < kVecBlocks ? b : kVecBlocks;
}

// r = b - A x0 (or b), z = M^-1 r, p = z; partials rz -> RZ[0], rr, bb.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_pcg_init(int nn, const double* __restrict__ f,
                                                    const T* __restrict__ Ax, int use_x0,
                                                    T* __restrict__ x, T* __restrict__ r,
                                                    T* __restrict__ z, T* __restrict__ p,
                                                    const T* __restrict__ minv, Parts parts) {
  double rz = 0.0, rr = 0.0, bb = 0.0;
  for (long long d = blockIdx.x * (long long)kBlock + threadIdx.x; d < 3ll * nn;
       d += (long long)gridDim.x * kBlock) {
    const T mi = minv[d];
    const bool freed = mi != T(0);
    const double b = freed ? f[d] : 0.0;
    const T xd = (use_x0 && freed) ? x[d] : T(0);
    x[d] = xd;
    const T rd = freed ? (T)(b - (use_x0 ? (double)Ax[d] : 0.0)) : T(0);
    const T zd = mi * rd;
    r[d] = rd;
    z[d] = zd;
    p[d] = zd;
    rz += (double)rd * (double)zd;
    rr += (double)rd * (double)rd;
    bb += b * b;
  }
  rz = block_sum0(rz);
  rr = block_sum0(rr);
  bb = block_sum0(bb);
  if (threadIdx.x == 0) {
    parts.RZ[0][blockIdx.x] = rz;
    parts.RR[blockIdx.x] = rr;
    parts.BB[blockIdx.x] = bb;
  }
}

__global__ __launch_bounds__(kBlock) void k_pcg_init_fin(Parts parts, int nV, double tol,
                                                        int max_iter, DevScalars* sc) {
  body_pcg_init_fin(parts, nV, tol, max_iter, sc);
}

void launch_pcg_init_fin(Parts parts, int nV, double tol, int max_iter, DevScalars* sc,
                         hipStream_t st) {
  hipLaunchKernelGGL(k_pcg_init_fin, dim3(1), dim3(kBlock), 0, st, parts, nV, tol, max_iter, sc);
}

template <typename T>
void launch_pcg_init(const Grid& g, const double* f, const T* Ax, bool use_x0, T* x, T* r, T* z,
                     T* p, const T* minv, Parts parts, DevScalars* sc, double tol, int max_iter,
                     hipStream_t st) {
  const int nV = vec_blocks(g);
  hipLaunchKernelGGL((k_pcg_init<T>), dim3(nV), dim3(kBlock), 0, st, g.nn, f, Ax, use_x0 ? 1 : 0,
                     x, r, z, p, minv, parts);
  hipLaunchKernelGGL(k_pcg_init_fin, dim3(1), dim3(kBlock), 0, st, parts, nV, tol, max_iter, sc);
}

// Iteration `it`: alpha = rz / pAp from the partials of the previous
// kernels; x += a p; r -= a Ap; z = M^-1 r; partials rz' -> RZ[(it+1)&1], rr.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_pcg_b(int it, int nn, int nA, T* __restrict__ x,
                                                 T* __restrict__ r, T* __restrict__ z,
                                                 const T* __restrict__ p,
                                                 const T* __restrict__ Ap,
                                                 const T* __restrict__ minv, Parts parts,
                                                 DevScalars* sc) {
  if (sc->done) return;
  const int q = it & 1;
  const int nV = gridDim.x;
  const double pAp = reduce_all(parts.A, nA);
  const double rz = reduce_all(parts.RZ[q], nV);
  if (!(pAp > 0.0)) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      sc->pAp = pAp;
      sc->status = GP_BREAKDOWN;
      sc->iter = it;
      sc->done = 1;
    }
    return;
  }
  const T alpha = (T)(rz / pAp);
  double rzn = 0.0, rr = 0.0;
  for (long long d = blockIdx.x * (long long)kBlock + threadIdx.x; d < 3ll * nn;
       d += (long long)gridDim.x * kBlock) {
    const T pd = p[d];
    x[d] += alpha * pd;
    const T rd = r[d] - alpha * Ap[d];
    r[d] = rd;
    const T zd = minv[d] * rd;
    z[d] = zd;
    rzn += (double)rd * (double)zd;
    rr += (double)rd * (double)rd;
  }
  rzn = block_sum0(rzn);
  rr = block_sum0(rr);
  if (threadIdx.x == 0) {
    parts.RZ[q ^ 1][blockIdx.x] = rzn;
    parts.RR[blockIdx.x] = rr;
  }
}

template <typename T>
void launch_pcg_b(const Grid& g, int it, int nA, T* x, T* r, T* z, const T* p, const T* Ap,
                  const T* minv, Parts parts, DevScalars* sc, hipStream_t st) {
  hipLaunchKernelGGL((k_pcg_b<T>), dim3(vec_blocks(g)), dim3(kBlock), 0, st, it, g.nn, nA, x, r,
                     z, p, Ap, minv, parts, sc);
}

// Iteration `it` epilogue: convergence test ||r||^2 <= tol^2 ||b||^2 and the
// iteration cap, then p = z + beta p with beta = rz' / rz.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_pcg_c(int it, int nn, T* __restrict__ p,
                                                 const T* __restrict__ z, Parts parts,
                                                 DevScalars* sc) {
  if (sc->done) return;
  const int q = it & 1;
  const int nV = gridDim.x;
  const double rzn = reduce_all(parts.RZ[q ^ 1], nV);
  const double rr = reduce_all(parts.RR, nV);
  const double rz = reduce_all(parts.RZ[q], nV);
  const bool conv = rr <= sc->tol2;
  const bool cap = it + 1 >= sc->max_iter;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    sc->iter = it + 1;
    sc->rr = rr;
    sc->rz = rzn;
    if (conv || cap) {
      sc->status = conv ? GP_OK : GP_NOT_CONVERGED;
      sc->done = 1;
    }
  }
  if (conv || cap) return;
  const T beta = (T)(rzn / rz);
  for (long long d = blockIdx.x * (long long)kBlock + threadIdx.x; d < 3ll * nn;
       d += (long long)gridDim.x * kBlock)
    p[d] = z[d] + beta * p[d];
}
template <typename T>
void launch_pcg_c(const Grid& g, int it, T* p, const T* z, Parts parts, DevScalars* sc,
                  hipStream_t st) {
  hipLaunchKernelGGL((k_pcg_c<T>), dim3(vec_blocks(g)), dim3(kBlock), 0, st, it, g.nn, p, z, parts,
                     sc);
}

// ------------------------------------------------------- FP64 refinement
__global__ __launch_bounds__(kBlock) void k_residual64(int nn, const double* __restrict__ f,
                                                      const double* __restrict__ Ku,
                                                      const uint8_t* __restrict__ fixmask,
                                                      double* __restrict__ r, double* partS) {
  double acc = 0.0;
  for (long long d = blockIdx.x * (long long)kBlock + threadIdx.x; d < 3ll * nn;
       d += (long long)gridDim.x * kBlock) {
    const bool fixed = (fixmask[d / 3] >> (d % 3)) & 1;
    const double rd = fixed ? 0.0 : f[d] - Ku[d];
    r[d] = rd;
    acc += rd * rd;
  }
  acc = block_sum0(acc);
  if (threadIdx.x == 0) partS[blockIdx.x] = acc;
}
void launch_residual64(const Grid& g, const double* f, const double* Ku, const uint8_t* fixmask,
                       double* r, Parts parts, DevScalars* sc, hipStream_t st) {
  const int nb = vec_blocks(g);
  hipLaunchKernelGGL(k_residual64, dim3(nb), dim3(kBlock), 0, st, g.nn, f, Ku, fixmask, r,
                     parts.S);
  hipLaunchKernelGGL(k_sum_fin, dim3(1), dim3(kBlock), 0, st, parts.S, nb, &sc->sum);
}

__global__ __launch_bounds__(kBlock) void k_res64(Grid g, const double* __restrict__ ke,
                                                 const float* __restrict__ E,
                                                 const double* __restrict__ u,
                                                 const double* __restrict__ f,
                                                 const uint8_t* __restrict__ fixmask,
                                                 double* __restrict__ r, Parts parts) {
  body_res64(g, ke, E, u, f, fixmask, r, parts.A, parts.RZ[1], blockIdx.x, gridDim.x);
}
__global__ __launch_bounds__(kBlock) void k_res_fin(Parts parts, int nV, DevScalars* sc) {
  body_res_fin(parts, nV, sc);
}
void launch_res64(const Grid& g, const double* ke, const float* E, const double* u,
                  const double* f, const uint8_t* fixmask, double* r, Parts parts, DevScalars* sc,
                  hipStream_t st) {
  const int nb = vec_blocks(g);
  hipLaunchKernelGGL(k_res64, dim3(nb), dim3(kBlock), 0, st, g, ke, E, u, f, fixmask, r, parts);
  hipLaunchKernelGGL(k_res_fin, dim3(1), dim3(kBlock), 0, st, parts, nb, sc);
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

template <typename S>
__global__ void k_add_correction_lp(long long n, double* __restrict__ u, const S* __restrict__ d,
                                    double scale) {
  const long long i = blockIdx.x * (long long)kBlock + threadIdx.x;
  if (i < n) u[i] += scale * (double)ld_s<float>(d[i]);
}
template <typename S>
void launch_add_correction_lp(const Grid& g, double* u, const S* d, double scale, hipStream_t st) {
  hipLaunchKernelGGL((k_add_correction_lp<S>), dim3(grid_blocks(g.ndof)), dim3(kBlock), 0, st,
                     g.ndof, u, d, scale);
}
template void launch_add_correction_lp<_Float16>(const Grid&, double*, const _Float16*, double,
                                                 hipStream_t);
template void launch_add_correction_lp<bf16>(const Grid&, double*, const bf16*, double,
                                             hipStream_t);

__global__ void k_scale64(long long n, double* __restrict__ v, double a) {
  const long long i = blockIdx.x * (long long)kBlock + threadIdx.x;
  if (i < n) v[i] *= a;
}
void launch_scale64(const Grid& g, double* v, double a, hipStream_t st) {
  hipLaunchKernelGGL(k_scale64, dim3(grid_blocks(g.ndof)), dim3(kBlock), 0, st, g.ndof, v, a);
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
__global__ void k_stiffness(int ne, const double* __restrict__ x, float* __restrict__ E, int law,
                            double p, double E0, double Emin) {
  body_stiffness(ne, x, E, law, p, E0, Emin, blockIdx.x, gridDim.x);
}
void launch_stiffness(const Grid& g, const double* xPhys, float* E, int law, double penal,
                      double E0, double Emin, hipStream_t st) {
  hipLaunchKernelGGL(k_stiffness, dim3(grid_blocks(g.ne)), dim3(kBlock), 0, st, g.ne, xPhys, E,
                     law, penal, E0, Emin);
}

// ----------------------------------------------------------- sensitivity
template <typename T, bool MODAL>
__global__ __launch_bounds__(kBlock) void k_sens(Grid g, const double* __restrict__ ke,
                                                ModalCoef<double> mk,
                                                const double* __restrict__ xPhys,
                                                const T* __restrict__ u, int law, double p,
                                                double E0, double Emin, double* __restrict__ dc,
                                                int want_sum, double* partS, DevScalars* sc) {
  body_sens<T, MODAL>(g, ke, mk, xPhys, u, law, p, E0, Emin, dc, want_sum, partS, sc, blockIdx.x,
                      gridDim.x);
}
template <typename T>
void launch_sens(const Grid& g, const double* ke64, const ModalCoef<double>* mk,
                 const double* xPhys, const T* u, int law, double penal, double E0, double Emin,
                 double* dc, double* ce_sum, Parts parts, DevScalars* sc, hipStream_t st) {
  const int nb = grid_blocks(g.ne);  // parts.cap >= nb
  const ModalCoef<double> m = mk ? *mk : ModalCoef<double>{};
  if (mk)
    hipLaunchKernelGGL((k_sens<T, true>), dim3(nb), dim3(kBlock), 0, st, g, ke64, m, xPhys, u,
                       law, penal, E0, Emin, dc, ce_sum != nullptr ? 1 : 0, parts.S, sc);
  else
    hipLaunchKernelGGL((k_sens<T, false>), dim3(nb), dim3(kBlock), 0, st, g, ke64, m, xPhys, u,
                       law, penal, E0, Emin, dc, ce_sum != nullptr ? 1 : 0, parts.S, sc);
  if (ce_sum != nullptr)
    hipLaunchKernelGGL(k_sum_fin, dim3(1), dim3(kBlock), 0, st, parts.S, nb, ce_sum);
}

// ---------------------------------------------------------------- filter
__global__ __launch_bounds__(kBlock) void k_filter(Grid g, const Stencil* __restrict__ s,
                                                  const double* __restrict__ Hs,
                                                  const double* __restrict__ in,
                                                  double* __restrict__ out, int mode) {
  body_filter(g, s, Hs, in, out, mode, blockIdx.x, gridDim.x);
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
int oc_blocks(const Grid& g) {
  const int b = grid_blocks(g.ne);
  return b < kVecBlocks ? b : kVecBlocks;
}

__global__ __launch_bounds__(kBlock) void k_oc_max(int ne, const double* __restrict__ dc,
                                                  double* partM) {
  body_oc_max(ne, dc, partM, blockIdx.x, gridDim.x);
}

__global__ __launch_bounds__(kBlock) void k_oc_init_fin(Parts parts, int nb, int variant,
                                                       double volfrac, double move,
                                                       DevScalars* sc) {
  body_oc_init_fin(parts, nb, variant, volfrac, move, sc);
}

void launch_oc_init(const Grid& g, const double* dc, int variant, double volfrac, double move,
                    Parts parts, DevScalars* sc, hipStream_t st) {
  const int nb = oc_blocks(g);
  if (variant == GP_OC_GPU)
    hipLaunchKernelGGL(k_oc_max, dim3(nb), dim3(kBlock), 0, st, g.ne, dc, parts.M);
  hipLaunchKernelGGL(k_oc_init_fin, dim3(1), dim3(kBlock), 0, st, parts, nb, variant, volfrac,
                     move, sc);
}

__global__ __launch_bounds__(kBlock) void k_oc_step(int s, int ne, const double* __restrict__ x,
                                                   const double* __restrict__ dc,
                                                   const double* __restrict__ w,
                                                   const uint8_t* __restrict__ pres,
                                                   double n_pres, Parts parts, DevScalars* sc) {
  body_oc_step(s, ne, x, dc, w, pres, n_pres, parts, sc, blockIdx.x, gridDim.x);
}
void launch_oc_step(const Grid& g, int s, const double* x, const double* dc, const double* w,
                    const uint8_t* pres, double n_pres, Parts parts, DevScalars* sc,
                    hipStream_t st) {
  hipLaunchKernelGGL(k_oc_step, dim3(oc_blocks(g)), dim3(kBlock), 0, st, s, g.ne, x, dc, w, pres,
                     n_pres, parts, sc);
}

__global__ void k_oc_final(int ne, const double* __restrict__ x, const double* __restrict__ dc,
                           const uint8_t* __restrict__ pres, const DevScalars* sc,
                           double* __restrict__ xnew) {
  body_oc_final(ne, x, dc, pres, sc, xnew, blockIdx.x, gridDim.x);
}
void launch_oc_final(const Grid& g, const double* x, const double* dc, const uint8_t* pres,
                     const DevScalars* sc, double* xnew, hipStream_t st) {
  hipLaunchKernelGGL(k_oc_final, dim3(grid_blocks(g.ne)), dim3(kBlock), 0, st, g.ne, x, dc, pres,
                     sc, xnew);
}

__global__ void k_set_preserved(int ne, const uint8_t* __restrict__ pres, double* xPhys) {
  body_set_preserved(ne, pres, xPhys, blockIdx.x, gridDim.x);
}
void launch_set_preserved(const Grid& g, const uint8_t* pres, double* xPhys, hipStream_t st) {
  if (pres == nullptr) return;
  hipLaunchKernelGGL(k_set_preserved, dim3(grid_blocks(g.ne)), dim3(kBlock), 0, st, g.ne, pres,
                     xPhys);
}

// ------------------------------------------------------------ dot (FP64)
template <typename TA, typename T>
__global__ __launch_bounds__(kBlock) void k_dot(long long n, const TA* __restrict__ a,
                                               const T* __restrict__ b, double* partS) {
  body_dot<TA, T>(n, a, b, partS, blockIdx.x, gridDim.x);
}
template <typename TA, typename T>
void launch_dot(const TA* a, const T* b, long long n, Parts parts, DevScalars* sc,
                hipStream_t st) {
  int nb = grid_blocks(n);
  nb = nb < kVecBlocks ? nb : kVecBlocks;
  hipLaunchKernelGGL((k_dot<TA, T>), dim3(nb), dim3(kBlock), 0, st, n, a, b, parts.S);
  hipLaunchKernelGGL(k_sum_fin, dim3(1), dim3(kBlock), 0, st, parts.S, nb, &sc->sum);
}

__global__ void k_reset_flags(DevScalars* sc) { sc->nan_flag = 0; }

__global__ void k_fill_int(int* __restrict__ v, long long n, int val) {
  const long long i = blockIdx.x * (long long)kBlock + threadIdx.x;
  if (i < n) v[i] = val;
}

__global__ void k_zero_scalars(DevScalars* sc) {
  constexpr int n = (int)(sizeof(DevScalars) / sizeof(int));
  for (int i = threadIdx.x; i < n; i += blockDim.x) reinterpret_cast<int*>(sc)[i] = 0;
}
void launch_zero_scalars(DevScalars* sc, hipStream_t st) {
  hipLaunchKernelGGL(k_zero_scalars, dim3(1), dim3(64), 0, st, sc);
}
void launch_reset_flags(DevScalars* sc, hipStream_t st) {
  hipLaunchKernelGGL(k_reset_flags, dim3(1), dim3(1), 0, st, sc);
}

// ------------------------------------------------- generic-table kernels
// Reference semantics (cuda_kernels/simp_sensitivity_ptx.cu:163-270, 301-373):
// arbitrary edof rows, E = max(x,1e-3)^p, FP64 element arithmetic.
// The reference kernel API (cuda_kernels/__init__.py:286-368) on explicit
// tables: element products in FP64 whatever the storage type T (float or
// double, the reference kernels take double).
template <typename T>
__global__ __launch_bounds__(kBlock) void k_generic_sens(const T* __restrict__ xPhys,
                                                        const T* __restrict__ u,
                                                        const T* __restrict__ ke,
                                                        const int32_t* __restrict__ edof,
                                                        long long ne, double p,
                                                        T* __restrict__ dc) {
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
  dc[e] = (T)(-p * pow(fmax((double)xPhys[e], 1e-3), p - 1.0) * ce);
}
template <typename T>
void launch_generic_sens(const T* xPhys, const T* u, const T* ke, const int32_t* edof,
                         long long ne, double penal, T* dc, hipStream_t st) {
  hipLaunchKernelGGL((k_generic_sens<T>), dim3(grid_blocks(ne)), dim3(kBlock), 0, st, xPhys, u,
                     ke, edof, ne, penal, dc);
}
template void launch_generic_sens<float>(const float*, const float*, const float*, const int32_t*,
                                         long long, double, float*, hipStream_t);
template void launch_generic_sens<double>(const double*, const double*, const double*,
                                          const int32_t*, long long, double, double*, hipStream_t);

// y = sum_e max(x_e, 1e-3)^p KE u_e scattered by an arbitrary edof, in gather
// form so that every DOF sums its terms in one fixed order (element, then
// local index). The DOF -> (element, local index) lists are built on the
// device for each call: integer counts, an exclusive scan, a fill, and a
// per-DOF sort of its (short) list by key 24 e + a.
__global__ void k_t_count(const int32_t* __restrict__ edof, long long nk, int* __restrict__ cnt) {
  const long long q = blockIdx.x * (long long)kBlock + threadIdx.x;
  if (q < nk) atomicAdd(&cnt[edof[q]], 1);
}
// per-block sums of cnt[0..n) into bs[blk]
__global__ __launch_bounds__(kBlock) void k_t_bsum(const int* __restrict__ cnt, long long n,
                                                  int* __restrict__ bs) {
  const long long q = blockIdx.x * (long long)kBlock + threadIdx.x;
  __shared__ int sh[kBlock];
  sh[threadIdx.x] = q < n ? cnt[q] : 0;
  __syncthreads();
  if (threadIdx.x == 0) {
    int t = 0;
    for (int i = 0; i < kBlock; ++i) t += sh[i];
    bs[blockIdx.x] = t;
  }
}
// exclusive scan of bs[0..nb) in place (one block, contiguous chunks)
__global__ __launch_bounds__(kBlock) void k_t_bscan(int* __restrict__ bs, int nb) {
  __shared__ int sh[kBlock];
  const int t = threadIdx.x;
  const int chunk = (nb + kBlock - 1) / kBlock;
  const int b0 = t * chunk, b1 = b0 + chunk < nb ? b0 + chunk : nb;
  int s = 0;
  for (int b = b0; b < b1; ++b) s += bs[b];
  sh[t] = s;
  __syncthreads();
  if (t == 0) {
    int acc = 0;
    for (int u = 0; u < kBlock; ++u) {
      const int x = sh[u];
      sh[u] = acc;
      acc += x;
    }
  }
  __syncthreads();
  int acc = sh[t];
  for (int b = b0; b < b1; ++b) {
    const int x = bs[b];
    bs[b] = acc;
    acc += x;
  }
}
// off[d] = exclusive prefix of cnt; fillpos[d] = 0
__global__ __launch_bounds__(kBlock) void k_t_apply(const int* __restrict__ cnt, long long n,
                                                   const int* __restrict__ bs,
                                                   int* __restrict__ off, int* __restrict__ pos) {
  const long long q = blockIdx.x * (long long)kBlock + threadIdx.x;
  __shared__ int sh[kBlock];
  sh[threadIdx.x] = q < n ? cnt[q] : 0;
  __syncthreads();
  if (threadIdx.x == 0) {
    int acc = bs[blockIdx.x];
    for (int i = 0; i < kBlock; ++i) {
      const int x = sh[i];
      sh[i] = acc;
      acc += x;
    }
  }
  __syncthreads();
  if (q < n) {
    off[q] = sh[threadIdx.x];
    pos[q] = 0;
  }
}
__global__ void k_t_fill(const int32_t* __restrict__ edof, long long nk,
                         const int* __restrict__ off, int* __restrict__ pos,
                         int* __restrict__ keys) {
  const long long q = blockIdx.x * (long long)kBlock + threadIdx.x;
  if (q >= nk) return;
  const int d = edof[q];
  keys[off[d] + atomicAdd(&pos[d], 1)] = (int)q;  // key = 24 e + a
}
__global__ void k_t_sort(const int* __restrict__ off, const int* __restrict__ cnt, long long n,
                         int* __restrict__ keys) {
  const long long d = blockIdx.x * (long long)kBlock + threadIdx.x;
  if (d >= n) return;
  int* k = keys + off[d];
  const int m = cnt[d];
  for (int i = 1; i < m; ++i) {
    const int v = k[i];
    int j = i - 1;
    while (j >= 0 && k[j] > v) {
      k[j + 1] = k[j];
      --j;
    }
    k[j + 1] = v;
  }
}
template <typename T>
__global__ __launch_bounds__(kBlock) void k_t_spmv(const T* __restrict__ xPhys,
                                                  const T* __restrict__ u,
                                                  const T* __restrict__ ke,
                                                  const int32_t* __restrict__ edof,
                                                  const int* __restrict__ off,
                                                  const int* __restrict__ cnt,
                                                  const int* __restrict__ keys, long long n,
                                                  double p, T* __restrict__ y) {
  const long long d = blockIdx.x * (long long)kBlock + threadIdx.x;
  if (d >= n) return;
  double acc = 0.0;
  const int o = off[d], m = cnt[d];
  for (int i = 0; i < m; ++i) {
    const int key = keys[o + i];
    const long long e = key / 24;
    const int a = key - (int)(e * 24);
    double t = 0.0;
    for (int b = 0; b < 24; ++b) t += (double)ke[a * 24 + b] * (double)u[edof[e * 24 + b]];
    acc += pow(fmax((double)xPhys[e], 1e-3), p) * t;
  }
  y[d] = (T)acc;
}

long long generic_spmv_ws_ints(long long ne, long long ndof) {
  const long long nb = (ndof + kBlock - 1) / kBlock;
  return 3 * ndof + 24 * ne + nb + 1;
}

template <typename T>
void launch_generic_spmv(const T* xPhys, const T* u, const T* ke, const int32_t* edof,
                         long long ne, long long ndof, double penal, int* ws, T* y,
                         hipStream_t st) {
  const long long nk = 24 * ne;
  const int nb = (int)((ndof + kBlock - 1) / kBlock);
  int* cnt = ws;
  int* off = ws + ndof;
  int* pos = ws + 2 * ndof;
  int* keys = ws + 3 * ndof;
  int* bs = keys + nk;
  hipLaunchKernelGGL(k_fill_int, dim3(grid_blocks(ndof)), dim3(kBlock), 0, st, cnt, ndof, 0);
  hipLaunchKernelGGL(k_t_count, dim3(grid_blocks(nk)), dim3(kBlock), 0, st, edof, nk, cnt);
  hipLaunchKernelGGL(k_t_bsum, dim3(nb), dim3(kBlock), 0, st, cnt, ndof, bs);
  hipLaunchKernelGGL(k_t_bscan, dim3(1), dim3(kBlock), 0, st, bs, nb);
  hipLaunchKernelGGL(k_t_apply, dim3(nb), dim3(kBlock), 0, st, cnt, ndof, bs, off, pos);
  hipLaunchKernelGGL(k_t_fill, dim3(grid_blocks(nk)), dim3(kBlock), 0, st, edof, nk, off, pos,
                     keys);
  hipLaunchKernelGGL(k_t_sort, dim3(grid_blocks(ndof)), dim3(kBlock), 0, st, off, cnt, ndof,
                     keys);
  hipLaunchKernelGGL((k_t_spmv<T>), dim3(grid_blocks(ndof)), dim3(kBlock), 0, st, xPhys, u, ke,
                     edof, off, cnt, keys, ndof, penal, y);
}
template void launch_generic_spmv<float>(const float*, const float*, const float*,
                                         const int32_t*, long long, long long, double, int*,
                                         float*, hipStream_t);
template void launch_generic_spmv<double>(const double*, const double*, const double*,
                                          const int32_t*, long long, long long, double, int*,
                                          double*, hipStream_t);

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
                               const uint8_t*, const DevScalars*, double*, bool, hipStream_t);
template void launch_ku<double>(const Grid&, const double*, const float*, const double*, double*,
                                const uint8_t*, const DevScalars*, double*, bool, hipStream_t);
template void launch_diag_minv<float>(const Grid&, const float*, const float*, const uint8_t*,
                                      float*, bool, hipStream_t);
template void launch_diag_minv<double>(const Grid&, const double*, const float*, const uint8_t*,
                                       double*, bool, hipStream_t);
template void launch_mask_fixed<float>(const Grid&, float*, const uint8_t*, hipStream_t);
template void launch_mask_fixed<double>(const Grid&, double*, const uint8_t*, hipStream_t);
template void launch_pcg_init<float>(const Grid&, const double*, const float*, bool, float*,
                                     float*, float*, float*, const float*, Parts, DevScalars*,
                                     double, int, hipStream_t);
template void launch_pcg_init<double>(const Grid&, const double*, const double*, bool, double*,
                                      double*, double*, double*, const double*, Parts,
                                      DevScalars*, double, int, hipStream_t);
template void launch_pcg_b<float>(const Grid&, int, int, float*, float*, float*, const float*,
                                  const float*, const float*, Parts, DevScalars*, hipStream_t);
template void launch_pcg_b<double>(const Grid&, int, int, double*, double*, double*,
                                   const double*, const double*, const double*, Parts,
                                   DevScalars*, hipStream_t);
template void launch_pcg_c<float>(const Grid&, int, float*, const float*, Parts, DevScalars*,
                                  hipStream_t);
template void launch_pcg_c<double>(const Grid&, int, double*, const double*, Parts, DevScalars*,
                                   hipStream_t);
template void launch_add_correction<float>(const Grid&, double*, const float*, hipStream_t);
template void launch_add_correction<double>(const Grid&, double*, const double*, hipStream_t);
template void launch_sens<float>(const Grid&, const double*, const ModalCoef<double>*,
                                 const double*, const float*, int, double, double, double,
                                 double*, double*, Parts, DevScalars*, hipStream_t);
template void launch_sens<double>(const Grid&, const double*, const ModalCoef<double>*,
                                  const double*, const double*, int, double, double, double,
                                  double*, double*, Parts, DevScalars*, hipStream_t);
template void launch_dot<double, float>(const double*, const float*, long long, Parts,
                                        DevScalars*, hipStream_t);
template void launch_dot<double, double>(const double*, const double*, long long, Parts,
                                         DevScalars*, hipStream_t);
template void launch_dot<float, float>(const float*, const float*, long long, Parts, DevScalars*,
                                      hipStream_t);

}  // namespace gp
