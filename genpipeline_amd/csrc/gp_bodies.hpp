// Device bodies of the element/node/vector kernels, written against a local
// block index `lb` out of `nb` blocks so that the same code runs as a
// single-problem kernel (lb = blockIdx.x) or as one problem's share of a
// batched launch (lb = blockIdx.x - first block of that problem).
#pragma once
// Multiply-adds are fused only where one source expression allows it, so the
// rounding of a body does not depend on where the compiler inlines it (the
// single-problem and batched kernels then agree bit for bit).
#pragma clang fp contract(on)
#include "gp_internal.hpp"
#include "gp_modal.hpp"

namespace gp {

// C3D8 corner offsets (simp_solver.py:255-267) and the inverse map
// (cx,cy,cz) -> local corner index, keyed by cx*4 + cy*2 + cz.
__device__ constexpr int CX[8] = {0, 1, 1, 0, 0, 1, 1, 0};
__device__ constexpr int CY[8] = {0, 0, 1, 1, 0, 0, 1, 1};
__device__ constexpr int CZ[8] = {0, 0, 0, 0, 1, 1, 1, 1};
__device__ constexpr int LOC[8] = {0, 4, 3, 7, 1, 5, 2, 6};

// ------------------------------------------------------------ reductions
__device__ inline double wave_max(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
  return v;
}

// Block sum (blockDim == kBlock); result in thread 0.
__device__ inline double block_sum0(double v) {
  __shared__ double sh[kBlock / 64];
  v = wave_sum_nb<kBlock>(v);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0)
    for (int i = 0; i < kBlock / 64; ++i) s += sh[i];
  __syncthreads();
  return s;
}

// Sum of p[0..n) in a fixed order, returned to every thread of the block.
__device__ inline double reduce_all(const double* __restrict__ p, int n) {
  __shared__ double sh[kBlock / 64];
  double s = 0.0;
  for (int i = threadIdx.x; i < n; i += kBlock) s += p[i];
  s = wave_sum_nb<kBlock>(s);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
  __syncthreads();
  double t = 0.0;
  for (int i = 0; i < kBlock / 64; ++i) t += sh[i];
  __syncthreads();
  return t;
}

__device__ inline double reduce_max_all(const double* __restrict__ p, int n) {
  __shared__ double sh[kBlock / 64];
  double s = -INFINITY;
  for (int i = threadIdx.x; i < n; i += kBlock) s = fmax(s, p[i]);
  s = wave_max(s);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
  __syncthreads();
  double t = -INFINITY;
  for (int i = 0; i < kBlock / 64; ++i) t = fmax(t, sh[i]);
  __syncthreads();
  return t;
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

__device__ inline double law_E(double x, int law, double p, double E0, double Emin) {
  if (law == GP_LAW_CLAMP) return pow(fmax(x, 1e-3), p);
  return Emin + pow(x, p) * (E0 - Emin);
}
__device__ inline double law_dE(double x, int law, double p, double E0, double Emin) {
  if (law == GP_LAW_CLAMP) return p * pow(fmax(x, 1e-3), p - 1.0);
  return p * pow(x, p - 1.0) * (E0 - Emin);
}

__device__ inline double oc_xnew(double x, double dc, double lmid, double floor_, double move) {
  double v = x * sqrt(fmax(-dc / lmid, floor_));
  v = fmin(fmax(v, 0.0), 1.0);
  return fmin(fmax(v, x - move), x + move);
}

// Index of the entry whose block range [off[i], off[i+1]) holds block blk.
__device__ inline int find_prob(const int* __restrict__ off, int n, int blk) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (off[mid] <= blk)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

// ------------------------------------------------------------- bodies
__device__ inline void body_sum_fin(const double* __restrict__ p, int n, double* out) {
  const double s = reduce_all(p, n);
  if (threadIdx.x == 0) *out = s;
}

__device__ inline void body_stiffness(int ne, const double* __restrict__ x, float* __restrict__ E,
                                      int law, double p, double E0, double Emin, int lb, int nb) {
  for (int e = lb * kBlock + threadIdx.x; e < ne; e += nb * kBlock)
    E[e] = (float)law_E(x[e], law, p, E0, Emin);
}

// diag(K) per DOF; with invert 1: Jacobi M^-1 (0 on fixed DOFs and on DOFs
// with a non-positive diagonal). invert 2: one M^-1 per node, from the x
// component, unmasked (out has n_nodes entries). For the isotropic element
// the three diagonal entries of a node are equal; the float sweep reads this
// form and the support bits stay with the recurrence (r, and so z, is 0 on
// a fixed DOF from the set-up on).
template <typename T>
__device__ inline void body_diag(const Grid& g, const T* __restrict__ ke,
                                 const float* __restrict__ E, const uint8_t* __restrict__ fixmask,
                                 T* __restrict__ out, int invert, int lb, int nb) {
  for (int n = lb * kBlock + threadIdx.x; n < g.nn; n += nb * kBlock) {
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
    if (invert == 2) {
      out[n] = d0 > 0 ? T(1) / d0 : T(0);
      continue;
    }
    if (invert) {
      const uint8_t fm = fixmask ? fixmask[n] : 0;
      d0 = ((fm & 1) || !(d0 > 0)) ? T(0) : T(1) / d0;
      d1 = ((fm & 2) || !(d1 > 0)) ? T(0) : T(1) / d1;
      d2 = ((fm & 4) || !(d2 > 0)) ? T(0) : T(1) / d2;
    }
    st3(out, n, d0, d1, d2);
  }
}

// FP64 refinement residual r = f - K(E) u on the free DOFs (0 on a fixed
// one). One thread per node sums, over the eight elements around it, E_e
// times rows 3L..3L+2 of the FP64 table (L: the node's corner in that
// element) applied to the element's corner values, in a fixed order. The
// table addresses are uniform across a wave, so they are scalar loads; an
// element outside the grid reads corner node 0 with modulus 0. Block sums of
// r.r and of the free part of f.f go to partR[lb], partB[lb]. Runs once per
// refinement sweep, not per PCG iteration.
__device__ inline void body_res64(const Grid& g, const double* __restrict__ ke,
                                  const float* __restrict__ E, const double* __restrict__ u,
                                  const double* __restrict__ f,
                                  const uint8_t* __restrict__ fixmask, double* __restrict__ r,
                                  double* partR, double* partB, int lb, int nb) {
  double rr = 0.0, bb = 0.0;
  for (int n = lb * kBlock + threadIdx.x; n < g.nn; n += nb * kBlock) {
    const int i = n / g.sI, rem = n - i * g.sI, j = rem / g.sJ, k = rem - j * g.sJ;
    double y0 = 0.0, y1 = 0.0, y2 = 0.0;
#pragma unroll 1
    for (int abc = 0; abc < 8; ++abc) {  // the eight elements around the node
          const int a = abc >> 2, b = (abc >> 1) & 1, c = abc & 1;
          const int ei = i - 1 + a, ej = j - 1 + b, ek = k - 1 + c;
          const bool ok = (unsigned)ei < (unsigned)g.nx && (unsigned)ej < (unsigned)g.ny &&
                          (unsigned)ek < (unsigned)g.nz;
          const int e = ok ? ei * g.eI + ej * g.eJ + ek : 0;
          const int n0 = ok ? ei * g.sI + ej * g.sJ + ek : 0;
          const double Ee = ok ? (double)E[e] : 0.0;
          const int L = LOC[(1 - a) * 4 + (1 - b) * 2 + (1 - c)];
          double s0 = 0.0, s1 = 0.0, s2 = 0.0;
#pragma unroll
          for (int m = 0; m < 8; ++m) {
            const V3<double> v = ld3(u, n0 + CX[m] * g.sI + CY[m] * g.sJ + CZ[m]);
            const double* r0 = ke + (3 * L + 0) * 24 + 3 * m;
            const double* r1 = ke + (3 * L + 1) * 24 + 3 * m;
            const double* r2 = ke + (3 * L + 2) * 24 + 3 * m;
            s0 += r0[0] * v.x + r0[1] * v.y + r0[2] * v.z;
            s1 += r1[0] * v.x + r1[1] * v.y + r1[2] * v.z;
            s2 += r2[0] * v.x + r2[1] * v.y + r2[2] * v.z;
          }
          y0 += Ee * s0;
          y1 += Ee * s1;
          y2 += Ee * s2;
        }
    const uint8_t fm = fixmask[n];
    const double f0 = (fm & 1) ? 0.0 : f[3ll * n];
    const double f1 = (fm & 2) ? 0.0 : f[3ll * n + 1];
    const double f2 = (fm & 4) ? 0.0 : f[3ll * n + 2];
    const double r0 = (fm & 1) ? 0.0 : f0 - y0;
    const double r1 = (fm & 2) ? 0.0 : f1 - y1;
    const double r2 = (fm & 4) ? 0.0 : f2 - y2;
    st3(r, n, r0, r1, r2);
    rr += r0 * r0 + r1 * r1 + r2 * r2;
    bb += f0 * f0 + f1 * f1 + f2 * f2;
  }
  rr = block_sum0(rr);
  bb = block_sum0(bb);
  if (threadIdx.x == 0) {
    partR[lb] = rr;
    partB[lb] = bb;
  }
}

// One block: the refinement residual's sums (partials of body_res64).
__device__ inline void body_res_fin(const Parts& parts, int nV, DevScalars* sc) {
  const double rr = reduce_all(parts.A, nV);
  const double bb = reduce_all(parts.RZ[1], nV);
  if (threadIdx.x == 0) {
    sc->ir_rr = rr;
    sc->ir_bb = bb;
  }
}

// Single-kernel PCG set-up: r_{-1} = b - A x0 on free DOFs (or b), x masked,
// s_{-2} = 0, w_{-1} = 0, p = 0; partials bb -> BB[lb], rr -> RR[lb].
// S: storage of x and p (T for FP32/FP64; a 16-bit type with T = float).
// A DOF is free when its M^-1 is nonzero: minv per DOF, or with node_m one
// value per node (body_diag's invert 2) and the support bits.
template <typename T, typename S = T>
__device__ inline void body_cg_init(int nn, const double* __restrict__ f, const T* __restrict__ Ax,
                                    int use_x0, S* __restrict__ x, T* __restrict__ r,
                                    T* __restrict__ s0, T* __restrict__ w1, S* __restrict__ p,
                                    const T* __restrict__ minv, const uint8_t* __restrict__ fixmask,
                                    int node_m, const Parts& parts, int lb, int nb) {
  double rr = 0.0, bb = 0.0;
  for (long long d = lb * (long long)kBlock + threadIdx.x; d < 3ll * nn;
       d += (long long)nb * kBlock) {
    bool freed;
    if (node_m) {
      const long long n = d / 3;
      const int c = (int)(d - 3 * n);
      const bool fixed = fixmask != nullptr && ((fixmask[n] >> c) & 1);
      freed = !fixed && minv[n] != T(0);
    } else {
      freed = minv[d] != T(0);
    }
    const double bv = freed ? f[d] : 0.0;
    const S xd = (use_x0 && freed) ? x[d] : st_s<S>(T(0));
    const T rd = freed ? (T)(bv - (use_x0 ? (double)Ax[d] : 0.0)) : T(0);
    x[d] = xd;
    r[d] = rd;
    s0[d] = T(0);
    w1[d] = T(0);
    p[d] = st_s<S>(T(0));
    rr += (double)rd * (double)rd;
    bb += bv * bv;
  }
  rr = block_sum0(rr);
  bb = block_sum0(bb);
  if (threadIdx.x == 0) {
    parts.RR[lb] = rr;
    parts.BB[lb] = bb;
    parts.RZ[0][lb] = 0.0;
  }
}

// One block: finishes the set-up sums and the initial convergence test.
__device__ inline void body_pcg_init_fin(const Parts& parts, int nV, double tol, int max_iter,
                                         DevScalars* sc) {
  const double bb = reduce_all(parts.BB, nV);
  const double rr = reduce_all(parts.RR, nV);
  const double rz = reduce_all(parts.RZ[0], nV);
  if (threadIdx.x == 0) {
    sc->bb = bb;
    sc->rr = rr;
    sc->rz = rz;
    sc->tol2 = tol * tol * bb;
    sc->max_iter = max_iter;
    sc->iter = 0;
    sc->status = GP_OK;
    sc->done = (bb == 0.0 || rr <= tol * tol * bb) ? 1 : 0;
  }
}

// Element energy and sensitivity, FP64 (simp_solver.py:215-218). With
// want_sum the block's sum of E_e u_e^T KE u_e goes to partS[lb].
// u_e^T KE u_e in the modal basis (KE = T Khat T^T / 64, coefficients of k
// already divided by 64): with b = T^T u_e and v = Khat b / 64 the energy is
// b.v, 45 products instead of the 576 of the dense form.
__device__ inline double elem_energy_modal(const ModalCoef<double>& k, const double (&ue)[24]) {
  Plane<double> pl[2];
#pragma unroll
  for (int cx = 0; cx < 2; ++cx)
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      double c[2][2];
#pragma unroll
      for (int cy = 0; cy < 2; ++cy)
#pragma unroll
        for (int cz = 0; cz < 2; ++cz) c[cy][cz] = ue[3 * LOC[cx * 4 + cy * 2 + cz] + q];
      const double a0 = c[0][0] + c[0][1], d0 = c[0][1] - c[0][0];
      const double a1 = c[1][0] + c[1][1], d1 = c[1][1] - c[1][0];
      pl[cx].P[q][0] = a0 + a1;
      pl[cx].P[q][1] = d0 + d1;
      pl[cx].P[q][2] = a1 - a0;
      pl[cx].P[q][3] = d1 - d0;
    }
  double b0[3][4], b1[3][4], v0[3][4], v1[3][4];
  element_modes(k, 1.0, pl[0], pl[1], b0, b1, v0, v1);
  double ce = 0.0;
#pragma unroll
  for (int q = 0; q < 3; ++q)
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      ce += b0[q][m] * v0[q][m];
      ce += b1[q][m] * v1[q][m];
    }
  return ce;
}

// dc_e = -dE/dx u_e^T KE u_e. MODAL: the element table is the isotropic hex8
// matrix held as modal coefficients mk; otherwise the dense table ke.
template <typename T, bool MODAL>
__device__ inline void body_sens(const Grid& g, const double* __restrict__ ke,
                                 const ModalCoef<double>& mk, const double* __restrict__ xPhys,
                                 const T* __restrict__ u, int law, double p, double E0,
                                 double Emin, double* __restrict__ dc, int want_sum,
                                 double* partS, DevScalars* sc, int lb, int nb) {
  double esum = 0.0;
  const int nloop = (g.ne + nb * kBlock - 1) / (nb * kBlock);
  for (int it = 0; it < nloop; ++it) {
    const int e = (it * nb + lb) * kBlock + threadIdx.x;
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
    if (MODAL) {
      ce = elem_energy_modal(mk, ue);
    } else {
#pragma unroll
      for (int b = 0; b < 24; ++b) {
        double t = 0.0;
#pragma unroll
        for (int a = 0; a < 24; ++a) t += ue[a] * ke[a * 24 + b];
        ce += t * ue[b];
      }
    }
    const double x = valid ? xPhys[e] : 0.0;
    const double dcv = -law_dE(x, law, p, E0, Emin) * ce;
    if (valid) {
      dc[e] = dcv;
      if (dcv != dcv) sc->nan_flag = 1;
      esum += law_E(x, law, p, E0, Emin) * ce;
    }
  }
  if (want_sum) {
    const double s = block_sum0(esum);
    if (threadIdx.x == 0) partS[lb] = s;
  }
}

// out_e = (sum_o w_o in_{e+o}) / Hs_e            (mode 0, simp_solver.py:240)
// out_e =  sum_o w_o (in_{e+o} / Hs_{e+o})       (mode 1, simp_solver.py:223)
// Offsets run in lexicographic (di,dj,dk) order, i.e. ascending neighbour
// index, and products are not contracted into FMAs: the same operation order
// as a CSR row product of the assembled filter matrix.
__device__ inline void body_filter(const Grid& g, const Stencil* __restrict__ s,
                                   const double* __restrict__ Hs, const double* __restrict__ in,
                                   double* __restrict__ out, int mode, int lb, int nb) {
#pragma clang fp contract(off)
  const int no = s->n;
  for (int e = lb * kBlock + threadIdx.x; e < g.ne; e += nb * kBlock) {
    const int i = e / g.eI, rem = e - i * g.eI, j = rem / g.eJ, k = rem - j * g.eJ;
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
}

// per-block max(-dc) (gpu variant: l2 = 2 max(max(-dc), 1))
__device__ inline void body_oc_max(int ne, const double* __restrict__ dc, double* partM, int lb,
                                   int nb) {
  double m = -INFINITY;
  for (int e = lb * kBlock + threadIdx.x; e < ne; e += nb * kBlock) m = fmax(m, -dc[e]);
  m = wave_max(m);
  __shared__ double sh[kBlock / 64];
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int i = 0; i < kBlock / 64; ++i) m = fmax(m, sh[i]);
    partM[lb] = m;
  }
}

// One block: l1 = 0, l2 = 1e9 (cpu constants) or 2 max(max(-dc), 1) (gpu).
__device__ inline void body_oc_init_fin(const Parts& parts, int nb, int variant, double volfrac,
                                        double move, DevScalars* sc) {
  const double mm = variant == GP_OC_GPU ? reduce_max_all(parts.M, nb) : 0.0;
  if (threadIdx.x == 0) {
    sc->ocl1[0] = 0.0;
    sc->ocl2[0] = (variant == GP_OC_GPU) ? fmax(mm, 1.0) * 2.0 : 1e9;
    sc->oc_eps = (variant == GP_OC_GPU) ? 1e-15 : 1e-12;
    sc->oc_tol = (variant == GP_OC_GPU) ? 1e-5 : 1e-4;
    sc->oc_floor = (variant == GP_OC_GPU) ? 1e-15 : 1e-12;
    sc->oc_max = (variant == GP_OC_GPU) ? 30 : 50;
    sc->volfrac = volfrac;
    sc->move = move;
    sc->oc_steps = 0;
    sc->oc_done = 0;
  }
}

// OC bisection (simp_solver.py:230-248), kOcLevels steps per launch. The
// one-step loop is: at step s with interval (l1, l2), lmid = (l2 + l1) / 2; it
// stops with lmid if the interval is narrow enough or s is the last permitted
// step, else it evaluates the volume at lmid and keeps the upper half
// (l1 = lmid) when the mean density exceeds volfrac, the lower half otherwise.
// Launch j starts at step s = kOcLevels*j and evaluates the volume at the
// midpoint of every interval the next kOcLevels steps can reach: a binary
// tree in heap order, node c with children 2c+1 (lower half) and 2c+2 (upper
// half); a node where the loop would stop is a leaf and is not evaluated.
// Launch j+1 reduces those sums and follows the loop's path through the
// tree. Every interval, midpoint, test and volume sum is formed by the same
// operations as in the one-step loop (each volume by the same per-block
// partials and fixed-order reduction), so the multiplier and the step count
// are bitwise the one-step loop's, in a third of the dependent launches.
// Returns true when the multiplier is final (every block of the launch
// reaches the same answer); oc_launches(oc_max) launches always suffice.
// The one-step loop stops at step s with interval (l1, l2).
__device__ inline bool oc_leaf(double l1, double l2, int s, const DevScalars* sc) {
  return (l2 - l1) / (l1 + l2 + sc->oc_eps) < sc->oc_tol || s + 1 >= sc->oc_max;
}

__device__ inline bool body_oc_step(int j, int ne, const double* __restrict__ x,
                                    const double* __restrict__ dc, const double* __restrict__ w,
                                    const uint8_t* __restrict__ pres, double n_pres,
                                    const Parts& parts, DevScalars* sc, int lb, int nb) {
  if (sc->oc_done) return true;
  const int s = kOcLevels * j;
  double l1, l2;
  if (j == 0) {
    l1 = sc->ocl1[0];
    l2 = sc->ocl2[0];
  } else {
    // the loop's path through launch j-1's tree
    const int pq = (j - 1) & 1;
    double a = sc->ocl1[pq], b = sc->ocl2[pq];
    int c = 0;
    for (int d = 0; d < kOcLevels; ++d) {
      const double mid = 0.5 * (b + a);
      if (oc_leaf(a, b, s - kOcLevels + d, sc)) {
        if (lb == 0 && threadIdx.x == 0) {
          sc->lmid = mid;
          sc->oc_steps = s - kOcLevels + d + 1;
          sc->oc_done = 1;
        }
        return true;
      }
      const double mean = (reduce_all(parts.O[pq] + (long long)c * nb, nb) + n_pres) / (double)ne;
      if (mean > sc->volfrac) {
        a = mid;
        c = 2 * c + 2;
      } else {
        b = mid;
        c = 2 * c + 1;
      }
    }
    l1 = a;
    l2 = b;
    if (lb == 0 && threadIdx.x == 0) {
      sc->ocl1[j & 1] = l1;
      sc->ocl2[j & 1] = l2;
    }
  }
  if (oc_leaf(l1, l2, s, sc)) {
    if (lb == 0 && threadIdx.x == 0) {
      sc->lmid = 0.5 * (l2 + l1);
      sc->oc_steps = s + 1;
      sc->oc_done = 1;
    }
    return true;
  }
  // this launch's tree: intervals, midpoints, evaluated nodes
  double lo[kOcCand], hi[kOcCand], mid[kOcCand];
  bool ev[kOcCand];
  lo[0] = l1;
  hi[0] = l2;
#pragma unroll
  for (int c = 0; c < kOcCand; ++c) {
    const int d = c >= 3 ? 2 : (c >= 1 ? 1 : 0);  // depth in the tree
    const bool live = c == 0 || ev[(c - 1) / 2];
    mid[c] = 0.5 * (hi[c] + lo[c]);
    ev[c] = live && !oc_leaf(lo[c], hi[c], s + d, sc);
    if (2 * c + 2 < kOcCand) {
      lo[2 * c + 1] = lo[c];
      hi[2 * c + 1] = mid[c];
      lo[2 * c + 2] = mid[c];
      hi[2 * c + 2] = hi[c];
    }
  }
  const double fl = sc->oc_floor, mv = sc->move;
  double acc[kOcCand];
#pragma unroll
  for (int c = 0; c < kOcCand; ++c) acc[c] = 0.0;
  for (int e = lb * kBlock + threadIdx.x; e < ne; e += nb * kBlock) {
    const bool keep = pres != nullptr && pres[e];
    const double xe = x[e], de = dc[e], we = w[e];
#pragma unroll
    for (int c = 0; c < kOcCand; ++c)
      if (ev[c]) {
        const double xn = keep ? 1.0 : oc_xnew(xe, de, mid[c], fl, mv);
        acc[c] += we * xn;
      }
  }
#pragma unroll
  for (int c = 0; c < kOcCand; ++c)
    if (ev[c]) {
      const double t = block_sum0(acc[c]);
      if (threadIdx.x == 0) parts.O[j & 1][(long long)c * nb + lb] = t;
    }
  return false;
}
static_assert(kOcLevels == 3, "the depth rule in body_oc_step's tree assumes three levels");

__device__ inline void body_oc_final(int ne, const double* __restrict__ x,
                                     const double* __restrict__ dc,
                                     const uint8_t* __restrict__ pres, const DevScalars* sc,
                                     double* __restrict__ xnew, int lb, int nb) {
  for (int e = lb * kBlock + threadIdx.x; e < ne; e += nb * kBlock)
    xnew[e] = (pres != nullptr && pres[e]) ? 1.0
                                           : oc_xnew(x[e], dc[e], sc->lmid, sc->oc_floor, sc->move);
}

__device__ inline void body_set_preserved(int ne, const uint8_t* __restrict__ pres, double* xPhys,
                                          int lb, int nb) {
  for (int e = lb * kBlock + threadIdx.x; e < ne; e += nb * kBlock)
    if (pres[e]) xPhys[e] = 1.0;
}

template <typename TA, typename T>
__device__ inline void body_dot(long long n, const TA* __restrict__ a, const T* __restrict__ b,
                                double* partS, int lb, int nb) {
  double acc = 0.0;
  for (long long i = lb * (long long)kBlock + threadIdx.x; i < n; i += (long long)nb * kBlock)
    acc += (double)a[i] * (double)b[i];
  acc = block_sum0(acc);
  if (threadIdx.x == 0) partS[lb] = acc;
}

}  // namespace gp
