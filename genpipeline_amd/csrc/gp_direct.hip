// Exact FP64 direct solve of the free-DOF system K(xPhys) u = f on the device:
// the reference CPU solver's method (SciPy SuperLU, simp_solver.py:195-213).
// It works with element tables that no Krylov method converges on, such as
// the reference's own asymmetric, indefinite table (simp_solver.py:284-1543).
//
// Method: banded LU with partial pivoting (row interchanges), right-looking,
// one column per launch. Free DOFs are numbered in DOF order, so the free-free
// block of a structured grid has lower and upper bandwidth
// bw <= 3 (sI + sJ + 1) + 2 (sI, sJ: node strides). Interchanges widen the
// upper band of U to 2 bw.
//
// Storage: row r of the band lives in a physical slot of W = 2 bw + 2
// doubles; column j of that row sits at position j mod W. A row active at
// step k holds the live columns [k, k + 2 bw] (2 bw + 1 positions); the one
// spare position is that of column k - 1, which step k zeroes so that column
// k + 2 bw + 1 starts from zero when fill reaches it at step k + 1.
// Interchanges move no data: a map from logical row to physical slot is
// updated instead (two maps, read one / write the other per step, so blocks
// of a step never race on it). The right-hand side is eliminated alongside
// (it is stored per physical slot), so L is never kept: one factorisation per
// right-hand side, as SuperLU's spsolve does.
//
// Every value is produced by one thread in a fixed order: the result is the
// same bits on every call. Data in HBM for nf free DOFs: nf W doubles of band
// plus O(nf) vectors.
#include "gp_bodies.hpp"

namespace gp {

namespace {

constexpr int kStepBlock = 256;  // 4 waves, one band row each
constexpr int kSubBlock = 1024;  // back substitution: one workgroup

__global__ void k_zero_f64(double* __restrict__ v, long long n) {
  for (long long i = blockIdx.x * (long long)kBlock + threadIdx.x; i < n;
       i += (long long)gridDim.x * kBlock)
    v[i] = 0.0;
}

__global__ void k_iota2(int* __restrict__ m, int nf) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < nf) {
    m[i] = i;
    m[nf + i] = i;
  }
}

// Entry (row (n,r), column (m,c)) of the free-free block, m in the 27-node
// neighbourhood of n: sum over the elements holding both nodes of
// E(xPhys_e) ke[3 Ln + r][3 Lm + c] (y_n = sum_e E_e ke u_e, the orientation
// of every other operator in the library), elements in a fixed order.
__global__ __launch_bounds__(kBlock) void k_band_assemble(Grid g, const double* __restrict__ ke,
                                                          const double* __restrict__ xPhys,
                                                          int law, double penal, double E0,
                                                          double Emin,
                                                          const int* __restrict__ fi, int W,
                                                          double* __restrict__ band) {
  const long long q = blockIdx.x * (long long)kBlock + threadIdx.x;
  if (q >= (long long)g.nn * 243) return;
  const int rc = (int)(q % 9), o = (int)((q / 9) % 27);
  const int n = (int)(q / 243);
  const int r = rc / 3, c = rc % 3;
  const int row = fi[3ll * n + r];
  if (row < 0) return;
  const int i = n / g.sI, rem = n - i * g.sI, j = rem / g.sJ, k = rem - j * g.sJ;
  const int mi = i + o / 9 - 1, mj = j + (o / 3) % 3 - 1, mk = k + o % 3 - 1;
  if ((unsigned)mi > (unsigned)g.nx || (unsigned)mj > (unsigned)g.ny ||
      (unsigned)mk > (unsigned)g.nz)
    return;
  const int col = fi[3ll * (mi * g.sI + mj * g.sJ + mk) + c];
  if (col < 0) return;
  double s = 0.0;
  bool any = false;
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b)
      for (int cc = 0; cc < 2; ++cc) {
        // element (i-1+a, j-1+b, k-1+cc) holds node n at corner (1-a, 1-b, 1-cc)
        const int ei = i - 1 + a, ej = j - 1 + b, ek = k - 1 + cc;
        if ((unsigned)ei >= (unsigned)g.nx || (unsigned)ej >= (unsigned)g.ny ||
            (unsigned)ek >= (unsigned)g.nz)
          continue;
        const int mx = mi - ei, my = mj - ej, mz = mk - ek;  // m's corner in that element
        if ((unsigned)mx > 1u || (unsigned)my > 1u || (unsigned)mz > 1u) continue;
        const int Ln = LOC[(1 - a) * 4 + (1 - b) * 2 + (1 - cc)];
        const int Lm = LOC[mx * 4 + my * 2 + mz];
        const double Ee = law_E(xPhys[ei * g.eI + ej * g.eJ + ek], law, penal, E0, Emin);
        s += Ee * ke[(3 * Ln + r) * 24 + 3 * Lm + c];
        any = true;
      }
  if (any) band[(long long)row * W + col % W] = s;
}

__global__ void k_gather_free(const double* __restrict__ f, const int* __restrict__ fi,
                              long long ndof, double* __restrict__ bf) {
  const long long d = blockIdx.x * (long long)kBlock + threadIdx.x;
  if (d < ndof && fi[d] >= 0) bf[fi[d]] = f[d];
}

__global__ void k_scatter_free(const double* __restrict__ xf, const int* __restrict__ fi,
                               long long ndof, double* __restrict__ u) {
  const long long d = blockIdx.x * (long long)kBlock + threadIdx.x;
  if (d < ndof) u[d] = fi[d] >= 0 ? xf[fi[d]] : 0.0;
}

// Column 0 of rows 0..min(bw, nf-1): the candidates of the first pivot.
__global__ void k_cand0(const double* __restrict__ band, int W, int nf, int bw,
                        double* __restrict__ cand, int* __restrict__ info) {
  const int t = blockIdx.x * kBlock + threadIdx.x;
  if (t == 0) *info = 0;
  if (t <= bw && t < nf) cand[t] = band[(long long)t * W];
}

// Step k: pick the pivot among logical rows k..kend (the first largest
// |column k| value; cand[t] holds column k of row k + t), record the slot
// that now holds U row k, and subtract l_i times the pivot row from each
// row i in (k, kend] and from its right-hand side. One wave per row; wave gw
// always takes the rows i = gw (mod 4 gridDim.x), so a row stays with one
// workgroup (and one L2) for the steps it is active in.
__global__ __launch_bounds__(kStepBlock) void k_lu_step(
    double* __restrict__ band, int W, int nf, int bw, int k, const int* __restrict__ sm,
    int* __restrict__ dm, int* __restrict__ urow, double* __restrict__ b,
    const double* __restrict__ cand, double* __restrict__ cand_next, int* __restrict__ info) {
  extern __shared__ double s_u[];  // pivot row, columns k+1 .. kendc
  __shared__ double s_v[kStepBlock / 64];
  __shared__ int s_t[kStepBlock / 64];
  const int uw = 2 * bw;
  const int kend = min(k + bw, nf - 1), kendc = min(k + uw, nf - 1);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  // pivot search (every workgroup, same order: the same answer)
  double best = -1.0;
  int bt = 0;
  for (int t = threadIdx.x; t <= kend - k; t += kStepBlock) {
    const double v = fabs(cand[t]);
    if (v > best) {
      best = v;
      bt = t;
    }
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const double ov = __shfl_xor(best, off, 64);
    const int ot = __shfl_xor(bt, off, 64);
    if (ov > best || (ov == best && ot < bt)) {
      best = ov;
      bt = ot;
    }
  }
  if (lane == 0) {
    s_v[wave] = best;
    s_t[wave] = bt;
  }
  __syncthreads();
  best = s_v[0];
  bt = s_t[0];
  for (int w = 1; w < kStepBlock / 64; ++w)
    if (s_v[w] > best || (s_v[w] == best && s_t[w] < bt)) {
      best = s_v[w];
      bt = s_t[w];
    }
  const int p = k + bt;
  const double piv = cand[bt];
  const int slot_k = sm[k], slot_p = sm[p];
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    urow[k] = slot_p;
    if (!(fabs(piv) > 0.0) && *info == 0) *info = k + 1;  // zero or NaN pivot
    if (k + 1 + bw <= nf - 1)  // the row entering the window at step k+1
      cand_next[bw] = band[(long long)(k + 1 + bw) * W + (k + 1) % W];
  }
  if (!(fabs(piv) > 0.0)) return;  // uniform: every thread read the same piv
  const double* urow_p = band + (long long)slot_p * W;
  for (int j = k + 1 + threadIdx.x; j <= kendc; j += kStepBlock) s_u[j - k - 1] = urow_p[j % W];
  const double bp = b[slot_p];
  __syncthreads();
  const int nw = 4 * gridDim.x;
  const int gw = blockIdx.x * 4 + wave;
  const int i = k + 1 + ((gw - (k + 1)) % nw + nw) % nw;
  if (i > kend) return;
  const int slot = i == p ? slot_k : sm[i];
  const double a = i == p ? cand[0] : cand[i - k];
  const double l = a / piv;
  double* row = band + (long long)slot * W;
  if (lane == 0 && k >= 1) row[(k - 1) % W] = 0.0;
  if (l != 0.0) {
    for (int j = k + 1 + lane; j <= kendc; j += 64) {
      const int pos = j % W;
      row[pos] = row[pos] - l * s_u[j - k - 1];
    }
  }
  if (lane == 0) {
    if (l != 0.0) b[slot] = b[slot] - l * bp;
    dm[i] = slot;
    cand_next[i - k - 1] = row[(k + 1) % W];  // this lane wrote column k+1 itself
  }
}

// U x = y from the bottom row up; y and the rows of U are in the slots urow
// names. x also goes to a ring of W doubles in LDS (x[j] at j mod W), which
// holds every x[j] the dot of row k reads (j <= k + 2 bw < k + W).
__global__ __launch_bounds__(kSubBlock) void k_lu_backsub(const double* __restrict__ band, int W,
                                                          int nf, int bw,
                                                          const int* __restrict__ urow,
                                                          const double* __restrict__ b,
                                                          double* __restrict__ x) {
  extern __shared__ double s_x[];
  __shared__ double s_part[kSubBlock / 64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int k = nf - 1; k >= 0; --k) {
    const int slot = urow[k];
    const double* row = band + (long long)slot * W;
    const int kendc = min(k + 2 * bw, nf - 1);
    double s = 0.0;
    for (int j = k + 1 + threadIdx.x; j <= kendc; j += kSubBlock) s += row[j % W] * s_x[j % W];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
    if (lane == 0) s_part[wave] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
      double t = 0.0;
      for (int w = 0; w < kSubBlock / 64; ++w) t += s_part[w];
      const double xk = (b[slot] - t) / row[k % W];
      s_x[k % W] = xk;
      x[k] = xk;
    }
    __syncthreads();
  }
}

}  // namespace

int direct_band_width(const Grid& g) { return 3 * (g.sI + g.sJ + 1) + 2; }

void launch_direct(const Grid& g, const double* ke, const double* xPhys, int law, double penal,
                   double E0, double Emin, const double* f, const DirectWs& w, double* u,
                   hipStream_t st) {
  const int nf = w.nf, bw = w.bw, W = w.W;
  if (nf == 0) {
    hipLaunchKernelGGL(k_zero_f64, dim3(grid_blocks(g.ndof)), dim3(kBlock), 0, st, u,
                       (long long)g.ndof);
    return;
  }
  const long long nb = (long long)nf * W;
  hipLaunchKernelGGL(k_zero_f64, dim3(std::min<long long>(grid_blocks(nb), 8192)), dim3(kBlock), 0,
                     st, w.band, nb);
  hipLaunchKernelGGL(k_iota2, dim3(grid_blocks(nf)), dim3(kBlock), 0, st, w.map, nf);
  hipLaunchKernelGGL(k_band_assemble, dim3(grid_blocks((long long)g.nn * 243)), dim3(kBlock), 0,
                     st, g, ke, xPhys, law, penal, E0, Emin, w.fi, W, w.band);
  hipLaunchKernelGGL(k_gather_free, dim3(grid_blocks(g.ndof)), dim3(kBlock), 0, st, f, w.fi,
                     (long long)g.ndof, w.b);
  hipLaunchKernelGGL(k_cand0, dim3(grid_blocks(bw + 1)), dim3(kBlock), 0, st, w.band, W, nf, bw,
                     w.cand, w.info);
  const int blocks = (bw + 3) / 4;
  const size_t lds = sizeof(double) * (size_t)(2 * bw);
  for (int k = 0; k < nf; ++k) {
    const int* sm = w.map + (k & 1) * nf;
    int* dm = w.map + ((k + 1) & 1) * nf;
    const double* c0 = w.cand + (k & 1) * (bw + 1);
    double* c1 = w.cand + ((k + 1) & 1) * (bw + 1);
    hipLaunchKernelGGL(k_lu_step, dim3(blocks), dim3(kStepBlock), lds, st, w.band, W, nf, bw, k,
                       sm, dm, w.urow, w.b, c0, c1, w.info);
  }
  hipLaunchKernelGGL(k_lu_backsub, dim3(1), dim3(kSubBlock), sizeof(double) * (size_t)W, st,
                     w.band, W, nf, bw, w.urow, w.b, w.x);
  hipLaunchKernelGGL(k_scatter_free, dim3(grid_blocks(g.ndof)), dim3(kBlock), 0, st, w.x, w.fi,
                     (long long)g.ndof, u);
}

}  // namespace gp
